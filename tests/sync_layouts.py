"""A two-level periodic layout with face-flux fields, for the flux-register
and average-down tests (CPU pins and GPU parity)."""
import numpy as np

from fill_layouts import fab_shape, level_params, split

N_FINE = (32, 32, 32)
R = 2
FINE = [((8, 8, 8), (15, 15, 15)), ((16, 8, 8), (23, 15, 15)),
        ((0, 0, 24), (7, 7, 31))]          # the last one wraps around the domain
COARSE = split((16, 16, 16), (2, 2, 2), 3)


def params(nfluids=1):
    return level_params(N_FINE, 3, (1, 1, 1), nfluids=nfluids)


def face_fabs(boxes, field, dims=3):
    """Per box, per direction: the face FAB (box grown by one face in that
    direction) cut from a periodic global face field field[d] (shape
    (nc, nz, ny, nx) over the domain's faces, face n == face 0)."""
    out, los = [], []
    for lo, hi in boxes:
        for d in range(dims):
            fhi = list(hi)
            fhi[d] += 1
            nc = field[d].shape[0]
            shp = fab_shape((lo, tuple(fhi)), nc)
            zz, yy, xx = np.meshgrid(*[np.arange(lo[e], fhi[e] + 1) for e in (2, 1, 0)],
                                     indexing="ij")
            n = field[d].shape[1:]
            a = field[d][:, zz % n[0], yy % n[1], xx % n[2]]
            assert a.shape == shp
            out.append(np.ascontiguousarray(a))
            los.append(lo)
    return out, los


def cell_fabs(boxes, field):
    """Per box: its valid cells cut from a global cell field (nc, nz, ny, nx)."""
    out, los = [], []
    for lo, hi in boxes:
        out.append(np.ascontiguousarray(
            field[:, lo[2]:hi[2] + 1, lo[1]:hi[1] + 1, lo[0]:hi[0] + 1]))
        los.append(lo)
    return out, los


def divergence(field, dx):
    """Periodic flux divergence of global face fields field[d] on cells."""
    out = np.zeros_like(field[0])
    for d in range(3):
        ax = 3 - d
        out += (np.roll(field[d], -1, axis=ax) - field[d]) / dx
    return out
