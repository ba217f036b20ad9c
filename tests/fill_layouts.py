"""Level layouts and data for the ghost-fill tests (exchange and
coarse-fine interpolation).  Boxes are (lo, hi) inclusive cell ranges."""
import numpy as np

import mfmhd

GHOST = np.nan  # value of ghost cells nothing has written


def level_params(n, dims=3, periodic=(0, 0, 0), nfluids=1, dedner=False):
    """Parameters of a level whose problem domain is [0, n)."""
    p = mfmhd.default_params()
    p.spacedim = dims
    p.nfluids = nfluids
    p.dedner = int(dedner)
    if dedner:
        p.eight_wave = 0  # GLM cleaning excludes the 8-wave source (main/MHDAM.cpp:1042)
    for d in range(3):
        p.domain_lo[d] = 0
        p.domain_hi[d] = n[d] - 1 if d < dims else 0
        p.periodic[d] = int(periodic[d]) if d < dims else 0
    return p


def split(n, parts, dims=3):
    """The box [0, n) cut into parts[d] equal pieces per direction."""
    out = []
    sizes = [n[d] // parts[d] if d < dims else 1 for d in range(3)]
    rng = [range(parts[d]) if d < dims else range(1) for d in range(3)]
    for c in rng[2]:
        for b in rng[1]:
            for a in rng[0]:
                idx = (a, b, c)
                lo = tuple(idx[d] * sizes[d] if d < dims else 0 for d in range(3))
                hi = tuple(lo[d] + sizes[d] - 1 if d < dims else 0 for d in range(3))
                out.append((lo, hi))
    return out


def grown(box, g, dims=3):
    lo, hi = box
    return (tuple(lo[d] - (g if d < dims else 0) for d in range(3)),
            tuple(hi[d] + (g if d < dims else 0) for d in range(3)))


def fab_shape(box, nc):
    lo, hi = box
    return (nc, hi[2] - lo[2] + 1, hi[1] - lo[1] + 1, hi[0] - lo[0] + 1)


def centres(box, dx):
    """Cell-centre coordinates (z, y, x arrays) of a box."""
    lo, hi = box
    z, y, x = np.meshgrid(*[dx * (np.arange(lo[d], hi[d] + 1) + 0.5) for d in (2, 1, 0)],
                          indexing="ij")
    return x, y, z


def linear_data(box, nc, dx, coef):
    """u_c = a_c + b_c . x at the cell centres of box; coef = (a, b)."""
    a, b = coef
    x, y, z = centres(box, dx)
    return np.stack([a[c] + b[c][0] * x + b[c][1] * y + b[c][2] * z for c in range(nc)])


def rough_data(box, nc, seed, dx):
    """Smooth waves plus cell noise, so that the interpolation limiter acts
    in some cells and not in others."""
    rng = np.random.default_rng(seed)
    x, y, z = centres(box, dx)
    out = []
    for c in range(nc):
        k = rng.uniform(1.0, 4.0, size=3)
        base = 1.0 + 0.5 * np.sin(2 * np.pi * (k[0] * x + k[1] * y + k[2] * z) + c)
        out.append(base + 0.05 * rng.standard_normal(x.shape))
    return np.stack(out)


def fine_fabs(boxes, nc, ng, dims, dx, seed=0):
    """Fine FABs: rough data on the valid cells, GHOST elsewhere."""
    fabs, los = [], []
    for i, b in enumerate(boxes):
        g = grown(b, ng, dims)
        a = np.full(fab_shape(g, nc), GHOST)
        v = rough_data(b, nc, seed + 101 * i, dx)
        lo = g[0]
        sl = tuple(slice(b[0][d] - lo[d], b[1][d] - lo[d] + 1) for d in (2, 1, 0))
        a[(slice(None),) + sl] = v
        fabs.append(a)
        los.append(lo)
    return fabs, los


def owner(boxes, cell, dom_n, periodic, dims):
    """(box index, wrapped cell) of the box holding a cell or its periodic
    image, or None."""
    w = list(cell)
    for d in range(dims):
        if periodic[d]:
            w[d] %= dom_n[d]
        elif not 0 <= w[d] < dom_n[d]:
            return None
    for j, (lo, hi) in enumerate(boxes):
        if all(lo[d] <= w[d] <= hi[d] for d in range(3)):
            return j, tuple(w)
    return None


def interp_cells(boxes, i, ng, r, dom_n, periodic, dims):
    """Fine ghost cells of box i the interpolation fills: inside the box
    grown by ng (clipped to the non-periodic domain), with a coarse parent
    under no coarsened fine box and no periodic image of one."""
    lo, hi = grown(boxes[i], ng, dims)
    cboxes = [(tuple(l[d] // r if d < dims else l[d] for d in range(3)),
               tuple(h[d] // r if d < dims else h[d] for d in range(3))) for l, h in boxes]
    cn = [dom_n[d] // r if d < dims else 1 for d in range(3)]
    out = set()
    rngs = [range(lo[d], hi[d] + 1) for d in range(3)]
    for k in rngs[2]:
        for j in rngs[1]:
            for a in rngs[0]:
                c = (a, j, k)
                if any(not periodic[d] and not 0 <= c[d] < dom_n[d] for d in range(dims)):
                    continue
                par = tuple(c[d] // r if d < dims else c[d] for d in range(3))
                if owner(cboxes, par, cn, periodic, dims) is None:
                    out.add(c)
    return out
