"""Shared helpers for the tests: synthetic states, periodic ghost fill,
comparison metrics and a GPU runner over the C ABI."""
import ctypes as C

import numpy as np

import mfmhd

S4PI = np.sqrt(4.0 * np.pi)


def grid_shape(nc, valid_lo, valid_hi, ng, dims=3):
    lo = [valid_lo[d] - (ng if d < dims else 0) for d in range(3)]
    hi = [valid_hi[d] + (ng if d < dims else 0) for d in range(3)]
    shape = (nc, hi[2] - lo[2] + 1, hi[1] - lo[1] + 1, hi[0] - lo[0] + 1)
    return tuple(lo), tuple(hi), shape


def smooth_state(p, shape, lo, seed=0, amp=0.3, b0=1.0):
    """A smooth, positive multi-fluid state (conserved variables)."""
    rng = np.random.default_rng(seed)
    nc, nz, ny, nx = shape
    z, y, x = np.meshgrid(np.arange(nz) + lo[2], np.arange(ny) + lo[1],
                          np.arange(nx) + lo[0], indexing="ij")
    g = p.gamma

    def wave():
        kx, ky, kz = rng.integers(1, 3, size=3)
        ph = rng.uniform(0, 2 * np.pi)
        return np.sin(2 * np.pi * (kx * x / 16.0 + ky * y / 16.0 + kz * z / 16.0) + ph)

    U = np.zeros(shape)
    rho = 1.0 + amp * wave()
    vx, vy, vz = (0.5 * wave() for _ in range(3))
    pr = 1.0 + amp * wave()
    bx, by, bz = (b0 * (0.7 + 0.3 * wave()) for _ in range(3))
    U[0] = rho
    U[1], U[2], U[3] = rho * vx, rho * vy, rho * vz
    U[4] = pr / (g - 1) + 0.5 * rho * (vx * vx + vy * vy + vz * vz) + \
        (bx * bx + by * by + bz * bz) / (8 * np.pi)
    U[5], U[6], U[7] = bx, by, bz
    c = 8
    if p.dedner:
        U[8] = 0.05 * wave()
        c = 9
    for f in range(1, p.nfluids):
        r = 0.5 + 0.2 * wave()
        ux, uy, uz = (0.4 * wave() for _ in range(3))
        pn = 0.3 + 0.1 * wave()
        U[c] = r
        U[c + 1], U[c + 2], U[c + 3] = r * ux, r * uy, r * uz
        U[c + 4] = pn / (g - 1) + 0.5 * r * (ux * ux + uy * uy + uz * uz)
        c += 5
    return U


def periodic_fill(U, ng, dims=3):
    """One-box periodic ghost fill on a numpy FAB array (nc, nz, ny, nx):
    every ghost cell takes the valid cell congruent to it modulo the valid
    extent (which may be shorter than ng)."""
    V = U
    for d in range(dims):
        ax = 3 - d
        m = V.shape[ax]
        n = m - 2 * ng
        src = ng + (np.arange(m) - ng) % n
        V[...] = np.take(V, src, axis=ax)
    return V


def rel_err_per_comp(a, b):
    """max |a-b| / max |b| for every component (array leading axis)."""
    out = []
    for c in range(b.shape[0]):
        den = np.max(np.abs(b[c]))
        num = np.max(np.abs(a[c] - b[c]))
        out.append(num / den if den > 0 else num)
    return np.array(out)


def l1_rel_per_comp(a, b):
    out = []
    for c in range(b.shape[0]):
        den = np.sum(np.abs(b[c]))
        num = np.sum(np.abs(a[c] - b[c]))
        out.append(num / den if den > 0 else num)
    return np.array(out)


def vector_groups(nf, dedner=False):
    """Component triples that form vectors: plasma momentum, B, and each
    neutral's momentum."""
    g = [(1, 2, 3), (5, 6, 7)]
    base = 9 if dedner else 8
    for f in range(nf - 1):
        b = base + 5 * f
        g.append((b + 1, b + 2, b + 3))
    return g


def l1_rel_vector_scaled(a, b, groups):
    """Relative L1 error per component; a vector component is scaled by the
    L1 norm of its vector's magnitude (a component that is ~0 everywhere,
    like B_z of a field in the x-y plane, is not divided by ~0)."""
    out = l1_rel_per_comp(a, b)
    for grp in groups:
        mag = np.sqrt(sum(b[c] ** 2 for c in grp))
        den = np.sum(mag)
        for c in grp:
            out[c] = np.sum(np.abs(a[c] - b[c])) / den if den > 0 else 0.0
    return out


def valid_view(U, lo, vlo, vhi):
    return U[:, vlo[2] - lo[2]:vhi[2] - lo[2] + 1,
             vlo[1] - lo[1]:vhi[1] - lo[1] + 1,
             vlo[0] - lo[0]:vhi[0] - lo[0] + 1]


class GpuPatch:
    """A device FAB with its mfmhd.Fab view."""

    def __init__(self, U, lo):
        U = np.ascontiguousarray(U)
        self.t = mfmhd.DeviceArray(U.shape, data=U)
        nc, nz, ny, nx = U.shape
        self.lo = tuple(lo)
        self.hi = (lo[0] + nx - 1, lo[1] + ny - 1, lo[2] + nz - 1)
        self.fab = mfmhd.fab(self.t.ptr, self.lo, self.hi, nc)

    def numpy(self):
        mfmhd.synchronize()
        return self.t.numpy()


def gpu_update(ctx, U, lo, vlo, vhi, dt, dx, time=0.0, want_F=False,
               want_divB=False, noalloc=False):
    """Run mfm_patch_update_dev (or, with noalloc, the allocation-free
    mfm_patch_update_noalloc_dev) on a copy of U; returns a dict like
    oracle.patch_update."""
    g = GpuPatch(U, lo)
    nc = U.shape[0]
    dims = ctx.params.spacedim
    Ft, Ffabs = [], None
    if want_F:
        Ffabs = []
        for d in range(3):
            flo, fhi = list(vlo), list(vhi)
            if d < dims:
                fhi[d] += 1
            t = mfmhd.DeviceArray((nc, fhi[2] - flo[2] + 1, fhi[1] - flo[1] + 1,
                                   fhi[0] - flo[0] + 1), zero=True)
            Ft.append(t)
            Ffabs.append(mfmhd.fab(t.ptr, flo, fhi, nc))
    dfab = None
    dt_t = None
    if want_divB:
        dt_t = mfmhd.DeviceArray((1, vhi[2] - vlo[2] + 1, vhi[1] - vlo[1] + 1,
                                  vhi[0] - vlo[0] + 1), zero=True)
        dfab = mfmhd.fab(dt_t.ptr, vlo, vhi, 1)
    rc, nd, cell = ctx.update_state(g.fab, mfmhd.Box.of(vlo, vhi), dt, dx, time,
                                    F=Ffabs, divB=dfab, check=False, noalloc=noalloc)
    mfmhd.synchronize()
    return {"rc": rc, "U": g.numpy(), "dt": nd, "cell": cell,
            "F": [t.numpy() for t in Ft[:dims]] if want_F else None,
            "divB": dt_t.numpy() if want_divB else None}
