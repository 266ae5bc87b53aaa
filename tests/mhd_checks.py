"""Shared numerical checks for the test suite (numpy, float64 accumulation)."""
import numpy as np

FIELDS = ("rho", "vx", "vy", "vz", "p", "bx", "by", "bz")


def div_b(W, d):
    """Discrete divergence of the face field at every cell (periodic wrap)."""
    bx, by, bz = (W[i].astype(np.float64) for i in (5, 6, 7))
    return ((np.roll(bx, -1, 2) - bx) / d[0] + (np.roll(by, -1, 1) - by) / d[1] +
            (np.roll(bz, -1, 0) - bz) / d[2])


def totals(W, gamma, sl=(slice(None),) * 3):
    """Sums of rho, m_x, m_y, m_z, E (face-averaged B^2) and B_x, B_y, B_z."""
    W = W.astype(np.float64)
    rho, v, p, b = W[0], W[1:4], W[4], W[5:8]
    bb = 0.25 * (b[0] ** 2 + np.roll(b[0], -1, 2) ** 2 + b[1] ** 2 +
                 np.roll(b[1], -1, 1) ** 2 + b[2] ** 2 + np.roll(b[2], -1, 0) ** 2)
    e = p / (gamma - 1.0) + 0.5 * rho * (v ** 2).sum(0) + bb
    q = [rho, rho * v[0], rho * v[1], rho * v[2], e, b[0], b[1], b[2]]
    return np.array([x[sl].sum() for x in q])


def field_rel_err(a, b):
    """Per-field max |a-b| / max|b| (max|b| floored at 1e-30)."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    out = []
    for i in range(a.shape[0]):
        scale = max(np.abs(b[i]).max(), 1e-30)
        out.append(np.abs(a[i] - b[i]).max() / scale)
    return np.array(out)


def frac_within(a, b, tol):
    """Fraction of cells whose every field agrees within tol * max|field|."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    ok = np.ones(a.shape[1:], bool)
    for i in range(a.shape[0]):
        scale = max(np.abs(b[i]).max(), 1e-30)
        ok &= np.abs(a[i] - b[i]) <= tol * scale
    return ok.mean()


STATE_GROUPS = ([0], [1, 2, 3], [4], [5, 6, 7])  # rho | v | p | B
PACE_GROUPS = ([0], [1, 2, 3], [4], [5, 6, 7])   # rho | m | E | B


def group_rel_err(a, b, groups=STATE_GROUPS):
    """max |a-b| over each group of fields divided by the group's max |b|
    (a vector component is measured against the vector's magnitude)."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    out = []
    for g in groups:
        scale = max(np.abs(b[g]).max(), 1e-30)
        out.append(np.abs(a[g] - b[g]).max() / scale)
    return np.array(out)


def bitwise_equal(a, b):
    a = np.ascontiguousarray(a, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    return np.array_equal(a.view(np.uint32), b.view(np.uint32))


def perturb_ulps(W, seed=0, frac=0.5):
    """Move a random half of the values by one ulp (a probe of round-off
    sensitivity)."""
    rng = np.random.default_rng(seed)
    mask = rng.random(W.shape) < frac
    return np.where(mask, np.nextafter(W, np.float32(np.inf)), W).astype(np.float32)


def group_frac_within(a, b, tol, groups=STATE_GROUPS):
    """Fraction of cells where every group agrees within tol * group max."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    ok = np.ones(a.shape[1:], bool)
    for g in groups:
        scale = max(np.abs(b[g]).max(), 1e-30)
        ok &= (np.abs(a[g] - b[g]) <= tol * scale).all(axis=0)
    return ok.mean()


def group_stats(a, b, tol, groups=STATE_GROUPS, zchunk=16):
    """(group_rel_err(a, b), group_frac_within(a, b, tol)) evaluated slab by
    slab along z, so that arrays of billions of values need no float64
    copies of their own size."""
    nz = b.shape[1]
    slabs = [slice(z, min(z + zchunk, nz)) for z in range(0, nz, zchunk)]
    scale = []
    for g in groups:
        m = 0.0
        for f in g:
            for s in slabs:
                m = max(m, float(np.abs(b[f, s]).max()))
        scale.append(max(m, 1e-30))
    err = np.zeros(len(groups))
    ok_cells = 0
    for s in slabs:
        ok = None
        for gi, g in enumerate(groups):
            for f in g:
                d = np.abs(a[f, s].astype(np.float64) - b[f, s])
                err[gi] = max(err[gi], float(d.max()))
                within = d <= tol * scale[gi]
                ok = within if ok is None else ok & within
        ok_cells += int(ok.sum())
    return err / np.array(scale), ok_cells / float(np.prod(b.shape[1:]))
