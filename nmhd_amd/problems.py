"""Grid configurations and initial conditions for the benchmark workloads.

A configuration is a plain dict describing one box the way the reference's
compile-time constants do (library/simulation_config_default.h:25-122,
application/kepler_disk_demo/simulation_config.hrb:15-74), with every
derived float computed in FP32 exactly as those constant expressions are:

  kExtent  = kGlobalExtent / Real(kGlobalSize) * Real(kSize)   (:79-81)
  kD       = kExtent / kSize                                    (:101-103)
  kLocalExtent = kGlobalExtent / Real(kGlobalSize) * Real(kLocalSize) (:87-89)

Array layout: x fastest, shape (nz, ny, nx) per field, 8 fields in the order
rho, vx, vy, vz, p, Bx, By, Bz; face B_d[i] on the lower d-face of cell i.

Shearing box orientation follows the reference (application/
kepler_disk_demo): the radial axis is Y (ghost layers along Y), azimuth is X,
vertical is Z.

The counter-based RNG is splitmix64 of (seed ^ (idx*8 + component)), top 24
bits -> [0, 1); identical in every language that needs it.
"""
import math

import numpy as np

F32 = np.float32
NFIELD = 8
FIELDS = ("rho", "vx", "vy", "vz", "p", "bx", "by", "bz")


def make_config(global_size, margin=(0, 0, 0), global_extent=(1.0, 1.0, 1.0),
                gamma=5.0 / 3.0, cfl=0.4, source=0, omega=1.0, qshear=1.5,
                shear_y=0, recon=0, integrator=0):
    gs = [int(v) for v in global_size]
    m = [int(v) for v in margin]
    if m[0] != 0:
        raise ValueError("the reference never exchanges X ghosts; margin[0] must be 0")
    n = [gs[i] + 2 * m[i] for i in range(3)]
    ext, d, lext = [], [], []
    for i in range(3):
        cell = F32(F32(global_extent[i]) / F32(gs[i]))
        e = F32(cell * F32(n[i]))
        ext.append(float(e))
        d.append(float(F32(e / F32(n[i]))))
        lext.append(float(F32(cell * F32(gs[i]))))
    return {
        "global_size": gs, "m": m, "n": n,
        "global_extent": [float(F32(v)) for v in global_extent],
        "ext": ext, "d": d, "local_ext": lext,
        "gamma": float(F32(gamma)), "cfl": float(F32(cfl)),
        "source": int(source), "omega": float(F32(omega)),
        "qshear": float(F32(qshear)), "shear_y": int(shear_y),
        "recon": int(recon), "integrator": int(integrator),
    }


def ncell(cfg):
    n = cfg["n"]
    return n[0] * n[1] * n[2]


def interior_cells(cfg):
    g = cfg["global_size"]
    return g[0] * g[1] * g[2]


def empty_state(cfg):
    n = cfg["n"]
    return np.zeros((NFIELD, n[2], n[1], n[0]), dtype=F32)


# ---------------------------------------------------------------- RNG

_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def splitmix64(x):
    """splitmix64 finaliser of a uint64 array (wrapping arithmetic)."""
    x = np.asarray(x, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = x + np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def uniform01(seed, idx, component):
    """u = (splitmix64(seed ^ (idx*8 + component)) >> 40) * 2^-24."""
    idx = np.asarray(idx, dtype=np.uint64)
    with np.errstate(over="ignore"):
        key = np.uint64(seed) ^ (idx * np.uint64(8) + np.uint64(component))
    return (splitmix64(key) >> np.uint64(40)).astype(np.float64) * (2.0 ** -24)


# ------------------------------------------------------ initial conditions

def cell_centres(cfg, axis, corner=0.0):
    """Cell-centre coordinates along an axis of the array (ghosts included),
    corner + d*(i + 0.5), evaluated in float64."""
    n = cfg["n"][axis]
    d = cfg["d"][axis]
    return corner + d * (np.arange(n) + 0.5)


def orszag_tang(cfg, b0=None, amp_vz=0.01):
    """3D Orszag-Tang vortex extruded in z on [0,1)^3 (periodic, no ghosts).

    rho = 25/(36 pi), p = 5/(12 pi), v = (-sin 2pi y, sin 2pi x, amp_vz sin 2pi z),
    face B from A_z = B0 (cos 4pi x/(4pi) + cos 2pi y/(2pi)) by exact
    differences at the cell corners, so the discrete div B is zero.
    """
    if b0 is None:
        b0 = 1.0 / math.sqrt(4.0 * math.pi)
    n = cfg["n"]
    dx, dy, dz = (float(v) for v in cfg["d"])
    W = empty_state(cfg)
    x = (np.arange(n[0]) + 0.5) * dx
    y = (np.arange(n[1]) + 0.5) * dy
    z = (np.arange(n[2]) + 0.5) * dz
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    W[0] = 25.0 / (36.0 * math.pi)
    W[4] = 5.0 / (12.0 * math.pi)
    W[1] = -np.sin(2 * math.pi * Y)
    W[2] = np.sin(2 * math.pi * X)
    W[3] = amp_vz * np.sin(2 * math.pi * Z)

    def Az(xc, yc):
        return b0 * (np.cos(4 * math.pi * xc) / (4 * math.pi) +
                     np.cos(2 * math.pi * yc) / (2 * math.pi))

    xf = np.arange(n[0]) * dx  # lower x-faces / corners
    yf = np.arange(n[1]) * dy
    Zf, Yf, Xf = np.meshgrid(z, yf, xf, indexing="ij")
    # B_x on the lower x-face (x = i dx, y = (j+1/2) dy): dAz/dy
    W[5] = (Az(Xf, Yf + dy) - Az(Xf, Yf)) / dy
    # B_y on the lower y-face (x = (i+1/2) dx, y = j dy): -dAz/dx
    W[6] = -(Az(Xf + dx, Yf) - Az(Xf, Yf)) / dx
    W[7] = 0.0
    return W


def orszag_tang_noisy(cfg, seed=7, amp=1e-3, b0=None):
    """Orszag-Tang with seeded relative noise of size amp on rho, v and p from
    the counter RNG.  The noise breaks the vortex's exact mirror symmetries,
    under which sums of upwind velocities vanish exactly and the CT upwind
    choice becomes a tie decided by round-off."""
    W = orszag_tang(cfg, b0=b0)
    n = cfg["n"]
    idx = np.arange(n[0] * n[1] * n[2], dtype=np.uint64)
    for comp in (0, 1, 2, 3, 4):
        u = (2.0 * uniform01(seed, idx, comp) - 1.0).reshape(n[2], n[1], n[0])
        if comp in (1, 2, 3):
            W[comp] = W[comp] + amp * u
        else:
            W[comp] = W[comp] * (1.0 + amp * u)
    return W


def shearing_box_mri(cfg, seed=1002, beta=400.0, amp=0.01, rho0=1.0, p0=1.0):
    """MRI shearing box: rho, p uniform, vertical B with plasma beta,
    v_az = -q Omega x_rad + delta, delta ~ U[-amp, amp] per component
    from the counter RNG.  Radial = Y (with ghosts), azimuth = X, vertical Z.
    Ghost cells hold the same formula; the ghost fill overwrites them."""
    n = cfg["n"]
    W = empty_state(cfg)
    corner_y = -cfg["ext"][1] / 2.0
    yloc = cell_centres(cfg, 1, corner_y)  # radial position of every row
    idx = np.arange(n[0] * n[1] * n[2], dtype=np.uint64)
    W[0] = rho0
    W[4] = p0
    shear = (-cfg["qshear"] * cfg["omega"] * yloc)[None, :, None]
    for comp in range(3):
        dv = (2.0 * uniform01(seed, idx, comp) - 1.0) * amp
        W[1 + comp] = dv.reshape(n[2], n[1], n[0])
    W[1] += shear
    W[7] = math.sqrt(2.0 * p0 / beta)
    return W


def config_c1(n=32):
    """Orszag-Tang, n^3 periodic unit cube, gamma 5/3, CFL 0.4."""
    return make_config((n, n, n), cfl=0.4)


def config_c2(n=(256, 256, 256), cfl=0.4):
    """Shearing box n^3 with 4 radial (Y) ghost layers per side."""
    return make_config(n, margin=(0, 4, 0), global_extent=(1.0, 1.0, 1.0),
                       cfl=cfl, source=1, omega=1.0, qshear=1.5, shear_y=1)


# ------------------------------------------------------------ box batches

BATCH_EDGES = (16, 24, 32, 48, 64, 96)


def shock_tube(axis=0, n=512, recon=0, integrator=0):
    """The reference's MHD shock-tube demo (application/mhd_shock_demo:
    simulation_config.hrb, mhd_solver.curb:9-27): n cells along `axis` (the
    normal, extent 2, corner -1), one cell across, periodic, gamma 5/3, CFL
    0.5; rho, p = 1, 1 for x < 0 and 0.125, 0.1 beyond; B_normal = 0.75,
    B along the next axis +1 / -1, along the previous axis 0; v = 0 (the
    Brio-Wu states).  The demo runs to t = 0.1.  Returns (cfg, W)."""
    size = [1, 1, 1]
    size[axis] = n
    extent = [1.0, 1.0, 1.0]
    extent[axis] = 2.0
    cfg = make_config(size, global_extent=extent, cfl=0.5, recon=recon, integrator=integrator)
    W = empty_state(cfg)
    x = -1.0 + (np.arange(n, dtype=np.float64) + 0.5) * (2.0 / n)  # cell centres
    left = x < 0
    shape = [1, 1, 1]
    shape[2 - axis] = n  # arrays are (z, y, x)
    t1 = (axis + 1) % 3
    W[0] = np.where(left, 1.0, 0.125).reshape(shape)
    W[4] = np.where(left, 1.0, 0.1).reshape(shape)
    W[5 + axis] = 0.75
    W[5 + t1] = np.where(left, 1.0, -1.0).reshape(shape)
    return cfg, W


def batch_shapes(seed=1004, nbox=64, edges=BATCH_EDGES):
    """Edge lengths (nx, ny, nz) of each box, i.i.d. uniform over `edges`
    per axis from the counter RNG (component = axis, idx = box)."""
    out = []
    for b in range(nbox):
        e = []
        for axis in range(3):
            u = float(uniform01(seed, b, axis))
            e.append(int(edges[min(int(u * len(edges)), len(edges) - 1)]))
        out.append(tuple(e))
    return out


def blast(cfg, direction, p_in=10.0, p_out=0.1, radius=0.1, bmag=1.0 / math.sqrt(2.0)):
    """Blast wave in the periodic unit cube: rho = 1, v = 0, p = p_in within
    `radius` of the centre (cell centres) and p_out outside, uniform B of
    magnitude bmag along the unit vector `direction` (face values equal the
    components, so div B = 0)."""
    n = cfg["n"]
    W = empty_state(cfg)
    x = cell_centres(cfg, 0)
    y = cell_centres(cfg, 1)
    z = cell_centres(cfg, 2)
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    r2 = (X - 0.5) ** 2 + (Y - 0.5) ** 2 + (Z - 0.5) ** 2
    W[0] = 1.0
    W[4] = np.where(r2 < radius * radius, p_in, p_out)
    for d in range(3):
        W[5 + d] = bmag * direction[d]
    assert W.shape == (NFIELD, n[2], n[1], n[0])
    return W


def batch_boxes(seed=1004, nbox=64, edges=BATCH_EDGES):
    """The multi-box workload: box b has the edges of batch_shapes; even
    boxes hold an Orszag-Tang vortex with B0 ~ U[0.5, 1.5]/sqrt(4 pi), odd
    boxes a blast wave with |B| = 1/sqrt(2) along an isotropic random
    direction.  Periodic unit cubes, gamma 5/3, CFL 0.4.
    Returns a list of (cfg, W)."""
    out = []
    for b, shape in enumerate(batch_shapes(seed, nbox, edges)):
        cfg = make_config(shape, cfl=0.4)
        if b % 2 == 0:
            b0 = (0.5 + float(uniform01(seed, b, 3))) / math.sqrt(4.0 * math.pi)
            W = orszag_tang(cfg, b0=b0)
        else:
            cos_t = 2.0 * float(uniform01(seed, b, 4)) - 1.0
            phi = 2.0 * math.pi * float(uniform01(seed, b, 5))
            sin_t = math.sqrt(max(0.0, 1.0 - cos_t * cos_t))
            W = blast(cfg, (sin_t * math.cos(phi), sin_t * math.sin(phi), cos_t))
        out.append((cfg, W))
    return out


# ---------------------------------------------------------- adjoint case

def fourier_modes(seed, count, kmax=4, stream=0):
    """`count` integer wave vectors k (0 < |k| <= kmax) and phases from the
    counter RNG (idx = stream*1000 + mode)."""
    modes = []
    for m in range(count):
        j = 0
        while True:
            idx = stream * 1000 + m
            u = [float(uniform01(seed, idx * 64 + j, c)) for c in range(4)]
            k = tuple(int(math.floor(v * (2 * kmax + 1))) - kmax for v in u[:3])
            j += 1
            n2 = k[0] ** 2 + k[1] ** 2 + k[2] ** 2
            if 0 < n2 <= kmax * kmax:
                break
        modes.append((k, 2.0 * math.pi * u[3]))
    return modes


def mode_sum(cfg, modes):
    """sum_m cos(2 pi k_m . x + phase_m) at the cell centres."""
    n = cfg["n"]
    x = cell_centres(cfg, 0)
    y = cell_centres(cfg, 1)
    z = cell_centres(cfg, 2)
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    out = np.zeros((n[2], n[1], n[0]))
    for k, ph in modes:
        out += np.cos(2.0 * math.pi * (k[0] * X + k[1] * Y + k[2] * Z) + ph)
    return out


def adjoint_case(cfg, seed=1005, amp=1e-2, nmodes=8, vamp=1e-3):
    """Orszag-Tang vortex with a smooth density perturbation: rho times
    (1 + amp * sum of `nmodes` Fourier modes with |k| <= 4), plus smooth
    velocity modes of amplitude vamp.  The velocity modes break the vortex's
    exact mirror symmetry: with it, the four-cell velocity sums that choose
    the CT upwind side are exactly zero on whole rows of edges, where the
    discrete step is not differentiable (the choice jumps under any
    perturbation), so finite differences there measure the jump."""
    W = orszag_tang(cfg)
    W[0] = W[0] * (1.0 + amp * mode_sum(cfg, fourier_modes(seed, nmodes)))
    if vamp:
        for c in range(3):
            W[1 + c] = W[1 + c] + vamp * mode_sum(cfg, fourier_modes(seed, 2, stream=1 + c))
    return W


def smooth_direction(cfg, seed, index, fields=(0, 1, 2, 3), nmodes=4):
    """A smooth random perturbation direction of the initial state: on each
    listed field a sum of `nmodes` Fourier modes (|k| <= 4) with random
    signs, scaled to unit RMS; zero on the other fields."""
    n = cfg["n"]
    D = np.zeros((NFIELD, n[2], n[1], n[0]))
    for f in fields:
        s = mode_sum(cfg, fourier_modes(seed, nmodes, stream=index * 8 + f))
        D[f] = s / max(np.sqrt((s * s).mean()), 1e-30)
    return D


# ------------------------------------------------- large shearing box (C3)

def _splitmix64_torch(x):
    """splitmix64 on int64 tensors (two's-complement wrap, logical shifts)."""
    import torch

    def shr(v, k):
        return (v >> k) & ((1 << (64 - k)) - 1)

    def c(u):  # uint64 constant as a wrapped int64
        return u - (1 << 64) if u >= (1 << 63) else u

    z = x + c(0x9E3779B97F4A7C15)
    z = (z ^ shr(z, 30)) * c(0xBF58476D1CE4E5B9)
    z = (z ^ shr(z, 27)) * c(0x94D049BB133111EB)
    return z ^ shr(z, 31)


def shearing_box_noise_device(cfg, seed=1003, beta=1000.0, amp=1e-3, k0=8.0, width=2.0,
                              rho0=1.0, p0=1.0, device="cuda"):
    """The large MRI box on the device (torch tensors, (8, nz, ny, nx) fp32):
    rho0 (1 + amp * n), n = white noise u in [-1, 1) from the counter RNG
    (component 0), filtered in Fourier space by exp(-(|k|-k0)^2 / (2 width^2))
    (|k| in units of the box's fundamental) and scaled to unit RMS; p = p0,
    v_az = -q Omega x_rad, B = (0, 0, sqrt(2 p0 / beta)).  The noise lives on
    the interior rows; ghost rows are filled by the first ghost fill."""
    import torch
    n = cfg["n"]
    m = cfg["m"]
    gx, gy, gz = cfg["global_size"]
    W = torch.zeros((NFIELD, n[2], n[1], n[0]), dtype=torch.float32, device=device)
    # white noise over the interior, one plane at a time (int64 counters)
    noise = torch.empty((gz, gy, gx), dtype=torch.float32, device=device)
    plane = gx * gy
    base = torch.arange(plane, dtype=torch.int64, device=device)
    for z in range(gz):
        idx = base + z * plane
        key = (idx * 8 + 0) ^ int(seed)
        u = ((_splitmix64_torch(key) >> 40) & ((1 << 24) - 1)).to(torch.float64) * 2.0 ** -24
        noise[z] = (2.0 * u - 1.0).to(torch.float32).view(gy, gx)
    F = torch.fft.rfftn(noise)
    kz = torch.fft.fftfreq(gz, d=1.0 / gz, device=device).view(-1, 1, 1)
    ky = torch.fft.fftfreq(gy, d=1.0 / gy, device=device).view(1, -1, 1)
    kx = torch.fft.rfftfreq(gx, d=1.0 / gx, device=device).view(1, 1, -1)
    kk = torch.sqrt(kx * kx + ky * ky + kz * kz)
    F *= torch.exp(-(kk - k0) ** 2 / (2.0 * width * width))
    del kk
    filt = torch.fft.irfftn(F, s=(gz, gy, gx))
    del F
    filt /= torch.sqrt((filt * filt).mean())
    W[0, :, m[1]:m[1] + gy, :] = rho0 * (1.0 + amp * filt)
    W[0, :, :m[1], :] = rho0
    W[0, :, m[1] + gy:, :] = rho0
    del filt, noise
    W[4] = p0
    corner_y = -cfg["ext"][1] / 2.0
    yloc = torch.tensor(cell_centres(cfg, 1, corner_y), dtype=torch.float64, device=device)
    W[1] = (-cfg["qshear"] * cfg["omega"] * yloc).to(torch.float32).view(1, -1, 1)
    W[7] = math.sqrt(2.0 * p0 / beta)
    return W
