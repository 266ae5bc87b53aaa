"""Plain PyTorch transcription of the periodic step (test infrastructure).

A direct, vectorised restatement of the step that the device kernels and
the oracle compute, for periodic boxes without source terms: minmod MUSCL
reconstruction of (rho, v_n, v_t1, v_t2, p, Bbar_t1, Bbar_t2) at every lower
face, the HLLD flux (Miyoshi & Kusano 2005 with the reference's choices:
fast-speed outer waves with the sqrt(2) factor, 1e-10 guard, signBx = 0 at
Bx = 0; library/hlld.inl:7-284), the flux divergence, the upwinded CT edge
EMFs (library/mhd_solver.inl:819-897) and the midpoint2 update with the
conserved base recomputed from W^n (mhd_solver.inl:990-1071).

It runs in any dtype; the adjoint tests evaluate it in float64 and
differentiate it with torch.autograd.  Arrays are (8, nz, ny, nx): rho, vx,
vy, vz, p, Bx, By, Bz, face B on the lower faces.
"""
import torch

AX = (2, 1, 0)  # tensor dim of axis x, y, z in a (nz, ny, nx) field


def sh(a, s, d):
    """a shifted so that out[c] = a[c + s] along axis d (periodic)."""
    return torch.roll(a, shifts=-s, dims=AX[d])


def minmod(a, b):
    return torch.where(a * b <= 0, torch.zeros_like(a), torch.where(a.abs() < b.abs(), a, b))


def recon(q, d, mode=0):
    """Left/right states at the lower d-face of every cell (cells c-2..c+1)."""
    x0, x1, x2, x3 = sh(q, -2, d), sh(q, -1, d), q, sh(q, 1, d)
    if mode == 1:
        return x1, x2
    d12 = x2 - x1
    return x1 + 0.5 * minmod(x1 - x0, d12), x2 - 0.5 * minmod(d12, x3 - x2)


def hlld(gamma, bx, rL, rR, unL, unR, u1L, u1R, u2L, u2R, pL, pR, b1L, b1R, b2L, b2R):
    ig = 1.0 / (gamma - 1.0)
    tiny = 1e-10
    bx2 = bx * bx
    abx = bx.abs()
    sgn = torch.sign(bx)
    B2L = bx2 + b1L * b1L + b2L * b2L
    B2R = bx2 + b1R * b1R + b2R * b2R
    eL = pL * ig + 0.5 * rL * (unL * unL + u1L * u1L + u2L * u2L) + 0.5 * B2L
    eR = pR * ig + 0.5 * rR * (unR * unR + u1R * u1R + u2R * u2R) + 0.5 * B2R
    ptL = pL + 0.5 * B2L
    ptR = pR + 0.5 * B2R
    gpL, gpR = gamma * pL, gamma * pR
    gpbL, gpbR = gpL + B2L, gpR + B2R
    cfL = torch.sqrt((gpbL + torch.sqrt(torch.clamp(gpbL * gpbL - 4 * gpL * bx2, min=0))) / (2 * rL))
    cfR = torch.sqrt((gpbR + torch.sqrt(torch.clamp(gpbR * gpbR - 4 * gpR * bx2, min=0))) / (2 * rR))
    cmax = 1.41421356 * torch.maximum(cfL, cfR)
    SL = torch.minimum(unL, unR) - cmax
    SR = torch.maximum(unL, unR) + cmax
    sLu, sRu = SL - unL, SR - unR
    den = sRu * rR - sLu * rL
    SM = (sRu * rR * unR - sLu * rL * unL - ptR + ptL) / den
    pts = (sRu * rR * ptL - sLu * rL * ptR + rL * rR * sRu * sLu * (unR - unL)) / den
    sLm, sRm = SL - SM, SR - SM
    rLs = rL * sLu / sLm
    rRs = rR * sRu / sRm
    dL = rL * sLu * sLm - bx2 + tiny
    dR = rR * sRu * sRm - bx2 + tiny
    kvL = bx * (SM - unL) / dL
    kvR = bx * (SM - unR) / dR
    kbL = (rL * sLu * sLu - bx2) / dL
    kbR = (rR * sRu * sRu - bx2) / dR
    u1Ls, u2Ls = u1L - b1L * kvL, u2L - b2L * kvL
    u1Rs, u2Rs = u1R - b1R * kvR, u2R - b2R * kvR
    b1Ls, b2Ls = b1L * kbL, b2L * kbL
    b1Rs, b2Rs = b1R * kbR, b2R * kbR
    vbL = unL * bx + u1L * b1L + u2L * b2L
    vbR = unR * bx + u1R * b1R + u2R * b2R
    vbLs = SM * bx + u1Ls * b1Ls + u2Ls * b2Ls
    vbRs = SM * bx + u1Rs * b1Rs + u2Rs * b2Rs
    eLs = (sLu * eL - ptL * unL + pts * SM + bx * (vbL - vbLs)) / sLm
    eRs = (sRu * eR - ptR * unR + pts * SM + bx * (vbR - vbRs)) / sRm
    srL, srR = torch.sqrt(rLs), torch.sqrt(rRs)
    SLs = SM - abx / srL
    SRs = SM + abx / srR
    f = 1.0 / (srL + srR)
    u1ss = f * (srL * u1Ls + srR * u1Rs + (b1Rs - b1Ls) * sgn)
    u2ss = f * (srL * u2Ls + srR * u2Rs + (b2Rs - b2Ls) * sgn)
    srLR = torch.sqrt(rLs * rRs)
    b1ss = f * (srL * b1Rs + srR * b1Ls + srLR * (u1Rs - u1Ls) * sgn)
    b2ss = f * (srL * b2Rs + srR * b2Ls + srLR * (u2Rs - u2Ls) * sgn)
    vbss = SM * bx + u1ss * b1ss + u2ss * b2ss
    eLss = eLs - srL * (vbLs - vbss) * sgn
    eRss = eRs + srR * (vbRs - vbss) * sgn

    left = SM >= 0

    def pick(a, b):
        return torch.where(left, a, b)

    r, un, u1, u2 = pick(rL, rR), pick(unL, unR), pick(u1L, u1R), pick(u2L, u2R)
    b1, b2, e, pt, vb = pick(b1L, b1R), pick(b2L, b2R), pick(eL, eR), pick(ptL, ptR), pick(vbL, vbR)
    S, Ss, rs = pick(SL, SR), pick(SLs, SRs), pick(rLs, rRs)
    u1s, u2s, b1s, b2s = pick(u1Ls, u1Rs), pick(u2Ls, u2Rs), pick(b1Ls, b1Rs), pick(b2Ls, b2Rs)
    es, ess = pick(eLs, eRs), pick(eLss, eRss)
    region = torch.where(left,
                         torch.where(SL > 0, 0, torch.where(SLs >= 0, 1, 2)),
                         torch.where(SR < 0, 0, torch.where(SRs < 0, 1, 2)))
    r1 = (region >= 1).to(bx.dtype)
    r2 = (region == 2).to(bx.dtype)
    Frho = r * un + r1 * S * (rs - r)
    Fmn = r * un * un + pt - bx2 + r1 * S * (rs * SM - r * un)
    Fm1 = r * u1 * un - bx * b1 + r1 * S * (rs * u1s - r * u1) + r2 * Ss * rs * (u1ss - u1s)
    Fm2 = r * u2 * un - bx * b2 + r1 * S * (rs * u2s - r * u2) + r2 * Ss * rs * (u2ss - u2s)
    Fe = (e + pt) * un - bx * vb + r1 * S * (es - e) + r2 * Ss * (ess - es)
    Fb1 = b1 * un - bx * u1 + r1 * S * (b1s - b1) + r2 * Ss * (b1ss - b1s)
    Fb2 = b2 * un - bx * u2 + r1 * S * (b2s - b2) + r2 * Ss * (b2ss - b2s)
    return Frho, Fmn, Fm1, Fm2, Fe, Fb1, Fb2


def upwind(s, f0, f1):
    return torch.where(s > 0, f0, torch.where(s < 0, f1, 0.5 * (f0 + f1)))


def rhs(W, cfg):
    """P(W): cell paces (rho, m_x, m_y, m_z, E) and face paces (B_x, B_y, B_z)."""
    gamma = cfg["gamma"]
    mode = cfg.get("recon", 0)
    inv = [1.0 / float(v) for v in cfg["d"]]
    rho, vx, vy, vz, p, Bx, By, Bz = W
    v = (vx, vy, vz)
    B = (Bx, By, Bz)
    Bbar = tuple(0.5 * (B[t] + sh(B[t], 1, t)) for t in range(3))
    P = [torch.zeros_like(rho) for _ in range(5)]
    FB = {}
    for d in range(3):
        d1, d2 = (d + 1) % 3, (d + 2) % 3
        q = (rho, v[d], v[d1], v[d2], p, Bbar[d1], Bbar[d2])
        LR = [recon(x, d, mode) for x in q]
        F = hlld(gamma, B[d], *[t for pair in LR for t in pair])
        comp = (0, 1 + d, 1 + d1, 1 + d2, 4)
        for k in range(5):
            P[comp[k]] = P[comp[k]] + (F[k] - sh(F[k], 1, d)) * inv[d]
        FB[(d, d1)] = F[5]  # d-face flux of B_d1
        FB[(d, d2)] = F[6]  # d-face flux of B_d2

    def edge_sum(vc, p_, q_, comp):
        # sum of the comp-velocity of the 4 cells around the edge on the lower
        # p_- and q_-faces, in the device kernels' order
        c, cq, cp = vc, sh(vc, -1, q_), sh(vc, -1, p_)
        cpq = sh(cp, -1, q_)
        return c + cq + cp + cpq if comp == q_ else c + cp + cq + cpq

    # edge EMFs: E[(sweep, edge_face_axis)] on the lower faces of the cell
    E = {}
    for (a, b_) in ((0, 1), (0, 2), (1, 2)):
        # edge on lower a- and b_-faces; a-sweep member: a-face flux of B_b_,
        # upwinded along b_ by v_b_; b_-sweep member: b_-face flux of B_a,
        # upwinded along a by v_a
        sa = edge_sum(v[b_], a, b_, b_)
        E[(a, b_)] = upwind(sa, sh(FB[(a, b_)], -1, b_), FB[(a, b_)])
        sb = edge_sum(v[a], a, b_, a)
        E[(b_, a)] = upwind(sb, sh(FB[(b_, a)], -1, a), FB[(b_, a)])
    idx = inv
    # face paces (the update kernels' formulas, lower faces)
    Ez_x, Ez_y = E[(0, 1)], E[(1, 0)]
    Ey_x, Ey_z = E[(0, 2)], E[(2, 0)]
    Ex_y, Ex_z = E[(1, 2)], E[(2, 1)]
    PBx = (0.5 * (sh(Ez_x, 1, 1) - Ez_x) * idx[1] + 0.5 * (sh(Ey_x, 1, 2) - Ey_x) * idx[2] -
           0.5 * (sh(Ez_y, 1, 1) - Ez_y) * idx[1] - 0.5 * (sh(Ey_z, 1, 2) - Ey_z) * idx[2])
    PBy = (-0.5 * (sh(Ez_x, 1, 0) - Ez_x) * idx[0] + 0.5 * (sh(Ex_y, 1, 2) - Ex_y) * idx[2] +
           0.5 * (sh(Ez_y, 1, 0) - Ez_y) * idx[0] - 0.5 * (sh(Ex_z, 1, 2) - Ex_z) * idx[2])
    PBz = (-0.5 * (sh(Ey_x, 1, 0) - Ey_x) * idx[0] - 0.5 * (sh(Ex_y, 1, 1) - Ex_y) * idx[1] +
           0.5 * (sh(Ey_z, 1, 0) - Ey_z) * idx[0] + 0.5 * (sh(Ex_z, 1, 1) - Ex_z) * idx[1])
    return P, (PBx, PBy, PBz)


def face_b2(B):
    return sum(B[t] * B[t] + sh(B[t], 1, t) * sh(B[t], 1, t) for t in range(3))


def update(Win, Wbase, h, cfg):
    gamma = cfg["gamma"]
    P, PB = rhs(Win, cfg)
    rho0, vx0, vy0, vz0, p0 = Wbase[0], Wbase[1], Wbase[2], Wbase[3], Wbase[4]
    B0 = (Wbase[5], Wbase[6], Wbase[7])
    e0 = p0 / (gamma - 1.0) + 0.5 * rho0 * (vx0 * vx0 + vy0 * vy0 + vz0 * vz0) + 0.25 * face_b2(B0)
    rho = rho0 + h * P[0]
    mx = rho0 * vx0 + h * P[1]
    my = rho0 * vy0 + h * P[2]
    mz = rho0 * vz0 + h * P[3]
    e = e0 + h * P[4]
    B = tuple(B0[t] + h * PB[t] for t in range(3))
    p = (e - 0.5 * (mx * mx + my * my + mz * mz) / rho - 0.25 * face_b2(B)) * (gamma - 1.0)
    return torch.stack([rho, mx / rho, my / rho, mz / rho, p, B[0], B[1], B[2]])


def step(W, dt, cfg):
    """One step with a given dt (midpoint2, or euler1 when integrator == 1)."""
    if cfg.get("integrator", 0) == 1:
        return update(W, W, dt, cfg)
    Ws = update(W, W, dt / 2, cfg)
    return update(Ws, W, dt, cfg)


def magnetic_loss(W):
    """Total magnetic energy 1/2 sum over faces of B^2."""
    return 0.5 * (W[5:8] ** 2).sum()


def cfl_dt(W, cfg, goal=1e30):
    """CFL dt (|v_d + c| with b^2 = sum_d max(face^2)), as the device computes it."""
    gamma = cfg["gamma"]
    B = (W[5], W[6], W[7])
    b2 = sum(torch.maximum(B[t] ** 2, sh(B[t], 1, t) ** 2) for t in range(3))
    c = torch.sqrt((gamma * W[4] + b2) / W[0])
    t = torch.full_like(W[0], goal)
    for d in range(3):
        t = torch.minimum(t, cfg["cfl"] * float(cfg["d"][d]) / (W[1 + d] + c).abs())
    return float(t.min())

