"""CPU checks of the adjoint's parts (no GPU):

* the hand-written HLLD reverse sweep (nmhd_amd/csrc/adjoint_hlld.h, built
  for the host in double precision) against central finite differences of
  its own forward solve, in all six fan regions (outer, single-star and
  double-star state on either side), with Bx = 0 included;
* its forward solve against the oracle's HLLD (the reference's HLLD
  restated in C, pinned to the reference's golden vectors);
* the plain PyTorch transcription of the step (the FP64 autograd reference
  of the GPU tests) against the oracle's step;
* the optimal (binomial) checkpointing schedule: its recomputation count
  equals Griewank's closed form t*l - beta(s+1, t-1), and a reverse sweep
  driven by it reverses every step exactly once, in order, from the right
  state, never holding more than the allowed number of states.
"""
import ctypes
import os
from math import comb

import numpy as np
import pytest
import torch

import pyoracle as po
import torch_mhd as T
from mhd_checks import group_rel_err
from nmhd_amd import adjoint as A, problems as P

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST_LIB = os.path.join(ROOT, "tests", "cpp", "bin", "libadj_hlld_host.so")
DP = ctypes.POINTER(ctypes.c_double)


def host():
    if not os.path.exists(HOST_LIB):
        pytest.fail("tests/cpp/bin/libadj_hlld_host.so is not built (run the build)")
    return ctypes.CDLL(HOST_LIB)


def fwd(L, x, g=5.0 / 3.0):
    n = x.shape[0]
    out = np.zeros((n, 7))
    reg = np.zeros(n, np.int32)
    L.adj_hlld_forward(ctypes.c_long(n), ctypes.c_double(g), np.ascontiguousarray(x).ctypes.data_as(DP),
                       out.ctypes.data_as(DP), reg.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
    return out, reg


def rev(L, x, a, g=5.0 / 3.0):
    n = x.shape[0]
    out = np.zeros((n, 15))
    L.adj_hlld_reverse(ctypes.c_long(n), ctypes.c_double(g), np.ascontiguousarray(x).ctypes.data_as(DP),
                       np.ascontiguousarray(a).ctypes.data_as(DP), out.ctypes.data_as(DP))
    return out


def face_states(n, seed):
    rng = np.random.default_rng(seed)
    x = np.zeros((n, 15))
    x[:, 0] = rng.uniform(-2, 2, n)
    for i, (lo, hi) in enumerate([(0.1, 10), (-3, 3), (-3, 3), (-3, 3), (0.01, 10),
                                  (-2, 2), (-2, 2)]):
        x[:, 1 + 2 * i] = rng.uniform(lo, hi, n)
        x[:, 2 + 2 * i] = rng.uniform(lo, hi, n)
    k = n // 5
    x[:k, 0] = 0.0
    x[k:2 * k, 3:5] += 15.0
    x[2 * k:3 * k, 3:5] -= 15.0
    return x, k


def test_hlld_reverse_matches_finite_differences():
    L = host()
    x, k = face_states(6000, 3)
    F, reg = fwd(L, x)
    counts = np.bincount(reg, minlength=6)
    assert (counts > 100).all(), counts  # every region and side exercised
    rng = np.random.default_rng(4)
    a = rng.normal(size=(x.shape[0], 7))
    g = rev(L, x, a)
    d = rng.normal(size=x.shape)
    d[:k, 0] = 0.0  # the Bx = 0 states keep Bx = 0 (sign(Bx) jumps there)
    eps = 1e-6
    Fp, rp = fwd(L, x + eps * d)
    Fm, rm = fwd(L, x - eps * d)
    fd = ((Fp - Fm) / (2 * eps) * a).sum(1)
    an = (g * d).sum(1)
    same = (rp == reg) & (rm == reg)
    assert same.mean() > 0.99
    # error relative to |grad| |d| (a plain relative error of the dot product
    # is meaningless where it cancels)
    err = np.abs(fd - an) / (np.linalg.norm(g, axis=1) * np.linalg.norm(d, axis=1))
    assert err[same].max() < 1e-7, err[same].max()


def test_hlld_reverse_full_jacobian_columns():
    """J^T e_k against the finite-difference Jacobian row k, per input."""
    L = host()
    x, _ = face_states(300, 9)
    F, reg = fwd(L, x)
    eps = 1e-6
    for k in range(7):
        a = np.zeros((x.shape[0], 7))
        a[:, k] = 1.0
        g = rev(L, x, a)
        for j in range(15):
            dx = np.zeros_like(x)
            dx[:, j] = eps
            if j == 0:
                dx[x[:, 0] == 0.0, 0] = 0.0
            Fp, rp = fwd(L, x + dx)
            Fm, rm = fwd(L, x - dx)
            ok = (rp == reg) & (rm == reg) & (dx[:, j] != 0)
            fd = (Fp[:, k] - Fm[:, k]) / (2 * eps)
            scale = np.abs(fd[ok]).max() + 1.0
            assert np.abs(fd[ok] - g[ok, j]).max() < 1e-5 * scale, (k, j)


def test_hlld_forward_matches_oracle():
    L = host()
    x, _ = face_states(4000, 5)
    F, _ = fwd(L, x)
    fo = po.hlld_batch(x.T.astype(np.float32), 5.0 / 3.0).T
    scale = np.abs(F).max(axis=0)
    assert (np.abs(fo - F) / scale).max() < 2e-6


@pytest.mark.parametrize("case", ["ot16", "blast", "ot_euler_pc"])
def test_torch_transcription_matches_oracle(case):
    if case == "ot16":
        cfg = P.config_c1(16)
        W = P.orszag_tang_noisy(cfg, seed=4)
    elif case == "blast":
        cfg = P.make_config((16, 8, 12), cfl=0.4)
        W = P.blast(cfg, (0.48, 0.6, 0.64))
    else:
        cfg = P.make_config((12, 12, 8), cfl=0.3, recon=1, integrator=1)
        W = P.orszag_tang_noisy(cfg, seed=6)
    o = po.Oracle(cfg, W)
    dt = o.cfl_dt(1e30)
    Wt = torch.tensor(W, dtype=torch.float64)
    assert T.cfl_dt(Wt, cfg) == pytest.approx(dt, rel=1e-6)
    for _ in range(2):
        dt = o.cfl_dt(1e30)
        o.step()
        Wt = T.step(Wt, float(dt), cfg)
    errs = group_rel_err(Wt.numpy(), o.get_state())
    assert errs.max() < 5e-6, errs


class MockStepper:
    """States are step indices; records every reverse call."""

    def __init__(self):
        self.fwd = 0
        self.calls = []
        self.live = 0
        self.max_live = 0

    def forward(self, W, dt):
        self.fwd += 1
        return W + 1

    def reverse(self, W, dt, lam):
        self.calls.append(W)
        return lam + [W]


def griewank(l, s):
    """Fewest forward steps to reverse l steps with s checkpoints (the first
    one holding the initial state): t*l - beta(s+1, t-1)."""
    def beta(a, b):
        return comb(a + b, a) if b >= 0 else 0
    t = 0
    while beta(s, t) < l:
        t += 1
    return t * l - beta(s + 1, t - 1)


@pytest.mark.parametrize("l", [1, 2, 3, 7, 16, 33, 64])
@pytest.mark.parametrize("s", [0, 1, 2, 3, 6])
def test_revolve_schedule(l, s):
    assert A.revolve_cost(l, s) == griewank(l, s + 1)
    m = MockStepper()
    counter = [0]
    lam = A.reverse_segment(m, 0, 0, l, [0.0] * l, [], s, counter)
    assert lam == list(range(l - 1, -1, -1))  # each step once, last first, right state
    assert counter[0] == A.revolve_cost(l, s)


def fd_vs_autograd(cfg, W, nsteps, eps, ndir=3):
    """max over smooth directions of |FD - AD| / |AD| for the magnetic energy
    after nsteps steps of the FP64 transcription (dt frozen)."""
    Wt = torch.tensor(W, dtype=torch.float64)
    dts = []
    X = Wt
    for _ in range(nsteps):
        dt = T.cfl_dt(X, cfg)
        dts.append(dt)
        X = T.step(X, dt, cfg)
    W0 = Wt.clone().requires_grad_(True)
    X = W0
    for dt in dts:
        X = T.step(X, dt, cfg)
    T.magnetic_loss(X).backward()

    def loss(Wi):
        X = Wi
        for dt in dts:
            X = T.step(X, dt, cfg)
        return float(T.magnetic_loss(X))

    rel = []
    for i in range(ndir):
        D = P.smooth_direction(cfg, 1006, i)
        D[0] *= W[0].mean()
        Dt = torch.tensor(D)
        fd = (loss(Wt + eps * Dt) - loss(Wt - eps * Dt)) / (2 * eps)
        ad = float((W0.grad * Dt).sum())
        rel.append(abs(fd - ad) / abs(ad))
    return max(rel)


def test_transcription_differentiable_regime():
    """Where the discrete step is differentiable along the directions (a bulk
    flow fixes every CT upwind side; piecewise-constant reconstruction has no
    limiter kinks) central differences converge to the autograd derivative
    of the transcription; on the Orszag-Tang data itself they do not (the
    upwind choice jumps where a four-cell velocity sum changes sign), which
    is why the GPU finite-difference checks use the first kind of data."""
    cfg = P.make_config((12, 12, 12), cfl=0.4, recon=1)
    W = P.adjoint_case(cfg)
    Wb = W.copy()
    Wb[1:4] += 1.5
    e4 = fd_vs_autograd(cfg, Wb, 6, 1e-4)
    e5 = fd_vs_autograd(cfg, Wb, 6, 1e-5)
    assert e5 < 1e-3 and e5 < e4 / 5  # converges (first order: the fan's kinks)
    assert fd_vs_autograd(cfg, W, 6, 1e-5) > 1e-1
