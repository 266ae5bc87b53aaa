"""GPU checks of the hand-written adjoint (nmhd_adj_* through the torch
binding nmhd_amd.adjoint):

* the forward step with a frozen dt equals the production step bit for bit;
* one reverse step against the FP64 torch-autograd gradient of the plain
  PyTorch transcription of the step (tests/torch_mhd.py, itself checked
  against the oracle): per field group, the error (2-norm, relative) is at
  most twice that of FP32 torch autograd of the same transcription, i.e. the
  hand-written adjoint is as exact as automatic differentiation at the same
  precision (the pressure gradient carries the largest FP32 error, ~5e-3,
  from the cancellation in p = (gamma-1)(E - rho v^2/2 - B^2/2));
* the gradient of the magnetic energy after several steps (through
  torch.autograd with checkpointing) against the same FP64 reference;
* the stage-keeping forward/reverse pair equals the recomputing one bit for
  bit, and every checkpointing scheme gives the same gradient bit for bit;
* central finite differences (eps = 1e-3) along smooth random directions
  agree with the adjoint's directional derivatives per direction within
  1e-3 relative plus 4 sigma, sigma being the measured FP32 round-off noise
  of a central difference (nmhd_amd.adjoint.finite_difference_check), on
  data where the discrete step is differentiable along them: a bulk flow
  keeps every CT upwind choice on one side (that choice jumps where a
  four-cell velocity sum changes sign) and piecewise-constant reconstruction
  removes the minmod kinks, leaving only the kinks of the HLLD fan.  (On the
  Orszag-Tang data itself finite differences disagree with the exact
  derivative of the FP64 transcription just as much: the step is not
  differentiable there.)
"""
import numpy as np
import pytest
import torch

import torch_mhd as T
from nmhd_amd import adjoint as A, capi, problems as P

pytestmark = pytest.mark.gpu


def announce(msg):
    print("\n[case] " + msg, flush=True)


def cuda(W):
    return torch.tensor(np.ascontiguousarray(W), dtype=torch.float32, device="cuda")


def group_norm_err(got, ref):
    """max over field groups (rho | v | p | B) of |got - ref| / |ref| (2-norms)."""
    out = []
    for g in ([0], [1, 2, 3], [4], [5, 6, 7]):
        r = ref[g].double()
        out.append(float((got[g].double() - r).norm() / max(float(r.norm()), 1e-30)))
    return np.array(out)


def test_forward_equals_production_step(gpu_ready):
    announce("adjoint forward == production step (16^3 OT, 3 steps)")
    cfg = P.config_c1(16)
    W = P.adjoint_case(cfg)
    S = A.Stepper(cfg)
    g = capi.Context(cfg, path=capi.PATH_TILED)
    g.set_state(W)
    Wt = cuda(W)
    for _ in range(3):
        dt = S.cfl_dt(Wt)
        Wt = S.forward(Wt, dt)
        gdt, _ = g.step()
        assert gdt == dt
    torch.cuda.synchronize()
    assert np.array_equal(Wt.cpu().numpy().view(np.uint32), g.get_state().view(np.uint32))


@pytest.mark.parametrize("case", ["ot", "blast", "pc_euler"])
def test_reverse_step_vs_fp64_autograd(gpu_ready, case):
    announce("one reverse step vs FP64 autograd: " + case)
    if case == "ot":
        cfg = P.config_c1(16)
        W = P.adjoint_case(cfg)
    elif case == "blast":
        cfg = P.make_config((16, 12, 8), cfl=0.4)
        W = P.blast(cfg, (0.48, 0.6, 0.64))
        W[1:4] += 1e-3 * P.smooth_direction(cfg, 78, 0, fields=(1, 2, 3))[1:4]
    else:
        cfg = P.make_config((12, 16, 8), cfl=0.3, recon=1, integrator=1)
        W = P.adjoint_case(cfg)
    S = A.Stepper(cfg)
    Wt = cuda(W)
    dt = S.cfl_dt(Wt)
    lam_out = P.smooth_direction(cfg, 77, 0, fields=tuple(range(8)))
    lam_in = S.reverse(Wt, dt, cuda(lam_out)).cpu()
    grads = {}
    for dtype in (torch.float64, torch.float32):
        Wd = torch.tensor(W, dtype=dtype, requires_grad=True)
        out = T.step(Wd, float(dt), cfg)
        (out * torch.tensor(lam_out, dtype=dtype)).sum().backward()
        grads[dtype] = Wd.grad
    errs = group_norm_err(lam_in, grads[torch.float64])
    floor = group_norm_err(grads[torch.float32], grads[torch.float64])
    assert (errs <= np.maximum(2.0 * floor, 1e-4)).all(), (errs, floor)
    assert errs.max() < 1e-2, errs


def test_rollout_gradient_vs_fp64_autograd(gpu_ready):
    announce("8-step magnetic-energy gradient vs FP64 autograd (16^3)")
    cfg = P.config_c1(16)
    W = P.adjoint_case(cfg)
    S = A.Stepper(cfg)
    W0 = cuda(W).requires_grad_(True)
    WN, dts = A.rollout(W0, S, nsteps=8, checkpoint="uniform", every=3)
    (0.5 * (WN[5:8] ** 2).sum()).backward()
    W64 = torch.tensor(W, dtype=torch.float64, requires_grad=True)
    X = W64
    for dt in dts:
        X = T.step(X, float(dt), cfg)
    T.magnetic_loss(X).backward()
    errs = group_norm_err(W0.grad.cpu(), W64.grad)
    assert errs.max() < 1e-3, errs


@pytest.mark.parametrize("integrator", [0, 1])
def test_stage_keeping_forward_and_reverse(gpu_ready, integrator):
    announce("forward_stages/reverse_stages == forward/reverse bitwise, integrator %d"
             % integrator)
    cfg = P.make_config((16, 12, 8), cfl=0.4, integrator=integrator)
    W = P.adjoint_case(cfg)
    S = A.Stepper(cfg)
    Wt = cuda(W)
    dt = S.cfl_dt(Wt)
    ws, wn = S.forward_stages(Wt, dt)
    assert (ws is None) == (integrator == 1)
    ref = S.forward(Wt, dt)
    lam_out = cuda(P.smooth_direction(cfg, 79, 0, fields=tuple(range(8))))
    got = S.reverse_stages(Wt, ws, wn, dt, lam_out)
    want = S.reverse(Wt, dt, lam_out)
    torch.cuda.synchronize()
    bits = lambda X: X.cpu().numpy().view(np.uint32)
    assert np.array_equal(bits(wn), bits(ref))
    assert np.array_equal(bits(got), bits(want))
    if ws is not None:  # rejected on the host, before any launch
        L = A._lib()
        assert L.nmhd_adj_reverse_stages(S.h, A._ptrs(Wt), A._ptrs(ws), A._ptrs(wn), dt,
                                         A._ptrs(lam_out), A._ptrs(wn)) == 1  # lam_in = w_out
        assert L.nmhd_adj_reverse_stages(S.h, A._ptrs(Wt), None, A._ptrs(wn), dt,
                                         A._ptrs(lam_out), A._ptrs(got)) == 1  # no w_stage


def test_checkpoint_schemes_bit_identical(gpu_ready):
    announce("checkpointing all / uniform / revolve give the same gradient")
    cfg = P.make_config((16, 16, 8), cfl=0.4)
    W = P.adjoint_case(cfg)
    S = A.Stepper(cfg)
    grads = []
    dts = None
    for scheme, kw in (("all", {}), ("uniform", {"every": 3}), ("revolve", {"snapshots": 2}),
                       ("revolve", {"snapshots": 0})):
        W0 = cuda(W).requires_grad_(True)
        WN, dts = A.rollout(W0, S, nsteps=10 if dts is None else None, dts=dts,
                            checkpoint=scheme, **kw)
        (0.5 * (WN[5:8] ** 2).sum()).backward()
        grads.append(W0.grad.cpu().numpy())
    for g in grads[1:]:
        assert np.array_equal(g.view(np.uint32), grads[0].view(np.uint32))


def test_finite_differences_smooth_directions(gpu_ready):
    announce("central differences (eps 1e-3) on 8 smooth directions, 32^3, 16 steps")
    cfg = P.make_config((32, 32, 32), cfl=0.4, recon=1)
    W = P.adjoint_case(cfg)
    W[1:4] += 1.5  # bulk flow: every CT upwind choice on one side
    S = A.Stepper(cfg)
    Wc = cuda(W)
    W0 = Wc.clone().requires_grad_(True)
    WN, dts = A.rollout(W0, S, nsteps=16, checkpoint="uniform", every=4)
    (0.5 * (WN[5:8].double() ** 2).sum()).backward()
    dirs = []
    for i in range(8):
        D = P.smooth_direction(cfg, 1006, i)
        D[0] *= W[0].mean()  # density direction in units of the density
        dirs.append(torch.tensor(D, dtype=torch.float64, device="cuda"))
    r = A.finite_difference_check(S, Wc, dts, dirs, W0.grad.double(),
                                  lambda X: float(0.5 * (X[5:8].double() ** 2).sum()))
    print("FD sigma %.3g, rel errors %s, err/sigma %s" % (
        r["sigma"], ["%.2e" % v for v in r["rel_err"]], ["%.2f" % v for v in r["err_over_sigma"]]))
    # per direction: |fd - ad| <= 1e-3 |ad| + 4 sigma (sigma: the measured
    # FP32 round-off noise of a central difference), i.e. a plain relative
    # error of at most 1e-3 wherever the difference can resolve it
    assert r["ok"], r
    assert r["resolvable"] >= 6, r
    assert r["rel_err_max_resolvable"] <= 1e-3, r
