"""nmhd_adj_forward_noalloc (the frozen-dt step with no allocation: its
scratch is a fixed device array of the library) against nmhd_adj_forward
(the same step with scratch allocated by nmhd_adj_create): bit-identical
on two calls in a row, the second on a smaller box, so the second call
reuses the scratch the first one left behind; and sizes beyond the
scratch are rejected with a clear error."""
import ctypes

import numpy as np
import pytest
import torch

from nmhd_amd import adjoint as A, capi, problems as P

pytestmark = pytest.mark.gpu


def _case(shape, integrator, seed):
    cfg = P.make_config(shape, cfl=0.4, integrator=integrator)
    W = P.adjoint_case(cfg)
    rng = np.random.default_rng(seed)
    W[1:4] += 0.05 * rng.standard_normal(W[1:4].shape).astype(np.float32)
    return cfg, torch.tensor(np.ascontiguousarray(W), dtype=torch.float32, device="cuda")


@pytest.mark.parametrize("integrator", [0, 1])
def test_noalloc_step_equals_allocating_step(gpu_ready, integrator):
    for shape, seed in (((48, 40, 32), 1), ((24, 16, 8), 2)):
        print("\n[case] noalloc vs nmhd_adj_forward, %s, integrator %d" % (shape, integrator),
              flush=True)
        cfg, W = _case(shape, integrator, seed)
        S = A.Stepper(cfg)
        dt = S.cfl_dt(W)
        ref = S.forward(W, dt)
        got = torch.full_like(W, float("nan"))
        A.forward_noalloc(cfg, W, dt, got)
        torch.cuda.synchronize()
        S.close()
        assert torch.isfinite(got).all()
        assert np.array_equal(got.cpu().numpy().view(np.uint32), ref.cpu().numpy().view(np.uint32))


def test_noalloc_rejects_boxes_beyond_the_scratch(gpu_ready):
    print("\n[case] noalloc size limit", flush=True)
    # 128 x 128 x 65 = 1,064,960 cells > 2^20; the arrays have the box's
    # size, so the call would be a valid step but for the limit
    big, W = _case((128, 128, 65), 0, 3)
    out = torch.zeros_like(W)
    L = A._lib()
    d = capi.desc_from_config(big)
    rc = L.nmhd_adj_forward_noalloc(ctypes.byref(d), A._ptrs(W), 1e-4, A._ptrs(out), None)
    torch.cuda.synchronize()
    assert rc == 1
    assert b"exceeds the no-allocation scratch" in L.nmhd_adj_last_error(None)
    assert not out.any()  # nothing was launched
    with pytest.raises(capi.NmhdError, match="exceeds the no-allocation scratch"):
        A.forward_noalloc(big, W, 1e-4, out)
