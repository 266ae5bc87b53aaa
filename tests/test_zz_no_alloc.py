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


def test_noalloc_step_replayed_from_a_captured_graph(gpu_ready):
    """The use the entry point is for: a step captured into a HIP graph (a
    torch CUDA graph here) after the first call on the device has done the
    set-up outside the capture.  Each replay, with the program's own eager
    work and a new input between replays, equals the allocating step on the
    same input bit for bit."""
    print("\n[case] noalloc step captured in a graph, 3 replays", flush=True)
    cfg, W0 = _case((32, 24, 16), 0, 4)
    S = A.Stepper(cfg)
    dt = S.cfl_dt(W0)
    A.forward_noalloc(cfg, W0, dt, torch.empty_like(W0))  # set-up outside the capture
    torch.cuda.synchronize()
    static_in = W0.clone()
    static_out = torch.zeros_like(W0)
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        A.forward_noalloc(cfg, static_in, dt, static_out)
    torch.cuda.synchronize()
    for rep in range(3):
        _, Wr = _case((32, 24, 16), 0, 10 + rep)
        ref = S.forward(Wr, dt)
        static_in.copy_(Wr)
        static_out.fill_(float("nan"))
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        assert np.array_equal(static_out.cpu().numpy().view(np.uint32),
                              ref.cpu().numpy().view(np.uint32)), rep
    S.close()
