"""Differentiable time stepping: the hand-written adjoint (nmhd_adj_* in
include/nmhd.h) behind torch.autograd.

A thin PyTorch-ROCm binding: torch owns the device arrays and the stream;
the library's kernels run on torch's current stream through the C ABI, so
no copies or extra synchronisation are needed.

    stepper = Stepper(cfg)                      # periodic box, source none
    W0 = torch.tensor(state, device="cuda", requires_grad=True)
    WN, dts = rollout(W0, stepper, nsteps=64, checkpoint="uniform", every=8)
    loss = 0.5 * (WN[5:8] ** 2).sum()           # magnetic energy
    loss.backward()                             # W0.grad via the adjoint

dt is frozen: the forward pass records each step's CFL dt (or takes the
given ones), and the backward pass reuses them.  Checkpointing keeps a
subset of the states of the forward pass and recomputes the others in the
backward pass:
  "all"      every state and every step's midpoint stage state (memory
             2N, no recomputation);
  "uniform"  every `every`-th state (default round(sqrt(N))): memory
             N/every + 4 every, one recomputed forward pass that keeps the
             segment's stage states, so the reverse steps run no forward
             kernels; the recomputation of the next segment runs on a side
             stream while the current one is reversed;
  "revolve"  binomial checkpointing with `snapshots` slots (default
             ceil(log2 N)): the split points minimise the number of
             recomputed steps for that memory (the offline optimum that
             Griewank's revolve schedule attains), found by dynamic
             programming over (steps, slots).
"""
import ctypes
import functools
import math

import torch

from . import capi

_FPA = ctypes.c_void_p * 8
_OPA = ctypes.POINTER(ctypes.c_void_p)  # an optional array of 8 pointers (None = NULL)


def _sig(L):
    if getattr(L, "_adj_sig", False):
        return L
    V = ctypes.c_void_p
    FP = ctypes.POINTER(ctypes.c_float)
    sig = {
        "nmhd_adj_create": [ctypes.POINTER(capi.NmhdDesc), ctypes.POINTER(V)],
        "nmhd_adj_destroy": [V],
        "nmhd_adj_get_geometry": [V, ctypes.POINTER(capi.NmhdGeometry)],
        "nmhd_adj_set_stream": [V, V],
        "nmhd_adj_synchronize": [V],
        "nmhd_adj_cfl_dt": [V, _FPA, ctypes.c_float, FP],
        "nmhd_adj_forward": [V, _FPA, ctypes.c_float, _FPA],
        "nmhd_adj_reverse": [V, _FPA, ctypes.c_float, _FPA, _FPA],
        "nmhd_adj_forward_stages": [V, _FPA, ctypes.c_float, _OPA, _FPA],
        "nmhd_adj_reverse_stages": [V, _FPA, _OPA, _FPA, ctypes.c_float, _FPA, _FPA],
        "nmhd_adj_forward_noalloc": [ctypes.POINTER(capi.NmhdDesc), _FPA, ctypes.c_float, _FPA,
                                     V],
    }
    for name, args in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = ctypes.c_int
    L.nmhd_adj_last_error.argtypes = [V]
    L.nmhd_adj_last_error.restype = ctypes.c_char_p
    L._adj_sig = True
    return L


def _lib():
    return _sig(capi.load())


def _ptrs(W):
    if not (W.is_cuda and W.dtype == torch.float32 and W.is_contiguous()):
        raise ValueError("state arrays must be contiguous float32 CUDA tensors")
    return _FPA(*[W[i].data_ptr() for i in range(8)])


class Stepper:
    """S(W; dt) of one periodic box and its reverse, on torch's stream."""

    def __init__(self, cfg):
        L = _lib()
        self.cfg = cfg
        self._desc = capi.desc_from_config(cfg)
        h = ctypes.c_void_p()
        self._check(L.nmhd_adj_create(ctypes.byref(self._desc), ctypes.byref(h)), None)
        self.h = h
        g = capi.NmhdGeometry()
        self._check(L.nmhd_adj_get_geometry(self.h, ctypes.byref(g)), self.h)
        self.shape = (8, g.size[2], g.size[1], g.size[0])

    @staticmethod
    def _check(rc, h):
        if rc != 0:
            msg = _lib().nmhd_adj_last_error(h)
            raise capi.NmhdError("nmhd adjoint error %s: %s"
                                 % (capi.ERRORS.get(rc, rc), msg.decode() if msg else ""))

    def close(self):
        twin = getattr(self, "_twin", None)
        if twin is not None:
            twin.close()
            self._twin = None
        if getattr(self, "h", None):
            _lib().nmhd_adj_destroy(self.h)
            self.h = None

    def side(self):
        """(side stream, second stepper of the same box) for work that runs
        beside this stepper's: the checkpointed backward pass recomputes the
        next segment there.  Both are made once and kept."""
        if getattr(self, "_twin", None) is None:
            self._twin = Stepper(self.cfg)
            self._side_stream = torch.cuda.Stream()
        return self._side_stream, self._twin

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _on_torch_stream(self):
        s = torch.cuda.current_stream().cuda_stream
        self._check(_lib().nmhd_adj_set_stream(self.h, ctypes.c_void_p(s)), self.h)

    def _shape_ok(self, W):
        if tuple(W.shape) != self.shape:
            raise ValueError("state shape %s != %s" % (tuple(W.shape), self.shape))

    def cfl_dt(self, W, goal=1e30):
        self._shape_ok(W)
        self._on_torch_stream()
        dt = ctypes.c_float()
        self._check(_lib().nmhd_adj_cfl_dt(self.h, _ptrs(W), goal, ctypes.byref(dt)), self.h)
        return dt.value

    def forward(self, W, dt):
        self._shape_ok(W)
        self._on_torch_stream()
        out = torch.empty_like(W)
        self._check(_lib().nmhd_adj_forward(self.h, _ptrs(W), float(dt), _ptrs(out)), self.h)
        return out

    def reverse(self, W, dt, lam_out):
        self._shape_ok(W)
        self._shape_ok(lam_out)
        self._on_torch_stream()
        lam_out = lam_out.contiguous()
        lam_in = torch.empty_like(W)
        self._check(_lib().nmhd_adj_reverse(self.h, _ptrs(W), float(dt), _ptrs(lam_out),
                                            _ptrs(lam_in)), self.h)
        return lam_in

    def two_stage(self):
        return self._desc.integrator != capi.EULER1

    def forward_stages(self, W, dt):
        """(W*, S(W; dt)): the step's midpoint stage state (None for the
        one-stage integrator) and its result."""
        self._shape_ok(W)
        self._on_torch_stream()
        out = torch.empty_like(W)
        ws = torch.empty_like(W) if self.two_stage() else None
        self._check(_lib().nmhd_adj_forward_stages(self.h, _ptrs(W), float(dt),
                                                   _ptrs(ws) if ws is not None else None,
                                                   _ptrs(out)), self.h)
        return ws, out

    def reverse_stages(self, W, Ws, Wout, dt, lam_out):
        """reverse() with the step's stage states (from forward_stages) given
        instead of recomputed; bitwise the same result."""
        for X in (W, Wout, lam_out):
            self._shape_ok(X)
        if self.two_stage():
            self._shape_ok(Ws)
        self._on_torch_stream()
        lam_out = lam_out.contiguous()
        lam_in = torch.empty_like(W)
        self._check(_lib().nmhd_adj_reverse_stages(
            self.h, _ptrs(W), _ptrs(Ws) if Ws is not None else None, _ptrs(Wout), float(dt),
            _ptrs(lam_out), _ptrs(lam_in)), self.h)
        return lam_in


def forward_noalloc(cfg, W, dt, out):
    """out = S(W; dt) through nmhd_adj_forward_noalloc on torch's current
    stream: the library allocates nothing (its scratch is a fixed device
    array, boxes of at most NMHD_NOALLOC_MAX_CELLS cells).  W and out are
    caller-owned (8, nz, ny, nx) float32 CUDA tensors."""
    L = _lib()
    desc = capi.desc_from_config(cfg)
    if tuple(W.shape) != tuple(out.shape):
        raise ValueError("W and out differ in shape")
    s = torch.cuda.current_stream().cuda_stream
    Stepper._check(L.nmhd_adj_forward_noalloc(ctypes.byref(desc), _ptrs(W), float(dt), _ptrs(out),
                                              ctypes.c_void_p(s)), None)
    return out


# ------------------------------------------------------------ checkpointing

@functools.lru_cache(maxsize=None)
def revolve_cost(steps, slots):
    """Fewest forward steps needed to reverse `steps` steps when the first
    state is held and `slots` more states may be stored at once."""
    if steps <= 1:
        return 0
    if slots == 0:
        return steps * (steps - 1) // 2
    return min(m + revolve_cost(steps - m, slots - 1) + revolve_cost(m, slots)
               for m in range(1, steps))


@functools.lru_cache(maxsize=None)
def revolve_split(steps, slots):
    best, arg = None, 1
    for m in range(1, steps):
        c = m + revolve_cost(steps - m, slots - 1) + revolve_cost(m, slots)
        if best is None or c < best:
            best, arg = c, m
    return arg


def reverse_segment(stepper, W_a, a, b, dts, lam, slots, counter=None):
    """Reverse steps a..b-1 given the state W_a = W^a; returns dL/dW^a."""
    fwd = stepper.forward
    if b - a == 1:
        return stepper.reverse(W_a, dts[a], lam)
    if slots == 0:
        for n in range(b - 1, a - 1, -1):
            W = W_a
            for j in range(a, n):
                W = fwd(W, dts[j])
                if counter is not None:
                    counter[0] += 1
            lam = stepper.reverse(W, dts[n], lam)
        return lam
    m = revolve_split(b - a, slots)
    W = W_a
    for j in range(a, a + m):
        W = fwd(W, dts[j])
        if counter is not None:
            counter[0] += 1
    lam = reverse_segment(stepper, W, a + m, b, dts, lam, slots - 1, counter)
    del W
    return reverse_segment(stepper, W_a, a, a + m, dts, lam, slots, counter)


class _Rollout(torch.autograd.Function):
    @staticmethod
    def forward(ctx, W0, stepper, dts, checkpoint, every, snapshots):
        n = len(dts)
        keep = {0: W0.detach()}
        stage = {}
        W = W0.detach()
        for k in range(n):
            if checkpoint == "all":
                stage[k], W = stepper.forward_stages(W, dts[k])
            else:
                W = stepper.forward(W, dts[k])
            if checkpoint == "all" or (checkpoint == "uniform" and (k + 1) % every == 0
                                       and k + 1 < n):
                keep[k + 1] = W
        ctx.stepper, ctx.dts, ctx.keep, ctx.stage = stepper, list(dts), keep, stage
        ctx.checkpoint, ctx.every, ctx.snapshots = checkpoint, every, snapshots
        return W

    @staticmethod
    def backward(ctx, gN):
        stepper, dts, keep = ctx.stepper, ctx.dts, ctx.keep
        n = len(dts)
        lam = gN.contiguous()
        if ctx.checkpoint == "revolve":
            lam = reverse_segment(stepper, keep[0], 0, n, dts, lam, ctx.snapshots)
        elif ctx.checkpoint == "all":
            for j in range(n - 1, -1, -1):
                lam = stepper.reverse_stages(keep[j], ctx.stage[j], keep[j + 1], dts[j], lam)
        else:
            # one forward pass over each segment keeps every step's stage
            # states, so no step's reverse runs the forward kernels again.
            # The recomputation of segment i-1 does not depend on the
            # reverse of segment i: it runs on a side stream (default
            # priority) with a second stepper, whose scratch is its own,
            # while segment i is reversed on the caller's stream.
            starts = sorted(keep)
            main = torch.cuda.current_stream()
            side, helper = stepper.side()
            side.wait_stream(main)  # the checkpoints are the forward's outputs

            def recompute(i):
                a = starts[i]
                b = starts[i + 1] if i + 1 < len(starts) else n
                with torch.cuda.stream(side):
                    states, stages = [keep[a]], []
                    for j in range(a, b):
                        ws, w = helper.forward_stages(states[-1], dts[j])
                        stages.append(ws)
                        states.append(w)
                    done = torch.cuda.Event()
                    done.record(side)
                return a, b, states, stages, done

            pending = recompute(len(starts) - 1)
            for i in range(len(starts) - 1, -1, -1):
                a, b, states, stages, done = pending
                main.wait_event(done)
                if i > 0:
                    pending = recompute(i - 1)
                # made on the side stream, read on this one: not reused
                # before this stream's work on them is done
                for t in states[1:] + [t for t in stages if t is not None]:
                    t.record_stream(main)
                for j in range(b - 1, a - 1, -1):
                    lam = stepper.reverse_stages(states[j - a], stages[j - a],
                                                 states[j - a + 1], dts[j], lam)
                    states[j - a + 1] = stages[j - a] = None
                del states, stages
        ctx.keep = ctx.stage = None
        return lam, None, None, None, None, None


def record_dts(W0, stepper, nsteps, goal=1e30):
    """The CFL dt of every step of the forward trajectory from W0."""
    dts = []
    W = W0.detach()
    t = 0.0
    for _ in range(nsteps):
        dt = stepper.cfl_dt(W, goal)
        if t + dt > goal:
            dt = float(goal - t)
        dts.append(dt)
        W = stepper.forward(W, dt)
        t += dt
    return dts


def rollout(W0, stepper, nsteps=None, dts=None, checkpoint="uniform", every=None,
            snapshots=None):
    """W_N = S(...S(W0; dt_0)...; dt_{N-1}), differentiable w.r.t. W0.

    dts: the step sizes (frozen); when None they are recorded from the CFL
    condition along the trajectory (an extra forward pass without grad).
    Returns (W_N, dts)."""
    if dts is None:
        if nsteps is None:
            raise ValueError("give nsteps or dts")
        with torch.no_grad():
            dts = record_dts(W0, stepper, nsteps)
    n = len(dts)
    if checkpoint not in ("all", "uniform", "revolve"):
        raise ValueError("checkpoint must be 'all', 'uniform' or 'revolve'")
    if every is None:
        every = max(1, int(round(math.sqrt(n))))
    if snapshots is None:
        snapshots = max(1, int(math.ceil(math.log2(max(n, 2)))))
    return _Rollout.apply(W0.contiguous(), stepper, tuple(dts), checkpoint, every,
                          snapshots), list(dts)


def finite_difference_check(stepper, W0, dts, directions, grad, loss, eps=1e-3, nperturb=8,
                            seed=5):
    """Central differences of `loss` over the rollout of `dts` from W0 along
    each direction, against the adjoint's directional derivatives grad . D.

    The FP32 forward makes every evaluation of the loss carry round-off
    noise: the perturbed inputs W0 +- eps D are rounded to FP32, and the
    step's own rounding responds to any change of its input.  sigma, the
    standard deviation of that noise in a central difference, is measured:
    the loss at nperturb copies of W0 moved by random amounts of at most
    half an ulp per value, spread s, gives sigma = s sqrt(2) / (2 eps).  A
    central difference cannot resolve a derivative better than a few sigma,
    so the per-direction criterion is |fd - ad| <= 1e-3 |ad| + 4 sigma: a
    plain relative error of at most 1e-3 on every direction whose
    derivative exceeds 4 sigma / 1e-3, and agreement to the noise on the
    others.

    directions: float64 CUDA tensors shaped like W0; grad: dloss/dW0 (float64);
    loss: a function of the final state returning a float.
    Returns a dict of per-direction lists and the summary."""
    import numpy as np

    def run(Wi):
        X = Wi
        with torch.no_grad():
            for dt in dts:
                X = stepper.forward(X, dt)
        return loss(X)

    Wd = W0.detach().double()
    L0 = run(W0.detach())
    gen = torch.Generator(device=W0.device).manual_seed(seed)
    spread = []
    for _ in range(nperturb):
        r = torch.rand(W0.shape, generator=gen, device=W0.device, dtype=torch.float64) * 2 - 1
        spread.append(run((Wd * (1 + r * 2.0 ** -24)).float().contiguous()) - L0)
    sigma = float(np.std(spread)) * math.sqrt(2.0) / (2 * eps)
    fd, ad = [], []
    for D in directions:
        fp = run((Wd + eps * D).float().contiguous())
        fm = run((Wd - eps * D).float().contiguous())
        fd.append((fp - fm) / (2 * eps))
        ad.append(float((grad * D).sum()))
    fd, ad = np.array(fd), np.array(ad)
    err = np.abs(fd - ad)
    rel = err / np.maximum(np.abs(ad), 1e-300)
    resolvable = np.abs(ad) >= 4 * sigma / 1e-3
    return {
        "sigma": sigma, "fd": fd.tolist(), "ad": ad.tolist(), "rel_err": rel.tolist(),
        "err_over_sigma": (err / max(sigma, 1e-300)).tolist(),
        "ok": bool(np.all(err <= 1e-3 * np.abs(ad) + 4 * sigma)),
        "resolvable": int(resolvable.sum()),
        "rel_err_max_resolvable": float(rel[resolvable].max()) if resolvable.any() else None,
        "rel_err_median": float(np.median(rel)),
        "err_over_rms_derivative_max": float(err.max() / np.sqrt((ad ** 2).mean())),
    }
