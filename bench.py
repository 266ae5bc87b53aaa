"""Benchmark: zone-updates/s of the MHD step on the 256^3 shearing box.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]

One "step" is one full time step of the shearing-box MRI configuration
(ghost fill, CFL reduction, two RK stages) on a 256^3 box (array 256x264x256
with 4 radial ghost layers per side), Omega = 1, q = 1.5, gamma = 5/3,
beta = 400, seeded velocity noise.  The state stays resident in HBM; the
timed region covers exactly K steps enqueued back to back, bracketed by
device synchronisation.  value = interior cells * K / seconds.

The JSON line also carries
  roofline     -- the dominant kernel's algorithmic HBM bytes per launch over
                  its mean launch time (HIP events on the library's stream),
                  against the 8 TB/s HBM peak;
  step_roofline -- the whole step: 160 algorithmic bytes per zone-update;
  cpu_baseline -- the oracle (CPU restatement of the reference step) on the
                  host cores of the machine that runs the bench
                  (OMP_NUM_THREADS when set, else every core of the process's
                  affinity mask), a 10-step sample of the same box, with a
                  one-thread figure beside it;
  batched      -- 64 periodic boxes of uneven sizes (edges from
                  {16,24,32,48,64,96}, seed 1004; Orszag-Tang and blast
                  states) advanced 100 steps with per-box dt by the batched
                  path (persistent tile launches, HIP-graph replay, autotuned
                  tiles), against the same boxes advanced one after another
                  by single-box contexts;
  adjoint      -- the hand-written adjoint on the 64^3 Orszag-Tang box with
                  a smooth density perturbation (seed 1005): 64 steps with dt
                  recorded from the forward pass, loss = magnetic energy,
                  checkpoints every 8 steps; backward/forward GPU-time ratio
                  (HIP events, median of 5) and central finite differences
                  (eps 1e-3) along 32 smooth random directions (seed 1006),
                  per direction against a measured FP32 noise floor, on the
                  same state in a bulk flow with piecewise-constant
                  reconstruction (where the discrete step is differentiable
                  along them), and, reported beside it, on the workload
                  state itself.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12
# non-packed FP32 VALU issue peak: 256 CUs x 4 SIMDs x 16 lanes x 2.4 GHz
VALU_PEAK = 256 * 4 * 16 * 2.4e9
STEP_BYTES = 160.0  # algorithmic bytes per zone-update (8 fields x 5 passes x 4 B)
# Algorithmic bytes per interior cell of one launch of each kernel, per path.
KERNEL_BYTES = {
    "tiled": {
        "faces_s1": 76.0,    # read W (8), write the hydro divergence (5) + 6 B fluxes
        "faces_s2": 76.0,
        "update_s1": 108.0,  # read divergence (5), B fluxes (6), W^n (8); write 8
        "update_s2": 120.0,  # read divergence (5), B fluxes (6), winds of W* (3), W^n (8); write 8
        "cfl": 32.0,         # read 8
    },
    "unfused": {
        "flux_x": 60.0,      # read W (8), write 7 face fluxes
        "flux_y": 60.0,
        "flux_z": 60.0,
        "update_s1": 148.0,  # read 21 fluxes, W^n (8); write 8
        "update_s2": 160.0,  # read 21 fluxes, winds of W* (3), W^n (8); write 8
        "cfl": 32.0,
    },
}
KERNEL_BYTES["strict"] = KERNEL_BYTES["unfused"]
KERNEL_BYTES["auto"] = KERNEL_BYTES["tiled"]


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--size", type=int, nargs=3, default=None,
                    help="interior cells (default 256^3; 512 1024 512 for --workload large)")
    ap.add_argument("--workload", choices=("mri", "large"), default="mri",
                    help="mri: the 256^3 beta-400 shearing box with velocity noise; large: "
                         "the 512x1024x512 beta-1000 box with the filtered density noise "
                         "(per-stage GB/s; no CPU baseline, batched or adjoint legs)")
    ap.add_argument("--path", choices=("auto", "unfused", "tiled", "strict"), default="auto")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=10)
    ap.add_argument("--profile-steps", type=int, default=10)
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--batch-boxes", type=int, default=64)
    ap.add_argument("--batch-steps", type=int, default=100)
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--adjoint-size", type=int, default=64)
    ap.add_argument("--adjoint-steps", type=int, default=64)
    ap.add_argument("--fd-directions", type=int, default=32)
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="also run one device copy of this many MiB (a known-size kernel "
                         "that calibrates HBM byte counters in a profiler run)")
    return ap.parse_args()


def host_threads():
    """Threads for the CPU baseline: OMP_NUM_THREADS when set (the share of
    the host this process is given), else every core this process may run
    on.  Returns (threads, a record of what the host reports)."""
    omp = os.environ.get("OMP_NUM_THREADS")
    nproc = os.cpu_count() or 1
    try:
        affinity = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        affinity = nproc
    threads = int(omp) if omp and omp.strip().isdigit() and int(omp) > 0 else affinity
    return threads, {"nproc": nproc, "affinity_cpus": affinity, "OMP_NUM_THREADS": omp}


def c1_leg():
    """The 32^3 Orszag-Tang configuration (periodic, gamma 5/3, CFL 0.4) for
    its full 100 steps from the initial condition on the default path: a
    small box, where launch latency and the first step's full CFL pass
    weigh (its parity with the oracle, bit for bit on the strict path, is a
    GPU test)."""
    from nmhd_amd import capi, problems as P
    cfg = P.config_c1(32)
    g = capi.Context(cfg)
    g.set_state(P.orszag_tang(cfg))
    g.synchronize()
    t0 = time.perf_counter()
    k, _ = g.run(1e30, 100)
    el = time.perf_counter() - t0
    t_end = g.get_time()[0]
    g.close()
    return {"value": P.interior_cells(cfg) * k / el, "unit": "zone-updates/s", "steps": k,
            "ms_per_step": el / k * 1e3, "t_end": t_end}


def cpu_baseline(cfg, W, steps):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    threads, host = host_threads()
    orc = pyoracle.Oracle(cfg, W, threads=threads)
    t0 = time.perf_counter()
    for _ in range(steps):
        orc.step()
    el = time.perf_counter() - t0
    orc.close()
    from nmhd_amd import problems as P
    # the same step on one core, on a smaller box of the same setup
    c1 = P.config_c2((128, 128, 128))
    one = pyoracle.Oracle(c1, P.shearing_box_mri(c1, seed=1002), threads=1)
    t1 = time.perf_counter()
    one.step()
    el1 = time.perf_counter() - t1
    one.close()
    # the 32^3 Orszag-Tang configuration for its full 100 steps, all threads
    oc = P.config_c1(32)
    ot = pyoracle.Oracle(oc, P.orszag_tang(oc), threads=threads)
    t2 = time.perf_counter()
    k2, _ = ot.run(1e30, 100)
    el2 = time.perf_counter() - t2
    ot.close()
    return {
        "value": P.interior_cells(cfg) * steps / el,
        "unit": "zone-updates/s",
        "cores": threads,
        "host": host,
        "kind": "port",
        "sample": "%d steps of the same %dx%dx%d shearing box on the oracle "
                  "(C restatement of the reference step, OpenMP, %d threads)"
                  % ((steps,) + tuple(cfg["global_size"]) + (threads,)),
        "single_thread": {"value": P.interior_cells(c1) / el1, "unit": "zone-updates/s",
                          "sample": "1 step of the 128^3 box of the same setup, 1 thread"},
        "c1_ot32": {"value": P.interior_cells(oc) * k2 / el2, "unit": "zone-updates/s",
                    "seconds": el2,
                    "sample": "32^3 Orszag-Tang, %d steps, %d threads" % (k2, threads)},
    }


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "hbm_traffic_shearing_box_256.json")


def measured_profile(kernel, size, path):
    """The committed rocprofv3 record of `kernel` for this workload: HBM
    bytes per launch (FETCH_SIZE / WRITE_SIZE, scaled) and VALU wave
    instructions per launch (SQ_INSTS_VALU); {} when the summary does not
    cover this size, path and kernel."""
    try:
        with open(TRAFFIC_FILE) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return {}
    if list(t.get("global_size", [])) != list(size) or t.get("path") != path:
        return {}
    return t.get("kernels", {}).get(kernel) or {}


def batched_leg(a, torch):
    from nmhd_amd import capi, problems as P
    boxes = P.batch_boxes(seed=1004, nbox=a.batch_boxes)
    cells = sum(P.interior_cells(c) for c, _ in boxes)
    B = capi.Batch([c for c, _ in boxes])
    for i, (_, W) in enumerate(boxes):
        B.set_state(i, W)
    t0 = time.perf_counter()
    timed_shapes = B.autotune()
    tune_s = time.perf_counter() - t0
    B.run(1e30, a.warmup, use_graph=True)  # the full-CFL first step, graph capture
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    kmin, _ = B.run(1e30, a.batch_steps, use_graph=True)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    assert kmin == a.batch_steps, kmin
    batched = cells * a.batch_steps / el
    launch = B.get_launch()
    B.close()
    ctxs = []
    for cfg, W in boxes:
        g = capi.Context(cfg, path=capi.PATH_TILED)
        g.set_state(W)
        g.run(1e30, a.warmup)
        ctxs.append(g)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for g in ctxs:
        k, _ = g.run(1e30, a.batch_steps)
        assert k == a.batch_steps
    torch.cuda.synchronize()
    el2 = time.perf_counter() - t0
    for g in ctxs:
        g.close()
    plain = cells * a.batch_steps / el2
    return {
        "value": batched, "unit": "zone-updates/s", "boxes": len(boxes), "cells": cells,
        "distinct_shapes": len({tuple(c["global_size"]) for c, _ in boxes}),
        "steps": a.batch_steps, "ms_per_step": el / a.batch_steps * 1e3,
        "plain_sequential": plain, "plain_ms_per_step": el2 / a.batch_steps * 1e3,
        "vs_plain_sequential": batched / plain,
        "launch": {"faces_zchunk": launch[0], "update_zchunk": launch[1],
                   "faces_wg_per_cu": launch[2], "update_wg_per_cu": launch[3]},
        "autotune_s": tune_s, "autotune_timed_shapes": timed_shapes,
        "workload": "64 periodic boxes, edges i.i.d. from {16,24,32,48,64,96} (seed 1004), "
                    "Orszag-Tang (B0 ~ U[0.5,1.5]/sqrt(4 pi)) and blast states, per-box dt, "
                    "HIP-graph replay; baseline: the same boxes stepped one after another by "
                    "single-box contexts (tiled path)",
    }


def adjoint_leg(a, torch):
    import numpy as np
    from nmhd_amd import adjoint as A, problems as P
    cfg = P.config_c1(a.adjoint_size)
    W = P.adjoint_case(cfg, seed=1005)
    S = A.Stepper(cfg)
    Wc = torch.tensor(W, dtype=torch.float32, device="cuda")
    with torch.no_grad():
        dts = A.record_dts(Wc, S, a.adjoint_steps)
    reps = 5

    def passes(lead_cycles=0):
        """One forward (rollout + loss) and one backward.  Returns the host
        wall times of both (each until the device finished it, when no lead
        is given), the HIP-event times of both on torch's stream, and the
        host time spent enqueueing the pair.  With lead_cycles a spin kernel
        is enqueued first: the device is then busy while the host enqueues
        both passes, so the events time the device's own work."""
        W0 = Wc.clone().requires_grad_(True)
        torch.cuda.synchronize()
        if lead_cycles:
            torch.cuda._sleep(lead_cycles)
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
        h0 = time.perf_counter()
        ev[0].record()
        WN, _ = A.rollout(W0, S, dts=dts, checkpoint="uniform", every=8)
        loss = 0.5 * (WN[5:8].double() ** 2).sum()
        ev[1].record()
        if not lead_cycles:
            torch.cuda.synchronize()
        h1 = time.perf_counter()
        loss.backward()
        ev[2].record()
        h2 = time.perf_counter()
        torch.cuda.synchronize()
        h3 = time.perf_counter()
        return {"wall_fwd": h1 - h0, "wall_bwd": h3 - h1, "enqueue": h2 - h0,
                "gpu_fwd": ev[0].elapsed_time(ev[1]) * 1e-3,
                "gpu_bwd": ev[1].elapsed_time(ev[2]) * 1e-3, "grad": W0.grad}

    passes()  # warm-up of both passes
    wall = [passes() for _ in range(reps)]
    grad = wall[-1]["grad"].double()
    # spin-kernel rate, to size the lead: a spin of 1e7 cycles timed by events
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    torch.cuda._sleep(10 ** 7)
    e1.record()
    torch.cuda.synchronize()
    cycles_per_s = 1e7 / max(e0.elapsed_time(e1) * 1e-3, 1e-9)
    lead_s = min(1.0, 3.0 * max(w["wall_fwd"] + w["wall_bwd"] for w in wall))
    dev = [passes(int(lead_s * cycles_per_s)) for _ in range(reps)]
    med = lambda rows, k: float(np.median([r[k] for r in rows]))
    t_fwd, t_bwd = med(wall, "wall_fwd"), med(wall, "wall_bwd")
    g_fwd, g_bwd = med(dev, "gpu_fwd"), med(dev, "gpu_bwd")
    host_ahead = all(r["enqueue"] < lead_s for r in dev)

    def magnetic_energy(X):
        return float(0.5 * (X[5:8].double() ** 2).sum())

    eps = 1e-3

    def directions(cfg_, W_):
        out = []
        for i in range(a.fd_directions):
            D = P.smooth_direction(cfg_, 1006, i)
            D[0] *= W_[0].mean()  # density direction in units of the density
            out.append(torch.tensor(D, dtype=torch.float64, device="cuda"))
        return out

    # central differences on the workload state itself: reported, not a
    # criterion (the discrete step is not differentiable there: four-cell
    # velocity sums that choose the CT upwind side sit at or near zero and
    # flip under the perturbation, and the minmod limiter has kinks)
    fd_work = A.finite_difference_check(S, Wc, dts, directions(cfg, W), grad, magnetic_energy,
                                        eps=eps)
    # where the step is differentiable along the directions: the same state
    # in a bulk flow of 1.5 (every CT upwind choice on one side) with
    # piecewise-constant reconstruction (no limiter kinks)
    cfg_s = P.make_config((a.adjoint_size,) * 3, cfl=0.4, recon=1)
    Ws = P.adjoint_case(cfg_s, seed=1005)
    Ws[1:4] += 1.5
    S_s = A.Stepper(cfg_s)
    Wsc = torch.tensor(Ws, dtype=torch.float32, device="cuda")
    with torch.no_grad():
        dts_s = A.record_dts(Wsc, S_s, a.adjoint_steps)
    W0s = Wsc.clone().requires_grad_(True)
    WNs, _ = A.rollout(W0s, S_s, dts=dts_s, checkpoint="uniform", every=8)
    (0.5 * (WNs[5:8].double() ** 2).sum()).backward()
    fd = A.finite_difference_check(S_s, Wsc, dts_s, directions(cfg_s, Ws), W0s.grad.double(),
                                   magnetic_energy, eps=eps)
    S_s.close()
    zones = P.interior_cells(cfg) * a.adjoint_steps
    return {
        "gpu_forward_ms": g_fwd * 1e3, "gpu_backward_ms": g_bwd * 1e3,
        "gpu_backward_over_forward": g_bwd / g_fwd,
        "gpu_timing": "median of %d; HIP events on torch's stream, the device kept busy by a "
                      "%.0f ms spin kernel while the host enqueued both passes (host ahead of "
                      "the device: %s)" % (reps, lead_s * 1e3, host_ahead),
        "wall_forward_ms": t_fwd * 1e3, "wall_backward_ms": t_bwd * 1e3,
        "wall_backward_over_forward": t_bwd / t_fwd,
        "wall_forward_ms_range": [min(r["wall_fwd"] for r in wall) * 1e3,
                                  max(r["wall_fwd"] for r in wall) * 1e3],
        "wall_timing": "median of %d; host clock until the device finished each pass "
                       "(includes the host's launch overhead)" % reps,
        "forward_zone_updates_per_s": zones / g_fwd,
        "adjoint_zone_updates_per_s": zones / g_bwd,
        "fd_directions": a.fd_directions, "fd_eps": eps,
        "fd_noise_sigma": fd["sigma"],
        "fd_criterion": "|fd - ad| <= 1e-3 |ad| + 4 sigma per direction, sigma = the FP32 "
                        "round-off noise of a central difference, measured from the loss at 8 "
                        "inputs moved by at most half an ulp",
        "fd_pass": fd["ok"],
        "fd_resolvable_directions": fd["resolvable"],
        "fd_rel_err_max_resolvable": fd["rel_err_max_resolvable"],
        "fd_rel_err_median": fd["rel_err_median"],
        "fd_rel_err": [float("%.3g" % v) for v in fd["rel_err"]],
        "fd_err_over_sigma": [float("%.3g" % v) for v in fd["err_over_sigma"]],
        "fd_err_over_rms_derivative_max": fd["err_over_rms_derivative_max"],
        "fd_on_workload_state": {
            "rel_err_median": fd_work["rel_err_median"],
            "rel_err_max": max(fd_work["rel_err"]),
            "pass": fd_work["ok"],
            "note": "the discrete step is not differentiable on this state (CT upwind sums at "
                    "zero, minmod kinks); an exact FP64 derivative of the same step fails the "
                    "same way"},
        "workload": "%d^3 Orszag-Tang + 8 Fourier density modes (|k|<=4, 1e-2, seed 1005) + "
                    "smooth velocity modes (1e-3), %d steps, dt recorded from the forward pass, "
                    "loss = 1/2 sum B^2, checkpoints every 8 steps (backward includes "
                    "recomputation)" % (a.adjoint_size, a.adjoint_steps),
        "fd_workload": "the same state in a bulk flow of 1.5 per axis with piecewise-constant "
                       "reconstruction (the step is differentiable along the directions), "
                       "same steps; directions: smooth Fourier modes in rho, v (seed 1006)",
    }


def stage_rates(cfg, tim, kbytes, zones):
    """Algorithmic GB/s of each stage of the step: reconstruction + HLLD (the
    face kernels), CT + update (the update kernels), CFL reduction (the
    ghost-cell pass; the interior is reduced in the update's epilogue) and the
    boundary (shearing-periodic ghost fill, 8 fields read + written per ghost
    cell)."""
    m = cfg["m"][1]
    ghost_cells = 2 * m * cfg["n"][0] * cfg["n"][2]

    def rate(names, nbytes):
        ms = sum(tim[k][0] for k in names if k in tim)
        n = sum(tim[k][1] for k in names if k in tim)
        return None if not n else nbytes * n / (ms * 1e-3) / 1e9

    return {
        "reconstruct_hlld": rate(["faces_s1", "faces_s2"], kbytes["faces_s1"] * zones),
        "ct_update": rate(["update_s1", "update_s2"],
                          0.5 * (kbytes["update_s1"] + kbytes["update_s2"]) * zones),
        "cfl_reduce": rate(["cfl_ghosts"], 32.0 * ghost_cells),
        "boundary": rate(["ghost_y"], 64.0 * ghost_cells),
    }


def main():
    a = parse()
    if a.gpus != 1:
        print("bench.py: the step runs on one GPU; --gpus %d is not supported (a "
              "multi-GPU run would shard the box in z-slabs with a 4-plane halo "
              "exchange and one dt all-reduce per step)" % a.gpus,
              file=sys.stderr)
        sys.exit(2)
    import torch
    from nmhd_amd import capi, problems as P

    torch.cuda.set_device(0)
    if a.copy_mib:
        src = torch.empty(a.copy_mib * (1 << 18), dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        src.fill_(1.0)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    large = a.workload == "large"
    if a.size is None:
        a.size = (512, 1024, 512) if large else (256, 256, 256)
    if large:
        a.no_cpu = a.no_batched = a.no_adjoint = True
    cfg = P.config_c2(tuple(a.size))
    path = {"auto": capi.PATH_AUTO, "unfused": capi.PATH_UNFUSED,
            "tiled": capi.PATH_TILED, "strict": capi.PATH_STRICT}[a.path]
    ctx = capi.Context(cfg, path=path)
    if large:
        Wd = P.shearing_box_noise_device(cfg, seed=1003, beta=1000.0)
        torch.cuda.synchronize()
        ctx.set_state_device([Wd[i].data_ptr() for i in range(8)])
        del Wd
        torch.cuda.empty_cache()
        W = None
    else:
        W = P.shearing_box_mri(cfg, seed=1002)
        ctx.set_state(W)
    goal = 1e30
    zones = P.interior_cells(cfg)

    ctx.reset_run()
    ctx.enqueue_steps(goal, a.warmup)
    ctx.synchronize()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ctx.enqueue_steps(goal, a.steps)
    done, _ = ctx.synchronize()
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    assert done == a.warmup + a.steps, (done, a.warmup, a.steps)
    value = zones * a.steps / el

    # per-kernel times with HIP events on the library's stream
    ctx.timing_enable(True)
    ctx.timing_reset()
    ctx.enqueue_steps(goal, a.profile_steps)
    ctx.synchronize()
    tim = ctx.timing()
    ctx.timing_enable(False)
    path_name = ["auto", "unfused", "tiled", "strict"][ctx.get_path()]
    kbytes = KERNEL_BYTES[path_name]
    kern = {k: v for k, v in tim.items() if v[1] > 0 and k in kbytes}
    dom = max(kern, key=lambda k: kern[k][0])
    mean_s = kern[dom][0] / kern[dom][1] / 1e3
    achieved = kbytes[dom] * zones / mean_s
    prof = measured_profile(dom, cfg["global_size"], path_name)
    roofline = {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9,
                "unit": "GB/s", "frac": achieved / HBM_PEAK,
                "traffic": prof.get("bytes_per_launch"),
                "kernel": dom, "mean_launch_ms": mean_s * 1e3,
                "algorithmic_bytes_per_launch": kbytes[dom] * zones}
    if "valu_insts" in prof:
        # the bound that binds this kernel: its VALU issue rate (instruction
        # count per launch from the profile, time live) against the
        # non-packed FP32 peak.  The count belongs to the kernel the profile
        # traced; a live time far from the traced one means the profile is
        # of another build, and its count is not used.
        lane_ops = prof["valu_insts"] * 64.0
        traced = prof.get("trace_mean_us")
        if traced and abs(mean_s * 1e6 - traced) > 0.05 * traced:
            roofline["valu"] = None
            roofline["valu_note"] = ("committed profile traced this kernel at %.1f us, "
                                     "measured now %.1f us: count not used"
                                     % (traced, mean_s * 1e6))
        else:
            roofline["valu"] = {"lane_ops_per_launch": lane_ops,
                                "achieved": lane_ops / mean_s / 1e12,
                                "peak": VALU_PEAK / 1e12, "unit": "Tlane-op/s",
                                "frac": lane_ops / mean_s / VALU_PEAK,
                                "profile_trace_mean_us": traced}
    step_achieved = value * STEP_BYTES
    out = {
        "metric": "zone-updates/s at 256^3 fp32, % HBM roofline; batched & adjoint vs plain",
        "value": value,
        "unit": "zone-updates/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": el / a.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded shearing-box MRI initial state)",
        "config": {"workload": "shearing-box MRI step, %dx%dx%d cells (array %dx%dx%d incl. "
                               "4 radial ghost layers/side), gamma 5/3, %s, CFL 0.4"
                               % (tuple(cfg["global_size"]) + tuple(cfg["n"]) +
                                  ("beta 1000, filtered density noise 1e-3 peaked at k=8 "
                                   "(seed 1003)" if large else
                                   "beta 400, velocity noise 0.01 (seed 1002)",)),
                   "global_size": cfg["global_size"], "steps": a.steps,
                   "path": path_name},
        "roofline": roofline,
        "step_roofline": {"achieved": step_achieved / 1e9, "peak": HBM_PEAK / 1e9,
                          "unit": "GB/s", "frac": step_achieved / HBM_PEAK,
                          "bytes_per_zone_update": STEP_BYTES},
        "kernel_ms_per_step": {k: v[0] / a.profile_steps for k, v in tim.items() if v[1]},
    }
    if large:
        out["stage_gbps"] = stage_rates(cfg, tim, kbytes, zones)
    ctx.close()
    if not large:
        out["c1_ot32"] = c1_leg()
    out["batched"] = None if a.no_batched else batched_leg(a, torch)
    out["adjoint"] = None if a.no_adjoint else adjoint_leg(a, torch)
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline(cfg, W, a.cpu_steps)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
