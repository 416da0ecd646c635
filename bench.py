"""Benchmark: 64-shot full-matrix-capture W2 gradient evaluations per second.

Workload (one "step" = one whole gradient evaluation): 1024 x 1024 grid,
dx = 0.05 mm, 64-element array with pitch 12 cells on row 20, each element
firing in turn and all 64 receiving, 4000 time steps at CFL 0.5, 20-cell PML
with R = 1e-3, background 5900 m/s with 1-5 random defects (observed data
from that model, gradient at the defect-free background), W2 misfit.  Every
shot of the evaluation is resident on the GPU at once.

Printed: one JSON line.  ``roofline`` is for the dominant kernel (the fused
adjoint step, 60 algorithmic bytes per cell-step); its per-launch time comes
from HIP events around the adjoint time loop on the library's stream.
``cpu_baseline`` times the CPU oracle (a restatement of the reference scheme)
on a bounded sample on this host.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

MFMA_BF16_PEAK = 2.5e15      # dense BF16 matrix-core peak, FLOP/s
HBM_PEAK = 8.0e12           # MI355X HBM3E, bytes/s
FWD_BYTES = 28              # read P, Ax, Az, c^2; write P, Ax, Az (FP32)
ADJ_BYTES = 60              # lambda 3R+3W, u_b 3R+3W, c^2 R, grad R+W


def build_problem(N, T, n_elem, seed):
    """The configuration-2 setup through the drop-in wrapper, as a reference
    driver builds it (regression.py:34-44): Cuda_interface_PML computes the
    PML profile (_calcPML) and, in init_regression with prec_deriv=4, the
    boundary ring; the observed traces come from a forward run of the true
    model.  Returns (wrapper, config)."""
    from otfwi_amd import adjoint_functions, configs
    from otfwi_amd.cuda_interface_pml import Cuda_interface_PML

    cfg = configs.config2(model_seed=seed, N=N, T=T, n_elem=n_elem)
    m = cfg["src_mask"]
    n = int(m.sum())
    iface = Cuda_interface_PML(N, N, T, cfg["pml"], cfg["c_true"], n, m, n, m, cfg["src"],
                               multishot=True, R=cfg["R"])
    return iface, cfg, adjoint_functions


def cpu_baseline(N, T, n_shots, n_rx, sample_steps):
    """Oracle (C/OpenMP restatement of the reference scheme) on this host:
    one shot, `sample_steps` forward + adjoint steps at N x N (the whole run
    by default), plus the W2 oracle on 64 traces (one shot's); extrapolated
    linearly in shots (and steps, for a shorter sample) to the full gradient.
    About 25 s of CPU on 16 cores."""
    from oracle import oracle as O
    from otfwi_amd import configs

    cores = len(os.sched_getaffinity(0))
    O.lib().orc_set_threads(cores)
    Ts = sample_steps + 1
    cfg = configs.config2(model_seed=0, N=N, T=Ts, n_elem=n_rx)
    d_x, d_z = O.calc_pml(N, N, cfg["pml"], R=cfg["R"])
    src_xz = O.point_order(cfg["src_mask"])
    ring_xz = O.point_order(O.revert_mask(N, N, cfg["pml"], 4, cfg["src_mask"]))
    cq = (cfg["c_bg"] ** 2).astype(np.float32)
    t0 = time.perf_counter()
    fw = O.forward(cq, d_x, d_z, src_xz, cfg["src"], src_xz, Ts, idx=0, ring_xz=ring_xz,
                   want_seed=True)
    t1 = time.perf_counter()
    adj = np.zeros((n_rx, Ts), np.float32)
    adj[:, : Ts // 2] = 1e-3
    O.adjoint(cq, d_x, d_z, src_xz, adj, ring_xz, fw["ring"], fw["seed"])
    t2 = time.perf_counter()
    rng = np.random.default_rng(0)
    n_w2 = 64
    sim = rng.standard_normal((n_w2, T)).cumsum(1) * 1e-3
    obs = sim + 1e-4 * rng.standard_normal((n_w2, T))
    t3 = time.perf_counter()
    O.wasserstein(sim, obs)
    t4 = time.perf_counter()
    per_pair = (t1 - t0 + t2 - t1) / sample_steps          # one fwd + one adj step, one shot
    w2_per_trace = (t4 - t3) / n_w2
    t_grad = per_pair * n_shots * (T - 1) + w2_per_trace * n_shots * n_rx
    return dict(value=1.0 / t_grad, unit="gradient evals/s", cores=cores, kind="port",
                sample=(f"1 shot x {sample_steps} forward+adjoint steps at {N}x{N} and "
                        f"{n_w2} W2 traces of T={T}, extrapolated linearly to {n_shots} shots x "
                        f"{T - 1} steps and {n_shots * n_rx} traces"),
                shot_steps_per_s=1.0 / per_pair,
                sample_seconds=round(t4 - t0, 2))


def calib_copy(nbytes):
    """One device-to-device hipMemcpy of nbytes (the runtime's copy kernel):
    the reference point that scales HBM byte counters."""
    import ctypes

    hip = ctypes.CDLL("libamdhip64.so")
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMemset(a, 1, ctypes.c_size_t(nbytes)) == 0
    assert hip.hipDeviceSynchronize() == 0
    assert hip.hipMemcpy(b, a, ctypes.c_size_t(nbytes), 3) == 0   # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    print(f"[bench] copied {nbytes} bytes device-to-device", file=sys.stderr, flush=True)


def runtime_copy(nbytes=2 << 30, reps=10):
    """Measured copy bandwidth of this GPU through the HIP runtime's own copy
    (a device-to-device copy of nbytes on torch's stream, timed with HIP
    events); bytes read + written per second (GB/s)."""
    import torch

    a = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    b.copy_(a)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        b.copy_(a)
    e1.record()
    torch.cuda.synchronize()
    sec = e0.elapsed_time(e1) / 1e3 / reps
    del a, b
    return 2 * nbytes / sec / 1e9


def sum_traffic(kernels, N, S):
    """Summed HBM bytes per launch of the listed kernels (None if any is missing)."""
    vals = [profiled_traffic(k, N, S) for k in kernels]
    return None if any(v is None for v in vals) else float(sum(vals))


def profiled_traffic(kernel, N, S):
    """HBM bytes per launch of `kernel` (e.g. "adj_step_kernel<float,32>") from
    the committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary of the 64-shot
    1024^2 workload (scaled by the 1 GiB copy calibration); None for other
    workloads or when the summary has no line for the kernel."""
    import re

    if (N, S) != (1024, 64):
        return None
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                        "hbm_traffic_fmc64_1024x1024.md")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        mt = re.search(r"traffic_bytes_per_launch " + re.escape(kernel) + r": (\S+)", f.read())
    return float(mt.group(1)) if mt else None


def unet_timing(N=8, H=1024, reps=5):
    """Training step and forward of the U-Net at the learned preconditioner's
    full-size shape (batch 8 of 1024^2, two input channels) on the MFMA
    kernels: TFLOP/s of the 3x3 convolutions against the dense BF16 peak."""
    import torch

    from otfwi_amd import unet as U

    net = U.UNet(seed=0)
    x = torch.randn(N, 2, H, H, device="cuda")
    t = 0.1 * torch.randn(N, H, H, device="cuda")
    for _ in range(2):
        net.train_step(x, t)
    e0, e1, e2 = (torch.cuda.Event(enable_timing=True) for _ in range(3))
    e0.record()
    for _ in range(reps):
        net.train_step(x, t)
    e1.record()
    for _ in range(reps):
        net.predict(x)
    e2.record()
    torch.cuda.synchronize()
    train_s = e0.elapsed_time(e1) / 1e3 / reps
    fwd_s = e1.elapsed_time(e2) / 1e3 / reps
    ftrain, ffwd = U.train_flops(N, H, H), sum(U.conv_flops(N, H, H).values())
    roof = U.train_roofline(N, H, H, hbm=HBM_PEAK, mfma=MFMA_BF16_PEAK)
    print(f"[bench] U-Net {N}x{H}^2: train step {train_s * 1e3:.2f} ms, forward {fwd_s * 1e3:.2f} ms",
          file=sys.stderr, flush=True)
    return {"bound": "mfma", "kernel": "conv3x3 + wgrad (BF16, FP32 accumulate)",
            "workload": f"U-Net 4 levels / 32 base channels, batch {N} of {H}x{H}",
            "achieved": ftrain / train_s / 1e12, "peak": MFMA_BF16_PEAK / 1e12,
            "unit": "TFLOP/s", "frac": ftrain / train_s / MFMA_BF16_PEAK,
            "train_step_ms": train_s * 1e3, "train_flops": ftrain,
            "forward_ms": fwd_s * 1e3, "forward_tflops": ffwd / fwd_s / 1e12,
            # per-kernel roofline: each convolution bound by max(HBM bytes, FLOPs), the
            # elementwise kernels by their bytes (unet.train_roofline)
            "kernel_roofline_ms": roof["ms"], "frac_of_kernel_roofline": roof["ms"] / 1e3 / train_s,
            "hbm_bound_convs": roof["hbm_bound_convs"], "mfma_bound_convs": roof["mfma_bound_convs"]}


def inversion_summary(summ, res, n_iter):
    """Time-to-tolerance figures added to a learned.summarize() dict `summ` of
    the per-model results `res` of learned.compare() (n_iter iterations).
    Where the plain method never reaches the tolerance, the learned method's
    lead is at least the plain run's whole wall time over the learned time,
    and unbounded where the plain run stalled (it ended, no step size lowering
    the misfit, before its iteration budget)."""
    summ = dict(summ)
    summ["time_to_tol_ratio"] = (summ["plain_time_to_tol"] / summ["learned_time_to_tol"]
                                 if summ["plain_time_to_tol"] and summ["learned_time_to_tol"]
                                 else None)
    lower = [r["plain"]["wall"] / r["learned"]["time_to_tol"] for r in res
             if r["learned"]["time_to_tol"] and r["plain"]["time_to_tol"] is None]
    summ["time_to_tol_ratio_lower_bound"] = min(lower) if lower else None
    summ["time_to_tol_ratio_lower_bound_per_model"] = lower
    summ["plain_stalled_above_tol"] = sum(
        1 for r in res if r["plain"]["time_to_tol"] is None and r["plain"]["iters"] < n_iter)
    summ["per_model"] = [{"seed": r["seed"], "err0": r["plain"]["err0"],
                          "plain_best_err": r["plain"]["best_err"],
                          "learned_best_err": r["learned"]["best_err"],
                          "plain_time_to_tol": r["plain"]["time_to_tol"],
                          "learned_time_to_tol": r["learned"]["time_to_tol"],
                          "plain_wall": r["plain"]["wall"],
                          "learned_wall": r["learned"]["wall"]} for r in res]
    return summ


def inversion_compare(n_models, n_iter, geometry="fmc1024", max_err0=0.025):
    """Learned vs plain inversion (30 iterations, W2 acceptance) with the
    shipped network weights, on the first n_models held-out random-defect
    models (seeds from 100000 up) whose starting velocity error lies in
    (2 %, max_err0]: the 2 % tolerance is within reach of a few steps only
    from there (at the 64-shot 1024^2 geometry the starting errors of the
    held-out models run from about 1 % to 11 %, median 6.7 %)."""
    from otfwi_amd import learned as L
    from otfwi_amd.unet import UNet

    weights = L.WEIGHTS_1024 if geometry == "fmc1024" else L.WEIGHTS
    if not os.path.exists(weights):
        return {"skipped": "no trained weights"}
    prob = L.Problem(**L.GEOMETRIES[geometry])
    seeds, s = [], 100000
    while len(seeds) < n_models and s < 101000:
        if 0.02 < L.start_error(prob, s) <= max_err0:
            seeds.append(s)
        s += 1
    net = UNet(params=L.load_weights(weights))
    res = L.compare(prob, net, seeds, n_iter=n_iter,
                    log=lambda s: print(s, file=sys.stderr, flush=True),
                    learned_scales=L.LEARNED_SCALES, patience=L.LEARNED_PATIENCE)
    summ = inversion_summary(L.summarize(res), res, n_iter)
    meta = os.path.splitext(weights)[0] + ".json"
    if os.path.exists(meta):
        with open(meta) as f:
            summ["training"] = json.load(f)
    summ["iterations"] = n_iter
    summ["learned_scales"] = list(L.LEARNED_SCALES)
    summ["learned_patience"] = L.LEARNED_PATIENCE
    summ["workload"] = (f"{geometry}: {prob.N}x{prob.N}, {prob.ns} shots, T={prob.T}, W2; "
                        f"held-out seeds {seeds} (the first from 100000 up with starting error "
                        f"in (2 %, {max_err0:.1%}])")
    return summ


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--N", type=int, default=1024)
    ap.add_argument("--T", type=int, default=4000)
    ap.add_argument("--shots", type=int, default=64)
    ap.add_argument("--th-fwd", type=int, default=0)
    ap.add_argument("--th-adj", type=int, default=0)
    ap.add_argument("--misfit", default="w2", choices=["w2", "l2", "env"])
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--pair-streams", type=int, default=0, choices=[0, 1, 2, 4],
                    help="streams of the two-step phases (0: by launch size)")
    ap.add_argument("--direct", action="store_true",
                    help="launch every time loop directly instead of replaying captured graphs "
                         "(loops whose two-step phase splits over streams are launched "
                         "directly either way)")
    ap.add_argument("--steps-per-launch", type=int, default=0, choices=[0, 1, 2],
                    help="time steps per launch of both loops (0: the autotuner's choice)")
    ap.add_argument("--cpu-steps", type=int, default=0,
                    help="forward+adjoint steps of the CPU sample (0: the whole T-1)")
    ap.add_argument("--inversion-geometry", default="fmc1024", choices=["fmc256", "fmc1024"])
    ap.add_argument("--inversion-max-err0", type=float, default=0.025)
    ap.add_argument("--inversion-models", type=int, default=3,
                    help="held-out models for the learned vs plain inversion (0: skip)")
    ap.add_argument("--inversion-iters", type=int, default=30)
    ap.add_argument("--no-unet", action="store_true", help="skip the U-Net MFMA timing")
    ap.add_argument("--calib-copy", action="store_true",
                    help="after the run, copy 1 GiB device-to-device (scale for HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path "
              f"(--gpus {args.gpus} requested); shots would shard across GPUs with one FP64 "
              f"all-reduce of the gradient per evaluation.", file=sys.stderr)
        sys.exit(2)

    from otfwi_amd import PMLModel, _lib

    N, T, S = args.N, args.T, args.shots
    iface, cfg, AF = build_problem(N, T, S, seed=0)
    m = iface.model
    shots = np.arange(S)
    iface.simulate()                                  # observed traces of the true model
    misfit_fn = {"w2": AF.wasserstein, "l2": AF.L2, "env": AF.quad_envelope}[args.misfit]
    iface.init_regression(iface.recording.copy(), misfit_fn, prec_deriv=4)
    iface._set_c_quad(cfg["c_bg"] ** 2)               # gradient at the defect-free background
    tune = None
    if args.th_fwd or args.th_adj:
        m.set_tuning(args.th_fwd, args.th_adj, 0)
    elif not args.no_autotune:
        tune = m.autotune(S, steps=20)
    if args.steps_per_launch:
        m.set_steps_per_launch(args.steps_per_launch, args.steps_per_launch)
    if args.direct:
        m.set_graph(False)
    if args.pair_streams:
        m.set_pair_streams(args.pair_streams)
    ktimes = m.kernel_times(S, reps=10)
    fs, ads = m.steps_per_launch(S)
    graphs = m.graph_modes(S)
    print(f"[bench] device {_lib.device_info()}; {S} shots {N}x{N} T={T}; tiles {m.get_tuning()}"
          f" persistent {m.get_launch()} autotune {tune}; steps per launch fwd {fs} adj {ads};"
          f" graphs {graphs}; kernel ms {ktimes}", file=sys.stderr, flush=True)

    for i in range(args.warmup):
        iface.calc_grad()
        print(f"[bench] warmup {i} done", file=sys.stderr, flush=True)
    t_ev = dict(fwd_ms=0.0, misfit_ms=0.0, adj_ms=0.0, n_fwd=0, n_adj=0)
    t0 = time.perf_counter()
    for i in range(args.steps):
        iface.calc_grad()                             # whole gradient, device misfit
        tm = m.timing()
        for k in t_ev:
            t_ev[k] += tm[k]
        print(f"[bench] step {i}: misfit {iface.mse[0]:.6e}", file=sys.stderr, flush=True)
    wall = time.perf_counter() - t0
    assert np.isfinite(iface.grad).all() and np.isfinite(iface.mse).all()
    ms_step = 1e3 * wall / args.steps
    value = args.steps / wall
    cells = S * N * N
    # Launch schedule of one gradient (the loops' own rule): the forward runs
    # T-1 steps as lead singles, an even number of pairs and a last single;
    # the adjoint runs T-1 steps as lead singles and an even number of pairs.
    fp = ((T - 2) // 2) & ~1 if fs == 2 else 0
    f1 = (T - 1) - 2 * fp
    ap_ = ((T - 1) // 2) & ~1 if ads == 2 else 0
    a1 = (T - 1) - 2 * ap_
    fwd_phase_s = t_ev["fwd_ms"] / 1e3 / args.steps
    adj_phase_s = t_ev["adj_ms"] / 1e3 / args.steps
    # Dominant unit: the adjoint step pair (reconstruction kernel + lambda
    # kernel, 2 x 60 B per cell) or the one-step adjoint kernel (60 B per
    # cell).  Its average duration comes from the timed region's adjoint-phase
    # HIP events, less the lead single steps at their own measured duration.
    if ap_:
        unit_s = (adj_phase_s - a1 * ktimes["adj"] / 1e3) / ap_
        unit_bytes = 2 * ADJ_BYTES * cells
        unit_name = (f"adjoint step pair: pair2_kernel<float, 2, 256> (reconstruction) + "
                     f"pair2_kernel<float, 1, 256> (lambda, gradient)")
        traffic = sum_traffic(["pair2_kernel<float,2,256>", "pair2_kernel<float,1,256>"], N, S)
        split = {"rec_pair_ms": ktimes["rec_pair"], "lam_pair_ms": ktimes["lam_pair"]}
    else:
        unit_s = adj_phase_s / (T - 1)
        unit_bytes = ADJ_BYTES * cells
        unit_name = f"adj_step_kernel<float,{m.get_tuning()[1]}>"
        traffic = sum_traffic([unit_name], N, S)
        split = {}
    if fp:
        fwd_unit_s = (fwd_phase_s - f1 * ktimes["fwd"] / 1e3) / fp
        fwd_bytes = 2 * FWD_BYTES * cells
        fwd_name = "pair2_kernel<float, 0, 256>"
        fwd_traffic = sum_traffic(["pair2_kernel<float,0,256>"], N, S)
    else:
        fwd_unit_s = fwd_phase_s / (T - 1)
        fwd_bytes = FWD_BYTES * cells
        fwd_name = f"fwd_step_kernel<float,{m.get_tuning()[0]}>"
        fwd_traffic = sum_traffic([fwd_name], N, S)
    unit_ach = unit_bytes / unit_s
    fwd_ach = fwd_bytes / fwd_unit_s
    alg_total = (FWD_BYTES + ADJ_BYTES) * cells * (T - 1)
    out = {
        "metric": "64-shot gradient evals/s @1024² (% HBM peak); learned vs plain time-to-tol",
        "value": value,
        "unit": "gradient evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded random-defect velocity model; observed traces simulated)",
        "config": {"workload": f"fmc{S}_{N}x{N}_T{T}_{args.misfit}_gradient", "grid": [N, N],
                   "T": T, "shots": S, "receivers": S, "misfit": args.misfit, "pml": 20,
                   "tiles": list(m.get_tuning()[:2]), "persistent": list(m.get_launch()),
                   "steps_per_launch": [fs, ads], "autotune": tune,
                   "graphs": graphs, "pair_streams": args.pair_streams},
        "roofline": {"bound": "hbm", "kernel": unit_name,
                     "achieved": unit_ach / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": unit_ach / HBM_PEAK,
                     "traffic": traffic,
                     "traffic_unit": "bytes per unit (rocprofv3 FETCH_SIZE x2 + WRITE_SIZE x1, "
                                     "profiles/hbm_traffic_fmc64_1024x1024.md)",
                     "bytes_per_cell_step": ADJ_BYTES, "cells_per_launch": cells,
                     "steps_per_unit": 2 if ap_ else 1,
                     "unit_ms": unit_s * 1e3, **split},
        "fwd_kernel": {"kernel": fwd_name, "achieved_GBps": fwd_ach / 1e9,
                       "frac": fwd_ach / HBM_PEAK, "unit_ms": fwd_unit_s * 1e3,
                       "steps_per_unit": 2 if fp else 1,
                       "traffic": fwd_traffic, "bytes_per_cell_step": FWD_BYTES},
        "kernel_ms": ktimes,
        "whole_gradient_hbm_frac": alg_total / (ms_step / 1e3) / HBM_PEAK,
        "phase_ms": {"forward": t_ev["fwd_ms"] / args.steps,
                     "misfit": t_ev["misfit_ms"] / args.steps,
                     "adjoint": t_ev["adj_ms"] / args.steps},
        "shot_steps_per_s": S * (T - 1) * value,
        "cpu_baseline": None,
    }
    if not args.no_cpu:
        cb = cpu_baseline(N, T, S, S, args.cpu_steps or T - 1)
        out["cpu_baseline"] = cb
    del m, iface
    cp = runtime_copy()
    out["runtime_copy_GBps"] = cp
    out["roofline"]["frac_of_runtime_copy"] = out["roofline"]["achieved"] / cp
    out["fwd_kernel"]["frac_of_runtime_copy"] = out["fwd_kernel"]["achieved_GBps"] / cp
    out["whole_gradient_frac_of_runtime_copy"] = out["whole_gradient_hbm_frac"] * HBM_PEAK / 1e9 / cp
    print(f"[bench] device-to-device copy {cp:.0f} GB/s (read + write)", file=sys.stderr, flush=True)
    if not args.no_unet:
        out["unet_roofline"] = unet_timing()
    if args.inversion_models > 0:
        out["inversion"] = inversion_compare(args.inversion_models, args.inversion_iters,
                                             args.inversion_geometry, args.inversion_max_err0)
    if args.calib_copy:
        calib_copy(1 << 30)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
