"""Throughput runs of the uneven multi-model batch and of the checkpointed
exact gradient on one GPU (python -m otfwi_amd.scale c3|exact).  Each prints
one JSON summary line; progress goes to stderr.

c3     16 independent models with grid sizes drawn from {512, 768, 1024,
       1536, 2048}^2, a 32-element array and 32 shots each, T scaled with the
       grid depth as T = round(2500 + 3500 (N - 512) / 1536) (times --t-scale),
       one gradient of every model (--misfit) through otfwi_gradient_multi with
       1 and with --workers host threads, and through otfwi_gradient_fused
       (every two-step launch covering all models, each phase one graph).
exact  one checkpointed exact-adjoint gradient (L2) of the 1024^2 FMC block,
       beside the reference-mode gradient of the same shots.
c4     the 2048^2 differentiation configuration (16 shots, T = --T): FP32 and
       FP64 checkpointed exact gradients, FP64 central differences along 8
       Gaussian bumps, the reference-mode gradient (run_c4).
cpml   the second-order C-PML scheme (regression_CPML.cu): --shots runs of a
       --N^2 grid, T = --T steps, each shot firing its own source, advanced in
       one batch; 56 algorithmic bytes per cell-step (both frames in one
       pass: P(t-1) and P(t-2) read, P(t) written, kappa and the 2 densities
       read, the 4 memory variables read and written).
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np

from . import configs
from .cuda_interface_pml import Cuda_interface_PML
from .model import gradient_multi

FWD_B, ADJ_B = 28, 60


def log(s):
    print(s, file=sys.stderr, flush=True)


def _model(N, T, n_elem, pitch, seed, max_batch):
    p = configs.config2(model_seed=seed, N=N, T=T, n_elem=n_elem, pitch=pitch)
    cp = Cuda_interface_PML(N, N, T, p["pml"], p["c_true"], n_elem, p["src_mask"], n_elem,
                            p["rx_mask"], p["src"], multishot=True, R=p["R"],
                            max_batch=max_batch)
    cp.simulate()
    from . import adjoint_functions as adj
    cp.init_regression(cp.recording.copy(), adj.wasserstein, prec_deriv=4)
    cp.model.set_cquad((p["c_bg"] ** 2).astype(np.float32))
    return cp


def run_c3(n_models=16, shots=32, t_scale=1.0, workers=4, max_batch=0, seed=0, misfit="w2"):
    rng = np.random.default_rng(seed)
    sizes = rng.choice(configs.C3_SIZES, size=n_models)
    models, cps, alg = [], [], 0.0
    for i, N in enumerate(sizes):
        T = configs.config3_T(int(N), t_scale)
        pitch = configs.config3_pitch(int(N), shots)
        cp = _model(int(N), T, shots, pitch, seed=1000 + i, max_batch=max_batch)
        cps.append(cp)
        models.append(cp.model)
        alg += (FWD_B + ADJ_B) * float(N) * N * shots * (T - 1)
        log(f"[c3] model {i}: {N}^2, T={T}, {shots} shots")
    shot_lists = [np.arange(shots)] * n_models
    out = {"workload": f"C3: {n_models} models, sizes {sorted(int(s) for s in sizes)}, "
                       f"{shots} shots each, T scale {t_scale}, {misfit}",
           "algorithmic_bytes": alg}
    ref = None
    # the fused launch needs every model's shots in one batch
    for w in (1, workers) + (() if max_batch else ("fused",)):
        kw = dict(fused=True) if w == "fused" else dict(workers=w)
        gradient_multi(models, shot_lists, misfit, **kw)      # graphs, buffers
        t0 = time.perf_counter()
        res = gradient_multi(models, shot_lists, misfit, **kw)
        dt = time.perf_counter() - t0
        assert all(np.isfinite(g).all() for g, _ in res)
        if ref is None:
            ref = res
        same = all(np.array_equal(g, g0) and np.array_equal(m, m0)
                   for (g, m), (g0, m0) in zip(res, ref))
        key = "fused" if w == "fused" else f"workers_{w}"
        out[key] = {"seconds": dt, "hbm_frac": alg / dt / 8e12,
                    "model_gradients_per_s": n_models / dt, "bitwise_equal_to_first": same}
        log(f"[c3] {key}: {dt:.2f} s ({alg / dt / 8e12:.1%} of HBM peak, algorithmic), "
            f"bitwise equal to the first run: {same}")
    return out


def run_exact(N=1024, T=4000, shots=8, snap=0):
    cp = _model(N, T, 64, 12, seed=0, max_batch=0)
    m = cp.model
    s = np.arange(shots)
    m.gradient(s, "l2")
    t0 = time.perf_counter()
    g_ref, _ = m.gradient(s, "l2")
    t_ref = time.perf_counter() - t0
    m.gradient_exact(s, "l2", snap_every=snap)
    t0 = time.perf_counter()
    g_ex, mis = m.gradient_exact(s, "l2", snap_every=snap)
    t_ex = time.perf_counter() - t0
    k = snap or int(round(np.sqrt(T - 1)))
    pml = 20
    inner = np.zeros((N, N), bool)
    inner[pml + 16:N - pml - 16, pml + 16:N - pml - 16] = True
    a, b = g_ref[inner].astype(np.float64), g_ex[inner].astype(np.float64)
    cos = float(a @ b / np.linalg.norm(a) / np.linalg.norm(b))
    plane = (N + 128) * (N + 16) * 4
    return {"workload": f"{shots}-shot {N}^2 T={T} L2 gradient", "snap_every": k,
            "exact_seconds": t_ex, "reference_mode_seconds": t_ref,
            "exact_over_reference": t_ex / t_ref,
            "checkpoint_bytes": (int(np.ceil((T - 1) / k)) * 4 + k * 2) * plane * shots,
            "full_history_bytes": (T - 1) * 2 * plane * shots,
            "interior_cosine_exact_vs_reference": cos}


def run_c4(T=8000, shots=16, n_bumps=8, misfit="l2", fp64_batch=8, seed=0):
    """The differentiation configuration: 2048^2 at dx = 0.025 mm, 16 shots,
    T = 8000.  Observed traces from the FP64 build on the true model; at the
    background + half the model error: the FP32 exact gradient (snapshots
    every sqrt(T-1) steps, all shots resident), the FP64 exact gradient
    (fp64_batch shots resident at a time, to fit HBM) as its oracle, the
    reference-mode FP32 gradient, and central differences of the FP64
    objective (step 1e-3 m/s of velocity) along n_bumps Gaussian bumps of
    sigma 1 mm (40 cells) placed at seeded positions below the array."""
    from .model import PMLModel

    cfg = configs.config4(model_seed=seed, T=T, n_elem=shots)
    N, pml = cfg["X"], cfg["pml"]
    cp = Cuda_interface_PML(N, N, T, pml, cfg["c_true"], shots, cfg["src_mask"], shots,
                            cfg["rx_mask"], cfg["src"], multishot=True, R=cfg["R"])
    from . import adjoint_functions as adj
    cp.init_regression(np.zeros((shots, shots, T), np.float32), adj.L2, prec_deriv=4)
    d_x, d_z = cp.d_x, cp.d_z
    sx, sz = cp.pos_source_x, cp.pos_source_z
    rx, rz = cp.pos_sensor_x, cp.pos_sensor_z
    vx, vz = cp._pos_revert_x, cp._pos_revert_z
    cp.model.close()

    def model(prec, max_batch=0):
        m = PMLModel(N, N, T, precision=prec)
        m.set_damping(d_x, d_z)
        m.set_sources(sx, sz, cfg["src"].astype(np.float32))
        m.set_receivers(rx, rz)
        m.set_ring(vx, vz)
        if max_batch:
            m.set_tuning(0, 0, max_batch)
        return m

    s = np.arange(shots)
    cq_true = cfg["c_true"] ** 2
    cq = 0.5 * (cfg["c_bg"] ** 2 + cq_true)
    out = {"workload": f"C4: {shots}-shot {N}^2 T={T} {misfit} gradient, dx 0.025 mm"}
    m64 = model(8, fp64_batch)
    m64.set_cquad64(cq_true)
    t0 = time.perf_counter()
    obs = m64.forward(s)
    log(f"[c4] FP64 observed traces in {time.perf_counter() - t0:.1f} s")
    m64.set_observed(obs)
    m64.set_cquad64(cq)
    k = int(round(np.sqrt(T - 1)))
    t0 = time.perf_counter()
    g64, mis64 = m64.gradient_exact(s, misfit, snap_every=k, fp64=True)
    out["fp64_exact_seconds"] = time.perf_counter() - t0
    log(f"[c4] FP64 exact gradient in {out['fp64_exact_seconds']:.1f} s")
    rng = np.random.default_rng(seed)
    h = 2 * 0.5 * float(configs.courant(1e-3))
    zz, xx = np.mgrid[0:N, 0:N]
    fd = []
    t0 = time.perf_counter()
    for b in range(n_bumps):
        cz, cx = rng.uniform(0.2 * N, N - pml - 60), rng.uniform(pml + 60, N - pml - 60)
        d = np.exp(-((zz - cz) ** 2 + (xx - cx) ** 2) / (2 * 40.0 ** 2))
        m64.set_cquad64(cq + h * d)
        jp = m64.objective(s, misfit).mean()
        m64.set_cquad64(cq - h * d)
        jm = m64.objective(s, misfit).mean()
        f = (jp - jm) / (2 * h)
        a = float((g64 * d).sum())
        fd.append({"bump": [round(cz, 1), round(cx, 1)], "fd": f, "adjoint": a,
                   "rel": abs(f - a) / abs(f)})
        log(f"[c4] bump {b}: fd {f:.9e} adjoint {a:.9e} rel {abs(f - a) / abs(f):.2e}")
    out["fd_seconds"] = time.perf_counter() - t0
    out["fd"] = fd
    out["fd_worst_rel"] = max(r["rel"] for r in fd)
    m64.close()
    del m64
    m32 = model(4)
    m32.set_observed(obs.astype(np.float32))
    m32.set_cquad((cq).astype(np.float32))
    t0 = time.perf_counter()
    g32, mis32 = m32.gradient_exact(s, misfit, snap_every=k)
    out["fp32_exact_seconds"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    g_ref, _ = m32.gradient(s, misfit)
    out["fp32_reference_mode_seconds"] = time.perf_counter() - t0
    m32.close()
    rel = lambda a, b: float(np.linalg.norm(a - b) / np.linalg.norm(b))
    inner = np.zeros((N, N), bool)     # away from the ring and the PML edges, as run_exact
    inner[pml + 16:N - pml - 16, pml + 16:N - pml - 16] = True
    a, b = g32[inner].astype(np.float64), g_ref[inner].astype(np.float64)
    out.update(snap_every=k, fp32_exact_vs_fp64=rel(g32.astype(np.float64), g64),
               fp32_misfit_vs_fp64=abs(mis32.mean() - mis64.mean()) / abs(mis64.mean()),
               cosine_exact_vs_reference_mode=float(a @ b / np.linalg.norm(a) / np.linalg.norm(b)),
               checkpoint_bytes_fp32=(int(np.ceil((T - 1) / k)) * 4 + k * 2)
               * (N + 128) * (N + 16) * 4 * shots,
               full_history_bytes_fp32=(T - 1) * 2 * (N + 128) * (N + 16) * 4 * shots)
    log(f"[c4] FP32 exact {out['fp32_exact_seconds']:.1f} s vs FP64 {out['fp32_exact_vs_fp64']:.2e}")
    return out


# Bytes per cell-step of the second-order C-PML step with both frames in one
# pass: P(t-1), P(t-2) read, P(t) written, kappa, rho_hx, rho_hz read, the four
# memory variables read and written (the x / z derivative fields stay on chip).
CPML_B = 56


def run_cpml(N=1024, T=1000, shots=64):
    from . import cpml

    p = configs.cpml_config(N, N, T, n_src=shots, n_rx=shots, L=20)
    m = cpml.CPMLModel.from_problem(p)
    idx = np.arange(shots)
    m.forward(idx[:1])
    t0 = time.perf_counter()
    rec = m.forward(idx)
    dt = time.perf_counter() - t0
    assert np.isfinite(rec).all()
    cells = float(N) * N * shots * (T - 2)
    log(f"[cpml] {shots} shots {N}^2 T={T}: {dt:.2f} s, "
        f"{CPML_B * cells / dt / 8e12:.1%} of HBM peak (algorithmic)")
    return {"workload": f"C-PML second-order scheme, {shots} shots of {N}^2, T={T}",
            "seconds": dt, "cell_steps_per_s": cells / dt, "bytes_per_cell_step": CPML_B,
            "hbm_frac": CPML_B * cells / dt / 8e12}


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("what", choices=["c3", "exact", "cpml", "c4"])
    ap.add_argument("--t-scale", type=float, default=1.0)
    ap.add_argument("--shots", type=int, default=32)
    ap.add_argument("--workers", type=int, default=4)
    ap.add_argument("--models", type=int, default=16)
    ap.add_argument("--T", type=int, default=4000)
    ap.add_argument("--N", type=int, default=1024, help="cpml: grid size")
    ap.add_argument("--misfit", default="w2", choices=["w2", "l2"])
    ap.add_argument("--max-batch", type=int, default=0,
                    help="c3: shots resident per model (0, the default: all of its shots)")
    a = ap.parse_args(argv)
    if a.what == "c3":
        out = run_c3(a.models, a.shots, a.t_scale, a.workers, max_batch=a.max_batch,
                     misfit=a.misfit)
    elif a.what == "cpml":
        out = run_cpml(N=a.N, T=a.T, shots=a.shots)
    elif a.what == "c4":
        out = run_c4(T=a.T, shots=a.shots, misfit=a.misfit)
    else:
        out = run_exact(T=a.T, shots=min(a.shots, 64))
    print(json.dumps(out))


if __name__ == "__main__":
    main()
