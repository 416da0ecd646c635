"""Learned-preconditioned inversion: training pairs generated by the
propagator, U-Net training on the GPU kernels, and the learned step that the
inversion driver judges with its misfit-decrease acceptance test.

A training pair is (input, target) on one random-defect model:

* input channel 0: the W2 gradient at the current model, restricted to the
  update mask, divided by the 99th percentile of its magnitude and clipped
  to +-5;
* input channel 1: the source-side illumination of the same run,
  log10(ill / max ill) mapped from [-ILL_DECADES, 0] to [0, 1];
* target: (c_true - c_current) / C_SCALE on the update mask.

The current model is the background (half of the pairs) or the background
plus a random fraction of the blurred true model error, so that the network
also sees the gradients of later iterations.
"""
from __future__ import annotations

import argparse
import json
import os
import time

import numpy as np

from . import configs, inversion
from .cuda_interface_pml import Cuda_interface_PML
from . import adjoint_functions as adj

WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "unet_fmc256.npz")
# the network fine-tuned at the 64-shot 1024^2 full-matrix-capture geometry
WEIGHTS_1024 = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "unet_fmc1024.npz")
# Problem keywords of the two geometries: the 256^2 block (32 elements, T = 1200, 1-3 defects
# of semi-axes 6-30 cells) and the 64-shot 1024^2 configuration itself (T = 4000, 1-5 defects
# of semi-axes 6-60 cells)
GEOMETRIES = {"fmc256": {},
              "fmc1024": dict(N=1024, T=4000, n_elem=64, pitch=12, n_max=5, semi_min=6.0,
                              semi_max=60.0)}
C_SCALE = 0.25        # target = model error / C_SCALE
# A learned proposal is tried at these scales in turn (the first that lowers
# the W2 misfit is taken); after LEARNED_PATIENCE rejected proposals the run
# takes plain steps only.  At the 64-shot 1024^2 geometry the W2 test rejected
# the full and the half proposal on 4 of 7 held-out models and accepted a
# quarter or an eighth of it on 3 of those 4.
LEARNED_SCALES = (1.0, 0.5, 0.25, 0.125)
LEARNED_PATIENCE = 2
ILL_DECADES = 6.0
G_PCT = 99.0          # gradient scaled by this percentile of |g| on the mask ...
G_CLIP = 5.0          # ... and clipped to +-G_CLIP


def features(grad, illum, mask):
    """(2, Z, X) FP32 network input from a raw gradient and an illumination map."""
    g = np.asarray(grad, np.float64) * mask
    a = np.abs(g[mask > 0])
    s = np.percentile(a, G_PCT) if a.size else 0.0
    g = np.clip(g / s, -G_CLIP, G_CLIP) if s > 0 else g
    ill = np.asarray(illum, np.float64)
    li = np.log10(np.maximum(ill / max(ill.max(), 1e-300), 10.0 ** -ILL_DECADES))
    li = 1.0 + li / ILL_DECADES
    return np.stack([g, li]).astype(np.float32)


class Problem:
    """One geometry (grid, array, T) of learn_config, reused across models:
    the propagator, its boundary ring and the update mask."""

    def __init__(self, N=256, T=1200, n_elem=32, pitch=6, pml=20, illum_stride=4, **defects):
        self.kw = dict(N=N, T=T, n_elem=n_elem, pitch=pitch, pml=pml, **defects)
        p = configs.learn_config(0, **self.kw)
        self.N, self.T, self.pml, self.ns = N, T, pml, p["n_shots"]
        self.c_bg = p["c_bg"]
        self.cp = Cuda_interface_PML(N, N, T, pml, p["c_bg"], self.ns, p["src_mask"], self.ns,
                                     p["rx_mask"], p["src"], multishot=True, R=p["R"])
        self.mask = inversion.array_mask(N, N, pml, pml + 10, taper=8)
        # the array's mirror axis: x -> mirror - x maps its elements onto each other
        xs = np.where(np.asarray(p["src_mask"]).any(0))[0]
        self.mirror = int(xs.min() + xs.max())
        self.vmin = float(configs.courant(1400.0))
        self.vmax = float(configs.courant(5900.0))
        self.shots = np.arange(self.ns, dtype=np.int32)
        self._regression_ready = False
        self.cp.model.set_illum(illum_stride)

    def model(self, seed):
        return configs.learn_config(seed, **self.kw)["c_true"]

    def observe(self, c_true):
        """Observed traces of all shots on the true model; installs them."""
        self.cp.simulate(c=c_true)
        obs = self.cp.recording.copy()
        if not self._regression_ready:
            self.cp.init_regression(obs, adj.wasserstein, prec_deriv=4)
            self._regression_ready = True
        else:
            self.cp.model.set_observed(obs)
            self.cp.observed = obs
        return obs

    def gradient(self, c):
        m = self.cp.model
        m.set_cquad((np.asarray(c, np.float64) ** 2).astype(np.float32))
        g, mis = m.gradient(self.shots, "w2")
        return g, m.illum(), float(np.mean(mis))

    def fwi(self, c_true=None, **kw):
        return inversion.FWI(self.cp, self.mask, self.vmin, self.vmax, misfit="w2",
                             c_true=c_true, **kw)


def make_pairs(prob: Problem, seeds, augment=0.5, rng_seed=0, log_every=100, log=print):
    """Training pairs on the given model seeds.  Returns (inp (n,2,N,N),
    target (n,N,N)) FP32 and the seconds spent."""
    from scipy.ndimage import gaussian_filter

    rng = np.random.default_rng(rng_seed)
    n = len(seeds)
    N = prob.N
    X = np.zeros((n, 2, N, N), np.float32)
    Y = np.zeros((n, N, N), np.float32)
    t0 = time.perf_counter()
    for k, s in enumerate(seeds):
        c_true = prob.model(int(s))
        prob.observe(c_true)
        c = prob.c_bg.copy()
        if rng.random() < augment:
            beta = rng.uniform(0.2, 0.9)
            sig = rng.uniform(1.0, 4.0)
            c = c + beta * gaussian_filter(c_true - prob.c_bg, sig)
        g, ill, _ = prob.gradient(c)
        X[k] = features(g, ill, prob.mask)
        Y[k] = ((c_true - c) * prob.mask / C_SCALE).astype(np.float32)
        if log and log_every and (k + 1) % log_every == 0:
            log(f"[pairs] {k + 1}/{n} in {time.perf_counter() - t0:.1f} s")
    return X, Y, time.perf_counter() - t0


def train(net, X, Y, epochs, batch=8, lr=1e-4, seed=0, log=print, mirror=None, mask=None):
    """Adam on mean squared error over the pairs, data resident on the GPU.
    mirror (with mask): half of the batches, drawn at random, are mirrored
    across the array's axis, x -> mirror - x (the geometry maps onto itself,
    so a mirrored pair is the pair of the mirrored model), their gradient
    channel and target multiplied by the update mask again.  Returns the
    per-epoch mean losses and the seconds spent."""
    import torch

    dev = net.dev
    Xd = torch.tensor(X, device=dev)
    Yd = torch.tensor(Y, device=dev)
    n = len(X)
    rng = np.random.default_rng(seed)
    if mirror is not None:
        W = X.shape[-1]
        mcols = torch.tensor((mirror - np.arange(W)) % W, device=dev)
        md = torch.tensor(np.asarray(mask, np.float32), device=dev)
    hist = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for ep in range(epochs):
        perm = rng.permutation(n)
        acc = []
        for i in range(0, n - batch + 1, batch):
            idx = torch.tensor(perm[i:i + batch], device=dev)
            xb, yb = Xd.index_select(0, idx), Yd.index_select(0, idx)
            if mirror is not None and rng.random() < 0.5:
                xb = xb.index_select(-1, mcols)
                yb = yb.index_select(-1, mcols) * md
                xb[:, 0] *= md
            acc.append(net.train_step(xb.contiguous(), yb.contiguous(), lr=lr))
        loss = float(torch.stack(acc).sum().item()) / (len(acc) * batch * X.shape[-1] ** 2)
        hist.append(loss)
        if log:
            log(f"[train] epoch {ep + 1}/{epochs} loss {loss:.4e} "
                f"({time.perf_counter() - t0:.1f} s)")
    torch.cuda.synchronize()
    return hist, time.perf_counter() - t0


def step_pairs(prob: Problem, net, seeds, log_every=100, log=print):
    """Second-stage pairs: on each model the network's own first step from
    the background, c1 = clip(c_bg + step), then (features at c1,
    (c_true - c1) / C_SCALE)."""
    n = len(seeds)
    N = prob.N
    X = np.zeros((n, 2, N, N), np.float32)
    Y = np.zeros((n, N, N), np.float32)
    step = learned_step(net, prob.mask)
    t0 = time.perf_counter()
    for k, s in enumerate(seeds):
        c_true = prob.model(int(s))
        prob.observe(c_true)
        g, ill, _ = prob.gradient(prob.c_bg)
        c1 = np.clip(prob.c_bg + step(g, ill, prob.c_bg), prob.vmin, prob.vmax)
        g1, ill1, _ = prob.gradient(c1)
        X[k] = features(g1, ill1, prob.mask)
        Y[k] = ((c_true - c1) * prob.mask / C_SCALE).astype(np.float32)
        if log and log_every and (k + 1) % log_every == 0:
            log(f"[pairs2] {k + 1}/{n} in {time.perf_counter() - t0:.1f} s")
    return X, Y, time.perf_counter() - t0


def chain_pairs(prob: Problem, net, seeds, steps=3, scales=(1.0, 0.5), log_every=100, log=print):
    """Pairs along the network's own step chain: from the background, take up
    to `steps` network steps, each kept (at the first of `scales` that
    lowers the W2 misfit, as inversion.FWI.run keeps learned steps) or ending
    the chain; every state reached gives (features there, (c_true - c) /
    C_SCALE)."""
    X, Y = [], []
    step = learned_step(net, prob.mask)
    t0 = time.perf_counter()
    for k, s in enumerate(seeds):
        c_true = prob.model(int(s))
        prob.observe(c_true)
        c = prob.c_bg.copy()
        g, ill, J = prob.gradient(c)
        for _ in range(steps):
            dc = step(g, ill, c)
            for sc in scales:
                ct = np.clip(c + sc * dc, prob.vmin, prob.vmax)
                gt, illt, Jt = prob.gradient(ct)
                if Jt < J:
                    break
            else:
                break
            c, g, ill, J = ct, gt, illt, Jt
            X.append(features(g, ill, prob.mask))
            Y.append(((c_true - c) * prob.mask / C_SCALE).astype(np.float32))
        if log and log_every and (k + 1) % log_every == 0:
            log(f"[chain] {k + 1}/{len(seeds)}: {len(X)} pairs in {time.perf_counter() - t0:.1f} s")
    N = prob.N
    X = np.stack(X) if X else np.zeros((0, 2, N, N), np.float32)
    Y = np.stack(Y) if Y else np.zeros((0, N, N), np.float32)
    return X, Y, time.perf_counter() - t0


def refine_recipe(prob: Problem, net, n_first=1000, n_chain=1000, steps=3, epochs=15, lr=5e-5,
                  batch=8, seed=2, log=print):
    """Fine-tune a trained network on first-step pairs (seeds 0..n_first-1)
    plus pairs along its own step chains (seeds n_first..n_first+n_chain-1).
    Returns (net, info)."""
    X1, Y1, t1 = make_pairs(prob, range(n_first), rng_seed=seed, log=log)
    X2, Y2, t2 = chain_pairs(prob, net, range(n_first, n_first + n_chain), steps=steps, log=log)
    h, tt = train(net, np.concatenate([X1, X2]), np.concatenate([Y1, Y2]), epochs, batch=batch,
                  lr=lr, seed=seed, log=log)
    return net, dict(first_pairs=n_first, chain_seeds=n_chain, chain_pairs=len(X2),
                     chain_steps=steps, epochs=epochs, lr=lr, t_pairs=t1 + t2, t_train=tt, loss=h)


def save_weights(net, path=WEIGHTS):
    """Weights as BF16 bit patterns (what the kernels compute with), biases
    and the head in FP32."""
    import torch

    p = net.params()
    from .unet import CONVS, param_layout
    lay, n = param_layout()
    keep = np.ones(n, bool)
    for name, *_ in CONVS:
        off, nw, _, _ = lay[name]
        keep[off:off + nw] = False
    bits = torch.tensor(p).to(torch.bfloat16).view(torch.int16).numpy().view(np.uint16)
    idx = np.nonzero(keep)[0].astype(np.int32)
    os.makedirs(os.path.dirname(path), exist_ok=True)
    np.savez_compressed(path, bf16=bits, idx=idx, fp32=p[idx])


def load_weights(path=WEIGHTS):
    """Flat FP32 parameter vector from save_weights."""
    d = np.load(path)
    p = (d["bf16"].astype(np.uint32) << 16).view(np.float32)
    p[d["idx"]] = d["fp32"]
    return p


def train_recipe(prob: Problem, n_pairs=2000, epochs=30, n_stage2=1000, epochs2=15, lr=1e-4,
                 batch=8, seed=0, log=print):
    """Stage 1: pairs on seeds 0..n_pairs-1, trained epochs; stage 2: pairs
    at the network's own first step on the first n_stage2 seeds, trained on
    the union for epochs2.  Returns (net, info)."""
    from .unet import UNet

    X, Y, t_pairs = make_pairs(prob, range(n_pairs), log=log)
    net = UNet(seed=seed)
    h1, t_train = train(net, X, Y, epochs, batch=batch, lr=lr, seed=seed, log=log)
    info = dict(pairs=n_pairs, epochs=epochs, lr=lr, batch=batch, t_pairs=t_pairs,
                t_train=t_train, loss=h1)
    if n_stage2:
        X2, Y2, t_p2 = step_pairs(prob, net, range(n_stage2), log=log)
        h2, t_train2 = train(net, np.concatenate([X, X2]), np.concatenate([Y, Y2]), epochs2,
                             batch=batch, lr=lr, seed=seed + 1, log=log)
        info.update(pairs2=n_stage2, epochs2=epochs2, t_pairs2=t_p2, t_train2=t_train2, loss2=h2)
    return net, info


def learned_step(net, mask):
    """Step proposal for inversion.FWI.run: the network's model update."""
    import torch

    def step(grad, illum, c):
        x = torch.tensor(features(grad, illum, mask)[None], device=net.dev)
        out = net.predict(x)
        return (out[0].float().cpu().numpy() * C_SCALE * mask).astype(np.float64)

    return step


def compare(prob: Problem, net, seeds, n_iter=30, tol=0.02, log=print, methods=("plain", "learned"),
            **run_kw):
    """Plain vs learned inversion on held-out models.  Returns a list of
    per-model dicts with the velocity-error histories and times to tol;
    "matched" is the time each method needs to reach the plain method's
    final error.  run_kw (patience, learned_scales) go to the learned runs."""
    out = []
    for s in seeds:
        c_true = prob.model(int(s))
        prob.observe(c_true)
        res = {"seed": int(s)}
        for kind in methods:
            fwi = prob.fwi(c_true=c_true)
            prog = (lambda r, s=s, kind=kind: log(
                f"[compare] seed {s} {kind} it {r.it}: {r.kind} step, err {r.err:.4f}, "
                f"{r.t:.2f} s")) if log else None
            _, h = fwi.run(prob.c_bg, n_iter=n_iter, tol=tol,
                           learned=learned_step(net, prob.mask) if kind == "learned" else None,
                           progress=prog, **(run_kw if kind == "learned" else {}))
            res[kind] = dict(time_to_tol=h.time_to(tol), best_err=h.best_err,
                             err0=h.records[0].err, n_grad=h.n_grad, n_obj=h.n_obj,
                             iters=len(h.records) - 1,
                             wall=h.records[-1].t,
                             kinds=[r.kind for r in h.records[1:]],
                             errs=[round(r.err, 5) for r in h.records],
                             times=[round(r.t, 4) for r in h.records])
            res[kind]["_h"] = h
            if log:
                log(f"[compare] seed {s} {kind}: err {h.records[0].err:.4f} -> {h.best_err:.4f} "
                    f"in {len(h.records) - 1} it ({h.records[-1].t:.2f} s), "
                    f"time to {tol:.0%}: {h.time_to(tol)}")
        if "plain" in res:
            target = res["plain"]["_h"].records[-1].err
            for kind in methods:
                res[kind]["time_to_plain_final"] = res[kind]["_h"].time_to(target)
        for kind in methods:
            res[kind].pop("_h")
        out.append(res)
    return out


def start_error(prob: Problem, seed):
    """Relative L2 velocity error of the background model against model `seed`."""
    c_true = prob.model(int(seed))
    return float(np.linalg.norm(prob.c_bg - c_true) / np.linalg.norm(c_true))


def summarize(res, tol=0.02):
    """Aggregate of compare(): mean best errors, times to tol (None where a
    method never reached it) and the matched-error speed-up."""
    def mean(v):
        v = [x for x in v if x is not None]
        return float(np.mean(v)) if v else None

    sp = [r["plain"]["time_to_plain_final"] / r["learned"]["time_to_plain_final"]
          for r in res if r["learned"]["time_to_plain_final"] and r["plain"]["time_to_plain_final"]]
    return dict(
        models=len(res), tol=tol,
        plain_err0=mean([r["plain"]["err0"] for r in res]),
        plain_best_err=mean([r["plain"]["best_err"] for r in res]),
        learned_best_err=mean([r["learned"]["best_err"] for r in res]),
        plain_reached_tol=sum(r["plain"]["time_to_tol"] is not None for r in res),
        learned_reached_tol=sum(r["learned"]["time_to_tol"] is not None for r in res),
        plain_time_to_tol=mean([r["plain"]["time_to_tol"] for r in res]),
        learned_time_to_tol=mean([r["learned"]["time_to_tol"] for r in res]),
        plain_wall=mean([r["plain"]["wall"] for r in res]),
        learned_wall=mean([r["learned"]["wall"] for r in res]),
        matched_error_speedup=mean(sp),
        matched_error_speedup_min=min(sp) if sp else None,
        learned_reached_plain_final=len(sp))


def main(argv=None):
    ap = argparse.ArgumentParser(description="train / evaluate the learned preconditioner")
    ap.add_argument("cmd", choices=["train", "refine", "finetune", "compare"])
    ap.add_argument("--geometry", default="fmc256", choices=sorted(GEOMETRIES))
    ap.add_argument("--pairs", type=int, default=2000)
    ap.add_argument("--epochs", type=int, default=30)
    ap.add_argument("--stage2", type=int, default=1000)
    ap.add_argument("--epochs2", type=int, default=15)
    ap.add_argument("--lr", type=float, default=1e-4)
    ap.add_argument("--weights", default=WEIGHTS)
    ap.add_argument("--models", type=int, default=4)
    ap.add_argument("--iters", type=int, default=30)
    ap.add_argument("--out", default=None, help="JSON summary file")
    ap.add_argument("--chain", type=int, default=1000, help="refine: seeds of step-chain pairs")
    ap.add_argument("--chain-steps", type=int, default=3)
    ap.add_argument("--first", type=int, default=1000, help="refine: first-step pairs")
    ap.add_argument("--save", default=None, help="refine / finetune: write the weights here")
    ap.add_argument("--seed0", type=int, default=0,
                    help="finetune: first model seed of the pairs (seeds seed0 .. seed0+pairs-1)")
    ap.add_argument("--no-compare", action="store_true",
                    help="finetune: skip the held-out comparison")
    ap.add_argument("--seed-range", type=int, nargs=2, default=None, metavar=("FIRST", "END"),
                    help="compare: held-out seeds FIRST .. END-1 (default 100000 .. +models)")
    ap.add_argument("--max-err0", type=float, default=None,
                    help="compare: only the seeds whose starting velocity error is at most this")
    ap.add_argument("--min-err0", type=float, default=None,
                    help="compare: only the seeds whose starting velocity error exceeds this")
    ap.add_argument("--seeds", type=int, nargs="+", default=None,
                    help="compare: these held-out seeds (overrides --seed-range)")
    ap.add_argument("--scales", type=float, nargs="+", default=list(LEARNED_SCALES),
                    help="compare: scales a learned proposal is tried at, in turn")
    ap.add_argument("--patience", type=int, default=LEARNED_PATIENCE,
                    help="compare: rejected learned proposals before plain steps only")
    ap.add_argument("--mirror", action="store_true",
                    help="finetune: mirror half of the batches across the array's axis")
    ap.add_argument("--learned-only", action="store_true",
                    help="compare: run the learned inversions only (no summary)")
    a = ap.parse_args(argv)
    log = lambda s: print(s, flush=True)  # noqa: E731
    prob = Problem(**GEOMETRIES[a.geometry])
    summary = {}
    if a.cmd == "train":
        net, info = train_recipe(prob, a.pairs, a.epochs, a.stage2, a.epochs2, a.lr, log=log)
        save_weights(net, a.weights)
        summary["train"] = info
        with open(os.path.splitext(a.weights)[0] + ".json", "w") as f:
            json.dump({k: v for k, v in info.items() if not k.startswith("loss")}, f, indent=1)
    elif a.cmd == "finetune":
        # pairs at this geometry (seeds seed0 .. seed0+pairs-1), the network in --weights as
        # the start, written to --save (default: the 1024^2 weights)
        from .unet import UNet
        X, Y, t_pairs = make_pairs(prob, range(a.seed0, a.seed0 + a.pairs), rng_seed=a.seed0,
                                   log=log, log_every=10)
        net = UNet(params=load_weights(a.weights))
        h, t_train = train(net, X, Y, a.epochs, lr=a.lr, seed=a.seed0, log=log,
                           mirror=prob.mirror if a.mirror else None, mask=prob.mask)
        out_w = a.save or WEIGHTS_1024
        save_weights(net, out_w)
        info = dict(pairs=a.pairs, seeds=[a.seed0, a.seed0 + a.pairs - 1], epochs=a.epochs,
                    lr=a.lr, mirror=bool(a.mirror), t_pairs=t_pairs, t_train=t_train,
                    init=os.path.basename(a.weights),
                    loss_first_epoch=h[0], loss_last_epoch=h[-1])
        summary["finetune"] = info
        with open(os.path.splitext(out_w)[0] + ".json", "w") as f:
            json.dump(info, f, indent=1)
        a.weights = out_w
    elif a.cmd == "refine":
        from .unet import UNet
        net = UNet(params=load_weights(a.weights))
        net, info = refine_recipe(prob, net, a.first, a.chain, a.chain_steps, a.epochs2,
                                  a.lr, log=log)
        out_w = a.save or a.weights
        save_weights(net, out_w)
        summary["refine"] = info
        with open(os.path.splitext(out_w)[0] + ".refine.json", "w") as f:
            json.dump({k: v for k, v in info.items() if not k.startswith("loss")}, f, indent=1)
    if a.cmd == "finetune" and a.no_compare:
        log(json.dumps(summary["finetune"]))
        return
    if a.cmd in ("compare", "finetune"):
        from .unet import UNet
        net = UNet(params=load_weights(a.weights))
    seeds = (list(a.seeds) if a.seeds else list(range(*a.seed_range)) if a.seed_range
             else [100000 + k for k in range(a.models)])
    if a.max_err0 is not None or a.min_err0 is not None:
        hi = np.inf if a.max_err0 is None else a.max_err0
        lo = -np.inf if a.min_err0 is None else a.min_err0
        seeds = [s for s in seeds if lo < start_error(prob, s) <= hi]
        log(f"[compare] seeds with starting error in ({lo}, {hi}]: {seeds}")
    res = compare(prob, net, seeds, n_iter=a.iters, log=log, patience=a.patience,
                  learned_scales=tuple(a.scales),
                  methods=("learned",) if a.learned_only else ("plain", "learned"))
    summary["compare"] = {} if a.learned_only else summarize(res)
    summary["models"] = res
    log(json.dumps(summary["compare"]))
    if a.out:
        with open(a.out, "w") as f:
            json.dump(summary, f)


if __name__ == "__main__":
    main()
