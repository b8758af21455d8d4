"""Training of the learned preconditioner's graph network (preconditioner = 2).

Samples come from the plain solver itself: synthetic scenes with seeds
disjoint from every test and bench seed are solved with block-Jacobi PCG, and
after each LM iteration the camera graph of that iteration's S is logged
together with the PCG iteration count of the same S and b for a grid of
Neumann weights w (rpba_ctx_lp_sample).  The label of a sample is the weight
with the fewest iterations (the mean of the grid weights that tie for it).
The network is then trained on the GPU by the library's own forward,
backward and Adam kernels (rpba_lp_train) and its parameters are written as
little-endian float32 (lp_weights.f32 beside this file by default).

    python -m rpba_amd.lp_train [--out PATH] [--scenes N] [--steps K]
"""
import argparse
import json
import os
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
WEIGHTS = os.path.join(HERE, "lp_weights.f32")
OMEGAS = [0.0, 0.2, 0.4, 0.5, 0.6, 0.7, 0.8, 0.85, 0.9, 0.95]
TRAIN_SEED0 = 1000   # training scenes: seeds 1000 ..
HOLDOUT_SEED0 = 2000  # held-out scenes: seeds 2000 ..


def scene_for(seed):
    """Synthetic scene of the sample set: 30-160 cameras, 120 points per
    camera, track length 2 + geometric with mean 6..9, 2% outliers."""
    from rpba_amd import scenes
    rng = np.random.default_rng(seed)
    nc = int(rng.integers(30, 161))
    tm = float(rng.uniform(6.0, 9.0))
    return scenes.generate(None, seed=seed, n_cams=nc, n_pts=120 * nc, track="geometric", track_mean=tm,
                           outlier_frac=0.02)


def label_of(iters):
    it = np.asarray(iters)
    best = it.min()
    return float(np.mean(np.asarray(OMEGAS)[it == best]))


def collect(seeds, omegas=OMEGAS, max_iters=40):
    """One sample per LM iteration of the plain robust solve of each scene."""
    from rpba_amd import api, capi
    samples = []
    for seed in seeds:
        sc = scene_for(seed)
        ctx = api.Context(sc, capi.make_flags(robust=1))
        for _ in range(max_iters):
            rc, done, fin = ctx.run(1)
            if rc != 0 or done == 0:
                break
            smp = ctx.lp_sample(omegas)
            smp["seed"] = seed
            smp["label"] = label_of(smp["iters"])
            samples.append(smp)
            if fin:
                break
        ctx.close()
        print(f"lp_train: scene {seed} ({sc['n_cams']} cameras): {len(samples)} samples", flush=True)
    return samples


def train(samples, steps=600, lr=2e-3, seed=0):
    from rpba_amd import api
    model = api.LpModel(seed=seed)
    labels = np.array([s["label"] for s in samples], dtype=np.float32)
    trace = model.train(samples, labels, steps, lr)
    return model, trace


def evaluate(model, seeds, max_iters=40):
    """PCG iterations on the held-out scenes' S matrices: block-Jacobi, the
    best fixed weight of the grid, and the network's weight."""
    from rpba_amd import api, capi
    tot = {"block_jacobi": 0, "fixed_best": None, "learned": 0, "oracle_label": 0}
    per_w = np.zeros(len(OMEGAS), dtype=np.int64)
    n = 0
    for seed in seeds:
        sc = scene_for(seed)
        ctx = api.Context(sc, capi.make_flags(robust=1))
        for _ in range(max_iters):
            rc, done, fin = ctx.run(1)
            if rc != 0 or done == 0:
                break
            g = ctx.lp_sample(OMEGAS)
            w = float(model.predict([g])[0])
            it_l = ctx.lp_sample([w])["iters"][0] if w > 0 else g["iters"][0]
            per_w += g["iters"]
            tot["block_jacobi"] += int(g["iters"][0])
            tot["learned"] += int(it_l)
            tot["oracle_label"] += int(g["iters"].min())
            n += 1
            if fin:
                break
        ctx.close()
    tot["fixed_best"] = int(per_w[1:].min())
    tot["fixed_best_omega"] = OMEGAS[1 + int(np.argmin(per_w[1:]))]
    tot["samples"] = n
    return tot


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=WEIGHTS)
    ap.add_argument("--scenes", type=int, default=60)
    ap.add_argument("--holdout", type=int, default=8)
    ap.add_argument("--steps", type=int, default=600)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--report", default=None, help="write a JSON summary here")
    args = ap.parse_args()
    t0 = time.time()
    samples = collect(range(TRAIN_SEED0, TRAIN_SEED0 + args.scenes))
    t1 = time.time()
    print(f"lp_train: {len(samples)} samples from {args.scenes} scenes in {t1 - t0:.1f} s", flush=True)
    model, trace = train(samples, args.steps, args.lr)
    t2 = time.time()
    print(f"lp_train: {args.steps} Adam steps in {t2 - t1:.1f} s, loss {trace[0]:.5f} -> {trace[-1]:.5f}", flush=True)
    model.save(args.out)
    ev = evaluate(model, range(HOLDOUT_SEED0, HOLDOUT_SEED0 + args.holdout))
    print("lp_train: held-out PCG iterations", json.dumps(ev), flush=True)
    if args.report:
        labels = [s["label"] for s in samples]
        with open(args.report, "w") as fh:
            json.dump({"samples": len(samples), "label_mean": float(np.mean(labels)),
                       "label_hist": {str(w): int(sum(1 for x in labels if abs(x - w) < 1e-6)) for w in OMEGAS},
                       "loss_first": float(trace[0]), "loss_last": float(trace[-1]), "steps": args.steps,
                       "lr": args.lr, "collect_s": round(t1 - t0, 1), "train_s": round(t2 - t1, 1),
                       "holdout": ev}, fh, indent=1)


if __name__ == "__main__":
    main()
