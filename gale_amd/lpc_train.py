"""Train the learned Schur preconditioner on the GPU.

Pairs come from plain GPU Uzawa solves of random_viscosity_3d fields (the
generator of the 96^3 workload, other seeds) at a training resolution; the
network is trained by the library's own BF16 forward / backward kernels and
Adam, then saved as FP32 parameters (.npy) with a JSON summary beside them.

    python -m gale_amd.lpc_train --res 32 --solves 200 --steps 3000 --out gale_amd/lpc_weights.npy
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
if os.path.dirname(HERE) not in sys.path:
    sys.path.insert(0, os.path.dirname(HERE))

import gale_amd as G  # noqa: E402
from gale_amd import configs  # noqa: E402


def generate(ds, res, ppc, solves, krylov_every, tol, seed0, log=print):
    """Plain FP64 Uzawa solves recording (r0, p* - p0) and Krylov pairs."""
    t0 = time.perf_counter()
    bounds = [0]
    its = []
    for i in range(solves):
        cfg = configs.random_viscosity_3d(res=res, ppc=ppc, seed=seed0 + i, field_seed=seed0 + 7919 + i)
        pb = G.Problem(**cfg.problem_kwargs())
        s = G.Stokes_SLE_UzawaSolver(pb, G.uzawa_params(tolerance=tol, maxIterations=20000))
        s.record_pairs(ds, krylov_every)
        _, _, rep = s.Solve(raise_on_error=False)
        s.record_pairs(None)
        s.close()
        pb.close()
        its.append(int(rep.uzawa_iterations))
        bounds.append(len(ds))
        if (i + 1) % 10 == 0 or i == solves - 1:
            log(f"# solve {i + 1}/{solves}: status {rep.status}, {rep.uzawa_iterations} Uzawa its, "
                f"{rep.solve_seconds:.2f}s; pairs {len(ds)}; {time.perf_counter() - t0:.0f}s")
    return bounds, its, time.perf_counter() - t0


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--res", type=int, default=32)
    ap.add_argument("--ppc", type=int, default=16)
    ap.add_argument("--solves", type=int, default=200)
    ap.add_argument("--krylov-every", type=int, default=8)
    ap.add_argument("--tol", type=float, default=1e-6)
    ap.add_argument("--capacity", type=int, default=20000)
    ap.add_argument("--steps", type=int, default=3000)
    ap.add_argument("--batch", type=int, default=8)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--val-frac", type=float, default=0.1)
    ap.add_argument("--out", default=os.path.join(HERE, "lpc_weights.npy"))
    a = ap.parse_args(argv)

    def log(m):
        print(m, file=sys.stderr, flush=True)

    ds = G.PairDataset(a.res, a.capacity)
    bounds, its, t_gen = generate(ds, a.res, a.ppc, a.solves, a.krylov_every, a.tol, 20000 + 1000 * a.seed, log)
    n = len(ds)
    nval_solves = max(1, int(round(a.val_frac * a.solves)))
    split = bounds[a.solves - nval_solves]
    net = G.LearnedPreconditioner(seed=a.seed)
    l_train0 = net.evaluate(ds, 0, min(split, 512))
    l_val0 = net.evaluate(ds, split, n - split)
    t = time.perf_counter()
    hist = []
    chunk = 250
    done = 0
    while done < a.steps:
        k = min(chunk, a.steps - done)
        # one cosine schedule over all steps: pass the global position through warmup/count
        h = net.train(ds, steps=k, batch=a.batch, lr=a.lr * _cos(done, a.steps), lr_final=a.lr * _cos(done + k, a.steps),
                      warmup=50 if done == 0 else 0, count=split, seed=a.seed * 7777 + done)
        hist += list(map(float, h))
        done += k
        log(f"# train step {done}/{a.steps}: batch loss {np.mean(h[-50:]):.4f} ({time.perf_counter() - t:.0f}s)")
    t_train = time.perf_counter() - t
    l_train1 = net.evaluate(ds, 0, min(split, 512))
    l_val1 = net.evaluate(ds, split, n - split)
    P = net.get_params()
    np.save(a.out, P)
    summary = dict(res=a.res, ppc=a.ppc, solves=a.solves, pairs=n, val_pairs=n - split,
                   krylov_every=a.krylov_every, uzawa_its_mean=float(np.mean(its)), gen_seconds=t_gen,
                   steps=a.steps, batch=a.batch, lr=a.lr, train_seconds=t_train,
                   loss_train=[l_train0, l_train1], loss_val=[l_val0, l_val1], params=int(P.size),
                   loss_curve=[float(np.mean(hist[i:i + 50])) for i in range(0, len(hist), 50)])
    with open(os.path.splitext(a.out)[0] + ".json", "w") as f:
        json.dump(summary, f, indent=1)
    print(json.dumps({k: v for k, v in summary.items() if k != "loss_curve"}), flush=True)


def _cos(step, total):
    return 0.5 * (1.0 + np.cos(np.pi * min(step, total) / max(total, 1))) * 0.99 + 0.01


if __name__ == "__main__":
    main()
