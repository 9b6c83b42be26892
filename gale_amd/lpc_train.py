"""Train the learned Schur preconditioner on the GPU.

Phase 1: plain FP64 Uzawa solves of random_viscosity_3d fields (the generator
of the 96^3 workload, other seeds) at the training resolution record
(r0, p* - p0) and Krylov pairs; the network is trained on them by the
library's own BF16 forward / backward kernels and Adam.
Model selection: with --candidates K, K networks (different initialisations
and batch orders) are trained on the same pairs and each is timed on held-out
learned solves at --select-res (fields of other seeds); the fastest is kept.
FP32 parameters are saved as .npy with a JSON summary beside them.

    python -m gale_amd.lpc_train --res 32 --ppc 8 --solves 50 --steps 3000 \
        --candidates 4 --select-res 64 --out gale_amd/lpc_weights.npy
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


def log(m):
    print(m, file=sys.stderr, flush=True)


def record_solves(ds, res, ppc, solves, krylov_every, tol, seed0, net=None, tag=""):
    """Uzawa solves recording pairs into ds (plain FP64 when net is None,
    else learned + mixed precision).  Returns (pair index after each solve,
    Uzawa iteration counts, seconds)."""
    t0 = time.perf_counter()
    bounds, its = [len(ds)], []
    for i in range(solves):
        cfg = configs.random_viscosity_3d(res=res, ppc=ppc, seed=seed0 + i, field_seed=seed0 + 7919 + i)
        pb = G.Problem(**cfg.problem_kwargs())
        if net is None:
            prm = G.uzawa_params(tolerance=tol, maxIterations=20000)
        else:
            prm = G.uzawa_params(tolerance=tol, maxIterations=20000, learned=True, precision=G.PREC_MIXED)
        s = G.Stokes_SLE_UzawaSolver(pb, prm)
        if net is not None:
            s.set_learned(net)
        s.record_pairs(ds, krylov_every)
        _, _, rep = s.Solve(raise_on_error=False)
        s.record_pairs(None)
        s.close()
        pb.close()
        its.append(int(rep.uzawa_iterations))
        bounds.append(len(ds))
        if (i + 1) % 10 == 0 or i == solves - 1:
            log(f"# {tag} solve {i + 1}/{solves} at {res}^3: status {rep.status}, {rep.uzawa_iterations} Uzawa its, "
                f"{rep.solve_seconds:.2f}s; pairs {len(ds)}; {time.perf_counter() - t0:.0f}s")
    return bounds, its, time.perf_counter() - t0


def _cos(step, total):
    return 0.5 * (1.0 + np.cos(np.pi * min(step, total) / max(total, 1))) * 0.99 + 0.01


def train(net, sets, steps, lr, seed, chunk=250):
    """Cosine-scheduled Adam over `steps`, cycling through (dataset, batch,
    first, count) entries chunk by chunk."""
    hist, done, k = [], 0, 0
    t = time.perf_counter()
    while done < steps:
        ds, batch, first, count = sets[k % len(sets)]
        n = min(chunk, steps - done)
        h = net.train(ds, steps=n, batch=batch, lr=lr * _cos(done, steps), lr_final=lr * _cos(done + n, steps),
                      warmup=50 if done == 0 else 0, first=first, count=count, seed=seed * 7777 + done)
        hist += list(map(float, h))
        done += n
        k += 1
        log(f"# train step {done}/{steps} ({ds.shape[0]}^3 batch {batch}): loss {np.mean(h[-50:]):.4f} "
            f"({time.perf_counter() - t:.0f}s)")
    return hist, time.perf_counter() - t


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--res", type=int, default=32)
    ap.add_argument("--ppc", type=int, default=8)
    ap.add_argument("--solves", type=int, default=50)
    ap.add_argument("--krylov-every", type=int, default=8)
    ap.add_argument("--tol", type=float, default=1e-6)
    ap.add_argument("--capacity", type=int, default=16000)
    ap.add_argument("--steps", type=int, default=3000)
    ap.add_argument("--batch", type=int, default=8)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--val-frac", type=float, default=0.1)
    ap.add_argument("--candidates", type=int, default=1, help="networks trained; the fastest on held-out solves is kept")
    ap.add_argument("--select-res", type=int, default=64)
    ap.add_argument("--select-solves", type=int, default=1)
    ap.add_argument("--select-ppc", type=int, default=16)
    ap.add_argument("--out", default=os.path.join(HERE, "lpc_weights.npy"))
    a = ap.parse_args(argv)

    seed0 = 20000 + 1000 * a.seed
    ds = G.PairDataset(a.res, a.capacity)
    bounds, its, t_gen = record_solves(ds, a.res, a.ppc, a.solves, a.krylov_every, a.tol, seed0, tag="plain")
    nval = max(1, int(round(a.val_frac * a.solves)))
    split = bounds[a.solves - nval]
    n1 = len(ds)
    net = G.LearnedPreconditioner(seed=a.seed)
    l_tr0, l_val0 = net.evaluate(ds, 0, min(split, 512)), net.evaluate(ds, split, n1 - split)
    hist, t_train = train(net, [(ds, a.batch, 0, split)], a.steps, a.lr, a.seed)
    l_tr1, l_val1 = net.evaluate(ds, 0, min(split, 512)), net.evaluate(ds, split, n1 - split)
    summary = dict(res=a.res, ppc=a.ppc, solves=a.solves, pairs=n1, val_pairs=n1 - split,
                   krylov_every=a.krylov_every, uzawa_its_mean=float(np.mean(its)), gen_seconds=t_gen,
                   steps=a.steps, batch=a.batch, lr=a.lr, train_seconds=t_train,
                   loss_train=[l_tr0, l_tr1], loss_val=[l_val0, l_val1], params=int(net.num_params()))
    if a.candidates > 1:
        # model selection on solver performance: K networks (different
        # initialisations and batch orders) on the same pairs, each timed on
        # held-out solves at --select-res; the fastest is kept
        cands = [(net, hist, t_train, l_val1)]
        for k in range(1, a.candidates):
            nk = G.LearnedPreconditioner(seed=a.seed + 100 * k)
            hk, tk = train(nk, [(ds, a.batch, 0, split)], a.steps, a.lr, a.seed + 100 * k)
            cands.append((nk, hk, tk, nk.evaluate(ds, split, n1 - split)))
        sel = []
        for i in range(a.select_solves):
            cfg = configs.random_viscosity_3d(res=a.select_res, ppc=a.select_ppc, seed=seed0 + 90000 + i,
                                              field_seed=seed0 + 97919 + i)
            pb = G.Problem(**cfg.problem_kwargs())
            times = []
            for (nk, _, _, _) in cands:
                s = G.Stokes_SLE_UzawaSolver(pb, G.uzawa_params(tolerance=a.tol, maxIterations=20000, learned=True,
                                                                precision=G.PREC_MIXED))
                s.set_learned(nk)
                _, _, rep = s.Solve(raise_on_error=False)
                s.close()
                times.append(rep.solve_seconds if rep.status == 0 else float("inf"))
            pb.close()
            sel.append(times)
            log(f"# selection solve {i + 1}/{a.select_solves} at {a.select_res}^3: " +
                ", ".join(f"{t:.2f}s" for t in times))
        score = np.sum(np.asarray(sel), axis=0)
        best = int(np.argmin(score))
        net, hist, t_train, _ = cands[best]
        summary.update(candidates=a.candidates, select_res=a.select_res, select_solves=a.select_solves,
                       select_ppc=a.select_ppc,
                       select_seconds=[float(v) for v in score], selected=best,
                       candidate_loss_val=[float(c[3]) for c in cands],
                       train_seconds_all=float(sum(c[2] for c in cands)))
        summary["loss_val"] = [l_val0, float(cands[best][3])]
    P = net.get_params()
    np.save(a.out, P)
    summary["loss_curve"] = [float(np.mean(hist[i:i + 50])) for i in range(0, len(hist), 50)]
    with open(os.path.splitext(a.out)[0] + ".json", "w") as f:
        json.dump(summary, f, indent=1)
    print(json.dumps({k: v for k, v in summary.items() if k != "loss_curve"}), flush=True)


if __name__ == "__main__":
    main()
