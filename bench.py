"""Benchmark of the MI355X Stokes hot path.

One step = one Stokes solve (Uzawa to 1e-6 relative residual, reference
settings) of the 3D random-viscosity workload (96^3 Q1-P0 elements, 16
particles per cell, log10 eta in [-3, 3]) with all inputs resident in HBM.
Prints one JSON line (see the "metric" field).  Extra objects:

  roofline      the dominant kernel (the FP64 matrix-free K apply) timed with
                HIP events on the solver's stream: algorithmic bytes per apply
                / measured apply time, against the 8 TB/s HBM peak.
  cpu_baseline  the CPU oracle (assembled CSR + reference Uzawa with the
                PETSc-default GMRES(30)+ILU(0) inner solve) on a bounded
                sample of the same generator, timed on the host.
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

METRIC = "Stokes solve time at 1e-6 residual (learned-PC vs plain Uzawa); HBM GB/s %peak"
HBM_PEAK_GBS = 8000.0


def k_apply_bytes(dim, N, E, P, b=8):
    """Algorithmic HBM bytes of one particle-form K apply: read u and write y
    (dim*N values each), the particle stream (xi[dim] and omega = w detJ eta,
    padded to 4 values), the element offsets (int32)."""
    return 2 * b * dim * N + b * 4 * P + 4 * (E + 1)


def k_moments_bytes(dim, N, E, b=8):
    """Algorithmic HBM bytes of one moment-form K apply: read u and write y,
    and 27 (3D) / 9 (2D) moments per element."""
    return 2 * b * dim * N + b * (27 if dim == 3 else 9) * E


def cpu_baseline(res, threads):
    import oracle as O
    from gale_amd import configs
    cfg = configs.random_viscosity_3d(res=res, ppc=16)
    pb = O.Problem(**cfg.oracle_kwargs())
    up = O.uzawa_params(tolerance=1e-6, inner=1, nthreads=threads, inner_max_iterations=100000)
    t0 = time.perf_counter()
    st = pb.assemble(nthreads=threads)
    t_asm = time.perf_counter() - t0
    st, u, p, rep = pb.uzawa(up)
    return dict(status=int(st), solve_s=rep.solve_seconds, assemble_s=t_asm, uzawa_iterations=rep.uzawa_iterations,
                inner_iterations=int(rep.inner_iterations), E=cfg.E, P=cfg.P)


def bench_dvc(G, pb, cfg, sample_cells=2000, res=10, population_control=False):
    """Discrete Voronoi cell weights (DVCWeights, resolution 10) of every
    element of the workload: GPU time of a relocation with DVC weights minus
    one with constant weights (the DVC growth kernel), and the oracle's
    restatement on a sample of the same elements on one host core."""
    import oracle as O
    _, owner, xi, _ = pb.particles()
    rng = np.random.default_rng(3)
    cells = rng.choice(cfg.E, size=min(sample_cells, cfg.E), replace=False)
    order = np.argsort(owner, kind="stable")
    starts = np.searchsorted(owner[order], np.arange(cfg.E + 1))
    t = time.perf_counter()
    for e in cells:
        O.dvc_cell(cfg.dim, (res,) * cfg.dim, xi[order[starts[e]:starts[e + 1]]])
    cpu_s = time.perf_counter() - t
    times = {}
    for rep in range(3):
        for kind in (G.WEIGHTS_CONSTANT, G.WEIGHTS_DVC):
            pb.set_weights(G.WEIGHTS_CONSTANT)  # same starting order for both
            t = time.perf_counter()
            pb.set_weights(kind, res)
            times.setdefault(kind, []).append(time.perf_counter() - t)
    pb.set_weights(G.WEIGHTS_CONSTANT)
    gpu_ms = 1e3 * (min(times[G.WEIGHTS_DVC]) - min(times[G.WEIGHTS_CONSTANT]))
    pc = None
    if population_control:
        # one PCDVC pass with the reference defaults (changes the particle set)
        t = time.perf_counter()
        added, deleted = pb.population_control(G.pcdvc_params(resolution=res))
        pc = {"ms": 1e3 * (time.perf_counter() - t), "added": added, "deleted": deleted, "particles_after": pb.P,
              "params": "upperT 25, lowerT 0.6, maxSplits 3, maxDeletions 2, resolution 10"}
        print(f"# population control: {pc['ms']:.1f} ms, +{added} -{deleted} particles", file=sys.stderr, flush=True)
    cpu_ms_all = 1e3 * cpu_s / len(cells) * cfg.E
    out = {"elements": cfg.E, "particles": cfg.P, "resolution": res, "gpu_ms": gpu_ms,
           "gpu_elements_per_s": cfg.E / (gpu_ms * 1e-3) if gpu_ms > 0 else None,
           "relocate_constant_ms": 1e3 * min(times[G.WEIGHTS_CONSTANT]),
           "cpu_oracle_ms_projected": cpu_ms_all, "cpu_sample_elements": int(len(cells)), "cpu_cores": 1,
           "population_control": pc}
    print(f"# dvc weights: GPU {gpu_ms:.2f} ms for {cfg.E} elements, oracle (1 core) {cpu_ms_all:.0f} ms projected",
          file=sys.stderr, flush=True)
    return out


def bench_batch(G, configs, count):
    """Batched workload: count independent 2D sinkers (meshes 16^2..128^2 in
    steps of 8, 1-5 discs with log10 viscosity contrast in [-2, 4], 20 ppc),
    FP32 multigrid applies inside FP64 inner CG and FP64 Uzawa, each to 1e-6,
    in one persistent launch (eager and HIP-graph replay).  For scale, every
    16th instance is also solved alone with the single-problem solver."""
    t = time.perf_counter()
    cfgs = configs.sinker_batch(count=count)
    t_gen = time.perf_counter() - t
    t = time.perf_counter()
    bs = G.BatchSolver([c.problem_kwargs() for c in cfgs])
    t_setup = time.perf_counter() - t
    p = G.batch_params(tolerance=1e-6)
    best, ms = bs.autotune(p)
    st, res, sec = bs.solve(p)
    stg, _, _ = bs.solve(G.batch_params(tolerance=1e-6, use_graph=True))  # capture
    stg, resg, secg = bs.solve(G.batch_params(tolerance=1e-6, use_graph=True))  # replay
    outer = [r["uzawa_iterations"] for r in res]
    inner = [r["inner_iterations"] for r in res]
    step = max(1, count // 16)
    seq = 0.0
    for i in range(0, count, step):
        pr = G.Problem(**cfgs[i].problem_kwargs())
        sv = G.Stokes_SLE_UzawaSolver(pr, G.uzawa_params(tolerance=1e-6, maxIterations=20000))
        _, _, rep = sv.Solve(raise_on_error=False)
        seq += rep.solve_seconds
        sv.close()
        pr.close()
    bs.close()
    print(f"# batch {count}: {sec:.3f}s eager, {secg:.3f}s graph, tile {best}", file=sys.stderr, flush=True)
    return {"workload": f"{count} independent 2D sinkers, meshes 16^2..128^2 (multiples of 8), 1-5 discs, "
                        f"log10 eta contrast [-2,4], 20 ppc, free slip, Uzawa to 1e-6",
            "instances": count, "particles": int(sum(c.P for c in cfgs)),
            "elements": int(sum(c.E for c in cfgs)), "generate_s": t_gen, "setup_s": t_setup,
            "tile_autotuned": list(best), "tile_ms": {f"{b}x{t}": v for (b, t), v in ms.items()},
            "solve_s": sec, "graph_solve_s": secg, "status": int(st), "graph_status": int(stg),
            "all_converged": all(r["status"] == 0 and r["rel_residual"] <= 1e-6 for r in res),
            "uzawa_iterations_mean": float(np.mean(outer)), "uzawa_iterations_max": int(max(outer)),
            "inner_iterations_mean": float(np.mean(inner)),
            "single_problem_sequential_estimate_s": seq * step,
            "single_problem_sample": f"every {step}th instance solved alone, x{step}"}


def profile_apply(args):
    """Calibration copy of known size (torch elementwise copy, 256 MiB read +
    256 MiB written per call) followed by the four K apply forms, so that
    rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE per dispatch can be scaled to
    bytes.  torch is imported before the library so both share one HIP runtime."""
    import torch
    n = 32 * 1024 * 1024
    a = torch.ones(n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    import gale_amd as G
    from gale_amd import configs
    cfg = configs.random_viscosity_3d(res=args.res, ppc=args.ppc)
    pb = G.Problem(**cfg.problem_kwargs())
    s = G.Stokes_SLE_UzawaSolver(pb, G.uzawa_params(tolerance=1e-6))
    s.setup()
    for form in (G.APPLY_MOMENTS_FP64, G.APPLY_MOMENTS_FP32, G.APPLY_PARTICLES_FP64, G.APPLY_PARTICLES_FP32):
        ms = s.time_apply_K(form, 3)
        print(f"# form {form}: {ms:.4f} ms", file=sys.stderr, flush=True)
    print(json.dumps({"copy_bytes_each_way": n * 8, "res": args.res, "ppc": args.ppc,
                      "moments_bytes": k_moments_bytes(cfg.dim, cfg.N, cfg.E),
                      "particles_bytes": k_apply_bytes(cfg.dim, cfg.N, cfg.E, cfg.P)}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=0)
    ap.add_argument("--res", type=int, default=96)
    ap.add_argument("--ppc", type=int, default=16)
    ap.add_argument("--apply-reps", type=int, default=20)
    ap.add_argument("--cpu-res", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--max-outer", type=int, default=20000)
    ap.add_argument("--plain-max-outer", type=int, default=0,
                    help="cap of the plain solve only (profiling the learned solve); 0 = --max-outer")
    ap.add_argument("--mg-smooth", type=int, default=2)
    ap.add_argument("--weights", default=os.path.join(ROOT, "gale_amd", "lpc_weights.npy"))
    ap.add_argument("--no-learned", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--batch-count", type=int, default=256)
    ap.add_argument("--no-dvc", action="store_true", help="skip the DVC weights measurement")
    ap.add_argument("--only-dvc", action="store_true", help="measure only the DVC weights of the workload")
    ap.add_argument("--inner-relax", type=float, default=0.1,
                    help="learned solve: cap of the relaxed inner tolerance (0 = reference 0.1 x tol throughout)")
    ap.add_argument("--profile-apply", action="store_true",
                    help="only run a calibration copy and the K applies (for rocprofv3 --pmc traffic counts)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU only; --gpus {args.gpus} is not supported "
              "(the multi-GPU sharding is a design only, not built).", file=sys.stderr)
        sys.exit(2)

    if args.profile_apply:
        profile_apply(args)
        return

    import gale_amd as G
    from gale_amd import configs

    t = time.perf_counter()
    cfg = configs.random_viscosity_3d(res=args.res, ppc=args.ppc)
    t_gen = time.perf_counter() - t
    print(f"# generated {cfg.name} res={args.res} E={cfg.E} P={cfg.P} in {t_gen:.1f}s", file=sys.stderr, flush=True)
    pb = G.Problem(**cfg.problem_kwargs())
    if args.only_dvc:
        print(json.dumps({"dvc_weights": bench_dvc(G, pb, cfg, population_control=True)}), flush=True)
        return
    params = G.uzawa_params(tolerance=1e-6, maxIterations=args.plain_max_outer or args.max_outer, inner_max_iterations=20000,
                            mg_smooth=args.mg_smooth, monitor=100)
    s = G.Stokes_SLE_UzawaSolver(pb, params)
    s.setup()  # first linear solve of the Picard sequence: no yielding yet

    # roofline: the solve's dominant kernel is the FP64 moment-form K apply
    # (every inner FCG iteration and every fine-level smoothing step); the
    # particle-form apply (the per-material-point integration) is reported
    # beside it.  Both timed with HIP events on the solver's stream.
    ms_m = s.time_apply_K(G.APPLY_MOMENTS_FP64, args.apply_reps)
    bytes_m = k_moments_bytes(cfg.dim, cfg.N, cfg.E)
    ms_m32 = s.time_apply_K(G.APPLY_MOMENTS_FP32, args.apply_reps)
    bytes_m32 = k_moments_bytes(cfg.dim, cfg.N, cfg.E, b=4)
    ms = s.time_apply_K(G.APPLY_PARTICLES_FP64, args.apply_reps)
    nbytes = k_apply_bytes(cfg.dim, cfg.N, cfg.E, cfg.P)
    ms32 = s.time_apply_K(G.APPLY_PARTICLES_FP32, args.apply_reps)
    nbytes32 = k_apply_bytes(cfg.dim, cfg.N, cfg.E, cfg.P, b=4)
    # the pressure operators of every Uzawa iteration: G p reads p (8 B per
    # element) and the node flags (1 B per node) and writes dim values per node;
    # G^T u reads dim values per node and writes one per element
    ms_g = s.time_apply_K(G.APPLY_G, args.apply_reps)
    bytes_g = 8 * cfg.E + 8 * cfg.dim * cfg.N + cfg.N
    ms_gt = s.time_apply_K(G.APPLY_GT, args.apply_reps)
    bytes_gt = 8 * cfg.dim * cfg.N + 8 * cfg.E
    achieved = bytes_m / (ms_m * 1e-3) / 1e9
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "k_apply_traffic.json")
    if os.path.exists(tpath):
        try:
            tr = json.load(open(tpath))
            if tr.get("res") == args.res and tr.get("ppc") == args.ppc:
                traffic = tr.get("bytes_per_apply")
        except Exception:
            traffic = None
    print(f"# K apply moments fp64 {ms_m:.4f} ms ({achieved:.0f} GB/s), fp32 {ms_m32:.4f} ms; particles fp64 "
          f"{ms:.4f} ms ({nbytes / ms / 1e6:.0f} GB/s), fp32 {ms32:.4f} ms", file=sys.stderr, flush=True)

    for w in range(args.warmup):
        s.set_pressure(None)
        u, p, rep = s.Solve()
        print(f"# warmup {w}: {rep.solve_seconds:.3f}s uzawa={rep.uzawa_iterations} inner={rep.inner_iterations}",
              file=sys.stderr, flush=True)
    times, reps = [], []
    s.sync()
    t0 = time.perf_counter()
    for k in range(args.steps):
        s.set_pressure(None)
        u, p, rep = s.Solve(raise_on_error=False)
        times.append(rep.solve_seconds)
        reps.append(rep.as_dict())
        print(f"# step {k}: {rep.solve_seconds:.3f}s uzawa={rep.uzawa_iterations} inner={rep.inner_iterations} "
              f"rel={rep.rel_residual:.3e}", file=sys.stderr, flush=True)
    s.sync()
    wall = time.perf_counter() - t0
    solve_s = float(np.mean(times))
    last = reps[-1]
    n_u = pb.N * cfg.dim
    dofs = n_u + cfg.E

    # learned Schur preconditioner + mixed precision (FP32 V-cycle inside the
    # FP64 flexible inner CG), trained offline on 32^3 fields of the same
    # generator (gale_amd/lpc_train.py); the plain solve above is the
    # check every learned solve must meet.
    learned = None
    wpath = args.weights
    if not args.no_learned and os.path.exists(wpath):
        net = G.LearnedPreconditioner(seed=1)
        net.set_params(np.load(wpath))
        meta = {}
        mpath = os.path.splitext(wpath)[0] + ".json"
        if os.path.exists(mpath):
            meta = {k: v for k, v in json.load(open(mpath)).items() if k != "loss_curve"}
        sl = G.Stokes_SLE_UzawaSolver(pb, G.uzawa_params(tolerance=1e-6, maxIterations=args.max_outer,
                                                         inner_max_iterations=20000, mg_smooth=args.mg_smooth,
                                                         monitor=100, learned=True, precision=G.PREC_MIXED,
                                                         inner_relax_max=args.inner_relax))
        sl.set_learned(net)
        for w in range(args.warmup):
            sl.set_pressure(None)
            _, _, rl = sl.Solve(raise_on_error=False)
            print(f"# learned warmup {w}: {rl.solve_seconds:.3f}s uzawa={rl.uzawa_iterations}",
                  file=sys.stderr, flush=True)
        ltimes = []
        sl.sync()
        t0 = time.perf_counter()
        for k in range(args.steps):
            sl.set_pressure(None)
            ul, pl, rl = sl.Solve(raise_on_error=False)
            ltimes.append(rl.solve_seconds)
            print(f"# learned step {k}: {rl.solve_seconds:.3f}s uzawa={rl.uzawa_iterations} "
                  f"inner={rl.inner_iterations} fallback={rl.learned_fallback} rel={rl.rel_residual:.3e}",
                  file=sys.stderr, flush=True)
        sl.sync()
        lwall = time.perf_counter() - t0
        pm, plm = p - p.mean(), pl - pl.mean()
        learned = {"solve_s": float(np.mean(ltimes)), "status": int(rl.status),
                   "uzawa_iterations": int(rl.uzawa_iterations), "inner_iterations": int(rl.inner_iterations),
                   "fallback_at": int(rl.learned_fallback), "rel_residual_true": float(rl.rel_residual),
                   "verify_rounds": int(rl.verify_rounds), "inner_relax_max": args.inner_relax,
                   "u_rel_diff_vs_plain": float(np.linalg.norm(ul - u) / np.linalg.norm(u)),
                   "p_rel_diff_vs_plain": float(np.linalg.norm(plm - pm) / np.linalg.norm(pm)),
                   "speedup_vs_plain": solve_s / float(np.mean(ltimes)), "wall_s": lwall, "training": meta}
        sl.close()

    batch = None
    if not args.no_batch:
        batch = bench_batch(G, configs, args.batch_count)
    dvc = None if args.no_dvc else bench_dvc(G, pb, cfg)

    cpu = None
    if not args.no_cpu:
        threads = min(16, os.cpu_count() or 1)
        t = time.perf_counter()
        cb = cpu_baseline(args.cpu_res, threads)
        print(f"# cpu baseline {cb} in {time.perf_counter() - t:.1f}s", file=sys.stderr, flush=True)
        # the GPU on the same sample, for a like-for-like ratio
        cfg_s = configs.random_viscosity_3d(res=args.cpu_res, ppc=16)
        ps = G.Problem(**cfg_s.problem_kwargs())
        ss = G.Stokes_SLE_UzawaSolver(ps, G.uzawa_params(tolerance=1e-6, inner_max_iterations=20000))
        _, _, rs = ss.Solve()
        cpu = {"value": cb["solve_s"], "unit": "s", "cores": threads, "kind": "port",
               "sample": f"random_viscosity_3d at {args.cpu_res}^3 ({cb['E']} elements, {cb['P']} particles): "
                         f"oracle assembled CSR + reference Uzawa, GMRES(30)+ILU(0) inner (PETSc defaults), "
                         f"{cb['uzawa_iterations']} Uzawa its; assembly {cb['assemble_s']:.2f}s not included",
               "gpu_same_sample_s": rs.solve_seconds,
               "status": cb["status"]}
    ok_learned = learned is not None and learned["status"] == 0
    line = {
        "metric": METRIC,
        "value": learned["solve_s"] if ok_learned else solve_s,
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": (learned["wall_s"] if ok_learned else wall) / max(args.steps, 1) * 1e3,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"random_viscosity_3d {args.res}^3 Q1-P0, {args.ppc} ppc, log10 eta GRF [-3,3] "
                               f"+ plastic lid (first, linear solve), free slip, Uzawa to 1e-6",
                   "value_is": ("learned Schur preconditioner + mixed precision (FP32 V-cycle in FP64 flexible "
                                "inner CG) + relaxed inner tolerance with a true-residual check"
                                if ok_learned else "plain FP64 Uzawa (diagonal Q)"),
                   "plain_fp64_solve_s": solve_s,
                   "plain": "diagonal Q, inner FP64 flexible CG + multigrid V-cycle at 0.1 x tol",
                   "learned": learned,
                   "elements": cfg.E, "particles": cfg.P, "dofs": dofs,
                   "uzawa_iterations": last["uzawa_iterations"], "inner_iterations": last["inner_iterations"],
                   "rel_residual": last["rel_residual"], "dof_per_s": dofs / solve_s,
                   "setup_s": last["setup_seconds"], "k_applies": last["k_applies"],
                   "k_moments_fp32_ms": ms_m32, "k_moments_fp32_gbs": bytes_m32 / (ms_m32 * 1e-3) / 1e9},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": "gs::k_tile_apply3<double,0> + gs::k_tile_halo3<double> (moment-form K apply, fine level)",
                     "ms_per_apply": ms_m, "algorithmic_bytes": bytes_m},
        "roofline_particle_apply": {"bound": "hbm", "achieved": nbytes / (ms * 1e-3) / 1e9, "peak": HBM_PEAK_GBS,
                                    "unit": "GB/s", "frac": nbytes / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                                    "kernel": "gs::k_tile_apply3<double,1> + gs::k_tile_halo3<double> (particle-form K apply)",
                                    "ms_per_apply": ms, "algorithmic_bytes": nbytes,
                                    "fp32_ms_per_apply": ms32, "fp32_algorithmic_bytes": nbytes32},
        "roofline_G": {"bound": "hbm", "achieved": bytes_g / (ms_g * 1e-3) / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                       "frac": bytes_g / (ms_g * 1e-3) / 1e9 / HBM_PEAK_GBS, "kernel": "gs::k_apply_G (G p)",
                       "ms_per_apply": ms_g, "algorithmic_bytes": bytes_g},
        "roofline_Gt": {"bound": "hbm", "achieved": bytes_gt / (ms_gt * 1e-3) / 1e9, "peak": HBM_PEAK_GBS,
                        "unit": "GB/s", "frac": bytes_gt / (ms_gt * 1e-3) / 1e9 / HBM_PEAK_GBS,
                        "kernel": "gs::k_apply_Gt (G^T u)", "ms_per_apply": ms_gt, "algorithmic_bytes": bytes_gt},
        "cpu_baseline": cpu,
        "batch": batch,
        "dvc_weights": dvc,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
