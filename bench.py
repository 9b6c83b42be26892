"""Benchmark of the MI355X Stokes hot path.

One step = one Stokes solve (Uzawa to 1e-6 relative residual, reference
settings) of the 3D random-viscosity workload (96^3 Q1-P0 elements, 16
particles per cell, log10 eta in [-3, 3]) with all inputs resident in HBM.
--steps / --warmup apply to the headline solve (learned Schur preconditioner
+ mixed precision); the plain FP64 Uzawa it is compared with runs once.
Prints one JSON line (see the "metric" field) as soon as the headline and the
roofline exist, and prints it again, extended, after each later leg (plain
solve, ablation, yielded solve, CPU baseline, batched solves, DVC weights);
the last line printed is the complete one.  A wall-clock budget (--budget)
skips the extra legs that would not fit and records why.  Extra objects:

  roofline      the dominant kernel (the FP64 matrix-free K apply) timed with
                HIP events on the solver's stream, cold (operand copies cycled
                past the last-level cache): algorithmic bytes per apply /
                measured apply time, against the 8 TB/s HBM peak.
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


def bench_dvc(G, pb, cfg, sample_cells=2000, res=10, population_control=False, ppc=16):
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
    pc = pci = None
    if population_control:
        # one PCDVC pass with the reference defaults (changes the particle set)
        t = time.perf_counter()
        added, deleted = pb.population_control(G.pcdvc_params(resolution=res))
        pc = {"ms": 1e3 * (time.perf_counter() - t), "added": added, "deleted": deleted, "particles_after": pb.P,
              "params": "upperT 25, lowerT 0.6, maxSplits 3, maxDeletions 2, resolution 10"}
        print(f"# population control: {pc['ms']:.1f} ms, +{added} -{deleted} particles", file=sys.stderr, flush=True)
        # then a pass with the Inflow branch on (defaults: Threshold 0.8,
        # CentPosRatio 0.01, particlesPerCell = the layout's), and the oracle
        # on a 16^3 field of the same generator for a per-cell CPU rate
        t = time.perf_counter()
        added, deleted = pb.population_control(G.pcdvc_params(resolution=res, Inflow=True,
                                                              particlesPerCell=ppc))
        pci = {"ms": 1e3 * (time.perf_counter() - t), "added": added, "deleted": deleted, "particles_after": pb.P,
               "rand_calls": pb.rand_draws, "repopulated_cells": pb.repopulated_cells,
               "repopulation_launches": pb.repopulation_launches,
               "params": f"defaults + Inflow, particlesPerCell {ppc}"}
        from gale_amd import configs as _configs
        sc = _configs.random_viscosity_3d(res=16, ppc=ppc)
        o = O.Problem(**sc.oracle_kwargs())
        t = time.perf_counter()
        o.population_control(Inflow=True, particlesPerCell=ppc)
        cpu_cell_ms = 1e3 * (time.perf_counter() - t) / sc.E
        pci.update(cpu_oracle_ms_per_cell=cpu_cell_ms, cpu_oracle_ms_projected=cpu_cell_ms * cfg.E,
                   cpu_sample="oracle, 16^3 field of the same generator, 1 core")
        print(f"# population control with Inflow: {pci['ms']:.1f} ms, +{added} -{deleted} particles, "
              f"{pci['repopulated_cells']} re-populated cells in {pci['repopulation_launches']} launches; "
              f"oracle {cpu_cell_ms * cfg.E:.0f} ms projected", file=sys.stderr, flush=True)
    cpu_ms_all = 1e3 * cpu_s / len(cells) * cfg.E
    out = {"elements": cfg.E, "particles": cfg.P, "resolution": res, "gpu_ms": gpu_ms,
           "gpu_elements_per_s": cfg.E / (gpu_ms * 1e-3) if gpu_ms > 0 else None,
           "relocate_constant_ms": 1e3 * min(times[G.WEIGHTS_CONSTANT]),
           "cpu_oracle_ms_projected": cpu_ms_all, "cpu_sample_elements": int(len(cells)), "cpu_cores": 1,
           "population_control": pc, "population_control_inflow": pci}
    print(f"# dvc weights: GPU {gpu_ms:.2f} ms for {cfg.E} elements, oracle (1 core) {cpu_ms_all:.0f} ms projected",
          file=sys.stderr, flush=True)
    return out


def bench_batch(G, configs, count):
    """Batched workload: count independent 2D sinkers (meshes 16^2..128^2 in
    steps of 8, 1-5 discs with log10 viscosity contrast in [-2, 4], 20 ppc),
    FP32 multigrid applies inside FP64 inner CG and FP64 Uzawa, each to 1e-6,
    in one persistent launch: the reference Uzawa loop (eager and HIP-graph
    replay) and the inexact loop (relaxed inner tolerance, flexible outer
    iteration, true-residual check).  For scale, every 16th instance is also
    solved alone with the single-problem solver."""
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
    pr_ = G.batch_params(tolerance=1e-6, inner_relax_max=0.03, flexible_window=8)
    str_, resr, secr = bs.solve(pr_, raise_on_error=False)

    def stats(rs):
        t_i = np.array([r["seconds"] for r in rs])
        return {"instance_s_max": float(t_i.max()), "instance_s_mean": float(t_i.mean()),
                "tail_over_mean": float(t_i.max() / t_i.mean()),
                "uzawa_iterations_mean": float(np.mean([r["uzawa_iterations"] for r in rs])),
                "uzawa_iterations_max": int(max(r["uzawa_iterations"] for r in rs)),
                "inner_iterations_mean": float(np.mean([r["inner_iterations"] for r in rs])),
                "all_converged": all(r["status"] == 0 and r["rel_residual"] <= 1e-6 for r in rs)}
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
    print(f"# batch {count}: reference loop {sec:.3f}s eager, {secg:.3f}s graph; relaxed loop {secr:.3f}s; "
          f"tile {best}", file=sys.stderr, flush=True)
    out = {"workload": f"{count} independent 2D sinkers, meshes 16^2..128^2 (multiples of 8), 1-5 discs, "
                       f"log10 eta contrast [-2,4], 20 ppc, free slip, Uzawa to 1e-6",
           "instances": count, "particles": int(sum(c.P for c in cfgs)),
           "elements": int(sum(c.E for c in cfgs)), "generate_s": t_gen, "setup_s": t_setup,
           "tile_autotuned": list(best), "tile_ms": {f"{b}x{t}": v for (b, t), v in ms.items()},
           "solve_s": sec, "graph_solve_s": secg, "status": int(st), "graph_status": int(stg),
           "reference_loop": stats(res),
           "relaxed_solve_s": secr, "relaxed_status": int(str_),
           "relaxed_loop": dict(stats(resr), params="inner rtol relaxed up to 0.03, flexible outer CG window 8, "
                                                    "true-residual check"),
           "single_problem_sequential_estimate_s": seq * step,
           "single_problem_sample": f"every {step}th instance solved alone, x{step}"}
    return out


def profile_apply(args):
    """Calibration copy of known size (torch elementwise copy, 1 GiB read +
    1 GiB written) followed by the four K apply forms and the two
    pressure operators, each timed cold (operand copies cycled so every apply
    streams from HBM), so that rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE per
    dispatch can be scaled to bytes and the kernel-trace averages compare with
    the bench's roofline.  torch is imported before the library so both share
    one HIP runtime."""
    import torch
    # calibration: one elementwise copy of 1 GiB (4x the last-level cache, so
    # every byte moves to and from HBM): FETCH_SIZE and WRITE_SIZE of that one
    # dispatch against 1 GiB read and 1 GiB written
    n = 128 * 1024 * 1024
    a = torch.ones(n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    b.copy_(a)
    torch.cuda.synchronize()
    del a, b
    import gale_amd as G
    from gale_amd import configs
    cfg = configs.random_viscosity_3d(res=args.res, ppc=args.ppc)
    pb = G.Problem(**cfg.problem_kwargs())
    s = G.Stokes_SLE_UzawaSolver(pb, G.uzawa_params(tolerance=1e-6))
    s.setup()
    for form in (G.APPLY_MOMENTS_FP64, G.APPLY_MOMENTS_FP32, G.APPLY_PARTICLES_FP64, G.APPLY_PARTICLES_FP32,
                 G.APPLY_G, G.APPLY_GT):
        ms = s.time_apply(form, args.apply_reps, cold=True)
        print(f"# form {form}: {ms:.4f} ms (cold)", file=sys.stderr, flush=True)
    # a few Uzawa iterations of the plain solve, so the solve's vector,
    # reduction and multigrid kernels are profiled as they run inside it
    s.set_params(G.uzawa_params(tolerance=1e-6, maxIterations=args.profile_outer, inner_max_iterations=20000))
    _, _, rep = s.Solve(raise_on_error=False)
    print(f"# capped solve: {rep.uzawa_iterations} Uzawa its, {rep.inner_iterations} inner, "
          f"{rep.solve_seconds:.3f} s", file=sys.stderr, flush=True)
    print(json.dumps({"calibration_copy_bytes_each_way": n * 8, "res": args.res, "ppc": args.ppc,
                      "moments_bytes": k_moments_bytes(cfg.dim, cfg.N, cfg.E),
                      "particles_bytes": k_apply_bytes(cfg.dim, cfg.N, cfg.E, cfg.P)}), flush=True)


def cpu_info():
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    model = ln.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count() or 1
    return model, os.cpu_count() or 1, usable


def roofline_obj(nbytes, ms, kernel, **extra):
    gbs = nbytes / (ms * 1e-3) / 1e9
    d = {"bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": gbs / HBM_PEAK_GBS,
         "kernel": kernel, "ms_per_apply": ms, "algorithmic_bytes": nbytes}
    d.update(extra)
    return d


class Emitter:
    """Holds the one JSON line.  It is printed as soon as the headline and the
    roofline exist and printed again, complete so far, after each later leg,
    so a run stopped by a time limit still leaves a valid line."""

    def __init__(self, budget_s):
        self.t0 = time.perf_counter()
        self.budget = budget_s
        self.line = None
        self.skipped = {}

    def left(self):
        return self.budget - (time.perf_counter() - self.t0)

    def fits(self, name, est_s):
        if self.left() >= est_s:
            return True
        self.skipped[name] = f"skipped: {self.left():.0f} s of the {self.budget:.0f} s bench budget left, " \
                             f"leg needs ~{est_s:.0f} s"
        print(f"# {name}: {self.skipped[name]}", file=sys.stderr, flush=True)
        return False

    def emit(self):
        if self.line is None:
            return
        self.line["config"]["skipped_legs"] = dict(self.skipped)
        self.line["config"]["bench_wall_s"] = time.perf_counter() - self.t0
        print(json.dumps(self.line), flush=True)


def solve_once(sv, p0=None):
    sv.set_pressure(p0)
    u, p, rep = sv.Solve(raise_on_error=False)
    return u, p, rep


def leg_dict(rep, precision, **extra):
    d = {"solve_s": float(rep.solve_seconds), "status": int(rep.status), "uzawa_iterations": int(rep.uzawa_iterations),
         "inner_iterations": int(rep.inner_iterations), "rel_residual": float(rep.rel_residual),
         "div_u": float(rep.div_u), "momentum_residual": float(rep.momentum_residual),
         "learned_fallback": int(rep.learned_fallback), "verify_rounds": int(rep.verify_rounds),
         "precision": precision}
    d.update(extra)
    return d


def rel_diff(a, b, mean_free=False):
    if mean_free:
        a, b = a - a.mean(), b - b.mean()
    den = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (den if den > 0 else 1.0))


PREC_PLAIN = "FP64 throughout (Uzawa, inner flexible CG, multigrid V-cycle, K apply)"
PREC_MIXED = "FP64 Uzawa, inner CG vectors, dots and residuals; FP32 multigrid V-cycle (K applies) as preconditioner"
PREC_NET = "network BF16 (MFMA, FP32 accumulation)"


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3, help="timed steps of the headline (learned) solve")
    ap.add_argument("--warmup", type=int, default=1, help="untimed warmup solves of the headline (learned) solve")
    ap.add_argument("--res", type=int, default=96)
    ap.add_argument("--ppc", type=int, default=16)
    ap.add_argument("--apply-reps", type=int, default=30)
    ap.add_argument("--cpu-res", type=int, default=16)
    ap.add_argument("--budget", type=float, default=360.0,
                    help="wall-clock budget (s) of the whole run; extra legs that would not fit are skipped")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--max-outer", type=int, default=20000)
    ap.add_argument("--plain-max-outer", type=int, default=0,
                    help="cap of the plain solve only (profiling the learned solve); 0 = --max-outer")
    ap.add_argument("--mg-smooth", type=int, default=2)
    ap.add_argument("--weights", default=os.path.join(ROOT, "gale_amd", "lpc_weights.npy"))
    ap.add_argument("--no-ablation", action="store_true")
    ap.add_argument("--no-yield", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--batch-count", type=int, default=256)
    ap.add_argument("--no-dvc", action="store_true", help="skip the DVC weights measurement")
    ap.add_argument("--only-dvc", action="store_true", help="measure only the DVC weights of the workload")
    ap.add_argument("--only-learned", action="store_true",
                    help="only the headline learned solve and the applies (profiling)")
    ap.add_argument("--inner-relax", type=float, default=0.1,
                    help="learned solve: cap of the relaxed inner tolerance (0 = reference 0.1 x tol throughout)")
    ap.add_argument("--profile-apply", action="store_true",
                    help="only run a calibration copy, the operator applies (cold) and a capped plain solve "
                         "(for rocprofv3 kernel traces and --pmc traffic counts)")
    ap.add_argument("--profile-outer", type=int, default=3,
                    help="Uzawa iterations of the capped solve of --profile-apply")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU only; --gpus {args.gpus} is not supported "
              "(the multi-GPU sharding is a design only, not built).", file=sys.stderr)
        sys.exit(2)

    if args.profile_apply:
        profile_apply(args)
        return

    em = Emitter(args.budget)
    import gale_amd as G
    from gale_amd import configs

    t = time.perf_counter()
    cfg = configs.random_viscosity_3d(res=args.res, ppc=args.ppc)
    t_gen = time.perf_counter() - t
    print(f"# generated {cfg.name} res={args.res} E={cfg.E} P={cfg.P} in {t_gen:.1f}s", file=sys.stderr, flush=True)
    pb = G.Problem(**cfg.problem_kwargs())
    if args.only_dvc:
        print(json.dumps({"dvc_weights": bench_dvc(G, pb, cfg, population_control=True, ppc=args.ppc)}), flush=True)
        return
    params = G.uzawa_params(tolerance=1e-6, maxIterations=args.plain_max_outer or args.max_outer,
                            inner_max_iterations=20000, mg_smooth=args.mg_smooth, monitor=500)
    s = G.Stokes_SLE_UzawaSolver(pb, params)
    s.setup()  # first linear solve of the Picard sequence: no yielding yet

    # roofline: the solve's dominant kernel is the FP64 moment-form K apply
    # (every inner FCG iteration and every fine-level smoothing step), timed
    # cold with HIP events on the solver's stream (operand copies cycled so
    # each apply streams from HBM, as inside a solve); the other forms and
    # the pressure operators are reported beside it the same way.
    ms_m = s.time_apply(G.APPLY_MOMENTS_FP64, args.apply_reps, cold=True)
    ms_m_warm = s.time_apply(G.APPLY_MOMENTS_FP64, args.apply_reps, cold=False)
    bytes_m = k_moments_bytes(cfg.dim, cfg.N, cfg.E)
    ms_m32 = s.time_apply(G.APPLY_MOMENTS_FP32, args.apply_reps, cold=True)
    bytes_m32 = k_moments_bytes(cfg.dim, cfg.N, cfg.E, b=4)
    ms_p = s.time_apply(G.APPLY_PARTICLES_FP64, args.apply_reps, cold=True)
    bytes_p = k_apply_bytes(cfg.dim, cfg.N, cfg.E, cfg.P)
    ms_p32 = s.time_apply(G.APPLY_PARTICLES_FP32, args.apply_reps, cold=True)
    bytes_p32 = k_apply_bytes(cfg.dim, cfg.N, cfg.E, cfg.P, b=4)
    # G p reads p (8 B per element) and the node flags (1 B per node) and
    # writes dim values per node; G^T u reads dim values per node and writes
    # one per element
    ms_g = s.time_apply(G.APPLY_G, args.apply_reps, cold=True)
    bytes_g = 8 * cfg.E + 8 * cfg.dim * cfg.N + cfg.N
    ms_gt = s.time_apply(G.APPLY_GT, args.apply_reps, cold=True)
    bytes_gt = 8 * cfg.dim * cfg.N + 8 * cfg.E
    # measured HBM bytes per launch (FETCH_SIZE + WRITE_SIZE passes of
    # rocprofv3, scaled by a 1 GiB copy) of this workload, per form
    traffic = {}
    tpath = os.path.join(ROOT, "profiles", "k_apply_traffic.json")
    if os.path.exists(tpath):
        try:
            tr = json.load(open(tpath))
            if tr.get("res") == args.res and tr.get("ppc") == args.ppc:
                traffic = {k: v.get("bytes_per_apply") for k, v in tr.get("forms", {}).items()}
        except Exception:
            traffic = {}
    print(f"# K apply moments fp64 {ms_m:.4f} ms cold ({bytes_m / ms_m / 1e6:.0f} GB/s), {ms_m_warm:.4f} ms warm; "
          f"fp32 {ms_m32:.4f} ms; particles fp64 {ms_p:.4f} ms, fp32 {ms_p32:.4f} ms; G {ms_g:.4f} ms, "
          f"G^T {ms_gt:.4f} ms", file=sys.stderr, flush=True)
    n_u = pb.N * cfg.dim
    dofs = n_u + cfg.E
    model, nproc, usable = cpu_info()

    line = {
        "metric": METRIC, "value": None, "unit": "s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": None, "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64+f32+bf16", "data": "synthetic",
        "config": {"workload": f"random_viscosity_3d {args.res}^3 Q1-P0, {args.ppc} ppc, log10 eta GRF [-3,3] "
                               f"+ plastic lid (first, linear solve of the Picard sequence), free slip, "
                               f"Uzawa to 1e-6",
                   "elements": cfg.E, "particles": cfg.P, "dofs": dofs, "generate_s": t_gen},
        "roofline": roofline_obj(bytes_m, ms_m, "gs::k_tile_apply3<double,0> + gs::k_tile_face3<double> "
                                 "(moment-form FP64 K apply, fine level; cold: operand copies cycled past the "
                                 "256 MB last-level cache)", traffic=traffic.get("moments_fp64"),
                                 ms_per_apply_warm=ms_m_warm),
        "roofline_moments_fp32": roofline_obj(bytes_m32, ms_m32, "gs::k_tile_apply3<float,0> + "
                                              "gs::k_tile_face3<float> (moment-form FP32 K apply, V-cycle fine "
                                              "level)", traffic=traffic.get("moments_fp32")),
        "roofline_particle_apply": roofline_obj(bytes_p, ms_p, "gs::k_tile_apply3<double,1> + "
                                                "gs::k_tile_face3<double> (particle-form FP64 K apply)",
                                                traffic=traffic.get("particles_fp64"), fp32_ms_per_apply=ms_p32,
                                                fp32_algorithmic_bytes=bytes_p32,
                                                fp32_traffic=traffic.get("particles_fp32")),
        "roofline_G": roofline_obj(bytes_g, ms_g, "gs::k_apply_G<3> (G p)", traffic=traffic.get("G")),
        "roofline_Gt": roofline_obj(bytes_gt, ms_gt, "gs::k_apply_Gt (G^T u)", traffic=traffic.get("Gt")),
        "cpu_baseline": None,
    }
    line["roofline"]["frac"] = line["roofline"]["achieved"] / HBM_PEAK_GBS
    em.line = line
    C = line["config"]

    # headline: learned Schur preconditioner + mixed precision (FP32 V-cycle
    # inside the FP64 flexible inner CG) + relaxed inner tolerance with a
    # true-residual check, network trained on 32^3 fields of the same
    # generator (gale_amd/lpc_train.py).  The plain FP64 solve below is the
    # check every learned solve must meet.
    wpath = args.weights
    net = None
    if os.path.exists(wpath):
        net = G.LearnedPreconditioner(seed=1)
        net.set_params(np.load(wpath))
        mpath = os.path.splitext(wpath)[0] + ".json"
        if os.path.exists(mpath):
            C["training"] = {k: v for k, v in json.load(open(mpath)).items() if k != "loss_curve"}

    def learned_solver(q_every=0, precision=G.PREC_MIXED, relax=args.inner_relax):
        sv = G.Stokes_SLE_UzawaSolver(pb, G.uzawa_params(tolerance=1e-6, maxIterations=args.max_outer,
                                                         inner_max_iterations=20000, mg_smooth=args.mg_smooth,
                                                         monitor=500, learned=True, precision=precision,
                                                         inner_relax_max=relax, learned_q_every=q_every))
        sv.set_learned(net)
        return sv

    ul = pl = None
    learned = None
    if net is not None:
        sl = learned_solver()
        for w in range(args.warmup):
            _, _, rl = solve_once(sl)
            print(f"# learned warmup {w}: {rl.solve_seconds:.3f}s uzawa={rl.uzawa_iterations} status={rl.status}",
                  file=sys.stderr, flush=True)
        ltimes = []
        sl.sync()
        t0 = time.perf_counter()
        for k in range(args.steps):
            ul, pl, rl = solve_once(sl)
            ltimes.append(rl.solve_seconds)
            print(f"# learned step {k}: {rl.solve_seconds:.3f}s uzawa={rl.uzawa_iterations} "
                  f"inner={rl.inner_iterations} fallback={rl.learned_fallback} rel={rl.rel_residual:.3e}",
                  file=sys.stderr, flush=True)
        sl.sync()
        lwall = time.perf_counter() - t0
        learned = leg_dict(rl, PREC_MIXED + "; " + PREC_NET + "; inner rtol relaxed up to "
                           f"{args.inner_relax} with a true-residual check", wall_s=lwall,
                           solve_s_each=[float(v) for v in ltimes])
        learned["solve_s"] = float(np.mean(ltimes))
        C["learned"] = learned
        if learned["status"] == 0:
            line["value"] = learned["solve_s"]
            line["ms_per_step"] = lwall / max(args.steps, 1) * 1e3
            C["value_is"] = ("learned Schur preconditioner + mixed precision (FP32 V-cycle in FP64 flexible inner "
                             "CG) + relaxed inner tolerance with a true-residual check; mean device time per solve")
            C["dof_per_s"] = dofs / learned["solve_s"]
            em.emit()
        sl.close()
    if args.only_learned:
        em.emit()
        return

    # plain FP64 GPU Uzawa (the reference algorithm), timed once: its ~40 s
    # would otherwise dominate the run
    u, p, rep = solve_once(s)
    C["plain_runs"] = 1
    C["plain"] = leg_dict(rep, PREC_PLAIN, method="reference Uzawa, diagonal Q, inner FP64 flexible CG + "
                          "multigrid V-cycle at 0.1 x tol", dof_per_s=dofs / rep.solve_seconds)
    C["plain_fp64_solve_s"] = float(rep.solve_seconds)
    print(f"# plain: {rep.solve_seconds:.3f}s uzawa={rep.uzawa_iterations} inner={rep.inner_iterations} "
          f"rel={rep.rel_residual:.3e} status={rep.status}", file=sys.stderr, flush=True)
    if line["value"] is None:
        line["value"] = float(rep.solve_seconds)
        line["ms_per_step"] = rep.solve_seconds * 1e3
        line["steps"], line["warmup"] = 1, 0
        C["value_is"] = "plain FP64 Uzawa (the learned solve did not converge or no network was found)"
        C["dof_per_s"] = dofs / rep.solve_seconds
    if learned is not None and learned["status"] == 0:
        learned["speedup_vs_plain"] = rep.solve_seconds / learned["solve_s"]
        learned["u_rel_diff_vs_plain"] = rel_diff(ul, u)
        learned["p_rel_diff_vs_plain"] = rel_diff(pl, p, mean_free=True)
    em.emit()

    # ablation: what each change contributes, each solve once
    if net is not None and not args.no_ablation:
        abl = {}
        legs = [
            ("no_network_mixed_relaxed", dict(q_every=1), PREC_MIXED + "; no network (every iteration Q^-1 r); "
             "relaxed inner rtol; same truncated flexible outer loop as the headline"),
            ("network_mixed_no_relax", dict(relax=0.0), PREC_MIXED + "; " + PREC_NET + "; inner rtol 0.1 x tol"),
            ("network_fp64_no_relax", dict(relax=0.0, precision=G.PREC_FP64), PREC_PLAIN + "; " + PREC_NET +
             "; inner rtol 0.1 x tol"),
        ]
        est = {"no_network_mixed_relaxed": 30.0, "network_mixed_no_relax": 15.0, "network_fp64_no_relax": 20.0}
        for name, kw, prec in legs:
            if not em.fits(f"ablation.{name}", est[name]):
                continue
            sv = learned_solver(**kw)
            ua, pa, ra = solve_once(sv)
            sv.close()
            abl[name] = leg_dict(ra, prec, speedup_vs_plain=rep.solve_seconds / ra.solve_seconds,
                                 u_rel_diff_vs_plain=rel_diff(ua, u), p_rel_diff_vs_plain=rel_diff(pa, p, True))
            print(f"# ablation {name}: {ra.solve_seconds:.3f}s uzawa={ra.uzawa_iterations} "
                  f"inner={ra.inner_iterations} status={ra.status}", file=sys.stderr, flush=True)
            C["ablation"] = abl
            if learned is not None and name == "no_network_mixed_relaxed" and ra.status == 0:
                learned["network_speedup"] = ra.solve_seconds / learned["solve_s"]
            em.emit()

    # a yielded solve: the second (Picard) solve of the sequence, with the
    # Drucker-Prager lid yielding from the first solution (the plain FP64
    # one, for both solvers, so both solve the same problem), warm-started
    # from its pressure as the reference keeps pSolnVec
    if not args.no_yield and em.fits("yielded", 10.0):
        yl = {}
        if net is not None:
            sv = learned_solver()
            sv.setup(u, p, yield_on=True)
            eta_y = sv.setup_arrays()["eta"]
            lid = cfg.material == 1
            yl["lid_points_yielded"] = int(np.sum(eta_y[lid] < 1e2))
            yl["lid_points"] = int(np.sum(lid))
            uy, py, ry = solve_once(sv, p)
            sv.close()
            yl["learned"] = leg_dict(ry, PREC_MIXED + "; " + PREC_NET + "; relaxed inner rtol")
            print(f"# yielded learned: {ry.solve_seconds:.3f}s uzawa={ry.uzawa_iterations} status={ry.status}",
                  file=sys.stderr, flush=True)
        C["yielded"] = yl
        em.emit()
        if em.fits("yielded.plain", 1.3 * rep.solve_seconds + 10):
            s.setup(u, p, yield_on=True)
            uyp, pyp, ryp = solve_once(s, p)
            yl["plain"] = leg_dict(ryp, PREC_PLAIN)
            if "learned" in yl and yl["learned"]["status"] == 0 and ryp.status == 0:
                yl["learned"]["speedup_vs_plain"] = ryp.solve_seconds / yl["learned"]["solve_s"]
                yl["learned"]["u_rel_diff_vs_plain"] = rel_diff(uy, uyp)
                yl["learned"]["p_rel_diff_vs_plain"] = rel_diff(py, pyp, True)
            print(f"# yielded plain: {ryp.solve_seconds:.3f}s uzawa={ryp.uzawa_iterations} status={ryp.status}",
                  file=sys.stderr, flush=True)
            em.emit()

    if not args.no_cpu and em.fits("cpu_baseline", 60.0):
        threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or usable
        threads = max(1, min(threads, usable))
        t = time.perf_counter()
        cb = cpu_baseline(args.cpu_res, threads)
        print(f"# cpu baseline {cb} in {time.perf_counter() - t:.1f}s", file=sys.stderr, flush=True)
        # the GPU on the same sample, for a like-for-like ratio
        cfg_s = configs.random_viscosity_3d(res=args.cpu_res, ppc=16)
        ps = G.Problem(**cfg_s.problem_kwargs())
        ss = G.Stokes_SLE_UzawaSolver(ps, G.uzawa_params(tolerance=1e-6, inner_max_iterations=20000))
        _, _, rs = ss.Solve(raise_on_error=False)
        ss.close()
        ps.close()
        line["cpu_baseline"] = {
            "value": cb["solve_s"], "unit": "s", "cores": threads, "kind": "port",
            "sample": f"random_viscosity_3d at {args.cpu_res}^3 ({cb['E']} elements, {cb['P']} particles), same "
                      f"generator as the workload: oracle assembled CSR + reference Uzawa, GMRES(30)+ILU(0) inner "
                      f"(PETSc defaults), {cb['uzawa_iterations']} Uzawa its, OpenMP; assembly "
                      f"{cb['assemble_s']:.2f}s not included",
            "gpu_same_sample_plain_fp64_s": rs.solve_seconds, "status": cb["status"],
            "host_cpu_model": model, "host_nproc": nproc, "host_usable_cpus": usable}
        em.emit()

    if not args.no_batch and em.fits("batch", 60.0):
        C["batch"] = bench_batch(G, configs, args.batch_count)
        em.emit()
    if not args.no_dvc and em.fits("dvc_weights", 20.0):
        C["dvc_weights"] = bench_dvc(G, pb, cfg)
    em.emit()


if __name__ == "__main__":
    main()
