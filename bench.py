"""Benchmark: FDV time steps/s at 2.1M hex8 on one MI355X.

Workload: the 3D hypersonic 15-degree compression ramp, structured hex8
256x128x64 (2,097,152 elements, 2,154,945 nodes, 5 equations), M=6,
Re=5e5, adiabatic wall, CFL 0.5, synthetic initial state (uniform
freestream, wall at rest).  A step is one full FDV time step on the GPU:
CFL, element build (FDVParam + domain + boundary + ApplyBC, FP64 arithmetic),
GMRES to a 1e-8 relative true residual (both paths right-preconditioned by
node-block Jacobi unless --precond none; each solve after the first two
starts from dt (2 du_1 / dt_1 - du_2 / dt_2), the increment rate of the two
previous steps extrapolated, the second from du_1 dt / dt_1; --warm-start 1
uses the latter throughout, --warm-start 0 starts every solve from 0),
update and RMS monitor.

Two paths are timed from the same initial state:
  fp64  : FP64 blocks, FP64 restarted GMRES(m)
  mixed : FP32 blocks with FP32 accumulation inside FP64 iterative
          refinement (outer residual with FP64 blocks)
The solver's products K x run on the assembled node rows (the element
blocks summed per node pair after each build: 27 blocks per node instead of
64 per element); ``--storage ebe`` multiplies element by element as the
reference does.
``value`` is the mixed path's steps/s; the FP64 path and the ratio are
reported beside it.  The CPU baseline times one full step of the CPU
oracle (a port of the reference algorithm, OpenMP over elements and element
colors) on the 128 x 64 x 32 supersonic flat plate (262,144 hex8) at the
host's allowed thread count (``--cpu-threads 1,all`` adds one thread), and
the GPU on that same mesh beside it.
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

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFS = 78.6
BUILD_FLOP_PER_ELEMENT = 0.70541e6
# repeated nodes of a refinement are interpolated to the same coordinates up
# to rounding; the reference's absolute 1e-3 (MeshRefine2D.h:337-361) would
# merge distinct nodes of the level-2 wall cells (2.8e-4 high)
REFINE_NODE_TOL = 1e-9

# HBM bytes per full product measured with rocprofv3 FETCH_SIZE / WRITE_SIZE
# on MI355X, each scaled by the counter's reading for a plain 1 GiB copy
# kernel (fetch x2.0, write x1.0); summaries and commands in
# profiles/ebe_matvec_ramp_256x128x64_hbm_counters.txt (EBE, 8 color
# launches), profiles/rows_matvec_ramp_256x128x64_hbm_counters.txt (node
# rows, one launch) and profiles/rows_matvec_ramp_refined_4390912_hbm_counters.txt
# (node rows of the ramp after two sensor refinement passes).  Keyed by
# (storage, base size, refinement levels); other workloads have none.
MEASURED_TRAFFIC = {("ebe", "256x128x64", 0): {"fp32": 14.607e9, "fp64": 29.121e9},
                    ("rows", "256x128x64", 0): {"fp32": 6.342e9, "fp64": 12.502e9},
                    ("rows", "256x128x64", 2): {"fp32": 12.918e9, "fp64": 25.348e9}}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    # a time loop of 10 steps from the initial state (the first solve cold,
    # the rest warm-started)
    p.add_argument("--steps", type=int, default=10)
    p.add_argument("--warmup", type=int, default=1)
    p.add_argument("--size", default="256x128x64")
    p.add_argument("--cfl", type=float, default=0.5)
    p.add_argument("--m", type=int, default=30)
    p.add_argument("--rtol", type=float, default=1e-8)
    p.add_argument("--inner-m", type=int, default=30)
    # the mixed path's inner FP32 GMRES tolerance: 1e-4 gives two outer FP64
    # sweeps to 1e-8 with the fewest FP32 matvecs on the ramp (1e-3: three
    # sweeps, 1e-5: more inner iterations; 3e-4, 2e-4 and 5e-5: 25.0-25.1
    # FP32 products per step and 5.88-5.90 steps/s against 24.8 and 5.93)
    p.add_argument("--inner-rtol", type=float, default=1e-4)
    p.add_argument("--precond", choices=["nodeblock", "none"], default="nodeblock")
    # rows: FP64 element blocks, node rows summed from them after the build;
    # rows-inplace: node rows built in place by the build (no element
    # blocks); ebe: element blocks only, products element by element
    p.add_argument("--storage", choices=["rows", "rows-inplace", "ebe"], default="rows")
    # 2: each solve starts from the increment rate of the two previous steps
    # extrapolated (the first step after the initial state starts from 0,
    # the second from du_1 dt / dt_1); 3: quadratic in the rate of the last
    # three; 1: always du_1 dt / dt_1; 0: every solve from 0 as the
    # reference does
    p.add_argument("--warm-start", type=int, default=2, choices=[0, 1, 2, 3])
    # sensor-driven refinement before the timed steps: passes of the GPU
    # flag pass (s1 > 0.8) + host AdaptiveRefineList / MeshRefine2D
    # (every coordinate compared for repeated nodes), each on the mesh and
    # state the previous pass left
    p.add_argument("--refine-levels", type=int, default=2)
    p.add_argument("--no-cpu", action="store_true")
    # thread counts of the CPU oracle's full step ("all" = the threads this
    # process may use); "1,all" adds the single-thread step (minutes)
    p.add_argument("--cpu-threads", default="all")
    return p.parse_args()


def sync():
    try:
        import torch
        if torch.cuda.is_available():
            torch.cuda.synchronize()
    except Exception:
        pass


def run_steps(ctx, U0, mode, n, args):
    ctx.set_state(U0)
    out = []
    for i in range(n):
        st = ctx.step(args.cfl, mode=mode, restarts=60 if mode == "fp64" else 20, m=args.m,
                      rtol=args.rtol, inner_m=args.inner_m, inner_rtol=args.inner_rtol,
                      inner_restarts=3, precond=args.precond, warm_start=int(args.warm_start))
        out.append(st)
        log(f"  {mode} step {i}: build {st.ms_build:.1f} ms assemble {st.ms_assemble:.1f} ms "
            f"solve {st.ms_solve:.1f} ms "
            f"mv64 {st.solve.matvecs64} mv32 {st.solve.matvecs32} rel {st.solve.rel_resid:.2e}")
    return out


def cpu_info():
    """CPU model, physical cores and the threads this process may use."""
    model, cores, fam, mod = None, set(), None, None
    try:
        phys = core = None
        with open("/proc/cpuinfo") as f:
            for line in f:
                k, _, v = line.partition(":")
                k, v = k.strip(), v.strip()
                if k == "model name" and model is None:
                    model = v
                elif k == "cpu family" and fam is None:
                    fam = v
                elif k == "model" and mod is None:
                    mod = v
                elif k == "physical id":
                    phys = v
                elif k == "core id":
                    core = v
                elif not k and phys is not None:
                    cores.add((phys, core))
                    phys = core = None
    except OSError:
        pass
    avail = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    env = os.environ.get("OMP_NUM_THREADS")
    threads = min(avail, int(env)) if env and env.isdigit() and int(env) > 0 else avail
    if model and fam:
        model = f"{model} (family {fam}, model {mod})"
    return {"cpu_model": model, "physical_cores": len(cores) or None, "logical_cpus": os.cpu_count(),
            "affinity_cpus": avail, "threads_allowed": threads}


def cpu_baseline(threads_list, gpu_steps=3):
    """One full FDV step of the CPU oracle (a port of the reference
    algorithm: build + ApplyBC + restarted GMRES(30) to a 1e-8 relative
    true residual, unpreconditioned, from 0 + update) on the 128 x 64 x 32
    supersonic flat plate (262,144 hex8), timed at each thread count; and the
    GPU on the same mesh and state, once with the same algorithm and once
    with the benchmark's (node-block Jacobi, warm start)."""
    from oracle import oracle as O
    from cfdv_amd import mesh as M
    from cfdv_amd.fdv import FDVContext
    mesh, th, U = M.flat_plate(128, 64, 32)
    colors, ncol = O.color_greedy(mesh)
    runs = []
    for t in threads_list:
        log(f"cpu baseline: oracle step on the 128x64x32 plate, {t} threads")
        t0 = time.perf_counter()
        _, _, st, ph = O.step_par(th, mesh, U, threads=t, cfl=0.5, mdim=30, rtol=1e-8, maxrestart=60,
                                  colors=colors, verbose=True)
        wall = time.perf_counter() - t0
        runs.append({"threads": t, "steps_per_s": 1.0 / wall, "s_per_step": wall,
                     "s_build": float(ph[1]), "s_bc_rhs": float(ph[2]), "s_gmres_update": float(ph[3]),
                     "products": int(st.gm.matvecs),
                     "rel_resid": float(st.gm.true_resid / st.gm.bnorm)})
        log(f"  {wall:.1f} s: build {ph[1]:.1f} s, GMRES {ph[3]:.1f} s, {st.gm.matvecs} products, "
            f"rel resid {st.gm.true_resid / st.gm.bnorm:.2e}")
    gpu = {}
    ctx = FDVContext(mesh, th, U, blocks=("fp64", "rows"))
    for name, mode, pre, warm in (("fp64_same_algorithm", "fp64", "none", 0),
                                  ("fp64", "fp64", "nodeblock", 2), ("mixed", "mixed", "nodeblock", 2)):
        ctx.set_state(U)
        kw = dict(restarts=60 if mode == "fp64" else 20, m=30, rtol=1e-8, inner_m=30, inner_rtol=1e-4,
                  inner_restarts=3, precond=pre, warm_start=warm)
        ctx.step(0.5, mode=mode, **kw)
        sync()
        t0 = time.perf_counter()
        sts = [ctx.step(0.5, mode=mode, **kw) for _ in range(gpu_steps)]
        sync()
        gpu[name] = {"steps_per_s": gpu_steps / (time.perf_counter() - t0),
                     "fp64_products_per_step": sum(s.solve.matvecs64 for s in sts) / gpu_steps,
                     "fp32_products_per_step": sum(s.solve.matvecs32 for s in sts) / gpu_steps,
                     "rel_resid": max(s.solve.rel_resid for s in sts)}
    del ctx
    return {"mesh": "128x64x32 flat plate (262,144 hex8, 274,625 nodes), M 3, Re 1e5, T_w/T_inf 2",
            "colors": int(ncol), "runs": runs, "gpu_same_mesh": gpu}


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the FDV path "
              "is built and measured on a single MI355X", file=sys.stderr)
        sys.exit(2)
    from cfdv_amd import mesh as M
    from cfdv_amd.fdv import FDVContext, field_differences

    nx, ny, nz = (int(v) for v in args.size.split("x"))
    t0 = time.time()
    mesh, th, U0 = M.compression_ramp(nx, ny, nz)
    log(f"mesh {nx}x{ny}x{nz}: {mesh.n_elem} elements, {mesh.n_node} nodes ({time.time() - t0:.1f}s)")
    base_elem, base_node = mesh.n_elem, mesh.n_node
    refine_info = None
    if args.refine_levels > 0:
        from cfdv_amd.refine import sensor_refine
        mesh, U0, refine_info = sensor_refine(mesh, th, U0, args.refine_levels, tol=REFINE_NODE_TOL)
        for p in range(args.refine_levels):
            log(f"refine pass {p}: {refine_info['flagged'][p]} flagged, {refine_info['refined'][p]} refined -> "
                f"{refine_info['elements'][p + 1]} elements, {refine_info['nodes'][p + 1]} nodes; host list "
                f"{refine_info['host_list_s'][p]:.1f} s, apply {refine_info['host_apply_s'][p]:.1f} s")
    t0 = time.time()
    blocks = {"rows": ("fp64", "rows"), "rows-inplace": ("rows",), "ebe": ("fp64", "fp32")}[args.storage]
    ctx = FDVContext(mesh, th, U0, blocks=blocks)
    info = ctx.info()
    log(f"context: {info} ({time.time() - t0:.1f}s)")

    results, finals = {}, {}
    for mode in ("fp64", "mixed"):
        log(f"{mode}: warmup {args.warmup}")
        run_steps(ctx, U0, mode, args.warmup, args)
        sync()
        t0 = time.perf_counter()
        sts = run_steps(ctx, U0, mode, args.steps, args)
        sync()
        el = time.perf_counter() - t0
        results[mode] = (el, sts)
        finals[mode] = ctx.get_state()[0]
    ferr = field_differences(mesh.ndim, U0, finals["fp64"], finals["mixed"])
    log(f"mixed vs fp64 after {args.steps} steps: {ferr}")

    # the mixed path again with every solve started from 0, for comparison
    cold = None
    if args.warm_start:
        cargs = argparse.Namespace(**vars(args))
        cargs.warm_start = 0
        run_steps(ctx, U0, "mixed", args.warmup, cargs)
        sync()
        t0 = time.perf_counter()
        sts_c = run_steps(ctx, U0, "mixed", args.steps, cargs)
        sync()
        cold = {"steps_per_s": args.steps / (time.perf_counter() - t0),
                "fp32_products_per_step": sum(s.n_matvec32 for s in sts_c) / args.steps}
    el_mx, st_mx = results["mixed"]
    el_64, st_64 = results["fp64"]
    value = args.steps / el_mx
    fp64_sps = args.steps / el_64

    ne, nn = mesh.n_elem, mesh.n_node
    if args.storage.startswith("rows"):
        # one block per node pair sharing an element (on the unrefined
        # structured grid (3nx+1)(3ny+1)(3nz+1) of them); each block's 25
        # values plus its column index, x read and y written once
        nnzb, _ = ctx.rows_info()
        b32 = nnzb * (25 * 4 + 4) + nn * 5 * 2 * 4
        b64 = nnzb * (25 * 8 + 4) + nn * 5 * 2 * 8
        kname = "rows_matvec_kernel<float,float,float,5> (assembled node rows, FP32)"
        per = f"{nnzb} node blocks"
    else:
        b32 = ne * 1600 * 4 + nn * 5 * 2 * 4
        b64 = ne * 1600 * 8 + nn * 5 * 2 * 8
        kname = "matvec_kernel<float,float,float,8,5> (EBE matvec, FP32 blocks)"
        per = f"{info['ncolors']} color launches"
    ms32 = sum(s.ms_matvec32 for s in st_mx)
    n32 = sum(s.n_matvec32 for s in st_mx)
    ms64 = sum(s.ms_matvec64 for s in st_64)
    n64 = sum(s.n_matvec64 for s in st_64)
    ach32 = b32 * n32 / (ms32 * 1e-3) / 1e9 if ms32 > 0 else None
    ms_build = float(np.mean([s.ms_build for s in st_mx]))
    ach64 = b64 * n64 / (ms64 * 1e-3) / 1e9 if ms64 > 0 else None

    cpu = None
    if not args.no_cpu:
        ci = cpu_info()
        tl = [ci["threads_allowed"] if t.strip() == "all" else int(t) for t in args.cpu_threads.split(",")]
        cb = cpu_baseline(tl)
        main_run = max(cb["runs"], key=lambda r: r["threads"])
        # secondary: the measured per-element build time and per-product time
        # of the plate step scaled to this workload's elements and the FP64
        # GPU path's product count (an extrapolation, not a measurement)
        e_c2 = 262144
        mv_per_step = n64 / max(1, args.steps)
        t_build = main_run["s_build"] * ne / e_c2
        t_mv = main_run["s_gmres_update"] / max(1, main_run["products"]) * ne / e_c2
        cpu = {"value": main_run["steps_per_s"], "unit": "steps/s", "cores": main_run["threads"],
               "kind": "port", "measured": True, **ci,
               "sample": (f"one full FDV step of the CPU oracle (FP64 port of the reference algorithm: "
                          f"build + ApplyBC + GMRES(30) to a 1e-8 relative true residual, "
                          f"unpreconditioned, from 0 + update; OpenMP over elements for the build and over "
                          f"the {cb['colors']} element colors for the EBE product) on the {cb['mesh']}, "
                          f"{main_run['threads']} threads, {main_run['products']} products"),
               "runs": cb["runs"], "gpu_same_mesh": cb["gpu_same_mesh"],
               "secondary_extrapolated_to_workload": {
                   "steps_per_s": 1.0 / (t_build + mv_per_step * t_mv),
                   "basis": (f"plate step's build time per element and product time per element scaled to "
                             f"{ne} elements with {mv_per_step:.0f} FP64 products per step (the FP64 GPU "
                             f"path's count), at {main_run['threads']} threads")}}

    traffic = MEASURED_TRAFFIC.get(("rows" if args.storage.startswith("rows") else args.storage, args.size,
                                    args.refine_levels), {})
    wl = (f"3D hypersonic 15-deg compression ramp hex8 {nx}x{ny}x{nz} ({base_elem} elements, "
          f"{base_node} nodes)")
    if refine_info is not None:
        wl += (f", sensor-refined (s1 > 0.8) {args.refine_levels} levels to {ne} elements, {nn} nodes "
               f"({refine_info['adap_elements']} with hanging nodes)")
    wl += (f", M=6, Re=5e5, adiabatic, CFL {args.cfl}, GMRES({args.m}) to rel. residual {args.rtol}, "
           f"{args.precond} right preconditioner on both paths")
    line = {
        "metric": "FDV steps/s at 2.1M hex8: FP32+FP64-refine vs FP64 GPU, equal 1e-8 residual",
        "value": value,
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 / value,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 blocks/matvec + f64 build, residual, Givens on GPU, update",
        "data": "synthetic (generated ramp mesh and freestream initial state)",
        "config": {"workload": wl, "base_elements": base_elem, "base_nodes": base_node,
                   "elements": ne, "nodes": nn, "colors": info["ncolors"],
                   "precond": args.precond, "inner_m": args.inner_m, "inner_rtol": args.inner_rtol,
                   "storage": args.storage, "warm_start": int(args.warm_start)},
        "fp64_steps_per_s": fp64_sps,
        "mixed_cold_start": cold,
        "mixed_over_fp64": value / fp64_sps,
        "rel_resid_mixed": max(s.solve.rel_resid for s in st_mx),
        "rel_resid_fp64": max(s.solve.rel_resid for s in st_64),
        # mixed minus FP64 after the timed steps from the same state, per
        # field, relative to the FP64 path's change of the field
        "mixed_vs_fp64_fields": ferr,
        "refine": refine_info,
        "ms_build": float(np.mean([s.ms_build for s in st_mx])),
        "ms_assemble": float(np.mean([s.ms_assemble for s in st_mx])),
        "ms_solve_mixed": float(np.mean([s.ms_solve for s in st_mx])),
        "ms_solve_fp64": float(np.mean([s.ms_solve for s in st_64])),
        "matvecs_per_step_mixed": {"fp32": n32 / args.steps,
                                   "fp64": sum(s.n_matvec64 for s in st_mx) / args.steps},
        "matvecs_per_step_fp64": n64 / args.steps,
        "roofline": {"bound": "hbm", "achieved": ach32, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": (ach32 / HBM_PEAK_GBS) if ach32 else None,
                     "traffic": traffic.get("fp32"),
                     "kernel": f"{kname}, {b32} algorithmic bytes per full product ({per})"},
        "roofline_fp64_matvec": {"achieved": ach64, "frac": (ach64 / HBM_PEAK_GBS) if ach64 else None,
                                 "bytes_per_matvec": b64, "traffic": traffic.get("fp64")},
        # the element build (interior and face kernels): algorithmic FP64
        # flops of the contraction form, 0.705 MFLOP per hex8 element
        # (counted on the all-VALU kernel: 64 (2 FMA + MUL + ADD) per wave)
        "roofline_build": {"bound": "fp64", "unit": "TFLOP/s", "peak": FP64_PEAK_TFS,
                           "achieved": BUILD_FLOP_PER_ELEMENT * ne / (ms_build * 1e-3) / 1e12,
                           "frac": BUILD_FLOP_PER_ELEMENT * ne / (ms_build * 1e-3) / 1e12 / FP64_PEAK_TFS,
                           "ms": ms_build},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
