"""Stand-alone run of the element build and the products K x, for profilers.

    python -m cfdv_amd.profile_matvec --size 256x128x64 --reps 20 [--storage rows] [--refine-levels 2]

Builds the compression-ramp mesh of the given size (with ``--refine-levels``
refined by that many sensor passes, as the benchmark does), runs one element build
(with ``--storage rows`` followed by the assembly of the node rows), then
``--reps`` full products in FP32 and in FP64 (element by element, or over
the assembled rows) and ``--reps`` 8-vector FP32 EBE products, and prints
one JSON line with the event-timed milliseconds and algorithmic bytes of
each.
With ``--copy-mib N`` it first runs one plain device copy of N MiB (an
elementwise ``mul(src, 1)`` through torch) whose HBM bytes are known, so that
hardware byte counters can be scaled against it.
"""
from __future__ import annotations

import argparse
import json
import sys
import time

from . import mesh as M
from .fdv import FDVContext


def main(argv=None):
    p = argparse.ArgumentParser()
    p.add_argument("--size", default="256x128x64")
    p.add_argument("--reps", type=int, default=20)
    p.add_argument("--cfl", type=float, default=0.5)
    p.add_argument("--copy-mib", type=int, default=0)
    p.add_argument("--precision", choices=["32", "64", "both"], default="both")
    p.add_argument("--storage", choices=["ebe", "rows"], default="ebe")
    p.add_argument("--refine-levels", type=int, default=0)
    p.add_argument("--no-multi", action="store_true", help="skip the 8-vector product")
    p.add_argument("--adjoint8", action="store_true",
                   help="also time 8 adjoint solves at once against 8 single ones (rtol 1e-10)")
    a = p.parse_args(argv)
    nx, ny, nz = (int(v) for v in a.size.split("x"))
    out = {"size": a.size}
    if a.copy_mib:
        import torch
        n = a.copy_mib * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        torch.mul(src, 1.0, out=dst)
        torch.cuda.synchronize()
        out["copy_bytes_read"] = n * 4
        out["copy_bytes_written"] = n * 4
        del src, dst
    mesh, th, U = M.compression_ramp(nx, ny, nz)
    if a.refine_levels > 0:
        from .refine import sensor_refine
        mesh, U, _ = sensor_refine(mesh, th, U, a.refine_levels)
    rows = a.storage == "rows"
    ctx = FDVContext(mesh, th, U, blocks=("fp64", "fp32", "rows") if rows else ("fp64", "fp32"))
    dt = ctx.cfl(a.cfl)
    t0 = time.perf_counter()
    ctx.build(dt)
    ctx.rhs_vector()
    out["build_wall_ms"] = (time.perf_counter() - t0) * 1e3
    ne, nn = mesh.n_elem, mesh.n_node
    k = mesh.k
    # node blocks of the assembled rows (on the structured grid
    # (3nx+1)(3ny+1)(3nz+1)): 25 values and a column index each
    nnzb = ctx.rows_info()[0] if rows else 0
    if rows:
        out["row_blocks"], out["row_slots_stored"] = ctx.rows_info()
    for prec in ((32, 64) if a.precision == "both" else (int(a.precision),)):
        ms = ctx.profile_matvec(prec, a.reps)
        w = prec // 8
        if rows:
            # x read and y written once
            nbytes = nnzb * (mesh.neqn ** 2 * w + 4) + nn * mesh.neqn * 2 * w
        else:
            nbytes = ne * k * k * w + nn * mesh.neqn * 2 * w
        out[f"matvec{prec}_ms"] = ms
        out[f"matvec{prec}_bytes"] = nbytes
        out[f"matvec{prec}_GBps"] = nbytes / (ms * 1e-3) / 1e9
    if a.no_multi:
        out.update(elements=ne, nodes=nn, colors=ctx.info()["ncolors"], storage=a.storage)
        print(json.dumps(out), flush=True)
        return 0
    ms8 = ctx.profile_matvec_multi(a.reps)
    if rows:
        b8 = nnzb * (mesh.neqn ** 2 * 4 + 4) + nn * mesh.neqn * 8 * 2 * 4
    else:
        b8 = ne * k * k * 4 + nn * mesh.neqn * 8 * 2 * 4
    out["matvec32x8_ms"] = ms8
    out["matvec32x8_bytes"] = b8
    out["matvec32x8_GBps"] = b8 / (ms8 * 1e-3) / 1e9
    out["matvec32x8_speedup_vs_8_single"] = 8 * out.get("matvec32_ms", ms8) / ms8
    if a.adjoint8:
        import numpy as np
        G = np.random.default_rng(3).standard_normal((8, ctx.ndof))
        ctx.adjoint_solve_multi(G[:, :], rtol=1e-10)   # warm
        t0 = time.perf_counter()
        _, rel, st = ctx.adjoint_solve_multi(G, rtol=1e-10)
        out["adjoint8_s"] = time.perf_counter() - t0
        out["adjoint8_products"] = {"x8_fp32": st.matvecs32, "fp64": st.matvecs64, "sweeps": st.restarts}
        out["adjoint8_max_rel"] = float(rel.max())
        for mode in ("mixed", "fp64"):
            t0 = time.perf_counter()
            n32 = n64 = 0
            for s in range(8):
                _, st1 = ctx.adjoint_solve(G[s], mode=mode, rtol=1e-10, restarts=200 if mode == "fp64" else 20)
                n32 += st1.matvecs32
                n64 += st1.matvecs64
            out[f"adjoint_single_{mode}_x8_s"] = time.perf_counter() - t0
            out[f"adjoint_single_{mode}_x8_products"] = {"fp32": n32, "fp64": n64}
    out["elements"] = ne
    out["nodes"] = nn
    out["colors"] = ctx.info()["ncolors"]
    out["storage"] = a.storage
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
