"""Benchmark: Kantorovich evaluations (masses + sparse Hessian) per second at
N = 1M Diracs on one MI355X, plus the Newton solve time to max relative mass
error 1e-8 in FP64 and mixed precision.

Workload (mixed precision at scale): unit square as a 1024 x 1024 vertex grid
(2,093,058 triangles), P1 density 0.05 + 32 Gaussians (sigma in [0.01, 0.1]),
N = 1,000,000 Diracs rejection-sampled from the density, lognormal(0, 0.5)
target masses normalised to 1 (mongeampere_amd.configs.config3, seed 3).

A step is one evaluation (value, masses and the CSR Hessian on the device) at
device-resident weights.  Prints one JSON line.
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

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E


def algorithmic_bytes(N, V, T, nnz, s_p=8, s_v=8, s_h=8):
    """Bytes one evaluation must move at minimum: positions + weights in,
    masses out, mesh coordinates + values, triangle indices, CSR values +
    columns + row pointers."""
    return N * (2 * s_p + 8 + 8) + 3 * V * s_v + 12 * T + nnz * (s_h + 4) + 4 * (N + 1)


def cpu_baseline_sample(p, frac=0.5):
    """Times the FP64 CPU oracle (one thread) on a window [0, frac]^2 of the
    same workload and scales the per-Dirac cost to the full problem."""
    import oracle
    from mongeampere_amd.configs import Mesh

    n = p.grid_n
    keep = p.mesh.vxy.max(axis=1) <= frac + 1e-12
    vid = -np.ones(p.mesh.vxy.shape[0], np.int64)
    vid[keep] = np.arange(int(keep.sum()))
    tk = keep[p.mesh.tri].all(axis=1)
    tri = vid[p.mesh.tri[tk]].astype(np.int32)
    mesh = Mesh(p.mesh.vxy[keep], p.mesh.vval[keep], tri)
    inside = (p.X <= frac).all(axis=1)
    X = p.X[inside]
    w = np.zeros(X.shape[0])
    t0 = time.perf_counter()
    oracle.kantorovich(mesh, X, w)
    dt = time.perf_counter() - t0
    per_dirac = dt / X.shape[0]
    return {
        "value": 1.0 / (per_dirac * p.X.shape[0]),
        "unit": "evals/s",
        "cores": 1,
        "kind": "port",
        "sample": (f"FP64 oracle (C restatement, exact predicates) on the window [0,{frac}]^2 of "
                   f"the same workload: {X.shape[0]} Diracs, {tri.shape[0]} triangles, "
                   f"{dt:.2f} s; scaled by Dirac count to N={p.X.shape[0]}"),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--precision", default="fp64", choices=["fp64", "mixed"])
    ap.add_argument("--no-solve", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--N", type=int, default=1_000_000)
    ap.add_argument("--grid", type=int, default=1024)
    ap.add_argument("--solve-precisions", default="fp64,mixed")
    ap.add_argument("--calib-copy-mb", type=int, default=0,
                    help="run one device copy of this many MiB first (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py runs on one GPU only: this engine has no multi-GPU path "
              "(see DESIGN.md for how the evaluation would shard).", file=sys.stderr)
        sys.exit(2)

    import torch
    from mongeampere_amd import configs
    from mongeampere_amd.api import Engine

    p = configs.config3(N=args.N, n=args.grid)
    N = p.X.shape[0]
    e = Engine(0)
    e.set_mesh(p.mesh).set_diracs(p.X)
    dev = torch.device("cuda", 0)
    w_dev = torch.zeros(N, dtype=torch.float64, device=dev)
    if args.calib_copy_mb > 0:
        a = torch.ones(args.calib_copy_mb * (1 << 20) // 8, dtype=torch.float64, device=dev)
        b = torch.empty_like(a)
        b.copy_(a)
        del a, b
    torch.cuda.synchronize()

    for _ in range(args.warmup):
        e.kantorovich_device(w_dev.data_ptr(), args.precision)
    torch.cuda.synchronize()
    cells_ms, tot_ms = [], []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        e.kantorovich_device(w_dev.data_ptr(), args.precision)
        info = e.eval_info()
        cells_ms.append(info["ms_cells"])
        tot_ms.append(info["ms_total"])
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_per_step = 1000.0 * wall / args.steps
    nnz = int(info["nnz"])
    V, T = p.mesh.vxy.shape[0], p.mesh.tri.shape[0]
    B = algorithmic_bytes(N, V, T, nnz)
    k_ms = float(np.mean(cells_ms))
    achieved = B / (k_ms * 1e-3) / 1e9
    traffic = None
    tfile = os.path.join(ROOT, "profiles", "k_cells_C3_traffic.json")
    if os.path.exists(tfile):
        with open(tfile) as f:
            tj = json.load(f)
        if tj.get("N") == N and tj.get("precision") == args.precision:
            traffic = tj.get("bytes_per_dispatch")

    out = {
        "metric": "N=1M: mass+Hessian evals/s and Newton solve time to 1e-8 mass err, mixed vs FP64",
        "value": 1000.0 / ms_per_step,
        "unit": "evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32+f64",
        "data": "synthetic (seeded generator, config 3)",
        "config": {"workload": f"C3: {args.grid}x{args.grid} P1 grid ({T} triangles), "
                               f"N={N} Diracs, w=0, {args.precision}",
                   "N": N, "triangles": T, "vertices": V, "nnz": nnz,
                   "precision": args.precision},
        "roofline": {"bound": "hbm", "kernel": "k_cells", "achieved": achieved,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": traffic, "algorithmic_bytes": B, "kernel_ms": k_ms},
        "eval_ms_device": float(np.mean(tot_ms)),
        "phases_ms": {k: info[k] for k in ("ms_weights", "ms_candidates", "ms_cells",
                                           "ms_assemble")},
    }
    if not args.no_solve:
        for prec in [q for q in args.solve_precisions.split(",") if q]:
            t0 = time.perf_counter()
            w, st = e.ot_solve(p.nu, stop="maxrel", tol=1e-8, precision=prec,
                               raise_on_error=False)
            dt = time.perf_counter() - t0
            out[f"solve_{prec}"] = {"seconds": dt, "status": st.status, "newton": st.niter,
                                    "evals": st.neval, "cg_iters": st.cg_iter,
                                    "max_rel_err": st.max_rel_err,
                                    "ms_eval": st.ms_eval, "ms_linear": st.ms_linear}
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline_sample(p)
    e.close()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
