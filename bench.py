"""Benchmark: demag + LLG RHS evaluations per second on one MI355X.

Workload (default): the dense-BEM cube, 150 nm per side, 30^3 hexes x 6 Kuhn
tets (29 791 nodes, 5 402 boundary nodes, dense B 5402^2), Permalloy
(Ms 8e5 A/m, A 1.3e-11 J/m, alpha 0.5), m = vortex + N(0, 0.05) noise,
renormalized (seeded).  One step = one full LLG RHS evaluation
(EffectiveField.llg_rhs_vec: Div SpMV, u1 CG, boundary matvec, u2 CG, fused
field + RHS kernel) at a magnetization taken from a pool of distinct states,
with the reference's solver settings (CG rtol 1e-6, atol 1e-12, zero initial
guess).  Inputs are resident in HBM before the timed region.

Precision modes: "fp64" (B in FP64) and "mixed" (B stored in FP32, FP64
accumulation; everything else FP64).  `value` is the mixed-mode rate; the FP64
rate and the relative RHS error of mixed vs FP64 are reported beside it.

The JSON line carries `roofline` for the boundary matvec kernel (algorithmic
bytes Nb^2*s_B + 16 Nb per launch over its event-timed duration, B flushed
from the 256 MB Infinity Cache between launches) and `cpu_baseline` (the
oracle's FP64 numpy/scipy path, LU Poisson solves, timed on the host).
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "fenicsx-micromagnetics_amd"))

HBM_PEAK_GBS = 8000.0


def vortex_m(X, seed=0, center=(75.0, 75.0), rc=10.0, noise=0.05):
    import numpy as np
    rng = np.random.default_rng(seed)
    x = X[:, 0] - center[0]
    y = X[:, 1] - center[1]
    m = np.stack([-y, x, np.full_like(x, rc)], axis=1)
    m /= np.linalg.norm(m, axis=1)[:, None]
    m += rng.normal(0.0, noise, m.shape)
    m /= np.linalg.norm(m, axis=1)[:, None]
    return m.ravel()


def time_matvec(d, reps=20, flush_mb=512):
    """Average duration of the boundary GEMV kernel (HIP events on the launch
    stream), with a cache flush (write of flush_mb) before every launch."""
    import torch
    from smx import _capi
    L = _capi.lib()
    st = torch.cuda.current_stream()
    flush = torch.empty(flush_mb * 1024 * 1024 // 4, dtype=torch.float32, device="cuda")
    x = torch.randn(d.Nb, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    B = d.B_dense_gpu
    code = _capi.DTYPE_CODE[B.dtype]
    ts = []
    for _ in range(reps + 3):
        flush.fill_(1.0)
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(st)
        _capi.check(L.smx_dense_bmatvec(B.data_ptr(), code, d.Nb, d.B_ld, x.data_ptr(), None, None, 0.0,
                                        None, y.data_ptr(), st.cuda_stream))
        e1.record(st)
        e1.synchronize()
        ts.append(e0.elapsed_time(e1) * 1e-3)
    ts = ts[3:]
    warm = []
    for _ in range(reps):
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(st)
        _capi.check(L.smx_dense_bmatvec(B.data_ptr(), code, d.Nb, d.B_ld, x.data_ptr(), None, None, 0.0,
                                        None, y.data_ptr(), st.cuda_stream))
        e1.record(st)
        e1.synchronize()
        warm.append(e0.elapsed_time(e1) * 1e-3)
    del flush
    return sum(ts) / len(ts), sum(warm) / len(warm)


def cpu_baseline(mesh, B, m_pool, Ms, A_ex, alpha, budget_s=20.0):
    """FP64 oracle RHS evaluations on the host (numpy/scipy, LU solves)."""
    import numpy as np
    sys.path.insert(0, str(ROOT / "oracle"))
    import smx_oracle as O
    t0 = time.perf_counter()
    bd = O.boundary_data(mesh.X, mesh.tets)
    dem = O.DemagOracle(mesh.X, mesh.tets, Ms, B=B, bd=bd)
    fem = O.fem_operators(mesh.X, mesh.tets, 1.0)
    Kex = O.exchange_matrix(mesh.X, mesh.tets, fem.V, A_ex, Ms)
    setup = time.perf_counter() - t0
    n = 0
    t0 = time.perf_counter()
    while True:
        mv = m_pool[n % len(m_pool)]
        H = Kex @ mv + dem.compute(mv)
        O.llg_rhs(mv, H, 2.211e5, alpha, Ms)
        n += 1
        el = time.perf_counter() - t0
        if el >= budget_s or n >= 200:
            break
    threads = os.environ.get("OMP_NUM_THREADS") or os.environ.get("OPENBLAS_NUM_THREADS") or str(os.cpu_count())
    return {"value": n / el, "unit": "evals/s", "cores": int(threads), "kind": "port",
            "sample": f"{n} FP64 RHS evals of the 150 nm cube (oracle: scipy splu Poisson solves, numpy B matvec; "
                      f"setup {setup:.1f}s excluded)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=20.0)
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1"}))
        sys.exit(2)
    import numpy as np
    import torch
    from smx import mesh as M
    from smx.llg import EffectiveField
    from smx.demag import OperatorSet

    dev = torch.device("cuda", 0)
    Ms, A_ex, alpha = 8e5, 1.3e-11, 0.5
    mesh = M.cube_c2()
    t0 = time.perf_counter()
    ops = OperatorSet(mesh, dev)
    common = dict(ksp_rtol_u1=1e-6, ksp_rtol_u2=1e-6, ksp_atol=1e-12)
    h64 = EffectiveField(mesh, Ms, A_ex, alpha=alpha, use_demag=True, operators=ops,
                         demag_kwargs=dict(B_dtype="float64", **common))
    hmx = EffectiveField(mesh, Ms, A_ex, alpha=alpha, use_demag=True, operators=ops,
                         demag_kwargs=dict(B_dtype="float32", geometry=h64.demag_field.opB, **common))
    torch.cuda.synchronize()
    setup_s = time.perf_counter() - t0
    pool = [torch.from_numpy(vortex_m(mesh.X, seed=s)).to(dev) for s in range(8)]
    out = torch.empty(3 * mesh.num_nodes, dtype=torch.float64, device=dev)

    def run(hef, steps, warmup):
        for i in range(warmup):
            hef.llg_rhs_vec(pool[i % len(pool)], out)
        torch.cuda.synchronize()
        its = []
        t0 = time.perf_counter()
        for i in range(steps):
            hef.llg_rhs_vec(pool[i % len(pool)], out)
            its.append((hef.demag_field.last_its_u1, hef.demag_field.last_its_u2))
        torch.cuda.synchronize()
        return time.perf_counter() - t0, its

    t64, its64 = run(h64, args.steps, args.warmup)
    tmx, itsmx = run(hmx, args.steps, args.warmup)
    # accuracy of mixed vs fp64 at the same m
    r64 = torch.empty_like(out)
    rmx = torch.empty_like(out)
    errs = []
    for mv in pool[:4]:
        h64.llg_rhs_vec(mv, r64)
        hmx.llg_rhs_vec(mv, rmx)
        errs.append(float(torch.linalg.norm(rmx - r64) / torch.linalg.norm(r64)))
    mv_avg64, mv_warm64 = time_matvec(h64.demag_field)
    mv_avgmx, mv_warmmx = time_matvec(hmx.demag_field)
    Nb = h64.demag_field.Nb
    bytes64 = Nb * Nb * 8 + 16 * Nb
    bytes32 = Nb * Nb * 4 + 16 * Nb
    ach = bytes32 / mv_avgmx / 1e9
    res = {
        "metric": "demag+LLG RHS evals/s (mixed: FP32 boundary operator, FP64 elsewhere)",
        "value": args.steps / tmx,
        "unit": "evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * tmx / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64 (B f32)",
        "data": "synthetic (seeded vortex + noise)",
        "config": {"workload": "cube150nm_30^3_kuhn (N=29791, Nb=5402, dense B)",
                   "cg_rtol": 1e-6, "initial_guess": "zero", "setup_s": round(setup_s, 2)},
        "fp64": {"value": args.steps / t64, "ms_per_step": 1e3 * t64 / args.steps},
        "mixed_vs_fp64_rel_err": max(errs),
        "cg_iterations_u1_u2": [float(np.mean([a for a, _ in itsmx])), float(np.mean([b for _, b in itsmx]))],
        "roofline": {"bound": "hbm", "kernel": "gemv_rows_kernel<float,4> (boundary matvec, FP32 B)",
                     "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": ach / HBM_PEAK_GBS,
                     "traffic": None, "algorithmic_bytes": bytes32, "avg_us_flushed": mv_avgmx * 1e6,
                     "avg_us_cache_warm": mv_warmmx * 1e6,
                     "fp64_achieved": bytes64 / mv_avg64 / 1e9},
    }
    if not args.no_cpu:
        Bh = h64.demag_field.B_dense_gpu.cpu().numpy()[:, :Nb]
        res["cpu_baseline"] = cpu_baseline(mesh, Bh, [p.cpu().numpy() for p in pool[:4]], Ms, A_ex, alpha,
                                           args.cpu_budget)
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
