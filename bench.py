"""Benchmark: demag + LLG RHS evaluations per second on one MI355X.

Default workload ("film"): the ACA H-matrix thin film, 500 x 500 x 10 nm at
2.5 nm spacing (200 x 200 x 4 hexes x 6 Kuhn tets: 202 005 nodes, 83 202
boundary nodes, 166 400 boundary triangles), cluster tree leaf 64, eta 2, ACA
tolerance 1e-6; Ms 8e5 A/m, A 1.3e-11 J/m, uniaxial Ku 5e5 J/m^3 along z,
interfacial DMI 3e-3 J/m^2 (n0 = z), alpha 0.5; m = seeded random unit
vectors.  "--config cube" runs the dense-BEM cube (150 nm, 30^3 hexes,
N 29 791, Nb 5 402, vortex + N(0, 0.05) noise).

One step = one LLG RHS evaluation (EffectiveField.llg_rhs_vec: Div SpMV, u1
CG, boundary matvec, u2 CG, fused field + RHS kernel) at a magnetization
taken from a pool of distinct states, with the reference's solver settings
(CG rtol 1e-6, atol 1e-12, zero initial guess).  Inputs are resident in HBM
before the timed region.  Precision modes: "fp64" (FP64 boundary operator)
and "mixed" (FP32 H-matrix factors / FP32 dense B, FP64 accumulation);
`value` is the mixed-mode rate, the FP64 rate and the relative RHS error of
mixed vs FP64 are reported beside it.

`roofline` describes the boundary matvec, the metric's named kernel and the
largest single operation of a film step (hm_gather + hm_far_z + hm_leaf):
algorithmic bytes per matvec (factor entries x 4 B + near nonzeros x 12 B +
the vectors) divided by its average duration, measured with HIP events on the
launch stream after a read-only sweep that evicts the Infinity Cache;
`traffic` is the HBM bytes of the same matvec from two rocprofv3 counter
passes (FETCH_SIZE, WRITE_SIZE) of `--pmc-probe matvec`, scaled by a plain
copy of known size.  `cg_roofline` gives one AMG-CG iteration of the u1 solve.
`cpu_baseline` times the FP64 oracle path on the host on a bounded sample.

Other modes: `--config cube` (dense B, plus the 64-RHS matrix-core product),
`--config disk` (energy-gradient check), `--config adjoint` (gradient through
500 Dormand-Prince steps), `--config paths` (torque terms, implicit BDF relax,
minimizer, H-matrix build per compressor, with their kernel rooflines and host
baselines).
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "fenicsx-micromagnetics_amd"))

HBM_PEAK_GBS = 8000.0
GAMMA = 2.211e5


PC_TYPE = "auto"
PC_TYPE_U2 = None


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


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


def random_m(n, seed=0):
    import numpy as np
    g = np.random.default_rng(seed).standard_normal((n, 3))
    return (g / np.linalg.norm(g, axis=1)[:, None]).ravel()


def event_time(fn, reps, flush=None):
    """Average seconds of fn() measured by HIP events on the current stream."""
    import torch
    st = torch.cuda.current_stream()
    ts = []
    for _ in range(reps):
        if flush is not None:
            flush.sum()  # read-only sweep: evicts the Infinity Cache without leaving dirty lines
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(st)
        fn()
        e1.record(st)
        e1.synchronize()
        ts.append(e0.elapsed_time(e1) * 1e-3)
    ts = sorted(ts)[: max(1, len(ts) - len(ts) // 5)]
    return sum(ts) / len(ts)


def amg_iteration_bytes(d):
    """Algorithmic bytes of one AMG-CG iteration of the u1 solve: every
    operator entry once (4-byte column + value), every vector read or written
    by a launch once (the restriction's reads of r, w, s are the update's)."""
    h = d.amg_u1
    vs = 4 if h.fp32 else 8
    N = d.N
    lv = h.host
    op = lambda M: M.nnz * (4 + vs) + 4 * (M.shape[0] + 1)
    nb = 80 * N                                        # update: r, w, s, u, p, x read; r, s, p, x written
    nb += d.A_u1.nnz * 12 + 16 * N                      # w = A u: matrix, u read, w written
    for l, L in enumerate(lv[:-1]):
        if l == 0:   # restriction by the carried coarse vectors: b_1 and z_1 read and written
            nb += op(L.R) + 32 * lv[1].n
        else:
            nb += op(L.R) + 8 * lv[l + 1].n + 8 * L.n
        nb += op(L.S) + op(L.Q) + 8 * (2 * L.n + lv[l + 1].n)                # up: S b + Q x'
    c = lv[-1]
    nb += c.S.nnz * vs + 4 * (c.n + 1) + 16 * c.n       # coarse: dense rows, no column indices
    return nb


def eval_breakdown(hef, pool, reps=10):
    """Device time of each stage of one demag + RHS evaluation (events on the
    launch stream between the stages of the product's own sequence, as
    DemagFieldLindholmGPU.solve_potential and the fused field kernel run it,
    speculative solves): divergence, u1 solve, boundary matvec, u2 solve,
    field kernel.  Averages over `reps` evaluations (ms)."""
    import torch
    d = hef.demag_field
    names = ("div", "u1_solve", "boundary_matvec", "u2_solve", "field_rhs")
    acc = dict.fromkeys(names, 0.0)
    out = torch.empty_like(pool[0])
    for i in range(reps + 1):
        m = pool[i % len(pool)]
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(6)]
        spec = d.cg_speculative and d._its_known
        ev[0].record()
        d.Div.spmv(m, d.b_u1, kind=1, partials=d.b_part)
        ev[1].record()
        d.solve("u1", 0, d.u1, d.ws1, b=d.b_u1, b_part=d.b_part, expected=d.last_its_u1, speculative=spec)
        ev[2].record()
        d.boundary_matvec(d.u1, d.y_rows, gather=d.bverts_gpu, shift_ptr=d.u1_sum_ptr, shift_scale=1.0 / d.N)
        ev[3].record()
        d.solve("u2", 1, d.u2, d.ws2, gb=d.y_rows, project=0, expected=d.last_its_u2, speculative=spec)
        ev[4].record()
        hef._launch(hef._args(m, h_eff=hef.H_eff_gpu, rhs=out))
        ev[5].record()
        ev[5].synchronize()
        if i == 0:
            continue                                   # warm-up
        for k, nm in enumerate(names):
            acc[nm] += ev[k].elapsed_time(ev[k + 1]) / reps
    acc["total"] = sum(acc[nm] for nm in names)
    return acc


def cg_iteration_timing(d, K=200):
    """Average duration of one CG iteration (u1 operator), events on the
    launch stream: (solve with K iterations - solve with 0) / K.  For the AMG
    preconditioner an iteration is its whole launch sequence."""
    import torch
    from smx import _capi
    L = _capi.lib()
    A = d.A_u1
    raw = d.A_raw
    x = torch.zeros_like(d.u1)
    its = C.c_int32(0)
    rn = C.c_double(0.0)
    sums = (C.c_double * 1)()
    if d.pc_type == "gamg":
        K = min(K, 20)                                  # stays well above round-off for both workloads
        h = d.amg_u1

        def solve_amg(k):
            L.smx_pcg_amg(d.N, 0, A.width, A.ld, A.cols.data_ptr(), A.vals.data_ptr(), 0,
                          d.b_u1.data_ptr(), d.b_part.data_ptr(), d.b_part.numel(), 1, raw.width, raw.ld,
                          raw.cols.data_ptr(), raw.vals.data_ptr(), None, None, x.data_ptr(), 0,
                          _capi.ptr(d.cg_perm), C.addressof(h.c_levels), h.nlev, h.val_dtype, 0.0, 0.0, k,
                          10 ** 6, k, d.ws1.data_ptr(), C.byref(its), C.byref(rn), sums, _capi.stream_ptr())

        t_k = event_time(lambda: solve_amg(K), 5)
        t_0 = event_time(lambda: solve_amg(0), 5)
        return (t_k - t_0) / K, amg_iteration_bytes(d)

    def solve(k):
        L.smx_pcg_jacobi(d.N, 0, A.width, A.ld, A.cols.data_ptr(), A.vals.data_ptr(), d.dinv_u1.data_ptr(), 0,
                         d.b_u1.data_ptr(), d.b_part.data_ptr(), d.b_part.numel(), 1, raw.width, raw.ld,
                         raw.cols.data_ptr(), raw.vals.data_ptr(), None, None, x.data_ptr(), 0,
                         _capi.ptr(d.cg_perm), 0.0, 0.0, k, 10 ** 6,
                         k, d.ws1.data_ptr(), C.byref(its), C.byref(rn), sums, _capi.stream_ptr())

    t_k = event_time(lambda: solve(K), 5)
    t_0 = event_time(lambda: solve(0), 5)
    per = (t_k - t_0) / K
    nbytes = A.nnz * 12 + 88 * d.N
    return per, nbytes


def matvec_timing(d, flush_mb=512, reps=12):
    """Boundary matvec: (avg s after a cache-evicting sweep, avg s warm,
    algorithmic bytes, kernel names).  reps=0: bytes and names only."""
    import torch
    flush = torch.ones(flush_mb * 1024 * 1024 // 4, dtype=torch.float32, device="cuda") if reps else None
    x = torch.randn(d.Nb, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    if d.boundary_backend == "dense":
        from smx import _capi
        L = _capi.lib()
        B = d.B_dense_gpu
        code = _capi.DTYPE_CODE[B.dtype]
        fn = lambda: _capi.check(L.smx_dense_bmatvec(B.data_ptr(), code, d.Nb, d.B_ld, x.data_ptr(), None, None,
                                                     0.0, None, y.data_ptr(), _capi.stream_ptr()))
        nbytes = d.Nb * d.Nb * B.element_size() + 16 * d.Nb
        name = f"gemv_rows_kernel<{str(B.dtype).split('.')[-1]}> (dense boundary matvec)"
    else:
        H = d.B_hmatrix_gpu
        fn = lambda: H.matvec(x, y)
        nbytes = H.algorithmic_bytes()
        where = "near launch" if H.near_in_z else "leaf launch"
        log(f"near rows in the {where}; set-up times {H.near_mode_seconds}")
        near_k = ("hm_near_lds_kernel + " if H._near_by_task() else "hm_near_kernel + ") if H.near_in_z else ""
        name = (f"hm_gather_kernel + {near_k}hm_far_z_kernel + hm_leaf_kernel "
                f"(H-matrix boundary matvec; near rows in the {where}, chosen at set-up)")
    if not reps:
        return None, None, nbytes, name
    cold = event_time(fn, reps, flush)
    warm = event_time(fn, reps)
    del flush
    return cold, warm, nbytes, name


def energy_gradient_disk(n_dir=64, h=1e-6):
    """Config 4: gradient of E = E_ex + E_an + E_d w.r.t. m, per-cell Ms and A on
    the 200 nm x 20 nm disk (3 nm edges), Ms_c ~ U[6e5, 1e6], A_c ~ U[1e-11,
    1.5e-11], Ku 5e5 along z; central differences (step h) on `n_dir` random
    directions in (m, Ms, A) jointly; FP64 and FP32 dense boundary operators."""
    import numpy as np
    import torch
    from smx import mesh as M
    from smx.adjoint import MicromagneticEnergy
    dev = torch.device("cuda", 0)
    mesh = M.disk(diameter=200.0, height=20.0, h=3.0)
    T = int(mesh.tets.shape[0])
    rng = np.random.default_rng(0)
    Ms = torch.from_numpy(rng.uniform(6e5, 1e6, T)).to(dev)
    A = torch.from_numpy(rng.uniform(1e-11, 1.5e-11, T)).to(dev)
    m = torch.from_numpy(random_m(mesh.num_nodes, seed=1)).to(dev)
    out = {"metric": "energy adjoint: dE/d(m, Ms_cell, A_cell) vs central differences", "config": {
        "workload": f"disk200x20nm_h3 (N={mesh.num_nodes}, T={T})", "directions": n_dir, "h": h,
        "Ku": 5e5, "cg_rtol": 1e-13}}
    grads = {}
    for name, bdt in (("fp64", "float64"), ("fp32_boundary", "float32")):
        model = MicromagneticEnergy(mesh, Ku=5e5, n_ani_vec=(0.0, 0.0, 1.0), ksp_rtol=1e-13, device=dev,
                                    demag_kwargs=dict(B_dtype=bdt))
        model.value_and_grad(m, Ms, A)                  # warm-up (allocations)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        E, (gm, gMs, gA) = model.value_and_grad(m, Ms, A)
        torch.cuda.synchronize()
        t_grad = time.perf_counter() - t0
        grads[name] = (gm.clone(), gMs.clone(), gA.clone())
        out[name] = {"E_total_J": E["E_total"], "value_and_grad_s": t_grad, "Nb": model.demag.Nb,
                     "boundary": model.demag.boundary_backend}
        if name == "fp64":
            errs = []
            for _ in range(n_dir):
                dm = torch.from_numpy(rng.standard_normal(3 * mesh.num_nodes)).to(dev)
                dM = torch.from_numpy(rng.standard_normal(T)).to(dev) * 1e5
                dA = torch.from_numpy(rng.standard_normal(T)).to(dev) * 1e-12
                Ep = model.forward(m + h * dm, Ms + h * dM, A + h * dA)["E_total"]
                Em = model.forward(m - h * dm, Ms - h * dM, A - h * dA)["E_total"]
                fd = (Ep - Em) / (2 * h)
                an = float((gm * dm).sum() + (gMs * dM).sum() + (gA * dA).sum())
                errs.append(abs(fd - an) / abs(an))
            out["fp64"]["fd_rel_err_max"] = max(errs)
            out["fp64"]["fd_rel_err_median"] = float(np.median(errs))
            log(f"disk FD over {n_dir} directions: max {max(errs):.2e} median {np.median(errs):.2e}")
        del model
        torch.cuda.empty_cache()
    rel = lambda a, b: float(torch.linalg.norm(a - b) / torch.linalg.norm(b))
    out["fp32_boundary_vs_fp64_grad_rel_err"] = {
        k: rel(a, b) for k, a, b in zip(("m", "Ms", "A"), grads["fp32_boundary"], grads["fp64"])}
    print(json.dumps(out), flush=True)


def adjoint_through_time(steps=500, dt=1e-13, every=50, n_fd=2):
    """Gradient of <m_z>(t_end) w.r.t. per-cell Ms after `steps` fixed DP5 steps
    on the 150 nm cube (alpha 0.02, vortex start, Ms 8e5 A/m in every cell),
    checkpoint every `every` steps.  FP64 boundary operator (the reference
    gradient) and FP32 boundary operator (mixed forward); central differences
    (relative step 1e-6) on `n_fd` random directions of the FP64 model."""
    import numpy as np
    import torch
    from smx import mesh as M
    from smx.llg_adjoint import LLGAdjointModel, integrate, mean_mz
    dev = torch.device("cuda", 0)
    mesh = M.cube_c2()
    m0 = torch.from_numpy(vortex_m(mesh.X, seed=0)).to(dev)
    out = {"metric": "adjoint through time: dJ/dMs_cell, J = <m_z>(t_end)", "config": {
        "workload": "cube150nm_30x30x30_kuhn (N=29791, T=162000 cells, Nb=5402, dense B)", "steps": steps,
        "dt_s": dt, "alpha": 0.02, "checkpoint_every": every, "Ms_A_per_m": 8e5, "A_J_per_m": 1.3e-11,
        "cg_rtol": 1e-10}}
    grads = {}
    for name, bdt in (("fp64", "float64"), ("fp32_boundary", "float32")):
        model = LLGAdjointModel(mesh, A_ex=1.3e-11, alpha=0.02, ksp_rtol=1e-10, device=dev,
                                demag_kwargs=dict(B_dtype=bdt, ksp_max_it_u1=20000, ksp_max_it_u2=20000))
        Ms = torch.full((model.T,), 8e5, dtype=torch.float64, device=dev, requires_grad=True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        with torch.no_grad():
            Jf = float(mean_mz(model, integrate(model, m0, Ms.detach(), dt, steps, every)))
        torch.cuda.synchronize()
        t_fwd = time.perf_counter() - t0
        model.rhs_evals = model.vjp_evals = 0
        t0 = time.perf_counter()
        J = mean_mz(model, integrate(model, m0, Ms, dt, steps, every))
        (g,) = torch.autograd.grad(J, Ms)
        torch.cuda.synchronize()
        t_grad = time.perf_counter() - t0
        grads[name] = g
        out[name] = {"J": Jf, "forward_s": t_fwd, "forward_and_gradient_s": t_grad,
                     "rhs_evals": model.rhs_evals, "rhs_vjps": model.vjp_evals,
                     "grad_norm": float(torch.linalg.norm(g))}
        log(f"{name}: J={Jf:.12f} forward {t_fwd:.1f}s gradient {t_grad:.1f}s")
        if name == "fp64":
            rng = np.random.default_rng(11)
            errs = []
            with torch.no_grad():
                for _ in range(n_fd):
                    w = torch.from_numpy(rng.standard_normal(model.T)).to(dev) * 8e5
                    h = 1e-6
                    Jp = float(mean_mz(model, integrate(model, m0, Ms.detach() + h * w, dt, steps, every)))
                    Jm = float(mean_mz(model, integrate(model, m0, Ms.detach() - h * w, dt, steps, every)))
                    fd = (Jp - Jm) / (2 * h)
                    an = float(g @ w)
                    errs.append(abs(fd - an) / abs(an))
            out["fp64"]["fd_rel_err"] = errs
            log(f"fp64 central-difference check: {errs}")
        del model
        torch.cuda.empty_cache()
    out["fp32_boundary_vs_fp64_grad_rel_err"] = float(torch.linalg.norm(grads["fp32_boundary"] - grads["fp64"])
                                                      / torch.linalg.norm(grads["fp64"]))
    print(json.dumps(out), flush=True)


def _kernel_roofline(name, fn, nbytes, flush, reps=20):
    """HIP-event time of one launch (Infinity Cache evicted before each) and
    its HBM roofline from the algorithmic bytes."""
    sec = event_time(fn, reps, flush=flush)
    return {"kernel": name, "bound": "hbm", "avg_us": sec * 1e6, "algorithmic_bytes": int(nbytes),
            "achieved": nbytes / sec / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": nbytes / sec / 1e9 / HBM_PEAK_GBS}


def extended_paths(cpu_budget=20.0, no_cpu=False):
    """Measurements of the paths beside the demag + RHS evaluation: torque
    terms in the fused field kernel, the implicit BDF Newton-Krylov relax and
    its kernels, the minimizer, and H-matrix construction per compressor."""
    import numpy as np
    import torch
    from smx import mesh as M
    from smx.demag import OperatorSet
    from smx.implicit import GMRES, JvContext, relax_implicit
    from smx.llg import EffectiveField
    from smx.minimizer import LaBonteBBMinimizerGPU
    dev = torch.device("cuda", 0)
    flush = torch.empty(64 * 1024 * 1024, dtype=torch.float64, device=dev)
    flush.fill_(1.0)
    out = {"metric": "extended paths: torque RHS, implicit relax, minimizer, H-matrix build", "n_gpus": 1,
           "data": "synthetic (seeded)"}

    # ---- torque terms in the fused field kernel (film, no demag) ----
    film = M.film_c3()
    ops = OperatorSet(film, dev)
    base = dict(Ms=8e5, Aex=1.3e-11, Ku=5e5, n_ani_vec=(0.0, 0.0, 1.0), D_int=3e-3, n0_int_vec=(0.0, 0.0, 1.0),
                alpha=0.5)
    extra_terms = dict(Kc1=-4.5e4, u1_cub=(1.0, 0.0, 0.0), u2_cub=(0.0, 1.0, 0.0), Jmagnitude=1e12,
                       Jdir_vec=(1.0, 0.0, 0.0), P=0.6, beta=0.04, H_dl=2e4, H_fl=-7e3,
                       polarization_vec=(0.0, 1.0, 0.0))
    hb = EffectiveField(film, operators=ops, device=dev, **base)
    ha = EffectiveField(film, operators=ops, device=dev, **base, **extra_terms)
    N = film.num_nodes
    nnz = int(ops.graph.nnz)
    m = torch.from_numpy(random_m(N, seed=1)).to(dev)
    r = torch.empty_like(m)
    v = torch.from_numpy(random_m(N, seed=2)).to(dev)
    jv = torch.empty_like(m)
    dmi_b = 8 * int(np.count_nonzero(hb.dmi_off >= 0))   # DMI bytes per graph slot (compact layout)
    rb = _kernel_roofline("field_rhs_fast_kernel (exchange, DMI, uniaxial)", lambda: hb.llg_rhs_vec(m, r),
                          nnz * (4 + 8 + dmi_b) + N * 72, flush)
    ra = _kernel_roofline("field_rhs_kernel (+ cubic, Zhang-Li STT, SOT)", lambda: ha.llg_rhs_vec(m, r),
                          nnz * (4 + 8 + dmi_b + 8) + N * 72, flush)
    ha.update_jac_state(m)
    rj = _kernel_roofline("llg_jvp_kernel (all terms)", lambda: ha.jac_vec_times_vec(v, jv),
                          nnz * (4 + 8 + dmi_b + 8) + N * 96, flush)
    hb.update_jac_state(m)
    jb = _kernel_roofline("llg_jvp_fast_kernel (exchange, DMI, uniaxial)", lambda: hb.jac_vec_times_vec(v, jv),
                          nnz * (4 + 8 + dmi_b) + N * 96, flush)
    rb["dmi_components_stored"] = dmi_b // 8
    tor = {"workload": f"film500x500x10nm (N={N}, graph nnz {nnz}), no demag", "rhs_base": rb, "rhs_all_terms": ra,
           "jvp_base": jb, "jvp_all_terms": rj, "value": 1.0 / (ra["avg_us"] * 1e-6),
           "unit": "field+RHS evals/s (all terms)"}
    if not no_cpu:
        sys.path.insert(0, str(ROOT / "oracle"))
        import smx_oracle as O
        log("torques: host operator assembly for the CPU port")
        t0 = time.perf_counter()
        fem = O.fem_operators(film.X, film.tets, 8e5)
        Kl = (O.exchange_matrix(film.X, film.tets, fem.V, 1.3e-11, 8e5)
              + O.dmi_interfacial_matrix(film.X, film.tets, fem.V, 3e-3, 8e5, (0.0, 0.0, 1.0))).tocsr()
        KJ = O.zhang_li_matrix(film.X, film.tets, fem.V, np.tile([1.0, 0.0, 0.0], (N, 1))).tocsr()
        axes = O.cubic_axes_oracle((1.0, 0.0, 0.0), (0.0, 1.0, 0.0))
        setup = time.perf_counter() - t0
        mh = m.cpu().numpy()
        n_ev, t0 = 0, time.perf_counter()
        while time.perf_counter() - t0 < min(cpu_budget, 10.0) and n_ev < 200:
            H = Kl @ mh + O.anisotropy_field(mh, (0.0, 0.0, 1.0), 5e5, 8e5) + O.cubic_field(mh, axes, -4.5e4, 8e5)
            Z = np.column_stack([KJ @ mh.reshape(-1, 3)[:, k] for k in range(3)]).ravel()
            O.llg_rhs_torques(mh, H, GAMMA, 0.5, 8e5, Z=Z, stt=(1e12, 0.6, 0.04), p=(0.0, 1.0, 0.0),
                              H_dl=2e4, H_fl=-7e3)
            n_ev += 1
        el = time.perf_counter() - t0
        tor["cpu_baseline"] = {"value": n_ev / el, "unit": "field+RHS evals/s (all terms)", "cores": 1,
                               "kind": "port", "sample": f"{n_ev} oracle evaluations (scipy sparse + numpy) in "
                                                         f"{el:.1f}s, operator assembly {setup:.1f}s excluded"}
    out["torques"] = tor
    log(f"torques: rhs {ra['avg_us']:.1f} us ({ra['frac']:.2f} of HBM), jvp {rj['avg_us']:.1f} us")

    # ---- implicit-path kernels at film size ----
    ctx = JvContext(ha)
    ctx.update_pc_full_fast_gpu(2e12, exact=True, use_abs_kappa=False)
    y = torch.empty_like(m)
    kr = {"pc_build": _kernel_roofline("pc_build_kernel", lambda: ctx.update_pc_full_fast_gpu(
              2e12, exact=True, use_abs_kappa=False), nnz * 4 + N * (8 + 72 + 24 + 24 + 72), flush),
          "block3_apply": _kernel_roofline("block3_apply_kernel", lambda: ctx.apply(v, y), N * 120, flush)}
    g = GMRES(3 * N, dev)
    from smx import _capi
    L = _capi.lib()
    k = 31
    g.V[:k].normal_()
    kr["multi_dot_31"] = _kernel_roofline(
        "multi_dot_kernel + reduce (31 vectors)",
        lambda: _capi.check(L.smx_multi_dot(3 * N, k, g.V.data_ptr(), 3 * N, v.data_ptr(), g.h_dev.data_ptr(),
                                            g.work.data_ptr(), _capi.stream_ptr())), (k + 1) * 8 * 3 * N, flush)
    coef = torch.randn(k, dtype=torch.float64, device=dev)       # one coefficient per vector
    assert g.V.shape[0] >= k and g.h_dev.numel() >= k and g.V.shape[1] == 3 * N
    kr["lincomb_31"] = _kernel_roofline(
        "lincomb_kernel (31 vectors)",
        lambda: _capi.check(L.smx_lincomb(3 * N, k, g.V.data_ptr(), 3 * N, coef.data_ptr(), 1.0, y.data_ptr(),
                                          _capi.stream_ptr())), (k + 2) * 8 * 3 * N, flush)
    del g, coef

    # ---- implicit relax (150 nm cube, dense FP64 B) ----
    cube = M.cube_c2()
    mat = dict(Ms=8e5, Aex=1.3e-11, Ku=5e5, n_ani_vec=(0.0, 0.0, 1.0), alpha=0.5)
    hc = EffectiveField(cube, use_demag=True, device=dev, demag_kwargs=dict(
        B_dtype="float64", ksp_rtol_u1=1e-6, ksp_rtol_u2=1e-6, ksp_atol=1e-12), **mat)
    m0 = vortex_m(cube.X, seed=0)
    t_final = 50e-12
    torch.cuda.synchronize()
    st = relax_implicit(hc, t_final, dt_init=1e-14, m0_array=m0)
    torch.cuda.synchronize()
    imp = {"workload": "cube150nm_30x30x30_kuhn (N=29791, Nb=5402, dense FP64 B), exchange + uniaxial + demag, "
                       "alpha 0.5, vortex start, BDF to 50 ps (ts_rtol = ts_atol = 1e-6, snes_rtol 1e-2, "
                       "ksp_rtol 1e-4, exact block preconditioner)",
           "stats": {k2: (None if isinstance(v2, float) and v2 != v2 else v2) for k2, v2 in st.items()},
           "value": st["steps"] / st["elapsed_s"], "unit": "BDF steps/s", "kernels_at_film_size": kr}
    if not no_cpu:
        sys.path.insert(0, str(ROOT / "oracle"))
        import smx_oracle as O
        B_host = hc.demag_field.B_dense_gpu.cpu().numpy()[:, :hc.demag_field.opB.Nb]
        dem = O.DemagOracle(cube.X, cube.tets, mat["Ms"], B=B_host)
        t0 = time.perf_counter()
        _, t_c, cst = O.llg_bdf_relax(cube.X, cube.tets, m0, t_final, 1e-14, mat["Ms"], mat["Aex"], mat["alpha"],
                                      Ku=mat["Ku"], n_ani=mat["n_ani_vec"], dem=dem, budget_s=cpu_budget)
        el_c = time.perf_counter() - t0
        imp["cpu_baseline"] = {"value": cst["steps"] / el_c, "unit": "BDF steps/s", "cores": 1, "kind": "port",
                               "sample": f"first {cst['steps']} BDF steps (to t = {t_c:.3e} s) of the same run by "
                                         f"the oracle's BDF Newton-Krylov (splu Poisson solves, numpy GMRES and "
                                         f"block preconditioner) in {el_c:.1f}s; {cst}"}
    out["implicit"] = imp
    log(f"implicit: {st['steps']} steps in {st['elapsed_s']:.2f}s ({imp['value']:.0f} steps/s)")

    # ---- minimizer (150 nm cube with demag) ----
    hc.set_m_from_cpu(random_m(cube.num_nodes, seed=5))
    mz = LaBonteBBMinimizerGPU(hc)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    res = mz.minimize(max_iter=400, tol=1e-3)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    mr = {"projected_direction": _kernel_roofline("projdir_kernel", lambda: _capi.check(L.smx_projected_direction(
              N, m.data_ptr(), v.data_ptr(), y.data_ptr(), _capi.stream_ptr())), N * 72, flush),
          "cayley_step": _kernel_roofline("cayley_kernel", lambda: _capi.check(L.smx_cayley_step(
              N, m.data_ptr(), v.data_ptr(), 0.1, y.data_ptr(), _capi.stream_ptr())), N * 72, flush)}
    out["minimizer"] = {"workload": "cube150nm_30x30x30_kuhn with demag, random start, LaBonte/BB adaptive, "
                                    "400 iterations at most", "iterations": res["iterations"],
                        "stop_reason": res["stop_reason"], "elapsed_s": el,
                        "value": res["iterations"] / el, "unit": "minimizer iterations/s",
                        "kernels_at_film_size": mr}
    log(f"minimizer: {res['iterations']} iterations in {el:.2f}s")
    del hc

    # ---- H-matrix construction per compressor (150 nm cube) ----
    from smx.boundary import build_boundary_geometry
    from smx.demag import assemble_dense_B
    from smx.hmatrix import HMatrixBuildConfig, HMatrixGPUPackedFused, build_hmatrix_device
    geo = build_boundary_geometry(cube.X, cube.tets)
    Bd, _ = assemble_dense_B(geo, dev)
    Bh = Bd[:, :geo.Nb]
    hb_out = {}
    for comp in ("fullaca", "svd", "rsvd"):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        data = build_hmatrix_device(geo, HMatrixBuildConfig(epsilon=1e-6, leaf_size=64, compressor=comp), dev,
                                    verbose=True)
        torch.cuda.synchronize()
        tb = time.perf_counter() - t0
        Hm_ = HMatrixGPUPackedFused(data, use_fp32=False)
        x = torch.from_numpy(np.random.default_rng(1234).standard_normal(geo.Nb)).to(dev)
        ref = Bh @ x
        err = float(torch.linalg.norm(Hm_.matvec(x) - ref) / torch.linalg.norm(ref))
        md = data.metadata
        hb_out[comp] = {"build_s": tb, "far_blocks": md["low_rank_blocks"], "dense_blocks": md["dense_blocks"],
                        "average_far_rank": md["average_far_rank"], "far_storage_entries": md["far_storage_entries"],
                        "matvec_rel_err_vs_dense": err}
        log(f"hmatrix {comp}: build {tb:.2f}s far {md['low_rank_blocks']} rank {md['average_far_rank']:.1f} "
            f"err {err:.2e}")
        del Hm_, data
    out["hmatrix_build"] = {"workload": "cube150nm (Nb=5402), leaf 64, eta 2, eps 1e-6", **hb_out}
    print(json.dumps(out), flush=True)


def measured_traffic(config, what="matvec", variant=None):
    """HBM-side bytes per launch (or per boundary matvec) from the committed
    counter passes (FETCH_SIZE and WRITE_SIZE of `bench.py --pmc-probe WHAT`,
    each scaled by the same counter's reading for a 256 MiB device copy), or
    None when no pass for this workload is committed.  `variant` names a launch
    variant of the same workload (the H-matrix matvec's "near_launch")."""
    f = ROOT / "profiles" / {"matvec": "matvec_traffic.json", "cg": "cg_iter_traffic.json",
                             "amg": "amg_iter_traffic.json"}[what]
    if not f.exists():
        return None
    rec = json.loads(f.read_text()).get(config if variant is None else f"{config}:{variant}")
    return None if rec is None else int(rec["traffic_bytes_per_launch"])


def pmc_probe_amg(config, K=40):
    """Counter-pass program for the AMG-CG iteration of the u1 solve: the
    workload's Poisson operator and AMG hierarchy (FP32 values, as in the mixed
    evaluation) are built, one 256 MiB device copy runs (calibration), then one
    solve of exactly K iterations (smx_pcg_amg, the initial V-cycle included;
    the parser drops the solve's first seven launches, its initialisation)."""
    import numpy as np
    import torch
    from smx import _capi
    from smx import mesh as M
    from smx.amg import AmgHierarchy
    from smx.demag import OperatorSet
    from smx.fem import assemble_fem
    dev = torch.device("cuda", 0)
    src = torch.ones(64 * 1024 * 1024, dtype=torch.float32, device=dev)
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    dst.copy_(src)
    torch.cuda.synchronize()
    mesh = M.film_c3() if config == "film" else M.cube_c2()
    fem = assemble_fem(mesh.X, mesh.tets)
    ops = OperatorSet(mesh, dev, fem)
    perm = ops.solver_order()
    A = ops.compact(fem.stiff, perm)
    h = AmgHierarchy(ops.compact_host(fem.stiff, perm), dev, fp32=True)
    b = np.random.default_rng(0).standard_normal(fem.N)
    b -= b.mean()
    bd = torch.from_numpy(b).to(dev)
    bpart = torch.zeros(1, dtype=torch.float64, device=dev)
    pd = torch.from_numpy(perm.astype(np.int32)).to(dev)
    x = torch.zeros(fem.N, dtype=torch.float64, device=dev)
    L = _capi.lib()
    ws = torch.zeros(int(L.smx_pcg_amg_ws_size(fem.N)), dtype=torch.float64, device=dev)
    its, rn, sums = C.c_int32(0), C.c_double(0.0), (C.c_double * 1)()
    L.smx_pcg_amg(fem.N, 0, A.width, A.ld, A.cols.data_ptr(), A.vals.data_ptr(), 0, bd.data_ptr(), bpart.data_ptr(), 1,
                  1, A.width, A.ld, A.cols.data_ptr(), A.vals.data_ptr(), None, None, x.data_ptr(), 0, pd.data_ptr(),
                  C.addressof(h.c_levels), h.nlev, h.val_dtype, 0.0, 0.0, K, 10 ** 6, K, ws.data_ptr(),
                  C.byref(its), C.byref(rn), sums, _capi.stream_ptr())
    torch.cuda.synchronize()
    print(json.dumps({"pmc_probe": "amg", "config": config, "N": fem.N, "cg_iterations": K,
                      "levels": h.describe(), "calibration_copy_bytes": 2 * src.numel() * 4}), flush=True)


def pmc_probe_matvec(config, reps=20):
    """Counter-pass program for the boundary matvec: the workload's operator is
    built (mixed precision: FP32 factors / FP32 dense B), then one 256 MiB
    device copy (the calibration kernel: 256 MiB read + 256 MiB written), then
    `reps` boundary matvecs."""
    import torch
    dev = torch.device("cuda", 0)
    mesh, mat, h64, hmx, pool, work, extra, setup_s = build_case(config, dev)
    d = hmx.demag_field
    src = torch.ones(64 * 1024 * 1024, dtype=torch.float32, device=dev)
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    dst.copy_(src)
    torch.cuda.synchronize()
    x = torch.randn(d.Nb, dtype=torch.float64, device=dev)
    y = torch.empty_like(x)
    # the H-matrix runs both near-row variants: `reps` matvecs with the near
    # rows in the leaf launch, then `reps` with them in a launch of their own
    H = d.B_hmatrix_gpu if d.boundary_backend == "hmatrix" else None
    modes = (False, True) if H is not None else (None,)
    for mode in modes:
        if H is not None:
            H.set_near_mode(mode)
        for _ in range(reps):
            d.boundary_matvec(x, y)
        torch.cuda.synchronize()
    _, _, nbytes, name = matvec_timing(d, reps=0)
    print(json.dumps({"pmc_probe": config, "what": "matvec", "matvecs_per_variant": reps,
                      "variants": (["near rows in the leaf launch", "near rows in their own launch"]
                                   if H is not None else [name]),
                      "algorithmic_bytes_per_matvec": int(nbytes),
                      "calibration_copy_bytes": 2 * src.numel() * 4}), flush=True)


def pmc_probe(config):
    """Short program for hardware-counter passes: one 256 MiB device copy (the
    calibration kernel of known traffic: 256 MiB read + 256 MiB written), then
    200 Jacobi CG iterations of the workload's u1 operator (the same ELL matrix
    and launch shape as an evaluation; no boundary operator is built)."""
    import numpy as np
    import torch
    from smx import _capi
    from smx import mesh as M
    from smx.device import to_dev
    from smx.fem import assemble_fem
    dev = torch.device("cuda", 0)
    src = torch.ones(64 * 1024 * 1024, dtype=torch.float32, device=dev)
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    dst.copy_(src)
    torch.cuda.synchronize()
    mesh = M.film_c3() if config == "film" else M.cube_c2()
    fem = assemble_fem(mesh.X, mesh.tets)
    from smx.demag import OperatorSet, solver_operator
    ops = OperatorSet(mesh, dev, fem)
    perm = ops.solver_order()                     # the demag solver's row order
    A_plain = ops.compact(fem.stiff, perm)
    g = fem.graph
    r = np.repeat(np.arange(fem.N), np.diff(g.indptr))
    dg = np.zeros(fem.N)
    on = r == g.indices
    dg[r[on]] = fem.stiff[on]
    dinv = to_dev(1.0 / dg[perm], dev)
    A = solver_operator(A_plain, dinv)            # A D^-1, diagonal in slot 0
    pd = to_dev(perm.astype(np.int32), dev)
    b = np.random.default_rng(0).standard_normal(fem.N)
    b -= b.mean()
    bd = to_dev(b, dev)
    bpart = to_dev(np.zeros(1), dev)
    x = torch.zeros(fem.N, dtype=torch.float64, device=dev)
    L = _capi.lib()
    ws = torch.zeros(int(L.smx_pcg_ws_size(fem.N)), dtype=torch.float64, device=dev)
    its, rn, sums = C.c_int32(0), C.c_double(0.0), (C.c_double * 1)()
    K = 200
    L.smx_pcg_jacobi(fem.N, 0, A.width, A.ld, A.cols.data_ptr(), A.vals.data_ptr(), dinv.data_ptr(), 0,
                     bd.data_ptr(), bpart.data_ptr(), 1, 1, A_plain.width, A_plain.ld, A_plain.cols.data_ptr(),
                     A_plain.vals.data_ptr(),
                     None, None, x.data_ptr(), 0, pd.data_ptr(), 0.0, 0.0, K, 10 ** 6, K, ws.data_ptr(), C.byref(its),
                     C.byref(rn), sums, _capi.stream_ptr())
    torch.cuda.synchronize()
    print(json.dumps({"pmc_probe": config, "N": fem.N, "ell_nnz": A.nnz, "cg_iterations": K,
                      "algorithmic_bytes_per_iteration": A.nnz * 12 + 88 * fem.N,
                      "calibration_copy_bytes": 2 * src.numel() * 4}), flush=True)


def multi_rhs_timing(d, k=64, flush_mb=512):
    """Dense boundary product for k right-hand sides on the matrix cores
    (FP64, FP32 and BF16 storage of the same B), flushed, HIP events."""
    import numpy as np
    import torch
    from smx import _capi
    from smx.demag import dense_bmatmul
    L = _capi.lib()
    n, ld = d.Nb, d.B_ld
    B64 = d.B_dense_gpu if d.B_dense_gpu.dtype == torch.float64 else None
    if B64 is None:
        return None
    X = torch.from_numpy(np.random.default_rng(0).standard_normal((n, k))).to(B64.device)
    ref = B64[:, :n] @ X
    flush = torch.ones(flush_mb * 1024 * 1024 // 4, dtype=torch.float32, device=B64.device)
    out = {}
    for dt in (torch.float64, torch.float32, torch.bfloat16):
        B = B64 if dt == torch.float64 else B64.to(dt)
        work = torch.empty(int(L.smx_dense_bmatmul_ws_size(n, _capi.DTYPE_CODE[dt])), dtype=torch.uint8,
                           device=B64.device)
        Y = dense_bmatmul(B, ld, X, work=work)
        t = event_time(lambda: dense_bmatmul(B, ld, X, work=work), 10, flush)
        nbytes = n * n * B.element_size()
        out[str(dt).split(".")[-1]] = {
            "us": t * 1e6, "B_GBs": nbytes / t / 1e9, "hbm_frac": nbytes / t / 1e9 / HBM_PEAK_GBS,
            "TFLOPs": 2.0 * n * n * k / t / 1e12,
            "rel_err_vs_fp64": float(torch.linalg.norm(Y - ref) / torch.linalg.norm(ref))}
        del B, work
    return {"nrhs": k, "kernel": "mm_tile_lds_kernel (X slice staged in LDS; v_mfma_f64_16x16x4 / f32_16x16x4 / f32_16x16x32_bf16)",
            "note": "times include the X staging and split-K reduction launches", **out}


def hmatrix_multi_rhs_timing(h64, hmx, k=64, flush_mb=512):
    """H-matrix product for k right-hand sides (smx_hmat_matmat: the z and
    leaf stages as MFMA tiles over the columns) against k matvecs, FP64 and
    FP32 factors, flushed, HIP events; error against the FP64 operator's k
    matvecs."""
    import numpy as np
    import torch
    dev = torch.device("cuda", 0)
    H64, Hmx = h64.demag_field.B_hmatrix_gpu, hmx.demag_field.B_hmatrix_gpu
    X = torch.from_numpy(np.random.default_rng(0).standard_normal((k, H64.nb))).to(dev)
    ref = torch.stack([H64.matvec(X[j]) for j in range(k)])
    flush = torch.ones(flush_mb * 1024 * 1024 // 4, dtype=torch.float32, device=dev)
    out = {}
    for name, H in (("fp64", H64), ("fp32", Hmx)):
        Y = H.matmat(X)
        t = event_time(lambda: H.matmat(X, Y), 8, flush)
        tc = event_time(lambda: H.matmat(X, Y, near_tiles=False), 8, flush)
        H.matmat(X, Y)
        tv = event_time(lambda: H.matvec(X[0]), 8, flush)
        flops = 2.0 * H.far_storage_entries * k
        out[name] = {"us": t * 1e6, "us_per_rhs": t * 1e6 / k, "matvec_us": tv * 1e6,
                     "us_near_csr_rows": tc * 1e6,
                     "speedup_vs_k_matvecs": k * tv / t, "lowrank_TFLOPs": flops / t / 1e12,
                     "rel_err_vs_fp64_matvecs": float((torch.linalg.norm(Y - ref, dim=1)
                                                       / torch.linalg.norm(ref, dim=1)).max()),
                     "near_tiles": dict(H.near_tile_stats)}
        log(f"H-matrix {k}-RHS product {name}: {t * 1e6:.0f} us ({k * tv / t:.1f}x {k} matvecs; "
            f"near as CSR rows {tc * 1e6:.0f} us) tiles {H.near_tile_stats}")
    return {"nrhs": k, "kernel": "hmm_z_kernel + hmm_leaf_kernel (v_mfma_f32_16x16x4 / f64_16x16x4) + "
                                 "hmm_near_tile_kernel (v_mfma_f64_16x16x4 over dense 16-row near tiles)",
            "note": "times include the X staging launch", **out}


def build_case(config, dev):
    import numpy as np
    import torch
    from smx import mesh as M
    from smx.demag import OperatorSet
    from smx.llg import EffectiveField
    common = dict(ksp_rtol_u1=1e-6, ksp_rtol_u2=1e-6, ksp_atol=1e-12, ksp_pc_type=PC_TYPE,
                  ksp_pc_type_u2=PC_TYPE_U2)
    t0 = time.perf_counter()
    if config == "cube":
        mesh = M.cube_c2()
        mat = dict(Ms=8e5, Aex=1.3e-11, alpha=0.5)
        ops = OperatorSet(mesh, dev)
        h64 = EffectiveField(mesh, use_demag=True, operators=ops, demag_kwargs=dict(B_dtype="float64", **common),
                             **mat)
        hmx = EffectiveField(mesh, use_demag=True, operators=ops,
                             demag_kwargs=dict(B_dtype="float32", geometry=h64.demag_field.opB, **common), **mat)
        h_n1 = EffectiveField(mesh, use_demag=True, operators=ops,
                              demag_kwargs=dict(B_dtype="float32", geometry=h64.demag_field.opB, cg_mixed=True,
                                                **common), **mat)
        pool = [torch.from_numpy(vortex_m(mesh.X, seed=s)).to(dev) for s in range(8)]
        work = "cube150nm_30x30x30_kuhn (N=29791, Nb=5402, dense B)"
        extra = {}
    else:
        from smx.boundary import build_boundary_geometry
        from smx.hmatrix import HMatrixBuildConfig, build_hmatrix_device
        mesh = M.film_c3()
        mat = dict(Ms=8e5, Aex=1.3e-11, alpha=0.5, Ku=5e5, n_ani_vec=(0.0, 0.0, 1.0), D_int=3e-3,
                   n0_int_vec=(0.0, 0.0, 1.0))
        ops = OperatorSet(mesh, dev)
        log(f"film mesh + FEM {time.perf_counter() - t0:.1f}s")
        geo = build_boundary_geometry(mesh.X, mesh.tets)
        tb = time.perf_counter()
        data = build_hmatrix_device(geo, HMatrixBuildConfig(epsilon=1e-6, eta=2.0, leaf_size=64), dev)
        t_build = time.perf_counter() - tb
        log(f"H-matrix build {t_build:.1f}s far={len(data.far_blocks)} near_nnz={data.near_csr.nnz}")
        hk = dict(boundary_backend="hmatrix", geometry=geo, hmatrix_prebuilt=data, **common)
        h64 = EffectiveField(mesh, use_demag=True, operators=ops, demag_kwargs=dict(hmatrix_use_fp32=False, **hk),
                             **mat)
        hmx = EffectiveField(mesh, use_demag=True, operators=ops,
                             demag_kwargs=dict(hmatrix_use_fp32=True, **hk), **mat)
        h_n1 = EffectiveField(mesh, use_demag=True, operators=ops,
                              demag_kwargs=dict(hmatrix_use_fp32=True, cg_mixed=True, **hk), **mat)
        pool = [torch.from_numpy(random_m(mesh.num_nodes, seed=s)).to(dev) for s in range(4)]
        work = "film500x500x10nm_200x200x4_kuhn (N=202005, Nb=83202, H-matrix leaf 64, eta 2, eps 1e-6)"
        md = data.metadata
        extra = {"hmatrix_build_s": round(t_build, 2), "far_blocks": len(data.far_blocks),
                 "near_nnz": int(md["near_nnz"]), "far_storage_entries": int(md["far_storage_entries"]),
                 "average_far_rank": round(float(md["average_far_rank"]), 2),
                 "lindholm_pair_evaluations": int(md["lindholm_pair_evaluations"])}
        extra["_data"] = data
    extra["_h_n1"] = h_n1
    torch.cuda.synchronize()
    return mesh, mat, h64, hmx, pool, work, extra, time.perf_counter() - t0


def cpu_baseline(config, mesh, mat, pool, extra, h64, budget_s):
    """FP64 oracle RHS evaluations on the host (bounded sample)."""
    import numpy as np
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla
    sys.path.insert(0, str(ROOT / "oracle"))
    import smx_oracle as O
    t0 = time.perf_counter()
    Ms, A_ex, alpha = mat["Ms"], mat["Aex"], mat["alpha"]
    fem = O.fem_operators(mesh.X, mesh.tets, Ms)
    Kex = O.exchange_matrix(mesh.X, mesh.tets, fem.V, A_ex, Ms)
    Klin = Kex
    if mat.get("D_int"):
        Klin = Klin + O.dmi_interfacial_matrix(mesh.X, mesh.tets, fem.V, mat["D_int"], Ms, mat["n0_int_vec"])
    geo = h64.demag_field.opB
    if config == "cube":
        B = h64.demag_field.B_dense_gpu.cpu().numpy()[:, :geo.Nb]
        bmat = lambda x: B @ x
        lu1 = spla.splu(_pinned(fem.A).tocsc())
        lu2 = spla.splu(fem.A_bc.tocsc())
        solve1 = lambda b: _pin_solve(lu1, b)
        solve2 = lambda b: lu2.solve(b)
        how = "scipy splu Poisson solves (N < 60000, as the reference CPU path), numpy dense B matvec"
    else:
        data = extra["_data"]
        H = O.HMatrixOracle(perm=data.perm, diag_jump=data.diag_jump, near=data.near_csr,
                            far=[(b.row_start, b.row_stop, b.col_start, b.col_stop, b.U, b.V, 0.0)
                                 for b in data.far_blocks], stats={})
        bmat = H.matvec
        from smx.amg import apply_vcycle, build_levels
        A1 = sp.csr_matrix(fem.A)
        A2 = sp.csr_matrix(fem.A_bc)
        lv1 = build_levels(A1)
        lv2 = build_levels(A2)
        M1 = spla.LinearOperator(A1.shape, matvec=lambda v: apply_vcycle(lv1, v))
        M2 = spla.LinearOperator(A2.shape, matvec=lambda v: apply_vcycle(lv2, v))
        solve1 = lambda b: spla.cg(A1, b, rtol=1e-6, atol=1e-12, maxiter=2000, M=M1)[0]
        solve2 = lambda b: spla.cg(A2, b, rtol=1e-6, atol=1e-12, maxiter=12000, M=M2)[0]
        how = ("scipy CG preconditioned by the same smoothed-aggregation V-cycle on the host (rtol 1e-6; the "
               "reference CPU path uses GAMG-CG above 60000 dofs), H-matrix matvec with the same factors in "
               "numpy FP64")
    setup = time.perf_counter() - t0
    Vn = np.repeat(fem.V, 3)

    def rhs(mv):
        b = fem.Div @ mv
        b -= b.mean()
        u1 = solve1(b)
        u1 -= u1.mean()
        g = np.zeros(fem.N)
        g[geo.bverts] = bmat(u1[geo.bverts])
        bb = -(fem.A @ g)
        bb[fem.bnodes] = g[fem.bnodes]
        u2 = solve2(bb)
        H = Klin @ mv + (fem.G @ (u1 + u2)) / Vn
        if mat.get("Ku"):
            H = H + O.anisotropy_field(mv, mat["n_ani_vec"], mat["Ku"], Ms)
        return O.llg_rhs(mv, H, GAMMA, alpha, Ms)

    def timed(budget):
        n = 0
        t0 = time.perf_counter()
        while True:
            rhs(pool[n % len(pool)])
            n += 1
            el = time.perf_counter() - t0
            if el >= budget or n >= 200:
                return n, el

    from threadpoolctl import threadpool_info, threadpool_limits
    threads = max((p.get("num_threads", 1) for p in threadpool_info()), default=1)
    n, el = timed(budget_s)
    with threadpool_limits(limits=1):
        n1, el1 = timed(budget_s / 2)
    cpu = ""
    try:
        cpu = next(l.split(":", 1)[1].strip() for l in open("/proc/cpuinfo") if l.startswith("model name"))
    except (OSError, StopIteration):
        pass
    return {"value": n / el, "unit": "evals/s", "cores": int(threads), "kind": "port",
            "sample": f"{n} FP64 RHS evals of the same workload in {el:.1f}s ({how}; setup {setup:.1f}s excluded)",
            "cpu_model": cpu,
            "single_thread": {"value": n1 / el1, "unit": "evals/s", "cores": 1,
                              "sample": f"{n1} evals in {el1:.1f}s with the BLAS pool limited to 1 thread"}}


def _pinned(A):
    A = A.tolil()
    A[0, :] = 0.0
    A[0, 0] = 1.0
    return A.tocsr()


def _pin_solve(lu, b):
    bb = b.copy()
    bb[0] = 0.0
    return lu.solve(bb)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", choices=["film", "cube", "adjoint", "disk", "paths"], default="film")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--pc", choices=["gamg", "jacobi", "auto"], default="auto",
                    help="CG preconditioner of both Poisson solves (the reference uses GAMG; auto: the faster "
                         "of GAMG and Jacobi, timed at set-up)")
    ap.add_argument("--pc-u2", choices=["gamg", "jacobi", "auto"], default=None,
                    help="preconditioner of the u2 solve when it differs from --pc")
    ap.add_argument("--cpu-budget", type=float, default=20.0)
    ap.add_argument("--pmc-probe", nargs="?", const="cg", default=None, choices=["cg", "matvec", "amg"],
                    help="short program for rocprofv3 --pmc passes: calibration copy + Jacobi CG iterations "
                         "(cg), + boundary matvecs (matvec) or + AMG-CG iterations (amg)")
    ap.add_argument("--evals-only", action="store_true",
                    help="after set-up run only the mixed-precision evaluations (for kernel traces)")
    ap.add_argument("--multi-rhs-only", action="store_true",
                    help="film: after set-up time only the 64-column H-matrix product (for kernel traces)")
    args = ap.parse_args()
    global PC_TYPE, PC_TYPE_U2
    PC_TYPE = args.pc
    PC_TYPE_U2 = args.pc_u2
    if args.pmc_probe == "matvec":
        pmc_probe_matvec(args.config)
        return
    if args.pmc_probe == "amg":
        pmc_probe_amg(args.config)
        return
    if args.pmc_probe:
        pmc_probe(args.config)
        return
    if args.config == "adjoint":
        adjoint_through_time()
        return
    if args.config == "disk":
        energy_gradient_disk()
        return
    if args.config == "paths":
        extended_paths(args.cpu_budget, args.no_cpu)
        return
    if args.gpus != 1:
        print(json.dumps({"error": "the benchmark runs on one GPU; --gpus must be 1"}))
        sys.exit(2)
    import numpy as np
    import torch
    dev = torch.device("cuda", 0)
    mesh, mat, h64, hmx, pool, work, extra, setup_s = build_case(args.config, dev)
    log(f"setup {setup_s:.1f}s")
    if args.multi_rhs_only:
        if args.config != "film":
            raise SystemExit("--multi-rhs-only needs the film H-matrix operator")
        print(json.dumps({"workload": work, "multi_rhs_boundary_product": hmatrix_multi_rhs_timing(h64, hmx)}))
        return
    out = torch.empty(3 * mesh.num_nodes, dtype=torch.float64, device=dev)

    def run(hef, steps, warmup):
        for i in range(warmup):
            hef.llg_rhs_vec(pool[i % len(pool)], out)
        torch.cuda.synchronize()
        its = []
        redo0 = hef.demag_field.speculative_redos
        t0 = time.perf_counter()
        for i in range(steps):
            hef.llg_rhs_vec(pool[i % len(pool)], out)
            its.append((hef.demag_field.last_its_u1, hef.demag_field.last_its_u2))
        torch.cuda.synchronize()
        run.redos = hef.demag_field.speculative_redos - redo0
        return time.perf_counter() - t0, its

    if args.evals_only:
        tmx, itsmx = run(hmx, args.steps, args.warmup)
        log(f"mixed {args.steps / tmx:.2f} evals/s (evaluations only)")
        return
    t64, _ = run(h64, args.steps, args.warmup)
    log(f"fp64 {args.steps / t64:.2f} evals/s")
    tmx, itsmx = run(hmx, args.steps, args.warmup)
    redos_mx = run.redos
    log(f"mixed {args.steps / tmx:.2f} evals/s ({redos_mx} speculative solve repairs)")
    h_n1 = extra["_h_n1"]
    tn1, _ = run(h_n1, args.steps, args.warmup)
    redos_n1 = run.redos
    log(f"mixed with FP32 inner CG and FP64 refinement {args.steps / tn1:.2f} evals/s ({redos_n1} repairs)")
    bd64 = eval_breakdown(h64, pool)
    bdmx = eval_breakdown(hmx, pool)
    bdn1 = eval_breakdown(h_n1, pool)
    log(f"stage times (ms) fp64 {bd64} mixed {bdmx} fp32 inner cg {bdn1}")
    r64 = torch.empty_like(out)
    rmx = torch.empty_like(out)
    errs, herr = [], []
    herr_n1, errs_n1 = [], []
    rn1 = torch.empty_like(out)
    for mv in pool[:2]:
        h64.llg_rhs_vec(mv, r64)
        hmx.llg_rhs_vec(mv, rmx)
        h_n1.llg_rhs_vec(mv, rn1)
        errs.append(float(torch.linalg.norm(rmx - r64) / torch.linalg.norm(r64)))
        herr.append(float(torch.linalg.norm(hmx.H_demag_gpu - h64.H_demag_gpu) / torch.linalg.norm(h64.H_demag_gpu)))
        errs_n1.append(float(torch.linalg.norm(rn1 - r64) / torch.linalg.norm(r64)))
        herr_n1.append(float(torch.linalg.norm(h_n1.H_demag_gpu - h64.H_demag_gpu)
                             / torch.linalg.norm(h64.H_demag_gpu)))
    dn1 = h_n1.demag_field
    krylov64 = (bd64["u1_solve"] + bd64["u2_solve"]) / bd64["total"]
    krylovmx = (bdmx["u1_solve"] + bdmx["u2_solve"]) / bdmx["total"]
    precision = {
        "mixed": ("the value: FP32 boundary operator (H-matrix factors or dense B), FP32 AMG preconditioner "
                  "values, CG vectors, scalars and inner products FP64"),
        "fp32_inner_cg": {
            "what": ("FP32 boundary operator and FP32 inner CG (vectors and iteration operator) with FP64 x, scalars "
                     "and inner products (exact products, compensated cross-wave and cross-workgroup sums), the "
                     "FP32 solve to rtol/2, then one FP64 iterative-refinement step per solve (FP64 residual check, "
                     "correction solve if it fails); FP32 AMG values"),
            "evals_per_s": args.steps / tn1, "over_fp64": t64 / tn1,
            "rel_err_hdemag_vs_fp64": max(herr_n1), "rel_err_rhs_vs_fp64": max(errs_n1),
            "speculative_repairs_in_timed_steps": int(redos_n1),
            "stage_ms": {k: round(v, 4) for k, v in bdn1.items()},
            "cg_iterations_before_and_in_last_solve": {"u1": list(dn1._its_split["u1"]),
                                                       "u2": list(dn1._its_split["u2"])}},
        "stage_ms_fp64": {k: round(v, 4) for k, v in bd64.items()},
        "stage_ms_mixed": {k: round(v, 4) for k, v in bdmx.items()},
        "krylov_share_fp64": krylov64,
        "krylov_share_mixed": krylovmx,
        "amdahl_ratio_if_mixed_matvec_free": bd64["total"] / (bdmx["total"] - bdmx["boundary_matvec"]),
        "amdahl_ratio_if_mixed_krylov_free": bd64["total"] / (bdmx["total"] - bdmx["u1_solve"] - bdmx["u2_solve"]),
    }
    cg_t, cg_bytes = cg_iteration_timing(hmx.demag_field)
    mv_cold, mv_warm, mv_bytes, mv_name = matvec_timing(hmx.demag_field)
    mv_variant = ("near_launch" if hmx.demag_field.boundary_backend == "hmatrix"
                  and hmx.demag_field.B_hmatrix_gpu.near_in_z else None)
    mv64_cold, _, mv64_bytes, _ = matvec_timing(h64.demag_field)
    ach = cg_bytes / cg_t / 1e9
    mv_ach = mv_bytes / mv_cold / 1e9
    cfg = {"workload": work, "cg_rtol": 1e-6, "initial_guess": "zero", "setup_s": round(setup_s, 2),
           "cg_launch": "speculative (last iteration count + margin, one status check per evaluation)",
           "speculative_repairs_in_timed_steps": int(redos_mx)}
    cfg.update({k: v for k, v in extra.items() if not k.startswith("_")})
    res = {
        "metric": "demag+LLG RHS evals/s (mixed precision boundary operator vs FP64)",
        "value": args.steps / tmx,
        "unit": "evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * tmx / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64 (boundary operator and AMG values f32)",
        "data": "synthetic (seeded)",
        "config": cfg,
        "fp64": {"value": args.steps / t64, "ms_per_step": 1e3 * t64 / args.steps},
        "mixed_over_fp64": t64 / tmx,
        "mixed_vs_fp64_rel_err_rhs": max(errs),
        "mixed_vs_fp64_rel_err_hdemag": max(herr),
        "precision_study": precision,
        "cg_iterations_u1_u2": [float(np.mean([a for a, _ in itsmx])), float(np.mean([b for _, b in itsmx]))],
        "pc": [hmx.demag_field.pc_type, hmx.demag_field.pc_type_u2],
        "amg_levels_u1": hmx.demag_field.amg_u1.describe() if hmx.demag_field.amg_u1 else None,
        "roofline": {"bound": "hbm", "kernel": mv_name, "achieved": mv_ach, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": mv_ach / HBM_PEAK_GBS,
                     "traffic": measured_traffic(args.config, "matvec", mv_variant),
                     "algorithmic_bytes": int(mv_bytes),
                     "streamed_bytes": (int(hmx.demag_field.B_hmatrix_gpu.stored_bytes())
                                        if hmx.demag_field.boundary_backend == "hmatrix" else int(mv_bytes)),
                     "avg_us": mv_cold * 1e6, "avg_us_cache_warm": mv_warm * 1e6,
                     "note": "mixed precision boundary operator; timed after a 512 MB read sweep that evicts the "
                             "Infinity Cache",
                     "fp64_operator_GBs": mv64_bytes / mv64_cold / 1e9,
                     "fp64_operator_avg_us": mv64_cold * 1e6},
        "cg_roofline": {"bound": "hbm", "kernel": ("AMG-CG iteration of the u1 solve (update, V-cycle and A u "
                                                   "launches)" if hmx.demag_field.pc_type == "gamg"
                                                   else "cg_iter_kernel (Jacobi CG iteration, u1 operator)"),
                        "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": ach / HBM_PEAK_GBS,
                        "traffic": measured_traffic(args.config, "cg" if hmx.demag_field.pc_type == "jacobi"
                                                    else "amg"),
                        "algorithmic_bytes": int(cg_bytes), "avg_us": cg_t * 1e6},
    }
    if h64.demag_field.boundary_backend == "dense":
        res["multi_rhs_boundary_product"] = multi_rhs_timing(h64.demag_field)
    if args.config == "film":
        res["multi_rhs_boundary_product"] = hmatrix_multi_rhs_timing(h64, hmx)
    if not args.no_cpu:
        res["cpu_baseline"] = cpu_baseline(args.config, mesh, mat, [p.cpu().numpy() for p in pool[:2]], extra, h64,
                                           args.cpu_budget)
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
