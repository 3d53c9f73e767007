#!/usr/bin/env python3
"""Benchmark of the BoR monostatic hot path on one MI355X.

Workload: the ogive-cylinder-cone body (10 wavelengths long, 1 wavelength
max radius, 1200 elements), |m| <= 70 (141 mode systems of size ~2400),
361 aspects x 2 polarisations at one frequency.  One step = fill + batched
LU/solve (+ FP64 refinement) + far-field sum of all 141 x 361 solves, with
the geometry already on the device.  The headline path is the mixed
precision one (FP32 far fill and LU, FP64 near band, residual refinement and
far field); the plain FP64 path runs once beside it to give the dB error.

Prints one JSON line (the driver's contract).
"""

import argparse
import importlib
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

MI355X_PEAK = {"fp64_mfma_tflops": 78.6, "fp64_vector_tflops": 78.6, "fp32_mfma_tflops": 157.3,
               "bf16_mfma_tflops": 2500.0, "hbm_gbs": 8000.0}
SPLIT_BF16_PRODUCTS = 6   # bf16 MFMA products per FP32 product in the split-bf16 GEMM (variant 5)


def pkg(m):
    return importlib.import_module("20260722_ghost_amd." + m)


def db(x):
    return 10.0 * np.log10(np.maximum(np.abs(x) ** 2 * 4 * math.pi, 1e-300))


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def tuned_tiles(precision, nodes, thetas, m_max):
    """The autotuner's (LU block size, GEMM tile) for this sweep's shape,
    measured (untimed, before the benchmark) on first use per device."""
    L = pkg("_lib")
    at = pkg("autotune")
    prec = L.GHOST_MIXED if precision == "mixed" else L.GHOST_FP64
    t0 = time.perf_counter()
    tiles = at.best(prec, 2 * len(nodes), m_max + 1, 2 * len(thetas))
    log(f"{precision}: autotuned (nb, variant) = {tiles} in {time.perf_counter() - t0:.1f} s")
    return tiles


def run_gpu(nodes, freq, thetas, m_max, precision, steps, warmup, refine, profile=True, tiles=(64, -1),
            lu_super=None):
    log(f"{precision}: build (m_max={m_max}, n_elems={len(nodes) - 1}, aspects={len(thetas)})")
    L = pkg("_lib")
    bs = pkg("bor_solver")
    k = 2 * math.pi * freq / bs.C0
    _, ne, nbk, streaming = bs.discretisation(nodes, k, thetas, m_max, "cfie")
    h = L.Handle(L.GHOST_MIXED if precision == "mixed" else L.GHOST_FP64, nb=tiles[0], residual=False)
    h.set_gemm_variant(tiles[1])
    if lu_super is not None:
        h.set_lu_super(lu_super)
    h.add_problem(nodes, k, m_max, ne, nbk, 0.5, 0.5 * bs.ETA0, thetas)
    h.build()
    for i in range(warmup):
        h.run(refine)
        h.sync()
        log(f"{precision}: warmup {i + 1}/{warmup} done")
    # device-wide barrier on both sides of the timed region (libghost_bor owns the
    # device work; torch's own HIP runtime is not loaded into this process)
    L.device_synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        h.run(refine)
    h.sync()
    L.device_synchronize()
    dt = (time.perf_counter() - t0) / steps
    log(f"{precision}: {steps} timed steps, {dt * 1e3:.1f} ms/step")
    phase = h.timings()
    ktimes = {}
    if profile:
        # one more step, untimed, with HIP events around every launch (which
        # serialises the near band and the far fill) for the per-kernel table
        h.set_profile(True)
        h.run(refine)
        h.sync()
        ktimes = h.kernel_times()
        log(f"{precision}: profiled step done")
    if precision == "mixed":
        phase["probe_errors"] = h.refine_history()
    dF = h.dF(0)
    h.close()
    return dt, dF, ktimes, phase, (ne, nbk)


def cpu_baseline(nodes, freq, thetas, m_max):
    """The FP64 oracle (NumPy/SciPy + OpenBLAS) on a bounded sample of the
    same sweep, with the reference's thread model: a pool of nproc - 1
    workers over independent units (bor_dispatch.py:386-387), one BLAS thread
    each (run_hpc_bor_monostatic.py:107, 155-160).  Each leg runs one unit per
    worker concurrently and is scaled by (units in the sweep) / (units run):
    far fill of 2-element test-row tiles, near element pairs, mode LU +
    solve, and the per-aspect excitation and far-field projection of a mode.
    A whole mode cannot be timed at this size: the far sampling is shared by
    all modes and alone takes the better part of an hour of CPU time."""
    from concurrent.futures import ThreadPoolExecutor
    from threadpoolctl import threadpool_limits
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import bor_oracle as bo
    log("cpu baseline: oracle sample")
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 2
    workers = max(1, min(cores, 16) - 1)
    body = bo.Body(nodes)
    k = 2 * math.pi * freq / bo.C0
    osc = bo.osc_for(body, k, streaming=True)
    ne, nbk = bo.n_xi_efie(osc, m_max), bo.n_xi_bracket(osc, m_max)
    Ne = body.Ne
    g = body.g
    xe, xb = bo.xi_grid(ne), bo.xi_grid(nbk)
    n = 2 * body.Nn - 4
    nrhs = 2 * len(thetas)

    def far_tile(e0):
        sl = slice(e0 * 4, (e0 + 2) * 4)
        P_ = lambda v: v[sl][:, None]
        Q_ = lambda v: v[None, :]
        G = bo.green_orders_fft(bo.green_samples(P_(g.rho), P_(g.z), Q_(g.rho), Q_(g.z), k, xe), m_max + 2)
        bo._zero_near_rows(G, e0, 2, Ne, 4)
        Fs = bo.bracket_samples(P_(g.rho), P_(g.z), P_(g.trho), P_(g.tz), Q_(g.rho), Q_(g.z), Q_(g.trho),
                                Q_(g.tz), k, xb)
        return [bo.bracket_modes_fft(F, m_max) for F in Fs]

    def near_pair(i):
        e = Ne // 4 + i
        f = e + (i % 3) - 1
        s, sp, w = bo.cell_points(bo.near_kind(e, f))
        p, q = body.element_points(e, s), body.element_points(f, sp)
        bo.green_orders_near(p["rho"], p["z"], q["rho"], q["z"], k, m_max)
        bo.bracket_modes_near(p["rho"], p["z"], p["trho"], p["tz"], q["rho"], q["z"], q["trho"], q["tz"], k, m_max)

    def mode_solve(i):
        rng = np.random.default_rng(i)
        A = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)) + n * np.eye(n)
        B = rng.standard_normal((n, nrhs)) + 1j * rng.standard_normal((n, nrhs))
        t0 = time.perf_counter()
        np.linalg.solve(A, B)
        return time.perf_counter() - t0

    nth = 4

    def excite(i):
        m = i % (m_max + 1)
        for th in thetas[i * nth:(i + 1) * nth]:
            for pol in ("VV", "HH"):
                bo.rhs_efie(body, k, m, th, pol)
                bo.rhs_mfie(body, k, m, th, pol)
            sol = np.zeros(2 * body.Nn, dtype=complex)
            bo.farfield(body, k, m, sol, th)

    def leg(fn, units):
        t0 = time.perf_counter()
        with ThreadPoolExecutor(workers) as ex:
            out = list(ex.map(fn, units))
        return time.perf_counter() - t0, out

    with threadpool_limits(limits=1):
        t_far, _ = leg(far_tile, [2 * Ne // (2 * workers + 2) * (i + 1) for i in range(workers)])
        t_near, _ = leg(near_pair, range(workers))
        t_sol, per = leg(mode_solve, range(workers))
        t_exc, _ = leg(excite, range(workers))
    est = {"far": t_far * Ne / (2 * workers), "near": t_near * (3 * Ne - 2) / workers,
           "solve": t_sol * (2 * m_max + 1) / workers,
           "rhs_farfield": t_exc * (2 * m_max + 1) * len(thetas) / (workers * nth)}
    total = sum(est.values())
    solves = (2 * m_max + 1) * len(thetas)
    return {
        "value": solves / total,
        "unit": "solves/s",
        "cores": workers,
        "kind": "port",
        "sample": (f"FP64 oracle, {workers} worker threads x 1 BLAS thread (the reference's mode pool); each "
                   f"leg runs {workers} units concurrently and is scaled to the sweep: far fill "
                   f"{2 * workers} of {Ne} test rows, near {workers} of {3 * Ne - 2} element pairs, "
                   f"{workers} of {2 * m_max + 1} mode LU+solves (n={n}, nrhs={nrhs}, "
                   f"{np.mean(per):.1f} s each), RHS/far field {workers * nth} of "
                   f"{(2 * m_max + 1) * len(thetas)} (mode, aspect) pairs; est. total {total:.0f} s "
                   f"(far {est['far']:.0f}, near {est['near']:.0f}, solve {est['solve']:.0f}, "
                   f"rhs/far-field {est['rhs_farfield']:.0f}); {time.perf_counter() - T_CPU0:.0f} s of wall time"),
    }


T_CPU0 = 0.0


def cpu_baseline_timed(*args):
    global T_CPU0
    T_CPU0 = time.perf_counter()
    return cpu_baseline(*args)


def run_adjoint(steps):
    """Config 4: forward RCS + adjoint gradient of J = sum sigma_dBsm over the
    node radii (FP64), with central differences of the full solve on a few
    nodes as the check.  Prints its own JSON line."""
    import torch  # noqa: F401  (the autograd front end needs it importable)
    wl = pkg("workloads")
    ad = pkg("bor_adjoint")
    name, nodes, freq, thetas, m_max = wl.config(4)
    spec = ad.RcsSpec(nodes, freq, thetas, m_max, cfie_alpha=0.5)
    log(f"adjoint: {name}, m_max={m_max}, aspects={len(thetas)}")
    L = pkg("_lib")
    # warm-up (module load, first build)
    J, g = ad.objective_and_gradient(spec)
    L.device_synchronize()
    t_fwd, t_adj, t_fd = [], [], []
    g_fd = None
    for i in range(steps):
        t0 = time.perf_counter()
        F, h = ad.solve_far_field(spec, spec.nodes[:, 0])
        L.device_synchronize()
        t1 = time.perf_counter()
        s = ad.sigma_db(F)
        a = np.ascontiguousarray((20.0 / math.log(10.0) * np.conj(F) / np.abs(F) ** 2).T).ravel()
        g = h.adjoint(a, method="analytic")
        L.device_synchronize()
        t2 = time.perf_counter()
        if i == 0:
            # the central-difference adjoint (12 perturbed fills) on the same handle, for comparison
            g_fd = h.adjoint(a, 0.0, method="fd")
            L.device_synchronize()
            t_fd.append(time.perf_counter() - t2)
        h.close()
        t_fwd.append(t1 - t0)
        t_adj.append(t2 - t1)
    log(f"adjoint: forward {min(t_fwd) * 1e3:.1f} ms, analytic gradient {min(t_adj) * 1e3:.1f} ms, "
        f"difference adjoint {t_fd[0] * 1e3:.1f} ms")
    lam = bench_c0() / freq
    free = int(np.sum(spec.nodes[:, 0] > 1e-9))
    errs = []
    for q in (5, len(nodes) // 3, len(nodes) // 2, 2 * len(nodes) // 3, len(nodes) - 6):
        vals = []
        for sgn in (1.0, -1.0):
            r = spec.nodes[:, 0].copy()
            r[q] += sgn * 1e-6 * lam
            Fq, hq = ad.solve_far_field(spec, r)
            hq.close()
            vals.append(float(np.sum(ad.sigma_db(Fq))))
        fd = (vals[0] - vals[1]) / (2e-6 * lam)
        errs.append(abs(g[q] - fd) / max(abs(fd), 1e-30))
        log(f"adjoint: node {q} adjoint {g[q]:.9e} fd {fd:.9e}")
    out = {
        "metric": "RCS gradient (J = sum sigma_dBsm over aspects x pols) w.r.t. node radii, forward + adjoint",
        "config": {"workload": name, "n_elems": int(len(nodes) - 1), "m_max": m_max, "modes": 2 * m_max + 1,
                   "aspects": len(thetas), "polarizations": 2, "free_nodes": free, "precision": "fp64"},
        "forward_ms": round(min(t_fwd) * 1e3, 3),
        "adjoint_ms": round(min(t_adj) * 1e3, 3),
        "adjoint_method": "analytic (dual-number fill derivative contracted with the adjoint weights)",
        "difference_adjoint_ms": round(t_fd[0] * 1e3, 3),
        "analytic_vs_difference_adjoint_max_rel": float(np.max(np.abs(g - g_fd)) / np.max(np.abs(g_fd))),
        "central_difference_ms_estimate": round(2 * free * min(t_fwd) * 1e3, 1),
        "speedup_vs_central_differences": round(2 * free * min(t_fwd) / (min(t_fwd) + min(t_adj)), 1),
        "fd_check_max_rel_err": float(max(errs)),
        "J": float(np.sum(s)),
        "steps": steps,
    }
    print(json.dumps(out), flush=True)


def batch_bytes(nodes, m_max, n_theta):
    """Device bytes of one problem in a mixed-precision handle with mode
    pairing off and the adjoint's buffers: FP32 augmented systems, their FP64
    copy, the FP32 Q, the FP64 n x n_rhs blocks, far blocks and near
    point-pair values."""
    Nn = len(nodes)
    n, nsys, nrhs = 2 * Nn, 2 * m_max + 1, 2 * n_theta
    ld = n + nrhs
    zmodes = 16 * ((nsys + 15) // 16)
    total = (nsys * n * ld * (8 + 8) + nsys * n * n * 16 + 8 * nsys * n * nrhs * 16
             + (2 * Nn) ** 2 * zmodes * 8 + 1440 * (Nn - 1) * (m_max + 2) * 80)
    return 1.1 * total


def run_shape_batch(n_shapes, budget_gb):
    """Config 5: forward RCS and the gradient of J = sum sigma_dBsm w.r.t.
    every node radius for n_shapes Fourier perturbations of the ogive body at
    three frequencies (600 / 1200 / 1800 elements, |m| <= 35 / 70 / 105), mixed
    precision (FP32 fill and LU, one FP64 refinement sweep for the currents and
    one for lambda).  Per frequency one handle holds as many shapes as fit the
    memory budget and ONE captured graph (fill + solve + far field + a' +
    transposed solve + Q + analytic fill derivative) serves all its shapes in
    turn: ghost_bor_set_nodes swaps the geometry in place and every graph
    launch returns the forward RCS and the gradients of a group of shapes.
    The quadrature rules (n_xi, near tail orders) are those of the handle's
    first shapes.  Prints its own JSON line."""
    wl = pkg("workloads")
    bs = pkg("bor_solver")
    L = pkg("_lib")
    probs = wl.perturbed_batch(n_shapes)
    scales = sorted({round(f / bs.C0, 6) for (_, _, f, _, _) in probs})
    by_scale = {sc: [p for p in probs if round(p[2] / bs.C0, 6) == sc] for sc in scales}
    solves = sum((2 * mm + 1) * len(th) for (_, _, _, th, mm) in probs)
    t_gpu, t_wall, rows, n_launch = 0.0, time.perf_counter(), [], 0
    results = {}
    log(f"shape batch: {n_shapes} shapes x {len(scales)} frequencies = {len(probs)} problems, {solves} solves")
    for sc in scales:
        group = by_scale[sc]
        _, nodes0, f, th, mm = group[0]
        k = 2 * math.pi * f / bs.C0
        _, ne, nb, _ = bs.discretisation(nodes0, k, th, mm, "cfie")
        slots = max(1, min(len(group), int(budget_gb * 1e9 // batch_bytes(nodes0, mm, len(th)))))
        t0 = time.perf_counter()
        h = L.Handle(L.GHOST_MIXED, nb=64, residual=False, pairing=False)
        for j in range(slots):
            h.add_problem(group[j][1], k, mm, ne, nb, 0.5, 0.5 * bs.ETA0, th)
        h.build()
        h.set_objective(np.ones(2 * len(th) * slots))
        h.capture_adjoint(1)
        t_cap = time.perf_counter() - t0
        for c0 in range(0, len(group), slots):
            chunk = group[c0:c0 + slots]
            if c0 > 0:
                for j, (_, nodes, _, _, _) in enumerate(chunk):
                    h.set_nodes(j, nodes)
            h.launch()
            h.sync()
            ms = h.graph_ms()
            g = h.gradient()
            n_launch += 1
            t_gpu += ms * 1e-3
            off = 0
            for j, (si, nodes, _, _, _) in enumerate(chunk):
                F = h.dF(j).sum(axis=0)
                results[(si, sc)] = (float(np.sum(10 * np.log10(4 * math.pi * np.abs(F) ** 2))),
                                     g[off:off + len(nodes)], (ne, nb))
                off += len(nodes)
            rows.append({"scale": sc, "elements": len(nodes0) - 1, "shapes": len(chunk), "graph_ms": round(ms, 1)})
            log(f"shape batch: {sc} f0 ({len(nodes0) - 1} elements), shapes {c0}-{c0 + len(chunk) - 1}: "
                f"graph {ms:.1f} ms")
        h.close()
        log(f"shape batch: {sc} f0 handle of {slots} slots, build + capture {t_cap:.1f} s")
    t_wall = time.perf_counter() - t_wall
    # check: the smallest problem of shape 0 against its FP64 analytic gradient and
    # central differences of the FP64 forward on two nodes (same quadrature rules)
    sc0 = scales[0]
    _, nodes, f, th, mm = by_scale[sc0][0]
    J0, g0, (ne, nb) = results[(0, sc0)]
    ad = pkg("bor_adjoint")
    spec = ad.RcsSpec(nodes, f, th, mm, cfie_alpha=0.5, n_xi=(ne, nb))
    J64, g64 = ad.objective_and_gradient(spec)
    rel64 = float(np.max(np.abs(g0 - g64)) / np.max(np.abs(g64)))
    lam = bs.C0 / f
    fd_err = []
    for q in (len(nodes) // 3, 2 * len(nodes) // 3):
        vals = []
        for sgn in (1.0, -1.0):
            r = nodes[:, 0].copy()
            r[q] += sgn * 1e-6 * lam
            Fq, hq = ad.solve_far_field(spec, r)
            hq.close()
            vals.append(float(np.sum(ad.sigma_db(Fq))))
        fd = (vals[0] - vals[1]) / (2e-6 * lam)
        fd_err.append(abs(g0[q] - fd) / max(abs(fd), 1e-30))
        log(f"shape batch check: node {q} mixed graph {g0[q]:.9e} fp64 {g64[q]:.9e} fd {fd:.9e}")
    out = {
        "metric": "forward RCS + gradient (J = sum sigma_dBsm) of perturbed shapes, batched graph launches",
        "config": {"workload": f"ogive-1200seg-fourier16-perturbed-x{n_shapes}-3freq", "shapes": n_shapes,
                   "frequencies": scales, "elements": [len(by_scale[sc][0][1]) - 1 for sc in scales],
                   "m_max": [by_scale[sc][0][4] for sc in scales], "aspects": len(probs[0][3]),
                   "polarizations": 2,
                   "precision": "mixed (FP32 fill/LU, FP64 refinement: 1 sweep currents, 1 sweep lambda)",
                   "problems": len(probs), "graph_launches": n_launch, "budget_gb": budget_gb},
        "gpu_s": round(t_gpu, 3), "wall_s": round(t_wall, 1),
        "solves_per_s_gpu": round(solves / t_gpu, 1),
        "gradients_per_s_gpu": round(len(probs) / t_gpu, 3),
        "check_vs_fp64_max_rel": rel64, "check_vs_central_differences_max_rel": float(max(fd_err)),
        "J_shape0_vs_fp64": [J0, J64],
        "launches": rows,
    }
    print(json.dumps(out), flush=True)


def run_sweep(steps, warmup):
    """Config 3: 64 frequencies of one body in ONE batch (uneven systems of
    ~100-2000 unknowns), eagerly and as a replayed hipGraph, against the same
    problems run one handle per frequency.  Mixed precision, no refinement
    sweep (a captured graph has a fixed schedule); dB error vs the FP64 batch.
    Prints its own JSON line."""
    wl = pkg("workloads")
    bs = pkg("bor_solver")
    L = pkg("_lib")
    jobs = wl.frequency_sweep()
    specs = []
    for j in jobs:
        k = 2 * math.pi * j["freq_hz"] / bs.C0
        m_max, ne, nbk, _ = bs.discretisation(j["points"], k, j["thetas_deg"], j["n_modes"], "cfie")
        specs.append((j["points"], k, m_max, ne, nbk, j["thetas_deg"]))
    solves = sum((2 * s[2] + 1) * len(s[5]) for s in specs)
    nmax = max(2 * len(s[0]) for s in specs)
    log(f"sweep: {len(specs)} frequencies, {sum(2 * s[2] + 1 for s in specs)} systems (n up to ~{nmax})")

    def handle(prec, subset):
        h = L.Handle(prec, nb=64, residual=False)
        for (pts, k, m_max, ne, nbk, th) in subset:
            h.add_problem(pts, k, m_max, ne, nbk, 0.5, 0.5 * bs.ETA0, th)
        h.build()
        return h

    h = handle(L.GHOST_MIXED, specs)
    for _ in range(warmup):
        h.run(0)
    h.sync()
    L.device_synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        h.run(0)
    h.sync()
    L.device_synchronize()
    dt = (time.perf_counter() - t0) / steps
    dF_mx = [h.dF(i) for i in range(len(specs))]
    h.capture(0)
    h.launch()
    h.sync()
    L.device_synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        h.launch()
    h.sync()
    L.device_synchronize()
    dt_graph = (time.perf_counter() - t0) / steps
    dF_graph = [h.dF(i) for i in range(len(specs))]
    h.close()
    log(f"sweep: batch {dt * 1e3:.1f} ms, graph {dt_graph * 1e3:.1f} ms")
    # one handle per frequency, run one after the other (the reference's
    # frequency loop), each after a warm-up run
    hs = [handle(L.GHOST_MIXED, [s]) for s in specs]
    for hh in hs:
        hh.run(0)
    L.device_synchronize()
    t0 = time.perf_counter()
    for hh in hs:
        hh.run(0)
        hh.sync()
    L.device_synchronize()
    dt_loop = time.perf_counter() - t0
    for hh in hs:
        hh.close()
    log(f"sweep: per-frequency loop {dt_loop * 1e3:.1f} ms")
    h64 = handle(L.GHOST_FP64, specs)
    h64.run(0)
    h64.sync()
    err = 0.0
    graph_same = True
    for i, s in enumerate(specs):
        F64 = bs.truncate_modes(h64.dF(i), s[2], 0.0)[0]
        Fmx = bs.truncate_modes(dF_mx[i], s[2], 0.0)[0]
        err = max(err, float(np.max(np.abs(db(Fmx) - db(F64)))))
        graph_same = graph_same and bool(np.array_equal(dF_graph[i], dF_mx[i]))
    h64.close()
    out = {
        "metric": "BOR frequency-sweep solves/s (mode x freq x angle), 64 frequencies in one batch",
        "value": round(solves / dt, 2), "unit": "solves/s", "ms_per_step": round(dt * 1e3, 3),
        "graph_ms_per_step": round(dt_graph * 1e3, 3), "graph_solves_per_s": round(solves / dt_graph, 2),
        "per_frequency_loop_ms": round(dt_loop * 1e3, 3),
        "speedup_batch_vs_loop": round(dt_loop / dt, 2), "graph_replay_equals_eager": graph_same,
        "err_vs_fp64_max_abs_db": err, "steps": steps, "warmup": warmup,
        "config": {"workload": "smooth-body-1m x 64 frequencies 0.2-2 GHz", "frequencies": len(specs),
                   "systems": sum(2 * s[2] + 1 for s in specs), "solves_per_step": solves,
                   "n_elems": [len(specs[0][0]) - 1, len(specs[-1][0]) - 1],
                   "m_max": [specs[0][2], specs[-1][2]], "aspects": len(specs[0][5]),
                   "precision": "mixed", "refine_iters": 0},
    }
    print(json.dumps(out), flush=True)


def dielectric_cpu_baseline(nodes, freq, thetas, eps, mu, m_max, stride=10, near_sample=9):
    """The FP64 oracle's PMCHWT sweep (bor_solver.py:1019-1094 restated), one
    thread, timed per leg on a bounded sample and scaled to the full sweep:
    both media's far tables in full (shared by every mode and aspect), the
    near-band projections of ``near_sample`` element pairs per medium
    (scaled by the pair count), and per mode the block assembly (far part)
    and LU + solve with one aspect in ``stride`` plus the excitation and far
    field of those aspects (scaled by the stride)."""
    from threadpoolctl import threadpool_limits
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import bor_oracle as bo
    log("dielectric cpu baseline: oracle sample")
    th = np.asarray(thetas)[::stride]
    k0 = 2 * math.pi * freq / bo.C0
    body = bo.Body(nodes)
    Nn, Ne = body.Nn, body.Ne
    mi, er = bo.causal_medium(eps, mu)
    media = [(k0, bo.ETA0), (k0 * mi, bo.ETA0 * er)]
    pairs = bo.near_pairs(Ne)
    # stratified by cell kind (a self pair has 46 cells, a neighbour pair 13)
    by_kind = {}
    for pr in pairs:
        by_kind.setdefault(bo.near_kind(*pr), []).append(pr)
    per = max(1, near_sample // len(by_kind))
    pick = {kd: [v[i] for i in np.linspace(len(v) // 4, 3 * len(v) // 4, per).astype(int)]
            for kd, v in by_kind.items()}
    with threadpool_limits(limits=1):
        t0 = time.perf_counter()
        fars = []
        for k, _ in media:
            osc = bo.osc_for(body, k, streaming=False)
            fars.append(bo.FarBlocks(body, k, m_max, bo.n_xi_efie(osc, m_max), bo.n_xi_bracket(osc, m_max),
                                     efie=True, mfie=False, zs_pt=np.ones(body.P, dtype=np.complex128)))
        t_far = time.perf_counter() - t0
        t_near = 0.0
        for kd, sel in pick.items():
            t0 = time.perf_counter()
            for k, _ in media:
                for (e, f) in sel:
                    s_, sp, _w = bo.cell_points(kd)
                    p, q = body.element_points(e, s_), body.element_points(f, sp)
                    bo.green_orders_near(p["rho"], p["z"], q["rho"], q["z"], k, m_max)
                    bo.bracket_modes_near(p["rho"], p["z"], p["trho"], p["tz"], q["rho"], q["z"], q["trho"],
                                          q["tz"], k, m_max, ibc=True)
            t_near += (time.perf_counter() - t0) * len(by_kind[kd]) / len(sel)
        r2 = (bo.ETA0 / media[1][1]) ** 2
        t_sys = t_exc = 0.0
        for m in range(-m_max, m_max + 1):
            t0 = time.perf_counter()
            T, P = [], []
            for (k, eta), fb in zip(media, fars):
                b = fb.efie(m)
                T.append(1j * k * eta * 2.0 * np.pi * np.block([[b[0], b[1]], [b[2], b[3]]]))
                c = fb.ibc(m)
                P.append(np.block([[-c[1], c[0]], [-c[3], c[2]]]))
            Ps = bo.ETA0 * (P[0] + P[1])
            mask = np.tile(bo.basis_mask(body, m), 2)
            A = np.block([[T[0] + T[1], -Ps], [Ps, T[0] + r2 * T[1]]])[np.ix_(mask, mask)]
            t1 = time.perf_counter()
            cols = []
            for a in th:
                for pol in ("VV", "HH"):
                    hp = bo.rhs_efie(body, k0, m, a, "HH" if pol == "VV" else "VV")
                    cols.append(np.concatenate([bo.rhs_efie(body, k0, m, a, pol), -hp if pol == "VV" else hp])[mask])
            t2 = time.perf_counter()
            X = np.linalg.solve(A, np.stack(cols, axis=1))
            t3 = time.perf_counter()
            for it, a in enumerate(th):
                for ip in range(2):
                    full = np.zeros(4 * Nn, dtype=complex)
                    full[mask] = X[:, 2 * it + ip]
                    bo.farfield(body, k0, m, full[:2 * Nn], a, msol=bo.ETA0 * full[2 * Nn:])
            t4 = time.perf_counter()
            t_sys += (t1 - t0) + (t3 - t2)
            t_exc += (t2 - t1) + (t4 - t3)
    scale = len(thetas) / len(th)
    total = t_far + t_near + t_sys + t_exc * scale
    solves = (2 * m_max + 1) * len(thetas)
    return {"value": solves / total, "unit": "solves/s", "cores": 1, "kind": "port",
            "sample": (f"FP64 oracle PMCHWT, 1 thread: far tables of both media in full ({t_far:.1f} s); near "
                       f"projections of {sum(len(v) for v in pick.values())} of {len(pairs)} element pairs per medium "
                       f"(stratified by cell kind), scaled "
                       f"({t_near:.1f} s); {2 * m_max + 1} mode assemblies + LU/solves with {len(th)} of "
                       f"{len(thetas)} aspects ({t_sys:.1f} s, not rescaled); excitation + far field of "
                       f"{len(th)} aspects x {scale:.1f} ({t_exc * scale:.1f} s); est. total {total:.0f} s")}


def run_dielectric(steps, warmup, no_cpu):
    """Homogeneous penetrable body on PMCHWT: a lossy dielectric sphere
    (a = 0.5 m, eps_r = 2.5 - 0.1j, 300 MHz, 120 segments, 181 aspects, the
    reference's default mode count).  One step = both media's fills, the
    block assembly, batched LU/solve of the (J, M) systems and the far field,
    in FP64.  Prints its own JSON line."""
    wl = pkg("workloads")
    bs = pkg("bor_solver")
    L = pkg("_lib")
    at = pkg("autotune")
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import mie
    a, freq, eps, mu = 0.5, 3e8, 2.5 - 0.1j, 1.0
    nodes = wl.sphere(a, 120)
    thetas = np.arange(0.0, 181.0, 1.0)
    k0 = 2 * math.pi * freq / bs.C0
    m_max, ne, nbk, _ = bs.discretisation(nodes, k0, thetas, None, "efie", "tables")
    mi, er = bs.causal_medium(eps, mu)
    osc_i = 2.0 * abs(k0 * mi) * bs.gauss_rho_max(nodes)
    nxi_in = (bs.n_xi_efie(osc_i, m_max), bs.n_xi_bracket(osc_i, m_max))
    tiles = at.best(L.GHOST_FP64, 4 * len(nodes), m_max + 1, 2 * len(thetas))
    log(f"dielectric: m_max={m_max}, n_xi {ne}/{nbk} exterior, {nxi_in} interior, tiles {tiles}")
    h = L.Handle(L.GHOST_FP64, nb=tiles[0], residual=False)
    h.set_gemm_variant(tiles[1])
    h.add_problem(nodes, k0, m_max, ne, nbk, 1.0, 0.0, thetas)
    h.set_interior(0, k0 * mi, bs.ETA0 * er, *nxi_in)
    h.build()
    n_sys = h.system_info(0, 0)[0]
    for i in range(warmup):
        h.run(0)
        h.sync()
        log(f"dielectric: warmup {i + 1}/{warmup} done")
    L.device_synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        h.run(0)
    h.sync()
    L.device_synchronize()
    dt = (time.perf_counter() - t0) / steps
    log(f"dielectric: {steps} timed steps, {dt * 1e3:.1f} ms/step")
    phase = h.timings()
    h.set_profile(True)
    h.run(0)
    h.sync()
    kt = h.kernel_times()
    dF = h.dF(0)
    h.close()
    F = bs.truncate_modes(dF, m_max, 0.0)[0]
    ref = 10.0 * math.log10(mie.sigma_dielectric_sphere(a, eps, mu, freq))
    err = float(np.max(np.abs(db(F) - ref)))
    solves = (2 * m_max + 1) * len(thetas)
    # the assembly kernel against HBM: per assembled mode it reads the four
    # node-based operators' far blocks (2 Nn)^2 and writes the n x n system,
    # complex FP64 (the near band's few entries aside)
    Nn = len(nodes)
    n_asm = m_max + 1
    asm_bytes = n_asm * (4 * (2 * Nn) ** 2 + n_sys ** 2) * 16.0
    rows = {}
    for fam, peak_key, bound in (("far_fill", "fp64_mfma_tflops", "mfma"), ("near_modal", "fp64_vector_tflops", "valu"),
                                 ("assemble", "hbm_gbs", "hbm")):
        ms_f, n_f, fl_f = kt.get(fam, (0.0, 0, 0.0))
        if ms_f <= 0:
            continue
        if bound == "hbm":
            ach = asm_bytes / (ms_f * 1e-3) / 1e9
            rows[fam] = {"bound": "hbm", "achieved": round(ach, 1), "peak": MI355X_PEAK[peak_key], "unit": "GB/s",
                         "frac": round(ach / MI355X_PEAK[peak_key], 4), "ms_per_step": round(ms_f, 3),
                         "bytes_per_step": asm_bytes}
        elif fl_f > 0:
            ach = fl_f / (ms_f * 1e-3) / 1e12
            rows[fam] = {"bound": bound, "achieved": round(ach, 3), "peak": MI355X_PEAK[peak_key], "unit": "TFLOP/s",
                         "frac": round(ach / MI355X_PEAK[peak_key], 4), "ms_per_step": round(ms_f, 3)}
    gl = [v for kk, v in kt.items() if kk.startswith("gemm")]
    gemm_ms, gemm_fl = sum(v[0] for v in gl), sum(v[2] for v in gl)
    if gemm_ms > 0:
        ach = gemm_fl / (gemm_ms * 1e-3) / 1e12
        rows["gemm (all uses)"] = {"bound": "mfma", "achieved": round(ach, 3), "peak": MI355X_PEAK["fp64_mfma_tflops"],
                                   "unit": "TFLOP/s", "frac": round(ach / MI355X_PEAK["fp64_mfma_tflops"], 4),
                                   "ms_per_step": round(gemm_ms, 3)}
    out = {
        "metric": "BOR PMCHWT (homogeneous dielectric) monostatic solves/s (mode x angle)",
        "value": round(solves / dt, 2), "unit": "solves/s", "n_gpus": 1, "steps": steps, "warmup": warmup,
        "ms_per_step": round(dt * 1e3, 3), "higher_is_better": True, "dtype": "fp64", "data": "synthetic",
        "config": {"workload": "dielectric-sphere-a0.5m-eps2.5-0.1j-300MHz-120seg", "m_max": m_max,
                   "modes": 2 * m_max + 1, "aspects": len(thetas), "polarizations": 2, "system_n": n_sys,
                   "n_xi_exterior": [ne, nbk], "n_xi_interior": list(nxi_in), "lu_tiles": list(tiles)},
        "err_vs_mie_max_abs_db": err,
        "kernel_ms_per_step": {k: round(v[0], 3) for k, v in kt.items() if v[1]},
        "phase_ms_last_step": phase,
        "rooflines": rows,
        "cpu_baseline": None if no_cpu else dielectric_cpu_baseline(nodes, freq, thetas, eps, mu, m_max),
    }
    print(json.dumps(out), flush=True)


def bench_c0():
    return pkg("bor_solver").C0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--refine", type=int, default=-2,
                    help="FP64 refinement sweeps; negative = adaptive (error probe), up to |n|")
    ap.add_argument("--config", type=int, default=2)
    ap.add_argument("--shapes", type=int, default=32, help="config 5: number of perturbed shapes")
    ap.add_argument("--budget-gb", type=float, default=200.0,
                    help="config 5: device memory per batch handle (problems are grouped under it)")
    ap.add_argument("--gemm-variant", type=int, default=None,
                    help="GEMM tile variant of the headline path in place of the autotuner's")
    ap.add_argument("--lu-super", type=int, default=None,
                    help="LU panels per super-block of the headline path (default: the library's 4)")
    ap.add_argument("--precision", choices=("mixed", "fp64"), default="mixed",
                    help="path of the timed steps (fp64: for profiling the plain FP64 path)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--copy-calibration", type=float, default=0.0, metavar="GIB",
                    help="run one plain device copy of GIB GiB (byte-counter calibration) and exit")
    ap.add_argument("--dielectric", action="store_true",
                    help="time the PMCHWT sweep of a lossy dielectric sphere (120 segments) instead")
    ap.add_argument("--adjoint", action="store_true",
                    help="time forward + adjoint RCS gradient on the 600-element cone-sphere instead")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one MI355X only: runs on several GPUs are not supported "
              "(the sweep would shard by frequency and mode, one process per GPU).", file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration > 0:
        nbytes = args.copy_calibration * 2 ** 30
        ms = pkg("_lib").copy_calibration(nbytes)
        print(json.dumps({"copy_bytes": nbytes, "ms": ms, "read_plus_write_GBps": 2 * nbytes / ms / 1e6}))
        return
    if args.adjoint:
        run_adjoint(max(1, args.steps))
        return
    if args.dielectric:
        run_dielectric(max(1, args.steps), args.warmup, args.no_cpu)
        return
    if args.config == 3:
        run_sweep(max(1, args.steps), args.warmup)
        return
    if args.config == 5:
        run_shape_batch(args.shapes, args.budget_gb)
        return
    wl = pkg("workloads")
    name, nodes, freq, thetas, m_max = wl.config(args.config)
    solves = (2 * m_max + 1) * len(thetas)

    tiles_mx = tuned_tiles(args.precision, nodes, thetas, m_max)
    if args.gemm_variant is not None:
        tiles_mx = (tiles_mx[0], args.gemm_variant)
        log(f"{args.precision}: GEMM variant set to {args.gemm_variant} by --gemm-variant")
    dt_mx, dF_mx, kt_mx, ph_mx, nxi = run_gpu(nodes, freq, thetas, m_max, args.precision, args.steps,
                                               args.warmup, args.refine if args.precision == "mixed" else 0,
                                               tiles=tiles_mx, lu_super=args.lu_super)
    F_mx = pkg("bor_solver").truncate_modes(dF_mx, m_max, 0.0)[0]
    err_db = None
    fp64_rate = None
    kt_64 = {}
    tiles_64 = None
    if not args.no_fp64:
        tiles_64 = tuned_tiles("fp64", nodes, thetas, m_max)
        dt_64, dF_64, kt_64, ph_64, _ = run_gpu(nodes, freq, thetas, m_max, "fp64", 1, 1, 0, tiles=tiles_64)
        F_64 = pkg("bor_solver").truncate_modes(dF_64, m_max, 0.0)[0]
        s64, smx = db(F_64), db(F_mx)
        peak = np.max(s64)
        err_db = {
            "max_abs_db": float(np.max(np.abs(smx - s64))),
            "max_abs_db_above_peak_minus_40db": float(np.max(np.abs(smx - s64)[s64 > peak - 40.0])),
        }
        fp64_rate = solves / dt_64

    # rooflines from live HIP-event timing of the profiled step: the GEMM over all
    # its uses in the step (LU updates, back substitution, refinement solves; the
    # FP64 residual kernel aside), the far fill and the near band; "roofline" is
    # the one of the family that takes the most time per step
    gl = [v for kk, v in kt_mx.items() if kk.startswith("gemm") and kk != "gemm_fp64_residual"]
    gemm_ms, gemm_n, gemm_fl = (sum(v[0] for v in gl), sum(v[1] for v in gl), sum(v[2] for v in gl))
    merged = {kk: v for kk, v in kt_mx.items() if not (kk.startswith("gemm") and kk != "gemm_fp64_residual")}
    merged["gemm (all uses)"] = (gemm_ms, gemm_n, gemm_fl)
    fam, (ms, cnt, fl) = max(merged.items(), key=lambda kv: kv[1][0])
    fp32 = args.precision == "mixed"
    per = {}
    tpath = os.path.join(ROOT, "profiles", "ogive1200_m70_mixed_hbm_traffic.json")
    if args.config == 2 and fp32 and os.path.exists(tpath):
        with open(tpath) as f:
            per = json.load(f)["per_launch"]
    traffic_unit = "bytes/launch (rocprofv3 FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0)"
    kinds = {}
    if gemm_n > 0 and gemm_ms > 0:
        achieved = gemm_fl / (gemm_ms * 1e-3) / 1e12
        split = fp32 and tiles_mx[1] == 5
        key = "gemm_bf3" if split else "gemm"
        traffic = per[key]["hbm_bytes"] if key in per else None
        if split:
            # the pipe doing the work is the bf16 MFMA: each FP32 product is six bf16 products
            issued = SPLIT_BF16_PRODUCTS * achieved
            pk = MI355X_PEAK["bf16_mfma_tflops"]
            g = {"kernel": ("k_gemm_bf3 (complex FP32 GEMM on the bf16 MFMA 32x32x16: operands split into "
                            "three bf16 parts, six products each; 64x64 tiles; LU updates and solves)"),
                 "bound": "mfma", "achieved": round(issued, 3), "peak": pk, "unit": "TFLOP/s",
                 "frac": round(issued / pk, 4),
                 "achieved_note": "bf16 MFMA flops issued = 6 x the algorithmic complex FP32 GEMM flops",
                 "fp32_equivalent": {"achieved": round(achieved, 3), "peak": MI355X_PEAK["fp32_mfma_tflops"],
                                     "frac": round(achieved / MI355X_PEAK["fp32_mfma_tflops"], 4)}}
        else:
            pk = MI355X_PEAK["fp32_mfma_tflops" if fp32 else "fp64_mfma_tflops"]
            g = {"kernel": ("k_gemm2<float,2,2> (complex FP32 MFMA 16x16x4, 64x64 tiles; LU updates and solves)"
                            if fp32 else "k_gemm2<double,...> (complex FP64 MFMA 16x16x4; LU updates and solves)"),
                 "bound": "mfma", "achieved": round(achieved, 3), "peak": pk,
                 "unit": "TFLOP/s", "frac": round(achieved / pk, 4)}
        g.update({"traffic": traffic, "traffic_unit": traffic_unit,
                  "launches": gemm_n, "avg_launch_ms": round(gemm_ms / gemm_n, 4),
                  "flops_per_launch": gemm_fl / gemm_n, "ms_per_step": round(gemm_ms, 3)})
        kinds["gemm"] = g
    # the far fill's MFMA flops issued against the FP32 MFMA peak (it also issues the sample
    # arithmetic on the VALU, which does not co-issue with MFMA on gfx950), the near band's
    # projection flops against the FP64 vector peak (its FP32 tail runs at the same issue rate
    # with packed FP32 off)
    for fk, peak_key, what in (("far_fill", "fp32_mfma_tflops" if fp32 else "fp64_mfma_tflops",
                                "k_far_fill (MFMA azimuth transform + VALU samples)"),
                               ("near_modal", "fp64_vector_tflops", "k_near_modal (FP64 core / FP32 tail VALU projection)")):
        ms_f, n_f, fl_f = kt_mx.get(fk, (0.0, 0, 0.0))
        if ms_f > 0 and fl_f > 0 and n_f > 0:
            ach = fl_f / (ms_f * 1e-3) / 1e12
            kinds[fk] = {"kernel": what, "bound": "mfma" if fk == "far_fill" else "valu",
                         "achieved": round(ach, 3), "peak": MI355X_PEAK[peak_key], "unit": "TFLOP/s",
                         "frac": round(ach / MI355X_PEAK[peak_key], 4),
                         "traffic": per[fk]["hbm_bytes"] if fk in per else None, "traffic_unit": traffic_unit,
                         "launches": n_f, "avg_launch_ms": round(ms_f / n_f, 4), "flops_per_launch": fl_f / n_f,
                         "ms_per_step": round(ms_f, 3)}
    dom = "gemm" if fam == "gemm (all uses)" else fam
    roof = kinds.get(dom, kinds.get("gemm"))
    out = {
        "metric": "BOR monostatic solves/s (mode x freq x angle) at <=0.01 dB vs FP64; MFMA/HBM %",
        "value": round(solves / dt_mx, 2),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(dt_mx * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32+fp64" if args.precision == "mixed" else "fp64",
        "data": "synthetic (seeded generatrix)",
        "config": {"workload": name, "modes": 2 * m_max + 1, "m_max": m_max, "aspects": len(thetas),
                   "polarizations": 2, "frequencies": 1, "n_elems": int(len(nodes) - 1),
                   "n_xi": list(nxi), "formulation": "cfie(alpha=0.5)", "precision": args.precision,
                   "refine_iters": args.refine, "solves_per_step": solves},
        "fp64_path_solves_per_s": round(fp64_rate, 2) if fp64_rate else None,
        "speedup_vs_fp64_path": round((solves / dt_mx) / fp64_rate, 3) if fp64_rate else None,
        "err_vs_fp64": err_db,
        "dominant_kernel": {"family": fam, "ms_per_step": round(ms, 3), "launches_per_step": cnt},
        "kernel_ms_per_step": {k: round(v[0], 3) for k, v in kt_mx.items() if v[1]},
        "fp64_kernel_ms": {k: round(v[0], 3) for k, v in kt_64.items() if v[1]},
        "phase_ms_last_step": ph_mx,
        "autotuned_lu_tiles": {args.precision: list(tiles_mx), "fp64": list(tiles_64) if tiles_64 else None,
                               "cache": pkg("autotune").cache_path()},
        "roofline": roof,
        "roofline_kernels": kinds,
        "cpu_baseline": None if args.no_cpu else cpu_baseline_timed(nodes, freq, thetas, m_max),
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
