"""GPU drop-in for the reference's bor_solver PEC call surface.

Same signatures and result keys as bor_solver.solve_bor (bor_solver.py:820-1004)
for PEC bodies (EFIE / MFIE / CFIE), with the fill, the per-mode solves and
the far-field sum running in libghost_bor on the current HIP device.  All 2
m_max + 1 mode systems of a frequency are filled, factored and solved as one
batch; the reference's mode truncation (two consecutive quiet |m|,
bor_solver.py:782-793) is then applied to the per-mode far-field
contributions in the same order, so ``modes_used`` and the summed amplitudes
follow the reference's rule.

Extra keyword arguments (not in the reference): ``precision`` ("fp64" or
"mixed"), ``refine_iters`` (mixed precision FP64 refinement sweeps; negative
= adaptive, up to |refine_iters| sweeps while the estimated relative error
of the currents exceeds 1e-5), ``residual`` (compute linear_residual,
default True).
"""

from __future__ import annotations

import math
from typing import Callable, Dict, List, Optional, Sequence

import numpy as np

from . import _lib

C0 = 299_792_458.0
ETA0 = 376.730313668
AXIS_TOL = 1e-12


# ─────────────────────────────────────────────────────────────────────────────
# canonical generatrices (bor_solver.py:2179-2190)
# ─────────────────────────────────────────────────────────────────────────────

def sphere_generatrix(a: float, n: int) -> np.ndarray:
    """North pole (+z) to south pole, n elements."""
    th = np.linspace(0.0, math.pi, n + 1)
    return np.column_stack([a * np.sin(th), a * np.cos(th)])


def cylinder_generatrix(a: float, L: float, n_rad: int, n_len: int) -> np.ndarray:
    """Closed cylinder: top cap centre -> rim -> side -> bottom rim -> centre."""
    top = np.column_stack([np.linspace(0.0, a, n_rad + 1), np.full(n_rad + 1, L / 2)])
    side = np.column_stack([np.full(n_len - 1, a), np.linspace(L / 2, -L / 2, n_len + 1)[1:-1]])
    bot = np.column_stack([np.linspace(a, 0.0, n_rad + 1), np.full(n_rad + 1, -L / 2)])
    return np.vstack([top, side, bot])


# ─────────────────────────────────────────────────────────────────────────────
# discretisation rules shared with the reference
# ─────────────────────────────────────────────────────────────────────────────

def n_xi_efie(osc: float, m_max: int) -> int:
    """bor_kernels.py:149-152."""
    return int(2 ** math.ceil(math.log2(max(64, 4 * (m_max + 2), 6 * (osc + 4)))))


def n_xi_bracket(osc: float, m_max: int) -> int:
    """bor_kernels.py:315-319."""
    return int(2 ** math.ceil(math.log2(max(128, 6 * (m_max + 2), 8 * (osc + 4)))))


def estimate_bor_table_gb(n_elems: int, m_max: int, formulation: str = "cfie", has_ibc: bool = False,
                          gauss_order: int = 4, single_tables: bool = False) -> float:
    """Far-table memory the reference would hold (bor_solver.py:804-817); it
    decides which azimuth-grid rule the reference applies."""
    P = float(gauss_order * n_elems)
    per = 8.0 if single_tables else 16.0
    total = P * P * (m_max + 2) * per
    if formulation in ("cfie", "mfie"):
        total += 4.0 * P * P * (2 * m_max + 1) * per
    if has_ibc:
        total += 4.0 * P * P * (2 * m_max + 1) * per
    return total / 1e9


def _validate_nodes(points) -> np.ndarray:
    pts = np.array(points, dtype=float)
    if pts.ndim != 2 or pts.shape[1] != 2 or pts.shape[0] < 2:
        raise ValueError("Generatrix needs an (Nn, 2) array of (rho, z) nodes.")
    if np.any(pts[:, 0] < -1e-12):
        raise ValueError("Generatrix rho coordinates must be >= 0.")
    pts[:, 0] = np.maximum(pts[:, 0], 0.0)
    d = np.diff(pts, axis=0)
    if np.any(np.hypot(d[:, 0], d[:, 1]) <= 0):
        raise ValueError("Generatrix has a zero-length element.")
    return pts


def gauss_rho_max(pts: np.ndarray, order: int = 4) -> float:
    """Largest rho over the order-4 Gauss points (the table path's oscillation source)."""
    x, _ = np.polynomial.legendre.leggauss(order)
    s = 0.5 * (x + 1.0)
    r0, r1 = pts[:-1, 0], pts[1:, 0]
    return float(np.max(r0[:, None] + s[None, :] * (r1 - r0)[:, None]))


def discretisation(pts: np.ndarray, k: float, thetas: np.ndarray, n_modes: Optional[int], form: str,
                   assembly: str = "auto", has_ibc: bool = False):
    """(m_max, n_xi_efie, n_xi_bracket, streaming) as the reference picks them
    (bor_solver.py:842-847, 870-885; bor_kernels.py:149-152, 315-319;
    bor_streaming.py:179-183).  has_ibc: the impedance tables count toward
    the table-size estimate that selects the streamed rule."""
    rho_max = float(np.max(pts[:, 0]))
    if n_modes is None:
        st_max = float(np.max(np.abs(np.sin(np.radians(thetas)))))
        n_modes = int(math.ceil(k * rho_max * max(st_max, 0.05))) + 12
    m_max = int(n_modes)
    asm = str(assembly).strip().lower()
    if asm not in ("auto", "tables", "streaming"):
        raise ValueError("assembly must be 'auto', 'tables', or 'streaming'.")
    streaming = asm == "streaming" or (asm == "auto" and
                                       estimate_bor_table_gb(len(pts) - 1, m_max, form, has_ibc) > 2.0)
    osc = 2.0 * abs(k) * (rho_max if streaming else gauss_rho_max(pts))
    return m_max, n_xi_efie(osc, m_max), n_xi_bracket(osc, m_max), streaming


def truncate_modes(dF: np.ndarray, m_max: int, mode_tol: float):
    """Per-|m| accumulation in mode order with the two-quiet-modes stop
    (bor_solver.py:771-796).  dF: [2 m_max + 1, 2, n_theta] (index m + m_max).
    Returns (F [2, n_theta], modes_used)."""
    F = np.zeros(dF.shape[1:], dtype=np.complex128)
    quiet = 0
    used = 0
    for am in range(m_max + 1):
        d = dF[m_max].copy() if am == 0 else dF[m_max + am] + dF[m_max - am]
        F = F + d
        used = am
        scale = max(float(np.max(np.abs(F))), 1e-30)
        if float(np.max(np.abs(d))) < mode_tol * scale:
            quiet += 1
            if quiet >= 2:
                break
        else:
            quiet = 0
    return F, used


_PRECISION = {"fp64": _lib.GHOST_FP64, "double": _lib.GHOST_FP64, "mixed": _lib.GHOST_MIXED}


def impedance_per_element(zs, n_elems: int) -> Optional[np.ndarray]:
    """Z_s as one complex value per element, or None when the body is PEC
    (bor_solver.py:850-866: a scalar applies to every element)."""
    if zs is None:
        return None
    z = np.asarray(zs, dtype=np.complex128)
    if z.ndim == 0:
        z = np.full(n_elems, complex(z))
    elif z.shape != (n_elems,):
        raise ValueError(f"zs needs one complex value per generatrix element ({n_elems}), got shape {z.shape}.")
    if not np.all(np.isfinite(z)):
        raise ValueError("zs holds non-finite values.")
    return None if np.all(z == 0.0) else z


def causal_medium(eps_r, mu_r):
    """(m, eta_r) of a homogeneous medium: the refractive index sqrt(eps mu)
    on the branch with Im(m) <= 0 (a passive medium under e^{j w t}) and
    Re(m) >= 0, and the relative wave impedance mu_r / m (bor_solver.py:
    161-172)."""
    m = complex(np.sqrt(complex(eps_r) * complex(mu_r)))
    if m.imag > 0.0:
        m = -m
    if m.real < 0.0:
        m = -m
    return m, complex(mu_r) / m


def _ibc_warnings(pts: np.ndarray, zs: Optional[np.ndarray], form: str) -> List[str]:
    """A closed body on the EFIE with a purely reactive Z_s has undamped
    cavity resonances (the reference's guard, bor_solver.py:944-968)."""
    if zs is None or form != "efie":
        return []
    tol = AXIS_TOL * max(1.0, float(np.max(pts[:, 0])))
    closed = pts[0, 0] <= tol and pts[-1, 0] <= tol
    if closed and float(np.max(np.abs(zs.real))) < 1e-3 * float(np.max(np.abs(zs))):
        return ["Closed body with a lossless (purely reactive) surface impedance on the EFIE: its "
                "interior resonances are undamped, so results near a resonant frequency may be "
                "wrong without any sign; add loss to Z_s or check the sweep for spikes."]
    return []


def _coefficients(form: str, cfie_alpha: float):
    alpha = float(cfie_alpha) if form == "cfie" else (1.0 if form == "efie" else 0.0)
    return alpha, (1.0 - alpha) * ETA0


# ─────────────────────────────────────────────────────────────────────────────
# batched sweeps
# ─────────────────────────────────────────────────────────────────────────────

def _result(pts, thetas, form, m_max, ne, nbk, streaming, dF, res, mode_tol, prec, refine_iters, times,
            probe=None, warnings=()):
    F, used = truncate_modes(dF, m_max, mode_tol)
    out = {
        "theta_deg": thetas.tolist(),
        "sigma_vv": (4.0 * math.pi * np.abs(F[0]) ** 2).tolist(),
        "sigma_hh": (4.0 * math.pi * np.abs(F[1]) ** 2).tolist(),
        "amp_vv": F[0].tolist(),
        "amp_hh": F[1].tolist(),
        "modes_used": used,
        "n_unknowns": int(2 * len(pts)),
        "formulation": form,
        "assembly": "streaming" if streaming else "tables",
        "table_precision": "single" if prec == "mixed" else "double",
        "stream_mode_block": None,
        "stream_sweeps": 0,
        "warnings": list(warnings),
        "precision": prec,
        "refine_iters": refine_iters if prec == "mixed" else 0,
        "probe_errors": probe,
        "n_xi": (ne, nbk),
        "m_max": m_max,
        "timings_ms": times,
    }
    if res is not None:
        # max over the modes the truncated sweep used (bor_solver.py:754-756, 784)
        ms = np.arange(-m_max, m_max + 1)
        out["linear_residual"] = float(np.max(res[np.abs(ms) <= used]))
    else:
        out["linear_residual"] = float("nan")
    return out


def solve_bor_batch(jobs: Sequence[Dict], *, precision: str = "fp64", refine_iters: int = -3,
                    residual: bool = True, nb: Optional[int] = None, mode_tol: float = 1e-6,
                    graph: bool = False, check_abort: Optional[Callable] = None,
                    tune: bool = True) -> List[Dict]:
    """Solve several (generatrix, frequency) problems as ONE batch: every mode
    system of every problem goes through the same batched fill, LU/solve and
    far-field launches (systems of uneven size share each launch).  Each job
    is a dict with keys points, freq_hz, thetas_deg and optionally
    formulation ("cfie"), cfie_alpha (0.5), n_modes, assembly, zs (surface
    impedance: scalar or one complex value per element, EFIE only), medium
    ((eps_r, mu_r): a homogeneous penetrable body on PMCHWT, FP64 only; the
    formulation key is then ignored).  Returns one solve_bor-style dict per
    job (solve_bor_dielectric's for a penetrable body).  graph=True captures the batch in a
    hipGraph and replays it.  nb = None takes the LU block size and GEMM tile
    from the autotuner (``autotune.best`` for the batch's largest system,
    system count and column count; measured once per shape bucket and
    device when tune is true, else its defaults); an explicit nb keeps the
    default tile.  The choice is reported as ``lu_tiles``."""
    prec = str(precision).strip().lower()
    if prec not in _PRECISION:
        raise ValueError("precision must be 'fp64' or 'mixed'.")
    specs = []
    for j in jobs:
        pts = _validate_nodes(j["points"])
        thetas = np.atleast_1d(np.asarray(j["thetas_deg"], dtype=float))
        k = 2.0 * math.pi * float(j["freq_hz"]) / C0
        med = None
        if j.get("medium") is not None:
            # PMCHWT (bor_solver.py:1019-1094): both media on the table rules
            if j.get("zs") is not None:
                raise ValueError("A penetrable body takes no surface impedance.")
            if prec != "fp64":
                raise ValueError("Penetrable bodies (complex k) run in FP64: use precision='fp64'.")
            eps, mu = complex(j["medium"][0]), complex(j["medium"][1])
            mi, er = causal_medium(eps, mu)
            form, zs = "pmchwt", None
            m_max, ne, nbk, streaming = discretisation(pts, k, thetas, j.get("n_modes"), "efie", "tables")
            osc_i = 2.0 * abs(k * mi) * gauss_rho_max(pts)
            med = {"k_in": k * mi, "eta_in": ETA0 * er, "eps_r": eps, "mu_r": mu,
                   "n_xi_in": (n_xi_efie(osc_i, m_max), n_xi_bracket(osc_i, m_max))}
            cE, cM = 1.0, 0.0
        else:
            form = str(j.get("formulation", "cfie")).strip().lower()
            if form not in ("efie", "mfie", "cfie"):
                raise ValueError("formulation must be 'efie', 'mfie' or 'cfie'.")
            zs = impedance_per_element(j.get("zs"), len(pts) - 1)
            if zs is not None and form == "cfie":
                raise ValueError("A surface impedance needs formulation='efie' (CFIE with Z_s is not "
                                 "defined here; a lossy Z_s already damps interior resonances).")
            m_max, ne, nbk, streaming = discretisation(pts, k, thetas, j.get("n_modes"), form,
                                                       j.get("assembly", "auto"), has_ibc=zs is not None)
            cE, cM = _coefficients(form, j.get("cfie_alpha", 0.5))
        specs.append((pts, thetas, form, k, m_max, ne, nbk, streaming, cE, cM, zs, med))
    if check_abort is not None:
        check_abort()
    variant = -1
    if nb is None:
        from . import autotune
        n_max = max((4 if sp[11] else 2) * len(sp[0]) for sp in specs)
        nsys = sum(sp[4] + 1 for sp in specs)
        nrhs = max(2 * len(sp[1]) for sp in specs)
        nb, variant = autotune.best(_PRECISION[prec], n_max, nsys, nrhs, tune=tune)
    h = _lib.Handle(_PRECISION[prec], nb=nb, residual=residual)
    try:
        h.set_gemm_variant(variant)
        for (pts, thetas, form, k, m_max, ne, nbk, streaming, cE, cM, zs, med) in specs:
            idx = h.add_problem(pts, k, m_max, ne, nbk, cE, cM, thetas)
            if zs is not None:
                # the impedance operator carries the EFIE's weight (bor_solver.py:929-934)
                h.set_impedance(idx, zs, cE)
            if med is not None:
                h.set_interior(idx, med["k_in"], med["eta_in"], *med["n_xi_in"])
        h.build()
        it = refine_iters if prec == "mixed" else 0
        if graph:
            h.capture(abs(it))
            h.launch()
        else:
            h.run(it)
        dFs = [h.dF(i) for i in range(len(specs))]
        probe = h.refine_history() if prec == "mixed" else []
        res = h.residuals() if residual else None
        times = h.timings() if not graph else {"graph_ms": h.graph_ms()}
    finally:
        h.close()
    if check_abort is not None:
        check_abort()
    outs, off = [], 0
    for i, (pts, thetas, form, k, m_max, ne, nbk, streaming, cE, cM, zs, med) in enumerate(specs):
        nm = 2 * m_max + 1
        r = res[off:off + nm] if res is not None else None
        off += nm
        outs.append(_result(pts, thetas, form, m_max, ne, nbk, streaming, dFs[i], r, mode_tol, prec,
                            refine_iters, times, probe, _ibc_warnings(pts, zs, form)))
        outs[-1]["lu_tiles"] = (int(nb), int(variant))
        if med is not None:
            outs[-1].update({"n_unknowns": int(4 * len(pts)), "eps_r": med["eps_r"], "mu_r": med["mu_r"],
                             "n_xi_interior": med["n_xi_in"]})
    return outs


# ─────────────────────────────────────────────────────────────────────────────
# solve_bor
# ─────────────────────────────────────────────────────────────────────────────

def solve_bor(points, freq_hz: float, thetas_deg, formulation: str = "efie",
              cfie_alpha: float = 0.5, zs=None, n_modes: Optional[int] = None,
              gauss_order: int = 4, mode_tol: float = 1e-6, workers: int = 1,
              progress: Optional[Callable] = None,
              check_abort: Optional[Callable] = None,
              table_precision: str = "auto", assembly: str = "auto",
              stream_budget_gb: float = 8.0, *, precision: str = "fp64",
              refine_iters: int = -3, residual: bool = True, nb: Optional[int] = None) -> Dict:
    """Monostatic RCS of a PEC body of revolution at aspects thetas_deg (from +z).

    Mirrors bor_solver.solve_bor (bor_solver.py:820-1004): PEC bodies
    (zs = None) on EFIE / MFIE / CFIE, and impedance bodies (zs a complex
    scalar or one value per element, IBC-EFIE) on the EFIE.  ``workers``,
    ``table_precision`` and
    ``stream_budget_gb`` are accepted for signature compatibility: the GPU
    path holds every table in FP64 (or FP32 far blocks under
    precision="mixed") and needs no host threads or re-sweeps.
    """
    form = formulation.strip().lower()
    if form not in ("efie", "mfie", "cfie"):
        raise ValueError("formulation must be 'efie', 'mfie' or 'cfie'.")
    if int(gauss_order) != 4:
        raise ValueError("the GPU BoR path is built for gauss_order = 4 (the reference default).")
    tp = str(table_precision).strip().lower()
    if tp not in ("auto", "single", "double"):
        raise ValueError("table_precision must be 'auto', 'single', or 'double'.")
    out = solve_bor_batch([{"points": points, "freq_hz": freq_hz, "thetas_deg": thetas_deg,
                            "formulation": form, "cfie_alpha": cfie_alpha, "n_modes": n_modes,
                            "assembly": assembly, "zs": zs}],
                          precision=precision, refine_iters=refine_iters, residual=residual, nb=nb,
                          mode_tol=mode_tol, check_abort=check_abort)[0]
    if progress is not None:
        progress(out["modes_used"], out["m_max"])
    return out


def solve_bor_pec(points, freq_hz: float, thetas_deg, n_modes: Optional[int] = None,
                  gauss_order: int = 4, mode_tol: float = 1e-6) -> Dict:
    """PEC EFIE entry point (bor_solver.py:1007-1012)."""
    return solve_bor(points, freq_hz, thetas_deg, formulation="efie", n_modes=n_modes,
                     gauss_order=gauss_order, mode_tol=mode_tol)


def solve_bor_dielectric(points, freq_hz: float, thetas_deg, eps_r: complex, mu_r: complex = 1.0,
                         n_modes: Optional[int] = None, gauss_order: int = 4, mode_tol: float = 1e-6,
                         workers: int = 1, progress: Optional[Callable] = None,
                         check_abort: Optional[Callable] = None, *, residual: bool = True,
                         nb: Optional[int] = None) -> Dict:
    """Monostatic RCS of a closed homogeneous penetrable body of revolution
    by per-mode PMCHWT (bor_solver.py:1019-1094): unknowns J and M / eta0,
    the exterior (free space) and interior (eps_r, mu_r; lossy when Im < 0)
    operators filled on the device, the 2 x 2 block systems assembled,
    factored and solved there, and the far field radiating both currents.
    ``workers`` is accepted for signature compatibility."""
    if int(gauss_order) != 4:
        raise ValueError("the GPU BoR path is built for gauss_order = 4 (the reference default).")
    out = solve_bor_batch([{"points": points, "freq_hz": freq_hz, "thetas_deg": thetas_deg,
                            "n_modes": n_modes, "medium": (eps_r, mu_r)}],
                          precision="fp64", residual=residual, nb=nb, mode_tol=mode_tol,
                          check_abort=check_abort)[0]
    if progress is not None:
        progress(out["modes_used"], out["m_max"])
    return out


# ─────────────────────────────────────────────────────────────────────────────
# generic mode loop (bor_solver.py:711-801) on the batched device solve
# ─────────────────────────────────────────────────────────────────────────────

def _mode_sweep(n_dofs: int, thetas, pols, m_max: int, mode_tol: float,
                assemble: Callable, rhs: Callable, farfield: Callable,
                prepare: Optional[Callable] = None, workers: int = 1,
                progress: Optional[Callable] = None,
                check_abort: Optional[Callable] = None,
                monitor_cond: bool = False):
    """The reference's closure-driven adaptive mode loop with the same
    contract (bor_solver.py:711-735): assemble(m) -> (A, mask), rhs(m, theta,
    pol) -> V (full, or reduced when mask is None), farfield(m, sol, theta,
    pol) -> complex.  The dense solves of a wave of |m| go to the device as
    one batched LU (ghost_bor_add_dense / set_dense / solve) in place of
    np.linalg.solve per mode; the accumulation and the two-quiet-modes stop
    run in mode order as in the reference, so F, modes_used and the linear
    residual follow the same rule.  A wave holds max(workers, 8) values of
    |m|.  monitor_cond evaluates numpy.linalg.cond of each assembled system
    on the host (a diagnostic, as in the reference)."""
    thetas = np.atleast_1d(np.asarray(thetas, dtype=float))
    pols = list(pols)
    F = np.zeros((len(pols), len(thetas)), dtype=np.complex128)
    if prepare is not None:
        prepare(m_max)
    wave_n = max(8, int(workers))
    modes_used, quiet, am, max_res = 0, 0, 0, 0.0
    conds: List[float] = []
    done = False
    while am <= m_max and not done:
        if check_abort is not None:
            check_abort()
        wave = list(range(am, min(am + wave_n, m_max + 1)))
        systems = []
        for w_am in wave:
            for m in ([0] if w_am == 0 else [w_am, -w_am]):
                A, mask = assemble(m)
                cols = [rhs(m, th, pol) if mask is None else np.asarray(rhs(m, th, pol))[mask]
                        for th in thetas for pol in pols]
                systems.append((w_am, m, np.asarray(A, dtype=np.complex128), np.stack(cols, axis=1), mask))
        h = _lib.Handle(_lib.GHOST_FP64, residual=True)
        try:
            idx = [h.add_dense(A.shape[0], B.shape[1]) for (_, _, A, B, _) in systems]
            h.build()
            for i, (_, _, A, B, _) in zip(idx, systems):
                h.set_dense(i, A, B)
            h.solve(0)
            h.sync()
            res = h.residuals()
            sols = [h.get_dense(i) for i in idx]
        finally:
            h.close()
        dF_am = {w: np.zeros_like(F) for w in wave}
        res_am = {w: 0.0 for w in wave}
        cond_am = {w: 0.0 for w in wave}
        for (w_am, m, A, B, mask), X, r in zip(systems, sols, res):
            if float(np.linalg.norm(B)) > 0.0:
                res_am[w_am] = max(res_am[w_am], float(r))
            if monitor_cond:
                cond_am[w_am] = max(cond_am[w_am], float(np.linalg.cond(A)))
            ci = 0
            for it, th in enumerate(thetas):
                for ip, pol in enumerate(pols):
                    if mask is None:
                        sol = X[:, ci]
                    else:
                        sol = np.zeros(n_dofs, dtype=np.complex128)
                        sol[mask] = X[:, ci]
                    ci += 1
                    dF_am[w_am][ip, it] += farfield(m, sol, th, pol)
        for w_am in wave:
            dF = dF_am[w_am]
            F += dF
            max_res = max(max_res, res_am[w_am])
            if monitor_cond:
                conds.append(cond_am[w_am])
            modes_used = w_am
            scale = max(float(np.max(np.abs(F))), 1e-30)
            if float(np.max(np.abs(dF))) < mode_tol * scale:
                quiet += 1
                if quiet >= 2:
                    done = True
                    break
            else:
                quiet = 0
        am = wave[-1] + 1
        if progress is not None:
            progress(modes_used, m_max)
    stats = {"linear_residual": max_res}
    if monitor_cond and conds:
        stats["max_cond"] = max(conds)
        stats["median_cond"] = float(np.median(conds))
    return F, modes_used, stats


# ─────────────────────────────────────────────────────────────────────────────
# per-mode operator factory (bor_solver.py:175-708, PEC in air)
# ─────────────────────────────────────────────────────────────────────────────

class BorPecSolver:
    """The reference's single-surface operator factory with its per-mode
    methods, each answered by the device: every assemble_* fills all modes
    |m| <= m_max of one formulation in one batch (unmasked node-based
    [2Nn, 2Nn] systems, cached per m_max), rhs_* run the excitation kernels
    for one aspect, farfield_mode runs the far-field kernels on a caller's
    solution (with the impedance body's radiating magnetic current when
    zs_pt is given, or an independent one, msol).  medium=(eps_r, mu_r)
    fills the operators of that homogeneous medium: k = k0 m with the
    causal index m and eta = eta0 mu_r / m (bor_solver.py:175-199)."""

    def __init__(self, points, freq_hz: float, gauss_order: int = 4,
                 near_depth: int = 4, medium=None, single_tables: bool = False):
        if int(gauss_order) != 4 or int(near_depth) != 4:
            raise NotImplementedError("the GPU BoR path is built for gauss_order = 4 and near_depth = 4.")
        self.pts = _validate_nodes(points)
        self.freq_hz = float(freq_hz)
        self.k0 = 2.0 * math.pi * self.freq_hz / C0
        self.medium = None if medium is None else (complex(medium[0]), complex(medium[1]))
        if self.medium is None:
            self.k, self.eta = self.k0, ETA0
        else:
            mi, er = causal_medium(*self.medium)
            self.k, self.eta = self.k0 * mi, ETA0 * er
        self.gauss_order = 4
        self.near_depth = 4
        self.Nn = int(self.pts.shape[0])
        self.P = 4 * (self.Nn - 1)
        self._m_max = None
        self._ops: Dict = {}
        self._rhs: Dict = {}

    # set-up (the device fill has one path: both are hints)
    def enable_streaming(self, m_max: int, efie: bool = True, mfie: bool = False,
                         ibc_zs_pt=None, single_blocks: bool = False, tile_budget_gb: float = 1.0,
                         workers: int = 1, mode_block=None) -> None:
        self._m_max = max(int(m_max), self._m_max or 0)

    def prepare_operators(self, m_max: int, efie: bool = True, mfie: bool = False,
                          ibc: bool = False, workers: int = 1) -> None:
        """The impedance operator depends on Z_s, so it is filled at
        assemble_ibc_extra time."""
        if efie:
            self._systems("efie", m_max)
        if mfie:
            self._systems("mfie", m_max)

    def _n_xi(self, m_max: int):
        osc = 2.0 * abs(self.k) * gauss_rho_max(self.pts)
        return n_xi_efie(osc, m_max), n_xi_bracket(osc, m_max)

    def _elem_zs(self, zs_pt) -> np.ndarray:
        """Per-element Z_s from the per-Gauss-point array (constant on each element)."""
        z = np.asarray(zs_pt, dtype=np.complex128).reshape(self.Nn - 1, self.gauss_order)
        return z[:, 0].copy()

    def _systems(self, kind: str, m_max: int, zs_elem=None) -> Dict[int, np.ndarray]:
        key = (kind, int(m_max)) if zs_elem is None else (kind, int(m_max), np.asarray(zs_elem).tobytes())
        if key not in self._ops:
            ne, nb = self._n_xi(m_max)
            cE, cM = {"efie": (1.0, 0.0), "mfie": (0.0, 1.0), "ibc": (0.0, 0.0), "pv": (0.0, 0.0)}[kind]
            h = _lib.Handle(_lib.GHOST_FP64, residual=False)
            try:
                h.add_problem(self.pts, complex(self.k).real, m_max, ne, nb, cE, cM, [0.0], unmasked=True)
                if kind == "ibc":
                    h.set_impedance(0, zs_elem, 1.0)
                elif kind == "pv":   # bare rotated-PV brackets: unit Z_s, no mass term
                    h.set_impedance(0, np.ones(self.Nn - 1), 1.0, 0.0)
                if self.medium is not None:
                    h.set_medium(0, complex(self.k).imag, self.eta)
                h.build()
                h.fill()
                h.sync()
                self._ops[key] = {m: h.system(0, m)[0] for m in range(-m_max, m_max + 1)}
            finally:
                h.close()
        return self._ops[key]

    def assemble_mode(self, m: int, m_max: int) -> np.ndarray:
        """EFIE [2Nn, 2Nn] with C = j k eta 2 pi (bor_solver.py:293-345)."""
        return self._systems("efie", m_max)[int(m)].copy()

    def assemble_mfie_mode(self, m: int, m_max: int) -> np.ndarray:
        """Z_MFIE = (1/2) M - K [2Nn, 2Nn] (bor_solver.py:399-435)."""
        return self._systems("mfie", m_max)[int(m)].copy()

    def assemble_ibc_extra(self, m: int, m_max: int, zs_pt, zs_elem) -> np.ndarray:
        """IBC-EFIE additions (1/2) <W, Z_s J> + <W, PV curl int G M>,
        M = -Z_s n x J, Z_s at the source [2Nn, 2Nn] (bor_solver.py:534-552).
        zs_pt (per Gauss point) weights the mass and the far brackets, zs_elem
        the near brackets; both come from one per-element impedance here."""
        ze = np.asarray(zs_elem, dtype=np.complex128).reshape(self.Nn - 1)
        if not np.array_equal(self._elem_zs(zs_pt), ze):
            raise ValueError("zs_pt must repeat zs_elem over each element's Gauss points.")
        return self._systems("ibc", m_max, ze)[int(m)].copy()

    def assemble_pmchwt_P(self, m: int, m_max: int) -> np.ndarray:
        """The PMCHWT coupling operator of this medium [2Nn, 2Nn]: the
        rotated-PV brackets B (unit weight, no mass term) with their columns
        remapped, P = [[-Btf, Btt], [-Bff, Bft]] (bor_solver.py:554-575)."""
        B = self._systems("pv", m_max)[int(m)]
        n = self.Nn
        P = np.empty_like(B)
        P[:, :n] = -B[:, n:]
        P[:, n:] = B[:, :n]
        return P

    def mass_blocks(self, weight=None) -> np.ndarray:
        """2 pi int w rho T_i T_j dt, node-based [Nn, Nn] (bor_solver.py:389-397):
        tridiagonal host arithmetic on the order-4 Gauss points."""
        x, w = np.polynomial.legendre.leggauss(4)
        s = 0.5 * (x + 1.0)
        p0, p1 = self.pts[:-1], self.pts[1:]
        L = np.hypot(*(p1 - p0).T)
        rho = p0[:, 0, None] + s[None, :] * (p1[:, 0] - p0[:, 0])[:, None]
        wt = np.ones(self.P) if weight is None else np.asarray(weight).reshape(self.P)
        K = (0.5 * w[None, :] * L[:, None] * rho).ravel() * wt
        T = np.stack([np.tile(1.0 - s, self.Nn - 1), np.tile(s, self.Nn - 1)])
        e = np.repeat(np.arange(self.Nn - 1), 4)
        M = np.zeros((self.Nn, self.Nn), dtype=np.result_type(K, float))
        for a in range(2):
            for b in range(2):
                np.add.at(M, (e + a, e + b), 2.0 * np.pi * K * T[a] * T[b])
        return M

    def basis_mask(self, m: int) -> np.ndarray:
        """Active bases of mode m (bor_solver.py:611-622)."""
        tol = AXIS_TOL * max(1.0, float(np.max(self.pts[:, 0])))
        t_act = np.ones(self.Nn, dtype=bool)
        f_act = np.ones(self.Nn, dtype=bool)
        for end in (0, self.Nn - 1):
            if self.pts[end, 0] <= tol:
                t_act[end] = abs(m) == 1
                f_act[end] = False
            else:
                t_act[end] = False
                f_act[end] = True
        return np.concatenate([t_act, f_act])

    def _free_space(self, what: str) -> None:
        if self.medium is not None:
            raise ValueError(f"{what} is defined for the exterior (free-space) solver only.")

    def _rhs_all(self, kind: str, m: int, theta_inc_deg: float) -> np.ndarray:
        self._free_space("the plane-wave excitation")
        key = (kind, float(theta_inc_deg))
        cache = self._rhs.get(key)
        if cache is None or abs(m) > cache[0]:
            mm = max(abs(int(m)), 1, cache[0] if cache else 0)
            cE, cM = (1.0, 0.0) if kind == "efie" else (0.0, 1.0)
            h = _lib.Handle(_lib.GHOST_FP64, residual=False)
            try:
                h.add_problem(self.pts, self.k, mm, 64, 128, cE, cM, [float(theta_inc_deg)], unmasked=True)
                h.build()
                h.fill_rhs()
                h.sync()
                cache = (mm, {q: h.system(0, q)[1] for q in range(-mm, mm + 1)})
            finally:
                h.close()
            self._rhs[key] = cache
        return cache[1][int(m)]

    def rhs_mode(self, m: int, theta_inc_deg: float, pol: str) -> np.ndarray:
        """<W, E_inc> for the plane wave from theta (bor_solver.py:625-645), [2Nn]."""
        col = 0 if pol.upper() in ("VV", "THETA", "TM") else 1
        return self._rhs_all("efie", m, theta_inc_deg)[:, col].copy()

    def rhs_mfie_mode(self, m: int, theta_inc_deg: float, pol: str) -> np.ndarray:
        """<W, n x H_inc> (bor_solver.py:437-461), [2Nn]."""
        col = 0 if pol.upper() in ("VV", "THETA", "TM") else 1
        return self._rhs_all("mfie", m, theta_inc_deg)[:, col].copy()

    def rhs_h_mode(self, m: int, theta_inc_deg: float, pol: str) -> np.ndarray:
        """Unrotated <W, H_inc> (bor_solver.py:647-657)."""
        if pol.upper() in ("VV", "THETA", "TM"):
            return -self.rhs_mode(m, theta_inc_deg, "HH") / ETA0
        return self.rhs_mode(m, theta_inc_deg, "VV") / ETA0

    def farfield_mode(self, m: int, sol: np.ndarray, theta_s_deg: float, zs_pt=None, msol=None):
        """(F_theta, F_phi) of one mode's full [2Nn] solution (bor_solver.py:659-708);
        zs_pt adds the radiation of M = -Z_s n x J."""
        self._free_space("the far field")
        sol = np.asarray(sol, dtype=np.complex128).reshape(2 * self.Nn)
        if msol is not None:
            if zs_pt is not None:
                raise ValueError("give zs_pt or msol, not both.")
            # a penetrable pair's rows: J, then M / eta0
            sol = np.concatenate([sol, np.asarray(msol, dtype=np.complex128).reshape(2 * self.Nn) / ETA0])
        mm = max(abs(int(m)), 1)
        h = _lib.Handle(_lib.GHOST_FP64, residual=False, pairing=False)   # one mode's own currents
        try:
            h.add_problem(self.pts, self.k, mm, 64, 128, 1.0, 0.0, [float(theta_s_deg)], unmasked=True)
            if zs_pt is not None:
                h.set_impedance(0, self._elem_zs(zs_pt), 0.0)   # far field only
            if msol is not None:
                h.set_interior(0, self.k, ETA0, 64, 128)   # only the pair's row layout is used
            h.build()
            h.fill_rhs()   # the far field reads the Bessel table of the aspect
            h.set_solution(0, int(m), np.stack([sol, sol], axis=1))
            h.farfield()
            h.sync()
            dF = h.dF(0)[int(m) + mm]
        finally:
            h.close()
        return complex(dF[0, 0]), complex(dF[1, 0])
