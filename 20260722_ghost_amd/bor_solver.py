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
                   assembly: str = "auto"):
    """(m_max, n_xi_efie, n_xi_bracket, streaming) as the reference picks them
    (bor_solver.py:842-847, 870-885; bor_kernels.py:149-152, 315-319;
    bor_streaming.py:179-183)."""
    rho_max = float(np.max(pts[:, 0]))
    if n_modes is None:
        st_max = float(np.max(np.abs(np.sin(np.radians(thetas)))))
        n_modes = int(math.ceil(k * rho_max * max(st_max, 0.05))) + 12
    m_max = int(n_modes)
    asm = str(assembly).strip().lower()
    if asm not in ("auto", "tables", "streaming"):
        raise ValueError("assembly must be 'auto', 'tables', or 'streaming'.")
    streaming = asm == "streaming" or (asm == "auto" and
                                       estimate_bor_table_gb(len(pts) - 1, m_max, form) > 2.0)
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


def _coefficients(form: str, cfie_alpha: float):
    alpha = float(cfie_alpha) if form == "cfie" else (1.0 if form == "efie" else 0.0)
    return alpha, (1.0 - alpha) * ETA0


# ─────────────────────────────────────────────────────────────────────────────
# batched sweeps
# ─────────────────────────────────────────────────────────────────────────────

def _result(pts, thetas, form, m_max, ne, nbk, streaming, dF, res, mode_tol, prec, refine_iters, times,
            probe=None):
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
        "warnings": [],
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
                    residual: bool = True, nb: int = 64, mode_tol: float = 1e-6,
                    graph: bool = False, check_abort: Optional[Callable] = None) -> List[Dict]:
    """Solve several (generatrix, frequency) problems as ONE batch: every mode
    system of every problem goes through the same batched fill, LU/solve and
    far-field launches (systems of uneven size share each launch).  Each job
    is a dict with keys points, freq_hz, thetas_deg and optionally
    formulation ("cfie"), cfie_alpha (0.5), n_modes, assembly.  Returns one
    solve_bor-style dict per job.  graph=True captures the batch in a
    hipGraph and replays it."""
    prec = str(precision).strip().lower()
    if prec not in _PRECISION:
        raise ValueError("precision must be 'fp64' or 'mixed'.")
    specs = []
    for j in jobs:
        form = str(j.get("formulation", "cfie")).strip().lower()
        if form not in ("efie", "mfie", "cfie"):
            raise ValueError("formulation must be 'efie', 'mfie' or 'cfie'.")
        pts = _validate_nodes(j["points"])
        thetas = np.atleast_1d(np.asarray(j["thetas_deg"], dtype=float))
        k = 2.0 * math.pi * float(j["freq_hz"]) / C0
        m_max, ne, nbk, streaming = discretisation(pts, k, thetas, j.get("n_modes"), form,
                                                   j.get("assembly", "auto"))
        cE, cM = _coefficients(form, j.get("cfie_alpha", 0.5))
        specs.append((pts, thetas, form, k, m_max, ne, nbk, streaming, cE, cM))
    if check_abort is not None:
        check_abort()
    h = _lib.Handle(_PRECISION[prec], nb=nb, residual=residual)
    try:
        for (pts, thetas, form, k, m_max, ne, nbk, streaming, cE, cM) in specs:
            h.add_problem(pts, k, m_max, ne, nbk, cE, cM, thetas)
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
    for i, (pts, thetas, form, k, m_max, ne, nbk, streaming, cE, cM) in enumerate(specs):
        nm = 2 * m_max + 1
        r = res[off:off + nm] if res is not None else None
        off += nm
        outs.append(_result(pts, thetas, form, m_max, ne, nbk, streaming, dFs[i], r, mode_tol, prec,
                            refine_iters, times, probe))
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
              refine_iters: int = -3, residual: bool = True, nb: int = 64) -> Dict:
    """Monostatic RCS of a PEC body of revolution at aspects thetas_deg (from +z).

    Mirrors bor_solver.solve_bor (bor_solver.py:820-1004) for zs = None.
    Surface impedance (IBC-EFIE) is not part of this build and raises
    NotImplementedError.  ``workers``, ``table_precision`` and
    ``stream_budget_gb`` are accepted for signature compatibility: the GPU
    path holds every table in FP64 (or FP32 far blocks under
    precision="mixed") and needs no host threads or re-sweeps.
    """
    form = formulation.strip().lower()
    if form not in ("efie", "mfie", "cfie"):
        raise ValueError("formulation must be 'efie', 'mfie' or 'cfie'.")
    if zs is not None and np.any(np.abs(np.asarray(zs, dtype=complex)) != 0.0):
        raise NotImplementedError("Surface impedance (IBC-EFIE) is not part of the GPU BoR path.")
    if int(gauss_order) != 4:
        raise ValueError("the GPU BoR path is built for gauss_order = 4 (the reference default).")
    tp = str(table_precision).strip().lower()
    if tp not in ("auto", "single", "double"):
        raise ValueError("table_precision must be 'auto', 'single', or 'double'.")
    out = solve_bor_batch([{"points": points, "freq_hz": freq_hz, "thetas_deg": thetas_deg,
                            "formulation": form, "cfie_alpha": cfie_alpha, "n_modes": n_modes,
                            "assembly": assembly}],
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
