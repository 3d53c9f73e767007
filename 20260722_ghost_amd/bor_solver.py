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
from typing import Callable, Dict, List, Optional, Sequence, Tuple

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
# multi-region bodies (bor_solver.py:1100-1297, 1805-2172)
# ─────────────────────────────────────────────────────────────────────────────

def _medium(freq_hz: float, medium):
    """(k, eta) of free space (None) or of (eps_r, mu_r)."""
    k0 = 2.0 * math.pi * freq_hz / C0
    if medium is None:
        return complex(k0), complex(ETA0)
    mi, er = causal_medium(*medium)
    return k0 * mi, ETA0 * er


class MultiRegionBor:
    """Surfaces (generatrix, is_conductor) bounding homogeneous regions
    ({"medium": None | (eps_r, mu_r), "bounds": [(surface, sigma), ...],
    "exterior": bool}, sigma = +1 where the surface's normal points into the
    region), solved per mode on PMCHWT (bor_solver.py:1805-2041).

    The host keeps the layout: per surface its J unknowns (2 Nn) and, for an
    interface, M / eta0 (2 Nn); the sign sigma_i sigma_j of each block; the
    junctions (coinciding off-axis generatrix ends) and the reduction that
    ties a follower's ends to the junction's lead and drops the bases the
    axis or a conductor line removes.  The device fills every surface's
    operators in each region it bounds, the cross-surface operators between
    the surfaces of a region, sums the blocks, applies the reduction,
    factors, solves and radiates the exterior surfaces' currents."""

    def __init__(self, surfaces, regions, freq_hz: float, gauss_order: int = 4, near_factor: float = 2.0,
                 near_order: int = 12):
        if int(gauss_order) != 4:
            raise ValueError("the GPU BoR path is built for gauss_order = 4 (the reference default).")
        self.freq_hz = float(freq_hz)
        self.pts = [_validate_nodes(p) for (p, _) in surfaces]
        self.cond = [bool(c) for (_, c) in surfaces]
        self.regions = regions
        self.near_factor, self.near_order = float(near_factor), int(near_order)
        ns = len(surfaces)
        self.adj: List[List[int]] = [[] for _ in range(ns)]
        self.sigma: Dict = {}
        for ri, reg in enumerate(regions):
            for (si, sg) in reg["bounds"]:
                self.adj[si].append(ri)
                self.sigma[(ri, si)] = int(sg)
        ext = [ri for ri, r in enumerate(regions) if r.get("exterior")]
        if len(ext) != 1:
            raise ValueError("exactly one region must be the exterior.")
        self.ext = ext[0]
        if regions[self.ext].get("medium") is not None:
            raise ValueError("the exterior region is free space.")
        if any(sg != 1 for (_, sg) in regions[self.ext]["bounds"]):
            raise ValueError("surfaces bounding the exterior need sigma = +1 there.")
        for si in range(ns):
            if not self.adj[si]:
                raise ValueError(f"surface {si} bounds no region.")
        self.media = [_medium(self.freq_hz, r.get("medium")) for r in regions]
        self.Nn = [len(p) for p in self.pts]
        self.off_J, self.off_M, acc = [], [], 0
        for si in range(ns):
            self.off_J.append(acc)
            acc += 2 * self.Nn[si]
            self.off_M.append(None if self.cond[si] else acc)
            acc += 0 if self.cond[si] else 2 * self.Nn[si]
        self.n_full = acc
        self.junctions = self._junctions()

    def _on_axis(self, si: int, node: int) -> bool:
        p = self.pts[si]
        return p[node, 0] <= AXIS_TOL * max(1.0, float(np.max(p[:, 0])))

    def _junctions(self):
        """Coinciding off-axis ends, grouped; the first interface of a group
        leads it (bor_solver.py:1856-1895)."""
        allp = np.vstack(self.pts)
        tol = 1e-8 * max(float(np.ptp(allp[:, 0])) + float(np.ptp(allp[:, 1])), 1e-9)
        groups: List[List[Tuple[int, int]]] = []
        for si, p in enumerate(self.pts):
            for node in (0, len(p) - 1):
                if self._on_axis(si, node):
                    continue
                for g in groups:
                    s0, n0 = g[0]
                    if float(np.hypot(*(self.pts[s0][n0] - p[node]))) <= tol:
                        g.append((si, node))
                        break
                else:
                    groups.append([(si, node)])
        for g in groups:
            if len(g) < 2:
                raise ValueError(f"surface {g[0][0]} ends off the axis at a point no other surface reaches.")
            lead = next((i for i, (si, _) in enumerate(g) if not self.cond[si]), 0)
            g[0], g[lead] = g[lead], g[0]
            for (si, _) in g[1:]:
                if not set(self.adj[si]) & set(self.adj[g[0][0]]):
                    raise ValueError("a junction surface shares no region with the junction's lead.")
        return groups

    def reduction(self, m: int):
        """(red_of_full, coef, n_red): the reduced unknown and coefficient of
        each full unknown (-1: dropped).  Axis ends keep J_t only for |m| = 1
        and never J_phi; a follower's end takes its lead's unknowns with the
        signs of the shared region's sigma and of the traversal direction
        (t: -sigma_l dir_l / (sigma_f dir_f), phi: sigma_l / sigma_f); where a
        conductor meets the junction, M_t of the interfaces there is 0
        (bor_solver.py:1908-1986)."""
        members = {x for g in self.junctions for x in g}
        follower = {x for g in self.junctions for x in g[1:]}
        mt_off = {x for g in self.junctions if any(self.cond[s] for (s, _) in g)
                  for x in g if not self.cond[x[0]]}
        red = np.full(self.n_full, -1, dtype=np.int64)
        coef = np.zeros(self.n_full)
        n = 0
        for si in range(len(self.pts)):
            Nn = self.Nn[si]
            for off, is_m in ((self.off_J[si], False), (self.off_M[si], True)):
                if off is None:
                    continue
                t_on, f_on = np.ones(Nn, dtype=bool), np.ones(Nn, dtype=bool)
                for end in (0, Nn - 1):
                    if self._on_axis(si, end):
                        t_on[end], f_on[end] = abs(m) == 1, False
                    elif (si, end) in follower:
                        t_on[end] = f_on[end] = False
                    elif (si, end) in members:
                        if is_m and (si, end) in mt_off:
                            t_on[end] = False
                    else:
                        t_on[end] = False
                for base, on in ((off, t_on), (off + Nn, f_on)):
                    idx = base + np.flatnonzero(on)
                    red[idx] = np.arange(n, n + idx.size)
                    coef[idx] = 1.0
                    n += idx.size
        for g in self.junctions:
            sl, nl = g[0]
            for (sf, nf) in g[1:]:
                r = sorted(set(self.adj[sf]) & set(self.adj[sl]))[0]
                dl, df = (1 if nl != 0 else -1), (1 if nf != 0 else -1)
                sgl, sgf = self.sigma[(r, sl)], self.sigma[(r, sf)]
                ties = [(self.off_J[sl], self.off_J[sf])]
                if self.off_M[sl] is not None and self.off_M[sf] is not None:
                    ties.append((self.off_M[sl], self.off_M[sf]))
                for ol, of in ties:
                    for half, c in ((0, -(sgl * dl) / (sgf * df)), (1, sgl / sgf)):
                        lead = red[ol + half * self.Nn[sl] + nl]
                        if lead >= 0:
                            red[of + half * self.Nn[sf] + nf] = lead
                            coef[of + half * self.Nn[sf] + nf] = c
        return red, coef, n

    def rho_max(self) -> float:
        return max(float(np.max(p[:, 0])) for p in self.pts)

    def build(self, h, thetas, m_max: int):
        """Add this body's operators, cross operators and region set to
        handle h; returns the set index."""
        ops = {}
        for si, p in enumerate(self.pts):
            for ri in self.adj[si]:
                k, eta = self.media[ri]
                osc = 2.0 * abs(k) * gauss_rho_max(p)
                ops[(si, ri)] = h.add_operator(p, k, eta, m_max, n_xi_efie(osc, m_max), n_xi_bracket(osc, m_max),
                                               not self.cond[si], thetas if ri == self.ext else ())
        cross = {}
        for ri, reg in enumerate(self.regions):
            ids = [si for (si, _) in reg["bounds"]]
            for si in ids:
                for sj in ids:
                    if si != sj:
                        cross[(ri, si, sj)] = h.add_cross(ops[(si, ri)], ops[(sj, ri)], self.near_factor,
                                                          self.near_order)
        rs = h.add_region(m_max, self.n_full, thetas)
        for ri, reg in enumerate(self.regions):
            r2 = (ETA0 / self.media[ri][1]) ** 2
            for (si, sgi) in reg["bounds"]:
                for (sj, sgj) in reg["bounds"]:
                    ss = float(sgi * sgj)
                    src, idx = (0, ops[(si, ri)]) if si == sj else (1, cross[(ri, si, sj)])
                    h.region_block(rs, src, idx, 0, ss, self.off_J[si], self.off_J[sj])
                    if self.off_M[sj] is not None:
                        h.region_block(rs, src, idx, 1, -ETA0 * ss, self.off_J[si], self.off_M[sj])
                    if self.off_M[si] is not None:
                        h.region_block(rs, src, idx, 1, ETA0 * ss, self.off_M[si], self.off_J[sj])
                        if self.off_M[sj] is not None:
                            h.region_block(rs, src, idx, 0, r2 * ss, self.off_M[si], self.off_M[sj])
        for m in range(-m_max, m_max + 1):
            red, coef, _ = self.reduction(m)
            h.region_map(rs, m, red, coef)
        for (si, _) in self.regions[self.ext]["bounds"]:
            h.region_exterior(rs, ops[(si, self.ext)], self.off_J[si],
                              -1 if self.off_M[si] is None else self.off_M[si])
        return rs

    def solve(self, thetas_deg, n_modes: Optional[int] = None, mode_tol: float = 1e-6, *, residual: bool = True,
              nb: Optional[int] = None, tune: bool = True, pairing: bool = True,
              check_abort: Optional[Callable] = None) -> Dict:
        """Monostatic far field over thetas_deg (bor_solver.py:2044-2071)."""
        thetas = np.atleast_1d(np.asarray(thetas_deg, dtype=float))
        k0 = 2.0 * math.pi * self.freq_hz / C0
        if n_modes is None:
            st = float(np.max(np.abs(np.sin(np.radians(thetas)))))
            n_modes = int(math.ceil(k0 * self.rho_max() * max(st, 0.05))) + 12
        m_max = int(n_modes)
        variant = -1
        if nb is None:
            from . import autotune
            n_max = max(self.reduction(m)[2] for m in (0, 1))
            nb, variant = autotune.best(_lib.GHOST_FP64, n_max, 2 * m_max + 1, 2 * len(thetas), tune=tune)
        if check_abort is not None:
            check_abort()
        h = _lib.Handle(_lib.GHOST_FP64, nb=nb, residual=residual, pairing=pairing)
        try:
            h.set_gemm_variant(variant)
            rs = self.build(h, thetas, m_max)
            h.build()
            h.run(0)
            dF = h.region_dF(rs)
            res = h.residuals() if residual else None
            times = h.timings()
        finally:
            h.close()
        if check_abort is not None:
            check_abort()
        F, used = truncate_modes(dF, m_max, mode_tol)
        out = {"theta_deg": thetas.tolist(), "sigma_vv": (4.0 * math.pi * np.abs(F[0]) ** 2).tolist(),
               "sigma_hh": (4.0 * math.pi * np.abs(F[1]) ** 2).tolist(), "amp_vv": F[0].tolist(),
               "amp_hh": F[1].tolist(), "modes_used": used, "n_unknowns": int(self.n_full),
               "n_junctions": len(self.junctions), "m_max": m_max, "timings_ms": times,
               "lu_tiles": (int(nb), int(variant))}
        if res is not None:
            ms = np.arange(-m_max, m_max + 1)
            out["linear_residual"] = float(np.max(res[np.abs(ms) <= used]))
        else:
            out["linear_residual"] = float("nan")
        return out


def _multiregion(surfaces, regions, freq_hz, thetas_deg, n_modes, mode_tol, gauss_order, near_factor, near_order,
                 progress, check_abort, formulation: str, extra: Dict) -> Dict:
    sys_ = MultiRegionBor(surfaces, regions, freq_hz, gauss_order, near_factor, near_order)
    out = sys_.solve(thetas_deg, n_modes, mode_tol, check_abort=check_abort)
    out["formulation"] = formulation
    out.update(extra)
    if progress is not None:
        progress(out["modes_used"], out["m_max"])
    return out


def coated_layers(interface_points, points_core, eps_list, mu_list):
    """Surfaces and regions of a PEC core under N full layers: interfaces
    outermost first, media innermost first (bor_solver.py:2102-2131)."""
    N = len(interface_points)
    if len(eps_list) != N or len(mu_list) != N:
        raise ValueError("One (eps, mu) per layer, innermost first.")
    surfaces = [(p, False) for p in interface_points] + [(points_core, True)]
    regions = [{"medium": None, "bounds": [(0, +1)], "exterior": True}]
    for i in range(N):
        lay = N - i
        regions.append({"medium": (eps_list[lay - 1], mu_list[lay - 1]), "bounds": [(i, -1), (i + 1, +1)]})
    return surfaces, regions


def solve_bor_coated_pec(points_outer, points_core, freq_hz: float, thetas_deg, eps_r: complex,
                         mu_r: complex = 1.0, n_modes: Optional[int] = None, gauss_order: int = 4,
                         mode_tol: float = 1e-6, near_factor: float = 2.0, near_order: int = 12,
                         workers: int = 1, progress: Optional[Callable] = None,
                         check_abort: Optional[Callable] = None) -> Dict:
    """PEC core fully covered by one homogeneous layer (bor_solver.py:1299-1393):
    unknowns J, M / eta0 on the outer interface and J on the core; the region
    system of that layout (the reference's explicit 3 x 3 block form differs
    from it by the sign of the core rows, which leaves the currents alone)."""
    surfaces, regions = coated_layers([points_outer], points_core, [eps_r], [mu_r])
    return _multiregion(surfaces, regions, freq_hz, thetas_deg, n_modes, mode_tol, gauss_order, near_factor,
                        near_order, progress, check_abort, "pmchwt-coated",
                        {"eps_r": complex(eps_r), "mu_r": complex(mu_r)})


def solve_bor_coated2_pec(points_outer, points_mid, points_core, freq_hz: float, thetas_deg,
                          eps_inner: complex, mu_inner: complex, eps_outer: complex, mu_outer: complex,
                          n_modes: Optional[int] = None, gauss_order: int = 4, mode_tol: float = 1e-6,
                          near_factor: float = 2.0, near_order: int = 12, workers: int = 1,
                          progress: Optional[Callable] = None, check_abort: Optional[Callable] = None) -> Dict:
    """PEC core under two full layers (bor_solver.py:2074-2099)."""
    surfaces, regions = coated_layers([points_outer, points_mid], points_core, [eps_inner, eps_outer],
                                      [mu_inner, mu_outer])
    return _multiregion(surfaces, regions, freq_hz, thetas_deg, n_modes, mode_tol, gauss_order, near_factor,
                        near_order, progress, check_abort, "pmchwt-coated-2layer",
                        {"eps_inner": complex(eps_inner), "eps_outer": complex(eps_outer)})


def solve_bor_coated_n_pec(interface_points, points_core, freq_hz: float, thetas_deg, eps_list, mu_list,
                           n_modes: Optional[int] = None, gauss_order: int = 4, mode_tol: float = 1e-6,
                           near_factor: float = 2.0, near_order: int = 12, workers: int = 1,
                           progress: Optional[Callable] = None, check_abort: Optional[Callable] = None) -> Dict:
    """PEC core under N full layers; interfaces outermost first, media
    innermost first (bor_solver.py:2102-2131)."""
    surfaces, regions = coated_layers(list(interface_points), points_core, list(eps_list), list(mu_list))
    N = len(interface_points)
    return _multiregion(surfaces, regions, freq_hz, thetas_deg, n_modes, mode_tol, gauss_order, near_factor,
                        near_order, progress, check_abort, f"pmchwt-coated-{N}layer",
                        {"eps_layers": [complex(e) for e in eps_list]})


def partial_coating_layout(points_interface, points_covered, bare_pieces, eps_r, mu_r):
    """Surfaces and regions of a PEC body partly under one coating: the
    interface and the bare conductor pieces bound the exterior, the interface
    (reversed) and the covered core bound the coating; every off-axis end is
    a junction of the interface, the covered core and one bare piece
    (bor_solver.py:1402-1450)."""
    bare = list(bare_pieces)
    surfaces = [(points_interface, False), (points_covered, True)] + [(p, True) for p in bare]
    regions = [{"medium": None, "exterior": True, "bounds": [(0, +1)] + [(2 + i, +1) for i in range(len(bare))]},
               {"medium": (eps_r, mu_r), "bounds": [(0, -1), (1, +1)]}]
    return surfaces, regions


def solve_bor_partial_coating(points_interface, points_covered, bare_pieces, freq_hz: float, thetas_deg,
                              eps_r: complex, mu_r: complex = 1.0, bare_zs=None, n_modes: Optional[int] = None,
                              gauss_order: int = 4, mode_tol: float = 1e-6, near_factor: float = 2.0,
                              near_order: int = 12, workers: int = 1, progress: Optional[Callable] = None,
                              check_abort: Optional[Callable] = None) -> Dict:
    """PEC body partly covered by a homogeneous coating that ends on the
    conductor at junction circles (cap or band coatings;
    bor_solver.py:1402-1803), solved as the region system of
    ``partial_coating_layout``: J and M / eta0 on the interface, J on the
    covered core and on each bare piece, the three chain ends of a junction
    tied to the interface's, M_t = 0 there.  Bare pieces are perfect
    conductors here: a nonzero ``bare_zs`` raises NotImplementedError."""
    bare = list(bare_pieces)
    if bare_zs is not None:
        if len(bare_zs) != len(bare):
            raise ValueError("bare_zs needs one entry (None, a scalar or per-element values) per bare piece.")
        if any(z is not None and np.any(np.abs(np.asarray(z, dtype=complex)) > 0.0) for z in bare_zs):
            raise NotImplementedError("surface impedance on the bare pieces of a partial coating is not part of "
                                      "the GPU BoR path; the bare pieces are solved as perfect conductors.")
    surfaces, regions = partial_coating_layout(points_interface, points_covered, bare, eps_r, mu_r)
    out = _multiregion(surfaces, regions, freq_hz, thetas_deg, n_modes, mode_tol, gauss_order, near_factor,
                       near_order, progress, check_abort, "pmchwt-partial-coating",
                       {"eps_r": complex(eps_r), "mu_r": complex(mu_r)})
    out.setdefault("warnings", [])
    return out


def solve_bor_coating_patch(points_patch, points_mid_covered, points_mid_bare, points_core, freq_hz: float,
                            thetas_deg, eps_inner: complex, mu_inner: complex, eps_patch: complex,
                            mu_patch: complex, n_modes: Optional[int] = None, gauss_order: int = 4,
                            mode_tol: float = 1e-6, near_factor: float = 2.0, near_order: int = 12,
                            workers: int = 1, progress: Optional[Callable] = None,
                            check_abort: Optional[Callable] = None) -> Dict:
    """A second coating layer over part of a fully coated PEC body, meeting
    the inner coating's interface at dielectric triple junctions
    (bor_solver.py:2134-2176)."""
    bare = list(points_mid_bare) if isinstance(points_mid_bare, (list, tuple)) else [points_mid_bare]
    surfaces = [(points_patch, False), (points_mid_covered, False)] + [(p, False) for p in bare]
    surfaces.append((points_core, True))
    core = len(surfaces) - 1
    bare_ids = list(range(2, 2 + len(bare)))
    regions = [{"medium": None, "exterior": True, "bounds": [(0, +1)] + [(b, +1) for b in bare_ids]},
               {"medium": (eps_patch, mu_patch), "bounds": [(0, -1), (1, +1)]},
               {"medium": (eps_inner, mu_inner), "bounds": [(1, -1)] + [(b, -1) for b in bare_ids] + [(core, +1)]}]
    return _multiregion(surfaces, regions, freq_hz, thetas_deg, n_modes, mode_tol, gauss_order, near_factor,
                        near_order, progress, check_abort, "pmchwt-coating-patch",
                        {"eps_inner": complex(eps_inner), "eps_patch": complex(eps_patch)})


class BorCrossOperators:
    """T and P between the test bases of one BorPecSolver's surface and the
    source bases of another's in one medium (bor_solver.py:1115-1297),
    filled on the device for every mode |m| <= m_max at once and cached."""

    def __init__(self, sp: "BorPecSolver", sq: "BorPecSolver", near_factor: float = 2.0, near_order: int = 12):
        if not np.isclose(complex(sp.k), complex(sq.k)):
            raise ValueError("Cross operators need both solvers in the same medium.")
        self.sp, self.sq = sp, sq
        self.near_factor, self.near_order = float(near_factor), int(near_order)
        self._cache: Dict = {}

    def prepare(self, m_max: int) -> None:
        self._blocks(m_max)

    def _blocks(self, m_max: int):
        if m_max not in self._cache:
            sp, sq = self.sp, self.sq
            Np, Nq = sp.Nn, sq.Nn
            n = 2 * (Np + Nq)
            h = _lib.Handle(_lib.GHOST_FP64, residual=False)
            try:
                ops = []
                for s_ in (sp, sq):
                    osc = 2.0 * abs(s_.k) * gauss_rho_max(s_.pts)
                    ops.append(h.add_operator(s_.pts, s_.k, s_.eta, m_max, n_xi_efie(osc, m_max),
                                              n_xi_bracket(osc, m_max), True))
                x = h.add_cross(ops[0], ops[1], self.near_factor, self.near_order)
                sets = []
                for which in (0, 1):
                    rs = h.add_region(m_max, n, [0.0])
                    h.region_block(rs, 1, x, which, 1.0, 0, 2 * Np)
                    for m in range(-m_max, m_max + 1):
                        h.region_map(rs, m, np.arange(n), np.ones(n))
                    sets.append(rs)
                h.build()
                h.fill()
                h.sync()
                self._cache[m_max] = {m: tuple(h.region_system(rs, m)[0][:2 * Np, 2 * Np:] for rs in sets)
                                      for m in range(-m_max, m_max + 1)}
            finally:
                h.close()
        return self._cache[m_max]

    def assemble_T(self, m: int, m_max: int) -> np.ndarray:
        """Cross EFIE operator [2 Np, 2 Nq] with C = j k eta 2 pi of the medium."""
        return self._blocks(m_max)[int(m)][0].copy()

    def assemble_P(self, m: int, m_max: int) -> np.ndarray:
        """Cross rotated-PV operator [2 Np, 2 Nq] (columns remapped as assemble_pmchwt_P)."""
        return self._blocks(m_max)[int(m)][1].copy()


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
