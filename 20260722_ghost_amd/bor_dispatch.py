"""GPU drop-in for bor_dispatch.solve_monostatic_rcs_bor (conductor bodies).

Same signature, validation, meshing and result format as the reference
(bor_dispatch.py:335-702) for geometry snapshots whose segments are all
TYPE 2 (PEC) without IBC flags.  Every frequency of the request becomes one
problem of a single GPU batch (bor_solver.solve_bor_batch), instead of the
reference's frequency-by-frequency loop (bor_dispatch.py:504-539).

Penetrable layouts (TYPE 3/4/5: PMCHWT) and impedance flags (IBC-EFIE) are
outside this build and raise NotImplementedError.
"""

from __future__ import annotations

import cmath
import math
import os
import threading
from typing import Any, Callable, Dict, List, Optional, Tuple

import numpy as np

from . import bor_solver

C0 = bor_solver.C0
DEFAULT_ELEMENTS_PER_WAVELENGTH = 20
MAX_ELEMENTS_DEFAULT = 50_000


def _unit_scale(units: str) -> float:
    """bor_dispatch.py:48-54."""
    v = str(units or "").strip().lower()
    if v in {"inch", "inches", "in"}:
        return 0.0254
    if v in {"meter", "meters", "m"}:
        return 1.0
    raise ValueError(f"Unsupported geometry units '{units}'. Use inches or meters.")


def _polarization(label: Optional[str]) -> str:
    """VV (theta-pol) / HH (phi-pol) with the 2-D aliases TE -> VV, TM -> HH (bor_dispatch.py:57-68)."""
    t = str(label or "").strip().upper()
    if t in {"VV", "V", "TE", "VERTICAL"} or not t:
        return "VV"
    if t in {"HH", "H", "TM", "HORIZONTAL"}:
        return "HH"
    raise ValueError(f"Unsupported polarization '{label}'. Use VV or HH.")


def _flag(tok: Any, default: int = 0) -> int:
    try:
        t = str(tok).strip().lower()
        if t.startswith("mat."):
            t = t[4:]
        return int(float(t)) if t else default
    except (TypeError, ValueError):
        return default


def _int(tok: Any, default: int = 0) -> int:
    try:
        t = str(tok).strip()
        return int(float(t)) if t else default
    except (TypeError, ValueError):
        return default


class Chain:
    """One drawn segment as a (rho, z) polyline (bor_dispatch.py:91-106)."""

    def __init__(self, name, seg_type, n_prop, ibc_flag, pos_mat, neg_mat, pts):
        self.name, self.seg_type, self.n_prop = name, seg_type, n_prop
        self.ibc_flag, self.pos_mat, self.neg_mat = ibc_flag, pos_mat, neg_mat
        self.pts = pts
        d = np.diff(pts, axis=0)
        self.prim_lengths = np.hypot(d[:, 0], d[:, 1])
        self.length = float(np.sum(self.prim_lengths))


def chains_from_snapshot(snapshot: Dict[str, Any], scale: float) -> List[Chain]:
    """Segments -> chains; primitives inside a segment must chain head-to-tail
    (bor_dispatch.py:109-140)."""
    out: List[Chain] = []
    for si, seg in enumerate(snapshot.get("segments", []) or []):
        props = list(seg.get("properties", []) or [])
        first = props[0] if props and str(props[0]).strip() else seg.get("seg_type", 2)
        seg_type = _flag(first, 2)
        n_prop = _int(props[1] if len(props) > 1 else 0)
        ibc = _flag(props[2] if len(props) > 2 else 0)
        pos = _flag(props[3] if len(props) > 3 else 0)
        neg = _flag(props[4] if len(props) > 4 else 0)
        pts: List[Tuple[float, float]] = []
        for i, pp in enumerate(list(seg.get("point_pairs", []) or [])):
            x1, y1 = float(pp.get("x1", 0.0)) * scale, float(pp.get("y1", 0.0)) * scale
            x2, y2 = float(pp.get("x2", 0.0)) * scale, float(pp.get("y2", 0.0)) * scale
            if i == 0:
                pts.append((x1, y1))
            elif math.hypot(x1 - pts[-1][0], y1 - pts[-1][1]) > 0:
                raise ValueError(f"Segment '{seg.get('name', si)}': primitives do not chain "
                                 f"head-to-tail at ({x1 / scale:.6g}, {y1 / scale:.6g}).")
            pts.append((x2, y2))
        if len(pts) >= 2:
            out.append(Chain(str(seg.get("name", f"segment_{si + 1}")), seg_type, n_prop, ibc, pos, neg,
                             np.asarray(pts, dtype=float)))
    if not out:
        raise ValueError("Geometry contains no usable segments.")
    return out


def classify(chains: List[Chain]) -> str:
    """Material layout (bor_dispatch.py:254-298); this build solves 'conductor'."""
    types = {c.seg_type for c in chains}
    if types == {2}:
        return "conductor"
    if types <= {2, 3, 4, 5}:
        raise NotImplementedError(
            "Penetrable BoR layouts (TYPE 3/4/5, PMCHWT) are not part of the GPU BoR path.")
    raise ValueError(f"Segment TYPE(s) {sorted(types - {2, 3, 4, 5})} are not supported by the BoR "
                     "solver (supported: TYPE 2 PEC/IBC, TYPE 3 dielectric, TYPE 3+4 coated, "
                     "TYPE 2+3+4 partially coated, TYPE 3+5+4 layered).")


def stitch(chains: List[Chain], what: str, tol: float) -> List[Chain]:
    """Order chains head-to-tail as drawn into one run (bor_dispatch.py:143-185)."""
    if len(chains) == 1:
        return chains
    key = lambda p: (int(round(p[0] / tol)), int(round(p[1] / tol)))
    starts, ends = {}, {}
    for c in chains:
        ks, ke = key(c.pts[0]), key(c.pts[-1])
        if ks in starts or ke in ends:
            raise ValueError(f"The {what} segments do not form a single chain "
                             "(two segments start or end at the same point).")
        starts[ks] = c
        ends[ke] = c
    heads = [c for c in chains if key(c.pts[0]) not in ends]
    if len(heads) != 1:
        raise ValueError(f"The {what} segments do not chain head-to-tail into one generatrix. "
                         "Check that consecutive segments share endpoints and that each segment is "
                         "drawn in the same traversal direction.")
    run = [heads[0]]
    while True:
        nxt = starts.get(key(run[-1].pts[-1]))
        if nxt is None:
            break
        if nxt is run[0]:
            raise ValueError(f"The {what} segments form a closed loop in the (rho, z) plane; a BoR "
                             "generatrix must be an open polyline with both endpoints on the axis.")
        run.append(nxt)
    if len(run) != len(chains):
        raise ValueError(f"The {what} segments split into multiple disconnected chains; "
                         "expected one generatrix.")
    return run


def preflight(run: List[Chain], what: str, tol: float) -> None:
    """Axis endpoints and +z -> -z traversal (bor_dispatch.py:188-207)."""
    pts = np.vstack([run[0].pts] + [c.pts[1:] for c in run[1:]])
    rho, z = pts[:, 0], pts[:, 1]
    if np.any(rho < -tol):
        bad = pts[np.argmin(rho)]
        raise ValueError(f"The {what} generatrix crosses the rotation axis (rho = x = {bad[0]:.6g} < 0). "
                         "Draw the half-profile entirely at x >= 0.")
    if rho[0] > tol or rho[-1] > tol:
        raise ValueError(f"The {what} generatrix endpoints must lie ON the rotation axis (x = 0) to "
                         f"close the body of revolution; got start rho = {rho[0]:.6g}, end rho = "
                         f"{rho[-1]:.6g}. Open BoR shells are not supported in phase 4.")
    if z[0] <= z[-1]:
        raise ValueError(f"The {what} generatrix must be traversed from the +z (top) axis end to the "
                         f"-z (bottom) axis end so the left-of-travel normal faces the exterior; it is "
                         f"drawn bottom-to-top (z {z[0]:.6g} -> {z[-1]:.6g}). Reverse the segment "
                         "endpoint order.")


def element_count(n_prop: int, prim_len: float, lam: float) -> int:
    """bor_dispatch.py:210-217."""
    if prim_len <= 0.0:
        return 1
    if n_prop > 0:
        return max(1, n_prop)
    n_wave = abs(n_prop) if n_prop < 0 else DEFAULT_ELEMENTS_PER_WAVELENGTH
    target = max(lam / max(1, n_wave), prim_len / 2000.0)
    return max(1, int(math.ceil(prim_len / target)))


def mesh(run: List[Chain], lam: float, max_elements: int, axis_tol: float) -> np.ndarray:
    """Subdivide the ordered chains; snap near-axis nodes onto the axis (bor_dispatch.py:220-251)."""
    points: List[np.ndarray] = []
    for c in run:
        for pi in range(len(c.pts) - 1):
            p0, p1 = c.pts[pi], c.pts[pi + 1]
            cnt = element_count(c.n_prop, c.prim_lengths[pi], lam)
            for i in range(cnt):
                if not points:
                    points.append(p0 + (p1 - p0) * (i / cnt))
                points.append(p0 + (p1 - p0) * ((i + 1) / cnt))
    pts = np.asarray(points, dtype=float)
    pts[np.abs(pts[:, 0]) <= axis_tol, 0] = 0.0
    pts[:, 0] = np.maximum(pts[:, 0], 0.0)
    if len(pts) - 1 > max_elements:
        raise ValueError(f"BoR mesh would need {len(pts) - 1} elements (> max {max_elements}). "
                         "Reduce frequency or density.")
    return pts


def solve_monostatic_rcs_bor(
    geometry_snapshot: Dict[str, Any],
    frequencies_ghz: List[float],
    elevations_deg: List[float],
    polarization: str,
    geometry_units: str = "inches",
    material_base_dir: Optional[str] = None,
    progress_callback: Optional[Callable[[int, int, str], None]] = None,
    mesh_reference_ghz: Optional[float] = None,
    cfie_alpha: float = 0.5,
    n_modes: Optional[int] = None,
    mode_tol: float = 1e-6,
    max_elements: int = MAX_ELEMENTS_DEFAULT,
    workers: Optional[int] = None,
    abort_event: Optional[threading.Event] = None,
    table_precision: str = "auto",
    assembly: str = "auto",
    expand_to_360: bool = False,
    *,
    precision: str = "fp64",
    refine_iters: int = 2,
) -> Dict[str, Any]:
    """Monostatic 3-D RCS (m^2 / dBsm) of a PEC body of revolution given as a
    .geo snapshot; ``elevations_deg`` are aspects from +z (bor_dispatch.py:335-702)."""
    if not frequencies_ghz:
        raise ValueError("At least one frequency is required.")
    if not elevations_deg:
        raise ValueError("At least one aspect angle is required.")
    freqs = [float(f) for f in frequencies_ghz]
    if any(f <= 0.0 for f in freqs):
        raise ValueError("Frequencies must be positive GHz values.")
    aspects = [float(a) for a in elevations_deg]
    if any(a < 0.0 or a > 180.0 for a in aspects):
        raise ValueError("Aspect angles must lie in [0, 180] degrees from +z.")
    pol = _polarization(polarization)
    scale = _unit_scale(geometry_units)
    if workers is None:
        workers = max(1, (os.cpu_count() or 2) - 1)
    chains = chains_from_snapshot(geometry_snapshot, scale)
    kind = classify(chains)
    if any(c.ibc_flag > 0 for c in chains):
        raise NotImplementedError("Impedance (IBC) flags are not part of the GPU BoR path.")
    diag = max(float(np.ptp(np.vstack([c.pts for c in chains]), axis=0).max()), 1e-9)
    tol = max(1e-12, 1e-9 * diag)
    axis_tol = 1e-6 * diag
    run = stitch(chains, "TYPE 2", tol)
    preflight(run, "body", axis_tol)

    def check_abort():
        if abort_event is not None and abort_event.is_set():
            raise InterruptedError("Solve cancelled by user.")

    check_abort()
    form = "cfie" if cfie_alpha > 0.0 else "efie"
    jobs = []
    for fg in freqs:
        mesh_ghz = float(mesh_reference_ghz) if mesh_reference_ghz else fg
        pts = mesh(run, C0 / (mesh_ghz * 1e9), max_elements, axis_tol)
        jobs.append({"points": pts, "freq_hz": fg * 1e9, "thetas_deg": aspects, "formulation": form,
                     "cfie_alpha": cfie_alpha, "n_modes": n_modes, "assembly": assembly})
    outs = bor_solver.solve_bor_batch(jobs, precision=precision, refine_iters=refine_iters,
                                      mode_tol=mode_tol, check_abort=check_abort)
    samples: List[Dict[str, Any]] = []
    per_freq: List[Dict[str, Any]] = []
    for fi, (fg, out) in enumerate(zip(freqs, outs)):
        sig = out["sigma_vv"] if pol == "VV" else out["sigma_hh"]
        amp = out["amp_vv"] if pol == "VV" else out["amp_hh"]
        residual = float(out.get("linear_residual", 0.0))
        for ai, aspect in enumerate(aspects):
            lin = max(float(sig[ai]), 1e-30)
            a = complex(amp[ai])
            samples.append({
                "frequency_ghz": fg, "theta_inc_deg": aspect, "theta_scat_deg": aspect,
                "rcs_linear": lin, "rcs_db": 10.0 * math.log10(lin),
                "rcs_amp_real": float(a.real), "rcs_amp_imag": float(a.imag),
                "rcs_amp_phase_deg": float(math.degrees(cmath.phase(a))),
                "linear_residual": residual,
            })
        per_freq.append({"frequency_ghz": fg, "modes_used": int(out["modes_used"]),
                         "n_unknowns": int(out["n_unknowns"]), "linear_residual": residual,
                         "max_cond": None})
        if progress_callback is not None:
            try:
                progress_callback(fi + 1, len(freqs), f"Solved {fg:g} GHz")
            except Exception:
                pass
    if expand_to_360:
        mirrored = []
        for smp in samples:
            th = float(smp["theta_inc_deg"])
            if 0.0 < th < 180.0:
                m = dict(smp)
                m["theta_inc_deg"] = m["theta_scat_deg"] = 360.0 - th
                mirrored.append(m)
        samples = sorted(samples + mirrored, key=lambda s: (s["frequency_ghz"], s["theta_inc_deg"]))
    return {
        "solver": "bor_mom_rcs",
        "scattering_mode": "monostatic",
        "polarization": pol,
        "polarization_export": pol,
        "rcs_log_unit": "dBsm",
        "rcs_linear_quantity": "sigma_3d",
        "samples": samples,
        "metadata": {
            "formulation": f"BoR-MoM PEC {form.upper()}",
            "geometry_kind": kind,
            "expanded_to_360": bool(expand_to_360),
            "per_frequency": per_freq,
            "cfie_alpha": float(cfie_alpha),
            "workers": int(workers),
            "warnings": [],
            "precision": precision,
        },
    }
