"""GPU drop-in for the reference's BoR dispatch (bor_dispatch.py:335-925).

``solve_monostatic_rcs_bor`` keeps the reference's signature, validation
contract (ValueError for bad input or geometry, InterruptedError on abort)
and result format (sample rows and metadata), for conductor bodies (every
segment TYPE 2, PEC or with impedance flags: IBC-EFIE, tapered and tabulated
Z_s through ``materials.MaterialLibrary``) and homogeneous penetrable bodies
(every segment TYPE 3 with one pos_mat dielectric flag: PMCHWT, meshed at the
interior wavelength) and multi-region bodies (a PEC core fully or partly
under a coating, TYPE 3 + 4 and 2 + 3 + 4; two layers or a patch over a
coating, TYPE 3 + 4 + 5; a stack of layers, several TYPE 5 pairs: the region
PMCHWT of ``bor_solver.MultiRegionBor``).  All frequencies of a conductor or
homogeneous request go to the GPU as one batch
(``bor_solver.solve_bor_batch``); multi-region bodies are solved one
frequency at a time.

Geometry path: the snapshot's segments become polylines (``Piece``) in
metres; the conductor's pieces must join end-to-start as drawn into one run
(``order_runs``) that starts and ends on the axis and runs from +z to -z
(``check_generatrix``); the run is subdivided per primitive with the
reference's element-count rule (``element_count``) and near-axis nodes are
put on the axis (``mesh_run``).  The element counts and node coordinates
must equal the reference's bit for bit, so the count formula and the node
interpolation are kept in the reference's arithmetic form.

Post-processing of solved sweeps: ``bor_az_el_grid`` (radar-frame
polarimetric az/el grid from the two co-polarised aspect sweeps) and
``solve_adaptive_frequency_sweep_bor`` (midpoint refinement where adjacent
frequencies differ by more than a dB threshold).
"""

from __future__ import annotations

import cmath
import math
import os
import threading
from dataclasses import dataclass, field
from typing import Any, Callable, Dict, List, Optional, Sequence, Tuple

import numpy as np

from . import bor_solver
from .materials import MaterialLibrary, flag_value

C0 = bor_solver.C0
DEFAULT_ELEMENTS_PER_WAVELENGTH = 20
MAX_ELEMENTS_DEFAULT = 50_000

_METRES_PER_UNIT = {"inch": 0.0254, "inches": 0.0254, "in": 0.0254, "meter": 1.0, "meters": 1.0, "m": 1.0}
_POL_NAMES = {"VV": ("VV", "V", "TE", "VERTICAL", ""), "HH": ("HH", "H", "TM", "HORIZONTAL")}


def metres_per_unit(units: str) -> float:
    key = str(units or "").strip().lower()
    if key not in _METRES_PER_UNIT:
        raise ValueError(f"geometry units {units!r} are not known: give inches or meters.")
    return _METRES_PER_UNIT[key]


def canonical_polarization(label: Optional[str]) -> str:
    """'VV' (theta-polarised) or 'HH' (phi-polarised); the 2-D aliases TE and
    TM map to VV and HH, an empty label to VV (bor_dispatch.py:57-68)."""
    key = str(label or "").strip().upper()
    for pol, names in _POL_NAMES.items():
        if key in names:
            return pol
    raise ValueError(f"polarization {label!r} is neither VV nor HH.")


def _type_token(tok: Any) -> int:
    """Segment TYPE of a token; 2 (conductor) when it is empty or unreadable."""
    text = str(tok).strip().lower()
    text = text[4:] if text.startswith("mat.") else text
    try:
        return int(float(text)) if text else 2
    except ValueError:
        return 2


def _count_token(tok: Any) -> int:
    try:
        t = str(tok).strip()
        return int(float(t)) if t else 0
    except (TypeError, ValueError):
        return 0


@dataclass
class Piece:
    """One drawn segment as a polyline in (rho, z) metres, with its
    properties: TYPE, element density n (>0 fixed count per primitive, <0
    elements per wavelength, 0 the default), impedance flag and material flags."""

    name: str
    kind: int
    density: int
    ibc: int
    pos: int
    neg: int
    pts: np.ndarray
    lengths: np.ndarray = field(init=False)
    total: float = field(init=False)

    def __post_init__(self):
        self.lengths = np.hypot(*np.diff(self.pts, axis=0).T)
        self.total = float(self.lengths.sum())

    @property
    def head(self):
        return self.pts[0]

    @property
    def tail(self):
        return self.pts[-1]


def read_pieces(snapshot: Dict[str, Any], scale: float) -> List[Piece]:
    """The snapshot's segments as Pieces; primitives inside a segment must
    follow one another end to start.  Segments with no primitive are skipped."""
    pieces: List[Piece] = []
    for i, seg in enumerate(snapshot.get("segments") or []):
        props = list(seg.get("properties") or [])
        prop = lambda j: props[j] if len(props) > j else 0
        first = props[0] if props and str(props[0]).strip() else seg.get("seg_type", 2)
        kind = _type_token(first)
        name = str(seg.get("name", f"segment_{i + 1}"))
        pts: List[Tuple[float, float]] = []
        for pp in seg.get("point_pairs") or []:
            a = (float(pp.get("x1", 0.0)) * scale, float(pp.get("y1", 0.0)) * scale)
            b = (float(pp.get("x2", 0.0)) * scale, float(pp.get("y2", 0.0)) * scale)
            if pts and (a[0] != pts[-1][0] or a[1] != pts[-1][1]):
                raise ValueError(f"segment {name!r}: a primitive starts at ({a[0] / scale:.6g}, "
                                 f"{a[1] / scale:.6g}), away from where the previous one ended.")
            if not pts:
                pts.append(a)
            pts.append(b)
        if len(pts) >= 2:
            pieces.append(Piece(name, kind, _count_token(prop(1)), flag_value(prop(2)), flag_value(prop(3)),
                                flag_value(prop(4)), np.asarray(pts, dtype=float)))
    if not pieces:
        raise ValueError("the geometry has no segment with a primitive.")
    return pieces


def order_runs(pieces: Sequence[Piece], tol: float, what: str = "conductor") -> List[List[Piece]]:
    """Join pieces end to start, as drawn, into maximal runs.  Endpoints are
    compared on a grid of pitch tol.  Two pieces may not start (or end) at
    the same point, and every piece must be reached from a run's first piece
    (a piece whose start is no other piece's end): otherwise the drawing has
    a loop or a fork."""
    snap = lambda p: (int(round(p[0] / tol)), int(round(p[1] / tol)))
    by_head: Dict[Tuple[int, int], Piece] = {}
    tails = set()
    for pc in pieces:
        h, t = snap(pc.head), snap(pc.tail)
        if h in by_head or t in tails:
            raise ValueError(f"two {what} segments begin or finish at the same point; a generatrix "
                             "is one chain of segments, each drawn start to end.")
        by_head[h] = pc
        tails.add(t)
    runs, seen = [], set()
    for first in (pc for pc in pieces if snap(pc.head) not in tails):
        run = [first]
        seen.add(id(first))
        while (nxt := by_head.get(snap(run[-1].tail))) is not None and id(nxt) not in seen:
            run.append(nxt)
            seen.add(id(nxt))
        runs.append(run)
    if len(seen) != len(pieces):
        raise ValueError(f"some {what} segments close a loop in the (rho, z) plane or are not reached "
                         "from an open end; a generatrix is an open chain.")
    return runs


def single_run(pieces: Sequence[Piece], what: str, tol: float) -> List[Piece]:
    """The one run all pieces form (ValueError otherwise)."""
    if len(pieces) == 1:
        return list(pieces)
    runs = order_runs(pieces, tol, what)
    if len(runs) != 1:
        raise ValueError(f"the {what} segments form {len(runs)} separate chains; they must join end to "
                         "start into one generatrix (a chain that meets another start-to-start or "
                         "end-to-end has one segment drawn in the reverse direction).")
    return runs[0]


def run_points(run: Sequence[Piece]) -> np.ndarray:
    return np.vstack([run[0].pts] + [pc.pts[1:] for pc in run[1:]])


def check_generatrix(run: Sequence[Piece], what: str, axis_tol: float) -> None:
    """Closed-body conditions (bor_dispatch.py:188-207): rho >= 0, both ends
    on the axis, traversed from the +z end to the -z end."""
    pts = run_points(run)
    rho, z = pts[:, 0], pts[:, 1]
    if np.any(rho < -axis_tol):
        raise ValueError(f"the {what} generatrix reaches rho = {float(rho.min()):.6g} < 0; draw the "
                         "half-profile at x >= 0.")
    if rho[0] > axis_tol or rho[-1] > axis_tol:
        raise ValueError(f"the {what} generatrix must start and finish on the axis (x = 0) to close "
                         f"the body; its ends are at rho = {rho[0]:.6g} and {rho[-1]:.6g}.")
    if not z[0] > z[-1]:
        raise ValueError(f"the {what} generatrix must be traversed from its +z end to its -z end (so "
                         f"its left-hand normal faces outward); it runs from z = {z[0]:.6g} to "
                         f"{z[-1]:.6g}: reverse the segments.")


def element_count(density: int, length: float, wavelength: float) -> int:
    """Elements on one primitive (bor_dispatch.py:210-217): density > 0 is a
    fixed count; otherwise |density| (or 20) per wavelength, never finer than
    length / 2000."""
    if length <= 0.0:
        return 1
    if density > 0:
        return density
    per_wave = -density if density < 0 else DEFAULT_ELEMENTS_PER_WAVELENGTH
    target = max(wavelength / max(1, per_wave), length / 2000.0)
    return max(1, int(math.ceil(length / target)))


def mesh_run(run: Sequence[Piece], wavelength: float, max_elements: int, axis_tol: float):
    """Nodes [Nn, 2] of the subdivided run, the piece index of every element
    and the element midpoint's arc position along its own piece (0..1, the
    taper coordinate) (bor_dispatch.py:220-251)."""
    counts = [[element_count(pc.density, float(L), wavelength) for L in pc.lengths] for pc in run]
    total = sum(sum(c) for c in counts)
    if total > max_elements:
        raise ValueError(f"the mesh needs {total} elements, above max_elements = {max_elements}; lower the "
                         "frequency or the element density.")
    nodes = [run[0].pts[0]]
    owner, arc = [], []
    for ip, (pc, cnts) in enumerate(zip(run, counts)):
        done = 0.0
        for ia, cnt in enumerate(cnts):
            p0, p1, L = pc.pts[ia], pc.pts[ia + 1], float(pc.lengths[ia])
            for i in range(cnt):
                nodes.append(p0 + (p1 - p0) * ((i + 1) / cnt))
                owner.append(ip)
                arc.append((done + L * (i + 0.5) / cnt) / max(pc.total, 1e-300))
            done += L
    pts = np.array(nodes, dtype=float)
    pts[np.abs(pts[:, 0]) <= axis_tol, 0] = 0.0
    pts[:, 0] = np.maximum(pts[:, 0], 0.0)
    return pts, np.array(owner, dtype=int), np.array(arc, dtype=float)


MULTI_REGION = ("coated", "partial", "layered", "layered_n")


def layout(pieces: Sequence[Piece]) -> str:
    """Which body the segment TYPEs describe (bor_dispatch.py:254-298):
    'conductor' (all TYPE 2), 'dielectric' (all TYPE 3 on one pos_mat flag),
    'coated' (a TYPE 3 interface over a TYPE 4 core, one coating flag),
    'partial' (the same with bare TYPE 2 conductor pieces), 'layered' (two
    layers: TYPE 3 outer interface, one TYPE 5 interface pair, TYPE 4 core;
    TYPE 3 pieces on the inner flag are an exposed inner interface under a
    patch) or 'layered_n' (several TYPE 5 pairs: a full stack)."""
    kinds = {pc.kind for pc in pieces}
    odd = sorted(kinds - {2, 3, 4, 5})
    if odd:
        raise ValueError(f"segment TYPE {odd} cannot be solved as a body of revolution (TYPE 2 conductor "
                         "or impedance, 3 dielectric interface, 4 coated conductor, 5 layer interface).")
    flags = lambda t: {pc.pos for pc in pieces if pc.kind == t}
    if kinds == {2}:
        return "conductor"
    if kinds == {3}:
        if len(flags(3)) != 1:
            raise ValueError("a homogeneous penetrable body needs the same pos_mat flag on every TYPE 3 segment.")
        return "dielectric"
    if kinds in ({3, 4}, {2, 3, 4}):
        if len(flags(3)) != 1 or flags(3) != flags(4):
            raise ValueError("a coated conductor has one coating: every TYPE 3 interface segment and every "
                             "TYPE 4 covered-conductor segment must carry the same pos_mat flag.")
        return "partial" if 2 in kinds else "coated"
    if kinds == {3, 4, 5}:
        pairs = {(pc.pos, pc.neg) for pc in pieces if pc.kind == 5}
        if len(pairs) > 1:
            return "layered_n"
        outer, inner = next(iter(pairs))
        if outer <= 0 or inner <= 0 or outer == inner:
            raise ValueError("a TYPE 5 layer interface needs two different positive flags: pos_mat the outer "
                             "layer, neg_mat the inner one.")
        if flags(4) != {inner}:
            raise ValueError("the TYPE 4 core of a two-layer coating must carry the inner layer's flag "
                             "(the TYPE 5 interface's neg_mat).")
        if not flags(3) <= {outer, inner}:
            raise ValueError("every TYPE 3 segment of a two-layer coating must carry the outer or the inner "
                             "layer's flag.")
        return "layered"
    raise ValueError("TYPE 2 conductors combine with penetrable layers only as a partial coating "
                     "(TYPE 2 + 3 + 4).")


@dataclass
class RegionPlan:
    """The runs of a multi-region body and how each is meshed: ``runs``
    maps a role to its runs, ``flags`` the layer flags outermost first."""

    kind: str
    runs: Dict[str, List[List[Piece]]]
    flags: List[int]


def region_plan(kind: str, pieces: Sequence[Piece], tol: float, axis_tol: float) -> RegionPlan:
    """Group a multi-region body's pieces into runs by role and check that
    the closed surfaces close (bor_dispatch.py:400-487)."""
    of = lambda pred: [pc for pc in pieces if pred(pc)]
    closed = lambda run, what: (check_generatrix(run, what, axis_tol), run)[1]
    if kind in ("coated", "partial"):
        flag = pieces[[pc.kind for pc in pieces].index(3)].pos
        iface = single_run(of(lambda p: p.kind == 3), "coating interface (TYPE 3)", tol)
        covered = single_run(of(lambda p: p.kind == 4), "covered conductor (TYPE 4)", tol)
        if kind == "coated":
            return RegionPlan(kind, {"interface": [closed(iface, "coating interface")],
                                     "core": [closed(covered, "core")]}, [flag])
        bare_pcs = of(lambda p: p.kind == 2)
        bare = order_runs(bare_pcs, tol, "bare conductor (TYPE 2)")
        # the conductor as a whole (covered and bare pieces) is one closed body
        closed(single_run(of(lambda p: p.kind == 4) + bare_pcs, "conductor (TYPE 2 + 4)", tol), "conductor")
        return RegionPlan(kind, {"interface": [iface], "core": [covered], "bare": bare}, [flag])
    if kind == "layered":
        outer, inner = next(iter({(pc.pos, pc.neg) for pc in pieces if pc.kind == 5}))
        mid = single_run(of(lambda p: p.kind == 5), "layer interface (TYPE 5)", tol)
        core = closed(single_run(of(lambda p: p.kind == 4), "core (TYPE 4)", tol), "core")
        patch = single_run(of(lambda p: p.kind == 3 and p.pos == outer), "outer interface (TYPE 3)", tol)
        exposed_pcs = of(lambda p: p.kind == 3 and p.pos == inner)
        exposed = order_runs(exposed_pcs, tol, "exposed inner interface (TYPE 3)") if exposed_pcs else []
        closed(single_run(of(lambda p: p.kind == 5) + exposed_pcs, "inner layer interface (TYPE 5 + 3)", tol),
               "inner layer interface")
        if not exposed:
            closed(patch, "outer interface")
        return RegionPlan(kind, {"patch": [patch], "mid": [mid], "exposed": exposed, "core": [core]},
                          [outer, inner])
    # layered_n: walk the TYPE 5 pairs from the outer interface's flag to the core's
    top = {pc.pos for pc in pieces if pc.kind == 3}
    if len(top) != 1:
        raise ValueError("a stack of more than two layers covers the whole core: every TYPE 3 segment must "
                         "carry the outermost layer's flag.")
    bottom = {pc.pos for pc in pieces if pc.kind == 4}
    if len(bottom) != 1:
        raise ValueError("the TYPE 4 core segments must all carry one flag (the innermost layer's).")
    below: Dict[int, int] = {}
    for po, ne in {(pc.pos, pc.neg) for pc in pieces if pc.kind == 5}:
        if po in below:
            raise ValueError(f"two TYPE 5 interfaces have the same outer-layer flag {po}.")
        below[po] = ne
    order = [next(iter(top))]
    target = next(iter(bottom))
    while order[-1] != target:
        if order[-1] not in below or len(order) > len(pieces):
            raise ValueError(f"no TYPE 5 interface leads from layer flag {order[-1]} toward the core's flag "
                             f"{target}.")
        order.append(below.pop(order[-1]))
    if below:
        raise ValueError(f"TYPE 5 interfaces on layer flags {sorted(below)} are not between the outer "
                         "interface and the core.")
    ifaces = [closed(single_run(of(lambda p: p.kind == 3), "outer interface (TYPE 3)", tol), "outer interface")]
    for po, ne in zip(order[:-1], order[1:]):
        ifaces.append(closed(single_run(of(lambda p, a=po, b=ne: p.kind == 5 and (p.pos, p.neg) == (a, b)),
                                        f"layer interface {po}|{ne} (TYPE 5)", tol), f"layer interface {po}|{ne}"))
    core = closed(single_run(of(lambda p: p.kind == 4), "core (TYPE 4)", tol), "core")
    return RegionPlan(kind, {"interfaces": ifaces, "core": [core]}, order)


def solve_region_plan(plan: RegionPlan, materials: MaterialLibrary, fg: float, wavelength: float,
                      aspects: Sequence[float], n_modes, mode_tol: float, max_elements: int, axis_tol: float,
                      check_abort: Callable) -> Tuple[Dict[str, Any], str]:
    """Mesh a multi-region body at one frequency (the layers at the shortest
    interior wavelength, bare conductor pieces at the free-space one) and
    solve it on the GPU (bor_dispatch.py:565-640)."""
    media = [materials.get_medium(fl, fg) for fl in plan.flags]
    index = max([1.0] + [abs(bor_solver.causal_medium(e, u)[0]) for (e, u) in media])
    mesh = lambda run, lam=wavelength / index: mesh_run(run, lam, max_elements, axis_tol)[0]
    kw = {"n_modes": n_modes, "mode_tol": mode_tol, "check_abort": check_abort}
    f_hz = fg * 1e9
    R = plan.runs
    if plan.kind == "coated":
        (eps, mu), = media
        out = bor_solver.solve_bor_coated_pec(mesh(R["interface"][0]), mesh(R["core"][0]), f_hz, aspects, eps,
                                              mu, **kw)
        return out, "BoR-MoM PMCHWT coated PEC (multi-region)"
    if plan.kind == "partial":
        (eps, mu), = media
        bare = [mesh(run, wavelength) for run in R["bare"]]
        for run in R["bare"]:
            if any(pc.ibc > 0 for pc in run):
                raise NotImplementedError("impedance flags on the bare conductor of a partial coating are not "
                                          "part of the GPU BoR path.")
        out = bor_solver.solve_bor_partial_coating(mesh(R["interface"][0]), mesh(R["core"][0]), bare, f_hz,
                                                   aspects, eps, mu, **kw)
        return out, f"BoR-MoM PMCHWT partial coating ({out['n_junctions']} junction(s))"
    if plan.kind == "layered":
        (eps_o, mu_o), (eps_i, mu_i) = media
        patch, mid, core = mesh(R["patch"][0]), mesh(R["mid"][0]), mesh(R["core"][0])
        if R["exposed"]:
            exposed = [mesh(run) for run in R["exposed"]]
            out = bor_solver.solve_bor_coating_patch(patch, mid, exposed, core, f_hz, aspects, eps_i, mu_i,
                                                     eps_o, mu_o, **kw)
            return out, f"BoR-MoM PMCHWT coating patch ({out['n_junctions']} junction(s))"
        out = bor_solver.solve_bor_coated2_pec(patch, mid, core, f_hz, aspects, eps_i, mu_i, eps_o, mu_o, **kw)
        return out, "BoR-MoM PMCHWT two-layer coated PEC"
    ifaces = [mesh(run) for run in R["interfaces"]]
    out = bor_solver.solve_bor_coated_n_pec(ifaces, mesh(R["core"][0]), f_hz, aspects,
                                            [e for (e, _) in media[::-1]], [u for (_, u) in media[::-1]], **kw)
    return out, f"BoR-MoM PMCHWT {len(ifaces)}-layer coated PEC"


def _sample_rows(fg: float, aspects: Sequence[float], out: Dict[str, Any], pol: str) -> List[Dict[str, Any]]:
    sig = out["sigma_vv"] if pol == "VV" else out["sigma_hh"]
    amp = out["amp_vv"] if pol == "VV" else out["amp_hh"]
    res = float(out.get("linear_residual", 0.0))
    rows = []
    for th, s, a in zip(aspects, sig, amp):
        lin = max(float(s), 1e-30)
        a = complex(a)
        rows.append({"frequency_ghz": fg, "theta_inc_deg": th, "theta_scat_deg": th, "rcs_linear": lin,
                     "rcs_db": 10.0 * math.log10(lin), "rcs_amp_real": a.real, "rcs_amp_imag": a.imag,
                     "rcs_amp_phase_deg": math.degrees(cmath.phase(a)), "linear_residual": res})
    return rows


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
    refine_iters: int = -3,
) -> Dict[str, Any]:
    """Monostatic 3-D RCS (m^2, dBsm) of a conductor body of revolution from
    a .geo snapshot; ``elevations_deg`` are aspects from +z (0 nose-on, 180
    tail-on).  Both polarisations are solved; ``polarization`` picks the one
    reported.  ``expand_to_360`` adds the mirror aspects 360 - theta (exact
    for a body of revolution).  ``precision`` / ``refine_iters`` select the
    GPU path's FP64 or mixed solve (not in the reference)."""
    freqs = [float(f) for f in (frequencies_ghz or [])]
    if not freqs:
        raise ValueError("give at least one frequency.")
    if any(not f > 0.0 for f in freqs):
        raise ValueError("frequencies must be positive (GHz).")
    aspects = [float(a) for a in (elevations_deg or [])]
    if not aspects:
        raise ValueError("give at least one aspect angle.")
    if any(a < 0.0 or a > 180.0 for a in aspects):
        raise ValueError("aspect angles are measured from +z and must lie in [0, 180] degrees.")
    pol = canonical_polarization(polarization)
    scale = metres_per_unit(geometry_units)
    workers = max(1, (os.cpu_count() or 2) - 1) if workers is None else workers
    materials = MaterialLibrary.from_entries(geometry_snapshot.get("ibcs") or [],
                                             geometry_snapshot.get("dielectrics") or [],
                                             base_dir=material_base_dir or os.getcwd())
    pieces = read_pieces(geometry_snapshot, scale)
    kind = layout(pieces)
    span = float(np.ptp(np.vstack([pc.pts for pc in pieces]), axis=0).max())
    diag = max(span, 1e-9)
    tol = max(1e-12, 1e-9 * diag)
    axis_tol = 1e-6 * diag

    def check_abort():
        if abort_event is not None and abort_event.is_set():
            raise InterruptedError("Solve cancelled by user.")

    if kind in MULTI_REGION:
        plan = region_plan(kind, pieces, tol, axis_tol)
        check_abort()
        return _solve_multi_region(plan, materials, freqs, aspects, pol, mesh_reference_ghz, n_modes, mode_tol,
                                   max_elements, axis_tol, check_abort, progress_callback, cfie_alpha, workers,
                                   expand_to_360)
    run = single_run(pieces, "TYPE 2" if kind == "conductor" else "TYPE 3", tol)
    check_generatrix(run, "body", axis_tol)
    check_abort()
    jobs, forms = [], []
    # a penetrable body runs in FP64 (complex interior wavenumber)
    prec = "fp64" if kind == "dielectric" else precision
    for fg in freqs:
        wavelength = C0 / ((float(mesh_reference_ghz) if mesh_reference_ghz else fg) * 1e9)
        if kind == "dielectric":
            # meshed at the interior wavelength (bor_dispatch.py:543-547)
            eps, mu = materials.get_medium(run[0].pos, fg)
            m_idx, _ = bor_solver.causal_medium(eps, mu)
            nodes, _, _ = mesh_run(run, wavelength / max(1.0, abs(m_idx)), max_elements, axis_tol)
            forms.append(("pmchwt", False))
            jobs.append({"points": nodes, "freq_hz": fg * 1e9, "thetas_deg": aspects, "n_modes": n_modes,
                         "medium": (eps, mu)})
            continue
        nodes, owner, arc = mesh_run(run, wavelength, max_elements, axis_tol)
        zs = np.array([materials.get_impedance(run[o].ibc, fg, arc_s=float(s)) if run[o].ibc > 0 else 0.0
                       for o, s in zip(owner, arc)], dtype=np.complex128)
        has_zs = bool(np.any(zs != 0.0))
        form = "efie" if (has_zs or cfie_alpha <= 0.0) else "cfie"
        forms.append((form, has_zs))
        jobs.append({"points": nodes, "freq_hz": fg * 1e9, "thetas_deg": aspects, "formulation": form,
                     "cfie_alpha": cfie_alpha, "n_modes": n_modes, "assembly": assembly,
                     "zs": zs if has_zs else None})
    outs = bor_solver.solve_bor_batch(jobs, precision=prec, refine_iters=refine_iters,
                                      mode_tol=mode_tol, check_abort=check_abort)
    samples: List[Dict[str, Any]] = []
    per_freq: List[Dict[str, Any]] = []
    for fi, (fg, out) in enumerate(zip(freqs, outs)):
        samples += _sample_rows(fg, aspects, out, pol)
        for w in out.get("warnings") or []:
            materials.warn_once(str(w))
        per_freq.append({"frequency_ghz": fg, "modes_used": int(out["modes_used"]),
                         "n_unknowns": int(out["n_unknowns"]),
                         "linear_residual": float(out.get("linear_residual", 0.0)), "max_cond": None})
        if progress_callback is not None:
            try:
                progress_callback(fi + 1, len(freqs), f"Solved {fg:g} GHz")
            except Exception:
                pass
    any_zs = any(h for _, h in forms)
    if kind == "dielectric":
        label = "BoR-MoM PMCHWT (homogeneous dielectric)"
    else:
        label = ("BoR-MoM IBC-EFIE (Leontovich)" if any_zs
                 else f"BoR-MoM PEC {forms[0][0].upper()}")
    return _result(samples, per_freq, pol, label, kind, expand_to_360, cfie_alpha, workers, materials, prec)


def _mirror_360(samples: List[Dict[str, Any]]) -> List[Dict[str, Any]]:
    """Add the aspects 360 - theta (exact for a body of revolution)."""
    mirror = [dict(s, theta_inc_deg=360.0 - s["theta_inc_deg"], theta_scat_deg=360.0 - s["theta_inc_deg"])
              for s in samples if 0.0 < s["theta_inc_deg"] < 180.0]
    return sorted(samples + mirror, key=lambda s: (s["frequency_ghz"], s["theta_inc_deg"]))


def _solve_multi_region(plan, materials, freqs, aspects, pol, mesh_reference_ghz, n_modes, mode_tol, max_elements,
                        axis_tol, check_abort, progress_callback, cfie_alpha, workers, expand_to_360):
    """One GPU region solve per frequency (FP64: the layers' wavenumbers are complex)."""
    samples: List[Dict[str, Any]] = []
    per_freq: List[Dict[str, Any]] = []
    label = ""
    for fi, fg in enumerate(freqs):
        check_abort()
        wavelength = C0 / ((float(mesh_reference_ghz) if mesh_reference_ghz else fg) * 1e9)
        out, label = solve_region_plan(plan, materials, fg, wavelength, aspects, n_modes, mode_tol, max_elements,
                                       axis_tol, check_abort)
        for w in out.get("warnings") or []:
            materials.warn_once(str(w))
        samples += _sample_rows(fg, aspects, out, pol)
        per_freq.append({"frequency_ghz": fg, "modes_used": int(out["modes_used"]),
                         "n_unknowns": int(out["n_unknowns"]),
                         "linear_residual": float(out.get("linear_residual", 0.0)), "max_cond": None})
        if progress_callback is not None:
            try:
                progress_callback(fi + 1, len(freqs), f"Solved {fg:g} GHz")
            except Exception:
                pass
    return _result(samples, per_freq, pol, label, plan.kind, expand_to_360, cfie_alpha, workers, materials, "fp64")


def _result(samples, per_freq, pol, label, kind, expand_to_360, cfie_alpha, workers, materials, prec):
    if expand_to_360:
        samples = _mirror_360(samples)
    return {
        "solver": "bor_mom_rcs",
        "scattering_mode": "monostatic",
        "polarization": pol,
        "polarization_export": pol,
        "rcs_log_unit": "dBsm",
        "rcs_linear_quantity": "sigma_3d",
        "samples": samples,
        "metadata": {
            "formulation": label,
            "geometry_kind": kind,
            "expanded_to_360": bool(expand_to_360),
            "per_frequency": per_freq,
            "cfie_alpha": float(cfie_alpha),
            "workers": int(workers),
            "warnings": list(materials.warnings),
            "precision": prec,
        },
    }


# ─────────────────────────────────────────────────────────────────────────────
# post-processing of solved sweeps (bor_dispatch.py:705-925)
# ─────────────────────────────────────────────────────────────────────────────

def _amplitudes_by_frequency(res: Dict[str, Any]) -> Dict[float, Tuple[np.ndarray, np.ndarray]]:
    rows: Dict[float, List[Tuple[float, complex]]] = {}
    for s in res["samples"]:
        rows.setdefault(float(s["frequency_ghz"]), []).append(
            (float(s["theta_inc_deg"]), complex(s["rcs_amp_real"], s["rcs_amp_imag"])))
    out = {}
    for f, v in rows.items():
        v.sort(key=lambda t: t[0])
        out[f] = (np.array([t for t, _ in v]), np.array([a for _, a in v]))
    return out


def _unit_rows(v: np.ndarray) -> np.ndarray:
    return v / np.linalg.norm(v, axis=-1, keepdims=True)


def bor_az_el_grid(res_vv: Dict[str, Any], res_hh: Dict[str, Any], azimuths_deg, elevations_deg,
                   axis_az_deg: float = 0.0, axis_el_deg: float = 0.0) -> Dict[str, Any]:
    """Radar-frame monostatic VV, HH and VH over an (azimuth, elevation) grid
    from the VV and HH aspect sweeps of one body (bor_dispatch.py:705-824).

    A monostatic look at a body of revolution sees only the aspect theta
    between the look direction d and the axis, and in the look's meridian
    plane the scattering matrix is diag(F_VV(theta), F_HH(theta)).  Each
    look's theta is interpolated (linearly in the complex amplitude) from the
    sweeps, and the diagonal matrix is rotated into the earth frame: with the
    meridian basis (v_t, h_t), h_t along axis x d, and the radar basis
    (v_r, h_r), h_r along z x d, M_ij = (target_i . radar_j) and
    S = M^T diag(F_VV, F_HH) M.  Looks along the axis or straight up/down
    keep M = I.  The returned VV/HH are earth-frame channels (for a
    horizontal axis on the el = 0 cut, radar VV is the solver's HH)."""
    if canonical_polarization(res_vv.get("polarization")) != "VV" or \
            canonical_polarization(res_hh.get("polarization")) != "HH":
        raise ValueError("pass the VV sweep first and the HH sweep second.")
    az = np.atleast_1d(np.asarray(azimuths_deg, dtype=float))
    el = np.atleast_1d(np.asarray(elevations_deg, dtype=float))
    fv, fh = _amplitudes_by_frequency(res_vv), _amplitudes_by_frequency(res_hh)
    freqs = sorted(fv)
    if sorted(fh) != freqs:
        raise ValueError("the VV and HH sweeps cover different frequencies.")

    def direction(a_deg, e_deg):
        a, e = np.radians(a_deg), np.radians(e_deg)
        return np.stack([np.cos(e) * np.cos(a), np.cos(e) * np.sin(a), np.sin(e)], axis=-1)

    A, E = np.meshgrid(az, el, indexing="ij")
    d = direction(A, E)                                            # [n_az, n_el, 3]
    ax = direction(np.float64(axis_az_deg), np.float64(axis_el_deg))
    theta = np.degrees(np.arccos(np.clip(d @ ax, -1.0, 1.0)))
    c = np.cross(ax, d)
    r = np.cross(np.array([0.0, 0.0, 1.0]), d)
    degenerate = (np.linalg.norm(c, axis=-1) < 1e-9) | (np.linalg.norm(r, axis=-1) < 1e-9)
    safe = lambda v: np.where(degenerate[..., None], np.array([1.0, 0.0, 0.0]), v)
    h_t = _unit_rows(safe(c))
    v_t = _unit_rows(np.cross(h_t, d))
    h_r = _unit_rows(safe(r))
    v_r = _unit_rows(np.cross(h_r, d))
    dot = lambda a, b: np.sum(a * b, axis=-1)
    M = np.stack([np.stack([dot(v_t, v_r), dot(v_t, h_r)], -1), np.stack([dot(h_t, v_r), dot(h_t, h_r)], -1)], -2)
    M[degenerate] = np.eye(2)
    shape = A.shape + (len(freqs),)
    amp = {ch: np.zeros(shape, dtype=complex) for ch in ("VV", "HH", "VH")}
    for kf, f in enumerate(freqs):
        tv, av = fv[f]
        th_, ah = fh[f]
        Fv = np.interp(theta, tv, av.real) + 1j * np.interp(theta, tv, av.imag)
        Fh = np.interp(theta, th_, ah.real) + 1j * np.interp(theta, th_, ah.imag)
        m00, m01, m10, m11 = M[..., 0, 0], M[..., 0, 1], M[..., 1, 0], M[..., 1, 1]
        amp["VV"][..., kf] = m00 * m00 * Fv + m10 * m10 * Fh
        amp["HH"][..., kf] = m01 * m01 * Fv + m11 * m11 * Fh
        amp["VH"][..., kf] = m00 * m01 * Fv + m10 * m11 * Fh
    return {
        "azimuths_deg": az.tolist(),
        "elevations_deg": el.tolist(),
        "frequencies_ghz": freqs,
        "axis_az_deg": float(axis_az_deg),
        "axis_el_deg": float(axis_el_deg),
        "theta_map_deg": theta,
        "amp": amp,
        "sigma": {ch: 4.0 * math.pi * np.abs(a) ** 2 for ch, a in amp.items()},
        "rcs_log_unit": "dBsm",
        "rcs_linear_quantity": "sigma_3d",
    }


def solve_adaptive_frequency_sweep_bor(geometry_snapshot: Dict[str, Any], freq_start_ghz: float,
                                       freq_stop_ghz: float, elevations_deg: List[float], polarization: str,
                                       initial_points: int = 11, max_refinements: int = 3,
                                       db_threshold: float = 1.0, max_total_points: int = 201,
                                       **kwargs) -> Dict[str, Any]:
    """Broadband sweep (bor_dispatch.py:827-925): a uniform seed of
    frequencies, then rounds of midpoints between neighbours whose RCS at any
    requested aspect differs by more than db_threshold, until a round adds
    none, max_refinements rounds ran or max_total_points is reached.
    Frequencies are keyed rounded to 1e-12 GHz, so none is solved twice;
    every round is one GPU batch."""
    if freq_start_ghz <= 0 or freq_stop_ghz <= 0:
        raise ValueError("frequencies must be positive.")
    if freq_start_ghz >= freq_stop_ghz:
        raise ValueError("freq_start_ghz must be below freq_stop_ghz.")
    n0 = max(3, int(initial_points))
    key = lambda f: round(float(f), 12)
    abort = kwargs.get("abort_event")
    by_freq: Dict[float, List[Dict[str, Any]]] = {}

    def solve(fs: List[float]) -> None:
        if fs:
            res = solve_monostatic_rcs_bor(geometry_snapshot=geometry_snapshot, frequencies_ghz=fs,
                                           elevations_deg=elevations_deg, polarization=polarization, **kwargs)
            for s in res["samples"]:
                by_freq.setdefault(key(s["frequency_ghz"]), []).append(s)

    freqs = sorted({key(f) for f in np.linspace(freq_start_ghz, freq_stop_ghz, n0)})
    solve(freqs)
    rounds = 0
    for _ in range(max_refinements):
        if (abort is not None and abort.is_set()) or len(freqs) >= max_total_points:
            break
        fresh = set()
        for aspect in elevations_deg:
            level = {}
            for f in freqs:
                hit = next((s for s in by_freq.get(f, []) if abs(s["theta_inc_deg"] - aspect) < 0.01), None)
                if hit is not None:
                    level[f] = hit["rcs_db"]
            for f0, f1 in zip(freqs[:-1], freqs[1:]):
                if f0 in level and f1 in level and abs(level[f1] - level[f0]) > db_threshold:
                    mid = key(0.5 * (f0 + f1))
                    if mid not in by_freq and mid != f0 and mid != f1:
                        fresh.add(mid)
        room = max_total_points - len(freqs)
        if not fresh or room <= 0:
            break
        new = sorted(fresh)[:room]
        solve(new)
        freqs = sorted(set(freqs) | set(new))
        rounds += 1
    pol = canonical_polarization(polarization)
    samples = sorted((s for v in by_freq.values() for s in v), key=lambda s: (s["frequency_ghz"], s["theta_inc_deg"]))
    return {
        "solver": "bor_mom_rcs",
        "scattering_mode": "monostatic_adaptive",
        "polarization": pol,
        "polarization_export": pol,
        "rcs_log_unit": "dBsm",
        "rcs_linear_quantity": "sigma_3d",
        "samples": samples,
        "metadata": {
            "formulation": "BoR adaptive frequency sweep",
            "initial_points": n0,
            "final_point_count": len(freqs),
            "refinement_count": rounds,
            "db_threshold": db_threshold,
            "freq_start_ghz": freq_start_ghz,
            "freq_stop_ghz": freq_stop_ghz,
        },
    }
