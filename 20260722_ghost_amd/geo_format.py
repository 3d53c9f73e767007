"""The .geo geometry text format and the solver snapshot built from it.

Drop-in for the reference's ``geometry_io`` reading and writing surface
(geometry_io.py:73-250, 511-550) used by the BoR front end: a ``.geo`` file
holds a title, segments (a header ``Segment: name [type]``, a
``properties: type n ibc pos_mat neg_mat`` line and primitive lines
``x1 y1 x2 y2``), then an ``IBCS:`` (or ``IBCS_Resistances:``) section and a
``Dielectrics:`` section of whitespace-separated rows.  For a body of
revolution x is rho and y is z.

    title, segments, ibcs, dielectrics = parse_geometry(text)
    snapshot = build_geometry_snapshot(title, segments, ibcs, dielectrics)
    text = build_geometry_text(title, segments, ibcs, dielectrics)

Comment lines start with '#'.  Errors are ValueError naming the offending
line.
"""

from __future__ import annotations

from dataclasses import dataclass, field
from typing import Any, Dict, List, Optional, Tuple

from .materials import TABLE_FLAG_MIN, TAPER_KINDS

IBC_KINDS = TAPER_KINDS
TABULATED_FLAG_THRESHOLD = TABLE_FLAG_MIN - 1


@dataclass
class Segment:
    name: str
    seg_type: Optional[str]
    properties: List[str]
    x: List[float] = field(default_factory=list)
    y: List[float] = field(default_factory=list)


def _check_ibc_row(tokens: List[str], line: str) -> None:
    try:
        flag = int(tokens[0])
    except ValueError:
        raise ValueError(f"impedance row needs an integer flag first: {line!r}") from None
    if flag > TABULATED_FLAG_THRESHOLD:
        if len(tokens) != 1:
            raise ValueError(f"a table impedance row (flag > {TABULATED_FLAG_THRESHOLD}) is the flag alone: {line!r}")
        return
    if len(tokens) != 6:
        raise ValueError(f"an inline impedance row is 'flag kind R_start X_start R_end X_end': {line!r}")
    if tokens[1].strip().lower() not in IBC_KINDS:
        raise ValueError(f"impedance kind must be one of {IBC_KINDS}: {line!r}")


class _Reader:
    """Line-by-line state machine over the three sections of a .geo file."""

    def __init__(self):
        self.title = "Geometry"
        self.segments: List[Segment] = []
        self.ibcs: List[List[str]] = []
        self.dielectrics: List[List[str]] = []
        self.section = "segments"
        self.current: Optional[Segment] = None

    def _close_segment(self) -> None:
        if self.current is not None:
            self.segments.append(self.current)
            self.current = None

    def feed(self, line: str) -> None:
        low = line.lower()
        if low.startswith("title"):
            if ":" not in line:
                raise ValueError(f"title line needs a ':': {line!r}")
            self.title = line.split(":", 1)[1].strip() or self.title
            return
        if low.startswith("ibcs:") or low.startswith("ibcs_resistances:"):
            self._close_segment()
            self.section = "ibcs"
            return
        if low.startswith("dielectrics:"):
            self._close_segment()
            self.section = "dielectrics"
            return
        if self.section == "ibcs":
            toks = line.split()
            _check_ibc_row(toks, line)
            self.ibcs.append(toks)
            return
        if self.section == "dielectrics":
            self.dielectrics.append(line.split())
            return
        if low.startswith("segment:"):
            self._close_segment()
            head = line.split(":", 1)[1].split()
            self.current = Segment(name=head[0] if head else "Unnamed",
                                   seg_type=head[1] if len(head) > 1 else None, properties=[])
            return
        if low.startswith("properties:"):
            props = line.split(":", 1)[1].split()
            if len(props) != 5:
                raise ValueError(f"properties line has {len(props)} fields, not 5 (type n ibc pos_mat neg_mat): {line!r}")
            if self.current is None:
                raise ValueError(f"properties line before any 'Segment:' header: {line!r}")
            self.current.properties = props
            return
        toks = line.split()
        if len(toks) != 4:
            raise ValueError(f"primitive line needs 4 numbers (x1 y1 x2 y2), found {len(toks)}: {line!r}")
        if self.current is None:
            raise ValueError(f"primitive line before any 'Segment:' header: {line!r}")
        try:
            x1, y1, x2, y2 = (float(t) for t in toks)
        except ValueError:
            raise ValueError(f"primitive line holds a non-number: {line!r}") from None
        self.current.x += [x1, x2]
        self.current.y += [y1, y2]


def parse_geometry(text: str) -> Tuple[str, List[Segment], List[List[str]], List[List[str]]]:
    """(title, segments, ibc rows, dielectric rows) of a .geo text."""
    r = _Reader()
    for raw in text.splitlines():
        line = raw.strip()
        if line and not line.startswith("#"):
            r.feed(line)
    r._close_segment()
    return r.title, r.segments, r.ibcs, r.dielectrics


def _declared_type(seg: Segment) -> Optional[str]:
    if seg.seg_type:
        return seg.seg_type
    if seg.properties and str(seg.properties[0]).strip():
        return str(seg.properties[0]).strip()
    return None


def build_geometry_text(title: str, segments: List[Segment], ibcs_entries: List[List[str]],
                        dielectric_entries: List[List[str]]) -> str:
    """The .geo text of a geometry; coordinates are written with repr so a
    read-back gives the same floats, blank property fields as concrete tokens."""
    out = [f"Title: {title}"]
    for seg in segments:
        st = _declared_type(seg)
        out.append(f"Segment: {seg.name} {st}" if st else f"Segment: {seg.name}")
        props = [("" if p is None else str(p).strip()) for p in list(seg.properties)[:5]]
        props += [""] * (5 - len(props))
        props[0] = props[0] or (st or "2")
        props[1:] = [p or "0" for p in props[1:]]
        out.append("properties: " + " ".join(props))
        if len(seg.x) != len(seg.y) or len(seg.x) % 2:
            raise ValueError(f"segment {seg.name!r}: x / y lists must be equal and even in length.")
        for i in range(0, len(seg.x), 2):
            out.append(" ".join(repr(float(v)) for v in (seg.x[i], seg.y[i], seg.x[i + 1], seg.y[i + 1])))
    out.append("IBCS_Resistances:")
    out += [" ".join(map(str, row)) for row in ibcs_entries]
    out.append("Dielectrics:")
    out += [" ".join(map(str, row)) for row in dielectric_entries]
    return "\n".join(out) + "\n"


def build_geometry_snapshot(title: str, segments: List[Segment], ibcs_entries: List[List[str]],
                            dielectric_entries: List[List[str]]) -> Dict[str, Any]:
    """The solver snapshot dict: segments with point_pairs {x1, y1, x2, y2}."""
    segs = []
    for seg in segments:
        n = min(len(seg.x), len(seg.y)) // 2
        pairs = [{"x1": seg.x[2 * i], "y1": seg.y[2 * i], "x2": seg.x[2 * i + 1], "y2": seg.y[2 * i + 1]}
                 for i in range(n)]
        segs.append({"name": seg.name, "seg_type": _declared_type(seg), "properties": list(seg.properties),
                     "point_pairs": pairs})
    return {"title": title, "segment_count": len(segments), "segments": segs,
            "ibcs": [list(r) for r in ibcs_entries], "dielectrics": [list(r) for r in dielectric_entries]}


def load_geometry_snapshot(path: str) -> Dict[str, Any]:
    """Read a .geo file straight into a solver snapshot."""
    with open(path, "r", encoding="utf-8") as fh:
        return build_geometry_snapshot(*parse_geometry(fh.read()))
