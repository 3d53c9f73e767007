"""Surface-impedance and dielectric material lookup for .geo snapshots.

The drop-in for the reference's ``rcs_solver.MaterialLibrary`` surface
(rcs_solver.py:303-449) that ``bor_dispatch`` uses: impedance flags of TYPE 2
segments (inline constant or tapered Z_s, or ``mat.<flag>`` frequency
tables) and dielectric flags (inline (eps, mu), or tables).  Host-side
bookkeeping only; the values feed the GPU fill through
``ghost_bor_set_impedance``.

Row and file formats (fixed by the .geo format; geometry_io.py:1-60,
rcs_solver.py:196-268, 803-860):

* IBC row, inline: ``flag kind R_start X_start R_end X_end`` with kind in
  constant / linear / cosine / exp; the impedance along a segment follows the
  kind between the start and end values over the segment's normalised arc
  coordinate s in [0, 1] (``constant`` ignores the end values).
* IBC row, tabulated: a single flag above 50; ``mat.<flag>`` next to the
  geometry (or in the working directory) holds rows ``f_GHz R X``.
* Dielectric row: ``flag eps' eps'' mu' mu''`` (missing values default to
  1, 0, 1, 0), or a flag above 50 with ``mat.<flag>`` rows
  ``f_GHz eps' eps'' mu' mu''``.  Imaginary parts are taken as written
  (e^{+jwt}: lossy media have negative imaginary parts); a non-finite or
  zero value falls back to 1.
"""

from __future__ import annotations

import cmath
import math
import os
from dataclasses import dataclass
from typing import Dict, Iterable, List, Optional, Sequence, Tuple, Union

import numpy as np

TABLE_FLAG_MIN = 51           # flags above 50 name a mat.<flag> table
TAPER_KINDS = ("constant", "linear", "cosine", "exp")
_TINY = 1e-12


def flag_value(token) -> int:
    """Integer flag of a token ('3', '3.0', 'mat.57'); 0 when unreadable."""
    text = str(token).strip().lower()
    if text.startswith("mat."):
        text = text[4:]
    try:
        return int(float(text)) if text else 0
    except ValueError:
        return 0


def _number(token, default: float) -> float:
    try:
        return float(token)
    except (TypeError, ValueError):
        return default


def _finite(z: complex, what: str) -> complex:
    z = complex(z)
    if not (math.isfinite(z.real) and math.isfinite(z.imag)):
        raise ValueError(f"{what} is not finite ({z!r}).")
    return z


def _or_unity(z: complex) -> complex:
    if not (math.isfinite(z.real) and math.isfinite(z.imag)) or abs(z) < _TINY:
        return 1.0 + 0.0j
    return z


@dataclass
class ImpedanceTaper:
    """Z_s along a segment's arc coordinate s in [0, 1] (start as drawn = 0)."""

    kind: str
    z_start: complex
    z_end: complex

    def __post_init__(self):
        if self.kind not in TAPER_KINDS:
            raise ValueError(f"impedance kind '{self.kind}' is not one of {TAPER_KINDS}.")
        self.z_start = _finite(self.z_start, "taper start impedance")
        self.z_end = self.z_start if self.kind == "constant" else _finite(self.z_end, "taper end impedance")

    def evaluate(self, arc_s: float) -> complex:
        s = min(1.0, max(0.0, float(arc_s)))
        a, b = self.z_start, self.z_end
        if self.kind == "constant":
            return a
        if self.kind == "exp":
            # geometric interpolation; zero ends are lifted to a tiny resistance
            la = cmath.log(a if abs(a) > _TINY else complex(_TINY, 0.0))
            lb = cmath.log(b if abs(b) > _TINY else complex(_TINY, 0.0))
            return cmath.exp((1.0 - s) * la + s * lb)
        w = s if self.kind == "linear" else 0.5 * (1.0 - math.cos(math.pi * s))
        return (1.0 - w) * a + w * b


@dataclass
class ComplexTable:
    """Frequency (GHz) -> complex value, linear in real and imaginary parts."""

    freqs_ghz: np.ndarray
    values: np.ndarray

    def sample(self, f_ghz: float) -> complex:
        if self.freqs_ghz.size == 1:
            return complex(self.values[0])
        return complex(np.interp(f_ghz, self.freqs_ghz, self.values.real),
                       np.interp(f_ghz, self.freqs_ghz, self.values.imag))


@dataclass
class MediumTable:
    """Frequency (GHz) -> (eps_r, mu_r), linear in real and imaginary parts."""

    freqs_ghz: np.ndarray
    eps_values: np.ndarray
    mu_values: np.ndarray

    def sample(self, f_ghz: float) -> Tuple[complex, complex]:
        e = ComplexTable(self.freqs_ghz, self.eps_values).sample(f_ghz)
        m = ComplexTable(self.freqs_ghz, self.mu_values).sample(f_ghz)
        return e, m


def material_file(base_dir: str, flag: int) -> str:
    """Path of mat.<flag>: next to the geometry, else in the working directory."""
    name = f"mat.{flag}"
    for d in (base_dir, os.getcwd()):
        p = os.path.join(d, name)
        if os.path.isfile(p):
            return p
    raise FileNotFoundError(f"material table {name} not found in {base_dir!r} or the working directory.")


def read_table_rows(path: str, columns: int) -> np.ndarray:
    """Numeric rows of a material table: '#' comments and short or
    non-numeric lines skipped, sorted by frequency, the last of equal
    frequencies kept.  Returns [rows, columns]."""
    by_freq: Dict[float, List[float]] = {}
    with open(path, "r") as fh:
        for lineno, line in enumerate(fh, start=1):
            toks = line.split("#", 1)[0].split()
            if len(toks) < columns:
                continue
            try:
                vals = [float(t) for t in toks[:columns]]
            except ValueError:
                continue
            if not all(math.isfinite(v) for v in vals):
                raise ValueError(f"{path}:{lineno}: non-finite value in {toks[:columns]}.")
            by_freq[vals[0]] = vals
    if not by_freq:
        raise ValueError(f"{path} holds no numeric rows.")
    return np.array([by_freq[f] for f in sorted(by_freq)], dtype=float)


def load_impedance_table(path: str) -> ComplexTable:
    r = read_table_rows(path, 3)
    return ComplexTable(r[:, 0].copy(), r[:, 1] + 1j * r[:, 2])


def load_dielectric_table(path: str) -> MediumTable:
    r = read_table_rows(path, 5)
    return MediumTable(r[:, 0].copy(), r[:, 1] + 1j * r[:, 2], r[:, 3] + 1j * r[:, 4])


ImpedanceModel = Union[complex, ComplexTable, ImpedanceTaper]
MediumModel = Union[Tuple[complex, complex], MediumTable]


class MaterialLibrary:
    """Impedance and medium models by flag; out-of-range table lookups and
    solver notes are collected once each in ``warnings``."""

    def __init__(self, impedance_models: Dict[int, ImpedanceModel],
                 dielectric_models: Dict[int, MediumModel]):
        self.impedance_models = impedance_models
        self.dielectric_models = dielectric_models
        self.warnings: List[str] = []

    @classmethod
    def from_entries(cls, ibcs_entries: Iterable[Sequence], dielectric_entries: Iterable[Sequence],
                     base_dir: str) -> "MaterialLibrary":
        imp: Dict[int, ImpedanceModel] = {}
        med: Dict[int, MediumModel] = {}
        for row in ibcs_entries:
            if not row:
                continue
            flag = flag_value(row[0])
            if flag <= 0:
                continue
            if flag >= TABLE_FLAG_MIN:
                imp[flag] = load_impedance_table(material_file(base_dir, flag))
                continue
            data = [str(t).strip() for t in row[1:] if str(t).strip()]
            if len(data) != 5:
                raise ValueError(f"impedance flag {flag}: an inline row is 'kind R_start X_start R_end X_end' "
                                 f"after the flag; found {len(data)} values.")
            r0, x0, r1, x1 = (_number(t, 0.0) for t in data[1:])
            imp[flag] = ImpedanceTaper(data[0].lower(), complex(r0, x0), complex(r1, x1))
        for row in dielectric_entries:
            if not row:
                continue
            flag = flag_value(row[0])
            if flag <= 0:
                continue
            if flag >= TABLE_FLAG_MIN:
                med[flag] = load_dielectric_table(material_file(base_dir, flag))
                continue
            v = [_number(row[i], dflt) if len(row) > i else dflt for i, dflt in ((1, 1.0), (2, 0.0), (3, 1.0), (4, 0.0))]
            eps = _or_unity(_finite(complex(v[0], v[1]), f"dielectric flag {flag} permittivity"))
            mu = _or_unity(_finite(complex(v[2], v[3]), f"dielectric flag {flag} permeability"))
            med[flag] = (eps, mu)
        return cls(imp, med)

    def _note_range(self, what: str, flag: int, table, f_ghz: float) -> None:
        lo, hi = float(np.min(table.freqs_ghz)), float(np.max(table.freqs_ghz))
        if f_ghz < lo or f_ghz > hi:
            self.warn_once(f"{what} flag {flag} sampled at {f_ghz:g} GHz outside its table [{lo:g}, {hi:g}] GHz.")

    def get_impedance(self, flag: int, freq_ghz: float, arc_s: Optional[float] = None) -> complex:
        """Z_s (ohm) of an impedance flag at a frequency and, for tapers, at
        the normalised arc position along its segment (0.5 if not given);
        0 for flag <= 0 or an unknown flag (PEC)."""
        model = self.impedance_models.get(flag) if flag > 0 else None
        if model is None:
            return 0.0 + 0.0j
        if isinstance(model, ComplexTable):
            self._note_range("impedance", flag, model, freq_ghz)
            return _finite(model.sample(freq_ghz), f"impedance flag {flag} at {freq_ghz:g} GHz")
        if isinstance(model, ImpedanceTaper):
            s = 0.5 if arc_s is None else float(arc_s)
            return _finite(model.evaluate(s), f"impedance flag {flag} at s = {s:g}")
        return _finite(model, f"impedance flag {flag}")

    def is_tapered_impedance(self, flag: int) -> bool:
        return isinstance(self.impedance_models.get(flag), ImpedanceTaper)

    def get_medium(self, flag: int, freq_ghz: float) -> Tuple[complex, complex]:
        """(eps_r, mu_r) of a dielectric flag; free space for flag <= 0 or unknown."""
        model = self.dielectric_models.get(flag) if flag > 0 else None
        if model is None:
            return 1.0 + 0.0j, 1.0 + 0.0j
        if isinstance(model, MediumTable):
            self._note_range("dielectric", flag, model, freq_ghz)
            eps, mu = model.sample(freq_ghz)
        else:
            eps, mu = model
        return _or_unity(eps), _or_unity(mu)

    def warn_once(self, message: str) -> None:
        if message not in self.warnings:
            self.warnings.append(message)
