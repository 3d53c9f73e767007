"""``.grim`` export of BoR monostatic results (and of radar-frame az/el grids).

A ``.grim`` file is a NumPy ``.npz`` archive with a fixed set of arrays and
string fields (the format of the reference's grim_io.py:80-311, which its
viewer and downstream tools read):

* axes ``azimuths``, ``elevations``, ``frequencies`` (deg, deg, GHz) and
  ``polarizations`` (string array);
* ``rcs_power`` (linear sigma) and ``rcs_phase`` (rad), float32
  [n_az, n_el, n_f, n_pol], NaN where no sample exists;
* optionally ``rcs_amp_real`` / ``rcs_amp_imag`` (the solver's complex
  far-field amplitude, float32), flagged by
  ``raw_complex_amplitude_preserved`` and ``complex_field_domain``;
* ``polarization_alias_primary``, ``polarization_aliases_json``,
  ``rcs_domain``, ``power_domain``, ``units`` (JSON), ``phase_reference``,
  ``source_path``, ``history``.

For a monostatic sweep the azimuth axis holds the aspect angles theta and
the elevation axis the single value 0.  BoR results are sigma_3d in m^2,
shown as dBsm.
"""

from __future__ import annotations

import cmath
import json
import math
import os
from typing import Any, Dict, List, Optional

import numpy as np

_VV_NAMES = ["TE", "VV", "V", "VERTICAL"]
_HH_NAMES = ["TM", "HH", "H", "HORIZONTAL"]
_MIN_AMP = 1e-12


def _pol_label(label: Optional[str]) -> str:
    """'TE' (VV family) or 'TM' (HH family); empty means TM."""
    t = str(label or "").strip().upper()
    if not t or t in _HH_NAMES:
        return "TM"
    if t in _VV_NAMES:
        return "TE"
    raise ValueError(f"polarization {label!r} is neither a VV (TE, V, VERTICAL) nor an HH (TM, H, HORIZONTAL) name.")


def _grim_path(path: str) -> str:
    return path if path.lower().endswith(".grim") else path + ".grim"


def _write(path: str, fields: Dict[str, Any]) -> str:
    out = _grim_path(path)
    with open(out, "wb") as fh:
        np.savez(fh, **fields)
    return os.path.abspath(out)


def _units(log_unit: str, lin_quantity: str) -> str:
    return json.dumps({"azimuth": "deg", "elevation": "deg", "frequency": "GHz",
                       "rcs_log_unit": str(log_unit), "rcs_linear_quantity": str(lin_quantity)})


def _grid(samples: List[Dict[str, Any]], pol: str, source_path: str, history: str, keep_amp: bool,
          log_unit: str, lin_quantity: str) -> Dict[str, Any]:
    if not samples:
        raise ValueError("nothing to export: the result holds no samples.")
    az = np.array(sorted({float(s["theta_scat_deg"]) for s in samples}))
    fr = np.array(sorted({float(s["frequency_ghz"]) for s in samples}))
    ia = {v: i for i, v in enumerate(az)}
    jf = {v: i for i, v in enumerate(fr)}
    shape = (az.size, 1, fr.size, 1)
    power = np.full(shape, np.nan, dtype=np.float32)
    phase = np.full(shape, np.nan, dtype=np.float32)
    are = np.full(shape, np.nan, dtype=np.float32)
    aim = np.full(shape, np.nan, dtype=np.float32)
    for s in samples:
        idx = (ia[float(s["theta_scat_deg"])], 0, jf[float(s["frequency_ghz"])], 0)
        lin = float(s.get("rcs_linear", 0.0))
        lin = lin if math.isfinite(lin) and lin >= 0.0 else 0.0
        a = complex(*(v if math.isfinite(v) else 0.0
                      for v in (float(s.get("rcs_amp_real", 0.0)), float(s.get("rcs_amp_imag", 0.0)))))
        ph = cmath.phase(a) if abs(a) > _MIN_AMP else 0.0
        if not np.isnan(power[idx]):
            # a repeated sample must agree with the first (power, and phase modulo 2 pi)
            if abs(float(power[idx]) - lin) > _MIN_AMP:
                raise ValueError(f"two different powers at theta={az[idx[0]]:g}, f={fr[idx[2]]:g} GHz.")
            d = (ph - float(phase[idx]) + math.pi) % (2.0 * math.pi) - math.pi
            if abs(d) > 1e-5:
                raise ValueError(f"two different phases at theta={az[idx[0]]:g}, f={fr[idx[2]]:g} GHz.")
            continue
        power[idx], phase[idx] = lin, ph
        are[idx], aim[idx] = a.real, a.imag
    label = _pol_label(pol)
    fields = {
        "azimuths": az, "elevations": np.array([0.0]), "frequencies": fr,
        "polarizations": np.array(["HH" if label == "TM" else "VV"], dtype=str),
        "polarization_alias_primary": label,
        "polarization_aliases_json": json.dumps(_HH_NAMES if label == "TM" else _VV_NAMES),
        "rcs_power": power, "rcs_phase": phase,
        "rcs_domain": "power_phase", "power_domain": "linear_rcs",
        "source_path": source_path, "history": history,
        "units": _units(log_unit, lin_quantity),
        "phase_reference": "origin=(0,0), convention=exp(+jwt), monostatic far-field amplitude",
        "raw_complex_amplitude_preserved": bool(keep_amp),
    }
    if keep_amp:
        fields["rcs_amp_real"] = are
        fields["rcs_amp_imag"] = aim
        fields["complex_field_domain"] = "solver_raw_far_field_amplitude"
    return fields


def export_result_to_grim(result: Dict[str, Any], output_path: str, polarization: Optional[str] = None,
                          source_path: str = "", history: str = "",
                          preserve_raw_complex_amplitude: bool = True) -> List[str]:
    """Write a solve_monostatic_rcs_bor result (or any result with the same
    sample format) as .grim; a bistatic result becomes one file per incidence
    angle (``<root>_inc_<angle>.grim``).  Returns the absolute paths."""
    samples = result.get("samples") or []
    if not samples:
        raise ValueError("nothing to export: the result holds no samples.")
    pol = polarization or result.get("polarization_export") or result.get("polarization") or "TM"
    log_unit = str(result.get("rcs_log_unit", "dBke"))
    lin_q = str(result.get("rcs_linear_quantity", "sigma_2d"))
    if str(result.get("scattering_mode", "monostatic")).strip().lower() != "bistatic":
        return [_write(output_path, _grid(samples, pol, source_path, history, preserve_raw_complex_amplitude,
                                          log_unit, lin_q))]
    groups: Dict[float, List[Dict[str, Any]]] = {}
    for s in samples:
        groups.setdefault(float(s.get("theta_inc_deg", 0.0)), []).append(s)
    root = _grim_path(output_path)[:-5]
    paths = []
    for inc in sorted(groups):
        tag = (f"{inc:.6f}".rstrip("0").rstrip(".") or "0").replace("-", "m").replace(".", "p")
        hist = (history + f" | theta_inc_deg={inc:g}").strip(" |")
        paths.append(_write(f"{root}_inc_{tag}", _grid(groups[inc], pol, source_path, hist,
                                                          preserve_raw_complex_amplitude, log_unit, lin_q)))
    return paths


def save_bor_az_el_grim(grid: Dict[str, Any], output_path: str, source_path: str = "",
                        history: str = "") -> List[str]:
    """Write a bor_dispatch.bor_az_el_grid result as three .grim files
    (``<root>_VV``, ``_HH``, ``_VH``) with real azimuth and elevation axes."""
    az = np.asarray(grid["azimuths_deg"], dtype=float)
    el = np.asarray(grid["elevations_deg"], dtype=float)
    fr = np.asarray(grid["frequencies_ghz"], dtype=float)
    root = _grim_path(output_path)[:-5]
    aliases = {"VV": _VV_NAMES, "HH": _HH_NAMES, "VH": ["VH", "HV"]}
    note = f"bor_az_el_grid axis_az={grid['axis_az_deg']:g} axis_el={grid['axis_el_deg']:g}"
    paths = []
    for ch in ("VV", "HH", "VH"):
        amp = np.asarray(grid["amp"][ch])[..., None]
        paths.append(_write(f"{root}_{ch}", {
            "azimuths": az, "elevations": el, "frequencies": fr,
            "polarizations": np.array([ch], dtype=str),
            "polarization_alias_primary": ch, "polarization_aliases_json": json.dumps(aliases[ch]),
            "rcs_power": (4.0 * math.pi * np.abs(amp) ** 2).astype(np.float32),
            "rcs_phase": np.angle(amp).astype(np.float32),
            "rcs_domain": "power_phase", "power_domain": "linear_rcs",
            "source_path": source_path, "history": (history + " | " + note).strip(" |"),
            "units": _units("dBsm", "sigma_3d"),
            "phase_reference": "origin=(0,0), convention=exp(+jwt), monostatic radar-frame amplitude",
            "raw_complex_amplitude_preserved": True,
            "rcs_amp_real": amp.real.astype(np.float32), "rcs_amp_imag": amp.imag.astype(np.float32),
            "complex_field_domain": "solver_raw_far_field_amplitude",
        }))
    return paths
