"""Deterministic synthetic generatrices for the benchmark configurations.

All bodies are closed (both ends on the axis) and traversed from the +z end,
the reference's generatrix convention (bor_kernels.py:38-47).
"""

from __future__ import annotations

import math

import numpy as np
from scipy.interpolate import CubicSpline

C0 = 299_792_458.0


def _resample(profile: np.ndarray, n_elems: int) -> np.ndarray:
    """Uniform arc-length resampling of a dense (rho, z) polyline to n_elems elements."""
    seg = np.hypot(*np.diff(profile, axis=0).T)
    s = np.concatenate([[0.0], np.cumsum(seg)])
    t = np.linspace(0.0, s[-1], n_elems + 1)
    out = np.column_stack([np.interp(t, s, profile[:, 0]), np.interp(t, s, profile[:, 1])])
    out[0, 0] = 0.0
    out[-1, 0] = 0.0
    return out


def sphere(radius: float = 0.5, n_elems: int = 120) -> np.ndarray:
    """Pole-to-pole semicircle (bor_solver.py:2179-2182)."""
    th = np.linspace(0.0, math.pi, n_elems + 1)
    return np.column_stack([radius * np.sin(th), radius * np.cos(th)])


def ogive_cylinder_cone(n_elems: int = 1200, wavelength: float = 1.0, seed: int = 20260722) -> np.ndarray:
    """Rocket-like body: tangent-ogive nose, cylinder, conical boat-tail and a
    flat base; length 10 wavelengths, max radius 1 wavelength.  A seeded cubic
    spline of small radial offsets (<= 0.02 wavelength) perturbs the profile;
    nodes are uniform in arc length."""
    rng = np.random.default_rng(seed)
    lam = wavelength
    L, R = 10.0 * lam, 1.0 * lam
    Ln = rng.uniform(2.5, 3.5) * lam       # nose
    La = rng.uniform(1.5, 2.5) * lam       # boat-tail
    Rb = rng.uniform(0.35, 0.55) * R       # base radius
    Lc = L - Ln - La
    z_tip = 0.5 * L
    x = np.linspace(0.0, L, 40001)         # axial distance from the tip
    rho = np.empty_like(x)
    Ro = (R * R + Ln * Ln) / (2.0 * R)
    nose = x <= Ln
    rho[nose] = np.sqrt(np.maximum(Ro * Ro - (Ln - x[nose]) ** 2, 0.0)) + R - Ro
    cyl = (x > Ln) & (x <= Ln + Lc)
    rho[cyl] = R
    aft = x > Ln + Lc
    rho[aft] = R - (R - Rb) * (x[aft] - Ln - Lc) / La
    # seeded smooth ripple, zero at the tip and at the base
    knots = np.linspace(0.0, L, 9)
    offs = rng.uniform(-0.02, 0.02, 9) * lam
    offs[0] = offs[-1] = 0.0
    ripple = CubicSpline(knots, offs, bc_type="clamped")(x)
    rho = np.maximum(rho + ripple * np.clip(rho / (0.2 * R), 0.0, 1.0), 0.0)
    rho[0] = 0.0
    side = np.column_stack([rho, z_tip - x])
    base = np.column_stack([np.linspace(rho[-1], 0.0, 200)[1:], np.full(199, z_tip - L)])
    return _resample(np.vstack([side, base]), n_elems)


def random_smooth(length: float = 1.0, rmin: float = 0.05, rmax: float = 0.2, n_elems: int = 200,
                  seed: int = 7) -> np.ndarray:
    """Seeded smooth body: rho(z) spline in [rmin, rmax] with rounded ends."""
    rng = np.random.default_rng(seed)
    z = np.linspace(0.0, length, 4001)
    knots = np.linspace(0.0, length, 7)
    vals = rng.uniform(rmin, rmax, 7)
    core = CubicSpline(knots, vals)(z)
    core = np.clip(core, rmin, rmax)
    env = np.sqrt(np.clip(1.0 - ((2.0 * z / length) - 1.0) ** 8, 0.0, 1.0))
    rho = core * env
    rho[0] = rho[-1] = 0.0
    prof = np.column_stack([rho, 0.5 * length - z])
    return _resample(prof, n_elems)


def cone_sphere(n_elems: int = 600, half_angle_deg: float = 15.0, radius: float = 1.0) -> np.ndarray:
    """Cone tip at +z, flank tangent to a sphere cap of the given radius, pole at -z;
    nodes uniform in arc length."""
    al = math.radians(half_angle_deg)
    tp = math.pi / 2 - al
    zt, rt = radius * math.cos(tp), radius * math.sin(tp)
    tip = zt + rt / math.tan(al)
    cone = np.column_stack([np.linspace(0.0, rt, 2001), np.linspace(tip, zt, 2001)])
    th = np.linspace(tp, math.pi, 4001)[1:]
    sph = np.column_stack([radius * np.sin(th), radius * np.cos(th)])
    return _resample(np.vstack([cone, sph]), n_elems)


def frequency_sweep(n_freq: int = 64, seed: int = 7, n_theta: int = 181):
    """Config 3: one seeded smooth body (rho in [0.05, 0.2] m over 1 m) at
    n_freq frequencies geomspace(0.2, 2 GHz).  The mesh grows with frequency
    (50 -> ~1000 elements: systems of ~100 -> ~2000 unknowns), m_max =
    ceil(1.2 k rho_max) + 5 (7 -> 16), n_theta aspects x 2 pols.  Returns
    solve_bor_batch job dicts."""
    jobs = []
    for f in np.geomspace(0.2e9, 2e9, n_freq):
        ne = int(round(50 * (f / 0.2e9) ** 1.3))
        nodes = random_smooth(1.0, 0.05, 0.2, ne, seed)
        k = 2.0 * math.pi * f / C0
        m_max = int(math.ceil(1.2 * k * 0.2)) + 5     # rho bound 0.2 m
        jobs.append({"points": nodes, "freq_hz": float(f), "thetas_deg": np.linspace(0.0, 180.0, n_theta),
                     "formulation": "cfie", "cfie_alpha": 0.5, "n_modes": m_max})
    return jobs


def config(index: int):
    """(name, nodes, freq_hz, thetas_deg, m_max) of a benchmark configuration
    (configuration 5, a batch of problems, is perturbed_batch)."""
    if index == 1:
        return ("pec-sphere-a0.5m-300MHz-120seg", sphere(0.5, 120), 3e8,
                np.arange(0.0, 181.0, 1.0), 12)
    if index == 2:
        lam = 1.0
        return ("ogive-cylinder-cone-10lambda-1200seg", ogive_cylinder_cone(1200, lam), C0 / lam,
                np.arange(0.0, 180.25, 0.5), 70)
    if index == 4:
        lam = 1.0
        return ("cone-sphere-15deg-600seg", cone_sphere(600, 15.0, lam), C0 / lam,
                np.arange(0.0, 181.0, 2.0), 40)
    raise ValueError(f"no configuration {index}")


def fourier_perturbation(nodes: np.ndarray, coef: np.ndarray, amplitude: float, taper: float = 0.2) -> np.ndarray:
    """Radial Fourier perturbation of a generatrix: d(t) = sum_k coef_k sin(k pi t)
    over the normalised arc length t in [0, 1] (zero at both ends), scaled to
    max |d| = amplitude, tapered by min(1, rho / taper) so thin ends stay off
    the axis (rho + d >= rho (1 - amplitude / taper))."""
    seg = np.hypot(*np.diff(nodes, axis=0).T)
    t = np.concatenate([[0.0], np.cumsum(seg)]) / np.sum(seg)
    k = np.arange(1, len(coef) + 1)
    d = np.sin(np.pi * np.outer(t, k)) @ np.asarray(coef, dtype=np.float64)
    d *= amplitude / np.max(np.abs(d))
    out = nodes.copy()
    out[:, 0] = nodes[:, 0] + d * np.minimum(1.0, nodes[:, 0] / taper)
    out[0, 0] = nodes[0, 0]
    out[-1, 0] = nodes[-1, 0]
    return out


def perturbed_batch(n_shapes: int = 32, n_coef: int = 16, amplitude: float = 0.05, seed: int = 5,
                    freq_scales=(0.5, 1.0, 1.5), n_theta: int = 91, base_elems: int = 1200, base_m: int = 70):
    """Config 5: n_shapes seeded Fourier radial perturbations (n_coef
    coefficients c_k ~ U(-1, 1) / k, max offset = amplitude wavelengths) of the
    ogive-cylinder-cone body (wavelength 1 m at f0 = c / 1 m), each at the
    frequencies f0 x freq_scales with lambda / 10 meshing (600, 1200, 1800
    elements) and m_max = base_m x scale; n_theta aspects x 2 pols.  Returns a
    list of (shape index, nodes, freq_hz, thetas_deg, m_max), shape-major."""
    rng = np.random.default_rng(seed)
    coefs = [rng.uniform(-1.0, 1.0, n_coef) / np.arange(1, n_coef + 1) for _ in range(n_shapes)]
    thetas = np.linspace(0.0, 180.0, n_theta)
    bases = {sc: ogive_cylinder_cone(int(round(base_elems * sc)), 1.0) for sc in freq_scales}
    out = []
    for i, c in enumerate(coefs):
        for sc in freq_scales:
            nodes = fourier_perturbation(bases[sc], c, amplitude)
            out.append((i, nodes, C0 * sc, thetas, int(round(base_m * sc))))
    return out
