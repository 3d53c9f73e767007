"""Mie-series backscatter of a PEC sphere (analytic oracle).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the product path.

Restates mie_sphere.py:42-125, 178-183, 231-242 in the e^{+jwt} convention
(outgoing Riccati-Hankel xi_n = psi_n - j zeta_n):
    A_n = -psi_n'(x)/xi_n'(x)  (TM: U' = 0),  B_n = -psi_n(x)/xi_n(x)  (TE: U = 0)
    sigma = lambda^2/(4 pi) |sum_{n>=1} (-1)^n (2n+1)(A_n - B_n)|^2
Leontovich (impedance) sphere, g = j Z_s / eta0 (mie_sphere.py:105-121):
    A_n = -(psi' - g psi)/(xi' - g xi),  B_n = -(psi + g psi')/(xi + g xi')
which reduces to the PEC pair at Z_s = 0 and gives A_n = B_n (zero
backscatter, Weston's theorem) at Z_s = eta0.
Homogeneous dielectric sphere (mie_sphere.py:124-141), index m = sqrt(eps mu)
with Im(m) <= 0: A_n, B_n from continuity with the interior psi_n(m x),
which vanish for eps = mu = 1 and tend to the PEC pair as eps -> infinity.
Riccati functions are built from scipy's spherical Bessel routines.
"""

import math

import numpy as np
from scipy import special

C0 = 299_792_458.0


def _nmax(x: float, pad: int = 12) -> int:
    """Wiscombe-style truncation (mie_sphere.py:74-77)."""
    x = abs(x)
    return max(6, int(math.ceil(x + 4.05 * x ** (1.0 / 3.0))) + pad)


def _riccati(nmax: int, x: float):
    n = np.arange(nmax + 1)
    j = special.spherical_jn(n, x)
    y = special.spherical_yn(n, x)
    dj = special.spherical_jn(n, x, derivative=True)
    dy = special.spherical_yn(n, x, derivative=True)
    psi, dpsi = x * j, j + x * dj
    zeta, dzeta = x * y, y + x * dy
    return psi, dpsi, psi - 1j * zeta, dpsi - 1j * dzeta


def pec_sphere_coeffs(x: float, nmax: int):
    psi, dpsi, xi, dxi = _riccati(nmax, x)
    return -dpsi / dxi, -psi / xi


def sigma_pec_sphere(radius_m: float, freq_hz: float) -> float:
    """Monostatic RCS (m^2) of a PEC sphere."""
    k = 2.0 * math.pi * freq_hz / C0
    x = k * radius_m
    a, b = pec_sphere_coeffs(x, _nmax(x))
    n = np.arange(1, len(a))
    s = np.sum(((-1.0) ** n) * (2 * n + 1) * (a[1:] - b[1:]))
    lam = 2.0 * math.pi / k
    return float(lam ** 2 / (4.0 * math.pi) * abs(s) ** 2)


ETA0 = 376.730313668


def impedance_sphere_coeffs(x: float, zs: complex, nmax: int):
    psi, dpsi, xi, dxi = _riccati(nmax, x)
    g = 1j * complex(zs) / ETA0
    return -(dpsi - g * psi) / (dxi - g * xi), -(psi + g * dpsi) / (xi + g * dxi)


def sigma_impedance_sphere(radius_m: float, freq_hz: float, zs_ohm: complex) -> float:
    """Monostatic RCS (m^2) of a sphere with surface impedance zs_ohm."""
    k = 2.0 * math.pi * freq_hz / C0
    x = k * radius_m
    a, b = impedance_sphere_coeffs(x, zs_ohm, _nmax(x))
    n = np.arange(1, len(a))
    s = np.sum(((-1.0) ** n) * (2 * n + 1) * (a[1:] - b[1:]))
    lam = 2.0 * math.pi / k
    return float(lam ** 2 / (4.0 * math.pi) * abs(s) ** 2)


def _riccati_complex(nmax: int, z: complex):
    """psi_n(z) = z j_n(z) and psi_n'(z) for complex z."""
    n = np.arange(nmax + 1)
    j = special.spherical_jn(n, z)
    dj = special.spherical_jn(n, z, derivative=True)
    return z * j, j + z * dj


def causal_index(eps_r: complex, mu_r: complex) -> complex:
    m = np.sqrt(complex(eps_r) * complex(mu_r))
    if m.imag > 0.0:
        m = -m
    if m.real < 0.0:
        m = -m
    return m


def dielectric_sphere_coeffs(x: float, eps_r: complex, mu_r: complex, nmax: int):
    """Homogeneous sphere (mie_sphere.py:124-141): interior psi_n(m x),
    continuity of U and U'/(eps or mu) scaled by the index m."""
    m = causal_index(eps_r, mu_r)
    psi, dpsi, xi, dxi = _riccati(nmax, x)
    pm, dpm = _riccati_complex(nmax, m * x)
    wa, wb = m / complex(eps_r), m / complex(mu_r)
    a = -(dpsi * pm - wa * dpm * psi) / (dxi * pm - wa * dpm * xi)
    b = -(dpsi * pm - wb * dpm * psi) / (dxi * pm - wb * dpm * xi)
    return a, b


def sigma_dielectric_sphere(radius_m: float, eps_r: complex, mu_r: complex, freq_hz: float) -> float:
    """Monostatic RCS (m^2) of a homogeneous dielectric / magnetic sphere."""
    k = 2.0 * math.pi * freq_hz / C0
    x = k * radius_m
    m = causal_index(eps_r, mu_r)
    a, b = dielectric_sphere_coeffs(x, eps_r, mu_r, _nmax(max(x, abs(m) * x)))
    n = np.arange(1, len(a))
    s = np.sum(((-1.0) ** n) * (2 * n + 1) * (a[1:] - b[1:]))
    lam = 2.0 * math.pi / k
    return float(lam ** 2 / (4.0 * math.pi) * abs(s) ** 2)


def _riccati_pair(nmax: int, z: complex):
    """psi_n(z) = z j_n(z), chi_n(z) = z y_n(z) and their derivatives for
    complex z (two independent radial solutions inside a layer)."""
    n = np.arange(nmax + 1)
    j = special.spherical_jn(n, z)
    dj = special.spherical_jn(n, z, derivative=True)
    y = special.spherical_yn(n, z)
    dy = special.spherical_yn(n, z, derivative=True)
    return z * j, j + z * dj, z * y, y + z * dy


def multilayer_pec_coeffs(radii, eps_list, mu_list, k0: float, nmax: int):
    """PEC core of radius radii[0] under layers j = 1..L (outer radii
    radii[1:], media inner to outer).  In layer j the radial function is
    U = psi(k_j r) + R_j chi(k_j r); the admittance Y = w_j U'/U (w_j =
    k_j / (k0 eps_j) for TM, k_j / (k0 mu_j) for TE) is continuous across
    every interface, so R_{j+1} follows from Y at r_j and the exterior
    U = psi(k0 r) + A xi(k0 r) from Y at the outer radius:
    A = (Y psi - psi') / (xi' - Y xi).  The core sets U' = 0 (TM) or U = 0
    (TE) at radii[0].  The same convention as dielectric_sphere_coeffs (the
    e^{+jwt} outgoing xi)."""
    L = len(eps_list)
    ks = [k0 * causal_index(e, u) for e, u in zip(eps_list, mu_list)]
    out = []
    for tm in (True, False):
        w = [ks[j] / (k0 * complex(eps_list[j] if tm else mu_list[j])) for j in range(L)]
        ps, dps, ch, dch = _riccati_pair(nmax, ks[0] * radii[0])
        R = -(dps / dch) if tm else -(ps / ch)
        for j in range(L):
            r = radii[j + 1]
            ps, dps, ch, dch = _riccati_pair(nmax, ks[j] * r)
            Y = w[j] * (dps + R * dch) / (ps + R * ch)
            if j + 1 < L:
                ps, dps, ch, dch = _riccati_pair(nmax, ks[j + 1] * r)
                R = (Y * ps - w[j + 1] * dps) / (w[j + 1] * dch - Y * ch)
        psi, dpsi, xi, dxi = _riccati(nmax, k0 * radii[-1])
        out.append((Y * psi - dpsi) / (dxi - Y * xi))
    return out[0], out[1]


def sigma_multilayer_pec_sphere(radii, eps_list, mu_list, freq_hz: float) -> float:
    """Monostatic RCS (m^2) of a PEC sphere under concentric layers
    (radii = [core, r1, ..., rL] ascending, media inner to outer)."""
    k = 2.0 * math.pi * freq_hz / C0
    scale = max([abs(causal_index(e, u)) for e, u in zip(eps_list, mu_list)] + [1.0])
    a, b = multilayer_pec_coeffs([float(r) for r in radii], [complex(e) for e in eps_list],
                                 [complex(u) for u in mu_list], k, _nmax(scale * k * radii[-1]))
    n = np.arange(1, len(a))
    s = np.sum(((-1.0) ** n) * (2 * n + 1) * (a[1:] - b[1:]))
    lam = 2.0 * math.pi / k
    return float(lam ** 2 / (4.0 * math.pi) * abs(s) ** 2)


def sigma_coated_pec_sphere(core_radius: float, outer_radius: float, eps_r: complex, mu_r: complex,
                            freq_hz: float) -> float:
    """One coating layer (mie_sphere.py:254-264)."""
    return sigma_multilayer_pec_sphere([core_radius, outer_radius], [eps_r], [mu_r], freq_hz)
