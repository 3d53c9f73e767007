"""The oracle's PMCHWT (homogeneous penetrable body) path, pinned against the
reference's own vectors (tests/golden/make_golden_pmchwt.py: a 12-segment
sphere and a 16-segment cone-sphere, lossy dielectric and lossy magnetic
interiors) and against the dielectric Mie series (a closed form that does
not share the integral-equation discretisation)."""

import math
import os
import sys

import numpy as np
import pytest

from conftest import ROOT

sys.path.insert(0, ROOT)
from oracle import bor_oracle as bo  # noqa: E402
from oracle import mie  # noqa: E402

GOLD = os.path.join(ROOT, "tests", "golden")


def _load(name):
    return np.load(os.path.join(GOLD, name + ".npz"))


def _rel(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


@pytest.mark.parametrize("name", ["sphere12_pmchwt", "conesphere16_pmchwt"])
def test_medium_operators_and_system_match_reference(name):
    d = _load(name)
    nodes, f, nm = d["nodes"], float(d["freq_hz"]), int(d["m_max"])
    ext = bo.MediumOperators(nodes, f, nm)
    inner = bo.MediumOperators(nodes, f, nm, (complex(d["eps_r"]), complex(d["mu_r"])))
    assert abs(inner.k - complex(d["k_in"])) < 1e-13 * abs(inner.k)
    assert abs(inner.eta - complex(d["eta_in"])) < 1e-13 * abs(inner.eta)
    for m in range(-nm, nm + 1):
        assert _rel(inner.T(m), d[f"Ti_{m}"]) < 1e-12, m
        assert _rel(inner.P(m), d[f"Pi_{m}"]) < 1e-12, m
        assert _rel(ext.P(m), d[f"Pe_{m}"]) < 1e-12, m
        mask = d[f"mask_{m}"]
        assert np.array_equal(mask, np.tile(bo.basis_mask(ext.body, m), 2))
        A = bo.pmchwt_system(ext, inner, m)[np.ix_(mask, mask)]
        assert _rel(A, d[f"A_{m}"]) < 1e-12, m


@pytest.mark.parametrize("name", ["sphere12_pmchwt", "conesphere16_pmchwt"])
def test_solve_bor_dielectric_matches_reference(name):
    d = _load(name)
    out = bo.solve_bor_dielectric(d["nodes"], float(d["freq_hz"]), d["thetas"], complex(d["eps_r"]),
                                  complex(d["mu_r"]), n_modes=int(d["m_max"]), keep=True)
    for key in ("amp_vv", "amp_hh"):
        assert _rel(np.array(out[key]), d[key]) < 1e-12, key
    assert out["modes_used"] == int(d["modes_used"])
    for m, (X, B, _, _) in out["_modes"].items():
        assert _rel(B, d[f"B_{m}"]) < 1e-12, m
        assert _rel(X, d[f"X_{m}"]) < 1e-10, m


def test_causal_medium():
    """Im(m) <= 0 and Re(m) >= 0 whatever branch sqrt takes; eta_r = mu_r / m."""
    for eps, mu in ((2.5 - 0.1j, 1.0), (3.0 - 0.4j, 1.3 - 0.05j), (4.0, 1.0), (1.0, 1.0)):
        m, er = bo.causal_medium(eps, mu)
        assert m.imag <= 0.0 and m.real >= 0.0
        assert abs(m * m - complex(eps) * complex(mu)) < 1e-14 * abs(m * m)
        assert abs(er - complex(mu) / m) < 1e-15
    assert bo.causal_medium(1.0, 1.0) == (1.0 + 0j, 1.0 + 0j)


def test_free_space_medium_reduces_to_the_pec_operators():
    """A medium of eps = mu = 1 is free space: its T and P are the exterior's."""
    nodes = bo.sphere_nodes(0.5, 8)
    ext = bo.MediumOperators(nodes, 300e6, 2)
    same = bo.MediumOperators(nodes, 300e6, 2, (1.0, 1.0))
    for m in (-2, 0, 1):
        assert _rel(same.T(m), ext.T(m)) < 1e-14
        assert _rel(same.P(m), ext.P(m)) < 1e-14


@pytest.mark.parametrize("eps, mu, n, tol_db", [
    (4.0, 1.0, 32, 0.15),
    (3.0 - 0.4j, 1.3 - 0.05j, 32, 0.2),
])
def test_dielectric_sphere_against_mie(eps, mu, n, tol_db):
    """0.2 m sphere at 300 MHz (k0 a = 1.26): monostatic RCS within the
    discretisation error of the Mie series at every aspect, both
    polarisations (a sphere's monostatic RCS does not depend on the aspect)."""
    a, f = 0.2, 300e6
    out = bo.solve_bor_dielectric(bo.sphere_nodes(a, n), f, [0.0, 60.0, 120.0, 180.0], eps, mu, n_modes=6)
    ref = 10.0 * math.log10(mie.sigma_dielectric_sphere(a, eps, mu, f))
    for key in ("sigma_vv", "sigma_hh"):
        db = 10.0 * np.log10(np.array(out[key]))
        assert np.max(np.abs(db - ref)) < tol_db, (key, db, ref)


def test_dielectric_mie_limits():
    """The Mie series of the oracle: eps -> 1, mu -> 1 scatters nothing, and
    in the Rayleigh limit sigma -> 4 pi k^4 a^6 |(eps - 1)/(eps + 2)|^2."""
    assert mie.sigma_dielectric_sphere(0.2, 1.0, 1.0, 300e6) < 1e-30
    a, f, eps = 0.01, 100e6, 3.0 - 0.2j
    k = 2.0 * math.pi * f / bo.C0
    ray = 4.0 * math.pi * k ** 4 * a ** 6 * abs((eps - 1.0) / (eps + 2.0)) ** 2
    assert abs(mie.sigma_dielectric_sphere(a, eps, 1.0, f) / ray - 1.0) < 1e-3
