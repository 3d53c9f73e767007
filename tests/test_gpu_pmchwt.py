"""Homogeneous penetrable bodies (PMCHWT) on the HIP path against the
reference's golden vectors (tests/golden/make_golden_pmchwt.py), the FP64
oracle and the dielectric Mie series (run on an MI355X)."""

import math
import os

import numpy as np
import pytest

from conftest import pkg

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
CASES = ["sphere12_pmchwt", "conesphere16_pmchwt"]


def load(name):
    return np.load(os.path.join(GOLD, name + ".npz"))


def rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


@pytest.mark.parametrize("name", CASES)
def test_medium_operators_match_reference(name):
    """The interior medium's EFIE operator T (complex k, complex eta) and the
    rotated-PV coupling operators P of both media, unmasked, every mode."""
    d = load(name)
    bs = pkg("bor_solver")
    nm = int(d["m_max"])
    print(f"[case] {name} medium operators", flush=True)
    se = bs.BorPecSolver(d["nodes"], float(d["freq_hz"]))
    si = bs.BorPecSolver(d["nodes"], float(d["freq_hz"]), medium=(complex(d["eps_r"]), complex(d["mu_r"])))
    assert abs(si.k - complex(d["k_in"])) < 1e-13 * abs(si.k)
    assert abs(si.eta - complex(d["eta_in"])) < 1e-13 * abs(si.eta)
    for m in range(-nm, nm + 1):
        assert rel(si.assemble_mode(m, nm), d[f"Ti_{m}"]) < 1e-11, m
        assert rel(si.assemble_pmchwt_P(m, nm), d[f"Pi_{m}"]) < 1e-11, m
        assert rel(se.assemble_pmchwt_P(m, nm), d[f"Pe_{m}"]) < 1e-11, m


def pen_handle(d, pairing=True):
    L = pkg("_lib")
    bs = pkg("bor_solver")
    nodes, th = d["nodes"], d["thetas"]
    k0 = 2 * math.pi * float(d["freq_hz"]) / bs.C0
    nm = int(d["m_max"])
    _, ne, nbk, _ = bs.discretisation(nodes, k0, th, nm, "efie", "tables")
    kin = complex(d["k_in"])
    osc = 2.0 * abs(kin) * bs.gauss_rho_max(nodes)
    h = L.Handle(L.GHOST_FP64, pairing=pairing)
    h.add_problem(nodes, k0, nm, ne, nbk, 1.0, 0.0, th)
    h.set_interior(0, kin, complex(d["eta_in"]), bs.n_xi_efie(osc, nm), bs.n_xi_bracket(osc, nm))
    h.build()
    return h


@pytest.mark.parametrize("pairing", [True, False])
@pytest.mark.parametrize("name", CASES)
def test_pmchwt_systems_currents_farfield_match_reference(name, pairing):
    """The masked 2 x 2 block systems and right-hand sides as assembled on the
    device, then the currents (J, M / eta0) and each mode's far field."""
    d = load(name)
    nm = int(d["m_max"])
    print(f"[case] {name} PMCHWT pairing={pairing}", flush=True)
    h = pen_handle(d, pairing)
    try:
        h.fill()
        h.sync()
        for m in range(-nm, nm + 1):
            A, B = h.system(0, m)
            assert A.shape == d[f"A_{m}"].shape, m
            assert rel(A, d[f"A_{m}"]) < 1e-11, (m, rel(A, d[f"A_{m}"]))
            assert rel(B, d[f"B_{m}"]) < 1e-13, (m, rel(B, d[f"B_{m}"]))
        h.solve(0)
        h.farfield()
        h.sync()
        dF = h.dF(0)
        res = h.residuals()
        assert np.max(res) < 1e-12, res
        for m in range(-nm, nm + 1):
            X = h.solution(0, m)
            assert rel(X, d[f"X_{m}"]) < 1e-9, (m, rel(X, d[f"X_{m}"]))
            assert rel(dF[m + nm], d[f"dF_{m}"]) < 1e-9, (m, rel(dF[m + nm], d[f"dF_{m}"]))
    finally:
        h.close()


@pytest.mark.parametrize("name", CASES)
def test_solve_bor_dielectric_matches_reference(name):
    d = load(name)
    bs = pkg("bor_solver")
    print(f"[case] {name} solve_bor_dielectric", flush=True)
    out = bs.solve_bor_dielectric(d["nodes"], float(d["freq_hz"]), d["thetas"], complex(d["eps_r"]),
                                  complex(d["mu_r"]), n_modes=int(d["m_max"]), mode_tol=1e-6)
    for key in ("amp_vv", "amp_hh"):
        assert rel(np.array(out[key]), d[key]) < 1e-9, (key, rel(np.array(out[key]), d[key]))
    for key in ("sigma_vv", "sigma_hh"):
        assert rel(np.array(out[key]), d[key]) < 2e-9, key
    assert out["modes_used"] == int(d["modes_used"])
    assert out["formulation"] == "pmchwt" and out["n_unknowns"] == 4 * len(d["nodes"])
    assert out["linear_residual"] < 1e-12


def test_farfield_mode_with_magnetic_current_matches_reference():
    """BorPecSolver.farfield_mode(msol=M) on the reference's currents."""
    d = load("conesphere16_pmchwt")
    bs = pkg("bor_solver")
    nodes = d["nodes"]
    Nn = len(nodes)
    se = bs.BorPecSolver(nodes, float(d["freq_hz"]))
    for m in (-2, 0, 1, 3):
        X, mask = d[f"X_{m}"], d[f"mask_{m}"]
        for it in (0, 3, 6):
            th = float(d["thetas"][it])
            for ip in range(2):
                full = np.zeros(4 * Nn, dtype=complex)
                full[mask] = X[:, 2 * it + ip]
                fth, fph = se.farfield_mode(m, full[:2 * Nn], th, msol=bs.ETA0 * full[2 * Nn:])
                ref = d[f"dF_{m}"][ip, it]
                got = fth if ip == 0 else fph
                assert abs(got - ref) <= 1e-9 * np.max(np.abs(d[f"dF_{m}"])), (m, it, ip)


@pytest.mark.parametrize("eps, mu", [(4.0, 1.0), (3.0 - 0.4j, 1.3 - 0.05j)])
def test_dielectric_sphere_against_oracle_and_mie(eps, mu):
    """0.2 m sphere at 300 MHz, 48 segments: the device result equals the
    FP64 oracle and lies within the discretisation error of the Mie series."""
    import bor_oracle as bo
    import mie
    bs = pkg("bor_solver")
    a, f = 0.2, 300e6
    nodes = bs.sphere_generatrix(a, 48)
    th = [0.0, 45.0, 90.0, 135.0, 180.0]
    print(f"[case] dielectric sphere eps={eps} mu={mu}", flush=True)
    out = bs.solve_bor_dielectric(nodes, f, th, eps, mu, n_modes=6, mode_tol=0.0)
    ref = bo.solve_bor_dielectric(nodes, f, th, eps, mu, n_modes=6, mode_tol=0.0)
    for key in ("amp_vv", "amp_hh"):
        assert rel(np.array(out[key]), np.array(ref[key])) < 1e-9, key
    mie_db = 10.0 * math.log10(mie.sigma_dielectric_sphere(a, eps, mu, f))
    for key in ("sigma_vv", "sigma_hh"):
        db = 10.0 * np.log10(np.array(out[key]))
        assert np.max(np.abs(db - mie_db)) < 0.1, (key, db, mie_db)
