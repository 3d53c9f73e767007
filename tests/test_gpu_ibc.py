"""IBC-EFIE (surface impedance) and open generatrices on the HIP path against
the reference's golden vectors, the FP64 oracle and the impedance Mie series
(run on an MI355X)."""

import math
import os

import numpy as np
import pytest

import bor_oracle as bo
import mie
from conftest import pkg

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
CASES = ["sphere12_ibc", "conesphere16_ibc_taper", "cup10_efie_open"]


def load(name):
    return np.load(os.path.join(GOLD, name + ".npz"))


def rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def db(s):
    return 10.0 * np.log10(np.asarray(s, dtype=float))


def handle(d, precision=0, pairing=True, unmasked=False):
    L = pkg("_lib")
    bs = pkg("bor_solver")
    nodes = d["nodes"]
    k = 2 * math.pi * float(d["freq_hz"]) / bs.C0
    mm = int(d["m_max"])
    zs = d["zs"] if "zs" in d else None
    _, ne, nbk, _ = bs.discretisation(nodes, k, d["thetas"], mm, "efie", "tables", has_ibc=zs is not None)
    h = L.Handle(precision, pairing=pairing)
    h.add_problem(nodes, k, mm, ne, nbk, 1.0, 0.0, d["thetas"], unmasked=unmasked)
    if zs is not None:
        h.set_impedance(0, zs, 1.0)
    h.build()
    return h


@pytest.mark.parametrize("pairing", [True, False])
@pytest.mark.parametrize("name", CASES)
def test_systems_currents_farfield_match_reference(name, pairing):
    d = load(name)
    mm = int(d["m_max"])
    print(f"[case] {name} pairing={pairing}", flush=True)
    h = handle(d, pairing=pairing)
    h.fill()
    h.sync()
    for m in range(-mm, mm + 1):
        A, B = h.system(0, m)
        assert rel(A, d[f"A_{m}"]) < 1e-11, (m, rel(A, d[f"A_{m}"]))
        assert rel(B, d[f"B_{m}"]) < 1e-12, (m, rel(B, d[f"B_{m}"]))
    h.solve()
    h.farfield()
    h.sync()
    dF = h.dF(0)
    for m in range(-mm, mm + 1):
        X = h.solution(0, m)
        assert rel(X, d[f"X_{m}"]) < 1e-9, (m, rel(X, d[f"X_{m}"]))
        assert rel(dF[m + mm], d[f"dF_{m}"]) < 1e-9, (m, rel(dF[m + mm], d[f"dF_{m}"]))
    assert np.max(h.residuals()) < 1e-12
    h.close()


@pytest.mark.parametrize("name", CASES)
def test_solve_bor_matches_reference(name, bs):
    d = load(name)
    zs = d["zs"] if "zs" in d else None
    out = bs.solve_bor(d["nodes"], float(d["freq_hz"]), d["thetas"], formulation="efie", zs=zs,
                       n_modes=int(d["m_max"]), mode_tol=1e-6)
    assert out["modes_used"] == int(d["modes_used"])
    assert np.max(np.abs(db(out["sigma_vv"]) - db(d["sigma_vv"]))) < 1e-7
    assert np.max(np.abs(db(out["sigma_hh"]) - db(d["sigma_hh"]))) < 1e-7
    assert rel(np.array(out["amp_vv"]), d["amp_vv"]) < 1e-9
    assert rel(np.array(out["amp_hh"]), d["amp_hh"]) < 1e-9


def test_mixed_precision_ibc_within_0p01_db(bs):
    d = load("conesphere16_ibc_taper")
    th = np.linspace(0, 180, 19)
    kw = dict(formulation="efie", zs=d["zs"], n_modes=int(d["m_max"]), mode_tol=0.0)
    a = bs.solve_bor(d["nodes"], float(d["freq_hz"]), th, **kw)
    b = bs.solve_bor(d["nodes"], float(d["freq_hz"]), th, precision="mixed", refine_iters=-3, **kw)
    assert np.max(np.abs(db(a["sigma_vv"]) - db(b["sigma_vv"]))) < 0.01
    assert np.max(np.abs(db(a["sigma_hh"]) - db(b["sigma_hh"]))) < 0.01


def test_operator_factory_ibc_blocks_and_farfield_match_oracle(bs):
    """BorPecSolver.assemble_ibc_extra and farfield_mode(zs_pt=...) against
    the oracle's unmasked impedance blocks and magnetic-current far field."""
    nodes = bo.sphere_nodes(0.45, 14)
    f, mm = 3.2e8, 4
    ze = (0.2 + 0.05 * np.arange(14)) * bo.ETA0 - 30j
    sysm = bo.ModeSystem(nodes, f, mm, "efie", 1.0, streaming=False, zs=ze)
    sv = bs.BorPecSolver(nodes, f)
    zpt = np.repeat(ze, 4)
    for m in (-3, 0, 1, 4):
        got = sv.assemble_ibc_extra(m, mm, zpt, ze)
        assert rel(got, sysm.ibc_full(m)) < 1e-11, m
    rng = np.random.default_rng(4)
    sol = rng.standard_normal(2 * len(nodes)) + 1j * rng.standard_normal(2 * len(nodes))
    for m, th in ((0, 20.0), (2, 75.0), (-3, 140.0)):
        got = sv.farfield_mode(m, sol, th, zs_pt=zpt)
        ref = bo.farfield(sysm.body, sysm.k, m, sol, th, zpt)
        assert abs(got[0] - ref[0]) < 1e-11 * abs(ref[0]) and abs(got[1] - ref[1]) < 1e-11 * abs(ref[1])


def test_weston_null_and_lossy_mie(bs):
    """Z_s = eta0 on a sphere: the backscatter falls at least 35 dB below the
    PEC sphere's (zero in the continuum); a lossy Z_s agrees with the
    impedance Mie series within 0.15 dB and with the oracle within 1e-8 dB."""
    nodes = bo.sphere_nodes(0.5, 20)
    th = [0.0, 45.0, 90.0]
    pec = db(mie.sigma_pec_sphere(0.5, 3e8))
    w = bs.solve_bor(nodes, 3e8, th, formulation="efie", zs=bo.ETA0, n_modes=7, mode_tol=0.0)
    assert np.max(db(w["sigma_vv"])) < pec - 35 and np.max(db(w["sigma_hh"])) < pec - 35
    zs = complex(0.3 * bo.ETA0, -0.2 * bo.ETA0)
    out = bs.solve_bor(nodes, 3e8, th, formulation="efie", zs=zs, n_modes=7, mode_tol=0.0)
    ref = bo.solve_bor(nodes, 3e8, th, formulation="efie", zs=zs, n_modes=7, mode_tol=0.0)
    assert np.max(np.abs(db(out["sigma_vv"]) - db(ref["sigma_vv"]))) < 1e-8
    assert abs(db(out["sigma_vv"][0]) - db(mie.sigma_impedance_sphere(0.5, 3e8, zs))) < 0.15


def test_cfie_with_impedance_is_rejected(bs):
    with pytest.raises(ValueError, match="efie"):
        bs.solve_bor(bo.sphere_nodes(0.5, 8), 3e8, [0.0], formulation="cfie", zs=100.0)
