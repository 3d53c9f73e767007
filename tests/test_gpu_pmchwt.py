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
