"""Aspects beyond [0, 180] degrees, and the largest azimuth grids the FP64
far fill accepts (run on an MI355X)."""

import math
import os

import numpy as np
import pytest

import bor_oracle as bo
from conftest import pkg

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def test_aspects_with_negative_sine_match_oracle(bs):
    """sin(theta) < 0 (theta in (180, 360) or below 0): the Bessel table
    holds J_n of the signed argument, J_n(-x) = (-1)^n J_n(x), as the
    reference's scipy.special.jv does (bor_solver.py:632, 679).  By the
    body's rotational symmetry sigma(360 - theta) = sigma(theta) =
    sigma(-theta)."""
    nodes = bo.sphere_nodes(0.5, 12)
    thetas = np.array([30.0, -30.0, 330.0, 150.0, 210.0, -150.0])
    print("[case] signed-sine aspects, sphere 12, cfie", flush=True)
    out = bs.solve_bor(nodes, 3e8, thetas, formulation="cfie", n_modes=8, mode_tol=0.0)
    ref = bo.solve_bor(nodes, 3e8, thetas, formulation="cfie", n_modes=8, mode_tol=0.0)
    for key in ("amp_vv", "amp_hh"):
        assert rel(np.array(out[key]), np.array(ref[key])) < 1e-9, key
    for key in ("sigma_vv", "sigma_hh"):
        s = np.asarray(out[key])
        assert abs(s[1] / s[0] - 1) < 1e-9 and abs(s[2] / s[0] - 1) < 1e-9, (key, s)
        assert abs(s[4] / s[3] - 1) < 1e-9 and abs(s[5] / s[3] - 1) < 1e-9, (key, s)


def test_fp64_far_fill_with_2048_sample_grids_at_m126():
    """Eight 16-order tiles with 2048-sample azimuth grids: the FP64 tables
    leave room for one wavefront's order stash only; the blocks equal the
    oracle's (computed on 512 / 1024 samples, whose aliasing error at these
    orders is far below 1e-11 for |e - f| >= 2)."""
    L = pkg("_lib")
    bs = pkg("bor_solver")
    lib = L.load()
    assert lib.ghost_far_fill_waves(L.GHOST_FP64, 126, 2048, 2048) == 1
    d = np.load(os.path.join(GOLD, "oracle_sphere8_high_orders.npz"))
    mm = 126
    nodes, f, alpha = d["nodes"], float(d["freq_hz"]), float(d["alpha"])
    k = 2 * math.pi * f / bs.C0
    print("[case] FP64 far fill m_max=126 n_xi=2048/2048", flush=True)
    h = L.Handle(L.GHOST_FP64)
    h.add_problem(nodes, k, mm, 2048, 2048, alpha, (1 - alpha) * bs.ETA0, [0.0, 90.0], unmasked=True)
    h.build()
    h.fill()
    h.sync()
    for m in d[f"m_{mm}"]:
        A, _ = h.system(0, int(m))
        err = rel(A, d[f"A_{mm}_{m}"])
        assert err < 1e-11, (m, err)
    h.close()


def test_fp64_grids_beyond_lds_are_rejected_at_add():
    L = pkg("_lib")
    h = L.Handle(L.GHOST_FP64)
    nodes = bo.sphere_nodes(0.5, 8)
    with pytest.raises(ValueError, match="LDS"):
        h.add_problem(nodes, 6.0, 126, 4096, 4096, 1.0, 0.0, [0.0])
    h.close()
