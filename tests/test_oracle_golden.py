"""The FP64 oracle against the reference's own values (golden fixtures written
by tests/golden/make_golden.py), closed forms and invariants."""

import math
import os

import numpy as np
import pytest
from scipy import integrate, special

import bor_oracle as bo
import mie

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    return np.load(os.path.join(GOLD, name + ".npz"))


def rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


@pytest.mark.parametrize("pair", [(0, 0), (0, 1), (1, 0), (5, 5), (5, 6), (11, 11)])
def test_near_quadrature_matches_reference(pair):
    d = load("near_sphere12")
    body = bo.Body(d["nodes"])
    k = 2 * math.pi * float(d["freq_hz"]) / bo.C0
    mm, st = int(d["m_max"]), int(d["stride"])
    e, f = pair
    s, sp, w = bo.cell_points(bo.near_kind(e, f))
    p, q = body.element_points(e, s), body.element_points(f, sp)
    G = bo.green_orders_near(p["rho"], p["z"], q["rho"], q["z"], k, mm)[::st]
    assert rel(G, d[f"G_{e}_{f}"]) < 1e-13
    K = bo.bracket_modes_near(p["rho"], p["z"], p["trho"], p["tz"], q["rho"], q["z"], q["trho"], q["tz"], k, mm)
    for uv in range(4):
        assert rel(K[uv][::st], d[f"K{uv}_{e}_{f}"]) < 1e-13


@pytest.mark.parametrize("name", ["sphere12_cfie", "sphere12_efie", "conesphere16_cfie", "sphere10_stream"])
def test_end_to_end_matches_reference(name):
    d = load(name)
    form = "cfie" if float(d["alpha"]) < 1 else "efie"
    out = bo.solve_bor(d["nodes"], float(d["freq_hz"]), d["thetas"], formulation=form,
                       n_modes=int(d["m_max"]), keep=True, streaming="stream" in name)
    for m in range(-int(d["m_max"]), int(d["m_max"]) + 1):
        X, B, A, mask = out["_modes"][m]
        assert np.array_equal(mask, d[f"mask_{m}"])
        assert rel(A, d[f"A_{m}"]) < 1e-13
        assert rel(B, d[f"B_{m}"]) < 1e-13
        assert rel(X, d[f"X_{m}"]) < 1e-10
    assert out["modes_used"] == int(d["modes_used"])
    db = lambda s: 10 * np.log10(np.asarray(s))
    assert np.max(np.abs(db(out["sigma_vv"]) - db(d["sigma_vv"]))) < 1e-10
    assert np.max(np.abs(db(out["sigma_hh"]) - db(d["sigma_hh"]))) < 1e-10


def test_mie_matches_reference():
    d = load("mie_pec")
    got = np.array([mie.sigma_pec_sphere(float(d["radius"]), f) for f in d["freq_hz"]])
    assert rel(got, d["sigma"]) < 1e-12


def test_mie_limits():
    a = 0.5
    # Rayleigh: 9 pi k^4 a^6
    f = 1e6
    k = 2 * math.pi * f / mie.C0
    assert abs(mie.sigma_pec_sphere(a, f) / (9 * math.pi * k ** 4 * a ** 6) - 1) < 1e-3
    # geometric optics: sigma -> pi a^2
    assert abs(mie.sigma_pec_sphere(a, 60e9) / (math.pi * a * a) - 1) < 2e-2


def test_gauss_exactness():
    body = bo.Body(bo.sphere_nodes(1.0, 5))
    g = body.g
    # sum of weights per element = length; 4-point rule integrates s^7 exactly
    for e in range(body.Ne):
        sl = slice(4 * e, 4 * e + 4)
        assert abs(np.sum(g.w[sl]) - body.L[e]) < 1e-14
        assert abs(np.sum(g.w[sl] * g.s[sl] ** 7) - body.L[e] / 8) < 1e-14


def test_far_orders_brute_force_and_static_limit():
    rho_p, z_p, rho_q, z_q = 0.4, 0.1, 0.3, -0.25
    k, mm = 5.0, 6
    n = bo.n_xi_efie(2 * k * math.sqrt(rho_p * rho_q), mm)
    G = bo.green_orders_fft(bo.green_samples(np.array([rho_p]), np.array([z_p]), np.array([rho_q]),
                                             np.array([z_q]), k, bo.xi_grid(n)), mm + 2)[0]
    for m in range(mm + 2):
        f = lambda xi, part: getattr(np.exp(-1j * k * math.sqrt((rho_p - rho_q) ** 2 + (z_p - z_q) ** 2
                                                                 + 4 * rho_p * rho_q * math.sin(xi / 2) ** 2))
                                     / (4 * math.pi * math.sqrt((rho_p - rho_q) ** 2 + (z_p - z_q) ** 2
                                                                + 4 * rho_p * rho_q * math.sin(xi / 2) ** 2))
                                     * np.exp(-1j * m * xi), part)
        re = integrate.quad(f, -math.pi, math.pi, args=("real",), epsabs=1e-14, limit=200)[0]
        im = integrate.quad(f, -math.pi, math.pi, args=("imag",), epsabs=1e-14, limit=200)[0]
        assert abs(G[m] - (re + 1j * im)) < 1e-10 * abs(G[0])
    # static limit k -> 0: G_0 = 4 K(kappa) / (4 pi sqrt((rho+rho')^2 + dz^2))
    G0 = bo.green_orders_fft(bo.green_samples(np.array([rho_p]), np.array([z_p]), np.array([rho_q]),
                                              np.array([z_q]), 0.0, bo.xi_grid(256)), 1)[0, 0]
    den = (rho_p + rho_q) ** 2 + (z_p - z_q) ** 2
    kap2 = 4 * rho_p * rho_q / den
    assert abs(G0 - 4 * special.ellipk(kap2) / (4 * math.pi * math.sqrt(den))) < 1e-12


def test_bracket_modes_brute_dft():
    args = (np.array([0.4]), np.array([0.1]), np.array([0.6]), np.array([0.8]),
            np.array([0.3]), np.array([-0.25]), np.array([-0.2]), np.array([0.98]))
    k, mm, n = 4.0, 5, 128
    xi = bo.xi_grid(n)
    F = bo.bracket_samples(*args, k, xi)
    K = bo.bracket_modes_fft(F[1], mm)[0]
    for i, m in enumerate(range(-mm, mm + 1)):
        direct = (2 * math.pi / n) * np.sum(F[1][0] * np.exp(-1j * m * xi))
        assert abs(K[i] - direct) < 1e-12 * np.max(np.abs(K))
    # parity: tt, ff even in xi; tf, ft odd
    Fm = bo.bracket_samples(*args, k, -xi[1:])
    Fp = bo.bracket_samples(*args, k, xi[1:])
    assert np.allclose(Fm[0], Fp[0], rtol=1e-13, atol=0) and np.allclose(Fm[3], Fp[3], rtol=1e-13, atol=0)
    assert np.allclose(Fm[1], -Fp[1], rtol=1e-13, atol=0) and np.allclose(Fm[2], -Fp[2], rtol=1e-13, atol=0)


def test_streamed_contraction_equals_table_contraction():
    body = bo.Body(bo.sphere_nodes(0.5, 10))
    k, mm = 2 * math.pi * 3e8 / bo.C0, 3
    osc = bo.osc_for(body, k, streaming=False)
    ne, nbk = bo.n_xi_efie(osc, mm), bo.n_xi_bracket(osc, mm)
    far = bo.FarBlocks(body, k, mm, ne, nbk, efie=True, mfie=False, rows_per_tile=3)
    G = bo.far_green_table(body, k, mm, ne)
    for m in (-3, -1, 0, 1, 2, 3):
        for a, b in zip(far.efie(m), bo.table_efie_blocks(body, k, m, G)):
            assert rel(a, b) < 1e-12


def test_sphere_invariants():
    out = bo.solve_bor(bo.sphere_nodes(0.5, 24), 3e8, [0, 40, 90, 135, 180], formulation="cfie",
                       n_modes=6, mode_tol=0, keep=True)
    vv, hh = np.array(out["sigma_vv"]), np.array(out["sigma_hh"])
    # axial incidence: only |m| = 1 contributes
    for m in (0, 2, -2, 3):
        assert np.max(np.abs(out["_dF"][abs(m)][:, [0, -1]])) < 1e-12 * np.max(np.abs(out["_dF"][1]))
    # sphere: sigma independent of aspect, VV = HH to discretisation level
    assert np.ptp(10 * np.log10(vv)) < 0.05 and np.ptp(10 * np.log10(hh)) < 0.05
    assert out["linear_residual"] < 1e-13


@pytest.mark.parametrize("name", ["sphere12_ibc", "conesphere16_ibc_taper", "cup10_efie_open"])
def test_ibc_and_open_shell_match_reference(name):
    """IBC-EFIE systems (rotated-PV brackets with Z_s at the source, the Z_s
    mass), currents and the far field with the radiating magnetic current,
    for a scalar and a tapered impedance; and the open-end basis mask."""
    d = load(name)
    mm = int(d["m_max"])
    zs = d["zs"] if "zs" in d else None
    sysm = bo.ModeSystem(d["nodes"], float(d["freq_hz"]), mm, "efie", 1.0, streaming=False, zs=zs)
    for m in range(-mm, mm + 1):
        X, B, A, mask, _ = sysm.solve_mode(m, d["thetas"])
        assert np.array_equal(mask, d[f"mask_{m}"])
        assert rel(A, d[f"A_{m}"]) < 1e-13
        assert rel(B, d[f"B_{m}"]) < 1e-13
        assert rel(X, d[f"X_{m}"]) < 1e-10
        assert rel(sysm.mode_farfield(m, X, mask, d["thetas"]), d[f"dF_{m}"]) < 1e-10


def test_impedance_mie_closed_forms():
    """Z_s = 0 gives the PEC coefficients; Z_s = eta0 gives A_n = B_n
    (Weston: zero backscatter) for every order."""
    x = 2.7
    a0, b0 = mie.pec_sphere_coeffs(x, 20)
    a, b = mie.impedance_sphere_coeffs(x, 0.0, 20)
    assert np.allclose(a, a0, rtol=1e-14, atol=0) and np.allclose(b, b0, rtol=1e-14, atol=0)
    a, b = mie.impedance_sphere_coeffs(x, mie.ETA0, 20)
    assert np.max(np.abs(a - b)) < 1e-14 * np.max(np.abs(a))
    assert mie.sigma_impedance_sphere(0.5, 3e8, mie.ETA0) < 1e-25


def test_ibc_oracle_against_mie_and_weston():
    """End to end on a 20-segment sphere: a lossy Z_s within 0.15 dB of the
    impedance Mie series, and Z_s = eta0 at least 35 dB below the PEC sphere
    (the discretised Weston null); the sign of the magnetic-current far field
    or of the rotated-PV operator flips either check."""
    nodes = bo.sphere_nodes(0.5, 20)
    zs = complex(0.3 * bo.ETA0, -0.2 * bo.ETA0)
    th = [0.0, 90.0]
    out = bo.solve_bor(nodes, 3e8, th, formulation="efie", n_modes=7, mode_tol=0.0, zs=zs)
    ref = 10 * np.log10(mie.sigma_impedance_sphere(0.5, 3e8, zs))
    assert abs(10 * np.log10(out["sigma_vv"][0]) - ref) < 0.15
    w = bo.solve_bor(nodes, 3e8, th, formulation="efie", n_modes=7, mode_tol=0.0, zs=bo.ETA0)
    pec = 10 * np.log10(mie.sigma_pec_sphere(0.5, 3e8))
    assert 10 * np.log10(max(w["sigma_vv"])) < pec - 35
    assert 10 * np.log10(max(w["sigma_hh"])) < pec - 35


def test_cfg1_sphere120_end_to_end_against_mie():
    """The reference's 120-segment sphere result (fixture) sits within
    0.05 dB of the Mie series, so the GPU's agreement with it is also
    agreement with the analytic oracle at that discretisation."""
    d = load("sphere120_cfg1")
    s_mie = 10 * np.log10(mie.sigma_pec_sphere(0.5, float(d["freq_hz"])))
    for form in ("cfie", "efie"):
        for pol in ("sigma_vv", "sigma_hh"):
            err = np.max(np.abs(10 * np.log10(d[f"{form}_{pol}"]) - s_mie))
            assert err < 0.05, (form, pol, err)


def test_row_tile_oracle_equals_full_oracle():
    """The row-tile restriction that writes the 1200-element ogive rows
    (tests/golden/make_oracle_cfg2_rows.py) gives the full oracle's rows on
    a small body."""
    import sys
    sys.path.insert(0, GOLD)
    import make_oracle_cfg2_rows as mr
    nodes = bo.sphere_nodes(0.5, 16)
    f, mm, E = 4e8, 5, 7
    body = bo.Body(nodes)
    k = 2 * math.pi * f / bo.C0
    sysm = bo.ModeSystem(nodes, f, mm, "cfie", 0.5, streaming=False)
    far = mr.RowFar(body, k, mm, sysm.n_e, sysm.n_b, E - 1)
    near = mr.RowNear(body, k, mm, E - 1)
    C = 1j * k * bo.ETA0 * 2.0 * np.pi
    Ml = sysm.M[E - 1:E + 3]
    for m in (-2, 0, 3):
        eb, kb = list(far.efie(m)), list(far.mfie(m))
        for (e, ff), blk in near.efie.items():
            for b in range(4):
                eb[b][e - E + 1:e - E + 3, ff:ff + 2] += blk[b, m + mm]
        for (e, ff), blk in near.mfie.items():
            for b in range(4):
                kb[b][e - E + 1:e - E + 3, ff:ff + 2] += blk[b, m + mm]
        Z = 0.5 * C * np.block([[eb[0], eb[1]], [eb[2], eb[3]]]) + 0.5 * bo.ETA0 * np.block(
            [[0.5 * Ml - kb[0], -kb[1]], [-kb[2], 0.5 * Ml - kb[3]]])
        full = 0.5 * sysm.efie_full(m) + 0.5 * bo.ETA0 * sysm.mfie_full(m)
        Nn = body.Nn
        assert rel(Z[[1, 2, 5, 6]], full[[E, E + 1, Nn + E, Nn + E + 1]]) < 1e-13
