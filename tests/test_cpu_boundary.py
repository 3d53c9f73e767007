"""CPU-side checks of the C-ABI boundary: the library loads, exports every
symbol include/ghost_bor.h declares, and the host logic is consistent."""

import ctypes
import math
import os
import re

import numpy as np
import pytest

from conftest import ROOT, pkg

HEADER = os.path.join(ROOT, "include", "ghost_bor.h")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"^\s*#.*$", "", text, flags=re.M)
    return sorted(set(re.findall(r"\b(\w+)\s*\([^;{]*\)\s*;", text)))


def test_library_exports_every_declared_symbol():
    L = pkg("_lib")
    lib = L.load()
    names = declared_symbols()
    assert len(names) >= 20
    for n in names:
        assert hasattr(lib, n), n
    bound = {n for n, _, _ in L.SIGNATURES}
    assert set(names) == bound


def test_missing_library_fails_loudly(tmp_path):
    L = pkg("_lib")
    with pytest.raises(L.GhostLibraryError):
        L._LIB_saved = L._LIB
        try:
            L._LIB = None
            L.load(str(tmp_path / "nope.so"))
        finally:
            L._LIB = L._LIB_saved


def test_discretisation_rules_match_oracle():
    import bor_oracle as bo
    bs = pkg("bor_solver")
    for nodes, f in ((bs.sphere_generatrix(0.5, 120), 3e8), (bs.sphere_generatrix(2.0, 40), 9e8)):
        k = 2 * math.pi * f / bs.C0
        body = bo.Body(nodes)
        th = np.arange(0, 181.0)
        m, ne, nb, st = bs.discretisation(nodes, k, th, None, "cfie")
        assert m == bo.default_m_max(body, k, th)
        osc = bo.osc_for(body, k, st)
        assert (ne, nb) == (bo.n_xi_efie(osc, m), bo.n_xi_bracket(osc, m))
        assert st == (bo.table_gb(body.Ne, m, "cfie") > 2.0)


def test_truncation_rule_matches_oracle():
    import bor_oracle as bo
    bs = pkg("bor_solver")
    rng = np.random.default_rng(1)
    mm = 8
    dF = rng.standard_normal((2 * mm + 1, 2, 5)) * (10.0 ** -np.abs(np.arange(-mm, mm + 1)))[:, None, None]
    dF = dF.astype(complex)
    F, used = bs.truncate_modes(dF, mm, 1e-6)
    per_am = [dF[mm]] + [dF[mm + a] + dF[mm - a] for a in range(1, mm + 1)]
    F2, used2 = bo.truncate_modes(per_am, 1e-6)
    assert used == used2 and np.allclose(F, F2, rtol=0, atol=0)


def test_workload_generatrices_are_valid():
    wl = pkg("workloads")
    for idx in (1, 2, 4):
        name, nodes, f, th, mm = wl.config(idx)
        assert nodes[0, 0] < 1e-12 and nodes[-1, 0] < 1e-12
        assert nodes[0, 1] > nodes[-1, 1]
        seg = np.hypot(*np.diff(nodes, axis=0).T)
        assert np.all(seg > 0)
    _, n2, f2, th2, m2 = wl.config(2)
    assert len(n2) - 1 == 1200 and m2 == 70 and len(th2) == 361
    assert abs(np.max(n2[:, 0]) - 1.0) < 0.05
    assert abs((n2[0, 1] - n2[-1, 1]) - 10.0) < 1e-9


def test_reference_c_sampler_agrees_with_oracle():
    """oracle/_ref (the reference's own C sampler, built from its source) pins
    the oracle's integrand restatement when it is present."""
    import bor_oracle as bo
    so = os.path.join(ROOT, "oracle", "_ref", "bor_stream_kernel_ref.so")
    if not os.path.exists(so):
        pytest.skip("oracle/_ref not built (needs the reference sources)")
    lib = ctypes.CDLL(so)
    dp = ctypes.POINTER(ctypes.c_double)
    rng = np.random.default_rng(5)
    nr, np_, nxi = 3, 4, 32
    rp, zp, rq, zq = (rng.uniform(0.05, 1, nr), rng.uniform(-1, 1, nr), rng.uniform(0.05, 1, np_),
                      rng.uniform(-1, 1, np_))
    k = 6.1
    xi = bo.xi_grid(nxi)
    sin2 = np.sin(0.5 * xi) ** 2
    out = np.zeros((nr, np_, nxi), dtype=np.complex128)
    P = lambda a: np.ascontiguousarray(a).ctypes.data_as(dp)
    lib.sample_g(nr, np_, nxi, P(rp), P(zp), P(rq), P(zq), ctypes.c_double(k), P(sin2), out.ctypes.data_as(dp))
    ref = bo.green_samples(rp[:, None], zp[:, None], rq[None, :], zq[None, :], k, xi, r_min=1e-30)
    assert np.max(np.abs(out - ref)) < 1e-13 * np.max(np.abs(ref))
    trp, tzp, trq, tzq = (rng.uniform(-1, 1, nr), rng.uniform(-1, 1, nr), rng.uniform(-1, 1, np_),
                          rng.uniform(-1, 1, np_))
    cx, sx = np.cos(xi), np.sin(xi)
    for fname, refn in (("sample_mfie", bo.bracket_samples), ("sample_ibc", bo.ibc_bracket_samples)):
        outs = [np.zeros((nr, np_, nxi), dtype=np.complex128) for _ in range(4)]
        getattr(lib, fname)(nr, np_, nxi, P(rp), P(zp), P(trp), P(tzp), P(rq), P(zq), P(trq), P(tzq),
                            ctypes.c_double(k), P(cx), P(sx), *[o.ctypes.data_as(dp) for o in outs])
        refs = refn(rp[:, None], zp[:, None], trp[:, None], tzp[:, None], rq[None, :], zq[None, :],
                    trq[None, :], tzq[None, :], k, xi, r_min=1e-30)
        for o, r in zip(outs, refs):
            assert np.max(np.abs(o - r)) < 1e-12 * np.max(np.abs(r))


def test_far_fill_lds_budget_covers_accepted_problems():
    """Every (precision, m_max, n_xi_efie, n_xi_bracket) that add_problem
    accepts gets a far-fill launch whose azimuth tables and order stash fit
    the 160 KiB of LDS; FP64 at m_max = 126 with 2048-sample grids (a default
    discretisation for k rho_max ~ 110) must fit with one wavefront per group."""
    lib = pkg("_lib").load()
    lds = 160 * 1024
    grids = [64, 128, 256, 512, 1024, 2048, 4096]
    for prec, rb in ((0, 8), (1, 4)):
        for m_max in range(0, 127, 5):
            nt = (m_max + 2 + 15) // 16
            for ne in grids:
                for nb in grids:
                    w = lib.ghost_far_fill_waves(prec, m_max, ne, nb)
                    assert w in (0, 1, 2, 4)
                    tab = rb * (2 * nb + 4 * (nb // 2 + 4) + 2 * (ne // 2 + 4) + ne)
                    stash = rb * max(w, 1) * nt * 8 * 64
                    if w:
                        assert tab + stash <= lds, (prec, m_max, ne, nb, w)
                    else:
                        assert tab + rb * nt * 8 * 64 > lds, (prec, m_max, ne, nb)
    assert lib.ghost_far_fill_waves(0, 126, 2048, 2048) == 1
    assert lib.ghost_far_fill_waves(0, 110, 2048, 2048) == 2
    for ne in grids:
        for nb in grids:
            assert lib.ghost_far_fill_waves(1, 126, ne, nb) >= 1   # FP32 always fits


def test_void_compat_samplers_fill_nan_on_failure():
    """The reference ABI has no error channel: a failing sample_g /
    sample_mfie / sample_ibc leaves every output NaN (here a null input array,
    rejected on the host before any device call)."""
    lib = pkg("_lib").load()
    dp = ctypes.POINTER(ctypes.c_double)
    nr, np_, nxi = 2, 3, 8
    a = np.ones(max(nr, np_, nxi))
    P = lambda x: x.ctypes.data_as(dp)
    null = ctypes.cast(None, dp)
    out = np.zeros((nr, np_, nxi), dtype=np.complex128)
    lib.sample_g(nr, np_, nxi, null, P(a), P(a), P(a), ctypes.c_double(1.0), P(a), out.ctypes.data_as(dp))
    assert np.all(np.isnan(out.real)) and np.all(np.isnan(out.imag))
    for fname in ("sample_mfie", "sample_ibc"):
        outs = [np.zeros((nr, np_, nxi), dtype=np.complex128) for _ in range(4)]
        getattr(lib, fname)(nr, np_, nxi, P(a), P(a), P(a), P(a), P(a), null, P(a), P(a),
                            ctypes.c_double(1.0), P(a), P(a), *[o.ctypes.data_as(dp) for o in outs])
        for o in outs:
            assert np.all(np.isnan(o.real)) and np.all(np.isnan(o.imag))
