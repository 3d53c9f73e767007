"""The benchmark workloads themselves on the HIP path (run on an MI355X):

* the 120-segment PEC sphere at 300 MHz (|m| <= 12, 181 aspects) against
  the reference's own result (fixture) and the Mie series, FP64 and mixed;
* the 1200-element ogive-cylinder-cone (|m| <= 70, 361 aspects): complete
  system rows against the oracle, and full-size invariants (residual, mixed
  vs FP64, mode pairing on = off);
* the 64-frequency batched sweep: batch = single problems, graph = eager,
  mixed vs FP64;
* the 600-element cone-sphere adjoint against central differences.
"""

import math
import os

import numpy as np
import pytest

import mie
from conftest import pkg

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    return np.load(os.path.join(GOLD, name + ".npz"))


def rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def db(s):
    return 10.0 * np.log10(np.asarray(s, dtype=float))


# ── 120-segment sphere ─────────────────────────────────────────────────────

@pytest.mark.parametrize("form", ["cfie", "efie"])
def test_sphere120_matches_reference_and_mie(form, bs):
    d = load("sphere120_cfg1")
    print(f"[case] sphere120 {form}", flush=True)
    kw = dict(formulation=form, n_modes=12, mode_tol=0.0)
    out = bs.solve_bor(d["nodes"], float(d["freq_hz"]), d["thetas"], **kw)
    assert out["modes_used"] == int(d[f"{form}_modes_used"])
    for pol in ("amp_vv", "amp_hh"):
        assert rel(np.array(out[pol]), d[f"{form}_{pol}"]) < 1e-9, pol
    for pol in ("sigma_vv", "sigma_hh"):
        assert np.max(np.abs(db(out[pol]) - db(d[f"{form}_{pol}"]))) < 1e-7, pol
        assert np.max(np.abs(db(out[pol]) - db(mie.sigma_pec_sphere(0.5, float(d["freq_hz"]))))) < 0.05
    assert out["linear_residual"] < 1e-12
    dflt = bs.solve_bor(d["nodes"], float(d["freq_hz"]), d["thetas"], formulation=form, n_modes=12, mode_tol=1e-6)
    assert dflt["modes_used"] == int(d[f"{form}_modes_used_default_tol"])
    mx = bs.solve_bor(d["nodes"], float(d["freq_hz"]), d["thetas"], precision="mixed", refine_iters=-3, **kw)
    for pol in ("sigma_vv", "sigma_hh"):
        assert np.max(np.abs(db(mx[pol]) - db(d[f"{form}_{pol}"]))) < 0.01, pol


def _sphere120_handle(precision=0, nb=64, variant=None, lu_super=None):
    L = pkg("_lib")
    bs = pkg("bor_solver")
    d = load("sphere120_cfg1")
    k = 2 * math.pi * float(d["freq_hz"]) / bs.C0
    th = d["x_thetas"]
    _, ne, nbk, _ = bs.discretisation(d["nodes"], k, th, 12, "cfie")
    h = L.Handle(precision, nb=nb)
    if variant is not None:
        h.set_gemm_variant(variant)
    if lu_super is not None:
        h.set_lu_super(lu_super)
    h.add_problem(d["nodes"], k, 12, ne, nbk, 0.5, 0.5 * bs.ETA0, th)
    h.build()
    return h, d


@pytest.mark.parametrize("lu_super", [1, 3, 4])
@pytest.mark.parametrize("variant", [0, 1, 2, 3, 4, 5])
@pytest.mark.parametrize("nb", [16, 32, 64])
@pytest.mark.parametrize("precision", [0, 1])
def test_lu_tile_variants_on_sphere120(variant, nb, precision, lu_super):
    """Every GEMM tile variant, LU block size and super-block depth on the
    238 / 240-unknown systems (4 to 15 panels): the reference's currents to
    1e-9 in FP64.  Mixed (FP32 far fill and LU, two FP64 refinement sweeps):
    1e-5; the sweeps converge to the solution of the FP32-filled matrix,
    whose far entries carry FP32 rounding (2.4e-6 in the currents here, the
    fill's rounding times the conditioning), far inside the 0.01 dB bar."""
    print(f"[case] sphere120 variant={variant} nb={nb} precision={precision} lu_super={lu_super}", flush=True)
    h, d = _sphere120_handle(precision, nb, variant, lu_super)
    h.fill()
    h.solve(2 if precision == 1 else 0)
    h.sync()
    for m in (0, 1, -1, 12):
        X = h.solution(0, m)
        err = rel(X, d[f"X_{m}"])
        assert err < (1e-5 if precision == 1 else 1e-9), (m, err)
    h.close()


# ── 1200-element ogive-cylinder-cone ───────────────────────────────────────

def _ogive(precision=0, pairing=True, unmasked=False, thetas=None, residual=True):
    L = pkg("_lib")
    bs = pkg("bor_solver")
    wl = pkg("workloads")
    name, nodes, f, th, mm = wl.config(2)
    th = th if thetas is None else np.asarray(thetas, dtype=float)
    k = 2 * math.pi * f / bs.C0
    _, ne, nbk, _ = bs.discretisation(nodes, k, th, mm, "cfie")
    h = L.Handle(precision, pairing=pairing, residual=residual)
    h.add_problem(nodes, k, mm, ne, nbk, 0.5, 0.5 * bs.ETA0, th, unmasked=unmasked)
    h.build()
    return h, nodes, (ne, nbk)


def test_ogive1200_rows_match_oracle():
    """Complete rows (nodes E, E+1; J_t and J_phi) of the unmasked CFIE
    systems of modes 0, 1, 35, 70 against the oracle, FP64, 1e-11."""
    d = load("oracle_ogive1200_rows")
    print("[case] ogive1200 rows FP64", flush=True)
    h, nodes, nxi = _ogive(unmasked=True, thetas=[0.0, 90.0], residual=False)
    assert np.array_equal(nodes, d["nodes"])
    assert tuple(nxi) == tuple(d["n_xi"])
    h.fill()
    h.sync()
    Nn = len(nodes)
    for m in d["modes"]:
        A, _ = h.system(0, int(m))
        for E in d["elems"]:
            got = A[[E, E + 1, Nn + E, Nn + E + 1]]
            err = rel(got, d[f"rows_{E}_{m}"])
            print(f"[case] m={m} E={E} rel {err:.2e}", flush=True)
            assert err < 1e-11, (m, E, err)
        del A
    h.close()


def test_ogive1200_full_size_residual_mixed_and_pairing():
    """The whole sweep (141 modes, 361 aspects x 2 pols): FP64 residual
    below 1e-12, mixed within 0.01 dB of FP64, and mode pairing on = off
    (currents of modes +-35 and +-70)."""
    bs = pkg("bor_solver")
    print("[case] ogive1200 FP64 paired", flush=True)
    h, nodes, _ = _ogive(0, pairing=True)
    h.run(0)
    h.sync()
    res = h.residuals()
    assert np.max(res) < 1e-12, np.max(res)
    F64 = h.dF(0).sum(axis=0)
    X_on = {m: h.solution(0, m) for m in (35, -35, 70, -70)}
    h.close()
    print("[case] ogive1200 mixed", flush=True)
    h, _, _ = _ogive(1, pairing=True, residual=False)
    h.run(-2)
    h.sync()
    F32 = h.dF(0).sum(axis=0)
    h.close()
    err = np.max(np.abs(db(4 * math.pi * np.abs(F32) ** 2) - db(4 * math.pi * np.abs(F64) ** 2)))
    print(f"[case] mixed vs FP64 {err:.2e} dB", flush=True)
    assert err < 0.01
    print("[case] ogive1200 FP64 unpaired", flush=True)
    h, _, _ = _ogive(0, pairing=False, residual=False)
    h.fill()
    h.solve(0)
    h.sync()
    for m, X in X_on.items():
        assert rel(h.solution(0, m), X) < 1e-12, m
    h.close()


# ── 64-frequency batched sweep ─────────────────────────────────────────────

def test_frequency_sweep_batch_graph_and_mixed(bs):
    wl = pkg("workloads")
    jobs = wl.frequency_sweep(64)
    print("[case] sweep64 FP64 batch", flush=True)
    batch = bs.solve_bor_batch(jobs, mode_tol=0.0, residual=False)
    print("[case] sweep64 FP64 graph", flush=True)
    graph = bs.solve_bor_batch(jobs, mode_tol=0.0, residual=False, graph=True)
    for b, g in zip(batch, graph):
        assert np.array_equal(np.array(b["amp_vv"]), np.array(g["amp_vv"]))
        assert np.array_equal(np.array(b["amp_hh"]), np.array(g["amp_hh"]))
    for i in (0, 31, 63):
        j = jobs[i]
        one = bs.solve_bor(j["points"], j["freq_hz"], j["thetas_deg"], formulation="cfie",
                           n_modes=j["n_modes"], mode_tol=0.0)
        assert rel(np.array(batch[i]["amp_vv"]), np.array(one["amp_vv"])) < 1e-12, i
        assert rel(np.array(batch[i]["amp_hh"]), np.array(one["amp_hh"])) < 1e-12, i
    print("[case] sweep64 mixed", flush=True)
    mixed = bs.solve_bor_batch(jobs, precision="mixed", refine_iters=-3, mode_tol=0.0, residual=False)
    worst = 0.0
    for a, b in zip(batch, mixed):
        for pol in ("sigma_vv", "sigma_hh"):
            worst = max(worst, float(np.max(np.abs(db(a[pol]) - db(b[pol])))))
    print(f"[case] sweep64 mixed vs FP64 {worst:.2e} dB", flush=True)
    assert worst < 0.01


# ── 600-element cone-sphere adjoint ────────────────────────────────────────

def test_cone_sphere600_adjoint_matches_central_differences():
    """J = sum sigma_dBsm over 91 aspects x 2 pols; dJ/d rho at three interior
    nodes (the largest |g| of each third of the generatrix) against central
    differences with h = 1e-6 wavelength, 1e-4 relative."""
    ad = pkg("bor_adjoint")
    wl = pkg("workloads")
    name, nodes, f, th, mm = wl.config(4)
    spec = ad.RcsSpec(nodes, f, th, mm, cfie_alpha=0.5)
    print("[case] cone-sphere600 adjoint", flush=True)
    J, g = ad.objective_and_gradient(spec)
    Nn = len(nodes)
    thirds = [(2, Nn // 3), (Nn // 3, 2 * Nn // 3), (2 * Nn // 3, Nn - 2)]
    picks = [a + int(np.argmax(np.abs(g[a:b]))) for a, b in thirds]
    h = 1e-6 * (299_792_458.0 / f)
    for q in picks:
        vals = []
        for sgn in (1.0, -1.0):
            r = nodes[:, 0].copy()
            r[q] += sgn * h
            F, hd = ad.solve_far_field(spec, r)
            hd.close()
            vals.append(float(np.sum(ad.sigma_db(F))))
        fd = (vals[0] - vals[1]) / (2 * h)
        print(f"[case] node {q}: adjoint {g[q]:.9e} fd {fd:.9e}", flush=True)
        assert abs(g[q] - fd) <= 1e-4 * abs(fd), (q, g[q], fd)
