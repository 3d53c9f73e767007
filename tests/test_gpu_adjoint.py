"""Adjoint RCS gradient (HIP path) against central differences of the forward
solve, on the GPU and through the FP64 oracle (run on an MI355X)."""

import math

import numpy as np
import pytest
import torch

import bor_oracle as bo
from conftest import pkg

pytestmark = pytest.mark.gpu


def cone_sphere_spec(n_elems, m_max, n_theta, alpha):
    wl = pkg("workloads")
    ad = pkg("bor_adjoint")
    nodes = wl.cone_sphere(n_elems, 15.0, 0.3)
    return ad.RcsSpec(nodes, 3e8, np.linspace(2.0, 178.0, n_theta), m_max, cfie_alpha=alpha)


def fd_objective(spec, w, q, h):
    ad = pkg("bor_adjoint")
    vals = []
    for sgn in (1.0, -1.0):
        r = spec.nodes[:, 0].copy()
        r[q] += sgn * h
        F, hd = ad.solve_far_field(spec, r)
        hd.close()
        vals.append(float(np.sum(w * ad.sigma_db(F))))
    return (vals[0] - vals[1]) / (2 * h)


@pytest.mark.parametrize("alpha", [0.5, 1.0])
def test_adjoint_gradient_matches_central_differences(alpha):
    ad = pkg("bor_adjoint")
    spec = cone_sphere_spec(40, 6, 9, alpha)
    rng = np.random.default_rng(11)
    w = rng.uniform(0.5, 1.5, size=(2, spec.thetas_deg.size))
    x = torch.zeros(spec.nodes.shape[0], dtype=torch.float64, requires_grad=True)
    s = ad.rcs_db(x, spec)
    J = (s * torch.from_numpy(w)).sum()
    J.backward()
    g = x.grad.numpy()
    print(f"[case] alpha={alpha} |g|max={np.max(np.abs(g)):.3e}", flush=True)
    # the axis end nodes do not move
    assert g[0] == 0.0 and g[-1] == 0.0
    lam = 299_792_458.0 / spec.freq_hz
    for q in (1, 2, 7, 13, 20, 27, 33, 38):
        fd = fd_objective(spec, w, q, 1e-6 * lam)
        print(f"[case] node {q}: adjoint {g[q]:.9e} fd {fd:.9e}", flush=True)
        assert abs(g[q] - fd) <= 2e-7 * np.max(np.abs(g)) + 1e-6 * abs(fd), (q, g[q], fd)


def test_adjoint_gradient_matches_oracle_differences():
    """One node against central differences of the FP64 NumPy oracle: an
    independent implementation of the whole forward path."""
    ad = pkg("bor_adjoint")
    spec = cone_sphere_spec(24, 4, 5, 0.5)
    w = np.ones((2, spec.thetas_deg.size))
    J, g = ad.objective_and_gradient(spec, w)
    q = 11
    h = 1e-5 * 299_792_458.0 / spec.freq_hz
    vals = []
    for sgn in (1.0, -1.0):
        nodes = spec.nodes.copy()
        nodes[q, 0] += sgn * h
        out = bo.solve_bor(nodes, spec.freq_hz, spec.thetas_deg, formulation="cfie", n_modes=spec.m_max,
                           mode_tol=0.0)
        vals.append(float(np.sum(10 * np.log10(out["sigma_vv"])) + np.sum(10 * np.log10(out["sigma_hh"]))))
    fd = (vals[0] - vals[1]) / (2 * h)
    print(f"[case] oracle fd {fd:.9e} adjoint {g[q]:.9e}", flush=True)
    assert abs(g[q] - fd) <= 1e-6 * abs(fd)


def test_adjoint_of_a_batch_equals_single_problem_adjoints():
    """Two problems (different bodies and frequencies) in one handle: the
    batched adjoint returns each problem's gradient as its own handle would."""
    L = pkg("_lib")
    bs = pkg("bor_solver")
    wl = pkg("workloads")
    probs = [(wl.cone_sphere(30, 15.0, 0.3), 3e8, 5, np.linspace(5.0, 175.0, 7)),
             (wl.random_smooth(1.0, 0.05, 0.2, 36, 3), 4.5e8, 6, np.linspace(0.0, 180.0, 5))]

    def setup(h, nodes, f, mm, th):
        k = 2 * math.pi * f / bs.C0
        _, ne, nb, _ = bs.discretisation(nodes, k, th, mm, "cfie")
        h.add_problem(nodes, k, mm, ne, nb, 0.5, 0.5 * bs.ETA0, th)

    def acoef(dF, w):
        F = dF.sum(axis=0)
        a = w * (20.0 / math.log(10.0)) * np.conj(F) / np.abs(F) ** 2
        return np.ascontiguousarray(a.T).ravel()

    rng = np.random.default_rng(4)
    ws = [rng.uniform(0.5, 1.5, size=(2, p[3].size)) for p in probs]
    single = []
    for p, w in zip(probs, ws):
        h = L.Handle(L.GHOST_FP64, residual=False, pairing=False)
        setup(h, *p)
        h.build()
        h.run(0)
        h.sync()
        single.append(h.adjoint(acoef(h.dF(0), w), 0.0))
        h.close()
    h = L.Handle(L.GHOST_FP64, residual=False, pairing=False)
    for p in probs:
        setup(h, *p)
    h.build()
    h.run(0)
    h.sync()
    a = np.concatenate([acoef(h.dF(i), w) for i, w in enumerate(ws)])
    g = h.adjoint(a, 0.0)
    h.close()
    n0 = probs[0][0].shape[0]
    for got, want in ((g[:n0], single[0]), (g[n0:], single[1])):
        assert np.max(np.abs(got - want)) <= 1e-9 * np.max(np.abs(want))


@pytest.mark.parametrize("alpha", [0.5, 1.0, 0.0])
def test_analytic_adjoint_matches_difference_adjoint(alpha):
    """The analytic fill derivative (dual numbers in the kernels) against the
    central-difference adjoint (12 perturbed fills of the same handle), CFIE,
    EFIE and MFIE: the same discrete gradient, so they agree to the
    difference adjoint's own truncation error."""
    L = pkg("_lib")
    bs = pkg("bor_solver")
    wl = pkg("workloads")
    nodes = wl.cone_sphere(30, 15.0, 0.3)
    th = np.linspace(3.0, 177.0, 7)
    f, mm = 3e8, 5
    k = 2 * math.pi * f / bs.C0
    form = "cfie" if 0.0 < alpha < 1.0 else ("efie" if alpha == 1.0 else "mfie")
    _, ne, nb, _ = bs.discretisation(nodes, k, th, mm, form)
    h = L.Handle(L.GHOST_FP64, residual=False, pairing=False)
    h.add_problem(nodes, k, mm, ne, nb, alpha, (1.0 - alpha) * bs.ETA0, th)
    h.build()
    h.run(0)
    h.sync()
    F = h.dF(0).sum(axis=0)
    rng = np.random.default_rng(7)
    w = rng.uniform(0.5, 1.5, size=F.shape)
    a = np.ascontiguousarray((w * np.conj(F) / np.abs(F) ** 2).T).ravel()
    ga = h.adjoint(a, method="analytic")
    ga2 = h.adjoint(a, method="analytic")
    gf = h.adjoint(a, 0.0, method="fd")
    h.close()
    print(f"[case] alpha={alpha} |ga|max={np.max(np.abs(ga)):.3e} |gf|max={np.max(np.abs(gf)):.3e}", flush=True)
    for q in range(0, nodes.shape[0], 3):
        print(f"[case] node {q}: analytic {ga[q]:.10e} fd {gf[q]:.10e}", flush=True)
    assert np.array_equal(ga, ga2), "the analytic adjoint is deterministic and leaves the handle solved"
    assert np.max(np.abs(ga - gf)) <= 1e-6 * np.max(np.abs(gf))


def test_mixed_precision_adjoint_matches_fp64():
    """FP32 fill and LU under FP64 refinement, forward and adjoint (lambda on
    the FP32 factors with one FP64 refinement sweep): the gradient agrees with
    the FP64 path to the 1e-4 relative bar, and with central differences of
    the FP64 forward."""
    ad = pkg("bor_adjoint")
    spec = cone_sphere_spec(40, 6, 9, 0.5)
    w = np.ones((2, spec.thetas_deg.size))
    J64, g64 = ad.objective_and_gradient(spec, w)
    spec.precision = "mixed"
    J32, g32 = ad.objective_and_gradient(spec, w)
    err = np.max(np.abs(g32 - g64)) / np.max(np.abs(g64))
    print(f"[case] mixed vs fp64 gradient: max rel err {err:.3e}, J {J32:.9e} vs {J64:.9e}", flush=True)
    assert err <= 1e-4
    assert abs(J32 - J64) <= 1e-4 * abs(J64)


def test_captured_forward_and_adjoint_equal_eager():
    """Forward RCS and gradient of two perturbed bodies at two frequencies in
    one graph launch (mixed precision, fixed refinement, the objective's a'
    formed on the device) against the same handles run eagerly, and the mixed
    gradient against the FP64 one."""
    L = pkg("_lib")
    bs = pkg("bor_solver")
    wl = pkg("workloads")
    probs = wl.perturbed_batch(2, freq_scales=(0.5, 0.8), n_theta=13, base_elems=120, base_m=8)
    specs = []
    for (_, nodes, f, th, mm) in probs:
        k = 2 * math.pi * f / bs.C0
        _, ne, nb, _ = bs.discretisation(nodes, k, th, mm, "cfie")
        specs.append((nodes, k, mm, ne, nb, th))

    def handle(prec):
        h = L.Handle(prec, residual=False, pairing=False)
        for (nodes, k, mm, ne, nb, th) in specs:
            h.add_problem(nodes, k, mm, ne, nb, 0.5, 0.5 * bs.ETA0, th)
        h.build()
        return h

    rng = np.random.default_rng(3)
    ws = [rng.uniform(0.5, 1.5, size=2 * sp[5].size) for sp in specs]
    wflat = np.concatenate(ws)

    def acoef(h):
        out = []
        for i, w in enumerate(ws):
            F = h.dF(i).sum(axis=0)                       # [2, n_theta]
            Fc = np.ascontiguousarray(F.T).ravel()        # column 2 it + pol
            out.append(w * (20.0 / math.log(10.0)) * np.conj(Fc) / np.abs(Fc) ** 2)
        return np.concatenate(out)

    h = handle(L.GHOST_MIXED)
    h.set_objective(wflat)
    h.capture_adjoint(1)
    h.launch()
    h.sync()
    g_graph = h.gradient()
    dF_graph = [h.dF(i) for i in range(len(specs))]
    h.launch()
    h.sync()
    g_graph2 = h.gradient()
    h.close()
    h = handle(L.GHOST_MIXED)
    h.run(1)
    h.sync()
    g_eager = h.adjoint(acoef(h), method="analytic")
    dF_eager = [h.dF(i) for i in range(len(specs))]
    h.close()
    h = handle(L.GHOST_FP64)
    h.run(0)
    h.sync()
    g64 = h.adjoint(acoef(h), method="analytic")
    h.close()
    rel = np.max(np.abs(g_graph - g_eager)) / np.max(np.abs(g_eager))
    rel64 = np.max(np.abs(g_graph - g64)) / np.max(np.abs(g64))
    print(f"[case] graph vs eager {rel:.3e}, graph vs fp64 {rel64:.3e}", flush=True)
    assert np.array_equal(g_graph, g_graph2), "graph replays agree bitwise"
    for a, b in zip(dF_graph, dF_eager):
        assert np.array_equal(a, b)
    # the graph forms a' on the device, the eager run on the host: their last-bit
    # differences pass through the FP32 Q of a mixed handle
    assert rel <= 1e-8
    assert rel64 <= 1e-4


def test_set_nodes_reuses_a_built_handle_and_its_graph():
    """ghost_bor_set_nodes swaps a built problem's geometry in place: the
    captured forward + gradient graph then returns the new shape's far field
    and gradient, as a handle built on that shape does (same quadrature
    rules: the shapes differ by 1e-3 wavelengths)."""
    L = pkg("_lib")
    bs = pkg("bor_solver")
    wl = pkg("workloads")
    probs = wl.perturbed_batch(2, freq_scales=(0.6,), n_theta=9, base_elems=100, base_m=6, amplitude=1e-3)
    (_, nodes_a, f, th, mm), (_, nodes_b, _, _, _) = probs
    k = 2 * math.pi * f / bs.C0
    _, ne, nb, _ = bs.discretisation(nodes_a, k, th, mm, "cfie")

    def run(nodes_build, nodes_set):
        h = L.Handle(L.GHOST_MIXED, residual=False, pairing=False)
        h.add_problem(nodes_build, k, mm, ne, nb, 0.5, 0.5 * bs.ETA0, th)
        h.build()
        h.set_objective(np.ones(2 * th.size))
        h.capture_adjoint(1)
        if nodes_set is not None:
            h.set_nodes(0, nodes_set)
        h.launch()
        h.sync()
        out = (h.dF(0), h.gradient())
        h.close()
        return out

    dF_b, g_b = run(nodes_b, None)
    dF_ab, g_ab = run(nodes_a, nodes_b)
    dF_a, g_a = run(nodes_a, None)
    rel_f = np.max(np.abs(dF_ab - dF_b)) / np.max(np.abs(dF_b))
    rel_g = np.max(np.abs(g_ab - g_b)) / np.max(np.abs(g_b))
    moved = np.max(np.abs(dF_a - dF_b)) / np.max(np.abs(dF_b))
    print(f"[case] set_nodes: dF {rel_f:.3e} grad {rel_g:.3e} (shapes differ by {moved:.3e})", flush=True)
    assert moved > 1e-6
    assert rel_f <= 1e-9 and rel_g <= 1e-7
