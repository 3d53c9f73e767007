"""Per-mode operator API (BorPecSolver) and the generic mode loop (_mode_sweep)
on the device, against the FP64 oracle and the batched solve_bor (run on an
MI355X)."""

import math

import numpy as np
import pytest

import bor_oracle as bo
from conftest import pkg

pytestmark = pytest.mark.gpu

ETA0 = 376.730313668


def rel(a, b):
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b))) / max(np.max(np.abs(b)), 1e-300))


@pytest.fixture(scope="module")
def case():
    wl = pkg("workloads")
    nodes = wl.cone_sphere(16, 15.0, 0.3)
    return nodes, 3e8, 4


def test_bor_pec_solver_operators_match_oracle(case):
    nodes, f, mm = case
    bs = pkg("bor_solver")
    s = bs.BorPecSolver(nodes, f)
    ms = bo.ModeSystem(nodes, f, mm, "cfie", 0.5, streaming=False)
    body, k = ms.body, ms.k
    assert rel(s.mass_blocks(), bo.mass_matrix(body)) < 1e-13
    for m in (-3, 0, 1, 2):
        print(f"[case] m={m}", flush=True)
        assert rel(s.assemble_mode(m, mm), ms.efie_full(m)) < 1e-11, m
        assert rel(s.assemble_mfie_mode(m, mm), ms.mfie_full(m)) < 1e-11, m
        assert np.array_equal(s.basis_mask(m), bo.basis_mask(body, m))
        for th in (0.0, 37.0, 120.0):
            for pol in ("VV", "HH"):
                assert rel(s.rhs_mode(m, th, pol), bo.rhs_efie(body, k, m, th, pol)) < 1e-12
                assert rel(s.rhs_mfie_mode(m, th, pol), bo.rhs_mfie(body, k, m, th, pol)) < 1e-12
    rng = np.random.default_rng(5)
    sol = rng.standard_normal(2 * body.Nn) + 1j * rng.standard_normal(2 * body.Nn)
    for m, th in ((0, 30.0), (2, 75.0), (-1, 150.0)):
        got = s.farfield_mode(m, sol, th)
        want = bo.farfield(body, k, m, sol, th)
        assert abs(got[0] - want[0]) <= 1e-12 * abs(want[0]) + 1e-300
        assert abs(got[1] - want[1]) <= 1e-12 * abs(want[1]) + 1e-300


def test_mode_sweep_with_device_operators_matches_solve_bor(case):
    nodes, f, mm = case
    bs = pkg("bor_solver")
    s = bs.BorPecSolver(nodes, f)
    thetas = np.array([10.0, 60.0, 110.0, 170.0])
    a = 0.5

    def assemble(m):
        Z = a * s.assemble_mode(m, mm) + (1 - a) * ETA0 * s.assemble_mfie_mode(m, mm)
        mask = s.basis_mask(m)
        return Z[np.ix_(mask, mask)], mask

    def rhs(m, th, pol):
        return a * s.rhs_mode(m, th, pol) + (1 - a) * ETA0 * s.rhs_mfie_mode(m, th, pol)

    def farfield(m, sol, th, pol):
        fth, fph = s.farfield_mode(m, sol, th)
        return fth if pol == "VV" else fph

    F, used, stats = bs._mode_sweep(2 * s.Nn, thetas, ["VV", "HH"], mm, 0.0, assemble, rhs, farfield)
    ref = bs.solve_bor(nodes, f, thetas, formulation="cfie", n_modes=mm, mode_tol=0.0)
    assert used == ref["modes_used"] == mm
    assert rel(F[0], ref["amp_vv"]) < 1e-10
    assert rel(F[1], ref["amp_hh"]) < 1e-10
    assert stats["linear_residual"] < 1e-12


@pytest.mark.parametrize("nb,lu_super", [(64, 1), (64, 2), (64, 4), (32, 5), (16, 4)])
def test_dense_batch_solves_caller_systems(nb, lu_super):
    """Uneven caller systems in one batched LU, FP64 and mixed precision, for
    several block sizes and LU super-blocks (the catch-up GEMMs of a
    super-block run when a system spans several of its panels)."""
    L = pkg("_lib")
    rng = np.random.default_rng(9)
    sizes = [(37, 5), (130, 17), (64, 64), (201, 3), (517, 40)]
    mats = []
    for n, r in sizes:
        A = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)) + 3 * math.sqrt(n) * np.eye(n)
        B = rng.standard_normal((n, r)) + 1j * rng.standard_normal((n, r))
        mats.append((A, B, np.linalg.solve(A, B)))
    for prec, tol, it in ((L.GHOST_FP64, 1e-12, 0), (L.GHOST_MIXED, 1e-12, 3)):
        print(f"[case] nb={nb} lu_super={lu_super} precision={prec}", flush=True)
        h = L.Handle(prec, nb=nb)
        h.set_lu_super(lu_super)
        idx = [h.add_dense(n, r) for n, r in sizes]
        h.build()
        for i, (A, B, _) in zip(idx, mats):
            h.set_dense(i, A, B)
        h.solve(it)
        h.sync()
        for i, (_, _, X) in zip(idx, mats):
            assert rel(h.get_dense(i), X) < tol, (prec, i)
        h.close()


@pytest.mark.parametrize("nb,lu_super", [(64, 4), (32, 1)])
def test_pivot_order_matches_lapack_getrf(nb, lu_super):
    """The batched LU picks LAPACK's pivots (first maximum of |Re| + |Im|, as
    zgetrf's izamax).  Random systems, FP64: the row interchanges equal
    scipy.linalg.lu_factor's exactly.  Row-permuted, strongly diagonally
    dominant systems, FP64 and the FP32 factor of the mixed path: every
    column's pivot is unambiguous, so both equal LAPACK's.  The sizes use one,
    two and three panel rows per thread."""
    import scipy.linalg as sla
    L = pkg("_lib")
    rng = np.random.default_rng(21)
    sizes = [(37, 2), (517, 3), (1500, 2), (2100, 2)]

    def random_sys(n, r):
        return rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))

    def dominant_sys(n, r):
        D = np.diag(10.0 + rng.uniform(0, 1, n)) * np.exp(1j * rng.uniform(0, 2 * math.pi, n))
        N = 1e-3 * (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))
        return (D + N)[rng.permutation(n)]

    cases = [(L.GHOST_FP64, random_sys), (L.GHOST_FP64, dominant_sys), (L.GHOST_MIXED, dominant_sys)]
    for prec, make in cases:
        print(f"[case] nb={nb} lu_super={lu_super} precision={prec} {make.__name__}", flush=True)
        mats = [(make(n, r), rng.standard_normal((n, r)) + 1j * rng.standard_normal((n, r))) for n, r in sizes]
        h = L.Handle(prec, nb=nb)
        h.set_lu_super(lu_super)
        idx = [h.add_dense(n, r) for n, r in sizes]
        h.build()
        for i, (A, B) in zip(idx, mats):
            h.set_dense(i, A, B)
        h.solve(0)
        h.sync()
        for i, (A, _) in zip(idx, mats):
            got = h.pivots(i)
            want = sla.lu_factor(A, check_finite=False)[1]
            assert np.array_equal(got, want), (make.__name__, i, int(np.argmax(got != want)))
        h.close()


@pytest.mark.parametrize("precision", [0, 1])
def test_mode_pairing_matches_separate_mode_solves(precision):
    """Pairing (only m >= 0 factored; mode -m from mode m's currents by the
    signs of S [VV, -HH]) against every mode solved on its own: systems,
    right-hand sides, currents and far field of each mode agree to rounding."""
    L = pkg("_lib")
    bs = pkg("bor_solver")
    nodes = bo.sphere_nodes(0.5, 40)
    k = 2 * math.pi * 3e8 / bs.C0
    thetas = np.array([0.0, 25.0, 90.0, 140.0, 180.0])
    mm = 9
    _, ne, nbk, _ = bs.discretisation(nodes, k, thetas, mm, "cfie", "tables")
    print(f"[case] pairing precision={precision}", flush=True)
    out = []
    for pairing in (True, False):
        h = L.Handle(precision, pairing=pairing)
        h.add_problem(nodes, k, mm, ne, nbk, 0.5, 0.5 * bs.ETA0, thetas)
        h.build()
        h.fill()
        h.sync()
        AB = [h.system(0, m) for m in range(-mm, mm + 1)]
        h.solve(2 if precision else 0)
        h.farfield()
        h.sync()
        X = [h.solution(0, m) for m in range(-mm, mm + 1)]
        out.append((AB, X, h.dF(0)))
        h.close()
    (ab1, x1, f1), (ab0, x0, f0) = out
    tol = 1e-12 if precision == 0 else 1e-9
    for m, (a, b, xa, xb) in enumerate(zip(ab1, ab0, x1, x0)):
        assert np.max(np.abs(a[0] - b[0])) <= 1e-13 * np.max(np.abs(b[0])), m - mm
        assert np.max(np.abs(a[1] - b[1])) <= 1e-13 * np.max(np.abs(b[1])), m - mm
        assert np.max(np.abs(xa - xb)) <= tol * np.max(np.abs(xb)), m - mm
    assert np.max(np.abs(f1 - f0)) <= tol * np.max(np.abs(f0))


@pytest.mark.parametrize("precision", [0, 1])
def test_two_stream_lu_equals_one_stream(precision):
    """The two interleaved halves of the systems on two streams run the same
    per-system operations in the same order as the whole-batch schedule on
    one stream: the currents are bitwise equal, also on uneven batches
    (two frequencies, so the halves hold systems of different sizes)."""
    L = pkg("_lib")
    bs = pkg("bor_solver")
    nodes = bo.sphere_nodes(0.5, 30)
    thetas = np.array([0.0, 60.0, 120.0])
    print(f"[case] lu streams precision={precision}", flush=True)
    h = L.Handle(precision)
    for f in (3e8, 4.5e8):
        k = 2 * math.pi * f / bs.C0
        _, ne, nbk, _ = bs.discretisation(nodes, k, thetas, 7, "cfie", "tables")
        h.add_problem(nodes, k, 7, ne, nbk, 0.5, 0.5 * bs.ETA0, thetas)
    h.build()
    out = []
    for n in (2, 1):
        h.set_lu_streams(n)
        h.fill()
        h.solve(1 if precision else 0)
        h.farfield()
        h.sync()
        out.append([h.solution(p, m) for p in range(2) for m in range(-7, 8)])
    for a, b in zip(*out):
        assert np.array_equal(a, b)
    h.close()


def test_pairing_with_dense_systems_in_the_batch():
    """Paired mode systems and caller-filled dense systems in one batch: the
    dense solve is np.linalg.solve's, the mode currents equal an unpaired
    handle's, and the dense system keeps its own pivots and zero-pivot count."""
    L = pkg("_lib")
    bs = pkg("bor_solver")
    rng = np.random.default_rng(3)
    nodes = bo.sphere_nodes(0.5, 24)
    k = 2 * math.pi * 3e8 / bs.C0
    thetas = np.array([10.0, 80.0])
    _, ne, nbk, _ = bs.discretisation(nodes, k, thetas, 5, "cfie", "tables")
    n, r = 150, 7
    A = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)) + 4 * math.sqrt(n) * np.eye(n)
    B = rng.standard_normal((n, r)) + 1j * rng.standard_normal((n, r))
    print("[case] pairing with a dense system", flush=True)
    X = []
    for pairing in (True, False):
        h = L.Handle(L.GHOST_FP64, pairing=pairing)
        h.add_problem(nodes, k, 5, ne, nbk, 0.5, 0.5 * bs.ETA0, thetas)
        d = h.add_dense(n, r)
        h.build()
        h.fill()
        h.set_dense(d, A, B)
        h.solve(0)
        h.sync()
        Xd = h.get_dense(d)
        assert np.max(np.abs(Xd - np.linalg.solve(A, B))) <= 1e-12 * np.max(np.abs(Xd))
        assert not np.any(h.zero_pivots())
        X.append([h.solution(0, m) for m in range(-5, 6)])
        h.close()
    for a, b in zip(*X):
        assert np.max(np.abs(a - b)) <= 1e-12 * np.max(np.abs(b))
