"""The perturbed-shape batch of the batched-adjoint benchmark (host only):
Fourier radial perturbations vanish at the axis ends, keep every interior
node off the axis, reach the stated amplitude, are seeded, and mesh the three
frequencies at lambda / 10."""

import numpy as np

from conftest import pkg


def test_fourier_perturbation_properties():
    wl = pkg("workloads")
    base = wl.ogive_cylinder_cone(600, 1.0)
    rng = np.random.default_rng(0)
    for _ in range(5):
        c = rng.uniform(-1, 1, 16) / np.arange(1, 17)
        out = wl.fourier_perturbation(base, c, 0.05)
        d = out[:, 0] - base[:, 0]
        assert out[0, 0] == base[0, 0] and out[-1, 0] == base[-1, 0]
        assert np.array_equal(out[:, 1], base[:, 1])             # radial only
        assert np.all(out[1:-1, 0] > 0.0)
        assert np.max(np.abs(d)) <= 0.05 + 1e-15
        # the taper is 1 where rho >= 0.2: there the offset reaches the amplitude's scale
        assert np.max(np.abs(d[base[:, 0] >= 0.2])) > 0.02


def test_perturbed_batch_is_seeded_and_meshed_per_frequency():
    wl = pkg("workloads")
    a = wl.perturbed_batch(3)
    b = wl.perturbed_batch(3)
    assert len(a) == 9
    for (ia, na, fa, ta, ma), (ib, nb, fb, tb, mb) in zip(a, b):
        assert ia == ib and fa == fb and ma == mb and np.array_equal(na, nb) and np.array_equal(ta, tb)
    elems = sorted({len(p[1]) - 1 for p in a})
    assert elems == [600, 1200, 1800]
    for (_, nodes, f, _, m) in a:
        lam = wl.C0 / f
        seg = np.hypot(*np.diff(nodes, axis=0).T)
        assert np.max(seg) < 0.11 * lam                         # lambda / 10 meshing
        assert m == int(round(70 * f / wl.C0))
    # shapes differ from each other
    assert not np.array_equal(a[0][1], a[3][1])
