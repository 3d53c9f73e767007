"""The EFIE symmetry that the far fill uses for unordered element pairs.

k_far_fill (csrc/fill.hip) evaluates the per-order EFIE blocks of a far pair
(e, f) once and writes the (f, e) blocks from the same values by lane
exchanges: tt and ff transposed, tf(f, e) = -ft(e, f)^T and ft(f, e) =
-tf(e, f)^T.  Here the oracle's far EFIE blocks (FarBlocks.efie, following
bor_streaming.py:388-407) are checked for exactly those identities on a body
without mirror symmetry, at positive and negative orders, together with the
checks that keep the test from passing vacuously: the tf block is not small,
and the opposite sign of the identity fails.  The near band's singular
quadrature is not symmetric to rounding, which is why the kernel applies the
identity to far pairs only."""

import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "oracle"))

import bor_oracle as bo  # noqa: E402


def lopsided_body(n):
    """A pole-to-pole generatrix that is not symmetric about z = 0."""
    th = np.linspace(0.0, np.pi, n + 1)
    rho = 0.5 * np.sin(th) * (1.0 + 0.3 * np.cos(th))
    return np.column_stack([rho, 0.5 * np.cos(th) + 0.1 * np.cos(th) ** 2])


@pytest.fixture(scope="module")
def system():
    return bo.ModeSystem(lopsided_body(14), 3.5e8, 4, "cfie", 0.5, streaming=False)


@pytest.mark.parametrize("m", [1, 2, 4, -3])
def test_far_efie_blocks_are_symmetric(system, m):
    tt, tf, ft, ff = system.far.efie(m)
    scale = max(abs(tt).max(), abs(ff).max())
    assert abs(tf).max() > 0.05 * scale                 # the identity is not vacuous
    assert abs(tt - tt.T).max() <= 1e-14 * scale
    assert abs(ff - ff.T).max() <= 1e-14 * scale
    assert abs(tf + ft.T).max() <= 1e-14 * scale        # tf(f, e) = -ft(e, f)^T
    assert abs(tf - ft.T).max() > 0.05 * scale          # the other sign fails


def test_order_zero_has_no_cross_blocks(system):
    _, tf, ft, _ = system.far.efie(0)
    assert abs(tf).max() == 0.0 and abs(ft).max() == 0.0


def test_assembled_efie_is_symmetric_only_to_near_quadrature_accuracy(system):
    """The assembled EFIE (far + near band) is symmetric only to the near
    quadrature's accuracy, far above rounding: the far fill must not be
    given the near pairs (|e - f| <= 1), which it never is."""
    Z = system.efie_full(2)
    n = Z.shape[0] // 2
    asym = abs(Z[:n, :n] - Z[:n, :n].T).max() / abs(Z).max()
    assert 1e-12 < asym < 1e-3
