"""The no-allocation bracket sampler against the allocating one (run on an MI355X)."""

import ctypes

import numpy as np
import pytest

import bor_oracle as bo
from conftest import pkg

pytestmark = pytest.mark.gpu

SCRATCH_BYTES = 48 << 20


def _inputs(rng, nr, np_, nxi):
    rp, zp = rng.uniform(0.05, 1, nr), rng.uniform(-1, 1, nr)
    rq, zq = rng.uniform(0.05, 1, np_), rng.uniform(-1, 1, np_)
    trp, tzp = rng.uniform(-1, 1, nr), rng.uniform(-1, 1, nr)
    trq, tzq = rng.uniform(-1, 1, np_), rng.uniform(-1, 1, np_)
    xi = bo.xi_grid(nxi)
    return [rp, zp, trp, tzp, rq, zq, trq, tzq], np.cos(xi), np.sin(xi)


def _call(fn, L, nr, np_, nxi, arrs, cx, sx, k):
    outs = [np.full((nr, np_, nxi), np.nan, dtype=np.complex128) for _ in range(4)]
    rc = fn(nr, np_, nxi, *[L.dptr(a) for a in arrs], k, L.dptr(cx), L.dptr(sx),
            *[o.ctypes.data_as(L._dp) for o in outs])
    return rc, outs


def test_noalloc_sampler_equals_allocating_sampler_twice():
    """Two calls in a row, the second smaller: bitwise the allocating
    ghost_sample_mfie's outputs each time (the array keeps the first call's
    contents, which the second must not read)."""
    L = pkg("_lib")
    lib = L.load()
    rng = np.random.default_rng(11)
    for (nr, np_, nxi, k) in ((12, 128, 256, 9.1), (5, 7, 64, 3.3)):
        print(f"[case] noalloc nr={nr} np={np_} nxi={nxi}", flush=True)
        arrs, cx, sx = _inputs(rng, nr, np_, nxi)
        rc0, ref = _call(lib.ghost_sample_mfie, L, nr, np_, nxi, arrs, cx, sx, k)
        rc1, got = _call(lib.ghost_sample_mfie_noalloc, L, nr, np_, nxi, arrs, cx, sx, k)
        assert rc0 == 0 and rc1 == 0, L.load().ghost_last_error()
        for a, b in zip(got, ref):
            assert np.array_equal(a, b)
        oracle = bo.bracket_samples(arrs[0][:, None], arrs[1][:, None], arrs[2][:, None], arrs[3][:, None],
                                    arrs[4][None, :], arrs[5][None, :], arrs[6][None, :], arrs[7][None, :],
                                    k, bo.xi_grid(nxi), r_min=1e-30)
        for a, r in zip(got, oracle):
            assert np.max(np.abs(a - r)) <= 1e-12 * np.max(np.abs(r))


def test_noalloc_sampler_rejects_sizes_beyond_the_array():
    L = pkg("_lib")
    lib = L.load()
    nr, np_, nxi = 64, 1024, 256          # 64 nr np_ nxi = 1 GiB of staging
    assert 64 * nr * np_ * nxi > SCRATCH_BYTES
    small = np.zeros(max(nr, np_, nxi))
    one = np.zeros(1, dtype=np.complex128)
    dp = L._dp
    rc = lib.ghost_sample_mfie_noalloc(nr, np_, nxi, *[L.dptr(small)] * 8, 1.0, L.dptr(small), L.dptr(small),
                                       *[one.ctypes.data_as(dp)] * 4)
    assert rc == 1
    assert "scratch" in lib.ghost_last_error().decode()
