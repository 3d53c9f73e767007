"""The three-way bf16 split behind the FP32 GEMM on the bf16 MFMA
(k_gemm_bf3), restated on the CPU: x = p0 + p1 + p2 exactly for FP32 x,
each part a bf16 value.  With the parts rounded to nearest (|p1| <= 2^-8 |x|,
|p2| <= 2^-16 |x|) the six kept products reproduce x y to 2 x 2^-24
relative, the FP32 product's own rounding bound twice over, and the dropped
products are unbiased; truncated parts (|p1| < 2^-7 |x|, |p2| < 2^-14 |x|)
bound it only by 2^-20 and shrink every product (the reason the kernel
rounds)."""

import numpy as np


def bf16_rn(x):
    """float32 -> nearest bf16 (ties to even), returned as float32."""
    u = np.asarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000).astype(np.uint32)
    return r.view(np.float32)


def bf16_trunc(x):
    u = np.asarray(x, dtype=np.float32).view(np.uint32)
    return (u & np.uint32(0xFFFF0000)).view(np.float32)


def split3(x, rnd):
    p0 = rnd(x)
    r = (x - p0).astype(np.float32)
    p1 = rnd(r)
    p2 = (r - p1).astype(np.float32)
    return p0, p1, p2


def test_split_is_exact_and_parts_are_bf16():
    rng = np.random.default_rng(5)
    x = (rng.standard_normal(200_000) * np.exp2(rng.integers(-60, 60, 200_000))).astype(np.float32)
    for rnd in (bf16_rn, bf16_trunc):
        p0, p1, p2 = split3(x, rnd)
        for p in (p0, p1, p2):   # every part is a bf16 number
            assert np.all(p.view(np.uint32) & 0xFFFF == 0)
        assert np.array_equal(p0.astype(np.float64) + p1 + p2, x.astype(np.float64))


def test_six_products_carry_fp32_accuracy_and_rounding_removes_the_bias():
    rng = np.random.default_rng(6)
    n = 400_000
    x = rng.uniform(0.5, 1.0, n).astype(np.float32) * rng.choice([-1, 1], n).astype(np.float32)
    y = rng.uniform(0.5, 1.0, n).astype(np.float32) * rng.choice([-1, 1], n).astype(np.float32)
    exact = x.astype(np.float64) * y.astype(np.float64)
    rel = {}
    for name, rnd, bound in (("rn", bf16_rn, 2.01 * 2.0 ** -24), ("trunc", bf16_trunc, 2.0 ** -20)):
        a = [p.astype(np.float64) for p in split3(x, rnd)]
        b = [p.astype(np.float64) for p in split3(y, rnd)]
        kept = sum(a[i] * b[j] for i in range(3) for j in range(3) if i + j <= 2)
        err = (kept - exact) / np.abs(exact)
        assert np.max(np.abs(err)) < bound, name
        # signed error relative to the product's own sign: the bias
        rel[name] = float(np.mean(err * np.sign(exact)))
    assert abs(rel["rn"]) < 2e-10
    assert rel["trunc"] < -2e-9   # truncated parts shrink every product
