"""The far fill's unordered element-pair schedule (host side in csrc/api.hip
far_fill, wavefront side at the top of k_far_fill in csrc/fill.hip): four
launches over the parity classes (e mod 2, f mod 2) of the pairs e < f with
f - e >= 2, wavefront w of a class decoding its pair from the triangular row
sums S(i) = i (nfp - delta) - i (i - 1) / 2 by bisection.  Each wavefront
writes the 2 x 2 node footprints of (e, f) and of (f, e) without atomics, so
the classes must cover every far pair exactly once and keep all footprints
of one launch disjoint."""

import pytest


def class_pairs(ne, pe, pf):
    nep, nfp = (ne - pe + 1) // 2, (ne - pf + 1) // 2
    if nep <= 0 or nfp <= 0:
        return []
    delta = 1 if pf >= pe else 2
    imax = min(nep, max(0, nfp - delta))
    base = nfp - delta
    npairs = imax * base - imax * (imax - 1) // 2

    def s(i):
        return i * base - i * (i - 1) // 2

    out = []
    for wid in range(max(npairs, 0)):
        lo, hi = 0, imax
        while hi - lo > 1:
            mid = (lo + hi) >> 1
            if s(mid) <= wid:
                lo = mid
            else:
                hi = mid
        out.append((pe + 2 * lo, pf + 2 * (lo + delta + wid - s(lo))))
    return out


@pytest.mark.parametrize("ne", list(range(1, 24)) + [119, 120, 601])
def test_classes_cover_every_far_pair_once_with_disjoint_footprints(ne):
    seen = []
    for pe in (0, 1):
        for pf in (0, 1):
            pairs = class_pairs(ne, pe, pf)
            owner = {}
            for e, f in pairs:
                assert 0 <= e < f < ne and f - e >= 2
                for a, b in ((e, f), (f, e)):
                    for r in (a, a + 1):
                        for c in (b, b + 1):
                            assert owner.setdefault((r, c), (e, f)) == (e, f), (pe, pf, e, f, r, c)
            seen += pairs
    assert len(seen) == len(set(seen))
    assert set(seen) == {(e, f) for e in range(ne) for f in range(e + 2, ne)}
