"""HIP-backed replacement of the reference's native sampling kernel loader.

The reference binds ``bor_stream_kernel.{system}-{machine}.so`` with ctypes
(bor_streaming.py:50-76) and calls ``sample_g`` / ``sample_mfie`` /
``sample_ibc`` (bor_stream_kernel.c:26-150) on host arrays.  libghost_bor
exports the same three symbols with identical signatures and semantics
(host pointers, caller-allocated complex128 [nr, np, nxi] outputs, R clamped
at 1e-30), computed on the GPU.  ``load_native()`` returns the library with
the argtypes the reference's loader sets, so a caller written against the
reference's ``_NATIVE`` object works unchanged.

The numeric helpers below wrap those calls for NumPy callers.
"""

from __future__ import annotations

import numpy as np

from . import _lib


def load_native():
    """The ctypes library exposing sample_g / sample_mfie / sample_ibc (bor_streaming.py:50-73)."""
    return _lib.load()


def _c(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


def sample_g(rho_p, z_p, rho_q, z_q, k: float, sin2_tab) -> np.ndarray:
    """e^{-jkR}/(4 pi R) on [nr, np, nxi] (bor_stream_kernel.c:26-48)."""
    rp, zp, rq, zq, s2 = (_c(v) for v in (rho_p, z_p, rho_q, z_q, sin2_tab))
    out = np.empty((rp.size, rq.size, s2.size), dtype=np.complex128)
    _lib.check(_lib.load().ghost_sample_g(rp.size, rq.size, s2.size, _lib.dptr(rp), _lib.dptr(zp),
                                          _lib.dptr(rq), _lib.dptr(zq), float(k), _lib.dptr(s2),
                                          out.ctypes.data_as(_lib._dp)))
    return out


def _brackets(fn, rho_p, z_p, tr_p, tz_p, rho_q, z_q, tr_q, tz_q, k, cx_tab, sx_tab):
    arr = [_c(v) for v in (rho_p, z_p, tr_p, tz_p, rho_q, z_q, tr_q, tz_q, cx_tab, sx_tab)]
    nr, np_, nxi = arr[0].size, arr[4].size, arr[8].size
    outs = [np.empty((nr, np_, nxi), dtype=np.complex128) for _ in range(4)]
    _lib.check(fn(nr, np_, nxi, *[_lib.dptr(a) for a in arr[:8]], float(k), _lib.dptr(arr[8]),
                  _lib.dptr(arr[9]), *[o.ctypes.data_as(_lib._dp) for o in outs]))
    return tuple(outs)


def sample_mfie(rho_p, z_p, tr_p, tz_p, rho_q, z_q, tr_q, tz_q, k: float, cx_tab, sx_tab):
    """The four MFIE bracket integrands (bor_stream_kernel.c:65-106)."""
    return _brackets(_lib.load().ghost_sample_mfie, rho_p, z_p, tr_p, tz_p, rho_q, z_q, tr_q, tz_q, k,
                     cx_tab, sx_tab)


def sample_ibc(rho_p, z_p, tr_p, tz_p, rho_q, z_q, tr_q, tz_q, k: float, cx_tab, sx_tab):
    """The four rotated-PV (IBC) bracket integrands (bor_stream_kernel.c:108-150)."""
    return _brackets(_lib.load().ghost_sample_ibc, rho_p, z_p, tr_p, tz_p, rho_q, z_q, tr_q, tz_q, k,
                     cx_tab, sx_tab)


def estimate_streaming_gb(n_elems: int, m_max: int, formulation: str = "cfie", has_ibc: bool = False,
                          single_blocks: bool = False) -> float:
    """Nodal-block memory of the reference's streamed assembly (bor_streaming.py:424-435)."""
    Nn = float(n_elems + 1)
    per = 8.0 if single_blocks else 16.0
    total = 4.0 * Nn * Nn * (m_max + 1) * per
    if formulation in ("cfie", "mfie"):
        total += 4.0 * Nn * Nn * (2 * m_max + 1) * per
    if has_ibc:
        total += 4.0 * Nn * Nn * (2 * m_max + 1) * per
    return total / 1e9
