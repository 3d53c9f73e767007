"""FP64 CPU restatement of GHOST's body-of-revolution (BoR) MoM hot path.

TEST INFRASTRUCTURE ONLY.  This module is the oracle: the checker that the
HIP path is compared against.  Only ``tests/``, ``__graft_entry__.smoke()``
and ``bench.py``'s ``cpu_baseline`` leg may import it, and only as the
checker / the timed CPU baseline -- never as the product path.

It restates, in NumPy/SciPy, the discretization of the reference solver:

* a1  generatrix + Gauss data ........ bor_kernels.py:38-125, bor_solver.py:230-242
* a2  far modal Green's orders ....... bor_kernels.py:132-167, bor_solver.py:244-267
* a3  far MFIE bracket modes ......... bor_kernels.py:273-331, bor_solver.py:348-370
* a4  integrand samplers ............. bor_stream_kernel.c:26-150
* a5  nodal primitive contraction .... bor_streaming.py:144-155, 352-414
* a6  near element pairs ............. bor_solver.py:54-110, 270-290, 372-387,
                                       bor_kernels.py:170-257, 334-431
* a7  Gc/Gs order relations .......... bor_kernels.py:563-584
* a8  EFIE mode blocks ............... bor_solver.py:130-154, 293-345
* a9  MFIE mode blocks + mass ........ bor_solver.py:389-435
* a10 CFIE mix + axis mask ........... bor_solver.py:611-622, 929-938
* a11 plane-wave RHS ................. bor_solver.py:437-461, 625-645
* a12 dense solve + residual ......... bor_solver.py:744-758
* a13 far field + mode truncation .... bor_solver.py:659-708, 771-796, 987-992
* a14 orchestration defaults ......... bor_solver.py:804-916
* f1  IBC-EFIE (Z_s brackets, Z_s mass, magnetic-current far field)
                                       bor_kernels.py:440-560, bor_solver.py:463-552,
                                       697-707, 850-869
* f3  PMCHWT, homogeneous body ....... bor_solver.py:161-230, 554-575, 647-657,
                                       1019-1094
* f3  PMCHWT, multi-region (cross-surface operators, region system, junction
      reduction; coated, layered, partial-coating and coating-patch
      layouts) ....................... bor_solver.py:1100-1297, 1402-1450,
                                       1805-2176

Pinning (what checks each function, besides golden vectors written from the
reference itself under tests/golden/):
* a1  Gauss-Legendre exactness on polynomials.
* a2  brute-force quadrature of the azimuthal integral; static limit against
      the complete elliptic integral K; G_{-m} = G_m symmetry.
* a3  direct evaluation of the bracket formula + brute-force DFT.
* a5  equality with the dense point-table contraction (a8/a9).
* a6  convergence to adaptive quadrature; golden vectors.
* a11/a13 Bessel identities; sphere invariants; Mie series end to end.
* a12 residual ~1e-15; agreement with LAPACK.
* f1  golden vectors (scalar and tapered Z_s); the impedance Mie series and
      Weston's null (Z_s = eta0) end to end (oracle/mie.py).
* f3  golden vectors (every mode's operators, systems, currents and far
      field for a lossy dielectric and a lossy magnetic body, a coated PEC
      sphere with tensor-grid near pairs, two layers; the reference's
      partial-coating (cap and band) and coating-patch far fields); the dielectric and
      layered Mie series end to end (oracle/mie.py, whose layered series is
      itself checked by the free-space-layer and split-layer identities).
"""

from __future__ import annotations

import math
from dataclasses import dataclass
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np
from scipy import special

C0 = 299_792_458.0
ETA0 = 376.730313668
AXIS_TOL = 1e-12


# ─────────────────────────────────────────────────────────────────────────────
# a1  generatrix and Gauss data
# ─────────────────────────────────────────────────────────────────────────────

@dataclass
class GaussSet:
    """Gauss points over the whole generatrix, element-major (bor_kernels.py:103-125)."""

    elem: np.ndarray
    s: np.ndarray
    w: np.ndarray       # quadrature weight x element length
    rho: np.ndarray
    z: np.ndarray
    trho: np.ndarray
    tz: np.ndarray
    T0: np.ndarray
    T1: np.ndarray
    dRT0: np.ndarray    # d(rho T0)/dt
    dRT1: np.ndarray    # d(rho T1)/dt


class Body:
    """(rho, z) polyline with element tangents (bor_kernels.py:38-81)."""

    def __init__(self, nodes, order: int = 4):
        pts = np.array(nodes, dtype=float)
        if pts.ndim != 2 or pts.shape[1] != 2 or pts.shape[0] < 2:
            raise ValueError("Generatrix needs an (Nn, 2) array of (rho, z) nodes.")
        if np.any(pts[:, 0] < -1e-12):
            raise ValueError("Generatrix rho coordinates must be >= 0.")
        pts[:, 0] = np.maximum(pts[:, 0], 0.0)
        seg = np.diff(pts, axis=0)
        length = np.hypot(seg[:, 0], seg[:, 1])
        if np.any(length <= 0.0):
            raise ValueError("Generatrix has a zero-length element.")
        self.nodes = pts
        self.L = length
        self.trho = seg[:, 0] / length
        self.tz = seg[:, 1] / length
        self.Ne = len(length)
        self.Nn = self.Ne + 1
        self.order = order
        self.g = self._gauss(order)
        self.P = len(self.g.rho)

    def on_axis(self, i: int) -> bool:
        """bor_kernels.py:79-80."""
        return self.nodes[i, 0] <= AXIS_TOL * max(1.0, float(np.max(self.nodes[:, 0])))

    def _gauss(self, order: int) -> GaussSet:
        x, wq = np.polynomial.legendre.leggauss(order)
        u = 0.5 * (x + 1.0)
        wu = 0.5 * wq
        e = np.repeat(np.arange(self.Ne), order)
        s = np.tile(u, self.Ne)
        r0 = self.nodes[:-1]
        dr = np.diff(self.nodes, axis=0)
        rho = r0[e, 0] + s * dr[e, 0]
        z = r0[e, 1] + s * dr[e, 1]
        Le = self.L[e]
        return GaussSet(
            elem=e, s=s, w=np.tile(wu, self.Ne) * Le, rho=rho, z=z,
            trho=self.trho[e], tz=self.tz[e], T0=1.0 - s, T1=s,
            dRT0=(dr[e, 0] * (1.0 - s) - rho) / Le,
            dRT1=(dr[e, 0] * s + rho) / Le)

    def element_points(self, e: int, s: np.ndarray):
        """Points at local coordinates s on element e (bor_solver.py:117-127)."""
        r0, r1 = self.nodes[e], self.nodes[e + 1]
        rho = r0[0] + s * (r1[0] - r0[0])
        z = r0[1] + s * (r1[1] - r0[1])
        drho = r1[0] - r0[0]
        Le = self.L[e]
        return dict(rho=rho, z=z, trho=self.trho[e], tz=self.tz[e],
                    T=np.vstack([1.0 - s, s]),
                    D=np.vstack([(drho * (1.0 - s) - rho) / Le, (drho * s + rho) / Le]),
                    L=Le)

    def nodal_matrices(self) -> Tuple[np.ndarray, np.ndarray]:
        """B_T, B_D [Nn, P] with two nonzeros per column (bor_solver.py:230-242)."""
        g = self.g
        BT = np.zeros((self.Nn, self.P))
        BD = np.zeros((self.Nn, self.P))
        cols = np.arange(self.P)
        BT[g.elem, cols] = g.T0
        BT[g.elem + 1, cols] = g.T1
        BD[g.elem, cols] = g.dRT0
        BD[g.elem + 1, cols] = g.dRT1
        return BT, BD


def sphere_nodes(a: float, n: int) -> np.ndarray:
    """Pole-to-pole semicircle, +z end first (bor_solver.py:2179-2182)."""
    th = np.linspace(0.0, math.pi, n + 1)
    return np.column_stack([a * np.sin(th), a * np.cos(th)])


# ─────────────────────────────────────────────────────────────────────────────
# a2/a3/a4  far azimuthal integrands and their FFT orders
# ─────────────────────────────────────────────────────────────────────────────

def n_xi_efie(osc: float, m_max: int) -> int:
    """bor_kernels.py:149-152 / bor_streaming.py:109-111."""
    return int(2 ** math.ceil(math.log2(max(64, 4 * (m_max + 2), 6 * (osc + 4)))))


def n_xi_bracket(osc: float, m_max: int) -> int:
    """bor_kernels.py:315-319 / bor_streaming.py:114-116."""
    return int(2 ** math.ceil(math.log2(max(128, 6 * (m_max + 2), 8 * (osc + 4)))))


def xi_grid(n: int) -> np.ndarray:
    """Uniform azimuth offsets xi_l = -pi + 2 pi l / N (bor_kernels.py:153)."""
    return 2.0 * np.pi * np.arange(n) / n - np.pi


def green_samples(rho_p, z_p, rho_q, z_q, k: float, xi, r_min: float = 1e-300):
    """g(xi) = exp(-jkR)/(4 pi R), R^2 = d^2 + 4 rho rho' sin^2(xi/2)
    (bor_kernels.py:154-162; bor_stream_kernel.c:26-48 clamps at 1e-30)."""
    sin2 = np.sin(0.5 * xi) ** 2
    d2 = (rho_p - rho_q) ** 2 + (z_p - z_q) ** 2
    R = np.sqrt(d2[..., None] + (4.0 * rho_p * rho_q)[..., None] * sin2)
    R = np.maximum(R, r_min)
    return np.exp(-1j * k * R) / (4.0 * np.pi * R)


def bracket_samples(rho_p, z_p, tr_p, tz_p, rho_q, z_q, tr_q, tz_q, k: float, xi,
                    r_min: float = 1e-300):
    """The four MFIE brackets -p(R)[(W_u.R)(n.f_v) - (W_u.f_v)(n.R)],
    p(R) = (1+jkR)e^{-jkR}/(4 pi R^3); test point at phi=0, source at
    phi'=-xi (bor_kernels.py:273-306; bor_stream_kernel.c:65-106).
    Point arrays broadcast against each other; xi is appended as last axis."""
    ex = lambda a: np.asarray(a)[..., None]
    cx, sx = np.cos(xi), np.sin(xi)
    Rx = ex(rho_p) - ex(rho_q) * cx
    Ry = ex(rho_q) * sx
    Rz = ex(np.asarray(z_p) - np.asarray(z_q)) + 0.0 * cx
    R = np.maximum(np.sqrt(Rx * Rx + Ry * Ry + Rz * Rz), r_min)
    with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
        p = (1.0 + 1j * k * R) * np.exp(-1j * k * R) / (4.0 * np.pi * R ** 3)
        WtR = ex(tr_p) * Rx + ex(tz_p) * Rz
        nR = -ex(tz_p) * Rx + ex(tr_p) * Rz
        n_t = -ex(np.asarray(tz_p) * tr_q) * cx + ex(np.asarray(tr_p) * tz_q)
        n_f = -ex(tz_p) * sx
        Wt_t = ex(np.asarray(tr_p) * tr_q) * cx + ex(np.asarray(tz_p) * tz_q)
        Wt_f = ex(tr_p) * sx
        Wf_t = -ex(tr_q) * sx
        Wf_f = cx + 0.0 * Rx
        return (-p * (WtR * n_t - Wt_t * nR), -p * (WtR * n_f - Wt_f * nR),
                -p * (Ry * n_t - Wf_t * nR), -p * (Ry * n_f - Wf_f * nR))


def ibc_bracket_samples(rho_p, z_p, tr_p, tz_p, rho_q, z_q, tr_q, tz_q, k: float, xi,
                        r_min: float = 1e-300):
    """The four rotated-PV (IBC) brackets p(R)[(W_u.n_q)(R.f_v) - (W_u.f_v)(R.n_q)]
    (bor_kernels.py:449-478; bor_stream_kernel.c:108-150)."""
    ex = lambda a: np.asarray(a)[..., None]
    cx, sx = np.cos(xi), np.sin(xi)
    Rx = ex(rho_p) - ex(rho_q) * cx
    Ry = ex(rho_q) * sx
    Rz = ex(np.asarray(z_p) - np.asarray(z_q)) + 0.0 * cx
    R = np.maximum(np.sqrt(Rx * Rx + Ry * Ry + Rz * Rz), r_min)
    with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
        p = (1.0 + 1j * k * R) * np.exp(-1j * k * R) / (4.0 * np.pi * R ** 3)
        Wt_n = -ex(np.asarray(tr_p) * tz_q) * cx + ex(np.asarray(tz_p) * tr_q)
        Wf_n = ex(tz_q) * sx
        D = ex(rho_p) * cx - ex(rho_q)
        R_t = ex(tr_q) * D + ex(tz_q) * Rz
        R_f = ex(rho_p) * sx
        R_n = -ex(tz_q) * D + ex(tr_q) * Rz
        Wt_t = ex(np.asarray(tr_p) * tr_q) * cx + ex(np.asarray(tz_p) * tz_q)
        Wt_f = ex(tr_p) * sx
        Wf_t = -ex(tr_q) * sx
        Wf_f = cx + 0.0 * Rx
        return (p * (Wt_n * R_t - Wt_t * R_n), p * (Wt_n * R_f - Wt_f * R_n),
                p * (Wf_n * R_t - Wf_t * R_n), p * (Wf_n * R_f - Wf_f * R_n))


def green_orders_fft(samples: np.ndarray, n_orders: int) -> np.ndarray:
    """G_n = (2 pi/N) e^{j pi n} FFT_l[g_l], n = 0..n_orders-1 (bor_kernels.py:163-167)."""
    N = samples.shape[-1]
    spec = np.fft.fft(samples, axis=-1)
    n = np.arange(n_orders)
    return spec[..., :n_orders] * (np.exp(1j * np.pi * n) * (2.0 * np.pi / N))


def bracket_modes_fft(samples: np.ndarray, m_max: int) -> np.ndarray:
    """K_m for m = -m_max..m_max at centred index m + m_max (bor_kernels.py:321-331)."""
    N = samples.shape[-1]
    spec = np.fft.fft(samples, axis=-1)
    m = np.arange(-m_max, m_max + 1)
    bins = np.where(m >= 0, m, N + m)
    return spec[..., bins] * (np.exp(1j * np.pi * m) * (2.0 * np.pi / N))


# ─────────────────────────────────────────────────────────────────────────────
# a6  near element pairs: graded cells + sinh core + oscillation tail
# ─────────────────────────────────────────────────────────────────────────────

def graded_cells(kind: str, depth: int = 4) -> List[Tuple[float, float, float, float]]:
    """Quadtree cells of [0,1]^2 refined toward s = s' ('diag') or toward a
    shared corner ('cornerAB': s = A, s' = B) (bor_solver.py:54-79)."""
    out: List[Tuple[float, float, float, float]] = []
    if kind == "diag":
        def hit(a0, a1, b0, b1):
            return not (a1 <= b0 or b1 <= a0)
    else:
        ca = float(kind[6])
        cb = float(kind[7])

        def hit(a0, a1, b0, b1):
            return a0 <= ca <= a1 and b0 <= cb <= b1

    def split(a0, a1, b0, b1, lev):
        if lev >= depth or not hit(a0, a1, b0, b1):
            out.append((a0, a1, b0, b1))
            return
        am, bm = 0.5 * (a0 + a1), 0.5 * (b0 + b1)
        for (x0, x1) in ((a0, am), (am, a1)):
            for (y0, y1) in ((b0, bm), (bm, b1)):
                split(x0, x1, y0, y1, lev + 1)

    split(0.0, 1.0, 0.0, 1.0, 0)
    return out


_CELLS: Dict[Tuple[str, int], Tuple[np.ndarray, np.ndarray, np.ndarray]] = {}


def cell_points(kind: str, order: int = 4):
    """Tensor Gauss points on each graded cell: `order` on the test axis,
    `order+1` on the source axis so points never coincide (bor_solver.py:85-110)."""
    key = (kind, order)
    if key not in _CELLS:
        x, wx = np.polynomial.legendre.leggauss(order)
        y, wy = np.polynomial.legendre.leggauss(order + 1)
        ux, wux = 0.5 * (x + 1.0), 0.5 * wx
        uy, wuy = 0.5 * (y + 1.0), 0.5 * wy
        S, SP, W = [], [], []
        for (a0, a1, b0, b1) in graded_cells(kind):
            ha, hb = a1 - a0, b1 - b0
            S.append(np.repeat(a0 + ux * ha, len(uy)))
            SP.append(np.tile(b0 + uy * hb, len(ux)))
            W.append(np.outer(wux * ha, wuy * hb).ravel())
        _CELLS[key] = (np.concatenate(S), np.concatenate(SP), np.concatenate(W))
    return _CELLS[key]


def near_kind(e: int, f: int) -> str:
    """bor_solver.py:276-283."""
    if e == f:
        return "diag"
    return "corner10" if f == e + 1 else "corner01"


def _tail_count(k: float, a: np.ndarray, m_max: int) -> int:
    """Gauss order of the [xi0, pi] tail (bor_kernels.py:238-240)."""
    osc = float(np.max(abs(k) * a)) / math.pi + (m_max + 2)
    return int(min(1024, max(64, math.ceil(4.0 * osc))))


def _core_nodes(d, a, s0, order: int = 48):
    """sinh-graded nodes in s = sin(xi/2) on [0, s0]; returns (xi, weight dxi)
    (bor_kernels.py:216-230)."""
    x, wx = np.polynomial.legendre.leggauss(order)
    u, wu = 0.5 * (x + 1.0), 0.5 * wx
    vmax = np.arcsinh((a / d) * s0)
    v = u[None, :] * vmax[:, None]
    wv = wu[None, :] * vmax[:, None]
    s = np.minimum((d / a)[:, None] * np.sinh(v), 1.0)
    xi = 2.0 * np.arcsin(s)
    dsdv = (d / a)[:, None] * np.cosh(v)
    return xi, wv * 2.0 * dsdv / np.sqrt(np.maximum(1.0 - s * s, 1e-15))


def _tail_nodes(s0, n_tail: int):
    """Plain Gauss on [xi0, pi] (bor_kernels.py:241-246)."""
    x, wx = np.polynomial.legendre.leggauss(n_tail)
    xi0 = 2.0 * np.arcsin(np.minimum(s0, 1.0))
    span = np.pi - xi0
    return (xi0[:, None] + 0.5 * (x + 1.0)[None, :] * span[:, None],
            0.5 * wx[None, :] * span[:, None])


def green_orders_near(rho_p, z_p, rho_q, z_q, k: float, m_max: int) -> np.ndarray:
    """G_n, n = 0..m_max+1, for near point pairs (bor_kernels.py:170-257).
    Axis pairs (4 rho rho' <= 1e-30) only have n = 0, equal to 2 pi g(d)."""
    rho_p, rho_q = np.atleast_1d(rho_p).astype(float), np.atleast_1d(rho_q).astype(float)
    z_p, z_q = np.atleast_1d(z_p).astype(float), np.atleast_1d(z_q).astype(float)
    d2 = (rho_p - rho_q) ** 2 + (z_p - z_q) ** 2
    rr4 = 4.0 * rho_p * rho_q
    n = np.arange(m_max + 2)
    out = np.zeros((rho_p.size, m_max + 2), dtype=np.complex128)
    axis = rr4 <= 1e-30
    if np.any(axis):
        R0 = np.sqrt(d2[axis])
        out[axis, 0] = 2.0 * np.pi * np.exp(-1j * k * R0) / (4.0 * np.pi * np.maximum(R0, 1e-300))
    idx = np.flatnonzero(~axis)
    if idx.size == 0:
        return out
    d = np.sqrt(np.maximum(d2[idx], 1e-300))
    a = np.sqrt(rr4[idx])
    s0 = np.minimum(0.25, 20.0 * d / a)
    xc, wc = _core_nodes(d, a, s0)
    Rc = _core_R(d, a, s0)   # R from the clipped node s itself (s = sin(xi/2))
    xt, wt = _tail_nodes(s0, _tail_count(k, a, m_max))
    Rt = np.sqrt(d[:, None] ** 2 + (a[:, None] * np.sin(0.5 * xt)) ** 2)
    acc = np.zeros((idx.size, m_max + 2), dtype=np.complex128)
    for xi, wq, R in ((xc, wc, Rc), (xt, wt, Rt)):
        gw = np.exp(-1j * k * R) / (4.0 * np.pi * R) * wq
        c = np.cos(xi[:, :, None] * n[None, None, :])
        acc += 2.0 * (np.einsum("pv,pvm->pm", gw.real, c) + 1j * np.einsum("pv,pvm->pm", gw.imag, c))
    out[idx] = acc
    return out


def _core_R(d, a, s0, order: int = 48):
    """R at the core nodes, R^2 = d^2 + (a s)^2 with s the clipped sinh node
    (bor_kernels.py:224-226)."""
    x, _ = np.polynomial.legendre.leggauss(order)
    u = 0.5 * (x + 1.0)
    vmax = np.arcsinh((a / d) * s0)
    s = np.minimum((d / a)[:, None] * np.sinh(u[None, :] * vmax[:, None]), 1.0)
    return np.sqrt(d[:, None] ** 2 + (a[:, None] * s) ** 2)


def _pm_projection(Fp, Fm, wq, xi, m):
    """int_0^pi [F(+xi) e^{-jm xi} + F(-xi) e^{+jm xi}] dxi over the node set
    (bor_kernels.py:334-357)."""
    arg = xi[:, :, None] * m[None, None, :]
    c, s = np.cos(arg), np.sin(arg)
    res = []
    for fp, fm in zip(Fp, Fm):
        S = (fp + fm) * wq
        D = (fp - fm) * wq
        pc = np.einsum("pv,pvm->pm", S.real, c) + 1j * np.einsum("pv,pvm->pm", S.imag, c)
        ps = np.einsum("pv,pvm->pm", D.real, s) + 1j * np.einsum("pv,pvm->pm", D.imag, s)
        res.append(pc - 1j * ps)
    return res


def bracket_modes_near(rho_p, z_p, tr_p, tz_p, rho_q, z_q, tr_q, tz_q, k: float,
                       m_max: int, ibc: bool = False):
    """Modal brackets K_m, m = -m_max..m_max, on near point pairs, with the
    same core + tail grid mirrored to -xi (bor_kernels.py:360-431, 515-560)."""
    rho_p, rho_q = np.atleast_1d(rho_p).astype(float), np.atleast_1d(rho_q).astype(float)
    z_p, z_q = np.atleast_1d(z_p).astype(float), np.atleast_1d(z_q).astype(float)
    tr_p = np.broadcast_to(np.asarray(tr_p, float), rho_p.shape)
    tz_p = np.broadcast_to(np.asarray(tz_p, float), rho_p.shape)
    tr_q = np.broadcast_to(np.asarray(tr_q, float), rho_q.shape)
    tz_q = np.broadcast_to(np.asarray(tz_q, float), rho_q.shape)
    d2 = (rho_p - rho_q) ** 2 + (z_p - z_q) ** 2
    rr4 = 4.0 * rho_p * rho_q
    d = np.sqrt(np.maximum(d2, 1e-300))
    a = np.sqrt(np.maximum(rr4, 1e-300))
    s0 = np.minimum(0.25, 20.0 * d / np.maximum(a, 1e-300))
    s0 = np.where(rr4 <= 1e-30, 1.0, s0)
    xc, wc = _core_nodes(d, a, s0)
    xt, wt = _tail_nodes(s0, _tail_count(k, a, m_max))
    xi = np.concatenate([xc, xt], axis=1)
    wq = np.concatenate([wc, wt], axis=1)
    fn = ibc_bracket_samples if ibc else bracket_samples

    def on_grid(x):
        # per-pair xi grids: evaluate with the pair axis leading
        return _bracket_grid(fn, rho_p, z_p, tr_p, tz_p, rho_q, z_q, tr_q, tz_q, k, x)

    Fp = on_grid(xi)
    Fm = on_grid(-xi)
    return _pm_projection(Fp, Fm, wq, xi, np.arange(-m_max, m_max + 1))


def _bracket_grid(fn, rho_p, z_p, tr_p, tz_p, rho_q, z_q, tr_q, tz_q, k, xi):
    """Evaluate a bracket sampler on a per-pair grid xi[n_pairs, nq]."""
    n, nq = xi.shape
    rep = lambda v: np.repeat(v, nq)
    flat = fn(rep(rho_p), rep(z_p), rep(tr_p), rep(tz_p), rep(rho_q), rep(z_q),
              rep(tr_q), rep(tz_q), k, xi.reshape(-1, 1))
    return tuple(F.reshape(n, nq) for F in flat)


# ─────────────────────────────────────────────────────────────────────────────
# a5/a8/a9  far nodal blocks (streamed primitive contraction)
# ─────────────────────────────────────────────────────────────────────────────

EFIE_PRIMS = (("r", "r"), ("z", "z"), ("r", "1"), ("1", "r"), ("1", "1"),
              ("d", "d"), ("d", "s"), ("s", "d"), ("s", "s"))


def nodal_weights(g: GaussSet) -> Dict[str, np.ndarray]:
    """Per-point weights [2, P] for the two nodes of each point's element
    (bor_streaming.py:145-155)."""
    wr = g.w * g.rho
    return {
        "r": np.stack([g.T0 * wr * g.trho, g.T1 * wr * g.trho]),
        "z": np.stack([g.T0 * wr * g.tz, g.T1 * wr * g.tz]),
        "1": np.stack([g.T0 * wr, g.T1 * wr]),
        "s": np.stack([g.T0 * g.w, g.T1 * g.w]),
        "d": np.stack([g.dRT0 * g.w, g.dRT1 * g.w]),
    }


def _contract(K, Lw, Rw, e0, ne_rows, Ne, order, out):
    """out[n, i, j] += sum over point pairs of L[a, p] K[p, q, n] R[b, q] for
    the two nodes a, b of each point's element (bor_streaming.py:352-368)."""
    Kr = K.reshape(ne_rows, order, Ne, order, -1)
    Lr = Lw.reshape(2, ne_rows, order)
    Rr = Rw.reshape(2, Ne, order)
    M = np.einsum("aeg,egfhm->aefhm", Lr, Kr, optimize=True)
    M = np.einsum("aefhm,bfh->maebf", M, Rr, optimize=True)
    out[:, e0:e0 + ne_rows, 0:Ne] += M[:, 0, :, 0, :]
    out[:, e0:e0 + ne_rows, 1:Ne + 1] += M[:, 0, :, 1, :]
    out[:, e0 + 1:e0 + ne_rows + 1, 0:Ne] += M[:, 1, :, 0, :]
    out[:, e0 + 1:e0 + ne_rows + 1, 1:Ne + 1] += M[:, 1, :, 1, :]


def _zero_near_rows(K, e0, ne_rows, Ne, order):
    """Zero |e - f| <= 1 entries of a [rows, P, ...] tile (bor_streaming.py:343-349)."""
    for e in range(e0, e0 + ne_rows):
        c0 = max(0, e - 1) * order
        c1 = min(Ne, e + 2) * order
        K[(e - e0) * order:(e - e0 + 1) * order, c0:c1] = 0.0


class FarBlocks:
    """EFIE primitives per order [9, m_max+2, Nn, Nn] and MFIE bracket blocks
    per mode [4, 2m_max+1, Nn, Nn] from far point pairs (|e-f| >= 2)."""

    def __init__(self, body: Body, k: float, m_max: int, n_e: int, n_b: int,
                 efie: bool = True, mfie: bool = True, rows_per_tile: int = 8,
                 zs_pt: Optional[np.ndarray] = None):
        """zs_pt (per Gauss point, complex): also the IBC rotated-PV bracket
        blocks with Z_s applied at the source point (bor_solver.py:463-478,
        508-518; bor_streaming.py:153-155, 255-258)."""
        g = body.g
        Ne, Nn, go = body.Ne, body.Nn, body.order
        lv = nodal_weights(g)
        self.m_max = m_max
        self.k = k
        self.E = np.zeros((9, m_max + 2, Nn, Nn), dtype=np.complex128) if efie else None
        self.K = np.zeros((4, 2 * m_max + 1, Nn, Nn), dtype=np.complex128) if mfie else None
        self.KI = None
        if zs_pt is not None:
            self.KI = np.zeros((4, 2 * m_max + 1, Nn, Nn), dtype=np.complex128)
            rw_ibc = lv["1"] * np.asarray(zs_pt, dtype=np.complex128)[None, :]
        xe, xb = xi_grid(n_e), xi_grid(n_b)
        for e0 in range(0, Ne, rows_per_tile):
            ne_rows = min(rows_per_tile, Ne - e0)
            rows = slice(e0 * go, (e0 + ne_rows) * go)
            P_ = lambda v: v[rows][:, None]
            Q_ = lambda v: v[None, :]
            if efie:
                G = green_orders_fft(green_samples(P_(g.rho), P_(g.z), Q_(g.rho), Q_(g.z), k, xe),
                                     m_max + 2)
                _zero_near_rows(G, e0, ne_rows, Ne, go)
                for ci, (lx, rx) in enumerate(EFIE_PRIMS):
                    _contract(G, lv[lx][:, rows], lv[rx], e0, ne_rows, Ne, go, self.E[ci])
            if mfie:
                Fs = bracket_samples(P_(g.rho), P_(g.z), P_(g.trho), P_(g.tz),
                                     Q_(g.rho), Q_(g.z), Q_(g.trho), Q_(g.tz), k, xb)
                for uv in range(4):
                    Km = bracket_modes_fft(Fs[uv], m_max) * (2.0 * np.pi)
                    _zero_near_rows(Km, e0, ne_rows, Ne, go)
                    _contract(Km, lv["1"][:, rows], lv["1"], e0, ne_rows, Ne, go, self.K[uv])
            if self.KI is not None:
                Fs = ibc_bracket_samples(P_(g.rho), P_(g.z), P_(g.trho), P_(g.tz),
                                         Q_(g.rho), Q_(g.z), Q_(g.trho), Q_(g.tz), k, xb)
                for uv in range(4):
                    Km = bracket_modes_fft(Fs[uv], m_max) * (2.0 * np.pi)
                    _zero_near_rows(Km, e0, ne_rows, Ne, go)
                    _contract(Km, lv["1"][:, rows], rw_ibc, e0, ne_rows, Ne, go, self.KI[uv])

    def efie(self, m: int):
        """EFIE blocks (no C factor) for mode m (bor_streaming.py:388-407)."""
        am, k = abs(m), self.k
        E = self.E
        lo, hi = abs(am - 1), am + 1
        ztt = 0.5 * (E[0, lo] + E[0, hi]) + E[1, am] - E[5, am] / k ** 2
        ztf = (E[2, lo] - E[2, hi]) / 2j - (1j * am / k ** 2) * E[6, am]
        zft = -(E[3, lo] - E[3, hi]) / 2j + (1j * am / k ** 2) * E[7, am]
        zff = 0.5 * (E[4, lo] + E[4, hi]) - (am ** 2 / k ** 2) * E[8, am]
        if m < 0:
            ztf, zft = -ztf, -zft
        return ztt.copy(), ztf, zft, zff.copy()

    def mfie(self, m: int):
        i = m + self.m_max
        return tuple(self.K[uv, i].copy() for uv in range(4))

    def ibc(self, m: int):
        """IBC bracket blocks (tt, tf, ft, ff) of mode m, Z_s at the source, 2 pi included."""
        i = m + self.m_max
        return tuple(self.KI[uv, i].copy() for uv in range(4))


def table_efie_blocks(body: Body, k: float, m: int, G_table: np.ndarray):
    """Dense point-table contraction (bor_solver.py:130-154): the equivalence
    check for FarBlocks.efie.  G_table is [P, P, >= |m|+2] with near pairs zeroed."""
    g = body.g
    BT, BD = body.nodal_matrices()
    am = abs(m)
    G = G_table[..., am]
    Gc = 0.5 * (G_table[..., abs(am - 1)] + G_table[..., am + 1])
    Gs = (G_table[..., abs(am - 1)] - G_table[..., am + 1]) / 2j
    if m < 0:
        Gs = -Gs
    rrw = (g.rho * g.w)[:, None] * (g.rho * g.w)[None, :]
    Ktt = rrw * (np.outer(g.trho, g.trho) * Gc + np.outer(g.tz, g.tz) * G)
    Ksc = np.outer(g.w, g.w) * G
    Ktf = rrw * (g.trho[:, None] * Gs)
    Kft = -rrw * (g.trho[None, :] * Gs)
    Kff = rrw * Gc
    ztt = BT @ Ktt @ BT.T - (BD @ Ksc @ BD.T) / k ** 2
    ztf = BT @ Ktf @ BT.T - (1j * m / k ** 2) * (BD @ Ksc @ BT.T)
    zft = BT @ Kft @ BT.T + (1j * m / k ** 2) * (BT @ Ksc @ BD.T)
    zff = BT @ Kff @ BT.T - (m ** 2 / k ** 2) * (BT @ Ksc @ BT.T)
    return ztt, ztf, zft, zff


def far_green_table(body: Body, k: float, m_max: int, n_e: int) -> np.ndarray:
    """[P, P, m_max+2] far table with |e-f| <= 1 zeroed (bor_solver.py:244-267)."""
    g = body.g
    G = green_orders_fft(green_samples(g.rho[:, None], g.z[:, None], g.rho[None, :],
                                       g.z[None, :], k, xi_grid(n_e)), m_max + 2)
    near = np.abs(g.elem[:, None] - g.elem[None, :]) <= 1
    G[near] = 0.0
    return G


# ─────────────────────────────────────────────────────────────────────────────
# a6  near-band Galerkin blocks
# ─────────────────────────────────────────────────────────────────────────────

def near_pairs(Ne: int) -> List[Tuple[int, int]]:
    """Element pairs (e, f), |e-f| <= 1, in the reference's loop order (bor_solver.py:305-308)."""
    return [(e, f) for e in range(Ne) for f in (e - 1, e, e + 1) if 0 <= f < Ne]


class NearBlocks:
    """Per near element pair, 2x2 nodal blocks for every mode:
    efie[(e,f)] -> [4 (tt,tf,ft,ff), 2m_max+1, 2, 2] (no C factor);
    mfie[(e,f)] -> [4, 2m_max+1, 2, 2] (2 pi factor included)."""

    def __init__(self, body: Body, k: float, m_max: int, efie: bool = True,
                 mfie: bool = True, zs_elem: Optional[np.ndarray] = None):
        """zs_elem (per element, complex): also ibc[(e,f)], the IBC brackets
        with the source element's Z_s (bor_solver.py:480-495, 519-530)."""
        self.m_max = m_max
        self.efie: Dict[Tuple[int, int], np.ndarray] = {}
        self.mfie: Dict[Tuple[int, int], np.ndarray] = {}
        self.ibc: Dict[Tuple[int, int], np.ndarray] = {}
        ms = np.arange(-m_max, m_max + 1)
        for (e, f) in near_pairs(body.Ne):
            s, sp, w = cell_points(near_kind(e, f))
            p = body.element_points(e, s)
            q = body.element_points(f, sp)
            W = w * p["L"] * q["L"]
            rr = p["rho"] * q["rho"] * W
            Tp, Tq, Dp, Dq = p["T"], q["T"], p["D"], q["D"]
            if efie:
                Gn = green_orders_near(p["rho"], p["z"], q["rho"], q["z"], k, m_max)
                blk = np.zeros((4, len(ms), 2, 2), dtype=np.complex128)
                for mi, m in enumerate(ms):
                    am = abs(m)
                    G = Gn[:, am]
                    Gc = 0.5 * (Gn[:, abs(am - 1)] + Gn[:, am + 1])
                    Gs = (Gn[:, abs(am - 1)] - Gn[:, am + 1]) / 2j
                    if m < 0:
                        Gs = -Gs
                    ktt = rr * (p["trho"] * q["trho"] * Gc + p["tz"] * q["tz"] * G)
                    ksc = W * G
                    ktf = rr * (p["trho"] * Gs)
                    kft = -rr * (q["trho"] * Gs)
                    kff = rr * Gc
                    bil = lambda A, v, B: np.einsum("ip,p,jp->ij", A, v, B)
                    blk[0, mi] = bil(Tp, ktt, Tq) - bil(Dp, ksc, Dq) / k ** 2
                    blk[1, mi] = bil(Tp, ktf, Tq) - (1j * m / k ** 2) * bil(Dp, ksc, Tq)
                    blk[2, mi] = bil(Tp, kft, Tq) + (1j * m / k ** 2) * bil(Tp, ksc, Dq)
                    blk[3, mi] = bil(Tp, kff, Tq) - (m ** 2 / k ** 2) * bil(Tp, ksc, Tq)
                self.efie[(e, f)] = blk
            if mfie:
                Kn = bracket_modes_near(p["rho"], p["z"], p["trho"], p["tz"],
                                        q["rho"], q["z"], q["trho"], q["tz"], k, m_max)
                blk = np.zeros((4, len(ms), 2, 2), dtype=np.complex128)
                for uv in range(4):
                    blk[uv] = 2.0 * np.pi * np.einsum("ip,pm,jp->mij", Tp, rr[:, None] * Kn[uv], Tq)
                self.mfie[(e, f)] = blk
            if zs_elem is not None and abs(zs_elem[f]) != 0.0:
                Kn = bracket_modes_near(p["rho"], p["z"], p["trho"], p["tz"],
                                        q["rho"], q["z"], q["trho"], q["tz"], k, m_max, ibc=True)
                blk = np.zeros((4, len(ms), 2, 2), dtype=np.complex128)
                for uv in range(4):
                    blk[uv] = 2.0 * np.pi * np.einsum("ip,pm,jp->mij", Tp,
                                                      (rr * zs_elem[f])[:, None] * Kn[uv], Tq)
                self.ibc[(e, f)] = blk


# ─────────────────────────────────────────────────────────────────────────────
# a9-a13  mode systems, excitation, far field, sweep
# ─────────────────────────────────────────────────────────────────────────────

def basis_mask(body: Body, m: int) -> np.ndarray:
    """Active (J_t, J_phi) node bases for mode m (bor_solver.py:611-622)."""
    t = np.ones(body.Nn, dtype=bool)
    f = np.ones(body.Nn, dtype=bool)
    for end in (0, body.Nn - 1):
        if body.on_axis(end):
            t[end] = abs(m) == 1
            f[end] = False
        else:
            t[end] = False
            f[end] = True
    return np.concatenate([t, f])


def _bessel_terms(k, g: GaussSet, m: int, theta_deg: float):
    th = math.radians(theta_deg)
    st, ct = math.sin(th), math.cos(th)
    u = k * g.rho * st
    ph = np.exp(1j * k * ct * g.z)
    jn = lambda n: (1j) ** n * special.jv(n, u)
    return st, ct, ph, jn(m), jn(m - 1), jn(m + 1)


def rhs_efie(body: Body, k: float, m: int, theta_deg: float, pol: str) -> np.ndarray:
    """<W, E_inc> for the plane wave from (sin th, 0, cos th) (bor_solver.py:625-645)."""
    g = body.g
    st, ct, ph, J0, Jm1, Jp1 = _bessel_terms(k, g, m, theta_deg)
    Ic = math.pi * (Jm1 + Jp1)
    Is = (math.pi / 1j) * (Jm1 - Jp1)
    I1 = 2.0 * math.pi * J0
    if pol == "VV":
        et, ef = ct * g.trho * Ic - st * g.tz * I1, -ct * Is
    else:
        et, ef = g.trho * Is, Ic
    BT, _ = body.nodal_matrices()
    base = g.w * g.rho * ph
    return np.concatenate([BT @ (base * et), BT @ (base * ef)])


def rhs_mfie(body: Body, k: float, m: int, theta_deg: float, pol: str) -> np.ndarray:
    """<W, n x H_inc> (bor_solver.py:437-461)."""
    g = body.g
    st, ct, ph, J0, Jm1, Jp1 = _bessel_terms(k, g, m, theta_deg)
    Ic = math.pi * (Jm1 + Jp1)
    Is = (math.pi / 1j) * (Jm1 - Jp1)
    I1 = 2.0 * math.pi * J0
    if pol == "VV":
        et, ef = Ic / ETA0, -(g.trho * Is) / ETA0
    else:
        et, ef = (ct * Is) / ETA0, (ct * g.trho * Ic - st * g.tz * I1) / ETA0
    BT, _ = body.nodal_matrices()
    base = g.w * g.rho * ph
    return np.concatenate([BT @ (base * et), BT @ (base * ef)])


def farfield(body: Body, k: float, m: int, sol: np.ndarray, theta_deg: float,
             zs_pt: Optional[np.ndarray] = None, msol: Optional[np.ndarray] = None):
    """Modal (F_theta, F_phi) of the full-length solution vector
    (bor_solver.py:659-696; the far-field pairs (j_{m+1}, j_{m-1}) are
    reversed relative to the excitation).  zs_pt: the eliminated magnetic
    current M = -Z_s n x J also radiates, M_t = Z_s J_phi, M_phi = -Z_s J_t
    (bor_solver.py:697-707)."""
    g = body.g
    st, ct, ph, J0, Jm1, Jp1 = _bessel_terms(k, g, m, theta_deg)
    Icos = math.pi * (Jp1 + Jm1)
    Isin = (math.pi / 1j) * (Jp1 - Jm1)
    I1 = 2.0 * math.pi * J0
    BT, _ = body.nodal_matrices()
    Nn = body.Nn
    Jt = BT.T @ sol[:Nn]
    Jf = BT.T @ sol[Nn:]
    base = g.w * g.rho * ph
    pref = -1j * k * ETA0 / (4.0 * math.pi)
    proj_th = lambda Xt, Xf: np.sum(base * (Xt * (ct * g.trho * Icos - st * g.tz * I1) + Xf * (-ct * Isin)))
    proj_ph = lambda Xt, Xf: np.sum(base * (Xt * g.trho * Isin + Xf * Icos))
    fth = pref * proj_th(Jt, Jf)
    fph = pref * proj_ph(Jt, Jf)
    Mt = Mf = None
    if zs_pt is not None:
        Mt, Mf = zs_pt * Jf, -zs_pt * Jt
    elif msol is not None:
        # an independent magnetic current on the same bases (PMCHWT, bor_solver.py:700-702)
        Mt, Mf = BT.T @ msol[:Nn], BT.T @ msol[Nn:]
    if Mt is not None:
        pm = 1j * k / (4.0 * math.pi)
        fth = fth - pm * proj_ph(Mt, Mf)
        fph = fph + pm * proj_th(Mt, Mf)
    return fth, fph


def mass_matrix(body: Body, weight: Optional[np.ndarray] = None) -> np.ndarray:
    """2 pi int rho T_i T_j (weight) dt (bor_solver.py:389-397); weight per Gauss point."""
    g = body.g
    BT, _ = body.nodal_matrices()
    wt = g.w * g.rho if weight is None else g.w * g.rho * weight
    return 2.0 * np.pi * (BT * wt[None, :]) @ BT.T


def default_m_max(body: Body, k: float, thetas_deg) -> int:
    """bor_solver.py:842-847."""
    st = float(np.max(np.abs(np.sin(np.radians(np.atleast_1d(thetas_deg))))))
    return int(math.ceil(k * float(np.max(body.nodes[:, 0])) * max(st, 0.05))) + 12


def table_gb(n_elems: int, m_max: int, form: str, gauss_order: int = 4, has_ibc: bool = False) -> float:
    """bor_solver.py:804-817 (double)."""
    P = float(gauss_order * n_elems)
    t = P * P * (m_max + 2) * 16.0
    if form in ("cfie", "mfie"):
        t += 4.0 * P * P * (2 * m_max + 1) * 16.0
    if has_ibc:
        t += 4.0 * P * P * (2 * m_max + 1) * 16.0
    return t / 1e9


def impedance_per_element(body: Body, zs) -> Optional[np.ndarray]:
    """Z_s as a per-element complex array, or None for a PEC body
    (bor_solver.py:850-866: a scalar applies to every element)."""
    if zs is None:
        return None
    z = np.asarray(zs, dtype=np.complex128)
    if z.ndim == 0:
        z = np.full(body.Ne, complex(z))
    elif len(z) != body.Ne:
        raise ValueError("zs array must have one entry per generatrix element.")
    return None if np.all(np.abs(z) == 0.0) else z


def osc_for(body: Body, k: float, streaming: bool) -> float:
    """Kernel oscillation 2|k| max sqrt(rho rho'): over Gauss points on the
    table path (bor_kernels.py:151), over nodes on the streamed path
    (bor_streaming.py:179-183)."""
    rmax = float(np.max(body.nodes[:, 0])) if streaming else float(np.max(body.g.rho))
    return 2.0 * abs(k) * rmax


class ModeSystem:
    """Everything for one (body, frequency): assembled CFIE/EFIE systems for
    m = -m_max..m_max, excitation and far-field projection."""

    def __init__(self, nodes, freq_hz: float, m_max: int, formulation: str = "cfie",
                 cfie_alpha: float = 0.5, streaming: Optional[bool] = None, zs=None):
        self.body = Body(nodes)
        self.k = 2.0 * math.pi * freq_hz / C0
        self.m_max = m_max
        form = formulation.lower()
        self.alpha = float(cfie_alpha) if form == "cfie" else (1.0 if form == "efie" else 0.0)
        # surface impedance (IBC-EFIE, bor_solver.py:850-866, 929-934)
        self.zs_elem = impedance_per_element(self.body, zs)
        self.zs_pt = None if self.zs_elem is None else self.zs_elem[self.body.g.elem]
        if self.zs_elem is not None and form == "cfie":
            raise ValueError("CFIE with a surface impedance is not supported; use formulation='efie'.")
        ibc = self.zs_elem is not None and self.alpha > 0.0
        if streaming is None:
            streaming = table_gb(self.body.Ne, m_max, form, has_ibc=self.zs_elem is not None) > 2.0
        self.streaming = streaming
        osc = osc_for(self.body, self.k, streaming)
        self.n_e = n_xi_efie(osc, m_max)
        self.n_b = n_xi_bracket(osc, m_max)
        efie, mfie = self.alpha > 0.0, self.alpha < 1.0
        self.far = FarBlocks(self.body, self.k, m_max, self.n_e, self.n_b, efie, mfie,
                             zs_pt=self.zs_pt if ibc else None)
        self.near = NearBlocks(self.body, self.k, m_max, efie, mfie,
                               zs_elem=self.zs_elem if ibc else None)
        self.M = mass_matrix(self.body)
        self.Mz = mass_matrix(self.body, self.zs_pt) if ibc else None

    def efie_full(self, m: int) -> np.ndarray:
        """Unmasked Z_EFIE [2Nn, 2Nn] (bor_solver.py:293-345)."""
        Nn, k = self.body.Nn, self.k
        blocks = list(self.far.efie(m))
        mi = m + self.m_max
        for (e, f), blk in self.near.efie.items():
            for b in range(4):
                blocks[b][e:e + 2, f:f + 2] += blk[b, mi]
        C = 1j * k * ETA0 * 2.0 * np.pi
        return C * np.block([[blocks[0], blocks[1]], [blocks[2], blocks[3]]])

    def mfie_full(self, m: int) -> np.ndarray:
        """Unmasked Z_MFIE = (1/2)M - K (bor_solver.py:399-435)."""
        blocks = list(self.far.mfie(m))
        mi = m + self.m_max
        for (e, f), blk in self.near.mfie.items():
            for b in range(4):
                blocks[b][e:e + 2, f:f + 2] += blk[b, mi]
        M = self.M
        return np.block([[0.5 * M - blocks[0], -blocks[1]], [-blocks[2], 0.5 * M - blocks[3]]])

    def ibc_full(self, m: int) -> np.ndarray:
        """Unmasked IBC-EFIE additions (1/2) <W, Z_s J> + K'_Zs [2Nn, 2Nn]
        (bor_solver.py:534-552)."""
        blocks = list(self.far.ibc(m))
        mi = m + self.m_max
        for (e, f), blk in self.near.ibc.items():
            for b in range(4):
                blocks[b][e:e + 2, f:f + 2] += blk[b, mi]
        Mz = self.Mz
        return np.block([[0.5 * Mz + blocks[0], blocks[1]], [blocks[2], 0.5 * Mz + blocks[3]]])

    def system(self, m: int) -> Tuple[np.ndarray, np.ndarray]:
        """Masked A_m and its mask (bor_solver.py:929-938)."""
        n = 2 * self.body.Nn
        Z = np.zeros((n, n), dtype=np.complex128)
        if self.alpha > 0.0:
            Z += self.alpha * self.efie_full(m)
            if self.Mz is not None:
                Z += self.alpha * self.ibc_full(m)
        if self.alpha < 1.0:
            Z += (1.0 - self.alpha) * ETA0 * self.mfie_full(m)
        mask = basis_mask(self.body, m)
        return Z[np.ix_(mask, mask)], mask

    def rhs_full(self, m: int, th: float, pol: str) -> np.ndarray:
        """bor_solver.py:940-946."""
        V = np.zeros(2 * self.body.Nn, dtype=np.complex128)
        if self.alpha > 0.0:
            V += self.alpha * rhs_efie(self.body, self.k, m, th, pol)
        if self.alpha < 1.0:
            V += (1.0 - self.alpha) * ETA0 * rhs_mfie(self.body, self.k, m, th, pol)
        return V

    def rhs_block(self, m: int, thetas, mask) -> np.ndarray:
        """Columns ordered theta-major, then (VV, HH) (bor_solver.py:750-752)."""
        cols = [self.rhs_full(m, th, pol)[mask] for th in thetas for pol in ("VV", "HH")]
        return np.stack(cols, axis=1)

    def solve_mode(self, m: int, thetas):
        """A X = B for one mode; returns (X, B, A, mask, residual) (bor_solver.py:744-756)."""
        A, mask = self.system(m)
        B = self.rhs_block(m, thetas, mask)
        X = np.linalg.solve(A, B)
        bn = float(np.linalg.norm(B))
        res = float(np.linalg.norm(A @ X - B)) / bn if bn > 0 else 0.0
        return X, B, A, mask, res

    def mode_farfield(self, m: int, X: np.ndarray, mask, thetas) -> np.ndarray:
        """dF [2 (VV, HH), n_theta] for one mode (bor_solver.py:759-768)."""
        dF = np.zeros((2, len(thetas)), dtype=np.complex128)
        n = 2 * self.body.Nn
        for it, th in enumerate(thetas):
            for ip in range(2):
                sol = np.zeros(n, dtype=np.complex128)
                sol[mask] = X[:, 2 * it + ip]
                fth, fph = farfield(self.body, self.k, m, sol, th, self.zs_pt)
                dF[ip, it] = fth if ip == 0 else fph
        return dF


def truncate_modes(dF_by_am: Sequence[np.ndarray], mode_tol: float):
    """Sum per-|m| contributions in order, stopping after two consecutive quiet
    |m| (max|dF| < mode_tol max|F|) (bor_solver.py:771-796 with workers=1).
    Returns (F, modes_used)."""
    F = np.zeros_like(dF_by_am[0])
    quiet = 0
    used = 0
    for am, dF in enumerate(dF_by_am):
        F = F + dF
        used = am
        scale = max(float(np.max(np.abs(F))), 1e-30)
        if float(np.max(np.abs(dF))) < mode_tol * scale:
            quiet += 1
            if quiet >= 2:
                break
        else:
            quiet = 0
    return F, used


def solve_bor(nodes, freq_hz: float, thetas_deg, formulation: str = "efie",
              cfie_alpha: float = 0.5, n_modes: Optional[int] = None,
              mode_tol: float = 1e-6, streaming: Optional[bool] = None,
              keep: bool = False, zs=None) -> Dict:
    """FP64 monostatic sweep of a PEC or impedance body (bor_solver.py:820-1004)."""
    thetas = np.atleast_1d(np.asarray(thetas_deg, dtype=float))
    body = Body(nodes)
    k = 2.0 * math.pi * freq_hz / C0
    m_max = default_m_max(body, k, thetas) if n_modes is None else int(n_modes)
    sysm = ModeSystem(nodes, freq_hz, m_max, formulation, cfie_alpha, streaming, zs=zs)
    dFs, res, sols = [], 0.0, {}
    for am in range(m_max + 1):
        dF = np.zeros((2, len(thetas)), dtype=np.complex128)
        for m in ([0] if am == 0 else [am, -am]):
            X, B, A, mask, r = sysm.solve_mode(m, thetas)
            res = max(res, r)
            dF += sysm.mode_farfield(m, X, mask, thetas)
            if keep:
                sols[m] = (X, B, A, mask)
        dFs.append(dF)
    F, used = truncate_modes(dFs, mode_tol)
    out = {
        "theta_deg": thetas.tolist(),
        "sigma_vv": (4.0 * math.pi * np.abs(F[0]) ** 2).tolist(),
        "sigma_hh": (4.0 * math.pi * np.abs(F[1]) ** 2).tolist(),
        "amp_vv": F[0].tolist(), "amp_hh": F[1].tolist(),
        "modes_used": used, "n_unknowns": 2 * body.Nn,
        "formulation": formulation.lower(), "linear_residual": res,
        "m_max": m_max, "n_xi_efie": sysm.n_e, "n_xi_bracket": sysm.n_b,
    }
    if keep:
        out["_modes"] = sols
        out["_dF"] = dFs
    return out


# ─────────────────────────────────────────────────────────────────────────────
# f3  PMCHWT: homogeneous penetrable body (bor_solver.py:161-172, 175-230,
#     554-575, 1019-1094)
# ─────────────────────────────────────────────────────────────────────────────

def causal_medium(eps_r, mu_r):
    """(m, eta_r): refractive index with Im(m) <= 0 and eta_r = mu_r / m
    (bor_solver.py:161-172)."""
    m = np.sqrt(complex(eps_r) * complex(mu_r))
    if m.imag > 0.0:
        m = -m
    if m.real < 0.0:
        m = -m
    return m, complex(mu_r) / m


class MediumOperators:
    """The EFIE operator T (with C = j k eta 2 pi) and the rotated-PV operator
    P of one homogeneous medium, node-based [2Nn, 2Nn], on the table rules
    (bor_solver.py:293-345, 497-575).  medium None is free space."""

    def __init__(self, nodes, freq_hz: float, m_max: int, medium=None):
        self.body = Body(nodes)
        k0 = 2.0 * math.pi * freq_hz / C0
        if medium is None:
            self.k, self.eta = k0, ETA0
        else:
            mi, er = causal_medium(*medium)
            self.k, self.eta = k0 * mi, ETA0 * er
        self.m_max = m_max
        osc = osc_for(self.body, self.k, streaming=False)
        self.n_e, self.n_b = n_xi_efie(osc, m_max), n_xi_bracket(osc, m_max)
        b = self.body
        self.far = FarBlocks(b, self.k, m_max, self.n_e, self.n_b, efie=True, mfie=False,
                             zs_pt=np.ones(b.P, dtype=np.complex128))
        self.near = NearBlocks(b, self.k, m_max, efie=True, mfie=False,
                               zs_elem=np.ones(b.Ne, dtype=np.complex128))

    def T(self, m: int) -> np.ndarray:
        blocks = list(self.far.efie(m))
        mi = m + self.m_max
        for (e, f), blk in self.near.efie.items():
            for q in range(4):
                blocks[q][e:e + 2, f:f + 2] += blk[q, mi]
        C = 1j * self.k * self.eta * 2.0 * np.pi
        return C * np.block([[blocks[0], blocks[1]], [blocks[2], blocks[3]]])

    def P(self, m: int) -> np.ndarray:
        """Columns remapped from the rotated-PV brackets B (unit weight):
        P = [[-Btf, Btt], [-Bff, Bft]] (bor_solver.py:554-575)."""
        B = list(self.far.ibc(m))
        mi = m + self.m_max
        for (e, f), blk in self.near.ibc.items():
            for q in range(4):
                B[q][e:e + 2, f:f + 2] += blk[q, mi]
        return np.block([[-B[1], B[0]], [-B[3], B[2]]])


def pmchwt_system(ext: MediumOperators, inner: MediumOperators, m: int) -> np.ndarray:
    """[[Te + Ti, -eta0 (Pe + Pi)], [eta0 (Pe + Pi), Te + (eta0/eta_i)^2 Ti]]
    for unknowns (J, M / eta0) (bor_solver.py:1047-1056)."""
    Te, Ti = ext.T(m), inner.T(m)
    Ps = ETA0 * (ext.P(m) + inner.P(m))
    r2 = (ETA0 / inner.eta) ** 2
    return np.block([[Te + Ti, -Ps], [Ps, Te + r2 * Ti]])


def solve_bor_dielectric(nodes, freq_hz: float, thetas_deg, eps_r, mu_r=1.0,
                         n_modes: Optional[int] = None, mode_tol: float = 1e-6, keep: bool = False) -> Dict:
    """Monostatic RCS of a closed homogeneous penetrable body by per-mode
    PMCHWT (bor_solver.py:1019-1094): RHS [<W, E_inc>; eta0 <W, H_inc>], the
    far field radiating J and M = eta0 M'."""
    thetas = np.atleast_1d(np.asarray(thetas_deg, dtype=float))
    body = Body(nodes)
    k0 = 2.0 * math.pi * freq_hz / C0
    m_max = default_m_max(body, k0, thetas) if n_modes is None else int(n_modes)
    ext = MediumOperators(nodes, freq_hz, m_max)
    inner = MediumOperators(nodes, freq_hz, m_max, (eps_r, mu_r))
    Nn = body.Nn
    dFs, res, sols = [], 0.0, {}
    for am in range(m_max + 1):
        dF = np.zeros((2, len(thetas)), dtype=np.complex128)
        for m in ([0] if am == 0 else [am, -am]):
            mask = np.tile(basis_mask(body, m), 2)
            A = pmchwt_system(ext, inner, m)[np.ix_(mask, mask)]
            cols = []
            for th in thetas:
                for pol in ("VV", "HH"):
                    hpol = rhs_efie(body, k0, m, th, "HH" if pol == "VV" else "VV")
                    h = -hpol if pol == "VV" else hpol        # eta0 <W, H_inc> (bor_solver.py:647-657)
                    cols.append(np.concatenate([rhs_efie(body, k0, m, th, pol), h])[mask])
            B = np.stack(cols, axis=1)
            X = np.linalg.solve(A, B)
            bn = float(np.linalg.norm(B))
            res = max(res, float(np.linalg.norm(A @ X - B)) / bn if bn > 0 else 0.0)
            for it, th in enumerate(thetas):
                for ip in range(2):
                    full = np.zeros(4 * Nn, dtype=np.complex128)
                    full[mask] = X[:, 2 * it + ip]
                    fth, fph = farfield(body, k0, m, full[:2 * Nn], th, msol=ETA0 * full[2 * Nn:])
                    dF[ip, it] += fth if ip == 0 else fph
            if keep:
                sols[m] = (X, B, A, mask)
        dFs.append(dF)
    F, used = truncate_modes(dFs, mode_tol)
    out = {"theta_deg": thetas.tolist(), "sigma_vv": (4.0 * math.pi * np.abs(F[0]) ** 2).tolist(),
           "sigma_hh": (4.0 * math.pi * np.abs(F[1]) ** 2).tolist(), "amp_vv": F[0].tolist(),
           "amp_hh": F[1].tolist(), "modes_used": used, "n_unknowns": 4 * Nn, "linear_residual": res,
           "formulation": "pmchwt", "m_max": m_max}
    if keep:
        out["_modes"] = sols
    return out


# ─────────────────────────────────────────────────────────────────────────────
# f3  multi-region PMCHWT: cross-surface operators and the region system
#     (bor_solver.py:1100-1297, 1805-2172)
# ─────────────────────────────────────────────────────────────────────────────

def segment_distance(p0, p1, q0, q1) -> float:
    """Least distance between two segments that do not cross: one of the
    four endpoint-to-segment distances (bor_solver.py:1100-1112)."""
    def to_seg(c, a, b):
        ab = b - a
        t = float(np.clip(np.dot(c - a, ab) / max(float(np.dot(ab, ab)), 1e-300), 0.0, 1.0))
        return float(np.hypot(*(c - (a + t * ab))))
    return min(to_seg(p0, q0, q1), to_seg(p1, q0, q1), to_seg(q0, p0, p1), to_seg(q1, p0, p1))


def cross_pair_kinds(bp: Body, bq: Body, near_factor: float = 2.0) -> Dict[Tuple[int, int], Optional[str]]:
    """Near element pairs between two generatrices (bor_solver.py:1131-1164):
    a pair sharing one endpoint (a junction) is 'corner<a><b>' (a, b the
    shared ends), a pair closer than near_factor x the longer element is
    None (re-integrated on a tensor grid); touching elsewhere is an error."""
    diag = max(float(np.max(bp.nodes)) - float(np.min(bp.nodes)),
               float(np.max(bq.nodes)) - float(np.min(bq.nodes)), 1e-9)
    tol = 1e-9 * diag
    kinds: Dict[Tuple[int, int], Optional[str]] = {}
    for e in range(bp.Ne):
        pe = (bp.nodes[e], bp.nodes[e + 1])
        for f in range(bq.Ne):
            qf = (bq.nodes[f], bq.nodes[f + 1])
            shared = [(a, b) for a in (0, 1) for b in (0, 1) if float(np.hypot(*(pe[a] - qf[b]))) <= tol]
            if len(shared) > 1:
                raise ValueError("two surfaces share a whole element")
            d = segment_distance(pe[0], pe[1], qf[0], qf[1])
            if shared:
                kinds[(e, f)] = "corner%d%d" % shared[0]
            elif d <= tol:
                raise ValueError("two surfaces cross or touch away from a junction endpoint")
            elif d < near_factor * max(bp.L[e], bq.L[f]):
                kinds[(e, f)] = None
    return kinds


def tensor_points(order: int = 12):
    """order x order Gauss tensor grid on the unit square, test axis slowest
    (bor_solver.py:1203-1209)."""
    x, w = np.polynomial.legendre.leggauss(order)
    u, w1 = 0.5 * (x + 1.0), 0.5 * w
    S, SP = np.meshgrid(u, u, indexing="ij")
    return S.ravel(), SP.ravel(), np.outer(w1, w1).ravel()


class CrossOperators:
    """EFIE (T) and rotated-PV (P) Galerkin operators between the test bases
    of generatrix bp and the source bases of bq in one homogeneous medium
    (k, eta), node-based [2 Nn_p, 2 Nn_q] (bor_solver.py:1115-1297): far
    point pairs on the uniform-xi tables (n_xi from 2|k| max sqrt(rho rho')
    over the two point sets), near pairs re-integrated with the near
    kernels on corner cells or a tensor grid."""

    def __init__(self, bp: Body, bq: Body, k, eta, m_max: int, near_factor: float = 2.0,
                 near_order: int = 12):
        self.bp, self.bq, self.k, self.eta, self.m_max = bp, bq, k, eta, m_max
        self.kinds = cross_pair_kinds(bp, bq, near_factor)
        gp, gq = bp.g, bq.g
        osc = float(np.max(2.0 * abs(k) * np.sqrt(np.outer(gp.rho, gq.rho))))
        self.n_e, self.n_b = n_xi_efie(osc, m_max), n_xi_bracket(osc, m_max)
        G = green_orders_fft(green_samples(gp.rho[:, None], gp.z[:, None], gq.rho[None, :], gq.z[None, :],
                                           k, xi_grid(self.n_e)), m_max + 2)
        Fs = ibc_bracket_samples(gp.rho[:, None], gp.z[:, None], gp.trho[:, None], gp.tz[:, None],
                                 gq.rho[None, :], gq.z[None, :], gq.trho[None, :], gq.tz[None, :],
                                 k, xi_grid(self.n_b))
        B = [bracket_modes_fft(F, m_max) * (2.0 * np.pi) for F in Fs]
        near = np.zeros((bp.P, bq.P), dtype=bool)
        for (e, f) in self.kinds:
            near[np.ix_(gp.elem == e, gq.elem == f)] = True
        G[near] = 0.0
        for Bu in B:
            Bu[near] = 0.0
        self.G, self.B = G, B
        self.near = {}
        for (e, f), kind in self.kinds.items():
            s, sp, W = cell_points(kind) if kind is not None else tensor_points(near_order)
            p, q = bp.element_points(e, s), bq.element_points(f, sp)
            Gn = green_orders_near(p["rho"], p["z"], q["rho"], q["z"], k, m_max)
            Bn = bracket_modes_near(p["rho"], p["z"], p["trho"], p["tz"], q["rho"], q["z"], q["trho"],
                                    q["tz"], k, m_max, ibc=True)
            self.near[(e, f)] = (W * p["L"] * q["L"], p, q, Gn, Bn)

    def T(self, m: int) -> np.ndarray:
        bp, bq, k = self.bp, self.bq, self.k
        gp, gq = bp.g, bq.g
        BTp, BDp = bp.nodal_matrices()
        BTq, BDq = bq.nodal_matrices()
        am = abs(m)
        sg = -1.0 if m < 0 else 1.0

        def orders(Gt):
            G = Gt[..., am]
            Gc = 0.5 * (Gt[..., abs(am - 1)] + Gt[..., am + 1])
            Gs = sg * (Gt[..., abs(am - 1)] - Gt[..., am + 1]) / 2j
            return G, Gc, Gs

        G, Gc, Gs = orders(self.G)
        rrw = (gp.rho * gp.w)[:, None] * (gq.rho * gq.w)[None, :]
        Ksc = np.outer(gp.w, gq.w) * G
        ztt = BTp @ (rrw * (np.outer(gp.trho, gq.trho) * Gc + np.outer(gp.tz, gq.tz) * G)) @ BTq.T \
            - (BDp @ Ksc @ BDq.T) / k ** 2
        ztf = BTp @ (rrw * gp.trho[:, None] * Gs) @ BTq.T - (1j * m / k ** 2) * (BDp @ Ksc @ BTq.T)
        zft = BTp @ (-rrw * gq.trho[None, :] * Gs) @ BTq.T + (1j * m / k ** 2) * (BTp @ Ksc @ BDq.T)
        zff = BTp @ (rrw * Gc) @ BTq.T - (m ** 2 / k ** 2) * (BTp @ Ksc @ BTq.T)
        bil = lambda A, v, Bm: np.einsum("ip,p,jp->ij", A, v, Bm)
        for (e, f), (w, p, q, Gt, _) in self.near.items():
            G, Gc, Gs = orders(Gt)
            rr = p["rho"] * q["rho"] * w
            Tp, Tq, Dp, Dq = p["T"], q["T"], p["D"], q["D"]
            ksc = w * G
            r, c = slice(e, e + 2), slice(f, f + 2)
            ztt[r, c] += bil(Tp, rr * (p["trho"] * q["trho"] * Gc + p["tz"] * q["tz"] * G), Tq) \
                - bil(Dp, ksc, Dq) / k ** 2
            ztf[r, c] += bil(Tp, rr * p["trho"] * Gs, Tq) - (1j * m / k ** 2) * bil(Dp, ksc, Tq)
            zft[r, c] += bil(Tp, -rr * q["trho"] * Gs, Tq) + (1j * m / k ** 2) * bil(Tp, ksc, Dq)
            zff[r, c] += bil(Tp, rr * Gc, Tq) - (m ** 2 / k ** 2) * bil(Tp, ksc, Tq)
        return 1j * k * self.eta * 2.0 * np.pi * np.block([[ztt, ztf], [zft, zff]])

    def P(self, m: int) -> np.ndarray:
        bp, bq = self.bp, self.bq
        gp, gq = bp.g, bq.g
        BTp, _ = bp.nodal_matrices()
        BTq, _ = bq.nodal_matrices()
        mi = m + self.m_max
        Lw, Rw = BTp * (gp.w * gp.rho)[None, :], BTq * (gq.w * gq.rho)[None, :]
        blk = [Lw @ Bu[..., mi] @ Rw.T for Bu in self.B]
        for (e, f), (w, p, q, _, Bn) in self.near.items():
            rr = p["rho"] * q["rho"] * w
            for uv in range(4):
                blk[uv][e:e + 2, f:f + 2] += 2.0 * np.pi * np.einsum("ip,p,jp->ij", p["T"], rr * Bn[uv][:, mi],
                                                                     q["T"])
        Btt, Btf, Bft, Bff = blk
        return np.block([[-Btf, Btt], [-Bff, Bft]])


class MultiRegion:
    """Surfaces (generatrix, is_conductor) bounding homogeneous regions
    ({"medium": None | (eps, mu), "bounds": [(surface, sigma)], "exterior":
    bool}); unknowns per surface J and, for an interface, M / eta0; the
    system sum over regions of sigma_i sigma_j x [[T, -eta0 P], [eta0 P,
    (eta0/eta_r)^2 T]] blocks, reduced by the junction map Q
    (bor_solver.py:1805-2041)."""

    def __init__(self, surfaces, regions, freq_hz: float, m_max: int, near_factor: float = 2.0,
                 near_order: int = 12):
        self.regions = regions
        self.bodies = [Body(p) for (p, _) in surfaces]
        self.cond = [bool(c) for (_, c) in surfaces]
        ns = len(surfaces)
        self.adj = [[] for _ in range(ns)]
        self.sigma = {}
        for ri, reg in enumerate(regions):
            for (si, sg) in reg["bounds"]:
                self.adj[si].append(ri)
                self.sigma[(ri, si)] = int(sg)
        self.ext = next(ri for ri, r in enumerate(regions) if r.get("exterior"))
        if any(sg != 1 for (_, sg) in regions[self.ext]["bounds"]):
            raise ValueError("surfaces bounding the exterior need sigma = +1 there")
        self.k0 = 2.0 * math.pi * freq_hz / C0
        self.m_max = m_max
        self.media = []
        for reg in regions:
            if reg["medium"] is None:
                self.media.append((self.k0, ETA0))
            else:
                mi, er = causal_medium(*reg["medium"])
                self.media.append((self.k0 * mi, ETA0 * er))
        self.ops = {(si, ri): MediumOperators(surfaces[si][0], freq_hz, m_max, regions[ri]["medium"])
                    for si in range(ns) for ri in self.adj[si]}
        self.X = {}
        for ri, reg in enumerate(regions):
            ids = [si for (si, _) in reg["bounds"]]
            for si in ids:
                for sj in ids:
                    if si != sj:
                        self.X[(ri, si, sj)] = CrossOperators(self.bodies[si], self.bodies[sj], *self.media[ri],
                                                              m_max, near_factor, near_order)
        self.Nn = [b.Nn for b in self.bodies]
        self.off_J, self.off_M, acc = [], [], 0
        for si in range(ns):
            self.off_J.append(acc)
            acc += 2 * self.Nn[si]
            self.off_M.append(None if self.cond[si] else acc)
            acc += 0 if self.cond[si] else 2 * self.Nn[si]
        self.n_full = acc
        self.junctions = self._junctions()

    def _junctions(self):
        """Off-axis generatrix ends that coincide, grouped; the first
        interface of a group leads it (bor_solver.py:1856-1895)."""
        allp = np.vstack([b.nodes for b in self.bodies])
        tol = 1e-8 * max(float(np.ptp(allp[:, 0])) + float(np.ptp(allp[:, 1])), 1e-9)
        groups = []
        for si, b in enumerate(self.bodies):
            for node in (0, b.Nn - 1):
                if b.on_axis(node):
                    continue
                for g in groups:
                    s0, n0 = g[0]
                    if float(np.hypot(*(self.bodies[s0].nodes[n0] - b.nodes[node]))) <= tol:
                        g.append((si, node))
                        break
                else:
                    groups.append([(si, node)])
        for g in groups:
            if len(g) < 2:
                raise ValueError(f"surface {g[0][0]} has a free off-axis end that is no junction")
            lead = next((i for i, (si, _) in enumerate(g) if not self.cond[si]), 0)
            g[0], g[lead] = g[lead], g[0]
            for (si, _) in g[1:]:
                if not set(self.adj[si]) & set(self.adj[g[0][0]]):
                    raise ValueError("a junction member shares no region with its lead")
        return groups

    def reduction(self, m: int):
        """(red_of_full [n_full] (-1: dropped), coef [n_full], n_red): each
        kept full index maps to one reduced index with a real coefficient;
        junction followers tie to their lead's index (bor_solver.py:1908-1986)."""
        members = {x for g in self.junctions for x in g}
        follower = {x for g in self.junctions for x in g[1:]}
        mt_off = {x for g in self.junctions if any(self.cond[s] for (s, _) in g)
                  for x in g if not self.cond[x[0]]}
        red = np.full(self.n_full, -1, dtype=int)
        coef = np.zeros(self.n_full)
        n = 0
        for si, b in enumerate(self.bodies):
            for off, is_m in ((self.off_J[si], False), (self.off_M[si], True)):
                if off is None:
                    continue
                Nn = self.Nn[si]
                t_on, f_on = np.ones(Nn, dtype=bool), np.ones(Nn, dtype=bool)
                for end in (0, Nn - 1):
                    if b.on_axis(end):
                        t_on[end], f_on[end] = abs(m) == 1, False
                    elif (si, end) in follower:
                        t_on[end] = f_on[end] = False
                    elif (si, end) in members:
                        if is_m and (si, end) in mt_off:
                            t_on[end] = False
                    else:
                        t_on[end] = False
                for blk, on in ((0, t_on), (Nn, f_on)):
                    for i in np.flatnonzero(on):
                        red[off + blk + i] = n
                        coef[off + blk + i] = 1.0
                        n += 1
        direction = lambda node: 1 if node != 0 else -1
        for g in self.junctions:
            sl, nl = g[0]
            for (sf, nf) in g[1:]:
                r = next(iter(set(self.adj[sf]) & set(self.adj[sl])))
                sgl, sgf = self.sigma[(r, sl)], self.sigma[(r, sf)]
                ct = -(sgl * direction(nl)) / (sgf * direction(nf))
                cf = sgl / sgf
                pairs = [(self.off_J[sl], self.off_J[sf])]
                if self.off_M[sl] is not None and self.off_M[sf] is not None:
                    pairs.append((self.off_M[sl], self.off_M[sf]))
                for ol, of in pairs:
                    for blk, c in ((0, ct), (1, cf)):
                        lead = red[ol + blk * self.Nn[sl] + nl]
                        if lead >= 0:
                            red[of + blk * self.Nn[sf] + nf] = lead
                            coef[of + blk * self.Nn[sf] + nf] = c
        return red, coef, n

    def Q(self, m: int) -> np.ndarray:
        red, coef, n = self.reduction(m)
        Q = np.zeros((self.n_full, n))
        on = red >= 0
        Q[np.flatnonzero(on), red[on]] = coef[on]
        return Q

    def full_system(self, m: int) -> np.ndarray:
        A = np.zeros((self.n_full, self.n_full), dtype=np.complex128)
        for ri, reg in enumerate(self.regions):
            r2 = (ETA0 / self.media[ri][1]) ** 2
            for (si, sgi) in reg["bounds"]:
                for (sj, sgj) in reg["bounds"]:
                    ss = sgi * sgj
                    if si == sj:
                        T = self.ops[(si, ri)].T(m)
                        P = None if self.cond[si] else self.ops[(si, ri)].P(m)
                    else:
                        T, P = self.X[(ri, si, sj)].T(m), self.X[(ri, si, sj)].P(m)
                    Ji = slice(self.off_J[si], self.off_J[si] + 2 * self.Nn[si])
                    Jj = slice(self.off_J[sj], self.off_J[sj] + 2 * self.Nn[sj])
                    A[Ji, Jj] += ss * T
                    if self.off_M[sj] is not None:
                        Mj = slice(self.off_M[sj], self.off_M[sj] + 2 * self.Nn[sj])
                        A[Ji, Mj] += -ETA0 * ss * P
                    if self.off_M[si] is not None:
                        Mi = slice(self.off_M[si], self.off_M[si] + 2 * self.Nn[si])
                        A[Mi, Jj] += ETA0 * ss * P
                        if self.off_M[sj] is not None:
                            A[Mi, Mj] += r2 * ss * T
        return A

    def full_rhs(self, m: int, th: float, pol: str) -> np.ndarray:
        V = np.zeros(self.n_full, dtype=np.complex128)
        for (si, _) in self.regions[self.ext]["bounds"]:
            b, oj, om = self.bodies[si], self.off_J[si], self.off_M[si]
            V[oj:oj + 2 * b.Nn] = rhs_efie(b, self.k0, m, th, pol)
            if om is not None:
                h = rhs_efie(b, self.k0, m, th, "HH" if pol == "VV" else "VV")
                V[om:om + 2 * b.Nn] = -h if pol == "VV" else h
        return V

    def farfield(self, m: int, x_full: np.ndarray, th: float):
        ft = fp = 0.0
        for (si, _) in self.regions[self.ext]["bounds"]:
            b, oj, om = self.bodies[si], self.off_J[si], self.off_M[si]
            ms = None if om is None else ETA0 * x_full[om:om + 2 * b.Nn]
            a, c = farfield(b, self.k0, m, x_full[oj:oj + 2 * b.Nn], th, msol=ms)
            ft, fp = ft + a, fp + c
        return ft, fp

    def rho_max(self) -> float:
        return max(float(np.max(b.nodes[:, 0])) for b in self.bodies)


def default_modes_multiregion(rho_max: float, k0: float, thetas) -> int:
    """bor_solver.py:2050-2053."""
    st = float(np.max(np.abs(np.sin(np.radians(np.atleast_1d(thetas))))))
    return int(math.ceil(k0 * rho_max * max(st, 0.05))) + 12


def solve_multiregion(sys_: "MultiRegion", thetas_deg, mode_tol: float = 1e-6, keep: bool = False) -> Dict:
    """Per-mode solve of Q^T A Q x = Q^T V, far field of Q x
    (bor_solver.py:2044-2071)."""
    thetas = np.atleast_1d(np.asarray(thetas_deg, dtype=float))
    dFs, res, sols = [], 0.0, {}
    for am in range(sys_.m_max + 1):
        dF = np.zeros((2, len(thetas)), dtype=np.complex128)
        for m in ([0] if am == 0 else [am, -am]):
            Q = sys_.Q(m)
            A = Q.T @ sys_.full_system(m) @ Q
            B = np.stack([Q.T @ sys_.full_rhs(m, th, pol) for th in thetas for pol in ("VV", "HH")], axis=1)
            X = np.linalg.solve(A, B)
            bn = float(np.linalg.norm(B))
            res = max(res, float(np.linalg.norm(A @ X - B)) / bn if bn > 0 else 0.0)
            for it, th in enumerate(thetas):
                for ip in range(2):
                    ft, fp = sys_.farfield(m, Q @ X[:, 2 * it + ip], th)
                    dF[ip, it] += ft if ip == 0 else fp
            if keep:
                sols[m] = (X, B, A)
        dFs.append(dF)
    F, used = truncate_modes(dFs, mode_tol)
    out = {"theta_deg": thetas.tolist(), "sigma_vv": (4.0 * math.pi * np.abs(F[0]) ** 2).tolist(),
           "sigma_hh": (4.0 * math.pi * np.abs(F[1]) ** 2).tolist(), "amp_vv": F[0].tolist(),
           "amp_hh": F[1].tolist(), "modes_used": used, "n_unknowns": sys_.n_full,
           "n_junctions": len(sys_.junctions), "linear_residual": res, "m_max": sys_.m_max}
    if keep:
        out["_modes"] = sols
    return out


def coated_layers(interface_points, core_points, eps_list, mu_list):
    """Surfaces and regions of a PEC core under N full layers, interfaces
    outermost first, media innermost first (bor_solver.py:2102-2131)."""
    N = len(interface_points)
    surfaces = [(p, False) for p in interface_points] + [(core_points, True)]
    regions = [{"medium": None, "bounds": [(0, 1)], "exterior": True}]
    for i in range(N):
        lay = N - i
        regions.append({"medium": (eps_list[lay - 1], mu_list[lay - 1]), "bounds": [(i, -1), (i + 1, 1)]})
    return surfaces, regions


def partial_coating_layout(interface_points, covered_points, bare_pieces, eps_r, mu_r):
    """Surfaces and regions of a PEC body partly under one coating: air
    bounded by the interface and the bare conductor pieces, the coating by the
    interface (reversed) and the covered core; each off-axis end is a
    junction of the three chains (bor_solver.py:1402-1450).  The region
    system with this layout gives the reference's partial-coating far field."""
    surfaces = [(interface_points, False), (covered_points, True)] + [(p, True) for p in bare_pieces]
    regions = [{"medium": None, "exterior": True, "bounds": [(0, 1)] + [(2 + i, 1) for i in range(len(bare_pieces))]},
               {"medium": (eps_r, mu_r), "bounds": [(0, -1), (1, 1)]}]
    return surfaces, regions


def coating_patch_layout(patch_points, mid_covered_points, mid_bare_points, core_points, eps_inner, mu_inner,
                         eps_patch, mu_patch):
    """A second layer (the patch) over part of a fully coated PEC core; the
    inner layer's interface is split into the covered part and the exposed
    part, which meet the patch at dielectric triple junctions
    (bor_solver.py:2134-2176)."""
    surfaces = [(patch_points, False), (mid_covered_points, False), (mid_bare_points, False), (core_points, True)]
    regions = [{"medium": None, "exterior": True, "bounds": [(0, 1), (2, 1)]},
               {"medium": (eps_patch, mu_patch), "bounds": [(0, -1), (1, 1)]},
               {"medium": (eps_inner, mu_inner), "bounds": [(1, -1), (2, -1), (3, 1)]}]
    return surfaces, regions
