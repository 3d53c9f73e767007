"""Monostatic RCS (dBsm) as a differentiable function of the generatrix radii.

``rcs_db(offsets, spec)`` returns sigma in dBsm, shape [2 (VV, HH), n_theta],
of the PEC body whose node radii are ``spec.nodes[:, 0] + offsets``.  Its
forward pass is the HIP path in FP64 (fill, batched LU, far field: the
solve of bor_solver.py:820-1004 with every mode kept, mode_tol = 0); its
backward pass is the library's adjoint (ghost_bor_adjoint_analytic): one
transposed solve through the LU factors of the forward pass, then the
fill's, right-hand side's and far-field projection's analytic shape
derivative contracted with the adjoint weights in the kernels, whatever the
number of nodes (``method="fd"``: FP64 central differences of 12 batched
fills instead, the check path).

The reference solver has no gradient; this is the differentiable front end
of the same discretisation.  The azimuth grids (n_xi) are fixed by ``spec``
from the base geometry, so the function is smooth in the offsets (the
reference would re-round them per geometry).  Only CPU tensors are used on
the torch side: the device work belongs to libghost_bor.
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _lib
from .bor_solver import C0, ETA0, discretisation

_DB = 20.0 / math.log(10.0)   # d(10 log10 |F|^2) = _DB Re(conj(F) dF) / |F|^2


@dataclass
class RcsSpec:
    """A body and sweep: base nodes [Nn, 2] (rho, z), frequency, aspects and
    modes |m| <= m_max; CFIE with alpha (1 = EFIE, 0 = MFIE)."""
    nodes: np.ndarray
    freq_hz: float
    thetas_deg: np.ndarray
    m_max: int
    cfie_alpha: float = 0.5
    step: float = 0.0             # "fd" half width [m]; 0: 1e-6 shortest element
    method: str = "analytic"      # adjoint: "analytic" (dual-number fill derivative) or "fd"
    precision: str = "fp64"       # "fp64", or "mixed": FP32 fill and LU under FP64 refinement
                                  # (forward and adjoint solves; analytic adjoint only)
    n_xi: tuple = field(default=(0, 0))

    def __post_init__(self):
        self.nodes = np.ascontiguousarray(self.nodes, dtype=np.float64)
        self.thetas_deg = np.ascontiguousarray(np.atleast_1d(self.thetas_deg), dtype=np.float64)
        self.k = 2.0 * math.pi * self.freq_hz / C0
        if self.n_xi == (0, 0):
            a = self.cfie_alpha
            form = "cfie" if 0.0 < a < 1.0 else ("efie" if a == 1.0 else "mfie")
            _, ne, nb, _ = discretisation(self.nodes, self.k, self.thetas_deg, self.m_max, form)
            self.n_xi = (int(ne), int(nb))


def solve_far_field(spec: RcsSpec, radii: np.ndarray):
    """Mode-summed far field F [2, n_theta] of the body with the given node
    radii, and the solved handle (kept for the adjoint; FP64 or mixed as the
    spec says, mode pairing off)."""
    nodes = spec.nodes.copy()
    nodes[:, 0] = radii
    mixed = spec.precision == "mixed"
    if spec.precision not in ("fp64", "mixed"):
        raise ValueError(f"precision must be 'fp64' or 'mixed', got {spec.precision!r}")
    h = _lib.Handle(_lib.GHOST_MIXED if mixed else _lib.GHOST_FP64, nb=64, residual=False, pairing=False)
    a = spec.cfie_alpha
    h.add_problem(nodes, spec.k, spec.m_max, spec.n_xi[0], spec.n_xi[1], a, (1.0 - a) * ETA0, spec.thetas_deg)
    h.build()
    h.run(-2 if mixed else 0)   # mixed: adaptive refinement (error probe, up to 2 sweeps)
    h.sync()
    F = h.dF(0).sum(axis=0)
    return F, h


def sigma_db(F: np.ndarray) -> np.ndarray:
    return 10.0 * np.log10(4.0 * math.pi * np.abs(F) ** 2)


class BorRcsDb(torch.autograd.Function):
    """sigma_dB [2, n_theta] of the radii spec.nodes[:, 0] + offsets."""

    @staticmethod
    def forward(ctx, offsets: torch.Tensor, spec: RcsSpec):
        off = offsets.detach().cpu().double().numpy()
        if off.shape != (spec.nodes.shape[0],):
            raise ValueError(f"offsets must have shape ({spec.nodes.shape[0]},)")
        F, h = solve_far_field(spec, spec.nodes[:, 0] + off)
        ctx.spec, ctx.F, ctx.handle = spec, F, h
        return torch.from_numpy(sigma_db(F))

    @staticmethod
    def backward(ctx, grad_out: torch.Tensor):
        F, h, spec = ctx.F, ctx.handle, ctx.spec
        w = grad_out.detach().cpu().double().numpy()
        a = w * _DB * np.conj(F) / np.abs(F) ** 2          # dJ = Re sum a dF
        acoef = np.ascontiguousarray(a.T).ravel()           # column 2 theta_index + pol
        g = h.adjoint(acoef, spec.step, spec.method)
        h.close()
        ctx.handle = None
        return torch.from_numpy(g), None


def rcs_db(offsets: torch.Tensor, spec: RcsSpec) -> torch.Tensor:
    """Differentiable sigma_dB [2 (VV, HH), n_theta] (see BorRcsDb)."""
    return BorRcsDb.apply(offsets, spec)


def objective_and_gradient(spec: RcsSpec, weights=None):
    """J = sum weights * sigma_dB (weights [2, n_theta], default all ones) and
    dJ/d rho_node for every node, by one forward and one adjoint pass."""
    F, h = solve_far_field(spec, spec.nodes[:, 0])
    s = sigma_db(F)
    w = np.ones_like(s) if weights is None else np.asarray(weights, dtype=np.float64)
    a = w * _DB * np.conj(F) / np.abs(F) ** 2
    g = h.adjoint(np.ascontiguousarray(a.T).ravel(), spec.step, spec.method)
    h.close()
    return float(np.sum(w * s)), g
