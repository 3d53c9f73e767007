"""ctypes binding of libghost_bor.so (the gfx950 BoR-MoM hot path).

The library is built in-tree next to this file (``__graft_entry__.build()``).
There is no CPU fallback: if the shared object is missing or cannot be
loaded, every call raises GhostLibraryError.
"""

import ctypes
import os
import threading
from typing import Optional

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libghost_bor.so")

GHOST_FP64 = 0
GHOST_MIXED = 1
GHOST_UNMASKED = 1

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_fp = ctypes.POINTER(ctypes.c_float)
_ci = ctypes.c_int
_cd = ctypes.c_double
_vp = ctypes.c_void_p

# (name, restype, argtypes) of every symbol include/ghost_bor.h declares
SIGNATURES = [
    ("ghost_last_error", ctypes.c_char_p, []),
    ("sample_g", None, [_ci, _ci, _ci, _dp, _dp, _dp, _dp, _cd, _dp, _dp]),
    ("sample_mfie", None, [_ci, _ci, _ci] + [_dp] * 8 + [_cd, _dp, _dp] + [_dp] * 4),
    ("sample_ibc", None, [_ci, _ci, _ci] + [_dp] * 8 + [_cd, _dp, _dp] + [_dp] * 4),
    ("ghost_sample_g", _ci, [_ci, _ci, _ci, _dp, _dp, _dp, _dp, _cd, _dp, _dp]),
    ("ghost_sample_mfie", _ci, [_ci, _ci, _ci] + [_dp] * 8 + [_cd, _dp, _dp] + [_dp] * 4),
    ("ghost_sample_ibc", _ci, [_ci, _ci, _ci] + [_dp] * 8 + [_cd, _dp, _dp] + [_dp] * 4),
    ("ghost_bor_create", _ci, [ctypes.POINTER(_vp), _ci, _ci]),
    ("ghost_bor_destroy", _ci, [_vp]),
    ("ghost_bor_set_residual", _ci, [_vp, _ci]),
    ("ghost_bor_set_pairing", _ci, [_vp, _ci]),
    ("ghost_bor_set_lu_streams", _ci, [_vp, _ci]),
    ("ghost_bor_add_problem", _ci, [_vp, _dp, _ci, _cd, _ci, _ci, _ci, _cd, _cd, _dp, _ci, _ci, _ip]),
    ("ghost_bor_build", _ci, [_vp]),
    ("ghost_bor_fill", _ci, [_vp]),
    ("ghost_bor_solve", _ci, [_vp, _ci]),
    ("ghost_bor_farfield", _ci, [_vp]),
    ("ghost_bor_set_gemm_variant", _ci, [_vp, _ci]),
    ("ghost_bor_set_probe_tol", _ci, [_vp, _cd]),
    ("ghost_bor_refine_history", _ci, [_vp, _dp, _ip]),
    ("ghost_bor_run", _ci, [_vp, _ci]),
    ("ghost_bor_sync", _ci, [_vp]),
    ("ghost_device_synchronize", _ci, []),
    ("ghost_bor_get_dF", _ci, [_vp, _ci, _dp]),
    ("ghost_bor_residual", _ci, [_vp, _dp]),
    ("ghost_bor_system_info", _ci, [_vp, _ci, _ci, _ip, _ip, _ip, _ip]),
    ("ghost_bor_num_systems", _ci, [_vp, _ci, _ip, _ip]),
    ("ghost_bor_get_system", _ci, [_vp, _ci, _ci, _ci, _dp, _dp]),
    ("ghost_bor_info", _ci, [_vp, _ip]),
    ("ghost_bor_timings", _ci, [_vp, _fp]),
    ("ghost_bor_capture", _ci, [_vp, _ci]),
    ("ghost_bor_launch", _ci, [_vp]),
    ("ghost_bor_graph_ms", _ci, [_vp, _fp]),
    ("ghost_bor_set_profile", _ci, [_vp, _ci]),
    ("ghost_bor_kernel_count", _ci, []),
    ("ghost_bor_kernel_name", ctypes.c_char_p, [_ci]),
    ("ghost_bor_kernel_times", _ci, [_vp, _dp, _ip, _dp]),
    ("ghost_bor_adjoint", _ci, [_vp, _dp, _cd, _dp]),
    ("ghost_bor_adjoint_analytic", _ci, [_vp, _dp, _dp]),
    ("ghost_bor_set_objective", _ci, [_vp, _dp]),
    ("ghost_bor_set_nodes", _ci, [_vp, _ci, _dp]),
    ("ghost_bor_capture_adjoint", _ci, [_vp, _ci]),
    ("ghost_bor_get_gradient", _ci, [_vp, _dp]),
    ("ghost_copy_calibration", _ci, [_cd, _fp]),
    ("ghost_bor_add_dense", _ci, [_vp, _ci, _ci, _ip]),
    ("ghost_bor_set_dense", _ci, [_vp, _ci, _dp, _dp]),
    ("ghost_bor_get_dense", _ci, [_vp, _ci, _dp]),
    ("ghost_bor_fill_rhs", _ci, [_vp]),
    ("ghost_bor_set_solution", _ci, [_vp, _ci, _ci, _dp]),
    ("ghost_bor_system_count", _ci, [_vp, _ip]),
    ("ghost_bor_get_pivots", _ci, [_vp, _ci, _ip]),
    ("ghost_bor_set_lu_super", _ci, [_vp, _ci]),
    ("ghost_far_fill_waves", _ci, [_ci, _ci, _ci, _ci]),
    ("ghost_sample_mfie_noalloc", _ci, [_ci, _ci, _ci] + [_dp] * 8 + [_cd, _dp, _dp] + [_dp] * 4),
    ("ghost_bor_set_impedance", _ci, [_vp, _ci, _dp, _cd, _cd]),
    ("ghost_bor_set_medium", _ci, [_vp, _ci, _cd, _cd, _cd]),
    ("ghost_bor_set_interior", _ci, [_vp, _ci, _cd, _cd, _cd, _cd, _ci, _ci]),
    ("ghost_bor_add_operator", _ci, [_vp, _dp, _ci, _cd, _cd, _cd, _cd, _ci, _ci, _ci, _ci, _dp, _ci, _ip]),
    ("ghost_bor_add_cross", _ci, [_vp, _ci, _ci, _cd, _ci, _ip]),
    ("ghost_bor_add_region", _ci, [_vp, _ci, _ci, _dp, _ci, _ip]),
    ("ghost_bor_region_block", _ci, [_vp, _ci, _ci, _ci, _ci, _cd, _cd, _ci, _ci]),
    ("ghost_bor_region_map", _ci, [_vp, _ci, _ci, _ci, _ip, _dp]),
    ("ghost_bor_region_exterior", _ci, [_vp, _ci, _ci, _ci, _ci]),
    ("ghost_bor_region_get_dF", _ci, [_vp, _ci, _dp]),
    ("ghost_bor_region_system", _ci, [_vp, _ci, _ci, _ci, _ip, _dp, _dp]),
    ("ghost_device_name", _ci, [ctypes.c_char_p, _ci]),
]


def device_name() -> str:
    """'name|gfx arch|cuN' of the current HIP device (from the library's own runtime)."""
    buf = ctypes.create_string_buffer(256)
    check(load().ghost_device_name(buf, 256))
    return buf.value.decode(errors="replace")


def device_synchronize() -> None:
    check(load().ghost_device_synchronize())


class GhostLibraryError(RuntimeError):
    """libghost_bor is missing, failed to load, or a call returned an error."""


_LIB = None
_LOCK = threading.Lock()


def load(path: str = LIB_PATH):
    """Load (once) and return the ctypes library; raises if it is absent."""
    global _LIB
    with _LOCK:
        if _LIB is not None:
            return _LIB
        if not os.path.exists(path):
            raise GhostLibraryError(
                f"{path} not found: build it first with __graft_entry__.build() "
                "(there is no CPU fallback for the BoR hot path)")
        try:
            lib = ctypes.CDLL(path)
        except OSError as exc:
            raise GhostLibraryError(f"cannot load {path}: {exc}") from exc
        for name, res, args in SIGNATURES:
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = lib
        return lib


def check(rc: int) -> None:
    if rc != 0:
        msg = load().ghost_last_error().decode(errors="replace")
        if rc == 4:
            raise MemoryError(msg)
        if rc == 1:
            raise ValueError(msg)
        raise GhostLibraryError(f"libghost_bor error {rc}: {msg}")


def copy_calibration(nbytes: float) -> float:
    """Milliseconds of one plain device copy of nbytes (see ghost_copy_calibration)."""
    ms = ctypes.c_float(0.0)
    check(load().ghost_copy_calibration(float(nbytes), ctypes.byref(ms)))
    return float(ms.value)


def dptr(a: np.ndarray):
    return a.ctypes.data_as(_dp)


def iptr(a: np.ndarray):
    return a.ctypes.data_as(_ip)


class Handle:
    """Owner of one ghost_bor handle (a batch of problems on the current device)."""

    def __init__(self, precision: int = GHOST_FP64, nb: int = 64, residual: bool = True, pairing: bool = True):
        """pairing: solve mode -m with mode m's LU factors (Z_{-m} = S Z_m S);
        a handle that will run the adjoint needs it off."""
        self.lib = load()
        self._h = _vp()
        check(self.lib.ghost_bor_create(ctypes.byref(self._h), int(precision), int(nb)))
        check(self.lib.ghost_bor_set_residual(self._h, 1 if residual else 0))
        check(self.lib.ghost_bor_set_pairing(self._h, 1 if pairing else 0))
        self.precision = precision
        self.problems = []
        self.nodes_per_problem = []
        self.ops = []
        self.region_sets = []
        self.dense = {}

    def close(self):
        if self._h:
            self.lib.ghost_bor_destroy(self._h)
            self._h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add_problem(self, nodes, k, m_max, n_xi_efie, n_xi_bracket, c_efie, c_mfie, thetas_deg,
                    unmasked: bool = False) -> int:
        nodes = np.ascontiguousarray(nodes, dtype=np.float64)
        th = np.ascontiguousarray(np.atleast_1d(thetas_deg), dtype=np.float64)
        idx = ctypes.c_int(-1)
        check(self.lib.ghost_bor_add_problem(self._h, dptr(nodes), int(nodes.shape[0]), float(k), int(m_max),
                                             int(n_xi_efie), int(n_xi_bracket), float(c_efie), float(c_mfie),
                                             dptr(th), int(th.size), GHOST_UNMASKED if unmasked else 0,
                                             ctypes.byref(idx)))
        self.problems.append((int(m_max), int(th.size)))
        self.nodes_per_problem.append(int(nodes.shape[0]))
        return idx.value

    def set_impedance(self, prob: int, zs, c_ibc: float = 1.0, c_mass: Optional[float] = None):
        """Per-element complex surface impedance of problem prob (before
        build); c_mass (default c_ibc) weighs the (1/2) Z_s mass term."""
        ne = self.nodes_per_problem[prob] - 1
        z = np.asarray(zs, dtype=np.complex128)
        z = np.full(ne, complex(z)) if z.ndim == 0 else z.reshape(ne)
        zz = np.ascontiguousarray(np.column_stack([z.real, z.imag]).ravel())
        cm = float(c_ibc) if c_mass is None else float(c_mass)
        check(self.lib.ghost_bor_set_impedance(self._h, int(prob), dptr(zz), float(c_ibc), cm))

    def set_medium(self, prob: int, k_im: float, eta: complex):
        """Homogeneous medium of problem prob's fill (before build): Im k and
        the complex wave impedance."""
        eta = complex(eta)
        check(self.lib.ghost_bor_set_medium(self._h, int(prob), float(k_im), eta.real, eta.imag))

    def set_interior(self, prob: int, k_in: complex, eta_in: complex, n_xi_efie: int, n_xi_bracket: int):
        """Make problem prob a homogeneous penetrable body with this interior
        medium (complex wavenumber and wave impedance; before build)."""
        k_in, eta_in = complex(k_in), complex(eta_in)
        check(self.lib.ghost_bor_set_interior(self._h, int(prob), k_in.real, k_in.imag, eta_in.real, eta_in.imag,
                                              int(n_xi_efie), int(n_xi_bracket)))

    # ── multi-region bodies ──
    def add_operator(self, nodes, k, eta, m_max: int, n_xi_efie: int, n_xi_bracket: int, brackets: bool,
                     thetas_deg=()) -> int:
        """One generatrix in one medium (complex k, eta), operators only."""
        nodes = np.ascontiguousarray(nodes, dtype=np.float64)
        th = np.ascontiguousarray(np.atleast_1d(np.asarray(thetas_deg, dtype=np.float64)))
        k, eta = complex(k), complex(eta)
        idx = ctypes.c_int(-1)
        check(self.lib.ghost_bor_add_operator(self._h, dptr(nodes), int(nodes.shape[0]), k.real, k.imag, eta.real,
                                              eta.imag, int(m_max), int(n_xi_efie), int(n_xi_bracket),
                                              1 if brackets else 0, dptr(th), int(th.size), ctypes.byref(idx)))
        self.ops.append((int(nodes.shape[0]), int(m_max)))
        return idx.value

    def add_cross(self, op_test: int, op_src: int, near_factor: float = 2.0, near_order: int = 12) -> int:
        idx = ctypes.c_int(-1)
        check(self.lib.ghost_bor_add_cross(self._h, int(op_test), int(op_src), float(near_factor), int(near_order),
                                           ctypes.byref(idx)))
        return idx.value

    def add_region(self, m_max: int, n_full: int, thetas_deg) -> int:
        th = np.ascontiguousarray(np.atleast_1d(np.asarray(thetas_deg, dtype=np.float64)))
        idx = ctypes.c_int(-1)
        check(self.lib.ghost_bor_add_region(self._h, int(m_max), int(n_full), dptr(th), int(th.size),
                                            ctypes.byref(idx)))
        self.region_sets.append((int(m_max), int(th.size)))
        return idx.value

    def region_block(self, rset: int, src: int, idx: int, which: int, coef: complex, row_off: int, col_off: int):
        coef = complex(coef)
        check(self.lib.ghost_bor_region_block(self._h, int(rset), int(src), int(idx), int(which), coef.real,
                                              coef.imag, int(row_off), int(col_off)))

    def region_map(self, rset: int, m: int, red_of_full, coef):
        red = np.ascontiguousarray(red_of_full, dtype=np.int32)
        cf = np.ascontiguousarray(coef, dtype=np.float64)
        n_red = int(red.max()) + 1
        check(self.lib.ghost_bor_region_map(self._h, int(rset), int(m), n_red, iptr(red), dptr(cf)))

    def region_exterior(self, rset: int, op: int, off_J: int, off_M: int = -1):
        check(self.lib.ghost_bor_region_exterior(self._h, int(rset), int(op), int(off_J), int(off_M)))

    def region_dF(self, rset: int) -> np.ndarray:
        """complex [2 m_max + 1, 2 (VV, HH), n_theta]"""
        m_max, nth = self.region_sets[rset]
        out = np.zeros((2 * m_max + 1, 2, nth, 2), dtype=np.float64)
        check(self.lib.ghost_bor_region_get_dF(self._h, int(rset), dptr(out)))
        return out[..., 0] + 1j * out[..., 1]

    def region_system(self, rset: int, m: int):
        """(A [n, n], B [n, nrhs]) of one mode as assembled (after fill, before solve)."""
        nth = self.region_sets[rset][1]
        n = ctypes.c_int(0)
        check(self.lib.ghost_bor_region_system(self._h, int(rset), int(m), 0, ctypes.byref(n), None, None))
        A = np.zeros((n.value, n.value, 2))
        B = np.zeros((n.value, 2 * nth, 2))
        check(self.lib.ghost_bor_region_system(self._h, int(rset), int(m), 0, ctypes.byref(n), dptr(A), dptr(B)))
        return A[..., 0] + 1j * A[..., 1], B[..., 0] + 1j * B[..., 1]

    def region_solution(self, rset: int, m: int) -> np.ndarray:
        nth = self.region_sets[rset][1]
        n = ctypes.c_int(0)
        check(self.lib.ghost_bor_region_system(self._h, int(rset), int(m), 1, ctypes.byref(n), None, None))
        X = np.zeros((n.value, 2 * nth, 2))
        check(self.lib.ghost_bor_region_system(self._h, int(rset), int(m), 1, ctypes.byref(n), dptr(X), None))
        return X[..., 0] + 1j * X[..., 1]

    def add_dense(self, n: int, nrhs: int) -> int:
        """A caller-filled n x n system with nrhs columns; returns its system index."""
        idx = ctypes.c_int(-1)
        check(self.lib.ghost_bor_add_dense(self._h, int(n), int(nrhs), ctypes.byref(idx)))
        self.dense[idx.value] = (int(n), int(nrhs))
        return idx.value

    def set_dense(self, idx: int, A, B):
        n, nrhs = self.dense[idx]
        A = np.ascontiguousarray(A, dtype=np.complex128)
        B = np.ascontiguousarray(B, dtype=np.complex128).reshape(n, nrhs)
        if A.shape != (n, n):
            raise ValueError(f"A must be {n} x {n}")
        check(self.lib.ghost_bor_set_dense(self._h, int(idx), A.ctypes.data_as(_dp), B.ctypes.data_as(_dp)))

    def get_dense(self, idx: int) -> np.ndarray:
        n, nrhs = self.dense[idx]
        X = np.zeros((n, nrhs), dtype=np.complex128)
        check(self.lib.ghost_bor_get_dense(self._h, int(idx), X.ctypes.data_as(_dp)))
        return X

    def fill_rhs(self):
        check(self.lib.ghost_bor_fill_rhs(self._h))

    def set_solution(self, prob: int, m: int, X):
        n = self.system_info(prob, m)[0]
        nrhs = 2 * self.problems[prob][1]
        X = np.ascontiguousarray(X, dtype=np.complex128).reshape(n, nrhs)
        check(self.lib.ghost_bor_set_solution(self._h, int(prob), int(m), X.ctypes.data_as(_dp)))

    def build(self):
        check(self.lib.ghost_bor_build(self._h))

    def fill(self):
        check(self.lib.ghost_bor_fill(self._h))

    def solve(self, refine_iters: int = 0):
        check(self.lib.ghost_bor_solve(self._h, int(refine_iters)))

    def set_lu_super(self, panels: int):
        check(self.lib.ghost_bor_set_lu_super(self._h, int(panels)))

    def set_lu_streams(self, n: int):
        check(self.lib.ghost_bor_set_lu_streams(self._h, int(n)))

    def set_gemm_variant(self, v: int):
        check(self.lib.ghost_bor_set_gemm_variant(self._h, int(v)))

    def set_probe_tol(self, tol: float):
        check(self.lib.ghost_bor_set_probe_tol(self._h, float(tol)))

    def refine_history(self):
        out = np.zeros(64)
        n = ctypes.c_int(64)
        check(self.lib.ghost_bor_refine_history(self._h, dptr(out), ctypes.byref(n)))
        return out[:min(n.value, 64)].tolist()

    def farfield(self):
        check(self.lib.ghost_bor_farfield(self._h))

    def run(self, refine_iters: int = 0):
        check(self.lib.ghost_bor_run(self._h, int(refine_iters)))

    def sync(self):
        check(self.lib.ghost_bor_sync(self._h))

    def dF(self, prob: int) -> np.ndarray:
        """complex [2 m_max + 1, 2 (VV, HH), n_theta]"""
        m_max, nth = self.problems[prob]
        out = np.zeros((2 * m_max + 1, 2, nth, 2), dtype=np.float64)
        check(self.lib.ghost_bor_get_dF(self._h, int(prob), dptr(out)))
        return out[..., 0] + 1j * out[..., 1]

    def adjoint(self, acoef, step: float = 0.0, method: str = "analytic") -> np.ndarray:
        """Gradient w.r.t. the radial coordinate of every node of every problem
        of dJ = Re sum_c a_c dF_c, with acoef [problems][n_rhs] complex in
        column order 2 * theta_index + (0 VV, 1 HH) and F the mode-summed
        far-field amplitude.  Needs a solved FP64 handle.  method "analytic":
        the fill's shape derivative in the kernels (the systems keep their
        factors); "fd": central differences of the fill with half-width step
        (0: 1e-6 of the shortest element), after which the systems hold a
        perturbed fill (fill again before the next solve)."""
        a = np.ascontiguousarray(np.asarray(acoef, dtype=np.complex128).ravel())
        ncol = sum(2 * nth for _, nth in self.problems)
        if a.size != ncol:
            raise ValueError(f"acoef must have {ncol} entries, got {a.size}")
        ab = np.ascontiguousarray(np.column_stack([a.real, a.imag]).ravel())
        out = np.zeros(sum(self.nodes_per_problem), dtype=np.float64)
        if method == "analytic":
            check(self.lib.ghost_bor_adjoint_analytic(self._h, dptr(ab), dptr(out)))
        elif method == "fd":
            check(self.lib.ghost_bor_adjoint(self._h, dptr(ab), float(step), dptr(out)))
        else:
            raise ValueError(f"method must be 'analytic' or 'fd', got {method!r}")
        return out

    def set_nodes(self, prob: int, nodes) -> None:
        """New (rho, z) nodes for a built problem (same count, same axis ends);
        the build's quadrature rules and a captured graph stay valid."""
        a = np.ascontiguousarray(np.asarray(nodes, dtype=np.float64))
        if a.ndim != 2 or a.shape[1] != 2:
            raise ValueError("nodes must be an (Nn, 2) array")
        check(self.lib.ghost_bor_set_nodes(self._h, int(prob), dptr(a.ravel())))

    def set_objective(self, weights) -> None:
        """J = sum_c w_c sigma_dBsm(c): weights [problems][n_rhs] real, column
        order 2 * theta_index + (0 VV, 1 HH); used by capture_adjoint."""
        w = np.ascontiguousarray(np.asarray(weights, dtype=np.float64).ravel())
        ncol = sum(2 * nth for _, nth in self.problems)
        if w.size != ncol:
            raise ValueError(f"weights must have {ncol} entries, got {w.size}")
        check(self.lib.ghost_bor_set_objective(self._h, dptr(w)))

    def capture_adjoint(self, refine_iters: int = 0) -> None:
        """Capture fill + solve + far field + the objective's adjoint into one
        graph (launch() runs it; dF() and gradient() read the results)."""
        check(self.lib.ghost_bor_capture_adjoint(self._h, int(refine_iters)))

    def gradient(self) -> np.ndarray:
        """dJ/d rho_node of every problem's nodes from the last analytic adjoint."""
        out = np.zeros(sum(self.nodes_per_problem), dtype=np.float64)
        check(self.lib.ghost_bor_get_gradient(self._h, dptr(out)))
        return out

    def num_systems(self) -> int:
        """Mode systems of every problem plus dense systems (the library's system order)."""
        n = ctypes.c_int(0)
        check(self.lib.ghost_bor_system_count(self._h, ctypes.byref(n)))
        return n.value

    def pivots(self, idx: int, n: int | None = None) -> np.ndarray:
        """Row interchanges of system idx's factor (LAPACK getrf ipiv - 1);
        n: the system size (known here for dense systems)."""
        if n is None:
            if idx not in self.dense:
                raise ValueError("pass n for a mode system")
            n = self.dense[idx][0]
        out = np.zeros(int(n), dtype=np.int32)
        check(self.lib.ghost_bor_get_pivots(self._h, int(idx), out.ctypes.data_as(_ip)))
        return out

    def residuals(self) -> np.ndarray:
        """||A X - B|| / ||B|| per system, in the order systems were added."""
        out = np.zeros(self.num_systems(), dtype=np.float64)
        check(self.lib.ghost_bor_residual(self._h, dptr(out)))
        return out

    def system_info(self, prob: int, m: int):
        n, nt, t0, f0 = (ctypes.c_int() for _ in range(4))
        check(self.lib.ghost_bor_system_info(self._h, int(prob), int(m), ctypes.byref(n), ctypes.byref(nt),
                                             ctypes.byref(t0), ctypes.byref(f0)))
        return n.value, nt.value, t0.value, f0.value

    def system(self, prob: int, m: int):
        """(A [n, n], B [n, nrhs]) as assembled; call after fill and before solve."""
        n = self.system_info(prob, m)[0]
        nrhs = 2 * self.problems[prob][1]
        A = np.zeros((n, n, 2))
        B = np.zeros((n, nrhs, 2))
        check(self.lib.ghost_bor_get_system(self._h, int(prob), int(m), 0, dptr(A), dptr(B)))
        return A[..., 0] + 1j * A[..., 1], B[..., 0] + 1j * B[..., 1]

    def solution(self, prob: int, m: int) -> np.ndarray:
        n = self.system_info(prob, m)[0]
        nrhs = 2 * self.problems[prob][1]
        X = np.zeros((n, nrhs, 2))
        check(self.lib.ghost_bor_get_system(self._h, int(prob), int(m), 1, dptr(X), None))
        return X[..., 0] + 1j * X[..., 1]

    def zero_pivots(self) -> np.ndarray:
        out = np.zeros(self.num_systems(), dtype=np.int32)
        check(self.lib.ghost_bor_info(self._h, iptr(out)))
        return out

    def capture(self, refine_iters: int = 0):
        check(self.lib.ghost_bor_capture(self._h, int(refine_iters)))

    def launch(self):
        check(self.lib.ghost_bor_launch(self._h))

    def graph_ms(self) -> float:
        out = np.zeros(1, dtype=np.float32)
        check(self.lib.ghost_bor_graph_ms(self._h, out.ctypes.data_as(_fp)))
        return float(out[0])

    def set_profile(self, on: bool = True):
        check(self.lib.ghost_bor_set_profile(self._h, 1 if on else 0))

    def kernel_times(self):
        '''{family: (ms, launches, algorithmic flops)} since set_profile.'''
        n = self.lib.ghost_bor_kernel_count()
        ms = np.zeros(n)
        cnt = np.zeros(n, dtype=np.int32)
        fl = np.zeros(n)
        check(self.lib.ghost_bor_kernel_times(self._h, dptr(ms), iptr(cnt), dptr(fl)))
        return {self.lib.ghost_bor_kernel_name(i).decode(): (float(ms[i]), int(cnt[i]), float(fl[i]))
                for i in range(n)}

    def timings(self):
        out = np.zeros(3, dtype=np.float32)
        check(self.lib.ghost_bor_timings(self._h, out.ctypes.data_as(_fp)))
        return {"fill_ms": float(out[0]), "solve_ms": float(out[1]), "farfield_ms": float(out[2])}
