"""Write golden-vector fixtures from the reference solver for the impedance
(IBC-EFIE) path, an open generatrix, and the 120-segment sphere workload.

Run in the development container only (the reference sources are not
present where the GPU tests run):

    PYTHONDONTWRITEBYTECODE=1 python tests/golden/make_golden_ibc_cfg1.py /root/reference

It imports the reference's bor_solver / bor_kernels and writes small .npz
files of inputs and outputs next to this script.  They pin what no closed
form fixes exactly:

* the IBC-EFIE mode systems with the rotated-PV brackets and the Z_s mass
  (bor_solver.py:463-552, 929-934), their currents, and the far field with
  the radiating magnetic current (bor_solver.py:697-707), for a scalar and a
  per-element (tapered) impedance;
* the open-end basis mask (bor_solver.py:611-622) on an open shell;
* the 120-segment PEC sphere at 300 MHz (|m| <= 12, 181 aspects, CFIE and
  EFIE, FP64 tables): amplitudes, modes_used and a column sample of the
  currents of modes 0, +-1 and 12.
"""

import math
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
THETAS = np.array([0.0, 30.0, 60.0, 90.0, 120.0, 150.0, 180.0])

sys.path.insert(0, HERE)
from make_golden import cone_sphere_nodes  # noqa: E402


def open_cup_nodes(n: int, a: float = 0.4):
    """Hemispherical shell: pole on the axis at +z, open rim at z = 0."""
    th = np.linspace(0.0, math.pi / 2, n + 1)
    return np.column_stack([a * np.sin(th), a * np.cos(th)])


def taper_zs(ne: int):
    """A per-element complex impedance ramp (ohm) from 0.1 eta0 to 0.6 eta0
    with a reactive part, the shape of a tapered absorber."""
    s = (np.arange(ne) + 0.5) / ne
    return (0.1 + 0.5 * s) * 376.730313668 - 1j * (40.0 + 60.0 * s)


def mode_case(bs, bk, nodes, f_hz, nm, alpha, zs_elem, formulation):
    sv = bs.BorPecSolver(nodes, f_hz)
    ibc = zs_elem is not None
    zs_pt = zs_elem[sv.g.elem] if ibc else None
    sv.prepare_operators(nm, efie=True, mfie=alpha < 1.0, ibc=ibc)
    data = {"nodes": nodes, "freq_hz": f_hz, "alpha": alpha, "m_max": nm, "thetas": THETAS}
    if ibc:
        data["zs"] = zs_elem
    for m in range(-nm, nm + 1):
        Z = alpha * sv.assemble_mode(m, nm)
        if ibc:
            Z = Z + alpha * sv.assemble_ibc_extra(m, nm, zs_pt, zs_elem)
        if alpha < 1.0:
            Z = Z + (1 - alpha) * bk.ETA0 * sv.assemble_mfie_mode(m, nm)
        mask = sv.basis_mask(m)
        A = Z[np.ix_(mask, mask)]
        cols = []
        for th in THETAS:
            for pol in ("VV", "HH"):
                V = alpha * sv.rhs_mode(m, th, pol)
                if alpha < 1.0:
                    V = V + (1 - alpha) * bk.ETA0 * sv.rhs_mfie_mode(m, th, pol)
                cols.append(V[mask])
        B = np.stack(cols, axis=1)
        X = np.linalg.solve(A, B)
        dF = np.zeros((2, len(THETAS)), dtype=complex)
        for it, th in enumerate(THETAS):
            for ip in range(2):
                sol = np.zeros(2 * sv.Nn, dtype=complex)
                sol[mask] = X[:, 2 * it + ip]
                fth, fph = sv.farfield_mode(m, sol, th, zs_pt=zs_pt)
                dF[ip, it] = fth if ip == 0 else fph
        data[f"A_{m}"] = A
        data[f"B_{m}"] = B
        data[f"X_{m}"] = X
        data[f"dF_{m}"] = dF
        data[f"mask_{m}"] = mask
    out = bs.solve_bor(nodes, f_hz, THETAS, formulation=formulation, zs=zs_elem, n_modes=nm,
                       mode_tol=1e-6, table_precision="double")
    for key in ("sigma_vv", "sigma_hh", "amp_vv", "amp_hh"):
        data[key] = np.array(out[key])
    data["modes_used"] = out["modes_used"]
    data["linear_residual"] = out["linear_residual"]
    return data


def main(ref_dir: str):
    sys.path.insert(0, ref_dir)
    import bor_kernels as bk
    import bor_solver as bs

    # ── IBC-EFIE: scalar lossy impedance on a sphere, tapered impedance on a cone-sphere ──
    sph = bs.sphere_generatrix(0.5, 12)
    zs_sph = np.full(12, complex(0.3 * bk.ETA0, -0.2 * bk.ETA0))
    np.savez_compressed(os.path.join(HERE, "sphere12_ibc.npz"),
                        **mode_case(bs, bk, sph, 300e6, 4, 1.0, zs_sph, "efie"))
    cs = cone_sphere_nodes(7, 9)
    np.savez_compressed(os.path.join(HERE, "conesphere16_ibc_taper.npz"),
                        **mode_case(bs, bk, cs, 450e6, 5, 1.0, taper_zs(len(cs) - 1), "efie"))

    # ── open generatrix (open-end mask), PEC EFIE ──
    np.savez_compressed(os.path.join(HERE, "cup10_efie_open.npz"),
                        **mode_case(bs, bk, open_cup_nodes(10), 400e6, 3, 1.0, None, "efie"))

    # ── 120-segment PEC sphere, 300 MHz, |m| <= 12, theta 0..180 step 1 ──
    nodes = bs.sphere_generatrix(0.5, 120)
    f_hz = 300e6
    th = np.arange(0.0, 181.0)
    data = {"nodes": nodes, "freq_hz": f_hz, "thetas": th, "m_max": 12}
    for form in ("cfie", "efie"):
        out = bs.solve_bor(nodes, f_hz, th, formulation=form, n_modes=12, mode_tol=0.0,
                           table_precision="double", assembly="tables")
        for key in ("sigma_vv", "sigma_hh", "amp_vv", "amp_hh"):
            data[f"{form}_{key}"] = np.array(out[key])
        data[f"{form}_modes_used"] = out["modes_used"]
        data[f"{form}_linear_residual"] = out["linear_residual"]
        dflt = bs.solve_bor(nodes, f_hz, th, formulation=form, n_modes=12, mode_tol=1e-6,
                            table_precision="double", assembly="tables")
        data[f"{form}_modes_used_default_tol"] = dflt["modes_used"]
        print(form, "done", flush=True)
    # currents of a few CFIE modes, every 8th aspect (both polarisations)
    sv = bs.BorPecSolver(nodes, f_hz)
    sv.prepare_operators(12, efie=True, mfie=True)
    col_th = th[::8]
    data["x_thetas"] = col_th
    for m in (0, 1, -1, 12):
        Z = 0.5 * sv.assemble_mode(m, 12) + 0.5 * bk.ETA0 * sv.assemble_mfie_mode(m, 12)
        mask = sv.basis_mask(m)
        A = Z[np.ix_(mask, mask)]
        cols = []
        for t in col_th:
            for pol in ("VV", "HH"):
                V = 0.5 * sv.rhs_mode(m, t, pol) + 0.5 * bk.ETA0 * sv.rhs_mfie_mode(m, t, pol)
                cols.append(V[mask])
        data[f"X_{m}"] = np.linalg.solve(A, np.stack(cols, axis=1))
    np.savez_compressed(os.path.join(HERE, "sphere120_cfg1.npz"), **data)
    print("golden fixtures written to", HERE)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
