"""Write golden vectors of the reference's PMCHWT (homogeneous penetrable
body) path.

Run in the development container only (the reference sources are not
present where the GPU tests run):

    PYTHONDONTWRITEBYTECODE=1 python tests/golden/make_golden_pmchwt.py /root/reference

For a 12-segment sphere and a 16-segment cone-sphere it stores, per mode,
the interior medium's EFIE operator T_i and rotated-PV operator P_i
(unmasked, bor_solver.py:293-345, 554-575 with complex k), the masked PMCHWT
system, its right-hand sides, currents and far field
(bor_solver.py:1019-1094), and the end-to-end result of solve_bor_dielectric.
dispatch_dielectric.npz holds solve_monostatic_rcs_bor on a TYPE 3 sphere
(one inline dielectric flag, elements per interior wavelength) at
two frequencies.  A second argument "dispatch" writes only that file.
"""

import json

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
THETAS = np.array([0.0, 30.0, 60.0, 90.0, 120.0, 150.0, 180.0])
sys.path.insert(0, HERE)
from make_golden import cone_sphere_nodes  # noqa: E402


def case(bs, bk, nodes, f_hz, nm, eps, mu):
    se = bs.BorPecSolver(nodes, f_hz)
    si = bs.BorPecSolver(nodes, f_hz, medium=(eps, mu))
    se.prepare_operators(nm, efie=True, ibc=True)
    si.prepare_operators(nm, efie=True, ibc=True)
    Nn = se.Nn
    r2 = (bk.ETA0 / si.eta) ** 2
    data = {"nodes": nodes, "freq_hz": f_hz, "m_max": nm, "thetas": THETAS, "eps_r": eps, "mu_r": mu,
            "k_in": si.k, "eta_in": si.eta}
    for m in range(-nm, nm + 1):
        Te, Ti = se.assemble_mode(m, nm), si.assemble_mode(m, nm)
        Pe, Pi = se.assemble_pmchwt_P(m, nm), si.assemble_pmchwt_P(m, nm)
        Ps = bk.ETA0 * (Pe + Pi)
        A = np.block([[Te + Ti, -Ps], [Ps, Te + r2 * Ti]])
        mask = np.tile(se.basis_mask(m), 2)
        A = A[np.ix_(mask, mask)]
        cols = []
        for th in THETAS:
            for pol in ("VV", "HH"):
                cols.append(np.concatenate([se.rhs_mode(m, th, pol), bk.ETA0 * se.rhs_h_mode(m, th, pol)])[mask])
        B = np.stack(cols, axis=1)
        X = np.linalg.solve(A, B)
        dF = np.zeros((2, len(THETAS)), dtype=complex)
        for it, th in enumerate(THETAS):
            for ip in range(2):
                full = np.zeros(4 * Nn, dtype=complex)
                full[mask] = X[:, 2 * it + ip]
                fth, fph = se.farfield_mode(m, full[:2 * Nn], th, msol=bk.ETA0 * full[2 * Nn:])
                dF[ip, it] = fth if ip == 0 else fph
        data.update({f"Ti_{m}": Ti, f"Pi_{m}": Pi, f"Pe_{m}": Pe, f"A_{m}": A, f"B_{m}": B,
                     f"X_{m}": X, f"dF_{m}": dF, f"mask_{m}": mask})
    out = bs.solve_bor_dielectric(nodes, f_hz, THETAS, eps, mu, n_modes=nm, mode_tol=1e-6)
    for key in ("sigma_vv", "sigma_hh", "amp_vv", "amp_hh"):
        data[key] = np.array(out[key])
    data["modes_used"] = out["modes_used"]
    return data


def dielectric_snapshot():
    """The 0.4 m sphere of make_golden.dispatch_snapshot as TYPE 3 segments
    on dielectric flag 1, 10 elements per (interior) wavelength."""
    from make_golden import dispatch_snapshot
    snap = dispatch_snapshot()
    for seg in snap["segments"]:
        seg["seg_type"] = 3
        seg["properties"] = [3, -10, 0, 1, 0]
    snap["dielectrics"] = [["1", "2.5", "-0.1", "1.0", "0.0"]]
    return snap


def dispatch_case(bd):
    snap = dielectric_snapshot()
    data = {"snapshot": json.dumps(snap)}
    for pol in ("VV", "HH"):
        out = bd.solve_monostatic_rcs_bor(snap, [0.3, 0.45], [0.0, 45.0, 90.0, 135.0, 180.0], pol,
                                          geometry_units="meters", n_modes=4, workers=1)
        data[f"{pol}_amp"] = np.array([s["rcs_amp_real"] + 1j * s["rcs_amp_imag"] for s in out["samples"]])
        data[f"{pol}_rcs_db"] = np.array([s["rcs_db"] for s in out["samples"]])
        data["modes_used"] = np.array([m["modes_used"] for m in out["metadata"]["per_frequency"]])
        data["n_unknowns"] = np.array([m["n_unknowns"] for m in out["metadata"]["per_frequency"]])
        data["formulation"] = out["metadata"]["formulation"]
    np.savez_compressed(os.path.join(HERE, "dispatch_dielectric.npz"), **data)


def main(ref_dir: str, parts: str = "all"):
    sys.path.insert(0, ref_dir)
    import bor_dispatch as bd
    import bor_kernels as bk
    import bor_solver as bs
    if parts == "dispatch":
        dispatch_case(bd)
        print("dispatch dielectric fixture written to", HERE)
        return
    dispatch_case(bd)
    np.savez_compressed(os.path.join(HERE, "sphere12_pmchwt.npz"),
                        **case(bs, bk, bs.sphere_generatrix(0.5, 12), 300e6, 4, 2.5 - 0.1j, 1.0 + 0j))
    np.savez_compressed(os.path.join(HERE, "conesphere16_pmchwt.npz"),
                        **case(bs, bk, cone_sphere_nodes(7, 9), 350e6, 4, 3.0 - 0.4j, 1.3 - 0.05j))
    print("PMCHWT fixtures written to", HERE)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference", sys.argv[2] if len(sys.argv) > 2 else "all")
