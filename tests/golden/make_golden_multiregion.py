"""Write golden vectors of the reference's multi-region PMCHWT path (PEC
cores under coating layers).

Run in the development container only (the reference sources are not
present where the GPU tests run):

    PYTHONDONTWRITEBYTECODE=1 python tests/golden/make_golden_multiregion.py /root/reference

coated1_multiregion.npz: a PEC sphere (a = 0.35 m, 8 segments) under one
lossy magnetodielectric layer (outer a = 0.5 m, 10 segments; the gap is
under twice the element length, so tensor-grid near pairs occur), 300 MHz,
|m| <= 3.  Per mode: the reduced system, its right-hand sides, currents and
far field from the region system (bor_solver.py:1805-2041), the two cross
operators (bor_solver.py:1115-1297), and the end-to-end results of
solve_bor_coated_pec and solve_bor_coated_n_pec.
coated2_multiregion.npz: two layers (radii 0.33 / 0.42 / 0.5 m), end-to-end
solve_bor_coated2_pec and the currents of modes 0 and 1.
"""

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
THETAS = np.array([0.0, 45.0, 90.0, 135.0, 180.0])


def regions_n(bs, outer_list, core, eps_list, mu_list, f_hz):
    N = len(outer_list)
    surfaces = [(p, False) for p in outer_list] + [(core, True)]
    regions = [{"medium": None, "bounds": [(0, +1)], "exterior": True}]
    for i in range(N):
        lay = N - i
        regions.append({"medium": (eps_list[lay - 1], mu_list[lay - 1]), "bounds": [(i, -1), (i + 1, +1)]})
    return bs._MultiRegionBor(surfaces, regions, f_hz)


def per_mode(sys_, nm, data, prefix):
    sys_.prepare(nm)
    for m in range(-nm, nm + 1):
        A = sys_.assemble(m, nm)[0]
        Q = sys_.build_Q(m)
        B = np.stack([sys_.rhs(m, th, pol) for th in THETAS for pol in ("VV", "HH")], axis=1)
        X = np.linalg.solve(A, B)
        dF = np.zeros((2, len(THETAS)), dtype=complex)
        for it, th in enumerate(THETAS):
            for ip, pol in enumerate(("VV", "HH")):
                dF[ip, it] = sys_.farfield(m, X[:, 2 * it + ip], th, pol)
        data.update({f"{prefix}A_{m}": A, f"{prefix}Q_{m}": Q, f"{prefix}B_{m}": B, f"{prefix}X_{m}": X,
                     f"{prefix}dF_{m}": dF})


def main(ref_dir: str):
    sys.path.insert(0, ref_dir)
    import bor_solver as bs
    f_hz, nm = 300e6, 3
    outer, core = bs.sphere_generatrix(0.5, 10), bs.sphere_generatrix(0.35, 8)
    eps, mu = 2.5 - 0.2j, 1.2 - 0.05j
    data = {"outer": outer, "core": core, "freq_hz": f_hz, "m_max": nm, "thetas": THETAS, "eps": eps, "mu": mu}
    sys_ = regions_n(bs, [outer], core, [eps], [mu], f_hz)
    per_mode(sys_, nm, data, "")
    for key, X in ((("oc"), sys_.X[(1, 0, 1)]), (("co"), sys_.X[(1, 1, 0)])):
        for m in range(-nm, nm + 1):
            data[f"T{key}_{m}"] = X.assemble_T(m, nm)
            data[f"P{key}_{m}"] = X.assemble_P(m, nm)
        data[f"near_{key}"] = np.array([[e, f, -1 if X.pair_kind[(e, f)] is None else
                                         int(X.pair_kind[(e, f)][6:8])] for (e, f) in X.near_pairs])
    out = bs.solve_bor_coated_pec(outer, core, f_hz, THETAS, eps, mu, n_modes=nm, mode_tol=1e-6)
    out_n = bs.solve_bor_coated_n_pec([outer], core, f_hz, THETAS, [eps], [mu], n_modes=nm, mode_tol=1e-6)
    for key in ("amp_vv", "amp_hh"):
        data["coated_" + key] = np.array(out[key])
        data["coated_n_" + key] = np.array(out_n[key])
    data["coated_modes_used"] = out["modes_used"]
    data["coated_n_unknowns"] = out["n_unknowns"]
    np.savez_compressed(os.path.join(HERE, "coated1_multiregion.npz"), **data)

    mid = bs.sphere_generatrix(0.42, 9)
    core2 = bs.sphere_generatrix(0.33, 8)
    e_in, m_in, e_out, m_out = 4.0 - 0.5j, 1.0, 2.2 - 0.1j, 1.1 - 0.02j
    d2 = {"outer": outer, "mid": mid, "core": core2, "freq_hz": f_hz, "m_max": nm, "thetas": THETAS,
          "eps_inner": e_in, "mu_inner": m_in, "eps_outer": e_out, "mu_outer": m_out}
    out2 = bs.solve_bor_coated2_pec(outer, mid, core2, f_hz, THETAS, e_in, m_in, e_out, m_out, n_modes=nm,
                                    mode_tol=1e-6)
    for key in ("amp_vv", "amp_hh"):
        d2[key] = np.array(out2[key])
    d2["modes_used"] = out2["modes_used"]
    d2["n_unknowns"] = out2["n_unknowns"]
    sys2 = regions_n(bs, [outer, mid], core2, [e_in, e_out], [m_in, m_out], f_hz)
    sys2.prepare(nm)
    for m in (0, 1):
        A = sys2.assemble(m, nm)[0]
        B = np.stack([sys2.rhs(m, th, pol) for th in THETAS for pol in ("VV", "HH")], axis=1)
        d2[f"X_{m}"] = np.linalg.solve(A, B)
    np.savez_compressed(os.path.join(HERE, "coated2_multiregion.npz"), **d2)
    print("multi-region fixtures written to", HERE)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
