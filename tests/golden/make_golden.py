"""Write the golden-vector fixtures from the reference solver itself.

Run in the development container only (the reference sources are not
present where the GPU tests run):

    PYTHONDONTWRITEBYTECODE=1 python tests/golden/make_golden.py /root/reference

It imports the reference's bor_kernels / bor_solver / mie_sphere and writes
small .npz files of inputs and outputs next to this script.  The fixtures
pin what neither closed forms nor brute-force quadrature fix exactly: the
near-pair quadrature values (bor_kernels.py:170-257, 360-431), the assembled
masked mode systems (bor_solver.py:293-345, 399-435, 611-622, 929-938),
the per-mode currents of np.linalg.solve (bor_solver.py:744-753) and the
end-to-end monostatic result (bor_solver.py:820-1004).
"""

import math
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
THETAS = np.array([0.0, 30.0, 60.0, 90.0, 120.0, 150.0, 180.0])


def cone_sphere_nodes(n_cone: int, n_sph: int, r: float = 0.3, half_deg: float = 15.0):
    """Nose tip at +z, straight cone flank tangent to a sphere cap, pole at -z."""
    al = math.radians(half_deg)
    # tangent point on the sphere (centre at origin) where the flank meets it
    tp = math.pi / 2 - al                 # polar angle of tangency measured from +z
    zt, rt = r * math.cos(tp), r * math.sin(tp)
    tip = zt + rt / math.tan(al)
    cone = np.column_stack([np.linspace(0.0, rt, n_cone + 1), np.linspace(tip, zt, n_cone + 1)])
    th = np.linspace(tp, math.pi, n_sph + 1)[1:]
    sph = np.column_stack([r * np.sin(th), r * np.cos(th)])
    return np.vstack([cone, sph])


def dispatch_snapshot():
    """A 0.4 m PEC sphere drawn as two TYPE 2 segments (+z pole -> equator -> -z pole),
    10 primitives each, with 2 elements per primitive (N = 2)."""
    def arc(t0, t1, n):
        th = np.linspace(t0, t1, n + 1)
        x, y = 0.4 * np.sin(th), 0.4 * np.cos(th)
        return [{"x1": x[i], "y1": y[i], "x2": x[i + 1], "y2": y[i + 1]} for i in range(n)]
    return {"title": "sphere", "segments": [
        {"name": "north", "seg_type": 2, "properties": [2, 2, 0, 0, 0], "point_pairs": arc(0.0, math.pi / 2, 10)},
        {"name": "south", "seg_type": 2, "properties": [2, 2, 0, 0, 0], "point_pairs": arc(math.pi / 2, math.pi, 10)},
    ], "ibcs": [], "dielectrics": []}


def main(ref_dir: str):
    sys.path.insert(0, ref_dir)
    import bor_kernels as bk
    import bor_solver as bs
    import mie_sphere as ms

    # ── near-pair modal quadrature on a small sphere (a6) ──
    nodes = bs.sphere_generatrix(0.5, 12)
    f_hz = 300e6
    k = 2 * math.pi * f_hz / bk.C0
    m_max = 4
    solver = bs.BorPecSolver(nodes, f_hz)
    near = {}
    stride = 5      # every 5th cell point pair keeps the fixture small
    for (e, f) in ((0, 0), (0, 1), (1, 0), (5, 5), (5, 6), (11, 11)):
        d = solver._near_pair_data(e, f, m_max)
        near[f"G_{e}_{f}"] = d[-1][::stride]
        dm = solver._near_mfie_data(e, f, m_max)
        for uv in range(4):
            near[f"K{uv}_{e}_{f}"] = dm[-1][uv][::stride]
    np.savez_compressed(os.path.join(HERE, "near_sphere12.npz"), nodes=nodes, freq_hz=f_hz,
                        m_max=m_max, stride=stride, **near)

    # ── assembled masked systems, currents and far fields (a8-a13) ──
    cases = {
        "sphere12_cfie": (bs.sphere_generatrix(0.5, 12), 300e6, "cfie", 4, "tables"),
        "sphere12_efie": (bs.sphere_generatrix(0.5, 12), 300e6, "efie", 4, "tables"),
        "conesphere16_cfie": (cone_sphere_nodes(7, 9), 450e6, "cfie", 5, "tables"),
        "sphere10_stream": (bs.sphere_generatrix(0.4, 10), 500e6, "cfie", 3, "streaming"),
    }
    for name, (nodes, f_hz, form, nm, asm) in cases.items():
        alpha = 0.5 if form == "cfie" else 1.0
        sv = bs.BorPecSolver(nodes, f_hz)
        if asm == "streaming":
            sv.enable_streaming(nm, efie=True, mfie=alpha < 1.0)
        sv.prepare_operators(nm, efie=True, mfie=alpha < 1.0)
        data = {"nodes": nodes, "freq_hz": f_hz, "alpha": alpha, "m_max": nm,
                "thetas": THETAS}
        for m in range(-nm, nm + 1):
            Z = alpha * sv.assemble_mode(m, nm)
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
                    fth, fph = sv.farfield_mode(m, sol, th)
                    dF[ip, it] = fth if ip == 0 else fph
            data[f"A_{m}"] = A
            data[f"B_{m}"] = B
            data[f"X_{m}"] = X
            data[f"dF_{m}"] = dF
            data[f"mask_{m}"] = mask
        out = bs.solve_bor(nodes, f_hz, THETAS, formulation=form, n_modes=nm,
                           mode_tol=1e-6, assembly=asm, table_precision="double")
        data["sigma_vv"] = np.array(out["sigma_vv"])
        data["sigma_hh"] = np.array(out["sigma_hh"])
        data["amp_vv"] = np.array(out["amp_vv"])
        data["amp_hh"] = np.array(out["amp_hh"])
        data["modes_used"] = out["modes_used"]
        data["linear_residual"] = out["linear_residual"]
        np.savez_compressed(os.path.join(HERE, f"{name}.npz"), **data)

    # ── dispatch: snapshot -> chains -> mesh -> samples (bor_dispatch.py:109-251, 335-702) ──
    import bor_dispatch as bd
    snap = dispatch_snapshot()
    out = bd.solve_monostatic_rcs_bor(snap, [0.3], [0.0, 60.0, 90.0, 180.0], "VV",
                                      geometry_units="meters", cfie_alpha=0.5, n_modes=3,
                                      workers=1, table_precision="double")
    chains = bd._chains_from_snapshot(snap, 1.0)
    diag = max(float(np.ptp(np.vstack([c.pts for c in chains]), axis=0).max()), 1e-9)
    ordered = bd._stitch_generatrix(chains, "TYPE 2", max(1e-12, 1e-9 * diag))
    pts, _, _ = bd._mesh_generatrix(ordered, bk.C0 / 0.3e9, 50000, 1e-6 * diag)
    np.savez_compressed(os.path.join(HERE, "dispatch_sphere.npz"), nodes=pts,
                        rcs_db=np.array([s["rcs_db"] for s in out["samples"]]),
                        amp=np.array([s["rcs_amp_real"] + 1j * s["rcs_amp_imag"] for s in out["samples"]]),
                        theta=np.array([s["theta_inc_deg"] for s in out["samples"]]),
                        modes_used=out["metadata"]["per_frequency"][0]["modes_used"])

    # ── Mie anchors (analytic oracle, mie_sphere.py:231-235) ──
    ka = np.array([0.05, 0.5, 1.0, math.pi, 5.0, 10.0])
    a = 0.5
    freqs = ka / a * bk.C0 / (2 * math.pi)
    np.savez_compressed(os.path.join(HERE, "mie_pec.npz"), radius=a, freq_hz=freqs,
                        sigma=np.array([ms.sigma_pec_sphere(a, f) for f in freqs]))
    print("golden fixtures written to", HERE)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
