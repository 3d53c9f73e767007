"""Write golden vectors of the reference's partial-coating and coating-patch
solvers and of its dispatch over the multi-region layouts (TYPE 3 + 4,
2 + 3 + 4, 3 + 4 + 5 with and without an exposed inner interface, and an
N-layer TYPE 3 + 5 + 5 + 4 stack).

Run in the development container only (the reference sources are not
present where the GPU tests run):

    PYTHONDONTWRITEBYTECODE=1 python tests/golden/make_golden_multiregion_dispatch.py /root/reference

partial_multiregion.npz: a PEC sphere (a = 0.4 m) whose north cap (polar
angle < 70 deg) is under a lossy coating to r = 0.5 m, closed by a radial
step down to the core: interface, covered core and one bare piece meeting at
one junction circle; then a band coating (two bare caps, two junctions);
solve_bor_partial_coating end to end at 300 MHz, |m| <= 3.
patch_multiregion.npz: solve_bor_coating_patch, a second layer over the north
cap of a fully coated PEC sphere (triple junction of two dielectrics and air).
dispatch_multiregion.npz: solve_monostatic_rcs_bor on .geo snapshots of
the coated, partially coated, two-layer, patch and three-layer layouts.
"""

import math
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
THETAS = np.array([0.0, 40.0, 90.0, 140.0, 180.0])
F_HZ = 300e6
NM = 3


def arc(r, t0, t1, n):
    """Nodes (rho, z) on a circle of radius r from polar angle t0 to t1."""
    th = np.linspace(t0, t1, n + 1)
    return np.column_stack([r * np.sin(th), r * np.cos(th)])


def line(p0, p1, n):
    s = np.linspace(0.0, 1.0, n + 1)[:, None]
    return np.asarray(p0)[None, :] * (1.0 - s) + np.asarray(p1)[None, :] * s


def join(*parts):
    out = [parts[0]]
    for p in parts[1:]:
        out.append(p[1:])
    pts = np.vstack(out)
    pts[np.abs(pts[:, 0]) < 1e-15, 0] = 0.0
    return pts


def cap_geometry(rc=0.4, ro=0.5, tj=math.radians(70.0)):
    """Interface (outer cap + radial step), covered core cap, bare core."""
    iface = join(arc(ro, 0.0, tj, 6), line(arc(ro, tj, tj, 1)[0], arc(rc, tj, tj, 1)[0], 2))
    covered = arc(rc, 0.0, tj, 5)
    bare = arc(rc, tj, math.pi, 6)
    return iface, covered, bare


def band_geometry(rc=0.4, ro=0.5, t0=math.radians(60.0), t1=math.radians(120.0)):
    """A coating band between polar angles t0 and t1: the interface runs
    from the core up to ro and back (two junctions), two bare caps."""
    p0c, p0o = arc(rc, t0, t0, 1)[0], arc(ro, t0, t0, 1)[0]
    p1o, p1c = arc(ro, t1, t1, 1)[0], arc(rc, t1, t1, 1)[0]
    iface = join(line(p0c, p0o, 2), arc(ro, t0, t1, 4), line(p1o, p1c, 2))
    covered = arc(rc, t0, t1, 4)
    bares = [arc(rc, 0.0, t0, 4), arc(rc, t1, math.pi, 4)]
    return iface, covered, bares


def segment(name, props, pts):
    pp = [{"x1": float(pts[i, 0]), "y1": float(pts[i, 1]), "x2": float(pts[i + 1, 0]),
           "y2": float(pts[i + 1, 1])} for i in range(len(pts) - 1)]
    return {"name": name, "seg_type": int(props[0]), "properties": list(props), "point_pairs": pp}


def snapshots():
    """.geo snapshots (metres) of the multi-region layouts; every primitive
    carries 1 element (density 1), so the meshes equal the drawn polylines."""
    diel = [["1", "2.5", "-0.2", "1.2", "-0.05"], ["2", "3.2", "-0.4", "1.0", "0.0"],
            ["3", "1.8", "-0.1", "1.1", "-0.02"]]
    snaps = {}
    outer, core = arc(0.5, 0.0, math.pi, 9), arc(0.38, 0.0, math.pi, 7)
    snaps["coated"] = [segment("outer", [3, 1, 0, 1, 0], outer), segment("core", [4, 1, 0, 1, 0], core)]
    iface, covered, bare = cap_geometry()
    snaps["partial"] = [segment("iface", [3, 1, 0, 1, 0], iface), segment("covered", [4, 1, 0, 1, 0], covered),
                        segment("bare", [2, 1, 0, 0, 0], bare)]
    mid = arc(0.44, 0.0, math.pi, 8)
    snaps["layered"] = [segment("outer", [3, 1, 0, 2, 0], outer), segment("mid", [5, 1, 0, 2, 1], mid),
                        segment("core", [4, 1, 0, 1, 0], core)]
    tj = math.radians(80.0)
    patch = join(arc(0.5, 0.0, tj, 5), line(arc(0.5, tj, tj, 1)[0], arc(0.44, tj, tj, 1)[0], 2))
    snaps["patch"] = [segment("patch", [3, 1, 0, 2, 0], patch),
                      segment("mid_cov", [5, 1, 0, 2, 1], arc(0.44, 0.0, tj, 5)),
                      segment("mid_bare", [3, 1, 0, 1, 0], arc(0.44, tj, math.pi, 5)),
                      segment("core", [4, 1, 0, 1, 0], core)]
    snaps["layered_n"] = [segment("outer", [3, 1, 0, 3, 0], outer),
                          segment("mid_a", [5, 1, 0, 3, 2], arc(0.46, 0.0, math.pi, 8)),
                          segment("mid_b", [5, 1, 0, 2, 1], arc(0.42, 0.0, math.pi, 8)),
                          segment("core", [4, 1, 0, 1, 0], core)]
    return {k: {"title": k, "segments": v, "ibcs": [], "dielectrics": diel} for k, v in snaps.items()}


def main(ref_dir: str):
    sys.path.insert(0, ref_dir)
    import bor_dispatch as bd
    import bor_solver as bs

    eps, mu = 2.5 - 0.2j, 1.2 - 0.05j
    data = {"freq_hz": F_HZ, "m_max": NM, "thetas": THETAS, "eps": eps, "mu": mu}
    iface, covered, bare = cap_geometry()
    data.update(cap_iface=iface, cap_covered=covered, cap_bare=bare)
    out = bs.solve_bor_partial_coating(iface, covered, [bare], F_HZ, THETAS, eps, mu, n_modes=NM)
    for key in ("amp_vv", "amp_hh"):
        data["cap_" + key] = np.array(out[key])
    data["cap_modes_used"], data["cap_n_unknowns"] = out["modes_used"], out["n_unknowns"]
    data["cap_n_junctions"] = out["n_junctions"]
    iface, covered, bares = band_geometry()
    data.update(band_iface=iface, band_covered=covered, band_bare0=bares[0], band_bare1=bares[1])
    out = bs.solve_bor_partial_coating(iface, covered, bares, F_HZ, THETAS, eps, mu, n_modes=NM)
    for key in ("amp_vv", "amp_hh"):
        data["band_" + key] = np.array(out[key])
    data["band_modes_used"], data["band_n_unknowns"] = out["modes_used"], out["n_unknowns"]
    data["band_n_junctions"] = out["n_junctions"]
    np.savez_compressed(os.path.join(HERE, "partial_multiregion.npz"), **data)

    tj = math.radians(80.0)
    patch = join(arc(0.5, 0.0, tj, 5), line(arc(0.5, tj, tj, 1)[0], arc(0.44, tj, tj, 1)[0], 2))
    mid_cov, mid_bare, core = arc(0.44, 0.0, tj, 5), arc(0.44, tj, math.pi, 5), arc(0.38, 0.0, math.pi, 7)
    e_in, m_in, e_p, m_p = 3.2 - 0.4j, 1.0, 2.5 - 0.2j, 1.2 - 0.05j
    out = bs.solve_bor_coating_patch(patch, mid_cov, mid_bare, core, F_HZ, THETAS, e_in, m_in, e_p, m_p,
                                     n_modes=NM)
    d = {"freq_hz": F_HZ, "m_max": NM, "thetas": THETAS, "patch": patch, "mid_cov": mid_cov,
         "mid_bare": mid_bare, "core": core, "eps_inner": e_in, "mu_inner": m_in, "eps_patch": e_p,
         "mu_patch": m_p, "amp_vv": np.array(out["amp_vv"]), "amp_hh": np.array(out["amp_hh"]),
         "modes_used": out["modes_used"], "n_unknowns": out["n_unknowns"], "n_junctions": out["n_junctions"]}
    np.savez_compressed(os.path.join(HERE, "patch_multiregion.npz"), **d)

    dd = {}
    for name, snap in snapshots().items():
        res = bd.solve_monostatic_rcs_bor(snap, [0.3, 0.36], list(THETAS), "VV", geometry_units="meters",
                                          n_modes=NM, workers=1)
        dd[name + "_amp"] = np.array([s["rcs_amp_real"] + 1j * s["rcs_amp_imag"] for s in res["samples"]])
        dd[name + "_db"] = np.array([s["rcs_db"] for s in res["samples"]])
        dd[name + "_modes_used"] = np.array([m["modes_used"] for m in res["metadata"]["per_frequency"]])
        dd[name + "_n_unknowns"] = np.array([m["n_unknowns"] for m in res["metadata"]["per_frequency"]])
        dd[name + "_formulation"] = res["metadata"]["formulation"]
        dd[name + "_kind"] = res["metadata"]["geometry_kind"]
        print(name, res["metadata"]["formulation"], flush=True)
    np.savez_compressed(os.path.join(HERE, "dispatch_multiregion.npz"), **dd)
    print("multi-region dispatch fixtures written to", HERE)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
