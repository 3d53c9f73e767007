"""Write golden fixtures for the dispatch front end from the reference itself.

Run in the development container only (the reference sources are not
present where the GPU tests run):

    PYTHONDONTWRITEBYTECODE=1 python tests/golden/make_golden_dispatch.py /root/reference

It imports the reference's geometry_io, rcs_solver (MaterialLibrary),
bor_dispatch and grim_io and writes, next to this script:

* geo_parse.json: the reference parser's snapshot of a .geo text written
  here (three TYPE 2 segments with impedance rows, one tabulated);
* materials.npz: MaterialLibrary impedances and media at chosen flags,
  frequencies and arc positions (inline tapers of every kind, a table);
* dispatch_ibc.npz: solve_monostatic_rcs_bor on that geometry (an IBC
  sphere with a tapered and a tabulated flag) at two frequencies, VV and HH;
* az_el.npz: bor_az_el_grid of a PEC sweep's VV / HH results (their samples
  stored too, so the grid is pinned without a solve);
* adaptive.npz: solve_adaptive_frequency_sweep_bor on a small PEC sphere;
* grim_mono.npz: the arrays export_result_to_grim writes for the PEC sweep.
"""

import json
import math
import os
import shutil
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
MAT_DIR = os.path.join(HERE, "materials")


def arc_pairs(r, t0, t1, n, x0=0.0):
    th = np.linspace(t0, t1, n + 1)
    x, y = r * np.sin(th), r * np.cos(th)
    return [(float(x[i]), float(y[i]), float(x[i + 1]), float(y[i + 1])) for i in range(n)]


def geo_text():
    """A 0.3 m sphere in metres: a tapered-impedance cap, a PEC band and a
    tabulated-impedance cap (flags 2 and 57), plus a dielectric row."""
    lines = ["Title: ibc sphere", "# three conductor segments, nose to tail"]
    segs = [("cap_north", "2 3 2 0 0", arc_pairs(0.3, 0.0, 0.9, 6)),
            ("band", "2 -12 0 0 0", arc_pairs(0.3, 0.9, 2.2, 8)),
            ("cap_south", "2 3 57 0 0", arc_pairs(0.3, 2.2, math.pi, 5))]
    for name, props, prims in segs:
        lines.append(f"Segment: {name} 2")
        lines.append(f"properties: {props}")
        lines += [f"{a!r} {b!r} {c!r} {d!r}" for a, b, c, d in prims]
    lines += ["IBCS:", "2 cosine 60 -40 250 -90", "57", "Dielectrics:", "1 2.5 -0.1 1.0 0.0"]
    return "\n".join(lines) + "\n"


def write_tables():
    os.makedirs(MAT_DIR, exist_ok=True)
    with open(os.path.join(MAT_DIR, "mat.57"), "w") as f:
        f.write("# f_GHz  R  X\n0.2 120 -60\n0.5 160 -35\n1.0 210 -10\n")
    with open(os.path.join(MAT_DIR, "mat.61"), "w") as f:
        f.write("# f_GHz eps' eps'' mu' mu''\n0.1 4.0 -0.2 1.0 0.0\n1.0 3.6 -0.3 1.1 -0.05\n")


def main(ref_dir: str):
    sys.path.insert(0, ref_dir)
    import bor_dispatch as bd
    import geometry_io as gio
    import grim_io
    from rcs_solver import MaterialLibrary

    write_tables()
    text = geo_text()
    with open(os.path.join(HERE, "ibc_sphere.geo"), "w") as f:
        f.write(text)
    title, segs, ibcs, diel = gio.parse_geometry(text)
    snap = gio.build_geometry_snapshot(title, segs, ibcs, diel)
    with open(os.path.join(HERE, "geo_parse.json"), "w") as f:
        json.dump(snap, f, indent=1)

    # materials
    lib = MaterialLibrary.from_entries(
        [["3", "linear", "50", "-10", "300", "-80"], ["4", "exp", "20", "0", "400", "-50"],
         ["5", "constant", "77", "-7", "0", "0"], ["6", "cosine", "0", "0", "200", "100"], ["57"]],
        [["1", "2.5", "-0.1", "1.0", "0.0"], ["2", "0", "0"], ["61"]], base_dir=MAT_DIR)
    grid = [(fl, f, s) for fl in (3, 4, 5, 6, 57, 0, 9) for f in (0.1, 0.35, 0.8, 1.5) for s in (0.0, 0.3, 1.0)]
    imp = np.array([lib.get_impedance(fl, f, arc_s=s) for fl, f, s in grid])
    med = np.array([lib.get_medium(fl, f) for fl in (1, 2, 61, 0, 9) for f in (0.05, 0.4, 1.0, 2.0)])
    np.savez_compressed(os.path.join(HERE, "materials.npz"), imp_args=np.array(grid, dtype=float),
                        imp=imp, med=med, warnings=np.array(lib.warnings))

    # dispatch with impedance flags
    data = {}
    for pol in ("VV", "HH"):
        out = bd.solve_monostatic_rcs_bor(snap, [0.4, 0.7], [0.0, 45.0, 90.0, 135.0, 180.0], pol,
                                          geometry_units="meters", material_base_dir=MAT_DIR, n_modes=5,
                                          workers=1, table_precision="double")
        data[f"{pol}_rcs_db"] = np.array([s["rcs_db"] for s in out["samples"]])
        data[f"{pol}_amp"] = np.array([s["rcs_amp_real"] + 1j * s["rcs_amp_imag"] for s in out["samples"]])
        data[f"{pol}_modes_used"] = np.array([m["modes_used"] for m in out["metadata"]["per_frequency"]])
        data["formulation"] = out["metadata"]["formulation"]
    np.savez_compressed(os.path.join(HERE, "dispatch_ibc.npz"), **data)

    # PEC sweep for the az/el grid and the .grim export
    sys.path.insert(0, HERE)
    from make_golden import dispatch_snapshot
    pec = dispatch_snapshot()
    th = list(np.arange(0.0, 181.0, 10.0))
    rv = bd.solve_monostatic_rcs_bor(pec, [0.3, 0.45], th, "VV", geometry_units="meters", n_modes=4, workers=1)
    rh = bd.solve_monostatic_rcs_bor(pec, [0.3, 0.45], th, "HH", geometry_units="meters", n_modes=4, workers=1)
    g = bd.bor_az_el_grid(rv, rh, np.arange(-60.0, 61.0, 20.0), np.arange(-30.0, 31.0, 15.0),
                          axis_az_deg=10.0, axis_el_deg=20.0)
    keys = ["frequency_ghz", "theta_inc_deg", "rcs_amp_real", "rcs_amp_imag"]
    np.savez_compressed(os.path.join(HERE, "az_el.npz"),
                        vv=np.array([[s[k] for k in keys] for s in rv["samples"]]),
                        hh=np.array([[s[k] for k in keys] for s in rh["samples"]]),
                        theta_map=g["theta_map_deg"], VV=g["amp"]["VV"], HH=g["amp"]["HH"], VH=g["amp"]["VH"])
    tmp = tempfile.mkdtemp()
    try:
        path = grim_io.export_result_to_grim(rv, os.path.join(tmp, "sweep"), source_path="s", history="h")[0]
        with np.load(path) as z:
            np.savez_compressed(os.path.join(HERE, "grim_mono.npz"), **{k: z[k] for k in z.files})
    finally:
        shutil.rmtree(tmp)
    with open(os.path.join(HERE, "pec_sweep_vv.json"), "w") as f:
        json.dump(rv, f)

    # adaptive frequency sweep
    ad = bd.solve_adaptive_frequency_sweep_bor(pec, 0.2, 0.6, [0.0, 90.0], "VV", initial_points=4,
                                               max_refinements=2, db_threshold=0.5,
                                               geometry_units="meters", n_modes=4, workers=1)
    np.savez_compressed(os.path.join(HERE, "adaptive.npz"),
                        freqs=np.array([s["frequency_ghz"] for s in ad["samples"]]),
                        theta=np.array([s["theta_inc_deg"] for s in ad["samples"]]),
                        rcs_db=np.array([s["rcs_db"] for s in ad["samples"]]),
                        count=ad["metadata"]["final_point_count"],
                        rounds=ad["metadata"]["refinement_count"])
    print("dispatch fixtures written to", HERE)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
