"""Writes the golden fixtures tests/golden/<deck>.npz from the reference.

Runs oracle/_ref/ref_driver (the reference's own step routines compiled
from /root/reference/src by oracle/build_ref.sh, plus the dump driver) on
every deck in tests/golden/decks/ and stores, per deck:

* ``s0_<field>``  particle state after the input phase (InputPara +
  calcEnergy), the fields the deck reader must reproduce;
* ``s10_<field>`` particle state after 10 steps, every field;
* ``g10``         grid state (Mg, PXg, FXg) after 10 steps;
* ``icell10``     cell index of every particle in step 10;
* ``fix``         node fixities;
* ``gc10``        contact decks: the grids of components 2.. after 10 steps,
  ``cont10`` the contact force tot_cont_for of step 10;
* ``scal0`` / ``scal10``: DT, time, EngKinetic, EngInternal, Momentum(3),
  CutOff, DCell, Span(6), DTScale.

Only run in the container that has the reference:
    python tests/golden/make_golden.py [deck ...]
"""
import os
import shutil
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import pyoracle  # noqa: E402

DECKS = ["bar_j2_gimp_musl", "slope_dp_bspline_usf", "cube_elastic_linear_usl",
         "plate_neoh_musl_qulo", "dam_null_eos4_musl",
         "plate_neoh_sgmp_musl", "dam_null_sgmp_musl", "block_j2_sgmp_usf",
         "cube_elastic_linear_usf", "bar_j2_gimp_usf", "slope_dp_bspline_musl", "slope_dp_bspline_usl",
         "block_dp_bspline_shear_usf", "block_dp_bspline_pull_musl",
         "taylor_sjc_eos2_musl", "cube_sjc_eos1_gimp_usl", "impact_sjcf_eos2_bspline_usf", "tnt_hiex_jwl_usl",
         "cube_drda_linear_usf", "cube_smoo_linear_usl", "bar_smoo_gimp_usf", "slope_smoo_bspline_musl",
         "block_smoo_sgmp_usf", "blocks_contact_linear_usf", "blocks_contact_bspline_musl",
         "blocks_contact_gimp_usl", "beam_tl_linear_musl", "block_tl_bspline_usf", "block_tl_gimp_usl",
         "block_tl_sgmp_musl", "taylor_john_linear_musl", "impact_jcf_eos2_gimp_musl", "block_john_sgmp_usf",
         "blocks_tl_contact_bspline_musl", "blocks_outside_linear_usl",
         "blocks_contact_smoo_gimp_musl"]
S0_FIELDS = ["x", "y", "z", "vx", "vy", "vz", "fx", "fy", "fz", "mass", "vol", "sig_y", "ie", "cp", "temp", "lt",
             "xo", "yo", "zo", "vol0"]


def scal(d):
    return np.concatenate([[d["dt"], d["time"], d["ek"], d["ei"]], d["mom"], [d["cutoff"], d["dcell"]],
                           d["span"], [d["dtscale"]]])


def main(names):
    if not os.path.exists(pyoracle.REF_DRIVER):
        subprocess.run([os.path.join(ROOT, "oracle", "build_ref.sh")], check=True)
    for name in names:
        tmp = tempfile.mkdtemp(prefix="golden_")
        try:
            shutil.copy(os.path.join(HERE, "decks", name + ".mpm"), tmp)
            pyoracle.run_ref(os.path.join(tmp, name + ".mpm"), 10, os.path.join(tmp, "d"), [0, 10], cwd=tmp)
            d0 = pyoracle.read_ref_dump(os.path.join(tmp, "d_s00000.bin"))
            d10 = pyoracle.read_ref_dump(os.path.join(tmp, "d_s00010.bin"))
        finally:
            shutil.rmtree(tmp)
        out = {"scal0": scal(d0), "scal10": scal(d10), "g10": d10["grid"], "icell10": d10["icell"],
               "fix": d10["fix"], "ncell": np.array(d10["ncell"])}
        if len(d10["grids"]) > 1:
            out["gc10"] = np.stack(d10["grids"][1:])
            out["cont10"] = d10["cont_force"]
        for f in S0_FIELDS:
            out["s0_" + f] = d0["particles"][f]
        for f, v in d10["particles"].items():
            out["s10_" + f] = v
        path = os.path.join(HERE, name + ".npz")
        np.savez_compressed(path, **out)
        print(path, os.path.getsize(path))


if __name__ == "__main__":
    main(sys.argv[1:] or DECKS)   # deck names, default all
