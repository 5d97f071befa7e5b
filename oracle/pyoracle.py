"""ctypes wrapper of the CPU oracle (oracle/_build/libmpm_oracle.so) and a
reader for the reference dumps written by oracle/ref_driver.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker.  The product never loads it.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_build", "libmpm_oracle.so")
REF_DRIVER = os.path.join(HERE, "_ref", "ref_driver")

PF_COUNT = 54   # enum mpm_pfield (include/mpm3d_problem.h)
GF_COUNT = 7


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_L = None


def lib():
    global _L
    if _L is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        vp = C.c_void_p
        for name in ("oracle_begin_step", "oracle_grid_momentum_initial", "oracle_apply_bc",
                     "oracle_particle_stress_update", "oracle_grid_momentum_update",
                     "oracle_integrate_momentum", "oracle_particle_position_update",
                     "oracle_grid_momentum_musl", "oracle_calc_energy", "oracle_update_dt",
                     "oracle_smooth_stress", "oracle_contact", "oracle_tl_grid_mass_initial",
                     "oracle_tl_grid_momentum_initial", "oracle_tl_particle_stress_update",
                     "oracle_tl_grid_momentum_update", "oracle_destroy"):
            getattr(L, name).argtypes = [vp]
            getattr(L, name).restype = None
        L.oracle_create.argtypes = [vp]
        L.oracle_create.restype = vp
        L.oracle_step.argtypes = [vp, C.c_int]
        L.oracle_nb_node.argtypes = [vp]
        L.oracle_nb_node.restype = C.c_int
        for name in ("oracle_get_particles", "oracle_get_grid", "oracle_get_scalars",
                     "oracle_set_particles"):
            getattr(L, name).argtypes = [vp, C.POINTER(C.c_double)]
        L.oracle_get_icell.argtypes = [vp, C.POINTER(C.c_int32)]
        L.oracle_get_grid_component.argtypes = [vp, C.c_int, C.POINTER(C.c_double)]
        L.oracle_get_contact_force.argtypes = [vp, C.POINTER(C.c_double)]
        L.oracle_get_contact_counts.argtypes = [vp, C.POINTER(C.c_long)]
        L.oracle_nb_component.argtypes = [vp]
        L.oracle_nb_component.restype = C.c_int
        L.oracle_get_fix.argtypes = [vp, C.POINTER(C.c_int8)]
        L.oracle_get_dp_branches.argtypes = [vp, C.POINTER(C.c_int8), C.POINTER(C.c_long)]
        L.oracle_get_jc_counts.argtypes = [vp, C.POINTER(C.c_long)]
        L.oracle_set_fresh_depeff.argtypes = [vp, C.c_int]
        L.oracle_set_log_cr.argtypes = [vp, C.c_int]
        L.oracle_log_cr.argtypes = [C.c_double]
        L.oracle_log_cr.restype = C.c_double
        L.oracle_exp_cr.argtypes = [C.c_double]
        L.oracle_exp_cr.restype = C.c_double
        L.oracle_pow_cr.argtypes = [C.c_double, C.c_double]
        L.oracle_pow_cr.restype = C.c_double
        dp = C.POINTER(C.c_double)
        L.oracle_dp_point.argtypes = [vp, C.c_double, C.c_double, dp, dp, dp, dp]
        L.oracle_dp_point.restype = C.c_int
        _L = L
    return _L


class Oracle:
    """FP64 CPU restatement of one instance; method names follow update_step.f90."""

    def __init__(self, problem, log_cr=False):
        """log_cr: the neo-Hookean ln J from the correctly rounded logarithm
        (the device's reference-order mode) instead of the C library's."""
        self._s = problem.struct()
        self._prob = problem
        self.h = lib().oracle_create(C.cast(C.pointer(self._s), C.c_void_p))
        if log_cr:
            lib().oracle_set_log_cr(self.h, 1)
        self.n = problem.n
        self.nnode = lib().oracle_nb_node(self.h)

    def __del__(self):
        if getattr(self, "h", None):
            lib().oracle_destroy(self.h)
            self.h = None

    def step(self, n=1):
        lib().oracle_step(self.h, n)

    def phase(self, name):
        getattr(lib(), "oracle_" + name)(self.h)

    def particles(self):
        out = np.empty((PF_COUNT, self.n))
        lib().oracle_get_particles(self.h, out.ctypes.data_as(C.POINTER(C.c_double)))
        return out

    def set_particles(self, ps):
        ps = np.ascontiguousarray(ps, dtype=np.float64)
        lib().oracle_set_particles(self.h, ps.ctypes.data_as(C.POINTER(C.c_double)))

    def grid(self):
        out = np.empty((GF_COUNT, self.nnode))
        lib().oracle_get_grid(self.h, out.ctypes.data_as(C.POINTER(C.c_double)))
        return out

    def grid_component(self, c):
        """grid of component c (0-based) of a contact problem"""
        out = np.empty((GF_COUNT, self.nnode))
        lib().oracle_get_grid_component(self.h, int(c), out.ctypes.data_as(C.POINTER(C.c_double)))
        return out

    def nb_component(self):
        return lib().oracle_nb_component(self.h)

    def contact_force(self):
        """tot_cont_for of the last Lagr_NodContact"""
        out = np.zeros(3)
        lib().oracle_get_contact_force(self.h, out.ctypes.data_as(C.POINTER(C.c_double)))
        return out

    def contact_counts(self):
        """contact nodes over all steps: [frictionless, sticking, slipping]
        (oracle diagnostics of Lagr_NodContact's branches)"""
        out = np.zeros(3, dtype=np.int64)
        lib().oracle_get_contact_counts(self.h, out.ctypes.data_as(C.POINTER(C.c_long)))
        return out

    def icell(self):
        out = np.empty(self.n, dtype=np.int32)
        lib().oracle_get_icell(self.h, out.ctypes.data_as(C.POINTER(C.c_int32)))
        return out

    def fix(self):
        out = np.empty((3, self.nnode), dtype=np.int8)
        lib().oracle_get_fix(self.h, out.ctypes.data_as(C.POINTER(C.c_int8)))
        return out

    def dp_branches(self):
        """(branch of every particle's last Drucker-Prager update, running
        counts [elastic, shear with sm < tenf, shear with sm >= tenf, tension
        cap]) -- oracle diagnostics of M3DM9's return map."""
        last = np.zeros(self.n, dtype=np.int8)
        cnt = np.zeros(4, dtype=np.int64)
        lib().oracle_get_dp_branches(self.h, last.ctypes.data_as(C.POINTER(C.c_int8)),
                                     cnt.ctypes.data_as(C.POINTER(C.c_long)))
        return last, cnt

    def set_fresh_depeff(self, on=True):
        """test switch (not the reference): start every particle's update
        from depeff = 0 instead of the value the last update left"""
        lib().oracle_set_fresh_depeff(self.h, 1 if on else 0)

    def jc_counts(self):
        """running counts [john/jcf updates that read another particle's
        depeff, jcf failures by melting, by the failure strain, by damage]
        -- oracle diagnostics of the Johnson-Cook models."""
        cnt = np.zeros(4, dtype=np.int64)
        lib().oracle_get_jc_counts(self.h, cnt.ctypes.data_as(C.POINTER(C.c_long)))
        return cnt

    def scalars(self):
        out = np.empty(8)
        lib().oracle_get_scalars(self.h, out.ctypes.data_as(C.POINTER(C.c_double)))
        return out


def dp_point(material, young, poisson, sd, sm):
    """M3DM9 on one stress point with a zero strain increment.
    material: an mpm_material ctypes struct (q_fai, k_fai, q_psi, ten_f set);
    sd: 6 deviatoric components (xx yy zz yz xz xy).  Returns
    (branch, sd_new, sm_new, seqv, depeff)."""
    sdv = (C.c_double * 6)(*[float(v) for v in sd])
    smv, seqv, dep = C.c_double(sm), C.c_double(0.0), C.c_double(0.0)
    br = lib().oracle_dp_point(C.cast(C.pointer(material), C.c_void_p), young, poisson, sdv, C.byref(smv),
                               C.byref(seqv), C.byref(dep))
    return br, np.array(sdv[:]), smv.value, seqv.value, dep.value


# ---------------------------------------------------------------- ref dumps
# particle field order written by ref_driver.f90 -> index in PF layout
_REF_PFIELDS = (["x", "y", "z", "xx", "xy", "xz", "vx", "vy", "vz", "fx", "fy", "fz", "mass", "vol",
                 "sm", "seqv", "sdxx", "sdyy", "sdzz", "sdyz", "sdxz", "sdxy", "sig_y", "epeff",
                 "ie", "q", "cp", "temp", "dmg", "lt", "fail", "spre", "xo", "yo", "zo", "vol0"]
                + ["pk%d" % (i + 1) for i in range(9)]
                + ["fp%d%d" % (i + 1, j + 1) for j in range(3) for i in range(3)])


def read_ref_dump(path):
    """Parse one ref_driver dump -> dict(hdr, scal, particles{name: arr}, icell, grid[7,nn], fix[3,nn])."""
    raw = open(path, "rb").read()
    hdr = np.frombuffer(raw, dtype="<i4", count=16)
    assert hdr[0] == 1146113101 and hdr[1] == 5, "not a ref_driver dump (layout 5): %s" % path
    ncomp = int(hdr[11])
    n, nn = int(hdr[2]), int(hdr[3])
    scal = np.frombuffer(raw, dtype="<f8", count=16, offset=64)
    off = 64 + 128
    arrs = np.frombuffer(raw, dtype="<f8", offset=off)
    parts = {}
    pos = 0
    for name in _REF_PFIELDS:
        parts[name] = arrs[pos:pos + n].copy()
        pos += n
    icell = arrs[pos:pos + n].astype(np.int32)
    pos += n
    grid = arrs[pos:pos + 7 * nn].reshape(7, nn).copy()
    pos += 7 * nn
    fix = arrs[pos:pos + 3 * nn].reshape(3, nn).astype(np.int8)
    pos += 3 * nn
    grids = [grid]
    for _ in range(1, ncomp):
        grids.append(arrs[pos:pos + 7 * nn].reshape(7, nn).copy())
        pos += 7 * nn
    cont_force = arrs[pos:pos + 3].copy()
    pos += 3
    assert pos == arrs.size, "dump size mismatch"
    return {"hdr": hdr, "istep": int(hdr[4]), "phase": int(hdr[5]), "n": n, "nnode": nn,
            "ncell": (int(hdr[6]), int(hdr[7]), int(hdr[8])),
            "dt": scal[0], "time": scal[1], "ek": scal[2], "ei": scal[3], "mom": scal[4:7].copy(),
            "cutoff": scal[7], "dcell": scal[8], "span": scal[9:15].copy(), "dtscale": scal[15],
            "particles": parts, "icell": icell, "grid": grid, "fix": fix, "grids": grids,
            "cont_force": cont_force}


def run_ref(deck, nsteps, outprefix, dumpsteps, phasestep=None, cwd=None, timeout=600):
    """Run oracle/_ref/ref_driver (built from the reference sources) on a deck."""
    args = [REF_DRIVER, os.path.basename(deck), str(nsteps), outprefix,
            ",".join(str(s) for s in dumpsteps)]
    if phasestep is not None:
        args.append(str(phasestep))
    r = subprocess.run(args, cwd=cwd or os.path.dirname(os.path.abspath(deck)),
                       capture_output=True, text=True, timeout=timeout)
    if r.returncode != 0:
        raise RuntimeError("ref_driver failed: %s\n%s" % (r.stdout[-2000:], r.stderr[-2000:]))
    return r.stdout
