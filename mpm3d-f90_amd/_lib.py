"""ctypes bindings of libmpm3d_hip.so (the C-ABI declared in include/mpm3d_hip.h).

The library is built in-tree (``build.py``) and loaded from this directory.
There is no fallback: if the shared library is missing the import fails
loudly, naming the build step.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmpm3d_hip.so")

# enum mpm_pfield / mpm_gfield (include/mpm3d_problem.h)
PF_NAMES = (["x", "y", "z", "vx", "vy", "vz", "fx", "fy", "fz", "mass", "vol", "sm", "seqv",
             "sdxx", "sdyy", "sdzz", "sdyz", "sdxz", "sdxy", "sig_y", "epeff", "ie", "q", "cp"]
            + ["fp%d%d" % (i + 1, j + 1) for i in range(3) for j in range(3)]
            + ["xx", "xy", "xz", "temp", "dmg", "lt", "fail", "spre", "xo", "yo", "zo", "vol0"]
            + ["pk%d" % (i + 1) for i in range(9)])
PF = {n: i for i, n in enumerate(PF_NAMES)}
PF_COUNT = len(PF_NAMES)
GF_NAMES = ["mg", "px", "py", "pz", "fx", "fy", "fz"]
GF = {n: i for i, n in enumerate(GF_NAMES)}
GF_COUNT = len(GF_NAMES)

SHAPES = {"linear": 0, "gimp": 1, "bspline": 2, "sgmp": 3}
SCHEMES = {"usf": 0, "usl": 1, "musl": 2}
PREC_FP64, PREC_FP32, PREC_FP64_EXACT = 0, 1, 2

MPM_OK = 0
ERRORS = {-1: "MPM_ERR_ARG", -2: "MPM_ERR_IO", -3: "MPM_ERR_DECK", -4: "MPM_ERR_UNSUPPORTED",
          -5: "MPM_ERR_HIP", -6: "MPM_ERR_NOMEM", -7: "MPM_ERR_STATE", -8: "MPM_ERR_NUMERIC"}


class Material(C.Structure):
    _fields_ = [("type", C.c_int32), ("eos_type", C.c_int32), ("density", C.c_double),
                ("young", C.c_double), ("poisson", C.c_double), ("yield0", C.c_double),
                ("tangmod", C.c_double), ("miu", C.c_double), ("bulk_modulus", C.c_double),
                ("wavespd", C.c_double), ("q_fai", C.c_double), ("k_fai", C.c_double),
                ("q_psi", C.c_double), ("ten_f", C.c_double), ("ceos", C.c_double * 10),
                ("det_speed", C.c_double), ("b_jc", C.c_double), ("n_jc", C.c_double), ("c_jc", C.c_double),
                ("m_jc", C.c_double), ("epso", C.c_double), ("epf", C.c_double), ("roomt", C.c_double),
                ("melt", C.c_double), ("spec_heat", C.c_double)]


class Body(C.Structure):
    _fields_ = [("mat", C.c_int32), ("com", C.c_int32), ("par_begin", C.c_int32),
                ("par_end", C.c_int32), ("grav", C.c_double * 3)]


class ProblemStruct(C.Structure):
    _fields_ = [("nb_particle", C.c_int32), ("nb_body", C.c_int32), ("nb_mat", C.c_int32),
                ("shape", C.c_int32), ("scheme", C.c_int32), ("quasi_load", C.c_int32),
                ("quasi_damp", C.c_int32), ("gravity", C.c_int32), ("update_v_by_fp", C.c_int32),
                ("fix", C.c_int32 * 6), ("ncx", C.c_int32), ("ncy", C.c_int32), ("ncz", C.c_int32),
                ("dr_damping", C.c_int32), ("span", C.c_double * 6), ("dcell", C.c_double),
                ("cutoff", C.c_double), ("dt", C.c_double), ("dtscale", C.c_double),
                ("endtime", C.c_double), ("quasi_load_time", C.c_double),
                ("damp_factor", C.c_double), ("bq1", C.c_double), ("bq2", C.c_double),
                ("mats", C.POINTER(Material)), ("bodies", C.POINTER(Body)),
                ("pstate", C.POINTER(C.c_double)), ("pbody", C.POINTER(C.c_int32)),
                ("smooth_stress", C.c_int32), ("tlmpm", C.c_int32), ("contact", C.c_int32),
                ("normbody", C.c_int32), ("nb_component", C.c_int32), ("reserved0", C.c_int32),
                ("fricfa", C.c_double)]


MAX_CURVES, MAX_ANIM = 15, 10


class Outputs(C.Structure):
    """mpm_outputs (include/mpm3d_hip.h): the deck's result-file options"""
    _fields_ = [("title", C.c_char * 257), ("out_time", C.c_double), ("report_time", C.c_double),
                ("tecplot", C.c_int32), ("paraview", C.c_int32), ("n_curves", C.c_int32),
                ("curve_option", C.c_int32 * MAX_CURVES), ("curve_point", C.c_int32 * MAX_CURVES),
                ("n_anim", C.c_int32), ("anim_option", C.c_int32 * MAX_ANIM), ("plot2d", C.c_int32),
                ("skip_first", C.c_int32), ("plot2d_box", C.c_double * 6)]


class Tuning(C.Structure):
    """mpm_tuning (include/mpm3d_engine.h)"""
    _fields_ = [("block_particles", C.c_int32), ("block_nodes", C.c_int32), ("sort_every", C.c_int32),
                ("sort_brick", C.c_int32), ("interleave", C.c_int32), ("trials", C.c_int32),
                ("block_scatter", C.c_int32), ("tile_nodes", C.c_int32), ("step_ms", C.c_double), ("sort_ms", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class MpmError(RuntimeError):
    """A non-zero status from the C-ABI, with the library's message."""

    def __init__(self, code, where, msg):
        self.code = code
        super().__init__("%s failed: %s (%s)" % (where, msg, ERRORS.get(code, code)))


_lib = None


def lib():
    """Load libmpm3d_hip.so (once) and declare its signatures."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError("libmpm3d_hip.so not built: run __graft_entry__.build() "
                          "(hipcc --offload-arch=gfx950) first; expected at " + LIB_PATH)
    L = C.CDLL(LIB_PATH)
    P = C.POINTER
    vp = C.c_void_p
    sig = {
        "mpm_last_error": (C.c_char_p, []),
        "mpm_version": (C.c_char_p, []),
        "mpm_deck_read": (C.c_int, [C.c_char_p, P(P(ProblemStruct)), C.c_char_p, C.c_int]),
        "mpm_problem_free": (None, [P(ProblemStruct)]),
        "mpm_ctx_create": (C.c_int, [P(ProblemStruct), C.c_int, P(vp)]),
        "mpm_ctx_destroy": (C.c_int, [vp]),
        "mpm_begin_step": (C.c_int, [vp]),
        "mpm_grid_momentum_initial": (C.c_int, [vp]),
        "mpm_apply_boundary_conditions": (C.c_int, [vp]),
        "mpm_particle_stress_update": (C.c_int, [vp]),
        "mpm_grid_momentum_update": (C.c_int, [vp]),
        "mpm_integrate_momentum": (C.c_int, [vp]),
        "mpm_particle_position_update": (C.c_int, [vp]),
        "mpm_grid_momentum_musl": (C.c_int, [vp]),
        "mpm_smooth_stress_by_grid": (C.c_int, [vp]),
        "mpm_lagr_nod_contact": (C.c_int, [vp]),
        "mpm_download_grid_component": (C.c_int, [vp, C.c_int, P(C.c_double)]),
        "mpm_get_contact_force": (C.c_int, [vp, P(C.c_double)]),
        "mpm_calc_energy": (C.c_int, [vp, P(C.c_double), P(C.c_double), P(C.c_double)]),
        "mpm_update_dt": (C.c_int, [vp, P(C.c_double)]),
        "mpm_step": (C.c_int, [vp, C.c_int]),
        "mpm_download_particles": (C.c_int, [vp, P(C.c_double)]),
        "mpm_download_grid": (C.c_int, [vp, P(C.c_double)]),
        "mpm_download_icell": (C.c_int, [vp, P(C.c_int32)]),
        "mpm_download_particle_subset": (C.c_int, [vp, C.c_int, P(C.c_int32), P(C.c_double)]),
        "mpm_deck_read_outputs": (C.c_int, [C.c_char_p, P(Outputs), C.c_char_p, C.c_int]),
        "mpm_upload_particles": (C.c_int, [vp, P(C.c_double)]),
        "mpm_get_scalars": (C.c_int, [vp, P(C.c_double)]),
        "mpm_nb_node": (C.c_int, [vp]),
        "mpm_sync": (C.c_int, [vp]),
        # batched engine (include/mpm3d_engine.h)
        "mpm_engine_create": (C.c_int, [C.c_long, C.c_long, C.c_int, C.c_int, C.c_int, C.c_int, P(vp)]),
        "mpm_engine_create_budget": (C.c_int, [C.c_double, C.c_int, C.c_int, P(vp)]),
        "mpm_engine_destroy": (C.c_int, [vp]),
        "mpm_engine_device_bytes": (C.c_long, [vp]),
        "mpm_engine_add": (C.c_int, [vp, P(ProblemStruct), C.c_int, P(C.c_int)]),
        "mpm_engine_remove": (C.c_int, [vp, C.c_int]),
        "mpm_engine_set_step_limit": (C.c_int, [vp, C.c_int, C.c_int]),
        "mpm_engine_capacity": (C.c_int, [vp, P(C.c_long), P(C.c_long), P(C.c_int)]),
        "mpm_engine_step": (C.c_int, [vp, C.c_int, C.c_int, C.c_int]),
        "mpm_engine_sync": (C.c_int, [vp]),
        "mpm_engine_sort": (C.c_int, [vp]),
        "mpm_engine_counters": (C.c_int, [vp, P(C.c_long), P(C.c_long)]),
        "mpm_engine_numeric_flags": (C.c_int, [vp, C.c_int, P(C.c_int), P(C.c_int), P(C.c_int)]),
        "mpm_numeric_flags": (C.c_int, [vp, P(C.c_int), P(C.c_int)]),
        "mpm_cell_order": (C.c_int, [P(ProblemStruct), vp, vp, vp, C.c_long, vp, vp, vp]),
        "mpm_cell_order_noalloc": (C.c_int, [P(ProblemStruct), vp, vp, vp, C.c_long, vp, vp, vp]),
        "mpm_engine_set_tuning": (C.c_int, [vp, C.c_int, C.c_int, C.c_int]),
        "mpm_engine_set_interleave": (C.c_int, [vp, C.c_int]),
        "mpm_engine_get_tuning": (C.c_int, [vp, P(Tuning)]),
        "mpm_engine_apply_tuning": (C.c_int, [vp, P(Tuning)]),
        "mpm_engine_autotune": (C.c_int, [vp, C.c_int, P(Tuning)]),
        "mpm_engine_download_particles": (C.c_int, [vp, C.c_int, P(C.c_double)]),
        "mpm_engine_download_grid": (C.c_int, [vp, C.c_int, P(C.c_double)]),
        "mpm_engine_download_icell": (C.c_int, [vp, C.c_int, P(C.c_int32)]),
        "mpm_engine_get_scalars": (C.c_int, [vp, C.c_int, P(C.c_double)]),
        "mpm_engine_download_order": (C.c_int, [vp, C.c_int, P(C.c_int32)]),
        "mpm_engine_status": (C.c_int, [vp, C.c_int, P(C.c_int), P(C.c_int), P(C.c_int)]),
        "mpm_engine_snapshot_status": (C.c_int, [vp]),
        "mpm_engine_snapshot_read": (C.c_int, [vp, C.c_int, P(C.c_int), P(C.c_int), P(C.c_double)]),
        "mpm_ctx_engine": (C.c_int, [vp, P(vp)]),
        "mpm_engine_profile": (C.c_int, [vp, C.c_int]),
        "mpm_engine_kernel_times": (C.c_int, [vp, P(C.c_double), P(C.c_long)]),
        "mpm_engine_kernel_name": (C.c_char_p, [C.c_int]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def check(code, where):
    if code != MPM_OK:
        msg = lib().mpm_last_error()
        raise MpmError(code, where, msg.decode() if msg else "")
    return code


def dptr(a):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_double))


def iptr(a):
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_int32))
