"""MI355X engine for the MPM3D explicit updated-Lagrangian time step.

Host-side mirror of the reference's step interface.  ``Context`` exposes
one method per argument-less subroutine the reference driver calls
(update_step.f90, called from MPM3D.f90:69-113), with the same names, on
device-resident state; ``Engine`` holds many independent instances and
advances them together (batched ensembles, continuous-batching serving).

All compute runs in libmpm3d_hip.so (HIP, gfx950).  There is no CPU path:
importing works on a machine without a GPU (for deck reading), but creating
a context there raises.
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import (PF, PF_NAMES, PF_COUNT, GF, GF_NAMES, GF_COUNT, SHAPES, SCHEMES, PREC_FP64,
                   PREC_FP32, PREC_FP64_EXACT, MpmError, check, dptr, iptr)
from .problem import Problem
from . import serve

__all__ = ["Problem", "Context", "Engine", "MpmError", "PF", "PF_NAMES", "PF_COUNT", "GF",
           "GF_NAMES", "GF_COUNT", "PREC_FP64", "PREC_FP32", "PREC_FP64_EXACT", "read_deck"]


def read_deck(path):
    """Read a .mpm deck (InputPara semantics) into a Problem."""
    return Problem.from_deck(path)


def _prec(p):
    if p in (PREC_FP64, "fp64", "f64", "double"):
        return PREC_FP64
    if p in (PREC_FP32, "fp32", "f32", "float"):
        return PREC_FP32
    if p in (PREC_FP64_EXACT, "fp64exact", "exact"):
        return PREC_FP64_EXACT
    raise ValueError("precision must be fp64, fp32 or fp64exact")


class Context:
    """One problem instance resident on the current HIP device."""

    def __init__(self, problem, precision="fp64"):
        L = _lib.lib()
        self.problem = problem
        self.n = problem.n
        self.precision = _prec(precision)
        h = C.c_void_p()
        s = problem.struct()
        check(L.mpm_ctx_create(C.byref(s), self.precision, C.byref(h)), "mpm_ctx_create")
        self._h = h
        self.nb_node = L.mpm_nb_node(h)

    def close(self):
        if getattr(self, "_h", None):
            _lib.lib().mpm_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # ---- reference phases (update_step.f90) -----------------------------
    def _call(self, name):
        check(getattr(_lib.lib(), name)(self._h), name)

    def begin_step(self):
        """istep += 1; CurrentTime += DT (MPM3D.f90:72-74)."""
        self._call("mpm_begin_step")

    def GridMomentumInitial(self):
        self._call("mpm_grid_momentum_initial")

    def ApplyBoundaryConditions(self):
        self._call("mpm_apply_boundary_conditions")

    def ParticleStressUpdate(self):
        self._call("mpm_particle_stress_update")

    def GridMomentumUpdate(self):
        self._call("mpm_grid_momentum_update")

    def IntegrateMomentum(self):
        self._call("mpm_integrate_momentum")

    def LagrNodContact(self):
        self._call("mpm_lagr_nod_contact")

    def ParticlePositionUpdate(self):
        self._call("mpm_particle_position_update")

    def GridMomentumMUSL(self):
        self._call("mpm_grid_momentum_musl")

    def SmoothStressbyGrid(self):
        self._call("mpm_smooth_stress_by_grid")

    def calcEnergy(self):
        ek, ei, mom = C.c_double(), C.c_double(), (C.c_double * 3)()
        check(_lib.lib().mpm_calc_energy(self._h, C.byref(ek), C.byref(ei), mom), "mpm_calc_energy")
        return ek.value, ei.value, tuple(mom)

    def UpdateDT(self):
        dt = C.c_double()
        check(_lib.lib().mpm_update_dt(self._h, C.byref(dt)), "mpm_update_dt")
        return dt.value

    def step(self, n=1):
        """n whole steps in the reference order (fused kernels)."""
        check(_lib.lib().mpm_step(self._h, int(n)), "mpm_step")

    def step_phases(self, n=1):
        """n whole steps through the per-phase entry points, in the call
        order of MPM3D.f90:69-113 (what a drop-in host does)."""
        scheme = self.problem.scheme
        for _ in range(n):
            self.begin_step()
            self.GridMomentumInitial()
            if scheme == SCHEMES["usf"]:
                self.ApplyBoundaryConditions()
                self.ParticleStressUpdate()
            self.GridMomentumUpdate()
            self.IntegrateMomentum()
            if self.problem.contact == 1:
                self.LagrNodContact()
            self.ParticlePositionUpdate()
            if scheme == SCHEMES["musl"]:
                self.GridMomentumMUSL()
                self.ApplyBoundaryConditions()
            if scheme != SCHEMES["usf"]:
                self.ParticleStressUpdate()
            self.calcEnergy()
            self.UpdateDT()
            if self.problem.smooth_stress:
                self.SmoothStressbyGrid()

    # ---- state ---------------------------------------------------------
    def particles(self):
        out = np.empty((PF_COUNT, self.n))
        check(_lib.lib().mpm_download_particles(self._h, dptr(out)), "mpm_download_particles")
        return out

    def particle_subset(self, ids):
        """PF_COUNT x len(ids) state of the particles ids (1-based, as the
        reference numbers them), gathered on the device."""
        ids = np.ascontiguousarray(ids, dtype=np.int32)
        out = np.empty((PF_COUNT, len(ids)))
        check(_lib.lib().mpm_download_particle_subset(self._h, len(ids), iptr(ids), dptr(out)),
              "mpm_download_particle_subset")
        return out

    def upload_particles(self, ps):
        ps = np.ascontiguousarray(ps, dtype=np.float64)
        check(_lib.lib().mpm_upload_particles(self._h, dptr(ps)), "mpm_upload_particles")

    def grid(self):
        out = np.empty((GF_COUNT, self.nb_node))
        check(_lib.lib().mpm_download_grid(self._h, dptr(out)), "mpm_download_grid")
        return out

    def grid_component(self, comp):
        """grid of component comp (0-based) of a contact problem"""
        out = np.empty((GF_COUNT, self.nb_node))
        check(_lib.lib().mpm_download_grid_component(self._h, int(comp), dptr(out)), "mpm_download_grid_component")
        return out

    def contact_force(self):
        """tot_cont_for of the last contact pass"""
        out = np.zeros(3)
        check(_lib.lib().mpm_get_contact_force(self._h, dptr(out)), "mpm_get_contact_force")
        return out

    def icell(self):
        out = np.empty(self.n, dtype=np.int32)
        check(_lib.lib().mpm_download_icell(self._h, iptr(out)), "mpm_download_icell")
        return out

    def scalars(self):
        """istep, CurrentTime, DT, EngKinetic, EngInternal, Momentum(3)."""
        out = np.empty(8)
        check(_lib.lib().mpm_get_scalars(self._h, dptr(out)), "mpm_get_scalars")
        return out

    def sync(self):
        check(_lib.lib().mpm_sync(self._h), "mpm_sync")

    def engine_handle(self):
        e = C.c_void_p()
        check(_lib.lib().mpm_ctx_engine(self._h, C.byref(e)), "mpm_ctx_engine")
        return e

    def profile(self, enable=True):
        """Bracket every following kernel launch with HIP events."""
        _profile(self.engine_handle(), enable)

    def kernel_times(self):
        """{kernel class: (total ms, launches)} since profile(True)."""
        return _kernel_times(self.engine_handle())

    def sort(self):
        """Stable re-sort of the particles by cell (as the step does every
        few steps)."""
        check(_lib.lib().mpm_engine_sort(self.engine_handle()), "mpm_engine_sort")

    def counters(self):
        """(steps advanced, whole-pool re-sorts run) so far."""
        return _counters(self.engine_handle())

    def numeric_flags(self):
        """(negative-volume count, non-finite volume/stress count) of the
        stress updates so far (the reference only warns on the first)."""
        nv, nf = C.c_int(), C.c_int()
        check(_lib.lib().mpm_numeric_flags(self._h, C.byref(nv), C.byref(nf)), "mpm_numeric_flags")
        return nv.value, nf.value


def _counters(h):
    st, so = C.c_long(), C.c_long()
    check(_lib.lib().mpm_engine_counters(h, C.byref(st), C.byref(so)), "mpm_engine_counters")
    return st.value, so.value


def _profile(h, enable):
    check(_lib.lib().mpm_engine_profile(h, int(bool(enable))), "mpm_engine_profile")


def _kernel_times(h):
    ms = (C.c_double * 8)()
    n = (C.c_long * 8)()
    check(_lib.lib().mpm_engine_kernel_times(h, ms, n), "mpm_engine_kernel_times")
    return {_lib.lib().mpm_engine_kernel_name(k).decode(): (ms[k], n[k]) for k in range(8)}


class _Sizes:
    def __init__(self, n, nb_node):
        self.n = n
        self.nb_node = nb_node


class Engine:
    """Many independent instances in one device pool (batched ensembles and
    continuous-batching serving)."""

    def __init__(self, max_particles=None, max_nodes=None, max_instances=64, max_materials=None,
                 max_bodies=None, precision="fp64", hbm_bytes=None):
        L = _lib.lib()
        self.precision = _prec(precision)
        h = C.c_void_p()
        if hbm_bytes is not None:
            check(L.mpm_engine_create_budget(float(hbm_bytes), int(max_instances), self.precision,
                                             C.byref(h)), "mpm_engine_create_budget")
        else:
            mm = max_materials or 8 * max_instances
            mb = max_bodies or 8 * max_instances
            check(L.mpm_engine_create(int(max_particles), int(max_nodes), int(max_instances), int(mm),
                                      int(mb), self.precision, C.byref(h)), "mpm_engine_create")
        self._h = h
        self.inst = {}

    def close(self):
        if getattr(self, "_h", None):
            _lib.lib().mpm_engine_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def device_bytes(self):
        return _lib.lib().mpm_engine_device_bytes(self._h)

    def add(self, problem, stop_at_endtime=False):
        """Admit an instance; raises MpmError(code=-6) if it does not fit now."""
        s = problem.struct()
        i = C.c_int()
        check(_lib.lib().mpm_engine_add(self._h, C.byref(s), int(bool(stop_at_endtime)), C.byref(i)),
              "mpm_engine_add")
        # sizes only: the host copy of the state is not needed once uploaded
        self.inst[i.value] = _Sizes(problem.n, problem.nb_node)
        return i.value

    def try_add(self, problem, stop_at_endtime=False):
        try:
            return self.add(problem, stop_at_endtime)
        except MpmError as e:
            if e.code == -6:
                return None
            raise

    def set_step_limit(self, iid, max_steps):
        """Stop the instance after max_steps steps (0: no limit)."""
        check(_lib.lib().mpm_engine_set_step_limit(self._h, int(iid), int(max_steps)),
              "mpm_engine_set_step_limit")

    def remove(self, iid):
        check(_lib.lib().mpm_engine_remove(self._h, int(iid)), "mpm_engine_remove")
        self.inst.pop(iid, None)

    def capacity(self):
        fp, fn, fi = C.c_long(), C.c_long(), C.c_int()
        check(_lib.lib().mpm_engine_capacity(self._h, C.byref(fp), C.byref(fn), C.byref(fi)),
              "mpm_engine_capacity")
        return fp.value, fn.value, fi.value

    def step(self, n=1, graph=False, graph_steps=1):
        """graph: False (plain launches), True (capture once, replay) or
        "lazy" (capture only when the resident set did not change since the
        previous call)."""
        g = 2 if graph == "lazy" else int(bool(graph))
        check(_lib.lib().mpm_engine_step(self._h, int(n), g, int(graph_steps)), "mpm_engine_step")

    def sync(self):
        check(_lib.lib().mpm_engine_sync(self._h), "mpm_engine_sync")

    def sort(self):
        check(_lib.lib().mpm_engine_sort(self._h), "mpm_engine_sort")

    def counters(self):
        """(steps advanced, whole-pool re-sorts run) so far."""
        return _counters(self._h)

    def numeric_flags(self, ids):
        """{id: (negative-volume count, non-finite volume/stress count)}"""
        ids = [int(i) for i in ids]
        a = (C.c_int * len(ids))(*ids)
        nv, nf = (C.c_int * len(ids))(), (C.c_int * len(ids))()
        check(_lib.lib().mpm_engine_numeric_flags(self._h, len(ids), a, nv, nf), "mpm_engine_numeric_flags")
        return {i: (nv[k], nf[k]) for k, i in enumerate(ids)}

    def profile(self, enable=True):
        _profile(self._h, enable)

    def kernel_times(self):
        return _kernel_times(self._h)

    def set_tuning(self, block_particles=256, block_nodes=256, sort_every=0):
        check(_lib.lib().mpm_engine_set_tuning(self._h, int(block_particles), int(block_nodes),
                                               int(sort_every)), "mpm_engine_set_tuning")

    def tuning(self):
        t = _lib.Tuning()
        check(_lib.lib().mpm_engine_get_tuning(self._h, C.byref(t)), "mpm_engine_get_tuning")
        return t.as_dict()

    def apply_tuning(self, **kw):
        t = _lib.Tuning()
        check(_lib.lib().mpm_engine_get_tuning(self._h, C.byref(t)), "mpm_engine_get_tuning")
        for k, v in kw.items():
            setattr(t, k, v)
        check(_lib.lib().mpm_engine_apply_tuning(self._h, C.byref(t)), "mpm_engine_apply_tuning")

    def autotune(self, trial_steps=6):
        """Time launch shapes on the resident instances (state is restored
        afterwards) and apply the fastest; returns the chosen settings."""
        t = _lib.Tuning()
        check(_lib.lib().mpm_engine_autotune(self._h, int(trial_steps), C.byref(t)), "mpm_engine_autotune")
        return t.as_dict()

    def particles(self, iid):
        n = self.inst[iid].n
        out = np.empty((PF_COUNT, n))
        check(_lib.lib().mpm_engine_download_particles(self._h, int(iid), dptr(out)),
              "mpm_engine_download_particles")
        return out

    def grid(self, iid):
        out = np.empty((GF_COUNT, self.inst[iid].nb_node))
        check(_lib.lib().mpm_engine_download_grid(self._h, int(iid), dptr(out)),
              "mpm_engine_download_grid")
        return out

    def icell(self, iid):
        out = np.empty(self.inst[iid].n, dtype=np.int32)
        check(_lib.lib().mpm_engine_download_icell(self._h, int(iid), iptr(out)),
              "mpm_engine_download_icell")
        return out

    def order(self, iid):
        """original particle index held by each device slot of an instance"""
        out = np.empty(self.inst[iid].n, dtype=np.int32)
        check(_lib.lib().mpm_engine_download_order(self._h, int(iid), iptr(out)),
              "mpm_engine_download_order")
        return out

    def scalars(self, iid):
        out = np.empty(8)
        check(_lib.lib().mpm_engine_get_scalars(self._h, int(iid), dptr(out)), "mpm_engine_get_scalars")
        return out

    def snapshot(self):
        """Enqueue a copy of every instance's time state behind the queued
        work (non-blocking); read it with snapshot_read."""
        check(_lib.lib().mpm_engine_snapshot_status(self._h), "mpm_engine_snapshot_status")

    def snapshot_read(self, ids):
        """{id: (done, scalars)} as of the last snapshot (waits for it only)."""
        ids = list(ids)
        n = len(ids)
        a = (C.c_int * max(n, 1))(*ids)
        done = (C.c_int * max(n, 1))()
        out = np.zeros((max(n, 1), 8))
        check(_lib.lib().mpm_engine_snapshot_read(self._h, n, a, done, dptr(out)), "mpm_engine_snapshot_read")
        return {ids[k]: (done[k], out[k].copy()) for k in range(n)}

    def status(self, ids=None):
        ids = list(self.inst) if ids is None else list(ids)
        n = len(ids)
        a = (C.c_int * max(n, 1))(*ids)
        done = (C.c_int * max(n, 1))()
        st = (C.c_int * max(n, 1))()
        check(_lib.lib().mpm_engine_status(self._h, n, a, done, st), "mpm_engine_status")
        return {ids[k]: (done[k], st[k]) for k in range(n)}
