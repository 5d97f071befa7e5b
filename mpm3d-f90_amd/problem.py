"""Host-side problem description: one MPM3D instance after the reference's
input phase (InputPara, DataIn.f90:34-103), as numpy arrays.

``Problem.from_deck(path)`` runs the library's .mpm reader (mpm_deck_read),
which reproduces the reference's tokenizer and deck semantics; the result is
copied into numpy-owned arrays so it can be edited, batched and passed to a
device context or to the CPU oracle through the same C struct.
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import PF, PF_COUNT, ProblemStruct, Material, Body

SCALARS = ("shape", "scheme", "quasi_load", "quasi_damp", "gravity", "update_v_by_fp",
           "ncx", "ncy", "ncz", "dr_damping", "dcell", "cutoff", "dt", "dtscale", "endtime",
           "quasi_load_time", "damp_factor", "bq1", "bq2", "smooth_stress", "tlmpm", "contact", "normbody",
           "nb_component", "fricfa")


class Problem:
    def __init__(self, n, mats, bodies, **scal):
        self.n = int(n)
        self.mats = mats          # ctypes array of Material
        self.bodies = bodies      # ctypes array of Body
        self.pstate = np.zeros((PF_COUNT, self.n), dtype=np.float64)
        self.pbody = np.zeros(self.n, dtype=np.int32)
        self.fix = np.zeros(6, dtype=np.int32)
        self.span = np.zeros(6, dtype=np.float64)
        self.name = scal.pop("name", "")
        for k in SCALARS:
            setattr(self, k, scal.pop(k, 0))
        if scal:
            raise TypeError("unknown fields %s" % sorted(scal))
        self._keep = None

    # ------------------------------------------------------------------
    @classmethod
    def from_deck(cls, path):
        L = _lib.lib()
        out = C.POINTER(ProblemStruct)()
        err = C.create_string_buffer(512)
        rc = L.mpm_deck_read(str(path).encode(), C.byref(out), err, 512)
        if rc != 0:
            raise _lib.MpmError(rc, "mpm_deck_read(%s)" % path, err.value.decode())
        try:
            return cls._from_struct(out.contents, name=str(path))
        finally:
            L.mpm_problem_free(out)

    @classmethod
    def _from_struct(cls, s, name=""):
        mats = (Material * s.nb_mat)()
        for i in range(s.nb_mat):
            C.memmove(C.byref(mats[i]), C.byref(s.mats[i]), C.sizeof(Material))
        bodies = (Body * s.nb_body)()
        for i in range(s.nb_body):
            C.memmove(C.byref(bodies[i]), C.byref(s.bodies[i]), C.sizeof(Body))
        p = cls(s.nb_particle, mats, bodies, name=name,
                **{k: getattr(s, k) for k in SCALARS})
        p.fix[:] = list(s.fix)
        p.span[:] = list(s.span)
        n = s.nb_particle
        p.pstate[:] = np.ctypeslib.as_array(s.pstate, shape=(PF_COUNT * n,)).reshape(PF_COUNT, n)
        p.pbody[:] = np.ctypeslib.as_array(s.pbody, shape=(n,))
        return p

    # ------------------------------------------------------------------
    def struct(self):
        """C view (ctypes ProblemStruct) over this object's arrays; valid while
        ``self`` is alive and its arrays are not reallocated."""
        self.pstate = np.ascontiguousarray(self.pstate, dtype=np.float64)
        self.pbody = np.ascontiguousarray(self.pbody, dtype=np.int32)
        s = ProblemStruct()
        s.nb_particle = self.n
        s.nb_body = len(self.bodies)
        s.nb_mat = len(self.mats)
        for k in SCALARS:
            setattr(s, k, getattr(self, k))
        for i in range(6):
            s.fix[i] = int(self.fix[i])
            s.span[i] = float(self.span[i])
        s.mats = C.cast(self.mats, C.POINTER(Material))
        s.bodies = C.cast(self.bodies, C.POINTER(Body))
        s.pstate = _lib.dptr(self.pstate)
        s.pbody = _lib.iptr(self.pbody)
        self._keep = s
        return s

    @property
    def nb_node(self):
        return (self.ncx + 1) * (self.ncy + 1) * (self.ncz + 1)

    def field(self, name):
        return self.pstate[PF[name]]

    def copy(self):
        mats = (Material * len(self.mats))()
        C.memmove(mats, self.mats, C.sizeof(mats))
        bodies = (Body * len(self.bodies))()
        C.memmove(bodies, self.bodies, C.sizeof(bodies))
        p = Problem(self.n, mats, bodies, name=self.name, **{k: getattr(self, k) for k in SCALARS})
        p.fix[:] = self.fix
        p.span[:] = self.span
        p.pstate[:] = self.pstate
        p.pbody[:] = self.pbody
        return p

    def __repr__(self):
        return "Problem(%s: %d particles, %d nodes, shape=%d scheme=%d)" % (
            self.name, self.n, self.nb_node, self.shape, self.scheme)
