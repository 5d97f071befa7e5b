"""The reference's result files, written from device state by a Python host.

The reference's driver writes, besides its message file (<name>.out):
  <name>_curv.dat      one line per step: time, then one column per curve
                       (OutCurve, DataOut.f90:66-154)
  EnergyPlot.dat       time, total, kinetic and internal energy per step
  MomentumPlot.dat     time and momentum per step
  ContforcPlot.dat     contact force per step (contact decks; empty otherwise)
  <name>_anim.dat      TecPlot frames at the output times (OutAnim,
                       DataOut.f90:157-247), with ``tecp``
  <name>_anim_<k>.vtu  ParaView frames, collected in <name>_anim.pvd, and
  <name>_Grid.vtu      the grid (OutAnimPV, OutGrid, DataOut.f90:250-608),
                       with ``para``
on the schedule of MPM3D.f90:34-113: a frame at time 0, one curve line per
step, a frame whenever CurrentTime reaches plt (a default REAL advanced by
OutTime).  The Fortran drop-in runs those writers themselves; hosts without
Fortran use this module, which produces the same characters: Fortran's Ew.d
editing as flang writes it (mantissa 0.dddd, exponent letter dropped for
three-digit exponents, Inf/NaN right-justified, ties of the exact binary
value to even), Aw truncation and padding, and the list-directed records'
leading blank.

``read_outputs(path)`` gives the deck's options (the library's deck reader,
mpm_deck_read_outputs); ``run(ctx, outputs, name)`` steps a Context to the
end time writing every file; ``ResultWriter`` is the writer alone, for hosts
with their own loop.  Particle state comes from the device: per step only
the curve particles (mpm_download_particle_subset), everything at frames.

Temperature, damage, the failure flag and the smoothed pressure (smoo) are
particle state (PF temp, dmg, fail, spre; spre stays 0 without smoo).  TecPlot
frames leave out the particles pt2d marks (SkipThis, DataIn.f90:867-880).
One difference: a curve of option 12 (fail) writes the logical with an I
edit descriptor, which the flang-built reference refuses at run time; here it
prints the flag as 0 or 1.
"""
import ctypes as C
import math
import os

import numpy as np

from . import _lib
from ._lib import PF, Outputs

# DataOut.f90:30-34
OUTPUT_NAMES = ("seqv", "epef", "mat ", "pres", "volu", "engk", "engi", "velx", "vely", "velz",
                "cels", "fail", "sspd", "damg", "stsx", "disx", "disy", "disz")
INDENT = "    "


def read_outputs(path):
    """The result-file options of a .mpm deck (mpm_outputs)."""
    out = Outputs()
    err = C.create_string_buffer(512)
    rc = _lib.lib().mpm_deck_read_outputs(str(path).encode(), C.byref(out), err, 512)
    if rc != 0:
        raise _lib.MpmError(rc, "mpm_deck_read_outputs(%s)" % path, err.value.decode())
    return out


# ---- Fortran edit descriptors --------------------------------------------
def fe(v, w, d):
    """Ew.d: [-]0.<d digits>E+xx, or [-]0.<d digits>+xxx for |exponent| > 99."""
    v = float(v)
    if math.isnan(v):
        s = "NaN"
    elif math.isinf(v):
        s = "Inf" if v > 0 else "-Inf"
    else:
        sign = "-" if math.copysign(1.0, v) < 0 else ""
        if v == 0.0:
            digits, e = "0" * d, 0
        else:
            m, x = ("%.*e" % (d - 1, abs(v))).split("e")
            digits, e = m.replace(".", ""), int(x) + 1
        s = sign + "0." + digits + ("E%+03d" % e if abs(e) <= 99 else "%+04d" % e)
    return s.rjust(w) if len(s) <= w else "*" * w


def fa(s, w):
    """Aw: the leftmost w characters, or right-justified in w."""
    return s[:w] if len(s) >= w else s.rjust(w)


def fi(v, w):
    s = "%d" % v
    return s.rjust(w) if len(s) <= w else "*" * w


# ---- grid of the reference's SetGridData (Grid.f90:250-380) ---------------
def grid_nodes(problem):
    """node coordinates Xg, node order (iz, iy, ix) with ix fastest"""
    ngx, ngy, ngz = problem.ncx + 1, problem.ncy + 1, problem.ncz + 1
    iz, iy, ix = np.meshgrid(np.arange(ngz), np.arange(ngy), np.arange(ngx), indexing="ij")
    sp, dc = problem.span, problem.dcell
    return np.stack([ix.ravel() * dc + sp[0], iy.ravel() * dc + sp[2], iz.ravel() * dc + sp[4]])


def cells_node(problem):
    """CellsNode(NumCell, 8), 1-based node numbers"""
    ngx, ngxy = problem.ncx + 1, (problem.ncx + 1) * (problem.ncy + 1)
    iz, iy, ix = np.meshgrid(np.arange(problem.ncz), np.arange(problem.ncy), np.arange(problem.ncx),
                             indexing="ij")
    n1 = (iz * ngxy + iy * ngx + ix + 1).ravel()
    n5 = n1 + ngxy
    return np.stack([n1, n1 + 1, n1 + 1 + ngx, n1 + ngx, n5, n5 + 1, n5 + 1 + ngx, n5 + ngx], axis=1)


class ResultWriter:
    """The reference's output files for one problem.

    open() creates them as InputPara does (DataIn.f90:58-94); curve() is
    OutCurve, frame() is OutAnim and OutAnimPV; close() ends the ParaView
    collection as MPM3D.f90:289-293 does.  Particle arrays are PF_COUNT x n
    in the reference's particle order (Context.particles())."""

    def __init__(self, problem, outputs, name, directory="."):
        self.problem = problem
        self.o = outputs
        self.name = name
        self.dir = directory
        self.curves = [(outputs.curve_option[i], outputs.curve_point[i]) for i in range(outputs.n_curves)]
        self.anim = [outputs.anim_option[i] for i in range(outputs.n_anim)]
        for opt, pt in self.curves:
            if not 1 <= opt <= len(OUTPUT_NAMES) or not 1 <= pt <= problem.n:
                raise ValueError("curve option %d of particle %d out of range" % (opt, pt))
        self.xo = problem.pstate[[PF["x"], PF["y"], PF["z"]]].copy()   # Xo (DataIn.f90:859-863)
        # SkipThis (setskip, DataIn.f90:867-880): outside the pt2d box at generation
        self.skip = np.zeros(problem.n, dtype=bool)
        if outputs.plot2d:
            b = np.array(outputs.plot2d_box[:])
            x = self.xo[:, outputs.skip_first:]
            self.skip[outputs.skip_first:] = ((x[0] < b[0]) | (x[0] > b[1]) | (x[1] < b[2]) | (x[1] > b[3])
                                              | (x[2] < b[4]) | (x[2] > b[5]))
        self.mat = np.zeros(problem.n, dtype=np.int64)
        for b in problem.bodies:
            self.mat[b.par_begin:b.par_end] = b.mat
        self.n_cvs = 0
        self.n_anm = 0
        self.plot_step = 1
        self.files = {}

    def _path(self, f):
        return os.path.join(self.dir, f)

    def open(self):
        f = self.files
        f["curv"] = open(self._path(self.name + "_curv.dat"), "w")
        if self.o.tecplot:
            f["anim"] = open(self._path(self.name + "_anim.dat"), "w")
        if self.o.paraview:
            f["pvd"] = open(self._path(self.name + "_anim.pvd"), "w")
            f["pvd"].write('<?xml version="1.0"?>\n<VTKFile type="Collection" version="0.1" byte_order="BigEndian"'
                           ' compressor="vtkZLibDataCompressor">\n    <Collection>\n')
        f["energy"] = open(self._path("EnergyPlot.dat"), "w")
        f["momentum"] = open(self._path("MomentumPlot.dat"), "w")
        f["contact"] = open(self._path("ContforcPlot.dat"), "w")
        if self.o.paraview:
            self.write_grid()
        return self

    def close(self):
        if "pvd" in self.files:
            self.files["pvd"].write("    </Collection>\n</VTKFile>\n")
        for h in self.files.values():
            h.close()
        self.files = {}

    def __enter__(self):
        return self.open()

    def __exit__(self, *a):
        self.close()

    # ---- OutCurve --------------------------------------------------------
    def curve_points(self):
        """1-based particles whose state curve() reads, in curve order"""
        return [pt for _, pt in self.curves]

    def curve(self, time, eng_kinetic, eng_internal, momentum, points, contact_force=None):
        """one step's line; points: PF_COUNT x n_curves state of
        curve_points(); contact_force: tot_cont_for of a contact problem"""
        out = self.files["curv"]
        self.n_cvs += 1
        if self.n_cvs == 1:
            # the number left-justified in a character(20), written a8
            out.write("Time " + "".join(OUTPUT_NAMES[o - 1] + fa(("%d" % pt).ljust(20), 8) for o, pt in self.curves)
                      + "\n")
        line = [fe(time, 12, 4)]
        for k, (opt, pt) in enumerate(self.curves):
            line.append(self._curve_value(opt, points[:, k], self.xo[:, pt - 1], eng_kinetic, eng_internal))
        out.write("".join(line) + "\n")
        self.files["energy"].write("".join(fe(v, 12, 4) for v in (time, eng_kinetic + eng_internal, eng_kinetic,
                                                                  eng_internal)) + "\n")
        self.files["momentum"].write("".join(fe(v, 12, 4) for v in (time,) + tuple(momentum)) + "\n")
        if self.problem.contact:   # DataOut.f90:148-151
            cf = tuple(contact_force) if contact_force is not None else (0.0, 0.0, 0.0)
            self.files["contact"].write("".join(fe(v, 12, 4) for v in (time,) + cf) + "\n")

    def _curve_value(self, opt, p, xo, ek, ei):
        # DataOut.f90:94-137
        if opt == 12:
            return fi(int(p[PF["fail"]] != 0), 12)
        v = {1: lambda: p[PF["seqv"]], 2: lambda: p[PF["epeff"]], 3: lambda: p[PF["epeff"]],
             4: lambda: -p[PF["sm"]], 5: lambda: p[PF["vol"]], 6: lambda: ek, 7: lambda: ei,
             8: lambda: p[PF["vx"]], 9: lambda: p[PF["vy"]], 10: lambda: p[PF["vz"]], 11: lambda: p[PF["temp"]],
             13: lambda: p[PF["sdzz"]], 14: lambda: p[PF["dmg"]], 15: lambda: p[PF["sdxx"]] + p[PF["sm"]],
             16: lambda: p[PF["x"]] - xo[0], 17: lambda: p[PF["y"]] - xo[1], 18: lambda: p[PF["z"]] - xo[2]}
        return fe(v[opt](), 12, 4)

    # ---- frames -------------------------------------------------------------
    def frame(self, time, particles, icell=None):
        """the output-time block of MPM3D.f90:93-100: OutAnim with tecp,
        OutAnimPV with para; icell None means before the first step (the
        reference's initial icell 0: every particle inside)"""
        if self.o.tecplot:
            self.write_tecplot(time, particles, icell)
        if self.o.paraview:
            self.write_paraview(time, particles, self.plot_step)
        self.plot_step += 1

    def write_tecplot(self, time, ps, icell):
        out = self.files["anim"]
        self.n_anm += 1
        if self.n_anm == 1:
            title = self.o.title.decode("latin-1")
            out.write('TITLE = "' + fa(title, 60) + '"\n')
            items = ['"' + OUTPUT_NAMES[a - 1].rstrip() + '"' for a in self.anim]
            widths = (2, 4, 1)
            out.write('VARIABLES= "X"   "Y"   "Z"  ' + "".join(fa(s, widths[k % 3]) for k, s in enumerate(items))
                      + "\n")
        out.write('ZONE T="time' + fe(time, 10, 3) + '"\n')
        keep = ~self.skip if icell is None else (np.asarray(icell) >= 0) & ~self.skip
        cols = []
        for a in self.anim:
            cols.append(self._anim_column(a, ps))
        xyz = [ps[PF["x"]], ps[PF["y"]], ps[PF["z"]]]
        lines = []
        for p in np.nonzero(keep)[0]:
            s = fe(xyz[0][p], 12, 4) + fe(xyz[1][p], 12, 4) + fe(xyz[2][p], 12, 4)
            s += "".join(c(p) for c in cols if c is not None)
            lines.append(s + "\n")
        out.write("".join(lines))

    def _anim_column(self, a, ps):
        # DataOut.f90:194-235; options 17 and 18 write nothing
        g = lambda k: ps[PF[k]]
        if a == 3:
            return lambda p: fi(self.mat[p], 3)
        if a == 12:
            fail = ps[PF["fail"]] != 0
            return lambda p: fi(int(fail[p]), 3)
        arr = {1: lambda: g("seqv"), 2: lambda: g("epeff"), 4: lambda: -g("sm"), 5: lambda: g("vol"),
               6: lambda: 0.5 * g("mass") * (g("vx") ** 2 + g("vy") ** 2 + g("vz") ** 2),
               7: lambda: g("ie"), 8: lambda: g("vx"), 9: lambda: g("vy"), 10: lambda: g("vz"),
               11: lambda: g("temp"), 13: lambda: g("cp"),
               14: lambda: g("dmg"), 15: lambda: g("sdxx") + g("sm"),
               16: lambda: g("x") - self.xo[0]}.get(a)
        if arr is None:
            return None
        col = arr()
        return lambda p: fe(col[p], 12, 4)

    def write_paraview(self, time, ps, step_no):
        i1, i2, i3, i4 = INDENT, INDENT * 2, INDENT * 3, INDENT * 4
        n = self.problem.n
        fname = "%s_anim_%d.vtu" % (self.name, step_no)
        w = []
        w.append('<?xml version="1.0"?>\n<VTKFile type="UnstructuredGrid" version="0.1" byte_order="BigEndian">\n'
                 "<UnstructuredGrid>\n")
        w.append('%s<Piece NumberOfPoints="%d" NumberOfCells="%d">\n' % (i1, n, n))
        w.append(i2 + "<Points>\n")
        w.append(" " + i3 + '<DataArray type="Float32" NumberOfComponents="3" Format="ascii">\n')
        x, y, z = ps[PF["x"]], ps[PF["y"]], ps[PF["z"]]
        w.extend("%s%s  %s  %s\n" % (i4, fe(x[p], 12, 4), fe(y[p], 12, 4), fe(z[p], 12, 4)) for p in range(n))
        w.append(i3 + "</DataArray>\n" + i2 + "</Points>\n")
        w.append(i2 + '<PointData Scalars="scalars">\n')
        g = lambda k: ps[PF[k]]
        fields = [("seqv", g("seqv")), ("epeff", g("epeff")), ("pressure", -g("sm")), ("Smoothpressure", g("spre")),
                  ("MeanNormalStress", g("sm")), ("SmoothMeanNormalStress", -g("spre")), ("vx", g("vx")),
                  ("vy", g("vy")), ("vz", g("vz")), ("disx", x - self.xo[0]), ("disy", y - self.xo[1]),
                  ("disz", z - self.xo[2]), ("fxpx", g("fx")), ("fxpy", g("fy")), ("fxpz", g("fz")),
                  ("vol", g("vol")), ("celsius_t", g("temp")), ("damage", g("dmg")),
                  ("stressx", g("sdxx") + g("sm")), ("disx", x - self.xo[0])]
        for label, col in fields:
            w.append('%s<DataArray type="Float32" Name="%s" Format="ascii">\n' % (i3, label))
            w.extend(i4 + fe(v, 12, 4) + "\n" for v in col)
            w.append(i3 + "</DataArray>\n")
        w.append(i2 + "</PointData>\n" + i2 + "<Cells>\n")
        for label, kind, vals in (("connectivity", "Int32", range(n)), ("offsets", "Int32", range(1, n + 1)),
                                  ("types", "Float32", [1] * n)):
            w.append('%s<DataArray type="%s" Name="%s" Format="ascii">\n' % (i3, kind, label))
            w.extend(i4 + fi(v, 10) + "\n" for v in vals)
            w.append(i3 + "</DataArray>\n")
        w.append(i2 + "</Cells>\n" + i1 + "</Piece>\n</UnstructuredGrid>\n</VTKFile>\n")
        with open(self._path(fname), "w") as f:
            f.write("".join(w))
        self.files["pvd"].write('%s<DataSet timestep="%s" group="" part="0" file="%s"/>\n'
                                % (i1, fe(time, 12, 4), fname))

    def write_grid(self):
        i1, i2, i3, i4 = INDENT, INDENT * 2, INDENT * 3, INDENT * 4
        xg = grid_nodes(self.problem)
        cn = cells_node(self.problem)
        order = [0, 1, 3, 2, 4, 5, 7, 6]      # iNodeOrder (1,2,4,3,5,6,8,7)
        w = ['<?xml version="1.0"?>\n<VTKFile type="UnstructuredGrid" version="0.1" byte_order="BigEndian">\n'
             "<UnstructuredGrid>\n",
             '%s<Piece NumberOfPoints="%d" NumberOfCells="%d">\n' % (i1, xg.shape[1], cn.shape[0]),
             i2 + "<Points>\n", " " + i3 + '<DataArray type="Float32" NumberOfComponents="3" Format="ascii">\n']
        w.extend("%s%s  %s  %s\n" % (i4, fe(a, 12, 4), fe(b, 12, 4), fe(c, 12, 4)) for a, b, c in xg.T)
        w.append(i3 + "</DataArray>\n" + i2 + "</Points>\n" + i2 + "<Cells>\n")
        w.append('%s<DataArray type="Int32" Name="connectivity" Format="ascii">\n' % i3)
        w.extend(i4 + "".join("%d%s" % (row[k] - 1, INDENT) for k in order) + "\n" for row in cn)
        w.append(i3 + "</DataArray>\n")
        w.append('%s<DataArray type="Int32" Name="offsets" Format="ascii">\n' % i3)
        w.extend("%s%d\n" % (i4, 8 * (c + 1)) for c in range(cn.shape[0]))
        w.append(i3 + "</DataArray>\n")
        w.append('%s<DataArray type="Int32" Name="types" Format="ascii">\n' % i3)
        w.extend(i4 + fi(11, 5) + "\n" for _ in range(cn.shape[0]))
        w.append(i3 + "</DataArray>\n" + i2 + "</Cells>\n" + i1 + "</Piece>\n</UnstructuredGrid>\n</VTKFile>\n")
        with open(self._path(self.name + "_Grid.vtu"), "w") as f:
            f.write("".join(w))


def run(ctx, outputs, name, directory=".", max_steps=None):
    """Step ctx to the problem's end time on the reference driver's schedule
    (MPM3D.f90:34-113), writing the result files into directory.  ctx is a
    Context (or anything with its step/scalars/particles/icell methods);
    returns the number of steps taken."""
    prob = ctx.problem
    w = ResultWriter(prob, outputs, name, directory)
    subset = getattr(ctx, "particle_subset", None)
    pts = np.array(w.curve_points(), dtype=np.int32)
    steps = 0
    with w:
        plt = np.float32(np.float64(np.float32(0.0)) + outputs.out_time)   # plt = plt + OutTime
        w.frame(0.0, ctx.particles(), None)
        while max_steps is None or steps < max_steps:
            ctx.step(1)
            steps += 1
            sc = ctx.scalars()
            t, ek, ei = sc[1], sc[3], sc[4]
            sub = subset(pts) if subset is not None else ctx.particles()[:, pts - 1]
            cf = ctx.contact_force() if prob.contact == 1 else None
            w.curve(t, ek, ei, sc[5:8], sub, cf)
            if t >= np.float64(plt):
                plt = np.float32(np.float64(plt) + outputs.out_time)
                w.frame(t, ctx.particles(), ctx.icell())
            if prob.quasi_load:
                if t > prob.endtime and ek / (ek + ei) < 0.005 and t > 100:
                    break
            elif t > prob.endtime:
                break
    return steps
