"""Seeded synthetic .mpm decks for the benchmark workloads.

Each builder returns deck text in the reference's input format (read back
through mpm_deck_read, so every real goes through the same FP32 rounding as a
hand-written deck).  Units and geometry:

* ``taylor_bar``: copper Taylor bar (mm-g-N-ms-MPa): radius 3.2, length 32.4,
  0.2 mm particle lattice (2x2x2 per 0.4 mm cell), J2 linear hardening
  (rho 8.93e-3, E 117e3, nu 0.35, yield 400, H 100), GIMP, MUSL, impact
  onto the fixed plane z = 0 at 190 m/s (= mm/ms), CFL 0.1.
* ``slope``: Drucker-Prager soil slope (SI): 40 m long profile, 45-degree
  face rising from 5 m to 20 m, y-thickness 0.3 m (plane strain between
  roller faces), cell 0.1 m, 2x2x2 particles per cell, quadratic B-spline,
  USF, gravity applied as a body load ramped over 0.5 s (QuLo).
* ``slope_ensemble``: seeded variants of ``slope`` (size, angle, friction,
  cohesion) for the batched ensemble.
* ``cook_membrane``: tapered neo-Hookean plate (nu = 0.4999) with a shear
  traction ramp, SGMP; ``dam_break``: weakly compressible water column, SGMP.
"""
import math
import os
import tempfile

import numpy as np


def dp_params(c, phi_deg, psi_deg=0.0):
    """Inner-cone Drucker-Prager constants (q_phi, k_phi, q_psi, ten_f) from
    cohesion and friction/dilation angles."""
    phi = math.radians(phi_deg)
    psi = math.radians(psi_deg)
    s3 = math.sqrt(3.0)
    qf = 6 * math.sin(phi) / (s3 * (3 + math.sin(phi)))
    kf = 6 * c * math.cos(phi) / (s3 * (3 + math.sin(phi)))
    qp = 6 * math.sin(psi) / (s3 * (3 + math.sin(psi)))
    tf = c / math.tan(phi) if phi > 0 else 0.0
    return qf, kf, qp, tf


def taylor_bar(radius=3.2, length=32.4, dp=0.2, speed=190.0, dtscale=0.1, endtime=0.08, h=None):
    """Taylor bar as stacked 'cicl' discs (one body per disc)."""
    h = h or 2 * dp
    nr = int(round(radius / dp))
    nz = int(round(length / dp))
    # particles per disc (the cicl generator's test, DataIn.f90:814-842)
    k = np.arange(-nr, nr) + 0.5
    per = int(np.sum((k[:, None] ** 2 + k[None, :] ** 2) < nr * nr))
    n = per * nz
    half = max(2.0 * radius, radius + 4 * h)
    half = math.ceil(half / h) * h
    z0 = h   # bar starts one cell above the fixed plane z = 0
    ztop = math.ceil((z0 + length + 2 * h) / h) * h
    L = ["mpm3d *** Taylor bar R=%g L=%g dp=%g v=%g" % (radius, length, dp, speed),
         "nbco 1", "nbbo %d" % nz, "nbmp %d" % n, "nmat 1",
         "spx %.6f %.6f" % (-half, half), "spy %.6f %.6f" % (-half, half),
         "spz 0.0 %.6f" % ztop, "dcell %g" % h,
         "musl on", "gimp", "dtscale %g" % dtscale, "endtime %g" % endtime,
         "fixed 0 0 0 0 2 0", "material", " 1 pla2 8.93e-3 117.0d3 0.35 400 100"]
    for iz in range(nz):
        L.append("particle cicl 1 1 8.93e-3 %g 0.0 0.0 %.6f %d" % (dp, z0 + iz * dp, nr))
    L.append("velo")
    for b in range(1, nz + 1):
        L.append("body %d 0.0 0.0 %.6f" % (b, -speed))
    L += ["endv", "endi"]
    return "\n".join(L) + "\n", n


def slope(length=40.0, height=20.0, base=5.0, angle_deg=45.0, thickness=0.3, h=0.1, ppc=2,
          rho=2000.0, E=70e6, nu=0.3, c=10e3, phi=30.0, psi=0.0, ramp=0.5, dtscale=0.2,
          endtime=0.5, toe=10.0):
    """Slope profile built from one 'bloc' particle column per x position."""
    dp = h / ppc
    qf, kf, qp, tf = dp_params(c, phi, psi)
    run = (height - base) / math.tan(math.radians(angle_deg))
    nx = int(round(length / dp))
    ny = int(round(thickness / dp))
    cols = []
    for i in range(nx):
        x = (i + 0.5) * dp
        if x <= toe:
            z = base
        elif x >= toe + run:
            z = height
        else:
            z = base + (x - toe) * (height - base) / run
        nz = int(round(z / dp))
        cols.append((i, nz))
    n = sum(nz * ny for _, nz in cols)
    pm = rho * dp ** 3
    g = 9.8
    L = ["mpm3d *** slope %gx%gx%g m, %g deg, h=%g" % (length, thickness, height, angle_deg, h),
         "nbco 1", "nbbo %d" % len(cols), "nbmp %d" % n, "nmat 1",
         "spx %.6f %.6f" % (-h, length + h), "spy %.6f %.6f" % (-h, thickness + h),
         "spz %.6f %.6f" % (-h, height + 5 * h), "dcell %g" % h,
         "usf on", "bimp", "dtscale %g" % dtscale, "endtime %g" % endtime, "QuLo %g" % ramp,
         "fixed 2 2 2 2 1 0", "material",
         " 1 dpm %g %g %g %.6f %.4f %.6f %.2f" % (rho, E, nu, qf, kf, qp, tf)]
    for i, nz in cols:
        L.append("particle bloc 1 1 %g %g %.6f 0.0 0.0 1 %d %d" % (rho, dp, i * dp, ny, nz))
    L.append("load")
    for b in range(1, len(cols) + 1):
        L.append("body %d 0.0 0.0 %.6f" % (b, -pm * g))
    L += ["endl", "endi"]
    return "\n".join(L) + "\n", n


def slope_ensemble(count, seed=0, n_min=2e4, n_max=4e5, **kw):
    """Seeded slope instances: particle counts log-uniform in [n_min, n_max]
    (through the y-thickness), angle U(30,60), friction U(20,40), cohesion
    lognormal (median 10 kPa, sigma 0.5)."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(count):
        target = math.exp(rng.uniform(math.log(n_min), math.log(n_max)))
        ang = rng.uniform(30.0, 60.0)
        phi = rng.uniform(20.0, 40.0)
        coh = 10e3 * math.exp(0.5 * rng.standard_normal())
        # size the profile so that area * thickness * 8000 ~ target at h = 0.1
        scale = 1.0
        base_area = 537.5
        t = max(0.1, min(0.5, target / (base_area * 8000)))
        scale = math.sqrt(max(target / (t * 8000 * base_area), 0.05))
        L, H, B = 40.0 * scale, 20.0 * scale, 5.0 * scale
        L, H, B, t = round(L, 1), round(H, 1), round(B, 1), round(t, 1)
        text, n = slope(length=L, height=H, base=B, angle_deg=ang, thickness=t, c=coh, phi=phi,
                        toe=round(10.0 * scale, 1), **kw)
        h = kw.get("h", 0.1)
        # grid nodes as SetGridData counts them (Grid.f90:250-274), spans as in slope()
        cells = [int((L + 2 * h) / h + 0.5), int((t + 2 * h) / h + 0.5), int((H + 6 * h) / h + 0.5)]
        nodes = (cells[0] + 1) * (cells[1] + 1) * (cells[2] + 1)
        out.append((text, n, {"angle": ang, "phi": phi, "c": coh, "nodes": nodes,
                              "bodies": text.count("\nparticle ")}))
    return out


def cook_membrane(dp=0.25, h=0.5, thickness=10.0, rho=1.0e-3, E=240.0, nu=0.4999, shear=100.0, ramp=1.0,
                  damp=0.3, dtscale=0.9, endtime=5.0):
    """Cook-membrane-like tapered plate (mm-g-N-ms-MPa): vertices (0,0),
    (48,44), (48,60), (0,44) in x-y, `thickness` in z; nearly incompressible
    neo-Hookean; left face clamped with one extra grid layer; shear traction
    `shear` (N, total) on the right-edge particles in +y, ramped over `ramp`
    (QuLo); damping; SGMP, MUSL.  One 'bloc' column per x position."""
    nx = int(round(48.0 / dp))
    nz = int(round(thickness / dp))
    cols = []
    for i in range(nx):
        x = (i + 0.5) * dp
        ylo, yhi = 44.0 * x / 48.0, 44.0 + 16.0 * x / 48.0
        jlo = int(math.ceil(ylo / dp - 0.5))
        jhi = int(math.floor(yhi / dp - 0.5))
        cols.append((i, jlo, jhi - jlo + 1))
    n = sum(ny * nz for _, _, ny in cols)
    L = ["mpm3d *** Cook membrane plate 48x44(->16)x%g mm, h=%g, SGMP" % (thickness, h),
         "nbco 1", "nbbo %d" % len(cols), "nbmp %d" % n, "nmat 1",
         "spx %.6f %.6f" % (-h, 48.0 + 2 * h), "spy %.6f %.6f" % (-2 * h, 60.0 + 2 * h),
         "spz %.6f %.6f" % (-2 * h, thickness + 2 * h), "dcell %g" % h,
         "musl on", "sgmp", "dtscale %g" % dtscale, "endtime %g" % endtime, "QuLo %g" % ramp, "Damp %g" % damp,
         "fixed 1 0 0 0 0 0", "material", " 1 neoh %g %g %g" % (rho, E, nu)]
    for i, jlo, ny in cols:
        L.append("particle bloc 1 1 %g %g %.6f %.6f 0.0 1 %d %d" % (rho, dp, i * dp, jlo * dp, ny, nz))
    # right-edge particles: the last column's block, x fastest then y then z
    last = cols[-1][2] * nz
    first = n - last + 1
    f = shear / last
    L.append("load")
    for pid in range(first, n + 1):
        L.append("part %d 0.0 %.9e 0.0" % (pid, f))
    L += ["endl", "endi"]
    return "\n".join(L) + "\n", n


def dam_break(h=0.005, width=0.6, depth=0.3, height=0.3, tank=1.6, c=50.0, rho=1000.0, dtscale=0.5, endtime=1.0,
              sgmp=True):
    """3D dam break (SI): water column width x depth x height at the x = 0
    wall of a tank `tank` long, weakly compressible (null material, EOS 4,
    sound speed c, K = rho c^2), gravity; walls and floor symmetric (fixed 2)
    with one extra grid layer; 2x2x2 particles per cell; SGMP, MUSL."""
    dp = h / 2
    nx, ny, nz = int(round(width / dp)), int(round(depth / dp)), int(round(height / dp))
    n = nx * ny * nz
    L = ["mpm3d *** dam break %gx%gx%g m, h=%g" % (width, depth, height, h),
         "nbco 1", "nbbo 1", "nbmp %d" % n, "nmat 1",
         "spx %.6f %.6f" % (-h, tank + h), "spy %.6f %.6f" % (-h, depth + h), "spz %.6f %.6f" % (-h, 2 * height),
         "dcell %g" % h, "musl on"] + (["sgmp"] if sgmp else []) + [
         "dtscale %g" % dtscale, "endtime %g" % endtime, "fixed 2 2 2 2 2 0", "material",
         " 1 null %g %g 1.0e-3" % (rho, rho * c * c), "seos 1 4 %g" % c,
         "particle bloc 1 1 %g %g 0.0 0.0 0.0 %d %d %d" % (rho, dp, nx, ny, nz),
         "load", "grav 1 0.0 0.0 -9.81", "endl", "endi"]
    return "\n".join(L) + "\n", n


def write_deck(text, directory=None, name="synthetic"):
    d = directory or tempfile.mkdtemp(prefix="mpmdeck_")
    path = os.path.join(d, name + ".mpm")
    with open(path, "w") as f:
        f.write(text)
    return path


def load(text, mpm, name="synthetic"):
    """Write the deck to a temporary file and read it with mpm_deck_read."""
    d = tempfile.mkdtemp(prefix="mpmdeck_")
    path = write_deck(text, d, name)
    try:
        return mpm.read_deck(path)
    finally:
        try:
            os.remove(path)
            os.rmdir(d)
        except OSError:
            pass
