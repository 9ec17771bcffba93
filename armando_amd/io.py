"""Particle-state files in the reference's formats, and visualisation output.

Host-side utilities around the device state (`Simulation.state()` /
`cases.*` dictionaries); nothing here touches the GPU.

* Fortran restart files (CERNsource/output.f:1-74, input.f:197-314): real
  particles in f_xv.dat, f_state.dat, f_other.dat, virtual ones in v_xv.dat,
  v_state.dat, v_other.dat.  Format 1001 `i, x(dim), v(dim)`, 1002 `i, mass,
  rho, p, u`, 1003 `i, mat, h` (output.f:70-72); the first line of *_xv.dat
  is the particle count.  dim is 2 for this build's step; files written with
  dim = 3 (the reference as shipped) are read with z and vz dropped.
* C path state files (ArmandoList/armando2Dcpu.cu:1530-1572, read at
  :1480-1525): <prefix>pos.txt `x y`, <prefix>vel.txt `vx vy`,
  <prefix>info.txt `material mass h`, <prefix>field.txt `rho p u`; walls
  have material < 0.
* EnSight Gold ASCII time series (armando2Dcpu.cu:1575-1720 writes the same
  layout): armando.case plus per-step geometry and velocity, density,
  pressure and energy files.
* VTK legacy polydata, one file per step.

* FLUKA USRBIN energy-deposition maps (CERNsource/fluka_heat.f:1-99 binary,
  :100-199 text `fluka_binning.FORM`), Cartesian binning: the map the
  deposition heat source interpolates (`Simulation.set_deposition`).

The reference formats print 9 significant digits; `precise=True` writes 17
(round-trip exact doubles) in the same column layout.
"""
import os

import numpy as np

# ------------------------------------------------------------------ Fortran


def _ffmt(precise):
    return "%.16E" if precise else "%14.8E"


def write_fortran(dirpath, state, np_real, mat=None, h=None, mass=None, precise=False,
                  dim=2):
    """Writes f_*.dat (particles 0..np_real-1) and v_*.dat (the rest).

    state: dict with x, y, vx, vy, rho, p, u (p may be absent: zeros);
    mat: per-particle material (default 1 for real, -1 for virtual particles);
    h, mass: scalars or arrays.  Ids are 1-based as in the reference."""
    os.makedirs(dirpath, exist_ok=True)
    n = len(state["x"])
    zeros = np.zeros(n)
    p = np.asarray(state.get("p", zeros), float)
    if mat is None:
        mat = np.where(np.arange(n) < np_real, 1, -1)
    h = np.broadcast_to(np.asarray(h if h is not None else 0.0, float), (n,))
    mass = np.broadcast_to(np.asarray(mass if mass is not None else 0.0, float), (n,))
    f = _ffmt(precise)
    cols = [state["x"], state["y"]] + ([zeros] if dim == 3 else []) + \
        [state["vx"], state["vy"]] + ([zeros] if dim == 3 else [])
    for tag, lo, hi in (("f", 0, np_real), ("v", np_real, n)):
        with open(os.path.join(dirpath, f"{tag}_xv.dat"), "w") as fx, \
                open(os.path.join(dirpath, f"{tag}_state.dat"), "w") as fs, \
                open(os.path.join(dirpath, f"{tag}_other.dat"), "w") as fo:
            fx.write(f" {hi - lo:10d}\n")
            for i in range(lo, hi):
                fx.write(f" {i + 1:6d}" + "".join("  " + f % c[i] for c in cols) + "\n")
                fs.write(f" {i + 1:6d}" + "".join(
                    "  " + f % v for v in (mass[i], state["rho"][i], p[i], state["u"][i])) + "\n")
                fo.write(f" {i + 1:6d}  {int(mat[i]):4d}  " + f % h[i] + "\n")


def _read_rows(path, ncols=None):
    rows = []
    with open(path) as fh:
        for line in fh:
            tok = line.split()
            if tok:
                rows.append([float(t.replace("D", "E").replace("d", "e")) for t in tok])
    return rows


def read_fortran(dirpath, tag="f"):
    """Reads <tag>_xv.dat, <tag>_state.dat, <tag>_other.dat (any that exist).

    Returns a dict with id (1-based), x, y, vx, vy (dim 2 or 3 detected from
    the column count; z and vz are returned as z, vz when present), mass,
    rho, p, u, mat, h."""
    out = {}
    path = os.path.join(dirpath, f"{tag}_xv.dat")
    if os.path.exists(path):
        rows = _read_rows(path)
        count = int(rows[0][0])
        body = np.array(rows[1:1 + count])
        if body.shape[0] != count:
            raise ValueError(f"{path}: {count} particles announced, {body.shape[0]} rows")
        dim = (body.shape[1] - 1) // 2
        if dim not in (2, 3) or body.shape[1] != 1 + 2 * dim:
            raise ValueError(f"{path}: {body.shape[1]} columns")
        out["id"] = body[:, 0].astype(np.int64)
        out["x"], out["y"] = body[:, 1], body[:, 2]
        out["vx"], out["vy"] = body[:, 1 + dim], body[:, 2 + dim]
        if dim == 3:
            out["z"], out["vz"] = body[:, 3], body[:, 6]
        out["dim"] = dim
    path = os.path.join(dirpath, f"{tag}_state.dat")
    if os.path.exists(path):
        body = np.array(_read_rows(path))
        out.setdefault("id", body[:, 0].astype(np.int64))
        out["mass"], out["rho"], out["p"], out["u"] = body[:, 1], body[:, 2], body[:, 3], body[:, 4]
    path = os.path.join(dirpath, f"{tag}_other.dat")
    if os.path.exists(path):
        body = np.array(_read_rows(path))
        out.setdefault("id", body[:, 0].astype(np.int64))
        out["mat"] = body[:, 1].astype(np.int64)
        out["h"] = body[:, 2]
    return out


# ------------------------------------------------------------------ C path


def write_c(dirpath, state, np_real, mass, h, prefix="new_", precise=False):
    """pos / vel / info / field text files of the C path; walls (index >=
    np_real) get material -1, fluid 1."""
    os.makedirs(dirpath, exist_ok=True)
    f = "%+.16e" if precise else "%+14.8e"
    n = len(state["x"])
    p = state.get("p", np.zeros(n))
    with open(os.path.join(dirpath, prefix + "pos.txt"), "w") as fh:
        for i in range(n):
            fh.write(f"{f % state['x'][i]} {f % state['y'][i]}\n")
    with open(os.path.join(dirpath, prefix + "vel.txt"), "w") as fh:
        for i in range(n):
            fh.write(f"{f % state['vx'][i]} {f % state['vy'][i]} \n")
    with open(os.path.join(dirpath, prefix + "info.txt"), "w") as fh:
        for i in range(n):
            fh.write(f"{1 if i < np_real else -1} {f % mass} {f % h} \n")
    with open(os.path.join(dirpath, prefix + "field.txt"), "w") as fh:
        for i in range(n):
            fh.write(f"{f % state['rho'][i]} {f % p[i]} {f % state['u'][i]} \n")


def read_c(dirpath, prefix="in_"):
    """Reads <prefix>pos/vel/info/field.txt (field optional)."""
    def load(name):
        path = os.path.join(dirpath, prefix + name)
        return np.array(_read_rows(path)) if os.path.exists(path) else None

    out = {}
    pos = load("pos.txt")
    out["x"], out["y"] = pos[:, 0], pos[:, 1]
    vel = load("vel.txt")
    if vel is not None:
        out["vx"], out["vy"] = vel[:, 0], vel[:, 1]
    info = load("info.txt")
    if info is not None:
        out["mat"] = info[:, 0].astype(np.int64)
        out["mass"], out["h"] = info[:, 1], info[:, 2]
    field = load("field.txt")
    if field is not None:
        out["rho"], out["p"], out["u"] = field[:, 0], field[:, 1], field[:, 2]
    return out


def walls_last(state):
    """Reorders a C-path state so that fluid (material >= 0) comes first and
    walls (material < 0) last, as the boundary expects; returns (state, np)."""
    mat = state["mat"]
    order = np.concatenate([np.nonzero(mat >= 0)[0], np.nonzero(mat < 0)[0]])
    out = {k: (v[order] if isinstance(v, np.ndarray) and v.shape[:1] == mat.shape else v)
           for k, v in state.items()}
    return out, int(np.count_nonzero(mat >= 0))


# ------------------------------------------------------------------ EnSight / VTK


def write_ensight_case(dirpath, times, name="armando"):
    os.makedirs(dirpath, exist_ok=True)
    with open(os.path.join(dirpath, f"{name}.case"), "w") as fh:
        fh.write(f"# Ensight formatted case file for {name}\n\nFORMAT\ntype: ensight gold\n\n")
        fh.write(f"GEOMETRY\nmodel:    1           {name}_pos_*****.geo\n\nVARIABLE\n")
        fh.write(f"vector per node:    1 velocity {name}_vel_*****.dat\n")
        fh.write(f"scalar per node:    1 density  {name}_rho_*****.dat\n")
        fh.write(f"scalar per node:    1 pressure {name}_pre_*****.dat\n")
        fh.write(f"scalar per node:    1 energy   {name}_ene_*****.dat\n\n")
        fh.write(f"TIME\ntime set: 1\nnumber of steps: {len(times)}\n")
        fh.write("filename start number: 0\nfilename increment: 1\ntime values:\n")
        for t in times:
            fh.write(f"{t:14.8e}\n")


def write_ensight_step(dirpath, index, state, name="armando"):
    """One time step of the EnSight Gold ASCII series (particles as nodes of
    one part, z = 0)."""
    os.makedirs(dirpath, exist_ok=True)
    n = len(state["x"])

    def column(fh, values):
        fh.write("".join(f"{v:+e}\n" for v in values))

    with open(os.path.join(dirpath, f"{name}_pos_{index:05d}.geo"), "w") as fh:
        fh.write("Armando output in EnSight Gold format\nEnSight 8.0.7\n")
        fh.write("node id assign\nelement id assign\nextents\n")
        fh.write(" 1.00000e+38-1.00000e+38\n" * 3)
        fh.write(f"part\n{1:10d}\nSPH particles\ncoordinates\n{n:10d}\n")
        column(fh, state["x"])
        column(fh, state["y"])
        column(fh, np.zeros(n))
    with open(os.path.join(dirpath, f"{name}_vel_{index:05d}.dat"), "w") as fh:
        fh.write(f"particle velocity in EnSight Gold format\npart\n{1:10d}\ncoordinates\n")
        column(fh, state["vx"])
        column(fh, state["vy"])
        column(fh, np.zeros(n))
    for tag, key, label in (("rho", "rho", "density"), ("pre", "p", "pressure"),
                            ("ene", "u", "energy")):
        with open(os.path.join(dirpath, f"{name}_{tag}_{index:05d}.dat"), "w") as fh:
            fh.write(f"particle {label} in EnSight Gold format\npart\n{1:10d}\ncoordinates\n")
            column(fh, state[key])


def write_vtk(path, state, fields=("rho", "p", "u")):
    """VTK legacy ASCII polydata: particles as vertices, velocity as a vector
    and the named scalar fields."""
    n = len(state["x"])
    with open(path, "w") as fh:
        fh.write("# vtk DataFile Version 3.0\nSPH particles\nASCII\nDATASET POLYDATA\n")
        fh.write(f"POINTS {n} double\n")
        fh.write("".join(f"{x:.9e} {y:.9e} 0\n" for x, y in zip(state["x"], state["y"])))
        fh.write(f"VERTICES {n} {2 * n}\n")
        fh.write("".join(f"1 {i}\n" for i in range(n)))
        fh.write(f"POINT_DATA {n}\nVECTORS velocity double\n")
        fh.write("".join(f"{a:.9e} {b:.9e} 0\n" for a, b in zip(state["vx"], state["vy"])))
        for key in fields:
            if key in state:
                fh.write(f"SCALARS {key} double 1\nLOOKUP_TABLE default\n")
                fh.write("".join(f"{v:.9e}\n" for v in state[key]))


# ------------------------------------------------------------------ cases


def case_from_state(template, state, np_real):
    """A case dictionary (as armando_amd.cases builds) with the template's
    parameters and the particle arrays of `state` (x, y, vx, vy, rho, u)."""
    c = dict(template)
    for k in ("x", "y", "vx", "vy", "rho", "u"):
        c[k] = np.ascontiguousarray(state[k], np.float64)
    c["n"] = len(c["x"])
    c["np"] = int(np_real)
    c["nv"] = c["n"] - c["np"]
    return c


# ------------------------------------------------------------------ FLUKA
class Binning(dict):
    """A deposition map: origin (3), cell (3), n (3), type (FLUKA's tfluka:
    0 Cartesian x, y, z; 1 or 7 R-phi-z, r, phi in degrees, z), data (n[0]
    n[1] n[2] values in W/m^3, first axis fastest)."""


def _scale_binning(typ, minb, dbin, nbin, vals, translation, size_scale, energy_scale,
                   n_primaries, dep_time):
    # fluka_heat.f:76-95 / :176-195: sizes scaled by sf, then translated;
    # values (GeV/cm^3 per primary in FLUKA's units) times
    # np ef / (dt sf^3).  minb, dbin and vals are REAL(4) in the reference.
    # R-phi-z (tb 1 or 7 with more than one phi bin, :56-75 / :157-176): r and
    # z scaled by sf, phi (degrees) not; all three translated; the values
    # scaled from the ones read (the reference's branch scales the array it
    # was given instead of the values it read).
    tb = typ % 10
    if tb not in (0, 1, 7) or (tb in (1, 7) and nbin[1] < 2):
        raise ValueError(f"binning type {typ}: Cartesian (0) and R-phi-z (1, 7) binnings only")
    sf = float(size_scale)
    sc = [sf, 1.0 if tb else sf, sf]
    origin = [s * float(np.float32(m)) + float(t) for s, m, t in zip(sc, minb, translation)]
    cell = [s * float(np.float32(d)) for s, d in zip(sc, dbin)]
    factor = (float(n_primaries) * float(energy_scale)) / (float(dep_time) * sf ** 3)
    data = factor * np.asarray(vals, dtype=np.float32).astype(np.float64)
    return Binning(origin=origin, cell=cell, n=[int(k) for k in nbin], type=int(tb), data=data)


def read_fluka_text(path, translation=(0.0, 0.0, 0.0), size_scale=1.0, energy_scale=1.0,
                    n_primaries=1.0, dep_time=1.0):
    """read_binning_txt (fluka_heat.f:100-199): five header lines, the
    binning type, three `min max n width` lines, one skipped line, then the
    n0 n1 n2 values, free format."""
    with open(path) as f:
        lines = f.read().splitlines()
    typ = int(lines[5].replace(",", " ").split()[0])
    minb, nbin, dbin = [], [], []
    for k in range(3):
        t = lines[6 + k].replace(",", " ").split()
        minb.append(np.float32(t[0]))
        nbin.append(int(t[2]))
        dbin.append(np.float32(t[3]))
    count = nbin[0] * nbin[1] * nbin[2]
    vals = []
    for line in lines[10:]:
        vals.extend(line.replace(",", " ").split())
        if len(vals) >= count:
            break
    if len(vals) < count:
        raise ValueError(f"{path}: {len(vals)} values, {count} expected")
    vals = np.array([np.float32(v) for v in vals[:count]], dtype=np.float32)
    return _scale_binning(typ, minb, dbin, nbin, vals, translation, size_scale, energy_scale,
                          n_primaries, dep_time)


def _records(path):
    """Records of a Fortran sequential unformatted file (4-byte markers)."""
    raw = open(path, "rb").read()
    out, pos = [], 0
    while pos + 4 <= len(raw):
        n = int(np.frombuffer(raw, "<i4", 1, pos)[0])
        body = raw[pos + 4:pos + 4 + n]
        tail = int(np.frombuffer(raw, "<i4", 1, pos + 4 + n)[0])
        if tail != n:
            raise ValueError(f"{path}: bad record marker at byte {pos}")
        out.append(body)
        pos += 8 + n
    return out


def read_fluka_binary(path, translation=(0.0, 0.0, 0.0), size_scale=1.0, energy_scale=1.0,
                      n_primaries=1.0, dep_time=1.0):
    """read_binning_bin (fluka_heat.f:1-99): the USRBIN unformatted file,
    records (title, header, weight, ncase), (mbin, name, type, dist, then
    min max n width for each axis, lntzer, b1, b2, tc), (values)."""
    rec = _records(path)
    h = rec[1]
    typ = int(np.frombuffer(h, "<i4", 1, 14)[0])
    minb, nbin, dbin = [], [], []
    for k in range(3):
        o = 22 + 16 * k
        minb.append(np.frombuffer(h, "<f4", 1, o)[0])
        nbin.append(int(np.frombuffer(h, "<i4", 1, o + 8)[0]))
        dbin.append(np.frombuffer(h, "<f4", 1, o + 12)[0])
    count = nbin[0] * nbin[1] * nbin[2]
    vals = np.frombuffer(rec[2], "<f4", count)
    return _scale_binning(typ, minb, dbin, nbin, vals, translation, size_scale, energy_scale,
                          n_primaries, dep_time)
