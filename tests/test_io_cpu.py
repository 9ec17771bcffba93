"""Particle-state files: the reference's own data files parse, and this
build's writers round-trip."""
import os

import numpy as np
import pytest

from armando_amd import cases
from armando_amd import io as aio

REF = "/root/reference"
need_ref_files = pytest.mark.skipif(not os.path.isdir(REF), reason="reference tree not mounted")


@need_ref_files
def test_reads_the_reference_fortran_virtual_particle_files():
    """FORTRANsource/v_*.dat were written by output.f with dim = 3."""
    d = aio.read_fortran(os.path.join(REF, "FORTRANsource"), tag="v")
    assert d["dim"] == 3 and len(d["x"]) == 5636
    assert d["id"][0] == 42255 and np.all(np.diff(d["id"]) == 1)
    # first row of v_xv.dat, as printed
    assert d["x"][0] == -0.59878019e-02 and d["vz"][0] == 0.50205583e-01
    assert np.all(d["mat"] < 0) and np.all(d["h"] == 0.36000002e-03)
    assert np.all(d["rho"] == 0.13540000e05) and np.all(d["mass"] == 0.0)


@need_ref_files
def test_reads_the_reference_fortran_state_files():
    d = aio.read_fortran(os.path.join(REF, "FORTRANsource"), tag="f")
    # f_xv.dat is not shipped; state and other are
    assert "x" not in d and len(d["rho"]) == 42254 and len(d["mat"]) == 42254
    assert d["id"][0] == 1 and d["rho"][0] == 0.10300653e05 and d["mat"][0] == 21


@need_ref_files
def test_reads_the_reference_c_state_and_orders_walls_last():
    d = aio.read_c(os.path.join(REF, "ArmandoList"), prefix="new_")
    n = len(d["x"])
    assert n == 85773 and len(d["vx"]) == n and len(d["mat"]) == n
    assert np.all(d["mass"] == d["mass"][0]) and np.all(d["h"] == d["h"][0])
    s, nreal = aio.walls_last(d)
    assert np.all(s["mat"][:nreal] >= 0) and np.all(s["mat"][nreal:] < 0)
    assert nreal == np.count_nonzero(d["mat"] >= 0)
    # the first wall particle of the file keeps its position after reordering
    k = int(np.nonzero(d["mat"] < 0)[0][0])
    assert s["x"][nreal] == d["x"][k] and s["y"][nreal] == d["y"][k]


def state_of(c):
    rng = np.random.default_rng(3)
    s = {k: c[k].copy() for k in ("x", "y", "vx", "vy", "rho", "u")}
    s["p"] = rng.standard_normal(c["n"]) * 1e3
    return s


@pytest.mark.parametrize("precise", [False, True])
def test_fortran_round_trip(tmp_path, precise):
    c = cases.dam_break_small(scale=0.02)
    s = state_of(c)
    aio.write_fortran(tmp_path, s, c["np"], h=c["h"], mass=c["mass"], precise=precise)
    f = aio.read_fortran(tmp_path, "f")
    v = aio.read_fortran(tmp_path, "v")
    assert len(f["x"]) == c["np"] and len(v["x"]) == c["nv"]
    for k in ("x", "y", "vx", "vy", "rho", "p", "u"):
        got = np.concatenate([f[k], v[k]])
        if precise:
            assert np.array_equal(got, s[k]), k
        else:
            assert np.allclose(got, s[k], rtol=1e-8, atol=1e-300), k
    assert np.all(f["mat"] == 1) and np.all(v["mat"] == -1)
    assert v["id"][0] == c["np"] + 1


def test_c_round_trip_and_case(tmp_path):
    c = cases.dam_break_small(scale=0.02)
    s = state_of(c)
    aio.write_c(tmp_path, s, c["np"], c["mass"], c["h"], prefix="in_", precise=True)
    d = aio.read_c(tmp_path, prefix="in_")
    for k in ("x", "y", "vx", "vy", "rho", "p", "u"):
        assert np.array_equal(d[k], s[k]), k
    st, nreal = aio.walls_last(d)
    assert nreal == c["np"]
    c2 = aio.case_from_state(c, st, nreal)
    assert c2["n"] == c["n"] and c2["nv"] == c["nv"] and np.array_equal(c2["x"], c["x"])


def test_ensight_and_vtk_layout(tmp_path):
    c = cases.dam_break_small(scale=0.02)
    s = state_of(c)
    aio.write_ensight_case(tmp_path, [0.0, 1e-3])
    aio.write_ensight_step(tmp_path, 0, s)
    with open(tmp_path / "armando_pos_00000.geo") as fh:
        lines = fh.read().splitlines()
    assert lines[1].startswith("EnSight") and int(lines[12]) == c["n"]
    assert len(lines) == 13 + 3 * c["n"]
    assert float(lines[13]) == pytest.approx(s["x"][0], rel=1e-6)
    with open(tmp_path / "armando_rho_00000.dat") as fh:
        assert len(fh.read().splitlines()) == 4 + c["n"]
    aio.write_vtk(tmp_path / "s.vtk", s)
    with open(tmp_path / "s.vtk") as fh:
        txt = fh.read()
    assert f"POINTS {c['n']} double" in txt and txt.count("LOOKUP_TABLE") == 3
