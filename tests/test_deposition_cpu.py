"""The energy-deposition heat source of CERNsource/fluka_heat.f (Cartesian
binning) in the oracle, and the binning readers of armando_amd.io, against
the reference's own fluka_heat, read_binning_txt and read_binning_bin
compiled unmodified (oracle/_ref), and against what linear interpolation
between bin centres fixes: a constant map, a map linear in the bin centres,
the bin values at the centres and zero outside the binning.
"""
import os

import numpy as np
import pytest

from armando_amd import io
from oracle import orc

from . import refbind

ORIGIN = (-0.021, 0.013, -0.03)
CELL = (0.013, 0.011, 0.02)
NBIN = (7, 5, 4)  # z range [-0.03, 0.05] holds the slice z = 0


def positions(rng, n):
    """Points across the binning, its edges and outside it."""
    lo = np.array(ORIGIN[:2]) - 0.02
    hi = np.array(ORIGIN[:2]) + np.array(NBIN[:2]) * np.array(CELL[:2]) + 0.02
    x = rng.uniform(lo[0], hi[0], n)
    y = rng.uniform(lo[1], hi[1], n)
    # exact edges, corners and bin centres
    ex = [ORIGIN[0], ORIGIN[0] + NBIN[0] * CELL[0], ORIGIN[0] + 2.5 * CELL[0]]
    ey = [ORIGIN[1], ORIGIN[1] + NBIN[1] * CELL[1], ORIGIN[1] + 1.5 * CELL[1]]
    gx, gy = np.meshgrid(ex, ey)
    return np.concatenate([x, gx.ravel()]), np.concatenate([y, gy.ravel()])


def oracle_params(data, t_end=1.0):
    P = orc.params(dt=1e-3)
    return orc.set_deposition(P, ORIGIN, CELL, NBIN, data, t_end)


def centres(k):
    return ORIGIN[k] + (np.arange(NBIN[k]) + 0.5) * CELL[k]


@pytest.mark.skipif(not refbind.available(), reason="oracle/_ref not built")
def test_oracle_equals_reference_fluka_heat():
    rng = np.random.default_rng(11)
    x, y = positions(rng, 4000)
    rho = rng.uniform(900.0, 1100.0, len(x))
    data = rng.uniform(-1e6, 5e8, int(np.prod(NBIN)))
    want = refbind.fluka_heat(x, y, rho, ORIGIN, CELL, NBIN, data)
    got = orc.fluka_heat(oracle_params(data), x, y, rho)
    assert np.count_nonzero(want) > 1000
    np.testing.assert_array_equal(got, want)


@pytest.mark.skipif(not refbind.available(), reason="oracle/_ref not built")
def test_oracle_3d_equals_reference_fluka_heat():
    """The dim = 3 branch with real z, points above, inside and below the
    binning's z range."""
    rng = np.random.default_rng(17)
    x, y = positions(rng, 4000)
    z = rng.uniform(ORIGIN[2] - 0.02, ORIGIN[2] + NBIN[2] * CELL[2] + 0.02, len(x))
    z[:12] = [ORIGIN[2], ORIGIN[2] + NBIN[2] * CELL[2]] * 6
    rho = rng.uniform(900.0, 1100.0, len(x))
    data = rng.uniform(-1e6, 5e8, int(np.prod(NBIN)))
    want = refbind.fluka_heat(x, y, rho, ORIGIN, CELL, NBIN, data, z=z)
    got = orc.fluka_heat3(oracle_params(data), x, y, z, rho)
    assert np.count_nonzero(want) > 1000
    np.testing.assert_array_equal(got, want)
    # at z = 0 the 3D routine is the 2D one
    np.testing.assert_array_equal(orc.fluka_heat3(oracle_params(data), x, y, 0 * z, rho),
                                  orc.fluka_heat(oracle_params(data), x, y, rho))


def test_constant_map_and_outside():
    rng = np.random.default_rng(12)
    x, y = positions(rng, 2000)
    rho = rng.uniform(900.0, 1100.0, len(x))
    data = np.full(int(np.prod(NBIN)), 3.0e8)
    got = orc.fluka_heat(oracle_params(data), x, y, rho)
    inside = ((x >= ORIGIN[0]) & (x <= ORIGIN[0] + NBIN[0] * CELL[0]) &
              (y >= ORIGIN[1]) & (y <= ORIGIN[1] + NBIN[1] * CELL[1]))
    assert inside.sum() > 500 and (~inside).sum() > 200
    np.testing.assert_allclose(got[inside], 3.0e8 / rho[inside], rtol=1e-15)
    assert np.all(got[~inside] == 0.0)


def test_linear_map_is_reproduced_between_centres():
    """A map linear in the bin centres is interpolated exactly wherever the
    stencil stays between the first and last centres; the z slice enters
    through the same weights."""
    a, b, c, d = 2.0e8, 3.0e9, -1.5e9, 4.0e9
    xc, yc, zc = centres(0), centres(1), centres(2)
    Z, Y, X = np.meshgrid(zc, yc, xc, indexing="ij")
    data = (a + b * X + c * Y + d * Z).reshape(-1)
    rng = np.random.default_rng(13)
    x = rng.uniform(xc[0], xc[-1], 3000)
    y = rng.uniform(yc[0], yc[-1], 3000)
    rho = np.ones_like(x)
    got = orc.fluka_heat(oracle_params(data), x, y, rho)
    want = a + b * x + c * y + d * 0.0
    np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-12 * abs(a))


def test_bin_centres_give_the_bin_values_in_x_and_y():
    """At a bin centre in x and y the value is the z interpolation of that
    column alone (the x and y weights are 1 and 0)."""
    rng = np.random.default_rng(14)
    data = rng.uniform(0.0, 1e9, int(np.prod(NBIN)))
    m = data.reshape(NBIN[2], NBIN[1], NBIN[0])
    xc, yc = centres(0), centres(1)
    i, j = np.meshgrid(np.arange(1, NBIN[0] - 1), np.arange(1, NBIN[1] - 1))
    i, j = i.ravel(), j.ravel()
    got = orc.fluka_heat(oracle_params(data), xc[i], yc[j], np.ones(len(i)))
    # the slice z = 0 sits at s = (0 - oz)/dz = 1.5: the centre of z bin 1
    np.testing.assert_allclose(got, m[1, j, i], rtol=1e-14)


def write_form(path, minb, maxb, nbin, dbin, vals, typ=10):
    with open(path, "w") as f:
        f.write(" deposition test run\n")
        f.write(" synthetic binning\n")
        f.write("  1.0  1000\n")
        f.write("  1\n")
        f.write(" edep\n")
        f.write(f"  {typ}\n")
        for k in range(3):
            f.write(f"  {minb[k]:.9e}  {maxb[k]:.9e}  {nbin[k]}  {dbin[k]:.9e}\n")
        f.write(" values\n")
        for s in range(0, len(vals), 5):
            f.write(" " + " ".join(f"{v:.9e}" for v in vals[s:s + 5]) + "\n")


def write_bin(path, minb, maxb, nbin, dbin, vals, typ=10):
    def rec(b):
        n = np.int32(len(b)).tobytes()
        return n + b + n

    r1 = b"deposition test run".ljust(80) + b"synthetic".ljust(32) + \
        np.float32(1.0).tobytes() + np.int32(1000).tobytes()
    r2 = np.int32(1).tobytes() + b"edep".ljust(10) + np.int32(typ).tobytes() + \
        np.int32(208).tobytes()
    for k in range(3):
        r2 += (np.float32(minb[k]).tobytes() + np.float32(maxb[k]).tobytes() +
               np.int32(nbin[k]).tobytes() + np.float32(dbin[k]).tobytes())
    r2 += np.int32(0).tobytes() + np.float32(0).tobytes() * 3
    r3 = np.asarray(vals, np.float32).tobytes()
    with open(path, "wb") as f:
        f.write(rec(r1) + rec(r2) + rec(r3))


# the example scaling of CERNsource/sph.f:230-235 (cm -> m, GeV -> J,
# 3.03e13 protons over 140 ns, the binning shifted 15 cm along z)
SCALE = dict(translation=(0.0, 0.0, 0.15), size_scale=1.0e-2, energy_scale=1.602e-10,
             n_primaries=3.03e13, dep_time=140e-9)


def synthetic_map(rng):
    nbin = (6, 4, 3)
    minb = np.float32([-3.0, -2.0, -20.0])
    dbin = np.float32([0.5, 0.75, 5.0])
    maxb = minb + dbin * np.float32(nbin)
    vals = rng.uniform(0.0, 2e-3, int(np.prod(nbin))).astype(np.float32)
    return minb, maxb, nbin, dbin, vals


@pytest.mark.parametrize("kind", ["txt", "bin"])
def test_readers_equal_reference(tmp_path, kind):
    rng = np.random.default_rng(15)
    minb, maxb, nbin, dbin, vals = synthetic_map(rng)
    if kind == "txt":
        path = os.path.join(tmp_path, "fluka_binning.FORM")
        write_form(path, minb, maxb, nbin, dbin, vals)
        b = io.read_fluka_text(path, **SCALE)
    else:
        path = os.path.join(tmp_path, "fluka_binning")
        write_bin(path, minb, maxb, nbin, dbin, vals)
        b = io.read_fluka_binary(path, **SCALE)
    assert b["type"] == 0 and list(b["n"]) == list(nbin)
    if refbind.available():
        O, D, N, T, data = refbind.read_binning(
            kind, str(tmp_path), SCALE["translation"], SCALE["size_scale"],
            SCALE["energy_scale"], SCALE["n_primaries"], SCALE["dep_time"], len(vals))
        assert T == 0 and list(N) == list(nbin)
        np.testing.assert_array_equal(np.array(b["origin"]), O)
        np.testing.assert_array_equal(np.array(b["cell"]), D)
        np.testing.assert_array_equal(b["data"], data)
    # units: sizes in m, values per unit volume and time
    np.testing.assert_allclose(b["cell"], np.float64(dbin) * 1e-2)
    np.testing.assert_allclose(b["origin"][2], float(minb[2]) * 1e-2 + 0.15)
    f = 3.03e13 * 1.602e-10 / (140e-9 * 1e-6)
    np.testing.assert_allclose(b["data"], f * vals.astype(np.float64), rtol=1e-15)


def test_reader_rejects_cylindrical_binning(tmp_path):
    rng = np.random.default_rng(16)
    minb, maxb, nbin, dbin, vals = synthetic_map(rng)
    path = os.path.join(tmp_path, "fluka_binning.FORM")
    write_form(path, minb, maxb, nbin, dbin, vals, typ=11)
    with pytest.raises(ValueError, match="Cartesian"):
        io.read_fluka_text(path)
