"""Readers of pSi3D's input files (si3d_amd/inputs.py) on the reference's own
sample inputs (tests/golden/psi3d/, copied by make_psi3d_inputs.py), and the
writers' round trip."""
import os

import numpy as np
import pytest

from si3d_amd import inputs as I
from si3d_amd import lake

G = os.path.join(os.path.dirname(__file__), "golden", "psi3d")


def test_inp_sample_files():
    p = I.read_inp(os.path.join(G, "sample_files", "si3d_inp.txt"))
    assert p["title"].startswith("LAKE Hg Test RUN 2")
    assert (p["iyr0"], p["imon0"], p["iday0"], p["ihr0"]) == (2023, 1, 1, 0)
    assert (p["xl"], p["yl"], p["zl"], p["tl"]) == (100.0, 100.0, 10.0, 31536000.0)
    assert (p["idx"], p["idy"], p["idz"], p["dzmin"], p["idt"], p["ibathyf"]) == (10.0, 10.0, 0.5, 0.1, 10.0, 0)
    assert (p["itrap"], p["niter"], p["ismooth"], p["iturb"], p["iadv"], p["itrmom"], p["ihd"]) == (1, 2, 0, 1, 1, 1, 2)
    assert (p["Av0"], p["Ax0"], p["f"], p["theta"], p["cd"]) == (3e-2, 1e-2, 0.0, 1.0, 0.025)
    assert (p["ibc"], p["isal"], p["ifsurfbc"], p["dtsurfbc"], p["icw"]) == (1, 1, 0, 3600.0, 1)
    assert (p["cw"], p["wa"], p["phi"], p["num_threads"]) == (1.2e-3, 0.025, 0.0, 4)
    assert (p["ipt"], p["nnodes"], p["inode"], p["jnode"]) == (360, 3, [2, 6, 11], [6, 6, 6])
    assert (p["iop"], p["n_planes"], p["p_out"], p["iox"], p["ipxml"]) == (360, 1, [2], 0, 360)
    assert (p["nopen"], p["nxNBO"], p["ntr"], p["iopss"]) == (0, 0, 4, 0)
    assert (p["im"], p["jm"], p["nts"]) == (11, 11, 3153600)


def test_bathy_layers_init_surfbc_sample_files():
    imm, jmm, b = I.read_bathy(os.path.join(G, "sample_files", "si3d_bathy"))
    assert (imm, jmm) == (26, 44) and b.shape == (26, 44)
    assert b.min() == -99 and b.max() == 5010 and int((b > 0).sum()) == 774
    np.testing.assert_array_equal(b[12, :8], [-99, 96, 1705, 3288, 3807, 4128, 4340, 4470])
    # the sample inp (11 x 11 cells) does not match this bathymetry: the reference stops there too
    with pytest.raises(I.InputError, match="does not agree"):
        I.read_bathy(os.path.join(G, "sample_files", "si3d_bathy"), 11, 11)
    z = I.read_layers(os.path.join(G, "sample_files", "si3d_layer.txt"))
    assert len(z) == 172 and z[0] == z[1] == -100.0 and z[2] == 0.51 and z[-1] == 505.0
    assert np.all(np.diff(z[2:]) > 0)
    prof, tr = I.read_init(os.path.join(G, "sample_files", "si3d_init.txt"), 20)
    assert np.all(prof == 4.0) and tr.shape == (0, 20)
    sb = I.read_surfbc(os.path.join(G, "sample_files", "si3d_surfbc.txt"), 9)
    assert sb.shape == (48, 9)
    np.testing.assert_array_equal(sb[0], [0.1045, 0.0, -1.26, 81450.0, 0.815, 196.9, 0.0011, 1.5111, -3.5235])


def test_wind_driven_flow_sample_run():
    """The windDrivenFlow sample predates the icw record: the strict reader
    stops there (the reference reads it one record off); with missing_icw the
    domain, forcing and output nodes come out as the file states, and its 'h'
    bathymetry is the flat 40 m rectangle of the readme."""
    path = os.path.join(G, "windDrivenFlow", "si3d_inp.txt")
    with pytest.raises(I.InputError, match="icw"):
        I.read_inp(path)
    p = I.read_inp(path, missing_icw=True)
    assert (p["xl"], p["yl"], p["zl"], p["idx"], p["idt"]) == (40000.0, 2000.0, 40.25, 50.0, 50.0)
    assert (p["iturb"], p["iadv"], p["ihd"], p["isal"], p["ifsurfbc"]) == (0, 0, 0, 0, 0)
    assert (p["cd"], p["cw"], p["wa"], p["phi"]) == (0.625, 1.2e-3, 8.125, 270.0)
    assert (p["nnodes"], p["inode"], p["jnode"]) == (3, [401, 401, 3], [21, 20, 20])
    assert (p["im"], p["jm"], p["nts"]) == (801, 41, 3456)
    imm, jmm, b = I.read_bathy(os.path.join(G, "windDrivenFlow", "h"), p["im"], p["jm"])
    assert (imm, jmm) == (800, 40) and np.all(b == 400)
    prof, _ = I.read_init(os.path.join(G, "windDrivenFlow", "si3d_init.txt"), 82)
    assert np.all(prof == 4.0)


def test_clear_lake_met_station_series_is_read():
    """CLWaterQuality's si3d_surfbc.txt (ifsurfbc 11; the fixture keeps the
    first 48 of its 2252 hourly records): 7 stations at their grid positions,
    then records of attenuation, pressure and, per station, shortwave, air
    temperature, humidity, incoming longwave and wind, as its header says."""
    xy, s = I.read_met_surfbc(os.path.join(G, "CLWaterQuality", "si3d_surfbc.txt"))
    assert xy.shape == (7, 2) and tuple(xy[0]) == (143.0, 82.0) and tuple(xy[2]) == (227.0, 26.0)
    assert s.shape == (48, 6 * 7 + 2)
    assert (s[0, 0], s[0, 1]) == (0.5, 96630.603)
    assert tuple(s[0, 2:8]) == (0.0, 11.35, 0.705, 129.5495, 0.7534, -0.4771)
    assert s[2, 2] == 0.7271


def test_clear_lake_case_is_read_and_refused():
    """CLWaterQuality: met-station interpolation (ifsurfbc 11, read) and the
    water quality model with 11 tracers; its input records are read, and
    read_case names the option this path does not run."""
    p = I.read_inp(os.path.join(G, "CLWaterQuality", "si3d_inp.txt"))
    assert (p["im"], p["jm"], p["ifsurfbc"], p["ntr"], p["ecomod"], p["iinterp"]) == (240, 194, 11, 11, 1, 1)
    z = I.read_layers(os.path.join(G, "CLWaterQuality", "si3d_layer.txt"))
    assert len(z) == 37 and z[2] == 0.5
    prof, tr = I.read_init(os.path.join(G, "CLWaterQuality", "si3d_init.txt"), 37, 11)
    assert prof[0] == 20.6994 and tr[0, 0] == 6393.75 and tr.shape == (11, 37)
    with pytest.raises(I.InputError, match="ecomod = 1"):
        I.read_case(os.path.join(G, "CLWaterQuality"))


def _river():
    """Channel with a tidal water level at the west end, a river leaving
    through part of the east end and a tributary entering from the south."""
    d = lake.channel(nx=30, ny=12, ntr=1)
    t = np.arange(50) * 600.0
    d = lake.open_boundary(d, "W", 1, 2, 13, np.stack([0.02 + 0.05 * np.sin(t / 1146.0), np.full(50, 18.0),
                                                         np.full(50, 1.0)], axis=1))
    d = lake.open_boundary(d, "E", 2, 4, 11, lake.steady_series(5.0, 20.0, [0.0], npts=50))
    return lake.open_boundary(d, "S", 2, 3, 8, lake.steady_series(4.0, 16.5, [0.5], npts=50))


@pytest.mark.parametrize("which", ["ellipse", "rectangle", "heat3", "river", "met11"])
def test_write_read_round_trip(tmp_path, which):
    if which == "met11":
        base = lake.ellipse(nx=40, ny=24, dx=100.0, nlayers=12, nts=20, top=1.0, bottom=12.0, depth=40.0, seed=3)
        d = lake.met_stations(base, 11, [(6.0, 5.0), (33.5, 9.25)], iinterp=2, gammaB=0.3, delNfactor=0.8, icw=1)
    elif which == "ellipse":
        d = lake.ellipse(nx=40, ny=24, dx=100.0, nlayers=12, nts=20, top=1.0, bottom=12.0, depth=40.0, seed=3)
    elif which == "rectangle":
        d = lake.rectangle(nx=100, ny=50, dx=50.0, depth=20.0, dz=1.0, tau=0.1, dt=30.0, nts=210)
    elif which == "river":
        d = _river()
    else:
        d = lake.heat_budget(lake.ellipse(nx=40, ny=24, dx=100.0, nlayers=12, nts=20, top=1.0, bottom=12.0,
                                          depth=40.0, seed=3), 3)
    I.write_case(d, str(tmp_path))
    e = I.read_case(str(tmp_path))
    for k in ("imm", "jmm", "dx", "dy", "ddz", "zl", "dt", "nts", "ibathyf", "dzmin", "datadj", "itrap", "niter",
              "ismooth", "beta", "iturb", "Av0", "Dv0", "iadv", "itrmom", "ihd", "Ax0", "Ay0", "f", "theta", "ibc",
              "isal", "cd", "ifsurfbc", "dtsurfbc", "cw", "wa", "phi"):
        assert e[k] == d[k], k
    np.testing.assert_array_equal(e["bathy_dm"], d["bathy_dm"])
    if d["ibathyf"] < 0:  # layer depths are single precision in the reference
        np.testing.assert_array_equal(np.float32(e["zlevel"]), np.float32(d["zlevel"]))
    np.testing.assert_allclose(e["init_profile"], d["init_profile"], atol=5e-6)  # G11.5
    if d.get("surfbc") is not None:
        np.testing.assert_allclose(e["surfbc"], d["surfbc"], rtol=5e-5, atol=1e-12)  # G11.4
    if d.get("metxy") is not None:
        np.testing.assert_allclose(e["metxy"], d["metxy"], atol=5e-5)
        for k in ("nmetstat", "iinterp", "gammaB", "delNfactor", "icw"):
            assert e[k] == d[k], k
    if d.get("openbc") is not None:
        for k in ("iside", "itype", "isbc", "jsbc", "iebc", "jebc", "dtsec"):
            assert e["openbc"][k] == d["openbc"][k], k
        np.testing.assert_allclose(e["openbc"]["values"], d["openbc"]["values"], rtol=5e-5, atol=1e-12)
    else:
        assert e.get("openbc") is None
    assert lake.km1_of(e) == lake.km1_of(d)


def test_river_case_from_files_runs_like_the_deck(oracle_lib, tmp_path):
    """Open boundaries read back from si3d_inp.txt and openbcNN.txt drive the
    oracle like the deck they were written from (up to the files' 5 digits)."""
    d = _river()
    I.write_case(d, str(tmp_path))
    assert sorted(f for f in os.listdir(tmp_path) if f.startswith("openbc")) == [
        "openbc01.txt", "openbc02.txt", "openbc03.txt"]
    e = I.read_case(str(tmp_path))
    a, b = oracle_lib.Oracle(d), oracle_lib.Oracle(e)
    for n in range(1, 6):
        a.step(n)
        b.step(n)
    for name in ("s", "uh", "sal", "tracer1"):
        x, y = a.get(name), b.get(name)
        assert np.max(np.abs(x - y)) <= 1e-4 * np.max(np.abs(x)), name


def test_open_boundary_types_3_and_4_are_refused(tmp_path):
    d = lake.open_boundary(lake.channel(), "W", 3, 2, 13, lake.steady_series(5.0, 20.0))
    I.write_case(d, str(tmp_path))
    with pytest.raises(I.InputError, match="type 3 or 4"):
        I.read_case(str(tmp_path))


def test_case_from_files_runs_like_the_deck(oracle_lib, tmp_path):
    """A case written to files and read back gives the oracle the same grid and,
    up to the files' precision, the same flow."""
    d = lake.ellipse(nx=40, ny=24, dx=100.0, nlayers=12, nts=20, top=1.0, bottom=12.0, depth=40.0, seed=3)
    I.write_case(d, str(tmp_path))
    e = I.read_case(str(tmp_path))
    a, b = oracle_lib.Oracle(d), oracle_lib.Oracle(e)
    for n in range(1, 4):
        a.step(n)
        b.step(n)
    for name in ("kmz", "k1z", "k1u", "k1v"):
        np.testing.assert_array_equal(a.get(name), b.get(name))
    for name in ("s", "uh", "sal"):
        x, y = a.get(name), b.get(name)
        assert np.max(np.abs(x - y)) <= 1e-4 * np.max(np.abs(x)), name
