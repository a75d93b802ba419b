"""Surface heat budget (ifsurfbc 1-3: spatially uniform series, penetrative
shortwave, longwave, latent and sensible fluxes) on the GPU: parity with the
oracle on a small lake, and the closed-basin heat budget at the benchmark
size (the heat content follows the sources)."""
import numpy as np
import pytest

from si3d_amd import lake
from tests.helpers import relerr, small_ellipse

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("mode", [1, 2, 3])
def test_heat_budget_matches_oracle(oracle_lib, mode):
    from si3d_amd.model import GpuLake
    d = lake.heat_budget(small_ellipse(nts=10), mode)
    d["pcg_tol"] = 1e-12
    print(f"case: small ellipse, ifsurfbc {mode}, 6 steps", flush=True)
    o = oracle_lib.Oracle(d)
    g = GpuLake(d, precision="fp64", preconditioner="mg", tol=1e-12)
    for n in range(1, 7):
        o.step(n)
        g.step(n)
    for name in ("HeatSource", "sal", "rhop", "s", "uh", "vh", "Dv"):
        e = relerr(g.get(name), o.get(name))
        assert e < 1e-8, (name, e)
    assert np.max(np.abs(g.get("HeatSource"))) > 0


def _wet(g):
    kmz = g.get("kmz")[:g.lm].astype(int)
    k1z = g.get("k1z")[:g.lm].astype(int)
    kk = np.arange(g.km1)[None, :] + 1
    return (kk >= k1z[:, None]) & (kk <= kmz[:, None])


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-9), ("mixed", 1e-5)])
def test_heat_content_follows_sources_full_size(precision, tol):
    """1024x512x60, ifsurfbc = 2: after each step sum(h T) changes by
    dt sum(HeatSource) (closed basin; the sources of the step's leapfrog call
    enter the trapezoidal iteration with twodt1 = dt)."""
    from si3d_amd.model import GpuLake
    d = lake.heat_budget(lake.ellipse(nx=1024, ny=512, dx=100.0, nlayers=60, dt=60.0, nts=12, seed=1), 2)
    print(f"case: 1024x512x60, ifsurfbc 2, {precision}", flush=True)
    g = GpuLake(d, precision=precision, preconditioner="mg", tol=1e-12)
    g.step(1)
    for n in range(2, 5):
        w = _wet(g)
        h0 = float((g.get("h")[:g.lm] * g.get("sal")[:g.lm] * w).sum())
        g.step(n)
        w = _wet(g)
        src = float((g.get("HeatSource")[:g.lm] * w).sum())
        h1 = float((g.get("h")[:g.lm] * g.get("sal")[:g.lm] * w).sum())
        print(f"step {n}: d(sum h T) {h1 - h0:.6e}, dt sum(HeatSource) {d['dt'] * src:.6e}", flush=True)
        assert abs(src) > 0
        assert abs((h1 - h0) - d["dt"] * src) <= tol * abs(h0) + 1e-4 * abs(d["dt"] * src)


def test_heat_budget_from_the_reference_met_file(oracle_lib):
    """ifsurfbc = 3 driven by the first 48 records of the reference's Lake Tahoe
    met series (si3d_surfbc.txt, 10-minute records: attenuation, shortwave,
    air temperature, pressure, humidity, longwave, drag, wind), read with the
    reference's record layout: GPU = oracle."""
    import os
    from si3d_amd import inputs
    from si3d_amd.model import GpuLake
    path = os.path.join(os.path.dirname(__file__), "golden", "psi3d", "sample_files", "si3d_surfbc.txt")
    d = small_ellipse(nts=10)
    d.update(ifsurfbc=3, dtsurfbc=600.0, surfbc=inputs.read_surfbc(path, 9), pcg_tol=1e-12)
    print("case: small ellipse, Lake Tahoe met series (ifsurfbc 3), 6 steps", flush=True)
    o = oracle_lib.Oracle(d)
    g = GpuLake(d, precision="fp64", preconditioner="mg", tol=1e-12)
    for n in range(1, 7):
        o.step(n)
        g.step(n)
    for name in ("HeatSource", "sal", "s", "uh", "vh"):
        e = relerr(g.get(name), o.get(name))
        assert e < 1e-8, (name, e)
    assert np.max(np.abs(g.get("HeatSource"))) > 0


MET_STATIONS = [(6.0, 5.0), (33.0, 9.0), (18.0, 20.0), (28.0, 22.0)]


@pytest.mark.parametrize("mode,iinterp,icw", [(11, 1, 0), (11, 2, 1), (11, 3, 2), (10, 1, 0), (10, 2, 2)])
def test_met_station_modes_match_oracle(oracle_lib, mode, iinterp, icw):
    """ifsurfbc 10 / 11 (four met stations, each with its own shortwave, air
    temperature, humidity, longwave and turning wind, interpolated to every
    column by Barnes, two-pass Barnes or inverse distance; per-column wind
    and drag) against the oracle over 6 steps, FP64 to 1e-8."""
    from si3d_amd.model import GpuLake
    d = lake.met_stations(small_ellipse(nts=10), mode, MET_STATIONS, iinterp=iinterp, gammaB=0.3,
                          delNfactor=0.8, icw=icw)
    d["pcg_tol"] = 1e-12
    print(f"case: small ellipse, ifsurfbc {mode}, iinterp {iinterp}, icw {icw}, 6 steps", flush=True)
    o = oracle_lib.Oracle(d)
    g = GpuLake(d, precision="fp64", preconditioner="mg", tol=1e-12)
    for n in range(1, 7):
        o.step(n)
        g.step(n)
    for name in ("HeatSource", "sal", "rhop", "s", "uh", "vh", "q2", "Dv"):
        e = relerr(g.get(name), o.get(name))
        assert e < 1e-8, (name, e)
    hs = g.get("HeatSource")[:g.lm]
    assert np.max(np.abs(hs)) > 0


def test_met_station_mode_mixed_precision(oracle_lib):
    from si3d_amd.model import GpuLake
    d = lake.met_stations(small_ellipse(nts=10), 11, MET_STATIONS, iinterp=2, gammaB=0.3, delNfactor=0.8)
    o = oracle_lib.Oracle(d)
    g = GpuLake(d, precision="mixed", preconditioner="mg", tol=1e-12)
    for n in range(1, 7):
        o.step(n)
        g.step(n)
    errs = {n: relerr(g.get(n), o.get(n)) for n in ("s", "sal", "HeatSource", "uh")}
    print("mixed:", errs, flush=True)
    assert errs["s"] < 1e-4 and errs["sal"] < 1e-5 and errs["HeatSource"] < 1e-4 and errs["uh"] < 1e-3, errs


def test_clear_lake_met_stations_drive_gpu_and_oracle_alike(oracle_lib):
    """The reference's own Clear Lake met-station series (7 stations, hourly
    records of attenuation, pressure, shortwave, air temperature, humidity,
    incoming longwave and wind; ifsurfbc 11, Barnes interpolation, the
    Amorocho-DeVries drag of its si3d_inp.txt) over a synthetic lake on the
    Clear Lake grid (240 x 194 cells of 100 m, dt 20 s): GPU = oracle to 1e-8
    after 4 steps."""
    import os
    from si3d_amd import inputs as I
    from si3d_amd.model import GpuLake
    G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "psi3d", "CLWaterQuality")
    p = I.read_inp(os.path.join(G, "si3d_inp.txt"))
    xy, series = I.read_met_surfbc(os.path.join(G, "si3d_surfbc.txt"))
    d = lake.ellipse(nx=p["im"] - 1, ny=p["jm"] - 1, dx=p["idx"], nlayers=20, dt=p["idt"], nts=10, seed=4,
                     depth=30.0, top=0.5, bottom=3.0)
    d.update(ifsurfbc=11, surfbc=series, dtsurfbc=p["dtsurfbc"], metxy=xy.ravel(), nmetstat=len(xy),
             iinterp=p["iinterp"], icw=p["icw"], cw=p["cw"], pcg_tol=1e-12)
    print("case: Clear Lake met stations on a", p["im"] - 1, "x", p["jm"] - 1, "lake", flush=True)
    o = oracle_lib.Oracle(d)
    g = GpuLake(d, precision="fp64", preconditioner="mg", tol=1e-12)
    for n in range(1, 5):
        o.step(n)
        g.step(n)
    for name in ("HeatSource", "sal", "s", "uh", "vh", "q2"):
        e = relerr(g.get(name), o.get(name))
        assert e < 1e-8, (name, e)
