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
