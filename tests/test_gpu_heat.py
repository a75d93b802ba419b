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
