"""Conservative tracers (exTracer/ImTracer with zero sources) on the GPU:
parity with the oracle on a small lake, and mass conservation / bounds at the
benchmark size."""
import numpy as np
import pytest

from si3d_amd import lake
from tests.helpers import relerr, small_ellipse

pytestmark = pytest.mark.gpu


def _with_tracers(d):
    km1 = len(d["init_profile"])
    z = np.arange(km1)
    d["tracer_profiles"] = np.stack([np.where(z < km1 // 2, 10.0, 2.0), np.full(km1, 5.0),
                                     np.linspace(0.0, 1.0, km1)])
    return d


def test_tracers_match_oracle(oracle_lib):
    from si3d_amd.model import GpuLake
    d = _with_tracers(small_ellipse(nts=10))
    d["pcg_tol"] = 1e-12
    print("case: small ellipse, 3 tracers, 6 steps", flush=True)
    o = oracle_lib.Oracle(d)
    g = GpuLake(d, precision="fp64", preconditioner="mg", tol=1e-12)
    for n in range(1, 7):
        o.step(n)
        g.step(n)
    for name in ("tracer1", "tracer2", "tracer3", "tracerpp1", "tracerpp3", "sal"):
        e = relerr(g.get(name), o.get(name))
        assert e < 1e-8, (name, e)
    assert np.max(np.abs(g.get("tracer1") - g.get("tracer1")[0])) > 0  # not trivially uniform


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-11), ("mixed", 1e-5)])
def test_tracer_mass_and_bounds_full_size(precision, tol):
    from si3d_amd.model import GpuLake
    d = _with_tracers(lake.ellipse(nx=1024, ny=512, dx=100.0, nlayers=60, dt=60.0, nts=12, seed=1))
    d["tracer_profiles"] = d["tracer_profiles"][:1]
    print(f"case: 1024x512x60, one tracer, {precision}", flush=True)
    g = GpuLake(d, precision=precision, preconditioner="mg", tol=1e-12)

    def wet_mass():
        t, h = g.get("tracer1"), g.get("h")
        kmz = g.get("kmz")[:g.lm].astype(int)
        k1z = g.get("k1z")[:g.lm].astype(int)
        kk = np.arange(t.shape[1])[None, :] + 1
        wet = (kk >= k1z[:, None]) & (kk <= kmz[:, None])
        return float((h[:g.lm] * t[:g.lm] * wet).sum()), t[:g.lm][wet]

    m0, t0 = wet_mass()
    for n in range(1, 11):
        g.step(n)
    m1, t1 = wet_mass()
    print(f"mass change {(m1 - m0) / m0:.3e}, range {t1.min():.6f} .. {t1.max():.6f}", flush=True)
    assert abs(m1 - m0) <= tol * m0
    assert t1.min() >= 2.0 - 10 * tol and t1.max() <= 10.0 + 10 * tol
    assert np.max(np.abs(t1 - t0)) > 1e-3
