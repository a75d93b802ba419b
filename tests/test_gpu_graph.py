"""The time step replayed from captured HIP graphs: the kernel sequences of
fd before and after the free-surface solve, settrap, settrap2 and save are
captured once per (istep, lastiter) and replayed on every later call.  A
replayed run must equal the run that launches every kernel bit for bit, with
the wind changing every call (it reaches the kernels through device memory,
not through the captured arguments), tracers on the last trapezoidal
iteration only, the Asselin filter on leapfrog steps, an ensemble of lakes,
and mixed precision."""
import os
import time

import numpy as np
import pytest

from tests.helpers import FIELDS_2D, FIELDS_3D, FIELDS_TURB, INTS, small_ellipse

pytestmark = pytest.mark.gpu


def _pair(make):
    """The same context built twice: with the step graphs (default) and with
    S3G_STEP_GRAPH=0 (read when a context is created)."""
    a = make()
    os.environ["S3G_STEP_GRAPH"] = "0"
    try:
        b = make()
    finally:
        del os.environ["S3G_STEP_GRAPH"]
    return a, b


def _same(a, b, names):
    for n in names:
        x, y = a.get(n), b.get(n)
        assert np.array_equal(x, y), (n, float(np.max(np.abs(np.asarray(x, float) - np.asarray(y, float)))))


def _deck(kind):
    d = small_ellipse(nts=30)
    if kind == "tracers":
        km1 = len(d["init_profile"])
        d["tracer_profiles"] = np.stack([np.where(np.arange(km1) < km1 // 2, 10.0, 2.0),
                                         np.linspace(0.0, 1.0, km1)])
    elif kind == "asselin":
        d["itrap"], d["ismooth"] = 0, 1
    elif kind == "trap2":
        d["itrap"], d["niter"] = 2, 3
    return d


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("kind", ["plain", "tracers", "asselin", "trap2"])
def test_replayed_step_equals_launched_step(kind, precision):
    from si3d_amd.model import GpuLake
    print("case", kind, precision, flush=True)
    d = _deck(kind)
    a, b = _pair(lambda: GpuLake(d, precision=precision, preconditioner="mg", tol=1e-10))
    for n in range(1, 13):
        a.step(n)
        b.step(n)
    names = FIELDS_2D + FIELDS_3D + FIELDS_TURB + INTS
    if kind == "tracers":
        names += ["tracer1", "tracer2", "tracerpp1", "tracerpp2"]
    _same(a, b, names)
    sa, sb = a.stats(), b.stats()
    assert sa["graph_replays"] > 0 and sb["graph_replays"] == 0, (sa["graph_replays"], sb["graph_replays"])
    assert sa["pcg_iters_total"] == sb["pcg_iters_total"]


def test_replayed_ensemble_equals_launched_ensemble():
    """Three lakes of different sizes, layers and winds: the per-lake wind is
    uploaded before every fd call, outside the captured sequences."""
    from si3d_amd import lake
    from si3d_amd.model import GpuEnsemble
    decks = lake.ensemble(3, seed=5, side=(24, 40), layers=(8, 16), nts=20)
    a, b = _pair(lambda: GpuEnsemble(decks, precision="fp64", preconditioner="mg", tol=1e-10))
    for n in range(1, 9):
        a.step(n)
        b.step(n)
    _same(a, b, ["s", "uh", "vh", "wp", "sal", "q2", "Av", "k1z"])
    assert a.stats()["graph_replays"] > 0


def test_replay_gain_on_a_small_lake():
    """On a lake whose kernels last a few microseconds the step is bound by
    launch overhead; report both rates (printed) and require the replayed
    step not to be slower."""
    from si3d_amd.model import GpuLake
    d = small_ellipse(nts=200)
    a, b = _pair(lambda: GpuLake(d, precision="fp64", preconditioner="mg", tol=1e-10))
    rates = {}
    for name, g in (("replayed", a), ("launched", b)):
        for n in range(1, 11):
            g.step(n)
        g.synchronize()
        t0 = time.perf_counter()
        for n in range(11, 111):
            g.step(n)
        g.synchronize()
        rates[name] = 1e3 * (time.perf_counter() - t0) / 100
    print(f"ms per step on the 40x24x12 lake: replayed {rates['replayed']:.3f}, launched {rates['launched']:.3f}",
          flush=True)
    _same(a, b, ["s", "uh", "sal"])
    assert rates["replayed"] <= 1.05 * rates["launched"]
