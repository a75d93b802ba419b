"""Open boundaries (types 1 and 2) on the GPU, through the C-ABI
(s3g_set_open_boundaries + s3g_step / s3g_phase), against the CPU oracle on
the same decks: integer state exact, every FP64 field to 1e-8."""
import numpy as np
import pytest

from si3d_amd import lake
from tests.helpers import FIELDS_2D, FIELDS_3D, FIELDS_TURB, INTS, compare, relerr

pytestmark = pytest.mark.gpu

TOL_SOLVE = 1e-12
FIELD_TOL = 1e-8
Q = 20.0


def _wse(npts=200, dtsec=600.0, amp=0.05, mean=0.02, scalar=18.0, tracers=()):
    t = np.arange(npts) * dtsec
    cols = [mean + amp * np.sin(2 * np.pi * t / 7200.0), np.full(npts, scalar)] + [np.full(npts, v) for v in tracers]
    return np.stack(cols, axis=1)


def _pair(deck, oracle_lib, precision="fp64", precond="mg"):
    from si3d_amd.model import GpuLake
    deck = dict(deck)
    deck["pcg_tol"] = TOL_SOLVE
    print("case", deck["name"], "boundaries", deck["openbc"]["iside"], deck["openbc"]["itype"], precision, flush=True)
    return oracle_lib.Oracle(deck), GpuLake(deck, precision=precision, preconditioner=precond, tol=TOL_SOLVE)


def _fields(deck):
    names = FIELDS_2D + FIELDS_3D + (FIELDS_TURB if deck["iturb"] else [])
    ntr = 0 if deck.get("tracer_profiles") is None else len(deck["tracer_profiles"])
    return names + [f"tracer{i + 1}" for i in range(ntr)]


def _steps(o, g, deck, nsteps):
    for n in range(1, nsteps + 1):
        assert o.step(n) == 0
        g.step(n)
        for name in INTS:
            np.testing.assert_array_equal(g.get(name), o.get(name).astype(np.int32), err_msg=f"step {n} {name}")
    compare(g, o, _fields(deck), FIELD_TOL, f"{deck['name']} after {nsteps} steps")


def _flow_through(transpose=False):
    d = lake.channel(nx=30, ny=12, iturb=1, ntr=1)
    if transpose:
        d = lake.transpose_xy(d)
        d = lake.open_boundary(d, "S", 2, 2, 13, lake.steady_series(Q, 15.0, [1.0]))
        d = lake.open_boundary(d, "N", 2, 2, 13, lake.steady_series(Q, 15.0, [1.0]))
    else:
        d = lake.open_boundary(d, "W", 2, 2, 13, lake.steady_series(Q, 15.0, [1.0]))
        d = lake.open_boundary(d, "E", 2, 2, 13, lake.steady_series(Q, 15.0, [1.0]))
    return d


@pytest.mark.parametrize("transpose", [False, True])
def test_flow_through_matches_oracle(oracle_lib, transpose):
    """Flow in at one end, out at the other (the rows of both ends overlap,
    so the boundary rows shift twice per step, as in the reference)."""
    deck = _flow_through(transpose)
    o, g = _pair(deck, oracle_lib)
    _steps(o, g, deck, 10)


@pytest.mark.parametrize("side", ["W", "E", "N", "S"])
def test_water_level_boundary_matches_oracle(oracle_lib, side):
    d = lake.channel(nx=24, ny=10, iturb=1, ntr=1)
    if side in ("N", "S"):
        d = lake.transpose_xy(d)
    a1 = (d["jmm"] if side in ("W", "E") else d["imm"]) + 1
    d = lake.open_boundary(d, side, 1, 2, a1, _wse(tracers=[1.0]))
    o, g = _pair(d, oracle_lib)
    _steps(o, g, d, 10)


def test_water_level_and_outflow_matches_oracle(oracle_lib):
    """A tidal water level at the west end, a river leaving at the east end
    over part of the section, Jacobi-preconditioned solve."""
    d = lake.channel(nx=30, ny=12, iturb=0)
    d = lake.open_boundary(d, "W", 1, 2, 13, _wse())
    d = lake.open_boundary(d, "E", 2, 4, 11, lake.steady_series(5.0, 20.0))
    o, g = _pair(d, oracle_lib, precond="jacobi")
    _steps(o, g, d, 12)


def test_phase_by_phase_with_boundaries(oracle_lib):
    """After three steps, every phase of a leapfrog fd: the recomputed
    momentum terms of the boundary faces (phases 2, 4), the modified
    continuity system (6), the boundary zeta and flows after the solve (7),
    w (8) and the scalar (9, 10)."""
    d = lake.channel(nx=24, ny=10, iturb=1)
    d = lake.open_boundary(d, "W", 1, 2, 11, _wse())
    d = lake.open_boundary(d, "E", 2, 2, 11, lake.steady_series(Q, 15.0))
    d = lake.open_boundary(d, "S", 2, 3, 8, lake.steady_series(4.0, 16.0))
    o, g = _pair(d, oracle_lib)
    for n in range(1, 4):
        o.step(n)
        g.step(n)
    thrs = 4 * d["dt"] / 3600.0
    checks = {2: ["ex"], 3: ["agx", "arx", "eagx", "earx"], 5: ["agy", "ary", "eagy", "eary"],
              6: ["sx", "sy", "qq", "rr"], 7: ["s", "h", "hu", "hv"], 8: ["uh", "vh", "wp"], 9: ["bex"],
              10: ["sal"], 13: ["Av", "Dv", "q2", "q2l"]}
    for ph in [0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 13]:
        o.phase(ph, 4, 1, thrs)
        g.phase(ph, 4, 1, thrs)
        compare(g, o, checks.get(ph, []), 1e-11 if ph < 7 else FIELD_TOL, f"phase {ph}")


def test_long_run_with_tide_and_river_matches_oracle(oracle_lib):
    """300 steps (5 h at dt = 60 s) of a tidal water level at the west end and
    a river of colder, tracer-laden water entering over part of the south
    side, every field against the oracle: to 1e-8 after 20 steps, and after
    300 to 1e-6.  The difference grows as the two solvers' residuals (each
    below 1e-12, oracle MIC(0)-CG, GPU multigrid-CG) are amplified by the
    flow: the oracle against itself with solver tolerances 1e-12 and 1e-14
    differs by 2.0e-9, 2.4e-8 and 8.5e-8 after 100, 200 and 300 steps, the
    GPU against the oracle by 2.4e-9, 2.8e-8 and 9.9e-8.  The integer state
    stays exact throughout."""
    d = lake.channel(nx=30, ny=12, iturb=1, ntr=1, dt=60.0, nts=300)
    d = lake.open_boundary(d, "W", 1, 2, 13, _wse(npts=60, amp=0.03, mean=0.0, scalar=18.0, tracers=[1.0]))
    d = lake.open_boundary(d, "S", 2, 8, 14, lake.steady_series(3.0, 12.0, [2.0], npts=60))
    o, g = _pair(d, oracle_lib)
    errs = {}
    for n in range(1, 301):
        assert o.step(n) == 0
        g.step(n)
        if n in (20, 100, 200, 300):
            for name in INTS:
                np.testing.assert_array_equal(g.get(name), o.get(name).astype(np.int32), err_msg=f"step {n} {name}")
            errs[n] = max(relerr(g.get(f), o.get(f)) for f in ["s", "uh", "vh", "sal", "q2", "tracer1"])
            print("step", n, "max relative difference", errs[n], flush=True)
            if n == 20:
                compare(g, o, _fields(d), FIELD_TOL, "tide + river after 20 steps")
    compare(g, o, _fields(d), 1e-6, "tide + river after 300 steps")


def test_mixed_precision_close_to_oracle(oracle_lib):
    deck = _flow_through()
    o, g = _pair(deck, oracle_lib, precision="mixed")
    for n in range(1, 11):
        o.step(n)
        g.step(n)
    errs = {n: relerr(g.get(n), o.get(n)) for n in ["s", "uh", "wp", "sal", "tracer1"]}
    print("mixed-precision relative errors:", errs, flush=True)
    assert errs["s"] < 1e-4 and errs["sal"] < 1e-5 and errs["uh"] < 1e-3 and errs["tracer1"] < 1e-4, errs


def test_volume_grows_by_the_inflow(oracle_lib):
    """V(n) - V(0) = Q dt (n - 1/2) on the GPU as in the oracle's closed form."""
    from si3d_amd.model import GpuLake
    d = lake.open_boundary(lake.channel(), "W", 2, 2, 13, lake.steady_series(Q, 20.0))
    g = GpuLake(d, precision="fp64", preconditioner="mg", tol=TOL_SOLVE)
    lm = g.lm
    v0 = g.get("s")[:lm].sum() * d["dx"] * d["dy"]
    for n in range(1, 21):
        g.step(n)
    dv = g.get("s")[:lm].sum() * d["dx"] * d["dy"] - v0
    expect = Q * d["dt"] * 19.5
    assert abs(dv - expect) < 1e-8 * expect, (dv, expect)


def test_rejects_unsupported_boundaries():
    from si3d_amd.model import GpuLake
    base = lake.channel()
    with pytest.raises(RuntimeError, match="types 1"):
        GpuLake(lake.open_boundary(base, "W", 3, 2, 13, lake.steady_series(Q, 20.0)), precision="fp64")
    bad = lake.open_boundary(base, "W", 2, 2, 13, lake.steady_series(Q, 20.0))
    bad["openbc"]["isbc"] = [3]
    bad["openbc"]["iebc"] = [3]
    with pytest.raises(RuntimeError, match="wet edge cell"):
        GpuLake(bad, precision="fp64")
