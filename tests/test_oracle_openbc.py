"""Open boundaries in the CPU oracle (types 1 and 2 of
si3d_boundaryconditions.f90), pinned by what the mathematics fixes:

* volume: a flow boundary adds Q dt per step (half a step on step 1, whose
  n-1 boundary transport is still zero), so the lake volume after n steps is
  V0 + Q dt (n - 1/2) on every side; flow in at one end and out at the other
  leaves it unchanged;
* a scalar that enters at the lake's own value stays uniform, and its content
  grows with the volume;
* a water-level boundary holds the series value, and a lake closed elsewhere
  fills to that level;
* reflections: the same lake reflected in x (west <-> east flow boundary) or y
  (north <-> south) and exchanged in x and y (west <-> south) gives the
  reflected fields, which pins the per-side index and sign of every boundary
  term against its mirror image.  The inflow keeps the lake's temperature
  (matmom averages the density of the column west / south of a face only, so
  a density contrast at the boundary is not mirror-symmetric in the
  reference; what remains of it, the pressure term of the equation of state
  under a raised water level, stays below 1e-8) and carries a tracer at 1
  into a lake at 0;
* the reference's east water-level quirk: the face inside it keeps ex = 0.
"""
import numpy as np
import pytest

from si3d_amd import lake

DT = 30.0
Q = 20.0
MIRROR_TOL = 1e-8


def _vol(o):
    return float(o.get("s")[:o.lm].sum()) * o.deck["dx"] * o.deck["dy"]


def _run(o, nsteps):
    for n in range(1, nsteps + 1):
        assert o.step(n) == 0


def _inflow(side):
    """Flow (m3/s, positive toward +x/+y) that enters through `side`."""
    return Q if side in ("W", "S") else -Q


def _range(deck, side):
    return (2, deck["jmm"] + 1) if side in ("W", "E") else (2, deck["imm"] + 1)


def _tight(deck):
    d = dict(deck)
    d["pcg_tol"] = 1e-13  # reflected runs differ only by the order of the solver's sums
    return d


def _lake(side, btype=2, value=None, scalar=20.0, nx=30, ny=12, **kw):
    d = lake.channel(nx=nx, ny=ny, **kw) if side in ("W", "E") else lake.transpose_xy(lake.channel(nx=nx, ny=ny, **kw))
    a0, a1 = _range(d, side)
    v = _inflow(side) if value is None else value
    return lake.open_boundary(d, side, btype, a0, a1, lake.steady_series(v, scalar))


def test_flow_through_conserves_volume(oracle_lib):
    d = lake.channel()
    d = lake.open_boundary(d, "W", 2, 2, 13, lake.steady_series(Q, 20.0))
    d = lake.open_boundary(d, "E", 2, 2, 13, lake.steady_series(Q, 20.0))
    o = oracle_lib.Oracle(d)
    v0 = _vol(o)
    _run(o, 30)
    assert abs(_vol(o) - v0) < 1e-9 * Q * DT * 30
    assert o.get("uh")[:o.lm].mean() > 0.0  # the flow runs from west to east


@pytest.mark.parametrize("side", ["W", "E", "N", "S"])
def test_inflow_volume_closed_form(oracle_lib, side):
    o = oracle_lib.Oracle(_lake(side))
    v0 = _vol(o)
    for n in range(1, 21):
        assert o.step(n) == 0
        if n in (1, 7, 20):
            expect = Q * DT * (n - 0.5)
            assert abs((_vol(o) - v0) - expect) < 1e-9 * expect, (side, n, _vol(o) - v0, expect)


@pytest.mark.parametrize("side", ["W", "N"])
def test_outflow_volume_closed_form(oracle_lib, side):
    o = oracle_lib.Oracle(_lake(side, value=-_inflow(side)))
    v0 = _vol(o)
    _run(o, 10)
    expect = -Q * DT * 9.5
    assert abs((_vol(o) - v0) - expect) < 1e-9 * abs(expect)


@pytest.mark.parametrize("side", ["W", "E", "N", "S"])
def test_inflow_at_lake_value_keeps_scalar_uniform(oracle_lib, side):
    o = oracle_lib.Oracle(_lake(side, scalar=20.0))
    h0 = float((o.get("h")[:o.lm] * o.get("sal")[:o.lm]).sum())
    _run(o, 15)
    sal = o.get("sal")[:o.lm]
    wet = o.get("h")[:o.lm] > 0
    assert np.max(np.abs(sal[wet] - 20.0)) < 1e-10
    dh = (float((o.get("h")[:o.lm] * sal).sum()) - h0) * o.deck["dx"] * o.deck["dy"]
    expect = Q * 20.0 * DT * 14.5
    assert abs(dh - expect) < 1e-8 * expect


def test_inflow_scalar_enters(oracle_lib):
    o = oracle_lib.Oracle(_lake("W", scalar=10.0))
    _run(o, 20)
    sal = o.get("sal")[:o.lm]
    wet = o.get("h")[:o.lm] > 0
    assert sal[wet].min() < 19.8 and sal[wet].max() <= 20.0 + 1e-12
    # the coldest water sits next to the inflow (column i = 2)
    i = o.get("l2i")[:o.lm].astype(int)
    assert sal[(i == 2)[:, None] & wet].mean() < sal[(i == 20)[:, None] & wet].mean()


def _wse_series(npts=200, dtsec=600.0, amp=0.05, mean=0.02, tracer=None):
    t = np.arange(npts) * dtsec
    cols = [mean + amp * np.sin(2 * np.pi * t / 7200.0), np.full(npts, 20.0)]
    if tracer is not None:
        cols.append(np.full(npts, tracer))
    return np.stack(cols, axis=1)


@pytest.mark.parametrize("side", ["W", "E", "N", "S"])
def test_water_level_boundary_holds_series(oracle_lib, side):
    d = lake.channel() if side in ("W", "E") else lake.transpose_xy(lake.channel())
    a0, a1 = _range(d, side)
    ser = _wse_series()
    d = lake.open_boundary(d, side, 1, a0, a1, ser)
    o = oracle_lib.Oracle(d)
    i, j = o.get("l2i")[:o.lm].astype(int), o.get("l2j")[:o.lm].astype(int)
    cell = {"W": i == 2, "E": i == d["imm"] + 1, "S": j == 2, "N": j == d["jmm"] + 1}[side]
    for n in range(1, 13):
        assert o.step(n) == 0
        thrs = n * DT / 3600.0
        want = oracle_lib.parab(0.0, thrs, ser[:, 0], 600.0 / 3600.0)
        np.testing.assert_allclose(o.get("s")[:o.lm][cell], want, rtol=0, atol=1e-14)


def test_water_level_boundary_fills_lake(oracle_lib):
    """The lake fills from 0 and then seiches about the boundary level."""
    d = lake.channel(nx=12, ny=6, dt=30.0)
    d = lake.open_boundary(d, "W", 1, 2, 7, lake.steady_series(0.05, 20.0))
    o = oracle_lib.Oracle(d)
    level = []
    for n in range(1, 601):
        assert o.step(n) == 0
        level.append(float(o.get("s")[:o.lm].mean()))
    level = np.array(level)
    assert abs(level[200:].mean() - 0.05) < 0.002
    assert level[0] < 0.025 and level.max() > 0.05


def test_east_water_level_face_keeps_zero_ex(oracle_lib):
    d = lake.open_boundary(lake.channel(), "E", 1, 2, 13, _wse_series())
    o = oracle_lib.Oracle(d)
    _run(o, 3)
    o.phase(0, 4, 1, 4 * DT / 3600.0)
    o.phase(2, 4, 1, 4 * DT / 3600.0)
    i = o.get("l2i")[:o.lm].astype(int)
    ex = o.get("ex")[:o.lm]
    assert np.all(ex[i == d["imm"]] == 0.0)          # faces inside the east boundary
    assert np.any(ex[i == d["imm"] - 1] != 0.0)      # the next faces in are computed


def _cells(o):
    return o.get("l2i")[:o.lm].astype(int), o.get("l2j")[:o.lm].astype(int)


def _by_ij(o, name):
    """Field as a dict (i, j) -> column values."""
    i, j = _cells(o)
    f = o.get(name)[:o.lm]
    return {(a, b): f[n] for n, (a, b) in enumerate(zip(i, j))}


@pytest.mark.parametrize("btype", [2])
def test_mirror_x_west_east(oracle_lib, btype):
    base = lake.channel(nx=24, ny=10, iturb=1, ntr=1)
    dw = lake.open_boundary(base, "W", btype, 2, 11, lake.steady_series(Q, 20.0, [1.0]))
    de = lake.open_boundary(lake.mirror_x(base), "E", btype, 2, 11, lake.steady_series(-Q, 20.0, [1.0]))
    ow, oe = oracle_lib.Oracle(_tight(dw)), oracle_lib.Oracle(_tight(de))
    _run(ow, 12)
    _run(oe, 12)
    im = base["imm"] + 1
    for name, sign, shift in [("s", 1, 0), ("sal", 1, 0), ("wp", 1, 0), ("h", 1, 0), ("uh", -1, 1), ("vh", 1, 0),
                              ("q2", 1, 0), ("tracer1", 1, 0)]:
        fw, fe = _by_ij(ow, name), _by_ij(oe, name)
        scale = max(max(np.max(np.abs(v)) for v in fw.values()), 1e-30)
        worst = 0.0
        for (i, j), v in fw.items():
            ii = im + 2 - i - shift  # u-faces sit east of their column
            if (ii, j) not in fe:
                continue
            worst = max(worst, float(np.max(np.abs(sign * fe[(ii, j)] - v))))
        assert worst <= MIRROR_TOL * scale, (name, worst, scale)


@pytest.mark.parametrize("btype", [1, 2])
def test_mirror_y_north_south(oracle_lib, btype):
    base = lake.transpose_xy(lake.channel(nx=24, ny=10, iturb=1, ntr=1))
    val = (lambda s: _wse_series(tracer=1.0)) if btype == 1 else (lambda s: lake.steady_series(_inflow(s), 20.0, [1.0]))
    ds = lake.open_boundary(base, "S", btype, 2, 11, val("S"))
    dn = lake.open_boundary(lake.mirror_y(base), "N", btype, 2, 11, val("N"))
    os_, on = oracle_lib.Oracle(_tight(ds)), oracle_lib.Oracle(_tight(dn))
    _run(os_, 12)
    _run(on, 12)
    jm = base["jmm"] + 1
    for name, sign, shift in [("s", 1, 0), ("sal", 1, 0), ("wp", 1, 0), ("vh", -1, 1), ("uh", 1, 0), ("q2", 1, 0),
                              ("tracer1", 1, 0)]:
        fs, fn = _by_ij(os_, name), _by_ij(on, name)
        scale = max(max(np.max(np.abs(v)) for v in fs.values()), 1e-30)
        worst = 0.0
        for (i, j), v in fs.items():
            jj = jm + 2 - j - shift
            if (i, jj) not in fn:
                continue
            worst = max(worst, float(np.max(np.abs(sign * fn[(i, jj)] - v))))
        assert worst <= MIRROR_TOL * scale, (name, worst, scale)


@pytest.mark.parametrize("btype", [1, 2])
def test_transpose_west_south(oracle_lib, btype):
    base = lake.channel(nx=24, ny=10, iturb=1, ntr=1)
    val = _wse_series(tracer=1.0) if btype == 1 else lake.steady_series(Q, 20.0, [1.0])
    dw = lake.open_boundary(base, "W", btype, 2, 11, val)
    ds = lake.open_boundary(lake.transpose_xy(base), "S", btype, 2, 11, val)
    ow, os_ = oracle_lib.Oracle(_tight(dw)), oracle_lib.Oracle(_tight(ds))
    _run(ow, 12)
    _run(os_, 12)
    for nw, ns in [("s", "s"), ("sal", "sal"), ("wp", "wp"), ("uh", "vh"), ("vh", "uh"), ("q2", "q2"),
                   ("tracer1", "tracer1")]:
        fw, fs = _by_ij(ow, nw), _by_ij(os_, ns)
        scale = max(max(np.max(np.abs(v)) for v in fw.values()), 1e-30)
        worst = max(float(np.max(np.abs(fs[(j, i)] - v))) for (i, j), v in fw.items())
        assert worst <= MIRROR_TOL * scale, (nw, worst, scale)


def test_rejects_bad_boundaries(oracle_lib):
    d = lake.channel()
    with pytest.raises(RuntimeError):  # column 3 is not on the edge of the lake
        bad = lake.open_boundary(d, "W", 2, 2, 13, lake.steady_series(Q, 20.0))
        bad["openbc"]["isbc"] = [3]
        bad["openbc"]["iebc"] = [3]
        oracle_lib.Oracle(bad)
    with pytest.raises(RuntimeError):  # type 3 (submerged flow) is not restated
        bad = lake.open_boundary(d, "W", 3, 2, 13, lake.steady_series(Q, 20.0))
        oracle_lib.Oracle(bad)
