"""The met-station surface heat budget in the oracle (ifsurfbc 10 and 11:
surfbc0 4301-4450, surfbc 4596-4650, InitializeInterpolationMethods
5510-5646, CalculateMetResiduals 5650-5700, InterpMetData 5703-5748 and the
cases 10 / 11 of DistributeMomentumHeatSourcesH 5331-5483, all in
si3d_boundaryconditions.f90), pinned by what reduces to the already pinned
spatially uniform modes and by the interpolation's closed forms:

* stations that all report the same series (or a single station) give
  every column that series, so mode 11 must reproduce mode 3 and mode 10
  mode 2 with no cloud cover, for every interpolation scheme;
* the Barnes weights exp(-ln(100) d^2 / da^2) (Barnes 1964, with da the
  mean station distance), the inverse squared distance weights and the
  two-pass scheme (Koch et al. 1983) at chosen columns; a column half way
  between two stations gets their mean; at a station far from the others
  the two-pass scheme returns that station's value;
* the drag laws of icw 1 and 2 at chosen wind speeds."""
import math

import numpy as np
import pytest

from si3d_amd import lake

PI = 3.14159265358979


def _base(nts=12):
    d = lake.rectangle(nx=20, ny=10, dx=50.0, depth=8.0, dz=1.0, nts=nts)
    d.update(isal=1, ibc=1, iturb=1, ifsurfbc=20)
    d["init_profile"] = np.linspace(22.0, 18.0, lake.km1_of(d))
    d["surfbc"] = lake.wind_series(nts * d["dt"], 600.0, amp=0.15, period=86400.0, direction_deg=40.0)
    return d


def _uniform(d_met, mode):
    """The spatially uniform deck (mode 2 or 3) with station 0's series."""
    v = np.asarray(d_met["surfbc"])
    # eta Qsw Ta Pa Rh Cc|Qlw cdw uair vair
    cols = [v[:, 0], v[:, 2], v[:, 3], v[:, 1], v[:, 4], (np.zeros(len(v)) if mode == 2 else v[:, 5]),
            np.full(len(v), d_met["cw"]), v[:, 6], v[:, 7]]
    d = dict(d_met)
    for k in ("metxy", "nmetstat", "iinterp", "gammaB", "delNfactor", "icw"):
        d.pop(k, None)
    d.update(ifsurfbc=mode, surfbc=np.stack(cols, axis=1))
    return d


FIELDS = ["s", "uh", "vh", "wp", "sal", "HeatSource", "q2", "Av"]


@pytest.mark.parametrize("mode,uniform", [(11, 3), (10, 2)])
@pytest.mark.parametrize("iinterp,stations", [(1, [(3.0, 4.0), (15.5, 8.0), (9.0, 2.5)]),
                                              (2, [(3.0, 4.0), (15.5, 8.0), (9.0, 2.5)]),
                                              (3, [(3.0, 4.0), (15.5, 8.0)]),
                                              (2, [(7.0, 5.0)])])
def test_identical_stations_reproduce_the_uniform_mode(oracle_lib, mode, uniform, iinterp, stations):
    base = _base()
    dm = lake.met_stations(base, mode, stations, iinterp=iinterp, gammaB=0.3, delNfactor=1.2, same=True)
    du = _uniform(dm, uniform)
    a, b = oracle_lib.Oracle(dm), oracle_lib.Oracle(du)
    for n in range(1, 7):
        a.step(n)
        b.step(n)
    for f in FIELDS:
        x, y = a.get(f), b.get(f)
        scale = max(float(np.max(np.abs(y))), 1e-30)
        assert float(np.max(np.abs(x - y))) / scale < 1e-11, (f, float(np.max(np.abs(x - y))) / scale)
    assert float(np.max(np.abs(a.get("HeatSource")))) > 0.0


def _column_values(oracle_lib, deck, name):
    o = oracle_lib.Oracle(deck)
    o.phase(0, 1, 2, deck["dt"] / 3600.0)  # surfbc of step 1
    return o, o.get(name)


def _station_wind(deck, m, t_hours):
    v = np.asarray(deck["surfbc"])
    col = 2 + 6 * m + 4
    dt = deck["dtsurfbc"] / 3600.0
    # the series is smooth enough at the first record: step 1 lies within
    # the first interval, where parab is the quadratic through records 1-3
    from oracle.oracle import parab
    return parab(0.0, t_hours, v[:, col], dt)


@pytest.mark.parametrize("iinterp", [1, 3])
def test_interpolation_weights_closed_forms(oracle_lib, iinterp):
    base = _base()
    st = [(3.0, 4.0), (15.0, 8.0), (9.0, 2.0)]
    d = lake.met_stations(base, 11, st, iinterp=iinterp)
    o, ua = _column_values(oracle_lib, d, "uair")
    t = d["dt"] / 3600.0
    x = np.array([_station_wind(d, m, t) for m in range(3)])
    ii, jj = o.get("l2i").astype(int), o.get("l2j").astype(int)
    for l in range(0, o.lm, 7):
        dd = np.array([(sx - ii[l]) ** 2 + (sy - jj[l]) ** 2 for sx, sy in st])
        if iinterp == 1:
            da = np.mean(np.sqrt(dd))
            w = np.exp(-math.log(100.0) * dd / da ** 2)
        else:
            w = np.where(dd > 0, 1.0 / np.where(dd > 0, dd, 1.0), 1.0)
        assert ua[l] == pytest.approx(float(np.dot(w, x) / w.sum()), rel=1e-12, abs=1e-14), l


@pytest.mark.parametrize("iinterp", [1, 2, 3])
def test_column_between_two_stations_gets_their_mean(oracle_lib, iinterp):
    base = _base()
    st = [(4.0, 5.0), (14.0, 5.0)]  # column (9, 5) is half way
    d = lake.met_stations(base, 11, st, iinterp=iinterp, gammaB=0.3, delNfactor=1.0)
    o, ua = _column_values(oracle_lib, d, "uair")
    t = d["dt"] / 3600.0
    x = [_station_wind(d, m, t) for m in range(2)]
    ii, jj = o.get("l2i").astype(int), o.get("l2j").astype(int)
    l = int(np.nonzero((ii[:o.lm] == 9) & (jj[:o.lm] == 5))[0][0])
    assert ua[l] == pytest.approx(0.5 * (x[0] + x[1]), rel=1e-12)
    lo = int(np.nonzero((ii[:o.lm] == 4) & (jj[:o.lm] == 5))[0][0])
    assert abs(ua[lo] - x[0]) < abs(ua[lo] - x[1])  # nearer the first station's value


def test_two_pass_barnes_restores_a_remote_station(oracle_lib):
    """Two stations 16 cells apart with delNfactor 0.25: the first pass
    already gives the station's own cell 1 - 1e-10 of its weight; the second
    pass (weights to the power 1/gammaB) corrects the rest, so the column at
    a station returns that station's value, and the Koch smoothing length
    kappa = 3.5132 * 4/pi^2 * (delN da)^2 (gammaB 0.3) fixes the weight at a
    column in between."""
    base = _base()
    st = [(2.0, 5.0), (18.0, 5.0)]
    d = lake.met_stations(base, 11, st, iinterp=2, gammaB=0.3, delNfactor=0.25)
    o, ua = _column_values(oracle_lib, d, "uair")
    t = d["dt"] / 3600.0
    x = [_station_wind(d, m, t) for m in range(2)]
    ii, jj = o.get("l2i").astype(int), o.get("l2j").astype(int)
    at = lambda i, j: int(np.nonzero((ii[:o.lm] == i) & (jj[:o.lm] == j))[0][0])
    assert ua[at(2, 5)] == pytest.approx(x[0], rel=1e-9)
    assert ua[at(18, 5)] == pytest.approx(x[1], rel=1e-9)
    kappa = 3.5132 * 4.0 / PI ** 2 * (0.25 * 16.0) ** 2
    w = np.exp(-np.array([(6 - 2) ** 2, (18 - 6) ** 2]) / kappa)
    first = float(np.dot(w, x) / w.sum())
    wn = np.exp(-np.array([[0.0, 256.0], [256.0, 0.0]]) / kappa)
    resid = np.array(x) - wn @ np.array(x) / wn.sum(axis=1)
    g = w ** (1.0 / 0.3)
    assert ua[at(6, 5)] == pytest.approx(first + float(np.dot(g, resid) / g.sum()), rel=1e-12)


@pytest.mark.parametrize("icw", [1, 2])
def test_drag_laws(oracle_lib, icw):
    base = _base()
    d = lake.met_stations(base, 11, [(9.0, 5.0)], iinterp=1, icw=icw)
    o, cdw = _column_values(oracle_lib, d, "cdw")
    ua, va = o.get("uair"), o.get("vair")
    for l in range(0, o.lm, 11):
        ws = math.hypot(ua[l], va[l])
        if icw == 1:  # Amorocho & DeVries (1980): 1.04e-3 + 1.5e-3 / (1 + exp((12.5 - U) / 1.56))
            want = 0.00104 + 0.0015 / (1.0 + math.exp((12.5 - ws) / 1.56))
        else:  # Andreas et al. (2012) with the reference's (1/2) = 0: u* = 0.239 + 0.0433 (U - 8.271 + 1)
            want = ((0.239 + 0.0433 * (ws - 8.271 + 1.0)) / ws) ** 2
        assert cdw[l] == pytest.approx(want, rel=1e-13)
