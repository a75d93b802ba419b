"""Writers of pSi3D's output files for the Python driver -- node time files
(`outt`, si3d_utils.f90:1020-1202), horizontal plane files (`outh`,
2142-2334) and the others listed below -- from fields
copied back from the device (`GpuLake.get`, the library's sync_to_host).

With the Fortran driver these files come from the reference's own routines
after `s3g_get_field_*`; these writers give the same records to a run driven
from Python:

* tf<i>_<j>.txt: the title, a run line, the node/run parameters and two
  column-heading lines, then per output step one record of time (h), step,
  zeta (cm) and, for every layer k1..kmz+1, level depth (m), u, v, w (cm/s),
  Av, Dv (cm2/s) and the scalar (FORMAT 4: 1X,F10.4,I10,2PF9.2, then
  0PF9.2,2(2PF10.2),2PF9.4,2(4PF15.7),0PE15.5 per layer, continued on records
  that start with 30X).  Values of layers without water are -99 in output
  units, as the reference initialises them.
* plane_<k>: Fortran unformatted sequential records (4-byte length before and
  after each record): the frame count, the wet-cell count of the plane, then
  one record per frame: step, year, month, day (int32), hour (float32) and,
  per wet cell of the plane, (i, j, u, v, w, scalar, Av, s) on the first frame
  and (u, v, w, scalar, Av, s) with u/v/w averaged to the cell centre on the
  later ones (float32).

* si3d_out.txt (outr and the solver's lines), si3d_log.txt (outlog),
  section_<j> (outv), spacefile.xml with its empty binary companions (outs)
  and ScalarBalance.txt (OutScalarEnergyBalance): see their classes.

The run line of a time file carries the wall-clock date of the run, as the
reference's does; the model date of the plane records is the start date plus
n dt in exact calendar arithmetic (the reference steps its date in single
precision, which can differ in the last digit of the hundredths of an hour
after long runs).
"""
from __future__ import annotations

import datetime
import math
import os

import numpy as np

from . import lake


# ---- Fortran edit descriptors -------------------------------------------------
def fortran_f(x, w, d, scale=0):
    """Fw.d output of x with scale factor kP (x * 10**k), right-justified;
    asterisks when it does not fit."""
    v = float(x) * (10.0 ** scale)
    s = f"{v:.{d}f}" + ("." if d == 0 else "")
    if len(s) > w and s.startswith("0."):
        s = s[1:]
    elif len(s) > w and s.startswith("-0."):
        s = "-" + s[2:]
    return s.rjust(w) if len(s) <= w else "*" * w


def fortran_e(x, w, d):
    """Ew.d output with 0P: 0.ddddE+ee (mantissa in [0.1, 1))."""
    v = float(x)
    if v == 0.0 or not math.isfinite(v):
        s = f"0.{'0' * d}E+00" if v == 0.0 else ("NaN" if math.isnan(v) else ("-Inf" if v < 0 else "Inf"))
        return s.rjust(w)
    e = math.floor(math.log10(abs(v))) + 1
    m = round(abs(v) / 10.0 ** e, d)
    if m >= 1.0:
        m, e = m / 10.0, e + 1
    if m < 0.1:  # log10 rounding at an exact power of ten
        m, e = m * 10.0, e - 1
    mant = f"{m:.{d}f}"
    exp = f"E{'+' if e >= 0 else '-'}{abs(e):02d}" if abs(e) <= 99 else f"{'+' if e >= 0 else '-'}{abs(e):03d}"
    s = ("-" if v < 0 else "") + mant + exp
    if len(s) > w and s.startswith("-0."):  # the optional leading zero goes first
        s = "-" + s[2:]
    return s.rjust(w) if len(s) <= w else "*" * w


def hour_stamp(seconds_of_day):
    """The hour stamp of the time and plane files, ihr = NINT(hrs*100.) with
    hrs = isec/3600. in single precision (compute_date, si3d_utils.f90:
    3616-3618, 3627-3628).  NINT rounds halves away from zero: at 90 s
    (dt = 30 s, n = 3) hrs*100 = 2.5 and the stamp is 3."""
    v = np.float32(np.float32(int(seconds_of_day)) / np.float32(3600.0)) * np.float32(100.0)
    return int(math.floor(float(v) + 0.5))


def _layer_fields(z, u, v, w, av, dv, sc):
    return (fortran_f(z, 9, 2) + fortran_f(u, 10, 2, 2) + fortran_f(v, 10, 2, 2) + fortran_f(w, 9, 4, 2) +
            fortran_f(av, 15, 7, 4) + fortran_f(dv, 15, 7, 4) + fortran_e(sc, 15, 5))


# ---- grid helpers ---------------------------------------------------------------
def wet_columns(deck):
    """(i, j) of every wet column in the reference's l order (i outer, j inner;
    si3d_utils.f90:781-793), 1-based, from the FP32 depths of `bathy`."""
    h = deck["bathy_dm"].astype(np.float32) * np.float32(0.1) + np.float32(deck["datadj"])
    ii, jj = np.nonzero(h > 0)
    order = np.lexsort((jj, ii))
    return ii[order] + 2, jj[order] + 2


def west_south(deck):
    """0-based west and south neighbour of every wet column (lm, the dry
    sentinel, where the neighbour is dry), as bathy sets lWC / lSC
    (si3d_utils.f90:796-806)."""
    ii, jj = wet_columns(deck)
    lm = len(ii)
    index = {(int(i), int(j)): l for l, (i, j) in enumerate(zip(ii, jj))}
    lW = np.array([index.get((int(i) - 1, int(j)), lm) for i, j in zip(ii, jj)], dtype=np.int64)
    lS = np.array([index.get((int(i), int(j) - 1), lm) for i, j in zip(ii, jj)], dtype=np.int64)
    return lW, lS


def zlevels(deck):
    """zlevel(1..km1) as ZGridDimensions defines it (si3d_utils.f90:318-382)."""
    if deck["ibathyf"] < 0:
        return np.asarray(deck["zlevel"], dtype=np.float32).astype(np.float64)
    km1 = lake.km1_of(deck)
    z = np.zeros(km1, dtype=np.float32)
    dz = np.float32(deck["ddz"])
    for k in range(2, km1):  # zlevel(k1) = 0, then + ddz in single precision
        z[k] = z[k - 1] + dz
    z[0] = z[1] = -100.0
    return z.astype(np.float64)


# ---- time files ------------------------------------------------------------------
class TimeFiles:
    """tf<i>_<j>.txt for the nodes `nodes` [(i, j), ...] of a GpuLake run."""

    def __init__(self, g, deck, nodes, directory=".", title=None, sed_h=0.5):
        self.g, self.deck = g, deck
        self.dir = directory
        ii, jj = wet_columns(deck)
        self.l_of = {(int(i), int(j)): l for l, (i, j) in enumerate(zip(ii, jj))}
        self.nodes = []
        for (i, j) in nodes:
            if (i, j) not in self.l_of:
                raise ValueError(f"time-file node ({i}, {j}) is not a wet column")
            self.nodes.append((i, j, self.l_of[(i, j)]))
        self.z = zlevels(deck)
        self.sed_h = sed_h
        self.title = title if title is not None else deck.get("name", "si3d run")
        self.files = {}

    def _name(self, i, j):
        return os.path.join(self.dir, f"tf{i}_{j}.txt")

    def open(self):
        kmz = self.g.get("kmz")
        hhs = self.g.get("hhs")
        now = datetime.datetime.now()
        d = self.deck
        for i, j, l in self.nodes:
            f = open(self._name(i, j), "w")
            f.write(f"{self.title}\n")
            f.write(f"Run number = {now:%Y%m%d}{now:%H%M},  Start date of run:  "
                    f"{d.get('imon0', 1):2d}/{d.get('iday0', 1):2d}/{d.get('iyr0', 2000):4d} at "
                    f"{d.get('ihr0', 0):04d} hours\n")
            km = int(kmz[l]) - 2 + 1
            dtf = fortran_f(d["dt"], 5, 2) if d["dt"] >= 0.01 and d["ddz"] >= 0.01 else fortran_f(d["dt"], 5, 4)
            dzf = fortran_f(d["ddz"], 5, 2) if d["dt"] >= 0.01 and d["ddz"] >= 0.01 else fortran_f(d["ddz"], 5, 4)
            f.write(f"i ={i:4d}  j ={j:4d}  km ={km:4d}   dt ={dtf} sec  hhs ={fortran_f(hhs[l], 6, 3)} m"
                    f"   dz ={dzf} m\n")
            f.write(f"iexplt ={0:2d}  itrap ={int(d['itrap']):2d}  cd = {fortran_f(d['cd'], 7, 4)}  "
                    f"ismooth ={int(d['ismooth']):2d}  beta ={fortran_f(d['beta'], 6, 3)} niter ={int(d['niter']):2d}\n")
            f.write(f"iextrp ={0:2d}  f ={fortran_f(d['f'], 7, 4)}  tramp={fortran_f(0.0, 9, 1)}  "
                    f"iupwind ={0:2d}\n")
            f.write(" " + "   time     " + "  step     " + "  zeta " + "   depth   " + "    u       " +
                    "  v      " + " w       " + "   Av       " + "        Dv      " + "  scalar    " +
                    " Tracers-> " + "\n")
            f.write(" " + "    hrs     " + "   no      " + "   cm       " + " m    " + "   cm/s   " +
                    "   cm/s   " + "  cm/s      " + " cm2/s      " + "     cm2/s     " + "  oC        " +
                    "  g/l   -> " + "\n")
            self.files[(i, j)] = f

    def write(self, n, thrs):
        """One output record per node at step n, time thrs (hours)."""
        if not self.files:
            self.open()
        g = self.g
        s, u, v, wp, av, dv, sal, h = (g.get(x) for x in ("s", "u", "v", "wp", "Av", "Dv", "sal", "h"))
        kmz = g.get("kmz")
        for i, j, l in self.nodes:
            rows = []
            for k in range(2, int(kmz[l]) + 2):  # k = k1 .. kmz+1 (1-based)
                kk = k - 1
                if h[l, kk] <= 0.0:
                    rows.append((0.0, -99e-2, -99e-2, -99e-2, -99e-4, -99e-4, -99.0))
                    continue
                zexp = self.z[k] if k < len(self.z) else self.z[k - 1]
                if k == int(kmz[l]) + 1:
                    zexp = self.z[k - 1] + self.sed_h
                rows.append((zexp, u[l, kk], v[l, kk], wp[l, kk], 0.5 * (av[l, kk] + av[l, kk + 1]),
                             0.5 * (dv[l, kk] + dv[l, kk + 1]), sal[l, kk]))
            head = " " + fortran_f(thrs, 10, 4) + f"{n:10d}" + fortran_f(s[l], 9, 2, 2)
            out = [head + _layer_fields(*rows[0])]
            out += [" " * 30 + _layer_fields(*r) for r in rows[1:]]
            self.files[(i, j)].write("\n".join(out) + "\n")

    def close(self):
        for f in self.files.values():
            f.close()
        self.files = {}


# ---- plane files -------------------------------------------------------------------
def _record(f, *parts):
    payload = b"".join(np.ascontiguousarray(p).tobytes() for p in parts)
    n = np.array([len(payload)], dtype=np.int32).tobytes()
    f.write(n + payload + n)


class PlaneFiles:
    """plane_<k> files for the layers `planes` of a GpuLake run (outh)."""

    def __init__(self, g, deck, planes, nts, iop, itspfh=0, directory=".", start=(2000, 1, 1, 0)):
        self.g, self.deck = g, deck
        self.planes = list(planes)
        self.km = lake.km1_of(deck) - 1
        for kp in self.planes:
            if kp < 2:
                raise ValueError("ERROR: Output plane is not in DOMAIN")
        self.n_frames = int((nts - itspfh) / max(iop, 1))
        self.dir = directory
        self.start = datetime.datetime(start[0], start[1], start[2]) + datetime.timedelta(hours=start[3] / 100.0)
        ii, jj = wet_columns(deck)
        self.ij = (ii, jj)
        self.files = []

    def _cells(self, kp, kmz):
        lm = len(self.ij[0])
        if kp > self.km:
            return np.arange(lm), kmz[:lm].astype(int)
        sel = np.nonzero(kmz[:lm] >= kp)[0]
        return sel, np.full(len(sel), kp)

    def _stamp(self, n):
        t = self.start + datetime.timedelta(seconds=n * self.deck["dt"])
        hrs = (t - datetime.datetime(t.year, t.month, t.day)).total_seconds() / 3600.0
        return (np.array([n, t.year, t.month, t.day], dtype=np.int32), np.array([hour_stamp(hrs * 3600.0)], dtype=np.float32))

    def write(self, n):
        g = self.g
        kmz = g.get("kmz")
        up, vp, wp, salp, hp, av, s = (g.get(x) for x in ("up", "vp", "wp", "salp", "hp", "Av", "s"))
        first = not self.files
        if first:
            for kp in self.planes:
                name = os.path.join(self.dir, f"plane_{kp}")
                f = open(name, "wb")
                _record(f, np.array([self.n_frames], dtype=np.int32))
                _record(f, np.array([len(self._cells(kp, kmz)[0])], dtype=np.int32))
                self.files.append(f)
        ints, hour = self._stamp(n)
        for f, kp in zip(self.files, self.planes):
            cells, kb = self._cells(kp, kmz)
            kk = kb - 1
            sal = salp[cells, kk].copy()
            sal[hp[cells, kk] <= 0.0] = -99.0
            avm = 0.5 * (av[cells, kk] + av[cells, kk + 1])
            if first:
                cols = [self.ij[0][cells].astype(np.float64), self.ij[1][cells].astype(np.float64),
                        up[cells, kk], vp[cells, kk], wp[cells, kk], sal, avm, s[cells]]
            else:
                west, south = self._neighbours()
                cols = [(up[cells, kk] + up[west[cells], kk]) / 2.0, (vp[cells, kk] + vp[south[cells], kk]) / 2.0,
                        (wp[cells, kk] + wp[cells, kk + 1]) / 2.0, sal, avm, s[cells]]
            data = np.stack(cols, axis=1).astype(np.float32)  # (ipoints, nvar), written row by row
            _record(f, ints, hour, data)

    def _neighbours(self):
        """West and south neighbour (l index, dry -> the sentinel lm) of each wet
        column, from the (i, j) positions."""
        if not hasattr(self, "_nb"):
            ii, jj = self.ij
            lm = len(ii)
            pos = {(int(i), int(j)): l for l, (i, j) in enumerate(zip(ii, jj))}
            west = np.array([pos.get((int(i) - 1, int(j)), lm) for i, j in zip(ii, jj)] + [lm])
            south = np.array([pos.get((int(i), int(j) - 1), lm) for i, j in zip(ii, jj)] + [lm])
            self._nb = (west, south)
        return self._nb

    def close(self):
        for f in self.files:
            f.close()
        self.files = []


# ---- 3-D file for particle tracking ---------------------------------------------
class ThreeDFile:
    """si3d_3D (outp, si3d_utils.f90:2484-2602): every wet column in l order
    (the reference's c loop over (i, j) visits the wet cells in the same order,
    si3d_utils.f90:780-793), layers k1 = 2 .. kmz + 1.  Fortran unformatted
    sequential records: the frame count, the point count, then per frame
    (n, year, month, day, hour) and, point by point, the first frame's
    (i, j, k, hp, up, vp, wp, salp[, Dv, q2p, q2lp, kh, Av]) or a later
    frame's (hp, up, vp, wp, salp[, Dv, q2p, q2lp, kh, Av]) as REAL*4."""

    K1 = 2

    def __init__(self, g, deck, nts, ipxml, itspf=0, turb_vars=False, directory=".", start=(2000, 1, 1, 0)):
        self.g, self.deck = g, deck
        self.n_frames = int((nts - itspf) / max(abs(ipxml), 1))
        self.turb = bool(turb_vars)
        self.path = os.path.join(directory, "si3d_3D")
        self.start = datetime.datetime(start[0], start[1], start[2]) + datetime.timedelta(hours=start[3] / 100.0)
        self.ij = wet_columns(deck)
        self.f = None

    def _points(self, kmz):
        lm = len(self.ij[0])
        nk = kmz[:lm].astype(int) + 1 - self.K1 + 1  # layers K1 .. kmz + 1
        cols = np.repeat(np.arange(lm), nk)
        ks = np.concatenate([np.arange(self.K1, k + 2) for k in kmz[:lm].astype(int)]) if lm else np.zeros(0, int)
        return cols, ks

    def write(self, n):
        g = self.g
        kmz = g.get("kmz")
        cols, ks = self._points(kmz)
        names = ["hp", "up", "vp", "wp", "salp"] + (["Dv", "q2p", "q2lp", "kh", "Av"] if self.turb else [])
        vals = [g.get(x)[cols, ks - 1] for x in names]
        first = self.f is None
        if first:
            self.f = open(self.path, "wb")
            _record(self.f, np.array([self.n_frames], dtype=np.int32))
            _record(self.f, np.array([len(cols)], dtype=np.int32))
            ii, jj = self.ij
            vals = [ii[cols].astype(np.float64), jj[cols].astype(np.float64), ks.astype(np.float64)] + vals
        t = self.start + datetime.timedelta(seconds=n * self.deck["dt"])
        hrs = (t - datetime.datetime(t.year, t.month, t.day)).total_seconds() / 3600.0
        ints = np.array([n, t.year, t.month, t.day], dtype=np.int32)
        data = np.stack(vals, axis=1).astype(np.float32)  # (ipoints, nvar), written point by point
        _record(self.f, ints, np.array([hour_stamp(hrs * 3600.0)], dtype=np.float32), data)

    def close(self):
        if self.f is not None:
            self.f.close()
            self.f = None


# ---- ScalarBalance.txt ---------------------------------------------------------
def density(temperature, elevation, salinity=0.00004):
    """Water density (kg/m3) of densty_s (si3d_utils.f90:3865-3949), in FP64
    and element-wise over arrays: the UNESCO / Gill (1982) fresh-water
    polynomial with the pressure of `elevation` metres of water found by
    fixed-point iteration until the update is at most 1e-6.  salinity**(3/2)
    is the integer power salinity**1 there (3/2 = 1 in Fortran)."""
    t, z = np.broadcast_arrays(np.asarray(temperature, dtype=np.float64), np.asarray(elevation, dtype=np.float64))
    s = salinity
    t2, t3, t4 = t * t, t * t * t, t * t * t * t
    densw = (999.842594 + 6.793952e-2 * t - 9.095290e-3 * t2 + 1.001685e-4 * t3 - 1.120083e-6 * t4
             + 6.536332e-9 * t4 * t)
    densws = (densw + s * (0.824493 - 4.0899e-3 * t + 7.6438e-5 * t2 - 8.2467e-7 * t3 + 5.3875e-9 * t4)
              + s * (-5.72466e-3 + 1.0227e-4 * t - 1.6546e-6 * t2) + 4.8314e-4 * s * s)
    kw = 19652.21 + 148.4206 * t - 2.327105 * t2 + 1.360477e-2 * t3 - 5.155288e-5 * t4
    ks = kw + s * (54.6746 - 0.603459 * t + 1.09987e-2 * t2 - 6.1670e-5 * t3) + s * (7.944e-2 + 1.6483e-2 * t
                                                                                       - 5.3009e-4 * t2)
    rho = np.array(densws, dtype=np.float64, copy=True)
    active = np.ones(rho.shape, dtype=bool)
    for _ in range(10000):
        p = 1e-5 * rho * 9.806 * z
        k = (ks + p * (3.239908 + 1.43713e-3 * t + 1.16092e-4 * t2 - 5.77905e-7 * t3)
             + p * s * (2.2838e-3 - 1.0981e-5 * t - 1.6078e-6 * t2) + 1.91075e-4 * p * s
             + p * p * (8.50935e-5 - 6.12293e-6 * t + 5.2787e-8 * t2)
             + p * p * s * (-9.9348e-7 + 2.0816e-8 * t + 9.1697e-10 * t2))
        new = densws / (1.0 - p / k)
        done = np.abs(new - rho) <= 1e-6
        rho = np.where(active, new, rho)
        active &= ~done
        if not active.any():
            break
    return rho


class ScalarBalance:
    """ScalarBalance.txt as OutScalarEnergyBalance writes it (si3d_utils.f90:
    3698-3814), one record (9E20.11) per call: time (h), heat content
    sum(sal h), the last tracer's content, the basin's potential and kinetic
    energy (densty_s of the p-level scalar at each layer's mid depth), and the
    closure's TKinE, ShearCum, BuocyCum, DisspCum from s3g_energy_balance,
    whose integrals are cleared after each record as the reference does.
    The driver calls it at n = 0 and every ipt steps (si3d.f90:107, 442).
    Sums run in FP64 (the reference's REAL*8 and REAL mix)."""

    def __init__(self, g, deck, directory="."):
        self.g, self.deck = g, deck
        self.path = os.path.join(directory, "ScalarBalance.txt")
        self.f = open(self.path, "w")
        self.turb = deck.get("iturb", 0) >= 1
        self.ntr = 0 if deck.get("tracer_profiles") is None else len(deck["tracer_profiles"])
        self.lW, self.lS = west_south(deck)

    def values(self, n, thrs):
        g, d = self.g, self.deck
        lm = g.lm
        G = 9.806
        kmz = g.get("kmz")[:lm].astype(int)
        k1z = g.get("k1z")[:lm].astype(int)
        lW, lS = self.lW, self.lS
        sal, h, hp, salp = g.get("sal"), g.get("h"), g.get("hp"), g.get("salp")
        up, vp, wp = g.get("up"), g.get("vp"), g.get("wp")
        hhs = g.get("hhs")
        z = np.asarray(zlevels(d), dtype=np.float64)
        tr = g.get(f"tracer{self.ntr}") if self.ntr else None
        heat = oxyg = pot = kin = 0.0
        for l in range(lm):
            ks, k1 = kmz[l], k1z[l]
            rng = slice(k1 - 1, ks)
            heat += float(np.dot(sal[l, rng], h[l, rng]))
            if tr is not None:
                oxyg += float(np.dot(tr[l, rng], h[l, rng]))
            ks_ = np.arange(ks, k1 - 1, -1)                        # kms, kms-1, ..., k1s (1-based)
            zmid = np.where(z[ks_ - 1] == -100.0, 0.0, z[ks_ - 1]) + 0.5 * hp[l, ks_ - 1]
            rho = density(salp[l, ks_ - 1], zmid) + 1000.0
            elev = np.empty(len(ks_))
            elev[0] = hhs[l] - h[l, ks - 1] / 2.0
            if len(ks_) > 1:
                elev[1:] = elev[0] + np.cumsum((hp[l, ks_[1:]] + hp[l, ks_[1:] - 1]) / 2.0)
            hw = np.concatenate([[h[l, ks - 1]], hp[l, ks_[1:] - 1]])
            pot += float(np.sum(rho * G * elev * hw))
            uu = (up[l, ks_ - 1] + up[lW[l], ks_ - 1]) / 2.0
            vv = (vp[l, ks_ - 1] + vp[lS[l], ks_ - 1]) / 2.0
            ww = (wp[l, ks_ - 1] + wp[l, ks_]) / 2.0
            kin += float(np.sum(0.5 * rho * (uu * uu + vv * vv + ww * ww) * h[l, ks_ - 1]))
        pot *= d["dx"] * d["dy"]
        kin *= d["dx"] * d["dy"]
        tke = g.energy_balance(reset=True) if self.turb else (0.0, 0.0, 0.0, 0.0)
        return [thrs, heat, oxyg, pot, kin] + list(tke)

    def write(self, n, thrs):
        self.f.write("".join(fortran_e(v, 20, 11) for v in self.values(n, thrs)) + "\n")
        self.f.flush()

    def close(self):
        if self.f is not None:
            self.f.close()
            self.f = None


# ---- si3d_out.txt, si3d_log.txt ----------------------------------------------------
def fortran_g(x, w, d):
    """Gw.d output of a REAL: F editing with (w-4).(d-k) and four trailing
    blanks when 0.1 <= |x| (rounded to d significant digits) < 10**d, k its
    number of integer digits; Ew.d otherwise."""
    v = float(x)
    a = abs(v)
    if v == 0.0:
        return (fortran_f(0.0, w - 4, d - 1) + " " * 4) if d >= 1 else fortran_e(v, w, d)
    if 0.1 - 0.5 * 10.0 ** (-d - 1) <= a < 10.0 ** d - 0.5:
        k = 0
        while k < d and a >= 10.0 ** k - 0.5 * 10.0 ** (k - d):
            k += 1
        return fortran_f(v, w - 4, d - k) + " " * 4
    return fortran_e(v, w, d)


def run_parameters(deck, title=None, start=None):
    """The values of si3d_inp.txt that the run log prints (input,
    si3d_utils.f90:36-51): from the deck's own si3d_inp.txt when it was read
    from one (`inputs.read_case`), else derived from the deck as
    `inputs.write_inp` writes it."""
    p = dict(deck.get("inp") or {})
    if not p:
        imm, jmm = int(deck["imm"]), int(deck["jmm"])
        p = dict(xl=imm * deck["dx"], yl=jmm * deck["dy"], zl=deck["zl"], tl=deck["nts"] * deck["dt"],
                 idx=deck["dx"], idy=deck["dy"], idz=deck["ddz"], dzmin=deck["dzmin"], zetainit=deck["zetainit"],
                 idt=deck["dt"], itrsca=4, idbg=0,
                 title=title or "si3d case written by si3d_amd")
        for k in ("itrap", "niter", "ismooth", "beta", "iturb", "Av0", "Dv0", "iadv", "itrmom", "ihd", "Ax0", "Ay0",
                  "f", "theta", "ibc", "isal", "cd", "ifsurfbc", "dtsurfbc", "cw", "wa", "phi"):
            p[k] = deck.get(k, 0)
        st = start or (2000, 1, 1, 0)
        p["iyr0"], p["imon0"], p["iday0"], p["ihr0"] = st
    ob = deck.get("openbc")
    p["nopen"] = len(ob["iside"]) if ob is not None else 0
    if title:
        p["title"] = title
    return p


class RunLog:
    """si3d_out.txt: the run header of outr (si3d_utils.f90:944-1017) -- title,
    run number (wall-clock date and time) and start date, the domain and
    algorithm parameters -- and, per free-surface solve, the lines
    SolverSparse writes around its NSPCG call (si3d_procedures.f90:1997-2002)
    with the solver's own line between them at message level 1
    (nspcg.f:8420-8421): "cg  has converged in N iterations", N being this
    solver's iteration count (multigrid-preconditioned CG, not NSPCG's
    MIC(0) CG), and the error line when the solve stops at its iteration
    limit (ier = 1)."""

    def __init__(self, deck, directory=".", title=None, start=None, now=None):
        self.p = run_parameters(deck, title, start)
        self.path = os.path.join(directory, "si3d_out.txt")
        self.f = open(self.path, "w")
        self.f.write(self.header(now or datetime.datetime.now()))
        self.f.flush()

    def header(self, now):
        p = self.p
        date, time = now.strftime("%Y%m%d"), now.strftime("%H%M%S")
        L = [str(p["title"]).rstrip() + "\n",
             f"Run number = {date:8.8}{time[:4]},  Start date of run:  {int(p['imon0']):2d}/{int(p['iday0']):2d}/"
             f"{int(p['iyr0']):4d} at {int(p['ihr0']):04d} hours\n", "\n"]
        F, G = fortran_f, fortran_g
        L += [f" xl=   {F(p['xl'], 9, 1)}\n", f" yl=   {F(p['yl'], 9, 1)}\n", f" zl=   {F(p['zl'], 9, 3)}\n",
              f" idx=  {F(p['idx'], 9, 3)}\n", f" idy=  {F(p['idy'], 9, 3)}\n", f" idz=  {F(p['idz'], 9, 3)}\n",
              f" tl=   {G(p['tl'], 9, 2)}\n", f" idt=  {G(p['idt'], 9, 2)}\n", f" dzmin={F(p['dzmin'], 9, 2)}\n",
              f" zeta0={F(p['zetainit'], 9, 2)}\n", "\n", "\n"]
        rows = [("itrap= ", "itrap", "I"), ("niter= ", "niter", "I"), ("smooth=", "ismooth", "I"),
                ("beta=  ", "beta", "F4"), ("iturb= ", "iturb", "I"), ("Av0=   ", "Av0", "F4"),
                ("Dv0=   ", "Dv0", "F4"), ("iadv=  ", "iadv", "I"), ("trmom= ", "itrmom", "I"),
                ("ihd=   ", "ihd", "I"), ("Ax0=   ", "Ax0", "F4"), ("Ay0=   ", "Ay0", "F4"), ("f=     ", "f", "F4"),
                ("theta= ", "theta", "F4"), ("ibc=   ", "ibc", "I"), ("isal=  ", "isal", "I"),
                ("trsal= ", "itrsca", "I"), ("nopen= ", "nopen", "I"), ("cd=    ", "cd", "F4"),
                ("isbc=  ", "ifsurfbc", "I"), ("tsbc=  ", "dtsurfbc", "F2"), ("cw=    ", "cw", "F2"),
                ("wa=    ", "wa", "F2"), ("phi=   ", "phi", "F2"), ("idbg=  ", "idbg", "I")]
        for lab, k, e in rows:
            v = p.get(k, 0)
            s = f"{int(v):9d}" if e == "I" else F(v, 9, int(e[1]))
            L.append(f" {lab}{s}\n")
        L.append("\n")
        return "".join(L)

    def solve(self, n, iterations, ier=0):
        self.f.write(f"**Enter nspcg   n = {n:6d}\n")
        if ier == 0:
            self.f.write(f"\n cg  has converged in {iterations:5d} iterations\n")
        self.f.write(f"**Exit nspcg    n = {n:6d}\n")
        if ier != 0:
            self.f.write(f"**ERROR{ier:5d} from nspcg\n")
        self.f.flush()

    def close(self):
        if self.f is not None:
            self.f.close()
            self.f = None


class LogFile:
    """si3d_log.txt (outlog, si3d_utils.f90:3450-3483): opened, written and
    closed on every step, so it holds the latest " time=  hhh.hhhh hours
    n= nnnnnn" line while the run goes on, and deleted after the last step
    (n = nts), as the reference does."""

    def __init__(self, nts, directory="."):
        self.nts = int(nts)
        self.path = os.path.join(directory, "si3d_log.txt")

    def write(self, n, thrs):
        with open(self.path, "w") as f:
            f.write(f" time={fortran_f(np.float32(thrs), 10, 4)} hours    n={n:6d}\n")
        if n >= self.nts:
            os.remove(self.path)


def step_logged(g, deck, n, log):
    """One time step called phase by phase as si3d.f90:324-393 does (and
    s3g_step: step 1 is trapezoidal only, with at least two iterations; a
    leapfrog-only step when itrap = 0 or n is not a multiple of itrap), each
    free-surface solve reported to `log` (a RunLog)."""
    thrs = n * deck["dt"] / 3600.0
    niter = int(deck.get("niter", 2))
    itrap = int(deck.get("itrap", 1))

    def fd(istep, it, last):
        ier = g.fd(n, istep, lastiter=last, thrs=thrs, iter=it)
        log.solve(n, g.stats()["pcg_iters_last"], int(ier or 0))

    if n == 1:
        niter = max(niter, 2)
    else:
        lf_only = itrap == 0 or n % max(itrap, 1) != 0
        fd(1, 0, 1 if lf_only else 0)
        if lf_only:
            g.save(1)
            return
        g.settrap()
    for it in range(1, niter + 1):
        fd(2, it, 1 if it == niter else 0)
        if it < niter:
            g.settrap2()
    g.save(2)


# ---- cross-section files ------------------------------------------------------------
class SectionFiles:
    """section_<j> files (outv, si3d_utils.f90:1313-1463) for the vertical
    sections of si3d_inp.txt (a list of (i nodes, j nodes) per section).
    Fortran unformatted sequential records: the frame count, then per
    section the point count (wet cells of its columns, layers k1..kmz: kmz-1
    per column, CountSectionCells 1464-1498) and km1; per frame the
    time stamp (n, year, month, day, hour) and per point (i, j, node, depth
    of the layer's top, uhp, vhp, wp, salp, a tracer or the interface mean
    of Av, the interface mean of Dv) on the first frame, the last six on the
    later ones; salp (and the tracer) is -99 where hp <= 0.  The first frame
    takes the n-1 level tracer (tracerpp) and later frames the current one,
    as the reference does."""

    def __init__(self, g, deck, sections, nts, iox, directory=".", start=(2000, 1, 1, 0)):
        self.g, self.deck = g, deck
        self.sections = [(list(map(int, xs)), list(map(int, ys))) for xs, ys in sections]
        self.n_frames = int(nts / max(iox, 1))
        self.dir = directory
        self.start = datetime.datetime(start[0], start[1], start[2]) + datetime.timedelta(hours=start[3] / 100.0)
        ii, jj = wet_columns(deck)
        self.pos = {(int(i), int(j)): l for l, (i, j) in enumerate(zip(ii, jj))}
        self.ntr = 0 if deck.get("tracer_profiles") is None else len(deck["tracer_profiles"])
        self.files = []

    def _points(self, s, kmz):
        xs, ys = self.sections[s]
        rows = []
        for m, (i, j) in enumerate(zip(xs, ys), start=1):
            l = self.pos.get((i, j))
            if l is None:  # a dry cell of the section is skipped (and warned about)
                continue
            for k in range(2, int(kmz[l]) + 1):
                rows.append((i, j, m, l, k))
        return rows

    def write(self, n):
        g = self.g
        kmz = g.get("kmz")
        f = {x: g.get(x) for x in ("uhp", "vhp", "wp", "salp", "hp", "Av", "Dv")}
        z = np.asarray(zlevels(self.deck), dtype=np.float64)
        first = not self.files
        if self.ntr:
            trac = g.get(f"tracerpp{self.ntr}" if first else f"tracer{self.ntr}")
        if first:
            for s in range(len(self.sections)):
                fh = open(os.path.join(self.dir, f"section_{s + 1}"), "wb")
                _record(fh, np.array([self.n_frames], dtype=np.int32))
                _record(fh, np.array([len(self._points(s, kmz)), lake.km1_of(self.deck)], dtype=np.int32))
                self.files.append(fh)
        t = self.start + datetime.timedelta(seconds=n * self.deck["dt"])
        hrs = (t - datetime.datetime(t.year, t.month, t.day)).total_seconds() / 3600.0
        ints = np.array([n, t.year, t.month, t.day], dtype=np.int32)
        hour = np.array([hour_stamp(hrs * 3600.0)], dtype=np.float32)
        for s, fh in enumerate(self.files):
            out = []
            for i, j, m, l, k in self._points(s, kmz):
                kk = k - 1
                dry = f["hp"][l, kk] <= 0.0
                sal = -99.0 if dry else f["salp"][l, kk]
                if self.ntr:
                    c9 = -99.0 if dry else trac[l, kk]
                else:
                    c9 = 0.5 * (f["Av"][l, kk] + f["Av"][l, kk + 1])
                c10 = 0.5 * (f["Dv"][l, kk] + f["Dv"][l, kk + 1])
                vals = [f["uhp"][l, kk], f["vhp"][l, kk], f["wp"][l, kk], sal, c9, c10]
                if first:
                    vals = [float(i), float(j), float(m), z[k] - f["hp"][l, kk]] + vals
                out.append(vals)
            data = np.asarray(out, dtype=np.float32).reshape(-1, 10 if first else 6)
            _record(fh, ints, hour, data)

    def close(self):
        for fh in self.files:
            fh.close()
        self.files = []


# ---- spacefile.xml ------------------------------------------------------------------
class SpaceFileXml:
    """spacefile.xml with its binary companions spacefile2d.bin and
    spacefile3d.bin (outs, si3d_utils.f90:2605-2760; called every |ipxml|
    steps when ipxml < 0): on the first call, the xml header for the Gr
    viewer -- run parameters, title, run number (wall-clock date), the date of
    the first snapshot (the start date plus itspf rounded up to a multiple of
    ipxml dt) and its time step ipxml dt (negative, as the reference's
    INTEGER product of a negative ipxml), and the 2-D (zeta) and 3-D (u, v,
    w, dz) series descriptions.  The reference never writes records into the
    two binary files (outs_bin, 2763-2829, is not called), so they are
    created empty.  Layers of uniform thickness only (ibathyf >= 0), as the
    reference requires."""

    def __init__(self, deck, ipxml, itspf=0, directory=".", title=None, start=(2000, 1, 1, 0), now=None):
        if deck.get("ibathyf", 0) < 0:
            raise ValueError("STOP - Output to xml files only allowed for ibathyf > 0")
        self.deck, self.ipxml, self.itspf = deck, int(ipxml), int(itspf)
        self.dir, self.start, self.now = directory, start, now
        self.p = run_parameters(deck, title, start)
        self.written = False

    def text(self, now):
        d, p = self.deck, self.p
        idt = float(d["dt"])
        im, jm = int(d["imm"]) + 1, int(d["jmm"]) + 1
        F = fortran_f
        idt2 = int(self.ipxml * idt)
        t0 = datetime.datetime(*self.start[:3]) + datetime.timedelta(hours=self.start[3] / 100.0)
        if self.itspf == 0:
            first = 0
        else:
            first = int(self.itspf / idt2) * idt2
            if self.itspf % idt2 != 0:
                first += idt2
        t = t0 + datetime.timedelta(seconds=first)
        isec = int((t - datetime.datetime(t.year, t.month, t.day)).total_seconds())
        hh, mm, ss = isec // 3600, (isec % 3600) // 60, isec % 60
        ch = '"' + f"{idt2:8d}" .strip() + '"'
        L = ['<?xml version="1.0"?>', " ", "<!-- Si3D Output for Gr -->", "<!-- xml header file for binary spacefiles -->",
             f"<!-- Run Parameters:  dt ={F(idt, 5, 2)} s  dz ={F(d['ddz'], 5, 2)} m  im ={im:5d}  jm ={jm:5d}",
             f"     itrsca ={int(p.get('itrsca', 4)):2d}  itrap ={int(d.get('itrap', 1)):2d}  cd = {F(d['cd'], 7, 4)}"
             f"  ismooth ={int(d.get('ismooth', 0)):2d}  beta ={F(d.get('beta', 0.0), 6, 3)}"
             f"  niter ={int(d.get('niter', 2)):2d}",
             f"     itrmom ={int(d.get('itrmom', 1)):2d}  f ={F(d.get('f', 0.0), 7, 4)}  wa ={F(d.get('wa', 0.0), 8, 1)}"
             " -->", " ", "<gov.usgs.gr>", '   <obj class="gov.usgs.sfhydro.si3d.Si3dOutput"',
             f'        runTitle="{str(p["title"]).strip()}"', f'        runNumber="{now.strftime("%Y%m%d%H%M")}"',
             f'        startTime="{t.year:4d}/{t.month:02d}/{t.day:02d} {hh:02d}:{mm:02d}:{ss:02d}"'
             f" timeStepInSeconds={ch}>", " ",
             '      <obj class="gov.usgs.gr.GrObjectContainer" title="2-D x-y Variables">',
             '         <dataseries class="gov.usgs.gr.data.RafDataSeries"', '            file="spacefile2d.bin"',
             '            title="2-D Variables" numDimensions="1">',
             '            <dim num="0" title="Zeta" units="meters" unitScale="0.0001"/>', " ",
             "         </dataseries>", "      </obj>", " ",
             '      <obj class="gov.usgs.gr.GrObjectContainer" title="3-D Variables">',
             '         <dataseries class="gov.usgs.gr.data.RafDataSeries"', '            file="spacefile3d.bin"',
             '            title="3-D Variables" numDimensions="4">',
             '            <dim num="0" title="u" units="m/s" unitScale="0.0001"/>',
             '            <dim num="1" title="v" units="m/s" unitScale="0.0001"/>',
             '            <dim num="2" title="w" units="m/s" unitScale="0.00001"/>',
             '            <dim num="3" title="dz" units="m**2/s" unitScale="0.0001"/>', " ",
             "         </dataseries>", "      </obj>", "   </obj>", "</gov.usgs.gr>", " "]
        return "\n".join(L) + "\n"

    def write(self, n):
        if self.written:
            return
        self.written = True
        with open(os.path.join(self.dir, "spacefile.xml"), "w") as f:
            f.write(self.text(self.now or datetime.datetime.now()))
        for name in ("spacefile2d.bin", "spacefile3d.bin"):
            open(os.path.join(self.dir, name), "wb").close()
