"""ctypes wrapper of the CPU oracle (oracle/si3d_oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests, the smoke check and the
cpu_baseline leg of bench.py to check or time the reference algorithm.  The
product package never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "libsi3d_oracle.so")


class OrcDeck(C.Structure):
    _fields_ = [
        ("imm", C.c_int), ("jmm", C.c_int), ("ibathyf", C.c_int), ("km1_file", C.c_int),
        ("dx", C.c_double), ("dy", C.c_double), ("ddz", C.c_double), ("dzmin", C.c_double),
        ("datadj", C.c_double), ("zetainit", C.c_double), ("zl", C.c_double), ("dt", C.c_double),
        ("itrap", C.c_int), ("niter", C.c_int), ("ismooth", C.c_int), ("iturb", C.c_int),
        ("iadv", C.c_int), ("itrmom", C.c_int), ("ihd", C.c_int), ("ibc", C.c_int),
        ("isal", C.c_int), ("ifsurfbc", C.c_int),
        ("beta", C.c_double), ("Av0", C.c_double), ("Dv0", C.c_double), ("Ax0", C.c_double),
        ("Ay0", C.c_double), ("f", C.c_double), ("theta", C.c_double), ("cd", C.c_double),
        ("dtsurfbc", C.c_double), ("cw", C.c_double), ("wa", C.c_double), ("phi", C.c_double),
        ("zlevel_file", C.POINTER(C.c_double)), ("bathy_dm", C.POINTER(C.c_int)),
        ("init_profile", C.POINTER(C.c_double)), ("n_surfbc", C.c_int),
        ("surfbc", C.POINTER(C.c_double)), ("pcg_tol", C.c_double), ("pcg_maxit", C.c_int),
        ("ntr", C.c_int), ("tracer_profiles", C.POINTER(C.c_double)),
        ("nopen", C.c_int), ("dtsec_openbc", C.c_double), ("openbc", C.POINTER(C.c_int)),
        ("openbc_npts", C.c_int), ("openbc_values", C.POINTER(C.c_double)),
        ("icw", C.c_int), ("nmetstat", C.c_int), ("metxy", C.POINTER(C.c_double)), ("iinterp", C.c_int),
        ("gammaB", C.c_double), ("delNfactor", C.c_double),
    ]


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        L.orc_create.restype = C.c_void_p
        L.orc_create.argtypes = [C.POINTER(OrcDeck)]
        L.orc_destroy.argtypes = [C.c_void_p]
        L.orc_dims.argtypes = [C.c_void_p, C.POINTER(C.c_int)]
        L.orc_step.argtypes = [C.c_void_p, C.c_int]
        L.orc_fd.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_double]
        L.orc_settrap.argtypes = [C.c_void_p]
        L.orc_settrap2.argtypes = [C.c_void_p]
        L.orc_save.argtypes = [C.c_void_p, C.c_int]
        L.orc_phase.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_double]
        L.orc_get.restype = C.c_long
        L.orc_get.argtypes = [C.c_void_p, C.c_char_p, C.POINTER(C.c_double)]
        L.orc_set.restype = C.c_long
        L.orc_set.argtypes = [C.c_void_p, C.c_char_p, C.POINTER(C.c_double)]
        L.orc_solver_stats.argtypes = [C.c_void_p, C.POINTER(C.c_double)]
        L.orc_densty.restype = C.c_double
        L.orc_densty.argtypes = [C.c_double, C.c_double, C.c_double]
        L.orc_tvd_flux.restype = C.c_double
        L.orc_tvd_flux.argtypes = [C.c_double] * 7
        dp = C.POINTER(C.c_double)
        L.orc_trid.argtypes = [C.c_int, dp, dp, dp, dp, dp, dp, dp]
        L.orc_trid1.argtypes = [C.c_int, dp, dp, dp, dp, dp]
        L.orc_parab.restype = C.c_double
        L.orc_parab.argtypes = [C.c_double, C.c_double, dp, C.c_int, C.c_double]
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class Oracle:
    """One lake advanced by the CPU oracle."""

    def __init__(self, deck):
        L = lib()
        self._keep = []
        d = OrcDeck()
        for name, _ in OrcDeck._fields_:
            if name in ("zlevel_file", "bathy_dm", "init_profile", "surfbc", "km1_file", "n_surfbc", "ntr",
                        "tracer_profiles", "nopen", "dtsec_openbc", "openbc", "openbc_npts", "openbc_values",
                        "metxy"):
                continue
            if name in deck:
                setattr(d, name, deck[name])
        bathy = np.ascontiguousarray(deck["bathy_dm"], dtype=np.int32)
        self._keep.append(bathy)
        d.bathy_dm = bathy.ctypes.data_as(C.POINTER(C.c_int))
        if deck.get("ibathyf", 0) < 0:
            z = np.ascontiguousarray(deck["zlevel"], dtype=np.float64)
            self._keep.append(z)
            d.zlevel_file = _dp(z)
            d.km1_file = len(z)
        prof = np.ascontiguousarray(deck["init_profile"], dtype=np.float64)
        self._keep.append(prof)
        d.init_profile = _dp(prof)
        if deck.get("surfbc") is not None:
            sb = np.ascontiguousarray(deck["surfbc"], dtype=np.float64)
            self._keep.append(sb)
            d.surfbc = _dp(sb)
            d.n_surfbc = sb.shape[0]
        if deck.get("tracer_profiles") is not None:
            tp = np.ascontiguousarray(deck["tracer_profiles"], dtype=np.float64)
            self._keep.append(tp)
            d.ntr = tp.shape[0]
            d.tracer_profiles = _dp(tp)
        if deck.get("metxy") is not None:
            mx = np.ascontiguousarray(deck["metxy"], dtype=np.float64).ravel()
            self._keep.append(mx)
            d.metxy = _dp(mx)
            d.nmetstat = mx.size // 2
        ob = deck.get("openbc")
        if ob is not None:
            spec = np.ascontiguousarray(np.stack([ob["iside"], ob["itype"], ob["isbc"], ob["jsbc"], ob["iebc"],
                                                  ob["jebc"]], axis=1), dtype=np.int32)
            vals = np.ascontiguousarray(ob["values"], dtype=np.float64)  # (nopen, npts, ntr+2)
            self._keep += [spec, vals]
            d.nopen = spec.shape[0]
            d.dtsec_openbc = float(ob["dtsec"])
            d.openbc = spec.ctypes.data_as(C.POINTER(C.c_int))
            d.openbc_npts = vals.shape[1]
            d.openbc_values = _dp(vals)
        self.deck = deck
        self.h = L.orc_create(C.byref(d))
        if not self.h:
            raise RuntimeError("oracle: deck rejected")
        dims = (C.c_int * 8)()
        L.orc_dims(self.h, dims)
        self.im, self.jm, self.km1, self.lm, self.k1 = dims[0], dims[1], dims[2], dims[3], dims[4]
        self.im1, self.jm1 = dims[6], dims[7]
        self.lm1 = self.lm + 1

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_destroy(self.h)
            self.h = None

    # -- time stepping -------------------------------------------------
    def step(self, n):
        return lib().orc_step(self.h, n)

    def run(self, nsteps, start=1):
        rc = 0
        for n in range(start, start + nsteps):
            rc |= self.step(n)
        return rc

    def fd(self, n, istep, lastiter, thrs):
        return lib().orc_fd(self.h, n, istep, lastiter, thrs)

    def settrap(self):
        lib().orc_settrap(self.h)

    def settrap2(self):
        lib().orc_settrap2(self.h)

    def save(self, istep):
        lib().orc_save(self.h, istep)

    def phase(self, p, n, istep, thrs=0.0):
        return lib().orc_phase(self.h, p, n, istep, thrs)

    # -- fields ---------------------------------------------------------
    def get(self, name):
        n = max(self.km1 * self.lm1, self.im1 * self.jm1) + 8
        out = np.empty(n, dtype=np.float64)
        cnt = lib().orc_get(self.h, name.encode(), _dp(out))
        if cnt < 0:
            raise KeyError(name)
        out = out[:cnt].copy()
        if cnt == self.km1 * self.lm1 and cnt != self.lm1:
            return out.reshape(self.lm1, self.km1)
        return out

    def set(self, name, values):
        v = np.ascontiguousarray(values, dtype=np.float64).ravel()
        cnt = lib().orc_set(self.h, name.encode(), _dp(v))
        if cnt < 0:
            raise KeyError(name)

    def solver_stats(self):
        out = (C.c_double * 3)()
        lib().orc_solver_stats(self.h, out)
        return int(out[0]), float(out[1]), int(out[2])


def densty(t, s, z):
    return lib().orc_densty(t, s, z)


def tvd_flux(vel, s1, s2, s3, s4, cfdef, courant):
    return lib().orc_tvd_flux(vel, s1, s2, s3, s4, cfdef, courant)


def trid(a, b, c, g, r):
    n = len(b)
    a, b, c, g, r = (np.ascontiguousarray(x, dtype=np.float64) for x in (a, b, c, g, r))
    ag = np.empty(n)
    ar = np.empty(n)
    lib().orc_trid(n, _dp(a), _dp(b), _dp(c), _dp(g), _dp(r), _dp(ag), _dp(ar))
    return ag, ar


def trid1(a, b, c, d):
    n = len(b)
    a, b, c, d = (np.ascontiguousarray(x, dtype=np.float64) for x in (a, b, c, d))
    x = np.empty(n)
    lib().orc_trid1(n, _dp(a), _dp(b), _dp(c), _dp(d), _dp(x))
    return x


def parab(frstpt, x, fx, dx):
    fx = np.ascontiguousarray(fx, dtype=np.float64)
    return lib().orc_parab(frstpt, x, _dp(fx), len(fx), dx)
