"""Host-side driver of one lake on the GPU, mirroring the reference's call
sequence: the si3d.f90 time loop (324-393) calling fd / settrap / settrap2 /
save, with field access for outputs.

Two ways to create a lake:

* `GpuLake(deck)` -- standalone: the library restates input/bathy/init from
  an in-memory deck (see si3d_amd.lake).
* `GpuLake.from_state(grid, params, fields)` -- drop-in: the caller (the
  Fortran driver, or a test holding the reference-layout arrays) hands over
  the grid and every state array, exactly as si3d.f90 would after `init`.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import S3gDeck, S3gGrid, S3gLearnStats, S3gParams, S3gStats, check

_PREC = {"fp64": _lib.S3G_FP64, "mixed": _lib.S3G_MIXED}
_PRECOND = {"jacobi": _lib.S3G_PREC_JACOBI, "line": _lib.S3G_PREC_LINE, "mg": _lib.S3G_PREC_MG}
INT_FIELDS = ("kmz", "k1z", "k1u", "k1v")
TWO_D = ("s", "sp", "spp", "sx", "sy", "dd", "qq", "rr", "eagx", "earx", "eagy", "eary",
         "hhs", "hhu", "hhv") + INT_FIELDS


def make_params(deck, precision="fp64", preconditioner="jacobi", tol=None, maxit=None):
    p = S3gParams()
    for name, _ in S3gParams._fields_:
        if name in ("surfbc", "n_surfbc", "precision", "preconditioner", "pcg_tol", "pcg_maxit", "ntr", "metxy",
                    "nmetstat"):
            continue
        if name in deck:
            setattr(p, name, deck[name])
    keep = []
    if deck.get("surfbc") is not None:
        sb = np.ascontiguousarray(deck["surfbc"], dtype=np.float64)
        keep.append(sb)
        p.surfbc = sb.ctypes.data_as(_lib.dp)
        p.n_surfbc = sb.shape[0]
    if deck.get("tracer_profiles") is not None:
        p.ntr = int(np.asarray(deck["tracer_profiles"]).shape[0])
    if deck.get("metxy") is not None:
        mx = np.ascontiguousarray(deck["metxy"], dtype=np.float64).ravel()
        keep.append(mx)
        p.metxy = mx.ctypes.data_as(_lib.dp)
        p.nmetstat = mx.size // 2
    p.precision = _PREC[precision]
    p.preconditioner = _PRECOND[preconditioner]
    p.pcg_tol = tol if tol is not None else deck.get("pcg_tol", 1e-10)
    p.pcg_maxit = maxit if maxit is not None else deck.get("pcg_maxit", 4000)
    return p, keep


def _deck_struct(deck, keep):
    d = S3gDeck()
    for name in ("imm", "jmm", "ibathyf", "ddz", "dzmin", "datadj", "zetainit", "zl"):
        setattr(d, name, deck[name])
    bathy = np.ascontiguousarray(deck["bathy_dm"], dtype=np.int32)
    prof = np.ascontiguousarray(deck["init_profile"], dtype=np.float64)
    keep += [bathy, prof]
    d.bathy_dm = bathy.ctypes.data_as(_lib.ip)
    d.init_profile = prof.ctypes.data_as(_lib.dp)
    if deck["ibathyf"] < 0:
        z = np.ascontiguousarray(deck["zlevel"], dtype=np.float64)
        keep.append(z)
        d.zlevel_file = z.ctypes.data_as(_lib.dp)
        d.km1_file = len(z)
    if deck.get("tracer_profiles") is not None:
        tp = np.ascontiguousarray(deck["tracer_profiles"], dtype=np.float64)
        keep.append(tp)
        d.tracer_profiles = tp.ctypes.data_as(_lib.dp)
    return d


class GpuLake:
    def __init__(self, deck=None, precision="fp64", preconditioner="jacobi", tol=None,
                 maxit=None, _handle=None, _dims=None):
        self.L = _lib.load()
        self.precision = precision
        self._keep = []
        if _handle is not None:
            self.h = _handle
            self.km1, self.lm = _dims
        else:
            p, keep = make_params(deck, precision, preconditioner, tol, maxit)
            self._keep += keep
            d = _deck_struct(deck, self._keep)
            h = C.c_void_p()
            check(self.L.s3g_create_from_deck(C.byref(d), C.byref(p), C.byref(h)), "s3g_create_from_deck")
            self.h = h
            st = self.stats()
            self.km1, self.lm = st["km1"], st["lm"]
            if deck.get("openbc") is not None:
                self.set_open_boundaries(deck["openbc"], deck.get("dzmin", 0.1))
        self.lm1 = self.lm + 1

    def set_open_boundaries(self, ob, dzmin):
        """openbc0: the nopen block of si3d_inp.txt (`ob` as built by
        lake.open_boundary or read by inputs.read_case) and its series."""
        vals = np.ascontiguousarray(ob["values"], dtype=np.float64)  # (nopen, npts, 2+ntr)
        b = _lib.S3gOpenBc()
        b.nopen = vals.shape[0]
        b.dtsec = float(ob["dtsec"])
        b.dzmin = float(dzmin)
        for key in ("iside", "itype", "isbc", "jsbc", "iebc", "jebc"):
            a = np.ascontiguousarray(ob[key], dtype=np.int32)
            self._keep.append(a)
            setattr(b, key, a.ctypes.data_as(_lib.ip))
        b.npts, b.nvar = vals.shape[1], vals.shape[2]
        self._keep.append(vals)
        b.values = vals.ctypes.data_as(_lib.dp)
        check(self.L.s3g_set_open_boundaries(self.h, C.byref(b)), "s3g_set_open_boundaries")

    @classmethod
    def from_state(cls, grid, deck, fields, precision="fp64", preconditioner="jacobi", tol=None):
        """Drop-in creation: `grid` holds the bathy/init index arrays (1-based,
        reference layout), `fields` maps state-array names to reference-layout
        arrays ((lm1, km1) for 3-D, (lm1,) for 2-D)."""
        L = _lib.load()
        g = S3gGrid()
        keep = []
        for name in ("im1", "jm1", "km1", "lm"):
            setattr(g, name, int(grid[name]))
        for name in ("l2i", "l2j", "lEC", "lWC", "lNC", "lSC", "kmz", "k1z", "k1u", "k1v"):
            a = np.ascontiguousarray(grid[name], dtype=np.int32)
            keep.append(a)
            setattr(g, name, a.ctypes.data_as(_lib.ip))
        for name in ("zlevel", "hhs", "hhu", "hhv"):
            a = np.ascontiguousarray(grid[name], dtype=np.float32)
            keep.append(a)
            setattr(g, name, a.ctypes.data_as(_lib.fp))
        p, k2 = make_params(deck, precision, preconditioner, tol)
        keep += k2
        h = C.c_void_p()
        check(L.s3g_init(C.byref(g), C.byref(p), C.byref(h)), "s3g_init")
        obj = cls(_handle=h, _dims=(int(grid["km1"]), int(grid["lm"])), precision=precision)
        obj._keep = keep
        for name, arr in fields.items():
            obj.set(name, arr)
        return obj

    # -- lifecycle --------------------------------------------------------
    def close(self):
        if getattr(self, "h", None):
            self.L.s3g_finalize(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- the reference's calls -------------------------------------------
    def fd(self, n, istep, lastiter=1, thrs=None, iter=1):
        if thrs is None:
            thrs = 0.0
        return check(self.L.s3g_fd(self.h, n, istep, iter, lastiter, thrs), "s3g_fd")

    def settrap(self):
        return check(self.L.s3g_settrap(self.h), "s3g_settrap")

    def settrap2(self):
        return check(self.L.s3g_settrap2(self.h), "s3g_settrap2")

    def save(self, istep):
        return check(self.L.s3g_save(self.h, istep), "s3g_save")

    def step(self, n):
        return check(self.L.s3g_step(self.h, n), "s3g_step")

    def run(self, nsteps, start=1):
        return check(self.L.s3g_run(self.h, start, nsteps), "s3g_run")

    def phase(self, p, n, istep, thrs=0.0):
        return check(self.L.s3g_phase(self.h, p, n, istep, thrs), "s3g_phase")

    def synchronize(self):
        return check(self.L.s3g_synchronize(self.h), "s3g_synchronize")

    # -- fields -----------------------------------------------------------
    def get(self, name, dtype=np.float64):
        if name in INT_FIELDS:
            out = np.empty(self.lm1, dtype=np.int32)
            check(self.L.s3g_get_field_i32(self.h, name.encode(), out.ctypes.data_as(_lib.ip)), name)
            return out
        n = self.lm1 if name in TWO_D else self.lm1 * self.km1
        out = np.empty(n, dtype=np.float64)
        check(self.L.s3g_get_field_f64(self.h, name.encode(), out.ctypes.data_as(_lib.dp)), name)
        return out if name in TWO_D else out.reshape(self.lm1, self.km1)

    def set(self, name, values):
        if name in INT_FIELDS:
            v = np.ascontiguousarray(values, dtype=np.int32).ravel()
            check(self.L.s3g_set_field_i32(self.h, name.encode(), v.ctypes.data_as(_lib.ip)), name)
            return
        v = np.ascontiguousarray(values, dtype=np.float64).ravel()
        check(self.L.s3g_set_field_f64(self.h, name.encode(), v.ctypes.data_as(_lib.dp)), name)

    def stats(self):
        st = S3gStats()
        check(self.L.s3g_get_stats(self.h, C.byref(st)), "s3g_get_stats")
        return {name: getattr(st, name) for name, _ in S3gStats._fields_}

    def energy_balance(self, reset=False):
        """(TKinE, ShearCum, BuocyCum, DisspCum) of the closure's basin energy
        budget; `reset` clears the three integrals afterwards."""
        out = (C.c_double * 4)()
        check(self.L.s3g_energy_balance(self.h, out, 1 if reset else 0), "s3g_energy_balance")
        return tuple(out)

    def timers(self, enable=True):
        check(self.L.s3g_reset_timers(self.h, 1 if enable else 0), "s3g_reset_timers")

    # -- learned initial guess ---------------------------------------------
    def learn_setup(self, capacity=64, seed=1):
        return check(self.L.s3g_learn_setup(self.h, capacity, seed), "s3g_learn_setup")

    def learn_record(self, on=True):
        return check(self.L.s3g_learn_record(self.h, 1 if on else 0), "s3g_learn_record")

    def learn_train(self, steps, batch=4, lr=1e-3):
        loss = C.c_double()
        check(self.L.s3g_learn_train(self.h, steps, batch, lr, C.byref(loss)), "s3g_learn_train")
        return loss.value

    def learn_mode(self, mode):
        return check(self.L.s3g_learn_mode(self.h, mode), "s3g_learn_mode")

    def learn_flexible(self, period):
        """Learned correction stage inside flexible CG every `period`
        iterations (0: off)."""
        return check(self.L.s3g_learn_flexible(self.h, period), "s3g_learn_flexible")

    def learn_weights(self, values=None):
        n = self.L.s3g_learn_num_weights()
        if values is None:
            out = np.empty(n, dtype=np.float32)
            check(self.L.s3g_learn_get_weights(self.h, out.ctypes.data_as(_lib.fp), n), "s3g_learn_get_weights")
            return out
        v = np.ascontiguousarray(values, dtype=np.float32)
        check(self.L.s3g_learn_set_weights(self.h, v.ctypes.data_as(_lib.fp), n), "s3g_learn_set_weights")

    def learn_info(self):
        st = S3gLearnStats()
        check(self.L.s3g_learn_info(self.h, C.byref(st)), "s3g_learn_info")
        return {name: getattr(st, name) for name, _ in S3gLearnStats._fields_}

    TILE_NAMES = ("exmom_x", "exmom_y", "slayer_h", "vel", "exsal_hflux", "vel2", "exsal_rhs", "turb_cell", "settrap", "save")

    def tiles(self):
        kc = (C.c_int * 16)()
        n = check(self.L.s3g_get_tiles(self.h, 16, kc), "s3g_get_tiles")
        return {self.TILE_NAMES[i]: [kc[i] & 255, kc[i] >> 8] for i in range(n)}

    def kernel_times(self, maxn=64):
        names = (C.c_char * 32 * maxn)()
        avg = (C.c_double * maxn)()
        cnt = (C.c_long * maxn)()
        n = check(self.L.s3g_kernel_times(self.h, maxn, names, avg, cnt), "s3g_kernel_times")
        return {bytes(names[i]).split(b"\0")[0].decode(): (avg[i], cnt[i]) for i in range(n)}


def unet_num_weights():
    """Number of weights of the library's U-Net (the length of a weight vector)."""
    return int(_lib.load().s3g_learn_num_weights())


def unet_check(weights, x3, target, mask, static=False):
    """Forward + backward of the library's U-Net on one dense input (x3: 3 x H x
    W); returns (out H x W, weight gradients, loss).  `static` calls the
    variant that allocates no device memory (s3g_unet_check_static, at most
    4096 cells)."""
    L = _lib.load()
    fn = L.s3g_unet_check_static if static else L.s3g_unet_check
    _, H, W = x3.shape
    w = np.ascontiguousarray(weights, dtype=np.float32)
    x = np.ascontiguousarray(x3, dtype=np.float32)
    t = np.ascontiguousarray(target, dtype=np.float32)
    m = np.ascontiguousarray(mask, dtype=np.float32)
    out = np.empty((H, W), dtype=np.float32)
    grads = np.empty(L.s3g_learn_num_weights(), dtype=np.float32)
    loss = C.c_double()
    check(fn(H, W, w.ctypes.data_as(_lib.fp), x.ctypes.data_as(_lib.fp), t.ctypes.data_as(_lib.fp),
             m.ctypes.data_as(_lib.fp), out.ctypes.data_as(_lib.fp), grads.ctypes.data_as(_lib.fp), C.byref(loss)),
          "s3g_unet_check_static" if static else "s3g_unet_check")
    return out, grads, loss.value


class GpuEnsemble(GpuLake):
    """Independent lakes advanced together (s3g_create_ensemble): one launch per
    phase over the columns of all lakes.  `lake_get(name, i)` returns lake i's
    part of a field in the layout of a single lake (its columns, its km1
    levels; no sentinel column)."""

    def __init__(self, decks, precision="fp64", preconditioner="mg", tol=None, maxit=None):
        L = _lib.load()
        n = len(decks)
        if any(dk.get("openbc") is not None for dk in decks):
            raise ValueError("ensemble lakes take no open boundaries (run such a lake alone)")
        keep = []
        darr = (S3gDeck * n)()
        parr = (S3gParams * n)()
        for i, dk in enumerate(decks):
            darr[i] = _deck_struct(dk, keep)
            p, k2 = make_params(dk, precision, preconditioner, tol, maxit)
            keep += k2
            parr[i] = p
        h = C.c_void_p()
        check(L.s3g_create_ensemble(n, darr, parr, C.byref(h)), "s3g_create_ensemble")
        st = S3gStats()
        check(L.s3g_get_stats(h, C.byref(st)), "s3g_get_stats")
        super().__init__(_handle=h, _dims=(st.km1, st.lm), precision=precision)
        self._keep = keep + [darr, parr]
        off = np.zeros(n + 1, dtype=np.int32)
        km1 = np.zeros(n, dtype=np.int32)
        check(L.s3g_ensemble_info(h, n, off.ctypes.data_as(_lib.ip), km1.ctypes.data_as(_lib.ip)),
              "s3g_ensemble_info")
        self.offsets, self.lake_km1, self.nlakes = off, km1, n

    def lake_get(self, name, i, dtype=np.float64):
        a = self.get(name, dtype)
        sl = slice(int(self.offsets[i]), int(self.offsets[i + 1]))
        if a.ndim == 1:
            return a[sl]
        return a[sl, :int(self.lake_km1[i])]
