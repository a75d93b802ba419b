"""ctypes binding of libsi3dgpu.so (declared in include/si3d_gpu.h).

The library is built in-tree by si3d_amd/build.py.  Loading fails loudly when
it is missing: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libsi3dgpu.so")

S3G_FP64, S3G_MIXED = 0, 1
S3G_PREC_JACOBI, S3G_PREC_LINE, S3G_PREC_MG = 0, 1, 2

EXPORTS = [
    "s3g_init", "s3g_set_field_f32", "s3g_set_field_f64", "s3g_set_field_i32",
    "s3g_get_field_f32", "s3g_get_field_f64", "s3g_get_field_i32", "s3g_fd",
    "s3g_settrap", "s3g_settrap2", "s3g_save", "s3g_step", "s3g_run", "s3g_phase",
    "s3g_synchronize", "s3g_get_stats", "s3g_kernel_times", "s3g_reset_timers", "s3g_get_tiles",
    "s3g_finalize", "s3g_last_error", "s3g_create_from_deck",
    "s3g_learn_setup", "s3g_learn_record", "s3g_learn_train", "s3g_learn_mode", "s3g_learn_num_weights",
    "s3g_learn_get_weights", "s3g_learn_set_weights", "s3g_learn_info", "s3g_learn_flexible", "s3g_unet_check",
    "s3g_create_ensemble", "s3g_ensemble_info", "s3g_energy_balance", "s3g_set_open_boundaries",
    "s3g_unet_check_static", "s3g_probe_stream",
]

dp = C.POINTER(C.c_double)
ip = C.POINTER(C.c_int)
fp = C.POINTER(C.c_float)


class S3gGrid(C.Structure):
    _fields_ = [("im1", C.c_int), ("jm1", C.c_int), ("km1", C.c_int), ("lm", C.c_int),
                ("l2i", ip), ("l2j", ip), ("lEC", ip), ("lWC", ip), ("lNC", ip), ("lSC", ip),
                ("kmz", ip), ("k1z", ip), ("k1u", ip), ("k1v", ip),
                ("zlevel", fp), ("hhs", fp), ("hhu", fp), ("hhv", fp)]


class S3gParams(C.Structure):
    _fields_ = [("dt", C.c_double), ("dx", C.c_double), ("dy", C.c_double),
                ("itrap", C.c_int), ("niter", C.c_int), ("ismooth", C.c_int), ("iturb", C.c_int),
                ("iadv", C.c_int), ("itrmom", C.c_int), ("ihd", C.c_int), ("ibc", C.c_int),
                ("isal", C.c_int), ("ifsurfbc", C.c_int),
                ("beta", C.c_double), ("Av0", C.c_double), ("Dv0", C.c_double), ("Ax0", C.c_double),
                ("Ay0", C.c_double), ("f", C.c_double), ("theta", C.c_double), ("cd", C.c_double),
                ("cw", C.c_double), ("wa", C.c_double), ("phi", C.c_double), ("dtsurfbc", C.c_double),
                ("n_surfbc", C.c_int), ("surfbc", dp), ("precision", C.c_int),
                ("preconditioner", C.c_int), ("pcg_tol", C.c_double), ("pcg_maxit", C.c_int),
                ("ntr", C.c_int), ("icw", C.c_int), ("nmetstat", C.c_int), ("metxy", dp), ("iinterp", C.c_int),
                ("gammaB", C.c_double), ("delNfactor", C.c_double)]


class S3gDeck(C.Structure):
    _fields_ = [("imm", C.c_int), ("jmm", C.c_int), ("ibathyf", C.c_int), ("km1_file", C.c_int),
                ("ddz", C.c_double), ("dzmin", C.c_double), ("datadj", C.c_double),
                ("zetainit", C.c_double), ("zl", C.c_double), ("zlevel_file", dp),
                ("bathy_dm", ip), ("init_profile", dp), ("tracer_profiles", dp)]


class S3gOpenBc(C.Structure):
    _fields_ = [("nopen", C.c_int), ("dtsec", C.c_double), ("dzmin", C.c_double),
                ("iside", ip), ("itype", ip), ("isbc", ip), ("jsbc", ip), ("iebc", ip), ("jebc", ip),
                ("npts", C.c_int), ("nvar", C.c_int), ("values", dp)]


class S3gStats(C.Structure):
    _fields_ = [("steps", C.c_long), ("fd_calls", C.c_long), ("solves", C.c_long),
                ("pcg_iters_total", C.c_long), ("pcg_iters_last", C.c_int),
                ("pcg_relres_last", C.c_double), ("zeta_volume", C.c_double),
                ("lm", C.c_int), ("km1", C.c_int), ("n_wet", C.c_long), ("graph_replays", C.c_long)]


class S3gLearnStats(C.Structure):
    _fields_ = [("pairs", C.c_int), ("capacity", C.c_int), ("train_steps", C.c_long),
                ("last_loss", C.c_double), ("height", C.c_int), ("width", C.c_int), ("mode", C.c_int),
                ("guess_solves", C.c_long), ("guess_error_ratio", C.c_double), ("guess_rejected", C.c_long),
                ("corrections", C.c_long), ("flex_period", C.c_int)]


_lib = None


def load():
    """Load libsi3dgpu.so; raise if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'`")
    L = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    L.s3g_init.argtypes = [C.POINTER(S3gGrid), C.POINTER(S3gParams), C.POINTER(vp)]
    L.s3g_create_from_deck.argtypes = [C.POINTER(S3gDeck), C.POINTER(S3gParams), C.POINTER(vp)]
    for n, t in (("f32", fp), ("f64", dp), ("i32", ip)):
        getattr(L, f"s3g_set_field_{n}").argtypes = [vp, C.c_char_p, t]
        getattr(L, f"s3g_get_field_{n}").argtypes = [vp, C.c_char_p, t]
    L.s3g_fd.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double]
    L.s3g_settrap.argtypes = [vp]
    L.s3g_settrap2.argtypes = [vp]
    L.s3g_save.argtypes = [vp, C.c_int]
    L.s3g_step.argtypes = [vp, C.c_int]
    L.s3g_run.argtypes = [vp, C.c_int, C.c_int]
    L.s3g_phase.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_double]
    L.s3g_synchronize.argtypes = [vp]
    L.s3g_get_stats.argtypes = [vp, C.POINTER(S3gStats)]
    L.s3g_energy_balance.argtypes = [vp, dp, C.c_int]
    L.s3g_kernel_times.argtypes = [vp, C.c_int, C.c_void_p, dp, C.POINTER(C.c_long)]
    L.s3g_reset_timers.argtypes = [vp, C.c_int]
    L.s3g_get_tiles.argtypes = [vp, C.c_int, ip]
    L.s3g_probe_stream.argtypes = [C.c_size_t, C.c_int, dp]
    L.s3g_learn_setup.argtypes = [vp, C.c_int, C.c_uint]
    L.s3g_learn_record.argtypes = [vp, C.c_int]
    L.s3g_learn_train.argtypes = [vp, C.c_int, C.c_int, C.c_double, dp]
    L.s3g_learn_mode.argtypes = [vp, C.c_int]
    L.s3g_learn_flexible.argtypes = [vp, C.c_int]
    L.s3g_learn_num_weights.restype = C.c_long
    L.s3g_learn_num_weights.argtypes = []
    L.s3g_learn_get_weights.argtypes = [vp, fp, C.c_long]
    L.s3g_learn_set_weights.argtypes = [vp, fp, C.c_long]
    L.s3g_learn_info.argtypes = [vp, C.POINTER(S3gLearnStats)]
    L.s3g_unet_check.argtypes = [C.c_int, C.c_int, fp, fp, fp, fp, fp, fp, dp]
    L.s3g_unet_check_static.argtypes = [C.c_int, C.c_int, fp, fp, fp, fp, fp, fp, dp]
    L.s3g_create_ensemble.argtypes = [C.c_int, C.POINTER(S3gDeck), C.POINTER(S3gParams), C.POINTER(vp)]
    L.s3g_ensemble_info.argtypes = [vp, C.c_int, ip, ip]
    L.s3g_set_open_boundaries.argtypes = [vp, C.POINTER(S3gOpenBc)]
    L.s3g_finalize.argtypes = [vp]
    L.s3g_finalize.restype = None
    L.s3g_last_error.restype = C.c_char_p
    _lib = L
    return L


def check(rc, what="s3g"):
    if rc < 0:
        msg = load().s3g_last_error().decode(errors="replace")
        raise RuntimeError(f"{what} failed ({rc}): {msg}")
    return rc
