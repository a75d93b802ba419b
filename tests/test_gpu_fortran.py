"""The library called from Fortran through the si3d_gpu module
(ISO_C_BINDING), as the reference driver calls it: the 100x50x20 rectangle
created and stepped from Fortran (si3d_amd/fortran/s3g_fortran_check.f90)
gives the fields the same lake gives through the Python binding, which checks
the Fortran declarations of s3g_deck and s3g_params against the C structs."""
import ctypes as C
import os

import numpy as np
import pytest

from si3d_amd import lake

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "si3d_amd", "_build", "fortran", "libs3g_fortran_check.so")


def test_fortran_driver_matches_python():
    if not os.path.exists(LIB):
        pytest.skip("flang was not available when the library was built")
    from si3d_amd.model import GpuLake
    deck = lake.rectangle(nx=100, ny=50, dx=50.0, depth=20.0, dz=1.0, tau=0.1, dt=30.0, nts=20)
    print("case: 100x50x20 rectangle from Fortran and from Python, 10 steps", flush=True)
    g = GpuLake(deck, precision="fp64", preconditioner="mg", tol=1e-10)
    for n in range(1, 11):
        g.step(n)
    s_py, uh_py = g.get("s"), g.get("uh").ravel()
    s_f = np.zeros(g.lm1)
    uh_f = np.zeros(g.lm1 * g.km1)
    dp = C.POINTER(C.c_double)
    L = C.CDLL(LIB)
    L.s3g_fortran_check.restype = C.c_int
    L.s3g_fortran_check.argtypes = [C.c_int, dp, dp]
    rc = L.s3g_fortran_check(10, s_f.ctypes.data_as(dp), uh_f.ctypes.data_as(dp))
    assert rc >= 0, rc
    assert np.max(np.abs(s_py)) > 1e-4 and np.max(np.abs(uh_py)) > 1e-4  # the wind has set the lake moving
    np.testing.assert_allclose(s_f, s_py, rtol=0, atol=1e-12 * np.max(np.abs(s_py)))
    np.testing.assert_allclose(uh_f, uh_py, rtol=0, atol=1e-12 * np.max(np.abs(uh_py)))


DROPIN_3D = ["h", "hp", "hpp", "hu", "hup", "hupp", "hv", "hvp", "hvpp", "uh", "uhp", "uhpp", "u", "up", "upp", "vh",
             "vhp", "vhpp", "v", "vp", "vpp", "wp", "sal", "salp", "salpp", "rhop", "Av", "Dv"]
DROPIN_8 = ["q2", "q2p", "q2pp", "q2l", "q2lp", "q2lpp"]
GRID_INTS = ["l2i", "l2j", "lEC", "lWC", "lNC", "lSC", "kmz", "k1z", "k1u", "k1v"]


def _dropin_case(oracle_lib):
    """The small stratified ellipse (turbulence closure, TVD transport, wind
    series) and the state the reference's bathy / init leave, taken from the
    oracle at t = 0."""
    from tests.helpers import small_ellipse
    d = small_ellipse()
    d["pcg_tol"] = 1e-12
    o = oracle_lib.Oracle(d)
    grid = {n: o.get(n).astype(np.int32) for n in GRID_INTS}
    grid.update(im1=o.im1, jm1=o.jm1, km1=o.km1, lm=o.lm)
    for n in ("zlevel", "hhs", "hhu", "hhv"):
        grid[n] = o.get(n).astype(np.float32)
    fields = {n: o.get(n) for n in ["s", "sp", "spp"] + DROPIN_3D + DROPIN_8}
    return d, o, grid, fields


def test_from_state_matches_oracle(oracle_lib):
    """The drop-in creation from Python (GpuLake.from_state: s3g_init with the
    grid arrays, then s3g_set_field_* of every state array, FP64) advanced 5
    steps against the oracle from the same state."""
    from si3d_amd.model import GpuLake
    from tests.helpers import FIELDS_2D, FIELDS_3D, FIELDS_TURB, INTS, compare
    d, o, grid, fields = _dropin_case(oracle_lib)
    print("case: drop-in hand-over from Python, small ellipse", flush=True)
    g = GpuLake.from_state(grid, d, fields, precision="fp64", preconditioner="mg", tol=1e-12)
    for n in range(1, 6):
        o.step(n)
        g.step(n)
    for n in INTS:
        np.testing.assert_array_equal(g.get(n), o.get(n).astype(np.int32), err_msg=n)
    compare(g, o, FIELDS_2D + FIELDS_3D + FIELDS_TURB, 1e-8, "from_state after 5 steps")


def test_fortran_dropin_loop_matches_python_and_oracle(oracle_lib):
    """A Fortran caller that does what the patched si3d.f90 does
    (s3g_fortran_dropin: s3g_init with the driver's grid arrays,
    s3g_set_field_f32 of every REAL array and s3g_set_field_f64 of the REAL*8
    turbulence arrays, then the fd / settrap / settrap2 / save loop and
    s3g_get_field_f32 of the outputs) gives, after 5 steps, the fields of the
    Python drop-in path fed the same single-precision state, and stays within
    the single-precision hand-over's rounding of the FP64 oracle."""
    if not os.path.exists(LIB):
        pytest.skip("flang was not available when the library was built")
    from si3d_amd.model import GpuLake, make_params
    d, o, grid, fields = _dropin_case(oracle_lib)
    f32 = {n: (v.astype(np.float32) if n not in DROPIN_8 else v) for n, v in fields.items()}
    p, keep = make_params(d, "fp64", "mg", tol=1e-12)
    km1, lm, lm1 = grid["km1"], grid["lm"], grid["lm"] + 1
    f2d = np.ascontiguousarray(np.stack([f32[n] for n in ("s", "sp", "spp")]), dtype=np.float32)
    f3d = np.ascontiguousarray(np.stack([f32[n] for n in DROPIN_3D]), dtype=np.float32)
    f8 = np.ascontiguousarray(np.stack([f32[n] for n in DROPIN_8]), dtype=np.float64)
    s_f = np.zeros(lm1, np.float32)
    uh_f = np.zeros((lm1, km1), np.float32)
    sal_f = np.zeros((lm1, km1), np.float32)
    q2_f = np.zeros((lm1, km1), np.float64)
    L = C.CDLL(LIB)
    L.s3g_fortran_dropin.restype = C.c_int
    ptr = lambda a: a.ctypes.data_as(C.c_void_p)  # noqa: E731
    args = [C.c_int(grid["im1"]), C.c_int(grid["jm1"]), C.c_int(km1), C.c_int(lm)]
    args += [ptr(np.ascontiguousarray(grid[n])) for n in GRID_INTS]
    args += [ptr(np.ascontiguousarray(grid[n])) for n in ("zlevel", "hhs", "hhu", "hhv")]
    args += [C.byref(p), C.c_int(5), ptr(f2d), ptr(f3d), ptr(f8), ptr(s_f), ptr(uh_f), ptr(sal_f), ptr(q2_f)]
    print("case: drop-in hand-over and time loop from Fortran, small ellipse, 5 steps", flush=True)
    rc = L.s3g_fortran_dropin(*args)
    assert rc >= 0, rc
    g = GpuLake.from_state(grid, d, {n: v.astype(np.float64) for n, v in f32.items()}, precision="fp64",
                           preconditioner="mg", tol=1e-12)
    for n in range(1, 6):
        g.step(n)
        o.step(n)
    for name, got in (("s", s_f), ("uh", uh_f), ("sal", sal_f), ("q2", q2_f)):
        py, orc = g.get(name), o.get(name)
        scale = np.max(np.abs(orc))
        e_py = np.max(np.abs(got - py.astype(got.dtype))) / scale
        e_or = np.max(np.abs(got - orc)) / scale
        print(name, "Fortran vs Python", e_py, "vs oracle", e_or, flush=True)
        assert e_py < (1e-6 if got.dtype == np.float32 else 1e-10), (name, e_py)  # FP32 outputs: one rounding
        assert e_or < 1e-4, (name, e_or)
