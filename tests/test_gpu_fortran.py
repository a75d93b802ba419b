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
