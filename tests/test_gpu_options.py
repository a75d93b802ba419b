"""Parity of the HIP time step with the oracle across the options of the
scope's rows that the default lakes leave at one value: momentum advection
schemes (itrmom 2-4), Smagorinsky horizontal mixing (ihd 2), the Asselin filter
of leapfrog-only stepping (itrap 0, ismooth 1), trapezoidal iterations every
other step (itrap 2) and three of them (niter 3), no momentum advection
(iadv 0), constant vertical mixing with a baroclinic scalar (iturb 0), and no
Coriolis term."""
import numpy as np
import pytest

from tests.helpers import FIELDS_2D, FIELDS_3D, FIELDS_TURB, INTS, compare, small_ellipse

pytestmark = pytest.mark.gpu

VARIANTS = {
    "itrmom2": dict(itrmom=2),
    "itrmom3": dict(itrmom=3),
    "itrmom4": dict(itrmom=4),
    "smagorinsky": dict(ihd=2),
    "leapfrog_asselin": dict(itrap=0, ismooth=1, beta=0.1),
    "trapezoid_every_2nd": dict(itrap=2),
    "three_iterations": dict(niter=3),
    "no_advection": dict(iadv=0),
    "constant_mixing": dict(iturb=0, Av0=1e-3, Dv0=1e-4),
    "no_coriolis": dict(f=0.0),
}


@pytest.mark.parametrize("name", list(VARIANTS))
def test_option_matches_oracle(oracle_lib, name):
    from si3d_amd.model import GpuLake
    deck = small_ellipse(nts=12)
    deck.update(VARIANTS[name])
    deck["pcg_tol"] = 1e-12
    print(f"case: small ellipse, {name}", flush=True)
    o = oracle_lib.Oracle(deck)
    g = GpuLake(deck, precision="fp64", preconditioner="mg", tol=1e-12)
    for n in range(1, 7):
        o.step(n)
        g.step(n)
    for i in INTS:
        np.testing.assert_array_equal(g.get(i), o.get(i).astype(np.int32), err_msg=f"{name} {i}")
    names = FIELDS_2D + FIELDS_3D + (FIELDS_TURB if deck["iturb"] else [])
    if deck["ihd"] > 1:
        names = names + ["kh"]
    compare(g, o, names, 1e-8, name)
