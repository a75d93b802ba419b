"""The synthetic states of the oracle's transport pins, run through the GPU
library (s3g_set_field_f64 + s3g_phase) and compared with the oracle on the
same state: matmom on a cosine velocity mode with theta = 0.6 and on split
columns of unequal layers, the baroclinic pressure gradient over a sloping
bed (bclnc_km), exsal + imsal advecting a quadratic in both superbee
branches, imsal's diffusion of a cosine mode, and exmom's advection schemes
2-4.  FP64: to 1e-11 of each field's scale."""
import math

import numpy as np
import pytest

from si3d_amd import lake
from tests.helpers import relerr
from tests.test_oracle_pins import _full, _grid, _momentum_lake, _positions
from tests.test_oracle_pins_transport import _column_lake, _cosine, _transport_lake

pytestmark = pytest.mark.gpu

TOL = 1e-11


class _Both:
    """Sets every field on the oracle and the GPU lake alike."""

    def __init__(self, o, g):
        self.o, self.g = o, g
        self.lm1, self.km1, self.lm = o.lm1, o.km1, o.lm

    def set(self, name, value):
        self.o.set(name, value)
        self.g.set(name, value)

    def phase(self, p, istep=1):
        self.o.phase(p, 2, istep)
        self.g.phase(p, 2, istep)


def _pair(oracle_lib, deck):
    from si3d_amd.model import GpuLake
    print("case", deck["name"], flush=True)
    o = oracle_lib.Oracle(deck)
    g = GpuLake(deck, precision="fp64", preconditioner="mg", tol=1e-12)
    b = _Both(o, g)
    return b, _grid(o)


def _calm(b):
    """Flat free surface; both models keep the deck's (zero) wind of phase 0,
    so the surface drag of the moving top layer acts alike."""
    for n in ("s", "sp", "spp"):
        b.set(n, np.zeros(b.lm1))


def _cmp(b, names, where):
    for n in names:
        gn = {"ex_y": "ey"}.get(n, n)
        on = {"ex_y": "ex"}.get(n, n)
        e = relerr(b.g.get(gn), b.o.get(on))
        assert e <= TOL, (where, n, e)


@pytest.mark.parametrize("X", [True, False])
def test_matmom_cosine_mode_matches_oracle(oracle_lib, X):
    d = _column_lake()
    b, g = _pair(oracle_lib, d)
    b.phase(0)
    _calm(b)
    hq = b.o.get("hup" if X else "hvp")
    mode = 0.03 * _cosine(12, 3, 2, b.km1) + 0.01 * _cosine(12, 0, 2, b.km1)  # the offset keeps eagx off zero
    q = np.ones((b.lm1, 1)) * mode[None, :]
    b.set("Av", _full(b.o, d["Av0"]))
    b.set("upp" if X else "vpp", q)
    b.set("uhpp" if X else "vhpp", q * hq)
    b.o.set("ex", q * hq)
    b.g.set("ex" if X else "ey", q * hq)
    b.phase(3 if X else 5)
    _cmp(b, ["agx", "arx", "eagx", "earx"] if X else ["agy", "ary", "eagy", "eary"], "matmom cosine mode")


@pytest.mark.parametrize("X", [True, False])
def test_matmom_baroclinic_sloping_bed_matches_oracle(oracle_lib, X):
    d = lake.rectangle(nx=10, ny=10, dx=40.0, depth=8.0, dz=1.0, dt=60.0, nts=10, iadv=0, ihd=0, cd=0.0,
                       Av0=0.0, tau=0.0)
    d["f"], d["ibc"], d["zl"] = 0.0, 1, 12.0
    i, j = np.meshgrid(np.arange(10), np.arange(10), indexing="ij")
    d["bathy_dm"] = (53 + 7 * (i if X else j)).astype(np.int32)
    d["init_profile"] = np.full(lake.km1_of(d), 15.0)
    b, g = _pair(oracle_lib, d)
    b.phase(0)
    _calm(b)
    b.set("Av", _full(b.o, 0.0))
    hp = b.o.get("hp")
    pos = (g["l2i"] if X else g["l2j"]).astype(float) * d["dx"]
    zc = np.cumsum(hp, axis=1) - hp / 2
    b.set("rhop", -1.2 + 2.0e-4 * pos[:, None] + 1.5e-2 * zc)
    b.phase(3 if X else 5)
    _cmp(b, ["agx", "arx"] if X else ["agy", "ary"], "baroclinic over a sloping bed")


@pytest.mark.parametrize("X", [True, False])
@pytest.mark.parametrize("regime", ["lax_wendroff", "beam_warming"])
def test_exsal_quadratic_matches_oracle(oracle_lib, X, regime):
    d = _transport_lake()
    b, g = _pair(oracle_lib, d)
    sp = d["dx"] if X else d["dy"]
    pos = (g["l2i"] if X else g["l2j"]).astype(float) * sp
    n = d["imm"] if X else d["jmm"]
    x0 = pos.min() - 4 * sp if regime == "lax_wendroff" else pos.max() + 4 * sp + (n + 3) * sp
    s = 20.0 + 1.0e-3 * (pos - x0) ** 2
    S = np.ones((b.lm1, 1)) * np.ones(b.km1)[None, :] * s[:, None]
    for name in ("sal", "salp", "salpp"):
        b.set(name, S)
    hq, hn = b.o.get("hup"), b.o.get("hvp")
    zero = _full(b.o, 0.0)
    U = 0.12
    f = ({"uh": 1.5 * U * hq, "uhpp": 0.5 * U * hq, "vh": zero, "vhpp": zero} if X else
         {"vh": 1.5 * U * hn, "vhpp": 0.5 * U * hn, "uh": zero, "uhpp": zero})
    for name, v in list(f.items()) + [("wp", zero), ("Dv", zero)]:
        b.set(name, v)
    b.phase(9)
    _cmp(b, ["bex"], "exsal")
    b.phase(10)
    _cmp(b, ["sal"], "imsal")


@pytest.mark.parametrize("istep", [1, 2])
def test_imsal_cosine_mode_and_split_column_match_oracle(oracle_lib, istep):
    d = _transport_lake(nx=8, ny=6)
    b, g = _pair(oracle_lib, d)
    zero = _full(b.o, 0.0)
    for name in ("uh", "uhpp", "vh", "vhpp", "wp"):
        b.set(name, zero)
    Dv = _full(b.o, 3.0e-3)
    Dv[:, 3] = 0.0  # one interface closed
    b.set("Dv", Dv)
    mode = 15.0 + 2.0 * _cosine(int(g["kmz"][0]) - 1, 1, 2, b.km1)
    b.set("salpp", np.ones((b.lm1, 1)) * mode[None, :])
    b.phase(9, istep)
    b.phase(10, istep)
    _cmp(b, ["sal"], "imsal")


@pytest.mark.parametrize("itrmom", [2, 3, 4])
@pytest.mark.parametrize("X", [True, False])
def test_exmom_schemes_linear_fields_match_oracle(oracle_lib, itrmom, X):
    d = _momentum_lake(iadv=1, ihd=0)
    d["f"], d["itrmom"] = 0.0, itrmom
    b, g = _pair(oracle_lib, d)
    dx, dy = d["dx"], d["dy"]
    Xp, Yp, D = _positions(b.o, g, dx, dy)
    xo, yo = (dx / 2, 0.0) if X else (0.0, dy / 2)

    def field(ax, ay, az):
        return 0.05 + ax * (Xp[:, None] + xo) + ay * (Yp[:, None] + yo) + az * D[None, :]

    qn, qpp = field(2.0e-5, -3.0e-5, 4.0e-3), field(-1.2e-5, 2.5e-5, 6.0e-3)
    hu, hv = b.o.get("hup"), b.o.get("hvp")
    fields = ([("up", qn), ("upp", qpp), ("uhp", 0.04 * hu), ("uhpp", 0.7 * qn * hu), ("vhp", 0.03 * hv)] if X else
              [("vp", qn), ("vpp", qpp), ("vhp", 0.03 * hv), ("vhpp", 0.7 * qn * hv), ("uhp", 0.04 * hu)])
    for name, val in fields + [("wp", _full(b.o, 1.5e-4))]:
        b.set(name, val)
    b.phase(2 if X else 4)
    _cmp(b, ["ex" if X else "ex_y"], f"exmom itrmom {itrmom}")
