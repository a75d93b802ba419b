"""An ensemble of lakes of uneven size, layer count and wind, advanced together
(one launch per phase over all lakes), against each lake run on its own."""
import numpy as np
import pytest

from si3d_amd import lake
from tests.helpers import INTS, relerr

pytestmark = pytest.mark.gpu

FIELDS = ["s", "uh", "vh", "wp", "h", "sal", "rhop", "Av", "Dv", "q2", "q2l"]


def _decks(nts=8):
    return [
        lake.ellipse(nx=40, ny=24, dx=100.0, nlayers=12, nts=nts, top=1.0, bottom=12.0, depth=40.0, seed=3,
                     wind_amp=0.2, wind_dir=30.0),
        lake.ellipse(nx=32, ny=32, dx=100.0, nlayers=8, nts=nts, top=1.0, bottom=8.0, depth=30.0, seed=5,
                     wind_amp=0.1, wind_dir=200.0),
        lake.ellipse(nx=56, ny=20, dx=100.0, nlayers=16, nts=nts, top=0.5, bottom=6.0, depth=35.0, seed=8,
                     wind_amp=0.3, wind_dir=75.0),
    ]


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-9), ("mixed", 1e-9)])
def test_ensemble_matches_lakes_run_alone(precision, tol):
    from si3d_amd.model import GpuEnsemble, GpuLake
    decks = _decks()
    print(f"case: ensemble of 3 lakes, {precision}", flush=True)
    E = GpuEnsemble(decks, precision=precision, preconditioner="mg", tol=1e-13)
    alone = [GpuLake(d, precision=precision, preconditioner="mg", tol=1e-13) for d in decks]
    assert E.nlakes == 3
    assert list(E.lake_km1) == [a.km1 for a in alone]
    assert list(np.diff(E.offsets)) == [a.lm for a in alone]
    for n in range(1, 6):
        E.step(n)
        for a in alone:
            a.step(n)
    for i, a in enumerate(alone):
        lm, km1 = a.lm, a.km1
        for name in INTS:
            np.testing.assert_array_equal(E.lake_get(name, i), a.get(name)[:lm], err_msg=f"lake {i} {name}")
        for name in FIELDS:
            ref = a.get(name)
            ref = ref[:lm] if ref.ndim == 1 else ref[:lm, :km1]
            e = relerr(E.lake_get(name, i), ref)
            assert e < tol, (i, name, e)


def test_ensemble_lakes_match_oracle(oracle_lib):
    """Every lake of the ensemble against the FP64 oracle run on that lake's
    own deck: integer state identical, fields within 1e-8 after 5 steps.
    The ensemble's free-surface solve stops only when each lake's true
    residual is below the tolerance relative to that lake's right-hand side,
    so a small, weakly forced lake is solved as tightly as a large one."""
    from si3d_amd.model import GpuEnsemble
    from tests.helpers import FIELDS_2D, FIELDS_3D, FIELDS_TURB
    decks = _decks()
    for d in decks:
        d["pcg_tol"] = 1e-13
    print("case: ensemble of 3 lakes against the oracle", flush=True)
    E = GpuEnsemble(decks, precision="fp64", preconditioner="mg", tol=1e-13)
    orcs = [oracle_lib.Oracle(d) for d in decks]
    for n in range(1, 6):
        E.step(n)
        for o in orcs:
            assert o.step(n) == 0
    assert E.stats()["pcg_relres_last"] <= 1e-13
    for i, o in enumerate(orcs):
        lm, km1 = o.lm, o.km1
        for name in INTS:
            np.testing.assert_array_equal(E.lake_get(name, i), o.get(name)[:lm].astype(np.int32),
                                          err_msg=f"lake {i} {name}")
        for name in FIELDS_2D + FIELDS_3D + FIELDS_TURB:
            ref = o.get(name)
            ref = ref[:lm] if ref.ndim == 1 else ref[:lm, :km1]
            e = relerr(E.lake_get(name, i), ref)
            assert e < 1e-8, (i, name, e)


def test_ensemble_rejects_mismatched_time_steps():
    from si3d_amd.model import GpuEnsemble
    decks = _decks()
    decks[1] = dict(decks[1])
    decks[1]["dt"] = 30.0
    with pytest.raises(RuntimeError, match="differ only"):
        GpuEnsemble(decks)


def test_ensemble_rejects_heat_budget():
    from si3d_amd.model import GpuEnsemble
    decks = _decks()
    decks[2] = lake.heat_budget(decks[2], 1)
    with pytest.raises(RuntimeError, match="ifsurfbc"):
        GpuEnsemble(decks)
