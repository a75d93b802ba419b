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


def test_c4_64_lakes_volume_residual_and_one_lake_alone():
    """The 64-lake ensemble of the scope (square sides 128-768, 20-80 layers,
    winds 0.05-0.3 Pa from any direction; 12.1M columns, N_wet 4.2e8) in mixed
    precision, 4 steps: every lake's volume is conserved on its own (closed
    basins; zeta is FP64 in every mode), every solve meets the tolerance for
    the worst lake's own right-hand side, and the smallest lake's free surface
    equals the same lake run alone to what FP32 state arithmetic and the
    solver tolerance leave (the ensemble's CG takes other iterations)."""
    from si3d_amd.model import GpuEnsemble, GpuLake
    decks = lake.ensemble(64, seed=7, side=(128, 768), nts=12)
    print("case: 64-lake ensemble, mixed precision", sum(lake.n_wet_cells(d) for d in decks), "wet cells",
          flush=True)
    E = GpuEnsemble(decks, precision="mixed", preconditioner="mg", tol=1e-10)
    print("ensemble created", E.lm, "columns", flush=True)
    off = E.offsets
    s0 = E.get("s")
    v0 = np.array([s0[off[i]:off[i + 1]].sum() for i in range(64)])
    for n in range(1, 5):
        E.step(n)
        st = E.stats()
        print("step", n, "iterations", st["pcg_iters_last"], "worst lake relres", st["pcg_relres_last"], flush=True)
        assert st["pcg_relres_last"] <= 1e-10
    s = E.get("s")
    i0 = int(np.argmin([lake.n_wet_cells(d) for d in decks]))
    for i in range(64):
        si = s[off[i]:off[i + 1]]
        assert np.all(np.isfinite(si)), i
        assert np.max(np.abs(si)) > 0.0, i
        assert abs(si.sum() - v0[i]) <= 1e-9 * si.size * np.max(np.abs(si)), (i, si.sum() - v0[i])
    alone = GpuLake(decks[i0], precision="mixed", preconditioner="mg", tol=1e-10)
    for n in range(1, 5):
        alone.step(n)
    e = relerr(s[off[i0]:off[i0 + 1]], alone.get("s")[:alone.lm])
    print("smallest lake", i0, "zeta vs alone", e, flush=True)
    assert e < 1e-5
