"""Parity of the HIP time step (libsi3dgpu.so through its C-ABI) with the FP64
CPU oracle on identical decks.  Integer state (masks, kmz, top-layer indices)
must match exactly; FP64 fields within the free-surface tolerance."""
import numpy as np
import pytest

from tests.helpers import (FIELDS_2D, FIELDS_3D, FIELDS_TURB, INTS, compare, relerr,
                           small_ellipse, small_rectangle)

pytestmark = pytest.mark.gpu

TOL_SOLVE = 1e-12   # both free-surface solvers run to this relative residual
FIELD_TOL = 1e-8    # relative (max-norm) agreement of every FP64 field


def _pair(deck, oracle_lib, precond="jacobi"):
    from si3d_amd.model import GpuLake
    deck = dict(deck)
    deck["pcg_tol"] = TOL_SOLVE
    o = oracle_lib.Oracle(deck)
    g = GpuLake(deck, precision="fp64", preconditioner=precond, tol=TOL_SOLVE)
    return o, g


@pytest.mark.parametrize("mk", [small_rectangle, small_ellipse])
def test_initial_state_matches(mk, oracle_lib):
    deck = mk()
    o, g = _pair(deck, oracle_lib)
    print("case", deck["name"], "lm", o.lm, flush=True)
    assert g.lm == o.lm and g.km1 == o.km1
    for n in INTS:
        np.testing.assert_array_equal(g.get(n), o.get(n).astype(np.int32), err_msg=n)
    for n in ["hhs", "hhu", "hhv"]:
        np.testing.assert_array_equal(g.get(n), o.get(n), err_msg=n)
    compare(g, o, FIELDS_2D + FIELDS_3D, 1e-15, "init")
    if deck["iturb"] > 0:
        compare(g, o, FIELDS_TURB, 1e-15, "init turb")


@pytest.mark.parametrize("mk", [small_rectangle, small_ellipse])
def test_phase_by_phase_first_step(mk, oracle_lib):
    """Step 1 runs trapezoidal fd twice (si3d.f90:339-344); compare after
    every phase of the first fd call."""
    deck = mk()
    o, g = _pair(deck, oracle_lib)
    thrs = deck["dt"] / 3600.0
    checks = {
        2: ["ex"], 3: ["agx", "arx", "eagx", "earx"], 4: ["ex"], 5: ["agy", "ary", "eagy", "eary"],
        6: ["sx", "sy", "qq", "rr"], 7: ["s", "h", "hu", "hv"], 8: ["uh", "vh", "wp"] ,
        9: ["bex"], 10: ["sal"], 11: ["h", "hu", "hv", "uh", "vh"], 13: ["Av", "Dv"],
    }
    for ph in range(14):
        if ph == 12:
            continue
        if ph in (9, 10) and not deck["isal"]:
            continue
        if ph == 13 and deck["iturb"] == 0:
            continue
        if ph == 1 and deck["ihd"] <= 1:
            continue
        o.phase(ph, 1, 2, thrs)
        g.phase(ph, 1, 2, thrs)
        names = checks.get(ph, [])
        if ph == 13:
            names = names + ["q2", "q2l"]
        tol = 1e-11 if ph < 7 else FIELD_TOL
        compare(g, o, names, tol, f"{deck['name']} phase {ph}")
        for n in INTS:
            np.testing.assert_array_equal(g.get(n), o.get(n).astype(np.int32), err_msg=f"phase {ph} {n}")


@pytest.mark.parametrize("mk", [small_rectangle, small_ellipse])
def test_steps_match(mk, oracle_lib):
    deck = mk()
    o, g = _pair(deck, oracle_lib)
    for n in range(1, 6):
        o.step(n)
        g.step(n)
        for name in INTS:
            np.testing.assert_array_equal(g.get(name), o.get(name).astype(np.int32),
                                          err_msg=f"step {n} {name}")
    names = FIELDS_2D + FIELDS_3D + (FIELDS_TURB if deck["iturb"] else [])
    compare(g, o, names, FIELD_TOL, f"{deck['name']} after 5 steps")
    st = g.stats()
    assert st["pcg_relres_last"] <= TOL_SOLVE


def test_line_preconditioner_same_solution(oracle_lib):
    deck = small_ellipse()
    o, g = _pair(deck, oracle_lib, precond="line")
    for n in range(1, 4):
        o.step(n)
        g.step(n)
    compare(g, o, FIELDS_2D + ["uh", "vh", "sal"], FIELD_TOL, "line preconditioner")


@pytest.mark.parametrize("nx,ny", [(40, 24), (160, 96)])
def test_multigrid_preconditioner_same_solution(oracle_lib, nx, ny):
    """The multigrid-preconditioned CG reaches the same free surface and flow
    (one coarse level on the small lake, a four-level hierarchy on the larger)."""
    deck = small_ellipse(nx=nx, ny=ny)
    o, g = _pair(deck, oracle_lib, precond="mg")
    for n in range(1, 4):
        o.step(n)
        g.step(n)
    compare(g, o, FIELDS_2D + ["uh", "vh", "wp", "sal", "q2"], FIELD_TOL, f"multigrid {nx}x{ny}")
    assert g.stats()["pcg_relres_last"] <= TOL_SOLVE


@pytest.mark.parametrize("precond", ["jacobi", "mg"])
def test_mixed_precision_error_vs_fp64_oracle(oracle_lib, precond):
    """Mixed mode (FP32 state, stencils, column solves and inner SpMV; FP64
    zeta, dots, true residual and refinement) stays close to the FP64 oracle."""
    from si3d_amd.model import GpuLake
    deck = small_ellipse()
    deck["pcg_tol"] = 1e-12
    o = oracle_lib.Oracle(deck)
    g = GpuLake(deck, precision="mixed", preconditioner=precond, tol=1e-12)
    for n in range(1, 6):
        o.step(n)
        g.step(n)
    errs = {n: relerr(g.get(n), o.get(n)) for n in ["s", "uh", "vh", "wp", "sal", "q2"]}
    print("mixed-precision relative errors:", errs, flush=True)
    for name in INTS:
        np.testing.assert_array_equal(g.get(name), o.get(name).astype(np.int32), err_msg=name)
    assert errs["s"] < 1e-4 and errs["sal"] < 1e-5 and errs["uh"] < 1e-3, errs
    assert g.stats()["pcg_relres_last"] <= 1e-12


@pytest.mark.parametrize("precision,precond", [("fp64", "mg"), ("mixed", "mg"), ("fp64", "line")])
def test_cg_graph_replays_match_eager_launches(precision, precond):
    """The CG iterations replayed from the captured two-iteration graph give
    bit-identical fields to launching every kernel (S3G_GRAPH=0), step after
    step, with the host's own reads of the state between the steps."""
    import os
    from si3d_amd.model import GpuLake
    deck = small_ellipse(nx=160, ny=96)
    g = GpuLake(deck, precision=precision, preconditioner=precond, tol=1e-12)
    old = os.environ.get("S3G_GRAPH")
    os.environ["S3G_GRAPH"] = "0"
    try:
        e = GpuLake(deck, precision=precision, preconditioner=precond, tol=1e-12)
    finally:
        if old is None:
            del os.environ["S3G_GRAPH"]
        else:
            os.environ["S3G_GRAPH"] = old
    for n in range(1, 7):
        g.step(n)
        e.step(n)
        for name in ["s", "uh", "vh", "sal", "q2"]:
            np.testing.assert_array_equal(g.get(name), e.get(name), err_msg=f"step {n} {name}")
        assert g.stats()["pcg_iters_last"] == e.stats()["pcg_iters_last"]


def test_energy_budget_over_steps_matches_oracle(oracle_lib):
    """The closure's basin energy budget through 6 whole steps of the small
    ellipse (three fd calls per step: the integrals grow on each leapfrog
    call from the sums of the step before), read and cleared after step 3
    as the writer of ScalarBalance.txt does, against the oracle, to 1e-6:
    the dissipation and TKinE weights (rho h) are held in single precision
    between the two sweeps, as the reference computes rho h."""
    from si3d_amd.model import GpuLake
    deck = small_ellipse()
    deck["pcg_tol"] = TOL_SOLVE
    o = oracle_lib.Oracle(deck)
    g = GpuLake(deck, precision="fp64", preconditioner="mg", tol=TOL_SOLVE)
    for n in range(1, 7):
        o.step(n)
        g.step(n)
        want = o.get("energy_budget")
        got = g.energy_balance(reset=(n == 3))
        if n == 3:
            o.set("energy_budget", list(want[:4]) + [0.0, 0.0, 0.0])
        print("step", n, got, flush=True)
        for a, b in zip(got, want[3:]):
            assert a == pytest.approx(b, rel=1e-6, abs=1e-30), (n, got, want[3:])
    assert all(abs(v) > 0 for v in got)


def test_scalar_balance_records_match_oracle(oracle_lib, tmp_path):
    """ScalarBalance.txt written from the GPU fields (outputs.ScalarBalance)
    against the same writer fed the oracle's fields and budget: at n = 0 the
    lake is at rest (kinetic energy and the closure's terms zero) and its
    heat content is sum(T h); after 3 and 6 steps every column agrees to
    1e-8."""
    from si3d_amd import outputs
    from si3d_amd.model import GpuLake
    deck = small_ellipse()
    deck["pcg_tol"] = TOL_SOLVE
    o = oracle_lib.Oracle(deck)
    g = GpuLake(deck, precision="fp64", preconditioner="mg", tol=TOL_SOLVE)

    class OracleView:
        lm = o.lm

        def get(self, name):
            return o.get(name)

        def energy_balance(self, reset=False):
            v = o.get("energy_budget")
            if reset:
                o.set("energy_budget", list(v[:4]) + [0.0, 0.0, 0.0])
            return tuple(v[3:])

    wg = outputs.ScalarBalance(g, deck, directory=str(tmp_path))
    (tmp_path / "oracle").mkdir()
    wo = outputs.ScalarBalance(OracleView(), deck, directory=str(tmp_path / "oracle"))
    r0 = wg.values(0, 0.0)
    h, sal, k1z, kmz = g.get("h"), g.get("sal"), g.get("k1z").astype(int), g.get("kmz").astype(int)
    heat = sum(float(np.dot(sal[l, k1z[l] - 1:kmz[l]], h[l, k1z[l] - 1:kmz[l]])) for l in range(g.lm))
    assert r0[1] == pytest.approx(heat, rel=1e-14) and r0[4] == 0.0 and r0[5:] == [0.0] * 4
    np.testing.assert_allclose(r0, wo.values(0, 0.0), rtol=1e-12)
    for n in range(1, 7):
        o.step(n)
        g.step(n)
        if n % 3 == 0:
            a, b = wg.values(n, n * deck["dt"] / 3600.0), wo.values(n, n * deck["dt"] / 3600.0)
            print("step", n, a, flush=True)
            np.testing.assert_allclose(a[:5], b[:5], rtol=1e-8)
            np.testing.assert_allclose(a[5:], b[5:], rtol=1e-6)
            assert a[4] > 0 and a[5] > 0
    wg.write(6, 0.1)
    wg.close()
    wo.close()
    assert len(open(tmp_path / "ScalarBalance.txt").readline()) == 181


def test_section_files_and_run_log_from_the_gpu(oracle_lib, tmp_path):
    """section_<j> files written from the GPU fields against the same writer
    fed the oracle's fields after 3 steps (float32 records, the fields agree
    to the solver tolerance), and si3d_out.txt from a run driven call by call:
    one Enter/converged/Exit block per free-surface solve (2 on step 1, 3 on
    each later step) with the GPU's iteration counts."""
    from si3d_amd import outputs
    from si3d_amd.model import GpuLake
    deck = small_ellipse()
    deck["pcg_tol"] = TOL_SOLVE
    o = oracle_lib.Oracle(deck)
    g = GpuLake(deck, precision="fp64", preconditioner="mg", tol=TOL_SOLVE)
    ii, jj = outputs.wet_columns(deck)
    secs = [([int(ii[x]) for x in (100, 101, 102, 103)], [int(jj[x]) for x in (100, 101, 102, 103)]),
            ([int(ii[200])], [int(jj[200])])]
    (tmp_path / "gpu").mkdir()
    (tmp_path / "oracle").mkdir()
    log = outputs.RunLog(deck, directory=str(tmp_path / "gpu"))
    for n in range(1, 4):
        o.step(n)
        outputs.step_logged(g, deck, n, log)
    log.close()
    for src, sub in ((g, "gpu"), (o, "oracle")):
        sf = outputs.SectionFiles(src, deck, secs, nts=10, iox=3, directory=str(tmp_path / sub))
        sf.write(0)
        sf.write(3)
        sf.close()
    for s in (1, 2):
        a = (tmp_path / "gpu" / f"section_{s}").read_bytes()
        b = (tmp_path / "oracle" / f"section_{s}").read_bytes()
        assert len(a) == len(b)
        x, y = np.frombuffer(a, np.float32), np.frombuffer(b, np.float32)
        scale = np.maximum(np.abs(y), 1e-6)
        assert np.max(np.abs(x - y) / scale) < 1e-5, s
    text = (tmp_path / "gpu" / "si3d_out.txt").read_text().splitlines()
    enters = [l for l in text if l.startswith("**Enter nspcg")]
    done = [int(l.split()[4]) for l in text if "has converged in" in l]
    assert len(enters) == 2 + 3 + 3 and len(done) == 8 and all(0 < k < 200 for k in done)
