"""Writers of the node time files and plane files (si3d_amd/outputs.py): the
Fortran edit descriptors of their records, and what a run's files hold, read
back field by field (the writers take any object with the library's `get`
interface; the oracle provides one on the CPU)."""
import numpy as np
import pytest

from si3d_amd import lake, outputs as O
from tests.helpers import small_ellipse


@pytest.mark.parametrize("args,expected", [
    ((12.3456, 10, 4), "   12.3456"), ((0.012345, 9, 2, 2), "     1.23"), ((-0.5, 9, 2), "    -0.50"),
    ((1e-5, 15, 7, 4), "      0.1000000"), ((123456.0, 5, 2), "*****"), ((-99e-2, 10, 2, 2), "    -99.00")])
def test_fortran_f(args, expected):
    assert O.fortran_f(*args) == expected


@pytest.mark.parametrize("x,expected", [
    (12.345, "    0.12345E+02"), (-0.00123456, "   -0.12346E-02"), (0.0, "    0.00000E+00"),
    (1.0, "    0.10000E+01"), (0.999999, "    0.10000E+01"), (1e-120, "    0.10000-119"), (-99.0, "   -0.99000E+02")])
def test_fortran_e(x, expected):
    assert O.fortran_e(x, 15, 5) == expected


def _run(oracle_lib, nsteps=3):
    d = small_ellipse(nts=10)
    o = oracle_lib.Oracle(d)
    for n in range(1, nsteps + 1):
        o.step(n)
    return d, o


def test_time_file_records(oracle_lib, tmp_path):
    d, o = _run(oracle_lib)
    ii, jj = O.wet_columns(d)
    node = (int(ii[150]), int(jj[150]))
    tf = O.TimeFiles(o, d, [node], directory=str(tmp_path), title="test lake")
    tf.write(3, 3 * d["dt"] / 3600.0)
    tf.close()
    lines = open(tmp_path / f"tf{node[0]}_{node[1]}.txt").read().splitlines()
    assert lines[0] == "test lake" and lines[2].startswith(f"i ={node[0]:4d}  j ={node[1]:4d}")
    rec = lines[7:]
    l = 150
    kmz = int(o.get("kmz")[l])
    assert len(rec) == kmz  # layers k1 .. kmz+1, one record each
    assert float(rec[0][1:11]) == pytest.approx(3 * d["dt"] / 3600.0, abs=1e-4)
    assert int(rec[0][11:21]) == 3
    assert float(rec[0][21:30]) == pytest.approx(100 * o.get("s")[l], abs=0.006)
    u, av, sal, h = o.get("u"), o.get("Av"), o.get("sal"), o.get("h")
    for r, k in zip(rec, range(2, kmz + 2)):
        body = r[30:]
        assert r.startswith(" ") and (r is rec[0] or r[:30] == " " * 30)
        fields = [body[0:9], body[9:19], body[19:29], body[29:38], body[38:53], body[53:68], body[68:83]]
        if h[l, k - 1] <= 0:
            assert float(fields[1]) == -99.0
            continue
        assert float(fields[1]) == pytest.approx(100 * u[l, k - 1], abs=0.006)
        assert float(fields[4]) == pytest.approx(1e4 * 0.5 * (av[l, k - 1] + av[l, k]), abs=6e-8)
        assert float(fields[6]) == pytest.approx(sal[l, k - 1], rel=5e-5)  # E15.5: five digits


def _records(path):
    raw = open(path, "rb").read()
    out, p = [], 0
    while p < len(raw):
        n = int(np.frombuffer(raw[p:p + 4], np.int32)[0])
        out.append(raw[p + 4:p + 4 + n])
        assert int(np.frombuffer(raw[p + 4 + n:p + 8 + n], np.int32)[0]) == n
        p += n + 8
    return out


def test_plane_file_records(oracle_lib, tmp_path):
    d, o = _run(oracle_lib)
    kp = 4
    pf = O.PlaneFiles(o, d, [kp], nts=10, iop=1, directory=str(tmp_path), start=(2020, 5, 5, 1200))
    pf.write(0)
    pf.write(3)
    pf.close()
    recs = _records(tmp_path / f"plane_{kp}")
    assert len(recs) == 4
    assert np.frombuffer(recs[0], np.int32)[0] == 10
    kmz = o.get("kmz")[:o.lm].astype(int)
    cells = np.nonzero(kmz >= kp)[0]
    assert np.frombuffer(recs[1], np.int32)[0] == len(cells)
    first = recs[2]
    ints = np.frombuffer(first[:16], np.int32)
    assert list(ints) == [0, 2020, 5, 5]
    assert np.frombuffer(first[16:20], np.float32)[0] == 1200.0
    data = np.frombuffer(first[20:], np.float32).reshape(len(cells), 8)
    ii, jj = O.wet_columns(d)
    np.testing.assert_array_equal(data[:, 0], ii[cells])
    np.testing.assert_array_equal(data[:, 1], jj[cells])
    np.testing.assert_array_equal(data[:, 2], np.float32(o.get("up")[cells, kp - 1]))
    np.testing.assert_array_equal(data[:, 7], np.float32(o.get("s")[cells]))
    later = recs[3]
    ints = np.frombuffer(later[:16], np.int32)
    assert list(ints) == [3, 2020, 5, 5]
    hour = 12.0 + 3 * d["dt"] / 3600.0
    assert np.frombuffer(later[16:20], np.float32)[0] == np.float32(round(hour * 100))
    data = np.frombuffer(later[20:], np.float32).reshape(len(cells), 6)
    sal = o.get("salp")[cells, kp - 1]
    np.testing.assert_array_equal(data[:, 3], np.float32(np.where(o.get("hp")[cells, kp - 1] <= 0, -99.0, sal)))
    wp = o.get("wp")
    np.testing.assert_allclose(data[:, 2], (wp[cells, kp - 1] + wp[cells, kp]) / 2, rtol=1e-6, atol=1e-12)


@pytest.mark.parametrize("turb", [False, True])
def test_three_d_file_records(oracle_lib, tmp_path, turb):
    """si3d_3D: frame and point counts, then one record per frame with the
    stamp and every wet cell from layer 2 to kmz + 1 of every column in l
    order (i, j, k first; the fields as REAL*4)."""
    d, o = _run(oracle_lib)
    f3 = O.ThreeDFile(o, d, nts=10, ipxml=2, turb_vars=turb, directory=str(tmp_path), start=(2021, 3, 4, 600))
    f3.write(0)
    f3.write(2)
    f3.close()
    recs = _records(tmp_path / "si3d_3D")
    assert len(recs) == 4
    assert np.frombuffer(recs[0], np.int32)[0] == 5
    kmz = o.get("kmz")[:o.lm].astype(int)
    npts = int(kmz.sum())  # layers 2 .. kmz + 1
    assert np.frombuffer(recs[1], np.int32)[0] == npts
    nv = 13 if turb else 8
    first = recs[2]
    assert list(np.frombuffer(first[:16], np.int32)) == [0, 2021, 3, 4]
    assert np.frombuffer(first[16:20], np.float32)[0] == 600.0
    data = np.frombuffer(first[20:], np.float32).reshape(npts, nv)
    ii, jj = O.wet_columns(d)
    l, k = 37, 2  # the first point of column 37 is its layer 2
    row = int(kmz[:l].sum())
    assert data[row, 0] == ii[l] and data[row, 1] == jj[l] and data[row, 2] == k
    assert data[row, 3] == np.float32(o.get("hp")[l, k - 1])
    last = row + kmz[l] - 1  # its layer kmz + 1, the cell below the bed
    assert data[last, 2] == kmz[l] + 1
    assert data[last, 6] == np.float32(o.get("wp")[l, kmz[l]])
    later = np.frombuffer(recs[3][20:], np.float32).reshape(npts, nv - 3)
    assert list(np.frombuffer(recs[3][:16], np.int32)) == [2, 2021, 3, 4]
    np.testing.assert_array_equal(later[:, 0], data[:, 3])
    sal = o.get("salp")
    np.testing.assert_array_equal(later[row:last + 1, 4], np.float32(sal[l, 1:kmz[l] + 1]))
    if turb:
        np.testing.assert_array_equal(later[row:last + 1, 9], np.float32(o.get("Av")[l, 1:kmz[l] + 1]))



def test_hour_stamp_rounds_halves_away_from_zero():
    """NINT(hrs*100.) of compute_date: 90 s is 2.5 hundredths of an hour and
    stamps 3 (Python's round() would give 2); whole hours stamp 100 per hour."""
    from si3d_amd.outputs import hour_stamp
    assert hour_stamp(90) == 3
    assert hour_stamp(54) == 2          # 1.5 -> 2
    assert hour_stamp(18) == 1          # 0.5 -> 1
    assert hour_stamp(3600) == 100
    assert hour_stamp(86340) == 2398    # 23.98333 h
    assert hour_stamp(0) == 0


def test_density_matches_oracle_eos(oracle_lib):
    """The writer's vectorised densty_s against the oracle's (itself pinned by
    the UNESCO check values and a root finder) over temperatures and depths."""
    from si3d_amd.outputs import density
    T = np.linspace(0.0, 30.0, 31)
    Z = np.array([0.0, 0.25, 5.0, 50.0, 120.0])
    got = density(T[:, None], Z[None, :])
    want = np.array([[oracle_lib.densty(t, 0.00004, z) for z in Z] for t in T])
    np.testing.assert_allclose(got, want, rtol=1e-13, atol=0)
    assert abs(density(5.0, 0.0, 0.0) - 999.96675) < 1e-5


def test_scalar_balance_record_format(tmp_path):
    """A ScalarBalance.txt record: nine E20.11 fields."""
    from si3d_amd.outputs import fortran_e
    line = "".join(fortran_e(v, 20, 11) for v in [0.0, 1.5, 0.0, -2.5e10, 3.25e-3, 0.0, 1e-5, -2e-6, 4e-6])
    assert len(line) == 180
    assert line[:20] == "   0.00000000000E+00" and line[20:40] == "   0.15000000000E+01"


def test_west_south_neighbours_match_oracle(oracle_lib):
    from si3d_amd.outputs import west_south
    from tests.helpers import small_ellipse
    d = small_ellipse()
    o = oracle_lib.Oracle(d)
    lW, lS = west_south(d)
    np.testing.assert_array_equal(lW, o.get("lWC")[:o.lm].astype(int) - 1)
    np.testing.assert_array_equal(lS, o.get("lSC")[:o.lm].astype(int) - 1)


# ---- si3d_out.txt, si3d_log.txt, sections, spacefile.xml --------------------------
FLANG = "/opt/rocm/lib/llvm/bin/flang"


def _fortran_output(tmp_path, body):
    """Compile and run a small Fortran program (this test's own) and return
    what it prints: flang's formatted output is the golden reference for the
    edit descriptors the writers restate."""
    import os
    import shutil
    import subprocess
    fc = FLANG if os.path.exists(FLANG) else shutil.which("flang")
    if not fc:
        pytest.skip("no Fortran compiler in this image")
    src = tmp_path / "prog.f90"
    src.write_text("program prog\n" + body + "\nend program prog\n")
    exe = tmp_path / "prog"
    subprocess.run([fc, str(src), "-o", str(exe)], check=True, capture_output=True, cwd=tmp_path)
    return subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout


def test_fortran_g_matches_flang(tmp_path):
    vals = [60000.0, 86400.0, 60.0, 30.0, 0.5, 0.25, 123.4, 99.4, 99.6, 9.96, 0.04, 0.1, 0.0996, 1.0, -7.5, -250.0]
    body = "  real :: v(%d) = (/%s/)\n  integer :: i\n  do i = 1, %d\n    write (*, '(G9.2)') v(i)\n  end do" % (
        len(vals), ", ".join(f"{v!r}" for v in vals), len(vals))
    want = _fortran_output(tmp_path, body).splitlines()
    assert [O.fortran_g(v, 9, 2) for v in vals] == want


def _inp_deck():
    d = lake.rectangle(nx=20, ny=10, dx=50.0, depth=8.0, dz=1.0, nts=120)
    d.update(beta=0.05, Av0=0.0125, Dv0=0.0025, Ax0=1.5, Ay0=0.75, f=1e-4, theta=0.6, cd=0.0025, cw=0.0013,
             wa=4.25, phi=240.0, dtsurfbc=600.0, dzmin=0.1, zetainit=0.0)
    return d


def test_run_log_header_matches_flang(tmp_path):
    """The parameter block of si3d_out.txt against flang writing the same
    values through outr's two formats (si3d_utils.f90:981-1015), typed as
    the reference declares them (REAL and INTEGER)."""
    import datetime
    d = _inp_deck()
    p = O.run_parameters(d, title="test run", start=(2001, 3, 4, 1230))
    log = O.RunLog(d, directory=str(tmp_path), title="test run", start=(2001, 3, 4, 1230),
                   now=datetime.datetime(2026, 12, 4, 9, 8, 7))
    log.close()
    text = (tmp_path / "si3d_out.txt").read_text().splitlines()
    assert text[0] == "test run"
    assert text[1] == "Run number = 202612040908,  Start date of run:   3/ 4/2001 at 1230 hours"
    r = {k: repr(float(p[k])) for k in p if isinstance(p[k], (float, int)) and not isinstance(p[k], bool)}
    ints = ["itrap", "niter", "ismooth", "iturb", "iadv", "itrmom", "ihd", "ibc", "isal", "itrsca", "nopen",
            "ifsurfbc", "idbg"]
    body = "\n".join([
        "  real :: xl=%s, yl=%s, zl=%s, idx=%s, idy=%s, idz=%s, tl=%s, idt=%s, dzmin=%s, zetainit=%s" % tuple(
            r[k] for k in ("xl", "yl", "zl", "idx", "idy", "idz", "tl", "idt", "dzmin", "zetainit")),
        "  real :: beta=%s, Av0=%s, Dv0=%s, Ax0=%s, Ay0=%s, f=%s, theta=%s, cd=%s, dtsurfbc=%s, cw=%s, wa=%s, "
        "phi=%s" % tuple(r[k] for k in ("beta", "Av0", "Dv0", "Ax0", "Ay0", "f", "theta", "cd", "dtsurfbc", "cw",
                                        "wa", "phi")),
        "  integer :: " + ", ".join(f"{k}={int(p.get(k, 0))}" for k in ints),
        "  write (*, 1) xl,yl,zl,idx,idy,idz,tl,idt,dzmin,zetainit",
        "  write (*, 2) itrap,niter,ismooth,beta,iturb,Av0,Dv0,iadv,itrmom,ihd,Ax0,Ay0,f,theta, &",
        "    & ibc,isal,itrsca,nopen,cd,ifsurfbc,dtsurfbc,cw,wa,phi,idbg",
        "1 FORMAT (/' xl=   ',F9.1,/' yl=   ',F9.1,/' zl=   ',F9.3,/' idx=  ',F9.3,/' idy=  ',F9.3,/ &",
        "   &' idz=  ',F9.3,/' tl=   ',G9.2,/' idt=  ',G9.2,/' dzmin=',F9.2,/' zeta0=',F9.2/)",
        "2 FORMAT (/' itrap= ',I9,/' niter= ',I9,/' smooth=',I9,/' beta=  ',F9.4,/' iturb= ',I9,/ &",
        "   &' Av0=   ',F9.4,/' Dv0=   ',F9.4,/' iadv=  ',I9,/' trmom= ',I9,/' ihd=   ',I9,/ &",
        "   &' Ax0=   ',F9.4,/' Ay0=   ',F9.4,/' f=     ',F9.4,/' theta= ',F9.4,/' ibc=   ',I9,/ &",
        "   &' isal=  ',I9,/' trsal= ',I9,/' nopen= ',I9,/' cd=    ',F9.4,/' isbc=  ',I9,/ &",
        "   &' tsbc=  ',F9.2,/' cw=    ',F9.2,/' wa=    ',F9.2,/' phi=   ',F9.2,/' idbg=  ',I9/)"])
    want = _fortran_output(tmp_path, body).splitlines()
    assert text[2:] == want


def test_run_log_solver_lines(tmp_path):
    d = _inp_deck()
    log = O.RunLog(d, directory=str(tmp_path))
    log.solve(7, 15)
    log.solve(8, 4000, ier=1)
    log.close()
    lines = (tmp_path / "si3d_out.txt").read_text().splitlines()
    assert lines[-3:] == ["**Enter nspcg   n =      8", "**Exit nspcg    n =      8", "**ERROR    1 from nspcg"]
    assert lines[-7:-3] == ["**Enter nspcg   n =      7", "", " cg  has converged in    15 iterations",
                            "**Exit nspcg    n =      7"]


def test_log_file_holds_the_latest_line_and_goes_at_the_end(tmp_path):
    lf = O.LogFile(nts=3, directory=str(tmp_path))
    lf.write(1, 0.5 / 60)
    lf.write(2, 1.0 / 60)
    assert (tmp_path / "si3d_log.txt").read_text() == " time=    0.0167 hours    n=     2\n"
    lf.write(3, 1.5 / 60)
    assert not (tmp_path / "si3d_log.txt").exists()


def test_section_file_records(oracle_lib, tmp_path):
    """Two sections, one crossing a dry cell (skipped): the point counts are
    the wet layers k1..kmz of their columns, the first frame carries (i, j,
    node, depth of the layer top) and the fields, the later ones the fields."""
    d, o = _run(oracle_lib)
    ii, jj = O.wet_columns(d)
    c = [100, 101, 102]
    sec1 = ([int(ii[x]) for x in c], [int(jj[x]) for x in c])
    sec2 = ([int(ii[50]), 1], [int(jj[50]), 1])  # (1, 1) is a dry corner
    sf = O.SectionFiles(o, d, [sec1, sec2], nts=10, iox=2, directory=str(tmp_path), start=(2020, 5, 5, 0))
    sf.write(0)
    sf.write(3)
    sf.close()
    kmz = o.get("kmz").astype(int)
    z = O.zlevels(d)
    uhp, wp, salp, hp, av, dv = (o.get(x) for x in ("uhp", "wp", "salp", "hp", "Av", "Dv"))
    for s, cols in ((1, c), (2, [50])):
        recs = _records(tmp_path / f"section_{s}")
        assert len(recs) == 4
        assert np.frombuffer(recs[0], np.int32)[0] == 5
        npts = sum(kmz[l] - 1 for l in cols)
        assert list(np.frombuffer(recs[1], np.int32)) == [npts, lake.km1_of(d)]
        first = np.frombuffer(recs[2][20:], np.float32).reshape(npts, 10)
        later = np.frombuffer(recs[3][20:], np.float32).reshape(npts, 6)
        assert list(np.frombuffer(recs[3][:16], np.int32)) == [3, 2020, 5, 5]
        row = 0
        for m, l in enumerate(cols, start=1):
            for k in range(2, kmz[l] + 1):
                kk = k - 1
                assert (first[row, 0], first[row, 1], first[row, 2]) == (ii[l], jj[l], m)
                assert first[row, 3] == np.float32(z[k] - hp[l, kk])
                assert first[row, 4] == np.float32(uhp[l, kk]) and later[row, 2] == np.float32(wp[l, kk])
                sal = -99.0 if hp[l, kk] <= 0 else salp[l, kk]
                assert first[row, 7] == np.float32(sal) and later[row, 3] == np.float32(sal)
                assert later[row, 4] == np.float32(0.5 * (av[l, kk] + av[l, kk + 1]))
                assert later[row, 5] == np.float32(0.5 * (dv[l, kk] + dv[l, kk + 1]))
                row += 1


def test_spacefile_xml_matches_flang(tmp_path):
    """spacefile.xml's run-parameter comment against flang writing outs's
    format (si3d_utils.f90:2661-2665); the first-snapshot date, the
    (negative) time step and the empty binary companions."""
    import datetime
    d = _inp_deck()
    x = O.SpaceFileXml(d, ipxml=-4, itspf=0, directory=str(tmp_path), title="test run", start=(2001, 3, 4, 1230),
                       now=datetime.datetime(2026, 12, 4, 9, 8, 7))
    x.write(0)
    x.write(4)
    lines = (tmp_path / "spacefile.xml").read_text().splitlines()
    body = "\n".join([
        f"  real :: idt={float(d['dt'])!r}, ddz={float(d['ddz'])!r}, cd={float(d['cd'])!r}, beta={float(d['beta'])!r}, "
        f"f={float(d['f'])!r}, wa={float(d['wa'])!r}",
        f"  integer :: im={int(d['imm']) + 1}, jm={int(d['jmm']) + 1}, itrsca=4, itrap={int(d['itrap'])}, "
        f"ismooth={int(d['ismooth'])}, niter={int(d['niter'])}, itrmom={int(d['itrmom'])}",
        "  write (*, 1) idt,ddz,im,jm,itrsca,itrap,cd,ismooth,beta,niter,itrmom,f,wa",
        '1 FORMAT("<!-- ", "Run Parameters:  dt =",F5.2," s", "  dz =",F5.2, " m",  "  im =", I5, "  jm =", I5, &',
        '   & /"     itrsca =", I2, "  itrap =", I2, "  cd = ", F7.4, "  ismooth =", I2, "  beta =", F6.3, &',
        '   & "  niter =", I2/"     itrmom =", I2, "  f =", F7.4, "  wa =", F8.1, " -->" )'])
    want = _fortran_output(tmp_path, body).splitlines()
    assert lines[4:7] == want
    assert lines[0] == '<?xml version="1.0"?>' and lines[-2] == "</gov.usgs.gr>"
    assert '        runTitle="test run"' in lines and '        runNumber="202612040908"' in lines
    start = [s for s in lines if "startTime=" in s][0]
    assert start == f'        startTime="2001/03/04 12:18:00" timeStepInSeconds="{int(-4 * d["dt"])}">'
    for name in ("spacefile2d.bin", "spacefile3d.bin"):
        assert (tmp_path / name).read_bytes() == b""
