"""Closed-form pins of the oracle's vertical momentum system (matmom with
trid and bclnc_km), the TVD scalar transport (exsal), the implicit vertical
diffusion of the scalar (imsal with trid1) and the upwind momentum advection
schemes (exmom, itrmom 2-4).

Each test sets a synthetic state through the oracle's field setter, runs one
phase and compares with what the mathematics fixes for that state:

* a cosine mode cos(pi m (k - 1/2) / N) of a column of N equal layers is an
  eigenvector of the no-flux second difference, with eigenvalue
  -(2 - 2 cos(pi m / N)), so an implicit theta-scheme multiplies it by
  (1 - (1 - theta) L) / (1 + theta L), L = dt' K (2 - 2 cos(pi m / N)) / h^2;
* a column whose eddy coefficient vanishes on one interface splits in two
  columns that exchange nothing, so each part keeps its own content;
* a horizontal density gradient gamma gives the hydrostatic pressure
  gradient g gamma z at depth z, whatever the bed does (the reference's
  bottom-layer correction interpolates the density to a common depth);
* the Lax-Wendroff flux (superbee limiter at 1) and the Beam-Warming flux
  (limiter = r) both advect a quadratic profile exactly;
* first-order upwind and centred differences are exact on linear fields.
"""
import math

import numpy as np
import pytest

from si3d_amd import lake
from tests.test_oracle_pins import _full, _grid, _interior, _positions

G = 9.806


def _column_lake(theta=0.6, Av0=0.02, nx=8, ny=6, depth=12.0, stretched=False):
    d = lake.rectangle(nx=nx, ny=ny, dx=40.0, depth=depth, dz=1.0, dt=60.0, nts=10, iadv=0, ihd=0, cd=0.0,
                       Av0=Av0, tau=0.0)
    d["theta"] = theta
    d["f"] = 0.0
    d["dy"] = 55.0
    if stretched:  # layers 0.4 m at the top to 2.5 m at the bottom
        z = lake.stretched_levels(10, 0.4, 2.5)
        d["ibathyf"], d["zlevel"], d["zl"] = -1, z, float(z[-1])
        d["bathy_dm"] = np.full((nx, ny), int(round(10 * (z[-1] - 0.3))), dtype=np.int32)
        d["init_profile"] = np.full(len(z), 20.0)
    return d


def _calm(o):
    """No wind, no surface drag (cdw = 0), flat free surface."""
    o.set("uair", np.zeros(o.lm1))
    o.set("vair", np.zeros(o.lm1))
    o.set("cdw", np.zeros(o.lm1))
    for n in ("s", "sp", "spp"):
        o.set(n, np.zeros(o.lm1))


def _cosine(N, m, k1, km1):
    """cos(pi m (k - k1 + 1/2) / N) on layers k1..k1+N-1 (1-based), zero elsewhere."""
    v = np.zeros(km1)
    for k in range(k1, k1 + N):
        v[k - 1] = math.cos(math.pi * m * (k - k1 + 0.5) / N)
    return v


# ------------------------------------------------------------------ matmom

@pytest.mark.parametrize("X", [True, False])
@pytest.mark.parametrize("m", [1, 3])
def test_matmom_viscous_decay_of_cosine_mode(oracle_lib, X, m):
    """matmom (si3d_procedures.f90:1305-1344, trid 3210-3289) on a column
    of N = 12 one-metre layers with constant Av, theta = 0.6, no wind,
    bottom drag, pressure gradient or baroclinicity, ex = the n-1
    transport: the new transport of a cosine velocity mode is the mode times
    (1 - (1 - theta) L) / (1 + theta L), L = 2 dt Av (2 - 2 cos(pi m / N)) /
    h^2 (leapfrog, dt' = 2 dt)."""
    d = _column_lake()
    o = oracle_lib.Oracle(d)
    g = _grid(o)
    _calm(o)
    hq = o.get("hup" if X else "hvp")
    k1 = 2
    N = int(g["kmz"][0]) - k1 + 1
    assert N == 12
    mode = 0.03 * _cosine(N, m, k1, o.km1)
    q = np.ones((o.lm1, 1)) * mode[None, :]
    o.set("Av", _full(o, d["Av0"]))
    o.set("upp" if X else "vpp", q)
    o.set("uhpp" if X else "vhpp", q * hq)
    o.set("ex", q * hq)
    o.phase(3 if X else 5, 2, 1)
    ag = o.get("agx" if X else "agy")
    dt2, th = 2.0 * d["dt"], d["theta"]
    L = dt2 * d["Av0"] * (2 - 2 * math.cos(math.pi * m / N)) / 1.0 ** 2
    factor = (1 - (1 - th) * L) / (1 + th * L)
    # the viscosity and the theta weighting both matter here: fully implicit
    # or with theta and 1 - theta swapped the factor differs by > 1e-3
    assert abs(factor - 1 / (1 + L)) > 1e-3 and abs(factor - (1 - th * L) / (1 + (1 - th) * L)) > 1e-3
    nb = g["lEC"] if X else g["lNC"]
    checked = 0
    for l in range(o.lm):
        if nb[l] >= o.lm:
            continue
        for k in range(k1, k1 + N):
            assert ag[l, k - 1] == pytest.approx(factor * q[l, k - 1] * hq[l, k - 1], rel=1e-11, abs=1e-15), (l, k)
            checked += 1
    assert checked > 300


@pytest.mark.parametrize("X", [True, False])
def test_matmom_split_column_keeps_each_part(oracle_lib, X):
    """With Av = 0 on the interface above layer ks only (Av(ks) is the
    coefficient between layers ks-1 and ks, si3d_procedures.f90:1305-1344),
    the layers above and below exchange no momentum: each part keeps its
    own depth-integrated transport (the internal viscosity only moves it
    around within the part).  Layers of unequal thickness, so that each
    interface's geometry must be the same in the rows above and below it."""
    d = _column_lake(theta=1.0, Av0=0.05, stretched=True)
    o = oracle_lib.Oracle(d)
    g = _grid(o)
    _calm(o)
    hq = o.get("hup" if X else "hvp")
    ks = 6
    Av = _full(o, d["Av0"])
    Av[:, ks - 1] = 0.0
    o.set("Av", Av)
    rng = np.random.default_rng(4)
    q = np.ones((o.lm1, 1)) * rng.uniform(-0.05, 0.05, o.km1)[None, :]
    o.set("upp" if X else "vpp", q)
    o.set("uhpp" if X else "vhpp", q * hq)
    o.set("ex", q * hq)
    o.phase(3 if X else 5, 2, 1)
    ag = o.get("agx" if X else "agy")
    nb = g["lEC"] if X else g["lNC"]
    for l in range(o.lm):
        if nb[l] >= o.lm:
            continue
        km = g["kmz"][l]
        up0, up1 = (q[l, 1:ks - 1] * hq[l, 1:ks - 1]).sum(), ag[l, 1:ks - 1].sum()
        lo0, lo1 = (q[l, ks - 1:km] * hq[l, ks - 1:km]).sum(), ag[l, ks - 1:km].sum()
        assert up1 == pytest.approx(up0, rel=1e-12) and lo1 == pytest.approx(lo0, rel=1e-12), l
        assert np.max(np.abs(ag[l, 1:km] - q[l, 1:km] * hq[l, 1:km])) > 1e-4  # viscosity did act
        assert np.ptp(hq[l, 1:km]) > 0.5


@pytest.mark.parametrize("X", [True, False])
def test_matmom_baroclinic_pressure_gradient(oracle_lib, X):
    """From rest, with no viscosity, the one-step change of each layer's
    transport is the hydrostatic baroclinic pressure gradient of a density
    rho = rho0 + gamma x (or y) + beta z: uh = -2 dt (h / rho_m) g gamma z
    with z the depth of the layer's centre along the face and rho_m the
    column-mean density the reference divides by (si3d_procedures.f90:
    1277-1301).  The bed deepens along the gradient, so the deeper column's
    bottom layer is thicker than the face's and bclnc_km (1633-1746) must
    interpolate its density to the face's depth for the result to hold."""
    d = lake.rectangle(nx=10, ny=10, dx=40.0, depth=8.0, dz=1.0, dt=60.0, nts=10, iadv=0, ihd=0, cd=0.0,
                       Av0=0.0, tau=0.0)
    d["f"] = 0.0
    d["ibc"] = 1
    d["dy"] = 40.0
    i, j = np.meshgrid(np.arange(10), np.arange(10), indexing="ij")
    ramp = i if X else j
    d["bathy_dm"] = (53 + 7 * ramp).astype(np.int32)  # 5.3 m ... 11.6 m
    d["zl"] = 12.0
    d["init_profile"] = np.full(lake.km1_of(d), 15.0)
    o = oracle_lib.Oracle(d)
    g = _grid(o)
    _calm(o)
    o.set("Av", _full(o, 0.0))
    hp = o.get("hp")
    hq = o.get("hup" if X else "hvp")
    gamma, beta, rho0 = 2.0e-4, 1.5e-2, -1.2
    pos = (g["l2i"] if X else g["l2j"]).astype(float) * d["dx"]
    zc = np.zeros_like(hp)
    for l in range(o.lm1):
        top = 0.0
        for k in range(1, o.km1):
            zc[l, k] = top + hp[l, k] / 2
            top += hp[l, k]
    rho = rho0 + gamma * pos[:, None] + beta * zc
    o.set("rhop", rho)
    o.phase(3 if X else 5, 2, 1)
    ag = o.get("agx" if X else "agy")
    nb = g["lEC"] if X else g["lNC"]
    k1 = 2
    corrected = 0
    for l in range(o.lm):
        e = nb[l]
        if e >= o.lm:
            continue
        kmq = min(g["kmz"][l], g["kmz"][e])
        rho_m = np.mean(1000.0 + rho[l, k1 - 1:kmq])
        z = np.cumsum(hq[l, k1 - 1:kmq]) - hq[l, k1 - 1:kmq] / 2
        expect = -2 * d["dt"] * hq[l, k1 - 1:kmq] / rho_m * G * gamma * z
        np.testing.assert_allclose(ag[l, k1 - 1:kmq], expect, rtol=1e-10, atol=1e-16)
        corrected += int(abs(hp[e, kmq - 1] - hq[l, kmq - 1]) > 1e-3)
    assert corrected > 20  # faces whose bottom layer needed the interpolation


# ------------------------------------------------------------------- exsal

def _transport_lake(nx=24, ny=20):
    d = lake.rectangle(nx=nx, ny=ny, dx=50.0, depth=6.0, dz=1.0, dt=60.0, nts=10, iadv=1, ihd=0, tau=0.0)
    d["isal"] = 1
    d["ibc"] = 0
    d["dy"] = 40.0
    d["Ax0"] = d["Ay0"] = 0.0
    d["f"] = 0.0
    return d


@pytest.mark.parametrize("X", [True, False])
@pytest.mark.parametrize("regime", ["lax_wendroff", "beam_warming"])
def test_exsal_advects_a_quadratic_exactly(oracle_lib, X, regime):
    """exsal + imsal (si3d_procedures.f90:2885-3067, 3079-3207) with a
    uniform flow U along x (or y), no vertical motion or diffusion, and a
    scalar c (x - x0)^2 (the flow being the mean of the n+1 and n-1
    transports): where the superbee limiter is 1 (upwind ratio r in
    [1/2, 1], the increasing branch) the flux is Lax-Wendroff's, where it is
    r (r in [1, 2], the decreasing branch) Beam-Warming's, and both move a
    quadratic exactly: the new scalar is c (x - x0 - U dt')^2 with
    dt' = 2 dt.  Checked on the cells whose four-point stencils all sit in
    the chosen branch."""
    d = _transport_lake()
    o = oracle_lib.Oracle(d)
    g = _grid(o)
    sp = d["dx"] if X else d["dy"]
    pos = (g["l2i"] if X else g["l2j"]).astype(float) * sp
    U = 0.12
    dt2 = 2.0 * d["dt"]
    n = (d["imm"] if X else d["jmm"])
    if regime == "lax_wendroff":
        x0, cq = pos.min() - 4 * sp, 1.0e-3      # increasing, r = (2(x-x0) - dx)/(2(x-x0) + dx) in [1/2, 1)
    else:
        x0, cq = pos.max() + 4 * sp + (n + 3) * sp, 1.0e-3   # decreasing toward x0 > x: r in (1, 2]
    s = 20.0 + cq * (pos - x0) ** 2
    S = np.ones((o.lm1, 1)) * np.ones(o.km1)[None, :] * s[:, None]
    for name in ("sal", "salp", "salpp"):
        o.set(name, S)
    hq, hn = o.get("hup"), o.get("hvp")
    zero = _full(o, 0.0)
    # the face velocity is the mean of the n+1 and n-1 transports
    if X:
        fields = {"uh": 1.5 * U * hq, "uhpp": 0.5 * U * hq, "vh": zero, "vhpp": zero}
    else:
        fields = {"vh": 1.5 * U * hn, "vhpp": 0.5 * U * hn, "uh": zero, "uhpp": zero}
    for name, v in fields.items():
        o.set(name, v)
    o.set("wp", zero)
    o.set("Dv", zero)
    o.phase(9, 2, 1)
    o.phase(10, 2, 1)
    sal = o.get("sal")
    nb_p, nb_m = (g["lEC"], g["lWC"]) if X else (g["lNC"], g["lSC"])
    expect = 20.0 + cq * (pos - x0 - U * dt2) ** 2
    checked = 0
    for l in range(o.lm):
        chain = [l]
        for _ in range(2):
            chain.append(nb_p[chain[-1]] if chain[-1] < o.lm else o.lm)
        back = [l]
        for _ in range(2):
            back.append(nb_m[back[-1]] if back[-1] < o.lm else o.lm)
        if o.lm in chain or o.lm in back:
            continue  # walls within two cells
        for k in range(2, g["kmz"][l] + 1):
            assert sal[l, k - 1] == pytest.approx(expect[l], rel=1e-12), (l, k)
            checked += 1
    assert checked > 100
    # the limiter's branch matters: first-order upwind would be off by far more
    courant = U * dt2 / sp
    assert cq * sp ** 2 * courant * (1 - courant) > 1e-4 * 20.0


# ------------------------------------------------------------------- imsal

@pytest.mark.parametrize("m,istep", [(1, 1), (2, 1), (1, 2)])
def test_imsal_diffusive_decay_of_cosine_mode(oracle_lib, m, istep):
    """imsal (si3d_procedures.f90:3157-3189, trid1 3292-3352) with no
    transport: fully implicit vertical diffusion of a cosine mode with
    constant Dv over N one-metre layers divides it by 1 + dt' Dv (2 - 2
    cos(pi m / N)) / h^2, dt' = 2 dt on a leapfrog call and dt on a
    trapezoidal one."""
    d = _transport_lake(nx=8, ny=6)
    o = oracle_lib.Oracle(d)
    g = _grid(o)
    zero = _full(o, 0.0)
    for name in ("uh", "uhpp", "vh", "vhpp", "wp"):
        o.set(name, zero)
    Dv = 3.0e-3
    o.set("Dv", _full(o, Dv))
    k1 = 2
    N = int(g["kmz"][0]) - 1
    mode = 15.0 + 2.0 * _cosine(N, m, k1, o.km1)
    mode[N + 1:] = 15.0
    S = np.ones((o.lm1, 1)) * mode[None, :]
    o.set("salpp", S)
    o.phase(9, 2, istep)
    o.phase(10, 2, istep)
    sal = o.get("sal")
    L = 2.0 * d["dt"] / istep * Dv * (2 - 2 * math.cos(math.pi * m / N))
    for l in range(o.lm):
        for k in range(k1, k1 + N):
            want = 15.0 + 2.0 * _cosine(N, m, k1, o.km1)[k - 1] / (1 + L)
            assert sal[l, k - 1] == pytest.approx(want, rel=1e-12), (l, k)
    assert L > 0.02


def test_imsal_split_column_keeps_each_part(oracle_lib):
    """With Dv = 0 on the interface above layer ks only (Dv(k) couples layers
    k-1 and k in imsal's a1/a3), the heat content above and below it stays
    apart while diffusion mixes each part (layers of unequal thickness)."""
    d = _transport_lake(nx=8, ny=6)
    z = lake.stretched_levels(8, 0.3, 1.5)
    d["ibathyf"], d["zlevel"], d["zl"] = -1, z, float(z[-1])
    d["bathy_dm"] = np.full((d["imm"], d["jmm"]), int(round(10 * (z[-1] - 0.2))), dtype=np.int32)
    d["init_profile"] = np.full(len(z), 15.0)
    o = oracle_lib.Oracle(d)
    g = _grid(o)
    zero = _full(o, 0.0)
    for name in ("uh", "uhpp", "vh", "vhpp", "wp"):
        o.set(name, zero)
    ks = 4
    Dv = _full(o, 5.0e-2)
    Dv[:, ks - 1] = 0.0
    o.set("Dv", Dv)
    rng = np.random.default_rng(2)
    S = np.ones((o.lm1, 1)) * rng.uniform(10.0, 20.0, o.km1)[None, :]
    o.set("salpp", S)
    o.phase(9, 2, 1)
    o.phase(10, 2, 1)
    sal, h = o.get("sal"), o.get("h")
    for l in range(o.lm):
        km = g["kmz"][l]
        for a, b in ((1, ks - 1), (ks - 1, km)):
            assert (sal[l, a:b] * h[l, a:b]).sum() == pytest.approx((S[l, a:b] * h[l, a:b]).sum(), rel=1e-13)
        assert np.max(np.abs(sal[l, 1:km] - S[l, 1:km])) > 0.1


# -------------------------------------------------------- exmom, itrmom 2-4

@pytest.mark.parametrize("itrmom", [2, 3, 4])
@pytest.mark.parametrize("X", [True, False])
def test_exmom_advection_schemes_exact_on_linear_fields(oracle_lib, itrmom, X):
    """exmom's other momentum advection schemes (si3d_procedures.f90:885-929):
    2 = first-order upwind in x, y and z on the n-1 velocity, 3 = centred in
    x and y on the n velocity with thickness-weighted interface values in z,
    4 = upwind in x and y on the n-1 velocity with the centred vertical
    weights of scheme 1 on the n velocity.  With uniform advecting transports
    U h, V h, a uniform vertical velocity W > 0 and linear advected
    velocities at n and n-1 with different gradients, every one-sided and
    centred difference is exact:
    ex = uhpp - 2dt h (U ax + V ay - W az), each gradient taken from the
    time level that scheme's term uses."""
    from tests.test_oracle_pins import _momentum_lake
    d = _momentum_lake(iadv=1, ihd=0)
    d["f"] = 0.0
    d["itrmom"] = itrmom
    o = oracle_lib.Oracle(d)
    g = _grid(o)
    dx, dy = d["dx"], d["dy"]
    Xp, Yp, D = _positions(o, g, dx, dy)
    twodt1 = 2.0 * d["dt"]
    U, V, W = 0.04, 0.03, 1.5e-4
    grads = {"n": (2.0e-5, -3.0e-5, 4.0e-3), "n-1": (-1.2e-5, 2.5e-5, 6.0e-3)}
    xo, yo = (dx / 2, 0.0) if X else (0.0, dy / 2)

    def field(a):
        ax, ay, az = a
        return 0.05 + ax * (Xp[:, None] + xo) + ay * (Yp[:, None] + yo) + az * D[None, :]

    qn, qpp = field(grads["n"]), field(grads["n-1"])
    hu, hv = o.get("hup"), o.get("hvp")
    h = hu if X else hv
    fields = ([("up", qn), ("upp", qpp), ("uhp", U * hu), ("uhpp", 0.7 * qn * hu), ("vhp", V * hv)] if X else
              [("vp", qn), ("vpp", qpp), ("vhp", V * hv), ("vhpp", 0.7 * qn * hv), ("uhp", U * hu)])
    for name, val in fields + [("wp", _full(o, W))]:
        o.set(name, val)
    o.phase(2 if X else 4, 2, 1)
    ex = o.get("ex")
    horiz = grads["n-1"] if itrmom in (2, 4) else grads["n"]
    vert = grads["n-1"] if itrmom == 2 else grads["n"]
    rate = U * horiz[0] + V * horiz[1] - W * vert[2]
    k1q = g["k1u"] if X else g["k1v"]
    checked = 0
    for l in _interior(o, g):
        for k in range(k1q[l] + 1, g["kmz"][l]):
            expect = 0.7 * qn[l, k - 1] * h[l, k - 1] - twodt1 * h[l, k - 1] * rate
            assert ex[l, k - 1] == pytest.approx(expect, rel=1e-12, abs=1e-15), (itrmom, l, k)
            checked += 1
    assert checked > 200


@pytest.mark.parametrize("itrmom", [1, 3])
@pytest.mark.parametrize("X", [True, False])
def test_exmom_vertical_advection_conserves_momentum(oracle_lib, itrmom, X):
    """The vertical advection of the centred schemes (1: interface values
    weighted by th / th1, 3: by the layer thicknesses) is in flux form: the
    value at an interface is the same for the layers above and below it, so
    with a uniform vertical velocity, no horizontal gradient and w = 0 at the
    top and bottom of the face, the column sum of the advection term is zero
    for any velocity profile, on layers of unequal thickness."""
    d = _column_lake(stretched=True)
    d["iadv"], d["itrmom"] = 1, itrmom
    o = oracle_lib.Oracle(d)
    g = _grid(o)
    rng = np.random.default_rng(9)
    prof = rng.uniform(-0.05, 0.05, o.km1)
    q = np.ones((o.lm1, 1)) * prof[None, :]
    hu, hv = o.get("hup"), o.get("hvp")
    h = hu if X else hv
    zero = _full(o, 0.0)
    fields = ([("up", q), ("upp", q), ("uhp", 0.04 * hu), ("uhpp", zero), ("vhp", zero)] if X else
              [("vp", q), ("vpp", q), ("vhp", 0.04 * hv), ("vhpp", zero), ("uhp", zero)])
    for name, val in fields + [("wp", _full(o, 2.0e-4))]:
        o.set(name, val)
    o.phase(2 if X else 4, 2, 1)
    ex = o.get("ex")
    k1q = g["k1u"] if X else g["k1v"]
    nb = g["lEC"] if X else g["lNC"]
    checked = 0
    for l in _interior(o, g, margin=3):  # faces away from the walls: no horizontal term
        km = min(g["kmz"][l], g["kmz"][nb[l]])
        col = ex[l, k1q[l] - 1:km]
        checked += 1
        assert abs(col.sum()) < 1e-13 * np.abs(col).max(), l
        assert np.abs(col).max() > 1e-6
        assert np.ptp(h[l, k1q[l] - 1:km]) > 0.5
    assert checked >= 8


# ------------------------------------------------------ the TVD face flux

def _superbee(r):
    """Roe's superbee limiter (Roe 1985; Sweby 1984, fig. 4):
    phi(r) = max(0, min(2 r, 1), min(r, 2))."""
    return max(0.0, min(2.0 * r, 1.0), min(r, 2.0))


@pytest.mark.parametrize("vel", [0.3, -0.3])
@pytest.mark.parametrize("r", [-1.0, 0.0, 0.25, 0.5, 0.75, 1.0, 1.5, 2.0, 3.0])
def test_tvd_face_flux_is_superbee_limited_lax_wendroff(oracle_lib, vel, r):
    """The face flux of exsal is the flux-limited Lax-Wendroff flux of Sweby
    (1984), F = u (s2 + s3)/2 - ((1 - phi) |u| + phi u^2 dt/dx) (s3 - s2)/2,
    with phi the superbee limiter of the upwind gradient ratio
    r = (s2 - s1)/(s3 - s2) for u > 0 and (s4 - s3)/(s3 - s2) for u < 0;
    where s3 = s2 the limiter takes its default value."""
    courant = 0.8
    s2, s3 = 1.0, 3.0
    if vel > 0:
        s1, s4 = s2 - r * (s3 - s2), 11.0
    else:
        s1, s4 = -7.0, s3 + r * (s3 - s2)
    phi = _superbee(r)
    want = vel * (s2 + s3) / 2 - ((1 - phi) * abs(vel) + phi * vel * vel * courant) * (s3 - s2) / 2
    assert oracle_lib.tvd_flux(vel, s1, s2, s3, s4, 0.0, courant) == pytest.approx(want, rel=1e-14, abs=1e-15)
    # a level scalar on the face: the default limiter (0 horizontal, 1 vertical)
    for cfdef in (0.0, 1.0):
        want = vel * 2.0 - 0.0
        assert oracle_lib.tvd_flux(vel, 5.0, 2.0, 2.0, 9.0, cfdef, courant) == pytest.approx(want, rel=1e-14)
