"""Input decks for pSi3D lakes: the content of si3d_inp.txt, si3d_bathy,
si3d_layer.txt, si3d_init.txt and si3d_surfbc_W.txt, held in memory.

A deck is a plain dict.  Its keys follow the names the reference reads in
`input` (si3d_utils.f90:21-315), `ZGridDimensions` (318-382), `bathy`
(595-892), `InitializeScalarFields` (si3d_procedures.f90:292-363) and
`surfbc0` case 20 (si3d_boundaryconditions.f90:4450-4484).  Depths are
integer decimetres as in si3d_bathy.

The generators build the synthetic lakes used for measurement:

* rectangle  -- flat rectangular basin with a constant wind (uniform layers,
  constant eddy viscosity); the 100x50x20 case is the CPU-runnable workload.
* ellipse    -- elliptic basin, depth 100 m * (1 - r^2) plus a seeded Gaussian
  random field, geometrically stretched layers, tanh thermocline, a
  sinusoidal wind series, TVD transport and the two-equation turbulence
  closure on.  1024x512 cells with 60 layers is the single-GPU workload.
* channel    -- rectangular reach with a cross-channel depth profile, for
  open boundaries (`open_boundary` adds one: the nopen block of si3d_inp.txt
  and its openbcNN.txt series).
"""
from __future__ import annotations

import math

import numpy as np

G_RHOAIR = 1.3  # si3d_types.f90:33

_DEFAULTS = dict(
    ibathyf=0, zlevel=None, dzmin=0.1, datadj=0.0, zetainit=0.0,
    itrap=1, niter=2, ismooth=0, beta=0.05, iturb=0, Av0=3e-2, Dv0=3e-2,
    iadv=1, itrmom=1, ihd=1, Ax0=1.0, Ay0=1.0, f=0.0, theta=1.0, ibc=0,
    isal=0, cd=2.5e-3, ifsurfbc=0, dtsurfbc=600.0, cw=1.2e-3, wa=0.0,
    phi=0.0, surfbc=None, pcg_tol=1e-10, pcg_maxit=4000,
)


def _deck(**kw):
    d = dict(_DEFAULTS)
    d.update(kw)
    return d


def km1_of(deck) -> int:
    """Number of layer records km1 (ZGridDimensions, si3d_utils.f90:318-382)."""
    if deck["ibathyf"] < 0:
        return len(deck["zlevel"])
    zl = np.float32(deck["zl"])
    ddz = np.float32(deck["ddz"])
    return int(math.ceil(float(zl / ddz))) + 2


def rectangle(nx=100, ny=50, dx=50.0, depth=20.0, dz=1.0, tau=0.1, dt=30.0,
              nts=1000, iadv=1, ihd=1, Ax0=1.0, cd=2.5e-3, Av0=3e-2,
              temperature=20.0, phi=270.0):
    """Flat rectangular basin, homogeneous density, uniform wind stress `tau`
    (Pa) blowing toward +x (phi = 270 deg), closed boundaries."""
    cw = 1.2e-3
    wa = math.sqrt(tau / (cw * G_RHOAIR))
    d = _deck(
        name=f"rectangle_{nx}x{ny}x{int(round(depth / dz))}",
        imm=nx, jmm=ny, dx=dx, dy=dx, ddz=dz, zl=depth, dt=dt, nts=nts,
        bathy_dm=np.full((nx, ny), int(round(depth * 10)), dtype=np.int32),
        iadv=iadv, ihd=ihd, Ax0=Ax0, Ay0=Ax0, cd=cd, Av0=Av0, Dv0=Av0,
        iturb=0, isal=0, ibc=0, ifsurfbc=0, cw=cw, wa=wa, phi=phi,
    )
    d["init_profile"] = np.full(km1_of(d), temperature)
    return d


def stretched_levels(nlayers=60, top=0.5, bottom=5.0):
    """zlevel records of si3d_layer.txt: two -100 sentinels, then the depths of
    the layer interfaces for `nlayers` layers growing geometrically from `top`
    to `bottom` metres (si3d_utils.f90:330-358)."""
    r = (bottom / top) ** (1.0 / (nlayers - 1))
    dz = top * r ** np.arange(nlayers)
    z = np.concatenate([[-100.0, -100.0], np.cumsum(dz)])
    # zlevel(k) is the depth of the top of layer k; the last record is the
    # bottom of the deepest layer.
    return np.round(z, 4)


def gaussian_random_field(nx, ny, dx, corr_len, sigma, seed):
    """Seeded Gaussian random field with Gaussian covariance of length
    `corr_len` and standard deviation `sigma`, by FFT filtering of white noise."""
    rng = np.random.default_rng(seed)
    w = rng.standard_normal((nx, ny))
    kx = np.fft.fftfreq(nx, d=dx) * 2 * np.pi
    ky = np.fft.fftfreq(ny, d=dx) * 2 * np.pi
    k2 = kx[:, None] ** 2 + ky[None, :] ** 2
    f = np.real(np.fft.ifft2(np.fft.fft2(w) * np.exp(-k2 * corr_len ** 2 / 4.0)))
    f = f - f.mean()
    sd = f.std()
    return f * (sigma / sd) if sd > 0 else f


def ellipse_bathymetry(nx, ny, dx, depth=100.0, sigma=5.0, corr_len=2000.0,
                       seed=0, min_depth=1.0, max_depth=None):
    """Depths (dm, int32) of an elliptic basin spanning the box with a one-cell
    margin: depth*(1-r^2) + GRF; cells shallower than `min_depth` are dry."""
    i = np.arange(nx) + 0.5
    j = np.arange(ny) + 0.5
    a = nx / 2.0 - 1.0
    b = ny / 2.0 - 1.0
    X = (i[:, None] - nx / 2.0) / a
    Y = (j[None, :] - ny / 2.0) / b
    r2 = X ** 2 + Y ** 2
    h = depth * (1.0 - r2) + gaussian_random_field(nx, ny, dx, corr_len, sigma, seed)
    if max_depth is not None:
        h = np.minimum(h, max_depth)
    h = np.where((r2 < 1.0) & (h >= min_depth), h, -9.9)
    return np.round(h * 10.0).astype(np.int32)


def thermocline_profile(zlevel, t_top=24.0, t_bot=8.0, z_tc=15.0, width=2.0):
    """Temperature per layer k=1..km1 for si3d_init.txt: t_bot + (t_top-t_bot)
    * (1 + tanh((z_tc - z)/width))/2 at the layer centre."""
    z = np.asarray(zlevel, dtype=np.float64)
    km1 = len(z)
    prof = np.empty(km1)
    for k in range(km1):
        top = 0.0 if z[k] < -99 else z[k]
        bot = z[k + 1] if k + 1 < km1 else z[k] + 0.5
        zc = 0.5 * (top + bot)
        prof[k] = t_bot + (t_top - t_bot) * 0.5 * (1.0 + math.tanh((z_tc - zc) / width))
    return prof


def wind_series(tl, dtsbc=600.0, amp=0.2, period=86400.0, direction_deg=30.0,
                cdw=1.2e-3, phase=0.0):
    """si3d_surfbc_W.txt records (cdw, uair, vair) for a stress
    amp*sin(2 pi t/period + phase) Pa blowing from `direction_deg`."""
    n = int(tl / dtsbc) + 4
    t = np.arange(n) * dtsbc
    tau = amp * np.sin(2 * np.pi * t / period + phase)
    U = np.sqrt(np.abs(tau) / (cdw * G_RHOAIR))
    th = math.radians(direction_deg)
    ua = np.sign(tau) * U * (-math.sin(th))
    va = np.sign(tau) * U * (-math.cos(th))
    return np.stack([np.full(n, cdw), ua, va], axis=1)


def random_wind_series(tl, dtsbc=600.0, amp=(0.05, 0.3), seed=0, cdw=1.2e-3):
    """si3d_surfbc_W.txt records (cdw, uair, vair) whose stress changes every
    record to a random amplitude uniform in `amp` (Pa) blowing from a random
    direction uniform in [0, 360) deg: a sequence of different forcings."""
    n = int(tl / dtsbc) + 4
    rng = np.random.default_rng(seed)
    tau = rng.uniform(amp[0], amp[1], n)
    th = np.radians(rng.uniform(0.0, 360.0, n))
    U = np.sqrt(tau / (cdw * G_RHOAIR))
    return np.stack([np.full(n, cdw), -U * np.sin(th), -U * np.cos(th)], axis=1)


def heat_budget(deck, mode, eta=0.5, qsw_max=800.0, q_night=-150.0, period=86400.0):
    """The deck with a spatially uniform surface heat budget (ifsurfbc = mode,
    1, 2 or 3, si3d_surfbc.txt records) over its wind series: a diurnal
    shortwave cycle max(0, qsw_max cos(2 pi t/period)) (t = 0 at noon), attenuation `eta`
    (1/m); mode 1 adds a net flux Qn = Qsw + q_night, modes 2 and 3 an air
    temperature of 20 +- 5 C, 101325 Pa, 70 % humidity, and 30 % cloud cover
    (2) or 330 W/m2 of incoming longwave (3)."""
    if mode not in (1, 2, 3):
        raise ValueError("heat budget mode must be 1, 2 or 3")
    w = np.asarray(deck["surfbc"], dtype=np.float64)
    if w.ndim != 2 or w.shape[1] != 3:
        raise ValueError("heat_budget needs a deck with a wind series (ifsurfbc = 20)")
    n = w.shape[0]
    t = np.arange(n) * deck["dtsurfbc"]
    ph = 2 * np.pi * t / period
    qsw = np.maximum(0.0, qsw_max * np.cos(ph))
    if mode == 1:
        cols = [np.full(n, eta), qsw, qsw + q_night]
    else:
        ta = 20.0 + 5.0 * np.cos(ph - 0.5)
        last = np.full(n, 0.3) if mode == 2 else np.full(n, 330.0)
        cols = [np.full(n, eta), qsw, ta, np.full(n, 101325.0), np.full(n, 0.7), last]
    d = dict(deck)
    d["ifsurfbc"] = mode
    d["surfbc"] = np.concatenate([np.stack(cols, axis=1), w], axis=1)
    return d


def met_stations(deck, mode, stations, iinterp=1, gammaB=0.3, delNfactor=1.0, icw=0, eta=0.5, qsw_max=800.0,
                 period=86400.0, same=False):
    """The deck with the met-station surface heat budget (ifsurfbc = mode, 10
    or 11; si3d_surfbc.txt with 6 * nmetstat + 2 values per record): stations
    at the (x, y) grid positions `stations`, each with its own diurnal
    shortwave (max(0, qsw_max cos(2 pi t/period)) scaled by 1 - 0.1 m), air
    temperature (20 + 5 cos +- 1.5 m C), humidity (0.6 + 0.05 m), cloud cover
    0.3 (mode 10) or incoming longwave 320 + 10 m W/m2 (mode 11) and wind
    (the deck's wind series turned by 25 m degrees and scaled by 1 + 0.3 m),
    where m is the station index; `same` gives every station station 0's
    series.  Interpolation to the columns: iinterp 1 (Barnes), 2 (two-pass
    Barnes with gammaB, delNfactor) or 3 (inverse distance); drag icw."""
    if mode not in (10, 11):
        raise ValueError("met-station mode must be 10 or 11")
    w = np.asarray(deck["surfbc"], dtype=np.float64)
    if w.ndim != 2 or w.shape[1] != 3:
        raise ValueError("met_stations needs a deck with a wind series (ifsurfbc = 20)")
    n = w.shape[0]
    t = np.arange(n) * deck["dtsurfbc"]
    ph = 2 * np.pi * t / period
    cols = [np.full(n, eta), np.full(n, 101325.0)]
    for m in range(len(stations)):
        q = 0 if same else m
        qsw = np.maximum(0.0, qsw_max * np.cos(ph)) * (1.0 - 0.1 * q)
        ta = 20.0 + 5.0 * np.cos(ph - 0.5) + 1.5 * (q % 2) - 0.75 * q
        th = np.radians(25.0 * q)
        ua = (w[:, 1] * np.cos(th) - w[:, 2] * np.sin(th)) * (1.0 + 0.3 * q)
        va = (w[:, 1] * np.sin(th) + w[:, 2] * np.cos(th)) * (1.0 + 0.3 * q)
        last = np.full(n, 0.3) if mode == 10 else np.full(n, 320.0 + 10.0 * q)
        cols += [qsw, ta, np.full(n, 0.6 + 0.05 * q), last, ua, va]
    d = dict(deck)
    d.update(ifsurfbc=mode, surfbc=np.stack(cols, axis=1), metxy=np.asarray(stations, dtype=np.float64).ravel(),
             nmetstat=len(stations), iinterp=iinterp, gammaB=gammaB, delNfactor=delNfactor, icw=icw)
    return d


def ellipse(nx=1024, ny=512, dx=100.0, nlayers=60, dt=60.0, nts=2000, seed=1,
            wind_amp=0.2, wind_dir=30.0, f=1e-4, Ax0=1.0, depth=100.0,
            top=0.5, bottom=5.0, iturb=1):
    """Synthetic stratified elliptic lake with TVD transport and turbulence."""
    z = stretched_levels(nlayers, top, bottom)
    bathy = ellipse_bathymetry(nx, ny, dx, depth=depth, seed=seed,
                               max_depth=float(z[-1]) - 0.5)
    d = _deck(
        name=f"ellipse_{nx}x{ny}x{nlayers}",
        imm=nx, jmm=ny, dx=dx, dy=dx, ddz=1.0, zl=float(z[-1]), dt=dt, nts=nts,
        ibathyf=-1, zlevel=z, bathy_dm=bathy,
        iadv=1, itrmom=1, ihd=1, Ax0=Ax0, Ay0=Ax0, f=f, ibc=1, isal=1,
        iturb=iturb, ifsurfbc=20, dtsurfbc=600.0,
        surfbc=wind_series(nts * dt, 600.0, wind_amp, 86400.0, wind_dir),
    )
    d["init_profile"] = thermocline_profile(z)
    return d


def ensemble(n=64, seed=7, side=(128, 768), layers=(20, 80), wind_amp=(0.05, 0.3), dt=60.0, nts=500,
             dx=100.0):
    """Decks of an ensemble of `n` independent elliptic lakes: square boxes with
    sides uniform in `side`, layer counts uniform in `layers`, wind amplitude
    uniform in `wind_amp` (Pa) and direction uniform in [0, 360), independent
    bathymetry seeds."""
    rng = np.random.default_rng(seed)
    decks = []
    for _ in range(n):
        s_ = int(rng.integers(side[0], side[1] + 1))
        nl = int(rng.integers(layers[0], layers[1] + 1))
        amp = float(rng.uniform(*wind_amp))
        direc = float(rng.uniform(0.0, 360.0))
        decks.append(ellipse(nx=s_, ny=s_, dx=dx, nlayers=nl, dt=dt, nts=nts,
                             seed=int(rng.integers(1, 2 ** 31 - 1)), wind_amp=amp, wind_dir=direc))
    return decks


def n_wet_cells(deck) -> int:
    """N_wet = sum over wet columns of (kmz - k1 + 1) at t=0 for a level
    surface (the per-column layer count of bathy, si3d_utils.f90:812-867)."""
    h = deck["bathy_dm"].astype(np.float32) * np.float32(0.1) + np.float32(deck["datadj"])
    wet = h > 0
    if deck["ibathyf"] >= 0:
        ddz = np.float32(deck["ddz"])
        kz = np.floor(h / ddz).astype(np.int64)
        kz = kz + ((h - kz.astype(np.float32) * ddz) > np.float32(deck["dzmin"]))
        return int(kz[wet].sum())
    z = np.asarray(deck["zlevel"], dtype=np.float32)
    # layer count = index of the first level whose bottom reaches the depth
    bottoms = z[2:]
    kz = np.searchsorted(bottoms, h[wet], side="left") + 1
    return int(kz.sum())


def shallow_basin(nx=40, ny=8, dx=50.0, depth=0.9, tau=0.05, top=0.01, bottom=0.2, nlayers=15, dt=20.0,
                  nts=300, iturb=1, second=None, stratified=False):
    """Closed, shallow rectangular basin whose top layer is thinner than the
    wind set-down, so the upwind end dries and re-wets its top layers as the
    basin seiches: layers grow geometrically from `top` to `bottom` metres,
    the two side rows are 0.1 m shallower than the rest, and a uniform
    stress `tau` (Pa) toward +x starts at t = 0.  `second` (m) inserts a
    layer that thin right below the top one, so that the surface can fall
    through both in one call.  `stratified` transports temperature (22 C at
    the surface to 18 C at the bottom, density from it) instead of keeping
    a uniform 20 C."""
    z = stretched_levels(nlayers, top, bottom)
    if second is not None:
        z = np.concatenate([z[:3], [z[2] + second], z[3:]])
    bathy = np.full((nx, ny), int(round(depth * 10)), dtype=np.int32)
    bathy[:, 0] = bathy[:, -1] = int(round(depth * 10)) - 1
    cw = 1.2e-3
    d = _deck(
        name=f"shallow_basin_{nx}x{ny}x{nlayers}", imm=nx, jmm=ny, dx=dx, dy=dx, ddz=1.0, zl=float(z[-1]),
        dt=dt, nts=nts, ibathyf=-1, zlevel=z, bathy_dm=bathy, iadv=1, itrmom=1, ihd=1, Ax0=1.0, Ay0=1.0,
        f=0.0, ibc=0, isal=0, iturb=iturb, Av0=1e-3, Dv0=1e-4, ifsurfbc=0, cw=cw,
        wa=math.sqrt(tau / (cw * G_RHOAIR)), phi=270.0,
    )
    d["init_profile"] = np.linspace(22.0, 18.0, len(z)) if stratified else np.full(len(z), 20.0)
    if stratified:
        d["isal"], d["ibc"] = 1, 1
    return d


def tidal_flat(nx=40, ny=8, dx=50.0, dz=0.2, amp=0.45, period=7200.0, tau=0.1, dt=30.0, nts=300, iturb=1):
    """Channel (`channel`, 3 m deep on its axis, 2 m at the banks, layers of
    `dz` metres) with a tide of amplitude `amp` (m) on a water-level boundary
    across its west end and a wind stress `tau` (Pa) toward -x: the falling
    tide empties the top layers, the rising one refills them, and the wind
    tilts the surface so that the drying front crosses the channel over
    several steps.  The tide starts falling at t = 0."""
    d = channel(nx=nx, ny=ny, dx=dx, depth=3.0, shallow=2.0, dz=dz, dt=dt, nts=nts, iturb=iturb)
    d["cw"] = 1.2e-3
    d["wa"] = math.sqrt(tau / (d["cw"] * G_RHOAIR)) if tau > 0 else 0.0
    d["phi"] = 90.0
    npts = int(nts * dt / 60.0) + 8
    t = np.arange(npts) * 60.0
    series = np.stack([-amp * np.sin(2 * np.pi * t / period), np.full(npts, 20.0)], axis=1)
    return open_boundary(d, "W", 1, 2, ny + 1, series, dtsec=60.0)


SIDES = {"W": 1, "N": 2, "E": 3, "S": 4}


def channel(nx=30, ny=12, dx=50.0, depth=8.0, shallow=4.0, dz=1.0, dt=30.0, nts=200, iturb=0,
            isal=1, temperature=20.0, ntr=0):
    """Rectangular reach, deepest along its axis (x) and shoaling to `shallow`
    at the banks, no wind, no Coriolis, uniform temperature; `ntr` tracers
    start at zero."""
    jj = np.arange(ny) - 0.5 * (ny - 1)
    prof = depth - (depth - shallow) * (jj / (0.5 * ny)) ** 2
    bathy = np.repeat(np.round(prof * 10.0).astype(np.int32)[None, :], nx, axis=0)
    d = _deck(
        name=f"channel_{nx}x{ny}", imm=nx, jmm=ny, dx=dx, dy=dx, ddz=dz, zl=depth, dt=dt, nts=nts,
        bathy_dm=bathy, iadv=1, itrmom=1, ihd=1, Ax0=1.0, Ay0=1.0, f=0.0, ibc=1 if isal else 0, isal=isal,
        iturb=iturb, Av0=1e-3, Dv0=1e-4, ifsurfbc=0, wa=0.0,
    )
    d["init_profile"] = np.full(km1_of(d), temperature)
    if ntr:
        d["tracer_profiles"] = np.zeros((ntr, km1_of(d)))
    return d


def open_boundary(deck, side, btype, first, last, series, dtsec=600.0):
    """Deck with one more open boundary (si3d_utils.f90:123-144,
    si3d_boundaryconditions.f90:23-897): on the outer interior column (W at
    i = 2, E at i = imm+1) or row (S at j = 2, N at j = jmm+1) of the grid,
    cells first..last along it (grid indices as in si3d_inp.txt).  btype 1
    prescribes the water surface elevation (m), 2 the flow (m3/s, positive
    toward +x / +y).  series: records x (2 + ntr) values (elevation or flow,
    active scalar, tracers) dtsec seconds apart, as in openbcNN.txt."""
    sd = SIDES[side] if isinstance(side, str) else int(side)
    series = np.atleast_2d(np.asarray(series, dtype=np.float64))
    if sd in (1, 3):
        i = 2 if sd == 1 else deck["imm"] + 1
        spec = (sd, btype, i, first, i, last)
    else:
        j = 2 if sd == 4 else deck["jmm"] + 1
        spec = (sd, btype, first, j, last, j)
    d = dict(deck)
    ob = dict(d.get("openbc") or dict(dtsec=dtsec, iside=[], itype=[], isbc=[], jsbc=[], iebc=[], jebc=[],
                                      values=np.zeros((0,) + series.shape)))
    if ob["values"].shape[1:] != series.shape or ob["dtsec"] != dtsec:
        if ob["values"].shape[0]:
            raise ValueError("open boundaries share one record count, value count and dtsecopenbc")
    for key, v in zip(["iside", "itype", "isbc", "jsbc", "iebc", "jebc"], spec):
        ob[key] = list(ob[key]) + [v]
    ob["values"] = np.concatenate([ob["values"].reshape((-1,) + series.shape), series[None]], axis=0)
    ob["dtsec"] = dtsec
    d["openbc"] = ob
    return d


def steady_series(value, scalar=0.0, tracers=(), npts=200):
    """Boundary series holding every value constant."""
    row = [value, scalar] + list(tracers)
    return np.tile(np.asarray(row, dtype=np.float64), (npts, 1))


def mirror_x(deck):
    """The same lake reflected in x (i -> imm+3-i)."""
    d = dict(deck)
    d["bathy_dm"] = np.ascontiguousarray(deck["bathy_dm"][::-1, :])
    return d


def mirror_y(deck):
    """The same lake reflected in y (j -> jmm+3-j)."""
    d = dict(deck)
    d["bathy_dm"] = np.ascontiguousarray(deck["bathy_dm"][:, ::-1])
    return d


def transpose_xy(deck):
    """The same lake with x and y exchanged (needs dx = dy)."""
    d = dict(deck)
    d["bathy_dm"] = np.ascontiguousarray(deck["bathy_dm"].T)
    d["imm"], d["jmm"] = deck["jmm"], deck["imm"]
    return d
