"""Readers of pSi3D's own input files into a deck (the dict `GpuLake` and
`s3g_create_from_deck` take), so that an existing case runs on this path from
its files:

* si3d_inp.txt      -- `input` (si3d_utils.f90:21-315)
* si3d_bathy        -- `bathy` (si3d_utils.f90:595-640): depths in dm
* si3d_layer.txt    -- `ZGridDimensions` (si3d_utils.f90:318-382), ibathyf < 0
* si3d_init.txt     -- `InitializeScalarFields` (si3d_procedures.f90:292-363)
* si3d_surfbc.txt / si3d_surfbc_W.txt -- `surfbc0` (si3d_boundaryconditions.f90:
  4140-4487): ifsurfbc 1-3 (6 or 9 values per record) and 20 (3 values)

The files are read record by record with the reference's edit descriptors:
fixed-width fields (`14X,G20.2`, `10X,G11.2`, `5X,nG5.0`, ...), where a comma
ends a numeric field early and the next field starts after it (the bathymetry
header relies on this), blank fields read as zero, and the list reads
`(14X,20I)` take every integer on the record.  Each reader raises
`InputError` naming the file and record where the reference would stop with
an input error.
"""
from __future__ import annotations

import math
import os
import re

import numpy as np

from . import lake


class InputError(ValueError):
    pass


class _Record:
    """One record of a formatted sequential file and a read position in it."""

    def __init__(self, text, where):
        self.s = text.rstrip("\n")
        self.pos = 0
        self.where = where

    def skip(self, n):
        self.pos += n

    def _field(self, w):
        raw = self.s[self.pos:self.pos + w]
        cut = raw.find(",")
        if cut >= 0:  # a comma ends the field; the next one starts after it
            self.pos += cut + 1
            return raw[:cut]
        self.pos += w
        return raw

    def real(self, w):
        f = self._field(w).strip()
        if not f:
            return 0.0
        try:
            return float(f.replace("D", "E").replace("d", "e"))
        except ValueError:
            raise InputError(f"{self.where}: cannot read a number from {f!r}") from None

    def int(self, w):
        f = self._field(w).strip()
        if not f:
            return 0
        try:
            return int(f)
        except ValueError:
            try:  # a G or I field holding an integral real, e.g. "2020."
                v = float(f)
            except ValueError:
                raise InputError(f"{self.where}: cannot read an integer from {f!r}") from None
            if v != int(v):
                raise InputError(f"{self.where}: {f!r} is not an integer")
            return int(v)

    def int_list(self):
        """The `20I` list read: every integer after the current position."""
        return [int(t) for t in re.findall(r"[-+]?\d+", self.s[self.pos:])]

    def label(self):
        return self.s[:13].strip().split()[0] if self.s[:13].strip() else ""


class _File:
    def __init__(self, path):
        self.path = path
        try:
            with open(path, "r", errors="replace") as f:
                self.lines = f.readlines()
        except OSError as e:
            raise InputError(f"Error opening {path}: {e}") from None
        self.n = 0

    def record(self):
        if self.n >= len(self.lines):
            raise InputError(f"{os.path.basename(self.path)}: end of file at record {self.n + 1}")
        r = _Record(self.lines[self.n], f"{os.path.basename(self.path)} record {self.n + 1}")
        self.n += 1
        return r

    def skip(self, n):
        for _ in range(n):
            self.record()

    def values(self, n, skip_before=3, width=20, integer=False):
        """READ (FMT='(///(14X,G20.2))') of n items: `skip_before` records,
        then one item per record."""
        self.skip(skip_before)
        out = []
        for _ in range(n):
            r = self.record()
            r.skip(14)
            out.append(r.int(width) if integer else r.real(width))
        return out

    def value(self, integer=False):
        return self.values(1, skip_before=0, integer=integer)[0]

    def ilist(self):
        r = self.record()
        r.skip(14)
        return r.int_list()


_SIM_NAMES = ["itrap", "niter", "ismooth", "beta", "iturb", "Av0", "Dv0", "iadv", "itrmom", "ihd", "Ax0",
              "Ay0", "f", "theta", "ibc", "isal", "itrsca", "cd", "ifsurfbc", "dtsurfbc", "icw", "cw", "wa",
              "phi", "idbg", "num_threads"]
_INT_NAMES = {"itrap", "niter", "ismooth", "iturb", "iadv", "itrmom", "ihd", "ibc", "isal", "itrsca",
              "ifsurfbc", "icw", "idbg", "num_threads"}


def read_inp(path, missing_icw=False):
    """si3d_inp.txt as a dict of the variables `input` reads (si3d_utils.f90:
    21-315), plus the derived grid sizes im, jm, nts.  `missing_icw` accepts
    files written before the `icw` record existed (icw = 0 is assumed); the
    reference itself reads such a file shifted by one record and stops."""
    F = _File(path)
    p = {}
    F.skip(1)
    p["title"] = F.record().s
    p["iyr0"], p["imon0"], p["iday0"], p["ihr0"] = [int(v) for v in F.values(4)]
    vals = F.values(12)
    for k, v in zip(["xl", "yl", "zl", "tl", "idx", "idy", "idz", "dzmin", "datadj", "zetainit", "idt",
                     "ibathyf"], vals):
        p[k] = v
    p["ibathyf"] = int(p["ibathyf"])
    # the 26 solution parameters, one record each; detect a file without icw
    F.skip(3)
    names = list(_SIM_NAMES)
    q = 0
    while q < len(names):
        name = names[q]
        r = F.record()
        if name == "icw" and r.label().lower() == "cw":
            if not missing_icw:
                raise InputError(f"{r.where}: expected the 'icw' record, found 'cw' (a file written before icw "
                                 "existed; read it with missing_icw=True)")
            p["icw"] = 0
            q += 1
            name = names[q]
        r.skip(14)
        p[name] = r.int(20) if name in _INT_NAMES else r.real(20)
        q += 1
    # time-series nodes
    p["ipt"], p["nnodes"] = [int(v) for v in F.values(2, integer=True)]
    if p["nnodes"] > 0:
        p["inode"] = F.ilist()[:p["nnodes"]]
        p["jnode"] = F.ilist()[:p["nnodes"]]
    else:
        p["inode"] = p["jnode"] = []
        F.skip(2)
    # horizontal planes
    p["iop"] = F.values(1, integer=True)[0]
    p["itspfh"] = F.value(integer=True)
    p["n_planes"] = F.value(integer=True)
    p["p_out"] = F.ilist()[:p["n_planes"]] if p["n_planes"] > 0 else (F.skip(1) or [])
    # vertical sections
    p["iox"] = F.values(1, integer=True)[0]
    p["sections"] = []
    if p["iox"] != 0:
        ns = F.value(integer=True)
        for _ in range(ns):
            F.skip(1)
            nc = F.value(integer=True)
            xs, ys = [], []
            for dst in (xs, ys):
                while len(dst) < nc:  # (14X,10I5), continued on further records
                    r = F.record()
                    r.skip(14)
                    for _ in range(min(10, nc - len(dst))):
                        dst.append(r.int(5))
            p["sections"].append((xs, ys))
    # 3-D output
    p["ipxml"] = F.values(1, integer=True)[0]
    p["itspf"] = F.value(integer=True)
    p["iTurbVars"] = F.value(integer=True)
    # open boundaries
    p["nopen"] = F.values(1, integer=True)[0]
    if p["nopen"] > 0:
        p["dtsecopenbc"] = F.value()
        for k in ("iside", "itype", "isbc", "jsbc", "iebc", "jebc"):
            p[k] = F.ilist()[:p["nopen"]]
    else:
        F.skip(7)
    # nested-grid boundary output (ioNBTOGGLE = 1, si3d_types.f90:246)
    p["nxNBO"] = F.values(1, integer=True)[0]
    if p["nxNBO"] > 0:
        p["ioNBO"] = F.value()
        p["xxNBO"] = F.value()
        for k in ("isdNBO", "isbcNBO", "jsbcNBO", "iebcNBO", "jebcNBO"):
            p[k] = F.ilist()[:p["nxNBO"]]
    else:
        F.skip(7)
    # tracers
    p["ntr"] = F.values(1, integer=True)[0]
    if p["ntr"] > 0:
        for k in ("ecomod", "iotr", "itspftr", "ipwq", "nswq"):
            p[k] = F.value(integer=True)
    else:
        p["ecomod"] = 0
        F.skip(5)
    # point sources and sinks
    p["iopss"] = F.values(1, integer=True)[0]
    if p["iopss"] > 0:
        p["npssdev"] = F.value(integer=True)
        p["dtsecpss"] = F.value()
        for k in ("ipss", "jpss", "iodev"):
            p[k] = F.ilist()[:p["iopss"]]
    else:
        F.skip(5)
    if p["ifsurfbc"] >= 10:
        p["iinterp"] = F.values(1, integer=True)[0]
        if p["iinterp"] == 2:
            p["gammaB"] = F.value()
            p["delNfactor"] = F.value()
    # derived sizes (si3d_utils.f90:303-310)
    dt = p["idt"]
    p["im"] = int(round(p["xl"] / p["idx"])) + 1
    p["jm"] = int(round(p["yl"] / p["idy"])) + 1
    p["nts"] = int(p["tl"] / dt + 0.5)
    return p


def read_bathy(path, im=None, jm=None):
    """si3d_bathy: (imm, jmm, depths) with depths[i-2, j-2] = h4(i, j) in dm for
    i = 2..imm+1, j = 2..jmm+1.  With im/jm (from si3d_inp.txt) the header is
    checked as the reference does."""
    F = _File(path)
    r = F.record()
    r.skip(37)
    imm = r.int(5)
    r.skip(6)
    jmm = r.int(5)
    r.skip(8)
    ncols = r.int(5)
    F.skip(2)
    if im is not None and (im != imm + 1 or jm != jmm + 1):
        raise InputError(f"{os.path.basename(path)}: grid size from si3d_inp.txt (im={im}, jm={jm}) does not "
                         f"agree with the bathymetry header (imm+1={imm + 1}, jmm+1={jmm + 1})")
    if ncols <= 0:
        raise InputError(f"{os.path.basename(path)}: ncols = {ncols}")
    im_, jm_ = imm + 1, jmm + 1
    h4 = np.zeros((im_ + 2, jm_ + 2))  # 1-based (i, j)
    ia = -(ncols - 1) + 1
    for _ in range(imm // ncols + 1):
        ia += ncols
        ib = ia + ncols - 1
        if ia > im_:
            break
        ib = min(ib, im_)
        for j in range(jm_, 1, -1):
            r = F.record()
            r.skip(5)
            for i in range(ia, ib + 1):
                h4[i, j] = r.real(5)
    d = h4[2:im_ + 1, 2:jm_ + 1]
    if np.any(d != np.round(d)):
        raise InputError(f"{os.path.basename(path)}: depths must be whole decimetres")
    return imm, jmm, d.astype(np.int32)


def read_layers(path):
    """si3d_layer.txt: zlevel(1..km1), the depths of the layer tops."""
    F = _File(path)
    F.skip(3)
    r = F.record()
    r.skip(10)
    km1 = r.int(11)
    z = []
    for _ in range(km1):
        r = F.record()
        r.skip(10)
        z.append(r.real(11))
    return np.array(z)


def read_init(path, km1, ntr=0):
    """si3d_init.txt: the scalar profile (km1) and the ntr tracer profiles."""
    F = _File(path)
    F.skip(7)  # five header records, the tracer line, one more record
    out = np.zeros((ntr + 1, km1))
    for k in range(km1):
        r = F.record()
        r.skip(10)
        for nn in range(ntr + 1):
            out[nn, k] = r.real(11)
    return out[0], out[1:]


def read_surfbc(path, nvar):
    """si3d_surfbc.txt (ifsurfbc 1: nvar 6; 2, 3: 9) or si3d_surfbc_W.txt
    (ifsurfbc 20: nvar 3) as an (npsurfbc, nvar) array."""
    F = _File(path)
    F.skip(6)
    r = F.record()
    r.skip(10)
    npts = r.int(7)
    out = np.zeros((npts, nvar))
    for j in range(npts):
        r = F.record()
        r.skip(10)
        for q in range(nvar):
            out[j, q] = r.real(11)
    return out


_SURFBC_NVAR = {1: 6, 2: 9, 3: 9, 20: 3}


def read_met_surfbc(path):
    """si3d_surfbc.txt of the met-station modes (ifsurfbc 10, 11;
    si3d_boundaryconditions.f90:4301-4450): six header records, the station
    count in (10X,I7), the stations' x, y (grid units) in (10X, 2n G11.2),
    the record count in (10X,I7), then records of 6 n + 2 values in
    (10X, G11.2 ...): eta, Pa, and per station Qsw, Ta, Rh, Cc (10) or Qlw
    (11), uair, vair.  Returns (metxy as n x 2, series as npts x (6 n + 2))."""
    F = _File(path)
    F.skip(6)
    r = F.record()
    r.skip(10)
    n = r.int(7)
    r = F.record()
    r.skip(10)
    xy = np.array([r.real(11) for _ in range(2 * n)]).reshape(n, 2)
    r = F.record()
    r.skip(10)
    npts = r.int(7)
    nv = 6 * n + 2
    out = np.zeros((npts, nv))
    for j in range(npts):
        r = F.record()
        r.skip(10)
        for q in range(nv):
            out[j, q] = r.real(11)
    return xy, out


def read_openbc(path, nvar):
    """openbcNN.txt (si3d_boundaryconditions.f90:55-86): five header records,
    the record count in (10X,I7), then records of nvar = ntr + 2 values in
    (10X, nvar G11.2): elevation or flow, active scalar, tracers."""
    F = _File(path)
    F.skip(5)
    r = F.record()
    r.skip(10)
    npts = r.int(7)
    out = np.zeros((npts, nvar))
    for j in range(npts):
        r = F.record()
        r.skip(10)
        for q in range(nvar):
            out[j, q] = r.real(11)
    return out


def openbc_name(nn):
    """File of open boundary nn (1-based): openbc01.txt .. openbc99.txt."""
    return f"openbc{nn:02d}.txt"


def read_case(directory, bathy_file="si3d_bathy", missing_icw=False):
    """A deck from a pSi3D case directory (si3d_inp.txt, the bathymetry, and
    si3d_layer.txt / si3d_init.txt / the surface series as the options ask).
    Open boundaries of types 1 and 2 come with their openbcNN.txt series.
    Raises InputError for options this path does not run (submerged-flow and
    nested-grid boundaries, point sources, water-quality or tracer-cloud
    models, GOTM)."""
    p = read_inp(os.path.join(directory, "si3d_inp.txt"), missing_icw=missing_icw)
    why = []
    if p["nopen"] > 0 and any(t not in (1, 2) for t in p["itype"]):
        why.append("open boundaries of type 3 or 4 (submerged flow, nested grids)")
    if p["iopss"] > 0:
        why.append("point sources and sinks (iopss > 0)")
    if p["ntr"] > 0 and p["ecomod"] != 0:
        why.append(f"tracer model ecomod = {p['ecomod']}")
    if p["ifsurfbc"] not in (0, 1, 2, 3, 10, 11, 20):
        why.append(f"ifsurfbc = {p['ifsurfbc']}")
    if p["iturb"] < 0:
        why.append("GOTM turbulence (iturb < 0)")
    if why:
        raise InputError("case uses options this path does not run: " + ", ".join(why))
    imm, jmm, bathy = read_bathy(os.path.join(directory, bathy_file), p["im"], p["jm"])
    d = lake._deck(
        name=os.path.basename(os.path.normpath(directory)),
        imm=imm, jmm=jmm, dx=p["idx"], dy=p["idy"], ddz=p["idz"], zl=p["zl"], dt=p["idt"], nts=p["nts"],
        ibathyf=p["ibathyf"], dzmin=p["dzmin"], datadj=p["datadj"], zetainit=p["zetainit"],
        bathy_dm=bathy,
        **{k: p[k] for k in ("itrap", "niter", "ismooth", "beta", "iturb", "Av0", "Dv0", "iadv", "itrmom",
                             "ihd", "Ax0", "Ay0", "f", "theta", "ibc", "isal", "cd", "ifsurfbc", "dtsurfbc",
                             "cw", "wa", "phi")},
    )
    if p["ibathyf"] < 0:
        d["zlevel"] = read_layers(os.path.join(directory, "si3d_layer.txt"))
    km1 = lake.km1_of(d)
    prof, tr = read_init(os.path.join(directory, "si3d_init.txt"), km1, p["ntr"])
    d["init_profile"] = prof
    if p["ntr"] > 0:
        d["tracer_profiles"] = tr
    if p["ifsurfbc"] in _SURFBC_NVAR:
        name = "si3d_surfbc_W.txt" if p["ifsurfbc"] == 20 else "si3d_surfbc.txt"
        d["surfbc"] = read_surfbc(os.path.join(directory, name), _SURFBC_NVAR[p["ifsurfbc"]])
    elif p["ifsurfbc"] in (10, 11):
        xy, series = read_met_surfbc(os.path.join(directory, "si3d_surfbc.txt"))
        d.update(surfbc=series, metxy=xy.ravel(), nmetstat=len(xy), iinterp=p["iinterp"], icw=p["icw"],
                 gammaB=p.get("gammaB", 0.0), delNfactor=p.get("delNfactor", 0.0))
    if p["nopen"] > 0:
        series = [read_openbc(os.path.join(directory, openbc_name(nn + 1)), p["ntr"] + 2)
                  for nn in range(p["nopen"])]
        if len({a.shape[0] for a in series}) != 1:  # varsOpenBC is sized by the first file (74-77)
            raise InputError("openbcNN.txt files must hold the same number of records")
        d["openbc"] = dict(dtsec=p["dtsecopenbc"], values=np.stack(series),
                           **{k: list(p[k]) for k in ("iside", "itype", "isbc", "jsbc", "iebc", "jebc")})
    d["inp"] = p
    return d


# ---- writers: a deck as pSi3D input files -----------------------------------
def _fmt_real(v):
    s = repr(float(v))
    if len(s) > 19:
        s = f"{float(v):.12g}"
    return s


def _line(label, value, comment=""):
    return f" {label:<12}!{value:>20}  ! {comment}\n".rstrip() + "\n"


def _section(title):
    return ["-------------------------------------------------------\n", f"{title}\n",
            "-------------------------------------------------------\n"]


def write_inp(deck, path, title="si3d case written by si3d_amd", ntr=0, start=(2000, 1, 1, 0)):
    """si3d_inp.txt in the record layout `input` reads (si3d_utils.f90:21-315):
    output sections, nesting output and point sources off; the open
    boundaries of deck["openbc"], if any."""
    imm, jmm = int(deck["imm"]), int(deck["jmm"])
    xl, yl = imm * deck["dx"], jmm * deck["dy"]  # im = nint(xl/dx) + 1 = imm + 1
    L = ["*******************************************************\n", f"{title}\n",
         "=======================================================\n"]
    L += _section("Start date and time for simulations")[1:]
    L += [_line(k, str(v)) for k, v in zip(("year", "month", "day", "hour"), start)]
    L += _section("Space-time domains, cell size & time steps")
    for lab, v in (("xl", xl), ("yl", yl), ("zl", deck["zl"]), ("tl", deck["nts"] * deck["dt"]),
                   ("idx", deck["dx"]), ("idy", deck["dy"]), ("idz", deck["ddz"]), ("dzmin", deck["dzmin"]),
                   ("datadj", deck["datadj"]), ("zeta0", deck["zetainit"]), ("idt", deck["dt"])):
        L.append(_line(lab, _fmt_real(v)))
    L.append(_line("ibathf", str(int(deck["ibathyf"]))))
    L += _section("Parameters controlling solution algorithm")
    labels = dict(ismooth="smooth", Av0="az0", Dv0="dz0", Ax0="ax0", Ay0="ay0", itrsca="itrsch", ifsurfbc="ifsbc",
                  dtsurfbc="dtsbc", wa="ws", num_threads="nth")
    for name in _SIM_NAMES:
        v = {"itrsca": 4, "idbg": 0, "num_threads": 1}.get(name, deck.get(name, 0))
        L.append(_line(labels.get(name, name), str(int(v)) if name in _INT_NAMES else _fmt_real(v)))
    L += _section("Output specifications for time files (individual nodes)")
    L += [_line("ipt", "0"), _line("nnodes", "0"), " inodes      !\n", " jnodes      !\n"]
    L += _section("Output specifications for H-plane files")
    L += [_line("iop", "0"), _line("itspfh", "0"), _line("nplanes", "0"), " planes      !\n"]
    L += _section("Output specifications for X-section files")
    L += [_line("ixt", "0")]
    L += _section("Output specifications for 3D space files")
    L += [_line("ipxml", "0"), _line("itspf", "0"), _line("iTurbVars", "0")]
    L += _section("Open boundary conditions specifications")
    ob = deck.get("openbc")
    if ob is not None and len(ob["iside"]) > 0:
        L += [_line("nopen", str(len(ob["iside"]))), _line("dtsecopenbc", _fmt_real(ob["dtsec"]))]
        L += [_line(k, ",".join(str(int(v)) for v in ob[k])) for k in ("iside", "itype", "isbc", "jsbc", "iebc",
                                                                       "jebc")]
    else:
        L += [_line("nopen", "0")] + [f" {k:<12}!\n" for k in ("dtsecopenbc", "iside", "itype", "isbc", "jsbc",
                                                                "iebc", "jebc")]
    L += _section("Open boundary conditions for nesting procedures")
    L += [_line("nxBDO", "0")] + [f" {k:<12}!\n" for k in ("iob", "xxb", "iside", "isbc", "jsbc", "iebc",
                                                            "jebc")]
    L += _section("Specification for tracer (other than oxygen) simulation")
    L += [_line("ntr", str(ntr)), _line("ecomod", "0"), _line("iotr", "0"), _line("itspftr", "0"),
          _line("ipwq", "0"), _line("nswq", "0")]
    L += _section("Specification for oxygenation system simulations")
    L += [_line("iopts", "0")] + [f" {k:<12}!\n" for k in ("Nodev", "iodt", "ipss", "jpss", "iodev")]
    if deck.get("ifsurfbc", 0) >= 10:  # read for every ifsurfbc >= 10, the wind series (20) included
        L += _section("Specification for interpolation method")
        ii = int(deck.get("iinterp", 1))
        L += [_line("iinterp", str(ii))]
        if ii == 2:
            L += [_line("gammaB", _fmt_real(deck["gammaB"])), _line("delNfactor", _fmt_real(deck["delNfactor"]))]
    with open(path, "w") as f:
        f.writelines(L)


def write_bathy(bathy_dm, path, title="si3d_amd bathymetry"):
    """si3d_bathy with every column in one block (ncols = imm)."""
    b = np.asarray(bathy_dm, dtype=np.int64)
    imm, jmm = b.shape
    if b.min() < -9999 or b.max() > 99999:
        raise InputError("depths do not fit the 5-character fields of si3d_bathy")
    head = f"{title[:36]:<37}{imm:5d}{' jmx=':>6}{jmm:5d}{' ncols=':>8}{imm:5d}\n"
    L = [head, "HV   " + "    V" * imm + "\n", "     " + "".join(f"{i:5d}" for i in range(2, imm + 2)) + "\n"]
    for j in range(jmm + 1, 1, -1):
        L.append(f"{j:5d}" + "".join(f"{int(v):5d}" for v in b[:, j - 2]) + "\n")
    with open(path, "w") as f:
        f.writelines(L)


def write_layers(zlevel, path):
    z = np.asarray(zlevel, dtype=np.float64)
    L = ["Depths to top of layers in Si3D Grid\n", "** used if ibathyf in si3d_inp.txt is set to < 0\n",
         "------------------------------------------------------\n", f"{'   km1   =':<10}{len(z):11d}\n"]
    for k, v in enumerate(z, start=1):
        s = f"{float(np.float32(v)):11.9g}"
        if len(s) > 11:
            s = f"{v:11.4e}"
        L.append(f"{k:10.2f}{s}\n")
    with open(path, "w") as f:
        f.writelines(L)


def write_init(profile, path, tracers=None):
    cols = [np.asarray(profile, dtype=np.float64)] + ([] if tracers is None else list(np.asarray(tracers)))
    L = ["Initial condition file for si3d model\n", "written by si3d_amd\n", "-\n", "Depths (m) not used\n",
         "-\n", "--------------------------------------------------\n", "     dummy\n"]
    for k in range(len(cols[0])):
        L.append(f"{0.0:10.2f}" + "".join(f"{c[k]:11.5f}" for c in cols) + "\n")
    with open(path, "w") as f:
        f.writelines(L)


def write_surfbc(series, dtsurfbc, path, metxy=None):
    """si3d_surfbc*.txt; with `metxy` (n x 2) the met-station layout of
    ifsurfbc 10 / 11 (station count and positions before the record count)."""
    s = np.asarray(series, dtype=np.float64)
    L = ["Surface boundary condition file for si3d model\n", "written by si3d_amd\n", "-\n", "-\n", "-\n", "-\n"]
    if metxy is not None:
        xy = np.asarray(metxy, dtype=np.float64).ravel()
        L += [f"   nmet = {len(xy) // 2:7d}\n", " stations " + "".join(f"{v:11.4f}" for v in xy) + "\n"]
    L += [f"   npts = {s.shape[0]:7d}\n"]
    for j in range(s.shape[0]):
        L.append(f"{j * dtsurfbc / 3600.0:10.4f}" + "".join(f"{v:11.4e}" for v in s[j]) + "\n")
    with open(path, "w") as f:
        f.writelines(L)


def write_openbc(series, dtsec, path):
    s = np.asarray(series, dtype=np.float64)
    L = ["Open boundary condition file for si3d model\n", "written by si3d_amd\n", "-\n", "-\n", "-\n",
         f"   npts = {s.shape[0]:7d}\n"]
    for j in range(s.shape[0]):
        L.append(f"{j * dtsec / 3600.0:10.4f}" + "".join(f"{v:11.4e}" for v in s[j]) + "\n")
    with open(path, "w") as f:
        f.writelines(L)


def write_case(deck, directory):
    """The deck as a pSi3D case directory (si3d_inp.txt, si3d_bathy, and
    si3d_layer.txt / si3d_init.txt / the surface series as needed), readable
    by `read_case` and by the reference's own `input`."""
    os.makedirs(directory, exist_ok=True)
    ntr = 0 if deck.get("tracer_profiles") is None else int(np.asarray(deck["tracer_profiles"]).shape[0])
    write_inp(deck, os.path.join(directory, "si3d_inp.txt"), ntr=ntr)
    write_bathy(deck["bathy_dm"], os.path.join(directory, "si3d_bathy"))
    if deck["ibathyf"] < 0:
        write_layers(deck["zlevel"], os.path.join(directory, "si3d_layer.txt"))
    write_init(deck["init_profile"], os.path.join(directory, "si3d_init.txt"), deck.get("tracer_profiles"))
    if deck.get("surfbc") is not None:
        name = "si3d_surfbc_W.txt" if deck["ifsurfbc"] == 20 else "si3d_surfbc.txt"
        write_surfbc(deck["surfbc"], deck["dtsurfbc"], os.path.join(directory, name),
                     metxy=deck.get("metxy") if deck["ifsurfbc"] in (10, 11) else None)
    if deck.get("openbc") is not None:
        for nn, series in enumerate(np.asarray(deck["openbc"]["values"])):
            write_openbc(series, deck["openbc"]["dtsec"], os.path.join(directory, openbc_name(nn + 1)))
