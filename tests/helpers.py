"""Shared helpers of the test-suite: small lakes and field comparison."""
import numpy as np

from si3d_amd import lake

FIELDS_3D = ["h", "hp", "hpp", "hu", "hup", "hupp", "hv", "hvp", "hvpp", "uh", "uhp", "uhpp",
             "u", "up", "upp", "vh", "vhp", "vhpp", "v", "vp", "vpp", "wp", "sal", "salp",
             "salpp", "rhop", "Av", "Dv"]
FIELDS_TURB = ["q2", "q2p", "q2pp", "q2l", "q2lp", "q2lpp"]
FIELDS_2D = ["s", "sp", "spp"]
INTS = ["kmz", "k1z", "k1u", "k1v"]


def small_rectangle(nts=20, phi=240.0):
    return lake.rectangle(nx=20, ny=10, dx=50.0, depth=8.0, dz=1.0, nts=nts, phi=phi)


def small_ellipse(nts=20, seed=3, nx=40, ny=24, nlayers=12):
    d = lake.ellipse(nx=nx, ny=ny, dx=100.0, nlayers=nlayers, nts=nts, top=1.0, bottom=12.0,
                     depth=40.0, seed=seed)
    return d


def relerr(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    scale = max(np.max(np.abs(b)), 1e-30)
    return float(np.max(np.abs(a - b)) / scale) if a.size else 0.0


# u- and v-families share one scale: a component that is only rounding noise
# (e.g. v under a wind along x) is measured against the flow, not against itself.
_PARTNER = {"uh": "vh", "uhp": "vhp", "uhpp": "vhpp", "u": "v", "up": "vp", "upp": "vpp",
            "agx": "agy", "arx": "ary", "eagx": "eagy", "earx": "eary"}
_PARTNER.update({v: k for k, v in list(_PARTNER.items())})


def compare(gpu, orc, names, tol, where=""):
    bad = []
    for n in names:
        b = orc.get(n)
        if n in _PARTNER:
            scale = max(np.max(np.abs(b)), np.max(np.abs(orc.get(_PARTNER[n]))), 1e-30)
            e = float(np.max(np.abs(gpu.get(n) - b)) / scale)
        else:
            e = relerr(gpu.get(n), b)
        if not (e <= tol):
            bad.append((n, e))
    assert not bad, f"{where}: fields off beyond {tol}: {bad}"
