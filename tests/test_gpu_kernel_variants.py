"""Kernel variants must not change results.  The windowed per-cell kernels
(one lane per column and chunk of layers, buffer-resource loads, vertical
neighbours carried in registers) replace the generic ones (exmom for centred
momentum advection with ihd 0 or 1, exsal's face fluxes and right-hand side,
slayer_h, vel2's thicknesses and top-cell flux split, settrap, save);
S3G_WINDOW=0 keeps the generic ones.  Both evaluate the same expressions in
the same order, so every field must agree bit for bit after many steps, for
every layer-chunk size (S3G_TILES fixes it), in both precisions, with and
without horizontal diffusion, with tracers, and on a deck whose top layers
dry and wet (k1u, k1v move)."""
import os

import numpy as np
import pytest

from si3d_amd import lake
from tests.helpers import FIELDS_2D, FIELDS_3D, INTS, small_ellipse

pytestmark = pytest.mark.gpu


def _make(deck, precision, env):
    from si3d_amd.model import GpuLake
    old = {k: os.environ.get(k) for k in env}
    os.environ.update(env)
    try:
        return GpuLake(deck, precision=precision, preconditioner="mg", tol=1e-12)
    finally:
        for k, v in old.items():
            if v is None:
                del os.environ[k]
            else:
                os.environ[k] = v


def _tiles(kc):
    # tile words of exmom_x, exmom_y, slayer_h, vel, exsal_hflux, vel2,
    # exsal_rhs, turb_cell, settrap, save
    return ",".join([str(kc)] * 10)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("case", ["ellipse", "ellipse_ihd0", "tracers", "wetdry"])
def test_windowed_exmom_equals_generic(case, precision):
    print("case", case, precision, flush=True)
    if case == "wetdry":
        deck, nsteps = lake.shallow_basin(nts=200, stratified=True), 200
    else:
        deck, nsteps = small_ellipse(nts=20), 20
        if case == "ellipse_ihd0":
            deck["ihd"] = 0
        if case == "tracers":
            km1 = len(deck["init_profile"])
            deck["tracer_profiles"] = np.stack([np.where(np.arange(km1) < km1 // 2, 10.0, 2.0),
                                                np.linspace(0.0, 1.0, km1)])
    ref = _make(deck, precision, {"S3G_WINDOW": "0", "S3G_TILES": _tiles(1)})
    runs = [_make(deck, precision, {"S3G_TILES": _tiles(kc)}) for kc in (1, 3, 8)]
    moved = 0
    prev = ref.get("k1u").copy()
    for n in range(1, nsteps + 1):
        ref.step(n)
        for g in runs:
            g.step(n)
        k1u = ref.get("k1u")
        moved += int(np.sum(k1u != prev))
        prev = k1u.copy()
    names = FIELDS_3D + FIELDS_2D + INTS + (["tracer1", "tracer2"] if case == "tracers" else [])
    for g in runs:
        for name in names:
            a, b = g.get(name), ref.get(name)
            assert np.array_equal(a, b), (case, precision, name,
                                          float(np.max(np.abs(np.asarray(a, float) - np.asarray(b, float)))))
    if case == "wetdry":
        assert moved > 0, "the wet/dry deck must move the top u-layer"
