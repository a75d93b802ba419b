"""CPU-side checks of the product: the C-ABI library loads and exports every
entry point include/si3d_gpu.h declares, the deck helpers agree with the
oracle's grid, and the bench refuses multi-GPU runs."""
import ctypes
import os
import re
import subprocess
import sys

import numpy as np
import pytest

from si3d_amd import _lib, lake
from tests.helpers import small_ellipse, small_rectangle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _header_functions():
    txt = open(os.path.join(ROOT, "include", "si3d_gpu.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(s3g_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_entry_point():
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libsi3dgpu.so not built")
    lib = ctypes.CDLL(_lib.LIB_PATH)
    names = _header_functions()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert set(_lib.EXPORTS) <= set(names)


def test_product_loader_has_no_fallback():
    src = open(os.path.join(ROOT, "si3d_amd", "_lib.py")).read()
    assert "raise RuntimeError" in src
    for f in os.listdir(os.path.join(ROOT, "si3d_amd")):
        if f.endswith(".py"):
            assert "oracle" not in open(os.path.join(ROOT, "si3d_amd", f)).read().lower() or f == "lake.py"


@pytest.mark.parametrize("mk", [small_rectangle, small_ellipse])
def test_n_wet_matches_oracle_grid(mk, oracle_lib):
    d = mk()
    o = oracle_lib.Oracle(d)
    kmz = o.get("kmz")[:o.lm].astype(int)
    k1z = o.get("k1z")[:o.lm].astype(int)
    assert lake.n_wet_cells(d) == int(np.sum(kmz - k1z + 1))


def test_full_size_lake_shape():
    d = lake.ellipse(nts=2)
    assert d["bathy_dm"].shape == (1024, 512)
    assert len(d["zlevel"]) == 62
    wet = int((d["bathy_dm"] > 0).sum())
    assert 0.7 * np.pi / 4 * 1024 * 512 < wet < np.pi / 4 * 1024 * 512
    assert np.all(d["bathy_dm"][d["bathy_dm"] > 0] <= int(d["zlevel"][-1] * 10))


def test_bench_refuses_multi_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr


def test_ensemble_refuses_open_boundaries():
    """Open boundaries are per lake; the ensemble refuses a deck that has them
    (before any device work)."""
    from si3d_amd import lake
    from si3d_amd.model import GpuEnsemble
    d = lake.open_boundary(lake.channel(nx=12, ny=6), "W", 2, 2, 7, lake.steady_series(1.0, 20.0))
    with pytest.raises(ValueError, match="open boundaries"):
        GpuEnsemble([lake.channel(nx=12, ny=6), d])
