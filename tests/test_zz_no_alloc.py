"""s3g_unet_check_static, the variant of s3g_unet_check that allocates no
device memory (its buffers live in a static device array), against
s3g_unet_check through the C-ABI: two calls in a row, the second on a
smaller input, each identical to the allocating entry point bit for bit
(same kernels, same inputs, fixed-order reductions), and an input over the
4096-cell limit refused with a clear error."""
import numpy as np
import pytest

from oracle import unet_oracle as U

pytestmark = pytest.mark.gpu


def _case(H, W, seed):
    rng = np.random.default_rng(seed)
    w = U.init_weights(seed + 4)
    x3 = rng.normal(size=(3, H, W)).astype(np.float32)
    target = rng.normal(size=(H, W)).astype(np.float32)
    mask = (rng.random((H, W)) > 0.3).astype(np.float32)
    return w, x3, target, mask


def test_static_variant_matches_allocating_entry_point():
    from si3d_amd.model import unet_check
    for H, W, seed in [(64, 64, 11), (16, 24, 12)]:
        print("case: s3g_unet_check_static", H, "x", W, flush=True)
        case = _case(H, W, seed)
        out_s, g_s, loss_s = unet_check(*case, static=True)
        out_a, g_a, loss_a = unet_check(*case)
        assert np.all(np.isfinite(out_s)) and np.abs(g_s).max() > 0
        np.testing.assert_array_equal(out_s, out_a)
        np.testing.assert_array_equal(g_s, g_a)
        assert loss_s == loss_a


def test_static_variant_refuses_inputs_over_its_limit():
    from si3d_amd.model import unet_check
    with pytest.raises(RuntimeError, match="4096 cells"):
        unet_check(*_case(72, 64, 3), static=True)
