"""The device float -> int32 conversion every index computation of the path uses
(block and voxel coordinates, pixel projection) truncates toward zero, saturates
out-of-range values and maps NaN to 0, like the oracle's conversion; checked on edge
values (signed zeros, halves, the int32 limits, infinities, NaN payloads, subnormals)
and on random values against numpy."""
import numpy as np
import pytest

from vsg_amd import vsg

pytestmark = pytest.mark.gpu


def _reference(x: np.ndarray) -> np.ndarray:
    with np.errstate(invalid="ignore"):
        x64 = x.astype(np.float64)
    out = np.zeros(x.size, np.int64)
    ok = ~np.isnan(x64)
    out[ok] = np.clip(np.trunc(x64[ok]), -2.0 ** 31, 2.0 ** 31 - 1).astype(np.int64)
    return out.astype(np.int32)


def test_device_conversion_matches_saturating_truncation(torch_cuda):
    edge = np.array([0.0, -0.0, 0.5, -0.5, 0.99999994, -0.99999994, 1.0, -1.0, 1.5, -1.5, 2.5, -2.5,
                     8388607.5, -8388607.5, 16777217.0, 2147483520.0, -2147483520.0, 2147483648.0,
                     -2147483648.0, 3e9, -3e9, 1e30, -1e30, np.inf, -np.inf, 1e-40, -1e-40, 1e-45], np.float32)
    nans = np.array([0x7FC00000, 0xFFC00000, 0x7F800001, 0xFFFFFFFF], np.uint32).view(np.float32)
    rnd = np.random.default_rng(4).normal(0.0, 1e4, 4096).astype(np.float32)
    x = np.concatenate([edge, nans, rnd])
    got = vsg.diag_convert(x)
    assert np.array_equal(got, _reference(x))
