"""The binding's matmul: dtype and shape rules that are decided before any
device work (ferray-python/src/linalg.rs:1219-1284), and DeviceArray's
argument checks.  No GPU needed."""
from __future__ import annotations

import numpy as np
import pytest

from ferray_amd import binding
from ferray_amd.device import DeviceArray
from ferray_amd.products import ShapeMismatch


@pytest.mark.parametrize("dt", [np.int32, np.int64, np.bool_])
def test_integer_dtypes_without_a_kernel_raise(dt):
    a = np.ones((3, 4), dt)
    with pytest.raises(TypeError, match="no device kernel"):
        binding.matmul(a, a.T.copy())


def test_integer_promotion_to_a_wider_type_raises():
    # int8 x int32 promotes to int32, which has no matrix-core kernel
    with pytest.raises(TypeError):
        binding.matmul(np.ones((2, 2), np.int8), np.ones((2, 2), np.int32))


def test_integer_shape_errors_before_device_work():
    with pytest.raises(ShapeMismatch):
        binding.matmul(np.ones(3, np.int8), np.ones(4, np.int8))
    with pytest.raises(ShapeMismatch):
        binding.matmul(np.ones((2, 3), np.int16), np.ones((4, 2), np.int16))
    with pytest.raises(ShapeMismatch):
        binding.matmul(np.ones((2, 2, 3), np.int8), np.ones(3, np.int8))


def test_unsupported_float_dtype_raises():
    with pytest.raises(TypeError):
        binding.matmul(np.ones((2, 2), np.float16), np.ones((2, 2), np.float16))


def test_device_array_argument_checks():
    with pytest.raises(TypeError):
        DeviceArray((2, 2), np.float16)
    with pytest.raises(ValueError):
        DeviceArray((2, -1), np.float64)
