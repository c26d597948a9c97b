"""The binding's matmul on the device (ferray-python/src/linalg.rs:1219-1284).

Bars:
  * host float / complex results are bit-identical to the product layer's
    (same kernels), scalars come back as numpy scalars for 1-D x 1-D, and
    `b` is coerced to `a`'s dtype;
  * host integer results equal numpy's wrapping integer matmul exactly
    (int8, uint8, int16 and their promotions);
  * DeviceArray operands give bit-identical results to the host path and
    stay on the device (chains of products need no host copies); int8 /
    int16 device products are the exact wrapping int32.
"""
from __future__ import annotations

import numpy as np
import pytest

from ferray_amd import binding, products
from ferray_amd.device import DeviceArray
from tests.helpers import exact_int_matmul, rng, wrap_i32

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_host_float_ranks_and_scalars(gpu, dt):
    r = rng(1)
    a = r.standard_normal((37, 53)).astype(dt)
    b = r.standard_normal((53, 29)).astype(dt)
    v = r.standard_normal(53).astype(dt)
    w = r.standard_normal(37).astype(dt)
    assert binding.matmul(a, b).tobytes() == products.matmul(a, b).tobytes()
    assert binding.matmul(a, v).tobytes() == products.matmul(a, v).tobytes()
    assert binding.matmul(w, a).tobytes() == products.matmul(w, a).tobytes()
    s = binding.matmul(v, v)
    assert isinstance(s, np.generic) and s.dtype == dt
    assert abs(float(s) - float(v.astype(np.float64) @ v.astype(np.float64))) <= 1e-4 * float(v @ v)


def test_host_b_is_coerced_to_a_dtype(gpu):
    r = rng(2)
    a = r.standard_normal((16, 24)).astype(np.float32)
    b = r.standard_normal((24, 8))  # float64
    c = binding.matmul(a, b)
    assert c.dtype == np.float32
    assert c.tobytes() == products.matmul(a, b.astype(np.float32)).tobytes()


def test_host_batched_broadcast(gpu):
    r = rng(3)
    # trailing-batch broadcast, where the reference's i % batch indexing and
    # numpy's broadcasting pick the same operands
    a = r.standard_normal((2, 3, 20, 30))
    b = r.standard_normal((3, 30, 10))
    c = binding.matmul(a, b)
    assert c.shape == (2, 3, 20, 10)
    assert np.allclose(c, np.matmul(a, b), rtol=1e-12, atol=1e-12)
    # the reference's indexing (products/mod.rs:625-650) where the two differ
    a2 = r.standard_normal((3, 1, 20, 30))
    b2 = r.standard_normal((4, 30, 10))
    assert binding.matmul(a2, b2).tobytes() == products.matmul_batched(a2, b2).tobytes()


@pytest.mark.parametrize("dt", [np.complex128, np.complex64])
def test_host_complex(gpu, dt):
    r = rng(4)
    a = (r.standard_normal((12, 9)) + 1j * r.standard_normal((12, 9))).astype(dt)
    b = (r.standard_normal((9, 7)) + 1j * r.standard_normal((9, 7))).astype(dt)
    assert binding.matmul(a, b).tobytes() == products.matmul_complex(a, b).tobytes()
    u = (r.standard_normal(9) + 1j * r.standard_normal(9)).astype(dt)
    s = binding.matmul(u, u)  # unconjugated
    assert isinstance(s, np.generic)
    ref = np.sum(u.astype(np.complex128) * u.astype(np.complex128))
    assert abs(complex(s) - ref) <= 1e-4 * np.sum(np.abs(u) ** 2)


@pytest.mark.parametrize("da,db", [(np.int8, np.int8), (np.uint8, np.uint8), (np.int16, np.int16),
                                   (np.int8, np.int16), (np.uint8, np.int8)])
def test_host_integers_match_numpy_wrapping(gpu, da, db):
    r = rng(5)
    ia, ib = np.iinfo(da), np.iinfo(db)
    a = r.integers(ia.min, ia.max, (33, 70), endpoint=True).astype(da)
    b = r.integers(ib.min, ib.max, (70, 21), endpoint=True).astype(db)
    c = binding.matmul(a, b)
    ref = np.matmul(a, b)  # numpy: result dtype of the promotion, wrapping arithmetic
    assert c.dtype == ref.dtype and np.array_equal(c, ref)
    v = b[:, 0].copy()
    assert np.array_equal(binding.matmul(a, v), np.matmul(a, v))
    x = a[0].copy()
    s = binding.matmul(x, x)
    assert s == np.matmul(x, x) and np.asarray(s).dtype == np.matmul(x, x).dtype


@pytest.mark.parametrize("dt", [np.float64, np.float32, np.complex128, np.complex64])
def test_device_matches_host_bitwise(gpu, dt):
    r = rng(6)
    a = r.standard_normal((130, 257))
    b = r.standard_normal((257, 66))
    if np.dtype(dt).kind == "c":
        a = a + 1j * r.standard_normal(a.shape)
        b = b + 1j * r.standard_normal(b.shape)
    a, b = a.astype(dt), b.astype(dt)
    host = binding.matmul(a, b)
    dev = binding.matmul(DeviceArray.from_numpy(a), DeviceArray.from_numpy(b))
    assert isinstance(dev, DeviceArray) and dev.shape == (130, 66) and dev.dtype == dt
    assert dev.to_numpy().tobytes() == host.tobytes()


def test_device_chain_and_ranks(gpu):
    r = rng(7)
    a, b, c = (r.standard_normal(s) for s in ((64, 80), (80, 48), (48, 32)))
    da, db, dc = (DeviceArray.from_numpy(x) for x in (a, b, c))
    dev = binding.matmul(binding.matmul(da, db), dc)  # no host copy between the products
    host = binding.matmul(binding.matmul(a, b), c)
    assert dev.to_numpy().tobytes() == host.tobytes()
    v = r.standard_normal(80)
    dv = DeviceArray.from_numpy(v)
    assert binding.matmul(da, dv).to_numpy().tobytes() == binding.matmul(a, v).tobytes()
    s = binding.matmul(dv, dv)
    assert s.shape == () and s.to_numpy()[()] == binding.matmul(v, v)


def test_device_batched_matches_host(gpu):
    r = rng(8)
    a = r.standard_normal((5, 40, 24))
    b = r.standard_normal((1, 24, 16))
    dev = binding.matmul(DeviceArray.from_numpy(a), DeviceArray.from_numpy(b))
    assert dev.shape == (5, 40, 16)
    assert dev.to_numpy().tobytes() == products.matmul_batched(a, b).tobytes()


@pytest.mark.parametrize("dt", [np.int8, np.int16])
def test_device_integers_exact_int32(gpu, dt):
    r = rng(9)
    info = np.iinfo(dt)
    a = r.integers(info.min, info.max, (96, 300), endpoint=True).astype(dt)
    b = r.integers(info.min, info.max, (300, 40), endpoint=True).astype(dt)
    dev = binding.matmul(DeviceArray.from_numpy(a), DeviceArray.from_numpy(b))
    assert dev.dtype == np.int32
    assert np.array_equal(dev.to_numpy(), wrap_i32(exact_int_matmul(a, b)))


def test_device_operand_errors(gpu):
    a = DeviceArray.from_numpy(np.ones((4, 4)))
    with pytest.raises(TypeError):
        binding.matmul(a, np.ones((4, 4)))
    with pytest.raises(TypeError):
        binding.matmul(a, DeviceArray.from_numpy(np.ones((4, 4), np.float32)))
    with pytest.raises(TypeError):
        binding.matmul(DeviceArray.from_numpy(np.ones((4, 4), np.uint8)),
                       DeviceArray.from_numpy(np.ones((4, 4), np.uint8)))
    with pytest.raises(products.ShapeMismatch):
        binding.matmul(a, DeviceArray.from_numpy(np.ones((5, 2))))
