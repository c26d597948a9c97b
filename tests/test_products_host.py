"""Shape rules of the product layer (no device needed: errors are raised
before any device call), following products/mod.rs and gemm/mod.rs."""
from __future__ import annotations

import numpy as np
import pytest

from ferray_amd import products as pr


def z(*shape, dt=np.float64):
    return np.zeros(shape, dtype=dt)


@pytest.mark.parametrize("fn,a,b", [
    (pr.matmul, z(3), z(3)),                 # 1-D x 1-D: use dot (mod.rs:371-373)
    (pr.matmul, z(3), z(4, 2)),              # vector-matrix misaligned
    (pr.matmul, z(2, 3), z(4)),              # matrix-vector misaligned
    (pr.matmul, z(2, 3), z(4, 5)),           # inner dims
    (pr.matmul, z(2, 2, 3), z(2, 4, 5)),     # batched inner dims
    (pr.matmul, z(2, 2, 3), z(3, 3, 5)),     # batch shapes 2 vs 3
    (pr.dot, z(3), z(4)),
    (pr.dot, z(2, 3), z(4, 5)),
    (pr.dot, z(2, 2, 3), z(4)),
    (pr.dot, z(2, 2, 3), z(4, 5, 6)),
    (pr.matmul_complex, z(2, 3, dt=np.complex128), z(4, 2, dt=np.complex128)),
    (pr.gemm, z(2, 3), z(4, 2)),
    (pr.gemm, z(2, 3, 1), z(3, 2)),
])
def test_shape_mismatch(fn, a, b):
    with pytest.raises(pr.ShapeMismatch):
        fn(a, b)


def test_gemm_requires_contiguous():
    a = np.zeros((4, 6))[:, ::2]
    with pytest.raises(pr.ShapeMismatch):
        pr.gemm(a, np.zeros((3, 2)))


def test_dtype_rules():
    with pytest.raises(TypeError):
        pr.matmul(z(2, 2), z(2, 2, dt=np.float32))
    with pytest.raises(TypeError):
        pr.matmul_complex(z(2, 2), z(2, 2))


def test_batch_broadcast_shapes():
    assert pr._broadcast_batch((2, 1), (3,)) == (2, 3)
    assert pr._broadcast_batch((), (4, 5)) == (4, 5)
    assert pr._broadcast_batch((1,), (7,)) == (7,)
    with pytest.raises(pr.ShapeMismatch):
        pr._broadcast_batch((2,), (3,))


def test_empty_products_need_no_device():
    assert pr.gemm(z(0, 3), z(3, 4)).shape == (0, 4)
    assert pr.gemm(z(2, 0), z(0, 4)).tolist() == [[0.0] * 4] * 2
    assert pr.matmul_batched(z(0, 2, 3), z(3, 4)).shape == (0, 2, 4)


def test_consumer_argument_errors():
    with pytest.raises(pr.ShapeMismatch):
        pr.matrix_power(z(2, 3), 2)
    with pytest.raises(ValueError):
        pr.matrix_power(z(2, 2), -1)
    assert pr.matrix_power(z(3, 3), 0).tolist() == np.eye(3).tolist()
    with pytest.raises(ValueError):
        pr.multi_dot([z(2, 2)])
    with pytest.raises(pr.ShapeMismatch):
        pr.multi_dot([z(2, 3), z(4, 5), z(5, 2)])
    with pytest.raises(pr.ShapeMismatch):
        pr.multi_dot([z(2, 3), z(3, 4, 1), z(4, 2)])
    with pytest.raises(NotImplementedError):
        pr.einsum_matmul("ij,jk", z(2, 2), z(2, 2))
    with pytest.raises(NotImplementedError):
        pr.einsum_matmul("ij,kj->ik", z(2, 2), z(2, 2))
    with pytest.raises(NotImplementedError):
        pr.einsum_matmul("ii,ij->ij", z(2, 2), z(2, 2))
    with pytest.raises(TypeError):
        pr.matmul_int(z(2, 2, dt=np.int8), z(2, 2, dt=np.int16))
    with pytest.raises(pr.ShapeMismatch):
        pr.matmul_int(z(2, 3, dt=np.int8), z(2, 2, dt=np.int8))
