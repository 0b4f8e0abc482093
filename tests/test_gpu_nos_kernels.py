"""FP64 matrix-core products of the nearest-orthonormal polar iteration
(tt_nos.hip), through the test hook tt_testing_nos_product, against FP64
matrix products on the CPU.

The iteration (c++/src/circus.cpp:94-99 takes the polar factor of the
aligned sinogram) needs the Gram matrices X^T X, products of polynomials of
one symmetric matrix (symmetric results: only the lower blocks are computed
and mirrored), and the tall X Q.  Each case covers one kernel path: the
LDS-tiled kernel (A transposed, or a reduction length that is not a multiple
of 4) and the direct-load kernel (symmetric right operand), symmetric and
general outputs, and the polynomial step's epilogue.  FP64 rounding of a
k-term sum keeps the normwise error near 1e-16 * sqrt(k); the bound is 1e-13.
Symmetric outputs must be exactly symmetric.
"""
import ctypes as C

import numpy as np
import pytest
import torch

from tracetransform_amd import _lib

pytestmark = pytest.mark.gpu

TOL = 1e-13


def col_major(mats):
    """[count, rows, cols] float64 -> device buffer of column-major matrices."""
    return torch.from_numpy(np.ascontiguousarray(np.transpose(mats, (0, 2, 1)))).cuda()


def from_col_major(buf, rows, cols):
    return np.transpose(buf.cpu().numpy().reshape(-1, cols, rows), (0, 2, 1))


def run(gpu, trans_a, sym, a, b, m, n, k, coef=None):
    count = a.shape[0]
    da, db = col_major(a), (None if b is None else col_major(b))
    dc = torch.zeros(count * m * n, dtype=torch.float64, device="cuda")
    dt = dq = None
    cptr = None
    if coef is not None:
        dt = torch.zeros_like(dc)
        dq = torch.zeros_like(dc)
        cptr = (C.c_double * 5)(*coef)
    torch.cuda.synchronize()
    L = _lib.load()
    st = L.tt_testing_nos_product(
        gpu.handle, int(trans_a), int(sym), count, m, n, k, C.c_void_p(da.data_ptr()),
        C.c_void_p((da if db is None else db).data_ptr()), C.c_void_p(dc.data_ptr()), cptr,
        None if dt is None else C.c_void_p(dt.data_ptr()),
        None if dq is None else C.c_void_p(dq.data_ptr()))
    assert st == _lib.TT_OK, L.tt_last_error()
    out = from_col_major(dc, m, n)
    if coef is None:
        return out
    return out, from_col_major(dt, m, n), from_col_major(dq, m, n)


def rel(got, ref):
    return np.linalg.norm(got - ref) / np.linalg.norm(ref)


def sym_rand(rng, count, k):
    s = rng.standard_normal((count, k, k))
    return 0.5 * (s + np.transpose(s, (0, 2, 1)))


@pytest.mark.parametrize("k", [641, 96])
def test_gram_products(gpu, k):
    # X^T X as the iteration forms it: one row chunk of X per matrix
    rng = np.random.default_rng(k)
    x = rng.standard_normal((3, k, 360))
    got = run(gpu, True, True, x, None, 360, 360, k)
    ref = np.transpose(x, (0, 2, 1)) @ x
    err = rel(got, ref)
    print(f"[nos gram] k={k} rel={err:.2e}", flush=True)
    assert err <= TOL
    assert np.array_equal(got, np.transpose(got, (0, 2, 1)))


def test_transposed_general_product(gpu):
    rng = np.random.default_rng(5)
    a = rng.standard_normal((2, 50, 100))   # k x m
    b = rng.standard_normal((2, 50, 72))    # k x n
    got = run(gpu, True, False, a, b, 100, 72, 50)
    assert rel(got, np.transpose(a, (0, 2, 1)) @ b) <= TOL


@pytest.mark.parametrize("k,count", [(360, 8), (360, 16), (72, 3), (90, 2)])
def test_symmetric_products_of_commuting_matrices(gpu, k, count):
    # P and Q polynomials of one symmetric S: P Q is symmetric.  k = 90 is
    # not a multiple of 4 and takes the LDS-tiled kernel, the others the
    # direct-load one
    rng = np.random.default_rng(k + count)
    s = sym_rand(rng, count, k) / np.sqrt(k)
    eye = np.eye(k)
    p = s @ s + 0.5 * s + eye
    q = 2.0 * s - 0.25 * (s @ s) + eye
    got = run(gpu, False, True, p, q, k, k, k)
    ref = p @ q
    err = rel(got, ref)
    print(f"[nos sym] k={k} count={count} rel={err:.2e}", flush=True)
    assert err <= TOL
    assert np.array_equal(got, np.transpose(got, (0, 2, 1)))


def test_tall_product_with_symmetric_right_operand(gpu):
    rng = np.random.default_rng(9)
    x = rng.standard_normal((2, 1920, 360))
    q = sym_rand(rng, 2, 360)
    got = run(gpu, False, False, x, q, 1920, 360, 360)
    assert rel(got, x @ q) <= TOL


def test_general_product_without_multiple_of_four(gpu):
    # reduction length 30: the LDS-tiled kernel, B not symmetric
    rng = np.random.default_rng(11)
    a = rng.standard_normal((2, 50, 30))
    b = rng.standard_normal((2, 30, 30))
    got = run(gpu, False, False, a, b, 50, 30, 30)
    assert rel(got, a @ b) <= TOL


def test_polynomial_step_epilogue(gpu):
    rng = np.random.default_rng(13)
    k = 360
    s = sym_rand(rng, 4, k) / np.sqrt(k)
    c = [7.28, -66.06, 209.0, -256.85, 107.08]
    s2, t, q = run(gpu, False, True, s, None, k, k, k, coef=c)
    ref2 = s @ s
    assert rel(s2, ref2) <= TOL
    assert np.array_equal(s2, np.transpose(s2, (0, 2, 1)))
    # the epilogue's elementwise terms from the kernel's own S S
    eye = np.eye(k)
    assert rel(t, c[3] * s + c[4] * s2) <= TOL
    assert rel(q, c[0] * eye + c[1] * s + c[2] * s2) <= TOL


def test_hook_rejects_bad_shapes(gpu):
    L = _lib.load()
    buf = torch.zeros(16, dtype=torch.float64, device="cuda")
    p = C.c_void_p(buf.data_ptr())
    assert L.tt_testing_nos_product(gpu.handle, 0, 1, 1, 4, 3, 4, p, p, p, None, None,
                                    None) == _lib.TT_ERR_INVALID
    coef = (C.c_double * 5)(1, 0, 0, 0, 0)
    assert L.tt_testing_nos_product(gpu.handle, 0, 1, 1, 3, 3, 3, p, p, p, coef, p,
                                    p) == _lib.TT_ERR_INVALID
