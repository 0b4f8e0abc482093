"""GPU tests through the test hooks of include/tt_testing.h.

* The line kernel's short square root (v_sqrt_f32 plus one residual test per
  neighbour, used for images whose samples cannot fall in (0, 2^-96) or be
  negative) equals the correctly rounded sqrtf on every one of the 2^32
  float bit patterns outside those exceptions (NaN results compare equal);
  c++/src/functionals.cpp:203 feeds every square-root median with cwiseSqrt.
* A failing workspace allocation surfaces as TT_ERR_NOMEM (a status, not a
  crash), and the context keeps working once allocations succeed again.
"""
import ctypes as C

import numpy as np
import pytest
import torch

from tracetransform_amd import _lib, core

pytestmark = pytest.mark.gpu


def test_short_sqrt_equals_sqrtf_on_all_floats(gpu):
    L = _lib.load()
    m, first = C.c_uint64(), C.c_uint32()
    total = 0
    for half in range(2):  # the 2^32 patterns in two calls
        lo = half << 31
        st = L.tt_testing_short_sqrt(gpu.handle, lo, 1 << 31, C.byref(m), C.byref(first))
        assert st == _lib.TT_OK, L.tt_last_error()
        print(f"[short sqrt] patterns {lo:#x}+2^31: {m.value} mismatches", flush=True)
        assert m.value == 0, hex(first.value)
        total += m.value
    assert total == 0


def test_allocation_failure_is_a_status(gpu):
    L = _lib.load()
    img = torch.rand((301, 301), device="cuda")
    sino = torch.empty((2, 30, 301), device="cuda")
    torch.cuda.synchronize()  # the fresh context runs on its own stream
    fresh = core.Context(0)  # a context with no workspaces yet
    codes = (C.c_int32 * 2)(0, 3)
    try:
        L.tt_testing_fail_allocations(0)
        st = L.tt_sinograms(fresh.handle, C.c_void_p(img.data_ptr()), 301, 30, 12.0, codes,
                            2, C.c_void_p(sino.data_ptr()), None)
        assert st == _lib.TT_ERR_NOMEM, L.tt_last_error()
        assert b"out of memory" in L.tt_last_error().lower() or b"failed" in L.tt_last_error()
    finally:
        L.tt_testing_fail_allocations(-1)
    st = L.tt_sinograms(fresh.handle, C.c_void_p(img.data_ptr()), 301, 30, 12.0, codes, 2,
                        C.c_void_p(sino.data_ptr()), None)
    assert st == _lib.TT_OK, L.tt_last_error()
    fresh.synchronize()
    ref = core.sinograms(img, 30, 12.0, [0, 3])
    torch.cuda.synchronize()
    assert torch.equal(sino, ref)
