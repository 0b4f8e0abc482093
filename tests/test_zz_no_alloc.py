"""GPU test of tt_circus_noalloc (the allocation-free tt_circus): called
through the C ABI, its circus functions and P2 ranks equal tt_circus's bit
for bit on two calls in a row, the second on a smaller input (the shared
scratch keeps the first call's spectra), and requests beyond its scratch or
with Hermite are refused."""
import ctypes as C

import numpy as np
import pytest
import torch

from tracetransform_amd import _lib, core

pytestmark = pytest.mark.gpu

PF = [(1,), (3,), (2,), (3,)]


def noalloc(sino, pf):
    n_sino, n_cols, n = sino.shape
    out = torch.empty((n_sino, len(pf), n_cols), dtype=torch.float32, device=sino.device)
    rank = torch.empty((n_sino, len(pf), n_cols), dtype=torch.int32, device=sino.device)
    arr = (_lib.tt_pfunc * len(pf))(*[_lib.tt_pfunc(*(tuple(t) + (0, 0))[:3]) for t in pf])
    ctx = core.get_context(sino.device.index)
    ctx.bind_current_stream()
    st = _lib.load().tt_circus_noalloc(ctx.handle, C.c_void_p(sino.data_ptr()), n, n_cols,
                                       n_sino, arr, len(pf), C.c_void_p(out.data_ptr()),
                                       C.c_void_p(rank.data_ptr()))
    return st, out, rank


def test_noalloc_circus_matches_tt_circus(gpu):
    rng = np.random.default_rng(5)
    for n_sino, n_cols, n in ((4, 360, 1449), (2, 90, 363)):
        sino = torch.from_numpy(rng.random((n_sino, n_cols, n), dtype=np.float32)).cuda()
        print(f"[noalloc] n_sino={n_sino} n_cols={n_cols} n={n}", flush=True)
        st, got, grank = noalloc(sino, PF)
        assert st == _lib.TT_OK, _lib.load().tt_last_error()
        want, wrank = core.circus(sino, PF, with_p2_rank=True)
        torch.cuda.synchronize()
        assert torch.equal(got, want)
        assert torch.equal(grank[:, 2], wrank[:, 2])


def test_noalloc_circus_limits(gpu):
    sino = torch.zeros((9, 720, 1449), device="cuda")  # 9 * 720 * 1450 > 8 Mi floats
    st, _, _ = noalloc(sino, [(3,)])
    assert st == _lib.TT_ERR_INVALID
    st, _, _ = noalloc(sino[:1, :10], [(0, 2, 700)])
    assert st == _lib.TT_ERR_UNSUPPORTED
