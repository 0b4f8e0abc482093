"""CPU tests of the host side: the C-ABI library loads and exports every
symbol include/tt.h declares, the functional grammar of the reference, and
the seeded input generators."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from tracetransform_amd import PFunctional, PFunctionalWrapper, TFunctional, TFunctionalWrapper
from tracetransform_amd import _lib, inputs

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    """Entry points of include/tt.h and of the test hooks in
    include/tt_testing.h."""
    text = "".join(open(os.path.join(ROOT, "include", h)).read()
                   for h in ("tt.h", "tt_testing.h"))
    return sorted(set(re.findall(r"TT_API\s+[\w\s\*]+?\b(tt_\w+)\s*\(", text)))


def test_header_and_binding_agree():
    assert declared_symbols() == sorted(_lib.EXPORTED)


def test_library_exports_every_declared_symbol():
    L = _lib.load()
    for name in declared_symbols():
        assert hasattr(L, name), name
    # only the C ABI is exported (hidden visibility for everything else)
    out = os.popen(f"nm -D --defined-only {_lib.LIB_PATH}").read()
    exported = {ln.split()[-1] for ln in out.splitlines() if " T " in ln}
    assert {s for s in exported if s.startswith("tt_")} == set(declared_symbols())


def test_host_only_entry_points_without_gpu():
    L = _lib.load()
    n, dx, dy = C.c_int(), C.c_int(), C.c_int()
    assert L.tt_pad_geometry(256, 256, C.byref(n), C.byref(dx), C.byref(dy)) == 0
    assert (n.value, dx.value, dy.value) == (363, 54, 54)
    assert L.tt_pad_geometry(0, 5, C.byref(n), C.byref(dx), C.byref(dy)) == _lib.TT_ERR_INVALID
    assert b"non-empty" in L.tt_last_error()
    assert L.tt_angle_count(1.0) == 360 and L.tt_angle_count(0.5) == 720
    assert L.tt_angle_count(7.0) == 51 and L.tt_angle_count(0.0) == 0
    # a NULL context is rejected before any device call
    assert L.tt_sinograms(None, None, 7, 1, 1.0, None, 0, None, None) == _lib.TT_ERR_INVALID
    assert L.tt_circus(None, None, 7, 1, 1, None, 0, None, None) == _lib.TT_ERR_INVALID


def test_pad_geometry_matches_oracle(oracle):
    L = _lib.load()
    n, dx, dy = C.c_int(), C.c_int(), C.c_int()
    for rows, cols in [(1, 1), (3, 8), (64, 64), (100, 37), (512, 512), (1024, 1024)]:
        L.tt_pad_geometry(rows, cols, C.byref(n), C.byref(dx), C.byref(dy))
        assert (n.value, dx.value, dy.value) == oracle.pad_geometry(rows, cols)


def test_tfunctional_grammar():
    names = ["Radon", "T1", "T2", "T3", "T4", "T5", "T6", "T7"]
    for i, nm in enumerate(names):
        w = TFunctionalWrapper.parse(str(i))
        assert w.name == nm and w.functional == TFunctional(i)
    for bad in ("8", "", "T1", "-1", "01"):
        with pytest.raises(ValueError):
            TFunctionalWrapper.parse(bad)


def test_pfunctional_grammar():
    for i in (1, 2, 3):
        w = PFunctionalWrapper.parse(str(i))
        assert w.name == f"P{i}" and w.functional == PFunctional(i)
    w = PFunctionalWrapper.parse("H5")
    assert w.functional == PFunctional.Hermite and w.arguments.order == 5
    assert w.arguments.center is None
    for bad in ("0", "4", "H", "Hx", "P1", ""):
        with pytest.raises(ValueError):
            PFunctionalWrapper.parse(bad)


def test_splitmix64_known_values():
    # SplitMix64 reference outputs for seed 0 (Vigna's splitmix64.c)
    z = inputs.splitmix64(0, 3)
    assert [int(v) for v in z] == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4,
                                   0x06C45D188009454F]


def test_uniform_and_blob_images():
    u = inputs.uniform_image(16, 3)
    assert u.shape == (16, 16) and u.dtype == np.float32
    assert 0 <= u.min() and u.max() < 1
    assert np.array_equal(u, inputs.uniform_image(16, 3))
    b = inputs.blob_image(64, 2)
    assert b.shape == (64, 64) and b.min() >= 0 and b.max() > 0
    s = inputs.batch_sides(256, 64, 512, 3)
    assert s.min() >= 64 and s.max() <= 512 and len(set(s.tolist())) > 100
