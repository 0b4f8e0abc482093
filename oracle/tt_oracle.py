"""numpy front-end of the CPU oracle (test infrastructure only).

Only tests/, ``__graft_entry__.smoke()`` and bench.py's ``cpu_baseline`` leg
import this module, and only as the checker or as the timed CPU port of the
reference's OpenMP variant.  The product (``tracetransform_amd``) never does.

The C restatement lives in ``tt_oracle.c`` (built into ``liboracle.so`` by
``oracle/Makefile``); this module adds numpy wrappers plus the FP64
nearest-orthonormal-sinogram of c++/src/circus.cpp:68-102 (numpy SVD).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")


def build(force: bool = False) -> str:
    """Compile liboracle.so with oracle/Makefile."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-B" if force else "-s", "-C", _HERE],
                       check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        vp = C.c_void_p
        L.tto_pad_geometry.argtypes = [C.c_int, C.c_int, C.POINTER(C.c_int),
                                       C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.tto_pad.argtypes = [_f32p, C.c_int, C.c_int, _f32p]
        L.tto_resize.argtypes = [_f32p, C.c_int, C.c_int, _f32p, C.c_int, C.c_int]
        L.tto_deg2rad.argtypes = [C.c_float]
        L.tto_deg2rad.restype = C.c_float
        L.tto_angle_deg.argtypes = [C.c_int, C.c_float]
        L.tto_angle_deg.restype = C.c_float
        L.tto_rotation_trig.argtypes = [C.c_float, C.POINTER(C.c_float),
                                        C.POINTER(C.c_float)]
        L.tto_interpolate.argtypes = [_f32p, C.c_int, C.c_float, C.c_float]
        L.tto_interpolate.restype = C.c_float
        L.tto_rotate.argtypes = [_f32p, C.c_int, C.c_float, C.c_float, _f32p]
        for name in ("tto_eigen_sum", "tto_seq_sum", "tto_T0", "tto_P1", "tto_P3"):
            getattr(L, name).argtypes = [_f32p, C.c_int]
            getattr(L, name).restype = C.c_float
        L.tto_weighted_median.argtypes = [_f32p, C.c_int]
        L.tto_weighted_median.restype = C.c_int
        L.tto_stable_argsort.argtypes = [_f32p, C.c_int, _i32p, _i32p]
        L.tto_t345_table.argtypes = [C.c_int, C.c_int, _f32p, _f32p]
        ip = C.POINTER(C.c_int)
        for name in ("tto_T1", "tto_T2"):
            getattr(L, name).argtypes = [_f32p, C.c_int, ip]
            getattr(L, name).restype = C.c_float
        L.tto_T345.argtypes = [_f32p, C.c_int, _f32p, _f32p, ip]
        L.tto_T345.restype = C.c_float
        L.tto_T6.argtypes = [_f32p, C.c_int, ip, ip, ip]
        L.tto_T6.restype = C.c_float
        L.tto_T7.argtypes = [_f32p, C.c_int, ip, ip, ip]
        L.tto_T7.restype = C.c_float
        L.tto_P2.argtypes = [_f32p, C.c_int, ip, ip]
        L.tto_P2.restype = C.c_float
        L.tto_hermite_function.argtypes = [C.c_uint, C.c_float]
        L.tto_hermite_function.restype = C.c_float
        L.tto_Phermite.argtypes = [_f32p, C.c_int, C.c_uint, C.c_int]
        L.tto_Phermite.restype = C.c_float
        L.tto_sinograms.argtypes = [_f32p, C.c_int, C.c_int, C.c_float, _i32p,
                                    C.c_int, _f32p, vp, vp, vp, vp, vp, vp, C.c_int]
        L.tto_circus.argtypes = [_f32p, C.c_int, C.c_int, _i32p, _i32p, _i32p,
                                 C.c_int, _f32p, vp, vp, C.c_int]
        L.tto_zscore.argtypes = [_f32p, C.c_int, _f32p]
        L.tto_sinograms_f64.argtypes = [_f32p, C.c_int, C.c_int, C.c_float, _i32p,
                                        C.c_int, vp, vp, vp, vp, C.c_int, _f64p,
                                        C.c_int]
        L.tto_circus_f64.argtypes = [_f64p, C.c_int, C.c_int, _i32p, _i32p, _i32p,
                                     C.c_int, vp, C.c_int, _f64p, C.c_int]
        L.tto_dft_f64.argtypes = [_f64p, C.c_int, _f64p, _f64p]
        L.tto_backward_f64.argtypes = [_f32p, C.c_int, C.c_int, C.c_float, _i32p,
                                       C.c_int, _i32p, _i32p, _i32p, C.c_int, vp, vp,
                                       vp, vp, vp, _f64p, _f64p, C.c_int]
        L.tto_backward_f64.restype = C.c_int
        L.tto_sinograms_vjp_f64.argtypes = [_f32p, C.c_int, C.c_int, C.c_float, _i32p,
                                            C.c_int, vp, vp, vp, vp, _f64p, _f64p,
                                            C.c_int]
        L.tto_sinograms_vjp_f64.restype = C.c_int
        L.tto_transform.argtypes = [_f32p, C.c_int, C.c_int, C.c_float, _i32p,
                                    C.c_int, _i32p, _i32p, _i32p, C.c_int, _f32p,
                                    C.c_int]
        for name in ("tto_sinograms", "tto_circus", "tto_sinograms_f64",
                     "tto_circus_f64", "tto_transform", "tto_pad_geometry"):
            getattr(L, name).restype = C.c_int
        _lib = L
    return _lib


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _f32(x):
    return np.ascontiguousarray(x, dtype=np.float32)


# ---------------------------------------------------------------- geometry

def pad_geometry(rows: int, cols: int):
    n, dx, dy = C.c_int(), C.c_int(), C.c_int()
    if lib().tto_pad_geometry(rows, cols, C.byref(n), C.byref(dx), C.byref(dy)):
        raise ValueError("invalid image shape")
    return n.value, dx.value, dy.value


def pad(img_colmajor: np.ndarray, rows: int, cols: int) -> np.ndarray:
    """Pad a column-major rows x cols image (flat) to the reference's n x n."""
    n, _, _ = pad_geometry(rows, cols)
    out = np.empty(n * n, np.float32)
    lib().tto_pad(_f32(img_colmajor).ravel(), rows, cols, out)
    return out


def resize(img, in_rows, in_cols, rows, cols):
    out = np.empty(rows * cols, np.float32)
    lib().tto_resize(_f32(img).ravel(), in_rows, in_cols, out, rows, cols)
    return out


def rotation_trig(a_step: int, step_deg: float):
    L = lib()
    ang = L.tto_deg2rad(L.tto_angle_deg(a_step, step_deg))
    c, s = C.c_float(), C.c_float()
    L.tto_rotation_trig(ang, C.byref(c), C.byref(s))
    return np.float32(c.value), np.float32(s.value)


def rotate(img_colmajor: np.ndarray, n: int, c, s) -> np.ndarray:
    out = np.empty(n * n, np.float32)
    lib().tto_rotate(_f32(img_colmajor).ravel(), n, float(c), float(s), out)
    return out


# ------------------------------------------------------- column functionals

def eigen_sum(x):
    x = _f32(x)
    return np.float32(lib().tto_eigen_sum(x, x.size))


def weighted_median(x) -> int:
    x = _f32(x)
    return lib().tto_weighted_median(x, x.size)


def stable_argsort(key):
    key = _f32(key)
    idx = np.empty(key.size, np.int32)
    tmp = np.empty(max(key.size, 1), np.int32)
    lib().tto_stable_argsort(key, key.size, idx, tmp)
    return idx


def t345_table(which: int, n: int):
    re = np.empty(n, np.float32)
    im = np.empty(n, np.float32)
    lib().tto_t345_table(which, n, re, im)
    return re, im


def tfunctional(code: int, x):
    """One T-functional on one column: (value, dict of indices)."""
    L = lib()
    x = _f32(x)
    n = x.size
    a, b, c = C.c_int(-1), C.c_int(-1), C.c_int(-1)
    if code == 0:
        return np.float32(L.tto_T0(x, n)), {}
    if code in (1, 2):
        v = (L.tto_T1 if code == 1 else L.tto_T2)(x, n, C.byref(a))
        return np.float32(v), {"m": a.value}
    if code in (3, 4, 5):
        re, im = t345_table(code, n)
        v = L.tto_T345(x, n, re, im, C.byref(a))
        return np.float32(v), {"sm": a.value}
    if code == 6:
        v = L.tto_T6(x, n, C.byref(a), C.byref(b), C.byref(c))
        return np.float32(v), {"sm": a.value, "rank": b.value, "src": c.value}
    if code == 7:
        v = L.tto_T7(x, n, C.byref(a), C.byref(b), C.byref(c))
        return np.float32(v), {"m": a.value, "rank": b.value, "src": c.value}
    raise ValueError(code)


def pfunctional(code: int, x, order: int = 0, center: int = 0):
    L = lib()
    x = _f32(x)
    n = x.size
    if code == 1:
        return np.float32(L.tto_P1(x, n)), {}
    if code == 2:
        a, b = C.c_int(-1), C.c_int(-1)
        v = L.tto_P2(x, n, C.byref(a), C.byref(b))
        return np.float32(v), {"rank": a.value, "src": b.value}
    if code == 3:
        return np.float32(L.tto_P3(x, n)), {}
    if code == 0:
        return np.float32(L.tto_Phermite(x, n, order, center)), {}
    raise ValueError(code)


def hermite_function(order: int, z: float) -> np.float32:
    return np.float32(lib().tto_hermite_function(order, z))


def dft_f64(x):
    x = np.ascontiguousarray(x, np.float64)
    re = np.empty(x.size)
    im = np.empty(x.size)
    lib().tto_dft_f64(x, x.size, re, im)
    return re + 1j * im


# ------------------------------------------------------------------ drivers

def sinograms(img, n, n_angles, step_deg, tcodes, nthreads=1):
    """FP32-faithful sinograms: returns (sino[nt, A, n], aux dict of [A, n])."""
    tc = np.ascontiguousarray(tcodes, np.int32)
    nt = tc.size
    sino = np.empty(nt * n * n_angles, np.float32)
    aux = {k: np.full(n_angles * n, -1, np.int32)
           for k in ("med_x", "med_s", "t6_rank", "t6_src", "t7_rank", "t7_src")}
    rc = lib().tto_sinograms(_f32(img).ravel(), n, n_angles, step_deg, tc, nt, sino,
                             _ptr(aux["med_x"]), _ptr(aux["med_s"]),
                             _ptr(aux["t6_rank"]), _ptr(aux["t6_src"]),
                             _ptr(aux["t7_rank"]), _ptr(aux["t7_src"]), nthreads)
    if rc:
        raise ValueError("tto_sinograms rejected its arguments")
    return (sino.reshape(nt, n_angles, n),
            {k: v.reshape(n_angles, n) for k, v in aux.items()})


def sinograms_f64(img, n, n_angles, step_deg, tcodes, aux, nthreads=1, geometry=0):
    """FP64 sinograms with the FP32 run's indices.  geometry=1 keeps the
    reference's float sample positions (values in double)."""
    tc = np.ascontiguousarray(tcodes, np.int32)
    out = np.empty(tc.size * n * n_angles, np.float64)
    keep = {k: np.ascontiguousarray(aux[k], np.int32).ravel()
            for k in ("med_x", "med_s", "t6_src", "t7_src")}
    rc = lib().tto_sinograms_f64(_f32(img).ravel(), n, n_angles, step_deg, tc,
                                 tc.size, _ptr(keep["med_x"]), _ptr(keep["med_s"]),
                                 _ptr(keep["t6_src"]), _ptr(keep["t7_src"]),
                                 int(geometry), out, nthreads)
    if rc:
        raise ValueError("tto_sinograms_f64 rejected its arguments")
    return out.reshape(tc.size, n_angles, n)


def _pargs(pfuncs):
    codes = np.array([p[0] for p in pfuncs], np.int32)
    orders = np.array([p[1] if len(p) > 1 else 0 for p in pfuncs], np.int32)
    centers = np.array([p[2] if len(p) > 2 else 0 for p in pfuncs], np.int32)
    return codes, orders, centers


def circus(sino_one, n, n_cols, pfuncs, nthreads=1):
    """pfuncs: list of (code[, order, center]).  sino_one is [n_cols, n]."""
    codes, orders, centers = _pargs(pfuncs)
    np_ = codes.size
    out = np.empty(max(np_, 1) * n_cols, np.float32)
    p2r = np.full(max(np_, 1) * n_cols, -1, np.int32)
    p2s = np.full(max(np_, 1) * n_cols, -1, np.int32)
    rc = lib().tto_circus(_f32(sino_one).ravel(), n, n_cols, codes, orders, centers,
                          np_, out, _ptr(p2r), _ptr(p2s), nthreads)
    if rc:
        raise ValueError("tto_circus rejected its arguments")
    return (out[:np_ * n_cols].reshape(np_, n_cols),
            {"p2_rank": p2r[:np_ * n_cols].reshape(np_, n_cols),
             "p2_src": p2s[:np_ * n_cols].reshape(np_, n_cols)})


def circus_f64(sino_one, n, n_cols, pfuncs, p2_src, nthreads=1, geometry=0):
    codes, orders, centers = _pargs(pfuncs)
    out = np.empty(max(codes.size, 1) * n_cols, np.float64)
    src = np.ascontiguousarray(p2_src, np.int32).ravel()
    rc = lib().tto_circus_f64(np.ascontiguousarray(sino_one, np.float64).ravel(), n,
                              n_cols, codes, orders, centers, codes.size, _ptr(src),
                              int(geometry), out, nthreads)
    if rc:
        raise ValueError("tto_circus_f64 rejected its arguments")
    return out[:codes.size * n_cols].reshape(codes.size, n_cols)


def zscore(x):
    x = _f32(x)
    out = np.empty_like(x)
    lib().tto_zscore(x, x.size, out)
    return out


def transform(img, n, n_angles, step_deg, tcodes, pfuncs, nthreads=1):
    """Reference getTransform without I/O: signatures [nt*np, A]."""
    tc = np.ascontiguousarray(tcodes, np.int32)
    codes, orders, centers = _pargs(pfuncs)
    sig = np.empty(max(tc.size * codes.size, 1) * n_angles, np.float32)
    rc = lib().tto_transform(_f32(img).ravel(), n, n_angles, step_deg, tc, tc.size,
                             codes, orders, centers, codes.size, sig, nthreads)
    if rc:
        raise ValueError("tto_transform rejected its arguments")
    return sig[:tc.size * codes.size * n_angles].reshape(-1, n_angles)


# ---------------------------------------------- nearest orthonormal sinogram

def nearest_orthonormal_sinogram(sino_cols: np.ndarray):
    """c++/src/circus.cpp:68-102 in FP64.

    ``sino_cols`` is [A, n] (one row per sinogram column).  Column medians
    (FP32-faithful weighted median) give per-column offsets to
    floor((n-1)/2); the columns are aligned into (n + |max| + |min|) rows and
    the polar factor U[:, :A] V^T is taken with numpy's FP64 SVD.  Returns
    (nos[A, n'], center).
    """
    A, n = sino_cols.shape
    center = int(np.floor((n - 1) / 2.0))
    offs = np.array([weighted_median(sino_cols[a]) - center for a in range(A)])
    mn, mx = int(offs.min()), int(offs.max())
    padding = abs(mx) + abs(mn)
    new_center = center + mx
    aligned = np.zeros((n + padding, A), np.float64)
    for a in range(A):
        aligned[mx - offs[a]: mx - offs[a] + n, a] = sino_cols[a]
    u, _, vt = np.linalg.svd(aligned, full_matrices=False)
    nos = u @ vt
    return np.ascontiguousarray(nos.T), new_center



# ------------------------------------------------------------------ adjoint

def forward_f64_fixed(img, n, n_angles, step_deg, tcodes, pfuncs, aux, p2_src,
                      nthreads=1):
    """Circus functions [nt, np, A] in FP64, reference geometry, with the
    FP32 run's medians (aux) and P2 sources held fixed: the function whose
    derivative backward_f64 returns."""
    s = sinograms_f64(img, n, n_angles, step_deg, tcodes, aux, nthreads, geometry=1)
    out = np.empty((len(tcodes), len(pfuncs), n_angles))
    for t in range(len(tcodes)):
        out[t] = circus_f64(s[t], n, n_angles, pfuncs, p2_src[t], nthreads, geometry=1)
    return out


def backward_f64(img, n, n_angles, step_deg, tcodes, pfuncs, aux, p2_src, dcircus,
                 nthreads=1):
    """J^T dcircus (FP64, reference geometry, medians/P2 sources fixed):
    gradient of sum(dcircus * circus) with respect to the padded image,
    returned as a flat column-major n*n array."""
    tc = np.ascontiguousarray(tcodes, np.int32)
    codes, orders, centers = _pargs(pfuncs)
    dimg = np.empty(n * n, np.float64)
    mx = np.ascontiguousarray(aux["med_x"], np.int32).ravel()
    ms = np.ascontiguousarray(aux["med_s"], np.int32).ravel()
    s6, s7 = _t67_src(aux)
    src = np.ascontiguousarray(p2_src, np.int32).ravel()
    dc = np.ascontiguousarray(dcircus, np.float64).ravel()
    rc = lib().tto_backward_f64(_f32(img).ravel(), n, n_angles, step_deg, tc, tc.size,
                                codes, orders, centers, codes.size, _ptr(mx), _ptr(ms),
                                _ptr(s6), _ptr(s7), _ptr(src), dc, dimg, nthreads)
    if rc:
        raise ValueError("tto_backward_f64 rejected its arguments")
    return dimg


def _t67_src(aux):
    """T6 (relative to sm) and T7 (absolute row) sources of an aux dict, or
    None where absent."""
    out = []
    for k in ("t6_src", "t7_src"):
        v = aux.get(k)
        out.append(None if v is None else np.ascontiguousarray(v, np.int32).ravel())
    return out


def sinograms_vjp_f64(img, n, n_angles, step_deg, tcodes, aux, dsino, nthreads=1):
    """J_sinograms^T dsino ([nt, A, n]) in FP64 with fixed medians and T6/T7
    sources."""
    tc = np.ascontiguousarray(tcodes, np.int32)
    dimg = np.empty(n * n, np.float64)
    mx = np.ascontiguousarray(aux["med_x"], np.int32).ravel()
    ms = np.ascontiguousarray(aux["med_s"], np.int32).ravel()
    s6, s7 = _t67_src(aux)
    rc = lib().tto_sinograms_vjp_f64(_f32(img).ravel(), n, n_angles, step_deg, tc,
                                     tc.size, _ptr(mx), _ptr(ms), _ptr(s6), _ptr(s7),
                                     np.ascontiguousarray(dsino, np.float64).ravel(),
                                     dimg, nthreads)
    if rc:
        raise ValueError("tto_sinograms_vjp_f64 rejected its arguments")
    return dimg
