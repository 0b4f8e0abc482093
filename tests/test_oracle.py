"""CPU tests that pin the oracle (oracle/tt_oracle.c) before it is trusted.

Each check is chosen so that a plausible slip in the restatement (a dropped
term, a swapped operand, a wrong sign or index) fails: closed forms, library
routines (glibc qsort_r, numpy FFT), brute force on tiny inputs, and
hand-derived cases for Eigen's packet summation order.
"""
import ctypes as C
import math

import numpy as np
import pytest


# ----------------------------------------------------------------- geometry

@pytest.mark.parametrize("side,n,df", [(1, 7, 3), (64, 91, 14), (255, 361, 53),
                                       (256, 363, 54), (512, 725, 107),
                                       (1024, 1449, 213)])
def test_pad_sizes(oracle, side, n, df):
    # c++/src/auxiliary.cpp:257-275: n = 2*ceil(hypot(c-2-ox, r-2-oy)) + 3
    assert oracle.pad_geometry(side, side) == (n, df, df)
    ox = math.floor((side + 1) / 2) - 1
    assert n == 2 * (math.ceil(math.hypot(side - 2 - ox, side - 2 - ox)) + 1) + 1


def test_pad_nonsquare_places_image(oracle):
    rows, cols = 5, 9
    img = np.arange(rows * cols, dtype=np.float32) + 1  # column-major
    n, dx, dy = oracle.pad_geometry(rows, cols)
    out = oracle.pad(img, rows, cols).reshape(n, n)  # [col, row]
    assert out.sum() == img.sum()
    assert out[dx, dy] == img[0] and out[dx + cols - 1, dy + rows - 1] == img[-1]
    # the image is centred: both margins differ by at most one pixel
    assert abs(dx - (n - cols - dx)) <= 1 and abs(dy - (n - rows - dy)) <= 1


# ------------------------------------------------------------------- resize
# c++/src/auxiliary.cpp:198-221: out(row, col) = interpolate(in, p) with
# p = ((col, row) + 0.5) * diag(in_rows / rows, in_cols / cols) - 0.5 on the
# interior (borders stay 0).  p.x (the column coordinate) is scaled by the
# row ratio and p.y by the column ratio, as the reference's diagonal
# transform does.

def _resize_grid(in_rows, in_cols, rows, cols):
    """Exact (double) sample coordinates of resize's interior, and the mask
    of the samples whose four taps lie inside the input."""
    col = np.arange(cols)[None, :].astype(np.float64)
    row = np.arange(rows)[:, None].astype(np.float64)
    x = (col + 0.5) * (in_rows / rows) - 0.5 + 0 * row
    y = (row + 0.5) * (in_cols / cols) - 0.5 + 0 * col
    inner = np.zeros((rows, cols), bool)
    inner[1:rows - 1, 1:cols - 1] = True
    inside = (np.floor(x) >= 0) & (np.floor(x) + 1 <= in_cols - 1) & \
             (np.floor(y) >= 0) & (np.floor(y) + 1 <= in_rows - 1)
    return x, y, inner, inner & inside


@pytest.mark.parametrize("side", [7, 40, 255])
def test_resize_same_size_is_identity_inside(oracle, side):
    # scale 1: p = (col, row) exactly, fractions 0, so each interior sample
    # is its own pixel bit for bit; the one-pixel border is 0
    img = np.random.default_rng(side).random(side * side, dtype=np.float32)
    out = oracle.resize(img, side, side, side, side).reshape(side, side)  # [col, row]
    ref = img.reshape(side, side)
    assert np.array_equal(out[1:-1, 1:-1], ref[1:-1, 1:-1])
    assert not out[0, :].any() and not out[-1, :].any()
    assert not out[:, 0].any() and not out[:, -1].any()


@pytest.mark.parametrize("src,dst", [((64, 64), (32, 32)), ((32, 32), (64, 64)),
                                     ((64, 64), (32, 64)), ((40, 40), (80, 20))])
def test_resize_reproduces_dyadic_ramps(oracle, src, dst):
    # f(row, col) = 3 row + 5 col - 7: bilinear interpolation reproduces an
    # affine function, and with the dyadic scales 2 and 1/2 every coordinate,
    # fraction and product is exact in float, so the float result equals
    # 3 y + 5 x - 7 exactly wherever the four taps are inside the input.  A
    # swapped axis or scale, or a dropped 0.5 shift, moves x or y.
    in_rows, in_cols = src
    rows, cols = dst
    r = np.arange(in_rows)[:, None]
    c = np.arange(in_cols)[None, :]
    f = (3.0 * r + 5.0 * c - 7.0).astype(np.float32)          # [row, col]
    img = np.ascontiguousarray(f.T).ravel()                   # column-major
    out = oracle.resize(img, in_rows, in_cols, rows, cols).reshape(cols, rows).T
    x, y, inner, ok = _resize_grid(in_rows, in_cols, rows, cols)
    assert ok.sum() >= 0.2 * inner.sum()
    want = 3.0 * y + 5.0 * x - 7.0
    assert np.array_equal(out[ok].astype(np.float64), want[ok])
    assert not out[~inner].any()


def test_resize_constant_and_brute_force(oracle):
    # a constant image gives a constant interior (to float rounding of the
    # tap weights); random images match a double evaluation of the bilinear
    # taps at the exact coordinates (float coordinates differ by < 1e-5)
    in_rows = in_cols = 60
    rows = cols = 37
    const = np.full(in_rows * in_cols, 2.5, np.float32)
    out = oracle.resize(const, in_rows, in_cols, rows, cols).reshape(cols, rows).T
    x, y, inner, ok = _resize_grid(in_rows, in_cols, rows, cols)
    assert np.abs(out[ok] - 2.5).max() <= 4e-7 * 2.5
    img = np.random.default_rng(4).random((in_cols, in_rows), dtype=np.float32)
    out = oracle.resize(img.ravel(), in_rows, in_cols, rows, cols).reshape(cols, rows).T
    f = img.T.astype(np.float64)                               # [row, col]
    xi, yi = np.floor(x).astype(int), np.floor(y).astype(int)
    fx, fy = x - xi, y - yi
    xi, yi = np.clip(xi, 0, in_cols - 2), np.clip(yi, 0, in_rows - 2)
    want = (f[yi, xi] * (1 - fx) * (1 - fy) + f[yi, xi + 1] * fx * (1 - fy)
            + f[yi + 1, xi] * (1 - fx) * fy + f[yi + 1, xi + 1] * fx * fy)
    assert np.abs(out[ok] - want[ok]).max() <= 2e-5


# ------------------------------------------------------------------- rotate

def _rot(oracle, img, n, a_step, step=1.0):
    c, s = oracle.rotation_trig(a_step, step)
    return oracle.rotate(img, n, c, s).reshape(n, n), c, s


def test_rotation_trig_is_cosf_sinf_of_negated_float_angle(oracle):
    c, s = oracle.rotation_trig(30, 1.0)
    ang = np.float32(30.0 * math.pi / 180.0)
    assert c == np.float32(math.cos(-float(ang))) or abs(c - math.cos(-float(ang))) < 1e-7
    assert s < 0  # sin(-30 deg)
    c0, s0 = oracle.rotation_trig(0, 1.0)
    assert c0 == 1.0 and s0 == 0.0


def test_rotate_angle0_is_identity_but_last_row_and_column(oracle):
    n = 11
    img = np.random.default_rng(0).random(n * n, dtype=np.float32)
    out, _, _ = _rot(oracle, img, n, 0)
    ref = img.reshape(n, n).copy()  # [col, row]
    ref[n - 1, :] = 0  # last column (x < n-1 strict)
    ref[:, n - 1] = 0  # last row
    assert np.array_equal(out, ref)


def test_rotate_constant_and_centre(oracle):
    n = 15
    img = np.full(n * n, 0.75, np.float32)
    for a in (17, 45, 90, 200):
        out, _, _ = _rot(oracle, img, n, a)
        o = (n - 1) // 2
        assert out[o, o] == 0.75  # the origin maps to itself
        vals = out[out != 0]
        assert np.allclose(vals, 0.75, rtol=0, atol=1e-6)


def _rotate_f64(img_cm, n, c, s):
    """Independent FP64 bilinear rotation (vectorised numpy).  Also returns
    a mask of pixels whose source coordinate is not within 1e-4 of the
    strict bounds (where float and double may legitimately disagree)."""
    o = (n - 1) / 2.0
    col, row = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    px, py = col - o, row - o
    x = px * float(c) + py * float(s) + o
    y = -px * float(s) + py * float(c) + o
    inb = (x >= 0) & (x < n - 1) & (y >= 0) & (y < n - 1)
    X = np.where(inb, np.floor(x), 0).astype(int)
    Y = np.where(inb, np.floor(y), 0).astype(int)
    fx, fy = x - X, y - Y
    src = img_cm.reshape(n, n).astype(np.float64)  # [col, row]
    v = (src[X, Y] * (1 - fx) * (1 - fy) + src[X + 1, Y] * fx * (1 - fy)
         + src[X, Y + 1] * (1 - fx) * fy + src[X + 1, Y + 1] * fx * fy)
    eps = 1e-4
    safe = ((np.abs(x) > eps) & (np.abs(x - (n - 1)) > eps)
            & (np.abs(y) > eps) & (np.abs(y - (n - 1)) > eps))
    return np.where(inb, v, 0.0), safe


@pytest.mark.parametrize("n", [5, 9, 21])
def test_rotate_matches_fp64_bruteforce(oracle, n):
    rng = np.random.default_rng(n)
    img = rng.random(n * n, dtype=np.float32)
    for a in (0, 1, 29, 45, 90, 137, 180, 271, 359):
        out, c, s = _rot(oracle, img, n, a)
        ref, safe = _rotate_f64(img, n, c, s)
        assert safe.mean() > 0.3
        assert np.abs(out - ref)[safe].max() < 1e-5, a


def test_rotate_reproduces_affine_ramp(oracle):
    # Bilinear interpolation reproduces affine functions: rotated(row, col)
    # = f(x, y) with f(row, col) = 0.25*row + 0.5*col + 1 (dyadic, exact).
    n = 33
    col, row = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    img = (0.25 * row + 0.5 * col + 1).astype(np.float32).ravel()
    for a in (10, 33, 60, 150):
        out, c, s = _rot(oracle, img, n, a)
        o = (n - 1) / 2.0
        x = (col - o) * float(c) + (row - o) * float(s) + o
        y = -(col - o) * float(s) + (row - o) * float(c) + o
        inb = (x >= 0) & (x < n - 1) & (y >= 0) & (y < n - 1)
        expect = 0.25 * y + 0.5 * x + 1
        assert np.allclose(out[inb], expect[inb], rtol=0, atol=2e-5)
        assert np.all(out[~inb] == 0)


def test_rotate_90_is_a_permutation_up_to_cos_rounding(oracle):
    # theta = 90: x ~ n-1-row, y ~ col, i.e. rotated(row, col) ~ img(col, n-1-row)
    n = 13
    img = np.random.default_rng(1).random(n * n, dtype=np.float32)
    out, c, s = _rot(oracle, img, n, 90)
    assert s == -1.0 and abs(c) < 1e-7
    src = img.reshape(n, n)  # [col, row]
    for col in range(1, n - 2):
        for row in range(2, n - 1):
            assert abs(out[col, row] - src[n - 1 - row, col]) < 1e-5


# --------------------------------------------- Eigen sum and weighted median

def test_eigen_sum_packet_order_cases(oracle):
    # Hand-derived from Eigen's packet reduction (acc0/acc1, predux
    # (a0+a2)+(a1+a3)); sequential order gives different results.
    x = np.array([1e8, 1, -1e8, 1, 0, 0, 0, 0], np.float32)
    assert oracle.eigen_sum(x) == 2.0
    assert np.float32(oracle.lib().tto_seq_sum(x, 8)) == 1.0
    # n = 5: predux of x[0..3] then + x[4]
    x = np.array([1e8, 1, -1e8, 1, 3], np.float32)
    assert oracle.eigen_sum(x) == 5.0
    # n = 12: acc0 = x[0..3] + x[4..7], then + x[8..11] (tail packet)
    x = np.zeros(12, np.float32)
    x[0], x[4], x[8], x[2] = 1e8, 1, -1e8, 1
    # acc0[0] = (1e8 + 1) + -1e8 = 0 (the 1 is absorbed), acc0[2] = 1
    assert oracle.eigen_sum(x) == 1.0
    # n = 19: packets over 16, then 3 sequential tail adds
    x = np.zeros(19, np.float32)
    x[16], x[17], x[18] = 1e8, 1, -1e8
    assert oracle.eigen_sum(x) == 0.0
    # n = 3: purely sequential
    x = np.array([1e8, 1, -1e8], np.float32)
    assert oracle.eigen_sum(x) == 0.0
    # integers: every order is exact
    x = np.arange(1, 1000, dtype=np.float32)
    assert oracle.eigen_sum(x) == 999 * 1000 / 2


def test_weighted_median_closed_forms(oracle):
    for n in (1, 2, 7, 8, 9, 363):
        assert oracle.weighted_median(np.zeros(n, np.float32)) == 0
        assert oracle.weighted_median(np.ones(n, np.float32)) == math.ceil(n / 2) - 1
    x = np.zeros(50, np.float32)
    x[31] = 2.5
    assert oracle.weighted_median(x) == 31
    # no crossing: Eigen sum 3 exceeds twice every sequential prefix -> n-1
    x = np.array([-1e8, 1, 1, 1, 1e8, 0, 0, 0], np.float32)
    assert oracle.weighted_median(x) == 7


def test_stable_argsort_matches_glibc_qsort_r(oracle):
    libc = C.CDLL("libc.so.6")
    CMP = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_void_p, C.c_void_p)

    def cmp(a, b, arg):
        keys = C.cast(arg, C.POINTER(C.c_float))
        ia = C.cast(a, C.POINTER(C.c_int))[0]
        ib = C.cast(b, C.POINTER(C.c_int))[0]
        ka, kb = keys[ia], keys[ib]
        return -1 if ka < kb else (1 if ka > kb else 0)

    cb = CMP(cmp)
    libc.qsort_r.argtypes = [C.c_void_p, C.c_size_t, C.c_size_t, CMP, C.c_void_p]
    rng = np.random.default_rng(7)
    for n in (1, 2, 5, 64, 300, 1000):
        key = rng.integers(0, 4, n).astype(np.float32)  # tie-heavy
        key[rng.random(n) < 0.3] = 0.0
        idx = np.arange(n, dtype=np.int32)
        libc.qsort_r(idx.ctypes.data, n, 4, cb, key.ctypes.data)
        assert np.array_equal(oracle.stable_argsort(key), idx)


# ------------------------------------------------------------ T-functionals

def test_T0_T1_T2_box_closed_form(oracle):
    n, a, b = 40, 7, 22
    x = np.zeros(n, np.float32)
    x[a:b + 1] = 1
    L = b - a + 1
    m = a + math.ceil(L / 2) - 1
    k = b - m
    assert oracle.tfunctional(0, x)[0] == L
    v1, i1 = oracle.tfunctional(1, x)
    v2, i2 = oracle.tfunctional(2, x)
    assert i1["m"] == m and i2["m"] == m
    assert v1 == k * (k + 1) / 2
    assert v2 == k * (k + 1) * (2 * k + 1) / 6


def test_T345_two_spikes_closed_form(oracle):
    n = 60
    for r in (1, 5, 17):
        x = np.zeros(n, np.float32)
        x[10], x[10 + r] = 9.0, 4.0  # sqrt 3 >= sqrt 2 -> sqrt-median at 10
        expect = {3: r * 4.0, 4: 4.0, 5: math.sqrt(r) * 4.0}
        for code, e in expect.items():
            v, idx = oracle.tfunctional(code, x)
            assert idx["sm"] == 10
            assert abs(v - e) <= 2e-6 * e


def test_t345_tables(oracle):
    re, im = oracle.t345_table(3, 50)
    r = np.arange(1, 50)
    assert np.allclose(re[1:], r * np.cos(5 * np.log(r)), rtol=1e-6, atol=1e-6)
    assert np.allclose(im[1:], r * np.sin(5 * np.log(r)), rtol=1e-6, atol=1e-6)
    re, im = oracle.t345_table(5, 50)
    assert np.allclose(np.hypot(re[1:], im[1:]), np.sqrt(r), rtol=1e-6)


def _brute_T6(oracle, x):
    xs = np.sqrt(x.astype(np.float32))
    sm = oracle.weighted_median(xs)
    L = x.size - sm
    w = [np.float32(np.float32(r) * x[sm + r]) for r in range(L)]
    order = sorted(range(L), key=lambda r: w[r])  # Python's sort is stable
    ds = np.sqrt(np.array([x[sm + r] for r in order], np.float32))
    k = oracle.weighted_median(ds)
    return w[order[k]], sm, k, order[k]


def _brute_T7(oracle, x):
    m = oracle.weighted_median(x)
    tail = sorted(x[m:].tolist())
    k = oracle.weighted_median(np.sqrt(np.array(tail, np.float32)))
    return np.float32(tail[k]), m, k


def test_T6_T7_bruteforce_short_columns(oracle):
    rng = np.random.default_rng(3)
    cases = [np.zeros(5, np.float32), np.ones(6, np.float32)]
    for n in range(1, 9):
        for _ in range(40):
            x = rng.integers(0, 3, n).astype(np.float32)  # ties and zeros
            x[rng.random(n) < 0.3] *= 0.5
            cases.append(x)
    for x in cases:
        v6, i6 = oracle.tfunctional(6, x)
        e6, sm, k, src = _brute_T6(oracle, x)
        assert (v6, i6["sm"], i6["rank"], i6["src"]) == (e6, sm, k, src)
        v7, i7 = oracle.tfunctional(7, x)
        e7, m, k7 = _brute_T7(oracle, x)
        assert (v7, i7["m"], i7["rank"]) == (e7, m, k7)


# ------------------------------------------------------------ P-functionals

def test_P1_monotone_and_zigzag(oracle):
    x = np.linspace(0, 3, 17, dtype=np.float32)
    assert abs(oracle.pfunctional(1, x)[0] - 3.0) < 1e-6
    z = np.array([0, 1, 0, 1, 0], np.float32)
    assert oracle.pfunctional(1, z)[0] == 4.0


def test_P2_bruteforce(oracle):
    rng = np.random.default_rng(5)
    for n in (1, 3, 8, 33, 363):
        x = rng.random(n, dtype=np.float32)
        v, idx = oracle.pfunctional(2, x)
        srt = np.sort(x)
        m = oracle.weighted_median(srt)
        assert v == srt[m] and idx["rank"] == m and x[idx["src"]] == v


def test_P3_closed_forms(oracle):
    # The reference stores linspace as float, so the trapz spacings carry a
    # relative rounding error up to ~1e-5 at n = 363: tolerance 2e-5.
    for n in (7, 91, 363):
        c = 0.8
        v, _ = oracle.pfunctional(3, np.full(n, c, np.float32))
        assert abs(v - c ** 4 / (n - 1)) <= 2e-5 * c ** 4 / (n - 1)
        x = np.zeros(n, np.float32)
        x[n // 3] = 5.0
        v, _ = oracle.pfunctional(3, x)
        e = 2 * (5.0 / n) ** 4
        assert abs(v - e) <= 2e-5 * e


@pytest.mark.parametrize("n", [7, 64, 65, 91, 363, 1449])
def test_dft_f64_matches_numpy(oracle, n):
    x = np.random.default_rng(n).random(n)
    assert np.allclose(oracle.dft_f64(x), np.fft.fft(x), rtol=1e-12, atol=1e-10)


def test_hermite_basis(oracle):
    # h_n(0) closed forms and orthonormality on a fine symmetric grid
    assert abs(oracle.hermite_function(0, 0.0) - math.pi ** -0.25) < 1e-7
    assert abs(oracle.hermite_function(2, 0.0) + 2 / math.sqrt(8 * math.sqrt(math.pi))) < 1e-7
    assert oracle.hermite_function(3, 0.0) == 0.0
    z = np.linspace(-10, 10, 4001)
    dz = z[1] - z[0]
    H = np.array([[oracle.hermite_function(k, float(t)) for t in z] for k in range(6)],
                 np.float64)
    G = H @ H.T * dz
    assert np.allclose(G, np.eye(6), atol=1e-4)


def test_hermite_pfunctional_grid(oracle):
    # symmetric centre: z_p = p*10/c - 10 below, (p-c)*10/(n-1-c) above
    n, c = 21, 10
    x = np.zeros(n, np.float32)
    x[c] = 1.0
    v, _ = oracle.pfunctional(0, x, order=0, center=c)
    assert abs(v - math.pi ** -0.25) < 1e-6
    x[:] = 0
    x[0] = 1.0
    v, _ = oracle.pfunctional(0, x, order=2, center=c)
    assert abs(v - oracle.hermite_function(2, -10.0)) == 0


def test_zscore(oracle):
    x = np.random.default_rng(2).random(360, dtype=np.float32)
    z = oracle.zscore(x)
    assert abs(z.mean()) < 1e-6 and abs(z.std(ddof=1) - 1) < 1e-6


# ------------------------------------------------------------------ drivers

def test_sinogram_driver_matches_column_functionals(oracle):
    side = 9
    n, _, _ = oracle.pad_geometry(side, side)
    img = oracle.pad(np.random.default_rng(4).random(side * side, dtype=np.float32),
                     side, side)
    codes = [0, 1, 2, 3, 4, 5, 6, 7]
    sino, aux = oracle.sinograms(img, n, 12, 30.0, codes, nthreads=2)
    for a in (0, 5, 11):
        c, s = oracle.rotation_trig(a, 30.0)
        rot = oracle.rotate(img, n, c, s).reshape(n, n)
        for col in range(n):
            for t, code in enumerate(codes):
                v, idx = oracle.tfunctional(code, rot[col])
                assert sino[t, a, col] == v
                if code in (1, 2):
                    assert aux["med_x"][a, col] == idx["m"]
                if code in (3, 4, 5):
                    assert aux["med_s"][a, col] == idx["sm"]


def test_fp64_oracle_agrees_with_fp32(oracle):
    side = 24
    n, _, _ = oracle.pad_geometry(side, side)
    img = oracle.pad(np.random.default_rng(8).random(side * side, dtype=np.float32),
                     side, side)
    codes = [0, 1, 2, 3, 4, 5, 6, 7]
    s32, aux = oracle.sinograms(img, n, 36, 10.0, codes, nthreads=4)
    s64 = oracle.sinograms_f64(img, n, 36, 10.0, codes, aux, nthreads=4)
    for t in range(len(codes)):
        scale = np.abs(s64[t]).max()
        assert np.abs(s32[t] - s64[t]).max() <= 1e-5 * scale, codes[t]
    pf = [(1,), (2,), (3,), (0, 2, (n - 1) // 2)]
    for t in range(len(codes)):
        c32, paux = oracle.circus(s32[t], n, 36, pf)
        c64 = oracle.circus_f64(s64[t], n, 36, pf, paux["p2_src"])
        for p in range(len(pf)):
            scale = np.abs(c64[p]).max()
            assert np.abs(c32[p] - c64[p]).max() <= 1e-4 * scale + 1e-30, (t, p)


def test_fp64_reference_geometry(oracle):
    # Reference geometry keeps the float sample positions: on an image that
    # touches the padded border the all-double oracle disagrees with the FP32
    # reference at boundary flips, the reference-geometry one does not.
    rows, cols = 17, 3
    n, _, _ = oracle.pad_geometry(rows, cols)
    img = oracle.pad(np.random.default_rng(3).random(rows * cols, dtype=np.float32),
                     rows, cols)
    s32, aux = oracle.sinograms(img, n, 72, 5.0, [0], nthreads=2)
    g1 = oracle.sinograms_f64(img, n, 72, 5.0, [0], aux, nthreads=2, geometry=1)
    g0 = oracle.sinograms_f64(img, n, 72, 5.0, [0], aux, nthreads=2, geometry=0)
    assert np.abs(s32[0] - g1[0]).max() <= 1e-6 * np.abs(g1[0]).max()
    assert np.abs(s32[0] - g0[0]).max() > 1e-3 * np.abs(g0[0]).max()
    # on a padded square image with a zero margin both agree to float rounding
    side = 20
    n, _, _ = oracle.pad_geometry(side, side)
    img = oracle.pad(np.random.default_rng(4).random(side * side, dtype=np.float32),
                     side, side)
    s32, aux = oracle.sinograms(img, n, 36, 10.0, [0, 1, 3], nthreads=2)
    g1 = oracle.sinograms_f64(img, n, 36, 10.0, [0, 1, 3], aux, nthreads=2, geometry=1)
    g0 = oracle.sinograms_f64(img, n, 36, 10.0, [0, 1, 3], aux, nthreads=2, geometry=0)
    for t in range(3):
        assert np.abs(g1[t] - g0[t]).max() <= 1e-5 * np.abs(g0[t]).max()


# ------------------------------------------------------------------ adjoint

def _fixed_indices(oracle, img, n, A, step, codes, pf):
    s32, aux = oracle.sinograms(img, n, A, step, codes, nthreads=4)
    src = np.stack([oracle.circus(s32[t], n, A, pf)[1]["p2_src"]
                    for t in range(len(codes))])
    return aux, src


def test_backward_f64_matches_central_differences(oracle):
    # The analytic FP64 gradient of sum(dC * C) (medians and P2 ranks held
    # fixed, reference geometry) against FP64 central differences of the same
    # function, on pixels inside the image.
    side, A, step = 10, 24, 15.0
    n, dx, dy = oracle.pad_geometry(side, side)
    rng = np.random.default_rng(21)
    img = oracle.pad(rng.random(side * side, dtype=np.float32) + 0.1, side, side)
    herm = (0, 2, (n - 1) // 2)
    # T0-T5 with every P-functional; T6/T7 (sources fixed too) with the
    # smooth ones: adjacent T6/T7 values can lie within a step of each
    # other, where P1's |difference| has a kink the central difference
    # would straddle
    for codes, pf in (([0, 1, 2, 3, 4, 5], [(1,), (2,), (3,), herm]),
                      ([6, 7], [(2,), (3,), herm])):
        aux, src = _fixed_indices(oracle, img, n, A, step, codes, pf)
        dC = rng.standard_normal((len(codes), len(pf), A))
        grad = oracle.backward_f64(img, n, A, step, codes, pf, aux, src, dC, nthreads=4)
        pix = [(dy + r) + (dx + c) * n for r, c in
               ((0, 0), (3, 4), (9, 9), (5, 1), (7, 8), (2, 6))]
        for i in pix:
            h = 1e-3
            ip, im = img.copy(), img.copy()
            ip[i] += h
            im[i] -= h
            Lp = (dC * oracle.forward_f64_fixed(ip, n, A, step, codes, pf, aux, src)).sum()
            Lm = (dC * oracle.forward_f64_fixed(im, n, A, step, codes, pf, aux, src)).sum()
            fd = (Lp - Lm) / (float(ip[i]) - float(im[i]))
            assert abs(fd - grad[i]) <= 1e-6 * max(1.0, abs(grad[i])), (codes, i, fd, grad[i])
        # pixels of the zero margin receive gradient only through
        # interpolation weights, and the whole gradient is finite
        assert np.isfinite(grad).all()


def test_backward_f64_linear_parts(oracle):
    # T0 with a P2-free linear head: J^T dC for T0 alone and dC = 1 at one
    # angle is the transposed rotation of a constant column sum: every pixel
    # strictly inside the rotated support gets weight 1 (bilinear weights of a
    # constant field sum to one), so the gradient sums to the number of
    # in-bounds rotated samples.
    side, A = 9, 1
    n, _, _ = oracle.pad_geometry(side, side)
    img = oracle.pad(np.ones(side * side, np.float32), side, side)
    s32, aux = oracle.sinograms(img, n, A, 1.0, [0])
    # P1 of a sinogram column is not linear; use Hermite order 0 weights
    pf = [(0, 0, (n - 1) // 2)]
    src = np.zeros((1, 1, A), np.int32)
    dC = np.ones((1, 1, A))
    grad = oracle.backward_f64(img, n, A, 1.0, [0], pf, aux, src, dC)
    z = np.array([oracle.hermite_function(0, 0.0)])  # noqa: F841
    # gradient of sum_c h(z_c) * T0_c: each rotated sample (row, col) carries
    # h(z_col); at angle 0 the rotation is the identity (bar the last
    # row/column), so dimg[row, col] = h(z_col) for row, col < n - 1
    lo = np.float32(10.0 / ((n - 1) // 2))
    G = grad.reshape(n, n)  # [col, row]
    for col in range(n - 1):
        c0 = (n - 1) // 2
        zc = np.float32(col * lo - 10) if col < c0 else np.float32((col - c0) * np.float32(10.0 / (n - 1 - c0)))
        assert np.allclose(G[col, :n - 1], oracle.hermite_function(0, float(zc)), rtol=1e-6)
    assert np.all(G[n - 1, :] == 0) and np.all(G[:, n - 1] == 0)
