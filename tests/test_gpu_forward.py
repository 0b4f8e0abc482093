"""GPU parity of the forward path (tt_sinograms, tt_circus) against the oracle.

Bar:
* weighted-median indices (and P2's rank) bit-exact against the FP32-faithful
  oracle, and T0 (the Eigen-order float sum that also sets the median
  threshold) bit-exact too;
* values against the FP64 oracle in reference geometry (the reference's float
  sample positions, values in double): normwise relative error <= 1e-5 for
  every functional, elementwise <= 1e-5 for T1/T2/P1/P3 (non-negative sums,
  on values above 1e-3 of the maximum) and <= 2e-5 for T0 (a float sum of up
  to 1449 terms in Eigen's 8-lane order);
* against the all-double FP64 oracle (sample positions in double too), for
  images with a zero margin: normwise <= 1e-4 and no worse than the FP32
  reference's own error (its float sample coordinates move T1/T2 by ~1e-5),
  P3 <= 3e-5 (the reference stores its trapz abscissae as float).
The FP32 reference itself misses some of these bars (its sequential float
integrals); the GPU integrals accumulate in double and do not.
"""
import numpy as np
import pytest
import torch

from helpers import assert_same_indices, normwise_err, padded, rel_err, ulp_diff
from tracetransform_amd import core, inputs

pytestmark = pytest.mark.gpu

T05 = [0, 1, 2, 3, 4, 5]
T07 = [0, 1, 2, 3, 4, 5, 6, 7]
NTHREADS = 16


def gpu_sinograms(img_cm, n, A, step, codes):
    img = torch.from_numpy(np.ascontiguousarray(img_cm, np.float32).reshape(n, n)).cuda()
    sino, aux = core.sinograms(img, A, step, codes, with_aux=True)
    torch.cuda.synchronize()
    return sino.cpu().numpy(), {k: v.cpu().numpy() for k, v in aux.items()}


def check_forward(oracle, img, n, A, step, codes, f64=True, margin=True):
    print(f"[forward] n={n} A={A} step={step} codes={codes}", flush=True)
    g, gaux = gpu_sinograms(img, n, A, step, codes)
    o, oaux = oracle.sinograms(img, n, A, step, codes, nthreads=NTHREADS)
    if any(c in (1, 2, 7) for c in codes):
        assert_same_indices("median_x", gaux["median_x"], oaux["med_x"])
    if any(c in (3, 4, 5, 6) for c in codes):
        assert_same_indices("median_sqrt", gaux["median_sqrt"], oaux["med_s"])
    if 6 in codes:
        assert_same_indices("t6_rank", gaux["t6_rank"], oaux["t6_rank"])
        assert_same_indices("t6_src", gaux["t6_src"], oaux["t6_src"])
    if 7 in codes:
        assert_same_indices("t7_rank", gaux["t7_rank"], oaux["t7_rank"])
        assert_same_indices("t7_src", gaux["t7_src"], oaux["t7_src"])
    for t, c in enumerate(codes):
        if c in (0, 6, 7):  # float sum in Eigen order / selected elements
            assert np.array_equal(g[t], o[t]), (c, ulp_diff(g[t], o[t]))
    if not f64:
        return g, o
    d1 = oracle.sinograms_f64(img, n, A, step, codes, oaux, nthreads=NTHREADS,
                              geometry=1)
    d0 = oracle.sinograms_f64(img, n, A, step, codes, oaux, nthreads=NTHREADS,
                              geometry=0) if margin else None
    for t, c in enumerate(codes):
        err = normwise_err(g[t], d1[t])
        ref_err = normwise_err(o[t], d1[t])
        msg = f"  T{c}: normwise vs FP64 = {err:.2e} (FP32 reference {ref_err:.2e})"
        assert err <= 1e-5, (c, err)
        if c in (0, 1, 2, 6, 7):
            rel = rel_err(g[t], d1[t])
            msg += f", elementwise {rel:.2e}"
            assert rel <= (2e-5 if c == 0 else 1e-5), (c, rel)
        if d0 is not None:
            # all-double positions: the reference's float sample coordinates
            # alone move T1/T2 by up to ~1e-5 (the FP32 reference shows the
            # same error); the GPU must be no worse than the reference
            e0 = normwise_err(g[t], d0[t])
            r0 = normwise_err(o[t], d0[t])
            msg += f", vs all-double FP64 {e0:.2e} (reference {r0:.2e})"
            assert e0 <= max(1e-5, 1.05 * r0) and e0 <= 1e-4, (c, e0, r0)
        print(msg, flush=True)
    return g, o


def test_uniform_64(oracle, gpu):
    side = 64
    img, n = padded(oracle, inputs.uniform_image(side, 11), side, side)
    check_forward(oracle, img, n, 360, 1.0, T05)


def test_all_T_uniform_64(oracle, gpu):
    side = 64
    img, n = padded(oracle, inputs.uniform_image(side, 12), side, side)
    check_forward(oracle, img, n, 360, 1.0, T07)


def test_all_T_blob_1024_few_angles(oracle, gpu):
    # the large config's image (1024x1024 blobs, padded to 1449), 8 angles
    side = 1024
    img, n = padded(oracle, inputs.blob_image(side, 2), side, side)
    check_forward(oracle, img, n, 8, 45.0, T07)


@pytest.mark.parametrize("side", [300, 700, 1024])
def test_T6_T7_long_tails(oracle, gpu, side):
    # the mass sits in one corner, so along the diagonal lines the median is
    # near the line's start and the non-zero part of the tail is long: up to
    # 403 / 968 / 1426 samples for sides 300 / 700 / 1024 (n = 425 / 991 /
    # 1449).  Parts longer than 512 take the queued wide pass, longer than
    # 1024 its widest network.
    rng = np.random.default_rng(side)
    img = 0.01 + 1e-3 * rng.random((side, side))
    img[:30, :30] += 5.0
    img_cm = np.ascontiguousarray(img.T, np.float32).ravel()
    n, _, _ = oracle.pad_geometry(side, side)
    pimg = oracle.pad(img_cm, side, side)
    check_forward(oracle, pimg, n, 8, 45.0, [6, 7, 0], f64=False)


@pytest.mark.parametrize("rows,cols", [(1, 1), (3, 8), (30, 30)])
def test_T6_T7_small_and_ties(oracle, gpu, rows, cols):
    rng = np.random.default_rng(rows + 7 * cols)
    # quantised values: many equal keys exercise the stable tie order
    img_cm = (rng.integers(0, 4, rows * cols) / 4.0).astype(np.float32)
    n, _, _ = oracle.pad_geometry(rows, cols)
    img = oracle.pad(img_cm, rows, cols)
    check_forward(oracle, img, n, 24, 15.0, [6, 7, 0], margin=False)
    zero = np.zeros(n * n, np.float32)
    g, _ = check_forward(oracle, zero, n, 4, 90.0, [7, 6], f64=False)
    assert not g.any()


def test_config1_uniform_256(oracle, gpu):
    # 256x256 U[0,1) image padded to 363, 360 angles, T0-T5
    side = 256
    img, n = padded(oracle, inputs.uniform_image(side, 1), side, side)
    assert n == 363
    check_forward(oracle, img, n, 360, 1.0, T05)


def test_blob_1024_few_angles(oracle, gpu):
    side = 1024
    img, n = padded(oracle, inputs.blob_image(side, 2), side, side)
    assert n == 1449
    check_forward(oracle, img, n, 12, 30.0, T05)


@pytest.mark.parametrize("rows,cols", [(1, 1), (2, 2), (5, 9), (17, 3), (40, 41)])
def test_small_and_nonsquare(oracle, gpu, rows, cols):
    rng = np.random.default_rng(rows * 100 + cols)
    img_cm = rng.random(rows * cols, dtype=np.float32)
    n, _, _ = oracle.pad_geometry(rows, cols)
    img = oracle.pad(img_cm, rows, cols)
    # thin images touch the padded border (the reference's pad measures to
    # the second-to-last row), so only reference geometry is comparable
    check_forward(oracle, img, n, 72, 5.0, T05, margin=False)


def test_negative_and_zero_images(oracle, gpu):
    side = 48
    rng = np.random.default_rng(9)
    img_cm = (rng.random(side * side, dtype=np.float32) - 0.3).astype(np.float32)
    img, n = padded(oracle, img_cm, side, side)
    check_forward(oracle, img, n, 90, 4.0, T05, f64=False)
    # negative samples make the sqrt sums NaN: T6/T7 pick the last position
    check_forward(oracle, img, n, 90, 4.0, [6, 7], f64=False)
    zero = np.zeros(n * n, np.float32)
    g, _ = check_forward(oracle, zero, n, 8, 45.0, T05, f64=False)
    assert not g.any()


def test_walk_checkpoints_mixed_signs(oracle, gpu):
    # n = 285: lines of several 64-row blocks, where the line kernel resumes
    # the median walks from stored prefixes.  The left half is non-negative
    # (monotone prefixes: resumed walks), the right half holds negative
    # samples (those waves walk every row), and a zero band gives lines whose
    # medians sit right after a run of zeros.
    side = 200
    rng = np.random.default_rng(31)
    img = rng.random((side, side), dtype=np.float32)       # [col][row]
    img[side // 2:, :] -= 0.25
    img[:, 60:90] = 0.0
    img, n = padded(oracle, np.ascontiguousarray(img).reshape(-1), side, side)
    assert n == 285
    # both walks, the x walk alone (negatives flagged by the smallest
    # sample) and the square-root walk alone (by a NaN sum)
    for codes in (T05, [0, 1], [4, 0]):
        check_forward(oracle, img, n, 45, 4.0, codes, f64=False)


@pytest.mark.parametrize("scale", [1e-12, 1e-30])
def test_tiny_pixel_scales(oracle, gpu, scale):
    # Square roots of samples: 1e-12 keeps every pixel above 2^-45, where the
    # line kernel's short square root is exact for every sample; 1e-30 flags
    # the image (samples can fall below 2^-96) and takes the full sequence.
    side = 150
    img, n = padded(oracle, inputs.uniform_image(side, 13), side, side)
    img = (img * np.float32(scale)).astype(np.float32)
    check_forward(oracle, img, n, 30, 6.0, T05, f64=False)
    check_forward(oracle, img, n, 30, 6.0, [6, 7, 0], f64=False)


def test_fractional_step(oracle, gpu):
    side = 40
    img, n = padded(oracle, inputs.blob_image(side, 6, nblobs=4, sigma=(3, 10)),
                    side, side)
    check_forward(oracle, img, n, 720, 0.5, T05)


def test_subsets_and_duplicates(oracle, gpu):
    side = 32
    img, n = padded(oracle, inputs.uniform_image(side, 5), side, side)
    for codes in ([0], [1], [4], [2, 5], [3, 0, 3], [5, 4, 3, 2, 1, 0], [6], [7],
                  [7, 6, 7]):
        check_forward(oracle, img, n, 36, 10.0, codes, f64=False)


# ------------------------------------------------------------------ circus

def gpu_circus(sino_cols, pf):
    t = torch.from_numpy(np.ascontiguousarray(sino_cols, np.float32)).cuda()
    out, rank = core.circus(t, pf, with_p2_rank=True)
    torch.cuda.synchronize()
    return out.cpu().numpy(), rank.cpu().numpy()


def check_circus(oracle, sino_cols, pf):
    A, n = sino_cols.shape
    print(f"[circus] n={n} A={A} pf={pf}", flush=True)
    g, grank = gpu_circus(sino_cols, pf)
    o, oaux = oracle.circus(sino_cols, n, A, pf, nthreads=NTHREADS)
    sd = sino_cols.astype(np.float64)
    d1 = oracle.circus_f64(sd, n, A, pf, oaux["p2_src"], nthreads=NTHREADS, geometry=1)
    d0 = oracle.circus_f64(sd, n, A, pf, oaux["p2_src"], nthreads=NTHREADS, geometry=0)
    for q, p in enumerate(pf):
        if p[0] == 2:
            assert_same_indices("p2_rank", grank[q], oaux["p2_rank"][q])
            assert np.array_equal(g[q], o[q])
        e1, e0, e32 = (normwise_err(g[q], d1[q]), normwise_err(g[q], d0[q]),
                       normwise_err(o[q], d1[q]))
        print(f"  P{p[0]}: normwise vs FP64 {e1:.2e}, vs all-double FP64 {e0:.2e} "
              f"(FP32 reference {e32:.2e})", flush=True)
        assert e1 <= 1e-5, (p, e1)
        assert e0 <= (3e-5 if p[0] == 3 else 1e-5), (p, e0)
        if p[0] in (1, 3):
            assert rel_err(g[q], d1[q]) <= 1e-5, p


def test_circus_on_oracle_sinograms(oracle, gpu):
    side = 256
    img, n = padded(oracle, inputs.uniform_image(side, 1), side, side)
    sino, _ = oracle.sinograms(img, n, 360, 1.0, T05, nthreads=NTHREADS)
    pf = [(1,), (2,), (3,), (0, 1, (n - 1) // 2), (0, 4, 100)]
    for t in range(len(T05)):
        check_circus(oracle, sino[t], pf)


def test_circus_sizes(oracle, gpu):
    # P2 sorts columns of up to 2048 rows in one wave's registers (1, 2, 16
    # and 32 keys per lane here); 2049 rows take the workgroup sort
    rng = np.random.default_rng(12)
    for n, A in ((7, 3), (91, 180), (725, 40), (1449, 16), (2048, 3), (2049, 2)):
        sino = rng.random((A, n), dtype=np.float32)
        check_circus(oracle, sino, [(1,), (2,), (3,), (0, 3, n // 3)])
    # ties and negative values (P2's order of equal values is free)
    sino = (rng.integers(-4, 9, (6, 1449)) / 4.0).astype(np.float32)
    check_circus(oracle, sino, [(2,), (1,)])


def test_p2_signed_columns(oracle, gpu):
    # P2 walks the sorted column, whose prefix falls while the values are
    # negative: negative and zero column sums, all-negative columns and the
    # sorted column [-3, -1 x 8] (findWeightedMedian returns 0 there), at
    # heights that take 1, 2, 16 and 32 sort keys per lane
    rng = np.random.default_rng(41)
    for n, A in ((9, 4), (91, 64), (725, 16), (2048, 4)):
        sino = np.empty((A, n), np.float32)
        for a in range(A):
            shift = (0.5, 0.9, 1.5, 0.3)[a % 4]
            sino[a] = (rng.random(n) - shift).astype(np.float32)
        sino[0, :] = -1.0
        sino[0, min(3, n - 1)] = -3.0
        sino[1] = np.where(np.arange(n) % 2 == 0, 1.0, -1.0).astype(np.float32)
        check_circus(oracle, sino, [(2,), (1,)])


def test_hermite_tables_follow_height_and_centre(oracle, gpu):
    # the same (order, centre) at two heights, then alternating centres: the
    # weight tables depend on the height through the z grid, and the cached
    # tables must follow both
    rng = np.random.default_rng(14)
    cases = [(91, 30), (121, 30), (91, 30), (121, 40), (121, 30)]
    for n, c in cases:
        sino = rng.random((20, n), dtype=np.float32)
        check_circus(oracle, sino, [(0, 2, c), (0, 5, c)])


def test_circus_batched_sinograms(oracle, gpu):
    rng = np.random.default_rng(13)
    sino = rng.random((3, 50, 91), dtype=np.float32)
    t = torch.from_numpy(sino).cuda()
    out = core.circus(t, [(2,), (1,)]).cpu().numpy()  # no explicit synchronize
    for s in range(3):
        o, _ = oracle.circus(sino[s], 91, 50, [(2,), (1,)])
        assert np.array_equal(out[s, 0], o[0])
        assert normwise_err(out[s, 1], o[1]) < 1e-6


# --------------------------------------------------------- reference surface

def test_reference_surface(oracle, gpu):
    import tracetransform_amd as tt
    side = 50
    img_cm = inputs.uniform_image(side, 21)
    img_pad, n = padded(oracle, img_cm, side, side)
    mat = torch.from_numpy(img_pad.reshape(n, n).T.copy())  # [rows, cols]
    tf = [tt.TFunctionalWrapper.parse(s) for s in ("0", "3")]
    sinos = tt.getSinograms(mat, 2, tf)
    assert len(sinos) == 2 and tuple(sinos[0].shape) == (n, 180)
    o, _ = oracle.sinograms(img_pad, n, 180, 2.0, [0, 3], nthreads=NTHREADS)
    assert np.array_equal(sinos[0].cpu().numpy(), o[0].T)
    pf = [tt.PFunctionalWrapper.parse(s) for s in ("1", "2")]
    circ = tt.getCircusFunctions(sinos[1], pf)
    oc, _ = oracle.circus(np.ascontiguousarray(sinos[1].cpu().numpy().T), n, 180,
                          [(1,), (2,)])
    assert normwise_err(circ[0].cpu().numpy(), oc[0]) < 1e-6
    assert np.array_equal(circ[1].cpu().numpy(), oc[1])
    # Transformer pads on the device
    tr = tt.Transformer(torch.from_numpy(img_cm.reshape(side, side).T.copy()), "x", 2)
    assert np.array_equal(tr.image.cpu().numpy(), img_pad.reshape(n, n).T)
    sig = tr.getTransform(tf, pf)
    assert tuple(sig.shape) == (180, 4)


def test_invalid_arguments(gpu):
    img = torch.zeros((7, 7), device="cuda")
    with pytest.raises(ValueError):
        core.sinograms(img, 4, 90.0, [8])
    with pytest.raises(ValueError):
        core.sinograms(img, 0, 90.0, [0])
    with pytest.raises(ValueError):
        core.circus(torch.zeros((4, 7), device="cuda"), [(0, 2, -1)])
    with pytest.raises(ValueError):
        core.circus(torch.zeros((4, 7), device="cuda"), [(9,)])
