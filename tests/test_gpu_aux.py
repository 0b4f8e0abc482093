"""GPU tests of the steps either side of the path (resize, zscore) and of
the nearest orthonormal sinogram (NOS) used by the Hermite P-functionals.

Bar: resize bit-exact against the oracle's restatement (same float
operations); zscore within 1e-6 normwise (the double variance is summed in
another order); NOS: column offsets, aligned height and Hermite centre exact,
values within 1e-6 of numpy's FP64 polar factor on well-conditioned inputs
(the polar factor of a rank-deficient matrix is not unique, so sinograms of
smooth images are not compared elementwise).
"""
import numpy as np
import pytest
import torch

from helpers import normwise_err, padded
from tracetransform_amd import core, inputs

pytestmark = pytest.mark.gpu


def dev(x):
    return torch.from_numpy(np.ascontiguousarray(x, np.float32)).cuda()


@pytest.mark.parametrize("src,dst", [((256, 256), (255, 255)), ((100, 60), (255, 255)),
                                     ((1024, 1024), (255, 255)), ((40, 40), (97, 31))])
def test_resize_matches_oracle(oracle, gpu, src, dst):
    rows, cols = src
    img = inputs.blob_image(cols, 7, rows=rows).ravel()
    g = core.resize(dev(img.reshape(cols, rows)), rows, cols, dst[0], dst[1])
    o = oracle.resize(img, rows, cols, dst[0], dst[1])
    assert np.array_equal(g.cpu().numpy().ravel(), o)


def test_zscore_matches_oracle(oracle, gpu):
    x = np.random.default_rng(3).random((7, 360), dtype=np.float32) * 5 + 1
    g = core.zscore(dev(x)).cpu().numpy()
    for k in range(7):
        assert normwise_err(g[k], oracle.zscore(x[k])) <= 1e-6


def test_nos_matches_fp64_polar_factor(oracle, gpu):
    rng = np.random.default_rng(5)
    for n, A in ((400, 360), (255, 120), (91, 90)):
        sino = (rng.random((A, n)) + 0.05).astype(np.float32)
        nos, center = core.nearest_orthonormal_sinogram(dev(sino))
        ref, rc = oracle.nearest_orthonormal_sinogram(sino)
        print(f"[nos] n={n} A={A} rows={ref.shape[1]} center={rc}", flush=True)
        assert center == rc and tuple(nos.shape) == ref.shape
        err = normwise_err(nos.cpu().numpy(), ref)
        assert err <= 1e-6, err
        q = nos.double()
        gram = (q @ q.t()).cpu().numpy()       # columns orthonormal
        assert np.abs(gram - np.eye(A)).max() <= 1e-5


def test_nos_of_mixed_sign_sinograms(oracle, gpu):
    # the column medians that align the sinogram walk raw columns of any
    # sign: the T0 sinogram of an image with negative pixels (negative and
    # mixed column sums) must give the reference's offsets, hence its aligned
    # height and Hermite centre
    side = 40
    rng = np.random.default_rng(17)
    img_cm = (rng.random(side * side, dtype=np.float32) - 0.45).astype(np.float32)
    img, n = padded(oracle, img_cm, side, side)
    A = 24
    sino, _ = oracle.sinograms(img, n, A, 15.0, [0], nthreads=8)
    cols = sino[0]
    offs = np.array([oracle.weighted_median(cols[a]) for a in range(A)])
    print(f"[nos signed] n={n} offsets {offs.min()}..{offs.max()}", flush=True)
    nos, center = core.nearest_orthonormal_sinogram(dev(cols))
    ref, rc = oracle.nearest_orthonormal_sinogram(cols)
    assert center == rc and tuple(nos.shape) == ref.shape
    q = nos.double()
    assert np.abs((q @ q.t()).cpu().numpy() - np.eye(A)).max() <= 1e-5
    err = normwise_err(nos.cpu().numpy(), ref)
    print(f"[nos signed] rows={ref.shape[1]} center={rc} err={err:.2e}", flush=True)
    assert err <= 1e-5, err


@pytest.mark.parametrize("eps", [1e-3, 1e-4, 1e-5])
def test_nos_ill_conditioned(oracle, gpu, eps):
    # smooth bumps plus a small random part: condition numbers ~2e4, ~2e5,
    # ~2e6.  The polar iteration's error is ~1e-16 * kappa, so the float
    # output matches the FP64 SVD polar factor to its own rounding; a route
    # through the Gram matrix (error ~1e-16 * kappa^2, 6e-4 at 2e6) fails.
    rng = np.random.default_rng(21)
    n, A = 301, 90
    rows_idx = np.arange(n)
    sino = np.empty((A, n), np.float32)
    for a in range(A):
        w = 20 + 8 * np.sin(2 * np.pi * a / A)
        sino[a] = np.exp(-0.5 * ((rows_idx - n // 2) / w) ** 2) + eps * rng.random(n)
    nos, center = core.nearest_orthonormal_sinogram(dev(sino))
    ref, rc = oracle.nearest_orthonormal_sinogram(sino)
    s = np.linalg.svd(np.asarray(sino, np.float64).T, compute_uv=False)  # offsets are 0
    err = normwise_err(nos.cpu().numpy(), ref)
    print(f"[nos ill] eps={eps} kappa~{s[0] / s[-1]:.1e} err={err:.2e}", flush=True)
    assert center == rc and tuple(nos.shape) == ref.shape
    assert err <= 2e-7, err


def test_nos_batched_matches_single_and_fp64(oracle, gpu):
    # sinograms whose columns peak at different positions, so each has its
    # own offsets, aligned height and centre
    rng = np.random.default_rng(6)
    n, A = 181, 72
    rows_idx = np.arange(n)
    sino = np.empty((4, A, n), np.float32)
    for t in range(4):
        for a in range(A):
            peak = n // 2 + int(round((t + 1) * 9 * np.sin(2 * np.pi * a / A + t)))
            sino[t, a] = rng.random(n) * 0.3 + np.exp(-0.5 * ((rows_idx - peak) / 12.0) ** 2)
    views, centers = core.nearest_orthonormal_sinograms(dev(sino))
    assert len(set(v.shape[1] for v in views)) > 1   # uneven heights in one batch
    for t in range(4):
        single, c1 = core.nearest_orthonormal_sinogram(dev(sino[t]))
        ref, rc = oracle.nearest_orthonormal_sinogram(sino[t])
        print(f"[nos batch] t={t} rows={ref.shape[1]} center={rc}", flush=True)
        assert centers[t] == c1 == rc and tuple(views[t].shape) == ref.shape
        got = views[t].cpu().numpy()
        assert normwise_err(got, single.cpu().numpy()) <= 1e-6
        assert normwise_err(got, ref) <= 1e-6
        q = views[t].double()
        assert np.abs((q @ q.t()).cpu().numpy() - np.eye(A)).max() <= 1e-5


def test_nos_rank_deficient(oracle, gpu):
    # repeated columns: the aligned matrix has rank 20 of 40, its Gram matrix
    # is singular and the polar factor is not unique; any valid one has
    # orthonormal columns and makes H = P^T A symmetric positive semidefinite
    # (A = P H)
    rng = np.random.default_rng(8)
    n, A = 121, 40
    base = (rng.random((20, n)) + 0.05).astype(np.float32)
    sino = np.concatenate([base, base])
    nos, center = core.nearest_orthonormal_sinogram(dev(sino))
    ref, rc = oracle.nearest_orthonormal_sinogram(sino)
    assert center == rc and tuple(nos.shape) == ref.shape
    p = nos.double().cpu().numpy().T          # rows x A
    rows = p.shape[0]
    offs = np.array([oracle.weighted_median(sino[a]) for a in range(A)]) - (n - 1) // 2
    aligned = np.zeros((rows, A))
    for a in range(A):
        aligned[offs.max() - offs[a]: offs.max() - offs[a] + n, a] = sino[a]
    assert np.abs(p.T @ p - np.eye(A)).max() <= 1e-5
    h = p.T @ aligned
    scale = np.abs(h).max()
    assert np.abs(h - h.T).max() <= 1e-5 * scale
    assert np.linalg.eigvalsh(0.5 * (h + h.T)).min() >= -1e-5 * scale


def test_transformer_orthonormal_pipeline(oracle, gpu):
    import tracetransform_amd as tt
    img = torch.from_numpy(inputs.uniform_image(64, 9).T.copy())
    pf = [tt.PFunctionalWrapper.parse(s) for s in ("H1", "H3")]
    tr = tt.Transformer(img, "x", angle_stepsize=4, orthonormal=True)
    n = tr.image.shape[0]
    assert n == oracle.pad_geometry(64, 64)[0]  # ceil(ceil(90)/sqrt 2) = 64
    sig = tr.getTransform(["0", "3"], pf)
    assert tuple(sig.shape) == (90, 4) and torch.isfinite(sig).all()
    assert pf[0].arguments.center is not None
    # column 0 is zscore(H1 circus of T0 on its NOS): mean 0, std 1
    col = sig[:, 0].double()
    assert abs(col.mean().item()) < 1e-5 and abs(col.std().item() - 1) < 1e-5
