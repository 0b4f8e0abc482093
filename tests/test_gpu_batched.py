"""GPU tests of the batched path (tt_forward_batched): one persistent launch
per stage over a batch of uneven padded images, captured into a HIP graph.

Bar: sinograms equal the single-image path bit for bit (the same line
function runs on the same inputs); P1/P2/P3 circus functions equal too (both
paths form P3's spectra with the same DFT GEMM kernel).  A few images are
also checked
against the oracle's medians via the single-image path's aux outputs.
Graph replays are checked against the eager path after the inputs change.
"""
import numpy as np
import pytest
import torch

from helpers import normwise_err, padded
from tracetransform_amd import core, inputs

pytestmark = pytest.mark.gpu

T05 = [0, 1, 2, 3, 4, 5]
P123 = [(1,), (2,), (3,)]


def make_batch(oracle, sides, seed):
    imgs = []
    for k, side in enumerate(sides):
        img, n = padded(oracle, inputs.blob_image(int(side), seed + k, nblobs=8,
                                                  sigma=(3, 20)), side, side)
        imgs.append(torch.from_numpy(img.reshape(n, n)).cuda())
    return imgs


def eager(img, A, step, codes, pf):
    s = core.sinograms(img, A, step, codes)
    return s, core.circus(s, pf)


def check_against_eager(batch, A, step, codes, pf):
    sinos, circs = batch.sino, batch.circus
    torch.cuda.synchronize()
    for k, img in enumerate(batch.images):
        s, c = eager(img, A, step, codes, pf)
        torch.cuda.synchronize()
        assert torch.equal(sinos[k], s), k
        if not pf:
            assert circs[k] is None
            continue
        cb, ce = circs[k].cpu().numpy(), c.cpu().numpy()
        for q, p in enumerate(pf):
            assert np.array_equal(cb[:, q], ce[:, q]), (k, q, p)


def test_batched_matches_single_image(oracle, gpu):
    sides = inputs.batch_sides(24, 16, 160, 3)
    imgs = make_batch(oracle, sides, 100)
    print(f"[batched] {len(imgs)} images, n in {min(i.shape[0] for i in imgs)}.."
          f"{max(i.shape[0] for i in imgs)}", flush=True)
    b = core.Batch(imgs, 90, 4.0, T05, P123)
    b.run(autotune=False)
    check_against_eager(b, 90, 4.0, T05, P123)


def test_batched_p3_wide_dft_tiles_match_single_image(oracle, gpu):
    # 80 images give the batch's P3 product enough work for 128-row DFT
    # tiles while each single-image call uses 64-row tiles: the spectra,
    # hence P1-P3, are the same bits either way
    sides = inputs.batch_sides(80, 150, 250, 12)
    imgs = make_batch(oracle, sides, 900)
    b = core.Batch(imgs, 90, 2.0, T05, P123)
    b.run(autotune=False)
    check_against_eager(b, 90, 2.0, T05, P123)


def test_batched_replays_follow_input_changes(oracle, gpu):
    sides = [64, 33, 100, 7, 64]
    imgs = make_batch(oracle, sides, 200)
    b = core.Batch(imgs, 36, 10.0, [0, 3, 1], P123)
    for rep in range(4):
        b.run(autotune=False)
        check_against_eager(b, 36, 10.0, [0, 3, 1], P123)
        # eager work between replays: new image contents in place
        for im in imgs:
            im.mul_(0.5).add_(torch.rand_like(im) * (im != 0))
        torch.cuda.synchronize()


def test_batched_autotune_and_subsets(oracle, gpu):
    sides = inputs.batch_sides(12, 16, 96, 9)
    imgs = make_batch(oracle, sides, 300)
    for codes, pf in (([0], []), ([2, 4], [(2,)]), ([5, 0], [(3,), (1,)])):
        b = core.Batch(imgs, 45, 8.0, codes, pf)
        b.run(autotune=True)
        w, nb, ta, tc = b.launch_shape()
        print(f"[batched] autotuned shape: {w} waves/block, {nb} blocks, "
              f"tile {ta} angles x {tc} columns", flush=True)
        assert w in (1, 2, 4) and nb > 0 and tc in (16, 32, 64) and ta % (64 // tc) == 0
        check_against_eager(b, 45, 8.0, codes, pf)
    core.get_context(0)  # leave autotune off for later tests
    b = core.Batch(imgs[:2], 4, 90.0, [0], [])
    b.run(autotune=False)


def test_batched_medians_match_oracle(oracle, gpu):
    sides = [40, 17, 61]
    imgs = make_batch(oracle, sides, 400)
    b = core.Batch(imgs, 60, 6.0, T05, P123)
    b.run(autotune=False)
    torch.cuda.synchronize()
    for k, img in enumerate(imgs):
        n = img.shape[0]
        o, _ = oracle.sinograms(img.cpu().numpy().ravel(), n, 60, 6.0, T05, nthreads=8)
        g = b.sino[k].cpu().numpy()
        assert np.array_equal(g[0], o[0])          # T0 is bit-exact
        for t in range(1, 6):
            assert normwise_err(g[t], o[t]) <= 2e-5


def test_batched_rejects_unsupported(gpu):
    img = torch.zeros((7, 7), device="cuda")
    with pytest.raises(Exception):
        core.Batch([img], 4, 90.0, [6], []).run()
    with pytest.raises(Exception):
        core.Batch([img], 4, 90.0, [0], [(0, 1, 3)]).run()
    with pytest.raises(ValueError):
        core.Batch([img], 4, 90.0, [0, 0], []).run()


def test_batched_backward_matches_single_image(oracle, gpu):
    sides = [48, 20, 75, 33]
    imgs = make_batch(oracle, sides, 500)
    codes, pf, A, step = [0, 3, 4, 5], P123, 72, 5.0
    bb = core.BatchBackward(imgs, A, step, codes, pf)
    rng = np.random.default_rng(6)
    for rep in range(3):
        for dc in bb.dcircus:
            dc.copy_(torch.from_numpy(
                rng.standard_normal(tuple(dc.shape)).astype(np.float32)))
        out = [o.clone() for o in bb.run()]
        again = bb.run()          # the same inputs: the same bits
        torch.cuda.synchronize()
        for k, img in enumerate(imgs):
            assert torch.equal(out[k], again[k]), (rep, k)
            ref = core.backward(img, A, step, codes, pf, bb.dcircus[k])
            ref2 = core.backward(img, A, step, codes, pf, bb.dcircus[k])
            torch.cuda.synchronize()
            assert torch.equal(ref, ref2), (rep, k)
            # the single-image gather may split the angles into groups whose
            # partial images it sums in double: equal to float rounding
            err = normwise_err(out[k].cpu().numpy(), ref.cpu().numpy())
            assert err <= 1e-6, (rep, k, err)
        for im in imgs:  # eager work between replays
            im.mul_(0.9)
        torch.cuda.synchronize()


def test_batched_backward_from_saved_forward(oracle, gpu):
    # the backward reading the forward's saved line data (medians, T3..T5
    # integrals, P3 spectra, P2 sources) instead of recomputing the forward
    # gives the same gradients bit for bit, replay after replay
    sides = [60, 23, 41]
    imgs = make_batch(oracle, sides, 700)
    codes, pf, A, step = [0, 3, 4, 5], P123, 90, 4.0
    fwd = core.Batch(imgs, A, step, codes, pf, save_for_backward=True)
    saved = core.BatchBackward(imgs, A, step, codes, pf, forward=fwd)
    plain = core.BatchBackward(imgs, A, step, codes, pf)
    rng = np.random.default_rng(9)
    for rep in range(3):
        fwd.run()
        g = torch.from_numpy(rng.standard_normal(tuple(saved.dcircus_all.shape))
                             .astype(np.float32)).cuda()
        saved.dcircus_all.copy_(g)
        plain.dcircus_all.copy_(g)
        a = [x.clone() for x in saved.run()]
        b = plain.run()
        torch.cuda.synchronize()
        for k in range(len(imgs)):
            assert torch.equal(a[k], b[k]), (rep, k)
        for im in imgs:
            im.mul_(1.1)
    # a block written for other functionals is refused
    other = core.Batch(imgs, A, step, [0, 3], pf, save_for_backward=True)
    other.run()
    with pytest.raises(ValueError):
        core.BatchBackward(imgs, A, step, codes, pf, forward=other).run()
    with pytest.raises(ValueError):
        core.BatchBackward(imgs, A, step, codes, pf, forward=core.Batch(imgs, A, step, codes, pf))


LINE_TILES = [(1, 64), (2, 64), (4, 64), (2, 32), (4, 32), (4, 16), (8, 16)]
GATHER_TILES = [(32, 16), (32, 4), (16, 16), (16, 2)]


def test_every_tile_gives_the_same_bits(oracle, gpu):
    # every candidate line tile (angles x columns of a wave item) and gather
    # tile (pixel side, waves per workgroup) gives sinograms, circus
    # functions, saved medians and gradients (from the saved line data and
    # from the recompute) bit-identical to the default tiles; 90 angles at 2
    # degrees put the 45, 90 and 135 degree turns inside tiles
    sides = [200, 23, 97, 64, 131]
    imgs = make_batch(oracle, sides, 800)
    codes, pf, A, step = T05, P123, 90, 2.0
    fwd = core.Batch(imgs, A, step, codes, pf, save_for_backward=True)
    saved = core.BatchBackward(imgs, A, step, codes, pf, forward=fwd)
    plain = core.BatchBackward(imgs, A, step, codes, pf)
    g = torch.from_numpy(np.random.default_rng(11).standard_normal(
        tuple(saved.dcircus_all.shape)).astype(np.float32)).cuda()

    def run_all():
        fwd.run(autotune=False)
        saved.dcircus_all.copy_(g)
        plain.dcircus_all.copy_(g)
        a = [x.clone() for x in saved.run()]
        b = [x.clone() for x in plain.run()]
        torch.cuda.synchronize()
        meds = [tuple(m.clone() for m in fwd.saved_medians(k)) for k in range(len(imgs))]
        return ([x.clone() for x in fwd.sino], [x.clone() for x in fwd.circus], meds, a, b)

    def same(ref, got, what):
        for k in range(len(imgs)):
            assert torch.equal(ref[0][k], got[0][k]), (what, "sinograms", k)
            assert torch.equal(ref[1][k], got[1][k]), (what, "circus", k)
            for m0, m1 in zip(ref[2][k], got[2][k]):
                assert torch.equal(m0, m1), (what, "medians", k)
            assert torch.equal(ref[3][k], got[3][k]), (what, "saved-data gradient", k)
            assert torch.equal(ref[4][k], got[4][k]), (what, "recompute gradient", k)

    try:
        core.set_tiles(0)
        ref = run_all()
        t = core.batched_tiles(0)
        assert (t["line_angles"], t["line_columns"], t["gather_tile"]) == (1, 64, 32), t
        for k in range(len(imgs)):   # the recompute equals the saved-data backward
            assert torch.equal(ref[3][k], ref[4][k]), k
        for lt in LINE_TILES:
            core.set_tiles(0, line=lt)
            got = run_all()
            t = core.batched_tiles(0)
            print(f"[tiles] line {lt}: reported {t}", flush=True)
            assert (t["line_angles"], t["line_columns"]) == lt, t
            same(ref, got, lt)
        for gt in GATHER_TILES:
            core.set_tiles(0, gather=gt)
            got = run_all()
            t = core.batched_tiles(0)
            print(f"[tiles] gather {gt}: reported {t}", flush=True)
            assert (t["gather_tile"], t["gather_waves"]) == gt, t
            same(ref, got, gt)
        with pytest.raises(Exception):
            core.set_tiles(0, line=(3, 32))
        with pytest.raises(Exception):
            core.set_tiles(0, gather=(8, 4))
    finally:
        core.set_tiles(0)
