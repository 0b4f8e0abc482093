"""CPU check of the rules the T6/T7 sort kernel uses to avoid sorting whole
tails (sino_sort_kernel in tracetransform_amd/csrc/tt_kernels.hip): a NaN
shortcut for tails with a negative sample, and sorting only the part before
the trailing zeros, with the zero run placed by rule and the Eigen-order sum
and the walk read through the compact array.  A Python restatement of those
rules must give the oracle's full-sort T6/T7 (c++/src/functionals.cpp:226-276)
on columns with ties, zeros, -0 and negatives."""
import numpy as np


def ord_key(v):
    v = np.float32(v)
    if v == 0:
        v = np.float32(0.0)
    b = int(np.frombuffer(v.tobytes(), np.uint32)[0])
    return (~b & 0xFFFFFFFF) if b & 0x80000000 else (b | 0x80000000)


def walk_view(c, ins, zt, le, length, S):
    f = np.float32
    acc = f(0)
    for j in range(ins):
        acc = f(acc + c[j])
        if f(2) * acc >= S:
            return j
    if zt > 0 and f(2) * acc >= S:
        return ins
    for j in range(ins, le):
        acc = f(acc + c[j])
        if f(2) * acc >= S:
            return j + zt
    return length - 1


def trimmed_tail(oracle, x, t6):
    """Value, rank and source of T6 (t6) or T7 by the kernel's rules."""
    _, d = oracle.tfunctional(6 if t6 else 7, x)
    base = d["sm"] if t6 else d["m"]
    tail = x[base:].astype(np.float32)
    length = tail.size
    kv = (np.arange(length, dtype=np.float32) * tail).astype(np.float32) if t6 else tail
    keys = [(ord_key(kv[i]) << 32) | i for i in range(length)]
    if np.any(~(tail >= 0)):
        key, k = max(keys), length - 1
    else:
        nz = np.nonzero(tail != 0)[0]
        le = int(nz[-1]) + 1 if nz.size else 0
        first = sorted(keys[:le])
        ins = int(np.sum(kv[:le] == 0)) if t6 else 0
        zt = length - le
        c = np.array([np.sqrt(tail[k_ & 0xFFFFFFFF]) for k_ in first], np.float32)
        full = np.concatenate([c[:ins], np.zeros(zt, np.float32), c[ins:]])
        S = oracle.eigen_sum(full)  # Eigen order over the full sequence
        k = walk_view(c, ins, zt, le, length, S)
        assert k == oracle.weighted_median(full)
        if ins <= k < ins + zt:
            key = (ord_key(0.0) << 32) | (le + (k - ins))
        else:
            key = first[k if k < ins else k - zt]
    i = key & 0xFFFFFFFF
    return np.float32(kv[i]), k, i


def test_trimmed_tail_rules_match_full_sort(oracle):
    rng = np.random.default_rng(0)
    for trial in range(1200):
        n = int(rng.integers(1, 160))
        kind = trial % 6
        x = rng.random(n).astype(np.float32)
        if kind == 1:    # trailing zeros (the line leaves the image)
            x[int(rng.integers(0, n)):] = 0
        elif kind == 2:  # leading, interior and trailing zeros
            a, b = sorted(rng.integers(0, n, 2))
            x[:a] = 0
            x[b:] = 0
            x[rng.random(n) < 0.1] = 0
        elif kind == 3:  # ties
            x = (rng.integers(0, 4, n) / 4).astype(np.float32)
            x[int(rng.integers(0, n)):] = 0
        elif kind == 4:  # negatives
            x = (rng.random(n) - 0.3).astype(np.float32)
            x[int(rng.integers(0, n)):] = 0
        elif kind == 5:  # all zero, some -0
            x = np.zeros(n, np.float32)
            if trial % 2:
                x[rng.random(n) < 0.5] = -0.0
        for t6 in (True, False):
            v, d = oracle.tfunctional(6 if t6 else 7, x)
            got_v, got_k, got_i = trimmed_tail(oracle, x, t6)
            assert got_k == d["rank"], (trial, t6)
            assert got_v == v or (np.isnan(got_v) and np.isnan(v)), (trial, t6)
            if t6:
                assert got_i == d["src"], (trial, t6)


def block_walk(c, S, sign_test=True):
    """findWeightedMedian's walk as walk_segment runs it: blocks of 8 summed
    first and skipped when their last prefix fails the test, which is exact
    only when no value of the block is negative or NaN (sign_test keeps the
    element walk for such blocks)."""
    f = np.float32
    acc = f(0)
    j = 0
    while j + 8 <= c.size:
        t = acc
        for u in range(8):
            t = f(t + c[j + u])
        mono = bool(np.all(c[j:j + 8] >= 0)) if sign_test else True
        if mono and not (f(2) * t >= S):
            acc = t
            j += 8
            continue
        for u in range(8):
            acc = f(acc + c[j + u])
            if f(2) * acc >= S:
                return j + u
        j += 8
    for k in range(j, c.size):
        acc = f(acc + c[k])
        if f(2) * acc >= S:
            return k
    return c.size - 1


def test_block_walk_exact_for_signed_columns(oracle):
    # P2 walks its sorted column and the orthonormal step its raw columns,
    # both of any sign; the block skip must give findWeightedMedian's index
    # (c++/src/functionals.cpp:21-30) on every one of them
    rng = np.random.default_rng(3)
    cols = [np.array([-3.0] + [-1.0] * 8, np.float32),
            np.array([-1.0] * 16 + [5.0] * 3, np.float32),
            np.array([2.0, -1.0, -1.0, -1.0, 0.5, 0.5, 0.5, 0.5, 0.5, 0.5], np.float32),
            np.full(24, np.nan, np.float32), np.zeros(17, np.float32)]
    for trial in range(1500):
        n = int(rng.integers(1, 200))
        x = (rng.random(n) - (0.3, 0.5, 0.9)[trial % 3]).astype(np.float32)
        cols.append(np.sort(x) if trial % 2 else x)
    fails_without = 0
    for c in cols:
        S = oracle.eigen_sum(c)
        want = oracle.weighted_median(c)
        assert block_walk(c, S) == want, c
        fails_without += block_walk(c, S, sign_test=False) != want
    # the unguarded skip (monotone prefixes assumed) misses some of them,
    # starting with the sorted column [-3, -1 x 8]: index 0, not 8
    assert block_walk(cols[0], oracle.eigen_sum(cols[0]), sign_test=False) == 8
    assert oracle.weighted_median(cols[0]) == 0
    assert fails_without > 50
