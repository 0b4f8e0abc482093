"""CPU checks of the polar-iteration schedule behind the nearest orthonormal
sinogram (``kPolarSteps`` in tracetransform_amd/csrc/tt_capi.cpp).

The library runs X <- X q(X^T X) with the table's q; these tests read the
table from the source and check what the schedule promises, independently
of the GPU: the scalar map drives every singular value in [1e-8, 1.02] to 1
without crossing 0 (a crossing would flip a sign of the polar factor), and
the same matrix iteration in numpy float64 (the first kPolarTallSteps steps
on X, the rest on X^T X, as the library runs them) reproduces numpy's SVD
polar factor U V^T for an ill-conditioned matrix (kappa 3e6, the Radon sinogram
of the 1024 blob workload is 3.8e6).
"""
import os
import re

import numpy as np
import pytest

SRC = os.path.join(os.path.dirname(__file__), "..", "tracetransform_amd", "csrc",
                   "tt_capi.cpp")


def table():
    text = open(SRC).read()
    body = re.search(r"kPolarSteps\[\] = \{(.*?)\n\};", text, re.S).group(1)
    rows = re.findall(r"\{\{([^}]*)\}\}", body)
    return [np.array([float(v) for v in r.split(",")]) for r in rows]


def tall_steps():
    return int(re.search(r"kPolarTallSteps = (\d+);", open(SRC).read()).group(1))


def p(c, x):
    y = x * x
    return x * sum(cj * y ** j for j, cj in enumerate(c))


def test_table_shape():
    t = table()
    assert 8 <= len(t) <= 16 and all(c.size == 5 for c in t)


def test_scalar_map_converges_without_sign_flips():
    x = np.unique(np.concatenate([np.geomspace(1e-8, 1.02, 200001),
                                  np.linspace(1e-8, 1.02, 200001)]))
    for c in table():
        x = p(c, x)
        assert x.min() > 0.0
    assert np.abs(x - 1.0).max() <= 1e-14


def test_below_reach_does_not_converge():
    # the library must detect these (X^T X far from I) and fall back
    x = np.array([1e-12, 1e-10])
    for c in table():
        x = p(c, x)
    assert np.abs(x - 1.0).min() > 1e-3


@pytest.mark.parametrize("kappa", [1.0, 1e3, 3e6])
def test_matrix_iteration_matches_svd_polar(kappa):
    rng = np.random.default_rng(7)
    m, n = 300, 60
    u, _ = np.linalg.qr(rng.standard_normal((m, n)))
    v, _ = np.linalg.qr(rng.standard_normal((n, n)))
    s = np.geomspace(1.0, 1.0 / kappa, n) * 37.0
    a = (u * s) @ v.T
    g = a.T @ a
    x = a / np.sqrt(np.linalg.norm(g))         # X_0 = A ||A^T A||_F^(-1/2)

    def q_of(c, sm):
        s2 = sm @ sm
        return (c[0] * np.eye(n) + c[1] * sm + c[2] * s2) + s2 @ (c[3] * sm + c[4] * s2)

    # as the library: the first steps on X, the rest on S = X^T X
    steps = table()
    t = tall_steps()
    for c in steps[:t]:
        x = x @ q_of(c, x.T @ x)
    sm, m_acc = x.T @ x, np.eye(n)
    for c in steps[t:]:
        q = q_of(c, sm)
        m_acc = m_acc @ q
        sm = sm @ (q @ q)
    x = x @ m_acc
    assert np.abs(sm - np.eye(n)).max() <= 1e-10       # the library's test
    assert np.abs(x.T @ x - np.eye(n)).max() <= 1e-10
    ref = u @ v.T
    err = np.linalg.norm(x - ref) / np.linalg.norm(ref)
    # rounding perturbs the factor by about eps * kappa
    assert err <= 1e-14 * max(kappa, 10.0), err
