"""Generate the polynomial schedule of the nearest-orthonormal-sinogram
polar iteration (the ``kPolarSteps`` table in tt_capi.cpp).

The iteration is X <- X q(X^T X) with q of degree 4 in X^T X, so each step
maps every singular value x of X to p(x) = x q(x^2), an odd polynomial of
degree 9, and leaves the singular vectors (hence the polar factor) alone.
X starts as A / ||A^T A||_F^(1/2), whose singular values lie in
[sigma_min / sigma_max * c, 1] with c >= 360^(-1/4) for 360 columns.

Two phases, both found by linear programming over dense grids (scipy):

* growth: one polynomial, applied while the lower end l of the interval is
  small, that maximises the gain g subject to p(x) >= min(g x, THETA) and
  p(x) <= 1 on [L_MIN, 1 + SLACK].  The floor THETA keeps mid-range
  singular values away from 0 (the plain minimax polynomial touches its
  lower bound at interior points, and for tiny l rounding would push them
  through 0 and flip a sign of the polar factor); it costs little gain
  (7.41 instead of 7.53 per step at THETA = 0.1).
* convergence: greedy minimax steps, each the odd polynomial minimising
  max |1 - p(x)| on the current interval (widened by SLACK), while that
  maximum exceeds 1e-3; then the polynomial with p(1) = 1 and its first
  four derivatives 0 at 1 (fifth-order convergence, no interval to widen)
  until |p - 1| < TOL.

Singular values below L_MIN (condition numbers above about 1e8) do not
converge; the library detects that from X^T X and falls back.

Run ``python polar_schedule.py`` to print the table (a few minutes).
"""
import numpy as np
from scipy.optimize import linprog

DEG = 4          # q has degree 4 in X^T X: p(x) = x q(x^2) has degree 9
L_MIN = 1e-8     # smallest singular value (relative to the scale) handled
THETA = 0.1      # growth floor
SLACK = 0.02     # design intervals reach 2 % beyond their upper end
TOL = 1e-14


def grid(l, u, n=3000):
    return np.unique(np.concatenate([np.geomspace(l, u, n), np.linspace(l, u, n)]))


def apply(c, x):
    y = x * x
    return x * sum(cj * y ** j for j, cj in enumerate(c))


def growth_feasible(g, n=3000):
    xs = grid(L_MIN, 1.0 + SLACK, n)
    P = np.stack([xs ** (2 * j) for j in range(DEG + 1)], 1)   # p(x) / x
    lo = np.minimum(g * xs, THETA)
    A = np.vstack([-P, P * xs[:, None]])
    b = np.concatenate([-lo / xs, np.ones(len(xs))])
    r = linprog(np.zeros(DEG + 1), A_ub=A, b_ub=b, bounds=[(None, None)] * (DEG + 1),
                method="highs")
    return r.x if r.status == 0 else None


def growth_poly():
    lo, hi, best = 1.0, 12.0, None
    for _ in range(24):
        mid = 0.5 * (lo + hi)
        c = growth_feasible(mid)
        if c is not None:
            lo, best = mid, c
        else:
            hi = mid
    return lo, best


def minimax_poly(l, u):
    """Odd p minimising max |1 - p| on [l, u]: variables c, E."""
    xs = grid(l, u)
    V = np.stack([xs ** (2 * j + 1) for j in range(DEG + 1)], 1)
    cost = np.zeros(DEG + 2)
    cost[-1] = 1.0
    ones = np.ones((len(xs), 1))
    A = np.vstack([np.hstack([-V, -ones]), np.hstack([V, -ones])])
    b = np.concatenate([-np.ones(len(xs)), np.ones(len(xs))])
    r = linprog(cost, A_ub=A, b_ub=b, bounds=[(None, None)] * (DEG + 2), method="highs",
                options={"primal_feasibility_tolerance": 1e-10})
    assert r.status == 0, r.message
    return r.x[:-1]


def polish_poly():
    """p(1) = 1, p^(k)(1) = 0 for k = 1..4, solved in exact rationals
    (315/128, -420/128, 378/128, -180/128, 35/128)."""
    from fractions import Fraction
    n = DEG + 1
    M = [[Fraction(1)] * n for _ in range(n)]
    for k in range(n):
        for j in range(n):
            f = Fraction(1)
            for t in range(k):
                f *= 2 * j + 1 - t
            M[k][j] = f
    b = [Fraction(1)] + [Fraction(0)] * (n - 1)
    for i in range(n):                       # Gauss-Jordan
        piv = next(r for r in range(i, n) if M[r][i] != 0)
        M[i], M[piv], b[i], b[piv] = M[piv], M[i], b[piv], b[i]
        for r in range(n):
            if r != i and M[r][i] != 0:
                f = M[r][i] / M[i][i]
                M[r] = [a - f * c for a, c in zip(M[r], M[i])]
                b[r] -= f * b[i]
    return np.array([float(b[i] / M[i][i]) for i in range(n)])


def schedule():
    g, cg = growth_poly()
    steps = []
    xs = grid(L_MIN, 1.0 + SLACK, 20000)
    x = xs.copy()
    # the floor holds to the grid's resolution: grow until l >= THETA / 2
    while x.min() < 0.5 * THETA and len(steps) < 24:
        steps.append(cg)
        x = apply(cg, x)
    while max(1.0 - x.min(), x.max() - 1.0) > 1e-3 and len(steps) < 24:
        c = minimax_poly(x.min(), x.max() * (1.0 + SLACK))
        steps.append(c)
        x = apply(c, x)
    cp = polish_poly()
    while max(1.0 - x.min(), x.max() - 1.0) > TOL and len(steps) < 24:
        steps.append(cp)
        x = apply(cp, x)
    return g, steps, x


if __name__ == "__main__":
    g, steps, x = schedule()
    print(f"// growth gain {g:.4f}; {len(steps)} steps; max |p - 1| on "
          f"[{L_MIN:g}, {1 + SLACK:g}] after them: {np.abs(x - 1).max():.3e}")
    x = grid(L_MIN, 1.0 + SLACK, 20000)
    for c in steps:
        x = apply(c, x)
        print("    {" + ", ".join(f"{v:.17g}" for v in c) + "},"
              f"  // -> [{x.min():.6g}, {x.max():.6g}]")
