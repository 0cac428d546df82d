"""The optimisers of the inversion / UDE loop on the CPU (the reference's
Adam, Descent and LBFGS(linesearch = BackTracking()) choices,
swe_2D_inversion.jl:680-713, swe_2D_UDE.jl:278-300).  Optim.jl and
LineSearches.jl are not in the reference tree, so their iterates are not
reproduced bit for bit; the checks are the methods' mathematical
properties:

  * L-BFGS with exact line searches on a strictly convex quadratic is the
    conjugate-gradient method and ends within n iterations (this pins the
    two-loop recursion and its scaling: a transposed loop order, a missing
    rho or a wrong sign break it);
  * with the backtracking search it reaches a quadratic's minimiser to
    1e-10 and the Rosenbrock minimum to 1e-8;
  * the first backtracking reduction minimises the interpolating quadratic,
    which for a quadratic phi is phi's own minimiser; every accepted step
    satisfies the Armijo condition; a non-finite trial value is backed off.
"""
import numpy as np
import pytest

from hgswe import inversion as inv


def _quadratic(n, cond, seed):
    rng = np.random.default_rng(seed)
    Qm, _ = np.linalg.qr(rng.normal(size=(n, n)))
    A = Qm @ np.diag(np.geomspace(1.0, cond, n)) @ Qm.T
    b = rng.normal(size=n)
    return A, b, np.linalg.solve(A, b)


def test_two_loop_with_exact_line_search_terminates_in_n_steps():
    n = 8
    A, b, xs = _quadratic(n, 100.0, 0)
    opt = inv.LBFGS()
    x = np.zeros(n)
    g = A @ x - b
    S, Y = [], []
    for k in range(n):
        d = opt.direction(g, S, Y)
        a = -(g @ d) / (d @ A @ d)
        s = a * d
        gn = A @ (x + s) - b
        S.append(s)
        Y.append(gn - g)
        x, g = x + s, gn
    assert np.linalg.norm(x - xs) <= 1e-10 * np.linalg.norm(xs)


def test_direction_without_history_is_steepest_descent():
    g = np.array([0.3, -2.0, 1.5])
    assert np.array_equal(inv.LBFGS().direction(g, [], []), -g)


def test_lbfgs_backtracking_minimises_a_quadratic():
    n = 8
    A, b, xs = _quadratic(n, 50.0, 1)
    # written about the minimiser (f* = 0), so that loss decreases stay
    # resolvable in floating point down to tiny gradients
    f = lambda x: (0.5 * (x - xs) @ A @ (x - xs), A @ (x - xs))  # noqa: E731
    x, fx, rc, it = inv.LBFGS(g_tol=1e-13).minimize(f, np.zeros(n), 200)
    assert rc == "Success"
    assert np.linalg.norm(x - xs) <= 1e-10 * np.linalg.norm(xs)
    assert it <= 8 * n


def test_lbfgs_rosenbrock():
    def f(x):
        a, b = x
        v = (1 - a) ** 2 + 100 * (b - a * a) ** 2
        g = np.array([-2 * (1 - a) - 400 * a * (b - a * a), 200 * (b - a * a)])
        return v, g

    x, fx, rc, it = inv.LBFGS().minimize(f, np.array([-1.2, 1.0]), 500)
    assert rc == "Success" and np.max(np.abs(x - 1.0)) < 1e-8


def test_backtracking_first_reduction_is_the_quadratic_minimiser():
    # phi(a) = 3 (a - 0.2)^2 - 0.12: phi(1) fails Armijo; the interpolating
    # quadratic is phi itself, so the first reduction lands on a = 0.2
    phi = lambda a: (3.0 * (a - 0.2) ** 2 - 0.12 + 0.0, None)  # noqa: E731
    phi0, dphi0 = phi(0.0)[0], -1.2
    a, fa, _ = inv.backtracking(phi, phi0, dphi0)
    assert a == pytest.approx(0.2, rel=1e-14)
    assert fa <= phi0 + 1e-4 * a * dphi0


def test_backtracking_cubic_steps_keep_armijo_and_bounds():
    # phi falls steeply at 0 and rises very fast: several reductions
    trials = []

    def phi(a):
        trials.append(a)
        return -a + 1e4 * a ** 4, None

    a, fa, _ = inv.backtracking(phi, 0.0, -1.0)
    assert fa <= 1e-4 * a * -1.0
    assert len(trials) >= 3
    for t0, t1 in zip(trials, trials[1:]):
        assert 0.1 * t0 - 1e-15 <= t1 <= 0.5 * t0 + 1e-15


def test_backtracking_backs_off_a_non_finite_trial():
    phi = lambda a: (np.inf if a > 0.3 else -a + a * a, None)  # noqa: E731
    a, fa, _ = inv.backtracking(phi, 0.0, -1.0)
    assert a <= 0.3 and np.isfinite(fa) and fa <= 1e-4 * a * -1.0


def test_descent_step():
    p, g = np.array([1.0, -2.0]), np.array([0.5, 0.25])
    assert np.array_equal(inv.Descent(lr=0.2).step(p, g), p - 0.2 * g)


def test_staged_schedule_adam_then_lbfgs():
    def fg(x):
        a, b = x
        v = (1 - a) ** 2 + 100 * (b - a * a) ** 2
        return v, np.array([-2 * (1 - a) - 400 * a * (b - a * a), 200 * (b - a * a)])

    seen = []
    p, log = inv.run_optimizers(fg, np.array([-1.2, 1.0]), [("Adam", 1e-2, 50), ("Descent", 1e-4, 0),
                                                            ("LBFGS", None, 500), ("Adam", 1e-2, 10)],
                                callback=lambda k, it, L, x: seen.append(k))
    assert [s[0] for s in log] == ["Adam", "LBFGS"]  # the zero-iteration stage is skipped
    assert log[1][1] == "Success"  # LBFGS converged, so the last Adam stage does not run
    assert np.max(np.abs(p - 1.0)) < 1e-8
    assert set(seen) == {0, 2}
    with pytest.raises(ValueError):
        inv.run_optimizers(fg, p, [("SGD", 0.1, 3)])
