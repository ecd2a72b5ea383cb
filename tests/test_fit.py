"""Fitting layer on the CPU: the optimiser against problems whose answer is
known in closed form, the central-difference gradient against analytic
gradients, and a complete fit driven by the oracle's objective that recovers
the tau variables a target was generated from."""
import math

import numpy as np
import pytest

from shape_atrophy_amd import fit, workloads as W


def rosen(X):
    x, y = X[0], X[1]
    return (1.0 - x) ** 2 + 100.0 * (y - x * x) ** 2


def rosen_grad(x):
    return np.array([-2.0 * (1.0 - x[0]) - 400.0 * x[0] * (x[1] - x[0] ** 2), 200.0 * (x[1] - x[0] ** 2)])


def smooth(X):
    # sum_i exp(0.3 x_i) + x_i^2 sin(x_i) per candidate column
    return (np.exp(0.3 * X) + X * X * np.sin(X)).sum(axis=0)


def smooth_grad(x):
    return 0.3 * np.exp(0.3 * x) + 2.0 * x * np.sin(x) + x * x * np.cos(x)


def test_fd_candidates_layout():
    x = np.array([1.0, -2.0, 0.5])
    X = fit.fd_candidates(x, np.array([0.1, 0.2, 0.3]))
    assert X.shape == (3, 7)
    assert np.array_equal(X[:, 0], x)
    d = X[:, 1:] - x[:, None]
    assert np.allclose(d, np.array([[0.1, -0.1, 0, 0, 0, 0], [0, 0, 0.2, -0.2, 0, 0], [0, 0, 0, 0, 0.3, -0.3]]))


def test_fd_gradient_matches_analytic():
    x = np.array([0.3, -1.2, 2.5, 0.0, 4.0])
    f, g = fit.fd_gradient(smooth, x, fit.OptData(fdStep=1e-5))
    assert f == pytest.approx(float(smooth(x[:, None])[0]), rel=1e-15)
    ga = smooth_grad(x)
    # central difference: error O(h^2 f''') + O(eps f / h)
    assert np.abs(g - ga).max() <= 1e-8 * max(1.0, np.abs(ga).max())


def test_fd_gradient_reports_failed_neighbour():
    def obj(X):
        v = smooth(X)
        v[2] = np.finfo(np.float64).max   # one neighbour's solve failed
        return v
    f, g = fit.fd_gradient(obj, np.array([0.5, 0.5]))
    assert math.isfinite(f) and g is None


def test_wolfe_step_satisfies_strong_wolfe():
    # phi(a) = f(x + a p) along the steepest descent of the Rosenbrock function
    x = np.array([-1.2, 1.0])
    g = rosen_grad(x)
    p = -g / np.linalg.norm(g)
    opt = fit.OptData()

    def phi(a):
        xn = x + a * p
        return float(rosen(xn[:, None])[0]), float(rosen_grad(xn) @ p), xn

    f0, d0 = float(rosen(x[:, None])[0]), float(g @ p)
    a, fa, xn = fit.wolfe_search(phi, f0, d0, opt, a1=1.0)
    _, da, _ = phi(a)
    assert fa <= f0 + opt.wolfe1 * a * d0
    assert abs(da) <= -opt.wolfe2 * d0


def test_wolfe_search_backs_off_failed_solves():
    # a "solve" fails beyond a = 0.5: the search must return a step below it
    def phi(a):
        if a > 0.5:
            return math.inf, math.nan, None
        return (a - 0.4) ** 2, 2.0 * (a - 0.4), a
    a, fa, _ = fit.wolfe_search(phi, 0.16, -0.8, fit.OptData(), a1=4.0)
    assert a <= 0.5 and fa < 0.16


def test_bfgs_rosenbrock_analytic_and_fd():
    opt = fit.OptData(itrMax=200, tolVal=1e-10)
    r = fit.bfgs(rosen, [-1.2, 1.0], opt, grad=rosen_grad)
    assert r.status in ("gradient", "step")
    assert np.abs(r.x - 1.0).max() < 1e-6
    r2 = fit.bfgs(rosen, [-1.2, 1.0], fit.OptData(itrMax=200, tolVal=1e-9, fdStep=1e-6))
    assert np.abs(r2.x - 1.0).max() < 1e-5
    # every gradient is one batch of 2n + 1 candidates
    assert r2.evaluations == 5 * r2.batches


def test_bfgs_quadratic_reaches_the_linear_solve():
    rng = np.random.default_rng(3)
    Q = rng.standard_normal((6, 6))
    A = Q @ Q.T + 6.0 * np.eye(6)
    b = rng.standard_normal(6)

    def quad(X):
        return 0.5 * np.einsum("ij,ik,kj->j", X, A, X) - b @ X

    r = fit.bfgs(quad, np.zeros(6), fit.OptData(itrMax=100, tolVal=1e-11), grad=lambda x: A @ x - b)
    assert np.abs(r.x - np.linalg.solve(A, b)).max() < 1e-9
    # BFGS with exact line searches ends in at most n steps on a quadratic;
    # with Wolfe steps it still ends within a small multiple
    assert r.iterations <= 30


def test_fit_recovers_tau_variables_with_the_oracle(oracle):
    """Target = the varifold of the oracle's own final shape at theta*; a fit
    from a perturbed theta must return to theta*."""
    p = W.ring_problem(n_rings=5, n_per_ring=32)
    p["params"] = dict(p["params"], timeEnd=0.25)
    th = p["tauVar"][:, 0].copy()
    r = oracle.objfcn(p)
    assert r["status"] == 0
    p["target"] = fit.target_from_shape(p, r["lmkStk"][-1])
    assert oracle.objfcn(p)["obj"] < 1e-12

    def objective(X):
        out = np.empty(X.shape[1])
        for j in range(X.shape[1]):
            rr = oracle.objfcn(dict(p, tauVar=X[:, j:j + 1].copy()))
            out[j] = rr["obj"] if rr["status"] == 0 else math.inf
        return out

    x0 = th * np.array([1.05, 0.95, 1.08, 0.95])
    res = fit.bfgs(objective, x0, fit.OptData(itrMax=40, tolVal=1e-7))
    assert res.f < 1e-10
    assert np.abs(res.x - th).max() < 1e-4
