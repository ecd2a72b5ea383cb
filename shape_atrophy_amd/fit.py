"""Fitting the tau variables of the forward model to a target shape.

The reference evaluates the objective for a matrix of candidates
(``computeObjfcn``: ``h_tauVarMat`` is nVar x nCand, ``CUDA/computeObjfcn.m:2``)
and declares, but never fills, the state of a quasi-Newton optimiser with a
Wolfe line search (``struct optdata``, ``CUDA/2D_frame/struct.h:143-158``;
its arena ``assignOptStructMemory``, ``assignStructMemory.cu:5-31``).  This
module supplies that caller side:

* ``fd_gradient``: the gradient with respect to tauVar by central
  differences.  The 2 nVar + 1 candidates (centre and the +-h_i e_i
  neighbours) form one candidate matrix, so they are evaluated as one batch:
  concurrently on the request server, or in sequence on a session.  This is
  not an adjoint of the time loop: the reference has none, and a central
  difference needs nothing beyond the forward solve.
* ``bfgs``: BFGS on the inverse Hessian approximation (``h_apHMat``) with a
  line search for the strong Wolfe conditions (``wolfe1``, ``wolfe2``),
  following Nocedal & Wright, Numerical Optimization (2nd ed.), Algorithms
  3.5, 3.6 and 6.1, eq. (6.17) for the update and eq. (6.20) for the initial
  scaling.

An objective is any callable mapping an nVar x nCand candidate matrix to nCand
objective values; a failed solve reports ``inf`` (the reference's DBL_MAX).
``ServerObjective`` and ``SessionObjective`` evaluate the forward model on the
GPU.
"""
from __future__ import annotations

import dataclasses
import math
import time
from typing import Callable

import numpy as np

FAIL = np.finfo(np.float64).max   # objective of a failed solve (deform.cu:68-73)


@dataclasses.dataclass
class OptData:
    """Optimiser settings and work vectors; the fields of ``struct optdata``
    (struct.h:143-158) plus the finite-difference step.

    itrMax   maximum BFGS iterations
    tolVal   stop when |g| <= tolVal * max(|g0|, 1e-300) (relative gradient
             reduction) or when a step changes x by less than tolVal * (|x| + tolVal)
    wolfe1   sufficient-decrease constant c1
    wolfe2   curvature constant c2 (strong Wolfe: |phi'(a)| <= c2 |phi'(0)|)
    fdStep   central-difference step relative to max(|x_i|, fdFloor)
    """
    itrMax: int = 50
    tolVal: float = 1e-6
    wolfe1: float = 1e-4
    wolfe2: float = 0.9
    verbose: bool = False
    fdStep: float = 1e-4
    fdFloor: float = 0.1
    maxTrials: int = 20
    # work vectors (h_apHMat, h_dirVec, h_posNxt, h_grdNxt, h_dspVec, h_dgdVec)
    apHMat: np.ndarray | None = None
    dirVec: np.ndarray | None = None
    posNxt: np.ndarray | None = None
    grdNxt: np.ndarray | None = None
    dspVec: np.ndarray | None = None
    dgdVec: np.ndarray | None = None

    def assign(self, n: int):
        """Work vectors of an n-variable problem (assignOptStructMemory)."""
        self.apHMat = np.eye(n)
        self.dirVec = np.zeros(n)
        self.posNxt = np.zeros(n)
        self.grdNxt = np.zeros(n)
        self.dspVec = np.zeros(n)
        self.dgdVec = np.zeros(n)


@dataclasses.dataclass
class FitResult:
    x: np.ndarray
    f: float
    g: np.ndarray
    iterations: int
    evaluations: int        # objective values computed (candidates)
    batches: int            # candidate matrices evaluated
    status: str             # "gradient", "step", "itrMax", "line search"
    seconds: float
    history: list


def _finite(f: float) -> bool:
    return math.isfinite(f) and f < FAIL


def fd_steps(x: np.ndarray, opt: OptData) -> np.ndarray:
    return opt.fdStep * np.maximum(np.abs(x), opt.fdFloor)


def fd_candidates(x: np.ndarray, h: np.ndarray) -> np.ndarray:
    """Candidate matrix [x, x + h_1 e_1, x - h_1 e_1, ..., x - h_n e_n]
    (n x (2n + 1))."""
    n = x.size
    X = np.repeat(x[:, None], 2 * n + 1, axis=1)
    for i in range(n):
        X[i, 1 + 2 * i] += h[i]
        X[i, 2 + 2 * i] -= h[i]
    return X


def fd_gradient(objective: Callable[[np.ndarray], np.ndarray], x, opt: OptData | None = None):
    """Objective and central-difference gradient at x from one batch of
    2n + 1 candidates.  Returns (f, g); g is None when a neighbour failed."""
    opt = opt or OptData()
    x = np.asarray(x, dtype=np.float64).reshape(-1)
    h = fd_steps(x, opt)
    X = fd_candidates(x, h)
    # the steps actually represented in floating point
    hp = X[np.arange(x.size), 1 + 2 * np.arange(x.size)] - x
    hm = x - X[np.arange(x.size), 2 + 2 * np.arange(x.size)]
    vals = np.asarray(objective(X), dtype=np.float64).reshape(-1)
    f = float(vals[0])
    if not all(_finite(v) for v in vals):
        return f if _finite(f) else math.inf, None
    g = (vals[1::2] - vals[2::2]) / (hp + hm)
    return f, g


class _Counter:
    def __init__(self, objective):
        self.objective = objective
        self.evaluations = 0
        self.batches = 0

    def __call__(self, X):
        self.evaluations += X.shape[1]
        self.batches += 1
        vals = np.asarray(self.objective(X), dtype=np.float64).reshape(-1)
        return np.where(np.isfinite(vals) & (vals < FAIL), vals, math.inf)


def _cubic_min(a0, f0, d0, a1, f1, d1):
    """Minimiser of the cubic through (a0, f0, d0), (a1, f1, d1)
    (Nocedal & Wright eq. 3.59); None if it does not exist."""
    if not all(math.isfinite(v) for v in (f0, d0, f1, d1)) or a0 == a1:
        return None
    d1_ = d0 + d1 - 3.0 * (f0 - f1) / (a0 - a1)
    disc = d1_ * d1_ - d0 * d1
    if disc < 0:
        return None
    d2 = math.copysign(math.sqrt(disc), a1 - a0)
    den = d1 - d0 + 2.0 * d2
    if den == 0:
        return None
    return a1 - (a1 - a0) * (d1 + d2 - d1_) / den


def wolfe_search(phi, f0: float, d0: float, opt: OptData, a1: float = 1.0):
    """Step length satisfying the strong Wolfe conditions along a descent
    direction (Nocedal & Wright Algorithms 3.5 and 3.6).  phi(a) returns
    (f, phi'(a), payload); a non-finite f (failed solve) counts as a failed
    sufficient-decrease test.  Returns (a, f, payload) or None."""
    c1, c2 = opt.wolfe1, opt.wolfe2
    trials = 0

    def zoom(lo, flo, dlo, plo, hi, fhi, dhi):
        nonlocal trials
        while trials < opt.maxTrials:
            w = hi - lo
            a = _cubic_min(lo, flo, dlo, hi, fhi, dhi)
            if a is None or not (min(lo, hi) + 0.1 * abs(w) <= a <= max(lo, hi) - 0.1 * abs(w)):
                a = lo + 0.5 * w
            f, d, p = phi(a)
            trials += 1
            if not math.isfinite(f) or f > f0 + c1 * a * d0 or f >= flo:
                hi, fhi, dhi = a, f, d
            else:
                if abs(d) <= -c2 * d0:
                    return a, f, p
                if d * (hi - lo) >= 0:
                    hi, fhi, dhi = lo, flo, dlo
                lo, flo, dlo, plo = a, f, d, p
        # budget spent: the best point with sufficient decrease, if any
        return (lo, flo, plo) if lo > 0 else None

    a_prev, f_prev, d_prev, p_prev = 0.0, f0, d0, None
    a = a1
    while trials < opt.maxTrials:
        f, d, p = phi(a)
        trials += 1
        if not math.isfinite(f) or f > f0 + c1 * a * d0 or (a_prev > 0 and f >= f_prev):
            return zoom(a_prev, f_prev, d_prev, p_prev, a, f, d)
        if abs(d) <= -c2 * d0:
            return a, f, p
        if d >= 0:
            return zoom(a, f, d, p, a_prev, f_prev, d_prev)
        a_prev, f_prev, d_prev, p_prev = a, f, d, p
        a = 2.0 * a
    return (a_prev, f_prev, p_prev) if a_prev > 0 else None


def bfgs(objective, x0, opt: OptData | None = None, grad: Callable | None = None) -> FitResult:
    """Minimise objective over tauVar from x0 (Nocedal & Wright Algorithm 6.1).

    objective(X) takes an n x nCand candidate matrix and returns nCand values.
    The gradient is the central difference of fd_gradient unless grad(x) is
    given (analytic tests)."""
    opt = opt or OptData()
    t0 = time.perf_counter()
    x = np.asarray(x0, dtype=np.float64).reshape(-1).copy()
    n = x.size
    opt.assign(n)
    obj = _Counter(objective)

    def fg(xx):
        if grad is not None:
            f = float(obj(xx[:, None])[0])
            return f, (np.asarray(grad(xx), dtype=np.float64) if math.isfinite(f) else None)
        return fd_gradient(obj, xx, opt)

    f, g = fg(x)
    if not math.isfinite(f) or g is None:
        raise ValueError("the objective fails at the starting point")
    g0n = max(float(np.linalg.norm(g)), 1e-300)
    hist = [dict(it=0, f=f, gnorm=float(np.linalg.norm(g)), step=0.0)]
    scaled = False
    status = "itrMax"
    it = 0
    for it in range(1, opt.itrMax + 1):
        if np.linalg.norm(g) <= opt.tolVal * g0n:
            status, it = "gradient", it - 1
            break
        opt.dirVec[:] = -opt.apHMat @ g
        d0 = float(g @ opt.dirVec)
        if d0 >= 0:                      # not a descent direction: restart
            opt.apHMat[:] = np.eye(n)
            scaled = False
            opt.dirVec[:] = -g
            d0 = float(g @ opt.dirVec)

        def phi(a, x=x, p=opt.dirVec.copy()):
            xn = x + a * p
            fn, gn = fg(xn)
            if gn is None:
                return math.inf, math.nan, None
            return fn, float(gn @ p), (xn, gn)

        # first step: a tenth of |x| (or 0.1) until the curvature scales H
        a1 = 1.0 if scaled else min(1.0, 0.1 * max(float(np.linalg.norm(x)), 1.0)
                                    / max(float(np.linalg.norm(opt.dirVec)), 1e-300))
        res = wolfe_search(phi, f, d0, opt, a1)
        if res is None:
            status = "line search"
            break
        a, fn, (xn, gn) = res
        opt.posNxt[:], opt.grdNxt[:] = xn, gn
        opt.dspVec[:] = opt.posNxt - x
        opt.dgdVec[:] = opt.grdNxt - g
        sy = float(opt.dspVec @ opt.dgdVec)
        if sy > 1e-12 * np.linalg.norm(opt.dspVec) * np.linalg.norm(opt.dgdVec):
            if not scaled:               # eq. (6.20)
                opt.apHMat[:] = np.eye(n) * (sy / float(opt.dgdVec @ opt.dgdVec))
                scaled = True
            rho = 1.0 / sy
            Hy = opt.apHMat @ opt.dgdVec
            # eq. (6.17), expanded: H - rho (s Hy^T + Hy s^T) + (rho^2 y^T H y + rho) s s^T
            opt.apHMat[:] = (opt.apHMat - rho * (np.outer(opt.dspVec, Hy) + np.outer(Hy, opt.dspVec))
                             + (rho * rho * float(opt.dgdVec @ Hy) + rho) * np.outer(opt.dspVec, opt.dspVec))
        step = float(np.linalg.norm(opt.dspVec))
        x, f, g = opt.posNxt.copy(), fn, opt.grdNxt.copy()
        hist.append(dict(it=it, f=f, gnorm=float(np.linalg.norm(g)), step=step, alpha=a))
        if opt.verbose:
            print(f"[bfgs] it {it:3d}  f {f:.12e}  |g| {np.linalg.norm(g):.3e}  step {step:.3e}  a {a:.3g}",
                  flush=True)
        if step <= opt.tolVal * (float(np.linalg.norm(x)) + opt.tolVal):
            status = "step"
            break
    else:
        if np.linalg.norm(g) <= opt.tolVal * g0n:
            status = "gradient"
    return FitResult(x=x, f=f, g=g, iterations=it, evaluations=obj.evaluations, batches=obj.batches,
                     status=status, seconds=time.perf_counter() - t0, history=hist)


class ServerObjective:
    """Objective values of a candidate matrix from the request server: the
    candidates are submitted together and run concurrently on sessions of
    one registered mesh (sa_server_* in shape_atrophy.h)."""

    def __init__(self, prob: dict, precision: str = "fp64", workers: int = 8, server=None):
        from . import api
        self.own = server is None
        self.srv = server or api.Server(workers=workers)
        self.mesh = self.srv.add_mesh(prob, precision=precision)
        self.tauPrm = np.asarray(prob["tauPrm"], dtype=np.float64).reshape(-1)
        self.results = []

    def __call__(self, X: np.ndarray) -> np.ndarray:
        X = np.asarray(X, dtype=np.float64)
        ids = [self.srv.submit(self.mesh, X[:, j], self.tauPrm) for j in range(X.shape[1])]
        out = np.empty(X.shape[1])
        for j, rid in enumerate(ids):
            r = self.srv.wait(rid)
            self.results.append(r)
            out[j] = r["obj"] if r["rc"] == 0 and r["status"] == 0 else math.inf
        return out

    def close(self):
        if self.own:
            self.srv.close()


class SessionObjective:
    """Objective values of a candidate matrix, one candidate after another on
    one device-resident session (the reference's nCand loop, objfcn.cu)."""

    def __init__(self, prob: dict, precision: str = "fp64"):
        from . import api
        self.s = api.Session(prob, precision=precision)
        self.tauPrm = np.asarray(prob["tauPrm"], dtype=np.float64).reshape(-1)

    def __call__(self, X: np.ndarray) -> np.ndarray:
        X = np.asarray(X, dtype=np.float64)
        out = np.empty(X.shape[1])
        for j in range(X.shape[1]):
            r = self.s.run(X[:, j], self.tauPrm)
            out[j] = r["obj"] if r["status"] == 0 else math.inf
        return out

    def close(self):
        self.s.close()


def target_from_shape(prob: dict, qT: np.ndarray) -> dict:
    """Varifold target (centres, directions, measures) of the facets of
    prob's varifold layer placed at landmark positions qT (dim x V): the
    target of a fit whose true parameters are known."""
    from .workloads import facet_target
    return facet_target(np.asarray(qT).T, np.asarray(prob["elm"]["vfdVtxMat"]).T)
