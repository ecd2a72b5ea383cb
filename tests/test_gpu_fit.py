"""Fitting layer on the GPU: the central-difference gradient of the GPU
objective against the same difference of the oracle's objective, the served
and the serial candidate evaluations against each other, and a fit that
recovers the tau variables a target was generated from."""
import math

import numpy as np
import pytest

from shape_atrophy_amd import api, fit, workloads as W

pytestmark = pytest.mark.gpu


def small_ring():
    p = W.ring_problem(n_rings=6, n_per_ring=40)
    p["params"] = dict(p["params"], timeEnd=0.5)
    return p


def test_fd_gradient_matches_oracle(oracle):
    p = small_ring()
    x = p["tauVar"][:, 0] * np.array([1.03, 0.97, 1.05, 0.96])
    obj = fit.ServerObjective(p, precision="fp64", workers=8)
    try:
        f, g = fit.fd_gradient(obj, x)
    finally:
        obj.close()

    def oobj(X):
        return np.array([oracle.objfcn(dict(p, tauVar=X[:, j:j + 1].copy()))["obj"] for j in range(X.shape[1])])

    fo, go = fit.fd_gradient(oobj, x)
    print("gpu", f, g, "oracle", fo, go, flush=True)
    assert f == pytest.approx(fo, rel=1e-9)
    assert np.abs(g - go).max() <= 1e-5 * np.abs(go).max()


def test_served_and_serial_candidates_agree():
    p = small_ring()
    X = fit.fd_candidates(p["tauVar"][:, 0], np.full(4, 1e-3))
    so = fit.SessionObjective(p)
    sv = fit.ServerObjective(p, workers=4)
    try:
        a, b = so(X), sv(X)
    finally:
        so.close()
        sv.close()
    assert np.all(np.isfinite(a))
    assert np.allclose(a, b, rtol=1e-12, atol=0.0)


def test_fit_recovers_tau_variables(oracle):
    p = small_ring()
    th = p["tauVar"][:, 0].copy()
    s = api.Session(p)
    r = s.run(th, p["tauPrm"], want_traj=True)
    s.close()
    assert r["status"] == 0
    ro = oracle.objfcn(p)
    # the target shape itself agrees with the oracle's
    assert np.abs(r["lmkStk"][-1] - ro["lmkStk"][-1]).max() < 1e-8
    p["target"] = fit.target_from_shape(p, r["lmkStk"][-1])
    obj = fit.ServerObjective(p, precision="fp64", workers=8)
    try:
        x0 = th * np.array([1.08, 0.92, 1.12, 0.93])
        res = fit.bfgs(obj, x0, fit.OptData(itrMax=60, tolVal=1e-7))
    finally:
        obj.close()
    print("fit", res.status, res.iterations, res.evaluations, res.x, th, flush=True)
    assert res.f < 1e-10
    assert np.abs(res.x - th).max() < 1e-4
    assert math.isfinite(res.f)
