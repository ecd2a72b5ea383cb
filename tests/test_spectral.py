"""The spectral term of the control-point preconditioner, restated in numpy:
Rayleigh-Ritz on the Krylov basis one preconditioned CG solve leaves behind
(the recorded d_j, S d_j, r_j), as the engine's harvest does it.  Checked
against a dense eigensolver: the harvested vectors are orthonormal in the
M^-1 inner product, their Ritz values are the leading eigenvalues of M S, and
the preconditioner M + U diag(theta / lambda - 1) U^T cuts the iterations of
the next solve."""
import numpy as np
import scipy.linalg as sl


def system(n=400, seed=0):
    rng = np.random.default_rng(seed)
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    # a bulk of eigenvalues in [1, 5] plus 30 outliers up to 2000 (what the
    # elastic term adds on top of the K_cc-preconditioned kernel part)
    lam = np.concatenate([rng.uniform(1, 5, n - 30), np.geomspace(20, 2000, 30)])
    Kh = Q @ np.diag(np.sqrt(lam)) @ Q.T          # S = Kh Kh^T-like SPD factor
    Mi = np.eye(n) + 0.3 * np.diag(rng.uniform(0, 1, n))  # M^-1, SPD
    Lm = np.linalg.cholesky(Mi)
    S = Lm @ Kh @ Kh.T @ Lm.T                     # M S has eigenvalues lam
    return S, np.linalg.inv(Mi), Mi, np.sort(lam)[::-1]


def pcg(S, Mf, b, tol=1e-10, maxit=2000, rec=False):
    x = np.zeros_like(b); r = b.copy(); z = Mf(r); d = z.copy(); rz = r @ z
    D, SD, R, hist = [], [], [], [rz]
    for it in range(maxit):
        Sd = S @ d
        if rec:
            D.append(d.copy()); SD.append(Sd.copy()); R.append(r.copy())
        a = rz / (d @ Sd)
        x += a * d; r -= a * Sd
        if r @ r < tol * tol * (b @ b):
            return x, it + 1, (np.array(D).T, np.array(SD).T, np.array(R).T, np.array(hist))
        z = Mf(r); rzn = r @ z; be = rzn / rz; rz = rzn; hist.append(rz)
        d = z + be * d
    raise AssertionError("no convergence")


def harvest(D, SD, R, rz, k):
    """Ritz pairs of M S on span(Z), Z = D B, in the Lanczos scaling."""
    m = D.shape[1]
    beta = rz[1:m] / rz[:m - 1]
    B = np.eye(m)
    B[np.arange(m - 1), np.arange(1, m)] = -beta
    sc = 1.0 / np.sqrt(rz[:m])
    G = (B.T @ (D.T @ R)); G = 0.5 * (G + G.T) * np.outer(sc, sc)
    H = (B.T @ (D.T @ SD) @ B); H = 0.5 * (H + H.T) * np.outer(sc, sc)
    g, Q = np.linalg.eigh(G)
    keep = g > 1e-10 * g[-1]
    P = Q[:, keep] / np.sqrt(g[keep])
    lam, c = np.linalg.eigh(P.T @ H @ P)
    r = lam.size
    kk = min(k, 2 * r // 3)
    U = D @ (B @ (sc[:, None] * (P @ c[:, ::-1][:, :kk])))
    return lam[::-1][:kk + 1], U


def test_harvest_recovers_leading_eigenpairs():
    S, M, Mi, lam_true = system()
    n = S.shape[0]
    b = np.random.default_rng(1).standard_normal(n)
    _, it0, (D, SD, R, rz) = pcg(S, lambda r: M @ r, b, rec=True)
    lam, U = harvest(D, SD, R, rz, 30)
    kk = U.shape[1]
    # M^-1-orthonormal
    assert np.abs(U.T @ Mi @ U - np.eye(kk)).max() < 1e-8
    # Ritz values of the converged outliers equal the eigenvalues of M S
    ev = sl.eigh(S, Mi, eigvals_only=True)[::-1]
    assert np.allclose(ev[:20], lam_true[:20], rtol=1e-10)
    assert np.allclose(lam[:20], ev[:20], rtol=1e-8)
    # Ritz vectors: residual S u - lambda M^-1 u small for the leading ones
    res = np.linalg.norm(S @ U[:, :20] - (Mi @ U[:, :20]) * lam[:20], axis=0) / (lam[:20] * np.linalg.norm(Mi @ U[:, :20], axis=0))
    assert res.max() < 1e-6


def test_spectral_term_cuts_iterations():
    S, M, Mi, _ = system()
    n = S.shape[0]
    rng = np.random.default_rng(2)
    _, it0, (D, SD, R, rz) = pcg(S, lambda r: M @ r, rng.standard_normal(n), rec=True)
    lam, U = harvest(D, SD, R, rz, 40)
    kk = U.shape[1]
    theta = lam[kk]
    cf = theta / lam[:kk] - 1.0
    assert np.all(cf > -1.0) and np.all(cf < 0.0)
    Mf = lambda r: M @ r + U @ (cf * (U.T @ r))
    # the preconditioner stays symmetric positive definite
    Mp = M + U @ np.diag(cf) @ U.T
    assert np.linalg.eigvalsh(0.5 * (Mp + Mp.T)).min() > 0
    b = rng.standard_normal(n)
    x0, itp, _ = pcg(S, lambda r: M @ r, b)
    x1, its, _ = pcg(S, Mf, b)
    assert its < 0.6 * itp
    assert np.linalg.norm(x1 - x0) <= 1e-8 * np.linalg.norm(x0)


def carry(U, Mi_new):
    """The term carried to a new preconditioner (engine: spec_carry_finish):
    U G^-1/2 with G = U^T M_new^-1 U (symmetric square root)."""
    g, Q = np.linalg.eigh(U.T @ Mi_new @ U)
    return U @ (Q @ np.diag(1.0 / np.sqrt(g)) @ Q.T), g


def test_carried_term_stays_positive_definite_and_useful():
    S, M, Mi, _ = system()
    n = S.shape[0]
    rng = np.random.default_rng(5)
    _, _, (D, SD, R, rz) = pcg(S, lambda r: M @ r, rng.standard_normal(n), rec=True)
    lam, U = harvest(D, SD, R, rz, 40)
    kk = U.shape[1]
    cf = lam[kk] / lam[:kk] - 1.0
    # the operator moves a little and the new M^-1 grows by 1.6: without
    # re-orthonormalisation M_new + U diag(cf) U^T is indefinite
    # (1/1.6 + cf < 0 along the leading directions)
    E = 0.02 * rng.standard_normal((n, n))
    Mi2 = 1.6 * (Mi + E @ E.T)
    M2 = np.linalg.inv(Mi2)
    S2 = S + 0.01 * (E + E.T) @ (E + E.T)
    assert np.linalg.eigvalsh(M2 + U @ np.diag(cf) @ U.T).min() < 0
    U2, g = carry(U, Mi2)
    # within the engine's acceptance window for a carried term
    assert g.min() > 0.5 and g.max() < 2.0
    assert np.abs(U2.T @ Mi2 @ U2 - np.eye(kk)).max() < 1e-10
    Mp = M2 + U2 @ np.diag(cf) @ U2.T
    assert np.linalg.eigvalsh(0.5 * (Mp + Mp.T)).min() > 0
    # it still cuts the iterations of a solve with the new operator, and the
    # solution is the same
    b = rng.standard_normal(n)
    x0, itp, _ = pcg(S2, lambda r: M2 @ r, b)
    x1, its, _ = pcg(S2, lambda r: M2 @ r + U2 @ (cf * (U2.T @ r)), b)
    assert its < 0.7 * itp
    assert np.linalg.norm(x1 - x0) <= 1e-8 * np.linalg.norm(x0)
