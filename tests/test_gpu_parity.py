"""GPU parity: the HIP path (through the C ABI) against the FP64 oracle.

Tolerances: operators evaluated from the same inputs agree to 1e-12
relative (same arithmetic, different summation order and libm); full solves
agree to 1e-9 relative in displacement, concentration and objective against
the oracle running the same algorithm, and to 1e-7 against the oracle's
reference-exact solve path (dense Cholesky warm start and dense
reaction-diffusion solve), the gap there being set by the CG tolerance."""
import numpy as np
import pytest

from shape_atrophy_amd import api, workloads as W

pytestmark = pytest.mark.gpu


def rel(a, b):
    a = np.asarray(a, dtype=float)
    b = np.asarray(b, dtype=float)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def small_ring(**kw):
    return W.ring_problem(n_rings=6, n_per_ring=40, **kw)


def small_shell(**kw):
    return W.shell_problem(level=1, n_layers=3, time_steps=10, **kw)


PROBLEMS = {
    "ring2d": lambda: small_ring(),
    "ring2d_iso": lambda: small_ring(elaType="isotropic"),
    "shell3d": lambda: small_shell(),
    "shell3d_iso": lambda: small_shell(elaType="isotropic"),
}


@pytest.fixture(params=list(PROBLEMS))
def prob(request):
    p = PROBLEMS[request.param]()
    p["name"] = request.param
    return p


def _session(p):
    print(f"[gpu] session {p['name']} V={p['q0'].shape[1]} E={p['elm']['elmVtxMat'].shape[1]}", flush=True)
    return api.Session(p)


def test_probe_geometry_fem(prob, oracle):
    s = _session(prob)
    d, V = prob["q0"].shape
    E = prob["elm"]["elmVtxMat"].shape[1]
    geo = s.probe("geometry", prob["tauVar"][:, 0], prob["tauPrm"]).reshape(d * d + 1 + 2 * d, E)
    f1, f2 = oracle.frames(prob, prob["q0"])
    assert rel(geo[d * d + 1:d * d + 1 + d], f1) < 1e-13
    assert rel(geo[d * d + 1 + d:], f2) < 1e-13
    fem = s.probe("fem", prob["tauVar"][:, 0], prob["tauPrm"])
    Z = prob["elm"]["femVtxMat"].shape[1]
    ppd, ggd = oracle.fem_assemble(prob, prob["q0"])
    assert rel(fem[:Z], ppd) < 1e-13
    assert rel(fem[Z:], ppd + prob["params"]["timeStp"] * ggd) < 1e-13


def test_probe_elastic(prob, oracle):
    s = _session(prob)
    rng = np.random.default_rng(5)
    v = rng.standard_normal(prob["q0"].shape)
    out = s.probe("elastic", prob["tauVar"][:, 0], prob["tauPrm"], v).reshape(prob["q0"].shape)
    tau0 = oracle.tau_initial(prob)
    ref = oracle.elastic_apply(prob, prob["q0"], tau0, v)
    assert rel(out, ref) < 1e-12


def test_probe_yank(prob, oracle):
    s = _session(prob)
    d, V = prob["q0"].shape
    out = s.probe("yank", prob["tauVar"][:, 0], prob["tauPrm"])
    tau0 = oracle.tau_initial(prob)
    assert rel(out[d * V:], tau0) < 1e-14
    ya, _ = oracle.activities(tau0, prob["params"]["ynkActPrmVec"], prob["params"]["reaActPrmVec"])
    ref = oracle.external_yank(prob, prob["q0"], ya)
    assert rel(out[:d * V].reshape(d, V), ref) < 1e-12


@pytest.mark.parametrize("order", [-1, 0, 1, 2, 3, 4])
def test_probe_kmatvec(order, oracle):
    for p in (small_ring(), small_shell()):
        p["name"] = f"kmatvec{p['dim']}d_o{order}"
        p["params"]["knlOrder"] = order
        s = _session(p)
        rng = np.random.default_rng(7)
        x = rng.standard_normal(p["q0"].shape)
        out = s.probe("kmatvec", p["tauVar"][:, 0], p["tauPrm"], x).reshape(p["q0"].shape)
        prm = p["params"]
        ref = oracle.kernel_matvec(p["dim"], order, prm["knlWidth"], prm["knlEps"], p["q0"], x)
        assert rel(out, ref) < 1e-12, order


@pytest.mark.parametrize("order", [-1, 0, 1, 2, 3, 4])
def test_probe_kmatvec_matrix_free(order, oracle, monkeypatch):
    """The all-landmark product without the stored K (kmatvec_k: the form
    meshes above about 77,000 landmarks take; SA_MATRIX_FREE=1 forces it
    here) against the oracle, 2D and 3D, every kernel order."""
    monkeypatch.setenv("SA_MATRIX_FREE", "1")
    for p in (small_ring(), small_shell()):
        p["name"] = f"kmatvec{p['dim']}d_o{order}_matrix_free"
        p["params"]["knlOrder"] = order
        s = _session(p)
        rng = np.random.default_rng(7)
        x = rng.standard_normal(p["q0"].shape)
        out = s.probe("kmatvec", p["tauVar"][:, 0], p["tauPrm"], x).reshape(p["q0"].shape)
        s.close()
        prm = p["params"]
        ref = oracle.kernel_matvec(p["dim"], order, prm["knlWidth"], prm["knlEps"], p["q0"], x)
        assert rel(out, ref) < 1e-12, order


@pytest.mark.parametrize("which", ["ring2d", "shell3d"])
def test_full_solve_matrix_free_all_landmarks(which, oracle, monkeypatch):
    """A full all-landmark solve with K applied matrix-free (kmatvec_k in
    every CG iteration and for the warm start) against the oracle, at the
    full-solve tolerance."""
    monkeypatch.setenv("SA_MATRIX_FREE", "1")
    p = PROBLEMS[which]()
    p["name"] = which + "_matrix_free"
    s = _session(p)
    r = s.run(p["tauVar"][:, 0], p["tauPrm"], want_traj=True)
    o = oracle.objfcn(p, pcg_init=1, rd_solve=1)
    assert r["status"] == 0 and o["status"] == 0
    disp = np.linalg.norm(o["lmkStk"][-1] - o["lmkStk"][0])
    assert np.linalg.norm(r["lmkStk"] - o["lmkStk"]) / (disp * np.sqrt(len(o["lmkStk"]))) < 1e-9
    assert rel(r["tauMat"], o["tauMat"]) < 1e-9
    assert abs(r["obj"] - o["obj"]) <= 1e-9 * abs(o["obj"])


def test_full_solve_all_landmarks_staging_in_passes(oracle, monkeypatch):
    """The product that takes the all-landmark CG step stages its column
    slice 512 columns per pass.  kk=16 leaves 4 slices of 803 columns on
    this 3,210-landmark shell, so every launch takes two passes (and the
    first row panel's x and alpha updates both).  Full solve against the
    oracle, at the full-solve tolerance."""
    monkeypatch.setenv("SA_KS_TARGET", "kk=16")
    p = W.shell_problem(level=3, n_layers=5, time_steps=3)
    p["name"] = "shell3d_level3_4_slices"
    s = _session(p)
    r = s.run(p["tauVar"][:, 0], p["tauPrm"], want_traj=True)
    s.close()
    o = oracle.objfcn(p, pcg_init=1, rd_solve=1)
    assert r["status"] == 0 and o["status"] == 0
    disp = np.linalg.norm(o["lmkStk"][-1] - o["lmkStk"][0])
    assert np.linalg.norm(r["lmkStk"] - o["lmkStk"]) / (disp * np.sqrt(len(o["lmkStk"]))) < 1e-9
    assert rel(r["tauMat"], o["tauMat"]) < 1e-9
    assert abs(r["obj"] - o["obj"]) <= 1e-9 * abs(o["obj"])


def test_probe_varifold(prob, oracle):
    s = _session(prob)
    rng = np.random.default_rng(9)
    q = prob["q0"] + 0.01 * rng.standard_normal(prob["q0"].shape)
    out = s.probe("varifold", prob["tauVar"][:, 0], prob["tauPrm"], q)
    val, grad = oracle.varifold(prob, q, with_grad=True)
    assert abs(out[0] - val) <= 1e-12 * abs(val)
    assert rel(out[1:].reshape(q.shape), grad) < 1e-11


@pytest.mark.parametrize("ck,dk", [("C", "L"), ("G", "O"), ("C", "U")])
def test_probe_varifold_kernel_types(ck, dk, oracle):
    """The Cauchy centre kernel and the linear, oriented and unoriented
    direction kernels (varifold.cu's kernel switches), value and gradient."""
    for p in (small_ring(), small_shell()):
        p["vfd"] = dict(p["vfd"], cenKnlType=ck, dirKnlType=dk, dirKnlWidth=0.7)
        p["name"] = f"varifold{p['dim']}d_{ck}{dk}"
        s = _session(p)
        rng = np.random.default_rng(11)
        q = p["q0"] + 0.01 * rng.standard_normal(p["q0"].shape)
        out = s.probe("varifold", p["tauVar"][:, 0], p["tauPrm"], q)
        val, grad = oracle.varifold(p, q, with_grad=True)
        assert abs(out[0] - val) <= 1e-12 * abs(val), (ck, dk, p["dim"])
        assert rel(out[1:].reshape(q.shape), grad) < 1e-11, (ck, dk, p["dim"])


@pytest.mark.parametrize("ck,dk", [("G", "B"), ("C", "L"), ("G", "O"), ("C", "U")])
def test_varifold_value_on_mfma(ck, dk, oracle, monkeypatch):
    """The varifold value with its pair contractions on FP64 matrix cores
    (vfd_rows_mfma_k; SA_VFD_MFMA=1 turns it on in an FP64 session as the
    mixed form does by default) against the oracle, on facet counts that
    leave partial 16-row strips, partial 16-column tiles and several
    256-column chunks.  Tolerance 1e-11 relative: the expanded distance
    |c_i - o|^2 + |c_j - o|^2 - 2 (c_i - o).(c_j - o) (o the strip's first
    centre) loses the bits of |c - o|^2 / |c_i - c_j|^2; its FP64 restatement
    (test_oracle.py::test_expanded_distance_varifold_accuracy) stays within
    1.5e-13 of the oracle on these shapes."""
    monkeypatch.setenv("SA_VFD_MFMA", "1")
    shapes = (small_ring(), small_shell(), W.ring_problem(n_rings=4, n_per_ring=300),
              W.shell_problem(level=3, n_layers=2, time_steps=2))
    for p in shapes:
        p["vfd"] = dict(p["vfd"], cenKnlType=ck, dirKnlType=dk,
                        dirKnlWidth=p["vfd"]["dirKnlWidth"] if dk == "B" else 0.7)
        p["name"] = f"varifold{p['dim']}d_mfma_{ck}{dk}"
        s = _session(p)
        rng = np.random.default_rng(9)
        q = p["q0"] + 0.01 * rng.standard_normal(p["q0"].shape)
        out = s.probe("varifold", p["tauVar"][:, 0], p["tauPrm"], q)
        s.close()
        val, grad = oracle.varifold(p, q, with_grad=True)
        print(f"[gpu] {p['name']}: value rel {abs(out[0] - val) / abs(val):.2e}", flush=True)
        assert abs(out[0] - val) <= 1e-11 * abs(val), (ck, dk, p["dim"])
        # the gradient stays on the VALU kernel
        assert rel(out[1:].reshape(q.shape), grad) < 1e-11, (ck, dk, p["dim"])


def test_full_solve_matches_oracle(prob, oracle):
    s = _session(prob)
    r = s.run(prob["tauVar"][:, 0], prob["tauPrm"], want_traj=True)
    o = oracle.objfcn(prob, pcg_init=1, rd_solve=1)
    assert r["status"] == 0 and o["status"] == 0
    disp = np.linalg.norm(o["lmkStk"][-1] - o["lmkStk"][0])
    print(prob["name"], "pcg iters gpu", r["pcg_iters"], "oracle", int(o["pcg_iters"].sum()), "disp", disp)
    assert np.linalg.norm(r["lmkStk"] - o["lmkStk"]) / (disp * np.sqrt(len(o["lmkStk"]))) < 1e-9
    assert rel(r["tauMat"], o["tauMat"]) < 1e-9
    assert abs(r["obj"] - o["obj"]) <= 1e-9 * abs(o["obj"])


def test_full_solve_matches_reference_exact(oracle):
    p = small_shell()
    p["name"] = "shell3d_exact"
    s = _session(p)
    r = s.run(p["tauVar"][:, 0], p["tauPrm"], want_traj=True)
    o = oracle.objfcn(p, pcg_init=0, rd_solve=0)
    disp = np.linalg.norm(o["lmkStk"][-1] - o["lmkStk"][0])
    assert np.linalg.norm(r["lmkStk"][-1] - o["lmkStk"][-1]) / disp < 1e-7
    assert abs(r["obj"] - o["obj"]) <= 1e-7 * abs(o["obj"])


def test_c_abi_compute_objfcn_and_deformation(oracle):
    p = small_ring()
    args = api.problem_args(p)
    # two candidates: the seeded one and a shifted blob
    tv = np.concatenate([p["tauVar"], p["tauVar"] + np.array([[0.05], [0.0], [0.0], [0.0]])], axis=1)
    args[0] = tv
    objVec = api.computeObjfcn(*args)
    assert objVec.shape == (1, 2)
    for c in range(2):
        o = oracle.objfcn(p, cand=c) if c == 0 else oracle.objfcn(dict(p, tauVar=tv), cand=1)
        assert abs(objVec[0, c] - o["obj"]) <= 1e-9 * abs(o["obj"])
    dargs = [a for i, a in enumerate(args) if i not in (5, 6)]
    lmkStk, tauMat = api.deformation(*dargs)
    o = oracle.objfcn(dict(p, tauVar=tv), cand=1, with_target=False)
    T = o["tauMat"].shape[0]
    assert lmkStk.shape == (2, p["q0"].shape[1], T) and tauMat.shape == (p["q0"].shape[1], T)
    assert rel(np.transpose(lmkStk, (2, 0, 1)), o["lmkStk"]) < 1e-9
    assert rel(tauMat.T, o["tauMat"]) < 1e-9


@pytest.mark.parametrize("which", ["shell3d", "ring2d"])
def test_control_point_solve_matches_oracle(which, oracle):
    p = small_shell() if which == "shell3d" else small_ring()
    V = p["q0"].shape[1]
    p["ctlIdx"] = np.sort(np.random.default_rng(3).choice(V, V // 3, replace=False)).astype(np.int32)
    p["name"] = which + "_ctl"
    s = _session(p)
    r = s.run(p["tauVar"][:, 0], p["tauPrm"], want_traj=True)
    o = oracle.objfcn(p)
    assert r["status"] == 0 and o["status"] == 0
    disp = np.linalg.norm(o["lmkStk"][-1] - o["lmkStk"][0])
    print(p["name"], "pcg iters gpu", r["pcg_iters"], "oracle", int(o["pcg_iters"].sum()))
    # the control-point system inherits the conditioning of K_cc (knlEps =
    # 1e-10): rounding-level differences change the CG iterate path (the
    # iteration counts differ by ~1%), so the converged solutions agree to the
    # CG tolerance amplified by that conditioning, ~1e-9, not to rounding
    assert np.linalg.norm(r["lmkStk"][-1] - o["lmkStk"][-1]) / disp < 1e-8
    assert rel(r["tauMat"], o["tauMat"]) < 1e-8
    assert abs(r["obj"] - o["obj"]) <= 1e-8 * abs(o["obj"])


@pytest.mark.parametrize("which", ["shell3d", "ring2d"])
def test_mixed_precision_control_points(which, oracle):
    """FP32 stored kernel matrices inside FP64 iterative refinement: the
    refined velocity meets the same FP64 residual test, so the trajectory
    and objective agree with the FP64 oracle to the CG tolerance (<= 1e-6 is
    the stated bound for the objective; observed ~1e-9)."""
    p = small_shell() if which == "shell3d" else small_ring()
    V = p["q0"].shape[1]
    p["ctlIdx"] = np.sort(np.random.default_rng(3).choice(V, V // 3, replace=False)).astype(np.int32)
    p["name"] = which + "_mixed"
    print(f"[gpu] session {p['name']} V={V}", flush=True)
    s = api.Session(p, precision="mixed")
    r = s.run(p["tauVar"][:, 0], p["tauPrm"], want_traj=True)
    o = oracle.objfcn(p)
    assert r["status"] == 0 and o["status"] == 0
    print(p["name"], "iters", r["pcg_iters"], "sweeps", r["refine"], "fp64 steps", r["fp64_steps"],
          "oracle iters", int(o["pcg_iters"].sum()))
    # every step ran FP64 refinement sweeps, or (ill-conditioned K_cc) was finished in FP64
    assert r["refine"] > 0 and r["refine"] + r["fp64_steps"] >= r["steps"]
    disp = np.linalg.norm(o["lmkStk"][-1] - o["lmkStk"][0])
    assert np.linalg.norm(r["lmkStk"][-1] - o["lmkStk"][-1]) / disp < 1e-7
    assert rel(r["tauMat"], o["tauMat"]) < 1e-7
    assert abs(r["obj"] - o["obj"]) <= 1e-7 * abs(o["obj"])


def test_mixed_residual_products_split_and_matrix_free(oracle, monkeypatch):
    """The mixed form's refinement residuals use the FP32 stored matrices plus
    their FP32 remainders (each entry exact to ~2^-48); with SA_MATRIX_FREE=1
    they are matrix-free FP64 kernel products instead.  Both meet the same
    FP64 residual test, so the two runs agree with each other and with the
    oracle to the CG tolerance."""
    p = small_shell()
    V = p["q0"].shape[1]
    p["ctlIdx"] = np.sort(np.random.default_rng(3).choice(V, V // 3, replace=False)).astype(np.int32)
    p["name"] = "shell3d_mixed_split"
    print(f"[gpu] session {p['name']} V={V}", flush=True)
    rs = api.Session(p, precision="mixed").run(p["tauVar"][:, 0], p["tauPrm"], want_traj=True)
    monkeypatch.setenv("SA_MATRIX_FREE", "1")
    print(f"[gpu] session {p['name']} matrix-free V={V}", flush=True)
    rm = api.Session(p, precision="mixed").run(p["tauVar"][:, 0], p["tauPrm"], want_traj=True)
    o = oracle.objfcn(p)
    assert rs["status"] == 0 and rm["status"] == 0 and o["status"] == 0
    disp = np.linalg.norm(o["lmkStk"][-1] - o["lmkStk"][0])
    assert np.linalg.norm(rs["lmkStk"][-1] - rm["lmkStk"][-1]) / disp < 1e-8
    assert abs(rs["obj"] - rm["obj"]) <= 1e-8 * abs(o["obj"])
    for r in (rs, rm):
        assert np.linalg.norm(r["lmkStk"][-1] - o["lmkStk"][-1]) / disp < 1e-7
        assert abs(r["obj"] - o["obj"]) <= 1e-7 * abs(o["obj"])


def test_mixed_precision_falls_back_to_fp64():
    """2D Matern nu=4 on 10% of the landmarks: K_cc is so ill conditioned that
    FP32 kernel products cannot drive the FP64 residual down; such steps are
    finished by the FP64 form, so the mixed form still converges to the FP64
    solution (objective within 1e-7).  The inner FP32 solve that stagnates is
    stopped once its |r|^2 has not halved over 100 iterations (after 200),
    not after its first-step budget of 1,250: the mixed run costs at most
    400 CG iterations more than the FP64 one (223 measured)."""
    p = W.ring_problem(n_rings=8, n_per_ring=150)
    p["params"] = dict(p["params"], timeStp=0.1)
    p["timeNum"] = 6
    V = p["q0"].shape[1]
    p["ctlIdx"] = np.sort(np.random.default_rng(4).choice(V, V // 10, replace=False)).astype(np.int32)
    print(f"[gpu] session ring2d_fallback V={V}", flush=True)
    rm = api.Session(p, precision="mixed").run(p["tauVar"][:, 0], p["tauPrm"])
    r64 = api.Session(p, precision="fp64").run(p["tauVar"][:, 0], p["tauPrm"])
    print("fallback steps", rm["fp64_steps"], "sweeps", rm["refine"], "iters", rm["pcg_iters"], r64["pcg_iters"])
    assert rm["status"] == 0 and r64["status"] == 0
    assert rm["fp64_steps"] > 0
    assert abs(rm["obj"] - r64["obj"]) <= 1e-7 * abs(r64["obj"])
    assert rm["pcg_iters"] <= r64["pcg_iters"] + 400


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_spectral_preconditioner(precision, monkeypatch):
    """The spectral term of the control-point preconditioner (Ritz vectors of
    M S harvested after each refresh of M) changes the CG iterate path but not
    the converged solution: the objective agrees with the plain
    preconditioner's to the CG tolerance, in fewer iterations."""
    p = W.shell_problem(level=3, n_layers=5, time_steps=12, n_ctl=321)
    print(f"[gpu] spectral {precision} V={p['q0'].shape[1]}", flush=True)
    monkeypatch.setenv("SA_SPECTRAL", "0")
    plain = api.Session(p, precision=precision).run(p["tauVar"][:, 0], p["tauPrm"])
    monkeypatch.setenv("SA_SPECTRAL", "1")
    spec = api.Session(p, precision=precision).run(p["tauVar"][:, 0], p["tauPrm"])
    print("iterations plain", plain["pcg_iters"], "spectral", spec["pcg_iters"])
    assert plain["status"] == 0 and spec["status"] == 0
    assert abs(spec["obj"] - plain["obj"]) <= 1e-9 * abs(plain["obj"])
    assert spec["pcg_iters"] < 0.9 * plain["pcg_iters"]


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_control_point_budget_spent_ends_the_solve(precision):
    """A control-point solve whose CG budget (cgItrMax) runs out ends with a
    failed status (objective DBL_MAX, failReason budget) instead of waiting
    for a convergence that never comes."""
    p = small_shell()
    V = p["q0"].shape[1]
    p["ctlIdx"] = np.sort(np.random.default_rng(3).choice(V, V // 3, replace=False)).astype(np.int32)
    p["params"] = dict(p["params"], cgItrMax=3)
    print(f"[gpu] budget {precision} V={V}", flush=True)
    r = api.Session(p, precision=precision).run(p["tauVar"][:, 0], p["tauPrm"])
    print("status", r["status"], "fail", r["fail_reason"], "steps", r["steps"], "iters", r["pcg_iters"])
    assert r["status"] != 0
    assert r["obj"] == np.finfo(np.float64).max
    assert r["fail_reason"] == 2   # SA_FAIL_BUDGET


def test_matrix_free_control_point_form(oracle, monkeypatch):
    """The FP64 control-point form with matrix-free kernel products (the form
    used above the 48 GB store cap; SA_MATRIX_FREE=1 forces it here) agrees
    with the oracle to the CG tolerance."""
    p = small_shell()
    V = p["q0"].shape[1]
    p["ctlIdx"] = np.sort(np.random.default_rng(3).choice(V, V // 3, replace=False)).astype(np.int32)
    p["name"] = "shell3d_ctl_matrix_free"
    monkeypatch.setenv("SA_MATRIX_FREE", "1")
    r = _session(p).run(p["tauVar"][:, 0], p["tauPrm"], want_traj=True)
    o = oracle.objfcn(p)
    assert r["status"] == 0 and o["status"] == 0
    disp = np.linalg.norm(o["lmkStk"][-1] - o["lmkStk"][0])
    assert np.linalg.norm(r["lmkStk"][-1] - o["lmkStk"][-1]) / disp < 1e-8
    assert abs(r["obj"] - o["obj"]) <= 1e-8 * abs(o["obj"])


@pytest.mark.parametrize("which", ["ring2d", "shell3d", "shell3d_ctl"])
def test_boundary_yank(which, oracle):
    """sa_params.bdrYank = 1 adds the boundary yank the reference keeps
    commented out: the yank vector and a full solve against the oracle's
    same term."""
    p = small_ring() if which == "ring2d" else small_shell()
    if which == "shell3d_ctl":
        V = p["q0"].shape[1]
        p["ctlIdx"] = np.sort(np.random.default_rng(3).choice(V, V // 3, replace=False)).astype(np.int32)
    p["params"] = dict(p["params"], bdrYank=1)
    p["name"] = which + "_bdr_yank"
    s = _session(p)
    d, V = p["q0"].shape
    out = s.probe("yank", p["tauVar"][:, 0], p["tauPrm"])
    ya, _ = oracle.activities(oracle.tau_initial(p), p["params"]["ynkActPrmVec"], p["params"]["reaActPrmVec"])
    assert rel(out[:d * V].reshape(d, V), oracle.external_yank(p, p["q0"], ya)) < 1e-12
    r = s.run(p["tauVar"][:, 0], p["tauPrm"], want_traj=True)
    o = oracle.objfcn(p)
    assert r["status"] == 0 and o["status"] == 0
    disp = np.linalg.norm(o["lmkStk"][-1] - o["lmkStk"][0])
    assert np.linalg.norm(r["lmkStk"][-1] - o["lmkStk"][-1]) / disp < 1e-8
    assert abs(r["obj"] - o["obj"]) <= 1e-8 * abs(o["obj"])


# --------------------------------------------------------------------------
# bottom penalty (btmWgt != 0): 2D applyPenaltyOperator.cu:4-79 (segment
# length / 2 per vertex), 3D applyPenaltyOperator.cu:4-85 (xy-projected area
# / 3), added to the elastic operator inside the velocity solve
# (ipcgSecondMatrixMultiplication.cu:8-24)

def pen_strip():
    p = W.strip_problem()
    p["params"] = dict(p["params"], btmWgt=10.0)
    p["timeNum"] = 3
    return p


PENALTY = {
    "strip2d": pen_strip,
    "ring2d": lambda: small_ring(),
    "slab3d": lambda: W.slab_problem(n=7, n_layers=3, tilt=0.3, time_steps=6),
    "shell3d": lambda: small_shell(),
}


@pytest.mark.parametrize("which", list(PENALTY))
def test_probe_penalty(which, oracle):
    p = PENALTY[which]()
    p["name"] = which + "_penalty"
    s = _session(p)
    v = np.random.default_rng(12).standard_normal(p["q0"].shape)
    out = s.probe("penalty", p["tauVar"][:, 0], p["tauPrm"], v).reshape(p["q0"].shape)
    ref = oracle.penalty_apply(p, p["q0"], v)
    assert rel(out, ref) < 1e-13


@pytest.mark.parametrize("which,ctl", [("strip2d", False), ("slab3d", False), ("shell3d", False),
                                       ("slab3d", True), ("ring2d", True)])
def test_full_solve_with_penalty(which, ctl, oracle):
    """Full solves with btmWgt != 0 against the oracle: the all-landmark
    form to 1e-9 as the other full solves, the control-point form to 1e-8."""
    p = PENALTY[which]()
    p["params"] = dict(p["params"], btmWgt=p["params"]["btmWgt"] or 5.0)
    V = p["q0"].shape[1]
    if ctl:
        p["ctlIdx"] = np.sort(np.random.default_rng(3).choice(V, V // 3, replace=False)).astype(np.int32)
    p["name"] = which + ("_ctl" if ctl else "") + "_btmwgt"
    r = _session(p).run(p["tauVar"][:, 0], p["tauPrm"], want_traj=True)
    o = oracle.objfcn(p)
    assert r["status"] == 0 and o["status"] == 0
    disp = np.linalg.norm(o["lmkStk"][-1] - o["lmkStk"][0])
    tol = 1e-8 if ctl else 1e-9
    print(p["name"], "pcg iters gpu", r["pcg_iters"], "oracle", int(o["pcg_iters"].sum()))
    assert np.linalg.norm(r["lmkStk"][-1] - o["lmkStk"][-1]) / disp < tol
    assert rel(r["tauMat"], o["tauMat"]) < tol
    assert abs(r["obj"] - o["obj"]) <= tol * abs(o["obj"])
    # the penalty changed the solution (the same solve without it, on the GPU)
    q = dict(p, params=dict(p["params"], btmWgt=0.0), name=p["name"] + "0")
    r0 = _session(q).run(p["tauVar"][:, 0], p["tauPrm"], want_traj=True)
    assert r0["status"] == 0
    assert np.linalg.norm(r0["lmkStk"][-1] - o["lmkStk"][-1]) / disp > 1e-4


# --------------------------------------------------------------------------
# the fused operator apply (op_fused_k, one launch, forces in LDS) against the
# two-launch apply (elastic_elem_k + gather_op_k, SA_OP_FUSED=0): the same
# forces summed in the same order, Q^{-1} recomputed by the same function

@pytest.mark.parametrize("which", list(PROBLEMS) + ["slab3d_btm"])
def test_fused_operator_apply_matches_two_launch(which, monkeypatch):
    p = W.slab_problem(n=7, n_layers=3, tilt=0.3, time_steps=6, btmWgt=5.0) if which == "slab3d_btm" \
        else PROBLEMS[which]()
    p["name"] = which + "_fused"
    v = np.random.default_rng(5).standard_normal(p["q0"].shape)
    V = p["q0"].shape[1]
    out = {}
    for fused in ("1", "0"):
        monkeypatch.setenv("SA_OP_FUSED", fused)
        s = _session(p)
        e = s.probe("elastic", p["tauVar"][:, 0], p["tauPrm"], v)
        r = s.run(p["tauVar"][:, 0], p["tauPrm"])
        s.close()
        pc = dict(p, ctlIdx=np.sort(np.random.default_rng(3).choice(V, V // 3, replace=False)).astype(np.int32))
        sc = api.Session(pc, precision="mixed")
        rc = sc.run(p["tauVar"][:, 0], p["tauPrm"])
        sc.close()
        assert r["status"] == 0 and rc["status"] == 0
        out[fused] = (e, r["obj"], rc["obj"])
    a, b = out["1"], out["0"]
    print(f"[gpu] {which}: fused vs two-launch apply {rel(a[0], b[0]):.1e}, objectives {a[1]!r} {b[1]!r}, "
          f"control points {a[2]!r} {b[2]!r}", flush=True)
    assert rel(a[0], b[0]) < 1e-14
    assert abs(a[1] - b[1]) <= 1e-11 * abs(b[1]) and abs(a[2] - b[2]) <= 1e-9 * abs(b[2])


# --------------------------------------------------------------------------
# control-point kernel products K_xc u and K_cx y as the inner solve forms
# them: stored FP64 / matrix-free FP64 (plain form) to 1e-13, stored FP32
# (2D mixed form) and matrix-free FP32 on the matrix cores (3D mixed form,
# kx_mfma_k) to 5e-6 of the FP64 products (FP32 kernel values, FP32 inputs,
# FP32 sums over 256-column chunks added in FP64)

def dense_kernel(oracle, prob, a, b):
    prm = prob["params"]
    d2 = ((a[:, :, None] - b[:, None, :]) ** 2).sum(0)
    return oracle.kernel_values(prob["dim"], prm["knlOrder"], prm["knlWidth"], d2.ravel()).reshape(d2.shape)


PRODUCT_CASES = [
    ("shell3d", "mixed", {"SA_KX_MFMA": "1"}, 5e-6),
    ("shell3d_gauss", "mixed", {"SA_KX_MFMA": "1"}, 5e-6),
    ("shell3d", "mixed", {"SA_KX_MFMA": "0"}, 5e-6),
    ("ring2d", "mixed", {}, 5e-6),
    ("shell3d", "fp64", {}, 1e-13),
    ("shell3d", "fp64", {"SA_MATRIX_FREE": "1"}, 1e-13),
]


@pytest.mark.parametrize("which,precision,env,tol", PRODUCT_CASES)
def test_probe_control_point_products(which, precision, env, tol, oracle, monkeypatch):
    for k, v in env.items():
        monkeypatch.setenv(k, v)
    if which.startswith("shell3d"):
        p = W.shell_problem(level=3, n_layers=5, time_steps=2)
        if which == "shell3d_gauss":
            p["params"] = dict(p["params"], knlOrder=-1, knlWidth=0.15)
    else:
        p = W.ring_problem(n_rings=10, n_per_ring=120)
    V = p["q0"].shape[1]
    p["ctlIdx"] = np.sort(np.random.default_rng(21).choice(V, 700, replace=False)).astype(np.int32)
    p["name"] = f"{which}_{precision}_products_{env}"
    print(f"[gpu] session {p['name']} V={V} C=700", flush=True)
    s = api.Session(p, precision=precision)
    rng = np.random.default_rng(13)
    u = rng.standard_normal((p["dim"], 700))
    y = rng.standard_normal((p["dim"], V))
    kx = s.probe("kxc", p["tauVar"][:, 0], p["tauPrm"], u).reshape(p["dim"], V)
    kc = s.probe("kcx", p["tauVar"][:, 0], p["tauPrm"], y).reshape(p["dim"], 700)
    s.close()
    K = dense_kernel(oracle, p, p["q0"], p["q0"][:, p["ctlIdx"]])
    e1, e2 = rel(kx, u @ K.T), rel(kc, y @ K)
    print(f"[gpu] {p['name']}: K_xc u {e1:.2e}, K_cx y {e2:.2e}", flush=True)
    assert e1 < tol and e2 < tol


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_tile_targets_change_only_the_summation_order(precision, monkeypatch):
    """The block target of the stored products (the tile table's entry,
    forced here with SA_KS_TARGET) sets their column slices, so it changes
    only the order of the FP64 partial sums: products agree to 1e-12 across
    targets, a solve's objective to 1e-8; the product timer returns a
    positive time for every product."""
    p = W.shell_problem(level=3, n_layers=5, time_steps=4)
    V = p["q0"].shape[1]
    p["ctlIdx"] = np.sort(np.random.default_rng(21).choice(V, 700, replace=False)).astype(np.int32)
    rng = np.random.default_rng(13)
    u, y = rng.standard_normal((3, 700)), rng.standard_normal((3, V))
    tv, tp = p["tauVar"][:, 0], p["tauPrm"]
    out = {}
    for tgt in ("64", "16384"):
        monkeypatch.setenv("SA_KS_TARGET", tgt)
        print(f"[gpu] tile target {tgt}, {precision}, V={V} C=700", flush=True)
        s = api.Session(p, precision=precision)
        kx, kc = s.probe("kxc", tv, tp, u), s.probe("kcx", tv, tp, y)
        us = {w: s.time_product(w, 5) for w in ("kxc", "kcx", "precond")}
        r = s.run(tv, tp)
        s.close()
        print(f"[gpu] tile target {tgt}: {us} us, obj {r['obj']:.15e}", flush=True)
        assert r["status"] == 0
        assert all(np.isfinite(v) and v > 0 for v in us.values()), us
        out[tgt] = (kx, kc, r["obj"])
    a, b = out["64"], out["16384"]
    assert rel(a[0], b[0]) < 1e-12 and rel(a[1], b[1]) < 1e-12
    assert abs(a[2] - b[2]) <= 1e-8 * abs(b[2])


def test_tile_targets_all_landmark_product(oracle, monkeypatch):
    """The same for the stored all-landmark product: (K + eps I) x at two
    block targets against the oracle's (1e-12), and its timer."""
    p = small_shell()
    V = p["q0"].shape[1]
    x = np.random.default_rng(4).standard_normal(p["q0"].shape)
    prm = p["params"]
    ref = oracle.kernel_matvec(3, prm["knlOrder"], prm["knlWidth"], prm["knlEps"], p["q0"], x)
    for tgt in ("16", "4096"):
        monkeypatch.setenv("SA_KS_TARGET", f"kk={tgt}")
        print(f"[gpu] all-landmark product, target {tgt}, V={V}", flush=True)
        s = api.Session(p)
        got = s.probe("kmatvec", p["tauVar"][:, 0], p["tauPrm"], x).reshape(x.shape)
        us = s.time_product("kmatvec", 5)
        s.close()
        assert rel(got, ref) < 1e-12 and np.isfinite(us) and us > 0


# --------------------------------------------------------------------------
# the K_cc factorisation of the control-point preconditioner (chol_*_k):
# backward error |L L^T - A| / |A| (fixed by the mathematics whatever the
# conditioning), the factor against numpy's and against rocsolver_dpotrf run
# alone; sizes around the 64-column block; an indefinite matrix reported

CHOL_CASES = [("shell3d", 700), ("shell3d", 50), ("shell3d", 64), ("shell3d", 65), ("shell3d", 128),
              ("ring2d", 300)]


def chol_problem(which, C):
    p = W.shell_problem(level=3, n_layers=5, time_steps=2) if which == "shell3d" \
        else W.ring_problem(n_rings=10, n_per_ring=120)
    V = p["q0"].shape[1]
    p["ctlIdx"] = np.sort(np.random.default_rng(C).choice(V, C, replace=False)).astype(np.int32)
    return p


@pytest.mark.parametrize("which,C", CHOL_CASES)
def test_kcc_cholesky_factor(which, C, oracle, monkeypatch):
    p = chol_problem(which, C)
    qc = p["q0"][:, p["ctlIdx"]]
    A = dense_kernel(oracle, p, qc, qc) + p["params"]["knlEps"] * np.eye(C)
    tv, tp = p["tauVar"][:, 0], p["tauPrm"]
    print(f"[gpu] K_cc factor {which} C={C}", flush=True)
    got = {}
    for lib in ("", "rocsolver"):
        monkeypatch.setenv("SA_CHOL", lib)
        s = api.Session(p)
        got[lib] = s.probe("kccchol", tv, tp).reshape(C, C).T   # column-major out
        s.close()
    L = got[""]
    assert np.all(np.triu(L, 1) == 0.0) and np.all(np.diag(L) > 0)
    back = np.linalg.norm(L @ L.T - A) / np.linalg.norm(A)
    Lnp = np.linalg.cholesky(A)
    fwd, lib = rel(L, Lnp), rel(L, got["rocsolver"])
    cond = np.linalg.cond(A)
    print(f"[gpu] K_cc factor {which} C={C}: backward {back:.1e}, vs numpy {fwd:.1e}, vs rocsolver {lib:.1e}, "
          f"cond {cond:.1e}", flush=True)
    assert back < 1e-14
    # forward error: at most about cond x unit roundoff
    assert fwd < max(1e-12, 1e-15 * cond) and lib < max(1e-12, 1e-15 * cond)


def test_kcc_cholesky_reports_indefinite():
    """K_cc - I (knlEps = -1: a zero diagonal) fails at its first column: the
    probe raises with the column, a solve reports SA_FAIL_CHOLESKY."""
    p = chol_problem("shell3d", 200)
    p["params"] = dict(p["params"], knlEps=-1.0)
    tv, tp = p["tauVar"][:, 0], p["tauPrm"]
    s = api.Session(p)
    with pytest.raises(api.SaError, match="column 1"):
        s.probe("kccchol", tv, tp)
    r = s.run(tv, tp)
    s.close()
    assert r["status"] == 1 and r["fail_reason"] == 3, (r["status"], r["fail_reason"])


# --------------------------------------------------------------------------
# Jacobi preconditioning of the control-point CG (sa_params.velPrec): the
# diagonal of the control-point system against the oracle's, formed there by
# applying S to every unit vector (1e-12), and full solves against the
# oracle's default-preconditioned solve (the preconditioner changes the
# iterate path, not the solution: 1e-8 FP64, 1e-7 mixed)

JACOBI = {
    "shell3d": lambda: small_shell(),
    "shell3d_iso": lambda: small_shell(elaType="isotropic"),
    "ring2d": lambda: small_ring(),
    "slab3d_btm": lambda: W.slab_problem(n=7, n_layers=3, tilt=0.3, time_steps=6, btmWgt=5.0),
}


@pytest.mark.parametrize("which", list(JACOBI))
def test_probe_jacobi_diagonal(which, oracle):
    p = JACOBI[which]()
    V = p["q0"].shape[1]
    p["ctlIdx"] = np.sort(np.random.default_rng(3).choice(V, V // 3, replace=False)).astype(np.int32)
    p["params"] = dict(p["params"], velPrec="jacobi")
    p["name"] = which + "_jacobi"
    s = _session(p)
    out = s.probe("jacobi", p["tauVar"][:, 0], p["tauPrm"]).reshape(p["dim"], -1)
    s.close()
    ref = oracle.jacobi_diag(p, p["q0"], oracle.tau_initial(p))
    assert rel(out, ref) < 1e-12


# (not the 2D ring: with its Matern nu = 4 kernel the control-point system
# is too ill conditioned for a diagonal preconditioner, and the oracle's
# Jacobi CG does not converge within the 10,000-iteration budget either)
@pytest.mark.parametrize("which,precision", [("shell3d", "fp64"), ("shell3d", "mixed"), ("slab3d_btm", "fp64"),
                                             ("shell3d_iso", "fp64")])
def test_jacobi_solve_matches_oracle(which, precision, oracle):
    p = JACOBI[which]()
    V = p["q0"].shape[1]
    p["ctlIdx"] = np.sort(np.random.default_rng(3).choice(V, V // 3, replace=False)).astype(np.int32)
    o = oracle.objfcn(p)
    p["params"] = dict(p["params"], velPrec="jacobi")
    print(f"[gpu] jacobi {which} {precision} V={V}", flush=True)
    r = api.Session(p, precision=precision).run(p["tauVar"][:, 0], p["tauPrm"], want_traj=True)
    print(f"[gpu] jacobi {which} {precision}: iters {r['pcg_iters']} (oracle default {int(o['pcg_iters'].sum())})",
          flush=True)
    assert r["status"] == 0 and o["status"] == 0
    tol = 1e-8 if precision == "fp64" else 1e-7
    disp = np.linalg.norm(o["lmkStk"][-1] - o["lmkStk"][0])
    assert np.linalg.norm(r["lmkStk"][-1] - o["lmkStk"][-1]) / disp < tol
    assert abs(r["obj"] - o["obj"]) <= tol * abs(o["obj"])
