"""ctypes front end of the FP64 CPU oracle (oracle/sa_oracle.cpp).

Test infrastructure: imported by tests/, by __graft_entry__.smoke() and by
the CPU-baseline leg of bench.py, always as the checker.  The product path
never imports it.
"""
from __future__ import annotations

import ctypes as C
import math
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libsa_oracle.so")
_lib = None

_P = C.POINTER
_I = C.c_int
_D = C.c_double
_PI = _P(C.c_int)
_PD = _P(C.c_double)


class OMesh(C.Structure):
    _fields_ = [("dim", _I), ("lmkNum", _I), ("elmNum", _I), ("nzrNum", _I), ("btmElmNum", _I),
                ("btmLmkNum", _I), ("vfdElmNum", _I), ("vtxElmRows", _I), ("vtxBtmRows", _I),
                ("femIfoRows", _I), ("elmVtx", _PI), ("vtxElm", _PI), ("frmVtx", _PI), ("tsvVtx", _PI),
                ("btmVtx", _PI), ("vtxBtm", _PI), ("femVtx", _PI), ("femIfo", _PI), ("vfdVtx", _PI),
                ("elmLyr", _PI), ("ctlNum", _I), ("ctlIdx", _PI), ("bdrElmNum", _I), ("vtxBdrRows", _I),
                ("bdrVtx", _PI), ("vtxBdr", _PI), ("bdrYank", _I)]


class OParams(C.Structure):
    _fields_ = [("ldmWgt", _D), ("knlOrder", _I), ("knlWidth", _D), ("knlEps", _D), ("modVec", _D * 4),
                ("elaType", _I), ("nLyr", _I), ("modMat", _PD), ("spdTan", _D), ("spdTsv", _D),
                ("ynk", _D * 4), ("rea", _D * 5), ("btmWgt", _D), ("timeStp", _D), ("timeNum", _I),
                ("cgItrMax", _I), ("cgTolSqu", _D), ("pcgInit", _I), ("rdSolve", _I), ("rdTolSqu", _D),
                ("velPrec", _I)]


class OTarget(C.Structure):
    _fields_ = [("tgtNum", _I), ("cen", _PD), ("dir", _PD), ("vol", _PD), ("cenType", _I), ("dirType", _I),
                ("cenWidth", _D), ("dirWidth", _D)]


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (g++ -fopenmp)."""
    src = os.path.join(_HERE, "sa_oracle.cpp")
    if force or not os.path.exists(_LIB_PATH) or os.path.getmtime(_LIB_PATH) < os.path.getmtime(src):
        subprocess.run(["make", "-s", "-C", _HERE, "-B" if force else "all"], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = C.CDLL(_LIB_PATH)
        _lib.or_objfcn.restype = _I
        _lib.or_varifold.restype = _D
        _lib.or_num_threads.restype = _I
        _lib.or_time_elastic_apply.restype = _D
    return _lib


def _ip(a):
    return a.ctypes.data_as(_PI)


def _dp(a):
    return a.ctypes.data_as(_PD)


class _Keep(list):
    """Holds numpy buffers alive while ctypes structs point into them."""


def time_num(prm: dict) -> int:
    return int(math.floor(prm["timeEnd"] / prm["timeStp"])) + 1


def make_mesh(prob: dict, keep: _Keep) -> OMesh:
    e = prob["elm"]
    d = prob["dim"]

    def arr(name):
        a = np.ascontiguousarray(e[name], dtype=np.int32)
        keep.append(a)
        return a

    elmVtx = arr("elmVtxMat"); vtxElm = arr("vtxElmMat")
    frm = arr("tanVtxMat" if d == 2 else "nmlVtxMat"); tsv = arr("tsvVtxMat")
    btm = arr("btmVtxMat"); vtxBtm = arr("vtxBtmMat")
    femVtx = arr("femVtxMat"); femIfo = arr("femIfoMat"); vfd = arr("vfdVtxMat")
    lyr = arr("elmLyrVec")
    m = OMesh()
    m.dim = d
    m.lmkNum = prob["q0"].shape[1]
    m.elmNum = elmVtx.shape[1]
    m.nzrNum = femVtx.shape[1]
    m.btmElmNum = btm.shape[1]
    m.btmLmkNum = vtxBtm.shape[1]
    m.vfdElmNum = vfd.shape[1]
    m.vtxElmRows = vtxElm.shape[0]
    m.vtxBtmRows = vtxBtm.shape[0]
    m.femIfoRows = femIfo.shape[0]
    m.elmVtx, m.vtxElm, m.frmVtx, m.tsvVtx = _ip(elmVtx), _ip(vtxElm), _ip(frm), _ip(tsv)
    m.btmVtx, m.vtxBtm, m.femVtx, m.femIfo, m.vfdVtx = _ip(btm), _ip(vtxBtm), _ip(femVtx), _ip(femIfo), _ip(vfd)
    m.elmLyr = _ip(lyr)
    ctl = prob.get("ctlIdx")
    if ctl is not None and len(ctl):
        c = np.ascontiguousarray(ctl, dtype=np.int32)
        keep.append(c)
        m.ctlNum = c.size
        m.ctlIdx = _ip(c)
    else:
        m.ctlNum = 0
    if prob["params"].get("bdrYank", 0) and "bdrVtxMat" in e and "vtxBdrMat" in e:
        bdr = arr("bdrVtxMat"); vb = arr("vtxBdrMat")
        m.bdrElmNum = bdr.shape[1]
        m.vtxBdrRows = vb.shape[0]
        m.bdrVtx, m.vtxBdr = _ip(bdr), _ip(vb)
        m.bdrYank = 1
    return m


def make_params(prob: dict, keep: _Keep, pcg_init: int = 1, rd_solve: int = 1, rd_tol: float = 1e-13) -> OParams:
    p = prob["params"]
    o = OParams()
    o.ldmWgt = p["ldmWgt"]
    o.knlOrder = int(p["knlOrder"])
    o.knlWidth = p["knlWidth"]
    o.knlEps = p["knlEps"]
    mv = list(p["modVec"]) + [0.0] * (4 - len(p["modVec"]))
    o.modVec = (_D * 4)(*mv)
    o.elaType = 0 if p.get("elaType", "frame") == "frame" else 1
    mm = np.ascontiguousarray(np.asarray(p.get("modMat", [[0.0, 0.0]]), dtype=np.float64))
    keep.append(mm)
    o.nLyr = mm.shape[0]
    o.modMat = _dp(mm)
    o.spdTan = p["spdTanVal"]
    o.spdTsv = p["spdTsvVal"]
    o.ynk = (_D * 4)(*p["ynkActPrmVec"])
    o.rea = (_D * 5)(*p["reaActPrmVec"])
    o.btmWgt = p["btmWgt"]
    o.timeStp = p["timeStp"]
    o.timeNum = prob.get("timeNum", time_num(p))
    o.cgItrMax = int(p["cgItrMax"])
    o.cgTolSqu = p["cgTolVal"] ** 2
    o.pcgInit = pcg_init
    o.rdSolve = rd_solve
    o.rdTolSqu = rd_tol ** 2
    o.velPrec = 1 if p.get("velPrec", "default") == "jacobi" else 0
    return o


def make_target(prob: dict, keep: _Keep) -> OTarget:
    t = prob["target"]
    v = prob["vfd"]
    cen = np.ascontiguousarray(t["cen"], dtype=np.float64)
    dr = np.ascontiguousarray(t["dir"], dtype=np.float64)
    vol = np.ascontiguousarray(t["vol"], dtype=np.float64)
    keep.extend([cen, dr, vol])
    o = OTarget()
    o.tgtNum = vol.size
    o.cen, o.dir, o.vol = _dp(cen), _dp(dr), _dp(vol)
    o.cenType = ord(v["cenKnlType"])
    o.dirType = ord(v["dirKnlType"])
    o.cenWidth = v["cenKnlWidth"]
    o.dirWidth = v["dirKnlWidth"]
    return o


def objfcn(prob: dict, cand: int = 0, pcg_init: int = 1, rd_solve: int = 1, rd_tol: float = 1e-13,
           with_target: bool = True, with_grad: bool = False, threads: int | None = None) -> dict:
    """Run the forward solve (and the varifold objective) for candidate
    column `cand` of prob['tauVar']."""
    L = lib()
    if threads:
        L.or_set_threads(int(threads))
    keep = _Keep()
    m = make_mesh(prob, keep)
    p = make_params(prob, keep, pcg_init, rd_solve, rd_tol)
    d = prob["dim"]
    V = m.lmkNum
    T = p.timeNum
    q0 = np.ascontiguousarray(prob["q0"], dtype=np.float64)
    tv = np.ascontiguousarray(np.asarray(prob["tauVar"], dtype=np.float64)[:, cand])
    tp = np.ascontiguousarray(prob["tauPrm"], dtype=np.float64)
    lmk = np.zeros((T, d, V))
    tau = np.zeros((T, V))
    obj = np.zeros(1)
    grad = np.zeros((d, V)) if with_grad else None
    pit = np.zeros(max(T - 1, 1), dtype=np.int32)
    rit = np.zeros(max(T - 1, 1), dtype=np.int32)
    tgt = make_target(prob, keep) if with_target and "target" in prob else None
    st = L.or_objfcn(C.byref(m), C.byref(p), C.byref(tgt) if tgt is not None else None, _dp(q0), _dp(tv), _dp(tp),
                     _dp(lmk), _dp(tau), _dp(obj), _dp(grad) if grad is not None else None, _ip(pit), _ip(rit))
    return dict(status=st, lmkStk=lmk, tauMat=tau, obj=float(obj[0]), grad=grad, pcg_iters=pit, rd_iters=rit)


# --- unit-level helpers ----------------------------------------------------

def kernel_values(dim: int, order: int, width: float, d2: np.ndarray) -> np.ndarray:
    d2 = np.ascontiguousarray(d2, dtype=np.float64)
    out = np.empty_like(d2)
    lib().or_kernel_values(dim, order, C.c_double(width), _dp(d2), d2.size, _dp(out))
    return out


def bessel_k01(x: np.ndarray):
    x = np.ascontiguousarray(x, dtype=np.float64)
    k0 = np.empty_like(x)
    k1 = np.empty_like(x)
    lib().or_bessel_k01(_dp(x), x.size, _dp(k0), _dp(k1))
    return k0, k1


def activities(tau: np.ndarray, ynk, rea):
    tau = np.ascontiguousarray(tau, dtype=np.float64)
    y = np.ascontiguousarray(ynk, dtype=np.float64)
    r = np.ascontiguousarray(rea, dtype=np.float64)
    ya = np.empty_like(tau)
    ra = np.empty_like(tau)
    lib().or_activities(_dp(tau), tau.size, _dp(y), _dp(r), _dp(ya), _dp(ra))
    return ya, ra


def tau_initial(prob: dict, q=None, cand: int = 0) -> np.ndarray:
    q = np.ascontiguousarray(prob["q0"] if q is None else q, dtype=np.float64)
    tv = np.ascontiguousarray(np.asarray(prob["tauVar"], dtype=np.float64)[:, cand])
    tp = np.ascontiguousarray(prob["tauPrm"], dtype=np.float64)
    out = np.empty(q.shape[1])
    lib().or_tau_initial(prob["dim"], _dp(q), q.shape[1], _dp(tv), _dp(tp), _dp(out))
    return out


def _mesh_call(prob):
    keep = _Keep()
    return keep, make_mesh(prob, keep)


def frames(prob: dict, q) -> tuple:
    keep, m = _mesh_call(prob)
    q = np.ascontiguousarray(q, dtype=np.float64)
    f1 = np.empty((prob["dim"], m.elmNum))
    f2 = np.empty((prob["dim"], m.elmNum))
    lib().or_frames(C.byref(m), _dp(q), _dp(f1), _dp(f2))
    return f1, f2


def elastic_apply(prob: dict, q, tau, v) -> np.ndarray:
    keep, m = _mesh_call(prob)
    p = make_params(prob, keep)
    q = np.ascontiguousarray(q, dtype=np.float64)
    tau = np.ascontiguousarray(tau, dtype=np.float64)
    v = np.ascontiguousarray(v, dtype=np.float64)
    out = np.empty_like(v)
    lib().or_elastic_apply(C.byref(m), C.byref(p), _dp(q), _dp(tau), _dp(v), _dp(out))
    return out


def time_elastic_apply(prob: dict, reps: int = 100, threads: int | None = None) -> float:
    """Median wall time (s) of reps elastic-operator applies at q0 and the
    initial concentration, on the given OpenMP thread count (the CPU-baseline
    sample of the operator alone)."""
    L = lib()
    if threads:
        L.or_set_threads(int(threads))
    keep, m = _mesh_call(prob)
    p = make_params(prob, keep)
    q = np.ascontiguousarray(prob["q0"], dtype=np.float64)
    tau = tau_initial(prob)
    v = np.ascontiguousarray(np.random.default_rng(0).standard_normal(q.shape))
    return float(L.or_time_elastic_apply(C.byref(m), C.byref(p), _dp(q), _dp(tau), _dp(v), int(reps)))


def penalty_apply(prob: dict, q, v) -> np.ndarray:
    keep, m = _mesh_call(prob)
    q = np.ascontiguousarray(q, dtype=np.float64)
    v = np.ascontiguousarray(v, dtype=np.float64)
    out = np.empty_like(v)
    lib().or_penalty_apply(C.byref(m), _dp(q), _dp(v), _dp(out))
    return out


def external_yank(prob: dict, q, act) -> np.ndarray:
    keep, m = _mesh_call(prob)
    q = np.ascontiguousarray(q, dtype=np.float64)
    act = np.ascontiguousarray(act, dtype=np.float64)
    out = np.empty_like(q)
    lib().or_external_yank(C.byref(m), _dp(q), _dp(act), _dp(out))
    return out


def fem_assemble(prob: dict, q):
    keep, m = _mesh_call(prob)
    p = make_params(prob, keep)
    q = np.ascontiguousarray(q, dtype=np.float64)
    ppd = np.empty(m.nzrNum)
    ggd = np.empty(m.nzrNum)
    lib().or_fem_assemble(C.byref(m), C.byref(p), _dp(q), _dp(ppd), _dp(ggd))
    return ppd, ggd


def fem_vector(prob: dict, q, r) -> np.ndarray:
    keep, m = _mesh_call(prob)
    q = np.ascontiguousarray(q, dtype=np.float64)
    r = np.ascontiguousarray(r, dtype=np.float64)
    out = np.empty(q.shape[1])
    lib().or_fem_vector(C.byref(m), _dp(q), _dp(r), _dp(out))
    return out


def kernel_matvec(dim, order, width, eps, q, x) -> np.ndarray:
    q = np.ascontiguousarray(q, dtype=np.float64)
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.empty_like(x)
    lib().or_kernel_matvec(dim, order, C.c_double(width), C.c_double(eps), _dp(q), q.shape[1], _dp(x), _dp(y))
    return y


def varifold(prob: dict, q, with_grad: bool = False):
    keep, m = _mesh_call(prob)
    t = make_target(prob, keep)
    q = np.ascontiguousarray(q, dtype=np.float64)
    grad = np.zeros_like(q) if with_grad else None
    v = lib().or_varifold(C.byref(m), C.byref(t), _dp(q), _dp(grad) if grad is not None else None)
    return (v, grad) if with_grad else v


def jacobi_diag(prob: dict, q, tau) -> np.ndarray:
    """diag(S) of the control-point system (the Jacobi preconditioner) at
    positions q and concentration tau, dim x ctlNum."""
    keep, m = _mesh_call(prob)
    p = make_params(prob, keep)
    q = np.ascontiguousarray(q, dtype=np.float64)
    tau = np.ascontiguousarray(tau, dtype=np.float64)
    out = np.zeros((prob["dim"], len(prob["ctlIdx"])))
    lib().or_jacobi_diag(C.byref(m), C.byref(p), _dp(q), _dp(tau), _dp(out))
    return out


def facets(prob: dict, q):
    keep, m = _mesh_call(prob)
    q = np.ascontiguousarray(q, dtype=np.float64)
    d = prob["dim"]
    cen = np.empty((d, m.vfdElmNum))
    dr = np.empty((d, m.vfdElmNum))
    vol = np.empty(m.vfdElmNum)
    lib().or_facets(C.byref(m), _dp(q), _dp(cen), _dp(dr), _dp(vol))
    return cen, dr, vol


def pcg_sample(prob: dict, n_iter: int, cand: int = 0, threads: int | None = None) -> dict:
    """Time the set-up of the first step and n_iter velocity-CG iterations
    (the CPU-baseline sample)."""
    L = lib()
    if threads:
        L.or_set_threads(int(threads))
    keep = _Keep()
    m = make_mesh(prob, keep)
    p = make_params(prob, keep)
    q0 = np.ascontiguousarray(prob["q0"], dtype=np.float64)
    tv = np.ascontiguousarray(np.asarray(prob["tauVar"], dtype=np.float64)[:, cand])
    tp = np.ascontiguousarray(prob["tauPrm"], dtype=np.float64)
    out = np.zeros(3)
    L.or_pcg_sample(C.byref(m), C.byref(p), _dp(q0), _dp(tv), _dp(tp), int(n_iter), _dp(out))
    return dict(setup_s=float(out[0]), cg_start_s=float(out[1]), per_iter_s=float(out[2]), iters=int(n_iter))
