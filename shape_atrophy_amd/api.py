"""Python mirror of the reference's MATLAB entry points, over the C ABI of
libsa_hip.so (include/shape_atrophy.h).

    deformation(...)    = CUDA/deformation.m:1-62      (deformation_{2D,3D}_frame)
    computeObjfcn(...)  = CUDA/computeObjfcn.m:1-62    (computeObjfcn_{2D,3D}_frame)

Same positional arguments, same meaning, same shapes:
  h_lmkIniMat  dim x lmkNum
  elmObj       dict of the index tables built by shape_atrophy_amd.mesh
               (each as the MATLAB layout, see mesh.py)
  tgtObj       {'cenPos' | 'cen': dim x Ft, 'uniDir' | 'dir': dim x Ft, 'elmVol' | 'vol': Ft}
  vfdObj       {'cenKnlType', 'cenKnlWidth', 'dirKnlType', 'dirKnlWidth'}
and the same outputs: lmkStk dim x lmkNum x timeNum, tauMat lmkNum x
timeNum, objVec 1 x nCand.  Like the wrappers, cgTolVal is squared and
timeNum = floor(timeEnd / timeStp) + 1.  An unknown elasticity type raises
ValueError ('Unknown elasticity type ...', deformation.m:16), an unsupported
dimension too (deformation.m:57).  The device library must be present: there
is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import math
import os

import numpy as np

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("SA_HIP_LIB") or os.path.join(_PKG, "libsa_hip.so")   # SA_HIP_LIB: another build of the library
_lib = None

_P = C.POINTER
_PI = _P(C.c_int)
_PD = _P(C.c_double)

SA_ELA_FRAME, SA_ELA_ISOTROPIC = 0, 1
SA_PREC_FP64, SA_PREC_MIXED = 0, 1


class SaMesh(C.Structure):
    _fields_ = [("dim", C.c_int), ("on_device", C.c_int), ("lmkNum", C.c_int), ("elmNum", C.c_int),
                ("vtxElmCols", C.c_int), ("bdrElmNum", C.c_int), ("vtxBdrCols", C.c_int), ("btmElmNum", C.c_int),
                ("btmLmkNum", C.c_int), ("vtxBtmCols", C.c_int), ("nzrNum", C.c_int), ("femIfoCols", C.c_int),
                ("vfdElmNum", C.c_int), ("elmVtxMat", _PI), ("vtxElmMat", _PI), ("frmVtxMat", _PI),
                ("tsvVtxMat", _PI), ("bdrVtxMat", _PI), ("vtxBdrMat", _PI), ("btmVtxMat", _PI),
                ("vtxBtmMat", _PI), ("femVtxMat", _PI), ("femIfoMat", _PI), ("vfdVtxMat", _PI),
                ("elmLyrVec", _PI), ("ctlNum", C.c_int), ("ctlIdxVec", _PI)]


class SaParams(C.Structure):
    _fields_ = [("elaType", C.c_int), ("ldmWgt", C.c_double), ("knlOrder", C.c_int), ("knlWidth", C.c_double),
                ("knlEps", C.c_double), ("modVec", _PD), ("modNum", C.c_int), ("spdTanVal", C.c_double),
                ("spdTsvVal", C.c_double), ("ynkActPrmVec", _PD), ("reaActPrmVec", _PD), ("btmWgt", C.c_double),
                ("timeStp", C.c_double), ("timeNum", C.c_int), ("cgItrMax", C.c_int), ("cgTolSqu", C.c_double),
                ("precision", C.c_int), ("rdTolVal", C.c_double), ("bdrYank", C.c_int), ("velPrec", C.c_int)]


class SaTarget(C.Structure):
    _fields_ = [("tgtElmNum", C.c_int), ("cenPosMat", _PD), ("uniDirMat", _PD), ("elmVolVec", _PD),
                ("cenKnlType", C.c_char), ("cenKnlWidth", C.c_double), ("dirKnlType", C.c_char),
                ("dirKnlWidth", C.c_double)]


class SaStats(C.Structure):
    _fields_ = [("status", C.c_int), ("steps", C.c_int), ("pcgIters", C.c_int), ("rdIters", C.c_int),
                ("refineSweeps", C.c_int), ("seconds", C.c_double), ("kmatvecMs", C.c_double),
                ("kmatvecCount", C.c_int), ("operatorMs", C.c_double), ("operatorCount", C.c_int),
                ("fp64Steps", C.c_int), ("failReason", C.c_int), ("varifoldMs", C.c_double),
                ("varifoldPairs", C.c_longlong), ("varifoldMfma", C.c_int), ("deviceBytes", C.c_longlong),
                ("factorizations", C.c_int)]


class SaError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"shape_atrophy error {code}: {msg}")
        self.code = code


def lib():
    """Load libsa_hip.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with `python -m shape_atrophy_amd.build`")
        L = C.CDLL(LIB_PATH)
        L.sa_last_error.restype = C.c_char_p
        L.sa_session_footprint.restype = C.c_longlong
        L.sa_session_create.argtypes = [_P(SaMesh), _P(SaTarget), _P(SaParams), _PD, _P(C.c_void_p)]
        L.sa_session_run.argtypes = [C.c_void_p, _PD, C.c_int, _PD, C.c_int, _PD, _PD, _PD, _PD, _P(SaStats)]
        L.sa_session_destroy.argtypes = [C.c_void_p]
        L.sa_session_probe.argtypes = [C.c_void_p, C.c_int, _PD, C.c_int, _PD, C.c_int, _PD, _PD]
        L.sa_session_set_profiling.argtypes = [C.c_void_p, C.c_int]
        L.sa_session_set_time_num.argtypes = [C.c_void_p, C.c_int]
        L.sa_session_set_params.argtypes = [C.c_void_p, _P(SaParams)]
        L.sa_session_time_product.argtypes = [C.c_void_p, C.c_int, C.c_int, _PD]
        L.sa_tile_target.argtypes = [C.c_char_p, C.c_int, C.c_longlong, C.c_longlong]
        L.sa_server_create.argtypes = [C.c_int, C.c_longlong, _P(C.c_void_p)]
        L.sa_server_destroy.argtypes = [C.c_void_p]
        L.sa_server_add_mesh.argtypes = [C.c_void_p, _P(SaMesh), _P(SaTarget), _P(SaParams), _PD, _PI]
        L.sa_server_submit.argtypes = [C.c_void_p, C.c_int, _PD, C.c_int, _PD, C.c_int, _P(SaParams),
                                       _P(C.c_longlong)]
        L.sa_server_wait.argtypes = [C.c_void_p, C.c_longlong, _PD, _P(SaStats), _PD]
        L.sa_server_info.argtypes = [C.c_void_p, _P(C.c_longlong), _PI, _PI, _PI]
        L.sa_deformation.argtypes = [_P(SaMesh), _P(SaParams), _PD, C.c_int, C.c_int, _PD, C.c_int, _PD, _PD, _PD, _PD]
        L.sa_compute_objfcn.argtypes = [_P(SaMesh), _P(SaTarget), _P(SaParams), _PD, C.c_int, C.c_int, _PD, C.c_int,
                                        _PD, _PD, _PD]
        _lib = L
    return _lib


def tile_target(product: str, f32: bool, rows: int, cols: int) -> int:
    """Block target a new session takes for a stored product of this shape
    from the tile table (0: none, the built-in default)."""
    return lib().sa_tile_target(product.encode(), 1 if f32 else 0, int(rows), int(cols))


def _check(code: int):
    if code != 0:
        raise SaError(code, lib().sa_last_error().decode())


def _dp(a):
    return None if a is None else a.ctypes.data_as(_PD)


def _ip(a):
    return None if a is None else a.ctypes.data_as(_PI)


class _Keep(list):
    pass


def _i32(a, keep):
    a = np.ascontiguousarray(a, dtype=np.int32)
    keep.append(a)
    return a


def _f64(a, keep):
    a = np.ascontiguousarray(a, dtype=np.float64)
    keep.append(a)
    return a


def time_num(timeStp: float, timeEnd: float) -> int:
    return int(math.floor(timeEnd / timeStp)) + 1


def make_mesh(elmObj: dict, lmkIni: np.ndarray, keep: _Keep, with_vfd: bool) -> SaMesh:
    dim, V = lmkIni.shape
    m = SaMesh()
    m.dim = dim
    m.on_device = 0
    m.lmkNum = V
    e = _i32(elmObj["elmVtxMat"], keep)
    m.elmNum = e.shape[1]
    m.elmVtxMat = _ip(e)
    ve = _i32(elmObj["vtxElmMat"], keep)
    m.vtxElmCols = ve.shape[0]
    m.vtxElmMat = _ip(ve)
    frm = _i32(elmObj["tanVtxMat"] if dim == 2 else elmObj["nmlVtxMat"], keep)
    m.frmVtxMat = _ip(frm)
    m.tsvVtxMat = _ip(_i32(elmObj["tsvVtxMat"], keep))
    bdr = _i32(elmObj.get("bdrVtxMat", np.zeros((dim, 0))), keep)
    m.bdrElmNum = bdr.shape[1]
    m.bdrVtxMat = _ip(bdr)
    vb = _i32(elmObj.get("vtxBdrMat", np.zeros((1, V))), keep)
    m.vtxBdrCols = vb.shape[0]
    m.vtxBdrMat = _ip(vb)
    btm = _i32(elmObj["btmVtxMat"], keep)
    m.btmElmNum = btm.shape[1]
    m.btmVtxMat = _ip(btm)
    vtb = _i32(elmObj["vtxBtmMat"], keep)
    m.btmLmkNum = vtb.shape[1]
    m.vtxBtmCols = vtb.shape[0]
    m.vtxBtmMat = _ip(vtb)
    fv = _i32(elmObj["femVtxMat"], keep)
    m.nzrNum = fv.shape[1]
    m.femVtxMat = _ip(fv)
    fi = _i32(elmObj["femIfoMat"], keep)
    m.femIfoCols = fi.shape[0]
    m.femIfoMat = _ip(fi)
    if with_vfd:
        vf = _i32(elmObj["vfdVtxMat"], keep)
        m.vfdElmNum = vf.shape[1]
        m.vfdVtxMat = _ip(vf)
    if elmObj.get("elmLyrVec") is not None:
        m.elmLyrVec = _ip(_i32(elmObj["elmLyrVec"], keep))
    ctl = elmObj.get("ctlIdxVec")
    if ctl is not None and len(ctl):
        c = _i32(ctl, keep)
        m.ctlNum = c.size
        m.ctlIdxVec = _ip(c)
    return m


def make_params(elaType, ldmWgt, knlOrder, knlWidth, knlEps, modVec, ynkActPrmVec, spdTanVal, spdTsvVal,
                reaActPrmVec, btmWgt, timeStp, timeEnd, cgItrMax, cgTolVal, keep, precision="fp64",
                rdTolVal=1e-13, bdrYank=0, velPrec="default") -> SaParams:
    et = str(elaType).lower()
    if et not in ("frame", "isotropic"):
        raise ValueError(f"Unknown elasticity type {elaType}.")
    p = SaParams()
    p.elaType = SA_ELA_FRAME if et == "frame" else SA_ELA_ISOTROPIC
    p.ldmWgt = float(ldmWgt)
    p.knlOrder = int(knlOrder)
    p.knlWidth = float(knlWidth)
    p.knlEps = float(knlEps)
    mv = _f64(np.asarray(modVec, dtype=np.float64).reshape(-1), keep)
    p.modVec = _dp(mv)
    p.modNum = mv.size
    p.spdTanVal = float(spdTanVal)
    p.spdTsvVal = float(spdTsvVal)
    p.ynkActPrmVec = _dp(_f64(ynkActPrmVec, keep))
    p.reaActPrmVec = _dp(_f64(reaActPrmVec, keep))
    p.btmWgt = float(btmWgt)
    p.timeStp = float(timeStp)
    p.timeNum = time_num(timeStp, timeEnd)
    p.cgItrMax = int(cgItrMax)
    p.cgTolSqu = float(cgTolVal) ** 2
    p.precision = {"fp64": SA_PREC_FP64, "mixed": SA_PREC_MIXED}[precision]
    p.rdTolVal = float(rdTolVal)
    p.bdrYank = int(bool(bdrYank))
    p.velPrec = {"default": 0, "jacobi": 1}[velPrec]
    return p


def make_target(tgtObj: dict, vfdObj: dict, keep) -> SaTarget:
    cen = _f64(tgtObj.get("cenPosMat", tgtObj.get("cen")), keep)
    dr = _f64(tgtObj.get("uniDirMat", tgtObj.get("dir")), keep)
    vol = _f64(tgtObj.get("elmVolVec", tgtObj.get("vol")), keep)
    t = SaTarget()
    t.tgtElmNum = vol.size
    t.cenPosMat, t.uniDirMat, t.elmVolVec = _dp(cen), _dp(dr), _dp(vol)
    t.cenKnlType = str(vfdObj["cenKnlType"]).encode()[:1]
    t.cenKnlWidth = float(vfdObj["cenKnlWidth"])
    t.dirKnlType = str(vfdObj["dirKnlType"]).encode()[:1]
    t.dirKnlWidth = float(vfdObj["dirKnlWidth"])
    return t


def _mod_vector(elaType, modVec):
    # 'isotropic': the wrapper reshapes modVec to nLayer x 2 (lambda, mu) rows
    # (deformation.m:12-13); flattened row-major that is the pair sequence
    return np.asarray(modVec, dtype=np.float64).reshape(-1)


def deformation(h_tauVarVec, h_tauPrmVec, elaType, h_lmkIniMat, elmObj, ldmWgt, knlOrder, knlWidth, knlEps,
                h_modVec, h_ynkActPrmVec, spdTanVal, spdTsvVal, h_reaActPrmVec, btmWgt, timeStp, timeEnd,
                cgItrMax, cgTolVal, precision="fp64"):
    """CUDA/deformation.m: returns (lmkStk dim x lmkNum x timeNum, tauMat lmkNum x timeNum)."""
    lmk = np.ascontiguousarray(h_lmkIniMat, dtype=np.float64)
    dim, V = lmk.shape
    if dim not in (2, 3):
        raise ValueError(f"Dimension {dim} is not supported.")
    keep = _Keep()
    prm = make_params(elaType, ldmWgt, knlOrder, knlWidth, knlEps, _mod_vector(elaType, h_modVec), h_ynkActPrmVec,
                      spdTanVal, spdTsvVal, h_reaActPrmVec, btmWgt, timeStp, timeEnd, cgItrMax, cgTolVal, keep,
                      precision)
    mesh = make_mesh(elmObj, lmk, keep, with_vfd=False)
    tv = np.asarray(h_tauVarVec, dtype=np.float64)
    if tv.ndim == 1:
        tv = tv[:, None]
    tvF = _f64(tv.T, keep)                       # column-major nVar x nCand
    tp = _f64(np.asarray(h_tauPrmVec, dtype=np.float64).reshape(-1), keep)
    T = prm.timeNum
    lmkStk = np.zeros((T, dim, V))
    tauMat = np.zeros((T, V))
    obj = np.zeros(tv.shape[1])
    _check(lib().sa_deformation(C.byref(mesh), C.byref(prm), _dp(tvF), tv.shape[0], tv.shape[1], _dp(tp), tp.size,
                                _dp(lmk), _dp(lmkStk), _dp(tauMat), _dp(obj)))
    return np.transpose(lmkStk, (1, 2, 0)), tauMat.T


def computeObjfcn(h_tauVarMat, h_tauPrmVec, elaType, h_lmkIniMat, elmObj, tgtObj, vfdObj, ldmWgt, knlOrder,
                  knlWidth, knlEps, h_modVec, h_ynkActPrmVec, spdTanVal, spdTsvVal, h_reaActPrmVec, btmWgt,
                  timeStp, timeEnd, cgItrMax, cgTolVal, precision="fp64", return_grad=False):
    """CUDA/computeObjfcn.m: returns objVec (1 x nCand).  With return_grad
    also the gradient of the last candidate's objective with respect to its
    final landmarks (dim x lmkNum)."""
    lmk = np.ascontiguousarray(h_lmkIniMat, dtype=np.float64)
    dim, V = lmk.shape
    if dim not in (2, 3):
        raise ValueError(f"Dimension {dim} is not supported.")
    keep = _Keep()
    prm = make_params(elaType, ldmWgt, knlOrder, knlWidth, knlEps, _mod_vector(elaType, h_modVec), h_ynkActPrmVec,
                      spdTanVal, spdTsvVal, h_reaActPrmVec, btmWgt, timeStp, timeEnd, cgItrMax, cgTolVal, keep,
                      precision)
    mesh = make_mesh(elmObj, lmk, keep, with_vfd=True)
    tgt = make_target(tgtObj, vfdObj, keep)
    tv = np.asarray(h_tauVarMat, dtype=np.float64)
    if tv.ndim == 1:
        tv = tv[:, None]
    tvF = _f64(tv.T, keep)
    tp = _f64(np.asarray(h_tauPrmVec, dtype=np.float64).reshape(-1), keep)
    obj = np.zeros(tv.shape[1])
    grad = np.zeros((dim, V)) if return_grad else None
    _check(lib().sa_compute_objfcn(C.byref(mesh), C.byref(tgt), C.byref(prm), _dp(tvF), tv.shape[0], tv.shape[1],
                                   _dp(tp), tp.size, _dp(lmk), _dp(obj), _dp(grad)))
    objVec = obj[None, :]
    return (objVec, grad) if return_grad else objVec


class Session:
    """Device-resident solver for one mesh (and optional target): the state
    stays in HBM across candidates (sa_session_* in shape_atrophy.h)."""

    def __init__(self, prob: dict, precision: str = "fp64", with_target: bool = True):
        keep = _Keep()
        p = prob["params"]
        lmk = _f64(prob["q0"], keep)
        self.dim, self.V = lmk.shape
        elm = dict(prob["elm"])
        if prob.get("ctlIdx") is not None:
            elm["ctlIdxVec"] = prob["ctlIdx"]
        self.prm = make_params(p["elaType"], p["ldmWgt"], p["knlOrder"], p["knlWidth"], p["knlEps"],
                               p["modMat"] if p["elaType"] == "isotropic" else p["modVec"], p["ynkActPrmVec"],
                               p["spdTanVal"], p["spdTsvVal"], p["reaActPrmVec"], p["btmWgt"], p["timeStp"],
                               p["timeEnd"], p["cgItrMax"], p["cgTolVal"], keep, precision,
                               p.get("rdTolVal", 1e-13), p.get("bdrYank", 0), p.get("velPrec", "default"))
        if prob.get("timeNum"):
            self.prm.timeNum = int(prob["timeNum"])
        with_target = with_target and "target" in prob
        self.mesh = make_mesh(elm, lmk, keep, with_vfd=with_target)
        self.tgt = make_target(prob["target"], prob["vfd"], keep) if with_target else None
        self.T = self.prm.timeNum
        self._keep = keep
        h = C.c_void_p()
        _check(lib().sa_session_create(C.byref(self.mesh), C.byref(self.tgt) if self.tgt is not None else None,
                                       C.byref(self.prm), _dp(lmk), C.byref(h)))
        self._h = h
        self.footprint = int(lib().sa_session_footprint(C.byref(self.mesh),
                                                        C.byref(self.tgt) if self.tgt is not None else None,
                                                        C.byref(self.prm)))

    def run(self, tauVar, tauPrm, want_traj: bool = False, want_grad: bool = False) -> dict:
        tv = np.ascontiguousarray(np.asarray(tauVar, dtype=np.float64).reshape(-1))
        tp = np.ascontiguousarray(np.asarray(tauPrm, dtype=np.float64).reshape(-1))
        obj = np.zeros(1)
        lmk = np.zeros((self.T, self.dim, self.V)) if want_traj else None
        tau = np.zeros((self.T, self.V)) if want_traj else None
        grad = np.zeros((self.dim, self.V)) if want_grad else None
        st = SaStats()
        _check(lib().sa_session_run(self._h, _dp(tv), tv.size, _dp(tp), tp.size, _dp(obj), _dp(lmk), _dp(tau),
                                    _dp(grad), C.byref(st)))
        return dict(obj=float(obj[0]), lmkStk=lmk, tauMat=tau, grad=grad, status=st.status, steps=st.steps,
                    pcg_iters=st.pcgIters, rd_iters=st.rdIters, refine=st.refineSweeps, fp64_steps=st.fp64Steps,
                    fail_reason=st.failReason, seconds=st.seconds,
                    kmatvec_ms=st.kmatvecMs, kmatvec_count=st.kmatvecCount, operator_ms=st.operatorMs,
                    operator_count=st.operatorCount, varifold_ms=st.varifoldMs, varifold_pairs=st.varifoldPairs,
                    varifold_mfma=st.varifoldMfma, device_bytes=st.deviceBytes, factorizations=st.factorizations)

    def set_params(self, params: dict):
        """New scalar parameters for the following runs (sa_session_set_params)."""
        keep = _Keep()
        prm = params_struct(params, keep, "mixed" if self.prm.precision == SA_PREC_MIXED else "fp64", self.T)
        _check(lib().sa_session_set_params(self._h, C.byref(prm)))

    def set_time_num(self, timeNum: int):
        _check(lib().sa_session_set_time_num(self._h, int(timeNum)))
        self.T = int(timeNum)

    def set_profiling(self, on: bool = True):
        _check(lib().sa_session_set_profiling(self._h, 1 if on else 0))

    PROBES = {"elastic": 1, "kmatvec": 2, "yank": 3, "fem": 4, "geometry": 5, "varifold": 6, "penalty": 7,
              "kxc": 8, "kcx": 9, "jacobi": 10, "kccchol": 12}

    def probe(self, what: str, tauVar, tauPrm, x=None) -> np.ndarray:
        """One operator at the initial state (sa_session_probe)."""
        tv = np.ascontiguousarray(np.asarray(tauVar, dtype=np.float64).reshape(-1))
        tp = np.ascontiguousarray(np.asarray(tauPrm, dtype=np.float64).reshape(-1))
        d, V = self.dim, self.V
        E, Z = self.mesh.elmNum, self.mesh.nzrNum
        nc = self.mesh.ctlNum
        size = {"elastic": d * V, "kmatvec": d * V, "penalty": d * V, "kxc": d * V, "kcx": d * nc, "jacobi": d * nc,
                "kccchol": nc * nc,
                "yank": d * V + V, "fem": 2 * Z,
                "geometry": (d * d + 1 + 2 * d) * E, "varifold": 1 + d * V}[what]
        out = np.zeros(size)
        xin = None if x is None else np.ascontiguousarray(x, dtype=np.float64)
        _check(lib().sa_session_probe(self._h, self.PROBES[what], _dp(tv), tv.size, _dp(tp), tp.size, _dp(xin),
                                      _dp(out)))
        return out

    TIMED = {"kxc": 8, "kcx": 9, "kmatvec": 2, "precond": 11, "operator": 1}

    def time_product(self, what: str, reps: int = 20) -> float:
        """Device microseconds of one kernel product as a solve launches it
        (sa_session_time_product): "kxc", "kcx", "kmatvec", "precond" or
        "operator" (the operator apply at the last start's geometry)."""
        us = C.c_double()
        _check(lib().sa_session_time_product(self._h, self.TIMED[what], int(reps), C.byref(us)))
        return us.value

    def close(self):
        if getattr(self, "_h", None):
            lib().sa_session_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def params_struct(p: dict, keep, precision: str = "fp64", timeNum: int | None = None) -> SaParams:
    """SaParams of a workload parameter dict."""
    prm = make_params(p["elaType"], p["ldmWgt"], p["knlOrder"], p["knlWidth"], p["knlEps"],
                      p["modMat"] if p["elaType"] == "isotropic" else p["modVec"], p["ynkActPrmVec"],
                      p["spdTanVal"], p["spdTsvVal"], p["reaActPrmVec"], p["btmWgt"], p["timeStp"], p["timeEnd"],
                      p["cgItrMax"], p["cgTolVal"], keep, precision, p.get("rdTolVal", 1e-13),
                      p.get("bdrYank", 0), p.get("velPrec", "default"))
    if timeNum:
        prm.timeNum = int(timeNum)
    return prm


class Server:
    """Request server over device-resident sessions (sa_server_* in
    shape_atrophy.h): worker threads, per-session HIP streams, admission by
    device-memory footprint, per-request parameters."""

    def __init__(self, workers: int = 4, mem_budget: int = 0):
        h = C.c_void_p()
        _check(lib().sa_server_create(int(workers), int(mem_budget), C.byref(h)))
        self._h = h
        self._keep = []
        self._prec = {}

    def add_mesh(self, prob: dict, precision: str = "fp64") -> int:
        keep = _Keep()
        lmk = _f64(prob["q0"], keep)
        elm = dict(prob["elm"])
        if prob.get("ctlIdx") is not None:
            elm["ctlIdxVec"] = prob["ctlIdx"]
        prm = params_struct(prob["params"], keep, precision, prob.get("timeNum"))
        with_t = "target" in prob
        mesh = make_mesh(elm, lmk, keep, with_vfd=with_t)
        tgt = make_target(prob["target"], prob["vfd"], keep) if with_t else None
        mid = C.c_int()
        _check(lib().sa_server_add_mesh(self._h, C.byref(mesh), C.byref(tgt) if tgt is not None else None,
                                        C.byref(prm), _dp(lmk), C.byref(mid)))
        self._prec[mid.value] = (precision, prob.get("timeNum"))
        return mid.value

    def submit(self, mesh_id: int, tauVar, tauPrm, params: dict | None = None) -> int:
        keep = _Keep()
        tv = _f64(np.asarray(tauVar, dtype=np.float64).reshape(-1), keep)
        tp = _f64(np.asarray(tauPrm, dtype=np.float64).reshape(-1), keep)
        prm = None
        if params is not None:
            prec, tn = self._prec[mesh_id]
            prm = params_struct(params, keep, prec, tn)
        rid = C.c_longlong()
        _check(lib().sa_server_submit(self._h, int(mesh_id), _dp(tv), tv.size, _dp(tp), tp.size,
                                      C.byref(prm) if prm is not None else None, C.byref(rid)))
        return rid.value

    def wait(self, req_id: int) -> dict:
        obj = np.zeros(1)
        st = SaStats()
        times = np.zeros(3)
        rc = lib().sa_server_wait(self._h, C.c_longlong(req_id), _dp(obj), C.byref(st), _dp(times))
        return dict(rc=rc, error=lib().sa_last_error().decode() if rc else "", obj=float(obj[0]),
                    status=st.status, steps=st.steps, pcg_iters=st.pcgIters, refine=st.refineSweeps,
                    fp64_steps=st.fp64Steps, fail_reason=st.failReason, factorizations=st.factorizations,
                    seconds=st.seconds, t_submit=times[0], t_start=times[1], t_end=times[2])

    def info(self) -> dict:
        mem = C.c_longlong()
        ns, nq, nr = C.c_int(), C.c_int(), C.c_int()
        _check(lib().sa_server_info(self._h, C.byref(mem), C.byref(ns), C.byref(nq), C.byref(nr)))
        return dict(mem_used=mem.value, sessions=ns.value, queued=nq.value, running=nr.value)

    def close(self):
        if getattr(self, "_h", None):
            lib().sa_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def problem_args(prob: dict):
    """Positional argument list of computeObjfcn for a workload dict."""
    p = prob["params"]
    mod = p["modMat"] if p["elaType"] == "isotropic" else p["modVec"]
    return [prob["tauVar"], prob["tauPrm"], p["elaType"], prob["q0"], prob["elm"], prob.get("target"),
            prob.get("vfd"), p["ldmWgt"], p["knlOrder"], p["knlWidth"], p["knlEps"], mod, p["ynkActPrmVec"],
            p["spdTanVal"], p["spdTsvVal"], p["reaActPrmVec"], p["btmWgt"], p["timeStp"], p["timeEnd"],
            p["cgItrMax"], p["cgTolVal"]]
