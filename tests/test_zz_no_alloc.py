"""sa_varifold_objective_noalloc: the varifold objective of device-resident
landmarks, facets and target with the library's fixed device scratch (no
allocation), against the objective of sa_session_probe(SA_PROBE_VARIFOLD)
(an FP64 session, the same kernels on session buffers): bitwise equal, on
two calls in a row, the second with smaller facet sets; an oversized call is
refused with SA_E_ARG and a facet index out of range gives NaN."""
import ctypes as C

import numpy as np
import pytest

from shape_atrophy_amd import api, workloads as W

pytestmark = pytest.mark.gpu


class Hip:
    """Device buffers through the HIP runtime the library itself loads
    (/opt/rocm/lib): inputs must be device pointers, and a second runtime
    in the process (torch's) does not see the device the library uses."""

    def __init__(self):
        api.lib()
        self.L = C.CDLL("/opt/rocm/lib/libamdhip64.so")
        self.bufs = []

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = C.c_void_p()
        assert self.L.hipMalloc(C.byref(p), C.c_size_t(max(a.nbytes, 8))) == 0
        assert self.L.hipMemcpy(p, a.ctypes.data_as(C.c_void_p), C.c_size_t(a.nbytes), 1) == 0   # host to device
        self.bufs.append(p)
        return p

    def get(self, p, n):
        out = np.zeros(n)
        assert self.L.hipDeviceSynchronize() == 0
        assert self.L.hipMemcpy(out.ctypes.data_as(C.c_void_p), p, C.c_size_t(8 * n), 2) == 0   # device to host
        return out

    def free(self):
        for p in self.bufs:
            self.L.hipFree(p)
        self.bufs = []


def noalloc_objective(prob, q):
    hip = Hip()
    d = prob["dim"]
    try:
        qd = hip.put(np.asarray(q, dtype=np.float64))
        vf = hip.put(np.asarray(prob["elm"]["vfdVtxMat"], dtype=np.int32))
        tg = prob["target"]
        cen, dr = hip.put(np.asarray(tg["cen"], dtype=np.float64)), hip.put(np.asarray(tg["dir"], dtype=np.float64))
        vol = hip.put(np.asarray(tg["vol"], dtype=np.float64))
        obj = hip.put(np.zeros(1))
        t = api.SaTarget()
        t.tgtElmNum = np.asarray(tg["vol"]).size
        t.cenPosMat, t.uniDirMat, t.elmVolVec = C.cast(cen, api._PD), C.cast(dr, api._PD), C.cast(vol, api._PD)
        t.cenKnlType = prob["vfd"]["cenKnlType"].encode()
        t.cenKnlWidth = prob["vfd"]["cenKnlWidth"]
        t.dirKnlType = prob["vfd"]["dirKnlType"].encode()
        t.dirKnlWidth = prob["vfd"]["dirKnlWidth"]
        rc = api.lib().sa_varifold_objective_noalloc(d, q.shape[1], C.cast(qd, api._PD),
                                                     np.asarray(prob["elm"]["vfdVtxMat"]).shape[1], C.cast(vf, api._PI),
                                                     C.byref(t), C.cast(obj, api._PD), None)
        return rc, float(hip.get(obj, 1)[0])
    finally:
        hip.free()


def probe_objective(prob, q):
    s = api.Session(prob, precision="fp64")
    out = s.probe("varifold", prob["tauVar"][:, 0], prob["tauPrm"], q)
    s.close()
    return out[0]


def setup_module(module):
    L = api.lib()
    L.sa_varifold_objective_noalloc.argtypes = [C.c_int, C.c_int, api._PD, C.c_int, api._PI, C.POINTER(api.SaTarget),
                                                api._PD, C.c_void_p]
    L.sa_last_error.restype = C.c_char_p


def test_noalloc_matches_session_objective_twice():
    big = W.shell_problem(level=4, n_layers=3, time_steps=2)       # 5,120 facets each side
    small = W.ring_problem(n_rings=5, n_per_ring=90)               # 2D, 90 segments
    for p in (big, small):
        q = p["q0"] + 0.01 * np.random.default_rng(3).standard_normal(p["q0"].shape)
        print(f"[gpu] noalloc dim={p['dim']} F={p['elm']['vfdVtxMat'].shape[1]}", flush=True)
        rc, got = noalloc_objective(p, q)
        assert rc == 0, api.lib().sa_last_error().decode()
        ref = probe_objective(p, q)
        print(f"[gpu] noalloc {got!r} session {ref!r}", flush=True)
        assert got == ref


def test_noalloc_limits():
    p = W.shell_problem(level=1, n_layers=3, time_steps=2)
    bad = dict(p["elm"])
    vf = bad["vfdVtxMat"].copy()
    vf[0, 0] = p["q0"].shape[1] + 5
    bad["vfdVtxMat"] = vf
    rc, got = noalloc_objective(dict(p, elm=bad), p["q0"])
    assert rc == 0 and np.isnan(got)
    # facet sets past the fixed scratch are refused before any launch
    big = dict(p, target={k: np.tile(v, (1, 70000)) if v.ndim == 2 else np.tile(v, 70000)
                          for k, v in p["target"].items()})
    rc, _ = noalloc_objective(big, p["q0"])
    assert rc == -1 and "scratch" in api.lib().sa_last_error().decode()
