"""Write tests/golden/c2_oracle.npz: the FP64 oracle's first two time steps
of the C2 workload (51,210-landmark ellipsoidal shell of 245,760 tets,
Matern nu = 3/2 on 5,000 control points, varifold against 20,480 target
triangles), for the full-size GPU parity test.  The oracle needs a few
minutes for these steps on 8 host cores, too long to repeat in every GPU
test run, so its outputs are kept here:

  * an input fingerprint (sha256 over the landmarks, tables, control points,
    target and parameters): the test refuses a fixture of other inputs;
  * the objective after the two steps and its gradient with respect to the
    final landmarks, the final landmarks and concentration, all on every
    SUB-th landmark;
  * full-vector checks: the norms of the displacement and concentration and
    their dot products with a seeded random vector.

    python tests/golden/make_c2_oracle.py
"""
from __future__ import annotations

import hashlib
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
OUT = os.path.join(HERE, "c2_oracle.npz")
SUB = 7            # every SUB-th landmark is stored
TIME_NUM = 3       # two time steps
CHECK_SEED = 123   # the random vector of the full-vector checks


def workload():
    sys.path.insert(0, ROOT)
    from shape_atrophy_amd import workloads as W
    p = W.shell_problem(level=5, n_layers=5, time_steps=50, n_ctl=5000)
    p["timeNum"] = TIME_NUM
    return p


def fingerprint(p: dict) -> str:
    h = hashlib.sha256()
    for k in sorted(p["elm"]):
        v = p["elm"][k]
        if isinstance(v, np.ndarray):
            h.update(k.encode())
            h.update(np.ascontiguousarray(v).tobytes())
    for a in (p["q0"], p["ctlIdx"], p["tauVar"], p["tauPrm"], p["target"]["cen"], p["target"]["dir"],
              p["target"]["vol"]):
        h.update(np.ascontiguousarray(a).tobytes())
    h.update(json.dumps({k: v for k, v in p["params"].items()}, sort_keys=True, default=str).encode())
    h.update(json.dumps(p["vfd"], sort_keys=True).encode())
    h.update(str(p["timeNum"]).encode())
    return h.hexdigest()


def checks(q0, qT, tauT, grad):
    w = np.random.default_rng(CHECK_SEED).standard_normal(q0.shape)
    return np.array([np.linalg.norm(qT - q0), float((w * (qT - q0)).sum()), np.linalg.norm(tauT),
                     float((w[0] * tauT).sum()), np.linalg.norm(grad), float((w * grad).sum())])


def main():
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    O.build()
    p = workload()
    t0 = time.time()
    o = O.objfcn(p, with_grad=True)
    assert o["status"] == 0, o["status"]
    q0, qT, tauT = p["q0"], o["lmkStk"][-1], o["tauMat"][-1]
    idx = np.arange(0, q0.shape[1], SUB)
    np.savez_compressed(OUT, fingerprint=np.array(fingerprint(p)), obj=np.array(o["obj"]), idx=idx,
                        qT=qT[:, idx], tauT=tauT[idx], grad=o["grad"][:, idx],
                        checks=checks(q0, qT, tauT, o["grad"]), pcg_iters=o["pcg_iters"],
                        rd_iters=o["rd_iters"])
    print(f"wrote {OUT}: obj {o['obj']:.15e}, pcg {o['pcg_iters']}, {time.time() - t0:.1f} s")


if __name__ == "__main__":
    main()
