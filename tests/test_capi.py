"""The C ABI library loads without a GPU, exports every function that
include/shape_atrophy.h declares, and rejects malformed inputs with SA_E_ARG
before touching the device."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from shape_atrophy_amd import api, workloads as W

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    txt = open(os.path.join(ROOT, "include", "shape_atrophy.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(sa_[a-z_0-9]+)\s*\(", txt)))


def test_header_symbols_exported():
    lib = api.lib()
    names = declared_functions()
    assert {"sa_deformation", "sa_compute_objfcn", "sa_session_create", "sa_session_run"} <= set(names)
    for n in names:
        assert hasattr(lib, n), n


def test_invalid_tables_rejected_without_device():
    p = W.ring_problem(n_rings=4, n_per_ring=12)
    elm = dict(p["elm"])
    bad = elm["elmVtxMat"].copy()
    bad[0, 3] = 10_000
    elm["elmVtxMat"] = bad
    args = api.problem_args(dict(p, elm=elm))
    with pytest.raises(api.SaError) as ei:
        api.computeObjfcn(*args)
    assert ei.value.code == -1 and "out of range" in str(ei.value)


def test_unknown_elasticity_and_dimension():
    p = W.ring_problem(n_rings=4, n_per_ring=12)
    args = api.problem_args(p)
    args[2] = "neohookean"
    with pytest.raises(ValueError, match="Unknown elasticity type"):
        api.computeObjfcn(*args)
    args = api.problem_args(p)
    args[3] = np.zeros((4, p["q0"].shape[1]))
    with pytest.raises(ValueError, match="Dimension 4 is not supported"):
        api.computeObjfcn(*args)


def test_unsorted_fem_table_rejected():
    p = W.ring_problem(n_rings=4, n_per_ring=12)
    elm = dict(p["elm"])
    fv = elm["femVtxMat"].copy()
    fv[:, [3, 4]] = fv[:, [4, 3]]
    elm["femVtxMat"] = fv
    with pytest.raises(api.SaError, match="not sorted"):
        api.computeObjfcn(*api.problem_args(dict(p, elm=elm)))


def test_server_failure_path_without_device():
    """An invalid mesh is refused when it is registered (the sizing pass
    validates its tables); a request whose session cannot be created reports
    its error instead of hanging (here: no GPU in this container)."""
    p = W.ring_problem(n_rings=4, n_per_ring=12)
    elm = dict(p["elm"])
    bad = elm["elmVtxMat"].copy()
    bad[0, 0] = 99999
    elm["elmVtxMat"] = bad
    srv = api.Server(workers=1, mem_budget=1 << 30)
    with pytest.raises(api.SaError, match="range"):
        srv.add_mesh(dict(p, elm=elm))
    mid = srv.add_mesh(p)
    g = srv.wait(srv.submit(mid, p["tauVar"][:, 0], p["tauPrm"]))
    if api.lib().sa_device_count() == 0:
        assert g["rc"] == -3 and g["error"], g
    srv.close()


def test_footprint_sizing_without_device():
    """sa_session_footprint sizes a session without touching the device: it
    grows with the time levels and the control points, counts the stored
    all-landmark kernel matrix, and rejects an invalid mesh with -1."""
    keep = api._Keep()
    p = W.ring_problem(n_rings=4, n_per_ring=12)
    def fp(prob, precision="fp64", T=None):
        k = api._Keep()
        elm = dict(prob["elm"])
        if prob.get("ctlIdx") is not None:
            elm["ctlIdxVec"] = prob["ctlIdx"]
        m = api.make_mesh(elm, np.ascontiguousarray(prob["q0"]), k, with_vfd=True)
        t = api.make_target(prob["target"], prob["vfd"], k)
        prm = api.params_struct(prob["params"], k, precision, T)
        return api.lib().sa_session_footprint(C.byref(m), C.byref(t), C.byref(prm))
    V = p["q0"].shape[1]
    a, b = fp(p, T=5), fp(p, T=9)
    # four more time levels: positions and concentration of each (8 B x (D + 1) x V)
    assert b - a == 4 * 8 * 3 * V
    # the all-landmark form stores K (V rows padded to 128-row panels) and its slice partials
    assert a > 8 * 128 * V
    pc = dict(p, ctlIdx=np.arange(0, V, 4, dtype=np.int32))
    assert fp(pc, "mixed") > 0 and fp(pc, "fp64") > 0
    elm = dict(p["elm"])
    bad = elm["elmVtxMat"].copy()
    bad[0, 0] = 99999
    elm["elmVtxMat"] = bad
    assert fp(dict(p, elm=elm)) == -1
    del keep


@pytest.mark.parametrize("cname,pyname", [("sa_mesh", "SaMesh"), ("sa_params", "SaParams"),
                                          ("sa_target", "SaTarget"), ("sa_stats", "SaStats")])
def test_struct_layout_matches_header(cname, pyname, tmp_path):
    """The ctypes mirrors in api.py have the size and field offsets that the C
    compiler gives the header's structs (a field added on one side only would
    shift every later field)."""
    py = getattr(api, pyname)
    fields = [f[0] for f in py._fields_]
    src = tmp_path / "layout.c"
    lines = ["#include <stddef.h>", "#include <stdio.h>", '#include "shape_atrophy.h"', "int main(void) {",
             f'    printf("size %zu\\n", sizeof({cname}));']
    lines += [f'    printf("{f} %zu\\n", offsetof({cname}, {f}));' for f in fields]
    lines += ["    return 0;", "}"]
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    import subprocess
    subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = dict(l.split() for l in subprocess.run([str(exe)], check=True, capture_output=True,
                                                 text=True).stdout.splitlines())
    assert int(out["size"]) == C.sizeof(py), (cname, out["size"], C.sizeof(py))
    for f in fields:
        assert int(out[f]) == getattr(py, f).offset, (cname, f)


def test_tile_table_lookup(tmp_path):
    """The tile table: the entry of the product and element type nearest the
    shape (log distance of rows and columns), comments and malformed lines
    skipped, no entry -> 0 (the built-in default), SA_KS_TARGET over all.
    The table is read once per process, so each case runs in its own."""
    import subprocess
    import sys
    tbl = tmp_path / "tiles.txt"
    tbl.write_text("# product type rows cols target\n"
                   "xc f32 51210 5000 2048   # measured\n"
                   "xc f32 2568 256 512\n"
                   "xc f64 51210 5000 1024\n"
                   "cx f32 5000 51210 4096\n"
                   "kk f64 2000 2000 256\n"
                   "op f64 245760 51210 3\n"
                   "op f64 11520 2568 0\n"
                   "op f64 100 100 4\n"          # no such launch shape: skipped
                   "xc f32 100 100 7\n"          # target out of range: skipped
                   "zz f32 100 100 512\n"        # unknown product: skipped
                   "garbage\n")
    code = ("from shape_atrophy_amd import api; import json; print(json.dumps(["
            "api.tile_target('xc', True, 51210, 5000), api.tile_target('xc', True, 40000, 4000),"
            "api.tile_target('xc', True, 3000, 300), api.tile_target('xc', True, 100, 100),"
            "api.tile_target('xc', False, 2568, 256), api.tile_target('cx', True, 2000, 20000),"
            "api.tile_target('cx', False, 5000, 51210), api.tile_target('kk', False, 50000, 50000),"
            "api.tile_target('cc', True, 100, 100), api.tile_target('qq', True, 10, 10),"
            "api.tile_target('op', False, 983040, 204810), api.tile_target('op', False, 3600, 2000)]))")

    def run(**env):
        e = {k: v for k, v in os.environ.items() if k != "SA_KS_TARGET"}
        e.update(SA_TILE_TABLE=str(tbl), **env)
        out = subprocess.run([sys.executable, "-c", code], cwd=ROOT, env=e, check=True, capture_output=True,
                             text=True).stdout
        return __import__("json").loads(out.strip().splitlines()[-1])

    assert run() == [2048, 2048, 512, 512, 1024, 4096, 0, 256, 0, -1, 3, 0]
    # the block-target override leaves the operator's launch shape to the table
    assert run(SA_KS_TARGET="300") == [300, 300, 300, 300, 300, 300, 300, 300, 300, -1, 3, 0]
    assert run(SA_KS_TARGET="cx=640,kk=128") == [2048, 2048, 512, 512, 1024, 640, 640, 128, 0, -1, 3, 0]
    assert run(SA_OPF_CFG="2")[-2:] == [2, 2]


def test_committed_tile_table():
    """Every line of the tile table beside the library is well formed, and
    the library takes its entries: a shape in the table gets its own
    target."""
    path = os.path.join(ROOT, "shape_atrophy_amd", "tiles.txt")
    entries = []
    for line in open(path):
        line = line.split("#")[0].split()
        if not line:
            continue
        prod, ty, rows, cols, tgt = line
        assert prod in ("xc", "cx", "cc", "kk", "op") and ty in ("f32", "f64"), line
        assert int(rows) > 0 and int(cols) > 0, line
        assert (0 <= int(tgt) < 4) if prod == "op" else (16 <= int(tgt) <= 65536), line
        entries.append((prod, ty == "f32", int(rows), int(cols), int(tgt)))
    assert entries
    if all(os.environ.get(k) is None for k in ("SA_TILE_TABLE", "SA_KS_TARGET", "SA_OPF_CFG")):
        for prod, f32, rows, cols, tgt in entries:
            same = [e[4] for e in entries if e[:4] == (prod, f32, rows, cols)]
            assert api.tile_target(prod, f32, rows, cols) == same[0], (prod, f32, rows, cols)


def op_chunks(p):
    keep = api._Keep()
    q = np.ascontiguousarray(p["q0"], dtype=np.float64)   # dim x lmkNum row-major = lmkNum x dim column-major
    m = api.make_mesh(dict(p["elm"]), q, keep, with_vfd=False)
    n, u, h = C.c_int(), C.c_longlong(), C.c_longlong()
    rc = api.lib().sa_op_chunk_check(C.byref(m), q.ctypes.data_as(C.POINTER(C.c_double)), C.byref(n), C.byref(u),
                                     C.byref(h))
    return rc, n.value, u.value, h.value


@pytest.mark.parametrize("cfg", ["0", "1", "2", "3"])
@pytest.mark.parametrize("which", ["ring", "shell", "slab", "c2"])
def test_operator_chunk_tables(which, cfg, monkeypatch):
    """The vertex chunks of the fused operator apply, built on the host from
    the mesh as a session builds them, pass the checks the kernel's indexing
    relies on (every vertex in one chunk; every list entry one LDS slot,
    written once), and each element is evaluated at most as often as it is
    listed, at least once."""
    p = {"ring": lambda: W.ring_problem(),
         "shell": lambda: W.shell_problem(level=3, n_layers=4, time_steps=2),
         "slab": lambda: W.slab_problem(n=9, n_layers=3, tilt=0.3, time_steps=2),
         "c2": lambda: W.shell_problem(level=5, n_layers=5, time_steps=2)}[which]()
    monkeypatch.setenv("SA_OPF_CFG", cfg)
    rc, n, u, h = op_chunks(p)
    E = p["elm"]["elmVtxMat"].shape[1]
    S = int(p["elm"]["vtxElmMat"][0].sum())
    assert rc == 0, api.lib().sa_last_error()
    assert n >= 1 and E <= u <= S and p["q0"].shape[1] <= h, (n, u, E, S, h)
    print(f"{which} launch shape {cfg}: {n} chunks, {u / E:.2f} evaluations per element, "
          f"{h / p['q0'].shape[1]:.2f} halo loads per vertex")
