"""Index-table builders against a literal (loop-for-loop) restatement of the
MATLAB builders on tiny meshes, closed-form counts, and mesh invariants."""
import numpy as np
import pytest

from shape_atrophy_amd import mesh as M
from shape_atrophy_amd import workloads as W


def literal_2d(L, C, vfd):
    """generateElementObject2D.m:1-151 restated with its own loops and `find`
    (1-based, column-major), for tiny sizes only."""
    lmk = L * C
    nE = (L - 1) * (C - 1) * 2
    elm = np.zeros((3, nE), dtype=int)
    e = 0
    for lc in range(1, C):
        for ll in range(1, L):
            i = (lc - 1) * L + ll
            elm[:, e] = i + np.array([0, 1, L])
            elm[:, e + 1] = i + np.array([1, L + 1, L])
            e += 2
    cnt = np.array([(elm == v).sum() for v in range(1, lmk + 1)])
    vtxElm = np.zeros((1 + 2 * cnt.max(), lmk), dtype=int)
    for v in range(1, lmk + 1):
        lcl, el = np.nonzero((elm == v).T)[::-1]   # column-major find order: by element, then local
        order = np.lexsort((lcl, el))
        lcl, el = lcl[order], el[order]
        vtxElm[0, v - 1] = len(lcl)
        vtxElm[1:1 + 2 * len(lcl), v - 1] = np.stack([el, lcl], 1).reshape(-1)
    edges = set()
    for k in range(nE):
        t = elm[:, k]
        for a in range(3):
            for b in range(a + 1, 3):
                edges.add((min(t[a], t[b]), max(t[a], t[b])))
    edg = np.array(sorted(edges))
    fv = np.concatenate([np.stack([np.arange(1, lmk + 1)] * 2), edg.T, edg[:, ::-1].T], axis=1)
    Z = fv.shape[1]
    ifo = np.zeros((2 + 2 * cnt.max(), Z), dtype=int)
    ifo[1, :lmk] = 2
    ifo[1, lmk:] = 1
    glb = np.arange(1, nE * 6 + 1).reshape(6, nE).T
    slot = {(1, 1): 1, (1, 2): 2, (2, 1): 2, (1, 3): 3, (3, 1): 3, (2, 2): 4, (2, 3): 5, (3, 2): 5, (3, 3): 6}
    for k in range(nE):
        for l1 in range(1, 4):
            for l2 in range(1, 4):
                a, b = elm[l1 - 1, k], elm[l2 - 1, k]
                col = np.nonzero((fv[0] == a) & (fv[1] == b))[0][0]
                adj = ifo[0, col]
                ifo[2 + 2 * adj, col] = k
                ifo[3 + 2 * adj, col] = glb[k, slot[(l1, l2)] - 1] - 1
                ifo[0, col] = adj + 1
    order = np.lexsort((fv[1], fv[0]))
    fv = fv[:, order]
    ifo = ifo[:, order]
    # frame stencils of each element pair (generateElementObject2D.m:13-29)
    tan = np.zeros((4, nE), dtype=int)
    tsv = np.zeros((4, nE), dtype=int)
    e = 0
    for lc in range(1, C):
        for ll in range(1, L):
            i = (lc - 1) * L + ll
            tan[:, e] = tan[:, e + 1] = i + np.array([0, 1, L, L + 1])
            tsv[:, e] = tsv[:, e + 1] = i + np.array([0, L, 1, L + 1])
            e += 2
    # bottom segments and their vertex adjacency (:61-73)
    btm = np.stack([np.arange(1, L), np.arange(2, L + 1)])
    bcnt = np.array([(btm == v).sum() for v in range(1, L + 1)])
    vtxBtm = np.zeros((1 + 2 * bcnt.max(), L), dtype=int)
    for v in range(1, L + 1):
        lcl, bi = np.nonzero((btm == v).T)[::-1]
        order = np.lexsort((lcl, bi))
        lcl, bi = lcl[order], bi[order]
        vtxBtm[0, v - 1] = len(lcl)
        vtxBtm[1:1 + 2 * len(lcl), v - 1] = np.stack([bi, lcl], 1).reshape(-1)
    # varifold segments of the listed layers (:127-133)
    glbL = np.arange(1, lmk + 1).reshape(C, L).T          # glbIdxMat(lyrLmkIdx, lyrCntIdx)
    vfdM = np.concatenate([np.stack([glbL[:L - 1, c - 1], glbL[1:, c - 1]]) for c in vfd], axis=1)
    return dict(elmVtxMat=elm - 1, vtxElmMat=vtxElm, femVtxMat=fv - 1, femIfoMat=ifo, tanVtxMat=tan - 1,
                tsvVtxMat=tsv - 1, btmVtxMat=btm - 1, vtxBtmMat=vtxBtm, vfdVtxMat=vfdM - 1)


@pytest.mark.parametrize("L,C", [(4, 3), (5, 4), (3, 6)])
def test_2d_tables_match_literal_builder(L, C):
    got = M.generate_element_object_2d(L, C, [1, C])
    ref = literal_2d(L, C, [1, C])
    for k in ("elmVtxMat", "vtxElmMat", "femVtxMat", "femIfoMat", "tanVtxMat", "tsvVtxMat", "btmVtxMat",
              "vtxBtmMat", "vfdVtxMat"):
        assert got[k].dtype == np.int32
        np.testing.assert_array_equal(got[k], ref[k], err_msg=k)


def literal_tri2tetra(tri1, C):
    """tri2tetra.m:1-69 restated with its loop (1-based, column-major):
    tri1 is the 3 x T triangle table (1-based)."""
    lbl = (np.roll(tri1, -1, axis=0) - tri1) > 0      # circshift(triVtxMat, -1) - triVtxMat > 0
    L = tri1.max()
    T = tri1.shape[1]
    tet = np.zeros((4, 3 * T, C - 1), dtype=int)
    nml = np.zeros((6, 3 * T, C - 1), dtype=int)
    tsv = np.zeros((6, 3 * T, C - 1), dtype=int)
    for t in range(1, T + 1):
        lo = tri1[:, t - 1]
        up = lo + L
        key = tuple(int(b) for b in lbl[:, t - 1])
        if key == (0, 0, 1):
            rows = [[lo[0], lo[1], lo[2], up[0]], [lo[1], lo[2], up[0], up[1]], [lo[2], up[0], up[1], up[2]]]
        elif key == (0, 1, 0):
            rows = [[lo[0], lo[1], lo[2], up[2]], [up[2], lo[0], lo[1], up[0]], [lo[1], up[0], up[1], up[2]]]
        elif key == (0, 1, 1):
            rows = [[lo[0], lo[1], lo[2], up[0]], [lo[1], lo[2], up[0], up[2]], [up[2], up[0], lo[1], up[1]]]
        elif key == (1, 0, 0):
            rows = [[lo[0], lo[1], lo[2], up[1]], [lo[0], lo[2], up[2], up[1]], [up[1], up[2], lo[0], up[0]]]
        elif key == (1, 0, 1):
            rows = [[lo[0], lo[1], lo[2], up[1]], [lo[2], up[0], lo[0], up[1]], [up[0], lo[2], up[2], up[1]]]
        else:   # (1, 1, 0)
            rows = [[lo[0], lo[1], lo[2], up[2]], [up[2], lo[0], lo[1], up[1]], [up[1], up[2], lo[0], up[0]]]
        tet[:, 3 * (t - 1):3 * t, 0] = np.array(rows).T
        nml[:, 3 * (t - 1):3 * t, 0] = np.concatenate([lo, up])[:, None]
        tsv[:, 3 * (t - 1):3 * t, 0] = np.stack([lo, up], axis=1).reshape(6)[:, None]
    for k in range(1, C - 1):
        tet[:, :, k] = tet[:, :, k - 1] + L
        nml[:, :, k] = nml[:, :, k - 1] + L
        tsv[:, :, k] = tsv[:, :, k - 1] + L
    # reshape(..., 4, 3 T (C - 1)): column-major, layer slabs one after another
    return (tet.reshape(4, -1, order="F"), nml.reshape(6, -1, order="F"), tsv.reshape(6, -1, order="F"))


def literal_3d(tri1, C, vfd):
    """generateElementObject3D.m:1-128 (without the freeBoundary tables)
    restated with its loops and `find` on the tri2tetra tables."""
    elm, nml, tsv = literal_tri2tetra(tri1, C)
    lmk = elm.max()
    L = tri1.max()
    nE = elm.shape[1]

    def adjacency(tbl, num):
        cnt = np.array([(tbl == v).sum() for v in range(1, num + 1)])
        out = np.zeros((1 + 2 * cnt.max(), num), dtype=int)
        for v in range(1, num + 1):
            lcl, el = np.nonzero((tbl == v).T)[::-1]
            order = np.lexsort((lcl, el))
            lcl, el = lcl[order], el[order]
            out[0, v - 1] = len(lcl)
            out[1:1 + 2 * len(lcl), v - 1] = np.stack([el, lcl], 1).reshape(-1)
        return out, cnt

    vtxElm, cnt = adjacency(elm, lmk)
    vtxBtm, _ = adjacency(tri1, L)
    edges = set()
    for k in range(nE):
        t = elm[:, k]
        for a in range(4):
            for b in range(a + 1, 4):
                edges.add((min(t[a], t[b]), max(t[a], t[b])))
    edg = np.array(sorted(edges))
    fv = np.concatenate([np.stack([np.arange(1, lmk + 1)] * 2), edg.T, edg[:, ::-1].T], axis=1)
    Z = fv.shape[1]
    ifo = np.zeros((2 + 2 * cnt.max(), Z), dtype=int)
    ifo[1, :lmk] = 2
    ifo[1, lmk:] = 1
    glb = np.arange(1, nE * 10 + 1).reshape(10, nE).T      # reshape(1 : (elmNum * 10), elmNum, 10)
    slot = {}
    for a, b, sl in [(1, 1, 1), (1, 2, 2), (1, 3, 3), (1, 4, 4), (2, 2, 5), (2, 3, 6), (2, 4, 7), (3, 3, 8), (3, 4, 9),
                     (4, 4, 10)]:
        slot[(a, b)] = slot[(b, a)] = sl
    for k in range(nE):
        for l1 in range(1, 5):
            for l2 in range(1, 5):
                a, b = elm[l1 - 1, k], elm[l2 - 1, k]
                col = np.nonzero((fv[0] == a) & (fv[1] == b))[0][0]
                adj = ifo[0, col]
                ifo[2 + 2 * adj, col] = k
                ifo[3 + 2 * adj, col] = glb[k, slot[(l1, l2)] - 1] - 1
                ifo[0, col] = adj + 1
    order = np.lexsort((fv[1], fv[0]))
    fv, ifo = fv[:, order], ifo[:, order]
    vfdM = np.concatenate([tri1 + (l - 1) * L for l in vfd], axis=1)
    return dict(elmVtxMat=elm - 1, vtxElmMat=vtxElm, nmlVtxMat=nml - 1, tsvVtxMat=tsv - 1, btmVtxMat=tri1 - 1,
                vtxBtmMat=vtxBtm, femVtxMat=fv - 1, femIfoMat=ifo, vfdVtxMat=vfdM - 1)


# six triangles on disjoint vertex triples, one in each vertex order, so
# every case of tri2tetra's split table occurs (icospheres miss one)
ALL_ORDERS = np.array([[0, 1, 2], [3, 5, 4], [7, 6, 8], [10, 11, 9], [14, 12, 13], [17, 16, 15]])


@pytest.mark.parametrize("level,C,vfd", [(0, 3, [3]), (1, 4, [1, 4]), (0, 5, [2, 5]), (None, 3, [1, 3])])
def test_3d_tables_match_literal_builder(level, C, vfd):
    """tri2tetra and generateElementObject3D, table by table against the
    literal restatement (the freeBoundary tables are MATLAB's own order and
    are checked by test_3d_element_object instead)."""
    tri = ALL_ORDERS if level is None else W.icosphere(level)[1]
    if level is None:
        lbl = (np.roll(tri.T, -1, axis=0) - tri.T) > 0
        assert len({tuple(c) for c in lbl.T.astype(int)}) == 6
    tet, nml, tsv = M.tri2tetra(tri, C)
    got = M.generate_element_object_3d(tet, tri, nml, tsv, vfd)
    ref = literal_3d(tri.T + 1, C, vfd)
    for k in ("elmVtxMat", "vtxElmMat", "nmlVtxMat", "tsvVtxMat", "btmVtxMat", "vtxBtmMat", "femVtxMat", "femIfoMat",
              "vfdVtxMat"):
        assert got[k].dtype == np.int32, k
        np.testing.assert_array_equal(got[k], ref[k], err_msg=k)


def test_demo_strip_closed_form_counts():
    e = M.generate_element_object_2d(101, 31, [1, 31])
    assert e["elmVtxMat"].shape == (3, 6000)
    assert e["vtxElmMat"].shape == (13, 3131)          # max adjacency 6
    n_edges = 100 * 31 + 101 * 30 + 100 * 30
    assert e["femVtxMat"].shape == (2, 3131 + 2 * n_edges) == (2, 21391)
    assert e["femIfoMat"].shape == (14, 21391)
    assert e["vfdVtxMat"].shape == (2, 200)
    assert e["btmVtxMat"].shape == (2, 100) and e["vtxBtmMat"].shape[1] == 101
    key = e["femVtxMat"][0].astype(np.int64) * 3131 + e["femVtxMat"][1]
    assert np.all(np.diff(key) > 0)                    # sorted, unique
    # diagonal entries: mul 2 and adjacency = vertex degree
    diag = e["femVtxMat"][0] == e["femVtxMat"][1]
    assert np.all(e["femIfoMat"][1][diag] == 2) and np.all(e["femIfoMat"][1][~diag] == 1)
    np.testing.assert_array_equal(e["femIfoMat"][0][diag], e["vtxElmMat"][0])


def test_adjacency_lists_ascending():
    e = M.generate_element_object_2d(7, 5, [5])
    V = 35
    for v in range(V):
        n = e["vtxElmMat"][0, v]
        els = e["vtxElmMat"][1:1 + 2 * n:2, v]
        assert np.all(np.diff(els) > 0)
        for el, l in zip(els, e["vtxElmMat"][2:2 + 2 * n:2, v]):
            assert e["elmVtxMat"][l, el] == v


def _tet_faces(tet):
    f = []
    for drop in range(4):
        f.append(np.sort(np.delete(tet, drop, axis=1), axis=1))
    return np.concatenate(f)


@pytest.mark.parametrize("level", [0, 1, 2])
def test_tri2tetra_conforming_shell(level):
    sph, tri = W.icosphere(level)
    C = 3
    tet, nml, tsv = M.tri2tetra(tri, C)
    L = sph.shape[0]
    assert tet.shape == (3 * tri.shape[0] * (C - 1), 4)
    # each prism is split into 3 tets over its 6 vertices
    for p in range(tri.shape[0]):
        verts = set(tri[p]) | set(tri[p] + L)
        assert set(tet[3 * p:3 * p + 3].reshape(-1)) == verts
    faces, counts = np.unique(_tet_faces(tet), axis=0, return_counts=True)
    assert counts.max() == 2
    # a conforming shell of a closed surface has exactly the inner and outer
    # surfaces as boundary
    assert (counts == 1).sum() == 2 * tri.shape[0]
    # prism volumes are covered exactly: tet volumes sum to the shell volume
    pos = np.concatenate([s * sph for s in (0.8, 0.9, 1.0)])
    q = pos[tet]
    vol = np.abs(np.einsum("ij,ij->i", q[:, 1] - q[:, 0], np.cross(q[:, 2] - q[:, 0], q[:, 3] - q[:, 0]))) / 6
    # compare with the volume between the polyhedral surfaces
    def polyvol(s):
        v = s * sph[tri]
        return abs(np.einsum("ij,ij->i", v[:, 0], np.cross(v[:, 1], v[:, 2])).sum()) / 6
    assert abs(vol.sum() - (polyvol(1.0) - polyvol(0.8))) < 1e-12
    # stencils: nml = [lower tri, upper tri], tsv = 3 (lower, upper) pairs
    np.testing.assert_array_equal(nml[:, 3:], nml[:, :3] + L)
    np.testing.assert_array_equal(tsv[:, 1::2], tsv[:, 0::2] + L)


def test_3d_element_object():
    p = W.shell_problem(level=1, n_layers=3, time_steps=2)
    e = p["elm"]
    V = p["q0"].shape[1]
    L = 42
    assert V == 3 * L
    np.testing.assert_array_equal(e["vfdVtxMat"], e["btmVtxMat"] + 2 * L)   # vfd layer 3
    Z = e["femVtxMat"].shape[1]
    key = e["femVtxMat"][0].astype(np.int64) * V + e["femVtxMat"][1]
    assert np.all(np.diff(key) > 0)
    assert e["femIfoMat"].shape[0] == 2 + 2 * int(e["vtxElmMat"][0].max())
    # every element contributes to exactly 16 (row, col) pairs
    assert e["femIfoMat"][0].sum() == 16 * e["elmVtxMat"].shape[1]
    assert Z == V + 2 * (Z - V) // 2


def test_c2_shell_sizes():
    sph, tri = W.icosphere(5)
    assert sph.shape[0] == 10242 and tri.shape[0] == 20480
    tet, _, _ = M.tri2tetra(tri, 5)
    assert tet.shape[0] == 245760 and tet.max() + 1 == 51210


def test_demo_initial_shape_fixture():
    """The strip of getInitialShape.m against the reference's own
    demo/initialShape.mat (ndeIniMat), kept as a fixture."""
    import os
    d = np.load(os.path.join(os.path.dirname(__file__), "golden", "initial_shape.npz"))
    p = W.strip_problem()
    assert int(d["lyrNdeNum"]) == 101 and int(d["lyrCntNum"]) == 31
    np.testing.assert_array_equal(p["q0"], d["ndeIniMat"])
