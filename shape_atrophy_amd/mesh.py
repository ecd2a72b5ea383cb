"""Index-table builders for layered P1 meshes.

Every table is returned in the column-major layout the MEX entry points read:
a MATLAB table of size R x C (rows = items, columns = slots) is held here as a
C-contiguous numpy array of shape (C, R), so entry (r, c) sits at flat offset
c * R + r.  All indices are int32 and 0-based, exactly as the MATLAB builders
hand them to the MEX files (`int32(tbl' - 1)` for vertex/element lists,
`int32(tbl')` for the count-prefixed adjacency tables).

Algorithms follow:
  CUDA/generateElementObject2D.m:1-151   (strip mesh, adjacency, FEM sparsity)
  CUDA/generateElementObject3D.m:1-128   (tet mesh adjacency, FEM sparsity)
  CUDA/tri2tetra.m:1-69                  (prism -> 3 tets split by index parity)
The builders here are vectorised (sort/group) instead of the O(E*Z) `find`
loops of the .m files; the produced tables are the same integers.
"""
from __future__ import annotations

import numpy as np

__all__ = [
    "vertex_adjacency",
    "fem_sparsity",
    "free_boundary",
    "orient_boundary",
    "generate_element_object_2d",
    "generate_element_object_3d",
    "tri2tetra",
    "upper_slot_table",
]


def upper_slot_table(nvtx: int) -> np.ndarray:
    """slot[a, b] = index of the unordered local pair (a, b) in the row-major
    upper triangle (00,01,02,11,12,22 for triangles; 00..33 for tets).
    Mirrors the glbIdxMat case ladder of generateElementObject2D.m:95-117 and
    generateElementObject3D.m:66-99."""
    slot = np.zeros((nvtx, nvtx), dtype=np.int64)
    k = 0
    for a in range(nvtx):
        for b in range(a, nvtx):
            slot[a, b] = slot[b, a] = k
            k += 1
    return slot


def vertex_adjacency(items: np.ndarray, num: int) -> np.ndarray:
    """Count-prefixed incidence table, MATLAB layout (1 + 2*maxAdj, num).

    items: (N, k) int array, row n lists the vertices of item n (element or
    facet).  Column v of the result is [adjNum, (item, local)...] with items in
    ascending order, the order MATLAB's column-major `find` returns
    (generateElementObject2D.m:32-41).  Only vertices < num are listed
    (vtxBtmMat is built over the first layer only, :61-72)."""
    items = np.asarray(items, dtype=np.int64)
    n_items, k = items.shape
    flat = items.reshape(-1)                      # offset = item * k + local
    keep = flat < num
    pos = np.nonzero(keep)[0]
    vtx = flat[pos]
    order = np.argsort(vtx, kind="stable")        # stable: ascending (item, local)
    vtx = vtx[order]
    pos = pos[order]
    counts = np.bincount(vtx, minlength=num)
    max_adj = int(counts.max()) if counts.size else 0
    tbl = np.zeros((1 + 2 * max_adj, num), dtype=np.int32)
    tbl[0] = counts
    starts = np.concatenate(([0], np.cumsum(counts)[:-1]))
    rank = np.arange(vtx.size) - starts[vtx]
    tbl[1 + 2 * rank, vtx] = pos // k
    tbl[2 + 2 * rank, vtx] = pos % k
    return tbl


def _edges(elm: np.ndarray, num: int) -> np.ndarray:
    """Unique undirected edges (i < j) of a simplicial mesh, sorted."""
    nv = elm.shape[1]
    pairs = []
    for a in range(nv):
        for b in range(a + 1, nv):
            pairs.append(np.stack([elm[:, a], elm[:, b]], axis=1))
    p = np.concatenate(pairs, axis=0)
    p = np.sort(p, axis=1)
    key = np.unique(p[:, 0] * num + p[:, 1])
    return np.stack([key // num, key % num], axis=1)


def fem_sparsity(elm: np.ndarray, num: int):
    """femVtxMat (2, Z) and femIfoMat (2 + 2*maxAdj, Z), MATLAB layout.

    Z = num + 2 * #edges nonzeros sorted lexicographically by (row, col)
    (`sortrows`, generateElementObject2D.m:124-126).  Column z of femIfoMat is
    [adjNum, mul, (elm, slot*E + elm)...]: mul = 2 on the diagonal, 1 off it,
    and the contributing elements in ascending order, each with the flat index
    of its local upper-triangle entry (generateElementObject2D.m:82-122,
    generateElementObject3D.m:56-104)."""
    elm = np.asarray(elm, dtype=np.int64)
    n_elm, nv = elm.shape
    edges = _edges(elm, num)
    rows = np.concatenate([np.arange(num), edges[:, 0], edges[:, 1]])
    cols = np.concatenate([np.arange(num), edges[:, 1], edges[:, 0]])
    key_sorted = np.sort(rows * num + cols)
    nzr = key_sorted.size
    fem_vtx = np.stack([key_sorted // num, key_sorted % num]).astype(np.int32)

    slot = upper_slot_table(nv)
    # every (element, l1, l2) contribution, in element-major order
    e_idx = np.repeat(np.arange(n_elm), nv * nv)
    l1 = np.tile(np.repeat(np.arange(nv), nv), n_elm)
    l2 = np.tile(np.tile(np.arange(nv), nv), n_elm)
    r = elm[e_idx, l1]
    c = elm[e_idx, l2]
    z = np.searchsorted(key_sorted, r * num + c)
    order = np.argsort(z, kind="stable")          # ascending element within z
    z = z[order]
    e_s = e_idx[order]
    s_s = slot[l1[order], l2[order]]
    counts = np.bincount(z, minlength=nzr)
    vcount = np.bincount(elm.reshape(-1), minlength=num)
    max_adj = int(vcount.max())
    ifo = np.zeros((2 + 2 * max_adj, nzr), dtype=np.int32)
    ifo[0] = counts
    ifo[1] = np.where(fem_vtx[0] == fem_vtx[1], 2, 1)
    starts = np.concatenate(([0], np.cumsum(counts)[:-1]))
    rank = np.arange(z.size) - starts[z]
    ifo[2 + 2 * rank, z] = e_s
    ifo[3 + 2 * rank, z] = s_s * n_elm + e_s
    return fem_vtx, ifo


def free_boundary(elm: np.ndarray) -> np.ndarray:
    """Facets that belong to exactly one element, (nBdr, d) sorted rows.

    Stands in for MATLAB `freeBoundary`; the reference's active code never
    reads bdrVtxMat/vtxBdrMat (the boundary-yank kernel is commented out,
    computeExternalYank.cu:49-84), so only the set matters here; once
    positions exist, orient_boundary turns every facet outward, the
    orientation that kernel assumes."""
    elm = np.asarray(elm, dtype=np.int64)
    nv = elm.shape[1]
    faces = []
    for drop in range(nv):
        keep = [a for a in range(nv) if a != drop]
        faces.append(elm[:, keep])
    f = np.sort(np.concatenate(faces, axis=0), axis=1)
    uniq, cnt = np.unique(f, axis=0, return_counts=True)
    return uniq[cnt == 1]


def orient_boundary(elm_obj: dict, pos: np.ndarray) -> dict:
    """Orient the boundary facets of elm_obj outward for positions pos (V, d)
    and rebuild vtxBdrMat: the facet normal of the boundary-yank kernel,
    n = (t_y, -t_x) with t = q_1 - q_0 in 2D and (q_1 - q_0) x (q_2 - q_0) in
    3D (computeExternalYank.cu:49-84 | 3D :58-99), points away from the
    element the facet belongs to (MATLAB freeBoundary's outward orientation).
    Facets keep their rows; a facet pointing inward has its last two vertices
    swapped."""
    pos = np.asarray(pos, dtype=np.float64)
    d = pos.shape[1]
    bdr = np.asarray(elm_obj["bdrVtxMat"], dtype=np.int64).T.copy()      # (nB, d)
    elm = np.asarray(elm_obj["elmVtxMat"], dtype=np.int64).T             # (E, d + 1)
    if bdr.size == 0:
        return elm_obj
    # the vertex opposite each boundary facet in its element: facets as
    # integer keys of their sorted vertices, matched by binary search
    V = pos.shape[0]
    def keys(f):
        f = np.sort(f, axis=1)
        out = np.zeros(f.shape[0], dtype=np.int64)
        for c in range(f.shape[1]):
            out = out * V + f[:, c]
        return out
    fk, fo = [], []
    for drop in range(d + 1):
        fk.append(keys(np.delete(elm, drop, axis=1)))
        fo.append(elm[:, drop])
    fk = np.concatenate(fk)
    fo = np.concatenate(fo)
    order = np.argsort(fk, kind="stable")
    fk, fo = fk[order], fo[order]
    bk = keys(bdr)
    at = np.searchsorted(fk, bk)
    if np.any(at >= fk.size) or np.any(fk[np.minimum(at, fk.size - 1)] != bk):
        raise ValueError("a boundary facet is not a facet of any element")
    opp = fo[at]
    p = [pos[bdr[:, k]] for k in range(d)]
    if d == 2:
        t = p[1] - p[0]
        n = np.stack([t[:, 1], -t[:, 0]], axis=1)
    else:
        n = np.cross(p[1] - p[0], p[2] - p[0])
    out = (sum(p) / d) - pos[opp]
    flip = (n * out).sum(axis=1) < 0
    bdr[flip, d - 2], bdr[flip, d - 1] = bdr[flip, d - 1].copy(), bdr[flip, d - 2].copy()
    elm_obj = dict(elm_obj)
    elm_obj["bdrVtxMat"] = np.ascontiguousarray(bdr.T, dtype=np.int32)
    elm_obj["vtxBdrMat"] = vertex_adjacency(bdr, pos.shape[0])
    return elm_obj


def _pack(elm_obj: dict) -> dict:
    for k, v in list(elm_obj.items()):
        if isinstance(v, np.ndarray):
            elm_obj[k] = np.ascontiguousarray(v, dtype=np.int32)
    return elm_obj


def generate_element_object_2d(lyr_lmk_num: int, lyr_cnt_num: int, vfd_lyr_vec,
                               periodic: bool = False) -> dict:
    """Strip (or, with periodic=True, ring) triangulation of lyr_cnt_num stacked
    layers of lyr_lmk_num landmarks; generateElementObject2D.m:1-151.

    Quad (layer l, column c) with i = l*L + c gives triangles [i, i+1, i+L] and
    [i+1, i+L+1, i+L], tangent pairs (i, i+1), (i+L, i+L+1) and transverse pairs
    (i, i+L), (i+1, i+L+1) shared by both triangles (:13-30).  periodic=True
    wraps column L-1 onto column 0; it is a build-defined extension used for
    concentric-ring meshes (the reference only builds open strips).
    vfd_lyr_vec holds 1-based layer numbers, as in MATLAB."""
    L, C = int(lyr_lmk_num), int(lyr_cnt_num)
    lmk_num = L * C
    ncol = L if periodic else L - 1
    lay = np.repeat(np.arange(C - 1), ncol)
    col = np.tile(np.arange(ncol), C - 1)
    i0 = lay * L + col
    i1 = lay * L + (col + 1) % L
    i2 = i0 + L
    i3 = i1 + L
    n_elm = 2 * (C - 1) * ncol
    elm = np.empty((n_elm, 3), dtype=np.int64)
    elm[0::2] = np.stack([i0, i1, i2], axis=1)
    elm[1::2] = np.stack([i1, i3, i2], axis=1)
    tan = np.repeat(np.stack([i0, i1, i2, i3], axis=1), 2, axis=0)
    tsv = np.repeat(np.stack([i0, i2, i1, i3], axis=1), 2, axis=0)

    bdr = free_boundary(elm)
    if periodic:
        btm = np.stack([np.arange(L), (np.arange(L) + 1) % L], axis=1)
    else:
        btm = np.stack([np.arange(L - 1), np.arange(1, L)], axis=1)
    vfd = []
    for lyr in np.atleast_1d(vfd_lyr_vec):
        base = (int(lyr) - 1) * L
        vfd.append(btm + base)
    vfd = np.concatenate(vfd, axis=0)
    fem_vtx, fem_ifo = fem_sparsity(elm, lmk_num)
    return _pack({
        "dim": 2,
        "elmVtxMat": elm.T,
        "vtxElmMat": vertex_adjacency(elm, lmk_num),
        "tanVtxMat": tan.T,
        "tsvVtxMat": tsv.T,
        "bdrVtxMat": bdr.T,
        "vtxBdrMat": vertex_adjacency(bdr, lmk_num),
        "btmVtxMat": btm.T,
        "vtxBtmMat": vertex_adjacency(btm, L),
        "femVtxMat": fem_vtx,
        "femIfoMat": fem_ifo,
        "vfdVtxMat": vfd.T,
        "lmkNum": lmk_num,
        "lyrLmkNum": L,
        "lyrCntNum": C,
        "elmLyrVec": np.repeat(np.arange(C - 1), 2 * ncol),
    })


def tri2tetra(tri_vtx: np.ndarray, lyr_cnt_num: int):
    """Split the prisms between stacked copies of a surface triangulation into
    3 tets each; tri2tetra.m:1-69.

    tri_vtx: (T, 3) 0-based vertex indices of one layer (L = max + 1 vertices
    per layer; layer k holds vertices k*L .. k*L + L-1).  The split of each
    prism is chosen by which cyclic neighbours have the larger index
    (lblMat, :3), so quad faces shared by two prisms are cut along the same
    diagonal.  Returns tet (E, 4), nml (E, 6) = [lower tri, upper tri] and
    tsv (E, 6) = 3 (lower, upper) pairs, E = 3 * T * (lyr_cnt_num - 1)."""
    tri = np.asarray(tri_vtx, dtype=np.int64)
    L = int(tri.max()) + 1
    T = tri.shape[0]
    a, b, c = tri[:, 0], tri[:, 1], tri[:, 2]
    A, B, Cc = a + L, b + L, c + L
    lbl = np.stack([b > a, c > b, a > c], axis=1).astype(np.int64)
    code = lbl[:, 0] * 4 + lbl[:, 1] * 2 + lbl[:, 2]
    tets = np.empty((T, 3, 4), dtype=np.int64)
    # (lbl1,lbl2,lbl3) -> the three tets of tri2tetra.m:16-52
    cases = {
        0b001: lambda: [[a, b, c, A], [b, c, A, B], [c, A, B, Cc]],
        0b010: lambda: [[a, b, c, Cc], [Cc, a, b, A], [b, A, B, Cc]],
        0b011: lambda: [[a, b, c, A], [b, c, A, Cc], [Cc, A, b, B]],
        0b100: lambda: [[a, b, c, B], [a, c, Cc, B], [B, Cc, a, A]],
        0b101: lambda: [[a, b, c, B], [c, A, a, B], [A, c, Cc, B]],
        0b110: lambda: [[a, b, c, Cc], [Cc, a, b, B], [B, Cc, a, A]],
    }
    for k, fn in cases.items():
        m = code == k
        if not m.any():
            continue
        t = np.array([[np.asarray(x) for x in row] for row in fn()])  # (3,4,T)
        tets[m] = np.transpose(t, (2, 0, 1))[m]
    if np.any((code == 0) | (code == 7)):
        raise ValueError("degenerate triangle (repeated vertex index)")
    tet1 = tets.reshape(3 * T, 4)
    nml1 = np.repeat(np.stack([a, b, c, A, B, Cc], axis=1), 3, axis=0)
    tsv1 = np.repeat(np.stack([a, A, b, B, c, Cc], axis=1), 3, axis=0)
    shift = (np.arange(lyr_cnt_num - 1) * L)[:, None, None]
    tet = (tet1[None] + shift).reshape(-1, 4)
    nml = (nml1[None] + shift).reshape(-1, 6)
    tsv = (tsv1[None] + shift).reshape(-1, 6)
    return tet, nml, tsv


def generate_element_object_3d(tet: np.ndarray, btm_vtx: np.ndarray, nml_vtx: np.ndarray,
                               tsv_vtx: np.ndarray, vfd_lyr_vec) -> dict:
    """generateElementObject3D.m:1-128 for a stacked-layer tet mesh.

    tet (E, 4), btm_vtx (T, 3) = first-layer triangulation, nml_vtx/tsv_vtx
    (E, 6) from tri2tetra; all 0-based.  vfd_lyr_vec holds 1-based layer
    numbers whose surface triangles form the varifold facets (:112-113)."""
    tet = np.asarray(tet, dtype=np.int64)
    btm = np.asarray(btm_vtx, dtype=np.int64)
    lmk_num = int(tet.max()) + 1
    L = int(btm.max()) + 1
    C = lmk_num // L
    bdr = free_boundary(tet)
    vfd = np.concatenate([btm + (int(l) - 1) * L for l in np.atleast_1d(vfd_lyr_vec)], axis=0)
    fem_vtx, fem_ifo = fem_sparsity(tet, lmk_num)
    n_elm = tet.shape[0]
    return _pack({
        "dim": 3,
        "elmVtxMat": tet.T,
        "vtxElmMat": vertex_adjacency(tet, lmk_num),
        "nmlVtxMat": np.asarray(nml_vtx).T,
        "tsvVtxMat": np.asarray(tsv_vtx).T,
        "bdrVtxMat": bdr.T,
        "vtxBdrMat": vertex_adjacency(bdr, lmk_num),
        "btmVtxMat": btm.T,
        "vtxBtmMat": vertex_adjacency(btm, L),
        "femVtxMat": fem_vtx,
        "femIfoMat": fem_ifo,
        "vfdVtxMat": vfd.T,
        "lmkNum": lmk_num,
        "lyrLmkNum": L,
        "lyrCntNum": C,
        "elmLyrVec": np.arange(n_elm) // (n_elm // (C - 1)),
    })
