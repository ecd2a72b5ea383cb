"""Seeded synthetic workloads: layered 2D rings / strips and 3D ellipsoidal
shells, their varifold targets and parameter sets.

The reference ships one input (demo/initialShape.mat, a 101 x 31 strip built
by demo/getInitialShape.m:1-15) and the demo parameters of
demo/run_demo.m:1-28.  The larger problems below are build-defined,
deterministic functions of their seeds:

  ring      concentric rings (periodic layered triangulation), 2D
  strip     the demo strip of getInitialShape.m
  shell     icosphere surface copied onto stacked scaled layers and split into
            tets by tri2tetra (CUDA/tri2tetra.m), 3D
"""
from __future__ import annotations

import math

import numpy as np

from . import mesh as _mesh

__all__ = ["icosphere", "ring_problem", "strip_problem", "shell_problem", "facet_target"]


# ---------------------------------------------------------------------------
# geometry helpers

def icosphere(level: int):
    """Unit icosphere: (verts (N, 3), tris (T, 3)), N = 10*4^level + 2.
    Vertex order: the 12 icosahedron vertices, then edge midpoints in the
    order they are created, level by level."""
    t = (1.0 + math.sqrt(5.0)) / 2.0
    v = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t),
         (0, -1, -t), (0, 1, -t), (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    verts = [np.array(p, dtype=float) / np.linalg.norm(p) for p in v]
    tris = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4),
            (11, 10, 2), (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8),
            (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    for _ in range(level):
        cache = {}
        new = []

        def mid(a, b):
            key = (a, b) if a < b else (b, a)
            if key not in cache:
                m = verts[a] + verts[b]
                verts.append(m / np.linalg.norm(m))
                cache[key] = len(verts) - 1
            return cache[key]

        for a, b, c in tris:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            new += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        tris = new
    return np.array(verts), np.array(tris, dtype=np.int64)


def facet_target(pos: np.ndarray, facets: np.ndarray) -> dict:
    """Centres, unit directions (2D) / unit normals (3D) and lengths / areas of
    facets, the quantities landmarksToVarifoldKernel computes
    (2D varifold.cu:9-36, 3D varifold.cu:9-43).  pos (N, d), facets (F, d)."""
    d = pos.shape[1]
    p = [pos[facets[:, k]] for k in range(d)]
    if d == 2:
        cen = 0.5 * (p[0] + p[1])
        t = p[1] - p[0]
        ln = np.sqrt((t * t).sum(1))
        return {"cen": cen.T.copy(), "dir": (t / ln[:, None]).T.copy(), "vol": ln}
    cen = (p[0] + p[1] + p[2]) / 3.0
    n = np.cross(p[1] - p[0], p[2] - p[0])
    ln = np.sqrt((n * n).sum(1))
    return {"cen": cen.T.copy(), "dir": (n / ln[:, None]).T.copy(), "vol": 0.5 * ln}


def _real_sh_field(xyz: np.ndarray, lmax: int, rng) -> np.ndarray:
    """Random combination of real spherical harmonics of degree 1..lmax
    evaluated at unit vectors xyz (N, 3); normalised to max |f| = 1."""
    from scipy.special import sph_harm_y
    theta = np.arccos(np.clip(xyz[:, 2], -1.0, 1.0))
    phi = np.arctan2(xyz[:, 1], xyz[:, 0])
    f = np.zeros(xyz.shape[0])
    for ell in range(1, lmax + 1):
        for m in range(-ell, ell + 1):
            y = sph_harm_y(ell, abs(m), theta, phi)
            basis = y.real if m >= 0 else y.imag
            f += rng.standard_normal() * basis
    return f / np.abs(f).max()


# ---------------------------------------------------------------------------
# parameter sets

def demo_params() -> dict:
    """demo/run_demo.m:1-28 (frame model, 2D)."""
    return dict(elaType="frame", modVec=[3.0, 3.0, 15.0], spdTanVal=0.625, spdTsvVal=0.125,
                ynkActPrmVec=[0.01, 1.0, 4.0, 0.5], reaActPrmVec=[0.01, 1.0, 4.0, 0.5, 0.3],
                ldmWgt=1e-2, knlOrder=3, knlWidth=0.2, knlEps=1e-10, btmWgt=0.0,
                timeStp=0.01, timeEnd=1.0, cgItrMax=10000, cgTolVal=1e-12)


def strip_problem(vfd_layers=(1, 31)) -> dict:
    """The demo strip: 101 x 31 landmarks on [-3,3] x [-1,1]
    (demo/getInitialShape.m:1-15), run_demo.m parameters and tau bump."""
    L, C = 101, 31
    # MATLAB linspace(d1, d2, n) = d1 + (0:n-1) * (d2 - d1) / (n - 1), which
    # rounds differently from numpy.linspace in the last bit
    x = -3.0 + np.arange(L) * 6.0 / (L - 1)
    y = -1.0 + np.arange(C) * 2.0 / (C - 1)
    pos = np.stack([np.tile(x, C), np.repeat(y, L)], axis=1)
    elm = _mesh.orient_boundary(_mesh.generate_element_object_2d(L, C, list(vfd_layers)), pos)
    return dict(dim=2, elm=elm, q0=pos.T.copy(), params=demo_params(),
                tauVar=np.array([[0.5], [0.5], [0.2], [0.3]]), tauPrm=np.array([1.0]))


def ring_problem(n_rings: int = 10, n_per_ring: int = 200, seed: int = 1, target_seed: int = 2,
                 r_inner: float = 0.1, r_outer: float = 1.0, jitter: float = 0.15,
                 elaType: str = "frame") -> dict:
    """Concentric-ring layered disk mesh: n_rings x n_per_ring landmarks,
    2 * (n_rings-1) * n_per_ring triangles, radii linear r_inner -> r_outer,
    vertices jittered by +-jitter of the local spacing.  Material layers are
    cut at centroid radius 0.6 / 0.8.  The outer ring is the varifold curve;
    the target is a circle of radius 1 with a Fourier perturbation of
    amplitude 0.05 (modes 2..8) sampled at n_per_ring segments."""
    rng = np.random.default_rng(seed)
    L, C = n_per_ring, n_rings
    r = np.linspace(r_inner, r_outer, C)
    th = 2.0 * np.pi * np.arange(L) / L
    dr = (r_outer - r_inner) / (C - 1)
    dth = 2.0 * np.pi / L
    rr = np.repeat(r, L) + jitter * dr * rng.uniform(-1, 1, L * C)
    tt = np.tile(th, C) + jitter * dth * rng.uniform(-1, 1, L * C) * 0.5
    # keep the outer boundary on its radius so the varifold curve is a circle
    rr[(C - 1) * L:] = r_outer
    pos = np.stack([rr * np.cos(tt), rr * np.sin(tt)], axis=1)
    elm = _mesh.orient_boundary(_mesh.generate_element_object_2d(L, C, [C], periodic=True), pos)
    cen = pos[elm["elmVtxMat"].T].mean(axis=1)
    rad = np.sqrt((cen ** 2).sum(1))
    elm["elmLyrVec"] = np.ascontiguousarray(np.digitize(rad, [0.6, 0.8]).astype(np.int32))
    trng = np.random.default_rng(target_seed)
    a = trng.standard_normal(7)
    b = trng.standard_normal(7)
    tth = 2.0 * np.pi * np.arange(L) / L
    pert = sum(a[k - 2] * np.cos(k * tth) + b[k - 2] * np.sin(k * tth) for k in range(2, 9))
    rt = 1.0 + 0.05 * pert / np.abs(pert).max()
    tpos = np.stack([rt * np.cos(tth), rt * np.sin(tth)], axis=1)
    tfac = np.stack([np.arange(L), (np.arange(L) + 1) % L], axis=1)
    target = facet_target(tpos, tfac)
    prm = dict(elaType=elaType, modVec=[3.0, 3.0, 15.0],
               modMat=[[1.0, 1.0], [2.0, 1.5], [4.0, 3.0]],
               spdTanVal=0.01, spdTsvVal=0.01,
               ynkActPrmVec=[0.01, 1.0, 4.0, 0.5], reaActPrmVec=[0.0, 1.0, 5.0 / 4.0, 0.1, 1.0],
               ldmWgt=1e-2, knlOrder=3, knlWidth=0.2, knlEps=1e-10, btmWgt=0.0,
               timeStp=0.05, timeEnd=1.0, cgItrMax=10000, cgTolVal=1e-10)
    # seeded interior blob centre in the middle band
    crng = np.random.default_rng(seed + 100)
    ang = crng.uniform(0, 2 * np.pi)
    cr = 0.7
    tauVar = np.array([[cr * np.cos(ang)], [cr * np.sin(ang)], [1.0], [0.3]])
    vfd = dict(cenKnlType="G", cenKnlWidth=0.1, dirKnlType="B", dirKnlWidth=1.0)
    return dict(dim=2, elm=elm, q0=pos.T.copy(), params=prm, tauVar=tauVar, tauPrm=np.array([1.0]),
                target=target, vfd=vfd)


def shell_problem(level: int = 5, n_layers: int = 5, scales=None, axes=(1.0, 0.8, 0.7),
                  target_seed: int = 3, blob_seed: int = 7, time_steps: int = 50,
                  n_ctl: int = 0, ctl_seed: int = 11, elaType: str = "frame") -> dict:
    """Ellipsoidal shell: icosphere(level) copied onto n_layers layers scaled
    0.8 -> 1.0 and split by tri2tetra.  level 5 gives V = 51,210 landmarks,
    E = 245,760 tets and F = 20,480 outer-surface varifold triangles.  The
    target is the outer surface moved along the radius by a random real
    spherical-harmonic field (degree <= 6, amplitude 0.05).  n_ctl > 0 picks
    that many seeded landmarks as kernel control points."""
    if scales is None:
        scales = np.linspace(0.8, 1.0, n_layers)
    sph, tri = icosphere(level)
    ax = np.asarray(axes)
    L = sph.shape[0]
    pos = np.concatenate([s * sph * ax for s in scales], axis=0)
    tet, nml, tsv = _mesh.tri2tetra(tri, n_layers)
    elm = _mesh.orient_boundary(_mesh.generate_element_object_3d(tet, tri, nml, tsv, [n_layers]), pos)
    n_slab = n_layers - 1
    slab = np.arange(tet.shape[0]) // (3 * tri.shape[0])
    elm["elmLyrVec"] = np.ascontiguousarray(np.minimum(slab * 3 // n_slab, 2).astype(np.int32))
    trng = np.random.default_rng(target_seed)
    f = _real_sh_field(sph, 6, trng)
    tpos = (1.0 + 0.05 * f)[:, None] * sph * ax
    target = facet_target(tpos, tri)
    timeStp = 1.0 / time_steps
    prm = dict(elaType=elaType, modVec=[3.0, 3.0, 3.0, 15.0],
               modMat=[[1.0, 1.0], [2.0, 1.5], [4.0, 3.0]],
               spdTanVal=0.005, spdTsvVal=0.005,
               ynkActPrmVec=[0.01, 1.0, 4.0, 0.5], reaActPrmVec=[0.0, 1.0, 1.0, 0.1, 1.0],
               ldmWgt=1e-2, knlOrder=1, knlWidth=0.2, knlEps=1e-10, btmWgt=0.0,
               timeStp=timeStp, timeEnd=1.0, cgItrMax=10000, cgTolVal=1e-10)
    brng = np.random.default_rng(blob_seed)
    u = brng.standard_normal(3)
    u /= np.linalg.norm(u)
    c = 0.9 * u * ax
    tauVar = np.array([[c[0]], [c[1]], [c[2]], [1.0]])
    tauPrm = np.array([0.3, 1.0])
    vfd = dict(cenKnlType="G", cenKnlWidth=0.05, dirKnlType="B", dirKnlWidth=1.0)
    out = dict(dim=3, elm=elm, q0=pos.T.copy(), params=prm, tauVar=tauVar, tauPrm=tauPrm,
               target=target, vfd=vfd)
    if n_ctl:
        crng = np.random.default_rng(ctl_seed)
        out["ctlIdx"] = np.sort(crng.choice(pos.shape[0], size=n_ctl, replace=False)).astype(np.int32)
    return out


def slab_problem(n: int = 9, n_layers: int = 3, thickness: float = 0.2, tilt: float = 0.0, seed: int = 5,
                 time_steps: int = 10, btmWgt: float = 0.0) -> dict:
    """Flat 3D slab: an n x n grid on [0, 1]^2 (each square cut into two
    triangles along alternating diagonals, vertices jittered in the plane by
    +-10 % of the spacing) copied onto n_layers layers at heights
    0 .. thickness and split into tets by tri2tetra.  The bottom layer
    (btmVtxMat, the penalised facets) is the plane z = tilt * x; the top layer
    is the varifold surface, its target the top moved up by a seeded bump."""
    rng = np.random.default_rng(seed)
    h = 1.0 / (n - 1)
    ix, iy = np.meshgrid(np.arange(n), np.arange(n), indexing="xy")
    xy = np.stack([ix.ravel() * h, iy.ravel() * h], axis=1)
    inner = (ix.ravel() > 0) & (ix.ravel() < n - 1) & (iy.ravel() > 0) & (iy.ravel() < n - 1)
    xy[inner] += 0.1 * h * rng.uniform(-1, 1, (int(inner.sum()), 2))
    tris = []
    for j in range(n - 1):
        for i in range(n - 1):
            a, b, c, d = j * n + i, j * n + i + 1, (j + 1) * n + i, (j + 1) * n + i + 1
            tris += [(a, b, d), (a, d, c)] if (i + j) % 2 == 0 else [(a, b, c), (b, d, c)]
    tri = np.array(tris, dtype=np.int64)
    z = np.linspace(0.0, thickness, n_layers)
    pos = np.concatenate([np.column_stack([xy, np.full(len(xy), zk)]) for zk in z], axis=0)
    pos[:, 2] += tilt * pos[:, 0]
    tet, nml, tsv = _mesh.tri2tetra(tri, n_layers)
    elm = _mesh.orient_boundary(_mesh.generate_element_object_3d(tet, tri, nml, tsv, [n_layers]), pos)
    top = pos[(n_layers - 1) * n * n:].copy()
    top[:, 2] += 0.05 * np.exp(-((top[:, 0] - 0.5) ** 2 + (top[:, 1] - 0.5) ** 2) / 0.05)
    target = facet_target(top, tri)
    prm = dict(elaType="frame", modVec=[3.0, 3.0, 3.0, 15.0], modMat=[[1.0, 1.0], [2.0, 1.5], [4.0, 3.0]],
               spdTanVal=0.005, spdTsvVal=0.005, ynkActPrmVec=[0.01, 1.0, 4.0, 0.5],
               reaActPrmVec=[0.0, 1.0, 1.0, 0.1, 1.0], ldmWgt=1e-2, knlOrder=1, knlWidth=0.2, knlEps=1e-10,
               btmWgt=btmWgt, timeStp=1.0 / time_steps, timeEnd=1.0, cgItrMax=10000, cgTolVal=1e-10)
    tauVar = np.array([[0.5], [0.5], [0.5 * thickness], [1.0]])
    tauPrm = np.array([0.3, 1.0])
    vfd = dict(cenKnlType="G", cenKnlWidth=0.1, dirKnlType="B", dirKnlWidth=1.0)
    return dict(dim=3, elm=elm, q0=pos.T.copy(), params=prm, tauVar=tauVar, tauPrm=tauPrm, target=target, vfd=vfd)


def perturb_params(params: dict, rng, frac: float = 0.3) -> dict:
    """Seeded parameter candidate: elastic moduli, diffusion speeds and the
    reaction rate each scaled by an independent factor in [1-frac, 1+frac]."""
    p = dict(params)
    u = lambda: float(rng.uniform(1.0 - frac, 1.0 + frac))
    p["modVec"] = [m * u() for m in params["modVec"]]
    if "modMat" in params:
        p["modMat"] = [[lam * u(), mu * u()] for lam, mu in params["modMat"]]
    s = u()
    p["spdTanVal"] = params["spdTanVal"] * s
    p["spdTsvVal"] = params["spdTsvVal"] * s
    rea = list(params["reaActPrmVec"])
    rea[2] *= u()
    p["reaActPrmVec"] = rea
    return p


def batch_problems(n: int = 64, seed: int = 4, time_steps: int = 20, ctl_frac: float = 0.1,
                   max_ctl: int = 2000, v2d=(1000, 20000), v3d=(5000, 60000), mixed_dims: bool = False):
    """Independent layered meshes of uneven size: half 2D rings with
    log-uniform vertex counts in v2d, half 3D shells with counts in v3d (the
    first half 2D, or alternating when mixed_dims), each with seeded
    parameters (+-30%), its own blob and its own control points (ctl_frac of
    the landmarks, at least 50, at most max_ctl)."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n):
        two_d = (k % 2 == 0) if mixed_dims else k < n // 2
        lo, hi = v2d if two_d else v3d
        Vt = float(np.exp(rng.uniform(np.log(lo), np.log(hi))))
        if two_d:
            rings = int(np.clip(round(np.sqrt(Vt / 20.0)), 5, 120))
            per = max(20, int(round(Vt / rings)))
            p = ring_problem(n_rings=rings, n_per_ring=per, seed=100 + k, target_seed=200 + k)
        else:
            # icosphere level and layer count closest to the requested size
            best = None
            for lvl in (2, 3, 4, 5, 6):
                L = 10 * 4 ** lvl + 2
                layers = int(round(Vt / L))
                if 3 <= layers <= 12:
                    err = abs(layers * L - Vt)
                    if best is None or err < best[0]:
                        best = (err, lvl, layers)
            if best is None:
                best = (0, 2, 3)
            p = shell_problem(level=best[1], n_layers=best[2], target_seed=300 + k, blob_seed=400 + k)
        p["params"] = perturb_params(p["params"], rng)
        p["params"]["timeStp"] = 1.0 / time_steps
        p["timeNum"] = time_steps + 1
        V = p["q0"].shape[1]
        nc = int(min(max_ctl, max(50, ctl_frac * V)))
        p["ctlIdx"] = np.sort(rng.choice(V, nc, replace=False)).astype(np.int32)
        out.append(p)
    return out
