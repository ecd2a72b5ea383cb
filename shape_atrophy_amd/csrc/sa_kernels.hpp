// gfx950 kernels of the forward shape-evolution step.
//
// Element-parallel kernels write per-element contributions to SoA scratch
// and vertex/nonzero-parallel kernels gather them through the reference's
// count-prefixed adjacency tables in ascending element order, so every sum
// has a fixed order and results are bitwise reproducible run to run (no
// atomics anywhere).  Scalars of the Krylov solvers stay on the device: each
// kernel reduces the partial sums of its predecessor itself (every block
// computes the same value in the same order) and a done flag makes all later
// kernels of a launched chunk return at once.
#pragma once
#include <algorithm>
#include <type_traits>
#include "sa_device.hpp"

namespace sa {

struct Ctl {            // solver control block, device resident
    double bb, tol;     // |b|^2 and the stopping threshold on |r|^2
    int done, status, iters;
    int it, maxIt;      // control-point CG: current iteration (-1 before the first) and budget
    int pad;
    double rr;          // control-point CG: |r|^2 after the last iteration
};

// ---------------------------------------------------------------------------
// initial concentration, tauInitial.cu:4-64|4-66

__device__ inline double intpow(double b, int e) {
    double out = 1.0, pw = b;
    do {
        int quo = e >> 1, rem = e - 2 * quo;
        if (rem == 1) out *= pw;
        pw *= pw;
        e = quo;
    } while (e != 0);
    return out;
}

template <int D>
__global__ void tau_initial_k(int V, const double *__restrict__ q, double c0, double c1, double c2, double h,
                              double R, int pw, double *__restrict__ tau) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= V) return;
    double d0 = q[i] - c0, d1 = q[V + i] - c1, s = d0 * d0 + d1 * d1;
    if constexpr (D == 3) { double d2 = q[2 * V + i] - c2; s += d2 * d2; }
    double r2 = s / (R * R);
    double t = (r2 - 1.0) * (r2 - 1.0);
    tau[i] = h * (r2 < 1.0 ? 1.0 - intpow(1.0 - t, pw) : 0.0);
}

// ---------------------------------------------------------------------------
// geometry of every element at positions q: Q^{-1}, det Q, frames
// (vertexToEdge.cu, computeFrame.cu) and the element mass / anisotropic
// stiffness entries (assembleFemMatrix.cu:5-58|5-118)

template <int D>
__global__ __launch_bounds__(NT) void geometry_k(int E, int V, const int *__restrict__ elmVtx,
                                                 const int *__restrict__ frmVtx, const int *__restrict__ tsvVtx,
                                                 const double *__restrict__ q, double spdTan, double spdTsv,
                                                 double *__restrict__ geo, double *__restrict__ femElm) {
    using G = GeoLayout<D>;
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= E) return;
    int v0 = elmVtx[e];
    double p0[D];
#pragma unroll
    for (int c = 0; c < D; ++c) p0[c] = q[c * V + v0];
    double ed[D][D];
#pragma unroll
    for (int k = 0; k < D; ++k) {
        int vk = elmVtx[(k + 1) * E + e];
#pragma unroll
        for (int c = 0; c < D; ++c) ed[k][c] = q[c * V + vk] - p0[c];
    }
    double inv[D][D], det;
    mat_inv<D>(ed, inv, det);
    // frames
    double f1[D], f2[D];
    {
        double s1[D], s2[D];
#pragma unroll
        for (int c = 0; c < D; ++c) { s1[c] = 0.0; s2[c] = 0.0; }
        if constexpr (D == 2) {
#pragma unroll
            for (int k = 0; k < 2; ++k) {
                int a = frmVtx[(2 * k) * E + e], b = frmVtx[(2 * k + 1) * E + e];
                double d0 = q[b] - q[a], d1 = q[V + b] - q[V + a];
                double len = sqrt(d0 * d0 + d1 * d1);
                s1[0] += d0 / len; s1[1] += d1 / len;
            }
        } else {
#pragma unroll
            for (int k = 0; k < 2; ++k) {
                int a = frmVtx[(3 * k) * E + e], b = frmVtx[(3 * k + 1) * E + e], cc = frmVtx[(3 * k + 2) * E + e];
                double u[3], w[3];
#pragma unroll
                for (int i = 0; i < 3; ++i) { u[i] = q[i * V + b] - q[i * V + a]; w[i] = q[i * V + cc] - q[i * V + a]; }
                double cr0 = u[1] * w[2] - u[2] * w[1], cr1 = -u[0] * w[2] + u[2] * w[0], cr2 = u[0] * w[1] - u[1] * w[0];
                double len = sqrt(cr0 * cr0 + cr1 * cr1 + cr2 * cr2);
                s1[0] += cr0 / len; s1[1] += cr1 / len; s1[2] += cr2 / len;
            }
        }
        constexpr int NTSV = (D == 2 ? 2 : 3);
#pragma unroll
        for (int k = 0; k < NTSV; ++k) {
            int a = tsvVtx[(2 * k) * E + e], b = tsvVtx[(2 * k + 1) * E + e];
            double dd[D], l2 = 0.0;
#pragma unroll
            for (int c = 0; c < D; ++c) { dd[c] = q[c * V + b] - q[c * V + a]; l2 += dd[c] * dd[c]; }
            double len = sqrt(l2);
#pragma unroll
            for (int c = 0; c < D; ++c) s2[c] += dd[c] / len;
        }
        double l1 = 0.0, l2 = 0.0;
#pragma unroll
        for (int c = 0; c < D; ++c) { l1 += s1[c] * s1[c]; l2 += s2[c] * s2[c]; }
        l1 = sqrt(l1); l2 = sqrt(l2);
#pragma unroll
        for (int c = 0; c < D; ++c) { f1[c] = s1[c] / l1; f2[c] = s2[c] / l2; }
    }
#pragma unroll
    for (int a = 0; a < D; ++a)
#pragma unroll
        for (int b = 0; b < D; ++b) geo[(G::INV + a * D + b) * (size_t)E + e] = inv[a][b];
    geo[G::DET * (size_t)E + e] = det;
#pragma unroll
    for (int c = 0; c < D; ++c) { geo[(G::F1 + c) * (size_t)E + e] = f1[c]; geo[(G::F2 + c) * (size_t)E + e] = f2[c]; }

    // FEM element matrices
    double grad[D + 1][D];
#pragma unroll
    for (int c = 0; c < D; ++c) {
        double s = 0.0;
#pragma unroll
        for (int k = 0; k < D; ++k) { grad[k + 1][c] = inv[k][c]; s -= inv[k][c]; }
        grad[0][c] = s;
    }
    double a1[D + 1], a2[D + 1];
#pragma unroll
    for (int k = 0; k <= D; ++k) {
        double s1 = 0.0, s2 = 0.0;
#pragma unroll
        for (int c = 0; c < D; ++c) { s1 += grad[k][c] * f1[c]; s2 += grad[k][c] * f2[c]; }
        a1[k] = s1; a2[k] = s2;
    }
    double ad = fabs(det);
    femElm[e] = ad / (D == 2 ? 24.0 : 120.0);
    int sl = 0;
#pragma unroll
    for (int a = 0; a <= D; ++a)
#pragma unroll
        for (int b = a; b <= D; ++b, ++sl) {
            double v;
            if constexpr (D == 2) {
                v = 0.5 * ad * (spdTan * a1[a] * a1[b] + spdTsv * a2[a] * a2[b]);
            } else {
                double gg = grad[a][0] * grad[b][0] + grad[a][1] * grad[b][1] + grad[a][2] * grad[b][2];
                v = ad / 6.0 * (spdTan * (gg - a1[a] * a1[b]) + spdTsv * a2[a] * a2[b]);
            }
            femElm[(1 + sl) * (size_t)E + e] = v;
        }
}

// per-nonzero gather of the FEM matrices (femMatGatherKernel,
// assembleFemMatrix.cu:61-100|120-160): mass M(z) and lhs(z) = M + dt K_D
__global__ __launch_bounds__(NT) void fem_gather_k(int Z, int E, const int *__restrict__ femIfo,
                                                   const double *__restrict__ femElm, double dt,
                                                   double *__restrict__ ppd, double *__restrict__ lhs) {
    int z = blockIdx.x * blockDim.x + threadIdx.x;
    if (z >= Z) return;
    int adj = femIfo[z];
    double mul = femIfo[Z + z];
    double pv = 0.0, gv = 0.0;
    for (int a = 0; a < adj; ++a) {
        int e = femIfo[(2 + 2 * a) * Z + z];
        int gi = femIfo[(3 + 2 * a) * Z + z];
        pv += mul * femElm[e];
        gv += femElm[(size_t)E + gi];
    }
    ppd[z] = pv;
    lhs[z] = 1.0 * pv + dt * gv;
}

__global__ void activity_k(int V, const double *__restrict__ tau, double y0, double y1, double y2, double y3,
                           double r0, double r1, double r2, double r3, double r4, double *__restrict__ ya,
                           double *__restrict__ ra) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= V) return;
    double t = tau[i];
    ya[i] = yank_activity(t, y0, y1, y2, y3);
    ra[i] = reac_activity(t, r0, r1, r2, r3, r4);
}

// external yank per element, computeExternalYank.cu:5-47|5-56
template <int D>
__global__ __launch_bounds__(NT) void yank_elem_k(int E, const int *__restrict__ elmVtx, const double *__restrict__ geo,
                                                  const double *__restrict__ act, double *__restrict__ frc) {
    using G = GeoLayout<D>;
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= E) return;
    double fc = 0.0;
#pragma unroll
    for (int k = 0; k <= D; ++k) fc += act[elmVtx[k * E + e]];
    fc /= (D + 1);
    double vol = elem_vol<D>(geo[G::DET * (size_t)E + e]);
#pragma unroll
    for (int c = 0; c < D; ++c) {
        double s0 = 0.0;
#pragma unroll
        for (int k = 0; k < D; ++k) {
            double g = geo[(G::INV + k * D + c) * (size_t)E + e];
            s0 -= g;
            frc[((size_t)e * (D + 1) + k + 1) * D + c] = -fc * g * vol;
        }
        frc[(size_t)e * (D + 1) * D + c] = -fc * s0 * vol;
    }
}

// boundary yank, the term the reference keeps commented out
// (computeExternalYank.cu:49-84 | 3D :58-99): on boundary facet b with
// vertices q_0..q_{D-1}, n = (t_y, -t_x) with t = q_1 - q_0 (2D) or
// n = (q_1 - q_0) x (q_2 - q_0) (3D), and the force on its vertex k is
// act(q_k) n / D (VTXNUM - 1 = D).  bf[(k D + c) Bn + b].
template <int D>
__global__ void bdr_yank_k(int Bn, int V, const int *__restrict__ bdrVtx, const double *__restrict__ q,
                           const double *__restrict__ act, double *__restrict__ bf) {
    int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= Bn) return;
    int v[D];
#pragma unroll
    for (int k = 0; k < D; ++k) v[k] = bdrVtx[k * Bn + b];
    double n[D];
    if constexpr (D == 2) {
        const double t0 = q[v[1]] - q[v[0]], t1 = q[V + v[1]] - q[V + v[0]];
        n[0] = t1;
        n[1] = -t0;
    } else {
        double u[3], w[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) { u[c] = q[c * V + v[1]] - q[c * V + v[0]]; w[c] = q[c * V + v[2]] - q[c * V + v[0]]; }
        n[0] = u[1] * w[2] - u[2] * w[1];
        n[1] = u[2] * w[0] - u[0] * w[2];
        n[2] = u[0] * w[1] - u[1] * w[0];
    }
#pragma unroll
    for (int k = 0; k < D; ++k) {
        const double f = act[v[k]];
#pragma unroll
        for (int c = 0; c < D; ++c) bf[(size_t)(k * D + c) * Bn + b] = f * n[c] / (double)D;
    }
}

// vertex gather of element vectors (yankGatherKernel / frmElasticGatherKernel)
// with a partial sum of |out|^2 per block: the element forces in vtxElmMat
// order, then (with vtxBdr) the boundary forces in vtxBdrMat order
template <int D>
__global__ __launch_bounds__(NT) void gather_k(int V, int E, const int *__restrict__ vtxElm,
                                               const double *__restrict__ frc, double *__restrict__ out,
                                               double *__restrict__ P, const int *__restrict__ vtxBdr = nullptr,
                                               const double *__restrict__ bf = nullptr, int Bn = 0) {
    __shared__ double red[NT / 64];
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    double ss = 0.0;
    if (i < V) {
        double acc[D];
#pragma unroll
        for (int c = 0; c < D; ++c) acc[c] = 0.0;
        int adj = vtxElm[i];
        for (int a = 0; a < adj; ++a) {
            int e = vtxElm[(1 + 2 * a) * V + i], l = vtxElm[(2 + 2 * a) * V + i];
#pragma unroll
            for (int c = 0; c < D; ++c) acc[c] += frc[((size_t)e * (D + 1) + l) * D + c];
        }
        if (vtxBdr) {
            const int nb = vtxBdr[i];
            for (int a = 0; a < nb; ++a) {
                const int bb = vtxBdr[(1 + 2 * a) * V + i], l = vtxBdr[(2 + 2 * a) * V + i];
#pragma unroll
                for (int c = 0; c < D; ++c) acc[c] += bf[(size_t)(l * D + c) * Bn + bb];
            }
        }
#pragma unroll
        for (int c = 0; c < D; ++c) { out[c * V + i] = acc[c]; ss += acc[c] * acc[c]; }
    }
    double s = block_sum(ss, red);
    if (P && threadIdx.x == 0) P[blockIdx.x] = s;
}

// elastic element forces for a velocity field x (strain, stress, |T| Q^{-1} sigma):
// computeStrainTensor.cu:4-53|4-61 + frmElasticComputeKernel
// (applyFrameElasticOperator.cu:41-85|113-164).  ISO selects the per-layer
// isotropic law sigma = (1 - 0.5 u_bar)(lambda tr(eps) I + 2 mu eps).
// Element e with vertex velocities xv, Q^{-1} inv, volume vol, frames f1,
// f2 (frame model) or the mean concentration ub of its vertices (ISO); the
// force on its local vertex l lands in fo[l].
template <int D, bool ISO>
__device__ inline void elem_forces(int e, const double (&xv)[D + 1][D], double ub, const double (&inv)[D][D],
                                   double vol, const double (&f1)[D], const double (&f2)[D],
                                   const int *__restrict__ elmLyr, const double *__restrict__ mod,
                                   double (&fo)[D + 1][D]) {
    double Vm[D][D];
#pragma unroll
    for (int k = 0; k < D; ++k)
#pragma unroll
        for (int c = 0; c < D; ++c) Vm[c][k] = xv[k + 1][c] - xv[0][c];
    double Dv[D][D];
#pragma unroll
    for (int a = 0; a < D; ++a)
#pragma unroll
        for (int b = 0; b < D; ++b) {
            double s = 0.0;
#pragma unroll
            for (int k = 0; k < D; ++k) s += Vm[a][k] * inv[k][b];
            Dv[a][b] = s;
        }
    double eps[D][D];
#pragma unroll
    for (int a = 0; a < D; ++a)
#pragma unroll
        for (int b = 0; b < D; ++b) eps[a][b] = (a == b ? Dv[a][a] : 0.5 * (Dv[a][b] + Dv[b][a]));
    double sig[D][D];
    if constexpr (ISO) {
        int l = elmLyr ? elmLyr[e] : 0;
        double sc = 1.0 - 0.5 * ub;
        double lam = mod[2 * l] * sc, mu = mod[2 * l + 1] * sc;
        double tr = 0.0;
#pragma unroll
        for (int c = 0; c < D; ++c) tr += eps[c][c];
#pragma unroll
        for (int a = 0; a < D; ++a)
#pragma unroll
            for (int b = 0; b < D; ++b) sig[a][b] = (a == b ? lam * tr : 0.0) + 2.0 * mu * eps[a][b];
    } else {
        // moduli change between runs (host copies): read coherently
        double md[D + 1];
#pragma unroll
        for (int m = 0; m <= D; ++m) md[m] = cld(&mod[m]);
        frame_stress<D>(eps, f1, f2, md, sig);
    }
#pragma unroll
    for (int c = 0; c < D; ++c) {
        double s0 = 0.0;
#pragma unroll
        for (int k = 0; k < D; ++k) {
            double s = 0.0;
#pragma unroll
            for (int j = 0; j < D; ++j) s += inv[k][j] * sig[j][c];
            s0 -= s;
            fo[k + 1][c] = s * vol;
        }
        fo[0][c] = s0 * vol;
    }
}

// elem_forces over every element, Q^{-1} and det Q from the stored geometry.
// slot non-null: the force on local vertex l goes to frc[slot[l E + e] * D + c],
// the position of (e, l) in the vertex-ordered list gather_op_k reads
// contiguously (slot < 0: not listed).
template <int D, bool ISO, typename T>
__global__ __launch_bounds__(NT) void elastic_elem_k(int E, int V, const int *__restrict__ elmVtx,
                                                     const double *__restrict__ geo, const T *__restrict__ x,
                                                     const double *__restrict__ tau, const int *__restrict__ elmLyr,
                                                     const double *__restrict__ mod, T *__restrict__ frc,
                                                     const Ctl *ctl, const int *__restrict__ slot = nullptr) {
    using G = GeoLayout<D>;
    if (ctl && cld(&ctl->done)) return;
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= E) return;
    int vtx[D + 1];
#pragma unroll
    for (int k = 0; k <= D; ++k) vtx[k] = elmVtx[k * E + e];
    double inv[D][D];
#pragma unroll
    for (int a = 0; a < D; ++a)
#pragma unroll
        for (int b = 0; b < D; ++b) inv[a][b] = geo[(G::INV + a * D + b) * (size_t)E + e];
    const double vol = elem_vol<D>(geo[G::DET * (size_t)E + e]);
    double xv[D + 1][D];
#pragma unroll
    for (int k = 0; k <= D; ++k)
#pragma unroll
        for (int c = 0; c < D; ++c) xv[k][c] = (double)x[c * V + vtx[k]];
    double ub = 0.0;
    if constexpr (ISO) {
#pragma unroll
        for (int k = 0; k <= D; ++k) ub += tau[vtx[k]];
        ub /= (D + 1);
    }
    double f1[D], f2[D];
    if constexpr (!ISO) {
#pragma unroll
        for (int c = 0; c < D; ++c) { f1[c] = geo[(G::F1 + c) * (size_t)E + e]; f2[c] = geo[(G::F2 + c) * (size_t)E + e]; }
    }
    double fo[D + 1][D];
    elem_forces<D, ISO>(e, xv, ub, inv, vol, f1, f2, elmLyr, mod, fo);
    if (slot) {
#pragma unroll
        for (int l = 0; l <= D; ++l) {
            const int sl = slot[l * E + e];
            if (sl < 0) continue;
#pragma unroll
            for (int c = 0; c < D; ++c) frc[(size_t)sl * D + c] = (T)fo[l][c];
        }
    } else {
#pragma unroll
        for (int l = 0; l <= D; ++l)
#pragma unroll
            for (int c = 0; c < D; ++c) frc[((size_t)e * (D + 1) + l) * D + c] = (T)fo[l][c];
    }
}

// ---------------------------------------------------------------------------
// Jacobi preconditioner of the control-point system S = ldmWgt (K_cc + eps I)
// + K_cx (E + btmWgt P) K_xc: diag(S) for every control point j and
// component c.  With k_j = K(., c_j) over the landmarks and B_cc the
// (c, c) block of the elastic stiffness,
//   S_(jc,jc) = ldmWgt (k(0) + eps) + k_j^T B_cc k_j + btmWgt sum_v P_v k_j(v)^2.
//
// stiff_blocks_k: per element the (c, c) blocks of its stiffness,
// B_e[(l, c), (m, c)] = |T| g_l . sigma(sym(e_c g_m^T)) e_c with g_m the P1
// gradients (rows of Q^{-1}, g_0 = -sum), i.e. the force on vertex l,
// component c, of the element operator for a unit displacement of vertex m
// along c; symmetric, kept for l <= m in the femIfoMat slot order
// (00, 01, .. 0D, 11, ..): blk[(c NS + slot) E + e].
template <int D, bool ISO>
__global__ __launch_bounds__(NT) void stiff_blocks_k(int E, const int *__restrict__ elmVtx,
                                                     const double *__restrict__ geo, const double *__restrict__ tau,
                                                     const int *__restrict__ elmLyr, const double *__restrict__ mod,
                                                     double *__restrict__ blk) {
    using G = GeoLayout<D>;
    constexpr int NS = G::NS;
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= E) return;
    double g[D + 1][D];
#pragma unroll
    for (int b = 0; b < D; ++b) {
        double s = 0.0;
#pragma unroll
        for (int k = 0; k < D; ++k) {
            g[k + 1][b] = geo[(G::INV + k * D + b) * (size_t)E + e];
            s -= g[k + 1][b];
        }
        g[0][b] = s;
    }
    const double vol = elem_vol<D>(geo[G::DET * (size_t)E + e]);
    double f1[D], f2[D], md[D + 1], lam = 0.0, mu = 0.0;
    if constexpr (ISO) {
        const int l = elmLyr ? elmLyr[e] : 0;
        double ub = 0.0;
#pragma unroll
        for (int k = 0; k <= D; ++k) ub += tau[elmVtx[k * E + e]];
        ub /= (D + 1);
        const double sc = 1.0 - 0.5 * ub;
        lam = mod[2 * l] * sc;
        mu = mod[2 * l + 1] * sc;
    } else {
#pragma unroll
        for (int c = 0; c < D; ++c) { f1[c] = geo[(G::F1 + c) * (size_t)E + e]; f2[c] = geo[(G::F2 + c) * (size_t)E + e]; }
#pragma unroll
        for (int m = 0; m <= D; ++m) md[m] = cld(&mod[m]);
    }
#pragma unroll
    for (int c = 0; c < D; ++c) {
        int sl = 0;
#pragma unroll
        for (int l = 0; l <= D; ++l)
#pragma unroll
            for (int m = l; m <= D; ++m, ++sl) {
                double eps[D][D], sig[D][D];
#pragma unroll
                for (int a = 0; a < D; ++a)
#pragma unroll
                    for (int b = 0; b < D; ++b)
                        eps[a][b] = 0.5 * ((a == c ? g[m][b] : 0.0) + (b == c ? g[m][a] : 0.0));
                if constexpr (ISO) {
                    const double tr = eps[c][c];
#pragma unroll
                    for (int a = 0; a < D; ++a)
#pragma unroll
                        for (int b = 0; b < D; ++b) sig[a][b] = (a == b ? lam * tr : 0.0) + 2.0 * mu * eps[a][b];
                } else {
                    frame_stress<D>(eps, f1, f2, md, sig);
                }
                double v = 0.0;
#pragma unroll
                for (int j = 0; j < D; ++j) v += g[l][j] * sig[j][c];
                blk[((size_t)c * NS + sl) * E + e] = vol * v;
            }
    }
}

// per-nonzero gather of the (c, c) stiffness blocks onto the femVtxMat
// pattern (femIfoMat lists (element, slot E + element) per nonzero, as for
// the FEM matrices): bcc[c Z + z]
template <int D>
__global__ __launch_bounds__(NT) void stiff_gather_k(int Z, int E, const int *__restrict__ femIfo,
                                                     const double *__restrict__ blk, double *__restrict__ bcc) {
    constexpr int NS = GeoLayout<D>::NS;
    const int z = blockIdx.x * blockDim.x + threadIdx.x;
    if (z >= Z) return;
    const int adj = femIfo[z];
    double acc[D];
#pragma unroll
    for (int c = 0; c < D; ++c) acc[c] = 0.0;
    for (int a = 0; a < adj; ++a) {
        const int gi = femIfo[(3 + 2 * a) * Z + z];
#pragma unroll
        for (int c = 0; c < D; ++c) acc[c] += blk[(size_t)c * NS * E + gi];
    }
#pragma unroll
    for (int c = 0; c < D; ++c) bcc[(size_t)c * Z + z] = acc[c];
}

// lumped penalty weight of each bottom landmark: P_v = sum over its bottom
// facets of |F| / D (the diagonal of the penalty operator, pen_elem_k)
template <int D>
__global__ void pen_weight_k(int BL, const int *__restrict__ vtxBtm, const double *__restrict__ meas,
                             double *__restrict__ pw) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= BL) return;
    const int adj = vtxBtm[v];
    double s = 0.0;
    for (int a = 0; a < adj; ++a) s += meas[vtxBtm[(1 + 2 * a) * BL + v]] / (double)D;
    pw[v] = s;
}

// one block per control point j: diag(S) for its D components, the kernel
// values k_j(v) evaluated in FP64 from the positions; the threads' partial
// sums are added in a fixed order
template <int D, int ORD>
__global__ __launch_bounds__(NT) void jacobi_diag_k(int V, int C, const double *__restrict__ q,
                                                    const double *__restrict__ qc, const int *__restrict__ rowPtr,
                                                    const int *__restrict__ colIdx, int Z,
                                                    const double *__restrict__ bcc, int BL,
                                                    const double *__restrict__ pw, double btmWgt, double ldmWgt,
                                                    double eps, double invW, double *__restrict__ dS) {
    __shared__ double etab[64];
    __shared__ double red[NT / 64];
    init_exp_table(etab);
    __syncthreads();
    const int j = blockIdx.x;
    double cj[D];
#pragma unroll
    for (int c = 0; c < D; ++c) cj[c] = qc[(size_t)c * C + j];
    auto kval = [&](int v) {
        double d2 = 0.0;
#pragma unroll
        for (int c = 0; c < D; ++c) {
            const double dd = q[(size_t)c * V + v] - cj[c];
            d2 += dd * dd;
        }
        return rkhs_kernel<D, ORD>(d2, invW, etab);
    };
    double acc[D];
#pragma unroll
    for (int c = 0; c < D; ++c) acc[c] = 0.0;
    for (int v = threadIdx.x; v < V; v += NT) {
        const double kv = kval(v);
        double t[D];
#pragma unroll
        for (int c = 0; c < D; ++c) t[c] = 0.0;
        for (int z = rowPtr[v]; z < rowPtr[v + 1]; ++z) {
            const int w = colIdx[z];
            const double kw = w == v ? kv : kval(w);
#pragma unroll
            for (int c = 0; c < D; ++c) t[c] += bcc[(size_t)c * Z + z] * kw;
        }
        const double p = v < BL ? btmWgt * pw[v] * kv * kv : 0.0;
#pragma unroll
        for (int c = 0; c < D; ++c) acc[c] += kv * t[c] + p;
    }
    const double k0 = rkhs_kernel<D, ORD>(0.0, invW, etab);
#pragma unroll
    for (int c = 0; c < D; ++c) {
        const double s = block_sum(acc[c], red);
        if (threadIdx.x == 0) dS[(size_t)c * C + j] = ldmWgt * (k0 + eps) + s;
    }
}

// Jacobi step of the control-point CG: stop on |r|^2 (partials Prr) below
// the tolerance, else z = r / diag(S) and the partial of r.z for each block
// of 64 entries (the count the direction update reads)
__global__ __launch_bounds__(64) void jacobi_dot_k(int n, const double *__restrict__ Prr, int nPrr,
                                                   const double *__restrict__ r, const double *__restrict__ dS,
                                                   double *__restrict__ z, double *__restrict__ P3, Ctl *ctl) {
    __shared__ double red[1];
    if (cld(&ctl->done)) return;
    const int it = cld(&ctl->it);
    const double rr = reduce_partials(Prr, nPrr, red);
    if (blockIdx.x == 0 && threadIdx.x == 0) cst(&ctl->rr, rr);
    if (rr < cld(&ctl->tol)) {
        if (blockIdx.x == 0 && threadIdx.x == 0) { cst(&ctl->done, 1); cst(&ctl->iters, it + 1); }
        return;
    }
    const int i = blockIdx.x * 64 + threadIdx.x;
    double rz = 0.0;
    if (i < n) {
        const double zi = r[i] / dS[i];
        z[i] = zi;
        rz = r[i] * zi;
    }
    const double t = block_sum(rz, red);
    if (threadIdx.x == 0) P3[blockIdx.x] = t;
}

// bottom facet measure: 2D segment length, 3D xy-projected area
// (applyPenaltyOperator.cu:13-14 | 3D :12-14)
template <int D>
__global__ void btm_meas_k(int B, int V, const int *__restrict__ btmVtx, const double *__restrict__ q,
                           double *__restrict__ meas) {
    int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= B) return;
    int a0 = btmVtx[b], a1 = btmVtx[B + b];
    if constexpr (D == 2) {
        double d0 = q[a1] - q[a0], d1 = q[V + a1] - q[V + a0];
        meas[b] = sqrt(d0 * d0 + d1 * d1);
    } else {
        int a2 = btmVtx[2 * B + b];
        meas[b] = 0.5 * fabs((q[a1] - q[a0]) * (q[V + a2] - q[V + a0]) - (q[V + a1] - q[V + a0]) * (q[a2] - q[a0]));
    }
}

// penalty element values v_k |F| / D (penaltyComputeKernel)
template <int D>
__global__ void pen_elem_k(int B, int V, const int *__restrict__ btmVtx, const double *__restrict__ meas,
                           const double *__restrict__ x, double *__restrict__ pfrc, const Ctl *ctl) {
    if (ctl && cld(&ctl->done)) return;
    int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= B) return;
    double m = meas[b];
#pragma unroll
    for (int k = 0; k < D; ++k) {
        int v = btmVtx[k * B + b];
#pragma unroll
        for (int c = 0; c < D; ++c) pfrc[(k * D + c) * (size_t)B + b] = x[c * V + v] * m / (double)D;
    }
}

// Gather of E*x (+ btmWgt * P*x) and the PCG combinations around it:
// MODE 0 (iteration): Ad = ldmWgt*Kid + Bd; partial d.Ad
// MODE 1 (start):     r = b - (ldmWgt*a + Bx); Kid = r; partial r.r
template <int D, int MODE>
__global__ __launch_bounds__(NT) void gather_op_k(int V, int E, const int *__restrict__ vtxElm,
                                                  const double *__restrict__ frc, int BL, int B,
                                                  const int *__restrict__ vtxBtm, const double *__restrict__ pfrc,
                                                  double btmWgt, double ldmWgt, const double *__restrict__ u,
                                                  const double *__restrict__ w, double *__restrict__ out,
                                                  double *__restrict__ out2, double *__restrict__ P, const Ctl *ctl,
                                                  const int *__restrict__ vstart = nullptr) {
    __shared__ double red[NT / 64];
    if (cld(&ctl->done)) return;
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    double ss = 0.0;
    if (i < V) {
        double acc[D];
#pragma unroll
        for (int c = 0; c < D; ++c) acc[c] = 0.0;
        if (vstart) {
            // forces in vertex order (elastic_elem_k with slots): this
            // vertex's list is contiguous, in the order vtxElmMat lists it
            const int a1 = vstart[i + 1];
            for (int a = vstart[i]; a < a1; ++a)
#pragma unroll
                for (int c = 0; c < D; ++c) acc[c] += frc[(size_t)a * D + c];
        } else {
            int adj = vtxElm[i];
            for (int a = 0; a < adj; ++a) {
                int e = vtxElm[(1 + 2 * a) * V + i], l = vtxElm[(2 + 2 * a) * V + i];
#pragma unroll
                for (int c = 0; c < D; ++c) acc[c] += frc[((size_t)e * (D + 1) + l) * D + c];
            }
        }
        if (pfrc && i < BL) {
            double pa[D];
#pragma unroll
            for (int c = 0; c < D; ++c) pa[c] = 0.0;
            int adj2 = vtxBtm[i];
            for (int a = 0; a < adj2; ++a) {
                int b = vtxBtm[(1 + 2 * a) * BL + i], l = vtxBtm[(2 + 2 * a) * BL + i];
#pragma unroll
                for (int c = 0; c < D; ++c) pa[c] += pfrc[(l * D + c) * (size_t)B + b];
            }
#pragma unroll
            for (int c = 0; c < D; ++c) acc[c] = 1.0 * acc[c] + btmWgt * pa[c];
        }
#pragma unroll
        for (int c = 0; c < D; ++c) {
            int k = c * V + i;
            if constexpr (MODE == 0) {
                // u = Kid, w = d
                double ad = ldmWgt * u[k] + 1.0 * acc[c];
                out[k] = ad;
                ss += w[k] * ad;
            } else if constexpr (MODE == 2) {
                out[k] = acc[c];   // B x alone (control-point form)
            } else {
                // u = a (alpha), w = b (rhs); out = r, out2 = Kid
                double ax = ldmWgt * u[k] + acc[c];
                double r = 1.0 * w[k] - 1.0 * ax;
                out[k] = r;
                out2[k] = r;
                ss += r * r;
            }
        }
    }
    if constexpr (MODE != 2) {
        double s = block_sum(ss, red);
        if (threadIdx.x == 0) P[blockIdx.x] = s;
    }
}

// ---------------------------------------------------------------------------
// The operator apply in one launch: elastic_elem_k and gather_op_k fused, the
// element forces kept in LDS.  The vertices are grouped into chunks (built on
// the host from the vertex -> element lists: spatially close vertices, at
// most BT vertices, MS list entries and HM halo vertices per chunk).  A
// block takes one chunk.  It stages the positions and the input field x (and
// the concentration, ISO) of the chunk's halo (every vertex of an element
// its vertices' lists name) in LDS, computes each such element once (Q^{-1}
// and det Q from the positions, as geometry_k forms them), writes the force
// on each listed vertex into that list entry's LDS slot, and then each vertex
// sums its own entries in list order (ascending element, as gather_op_k does
// from HBM), adds the bottom penalty and applies the CG mode of gather_op_k.
// Tables (one chunk k):
//   chV[k] .. chV[k+1]   its vertex entries: vertex chVtx[j], whose list
//                        entries take LDS slots chLs[j] .. chLs[j] + count
//   chH[k] .. chH[k+1]   its halo: vertex chHv[h]
//   chE[k] .. chE[k+1]   its element entries: element chElm[u] (read for
//                        its layer, ISO), frames frE[c NU + u] (f1) and
//                        frE[(D + c) NU + u] (f2) in entry order (frame
//                        model, op_frames_k), and per local vertex l the word
//                        chLS[l NU + u] = halo position | (LDS slot + 1) << 16
//                        (slot -1: that vertex is not in the chunk)
// P[k]: the block's partial of the mode's dot product (MODE 0 and 1).
struct OpfShape {
    int nt, maxs, halo;   // BT, MS, HM
};
constexpr OpfShape OPF_SHAPES[] = {{256, 1024, 512}, {512, 1024, 512}, {512, 2048, 1024}, {1024, 2048, 1024}};
constexpr int OPF_NSHAPES = 4;

// frames of the chunks' element entries in entry order, from the geometry
// (once per geometry: the fused apply then reads them coalesced)
template <int D>
__global__ __launch_bounds__(NT) void op_frames_k(int NU, int E, const int *__restrict__ chElm,
                                                  const double *__restrict__ geo, double *__restrict__ frE) {
    using G = GeoLayout<D>;
    const int u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= NU) return;
    const int e = chElm[u];
#pragma unroll
    for (int c = 0; c < D; ++c) {
        frE[(size_t)c * NU + u] = geo[(G::F1 + c) * (size_t)E + e];
        frE[(size_t)(D + c) * NU + u] = geo[(G::F2 + c) * (size_t)E + e];
    }
}

template <int D, bool ISO, int MODE, int BT, int MS, int HM>
__global__ __launch_bounds__(BT) void op_fused_k(const int *__restrict__ chV, const int *__restrict__ chVtx,
                                                 const int *__restrict__ chLs, const int *__restrict__ chH,
                                                 const int *__restrict__ chHv, const int *__restrict__ chE,
                                                 const int *__restrict__ chElm, const int *__restrict__ chLS,
                                                 int NU, int V, const int *__restrict__ vstart,
                                                 const double *__restrict__ q, const double *__restrict__ frE,
                                                 const double *__restrict__ x,
                                                 const double *__restrict__ tau, const int *__restrict__ elmLyr,
                                                 const double *__restrict__ mod, int BL, int B,
                                                 const int *__restrict__ vtxBtm, const double *__restrict__ pfrc,
                                                 double btmWgt, double ldmWgt, const double *__restrict__ u,
                                                 const double *__restrict__ w, double *__restrict__ out,
                                                 double *__restrict__ out2, double *__restrict__ P, const Ctl *ctl) {
    __shared__ double f[MS * D];
    __shared__ double hq[D * HM], hx[D * HM];
    __shared__ double ht[ISO ? HM : 1];
    __shared__ double red[BT / 64];
    if (cld(&ctl->done)) return;
    const int k = blockIdx.x;
    const int h0 = chH[k], nh = chH[k + 1] - h0;
    for (int h = threadIdx.x; h < nh; h += BT) {
        const int v = chHv[h0 + h];
#pragma unroll
        for (int c = 0; c < D; ++c) {
            hq[c * HM + h] = q[c * V + v];
            hx[c * HM + h] = x[c * V + v];
        }
        if constexpr (ISO) ht[h] = tau[v];
    }
    __syncthreads();
    const int e0 = chE[k], e1 = chE[k + 1];
    for (int ue = e0 + (int)threadIdx.x; ue < e1; ue += BT) {
        int loc[D + 1], sl[D + 1];
#pragma unroll
        for (int l = 0; l <= D; ++l) {
            const int w = chLS[(size_t)l * NU + ue];
            loc[l] = w & 0xffff;
            sl[l] = (w >> 16) - 1;
        }
        double ed[D][D], xv[D + 1][D];
#pragma unroll
        for (int kk = 0; kk < D; ++kk)
#pragma unroll
            for (int c = 0; c < D; ++c) ed[kk][c] = hq[c * HM + loc[kk + 1]] - hq[c * HM + loc[0]];
#pragma unroll
        for (int l = 0; l <= D; ++l)
#pragma unroll
            for (int c = 0; c < D; ++c) xv[l][c] = hx[c * HM + loc[l]];
        double ub = 0.0;
        if constexpr (ISO) {
#pragma unroll
            for (int l = 0; l <= D; ++l) ub += ht[loc[l]];
            ub /= (D + 1);
        }
        double f1[D], f2[D];
        if constexpr (!ISO) {
#pragma unroll
            for (int c = 0; c < D; ++c) { f1[c] = frE[(size_t)c * NU + ue]; f2[c] = frE[(size_t)(D + c) * NU + ue]; }
        }
        double inv[D][D], det;
        mat_inv<D>(ed, inv, det);
        double fo[D + 1][D];
        elem_forces<D, ISO>(ISO ? chElm[ue] : 0, xv, ub, inv, elem_vol<D>(det), f1, f2, elmLyr, mod, fo);
#pragma unroll
        for (int l = 0; l <= D; ++l) {
            if (sl[l] < 0) continue;
#pragma unroll
            for (int c = 0; c < D; ++c) f[sl[l] * D + c] = fo[l][c];
        }
    }
    __syncthreads();
    const int j = chV[k] + (int)threadIdx.x;
    double ss = 0.0;
    if (j < chV[k + 1]) {
        const int i = chVtx[j];
        const int a0 = chLs[j], a1 = a0 + (vstart[i + 1] - vstart[i]);
        double acc[D];
#pragma unroll
        for (int c = 0; c < D; ++c) acc[c] = 0.0;
        for (int a = a0; a < a1; ++a)
#pragma unroll
            for (int c = 0; c < D; ++c) acc[c] += f[a * D + c];
        if (pfrc && i < BL) {
            double pa[D];
#pragma unroll
            for (int c = 0; c < D; ++c) pa[c] = 0.0;
            const int adj2 = vtxBtm[i];
            for (int a = 0; a < adj2; ++a) {
                const int b = vtxBtm[(1 + 2 * a) * BL + i], l = vtxBtm[(2 + 2 * a) * BL + i];
#pragma unroll
                for (int c = 0; c < D; ++c) pa[c] += pfrc[(l * D + c) * (size_t)B + b];
            }
#pragma unroll
            for (int c = 0; c < D; ++c) acc[c] = 1.0 * acc[c] + btmWgt * pa[c];
        }
#pragma unroll
        for (int c = 0; c < D; ++c) {
            const int kx = c * V + i;
            if constexpr (MODE == 0) {
                const double ad = ldmWgt * u[kx] + 1.0 * acc[c];
                out[kx] = ad;
                ss += w[kx] * ad;
            } else if constexpr (MODE == 2) {
                out[kx] = acc[c];
            } else {
                const double ax = ldmWgt * u[kx] + acc[c];
                const double r = 1.0 * w[kx] - 1.0 * ax;
                out[kx] = r;
                out2[kx] = r;
                ss += r * r;
            }
        }
    }
    if constexpr (MODE != 2) {
        const double sb = block_sum(ss, red);
        if (threadIdx.x == 0) P[k] = sb;
    }
}

// right-hand side |b|^2 and stopping threshold; bb == 0 -> x = 0 (ipcg.cu:25-30)
__global__ void pcg_start_k(const double *__restrict__ P, int nP, double tolSqu, int maxIt, Ctl *ctl) {
    __shared__ double red[NT / 64];
    double bb = reduce_partials(P, nP, red);
    if (threadIdx.x == 0) {
        cst(&ctl->bb, bb);
        cst(&ctl->tol, bb * tolSqu);
        cst(&ctl->done, (bb == 0.0) ? 1 : 0);
        cst(&ctl->status, 0);
        cst(&ctl->iters, 0);
        cst(&ctl->it, -1);
        cst(&ctl->maxIt, maxIt);
    }
}

// end of a control-point CG iteration: advance the iteration counter; the
// budget spent stops the solve with status 2
__global__ void ctl_next_k(Ctl *ctl) {
    if (cld(&ctl->done)) return;
    const int it = cld(&ctl->it) + 1;
    cst(&ctl->it, it);
    if (it >= cld(&ctl->maxIt)) { cst(&ctl->done, 1); cst(&ctl->status, 2); cst(&ctl->iters, it); }
}

__global__ void zero_if_empty_k(int n, double *__restrict__ x, double *__restrict__ a, const Ctl *ctl) {
    if (cld(&ctl->bb) != 0.0) return;
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) { x[i] = 0.0; if (a) a[i] = 0.0; }
}

// x += st d; a += st Kid; r -= st Ad with st = rKr / d.Ad; partial r.r
__global__ __launch_bounds__(NT) void pcg_update_k(int n, const double *__restrict__ P, int nP,
                                                   const double *__restrict__ hist, int it,
                                                   double *__restrict__ x, double *__restrict__ a,
                                                   double *__restrict__ r, const double *__restrict__ d,
                                                   const double *__restrict__ Kid, const double *__restrict__ Ad,
                                                   double *__restrict__ P2, Ctl *ctl) {
    __shared__ double red[NT / 64];
    if (cld(&ctl->done)) return;
    double dAd = reduce_partials(P, nP, red);
    if (dAd == 0.0) {   // ipcg.cu:63-68
        if (blockIdx.x == 0 && threadIdx.x == 0) { cst(&ctl->status, 1); cst(&ctl->done, 1); cst(&ctl->iters, it); }
        return;
    }
    double st = cld(&hist[it]) / dAd;
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    double ss = 0.0;
    if (i < n) {
        x[i] = 1.0 * x[i] + st * d[i];
        if (a) a[i] = a[i] + st * Kid[i];
        double ri = 1.0 * r[i] - st * Ad[i];
        r[i] = ri;
        ss = ri * ri;
    }
    double s = block_sum(ss, red);
    if (threadIdx.x == 0) P2[blockIdx.x] = s;
}

// beta = r.Kr_new / r.Kr_old; Kid = r + beta Kid; d = Kr + beta d
// it < 0: first direction d = Kr (ipcg.cu:41-53)
__global__ __launch_bounds__(NT) void pcg_dir_k(int n, const double *__restrict__ P, int nP, double *__restrict__ hist,
                                                int it, const double *__restrict__ r, const double *__restrict__ Kr,
                                                double *__restrict__ Kid, double *__restrict__ d, Ctl *ctl) {
    __shared__ double red[NT / 64];
    if (cld(&ctl->done)) return;
    double rKr = reduce_partials(P, nP, red);
    if (rKr == 0.0) {   // ipcg.cu:48-53, 83-88
        if (blockIdx.x == 0 && threadIdx.x == 0) { cst(&ctl->status, 1); cst(&ctl->done, 1); cst(&ctl->iters, it + 1); }
        return;
    }
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (it < 0) {
        if (blockIdx.x == 0 && threadIdx.x == 0) cst(&hist[0], rKr);
        if (i < n) d[i] = Kr[i];
        return;
    }
    double be = rKr / cld(&hist[it]);
    if (blockIdx.x == 0 && threadIdx.x == 0) cst(&hist[it + 1], rKr);
    if (i < n) {
        Kid[i] = 1.0 * r[i] + be * Kid[i];
        d[i] = 1.0 * Kr[i] + be * d[i];
    }
}

// pcg_dir_k for the stored all-landmark K, with kfin_k's work folded in:
// stop when |r|^2 (partials Prr) is below the tolerance; r.(K + eps I) r from
// the product's per-block partials of r.(its part of K r) (Pdot, nPdot of
// them) plus eps |r|^2; each thread forms its own entry of (K + eps I) r from
// the slice partials (slice order) and updates d and Kid with it
__global__ __launch_bounds__(NT) void pcg_dir_stored_k(int n, int slices, const double *__restrict__ part,
                                                       double eps, const double *__restrict__ Prr, int nPrr,
                                                       const double *__restrict__ Pdot, int nPdot,
                                                       double *__restrict__ hist, int it,
                                                       const double *__restrict__ r, double *__restrict__ Kid,
                                                       double *__restrict__ d, Ctl *ctl) {
    __shared__ double red[NT / 64];
    if (cld(&ctl->done)) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    // every operand is requested before the partials are reduced
    double kr = 0.0, rv = 0.0, kv = 0.0, dv = 0.0;
    if (i < n) {
        rv = r[i];
        if (it >= 0) { kv = Kid[i]; dv = d[i]; }
#pragma unroll 8
        for (int k = 0; k < slices; ++k) kr += part[(size_t)k * n + i];
    }
    const double rzOld = it >= 0 ? cld(&hist[it]) : 0.0;
    double pr = 0.0, pd = 0.0;
    for (int j = threadIdx.x; j < nPrr; j += NT) pr += Prr[j];
    for (int j = threadIdx.x; j < nPdot; j += NT) pd += Pdot[j];
    const double rr = block_sum(pr, red);
    if (rr < cld(&ctl->tol)) {
        if (blockIdx.x == 0 && threadIdx.x == 0) { cst(&ctl->done, 1); cst(&ctl->iters, it + 1); }
        return;
    }
    const double rKr = block_sum(pd, red) + eps * rr;
    if (rKr == 0.0) {   // ipcg.cu:48-53, 83-88
        if (blockIdx.x == 0 && threadIdx.x == 0) { cst(&ctl->status, 1); cst(&ctl->done, 1); cst(&ctl->iters, it + 1); }
        return;
    }
    kr += eps * rv;
    if (it < 0) {
        if (blockIdx.x == 0 && threadIdx.x == 0) cst(&hist[0], rKr);
        if (i < n) d[i] = kr;
        return;
    }
    const double be = rKr / rzOld;
    if (blockIdx.x == 0 && threadIdx.x == 0) cst(&hist[it + 1], rKr);
    if (i < n) {
        Kid[i] = 1.0 * rv + be * kv;
        d[i] = 1.0 * kr + be * dv;
    }
}

// ---------------------------------------------------------------------------
// y = (K + eps I) x, K_ij = k(|q_i - q_j| / width), matrix-free.
// Block = 64 rows x 4 column phases; columns are staged through LDS 256 at a
// time and every lane of a wave reads the same column (LDS broadcast).  The 4
// phase sums of a row are combined in a fixed order.
// CHECK: first reduce |r|^2 of the predecessor, stop when converged, and
// emit partials of x.y (r.Kr).
constexpr int MV_ROWS = 64;
constexpr int MV_PH = 4;

template <int D, int ORD, bool CHECK>
__global__ __launch_bounds__(NT) void kmatvec_k(int V, const double *__restrict__ q, const double *__restrict__ x,
                                                double *__restrict__ y, double eps, double invW,
                                                const double *__restrict__ Prr, int nPrr, double *__restrict__ P3,
                                                int it, Ctl *ctl) {
    __shared__ double sq[D][NT];
    __shared__ double sx[D][NT];
    __shared__ double racc[MV_PH][D][MV_ROWS];
    __shared__ double red[NT / 64];
    __shared__ double etab[64];
    if (ctl && cld(&ctl->done)) return;
    init_exp_table(etab);
    if constexpr (CHECK) {
        double rr = reduce_partials(Prr, nPrr, red);
        if (rr < cld(&ctl->tol)) {
            if (blockIdx.x == 0 && threadIdx.x == 0) { cst(&ctl->done, 1); cst(&ctl->iters, it + 1); }
            return;
        }
    }
    const int lr = threadIdx.x & (MV_ROWS - 1), ph = threadIdx.x / MV_ROWS;
    const int i = blockIdx.x * MV_ROWS + lr;
    double qi[D], acc[D];
#pragma unroll
    for (int c = 0; c < D; ++c) { qi[c] = (i < V) ? q[c * V + i] : 0.0; acc[c] = 0.0; }
    for (int j0 = 0; j0 < V; j0 += NT) {
        int j = j0 + threadIdx.x;
        __syncthreads();
#pragma unroll
        for (int c = 0; c < D; ++c) {
            sq[c][threadIdx.x] = (j < V) ? q[c * V + j] : 0.0;
            sx[c][threadIdx.x] = (j < V) ? x[c * V + j] : 0.0;
        }
        __syncthreads();
        int jn = min(NT, V - j0);
        for (int jj = ph; jj < jn; jj += MV_PH) {
            double d2 = 0.0;
#pragma unroll
            for (int c = 0; c < D; ++c) { double dd = qi[c] - sq[c][jj]; d2 += dd * dd; }
            double k = rkhs_kernel<D, ORD>(d2, invW, etab);
#pragma unroll
            for (int c = 0; c < D; ++c) acc[c] += k * sx[c][jj];
        }
    }
#pragma unroll
    for (int c = 0; c < D; ++c) racc[ph][c][lr] = acc[c];
    __syncthreads();
    double dotp = 0.0;
    if (ph == 0 && i < V) {
#pragma unroll
        for (int c = 0; c < D; ++c) {
            double s = racc[0][c][lr];
#pragma unroll
            for (int p = 1; p < MV_PH; ++p) s += racc[p][c][lr];
            double xi = x[c * V + i];
            s += eps * xi;
            y[c * V + i] = s;
            dotp += xi * s;
        }
    }
    if constexpr (CHECK) {
        double s = block_sum(dotp, red);
        if (threadIdx.x == 0) P3[blockIdx.x] = s;
    }
}

// rectangular y(i) = sum_j k(|a_i - b_j|) x(j) for the control-point form.
// blockIdx.y splits the column range into gridDim.y slices; slice s writes
// its partial sums to y + s * D * na (reduced afterwards in slice order) so
// short row ranges still fill the chip.
template <int D, int ORD>
__global__ __launch_bounds__(NT) void kcross_k(int na, const double *__restrict__ a, int nb,
                                               const double *__restrict__ b, const double *__restrict__ x,
                                               double *__restrict__ y, double invW, const Ctl *ctl) {
    __shared__ double sq[D][NT];
    __shared__ double sx[D][NT];
    __shared__ double racc[MV_PH][D][MV_ROWS];
    __shared__ double etab[64];
    if (ctl && cld(&ctl->done)) return;
    init_exp_table(etab);
    const int lr = threadIdx.x & (MV_ROWS - 1), ph = threadIdx.x / MV_ROWS;
    const int i = blockIdx.x * MV_ROWS + lr;
    const int per = (nb + gridDim.y - 1) / gridDim.y;
    const int jb = blockIdx.y * per, je = min(nb, jb + per);
    double qi[D], acc[D];
#pragma unroll
    for (int c = 0; c < D; ++c) { qi[c] = (i < na) ? a[c * na + i] : 0.0; acc[c] = 0.0; }
    for (int j0 = jb; j0 < je; j0 += NT) {
        int j = j0 + threadIdx.x;
        __syncthreads();
#pragma unroll
        for (int c = 0; c < D; ++c) {
            sq[c][threadIdx.x] = (j < je) ? b[c * nb + j] : 0.0;
            sx[c][threadIdx.x] = (j < je) ? x[c * nb + j] : 0.0;
        }
        __syncthreads();
        int jn = min(NT, je - j0);
        for (int jj = ph; jj < jn; jj += MV_PH) {
            double d2 = 0.0;
#pragma unroll
            for (int c = 0; c < D; ++c) { double dd = qi[c] - sq[c][jj]; d2 += dd * dd; }
            double k = rkhs_kernel<D, ORD>(d2, invW, etab);
#pragma unroll
            for (int c = 0; c < D; ++c) acc[c] += k * sx[c][jj];
        }
    }
#pragma unroll
    for (int c = 0; c < D; ++c) racc[ph][c][lr] = acc[c];
    __syncthreads();
    if (ph == 0 && i < na) {
        double *yo = y + (size_t)blockIdx.y * D * na;
#pragma unroll
        for (int c = 0; c < D; ++c) {
            double s = racc[0][c][lr];
#pragma unroll
            for (int p = 1; p < MV_PH; ++p) s += racc[p][c][lr];
            yo[c * na + i] = s;
        }
    }
}

// y = sum over slices of the kcross partials (slice order); optional |y|^2 partials
__global__ __launch_bounds__(NT) void slice_sum_k(int n, int slices, const double *__restrict__ part,
                                                  double *__restrict__ y, double *__restrict__ P, const Ctl *ctl) {
    __shared__ double red[NT / 64];
    if (ctl && cld(&ctl->done)) return;
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    double ss = 0.0;
    if (i < n) {
        double s = part[i];
        for (int k = 1; k < slices; ++k) s += part[(size_t)k * n + i];
        y[i] = s;
        ss = s * s;
    }
    if (P) {
        double t = block_sum(ss, red);
        if (threadIdx.x == 0) P[blockIdx.x] = t;
    }
}

// Panel-major storage of the dense products: rows are grouped in panels of
// panel_rows<T>() = 1024 / sizeof(T) rows (64 lanes x one 16-byte load), and
// a panel's columns follow one another, each holding the panel's rows, so a
// wavefront streaming consecutive columns reads one contiguous range.  The
// row count is padded to whole panels (padding rows hold zeros).
#ifndef SA_PANEL_BYTES
#define SA_PANEL_BYTES 1024
#endif
template <typename T>
#ifndef SA_PANEL_BYTES32
#define SA_PANEL_BYTES32 SA_PANEL_BYTES
#endif
__host__ __device__ constexpr int panel_rows() {
    return (sizeof(T) == 4 ? SA_PANEL_BYTES32 : SA_PANEL_BYTES) / (int)sizeof(T);
}

__host__ __device__ inline size_t panel_at(int i, int j, int PR, int ncols) {
    return (size_t)(i / PR) * PR * ncols + (size_t)j * PR + (i % PR);
}

template <int D>
__global__ void gather_ctl_k(int C, int V, const int *__restrict__ idx, const double *__restrict__ q,
                             double *__restrict__ qc) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= C) return;
    int v = idx[k];
#pragma unroll
    for (int c = 0; c < D; ++c) qc[c * C + k] = q[c * V + v];
}

// dense K_cc + eps I, column-major, for its Cholesky factorisation
template <int D, int ORD>
__global__ void kcc_build_k(int C, const double *__restrict__ qc, double invW, double eps, double *__restrict__ K,
                            double *__restrict__ K2) {
    __shared__ double etab[64];
    init_exp_table(etab);
    __syncthreads();
    long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long)C * C) return;
    int j = (int)(idx / C), i = (int)(idx - (long)j * C);
    double d2 = 0.0;
#pragma unroll
    for (int c = 0; c < D; ++c) { double dd = qc[c * C + i] - qc[c * C + j]; d2 += dd * dd; }
    double kv = rkhs_kernel<D, ORD>(d2, invW, etab);
    if (K) K[idx] = kv + (i == j ? eps : 0.0);
    if (K2) K2[panel_at(i, j, panel_rows<double>(), C)] = kv;   // K_cc without eps, panel-major
}

// Stored landmark x control-point kernel matrix, rebuilt once per time step:
// A (V x C) and its transpose At (C x V), both panel-major (panel_at), hold
// the same rounded values, so the two products of a CG iteration use exactly
// K and K^T.  T = double (plain FP64 form) or
// float (mixed form).  A 64 x 64 tile per block; At is written through LDS so
// both stores are coalesced.  For T = float, Alo / Atlo (optional) receive the
// FP32 remainder K - (double)A of each entry in the same layout: the pair
// gives the refinement residuals of the mixed form products accurate to about
// 2^-48 relative per entry from two FP32 matrix reads.
template <int D, int ORD, typename T>
__global__ __launch_bounds__(NT) void kstore_build_k(int V, const double *__restrict__ q, int C,
                                                     const double *__restrict__ qc, double invW, T *__restrict__ A,
                                                     T *__restrict__ At, T *__restrict__ Alo = nullptr,
                                                     T *__restrict__ Atlo = nullptr) {
    constexpr int PR = panel_rows<T>();
    __shared__ double etab[64];
    __shared__ T tile[64][65];
    __shared__ T tlo[sizeof(T) == 4 ? 64 : 1][65];
    init_exp_table(etab);
    __syncthreads();
    const int i0 = blockIdx.x * 64, j0 = blockIdx.y * 64;
    const int li = threadIdx.x & 63, lj0 = threadIdx.x >> 6;   // 4 column lanes
    const int i = i0 + li;
    double qi[D];
#pragma unroll
    for (int c = 0; c < D; ++c) qi[c] = i < V ? q[c * V + i] : 0.0;
    for (int lj = lj0; lj < 64; lj += 4) {
        int j = j0 + lj;
        T v = 0, vl = 0;
        if (i < V && j < C) {
            double d2 = 0.0;
#pragma unroll
            for (int c = 0; c < D; ++c) { double dd = qi[c] - qc[c * C + j]; d2 += dd * dd; }
            const double kv = rkhs_kernel<D, ORD>(d2, invW, etab);
            v = (T)kv;
            A[panel_at(i, j, PR, C)] = v;
            if constexpr (sizeof(T) == 4) {
                // FP32 remainder: v + vl equals kv to about 2^-48 relative
                vl = (T)(kv - (double)v);
                if (Alo) Alo[panel_at(i, j, PR, C)] = vl;
            }
        }
        tile[lj][li] = v;
        if constexpr (sizeof(T) == 4) tlo[lj][li] = vl;
    }
    __syncthreads();
    // transposed store: lanes run along j
    const int lj = threadIdx.x & 63, li0 = threadIdx.x >> 6;
    for (int l = li0; l < 64; l += 4) {
        int ii = i0 + l, jj = j0 + lj;
        if (At && ii < V && jj < C) At[panel_at(jj, ii, PR, V)] = tile[lj][l];
        if constexpr (sizeof(T) == 4)
            if (Atlo && ii < V && jj < C) Atlo[panel_at(jj, ii, PR, V)] = tlo[lj][l];
    }
}

// y = M x for a stored panel-major M (nr x nc, see panel_at), x and y with D
// columns (stride nc and nr).  A block owns one panel of 64 R rows (R per
// lane, one 16-byte load per column) and one slice of at most KS_MAX_COLS
// columns (blockIdx.y); it stages its slice of x in LDS (dynamic, D x slice
// doubles), then its wavefronts each stream a contiguous quarter of the slice
// (consecutive columns of a panel are adjacent in memory) with no further
// barrier, reading x from LDS as a broadcast, and are summed in wavefront
// order through LDS at the end (FP64; the FP32 wavefronts take interleaved
// columns instead, see the kernel).  The block's partial goes to
// part + s * D * nr; slice_rows_k sums the slices in slice order.
constexpr int KS_MAX_COLS = 1024;
// most Ritz vectors in the spectral term of the control-point preconditioner
#ifndef SA_SPEC_KMAX
#define SA_SPEC_KMAX 128
#endif
constexpr int SPEC_KMAX = SA_SPEC_KMAX;
// slice-sum consumers (slice_rows_k, slice_combine_k, slice_dot_k): a block
// sums the slice partials of 64 rows with SLW wavefronts, wavefront w adding
// slices w, w + SLW, ..., the wavefront sums then added in order.  Sixteen
// wavefronts keep enough loads in flight for the 100-250 slices of the
// control-point products (a block per 64 rows gives only ~1 wavefront per
// SIMD at C = 5,000); vectors of more than 64 Ki rows fill the chip with
// four (slice_rows_k<4>).
#ifndef SA_SLICE_WAVES
#define SA_SLICE_WAVES 16
#endif
constexpr int SLW = SA_SLICE_WAVES;
constexpr int NTS = 64 * SLW;
#ifndef SA_KS_BLOCKS
#define SA_KS_BLOCKS 1024
#endif
constexpr int KS_BLOCKS = SA_KS_BLOCKS;   // target blocks per stored product

// one 16-byte load of a stored matrix, streamed once per product: the
// non-temporal hint (SA_KS_NT, default on) keeps it from displacing cached
// data; measured 0.64 -> 0.69 (FP32) and 0.72 -> 0.80 (FP64) of HBM peak
#ifndef SA_KS_NT
#define SA_KS_NT 1
#endif
typedef float ks_f4 __attribute__((ext_vector_type(4)));
typedef double ks_d2 __attribute__((ext_vector_type(2)));
// (NTL = false: an ordinary load, for a matrix small enough to stay cached
// from one product to the next)
template <bool NTL = true>
__device__ inline float4 ks_load(const float4 *p) {
    if constexpr (NTL && SA_KS_NT) {
        ks_f4 v = __builtin_nontemporal_load(reinterpret_cast<const ks_f4 *>(p));
        return make_float4(v.x, v.y, v.z, v.w);
    } else {
        return *p;
    }
}
template <bool NTL = true>
__device__ inline double2 ks_load(const double2 *p) {
    if constexpr (NTL && SA_KS_NT) {
        ks_d2 v = __builtin_nontemporal_load(reinterpret_cast<const ks_d2 *>(p));
        return make_double2(v.x, v.y);
    } else {
        return *p;
    }
}

// a[r0 + k][c] += m.k * xv for the elements of one 16-byte load
template <int R, int D>
__device__ inline void fma_rows(double (&a)[R][D], int r0, int c, const float4 &m, double xv) {
    a[r0][c] += (double)m.x * xv; a[r0 + 1][c] += (double)m.y * xv;
    a[r0 + 2][c] += (double)m.z * xv; a[r0 + 3][c] += (double)m.w * xv;
}
template <int R, int D>
__device__ inline void fma_rows(double (&a)[R][D], int r0, int c, const double2 &m, double xv) {
    a[r0][c] += m.x * xv; a[r0 + 1][c] += m.y * xv;
}

// dotv / Pdot: the block's partial of dotv . (its slice's part of M x) over
// its rows, to Pdot[blockIdx.y * gridDim.x + blockIdx.x] (dotv has D columns
// of stride nr)
//
// upd (all-landmark CG, square M): the step of pcg_update_k taken while the
// input is staged.  st = hist[it] / d.Ad (partials P); the product's input
// is r' = rIn - st Ad, formed by every block for its columns (and, for the
// x.y partials, for its rows: dotv is then ignored); the first row panel
// writes r' to rOut (another buffer than rIn, which other blocks still
// read), x += st d, a += st Kid, and one |r'|^2 partial per slice to Prr.
struct KsUpd {
    const double *P = nullptr;
    int nP = 0;
    const double *hist = nullptr;
    int it = 0;
    const double *rIn = nullptr, *Ad = nullptr, *d = nullptr, *Kid = nullptr;
    double *rOut = nullptr, *xs = nullptr, *as = nullptr, *Prr = nullptr;
    Ctl *stop = nullptr;
};

template <int D, typename T, bool STEP = false, bool NTL = true>
__global__ __launch_bounds__(NT) void kstore_mv_k(int nr, int nc, const T *__restrict__ M,
                                                  const double *__restrict__ x, double *__restrict__ part,
                                                  const Ctl *ctl, Ctl *adv = nullptr,
                                                  const double *__restrict__ dz = nullptr,
                                                  const double *__restrict__ dP = nullptr, int dnP = 0,
                                                  double *__restrict__ dhist = nullptr, int dj = 0,
                                                  double *__restrict__ dout = nullptr,
                                                  const double *__restrict__ dotv = nullptr,
                                                  double *__restrict__ Pdot = nullptr, KsUpd upd = KsUpd()) {
    constexpr int PR = panel_rows<T>();
    constexpr int R = PR / 64;             // rows per lane
    constexpr int W = NT / 64;
#ifndef SA_KS_U
#define SA_KS_U 4
#endif
    constexpr int U = SA_KS_U;   // FP64: columns whose loads are issued together
    // FP32: columns of each wavefront in flight, a ring of KP loads.  The
    // 51,210 x 5,000 products of the mixed solve on one box: 2 -> 169.0 us
    // per product (16.77 steps/s), 4 -> 163.5 us (17.16), 8 -> 169.6 us
    // (16.66; 86 VGPRs, 5 waves per SIMD)
#ifndef SA_KS_P
#define SA_KS_P 4
#endif
    constexpr int KP = SA_KS_P;
    using V4 = typename std::conditional<sizeof(T) == 4, float4, double2>::type;
    constexpr int NV = R * (int)sizeof(T) / 16;   // 16-byte loads per lane and column
    constexpr int EV = 16 / (int)sizeof(T);       // elements per 16-byte load
    // D x (je - jb); FP32: then the (W-1) x D x PR wavefront partials
    // (ks_lds_bytes), which FP64 keeps in a separate array
    extern __shared__ double sx[];
    __shared__ double redd[W];
    if (ctl && cld(&ctl->done)) return;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int r0 = blockIdx.x * PR + R * lane;
    const int per = (nc + gridDim.y - 1) / gridDim.y;
    const int jb = blockIdx.y * per, je = min(nc, jb + per);
    const int ns = max(0, je - jb);
    const T *base = M + (size_t)blockIdx.x * PR * nc + (size_t)jb * PR + R * lane;
    // FP32: each wavefront's first KP columns (the load ring below) are
    // requested before the input is staged, so the stream starts while the
    // staging's loads and reductions are in flight
    V4 m[KP][NV];
    if constexpr (sizeof(T) == 4) {
#pragma unroll
        for (int k = 0; k < KP; ++k)
            if (wv + k * W < ns)
#pragma unroll
                for (int v = 0; v < NV; ++v)
                    m[k][v] = ks_load<NTL>(reinterpret_cast<const V4 *>(base + (size_t)(wv + k * W) * PR + v * EV));
    }
    // the staged input's operands (x; z and d_old; r and Ad), a thread's
    // columns j0 + s NT + threadIdx.x (s < SJ) requested together: the
    // first pass before the partials below are reduced
    constexpr int SJ = 2;
    // the CG step (upd) is taken by the STEP instantiation only (FP64)
    static_assert(!STEP || sizeof(T) == 8, "the CG step is taken by the FP64 product");
    constexpr bool UPD = STEP;
    const bool upd0 = UPD && upd.P && blockIdx.x == 0;
    double va[SJ][D], vb[SJ][D];
    double px[UPD ? SJ : 1][D], pd[UPD ? SJ : 1][D], pa[UPD ? SJ : 1][D], pk[UPD ? SJ : 1][D];
    // (block-uniform column bases, so the loads take 32-bit lane offsets)
    const double *__restrict__ srcA = dz ? dz : UPD && upd.P ? upd.rIn : x;
    const double *__restrict__ srcB = dz ? x : upd.Ad;
    const bool two = dz || (UPD && upd.P);
    auto fetch = [&](int j0) {
#pragma unroll
        for (int c = 0; c < D; ++c) {
            const size_t o0 = (size_t)c * nc + jb;
#pragma unroll
            for (int s = 0; s < SJ; ++s) {
                const int j = j0 + s * NT + (int)threadIdx.x;
                if (j < ns) {
                    va[s][c] = srcA[o0 + j];
                    if (two) vb[s][c] = srcB[o0 + j];
                    if constexpr (UPD) {
                        // the first row panel's x += st d and a += st Kid
                        if (upd0) {
                            px[s][c] = upd.xs[o0 + j];
                            pd[s][c] = upd.d[o0 + j];
                            if (upd.as) { pa[s][c] = upd.as[o0 + j]; pk[s][c] = upd.Kid[o0 + j]; }
                        }
                    }
                }
            }
        }
    };
    fetch(0);
    double be = 0.0;
    if (dz) {
        // the CG direction of iteration dj, d = z + beta d_old with beta =
        // rz_dj / rz_(dj-1) (rz_dj: the r.z partials dP), staged here and
        // written to dout by the first row panel (as ctl_dir_k would)
        const double rz = reduce_partials(dP, dnP, redd);
        if (rz == 0.0) {
            if (blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0) {
                cst(&adv->status, 1); cst(&adv->done, 1); cst(&adv->iters, dj);
            }
            return;
        }
        be = rz / cld(&dhist[dj - 1]);
        if (blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0) cst(&dhist[dj], rz);
    }
    if (adv && blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0) {
        // first kernel of a CG iteration: advance the iteration counter (no
        // kernel of this launch reads it); the budget spent stops the solve
        const int it = cld(&adv->it) + 1;
        cst(&adv->it, it);
        if (it >= cld(&adv->maxIt)) { cst(&adv->done, 1); cst(&adv->status, 2); cst(&adv->iters, it); }
    }
    double st = 0.0;
    if (UPD && upd.P) {
        // pcg_update_k's step (ipcg.cu:63-68 on zero curvature)
        const double dAd = reduce_partials(upd.P, upd.nP, redd);
        if (dAd == 0.0) {
            if (blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0) {
                cst(&upd.stop->status, 1); cst(&upd.stop->done, 1); cst(&upd.stop->iters, upd.it);
            }
            return;
        }
        st = cld(&upd.hist[upd.it]) / dAd;
    }
    double rrp = 0.0;
    for (int j0 = 0; j0 < ns; j0 += SJ * NT) {
        if (j0 > 0) fetch(j0);
#pragma unroll
        for (int c = 0; c < D; ++c)
#pragma unroll
            for (int s = 0; s < SJ; ++s) {
                const int j = j0 + s * NT + (int)threadIdx.x;
                if (j >= ns) continue;
                const size_t o = (size_t)c * nc + jb + j;
                double v = va[s][c];
                if (dz) v = v + be * vb[s][c];
                else if (UPD && upd.P) v = 1.0 * v - st * vb[s][c];
                sx[c * ns + j] = v;
                if (dz && blockIdx.x == 0) dout[o] = v;
                if constexpr (UPD) {
                    if (upd0) {
                        upd.rOut[o] = v;
                        upd.xs[o] = 1.0 * px[s][c] + st * pd[s][c];
                        if (upd.as) upd.as[o] = pa[s][c] + st * pk[s][c];
                        rrp += v * v;
                    }
                }
            }
    }
    if (upd0) {
        const double t = block_sum(rrp, redd);   // (its barriers also order the staging)
        if (threadIdx.x == 0) upd.Prr[blockIdx.y] = t;
    }
    __syncthreads();
    double a[R][D];
#pragma unroll
    for (int k = 0; k < R; ++k)
#pragma unroll
        for (int c = 0; c < D; ++c) a[k][c] = 0.0;
    double *red;   // [W - 1][D][PR]
    if constexpr (sizeof(T) == 4) {
        // FP32: the wavefronts of a block take interleaved columns (w, w + W,
        // ...), so the block reads one contiguous run of W KiB at a time, and
        // each wavefront keeps its next KP - 1 columns' loads in flight while
        // it uses the current one.  The DRAM then sees one sequential stream per
        // block instead of one per wavefront.  The wavefront partials reuse
        // the LDS of x, so a block needs at most 24 KiB and the tile table's
        // launch fits the chip in one round.  Measured on the 51,210 x 5,000
        // products and their transposes: 0.81-0.83 of HBM peak for the
        // kernel alone against 0.72-0.75 for contiguous quarter-slices per
        // wavefront; 0.79 inside the solve.
        for (int j = wv; j < ns; j += KP * W) {
#pragma unroll
            for (int k = 0; k < KP; ++k) {
                const int jj = j + k * W;
                if (jj < ns) {
#pragma unroll
                    for (int c = 0; c < D; ++c) {
                        const double xv = sx[c * ns + jj];
#pragma unroll
                        for (int v = 0; v < NV; ++v) fma_rows(a, v * EV, c, m[k][v], xv);
                    }
                    if (jj + KP * W < ns)
#pragma unroll
                        for (int v = 0; v < NV; ++v)
                            m[k][v] = ks_load<NTL>(reinterpret_cast<const V4 *>(base + (size_t)(jj + KP * W) * PR + v * EV));
                }
            }
        }
        __syncthreads();   // every wavefront is done with x
        red = sx;
    } else {
        __shared__ double red64[(W - 1) * D * PR];
        red = red64;
        const int pw = (ns + W - 1) / W;
        const int wb = min(ns, wv * pw), we = min(ns, wb + pw);   // slice-relative columns of this wavefront
        const T *col = M + (size_t)blockIdx.x * PR * nc + (size_t)(jb + wb) * PR + R * lane;
        int j = wb;
        for (; j + U <= we; j += U, col += U * PR) {
            V4 m[U][NV];
#pragma unroll
            for (int u = 0; u < U; ++u)
#pragma unroll
                for (int v = 0; v < NV; ++v) m[u][v] = ks_load<NTL>(reinterpret_cast<const V4 *>(col + u * PR + v * EV));
#pragma unroll
            for (int u = 0; u < U; ++u)
#pragma unroll
                for (int c = 0; c < D; ++c) {
                    const double xv = sx[c * ns + j + u];
#pragma unroll
                    for (int v = 0; v < NV; ++v) fma_rows(a, v * EV, c, m[u][v], xv);
                }
        }
        if (j < we) {
            // the last (we - j < U) columns: their loads also go out together
            V4 m[U][NV];
#pragma unroll
            for (int u = 0; u < U; ++u)
                if (j + u < we)
#pragma unroll
                    for (int v = 0; v < NV; ++v) m[u][v] = ks_load<NTL>(reinterpret_cast<const V4 *>(col + u * PR + v * EV));
#pragma unroll
            for (int u = 0; u < U; ++u)
                if (j + u < we)
#pragma unroll
                    for (int c = 0; c < D; ++c) {
                        const double xv = sx[c * ns + j + u];
#pragma unroll
                        for (int v = 0; v < NV; ++v) fma_rows(a, v * EV, c, m[u][v], xv);
                    }
        }
    }
    // FP64: wavefront 0 requests its rows of the dot product's vectors
    // before the barrier below (the FP32 products take no dot product)
    constexpr bool DOT = sizeof(T) == 8;
    double e0[DOT ? R : 1][D], e1[DOT ? R : 1][D];
    if constexpr (DOT) {
        if (wv == 0 && (UPD ? upd.P != nullptr : dotv != nullptr))
#pragma unroll
            for (int c = 0; c < D; ++c)
#pragma unroll
                for (int k = 0; k < R; ++k)
                    if (r0 + k < nr) {
                        const size_t o = (size_t)c * nr + r0 + k;
                        if (UPD) { e0[k][c] = upd.rIn[o]; e1[k][c] = upd.Ad[o]; }
                        else e0[k][c] = dotv[o];
                    }
    }
    // wavefront order: ((w0 + w1) + w2) + w3
    if (wv > 0) {
#pragma unroll
        for (int c = 0; c < D; ++c)
#pragma unroll
            for (int k = 0; k < R; ++k) red[((wv - 1) * D + c) * PR + lane * R + k] = a[k][c];
    }
    __syncthreads();
    if (wv == 0) {
        double *po = part + (size_t)blockIdx.y * D * nr;
        double dp = 0.0;
#pragma unroll
        for (int c = 0; c < D; ++c)
#pragma unroll
            for (int k = 0; k < R; ++k) {
                double sum = a[k][c];
#pragma unroll
                for (int w = 0; w < W - 1; ++w) sum += red[(w * D + c) * PR + lane * R + k];
                if (r0 + k < nr) {
                    po[c * nr + r0 + k] = sum;
                    if constexpr (DOT) {
                        if (UPD && upd.P) dp += (1.0 * e0[k][c] - st * e1[k][c]) * sum;
                        else if (!UPD && dotv) dp += e0[k][c] * sum;
                    }
                }
            }
        if (Pdot) {
            dp = wave_sum(dp);
            if (lane == 0) Pdot[(size_t)blockIdx.y * gridDim.x + blockIdx.x] = dp;
        }
    }
}

// dynamic LDS of a kstore_mv_k launch over ncol columns in `slices` slices:
// the block's D x slice copy of x, for FP32 at least the wavefront partials
template <int D, typename T>
inline size_t ks_lds_bytes(int ncol, int slices) {
    const size_t ns = (size_t)((ncol + slices - 1) / slices);
    const size_t red = sizeof(T) == 4 ? (size_t)(NT / 64 - 1) * D * panel_rows<T>() : 0;
    return sizeof(double) * std::max((size_t)D * ns, red);
}

// ---------------------------------------------------------------------------
// Matrix-free kernel-matrix product on FP32 matrix cores (mixed form, 3D and
// Gaussian kernels): part[s][c][i] = sum over the columns j of slice s of
// k(|a_i - b_j| / w) x_c(j), for y = K_xc u (rows landmarks, columns control
// points) and y = K_cx y (the roles swapped).  Rows and columns are visited
// in a Morton order of the points (rperm / cperm: position -> index), so the
// rows of a block and the columns near them lie close together; outputs and
// inputs stay in index order.
//
// Per 16 x 16 tile of (column j, row i) pairs, with coordinates scaled by
// 1 / w and shifted by the block's origin o (its first row point, FP64 before
// rounding to FP32):
//   one v_mfma_f32_16x16x4f32:  D[j][i] = |b_j|^2 - 2 b_j . a_i + |a_i|^2
//       (A[j][k] = [b_j, |b_j|^2], B[k][i] = [-2 a_i, 1], the accumulator
//       started at |a_i|^2), i.e. rho^2 of the pair;
//   the kernel value k(rho) on the VALU (sqrt and exp2 instructions);
//   four v_mfma_f32_16x16x4f32:  Y[i][c] += sum_j D[j][i] x_c(j), the tile's
//       kernel values (still in the accumulator layout) as the A operand.
// Y accumulates in FP32 over a staged chunk of 256 columns and is then added
// to FP64 per-lane sums.  A block is four wavefronts of KXM_RT row tiles each
// (KXM_ROWS rows); blockIdx.y selects the column slice.  Accumulator layout
// of the 16x16x4 FP32 instruction: lane l, element r holds row
// kxm_row(l / 16, r), column l % 16.  The direction / iteration-counter
// arguments are those of kstore_mv_k (the first product of a fused CG
// iteration).
#ifndef SA_KXM_GROUPED
#define SA_KXM_GROUPED 1
#endif
__device__ constexpr int kxm_row(int g4, int r) { return SA_KXM_GROUPED ? 4 * g4 + r : 4 * r + g4; }
constexpr int KXM_RT = 2;
#ifndef SA_KXM_BLOCKS
#define SA_KXM_BLOCKS 2048
#endif
constexpr int KXM_BLOCKS = SA_KXM_BLOCKS;   // target blocks per product (row blocks x column slices)
constexpr int KXM_ROWS = 4 * 16 * KXM_RT;
constexpr int KXM_CHUNK = 256;
typedef float kxm_f4 __attribute__((ext_vector_type(4)));

template <int ORD>
__device__ inline float kxm_kernel(float rho2) {
    constexpr float kLog2e = 1.4426950408889634f;
    rho2 = fmaxf(rho2, 0.0f);
    if constexpr (ORD == -1) {
        return __builtin_amdgcn_exp2f(-rho2 * kLog2e);
    } else {
        const float r = __builtin_amdgcn_sqrtf(rho2);
        const float e = __builtin_amdgcn_exp2f(-r * kLog2e);
        if constexpr (ORD == 0) return e;
        if constexpr (ORD == 1) return fmaf(r, e, e);
        if constexpr (ORD == 2) return (3.0f + r * (3.0f + r)) * (1.0f / 3.0f) * e;
        if constexpr (ORD == 3) return (15.0f + r * (15.0f + r * (6.0f + r))) * (1.0f / 15.0f) * e;
        return (105.0f + r * (105.0f + r * (45.0f + r * (10.0f + r)))) * (1.0f / 105.0f) * e;
    }
}

template <int D, int ORD>
__global__ __launch_bounds__(NT) void kx_mfma_k(int nr, const double *__restrict__ a, const int *__restrict__ rperm,
                                                int nc, const double *__restrict__ b, const int *__restrict__ cperm,
                                                const double *__restrict__ x, double invW, double *__restrict__ part,
                                                const Ctl *ctl, Ctl *adv = nullptr,
                                                const double *__restrict__ dz = nullptr,
                                                const double *__restrict__ dP = nullptr, int dnP = 0,
                                                double *__restrict__ dhist = nullptr, int dj = 0,
                                                double *__restrict__ dout = nullptr) {
    static_assert(NT == 256 && KXM_CHUNK == NT, "one staged column per thread");
    static_assert(D == 2 || D == 3, "K = 4 holds the D coordinates and |b|^2");
    __shared__ float colb[KXM_CHUNK * 4];   // [b_j - o (D, zero padded to 3), |b_j - o|^2]
    __shared__ float colx[KXM_CHUNK * 4];   // x_c(j), c < D, zero padded
    __shared__ double redd[NT / 64];
    if (ctl && cld(&ctl->done)) return;
    double be = 0.0;
    if (dz) {
        // the CG direction d = z + beta d_old, beta = rz_dj / rz_(dj-1), as kstore_mv_k forms it
        const double rz = reduce_partials(dP, dnP, redd);
        if (rz == 0.0) {
            if (blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0) {
                cst(&adv->status, 1); cst(&adv->done, 1); cst(&adv->iters, dj);
            }
            return;
        }
        be = rz / cld(&dhist[dj - 1]);
        if (blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0) cst(&dhist[dj], rz);
    }
    if (adv && blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0) {
        const int it = cld(&adv->it) + 1;
        cst(&adv->it, it);
        if (it >= cld(&adv->maxIt)) { cst(&adv->done, 1); cst(&adv->status, 2); cst(&adv->iters, it); }
    }
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int n16 = lane & 15, g4 = lane >> 4;
    const int row0 = blockIdx.x * KXM_ROWS;
    // the block's origin: its first row point (scaled)
    double o[D];
    {
        const int io = rperm[min(row0, nr - 1)];
#pragma unroll
        for (int c = 0; c < D; ++c) o[c] = a[(size_t)c * nr + io] * invW;
    }
    // B operand of the distance tiles (row i = tile base + n16, slot g4) and
    // the accumulator start |a_i|^2, per row tile of this wavefront
    float bOp[KXM_RT], na[KXM_RT];
#pragma unroll
    for (int t = 0; t < KXM_RT; ++t) {
        const int i = row0 + (wv * KXM_RT + t) * 16 + n16;
        float ac[3] = {0.0f, 0.0f, 0.0f};
        if (i < nr) {
            const int ii = rperm[i];
#pragma unroll
            for (int c = 0; c < D; ++c) ac[c] = (float)(a[(size_t)c * nr + ii] * invW - o[c]);
        }
        na[t] = ac[0] * ac[0] + ac[1] * ac[1] + ac[2] * ac[2];
        bOp[t] = g4 < 3 ? -2.0f * ac[g4] : 1.0f;
    }
    double y64[KXM_RT][4];
#pragma unroll
    for (int t = 0; t < KXM_RT; ++t)
#pragma unroll
        for (int r = 0; r < 4; ++r) y64[t][r] = 0.0;
    const int per = (nc + gridDim.y - 1) / gridDim.y;
    const int jb = blockIdx.y * per, je = min(nc, jb + per);
    for (int j0 = jb; j0 < je; j0 += KXM_CHUNK) {
        __syncthreads();
        {
            const int j = j0 + threadIdx.x;
            float bc[3] = {0.0f, 0.0f, 0.0f}, xv[4] = {0.0f, 0.0f, 0.0f, 0.0f};
            if (j < je) {
                const int jj = cperm[j];
#pragma unroll
                for (int c = 0; c < D; ++c) {
                    bc[c] = (float)(b[(size_t)c * nc + jj] * invW - o[c]);
                    const size_t k = (size_t)c * nc + jj;
                    double v;
                    if (dz) {
                        v = dz[k] + be * x[k];
                        if (blockIdx.x == 0) dout[k] = v;
                    } else {
                        v = x[k];
                    }
                    xv[c] = (float)v;
                }
            }
            float *cb = colb + 4 * threadIdx.x;
            cb[0] = bc[0]; cb[1] = bc[1]; cb[2] = bc[2];
            cb[3] = bc[0] * bc[0] + bc[1] * bc[1] + bc[2] * bc[2];
            float *cx = colx + 4 * threadIdx.x;
            cx[0] = xv[0]; cx[1] = xv[1]; cx[2] = xv[2]; cx[3] = xv[3];
        }
        __syncthreads();
        // columns past je carry x = 0: whole tiles of the chunk are evaluated
        const int ntile = (min(KXM_CHUNK, je - j0) + 15) >> 4;
        kxm_f4 y32[KXM_RT];
#pragma unroll
        for (int t = 0; t < KXM_RT; ++t) y32[t] = kxm_f4{0.0f, 0.0f, 0.0f, 0.0f};
        for (int tt = 0; tt < ntile; ++tt) {
            const float aOp = colb[(tt * 16 + n16) * 4 + g4];
            // x operands of the four contraction products: column
            // kxm_row(g4, r) (the accumulator row of element r), component n16
            float xr[4];
#pragma unroll
            for (int r = 0; r < 4; ++r) xr[r] = n16 < 4 ? colx[(tt * 16 + kxm_row(g4, r)) * 4 + n16] : 0.0f;
#pragma unroll
            for (int t = 0; t < KXM_RT; ++t) {
                const kxm_f4 c0 = {na[t], na[t], na[t], na[t]};
                const kxm_f4 d2 = __builtin_amdgcn_mfma_f32_16x16x4f32(aOp, bOp[t], c0, 0, 0, 0);
#pragma unroll
                for (int r = 0; r < 4; ++r)
                    y32[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(kxm_kernel<ORD>(d2[r]), xr[r], y32[t], 0, 0, 0);
            }
        }
#pragma unroll
        for (int t = 0; t < KXM_RT; ++t)
#pragma unroll
            for (int r = 0; r < 4; ++r) y64[t][r] += (double)y32[t][r];
    }
    // lane (g4, n16 < D), element r: row kxm_row(g4, r) of each tile, component n16
    if (n16 < D) {
        double *po = part + (size_t)blockIdx.y * D * nr + (size_t)n16 * nr;
#pragma unroll
        for (int t = 0; t < KXM_RT; ++t)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int i = row0 + (wv * KXM_RT + t) * 16 + kxm_row(g4, r);
                if (i < nr) po[rperm[i]] = y64[t][r];
            }
    }
}

// control-point CG, fused consumers of the stored-product slice partials
// (64 outputs per block, wavefront w adding slices w, w + SLW, ..., combined
// in wavefront order as in slice_rows_k):
//
// slice_combine_k: y = sum of the K_cx (B K_xc d) partials; out = S d = y +
// ldmWgt (K_cc d + eps d), K_cc d read from w = K_xc d at the control-point
// rows; block partial of d.(S d).  recD non-null: iteration it < recMax also
// records d, S d and the residual r entering the iteration in column it of
// recD, recSD, recR (n x recMax each), the Krylov basis the spectral
// preconditioner is harvested from.
template <int D>
__global__ __launch_bounds__(NTS) void slice_combine_k(int C, int V, int slices, const double *__restrict__ part,
                                                       const double *__restrict__ w, const int *__restrict__ idx,
                                                       const double *__restrict__ d, double ldmWgt, double eps,
                                                       double *__restrict__ out, double *__restrict__ P,
                                                       double *__restrict__ recD, double *__restrict__ recSD,
                                                       double *__restrict__ recR, const double *__restrict__ r,
                                                       int recMax, const Ctl *ctl) {
    __shared__ double red[SLW][64];
    __shared__ double red2[SLW];
    if (cld(&ctl->done)) return;
    const int n = C * D;
    const int it = recD ? cld(&ctl->it) : recMax;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int i = blockIdx.x * 64 + lane;
    // wavefront 0 requests its operands (K_cc d through the control-point
    // index, d, r) before the slice partials are summed
    double kd = 0.0, dv = 0.0, rv = 0.0;
    if (wv == 0 && i < n) {
        const int c = i / C;
        kd = w[(size_t)c * V + idx[i - c * C]];
        dv = d[i];
        if (it < recMax) rv = r[i];
    }
    double s = 0.0;
    if (i < n)
#pragma unroll 4
        for (int k = wv; k < slices; k += SLW) s += part[(size_t)k * n + i];
    red[wv][lane] = s;
    __syncthreads();
    double ss = 0.0;
    if (wv == 0 && i < n) {
        double y = red[0][lane];
#pragma unroll
        for (int ww = 1; ww < SLW; ++ww) y += red[ww][lane];
        const double sd = y + ldmWgt * (kd + eps * dv);
        out[i] = sd;
        ss = dv * sd;
        if (it < recMax) {
            const size_t o = (size_t)it * n + i;
            recD[o] = dv;
            recSD[o] = sd;
            recR[o] = rv;
        }
    }
    const double t = block_sum(ss, red2);
    if (threadIdx.x == 0) P[blockIdx.x] = t;
}

// slice_dot_k: after the residual update, record |r|^2 (partials Prr) in
// the control block and stop when it is below the tolerance; else z = M r
// from the preconditioner product's partials
// (plus, with U, the spectral term U diag(cf) w, w = U^T r summed from the
// nPw block partials wp of ctl_update_k) and the block partial of r.z
__global__ __launch_bounds__(NTS) void slice_dot_k(int n, int slices, const double *__restrict__ part,
                                                   const double *__restrict__ Prr, int nPrr,
                                                   const double *__restrict__ r, double *__restrict__ z,
                                                   double *__restrict__ P3, const double *__restrict__ U,
                                                   const double *__restrict__ cf, const double *__restrict__ wp,
                                                   int nPw, int k, Ctl *ctl) {
    constexpr int G = NTS / SPEC_KMAX;   // threads summing each w entry
    static_assert(G >= 1, "slice-consumer block smaller than the spectral term");
    __shared__ double red[SLW][64];
    __shared__ double red2[SLW];
    __shared__ double cwp[G][SPEC_KMAX];
    __shared__ double cw[SPEC_KMAX];
    if (cld(&ctl->done)) return;
    const int it = cld(&ctl->it);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int i = blockIdx.x * 64 + lane;
    // the slice partials and this thread's spectral basis entries are
    // requested before the reductions (they depend on neither)
    double s = 0.0;
    if (i < n)
#pragma unroll 4
        for (int q = wv; q < slices; q += SLW) s += part[(size_t)q * n + i];
    constexpr int UB = (SPEC_KMAX + SLW - 1) / SLW;
    double uv[UB];
#pragma unroll
    for (int t = 0; t < UB; ++t) {
        const int b = wv + t * SLW;
        uv[t] = U && i < n && b < k ? U[(size_t)b * n + i] : 0.0;
    }
    double ws = 0.0;
    const int b = threadIdx.x % SPEC_KMAX, g = threadIdx.x / SPEC_KMAX;
    if (U && b < k)
#pragma unroll 8
        for (int j = g; j < nPw; j += G) ws += wp[(size_t)j * k + b];
    const double rr = reduce_partials(Prr, nPrr, red2);
    if (blockIdx.x == 0 && threadIdx.x == 0) cst(&ctl->rr, rr);
    if (rr < cld(&ctl->tol)) {
        if (blockIdx.x == 0 && threadIdx.x == 0) { cst(&ctl->done, 1); cst(&ctl->iters, it + 1); }
        return;
    }
    if (U) {
        // w = U^T r from the nPw block partials: thread (g, b) sums partials
        // g, g + G, ...; the G sums are added in order (the same in every block)
        if (b < k) cwp[g][b] = ws;
        __syncthreads();
        if (threadIdx.x < k) {
            double s = cwp[0][threadIdx.x];
#pragma unroll
            for (int gg = 1; gg < G; ++gg) s += cwp[gg][threadIdx.x];
            cw[threadIdx.x] = cf[threadIdx.x] * s;
        }
        __syncthreads();
    }
    if (i < n && U) {
        // wavefront wv's share of the spectral term: columns wv, wv + SLW, ...
        double t = 0.0;
#pragma unroll
        for (int q = 0; q < UB; ++q)
            if (wv + q * SLW < k) t += uv[q] * cw[wv + q * SLW];
        s += t;
    }
    red[wv][lane] = s;
    __syncthreads();
    double ss = 0.0;
    if (wv == 0 && i < n) {
        double zi = red[0][lane];
#pragma unroll
        for (int ww = 1; ww < SLW; ++ww) zi += red[ww][lane];
        z[i] = zi;
        ss = r[i] * zi;
    }
    const double t = block_sum(ss, red2);
    if (threadIdx.x == 0) P3[blockIdx.x] = t;
}

// all-landmark form: y = (K + eps I) x from the slice partials of the stored
// K product (summed in slice order), and with CHECK the kmatvec_k contract:
// stop when |r|^2 (partials Prr) is below the tolerance, else partials of x.y
template <int D, bool CHECK>
__global__ __launch_bounds__(NT) void kfin_k(int V, int slices, const double *__restrict__ part,
                                             const double *__restrict__ x, double *__restrict__ y, double eps,
                                             const double *__restrict__ Prr, int nPrr, double *__restrict__ P3,
                                             int it, Ctl *ctl) {
    __shared__ double red[NT / 64];
    if (ctl && cld(&ctl->done)) return;
    // the slice partials and x are requested before the stopping test
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double sv[D], xv[D];
    if (i < V) {
#pragma unroll
        for (int c = 0; c < D; ++c) {
            double s = 0.0;
#pragma unroll 8
            for (int k = 0; k < slices; ++k) s += part[(size_t)k * D * V + (size_t)c * V + i];
            sv[c] = s;
            xv[c] = x[c * V + i];
        }
    }
    if constexpr (CHECK) {
        double rr = reduce_partials(Prr, nPrr, red);
        if (rr < cld(&ctl->tol)) {
            if (blockIdx.x == 0 && threadIdx.x == 0) { cst(&ctl->done, 1); cst(&ctl->iters, it + 1); }
            return;
        }
    }
    double dotp = 0.0;
    if (i < V) {
#pragma unroll
        for (int c = 0; c < D; ++c) {
            const double s = sv[c] + eps * xv[c];
            y[c * V + i] = s;
            dotp += xv[c] * s;
        }
    }
    if constexpr (CHECK) {
        double t = block_sum(dotp, red);
        if (threadIdx.x == 0) P3[blockIdx.x] = t;
    }
}

// FP64 panel-major n x n matrix -> FP32 panel-major copy (its own panel height)
__global__ void panel_to_float_k(int n, const double *__restrict__ src, float *__restrict__ dst) {
    long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long)n * n) return;
    int j = (int)(idx / n), i = (int)(idx - (long)j * n);
    dst[panel_at(i, j, panel_rows<float>(), n)] = (float)src[panel_at(i, j, panel_rows<double>(), n)];
}

// column-major (leading dimension ld) n x n matrix -> panel-major copy
// src: column-major symmetric matrix of which only the lower triangle is read
__global__ void panelize_k(int n, const double *__restrict__ src, int ld, double *__restrict__ dst) {
    long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long)n * n) return;
    int j = (int)(idx / n), i = (int)(idx - (long)j * n);
    dst[panel_at(i, j, panel_rows<double>(), n)] = i >= j ? src[(size_t)j * ld + i] : src[(size_t)i * ld + j];
}

// y[i] = sum of the slice partials part[s * n + i]: 64 outputs per block,
// wavefront w of W adding slices w, w + W, ..., combined in wavefront order
// (W = SLW for short vectors; 4 when the rows alone give enough blocks)
template <int W>
__global__ __launch_bounds__(64 * W) void slice_rows_k(int n, int slices, const double *__restrict__ part,
                                                       double *__restrict__ y, const Ctl *ctl) {
    __shared__ double red[W][64];
    if (ctl && cld(&ctl->done)) return;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int i = blockIdx.x * 64 + lane;
    double s = 0.0;
    if (i < n)
#pragma unroll 4
        for (int k = wv; k < slices; k += W) s += part[(size_t)k * n + i];
    red[wv][lane] = s;
    __syncthreads();
    if (wv == 0 && i < n) {
        double t = red[0][lane];
#pragma unroll
        for (int w = 1; w < W; ++w) t += red[w][lane];
        y[i] = t;
    }
}

// ---------------------------------------------------------------------------
// Cholesky factorisation A = L L^T in place (lower triangle, column-major,
// leading dimension lda), right-looking in CHB x CHB blocks: for each block
// column k0, chol_diag_k factors the diagonal block in LDS, chol_panel_k
// solves the rows below it against that factor, chol_update_k subtracts the
// panel's outer product from the lower triangle of the trailing matrix.  A
// pivot that is not positive and finite stops the factorisation: info =
// its 1-based column (LAPACK's convention) and every later kernel returns at
// once.  The upper triangle is not read; the diagonal tiles of the trailing
// update leave it alone.  Fixed summation order: deterministic.
constexpr int CHB = 64;

// lane l's copy of x (a wavefront-uniform value, through scalar registers)
__device__ inline double readlane_d(double x, int l) {
    const long long b = __double_as_longlong(x);
    const int lo = __builtin_amdgcn_readlane((int)b, l);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// one wavefront: lane i keeps row i of the diagonal block in registers
// (r[l] = L[i][l]).  Step j reads the pivot and then column j straight from
// the lanes that hold them (readlane: no LDS round trips), and every lane
// updates its row: the same operations, in the same order, as a
// right-looking update of the whole block.  The update runs on every lane
// and every column after j; what it writes above a lane's diagonal is
// never read as a pivot or column value, and is never stored.
__global__ __launch_bounds__(CHB) void chol_diag_k(int n, int k0, double *__restrict__ A, int lda, int *info) {
    if (cld(info) != 0) return;
    const int nb = min(CHB, n - k0), i = threadIdx.x;
    double r[CHB];
#pragma unroll
    for (int l = 0; l < CHB; ++l) r[l] = (i < nb && l <= i) ? A[(size_t)(k0 + i) + (size_t)(k0 + l) * lda] : 0.0;
    bool ok = true;   // the same in every lane
#pragma unroll
    for (int j = 0; j < CHB; ++j) {
        if (j < nb && ok) {
            const double d = readlane_d(r[j], j);   // the same in every lane
            if (!(d > 0.0) || !isfinite(d)) {
                if (i == 0) cst(info, k0 + j + 1);
                ok = false;
            } else {
                const double p = sqrt(d);
                if (i == j) r[j] = p;
                else if (i > j) r[j] = r[j] / p;
#pragma unroll
                for (int l = j + 1; l < CHB; ++l) r[l] -= r[j] * readlane_d(r[j], l);
            }
        }
    }
    if (!ok) return;
#pragma unroll
    for (int l = 0; l < CHB; ++l)
        if (i < nb && l <= i) A[(size_t)(k0 + i) + (size_t)(k0 + l) * lda] = r[l];
}

// rows k0 + nb .. n - 1 of block column k0: a <- a L_kk^{-T} (forward
// substitution, one row per thread, the row in registers)
__global__ __launch_bounds__(256) void chol_panel_k(int n, int k0, double *__restrict__ A, int lda, const int *info) {
    __shared__ double L[CHB][CHB + 1];
    if (cld(info) != 0) return;
    const int nb = min(CHB, n - k0), t = threadIdx.x;
    for (int idx = t; idx < nb * nb; idx += blockDim.x) {
        const int i = idx % nb, j = idx / nb;
        L[i][j] = i >= j ? A[(size_t)(k0 + i) + (size_t)(k0 + j) * lda] : 0.0;
    }
    __syncthreads();
    const int r = k0 + nb + blockIdx.x * blockDim.x + t;
    if (r >= n) return;
    double x[CHB];
#pragma unroll
    for (int j = 0; j < CHB; ++j) x[j] = j < nb ? A[(size_t)r + (size_t)(k0 + j) * lda] : 0.0;
#pragma unroll
    for (int j = 0; j < CHB; ++j) {
        if (j < nb) {
            // four partial sums (shorter dependency chains), added in a fixed order
            double p[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int l = 0; l < j; ++l) p[l & 3] += x[l] * L[j][l];
            x[j] = (x[j] - ((p[0] + p[1]) + (p[2] + p[3]))) / L[j][j];
        }
    }
#pragma unroll
    for (int j = 0; j < CHB; ++j)
        if (j < nb) A[(size_t)r + (size_t)(k0 + j) * lda] = x[j];
}

// trailing update A[c0:, c0:] -= P P^T (lower triangle), c0 = k0 + nb, P the
// solved panel: one CHB x CHB tile (bi >= bj) per block, blockIdx.x the
// tile's index in the lower triangle of tiles, 256 threads of 4 x 4 outputs
__global__ __launch_bounds__(256) void chol_update_k(int n, int k0, double *__restrict__ A, int lda, const int *info) {
    __shared__ double Pi[CHB][CHB + 1], Pj[CHB][CHB + 1];   // [panel column][tile row]
    if (cld(info) != 0) return;
    const int nb = min(CHB, n - k0), c0 = k0 + nb, t = threadIdx.x;
    // tile (bi, bj), bi >= bj, from the triangular index
    const int b = blockIdx.x;
    int bi = (int)((sqrt(8.0 * b + 1.0) - 1.0) * 0.5);
    while ((bi + 1) * (bi + 2) / 2 <= b) ++bi;
    while (bi * (bi + 1) / 2 > b) --bi;
    const int bj = b - bi * (bi + 1) / 2;
    const int r0 = c0 + bi * CHB, s0 = c0 + bj * CHB;
    for (int idx = t; idx < CHB * nb; idx += blockDim.x) {
        const int rr = idx % CHB, l = idx / CHB;
        Pi[l][rr] = r0 + rr < n ? A[(size_t)(r0 + rr) + (size_t)(k0 + l) * lda] : 0.0;
        Pj[l][rr] = s0 + rr < n ? A[(size_t)(s0 + rr) + (size_t)(k0 + l) * lda] : 0.0;
    }
    __syncthreads();
    const int tr = (t % 16) * 4, tc = (t / 16) * 4;
    double acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int c = 0; c < 4; ++c) acc[a][c] = 0.0;
    for (int l = 0; l < nb; ++l) {
        double pa[4], pb[4];
#pragma unroll
        for (int a = 0; a < 4; ++a) { pa[a] = Pi[l][tr + a]; pb[a] = Pj[l][tc + a]; }
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int c = 0; c < 4; ++c) acc[a][c] += pa[a] * pb[c];
    }
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const int i = r0 + tr + a, j = s0 + tc + c;
            if (i < n && j < n && i >= j) A[(size_t)i + (size_t)j * lda] -= acc[a][c];
        }
}

// zero the strict upper triangle of a column-major C x C matrix
__global__ void zero_upper_k(int C, double *__restrict__ M) {
    long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long)C * C) return;
    int j = (int)(idx / C), i = (int)(idx - (long)j * C);
    if (i < j) M[idx] = 0.0;
}

// S d = K_cx B K_xc d + ldmWgt (K_cc d + eps d), with K_cc d given (kd) or,
// since the control points are landmarks, read from w = K_xc d (V rows) at
// the control-point rows; both null: K_cc d taken as 0.
// MODE 0: out = S d, partial d.(S d).  MODE 1: out = g - S a, partial |out|^2.
template <int D, int MODE>
__global__ __launch_bounds__(NT) void ctl_combine_k(int C, int V, const double *__restrict__ y,
                                                    const double *__restrict__ kdv, const double *__restrict__ w,
                                                    const int *__restrict__ idx, const double *__restrict__ d,
                                                    const double *__restrict__ g,
                                                    double ldmWgt, double eps, double *__restrict__ out,
                                                    double *__restrict__ P, const Ctl *ctl) {
    __shared__ double red[NT / 64];
    if (cld(&ctl->done)) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double ss = 0.0;
    if (i < C * D) {
        const int c = i / C;
        const double kd = kdv ? kdv[i] : w ? w[(size_t)c * V + idx[i - c * C]] : 0.0;
        double sd = y[i] + ldmWgt * (kd + eps * d[i]);
        if constexpr (MODE == 0) { out[i] = sd; ss = d[i] * sd; }
        else { double r = g[i] - sd; out[i] = r; ss = r * r; }
    }
    double t = block_sum(ss, red);
    if (threadIdx.x == 0) P[blockIdx.x] = t;
}

// spectral term of the preconditioner, first half: partials of w = U^T r
// (U: n x k).  Block (c, g) covers rows [c * SPW_ROWS, (c + 1) * SPW_ROWS)
// and columns [g * SPW_COLS, (g + 1) * SPW_COLS): wp[c * k + b].
constexpr int SPW_ROWS = 4 * NT, SPW_COLS = 4;
__global__ __launch_bounds__(NT) void spec_wp_k(int n, const double *__restrict__ U, int k,
                                                const double *__restrict__ r, double *__restrict__ wp,
                                                const Ctl *ctl) {
    __shared__ double red[NT / 64][SPW_COLS];
    if (cld(&ctl->done)) return;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int b0 = blockIdx.y * SPW_COLS;
    double acc[SPW_COLS];
#pragma unroll
    for (int c = 0; c < SPW_COLS; ++c) acc[c] = 0.0;
#pragma unroll
    for (int q = 0; q < SPW_ROWS / NT; ++q) {
        const int i = blockIdx.x * SPW_ROWS + q * NT + threadIdx.x;
        if (i < n) {
            const double ri = r[i];
#pragma unroll
            for (int c = 0; c < SPW_COLS; ++c)
                if (b0 + c < k) acc[c] += U[(size_t)(b0 + c) * n + i] * ri;
        }
    }
#pragma unroll
    for (int c = 0; c < SPW_COLS; ++c) {
        const double v = wave_sum(acc[c]);
        if (lane == 0) red[wv][c] = v;
    }
    __syncthreads();
    if (threadIdx.x < SPW_COLS && b0 + (int)threadIdx.x < k) {
        double s = red[0][threadIdx.x];
#pragma unroll
        for (int w = 1; w < NT / 64; ++w) s += red[w][threadIdx.x];
        wp[(size_t)blockIdx.x * k + b0 + threadIdx.x] = s;
    }
}

// a += st d; rout = rin - st Sd (st = rz / d.Sd); partial |rout|^2 (rin and
// rout may be one buffer when gridDim.y == 1).  Blocks with blockIdx.y = g
// > 0 only recompute rout for the spectral term: with U (n x k), block
// (x, g) writes the partials of w = U^T rout for columns 4g .. 4g + 3 to
// wp[x * k + b].
__global__ __launch_bounds__(NT) void ctl_update_k(int n, const double *__restrict__ P, int nP,
                                                   const double *__restrict__ hist, double *__restrict__ a,
                                                   const double *__restrict__ rin, double *__restrict__ rout,
                                                   const double *__restrict__ d, const double *__restrict__ Sd,
                                                   double *__restrict__ P2, const double *__restrict__ U, int k,
                                                   double *__restrict__ wp, Ctl *ctl) {
    __shared__ double red[NT / 64];
    __shared__ double redw[NT / 64][SPW_COLS];
    if (cld(&ctl->done)) return;
    const int it = cld(&ctl->it);
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int b0 = blockIdx.y * SPW_COLS;
    // the operands are requested before the partials are reduced
    const double rz = cld(&hist[it]);
    double rv = 0.0, sv = 0.0, av = 0.0, dv = 0.0, uv[SPW_COLS];
    if (i < n) {
        rv = rin[i];
        sv = Sd[i];
        if (blockIdx.y == 0) { av = a[i]; dv = d[i]; }
    }
#pragma unroll
    for (int c = 0; c < SPW_COLS; ++c) uv[c] = U && i < n && b0 + c < k ? U[(size_t)(b0 + c) * n + i] : 0.0;
    double dSd = reduce_partials(P, nP, red);
    if (dSd == 0.0) {
        if (blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0) {
            cst(&ctl->status, 1); cst(&ctl->done, 1); cst(&ctl->iters, it);
        }
        return;
    }
    const double st = rz / dSd;
    const double ri = i < n ? rv - st * sv : 0.0;
    if (blockIdx.y == 0) {
        if (i < n) {
            a[i] = av + st * dv;
            rout[i] = ri;
        }
        double t = block_sum(ri * ri, red);
        if (threadIdx.x == 0) P2[blockIdx.x] = t;
    }
    if (U) {
        const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
        for (int c = 0; c < SPW_COLS; ++c) {
            const double v = wave_sum(uv[c] * ri);
            if (lane == 0) redw[wv][c] = v;
        }
        __syncthreads();
        if (threadIdx.x < SPW_COLS && b0 + (int)threadIdx.x < k) {
            double s = redw[0][threadIdx.x];
#pragma unroll
            for (int w = 1; w < NT / 64; ++w) s += redw[w][threadIdx.x];
            wp[(size_t)blockIdx.x * k + b0 + threadIdx.x] = s;
        }
    }
}

// beta = rz_new / rz_old; d = z + beta d (it < 0: d = z)
__global__ __launch_bounds__(NT) void ctl_dir_k(int n, const double *__restrict__ P, int nP, double *__restrict__ hist,
                                                const double *__restrict__ z, double *__restrict__ d, Ctl *ctl) {
    __shared__ double red[NT / 64];
    if (cld(&ctl->done)) return;
    const int it = cld(&ctl->it);
    double rz = reduce_partials(P, nP, red);
    if (rz == 0.0) {
        if (blockIdx.x == 0 && threadIdx.x == 0) { cst(&ctl->status, 1); cst(&ctl->done, 1); cst(&ctl->iters, it + 1); }
        return;
    }
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (it < 0) {
        if (blockIdx.x == 0 && threadIdx.x == 0) cst(&hist[0], rz);
        if (i < n) d[i] = z[i];
        return;
    }
    double be = rz / cld(&hist[it]);
    if (blockIdx.x == 0 && threadIdx.x == 0) cst(&hist[it + 1], rz);
    if (i < n) d[i] = z[i] + be * d[i];
}

// control-point form, start of a step: shift the solution history
// (p3 <- p2 <- p1 <- a) and extrapolate a to the new step with the given
// order (polynomial through the last order + 1 solutions): 0 keeps a,
// 1: 2 a - p1, 2: 3 a - 3 p1 + p2, 3: 4 a - 6 p1 + 4 p2 - p3
__global__ void extrap_k(int n, int order, double *__restrict__ a, double *__restrict__ p1,
                         double *__restrict__ p2, double *__restrict__ p3) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double a0 = a[i], a1 = p1[i], a2 = p2[i], a3 = p3[i];
    if (order == 1) a[i] = 2.0 * a0 - a1;
    else if (order == 2) a[i] = 3.0 * (a0 - a1) + a2;
    else if (order >= 3) a[i] = 4.0 * (a0 + a2) - 6.0 * a1 - a3;
    p3[i] = a2;
    p2[i] = a1;
    p1[i] = a0;
}

__global__ void axpby_k(int n, double *__restrict__ out, double a, const double *__restrict__ x, double b,
                        const double *__restrict__ y) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = a * x[i] + b * y[i];
}

// ---------------------------------------------------------------------------
// reaction-diffusion: rhs = M_t tau_t + dt * (M_T r) gathered
// (assembleFemVector.cu + the Dgemv of deform.cu:95-97)

template <int D>
__global__ __launch_bounds__(NT) void femvec_elem_k(int E, const int *__restrict__ elmVtx, const double *__restrict__ geo,
                                                    const double *__restrict__ r, double *__restrict__ frc) {
    using G = GeoLayout<D>;
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= E) return;
    double ad = fabs(geo[G::DET * (size_t)E + e]);
    double rv[D + 1], s = 0.0;
#pragma unroll
    for (int k = 0; k <= D; ++k) { rv[k] = r[elmVtx[k * E + e]]; s += rv[k]; }
#pragma unroll
    for (int k = 0; k <= D; ++k) frc[(size_t)e * (D + 1) + k] = (s + rv[k]) * ad / (D == 2 ? 24.0 : 120.0);
}

// rhs_i = sum_z M(z) tau(col z) + dt * F_i, and the Jacobi-CG start:
// r = rhs - A x, z = r / diag, d = z; partials r.r, r.z, rhs.rhs
__global__ __launch_bounds__(NT) void rd_start_k(int V, int E, int NV1, const int *__restrict__ vtxElm,
                                                 const double *__restrict__ frc, const int *__restrict__ rowPtr,
                                                 const int *__restrict__ colIdx, const int *__restrict__ diagIdx,
                                                 const double *__restrict__ ppd, const double *__restrict__ lhs,
                                                 const double *__restrict__ tau, double dt, double *__restrict__ x,
                                                 double *__restrict__ rhs, double *__restrict__ r,
                                                 double *__restrict__ zv, double *__restrict__ d,
                                                 double *__restrict__ P) {
    __shared__ double red[NT / 64];
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    double s_rr = 0.0, s_rz = 0.0, s_bb = 0.0;
    if (i < V) {
        double F = 0.0;
        int adj = vtxElm[i];
        for (int a = 0; a < adj; ++a) {
            int e = vtxElm[(1 + 2 * a) * V + i], l = vtxElm[(2 + 2 * a) * V + i];
            F += frc[(size_t)e * NV1 + l];
        }
        double mt = 0.0, ax = 0.0;
        int z0 = rowPtr[i], z1 = rowPtr[i + 1];
        for (int z = z0; z < z1; ++z) {
            double tc = tau[colIdx[z]];
            mt += ppd[z] * tc;
            ax += lhs[z] * tc;   // warm start x = tau_t
        }
        double b = 1.0 * mt + dt * F;
        double xi = tau[i];
        x[i] = xi;
        rhs[i] = b;
        double ri = b - ax;
        double zi = ri / lhs[diagIdx[i]];
        r[i] = ri;
        zv[i] = zi;
        d[i] = zi;
        s_rr = ri * ri;
        s_rz = ri * zi;
        s_bb = b * b;
    }
    double a0 = block_sum(s_rr, red);
    double a1 = block_sum(s_rz, red);
    double a2 = block_sum(s_bb, red);
    if (threadIdx.x == 0) {
        P[blockIdx.x] = a0;
        P[MAXP + blockIdx.x] = a1;
        P[2 * MAXP + blockIdx.x] = a2;
    }
}

__global__ void rd_fin_k(const double *__restrict__ P, int nP, double tolSqu, double *__restrict__ hist, Ctl *ctl) {
    __shared__ double red[NT / 64];
    double rr = reduce_partials(P, nP, red);
    double rz = reduce_partials(P + MAXP, nP, red);
    double bb = reduce_partials(P + 2 * MAXP, nP, red);
    if (threadIdx.x == 0) {
        cst(&ctl->bb, bb);
        cst(&ctl->tol, bb * tolSqu);
        cst(&ctl->status, 0);
        cst(&ctl->iters, 0);
        cst(&ctl->done, (bb == 0.0 || rr < bb * tolSqu) ? 1 : 0);
        cst(&hist[0], rz);
    }
}

__global__ __launch_bounds__(NT) void rd_spmv_k(int V, const int *__restrict__ rowPtr, const int *__restrict__ colIdx,
                                                const double *__restrict__ A, const double *__restrict__ d,
                                                double *__restrict__ Ad, double *__restrict__ P, const Ctl *ctl) {
    __shared__ double red[NT / 64];
    if (cld(&ctl->done)) return;
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    double ss = 0.0;
    if (i < V) {
        double s = 0.0;
        for (int z = rowPtr[i]; z < rowPtr[i + 1]; ++z) s += A[z] * d[colIdx[z]];
        Ad[i] = s;
        ss = d[i] * s;
    }
    double t = block_sum(ss, red);
    if (threadIdx.x == 0) P[blockIdx.x] = t;
}

__global__ __launch_bounds__(NT) void rd_update_k(int V, const double *__restrict__ P, int nP,
                                                  const double *__restrict__ hist, int it, const int *__restrict__ diagIdx,
                                                  const double *__restrict__ A, double *__restrict__ x,
                                                  double *__restrict__ r, double *__restrict__ zv,
                                                  const double *__restrict__ d, const double *__restrict__ Ad,
                                                  double *__restrict__ P2, Ctl *ctl) {
    __shared__ double red[NT / 64];
    if (cld(&ctl->done)) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    // the operands are requested before the partials are reduced
    const double rz = cld(&hist[it]);
    double xv = 0.0, dv = 0.0, rv = 0.0, av = 0.0, dg = 1.0;
    if (i < V) { xv = x[i]; dv = d[i]; rv = r[i]; av = Ad[i]; dg = A[diagIdx[i]]; }
    double dAd = reduce_partials(P, nP, red);
    double al = rz / dAd;
    double s_rr = 0.0, s_rz = 0.0;
    if (i < V) {
        x[i] = xv + al * dv;
        double ri = rv - al * av;
        double zi = ri / dg;
        r[i] = ri;
        zv[i] = zi;
        s_rr = ri * ri;
        s_rz = ri * zi;
    }
    double a0 = block_sum(s_rr, red);
    double a1 = block_sum(s_rz, red);
    if (threadIdx.x == 0) { P2[blockIdx.x] = a0; P2[MAXP + blockIdx.x] = a1; }
}

// the direction update of iteration it (beta, d = z + beta d) and rd_spmv_k of
// iteration it + 1 in one launch:
// the stopping test and beta = rz / hist[it] from the partials of
// rd_update_k (P: r.r, P + MAXP: r.z); each row forms the new direction at
// its columns, zv + beta dOld (the same value every row that reads it
// computes), writes its own entry to dNew (another buffer than dOld, which
// other rows still read) and the product row with its d.Ad partial
__global__ __launch_bounds__(NT) void rd_spmv_dir_k(int V, const int *__restrict__ rowPtr,
                                                    const int *__restrict__ colIdx, const double *__restrict__ A,
                                                    const double *__restrict__ P, int nP, double *__restrict__ hist,
                                                    int it, const double *__restrict__ zv,
                                                    const double *__restrict__ dOld, double *__restrict__ dNew,
                                                    double *__restrict__ Ad, double *__restrict__ Pd, Ctl *ctl) {
    __shared__ double red[NT / 64];
    if (cld(&ctl->done)) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    // the row's bounds and own entries and the step history are requested
    // before the partials are reduced
    int z0 = 0, z1 = 0;
    double zi = 0.0, di0 = 0.0;
    if (i < V) { z0 = rowPtr[i]; z1 = rowPtr[i + 1]; zi = zv[i]; di0 = dOld[i]; }
    const double rzOld = cld(&hist[it]);
    double pr = 0.0, pz = 0.0;
    for (int j = threadIdx.x; j < nP; j += NT) { pr += P[j]; pz += P[MAXP + j]; }
    const double rr = block_sum(pr, red);
    if (rr < cld(&ctl->tol)) {
        if (blockIdx.x == 0 && threadIdx.x == 0) { cst(&ctl->done, 1); cst(&ctl->iters, it + 1); }
        return;
    }
    const double rz = block_sum(pz, red);
    const double be = rz / rzOld;
    if (blockIdx.x == 0 && threadIdx.x == 0) cst(&hist[it + 1], rz);
    double ss = 0.0;
    if (i < V) {
        double s = 0.0;
#pragma unroll 8
        for (int z = z0; z < z1; ++z) {
            const int c = colIdx[z];
            s += A[z] * (zv[c] + be * dOld[c]);
        }
        const double di = zi + be * di0;
        dNew[i] = di;
        Ad[i] = s;
        ss = di * s;
    }
    const double t = block_sum(ss, red);
    if (threadIdx.x == 0) Pd[blockIdx.x] = t;
}

// ---------------------------------------------------------------------------
// varifold (2D varifold.cu:9-357, 3D :9-370)

// bad (optional): facets whose vertex indices lie outside [0, V) set
// *bad = 1 and are given zero measure instead of being read (callers whose
// tables the host never checked)
template <int D>
__global__ void facets_k(int F, int V, const int *__restrict__ vfdVtx, const double *__restrict__ q,
                         double *__restrict__ cen, double *__restrict__ dir, double *__restrict__ vol,
                         int *__restrict__ bad = nullptr) {
    int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= F) return;
    if (bad) {
        bool ok = true;
#pragma unroll
        for (int k = 0; k < D; ++k) {
            const int v = vfdVtx[k * F + f];
            ok = ok && v >= 0 && v < V;
        }
        if (!ok) {
            cst(bad, 1);
#pragma unroll
            for (int c = 0; c < D; ++c) { cen[c * F + f] = 0.0; dir[c * F + f] = 0.0; }
            vol[f] = 0.0;
            return;
        }
    }
    if constexpr (D == 2) {
        int a = vfdVtx[f], b = vfdVtx[F + f];
        double d0 = q[b] - q[a], d1 = q[V + b] - q[V + a];
        cen[f] = 0.5 * (q[a] + q[b]);
        cen[F + f] = 0.5 * (q[V + a] + q[V + b]);
        double len = sqrt(d0 * d0 + d1 * d1);
        dir[f] = d0 / len; dir[F + f] = d1 / len;
        vol[f] = len;
    } else {
        int a = vfdVtx[f], b = vfdVtx[F + f], c = vfdVtx[2 * F + f];
        double u[3], w[3];
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            cen[i * F + f] = (q[i * V + a] + q[i * V + b] + q[i * V + c]) / 3.0;
            u[i] = q[i * V + b] - q[i * V + a];
            w[i] = q[i * V + c] - q[i * V + a];
        }
        double cr0 = u[1] * w[2] - u[2] * w[1], cr1 = -u[0] * w[2] + u[2] * w[0], cr2 = u[0] * w[1] - u[1] * w[0];
        double len = sqrt(cr0 * cr0 + cr1 * cr1 + cr2 * cr2);
        dir[f] = cr0 / len; dir[F + f] = cr1 / len; dir[2 * F + f] = cr2 / len;
        vol[f] = 0.5 * len;
    }
}

struct VfdPrm {
    int cenType, dirType;
    double cenInvW2, dirInvW2;
};

// Row sums of the varifold pair kernel.  Blocks [0, nbD) take rows of the
// evolved set (columns: evolved set, then target set with factor -2); blocks
// [nbD, nbD + nbT) take target rows against the target set.  GRAD also
// accumulates d/dcentre and d/d(unnormalised direction) of each evolved row.
template <int D, bool GRAD>
__global__ __launch_bounds__(NT) void vfd_rows_k(int F, int Ft, int nbD, const double *__restrict__ dc_,
                                                 const double *__restrict__ du_, const double *__restrict__ dw_,
                                                 const double *__restrict__ tc_, const double *__restrict__ tu_,
                                                 const double *__restrict__ tw_, VfdPrm vp,
                                                 double *__restrict__ rowVal, double *__restrict__ gC,
                                                 double *__restrict__ gN) {
    __shared__ double sc[D][NT], su[D][NT], sw[NT];
    __shared__ double racc[MV_PH][1 + 2 * D][MV_ROWS];
    const int lr = threadIdx.x & (MV_ROWS - 1), ph = threadIdx.x / MV_ROWS;
    const bool drow = blockIdx.x < nbD;
    const int i = (drow ? blockIdx.x : blockIdx.x - nbD) * MV_ROWS + lr;
    const int nrow = drow ? F : Ft;
    const double *rc = drow ? dc_ : tc_, *ru = drow ? du_ : tu_, *rw = drow ? dw_ : tw_;
    double ci[D], ui[D], wi = 0.0;
#pragma unroll
    for (int c = 0; c < D; ++c) { ci[c] = i < nrow ? rc[c * nrow + i] : 0.0; ui[c] = i < nrow ? ru[c * nrow + i] : 0.0; }
    if (i < nrow) wi = rw[i];
    const double nlen = (D == 2 ? wi : 2.0 * wi);   // |T| (2D) or |N| = 2 area (3D)
    const double half = (D == 2 ? 1.0 : 0.5);
    double val = 0.0, gc[D], gn[D];
#pragma unroll
    for (int c = 0; c < D; ++c) { gc[c] = 0.0; gn[c] = 0.0; }
    const int npass = drow ? 2 : 1;
    for (int pass = 0; pass < npass; ++pass) {
        const bool tgt = (!drow) || pass == 1;
        const double *pc = tgt ? tc_ : dc_, *pu = tgt ? tu_ : du_, *pw = tgt ? tw_ : dw_;
        const int n = tgt ? Ft : F;
        const double sg = (drow && pass == 1) ? -2.0 : 1.0;
        double pv = 0.0, pgc[D], pgn[D];
#pragma unroll
        for (int c = 0; c < D; ++c) { pgc[c] = 0.0; pgn[c] = 0.0; }
        for (int j0 = 0; j0 < n; j0 += NT) {
            int j = j0 + threadIdx.x;
            __syncthreads();
#pragma unroll
            for (int c = 0; c < D; ++c) {
                sc[c][threadIdx.x] = j < n ? pc[c * n + j] : 0.0;
                su[c][threadIdx.x] = j < n ? pu[c * n + j] : 0.0;
            }
            sw[threadIdx.x] = j < n ? pw[j] : 0.0;
            __syncthreads();
            int jn = min(NT, n - j0);
            for (int jj = ph; jj < jn; jj += MV_PH) {
                double d2 = 0.0, ang = 0.0, dv[D];
#pragma unroll
                for (int c = 0; c < D; ++c) { dv[c] = ci[c] - sc[c][jj]; d2 += dv[c] * dv[c]; ang += ui[c] * su[c][jj]; }
                double kc = cen_kernel(vp.cenType, d2, vp.cenInvW2);
                double ku = dir_kernel(vp.dirType, ang, vp.dirInvW2);
                double wj = sw[jj];
                pv += kc * ku * wj;
                if constexpr (GRAD) {
                    double kcd = cen_kernel_d(vp.cenType, kc, vp.cenInvW2);
                    double kud = dir_kernel_d(vp.dirType, ang, ku, vp.dirInvW2);
#pragma unroll
                    for (int c = 0; c < D; ++c) {
                        double dku = kud / nlen * (-ang * ui[c] + su[c][jj]);
                        pgc[c] += kcd * dv[c] * ku * wj;
                        pgn[c] += kc * (dku * wi + ku * half * ui[c]) * wj;
                    }
                }
            }
        }
        val += sg * pv * wi;
        if constexpr (GRAD) {
            // d/dc_i of the DD double sum and of -2 DT both carry factor 2
            const double gs = (drow && pass == 1) ? -2.0 : 2.0;
#pragma unroll
            for (int c = 0; c < D; ++c) { gc[c] += gs * pgc[c] * wi; gn[c] += gs * pgn[c]; }
        }
    }
    racc[ph][0][lr] = val;
    if constexpr (GRAD) {
#pragma unroll
        for (int c = 0; c < D; ++c) { racc[ph][1 + c][lr] = gc[c]; racc[ph][1 + D + c][lr] = gn[c]; }
    }
    __syncthreads();
    if (ph == 0 && i < nrow) {
        double s = racc[0][0][lr];
#pragma unroll
        for (int p = 1; p < MV_PH; ++p) s += racc[p][0][lr];
        rowVal[(drow ? 0 : F) + i] = s;
        if constexpr (GRAD) {
            if (drow) {
#pragma unroll
                for (int c = 0; c < D; ++c) {
                    double a = racc[0][1 + c][lr], b = racc[0][1 + D + c][lr];
#pragma unroll
                    for (int p = 1; p < MV_PH; ++p) { a += racc[p][1 + c][lr]; b += racc[p][1 + D + c][lr]; }
                    gC[c * F + i] = a;
                    gN[c * F + i] = b;
                }
            }
        }
    }
}

// Row sums of the varifold pair kernel (value only; the same rows, columns
// and signs as vfd_rows_k<D, false>) with both pair contractions on FP64
// matrix cores.  Per strip of 16 rows and tile of 16 columns, three v_mfma_f64_16x16x4f64 (K = 4: D coordinates, padded) give
//   d2[i][j]  = |c_i - o|^2 + |c_j - o|^2 - 2 (c_i - o).(c_j - o)
//             (K slots [c_i - o, 1] x [-2 (c_j - o), |c_j - o|^2], then
//             [|c_i - o|^2, 0..] x [1, 0..] into the same accumulator)
//   ang[i][j] = u_i . u_j
// and the VALU evaluates only the centre and direction kernels.  A block
// takes one strip of VFD_MFMA_ROWS = 16 rows (F / 16 blocks fill the chip
// where F / 64 did not); its four wavefronts take every fourth column tile
// and their row sums are added in wavefront order.  The origin o is the
// strip's first row centre: for the pairs the kernel weighs (|c_i -
// c_j| a few widths) the expanded form then cancels no more than the direct
// one (|c_j - o| <= |c_j - c_i| + the strip's extent).  MFMA operand layout
// (16x16x4 FP64): lane l holds A[l % 16][l / 16] and B[l / 16][l % 16]; the
// result D[4 r + l / 16][l % 16] is element r of the lane's 4-vector (rows
// interleaved by register: the FP64 shape has one-row groups).
typedef double vfd_d4 __attribute__((ext_vector_type(4)));
constexpr int VFD_MFMA_ROWS = 16;

template <int D>
__global__ __launch_bounds__(NT) void vfd_rows_mfma_k(int F, int Ft, int nbD, const double *__restrict__ dc_,
                                                      const double *__restrict__ du_, const double *__restrict__ dw_,
                                                      const double *__restrict__ tc_, const double *__restrict__ tu_,
                                                      const double *__restrict__ tw_, VfdPrm vp,
                                                      double *__restrict__ rowVal) {
    static_assert(D == 2 || D == 3, "K = 4 holds D coordinates and the constant 1");
    static_assert(NT == 4 * 64, "four wavefronts per block");
    __shared__ double sc[D][NT], su[D][NT], sw[NT];
    __shared__ double wsum[4][VFD_MFMA_ROWS];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int li = lane & 15, lk = lane >> 4;
    const bool drow = blockIdx.x < nbD;
    const int nrow = drow ? F : Ft;
    const int i0 = (drow ? blockIdx.x : blockIdx.x - nbD) * VFD_MFMA_ROWS;
    const double *rc = drow ? dc_ : tc_, *ru = drow ? du_ : tu_, *rw = drow ? dw_ : tw_;
    // origin: the strip's first row centre (the last row of a strip past the end)
    const int io = min(i0, nrow - 1);
    double o[D];
#pragma unroll
    for (int c = 0; c < D; ++c) o[c] = rc[c * nrow + io];
    // A operands: row i0 + li, K slot lk
    const int ia = i0 + li;
    double aC = 0.0, aU = 0.0, aN = 0.0;
    if (ia < nrow) {
        if (lk < D) {
            aC = rc[lk * nrow + ia] - o[lk];
            aU = ru[lk * nrow + ia];
        } else if (lk == D) {
            aC = 1.0;
        }
        if (lk == 0) {
#pragma unroll
            for (int c = 0; c < D; ++c) {
                const double t = rc[c * nrow + ia] - o[c];
                aN += t * t;
            }
        }
    }
    const double bN = lk == 0 ? 1.0 : 0.0;
    double val[4] = {0.0, 0.0, 0.0, 0.0};
    const int npass = drow ? 2 : 1;
    for (int pass = 0; pass < npass; ++pass) {
        const bool tgt = (!drow) || pass == 1;
        const double *pc = tgt ? tc_ : dc_, *pu = tgt ? tu_ : du_, *pw = tgt ? tw_ : dw_;
        const int n = tgt ? Ft : F;
        const double sg = (drow && pass == 1) ? -2.0 : 1.0;
        double pv[4] = {0.0, 0.0, 0.0, 0.0};
        for (int j0 = 0; j0 < n; j0 += NT) {
            const int j = j0 + threadIdx.x;
            __syncthreads();
#pragma unroll
            for (int c = 0; c < D; ++c) {
                sc[c][threadIdx.x] = j < n ? pc[c * n + j] : 0.0;
                su[c][threadIdx.x] = j < n ? pu[c * n + j] : 0.0;
            }
            sw[threadIdx.x] = j < n ? pw[j] : 0.0;   // columns past the end weigh 0
            __syncthreads();
            const int ntile = (min(NT, n - j0) + 15) >> 4;
            for (int t = wv; t < ntile; t += 4) {
                const int jj = t * 16 + li;   // this lane's column, as B operand and in the result
                double bC = 0.0, bU = 0.0;
                if (lk < D) {
                    bC = -2.0 * (sc[lk][jj] - o[lk]);
                    bU = su[lk][jj];
                } else if (lk == D) {
#pragma unroll
                    for (int c = 0; c < D; ++c) {
                        const double tj = sc[c][jj] - o[c];
                        bC += tj * tj;
                    }
                }
                const vfd_d4 z = {0.0, 0.0, 0.0, 0.0};
                vfd_d4 d2 = __builtin_amdgcn_mfma_f64_16x16x4f64(aC, bC, z, 0, 0, 0);
                d2 = __builtin_amdgcn_mfma_f64_16x16x4f64(aN, bN, d2, 0, 0, 0);
                const vfd_d4 an = __builtin_amdgcn_mfma_f64_16x16x4f64(aU, bU, z, 0, 0, 0);
                const double wj = sw[jj];
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const double kc = cen_kernel(vp.cenType, fmax(d2[r], 0.0), vp.cenInvW2);
                    const double ku = dir_kernel(vp.dirType, an[r], vp.dirInvW2);
                    pv[r] += kc * ku * wj;
                }
            }
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) val[r] += sg * pv[r];
    }
    // row 4 r + lk: sum over the 16 lanes of the group (columns l % 16), fixed order
#pragma unroll
    for (int r = 0; r < 4; ++r) {
#pragma unroll
        for (int m = 8; m > 0; m >>= 1) val[r] += __shfl_xor(val[r], m, 64);
    }
    if (li == 0) {
#pragma unroll
        for (int r = 0; r < 4; ++r) wsum[wv][4 * r + lk] = val[r];
    }
    __syncthreads();
    if (threadIdx.x < VFD_MFMA_ROWS) {
        const int i = i0 + threadIdx.x;
        if (i < nrow)
            rowVal[(drow ? 0 : F) + i] =
                (((wsum[0][threadIdx.x] + wsum[1][threadIdx.x]) + wsum[2][threadIdx.x]) + wsum[3][threadIdx.x]) * rw[i];
    }
}

// facet -> vertex gradient (dqVfdGatherKernel, varifold.cu:472-498|493-530)
template <int D>
__global__ void vfd_grad_k(int V, int F, const int *__restrict__ adjTbl, const int *__restrict__ vfdVtx,
                           const double *__restrict__ q, const double *__restrict__ gC, const double *__restrict__ gN,
                           double *__restrict__ grad) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= V) return;
    double g[D];
#pragma unroll
    for (int c = 0; c < D; ++c) g[c] = 0.0;
    int adj = adjTbl[i];
    for (int a = 0; a < adj; ++a) {
        int f = adjTbl[(1 + 2 * a) * V + i], l = adjTbl[(2 + 2 * a) * V + i];
        if constexpr (D == 2) {
            double sgn = (l == 0 ? -1.0 : 1.0);
            g[0] += 0.5 * gC[f] + sgn * gN[f];
            g[1] += 0.5 * gC[F + f] + sgn * gN[F + f];
        } else {
            int qa = vfdVtx[((l + 2) % 3) * F + f], qb = vfdVtx[((l + 1) % 3) * F + f];
            double dd0 = q[qa] - q[qb], dd1 = q[V + qa] - q[V + qb], dd2 = q[2 * V + qa] - q[2 * V + qb];
            double n0 = gN[f], n1 = gN[F + f], n2 = gN[2 * F + f];
            g[0] += gC[f] / 3.0 + (n1 * dd2 - n2 * dd1);
            g[1] += gC[F + f] / 3.0 + (-n0 * dd2 + n2 * dd0);
            g[2] += gC[2 * F + f] / 3.0 + (n0 * dd1 - n1 * dd0);
        }
    }
#pragma unroll
    for (int c = 0; c < D; ++c) grad[c * V + i] = g[c];
}

// objective = scal[0] + scal[1] (NaN when a facet index was out of range)
__global__ void vfd_total_k(const double *__restrict__ scal, const int *__restrict__ bad, double *__restrict__ obj) {
    if (threadIdx.x == 0) obj[0] = cld(bad) ? (double)NAN : scal[0] + scal[1];
}

// zero n ints (a fill kernel: usable under graph capture, unlike a memset node)
__global__ void zero_ints_k(int *__restrict__ p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) cst(&p[i], 0);
}

// single-block fixed-order sum of n values into out[0]
__global__ void sum_k(const double *__restrict__ v, int n, double *__restrict__ out) {
    __shared__ double red[NT / 64];
    double s = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) s += v[i];
    s = block_sum(s, red);
    if (threadIdx.x == 0) out[0] = s;
}

}  // namespace sa
