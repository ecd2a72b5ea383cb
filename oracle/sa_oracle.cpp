// FP64 CPU oracle for the forward shape-evolution step and its varifold
// objective.  TEST INFRASTRUCTURE ONLY: the parity tests, the smoke check and
// the CPU-baseline leg of the benchmark call it as the checker.  The product
// path (the HIP library) never links, loads or calls it.
//
// It restates, from the reference sources, the per-element arithmetic and the
// solver/driver steps of the path.  Each function names the reference file
// and lines it follows (paths relative to the reference repository,
// 2D = CUDA/2D_frame, 3D = CUDA/3D_frame, inc = CUDA/include).
//
// Two linear-solve modes exist so the oracle can serve both as the
// reference-exact check (small meshes) and as the same-algorithm check of the
// matrix-free GPU path (any size):
//   pcgInit = 0  K^{-1} x0 of the warm start through a dense Cholesky of K,
//                as ipcgWholeMatrixMultiplication.cu:32-33 does with potrs.
//   pcgInit = 1  K^{-1} x0 is carried from the previous step (alpha = K^{-1}v
//                is updated alongside x inside the CG), the warm start being
//                x0 = K_t alpha_{t-1}; no factorisation of K is needed.
//   rdSolve = 0  dense Cholesky of M+dt*K_D (deform.cu:99-104).
//   rdSolve = 1  Jacobi-preconditioned CG on the same sparse matrix.
#include <algorithm>
#include <cfloat>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

extern "C" {

struct OMesh {
    int dim, lmkNum, elmNum, nzrNum, btmElmNum, btmLmkNum, vfdElmNum;
    int vtxElmRows, vtxBtmRows, femIfoRows;
    const int *elmVtx, *vtxElm, *frmVtx, *tsvVtx, *btmVtx, *vtxBtm, *femVtx, *femIfo, *vfdVtx;
    const int *elmLyr;     // per-element material layer (isotropic mode), may be null
    int ctlNum;            // control points (0 = kernel over all landmarks)
    const int *ctlIdx;     // landmark index of each control point
    // boundary facets (bdrElmNum x dim) and their vertex adjacency (count-
    // prefixed, vtxBdrRows x lmkNum); bdrYank = 1 adds the boundary yank
    int bdrElmNum, vtxBdrRows;
    const int *bdrVtx, *vtxBdr;
    int bdrYank;
};

struct OParams {
    double ldmWgt; int knlOrder; double knlWidth, knlEps;
    double modVec[4]; int elaType; int nLyr; const double *modMat;   // modMat[2*l], [2*l+1] = (lambda, mu)
    double spdTan, spdTsv; double ynk[4]; double rea[5];
    double btmWgt, timeStp; int timeNum, cgItrMax; double cgTolSqu;
    int pcgInit, rdSolve; double rdTolSqu;
    int velPrec;   // 1: the control-point CG preconditioned by diag(S) (Jacobi), else (K_cc + eps I)^-1
};

struct OTarget {
    int tgtNum; const double *cen, *dir, *vol;
    int cenType, dirType; double cenWidth, dirWidth;
};

}  // extern "C"

namespace {

// ---------------------------------------------------------------------------
// small dense algebra (inc/matvec.h:159-280 for 2D, :495-679 for 3D)

template <int D> struct Geo {
    double inv[D][D];   // Q^{-1}, Q = [q1-q0, ..., qD-q0]; row k = grad(lambda_{k+1})
    double det;         // det Q
};

// adjugate / det, with the cofactor expressions of matInv (matvec.h:196-203,
// :640-662)
inline void mat_inv2(const double q1[2], const double q2[2], Geo<2> &g) {
    double det = q1[0] * q2[1] - q1[1] * q2[0];
    g.det = det;
    g.inv[0][0] = q2[1] / det;  g.inv[0][1] = -q2[0] / det;
    g.inv[1][0] = -q1[1] / det; g.inv[1][1] = q1[0] / det;
}

inline void mat_inv3(const double x[3], const double y[3], const double z[3], Geo<3> &g) {
    double det = x[0] * y[1] * z[2] + x[1] * y[2] * z[0] + x[2] * y[0] * z[1]
               - x[0] * y[2] * z[1] - x[1] * y[0] * z[2] - x[2] * y[1] * z[0];
    g.det = det;
    g.inv[0][0] =  (y[1] * z[2] - y[2] * z[1]) / det;
    g.inv[0][1] = -(y[0] * z[2] - y[2] * z[0]) / det;
    g.inv[0][2] =  (y[0] * z[1] - y[1] * z[0]) / det;
    g.inv[1][0] = -(x[1] * z[2] - x[2] * z[1]) / det;
    g.inv[1][1] =  (x[0] * z[2] - x[2] * z[0]) / det;
    g.inv[1][2] = -(x[0] * z[1] - x[1] * z[0]) / det;
    g.inv[2][0] =  (x[1] * y[2] - x[2] * y[1]) / det;
    g.inv[2][1] = -(x[0] * y[2] - x[2] * y[0]) / det;
    g.inv[2][2] =  (x[0] * y[1] - x[1] * y[0]) / det;
}

template <int D> inline double elem_vol(double det) {
    return D == 2 ? 0.5 * std::fabs(det) : std::fabs(det) / 6.0;   // utility.h:11-27
}

template <int D>
inline void elem_geo(const double *q, int V, const int *vtx, Geo<D> &g) {
    double e[D][D];
    for (int k = 0; k < D; ++k)
        for (int c = 0; c < D; ++c) e[k][c] = q[c * V + vtx[k + 1]] - q[c * V + vtx[0]];
    if constexpr (D == 2) mat_inv2(e[0], e[1], g);
    else mat_inv3(e[0], e[1], e[2], g);
}

template <int D> struct MeshView {
    const OMesh &m;
    int V, E;
    explicit MeshView(const OMesh &mm) : m(mm), V(mm.lmkNum), E(mm.elmNum) {}
    void vtx(int e, int *out) const { for (int k = 0; k <= D; ++k) out[k] = m.elmVtx[k * E + e]; }
};

// ---------------------------------------------------------------------------
// frames: 2D computeFrame.cu:4-68 (t, s); 3D computeFrame.cu:6-78 (n, s)

template <int D>
void compute_frames(const OMesh &m, const double *q, std::vector<double> &f1, std::vector<double> &f2) {
    int V = m.lmkNum, E = m.elmNum;
    f1.assign((size_t)D * E, 0.0);
    f2.assign((size_t)D * E, 0.0);
#pragma omp parallel for schedule(static)
    for (int e = 0; e < E; ++e) {
        double s1[3] = {0, 0, 0}, s2[3] = {0, 0, 0};
        if constexpr (D == 2) {
            for (int k = 0; k < 2; ++k) {
                int a = m.frmVtx[(2 * k) * E + e], b = m.frmVtx[(2 * k + 1) * E + e];
                double d0 = q[b] - q[a], d1 = q[V + b] - q[V + a];
                double len = std::sqrt(d0 * d0 + d1 * d1);
                s1[0] += d0 / len; s1[1] += d1 / len;
            }
        } else {
            for (int k = 0; k < 2; ++k) {
                int a = m.frmVtx[(3 * k) * E + e], b = m.frmVtx[(3 * k + 1) * E + e], c = m.frmVtx[(3 * k + 2) * E + e];
                double u[3], w[3];
                for (int i = 0; i < 3; ++i) { u[i] = q[i * V + b] - q[i * V + a]; w[i] = q[i * V + c] - q[i * V + a]; }
                double cr[3] = {u[1] * w[2] - u[2] * w[1], -u[0] * w[2] + u[2] * w[0], u[0] * w[1] - u[1] * w[0]};
                double len = std::sqrt(cr[0] * cr[0] + cr[1] * cr[1] + cr[2] * cr[2]);
                for (int i = 0; i < 3; ++i) s1[i] += cr[i] / len;
            }
        }
        const int ntsv = (D == 2 ? 2 : 3);
        for (int k = 0; k < ntsv; ++k) {
            int a = m.tsvVtx[(2 * k) * E + e], b = m.tsvVtx[(2 * k + 1) * E + e];
            double d[3] = {0, 0, 0}, len2 = 0;
            for (int i = 0; i < D; ++i) { d[i] = q[i * V + b] - q[i * V + a]; len2 += d[i] * d[i]; }
            double len = std::sqrt(len2);
            for (int i = 0; i < D; ++i) s2[i] += d[i] / len;
        }
        double l1 = 0, l2 = 0;
        for (int i = 0; i < D; ++i) { l1 += s1[i] * s1[i]; l2 += s2[i] * s2[i]; }
        l1 = std::sqrt(l1); l2 = std::sqrt(l2);
        for (int i = 0; i < D; ++i) { f1[i * E + e] = s1[i] / l1; f2[i * E + e] = s2[i] / l2; }
    }
}

// ---------------------------------------------------------------------------
// activity functions, inc/activity.h:4-38 (yank template), :81-90 (yank),
// :106-167 (reaction template), :239-249 (reaction)

double yank_template(double x, double dw) {
    double a = 3.0 * (dw - 2.0) / (dw * dw * dw * (6.0 + dw * (-6.0 + dw)));
    double b = (6.0 + dw * (-15.0 + 8 * dw)) / (3.0 * (dw - 2.0));
    double c = 6.0 / (6.0 + dw * (-6.0 + dw));
    if (x < -1.0 || x >= 1.0) return 0.0;
    if (x < -1.0 + dw) return a * (x + 1.0) * (x + 1.0) * (x + 1.0) * (x - b);
    if (x < 1.0 - dw) return -c * x * x + 1.0;
    return a * (-x + 1.0) * (-x + 1.0) * (-x + 1.0) * (-x - b);
}

double reac_template(double x, double dw, double a) {
    double b1 = 8.0 * (6.0 * (a - 1.0) + a * dw * (-6.0 + dw)) / (3.0 * dw * dw * dw * dw);
    double b2 = -(144.0 * (a - 1.0) + dw * (24.0 * (5.0 - 11.0 * a) + dw * (156.0 * a - dw * 29.0 * a)))
                / (24.0 * (6.0 * (a - 1.0) + a * dw * (-6.0 + dw)));
    double c1 = -(24.0 * (a - 1.0) + a * dw * (-12.0 + dw)) / (9.0 * dw * dw * dw);
    double c2 = -(24.0 * (a - 1.0) + dw * (12.0 * (2.0 - 3.0 * a) + a * dw * (13.0 + dw * 2.0))) / (3.0 * dw * dw * dw);
    double c3 = -((dw - 1.0) * (dw - 1.0) * (24.0 * (a - 1.0) + a * dw * (-12.0 + dw))) / (3.0 * dw * dw * dw);
    double c4 = -(24.0 * (a - 1.0) + dw * (12.0 * (6.0 - 7.0 * a)
                 + dw * (-72.0 + 109.0 * a + dw * (3.0 * (5.0 - 21.0 * a) + a * dw * (15.0 - dw)))))
                / (9.0 * dw * dw * dw);
    if (x < -1.0 || x >= 1.0) return 0.0;
    if (x < -1.0 + 0.5 * dw) return b1 * (x + 1.0) * (x + 1.0) * (x + 1.0) * (x - b2);
    if (x < -1.0 + dw) return c4 + x * (c3 + x * (c2 + x * c1));
    if (x < 1.0 - dw) return -a * x * x + 1.0;
    if (x < 1.0 - 0.5 * dw) return c4 + x * (-c3 + x * (c2 - x * c1));
    return b1 * (-x + 1.0) * (-x + 1.0) * (-x + 1.0) * (-x - b2);
}

inline double yank_activity(double tau, const double *p) {
    double x = -1.0 + 2.0 * (tau - p[0]) / (p[1] - p[0]);
    return p[2] * yank_template(x, p[3]);
}
inline double reac_activity(double tau, const double *p) {
    double x = -1.0 + 2.0 * (tau - p[0]) / (p[1] - p[0]);
    return p[2] * reac_template(x, p[3], p[4]);
}

// tauInitial.cu:4-22 (intpow), :24-64 (2D/3D kernels)
double intpow(double base, int ex) {
    double out = 1, pw = base;
    do {
        int quo = ex >> 1, rem = ex - 2 * quo;
        if (rem == 1) out *= pw;
        pw *= pw;
        ex = quo;
    } while (ex != 0);
    return out;
}

template <int D>
void tau_initial(const double *q, int V, const double *tauVar, const double *tauPrm, double *tau) {
    double c[3], h, R; int p;
    if constexpr (D == 2) { c[0] = tauVar[0]; c[1] = tauVar[1]; h = tauVar[2]; R = tauVar[3]; p = (int)tauPrm[0]; }
    else { c[0] = tauVar[0]; c[1] = tauVar[1]; c[2] = tauVar[2]; h = tauVar[3]; R = tauPrm[0]; p = (int)tauPrm[1]; }
    double R2 = R * R;
    for (int i = 0; i < V; ++i) {
        double s = 0;
        for (int k = 0; k < D; ++k) { double d = q[k * V + i] - c[k]; s += d * d; }
        double r2 = s / R2;
        double t = (r2 - 1.0) * (r2 - 1.0);
        double pre = (r2 < 1.0 ? 1.0 - intpow(1.0 - t, p) : 0.0);
        tau[i] = h * pre;
    }
}

// ---------------------------------------------------------------------------
// stress of the frame model: 2D applyFrameElasticOperator.cu:7-39,
// 3D applyFrameElasticOperator.cu:7-111.  eps is the symmetric strain.

template <int D>
void frame_stress(const double eps[D][D], const double *f1, const double *f2, const double *mod, double sig[D][D]) {
    if constexpr (D == 2) {
        const double *t = f1, *s = f2;
        double es[2] = {eps[0][0] * s[0] + eps[0][1] * s[1], eps[1][0] * s[0] + eps[1][1] * s[1]};
        double TvT = t[0] * (eps[0][0] * t[0] + eps[0][1] * t[1]) + t[1] * (eps[1][0] * t[0] + eps[1][1] * t[1]);
        double SvS = s[0] * es[0] + s[1] * es[1];
        double TvS = t[0] * es[0] + t[1] * es[1];
        for (int a = 0; a < 2; ++a)
            for (int b = 0; b < 2; ++b)
                sig[a][b] = mod[0] * (t[a] * t[b]) * TvT + mod[1] * (s[a] * s[b]) * SvS
                          + mod[2] * (t[a] * s[b] + s[a] * t[b]) * TvS;
    } else {
        const double *n = f1, *s = f2;
        double en[3], es[3];
        for (int a = 0; a < 3; ++a) {
            en[a] = eps[a][0] * n[0] + eps[a][1] * n[1] + eps[a][2] * n[2];
            es[a] = eps[a][0] * s[0] + eps[a][1] * s[1] + eps[a][2] * s[2];
        }
        double NvN = n[0] * en[0] + n[1] * en[1] + n[2] * en[2];
        double SvS = s[0] * es[0] + s[1] * es[1] + s[2] * es[2];
        double NvS = n[0] * es[0] + n[1] * es[1] + n[2] * es[2];
        double ldv = eps[0][0] + eps[1][1] + eps[2][2] - NvN;
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) {
                double dab = (a == b ? 1.0 : 0.0);
                sig[a][b] = mod[0] * (dab - n[a] * n[b]) * ldv
                          + mod[1] * (eps[a][b] - (en[a] * n[b] + n[a] * en[b]) + NvN * n[a] * n[b])
                          + mod[2] * SvS * s[a] * s[b]
                          + mod[3] * ((es[a] * s[b] + s[a] * es[b]) - NvS * (n[a] * s[b] + s[a] * n[b]));
            }
    }
}

struct Workspace {
    // per-step geometry at the current positions
    std::vector<double> f1, f2;    // frames
    std::vector<double> frc;       // (D+1)*D*E element force scratch
};

// strain of a velocity field on one element, computeStrainTensor.cu:4-53|4-61:
// eps = sym([v_k - v_0] Q^{-1})
template <int D>
inline void elem_strain(const double *v, int V, const int *vtx, const Geo<D> &g, double eps[D][D]) {
    double Vm[D][D];   // Vm[a][k] = (v_{k+1} - v_0)_a
    for (int k = 0; k < D; ++k)
        for (int a = 0; a < D; ++a) Vm[a][k] = v[a * V + vtx[k + 1]] - v[a * V + vtx[0]];
    double Dv[D][D];
    for (int a = 0; a < D; ++a)
        for (int b = 0; b < D; ++b) {
            double s = 0;
            for (int k = 0; k < D; ++k) s += Vm[a][k] * g.inv[k][b];
            Dv[a][b] = s;
        }
    for (int a = 0; a < D; ++a)
        for (int b = 0; b < D; ++b) eps[a][b] = (a == b ? Dv[a][a] : 0.5 * (Dv[a][b] + Dv[b][a]));
}

// element -> vertex gather in ascending element order through vtxElmMat,
// applyFrameElasticOperator.cu:87-111|166-190
template <int D>
void gather_vertex(const OMesh &m, const std::vector<double> &frc, double *out) {
    int V = m.lmkNum, E = m.elmNum;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < V; ++i) {
        double acc[3] = {0, 0, 0};
        int adj = m.vtxElm[i];
        for (int a = 0; a < adj; ++a) {
            int e = m.vtxElm[(1 + 2 * a) * V + i];
            int l = m.vtxElm[(2 + 2 * a) * V + i];
            for (int c = 0; c < D; ++c) acc[c] += frc[(size_t)(l * D + c) * E + e];
        }
        for (int c = 0; c < D; ++c) out[c * V + i] = acc[c];
    }
}

// elastic operator E*v.  elaType 0 = frame (reference), 1 = isotropic
// per-layer Lame moduli scaled by (1 - 0.5 * mean tau) (build-defined; it fills
// the empty 'isotropic' branch of CUDA/deformation.m:20-24).
template <int D>
void elastic_apply(const OMesh &m, const OParams &p, const double *q, const double *tau,
                   const Workspace &ws, std::vector<double> &frc, const double *v, double *out) {
    int V = m.lmkNum, E = m.elmNum;
    frc.assign((size_t)(D + 1) * D * E, 0.0);
#pragma omp parallel for schedule(static)
    for (int e = 0; e < E; ++e) {
        int vtx[D + 1];
        for (int k = 0; k <= D; ++k) vtx[k] = m.elmVtx[k * E + e];
        Geo<D> g; elem_geo<D>(q, V, vtx, g);
        double vol = elem_vol<D>(g.det);
        double eps[D][D]; elem_strain<D>(v, V, vtx, g, eps);
        double sig[D][D];
        if (p.elaType == 0) {
            double f1[D], f2[D];
            for (int c = 0; c < D; ++c) { f1[c] = ws.f1[c * E + e]; f2[c] = ws.f2[c * E + e]; }
            frame_stress<D>(eps, f1, f2, p.modVec, sig);
        } else {
            int l = m.elmLyr ? m.elmLyr[e] : 0;
            double ub = 0;
            for (int k = 0; k <= D; ++k) ub += tau[vtx[k]];
            ub /= (D + 1);
            double sc = 1.0 - 0.5 * ub;
            double lam = p.modMat[2 * l] * sc, mu = p.modMat[2 * l + 1] * sc;
            double tr = 0;
            for (int c = 0; c < D; ++c) tr += eps[c][c];
            for (int a = 0; a < D; ++a)
                for (int b = 0; b < D; ++b) sig[a][b] = (a == b ? lam * tr : 0.0) + 2.0 * mu * eps[a][b];
        }
        // nodal forces |T| * (Q^{-1} sigma), rows k >= 1; row 0 = -sum
        double dv[D][D];
        for (int k = 0; k < D; ++k)
            for (int c = 0; c < D; ++c) {
                double s = 0;
                for (int j = 0; j < D; ++j) s += g.inv[k][j] * sig[j][c];
                dv[k][c] = s;
            }
        for (int c = 0; c < D; ++c) {
            double s0 = 0;
            for (int k = 0; k < D; ++k) s0 -= dv[k][c];
            frc[(size_t)(0 * D + c) * E + e] = s0 * vol;
            for (int k = 0; k < D; ++k) frc[(size_t)((k + 1) * D + c) * E + e] = dv[k][c] * vol;
        }
    }
    gather_vertex<D>(m, frc, out);
}

// bottom penalty P*v: 2D applyPenaltyOperator.cu:4-79 (segment length / 2),
// 3D :4-85 (xy-projected triangle area / 3)
template <int D>
void penalty_apply(const OMesh &m, const double *q, const double *v, double *out) {
    int V = m.lmkNum, B = m.btmElmNum, BL = m.btmLmkNum;
    std::vector<double> pb((size_t)D * D * B);
    for (int b = 0; b < B; ++b) {
        int vtx[D];
        for (int k = 0; k < D; ++k) vtx[k] = m.btmVtx[k * B + b];
        double meas;
        if constexpr (D == 2) {
            double d0 = q[vtx[1]] - q[vtx[0]], d1 = q[V + vtx[1]] - q[V + vtx[0]];
            meas = std::sqrt(d0 * d0 + d1 * d1);
        } else {
            meas = 0.5 * std::fabs((q[vtx[1]] - q[vtx[0]]) * (q[V + vtx[2]] - q[V + vtx[0]])
                                 - (q[V + vtx[1]] - q[V + vtx[0]]) * (q[vtx[2]] - q[vtx[0]]));
        }
        for (int k = 0; k < D; ++k)
            for (int c = 0; c < D; ++c) pb[(size_t)(k * D + c) * B + b] = v[c * V + vtx[k]] * meas / (double)D;
    }
    for (size_t i = 0; i < (size_t)D * V; ++i) out[i] = 0.0;
    for (int i = 0; i < BL; ++i) {
        double acc[3] = {0, 0, 0};
        int adj = m.vtxBtm[i];
        for (int a = 0; a < adj; ++a) {
            int b = m.vtxBtm[(1 + 2 * a) * BL + i], l = m.vtxBtm[(2 + 2 * a) * BL + i];
            for (int c = 0; c < D; ++c) acc[c] += pb[(size_t)(l * D + c) * B + b];
        }
        for (int c = 0; c < D; ++c) out[c * V + i] = acc[c];
    }
}

// external yank f_k = -abar |T| grad(lambda_k), computeExternalYank.cu:5-47|5-56
// and its gather :86-141|104-159
template <int D>
void external_yank(const OMesh &m, const double *q, const double *act, std::vector<double> &frc, double *out) {
    int V = m.lmkNum, E = m.elmNum;
    frc.assign((size_t)(D + 1) * D * E, 0.0);
#pragma omp parallel for schedule(static)
    for (int e = 0; e < E; ++e) {
        int vtx[D + 1];
        for (int k = 0; k <= D; ++k) vtx[k] = m.elmVtx[k * E + e];
        Geo<D> g; elem_geo<D>(q, V, vtx, g);
        double vol = elem_vol<D>(g.det);
        double fc = 0;
        for (int k = 0; k <= D; ++k) fc += act[vtx[k]];
        fc /= (D + 1);
        for (int c = 0; c < D; ++c) {
            double s0 = 0;
            for (int k = 0; k < D; ++k) s0 -= g.inv[k][c];
            frc[(size_t)(0 * D + c) * E + e] = -fc * s0 * vol;
            for (int k = 0; k < D; ++k) frc[(size_t)((k + 1) * D + c) * E + e] = -fc * g.inv[k][c] * vol;
        }
    }
    gather_vertex<D>(m, frc, out);
    if (!m.bdrYank || !m.bdrVtx) return;
    // boundary yank, commented out in the reference (computeExternalYank.cu
    // :49-84 | 3D :58-99): facet normal n = (t_y, -t_x), t = q1 - q0 (2D) or
    // (q1 - q0) x (q2 - q0) (3D); vertex k of the facet gets act(q_k) n /
    // (VTXNUM - 1), VTXNUM - 1 = D; gathered after the element forces in
    // vtxBdrMat order (:106-116 | :125-135)
    const int Bn = m.bdrElmNum;
    std::vector<double> bf((size_t)D * D * Bn);
    for (int b = 0; b < Bn; ++b) {
        int v[D];
        for (int k = 0; k < D; ++k) v[k] = m.bdrVtx[k * Bn + b];
        double n[D];
        if (D == 2) {
            const double t0 = q[v[1]] - q[v[0]], t1 = q[V + v[1]] - q[V + v[0]];
            n[0] = t1;
            n[1] = -t0;
        } else {
            double u[3], w[3];
            for (int c = 0; c < 3; ++c) { u[c] = q[c * V + v[1]] - q[c * V + v[0]]; w[c] = q[c * V + v[D - 1]] - q[c * V + v[0]]; }
            n[0] = u[1] * w[2] - u[2] * w[1];
            n[1] = u[2] * w[0] - u[0] * w[2];
            n[D - 1] = u[0] * w[1] - u[1] * w[0];
        }
        for (int k = 0; k < D; ++k)
            for (int c = 0; c < D; ++c) bf[(size_t)(k * D + c) * Bn + b] = act[v[k]] * n[c] / (double)D;
    }
    for (int i = 0; i < V; ++i) {
        const int adj = m.vtxBdr[i];
        for (int a = 0; a < adj; ++a) {
            const int b = m.vtxBdr[(size_t)(1 + 2 * a) * V + i], l = m.vtxBdr[(size_t)(2 + 2 * a) * V + i];
            for (int c = 0; c < D; ++c) out[c * V + i] += bf[(size_t)(l * D + c) * Bn + b];
        }
    }
}

// ---------------------------------------------------------------------------
// kernels k(rho), rho = |q_i - q_j| / width
// 3D: computeKernel.cu:5-113 (Gaussian, Matern nu = m + 1/2 closed forms)
// 2D: computeKernel.cu:30-168 with f0 = rho^2 K0(rho), f1 = rho K1(rho)
//     (polybesselk.h:17-99).  K0/K1 are evaluated here from their power
//     series (Abramowitz & Stegun 9.6.10-9.6.13) for x <= 2 and from
//     Temme's continued fraction for K_nu (Steed's algorithm, "CF2") above.

const double kEuler = 0.57721566490153286061;

void bessel_k01(double x, double &k0, double &k1) {
    if (x <= 2.0) {
        double y = 0.25 * x * x, lg = std::log(0.5 * x);
        double t0 = 1.0, t1 = 1.0, i0 = 0, i1s = 0, s0 = 0, s1 = 0, H = 0;
        for (int k = 0; k < 40; ++k) {
            if (k > 0) { t0 *= y / ((double)k * k); t1 *= y / ((double)k * (k + 1)); H += 1.0 / k; }
            i0 += t0;
            i1s += t1;
            s0 += H * t0;
            double psi1 = -kEuler + H, psi2 = -kEuler + H + 1.0 / (k + 1);
            s1 += (psi1 + psi2) * t1;
            if (t0 < 1e-18 * i0 && t1 < 1e-18 * i1s) break;
        }
        k0 = -(lg + kEuler) * i0 + s0;
        double i1 = 0.5 * x * i1s;
        k1 = 1.0 / x + lg * i1 - 0.25 * x * s1;
        return;
    }
    // CF2 for nu = 0: K0 = sqrt(pi/2x) e^{-x} / s, K1 = K0 (x + 1/2 - h) / x
    double b = 2.0 * (1.0 + x), d = 1.0 / b, h = d, delh = d;
    double q1 = 0.0, q2 = 1.0, a1 = 0.25, q = a1, c = a1, a = -a1;
    double s = 1.0 + q * delh;
    for (int i = 2; i < 2000; ++i) {
        a -= 2 * (i - 1);
        c = -a * c / i;
        double qn = (q1 - b * q2) / a;
        q1 = q2;
        q2 = qn;
        q += c * qn;
        b += 2.0;
        d = 1.0 / (b + a * d);
        delh = (b * d - 1.0) * delh;
        h += delh;
        double dels = q * delh;
        s += dels;
        if (std::fabs(dels / s) < 1e-17) break;
    }
    h = a1 * h;
    k0 = std::sqrt(M_PI / (2.0 * x)) * std::exp(-x) / s;
    k1 = k0 * (x + 0.5 - h) / x;
}

inline void p01(double x, double &f0, double &f1) {
    const double xMax = 700.0;
    if (x < DBL_EPSILON) { f0 = 0.0; f1 = 1.0; return; }
    if (x > xMax) { f0 = 0.0; f1 = 0.0; return; }
    double k0, k1;
    bessel_k01(x, k0, k1);
    f0 = x * x * k0;
    f1 = x * k1;
}

double kernel_value(int dim, int order, double d2, double width) {
    if (order == -1) return std::exp(-d2 / (width * width));
    double r = std::sqrt(d2) / width;
    if (dim == 3) {
        double e = std::exp(-r);
        switch (order) {
            case 0: return e;
            case 1: return (1.0 + r) * e;
            case 2: return (3.0 + r * (3.0 + r)) / 3.0 * e;
            case 3: return (15.0 + r * (15.0 + r * (6.0 + r))) / 15.0 * e;
            case 4: return (105.0 + r * (105.0 + r * (45.0 + r * (10.0 + r)))) / 105.0 * e;
        }
        return 0.0;
    }
    double f0, f1;
    p01(r, f0, f1);
    double r2 = r * r;
    switch (order) {
        case 0: return f1;
        case 1: return 0.5 * (f0 + 2.0 * f1);
        case 2: return (4.0 * f0 + (8.0 + r2) * f1) / 8.0;
        case 3: return ((24.0 + r2) * f0 + 8.0 * (6.0 + r2) * f1) / 48.0;
        case 4: return (12.0 * (16.0 + r2) * f0 + (384.0 + r2 * (72.0 + r2)) * f1) / 384.0;
    }
    return 0.0;
}

// y = K x over point sets: y(i) = sum_j k(|a_i - b_j|) x(j) (+ eps x(i) when
// diag), columns of x/y stored SoA with strides na / nb
template <int D>
void kernel_matvec(int order, double width, double eps, const double *a, int na, const double *b, int nb,
                   const double *x, double *y, bool diag) {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < na; ++i) {
        double acc[3] = {0, 0, 0};
        for (int j = 0; j < nb; ++j) {
            double d2 = 0;
            for (int c = 0; c < D; ++c) { double d = a[c * na + i] - b[c * nb + j]; d2 += d * d; }
            double k = kernel_value(D, order, d2, width);
            if (diag && i == j) k += eps;
            for (int c = 0; c < D; ++c) acc[c] += k * x[c * nb + j];
        }
        for (int c = 0; c < D; ++c) y[c * na + i] = acc[c];
    }
}

// dense Cholesky (lower, column-major L(i,k) = A[k*n+i]) and solve: the
// role of cusolverDnDpotrf / potrs (cholesky.cu:4-17, deform.cu:99-104).
// Left-looking by panels of CHB columns: every entry subtracts its terms in
// ascending k, so the factor does not depend on the thread count.
constexpr int CHB = 64;
bool cholesky(std::vector<double> &A, int n) {
    double *a = A.data();
    for (int j0 = 0; j0 < n; j0 += CHB) {
        const int j1 = std::min(n, j0 + CHB);
        // panel update A(j0:n, j0:j1) -= A(j0:n, 0:j0) A(j0:j1, 0:j0)^T
        if (j0 > 0) {
            const int nch = (n - j0 + CHB - 1) / CHB;
#pragma omp parallel for schedule(dynamic)
            for (int ch = 0; ch < nch; ++ch) {
                const int i0 = j0 + ch * CHB, i1 = std::min(n, i0 + CHB);
                for (int k = 0; k < j0; ++k) {
                    const double *ck = a + (size_t)k * n;
                    for (int jj = j0; jj < j1; ++jj) {
                        const double ljk = ck[jj];
                        double *cj = a + (size_t)jj * n;
                        for (int i = std::max(i0, jj); i < i1; ++i) cj[i] -= ljk * ck[i];
                    }
                }
            }
        }
        // the panel itself, column by column
        for (int jj = j0; jj < j1; ++jj) {
            double *cj = a + (size_t)jj * n;
            const int nch = (n - jj + 255) / 256;
#pragma omp parallel for schedule(static) if (nch > 4)
            for (int ch = 0; ch < nch; ++ch) {
                const int i0 = jj + ch * 256, i1 = std::min(n, i0 + 256);
                for (int k = j0; k < jj; ++k) {
                    const double *ck = a + (size_t)k * n;
                    const double ljk = ck[jj];
                    for (int i = i0; i < i1; ++i) cj[i] -= ljk * ck[i];
                }
            }
            if (!(cj[jj] > 0)) return false;
            const double l = std::sqrt(cj[jj]);
            cj[jj] = l;
            for (int i = jj + 1; i < n; ++i) cj[i] /= l;
        }
    }
    return true;
}

// L L^T x = b for m right-hand sides (b: n x m, column-major), by panels of
// CHB; the per-entry summation order is fixed (ascending k)
void chol_solve(const std::vector<double> &L, int n, double *b, int m = 1) {
    const double *l = L.data();
    for (int j0 = 0; j0 < n; j0 += CHB) {
        const int j1 = std::min(n, j0 + CHB);
        for (int c = 0; c < m; ++c) {
            double *bc = b + (size_t)c * n;
            for (int k = j0; k < j1; ++k) {
                const double *ck = l + (size_t)k * n;
                bc[k] /= ck[k];
                for (int i = k + 1; i < j1; ++i) bc[i] -= ck[i] * bc[k];
            }
        }
        const int nch = (n - j1 + 255) / 256;
#pragma omp parallel for schedule(static) if (nch > 4)
        for (int ch = 0; ch < nch; ++ch) {
            const int i0 = j1 + ch * 256, i1 = std::min(n, i0 + 256);
            for (int c = 0; c < m; ++c) {
                double *bc = b + (size_t)c * n;
                for (int k = j0; k < j1; ++k) {
                    const double *ck = l + (size_t)k * n;
                    const double bk = bc[k];
                    for (int i = i0; i < i1; ++i) bc[i] -= ck[i] * bk;
                }
            }
        }
    }
    for (int j1 = n; j1 > 0; j1 -= CHB) {
        const int j0 = std::max(0, j1 - CHB);
#pragma omp parallel for schedule(static) if (n - j1 > 1024)
        for (int k = j0; k < j1; ++k) {
            const double *ck = l + (size_t)k * n;
            for (int c = 0; c < m; ++c) {
                double *bc = b + (size_t)c * n;
                double s = 0.0;
                for (int i = j1; i < n; ++i) s += ck[i] * bc[i];
                bc[k] -= s;
            }
        }
        for (int c = 0; c < m; ++c) {
            double *bc = b + (size_t)c * n;
            for (int k = j1 - 1; k >= j0; --k) {
                const double *ck = l + (size_t)k * n;
                double s = bc[k];
                for (int i = k + 1; i < j1; ++i) s -= ck[i] * bc[i];
                bc[k] = s / ck[k];
            }
        }
    }
}

double dot(const double *a, const double *b, size_t n) {
    double s = 0;
    for (size_t i = 0; i < n; ++i) s += a[i] * b[i];
    return s;
}

// ---------------------------------------------------------------------------
// FEM mass / anisotropic stiffness per element and the per-nonzero gather:
// 2D assembleFemMatrix.cu:5-110, 3D :5-169.  Values land in the sorted
// (row, col) nonzero order of femVtxMat (a CSR order).
template <int D>
void fem_assemble(const OMesh &m, const OParams &p, const double *q, const std::vector<double> &f1,
                  const std::vector<double> &f2, std::vector<double> &ppd, std::vector<double> &ggd) {
    int V = m.lmkNum, E = m.elmNum, Z = m.nzrNum;
    const int NS = (D + 1) * (D + 2) / 2;
    std::vector<double> pe(E), ge((size_t)NS * E);
#pragma omp parallel for schedule(static)
    for (int e = 0; e < E; ++e) {
        int vtx[D + 1];
        for (int k = 0; k <= D; ++k) vtx[k] = m.elmVtx[k * E + e];
        Geo<D> g; elem_geo<D>(q, V, vtx, g);
        double ad = std::fabs(g.det);
        double grad[D + 1][D];   // grad lambda_k
        for (int c = 0; c < D; ++c) {
            double s = 0;
            for (int k = 0; k < D; ++k) { grad[k + 1][c] = g.inv[k][c]; s -= g.inv[k][c]; }
            grad[0][c] = s;
        }
        double a1[D + 1], a2[D + 1];
        for (int k = 0; k <= D; ++k) {
            double s1 = 0, s2 = 0;
            for (int c = 0; c < D; ++c) { s1 += grad[k][c] * f1[c * E + e]; s2 += grad[k][c] * f2[c * E + e]; }
            a1[k] = s1; a2[k] = s2;
        }
        pe[e] = ad / (D == 2 ? 24.0 : 120.0);
        int sl = 0;
        for (int a = 0; a <= D; ++a)
            for (int b = a; b <= D; ++b, ++sl) {
                double v;
                if constexpr (D == 2) {
                    v = 0.5 * ad * (p.spdTan * a1[a] * a1[b] + p.spdTsv * a2[a] * a2[b]);
                } else {
                    double gg = 0;
                    for (int c = 0; c < 3; ++c) gg += grad[a][c] * grad[b][c];
                    v = ad / 6.0 * (p.spdTan * (gg - a1[a] * a1[b]) + p.spdTsv * a2[a] * a2[b]);
                }
                ge[(size_t)sl * E + e] = v;
            }
    }
    ppd.assign(Z, 0.0);
    ggd.assign(Z, 0.0);
#pragma omp parallel for schedule(static)
    for (int z = 0; z < Z; ++z) {
        int adj = m.femIfo[z];
        double mul = m.femIfo[Z + z];
        double pv = 0, gv = 0;
        for (int a = 0; a < adj; ++a) {
            int e = m.femIfo[(2 + 2 * a) * Z + z];
            int gi = m.femIfo[(3 + 2 * a) * Z + z];
            pv += mul * pe[e];
            gv += ge[gi];
        }
        ppd[z] = pv;
        ggd[z] = gv;
    }
}

// element reaction load M_T r, assembleFemVector.cu:5-67|5-70
template <int D>
void fem_vector(const OMesh &m, const double *q, const double *r, double *out) {
    int V = m.lmkNum, E = m.elmNum;
    std::vector<double> frc((size_t)(D + 1) * E);
#pragma omp parallel for schedule(static)
    for (int e = 0; e < E; ++e) {
        int vtx[D + 1];
        for (int k = 0; k <= D; ++k) vtx[k] = m.elmVtx[k * E + e];
        Geo<D> g; elem_geo<D>(q, V, vtx, g);
        double ad = std::fabs(g.det);
        double s = 0;
        for (int k = 0; k <= D; ++k) s += r[vtx[k]];
        for (int k = 0; k <= D; ++k) frc[(size_t)k * E + e] = (s + r[vtx[k]]) * ad / (D == 2 ? 24.0 : 120.0);
    }
#pragma omp parallel for schedule(static)
    for (int i = 0; i < V; ++i) {
        double acc = 0;
        int adj = m.vtxElm[i];
        for (int a = 0; a < adj; ++a) {
            int e = m.vtxElm[(1 + 2 * a) * V + i], l = m.vtxElm[(2 + 2 * a) * V + i];
            acc += frc[(size_t)l * E + e];
        }
        out[i] = acc;
    }
}

struct Csr {
    std::vector<int> ptr, col;
};

void build_csr(const OMesh &m, Csr &c) {
    int V = m.lmkNum, Z = m.nzrNum;
    c.ptr.assign(V + 1, 0);
    c.col.resize(Z);
    for (int z = 0; z < Z; ++z) { c.ptr[m.femVtx[z] + 1]++; c.col[z] = m.femVtx[Z + z]; }
    for (int i = 0; i < V; ++i) c.ptr[i + 1] += c.ptr[i];
}

void csr_mv(const Csr &c, const double *a, const double *x, double *y, int V) {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < V; ++i) {
        double s = 0;
        for (int z = c.ptr[i]; z < c.ptr[i + 1]; ++z) s += a[z] * x[c.col[z]];
        y[i] = s;
    }
}

// reaction-diffusion step (M+ dt K_D+) tau+ = M tau + dt F, deform.cu:89-104
int rd_solve(const OMesh &m, const OParams &p, const Csr &c, const std::vector<double> &lhs, const double *rhs,
             double *x, int *iters) {
    int V = m.lmkNum;
    if (p.rdSolve == 0) {
        std::vector<double> A((size_t)V * V, 0.0);
        for (int i = 0; i < V; ++i)
            for (int z = c.ptr[i]; z < c.ptr[i + 1]; ++z) A[(size_t)c.col[z] * V + i] = lhs[z];
        if (!cholesky(A, V)) return 2;
        std::memcpy(x, rhs, sizeof(double) * V);
        chol_solve(A, V, x);
        if (iters) *iters = 0;
        return 0;
    }
    // Jacobi-PCG, warm start from the incoming x
    std::vector<double> dg(V), r(V), z(V), d(V), Ad(V);
    for (int i = 0; i < V; ++i)
        for (int k = c.ptr[i]; k < c.ptr[i + 1]; ++k)
            if (c.col[k] == i) dg[i] = lhs[k];
    double bb = dot(rhs, rhs, V);
    csr_mv(c, lhs.data(), x, Ad.data(), V);
    for (int i = 0; i < V; ++i) r[i] = rhs[i] - Ad[i];
    double rr = dot(r.data(), r.data(), V);
    int it = 0;
    if (bb == 0.0) { std::fill(x, x + V, 0.0); if (iters) *iters = 0; return 0; }
    for (int i = 0; i < V; ++i) { z[i] = r[i] / dg[i]; d[i] = z[i]; }
    double rz = dot(r.data(), z.data(), V);
    while (rr >= p.rdTolSqu * bb && it < 100000) {
        csr_mv(c, lhs.data(), d.data(), Ad.data(), V);
        double al = rz / dot(d.data(), Ad.data(), V);
        for (int i = 0; i < V; ++i) { x[i] += al * d[i]; r[i] -= al * Ad[i]; }
        rr = dot(r.data(), r.data(), V);
        for (int i = 0; i < V; ++i) z[i] = r[i] / dg[i];
        double rz2 = dot(r.data(), z.data(), V);
        double be = rz2 / rz;
        rz = rz2;
        for (int i = 0; i < V; ++i) d[i] = z[i] + be * d[i];
        ++it;
    }
    if (iters) *iters = it;
    return 0;
}

// ---------------------------------------------------------------------------
// varifold: landmarksToVarifoldKernel (2D varifold.cu:9-36, 3D :9-43),
// pair kernels (:38-179 | :45-192), sums (:181-357 | :194-370), gradient
// (:359-498 | :372-530)

template <int D>
void facets(const OMesh &m, const double *q, std::vector<double> &cen, std::vector<double> &dir, std::vector<double> &vol) {
    int V = m.lmkNum, F = m.vfdElmNum;
    cen.resize((size_t)D * F); dir.resize((size_t)D * F); vol.resize(F);
    for (int f = 0; f < F; ++f) {
        if constexpr (D == 2) {
            int a = m.vfdVtx[f], b = m.vfdVtx[F + f];
            double d0 = q[b] - q[a], d1 = q[V + b] - q[V + a];
            cen[f] = 0.5 * (q[a] + q[b]); cen[F + f] = 0.5 * (q[V + a] + q[V + b]);
            double len = std::sqrt(d0 * d0 + d1 * d1);
            dir[f] = d0 / len; dir[F + f] = d1 / len; vol[f] = len;
        } else {
            int a = m.vfdVtx[f], b = m.vfdVtx[F + f], c = m.vfdVtx[2 * F + f];
            double u[3], w[3];
            for (int i = 0; i < 3; ++i) {
                cen[i * F + f] = (q[i * V + a] + q[i * V + b] + q[i * V + c]) / 3.0;
                u[i] = q[i * V + b] - q[i * V + a]; w[i] = q[i * V + c] - q[i * V + a];
            }
            double cr[3] = {u[1] * w[2] - u[2] * w[1], -u[0] * w[2] + u[2] * w[0], u[0] * w[1] - u[1] * w[0]};
            double len = std::sqrt(cr[0] * cr[0] + cr[1] * cr[1] + cr[2] * cr[2]);
            for (int i = 0; i < 3; ++i) dir[i * F + f] = cr[i] / len;
            vol[f] = 0.5 * len;
        }
    }
}

inline double cen_kernel(int type, double d2, double w) {
    if (type == 'G') return std::exp(-d2 / (w * w));
    if (type == 'C') return 1.0 / (1.0 + d2 / (w * w));
    return 0.0;
}
inline double cen_kernel_d(int type, double k, double w) {   // dk/d(d2) * 2 ... returns factor of (c1-c2)
    if (type == 'G') return -2.0 * k / (w * w);
    if (type == 'C') return -2.0 * k * k / (w * w);
    return 0.0;
}
inline double dir_kernel(int type, double ang, double w) {
    if (type == 'B') return ang * ang;
    if (type == 'L') return ang;
    if (type == 'O') return std::exp(2.0 * (ang - 1.0) / (w * w));
    if (type == 'U') return std::exp(2.0 * (ang * ang - 1.0) / (w * w));
    return 0.0;
}
inline double dir_kernel_d(int type, double ang, double k, double w) {   // dk/d(ang)
    if (type == 'B') return 2.0 * ang;
    if (type == 'L') return 1.0;
    if (type == 'O') return 2.0 * k / (w * w);
    if (type == 'U') return 4.0 * ang * k / (w * w);
    return 0.0;
}

template <int D>
double pair_sum_row(const double *ci, const double *ui, double wi, const double *cen, const double *dir,
                    const double *vol, int n, const OTarget &t) {
    double s = 0;
    for (int j = 0; j < n; ++j) {
        double d2 = 0, ang = 0;
        for (int c = 0; c < D; ++c) { double d = ci[c] - cen[c * n + j]; d2 += d * d; ang += ui[c] * dir[c * n + j]; }
        s += cen_kernel(t.cenType, d2, t.cenWidth) * dir_kernel(t.dirType, ang, t.dirWidth) * wi * vol[j];
    }
    return s;
}

template <int D>
double varifold_value(const std::vector<double> &cen, const std::vector<double> &dir, const std::vector<double> &vol,
                      int F, const OTarget &t) {
    int Ft = t.tgtNum;
    std::vector<double> rows((size_t)F + Ft);
#pragma omp parallel for schedule(dynamic, 16)
    for (int i = 0; i < F; ++i) {
        double ci[3], ui[3];
        for (int c = 0; c < D; ++c) { ci[c] = cen[c * F + i]; ui[c] = dir[c * F + i]; }
        rows[i] = pair_sum_row<D>(ci, ui, vol[i], cen.data(), dir.data(), vol.data(), F, t)
                - 2.0 * pair_sum_row<D>(ci, ui, vol[i], t.cen, t.dir, t.vol, Ft, t);
    }
#pragma omp parallel for schedule(dynamic, 16)
    for (int i = 0; i < Ft; ++i) {
        double ci[3], ui[3];
        for (int c = 0; c < D; ++c) { ci[c] = t.cen[c * Ft + i]; ui[c] = t.dir[c * Ft + i]; }
        rows[F + i] = pair_sum_row<D>(ci, ui, t.vol[i], t.cen, t.dir, t.vol, Ft, t);
    }
    double s = 0;
    for (double r : rows) s += r;
    return s;
}

// gradient of the varifold distance with respect to the landmark positions
// of the varifold facets (dqVfd_DD_DT_Kernel + dqVfdGatherKernel); grad is
// V x D SoA, zero off the facets
template <int D>
void varifold_grad(const OMesh &m, const double *q, const std::vector<double> &cen, const std::vector<double> &dir,
                   const std::vector<double> &vol, const OTarget &t, double *grad) {
    int F = m.vfdElmNum, Ft = t.tgtNum, V = m.lmkNum;
    std::vector<double> dc((size_t)D * F), dn((size_t)D * F);
#pragma omp parallel for schedule(dynamic, 16)
    for (int i = 0; i < F; ++i) {
        double ci[3], ui[3], wi = vol[i];
        for (int c = 0; c < D; ++c) { ci[c] = cen[c * F + i]; ui[c] = dir[c * F + i]; }
        double nlen = (D == 2 ? wi : 2.0 * wi);      // |T| in 2D, |N| = 2 area in 3D
        double half = (D == 2 ? 1.0 : 0.5);          // dw/dT = u (2D), dw/dN = u/2 (3D)
        double gc[3] = {0, 0, 0}, gn[3] = {0, 0, 0};
        for (int pass = 0; pass < 2; ++pass) {
            const double *pc = pass == 0 ? cen.data() : t.cen;
            const double *pu = pass == 0 ? dir.data() : t.dir;
            const double *pw = pass == 0 ? vol.data() : t.vol;
            int n = pass == 0 ? F : Ft;
            double sg = pass == 0 ? 1.0 : -1.0;
            for (int j = 0; j < n; ++j) {
                double d2 = 0, ang = 0, dv[3];
                for (int c = 0; c < D; ++c) { dv[c] = ci[c] - pc[c * n + j]; d2 += dv[c] * dv[c]; ang += ui[c] * pu[c * n + j]; }
                double kc = cen_kernel(t.cenType, d2, t.cenWidth);
                double kcd = cen_kernel_d(t.cenType, kc, t.cenWidth);
                double ku = dir_kernel(t.dirType, ang, t.dirWidth);
                double kud = dir_kernel_d(t.dirType, ang, ku, t.dirWidth);
                for (int c = 0; c < D; ++c) {
                    double dku = kud / nlen * (-ang * ui[c] + pu[c * n + j]);
                    gc[c] += sg * 2.0 * (kcd * dv[c]) * ku * wi * pw[j];
                    gn[c] += sg * 2.0 * kc * (dku * wi + ku * half * ui[c]) * pw[j];
                }
            }
        }
        for (int c = 0; c < D; ++c) { dc[c * F + i] = gc[c]; dn[c * F + i] = gn[c]; }
    }
    for (size_t i = 0; i < (size_t)D * V; ++i) grad[i] = 0.0;
    // facet -> vertex, ascending facet order per vertex
    for (int f = 0; f < F; ++f) {
        if constexpr (D == 2) {
            int a = m.vfdVtx[f], b = m.vfdVtx[F + f];
            for (int c = 0; c < 2; ++c) {
                grad[c * V + a] += 0.5 * dc[c * F + f] - dn[c * F + f];
                grad[c * V + b] += 0.5 * dc[c * F + f] + dn[c * F + f];
            }
        } else {
            int v3[3] = {m.vfdVtx[f], m.vfdVtx[F + f], m.vfdVtx[2 * F + f]};
            double g[3] = {dn[f], dn[F + f], dn[2 * F + f]};
            for (int k = 0; k < 3; ++k) {
                int qa = v3[(k + 2) % 3], qb = v3[(k + 1) % 3];
                double d[3];
                for (int c = 0; c < 3; ++c) d[c] = q[c * V + qa] - q[c * V + qb];
                double cr[3] = {g[1] * d[2] - g[2] * d[1], -g[0] * d[2] + g[2] * d[0], g[0] * d[1] - g[1] * d[0]};
                for (int c = 0; c < 3; ++c) grad[c * V + v3[k]] += dc[c * F + f] / 3.0 + cr[c];
            }
        }
    }
}

// ---------------------------------------------------------------------------
// the velocity solve, ipcg.cu:11-100: CG on A = ldmWgt K^{-1} + E + btmWgt P
// preconditioned by K.  K applied matrix-free; with control points the
// velocity is v = K_xc a and the CG runs on (ldmWgt K_cc + K_cx B K_xc) a = K_cx f
// preconditioned by K_cc^{-1}... (control points: see run()).

template <int D>
struct Solver {
    const OMesh &m;
    const OParams &p;
    Workspace ws;
    std::vector<double> frc, tmp;
    int V;
    Solver(const OMesh &mm, const OParams &pp) : m(mm), p(pp), V(mm.lmkNum) {}

    void B_apply(const double *q, const double *tau, const double *x, double *out) {
        elastic_apply<D>(m, p, q, tau, ws, frc, x, out);
        if (p.btmWgt != 0.0) {
            tmp.resize((size_t)D * V);
            penalty_apply<D>(m, q, x, tmp.data());
            for (size_t i = 0; i < (size_t)D * V; ++i) out[i] += p.btmWgt * tmp[i];
        } else {
            // reference still computes btmWgt * P x; with btmWgt = 0 the sum
            // adds 0 * finite = 0 exactly
        }
    }
    // dense copy of K for the current step when it is small (the oracle then
    // pays the V^2 kernel evaluations once per step, as the reference does)
    std::vector<double> Kd;
    static constexpr int kDenseMax = 8000;
    static constexpr double kStoreMax = 6e9;   // bytes of a stored K_xc (control-point form)
    void K_build(const double *q) {
        Kd.clear();
        if (V > kDenseMax) return;
        Kd.resize((size_t)V * V);
#pragma omp parallel for schedule(static)
        for (int j = 0; j < V; ++j)
            for (int i = 0; i < V; ++i) {
                double d2 = 0;
                for (int c = 0; c < D; ++c) { double d = q[c * V + i] - q[c * V + j]; d2 += d * d; }
                Kd[(size_t)j * V + i] = kernel_value(D, p.knlOrder, d2, p.knlWidth) + (i == j ? p.knlEps : 0.0);
            }
    }
    void K_apply(const double *q, const double *x, double *y) {
        if (Kd.empty()) { kernel_matvec<D>(p.knlOrder, p.knlWidth, p.knlEps, q, V, q, V, x, y, true); return; }
#pragma omp parallel for schedule(static)
        for (int i = 0; i < V; ++i) {
            double acc[3] = {0, 0, 0};
            for (int j = 0; j < V; ++j) {
                double k = Kd[(size_t)j * V + i];   // symmetric: K(i,j) = K(j,i)
                for (int c = 0; c < D; ++c) acc[c] += k * x[c * V + j];
            }
            for (int c = 0; c < D; ++c) y[c * V + i] = acc[c];
        }
    }

    // returns 0 ok, 1 failure; x in: warm start, out: solution; alpha in/out
    // (alpha = K^{-1} x), Lk: Cholesky factor of K when pcgInit == 0
    int ipcg(const double *q, const double *tau, const double *b, double *x, double *alpha,
             const std::vector<double> *Lk, int *iters) {
        size_t n = (size_t)D * V;
        double bb = dot(b, b, n);
        if (iters) *iters = 0;
        if (bb == 0.0) {
            std::fill(x, x + n, 0.0);
            std::fill(alpha, alpha + n, 0.0);
            return 0;
        }
        double tol = bb * p.cgTolSqu;
        std::vector<double> Kix(n), Ax(n), r(n), Kid(n), Kr(n), d(n), Bd(n), Ad(n);
        if (p.pcgInit == 0) {
            std::memcpy(Kix.data(), x, sizeof(double) * n);
            for (int c = 0; c < D; ++c) chol_solve(*Lk, V, Kix.data() + (size_t)c * V);
        } else {
            std::memcpy(Kix.data(), alpha, sizeof(double) * n);
        }
        B_apply(q, tau, x, Bd.data());
        for (size_t i = 0; i < n; ++i) Ax[i] = p.ldmWgt * Kix[i] + Bd[i];
        for (size_t i = 0; i < n; ++i) r[i] = b[i] - Ax[i];
        double rr = dot(r.data(), r.data(), n);
        std::vector<double> a(Kix);
        auto finish = [&]() { std::memcpy(alpha, a.data(), sizeof(double) * n); };
        if (rr < tol) { finish(); return 0; }
        Kid = r;
        K_apply(q, r.data(), Kr.data());
        d = Kr;
        double rKr = dot(r.data(), Kr.data(), n);
        if (rKr == 0.0) return 1;
        for (int it = 0; it < p.cgItrMax; ++it) {
            B_apply(q, tau, d.data(), Bd.data());
            for (size_t i = 0; i < n; ++i) Ad[i] = p.ldmWgt * Kid[i] + Bd[i];
            double dAd = dot(d.data(), Ad.data(), n);
            if (dAd == 0.0) return 1;
            double st = rKr / dAd;
            for (size_t i = 0; i < n; ++i) { x[i] += st * d[i]; a[i] += st * Kid[i]; r[i] -= st * Ad[i]; }
            double rr2 = dot(r.data(), r.data(), n);
            if (iters) *iters = it + 1;
            if (rr2 < tol) { finish(); return 0; }
            K_apply(q, r.data(), Kr.data());
            double rKr2 = dot(r.data(), Kr.data(), n);
            if (rKr2 == 0.0) return 1;
            double be = rKr2 / rKr;
            for (size_t i = 0; i < n; ++i) { Kid[i] = r[i] + be * Kid[i]; d[i] = Kr[i] + be * d[i]; }
            rKr = rKr2;
        }
        return 1;
    }

    // control-point form (build-defined): v = K_xc a, a on the control
    // points, solve (ldmWgt K_cc + K_cx B K_xc) a = K_cx f by CG
    // preconditioned with K_cc^{-1} (applied through a dense Cholesky of
    // K_cc).  With the control points = all landmarks this is the reference
    // system multiplied on both sides by K.
    int ctl_pcg(const double *q, const double *tau, const double *f, double *v, double *a, int *iters,
                double *jdOut = nullptr) {
        int C = m.ctlNum;
        size_t nc = (size_t)D * C, nv = (size_t)D * V;
        std::vector<double> qc(nc);
        for (int c = 0; c < D; ++c)
            for (int k = 0; k < C; ++k) qc[c * C + k] = q[c * V + m.ctlIdx[k]];
        // K_cc Cholesky
        std::vector<double> Kcc((size_t)C * C);
#pragma omp parallel for schedule(static)
        for (int j = 0; j < C; ++j)
            for (int i = 0; i < C; ++i) {
                double d2 = 0;
                for (int c = 0; c < D; ++c) { double d = qc[c * C + i] - qc[c * C + j]; d2 += d * d; }
                Kcc[(size_t)j * C + i] = kernel_value(D, p.knlOrder, d2, p.knlWidth) + (i == j ? p.knlEps : 0.0);
            }
        std::vector<double> Lcc(Kcc);
        if (!cholesky(Lcc, C)) return 1;
        // K_xc stored row by row (K[i C + j] = k(|q_i - qc_j|)) when it fits
        // in kStoreMax bytes, else both products matrix-free
        std::vector<double> Kst;
        if ((double)V * C * sizeof(double) <= kStoreMax) {
            Kst.resize((size_t)V * C);
#pragma omp parallel for schedule(static)
            for (int i = 0; i < V; ++i)
                for (int j = 0; j < C; ++j) {
                    double d2 = 0;
                    for (int c = 0; c < D; ++c) { double d = q[c * V + i] - qc[c * C + j]; d2 += d * d; }
                    Kst[(size_t)i * C + j] = kernel_value(D, p.knlOrder, d2, p.knlWidth);
                }
        }
        auto Kxc = [&](const double *al, double *out) {
            if (Kst.empty()) { kernel_matvec<D>(p.knlOrder, p.knlWidth, 0.0, q, V, qc.data(), C, al, out, false); return; }
#pragma omp parallel for schedule(static)
            for (int i = 0; i < V; ++i) {
                const double *ki = Kst.data() + (size_t)i * C;
                for (int c = 0; c < D; ++c) {
                    double acc = 0.0;
                    for (int j = 0; j < C; ++j) acc += ki[j] * al[c * C + j];
                    out[c * V + i] = acc;
                }
            }
        };
        auto Kcx = [&](const double *y, double *out) {
            if (Kst.empty()) { kernel_matvec<D>(p.knlOrder, p.knlWidth, 0.0, qc.data(), C, q, V, y, out, false); return; }
            // fixed row chunks, each summed in row order, the chunks added in
            // chunk order: the result does not depend on the thread count
            const int nch = std::max(1, std::min(64, V / 256));
            std::vector<double> part((size_t)nch * nc, 0.0);
#pragma omp parallel for schedule(static)
            for (int ch = 0; ch < nch; ++ch) {
                double *pc = part.data() + (size_t)ch * nc;
                const int i0 = (int)((long long)V * ch / nch), i1 = (int)((long long)V * (ch + 1) / nch);
                for (int i = i0; i < i1; ++i) {
                    const double *ki = Kst.data() + (size_t)i * C;
                    for (int c = 0; c < D; ++c) {
                        const double yi = y[c * V + i];
                        double *pcc = pc + (size_t)c * C;
                        for (int j = 0; j < C; ++j) pcc[j] += ki[j] * yi;
                    }
                }
            }
            for (size_t k = 0; k < nc; ++k) {
                double acc = part[k];
                for (int ch = 1; ch < nch; ++ch) acc += part[(size_t)ch * nc + k];
                out[k] = acc;
            }
        };
        auto S = [&](const double *al, double *out) {   // (ldmWgt K_cc + K_cx B K_xc) al
            std::vector<double> w(nv), bw(nv);
            Kxc(al, w.data());
            B_apply(q, tau, w.data(), bw.data());
            Kcx(bw.data(), out);
#pragma omp parallel for schedule(static)
            for (int i = 0; i < C; ++i)
                for (int c = 0; c < D; ++c) {
                    double s = 0;
                    for (int j = 0; j < C; ++j) s += Kcc[(size_t)i * C + j] * al[c * C + j];   // symmetric
                    out[c * C + i] += p.ldmWgt * s;
                }
        };
        std::vector<double> rhs(nc), r(nc), z(nc), d(nc), Sd(nc);
        Kcx(f, rhs.data());
        double bb = dot(rhs.data(), rhs.data(), nc);
        if (iters) *iters = 0;
        if (bb == 0.0) { std::fill(a, a + nc, 0.0); std::fill(v, v + nv, 0.0); return 0; }
        S(a, Sd.data());
        for (size_t i = 0; i < nc; ++i) r[i] = rhs[i] - Sd[i];
        double rr = dot(r.data(), r.data(), nc);
        double tol = bb * p.cgTolSqu;
        // Jacobi: diag(S) from S applied to every unit vector (brute force,
        // small meshes only), no knowledge of how the GPU assembles it
        std::vector<double> jd;
        if (p.velPrec == 1) {
            jd.resize(nc);
            std::vector<double> e(nc, 0.0), Se(nc);
            for (size_t k = 0; k < nc; ++k) {
                e[k] = 1.0;
                S(e.data(), Se.data());
                jd[k] = Se[k];
                e[k] = 0.0;
            }
            if (jdOut) std::memcpy(jdOut, jd.data(), sizeof(double) * nc);
        }
        auto prec = [&](const double *in, double *out) {
            std::memcpy(out, in, sizeof(double) * nc);
            if (p.velPrec == 1) {
                for (size_t k = 0; k < nc; ++k) out[k] /= jd[k];
                return;
            }
            chol_solve(Lcc, C, out, D);
        };
        int it = 0;
        if (rr >= tol) {
            prec(r.data(), z.data());
            d = z;
            double rz = dot(r.data(), z.data(), nc);
            for (; it < p.cgItrMax; ++it) {
                S(d.data(), Sd.data());
                double dSd = dot(d.data(), Sd.data(), nc);
                if (dSd == 0.0) return 1;
                double st = rz / dSd;
                for (size_t i = 0; i < nc; ++i) { a[i] += st * d[i]; r[i] -= st * Sd[i]; }
                rr = dot(r.data(), r.data(), nc);
                if (rr < tol) { ++it; break; }
                prec(r.data(), z.data());
                double rz2 = dot(r.data(), z.data(), nc);
                double be = rz2 / rz;
                rz = rz2;
                for (size_t i = 0; i < nc; ++i) d[i] = z[i] + be * d[i];
            }
            if (rr >= tol) return 1;
        }
        if (iters) *iters = it;
        Kxc(a, v);
        return 0;
    }

    // one forward solve: deform.cu:22-119 / objfcn.cu:23-123
    int run(const double *q0, const double *tauVar, const double *tauPrm, double *lmkStk, double *tauMat,
            int *pcgIters, int *rdIters) {
        int T = p.timeNum, E = m.elmNum;
        size_t nv = (size_t)D * V;
        Csr csr; build_csr(m, csr);
        std::memcpy(lmkStk, q0, sizeof(double) * nv);
        tau_initial<D>(q0, V, tauVar, tauPrm, tauMat);
        std::vector<double> v(nv, 0.0), alpha(m.ctlNum > 0 ? (size_t)D * m.ctlNum : nv, 0.0);
        std::vector<double> f(nv), ya(V), ra(V), F(V), rhs(V), lhs;
        std::vector<double> ppdNow, ggdNow, ppdNxt, ggdNxt, f1n, f2n;
        compute_frames<D>(m, q0, ws.f1, ws.f2);
        fem_assemble<D>(m, p, q0, ws.f1, ws.f2, ppdNow, ggdNow);
        (void)E;
        for (int t = 0; t < T - 1; ++t) {
            const double *q = lmkStk + (size_t)t * nv;
            double *qn = lmkStk + (size_t)(t + 1) * nv;
            const double *tau = tauMat + (size_t)t * V;
            double *taun = tauMat + (size_t)(t + 1) * V;
            for (int i = 0; i < V; ++i) { ya[i] = yank_activity(tau[i], p.ynk); ra[i] = reac_activity(tau[i], p.rea); }
            external_yank<D>(m, q, ya.data(), frc, f.data());
            int st, its = 0;
            if (m.ctlNum > 0) {
                st = ctl_pcg(q, tau, f.data(), v.data(), alpha.data(), &its);
            } else {
                std::vector<double> L;
                K_build(q);
                if (p.pcgInit == 0) {
                    if (Kd.empty()) return 3;   // reference-exact mode needs the dense K
                    L = Kd;
                    if (!cholesky(L, V)) return 1;
                } else if (t > 0) {
                    // warm start x0 = K_t alpha_{t-1}
                    K_apply(q, alpha.data(), v.data());
                }
                st = ipcg(q, tau, f.data(), v.data(), alpha.data(), &L, &its);
            }
            if (pcgIters) pcgIters[t] = its;
            if (st != 0) return 1;
            for (size_t i = 0; i < nv; ++i) qn[i] = q[i] + p.timeStp * v[i];
            compute_frames<D>(m, qn, f1n, f2n);
            fem_assemble<D>(m, p, qn, f1n, f2n, ppdNxt, ggdNxt);
            fem_vector<D>(m, q, ra.data(), F.data());
            lhs.resize(m.nzrNum);
            for (int z = 0; z < m.nzrNum; ++z) lhs[z] = 1.0 * ppdNxt[z] + p.timeStp * ggdNxt[z];
            csr_mv(csr, ppdNow.data(), tau, rhs.data(), V);
            for (int i = 0; i < V; ++i) rhs[i] = 1.0 * rhs[i] + p.timeStp * F[i];
            std::memcpy(taun, tau, sizeof(double) * V);   // CG warm start
            int rit = 0;
            if (rd_solve(m, p, csr, lhs, rhs.data(), taun, &rit) != 0) return 1;
            if (rdIters) rdIters[t] = rit;
            ws.f1.swap(f1n); ws.f2.swap(f2n);
            ppdNow.swap(ppdNxt); ggdNow.swap(ggdNxt);
        }
        return 0;
    }
};

// CPU-baseline sample: the first time step's set-up (frames, FEM matrices,
// yank) and its velocity CG run twice, with 0 and with nIter iterations, so
// out = [set-up s, CG start s (kernel factorisation, first residual and
// preconditioned residual), s per CG iteration]
template <int D>
void pcg_sample_impl(const OMesh *m, const OParams *p, const double *q0, const double *tauVar, const double *tauPrm,
                     int nIter, double *out) {
    int V = m->lmkNum;
    size_t nv = (size_t)D * V;
    auto now = [] { return std::chrono::steady_clock::now(); };
    auto sec = [](auto a, auto b) { return std::chrono::duration<double>(b - a).count(); };
    double tcg[2] = {0, 0};
    double tsetup = 0;
    for (int pass = 0; pass < 2; ++pass) {
        OParams pp = *p;
        pp.cgItrMax = pass == 0 ? 0 : nIter;
        pp.cgTolSqu = 0.0;
        Solver<D> s(*m, pp);
        auto t0 = now();
        std::vector<double> tau(V), ya(V), f(nv), v(nv, 0.0);
        std::vector<double> alpha(m->ctlNum > 0 ? (size_t)D * m->ctlNum : nv, 0.0);
        tau_initial<D>(q0, V, tauVar, tauPrm, tau.data());
        compute_frames<D>(*m, q0, s.ws.f1, s.ws.f2);
        std::vector<double> ppd, ggd;
        fem_assemble<D>(*m, pp, q0, s.ws.f1, s.ws.f2, ppd, ggd);
        for (int i = 0; i < V; ++i) ya[i] = yank_activity(tau[i], pp.ynk);
        external_yank<D>(*m, q0, ya.data(), s.frc, f.data());
        auto t1 = now();
        int its = 0;
        if (m->ctlNum > 0) {
            s.ctl_pcg(q0, tau.data(), f.data(), v.data(), alpha.data(), &its);
        } else {
            std::vector<double> L;
            if (V <= 8000) s.K_build(q0);
            s.ipcg(q0, tau.data(), f.data(), v.data(), alpha.data(), &L, &its);
        }
        auto t2 = now();
        tsetup = sec(t0, t1);
        tcg[pass] = sec(t1, t2);
    }
    out[0] = tsetup;
    out[1] = tcg[0];
    out[2] = nIter > 0 ? (tcg[1] - tcg[0]) / nIter : 0.0;
}

template <int D>
int objfcn_impl(const OMesh *m, const OParams *p, const OTarget *t, const double *q0, const double *tauVar,
                const double *tauPrm, double *lmkStk, double *tauMat, double *obj, double *grad, int *pcgIters,
                int *rdIters) {
    Solver<D> s(*m, *p);
    int st = s.run(q0, tauVar, tauPrm, lmkStk, tauMat, pcgIters, rdIters);
    if (st != 0) { if (obj) *obj = DBL_MAX; return st; }
    if (t && obj) {
        const double *qT = lmkStk + (size_t)(p->timeNum - 1) * D * m->lmkNum;
        std::vector<double> cen, dir, vol;
        facets<D>(*m, qT, cen, dir, vol);
        *obj = varifold_value<D>(cen, dir, vol, m->vfdElmNum, *t);
        if (grad) varifold_grad<D>(*m, qT, cen, dir, vol, *t, grad);
    } else if (obj) {
        *obj = 0.0;
    }
    return 0;
}

}  // namespace

extern "C" {

int or_num_threads() {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void or_set_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

// full forward solve + (optionally) objective and its gradient w.r.t. q_T
int or_objfcn(const OMesh *m, const OParams *p, const OTarget *t, const double *q0, const double *tauVar,
              const double *tauPrm, double *lmkStk, double *tauMat, double *obj, double *grad, int *pcgIters,
              int *rdIters) {
    if (m->dim == 2) return objfcn_impl<2>(m, p, t, q0, tauVar, tauPrm, lmkStk, tauMat, obj, grad, pcgIters, rdIters);
    return objfcn_impl<3>(m, p, t, q0, tauVar, tauPrm, lmkStk, tauMat, obj, grad, pcgIters, rdIters);
}

void or_pcg_sample(const OMesh *m, const OParams *p, const double *q0, const double *tauVar, const double *tauPrm,
                   int nIter, double *out) {
    if (m->dim == 2) pcg_sample_impl<2>(m, p, q0, tauVar, tauPrm, nIter, out);
    else pcg_sample_impl<3>(m, p, q0, tauVar, tauPrm, nIter, out);
}

// --- unit-level entry points for the tests ---------------------------------

void or_kernel_values(int dim, int order, double width, const double *d2, int n, double *out) {
    for (int i = 0; i < n; ++i) out[i] = kernel_value(dim, order, d2[i], width);
}

void or_bessel_k01(const double *x, int n, double *k0, double *k1) {
    for (int i = 0; i < n; ++i) bessel_k01(x[i], k0[i], k1[i]);
}

void or_activities(const double *tau, int n, const double *ynk, const double *rea, double *ya, double *ra) {
    for (int i = 0; i < n; ++i) { ya[i] = yank_activity(tau[i], ynk); ra[i] = reac_activity(tau[i], rea); }
}

void or_tau_initial(int dim, const double *q, int V, const double *tauVar, const double *tauPrm, double *tau) {
    if (dim == 2) tau_initial<2>(q, V, tauVar, tauPrm, tau);
    else tau_initial<3>(q, V, tauVar, tauPrm, tau);
}

void or_frames(const OMesh *m, const double *q, double *f1, double *f2) {
    std::vector<double> a, b;
    size_t n = (size_t)m->dim * m->elmNum;
    if (m->dim == 2) compute_frames<2>(*m, q, a, b); else compute_frames<3>(*m, q, a, b);
    std::memcpy(f1, a.data(), sizeof(double) * n);
    std::memcpy(f2, b.data(), sizeof(double) * n);
}

void or_elastic_apply(const OMesh *m, const OParams *p, const double *q, const double *tau, const double *v, double *out) {
    std::vector<double> frc;
    Workspace ws;
    if (m->dim == 2) { compute_frames<2>(*m, q, ws.f1, ws.f2); elastic_apply<2>(*m, *p, q, tau, ws, frc, v, out); }
    else { compute_frames<3>(*m, q, ws.f1, ws.f2); elastic_apply<3>(*m, *p, q, tau, ws, frc, v, out); }
}

// CPU-baseline sample of the operator alone: median wall time (s) of reps
// elastic applies E(q) v (frames formed once, as a solve forms them once per
// step) on the current OpenMP thread count
double or_time_elastic_apply(const OMesh *m, const OParams *p, const double *q, const double *tau, const double *v,
                             int reps) {
    std::vector<double> frc, out((size_t)m->dim * m->lmkNum), t((size_t)std::max(reps, 1));
    Workspace ws;
    if (m->dim == 2) compute_frames<2>(*m, q, ws.f1, ws.f2); else compute_frames<3>(*m, q, ws.f1, ws.f2);
    for (auto &tk : t) {
        auto t0 = std::chrono::steady_clock::now();
        if (m->dim == 2) elastic_apply<2>(*m, *p, q, tau, ws, frc, v, out.data());
        else elastic_apply<3>(*m, *p, q, tau, ws, frc, v, out.data());
        tk = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
    std::nth_element(t.begin(), t.begin() + t.size() / 2, t.end());
    return t[t.size() / 2];
}

void or_penalty_apply(const OMesh *m, const double *q, const double *v, double *out) {
    if (m->dim == 2) penalty_apply<2>(*m, q, v, out); else penalty_apply<3>(*m, q, v, out);
}

void or_external_yank(const OMesh *m, const double *q, const double *act, double *out) {
    std::vector<double> frc;
    if (m->dim == 2) external_yank<2>(*m, q, act, frc, out); else external_yank<3>(*m, q, act, frc, out);
}

void or_fem_assemble(const OMesh *m, const OParams *p, const double *q, double *ppd, double *ggd) {
    std::vector<double> f1, f2, a, b;
    if (m->dim == 2) { compute_frames<2>(*m, q, f1, f2); fem_assemble<2>(*m, *p, q, f1, f2, a, b); }
    else { compute_frames<3>(*m, q, f1, f2); fem_assemble<3>(*m, *p, q, f1, f2, a, b); }
    std::memcpy(ppd, a.data(), sizeof(double) * m->nzrNum);
    std::memcpy(ggd, b.data(), sizeof(double) * m->nzrNum);
}

void or_fem_vector(const OMesh *m, const double *q, const double *r, double *out) {
    if (m->dim == 2) fem_vector<2>(*m, q, r, out); else fem_vector<3>(*m, q, r, out);
}

void or_kernel_matvec(int dim, int order, double width, double eps, const double *q, int V, const double *x, double *y) {
    if (dim == 2) kernel_matvec<2>(order, width, eps, q, V, q, V, x, y, true);
    else kernel_matvec<3>(order, width, eps, q, V, q, V, x, y, true);
}

double or_varifold(const OMesh *m, const OTarget *t, const double *q, double *grad) {
    std::vector<double> cen, dir, vol;
    double v;
    if (m->dim == 2) {
        facets<2>(*m, q, cen, dir, vol);
        v = varifold_value<2>(cen, dir, vol, m->vfdElmNum, *t);
        if (grad) varifold_grad<2>(*m, q, cen, dir, vol, *t, grad);
    } else {
        facets<3>(*m, q, cen, dir, vol);
        v = varifold_value<3>(cen, dir, vol, m->vfdElmNum, *t);
        if (grad) varifold_grad<3>(*m, q, cen, dir, vol, *t, grad);
    }
    return v;
}

// diag(S) of the control-point system at positions q and concentration tau
// (the Jacobi preconditioner), by S applied to unit vectors
void or_jacobi_diag(const OMesh *m, const OParams *p, const double *q, const double *tau, double *out) {
    OParams pp = *p;
    pp.velPrec = 1;
    pp.cgItrMax = 0;
    const int V = m->lmkNum, D = m->dim;
    // any nonzero right-hand side: the diagonal is formed before the (empty) iteration
    std::vector<double> f((size_t)D * V, 1.0), v((size_t)D * V), a((size_t)D * m->ctlNum, 0.0);
    int its = 0;
    if (D == 2) {
        Solver<2> s(*m, pp);
        compute_frames<2>(*m, q, s.ws.f1, s.ws.f2);
        s.ctl_pcg(q, tau, f.data(), v.data(), a.data(), &its, out);
    } else {
        Solver<3> s(*m, pp);
        compute_frames<3>(*m, q, s.ws.f1, s.ws.f2);
        s.ctl_pcg(q, tau, f.data(), v.data(), a.data(), &its, out);
    }
}

void or_facets(const OMesh *m, const double *q, double *cen, double *dir, double *vol) {
    std::vector<double> c, d, w;
    if (m->dim == 2) facets<2>(*m, q, c, d, w); else facets<3>(*m, q, c, d, w);
    std::memcpy(cen, c.data(), sizeof(double) * c.size());
    std::memcpy(dir, d.data(), sizeof(double) * d.size());
    std::memcpy(vol, w.data(), sizeof(double) * w.size());
}

}  // extern "C"
