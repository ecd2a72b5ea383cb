// Device-side building blocks shared by the gfx950 kernels: small dense
// algebra on per-element matrices, the material laws, activity functions,
// RKHS kernels and deterministic block reductions.
#pragma once
#include <hip/hip_runtime.h>
#include <cfloat>
#include <cmath>

#include "bessel_cheb.hpp"

namespace sa {

constexpr int NT = 256;          // threads per block of the streaming kernels
constexpr int MAXP = 32768;      // capacity of a partial-sum buffer

// geometry record per element, SoA over elements: [Q^{-1} (D*D, row-major),
// det Q, frame 1 (D), frame 2 (D)]
template <int D> struct GeoLayout {
    static constexpr int INV = 0;
    static constexpr int DET = D * D;
    static constexpr int F1 = D * D + 1;
    static constexpr int F2 = D * D + 1 + D;
    static constexpr int N = D * D + 1 + 2 * D;
    static constexpr int NS = (D + 1) * (D + 2) / 2;   // upper-triangle entries of an element matrix
};

// ---------------------------------------------------------------------------
// reductions: wavefront butterfly then a fixed-order sum over the 4 waves.
// The order depends only on the block shape, so results are reproducible.

// Solver control values (flags, counters, the CG step history) are produced
// by one kernel and read by the next, possibly on another XCD: they are
// accessed with agent-scope atomic loads and stores, which are coherent
// across the XCDs' L2 caches and never served by the scalar data cache.
template <typename T>
__device__ inline T cld(const T *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T, typename U>
__device__ inline void cst(T *p, U v) {
    __hip_atomic_store(p, (T)v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ inline double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    return v;
}

// every thread of the block gets the block total; must be reached by all threads
__device__ inline double block_sum(double v, double *lds) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    v = wave_sum(v);
    __syncthreads();
    if (lane == 0) lds[w] = v;
    __syncthreads();
    double s = 0.0;
    const int nw = (blockDim.x + 63) >> 6;
    for (int i = 0; i < nw; ++i) s += lds[i];
    return s;
}

// sum of n partials, computed identically by every block that calls it
__device__ inline double reduce_partials(const double *P, int n, double *lds) {
    double s = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) s += P[i];
    return block_sum(s, lds);
}

// ---------------------------------------------------------------------------
// per-element inverse of Q = [q1-q0, ..., qD-q0]; rows of the inverse are the
// barycentric gradients (inc/matvec.h:196-203 (2D), :640-662 (3D): adjugate
// divided by the determinant)

template <int D>
__device__ inline void mat_inv(const double (&e)[D][D], double (&inv)[D][D], double &det) {
    if constexpr (D == 2) {
        det = e[0][0] * e[1][1] - e[0][1] * e[1][0];
        inv[0][0] = e[1][1] / det;  inv[0][1] = -e[1][0] / det;
        inv[1][0] = -e[0][1] / det; inv[1][1] = e[0][0] / det;
    } else {
        const double *x = e[0], *y = e[1], *z = e[2];
        det = x[0] * y[1] * z[2] + x[1] * y[2] * z[0] + x[2] * y[0] * z[1]
            - x[0] * y[2] * z[1] - x[1] * y[0] * z[2] - x[2] * y[1] * z[0];
        inv[0][0] =  (y[1] * z[2] - y[2] * z[1]) / det;
        inv[0][1] = -(y[0] * z[2] - y[2] * z[0]) / det;
        inv[0][2] =  (y[0] * z[1] - y[1] * z[0]) / det;
        inv[1][0] = -(x[1] * z[2] - x[2] * z[1]) / det;
        inv[1][1] =  (x[0] * z[2] - x[2] * z[0]) / det;
        inv[1][2] = -(x[0] * z[1] - x[1] * z[0]) / det;
        inv[2][0] =  (x[1] * y[2] - x[2] * y[1]) / det;
        inv[2][1] = -(x[0] * y[2] - x[2] * y[0]) / det;
        inv[2][2] =  (x[0] * y[1] - x[1] * y[0]) / det;
    }
}

template <int D> __device__ inline double elem_vol(double det) {
    return D == 2 ? 0.5 * fabs(det) : fabs(det) / 6.0;   // inc/utility.h:11-27
}

// frame-model stress (2D applyFrameElasticOperator.cu:7-39, 3D :7-111)
template <int D>
__device__ inline void frame_stress(const double (&eps)[D][D], const double (&f1)[D], const double (&f2)[D],
                                    const double *mod, double (&sig)[D][D]) {
    if constexpr (D == 2) {
        double es0 = eps[0][0] * f2[0] + eps[0][1] * f2[1], es1 = eps[1][0] * f2[0] + eps[1][1] * f2[1];
        double TvT = f1[0] * (eps[0][0] * f1[0] + eps[0][1] * f1[1]) + f1[1] * (eps[1][0] * f1[0] + eps[1][1] * f1[1]);
        double SvS = f2[0] * es0 + f2[1] * es1;
        double TvS = f1[0] * es0 + f1[1] * es1;
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b)
                sig[a][b] = mod[0] * (f1[a] * f1[b]) * TvT + mod[1] * (f2[a] * f2[b]) * SvS
                          + mod[2] * (f1[a] * f2[b] + f2[a] * f1[b]) * TvS;
    } else {
        double en[3], es[3];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            en[a] = eps[a][0] * f1[0] + eps[a][1] * f1[1] + eps[a][2] * f1[2];
            es[a] = eps[a][0] * f2[0] + eps[a][1] * f2[1] + eps[a][2] * f2[2];
        }
        double NvN = f1[0] * en[0] + f1[1] * en[1] + f1[2] * en[2];
        double SvS = f2[0] * es[0] + f2[1] * es[1] + f2[2] * es[2];
        double NvS = f1[0] * es[0] + f1[1] * es[1] + f1[2] * es[2];
        double ldv = eps[0][0] + eps[1][1] + eps[2][2] - NvN;
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int b = 0; b < 3; ++b) {
                double dab = (a == b ? 1.0 : 0.0);
                sig[a][b] = mod[0] * (dab - f1[a] * f1[b]) * ldv
                          + mod[1] * (eps[a][b] - (en[a] * f1[b] + f1[a] * en[b]) + NvN * f1[a] * f1[b])
                          + mod[2] * SvS * f2[a] * f2[b]
                          + mod[3] * ((es[a] * f2[b] + f2[a] * es[b]) - NvS * (f1[a] * f2[b] + f2[a] * f1[b]));
            }
    }
}

// ---------------------------------------------------------------------------
// activities, inc/activity.h:4-38, 81-90 (yank), 106-167, 239-249 (reaction)

__device__ inline double yank_activity(double tau, double tmin, double tmax, double amax, double dw) {
    double x = -1.0 + 2.0 * (tau - tmin) / (tmax - tmin);
    double a = 3.0 * (dw - 2.0) / (dw * dw * dw * (6.0 + dw * (-6.0 + dw)));
    double b = (6.0 + dw * (-15.0 + 8 * dw)) / (3.0 * (dw - 2.0));
    double c = 6.0 / (6.0 + dw * (-6.0 + dw));
    double y;
    if (x < -1.0 || x >= 1.0) y = 0.0;
    else if (x < -1.0 + dw) y = a * (x + 1.0) * (x + 1.0) * (x + 1.0) * (x - b);
    else if (x < 1.0 - dw) y = -c * x * x + 1.0;
    else y = a * (-x + 1.0) * (-x + 1.0) * (-x + 1.0) * (-x - b);
    return amax * y;
}

__device__ inline double reac_activity(double tau, double tmin, double tmax, double amax, double dw, double a) {
    double x = -1.0 + 2.0 * (tau - tmin) / (tmax - tmin);
    double b1 = 8.0 * (6.0 * (a - 1.0) + a * dw * (-6.0 + dw)) / (3.0 * dw * dw * dw * dw);
    double b2 = -(144.0 * (a - 1.0) + dw * (24.0 * (5.0 - 11.0 * a) + dw * (156.0 * a - dw * 29.0 * a)))
                / (24.0 * (6.0 * (a - 1.0) + a * dw * (-6.0 + dw)));
    double c1 = -(24.0 * (a - 1.0) + a * dw * (-12.0 + dw)) / (9.0 * dw * dw * dw);
    double c2 = -(24.0 * (a - 1.0) + dw * (12.0 * (2.0 - 3.0 * a) + a * dw * (13.0 + dw * 2.0))) / (3.0 * dw * dw * dw);
    double c3 = -((dw - 1.0) * (dw - 1.0) * (24.0 * (a - 1.0) + a * dw * (-12.0 + dw))) / (3.0 * dw * dw * dw);
    double c4 = -(24.0 * (a - 1.0) + dw * (12.0 * (6.0 - 7.0 * a)
                 + dw * (-72.0 + 109.0 * a + dw * (3.0 * (5.0 - 21.0 * a) + a * dw * (15.0 - dw)))))
                / (9.0 * dw * dw * dw);
    double y;
    if (x < -1.0 || x >= 1.0) y = 0.0;
    else if (x < -1.0 + 0.5 * dw) y = b1 * (x + 1.0) * (x + 1.0) * (x + 1.0) * (x - b2);
    else if (x < -1.0 + dw) y = c4 + x * (c3 + x * (c2 + x * c1));
    else if (x < 1.0 - dw) y = -a * x * x + 1.0;
    else if (x < 1.0 - 0.5 * dw) y = c4 + x * (-c3 + x * (c2 - x * c1));
    else y = b1 * (-x + 1.0) * (-x + 1.0) * (-x + 1.0) * (-x - b2);
    return amax * y;
}

// ---------------------------------------------------------------------------
// RKHS kernels of rho = |q_i - q_j| / width.  3D: computeKernel.cu:5-113.
// 2D: computeKernel.cu:30-168 with f0 = rho^2 K0(rho), f1 = rho K1(rho);
// exp(-r) for r >= 0 to about 1 ulp: r / (ln2 / 64) = -k with k integer,
// exp(-r) = 2^(k >> 6) * 2^((k & 63) / 64) * exp(s), |s| <= ln2 / 128, the
// 64 powers 2^(j/64) held in LDS (filled by init_exp_table) and exp(s) by its
// degree-5 Taylor polynomial (truncation < 4e-17).  The remainder s is formed
// with two FMAs against ln2/64 split into a double and its residual.
__device__ inline void init_exp_table(double *tab) {
    if (threadIdx.x < 64) tab[threadIdx.x] = exp2((double)threadIdx.x / 64.0);
}

__device__ inline double exp_neg(double r, const double *__restrict__ tab) {
    const double kInvStep = 92.332482616893656877;        // 64 / ln 2
    const double kStepHi = 0.010830424696249145;          // ln 2 / 64 (double)
    const double kStepLo = 3.6235106466348434e-19;        // ln 2 / 64 - kStepHi
    r = fmin(r, 1000.0);
    double kd = rint(-r * kInvStep);
    double s = fma(-kd, kStepHi, -r);
    s = fma(-kd, kStepLo, s);
    int k = (int)kd;
    double p = fma(s, 1.0 / 120.0, 1.0 / 24.0);
    p = fma(s, p, 1.0 / 6.0);
    p = fma(s, p, 0.5);
    p = fma(s, p, 1.0);
    p = fma(s, p, 1.0);
    return ldexp(tab[k & 63] * p, k >> 6);
}

// sqrt(x) for x >= 0 (x = 0 gives ~1e-150): hardware reciprocal square root
// refined by Goldschmidt iterations and two Newton corrections
__device__ inline double sqrt_pos(double x) {
    x = fmax(x, 1e-300);
    double y = __builtin_amdgcn_rsq(x);
    double g = x * y, h = 0.5 * y;
    double e = fma(-h, g, 0.5);
    g = fma(g, e, g);
    h = fma(h, e, h);
    double d = fma(-g, g, x);
    g = fma(d, h, g);
    d = fma(-g, g, x);
    return fma(d, h, g);
}

// kernel value from the squared distance; invW = 1 / width; tab = LDS table
// of init_exp_table (3D and Gaussian kernels)
// K0, K1 of the 2D Matern kernels as f0 = x^2 K0(x), f1 = x K1(x):
// x <= 2 from the power series in y = x^2/4 (14 terms, Horner), x > 2 from
// Chebyshev series of sqrt(x) e^x K_nu(x) in u = 4/x - 1 (Clenshaw) times
// e^-x / sqrt(x).  Both tables are generated by gen_bessel_cheb.py; relative
// error about 2e-16 on (0, 700].  A fixed amount of work per pair, no
// data-dependent loops.  tab: the exp_neg table, or null for the library exp.
template <int N>
__device__ inline double horner(const double (&c)[N], double y) {
    double s = c[N - 1];
#pragma unroll
    for (int k = N - 2; k >= 0; --k) s = fma(s, y, c[k]);
    return s;
}

template <int N>
__device__ inline double clenshaw(const double (&c)[N], double u) {
    double b1 = 0.0, b2 = 0.0;
    const double u2 = 2.0 * u;
#pragma unroll
    for (int k = N - 1; k >= 1; --k) {
        double t = fma(u2, b1, c[k] - b2);
        b2 = b1;
        b1 = t;
    }
    return fma(u, b1, c[0] - b2);
}

__device__ inline void bessel_f01(double x, double &f0, double &f1, const double *tab) {
    if (x < DBL_EPSILON) { f0 = 0.0; f1 = 1.0; return; }
    if (x > 700.0) { f0 = 0.0; f1 = 0.0; return; }
    double k0, k1;
    if (x <= 2.0) {
        const double eg = 0.57721566490153286061;
        const double y = 0.25 * x * x, lg = log(0.5 * x);
        const double i0 = horner(kSerI0, y), s0 = horner(kSerS0, y);
        const double i1s = horner(kSerI1, y), s1 = horner(kSerS1, y);
        k0 = -(lg + eg) * i0 + s0;
        k1 = 1.0 / x + lg * (0.5 * x * i1s) - 0.25 * x * s1;
    } else {
        const double u = 4.0 / x - 1.0;
        const double e = (tab ? exp_neg(x, tab) : exp(-x)) / sqrt_pos(x);
        k0 = clenshaw(kCheb0, u) * e;
        k1 = clenshaw(kCheb1, u) * e;
    }
    f0 = x * x * k0;
    f1 = x * k1;
}

template <int D, int ORD>
__device__ inline double rkhs_kernel(double d2, double invW, const double *__restrict__ tab) {
    if constexpr (ORD == -1) {
        return exp_neg(d2 * (invW * invW), tab);
    } else if constexpr (D == 3) {
        double r = sqrt_pos(d2) * invW;
        double e = exp_neg(r, tab);
        if constexpr (ORD == 0) return e;
        if constexpr (ORD == 1) return (1.0 + r) * e;
        if constexpr (ORD == 2) return (3.0 + r * (3.0 + r)) / 3.0 * e;
        if constexpr (ORD == 3) return (15.0 + r * (15.0 + r * (6.0 + r))) / 15.0 * e;
        return (105.0 + r * (105.0 + r * (45.0 + r * (10.0 + r)))) / 105.0 * e;
    } else {
        double r = sqrt(d2) * invW;
        double f0, f1;
        bessel_f01(r, f0, f1, tab);
        double r2 = r * r;
        if constexpr (ORD == 0) return f1;
        if constexpr (ORD == 1) return 0.5 * (f0 + 2.0 * f1);
        if constexpr (ORD == 2) return (4.0 * f0 + (8.0 + r2) * f1) / 8.0;
        if constexpr (ORD == 3) return ((24.0 + r2) * f0 + 8.0 * (6.0 + r2) * f1) / 48.0;
        return (12.0 * (16.0 + r2) * f0 + (384.0 + r2 * (72.0 + r2)) * f1) / 384.0;
    }
}

// single-precision variant for the mixed-precision inner solve
template <int D, int ORD>
__device__ inline float rkhs_kernel_f(float d2, float invW) {
    if constexpr (ORD == -1) {
        return __expf(-d2 * (invW * invW));
    } else if constexpr (D == 3) {
        float r = __fsqrt_rn(d2) * invW;
        float e = __expf(-r);
        if constexpr (ORD == 0) return e;
        if constexpr (ORD == 1) return (1.0f + r) * e;
        if constexpr (ORD == 2) return (3.0f + r * (3.0f + r)) * (1.0f / 3.0f) * e;
        if constexpr (ORD == 3) return (15.0f + r * (15.0f + r * (6.0f + r))) * (1.0f / 15.0f) * e;
        return (105.0f + r * (105.0f + r * (45.0f + r * (10.0f + r)))) * (1.0f / 105.0f) * e;
    } else {
        return (float)rkhs_kernel<D, ORD>((double)d2, (double)invW, nullptr);
    }
}

// ---------------------------------------------------------------------------
// varifold pair kernels, 2D|3D varifold.cu:38-179 (centre) and (direction)

__device__ inline double cen_kernel(int type, double d2, double invW2) {
    if (type == 'G') return exp(-d2 * invW2);
    return 1.0 / (1.0 + d2 * invW2);   // 'C'
}
__device__ inline double cen_kernel_d(int type, double k, double invW2) {   // coefficient of (c_i - c_j)
    if (type == 'G') return -2.0 * k * invW2;
    return -2.0 * k * k * invW2;
}
__device__ inline double dir_kernel(int type, double ang, double invW2) {
    if (type == 'B') return ang * ang;
    if (type == 'L') return ang;
    if (type == 'O') return exp(2.0 * (ang - 1.0) * invW2);
    return exp(2.0 * (ang * ang - 1.0) * invW2);   // 'U'
}
__device__ inline double dir_kernel_d(int type, double ang, double k, double invW2) {   // d k / d ang
    if (type == 'B') return 2.0 * ang;
    if (type == 'L') return 1.0;
    if (type == 'O') return 2.0 * k * invW2;
    return 4.0 * ang * k * invW2;
}

}  // namespace sa
