// Host engine and C ABI of the gfx950 forward shape-evolution path.
//
// One Engine<D> holds a mesh, its derived tables and all per-step state in
// device memory (a session).  run() performs one candidate solve: the time
// loop of deform.cu:22-119 / objfcn.cu:23-123 with
//   * the RKHS velocity system (ldmWgt K^{-1} + E + btmWgt P) v = f solved by
//     the K-preconditioned CG of ipcg.cu:11-100 with K applied matrix-free;
//     K^{-1} of the warm start is carried along (alpha = K^{-1} v), so no
//     factorisation of K is ever formed;
//   * the reaction-diffusion system (M + dt K_D) tau = M tau + dt M r solved
//     by Jacobi-CG on the CSR matrix given by femVtxMat / femIfoMat instead
//     of a dense Cholesky;
//   * the varifold distance (and its gradient) by LDS-tiled pair sums.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <condition_variable>
#include <mutex>
#include <vector>
#include <type_traits>
#include <cstdlib>
#include <dlfcn.h>

#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>
#include <rocblas/rocblas.h>

#include "shape_atrophy.h"
#include "sa_kernels.hpp"

namespace sa {
namespace {

thread_local std::string g_err;

// Device gate.  rocSOLVER's Cholesky factorisation (ROCm 7.2, gfx950) returns
// wrong factors when kernels of another hardware queue run beside it:
// rocsolver_dpotrf of a 2,000 x 2,000 kernel matrix gave a different factor
// in 16 of 20 calls (and reported it indefinite in 7) while a memory-bound
// kernel ran on another stream, and the same factor in every call alone, with
// the kernels serialised (AMD_SERIALIZE_KERNEL=3, HIP_LAUNCH_BLOCKING=1) or
// with one hardware queue (GPU_MAX_HW_QUEUES=1).  dtrtri, dsyrk, dgemm and
// dsyevd gave bitwise the same results under the same load.  So the
// factorisation runs with no other session's work on the device: a session
// holds the gate shared while it has work queued and lets it go at its host
// synchronisations when a factorisation waits (gate_yield); the
// factorisation holds it exclusive, waiting for the others to drain.
class DeviceGate {
    std::mutex mu_;
    std::condition_variable cv_;
    int shared_ = 0, waiting_ = 0;
    bool excl_ = false;

  public:
    void lock_shared() {
        std::unique_lock<std::mutex> l(mu_);
        cv_.wait(l, [&] { return !excl_ && waiting_ == 0; });
        ++shared_;
    }
    void unlock_shared() {
        std::lock_guard<std::mutex> l(mu_);
        --shared_;
        cv_.notify_all();
    }
    void lock() {
        std::unique_lock<std::mutex> l(mu_);
        ++waiting_;
        cv_.wait(l, [&] { return !excl_ && shared_ == 0; });
        --waiting_;
        excl_ = true;
    }
    void unlock() {
        std::lock_guard<std::mutex> l(mu_);
        excl_ = false;
        cv_.notify_all();
    }
    bool exclusive_waiting() {
        std::lock_guard<std::mutex> l(mu_);
        return waiting_ > 0;
    }
};
DeviceGate g_gate;
std::mutex g_handleMu;              // rocBLAS handle creation and destruction
thread_local int g_gateDepth = 0;   // shared holds of this thread (nested scopes count once)
// SA_DEVICE_GATE=0 switches the gate off (to measure what it prevents)
const bool g_gateOn = [] {
    const char *e = std::getenv("SA_DEVICE_GATE");
    return !(e && e[0] == '0');
}();

// shared hold for the span of an entry point that queues device work (on =
// false: no hold, for the sizing pass)
struct GateShared {
    bool on;
    explicit GateShared(bool use = true) : on(use && g_gateOn) {
        if (on && g_gateDepth++ == 0) g_gate.lock_shared();
    }
    ~GateShared() {
        if (on && --g_gateDepth == 0) g_gate.unlock_shared();
    }
    GateShared(const GateShared &) = delete;
    GateShared &operator=(const GateShared &) = delete;
};

// at a host synchronisation (this thread's stream idle): step aside for a
// waiting library call
void gate_yield() {
    if (g_gateDepth > 0 && g_gate.exclusive_waiting()) {
        g_gate.unlock_shared();
        g_gate.lock_shared();
    }
}

// exclusive hold for the factorisation (its stream synchronised inside)
struct GateExclusive {
    bool held;
    GateExclusive() : held(g_gateDepth > 0) {
        if (!g_gateOn) return;
        if (held) g_gate.unlock_shared();
        g_gate.lock();
    }
    ~GateExclusive() {
        if (!g_gateOn) return;
        g_gate.unlock();
        if (held) g_gate.lock_shared();
    }
    GateExclusive(const GateExclusive &) = delete;
    GateExclusive &operator=(const GateExclusive &) = delete;
};

// Block-count targets of the stored kernel products (row panels x column
// slices), per product and matrix shape, from an offline sweep on the device
// kept in tiles.txt beside the library.  A line
//   <product> <f32|f64> <rows> <cols> <target blocks>
// names the product xc (K_xc: landmarks x control points), cx (K_cx), cc
// (the C x C preconditioner M) or kk (K over all landmarks), the stored
// element type and the shape it was measured on; '#' starts a comment.  A
// line "op f64 <elements> <landmarks> <shape>" gives instead the launch
// shape (index into OPF_SHAPES) of the fused operator apply.  A
// session takes the entry of its product and type nearest its shape (sum of
// the log distances of rows and columns), else the built-in default.
// SA_TILE_TABLE names another file ("" reads none); SA_KS_TARGET overrides
// every entry ("2048", or per product "xc=1024,cx=4096"), for the sweep.
// The slice counts are a function of the mesh and the table, so the partial
// sums of one mesh are added in one fixed order.
struct TileEntry {
    char prod[4];
    bool f32;
    double rows, cols;
    int target;
};

std::vector<TileEntry> load_tile_table() {
    std::vector<TileEntry> t;
    std::string path;
    if (const char *e = std::getenv("SA_TILE_TABLE")) {
        path = e;
    } else {
        Dl_info info;
        if (dladdr(reinterpret_cast<void *>(&load_tile_table), &info) && info.dli_fname) {
            path = info.dli_fname;
            const size_t sl = path.rfind('/');
            path = (sl == std::string::npos ? std::string(".") : path.substr(0, sl)) + "/tiles.txt";
        }
    }
    if (path.empty()) return t;
    FILE *f = std::fopen(path.c_str(), "r");
    if (!f) return t;
    char line[256];
    while (std::fgets(line, sizeof line, f)) {
        if (char *h = std::strchr(line, '#')) *h = 0;
        TileEntry en{};
        char prod[8], ty[8];
        long rows, cols;
        int tg;
        if (std::sscanf(line, "%7s %7s %ld %ld %d", prod, ty, &rows, &cols, &tg) != 5) continue;
        const std::string p(prod), y(ty);
        const bool op = p == "op";
        if ((p != "xc" && p != "cx" && p != "cc" && p != "kk" && !op) || (y != "f32" && y != "f64") || rows < 1 ||
            cols < 1 || (op ? (tg < 0 || tg >= OPF_NSHAPES) : (tg < 16 || tg > 65536)))
            continue;
        std::memcpy(en.prod, prod, 3);
        en.f32 = y == "f32";
        en.rows = (double)rows;
        en.cols = (double)cols;
        en.target = tg;
        t.push_back(en);
    }
    std::fclose(f);
    return t;
}

const std::vector<TileEntry> &tile_table() {
    static const std::vector<TileEntry> t = load_tile_table();
    return t;
}

// the table's entry of a product and type nearest a shape (dflt: none)
int tile_lookup(const char *prod, bool f32, long rows, long cols, int dflt) {
    const TileEntry *best = nullptr;
    double bd = 0.0;
    for (const auto &en : tile_table()) {
        if (std::strcmp(en.prod, prod) || en.f32 != f32) continue;
        const double d = std::fabs(std::log(en.rows / (double)rows)) + std::fabs(std::log(en.cols / (double)cols));
        if (!best || d < bd) { best = &en; bd = d; }
    }
    return best ? best->target : dflt;
}

// target block count of one stored product (see TileEntry)
int ks_target(const char *prod, bool f32, long rows, long cols, int dflt) {
    if (const char *e = std::getenv("SA_KS_TARGET")) {
        // "N" for every product, or "xc=N,cx=M,..."
        const std::string s(e);
        if (!s.empty() && s.find('=') == std::string::npos) {
            const int v = std::atoi(s.c_str());
            if (v >= 16 && v <= 65536) return v;
        }
        const std::string key = std::string(prod) + "=";
        const size_t k = s.find(key);
        if (k != std::string::npos && (k == 0 || s[k - 1] == ',')) {
            const int v = std::atoi(s.c_str() + k + key.size());
            if (v >= 16 && v <= 65536) return v;
        }
    }
    return tile_lookup(prod, f32, rows, cols, dflt);
}

#ifndef SA_MIXED_CC32
#define SA_MIXED_CC32 1   // mixed form: K_cc from the FP32 stored rows, FP32 preconditioner
#endif
// spectral preconditioner: iterations recorded, Ritz vectors kept
#ifndef SA_SPEC_M
#define SA_SPEC_M 384
#endif
// (measured at the 51k shell, mixed form: 128 Ritz vectors of 384 recorded
// iterations 16.95 steps/s; 192 or 256 of 384, 256 of 512: 16.60-16.61,
// 1.2 fewer CG iterations per step for a longer spectral term per iteration)
#ifndef SA_SPEC_K
#define SA_SPEC_K 128
#endif
// how M = X^T X is formed: 0 dgemm, 1 dsyrk, 2 lauum (in place)
#ifndef SA_MPRE_BUILD
#define SA_MPRE_BUILD 1
#endif
// mixed form: squared relative tolerance of an inner FP32 solve (at most;
// looser when the outer target needs less): 1e-9, a residual cut of about
// 3e-5 per sweep, measured best on the 51k-landmark shell (1e-10 and 1e-8
// took 1.5 % and 0.6 % longer per step)
#ifndef SA_INNER_TOL2
#define SA_INNER_TOL2 1e-9
#endif
// mixed form: squared relative tolerance of an inner solve that records the
// Krylov basis of a spectral harvest (never below the outer target): a
// longer basis gives a stronger term for all the solves that carry it.
// Measured at the 51k shell: 13.85 (1e-9, as the other inner solves) /
// 14.40 (1e-11) / 14.77 (1e-12) / 14.80 (1e-13) / 15.13 (1e-14) steps/s
#ifndef SA_REC_TOL2
#define SA_REC_TOL2 1e-14
#endif
// mixed form: refinement residuals from the FP32 matrices and their FP32
// remainders (1) or from matrix-free FP64 kernel products (0)
#ifndef SA_SPLIT_RESID
#define SA_SPLIT_RESID 1
#endif
// the remainders are stored when both fit in this many bytes
#ifndef SA_SPLIT_MAX_GB
#define SA_SPLIT_MAX_GB 8
#endif
// refreshes of the control-point preconditioner in a row that carry the
// spectral term over instead of harvesting a new one (default: all of a
// run's scheduled refreshes).  Measured at the 51k shell, FP64 form: 9.79
// (never) / 10.19 (1) / 10.36 (2) / 10.69 (always) steps/s; mixed form with
// a longer recording solve (SA_REC_TOL2): 14.28 (never, 1e-13) / 15.37 (4,
// 1e-16) / 15.62 (always, 1e-16)
#ifndef SA_SPEC_CARRY
#define SA_SPEC_CARRY 1000000
#endif
// steps a control-point preconditioner is reused at most (measured at the
// 51k shell, mixed form: 17.0 (8) / 16.6 (12) / 16.2 (16) steps/s)
#ifndef SA_PREC_MAX_AGE
#define SA_PREC_MAX_AGE 8
#endif
#ifndef SA_EXTRAP
#define SA_EXTRAP 3   // order of the initial-guess extrapolation across steps
#endif

struct SaError : std::runtime_error {
    int code;
    SaError(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

#define CK(x)                                                                                           \
    do {                                                                                                \
        hipError_t e_ = (x);                                                                            \
        if (e_ != hipSuccess)                                                                           \
            throw SaError(SA_E_DEVICE, std::string(#x) + ": " + hipGetErrorString(e_) + " (line " +   \
                                           std::to_string(__LINE__) + ")");                            \
    } while (0)

#define CHECK_LAUNCH() CK(hipGetLastError())

// the preconditioner product's matrix loads: ordinary (0, the default) or
// non-temporal (1) like the other stored products.  M (100 MB in FP32 at
// 5,000 control points) then stays in the 256 MB last-level cache from one
// CG iteration to the next, past the non-temporal streams of K_xc and K_cx:
// 17.51 against 17.41 steps/s on the 51k-landmark shell, mixed form
#ifndef SA_PRE_NT
#define SA_PRE_NT 0
#endif
constexpr bool PRE_NT = SA_PRE_NT != 0;
// all-landmark form: a stored K of at most this many MiB is read with
// ordinary loads in the CG iterations (0: always non-temporal).  The
// 2,000-landmark ring's 32 MB product: 11.2 against 12.7 us a launch; its
// whole solve, launch-latency bound, 116.5-117.6 against 114.0-117.9 steps/s
#ifndef SA_KST_CACHED_MB
#define SA_KST_CACHED_MB 64
#endif

inline int gridOf(long n, int nt = NT) { return (int)((n + nt - 1) / nt); }

// Set while a constructor only sizes a session (sa_session_footprint): every
// allocation is counted and nothing touches the device, so the footprint is
// exactly what a session of the same mesh and parameters allocates.
thread_local bool g_sizing = false;
struct SizingScope {
    SizingScope() { g_sizing = true; }
    ~SizingScope() { g_sizing = false; }
};

template <typename T>
struct DBuf {
    T *p = nullptr;
    size_t n = 0;
    DBuf() = default;
    DBuf(const DBuf &) = delete;
    DBuf &operator=(const DBuf &) = delete;
    ~DBuf() { if (p) (void)hipFree(p); }
    void alloc(size_t count, long long &total) {
        if (p) { (void)hipFree(p); p = nullptr; }
        n = count;
        if (count == 0) return;
        if (g_sizing) {
            total += (long long)(sizeof(T) * count);
            return;
        }
        if (hipMalloc((void **)&p, sizeof(T) * count) != hipSuccess) {
            p = nullptr;
            throw SaError(SA_E_ALLOC, "Fail to allocate device memory (" + std::to_string(sizeof(T) * count) + " bytes).");
        }
        total += (long long)(sizeof(T) * count);
        // SA_POISON_ALLOC=1: floating-point buffers start as NaN, so a read
        // of memory no kernel has written shows up in the results
        if constexpr (std::is_floating_point<T>::value) {
            static const bool poison = [] {
                const char *e = std::getenv("SA_POISON_ALLOC");
                return e && e[0] == '1';
            }();
            if (poison) {
                CK(hipMemset(p, 0xff, sizeof(T) * count));
                CK(hipDeviceSynchronize());   // before any upload on a session stream
            }
        }
    }
    void upload(const T *h, size_t count, hipStream_t s) {
        if (!p || count == 0) return;   // sizing, or nothing to copy
        CK(hipMemcpyAsync(p, h, sizeof(T) * count, hipMemcpyHostToDevice, s));
    }
    void zero(hipStream_t s) {
        if (p) CK(hipMemsetAsync(p, 0, sizeof(T) * n, s));
    }
};

// non-owning view of part of a DBuf (the stored kernel matrices share one
// allocation; p == nullptr means absent)
template <typename T>
struct View {
    T *p = nullptr;
};

// host copy of an input array that may live on the device
template <typename T>
std::vector<T> to_host(const T *src, size_t count, bool on_device) {
    std::vector<T> h(count);
    if (count == 0) return h;
    if (!src) throw SaError(SA_E_ARG, "null input array");
    if (on_device) {
        if (hipMemcpy(h.data(), src, sizeof(T) * count, hipMemcpyDeviceToHost) != hipSuccess)
            throw SaError(SA_E_ARG, "cannot read device input array");
    } else {
        std::memcpy(h.data(), src, sizeof(T) * count);
    }
    return h;
}

void require(bool ok, const std::string &msg) {
    if (!ok) throw SaError(SA_E_ARG, msg);
}

// count-prefixed incidence table (vertex -> (facet, local) in ascending facet
// order), the layout of vtxElmMat built by generateElementObject2D.m:32-41
std::vector<int> incidence(const std::vector<int> &items, int nItems, int k, int num, int &rows) {
    std::vector<int> cnt(num, 0);
    for (int l = 0; l < k; ++l)
        for (int f = 0; f < nItems; ++f) cnt[items[(size_t)l * nItems + f]]++;
    int mx = 0;
    for (int v = 0; v < num; ++v) mx = std::max(mx, cnt[v]);
    rows = 1 + 2 * mx;
    std::vector<int> tbl((size_t)rows * num, 0);
    std::vector<int> fill(num, 0);
    for (int f = 0; f < nItems; ++f)
        for (int l = 0; l < k; ++l) {
            int v = items[(size_t)l * nItems + f];
            int r = fill[v]++;
            tbl[(size_t)(1 + 2 * r) * num + v] = f;
            tbl[(size_t)(2 + 2 * r) * num + v] = l;
        }
    for (int v = 0; v < num; ++v) tbl[v] = cnt[v];
    return tbl;
}

// positions (SoA, n points, D coordinates) -> their indices in Morton
// (Z-curve) order of the coordinates quantised to 2^(63/D) cells per
// dimension over the bounding box; ties keep index order
std::vector<int> morton_order(const double *p, int n, int D) {
    std::vector<int> idx(n);
    for (int i = 0; i < n; ++i) idx[i] = i;
    if (n < 2) return idx;
    const int bits = D == 3 ? 21 : 31;
    double lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
    for (int c = 0; c < D; ++c) {
        lo[c] = hi[c] = p[(size_t)c * n];
        for (int i = 1; i < n; ++i) {
            lo[c] = std::min(lo[c], p[(size_t)c * n + i]);
            hi[c] = std::max(hi[c], p[(size_t)c * n + i]);
        }
    }
    double ext = 0.0;
    for (int c = 0; c < D; ++c) ext = std::max(ext, hi[c] - lo[c]);
    const double scale = ext > 0.0 ? (double)((1ull << bits) - 1) / ext : 0.0;
    std::vector<unsigned long long> code(n, 0ull);
    for (int i = 0; i < n; ++i) {
        unsigned long long q[3] = {0, 0, 0};
        for (int c = 0; c < D; ++c) q[c] = (unsigned long long)((p[(size_t)c * n + i] - lo[c]) * scale);
        unsigned long long m = 0;
        for (int b = bits - 1; b >= 0; --b)
            for (int c = 0; c < D; ++c) m = (m << 1) | ((q[c] >> b) & 1ull);
        code[i] = m;
    }
    std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return code[a] < code[b]; });
    return idx;
}

// Chunk tables of the fused operator apply (op_fused_k), built from the
// count-prefixed vertex -> element lists hVE (vtxElmMat layout), the element
// table hElm (elmVtxMat layout) and a vertex order ord: vertices taken in
// that order while a chunk has at most nt vertices, maxs list entries and
// hm halo vertices (the vertices of the elements its lists name); per chunk
// each such element once (ascending element id) with the halo position of
// each of its vertices and the LDS slot of each listed (element, local
// vertex) pair.  loc and slot are SoA by local vertex: [l NU + u].  An empty
// result: a vertex whose elements alone pass hm halo vertices.
struct OpTables {
    std::vector<int> chV{0}, chE{0}, chH{0}, chVtx, chLs, chHv, chElm, loc, slot;
};

OpTables op_tables(const std::vector<int> &hVE, const std::vector<int> &hElm, int V, int E, int D,
                   const std::vector<int> &ord, int nt, int maxs, int hm) {
    OpTables t;
    std::vector<int> aos, laos, mark(E, -1), hmark(V, -1), stamp(V, -1), perm, tmpE, tmpS, tmpL;
    int nv = 0, ns = 0, nh = 0, pass = 0;
    auto close_chunk = [&] {
        // the chunk's element entries in ascending element order
        const int u0 = t.chE.back(), u1 = (int)t.chElm.size();
        perm.resize(u1 - u0);
        for (int u = u0; u < u1; ++u) perm[u - u0] = u;
        std::sort(perm.begin(), perm.end(), [&](int a, int b) { return t.chElm[a] < t.chElm[b]; });
        tmpE.assign(perm.size(), 0);
        tmpS.assign(perm.size() * (D + 1), 0);
        tmpL.assign(perm.size() * (D + 1), 0);
        for (size_t r = 0; r < perm.size(); ++r) {
            tmpE[r] = t.chElm[perm[r]];
            for (int l = 0; l <= D; ++l) {
                tmpS[r * (D + 1) + l] = aos[(size_t)perm[r] * (D + 1) + l];
                tmpL[r * (D + 1) + l] = laos[(size_t)perm[r] * (D + 1) + l];
            }
        }
        for (size_t r = 0; r < perm.size(); ++r) {
            t.chElm[u0 + r] = tmpE[r];
            mark[tmpE[r]] = -1;
            for (int l = 0; l <= D; ++l) {
                aos[(size_t)(u0 + r) * (D + 1) + l] = tmpS[r * (D + 1) + l];
                laos[(size_t)(u0 + r) * (D + 1) + l] = tmpL[r * (D + 1) + l];
            }
        }
        for (int h = t.chH.back(); h < (int)t.chHv.size(); ++h) hmark[t.chHv[h]] = -1;
        t.chV.push_back((int)t.chVtx.size());
        t.chE.push_back(u1);
        t.chH.push_back((int)t.chHv.size());
        nv = ns = nh = 0;
    };
    // halo vertices vertex i would add to the open chunk
    auto growth = [&](int i) {
        ++pass;
        int g = 0;
        for (int a = 0; a < hVE[i]; ++a) {
            const int e = hVE[(size_t)(1 + 2 * a) * V + i];
            if (mark[e] >= 0) continue;
            for (int l = 0; l <= D; ++l) {
                const int w = hElm[(size_t)l * E + e];
                if (hmark[w] < 0 && stamp[w] != pass) { stamp[w] = pass; ++g; }
            }
        }
        return g;
    };
    for (int i : ord) {
        const int adj = hVE[i];
        int g = growth(i);
        if (nv > 0 && (nv == nt || ns + adj > maxs || nh + g > hm)) {
            close_chunk();
            g = growth(i);
        }
        if (adj > maxs || g > hm) return OpTables{};
        t.chVtx.push_back(i);
        t.chLs.push_back(ns);
        for (int a = 0; a < adj; ++a) {
            const int e = hVE[(size_t)(1 + 2 * a) * V + i], l = hVE[(size_t)(2 + 2 * a) * V + i];
            if (mark[e] < 0) {
                mark[e] = (int)t.chElm.size();
                t.chElm.push_back(e);
                aos.insert(aos.end(), D + 1, -1);
                for (int l2 = 0; l2 <= D; ++l2) {
                    const int w = hElm[(size_t)l2 * E + e];
                    if (hmark[w] < 0) {
                        hmark[w] = nh++;
                        t.chHv.push_back(w);
                    }
                    laos.push_back(hmark[w]);
                }
            }
            aos[(size_t)mark[e] * (D + 1) + l] = ns + a;
        }
        ns += adj;
        ++nv;
    }
    if (nv > 0) close_chunk();
    const size_t NU = t.chElm.size();
    t.slot.assign((size_t)(D + 1) * NU, -1);
    t.loc.assign((size_t)(D + 1) * NU, -1);
    for (size_t u = 0; u < NU; ++u)
        for (int l = 0; l <= D; ++l) {
            t.slot[(size_t)l * NU + u] = aos[u * (D + 1) + l];
            t.loc[(size_t)l * NU + u] = laos[u * (D + 1) + l];
        }
    return t;
}

// the invariants op_fused_k's indexing relies on: every vertex in exactly
// one chunk of at most nt vertices; a chunk's list entries (at most maxs)
// each own one LDS slot, written by exactly one (element, local vertex)
// pair of that chunk; its halo (at most hm vertices) holds each element
// entry's vertices at the positions loc names; ids in range
void check_op_tables(const OpTables &t, const std::vector<int> &hVE, const std::vector<int> &hElm, int V, int E,
                     int D, int nt, int maxs, int hm) {
    const int K = (int)t.chV.size() - 1;
    const int NU = (int)t.chElm.size();
    require(K >= 1 && t.chV[0] == 0 && t.chV[K] == V && (int)t.chVtx.size() == V && (int)t.chLs.size() == V &&
                (int)t.chE.size() == K + 1 && t.chE[0] == 0 && t.chE[K] == NU && (int)t.chH.size() == K + 1 &&
                t.chH[0] == 0 && t.chH[K] == (int)t.chHv.size() && t.slot.size() == (size_t)(D + 1) * NU &&
                t.loc.size() == (size_t)(D + 1) * NU,
            "operator chunk tables: bad shape");
    std::vector<char> seen(V, 0), wr;
    for (int k = 0; k < K; ++k) {
        const int v0 = t.chV[k], v1 = t.chV[k + 1], h0 = t.chH[k], nh = t.chH[k + 1] - h0;
        require(v1 > v0 && v1 - v0 <= nt && t.chE[k + 1] >= t.chE[k] && nh >= 0 && nh <= hm,
                "operator chunk tables: bad chunk");
        for (int h = h0; h < h0 + nh; ++h) require(t.chHv[h] >= 0 && t.chHv[h] < V, "operator chunk tables: bad halo");
        int ns = 0;
        for (int j = v0; j < v1; ++j) {
            const int i = t.chVtx[j];
            require(i >= 0 && i < V && !seen[i] && t.chLs[j] == ns, "operator chunk tables: bad vertex entry");
            seen[i] = 1;
            ns += hVE[i];
        }
        require(ns <= maxs, "operator chunk tables: chunk over its slots");
        wr.assign(ns, 0);
        for (int u = t.chE[k]; u < t.chE[k + 1]; ++u) {
            const int e = t.chElm[u];
            require(e >= 0 && e < E, "operator chunk tables: element out of range");
            for (int l = 0; l <= D; ++l) {
                const int lc = t.loc[(size_t)l * NU + u];
                require(lc >= 0 && lc < nh && t.chHv[h0 + lc] == hElm[(size_t)l * E + e],
                        "operator chunk tables: halo position wrong");
                const int sl = t.slot[(size_t)l * NU + u];
                if (sl < 0) continue;
                require(sl < ns && !wr[sl], "operator chunk tables: slot out of range or written twice");
                wr[sl] = 1;
            }
        }
        for (int a = 0; a < ns; ++a) require(wr[a], "operator chunk tables: slot never written");
    }
}

// the launch shape of the fused operator apply for a mesh of E elements
// and V vertices: SA_OPF_CFG (index into OPF_SHAPES), else the tile table's
// "op" entry nearest the mesh, else OPF_DEFAULT
constexpr int OPF_DEFAULT = 1;
int opf_cfg(long E, long V) {
    if (const char *e = std::getenv("SA_OPF_CFG")) {
        const int c = std::atoi(e);
        if (c >= 0 && c < OPF_NSHAPES) return c;
    }
    const int c = tile_lookup("op", false, std::max(E, 1L), std::max(V, 1L), OPF_DEFAULT);
    return c >= 0 && c < OPF_NSHAPES ? c : OPF_DEFAULT;
}

struct Stats {
    int status = 0, steps = 0, pcgIters = 0, rdIters = 0, refine = 0, fp64Steps = 0, failReason = 0;
    double kmvMs = 0.0, opMs = 0.0;   // summed durations of the timed kernels (profiling on)
    int kmvCount = 0, opCount = 0;
    double vfdMs = 0.0;               // the varifold row-sum launch of the objective (profiling on)
    long long vfdPairs = 0;           //   ... the facet pairs it summed
    int vfdMfma = 0;                  //   ... 1 when it ran on the matrix cores (vfd_rows_mfma_k)
    int factorizations = 0;           // K_cc factorisations (preconditioner rebuilds) of the run
};

// hipEvent pairs around the kernel-matrix product and the elastic operator
// of every velocity-CG iteration; only iterations that did work are summed
struct KernelTimer {
    bool on = false;
    std::vector<hipEvent_t> ev;   // per chunk slot: [kmv start, kmv end, op start, op end, kmv2 start, kmv2 end]
    void ensure(int slots) {
        while ((int)ev.size() < 6 * slots) {
            hipEvent_t e;
            if (hipEventCreate(&e) != hipSuccess) throw SaError(SA_E_DEVICE, "hipEventCreate failed");
            ev.push_back(e);
        }
    }
    hipEvent_t vfd[2] = {nullptr, nullptr};   // around the varifold row sums
    void ensure_vfd() {
        for (auto &e : vfd)
            if (!e && hipEventCreate(&e) != hipSuccess) throw SaError(SA_E_DEVICE, "hipEventCreate failed");
    }
    ~KernelTimer() {
        for (auto e : ev) (void)hipEventDestroy(e);
        for (auto e : vfd)
            if (e) (void)hipEventDestroy(e);
    }
};

struct EngineBase {
    long long bytes = 0;
    virtual ~EngineBase() = default;
    virtual void run(const double *tauVar, int nVar, const double *tauPrm, int nPrm, double *obj, double *lmkStk,
                     double *tauMat, double *grad, Stats &st) = 0;
    int Trun = 0;                      // time levels of the next run (<= allocated)
    virtual int timeLevels() const = 0;
    virtual void probe(int what, const double *tauVar, int nVar, const double *tauPrm, int nPrm, const double *in,
                       double *out) = 0;
    virtual void set_params(const sa_params &p) = 0;
    virtual double time_product(int what, int reps) = 0;
    KernelTimer timer;
};

template <int D>
struct Engine : EngineBase {
    using G = GeoLayout<D>;
    static constexpr int NS = G::NS;
    // sizes
    int V, E, Z, B, BL, F = 0, Ft = 0, T, NC = 0;
    int vtxElmRows, vtxBtmRows, femIfoRows, vfdAdjRows = 0;
    // parameters
    int elaType, knlOrder, itrMax, itrMax0, precision;
    double ldmWgt, knlWidth, knlEps, spdTan, spdTsv, btmWgt, dt, cgTolSqu, rdTolSqu;
    double ynk[4], rea[5];
    std::vector<double> mod;
    bool hasTarget = false;
    VfdPrm vp{};
    hipStream_t s = nullptr;

    // tables
    DBuf<int> elmVtx, vtxElm, frmVtx, tsvVtx, btmVtx, vtxBtm, femIfo, rowPtr, colIdx, diagIdx, vfdVtx, vfdAdj, elmLyr,
        ctlIdx;
    // vertex-ordered element forces of the operator apply: elmSlot[l E + e]
    // is the position of (e, l) in the concatenated vtxElmMat lists,
    // vtxStart[i] where vertex i's list starts (absent when an (e, l) is
    // listed twice: the element-ordered gather then runs instead)
    DBuf<int> elmSlot, vtxStart;
    // the operator apply in one launch (op_fused_k): vertex chunks and the
    // elements their lists name; SA_OP_FUSED=0 keeps the two-launch apply
    // (elastic_elem_k + gather_op_k)
    DBuf<int> opChV, opChVtx, opChLs, opChH, opChHv, opChE, opChElm, opChLS;
    DBuf<double> opFr[2];        // element-entry frames of geo[0], geo[1] (frame model)
    int opChunks = 0, opNU = 0, opCapU = 0, opCfg = 0;
    bool opFused = false;
    // mixed form on the matrix cores (kx_mfma_k): the kernel products of the
    // inner solve evaluated matrix-free, nothing stored; landmarks and
    // control points visited in Morton order (position -> index)
    bool kxMfma = false;
    DBuf<int> lmkPerm, ctlPerm;
    // Jacobi preconditioning of the control-point CG (sa_params.velPrec):
    // per-element stiffness blocks, their per-nonzero gather, the lumped
    // penalty weights and diag(S), rebuilt every step; no factorisation
    bool jacobi = false;
    DBuf<double> stiffBlk, bcc, pw, dS;
    int slicesAm = 1, slicesTm = 1;
    // boundary yank (sa_params.bdrYank): boundary facets, their vertex
    // adjacency, the facet forces; Bn = 0 when the tables are absent or unusable
    DBuf<int> bdrVtx, vtxBdr;
    DBuf<double> bfrc;
    int Bn = 0;
    bool bdrYank = false;
    DBuf<double> q0, dmod, tCen, tDir, tVol;
    // state
    DBuf<double> lmk, tau, geo[2], femElm, ppd[2], lhs, ya, ra, f, x, a, r, Kid, Kr, d, Ad, frc, pfrc, meas;
    DBuf<double> r2;   // the all-landmark CG's residual after odd iterations (step_product_dir)
    DBuf<double> rdRhs, rdR, rdZ, rdD, rdAd, rdD2;
    DBuf<double> P0, P1, P2, P3, hist, rdHist;
    DBuf<double> dCen, dDir, dVol, rowVal, gC, gN, grad, scal;
    DBuf<Ctl> ctl, rdCtl;
    // control-point form
    DBuf<double> qc, Kcc, Kc, Mpre, Mcol, cG, cR, cZ, cD, cS, cY, cK, cRout, cDelta, cZero, part, partS;
    DBuf<double> cA, cA0, cAm1, cAm2, cAm3;   // alpha, its value entering a step, the last three solutions
    DBuf<double> cR2, cD2;       // second residual / direction buffers (iteration parity, fused CG)
    // stored K_xc (V x C) and K_cx (C x V) of the control-point form, views
    // into kmem: FP64 (plain form) or FP32 with their FP32 remainders (mixed
    // form).  A mixed session whose remainders are stored keeps the FP64 pair
    // of its fallback steps in the same memory (alias64): the FP32 quartet
    // takes 8 bytes per entry pair, the FP64 pair too (use_fp64_store)
    DBuf<double> kmem;
    View<double> A64, At64;      // stored K_xc / K_cx, FP64
    size_t na32 = 0, nt32 = 0, na64 = 0, nt64 = 0;
    bool alias64 = false;
    bool matrixFree = false;     // SA_MATRIX_FREE=1: FP64 kernel products never stored
    DBuf<double> Kst;            // all-landmark form: stored K (V x V, panel-major FP64), rebuilt per step
    DBuf<double> aM1, aM2, aM3;  // all-landmark form: the last three alpha = K^-1 v
    int slicesK = 1;
    View<float> A32, At32;       // stored K_xc / K_cx, FP32 (mixed precision)
    View<float> A32lo, At32lo;   // their FP32 remainders (split products of the refinement residuals)
    DBuf<float> Mpre32;          // FP32 panel copy of the preconditioner (mixed, SA_MIXED_CC32)
    // spectral term of the control-point preconditioner, z = M r + U diag(cf) U^T r:
    // U holds Ritz vectors of M S for its largest eigenvalues lambda_l (orthonormal
    // in the M^-1 inner product), cf_l = theta / lambda_l - 1 maps them to theta,
    // the next Ritz value.  They are harvested from the Krylov basis of the first
    // solve after each refresh of M (Rayleigh-Ritz on the recorded d_j, S d_j, r_j)
    // and used until M is rebuilt.  SA_SPECTRAL=0 switches the term off.
    static constexpr int SPEC_M = SA_SPEC_M;   // iterations recorded
    static constexpr int SPEC_K = SA_SPEC_K;   // Ritz vectors kept at most
    static_assert(SPEC_K <= SPEC_KMAX, "SA_SPEC_K above the kernels' LDS capacity");
    DBuf<double> recD, recSD, recR, specU, specC, specW, specA, specScr;
    DBuf<int> specInfo;
    bool specEnabled = true, specOn = false, specRecord = false, recActive = false;
    // the next refresh of M carries the spectral term over (spec_carry_finish)
    // instead of recording a new basis; specCarried counts carried refreshes
    // since the last harvest
    bool specCarry = false;
    int specCarried = 0;
    static constexpr double SPEC_CARRY_GMIN = 0.5;   // Gram eigenvalues kept within [0.5, 2]
    int specK = 0, specMode = -1;
    int slicesA = 1, slicesT = 1, slicesM = 1, slicesM32 = 1;
    DBuf<int> potrfInfo;
    // K_cc factorisation: chol_*_k (default), or rocsolver_dpotrf behind the
    // device gate (SA_CHOL=rocsolver)
    bool cholLib = false;
    rocblas_handle blas = nullptr;
    int slicesC = 1;
    Ctl *hctl = nullptr;
    Stats *prof = nullptr;       // where the kernel timer accumulates during run()
    int lastIters = 0;           // CG iterations of the previous step (sizes the first launch chunk)
    // control-point preconditioner X^T X from the K_cc of an earlier step:
    // rebuilt on a run's first step, after PREC_MAX_AGE steps, or when a
    // step needs PREC_GROWTH times the CG iterations of the first step that
    // used it (any SPD preconditioner leaves the solution unchanged)
    static constexpr int PREC_MAX_AGE = SA_PREC_MAX_AGE;
    static constexpr double PREC_GROWTH = 1.2;
    int precAge = -1, precIters = 0;
    int fp64Steps = 0;           // mixed-form steps finished by the FP64 form (this run)
    bool mixedOff = false;       // a step of this run stalled: the rest run in FP64
    bool precEveryStep = false;  // a stale preconditioner failed in this run: refresh every step
    bool forceRefresh = false;   // refresh at the next step (after a stalled mixed step)
    // set while an inner FP32 solve of the mixed refinement runs: ctl_pcg_run
    // then stops (status 3) when |r|^2 has not halved over STALL_WIN
    // iterations, once STALL_MIN are done
    bool stallCheck = false;
    static constexpr int STALL_MIN = 200, STALL_WIN = 100;
    int failReason = 0;          // SA_FAIL_* of a failed run

    // mixed form, fallback to FP64: the FP64 pair takes the memory of the
    // FP32 quartet for the rest of the run (matrix-free FP64 products when
    // the session has no quartet); restore_fp32_store undoes it at the
    // start of the next run.  Padding rows of either layout must be zero.
    void use_fp64_store() {
        if (!alias64 || A64.p) return;
        kmem.zero(s);
        A64.p = kmem.p;
        At64.p = kmem.p + na64;
        A32.p = At32.p = A32lo.p = At32lo.p = nullptr;
    }
    void set_fp32_views() {
        float *f = reinterpret_cast<float *>(kmem.p);
        A32.p = f;
        At32.p = f + na32;
        A32lo.p = alias64 ? At32.p + nt32 : nullptr;
        At32lo.p = alias64 ? A32lo.p + na32 : nullptr;
    }
    void restore_fp32_store() {
        if (!alias64 || !A64.p) return;
        kmem.zero(s);
        A64.p = At64.p = nullptr;
        set_fp32_views();
    }
    int chunkHint = 4;           // first launch chunk of the next control-point CG
    int sweepIters[4] = {0, 0, 0, 0};   // inner iterations of the last step's refinement sweeps
    double ttSum = 0.0;
    bool ttReady = false;
    bool vfdMfma = false;        // varifold value pair contractions on MFMA

    int timeLevels() const override { return T; }

    // scalar model parameters for the following runs (sizes, the elasticity
    // model, precision and the allocated time levels stay)
    void set_params(const sa_params &p) override {
        require(p.elaType == elaType, "set_params cannot change the elasticity model");
        require(p.modVec && p.modNum == (int)mod.size(), "set_params needs a modVec of the session's length");
        require(p.ynkActPrmVec && p.reaActPrmVec, "null activity parameters");
        require(p.knlOrder == knlOrder, "set_params cannot change the kernel order");
        require(p.knlWidth > 0, "knlWidth must be positive");
        require(p.timeNum >= 1 && p.timeNum <= T, "timeNum exceeds the session's allocation");
        require(p.cgItrMax >= 1 && p.cgItrMax <= itrMax0, "cgItrMax exceeds the session's allocation");
        require(!p.bdrYank || Bn > 0, "bdrYank needs consistent bdrVtxMat / vtxBdrMat");
        require((p.velPrec == SA_VELPREC_JACOBI) == jacobi, "set_params cannot change the velocity preconditioner");
        bdrYank = p.bdrYank != 0;
        mod.assign(p.modVec, p.modVec + p.modNum);
        CK(hipMemcpyAsync(dmod.p, mod.data(), sizeof(double) * mod.size(), hipMemcpyHostToDevice, s));
        CK(hipStreamSynchronize(s));
        ldmWgt = p.ldmWgt; knlWidth = p.knlWidth; knlEps = p.knlEps;
        spdTan = p.spdTanVal; spdTsv = p.spdTsvVal; btmWgt = p.btmWgt; dt = p.timeStp;
        itrMax = p.cgItrMax; cgTolSqu = p.cgTolSqu;
        double rdTol = p.rdTolVal > 0 ? p.rdTolVal : 1e-13;
        rdTolSqu = rdTol * rdTol;
        for (int i = 0; i < 4; ++i) ynk[i] = p.ynkActPrmVec[i];
        for (int i = 0; i < 5; ++i) rea[i] = p.reaActPrmVec[i];
        Trun = p.timeNum;
    }

    ~Engine() override {
        if (blas) {
            std::lock_guard<std::mutex> lk(g_handleMu);
            (void)rocblas_destroy_handle(blas);
        }
        if (hctl) (void)hipHostFree(hctl);
        if (s) (void)hipStreamDestroy(s);
    }

    Engine(const sa_mesh &m, const sa_target *tg, const sa_params &p, const double *lmkIni) {
        GateShared gate(!g_sizing);
        const bool dev = m.on_device != 0;
        V = m.lmkNum; E = m.elmNum; Z = m.nzrNum; B = m.btmElmNum; BL = m.btmLmkNum;
        T = p.timeNum;
        require(V > 0 && E > 0 && Z >= V, "empty mesh or inconsistent nzrNum");
        require(T >= 1, "timeNum must be >= 1");
        require(p.knlOrder >= -1 && p.knlOrder <= 4, "knlOrder must be in -1..4");
        require(p.elaType == SA_ELA_FRAME || p.elaType == SA_ELA_ISOTROPIC, "unknown elasticity type");
        vtxElmRows = m.vtxElmCols; vtxBtmRows = m.vtxBtmCols; femIfoRows = m.femIfoCols;
        elaType = p.elaType; knlOrder = p.knlOrder; itrMax = p.cgItrMax; itrMax0 = p.cgItrMax; precision = p.precision;
        ldmWgt = p.ldmWgt; knlWidth = p.knlWidth; knlEps = p.knlEps; spdTan = p.spdTanVal; spdTsv = p.spdTsvVal;
        btmWgt = p.btmWgt; dt = p.timeStp; cgTolSqu = p.cgTolSqu;
        double rdTol = p.rdTolVal > 0 ? p.rdTolVal : 1e-13;
        rdTolSqu = rdTol * rdTol;
        require(p.ynkActPrmVec && p.reaActPrmVec && p.modVec, "null parameter vector");
        for (int i = 0; i < 4; ++i) ynk[i] = p.ynkActPrmVec[i];
        for (int i = 0; i < 5; ++i) rea[i] = p.reaActPrmVec[i];
        mod.assign(p.modVec, p.modVec + p.modNum);
        if (elaType == SA_ELA_FRAME) require(p.modNum >= (D == 2 ? 3 : 4), "modVec too short for the frame model");
        else require(p.modNum >= 2 && p.modNum % 2 == 0, "isotropic modVec needs (lambda, mu) pairs");
        require(knlWidth > 0, "knlWidth must be positive");

        // --- host copies, validation, derived tables
        auto hElm = to_host(m.elmVtxMat, (size_t)E * (D + 1), dev);
        auto hVtxElm = to_host(m.vtxElmMat, (size_t)V * vtxElmRows, dev);
        auto hFrm = to_host(m.frmVtxMat, (size_t)E * (D == 2 ? 4 : 6), dev);
        auto hTsv = to_host(m.tsvVtxMat, (size_t)E * (D == 2 ? 4 : 6), dev);
        auto hFemVtx = to_host(m.femVtxMat, (size_t)Z * 2, dev);
        auto hFemIfo = to_host(m.femIfoMat, (size_t)Z * femIfoRows, dev);
        std::vector<int> hBtm, hVtxBtm;
        if (B > 0) {
            hBtm = to_host(m.btmVtxMat, (size_t)B * D, dev);
            hVtxBtm = to_host(m.vtxBtmMat, (size_t)BL * vtxBtmRows, dev);
        }
        // boundary facets (only the optional boundary yank reads them; the
        // reference passes them and leaves them unused): kept when consistent
        std::vector<int> hBdr, hVtxBdr;
        if (m.bdrElmNum > 0 && m.bdrVtxMat && m.vtxBdrMat && m.vtxBdrCols > 0) {
            const int nb = m.bdrElmNum, rows = m.vtxBdrCols;
            hBdr = to_host(m.bdrVtxMat, (size_t)nb * D, dev);
            hVtxBdr = to_host(m.vtxBdrMat, (size_t)V * rows, dev);
            bool ok = true;
            for (int v : hBdr) ok = ok && v >= 0 && v < V;
            for (int i = 0; i < V && ok; ++i) {
                const int adj = hVtxBdr[i];
                ok = adj >= 0 && 1 + 2 * adj <= rows;
                for (int a = 0; a < adj && ok; ++a) {
                    const int b = hVtxBdr[(size_t)(1 + 2 * a) * V + i], l = hVtxBdr[(size_t)(2 + 2 * a) * V + i];
                    ok = b >= 0 && b < nb && l >= 0 && l < D && hBdr[(size_t)l * nb + b] == i;
                }
            }
            if (ok) Bn = nb;
            else { hBdr.clear(); hVtxBdr.clear(); }
        }
        bdrYank = p.bdrYank != 0;
        require(!bdrYank || Bn > 0, "bdrYank needs consistent bdrVtxMat / vtxBdrMat");
        require(p.velPrec == SA_VELPREC_DEFAULT || p.velPrec == SA_VELPREC_JACOBI, "unknown velPrec");
        require(p.velPrec != SA_VELPREC_JACOBI || m.ctlNum > 0,
                "Jacobi preconditioning applies to the control-point form (ctlNum > 0)");
        jacobi = p.velPrec == SA_VELPREC_JACOBI;
        for (int v : hElm) require(v >= 0 && v < V, "elmVtxMat index out of range");
        for (int v : hFrm) require(v >= 0 && v < V, "frame stencil index out of range");
        for (int v : hTsv) require(v >= 0 && v < V, "tsvVtxMat index out of range");
        for (int v : hBtm) require(v >= 0 && v < V, "btmVtxMat index out of range");
        require(BL <= V, "btmLmkNum exceeds lmkNum");
        for (int i = 0; i < V; ++i) {
            int adj = hVtxElm[i];
            require(adj >= 0 && 1 + 2 * adj <= vtxElmRows, "vtxElmMat count out of range");
            for (int a = 0; a < adj; ++a) {
                int e = hVtxElm[(size_t)(1 + 2 * a) * V + i], l = hVtxElm[(size_t)(2 + 2 * a) * V + i];
                require(e >= 0 && e < E && l >= 0 && l <= D && hElm[(size_t)l * E + e] == i, "vtxElmMat inconsistent with elmVtxMat");
            }
        }
        for (int i = 0; i < BL; ++i) {
            int adj = hVtxBtm[i];
            require(adj >= 0 && 1 + 2 * adj <= vtxBtmRows, "vtxBtmMat count out of range");
            for (int a = 0; a < adj; ++a) {
                int b = hVtxBtm[(size_t)(1 + 2 * a) * BL + i], l = hVtxBtm[(size_t)(2 + 2 * a) * BL + i];
                require(b >= 0 && b < B && l >= 0 && l < D, "vtxBtmMat entry out of range");
            }
        }
        // CSR from the sorted femVtxMat
        std::vector<int> hRow(V + 1, 0), hCol(Z), hDiag(V, -1);
        for (int z = 0; z < Z; ++z) {
            int rr = hFemVtx[z], cc = hFemVtx[(size_t)Z + z];
            require(rr >= 0 && rr < V && cc >= 0 && cc < V, "femVtxMat index out of range");
            require(z == 0 || rr > hFemVtx[z - 1] || (rr == hFemVtx[z - 1] && cc > hFemVtx[(size_t)Z + z - 1]),
                    "femVtxMat is not sorted by (row, col)");
            hRow[rr + 1]++;
            hCol[z] = cc;
            if (rr == cc) hDiag[rr] = z;
            int adj = hFemIfo[z];
            require(adj >= 0 && 2 + 2 * adj <= femIfoRows, "femIfoMat count out of range");
            for (int a = 0; a < adj; ++a) {
                int e = hFemIfo[(size_t)(2 + 2 * a) * Z + z], gi = hFemIfo[(size_t)(3 + 2 * a) * Z + z];
                require(e >= 0 && e < E && gi >= 0 && gi < NS * E, "femIfoMat entry out of range");
            }
        }
        for (int i = 0; i < V; ++i) { hRow[i + 1] += hRow[i]; require(hDiag[i] >= 0, "femVtxMat lacks a diagonal entry"); }
        std::vector<int> hLyr;
        if (elaType == SA_ELA_ISOTROPIC && m.elmLyrVec) {
            hLyr = to_host(m.elmLyrVec, (size_t)E, dev);
            int nl = p.modNum / 2;
            for (int l : hLyr) require(l >= 0 && l < nl, "elmLyrVec layer out of range");
        }
        std::vector<int> hCtl;
        if (m.ctlNum > 0) {
            NC = m.ctlNum;
            hCtl = to_host(m.ctlIdxVec, (size_t)NC, dev);
            for (int v : hCtl) require(v >= 0 && v < V, "ctlIdxVec index out of range");
        }

        {
            // SA_MATRIX_FREE=1: FP64 kernel products stay matrix-free (as
            // above the 48 GB store cap), for testing that form at small sizes
            const char *mf = std::getenv("SA_MATRIX_FREE");
            matrixFree = mf && mf[0] == '1';
        }
        const bool dryRun = g_sizing;   // sizing only: no stream, no handles, no copies
        if (!dryRun) {
            CK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
            CK(hipHostMalloc((void **)&hctl, sizeof(Ctl) * 2));
        }
        long long &tot = bytes;
        elmVtx.alloc(hElm.size(), tot); elmVtx.upload(hElm.data(), hElm.size(), s);
        vtxElm.alloc(hVtxElm.size(), tot); vtxElm.upload(hVtxElm.data(), hVtxElm.size(), s);
        {
            std::vector<int> hSlot((size_t)(D + 1) * E, -1), hStart((size_t)V + 1, 0);
            bool once = true;
            for (int i = 0; i < V; ++i) {
                const int adj = hVtxElm[i];
                hStart[i + 1] = hStart[i] + adj;
                for (int a = 0; a < adj && once; ++a) {
                    const int e = hVtxElm[(size_t)(1 + 2 * a) * V + i], l = hVtxElm[(size_t)(2 + 2 * a) * V + i];
                    int &sl = hSlot[(size_t)l * E + e];
                    if (sl >= 0) once = false;
                    sl = hStart[i] + a;
                }
            }
            // the slot layout must fit the force buffer ((D + 1) D doubles per element)
            if (once && hStart[V] <= (D + 1) * E) {
                elmSlot.alloc(hSlot.size(), tot); elmSlot.upload(hSlot.data(), hSlot.size(), s);
                vtxStart.alloc(hStart.size(), tot); vtxStart.upload(hStart.data(), hStart.size(), s);
                if (!dryRun) CK(hipStreamSynchronize(s));   // the host vectors go out of scope
            }
        }
        frmVtx.alloc(hFrm.size(), tot); frmVtx.upload(hFrm.data(), hFrm.size(), s);
        tsvVtx.alloc(hTsv.size(), tot); tsvVtx.upload(hTsv.data(), hTsv.size(), s);
        femIfo.alloc(hFemIfo.size(), tot); femIfo.upload(hFemIfo.data(), hFemIfo.size(), s);
        rowPtr.alloc(V + 1, tot); rowPtr.upload(hRow.data(), V + 1, s);
        colIdx.alloc(Z, tot); colIdx.upload(hCol.data(), Z, s);
        diagIdx.alloc(V, tot); diagIdx.upload(hDiag.data(), V, s);
        if (B > 0) {
            btmVtx.alloc(hBtm.size(), tot); btmVtx.upload(hBtm.data(), hBtm.size(), s);
            vtxBtm.alloc(hVtxBtm.size(), tot); vtxBtm.upload(hVtxBtm.data(), hVtxBtm.size(), s);
        }
        if (Bn > 0) {
            bdrVtx.alloc(hBdr.size(), tot); bdrVtx.upload(hBdr.data(), hBdr.size(), s);
            vtxBdr.alloc(hVtxBdr.size(), tot); vtxBdr.upload(hVtxBdr.data(), hVtxBdr.size(), s);
            bfrc.alloc((size_t)D * D * Bn, tot);
            if (!dryRun) CK(hipStreamSynchronize(s));   // the host tables go out of scope
        }
        if (!hLyr.empty()) { elmLyr.alloc(E, tot); elmLyr.upload(hLyr.data(), E, s); }
        dmod.alloc(mod.size(), tot); dmod.upload(mod.data(), mod.size(), s);
        q0.alloc((size_t)V * D, tot);
        std::vector<double> hq0;
        if (!dryRun) {
            hq0 = to_host(lmkIni, (size_t)V * D, dev);
            q0.upload(hq0.data(), hq0.size(), s);
        }
        if (vtxStart.n) build_op_chunks(hVtxElm, hElm, hq0, dryRun);

        if (tg) {
            hasTarget = true;
            F = m.vfdElmNum;
            Ft = tg->tgtElmNum;
            require(F > 0 && Ft > 0, "empty varifold facet set");
            auto hVfd = to_host(m.vfdVtxMat, (size_t)F * D, dev);
            for (int v : hVfd) require(v >= 0 && v < V, "vfdVtxMat index out of range");
            auto adj = incidence(hVfd, F, D, V, vfdAdjRows);
            vfdVtx.alloc(hVfd.size(), tot); vfdVtx.upload(hVfd.data(), hVfd.size(), s);
            vfdAdj.alloc(adj.size(), tot); vfdAdj.upload(adj.data(), adj.size(), s);
            auto hc = to_host(tg->cenPosMat, (size_t)Ft * D, dev);
            auto hd = to_host(tg->uniDirMat, (size_t)Ft * D, dev);
            auto hv = to_host(tg->elmVolVec, (size_t)Ft, dev);
            tCen.alloc(hc.size(), tot); tCen.upload(hc.data(), hc.size(), s);
            tDir.alloc(hd.size(), tot); tDir.upload(hd.data(), hd.size(), s);
            tVol.alloc(hv.size(), tot); tVol.upload(hv.data(), hv.size(), s);
            require(tg->cenKnlType == 'G' || tg->cenKnlType == 'C', "cenKnlType must be 'G' or 'C'");
            require(tg->dirKnlType == 'B' || tg->dirKnlType == 'L' || tg->dirKnlType == 'O' || tg->dirKnlType == 'U',
                    "dirKnlType must be 'B', 'L', 'O' or 'U'");
            vp.cenType = tg->cenKnlType;
            vp.dirType = tg->dirKnlType;
            vp.cenInvW2 = 1.0 / (tg->cenKnlWidth * tg->cenKnlWidth);
            vp.dirInvW2 = 1.0 / (tg->dirKnlWidth * tg->dirKnlWidth);
            dCen.alloc((size_t)F * D, tot); dDir.alloc((size_t)F * D, tot); dVol.alloc(F, tot);
            rowVal.alloc((size_t)F + Ft, tot);
            gC.alloc((size_t)F * D, tot); gN.alloc((size_t)F * D, tot);
            grad.alloc((size_t)V * D, tot);
            // the value on FP64 matrix cores (vfd_rows_mfma_k) in the mixed
            // form; SA_VFD_MFMA=0 / 1 forces it off / on in either form
            const char *e = std::getenv("SA_VFD_MFMA");
            vfdMfma = e && (e[0] == '0' || e[0] == '1') ? e[0] == '1' : precision == SA_PREC_MIXED;
        }
        size_t n = (size_t)V * D;
        lmk.alloc(n * T, tot);
        tau.alloc((size_t)V * T, tot);
        for (int k = 0; k < 2; ++k) { geo[k].alloc((size_t)G::N * E, tot); ppd[k].alloc(Z, tot); }
        femElm.alloc((size_t)(1 + NS) * E, tot);
        lhs.alloc(Z, tot);
        ya.alloc(V, tot); ra.alloc(V, tot);
        f.alloc(n, tot); x.alloc(n, tot); a.alloc(n, tot); r.alloc(n, tot); Kid.alloc(n, tot);
        if (NC == 0) r2.alloc(n, tot);
        Kr.alloc(n, tot); d.alloc(n, tot); Ad.alloc(n, tot);
        frc.alloc((size_t)(D + 1) * D * E, tot);
        if (B > 0) { pfrc.alloc((size_t)D * D * B, tot); meas.alloc(B, tot); }
        rdRhs.alloc(V, tot); rdR.alloc(V, tot); rdZ.alloc(V, tot); rdD.alloc(V, tot); rdAd.alloc(V, tot);
        rdD2.alloc(V, tot);
        P0.alloc(MAXP, tot); P1.alloc(MAXP, tot); P2.alloc(2 * MAXP, tot); P3.alloc(3 * MAXP, tot);
        hist.alloc((size_t)std::max(itrMax, 1) + 2, tot);
        rdHist.alloc(200002, tot);
        scal.alloc(8, tot);
        ctl.alloc(1, tot); rdCtl.alloc(1, tot);
        if (NC > 0) {
            size_t nc = (size_t)NC * D;
            ctlIdx.alloc(NC, tot); ctlIdx.upload(hCtl.data(), NC, s);
            qc.alloc(nc, tot);
            if (!jacobi) Kcc.alloc((size_t)NC * NC, tot);
            cA.alloc(nc, tot); cA0.alloc(nc, tot); cAm1.alloc(nc, tot); cAm2.alloc(nc, tot); cAm3.alloc(nc, tot);
            cG.alloc(nc, tot); cR.alloc(nc, tot); cZ.alloc(nc, tot); cD.alloc(nc, tot);
            cR2.alloc(nc, tot); cD2.alloc(nc, tot);
            cS.alloc(nc, tot); cY.alloc(nc, tot); cK.alloc(nc, tot);
            cRout.alloc(nc, tot); cDelta.alloc(nc, tot); cZero.alloc(nc, tot);
            cZero.zero(s);
            {
                // (the spectral term corrects the X^T X preconditioner, not the Jacobi one)
                const char *e = std::getenv("SA_SPECTRAL");
                specEnabled = !(e && e[0] == '0') && !jacobi;
            }
            if (specEnabled) {
                recD.alloc(nc * SPEC_M, tot); recSD.alloc(nc * SPEC_M, tot); recR.alloc(nc * SPEC_M, tot);
                specU.alloc(nc * SPEC_K, tot); specC.alloc(SPEC_K, tot);
                specW.alloc((size_t)SPEC_K * gridOf((long)nc), tot);
                specA.alloc(6 * (size_t)SPEC_M * SPEC_M + 2 * SPEC_M, tot);
                specScr.alloc((size_t)(SPEC_CHUNKS + 1) * SPEC_M * SPEC_M, tot);
                specInfo.alloc(1, tot);
            }
            // stored kernel matrices: FP64 for the plain form, FP32 for the mixed one
            // dense products are panel-major with rows padded to whole panels
            // (zero padding rows); the preconditioner leaves rocBLAS
            // column-major (Mcol) and is repacked
            const size_t pc = padded_rows<double>(NC) * (size_t)NC;
            if (jacobi) {
                stiffBlk.alloc((size_t)D * NS * E, tot);
                bcc.alloc((size_t)D * Z, tot);
                pw.alloc(std::max(BL, 1), tot);
                dS.alloc(nc, tot);
            } else {
                Mcol.alloc((size_t)NC * NC, tot);
                Mpre.alloc(pc, tot);
                Mpre.zero(s);
            }
            if (precision == SA_PREC_MIXED && !SA_MIXED_CC32 && !jacobi) {
                // FP64 K_cc for the FP32 inner solve: its FP32 rounding need
                // not be positive definite when K_cc is ill conditioned
                Kc.alloc(pc, tot);
                Kc.zero(s);
            }
            if (precision == SA_PREC_MIXED && SA_MIXED_CC32 && !jacobi) {
                // FP32 copy of the preconditioner for the FP32 inner solve (a
                // step whose FP32 operator is unfit falls back to FP64)
                const size_t pf = padded_rows<float>(NC) * (size_t)NC;
                Mpre32.alloc(pf, tot);
                Mpre32.zero(s);
            }
            na32 = padded_rows<float>(V) * (size_t)NC;
            nt32 = padded_rows<float>(NC) * (size_t)V;
            na64 = padded_rows<double>(V) * (size_t)NC;
            nt64 = padded_rows<double>(NC) * (size_t)V;
            {
                // the matrix-core products (3D and Gaussian kernels) replace the
                // stored FP32 matrices when those would pass 48 GB, or with
                // SA_KX_MFMA=1.  Measured on the 51,210-landmark shell with
                // 5,000 control points: 175 us per product against 179 us for the
                // stored FP32 product, but 3e-7 relative error against 2.5e-8,
                // which doubles the inner CG iterations; so the store is the default
                const char *e = std::getenv("SA_KX_MFMA");
                const bool big = (na32 + nt32) * sizeof(float) > ((size_t)48 << 30);
                kxMfma = precision == SA_PREC_MIXED && (D == 3 || knlOrder == -1) &&
                         (e && (e[0] == '0' || e[0] == '1') ? e[0] == '1' : big);
            }
            if (kxMfma) {
                // nothing stored: the inner products run on the matrix cores,
                // the refinement residuals and fallback steps are matrix-free FP64
                lmkPerm.alloc(V, tot);
                ctlPerm.alloc(NC, tot);
                const int rbA = gridOf(V, KXM_ROWS), rbT = gridOf(NC, KXM_ROWS);
                slicesAm = std::max(1, std::min(gridOf(NC, 256), (KXM_BLOCKS + rbA - 1) / rbA));
                slicesTm = std::max(1, std::min(gridOf(V, 256), (KXM_BLOCKS + rbT - 1) / rbT));
            } else if (precision == SA_PREC_MIXED) {
                // FP32 pair, plus the FP32 remainders when both fit in
                // SA_SPLIT_MAX_GB; the quartet then also holds the FP64 pair
                // of fallback steps (padded to 256 rows it is at least as
                // large as the FP64 pair padded to 128)
                alias64 = SA_SPLIT_RESID && !matrixFree &&
                          (na32 + nt32) * sizeof(float) <= ((size_t)SA_SPLIT_MAX_GB << 30);
                const size_t nf = (alias64 ? 2 : 1) * (na32 + nt32);
                kmem.alloc((nf + 1) / 2, tot);
                kmem.zero(s);
                if (kmem.p) set_fp32_views();
            } else if ((na64 + nt64) * sizeof(double) <= (size_t)48 << 30 && !matrixFree) {
                kmem.alloc(na64 + nt64, tot);
                kmem.zero(s);
                A64.p = kmem.p;
                At64.p = kmem.p ? kmem.p + na64 : nullptr;
            }
            // blocks per stored product (row blocks x column slices) from the
            // tile table, by default KS_BLOCKS, and at least enough slices
            // for the LDS copy of x (KS_MAX_COLS columns)
            const bool f32 = precision == SA_PREC_MIXED;
            const int prS = f32 ? panel_rows<float>() : panel_rows<double>();
            const int rbA = gridOf(V, prS), rbT = gridOf(NC, prS), rbM = gridOf(NC, panel_rows<double>());
            const int dflt = KS_BLOCKS;
            const int tgA = ks_target("xc", f32, V, NC, dflt), tgT = ks_target("cx", f32, NC, V, dflt);
            const int tgM = ks_target("cc", false, NC, NC, KS_BLOCKS), tgM32 = ks_target("cc", true, NC, NC, KS_BLOCKS);
            const int rbM32 = gridOf(NC, panel_rows<float>());
            slicesA = std::max(gridOf(NC, KS_MAX_COLS), std::min(64, (tgA + rbA - 1) / rbA));
            slicesT = std::max(gridOf(V, KS_MAX_COLS), std::min(256, (tgT + rbT - 1) / rbT));
            slicesM = std::max(gridOf(NC, KS_MAX_COLS), std::min(256, (tgM + rbM - 1) / rbM));
            slicesM32 = std::max(gridOf(NC, KS_MAX_COLS), std::min(256, (tgM32 + rbM32 - 1) / rbM32));
            // (twice the product slices for the split residual products)
            partS.alloc(std::max({kxMfma ? (size_t)slicesAm * D * V : (size_t)2 * slicesA * D * V,
                                  kxMfma ? (size_t)slicesTm * nc : (size_t)2 * slicesT * nc, (size_t)slicesM * nc,
                                  (size_t)slicesM32 * nc}), tot);
            int rb = gridOf(NC, MV_ROWS);
            slicesC = std::max(1, std::min(32, (1024 + rb - 1) / rb));
            part.alloc((size_t)slicesC * nc, tot);
            potrfInfo.alloc(2, tot);
            potrfInfo.zero(s);
            {
                const char *e = std::getenv("SA_CHOL");
                cholLib = e && std::string(e) == "rocsolver";
            }
            require(gridOf((long)nc, 64) <= MAXP, "too many control points for the partial-sum buffers");
            if (!dryRun && !jacobi) {
                std::lock_guard<std::mutex> lk(g_handleMu);
                if (rocblas_create_handle(&blas) != rocblas_status_success)
                    throw SaError(SA_E_DEVICE, "rocblas_create_handle failed");
                if (rocblas_set_stream(blas, s) != rocblas_status_success)
                    throw SaError(SA_E_DEVICE, "rocblas_set_stream failed");
                // deterministic factorisation and products: no atomics in rocBLAS/rocSOLVER kernels
                if (rocblas_set_atomics_mode(blas, rocblas_atomics_not_allowed) != rocblas_status_success)
                    throw SaError(SA_E_DEVICE, "rocblas_set_atomics_mode failed");
            }
        }
        if (NC == 0) {
            aM1.alloc(n, tot); aM2.alloc(n, tot); aM3.alloc(n, tot);
            // all-landmark form: K stored per step when it fits (21 GB at V = 51,210)
            // (matrix-free kmatvec_k products when it does not fit, with
            // SA_MATRIX_FREE=1, or when the device refuses the allocation)
            const size_t nk = padded_rows<double>(V) * (size_t)V;
            if (nk * sizeof(double) <= (size_t)48 << 30 && !matrixFree) {
                try {
                    Kst.alloc(nk, tot);
                } catch (const SaError &e) {
                    if (e.code != SA_E_ALLOC) throw;
                    Kst.n = 0;
                }
                if (Kst.n) {
                    Kst.zero(s);
                    const int rbK = gridOf(V, panel_rows<double>());
                    const int tgK = ks_target("kk", false, V, V, 2048);
                    slicesK = std::max(gridOf(V, KS_MAX_COLS), std::min(256, (tgK + rbK - 1) / rbK));
                    partS.alloc((size_t)slicesK * D * V, tot);
                }
            }
        }
        require(gridOf(n) <= MAXP && gridOf(V, MV_ROWS) <= MAXP, "mesh too large for the partial-sum buffers");
        if (!dryRun && kxMfma) {
            // Morton order of the initial positions (the points move little
            // over a solve, so consecutive rows stay close together)
            std::vector<double> hc((size_t)NC * D);
            for (int c = 0; c < D; ++c)
                for (int k = 0; k < NC; ++k) hc[(size_t)c * NC + k] = hq0[(size_t)c * V + hCtl[k]];
            const auto pl = morton_order(hq0.data(), V, D), pc = morton_order(hc.data(), NC, D);
            lmkPerm.upload(pl.data(), pl.size(), s);
            ctlPerm.upload(pc.data(), pc.size(), s);
            CK(hipStreamSynchronize(s));
        }
        if (!dryRun) CK(hipStreamSynchronize(s));   // host copies (hq0) go out of scope
    }

    // ---------------- kernel-order dispatch
    template <bool CHECK>
    void kmatvec(const double *qq, const double *xin, double *yout, const double *Prr, int nPrr, int it) {
        if (Kst.p) {
            // stored K (built from qq by kst_build this step): HBM-bound product
            dim3 g(gridOf(V, panel_rows<double>()), slicesK);
            kstore_mv_k<D, double><<<g, NT, ks_lds<double>(V, slicesK), s>>>(V, V, Kst.p, xin, partS.p, ctl.p);
            CHECK_LAUNCH();
            kfin_k<D, CHECK><<<gridOf(V), NT, 0, s>>>(V, slicesK, partS.p, xin, yout, knlEps, Prr, nPrr, P3.p, it,
                                                      ctl.p);
            CHECK_LAUNCH();
            return;
        }
        double invW = 1.0 / knlWidth;
        int g = gridOf(V, MV_ROWS);
        Ctl *c = ctl.p;
#define SA_MV(O) kmatvec_k<D, O, CHECK><<<g, NT, 0, s>>>(V, qq, xin, yout, knlEps, invW, Prr, nPrr, P3.p, it, c)
        switch (knlOrder) {
            case -1: SA_MV(-1); break;
            case 0: SA_MV(0); break;
            case 1: SA_MV(1); break;
            case 2: SA_MV(2); break;
            case 3: SA_MV(3); break;
            default: SA_MV(4); break;
        }
#undef SA_MV
        CHECK_LAUNCH();
    }

    // partial sums the kernel product leaves in P3 (consumed by pcg_dir_k)
    int kmv_parts() const { return Kst.p ? gridOf(V) : gridOf(V, MV_ROWS); }

    // the all-landmark CG iteration in three launches (step_product_dir):
    // stored K whose per-block partials fit P3
    bool step_fused() const {
        return Kst.p && r2.p && (long)gridOf(V, panel_rows<double>()) * slicesK <= 3L * MAXP;
    }

    // iteration it of the all-landmark CG after the operator apply, in two
    // launches: the stored-K product takes pcg_update_k's step while staging
    // its input (r_(it+1) = r_it - st Ad, x and alpha updated; the residual
    // alternates between r (even iterations) and r2 (odd)), then
    // pcg_dir_stored_k
    void step_product_dir(int it, hipEvent_t *ev) {
        const int n = V * D, gN = gridOf(n);
        double *rIn = (it & 1) ? r2.p : r.p, *rOut = (it & 1) ? r.p : r2.p;
        const dim3 g(gridOf(V, panel_rows<double>()), slicesK);
        KsUpd upd;
        upd.P = P1.p; upd.nP = op_parts(); upd.hist = hist.p; upd.it = it;
        upd.rIn = rIn; upd.Ad = Ad.p; upd.d = d.p; upd.Kid = Kid.p;
        upd.rOut = rOut; upd.xs = x.p; upd.as = a.p; upd.Prr = P2.p; upd.stop = ctl.p;
        if (ev) CK(hipEventRecord(ev[0], s));
        // a stored K of at most SA_KST_CACHED_MB is read with ordinary loads:
        // it stays in the last-level cache from one iteration to the next
        const bool cached = (double)V * V * sizeof(double) <= SA_KST_CACHED_MB * 1048576.0;
#define SA_KSTEP(NTL)                                                                                    \
    kstore_mv_k<D, double, true, NTL><<<g, NT, ks_lds<double>(V, slicesK), s>>>(                         \
        V, V, Kst.p, rIn, partS.p, ctl.p, nullptr, nullptr, nullptr, 0, nullptr, 0, nullptr, nullptr, P3.p, upd)
        if (cached) SA_KSTEP(false);
        else SA_KSTEP(true);
#undef SA_KSTEP
        CHECK_LAUNCH();
        if (ev) CK(hipEventRecord(ev[1], s));
        pcg_dir_stored_k<<<gN, NT, 0, s>>>(n, slicesK, partS.p, knlEps, P2.p, (int)g.y, P3.p, (int)(g.x * g.y),
                                           hist.p, it, rOut, Kid.p, d.p, ctl.p);
        CHECK_LAUNCH();
    }

    // the preconditioning half of an all-landmark CG iteration: K r (|r|^2
    // partials in P2, nPrr of them, for the stopping test), then the new
    // direction.  Stored K: the product leaves per-block partials of
    // r.(K r) and pcg_dir_stored_k forms (K + eps I) r itself (no kfin_k
    // launch); matrix-free: kmatvec_k, then pcg_dir_k.  ev: events around the
    // product (kernel timer).
    void precond_dir(const double *q, int nPrr, int it, hipEvent_t *ev) {
        const int n = V * D, gN = gridOf(n);
        if (ev) CK(hipEventRecord(ev[0], s));
        const dim3 g(gridOf(V, panel_rows<double>()), Kst.p ? slicesK : 1);
        if (Kst.p && (long)g.x * g.y <= 3L * MAXP) {
            kstore_mv_k<D, double><<<g, NT, ks_lds<double>(V, slicesK), s>>>(
                V, V, Kst.p, r.p, partS.p, ctl.p, nullptr, nullptr, nullptr, 0, nullptr, 0, nullptr, r.p, P3.p);
            CHECK_LAUNCH();
            if (ev) CK(hipEventRecord(ev[1], s));
            pcg_dir_stored_k<<<gN, NT, 0, s>>>(n, slicesK, partS.p, knlEps, P2.p, nPrr, P3.p, (int)(g.x * g.y),
                                               hist.p, it, r.p, Kid.p, d.p, ctl.p);
            CHECK_LAUNCH();
            return;
        }
        kmatvec<true>(q, r.p, Kr.p, P2.p, nPrr, it);
        if (ev) CK(hipEventRecord(ev[1], s));
        pcg_dir_k<<<gN, NT, 0, s>>>(n, P3.p, kmv_parts(), hist.p, it, r.p, Kr.p, Kid.p, d.p, ctl.p);
        CHECK_LAUNCH();
    }

    // stored K = k(|q_i - q_j|) (without eps) for the all-landmark form
    void kst_build(const double *qq) {
        const double invW = 1.0 / knlWidth;
        dim3 g(gridOf(V, 64), gridOf(V, 64));
#define SA_KS(O) kstore_build_k<D, O, double><<<g, NT, 0, s>>>(V, qq, V, qq, invW, Kst.p, nullptr)
        switch (knlOrder) {
            case -1: SA_KS(-1); break;
            case 0: SA_KS(0); break;
            case 1: SA_KS(1); break;
            case 2: SA_KS(2); break;
            case 3: SA_KS(3); break;
            default: SA_KS(4); break;
        }
#undef SA_KS
        CHECK_LAUNCH();
    }

    // external yank f (and its |f|^2 partials in P) at positions q: element
    // forces, plus the boundary forces when bdrYank is on
    void yank(const double *q, const double *geoNow, double *fOut, double *P) {
        yank_elem_k<D><<<gridOf(E), NT, 0, s>>>(E, elmVtx.p, geoNow, ya.p, frc.p);
        CHECK_LAUNCH();
        if (bdrYank) {
            bdr_yank_k<D><<<gridOf(Bn), NT, 0, s>>>(Bn, V, bdrVtx.p, q, ya.p, bfrc.p);
            CHECK_LAUNCH();
        }
        gather_k<D><<<gridOf(V), NT, 0, s>>>(V, E, vtxElm.p, frc.p, fOut, P, bdrYank ? vtxBdr.p : nullptr, bfrc.p,
                                             Bn);
        CHECK_LAUNCH();
    }

    // Vertex chunks of the fused operator apply: vertices in Morton order of
    // q0, taken greedily while a chunk has at most nt vertices and maxs list
    // entries (the launch shape); per chunk, each element its vertices' lists name once,
    // with the LDS slot of each of its listed vertices.  The buffers are
    // sized from the tables alone (at most one element entry per list entry;
    // chunk count bounded below), so the sizing pass, which has no positions,
    // counts the same bytes.
    void build_op_chunks(const std::vector<int> &hVE, const std::vector<int> &hElm, const std::vector<double> &hq,
                         bool dryRun) {
        const char *env = std::getenv("SA_OP_FUSED");
        if (env && env[0] == '0') return;
        opCfg = opf_cfg(E, V);
        const int nt = OPF_SHAPES[opCfg].nt, maxs = OPF_SHAPES[opCfg].maxs, hm = OPF_SHAPES[opCfg].halo;
        require(hm <= 0xffff && maxs < 0x7fff, "launch shape too large for the packed halo/slot words");
        int maxAdj = 0;
        long S = 0;
        for (int i = 0; i < V; ++i) { maxAdj = std::max(maxAdj, hVE[i]); S += hVE[i]; }
        // (a vertex's elements then have at most (D + 1) maxAdj <= hm vertices)
        if (V == 0 || maxAdj > maxs || (long)(D + 1) * maxAdj > hm || S > INT32_MAX / (D + 1)) return;
        // a chunk is closed only when the next vertex v would pass nt
        // vertices, maxs entries or hm halo vertices.  A chunk's halo has at
        // most D + 1 vertices per list entry, so in the last two cases the
        // closed chunk and the next (which starts with v) hold together more
        // than min(maxs, hm / (D + 1)) entries; each chunk is in at most two
        // such pairs, so there are fewer than 2 S / min(...) + V / nt + 1
        // chunks
        const long capCh = V / nt + 2 * (S / maxs) + 2 * ((long)(D + 1) * S / hm) + 6;
        // the all-landmark CG consumes one partial per chunk
        if (NC == 0 && capCh > MAXP) return;
        opCapU = (int)std::max(S, 1L);
        const long capH = std::min((long)(D + 1) * S, capCh * (long)hm);
        long long &tot = bytes;
        opChV.alloc(capCh + 1, tot); opChE.alloc(capCh + 1, tot); opChH.alloc(capCh + 1, tot);
        opChVtx.alloc(V, tot); opChLs.alloc(V, tot); opChHv.alloc(std::max(capH, 1L), tot);
        opChElm.alloc(opCapU, tot);
        opChLS.alloc((size_t)(D + 1) * opCapU, tot);
        if (elaType == SA_ELA_FRAME)
            for (auto &b : opFr) b.alloc((size_t)2 * D * opCapU, tot);
        if (dryRun) { opFused = true; return; }
        const OpTables t = op_tables(hVE, hElm, V, E, D, morton_order(hq.data(), V, D), nt, maxs, hm);
        if (t.chVtx.empty()) {
            // not reached (the halo bound above); keeps the two-launch apply
            for (DBuf<int> *b : {&opChV, &opChE, &opChH, &opChVtx, &opChLs, &opChHv, &opChElm, &opChLS})
                b->alloc(0, tot);
            for (auto &b : opFr) b.alloc(0, tot);
            return;
        }
        check_op_tables(t, hVE, hElm, V, E, D, nt, maxs, hm);
        opChunks = (int)t.chV.size() - 1;
        opNU = (int)t.chElm.size();
        require(opChunks <= capCh && opNU <= opCapU && (long)t.chHv.size() <= capH,
                "operator chunk tables exceed their bound");
        opChV.upload(t.chV.data(), t.chV.size(), s); opChE.upload(t.chE.data(), t.chE.size(), s);
        opChH.upload(t.chH.data(), t.chH.size(), s);
        opChVtx.upload(t.chVtx.data(), t.chVtx.size(), s); opChLs.upload(t.chLs.data(), t.chLs.size(), s);
        opChHv.upload(t.chHv.data(), t.chHv.size(), s);
        opChElm.upload(t.chElm.data(), t.chElm.size(), s);
        // halo position | (LDS slot + 1) << 16 (halo < 2^16, slot + 1 < 2^15)
        std::vector<int> ls(t.slot.size());
        for (size_t a = 0; a < ls.size(); ++a) ls[a] = t.loc[a] | ((t.slot[a] + 1) << 16);
        opChLS.upload(ls.data(), ls.size(), s);
        CK(hipStreamSynchronize(s));   // the host vectors go out of scope
        opFused = true;
    }

    // partials the operator apply leaves for the all-landmark CG (MODE 0, 1)
    int op_parts() const { return opFused ? opChunks : gridOf(V); }

    // after geometry_k wrote geo[g]: its frames in element-entry order
    void op_frames(int g) {
        if (!opFused || !opFr[g].p || opNU == 0) return;
        op_frames_k<D><<<gridOf(opNU), NT, 0, s>>>(opNU, E, opChElm.p, geo[g].p, opFr[g].p);
        CHECK_LAUNCH();
    }

    // y = (E + btmWgt P) x in the CG mode MODE of gather_op_k (positions q,
    // concentration tauNow, geometry geoNow of q; pen = false: E x alone)
    template <int MODE>
    void op_apply(const double *q, const double *xin, const double *tauNow, const double *geoNow, const double *u,
                  const double *w, double *out, double *out2, double *P, bool pen = true) {
        if (!opFused) {
            elastic(xin, tauNow, geoNow, ctl.p, pen);
            gather_op<MODE>(u, w, out, out2, P, pen);
            return;
        }
        const bool pn = pen && btmWgt != 0.0 && B > 0;
        if (pn) {
            pen_elem_k<D><<<gridOf(B), NT, 0, s>>>(B, V, btmVtx.p, meas.p, xin, pfrc.p, ctl.p);
            CHECK_LAUNCH();
        }
        require(geoNow == geo[0].p || geoNow == geo[1].p, "operator apply on a geometry without entry frames");
        const double *frE = geoNow == geo[1].p ? opFr[1].p : opFr[0].p;
#define SA_OPF(ISO, BT, MS, HM) op_fused_k<D, ISO, MODE, BT, MS, HM><<<opChunks, BT, 0, s>>>(                 \
            opChV.p, opChVtx.p, opChLs.p, opChH.p, opChHv.p, opChE.p, opChElm.p, opChLS.p, opNU, V, vtxStart.p,   \
            q, frE, xin, tauNow, elmLyr.p, dmod.p, BL, B, vtxBtm.p, pn ? pfrc.p : nullptr, btmWgt, ldmWgt, u, w,  \
            out, out2, P, ctl.p)
#define SA_OPF_SHAPE(BT, MS, HM) \
        if (elaType == SA_ELA_FRAME) SA_OPF(false, BT, MS, HM); else SA_OPF(true, BT, MS, HM)
        static_assert(OPF_NSHAPES == 4, "one case per launch shape");
        switch (opCfg) {
            case 0: SA_OPF_SHAPE(256, 1024, 512); break;
            case 1: SA_OPF_SHAPE(512, 1024, 512); break;
            case 2: SA_OPF_SHAPE(512, 2048, 1024); break;
            default: SA_OPF_SHAPE(1024, 2048, 1024); break;
        }
#undef SA_OPF_SHAPE
#undef SA_OPF
        CHECK_LAUNCH();
    }

    // element pass of E x with the geometry geoNow; the forces land in vertex
    // order when the slot map exists
    void elastic(const double *xin, const double *tauNow, const double *geoNow, const Ctl *c, bool pen = true) {
        int g = gridOf(E);
        if (elaType == SA_ELA_FRAME)
            elastic_elem_k<D, false, double><<<g, NT, 0, s>>>(E, V, elmVtx.p, geoNow, xin, tauNow, elmLyr.p, dmod.p,
                                                              frc.p, c, elmSlot.p);
        else
            elastic_elem_k<D, true, double><<<g, NT, 0, s>>>(E, V, elmVtx.p, geoNow, xin, tauNow, elmLyr.p, dmod.p,
                                                             frc.p, c, elmSlot.p);
        CHECK_LAUNCH();
        if (pen && btmWgt != 0.0 && B > 0) {
            pen_elem_k<D><<<gridOf(B), NT, 0, s>>>(B, V, btmVtx.p, meas.p, xin, pfrc.p, c);
            CHECK_LAUNCH();
        }
    }

    template <int MODE>
    void gather_op(const double *u, const double *w, double *out, double *out2, double *P, bool pen = true) {
        const double *pf = (pen && btmWgt != 0.0 && B > 0) ? pfrc.p : nullptr;
        gather_op_k<D, MODE><<<gridOf(V), NT, 0, s>>>(V, E, vtxElm.p, frc.p, BL, B, vtxBtm.p, pf, btmWgt, ldmWgt, u, w,
                                                     out, out2, P, ctl.p, vtxStart.p);
        CHECK_LAUNCH();
    }

    void sync_ctl(Ctl *dev, Ctl *host) {
        CK(hipMemcpyAsync(host, dev, sizeof(Ctl), hipMemcpyDeviceToHost, s));
        CK(hipStreamSynchronize(s));
        gate_yield();
    }

    // velocity solve of one step; returns 0 ok, 1 failure
    int velocity(const double *q, const double *tauNow, const double *geoNow, bool warm, int &iters) {
        const int n = V * D;
        const int gV = gridOf(V), gN = gridOf(n);
        if (Kst.p) kst_build(q);
        // f and |f|^2
        yank(q, geoNow, f.p, P0.p);
        pcg_start_k<<<1, NT, 0, s>>>(P0.p, gV, cgTolSqu, itrMax, ctl.p);
        CHECK_LAUNCH();
        zero_if_empty_k<<<gN, NT, 0, s>>>(n, x.p, a.p, ctl.p);
        CHECK_LAUNCH();
        if (warm) kmatvec<false>(q, a.p, x.p, nullptr, 0, -1);   // x0 = K alpha_{t-1}
        // r = b - A x0, Kid = r
        op_apply<1>(q, x.p, tauNow, geoNow, a.p, f.p, r.p, Kid.p, P2.p);
        precond_dir(q, op_parts(), -1, nullptr);
        int it0 = 0, chunk = 4;
        Ctl &h = hctl[0];
        while (true) {
            int itEnd = std::min(itrMax, it0 + chunk);
            if (timer.on) timer.ensure(itEnd - it0);
            for (int it = it0; it < itEnd; ++it) {
                // the kernel timer samples every TIMER_EVERY-th iteration
                hipEvent_t *ev = timer.on && it % TIMER_EVERY == 0 ? &timer.ev[6 * (it - it0)] : nullptr;
                if (ev) CK(hipEventRecord(ev[2], s));
                op_apply<0>(q, d.p, tauNow, geoNow, Kid.p, d.p, Ad.p, nullptr, P1.p);
                if (ev) CK(hipEventRecord(ev[3], s));
                if (step_fused()) {
                    step_product_dir(it, ev);
                } else {
                    pcg_update_k<<<gN, NT, 0, s>>>(n, P1.p, op_parts(), hist.p, it, x.p, a.p, r.p, d.p, Kid.p,
                                                   Ad.p, P2.p, ctl.p);
                    CHECK_LAUNCH();
                    precond_dir(q, gN, it, ev);
                }
            }
            sync_ctl(ctl.p, &h);
            if (timer.on) {
                // iteration it did the operator apply iff it < iters (or not yet done);
                // its product ran iff it + 1 < iters
                int lastOp = h.done ? h.iters : itEnd;
                for (int it = it0; it < itEnd; ++it) {
                    if (it % TIMER_EVERY) continue;
                    hipEvent_t *ev = &timer.ev[6 * (it - it0)];
                    float ms;
                    if (it < lastOp) {
                        CK(hipEventElapsedTime(&ms, ev[2], ev[3]));
                        prof->opMs += ms; prof->opCount++;
                    }
                    if (!h.done || it + 1 < h.iters) {
                        CK(hipEventElapsedTime(&ms, ev[0], ev[1]));
                        prof->kmvMs += ms; prof->kmvCount++;
                    }
                }
            }
            if (h.done) break;
            if (itEnd >= itrMax) { iters = itrMax; failReason = SA_FAIL_BUDGET; return 1; }
            it0 = itEnd;
            chunk = std::min(chunk * 2, 64);
        }
        iters = h.iters;
        return h.status;
    }

    // ---------------- control-point form
    // y = K(a_i, b_j) x over nb columns, split into `slices` column slices
    void kcross(int na, const double *A, int nb, const double *Bp, const double *xin, double *yout, int slices,
                double *P, const Ctl *c) {
        double invW = 1.0 / knlWidth;
        dim3 g(gridOf(na, MV_ROWS), slices);
        const bool viaPart = slices > 1 || P != nullptr;
        double *dst = viaPart ? part.p : yout;
#define SA_KC(O) kcross_k<D, O><<<g, NT, 0, s>>>(na, A, nb, Bp, xin, dst, invW, c)
        switch (knlOrder) {
            case -1: SA_KC(-1); break;
            case 0: SA_KC(0); break;
            case 1: SA_KC(1); break;
            case 2: SA_KC(2); break;
            case 3: SA_KC(3); break;
            default: SA_KC(4); break;
        }
#undef SA_KC
        CHECK_LAUNCH();
        if (viaPart) {
            slice_sum_k<<<gridOf((long)D * na), NT, 0, s>>>(D * na, slices, part.p, yout, P, c);
            CHECK_LAUNCH();
        }
    }

    // K_cc (+ eps I) and, when refresh is set, its preconditioner M = X^T X,
    // X = L^{-1} the inverted Cholesky factor (rocSOLVER potrf + trtri,
    // rocBLAS dgemm); Kc (FP32 inner solve only) keeps K_cc without eps for
    // the products K_cc u and is rebuilt every step
    void kcc_build(bool refresh) {
        kcc_build_once(refresh);
        if (!refresh) return;
        if (prof) prof->factorizations++;
        struct CarryDone {
            Engine *e;
            ~CarryDone() { e->specCarry = false; }
        } carryDone{this};
        if (specCarry) spec_carry_finish();
    }

    // diag(S) of the control-point system at positions q (the control points
    // already gathered into qc), concentration tauNow and geometry geoNow:
    // the Jacobi preconditioner, rebuilt every step
    void jacobi_build(const double *q, const double *tauNow, const double *geoNow) {
        if (elaType == SA_ELA_FRAME)
            stiff_blocks_k<D, false><<<gridOf(E), NT, 0, s>>>(E, elmVtx.p, geoNow, tauNow, elmLyr.p, dmod.p, stiffBlk.p);
        else
            stiff_blocks_k<D, true><<<gridOf(E), NT, 0, s>>>(E, elmVtx.p, geoNow, tauNow, elmLyr.p, dmod.p, stiffBlk.p);
        stiff_gather_k<D><<<gridOf(Z), NT, 0, s>>>(Z, E, femIfo.p, stiffBlk.p, bcc.p);
        const bool pen = btmWgt != 0.0 && B > 0;
        if (pen) pen_weight_k<D><<<gridOf(BL), NT, 0, s>>>(BL, vtxBtm.p, meas.p, pw.p);
        CHECK_LAUNCH();
        const double invW = 1.0 / knlWidth;
#define SA_JD(O) jacobi_diag_k<D, O><<<NC, NT, 0, s>>>(V, NC, q, qc.p, rowPtr.p, colIdx.p, Z, bcc.p, pen ? BL : 0, \
                                                       pw.p, btmWgt, ldmWgt, knlEps, invW, dS.p)
        switch (knlOrder) {
            case -1: SA_JD(-1); break;
            case 0: SA_JD(0); break;
            case 1: SA_JD(1); break;
            case 2: SA_JD(2); break;
            case 3: SA_JD(3); break;
            default: SA_JD(4); break;
        }
#undef SA_JD
        CHECK_LAUNCH();
    }

    // U <- U G^-1/2 (G = Q g Q^T, the Gram matrix kcc_build_once formed):
    // the carried Ritz vectors become orthonormal in the new M^-1 inner
    // product, so M + U diag(cf) U^T stays positive definite (each cf in
    // (-1, 0)); the symmetric (Loewdin) choice keeps each vector closest to
    // the one it replaces, so its coefficient still fits.  A Gram matrix far
    // from the identity (the subspace no longer near the old one) drops the
    // term and records a new basis instead.
    void spec_carry_finish() {
        const int nc = NC * D, k = specK;
        const size_t kk = (size_t)k * k, MM = (size_t)SPEC_M * SPEC_M;
        double *dG = specA.p, *dW = dG + MM, *dEv = dG + 6 * MM, *dE = dEv + SPEC_M;
        std::vector<double> Q(kk), ev(k), W(kk, 0.0);
        int info = 0;
        {
            if (rocsolver_dsyevd(blas, rocblas_evect_original, rocblas_fill_lower, k, dG, k, dEv, dE, specInfo.p) !=
                rocblas_status_success)
                throw SaError(SA_E_DEVICE, "rocsolver_dsyevd failed");
            CK(hipStreamSynchronize(s));
        }
        CK(hipMemcpyAsync(Q.data(), dG, sizeof(double) * kk, hipMemcpyDeviceToHost, s));
        CK(hipMemcpyAsync(ev.data(), dEv, sizeof(double) * k, hipMemcpyDeviceToHost, s));
        CK(hipMemcpyAsync(&info, specInfo.p, sizeof(int), hipMemcpyDeviceToHost, s));
        CK(hipStreamSynchronize(s));
        bool ok = info == 0;
        for (int i = 0; i < k && ok; ++i) ok = std::isfinite(ev[i]) && ev[i] > SPEC_CARRY_GMIN && ev[i] < 1.0 / SPEC_CARRY_GMIN;
        for (size_t i = 0; i < kk && ok; ++i) ok = std::isfinite(Q[i]);
        if (!ok) {
            specOn = false;
            specRecord = specEnabled;
            return;
        }
        // W = Q g^-1/2 Q^T
        for (int l = 0; l < k; ++l) {
            const double f = 1.0 / std::sqrt(ev[l]);
            for (int j = 0; j < k; ++j) {
                const double qj = Q[j + (size_t)l * k] * f;
                for (int i = 0; i < k; ++i) W[i + (size_t)j * k] += Q[i + (size_t)l * k] * qj;
            }
        }
        CK(hipMemcpyAsync(dW, W.data(), sizeof(double) * kk, hipMemcpyHostToDevice, s));
        const double one = 1.0, zero = 0.0;
        {
            if (rocblas_dgemm(blas, rocblas_operation_none, rocblas_operation_none, nc, k, k, &one, specU.p, nc, dW, k,
                              &zero, recD.p, nc) != rocblas_status_success)
                throw SaError(SA_E_DEVICE, "rocblas_dgemm failed");
            CK(hipMemcpyAsync(specU.p, recD.p, sizeof(double) * nc * k, hipMemcpyDeviceToDevice, s));
            CK(hipStreamSynchronize(s));
        }
        ++specCarried;
    }

    // A = L L^T in place (lower triangle of the column-major n x n A, upper
    // not read); info[0] = 0, or the 1-based column of the first pivot that
    // is not positive
    void cholesky(double *A, int n, int *info) {
        zero_ints_k<<<1, 64, 0, s>>>(info, 1);
        for (int k0 = 0; k0 < n; k0 += CHB) {
            chol_diag_k<<<1, CHB, 0, s>>>(n, k0, A, n, info);
            const int rows = n - k0 - CHB;
            if (rows > 0) {
                chol_panel_k<<<gridOf(rows, 256), 256, 0, s>>>(n, k0, A, n, info);
                const int nt = gridOf(rows, CHB);
                chol_update_k<<<nt * (nt + 1) / 2, 256, 0, s>>>(n, k0, A, n, info);
            }
        }
        CHECK_LAUNCH();
    }

    void kcc_build_once(bool refresh) {
        if (!refresh && !Kc.p) return;
        double invW = 1.0 / knlWidth;
        long nn = (long)NC * NC;
        int g = gridOf(nn);
        double *K = refresh ? Kcc.p : nullptr;
#define SA_KB(O) kcc_build_k<D, O><<<g, NT, 0, s>>>(NC, qc.p, invW, knlEps, K, Kc.p)
        switch (knlOrder) {
            case -1: SA_KB(-1); break;
            case 0: SA_KB(0); break;
            case 1: SA_KB(1); break;
            case 2: SA_KB(2); break;
            case 3: SA_KB(3); break;
            default: SA_KB(4); break;
        }
#undef SA_KB
        CHECK_LAUNCH();
        if (!refresh) return;
        if (specCarry) {
            // the spectral term carried over to the new M: its Gram matrix in
            // the new M^-1 = K_cc + eps I inner product, G = U^T (K_cc + eps I) U,
            // while Kcc still holds K_cc + eps I (spec_carry_finish
            // re-orthonormalises U with it)
            const int nc = NC * D, kd = specK * D;
            const double one = 1.0, zero = 0.0;
            if (rocblas_dgemm(blas, rocblas_operation_none, rocblas_operation_none, NC, kd, NC, &one, Kcc.p, NC,
                              specU.p, NC, &zero, recD.p, NC) != rocblas_status_success ||
                rocblas_dgemm(blas, rocblas_operation_transpose, rocblas_operation_none, specK, specK, nc, &one,
                              specU.p, nc, recD.p, nc, &zero, specA.p, specK) != rocblas_status_success)
                throw SaError(SA_E_DEVICE, "rocblas_dgemm failed");
        } else {
            // a new M: the spectral term (orthonormal in the old M's inner
            // product) is dropped and harvested again from the next solve
            specOn = false;
            specRecord = specEnabled;
        }
        if (cholLib) {
            // rocSOLVER's factorisation runs with no other session's work on the device
            GateExclusive lk;
            if (rocsolver_dpotrf(blas, rocblas_fill_lower, NC, Kcc.p, NC, potrfInfo.p) != rocblas_status_success)
                throw SaError(SA_E_DEVICE, "rocsolver_dpotrf failed");
            CK(hipStreamSynchronize(s));
        } else {
            cholesky(Kcc.p, NC, potrfInfo.p);
        }
        if (rocsolver_dtrtri(blas, rocblas_fill_lower, rocblas_diagonal_non_unit, NC, Kcc.p, NC, potrfInfo.p + 1) !=
            rocblas_status_success)
            throw SaError(SA_E_DEVICE, "rocsolver_dtrtri failed");
        zero_upper_k<<<g, NT, 0, s>>>(NC, Kcc.p);
        CHECK_LAUNCH();
        const double one = 1.0, zero = 0.0;
        // M = X^T X (lower triangle is all panelize_k reads)
#if SA_MPRE_BUILD == 2
        if (rocsolver_dlauum(blas, rocblas_fill_lower, NC, Kcc.p, NC) != rocblas_status_success)
            throw SaError(SA_E_DEVICE, "rocsolver_dlauum failed");
        const double *Mlow = Kcc.p;
#elif SA_MPRE_BUILD == 1
        if (rocblas_dsyrk(blas, rocblas_fill_lower, rocblas_operation_transpose, NC, NC, &one, Kcc.p, NC, &zero,
                          Mcol.p, NC) != rocblas_status_success)
            throw SaError(SA_E_DEVICE, "rocblas_dsyrk failed");
        const double *Mlow = Mcol.p;
#else
        if (rocblas_dgemm(blas, rocblas_operation_transpose, rocblas_operation_none, NC, NC, NC, &one, Kcc.p, NC,
                          Kcc.p, NC, &zero, Mcol.p, NC) != rocblas_status_success)
            throw SaError(SA_E_DEVICE, "rocblas_dgemm failed");
        const double *Mlow = Mcol.p;
#endif
        panelize_k<<<g, NT, 0, s>>>(NC, Mlow, NC, Mpre.p);
        CHECK_LAUNCH();
        if (Mpre32.p) {
            panel_to_float_k<<<g, NT, 0, s>>>(NC, Mpre.p, Mpre32.p);
            CHECK_LAUNCH();
        }
        CK(hipStreamSynchronize(s));
    }

    // dynamic LDS of kstore_mv_k
    template <typename T>
    static size_t ks_lds(int ncol, int slices) { return ks_lds_bytes<D, T>(ncol, slices); }

    template <typename T>
    static size_t padded_rows(int n) { return (size_t)gridOf(n, panel_rows<T>()) * panel_rows<T>(); }

    // kernel-matrix products of the control-point form: matrix-free FP64, or
    // the stored matrices (see kstore_build_k) in FP64 or FP32
    // Split32: the FP32 matrices plus their FP32 remainders, K = hi + lo to
    // about 2^-48 relative (residual products of the mixed form only)
    struct DirArgs {
        const double *z, *P;
        int nP;
        double *hist;
        int j;
        double *dout;
    };

    // Mfma32: matrix-free FP32 products on the matrix cores (kx_mfma_k)
    enum class KMode { MatrixFree64, Stored64, Stored32, Split32, Mfma32 };

    // column slices of the products of a mode (K_xc: xc = true)
    int kslices(KMode md, bool xc) const {
        if (md == KMode::Mfma32) return xc ? slicesAm : slicesTm;
        return xc ? slicesA : slicesT;
    }

    // y-slices = K(rows, columns) u on the matrix cores (partials in partS)
    template <int O>
    void kx_mfma_launch(bool xc, const double *q, const double *u, const Ctl *c, Ctl *adv, const DirArgs *dir) {
        if constexpr (D == 2 && O >= 0) {
            throw SaError(SA_E_ARG, "the 2D Matern kernels have no matrix-core product");
        } else {
            const int nr = xc ? V : NC, ncol = xc ? NC : V;
            dim3 g(gridOf(nr, KXM_ROWS), kslices(KMode::Mfma32, xc));
            kx_mfma_k<D, O><<<g, NT, 0, s>>>(nr, xc ? q : qc.p, xc ? lmkPerm.p : ctlPerm.p, ncol, xc ? qc.p : q,
                                              xc ? ctlPerm.p : lmkPerm.p, u, 1.0 / knlWidth, partS.p, c, adv,
                                              dir ? dir->z : nullptr, dir ? dir->P : nullptr, dir ? dir->nP : 0,
                                              dir ? dir->hist : nullptr, dir ? dir->j : 0, dir ? dir->dout : nullptr);
        }
    }

    // y[i] = sum of the slice partials part[k * n + i] (slice_rows_k; four
    // wavefronts per 64 rows once the rows alone give 1,024 blocks)
    void slice_rows(int n, int slices, const double *part, double *y, const Ctl *c) {
        if (gridOf(n, 64) >= 1024) slice_rows_k<4><<<gridOf(n, 64), 256, 0, s>>>(n, slices, part, y, c);
        else slice_rows_k<SLW><<<gridOf(n, 64), NTS, 0, s>>>(n, slices, part, y, c);
        CHECK_LAUNCH();
    }

    // y = M x for a dense FP64 C x C panel-major matrix
    void cc_product(const double *M, const double *x, double *y) {
        dim3 g(gridOf(NC, panel_rows<double>()), slicesM);
        kstore_mv_k<D, double><<<g, NT, ks_lds<double>(NC, slicesM), s>>>(NC, NC, M, x, partS.p, ctl.p);
        CHECK_LAUNCH();
        slice_rows(D * NC, slicesM, partS.p, y, ctl.p);
    }

    // z = M r, M = X^T X ~ K_cc^{-1} (its FP32 panel copy in the FP32 form),
    // fused with the CG test that follows: stop on |r|^2 (gC partials in P2)
    // under the tolerance, else the r.z partials into P3, gridOf(D NC, 64) of
    // them
    // the spectral term applies to this solve (harvested in the same mode,
    // not recording)
    bool spec_active(KMode md) const { return specOn && !recActive && specMode == (int)md; }

    // wpReady > 0: the U^T r partials (that many blocks) are already in specW
    // (ctl_update_k wrote them); else spec_wp_k computes them here
    void prec_dot(KMode md, const double *r, double *z, int wpReady = 0) {
        const int nc = NC * D;
        if (jacobi) {
            jacobi_dot_k<<<gridOf(nc, 64), 64, 0, s>>>(nc, P2.p, gridOf(nc), r, dS.p, z, P3.p, ctl.p);
            CHECK_LAUNCH();
            return;
        }
        const bool sp = spec_active(md);
        int nPw = wpReady;
        if (sp && !wpReady) {
            nPw = gridOf(nc, SPW_ROWS);
            spec_wp_k<<<dim3(nPw, gridOf(specK, SPW_COLS)), NT, 0, s>>>(nc, specU.p, specK, r, specW.p, ctl.p);
            CHECK_LAUNCH();
        }
        int sl;
        if ((md == KMode::Stored32 || md == KMode::Mfma32) && Mpre32.p) {
            sl = slicesM32;
            dim3 g(gridOf(NC, panel_rows<float>()), sl);
            kstore_mv_k<D, float, false, PRE_NT><<<g, NT, ks_lds<float>(NC, sl), s>>>(NC, NC, Mpre32.p, r, partS.p, ctl.p);
        } else {
            sl = slicesM;
            dim3 g(gridOf(NC, panel_rows<double>()), sl);
            kstore_mv_k<D, double, false, PRE_NT><<<g, NT, ks_lds<double>(NC, sl), s>>>(NC, NC, Mpre.p, r, partS.p, ctl.p);
        }
        CHECK_LAUNCH();
        slice_dot_k<<<gridOf(nc, 64), NTS, 0, s>>>(nc, sl, partS.p, P2.p, gridOf(nc), r, z, P3.p,
                                                  sp ? specU.p : nullptr, specC.p, specW.p, nPw, specK, ctl.p);
        CHECK_LAUNCH();
    }


    void kstore_build(const double *q, bool fp64, bool fp32) {
        double invW = 1.0 / knlWidth;
        dim3 g(gridOf(V, 64), gridOf(NC, 64));
        if (fp64 && A64.p) {
#define SA_KS(O) kstore_build_k<D, O, double><<<g, NT, 0, s>>>(V, q, NC, qc.p, invW, A64.p, At64.p)
            switch (knlOrder) {
                case -1: SA_KS(-1); break;
                case 0: SA_KS(0); break;
                case 1: SA_KS(1); break;
                case 2: SA_KS(2); break;
                case 3: SA_KS(3); break;
                default: SA_KS(4); break;
            }
#undef SA_KS
            CHECK_LAUNCH();
        }
        if (fp32 && A32.p) {
#define SA_KS(O) kstore_build_k<D, O, float><<<g, NT, 0, s>>>(V, q, NC, qc.p, invW, A32.p, At32.p, A32lo.p, At32lo.p)
            switch (knlOrder) {
                case -1: SA_KS(-1); break;
                case 0: SA_KS(0); break;
                case 1: SA_KS(1); break;
                case 2: SA_KS(2); break;
                case 3: SA_KS(3); break;
                default: SA_KS(4); break;
            }
#undef SA_KS
            CHECK_LAUNCH();
        }
    }

    // y = K_xc u (xc = true, V rows) or y = K_cx u (C rows) in the given mode;
    // |y|^2 partials in P when given.  ev (optional): events recorded around
    // the stored-matrix product kernel alone.
    // adv: control block whose iteration counter the stored product advances
    // (first kernel of a CG iteration); sum = false leaves the stored
    // product's slice partials in partS for a fused consumer.
    // dir: the product's input is the CG direction of iteration dir->j,
    // formed while staging it (kstore_mv_k's dz ... dout arguments)
    void kx_product(KMode md, bool xc, const double *q, const double *u, double *y, double *P, const Ctl *c,
                    hipEvent_t *ev = nullptr, Ctl *adv = nullptr, bool sum = true, const DirArgs *dir = nullptr) {
        if (md == KMode::MatrixFree64) {
            if (ev) CK(hipEventRecord(ev[0], s));
            if (xc) kcross(V, q, NC, qc.p, u, y, 1, P, c);
            else kcross(NC, qc.p, V, q, u, y, slicesC, P, c);
            if (ev) CK(hipEventRecord(ev[1], s));
            return;
        }
        const int nr = xc ? V : NC, ncol = xc ? NC : V;
        const int sl = kslices(md, xc);
        if (md == KMode::Mfma32) {
            if (ev) CK(hipEventRecord(ev[0], s));
            switch (knlOrder) {
                case -1: kx_mfma_launch<-1>(xc, q, u, c, adv, dir); break;
                case 0: kx_mfma_launch<0>(xc, q, u, c, adv, dir); break;
                case 1: kx_mfma_launch<1>(xc, q, u, c, adv, dir); break;
                case 2: kx_mfma_launch<2>(xc, q, u, c, adv, dir); break;
                case 3: kx_mfma_launch<3>(xc, q, u, c, adv, dir); break;
                default: kx_mfma_launch<4>(xc, q, u, c, adv, dir); break;
            }
            CHECK_LAUNCH();
            if (ev) CK(hipEventRecord(ev[1], s));
            if (!sum) return;
            if (P) slice_sum_k<<<gridOf((long)D * nr), NT, 0, s>>>(D * nr, sl, partS.p, y, P, c);
            else slice_rows(D * nr, sl, partS.p, y, c);
            CHECK_LAUNCH();
            return;
        }
        if (md == KMode::Split32) {
            // hi product into slices [0, sl), lo product into [sl, 2 sl), summed together
            dim3 g(gridOf(nr, panel_rows<float>()), sl);
            kstore_mv_k<D, float><<<g, NT, ks_lds<float>(ncol, sl), s>>>(nr, ncol, xc ? A32.p : At32.p, u, partS.p, c);
            kstore_mv_k<D, float><<<g, NT, ks_lds<float>(ncol, sl), s>>>(nr, ncol, xc ? A32lo.p : At32lo.p, u,
                                                                   partS.p + (size_t)sl * D * nr, c);
            CHECK_LAUNCH();
            if (P) slice_sum_k<<<gridOf((long)D * nr), NT, 0, s>>>(D * nr, 2 * sl, partS.p, y, P, c);
            else slice_rows(D * nr, 2 * sl, partS.p, y, c);
            CHECK_LAUNCH();
            return;
        }
        if (ev) CK(hipEventRecord(ev[0], s));
        if (md == KMode::Stored64) {
            dim3 g(gridOf(nr, panel_rows<double>()), sl);
            kstore_mv_k<D, double><<<g, NT, ks_lds<double>(ncol, sl), s>>>(
                nr, ncol, xc ? A64.p : At64.p, u, partS.p, c, adv, dir ? dir->z : nullptr, dir ? dir->P : nullptr,
                dir ? dir->nP : 0, dir ? dir->hist : nullptr, dir ? dir->j : 0, dir ? dir->dout : nullptr);
        } else {
            dim3 g(gridOf(nr, panel_rows<float>()), sl);
            kstore_mv_k<D, float><<<g, NT, ks_lds<float>(ncol, sl), s>>>(
                nr, ncol, xc ? A32.p : At32.p, u, partS.p, c, adv, dir ? dir->z : nullptr, dir ? dir->P : nullptr,
                dir ? dir->nP : 0, dir ? dir->hist : nullptr, dir ? dir->j : 0, dir ? dir->dout : nullptr);
        }
        CHECK_LAUNCH();
        if (ev) CK(hipEventRecord(ev[1], s));
        if (!sum) return;
        if (P) slice_sum_k<<<gridOf((long)D * nr), NT, 0, s>>>(D * nr, sl, partS.p, y, P, c);   // + |y|^2 partials
        else slice_rows(D * nr, sl, partS.p, y, c);
        CHECK_LAUNCH();
    }

    // S u = K_cx B K_xc u (-> cY); K_xc u stays in Kr.  Its control-point
    // rows are K_cc u, except in the FP32 form with an FP64 K_cc, which
    // takes K_cc u from that (-> cK).  ev: 6 events (product, operator,
    // product).  fused: S_apply opens a CG iteration whose stored products
    // feed slice_combine_k (see fused_iter); the K_cx product's partials stay
    // in partS.
    bool fused_iter(KMode md) const { return md != KMode::MatrixFree64 && !(md == KMode::Stored32 && Kc.p); }

    void S_apply(KMode md, const double *q, const double *tauNow, const double *geoNow, const double *u,
                 hipEvent_t *ev, bool fused = false, const DirArgs *dir = nullptr) {
        kx_product(md, true, q, u, Kr.p, nullptr, ctl.p, ev, fused ? ctl.p : nullptr, true, dir);   // w = K_xc u
        if (ev) CK(hipEventRecord(ev[2], s));
        op_apply<2>(q, Kr.p, tauNow, geoNow, nullptr, nullptr, Ad.p, nullptr, nullptr);   // B w
        if (ev) CK(hipEventRecord(ev[3], s));
        kx_product(md, false, q, Ad.p, cY.p, nullptr, ctl.p, ev ? ev + 4 : nullptr, nullptr, !fused);   // K_cx B w
        if (md == KMode::Stored32 && Kc.p) cc_product(Kc.p, u, cK.p);
    }

    // one control-point CG iteration on S x = rhs (x = xs): counter (the
    // iteration's first kernel advances it; the budget spent ends the solve),
    // S d, step, preconditioned residual, new direction; every kernel leaves
    // at once when the solve is done, so iterations queued past the end are
    // no-ops.  ev: optional per-iteration events (6).
    //
    // Fused form (stored products): iteration j keeps its direction and
    // residual in the buffers of parity j (cD/cR even, cD2/cR2 odd).  The
    // K_xc product forms d_j = z_j + beta d_(j-1) while staging it (the
    // start of the solve formed d_0), ctl_update_k writes r_(j+1) to the
    // other buffer and, with the spectral term, the partials of U^T r_(j+1);
    // no separate direction kernel.
    void cg_iter(KMode md, const double *q, const double *tauNow, const double *geoNow, double *xs, hipEvent_t *ev,
                 int j) {
        const int nc = NC * D, gC = gridOf(nc);
        if (fused_iter(md)) {
            double *dj = (j & 1) ? cD2.p : cD.p, *dprev = (j & 1) ? cD.p : cD2.p;
            double *rj = (j & 1) ? cR2.p : cR.p, *rnext = (j & 1) ? cR.p : cR2.p;
            DirArgs dir{cZ.p, P3.p, gridOf(nc, 64), hist.p, j, dj};
            S_apply(md, q, tauNow, geoNow, j > 0 ? dprev : dj, ev, true, j > 0 ? &dir : nullptr);
            const int gS = gridOf(nc, 64);   // d.Sd partials
            slice_combine_k<D><<<gS, NTS, 0, s>>>(NC, V, kslices(md, false), partS.p, Kr.p, ctlIdx.p, dj, ldmWgt,
                                                 knlEps, cS.p,
                                                 P1.p, recActive ? recD.p : nullptr, recSD.p, recR.p, rj, SPEC_M,
                                                 ctl.p);
            const bool sp = spec_active(md);
            ctl_update_k<<<dim3(gC, sp ? gridOf(specK, SPW_COLS) : 1), NT, 0, s>>>(
                nc, P1.p, gS, hist.p, xs, rj, rnext, dj, cS.p, P2.p, sp ? specU.p : nullptr, specK, specW.p, ctl.p);
            CHECK_LAUNCH();
            prec_dot(md, rnext, cZ.p, sp ? gC : 0);
            return;
        }
        {
            ctl_next_k<<<1, 1, 0, s>>>(ctl.p);
            const double *kdv = md == KMode::Stored32 && Kc.p ? cK.p : nullptr;
            S_apply(md, q, tauNow, geoNow, cD.p, ev);
            ctl_combine_k<D, 0><<<gC, NT, 0, s>>>(NC, V, cY.p, kdv, Kr.p, ctlIdx.p, cD.p, nullptr, ldmWgt, knlEps,
                                                 cS.p, P1.p, ctl.p);
        }
        ctl_update_k<<<gC, NT, 0, s>>>(nc, P1.p, gC, hist.p, xs, cR.p, cR.p, cD.p, cS.p, P2.p, nullptr, 0, nullptr,
                                       ctl.p);
        CHECK_LAUNCH();
        prec_dot(md, cR.p, cZ.p);
        ctl_dir_k<<<gC, NT, 0, s>>>(nc, P3.p, gridOf(nc, 64), hist.p, cZ.p, cD.p, ctl.p);
        CHECK_LAUNCH();
    }

    // the kernel timer's events serialise the launches around them: sampling
    // every 16th CG iteration costs under 0.5% of a solve on the 51k-landmark
    // shell (17.42 against 17.42-17.50 steps/s untimed) and about 6% on a
    // 2,000-landmark ring, whose launches are short (every 4th: 1.3% on the
    // shell; every iteration: 38% on the ring)
    static constexpr int TIMER_EVERY = 16;

    // PCG on S x = rhs (|rhs|^2 partials in P0, gC of them), preconditioned by
    // X^T X; x in/out (x0 = 0 when zeroStart); stop when |r|^2 < tolSqu |rhs|^2.
    // Returns 0 converged, 1 breakdown, 2 iteration budget spent.  The host
    // launches the iterations in chunks and reads the device control block
    // between chunks.
    // ctl_pcg_run, recording the Krylov basis when a spectral harvest is due
    // (first solve after a refresh of M) and harvesting from it afterwards
    int ctl_pcg(KMode md, const double *q, const double *tauNow, const double *geoNow, const double *rhs, double *xs,
                bool zeroStart, double tolSqu, int maxIt, int &iters) {
        recActive = specRecord && fused_iter(md);
        int st;
        try {
            st = ctl_pcg_run(md, q, tauNow, geoNow, rhs, xs, zeroStart, tolSqu, maxIt, iters);
        } catch (...) {
            recActive = false;
            throw;
        }
        if (recActive) {
            recActive = false;
            if (st != 1 && st != 3) spec_harvest(md, std::min(iters, SPEC_M));
        }
        return st;
    }

    // Rayleigh-Ritz for M S on the recorded Krylov basis (m iterations).
    // z_j = M r_j = d_j - beta_{j-1} d_{j-1}, so Z = D B with B upper
    // bidiagonal; the pencil is H = Z^T S Z = B^T (D^T SD) B against the Gram
    // matrix G = Z^T M^-1 Z = B^T (D^T R), both scaled to the Lanczos vectors
    // z_j / sqrt(r_j.z_j).  Directions where G is numerically singular (lost
    // orthogonality) are dropped; the leading Ritz vectors U = D B P c, with
    // G = Q g Q^T, P = Q g^-1/2 and c the eigenvectors of P^T H P, become the
    // spectral term.
    // C = recD^T X (m x m) over the nc recorded rows: row chunks as one
    // batched product (a single m x m x nc product leaves most of the chip
    // idle), the chunk results summed in chunk order by slice_rows_k
    static constexpr int SPEC_CHUNKS = 64;
    void spec_gram(const double *X, double *C, int m) {
        const int nc = NC * D;
        const int P = std::max(1, std::min(SPEC_CHUNKS, nc / 128));
        const int cs = nc / P, rem = nc - P * cs;
        const size_t mm = (size_t)m * m;
        double *scr = specScr.p;
        const double one = 1.0, zero = 0.0;
        if (rocblas_dgemm_strided_batched(blas, rocblas_operation_transpose, rocblas_operation_none, m, m, cs, &one,
                                          recD.p, nc, cs, X, nc, cs, &zero, scr, m, (rocblas_stride)mm, P) !=
            rocblas_status_success)
            throw SaError(SA_E_DEVICE, "rocblas_dgemm_strided_batched failed");
        int parts = P;
        if (rem > 0) {
            if (rocblas_dgemm(blas, rocblas_operation_transpose, rocblas_operation_none, m, m, rem, &one,
                              recD.p + (size_t)P * cs, nc, X + (size_t)P * cs, nc, &zero, scr + (size_t)P * mm, m) !=
                rocblas_status_success)
                throw SaError(SA_E_DEVICE, "rocblas_dgemm failed");
            ++parts;
        }
        slice_rows((int)mm, parts, scr, C, nullptr);
    }

    // Rayleigh-Ritz for M S on the recorded Krylov basis (m iterations).
    // z_j = M r_j = d_j - beta_{j-1} d_{j-1}, so Z = D B with B upper
    // bidiagonal; the pencil is H = Z^T S Z = B^T (D^T SD) B against the Gram
    // matrix G = Z^T M^-1 Z = B^T (D^T R), both scaled to the Lanczos vectors
    // z_j / sqrt(r_j.z_j).  Directions where G is numerically singular (lost
    // orthogonality) are dropped; the leading Ritz vectors U = D B P c, with
    // G = Q g Q^T, P = Q g^-1/2 and c the eigenvectors of P^T H P, become the
    // spectral term.  The m x m and m x r products run on the device.
    void spec_harvest(KMode md, int m) {
        const int nc = NC * D;
        if (m < 24) return;   // too short a basis: the next solve records again
        const size_t mm = (size_t)m * m, MM = (size_t)SPEC_M * SPEC_M;
        double *dA1 = specA.p, *dA2 = dA1 + MM, *dP = dA2 + MM, *dHP = dP + MM, *dHr = dHP + MM, *dY = dHr + MM;
        double *dEv = dY + MM, *dE = dEv + SPEC_M;
        std::vector<double> A1(mm), A2(mm), rz(m);
        const double one = 1.0, zero = 0.0;
        {
            spec_gram(recR.p, dA1, m);
            spec_gram(recSD.p, dA2, m);
            CK(hipStreamSynchronize(s));
        }
        CK(hipMemcpyAsync(A1.data(), dA1, sizeof(double) * mm, hipMemcpyDeviceToHost, s));
        CK(hipMemcpyAsync(A2.data(), dA2, sizeof(double) * mm, hipMemcpyDeviceToHost, s));
        CK(hipMemcpyAsync(rz.data(), hist.p, sizeof(double) * m, hipMemcpyDeviceToHost, s));
        CK(hipStreamSynchronize(s));
        specRecord = false;
        for (int j = 0; j < m; ++j)
            if (!(rz[j] > 0.0) || !std::isfinite(rz[j])) return;
        std::vector<double> beta(m, 0.0), sc(m);
        for (int j = 0; j + 1 < m; ++j) beta[j] = rz[j + 1] / rz[j];
        for (int j = 0; j < m; ++j) sc[j] = 1.0 / std::sqrt(rz[j]);
        auto at = [m](std::vector<double> &A, int i, int j) -> double & { return A[i + (size_t)j * m]; };
        // G = B^T A1, H = B^T A2 B (column-major), symmetrised and scaled by sc
        std::vector<double> G(mm), Tm(mm), H(mm);
        for (int j = 0; j < m; ++j)
            for (int i = 0; i < m; ++i) {
                at(G, i, j) = at(A1, i, j) - (i > 0 ? beta[i - 1] * at(A1, i - 1, j) : 0.0);
                at(Tm, i, j) = at(A2, i, j) - (j > 0 ? beta[j - 1] * at(A2, i, j - 1) : 0.0);
            }
        for (int j = 0; j < m; ++j)
            for (int i = 0; i < m; ++i) at(H, i, j) = at(Tm, i, j) - (i > 0 ? beta[i - 1] * at(Tm, i - 1, j) : 0.0);
        for (int j = 0; j < m; ++j)
            for (int i = 0; i <= j; ++i) {
                const double f = sc[i] * sc[j];
                const double g = 0.5 * (at(G, i, j) + at(G, j, i)) * f, h = 0.5 * (at(H, i, j) + at(H, j, i)) * f;
                at(G, i, j) = at(G, j, i) = g;
                at(H, i, j) = at(H, j, i) = h;
            }
        auto finite = [](const double *A, size_t n) {
            for (size_t i = 0; i < n; ++i)
                if (!std::isfinite(A[i])) return false;
            return true;
        };
        // a solve that went non-finite leaves nothing to harvest
        if (!finite(G.data(), mm) || !finite(H.data(), mm)) return;
        // eigenpairs of a symmetric n x n matrix on the device (in place in A);
        // eigenvalues ascending into ev
        std::vector<double> ev(m);
        auto syevd = [&](double *A, int n) -> bool {
            int info = 0;
            {
                if (rocsolver_dsyevd(blas, rocblas_evect_original, rocblas_fill_lower, n, A, n, dEv, dE,
                                     specInfo.p) != rocblas_status_success)
                    throw SaError(SA_E_DEVICE, "rocsolver_dsyevd failed");
                CK(hipStreamSynchronize(s));
            }
            CK(hipMemcpyAsync(ev.data(), dEv, sizeof(double) * n, hipMemcpyDeviceToHost, s));
            CK(hipMemcpyAsync(&info, specInfo.p, sizeof(int), hipMemcpyDeviceToHost, s));
            CK(hipStreamSynchronize(s));
            return info == 0 && finite(ev.data(), n);
        };
        // eigenpairs of the Gram matrix; keep directions above 1e-10 of its largest
        CK(hipMemcpyAsync(dA1, G.data(), sizeof(double) * mm, hipMemcpyHostToDevice, s));
        if (!syevd(dA1, m) || !(ev[m - 1] > 0.0)) return;
        CK(hipMemcpyAsync(G.data(), dA1, sizeof(double) * mm, hipMemcpyDeviceToHost, s));
        CK(hipStreamSynchronize(s));
        if (!finite(G.data(), mm)) return;
        std::vector<int> keep;
        for (int i = 0; i < m; ++i)
            if (ev[i] > 1e-10 * ev[m - 1]) keep.push_back(i);
        const int r = (int)keep.size();
        const int kk = std::min(SPEC_K, 2 * r / 3);
        if (kk < 8) return;
        std::vector<double> P((size_t)m * r);   // m x r: Q_keep / sqrt(g)
        for (int l = 0; l < r; ++l) {
            const double f = 1.0 / std::sqrt(ev[keep[l]]);
            for (int i = 0; i < m; ++i) P[i + (size_t)l * m] = G[i + (size_t)keep[l] * m] * f;
        }
        // P^T H P on the device, then its eigenpairs, then Y = P c (the kk
        // leading eigenvectors, ascending order: columns r - kk .. r - 1)
        CK(hipMemcpyAsync(dA2, H.data(), sizeof(double) * mm, hipMemcpyHostToDevice, s));
        CK(hipMemcpyAsync(dP, P.data(), sizeof(double) * m * r, hipMemcpyHostToDevice, s));
        {
            if (rocblas_dgemm(blas, rocblas_operation_none, rocblas_operation_none, m, r, m, &one, dA2, m, dP, m,
                              &zero, dHP, m) != rocblas_status_success ||
                rocblas_dgemm(blas, rocblas_operation_transpose, rocblas_operation_none, r, r, m, &one, dP, m, dHP, m,
                              &zero, dHr, r) != rocblas_status_success)
                throw SaError(SA_E_DEVICE, "rocblas_dgemm failed");
            CK(hipStreamSynchronize(s));
        }
        if (!syevd(dHr, r)) return;
        const double theta = ev[r - 1 - kk];
        if (!(theta > 0.0) || !(ev[r - 1] > theta)) return;
        {
            if (rocblas_dgemm(blas, rocblas_operation_none, rocblas_operation_none, m, kk, r, &one, dP, m,
                              dHr + (size_t)(r - kk) * r, r, &zero, dY, m) != rocblas_status_success)
                throw SaError(SA_E_DEVICE, "rocblas_dgemm failed");
            CK(hipStreamSynchronize(s));
        }
        std::vector<double> Y((size_t)m * kk), W((size_t)m * kk), cf(kk);
        CK(hipMemcpyAsync(Y.data(), dY, sizeof(double) * m * kk, hipMemcpyDeviceToHost, s));
        CK(hipStreamSynchronize(s));
        if (!finite(Y.data(), Y.size())) return;
        // W = B diag(sc) Y, largest Ritz value first; U = D W
        for (int l = 0; l < kk; ++l) {
            const double *y = &Y[(size_t)(kk - 1 - l) * m];
            for (int i = 0; i < m; ++i)
                W[i + (size_t)l * m] = sc[i] * y[i] - (i + 1 < m ? beta[i] * sc[i + 1] * y[i + 1] : 0.0);
            cf[l] = theta / ev[r - 1 - l] - 1.0;
        }
        CK(hipMemcpyAsync(dA2, W.data(), sizeof(double) * m * kk, hipMemcpyHostToDevice, s));
        CK(hipMemcpyAsync(specC.p, cf.data(), sizeof(double) * kk, hipMemcpyHostToDevice, s));
        {
            if (rocblas_dgemm(blas, rocblas_operation_none, rocblas_operation_none, nc, kk, m, &one, recD.p, nc, dA2, m,
                              &zero, specU.p, nc) != rocblas_status_success)
                throw SaError(SA_E_DEVICE, "rocblas_dgemm failed");
            CK(hipStreamSynchronize(s));
        }
        specOn = true;
        specMode = (int)md;
        specK = kk;
        specCarried = 0;
    }

    int ctl_pcg_run(KMode md, const double *q, const double *tauNow, const double *geoNow, const double *rhs, double *xs,
                    bool zeroStart, double tolSqu, int maxIt, int &iters) {
        const int nc = NC * D, gC = gridOf(nc);
        const double *kdv = md == KMode::Stored32 && Kc.p ? cK.p : nullptr;   // K_cc d (else gathered from K_xc d)
        pcg_start_k<<<1, NT, 0, s>>>(P0.p, gC, tolSqu, maxIt, ctl.p);
        CHECK_LAUNCH();
        zero_if_empty_k<<<gC, NT, 0, s>>>(nc, xs, nullptr, ctl.p);
        CHECK_LAUNCH();
        if (zeroStart) {
            CK(hipMemsetAsync(xs, 0, sizeof(double) * nc, s));
            ctl_combine_k<D, 1><<<gC, NT, 0, s>>>(NC, V, cZero.p, nullptr, nullptr, ctlIdx.p, cZero.p, rhs, 0.0, 0.0,
                                                 cR.p, P2.p, ctl.p);
        } else {
            S_apply(md, q, tauNow, geoNow, xs, nullptr);
            ctl_combine_k<D, 1><<<gC, NT, 0, s>>>(NC, V, cY.p, kdv, Kr.p, ctlIdx.p, xs, rhs, ldmWgt, knlEps, cR.p,
                                                 P2.p, ctl.p);
        }
        CHECK_LAUNCH();
        prec_dot(md, cR.p, cZ.p);
        ctl_dir_k<<<gC, NT, 0, s>>>(nc, P3.p, gridOf(nc, 64), hist.p, cZ.p, cD.p, ctl.p);   // it = -1: d = z
        CHECK_LAUNCH();
        if (maxIt < 1) ctl_next_k<<<1, 1, 0, s>>>(ctl.p);   // no budget: the counter check ends the solve here
        Ctl &h = hctl[0];
        int it0 = 0;
        int chunk = std::max(4, std::min(64, chunkHint));
        std::vector<std::pair<int, double>> rrAt;   // stallCheck: (iterations done, |r|^2) per chunk
        while (true) {
            // the counter check that ends a solve whose budget is spent runs
            // at the start of an iteration: iteration maxIt is launched too
            // (its first kernel marks the solve done, the rest are no-ops)
            int itEnd = std::min(maxIt + 1, it0 + chunk);
            if (timer.on) timer.ensure(itEnd - it0);
            // the kernel timer samples every TIMER_EVERY-th iteration (its
            // events serialise the launches around them)
            for (int it = it0; it < itEnd; ++it)
                cg_iter(md, q, tauNow, geoNow, xs,
                        timer.on && it % TIMER_EVERY == 0 ? &timer.ev[6 * (it - it0)] : nullptr, it);
            sync_ctl(ctl.p, &h);
            if (timer.on) {
                int last = h.done ? h.iters : itEnd;   // iterations it < last applied S
                for (int it = it0; it < std::min(itEnd, last); ++it) {
                    if (it % TIMER_EVERY) continue;
                    hipEvent_t *ev = &timer.ev[6 * (it - it0)];
                    float a0, a1, a2;
                    CK(hipEventElapsedTime(&a0, ev[0], ev[1]));
                    CK(hipEventElapsedTime(&a1, ev[2], ev[3]));
                    CK(hipEventElapsedTime(&a2, ev[4], ev[5]));
                    prof->kmvMs += a0 + a2; prof->kmvCount += 2;   // two product launches
                    prof->opMs += a1; prof->opCount++;
                }
            }
            if (h.done) break;
            if (stallCheck) {
                // an inner solve of the mixed form whose |r|^2 fell less than
                // half over the last STALL_WIN iterations (past STALL_MIN) no
                // longer converges: the FP32 operator is too far from the FP64
                // one, and the step is finished by the FP64 form at once
                // instead of after the inner solve's whole budget
                rrAt.emplace_back(itEnd, h.rr);
                if (itEnd >= STALL_MIN) {
                    double ref = -1.0;
                    for (const auto &pr : rrAt)
                        if (pr.first <= itEnd - STALL_WIN) ref = pr.second;
                    if (ref >= 0.0 && !(h.rr < 0.5 * ref)) {
                        iters = itEnd;
                        return 3;
                    }
                }
            }
            it0 = itEnd;
            chunk = 8;
        }
        iters = h.iters;
        return h.status;
    }

    // ctl_pcg guarded against a stale preconditioner: a solve that runs past
    // 4 x the iterations of the first step after the last refresh (+100)
    // rebuilds the preconditioner from this step's K_cc and continues from
    // its iterate with the rest of the budget
    int ctl_solve(KMode md, const double *q, const double *tauNow, const double *geoNow, const double *rhs, double *xs,
                  bool zeroStart, double tolSqu, int maxIt, int &iters) {
        const int cap = precAge > 0 && !jacobi ? std::min(std::max(1, maxIt / 2), 4 * precIters + 100) : maxIt;
        const bool spUsed = specOn && specMode == (int)md && !(specRecord && fused_iter(md));
        int st = ctl_pcg(md, q, tauNow, geoNow, rhs, xs, zeroStart, tolSqu, cap, iters);
        if (st == 1 && spUsed && iters < cap) {
            // a breakdown with the spectral term: continue from this iterate
            // with the plain preconditioner (the term returns at the next refresh)
            specOn = false;
            int more = 0;
            chunkHint = 16;
            st = ctl_pcg(md, q, tauNow, geoNow, rhs, xs, false, tolSqu, std::max(1, cap - iters), more);
            iters += more;
        }
        if (st == 2 && cap < maxIt) {
            kcc_build(true);
            precAge = 0;
            precEveryStep = true;   // K_cc too sensitive to reuse: fresh preconditioner each step
            int more = 0;
            chunkHint = 16;
            st = ctl_pcg(md, q, tauNow, geoNow, rhs, xs, false, tolSqu, maxIt - iters, more);
            iters += more;
        }
        return st;
    }

    int velocity_ctl(const double *qStep, const double *tauStep, const double *geoNow, int &iters, int &sweeps) {
        const int nc = NC * D, gC = gridOf(nc);
        const double *q = qStep, *tauNow = tauStep;
        yank(q, geoNow, f.p, P0.p);
        gather_ctl_k<D><<<gridOf(NC), NT, 0, s>>>(NC, V, ctlIdx.p, q, qc.p);
        CHECK_LAUNCH();
        const bool refresh = precAge < 0 || precAge >= PREC_MAX_AGE || lastIters > PREC_GROWTH * precIters + 4 ||
                             precEveryStep || forceRefresh;
        // a scheduled refresh inside a run keeps the spectral term (at most
        // SA_SPEC_CARRY times in a row); a run's first step, a growth-driven
        // refresh and every other case harvest a new one
        // this step in the mixed form (FP32 stored K inside FP64 refinement)
        // or in the FP64 form (stored FP64 K when it fits)
        const bool mixedStep = precision == SA_PREC_MIXED && (A32.p || kxMfma) && !mixedOff;
        const KMode cgMode = mixedStep ? (kxMfma ? KMode::Mfma32 : KMode::Stored32)
                                       : A64.p ? KMode::Stored64 : KMode::MatrixFree64;
        specCarry = refresh && precAge >= PREC_MAX_AGE && specOn && specMode == (int)cgMode &&
                    specCarried < SA_SPEC_CARRY;
        if (jacobi) jacobi_build(q, tauNow, geoNow);
        else kcc_build(refresh);
        if (!jacobi && refresh) {
            // a K_cc that does not factorise fails the step before any CG
            int info[2] = {0, 0};
            CK(hipMemcpyAsync(info, potrfInfo.p, sizeof(int) * 2, hipMemcpyDeviceToHost, s));
            CK(hipStreamSynchronize(s));
            if (info[0] != 0 || info[1] != 0) {
                failReason = info[0] != 0 ? SA_FAIL_CHOLESKY : SA_FAIL_TRTRI;
                iters = 0;
                sweeps = 0;
                return 1;
            }
        }
        if (refresh) { precAge = 0; forceRefresh = false; }
        if (!mixedStep && precision == SA_PREC_MIXED) use_fp64_store();
        kstore_build(q, !mixedStep, mixedStep);
        CK(hipMemsetAsync(ctl.p, 0, sizeof(Ctl), s));
        // g = K_cx f in FP64 (stored FP64 matrix when present), |g|^2 partials
        // exact products: the stored FP64 matrices (plain form), else the FP32
        // matrices with their remainders (mixed form), else matrix-free FP64
        KMode exact = !mixedStep ? (A64.p ? KMode::Stored64 : KMode::MatrixFree64)
                                 : (A32lo.p ? KMode::Split32 : KMode::MatrixFree64);
        kx_product(exact, false, q, f.p, cG.p, P0.p, nullptr);
        int status;
        sweeps = 0;
        bool kaReady = false;   // Kr holds K_xc alpha (FP64) for the final alpha
        if (!mixedStep) {
            if (precision == SA_PREC_MIXED) ++fp64Steps;
            chunkHint = lastIters - 4;
            status = ctl_solve(exact, q, tauNow, geoNow, cG.p, cA.p, false, cgTolSqu, itrMax, iters);
        } else {
            // iterative refinement: residual r = g - S alpha (FP64 arithmetic,
            // kernel entries exact to ~2^-48), inner
            // solve of S~ delta = r with the FP32 stored K (relative ~3e-5),
            // alpha += delta, until |r|^2 < cgTolSqu |g|^2
            double gg = 0.0;
            {
                pcg_start_k<<<1, NT, 0, s>>>(P0.p, gC, cgTolSqu, itrMax, ctl.p);
                CHECK_LAUNCH();
                sync_ctl(ctl.p, &hctl[0]);
                gg = hctl[0].bb;
            }
            iters = 0;
            status = 1;
            if (gg == 0.0) {
                CK(hipMemsetAsync(cA.p, 0, sizeof(double) * nc, s));
                status = 0;
            }
            double rrPrev = 0.0;
            bool stalled = false;
            // alpha entering the sweeps (the previous step's solution): the
            // FP64 form restarts from it if the sweeps stall, since a stalled
            // FP32 correction can carry alpha far from the solution
            CK(hipMemcpyAsync(cA0.p, cA.p, sizeof(double) * nc, hipMemcpyDeviceToDevice, s));
            for (int sw = 0; sw < 12 && gg != 0.0; ++sw) {
                CK(hipMemsetAsync(ctl.p, 0, sizeof(Ctl), s));
                S_apply(exact, q, tauNow, geoNow, cA.p, nullptr);
                ctl_combine_k<D, 1><<<gC, NT, 0, s>>>(NC, V, cY.p, nullptr, Kr.p, ctlIdx.p, cA.p, cG.p, ldmWgt, knlEps, cRout.p,
                                                     P0.p, ctl.p);
                CHECK_LAUNCH();
                pcg_start_k<<<1, NT, 0, s>>>(P0.p, gC, 1e-10, itrMax, ctl.p);
                CHECK_LAUNCH();
                sync_ctl(ctl.p, &hctl[0]);
                double rr = hctl[0].bb;
                if (rr < cgTolSqu * gg) { status = 0; kaReady = true; break; }
                // a sweep that does not cut the FP64 residual tenfold: the
                // FP32 operator is too far from the FP64 one (ill-conditioned
                // K_cc); the step is finished by the FP64 form
                if (sw > 0 && rr > 1e-2 * rrPrev) { stalled = true; break; }
                rrPrev = rr;
                ++sweeps;
                int its = 0;
                // inner tolerance: relative sqrt(SA_INNER_TOL2) of this residual, never below the outer target
                double inner = std::max(SA_INNER_TOL2, cgTolSqu * gg / rr);
                // a solve that records the Krylov basis of a spectral harvest runs longer
                if (specRecord) inner = std::max(std::min(inner, SA_REC_TOL2), cgTolSqu * gg / rr);
                // first chunk: about the previous step's iterations in this sweep
                chunkHint = sw < 4 && sweepIters[sw] > 0 && !specRecord ? std::max(4, sweepIters[sw] - 2) : 16;
                // an inner solve may not run past 3x the previous step's
                // iterations (+200; 1/8 of the budget on a run's first step):
                // an FP32 operator that does not converge counts as a stall
                const int capIn = lastIters > 0 ? 3 * lastIters + 200 : std::max(1000, itrMax / 8);
                int st;
                {
                    struct StallCheck {
                        bool &f;
                        explicit StallCheck(bool &b) : f(b) { f = true; }
                        ~StallCheck() { f = false; }
                    } sc(stallCheck);
                    st = ctl_solve(cgMode, q, tauNow, geoNow, cRout.p, cDelta.p, true, inner,
                                   std::max(1, std::min(itrMax - iters, capIn)), its);
                }
                iters += its;
                if (sw < 4) sweepIters[sw] = its;
                if (st != 0) { stalled = true; break; }   // breakdown or no convergence of the FP32 solve
                axpby_k<<<gC, NT, 0, s>>>(nc, cA.p, 1.0, cA.p, 1.0, cDelta.p);
                CHECK_LAUNCH();
                if (iters >= itrMax) break;
            }
            // not converged: the iteration budget is spent (status 2), or the
            // sweeps stalled or ran out (finished by the FP64 form below)
            if (status != 0 && iters >= itrMax) status = 2;
            if (status != 0 && status != 2) {
                // FP64 PCG from the current alpha, for this step and the rest
                // of the run (stored FP64 kernel matrices when they fit)
                ++fp64Steps;
                mixedOff = true;
                forceRefresh = true;   // the next step refreshes and takes its iteration count as the reference
                CK(hipMemcpyAsync(cA.p, cA0.p, sizeof(double) * nc, hipMemcpyDeviceToDevice, s));
                use_fp64_store();
                exact = KMode::MatrixFree64;
                if (A64.p) {
                    kstore_build(q, true, false);
                    exact = KMode::Stored64;
                }
                // g = K_cx f and its |g|^2 partials (P0) for the FP64 solve; the
                // sweeps left P0 holding the last residual's
                kx_product(exact, false, q, f.p, cG.p, P0.p, nullptr);
                chunkHint = 16;
                int its = 0;
                status = ctl_solve(exact, q, tauNow, geoNow, cG.p, cA.p, false, cgTolSqu,
                                 std::max(1, itrMax - iters), its);
                iters += its;
            }
        }
        if (status == 0) {
            // v = K_xc alpha; the converged residual check already formed it in FP64
            if (kaReady) CK(hipMemcpyAsync(x.p, Kr.p, sizeof(double) * V * D, hipMemcpyDeviceToDevice, s));
            else kx_product(exact, true, q, cA.p, x.p, nullptr, nullptr);
        }
        // a failed factorisation of K_cc explains a failed solve first
        int info[2] = {0, 0};
        CK(hipMemcpyAsync(info, potrfInfo.p, sizeof(int) * 2, hipMemcpyDeviceToHost, s));
        CK(hipStreamSynchronize(s));
        if (info[0] != 0) { failReason = SA_FAIL_CHOLESKY; return 1; }
        if (info[1] != 0) { failReason = SA_FAIL_TRTRI; return 1; }
        if (status == 0) return 0;
        failReason = status == 2 ? SA_FAIL_BUDGET : SA_FAIL_BREAKDOWN;
        return 1;
    }

    int reaction_diffusion(const double *tauNow, double *tauNxt, const double *geoNow, const double *ppdNow,
                           int &iters) {
        const int gV = gridOf(V);
        femvec_elem_k<D><<<gridOf(E), NT, 0, s>>>(E, elmVtx.p, geoNow, ra.p, frc.p);
        CHECK_LAUNCH();
        rd_start_k<<<gV, NT, 0, s>>>(V, E, D + 1, vtxElm.p, frc.p, rowPtr.p, colIdx.p, diagIdx.p, ppdNow, lhs.p, tauNow, dt,
                                    tauNxt, rdRhs.p, rdR.p, rdZ.p, rdD.p, P3.p);
        CHECK_LAUNCH();
        rd_fin_k<<<1, NT, 0, s>>>(P3.p, gV, rdTolSqu, rdHist.p, rdCtl.p);
        CHECK_LAUNCH();
        int it0 = 0, chunk = 8;
        const int rdMax = 200000;
        Ctl &h = hctl[1];
        while (true) {
            int itEnd = std::min(rdMax, it0 + chunk);
            // two launches per iteration: the direction of iteration it
            // is formed inside the next iteration's product (rd_spmv_dir_k);
            // the direction alternates between rdD (even) and rdD2 (odd)
            for (int it = it0; it < itEnd; ++it) {
                double *dIt = (it & 1) ? rdD2.p : rdD.p;
                if (it == 0) {
                    rd_spmv_k<<<gV, NT, 0, s>>>(V, rowPtr.p, colIdx.p, lhs.p, dIt, rdAd.p, P1.p, rdCtl.p);
                } else {
                    double *dPrev = (it & 1) ? rdD.p : rdD2.p;
                    rd_spmv_dir_k<<<gV, NT, 0, s>>>(V, rowPtr.p, colIdx.p, lhs.p, P2.p, gV, rdHist.p, it - 1, rdZ.p,
                                                   dPrev, dIt, rdAd.p, P1.p, rdCtl.p);
                }
                rd_update_k<<<gV, NT, 0, s>>>(V, P1.p, gV, rdHist.p, it, diagIdx.p, lhs.p, tauNxt, rdR.p, rdZ.p, dIt,
                                             rdAd.p, P2.p, rdCtl.p);
                CHECK_LAUNCH();
            }
            sync_ctl(rdCtl.p, &h);
            if (h.done) break;
            if (itEnd >= rdMax) { iters = rdMax; failReason = SA_FAIL_RD; return 1; }
            it0 = itEnd;
            chunk = std::min(chunk * 2, 128);
        }
        if (h.bb == 0.0) CK(hipMemsetAsync(tauNxt, 0, sizeof(double) * V, s));
        iters = h.iters;
        return 0;
    }

    double objective(const double *qT, double *gradOut, Stats *prof = nullptr) {
        facets_k<D><<<gridOf(F), NT, 0, s>>>(F, V, vfdVtx.p, qT, dCen.p, dDir.p, dVol.p);
        CHECK_LAUNCH();
        int nbD = gridOf(F, MV_ROWS), nbT = ttReady ? 0 : gridOf(Ft, MV_ROWS);
        const bool timed = prof && timer.on;
        if (timed) {
            timer.ensure_vfd();
            CK(hipEventRecord(timer.vfd[0], s));
        }
        if (gradOut)
            vfd_rows_k<D, true><<<nbD + nbT, NT, 0, s>>>(F, Ft, nbD, dCen.p, dDir.p, dVol.p, tCen.p, tDir.p, tVol.p, vp,
                                                        rowVal.p, gC.p, gN.p);
        else if (vfdMfma) {
            const int mD = gridOf(F, VFD_MFMA_ROWS), mT = ttReady ? 0 : gridOf(Ft, VFD_MFMA_ROWS);
            vfd_rows_mfma_k<D><<<mD + mT, NT, 0, s>>>(F, Ft, mD, dCen.p, dDir.p, dVol.p, tCen.p, tDir.p, tVol.p, vp,
                                                     rowVal.p);
        }
        else
            vfd_rows_k<D, false><<<nbD + nbT, NT, 0, s>>>(F, Ft, nbD, dCen.p, dDir.p, dVol.p, tCen.p, tDir.p, tVol.p, vp,
                                                         rowVal.p, gC.p, gN.p);
        CHECK_LAUNCH();
        if (timed) CK(hipEventRecord(timer.vfd[1], s));
        sum_k<<<1, NT, 0, s>>>(rowVal.p, F, scal.p);
        CHECK_LAUNCH();
        if (!ttReady) {
            sum_k<<<1, NT, 0, s>>>(rowVal.p + F, Ft, scal.p + 1);
            CHECK_LAUNCH();
        }
        double hs[2];
        CK(hipMemcpyAsync(hs, scal.p, sizeof(double) * 2, hipMemcpyDeviceToHost, s));
        if (gradOut) {
            vfd_grad_k<D><<<gridOf(V), NT, 0, s>>>(V, F, vfdAdj.p, vfdVtx.p, qT, gC.p, gN.p, grad.p);
            CHECK_LAUNCH();
            CK(hipMemcpyAsync(gradOut, grad.p, sizeof(double) * V * D, hipMemcpyDeviceToHost, s));
        }
        CK(hipStreamSynchronize(s));
        if (timed) {
            float ms = 0.f;
            CK(hipEventElapsedTime(&ms, timer.vfd[0], timer.vfd[1]));
            prof->vfdMs += ms;
            prof->vfdPairs += (long long)F * (F + Ft) + (ttReady ? 0LL : (long long)Ft * Ft);
            prof->vfdMfma = vfdMfma && !gradOut;
        }
        if (!ttReady) { ttSum = hs[1]; ttReady = true; }
        return hs[0] + ttSum;
    }

    // initial state of a solve: lmkStk[0] = q0, tau[0] from the bump
    // parameters (tauInitial.cu:52-63|53-64), geometry and FEM matrices at q0
    void start(const double *tauVar, int nVar, const double *tauPrm, int nPrm) {
        const size_t n = (size_t)V * D;
        require(tauVar && tauPrm, "null tau arguments");
        double c0, c1, c2 = 0.0, h, R;
        int pw;
        if (D == 2) {
            require(nVar >= 4 && nPrm >= 1, "2D tauVar needs [cx, cy, h, R] and tauPrm [p]");
            c0 = tauVar[0]; c1 = tauVar[1]; h = tauVar[2]; R = tauVar[3]; pw = (int)tauPrm[0];
        } else {
            require(nVar >= 4 && nPrm >= 2, "3D tauVar needs [cx, cy, cz, h] and tauPrm [R, p]");
            c0 = tauVar[0]; c1 = tauVar[1]; c2 = tauVar[2]; h = tauVar[3]; R = tauPrm[0]; pw = (int)tauPrm[1];
        }
        CK(hipMemcpyAsync(lmk.p, q0.p, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
        tau_initial_k<D><<<gridOf(V), NT, 0, s>>>(V, q0.p, c0, c1, c2, h, R, pw, tau.p);
        CHECK_LAUNCH();
        CK(hipMemsetAsync(x.p, 0, sizeof(double) * n, s));
        CK(hipMemsetAsync(a.p, 0, sizeof(double) * n, s));
        if (NC == 0) {
            CK(hipMemsetAsync(aM1.p, 0, sizeof(double) * n, s));
            CK(hipMemsetAsync(aM2.p, 0, sizeof(double) * n, s));
            CK(hipMemsetAsync(aM3.p, 0, sizeof(double) * n, s));
        }
        if (NC > 0) {
            CK(hipMemsetAsync(cA.p, 0, sizeof(double) * NC * D, s));
            CK(hipMemsetAsync(cAm1.p, 0, sizeof(double) * NC * D, s));
            CK(hipMemsetAsync(cAm2.p, 0, sizeof(double) * NC * D, s));
            CK(hipMemsetAsync(cAm3.p, 0, sizeof(double) * NC * D, s));
        }
        lastIters = 0;
        for (int &w : sweepIters) w = 0;
        restore_fp32_store();   // a fallback of the last run took the FP32 store
        precAge = -1;
        fp64Steps = 0;
        mixedOff = false;
        precEveryStep = false;
        forceRefresh = false;
        failReason = 0;
        geometry_k<D><<<gridOf(E), NT, 0, s>>>(E, V, elmVtx.p, frmVtx.p, tsvVtx.p, q0.p, spdTan, spdTsv, geo[0].p,
                                               femElm.p);
        CHECK_LAUNCH();
        op_frames(0);
        fem_gather_k<<<gridOf(Z), NT, 0, s>>>(Z, E, femIfo.p, femElm.p, dt, ppd[0].p, lhs.p);
        CHECK_LAUNCH();
    }

    // single operators at the initial state, for the parity tests
    void probe(int what, const double *tauVar, int nVar, const double *tauPrm, int nPrm, const double *in,
               double *out) override {
        GateShared gate;
        const size_t n = (size_t)V * D;
        start(tauVar, nVar, tauPrm, nPrm);
        CK(hipMemsetAsync(ctl.p, 0, sizeof(Ctl), s));
        switch (what) {
            case SA_PROBE_ELASTIC:
                CK(hipMemcpyAsync(x.p, in, sizeof(double) * n, hipMemcpyHostToDevice, s));
                // E x alone (no penalty term), as a solve applies it
                op_apply<2>(q0.p, x.p, tau.p, geo[0].p, nullptr, nullptr, f.p, nullptr, nullptr, false);
                CK(hipMemcpyAsync(out, f.p, sizeof(double) * n, hipMemcpyDeviceToHost, s));
                break;
            case SA_PROBE_PENALTY:
                // P(q0) in alone (the bottom penalty without btmWgt): the
                // element pass and the vertex gather of a solve, element forces zero
                require(B > 0, "the mesh has no bottom facets");
                CK(hipMemcpyAsync(x.p, in, sizeof(double) * n, hipMemcpyHostToDevice, s));
                btm_meas_k<D><<<gridOf(B), NT, 0, s>>>(B, V, btmVtx.p, q0.p, meas.p);
                pen_elem_k<D><<<gridOf(B), NT, 0, s>>>(B, V, btmVtx.p, meas.p, x.p, pfrc.p, ctl.p);
                frc.zero(s);
                gather_op_k<D, 2><<<gridOf(V), NT, 0, s>>>(V, E, vtxElm.p, frc.p, BL, B, vtxBtm.p, pfrc.p, 1.0, 0.0,
                                                           nullptr, nullptr, f.p, nullptr, nullptr, ctl.p, vtxStart.p);
                CHECK_LAUNCH();
                CK(hipMemcpyAsync(out, f.p, sizeof(double) * n, hipMemcpyDeviceToHost, s));
                break;
            case SA_PROBE_KXC:
            case SA_PROBE_KCX: {
                // the control-point kernel products at q0 as the inner solve
                // forms them (mixed form: FP32, stored or on the matrix cores;
                // FP64 form: stored or matrix-free FP64)
                require(NC > 0, "the session has no control points");
                const bool xc = what == SA_PROBE_KXC;
                gather_ctl_k<D><<<gridOf(NC), NT, 0, s>>>(NC, V, ctlIdx.p, q0.p, qc.p);
                CHECK_LAUNCH();
                restore_fp32_store();
                const KMode md = precision == SA_PREC_MIXED ? (kxMfma ? KMode::Mfma32 : KMode::Stored32)
                                                            : A64.p ? KMode::Stored64 : KMode::MatrixFree64;
                kstore_build(q0.p, md == KMode::Stored64, md == KMode::Stored32);
                double *xin = xc ? cD.p : Ad.p, *yout = xc ? Kr.p : cY.p;
                CK(hipMemcpyAsync(xin, in, sizeof(double) * (xc ? NC : V) * D, hipMemcpyHostToDevice, s));
                kx_product(md, xc, q0.p, xin, yout, nullptr, ctl.p);
                CK(hipMemcpyAsync(out, yout, sizeof(double) * (xc ? V : NC) * D, hipMemcpyDeviceToHost, s));
                break;
            }
            case SA_PROBE_KCC_CHOL: {
                // the factor L of K_cc(q0) + eps I as the preconditioner
                // build forms it (upper triangle zeroed)
                require(NC > 0 && !jacobi, "the session has no control-point preconditioner");
                gather_ctl_k<D><<<gridOf(NC), NT, 0, s>>>(NC, V, ctlIdx.p, q0.p, qc.p);
                CHECK_LAUNCH();
                const double invW = 1.0 / knlWidth;
                const int g = gridOf((long)NC * NC);
#define SA_KB(O) kcc_build_k<D, O><<<g, NT, 0, s>>>(NC, qc.p, invW, knlEps, Kcc.p, nullptr)
                switch (knlOrder) {
                    case -1: SA_KB(-1); break;
                    case 0: SA_KB(0); break;
                    case 1: SA_KB(1); break;
                    case 2: SA_KB(2); break;
                    case 3: SA_KB(3); break;
                    default: SA_KB(4); break;
                }
#undef SA_KB
                CHECK_LAUNCH();
                if (cholLib) {
                    GateExclusive lk;
                    if (rocsolver_dpotrf(blas, rocblas_fill_lower, NC, Kcc.p, NC, potrfInfo.p) !=
                        rocblas_status_success)
                        throw SaError(SA_E_DEVICE, "rocsolver_dpotrf failed");
                    CK(hipStreamSynchronize(s));
                } else {
                    cholesky(Kcc.p, NC, potrfInfo.p);
                }
                zero_upper_k<<<g, NT, 0, s>>>(NC, Kcc.p);
                CHECK_LAUNCH();
                int info = 0;
                CK(hipMemcpyAsync(&info, potrfInfo.p, sizeof(int), hipMemcpyDeviceToHost, s));
                CK(hipMemcpyAsync(out, Kcc.p, sizeof(double) * NC * NC, hipMemcpyDeviceToHost, s));
                CK(hipStreamSynchronize(s));
                require(info == 0, "K_cc + eps I is not positive definite (column " + std::to_string(info) + ")");
                break;
            }
            case SA_PROBE_JACOBI:
                require(jacobi, "the session does not use Jacobi preconditioning");
                gather_ctl_k<D><<<gridOf(NC), NT, 0, s>>>(NC, V, ctlIdx.p, q0.p, qc.p);
                if (btmWgt != 0.0 && B > 0) btm_meas_k<D><<<gridOf(B), NT, 0, s>>>(B, V, btmVtx.p, q0.p, meas.p);
                CHECK_LAUNCH();
                jacobi_build(q0.p, tau.p, geo[0].p);
                CK(hipMemcpyAsync(out, dS.p, sizeof(double) * NC * D, hipMemcpyDeviceToHost, s));
                break;
            case SA_PROBE_KMATVEC:
                CK(hipMemcpyAsync(r.p, in, sizeof(double) * n, hipMemcpyHostToDevice, s));
                if (Kst.p) kst_build(q0.p);
                kmatvec<false>(q0.p, r.p, Kr.p, nullptr, 0, -1);
                CK(hipMemcpyAsync(out, Kr.p, sizeof(double) * n, hipMemcpyDeviceToHost, s));
                break;
            case SA_PROBE_YANK:
                activity_k<<<gridOf(V), NT, 0, s>>>(V, tau.p, ynk[0], ynk[1], ynk[2], ynk[3], rea[0], rea[1], rea[2],
                                                   rea[3], rea[4], ya.p, ra.p);
                yank(q0.p, geo[0].p, f.p, P0.p);
                CK(hipMemcpyAsync(out, f.p, sizeof(double) * n, hipMemcpyDeviceToHost, s));
                CK(hipMemcpyAsync(out + n, tau.p, sizeof(double) * V, hipMemcpyDeviceToHost, s));
                break;
            case SA_PROBE_FEM:
                CK(hipMemcpyAsync(out, ppd[0].p, sizeof(double) * Z, hipMemcpyDeviceToHost, s));
                CK(hipMemcpyAsync(out + Z, lhs.p, sizeof(double) * Z, hipMemcpyDeviceToHost, s));
                break;
            case SA_PROBE_GEOMETRY:
                CK(hipMemcpyAsync(out, geo[0].p, sizeof(double) * G::N * E, hipMemcpyDeviceToHost, s));
                break;
            case SA_PROBE_VARIFOLD: {
                require(hasTarget, "no target in this session");
                double *qT = lmk.p + (size_t)(T - 1) * n;
                CK(hipMemcpyAsync(qT, in, sizeof(double) * n, hipMemcpyHostToDevice, s));
                // the value as a run computes it (no gradient), then the gradient
                out[0] = objective(qT, nullptr);
                (void)objective(qT, out + 1);
                break;
            }
            default:
                throw SaError(SA_E_ARG, "unknown probe");
        }
        CK(hipStreamSynchronize(s));
    }

    // Device time of one kernel product as a solve launches it (the product
    // kernel and its slice sum) at q0 on zero input: the mean over reps
    // launches after a warm-up one, in microseconds.
    double time_product(int what, int reps) override {
        require(reps >= 1 && reps <= 100000, "reps must be in [1, 100000]");
        GateShared gate;
        struct Ev {
            hipEvent_t e = nullptr;
            ~Ev() { if (e) (void)hipEventDestroy(e); }
        } e0, e1;
        CK(hipEventCreate(&e0.e));
        CK(hipEventCreate(&e1.e));
        CK(hipMemsetAsync(ctl.p, 0, sizeof(Ctl), s));
        auto timed = [&](auto &&one) {
            one();
            CK(hipEventRecord(e0.e, s));
            for (int k = 0; k < reps; ++k) one();
            CK(hipEventRecord(e1.e, s));
            CK(hipEventSynchronize(e1.e));
            float ms = 0.0f;
            CK(hipEventElapsedTime(&ms, e0.e, e1.e));
            return 1e3 * (double)ms / reps;
        };
        switch (what) {
            case SA_PROBE_KXC:
            case SA_PROBE_KCX: {
                require(NC > 0, "the session has no control points");
                const bool xc = what == SA_PROBE_KXC;
                gather_ctl_k<D><<<gridOf(NC), NT, 0, s>>>(NC, V, ctlIdx.p, q0.p, qc.p);
                CHECK_LAUNCH();
                restore_fp32_store();
                const KMode md = precision == SA_PREC_MIXED ? (kxMfma ? KMode::Mfma32 : KMode::Stored32)
                                                            : A64.p ? KMode::Stored64 : KMode::MatrixFree64;
                kstore_build(q0.p, md == KMode::Stored64, md == KMode::Stored32);
                double *xin = xc ? cD.p : Ad.p, *yout = xc ? Kr.p : cY.p;
                CK(hipMemsetAsync(xin, 0, sizeof(double) * (xc ? NC : V) * D, s));
                return timed([&] { kx_product(md, xc, q0.p, xin, yout, nullptr, ctl.p); });
            }
            case SA_PROBE_KMATVEC:
                require(NC == 0, "the session has control points (time SA_PROBE_KXC / SA_PROBE_KCX)");
                if (Kst.p) kst_build(q0.p);
                CK(hipMemsetAsync(r.p, 0, sizeof(double) * V * D, s));
                return timed([&] { kmatvec<false>(q0.p, r.p, Kr.p, nullptr, 0, -1); });
            case SA_PROBE_ELASTIC:
                // the operator apply as the control-point CG runs it
                // ((E + btmWgt P) x at q0, tau0 of the session's last start)
                CK(hipMemsetAsync(x.p, 0, sizeof(double) * V * D, s));
                if (btmWgt != 0.0 && B > 0) {
                    btm_meas_k<D><<<gridOf(B), NT, 0, s>>>(B, V, btmVtx.p, q0.p, meas.p);
                    CHECK_LAUNCH();
                }
                return timed([&] { op_apply<2>(q0.p, x.p, tau.p, geo[0].p, nullptr, nullptr, f.p, nullptr, nullptr); });
            case SA_TIME_PRECOND: {
                require(NC > 0 && !jacobi, "the session has no control-point preconditioner product");
                CK(hipMemsetAsync(cR.p, 0, sizeof(double) * NC * D, s));
                const bool f32 = precision == SA_PREC_MIXED && Mpre32.p;
                const int sl = f32 ? slicesM32 : slicesM;
                return timed([&] {
                    if (f32) {
                        dim3 g(gridOf(NC, panel_rows<float>()), sl);
                        kstore_mv_k<D, float, false, PRE_NT><<<g, NT, ks_lds<float>(NC, sl), s>>>(NC, NC, Mpre32.p, cR.p, partS.p, ctl.p);
                    } else {
                        dim3 g(gridOf(NC, panel_rows<double>()), sl);
                        kstore_mv_k<D, double, false, PRE_NT><<<g, NT, ks_lds<double>(NC, sl), s>>>(NC, NC, Mpre.p, cR.p, partS.p, ctl.p);
                    }
                    CHECK_LAUNCH();
                    slice_rows(D * NC, sl, partS.p, cZ.p, ctl.p);
                });
            }
            default:
                throw SaError(SA_E_ARG, "unknown product");
        }
    }

    void run(const double *tauVar, int nVar, const double *tauPrm, int nPrm, double *obj, double *lmkOut,
             double *tauOut, double *gradOut, Stats &st) override {
        GateShared gate;
        const size_t n = (size_t)V * D;
        st = Stats{};
        prof = &st;
        struct ProfReset {
            Engine *e;
            ~ProfReset() { e->prof = nullptr; }
        } profReset{this};
        const int TT = (Trun > 0 && Trun <= T) ? Trun : T;
        start(tauVar, nVar, tauPrm, nPrm);
        int cur = 0;
        int failed = 0;
        for (int t = 0; t < TT - 1; ++t) {
            const double *q = lmk.p + (size_t)t * n;
            double *qn = lmk.p + (size_t)(t + 1) * n;
            const double *tn = tau.p + (size_t)t * V;
            double *tx = tau.p + (size_t)(t + 1) * V;
            activity_k<<<gridOf(V), NT, 0, s>>>(V, tn, ynk[0], ynk[1], ynk[2], ynk[3], rea[0], rea[1], rea[2], rea[3],
                                               rea[4], ya.p, ra.p);
            CHECK_LAUNCH();
            if (btmWgt != 0.0 && B > 0) {
                btm_meas_k<D><<<gridOf(B), NT, 0, s>>>(B, V, btmVtx.p, q, meas.p);
                CHECK_LAUNCH();
            }
            int its = 0;
            int sweeps = 0;
            if (NC == 0 && t >= 1) {
                // all-landmark form: alpha = K^-1 v extrapolated the same way;
                // the step then starts from x0 = K alpha
                extrap_k<<<gridOf((long)n), NT, 0, s>>>((int)n, std::min(t - 1, SA_EXTRAP), a.p, aM1.p, aM2.p, aM3.p);
                CHECK_LAUNCH();
            }
            if (NC > 0 && t >= 1) {
                // initial guess: the previous solutions extrapolated to this
                // step (step t has t of them: the last one alone, then
                // linear from two, quadratic from three, cubic from four)
                extrap_k<<<gridOf((long)NC * D), NT, 0, s>>>(NC * D, std::min(t - 1, SA_EXTRAP), cA.p, cAm1.p, cAm2.p,
                                                             cAm3.p);
                CHECK_LAUNCH();
            }
            int vs = NC > 0 ? velocity_ctl(q, tn, geo[cur].p, its, sweeps) : velocity(q, tn, geo[cur].p, t > 0, its);
            st.refine += sweeps;
            st.fp64Steps = fp64Steps;
            st.pcgIters += its;
            lastIters = its;
            if (NC > 0) {
                if (precAge == 0) precIters = its;
                ++precAge;
            }
            if (vs != 0) { failed = 1; break; }
            axpby_k<<<gridOf((long)n), NT, 0, s>>>((int)n, qn, 1.0, q, dt, x.p);
            CHECK_LAUNCH();
            int nx = cur ^ 1;
            geometry_k<D><<<gridOf(E), NT, 0, s>>>(E, V, elmVtx.p, frmVtx.p, tsvVtx.p, qn, spdTan, spdTsv, geo[nx].p,
                                                   femElm.p);
            CHECK_LAUNCH();
            op_frames(nx);
            fem_gather_k<<<gridOf(Z), NT, 0, s>>>(Z, E, femIfo.p, femElm.p, dt, ppd[nx].p, lhs.p);
            CHECK_LAUNCH();
            int rits = 0;
            if (reaction_diffusion(tn, tx, geo[cur].p, ppd[cur].p, rits) != 0) { failed = 1; break; }
            st.rdIters += rits;
            st.steps = t + 1;
            cur = nx;
        }
        st.status = failed;
        st.failReason = failed ? failReason : 0;
        if (obj) {
            if (failed) *obj = DBL_MAX;
            else if (hasTarget) *obj = objective(lmk.p + (size_t)(TT - 1) * n, gradOut, &st);
            else *obj = 0.0;
        }
        if (lmkOut) CK(hipMemcpyAsync(lmkOut, lmk.p, sizeof(double) * n * TT, hipMemcpyDeviceToHost, s));
        if (tauOut) CK(hipMemcpyAsync(tauOut, tau.p, sizeof(double) * V * TT, hipMemcpyDeviceToHost, s));
        CK(hipStreamSynchronize(s));
    }
};

EngineBase *make_engine(const sa_mesh *m, const sa_target *t, const sa_params *p, const double *lmkIni) {
    if (!m || !p || (!lmkIni && !g_sizing)) throw SaError(SA_E_ARG, "null mesh, params or landmarks");
    if (m->dim == 2) return new Engine<2>(*m, t, *p, lmkIni);
    if (m->dim == 3) return new Engine<3>(*m, t, *p, lmkIni);
    throw SaError(SA_E_ARG, "Dimension " + std::to_string(m->dim) + " is not supported.");
}

template <typename Fn>
int guarded(Fn &&fn) {
    try {
        fn();
        return SA_OK;
    } catch (const SaError &e) {
        g_err = e.what();
        return e.code;
    } catch (const std::exception &e) {
        g_err = e.what();
        return SA_E_DEVICE;
    }
}

}  // namespace
}  // namespace sa

// Device scratch of sa_varifold_objective_noalloc (facet centres,
// directions and measures, the pair-sum rows, the two partial sums and the
// index flag): SA_NOALLOC_SCRATCH doubles, 8 MiB, held on every device the
// library runs on until the process ends.  In host code `scratch` names a
// host-side stand-in: kernels get the device address from
// hipGetSymbolAddress, looked up once per device.
#define SA_NOALLOC_SCRATCH (1 << 20)
static __device__ __align__(256) double scratch[SA_NOALLOC_SCRATCH];

namespace sa {
namespace {
std::mutex g_scratchMu;
constexpr int kMaxDevices = 64;
double *g_scratchDev[kMaxDevices] = {};

// device address of `scratch` on the current device (looked up on the
// device's first call, with the device current)
double *scratch_on_device() {
    int dev = -1;
    CK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) throw SaError(SA_E_DEVICE, "device index out of range");
    std::lock_guard<std::mutex> lk(g_scratchMu);
    if (!g_scratchDev[dev]) {
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            throw SaError(SA_E_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratchDev[dev] = static_cast<double *>(ptr);
    }
    return g_scratchDev[dev];
}

template <int D>
void varifold_noalloc(int V, const double *q, int F, const int *vfd, const sa_target &t, double *obj,
                      hipStream_t st) {
    double *base = scratch_on_device();
    double *cen = base, *dir = cen + (size_t)D * F, *vol = dir + (size_t)D * F, *rows = vol + F;
    double *scal = rows + F + t.tgtElmNum;
    int *bad = reinterpret_cast<int *>(scal + 2);
    VfdPrm vp{};
    vp.cenType = t.cenKnlType;
    vp.dirType = t.dirKnlType;
    vp.cenInvW2 = 1.0 / (t.cenKnlWidth * t.cenKnlWidth);
    vp.dirInvW2 = 1.0 / (t.dirKnlWidth * t.dirKnlWidth);
    const int nbD = gridOf(F, MV_ROWS), nbT = gridOf(t.tgtElmNum, MV_ROWS);
    zero_ints_k<<<1, 64, 0, st>>>(bad, 1);
    facets_k<D><<<gridOf(F), NT, 0, st>>>(F, V, vfd, q, cen, dir, vol, bad);
    vfd_rows_k<D, false><<<nbD + nbT, NT, 0, st>>>(F, t.tgtElmNum, nbD, cen, dir, vol, t.cenPosMat, t.uniDirMat,
                                                   t.elmVolVec, vp, rows, nullptr, nullptr);
    sum_k<<<1, NT, 0, st>>>(rows, F, scal);
    sum_k<<<1, NT, 0, st>>>(rows + F, t.tgtElmNum, scal + 1);
    vfd_total_k<<<1, 64, 0, st>>>(scal, bad, obj);
    CHECK_LAUNCH();
}
}  // namespace
}  // namespace sa

struct sa_session {
    std::unique_ptr<sa::EngineBase> eng;
};

#include "sa_server.hpp"

struct sa_server {
    std::unique_ptr<sa::Server> srv;
};

extern "C" {

const char *sa_last_error(void) { return sa::g_err.c_str(); }

int sa_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

int sa_session_create(const sa_mesh *mesh, const sa_target *tgt, const sa_params *prm, const double *lmkIniMat,
                      sa_session **out) {
    return sa::guarded([&] {
        if (!out) throw sa::SaError(SA_E_ARG, "null session output");
        *out = nullptr;
        auto *ss = new sa_session;
        try {
            ss->eng.reset(sa::make_engine(mesh, tgt, prm, lmkIniMat));
        } catch (...) {
            delete ss;
            throw;
        }
        *out = ss;
    });
}

int sa_session_run(sa_session *ss, const double *tauVarVec, int tauVarNum, const double *tauPrmVec, int tauPrmNum,
                   double *obj, double *lmkStk, double *tauMat, double *dqObjMat, sa_stats *stats) {
    return sa::guarded([&] {
        if (!ss || !ss->eng) throw sa::SaError(SA_E_NOTFOUND, "null session");
        sa::Stats st;
        auto t0 = std::chrono::steady_clock::now();
        ss->eng->run(tauVarVec, tauVarNum, tauPrmVec, tauPrmNum, obj, lmkStk, tauMat, dqObjMat, st);
        double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        if (stats) {
            stats->status = st.status;
            stats->steps = st.steps;
            stats->pcgIters = st.pcgIters;
            stats->rdIters = st.rdIters;
            stats->refineSweeps = st.refine;
            stats->fp64Steps = st.fp64Steps;
            stats->failReason = st.failReason;
            stats->seconds = sec;
            stats->kmatvecMs = st.kmvMs;
            stats->kmatvecCount = st.kmvCount;
            stats->operatorMs = st.opMs;
            stats->operatorCount = st.opCount;
            stats->varifoldMs = st.vfdMs;
            stats->varifoldPairs = st.vfdPairs;
            stats->varifoldMfma = st.vfdMfma;
            stats->deviceBytes = ss->eng->bytes;
            stats->factorizations = st.factorizations;
        }
    });
}

void sa_session_destroy(sa_session *ss) { delete ss; }

int sa_session_set_params(sa_session *ss, const sa_params *prm) {
    return sa::guarded([&] {
        if (!ss || !ss->eng) throw sa::SaError(SA_E_NOTFOUND, "null session");
        if (!prm) throw sa::SaError(SA_E_ARG, "null params");
        ss->eng->set_params(*prm);
    });
}

int sa_session_set_time_num(sa_session *ss, int timeNum) {
    return sa::guarded([&] {
        if (!ss || !ss->eng) throw sa::SaError(SA_E_NOTFOUND, "null session");
        if (timeNum < 1 || timeNum > ss->eng->timeLevels())
            throw sa::SaError(SA_E_ARG, "timeNum must be in 1..the session's timeNum");
        ss->eng->Trun = timeNum;
    });
}

int sa_session_set_profiling(sa_session *ss, int on) {
    return sa::guarded([&] {
        if (!ss || !ss->eng) throw sa::SaError(SA_E_NOTFOUND, "null session");
        ss->eng->timer.on = on != 0;
    });
}

int sa_session_probe(sa_session *ss, int what, const double *tauVarVec, int tauVarNum, const double *tauPrmVec,
                     int tauPrmNum, const double *in, double *out) {
    return sa::guarded([&] {
        if (!ss || !ss->eng) throw sa::SaError(SA_E_NOTFOUND, "null session");
        if (!out) throw sa::SaError(SA_E_ARG, "null output");
        ss->eng->probe(what, tauVarVec, tauVarNum, tauPrmVec, tauPrmNum, in, out);
    });
}

int sa_session_time_product(sa_session *ss, int what, int reps, double *usPerProduct) {
    return sa::guarded([&] {
        if (!ss || !ss->eng) throw sa::SaError(SA_E_NOTFOUND, "null session");
        if (!usPerProduct) throw sa::SaError(SA_E_ARG, "null output");
        *usPerProduct = ss->eng->time_product(what, reps);
    });
}

int sa_op_chunk_check(const sa_mesh *m, const double *lmkIniMat, int *chunks, long long *elementEntries,
                      long long *haloEntries) {
    return sa::guarded([&] {
        if (!m || !lmkIniMat || m->on_device) throw sa::SaError(SA_E_ARG, "host mesh and positions required");
        const int D = m->dim, V = m->lmkNum, E = m->elmNum, cols = m->vtxElmCols;
        if ((D != 2 && D != 3) || V < 1 || E < 1 || cols < 1 || !m->vtxElmMat)
            throw sa::SaError(SA_E_ARG, "bad mesh sizes");
        std::vector<int> hVE(m->vtxElmMat, m->vtxElmMat + (size_t)V * cols);
        const int cfg = sa::opf_cfg(E, V), nt = sa::OPF_SHAPES[cfg].nt, maxs = sa::OPF_SHAPES[cfg].maxs,
                  hm = sa::OPF_SHAPES[cfg].halo;
        if (!m->elmVtxMat) throw sa::SaError(SA_E_ARG, "bad mesh sizes");
        std::vector<int> hElm(m->elmVtxMat, m->elmVtxMat + (size_t)E * (D + 1));
        for (int v : hElm) sa::require(v >= 0 && v < V, "elmVtxMat index out of range");
        for (int i = 0; i < V; ++i) {
            sa::require(hVE[i] >= 0 && 1 + 2 * hVE[i] <= cols && hVE[i] <= maxs, "vtxElmMat count out of range");
            for (int a = 0; a < hVE[i]; ++a) {
                const int e = hVE[(size_t)(1 + 2 * a) * V + i], l = hVE[(size_t)(2 + 2 * a) * V + i];
                sa::require(e >= 0 && e < E && l >= 0 && l <= D, "vtxElmMat entry out of range");
            }
        }
        const sa::OpTables t = sa::op_tables(hVE, hElm, V, E, D, sa::morton_order(lmkIniMat, V, D), nt, maxs, hm);
        sa::require(!t.chVtx.empty(), "a vertex's elements pass the halo capacity");
        sa::check_op_tables(t, hVE, hElm, V, E, D, nt, maxs, hm);
        if (chunks) *chunks = (int)t.chV.size() - 1;
        if (elementEntries) *elementEntries = (long long)t.chElm.size();
        if (haloEntries) *haloEntries = (long long)t.chHv.size();
    });
}

int sa_tile_target(const char *product, int f32, long long rows, long long cols) {
    if (!product || rows < 1 || cols < 1) return -1;
    const std::string p(product);
    if (p == "op") return sa::opf_cfg((long)rows, (long)cols);
    if (p != "xc" && p != "cx" && p != "cc" && p != "kk") return -1;
    return sa::ks_target(product, f32 != 0, (long)rows, (long)cols, 0);
}

long long sa_session_footprint(const sa_mesh *m, const sa_target *t, const sa_params *p) {
    // a session constructor run in sizing mode: it validates the tables and
    // counts every allocation a session of this mesh makes, touching no device
    long long bytes = -1;
    const int rc = sa::guarded([&] {
        sa::SizingScope sizing;
        std::unique_ptr<sa::EngineBase> eng(sa::make_engine(m, t, p, nullptr));
        bytes = eng->bytes;
    });
    return rc == SA_OK ? bytes : -1;
}

int sa_deformation(const sa_mesh *mesh, const sa_params *prm, const double *tauVarMat, int tauVarNum, int candNum,
                   const double *tauPrmVec, int tauPrmNum, const double *lmkIniMat, double *lmkStk, double *tauMat,
                   double *objVec) {
    return sa::guarded([&] {
        if (candNum < 0 || (candNum > 0 && !tauVarMat)) throw sa::SaError(SA_E_ARG, "bad candidate matrix");
        std::unique_ptr<sa::EngineBase> eng(sa::make_engine(mesh, nullptr, prm, lmkIniMat));
        for (int c = 0; c < candNum; ++c) {
            sa::Stats st;
            double o = 0.0;
            bool last = (c == candNum - 1);
            eng->run(tauVarMat + (size_t)c * tauVarNum, tauVarNum, tauPrmVec, tauPrmNum, &o, last ? lmkStk : nullptr,
                     last ? tauMat : nullptr, nullptr, st);
            if (objVec) objVec[c] = st.status ? DBL_MAX : 0.0;
        }
    });
}

int sa_compute_objfcn(const sa_mesh *mesh, const sa_target *tgt, const sa_params *prm, const double *tauVarMat,
                      int tauVarNum, int candNum, const double *tauPrmVec, int tauPrmNum, const double *lmkIniMat,
                      double *objVec, double *dqObjMat) {
    return sa::guarded([&] {
        if (!tgt) throw sa::SaError(SA_E_ARG, "computeObjfcn needs a target");
        if (candNum < 0 || (candNum > 0 && !tauVarMat)) throw sa::SaError(SA_E_ARG, "bad candidate matrix");
        std::unique_ptr<sa::EngineBase> eng(sa::make_engine(mesh, tgt, prm, lmkIniMat));
        for (int c = 0; c < candNum; ++c) {
            sa::Stats st;
            double o = 0.0;
            bool last = (c == candNum - 1);
            eng->run(tauVarMat + (size_t)c * tauVarNum, tauVarNum, tauPrmVec, tauPrmNum, &o, nullptr, nullptr,
                     last ? dqObjMat : nullptr, st);
            if (objVec) objVec[c] = o;
        }
    });
}

int sa_varifold_objective_noalloc(int dim, int lmkNum, const double *lmkMat, int vfdElmNum, const int *vfdVtxMat,
                                  const sa_target *tgt, double *objDev, void *stream) {
    return sa::guarded([&] {
        if (dim != 2 && dim != 3) throw sa::SaError(SA_E_ARG, "dim must be 2 or 3");
        if (!tgt || !lmkMat || !vfdVtxMat || !objDev || !tgt->cenPosMat || !tgt->uniDirMat || !tgt->elmVolVec)
            throw sa::SaError(SA_E_ARG, "null argument");
        if (lmkNum <= 0 || vfdElmNum <= 0 || tgt->tgtElmNum <= 0) throw sa::SaError(SA_E_ARG, "empty input");
        if (tgt->cenKnlType != 'G' && tgt->cenKnlType != 'C')
            throw sa::SaError(SA_E_ARG, "cenKnlType must be 'G' or 'C'");
        if (tgt->dirKnlType != 'B' && tgt->dirKnlType != 'L' && tgt->dirKnlType != 'O' && tgt->dirKnlType != 'U')
            throw sa::SaError(SA_E_ARG, "dirKnlType must be 'B', 'L', 'O' or 'U'");
        if (!(tgt->cenKnlWidth > 0) || !(tgt->dirKnlWidth > 0)) throw sa::SaError(SA_E_ARG, "kernel widths must be positive");
        // centres, directions, measures, rows of both sets, 2 sums, the flag
        const long long need = (2LL * dim + 2) * vfdElmNum + (long long)tgt->tgtElmNum + 3;
        if (need > SA_NOALLOC_SCRATCH)
            throw sa::SaError(SA_E_ARG, "facet sets too large for the fixed scratch: (2 dim + 2) vfdElmNum + "
                                        "tgtElmNum + 3 = " + std::to_string(need) + " doubles > " +
                                        std::to_string(SA_NOALLOC_SCRATCH));
        hipStream_t st = static_cast<hipStream_t>(stream);
        if (dim == 2) sa::varifold_noalloc<2>(lmkNum, lmkMat, vfdElmNum, vfdVtxMat, *tgt, objDev, st);
        else sa::varifold_noalloc<3>(lmkNum, lmkMat, vfdElmNum, vfdVtxMat, *tgt, objDev, st);
    });
}

int sa_server_create(int workers, long long memBudgetBytes, sa_server **out) {
    return sa::guarded([&] {
        if (!out) throw sa::SaError(SA_E_ARG, "null server output");
        if (memBudgetBytes <= 0) {
            size_t fr = 0, tot = 0;
            if (hipMemGetInfo(&fr, &tot) != hipSuccess) throw sa::SaError(SA_E_DEVICE, "hipMemGetInfo failed");
            memBudgetBytes = (long long)(0.8 * (double)fr);
        }
        auto *sv = new sa_server;
        sv->srv.reset(new sa::Server(workers, memBudgetBytes));
        *out = sv;
    });
}

void sa_server_destroy(sa_server *sv) { delete sv; }

int sa_server_add_mesh(sa_server *sv, const sa_mesh *mesh, const sa_target *tgt, const sa_params *prm,
                       const double *lmkIniMat, int *meshId) {
    return sa::guarded([&] {
        if (!sv || !mesh || !prm || !lmkIniMat || !meshId) throw sa::SaError(SA_E_ARG, "null argument");
        *meshId = sv->srv->add_mesh(*mesh, tgt, *prm, lmkIniMat);
    });
}

int sa_server_submit(sa_server *sv, int meshId, const double *tauVarVec, int tauVarNum, const double *tauPrmVec,
                     int tauPrmNum, const sa_params *prm, long long *reqId) {
    return sa::guarded([&] {
        if (!sv || !reqId) throw sa::SaError(SA_E_ARG, "null argument");
        *reqId = sv->srv->submit(meshId, tauVarVec, tauVarNum, tauPrmVec, tauPrmNum, prm);
    });
}

int sa_server_wait(sa_server *sv, long long reqId, double *obj, sa_stats *stats, double *times) {
    int rc = SA_OK;
    int g = sa::guarded([&] {
        if (!sv) throw sa::SaError(SA_E_ARG, "null server");
        sa::Request r = sv->srv->wait(reqId);
        if (obj) *obj = r.obj;
        if (stats) *stats = r.st;
        if (times) { times[0] = r.tSubmit; times[1] = r.tStart; times[2] = r.tEnd; }
        rc = r.rc;
        if (rc != SA_OK) sa::g_err = r.err;
    });
    return g != SA_OK ? g : rc;
}

int sa_server_info(sa_server *sv, long long *memUsed, int *sessions, int *queued, int *running) {
    return sa::guarded([&] {
        if (!sv) throw sa::SaError(SA_E_ARG, "null server");
        sv->srv->info(memUsed, sessions, queued, running);
    });
}

}  // extern "C"
