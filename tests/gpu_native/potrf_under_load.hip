// rocSOLVER / rocBLAS calls of the control-point preconditioner, alone and
// while another stream keeps the device busy: does each give bitwise the same
// result?  Handle set up as the engine sets it (atomics not allowed).
//
//   hipcc --offload-arch=gfx950 -O2 -std=c++17 -Xclang -target-feature -Xclang -packed-fp32-ops \
//       tests/gpu_native/potrf_under_load.hip -o /tmp/potrf_under_load -L/opt/rocm/lib -lrocsolver -lrocblas
//   BD_STAGE=<1..7> /tmp/potrf_under_load <C> <reps> <sigma>
//
// C x C Gaussian kernel matrix of C random points in the unit cube (width
// sigma) plus 1e-6 I.  BD_STAGE: 1 dpotrf, 2 dpotrf + dtrtri, 3 dpotrf +
// dtrtri + dsyrk (the preconditioner X^T X); 4 dtrtri, 5 dsyrk, 6 dgemm
// (C x 128 x C), 7 dsyevd (384 x 384), each alone from inputs formed before
// the load.  BD_BLOCKING=1: default-flag streams; BD_LOADBLOCKS: blocks of
// the load kernel (4096); BD_ATOMICS=1: atomics allowed.  The load: 40
// launches of a memory-bound kernel over 512 MB on a second stream, queued
// before each call.  Prints how many calls differ from the first one (made
// alone) and in how many the info word changed.
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <vector>

#define CK(x) do { hipError_t e_ = (x); if (e_ != hipSuccess) { std::printf("HIP %s at %d\n", hipGetErrorString(e_), __LINE__); std::exit(1); } } while (0)
#define RB(x) do { rocblas_status e_ = (x); if (e_ != rocblas_status_success) { std::printf("rocblas %d at %d\n", (int)e_, __LINE__); std::exit(1); } } while (0)

__global__ void load_k(long n, double *a, const double *b) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        double v = b[i];
        for (int k = 0; k < 8; ++k) v = sqrt(v * v + 1.0) - 0.5;
        a[i] = v;
    }
}

int main(int argc, char **argv) {
    const int C = argc > 1 ? std::atoi(argv[1]) : 2000;
    const int reps = argc > 2 ? std::atoi(argv[2]) : 20;
    const double sigma = argc > 3 ? std::atof(argv[3]) : 0.25;
    std::mt19937_64 rng(7);
    std::uniform_real_distribution<double> U(0.0, 1.0);
    std::vector<double> p(3 * (size_t)C), K((size_t)C * C);
    for (auto &v : p) v = U(rng);
    for (int i = 0; i < C; ++i)
        for (int j = 0; j < C; ++j) {
            double d2 = 0;
            for (int c = 0; c < 3; ++c) { double d = p[3 * i + c] - p[3 * j + c]; d2 += d * d; }
            K[(size_t)j * C + i] = std::exp(-d2 / (2 * sigma * sigma)) + (i == j ? 1e-6 : 0.0);
        }
    const int stage = std::getenv("BD_STAGE") ? std::atoi(std::getenv("BD_STAGE")) : 3;
    const bool blocking = std::getenv("BD_BLOCKING") && std::getenv("BD_BLOCKING")[0] == '1';
    const int lblocks = std::getenv("BD_LOADBLOCKS") ? std::atoi(std::getenv("BD_LOADBLOCKS")) : 4096;
    hipStream_t sA, sB;
    CK(hipStreamCreateWithFlags(&sA, blocking ? hipStreamDefault : hipStreamNonBlocking));
    CK(hipStreamCreateWithFlags(&sB, blocking ? hipStreamDefault : hipStreamNonBlocking));
    rocblas_handle h;
    RB(rocblas_create_handle(&h));
    RB(rocblas_set_stream(h, sA));
    if (!(std::getenv("BD_ATOMICS") && std::getenv("BD_ATOMICS")[0] == '1'))
        RB(rocblas_set_atomics_mode(h, rocblas_atomics_not_allowed));
    double *dK0, *dW, *dM;
    int *dInfo;
    const size_t nn = (size_t)C * C;
    CK(hipMalloc(&dK0, nn * 8)); CK(hipMalloc(&dW, nn * 8)); CK(hipMalloc(&dM, nn * 8)); CK(hipMalloc(&dInfo, 8));
    CK(hipMemset(dM, 0, nn * 8));
    CK(hipMemcpy(dK0, K.data(), nn * 8, hipMemcpyHostToDevice));
    const long LN = 64L << 20;   // 512 MB per load buffer
    double *la, *lb;
    CK(hipMalloc(&la, LN * 8)); CK(hipMalloc(&lb, LN * 8));
    CK(hipMemset(lb, 0, LN * 8));
    std::vector<double> ref(nn), out(nn);
    int refInfo[2] = {0, 0};
    const double one = 1.0, zero = 0.0;
    // stages 4-7 time one call alone from inputs formed beforehand (no load):
    // 4 trtri of the factor, 5 syrk of the inverse factor, 6 dgemm C x 128 x C,
    // 7 dsyevd of the leading 384 x 384 block of M
    double *dL = nullptr, *dX = nullptr, *dE = nullptr, *dEv = nullptr;
    if (stage >= 4) {
        CK(hipMalloc(&dL, nn * 8)); CK(hipMalloc(&dX, nn * 8)); CK(hipMalloc(&dE, nn * 8)); CK(hipMalloc(&dEv, 8 * (size_t)C));
        CK(hipMemcpy(dL, dK0, nn * 8, hipMemcpyDeviceToDevice));
        RB(rocsolver_dpotrf(h, rocblas_fill_lower, C, dL, C, dInfo));
        CK(hipMemcpy(dX, dL, nn * 8, hipMemcpyDeviceToDevice));
        RB(rocsolver_dtrtri(h, rocblas_fill_lower, rocblas_diagonal_non_unit, C, dX, C, dInfo + 1));
        CK(hipDeviceSynchronize());
    }
    auto once = [&](std::vector<double> &o, int *info) {
        CK(hipMemsetAsync(dInfo, 0, 8, sA));
        if (stage <= 3) {
            CK(hipMemcpyAsync(dW, dK0, nn * 8, hipMemcpyDeviceToDevice, sA));
            RB(rocsolver_dpotrf(h, rocblas_fill_lower, C, dW, C, dInfo));
            if (stage >= 2) RB(rocsolver_dtrtri(h, rocblas_fill_lower, rocblas_diagonal_non_unit, C, dW, C, dInfo + 1));
            if (stage >= 3) RB(rocblas_dsyrk(h, rocblas_fill_lower, rocblas_operation_transpose, C, C, &one, dW, C, &zero, dM, C));
        } else if (stage == 4) {
            CK(hipMemcpyAsync(dW, dL, nn * 8, hipMemcpyDeviceToDevice, sA));
            RB(rocsolver_dtrtri(h, rocblas_fill_lower, rocblas_diagonal_non_unit, C, dW, C, dInfo + 1));
        } else if (stage == 5) {
            RB(rocblas_dsyrk(h, rocblas_fill_lower, rocblas_operation_transpose, C, C, &one, dX, C, &zero, dM, C));
        } else if (stage == 6) {
            RB(rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none, C, 128, C, &one, dK0, C, dX, C, &zero, dM, C));
        } else {
            const int m = 384;
            CK(hipMemcpy2DAsync(dW, (size_t)m * 8, dK0, (size_t)C * 8, (size_t)m * 8, m, hipMemcpyDeviceToDevice, sA));
            RB(rocsolver_dsyevd(h, rocblas_evect_original, rocblas_fill_lower, m, dW, m, dEv, dE, dInfo));
        }
        const double *src = stage == 3 || stage == 5 || stage == 6 ? dM : dW;
        CK(hipMemcpyAsync(o.data(), src, nn * 8, hipMemcpyDeviceToHost, sA));
        CK(hipMemcpyAsync(info, dInfo, 8, hipMemcpyDeviceToHost, sA));
        CK(hipStreamSynchronize(sA));
    };
    once(ref, refInfo);
    std::printf("C=%d sigma=%g stage %d blocking %d load blocks %d: reference info %d %d\n", C, sigma, stage, (int)blocking,
                lblocks, refInfo[0], refInfo[1]);
    int diffAlone = 0, diffLoad = 0, infoLoad = 0;
    double maxd = 0;
    for (int r = 0; r < reps; ++r) {
        int info[2];
        once(out, info);
        if (std::memcmp(out.data(), ref.data(), nn * 8)) ++diffAlone;
    }
    for (int r = 0; r < reps; ++r) {
        for (int k = 0; k < 40; ++k) load_k<<<lblocks, 256, 0, sB>>>(LN, la, lb);
        CK(hipGetLastError());
        int info[2];
        once(out, info);
        if (info[0] != refInfo[0] || info[1] != refInfo[1]) ++infoLoad;
        if (std::memcmp(out.data(), ref.data(), nn * 8)) {
            if (diffLoad == 0) {
                for (size_t i = 0; i < nn; ++i)
                    if (out[i] != ref[i]) {
                        std::printf("  first difference at row %zu col %zu: %.17g vs %.17g (info %d %d)\n", i % C, i / C,
                                    out[i], ref[i], info[0], info[1]);
                        break;
                    }
            }
            ++diffLoad;
            for (size_t i = 0; i < nn; ++i) maxd = std::fmax(maxd, std::fabs(out[i] - ref[i]) / (std::fabs(ref[i]) + 1e-300));
        }
        CK(hipStreamSynchronize(sB));
    }
    std::printf("alone: %d of %d differ; under load: %d of %d differ (max rel %.3g), info changed %d\n", diffAlone,
                reps, diffLoad, reps, maxd, infoLoad);
    std::fflush(stdout);
    return 0;
}
