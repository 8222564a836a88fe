// qp_noalloc.hip -- the periodic 2D solve on caller-owned device buffers with
// no device allocation (qp_solve2d_device_noalloc).
//
// qp_solve2d_device needs a qp_fourier2d whose creation allocates the half
// spectrum, the twiddle tables and the eigenvalue terms.  Callers that cannot
// allocate (code under HIP graph capture, tight serving loops) get the same
// three solve launches with those buffers in one file-scope device array,
// `scratch`, sized for the largest grid the solver takes (1024 x 1024):
//   half spectrum   1024 x 513 complex   8 404 992 B
//   twiddles        2 x 1024 complex        32 768 B
//   eigenvalues     1024 + 513 doubles      12 296 B (rounded up to whole complex values)
// 8.07 MiB per GPU, held from the library's load until the process ends.
// The tables are rebuilt by a small kernel on every call, so nothing depends
// on what an earlier call left in the array; the spectrum is fully written by
// the row pass before the column pass reads it, so nothing needs clearing.
// Every call on one GPU shares that GPU's array: calls that overlap in time on
// one GPU (other streams or threads) race on it.
#include <mutex>

#include "qp_internal.h"

namespace qp {
namespace {

constexpr size_t kSpecMax = (size_t)kMaxN * (kMaxN / 2 + 1);   // complex values of the largest half spectrum
constexpr size_t kTwOff = kSpecMax;                              // axis-1 twiddles, then axis-2 twiddles
constexpr size_t kLamOff = kTwOff + 2 * (size_t)kMaxN;           // lam1 (kMaxN doubles), then lam2
constexpr size_t kLamComplex = ((size_t)kMaxN + kMaxN / 2 + 1 + 1) / 2;
constexpr size_t kScratchComplex = kLamOff + kLamComplex;

static __device__ __align__(256) double2 scratch[kScratchComplex];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
double2 *g_ptr[kMaxDevices] = {};   // scratch's address on each GPU (null until set up)

// Twiddles e^{-2 pi i t / n} of both axes and the 5-point eigenvalue terms
// (2 cos(2 pi j / n) - 2) / d / d of System_3d::solve_poisson_equation
// (src/System_3d.cpp:699), split per axis as the allocating solver stores them.
__global__ void k_noalloc_tables(double2 *tw1, int n1, double2 *tw2, int n2, double *lam1, double *lam2, int nh,
                                 double d1, double d2) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    double s, c;
    if (t < n1) {
        sincospi(2.0 * t / n1, &s, &c);
        tw1[t] = make_double2(c, -s);
        lam1[t] = (2 * c - 2) / d1 / d1;
    }
    if (t < n2) {
        sincospi(2.0 * t / n2, &s, &c);
        tw2[t] = make_double2(c, -s);
        if (t < nh) lam2[t] = (2 * c - 2) / d2 / d2;
    }
}

}  // namespace

int noalloc_setup(int device) {
    if (device < 0 || device >= kMaxDevices) {
        set_error("noalloc_setup: device index " + std::to_string(device) + " beyond " +
                  std::to_string(kMaxDevices));
        return QP_ERR_ARG;
    }
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_ptr[device]) return QP_OK;
    double2 *p = nullptr;
    QP_HIP_CHECK(hipGetSymbolAddress((void **)&p, HIP_SYMBOL(scratch)));
    if (!p) {
        set_error("noalloc_setup: hipGetSymbolAddress returned a null address");
        return QP_ERR_HIP;
    }
    g_ptr[device] = p;
    return QP_OK;
}

}  // namespace qp

using namespace qp;

extern "C" int qp_solve2d_device_noalloc(qp_ctx ctx, int n1, int n2, double d1, double d2, double D,
                                         const double *src_dev, double *dst_dev) {
    if (!ctx || !src_dev || !dst_dev) {
        set_error("qp_solve2d_device_noalloc: null argument");
        return QP_ERR_ARG;
    }
    if (n1 < 1 || n2 < 1 || n1 > kMaxN || n2 > kMaxN) {
        set_error("qp_solve2d_device_noalloc: grid " + std::to_string(n1) + " x " + std::to_string(n2) +
                  " outside [1, " + std::to_string(kMaxN) + "] per side (the size of its fixed scratch array)");
        return QP_ERR_ARG;
    }
    if (!(d1 > 0 && d2 > 0)) {
        set_error("qp_solve2d_device_noalloc: steps must be positive");
        return QP_ERR_ARG;
    }
    double2 *base = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        if (ctx->device >= 0 && ctx->device < kMaxDevices) base = g_ptr[ctx->device];
    }
    if (!base) {
        set_error("qp_solve2d_device_noalloc: scratch not set up on this device (context not from qp_ctx_create)");
        return QP_ERR_STATE;
    }
    QP_HIP_CHECK(hipSetDevice(ctx->device));
    hipStream_t s = ctx->stream;
    Solver2d S;   // a view on the scratch array: never destroy()ed
    S.n1 = n1;
    S.n2 = n2;
    S.nh = n2 / 2 + 1;
    S.d1 = d1;
    S.d2 = d2;
    QP_TRY(S.ax1.plan(n1));
    QP_TRY(S.ax2.plan(n2));
    double2 *tw1 = base + kTwOff, *tw2 = base + kTwOff + kMaxN;
    double *lam1 = reinterpret_cast<double *>(base + kLamOff), *lam2 = lam1 + kMaxN;
    S.ax1.dev.tw = tw1;
    S.ax2.dev.tw = tw2;
    S.d_lam1 = lam1;
    S.d_lam2 = lam2;
    S.spec[0] = base;
    S.spec[1] = base;   // single solves use spectrum 0 only
    S.default_shape();
    const int nt = n1 > n2 ? n1 : n2;
    hipLaunchKernelGGL(k_noalloc_tables, dim3((nt + 255) / 256), dim3(256), 0, s, tw1, n1, tw2, n2, lam1, lam2,
                       S.nh, d1, d2);
    QP_HIP_CHECK(hipGetLastError());
    QP_TRY(solve_rows_fwd_plain(S, src_dev, S.spec[0], s));
    QP_TRY(solve_cols(S, D, 1, s, nullptr, 0));
    QP_TRY(solve_rows_inv_unscaled(S, S.spec[0], dst_dev, s));
    return QP_OK;
}
