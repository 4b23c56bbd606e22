// EM fit of a mixture of Gaussians to PSF stamps on gfx950: the Gaussian
// approximation of a pixelized PSF that HybridPixelizedPSF carries
// (GaussianMixturePSF.fromStamp, tractor/psf.py:537-597, with
// em_fit_2d_reg, tractor/emfit.i:274).  One 256-thread workgroup per stamp
// keeps the stamp in LDS and runs every EM step: E step (responsibilities
// of the K components, terms with Mahalanobis distance^2 >= 100 dropped,
// pixels with no term ignored), then the M step's two passes (weights and
// centroids, then the second moments about the new centroids, with the
// inverse-Wishart prior alpha I).  FP64 throughout; the per-step sums are
// reduced in a fixed order.
#include "internal.h"

using namespace trh;

namespace {

constexpr int EM_TB = 256;
constexpr int EM_MAXK = 8;

__device__ __forceinline__ double wave_sum_em(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// v[j] <- the block's sum of v[j] (the same value in every thread).
template <int NV>
__device__ void em_block_sums(double (&v)[NV], double (*red)[EM_TB / 64]) {
#pragma unroll
    for (int j = 0; j < NV; ++j) {
        const double w = wave_sum_em(v[j]);
        if ((threadIdx.x & 63) == 0) red[j][threadIdx.x >> 6] = w;
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < NV; ++j) {
        double s = 0.0;
        for (int w = 0; w < EM_TB / 64; ++w) s += red[j][w];
        v[j] = s;
    }
    __syncthreads();
}

// K components at compile time (the per-pixel arrays stay in registers).
template <int K>
__global__ void __launch_bounds__(EM_TB)
k_em_fit(const double *__restrict__ stamps, int NY, int NX, int x0, int y0, double alpha,
         int steps, double *__restrict__ amp_io, double *__restrict__ mean_io,
         double *__restrict__ var_io, int32_t *__restrict__ status) {
    extern __shared__ double img[];
    __shared__ double red[3 * EM_MAXK][EM_TB / 64];
    __shared__ double pm[EM_MAXK][2], piv[EM_MAXK][4], psc[EM_MAXK];
    __shared__ double amp[EM_MAXK], mean[EM_MAXK][2], var[EM_MAXK][4];
    __shared__ int bad;
    const int s = blockIdx.x;
    const int N = NX * NY;
    const double *st = stamps + (int64_t)s * N;
    for (int i = threadIdx.x; i < N; i += EM_TB) img[i] = st[i];
    if (threadIdx.x < K) {
        const int k = threadIdx.x;
        amp[k] = amp_io[K * s + k];
        mean[k][0] = mean_io[2 * (K * s + k)];
        mean[k][1] = mean_io[2 * (K * s + k) + 1];
        for (int j = 0; j < 4; ++j) var[k][j] = var_io[4 * (K * s + k) + j];
    }
    if (threadIdx.x == 0) bad = 0;
    __syncthreads();
    const double tpd = 4.0 * M_PI * M_PI;   // (2 pi)^D, D = 2
    for (int step = 0; step < steps; ++step) {
        if (threadIdx.x < K) {
            const int k = threadIdx.x;
            const double *V = var[k];
            const double det = V[0] * V[3] - V[1] * V[2];
            if (!(det > 0.0)) bad = 1;
            piv[k][0] = V[3] / det;
            piv[k][1] = -V[1] / det;
            piv[k][2] = -V[2] / det;
            piv[k][3] = V[0] / det;
            psc[k] = amp[k] / sqrt(tpd * det);
            pm[k][0] = mean[k][0];
            pm[k][1] = mean[k][1];
        }
        __syncthreads();
        if (bad) break;
        // the responsibilities of pixel i under the previous parameters
        auto resp = [&](int i, double (&z)[K]) -> bool {
            const int iy = i / NX, ix = i - iy * NX;
            const double x = x0 + ix, y = y0 + iy;
            double zs = 0.0;
#pragma unroll
            for (int k = 0; k < K; ++k) {
                const double dx = x - pm[k][0], dy = y - pm[k][1];
                const double dsq = piv[k][0] * dx * dx + piv[k][1] * dx * dy +
                                   piv[k][2] * dx * dy + piv[k][3] * dy * dy;
                z[k] = dsq >= 100.0 ? 0.0 : psc[k] * exp(-0.5 * dsq);
                zs += z[k];
            }
            if (zs == 0.0) return false;
#pragma unroll
            for (int k = 0; k < K; ++k) z[k] /= zs;
            return true;
        };
        // pass 1: weights and centroids
        double acc[3 * K];
#pragma unroll
        for (int j = 0; j < 3 * K; ++j) acc[j] = 0.0;
        for (int i = threadIdx.x; i < N; i += EM_TB) {
            double z[K];
            if (!resp(i, z)) continue;
            const int iy = i / NX, ix = i - iy * NX;
            const double x = x0 + ix, y = y0 + iy;
#pragma unroll
            for (int k = 0; k < K; ++k) {
                const double wi = img[i] * z[k];
                acc[3 * k] += wi;
                acc[3 * k + 1] += wi * x;
                acc[3 * k + 2] += wi * y;
            }
        }
        em_block_sums(acc, red);
        double wsum[K], mx[K], my[K];
#pragma unroll
        for (int k = 0; k < K; ++k) {
            wsum[k] = acc[3 * k];
            mx[k] = acc[3 * k + 1] / wsum[k];
            my[k] = acc[3 * k + 2] / wsum[k];
        }
        // pass 2: second moments about the new centroids
#pragma unroll
        for (int j = 0; j < 3 * K; ++j) acc[j] = 0.0;
        for (int i = threadIdx.x; i < N; i += EM_TB) {
            double z[K];
            if (!resp(i, z)) continue;
            const int iy = i / NX, ix = i - iy * NX;
            const double x = x0 + ix, y = y0 + iy;
#pragma unroll
            for (int k = 0; k < K; ++k) {
                const double wi = img[i] * z[k];
                const double dx = x - mx[k], dy = y - my[k];
                acc[3 * k] += wi * dx * dx;
                acc[3 * k + 1] += wi * dx * dy;
                acc[3 * k + 2] += wi * dy * dy;
            }
        }
        em_block_sums(acc, red);
        if (threadIdx.x < K) {
            const int k = threadIdx.x;
            const double den = wsum[k] + alpha;
            var[k][0] = (alpha + acc[3 * k]) / den;
            var[k][1] = acc[3 * k + 1] / den;
            var[k][2] = acc[3 * k + 1] / den;
            var[k][3] = (alpha + acc[3 * k + 2]) / den;
            mean[k][0] = mx[k];
            mean[k][1] = my[k];
            amp[k] = wsum[k];
        }
        __syncthreads();
    }
    if (threadIdx.x < K) {
        const int k = threadIdx.x;
        amp_io[K * s + k] = amp[k];
        mean_io[2 * (K * s + k)] = mean[k][0];
        mean_io[2 * (K * s + k) + 1] = mean[k][1];
        for (int j = 0; j < 4; ++j) var_io[4 * (K * s + k) + j] = var[k][j];
    }
    if (threadIdx.x == 0) status[s] = bad ? -1 : 0;
}

}  // namespace

extern "C" int trh_em_fit_psf(trh_handle *h, int n, const double *stamps, int ny, int nx, int x0,
                              int y0, int K, double alpha, int steps, double *amp, double *mean,
                              double *var, int32_t *status, void *stream) {
    const size_t lds = sizeof(double) * (size_t)nx * ny;
    if (!h || n < 0 || (n > 0 && (!stamps || !amp || !mean || !var || !status)) || nx <= 0 ||
        ny <= 0 || K < 1 || K > EM_MAXK || steps < 0 || !(alpha >= 0.0) ||
        lds + 8192 > (size_t)h->lds_max)
        return set_err(TRH_E_ARG, "trh_em_fit_psf: bad arguments (K %d, stamp %d x %d)", K, nx, ny);
    if (n == 0) return TRH_OK;
    hipStream_t st = (hipStream_t)stream;
#define EM_LAUNCH(KK)                                                                            \
    case KK:                                                                                     \
        hipLaunchKernelGGL(k_em_fit<KK>, dim3(n), dim3(EM_TB), lds, st, stamps, ny, nx, x0, y0, \
                           alpha, steps, amp, mean, var, status);                                \
        break;
    switch (K) {
        EM_LAUNCH(1) EM_LAUNCH(2) EM_LAUNCH(3) EM_LAUNCH(4)
        EM_LAUNCH(5) EM_LAUNCH(6) EM_LAUNCH(7) EM_LAUNCH(8)
    }
#undef EM_LAUNCH
    return hip_check(hipGetLastError(), "k_em_fit");
}
