// Device helpers shared by the render kernels: LDS-resident radix-2 FFT,
// Lanczos-3 taps, complex arithmetic, block reductions.  gfx950 only.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace trh {

__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cscale(float s, float2 a) { return make_float2(s * a.x, s * a.y); }

template <int N> struct Log2 { static constexpr int v = 1 + Log2<N / 2>::v; };
template <> struct Log2<1> { static constexpr int v = 0; };

template <int LOG2N>
__device__ __forceinline__ int bitrev(int j) {
    return (int)(__brev((unsigned)j) >> (32 - LOG2N));
}

// In-place radix-2 decimation-in-time FFT of `nseq` length-N sequences stored
// back to back in LDS (`buf[s*N + i]`), input already in bit-reversed order,
// output in natural order.  tw[m] = exp(+2 pi i m / N), m < N/2 (LDS).
// INVERSE=true computes sum_k X[k] exp(+2 pi i jk/N) (unnormalised), false
// uses exp(-...).  Every thread of the block must call this (barriers inside).
template <int N, bool INVERSE>
__device__ __forceinline__ void fft_dit(float2 *buf, int nseq, const float2 *tw) {
    constexpr int LOG2N = Log2<N>::v;
    constexpr int H2 = N / 2;
    const int nb = nseq * H2;
#pragma unroll 1
    for (int lg = 0; lg < LOG2N; ++lg) {
        const int half = 1 << lg;
        const int tsh = LOG2N - 1 - lg;
        for (int b = threadIdx.x; b < nb; b += blockDim.x) {
            const int s = b >> (LOG2N - 1);
            const int r = b & (H2 - 1);
            const int k = r & (half - 1);
            const int grp = r >> lg;
            const int i0 = s * N + (grp << (lg + 1)) + k;
            const int i1 = i0 + half;
            float2 w = tw[k << tsh];
            if (!INVERSE) w.y = -w.y;
            const float2 x0 = buf[i0];
            const float2 t = cmul(w, buf[i1]);
            buf[i0] = cadd(x0, t);
            buf[i1] = csub(x0, t);
        }
        __syncthreads();
    }
}

// Same transform as fft_dit, with pairs of radix-2 stages fused into one
// radix-4 pass (4 elements per thread in registers): ceil(log2 N / 2)
// barriers and half the LDS traffic.
template <int N, bool INVERSE>
__device__ __forceinline__ void fft_dit_r4(float2 *buf, int nseq, const float2 *tw) {
    constexpr int L = Log2<N>::v;
    int lg = 0;
    if (L & 1) {
        // leading radix-2 stage (span 1, twiddle 1)
        const int nb = nseq * (N / 2);
        for (int b = threadIdx.x; b < nb; b += blockDim.x) {
            const int i0 = 2 * b;
            const float2 x0 = buf[i0], x1 = buf[i0 + 1];
            buf[i0] = cadd(x0, x1);
            buf[i0 + 1] = csub(x0, x1);
        }
        __syncthreads();
        lg = 1;
    }
#pragma unroll 1
    for (; lg < L; lg += 2) {
        const int h = 1 << lg;
        const int nb = nseq * (N / 4);
        for (int b = threadIdx.x; b < nb; b += blockDim.x) {
            const int s = b >> (L - 2);
            const int r = b & (N / 4 - 1);
            const int k = r & (h - 1);
            const int grp = r >> lg;
            const int base = s * N + grp * 4 * h + k;
            const float2 a = buf[base], bq = buf[base + h], c = buf[base + 2 * h],
                         d = buf[base + 3 * h];
            float2 w1 = tw[k << (L - 1 - lg)];
            float2 w2 = tw[k << (L - 2 - lg)];
            float2 w3 = tw[(k + h) << (L - 2 - lg)];
            if (!INVERSE) {
                w1.y = -w1.y;
                w2.y = -w2.y;
                w3.y = -w3.y;
            }
            float2 t = cmul(w1, bq);
            const float2 a1 = cadd(a, t), b1 = csub(a, t);
            t = cmul(w1, d);
            const float2 c1 = cadd(c, t), d1 = csub(c, t);
            t = cmul(w2, c1);
            buf[base] = cadd(a1, t);
            buf[base + 2 * h] = csub(a1, t);
            t = cmul(w3, d1);
            buf[base + h] = cadd(b1, t);
            buf[base + 3 * h] = csub(b1, t);
        }
        __syncthreads();
    }
}

// lanczos_filter_3 (tractor/mp_fourier.i:355-402): 7 taps L3(i-3+dx), the one
// beyond +-3 set to zero, normalised to unit sum; computed in FP64, stored FP32.
__device__ __forceinline__ void lanczos3_taps(double dx, float *f) {
    if (dx == 0.0) {
        for (int i = 0; i < 7; ++i) f[i] = (i == 3) ? 1.f : 0.f;
        return;
    }
    double t[7];
    double s = 0.0;
    const int lo = dx > 0.0 ? 0 : 1;
    const int hi = dx > 0.0 ? 6 : 7;
    for (int i = 0; i < 7; ++i) t[i] = 0.0;
    for (int i = lo; i < hi; ++i) {
        const double pix = M_PI * (i - 3.0 + dx);
        const double val = 3.0 * sin(pix) * sin(pix / 3.0) / (pix * pix);
        t[i] = val;
        s += val;
    }
    for (int i = 0; i < 7; ++i) f[i] = (float)(t[i] / s);
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    return v;
}

// Sum of `v` over the block (blockDim.x multiple of 64); result valid in
// thread 0.  `red` is an LDS array of >= blockDim.x/64 doubles.
__device__ __forceinline__ double block_sum(double v, double *red) {
    v = wave_sum(v);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) red[wid] = v;
    __syncthreads();
    double r = 0.0;
    if (threadIdx.x == 0) {
        const int nw = blockDim.x >> 6;
        for (int i = 0; i < nw; ++i) r += red[i];
    }
    return r;
}

}  // namespace trh
