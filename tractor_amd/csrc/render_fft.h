// FFT-path and pixelized-PSF render device code, shared by the C-ABI batch
// entry points and the batched fitter.  gfx950 only.
#pragma once
#include "fft_common.h"
#include "../../include/tractor_hip.h"

namespace trh {

// -2 pi^2 log2(e): exp(-2 pi^2 q) = exp2(C * q)
#define TRH_NEG2PI2_LOG2E (-28.47765484177155)

// LDS bytes (dynamic) used by k_render_fft<SZ>: twiddles + work + row buffer.
__host__ __device__ inline size_t fft_lds_bytes(int sz, int ch, int maxrows) {
    return sizeof(float2) * ((size_t)sz / 2 + (size_t)ch * sz + (size_t)maxrows * (sz / 2 + 1));
}

struct RenderScratch {
    float4 comp[TRH_MAXK];
    float taps[14];      // fx[0..6], fy[7..13]
    double dtaps[14];
    double red[16];
};

// Lanczos taps for (mux, muy) into S.taps (all threads call; 2 barriers).
__device__ __forceinline__ void compute_taps(double mux, double muy, RenderScratch &S) {
    const int tid = threadIdx.x;
    if (tid < 14) {
        const double dx = tid < 7 ? mux : muy;
        const int i = tid % 7;
        double val = 0.0;
        if (dx == 0.0) {
            val = (i == 3) ? 1.0 : 0.0;
        } else {
            const bool zero = dx > 0.0 ? (i == 6) : (i == 0);
            if (!zero) {
                const double pix = M_PI * (i - 3.0 + dx);
                val = 3.0 * sin(pix) * sin(pix / 3.0) / (pix * pix);
            }
        }
        S.dtaps[tid] = val;
    }
    __syncthreads();
    if (tid < 14) {
        const int g = tid < 7 ? 0 : 7;
        const double dx = tid < 7 ? mux : muy;
        double s = 0.0;
        // same summation order as lanczos_filter_3 (mp_fourier.i:380-400)
        for (int i = 0; i < 7; ++i) s += S.dtaps[g + i];
        S.taps[tid] = (dx == 0.0) ? (float)S.dtaps[tid] : (float)(S.dtaps[tid] / s);
    }
    __syncthreads();
}

// Epilogue of one output pixel.  Returns the chi^2 contribution (TRH_CHI2).
__device__ __forceinline__ double epilogue(int mode, float val, float scale, int64_t idx,
                                           float *__restrict__ out,
                                           const float *__restrict__ data,
                                           const float *__restrict__ inverr) {
    if (mode == TRH_STORE) {
        out[idx] = scale * val;
    } else if (mode == TRH_ADD) {
        atomicAdd(out + idx, scale * val);
    } else {
        const float mod = val * scale;
        const float chi = (data[idx] - mod) * inverr[idx];
        return (double)chi * (double)chi;
    }
    return 0.0;
}

// Render one FFT job (see trh_fft_job).  Every thread of the block calls this.
// tw (LDS) must hold exp(+2 pi i m/SZ), m < SZ/2.  work: ch*SZ float2,
// hbuf: maxrows*(SZ/2+1) float2.
template <int SZ, int MODE>
__device__ void render_fft_one(const trh_fft_job *__restrict__ Jp,
                               const float2 *__restrict__ psf_spec, const float2 *tw,
                               float2 *work, float2 *hbuf, int ch, int maxrows,
                               RenderScratch &S, float *__restrict__ out,
                               const float *__restrict__ data, const float *__restrict__ inverr,
                               double *__restrict__ chi2) {
    constexpr int NV = SZ / 2 + 1;
    constexpr int LOG2 = Log2<SZ>::v;
    const int tid = threadIdx.x;
    const int nthr = blockDim.x;
    const int64_t off = Jp->out_offset;
    const int stride = Jp->out_stride;
    const int mw = Jp->mw, mh = Jp->mh, sx = Jp->sx, sy = Jp->sy;
    // MODE >= 0: epilogue fixed at compile time; -1: read from the job
    const int K = min(max(Jp->K, 0), TRH_MAXK), mode = MODE >= 0 ? MODE : Jp->mode;
    const float scale = Jp->scale;
    const float2 *__restrict__ P = psf_spec + (size_t)Jp->psf * SZ * NV;

    if (tid < K) {
        const double C = TRH_NEG2PI2_LOG2E;
        S.comp[tid] = make_float4(Jp->amp[tid], (float)(C * Jp->var[tid][0]),
                                  (float)(2.0 * C * Jp->var[tid][1]),
                                  (float)(C * Jp->var[tid][2]));
    }
    compute_taps(Jp->mux, Jp->muy, S);
    const float *fx = S.taps;
    const float *fy = S.taps + 7;
    const float inv_norm = 1.0f / (float)(SZ * SZ);

    double acc = 0.0;
    // G rows [ylo, yhi) land in the window; other window rows are empty.
    const int ylo = max(0, -sy);
    const int yhi = min(SZ, mh - sy);
    if (mode != TRH_ADD) {
        const int top = min(mh, max(0, sy));               // rows [0, top)
        const int bot = max(top, min(mh, SZ + sy));        // rows [bot, mh)
        const int nempty = (top + (mh - bot)) * mw;
        for (int idx = tid; idx < nempty; idx += nthr) {
            const int r = idx / mw;
            const int ox = idx - r * mw;
            const int oy = r < top ? r : bot + (r - top);
            acc += epilogue(mode, 0.f, scale, off + (int64_t)oy * stride + ox, out, data, inverr);
        }
    }

    for (int rb0 = ylo; rb0 < yhi; rb0 += maxrows) {
        const int nrb = min(maxrows, yhi - rb0);
        // ---- phase A: column IFFTs along w, y-Lanczos onto the needed rows
        for (int c0 = 0; c0 < NV; c0 += ch) {
            const int nc = min(ch, NV - c0);
            for (int idx = tid; idx < nc * SZ; idx += nthr) {
                const int j = idx / nc;
                const int cc = idx - j * nc;
                const int i = c0 + cc;
                const float v = (float)i * (1.0f / SZ);
                const float w = (float)(j < SZ / 2 ? j : j - SZ) * (1.0f / SZ);
                const float v2 = v * v, vw = v * w, w2 = w * w;
                float F = 0.f;
                for (int k = 0; k < K; ++k) {
                    const float4 c = S.comp[k];
                    F += c.x * __builtin_amdgcn_exp2f(c.y * v2 + c.z * vw + c.w * w2);
                }
                const float2 p = P[j * NV + i];
                work[cc * SZ + bitrev<LOG2>(j)] = make_float2(F * p.x, F * p.y);
            }
            __syncthreads();
            fft_dit_r4<SZ, true>(work, nc, tw);
            for (int idx = tid; idx < nrb * nc; idx += nthr) {
                const int q = idx / nc;
                const int cc = idx - q * nc;
                const int yr = rb0 + q;
                float2 a = make_float2(0.f, 0.f);
#pragma unroll
                for (int t = 0; t < 7; ++t) {
                    const int ys = yr - 3 + t;
                    if (ys >= 0 && ys < SZ) {
                        const float2 h = work[cc * SZ + ys];
                        a.x += fy[t] * h.x;
                        a.y += fy[t] * h.y;
                    }
                }
                hbuf[q * NV + c0 + cc] = a;
            }
            __syncthreads();
        }
        // ---- phase B: two real rows per complex IFFT, x-Lanczos, epilogue
        const int npairs = (nrb + 1) / 2;
        for (int p0 = 0; p0 < npairs; p0 += ch) {
            const int npp = min(ch, npairs - p0);
            for (int idx = tid; idx < npp * SZ; idx += nthr) {
                const int pp = idx >> LOG2;
                const int k = idx & (SZ - 1);
                const int qa = 2 * (p0 + pp);
                const bool hasb = (qa + 1) < nrb;
                const float2 *A = hbuf + qa * NV;
                const float2 *B = A + NV;
                float2 Z;
                if (k == 0 || k == SZ / 2) {
                    // irfft ignores the imaginary parts of DC and Nyquist
                    Z = make_float2(A[k].x, hasb ? B[k].x : 0.f);
                } else if (k < SZ / 2) {
                    const float2 a = A[k];
                    const float2 b = hasb ? B[k] : make_float2(0.f, 0.f);
                    Z = make_float2(a.x - b.y, a.y + b.x);
                } else {
                    const float2 a = A[SZ - k];
                    const float2 b = hasb ? B[SZ - k] : make_float2(0.f, 0.f);
                    Z = make_float2(a.x + b.y, b.x - a.y);
                }
                work[pp * SZ + bitrev<LOG2>(k)] = Z;
            }
            __syncthreads();
            fft_dit_r4<SZ, true>(work, npp, tw);
            const int rowlen = 2 * mw;
            for (int idx = tid; idx < npp * rowlen; idx += nthr) {
                const int pp = idx / rowlen;
                const int rem = idx - pp * rowlen;
                const int which = rem >= mw ? 1 : 0;
                const int ox = rem - which * mw;
                const int q = 2 * (p0 + pp) + which;
                if (q >= nrb) continue;
                const int oy = rb0 + q + sy;
                const int xs = ox - sx;
                float val = 0.f;
                if (xs >= 0 && xs < SZ) {
                    const float2 *row = work + pp * SZ;
#pragma unroll
                    for (int u = 0; u < 7; ++u) {
                        const int xx = xs - 3 + u;
                        if (xx >= 0 && xx < SZ) {
                            const float2 g = row[xx];
                            val += fx[u] * (which ? g.y : g.x);
                        }
                    }
                    val *= inv_norm;
                }
                acc += epilogue(mode, val, scale, off + (int64_t)oy * stride + ox, out, data, inverr);
            }
            __syncthreads();
        }
    }
    if (mode == TRH_CHI2) {
        const double tot = block_sum(acc, S.red);
        if (tid == 0) chi2[Jp->slot] = tot;
    }
    __syncthreads();
}

// Pixelized-PSF point-source render (PixelizedPSF.getPointSourcePatch with a
// modelMask, tractor/psf.py:137-190): the stamp, its (0,0) pixel placed at
// window coordinates (sx, sy), Lanczos-3 shifted by (mux, muy); stamp, tmp
// in LDS: tmp holds ph x mw floats.
static __device__ void render_point_one(const trh_fft_job *__restrict__ Jp,
                                 const float *__restrict__ stamp, int ph, int pw,
                                 float *lstamp, float *tmp, RenderScratch &S,
                                 float *__restrict__ out, const float *__restrict__ data,
                                 const float *__restrict__ inverr, double *__restrict__ chi2) {
    const int tid = threadIdx.x, nthr = blockDim.x;
    const int64_t off = Jp->out_offset;
    const int stride = Jp->out_stride;
    const int mw = Jp->mw, mh = Jp->mh, sx = Jp->sx, sy = Jp->sy, mode = Jp->mode;
    const float scale = Jp->scale;
    compute_taps(Jp->mux, Jp->muy, S);
    const float *fx = S.taps;
    const float *fy = S.taps + 7;
    for (int i = tid; i < ph * pw; i += nthr) lstamp[i] = stamp[i];
    __syncthreads();
    // x pass over stamp rows: tmp[r][ox] = sum_u fx[u] stamp[r][ox-sx-3+u]
    for (int idx = tid; idx < ph * mw; idx += nthr) {
        const int r = idx / mw;
        const int ox = idx - r * mw;
        float a = 0.f;
#pragma unroll
        for (int u = 0; u < 7; ++u) {
            const int c = ox - sx - 3 + u;
            if (c >= 0 && c < pw) a += fx[u] * lstamp[r * pw + c];
        }
        tmp[idx] = a;
    }
    __syncthreads();
    double acc = 0.0;
    for (int idx = tid; idx < mh * mw; idx += nthr) {
        const int oy = idx / mw;
        const int ox = idx - oy * mw;
        float a = 0.f;
#pragma unroll
        for (int t = 0; t < 7; ++t) {
            const int r = oy - sy - 3 + t;
            if (r >= 0 && r < ph) a += fy[t] * tmp[r * mw + ox];
        }
        acc += epilogue(mode, a, scale, off + (int64_t)oy * stride + ox, out, data, inverr);
    }
    if (mode == TRH_CHI2) {
        const double tot = block_sum(acc, S.red);
        if (tid == 0) chi2[Jp->slot] = tot;
    }
    __syncthreads();
}

// Grid-stride render kernel over a job list.  `count` (device, may be NULL:
// then njobs) bounds the list; `list` (may be NULL: identity) maps list
// entries to job indices.
template <int SZ, int MODE>
__global__ void __launch_bounds__(1024)
k_render_fft(const trh_fft_job *__restrict__ jobs, const int *__restrict__ list,
             const int *__restrict__ count, int njobs, const float2 *__restrict__ psf_spec,
             const float2 *__restrict__ twg, int ch, int maxrows, float *__restrict__ out,
             const float *__restrict__ data, const float *__restrict__ inverr,
             double *__restrict__ chi2) {
    extern __shared__ float2 smem[];
    __shared__ RenderScratch S;
    float2 *tw = smem;
    float2 *work = tw + SZ / 2;
    float2 *hbuf = work + ch * SZ;
    for (int i = threadIdx.x; i < SZ / 2; i += blockDim.x) tw[i] = twg[i];
    __syncthreads();
    const int n = count ? *count : njobs;
    for (int jj = blockIdx.x; jj < n; jj += gridDim.x) {
        const int jid = list ? list[jj] : jj;
        render_fft_one<SZ, MODE>(jobs + jid, psf_spec, tw, work, hbuf, ch, maxrows, S, out, data,
                                 inverr, chi2);
    }
}

template <int UNUSED = 0>
__global__ void __launch_bounds__(256)
k_render_point(const trh_fft_job *__restrict__ jobs, const int *__restrict__ list,
               const int *__restrict__ count, int njobs, const float *__restrict__ stamp,
               int ph, int pw, int maxw, float *__restrict__ out,
               const float *__restrict__ data, const float *__restrict__ inverr,
               double *__restrict__ chi2) {
    extern __shared__ float2 smem[];
    __shared__ RenderScratch S;
    float *lstamp = reinterpret_cast<float *>(smem);
    float *tmp = lstamp + ((ph * pw + 3) & ~3);
    const int n = count ? *count : njobs;
    for (int jj = blockIdx.x; jj < n; jj += gridDim.x) {
        const int jid = list ? list[jj] : jj;
        if (jobs[jid].mw > maxw) continue;  // host guarantees maxw >= every mw
        render_point_one(jobs + jid, stamp, ph, pw, lstamp, tmp, S, out, data, inverr, chi2);
    }
}

__host__ __device__ inline size_t point_lds_bytes(int ph, int pw, int maxw) {
    return sizeof(float) * ((((size_t)ph * pw + 3) & ~(size_t)3) + (size_t)ph * maxw);
}

}  // namespace trh
