// FFT-path and pixelized-PSF render device code, shared by the C-ABI batch
// entry points and the batched fitter.  gfx950 only.
#pragma once
#include "fft_common.h"
#include "../../include/tractor_hip.h"

namespace trh {

// -2 pi^2 log2(e): exp(-2 pi^2 q) = exp2(C * q)
#define TRH_NEG2PI2_LOG2E (-28.47765484177155)

// Geometry of the wavefront-level FFT: a wave holds NSEQ sequences of length
// SZ, each lane R elements (element index sl + SEQL * r, sl = lane % SEQL).
template <int SZ> struct WaveGeom {
    static constexpr int R = SZ >= 64 ? SZ / 64 : 1;
    static constexpr int SEQL = SZ >= 64 ? 64 : SZ;
    static constexpr int NSEQ = 64 / SEQL;
    static constexpr int NV = SZ / 2 + 1;
    static constexpr int LBUF = SZ + SZ / 32;   // padded per-sequence line buffer (float2)
    // The y-filtered half spectrum between the two passes holds up to SZ rows
    // of NV columns: in LDS for SZ <= 128; from SZ = 256 on (264 KB and more)
    // it lives in a per-block slot of global memory, row stride HS = NV + 1
    // so that rows start 16-byte aligned and phase A stores column pairs.
    static constexpr bool GH = SZ >= 256;
    static constexpr int HS = GH ? NV + 1 : NV;
};

// Floats of one block's global half-spectrum slot (SZ rows of HS columns).
__host__ __device__ inline size_t fft_hbuf_slot_floats(int sz) {
    return (size_t)2 * sz * (sz / 2 + 2);
}

// Dynamic LDS of k_render_fft<SZ> with `nwaves` waves per block and `maxrows`
// rows of the y-filtered half spectrum: per-wave line buffers (2 sequences
// per wave iteration) + rows (maxrows = 0 when the rows live in global
// memory, SZ >= 256, which keep the 2 SZ Lanczos responses there instead).
__host__ __device__ inline size_t fft_lds_bytes(int sz, int nwaves, int maxrows) {
    const int nseq = sz >= 64 ? 1 : 64 / sz;
    const int lbuf = sz + sz / 32;
    const size_t ltab = sz >= 128 ? 2 * (size_t)sz : 0;   // Lanczos responses
    return sizeof(float2) * ((size_t)nwaves * nseq * 2 * lbuf + (size_t)maxrows * (sz / 2 + 1) +
                             ltab);
}

struct RenderScratch {
    float4 comp[TRH_MAXK];
    float taps[14];      // fx[0..6], fy[7..13]
    double dtaps[14];
    double red[16];
};

// Lanczos taps for (mux, muy) into S.taps (all threads call; 1 barrier).
// Lanes 0..13 of wave 0 evaluate the 14 raw taps and normalise them with
// sums read across the wave, so no barrier separates the two steps.
__device__ __forceinline__ void compute_taps(double mux, double muy, RenderScratch &S) {
    const int tid = threadIdx.x;
    if (tid < 64) {
        const double dx = tid < 7 ? mux : muy;
        const int i = tid % 7;
        double val = 0.0;
        if (tid >= 14) {
        } else if (dx == 0.0) {
            val = (i == 3) ? 1.0 : 0.0;
        } else {
            const bool zero = dx > 0.0 ? (i == 6) : (i == 0);
            if (!zero) {
                const double pix = M_PI * (i - 3.0 + dx);
                val = 3.0 * sin(pix) * sin(pix / 3.0) / (pix * pix);
            }
        }
        const int g = tid < 7 ? 0 : 7;
        double s = 0.0;
        // same summation order as lanczos_filter_3 (mp_fourier.i:380-400)
        for (int k = 0; k < 7; ++k) s += __shfl(val, g + k);
        if (tid < 14) {
            S.dtaps[tid] = val;
            S.taps[tid] = (dx == 0.0) ? (float)val : (float)(val / s);
        }
    }
    __syncthreads();
}

// A render's chi^2 into *dst: the waves' sums added in wave order by thread
// 0 (block_sum's order) after one barrier, which also ends every wave's
// reads of the render's LDS.  S.red is written again only after the next
// render's first barrier, which thread 0 reaches after this read.
__device__ __forceinline__ void chi2_out(double acc, RenderScratch &S, double *dst) {
    const double v = wave_sum(acc);
    if ((threadIdx.x & 63) == 0) S.red[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double tot = 0.0;
        const int nw = blockDim.x >> 6;
        for (int i = 0; i < nw; ++i) tot += S.red[i];
        *dst = tot;
    }
}

// Image pixels of the FFT renders (data, inverr) are loaded non-temporally
// so they do not push the shared PSF spectrum out of L2 (measured on the
// sz = 128 chi^2 renders: 1.73 -> 1.42 GB read per LM iteration of 4096
// sources, against 1.14 GB of pixels).
__device__ __forceinline__ float ld_stream(const float *p) { return __builtin_nontemporal_load(p); }

// Whether the buffers a job's epilogue mode writes and reads were passed:
// out for TRH_STORE / TRH_ADD, data, inverr and chi2 slots for TRH_CHI2;
// unknown modes have none.
__device__ __forceinline__ bool epilogue_buffers_ok(int mode, const float *out, const float *data,
                                                    const float *inverr, const double *chi2) {
    if (mode == TRH_CHI2) return data && inverr && chi2;
    return (mode == TRH_STORE || mode == TRH_ADD) && out;
}

// Kernel mode of the chi^2 renders whose unit-flux model is the FFT render
// plus a pixel-space part already rendered into `base` (the MoG half of a
// hybrid-PSF galaxy): the job's window of it starts at base + pad0 - 1 with
// row stride mw when pad0 > 0; pad0 = 0: none.  Used by the batched fitter.
#define TRH_CHI2_BASE 18

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
        const float chi = (ld_stream(data + idx) - mod) * ld_stream(inverr + idx);
        return (double)chi * (double)chi;
    }
    return 0.0;
}

__device__ __forceinline__ int lpad(int i) { return i + (i >> 5); }

__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Exchange with lane (lane ^ D).  D = 1, 2: one DPP quad permutation;
// D = 4: row half-mirror then quad reverse; D = 8: row half-mirror then row
// mirror (DPP moves are VALU ops, no LDS round trip); D = 16: ds_swizzle
// (xor within 32-lane halves); D = 32: ds_bpermute.
template <int D>
__device__ __forceinline__ float xlane(float v) {
    const int i = __float_as_int(v);
    if constexpr (D == 1) {
        return __int_as_float(__builtin_amdgcn_update_dpp(0, i, 0xB1, 0xF, 0xF, true));
    } else if constexpr (D == 2) {
        return __int_as_float(__builtin_amdgcn_update_dpp(0, i, 0x4E, 0xF, 0xF, true));
    } else if constexpr (D == 4) {
        const int t = __builtin_amdgcn_update_dpp(0, i, 0x141, 0xF, 0xF, true);
        return __int_as_float(__builtin_amdgcn_update_dpp(0, t, 0x1B, 0xF, 0xF, true));
    } else if constexpr (D == 8) {
        const int t = __builtin_amdgcn_update_dpp(0, i, 0x141, 0xF, 0xF, true);
        return __int_as_float(__builtin_amdgcn_update_dpp(0, t, 0x140, 0xF, 0xF, true));
    } else if constexpr (D == 16) {
        return __int_as_float(__builtin_amdgcn_ds_swizzle(i, (0x10 << 10) | 0x1F));
    } else {
        return __shfl_xor(v, 32, 64);
    }
}

// Transpose lane bit B with a register bit: for the register pair (a, b),
// a' = a on lanes with bit B clear and b[lane - 2^B] on lanes with it set;
// b' = a[lane + 2^B] on lanes with bit B clear and b on lanes with it set.
// Afterwards a holds the elements whose index bit moved from the lane to the
// register was 0 and b those where it was 1, and lane bit B carries what the
// register bit carried.  B = 5, 4: gfx950's v_permlane32_swap /
// v_permlane16_swap (one instruction for both registers); B = 3, 2: two DPP
// row shifts, each writing only the 4-lane banks of one half.
template <int B>
__device__ __forceinline__ void lane_transpose(float &a, float &b) {
    if constexpr (B == 5 || B == 4) {
        const auto r = B == 5 ? __builtin_amdgcn_permlane32_swap(__float_as_uint(a),
                                                                  __float_as_uint(b), false, false)
                              : __builtin_amdgcn_permlane16_swap(__float_as_uint(a),
                                                                  __float_as_uint(b), false, false);
        a = __uint_as_float(r[0]);
        b = __uint_as_float(r[1]);
    } else {
        static_assert(B == 3 || B == 2, "lane bits 2..5");
        constexpr int D = 1 << B;
        constexpr int hi = B == 3 ? 0xC : 0xA, lo = B == 3 ? 0x3 : 0x5;   // bank masks
        const int ia = __float_as_int(a), ib = __float_as_int(b);
        a = __int_as_float(__builtin_amdgcn_update_dpp(ia, ib, 0x110 + D, 0xF, hi, false));
        b = __int_as_float(__builtin_amdgcn_update_dpp(ib, ia, 0x100 + D, 0xF, lo, false));
    }
}

// Per-lane twiddles of the wave FFT, loaded once per block.  In-register
// stages (span d >= 64) need tw[(sl + 64 q) * SZ/(2d)], q < d/64; the stages
// of span d = 32..4, done in registers after a lane transpose, need
// tw[(sl & (d-1)) * SZ/(2d)] (the lane bits below d are still the element
// bits below d).  The stages of span 2 and 1 stay across lanes: t = s x +
// x[lane ^ d] with s = +1 on the lower lane of each pair and -1 on the upper
// gives x_lo + x_hi on the lower lane and x_lo - x_hi on the upper, whose
// twiddle tw[(sl & (d-1)) * SZ/(2d)] is 1 at span 1 and 1 or i at span 2 (i on
// the odd pairs, `rot`): neither stage multiplies, and no branch diverges.
template <int SZ> struct WaveTw {
    static constexpr int NREG = SZ >= 128 ? SZ / 64 - 1 : 0;
    static constexpr int NLANE = (SZ >= 64 ? 6 : Log2<SZ>::v) - 2;
    float2 reg[NREG > 0 ? NREG : 1];
    float2 lane[NLANE];   // spans 32 (or 16) .. 4
    float sgn[2];         // spans 2, 1
    bool rot;             // span 2: upper lane of an odd pair (times i)
    __device__ __forceinline__ void load(const float2 *tw, int sl) {
        int n = 0;
#pragma unroll
        for (int d = SZ / 2; d >= 64; d >>= 1)
#pragma unroll
            for (int q = 0; q < d / 64; ++q) reg[n++] = tw[(sl + 64 * q) * (SZ / (2 * d))];
        n = 0;
#pragma unroll
        for (int d = (SZ >= 64 ? 32 : SZ / 2); d >= 4; d >>= 1)
            lane[n++] = tw[(sl & (d - 1)) * (SZ / (2 * d))];
        sgn[0] = (sl & 2) ? -1.f : 1.f;
        sgn[1] = (sl & 1) ? -1.f : 1.f;
        rot = (sl & 3) == 3;
    }
};

// DIF butterfly between the register pair (a: element bit 0, b: bit 1) after
// the transpose of lane bit B: a + b, (a - b) w.
template <int B>
__device__ __forceinline__ void transposed_butterfly(float2 &a, float2 &b, float2 w) {
    lane_transpose<B>(a.x, b.x);
    lane_transpose<B>(a.y, b.y);
    const float2 s = cadd(a, b), d = csub(a, b);
    a = s;
    b = cmul(d, w);
}

// The stage of span 2^B (B = 2..5).  The register bit it uses: bit 0 of the
// register index for SZ >= 128 (it held element bit 6, finished by then),
// the sequence bit c for SZ <= 64.
template <int SZ, int B, int C>
__device__ __forceinline__ void idif_transposed_stage(float2 (&x)[C][WaveGeom<SZ>::R], float2 w) {
    constexpr int R = WaveGeom<SZ>::R;
    if constexpr (SZ >= 128) {
#pragma unroll
        for (int c = 0; c < C; ++c)
#pragma unroll
            for (int m = 0; m < R / 2; ++m) transposed_butterfly<B>(x[c][2 * m], x[c][2 * m + 1], w);
    } else {
        static_assert(C == 2, "the sequence bit pairs two sequences");
#pragma unroll
        for (int r = 0; r < R; ++r) transposed_butterfly<B>(x[0][r], x[1][r], w);
    }
}

// One cross-lane DIF stage of span D = 2 or 1 on every register: lower lane
// x_lo + x_hi, upper lane (x_lo - x_hi) w with w = i on the `rot` lanes of
// span 2 and 1 elsewhere (see WaveTw).
template <int SZ, int D, int C>
__device__ __forceinline__ void idif_lane_stage(float2 (&x)[C][WaveGeom<SZ>::R], float s,
                                                bool rot) {
#pragma unroll
    for (int c = 0; c < C; ++c)
#pragma unroll
        for (int r = 0; r < WaveGeom<SZ>::R; ++r) {
            const float px = xlane<D>(x[c][r].x), py = xlane<D>(x[c][r].y);
            const float tx = fmaf(s, x[c][r].x, px), ty = fmaf(s, x[c][r].y, py);
            if constexpr (D == 2)
                x[c][r] = rot ? make_float2(-ty, tx) : make_float2(tx, ty);
            else
                x[c][r] = make_float2(tx, ty);
        }
}

// Inverse (exp(+2 pi i jk/SZ)) radix-2 decimation-in-frequency FFT of C
// independent sequences held by the lanes (interleaved for latency hiding).
// In: register (c, r) of lane l holds element sl + SEQL r of sequence
// c NSEQ + l / SEQL (sl = l % SEQL).  Stages of span >= 64 pair registers of
// a lane; spans 32..4 first transpose their lane bit into a register bit
// (lane_transpose) and then pair registers; spans 2 and 1 pair lanes.  Out:
// see wave_fft_out; the elements are in bit-reversed order.  No LDS, no
// barriers.
template <int SZ, int C>
__device__ __forceinline__ void wave_idif(float2 (&x)[C][WaveGeom<SZ>::R], const WaveTw<SZ> &T) {
    constexpr int R = WaveGeom<SZ>::R;
    int n = 0;
#pragma unroll
    for (int d = SZ / 2; d >= 64; d >>= 1) {
        const int rd = d / 64;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            if (r & rd) continue;
            const float2 w = T.reg[n + (r & (rd - 1))];
#pragma unroll
            for (int c = 0; c < C; ++c) {
                const float2 a = x[c][r], b = x[c][r + rd];
                x[c][r] = cadd(a, b);
                x[c][r + rd] = cmul(csub(a, b), w);
            }
        }
        n += rd;
    }
    constexpr int TOP = SZ >= 64 ? 5 : Log2<SZ>::v - 1;   // highest lane element bit
    if constexpr (TOP >= 5) idif_transposed_stage<SZ, 5, C>(x, T.lane[TOP - 5]);
    idif_transposed_stage<SZ, 4, C>(x, T.lane[TOP - 4]);
    idif_transposed_stage<SZ, 3, C>(x, T.lane[TOP - 3]);
    idif_transposed_stage<SZ, 2, C>(x, T.lane[TOP - 2]);
    idif_lane_stage<SZ, 2, C>(x, T.sgn[0], T.rot);
    idif_lane_stage<SZ, 1, C>(x, T.sgn[1], false);
}

// Where wave_idif leaves register (c, r) of lane l: the sequence slot
// (c NSEQ + sq of the input numbering) and the element index e (the output
// index is bitrev(e)).  Element bits 0, 1 stay on lane bits 0, 1; element bit
// 2 ends on the register bit of the transposes; element bits 3..5 (3..4 for
// SZ = 32) on lane bits 2..4 (2..3); what the register bit held first
// (element bit 6, or the sequence bit c) on the next lane bit.
template <int SZ>
__device__ __forceinline__ void wave_fft_out(int l, int c, int r, int &slot, int &e) {
    if constexpr (SZ >= 128) {
        slot = c;
        e = (l & 3) | ((r & 1) << 2) | (((l >> 2) & 15) << 3) | ((r >> 1) << 7);
    } else if constexpr (SZ == 64) {
        slot = l >> 5;
        e = (l & 3) | (c << 2) | (((l >> 2) & 7) << 3);
    } else {
        static_assert(SZ == 32, "SZ 32..512");
        slot = ((l >> 4) & 1) * 2 + (l >> 5);
        e = (l & 3) | (c << 2) | (((l >> 2) & 3) << 3);
    }
}

// Frequency response of the 7-tap Lanczos filter f at DFT index k of an SZ
// point transform: sum_u f[u] exp(+2 pi i k (u - 3) / SZ).  Filtering the
// spectrum by it is the circular convolution h[x] = sum_u f[u] g[x - 3 + u],
// equal to the reference's truncated one wherever all 7 taps fall inside the
// box.  tw: exp(+2 pi i m / SZ), m < SZ/2.
template <int SZ>
__device__ __forceinline__ float2 lanczos_response(const float *f, int k,
                                                   const float2 *__restrict__ tw) {
    float2 acc = make_float2(0.f, 0.f);
#pragma unroll
    for (int u = 0; u < 7; ++u) {
        const int m = (k * (u - 3)) & (SZ - 1);
        const float2 e = m < SZ / 2 ? tw[m] : make_float2(-tw[m - SZ / 2].x, -tw[m - SZ / 2].y);
        acc.x += f[u] * e.x;
        acc.y += f[u] * e.y;
    }
    return acc;
}

// Render one FFT job (see trh_fft_job).  Every thread of the block calls this.
// Phase A: each wave takes CW columns v_i of the half spectrum at a time: MoG
// Fourier transform x PSF spectrum for the SZ frequencies w_j, inverse FFT
// along w in registers, y-Lanczos onto the G rows this block needs ->
// hbuf[row][i].  Phase B: each wave takes CW pairs of G rows: the two half
// spectra packed into one complex sequence, inverse FFT along v, x-Lanczos,
// crop and epilogue.  Only the hand-off between the phases needs the block.
// psf_spec: [psf][SZ/2+1][SZ] complex (column-major half spectrum).
constexpr int CW = 2;

template <int SZ, int MODE, bool FD>
__device__ void render_fft_one(const trh_fft_job *__restrict__ Jp,
                               const float2 *__restrict__ psf_spec, const WaveTw<SZ> &T,
                               const float2 *__restrict__ twg, float2 *ltab,
                               float2 *lbuf, float2 *hbuf, int maxrows, RenderScratch &S,
                               float *__restrict__ out, const float *__restrict__ data,
                               const float *__restrict__ inverr, double *__restrict__ chi2,
                               const float *__restrict__ base) {
    using G = WaveGeom<SZ>;
    constexpr int NV = G::NV, R = G::R, SEQL = G::SEQL, NSEQ = G::NSEQ, HS = G::HS;
    constexpr int LOG2 = Log2<SZ>::v;
    constexpr int PER = NSEQ * CW;   // sequences per wave iteration
    static_assert(!G::GH || (NSEQ == 1 && CW == 2), "global rows store column pairs");
    const int tid = threadIdx.x;
    const int nthr = blockDim.x;
    // the wave index is the same on every lane: scalar, so everything
    // derived from it (column / row-pair numbers, their addresses) is too
    const int lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6),
              nw = nthr >> 6;
    const int sl = lane & (SEQL - 1), sq = lane / SEQL;
    float2 *lbw = lbuf + (size_t)wave * PER * G::LBUF;
    const int64_t off = Jp->out_offset;
    const int stride = Jp->out_stride;
    const int mw = Jp->mw, mh = Jp->mh, sx = Jp->sx, sy = Jp->sy;
    // MODE >= 0: epilogue fixed at compile time; -1: read from the job;
    // TRH_CHI2_BASE: chi^2 with the job's base plane added to the model
    constexpr bool BASE = MODE == TRH_CHI2_BASE;
    constexpr int EMODE = BASE ? TRH_CHI2 : MODE;
    const int K = min(max(Jp->K, 0), TRH_MAXK), mode = EMODE >= 0 ? EMODE : Jp->mode;
    const float *__restrict__ bp = BASE && Jp->pad0 > 0 ? base + (Jp->pad0 - 1) : nullptr;
    // the base value of window pixel (oy, ox)
    auto basev = [&](int oy, int ox) -> float {
        if constexpr (BASE) return bp ? bp[(int64_t)oy * mw + ox] : 0.f;
        return 0.f;
    };
    // a job whose epilogue has no buffer to use renders nothing (the same
    // decision in every thread of the block, so no barrier is left waiting)
    if (!epilogue_buffers_ok(mode, out, data, inverr, chi2)) return;
    const float scale = Jp->scale;
    const float2 *__restrict__ P = psf_spec + (size_t)Jp->psf * SZ * NV;

    if (tid < K) {
        const double Cc = TRH_NEG2PI2_LOG2E;
        S.comp[tid] = make_float4(Jp->amp[tid], (float)(Cc * Jp->var[tid][0]),
                                  (float)(2.0 * Cc * Jp->var[tid][1]),
                                  (float)(Cc * Jp->var[tid][2]));
    }
    compute_taps(Jp->mux, Jp->muy, S);
    const float *fx = S.taps;
    const float *fy = S.taps + 7;
    const float inv_norm = 1.0f / (float)(SZ * SZ);
    // FD: the Lanczos shift as a spectrum filter (the window keeps every tap
    // inside the box): y response at the lane's row frequencies (phase A),
    // x response at its column frequencies (phase B), both at sl + SEQL r.
    // Held in registers up to SZ = 64; from 128 on in the LDS table ltab (y at
    // [0, SZ), x at [SZ, 2 SZ)), filled by the block per job (registers there
    // cost more in spills than the table's reads: 69 -> 57 ns per sz = 128
    // chi^2 render).
    constexpr bool LREG = FD && SZ <= 64;
    float2 Ly[LREG ? R : 1], Lx[LREG ? R : 1];
    if constexpr (LREG) {
#pragma unroll
        for (int r = 0; r < R; ++r) {
            Ly[r] = lanczos_response<SZ>(fy, sl + SEQL * r, twg);
            Lx[r] = lanczos_response<SZ>(fx, sl + SEQL * r, twg);
        }
    } else if constexpr (FD) {
        for (int k = tid; k < 2 * SZ; k += nthr)
            ltab[k] = lanczos_response<SZ>(k < SZ ? fy : fx, k & (SZ - 1), twg);
        __syncthreads();
    }

    double acc = 0.0;
    // G rows [ylo, yhi) land in the window; other window rows are empty.
    const int ylo = max(0, -sy);
    const int yhi = mw > 0 ? min(SZ, mh - sy) : ylo;   // a window of width 0 has no pixels
    if (mode != TRH_ADD) {
        const int top = min(mh, max(0, sy));               // rows [0, top)
        const int bot = max(top, min(mh, SZ + sy));        // rows [bot, mh)
        const int nempty = (top + (mh - bot)) * mw;
        for (int idx = tid; idx < nempty; idx += nthr) {
            const int r = idx / mw;
            const int ox = idx - r * mw;
            const int oy = r < top ? r : bot + (r - top);
            acc += epilogue(mode, basev(oy, ox), scale, off + (int64_t)oy * stride + ox, out, data,
                            inverr);
        }
    }

    for (int rb0 = ylo; rb0 < yhi; rb0 += maxrows) {
        const int nrb = min(maxrows, yhi - rb0);
        // ---- phase A: columns of the half spectrum
        for (int i0 = wave * PER; i0 < NV; i0 += nw * PER) {
            float2 x[CW][R];
            // MoG Fourier transform: components outermost, so each one is read
            // from LDS once for all CW x R frequencies of the lane.  The PSF
            // spectrum loads are issued first; the exp work covers their latency.
            float v2[CW], vw[CW][R], F[CW][R], w[R], w2[R];
            float2 pv[CW][R];
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int j = sl + SEQL * r;
                w[r] = (float)(j < SZ / 2 ? j : j - SZ) * (1.0f / SZ);
                w2[r] = w[r] * w[r];
            }
#pragma unroll
            for (int c = 0; c < CW; ++c) {
                const int i = i0 + c * NSEQ + sq;
                const int ic = i < NV ? i : NV - 1;
                const float2 *__restrict__ Pc = P + (size_t)ic * SZ;
#pragma unroll
                for (int r = 0; r < R; ++r) pv[c][r] = Pc[sl + SEQL * r];
                const float v = (float)ic * (1.0f / SZ);
                v2[c] = v * v;
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    vw[c][r] = v * w[r];
                    F[c][r] = 0.f;
                }
            }
            for (int k = 0; k < K; ++k) {
                const float4 cp = S.comp[k];
#pragma unroll
                for (int c = 0; c < CW; ++c)
#pragma unroll
                    for (int r = 0; r < R; ++r)
                        F[c][r] += cp.x * __builtin_amdgcn_exp2f(cp.y * v2[c] + cp.z * vw[c][r] +
                                                                 cp.w * w2[r]);
            }
#pragma unroll
            for (int c = 0; c < CW; ++c) {
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    float2 p = pv[c][r];
                    if constexpr (LREG) p = cmul(p, Ly[r]);
                    else if constexpr (FD) p = cmul(p, ltab[sl + SEQL * r]);
                    x[c][r] = make_float2(F[c][r] * p.x, F[c][r] * p.y);
                }
            }
            wave_idif<SZ, CW>(x, T);
#pragma unroll
            for (int c = 0; c < CW; ++c)
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    int slot, e;
                    wave_fft_out<SZ>(lane, c, r, slot, e);
                    lbw[slot * G::LBUF + lpad(bitrev<LOG2>(e))] = x[c][r];
                }
            wave_sync();
            if constexpr (G::GH) {
                // columns i0, i0 + 1 of each row as one 16-byte store (the
                // pad column HS - 1 takes the clamped duplicate when i0 = NV - 1)
                for (int q = sl; q < nrb; q += SEQL) {
                    const int yr = rb0 + q;
                    float2 a[CW];
#pragma unroll
                    for (int c = 0; c < CW; ++c) {
                        const float2 *lb = lbw + c * G::LBUF;
                        if constexpr (FD) {
                            a[c] = lb[lpad(yr)];
                        } else {
                            a[c] = make_float2(0.f, 0.f);
#pragma unroll
                            for (int t = 0; t < 7; ++t) {
                                const int ys = yr - 3 + t;
                                if (ys >= 0 && ys < SZ) {
                                    const float2 h = lb[lpad(ys)];
                                    a[c].x += fy[t] * h.x;
                                    a[c].y += fy[t] * h.y;
                                }
                            }
                        }
                    }
                    *reinterpret_cast<float4 *>(hbuf + (size_t)q * HS + i0) =
                        make_float4(a[0].x, a[0].y, a[1].x, a[1].y);
                }
            } else {
#pragma unroll
                for (int c = 0; c < CW; ++c) {
                    const int i = i0 + c * NSEQ + sq;
                    if (i >= NV) continue;
                    const float2 *lb = lbw + (c * NSEQ + sq) * G::LBUF;
                    for (int q = sl; q < nrb; q += SEQL) {
                        const int yr = rb0 + q;
                        float2 a;
                        if constexpr (FD) {
                            a = lb[lpad(yr)];
                        } else {
                            a = make_float2(0.f, 0.f);
#pragma unroll
                            for (int t = 0; t < 7; ++t) {
                                const int ys = yr - 3 + t;
                                if (ys >= 0 && ys < SZ) {
                                    const float2 h = lb[lpad(ys)];
                                    a.x += fy[t] * h.x;
                                    a.y += fy[t] * h.y;
                                }
                            }
                        }
                        hbuf[q * HS + i] = a;
                    }
                }
            }
            wave_sync();
        }
        __syncthreads();
        // ---- phase B: two real rows per complex inverse FFT
        const int npairs = (nrb + 1) / 2;
        for (int p0 = wave * PER; p0 < npairs; p0 += nw * PER) {
            float2 x[CW][R];
#pragma unroll
            for (int c = 0; c < CW; ++c) {
                const int pp = p0 + c * NSEQ + sq;
                const bool act = pp < npairs;
                const int qa = 2 * (act ? pp : 0);
                const bool hasb = act && (qa + 1) < nrb;
                const float2 *A = hbuf + (size_t)qa * HS;
                const float2 *B = A + HS;
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    // Z_k = a_k + i b_k from the half spectra of rows a, b:
                    // k > SZ/2 takes the conjugates at SZ - k, and irfft
                    // ignores the imaginary parts at DC and Nyquist (factor
                    // f on the imaginary parts; no divergent branch)
                    const int k = sl + SEQL * r;
                    const int kk = k <= SZ / 2 ? k : SZ - k;
                    const float f = (k == 0 || k == SZ / 2) ? 0.f : (k < SZ / 2 ? 1.f : -1.f);
                    float2 a = make_float2(0.f, 0.f), b = make_float2(0.f, 0.f);
                    if (act) a = A[kk];
                    if (hasb) b = B[kk];
                    float2 Z = make_float2(a.x - f * b.y, f * a.y + b.x);
                    if constexpr (LREG) Z = cmul(Z, Lx[r]);
                    else if constexpr (FD) Z = cmul(Z, ltab[SZ + k]);
                    x[c][r] = Z;
                }
            }
            wave_idif<SZ, CW>(x, T);
#pragma unroll
            for (int c = 0; c < CW; ++c)
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    int slot, e;
                    wave_fft_out<SZ>(lane, c, r, slot, e);
                    lbw[slot * G::LBUF + lpad(bitrev<LOG2>(e))] = x[c][r];
                }
            wave_sync();
            // G rows a, b (real, imaginary part) at column xs of the box
            auto pair_value = [&](const float2 *lb, int xs) -> float2 {
                float va = 0.f, vb = 0.f;
                if constexpr (FD) {
                    // every xs of the window is inside [3, SZ - 4] on this path
                    const float2 g = lb[lpad(xs)];
                    va = g.x;
                    vb = g.y;
                } else if (xs >= 0 && xs < SZ) {
#pragma unroll
                    for (int u = 0; u < 7; ++u) {
                        const int xx = xs - 3 + u;
                        if (xx >= 0 && xx < SZ) {
                            const float2 g = lb[lpad(xx)];
                            va += fx[u] * g.x;
                            vb += fx[u] * g.y;
                        }
                    }
                }
                return make_float2(va * inv_norm, vb * inv_norm);
            };
            if (EMODE == TRH_CHI2 && mw <= SZ) {
                // (compile-time chi^2 kernels only: in the kernel that reads
                // the mode from the job it would raise every mode's register
                // pressure) the data / inverr loads of TB columns per lane
                // (both sequences, both rows) are issued before the arithmetic:
                // indices clamped into the window (mw <= SZ: SZ / SEQL columns
                // per lane at most; true of every modelMask window, whose
                // halfsize covers it), the clamped pixels' terms dropped
                constexpr int NT = SZ / SEQL, TB = SZ >= 256 ? 1 : (NT < 2 ? NT : 2);
                int64_t rowa[CW], rowb[CW];
                int oya[CW];
                bool act[CW], hasb[CW];
#pragma unroll
                for (int c = 0; c < CW; ++c) {
                    const int pp = p0 + c * NSEQ + sq;
                    act[c] = pp < npairs;
                    const int qa = act[c] ? 2 * pp : 0;
                    hasb[c] = act[c] && (qa + 1) < nrb;
                    oya[c] = rb0 + qa + sy;
                    rowa[c] = off + (int64_t)oya[c] * stride;
                    rowb[c] = hasb[c] ? rowa[c] + stride : rowa[c];
                }
#pragma unroll
                for (int t0 = 0; t0 < NT; t0 += TB) {
                    float dv[CW][TB][2], iv[CW][TB][2], bv[BASE ? CW : 1][BASE ? TB : 1][2];
#pragma unroll
                    for (int c = 0; c < CW; ++c)
#pragma unroll
                        for (int tt = 0; tt < TB; ++tt) {
                            const int oxc = min(sl + SEQL * (t0 + tt), mw - 1);
                            dv[c][tt][0] = ld_stream(data + rowa[c] + oxc);
                            iv[c][tt][0] = ld_stream(inverr + rowa[c] + oxc);
                            dv[c][tt][1] = ld_stream(data + rowb[c] + oxc);
                            iv[c][tt][1] = ld_stream(inverr + rowb[c] + oxc);
                            if constexpr (BASE) {
                                bv[c][tt][0] = basev(oya[c], oxc);
                                bv[c][tt][1] = basev(hasb[c] ? oya[c] + 1 : oya[c], oxc);
                            }
                        }
#pragma unroll
                    for (int c = 0; c < CW; ++c) {
                        const float2 *lb = lbw + (c * NSEQ + sq) * G::LBUF;
#pragma unroll
                        for (int tt = 0; tt < TB; ++tt) {
                            const int ox = sl + SEQL * (t0 + tt);
                            float2 v = pair_value(lb, min(ox, mw - 1) - sx);
                            if constexpr (BASE) {
                                v.x += bv[c][tt][0];
                                v.y += bv[c][tt][1];
                            }
                            const float ca = (dv[c][tt][0] - v.x * scale) * iv[c][tt][0];
                            const float cb = (dv[c][tt][1] - v.y * scale) * iv[c][tt][1];
                            if (act[c] && ox < mw) acc += (double)ca * (double)ca;
                            if (hasb[c] && ox < mw) acc += (double)cb * (double)cb;
                        }
                    }
                }
            } else {
#pragma unroll
                for (int c = 0; c < CW; ++c) {
                    const int pp = p0 + c * NSEQ + sq;
                    if (pp >= npairs) continue;
                    const int qa = 2 * pp;
                    const bool hasb = (qa + 1) < nrb;
                    const float2 *lb = lbw + (c * NSEQ + sq) * G::LBUF;
                    const int oya = rb0 + qa + sy;
                    for (int ox = sl; ox < mw; ox += SEQL) {
                        const float2 v = pair_value(lb, ox - sx);
                        acc += epilogue(mode, v.x + basev(oya, ox), scale,
                                        off + (int64_t)oya * stride + ox, out, data, inverr);
                        if (hasb)
                            acc += epilogue(mode, v.y + basev(oya + 1, ox), scale,
                                            off + (int64_t)(oya + 1) * stride + ox, out, data,
                                            inverr);
                    }
                }
            }
            wave_sync();
        }
        // The next pass's phase A overwrites the rows.  After the last pass
        // the next render's first LDS writes before a barrier are S.comp
        // (read in phase A only) and S.taps (read in phase B by the spatial
        // path): the last barrier is needed only there, and a chi^2 render
        // has chi2_out's.
        if (rb0 + maxrows < yhi || (!FD && mode != TRH_CHI2)) __syncthreads();
    }
    // No barrier follows the render: its LDS was last read before the
    // barrier that ends phase B, or chi2_out's.
    if (mode == TRH_CHI2) chi2_out(acc, S, chi2 + Jp->slot);
}

// Pixelized-PSF point-source render (PixelizedPSF.getPointSourcePatch with a
// modelMask, tractor/psf.py:137-190): the stamp, its (0,0) pixel placed at
// window coordinates (sx, sy), Lanczos-3 shifted by (mux, muy); stamp, tmp
// in LDS: tmp holds the x-filtered stamp rows over the window columns the
// shifted stamp reaches (at most pw + 6 of them, whatever the window width).
template <int MODE>
__device__ void render_point_one(const trh_fft_job *__restrict__ Jp,
                                 const float *__restrict__ stamp, bool load_stamp, int ph, int pw,
                                 float *lstamp, float *tmp, RenderScratch &S,
                                 float *__restrict__ out, const float *__restrict__ data,
                                 const float *__restrict__ inverr, double *__restrict__ chi2,
                                 const double *__restrict__ chi0) {
    const int tid = threadIdx.x, nthr = blockDim.x;
    const int64_t off = Jp->out_offset;
    const int stride = Jp->out_stride;
    const int mw = Jp->mw, mh = Jp->mh, sx = Jp->sx, sy = Jp->sy;
    const int mode = MODE >= 0 ? MODE : Jp->mode;
    if (!epilogue_buffers_ok(mode, out, data, inverr, chi2)) return;   // see render_fft_one
    const float scale = Jp->scale;
    compute_taps(Jp->mux, Jp->muy, S);
    const float *fx = S.taps;
    const float *fy = S.taps + 7;
    // the stamp stays in LDS while consecutive jobs of the block share it
    if (load_stamp)
        for (int i = tid; i < ph * pw; i += nthr) lstamp[i] = stamp[i];
    __syncthreads();
    // The shifted stamp covers window columns [cx0, cx1) and rows [ry0, ry1)
    // (taps reach 3 pixels past it); the model is exactly 0 elsewhere.
    const int cx0 = max(0, sx - 3), cx1 = min(mw, sx + pw + 3);
    const int ry0 = max(0, sy - 3), ry1 = min(mh, sy + ph + 3);
    const int cw = cx1 - cx0;
    // x pass over the stamp rows the y taps reach, [rlo, rhi), and the covered
    // columns: tmp[ox - cx0][r] = sum_u fx[u] stamp[r][ox-sx-3+u], stored by
    // column (row stride ts odd; cw <= pw + 6 columns).  A thread takes one stamp row and PB
    // consecutive columns (PB + 6 stamp reads for PB outputs); threads run
    // along r, which the stamp's odd row stride spreads over the LDS banks.
    constexpr int PB = 4;
    const int ts = ph | 1;
    const int rlo = max(0, -sy - 3), rhi = min(ph, mh - sy + 3);
    const int nrow = rhi - rlo, ncb = (cw + PB - 1) / PB;
    for (int idx = tid; cw > 0 && idx < nrow * ncb; idx += nthr) {
        const int cb = idx / nrow;
        const int r = rlo + (idx - cb * nrow);
        const int ox0 = cx0 + PB * cb;
        float v[PB + 6];
#pragma unroll
        for (int k = 0; k < PB + 6; ++k) {
            const int c = ox0 - sx - 3 + k;
            v[k] = (c >= 0 && c < pw) ? lstamp[r * pw + c] : 0.f;
        }
#pragma unroll
        for (int j = 0; j < PB; ++j) {
            float a = 0.f;
#pragma unroll
            for (int u = 0; u < 7; ++u) a += fx[u] * v[j + u];
            if (ox0 + j < cx1) tmp[(ox0 + j - cx0) * ts + r] = a;
        }
    }
    __syncthreads();
    double acc = 0.0;
    const int n = mh * mw;
    // the y taps of an output pixel read tmp rows inside [rlo, rhi)
    auto yval = [&](int oy, int ox) -> float {
        float a = 0.f;
        if (ox < cx0 || ox >= cx1 || oy < ry0 || oy >= ry1) return a;
#pragma unroll
        for (int t = 0; t < 7; ++t) {
            const int r = oy - sy - 3 + t;
            if (r >= 0 && r < ph) a += fy[t] * tmp[(ox - cx0) * ts + r];
        }
        return a;
    };
    if (MODE == TRH_CHI2 && chi0) {
        // chi0[pad0] = sum (data inverr)^2 over the whole window (the chi^2 of
        // a zero model), so only the shifted stamp's support is visited:
        // chi^2 = chi0 + sum_support [((d - m scale) ie)^2 - (d ie)^2].
        // A thread takes one column and PB consecutive rows (PB + 6 reads of
        // tmp for PB model pixels; threads run along the columns).
        const int rh = ry1 - ry0, nrb = (rh + PB - 1) / PB;
        const int nit = (cw > 0 && rh > 0) ? cw * nrb : 0;
        for (int idx = tid; idx < nit; idx += nthr) {
            const int rb = idx / cw;
            const int ox = cx0 + (idx - rb * cw);
            const int oy0 = ry0 + PB * rb;
            float dv[PB], iv[PB], v[PB + 6];
#pragma unroll
            for (int k = 0; k < PB; ++k) {
                const int64_t g = off + (int64_t)min(oy0 + k, ry1 - 1) * stride + ox;
                dv[k] = data[g];
                iv[k] = inverr[g];
            }
#pragma unroll
            for (int k = 0; k < PB + 6; ++k) {
                const int r = oy0 - sy - 3 + k;
                v[k] = (r >= rlo && r < rhi) ? tmp[(ox - cx0) * ts + r] : 0.f;
            }
#pragma unroll
            for (int j = 0; j < PB; ++j) {
                float m = 0.f;
#pragma unroll
                for (int t = 0; t < 7; ++t) m += fy[t] * v[j + t];
                const float chi = (dv[j] - m * scale) * iv[j];
                const float c0 = dv[j] * iv[j];
                if (oy0 + j < ry1) acc += (double)chi * (double)chi - (double)c0 * (double)c0;
            }
        }
        if (tid == 0) acc += chi0[Jp->pad0];
    } else if (MODE == TRH_CHI2) {
        // four pixels per thread per step, their data / inverr loads issued
        // together (indices clamped to the last pixel, extra terms dropped)
        for (int base = tid; base < n; base += 4 * nthr) {
            float dv[4], iv[4];
            int oyv[4], oxv[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int idc = min(base + u * nthr, n - 1);
                oyv[u] = idc / mw;
                oxv[u] = idc - oyv[u] * mw;
                const int64_t g = off + (int64_t)oyv[u] * stride + oxv[u];
                // (plain loads: the 13 trial renders of a source reuse these
                // pixels from L2, and this kernel reads no PSF spectrum)
                dv[u] = data[g];
                iv[u] = inverr[g];
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const float chi = (dv[u] - yval(oyv[u], oxv[u]) * scale) * iv[u];
                if (base + u * nthr < n) acc += (double)chi * (double)chi;
            }
        }
    } else {
        for (int idx = tid; idx < n; idx += nthr) {
            const int oy = idx / mw;
            const int ox = idx - oy * mw;
            acc += epilogue(mode, yval(oy, ox), scale, off + (int64_t)oy * stride + ox, out, data,
                            inverr);
        }
    }
    // the barrier of chi2_out, or this one, ends the reads of tmp and S.taps
    // before the next job writes them
    if (mode == TRH_CHI2) chi2_out(acc, S, chi2 + Jp->slot);
    else __syncthreads();
}

// The spatial-Lanczos render (a job whose taps reach past the box edges) as
// a call of its own in the SZ = 256 chi^2 kernels.  Inlined beside
// the spectrum-filter path it sets the register allocation of the whole
// kernel (sz = 256 chi^2: 380 bytes of spills per lane, 56 without it), and
// every render pays for it: as a call, the spectrum-filter renders take
// 182 ns instead of 192 and the spatial ones 175 instead of 186 (none of a
// batched fit's sz = 256 windows needs them).  Below 256 and in the STORE
// kernels the call costs more than it saves (sz = 128 chi^2: -1.4 ns on the
// spectrum-filter path, +17 ns on the spatial path, which 10-13% of a
// batched fit's windows take), so those keep it inline; at SZ = 512 (rare
// jobs) the call's stack frame costs 2 us per launch, so it stays inline
// there too.  The call builds
// its own twiddle registers; results are the same bit for bit.
template <int SZ, int MODE>
constexpr bool SPATIAL_CALL = SZ == 256 && (MODE == TRH_CHI2 || MODE == TRH_CHI2_BASE);

template <int SZ, int MODE>
__device__ __noinline__ void render_fft_spatial(const trh_fft_job *__restrict__ Jp,
                                                const float2 *__restrict__ psf_spec,
                                                const float2 *__restrict__ twg, float2 *ltab,
                                                float2 *lbuf, float2 *hbuf, int maxrows,
                                                RenderScratch *S, float *__restrict__ out,
                                                const float *__restrict__ data,
                                                const float *__restrict__ inverr,
                                                double *__restrict__ chi2,
                                                const float *__restrict__ base) {
    WaveTw<SZ> T;
    T.load(twg, (threadIdx.x & 63) & (WaveGeom<SZ>::SEQL - 1));
    render_fft_one<SZ, MODE, false>(Jp, psf_spec, T, twg, ltab, lbuf, hbuf, maxrows, *S, out,
                                    data, inverr, chi2, base);
}

// Grid-stride render kernel over a job list.  `count` (device, may be NULL:
// then njobs) bounds the list; `list` (may be NULL: identity) maps list
// entries to job indices.  base: MODE TRH_CHI2_BASE only.  blockDim.x is a multiple of 64; dynamic LDS holds
// fft_lds_bytes(SZ, blockDim.x / 64, maxrows).  For SZ >= 256 the rows live
// in hbuf_g, one slot of fft_hbuf_slot_floats(SZ) floats per block (gridDim.x
// slots), and every pass takes all SZ rows.
template <int SZ, int MODE>
__global__ void __launch_bounds__(1024)
k_render_fft(const trh_fft_job *__restrict__ jobs, const int *__restrict__ list,
             const int *__restrict__ count, int njobs, const float2 *__restrict__ psf_spec,
             const float2 *__restrict__ twg, int maxrows, float *__restrict__ out,
             const float *__restrict__ data, const float *__restrict__ inverr,
             double *__restrict__ chi2, float2 *__restrict__ hbuf_g,
             const float *__restrict__ base) {
    extern __shared__ float2 smem[];
    __shared__ RenderScratch S;
    using G = WaveGeom<SZ>;
    float2 *lbuf = smem;
    float2 *hbuf, *ltab = nullptr;
    if constexpr (G::GH) {
        hbuf = hbuf_g + (size_t)blockIdx.x * (fft_hbuf_slot_floats(SZ) / 2);
        ltab = lbuf + (size_t)(blockDim.x >> 6) * G::NSEQ * CW * G::LBUF;
        maxrows = SZ;
    } else {
        hbuf = lbuf + (size_t)(blockDim.x >> 6) * G::NSEQ * CW * G::LBUF;
        if constexpr (SZ >= 128) ltab = hbuf + (size_t)maxrows * G::HS;
    }
    WaveTw<SZ> T;
    T.load(twg, (threadIdx.x & 63) & (G::SEQL - 1));
    const int n = count ? *count : njobs;
    for (int jj = blockIdx.x; jj < n; jj += gridDim.x) {
        const int jid = list ? list[jj] : jj;
        // the spectrum-filter path when every Lanczos tap of every output
        // pixel stays inside the box (rows [3, SZ-3), columns likewise)
        const trh_fft_job *Jp = jobs + jid;
        const int sy = Jp->sy, sx = Jp->sx;
        const int ylo = max(0, -sy), yhi = min(SZ, Jp->mh - sy);
        const bool fd = ylo >= 3 && yhi <= SZ - 3 && -sx >= 3 && Jp->mw - sx <= SZ - 3;
        if (fd)
            render_fft_one<SZ, MODE, true>(Jp, psf_spec, T, twg, ltab, lbuf, hbuf, maxrows, S,
                                           out, data, inverr, chi2, base);
        else if constexpr (SPATIAL_CALL<SZ, MODE>)
            render_fft_spatial<SZ, MODE>(Jp, psf_spec, twg, ltab, lbuf, hbuf, maxrows, &S, out,
                                         data, inverr, chi2, base);
        else
            render_fft_one<SZ, MODE, false>(Jp, psf_spec, T, twg, ltab, lbuf, hbuf, maxrows, S,
                                            out, data, inverr, chi2, base);
    }
}

// MODE: TRH_STORE / TRH_CHI2 fixed at compile time, or -1 (read from each job).
// chi0 (CHI2 only, may be NULL): per-window chi^2 of a zero model, at
// chi0[job.pad0]; with it a chi^2 render visits only the stamp's support.
// Dynamic LDS: point_lds_bytes(ph, pw), for any window size.
template <int MODE>
__global__ void __launch_bounds__(256)
k_render_point(const trh_fft_job *__restrict__ jobs, const int *__restrict__ list,
               const int *__restrict__ count, int njobs, const float *__restrict__ stamp,
               int ph, int pw, float *__restrict__ out,
               const float *__restrict__ data, const float *__restrict__ inverr,
               double *__restrict__ chi2, const double *__restrict__ chi0) {
    extern __shared__ float2 smem[];
    __shared__ RenderScratch S;
    float *lstamp = reinterpret_cast<float *>(smem);
    float *tmp = lstamp + ((ph * pw + 3) & ~3);
    const int n = count ? *count : njobs;
    int cur = -1;   // stamp index held in LDS
    for (int jj = blockIdx.x; jj < n; jj += gridDim.x) {
        const int jid = list ? list[jj] : jj;
        // the job's psf field selects one of the stamps at stamp + psf * ph * pw
        const int psf = jobs[jid].psf;
        render_point_one<MODE>(jobs + jid, stamp + (size_t)psf * ph * pw, psf != cur, ph, pw,
                               lstamp, tmp, S, out, data, inverr, chi2, chi0);
        cur = psf;
    }
}

__host__ __device__ inline size_t point_lds_bytes(int ph, int pw) {
    return sizeof(float) * ((((size_t)ph * pw + 3) & ~(size_t)3) + (size_t)(ph | 1) * (pw + 6));
}

}  // namespace trh
