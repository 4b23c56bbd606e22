// Surrogate CNN kernels for gfx950: the learned initialiser that proposes each
// source's type and starting parameters from a 32x32 cutout.
//
// Network: four 3x3 / stride-2 / pad-1 convolutions with 32 output channels
// and ReLU (32x32x1 -> 16x16x32 -> 8x8x32 -> 4x4x32 -> 2x2x32), global average
// pool, MLP 32 -> 64 (ReLU) -> 9 (6 parameters + 3 type logits).
//
// Activations are NHWC bfloat16; every convolution (forward, backward-data,
// backward-weight) is an implicit GEMM on v_mfma_f32_32x32x16_bf16 with FP32
// accumulation: one wavefront owns a 32 (pixels or output channels) x 32 tile,
// 16 of the reduction index per instruction.  Operand lane layout
// (32x32x16): lane l supplies A[l % 32][8*(l/32) .. 8*(l/32)+7] and
// B[8*(l/32) .. +7][l % 32]; accumulator element i of lane l is
// D[8*(i/4) + 4*(l/32) + i%4][l % 32].
#include "internal.h"

using namespace trh;

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int C = 32;          // channels of every hidden layer
constexpr int KH = 9 * C;      // reduction length of a 32-channel layer
constexpr int HEAD_IN = 32, HEAD_HID = 64, HEAD_OUT = 9;
constexpr int HEAD_PARAMS = HEAD_HID * HEAD_IN + HEAD_HID + HEAD_OUT * HEAD_HID + HEAD_OUT;
constexpr int HEAD_SCR = HEAD_OUT + HEAD_HID + HEAD_HID + HEAD_IN;   // per-sample scratch

__device__ __forceinline__ float bf2f(uint16_t v) { return __uint_as_float((uint32_t)v << 16); }

// float -> bfloat16, round to nearest even (NaN kept quiet)
__device__ __forceinline__ uint16_t f2bf(float f) {
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

__device__ __forceinline__ bf16x8 load8(const uint16_t *p) {
    return *reinterpret_cast<const bf16x8 *>(p);
}

__device__ __forceinline__ bf16x8 zero8() {
    bf16x8 z;
    for (int j = 0; j < 8; ++j) z[j] = (__bf16)0.0f;
    return z;
}

__device__ __forceinline__ __bf16 raw_bf(uint16_t v) {
    return __builtin_bit_cast(__bf16, v);
}

__device__ __forceinline__ f32x16 mfma(bf16x8 a, bf16x8 b, f32x16 c) {
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
}

// ---------------------------------------------------------------------------
// forward: y[m][co] = relu(bias[co] + sum_k A[m][k] w[co][k]),
// A[m][k = tap*CIN + ci] = x[b][2oy+ky-1][2ox+kx-1][ci] (0 outside).
// CIN = 1 pads the 9 taps to K = 16 (w is [32][16]).
// ---------------------------------------------------------------------------
template <int CIN>
__global__ void __launch_bounds__(256)
k_conv_fwd(const uint16_t *__restrict__ x, const uint16_t *__restrict__ w,
           const float *__restrict__ bias, uint16_t *__restrict__ y, int B, int H, int W) {
    constexpr int K = (CIN == 1) ? 16 : 9 * CIN;
    const int OH = H / 2, OW = W / 2;
    const long M = (long)B * OH * OW;
    const int lane = threadIdx.x & 63, half = lane >> 5, r = lane & 31;
    const long ntiles = (M + 31) / 32;
    for (long t = (long)blockIdx.x * 4 + (threadIdx.x >> 6); t < ntiles; t += (long)gridDim.x * 4) {
        const long m = t * 32 + r;
        const bool mv = m < M;
        const long mm = mv ? m : 0;
        const int ox = (int)(mm % OW);
        const int oy = (int)((mm / OW) % OH);
        const long b = mm / ((long)OW * OH);
        f32x16 acc;
        for (int i = 0; i < 16; ++i) acc[i] = 0.f;
        for (int ks = 0; ks < K; ks += 16) {
            const int k0 = ks + 8 * half;
            bf16x8 a;
            if (CIN == 1) {
                for (int j = 0; j < 8; ++j) {
                    const int tap = k0 + j;
                    uint16_t v = 0;
                    if (mv && tap < 9) {
                        const int iy = 2 * oy + tap / 3 - 1, ix = 2 * ox + tap % 3 - 1;
                        if (iy >= 0 && iy < H && ix >= 0 && ix < W) v = x[(b * H + iy) * W + ix];
                    }
                    a[j] = raw_bf(v);
                }
            } else {
                const int tap = k0 / CIN, ci0 = k0 % CIN;
                const int iy = 2 * oy + tap / 3 - 1, ix = 2 * ox + tap % 3 - 1;
                if (mv && iy >= 0 && iy < H && ix >= 0 && ix < W)
                    a = load8(x + ((b * H + iy) * W + ix) * CIN + ci0);
                else
                    a = zero8();
            }
            const bf16x8 bv = load8(w + r * K + k0);
            acc = mfma(a, bv, acc);
        }
        const float bb = bias[r];
        for (int i = 0; i < 16; ++i) {
            const long row = t * 32 + 8 * (i / 4) + 4 * half + (i % 4);
            if (row < M) y[row * C + r] = f2bf(fmaxf(acc[i] + bb, 0.f));
        }
    }
}

// ---------------------------------------------------------------------------
// backward-data (32-channel input): dx[m][ci] = mask(x[m][ci] > 0) *
// sum_{tap,co} dy[b][oy][ox][co] w[co][tap][ci], with oy = (iy + 1 - ky) / 2
// when that is an integer in range.  wt is w re-laid as [ci][tap][co].
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256)
k_conv_bwd_data(const uint16_t *__restrict__ dy, const uint16_t *__restrict__ wt,
                const uint16_t *__restrict__ x, uint16_t *__restrict__ dx, int B, int H, int W) {
    const int OH = H / 2, OW = W / 2;
    const long M = (long)B * H * W;
    const int lane = threadIdx.x & 63, half = lane >> 5, r = lane & 31;
    const long ntiles = (M + 31) / 32;
    for (long t = (long)blockIdx.x * 4 + (threadIdx.x >> 6); t < ntiles; t += (long)gridDim.x * 4) {
        const long m = t * 32 + r;
        const bool mv = m < M;
        const long mm = mv ? m : 0;
        const int ix = (int)(mm % W);
        const int iy = (int)((mm / W) % H);
        const long b = mm / ((long)W * H);
        f32x16 acc;
        for (int i = 0; i < 16; ++i) acc[i] = 0.f;
        for (int ks = 0; ks < KH; ks += 16) {
            const int k0 = ks + 8 * half;
            const int tap = k0 / C, co0 = k0 % C;
            const int ty = iy + 1 - tap / 3, tx = ix + 1 - tap % 3;
            bf16x8 a;
            if (mv && ty >= 0 && tx >= 0 && !(ty & 1) && !(tx & 1) && (ty >> 1) < OH &&
                (tx >> 1) < OW)
                a = load8(dy + ((b * OH + (ty >> 1)) * OW + (tx >> 1)) * C + co0);
            else
                a = zero8();
            const bf16x8 bv = load8(wt + r * KH + k0);
            acc = mfma(a, bv, acc);
        }
        for (int i = 0; i < 16; ++i) {
            const long row = t * 32 + 8 * (i / 4) + 4 * half + (i % 4);
            if (row < M) {
                const bool on = bf2f(x[row * C + r]) > 0.f;
                dx[row * C + r] = f2bf(on ? acc[i] : 0.f);
            }
        }
    }
}

// ---------------------------------------------------------------------------
// backward-weight (32-channel input): dw[co][tap*32 + ci] += sum_m dy[m][co] *
// x[m at tap][ci]; db[co] += sum_m dy[m][co].  A block stages 64 output pixels
// of dy and their 9 input pixels in LDS; wave w owns taps w, w+4, w+8; each
// tap is one 32x32 MFMA tile reduced over pixels 16 at a time.
// ---------------------------------------------------------------------------
constexpr int WCH = 64;   // output pixels per LDS chunk

__global__ void __launch_bounds__(256)
k_conv_bwd_weight(const uint16_t *__restrict__ dy, const uint16_t *__restrict__ x,
                  float *__restrict__ dw, float *__restrict__ db, int B, int H, int W) {
    __shared__ uint16_t sdy[WCH * C];
    __shared__ uint16_t sx[9 * WCH * C];
    const int OH = H / 2, OW = W / 2;
    const long M = (long)B * OH * OW;
    const int tid = threadIdx.x, lane = tid & 63, half = lane >> 5, r = lane & 31;
    const int wave = tid >> 6;
    f32x16 acc[3];
    for (int s = 0; s < 3; ++s)
        for (int i = 0; i < 16; ++i) acc[s][i] = 0.f;
    float bsum = 0.f;
    const long nchunks = (M + WCH - 1) / WCH;
    for (long ch = blockIdx.x; ch < nchunks; ch += gridDim.x) {
        // stage: thread -> (pixel p = tid / 4, channels 8 * (tid % 4) .. +7)
        {
            const int p = tid >> 2, c0 = (tid & 3) * 8;
            const long m = ch * WCH + p;
            uint4 v = make_uint4(0, 0, 0, 0);
            if (m < M) v = *reinterpret_cast<const uint4 *>(dy + m * C + c0);
            *reinterpret_cast<uint4 *>(sdy + p * C + c0) = v;
            const long mm = m < M ? m : 0;
            const int ox = (int)(mm % OW), oy = (int)((mm / OW) % OH);
            const long b = mm / ((long)OW * OH);
            for (int tap = 0; tap < 9; ++tap) {
                const int iy = 2 * oy + tap / 3 - 1, ix = 2 * ox + tap % 3 - 1;
                uint4 u = make_uint4(0, 0, 0, 0);
                if (m < M && iy >= 0 && iy < H && ix >= 0 && ix < W)
                    u = *reinterpret_cast<const uint4 *>(x + ((b * H + iy) * W + ix) * C + c0);
                *reinterpret_cast<uint4 *>(sx + (tap * WCH + p) * C + c0) = u;
            }
        }
        __syncthreads();
        if (tid < C)
            for (int p = 0; p < WCH; ++p) bsum += bf2f(sdy[p * C + tid]);
        for (int s = 0; s < 3; ++s) {
            const int tap = wave + 4 * s;
            if (tap >= 9) break;
            for (int kk = 0; kk < WCH; kk += 16) {
                bf16x8 a, bv;
                for (int j = 0; j < 8; ++j) {
                    const int p = kk + 8 * half + j;
                    a[j] = raw_bf(sdy[p * C + r]);
                    bv[j] = raw_bf(sx[(tap * WCH + p) * C + r]);
                }
                acc[s] = mfma(a, bv, acc[s]);
            }
        }
        __syncthreads();
    }
    for (int s = 0; s < 3; ++s) {
        const int tap = wave + 4 * s;
        if (tap >= 9) break;
        for (int i = 0; i < 16; ++i) {
            const int co = 8 * (i / 4) + 4 * half + (i % 4);
            atomicAdd(dw + co * KH + tap * C + r, acc[s][i]);
        }
    }
    if (tid < C) atomicAdd(db + tid, bsum);
}

// backward-weight of the first layer (one input channel): dw[co][tap] and
// db[co] as FP32 dot products over 64-pixel LDS chunks.
__global__ void __launch_bounds__(256)
k_conv1_bwd_weight(const uint16_t *__restrict__ dy, const uint16_t *__restrict__ x,
                   float *__restrict__ dw, float *__restrict__ db, int B, int H, int W) {
    __shared__ float sdy[WCH * C];
    __shared__ float sx[9 * WCH];
    const int OH = H / 2, OW = W / 2;
    const long M = (long)B * OH * OW;
    const int tid = threadIdx.x;
    float a0 = 0.f, a1 = 0.f, bs = 0.f;   // pairs tid and tid + 256 of (co, tap)
    const long nchunks = (M + WCH - 1) / WCH;
    for (long ch = blockIdx.x; ch < nchunks; ch += gridDim.x) {
        for (int i = tid; i < WCH * C; i += 256) {
            const long m = ch * WCH + i / C;
            sdy[i] = m < M ? bf2f(dy[m * C + (i % C)]) : 0.f;
        }
        for (int i = tid; i < 9 * WCH; i += 256) {
            const int tap = i / WCH, p = i % WCH;
            const long m = ch * WCH + p;
            float v = 0.f;
            if (m < M) {
                const int ox = (int)(m % OW), oy = (int)((m / OW) % OH);
                const long b = m / ((long)OW * OH);
                const int iy = 2 * oy + tap / 3 - 1, ix = 2 * ox + tap % 3 - 1;
                if (iy >= 0 && iy < H && ix >= 0 && ix < W) v = bf2f(x[(b * H + iy) * W + ix]);
            }
            sx[i] = v;
        }
        __syncthreads();
        {
            const int co = tid / 9, tap = tid % 9;
            for (int p = 0; p < WCH; ++p) a0 += sdy[p * C + co] * sx[tap * WCH + p];
        }
        if (tid + 256 < 9 * C) {
            const int q = tid + 256, co = q / 9, tap = q % 9;
            for (int p = 0; p < WCH; ++p) a1 += sdy[p * C + co] * sx[tap * WCH + p];
        }
        if (tid < C)
            for (int p = 0; p < WCH; ++p) bs += sdy[p * C + tid];
        __syncthreads();
    }
    atomicAdd(dw + tid, a0);                       // dw layout [co][9], index co*9 + tap
    if (tid + 256 < 9 * C) atomicAdd(dw + tid + 256, a1);
    if (tid < C) atomicAdd(db + tid, bs);
}

// ---------------------------------------------------------------------------
// head: global average pool of y4 [B][2][2][32] (bf16), MLP 32 -> 64 (ReLU)
// -> 9.  One thread per sample.  With targets: loss = mean over the batch of
// 0.5 sum_i rw_i (o_i - t_i)^2 (parameters 3..5 only for galaxies, label > 0)
// + tw * cross-entropy(o[6..8], label); writes dy4 (bf16, ReLU-masked) and the
// per-sample scratch the weight-gradient kernel reduces.
// params: W1 [64][32], b1 [64], W2 [9][64], b2 [9].
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256)
k_head(const uint16_t *__restrict__ y4, const float *__restrict__ prm, int Bn,
       const float *__restrict__ targets, const int32_t *__restrict__ labels,
       const float *__restrict__ rw, float tw, float *__restrict__ out, float *__restrict__ loss,
       uint16_t *__restrict__ dy4, float *__restrict__ scr) {
    const float *W1 = prm, *b1 = W1 + HEAD_HID * HEAD_IN, *W2 = b1 + HEAD_HID,
                *b2 = W2 + HEAD_OUT * HEAD_HID;
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    float lsum = 0.f;
    if (b < Bn) {
        float f[HEAD_IN], z[HEAD_HID], o[HEAD_OUT];
        for (int c = 0; c < HEAD_IN; ++c) {
            float s = 0.f;
            for (int p = 0; p < 4; ++p) s += bf2f(y4[((long)b * 4 + p) * C + c]);
            f[c] = 0.25f * s;
        }
        for (int j = 0; j < HEAD_HID; ++j) {
            float s = b1[j];
            for (int c = 0; c < HEAD_IN; ++c) s += W1[j * HEAD_IN + c] * f[c];
            z[j] = s;
        }
        for (int i = 0; i < HEAD_OUT; ++i) {
            float s = b2[i];
            for (int j = 0; j < HEAD_HID; ++j) s += W2[i * HEAD_HID + j] * fmaxf(z[j], 0.f);
            o[i] = s;
            out[(long)b * HEAD_OUT + i] = s;
        }
        if (targets) {
            const float inv = 1.f / (float)Bn;
            const int lab = min(max(labels[b], 0), 2);
            float d[HEAD_OUT];
            for (int i = 0; i < 6; ++i) {
                const float wgt = (i < 3 || lab > 0) ? rw[i] : 0.f;
                const float e = o[i] - targets[(long)b * 6 + i];
                lsum += 0.5f * wgt * e * e;
                d[i] = wgt * e * inv;
            }
            float mx = fmaxf(o[6], fmaxf(o[7], o[8]));
            float se = 0.f, pr[3];
            for (int c = 0; c < 3; ++c) {
                pr[c] = expf(o[6 + c] - mx);
                se += pr[c];
            }
            for (int c = 0; c < 3; ++c) {
                pr[c] /= se;
                d[6 + c] = tw * (pr[c] - (c == lab ? 1.f : 0.f)) * inv;
            }
            lsum += -tw * logf(fmaxf(pr[lab], 1e-30f));
            float *sc = scr + (long)b * HEAD_SCR;
            for (int i = 0; i < HEAD_OUT; ++i) sc[i] = d[i];
            for (int j = 0; j < HEAD_HID; ++j) sc[HEAD_OUT + j] = fmaxf(z[j], 0.f);
            float dz[HEAD_HID];
            for (int j = 0; j < HEAD_HID; ++j) {
                float s = 0.f;
                for (int i = 0; i < HEAD_OUT; ++i) s += W2[i * HEAD_HID + j] * d[i];
                dz[j] = z[j] > 0.f ? s : 0.f;
                sc[HEAD_OUT + HEAD_HID + j] = dz[j];
            }
            for (int c = 0; c < HEAD_IN; ++c) sc[HEAD_OUT + 2 * HEAD_HID + c] = f[c];
            for (int c = 0; c < HEAD_IN; ++c) {
                float s = 0.f;
                for (int j = 0; j < HEAD_HID; ++j) s += W1[j * HEAD_IN + c] * dz[j];
                for (int p = 0; p < 4; ++p) {
                    const long idx = ((long)b * 4 + p) * C + c;
                    dy4[idx] = f2bf(bf2f(y4[idx]) > 0.f ? 0.25f * s : 0.f);
                }
            }
        }
    }
    if (targets && loss) {
        // wave sum, one atomic per wave
        for (int off = 32; off > 0; off >>= 1) lsum += __shfl_down(lsum, off, 64);
        if ((threadIdx.x & 63) == 0) atomicAdd(loss, lsum / (float)Bn);
    }
}

// head weight gradients: one thread per parameter, reduced over the batch.
__global__ void __launch_bounds__(256)
k_head_wgrad(const float *__restrict__ scr, int Bn, float *__restrict__ g) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= HEAD_PARAMS) return;
    int a, bo;   // scratch offsets of the two factors (bo < 0: factor 1)
    if (q < HEAD_HID * HEAD_IN) {
        a = HEAD_OUT + HEAD_HID + q / HEAD_IN;            // dz1[j]
        bo = HEAD_OUT + 2 * HEAD_HID + q % HEAD_IN;       // f[c]
    } else if (q < HEAD_HID * HEAD_IN + HEAD_HID) {
        a = HEAD_OUT + HEAD_HID + (q - HEAD_HID * HEAD_IN);
        bo = -1;
    } else if (q < HEAD_HID * HEAD_IN + HEAD_HID + HEAD_OUT * HEAD_HID) {
        const int t = q - HEAD_HID * HEAD_IN - HEAD_HID;
        a = t / HEAD_HID;                                  // d[i]
        bo = HEAD_OUT + t % HEAD_HID;                      // a1[j]
    } else {
        a = q - HEAD_HID * HEAD_IN - HEAD_HID - HEAD_OUT * HEAD_HID;
        bo = -1;
    }
    float s = 0.f;
    for (int b = 0; b < Bn; ++b) {
        const float *sc = scr + (long)b * HEAD_SCR;
        s += bo >= 0 ? sc[a] * sc[bo] : sc[a];
    }
    g[q] += s;
}

// Adam on FP32 master parameters (bias-corrected step size passed in).
__global__ void k_adam(float *__restrict__ p, const float *__restrict__ g, float *__restrict__ m,
                       float *__restrict__ v, int n, float lr_t, float b1, float b2, float eps,
                       float wd) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const float gi = g[i] + wd * p[i];
        const float mi = b1 * m[i] + (1.f - b1) * gi;
        const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        p[i] -= lr_t * mi / (sqrtf(vi) + eps);
    }
}

// bf16 copies of the conv weights: wf [co][K] (K = 16 for cin 1, taps 9..15
// zero) and, for cin 32, wt [ci][tap][co].
__global__ void k_pack(const float *__restrict__ w, int cin, uint16_t *__restrict__ wf,
                       uint16_t *__restrict__ wt) {
    const int K = cin == 1 ? 16 : 9 * cin;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < C * K; i += gridDim.x * blockDim.x) {
        const int co = i / K, k = i % K;
        if (cin == 1) {
            wf[i] = k < 9 ? f2bf(w[co * 9 + k]) : 0;
        } else {
            const uint16_t v = f2bf(w[i]);
            wf[i] = v;
            const int tap = k / cin, ci = k % cin;
            wt[(ci * 9 + tap) * C + co] = v;
        }
    }
}

// 32x32 cutouts of patches, centred at patch pixel (P/2, P/2) (cutout pixel
// (16, 16)), zero outside the patch, divided by the cutout maximum; bf16.
__global__ void __launch_bounds__(256)
k_cutouts(const float *__restrict__ data, const int64_t *__restrict__ off,
          const int32_t *__restrict__ P, int n, uint16_t *__restrict__ out,
          float *__restrict__ maxv) {
    __shared__ float red[256];
    for (int s = blockIdx.x; s < n; s += gridDim.x) {
        const int Ps = P[s];
        const int x0 = Ps / 2 - 16, y0 = Ps / 2 - 16;
        const float *d = data + off[s];
        float v[4], mx = -3.4e38f;
        for (int q = 0; q < 4; ++q) {
            const int i = threadIdx.x + 256 * q, yy = y0 + i / 32, xx = x0 + i % 32;
            v[q] = (yy >= 0 && yy < Ps && xx >= 0 && xx < Ps) ? d[(long)yy * Ps + xx] : 0.f;
            mx = fmaxf(mx, v[q]);
        }
        red[threadIdx.x] = mx;
        __syncthreads();
        for (int k = 128; k > 0; k >>= 1) {
            if (threadIdx.x < k) red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + k]);
            __syncthreads();
        }
        float m = red[0];
        if (!(m > 0.f)) m = 1.f;
        for (int q = 0; q < 4; ++q) out[(long)s * 1024 + threadIdx.x + 256 * q] = f2bf(v[q] / m);
        if (threadIdx.x == 0) maxv[s] = m;
        __syncthreads();
    }
}

int grid_for(trh_handle *h, long units, int per_block) {
    long g = (units + per_block - 1) / per_block;
    const long cap = 8L * h->n_cu;
    if (g > cap) g = cap;
    if (g < 1) g = 1;
    return (int)g;
}

bool even_hw(int H, int W) { return H >= 2 && W >= 2 && !(H & 1) && !(W & 1); }

}  // namespace

extern "C" int trh_cnn_conv_fwd(trh_handle *h, const uint16_t *x, const uint16_t *wf,
                                const float *bias, uint16_t *y, int B, int H, int W, int cin,
                                void *stream) {
    if (!h || !x || !wf || !bias || !y || B <= 0 || !even_hw(H, W) || (cin != 1 && cin != 32))
        return set_err(TRH_E_ARG, "trh_cnn_conv_fwd: bad arguments");
    const long tiles = ((long)B * (H / 2) * (W / 2) + 31) / 32;
    const int g = grid_for(h, tiles, 4);
    if (cin == 1)
        hipLaunchKernelGGL(k_conv_fwd<1>, dim3(g), dim3(256), 0, (hipStream_t)stream, x, wf, bias,
                           y, B, H, W);
    else
        hipLaunchKernelGGL(k_conv_fwd<32>, dim3(g), dim3(256), 0, (hipStream_t)stream, x, wf, bias,
                           y, B, H, W);
    return hip_check(hipGetLastError(), "k_conv_fwd");
}

extern "C" int trh_cnn_conv_bwd_data(trh_handle *h, const uint16_t *dy, const uint16_t *wt,
                                     const uint16_t *x, uint16_t *dx, int B, int H, int W,
                                     void *stream) {
    if (!h || !dy || !wt || !x || !dx || B <= 0 || !even_hw(H, W))
        return set_err(TRH_E_ARG, "trh_cnn_conv_bwd_data: bad arguments");
    const long tiles = ((long)B * H * W + 31) / 32;
    const int g = grid_for(h, tiles, 4);
    hipLaunchKernelGGL(k_conv_bwd_data, dim3(g), dim3(256), 0, (hipStream_t)stream, dy, wt, x, dx,
                       B, H, W);
    return hip_check(hipGetLastError(), "k_conv_bwd_data");
}

extern "C" int trh_cnn_conv_bwd_weight(trh_handle *h, const uint16_t *dy, const uint16_t *x,
                                       float *dw, float *db, int B, int H, int W, int cin,
                                       void *stream) {
    if (!h || !dy || !x || !dw || !db || B <= 0 || !even_hw(H, W) || (cin != 1 && cin != 32))
        return set_err(TRH_E_ARG, "trh_cnn_conv_bwd_weight: bad arguments");
    const long chunks = ((long)B * (H / 2) * (W / 2) + WCH - 1) / WCH;
    const int g = grid_for(h, chunks, 1) < h->n_cu ? grid_for(h, chunks, 1) : h->n_cu;
    if (cin == 1)
        hipLaunchKernelGGL(k_conv1_bwd_weight, dim3(g), dim3(256), 0, (hipStream_t)stream, dy, x,
                           dw, db, B, H, W);
    else
        hipLaunchKernelGGL(k_conv_bwd_weight, dim3(g), dim3(256), 0, (hipStream_t)stream, dy, x, dw,
                           db, B, H, W);
    return hip_check(hipGetLastError(), "k_conv_bwd_weight");
}

extern "C" int trh_cnn_head(trh_handle *h, const uint16_t *y4, const float *params, int B,
                            const float *targets, const int32_t *labels, const float *rw,
                            float tw, float *out, float *loss, uint16_t *dy4, float *scratch,
                            float *dparams, void *stream) {
    if (!h || !y4 || !params || !out || B <= 0)
        return set_err(TRH_E_ARG, "trh_cnn_head: bad arguments");
    if (targets && (!labels || !rw || !dy4 || !scratch || !dparams))
        return set_err(TRH_E_ARG, "trh_cnn_head: training needs labels, rw, dy4, scratch, dparams");
    hipStream_t st = (hipStream_t)stream;
    hipLaunchKernelGGL(k_head, dim3((B + 255) / 256), dim3(256), 0, st, y4, params, B, targets,
                       labels, rw, tw, out, loss, dy4, scratch);
    int rc = hip_check(hipGetLastError(), "k_head");
    if (rc || !targets) return rc;
    hipLaunchKernelGGL(k_head_wgrad, dim3((HEAD_PARAMS + 255) / 256), dim3(256), 0, st, scratch,
                       B, dparams);
    return hip_check(hipGetLastError(), "k_head_wgrad");
}

extern "C" int trh_cnn_adam(trh_handle *h, float *p, const float *g, float *m, float *v, int n,
                            float lr_t, float b1, float b2, float eps, float wd, void *stream) {
    if (!h || !p || !g || !m || !v || n <= 0) return set_err(TRH_E_ARG, "trh_cnn_adam: bad arguments");
    hipLaunchKernelGGL(k_adam, dim3(grid_for(h, n, 256)), dim3(256), 0, (hipStream_t)stream, p, g,
                       m, v, n, lr_t, b1, b2, eps, wd);
    return hip_check(hipGetLastError(), "k_adam");
}

extern "C" int trh_cnn_pack(trh_handle *h, const float *w, int cin, uint16_t *wf, uint16_t *wt,
                            void *stream) {
    if (!h || !w || !wf || (cin != 1 && cin != 32) || (cin == 32 && !wt))
        return set_err(TRH_E_ARG, "trh_cnn_pack: bad arguments");
    const int K = cin == 1 ? 16 : 9 * cin;
    hipLaunchKernelGGL(k_pack, dim3((C * K + 255) / 256), dim3(256), 0, (hipStream_t)stream, w, cin,
                       wf, wt);
    return hip_check(hipGetLastError(), "k_pack");
}

extern "C" int trh_cnn_cutouts(trh_handle *h, const float *data, const int64_t *pix_off,
                               const int32_t *P, int n, uint16_t *out, float *maxv, void *stream) {
    if (!h || !data || !pix_off || !P || !out || !maxv || n <= 0)
        return set_err(TRH_E_ARG, "trh_cnn_cutouts: bad arguments");
    hipLaunchKernelGGL(k_cutouts, dim3(grid_for(h, n, 1)), dim3(256), 0, (hipStream_t)stream, data,
                       pix_off, P, n, out, maxv);
    return hip_check(hipGetLastError(), "k_cutouts");
}
