// Batched forced photometry on gfx950: the linear flux solve of
// Optimizer.forced_photometry (tractor/optimize.py:16-145) with
// LsqrOptimizer's unscaled least-squares step (tractor/lsqr_optimizer.py:
// 11-251) for thousands of overlapping sources on one image.
//
// The least-squares fluxes solve the normal equations A f = b with
//   A_st = sum_pix u_s u_t inverr^2,   b_s = sum_pix u_s (data - mod0) inverr^2
// over the unit-flux patches u_s.  Only overlapping patches couple, so A is
// stored as CSR (one entry per overlapping pair, FP32 products, FP64 sums).
// It is solved in FP64 by preconditioned conjugate gradients whose
// preconditioner is the exact Cholesky factor of A's diagonal blocks, one
// block per spatial tile of sources.
#include "internal.h"

using namespace trh;

namespace {

__device__ __forceinline__ double wave_sum_d(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// out[s][e] = sum_{q < nq} w[s][q] * basis[idx[s][q]][e] (nq <= 16): a
// spatially varying PSF as a weighted sum of basis images -- bilinear
// interpolation of a PSF grid (nq = 4) or PSFEx's polynomial over its
// eigen-PSFs (psfex.py:326-346).  The same linear map applies to stamps and
// to their spectra (PixelizedPsfEx.getFourierTransform, psfex.py:481-520).
constexpr int PSF_MAXQ = 16;
// One 256-thread block per source; VT = float4 where the length and the
// pointers allow it (the spectra), float otherwise (odd-sized stamps).
__device__ __forceinline__ float4 operator*(float a, float4 v) {
    return make_float4(a * v.x, a * v.y, a * v.z, a * v.w);
}
__device__ __forceinline__ float4 &operator+=(float4 &a, float4 v) {
    a.x += v.x;
    a.y += v.y;
    a.z += v.z;
    a.w += v.w;
    return a;
}
__device__ __forceinline__ void vzero(float &a) { a = 0.f; }
__device__ __forceinline__ void vzero(float4 &a) { a = make_float4(0.f, 0.f, 0.f, 0.f); }
template <typename VT>
__global__ void __launch_bounds__(256)
k_psf_combine(const float *__restrict__ basis, int nbasis, int64_t len,
              const float *__restrict__ w, const int32_t *__restrict__ idx, int nq,
              float *__restrict__ out, int n) {
    const int s = blockIdx.x;
    if (s >= n) return;
    constexpr int V = sizeof(VT) / sizeof(float);
    float ws[PSF_MAXQ];
    const VT *b[PSF_MAXQ];
#pragma unroll
    for (int q = 0; q < PSF_MAXQ; ++q) {
        int k = q < nq ? idx[nq * s + q] : 0;
        k = k < 0 ? 0 : (k >= nbasis ? nbasis - 1 : k);
        ws[q] = q < nq ? w[nq * s + q] : 0.f;
        b[q] = reinterpret_cast<const VT *>(basis + (int64_t)k * len);
    }
    VT *o = reinterpret_cast<VT *>(out + (int64_t)s * len);
    const int64_t nv = len / V;
    for (int64_t e = threadIdx.x; e < nv; e += blockDim.x) {
        VT acc;
        vzero(acc);
#pragma unroll
        for (int q = 0; q < PSF_MAXQ; ++q)
            if (q < nq) acc += ws[q] * b[q][e];
        o[e] = acc;
    }
}

// PSFEx polynomial terms of source order[q] at its position (src rows x, y,
// ...; stride 6): dx = (x - x0) / xscale, dy = (y - y0) / yscale, term
// dx^j dy^k at index j + (degree + 1) k - k (k - 1) / 2 (PsfEx.polynomials,
// psfex.py:294-325), written with the identity basis index for each term.
__global__ void k_psfex_terms(int n, const int64_t *__restrict__ order,
                              const double *__restrict__ src, int degree, double x0, double y0,
                              double xscale, double yscale, float *__restrict__ w,
                              int32_t *__restrict__ idx) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const int64_t s = order ? order[q] : q;
    const double dx = (src[6 * s] - x0) / xscale, dy = (src[6 * s + 1] - y0) / yscale;
    const int nq = (degree + 1) * (degree + 2) / 2;
    double yk = 1.0;
    for (int k = 0; k <= degree; ++k) {
        double xj = 1.0;
        for (int j = 0; j + k <= degree; ++j) {
            const int ii = j + (degree + 1) * k - (k * (k - 1)) / 2;
            w[(int64_t)nq * q + ii] = (float)(xj * yk);
            idx[(int64_t)nq * q + ii] = ii;
            xj *= dx;
        }
        yk *= dy;
    }
}

// One wavefront per overlapping pair (s, t): A_st over the intersection of
// their windows; written at CSR positions pos[2p] and pos[2p+1].
__global__ void __launch_bounds__(256)
k_fp_pair_dots(const int32_t *__restrict__ pairs, const int64_t *__restrict__ pos, int npairs,
               const int32_t *__restrict__ win, const int64_t *__restrict__ poff,
               const float *__restrict__ patches, const float *__restrict__ inverr, int img_w,
               double *__restrict__ vals) {
    const int lane = threadIdx.x & 63;
    const int nwave = gridDim.x * (blockDim.x >> 6);
    for (int p = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); p < npairs; p += nwave) {
        const int s = pairs[2 * p], t = pairs[2 * p + 1];
        const int sx0 = win[4 * s], sy0 = win[4 * s + 1], sw = win[4 * s + 2], sh = win[4 * s + 3];
        const int tx0 = win[4 * t], ty0 = win[4 * t + 1], tw = win[4 * t + 2], th = win[4 * t + 3];
        const int X0 = max(sx0, tx0), X1 = min(sx0 + sw, tx0 + tw);
        const int Y0 = max(sy0, ty0), Y1 = min(sy0 + sh, ty0 + th);
        const float *us = patches + poff[s];
        const float *ut = patches + poff[t];
        float acc = 0.f;
        const int nx = X1 - X0;
        for (int y = Y0; y < Y1; ++y) {
            const float *rs = us + (int64_t)(y - sy0) * sw + (X0 - sx0);
            const float *rt = ut + (int64_t)(y - ty0) * tw + (X0 - tx0);
            const float *re = inverr + (int64_t)y * img_w + X0;
            for (int x = lane; x < nx; x += 64) {
                const float e = re[x];
                acc += rs[x] * rt[x] * (e * e);
            }
        }
        const double v = wave_sum_d((double)acc);
        if (lane == 0) {
            vals[pos[2 * p]] = v;
            vals[pos[2 * p + 1]] = v;
        }
    }
}

// (y, x) of element i of a row-major nx-wide block, stepped by `stride`
// elements at a time without a division per step.
struct FlatWalk {
    int y, x, dy, dx, nx;
    __device__ FlatWalk(int i, int nx_, int stride) : nx(nx_) {
        y = i / nx;
        x = i - y * nx;
        dy = stride / nx;
        dx = stride - dy * nx;
    }
    __device__ void step() {
        x += dx;
        y += dy;
        if (x >= nx) {
            x -= nx;
            ++y;
        }
    }
};

// One 256-thread workgroup per source, the window flattened over the
// threads, four elements per thread in flight:
// b_s = sum u_s (data - mod0) inverr^2 (mod0 may be NULL: zero; FP32
// products, FP64 sum) and, when wpatch is given, the weighted patch
// w_s = u_s inverr^2 at the patch's offset.
__global__ void __launch_bounds__(256)
k_fp_rhs(int n, const int32_t *__restrict__ win, const int64_t *__restrict__ poff,
         const float *__restrict__ patches, const float *__restrict__ data,
         const float *__restrict__ mod0, const float *__restrict__ inverr, int img_w,
         double *__restrict__ b, float *__restrict__ wpatch) {
    __shared__ double red[4];
    const int s = blockIdx.x;
    const int x0 = win[4 * s], y0 = win[4 * s + 1], w = win[4 * s + 2], h = win[4 * s + 3];
    double acc = 0.0;
    if (w > 0 && h > 0) {
        const float *u = patches + poff[s];
        float *wo = wpatch ? wpatch + poff[s] : nullptr;
        const int m = w * h;
        FlatWalk fw(threadIdx.x, w, 256);
        for (int i0 = threadIdx.x; i0 < m; i0 += 1024) {
            float uu[4], ee[4], dd[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const int i = i0 + 256 * k;
                const bool ok = i < m;
                const int64_t ib = (int64_t)(y0 + fw.y) * img_w + x0 + fw.x;
                uu[k] = ok ? u[i] : 0.f;
                ee[k] = ok ? inverr[ib] : 0.f;
                dd[k] = ok ? data[ib] - (mod0 ? mod0[ib] : 0.f) : 0.f;
                fw.step();
            }
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const float wu = uu[k] * (ee[k] * ee[k]);
                if (wo && i0 + 256 * k < m) wo[i0 + 256 * k] = wu;
                acc += (double)(wu * dd[k]);
            }
        }
    }
    acc = wave_sum_d(acc);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) b[s] = red[0] + red[1] + red[2] + red[3];
}

// One workgroup per tile of rows [t0, t1) (n = t1 - t0 rows, the block in
// LDS): copy A's diagonal block from the CSR and invert it in place by
// Gauss-Jordan elimination without pivoting (stable for a symmetric positive
// definite block: every pivot is a Schur complement's diagonal), every thread
// updating its share of the block at each pivot.  The inverse is symmetrised
// and written dense at blocks + block_off[T].  A zero pivot (a source with no
// pixels) becomes 1 so its flux solves to 0; a negative one sets info[T] =
// its index + 1 (and becomes 1); a tile of more than max_n rows sets -1.
__global__ void __launch_bounds__(256)
k_fp_block_inv(const int64_t *__restrict__ row_ptr, const int32_t *__restrict__ col,
               const double *__restrict__ vals, const int32_t *__restrict__ tile_start,
               const int64_t *__restrict__ block_off, int max_n, double *__restrict__ blocks,
               int32_t *__restrict__ info) {
    // LDS: B (n * n), then column k (n), then the tile's row_ptr (n + 1, as
    // offsets from its first entry)
    extern __shared__ double B[];
    const int T = blockIdx.x;
    const int t0 = tile_start[T], n = tile_start[T + 1] - t0;
    if (n > max_n) {   // the block does not fit the LDS reserved for max_n
        if (threadIdx.x == 0) info[T] = -1;
        return;
    }
    double *ck = B + n * n;
    int *rp = (int *)(ck + n);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nw = blockDim.x >> 6;
    const int64_t e0 = row_ptr[t0];
    for (int i = threadIdx.x; i <= n; i += blockDim.x) rp[i] = (int)(row_ptr[t0 + i] - e0);
    for (int i = threadIdx.x; i < n * n; i += blockDim.x) B[i] = 0.0;
    __syncthreads();
    // the tile's entries, 256 at a time; an entry's row by bisection of rp
    const int ne = rp[n];
    for (int e = threadIdx.x; e < ne; e += blockDim.x) {
        const int c = col[e0 + e] - t0;
        const double v = vals[e0 + e];
        int lo = 0, hi = n;   // rp[lo] <= e < rp[hi]
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (rp[mid] <= e) lo = mid;
            else hi = mid;
        }
        if (c >= 0 && c < n) B[lo * n + c] = v;
    }
    __syncthreads();
    // rows over waves, columns over lanes
    for (int k = 0; k < n; ++k) {
        double piv = B[k * n + k];
        if (piv == 0.0) piv = 1.0;
        if (!(piv > 0.0)) {
            if (threadIdx.x == 0) info[T] = k + 1;
            piv = 1.0;
        }
        const double ip = 1.0 / piv;
        for (int i = threadIdx.x; i < n; i += blockDim.x) ck[i] = i == k ? 0.0 : B[i * n + k];
        __syncthreads();
        // row k scaled by 1 / pivot (its pivot entry replaced by 1 first)
        for (int j = threadIdx.x; j < n; j += blockDim.x)
            B[k * n + j] = (j == k ? 1.0 : B[k * n + j]) * ip;
        __syncthreads();
        // a_ij -= a_ik a_kj for i != k, with a_ik = 0 first (so column k
        // becomes -a_ik / pivot)
        for (int j = lane; j < n; j += 64) {
            const double bkj = B[k * n + j];
            for (int i = wv; i < n; i += nw) {
                if (i == k) continue;
                B[i * n + j] = (j == k ? 0.0 : B[i * n + j]) - ck[i] * bkj;
            }
        }
        __syncthreads();
    }
    double *M = blocks + block_off[T];
    for (int i = wv; i < n; i += nw)
        for (int j = lane; j < n; j += 64) M[i * n + j] = 0.5 * (B[i * n + j] + B[j * n + i]);
}

// z = M r on one tile of n rows (M the block inverse, symmetric: z_i =
// sum_k M[k][i] r_k, so each wave reads rows of M with lanes along i).  Four
// waves split k; their partials meet in LDS in a fixed order.  rr (the tile's
// r) and part (4 * 64 doubles) are LDS; blockDim.x must be 256.  Returns this
// thread's share of sum_i r_i z_i.
__device__ double tile_precond(const double *__restrict__ M, int n, const double *rr, double *part,
                               double *__restrict__ z) {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    double rz = 0.0;
    for (int i0 = 0; i0 < n; i0 += 64) {
        const int i = i0 + lane;
        double acc = 0.0;
        if (i < n) {
#pragma unroll 8
            for (int k = wv; k < n; k += 4) acc += M[k * n + i] * rr[k];
        }
        part[wv * 64 + lane] = acc;
        __syncthreads();
        if (wv == 0 && i < n) {
            const double zi = part[lane] + part[64 + lane] + part[128 + lane] + part[192 + lane];
            z[i] = zi;
            rz += rr[i] * zi;
        }
        __syncthreads();
    }
    return rz;
}

__global__ void __launch_bounds__(256)
k_fp_precond(const int32_t *__restrict__ tile_start, const int64_t *__restrict__ block_off,
             int max_n, const double *__restrict__ blocks, const double *__restrict__ r,
             double *__restrict__ z) {
    extern __shared__ double part[];
    const int T = blockIdx.x;
    const int t0 = tile_start[T], n = min(tile_start[T + 1] - t0, max_n);
    double *rr = part + 256;
    for (int i = threadIdx.x; i < n; i += blockDim.x) rr[i] = r[t0 + i];
    __syncthreads();
    tile_precond(blocks + block_off[T], n, rr, part, z + t0);
}

// ---------------------------------------------------------------------------
// Preconditioned conjugate gradients with every scalar on the device, two
// launches per iteration:
//   step k: alpha = r.z / p.q; x += alpha p; r -= alpha q; z = M r; new r.z, r.r
//   dir k+1: stop if r.r <= tol^2 b.b; beta = r.z(new) / r.z(old);
//            p = z + beta p; q = A z + beta q (= A p); p.q
// Every launch has one 256-thread block per tile.  The dot products are
// per-tile partials (one double per tile) that every block of the next launch
// reduces in the same fixed order, so all blocks agree on alpha, beta and the
// stopping test bit for bit, with no host round trip and no atomics.
// status[0]: 0 running, 1 converged, 2 breakdown (p.Ap <= 0); status[1]:
// steps taken.  Once status[0] is set the remaining launches return at once.
constexpr int PCG_TB = 256;

// v[j] <- the block's sum of v[j] (the same value in every thread; fixed
// order, so every block reducing the same data gets the same bits).
template <int NV>
__device__ void block_reduce(double (&v)[NV]) {
    __shared__ double red[NV][PCG_TB / 64];
#pragma unroll
    for (int j = 0; j < NV; ++j) {
        const double w = wave_sum_d(v[j]);
        if ((threadIdx.x & 63) == 0) red[j][threadIdx.x >> 6] = w;
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < NV; ++j) {
        double s = 0.0;
        for (int w = 0; w < PCG_TB / 64; ++w) s += red[j][w];
        v[j] = s;
    }
    __syncthreads();
}

// out[j] = sum_i a[j][i] over m per-tile partials, and whether status[0] is
// set.  status[0] is read once per block (its load overlapping the partials'):
// another block may set it while this one runs, and every wave of a block
// must take the same branch.
template <int NV>
__device__ bool partial_sums(const double *const (&a)[NV], int m, const int32_t *status,
                             double (&out)[NV]) {
    __shared__ int st;
    if (threadIdx.x == 0) st = status[0];
#pragma unroll
    for (int j = 0; j < NV; ++j) out[j] = 0.0;
    for (int i = threadIdx.x; i < m; i += PCG_TB)
#pragma unroll
        for (int j = 0; j < NV; ++j) out[j] += a[j][i];
    block_reduce(out);   // its barriers publish st
    return st != 0;
}

// x = 0, r = b, z = M b; per-tile r.z, r.r (= b.b) and b.b.
__global__ void __launch_bounds__(PCG_TB)
k_pcg_init(const int32_t *__restrict__ tile_start, const int64_t *__restrict__ block_off,
           int max_n, const double *__restrict__ blocks, const double *__restrict__ b, double *__restrict__ x,
           double *__restrict__ r, double *__restrict__ z, double *__restrict__ rzp,
           double *__restrict__ rrp, double *__restrict__ bbp, int32_t *__restrict__ status) {
    extern __shared__ double part[];
    const int T = blockIdx.x;
    const int t0 = tile_start[T], n = min(tile_start[T + 1] - t0, max_n);
    double *rr = part + PCG_TB;
    double v[2] = {0.0, 0.0};
    for (int i = threadIdx.x; i < n; i += PCG_TB) {
        const double bi = b[t0 + i];
        rr[i] = bi;
        r[t0 + i] = bi;
        x[t0 + i] = 0.0;
        v[1] += bi * bi;
    }
    if (T == 0 && threadIdx.x == 0) {
        status[0] = 0;
        status[1] = 0;
    }
    __syncthreads();
    v[0] = tile_precond(blocks + block_off[T], n, rr, part, z + t0);
    block_reduce(v);
    if (threadIdx.x == 0) {
        rzp[T] = v[0];
        rrp[T] = v[1];
        bbp[T] = v[1];
    }
}

// The stopping test, then p = z + beta p and q = A z + beta q over the tile's
// rows (8 lanes per row) and the tile's p.q.  first: beta = 0 (p = z).
__global__ void __launch_bounds__(PCG_TB)
k_pcg_dir(const int32_t *__restrict__ tile_start, const int64_t *__restrict__ row_ptr,
          const int32_t *__restrict__ col, const double *__restrict__ vals, int ntile,
          const double *__restrict__ z, double *__restrict__ p, double *__restrict__ q,
          const double *__restrict__ rz_new, const double *__restrict__ rz_old,
          const double *__restrict__ rrp, const double *__restrict__ bbp,
          double *__restrict__ pqp, double tol2, int32_t *__restrict__ status,
          double *__restrict__ norms, int k, int first) {
    const double *const parts[4] = {rz_new, rz_old, rrp, bbp};
    double s[4];
    if (partial_sums(parts, ntile, status, s)) return;
    const bool done = s[2] <= tol2 * s[3];
    if (threadIdx.x == 0 && (done || blockIdx.x == 0)) {
        if (done) status[0] = 1;
        status[1] = k;
        norms[0] = s[2];
        norms[1] = s[3];
    }
    if (done) return;
    const double beta = first ? 0.0 : s[0] / s[1];
    const int T = blockIdx.x;
    const int t0 = tile_start[T], t1 = tile_start[T + 1];
    const int g = threadIdx.x >> 3, gl = threadIdx.x & 7;
    double v[1] = {0.0};
    for (int row = t0 + g; row < t1; row += PCG_TB / 8) {
        double acc = 0.0;
        const int64_t e1 = row_ptr[row + 1];
        // four entries per lane per pass: their loads issue together
        for (int64_t e = row_ptr[row] + gl; e < e1; e += 32) {
            int c[4];
            double a[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const bool ok = e + 8 * u < e1;
                c[u] = ok ? col[e + 8 * u] : row;
                a[u] = ok ? vals[e + 8 * u] : 0.0;
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) acc += a[u] * z[c[u]];
        }
#pragma unroll
        for (int o = 4; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
        if (gl == 0) {
            const double pi = first ? z[row] : z[row] + beta * p[row];
            const double qi = first ? acc : acc + beta * q[row];
            p[row] = pi;
            q[row] = qi;
            v[0] += pi * qi;
        }
    }
    block_reduce(v);
    if (threadIdx.x == 0) pqp[T] = v[0];
}

// alpha = r.z / p.q; x += alpha p; r -= alpha q; z = M r; the tile's new r.z
// (into rz_next) and r.r.
__global__ void __launch_bounds__(PCG_TB)
k_pcg_step(const int32_t *__restrict__ tile_start, const int64_t *__restrict__ block_off,
           int max_n, const double *__restrict__ blocks, int ntile, const double *__restrict__ p,
           const double *__restrict__ q, double *__restrict__ x, double *__restrict__ r,
           double *__restrict__ z, const double *__restrict__ rz_cur,
           double *__restrict__ rz_next, const double *__restrict__ pqp, double *__restrict__ rrp,
           int32_t *__restrict__ status, int k) {
    extern __shared__ double part[];
    const double *const parts[2] = {rz_cur, pqp};
    double s[2];
    if (partial_sums(parts, ntile, status, s)) return;
    if (!(s[1] > 0.0)) {   // A not positive on p (the stopping test caught r = 0)
        if (threadIdx.x == 0) {
            status[0] = 2;
            status[1] = k;
        }
        return;
    }
    const double alpha = s[0] / s[1];
    const int T = blockIdx.x;
    const int t0 = tile_start[T], n = min(tile_start[T + 1] - t0, max_n);
    double *rr = part + PCG_TB;
    double v[2] = {0.0, 0.0};
    for (int i = threadIdx.x; i < n; i += PCG_TB) {
        x[t0 + i] += alpha * p[t0 + i];
        const double ri = r[t0 + i] - alpha * q[t0 + i];
        r[t0 + i] = ri;
        rr[i] = ri;
        v[1] += ri * ri;
    }
    __syncthreads();
    v[0] = tile_precond(blocks + block_off[T], n, rr, part, z + t0);
    block_reduce(v);
    if (threadIdx.x == 0) {
        rz_next[T] = v[0];
        rrp[T] = v[1];
    }
}

// y = A x (CSR, FP64), one wavefront per row.
__global__ void __launch_bounds__(256)
k_fp_spmv(int n, const int64_t *__restrict__ row_ptr, const int32_t *__restrict__ col,
          const double *__restrict__ vals, const double *__restrict__ x, double *__restrict__ y) {
    const int lane = threadIdx.x & 63;
    const int nwave = gridDim.x * (blockDim.x >> 6);
    for (int r = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); r < n; r += nwave) {
        double acc = 0.0;
        for (int64_t k = row_ptr[r] + lane; k < row_ptr[r + 1]; k += 64) acc += vals[k] * x[col[k]];
        acc = wave_sum_d(acc);
        if (lane == 0) y[r] = acc;
    }
}

// img[y][x] += flux[s] * u_s (one wavefront per source, atomics where
// patches overlap).
__global__ void __launch_bounds__(256)
k_fp_add_patches(int n, const int32_t *__restrict__ win, const int64_t *__restrict__ poff,
                 const float *__restrict__ patches, const float *__restrict__ flux, int img_w,
                 float *__restrict__ img) {
    const int lane = threadIdx.x & 63;
    const int nwave = gridDim.x * (blockDim.x >> 6);
    for (int s = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); s < n; s += nwave) {
        const int x0 = win[4 * s], y0 = win[4 * s + 1], w = win[4 * s + 2], h = win[4 * s + 3];
        const float f = flux[s];
        const float *u = patches + poff[s];
        for (int y = 0; y < h; ++y)
            for (int x = lane; x < w; x += 64)
                atomicAdd(img + (int64_t)(y0 + y) * img_w + x0 + x, f * u[(int64_t)y * w + x]);
    }
}

int grid_waves(trh_handle *h, long units) {
    long g = (units + 3) / 4;
    const long cap = 8L * h->n_cu;
    return (int)(g < 1 ? 1 : (g > cap ? cap : g));
}


// The pattern of A: for each source s, the sources whose windows intersect
// s's (s itself always).  Sources are numbered in tile order (tile x tile
// pixel cells, ntx x nty of them; tile_first[T] .. tile_end[T] the sources of
// cell T, stile[s] the cell of s), and two windows can only meet when their
// cells are within `reach` of each other in x and in y.  One wavefront per
// source scans those cells in increasing order, which is increasing source
// order, so a row's columns come out sorted.  counts != NULL: write the row
// lengths; else write the columns at row_ptr[s] and the position of the
// diagonal entry at diag[s].
__device__ __forceinline__ bool win_meet(const int32_t *__restrict__ win, int s, int t) {
    const int sw = win[4 * s + 2], sh = win[4 * s + 3], tw = win[4 * t + 2], th = win[4 * t + 3];
    if (sw <= 0 || sh <= 0 || tw <= 0 || th <= 0) return false;
    const int sx0 = win[4 * s], sy0 = win[4 * s + 1], tx0 = win[4 * t], ty0 = win[4 * t + 1];
    return max(sx0, tx0) < min(sx0 + sw, tx0 + tw) && max(sy0, ty0) < min(sy0 + sh, ty0 + th);
}

__global__ void __launch_bounds__(256)
k_fp_overlaps(int n, const int32_t *__restrict__ win, const int32_t *__restrict__ stile, int ntx,
              int nty, const int32_t *__restrict__ tile_first, const int32_t *__restrict__ tile_end,
              int reach, int32_t *__restrict__ counts, const int64_t *__restrict__ row_ptr,
              int32_t *__restrict__ col, int64_t *__restrict__ diag, int32_t *__restrict__ erow) {
    const int lane = threadIdx.x & 63;
    const int nwave = gridDim.x * (blockDim.x >> 6);
    for (int s = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); s < n; s += nwave) {
        const int T = stile[s];
        const int ty = T / ntx, tx = T - ty * ntx;
        int64_t base = counts ? 0 : row_ptr[s];
        int found = 0;
        for (int yy = max(0, ty - reach); yy <= min(nty - 1, ty + reach); ++yy)
            for (int xx = max(0, tx - reach); xx <= min(ntx - 1, tx + reach); ++xx) {
                const int c = yy * ntx + xx;
                const int e = tile_end[c];
                for (int t0 = tile_first[c]; t0 < e; t0 += 64) {
                    const int t = t0 + lane;
                    const bool hit = t < e && (t == s || win_meet(win, s, t));
                    const unsigned long long m = __ballot(hit);
                    if (!counts && hit) {
                        const int64_t k = base + found + __popcll(m & ((1ull << lane) - 1ull));
                        col[k] = t;
                        if (erow) erow[k] = s;
                        if (t == s) diag[s] = k;
                    }
                    found += __popcll(m);
                }
            }
        if (counts && lane == 0) counts[s] = found;
    }
}

// One wavefront per entry (s, t) of A with t >= s (its row from erow): the
// sum of w_s u_t over the intersection of the two windows, w_s = u_s
// inverr^2 the weighted patch (FP32 products, FP64 sums).  The intersection
// is flattened over the lanes, four elements per lane in flight.  The value
// is also written at the mirror entry (t, s), found by bisection of row t.
__global__ void __launch_bounds__(256)
k_fp_pair_dots_csr(int n, const int64_t *__restrict__ row_ptr, const int32_t *__restrict__ col,
                   const int32_t *__restrict__ erow, const int32_t *__restrict__ win,
                   const int64_t *__restrict__ poff, const float *__restrict__ patches,
                   const float *__restrict__ wpatch, double *__restrict__ vals) {
    const int lane = threadIdx.x & 63;
    const int64_t nnz = row_ptr[n];
    const int64_t nwave = (int64_t)gridDim.x * (blockDim.x >> 6);
    for (int64_t k = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); k < nnz;
         k += nwave) {
        const int s = erow[k], t = col[k];
        if (t < s) continue;
        const int sx0 = win[4 * s], sy0 = win[4 * s + 1], sw = win[4 * s + 2], sh = win[4 * s + 3];
        const int tx0 = win[4 * t], ty0 = win[4 * t + 1], tw = win[4 * t + 2], th = win[4 * t + 3];
        const int X0 = max(sx0, tx0), X1 = min(sx0 + sw, tx0 + tw);
        const int Y0 = max(sy0, ty0), Y1 = min(sy0 + sh, ty0 + th);
        const int nx = X1 - X0, ny = Y1 - Y0;
        double acc = 0.0;
        if (nx > 0 && ny > 0) {
            const float *ws = wpatch + poff[s] + (int64_t)(Y0 - sy0) * sw + (X0 - sx0);
            const float *ut = patches + poff[t] + (int64_t)(Y0 - ty0) * tw + (X0 - tx0);
            const int m = nx * ny;
            FlatWalk fw(lane, nx, 64);
            for (int i = lane; i < m; i += 256) {
                float a[4], c[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const bool ok = i + 64 * u < m;
                    a[u] = ok ? ws[fw.y * sw + fw.x] : 0.f;
                    c[u] = ok ? ut[fw.y * tw + fw.x] : 0.f;
                    fw.step();
                }
#pragma unroll
                for (int u = 0; u < 4; ++u) acc += (double)(a[u] * c[u]);
            }
        }
        const double v = wave_sum_d(acc);
        if (lane == 0) {
            vals[k] = v;
            if (t != s) {
                int64_t lo = row_ptr[t], hi = row_ptr[t + 1] - 1;
                while (lo < hi) {
                    const int64_t mid = (lo + hi) >> 1;
                    if (col[mid] < s) lo = mid + 1;
                    else hi = mid;
                }
                vals[lo] = v;
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Planning on the device: sort keys, the unit-flux patch of every source and
// the render jobs grouped by render kind.

// Sort key of source s: the tile cell of its clipped position (tile x tile
// pixels, row-major over ntx cells) times 64 plus the Morton index of its
// 8 x 8 sub-cell, so that a run of consecutive sources is spatially compact.
__global__ void __launch_bounds__(256)
k_fp_keys(int n, const double *__restrict__ src, int img_h, int img_w, int tile, int ntx,
          int32_t *__restrict__ key) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const double x = fmin(fmax(src[6 * (int64_t)s], 0.0), (double)(img_w - 1));
    const double y = fmin(fmax(src[6 * (int64_t)s + 1], 0.0), (double)(img_h - 1));
    const int cx = (int)floor(x / tile), cy = (int)floor(y / tile);
    const int ux = min(7, (int)floor(fmod(x, (double)tile) * 8.0 / tile));
    const int uy = min(7, (int)floor(fmod(y, (double)tile) * 8.0 / tile));
    int m = 0;
#pragma unroll
    for (int b = 0; b < 3; ++b) m |= ((ux >> b) & 1) << (2 * b) | ((uy >> b) & 1) << (2 * b + 1);
    key[s] = (cy * ntx + cx) * 64 + m;
}

__device__ __forceinline__ int64_t pad_center(int64_t sz, int64_t p) {
    return sz >= p ? (sz - p) / 2 + p / 2 : p / 2 - (p - sz) / 2;
}

// Source order[q] -> its unit-flux patch window (clipped to the image), its
// render bucket, window area, reach extent and render job, written at q.
// src rows: x, y, kind, log r_e, ee1, ee2 (EllipseESoft).  mix rows (exp,
// dev): K, then TRH_MAXK amplitudes, then TRH_MAXK variances (circular, in
// units of r_e^2).
//  * point sources (PixelizedPSF.getPointSourcePatch, psf.py:137-207, no
//    modelMask): the Lanczos-shifted stamp at (round(x) - pw/2,
//    round(y) - ph/2), clipped;
//  * galaxies: HoggGalaxy._getUnitFluxPatchSize (galaxy.py:678-691, NullWCS
//    pixel scale 1; half-size nre r_e, nre = 4 exp / 8 dev, at least 1,
//    at most the distance to the far image edge, at least the PSF's
//    half-diagonal), the window of get_overlapping_region (patch.pyx:20-66),
//    and the FFT job of galaxy.py:336-508 (size the next power of two >= 2 hs,
//    the PSF centred by the same rule as the spectrum padding, the sheared
//    MoG var' = v T T^T, T the EllipseESoft basis in pixels,
//    ellipses.py:126-155, 232-255).
// bucket: 0 point, 1 + log2(sz / 32) for sz in 32..512, 6 empty window,
// 7 an FFT size outside 32..512.  ext: the largest distance from the clipped
// position to the window's far edges (0 for an empty window).
__global__ void __launch_bounds__(256)
k_fp_plan(int n, const int64_t *__restrict__ order, const double *__restrict__ src,
          const int32_t *__restrict__ key, int img_h, int img_w, int ph, int pw,
          const double *__restrict__ mix, int32_t *__restrict__ win, int32_t *__restrict__ bucket,
          int32_t *__restrict__ stile, int64_t *__restrict__ area, double *__restrict__ ext,
          trh_fft_job *__restrict__ jobs) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const int64_t s = order[q];
    const double *r = src + 6 * s;
    const double gx = r[0], gy = r[1];
    const int kind = (int)r[2];
    const int64_t W = img_w, H = img_h;
    trh_fft_job J;
    J.out_offset = 0;
    J.psf = 0;
    J.K = 0;
    J.mode = TRH_STORE;
    J.slot = 0;
    J.pad0 = 0;
    J.scale = 1.f;
    J.pad1 = 0.f;
    for (int k = 0; k < TRH_MAXK; ++k) {
        J.amp[k] = 0.f;
        J.var[k][0] = J.var[k][1] = J.var[k][2] = 0.f;
    }
    int64_t wx0, wy0, ww, wh;
    int b;
    if (kind == 0) {
        const double ix = rint(gx), iy = rint(gy);
        const int64_t x0 = (int64_t)ix - pw / 2, y0 = (int64_t)iy - ph / 2;
        const int64_t X0 = max(x0, (int64_t)0), Y0 = max(y0, (int64_t)0);
        const int64_t X1 = min(x0 + pw, W), Y1 = min(y0 + ph, H);
        wx0 = X0;
        wy0 = Y0;
        ww = max(X1 - X0, (int64_t)0);
        wh = max(Y1 - Y0, (int64_t)0);
        J.sx = (int)(x0 - X0);
        J.sy = (int)(y0 - Y0);
        J.mux = gx - ix;
        J.muy = gy - iy;
        b = 0;
    } else {
        const double *m = mix + (kind == 1 ? 0 : 1) * (1 + 2 * TRH_MAXK);
        const double nre = kind == 1 ? 4.0 : 8.0;
        const double re = exp(r[3]);
        double hs = fmax(1.0, nre * re);
        const double imsize = fmax(fmax(gx, (double)(W - 1) - gx), fmax(gy, (double)(H - 1) - gy)) + 1.0;
        hs = fmin(hs, imsize);
        hs = fmax(hs, hypot(ph / 2.0, pw / 2.0));
        const int64_t ihs = (int64_t)ceil(hs);
        const int64_t xlo = (int64_t)floor(gx - ihs), xhi = (int64_t)ceil(gx + ihs + 1);
        const int64_t ylo = (int64_t)floor(gy - ihs), yhi = (int64_t)ceil(gy + ihs + 1);
        const int64_t gx0 = max(xlo, (int64_t)0), gx1 = min(xhi, W - 1) + 1;
        const int64_t gy0 = max(ylo, (int64_t)0), gy1 = min(yhi, H - 1) + 1;
        int64_t sz = 1;
        while (sz < 2 * ihs) sz <<= 1;
        const double dx = gx - pad_center(sz, pw), dy = gy - pad_center(sz, ph);
        const double rx = rint(dx), ry = rint(dy);
        const int64_t ix0 = (int64_t)rx, iy0 = (int64_t)ry;
        const int64_t WX0 = max(ix0, gx0), WY0 = max(iy0, gy0);
        const int64_t WX1 = min(ix0 + sz, gx1), WY1 = min(iy0 + sz, gy1);
        const bool empty = xlo > W - 1 || xhi < 0 || ylo > H - 1 || yhi < 0;
        wx0 = empty ? 0 : WX0;
        wy0 = empty ? 0 : WY0;
        ww = empty ? 0 : max(WX1 - WX0, (int64_t)0);
        wh = empty ? 0 : max(WY1 - WY0, (int64_t)0);
        J.sx = (int)(ix0 - WX0);
        J.sy = (int)(iy0 - WY0);
        J.mux = dx - rx;
        J.muy = dy - ry;
        const double ee = hypot(r[4], r[5]);
        const double e = 1.0 - exp(-ee);
        const double th = atan2(r[5], r[4]) / 2.0;
        const double ct = cos(th), st = sin(th);
        const double ab = e >= 1.0 ? 1000.0 : fmin(1000.0, (1.0 + e) / fmax(1.0 - e, 1e-300));
        const double T00 = re * ct / ab, T01 = re * st, T10 = -re * st / ab, T11 = re * ct;
        const double M00 = T00 * T00 + T01 * T01, M01 = T00 * T10 + T01 * T11,
                     M11 = T10 * T10 + T11 * T11;
        const int K = (int)m[0];
        J.K = K;
        for (int k = 0; k < K; ++k) {
            const double v = m[1 + TRH_MAXK + k];
            J.amp[k] = (float)m[1 + k];
            J.var[k][0] = (float)(v * M00);
            J.var[k][1] = (float)(v * M01);
            J.var[k][2] = (float)(v * M11);
        }
        b = sz == 32 ? 1 : sz == 64 ? 2 : sz == 128 ? 3 : sz == 256 ? 4 : sz == 512 ? 5 : 7;
    }
    if (ww == 0 || wh == 0) b = 6;
    J.mw = (int)ww;
    J.mh = (int)wh;
    J.out_stride = (int)max(ww, (int64_t)1);
    win[4 * q] = (int)wx0;
    win[4 * q + 1] = (int)wy0;
    win[4 * q + 2] = (int)ww;
    win[4 * q + 3] = (int)wh;
    bucket[q] = b;
    stile[q] = key[s] >> 6;
    area[q] = ww * wh;
    double e = 0.0;
    if (ww > 0 && wh > 0) {
        const double px = fmin(fmax(gx, 0.0), (double)(W - 1));
        const double py = fmin(fmax(gy, 0.0), (double)(H - 1));
        e = fmax(fmax(px - wx0, wx0 + ww - px), fmax(py - wy0, wy0 + wh - py));
    }
    ext[q] = e;
    jobs[q] = J;
}

// Jobs regrouped by render bucket: dst[j] = src[border[j]] with the patch's
// offset poff[q], and psf = the job's rank inside its bucket (the index of its
// interpolated stamp or spectrum).
__global__ void __launch_bounds__(256)
k_fp_jobs_bucketed(int n, const int64_t *__restrict__ border, const int32_t *__restrict__ bucket,
                   const int64_t *__restrict__ bstart, const int64_t *__restrict__ poff,
                   const trh_fft_job *__restrict__ src, trh_fft_job *__restrict__ dst) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const int64_t q = border[j];
    trh_fft_job J = src[q];
    J.out_offset = poff[q];
    J.psf = (int)(j - bstart[bucket[q]]);
    dst[j] = J;
}

// Optimizer._get_fitstats (optimize.py:233-376) for one image, one unit-flux
// model per source: a 256-thread workgroup per source over its window (the
// window flattened over the threads), FP32 pixel products and FP64 sums.
// out[10 s + ...]: prochi2, pronpix, promasked, the fracflux ratio
// sum(|m/c - u| |u|) / sum(u^2), its denominator term sum(|u|) / |c|,
// proflux, npix, sum |u| over the window (fracin), 1 if the source was
// counted, and iv = sum (u scale inverr)^2 (_get_iv, optimize.py:378-423).
// The sums run over nz = {u != 0, inverr > 0}; a source with counts 0 or an
// empty nz gets zeros (iv aside).  chi = (data - (model + sky)) inverr;
// model is the image model without sky.
__global__ void __launch_bounds__(256)
k_fp_fitstats(int n, const int32_t *__restrict__ win, const int64_t *__restrict__ poff,
              const float *__restrict__ patches, const double *__restrict__ counts,
              const float *__restrict__ model, const float *__restrict__ data,
              const float *__restrict__ inverr, int img_w, double scale, double sky,
              double *__restrict__ out) {
    constexpr int NS = 10;
    __shared__ double red[NS][4];
    const int s = blockIdx.x;
    const int x0 = win[4 * s], y0 = win[4 * s + 1], w = win[4 * s + 2], h = win[4 * s + 3];
    const double c = counts[s];
    const float fs = (float)scale, fsky = (float)sky;
    double v[NS];
#pragma unroll
    for (int k = 0; k < NS; ++k) v[k] = 0.0;
    if (w > 0 && h > 0) {
        const float *u = patches + poff[s];
        const int m = w * h;
        FlatWalk fw(threadIdx.x, w, 256);
        for (int i = threadIdx.x; i < m; i += 256, fw.step()) {
            const int64_t ib = (int64_t)(y0 + fw.y) * img_w + x0 + fw.x;
            const float ui = u[i], e = inverr[ib], md = model[ib];
            const float t = ui * fs * e;
            v[9] += (double)t * (double)t;
            const float au = fabsf(ui);
            v[7] += au;
            if (e == 0.f) v[2] += au;
            if (ui != 0.f && e > 0.f) {
                const float chi = (data[ib] - (md + fsky)) * e;
                v[0] += (double)(au * chi * chi);
                v[1] += au;
                v[3] += fabs((double)md / c - (double)ui) * au;
                v[4] += (double)ui * (double)ui;
                v[5] += au / fabs(c);
                v[6] += fabs(((double)md - (double)ui * c) / scale) * au;
                v[8] += 1.0;
            }
        }
    }
#pragma unroll
    for (int k = 0; k < NS; ++k) {
        const double t = wave_sum_d(v[k]);
        if ((threadIdx.x & 63) == 0) red[k][threadIdx.x >> 6] = t;
    }
    __syncthreads();
    if (threadIdx.x < NS) {
        const int k = threadIdx.x;
        double t = red[k][0] + red[k][1] + red[k][2] + red[k][3];
        const double npix = red[8][0] + red[8][1] + red[8][2] + red[8][3];
        const double b2 = red[4][0] + red[4][1] + red[4][2] + red[4][3];
        const bool counted = c != 0.0 && npix > 0.0;
        double r;
        if (k == 9) r = t;
        else if (!counted) r = 0.0;
        else if (k == 3) r = t / b2;
        else if (k == 4) r = (red[5][0] + red[5][1] + red[5][2] + red[5][3]);
        else if (k == 5) r = (red[6][0] + red[6][1] + red[6][2] + red[6][3]);
        else if (k == 6) r = npix;
        else if (k == 8) r = 1.0;
        else r = t;   // 0 prochi2, 1 pronpix, 2 promasked, 7 fracin
        out[NS * (int64_t)s + k] = r;
    }
}

// Image chi^2 and the number of pixels with inverr > 0 (the per-image part
// of _get_fitstats, optimize.py:244-250): one workgroup, fixed order.
__global__ void __launch_bounds__(256)
k_fp_imstats_part(int64_t npix, const float *__restrict__ model, const float *__restrict__ data,
                  const float *__restrict__ inverr, double sky, double *__restrict__ part) {
    __shared__ double red[2][4];
    const float fsky = (float)sky;
    double c2 = 0.0, np = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < npix;
         i += (int64_t)gridDim.x * 256) {
        const float e = inverr[i];
        const float chi = (data[i] - (model[i] + fsky)) * e;
        c2 += (double)chi * (double)chi;
        np += e > 0.f ? 1.0 : 0.0;
    }
    c2 = wave_sum_d(c2);
    np = wave_sum_d(np);
    if ((threadIdx.x & 63) == 0) {
        red[0][threadIdx.x >> 6] = c2;
        red[1][threadIdx.x >> 6] = np;
    }
    __syncthreads();
    if (threadIdx.x < 2)
        part[2 * blockIdx.x + threadIdx.x] =
            (red[threadIdx.x][0] + red[threadIdx.x][1]) + (red[threadIdx.x][2] + red[threadIdx.x][3]);
}

// The per-block partials added up in block order (the same bits every run).
__global__ void __launch_bounds__(64)
k_fp_imstats_final(int nblk, const double *__restrict__ part, double *__restrict__ out) {
    const int lane = threadIdx.x;
    double c2 = 0.0, np = 0.0;
    for (int b = lane; b < nblk; b += 64) {
        c2 += part[2 * b];
        np += part[2 * b + 1];
    }
    c2 = wave_sum_d(c2);
    np = wave_sum_d(np);
    if (lane == 0) {
        out[0] = c2;
        out[1] = np;
    }
}

}  // namespace

extern "C" int trh_psf_combine(trh_handle *h, const float *basis, int nbasis, int64_t len,
                               const float *weights, const int32_t *idx, int nq, float *out, int n,
                               void *stream) {
    if (!h || !basis || !weights || !idx || !out || nbasis <= 0 || len <= 0 || n < 0 ||
        n > 65535 || nq < 1 || nq > PSF_MAXQ)
        return set_err(TRH_E_ARG, "trh_psf_combine: bad arguments (n <= 65535, 1 <= nq <= %d)",
                       PSF_MAXQ);
    if (n == 0) return TRH_OK;
    const bool vec = len % 4 == 0 && ((uintptr_t)basis % 16) == 0 && ((uintptr_t)out % 16) == 0;
    if (vec)
        hipLaunchKernelGGL(k_psf_combine<float4>, dim3(n), dim3(256), 0, (hipStream_t)stream, basis,
                           nbasis, len, weights, idx, nq, out, n);
    else
        hipLaunchKernelGGL(k_psf_combine<float>, dim3(n), dim3(256), 0, (hipStream_t)stream, basis,
                           nbasis, len, weights, idx, nq, out, n);
    return hip_check(hipGetLastError(), "k_psf_combine");
}

extern "C" int trh_psfex_terms(trh_handle *h, int n, const int64_t *order, const double *src,
                               int degree, double x0, double y0, double xscale, double yscale,
                               float *weights, int32_t *idx, void *stream) {
    if (!h || n < 0 || (n > 0 && (!src || !weights || !idx)) || degree < 0 ||
        (degree + 1) * (degree + 2) / 2 > PSF_MAXQ || !(xscale != 0.0) || !(yscale != 0.0))
        return set_err(TRH_E_ARG, "trh_psfex_terms: bad arguments (degree %d)", degree);
    if (n == 0) return TRH_OK;
    hipLaunchKernelGGL(k_psfex_terms, dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)stream, n,
                       order, src, degree, x0, y0, xscale, yscale, weights, idx);
    return hip_check(hipGetLastError(), "k_psfex_terms");
}

extern "C" int trh_fp_overlaps(trh_handle *h, int n, const int32_t *win, const int32_t *stile,
                               int ntx, int nty, const int32_t *tile_first,
                               const int32_t *tile_end, int reach, int32_t *counts,
                               const int64_t *row_ptr, int32_t *col, int64_t *diag,
                               int32_t *erow, void *stream) {
    if (!h || n <= 0 || !win || !stile || ntx <= 0 || nty <= 0 || !tile_first || !tile_end ||
        reach < 0 || (!counts && (!row_ptr || !col || !diag)))
        return set_err(TRH_E_ARG, "trh_fp_overlaps: bad arguments");
    hipLaunchKernelGGL(k_fp_overlaps, dim3(grid_waves(h, n)), dim3(256), 0, (hipStream_t)stream, n,
                       win, stile, ntx, nty, tile_first, tile_end, reach, counts, row_ptr, col,
                       diag, counts ? nullptr : erow);
    return hip_check(hipGetLastError(), "k_fp_overlaps");
}

extern "C" int trh_fp_pair_dots_csr(trh_handle *h, int n, const int64_t *row_ptr,
                                    const int32_t *col, const int32_t *erow, const int32_t *win,
                                    const int64_t *poff, const float *patches,
                                    const float *wpatch, double *vals, void *stream) {
    if (!h || n <= 0 || !row_ptr || !col || !erow || !win || !poff || !patches || !wpatch ||
        !vals)
        return set_err(TRH_E_ARG, "trh_fp_pair_dots_csr: bad arguments");
    hipLaunchKernelGGL(k_fp_pair_dots_csr, dim3(grid_waves(h, 64L * n)), dim3(256), 0,
                       (hipStream_t)stream, n, row_ptr, col, erow, win, poff, patches, wpatch,
                       vals);
    return hip_check(hipGetLastError(), "k_fp_pair_dots_csr");
}

extern "C" int trh_fp_pair_dots(trh_handle *h, const int32_t *pairs, const int64_t *pos,
                                int npairs, const int32_t *win, const int64_t *poff,
                                const float *patches, const float *inverr, int img_w,
                                double *vals, void *stream) {
    if (!h || (npairs > 0 && (!pairs || !pos || !win || !poff || !patches || !inverr || !vals)) ||
        img_w <= 0)
        return set_err(TRH_E_ARG, "trh_fp_pair_dots: bad arguments");
    if (npairs <= 0) return TRH_OK;
    hipLaunchKernelGGL(k_fp_pair_dots, dim3(grid_waves(h, npairs)), dim3(256), 0,
                       (hipStream_t)stream, pairs, pos, npairs, win, poff, patches, inverr, img_w,
                       vals);
    return hip_check(hipGetLastError(), "k_fp_pair_dots");
}

extern "C" int trh_fp_rhs(trh_handle *h, int n, const int32_t *win, const int64_t *poff,
                          const float *patches, const float *data, const float *mod0,
                          const float *inverr, int img_w, double *b, float *wpatch,
                          void *stream) {
    if (!h || (n > 0 && (!win || !poff || !patches || !data || !inverr || !b)) || img_w <= 0)
        return set_err(TRH_E_ARG, "trh_fp_rhs: bad arguments");
    if (n <= 0) return TRH_OK;
    hipLaunchKernelGGL(k_fp_rhs, dim3(n), dim3(256), 0, (hipStream_t)stream, n, win, poff,
                       patches, data, mod0, inverr, img_w, b, wpatch);
    return hip_check(hipGetLastError(), "k_fp_rhs");
}

extern "C" int trh_fp_block_inv(trh_handle *h, const int64_t *row_ptr, const int32_t *col,
                                const double *vals, const int32_t *tile_start,
                                const int64_t *block_off, int ntile, int max_n, double *blocks,
                                int32_t *info, void *stream) {
    const size_t lds = ((size_t)max_n * max_n + max_n) * sizeof(double) + (max_n + 1) * 4;
    if (!h || !row_ptr || !col || !vals || !tile_start || !block_off || !blocks || !info ||
        ntile <= 0 || max_n <= 0 || lds + 256 > (size_t)h->lds_max)
        return set_err(TRH_E_ARG, "trh_fp_block_inv: bad arguments (max_n %d)", max_n);
    hipLaunchKernelGGL(k_fp_block_inv, dim3(ntile), dim3(256), lds, (hipStream_t)stream, row_ptr,
                       col, vals, tile_start, block_off, max_n, blocks, info);
    return hip_check(hipGetLastError(), "k_fp_block_inv");
}

static bool pcg_lds_ok(const trh_handle *h, int max_n) {
    return max_n > 0 && ((size_t)PCG_TB + max_n) * sizeof(double) + 256 <= (size_t)h->lds_max;
}

extern "C" int trh_fp_precond(trh_handle *h, const int32_t *tile_start, const int64_t *block_off,
                              int ntile, int max_n, const double *blocks, const double *r,
                              double *z, void *stream) {
    if (!h || !tile_start || !block_off || !blocks || !r || !z || ntile <= 0 ||
        !pcg_lds_ok(h, max_n))
        return set_err(TRH_E_ARG, "trh_fp_precond: bad arguments (max_n %d)", max_n);
    hipLaunchKernelGGL(k_fp_precond, dim3(ntile), dim3(PCG_TB),
                       ((size_t)PCG_TB + max_n) * sizeof(double), (hipStream_t)stream, tile_start,
                       block_off, max_n, blocks, r, z);
    return hip_check(hipGetLastError(), "k_fp_precond");
}

extern "C" int trh_fp_spmv(trh_handle *h, int n, const int64_t *row_ptr, const int32_t *col,
                           const double *vals, const double *x, double *y, void *stream) {
    if (!h || n <= 0 || !row_ptr || !col || !vals || !x || !y)
        return set_err(TRH_E_ARG, "trh_fp_spmv: bad arguments");
    hipLaunchKernelGGL(k_fp_spmv, dim3(grid_waves(h, n)), dim3(256), 0, (hipStream_t)stream, n,
                       row_ptr, col, vals, x, y);
    return hip_check(hipGetLastError(), "k_fp_spmv");
}

extern "C" int trh_fp_pcg(trh_handle *h, int n, const int64_t *row_ptr, const int32_t *col,
                          const double *vals, const int32_t *tile_start, const int64_t *block_off,
                          int ntile, int max_n, const double *blocks, const double *b, double tol,
                          int it0, int iters, double *x, double *work, int32_t *status,
                          void *stream) {
    if (!h || n <= 0 || !row_ptr || !col || !vals || !tile_start || !block_off || ntile <= 0 ||
        ntile > n || !blocks || !b || !x || !work || !status || it0 < 0 || iters < 0 ||
        !(tol >= 0.0) || !pcg_lds_ok(h, max_n))
        return set_err(TRH_E_ARG, "trh_fp_pcg: bad arguments (n %d, ntile %d, max_n %d)", n,
                       ntile, max_n);
    hipStream_t st = (hipStream_t)stream;
    double *r = work, *z = r + n, *p = z + n, *q = p + n;
    double *rzp = q + n, *rrp = rzp + 2 * (size_t)ntile, *pqp = rrp + ntile,
           *bbp = pqp + ntile, *norms = bbp + ntile;
    const size_t lds = ((size_t)PCG_TB + max_n) * sizeof(double);
    const double tol2 = tol * tol;
    // r.z after k steps lives in rzp[k & 1]
    auto rz = [&](int k) { return rzp + (size_t)(k & 1) * ntile; };
    int rc;
    if (it0 == 0) {
        hipLaunchKernelGGL(k_pcg_init, dim3(ntile), dim3(PCG_TB), lds, st, tile_start, block_off,
                           max_n, blocks, b, x, r, z, rz(0), rrp, bbp, status);
        hipLaunchKernelGGL(k_pcg_dir, dim3(ntile), dim3(PCG_TB), 0, st, tile_start, row_ptr, col,
                           vals, ntile, z, p, q, rz(0), rz(1), rrp, bbp, pqp, tol2, status, norms,
                           0, 1);
        if ((rc = hip_check(hipGetLastError(), "k_pcg_init"))) return rc;
    }
    for (int k = it0; k < it0 + iters; ++k) {
        hipLaunchKernelGGL(k_pcg_step, dim3(ntile), dim3(PCG_TB), lds, st, tile_start, block_off,
                           max_n, blocks, ntile, p, q, x, r, z, rz(k), rz(k + 1), pqp, rrp,
                           status, k);
        hipLaunchKernelGGL(k_pcg_dir, dim3(ntile), dim3(PCG_TB), 0, st, tile_start, row_ptr, col,
                           vals, ntile, z, p, q, rz(k + 1), rz(k), rrp, bbp, pqp, tol2, status,
                           norms, k + 1, 0);
        if ((rc = hip_check(hipGetLastError(), "k_pcg"))) return rc;
    }
    return TRH_OK;
}

extern "C" int trh_fp_add_patches(trh_handle *h, int n, const int32_t *win, const int64_t *poff,
                                  const float *patches, const float *flux, int img_w, float *img,
                                  void *stream) {
    if (!h || (n > 0 && (!win || !poff || !patches || !flux || !img)) || img_w <= 0)
        return set_err(TRH_E_ARG, "trh_fp_add_patches: bad arguments");
    if (n <= 0) return TRH_OK;
    hipLaunchKernelGGL(k_fp_add_patches, dim3(grid_waves(h, n)), dim3(256), 0,
                       (hipStream_t)stream, n, win, poff, patches, flux, img_w, img);
    return hip_check(hipGetLastError(), "k_fp_add_patches");
}

extern "C" int trh_fp_keys(trh_handle *h, int n, const double *src, int img_h, int img_w,
                           int tile, int32_t *key, void *stream) {
    if (!h || n < 0 || (n > 0 && (!src || !key)) || img_h <= 0 || img_w <= 0 || tile <= 0)
        return set_err(TRH_E_ARG, "trh_fp_keys: bad arguments");
    const int ntx = (img_w + tile - 1) / tile, nty = (img_h + tile - 1) / tile;
    if ((int64_t)ntx * nty * 64 > INT32_MAX)
        return set_err(TRH_E_ARG, "trh_fp_keys: %d x %d cells exceed the key range", ntx, nty);
    if (n == 0) return TRH_OK;
    hipLaunchKernelGGL(k_fp_keys, dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)stream, n, src,
                       img_h, img_w, tile, ntx, key);
    return hip_check(hipGetLastError(), "k_fp_keys");
}

extern "C" int trh_fp_plan(trh_handle *h, int n, const int64_t *order, const double *src,
                           const int32_t *key, int img_h, int img_w, int ph, int pw,
                           const double *mix, int32_t *win, int32_t *bucket, int32_t *stile,
                           int64_t *area, double *ext, trh_fft_job *jobs, void *stream) {
    if (!h || n < 0 ||
        (n > 0 && (!order || !src || !key || !mix || !win || !bucket || !stile || !area || !ext ||
                   !jobs)) ||
        img_h <= 0 || img_w <= 0 || ph <= 0 || pw <= 0)
        return set_err(TRH_E_ARG, "trh_fp_plan: bad arguments");
    if (n == 0) return TRH_OK;
    hipLaunchKernelGGL(k_fp_plan, dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)stream, n,
                       order, src, key, img_h, img_w, ph, pw, mix, win, bucket, stile, area, ext,
                       jobs);
    return hip_check(hipGetLastError(), "k_fp_plan");
}

extern "C" int trh_fp_jobs_bucketed(trh_handle *h, int n, const int64_t *border,
                                    const int32_t *bucket, const int64_t *bstart,
                                    const int64_t *poff, const trh_fft_job *src, trh_fft_job *dst,
                                    void *stream) {
    if (!h || n < 0 || (n > 0 && (!border || !bucket || !bstart || !poff || !src || !dst)))
        return set_err(TRH_E_ARG, "trh_fp_jobs_bucketed: bad arguments");
    if (n == 0) return TRH_OK;
    hipLaunchKernelGGL(k_fp_jobs_bucketed, dim3((n + 255) / 256), dim3(256), 0,
                       (hipStream_t)stream, n, border, bucket, bstart, poff, src, dst);
    return hip_check(hipGetLastError(), "k_fp_jobs_bucketed");
}

extern "C" int trh_fp_fitstats(trh_handle *h, int n, const int32_t *win, const int64_t *poff,
                               const float *patches, const double *counts, const float *model,
                               const float *data, const float *inverr, int img_w, double scale,
                               double sky, double *out, void *stream) {
    if (!h || n < 0 ||
        (n > 0 && (!win || !poff || !patches || !counts || !model || !data || !inverr || !out)) ||
        img_w <= 0 || !(scale != 0.0))
        return set_err(TRH_E_ARG, "trh_fp_fitstats: bad arguments");
    if (n == 0) return TRH_OK;
    hipLaunchKernelGGL(k_fp_fitstats, dim3(n), dim3(256), 0, (hipStream_t)stream, n, win, poff,
                       patches, counts, model, data, inverr, img_w, scale, sky, out);
    return hip_check(hipGetLastError(), "k_fp_fitstats");
}

extern "C" int trh_fp_imstats(trh_handle *h, int64_t npix, const float *model, const float *data,
                              const float *inverr, double sky, double *out, void *stream) {
    if (!h || npix < 0 || (npix > 0 && (!model || !data || !inverr)) || !out)
        return set_err(TRH_E_ARG, "trh_fp_imstats: bad arguments");
    // about eight pixels per thread, at most 2048 blocks of 256
    int64_t nblk = (npix + 2047) / 2048;
    if (nblk < 1) nblk = 1;
    if (nblk > 2048) nblk = 2048;
    int rc;
    hipStream_t st = (hipStream_t)stream;
    if ((rc = shared_begin(h, st))) return rc;
    if ((rc = ensure_scratch(h, 16 * (size_t)nblk))) return rc;
    double *part = (double *)h->scratch;
    hipLaunchKernelGGL(k_fp_imstats_part, dim3((int)nblk), dim3(256), 0, st, npix, model, data,
                       inverr, sky, part);
    if ((rc = hip_check(hipGetLastError(), "k_fp_imstats_part"))) return rc;
    hipLaunchKernelGGL(k_fp_imstats_final, dim3(1), dim3(64), 0, st, (int)nblk, part, out);
    if ((rc = hip_check(hipGetLastError(), "k_fp_imstats_final"))) return rc;
    return shared_end(h, st);
}
