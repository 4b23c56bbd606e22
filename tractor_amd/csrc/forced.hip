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

// out[s][e] = sum_q w[s][q] * basis[idx[s][q]][e], q < 4 (bilinear PSF
// interpolation; the same linear map applies to stamps and to spectra).
__global__ void k_psf_combine(const float *__restrict__ basis, int nbasis, int64_t len,
                              const float *__restrict__ w, const int32_t *__restrict__ idx,
                              float *__restrict__ out, int n) {
    const int s = blockIdx.y;
    if (s >= n) return;
    float ws[4];
    const float *b[4];
    for (int q = 0; q < 4; ++q) {
        int k = idx[4 * s + q];
        k = k < 0 ? 0 : (k >= nbasis ? nbasis - 1 : k);
        ws[q] = w[4 * s + q];
        b[q] = basis + (int64_t)k * len;
    }
    float *o = out + (int64_t)s * len;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < len;
         e += (int64_t)gridDim.x * blockDim.x)
        o[e] = ws[0] * b[0][e] + ws[1] * b[1][e] + ws[2] * b[2][e] + ws[3] * b[3][e];
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

// One wavefront per source: b_s = sum u_s (data - mod0) inverr^2 over its
// window (mod0 may be NULL: zero).
__global__ void __launch_bounds__(256)
k_fp_rhs(int n, const int32_t *__restrict__ win, const int64_t *__restrict__ poff,
         const float *__restrict__ patches, const float *__restrict__ data,
         const float *__restrict__ mod0, const float *__restrict__ inverr, int img_w,
         double *__restrict__ b) {
    const int lane = threadIdx.x & 63;
    const int nwave = gridDim.x * (blockDim.x >> 6);
    for (int s = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); s < n; s += nwave) {
        const int x0 = win[4 * s], y0 = win[4 * s + 1], w = win[4 * s + 2], h = win[4 * s + 3];
        const float *u = patches + poff[s];
        float acc = 0.f;
        for (int y = 0; y < h; ++y) {
            const int64_t ib = (int64_t)(y0 + y) * img_w + x0;
            for (int x = lane; x < w; x += 64) {
                const float e = inverr[ib + x];
                const float d = data[ib + x] - (mod0 ? mod0[ib + x] : 0.f);
                acc += u[(int64_t)y * w + x] * d * (e * e);
            }
        }
        const double v = wave_sum_d((double)acc);
        if (lane == 0) b[s] = v;
    }
}

// One workgroup per tile of rows [t0, t1): copy A's diagonal block into a
// dense FP64 matrix and factor it (lower Cholesky L, right-looking); then
// invert the factor, so that applying the preconditioner is two dense
// triangular products with no per-row barrier.  Layout after the call: the
// diagonal holds L_ii, the lower triangle L, the strict upper triangle the
// transpose of the strict lower part of L^-1 (B[j][i] = (L^-1)_ij, i > j;
// (L^-1)_ii = 1 / L_ii).  A zero diagonal (a source with no pixels) becomes 1
// so its flux solves to 0.
__global__ void __launch_bounds__(256)
k_fp_block_chol(const int64_t *__restrict__ row_ptr, const int32_t *__restrict__ col,
                const double *__restrict__ vals, const int32_t *__restrict__ tile_start,
                const int64_t *__restrict__ block_off, double *__restrict__ blocks,
                int32_t *__restrict__ info) {
    const int T = blockIdx.x;
    const int t0 = tile_start[T], n = tile_start[T + 1] - t0;
    double *B = blocks + block_off[T];
    for (int i = threadIdx.x; i < n * n; i += blockDim.x) B[i] = 0.0;
    __syncthreads();
    for (int r = threadIdx.x; r < n; r += blockDim.x) {
        const int row = t0 + r;
        for (int64_t k = row_ptr[row]; k < row_ptr[row + 1]; ++k) {
            const int c = col[k] - t0;
            if (c >= 0 && c < n) B[r * n + c] = vals[k];
        }
    }
    __syncthreads();
    __shared__ double dj;
    for (int j = 0; j < n; ++j) {
        if (threadIdx.x == 0) {
            double d = B[j * n + j];
            if (d == 0.0) d = 1.0;
            if (!(d > 0.0)) {
                info[T] = j + 1;
                d = 1.0;
            }
            d = sqrt(d);
            B[j * n + j] = d;
            dj = d;
        }
        __syncthreads();
        for (int i = j + 1 + threadIdx.x; i < n; i += blockDim.x) B[i * n + j] /= dj;
        __syncthreads();
        const int m = n - j - 1;
        for (int q = threadIdx.x; q < m * m; q += blockDim.x) {
            const int i = j + 1 + q / m, k = j + 1 + q % m;
            if (k <= i) B[i * n + k] -= B[i * n + j] * B[k * n + j];
        }
        __syncthreads();
    }
    // column j of L^-1 by forward substitution (one thread per column); it
    // reads only the lower triangle and writes row j of the upper one
    for (int j = threadIdx.x; j < n; j += blockDim.x)
        for (int i = j + 1; i < n; ++i) {
            double acc = B[i * n + j] / B[j * n + j];   // k = j term: L_ij (L^-1)_jj
            for (int k = j + 1; k < i; ++k) acc += B[i * n + k] * B[j * n + k];
            B[j * n + i] = -acc / B[i * n + i];
        }
}

// z = (L L^T)^-1 r = L^-T (L^-1 r) on each tile, from the inverse factor
// stored by k_fp_block_chol: two dense triangular products.
__global__ void __launch_bounds__(256)
k_fp_precond(const int32_t *__restrict__ tile_start, const int64_t *__restrict__ block_off,
             const double *__restrict__ blocks, const double *__restrict__ r,
             double *__restrict__ z) {
    extern __shared__ double y[];
    double *rr = y + (tile_start[blockIdx.x + 1] - tile_start[blockIdx.x]);
    const int T = blockIdx.x;
    const int t0 = tile_start[T], n = tile_start[T + 1] - t0;
    const double *B = blocks + block_off[T];
    for (int i = threadIdx.x; i < n; i += blockDim.x) rr[i] = r[t0 + i];
    __syncthreads();
    // y_i = sum_{k < i} (L^-1)_ik r_k + r_i / L_ii,  (L^-1)_ik = B[k][i]
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        double acc = rr[i] / B[i * n + i];
        for (int k = 0; k < i; ++k) acc += B[k * n + i] * rr[k];
        y[i] = acc;
    }
    __syncthreads();
    // z_i = sum_{k > i} (L^-1)_ki y_k + y_i / L_ii,  (L^-1)_ki = B[i][k]
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        double acc = y[i] / B[i * n + i];
        for (int k = i + 1; k < n; ++k) acc += B[i * n + k] * y[k];
        z[t0 + i] = acc;
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
              int32_t *__restrict__ col, int64_t *__restrict__ diag) {
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
                        if (t == s) diag[s] = k;
                    }
                    found += __popcll(m);
                }
            }
        if (counts && lane == 0) counts[s] = found;
    }
}

// One wavefront per entry (s, t) of A with t >= s: the sum of u_s u_t
// inverr^2 over the intersection of the two windows, also written at the
// mirror entry (t, s), found by binary search in row t.  The row of entry k
// is found by binary search in row_ptr.
__global__ void __launch_bounds__(256)
k_fp_pair_dots_csr(int n, const int64_t *__restrict__ row_ptr, const int32_t *__restrict__ col,
                   const int32_t *__restrict__ win, const int64_t *__restrict__ poff,
                   const float *__restrict__ patches, const float *__restrict__ inverr, int img_w,
                   double *__restrict__ vals) {
    const int lane = threadIdx.x & 63;
    const int64_t nnz = row_ptr[n];
    const int64_t nwave = (int64_t)gridDim.x * (blockDim.x >> 6);
    for (int64_t k = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); k < nnz;
         k += nwave) {
        // s = the row holding entry k: the last r with row_ptr[r] <= k
        int lo = 0, hi = n - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (row_ptr[mid] <= k) lo = mid;
            else hi = mid - 1;
        }
        const int s = lo, t = col[k];
        if (t < s) continue;
        const int sx0 = win[4 * s], sy0 = win[4 * s + 1], sw = win[4 * s + 2];
        const int tx0 = win[4 * t], ty0 = win[4 * t + 1], tw = win[4 * t + 2], th = win[4 * t + 3];
        const int sh = win[4 * s + 3];
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
            vals[k] = v;
            if (t != s) {
                int64_t a = row_ptr[t], b = row_ptr[t + 1] - 1;
                while (a < b) {
                    const int64_t mid = (a + b) >> 1;
                    if (col[mid] < s) a = mid + 1;
                    else b = mid;
                }
                vals[a] = v;
            }
        }
    }
}
}  // namespace

extern "C" int trh_psf_combine(trh_handle *h, const float *basis, int nbasis, int64_t len,
                               const float *weights, const int32_t *idx, float *out, int n,
                               void *stream) {
    if (!h || !basis || !weights || !idx || !out || nbasis <= 0 || len <= 0 || n < 0 || n > 65535)
        return set_err(TRH_E_ARG, "trh_psf_combine: bad arguments (n <= 65535)");
    if (n == 0) return TRH_OK;
    long gx = (len + 255) / 256;
    if (gx > 64) gx = 64;
    hipLaunchKernelGGL(k_psf_combine, dim3((int)gx, n), dim3(256), 0, (hipStream_t)stream, basis,
                       nbasis, len, weights, idx, out, n);
    return hip_check(hipGetLastError(), "k_psf_combine");
}

extern "C" int trh_fp_overlaps(trh_handle *h, int n, const int32_t *win, const int32_t *stile,
                               int ntx, int nty, const int32_t *tile_first,
                               const int32_t *tile_end, int reach, int32_t *counts,
                               const int64_t *row_ptr, int32_t *col, int64_t *diag,
                               void *stream) {
    if (!h || n <= 0 || !win || !stile || ntx <= 0 || nty <= 0 || !tile_first || !tile_end ||
        reach < 0 || (!counts && (!row_ptr || !col || !diag)))
        return set_err(TRH_E_ARG, "trh_fp_overlaps: bad arguments");
    hipLaunchKernelGGL(k_fp_overlaps, dim3(grid_waves(h, n)), dim3(256), 0, (hipStream_t)stream, n,
                       win, stile, ntx, nty, tile_first, tile_end, reach, counts, row_ptr, col,
                       diag);
    return hip_check(hipGetLastError(), "k_fp_overlaps");
}

extern "C" int trh_fp_pair_dots_csr(trh_handle *h, int n, const int64_t *row_ptr,
                                    const int32_t *col, const int32_t *win, const int64_t *poff,
                                    const float *patches, const float *inverr, int img_w,
                                    double *vals, void *stream) {
    if (!h || n <= 0 || !row_ptr || !col || !win || !poff || !patches || !inverr || img_w <= 0 ||
        !vals)
        return set_err(TRH_E_ARG, "trh_fp_pair_dots_csr: bad arguments");
    hipLaunchKernelGGL(k_fp_pair_dots_csr, dim3(grid_waves(h, 64L * n)), dim3(256), 0,
                       (hipStream_t)stream, n, row_ptr, col, win, poff, patches, inverr, img_w,
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
                          const float *inverr, int img_w, double *b, void *stream) {
    if (!h || (n > 0 && (!win || !poff || !patches || !data || !inverr || !b)) || img_w <= 0)
        return set_err(TRH_E_ARG, "trh_fp_rhs: bad arguments");
    if (n <= 0) return TRH_OK;
    hipLaunchKernelGGL(k_fp_rhs, dim3(grid_waves(h, n)), dim3(256), 0, (hipStream_t)stream, n,
                       win, poff, patches, data, mod0, inverr, img_w, b);
    return hip_check(hipGetLastError(), "k_fp_rhs");
}

extern "C" int trh_fp_block_chol(trh_handle *h, const int64_t *row_ptr, const int32_t *col,
                                 const double *vals, const int32_t *tile_start,
                                 const int64_t *block_off, int ntile, double *blocks,
                                 int32_t *info, void *stream) {
    if (!h || !row_ptr || !col || !vals || !tile_start || !block_off || !blocks || !info ||
        ntile <= 0)
        return set_err(TRH_E_ARG, "trh_fp_block_chol: bad arguments");
    hipLaunchKernelGGL(k_fp_block_chol, dim3(ntile), dim3(256), 0, (hipStream_t)stream, row_ptr,
                       col, vals, tile_start, block_off, blocks, info);
    return hip_check(hipGetLastError(), "k_fp_block_chol");
}

extern "C" int trh_fp_precond(trh_handle *h, const int32_t *tile_start, const int64_t *block_off,
                              int ntile, int max_n, const double *blocks, const double *r,
                              double *z, void *stream) {
    if (!h || !tile_start || !block_off || !blocks || !r || !z || ntile <= 0 || max_n <= 0 ||
        2 * (size_t)max_n * sizeof(double) > (size_t)h->lds_max)
        return set_err(TRH_E_ARG, "trh_fp_precond: bad arguments (max_n %d)", max_n);
    hipLaunchKernelGGL(k_fp_precond, dim3(ntile), dim3(256), 2 * sizeof(double) * (size_t)max_n,
                       (hipStream_t)stream, tile_start, block_off, blocks, r, z);
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
