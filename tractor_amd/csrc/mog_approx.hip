// Adaptive ring-growth MoG evaluation (c_gauss_2d_approx3,
// tractor/approx3.c:1-465, with eval_all / eval_all_dxy, tractor/mix.i:31-82)
// for gfx950.
//
// One workgroup per job.  The reference grows square rings outward from a
// start pixel; a ring pixel is evaluated only if the previous ring marked it,
// and only pixels at or above `minval` (or inside `minradius`) mark their
// outward neighbours.  Each ring depends only on the previous ring's marks, so
// the ring loop is sequential but every ring's pixels (up to 2w + 2h) are
// evaluated in parallel by the workgroup.  Marks live in LDS as two byte
// buffers [top: w][bottom: w][left: h][right: h] used alternately; concurrent
// writes of a mark all store 1, and a corner pixel that two sides evaluate gets
// the same value from both.  The evaluation is FP64 like the reference's.
#include "internal.h"

using namespace trh;

namespace {

constexpr int kThreads = 256;

struct Comp {
    double i0, i1, i2;   // icov with -1/2 and the cross-term factor 2 folded in
    double scale;        // amp / (2 pi sqrt(det))
    double maxd;         // exponent floor (eval_all_dxy)
    double mx, my;       // mean
};

// eval_all: exponent floor -100, no derivatives (start-pixel search).
__device__ double eval_all(const Comp *C, int K, double x, double y) {
    double r = 0.0;
    for (int k = 0; k < K; ++k) {
        const double dx = x - C[k].mx, dy = y - C[k].my;
        const double d = C[k].i0 * dx * dx + C[k].i1 * dx * dy + C[k].i2 * dy * dy;
        if (d < -100.0) continue;
        r += C[k].scale * exp(d);
    }
    return r;
}

// eval_all_dxy: per-component floor maxd, non-finite exponents skipped;
// derivatives with respect to the mean.
__device__ double eval_dxy(const Comp *C, int K, double x, double y, double *gx, double *gy) {
    double r = 0.0, ax = 0.0, ay = 0.0;
    for (int k = 0; k < K; ++k) {
        const double dx = x - C[k].mx, dy = y - C[k].my;
        const double d = C[k].i0 * dx * dx + C[k].i1 * dx * dy + C[k].i2 * dy * dy;
        if (!(d >= C[k].maxd)) continue;
        if (!isfinite(d)) continue;
        const double G = C[k].scale * exp(d);
        r += G;
        ax += -G * (2.0 * C[k].i0 * dx + C[k].i1 * dy);
        ay += -G * (2.0 * C[k].i2 * dy + C[k].i1 * dx);
    }
    *gx = ax;
    *gy = ay;
    return r;
}

__device__ __forceinline__ void mark(unsigned char *a, int i, int n) {
    if (i >= 0 && i < n) a[i] = 1;
}

__global__ void __launch_bounds__(kThreads)
k_mog_approx3(const trh_mog_approx_job *__restrict__ jobs, int njobs, int max_wh,
              const double *__restrict__ amp, const double *__restrict__ mean,
              const double *__restrict__ var, double *__restrict__ out,
              double *__restrict__ dxo, double *__restrict__ dyo, int32_t *__restrict__ bbox) {
    extern __shared__ unsigned char lds_raw[];
    Comp *C = reinterpret_cast<Comp *>(lds_raw);
    double *redv = reinterpret_cast<double *>(C + TRH_APPROX_MAXK);
    int *redi = reinterpret_cast<int *>(redv + kThreads);
    int *st = redi + kThreads;    // [0..3] bbox, [4..6] ring flags, [7] xc, [8] yc
    unsigned char *marks = reinterpret_cast<unsigned char *>(st + 16);
    const int tid = threadIdx.x;

    for (int j = blockIdx.x; j < njobs; j += gridDim.x) {
        const trh_mog_approx_job J = jobs[j];
        const int W = J.w, H = J.h, K = J.K;
        const int x0 = J.x0, y0 = J.y0, x1 = x0 + W, y1 = y0 + H;
        // Invalid jobs are reported, not evaluated.
        if (W <= 0 || H <= 0 || K <= 0 || K > TRH_APPROX_MAXK || W + H > max_wh) {
            if (tid == 0) {
                bbox[4 * j + 0] = -1;
                bbox[4 * j + 1] = -1;
                bbox[4 * j + 2] = -1;
                bbox[4 * j + 3] = -1;
            }
            continue;
        }
        const bool derivs = J.want_derivs && dxo && dyo;
        double *o = out + J.out_offset;
        double *ox = derivs ? dxo + J.out_offset : nullptr;
        double *oy = derivs ? dyo + J.out_offset : nullptr;
        const int S = J.out_stride;

        // ---- component constants (approx3.c:96-126)
        if (tid == 0) {
            double ampsum = 0.0;
            for (int k = 0; k < K; ++k) ampsum += amp[J.comp0 + k];
            redv[0] = ampsum;
        }
        __syncthreads();
        const double ampsum = redv[0];
        const double tpd = 4.0 * M_PI * M_PI;
        for (int k = tid; k < K; k += blockDim.x) {
            const double *v = var + 4 * (size_t)(J.comp0 + k);
            const double V0 = v[0], V1 = (v[1] + v[2]) * 0.5, V2 = v[3];
            const double det = V0 * V2 - V1 * V1;
            const double isc = -0.5 / det;
            Comp c;
            c.i0 = V2 * isc;
            c.i1 = -V1 * isc * 2.0;
            c.i2 = V0 * isc;
            c.scale = amp[J.comp0 + k] / sqrt(tpd * det);
            if (J.minval > 0.0) {
                c.maxd = log(J.minval * sqrt(tpd * det) / ampsum);
                if (c.maxd < -100.0) c.maxd = -100.0;
            } else {
                c.maxd = -100.0;
            }
            c.mx = mean[2 * (size_t)(J.comp0 + k)];
            c.my = mean[2 * (size_t)(J.comp0 + k) + 1];
            C[k] = c;
        }
        // ---- clear the output window
        for (int p = tid; p < W * H; p += blockDim.x) {
            const int yy = p / W, xx = p - yy * W;
            o[(size_t)yy * S + xx] = 0.0;
            if (derivs) {
                ox[(size_t)yy * S + xx] = 0.0;
                oy[(size_t)yy * S + xx] = 0.0;
            }
        }
        __syncthreads();

        // ---- start pixel (approx3.c:128-170): first maximum along the nearest
        // edge(s), scanning the x edge (iy ascending) then the y edge (ix
        // ascending); the box centre when no value is positive.
        int xc = J.xc, yc = J.yc;
        if (xc < x0 || yc < y0 || xc >= x1 || yc >= y1) {
            const bool xout = (xc < x0 || xc >= x1), yout = (yc < y0 || yc >= y1);
            const int n1 = xout ? H : 0, n2 = yout ? W : 0;
            double best = 0.0;
            int bi = 0x7fffffff;
            for (int i = tid; i < n1 + n2; i += blockDim.x) {
                int ix, iy;
                if (i < n1) {
                    ix = (xc < x0) ? x0 : x1 - 1;
                    iy = y0 + i;
                } else {
                    ix = x0 + (i - n1);
                    iy = (yc < y0) ? y0 : y1 - 1;
                }
                const double v = eval_all(C, K, ix - J.fx, iy - J.fy);
                if (v > best) {
                    best = v;
                    bi = i;
                }
            }
            redv[tid] = best;
            redi[tid] = bi;
            __syncthreads();
            for (int s = blockDim.x / 2; s > 0; s >>= 1) {
                if (tid < s) {
                    const double vb = redv[tid + s];
                    const int ib = redi[tid + s];
                    if (vb > redv[tid] || (vb == redv[tid] && ib < redi[tid])) {
                        redv[tid] = vb;
                        redi[tid] = ib;
                    }
                }
                __syncthreads();
            }
            const int win = redi[0];
            if (redv[0] > 0.0 && win != 0x7fffffff) {
                if (win < n1) {
                    xc = (J.xc < x0) ? x0 : x1 - 1;
                    yc = y0 + win;
                } else {
                    xc = x0 + (win - n1);
                    yc = (J.yc < y0) ? y0 : y1 - 1;
                }
            } else {
                xc = (x0 + x1) / 2;
                yc = (y0 + y1) / 2;
            }
            __syncthreads();
        }

        // ---- marks: buffer b holds [T: W][B: W][L: H][R: H]
        const int nm = 2 * W + 2 * H;
        unsigned char *buf0 = marks, *buf1 = marks + max_wh * 2;
        for (int i = tid; i < nm; i += blockDim.x) {
            buf0[i] = 0;
            buf1[i] = 0;
        }
        if (tid < 3) st[4 + tid] = 0;
        __syncthreads();
        if (tid == 0) {
            unsigned char *T = buf0, *B = buf0 + W, *L = buf0 + 2 * W, *R = buf0 + 2 * W + H;
            if (yc > y0) {
                mark(B, xc + 1 - x0, W);
                mark(B, xc - x0, W);
                mark(B, xc - 1 - x0, W);
            }
            if (yc < y1 - 1) {
                mark(T, xc - x0, W);
                mark(T, xc + 1 - x0, W);
                mark(T, xc - 1 - x0, W);
            }
            if (xc > x0) mark(L, yc - y0, H);
            if (xc < x1 - 1) mark(R, yc - y0, H);
            double gx, gy;
            const double r = eval_dxy(C, K, xc - J.fx, yc - J.fy, &gx, &gy);
            const size_t off = (size_t)(yc - y0) * S + (xc - x0);
            o[off] = r;
            if (derivs) {
                ox[off] = gx;
                oy[off] = gy;
            }
            st[0] = xc - x0;  // sx0
            st[1] = xc - x0;  // sx1
            st[2] = yc - y0;  // sy0
            st[3] = yc - y0;  // sy1
        }
        __syncthreads();

        // ---- rings (approx3.c:200-440); at most W + H + 2 of them
        for (int R = 1; R <= W + H + 2; ++R) {
            unsigned char *dob = (R & 1) ? buf0 : buf1;
            unsigned char *nxb = (R & 1) ? buf1 : buf0;
            for (int i = tid; i < nm; i += blockDim.x) nxb[i] = 0;
            if (tid == 0) st[4 + (R + 1) % 3] = 0;
            __syncthreads();
            unsigned char *nT = nxb, *nB = nxb + W, *nL = nxb + 2 * W, *nR = nxb + 2 * W + H;
            for (int i = tid; i < nm; i += blockDim.x) {
                if (!dob[i]) continue;
                int side, xx, yy;
                if (i < W) {
                    side = 0;
                    xx = x0 + i;
                    yy = yc + R;
                    if (yy >= y1) continue;
                } else if (i < 2 * W) {
                    side = 1;
                    xx = x0 + (i - W);
                    yy = yc - R;
                    if (yy < y0) continue;
                } else if (i < 2 * W + H) {
                    side = 2;
                    xx = xc - R;
                    yy = y0 + (i - 2 * W);
                    if (xx < x0) continue;
                } else {
                    side = 3;
                    xx = xc + R;
                    yy = y0 + (i - 2 * W - H);
                    if (xx >= x1) continue;
                }
                double gx, gy;
                const double r = eval_dxy(C, K, xx - J.fx, yy - J.fy, &gx, &gy);
                const size_t off = (size_t)(yy - y0) * S + (xx - x0);
                o[off] = r;
                if (derivs) {
                    ox[off] = gx;
                    oy[off] = gy;
                }
                st[4 + R % 3] = 1;
                if (side == 0) atomicMax(&st[3], yy - y0);
                else if (side == 1) atomicMin(&st[2], yy - y0);
                else if (side == 2) atomicMin(&st[0], xx - x0);
                else atomicMax(&st[1], xx - x0);
                if (R > J.minradius && r < J.minval) continue;
                if (side == 0) {
                    if (xx == xc - R && xx > x0) {
                        mark(nL, yy - y0, H);
                        mark(nL, yy - 1 - y0, H);
                    }
                    if (xx == xc + R && xx < x1 - 1) {
                        mark(nR, yy - y0, H);
                        mark(nR, yy - 1 - y0, H);
                    }
                    if (yy < y1 - 1) {
                        mark(nT, xx - 1 - x0, W);
                        mark(nT, xx - x0, W);
                        mark(nT, xx + 1 - x0, W);
                    }
                } else if (side == 1) {
                    if (xx == xc - R && xx > x0) {
                        mark(nL, yy - y0, H);
                        mark(nL, yy + 1 - y0, H);
                    }
                    if (xx == xc + R && xx < x1 - 1) {
                        mark(nR, yy - y0, H);
                        mark(nR, yy + 1 - y0, H);
                    }
                    if (yy > y0) {
                        mark(nB, xx - 1 - x0, W);
                        mark(nB, xx - x0, W);
                        mark(nB, xx + 1 - x0, W);
                    }
                } else if (side == 2) {
                    if (xx > x0) {
                        mark(nL, yy - 1 - y0, H);
                        mark(nL, yy - y0, H);
                        mark(nL, yy + 1 - y0, H);
                    }
                } else {
                    if (xx < x1 - 1) {
                        mark(nR, yy - 1 - y0, H);
                        mark(nR, yy - y0, H);
                        mark(nR, yy + 1 - y0, H);
                    }
                }
            }
            __syncthreads();
            if (!st[4 + R % 3]) break;
        }
        if (tid == 0) {
            bbox[4 * j + 0] = st[0];
            bbox[4 * j + 1] = st[1] + 1;
            bbox[4 * j + 2] = st[2];
            bbox[4 * j + 3] = st[3] + 1;
        }
        __syncthreads();
    }
}

}  // namespace

extern "C" int trh_render_mog_approx3_batch(trh_handle *h, const trh_mog_approx_job *jobs,
                                            int njobs, int max_wh, const double *amp,
                                            const double *mean, const double *var, double *out,
                                            double *dx_out, double *dy_out, int32_t *bbox,
                                            void *stream) {
    if (!h || (njobs > 0 && (!jobs || !amp || !mean || !var || !out || !bbox)))
        return set_err(TRH_E_ARG, "trh_render_mog_approx3_batch: null argument");
    if (njobs <= 0) return TRH_OK;
    if (max_wh <= 0 || max_wh > TRH_APPROX_MAX_WH)
        return set_err(TRH_E_ARG, "trh_render_mog_approx3_batch: max_wh=%d outside [1, %d]", max_wh,
                       TRH_APPROX_MAX_WH);
    const size_t lds = sizeof(Comp) * TRH_APPROX_MAXK + (sizeof(double) + sizeof(int)) * kThreads +
                       16 * sizeof(int) + 4 * (size_t)max_wh;
    if (lds > (size_t)h->lds_max)
        return set_err(TRH_E_UNSUPPORTED, "trh_render_mog_approx3_batch: LDS %zu > %d", lds,
                       h->lds_max);
    int grid = njobs < 4 * h->n_cu ? njobs : 4 * h->n_cu;
    hipLaunchKernelGGL(k_mog_approx3, dim3(grid), dim3(kThreads), lds, (hipStream_t)stream, jobs,
                       njobs, max_wh, amp, mean, var, out, dx_out, dy_out, bbox);
    return hip_check(hipGetLastError(), "k_mog_approx3");
}
