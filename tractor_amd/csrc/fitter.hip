// Batched single-source fitter: one LM iteration over thousands of sources,
// each on its own P x P modelMask patch, entirely on the device.
//
// Per iteration (all launches on one stream, no host synchronisation):
//   prep_jac  -> build 1 (point) or 4 (galaxy: current + 3 shape steps)
//                render jobs per active source, bucketed by FFT size
//   render    -> k_render_fft<sz> / k_render_point write the unit-flux renders
//   normal_eq -> J^T W J, J^T W r, chi^2 per source (FP32 products, FP64 sums)
//   solve     -> column-scaled, damped FP64 Cholesky; the reference's line
//                search steps (ConstrainedOptimizer.getParameterSteps)
//   prep_ls   -> one chi^2 render job per (source, step)
//   render    -> fused render + chi^2 (no model written to HBM)
//   accept    -> tryUpdates acceptance rule + optimize_loop stopping rule
#include <hip/hip_runtime.h>
#include <math.h>

#include <algorithm>
#include <vector>

#include "../../include/tractor_hip.h"
#include "internal.h"
#include "render_fft.h"

using namespace trh;

#define NALPHA 13
#define NFFT 5     // FFT sizes 32, 64, 128, 256, 512
#define PTB NFFT   // bucket index of point sources
#define NBUCKET (NFFT + 1)
#define PREP_TB 128   // threads per block of the job-building kernels
#define LS_CNT 8   // line-search counters start at cnt[8]
// launch slots: 0-4 jacobian FFT renders (sz 32..512), 5 jacobian point
// renders, 6 normal equations, 7 solve, 8-12 line-search FFT renders,
// 13 line-search point renders, 14 job preparation
#define SL_JPT 5
#define SL_NEQ 6
#define SL_SOLVE 7
#define SL_LS 8
#define SL_LPT 13
#define SL_PREP 14
#define TRH_NSLOT 15
#define GRAPH_ITERS 4   // iterations per captured graph (BatchFitter.fit checks every 4)

// Normal equations of one source: neq[NEQ_W s + .] holds the packed upper
// triangle of the NC x NC matrix J^T W J (NC = 6; 7 for a Sersic galaxy),
// then the NC entries of J^T W r, chi^2 at NEQ_CHI2 and the zero-model chi^2
// of the window (point sources) at NEQ_CHI0.
#define NEQ_W 40
#define NEQ_CHI2 36
#define NEQ_CHI0 37

struct FitProfiles {
    double amp[2][TRH_MAXK];
    double var[2][TRH_MAXK];
    int K[2];
};

// Sersic index bounds and steps (sersic.py:344-359)
#define SER_LOWER 0.29
#define SER_UPPER 6.3
#define SER_STEP 0.01
#define SER_MAXSTEP 0.1
#define SER_MAXR 8
#define SER_MAXD 8192

// The Sersic mixture as piecewise cubics in the index (SersicMixture,
// sersic.py:246-342): per index range [lo, hi), K components whose amplitude
// and log-variance are cubic on each of nint pieces (breakpoints at
// d[boff ..], coefficients of component k, f = 0 amplitude / 1 log-variance,
// piece i at d[coff + ((2 k + f) nint + i) 4 ..], highest power first in
// powers of (n - breakpoint)); the core fraction likewise.
struct SersicTab {
    int nr;
    int K[SER_MAXR], nint[SER_MAXR], boff[SER_MAXR], coff[SER_MAXR];
    int core_nint, core_boff, core_coff;
    double lo[SER_MAXR], hi[SER_MAXR];
    double d[SER_MAXD];
};

// One piecewise cubic at x: the piece whose breakpoint is the last one <= x
// (the first piece below the first breakpoint, the last above the last).
__device__ __forceinline__ double ser_piece(const double *br, const double *c, int nint,
                                            double x) {
    int i = 0;
    while (i + 1 < nint && x >= br[i + 1]) ++i;
    const double t = x - br[i];
    const double *q = c + 4 * i;
    return ((q[0] * t + q[1]) * t + q[2]) * t + q[3];
}

// Every component of range r at n: the components share the range's
// breakpoints, so the piece is found once.
__device__ int ser_range(const SersicTab &T, int r, double n, double *amp, double *lvar) {
    const int K = T.K[r], m = T.nint[r];
    const double *br = T.d + T.boff[r];
    int i = 0;
    while (i + 1 < m && n >= br[i + 1]) ++i;
    const double t = n - br[i];
    const double *c = T.d + T.coff[r] + 4 * i;
    for (int k = 0; k < K; ++k) {
        const double *qa = c + (size_t)(2 * k) * m * 4, *ql = qa + (size_t)m * 4;
        amp[k] = ((qa[0] * t + qa[1]) * t + qa[2]) * t + qa[3];
        lvar[k] = ((ql[0] * t + ql[1]) * t + ql[2]) * t + ql[3];
    }
    return K;
}

// The range(s) holding index n (r1 >= 0 where two overlap), with n clamped
// to the table.
__device__ __forceinline__ void ser_ranges(const SersicTab &T, double &n, int &r0, int &r1) {
    r0 = -1;
    r1 = -1;
    if (!(n > T.lo[0])) {   // also a NaN index: the lowest range
        r0 = 0;
        n = T.lo[0];
    } else if (n >= T.hi[T.nr - 1]) {
        r0 = T.nr - 1;
        n = T.hi[T.nr - 1];
    } else {
        for (int r = 0; r < T.nr; ++r)
            if (T.lo[r] <= n && n < T.hi[r]) {
                if (r0 < 0)
                    r0 = r;
                else if (r1 < 0)
                    r1 = r;
            }
        // trh_fitter_set_sersic admits no gaps between ranges; keep every
        // index inside the table regardless
        if (r0 < 0)
            for (int r = 0; r < T.nr; ++r)
                if (T.lo[r] <= n) r0 = r;
    }
}

// Number of components of the mixture at n (no spline evaluated).
__device__ int sersic_K(const SersicTab &T, double n) {
    int r0, r1;
    ser_ranges(T, n, r0, r1);
    return T.K[r0] + (r1 >= 0 ? T.K[r1] : 0) + (n > 1.0 ? 1 : 0);
}

// SersicMixture._getProfile (sersic.py:286-342): amplitudes normalised per
// range (a linear ramp of weights where two ranges overlap), variances
// exp(log-variance), a zero-variance core above n = 1, sorted by variance
// (stable).  amp, var hold TRH_MAXK + 1 entries.  Returns K.
__device__ int sersic_profile(const SersicTab &T, double n, double *amp, double *var) {
    int r0, r1;
    ser_ranges(T, n, r0, r1);
    int K;
    if (r1 >= 0) {
        const double frac = (n - T.lo[r1]) / (T.hi[r0] - T.lo[r1]);
        const int K0 = ser_range(T, r0, n, amp, var);
        const int K1 = ser_range(T, r1, n, amp + K0, var + K0);
        double s0 = 0.0, s1 = 0.0;
        for (int k = 0; k < K0; ++k) s0 += amp[k];
        for (int k = 0; k < K1; ++k) s1 += amp[K0 + k];
        for (int k = 0; k < K0; ++k) amp[k] = (1.0 - frac) * (amp[k] / s0);
        for (int k = 0; k < K1; ++k) amp[K0 + k] = frac * (amp[K0 + k] / s1);
        K = K0 + K1;
    } else {
        K = ser_range(T, r0, n, amp, var);
        double s = 0.0;
        for (int k = 0; k < K; ++k) s += amp[k];
        for (int k = 0; k < K; ++k) amp[k] /= s;
    }
    for (int k = 0; k < K; ++k) var[k] = exp(var[k]);
    if (n > 1.0) {
        const double core = ser_piece(T.d + T.core_boff, T.d + T.core_coff, T.core_nint, n);
        double s = 0.0;
        for (int k = 0; k < K; ++k) s += amp[k];
        const double f = (1.0 - core) / s;
        for (int k = 0; k < K; ++k) amp[k] *= f;
        amp[K] = core;
        var[K] = 0.0;
        ++K;
    }
    for (int i = 1; i < K; ++i) {
        const double a = amp[i], v = var[i];
        int j = i - 1;
        while (j >= 0 && var[j] > v) {
            amp[j + 1] = amp[j];
            var[j + 1] = var[j];
            --j;
        }
        amp[j + 1] = a;
        var[j + 1] = v;
    }
    return K;
}

// Finite-difference step of the Sersic index for its Jacobian column, as
// SersicGalaxy.getDerivativeShearedProfiles (sersic.py:447-487): +0.01, or
// -0.01 where that would pass the upper bound; if the stepped profile has a
// different number of components, the opposite step when it stays inside
// the bounds and keeps the count.
__device__ double sersic_fd_step(const SersicTab &T, double n0) {
    double step = SER_STEP;
    if (n0 + step > SER_UPPER) step = -step;
    const int K0 = sersic_K(T, n0);
    if (sersic_K(T, n0 + step) != K0) {
        const double n2 = n0 - step;
        if (n2 < SER_UPPER && n2 > SER_LOWER && sersic_K(T, n2) == K0) step = -step;
    }
    return step;
}

// ---- Hybrid PSF (HybridPixelizedPSF): galaxy components split between the
// FFT render and a pixel-space mixture convolved with the PSF's own mixture
// (galaxy.py:412-455, 510-516; the same ramp weights as gpu_optimizer.py:
// 961-1011).  The pixel-space part is evaluated on the source's window at
// the galaxy's exact position (c_gauss_2d_grid, mix.i:452-524) and added to
// the FFT render's unit-flux patch.
#define FIT_PSF_MAXK 4
#define FIT_MOG_MAXC (TRH_MAXK * FIT_PSF_MAXK)

struct FitHybrid {
    int K;
    double amp[FIT_PSF_MAXK], mx[FIT_PSF_MAXK], my[FIT_PSF_MAXK];
    double v0[FIT_PSF_MAXK], v1[FIT_PSF_MAXK], v2[FIT_PSF_MAXK];
    double n1, n2;   // inner / outer real-space nsigma (3, 4)
};

// One pixel-space render: components k < K centred at (px, py) + m[k], with
// c_gauss_2d_grid's constants q[k] = (amp / (2 pi sqrt det), -V2 / (2 det),
// V1 / det, -V0 / (2 det)) (the exponent of pixel offset (dx, dy) is
// q.y dx^2 + q.z dx dy + q.w dy^2, and terms with exponent <= -50 are
// skipped, as the reference skips Mahalanobis distances >= 100).
struct FitMogJob {
    int64_t off;       // window offset in the output plane
    int32_t stride, P, K;
    float scale;       // counts for a model render, 1 for a unit-flux one
    double px, py;
    float4 q[FIT_MOG_MAXC];
    float2 m[FIT_MOG_MAXC];
};

struct trh_fitter {
    trh_handle *h = nullptr;
    int S = 0, n_gal = 0, n_pt = 0, ph = 0, pw = 0;
    double cdi[4] = {1, 0, 0, 1};
    int maxrows[NFFT] = {8, 8, 8, 8, 8};
    // smallest FFT size any galaxy render can use: halfsize >= max(pw, ph) / 2
    // (galaxy.py:337-351), so the buckets below it never hold a job
    int min_sz = 32;
    int cx[NFFT] = {0, 0, 0, 0, 0}, cy[NFFT] = {0, 0, 0, 0, 0};
    // device
    int32_t *kind = nullptr, *P = nullptr;
    int32_t *stride = nullptr;   // row stride of each source's data / inverr window
    int64_t *pix_off = nullptr, *render_off = nullptr;
    const float *data = nullptr, *inverr = nullptr;
    float *stamp = nullptr;
    float *spec[NFFT] = {nullptr, nullptr, nullptr, nullptr, nullptr};
    float *renders = nullptr;
    // Sersic galaxies (kind 3): their mixture table (set by
    // trh_fitter_set_sersic) and each one's index step of the last Jacobian
    int n_ser = 0;
    int nj = 4;   // Jacobian render slots per source: 5 when there are Sersic galaxies
    SersicTab *ser = nullptr;
    double *sstep = nullptr;
    double *params = nullptr, *trial = nullptr, *X = nullptr, *neq = nullptr, *chi2ls = nullptr;
    int32_t *nsteps = nullptr, *steplim = nullptr, *active = nullptr, *iters = nullptr;
    int32_t *flags = nullptr;    // TRH_FIT_F_* per source
    double *chisq = nullptr, *dlnp = nullptr;
    FitProfiles *prof = nullptr;
    trh_fft_job *jobs_jac = nullptr, *jobs_ls = nullptr;
    int32_t *list_jac = nullptr, *list_ls = nullptr, *cnt = nullptr;  // cnt[0..4] jac, [8..12] ls
    int32_t *neq_order = nullptr;   // sources by decreasing P (normal equations)
    // render lists without atomics: the job-building kernels write each job's
    // bucket (bk_*) and per-block bucket counts (bcnt_*), k_fit_lists turns
    // them into lists and counts
    int8_t *bk_jac = nullptr, *bk_ls = nullptr;
    int32_t *bcnt_jac = nullptr, *bcnt_ls = nullptr;
    hipEvent_t ev[6];
    bool have_ev = false;
    // per-launch profiling (trh_fitter_profile): events around every launch
    // slot, device accumulators of jobs and algorithmic bytes per slot
    bool profiling = false;
    hipEvent_t pev[TRH_NSLOT][2];
    unsigned long long *acc = nullptr;  // [TRH_NSLOT][2]: jobs, bytes
    double tot_ms[TRH_NSLOT];
    double tot_launch[TRH_NSLOT];
    double tot_extra_bytes[TRH_NSLOT];   // PSF spectrum reads, counted per launch
    // LM iterations captured as hipGraphs, replayed on gstream: one or
    // GRAPH_ITERS iterations per graph (the gap between two graph launches
    // is ~7.5 us, between the kernels inside one none), each on one stream
    // or with render launches of each phase forked onto a side stream
    // (parallel graph branches, replayed once the last count of active
    // sources was below branch_below; trh_fitter_set_branching, 0: never).
    // Slot (branched ? 2 : 0) + (GRAPH_ITERS iterations ? 1 : 0), each with
    // the lambda, dchisq and max_iter it was captured with.
    hipStream_t gstream = nullptr;
    struct GraphSlot {
        hipGraphExec_t x = nullptr;
        double lambda = 0.0, dchisq = 0.0;
        int max_iter = 0;
    } gs[4];
    hipEvent_t gev[2];
    bool have_gev = false;
    int branch_below = 0;
    int last_active = -1;   // -1: not read since the last set_params
    hipStream_t bside = nullptr;
    hipEvent_t bfork[2], bjoin[2];
    bool have_br = false;
    // active-source count read without synchronising the fit
    // (trh_fitter_count_active / trh_fitter_count_read): a one-block count
    // kernel on the caller's stream, its result copied to pinned host memory
    // on a side stream while the next replays run
    int32_t *nact_dev = nullptr;
    int32_t *nact_pin = nullptr;
    hipStream_t cstream = nullptr;
    hipEvent_t cev[2];
    bool have_count = false, count_pending = false;
    // the fitter's own row slots of the SZ >= 256 renders (sizes 256, 512),
    // so that fitters on different streams never share them
    float2 *hbuf[2] = {nullptr, nullptr};
    // hybrid PSF (trh_fitter_set_hybrid): the PSF mixture, the pixel-space
    // jobs beside the Jacobian / model and line-search FFT jobs (same
    // slots), and the line search's pixel-space renders: galaxy s's NALPHA
    // P x P windows at ls_base + boff[s] + a P^2
    FitHybrid *hyb = nullptr;       // null: pure pixelized PSF
    FitHybrid *hyb_dev = nullptr;   // its device copy, kept while hybrid is off
    // chained restarts (trh_fitter_set_chain): pair i starts source
    // ch_second[i] once source ch_first[i] has stopped with chi^2 above
    // ch_ref[i] (1 + ch_tol); ch_state[i]: 0 pending, 1 started, 2 not needed
    int n_chain = 0;
    int32_t *ch_first = nullptr, *ch_second = nullptr, *ch_state = nullptr;
    double *ch_ref = nullptr;
    double ch_tol = 0.0;
    FitMogJob *mog_jac = nullptr, *mog_ls = nullptr;
    float *ls_base = nullptr;
    int64_t *boff = nullptr;
    std::vector<void *> allocs;
};

static const int kSizes[NFFT] = {32, 64, 128, 256, 512};

__device__ __forceinline__ int sz_bucket(int sz) {
    switch (sz) {
    case 32: return 0;
    case 64: return 1;
    case 128: return 2;
    case 256: return 3;
    case 512: return 4;
    }
    return -1;
}

// Build the FFT render job of a galaxy with profile (amp, var)[K] at (x, y)
// with EllipseESoft shape (logre, ee1, ee2) on the modelMask (0, 0, P, P)
// (galaxy.py:336-406, 659-669; ellipses.py:126-155, 232-255; NullWCS
// pixscale 1).  Returns the FFT size.
__device__ int build_gal_job(trh_fft_job &J, const double *amp, const double *var, int K,
                             double x, double y, double logre, double ee1, double ee2, int P,
                             int ph, int pw, double (*dvar)[3] = nullptr) {
    const double re = exp(logre);
    const double ee = hypot(ee1, ee2);
    const double e = 1.0 - exp(-ee);
    const double theta = atan2(ee2, ee1) / 2.0;
    const double ct = cos(theta), st = sin(theta);
    const double ab = e >= 1.0 ? 1000.0 : fmin(1000.0, (1.0 + e) / (1.0 - e));
    const double r_deg = re / 3600.0;
    const double cdinv = 1.0 / (1.0 / 3600.0);
    // T = cdinv * G, G = r_deg [[ct/ab, st], [-st/ab, ct]]
    const double T00 = cdinv * r_deg * ct / ab, T01 = cdinv * r_deg * st;
    const double T10 = -cdinv * r_deg * st / ab, T11 = cdinv * r_deg * ct;
    const double M00 = T00 * T00 + T01 * T01;
    const double M01 = T00 * T10 + T01 * T11;
    const double M11 = T10 * T10 + T11 * T11;
    J.K = K;
    for (int k = 0; k < K; ++k) {
        const double v = var[k];
        J.amp[k] = (float)amp[k];
        J.var[k][0] = (float)(v * M00);
        J.var[k][1] = (float)(v * M01);
        J.var[k][2] = (float)(v * M11);
        if (dvar) {
            dvar[k][0] = v * M00;
            dvar[k][1] = v * M01;
            dvar[k][2] = v * M11;
        }
    }
    // halfsize with a modelMask (galaxy.py:337-351)
    double hs = fmax(P / 2.0, P / 2.0);
    hs = fmax(hs, fmax(fmax(1.0 + x - 0.0, 1.0 + P - x), fmax(1.0 + y - 0.0, 1.0 + P - y)));
    hs = fmax(hs, fmax(pw / 2.0, ph / 2.0));
    const int ihs = (int)ceil(hs);
    int sz = 1;
    while (sz < 2 * ihs) sz <<= 1;
    const int cx = sz >= pw ? (sz - pw) / 2 + pw / 2 : pw / 2 - (pw - sz) / 2;
    const int cy = sz >= ph ? (sz - ph) / 2 + ph / 2 : ph / 2 - (ph - sz) / 2;
    double mux = x - cx;
    double muy = y - cy;
    const double sx = rint(mux), sy = rint(muy);
    J.sx = (int)sx;
    J.sy = (int)sy;
    J.mux = mux - sx;
    J.muy = muy - sy;
    J.mw = P;
    J.mh = P;
    J.out_stride = P;
    J.psf = 0;
    return sz;
}

// Point source with PixelizedPSF on the modelMask (0, 0, P, P)
// (psf.py:137-190): the stamp's (0,0) lands at (ix - pw//2, iy - ph//2).
__device__ void build_point_job(trh_fft_job &J, double x, double y, int P, int ph, int pw) {
    const double ix = rint(x), iy = rint(y);
    J.mux = x - ix;
    J.muy = y - iy;
    J.sx = (int)ix - pw / 2;
    J.sy = (int)iy - ph / 2;
    J.mw = P;
    J.mh = P;
    J.out_stride = P;
    J.psf = 0;
    J.K = 0;
}

// Split the sheared components (FP64 pixel variances dv) of a render with
// FFT box size sz: the FFT job keeps the components narrow enough for the
// box (amplitudes ramped where both apply) and the MoG job gets the wide
// ones convolved with the PSF mixture (PSF component outermost, as
// MixtureOfGaussians.convolve orders them, mixture_profiles.py:158-172).
// A galaxy centred outside its window (sourceOut) is all pixel space
// (galaxy.py:354-372, run_mog on the modelMask).  Returns false when the MoG
// part has more than FIT_MOG_MAXC components.
__device__ bool hybrid_split(const FitHybrid &H, int sz, int K, const double *amp,
                             const double (*dv)[3], trh_fft_job &J, FitMogJob &M, double px,
                             double py, int P) {
    const double hw = sz / 2.0;   // pW / 2 of the padded PSF
    const bool out = px < 0.0 || px > P - 1.0 || py < 0.0 || py > P - 1.0;
    bool im[TRH_MAXK], iff[TRH_MAXK], ramp = false;
    double ns[TRH_MAXK];
    for (int k = 0; k < K; ++k) {
        const double vv = dv[k][0] + dv[k][2];
        im[k] = out || hw * hw < H.n2 * H.n2 * vv;
        iff[k] = !out && hw * hw > H.n1 * H.n1 * vv;
        ramp = ramp || (im[k] && iff[k]);
        ns[k] = hw / fmax(1e-6, sqrt(vv));
    }
    int nf = 0;
    for (int k = 0; k < K; ++k) {
        if (!iff[k]) continue;
        double a = amp[k];
        if (ramp) a *= fmin(1.0, (ns[k] - H.n1) / (H.n2 - H.n1));
        J.amp[nf] = (float)a;
        J.var[nf][0] = (float)dv[k][0];
        J.var[nf][1] = (float)dv[k][1];
        J.var[nf][2] = (float)dv[k][2];
        ++nf;
    }
    J.K = nf;
    int nm = 0;
    for (int k = 0; k < K; ++k) nm += im[k] ? 1 : 0;
    M.px = px;
    M.py = py;
    M.K = 0;
    M.scale = 1.f;
    if (nm * H.K > FIT_MOG_MAXC) return false;
    int c = 0;
    for (int j = 0; j < H.K; ++j)
        for (int k = 0; k < K; ++k) {
            if (!im[k]) continue;
            double a = amp[k];
            if (ramp) a *= fmin(1.0, (H.n2 - ns[k]) / (H.n2 - H.n1));
            a *= H.amp[j];
            const double V0 = dv[k][0] + H.v0[j], V1 = dv[k][1] + H.v1[j], V2 = dv[k][2] + H.v2[j];
            const double det = V0 * V2 - V1 * V1;
            float4 q = make_float4((float)(a / sqrt(4.0 * M_PI * M_PI * det)),
                                   (float)(-0.5 * V2 / det), (float)(V1 / det),
                                   (float)(-0.5 * V0 / det));
            if (!(det > 0.0)) q = make_float4(0.f, -1.f, 0.f, -1.f);
            M.q[c] = q;
            M.m[c] = make_float2((float)H.mx[j], (float)H.my[j]);
            ++c;
        }
    M.K = c;
    return true;
}

// Pixel-space parts of the galaxy jobs listed in the galaxy buckets
// lists[b * cap + i], i < cnt[b], b < NFFT (their FitMogJob at the same
// slot), into out: mode 0 stores, 1 adds (the FFT render of the same pixels
// was stored first, in stream order), 2 adds atomically.  Plane:
// out + (slot % nplanes) * plane (nplanes 1: one plane).  One block per job.
__global__ void __launch_bounds__(256)
k_fit_mog(const FitMogJob *__restrict__ jobs, const int32_t *__restrict__ lists, int cap,
          const int32_t *__restrict__ cnt, int mode, float *__restrict__ out, int nplanes,
          int64_t plane) {
    __shared__ int pre[NFFT + 1];
    __shared__ float4 Q[FIT_MOG_MAXC];
    __shared__ int2 CI[FIT_MOG_MAXC];
    __shared__ float2 CF[FIT_MOG_MAXC];
    if (threadIdx.x == 0) {
        pre[0] = 0;
        for (int b = 0; b < NFFT; ++b) pre[b + 1] = pre[b] + cnt[b];
    }
    __syncthreads();
    const int n = pre[NFFT];
    for (int jj = blockIdx.x; jj < n; jj += gridDim.x) {
        int b = 0;
        while (jj >= pre[b + 1]) ++b;
        const int slot = lists[(size_t)b * cap + (jj - pre[b])];
        const FitMogJob &J = jobs[slot];
        const int K = J.K, P = J.P;
        if (K <= 0) continue;
        __syncthreads();
        if ((int)threadIdx.x < K) {
            const double mx = J.px + J.m[threadIdx.x].x, my = J.py + J.m[threadIdx.x].y;
            const double fx = floor(mx), fy = floor(my);
            Q[threadIdx.x] = J.q[threadIdx.x];
            CI[threadIdx.x] = make_int2((int)fx, (int)fy);
            CF[threadIdx.x] = make_float2((float)(mx - fx), (float)(my - fy));
        }
        __syncthreads();
        float *o = out + (int64_t)(slot % nplanes) * plane + J.off;
        for (int pix = threadIdx.x; pix < P * P; pix += blockDim.x) {
            const int oy = pix / P, ox = pix - oy * P;
            float r = 0.f;
            for (int k = 0; k < K; ++k) {
                const float4 q = Q[k];
                if (q.x == 0.f) continue;
                const float dx = (float)(ox - CI[k].x) - CF[k].x;
                const float dy = (float)(oy - CI[k].y) - CF[k].y;
                const float e = q.y * dx * dx + q.z * dx * dy + q.w * dy * dy;
                if (!(e > -50.f)) continue;
                r += q.x * __expf(e);
            }
            const int64_t idx = (int64_t)oy * J.stride + ox;
            r *= J.scale;
            if (mode == 0) o[idx] = r;
            else if (mode == 1) o[idx] += r;
            else atomicAdd(o + idx, r);
        }
    }
}

__global__ void k_fit_reset(int32_t *cnt, int n) {
    if ((int)threadIdx.x < n) cnt[threadIdx.x] = 0;
}

// Bucket of this thread's job (b < 0: none) into bk[slot], and the block's
// job count per bucket into bcnt[blockIdx.x][.] (no global atomics: on this
// GPU, device-scope atomics on one address from thousands of waves
// serialise).  Every thread of the block calls this.
__device__ __forceinline__ void block_bucket_counts(int b, int slot, int njobs,
                                                    int8_t *__restrict__ bk,
                                                    int32_t *__restrict__ bcnt) {
    __shared__ int32_t wc[PREP_TB / 64][NBUCKET];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (slot < njobs) bk[slot] = (int8_t)b;
#pragma unroll
    for (int bb = 0; bb < NBUCKET; ++bb) {
        const unsigned long long m = __ballot(b == bb);
        if (lane == 0) wc[w][bb] = __popcll(m);
    }
    __syncthreads();
    if ((int)threadIdx.x < NBUCKET) {
        int t = 0;
        for (int ww = 0; ww < (int)(blockDim.x >> 6); ++ww) t += wc[ww][threadIdx.x];
        bcnt[blockIdx.x * NBUCKET + threadIdx.x] = t;
    }
}

// Render lists from the per-job buckets: block k (the job-building kernel's
// block k) adds up the counts of the blocks before it, which gives its
// offset in every list, then writes its jobs' slots in order; the last block
// also writes the list lengths.  Launched with the job-building kernel's grid.
__global__ void __launch_bounds__(PREP_TB)
k_fit_lists(int njobs, const int8_t *__restrict__ bk, const int32_t *__restrict__ bcnt,
            int32_t *__restrict__ lists, int cap, int32_t *__restrict__ cnt) {
    __shared__ int32_t part[PREP_TB / 64][NBUCKET];
    __shared__ int32_t base[NBUCKET];
    const int blk = blockIdx.x, lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    int sum[NBUCKET];
#pragma unroll
    for (int bb = 0; bb < NBUCKET; ++bb) sum[bb] = 0;
    for (int i = threadIdx.x; i < blk; i += blockDim.x)
#pragma unroll
        for (int bb = 0; bb < NBUCKET; ++bb) sum[bb] += bcnt[i * NBUCKET + bb];
#pragma unroll
    for (int bb = 0; bb < NBUCKET; ++bb) {
        int v = sum[bb];
#pragma unroll
        for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
        if (lane == 0) part[w][bb] = v;
    }
    __syncthreads();
    if ((int)threadIdx.x < NBUCKET) {
        int t = 0;
        for (int ww = 0; ww < (int)(blockDim.x >> 6); ++ww) t += part[ww][threadIdx.x];
        base[threadIdx.x] = t;
        if (blk == (int)gridDim.x - 1) cnt[threadIdx.x] = t + bcnt[blk * NBUCKET + threadIdx.x];
    }
    __syncthreads();
    const int slot = blk * blockDim.x + threadIdx.x;
    const int b = slot < njobs ? (int)bk[slot] : -1;
    int rank = 0;
#pragma unroll
    for (int bb = 0; bb < NBUCKET; ++bb) {
        const unsigned long long m = __ballot(b == bb);
        if (lane == 0) part[w][bb] = __popcll(m);
        if (b == bb) rank = __popcll(m & ((1ull << lane) - 1ull));
    }
    __syncthreads();
    if (b >= 0) {
        int off = base[b] + rank;
        for (int ww = 0; ww < w; ++ww) off += part[ww][b];
        lists[b * cap + off] = slot;
    }
}

// The job-building kernels run PREP_TB threads per block, one job record per
// thread: the records are assembled in LDS and copied out with coalesced
// 8-byte stores (a record is 328 bytes; written field by field from each
// thread, every store instruction would touch 64 cache lines).

// Copy the staged records whose flag is set (jobs that were built) to
// dst[0 .. count): consecutive threads write consecutive 8-byte words.
__device__ __forceinline__ void copy_jobs_out(const trh_fft_job *sj, const int8_t *built,
                                              trh_fft_job *dst, int count) {
    static_assert(sizeof(trh_fft_job) % 8 == 0, "8-byte words");
    constexpr int W = sizeof(trh_fft_job) / 8;
    const uint2 *s = reinterpret_cast<const uint2 *>(sj);
    uint2 *d = reinterpret_cast<uint2 *>(dst);
    for (int i = threadIdx.x; i < count * W; i += blockDim.x)
        if (built[i / W]) d[i] = s[i];
}

// The profile of a galaxy of kind kd (1 exp, 2 dev: the fixed tables; 3:
// the Sersic mixture at index n, evaluated into sa, sv).  Returns K.
__device__ __forceinline__ int gal_profile(const FitProfiles &pr, const SersicTab *ser, int kd,
                                           double n, const double *&amp, const double *&var,
                                           double *sa, double *sv) {
    if (kd == 3) {
        amp = sa;
        var = sv;
        return sersic_profile(*ser, n, sa, sv);
    }
    amp = pr.amp[kd - 1];
    var = pr.var[kd - 1];
    return pr.K[kd - 1];
}

// Jacobian render planes of a source: the unit render, the three EllipseE
// steps of a galaxy and the index step of a Sersic galaxy.
__device__ __forceinline__ int jac_planes(int kd) { return kd == 0 ? 1 : (kd == 3 ? 5 : 4); }

// Per-launch profiling sums (jobs and algorithmic bytes per launch slot) of
// one block of a job-building kernel: added up in LDS, then one global add
// per counter the block touched.  One global atomic per thread would hit
// the same 30 addresses from every wave and serialise (3 ms for a 213k-
// thread line-search launch against 35 us without profiling).
struct BlockAcc {
    unsigned long long *g;
    unsigned long long *l;
    // every thread of the block calls this (a barrier when g is set)
    __device__ __forceinline__ void init(unsigned long long *glob, unsigned long long *lds) {
        g = glob;
        l = lds;
        if (!g) return;
        for (int i = threadIdx.x; i < 2 * TRH_NSLOT; i += blockDim.x) l[i] = 0;
        __syncthreads();
    }
    __device__ __forceinline__ void add(int sl, unsigned long long bytes) const {
        atomicAdd(&l[2 * sl], 1ull);
        atomicAdd(&l[2 * sl + 1], bytes);
    }
    // after a barrier that follows the last add()
    __device__ __forceinline__ void flush() const {
        if (!g) return;
        for (int i = threadIdx.x; i < 2 * TRH_NSLOT; i += blockDim.x)
            if (l[i]) atomicAdd(&g[i], l[i]);
    }
};

// One thread per (source, render), nj slots per source: render 0 of every
// active source, renders 1..3 (EllipseE steps) of the galaxies and render 4
// (index step, whose size goes to sstep[s]) of the Sersic galaxies.
__global__ void __launch_bounds__(PREP_TB) k_fit_prep_jac(int S, int nj, const int32_t *__restrict__ kind,
                               const int32_t *__restrict__ Parr, const int64_t *__restrict__ roff,
                               const double *__restrict__ params,
                               const int32_t *active, const FitProfiles *__restrict__ pr,
                               const SersicTab *__restrict__ ser, double *__restrict__ sstep,
                               int ph, int pw, trh_fft_job *__restrict__ jobs,
                               int8_t *__restrict__ bk, int32_t *__restrict__ bcnt,
                               int32_t *__restrict__ flags, int32_t *__restrict__ act_out,
                               unsigned long long *__restrict__ acc,
                               const FitHybrid *__restrict__ hyb, FitMogJob *__restrict__ mog) {
    __shared__ trh_fft_job sj[PREP_TB];
    __shared__ unsigned long long lacc[2 * TRH_NSLOT];
    BlockAcc A;
    A.init(acc, lacc);
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int s = t / nj, k = t - s * nj;
    int b = -1;
    if (s < S && active[s] && k < jac_planes(kind[s])) {
        const int kd = kind[s];
        const int P = Parr[s];
        const double *p = params + 8 * s;
        const unsigned long long PP = (unsigned long long)P * P;
        if (acc && k == 0) {
            // normal equations read the renders + data + inverr and write
            // the packed sums; the solve reads them and writes X and trials
            const unsigned long long nc = kd == 3 ? 7 : 6, nq = nc * (nc + 1) / 2 + nc + 2;
            A.add(SL_NEQ, (4ull * jac_planes(kd) + 8) * PP + nq * 8);
            A.add(SL_SOLVE, nq * 8 + 8ull * 8 * NALPHA + 8 * 8);
        }
        trh_fft_job &J = sj[threadIdx.x];
        if (kd == 0) {
            build_point_job(J, p[0], p[1], P, ph, pw);
            J.out_offset = roff[s];
            b = PTB;
        } else {
            double sh[3] = {p[3], p[4], p[5]};
            if (k > 0 && k < 4) sh[k - 1] += 0.01;  // EllipseE stepsizes (ellipses.py:44)
            double n = p[6];
            if (k == 4) {
                const double st = sersic_fd_step(*ser, n);
                sstep[s] = st;
                n += st;
            }
            double sa[TRH_MAXK + 1], sv[TRH_MAXK + 1], dv[TRH_MAXK][3];
            const double *amp, *var;
            const int K = gal_profile(*pr, ser, kd, n, amp, var, sa, sv);
            const int sz = build_gal_job(J, amp, var, K, p[0], p[1], sh[0], sh[1], sh[2], P, ph,
                                         pw, hyb ? dv : nullptr);
            J.out_offset = roff[s] + (int64_t)k * P * P;
            b = sz_bucket(sz);
            if (b >= 0 && hyb) {
                FitMogJob &M = mog[t];
                M.off = J.out_offset;
                M.stride = P;
                M.P = P;
                if (!hybrid_split(*hyb, sz, K, amp, dv, J, M, p[0], p[1], P)) {
                    atomicOr(&flags[s], TRH_FIT_F_HYBRID);
                    b = -1;
                }
            }
            if (b < 0) {
                // no Jacobian for this source: it stops at its current
                // parameters, flagged (the other sources go on)
                if (sz_bucket(sz) < 0) atomicOr(&flags[s], TRH_FIT_F_JAC_SIZE);
                act_out[s] = 0;
            }
        }
        J.mode = TRH_STORE;
        J.scale = 1.f;
        J.slot = 0;
        if (acc && b >= 0) {
            A.add(b == PTB ? SL_JPT : b, 4ull * PP);  // render written
        }
    }
    __shared__ int8_t built[PREP_TB];
    built[threadIdx.x] = b >= 0;
    block_bucket_counts(b, t, nj * S, bk, bcnt);
    __syncthreads();
    A.flush();
    const int first = blockIdx.x * blockDim.x;
    copy_jobs_out(sj, built, jobs + first, min((int)blockDim.x, nj * S - first));
}

// Model renders (counts x unit-flux patch at the current parameters) into an
// image buffer laid out like `data` (patch s at pix_off[s]).
__global__ void k_fit_prep_model(int S, int nj, const int32_t *__restrict__ kind,
                                 const int32_t *__restrict__ Parr,
                                 const int32_t *__restrict__ stride, int mode,
                                 const int64_t *__restrict__ pix_off,
                                 const double *__restrict__ params,
                                 const FitProfiles *__restrict__ pr,
                                 const SersicTab *__restrict__ ser, int ph, int pw,
                                 trh_fft_job *__restrict__ jobs, int32_t *__restrict__ lists,
                                 int cap, int32_t *__restrict__ cnt,
                                 int32_t *__restrict__ flags, const FitHybrid *__restrict__ hyb,
                                 FitMogJob *__restrict__ mog) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= S) return;
    const int kd = kind[s];
    const int P = Parr[s];
    const double *p = params + 8 * s;
    trh_fft_job &J = jobs[(size_t)nj * s];
    int b;
    if (kd == 0) {
        build_point_job(J, p[0], p[1], P, ph, pw);
        b = PTB;
    } else {
        double sa[TRH_MAXK + 1], sv[TRH_MAXK + 1], dv[TRH_MAXK][3];
        const double *amp, *var;
        const int K = gal_profile(*pr, ser, kd, p[6], amp, var, sa, sv);
        const int sz = build_gal_job(J, amp, var, K, p[0], p[1], p[3], p[4], p[5], P, ph, pw,
                                     hyb ? dv : nullptr);
        b = sz_bucket(sz);
        if (b < 0) {
            atomicOr(&flags[s], TRH_FIT_F_MODEL_SIZE);
            return;
        }
        if (hyb) {
            FitMogJob &M = mog[(size_t)nj * s];
            M.off = pix_off[s];
            M.stride = stride[s];
            M.P = P;
            if (!hybrid_split(*hyb, sz, K, amp, dv, J, M, p[0], p[1], P)) {
                atomicOr(&flags[s], TRH_FIT_F_HYBRID);
                return;
            }
            M.scale = (float)p[2];
        }
    }
    J.out_offset = pix_off[s];
    J.out_stride = stride[s];
    J.mode = mode;
    J.scale = (float)p[2];
    J.slot = 0;
    const int i = atomicAdd(&cnt[b], 1);
    lists[b * cap + i] = nj * s;
}

// Sums over the 64 lanes of 32 per-lane values, scattered: afterwards lane l
// holds the sum of value q(l) = 16 b5 + 8 b4 + 4 b3 + 2 b2 + b1 (b_k = bit k of
// l; lanes l and l ^ 1 hold the same sum).  Each step keeps the half of the
// values selected by one lane bit and adds the partner's copy of it: 31
// exchanges instead of 32 x 6 for 32 separate all-sums.
__device__ __forceinline__ float wave_reduce_scatter32(const float (&v)[32]) {
    const int lane = threadIdx.x & 63;
    float a[16], b[8], c[4], d[2];
    bool up = (lane & 32) != 0;
#pragma unroll
    for (int i = 0; i < 16; ++i)
        a[i] = (up ? v[16 + i] : v[i]) + xlane<32>(up ? v[i] : v[16 + i]);
    up = (lane & 16) != 0;
#pragma unroll
    for (int i = 0; i < 8; ++i) b[i] = (up ? a[8 + i] : a[i]) + xlane<16>(up ? a[i] : a[8 + i]);
    up = (lane & 8) != 0;
#pragma unroll
    for (int i = 0; i < 4; ++i) c[i] = (up ? b[4 + i] : b[i]) + xlane<8>(up ? b[i] : b[4 + i]);
    up = (lane & 4) != 0;
#pragma unroll
    for (int i = 0; i < 2; ++i) d[i] = (up ? c[2 + i] : c[i]) + xlane<4>(up ? c[i] : c[2 + i]);
    up = (lane & 2) != 0;
    float e = (up ? d[1] : d[0]) + xlane<2>(up ? d[0] : d[1]);
    return e + xlane<1>(e);
}

__device__ __forceinline__ double wave_allsum_d(double v) {
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// One pixel's contribution to the NQ = NC (NC + 1) / 2 + NC FP32 sums (the
// packed J^T W J, then J^T W r; NC = 6, or 7 with a Sersic index column) and
// to chi^2 (FP64).  Columns (galaxy.py:137-223, pointsource.py:130-241,
// sersic.py:398-436, lsqr_optimizer.py:386-431):
//   pos_i = ((dxp*cdi[0][i] + dyp*cdi[1][i]) * counts) * inverr  (central
//           differences of the unit render, zero 1-px border)
//   flux  = render * inverr
//   shape = ((R_k - R_0) * (counts / 0.01)) * inverr
//   index = ((R_4 - R_0) * (counts / step)) * inverr  (step: sstep[s])
//   b     = (data - counts * R_0) * inverr
// One pixel's inputs to the normal equations: the model render, its four
// neighbours (central differences for the position columns), the NC - 3
// derivative renders of a galaxy, inverr and data.
template <int NC>
struct NeqIn {
    float r0, dx, dy, g[NC - 3], e, d;
};

template <int NC>
struct NeqPix {
    static constexpr int NT = NC * (NC + 1) / 2, NQ = NT + NC;
    const float *R0, *D, *E;
    int P, PP, kd, ds;   // ds: row stride of the data / inverr window
    float invP, counts, cstep, istep, c00, c01, c10, c11;
    __device__ __forceinline__ void load(int idx, NeqIn<NC> &in) const {
        // exact for PP < 2^24: (idx + 0.5) / P is within 0.5/P of y + 0.5
        const int y = (int)(((float)idx + 0.5f) * invP);
        const int x = idx - y * P;
        const bool inner = x > 0 && x < P - 1 && y > 0 && y < P - 1;
        in.r0 = R0[idx];
        const float l = inner ? R0[idx - 1] : 0.f, r = inner ? R0[idx + 1] : 0.f;
        const float u = inner ? R0[idx - P] : 0.f, w = inner ? R0[idx + P] : 0.f;
        in.dx = (l - r) / 2.f;
        in.dy = (u - w) / 2.f;
#pragma unroll
        for (int k = 0; k < NC - 3; ++k) in.g[k] = kd ? R0[(k + 1) * PP + idx] : 0.f;
        const int di = y * ds + x;
        in.e = E[di];
        in.d = D[di];
    }
    // Four pixels of one row (x .. x + 3, x a multiple of 4) from one 16-byte
    // load per plane and two scalar loads for the outer neighbours: the same
    // inputs as four load() calls, with a quarter of the load instructions.
    // Needs P, the row stride and the plane and window starts to be
    // multiples of 4 floats (vec4_ok).
    template <int I>
    __device__ static __forceinline__ float lane4(const float4 &v) {
        if constexpr (I == 0) return v.x;
        else if constexpr (I == 1) return v.y;
        else if constexpr (I == 2) return v.z;
        else return v.w;
    }
    template <int I>
    __device__ __forceinline__ void pix4(NeqIn<NC> &in, bool ry, int x, float l, float r,
                                         const float4 &c, const float4 &up, const float4 &dn,
                                         const float4 &e4, const float4 &d4) const {
        const bool inner = ry && x + I > 0 && x + I < P - 1;
        in.r0 = lane4<I>(c);
        const float li = inner ? (I == 0 ? l : lane4<(I > 0 ? I - 1 : 0)>(c)) : 0.f;
        const float ri = inner ? (I == 3 ? r : lane4<(I < 3 ? I + 1 : 3)>(c)) : 0.f;
        const float ui = inner ? lane4<I>(up) : 0.f, wi = inner ? lane4<I>(dn) : 0.f;
        in.dx = (li - ri) / 2.f;
        in.dy = (ui - wi) / 2.f;
        in.e = lane4<I>(e4);
        in.d = lane4<I>(d4);
    }
    __device__ __forceinline__ void load4(int idx, NeqIn<NC> &i0, NeqIn<NC> &i1, NeqIn<NC> &i2,
                                          NeqIn<NC> &i3) const {
        const int y = (int)(((float)idx + 0.5f) * invP);
        const int x = idx - y * P;
        const bool ry = y > 0 && y < P - 1;
        const float4 c = *reinterpret_cast<const float4 *>(R0 + idx);
        const float4 zero = make_float4(0.f, 0.f, 0.f, 0.f);
        float4 up = zero, dn = zero;
        float l = 0.f, r = 0.f;
        if (ry) {
            up = *reinterpret_cast<const float4 *>(R0 + idx - P);
            dn = *reinterpret_cast<const float4 *>(R0 + idx + P);
            if (x > 0) l = R0[idx - 1];
            if (x + 4 < P) r = R0[idx + 4];
        }
        const int di = y * ds + x;
        const float4 e4 = *reinterpret_cast<const float4 *>(E + di);
        const float4 d4 = *reinterpret_cast<const float4 *>(D + di);
        // the derivative planes of a galaxy (a point source has none)
#pragma unroll
        for (int k = 0; k < NC - 3; ++k) {
            float4 gq = zero;
            if (kd) gq = *reinterpret_cast<const float4 *>(R0 + (k + 1) * PP + idx);
            i0.g[k] = gq.x;
            i1.g[k] = gq.y;
            i2.g[k] = gq.z;
            i3.g[k] = gq.w;
        }
        pix4<0>(i0, ry, x, l, r, c, up, dn, e4, d4);
        pix4<1>(i1, ry, x, l, r, c, up, dn, e4, d4);
        pix4<2>(i2, ry, x, l, r, c, up, dn, e4, d4);
        pix4<3>(i3, ry, x, l, r, c, up, dn, e4, d4);
    }
    __device__ __forceinline__ bool vec4_ok() const {
        return (P & 3) == 0 && (ds & 3) == 0 && ((uintptr_t)R0 & 15) == 0 &&
               ((uintptr_t)D & 15) == 0 && ((uintptr_t)E & 15) == 0;
    }
    __device__ __forceinline__ void add(const NeqIn<NC> &in, float (&acc)[NQ], double &chi2,
                                        double &chi0) const {
        const float e = in.e, r0 = in.r0;
        float col[NC];
        col[0] = ((in.dx * c00 + in.dy * c10) * counts) * e;
        col[1] = ((in.dx * c01 + in.dy * c11) * counts) * e;
        col[2] = r0 * e;
#pragma unroll
        for (int k = 0; k < NC - 3; ++k)
            col[3 + k] = kd ? ((in.g[k] - r0) * (k < 3 ? cstep : istep)) * e : 0.f;
        const float b = (in.d - r0 * counts) * e;
        int q = 0;
#pragma unroll
        for (int i = 0; i < NC; ++i)
#pragma unroll
            for (int j = i; j < NC; ++j) acc[q++] += col[i] * col[j];
#pragma unroll
        for (int i = 0; i < NC; ++i) acc[NT + i] += col[i] * b;
        chi2 += (double)b * (double)b;
        if (!kd) {
            const float c0 = in.d * e;   // the chi of a zero model
            chi0 += (double)c0 * (double)c0;
        }
    }
};

// J^T W J, J^T W r (FP32 products and sums) and chi^2 (FP64) for one source
// per NEQ_TB-thread block, grid-stride over the sources in order of decreasing
// P (sources span 64x in pixels; big ones first keeps the tail short).  Each
// thread takes runs of four consecutive pixels of a row and issues all their
// loads before any of their arithmetic, as one 16-byte load per plane when
// the window is aligned (a quarter of the load instructions of per-pixel
// loads: 9.4 -> 8.6 ms of a 16384-source fit); each wave reduce-scatters its first 32 sums
// (wave_reduce_scatter32), adds up the rest lane-wide, and the waves combine
// through LDS.  NC = 6 takes every source but the Sersic galaxies, NC = 7
// only those.
#define NEQ_TB 512   // threads per source in k_fit_normal_eq

template <int NC>
__global__ void __launch_bounds__(NEQ_TB)
k_fit_normal_eq(int S, const int32_t *__restrict__ kind, const int32_t *__restrict__ Parr,
                const int32_t *__restrict__ stride,
                const int64_t *__restrict__ pix_off, const int64_t *__restrict__ roff,
                const float *__restrict__ data, const float *__restrict__ inverr,
                const float *__restrict__ renders, const double *__restrict__ counts,
                int cstride, double shape_step, const double *__restrict__ sstep,
                const int32_t *__restrict__ active, double cdi00, double cdi01, double cdi10,
                double cdi11, double *__restrict__ neq, const int32_t *__restrict__ order) {
    constexpr int NQ = NeqPix<NC>::NQ;
    static_assert(NQ <= NEQ_CHI2, "packed sums overlap chi^2");
    __shared__ double red[NEQ_TB / 64][NEQ_W];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    for (int i = blockIdx.x; i < S; i += gridDim.x) {
        const int s = order ? order[i] : i;
        if ((active && !active[s]) || ((kind[s] == 3) != (NC == 7))) continue;
        NeqPix<NC> px;
        px.kd = kind[s];
        px.P = Parr[s];
        px.PP = px.P * px.P;
        px.ds = stride[s];
        px.invP = 1.0f / (float)px.P;
        const double cs = counts[(size_t)cstride * s];
        px.counts = (float)cs;
        px.cstep = (float)(cs / shape_step);
        px.istep = NC == 7 ? (float)(cs / sstep[s]) : 0.f;
        px.c00 = (float)cdi00;
        px.c01 = (float)cdi01;
        px.c10 = (float)cdi10;
        px.c11 = (float)cdi11;
        px.R0 = renders + roff[s];
        px.D = data + pix_off[s];
        px.E = inverr + pix_off[s];
        float acc[NQ];
#pragma unroll
        for (int k = 0; k < NQ; ++k) acc[k] = 0.f;
        double chi2 = 0.0, chi0 = 0.0;
        // each thread takes runs of four consecutive pixels (runs tid, tid +
        // NEQ_TB, ...) and adds them in pixel order: with 16-byte loads when
        // the window allows them, else four scalar loads, so that the sums
        // (and the fit) are the same bit for bit wherever the window lies
        const bool v4 = px.vec4_ok();
        const int G4 = (px.PP + 3) >> 2;
        for (int g = threadIdx.x; g < G4; g += NEQ_TB) {
            const int idx = 4 * g;
            NeqIn<NC> a0, a1, a2, a3;
            if (v4) {
                px.load4(idx, a0, a1, a2, a3);
            } else {
                px.load(idx, a0);
                px.load(min(idx + 1, px.PP - 1), a1);
                px.load(min(idx + 2, px.PP - 1), a2);
                px.load(min(idx + 3, px.PP - 1), a3);
            }
            px.add(a0, acc, chi2, chi0);
            if (idx + 1 < px.PP) px.add(a1, acc, chi2, chi0);
            if (idx + 2 < px.PP) px.add(a2, acc, chi2, chi0);
            if (idx + 3 < px.PP) px.add(a3, acc, chi2, chi0);
        }
        float v32[32];
#pragma unroll
        for (int k = 0; k < 32; ++k) v32[k] = k < NQ ? acc[k] : 0.f;
        const float sum = wave_reduce_scatter32(v32);
        const int q = ((lane >> 5) & 1) * 16 + ((lane >> 4) & 1) * 8 + ((lane >> 3) & 1) * 4 +
                      ((lane >> 2) & 1) * 2 + ((lane >> 1) & 1);
        chi2 = wave_allsum_d(chi2);
        chi0 = wave_allsum_d(chi0);
        if (!(lane & 1) && q < NQ) red[wid][q] = (double)sum;
#pragma unroll
        for (int k = 32; k < NQ; ++k) {
            float v = acc[k];
#pragma unroll
            for (int o = 1; o < 64; o <<= 1) v += __shfl_xor(v, o, 64);
            if (lane == 0) red[wid][k] = (double)v;
        }
        if (lane == 0) {
            red[wid][NEQ_CHI2] = chi2;
            red[wid][NEQ_CHI0] = chi0;
        }
        __syncthreads();
        if (threadIdx.x < NEQ_W) {
            const int k = threadIdx.x;
            double t = 0.0;
            if (k < NQ || k == NEQ_CHI2 || k == NEQ_CHI0)
                for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t += red[w][k];
            neq[(size_t)NEQ_W * s + k] = t;
        }
        __syncthreads();
    }
}

// Column-scaled damped Cholesky solve (FP64) + getParameterSteps
// (constrained_optimizer.py:84-159).  One thread per source.
// Column-scaled, damped FP64 Cholesky solve of one source's NP x NP normal
// equations (q: the packed upper triangle of the NC x NC matrix, then the NC
// gradient entries; NP <= NC), then the line-search steps
// (getParameterSteps).  NP and NC are compile-time constants so every array
// stays in registers.
template <int NP, int NC>
__device__ __forceinline__ void solve_source(int s, const double *__restrict__ q,
                                             const double *__restrict__ p0, double lambda,
                                             double *__restrict__ X, double *__restrict__ trial,
                                             int32_t *__restrict__ nsteps,
                                             int32_t *__restrict__ steplim) {
    double A[NP][NP], g[NP], D[NP];
#pragma unroll
    for (int i = 0; i < NP; ++i)
#pragma unroll
        for (int j = i; j < NP; ++j) {
            // upper-triangle index of (i, j) in the NC x NC packing
            const double v = q[i * NC - i * (i - 1) / 2 + (j - i)];
            A[i][j] = v;
            A[j][i] = v;
        }
#pragma unroll
    for (int i = 0; i < NP; ++i) g[i] = q[NC * (NC + 1) / 2 + i];
    bool any = false;
#pragma unroll
    for (int i = 0; i < NP; ++i) {
        D[i] = sqrt(A[i][i]);
        if (D[i] > 0.0) any = true;
    }
    double x[NP];
#pragma unroll
    for (int i = 0; i < NP; ++i) x[i] = 0.0;
    if (!any) {
        // all derivatives zero: no update (gpu_optimizer.py:762-767)
        if (nsteps) nsteps[s] = 0;
        if (steplim) steplim[s] = 0;
        for (int i = 0; i < 8; ++i) X[8 * s + i] = 0.0;
        return;
    }
    double lam = lambda;
    for (int attempt = 0; attempt < 12; ++attempt) {
        double L[NP][NP], rhs[NP];
        bool ok = true;
#pragma unroll
        for (int i = 0; i < NP; ++i) {
#pragma unroll
            for (int j = 0; j < NP; ++j) {
                if (D[i] > 0.0 && D[j] > 0.0)
                    L[i][j] = A[i][j] / (D[i] * D[j]);
                else
                    L[i][j] = (i == j) ? 1.0 : 0.0;
            }
            if (D[i] > 0.0) L[i][i] += lam;
            rhs[i] = D[i] > 0.0 ? g[i] / D[i] : 0.0;
        }
#pragma unroll
        for (int j = 0; j < NP; ++j) {
            double d = L[j][j];
#pragma unroll
            for (int k = 0; k < j; ++k) d -= L[j][k] * L[j][k];
            if (!(d > 1e-300)) ok = false;
            d = sqrt(fmax(d, 1e-300));
            L[j][j] = d;
#pragma unroll
            for (int i = j + 1; i < NP; ++i) {
                double v = L[i][j];
#pragma unroll
                for (int k = 0; k < j; ++k) v -= L[i][k] * L[j][k];
                L[i][j] = v / d;
            }
        }
        if (ok) {
            double y[NP];
#pragma unroll
            for (int i = 0; i < NP; ++i) {
                double v = rhs[i];
#pragma unroll
                for (int k = 0; k < i; ++k) v -= L[i][k] * y[k];
                y[i] = v / L[i][i];
            }
#pragma unroll
            for (int i = NP - 1; i >= 0; --i) {
                double v = y[i];
#pragma unroll
                for (int k = i + 1; k < NP; ++k) v -= L[k][i] * x[k];
                x[i] = v / L[i][i];
            }
#pragma unroll
            for (int i = 0; i < NP; ++i) x[i] = D[i] > 0.0 ? x[i] / D[i] : 0.0;
            break;
        }
        lam = lam > 0.0 ? lam * 10.0 : 1e-12;
    }
#pragma unroll
    for (int i = 0; i < 8; ++i) X[8 * s + i] = i < NP ? x[i < NP ? i : 0] : 0.0;
    if (!trial) return;   // the step alone (trh_solve_damped_cholesky_f64_batch)

    // getParameterSteps: maxstep PixPos [1,1] (wcs.py:384), Flux none,
    // EllipseESoft [1, 0.1, 0.1] (ellipses.py:48), SersicIndex 0.1; bounds
    // only on the Sersic index, [0.29, 6.3] (sersic.py:344-359).
    const double maxstep[7] = {1.0, 1.0, -1.0, 1.0, 0.1, 0.1, SER_MAXSTEP};
    double p[NP];
#pragma unroll
    for (int i = 0; i < NP; ++i) p[i] = p0[i];
    int n = 0;
    int anylim = 0;
    for (int a = 0; a < NALPHA; ++a) {
        double alpha = a <= 10 ? ldexp(1.0, a - 10) : (a == 11 ? sqrt(2.0) : 2.0);
        bool do_break = false;
#pragma unroll
        for (int i = 0; i < NP; ++i) {
            if (maxstep[i] < 0.0) continue;
            const double step = alpha * fabs(x[i]);
            if (step > maxstep[i]) {
                do_break = true;
                anylim = 1;
                alpha = maxstep[i] / fabs(x[i]);
            }
        }
        if (NP == 7) {
            // reduce alpha to land on a bound the step would pass
            double px = p[NP - 1] + alpha * x[NP - 1];
            if (px < SER_LOWER) {
                alpha = (SER_LOWER - p[NP - 1]) / x[NP - 1];
                px = SER_LOWER;
            }
            if (px > SER_UPPER) alpha = (SER_UPPER - p[NP - 1]) / x[NP - 1];
        }
        if (alpha < 1e-8) break;
        double *tp = trial + ((size_t)s * NALPHA + n) * 8;
#pragma unroll
        for (int i = 0; i < 8; ++i) tp[i] = i < NP ? p[i < NP ? i : 0] + alpha * x[i < NP ? i : 0] : 0.0;
        if (NP == 7) tp[NP - 1] = fmin(fmax(tp[NP - 1], SER_LOWER), SER_UPPER);
        ++n;
        if (do_break) break;
    }
    nsteps[s] = n;
    steplim[s] = anylim;
}

__global__ void __launch_bounds__(64) k_fit_solve(int S, const int32_t *__restrict__ kind,
                            const double *__restrict__ neq, const double *__restrict__ params,
                            const int32_t *__restrict__ active, double lambda,
                            double *__restrict__ X, double *__restrict__ trial,
                            int32_t *__restrict__ nsteps, int32_t *__restrict__ steplim) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= S) return;
    if (active && !active[s]) {
        if (nsteps) nsteps[s] = 0;
        return;
    }
    const double *q = neq + (size_t)NEQ_W * s;
    const double *p0 = params ? params + 8 * s : nullptr;
    if (kind[s] == 3)
        solve_source<7, 7>(s, q, p0, lambda, X, trial, nsteps, steplim);
    else if (kind[s])
        solve_source<6, 6>(s, q, p0, lambda, X, trial, nsteps, steplim);
    else
        solve_source<3, 6>(s, q, p0, lambda, X, trial, nsteps, steplim);
}

// One thread per (source, trial step) of the line search.
__global__ void __launch_bounds__(PREP_TB) k_fit_prep_ls(int S, const int32_t *__restrict__ kind,
                              const int32_t *__restrict__ Parr,
                              const int32_t *__restrict__ stride,
                              const int64_t *__restrict__ pix_off,
                              const double *__restrict__ trial,
                              const int32_t *__restrict__ nsteps,
                              const int32_t *__restrict__ active,
                              const FitProfiles *__restrict__ pr,
                              const SersicTab *__restrict__ ser, int ph, int pw,
                              trh_fft_job *__restrict__ jobs, int8_t *__restrict__ bk,
                              int32_t *__restrict__ bcnt, int32_t *__restrict__ flags,
                              double *__restrict__ chi2ls,
                              unsigned long long *__restrict__ acc,
                              const FitHybrid *__restrict__ hyb, FitMogJob *__restrict__ mog,
                              const int64_t *__restrict__ boff) {
    __shared__ trh_fft_job sj[PREP_TB];
    __shared__ unsigned long long lacc[2 * TRH_NSLOT];
    BlockAcc A;
    A.init(acc, lacc);
    const int slot = blockIdx.x * blockDim.x + threadIdx.x;
    const int s = slot / NALPHA, a = slot - s * NALPHA;
    int b = -1;
    if (s < S && (!active || active[s]) && a < nsteps[s]) {
        const int kd = kind[s];
        const int P = Parr[s];
        const double *p = trial + (size_t)slot * 8;
        trh_fft_job &J = sj[threadIdx.x];
        if (kd == 0) {
            build_point_job(J, p[0], p[1], P, ph, pw);
            J.pad0 = NEQ_W * s + NEQ_CHI0;   // the window's zero-model chi^2 in neq
            b = PTB;
        } else {
            double sa[TRH_MAXK + 1], sv[TRH_MAXK + 1], dv[TRH_MAXK][3];
            const double *amp, *var;
            const int K = gal_profile(*pr, ser, kd, p[6], amp, var, sa, sv);
            const int sz = build_gal_job(J, amp, var, K, p[0], p[1], p[3], p[4], p[5], P, ph, pw,
                                         hyb ? dv : nullptr);
            b = sz_bucket(sz);
            J.pad0 = 0;   // no base plane
            if (b >= 0 && hyb) {
                FitMogJob &M = mog[slot];
                M.off = boff[s] + (int64_t)a * P * P;
                M.stride = P;
                M.P = P;
                if (!hybrid_split(*hyb, sz, K, amp, dv, J, M, p[0], p[1], P)) {
                    atomicOr(&flags[s], TRH_FIT_F_HYBRID);
                    b = -1;
                } else if (M.K > 0) {
                    // its pixel-space part, rendered first (fits int32:
                    // checked by trh_fitter_set_hybrid)
                    J.pad0 = (int32_t)(M.off + 1);
                }
            }
            if (b < 0) {
                // not rendered: the trial counts as non-finite, which ends
                // the walk over the steps there (tryUpdates' isfinite break)
                chi2ls[slot] = INFINITY;
                if (sz_bucket(sz) < 0) atomicOr(&flags[s], TRH_FIT_F_LS_SIZE);
            }
        }
        J.out_offset = pix_off[s];
        J.out_stride = stride[s];
        J.mode = TRH_CHI2;
        J.scale = (float)p[2];
        J.slot = slot;
        if (acc && b >= 0) {
            const int sl = b == PTB ? SL_LPT : SL_LS + b;
            // data + inverr read: the window, or for a point source the
            // stamp's support in it (3 pixels of Lanczos taps around it)
            unsigned long long px = (unsigned long long)P * P;
            if (b == PTB) {
                const int w = max(0, min(P, J.sx + pw + 3) - max(0, J.sx - 3));
                const int hh = max(0, min(P, J.sy + ph + 3) - max(0, J.sy - 3));
                px = (unsigned long long)w * hh;
            }
            A.add(sl, 8ull * px);
        }
    }
    __shared__ int8_t built[PREP_TB];
    built[threadIdx.x] = b >= 0;
    block_bucket_counts(b, slot, NALPHA * S, bk, bcnt);
    __syncthreads();
    A.flush();
    const int first = blockIdx.x * blockDim.x;
    copy_jobs_out(sj, built, jobs + first, min((int)blockDim.x, NALPHA * S - first));
}

// tryUpdates acceptance (constrained_optimizer.py:36-82) and the optimize_loop
// stopping rule (constrained_optimizer.py:12-34).
__global__ void k_fit_accept(int S, const double *__restrict__ neq,
                             const double *__restrict__ trial,
                             const int32_t *__restrict__ nsteps,
                             const int32_t *__restrict__ steplim,
                             const double *__restrict__ chi2ls, double *__restrict__ params,
                             int32_t *__restrict__ active, int32_t *__restrict__ iters,
                             double *__restrict__ chisq, double *__restrict__ dlnp_out,
                             int32_t *__restrict__ flags, double dchisq, int max_iter) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= S || !active[s]) return;
    const double chi0 = neq[(size_t)NEQ_W * s + NEQ_CHI2];
    const double lnp0 = -0.5 * chi0;
    double best = lnp0;
    int ib = -1;
    const int n = nsteps[s];
    for (int a = 0; a < n; ++a) {
        const double lnp = -0.5 * chi2ls[s * NALPHA + a];
        if (!isfinite(lnp)) break;
        if (lnp < best - 1.0) break;
        if (lnp > best) {
            best = lnp;
            ib = a;
        }
    }
    double dl = 0.0;
    if (ib >= 0) {
        const double *tp = trial + ((size_t)s * NALPHA + ib) * 8;
        for (int i = 0; i < 8; ++i) params[8 * s + i] = tp[i];
        dl = best - lnp0;
        chisq[s] = chi2ls[s * NALPHA + ib];
    } else {
        chisq[s] = chi0;
    }
    dlnp_out[s] = dl;
    const int it = iters[s] + 1;
    iters[s] = it;
    const bool limited = steplim[s] || n == 0;
    const bool rule = (!limited && dl <= dchisq) || (limited && dl <= 1e-6);
    int fl = flags[s] & ~(TRH_FIT_F_LIMITED | TRH_FIT_F_MAXITER);
    if (limited) fl |= TRH_FIT_F_LIMITED;
    if (!rule && it >= max_iter) fl |= TRH_FIT_F_MAXITER;
    flags[s] = fl;
    if (rule || it >= max_iter) active[s] = 0;
}

// Chained restarts, after k_fit_accept: a pending pair whose first source
// has stopped starts its second one (from the parameters set for it) when
// the first's chi^2 is above the reference, NaN included; otherwise the
// pair is settled without it.
__global__ void k_fit_chain(int n, const int32_t *__restrict__ first,
                            const int32_t *__restrict__ second, const double *__restrict__ ref,
                            double tol, int32_t *__restrict__ state, int32_t *__restrict__ active,
                            int32_t *__restrict__ iters, int32_t *__restrict__ flags,
                            const double *__restrict__ chisq) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || state[i] != 0 || active[first[i]]) return;
    const bool go = !(chisq[first[i]] <= ref[i] * (1.0 + tol));
    state[i] = go ? 1 : 2;
    if (go) {
        const int b = second[i];
        active[b] = 1;
        iters[b] = 0;
        flags[b] = 0;
    }
}

// Every pair pending again and its second source stopped until then.
__global__ void k_fit_chain_reset(int n, const int32_t *__restrict__ second,
                                  int32_t *__restrict__ state, int32_t *__restrict__ active) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    state[i] = 0;
    active[second[i]] = 0;
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
// Drop the captured iteration (its launches change, or the fitter goes):
// after the replays already issued on the fitter's stream.
static void drop_graph(trh_fitter *f) {
    bool any = false;
    for (auto &g : f->gs) any = any || g.x;
    if (!any) return;
    if (f->gstream) (void)hipStreamSynchronize(f->gstream);
    for (auto &g : f->gs) {
        if (g.x) (void)hipGraphExecDestroy(g.x);
        g.x = nullptr;
    }
}

template <typename T>
static int falloc(trh_fitter *f, T **p, size_t n) {
    void *q = nullptr;
    if (hipMalloc(&q, sizeof(T) * (n ? n : 1)) != hipSuccess)
        return set_err(TRH_E_NOMEM, "fitter allocation of %zu bytes failed", sizeof(T) * n);
    f->allocs.push_back(q);
    *p = (T *)q;
    return TRH_OK;
}

static void host_profiles(FitProfiles &pr) {
    // mixture_profiles.py:17-38, normalised as ExpGalaxy/DevGalaxy do
    // (galaxy.py:729, 748)
    const double ea[6] = {2.34853813e-03, 3.07995260e-02, 2.23364214e-01,
                          1.17949102e+00, 4.33873750e+00, 5.99820770e+00};
    const double ev[6] = {1.20078965e-03, 8.84526493e-03, 3.91463084e-02,
                          1.39976817e-01, 4.60962500e-01, 1.50159566e+00};
    const double da[8] = {4.26347652e-02, 2.40127183e-01, 6.85907632e-01, 1.51937350e+00,
                          2.83627243e+00, 4.46467501e+00, 5.72440830e+00, 5.60989349e+00};
    const double dv[8] = {2.23759216e-04, 1.00220099e-03, 4.18731126e-03, 1.69432589e-02,
                          6.84850479e-02, 2.87207080e-01, 1.33320254e+00, 8.40215071e+00};
    double s = 0;
    for (int i = 0; i < 6; ++i) s += ea[i];
    double a6[6], s2 = 0;
    for (int i = 0; i < 6; ++i) {
        a6[i] = ea[i] / s;
        s2 += a6[i];
    }
    for (int i = 0; i < 6; ++i) {
        pr.amp[0][i] = a6[i] / s2;
        pr.var[0][i] = ev[i];
    }
    pr.K[0] = 6;
    const double core = 0.010233;
    double sd = 0;
    for (int i = 0; i < 8; ++i) sd += da[i];
    double a9[9];
    a9[0] = core;
    for (int i = 0; i < 8; ++i) a9[i + 1] = da[i] * (1.0 - core) / sd;
    double s9 = 0;
    for (int i = 0; i < 9; ++i) s9 += a9[i];
    for (int i = 0; i < 9; ++i) pr.amp[1][i] = a9[i] / s9;
    pr.var[1][0] = 0.0;
    for (int i = 0; i < 8; ++i) pr.var[1][i + 1] = dv[i];
    pr.K[1] = 9;
}

extern "C" int trh_fitter_destroy(trh_fitter *f) {
    if (!f) return TRH_OK;
    (void)hipSetDevice(f->h->device);
    // replays still in flight read the buffers: let them finish first
    drop_graph(f);
    if (f->gstream) (void)hipStreamSynchronize(f->gstream);
    for (void *p : f->allocs) (void)hipFree(p);
    if (f->have_ev) {
        for (int i = 0; i < 6; ++i) (void)hipEventDestroy(f->ev[i]);
        for (int i = 0; i < TRH_NSLOT; ++i)
            for (int j = 0; j < 2; ++j) (void)hipEventDestroy(f->pev[i][j]);
    }
    if (f->have_gev)
        for (int i = 0; i < 2; ++i) (void)hipEventDestroy(f->gev[i]);
    if (f->have_br) {
        (void)hipStreamSynchronize(f->bside);
        (void)hipStreamDestroy(f->bside);
        for (int p = 0; p < 2; ++p) {
            (void)hipEventDestroy(f->bfork[p]);
            (void)hipEventDestroy(f->bjoin[p]);
        }
    }
    if (f->have_count) {
        (void)hipStreamSynchronize(f->cstream);
        (void)hipStreamDestroy(f->cstream);
        for (int i = 0; i < 2; ++i) (void)hipEventDestroy(f->cev[i]);
        (void)hipHostFree(f->nact_pin);
    }
    if (f->gstream) (void)hipStreamDestroy(f->gstream);
    delete f;
    return TRH_OK;
}

extern "C" int trh_fitter_create(trh_handle *h, trh_fitter **fp, int S, const int32_t *kind,
                                 const int32_t *P, const int64_t *pix_off,
                                 const int32_t *row_stride, const float *data,
                                 const float *inverr, const float *psf_stamp, int ph, int pw,
                                 const double *cdi) {
    if (!h || !fp || S <= 0 || !kind || !P || !pix_off || !data || !inverr || !psf_stamp)
        return set_err(TRH_E_ARG, "trh_fitter_create: bad arguments");
    if ((ph % 2) == 0 || (pw % 2) == 0)
        return set_err(TRH_E_ARG, "trh_fitter_create: PSF stamp must have odd size");
    int rc;
    if ((rc = hip_check(hipSetDevice(h->device), "hipSetDevice"))) return rc;
    trh_fitter *f = new trh_fitter();
    f->h = h;
    f->S = S;
    f->ph = ph;
    f->pw = pw;
    f->data = data;
    f->inverr = inverr;
    if (cdi)
        for (int i = 0; i < 4; ++i) f->cdi[i] = cdi[i];
    std::vector<int64_t> roff(S);
    int64_t tot = 0;
    int maxP_sz[NFFT] = {0, 0, 0, 0, 0};
    for (int s = 0; s < S; ++s) {
        if (kind[s] < 0 || kind[s] > 3 || P[s] < 3 || P[s] > 4096 ||
            (row_stride && row_stride[s] < P[s]) || pix_off[s] < 0) {
            trh_fitter_destroy(f);
            return set_err(TRH_E_ARG,
                           "trh_fitter_create: source %d has kind %d / P %d / row stride %d", s,
                           kind[s], P[s], row_stride ? row_stride[s] : P[s]);
        }
        roff[s] = tot;
        tot += (int64_t)(kind[s] == 3 ? 5 : (kind[s] ? 4 : 1)) * P[s] * P[s];
        if (kind[s]) {
            ++f->n_gal;
            if (kind[s] == 3) ++f->n_ser;
            for (int b = 0; b < NFFT; ++b)
                if (P[s] < kSizes[b] && P[s] > maxP_sz[b]) maxP_sz[b] = P[s];
        } else {
            ++f->n_pt;
        }
    }
    {
        const int hs = (int)ceil(fmax(pw / 2.0, ph / 2.0));
        while (f->min_sz < 2 * hs && f->min_sz < kSizes[NFFT - 1]) f->min_sz <<= 1;
    }
    for (int b = 0; b < NFFT; ++b) {
        int mr = max_rows_for(kSizes[b], h->lds_max);
        if (maxP_sz[b] > 0 && maxP_sz[b] < mr) mr = maxP_sz[b];
        if (mr < 8) mr = 8;
        f->maxrows[b] = mr;
    }
    if (f->n_pt && point_lds_bytes(ph, pw) + sizeof(RenderScratch) > (size_t)h->lds_max) {
        trh_fitter_destroy(f);
        return set_err(TRH_E_UNSUPPORTED, "point-source PSF stamp %dx%d too large for LDS", ph, pw);
    }
    f->nj = f->n_ser ? 5 : 4;
    const int cap_jac = f->nj * S, cap_ls = NALPHA * S;
    if ((rc = falloc(f, &f->kind, S)) || (rc = falloc(f, &f->P, S)) ||
        (rc = falloc(f, &f->stride, S)) || (rc = falloc(f, &f->flags, S)) ||
        (rc = falloc(f, &f->sstep, S)) ||
        (rc = falloc(f, &f->pix_off, S)) || (rc = falloc(f, &f->render_off, S)) ||
        (rc = falloc(f, &f->stamp, (size_t)ph * pw)) || (rc = falloc(f, &f->renders, tot)) ||
        (rc = falloc(f, &f->params, 8 * (size_t)S)) ||
        (rc = falloc(f, &f->trial, 8 * (size_t)S * NALPHA)) || (rc = falloc(f, &f->X, 8 * (size_t)S)) ||
        (rc = falloc(f, &f->neq, NEQ_W * (size_t)S)) || (rc = falloc(f, &f->chi2ls, (size_t)S * NALPHA)) ||
        (rc = falloc(f, &f->nsteps, S)) || (rc = falloc(f, &f->steplim, S)) ||
        (rc = falloc(f, &f->active, S)) || (rc = falloc(f, &f->iters, S)) ||
        (rc = falloc(f, &f->chisq, S)) || (rc = falloc(f, &f->dlnp, S)) ||
        (rc = falloc(f, &f->prof, 1)) ||
        (rc = falloc(f, &f->jobs_jac, (size_t)cap_jac)) || (rc = falloc(f, &f->jobs_ls, (size_t)cap_ls)) ||
        (rc = falloc(f, &f->list_jac, (size_t)NBUCKET * cap_jac)) ||
        (rc = falloc(f, &f->list_ls, (size_t)NBUCKET * cap_ls)) || (rc = falloc(f, &f->cnt, 16)) ||
        (rc = falloc(f, &f->neq_order, S)) || (rc = falloc(f, &f->bk_jac, (size_t)cap_jac)) ||
        (rc = falloc(f, &f->bk_ls, (size_t)cap_ls)) ||
        (rc = falloc(f, &f->bcnt_jac, (size_t)NBUCKET * ((cap_jac + PREP_TB - 1) / PREP_TB))) ||
        (rc = falloc(f, &f->bcnt_ls, (size_t)NBUCKET * ((cap_ls + PREP_TB - 1) / PREP_TB)))) {
        trh_fitter_destroy(f);
        return rc;
    }
    for (int b = 0; b < NFFT; ++b) {
        const int sz = kSizes[b];
        if ((rc = falloc(f, &f->spec[b], (size_t)2 * sz * (sz / 2 + 1)))) {
            trh_fitter_destroy(f);
            return rc;
        }
    }
    if (f->n_gal) {
        // own row slots of the 256 and 512 renders (every galaxy bucket is
        // launched from a captured graph, so both must exist)
        for (int i = 0; i < 2; ++i) {
            const int sz = i ? 512 : 256;
            const size_t n = fft_hbuf_slot_floats(sz) * (size_t)fft_hbuf_slots(h, sz) / 2;
            if ((rc = falloc(f, &f->hbuf[i], n))) {
                trh_fitter_destroy(f);
                return rc;
            }
        }
    }
    FitProfiles pr;
    host_profiles(pr);
    std::vector<int32_t> ones(S, 1), zeros(S, 0), order(S), strides(S);
    for (int s = 0; s < S; ++s) strides[s] = row_stride ? row_stride[s] : P[s];
    for (int s = 0; s < S; ++s) order[s] = s;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return P[a] > P[b]; });
    if ((rc = hip_check(hipMemcpy(f->kind, kind, 4 * (size_t)S, hipMemcpyHostToDevice), "copy kind")) ||
        (rc = hip_check(hipMemcpy(f->stride, strides.data(), 4 * (size_t)S, hipMemcpyHostToDevice),
                        "copy strides")) ||
        (rc = hip_check(hipMemcpy(f->flags, zeros.data(), 4 * (size_t)S, hipMemcpyHostToDevice),
                        "copy flags")) ||
        (rc = hip_check(hipMemcpy(f->P, P, 4 * (size_t)S, hipMemcpyHostToDevice), "copy P")) ||
        (rc = hip_check(hipMemcpy(f->neq_order, order.data(), 4 * (size_t)S, hipMemcpyHostToDevice),
                        "copy order")) ||
        (rc = hip_check(hipMemcpy(f->pix_off, pix_off, 8 * (size_t)S, hipMemcpyHostToDevice), "copy pix_off")) ||
        (rc = hip_check(hipMemcpy(f->render_off, roff.data(), 8 * (size_t)S, hipMemcpyHostToDevice), "copy roff")) ||
        (rc = hip_check(hipMemcpy(f->stamp, psf_stamp, 4 * (size_t)ph * pw, hipMemcpyHostToDevice), "copy stamp")) ||
        (rc = hip_check(hipMemcpy(f->prof, &pr, sizeof(pr), hipMemcpyHostToDevice), "copy profiles")) ||
        (rc = hip_check(hipMemcpy(f->active, ones.data(), 4 * (size_t)S, hipMemcpyHostToDevice), "copy active")) ||
        (rc = hip_check(hipMemcpy(f->iters, zeros.data(), 4 * (size_t)S, hipMemcpyHostToDevice), "copy iters"))) {
        trh_fitter_destroy(f);
        return rc;
    }
    for (int b = 0; b < NFFT; ++b) {
        if ((rc = trh_psf_spectrum_batch(h, f->stamp, 1, ph, pw, kSizes[b], f->spec[b], &f->cx[b],
                                         &f->cy[b], 0))) {
            trh_fitter_destroy(f);
            return rc;
        }
    }
    for (int i = 0; i < 6; ++i) (void)hipEventCreate(&f->ev[i]);
    for (int i = 0; i < TRH_NSLOT; ++i)
        for (int j = 0; j < 2; ++j) (void)hipEventCreate(&f->pev[i][j]);
    f->have_ev = true;
    if ((rc = falloc(f, &f->acc, 2 * TRH_NSLOT))) {
        trh_fitter_destroy(f);
        return rc;
    }
    if ((rc = hip_check(hipDeviceSynchronize(), "fitter create sync"))) {
        trh_fitter_destroy(f);
        return rc;
    }
    *fp = f;
    return TRH_OK;
}

extern "C" int trh_fitter_set_branching(trh_fitter *f, int below_active) {
    if (!f || below_active < 0)
        return set_err(TRH_E_ARG, "trh_fitter_set_branching: bad arguments");
    f->branch_below = below_active;
    return TRH_OK;
}

extern "C" int trh_fitter_set_params(trh_fitter *f, const double *params, void *stream) {
    if (!f || !params) return set_err(TRH_E_ARG, "trh_fitter_set_params: null argument");
    f->last_active = -1;   // every source starts again: the unbranched graph
    std::vector<double> p8((size_t)8 * f->S, 0.0);
    for (int s = 0; s < f->S; ++s)
        for (int i = 0; i < TRH_FIT_NPARAM; ++i)
            p8[8 * (size_t)s + i] = params[TRH_FIT_NPARAM * (size_t)s + i];
    std::vector<int32_t> ones(f->S, 1), zeros(f->S, 0);
    hipStream_t st = (hipStream_t)stream;
    int rc;
    // after every graph replay issued so far (they run on the fitter's own
    // stream); the copies are ordered on `stream`, and the host buffers are
    // released only once they are done
    if (f->have_gev && (rc = hip_check(hipStreamWaitEvent(st, f->gev[1], 0), "order after replays")))
        return rc;
    if ((rc = hip_check(hipMemcpyAsync(f->params, p8.data(), 8 * p8.size(), hipMemcpyHostToDevice, st),
                        "set params")) ||
        (rc = hip_check(hipMemcpyAsync(f->active, ones.data(), 4 * (size_t)f->S,
                                       hipMemcpyHostToDevice, st), "reset active")) ||
        (rc = hip_check(hipMemcpyAsync(f->iters, zeros.data(), 4 * (size_t)f->S,
                                       hipMemcpyHostToDevice, st), "reset iters")) ||
        (rc = hip_check(hipMemcpyAsync(f->flags, zeros.data(), 4 * (size_t)f->S,
                                       hipMemcpyHostToDevice, st), "reset flags")))
        return rc;
    if (f->n_chain > 0) {
        // the second source of every chained pair waits for its first
        hipLaunchKernelGGL(k_fit_chain_reset, dim3((f->n_chain + 255) / 256), dim3(256), 0, st,
                           f->n_chain, f->ch_second, f->ch_state, f->active);
        if ((rc = hip_check(hipGetLastError(), "k_fit_chain_reset"))) return rc;
    }
    return hip_check(hipStreamSynchronize(st), "set params sync");
}

extern "C" int trh_fitter_get_params(trh_fitter *f, double *params, void *stream) {
    if (!f || !params) return set_err(TRH_E_ARG, "trh_fitter_get_params: null argument");
    int rc;
    if ((rc = hip_check(hipStreamSynchronize((hipStream_t)stream), "sync"))) return rc;
    std::vector<double> p8((size_t)8 * f->S);
    if ((rc = hip_check(hipMemcpy(p8.data(), f->params, 8 * p8.size(), hipMemcpyDeviceToHost), "get params")))
        return rc;
    for (int s = 0; s < f->S; ++s)
        for (int i = 0; i < TRH_FIT_NPARAM; ++i)
            params[TRH_FIT_NPARAM * (size_t)s + i] = p8[8 * (size_t)s + i];
    return TRH_OK;
}

extern "C" int trh_fitter_set_sersic(trh_fitter *f, const double *tab, int ntab) {
    if (!f || !tab || ntab < 1) return set_err(TRH_E_ARG, "trh_fitter_set_sersic: bad arguments");
    SersicTab *T = new SersicTab();
    int pos = 0, nd = 0;
    // next value of the flat table, -1e300 past its end (rejected below)
    auto next = [&]() { return pos < ntab ? tab[pos++] : (++pos, -1e300); };
    // a count read from the table: an integer in [lo, hi]
    auto count = [&](int lo, int hi, int &out) {
        const double v = next();
        if (!(v >= lo && v <= hi) || v != (double)(int)v) return false;
        out = (int)v;
        return true;
    };
    auto copy = [&](int n) {
        if (nd + n > SER_MAXD || pos + n > ntab) return false;
        for (int i = 0; i < n; ++i) T->d[nd++] = tab[pos++];
        return true;
    };
    bool ok = count(1, SER_MAXR, T->nr);
    int maxK = 0;
    for (int r = 0; ok && r < T->nr; ++r) {
        T->lo[r] = next();
        T->hi[r] = next();
        // ranges in increasing lo, each starting at or before the previous
        // one's end (no gaps), at most two covering any index
        ok = T->lo[r] < T->hi[r] &&
             (r == 0 || (T->lo[r] > T->lo[r - 1] && T->lo[r] <= T->hi[r - 1] &&
                         T->hi[r] > T->hi[r - 1])) &&
             (r < 2 || T->lo[r] >= T->hi[r - 2]) &&
             count(1, TRH_MAXK, T->K[r]) && count(1, 256, T->nint[r]);
        if (!ok) break;
        T->boff[r] = nd;
        ok = copy(T->nint[r] + 1);
        T->coff[r] = nd;
        ok = ok && copy(8 * T->K[r] * T->nint[r]);
        int k = T->K[r];
        // a ramp between overlapping ranges carries both ranges' components
        if (r > 0 && T->lo[r] < T->hi[r - 1]) k += T->K[r - 1];
        if (k > maxK) maxK = k;
    }
    ok = ok && count(1, 256, T->core_nint);
    if (ok) {
        T->core_boff = nd;
        ok = copy(T->core_nint + 1);
        T->core_coff = nd;
        ok = ok && copy(4 * T->core_nint) && pos == ntab;
    }
    if (ok && maxK + 1 > TRH_MAXK) ok = false;
    int rc = TRH_OK;
    if (!ok)
        rc = set_err(TRH_E_ARG, "trh_fitter_set_sersic: malformed table (at value %d of %d)", pos,
                     ntab);
    else if (!f->ser)
        rc = falloc(f, &f->ser, 1);
    if (!rc)
        rc = hip_check(hipMemcpy(f->ser, T, sizeof(SersicTab), hipMemcpyHostToDevice), "copy sersic table");
    delete T;
    return rc;
}

static int launch_point(trh_fitter *f, const trh_fft_job *jobs, const int32_t *list,
                        const int32_t *count, int bound, float *out, double *chi2,
                        hipStream_t st, int mode = TRH_STORE, bool window_chi0 = true) {
    if (bound <= 0) return TRH_OK;
    const size_t lds = point_lds_bytes(f->ph, f->pw);
    int wg = (int)((160 * 1024) / (lds + sizeof(RenderScratch)));
    if (wg < 1) wg = 1;
    if (wg > 8) wg = 8;
    int grid = f->h->n_cu * wg;
    if (bound < grid) grid = bound;
    if (chi2)
        // the window's zero-model chi^2 comes from the normal equations of
        // the same iteration (neq[32 s + 28], indexed by the job's pad0)
        hipLaunchKernelGGL(k_render_point<TRH_CHI2>, dim3(grid), dim3(256), lds, st, jobs, list,
                           count, bound, f->stamp, f->ph, f->pw, out, f->data, f->inverr, chi2,
                           window_chi0 ? f->neq : nullptr);
    else if (mode == TRH_STORE)
        hipLaunchKernelGGL(k_render_point<TRH_STORE>, dim3(grid), dim3(256), lds, st, jobs, list,
                           count, bound, f->stamp, f->ph, f->pw, out, f->data, f->inverr, chi2,
                           nullptr);
    else
        hipLaunchKernelGGL(k_render_point<-1>, dim3(grid), dim3(256), lds, st, jobs, list,
                           count, bound, f->stamp, f->ph, f->pw, out, f->data, f->inverr, chi2,
                           nullptr);
    return hip_check(hipGetLastError(), "k_render_point");
}

// Record the start / end event of launch slot `k` when profiling.
#define SLOT_BEGIN(k) do { if (f->profiling) (void)hipEventRecord(f->pev[k][0], st); } while (0)
#define SLOT_END(k) do { if (f->profiling) (void)hipEventRecord(f->pev[k][1], st); } while (0)

// The launches of one LM iteration on `st`.  `capture`: the stream is being
// captured into a graph, so no events and no per-launch profiling.
static float2 *fitter_hbuf(trh_fitter *f, int sz) {
    return sz == 256 ? f->hbuf[0] : (sz == 512 ? f->hbuf[1] : nullptr);
}

// The pixel-space parts of a hybrid fitter's galaxy jobs (k_fit_mog).
static int launch_mog(trh_fitter *f, const FitMogJob *jobs, const int32_t *lists, int cap,
                      const int32_t *cnt, int mode, float *out, int nplanes, int64_t plane,
                      hipStream_t st) {
    if (!f->hyb || !f->n_gal) return TRH_OK;
    int grid = f->h->n_cu * 8;
    if (grid > NALPHA * f->n_gal) grid = NALPHA * f->n_gal;
    hipLaunchKernelGGL(k_fit_mog, dim3(grid), dim3(256), 0, st, jobs, lists, cap, cnt, mode, out,
                       nplanes, plane);
    return hip_check(hipGetLastError(), "k_fit_mog");
}

// Fork point of a render phase in a branched capture: the side stream
// starts after everything enqueued on `st` so far.
static int branch_fork(trh_fitter *f, int ph, hipStream_t st) {
    int rc = hip_check(hipEventRecord(f->bfork[ph], st), "branch fork");
    if (!rc) rc = hip_check(hipStreamWaitEvent(f->bside, f->bfork[ph], 0), "branch wait");
    return rc;
}

// Join of a render phase: `st` continues after the side stream's work.
static int branch_join(trh_fitter *f, int ph, hipStream_t st) {
    int rc = hip_check(hipEventRecord(f->bjoin[ph], f->bside), "branch join");
    if (!rc) rc = hip_check(hipStreamWaitEvent(st, f->bjoin[ph], 0), "branch join wait");
    return rc;
}

// Stream of render bucket b in a branched iteration: the sz = 128 bucket,
// the largest share of the renders, stays on the iteration's stream and
// every other bucket runs, in order, on the one side stream.  Measured on a
// 16384-source fit: 110.5 ms, against 112.4 with a side stream per bucket
// (each cross-stream edge of a captured graph costs ~10 us at the fork and
// the join, and streams beyond the hardware queues share them) and 116.0
// with no branching.
static hipStream_t bucket_stream(const trh_fitter *f, int b, bool branch, hipStream_t st) {
    const bool stays = b < NFFT && kSizes[b] == 128;
    return branch && !stays ? f->bside : st;
}

// The launches of one LM iteration.  `branch` (graph capture only): in each
// render phase the sz = 128 bucket runs beside the other buckets, which run
// on the side stream between a fork and a join (they write disjoint
// outputs: every job its own render window or chi^2 slot, every FFT size
// its own row slots), so under few active sources the latency-bound
// launches overlap instead of running one after another.  Every job's
// arithmetic is the same either way, so the results are identical.
static int iterate_body(trh_fitter *f, double lambda, double dchisq, int max_iter, hipStream_t st,
                        bool capture, bool branch = false) {
    const int S = f->S;
    const int cap_jac = f->nj * S, cap_ls = NALPHA * S;
    const int tb = 256, nb = (S + tb - 1) / tb;
    if (f->n_ser && !f->ser)
        return set_err(TRH_E_ARG, "trh_fitter_iterate: Sersic galaxies need trh_fitter_set_sersic");
    const bool prof = f->profiling && !capture;
    unsigned long long *acc = prof ? f->acc : nullptr;
    int rc;
    if (!capture) (void)hipEventRecord(f->ev[0], st);
    SLOT_BEGIN(SL_PREP);
    // (no counter reset: each k_fit_lists writes every bucket count it owns)
    const int nb_jac = (cap_jac + PREP_TB - 1) / PREP_TB, nb_ls = (cap_ls + PREP_TB - 1) / PREP_TB;
    hipLaunchKernelGGL(k_fit_prep_jac, dim3(nb_jac), dim3(PREP_TB), 0, st, S, f->nj, f->kind, f->P,
                       f->render_off, f->params, f->active, f->prof, f->ser, f->sstep, f->ph, f->pw, f->jobs_jac,
                       f->bk_jac, f->bcnt_jac, f->flags, f->active, acc, f->hyb, f->mog_jac);
    if ((rc = hip_check(hipGetLastError(), "k_fit_prep_jac"))) return rc;
    hipLaunchKernelGGL(k_fit_lists, dim3(nb_jac), dim3(PREP_TB), 0, st, cap_jac, f->bk_jac,
                       f->bcnt_jac, f->list_jac, cap_jac, f->cnt);
    if ((rc = hip_check(hipGetLastError(), "k_fit_lists"))) return rc;
    SLOT_END(SL_PREP);
    branch = branch && capture && f->have_br;
    if (branch && (rc = branch_fork(f, 0, st))) return rc;
    for (int b = 0; b < NFFT && f->n_gal; ++b) {
        if (kSizes[b] < f->min_sz) continue;
        SLOT_BEGIN(b);
        if ((rc = launch_fft_any(f->h, kSizes[b], TRH_STORE, f->jobs_jac,
                                 f->list_jac + (size_t)b * cap_jac, f->cnt + b, 4 * f->n_gal + f->n_ser, f->maxrows[b], f->spec[b], f->renders,
                                 nullptr, nullptr, nullptr, bucket_stream(f, b, branch, st),
                                 fitter_hbuf(f, kSizes[b]))))
            return rc;
        SLOT_END(b);
    }
    SLOT_BEGIN(SL_JPT);
    if ((rc = launch_point(f, f->jobs_jac, f->list_jac + (size_t)PTB * cap_jac, f->cnt + PTB, f->n_pt,
                           f->renders, nullptr, bucket_stream(f, PTB, branch, st))))
        return rc;
    SLOT_END(SL_JPT);
    if (branch && (rc = branch_join(f, 0, st))) return rc;
    // hybrid PSF: the pixel-space parts added to the galaxies' FFT renders
    // (after them: they add into the same planes)
    if ((rc = launch_mog(f, f->mog_jac, f->list_jac, cap_jac, f->cnt, 1, f->renders, 1, 0, st)))
        return rc;
    if (!capture) (void)hipEventRecord(f->ev[1], st);
    SLOT_BEGIN(SL_NEQ);
    {
        const int grid = S < f->h->n_cu * 16 ? S : f->h->n_cu * 16;
        if (S > f->n_ser) {
            hipLaunchKernelGGL(k_fit_normal_eq<6>, dim3(grid), dim3(NEQ_TB), 0, st, S, f->kind, f->P,
                               f->stride, f->pix_off, f->render_off, f->data, f->inverr,
                               f->renders, f->params + 2, 8, 0.01, f->sstep, f->active,
                               f->cdi[0], f->cdi[1], f->cdi[2], f->cdi[3], f->neq, f->neq_order);
            if ((rc = hip_check(hipGetLastError(), "k_fit_normal_eq<6>"))) return rc;
        }
        if (f->n_ser) {
            const int g7 = f->n_ser < f->h->n_cu * 8 ? f->n_ser : f->h->n_cu * 8;
            hipLaunchKernelGGL(k_fit_normal_eq<7>, dim3(g7), dim3(NEQ_TB), 0, st, S, f->kind, f->P,
                               f->stride, f->pix_off, f->render_off, f->data, f->inverr,
                               f->renders, f->params + 2, 8, 0.01, f->sstep, f->active,
                               f->cdi[0], f->cdi[1], f->cdi[2], f->cdi[3], f->neq, f->neq_order);
            if ((rc = hip_check(hipGetLastError(), "k_fit_normal_eq<7>"))) return rc;
        }
    }
    SLOT_END(SL_NEQ);
    if (!capture) (void)hipEventRecord(f->ev[2], st);
    SLOT_BEGIN(SL_SOLVE);
    hipLaunchKernelGGL(k_fit_solve, dim3((S + 63) / 64), dim3(64), 0, st, S, f->kind, f->neq,
                       f->params, f->active, lambda, f->X, f->trial, f->nsteps, f->steplim);
    if ((rc = hip_check(hipGetLastError(), "k_fit_solve"))) return rc;
    SLOT_END(SL_SOLVE);
    if (!capture) (void)hipEventRecord(f->ev[3], st);
    hipLaunchKernelGGL(k_fit_prep_ls, dim3(nb_ls), dim3(PREP_TB), 0, st, S, f->kind, f->P,
                       f->stride, f->pix_off, f->trial, f->nsteps, f->active, f->prof, f->ser,
                       f->ph, f->pw, f->jobs_ls, f->bk_ls, f->bcnt_ls, f->flags, f->chi2ls, acc,
                       f->hyb, f->mog_ls, f->boff);
    if ((rc = hip_check(hipGetLastError(), "k_fit_prep_ls"))) return rc;
    hipLaunchKernelGGL(k_fit_lists, dim3(nb_ls), dim3(PREP_TB), 0, st, cap_ls, f->bk_ls,
                       f->bcnt_ls, f->list_ls, cap_ls, f->cnt + LS_CNT);
    if ((rc = hip_check(hipGetLastError(), "k_fit_lists"))) return rc;
    // hybrid PSF: each trial's pixel-space part into its ls_base window,
    // which its chi^2 render adds to the FFT model
    if ((rc = launch_mog(f, f->mog_ls, f->list_ls, cap_ls, f->cnt + LS_CNT, 0, f->ls_base, 1, 0, st)))
        return rc;
    if (branch && (rc = branch_fork(f, 1, st))) return rc;
    for (int b = 0; b < NFFT && f->n_gal; ++b) {
        if (kSizes[b] < f->min_sz) continue;
        SLOT_BEGIN(SL_LS + b);
        if ((rc = launch_fft_any(f->h, kSizes[b], TRH_CHI2, f->jobs_ls,
                                 f->list_ls + (size_t)b * cap_ls, f->cnt + LS_CNT + b, NALPHA * f->n_gal, f->maxrows[b], f->spec[b],
                                 nullptr, f->data, f->inverr, f->chi2ls,
                                 bucket_stream(f, b, branch, st), fitter_hbuf(f, kSizes[b]),
                                 f->hyb ? f->ls_base : nullptr)))
            return rc;
        SLOT_END(SL_LS + b);
    }
    SLOT_BEGIN(SL_LPT);
    if ((rc = launch_point(f, f->jobs_ls, f->list_ls + (size_t)PTB * cap_ls, f->cnt + LS_CNT + PTB,
                           NALPHA * f->n_pt, nullptr, f->chi2ls, bucket_stream(f, PTB, branch, st))))
        return rc;
    SLOT_END(SL_LPT);
    if (branch && (rc = branch_join(f, 1, st))) return rc;
    if (!capture) (void)hipEventRecord(f->ev[4], st);
    hipLaunchKernelGGL(k_fit_accept, dim3(nb), dim3(tb), 0, st, S, f->neq, f->trial, f->nsteps,
                       f->steplim, f->chi2ls, f->params, f->active, f->iters, f->chisq, f->dlnp,
                       f->flags, dchisq, max_iter);
    if ((rc = hip_check(hipGetLastError(), "k_fit_accept"))) return rc;
    if (f->n_chain > 0) {
        hipLaunchKernelGGL(k_fit_chain, dim3((f->n_chain + 255) / 256), dim3(256), 0, st,
                           f->n_chain, f->ch_first, f->ch_second, f->ch_ref, f->ch_tol,
                           f->ch_state, f->active, f->iters, f->flags, f->chisq);
        if ((rc = hip_check(hipGetLastError(), "k_fit_chain"))) return rc;
    }
    if (!capture) (void)hipEventRecord(f->ev[5], st);
    if (prof) {
        // accumulate this iteration's per-slot times (synchronises the stream)
        if ((rc = hip_check(hipEventSynchronize(f->ev[5]), "profile sync"))) return rc;
        int32_t cnt[16];
        if ((rc = hip_check(hipMemcpy(cnt, f->cnt, sizeof(cnt), hipMemcpyDeviceToHost), "read counts")))
            return rc;
        for (int k = 0; k < TRH_NSLOT; ++k) {
            const bool fft = (k < NFFT) || (k >= SL_LS && k < SL_LS + NFFT);
            // a slot whose events were not recorded this iteration is skipped
            // (asking for their elapsed time would leave a sticky HIP error)
            const bool launched =
                fft ? f->n_gal > 0 && kSizes[k < NFFT ? k : k - SL_LS] >= f->min_sz
                    : ((k == SL_JPT || k == SL_LPT) ? f->n_pt > 0 : true);
            if (!launched) continue;
            float ms = 0.f;
            if (hipEventElapsedTime(&ms, f->pev[k][0], f->pev[k][1]) != hipSuccess) continue;
            f->tot_ms[k] += ms;
            f->tot_launch[k] += 1;
            if (fft) {
                const int b = k < NFFT ? k : k - SL_LS;
                const int njobs = k < NFFT ? cnt[b] : cnt[LS_CNT + b];
                const int sz = kSizes[b];
                if (njobs > 0) f->tot_extra_bytes[k] += 8.0 * sz * (sz / 2 + 1);  // PSF spectrum
            }
        }
    }
    return TRH_OK;
}

extern "C" int trh_fitter_iterate(trh_fitter *f, double lambda, double dchisq, int max_iter,
                                  void *stream) {
    if (!f || max_iter < 1) return set_err(TRH_E_ARG, "trh_fitter_iterate: bad arguments");
    return iterate_body(f, lambda, dchisq, max_iter, (hipStream_t)stream, false);
}

extern "C" int trh_fitter_iterate_graph(trh_fitter *f, double lambda, double dchisq, int max_iter,
                                        int n_iter, void *stream) {
    if (!f || n_iter < 0 || max_iter < 1)
        return set_err(TRH_E_ARG, "trh_fitter_iterate_graph: bad arguments");
    if (f->profiling)
        return set_err(TRH_E_ARG, "trh_fitter_iterate_graph: per-launch profiling needs "
                                  "trh_fitter_iterate");
    int rc;
    if (!f->gstream &&
        (rc = hip_check(hipStreamCreateWithFlags(&f->gstream, hipStreamNonBlocking), "graph stream")))
        return rc;
    if (!f->have_gev) {
        for (int i = 0; i < 2; ++i)
            if ((rc = hip_check(hipEventCreateWithFlags(&f->gev[i], hipEventDisableTiming),
                                "graph event")))
                return rc;
        f->have_gev = true;
    }
    // the branched iteration once few sources remain (see trh_fitter_set_branching)
    const bool br = f->branch_below > 0 && f->last_active >= 0 && f->last_active < f->branch_below;
    if (br && !f->have_br) {
        // the side stream and its four events, or none (what was made is released)
        int ne = 0;
        hipEvent_t evs[4];
        rc = hip_check(hipStreamCreateWithFlags(&f->bside, hipStreamNonBlocking), "branch stream");
        while (!rc && ne < 4)
            if (!(rc = hip_check(hipEventCreateWithFlags(&evs[ne], hipEventDisableTiming),
                                 "branch event")))
                ++ne;
        if (rc) {
            for (int i = 0; i < ne; ++i) (void)hipEventDestroy(evs[i]);
            if (f->bside) (void)hipStreamDestroy(f->bside);
            f->bside = nullptr;
            return rc;
        }
        f->bfork[0] = evs[0];
        f->bfork[1] = evs[1];
        f->bjoin[0] = evs[2];
        f->bjoin[1] = evs[3];
        f->have_br = true;
    }
    // GRAPH_ITERS iterations per graph launch when n_iter is a multiple
    const bool multi = n_iter > 0 && n_iter % GRAPH_ITERS == 0;
    const int per = multi ? GRAPH_ITERS : 1;
    trh_fitter::GraphSlot &slot = f->gs[(br ? 2 : 0) + (multi ? 1 : 0)];
    hipGraphExec_t &gx = slot.x;
    double &kl = slot.lambda;
    double &kd = slot.dchisq;
    int &km = slot.max_iter;
    if (!gx || lambda != kl || dchisq != kd || max_iter != km) {
        if (gx) {
            (void)hipStreamSynchronize(f->gstream);
            (void)hipGraphExecDestroy(gx);
            gx = nullptr;
        }
        // (the row slots of the SZ >= 256 renders are the fitter's own,
        // allocated at creation)
        if ((rc = hip_check(hipStreamBeginCapture(f->gstream, hipStreamCaptureModeThreadLocal),
                            "begin capture")))
            return rc;
        int brc = TRH_OK;
        for (int q = 0; q < per && !brc; ++q)
            brc = iterate_body(f, lambda, dchisq, max_iter, f->gstream, true, br);
        hipGraph_t g = nullptr;
        rc = hip_check(hipStreamEndCapture(f->gstream, &g), "end capture");
        if (brc) {
            if (g) (void)hipGraphDestroy(g);
            return brc;
        }
        if (rc) return rc;
        rc = hip_check(hipGraphInstantiate(&gx, g, nullptr, nullptr, 0), "graph instantiate");
        (void)hipGraphDestroy(g);
        if (rc) return rc;
        kl = lambda;
        kd = dchisq;
        km = max_iter;
    }
    hipStream_t st = (hipStream_t)stream;
    if ((rc = hip_check(hipEventRecord(f->gev[0], st), "graph order in")) ||
        (rc = hip_check(hipStreamWaitEvent(f->gstream, f->gev[0], 0), "graph wait in")))
        return rc;
    for (int i = 0; i < n_iter; i += per)
        if ((rc = hip_check(hipGraphLaunch(gx, f->gstream), "graph launch"))) return rc;
    if ((rc = hip_check(hipEventRecord(f->gev[1], f->gstream), "graph order out")) ||
        (rc = hip_check(hipStreamWaitEvent(st, f->gev[1], 0), "graph wait out")))
        return rc;
    return TRH_OK;
}

__global__ void k_fit_zero_u64(unsigned long long *a, int n) {
    if ((int)threadIdx.x < n) a[threadIdx.x] = 0ull;
}

extern "C" int trh_fitter_profile(trh_fitter *f, int enable) {
    if (!f) return set_err(TRH_E_ARG, "trh_fitter_profile: null fitter");
    f->profiling = enable != 0;
    for (int k = 0; k < TRH_NSLOT; ++k) f->tot_ms[k] = f->tot_launch[k] = f->tot_extra_bytes[k] = 0.0;
    hipLaunchKernelGGL(k_fit_zero_u64, dim3(1), dim3(64), 0, 0, f->acc, 2 * TRH_NSLOT);
    int rc = hip_check(hipGetLastError(), "k_fit_zero_u64");
    if (rc) return rc;
    return hip_check(hipDeviceSynchronize(), "profile reset sync");
}

extern "C" int trh_fitter_profile_read(trh_fitter *f, double *out) {
    if (!f || !out) return set_err(TRH_E_ARG, "trh_fitter_profile_read: null argument");
    unsigned long long a[2 * TRH_NSLOT];
    int rc;
    if ((rc = hip_check(hipDeviceSynchronize(), "profile read sync"))) return rc;
    if ((rc = hip_check(hipMemcpy(a, f->acc, sizeof(a), hipMemcpyDeviceToHost), "read acc"))) return rc;
    for (int k = 0; k < TRH_NSLOT; ++k) {
        out[4 * k + 0] = f->tot_ms[k];
        out[4 * k + 1] = f->tot_launch[k];
        out[4 * k + 2] = (double)a[2 * k];
        out[4 * k + 3] = (double)a[2 * k + 1] + f->tot_extra_bytes[k];
    }
    return TRH_OK;
}

extern "C" int trh_fitter_set_chain(trh_fitter *f, int n, const int32_t *first,
                                    const int32_t *second, const double *ref_chisq,
                                    double rel_tol, void *stream) {
    if (!f || n < 0 || (n > 0 && (!first || !second || !ref_chisq)) || !(rel_tol >= 0.0))
        return set_err(TRH_E_ARG, "trh_fitter_set_chain: bad arguments");
    const int S = f->S;
    if (n > S / 2)
        return set_err(TRH_E_ARG, "trh_fitter_set_chain: %d pairs for %d sources", n, S);
    // each source in at most one pair, and in one role
    std::vector<int8_t> used(S, 0);
    for (int i = 0; i < n; ++i) {
        const int a = first[i], b = second[i];
        if (a < 0 || a >= S || b < 0 || b >= S || a == b || used[a] || used[b])
            return set_err(TRH_E_ARG, "trh_fitter_set_chain: pair %d (%d, %d) out of range or "
                                      "reusing a source", i, a, b);
        used[a] = used[b] = 1;
    }
    int rc;
    if ((rc = hip_check(hipSetDevice(f->h->device), "hipSetDevice"))) return rc;
    // the captured iteration gains or loses the chain launch: capture again
    drop_graph(f);
    hipStream_t st = (hipStream_t)stream;
    if (f->have_gev && (rc = hip_check(hipStreamWaitEvent(st, f->gev[1], 0), "order after replays")))
        return rc;
    if (n > 0 && !f->ch_first) {
        const size_t cap = (size_t)S / 2;
        if ((rc = falloc(f, &f->ch_first, cap)) || (rc = falloc(f, &f->ch_second, cap)) ||
            (rc = falloc(f, &f->ch_state, cap)) || (rc = falloc(f, &f->ch_ref, cap)))
            return rc;
    }
    f->n_chain = 0;
    if (n == 0) return TRH_OK;
    if ((rc = hip_check(hipMemcpyAsync(f->ch_first, first, 4 * (size_t)n, hipMemcpyHostToDevice, st),
                        "chain firsts")) ||
        (rc = hip_check(hipMemcpyAsync(f->ch_second, second, 4 * (size_t)n, hipMemcpyHostToDevice,
                                       st), "chain seconds")) ||
        (rc = hip_check(hipMemcpyAsync(f->ch_ref, ref_chisq, 8 * (size_t)n, hipMemcpyHostToDevice,
                                       st), "chain references")))
        return rc;
    f->n_chain = n;
    f->ch_tol = rel_tol;
    hipLaunchKernelGGL(k_fit_chain_reset, dim3((n + 255) / 256), dim3(256), 0, st, n,
                       f->ch_second, f->ch_state, f->active);
    if ((rc = hip_check(hipGetLastError(), "k_fit_chain_reset"))) return rc;
    return hip_check(hipStreamSynchronize(st), "set chain sync");
}

extern "C" int trh_fitter_chain_state(trh_fitter *f, int32_t *state, void *stream) {
    if (!f || !state) return set_err(TRH_E_ARG, "trh_fitter_chain_state: null argument");
    int rc;
    if ((rc = hip_check(hipStreamSynchronize((hipStream_t)stream), "sync"))) return rc;
    if (f->n_chain == 0) return TRH_OK;
    return hip_check(hipMemcpy(state, f->ch_state, 4 * (size_t)f->n_chain, hipMemcpyDeviceToHost),
                     "read chain state");
}

extern "C" int trh_fitter_set_hybrid(trh_fitter *f, int K, const double *amp, const double *mean,
                                     const double *var, double inner_nsigma,
                                     double outer_nsigma) {
    if (!f || K < 0 || K > FIT_PSF_MAXK || (K > 0 && (!amp || !mean || !var)))
        return set_err(TRH_E_ARG, "trh_fitter_set_hybrid: K must be 0..%d with amp, mean, var",
                       FIT_PSF_MAXK);
    if (K > 0 && !(inner_nsigma > 0 && outer_nsigma > inner_nsigma))
        return set_err(TRH_E_ARG, "trh_fitter_set_hybrid: need 0 < inner_nsigma < outer_nsigma");
    int rc;
    if ((rc = hip_check(hipSetDevice(f->h->device), "hipSetDevice"))) return rc;
    // the captured iteration's launches change: capture again on next use
    drop_graph(f);
    if (K == 0) {
        f->hyb = nullptr;
        return TRH_OK;
    }
    FitHybrid H{};
    H.K = K;
    H.n1 = inner_nsigma;
    H.n2 = outer_nsigma;
    for (int k = 0; k < K; ++k) {
        H.amp[k] = amp[k];
        H.mx[k] = mean[2 * k];
        H.my[k] = mean[2 * k + 1];
        H.v0[k] = var[3 * k];
        H.v1[k] = var[3 * k + 1];
        H.v2[k] = var[3 * k + 2];
        if (!(var[3 * k] > 0 && var[3 * k] * var[3 * k + 2] - var[3 * k + 1] * var[3 * k + 1] > 0))
            return set_err(TRH_E_ARG, "trh_fitter_set_hybrid: PSF component %d variance is not "
                                      "positive definite", k);
    }
    const int S = f->S;
    if (!f->mog_jac) {
        std::vector<int32_t> kind(S), P(S);
        if ((rc = hip_check(hipMemcpy(kind.data(), f->kind, 4 * (size_t)S, hipMemcpyDeviceToHost),
                            "read kinds")) ||
            (rc = hip_check(hipMemcpy(P.data(), f->P, 4 * (size_t)S, hipMemcpyDeviceToHost),
                            "read sizes")))
            return rc;
        std::vector<int64_t> boff(S, 0);
        int64_t tot = 0;
        for (int s = 0; s < S; ++s) {
            boff[s] = tot;
            if (kind[s]) tot += (int64_t)NALPHA * P[s] * P[s];
        }
        // the chi^2 jobs carry their window's offset + 1 in an int32
        if (tot >= (int64_t)INT32_MAX)
            return set_err(TRH_E_UNSUPPORTED, "trh_fitter_set_hybrid: %lld line-search pixels "
                           "exceed 2^31 - 1", (long long)tot);
        FitMogJob *mj = nullptr, *ml = nullptr;
        FitHybrid *hy = nullptr;
        int64_t *bo = nullptr;
        float *lb = nullptr;
        if ((rc = falloc(f, &hy, 1)) || (rc = falloc(f, &mj, (size_t)f->nj * S)) ||
            (rc = falloc(f, &ml, (size_t)NALPHA * S)) || (rc = falloc(f, &bo, S)) ||
            (rc = falloc(f, &lb, tot > 0 ? tot : 1)))
            return rc;
        if ((rc = hip_check(hipMemcpy(bo, boff.data(), 8 * (size_t)S, hipMemcpyHostToDevice),
                            "window offsets")))
            return rc;
        f->mog_jac = mj;
        f->mog_ls = ml;
        f->boff = bo;
        f->ls_base = lb;
        f->hyb_dev = hy;
    }
    if ((rc = hip_check(hipMemcpy(f->hyb_dev, &H, sizeof(H), hipMemcpyHostToDevice), "hybrid PSF")))
        return rc;
    f->hyb = f->hyb_dev;
    return TRH_OK;
}

extern "C" int trh_fitter_render_models(trh_fitter *f, float *out, int accumulate,
                                        void *stream) {
    if (!f || !out) return set_err(TRH_E_ARG, "trh_fitter_render_models: null argument");
    hipStream_t st = (hipStream_t)stream;
    const int S = f->S;
    const int cap_jac = f->nj * S;
    int rc;
    if (f->n_ser && !f->ser)
        return set_err(TRH_E_ARG, "trh_fitter_render_models: Sersic galaxies need trh_fitter_set_sersic");
    hipLaunchKernelGGL(k_fit_reset, dim3(1), dim3(64), 0, st, f->cnt, 16);
    const int mode = accumulate ? TRH_ADD : TRH_STORE;
    hipLaunchKernelGGL(k_fit_prep_model, dim3((S + 255) / 256), dim3(256), 0, st, S, f->nj, f->kind, f->P,
                       f->stride, mode, f->pix_off, f->params, f->prof, f->ser, f->ph, f->pw,
                       f->jobs_jac, f->list_jac, cap_jac, f->cnt, f->flags, f->hyb, f->mog_jac);
    if ((rc = hip_check(hipGetLastError(), "k_fit_prep_model"))) return rc;
    for (int b = 0; b < NFFT && f->n_gal; ++b) {
        if ((rc = launch_fft_any(f->h, kSizes[b], accumulate ? -1 : TRH_STORE, f->jobs_jac,
                                 f->list_jac + (size_t)b * cap_jac, f->cnt + b, f->n_gal, f->maxrows[b], f->spec[b], out, nullptr,
                                 nullptr, nullptr, st, fitter_hbuf(f, kSizes[b]))))
            return rc;
    }
    if ((rc = launch_mog(f, f->mog_jac, f->list_jac, cap_jac, f->cnt, accumulate ? 2 : 1, out, 1, 0,
                         st)))
        return rc;
    return launch_point(f, f->jobs_jac, f->list_jac + (size_t)PTB * cap_jac, f->cnt + PTB, f->n_pt, out,
                        nullptr, st, mode);
}

extern "C" int trh_fitter_status(trh_fitter *f, int32_t *active, int32_t *iterations,
                                 double *chisq, double *dlnp, int32_t *n_active, void *stream) {
    if (!f) return set_err(TRH_E_ARG, "trh_fitter_status: null fitter");
    int rc;
    if ((rc = hip_check(hipStreamSynchronize((hipStream_t)stream), "sync"))) return rc;
    std::vector<int32_t> a(f->S);
    if ((rc = hip_check(hipMemcpy(a.data(), f->active, 4 * (size_t)f->S, hipMemcpyDeviceToHost), "read active")))
        return rc;
    if (active) std::copy(a.begin(), a.end(), active);
    int n = 0;
    for (int v : a) n += v ? 1 : 0;
    if (n_active) *n_active = n;
    f->last_active = n;
    if (iterations &&
        (rc = hip_check(hipMemcpy(iterations, f->iters, 4 * (size_t)f->S, hipMemcpyDeviceToHost), "read iters")))
        return rc;
    if (chisq && (rc = hip_check(hipMemcpy(chisq, f->chisq, 8 * (size_t)f->S, hipMemcpyDeviceToHost), "read chisq")))
        return rc;
    if (dlnp && (rc = hip_check(hipMemcpy(dlnp, f->dlnp, 8 * (size_t)f->S, hipMemcpyDeviceToHost), "read dlnp")))
        return rc;
    return TRH_OK;
}

// Number of active sources into out[0] (one block).
__global__ void __launch_bounds__(1024) k_fit_count_active(int S, const int32_t *__restrict__ active,
                                                           int32_t *__restrict__ out) {
    __shared__ int32_t part[16];
    int n = 0;
    for (int s = threadIdx.x; s < S; s += blockDim.x) n += active[s] ? 1 : 0;
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) n += __shfl_xor(n, d, 64);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = n;
    __syncthreads();
    if (threadIdx.x == 0) {
        int t = 0;
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t += part[w];
        out[0] = t;
    }
}

extern "C" int trh_fitter_count_active(trh_fitter *f, void *stream) {
    if (!f) return set_err(TRH_E_ARG, "trh_fitter_count_active: null fitter");
    int rc;
    // an unread earlier count is dropped (its copy finishes first: the
    // pinned word is written by one copy at a time)
    if (f->count_pending && (rc = hip_check(hipEventSynchronize(f->cev[1]), "count drop")))
        return rc;
    f->count_pending = false;
    if (!f->have_count) {
        if ((rc = falloc(f, &f->nact_dev, 1))) return rc;
        if ((rc = hip_check(hipHostMalloc((void **)&f->nact_pin, sizeof(int32_t)), "pinned count")))
            return rc;
        if ((rc = hip_check(hipStreamCreateWithFlags(&f->cstream, hipStreamNonBlocking),
                            "count stream"))) {
            (void)hipHostFree(f->nact_pin);
            return rc;
        }
        int ne = 0;
        while (!rc && ne < 2)
            if (!(rc = hip_check(hipEventCreateWithFlags(&f->cev[ne], hipEventDisableTiming),
                                 "count event")))
                ++ne;
        if (rc) {
            for (int i = 0; i < ne; ++i) (void)hipEventDestroy(f->cev[i]);
            (void)hipStreamDestroy(f->cstream);
            (void)hipHostFree(f->nact_pin);
            return rc;
        }
        f->have_count = true;
    }
    hipStream_t st = (hipStream_t)stream;
    // after the replays issued so far (trh_fitter_iterate_graph leaves
    // `stream` ordered after them; again here for a different stream)
    if (f->have_gev && (rc = hip_check(hipStreamWaitEvent(st, f->gev[1], 0), "count order")))
        return rc;
    hipLaunchKernelGGL(k_fit_count_active, dim3(1), dim3(1024), 0, st, f->S, f->active,
                       f->nact_dev);
    if ((rc = hip_check(hipGetLastError(), "k_fit_count_active")) ||
        (rc = hip_check(hipEventRecord(f->cev[0], st), "count event")) ||
        (rc = hip_check(hipStreamWaitEvent(f->cstream, f->cev[0], 0), "count wait")) ||
        (rc = hip_check(hipMemcpyAsync(f->nact_pin, f->nact_dev, sizeof(int32_t),
                                       hipMemcpyDeviceToHost, f->cstream), "count copy")) ||
        (rc = hip_check(hipEventRecord(f->cev[1], f->cstream), "count done")))
        return rc;
    f->count_pending = true;
    return TRH_OK;
}

extern "C" int trh_fitter_count_read(trh_fitter *f, int wait, int32_t *n_active) {
    if (!f || !n_active) return set_err(TRH_E_ARG, "trh_fitter_count_read: null argument");
    if (!f->count_pending)
        return set_err(TRH_E_ARG, "trh_fitter_count_read: no count enqueued");
    int rc;
    if (wait) {
        if ((rc = hip_check(hipEventSynchronize(f->cev[1]), "count sync"))) return rc;
    } else {
        const hipError_t q = hipEventQuery(f->cev[1]);
        if (q == hipErrorNotReady) {
            *n_active = -1;
            return TRH_OK;
        }
        if ((rc = hip_check(q, "count query"))) return rc;
    }
    *n_active = *f->nact_pin;
    f->last_active = *n_active;
    f->count_pending = false;
    return TRH_OK;
}

extern "C" int trh_fitter_flags(trh_fitter *f, int32_t *flags, void *stream) {
    if (!f || !flags) return set_err(TRH_E_ARG, "trh_fitter_flags: null argument");
    int rc;
    if ((rc = hip_check(hipStreamSynchronize((hipStream_t)stream), "sync"))) return rc;
    return hip_check(hipMemcpy(flags, f->flags, 4 * (size_t)f->S, hipMemcpyDeviceToHost),
                     "read flags");
}

extern "C" int trh_fitter_debug(trh_fitter *f, const float **renders, const int64_t **render_off,
                                const double **normal_eq, const double **step_x) {
    if (!f) return set_err(TRH_E_ARG, "trh_fitter_debug: null fitter");
    if (renders) *renders = f->renders;
    if (render_off) *render_off = f->render_off;
    if (normal_eq) *normal_eq = f->neq;
    if (step_x) *step_x = f->X;
    return TRH_OK;
}

extern "C" int trh_fitter_timing(trh_fitter *f, float *ms6) {
    if (!f || !ms6) return set_err(TRH_E_ARG, "trh_fitter_timing: null argument");
    int rc;
    if ((rc = hip_check(hipEventSynchronize(f->ev[5]), "event sync"))) return rc;
    for (int i = 0; i < 5; ++i) (void)hipEventElapsedTime(&ms6[i], f->ev[i], f->ev[i + 1]);
    (void)hipEventElapsedTime(&ms6[5], f->ev[0], f->ev[5]);
    return TRH_OK;
}

// ---------------------------------------------------------------------------
// The fitter's stages as standalone batched entry points
// ---------------------------------------------------------------------------
extern "C" int trh_jacobian_normal_eq_batch(trh_handle *h, int n, const int32_t *kind,
                                            const int32_t *P, const int32_t *row_stride,
                                            const int64_t *pix_off, const float *data,
                                            const float *inverr, const int64_t *render_off,
                                            const float *renders, const double *counts,
                                            const double *index_step, const double *cdi,
                                            double shape_step, double *neq, void *stream) {
    if (!h || n <= 0 || !kind || !P || !row_stride || !pix_off || !data || !inverr ||
        !render_off || !renders || !counts || !neq || !(shape_step != 0.0))
        return set_err(TRH_E_ARG, "trh_jacobian_normal_eq_batch: bad arguments");
    const double I2[4] = {1.0, 0.0, 0.0, 1.0};
    const double *c = cdi ? cdi : I2;
    hipStream_t st = (hipStream_t)stream;
    const int grid = n < h->n_cu * 8 ? n : h->n_cu * 8;
    hipLaunchKernelGGL(k_fit_normal_eq<6>, dim3(grid), dim3(NEQ_TB), 0, st, n, kind, P, row_stride,
                       pix_off, render_off, data, inverr, renders, counts, 1, shape_step,
                       index_step, nullptr, c[0], c[1], c[2], c[3], neq, nullptr);
    int rc = hip_check(hipGetLastError(), "k_fit_normal_eq<6>");
    if (rc || !index_step) return rc;
    hipLaunchKernelGGL(k_fit_normal_eq<7>, dim3(grid), dim3(NEQ_TB), 0, st, n, kind, P, row_stride,
                       pix_off, render_off, data, inverr, renders, counts, 1, shape_step,
                       index_step, nullptr, c[0], c[1], c[2], c[3], neq, nullptr);
    return hip_check(hipGetLastError(), "k_fit_normal_eq<7>");
}

extern "C" int trh_solve_damped_cholesky_f64_batch(trh_handle *h, int n, const int32_t *kind,
                                                   const double *neq, double lambda,
                                                   const double *params, double *X,
                                                   double *trial, int32_t *nsteps,
                                                   int32_t *steplim, void *stream) {
    if (!h || n <= 0 || !kind || !neq || !X || !(lambda >= 0.0) ||
        (trial && (!params || !nsteps || !steplim)))
        return set_err(TRH_E_ARG, "trh_solve_damped_cholesky_f64_batch: bad arguments");
    hipLaunchKernelGGL(k_fit_solve, dim3((n + 63) / 64), dim3(64), 0, (hipStream_t)stream, n, kind,
                       neq, params, nullptr, lambda, X, trial, trial ? nsteps : nullptr,
                       trial ? steplim : nullptr);
    return hip_check(hipGetLastError(), "k_fit_solve");
}

extern "C" int trh_line_search_chi2_batch(trh_fitter *f, const double *trial,
                                          const int32_t *nsteps, double *chi2, void *stream) {
    if (!f || !trial || !nsteps || !chi2)
        return set_err(TRH_E_ARG, "trh_line_search_chi2_batch: bad arguments");
    if (f->n_ser && !f->ser)
        return set_err(TRH_E_ARG, "trh_line_search_chi2_batch: Sersic galaxies need "
                                  "trh_fitter_set_sersic");
    hipStream_t st = (hipStream_t)stream;
    const int S = f->S, cap_ls = NALPHA * S;
    const int nb_ls = (cap_ls + PREP_TB - 1) / PREP_TB;
    int rc;
    hipLaunchKernelGGL(k_fit_reset, dim3(1), dim3(64), 0, st, f->cnt, 16);
    hipLaunchKernelGGL(k_fit_prep_ls, dim3(nb_ls), dim3(PREP_TB), 0, st, S, f->kind, f->P,
                       f->stride, f->pix_off, trial, nsteps, nullptr, f->prof, f->ser, f->ph,
                       f->pw, f->jobs_ls, f->bk_ls, f->bcnt_ls, f->flags, chi2, nullptr, f->hyb,
                       f->mog_ls, f->boff);
    if ((rc = hip_check(hipGetLastError(), "k_fit_prep_ls"))) return rc;
    hipLaunchKernelGGL(k_fit_lists, dim3(nb_ls), dim3(PREP_TB), 0, st, cap_ls, f->bk_ls,
                       f->bcnt_ls, f->list_ls, cap_ls, f->cnt + LS_CNT);
    if ((rc = hip_check(hipGetLastError(), "k_fit_lists"))) return rc;
    if ((rc = launch_mog(f, f->mog_ls, f->list_ls, cap_ls, f->cnt + LS_CNT, 0, f->ls_base, 1, 0, st)))
        return rc;
    for (int b = 0; b < NFFT && f->n_gal; ++b) {
        if (kSizes[b] < f->min_sz) continue;
        if ((rc = launch_fft_any(f->h, kSizes[b], TRH_CHI2, f->jobs_ls,
                                 f->list_ls + (size_t)b * cap_ls, f->cnt + LS_CNT + b,
                                 NALPHA * f->n_gal, f->maxrows[b], f->spec[b], nullptr, f->data,
                                 f->inverr, chi2, st, fitter_hbuf(f, kSizes[b]),
                                 f->hyb ? f->ls_base : nullptr)))
            return rc;
    }
    // point sources over their whole windows (no normal equations assumed)
    return launch_point(f, f->jobs_ls, f->list_ls + (size_t)PTB * cap_ls, f->cnt + LS_CNT + PTB,
                        NALPHA * f->n_pt, nullptr, chi2, st, TRH_STORE, false);
}
