// libtractor_hip: C ABI over the MI355X render kernels (handle, PSF spectra,
// batched FFT renders, pixel-space MoG evaluation, Lanczos shifts).
// The batched fitter lives in fitter.hip.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string>

#include "../../include/tractor_hip.h"
#include "internal.h"
#include "render_fft.h"

using namespace trh;

static thread_local std::string g_err;

// Scratch of the allocation-free Lanczos entry point (the x-pass
// intermediate): TRH_NOALLOC_SCRATCH_FLOATS floats, one copy per GPU, held
// until the process ends.  Host code never uses this name as an address:
// trh_create resolves the device address with hipGetSymbolAddress.
static __device__ __align__(256) float scratch[TRH_NOALLOC_SCRATCH_FLOATS];

namespace trh {
int set_err(int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

int hip_check(hipError_t e, const char *what) {
    if (e == hipSuccess) return TRH_OK;
    return set_err(TRH_E_HIP, "%s: %s", what, hipGetErrorString(e));
}

int tw_offset(int n) {
    // tables for N = 32, 64, 128, 256, 512 stored back to back (N/2 entries)
    switch (n) {
    case 32: return 0;
    case 64: return 16;
    case 128: return 48;
    case 256: return 112;
    case 512: return 240;
    }
    return -1;
}

int pad_center(int sz, int p) {
    // PixelizedPSF._padInImage (tractor/psf.py:225-238)
    if (sz >= p) return (sz - p) / 2 + p / 2;
    return p / 2 - (p - sz) / 2;
}

int cu_count(trh_handle *h) { return h->n_cu; }

static bool capturing(hipStream_t st) {
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    return hipStreamIsCapturing(st, &cs) == hipSuccess && cs != hipStreamCaptureStatusNone;
}

int shared_begin(trh_handle *h, hipStream_t st) {
    if (!h->order_recorded || st == h->order_stream || capturing(st)) return TRH_OK;
    return hip_check(hipStreamWaitEvent(st, h->order_ev, 0), "order shared buffers");
}

int shared_end(trh_handle *h, hipStream_t st) {
    if (capturing(st)) return TRH_OK;
    int rc = hip_check(hipEventRecord(h->order_ev, st), "record shared-buffer use");
    if (rc) return rc;
    h->order_recorded = true;
    h->order_stream = st;
    return TRH_OK;
}
}  // namespace trh

extern "C" const char *trh_last_error(void) { return g_err.c_str(); }
extern "C" int trh_version(void) { return 2; }

extern "C" int trh_struct_sizes(int *fft_job, int *mog_job, int *mog_approx_job) {
    if (fft_job) *fft_job = (int)sizeof(trh_fft_job);
    if (mog_job) *mog_job = (int)sizeof(trh_mog_job);
    if (mog_approx_job) *mog_approx_job = (int)sizeof(trh_mog_approx_job);
    return TRH_OK;
}

extern "C" int trh_copy_to_host(void *dst, const void *src, uint64_t bytes, void *stream) {
    if (!dst || !src) return set_err(TRH_E_ARG, "trh_copy_to_host: null pointer");
    int rc = hip_check(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, (hipStream_t)stream),
                       "hipMemcpyAsync D2H");
    if (rc) return rc;
    return hip_check(hipStreamSynchronize((hipStream_t)stream), "hipStreamSynchronize");
}

__global__ void k_twiddles(float2 *tw) {
    // exp(+2 pi i m / N) in FP64, for every table
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int Ns[5] = {32, 64, 128, 256, 512};
    const int offs[5] = {0, 16, 48, 112, 240};
    for (int q = 0; q < 5; ++q) {
        const int N = Ns[q];
        if (t < N / 2) {
            double s, c;
            sincospi(2.0 * t / N, &s, &c);
            tw[offs[q] + t] = make_float2((float)c, (float)s);
        }
    }
}

extern "C" int trh_create(trh_handle **hp, int device) {
    if (!hp) return set_err(TRH_E_ARG, "trh_create: null handle pointer");
    int rc;
    if ((rc = hip_check(hipSetDevice(device), "hipSetDevice"))) return rc;
    trh_handle *h = new trh_handle();
    h->device = device;
    hipDeviceProp_t prop;
    if ((rc = hip_check(hipGetDeviceProperties(&prop, device), "hipGetDeviceProperties"))) {
        delete h;
        return rc;
    }
    h->n_cu = prop.multiProcessorCount;
    h->lds_max = (int)prop.sharedMemPerBlock;
    if ((rc = hip_check(hipMalloc(&h->tw, sizeof(float2) * 496), "hipMalloc twiddles"))) {
        delete h;
        return rc;
    }
    if ((rc = hip_check(hipEventCreateWithFlags(&h->order_ev, hipEventDisableTiming),
                        "order event"))) {
        (void)hipFree(h->tw);
        delete h;
        return rc;
    }
    // this device's copy of the allocation-free entry point's scratch
    if ((rc = hip_check(hipGetSymbolAddress(&h->noalloc_scratch, HIP_SYMBOL(scratch)),
                        "hipGetSymbolAddress(scratch)"))) {
        (void)hipEventDestroy(h->order_ev);
        (void)hipFree(h->tw);
        delete h;
        return rc;
    }
    hipLaunchKernelGGL(k_twiddles, dim3(1), dim3(256), 0, 0, h->tw);
    if ((rc = hip_check(hipGetLastError(), "k_twiddles"))) return rc;
    if ((rc = hip_check(hipDeviceSynchronize(), "twiddle sync"))) return rc;
    *hp = h;
    return TRH_OK;
}

extern "C" int trh_destroy(trh_handle *h) {
    if (!h) return TRH_OK;
    (void)hipSetDevice(h->device);
    if (h->tw) (void)hipFree(h->tw);
    if (h->order_ev) (void)hipEventDestroy(h->order_ev);
    if (h->scratch) (void)hipFree(h->scratch);
    for (int i = 0; i < 5; ++i)
        if (h->hbuf[i]) (void)hipFree(h->hbuf[i]);
    delete h;
    return TRH_OK;
}

namespace trh {
int ensure_scratch(trh_handle *h, size_t bytes) {
    if (h->scratch_bytes >= bytes) return TRH_OK;
    if (h->scratch) (void)hipFree(h->scratch);
    h->scratch = nullptr;
    h->scratch_bytes = 0;
    if (hipMalloc(&h->scratch, bytes) != hipSuccess)
        return set_err(TRH_E_NOMEM, "scratch allocation of %zu bytes failed", bytes);
    h->scratch_bytes = bytes;
    return TRH_OK;
}
}  // namespace trh

// ---------------------------------------------------------------------------
// PSF spectra: rfft2 of the padded stamp (psf.py:249-280), rows then columns.
// ---------------------------------------------------------------------------
template <int SZ>
__global__ void __launch_bounds__(256)
k_psf_rows(const float *__restrict__ stamps, int ph, int pw, float2 *__restrict__ out,
           const float2 *__restrict__ twg, int rp) {
    constexpr int NV = SZ / 2 + 1;
    constexpr int LOG2 = Log2<SZ>::v;
    extern __shared__ float2 smem[];
    float2 *tw = smem;
    float2 *work = tw + SZ / 2;
    for (int i = threadIdx.x; i < SZ / 2; i += blockDim.x) tw[i] = twg[i];
    const int n = blockIdx.y;
    const int pair0 = blockIdx.x * rp;
    const int np = min(rp, SZ / 2 - pair0);
    const float *st = stamps + (size_t)n * ph * pw;
    const int y0 = SZ >= ph ? (SZ - ph) / 2 : -((ph - SZ) / 2);
    const int x0 = SZ >= pw ? (SZ - pw) / 2 : -((pw - SZ) / 2);
    for (int idx = threadIdx.x; idx < np * SZ; idx += blockDim.x) {
        const int pp = idx >> LOG2;
        const int x = idx & (SZ - 1);
        const int ya = 2 * (pair0 + pp);
        float v[2];
        for (int r = 0; r < 2; ++r) {
            const int sy = ya + r - y0, sx = x - x0;
            v[r] = (sy >= 0 && sy < ph && sx >= 0 && sx < pw) ? st[sy * pw + sx] : 0.f;
        }
        work[pp * SZ + bitrev<LOG2>(x)] = make_float2(v[0], v[1]);
    }
    __syncthreads();
    fft_dit<SZ, false>(work, np, tw);
    for (int idx = threadIdx.x; idx < np * NV; idx += blockDim.x) {
        const int pp = idx / NV;
        const int k = idx - pp * NV;
        const float2 zk = work[pp * SZ + (k & (SZ - 1))];
        const float2 zn = work[pp * SZ + ((SZ - k) & (SZ - 1))];
        const float2 A = make_float2(0.5f * (zk.x + zn.x), 0.5f * (zk.y - zn.y));
        const float2 B = make_float2(0.5f * (zk.y + zn.y), -0.5f * (zk.x - zn.x));
        const int ya = 2 * (pair0 + pp);
        out[((size_t)n * NV + k) * SZ + ya] = A;
        out[((size_t)n * NV + k) * SZ + ya + 1] = B;
    }
}

template <int SZ>
__global__ void __launch_bounds__(256)
k_psf_cols(float2 *__restrict__ spec, const float2 *__restrict__ twg, int cc) {
    constexpr int NV = SZ / 2 + 1;
    constexpr int LOG2 = Log2<SZ>::v;
    extern __shared__ float2 smem[];
    float2 *tw = smem;
    float2 *work = tw + SZ / 2;
    for (int i = threadIdx.x; i < SZ / 2; i += blockDim.x) tw[i] = twg[i];
    const int n = blockIdx.y;
    const int c0 = blockIdx.x * cc;
    const int nc = min(cc, NV - c0);
    float2 *S = spec + (size_t)n * SZ * NV;
    for (int idx = threadIdx.x; idx < nc * SZ; idx += blockDim.x) {
        const int c = idx >> LOG2;
        const int j = idx & (SZ - 1);
        work[c * SZ + bitrev<LOG2>(j)] = S[(size_t)(c0 + c) * SZ + j];
    }
    __syncthreads();
    fft_dit<SZ, false>(work, nc, tw);
    for (int idx = threadIdx.x; idx < nc * SZ; idx += blockDim.x) {
        const int c = idx >> LOG2;
        const int j = idx & (SZ - 1);
        S[(size_t)(c0 + c) * SZ + j] = work[c * SZ + j];
    }
}

template <int SZ>
static int psf_spectrum_sz(trh_handle *h, const float *stamps, int n, int ph, int pw, float *out,
                           hipStream_t st) {
    const int rp = SZ >= 256 ? 8 : 16;
    const int cc = SZ >= 256 ? 8 : 16;
    const float2 *tw = h->tw + tw_offset(SZ);
    size_t lds = sizeof(float2) * (SZ / 2 + (size_t)rp * SZ);
    hipLaunchKernelGGL(k_psf_rows<SZ>, dim3((SZ / 2 + rp - 1) / rp, n), dim3(256), lds, st,
                       stamps, ph, pw, (float2 *)out, tw, rp);
    int rc = hip_check(hipGetLastError(), "k_psf_rows");
    if (rc) return rc;
    lds = sizeof(float2) * (SZ / 2 + (size_t)cc * SZ);
    hipLaunchKernelGGL(k_psf_cols<SZ>, dim3((SZ / 2 + 1 + cc - 1) / cc, n), dim3(256), lds, st,
                       (float2 *)out, tw, cc);
    return hip_check(hipGetLastError(), "k_psf_cols");
}

extern "C" int trh_psf_spectrum_batch(trh_handle *h, const float *stamps, int n, int ph, int pw,
                                      int sz, float *out, int *cx, int *cy, void *stream) {
    if (!h || !stamps || !out || n <= 0 || ph <= 0 || pw <= 0)
        return set_err(TRH_E_ARG, "trh_psf_spectrum_batch: bad arguments");
    if ((ph % 2) == 0 || (pw % 2) == 0)
        return set_err(TRH_E_ARG, "trh_psf_spectrum_batch: PSF stamp must have odd size (psf.py:91-92)");
    hipStream_t st = (hipStream_t)stream;
    if (cx) *cx = pad_center(sz, pw);
    if (cy) *cy = pad_center(sz, ph);
    switch (sz) {
    case 32: return psf_spectrum_sz<32>(h, stamps, n, ph, pw, out, st);
    case 64: return psf_spectrum_sz<64>(h, stamps, n, ph, pw, out, st);
    case 128: return psf_spectrum_sz<128>(h, stamps, n, ph, pw, out, st);
    case 256: return psf_spectrum_sz<256>(h, stamps, n, ph, pw, out, st);
    case 512: return psf_spectrum_sz<512>(h, stamps, n, ph, pw, out, st);
    }
    return set_err(TRH_E_UNSUPPORTED, "trh_psf_spectrum_batch: sz=%d not in {32..512}", sz);
}

// ---------------------------------------------------------------------------
// Batched FFT renders
// ---------------------------------------------------------------------------
namespace trh {
static int sz_index(int sz) {
    switch (sz) {
    case 32: return 0;
    case 64: return 1;
    case 128: return 2;
    case 256: return 3;
    case 512: return 4;
    }
    return -1;
}

// Default launch shape per FFT size: threads (waves x 64) per block; the
// second knob caps the G rows kept per block (0: as many as LDS allows).
// trh_set_fft_config overrides them (the autotuner's knob).
// (from a sweep over threads x rows on MI355X; sz 128 keeps 120 rows so two
// blocks share a CU's 160 KB of LDS; from sz 256 on the rows are in global
// slots and the row knob is unused)
static const int kDefThreads[5] = {256, 128, 512, 512, 1024};
static const int kDefRows[5] = {0, 0, 120, 0, 0};

// Global row slots of the SZ >= 256 renders (one per block of the grid),
// as many per CU as blocks of the default shape are resident: 2 at SZ = 256
// (512 threads; 136 MB), 1 at SZ = 512 (1024 threads; 270 MB).  With no
// more slots than resident blocks the rows stay warmer in the caches (a
// sz = 256 chi^2 render 185 -> 179 ns against 4 per CU; 1 per CU: 205 ns).
int fft_hbuf_slots(const trh_handle *h, int sz) { return h->n_cu * (sz >= 512 ? 1 : 2); }

void fft_launch_shape(trh_handle *h, int sz, int maxrows, int *threads, int *rows, size_t *lds,
                      int *grid, int njobs) {
    const int si = sz_index(sz);
    int nt = h->fft_threads[si] > 0 ? h->fft_threads[si] : kDefThreads[si];
    int r = h->fft_ch[si] > 0 ? h->fft_ch[si] : (kDefRows[si] > 0 ? kDefRows[si] : sz);
    if (maxrows > 0 && r > maxrows) r = maxrows;
    if (r > sz) r = sz;
    const bool gh = sz >= 256;   // rows in global slots, all SZ of them per pass
    if (gh) r = sz;
    size_t b = fft_lds_bytes(sz, nt / 64, gh ? 0 : r);
    // fit LDS: fewer rows first (down to 32), then fewer waves, then rows to 8
    while (b + sizeof(RenderScratch) > (size_t)h->lds_max) {
        if (!gh && r > 32) r -= 8;
        else if (nt > 128) nt /= 2;
        else if (!gh && r > 8) r -= 8;
        else break;
        b = fft_lds_bytes(sz, nt / 64, gh ? 0 : r);
    }
    *rows = r;
    *lds = b;
    *threads = nt;
    const size_t per_cu = 160 * 1024;
    int wg_per_cu = (int)(per_cu / (b + sizeof(RenderScratch)));
    const int by_threads = 2048 / nt;    // 32 waves per CU
    if (wg_per_cu > by_threads) wg_per_cu = by_threads;
    if (wg_per_cu < 1) wg_per_cu = 1;
    int g = h->n_cu * wg_per_cu;
    if (gh && g > fft_hbuf_slots(h, sz)) g = fft_hbuf_slots(h, sz);
    if (njobs > 0 && njobs < g) g = njobs;
    *grid = g < 1 ? 1 : g;
}

int reserve_fft_hbuf(trh_handle *h, int sz) {
    if (sz < 256) return TRH_OK;
    const int si = sz_index(sz);
    if (si < 0) return set_err(TRH_E_ARG, "reserve_fft_hbuf: sz %d", sz);
    if (h->hbuf[si]) return TRH_OK;
    const size_t bytes = sizeof(float) * fft_hbuf_slot_floats(sz) * (size_t)fft_hbuf_slots(h, sz);
    if (hipMalloc(&h->hbuf[si], bytes) != hipSuccess) {
        h->hbuf[si] = nullptr;
        return set_err(TRH_E_NOMEM, "k_render_fft<%d>: row slots of %zu bytes", sz, bytes);
    }
    return TRH_OK;
}

int max_rows_for(int sz, int lds_max) {
    // upper bound on the G rows per block; fft_launch_shape fits it to LDS
    (void)lds_max;
    return sz;
}

template <int SZ>
int launch_fft(trh_handle *h, int mode, const trh_fft_job *jobs, const int *list,
               const int *count, int njobs_bound, int maxrows, const float *psf_spec, float *out,
               const float *data, const float *inverr, double *chi2, hipStream_t st,
               float2 *hbuf, const float *base) {
    int nt, rows, grid;
    size_t lds;
    fft_launch_shape(h, SZ, maxrows, &nt, &rows, &lds, &grid, njobs_bound);
    if (lds + sizeof(RenderScratch) > (size_t)h->lds_max)
        return set_err(TRH_E_UNSUPPORTED, "k_render_fft<%d>: LDS %zu exceeds %d", SZ, lds,
                       h->lds_max);
    float2 *hb = hbuf;
    if constexpr (WaveGeom<SZ>::GH) {
        const int si = sz_index(SZ);
        if (!hb && !h->hbuf[si]) {
            // the slots are allocated once, outside any stream capture
            hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
            if (hipStreamIsCapturing(st, &cs) == hipSuccess && cs != hipStreamCaptureStatusNone)
                return set_err(TRH_E_UNSUPPORTED,
                               "k_render_fft<%d>: row slots not reserved before capture", SZ);
            const int rc = reserve_fft_hbuf(h, SZ);
            if (rc) return rc;
        }
        if (!hb) hb = h->hbuf[si];
    }
    const float2 *tw = h->tw + tw_offset(SZ);
    const float2 *ps = (const float2 *)psf_spec;
    switch (mode) {
    case TRH_STORE:
        hipLaunchKernelGGL((k_render_fft<SZ, TRH_STORE>), dim3(grid), dim3(nt), lds, st, jobs, list,
                           count, njobs_bound, ps, tw, rows, out, data, inverr, chi2, hb, nullptr);
        break;
    case TRH_CHI2:
        if (base)
            hipLaunchKernelGGL((k_render_fft<SZ, TRH_CHI2_BASE>), dim3(grid), dim3(nt), lds, st,
                               jobs, list, count, njobs_bound, ps, tw, rows, out, data, inverr,
                               chi2, hb, base);
        else
            hipLaunchKernelGGL((k_render_fft<SZ, TRH_CHI2>), dim3(grid), dim3(nt), lds, st, jobs,
                               list, count, njobs_bound, ps, tw, rows, out, data, inverr, chi2,
                               hb, nullptr);
        break;
    default:
        hipLaunchKernelGGL((k_render_fft<SZ, -1>), dim3(grid), dim3(nt), lds, st, jobs, list, count,
                           njobs_bound, ps, tw, rows, out, data, inverr, chi2, hb, nullptr);
        break;
    }
    return hip_check(hipGetLastError(), "k_render_fft");
}

int launch_fft_any(trh_handle *h, int sz, int mode, const trh_fft_job *jobs, const int *list,
                   const int *count, int njobs_bound, int maxrows, const float *psf_spec,
                   float *out, const float *data, const float *inverr, double *chi2,
                   hipStream_t st, float2 *hbuf, const float *base) {
    if (base && mode != TRH_CHI2)
        return set_err(TRH_E_ARG, "launch_fft_any: a base plane needs chi^2 mode");
    switch (sz) {
    case 32:
        return launch_fft<32>(h, mode, jobs, list, count, njobs_bound, maxrows, psf_spec, out,
                              data, inverr, chi2, st, hbuf, base);
    case 64:
        return launch_fft<64>(h, mode, jobs, list, count, njobs_bound, maxrows, psf_spec, out,
                              data, inverr, chi2, st, hbuf, base);
    case 128:
        return launch_fft<128>(h, mode, jobs, list, count, njobs_bound, maxrows, psf_spec, out,
                               data, inverr, chi2, st, hbuf, base);
    case 256:
        return launch_fft<256>(h, mode, jobs, list, count, njobs_bound, maxrows, psf_spec, out,
                               data, inverr, chi2, st, hbuf, base);
    case 512:
        return launch_fft<512>(h, mode, jobs, list, count, njobs_bound, maxrows, psf_spec, out,
                               data, inverr, chi2, st, hbuf, base);
    }
    return set_err(TRH_E_UNSUPPORTED, "FFT render size %d not in {32,64,128,256,512}", sz);
}
}  // namespace trh

extern "C" int trh_set_fft_config(trh_handle *h, int sz, int threads, int rows) {
    if (!h) return set_err(TRH_E_ARG, "trh_set_fft_config: null handle");
    const int si = sz_index(sz);
    if (si < 0) return set_err(TRH_E_ARG, "trh_set_fft_config: sz %d not in {32..512}", sz);
    if (threads != 0 && (threads % 64 != 0 || threads < 64 || threads > 1024))
        return set_err(TRH_E_ARG, "trh_set_fft_config: threads must be a multiple of 64 in [64, 1024]");
    if (rows < 0) return set_err(TRH_E_ARG, "trh_set_fft_config: rows must be >= 0");
    h->fft_threads[si] = threads;
    h->fft_ch[si] = rows;
    return TRH_OK;
}

extern "C" int trh_render_fft_batch(trh_handle *h, int sz, const trh_fft_job *jobs, int njobs,
                                    const float *psf_spec, float *out, const float *data,
                                    const float *inverr, double *chi2, void *stream) {
    if (!h || (!jobs && njobs > 0) || !psf_spec)
        return set_err(TRH_E_ARG, "trh_render_fft_batch: null argument");
    if (njobs <= 0) return TRH_OK;
    const int maxrows = max_rows_for(sz, h->lds_max);
    // no output image but chi^2 slots: every job is a chi^2 job (a STORE or
    // ADD job would need `out`), so the chi^2-specialised kernel runs
    const int mode = (!out && chi2 && data && inverr) ? TRH_CHI2 : -1;
    hipStream_t st = (hipStream_t)stream;
    int rc;
    if (sz >= 256 && (rc = shared_begin(h, st))) return rc;
    if ((rc = launch_fft_any(h, sz, mode, jobs, nullptr, nullptr, njobs, maxrows, psf_spec, out,
                             data, inverr, chi2, st)))
        return rc;
    return sz >= 256 ? shared_end(h, st) : TRH_OK;
}

// ---------------------------------------------------------------------------
// Pixel-space MoG (c_gauss_2d_grid / c_gauss_2d_masked)
// ---------------------------------------------------------------------------
#define TRH_MOG_CHUNK 128

__global__ void __launch_bounds__(256)
k_render_mog(const trh_mog_job *__restrict__ jobs, int njobs, const float *__restrict__ amp,
             const int32_t *__restrict__ cxi, const int32_t *__restrict__ cyi,
             const float *__restrict__ cxf, const float *__restrict__ cyf,
             const float *__restrict__ var, const uint8_t *__restrict__ mask,
             float *__restrict__ out, float *__restrict__ dxo, float *__restrict__ dyo) {
    // per component: scale, e0, e1, e2 (exponent = e0 dx^2 + e1 dx dy + e2 dy^2)
    __shared__ float4 C[TRH_MOG_CHUNK];
    __shared__ int2 CI[TRH_MOG_CHUNK];
    __shared__ float2 CF[TRH_MOG_CHUNK];
    const int jid = blockIdx.y;
    if (jid >= njobs) return;
    const trh_mog_job J = jobs[jid];
    const int npix = J.w * J.h;
    const int pix = blockIdx.x * blockDim.x + threadIdx.x;
    const bool live = pix < npix;
    const int oy = live ? pix / J.w : 0;
    const int ox = live ? pix - oy * J.w : 0;
    const int64_t idx = J.out_offset + (int64_t)oy * J.out_stride + ox;
    bool on = live;
    if (live && J.mask_offset >= 0 && mask) on = mask[J.mask_offset + (int64_t)oy * J.out_stride + ox] != 0;
    const float cut = J.cutoff_mode == 0 ? -50.f : -30.f;
    float r = 0.f, xd = 0.f, yd = 0.f;
    for (int c0 = 0; c0 < J.K; c0 += TRH_MOG_CHUNK) {
        const int nc = min(TRH_MOG_CHUNK, J.K - c0);
        __syncthreads();
        if (threadIdx.x < nc) {
            const int c = J.comp0 + c0 + threadIdx.x;
            float4 q;
            if (J.cutoff_mode == 0) {
                // c_gauss_2d_grid constants in FP64 (mix.i:482-491)
                const double V0 = var[3 * c], V1 = var[3 * c + 1], V2 = var[3 * c + 2];
                const double det = V0 * V2 - V1 * V1;
                q.x = (float)(amp[c] / sqrt(4.0 * M_PI * M_PI * det));
                q.y = (float)(-0.5 * V2 / det);
                q.z = (float)(V1 / det);
                q.w = (float)(-0.5 * V0 / det);
                if (!(det > 0.0)) q = make_float4(0.f, -1.f, 0.f, -1.f);
            } else {
                // c_gauss_2d_masked constants in FP32 (gauss_masked.c:80-103)
                const float V0 = var[3 * c], V1 = var[3 * c + 1], V2 = var[3 * c + 2];
                const float det = V0 * V2 - V1 * V1;
                const float isc = -0.5f / det;
                q.y = V2 * isc;
                q.z = -V1 * isc * 2.0f;
                q.w = V0 * isc;
                q.x = amp[c] / sqrtf((float)(4.0 * M_PI * M_PI) * det);
                if (!(isfinite(V0) && isfinite(V1) && isfinite(V2) && isfinite(q.y) &&
                      isfinite(q.z) && isfinite(q.w) && isfinite(q.x)))
                    q = make_float4(0.f, -1.f, -1.f, -1.f);
            }
            C[threadIdx.x] = q;
            CI[threadIdx.x] = make_int2(cxi[c], cyi[c]);
            CF[threadIdx.x] = make_float2(cxf[c], cyf[c]);
        }
        __syncthreads();
        if (on) {
            const int gx = J.x0 + ox, gy = J.y0 + oy;
            for (int k = 0; k < nc; ++k) {
                const float4 q = C[k];
                if (q.x == 0.f) continue;
                const float dx = (float)(gx - CI[k].x) - CF[k].x;
                const float dy = (float)(gy - CI[k].y) - CF[k].y;
                const float e = q.y * dx * dx + q.z * dx * dy + q.w * dy * dy;
                const bool keep = J.cutoff_mode == 0 ? (e > cut) : (e >= cut);
                if (!keep || !isfinite(e)) continue;
                const float G = q.x * __expf(e);
                r += G;
                if (J.want_derivs) {
                    xd += -G * (2.f * q.y * dx + q.z * dy);
                    yd += -G * (2.f * q.w * dy + q.z * dx);
                }
            }
        }
    }
    if (!live) return;
    if (J.mode == TRH_STORE) {
        out[idx] = J.scale * r;
        if (J.want_derivs) {
            if (dxo) dxo[idx] = J.scale * xd;
            if (dyo) dyo[idx] = J.scale * yd;
        }
    } else if (on) {
        atomicAdd(out + idx, J.scale * r);
        if (J.want_derivs) {
            if (dxo) atomicAdd(dxo + idx, J.scale * xd);
            if (dyo) atomicAdd(dyo + idx, J.scale * yd);
        }
    }
}

// The caller passes the largest window pixel count so the grid can be shaped
// without reading jobs back.
extern "C" int trh_render_mog_batch(trh_handle *h, const trh_mog_job *jobs, int njobs,
                                    int max_pixels, const float *amp, const int32_t *cxi,
                                    const int32_t *cyi, const float *cxf, const float *cyf,
                                    const float *var, const uint8_t *mask, float *out,
                                    float *dx_out, float *dy_out, void *stream) {
    if (!h || (njobs > 0 && (!jobs || !amp || !cxi || !cyi || !cxf || !cyf || !var || !out)))
        return set_err(TRH_E_ARG, "trh_render_mog_batch: null argument");
    if (njobs <= 0 || max_pixels <= 0) return TRH_OK;
    if (njobs > 65535) return set_err(TRH_E_ARG, "trh_render_mog_batch: at most 65535 jobs per call");
    dim3 grid((max_pixels + 255) / 256, njobs);
    hipLaunchKernelGGL(k_render_mog, grid, dim3(256), 0, (hipStream_t)stream, jobs, njobs, amp,
                       cxi, cyi, cxf, cyf, var, mask, out, dx_out, dy_out);
    return hip_check(hipGetLastError(), "k_render_mog");
}

// ---------------------------------------------------------------------------
// Lanczos-3 shift (lanczos_shift_image): x pass into scratch, y pass into out.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256)
k_lanczos_pass(const float *__restrict__ in, float *__restrict__ out, int height, int width,
               const double *__restrict__ shifts, int along_y) {
    __shared__ RenderScratch S;
    const int n = blockIdx.y;
    const double d = shifts[n];
    compute_taps(along_y ? 0.0 : d, along_y ? d : 0.0, S);
    const float *f = along_y ? S.taps + 7 : S.taps;
    const size_t base = (size_t)n * height * width;
    for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < height * width;
         p += gridDim.x * blockDim.x) {
        const int y = p / width, x = p - y * width;
        float a = 0.f;
#pragma unroll
        for (int t = 0; t < 7; ++t) {
            const int yy = along_y ? y - 3 + t : y;
            const int xx = along_y ? x : x - 3 + t;
            if (yy >= 0 && yy < height && xx >= 0 && xx < width)
                a += f[t] * in[base + (size_t)yy * width + xx];
        }
        out[base + p] = a;
    }
}

// x pass into tmp (n * height * width floats), y pass into out
static int lanczos_passes(trh_handle *h, const float *in, float *tmp, float *out, int n,
                          int height, int width, const double *dx, const double *dy,
                          hipStream_t st) {
    (void)h;
    const int gx = (height * width + 255) / 256 < 1024 ? (height * width + 255) / 256 : 1024;
    hipLaunchKernelGGL(k_lanczos_pass, dim3(gx, n), dim3(256), 0, st, in, tmp, height, width, dx, 0);
    int rc = hip_check(hipGetLastError(), "k_lanczos_pass x");
    if (rc) return rc;
    hipLaunchKernelGGL(k_lanczos_pass, dim3(gx, n), dim3(256), 0, st, tmp, out, height, width, dy, 1);
    return hip_check(hipGetLastError(), "k_lanczos_pass y");
}

extern "C" int trh_lanczos_shift_batch(trh_handle *h, const float *in, float *out, int n,
                                       int height, int width, const double *dx, const double *dy,
                                       void *stream) {
    if (!h || !in || !out || !dx || !dy || n <= 0 || height <= 0 || width <= 0)
        return set_err(TRH_E_ARG, "trh_lanczos_shift_batch: bad arguments");
    const size_t bytes = sizeof(float) * (size_t)n * height * width;
    hipStream_t st = (hipStream_t)stream;
    int rc = shared_begin(h, st);
    if (rc) return rc;
    if ((rc = ensure_scratch(h, bytes))) return rc;
    if ((rc = lanczos_passes(h, in, (float *)h->scratch, out, n, height, width, dx, dy, st)))
        return rc;
    return shared_end(h, st);
}

// The same shift with no allocation: the intermediate goes to this device's
// static scratch array (through the address trh_create resolved), so the
// call makes no hipMalloc / hipFree.  Calls that overlap on one GPU share
// the array and race; nothing in it needs clearing (the x pass writes every
// element the y pass reads).
extern "C" int trh_lanczos_shift_batch_noalloc(trh_handle *h, const float *in, float *out, int n,
                                               int height, int width, const double *dx,
                                               const double *dy, void *stream) {
    if (!h || !in || !out || !dx || !dy || n <= 0 || height <= 0 || width <= 0)
        return set_err(TRH_E_ARG, "trh_lanczos_shift_batch_noalloc: bad arguments");
    const unsigned long long floats = (unsigned long long)n * height * width;
    if (floats > (unsigned long long)TRH_NOALLOC_SCRATCH_FLOATS)
        return set_err(TRH_E_ARG,
                       "trh_lanczos_shift_batch_noalloc: %d images of %d x %d need %llu floats "
                       "of scratch, more than the %d it has",
                       n, height, width, floats, TRH_NOALLOC_SCRATCH_FLOATS);
    if (!h->noalloc_scratch)
        return set_err(TRH_E_ARG, "trh_lanczos_shift_batch_noalloc: handle has no scratch address");
    int dev = -1;
    int rc = hip_check(hipGetDevice(&dev), "hipGetDevice");
    if (rc) return rc;
    if (dev != h->device)
        return set_err(TRH_E_ARG,
                       "trh_lanczos_shift_batch_noalloc: device %d is current, the handle is on %d",
                       dev, h->device);
    return lanczos_passes(h, in, (float *)h->noalloc_scratch, out, n, height, width, dx, dy,
                          (hipStream_t)stream);
}

// ---------------------------------------------------------------------------
// Batched pixelized-PSF point renders (generic API path)
// ---------------------------------------------------------------------------
extern "C" int trh_render_point_batch(trh_handle *h, const trh_fft_job *jobs, int njobs,
                                      const float *stamp, int ph, int pw, float *out,
                                      const float *data, const float *inverr, double *chi2,
                                      void *stream) {
    if (!h || !jobs || !stamp || njobs < 0 || ph <= 0 || pw <= 0)
        return set_err(TRH_E_ARG, "trh_render_point_batch: bad arguments");
    if (njobs == 0) return TRH_OK;
    const size_t lds = point_lds_bytes(ph, pw);
    if (lds + sizeof(RenderScratch) > (size_t)h->lds_max)
        return set_err(TRH_E_UNSUPPORTED, "point render: stamp %dx%d exceeds LDS", ph, pw);
    int grid = njobs < h->n_cu * 4 ? njobs : h->n_cu * 4;
    // no output image but chi^2 slots: every job is a chi^2 job
    if (!out && chi2 && data && inverr)
        hipLaunchKernelGGL(k_render_point<TRH_CHI2>, dim3(grid), dim3(256), lds,
                           (hipStream_t)stream, jobs, nullptr, nullptr, njobs, stamp, ph, pw,
                           out, data, inverr, chi2, nullptr);
    else
        hipLaunchKernelGGL(k_render_point<-1>, dim3(grid), dim3(256), lds, (hipStream_t)stream,
                           jobs, nullptr, nullptr, njobs, stamp, ph, pw, out, data, inverr,
                           chi2, nullptr);
    return hip_check(hipGetLastError(), "k_render_point");
}

// ---------------------------------------------------------------------------
// Dense normal equations and FP64 solve (the least-squares step of
// LsqrOptimizer.getUpdateDirection, tractor/lsqr_optimizer.py:328-644, for
// the object API path: A restricted to the rows its derivatives touch).
// ---------------------------------------------------------------------------
// Ab: column-major [ncols+1][nrows] float32 (the last column is b).
// out[(ncols+1)^2] += sum over rows of Ab[:, i] * Ab[:, j] (FP64 atomics).
__global__ void __launch_bounds__(256)
k_normal_eq_dense(const float *__restrict__ Ab, int nrows, int nc1, double *__restrict__ out) {
    __shared__ double red[4];
    const int p = blockIdx.y;          // pair index over the upper triangle
    int i = 0, rem = p;
    while (rem >= nc1 - i) {
        rem -= nc1 - i;
        ++i;
    }
    const int j = i + rem;
    const float *ci = Ab + (size_t)i * nrows;
    const float *cj = Ab + (size_t)j * nrows;
    double acc = 0.0;
    for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < nrows; r += gridDim.x * blockDim.x)
        acc += (double)(ci[r] * cj[r]);
    const double t = block_sum(acc, red);
    if (threadIdx.x == 0) {
        atomicAdd(out + i * nc1 + j, t);
        if (i != j) atomicAdd(out + j * nc1 + i, t);
    }
}

__global__ void k_zero_f64(double *a, int n) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) a[i] = 0.0;
}

extern "C" int trh_normal_eq_dense(trh_handle *h, const float *Ab, int nrows, int ncols, double *out,
                                   void *stream) {
    if (!h || !Ab || !out || nrows <= 0 || ncols <= 0 || ncols > 512)
        return set_err(TRH_E_ARG, "trh_normal_eq_dense: bad arguments");
    const int nc1 = ncols + 1;
    const int npairs = nc1 * (nc1 + 1) / 2;
    if (npairs > 65535) return set_err(TRH_E_ARG, "trh_normal_eq_dense: too many columns");
    hipStream_t st = (hipStream_t)stream;
    hipLaunchKernelGGL(k_zero_f64, dim3(64), dim3(256), 0, st, out, nc1 * nc1);
    int gx = (nrows + 255) / 256;
    if (gx > 64) gx = 64;
    hipLaunchKernelGGL(k_normal_eq_dense, dim3(gx, npairs), dim3(256), 0, st, Ab, nrows, nc1, out);
    return hip_check(hipGetLastError(), "k_normal_eq_dense");
}

// Solve (D^-1 N D^-1 + damp^2 I) y = D^-1 g, x = D^-1 y, with N = out[:n,:n],
// g = out[:n, n] from trh_normal_eq_dense; D = sqrt(diag N) when
// scale_columns, else I.  Columns with N_ii == 0 get x_i = 0.  One block, FP64
// Cholesky in LDS-free global scratch.  info: 0 ok, k>0 not positive definite
// at column k (x = 0).
__global__ void k_solve_normal_f64(const double *__restrict__ Nb, int n, double damp, int scale,
                                   double *__restrict__ L, double *__restrict__ x,
                                   int *__restrict__ info) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const int n1 = n + 1;
    for (int i = 0; i < n; ++i) {
        const double Dii = sqrt(Nb[i * n1 + i]);
        x[i] = 0.0;
        L[n * n + i] = (scale ? Dii : (Nb[i * n1 + i] > 0.0 ? 1.0 : 0.0));  // D
    }
    const double *D = L + n * n;
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
            double v;
            if (D[i] > 0.0 && D[j] > 0.0)
                v = Nb[i * n1 + j] / (D[i] * D[j]);
            else
                v = (i == j) ? 1.0 : 0.0;
            if (i == j && D[i] > 0.0) v += damp * damp;
            L[i * n + j] = v;
        }
    for (int j = 0; j < n; ++j) {
        double d = L[j * n + j];
        for (int k = 0; k < j; ++k) d -= L[j * n + k] * L[j * n + k];
        if (!(d > 0.0)) {
            *info = j + 1;
            return;
        }
        d = sqrt(d);
        L[j * n + j] = d;
        for (int i = j + 1; i < n; ++i) {
            double v = L[i * n + j];
            for (int k = 0; k < j; ++k) v -= L[i * n + k] * L[j * n + k];
            L[i * n + j] = v / d;
        }
    }
    // forward: L y = D^-1 g  (y stored in x)
    for (int i = 0; i < n; ++i) {
        double v = D[i] > 0.0 ? Nb[i * n1 + n] / D[i] : 0.0;
        for (int k = 0; k < i; ++k) v -= L[i * n + k] * x[k];
        x[i] = v / L[i * n + i];
    }
    for (int i = n - 1; i >= 0; --i) {
        double v = x[i];
        for (int k = i + 1; k < n; ++k) v -= L[k * n + i] * x[k];
        x[i] = v / L[i * n + i];
    }
    for (int i = 0; i < n; ++i) x[i] = D[i] > 0.0 ? x[i] / D[i] : 0.0;
    *info = 0;
}

extern "C" int trh_solve_normal_f64(trh_handle *h, const double *Nb, int n, double damp,
                                    int scale_columns, double *x, int *info_host, void *stream) {
    if (!h || !Nb || !x || n <= 0 || n > 512)
        return set_err(TRH_E_ARG, "trh_solve_normal_f64: bad arguments (1 <= n <= 512)");
    hipStream_t st = (hipStream_t)stream;
    int rc = shared_begin(h, st);
    if (rc) return rc;
    if ((rc = ensure_scratch(h, sizeof(double) * ((size_t)n * n + n) + 64))) return rc;
    double *L = (double *)h->scratch;
    int *info = (int *)(L + (size_t)n * n + n);
    hipLaunchKernelGGL(k_solve_normal_f64, dim3(1), dim3(64), 0, st, Nb, n, damp, scale_columns, L,
                       x, info);
    if ((rc = hip_check(hipGetLastError(), "k_solve_normal_f64"))) return rc;
    if ((rc = shared_end(h, st))) return rc;
    if (info_host) {
        if ((rc = hip_check(hipMemcpyAsync(info_host, info, sizeof(int), hipMemcpyDeviceToHost, st),
                            "copy info")))
            return rc;
        return hip_check(hipStreamSynchronize(st), "solve sync");
    }
    return TRH_OK;
}
