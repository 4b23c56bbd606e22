// Internal declarations shared by the library's translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>

#include "../../include/tractor_hip.h"

struct trh_handle {
    int device = 0;
    int n_cu = 256;
    int lds_max = 65536;
    float2 *tw = nullptr;      // twiddle tables for N = 32..512
    void *scratch = nullptr;   // grown on demand by entry points that need it
    size_t scratch_bytes = 0;
    int fft_threads[5] = {0, 0, 0, 0, 0};  // per FFT size 32..512; 0 = default
    int fft_ch[5] = {0, 0, 0, 0, 0};
    // per FFT size: global half-spectrum slots of the SZ >= 256 renders,
    // allocated once (a captured graph keeps the address)
    float2 *hbuf[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
    // ordering of the shared buffers above (scratch, hbuf) across streams:
    // every entry point that uses them records order_ev on its stream at the
    // end; a later call on another stream waits for it first
    hipEvent_t order_ev = nullptr;
    hipStream_t order_stream = nullptr;
    bool order_recorded = false;
    // device address of the allocation-free entry point's scratch array on
    // this handle's device
    void *noalloc_scratch = nullptr;
};

namespace trh {
int set_err(int code, const char *fmt, ...);
int hip_check(hipError_t e, const char *what);
int tw_offset(int n);
int pad_center(int sz, int p);
int ensure_scratch(trh_handle *h, size_t bytes);
int max_rows_for(int sz, int lds_max);
void fft_launch_shape(trh_handle *h, int sz, int maxrows, int *threads, int *rows, size_t *lds,
                      int *grid, int njobs);
int fft_hbuf_slots(const trh_handle *h, int sz);
int reserve_fft_hbuf(trh_handle *h, int sz);
// `hbuf`: the row slots of the SZ >= 256 renders (fft_hbuf_slots(h, sz)
// slots of fft_hbuf_slot_floats(sz) floats), or NULL for the handle's own
// `base` (TRH_CHI2 only): every job with pad0 > 0 adds the window at
// base + pad0 - 1 (row stride mw) to its model (the TRH_CHI2_BASE kernel).
int launch_fft_any(trh_handle *h, int sz, int mode, const trh_fft_job *jobs, const int *list,
                   const int *count, int njobs_bound, int maxrows, const float *psf_spec,
                   float *out, const float *data, const float *inverr, double *chi2,
                   hipStream_t st, float2 *hbuf = nullptr, const float *base = nullptr);
// Bracket every use of the handle's shared buffers (scratch, hbuf) on `st`:
// shared_begin orders `st` after the last use on another stream, shared_end
// records this use.  Streams being captured into a graph are not ordered
// (a graph that uses the shared buffers must own the handle's stream).
int shared_begin(trh_handle *h, hipStream_t st);
int shared_end(trh_handle *h, hipStream_t st);
}  // namespace trh
