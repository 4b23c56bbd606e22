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
int launch_fft_any(trh_handle *h, int sz, int mode, const trh_fft_job *jobs, const int *list,
                   const int *count, int njobs_bound, int maxrows, const float *psf_spec,
                   float *out, const float *data, const float *inverr, double *chi2,
                   hipStream_t st);
}  // namespace trh
