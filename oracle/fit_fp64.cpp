// FP64 C++ restatement of the oracle's one-source fit, threaded over sources.
//
// TEST INFRASTRUCTURE ONLY: the CPU baseline of the batched-fit benchmark and
// a parity reference with a larger sample than the NumPy oracle can fit in
// the same time.  Nothing in tractor_amd links or loads it.
//
// It computes what oracle/tractor_oracle.py's fit_source computes for
// SourceProblem kinds 0 (point source), 1 (ExpGalaxy) and 2 (DevGalaxy) on a
// P x P modelMask at (0, 0), with the same FP32 / FP64 split, so both follow
// the same reference code:
//   PixelizedPSF spectrum, size and centre   tractor/psf.py:209-280
//   MoG Fourier transform (delta core)        tractor/mp_fourier.i:431-478
//   irfft, float32 cast, Lanczos-3 shift      tractor/galaxy.py:336-519,
//                                             tractor/mp_fourier.i:267-429
//   pixelized point source with derivatives  tractor/psf.py:137-207
//   parameter derivatives                     tractor/galaxy.py:137-223,
//                                             tractor/pointsource.py:204-241
//   column-scaled least squares               tractor/lsqr_optimizer.py:470-634
//   ConstrainedOptimizer loop / steps / walk  tractor/constrained_optimizer.py:12-159
// Differences from the NumPy oracle are rounding only: its own radix-2 FFT
// instead of pocketfft, sequential instead of pairwise sums, a Householder QR
// plus a one-sided Jacobi SVD of R instead of LAPACK's lstsq (same rank cut,
// eps * max(M, N) * s_max).
//
// C ABI: orc_fit_sources (below).  Build: oracle/Makefile.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <map>
#include <mutex>
#include <thread>
#include <vector>

// glibc's vector exp (libmvec) for the simd loop of the MoG transform
extern "C" __attribute__((simd("notinbranch"))) double exp(double);

namespace {

using cd = std::complex<double>;

const double EXP_AMP_RAW[6] = {2.34853813e-03, 3.07995260e-02, 2.23364214e-01,
                               1.17949102e+00, 4.33873750e+00, 5.99820770e+00};
const double EXP_VAR[6] = {1.20078965e-03, 8.84526493e-03, 3.91463084e-02,
                           1.39976817e-01, 4.60962500e-01, 1.50159566e+00};
const double DEV_AMP_RAW[8] = {4.26347652e-02, 2.40127183e-01, 6.85907632e-01,
                               1.51937350e+00, 2.83627243e+00, 4.46467501e+00,
                               5.72440830e+00, 5.60989349e+00};
const double DEV_VAR[8] = {2.23759216e-04, 1.00220099e-03, 4.18731126e-03,
                           1.69432589e-02, 6.84850479e-02, 2.87207080e-01,
                           1.33320254e+00, 8.40215071e+00};
const double DEV_CORE = 0.010233;

struct Profile {
    int K;
    double amp[9], var[9];
};

// mixture_profiles.py:17-38 with the normalisations of the oracle's
// exp_mixture / dev_mixture (delta core first for dev)
Profile profile(int kind) {
    Profile p;
    if (kind == 1) {
        double s = 0;
        for (double a : EXP_AMP_RAW) s += a;
        double amp[6], s2 = 0;
        for (int k = 0; k < 6; ++k) s2 += (amp[k] = EXP_AMP_RAW[k] / s);
        p.K = 6;
        for (int k = 0; k < 6; ++k) {
            p.amp[k] = amp[k] / s2;
            p.var[k] = EXP_VAR[k];
        }
    } else {
        double s = 0;
        for (double a : DEV_AMP_RAW) s += a;
        double amp[9];
        amp[0] = DEV_CORE;
        double var[9];
        var[0] = 0.;
        for (int k = 0; k < 8; ++k) {
            amp[k + 1] = DEV_AMP_RAW[k] * (1. - DEV_CORE) / s;
            var[k + 1] = DEV_VAR[k];
        }
        double s2 = 0;
        for (int k = 0; k < 9; ++k) s2 += amp[k];
        p.K = 9;
        for (int k = 0; k < 9; ++k) {
            p.amp[k] = amp[k] / s2;
            p.var[k] = var[k];
        }
    }
    return p;
}

int round_half_even(double x) { return (int)std::nearbyint(x); }

// ---- radix-2 complex FFT, sign -1 forward (numpy's convention) ----
struct FFT {
    int n = 0, lg = 0;
    std::vector<cd> w;   // exp(-2 pi i k / n), k < n / 2
    std::vector<int> rev;
    explicit FFT(int n_) : n(n_) {
        while ((1 << lg) < n) ++lg;
        w.resize(n / 2 > 0 ? n / 2 : 1);
        for (int k = 0; k < n / 2; ++k) {
            const double a = -2. * M_PI * k / n;
            w[k] = cd(std::cos(a), std::sin(a));
        }
        rev.resize(n);
        for (int i = 0; i < n; ++i) {
            int r = 0;
            for (int b = 0; b < lg; ++b)
                if (i & (1 << b)) r |= 1 << (lg - 1 - b);
            rev[i] = r;
        }
    }
    // in place, stride s between elements; inverse: conjugate twiddles, no scaling
    void run(cd *x, int s, bool inverse, cd *tmp) const {
        for (int i = 0; i < n; ++i) tmp[rev[i]] = x[(size_t)i * s];
        // complex arithmetic written out (the textbook product, as NumPy's)
        double *t2 = reinterpret_cast<double *>(tmp);
        const double sg = inverse ? -1. : 1.;
        for (int len = 2; len <= n; len <<= 1) {
            const int half = len / 2, step = n / len;
            for (int i = 0; i < n; i += len)
                for (int j = 0; j < half; ++j) {
                    const double wr = w[j * step].real(), wi = sg * w[j * step].imag();
                    double *u = t2 + 2 * (i + j), *q = t2 + 2 * (i + j + half);
                    const double vr = q[0] * wr - q[1] * wi, vi = q[0] * wi + q[1] * wr;
                    const double ur = u[0], ui = u[1];
                    u[0] = ur + vr;
                    u[1] = ui + vi;
                    q[0] = ur - vr;
                    q[1] = ui - vi;
                }
        }
        for (int i = 0; i < n; ++i) x[(size_t)i * s] = tmp[i];
    }
};

// PixelizedPSF spectrum of one FFT size (psf.py:249-280; oracle psf_spectrum):
// the stamp padded into sz x sz (pad_psf), rfft2 in FP64, cast to complex64.
struct PsfSpec {
    int sz = 0, nv = 0, cx = 0, cy = 0;
    std::vector<cd> P;   // [w][v], values rounded to complex64
};

PsfSpec make_spec(const float *img, int ph, int pw, int sz) {
    PsfSpec S;
    S.sz = sz;
    S.nv = sz / 2 + 1;
    std::vector<double> pad((size_t)sz * sz, 0.);
    int y0, x0, cuty = 0, cutx = 0, sh = ph, sw = pw;
    if (sz >= ph) {
        y0 = (sz - ph) / 2;
        S.cy = y0 + ph / 2;
    } else {
        y0 = 0;
        cuty = (ph - sz) / 2;
        sh = sz;
        S.cy = ph / 2 - cuty;
    }
    if (sz >= pw) {
        x0 = (sz - pw) / 2;
        S.cx = x0 + pw / 2;
    } else {
        x0 = 0;
        cutx = (pw - sz) / 2;
        sw = sz;
        S.cx = pw / 2 - cutx;
    }
    for (int y = 0; y < sh; ++y)
        for (int x = 0; x < sw; ++x)
            pad[(size_t)(y0 + y) * sz + x0 + x] = img[(size_t)(y + cuty) * pw + x + cutx];
    FFT f(sz);
    std::vector<cd> a((size_t)sz * sz), tmp(sz);
    for (size_t i = 0; i < a.size(); ++i) a[i] = pad[i];
    for (int y = 0; y < sz; ++y) f.run(a.data() + (size_t)y * sz, 1, false, tmp.data());
    for (int x = 0; x < S.nv; ++x) f.run(a.data() + x, sz, false, tmp.data());
    S.P.resize((size_t)sz * S.nv);
    for (int w = 0; w < sz; ++w)
        for (int v = 0; v < S.nv; ++v) {
            const cd z = a[(size_t)w * sz + v];
            S.P[(size_t)w * S.nv + v] = cd((float)z.real(), (float)z.imag());
        }
    return S;
}

int fft_size(int halfsize) { return 1 << (int)std::ceil(std::log2(halfsize * 2.)); }

// ---- Lanczos-3 (mp_fourier.i:355-429; oracle lanczos_taps3 / correlate7f) ----
void lanczos_taps3(double dx, float taps[7]) {
    double f[7] = {0, 0, 0, 0, 0, 0, 0};
    if (dx == 0.) {
        f[3] = 1.;
    } else {
        const int lo = dx > 0 ? 0 : 1, hi = dx > 0 ? 6 : 7;
        double s = 0;
        for (int i = lo; i < hi; ++i) {
            const double pix = M_PI * (i - 3.0 + dx);
            f[i] = 3 * std::sin(pix) * std::sin(pix / 3.0) / (pix * pix);
            s += f[i];
        }
        for (int i = lo; i < hi; ++i) f[i] /= s;
    }
    for (int i = 0; i < 7; ++i) taps[i] = (float)f[i];
}

// rows with fx, then columns with fy, zero padding, FP32 products and sums
// in tap order (two roundings per tap, as the oracle's NumPy expression).
// Only outputs in rows [oy0, oy1) x columns [ox0, ox1) are computed (the
// rest of img is left stale); they read input rows oy0 - 3 .. oy1 + 2 only.
void lanczos_shift(std::vector<float> &img, int H, int W, double dx, double dy,
                   std::vector<float> &work, int oy0 = 0, int oy1 = -1, int ox0 = 0,
                   int ox1 = -1) {
    if (oy1 < 0) oy1 = H;
    if (ox1 < 0) ox1 = W;
    float fx[7], fy[7];
    lanczos_taps3(dx, fx);
    lanczos_taps3(dy, fy);
    work.assign((size_t)H * W, 0.f);
    const int ry0 = std::max(0, oy0 - 3), ry1 = std::min(H, oy1 + 3);
    for (int y = ry0; y < ry1; ++y) {
        const float *r = img.data() + (size_t)y * W;
        float *o = work.data() + (size_t)y * W;
        for (int x = ox0; x < ox1; ++x) {
            float acc = 0.f;
            for (int k = 0; k < 7; ++k) {
                const int xx = x - 3 + k;
                const float v = (xx >= 0 && xx < W) ? r[xx] : 0.f;
                const float p = fx[k] * v;
                acc = acc + p;
            }
            o[x] = acc;
        }
    }
    for (int y = oy0; y < oy1; ++y)
        for (int x = ox0; x < ox1; ++x) {
            float acc = 0.f;
            for (int k = 0; k < 7; ++k) {
                const int yy = y - 3 + k;
                const float v = (yy >= 0 && yy < H) ? work[(size_t)yy * W + x] : 0.f;
                const float p = fy[k] * v;
                acc = acc + p;
            }
            img[(size_t)y * W + x] = acc;
        }
}

// ---- shapes (ellipses.py:126-155, 232-255; galaxy.py:659-669) ----
// pixel covariance T var T^T of an isotropic component, T = 3600 I x basis
void sheared_var(double logre, double ee1, double ee2, double v, double out[3]) {
    const double re = std::exp(logre);
    const double ee = std::hypot(ee1, ee2);
    const double e = 1. - std::exp(-ee);
    const double th = std::atan2(ee2, ee1) / 2.;
    const double ct = std::cos(th), st = std::sin(th);
    const double ab = e >= 1. ? 1000. : std::min(1000., (1. + e) / (1. - e));
    const double r = re / 3600.;
    const double T[2][2] = {{3600. * (r * (ct / ab)), 3600. * (r * st)},
                            {3600. * (r * (-st / ab)), 3600. * (r * ct)}};
    // T (v I) T^T
    out[0] = v * (T[0][0] * T[0][0] + T[0][1] * T[0][1]);
    out[1] = v * (T[0][0] * T[1][0] + T[0][1] * T[1][1]);
    out[2] = v * (T[1][0] * T[1][0] + T[1][1] * T[1][1]);
}

// One component's terms of one row of the MoG transform:
// F[i] += amp exp(n2p2 (a v_i^2 + b2 v_i w + dww)), the exponent in the
// oracle's operation order.  Cloned for AVX2 hosts (libmvec's 4-wide exp)
// and for any x86-64.
__attribute__((target_clones("avx2", "default")))
void mog_row(const double *vv, int nv, double n2p2, double a, double b2, double dww, double w,
             double amp, double *ex, double *F) {
#pragma omp simd
    for (int i = 0; i < nv; ++i) {
        const double v = vv[i];
        ex[i] = ::exp(n2p2 * (a * v * v + b2 * v * w + dww));
    }
    for (int i = 0; i < nv; ++i) F[i] += amp * ex[i];
}

struct Work {
    std::vector<double> fv, fa, fe;   // MoG transform of one row
    std::vector<cd> a, tmp, row;
    std::vector<float> g, lw;
    std::map<int, FFT> plans;   // per transform size
    const FFT &plan(int n) {
        auto it = plans.find(n);
        if (it == plans.end()) it = plans.emplace(n, FFT(n)).first;
        return it->second;
    }
};

struct Problem {
    int kind, P;
    const float *data, *inverr;   // P x P
    const float *psf;
    int ph, pw;
    double cdi[4];
    std::map<int, PsfSpec> *specs;
    std::mutex *spec_mu;
};

const PsfSpec &spec_for(const Problem &pr, int sz) {
    std::lock_guard<std::mutex> g(*pr.spec_mu);
    auto it = pr.specs->find(sz);
    if (it == pr.specs->end())
        it = pr.specs->emplace(sz, make_spec(pr.psf, pr.ph, pr.pw, sz)).first;
    return it->second;
}

// Unit-flux galaxy render on the P x P window at (0, 0), force_sourceOut
// (oracle render_galaxy_fft with mm); false when the centre is outside the
// window (the reference raises there without a hybrid PSF).
bool render_galaxy(const Problem &pr, const double *p, int shape_step, float *out, Work &W) {
    const int P = pr.P;
    const double px = p[0], py = p[1];
    if (px < 0 || px > P - 1 || py < 0 || py > P - 1) return false;
    double sp[3] = {p[3], p[4], p[5]};
    if (shape_step >= 0) sp[shape_step] += 0.01;
    double hs = std::max(P / 2., P / 2.);
    hs = std::max(hs, std::max(std::max(1 + px, 1 + P - px), std::max(1 + py, 1 + P - py)));
    hs = std::max(hs, std::max(pr.pw / 2., pr.ph / 2.));
    const int halfsize = (int)std::ceil(hs);
    const int sz = fft_size(halfsize), nv = sz / 2 + 1;
    const PsfSpec &S = spec_for(pr, sz);
    double mux = px - S.cx, muy = py - S.cy;
    const int sx = round_half_even(mux), sy = round_half_even(muy);
    mux -= sx;
    muy -= sy;
    const Profile pf = profile(pr.kind);
    double var[9][3];
    for (int k = 0; k < pf.K; ++k) sheared_var(sp[0], sp[1], sp[2], pf.var[k], var[k]);
    // F[w][v] * P[w][v]; columns v then rows (irfftn: c2c along w, c2r along v).
    // Components outermost as in the oracle's sum; the exponent keeps its
    // operation order; exp is glibc's vector variant (libmvec).
    W.a.assign((size_t)sz * nv, cd(0, 0));
    const double n2p2 = -2. * M_PI * M_PI;
    W.fv.resize(nv);
    W.fa.resize(nv);
    W.fe.resize(nv);
    double *vv = W.fv.data(), *Fr = W.fa.data(), *ex = W.fe.data();
    for (int i = 0; i < nv; ++i) vv[i] = i / (double)sz;
    for (int j = 0; j < sz; ++j) {
        const double w = (j < sz / 2 ? j : j - sz) / (double)sz;
        for (int i = 0; i < nv; ++i) Fr[i] = 0.;
        for (int k = 0; k < pf.K; ++k) {
            const double a = var[k][0], b2 = 2. * var[k][1], d = var[k][2], amp = pf.amp[k];
            if (a == 0. && var[k][1] == 0. && d == 0.) {
                for (int i = 0; i < nv; ++i) Fr[i] += amp;
                continue;
            }
            mog_row(vv, nv, n2p2, a, b2, d * w * w, w, amp, ex, Fr);
        }
        for (int i = 0; i < nv; ++i) {
            const cd &pp = S.P[(size_t)j * nv + i];
            W.a[(size_t)j * nv + i] = cd(Fr[i] * pp.real(), Fr[i] * pp.imag());
        }
    }
    const FFT &f = W.plan(sz);
    W.tmp.resize(sz);
    for (int i = 0; i < nv; ++i) f.run(W.a.data() + i, nv, true, W.tmp.data());
    // c2r of the G rows the window needs (with the 3 rows either side the
    // y-Lanczos reads; the others stay zero and are never read for the
    // window), two real rows per complex inverse FFT: Z = A + i B with A, B
    // the Hermitian extensions of the two half rows (imaginary parts at DC
    // and Nyquist dropped), real(ifft Z) = row a, imag = row b.
    W.g.assign((size_t)sz * sz, 0.f);
    std::vector<cd> &row = W.row;
    row.resize(sz);
    const double norm = 1. / ((double)sz * sz);
    const int ylo = std::max(0, -sy - 3), yhi = std::min(sz, P - sy + 3);
    auto ext = [&](const cd *h, int k) -> cd {
        if (k == 0 || k == sz / 2) return cd(h[k].real(), 0.);
        return k < sz / 2 ? h[k] : std::conj(h[sz - k]);
    };
    for (int y = ylo; y < yhi; y += 2) {
        const bool two = y + 1 < yhi;
        const cd *ha = W.a.data() + (size_t)y * nv;
        const cd *hb = two ? ha + nv : nullptr;
        for (int k = 0; k < sz; ++k) {
            const cd A = ext(ha, k);
            const cd B = two ? ext(hb, k) : cd(0., 0.);
            row[k] = cd(A.real() - B.imag(), A.imag() + B.real());
        }
        f.run(row.data(), 1, true, W.tmp.data());
        for (int x = 0; x < sz; ++x) {
            W.g[(size_t)y * sz + x] = (float)(row[x].real() * norm);
            if (two) W.g[(size_t)(y + 1) * sz + x] = (float)(row[x].imag() * norm);
        }
    }
    // the window reads G rows [-sy, P - sy) and columns [-sx, P - sx)
    if (mux != 0. || muy != 0.)
        lanczos_shift(W.g, sz, sz, mux, muy, W.lw, std::max(0, -sy), std::min(sz, P - sy),
                      std::max(0, -sx), std::min(sz, P - sx));
    for (int y = 0; y < P; ++y)
        for (int x = 0; x < P; ++x) {
            const int gy = y - sy, gx = x - sx;
            out[(size_t)y * P + x] =
                (gy >= 0 && gy < sz && gx >= 0 && gx < sz) ? W.g[(size_t)gy * sz + gx] : 0.f;
        }
    return true;
}

// PixelizedPSF.getPointSourcePatch on the window (oracle pixpsf_point_patch),
// with the central-difference position derivatives when dx != nullptr.
void render_point(const Problem &pr, const double *p, float *out, float *ddx, float *ddy,
                  Work &W) {
    const int P = pr.P, H = pr.ph, Wd = pr.pw;
    const int ix = round_half_even(p[0]), iy = round_half_even(p[1]);
    const double dx = p[0] - ix, dy = p[1] - iy;
    const int x0 = ix - Wd / 2, y0 = iy - H / 2;
    std::fill(out, out + (size_t)P * P, 0.f);
    if (ddx) {
        std::fill(ddx, ddx + (size_t)P * P, 0.f);
        std::fill(ddy, ddy + (size_t)P * P, 0.f);
    }
    if (0 >= x0 + Wd || P <= x0 || 0 >= y0 + H || P <= y0) return;
    const int L = 3, ph = P + 2 * L, pw = P + 2 * L;
    W.g.assign((size_t)ph * pw, 0.f);
    for (int yy = 0; yy < ph; ++yy)
        for (int xx = 0; xx < pw; ++xx) {
            const int sy = yy - L - y0, sx = xx - L - x0;
            if (sy >= 0 && sy < H && sx >= 0 && sx < Wd)
                W.g[(size_t)yy * pw + xx] = pr.psf[(size_t)sy * Wd + sx];
        }
    lanczos_shift(W.g, ph, pw, dx, dy, W.lw);
    for (int y = 0; y < P; ++y)
        for (int x = 0; x < P; ++x) out[(size_t)y * P + x] = W.g[(size_t)(y + L) * pw + x + L];
    if (!ddx) return;
    for (int y = 1; y < P - 1; ++y)
        for (int x = 1; x < P - 1; ++x) {
            const float *o = out + (size_t)y * P + x;
            ddx[(size_t)y * P + x] = -(o[1] - o[-1]) / 2.f;
            ddy[(size_t)y * P + x] = -(o[P] - o[-P]) / 2.f;
        }
}

int nparams(int kind) { return kind == 0 ? 3 : 6; }

bool render(const Problem &pr, const double *p, float *out, Work &W) {
    if (pr.kind == 0) {
        render_point(pr, p, out, nullptr, nullptr, W);
        return true;
    }
    return render_galaxy(pr, p, -1, out, W);
}

// oracle source_chisq: float32 model and chi, FP64 sum of squares
double chisq_of(const Problem &pr, const double *p, const float *patch) {
    const float c = (float)p[2];
    double s = 0.;
    for (int i = 0; i < pr.P * pr.P; ++i) {
        const float mod = patch[i] * c;
        const float chi = (pr.data[i] - mod) * pr.inverr[i];
        s += (double)chi * (double)chi;
    }
    return s;
}

// min ||A x - b|| over A (m x n, column-major, already column scaled):
// Householder QR, then a one-sided Jacobi SVD of R with numpy's rank cut.
bool lstsq(std::vector<double> &A, int m, int n, std::vector<double> &b, double *x) {
    for (int k = 0; k < n; ++k) {
        double *a = A.data() + (size_t)k * m;
        double nrm = 0;
        for (int i = k; i < m; ++i) nrm += a[i] * a[i];
        nrm = std::sqrt(nrm);
        if (nrm == 0.) continue;
        const double alpha = a[k] > 0 ? -nrm : nrm;
        a[k] -= alpha;
        double vv = 0;
        for (int i = k; i < m; ++i) vv += a[i] * a[i];
        for (int j = k + 1; j < n; ++j) {
            double *c = A.data() + (size_t)j * m;
            double d = 0;
            for (int i = k; i < m; ++i) d += a[i] * c[i];
            const double f = 2. * d / vv;
            for (int i = k; i < m; ++i) c[i] -= f * a[i];
        }
        double d = 0;
        for (int i = k; i < m; ++i) d += a[i] * b[i];
        const double f = 2. * d / vv;
        for (int i = k; i < m; ++i) b[i] -= f * a[i];
        a[k] = alpha;   // R's diagonal
        for (int i = k + 1; i < m; ++i) a[i] = 0.;
    }
    // R (n x n, upper) and c = (Q^T b)[0:n]; one-sided Jacobi on the columns of R
    double R[8][8], V[8][8], c[8];
    for (int i = 0; i < n; ++i) {
        c[i] = b[i];
        for (int j = 0; j < n; ++j) {
            R[i][j] = i <= j ? A[(size_t)j * m + i] : 0.;
            V[i][j] = i == j ? 1. : 0.;
        }
    }
    for (int sweep = 0; sweep < 60; ++sweep) {
        double off = 0;
        for (int p = 0; p < n - 1; ++p)
            for (int q = p + 1; q < n; ++q) {
                double al = 0, be = 0, ga = 0;
                for (int i = 0; i < n; ++i) {
                    al += R[i][p] * R[i][p];
                    be += R[i][q] * R[i][q];
                    ga += R[i][p] * R[i][q];
                }
                if (ga == 0.) continue;
                off = std::max(off, std::fabs(ga) / std::sqrt(al * be));
                const double zeta = (be - al) / (2. * ga);
                const double t = (zeta > 0 ? 1. : -1.) / (std::fabs(zeta) + std::sqrt(1. + zeta * zeta));
                const double cs = 1. / std::sqrt(1. + t * t), sn = cs * t;
                for (int i = 0; i < n; ++i) {
                    const double rp = R[i][p], rq = R[i][q];
                    R[i][p] = cs * rp - sn * rq;
                    R[i][q] = sn * rp + cs * rq;
                    const double vp = V[i][p], vq = V[i][q];
                    V[i][p] = cs * vp - sn * vq;
                    V[i][q] = sn * vp + cs * vq;
                }
            }
        if (off < 1e-15) break;
    }
    // R V = U S: s_j = ||col j||, u_j = col j / s_j; x = V S^+ U^T c
    double s[8], smax = 0;
    for (int j = 0; j < n; ++j) {
        double t = 0;
        for (int i = 0; i < n; ++i) t += R[i][j] * R[i][j];
        s[j] = std::sqrt(t);
        smax = std::max(smax, s[j]);
    }
    const double cut = 2.220446049250313e-16 * std::max(m, n) * smax;
    for (int i = 0; i < n; ++i) x[i] = 0.;
    for (int j = 0; j < n; ++j) {
        if (s[j] <= cut || s[j] == 0.) continue;
        double ut = 0;
        for (int i = 0; i < n; ++i) ut += R[i][j] / s[j] * c[i];
        for (int i = 0; i < n; ++i) x[i] += V[i][j] * ut / s[j];
    }
    return true;
}

// oracle source_update: derivative columns x inverr (float32), column-scaled
// lstsq against chi; false when no column is live.
bool update_direction(const Problem &pr, const double *p, float *p0, double *X, Work &W,
                      std::vector<float> &cols) {
    const int P = pr.P, npix = P * P, np_ = nparams(pr.kind);
    cols.assign((size_t)np_ * npix, 0.f);
    const double counts = p[2];
    if (pr.kind == 0) {
        std::vector<float> dx(npix), dy(npix);
        render_point(pr, p, p0, dx.data(), dy.data(), W);
        for (int i = 0; i < npix; ++i) {
            for (int c = 0; c < 2; ++c)
                cols[(size_t)c * npix + i] =
                    (float)((dx[i] * pr.cdi[0 * 2 + c] + dy[i] * pr.cdi[1 * 2 + c]) * counts);
            cols[(size_t)2 * npix + i] = p0[i];
        }
    } else {
        if (!render_galaxy(pr, p, -1, p0, W)) return false;
        std::vector<float> pk(npix);
        for (int y = 0; y < P; ++y)
            for (int x = 0; x < P; ++x) {
                const int i = y * P + x;
                float gx = 0.f, gy = 0.f;
                if (y > 0 && y < P - 1 && x > 0 && x < P - 1) {
                    gx = (p0[i - 1] - p0[i + 1]) / 2.f;
                    gy = (p0[i - P] - p0[i + P]) / 2.f;
                }
                for (int c = 0; c < 2; ++c)
                    cols[(size_t)c * npix + i] =
                        (float)((gx * pr.cdi[0 * 2 + c] + gy * pr.cdi[1 * 2 + c]) * counts);
                cols[(size_t)2 * npix + i] = p0[i];
            }
        for (int k = 0; k < 3; ++k) {
            if (!render_galaxy(pr, p, k, pk.data(), W)) return false;
            for (int i = 0; i < npix; ++i)
                cols[(size_t)(3 + k) * npix + i] = (float)((pk[i] - p0[i]) * (counts / 0.01));
        }
    }
    std::vector<double> A, b(npix);
    double scale[8];
    int live[8], nl = 0;
    for (int c = 0; c < np_; ++c) {
        double s = 0;
        for (int i = 0; i < npix; ++i) {
            const double v = (double)(float)(cols[(size_t)c * npix + i] * pr.inverr[i]);
            s += v * v;
        }
        scale[c] = std::sqrt(s);
        if (scale[c] > 0) live[nl++] = c;
    }
    if (nl == 0) return false;
    A.resize((size_t)nl * npix);
    for (int j = 0; j < nl; ++j) {
        const int c = live[j];
        for (int i = 0; i < npix; ++i)
            A[(size_t)j * npix + i] = (double)(float)(cols[(size_t)c * npix + i] * pr.inverr[i]) / scale[c];
    }
    const float cnt = (float)counts;
    for (int i = 0; i < npix; ++i) {
        const float mod = p0[i] * cnt;
        b[i] = (double)(float)((pr.data[i] - mod) * pr.inverr[i]);
    }
    double xs[8];
    lstsq(A, npix, nl, b, xs);
    for (int c = 0; c < np_; ++c) X[c] = 0.;
    for (int j = 0; j < nl; ++j) X[live[j]] = xs[j] / scale[live[j]];
    return true;
}

const int NALPHA = 13;
double alpha_of(int i) { return i < 11 ? std::ldexp(1., i - 10) : (i == 11 ? std::sqrt(2.) : 2.); }

// one source's ConstrainedOptimizer loop (oracle fit_source)
int fit_one(const Problem &pr, double *p, double dchisq, int max_iter, double *chi_out,
            Work &W) {
    const int np_ = nparams(pr.kind), npix = pr.P * pr.P;
    const double maxstep[6] = {1., 1., -1., 1., 0.1, 0.1};   // < 0: none
    std::vector<float> p0(npix), pt(npix), cols;
    int it = 0, nit = 0;
    for (it = 0; it < max_iter; ++it) {
        nit = it + 1;
        double X[8];
        if (!update_direction(pr, p, p0.data(), X, W, cols)) {
            if (pr.kind != 0 && !render_galaxy(pr, p, -1, p0.data(), W)) return -1;
            break;
        }
        const double lnp0 = -0.5 * chisq_of(pr, p, p0.data());
        // every step first (getParameterSteps marks stepLimited for any
        // max-step clamp, walked or not), then the walk
        bool step_limited = false;
        double steps[NALPHA][8];
        int ns = 0;
        for (int ai = 0; ai < NALPHA; ++ai) {
            double a = alpha_of(ai);
            bool lim = false;
            for (int i = 0; i < np_; ++i)
                if (maxstep[i] >= 0 && a * std::fabs(X[i]) > maxstep[i]) {
                    a = maxstep[i] / std::fabs(X[i]);
                    lim = true;
                }
            if (lim) step_limited = true;
            if (a < 1e-8) break;   // no bounds on these parameters
            for (int i = 0; i < np_; ++i) steps[ns][i] = p[i] + a * X[i];
            ++ns;
            if (lim) break;
        }
        const bool any = ns > 0;
        double best = lnp0;
        int ib = -1;
        for (int k = 0; k < ns; ++k) {
            if (!render(pr, steps[k], pt.data(), W)) return -1;
            const double lnp = -0.5 * chisq_of(pr, steps[k], pt.data());
            if (!std::isfinite(lnp) || lnp < best - 1.) break;
            if (lnp > best) {
                best = lnp;
                ib = k;
            }
        }
        double *bestp = ib >= 0 ? steps[ib] : nullptr;
        if (!any) step_limited = true;
        const double dlnp = ib >= 0 ? best - lnp0 : 0.;
        if (ib >= 0) std::memcpy(p, bestp, sizeof(double) * np_);
        if (step_limited ? dlnp <= 1e-6 : dlnp <= dchisq) break;
    }
    if (!render(pr, p, p0.data(), W)) return -1;
    *chi_out = chisq_of(pr, p, p0.data());
    return nit;
}

}  // namespace

extern "C" {

// Fit n sources, each on its own P x P window (data / inverr at pix_off,
// row stride P), from start (n x 6; points use the first 3) with the
// ConstrainedOptimizer rules (dchisq, step-limited 1e-6, max_iter), on
// nthreads threads.  Sources are taken in order; after budget_s seconds
// (<= 0: no limit) no thread starts another one beyond its first.  Writes params_out (n x 6), iters_out
// (-1: centre left the window on a galaxy) and chisq_out for the first
// *done sources.  Returns 0, or -1 for an unsupported kind.
int orc_fit_sources(int n, const int32_t *kind, const int32_t *P, const int64_t *pix_off,
                    const float *data, const float *inverr, const double *start,
                    const float *psf, int ph, int pw, const double *cdi, double dchisq,
                    int max_iter, int nthreads, double budget_s, double *params_out,
                    int32_t *iters_out, double *chisq_out, int32_t *done) {
    for (int s = 0; s < n; ++s)
        if (kind[s] < 0 || kind[s] > 2) return -1;
    std::map<int, PsfSpec> specs;
    std::mutex mu;
    std::atomic<int> next(0);
    const auto t0 = std::chrono::steady_clock::now();
    auto worker = [&]() {
        Work W;
        for (bool first = true;; first = false) {
            if (budget_s > 0 && !first) {
                const double el = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
                if (el > budget_s) return;
            }
            const int s = next.fetch_add(1);
            if (s >= n) return;
            Problem pr;
            pr.kind = kind[s];
            pr.P = P[s];
            pr.data = data + pix_off[s];
            pr.inverr = inverr + pix_off[s];
            pr.psf = psf;
            pr.ph = ph;
            pr.pw = pw;
            for (int i = 0; i < 4; ++i) pr.cdi[i] = cdi[i];
            pr.specs = &specs;
            pr.spec_mu = &mu;
            double *p = params_out + (size_t)s * 6;
            for (int i = 0; i < 6; ++i) p[i] = start[(size_t)s * 6 + i];
            double chi = 0.;
            iters_out[s] = fit_one(pr, p, dchisq, max_iter, &chi, W);
            chisq_out[s] = chi;
        }
    };
    if (nthreads < 1) nthreads = 1;
    std::vector<std::thread> th;
    for (int t = 0; t < nthreads; ++t) th.emplace_back(worker);
    for (auto &t : th) t.join();
    *done = std::min(next.load(), n);
    return 0;
}

}  // extern "C"
