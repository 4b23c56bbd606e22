#!/usr/bin/env python
"""Benchmark: batched render-and-fit on one MI355X.

Default workload (config 4): 16384 sources (30% point, 40% Exp, 30% Dev), each
on its own P x P patch (P uniform on 16..128 step 8), 63x63 Moffat PSF
(beta 3.5, FWHM 4 px), noise sigma 1, fit from truth perturbed by 0.5 px /
+-30% flux and r_e / +-0.1 in e until every source meets the reference's
optimize_loop stopping rule.  One step = one complete batched fit from the
start point.  value = sources fit per second.

--workload c2: 4096 sources, one step = one LM iteration (Jacobian renders,
normal equations, FP64 solve, line search); value = source-iterations/s.

--workload c1: one ExpGalaxy in a 64x64 image fit by Tractor.optimize_loop
with the default LsqrOptimizer to dchisq 1e-3, through the drop-in API; one
step = one complete loop; value = fits/s; reports the NullWCS position-step
behaviour beside the same fit with PixelWCS.

--workload c3: forced photometry of 20000 sources on a 2048x2048 image with a
spatially varying PSF; one step = plan + render + normal equations + solve;
value = sources/s.

--workload sersic: the Sersic-galaxy fit that the config-4 line also reports
(4096 sources, the galaxies Sersic with seven parameters).

--workload c5: the config-4 sources fit from CNN proposals (the CNN is first
trained on the GPU on rendered cutouts, outside the timed region); one step =
cutouts + CNN inference + batched fit + fallback; the plain fit of the same
data is timed beside it.

Prints one JSON line.
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ('sources fit/sec at χ² ≤1e-4 rel of FP64; surrogate-init vs plain; '
          'HBM % peak')
HBM_PEAK_GBS = 8000.0
SLOT_NAMES = ['k_render_fft<32> (jacobian)', 'k_render_fft<64> (jacobian)',
              'k_render_fft<128> (jacobian)', 'k_render_fft<256> (jacobian)',
              'k_render_fft<512> (jacobian)',
              'k_render_point (jacobian)', 'k_fit_normal_eq', 'k_fit_solve',
              'k_render_fft<32> (line search chi2)', 'k_render_fft<64> (line search chi2)',
              'k_render_fft<128> (line search chi2)', 'k_render_fft<256> (line search chi2)',
              'k_render_fft<512> (line search chi2)',
              'k_render_point (line search chi2)', 'prep']


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def build_workload(name, torch):
    from tractor_amd import batch as B
    from tractor_amd import workloads as W
    cfg = 4 if name == 'c4' else 2
    n = 16384 if name == 'c4' else 4096
    kind, P, truth = W.draw_sources(n, config=cfg)
    psf = W.moffat_stamp(63, 4.0, 3.5)
    fitter, data, inverr, pix_off = B.make_problem(kind, P, truth, psf, W.seed(cfg, 2))
    start = W.perturb(kind, truth, np.random.default_rng(W.seed(cfg, 1)))
    return dict(cfg=cfg, n=n, kind=kind, P=P, truth=truth, psf=psf, fitter=fitter, data=data,
                inverr=inverr, pix_off=pix_off, start=start)


def tune_launch_shapes(wl):
    """Autotune the render launch shapes (waves per block, G rows per pass)
    of every FFT bucket this workload's patch sizes use, before the fitter
    records its graph; returns {sz: [threads, rows, ns per render]}."""
    from tractor_amd import autotune
    got = autotune.tune_fft(wl['psf'], np.unique(wl['P']))
    log('launch shapes', got)
    return {str(k): [int(v[0]), int(v[1]), round(float(v[2]), 2)] for k, v in got.items()}


def run_step(wl, torch, dchisq=0., graph=True):
    f = wl['fitter']
    if wl['cfg'] == 4:
        f.set_params(wl['start'])
        return f.fit(max_iter=100, dchisq=dchisq, graph=graph)
    f.set_params(wl['start'])
    f.iterate(0., 0.)
    return 1


def cpu_baseline(wl, budget_s=15.0, min_sources=4, dchisq=0.):
    """The FP64 NumPy oracle (single core) fitting the same sources from the
    same start, for about `budget_s` seconds."""
    import oracle as O
    from tractor_amd import batch as B
    data = wl['data'].cpu().numpy()
    inverr = wl['inverr'].cpu().numpy()
    n = 0
    t0 = time.time()
    results = []
    order = np.arange(wl['n'])
    while (time.time() - t0 < budget_s or n < min_sources) and n < wl['n']:
        s = int(order[n])
        Pz, o = int(wl['P'][s]), int(wl['pix_off'][s])
        prob = O.SourceProblem(int(wl['kind'][s]), wl['start'][s], Pz, wl['psf'],
                               data=data[o:o + Pz * Pz].reshape(Pz, Pz),
                               inverr=inverr[o:o + Pz * Pz].reshape(Pz, Pz), cdi=B.PIXEL_CDI)
        if wl['cfg'] == 4:
            p, nit, chi = O.fit_source(prob, wl['start'][s], dchisq=dchisq)
        else:
            p, nit, chi = O.fit_source(prob, wl['start'][s], steps=1)
        results.append((s, p, chi, nit))
        n += 1
    dt = time.time() - t0
    return n, dt, results


def cpu_threads():
    """Host threads for the C++ baseline: every core the process may use,
    at most 16 (a one-GPU share of the host)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def cpu_baseline_cpp(wl, budget_s=10.0, dchisq=0., nthreads=None):
    """The oracle's C++ FP64 restatement (oracle/fit_fp64.cpp) fitting the
    workload's sources in order from the same data and start, threaded over
    sources on nthreads (default cpu_threads()) host threads; no source is
    started after budget_s seconds.  -> (n_done, seconds, params,
    iterations, chisq, threads)."""
    from oracle import fit_fp64 as F
    from tractor_amd import batch as B
    nt = cpu_threads() if nthreads is None else nthreads
    data = wl['data'].cpu().numpy()
    inverr = wl['inverr'].cpu().numpy()
    t0 = time.time()
    done, p, it, chi = F.fit_sources(wl['kind'], wl['P'], wl['pix_off'], data, inverr,
                                     wl['start'], wl['psf'], cdi=B.PIXEL_CDI, dchisq=dchisq,
                                     max_iter=100 if wl['cfg'] == 4 else 1, nthreads=nt,
                                     budget_s=budget_s)
    dt = time.time() - t0
    return done, dt, p[:done], it[:done], chi[:done], nt


def cpu_baseline_record(wl, args, gpu_chisq, gpu_iters):
    """cpu_baseline of the bench line (the C++ FP64 port on the host cores,
    the same port on one thread and the NumPy oracle on one core beside it)
    and the parity of the GPU fit's
    chi^2 against both samples."""
    from oracle import fit_fp64 as F
    unit = 'sources/s' if wl['cfg'] == 4 else 'source-iterations/s'
    m, cdt, cp, cit, cchi, nt = cpu_baseline_cpp(wl, args.cpu_seconds, dchisq=args.dchisq)
    m1, cdt1 = cpu_baseline_cpp(wl, args.cpu_seconds / 4, dchisq=args.dchisq, nthreads=1)[:2]
    n, ndt, res = cpu_baseline(wl, args.cpu_seconds, dchisq=args.dchisq)
    cpu = dict(value=round(m / cdt, 4), unit=unit, cores=nt, kind='port',
               sample='first %d sources of the same workload (same data and start), C++ FP64 '
                      'restatement of the oracle (oracle/fit_fp64.cpp), %d threads over '
                      'sources, %.1f s' % (m, nt, cdt),
               cpu_model=F.cpu_model(), os_cpu_count=os.cpu_count(),
               single_thread=dict(value=round(m1 / cdt1, 4), unit=unit, cores=1,
                                  sample='first %d sources, the same C++ port on one thread, '
                                         '%.1f s' % (m1, cdt1)),
               numpy_port=dict(value=round(n / ndt, 4), unit=unit, cores=1,
                               sample='first %d sources, NumPy FP64 oracle, one process' % n))
    parity = None
    if wl['cfg'] == 4:
        ok = cit > 0
        signed = (gpu_chisq[:m][ok] - cchi[ok]) / cchi[ok]
        e = np.abs(signed)
        en = np.array([abs(gpu_chisq[s] - chi) / chi for (s, p, chi, nit) in res])
        for (s, p, chi, nit) in res:
            if abs(gpu_chisq[s] - chi) / chi > 1e-4:
                log('chi2 parity miss (NumPy oracle): source %d kind %d P %d gpu %.4f oracle %.4f '
                    '(iters %d)' % (s, wl['kind'][s], wl['P'][s], gpu_chisq[s], chi, nit))
        miss = e > 1e-4
        capped = (gpu_iters[:m][ok] >= 100) | (cit[ok] >= 100)
        parity = {'sources': int(ok.sum()), 'reference': 'C++ FP64 port',
                  'max_rel_chisq_vs_fp64': float(e.max()),
                  'frac_within_1e-4': float(np.mean(e <= 1e-4)),
                  'frac_not_worse_than_fp64_by_1e-4': float(np.mean(signed <= 1e-4)),
                  'misses': int(miss.sum()),
                  'misses_at_iteration_cap': int((miss & capped).sum()),
                  # the 100-iteration cap is the reference's own behaviour on
                  # these sources: the FP64 port stops there on (nearly) the
                  # same ones as the GPU fit
                  'gpu_at_iteration_cap': int((gpu_iters[:m][ok] >= 100).sum()),
                  'port_at_iteration_cap': int((cit[ok] >= 100).sum()),
                  'both_at_iteration_cap': int(((gpu_iters[:m][ok] >= 100) &
                                                (cit[ok] >= 100)).sum()),
                  'frac_within_1e-4_converged': float(np.mean(e[~capped] <= 1e-4)),
                  'max_rel_chisq_converged': float(e[~capped].max()) if np.any(~capped) else None,
                  'numpy_oracle': {'sources': len(res), 'max_rel_chisq_vs_fp64': float(en.max()),
                                   'frac_within_1e-4': float(np.mean(en <= 1e-4))}}
    return cpu, parity


def sersic_leg(torch, n=4096, steps=3, n_check=8):
    """Batched fit to convergence of n sources drawn as in the config-2
    workload with every galaxy a SersicGalaxy (index U(0.5, 5); seven
    parameters), from truth perturbed as config 4 plus U(-0.3, 0.3) in the
    index; timed over `steps` fits after one warm-up, with the chi^2 of the
    first n_check sources checked against the FP64 oracle's fit."""
    import oracle as O
    from tractor_amd import batch as B
    from tractor_amd import workloads as W
    kind, P, t6 = W.draw_sources(n, config=2)
    kind = np.where(kind > 0, 3, 0).astype(np.int32)
    rng = np.random.default_rng(W.seed(2, 9))
    truth = np.zeros((n, 7))
    truth[:, :6] = t6
    ser = kind == 3
    truth[ser, 6] = rng.uniform(0.5, 5.0, int(ser.sum()))
    psf = W.moffat_stamp(63, 4.0, 3.5)
    f, data, inverr, pix_off = B.make_problem(kind, P, truth, psf, W.seed(2, 10))
    start = np.zeros((n, 7))
    start[:, :6] = W.perturb(kind, t6, np.random.default_rng(W.seed(2, 11)))
    start[ser, 6] = np.clip(truth[ser, 6] + rng.uniform(-0.3, 0.3, int(ser.sum())), 0.3, 6.2)
    f.set_params(start)
    f.fit(max_iter=100)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        f.set_params(start)
        f.fit(max_iter=100)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / steps
    st = f.status()
    dh, ih = data.cpu().numpy(), inverr.cpu().numpy()
    errs = []
    for s in np.nonzero(ser)[0][:n_check]:
        Pz, o = int(P[s]), int(pix_off[s])
        prob = O.SourceProblem(3, start[s], Pz, psf, data=dh[o:o + Pz * Pz].reshape(Pz, Pz),
                               inverr=ih[o:o + Pz * Pz].reshape(Pz, Pz), cdi=B.PIXEL_CDI)
        _, _, chi = O.fit_source(prob, start[s])
        errs.append(abs(st['chisq'][s] - chi) / chi)
    return {'workload': '%d sources (%d Sersic, index U(0.5, 5); the rest points), P 16..128, '
                        'fit to convergence' % (n, int(ser.sum())),
            'ms_per_fit': round(dt * 1e3, 3), 'sources_per_s': round(n / dt, 1),
            'mean_iterations': round(float(st['iterations'].mean()), 2),
            'max_rel_chisq_vs_fp64_first_%d' % len(errs): float(max(errs))}


def hybrid_leg(torch, steps=3, cpu_seconds=3.0, n=16384):
    """The config-4 sources fitted with a hybrid PSF, the input the
    reference's GPU optimizer requires (HybridPixelizedPSF,
    tractor/psf.py:537-625; gpu_optimizer.py:827): the 63x63 Moffat stamp
    plus the two-Gaussian mixture that HybridPixelizedPSF(gauss=None) fits
    to it (the device EM fit).  Galaxy components too wide for the FFT box go
    to pixel space through that mixture (trh_fitter_set_hybrid).  Data are
    the same hybrid model of the truth plus unit noise; timed over `steps`
    fits after one warm-up, the NumPy FP64 oracle's hybrid fits of the first
    sources (for about cpu_seconds) are the parity check."""
    import oracle as O
    from tractor_amd import batch as B
    from tractor_amd import psf as PS
    from tractor_amd import workloads as W
    kind, P, truth = W.draw_sources(n, config=4)
    psf = W.moffat_stamp(63, 4.0, 3.5)
    amp, mean, var = PS.fit_mog_to_stamps([psf], N=2)[0]
    pix_off, total = W.patch_offsets(P)
    data = torch.zeros(total, dtype=torch.float32, device='cuda')
    inverr = torch.ones_like(data)
    f = B.BatchFitter(kind, P, pix_off, data, inverr, psf)
    f.set_hybrid(amp, mean, var)
    f.set_params(truth)
    f.render_models(data)
    g = torch.Generator(device='cuda')
    g.manual_seed(W.seed(4, 61))
    data.add_(torch.randn(total, generator=g, device='cuda'))
    start = W.perturb(kind, truth, np.random.default_rng(W.seed(4, 62)))
    f.set_params(start)
    f.fit(max_iter=100)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        f.set_params(start)
        f.fit(max_iter=100)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / steps
    st = f.status()
    fl = f.flags()
    f.close()
    out = {'workload': '16384 sources drawn as config 4 (30/40/30, P 16..128), hybrid PSF: '
                       '63x63 Moffat stamp + its EM-fitted 2-Gaussian mixture (amp %s)'
                       % np.round(amp, 4).tolist(),
           'ms_per_fit': round(dt * 1e3, 3), 'sources_per_s': round(n / dt, 1),
           'mean_iterations': round(float(st['iterations'].mean()), 2),
           'at_iteration_cap': int(np.sum(st['iterations'] >= 100)),
           'split_failures': int(np.sum((np.asarray(fl) & 32) != 0)),
           'parity': None}
    if cpu_seconds <= 0:
        return out
    dh = data.cpu().numpy()
    ih = inverr.cpu().numpy()
    hm = (np.asarray(amp), np.asarray(mean), np.asarray(var))
    e = []
    t0 = time.time()
    s = 0
    while s < n and (time.time() - t0 < cpu_seconds or s < 4):
        Pz, o = int(P[s]), int(pix_off[s])
        prob = O.SourceProblem(int(kind[s]), start[s], Pz, psf,
                               data=dh[o:o + Pz * Pz].reshape(Pz, Pz),
                               inverr=ih[o:o + Pz * Pz].reshape(Pz, Pz), cdi=B.PIXEL_CDI,
                               hybrid_mog=hm)
        _, nit, chi = O.fit_source(prob, start[s])
        if nit < 100 and st['iterations'][s] < 100:
            e.append(abs(st['chisq'][s] - chi) / chi)
        s += 1
    e = np.asarray(e)
    out['parity'] = {'sources': s, 'compared_converged': int(len(e)),
                     'reference': 'NumPy FP64 oracle, hybrid renders (galaxy.py:412-455, 510-516)',
                     'frac_within_1e-4': float(np.mean(e <= 1e-4)) if len(e) else None,
                     'max_rel_chisq': float(e.max()) if len(e) else None}
    return out


def tiles_leg(torch, steps=3, cpu_seconds=2.0, n_check=2000):
    """The config-4 sources in place on 64 independent 512 x 512 float32
    tiles, 256 per tile: each source's P x P window at a random position in
    its tile, the tiles holding noise plus every source's model (so a window
    also holds its neighbours' light, which no source's fit models), fitted
    in place (row stride 512) from the perturbed start; timed over `steps`
    fits after one warm-up.  The C++ FP64 port fits the same windows (copied
    out on the host) for about cpu_seconds, and the GPU fit's chi^2 is
    checked against it."""
    from oracle import fit_fp64 as F
    from tractor_amd import batch as B
    from tractor_amd import workloads as W
    n, ntile, side = 16384, 64, 512
    kind, P, truth = W.draw_sources(n, config=4)
    psf = W.moffat_stamp(63, 4.0, 3.5)
    rng = np.random.default_rng(W.seed(4, 50))
    tile = np.arange(n) % ntile
    x0 = rng.integers(0, side - P + 1)
    y0 = rng.integers(0, side - P + 1)
    off = (tile.astype(np.int64) * side * side + y0 * side + x0).astype(np.int64)
    data = torch.zeros(ntile * side * side, dtype=torch.float32, device='cuda')
    inverr = torch.ones_like(data)
    f = B.BatchFitter(kind, P, off, data, inverr, psf, row_stride=side)
    f.set_params(truth)
    f.render_models(data, accumulate=True)
    g = torch.Generator(device='cuda')
    g.manual_seed(W.seed(4, 51))
    data.add_(torch.randn(data.numel(), generator=g, device='cuda'))
    start = W.perturb(kind, truth, np.random.default_rng(W.seed(4, 52)))
    f.set_params(start)
    f.fit(max_iter=100)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        f.set_params(start)
        f.fit(max_iter=100)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / steps
    st = f.status()
    f.close()
    out = {'workload': '16384 sources drawn as config 4, 256 in place on each of 64 512x512 '
                       'tiles (noise + every model; windows overlap), row stride 512',
           'ms_per_fit': round(dt * 1e3, 3), 'sources_per_s': round(n / dt, 1),
           'mean_iterations': round(float(st['iterations'].mean()), 2),
           'at_iteration_cap': int(np.sum(st['iterations'] >= 100)),
           'us_per_source_iteration': round(dt * 1e6 / max(1, int(st['iterations'].sum())), 4),
           'note': 'every window also holds the light of the neighbours overlapping it, '
                   'which no fit models: a harder fit than the headline layout (one source '
                   'per window), so more iterations and more FP32/FP64 trajectories that part',
           'parity': None}
    if cpu_seconds <= 0:
        return out
    # the same windows, copied out for the CPU port
    m = min(n, n_check)
    dh = data.cpu().numpy().reshape(ntile, side, side)
    poff, tot = W.patch_offsets(P[:m])
    packed = np.zeros(tot, np.float32)
    for s in range(m):
        packed[poff[s]:poff[s] + P[s] * P[s]] = \
            dh[tile[s], y0[s]:y0[s] + P[s], x0[s]:x0[s] + P[s]].ravel()
    done, _, cit, cchi = F.fit_sources(kind[:m], P[:m], poff, packed, np.ones(tot, np.float32),
                                       start[:m], psf, cdi=B.PIXEL_CDI, nthreads=cpu_threads(),
                                       budget_s=cpu_seconds)
    ok = cit[:done] > 0
    e = np.abs(st['chisq'][:done][ok] - cchi[:done][ok]) / cchi[:done][ok]
    capped = (st['iterations'][:done][ok] >= 100) | (cit[:done][ok] >= 100)
    out['parity'] = {'sources': int(ok.sum()), 'reference': 'C++ FP64 port on the same windows',
                     'frac_within_1e-4': float(np.mean(e <= 1e-4)),
                     'max_rel_chisq': float(e.max()),
                     'frac_within_1e-4_converged': float(np.mean(e[~capped] <= 1e-4))
                     if np.any(~capped) else None}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=5)
    ap.add_argument('--warmup', type=int, default=1)
    ap.add_argument('--workload', default='c4',
                    choices=['c4', 'c1', 'c2', 'c3', 'c5', 'sersic', 'tiles', 'hybrid'])
    ap.add_argument('--cnn-train', type=int, default=1 << 20,
                    help='c5: number of rendered training cutouts')
    ap.add_argument('--cnn-epochs', type=int, default=4)
    ap.add_argument('--hbm-copy-probe', action='store_true',
                    help='only run three 256 MiB device-to-device copies (HBM counter '
                         'calibration under rocprofv3 --pmc)')
    ap.add_argument('--no-cpu-baseline', action='store_true')
    ap.add_argument('--no-sersic', action='store_true',
                    help='skip the Sersic-galaxy fit reported beside the config-4 line')
    ap.add_argument('--no-slot-profile', action='store_true',
                    help='skip the extra per-launch-timed step after the timed region (its '
                         'event timing and byte counters slow that step down; use this under '
                         'rocprofv3 so the trace holds the timed fit only); roofline is then null')
    ap.add_argument('--no-surrogate', action='store_true',
                    help='skip the surrogate-initialised fit (CNN) reported beside the config-4 '
                         'line')
    ap.add_argument('--no-tiles', action='store_true',
                    help='skip the in-place fit on 512x512 tiles reported beside the config-4 '
                         'line')
    ap.add_argument('--no-hybrid', action='store_true',
                    help='skip the hybrid-PSF fit reported beside the config-4 line')
    ap.add_argument('--no-autotune', action='store_true',
                    help='keep the default render launch shapes')
    ap.add_argument('--cpu-seconds', type=float, default=10.0,
                    help='wall-clock budget of each CPU baseline (C++ port, NumPy oracle)')
    ap.add_argument('--dchisq', type=float, default=0.,
                    help='optimize_loop stopping tolerance on dlnp (reference default 0)')
    args = ap.parse_args()
    if args.gpus != 1:
        print('bench.py measures one GPU only: the batched fit runs on a single '
              'MI355X (multi-GPU sharding is designed, not built); use --gpus 1', file=sys.stderr)
        sys.exit(2)

    import torch
    assert torch.cuda.is_available(), 'bench.py needs a HIP device'
    if args.hbm_copy_probe:
        a = torch.ones(64 << 20, dtype=torch.float32, device='cuda')
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({'copy_probe_bytes_per_copy': a.numel() * 4, 'copies': 3}), flush=True)
        return
    if args.workload == 'c5':
        return main_c5(args, torch)
    if args.workload == 'c1':
        return main_c1(args, torch)
    if args.workload == 'c3':
        return main_c3(args, torch)
    if args.workload == 'tiles':
        leg = tiles_leg(torch, steps=args.steps)
        print(json.dumps({
            'metric': 'in-place batched fit on 512x512 tiles (side measurement)',
            'value': leg['sources_per_s'], 'unit': 'sources/s', 'n_gpus': 1, 'steps': args.steps,
            'warmup': 1, 'ms_per_step': leg['ms_per_fit'], 'higher_is_better': True,
            'scaling': 'weak', 'vs_baseline': None,
            'dtype': 'f32 (renders, J^T W J) + f64 (Cholesky, chi2)', 'data': 'synthetic',
            'config': {'workload': leg['workload']}, 'tiles_fit': leg}), flush=True)
        return
    if args.workload == 'hybrid':
        leg = hybrid_leg(torch, steps=args.steps)
        print(json.dumps({
            'metric': 'hybrid-PSF batched fit (side measurement)', 'value': leg['sources_per_s'],
            'unit': 'sources/s', 'n_gpus': 1, 'steps': args.steps, 'warmup': 1,
            'ms_per_step': leg['ms_per_fit'], 'higher_is_better': True, 'scaling': 'weak',
            'vs_baseline': None, 'dtype': 'f32 (renders, J^T W J) + f64 (Cholesky, chi2)',
            'data': 'synthetic', 'config': {'workload': leg['workload']}, 'hybrid_fit': leg}),
            flush=True)
        return
    if args.workload == 'sersic':
        leg = sersic_leg(torch, steps=args.steps)
        print(json.dumps({
            'metric': 'Sersic-galaxy batched fit (side measurement)', 'value': leg['sources_per_s'],
            'unit': 'sources/s', 'n_gpus': 1, 'steps': args.steps, 'warmup': 1,
            'ms_per_step': leg['ms_per_fit'], 'higher_is_better': True, 'scaling': 'weak',
            'vs_baseline': None, 'dtype': 'f32 (renders, J^T W J) + f64 (Cholesky, chi2)',
            'data': 'synthetic', 'config': {'workload': leg['workload']}, 'sersic_fit': leg}),
            flush=True)
        return
    log('building workload', args.workload)
    wl = build_workload(args.workload, torch)
    shapes = None if args.no_autotune else tune_launch_shapes(wl)
    f = wl['fitter']
    for _ in range(args.warmup):
        run_step(wl, torch, args.dchisq)
    torch.cuda.synchronize()

    iters = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        iters.append(run_step(wl, torch, args.dchisq))
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    dt = (t1 - t0) / args.steps
    # one more step, outside the timed region, with per-kernel HIP-event
    # timing (it synchronises every iteration) for the roofline figures
    prof = None
    if not args.no_slot_profile:
        f.profile(True)
        run_step(wl, torch, args.dchisq, graph=False)
        prof = f.profile_read()
        f.profile(False)
    value = wl['n'] / dt
    st = f.status()
    log('per-step iterations', iters, 'max per-source iterations', int(st['iterations'].max()))
    hist = np.bincount(np.minimum(st['iterations'], 100), minlength=101)
    log('per-source iteration histogram (iterations: sources):',
        {int(i): int(c) for i, c in enumerate(hist) if c})

    if prof is not None:
        slots = {SLOT_NAMES[i]: dict(ms=round(float(prof[i, 0]), 3), launches=int(prof[i, 1]),
                                     jobs=int(prof[i, 2]), gbytes=round(float(prof[i, 3]) / 1e9, 4),
                                     gbs=round(float(prof[i, 3]) / max(float(prof[i, 0]), 1e-9) / 1e6,
                                               1))
                 for i in range(len(SLOT_NAMES)) if prof[i, 1] > 0}
        log('per-slot totals over one profiled step:', json.dumps(slots))

    cpu = None
    parity = None
    if not args.no_cpu_baseline:
        cpu, parity = cpu_baseline_record(wl, args, st['chisq'], st['iterations'])

    out = {
        'metric': METRIC,
        'value': round(value, 3),
        'unit': 'sources/s' if wl['cfg'] == 4 else 'source-iterations/s',
        'n_gpus': 1, 'steps': args.steps, 'warmup': args.warmup,
        'ms_per_step': round(dt * 1e3, 3),
        'higher_is_better': True, 'scaling': 'weak', 'vs_baseline': None,
        'dtype': 'f32 (renders, J^T W J) + f64 (Cholesky, chi2)',
        'data': 'synthetic',
        'config': {
            'workload': ('config4: batched LM fit of 16384 sources (30% point / 40% Exp / 30% Dev, '
                         'P 16..128), 63x63 Moffat PSF, sigma=1, from perturbed start to convergence'
                         if wl['cfg'] == 4 else
                         'config2: one LM iteration over 4096 sources (30/40/30), P 16..128, '
                         '63x63 Moffat PSF, sigma=1'),
            'sources': wl['n'], 'lm_iterations_per_step': float(np.mean(iters)),
            'dchisq': args.dchisq,
            'parallelism': 'single GPU',
            'launch_shapes': shapes,
        },
        'roofline': roofline_of(prof, args.workload) if prof is not None else None,
        'cpu_baseline': cpu,
    }
    if parity is not None:
        out['parity'] = parity
    if wl['cfg'] == 4 and not args.no_surrogate:
        log('surrogate leg')
        plain = f.status()
        out['surrogate'] = surrogate_leg(wl, torch, args, dt, plain)['summary']
    if wl['cfg'] == 4 and not args.no_sersic:
        log('sersic leg')
        out['sersic_fit'] = sersic_leg(torch)
    if wl['cfg'] == 4 and not args.no_hybrid:
        log('hybrid leg')
        out['hybrid_fit'] = hybrid_leg(torch, cpu_seconds=0.0 if args.no_cpu_baseline else 3.0)
    if wl['cfg'] == 4 and not args.no_tiles:
        log('tiles leg')
        out['tiles_fit'] = tiles_leg(torch, cpu_seconds=0.0 if args.no_cpu_baseline else 3.0)
    print(json.dumps(out), flush=True)


SLOT_SZ = {0: 32, 1: 64, 2: 128, 3: 256, 4: 512, 8: 32, 9: 64, 10: 128, 11: 256, 12: 512}
MEAN_K = (0.4 * 6 + 0.3 * 9) / 0.7   # MoG components per galaxy render (40% exp, 30% dev)
FP32_PEAK_TFLOPS = 157.3


def render_flops(sz, K=MEAN_K):
    """Arithmetic of one FFT-path render: MoG Fourier transform (8 flops + 1
    exp per component and frequency), spectrum product, the 2-D inverse FFT
    (2.5 sz^2 log2 sz^2) and the two Lanczos-3 passes (28 per pixel)."""
    nv = sz // 2 + 1
    return K * sz * nv * 9 + 6 * sz * nv + 2.5 * sz * sz * math.log2(sz * sz) + 28 * sz * sz


# kernel (as rocprofv3 names it, template arguments included) of each slot
SLOT_KERNEL = {0: 'k_render_fft<32, 0>', 1: 'k_render_fft<64, 0>', 2: 'k_render_fft<128, 0>',
               3: 'k_render_fft<256, 0>', 4: 'k_render_fft<512, 0>', 5: 'k_render_point<0>',
               6: 'k_fit_normal_eq<6>', 7: 'k_fit_solve', 8: 'k_render_fft<32, 2>',
               9: 'k_render_fft<64, 2>', 10: 'k_render_fft<128, 2>', 11: 'k_render_fft<256, 2>',
               12: 'k_render_fft<512, 2>', 13: 'k_render_point<2>'}
TRAFFIC_FILE = os.path.join(ROOT, 'profiles', 'render_fit_config4_hbm_traffic.json')


def measured_traffic(slot, workload):
    """HBM bytes per launch of a slot's kernel from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE summary of the same workload (copy-calibrated),
    or None when there is none for this workload."""
    if slot not in SLOT_KERNEL or not os.path.exists(TRAFFIC_FILE):
        return None
    with open(TRAFFIC_FILE) as fh:
        tr = json.load(fh)
    if tr.get('workload_key') != workload:
        return None
    for name, v in tr['kernels'].items():
        if SLOT_KERNEL[slot] + '(' in name.replace('void ', '').replace('trh::', ''):
            return dict(bytes=round(v['read_bytes_per_dispatch'] + v['write_bytes_per_dispatch']),
                        source='rocprofv3 PMC FETCH_SIZE + WRITE_SIZE of %s, mean over %d '
                               'dispatches (%s)' % (name.split('(')[0].replace('void ', ''),
                                                    v['dispatches'], tr['workload']))
    return None


def roofline_of(prof, workload):
    """Dominant kernel slot (largest total time over the profiled step) ->
    roofline dict.  Both ceilings are evaluated from the live per-launch HIP
    event timing: HBM (algorithmic bytes per launch / mean launch time, vs
    8 TB/s) and, for an FFT render slot, VALU FP32 (algorithmic flops of
    render_flops, vs the published 157.3 TFLOP/s).  `bound` is the one whose
    fraction of peak is higher, and the top-level figures are that one's."""
    k = int(np.argmax(prof[:, 0]))
    ms, launches, jobs, nbytes = prof[k]
    avg_s = ms / 1e3 / max(launches, 1)
    gbs = nbytes / max(launches, 1) / avg_s / 1e9
    hbm = {'achieved': round(gbs, 2), 'peak': HBM_PEAK_GBS, 'unit': 'GB/s',
           'frac': round(gbs / HBM_PEAK_GBS, 5)}
    valu = None
    if k in SLOT_SZ:
        tf = jobs * render_flops(SLOT_SZ[k]) / (ms / 1e3) / 1e12
        valu = {'achieved': round(tf, 2), 'peak': FP32_PEAK_TFLOPS, 'unit': 'TFLOP/s',
                'frac': round(tf / FP32_PEAK_TFLOPS, 4),
                'frac_of_unpacked_fp32': round(tf / (FP32_PEAK_TFLOPS / 2), 4)}
    bound = 'valu' if valu is not None and valu['frac'] > hbm['frac'] else 'hbm'
    top = valu if bound == 'valu' else hbm
    tr = measured_traffic(k, workload)
    out = {'bound': bound, 'kernel': SLOT_NAMES[k], 'achieved': top['achieved'],
           'peak': top['peak'], 'unit': top['unit'], 'frac': top['frac'],
           'traffic': tr['bytes'] if tr else None,
           'traffic_source': tr['source'] if tr else None,
           'avg_launch_ms': round(avg_s * 1e3, 4), 'launches': int(launches),
           'bytes_per_launch': round(nbytes / max(launches, 1)),
           'jobs_per_launch': round(jobs / max(launches, 1), 1), 'hbm': hbm}
    if valu is not None:
        out['valu'] = valu
    return out


def surrogate_leg(wl, torch, args, plain_dt, plain):
    """Surrogate-initialised fitting of a config-4 workload `wl` whose plain
    fit took `plain_dt` s per fit and ended at `plain` (BatchFitter.status):
    train the CNN on rendered cutouts (reported, not timed), then time the
    surrogate pipeline (cutouts + CNN + batched fit from the proposals +
    re-fit from the standard start of every source whose final chi^2 exceeds
    the plain fit's by more than 1e-4 relative) over args.steps runs."""
    from tractor_amd import surrogate as S
    n = wl['n']
    t0 = time.perf_counter()
    X, T, L = S.make_training_set(args.cnn_train, chunk=16384, seed=20261203 + 5000)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    net = S.SurrogateCNN(max_batch=1024, seed=5)
    hist = net.fit(X, T, L, epochs=args.cnn_epochs, batch=1024, lr=3e-3)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    log('cnn: %d cutouts rendered in %.2fs, %d epochs in %.2fs, loss %.4f -> %.4f' % (
        args.cnn_train, t1 - t0, args.cnn_epochs, t2 - t1, hist[0][1], hist[-1][1]))
    del X, T, L
    torch.cuda.empty_cache()
    kw = dict(plain_kind=wl['kind'], dchisq=args.dchisq, plain_chisq=plain['chisq'])
    args_fit = (net, wl['data'], wl['inverr'], wl['pix_off'], wl['P'], wl['psf'], wl['start'])
    for _ in range(args.warmup):
        S.fit_with_fallback(*args_fit, **kw)
    torch.cuda.synchronize()
    ta = time.perf_counter()
    for _ in range(args.steps):
        res = S.fit_with_fallback(*args_fit, **kw)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - ta) / args.steps
    prof = {}
    if not args.no_slot_profile:
        S.fit_with_fallback(*args_fit, profile=prof, **kw)
    # the same pipeline without the plain fit's chi^2 (what a pipeline that
    # has not run the plain fit can do): fall back on the chi-square
    # self-check instead
    kw_sc = dict(kw, plain_chisq=None)
    S.fit_with_fallback(*args_fit, **kw_sc)
    torch.cuda.synchronize()
    tb = time.perf_counter()
    for _ in range(args.steps):
        res_sc = S.fit_with_fallback(*args_fit, **kw_sc)
    torch.cuda.synchronize()
    dt_sc = (time.perf_counter() - tb) / args.steps
    rel_sc = (res_sc['chisq'] - plain['chisq']) / plain['chisq']
    rel = (res['chisq'] - plain['chisq']) / plain['chisq']
    acc = float(np.mean(res['proposed_kind'] == wl['kind']))
    log('surrogate %.0f sources/s vs plain %.0f; timings %s; iterations %.2f vs %.2f' % (
        n / dt, n / plain_dt, res['timings'], res['iterations'].mean(), plain['iterations'].mean()))
    summary = {
        'sources_per_s': round(n / dt, 3),
        'plain_sources_per_s': round(n / plain_dt, 3),
        'speedup_vs_plain': round(plain_dt / dt, 4),
        'fallback_rule': 'final chi2 > plain fit chi2 * (1 + 1e-4) -> re-fit from the standard start',
        'cnn_train_s': round(t2 - t1, 3), 'cnn_render_s': round(t1 - t0, 3),
        'cnn_train_cutouts': args.cnn_train, 'cnn_epochs': args.cnn_epochs,
        'loss_first_last': [round(hist[0][1], 4), round(hist[-1][1], 4)],
        'type_accuracy': round(acc, 4),
        'mean_iterations_plain': round(float(plain['iterations'].mean()), 3),
        'mean_iterations_surrogate': round(float(res['iterations'].mean()), 3),
        'timings_s': {k: (round(v, 4) if isinstance(v, float) else v)
                      for k, v in res['timings'].items()},
        'frac_not_worse_than_plain_by_1e-4': round(float(np.mean(rel <= 1e-4)), 5),
        'frac_within_1e-4_of_plain': round(float(np.mean(np.abs(rel) <= 1e-4)), 5),
        'selfcheck_fallback': {
            'rule': 'final chi2 > N_pix - N_par + 5 sqrt(2 N_pix) -> re-fit from the standard '
                    'start (no plain fit needed)',
            'sources_per_s': round(n / dt_sc, 3), 'speedup_vs_plain': round(plain_dt / dt_sc, 4),
            'fallback_sources': int(res_sc['timings']['fallback_sources']),
            'mean_iterations': round(float(res_sc['iterations'].mean()), 3),
            'frac_not_worse_than_plain_by_1e-4': round(float(np.mean(rel_sc <= 1e-4)), 5),
            'frac_within_1e-4_of_plain': round(float(np.mean(np.abs(rel_sc) <= 1e-4)), 5)},
    }
    return dict(summary=summary, res=res, dt=dt, prof=prof.get('slots'))


def main_c5(args, torch):
    """Surrogate-initialised fitting of the config-4 workload: the plain
    batched fit is timed first (its final chi^2 is the fallback reference),
    then the surrogate pipeline (surrogate_leg) on the same data, same
    stopping rule."""
    log('building workload c4 for c5')
    wl = build_workload('c4', torch)
    shapes = tune_launch_shapes(wl)
    n = wl['n']
    for _ in range(args.warmup):
        run_step(wl, torch, args.dchisq)
    torch.cuda.synchronize()
    ta = time.perf_counter()
    for _ in range(args.steps):
        run_step(wl, torch, args.dchisq)
    torch.cuda.synchronize()
    plain_dt = (time.perf_counter() - ta) / args.steps
    plain = wl['fitter'].status()
    leg = surrogate_leg(wl, torch, args, plain_dt, plain)
    res, dt = leg['res'], leg['dt']
    cpu = None
    parity = None
    if not args.no_cpu_baseline:
        # the CPU fits start from the standard start and type: the surrogate
        # pipeline's final chi^2 is compared with their converged chi^2
        cpu, parity = cpu_baseline_record(wl, args, res['chisq'], res['iterations'])

    out = {
        'metric': METRIC, 'value': round(n / dt, 3), 'unit': 'sources/s',
        'n_gpus': 1, 'steps': args.steps, 'warmup': args.warmup,
        'ms_per_step': round(dt * 1e3, 3), 'higher_is_better': True, 'scaling': 'weak',
        'vs_baseline': None,
        'dtype': 'bf16 MFMA / f32 acc (CNN); f32 (renders, J^T W J) + f64 (Cholesky, chi2)',
        'data': 'synthetic',
        'config': {
            'workload': 'config5: surrogate-initialised batched fit of the config-4 workload '
                        '(16384 sources, 30/40/30, P 16..128, 63x63 Moffat PSF, sigma=1); CNN '
                        '4x conv3x3/s2 32ch + MLP, trained on %d rendered cutouts (Moffat FWHM '
                        '3-6 px), %d epochs' % (args.cnn_train, args.cnn_epochs),
            'sources': n, 'dchisq': args.dchisq, 'parallelism': 'single GPU',
            'launch_shapes': shapes,
        },
        'roofline': roofline_of(leg['prof'], 'c5') if leg['prof'] is not None else None,
        'cpu_baseline': cpu,
        'surrogate': leg['summary'],
    }
    if parity is not None:
        out['parity'] = parity
    print(json.dumps(out), flush=True)


def main_c1(args, torch):
    """The single-galaxy scene (config 1) through the drop-in API: the data
    are the drop-in's own render of the truth plus the scene's noise; each
    step resets the galaxy to the start and runs Tractor.optimize_loop
    (LsqrOptimizer, dchisq 1e-3).  Timed with the Image default NullWCS
    (whose cd^-1 = 3600 scales every position step down 3600x) and with
    PixelWCS; the FP64 oracle's loop on the same data is the CPU baseline."""
    import tractor_amd as T
    from tractor_amd import workloads as W
    sc = W.config1_scene()
    psf, truth, start = sc['psf'], sc['truth'], sc['start']
    inv = np.ones((64, 64), np.float32)

    def galaxy(p):
        return T.ExpGalaxy(T.PixPos(p[0], p[1]), T.Flux(p[2]), T.EllipseESoft(p[3], p[4], p[5]))
    tim0 = T.Image(data=np.zeros((64, 64), np.float32), inverr=inv, psf=T.PixelizedPSF(psf))
    model = T.Tractor([tim0], [galaxy(truth)]).getModelImage(0)
    data = (model + sc['noise']).astype(np.float32)
    legs = {}
    for name, wcs in (('nullwcs', T.NullWCS), ('pixelwcs', T.PixelWCS)):
        tim = T.Image(data=data, inverr=inv, psf=T.PixelizedPSF(psf), wcs=wcs())
        gal = galaxy(start)
        tr = T.Tractor([tim], [gal])
        tr.freezeParam('images')

        def step():
            gal.setParams(start)
            return tr.optimize_loop(dchisq=1e-3)
        for _ in range(args.warmup):
            step()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            r = step()
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / args.steps
        p = np.array(gal.getParams())
        legs[name] = dict(ms_per_fit=round(dt * 1e3, 3), fits_per_s=round(1. / dt, 3),
                          steps=int(r['steps']), chisq=round(-2. * tr.getLogLikelihood(), 6),
                          params=[round(float(v), 6) for v in p],
                          position_moved_px=[round(float(v), 5) for v in p[:2] - start[:2]],
                          position_error_px=[round(float(v), 5) for v in p[:2] - truth[:2]])
        log(name, legs[name])
    cpu = None
    if not args.no_cpu_baseline:
        import oracle as O
        t0 = time.perf_counter()
        po, so, chio, _ = O.fit_image_lsqr(1, start, psf, data, inv, np.eye(2) * 3600.,
                                           dchisq=1e-3)
        cdt = time.perf_counter() - t0
        cpu = dict(value=round(1. / cdt, 4), unit='fits/s', cores=1, kind='port',
                   sample='one complete loop (NullWCS, %d steps), NumPy FP64 oracle, one '
                          'process; chi2 %.6f vs the drop-in %.6f' % (so + 1, chio,
                                                                       legs['nullwcs']['chisq']))
    lg = legs['nullwcs']
    print(json.dumps({
        'metric': 'single-galaxy LsqrOptimizer.optimize_loop fits/sec (drop-in API)',
        'value': lg['fits_per_s'], 'unit': 'fits/s', 'n_gpus': 1, 'steps': args.steps,
        'warmup': args.warmup, 'ms_per_step': lg['ms_per_fit'], 'higher_is_better': True,
        'scaling': 'weak', 'vs_baseline': None,
        'dtype': 'f32 (renders, normal equations) + f64 (solve, chi2)', 'data': 'synthetic',
        'config': {'workload': 'config1: ExpGalaxy flux 1000, r_e 3 px, EllipseE (0.2, -0.1) as '
                               'EllipseESoft, 64x64 image, 31x31 two-Gaussian PixelizedPSF, '
                               'sigma 1, LsqrOptimizer.optimize_loop(dchisq=1e-3)',
                   'parallelism': 'single GPU'},
        'roofline': None, 'cpu_baseline': cpu,
        'h1': {'nullwcs': legs['nullwcs'], 'pixelwcs': legs['pixelwcs'],
               'note': 'NullWCS scales position derivatives by cd^-1 = 3600, so each position '
                       'step is 3600x too small and the loop runs to its step limit'}}),
        flush=True)


def main_c3(args, torch):
    """Forced photometry of 20000 sources on a 2048x2048 image with a 3x3
    bilinear PSF grid: one step = plan + render unit-flux patches +
    overlap-pair normal equations + FP64 block-preconditioned CG."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spl
    from tractor_amd import forced as F, workloads as W
    size, n = 2048, 20000
    kind, pos, flux, shape = W.draw_forced_scene(n, size, rng=np.random.default_rng(W.seed(3, 0)))
    basis = W.psf_grid_stamps()
    idx, wts = W.bilinear_psf_weights(pos[:, 0], pos[:, 1], size, size)
    inverr = torch.ones((size, size), dtype=torch.float32, device='cuda')
    log('c3: synthesising the image')
    syn = F.BatchForcedPhotometry(torch.zeros_like(inverr), inverr, basis, idx, wts, kind, pos,
                                  shape).plan().render()
    g = torch.Generator(device='cuda')
    g.manual_seed(W.seed(3, 2))
    data = syn.model_image(flux) + torch.randn((size, size), generator=g, device='cuda')
    del syn
    torch.cuda.synchronize()

    # the sources and their PSF weights resident on the device, like the image
    cat = F.source_catalog(kind, pos, shape)
    d_idx = torch.from_numpy(np.ascontiguousarray(idx, np.int32)).cuda()
    d_wts = torch.from_numpy(np.ascontiguousarray(wts, np.float32)).cuda()

    def step():
        fp = F.BatchForcedPhotometry(data, inverr, basis, d_idx, d_wts, None, None, None,
                                     catalog=cat)
        fp.run(tol=1e-12)
        return fp
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        fp = step()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    value = n / dt
    ab = fp.algorithmic_bytes()
    dm = fp.device_ms
    log('c3: %.1f ms per step; host planning %.1f ms; device ms %s; pcg iterations %d, resid %.2e; '
        'nnz %d' % (dt * 1e3, fp.timing['plan_s'] * 1e3, dm, fp.iterations, fp.resid, fp.nnz))
    # the solver against an FP64 direct solve of the same device-built system
    A = sp.csr_matrix((fp.vals.cpu().numpy(), fp.d_col.cpu().numpy(), fp.row_ptr), shape=(n, n))
    bb = fp.b.cpu().numpy()
    dg = A.diagonal()
    A = A + sp.diags((dg == 0).astype(np.float64))
    xd = spl.spsolve(A.tocsc(), bb)
    xg = np.empty(n)
    xg[:] = fp.flux[fp.order]
    sig = 1. / np.sqrt(np.maximum(dg, 1e-300))
    solver_err = float(np.max(np.abs(xg - xd) / sig))
    ferr = (fp.flux - flux) * np.sqrt(np.maximum(fp.iv, 1e-300))
    stages = {'render': (dm['render'], ab['render']),
              'assemble (pair dots + rhs)': (dm['assemble'], ab['assemble']),
              'solve (block inverse + PCG)': (dm['solve'], ab['pcg_iteration'] * fp.iterations)}
    kname = max(stages, key=lambda k: stages[k][0])
    ms, nbytes = stages[kname]
    # FitStats of the fit (optimize.py:233-376), outside the timed step:
    # model image + per-source profile-weighted sums + image chi^2
    fp.fitstats()
    torch.cuda.synchronize()
    tf = time.perf_counter()
    fsx = fp.fitstats()
    torch.cuda.synchronize()
    fitstats_ms = (time.perf_counter() - tf) * 1e3
    # the same scene with a PSFEx-style PSF (degree-2 polynomial over six
    # eigen-PSFs built from the grid), one step after a warm-up
    grid = basis.astype(np.float64)
    mean = grid.mean(0)
    pbases = np.stack([mean, 0.5 * (grid[5] - grid[3]), 0.1 * (grid[4] - mean),
                       0.5 * (grid[7] - grid[1]), 0.05 * (grid[8] - grid[0]),
                       0.1 * (grid[6] - mean)]).astype(np.float32)
    pe = dict(degree=2, x0=size / 2., y0=size / 2., xscale=size / 2., yscale=size / 2.)
    for _ in range(2):
        torch.cuda.synchronize()
        tp = time.perf_counter()
        fpe = F.BatchForcedPhotometry(data, inverr, pbases, None, None, None, None, None, psfex=pe,
                                      catalog=cat)
        fpe.run(tol=1e-12)
        torch.cuda.synchronize()
        psfex_ms = (time.perf_counter() - tp) * 1e3
    del fpe
    # the Gaussian-mixture approximations of the nine grid PSFs
    # (HybridPixelizedPSF(gauss=None): EM, N = 2, 1000 steps), one batch
    from tractor_amd import psf as TP
    TP.fit_mog_to_stamps(list(basis), N=2, emsteps=1000)
    torch.cuda.synchronize()
    te = time.perf_counter()
    TP.fit_mog_to_stamps(list(basis), N=2, emsteps=1000)
    torch.cuda.synchronize()
    em_ms = (time.perf_counter() - te) * 1e3
    achieved = nbytes / (ms / 1e3) / 1e9
    cpu = None
    if not args.no_cpu_baseline:
        import oracle as O
        # bounded sample: the sources of a 256x256 corner of the same scene,
        # rendered and solved by the FP64 oracle on one core
        sel = np.flatnonzero((pos[:, 0] < 256) & (pos[:, 1] < 256))
        stamps = np.einsum('nq,nqij->nij', wts[sel].astype(np.float64),
                           basis[idx[sel]].astype(np.float64))
        dsub = data[:256, :256].cpu().numpy().astype(np.float64)
        tc = time.time()
        patches = [O.forced_unit_patch(int(kind[s]), pos[s, 0], pos[s, 1], shape[s], stamps[j],
                                       (256, 256)) for j, s in enumerate(sel)]
        O.forced_phot_lsq(patches, dsub, np.ones((256, 256)))
        ct = time.time() - tc
        cpu = dict(value=round(len(sel) / ct, 3), unit='sources/s', cores=1, kind='port',
                   sample='the %d sources of a 256x256 corner of the same scene (their patches '
                          'clipped to it), NumPy FP64 oracle renders + sparse FP64 solve, one '
                          'process' % len(sel))
    out = {
        'metric': METRIC, 'value': round(value, 3), 'unit': 'sources/s',
        'n_gpus': 1, 'steps': args.steps, 'warmup': args.warmup,
        'ms_per_step': round(dt * 1e3, 3), 'higher_is_better': True, 'scaling': 'weak',
        'vs_baseline': None, 'dtype': 'f32 (renders, J^T W J products) + f64 (sums, solve)',
        'data': 'synthetic',
        'config': {'workload': 'config3: forced photometry, 2048x2048 image, 20000 sources (60% '
                               'point, 40% exp/dev), 3x3 bilinear grid of 33x33 MoG PSF stamps '
                               '(FWHM 3-5 px), sigma=1', 'sources': n,
                   'parallelism': 'single GPU'},
        'roofline': {'bound': 'hbm', 'kernel': kname, 'achieved': round(achieved, 2),
                     'peak': HBM_PEAK_GBS, 'unit': 'GB/s', 'frac': round(achieved / HBM_PEAK_GBS, 5),
                     'traffic': None, 'stage_ms': {k: round(v[0], 3) for k, v in stages.items()},
                     'stage_gbytes': {k: round(v[1] / 1e9, 4) for k, v in stages.items()}},
        'cpu_baseline': cpu,
        'forced': {'host_plan_ms': round(fp.timing['plan_s'] * 1e3, 2),
                   'pcg_iterations': fp.iterations, 'pcg_rel_residual': fp.resid,
                   'nnz': int(fp.nnz), 'tiles': int(len(fp.tile_start) - 1),
                   'solver_max_err_over_sigma_vs_fp64_direct': solver_err,
                   'flux_pull_rms_vs_truth': float(np.sqrt(np.mean(ferr ** 2))),
                   'fitstats_ms': round(fitstats_ms, 3),
                   'psfex_degree2_step_ms': round(psfex_ms, 3),
                   'psf_em_fit_9_stamps_K2_1000_steps_ms': round(em_ms, 3),
                   'fitstats_median_prochi2_over_pronpix': float(np.median(
                       fsx.prochi2[fsx.pronpix > 0] / fsx.pronpix[fsx.pronpix > 0]))},
    }
    print(json.dumps(out), flush=True)


if __name__ == '__main__':
    main()
