#!/usr/bin/env python3
"""MORESANE hot-path benchmark on one MI355X.

Workload (one "step" = one image through the hot path, inputs resident in HBM):
  4096 x 4096 float32 image (N(0,1) + 200 point sources, seed 1)
  -> IUWT decomposition, scale_count=7, store_smoothed
  -> IUWT recomposition from the 7 detail planes + smoothed plane
  -> linear fft_convolve with the spectrum of an 8192 x 8192 Gaussian PSF (FWHM 6 px),
     precomputed once (timed separately).
value = images/s (steps/s).  Per-op HIP-event timings give algorithmic HBM GB/s:
  decomposition / recomposition: (S+2) * 4 N^2 bytes; fft_convolve (pruned 3-sweep):
  8 N^2 + 3 * 8 P (P/2+1) bytes.
The CPU baseline is the float64 oracle (single core) on a bounded sample: the same
step on a 2048 x 2048 crop (S=7, linear convolution with a 4096 PSF), reported as
4096^2-equivalent images/s (the IUWT time scaled by the pixel ratio 4, the convolution
time by the P^2 log P ratio).

Also in the line: the batched (c3) and served (c4) forms, and c0 -- one plain 256^2 deconvolution beside
the float64 oracle run of it on one core.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-batch] [--no-c0]
"""

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "images/s + p99 latency, batched vs plain HIP at 1e-4 rel err; IUWT/FFT HBM GB/s"
HBM_PEAK_GBS = 8000.0


def log(msg):
    sys.stderr.write(msg + "\n")
    sys.stderr.flush()


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "c1_hbm_traffic.json")


def measured_traffic(op, workload):
    """HBM bytes per operation from the committed rocprofv3 FETCH_SIZE/WRITE_SIZE summary of this
    workload (copy-calibrated), or None when there is none for this configuration."""
    try:
        with open(TRAFFIC_FILE) as f:
            doc = json.load(f)
    except (OSError, ValueError):
        return None
    rec = doc.get("ops", {}).get(op) if doc.get("workload") == workload else None
    return int(round(rec["traffic_bytes"])) if rec else None


def gbs(nbytes, ms):
    return nbytes / (ms * 1e-3) / 1e9


def host_info():
    """The CPU the baselines ran on: model name, usable cores, numpy/scipy versions."""
    import numpy as np
    import scipy
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            model = next((ln.split(":", 1)[1].strip() for ln in f if ln.startswith("model name")), None)
    except OSError:
        pass
    return {"cpu": model, "usable_cores": len(os.sched_getaffinity(0)), "numpy": np.__version__,
            "scipy": scipy.__version__}


def cpu_baseline(img, scales=7, sample_n=2048, psf_fwhm=6.0):
    """The float64 oracle (one core) on the central sample_n^2 crop of the c1 image: one IUWT
    decompose + recompose and one linear convolution with a 2*sample_n Gaussian PSF of the c1 FWHM.
    Returns the timing record and the crop's outputs, which double as the accuracy reference for
    the GPU's full-size outputs away from the crop border."""
    import numpy as np
    from oracle import convolution64 as oconv, iuwt64 as oiuwt
    from pymoresane_amd import synth
    n = img.shape[0]
    c0 = (n - sample_n) // 2
    crop = img[c0:c0 + sample_n, c0:c0 + sample_n]
    psf = synth.make_gaussian_psf(2 * sample_n, psf_fwhm)
    spec = oconv.psf_spectrum(psf)
    t0 = time.perf_counter()
    det, sm = oiuwt.iuwt_decomposition(crop, scales, store_smoothed=True)
    rec = oiuwt.iuwt_recomposition(det, 0, sm)
    t1 = time.perf_counter()
    out = oconv.fft_convolve(crop, spec, "linear")
    t2 = time.perf_counter()
    del rec
    # the IUWT's work grows with the pixels, the convolution's with P^2 log P (P = 2N)
    ratio = (float(n) / sample_n) ** 2
    fft_ratio = ratio * np.log2(2 * n) / np.log2(2 * sample_n)
    dt = (t1 - t0) * ratio + (t2 - t1) * fft_ratio
    rec_ = {"value": round(1.0 / dt, 6), "unit": "images/s", "cores": 1, "kind": "port",
            "sample": "float64 numpy oracle, one step on the central %dx%d crop of the c1 image: IUWT S=%d "
                      "decompose+recompose %.2f s, scaled by the pixel ratio %.0f, + linear fft_convolve P=%d %.2f s, "
                      "scaled by the P^2 log P ratio %.2f, = %.1f s per %dx%d image"
                      % (sample_n, sample_n, scales, t1 - t0, ratio, 2 * sample_n, t2 - t1, fft_ratio, dt, n, n),
            "host": host_info()}
    return rec_, (c0, det, sm, out)


def cpu_baseline_mp(crop, n, scales=7, psf_fwhm=6.0, workers=None):
    """The reference's 'mp' CPU mode restated (oracle/mp_iuwt.py, after pymoresane/iuwt.py:193-382:
    every a trous pass forks `workers` processes over strips of a shared float32 array; the
    convolution stays serial, as in the reference): one c1 step on the same central crop as the
    serial baseline, scaled by the same pixel ratio.  It runs in a child Python process that never
    touches the GPU, since its strip passes fork."""
    import subprocess
    import tempfile
    import numpy as np
    workers = workers or min(16, len(os.sched_getaffinity(0)))
    with tempfile.TemporaryDirectory() as tmp:
        path = os.path.join(tmp, "crop.npy")
        np.save(path, np.ascontiguousarray(crop, dtype=np.float32))
        res = subprocess.run([sys.executable, os.path.abspath(__file__), "--cpu-mp-child", path, str(scales),
                              str(psf_fwhm), str(workers)], capture_output=True, text=True, timeout=900)
    if res.returncode != 0:
        raise RuntimeError("mp baseline failed: %s" % res.stderr[-2000:])
    dt = json.loads(res.stdout.strip().splitlines()[-1])["seconds"]
    m = crop.shape[0]
    ratio = (float(n) / m) ** 2
    return {"value": round(1.0 / dt / ratio, 6), "unit": "images/s", "cores": workers, "kind": "port",
            "sample": "reference 'mp' mode restated: strip-parallel IUWT (%d forked workers per pass, float32 "
                      "shared array) + serial float64 linear fft_convolve P=%d, one step on the central %dx%d "
                      "crop of the c1 image in %.2f s, scaled by the pixel ratio %.0f" % (workers, 2 * m, m, m, dt,
                                                                                          ratio)}


def _cpu_mp_child(path, scales, psf_fwhm, workers):
    import numpy as np
    from oracle import convolution64 as oconv, mp_iuwt
    from pymoresane_amd import synth
    crop = np.load(path, allow_pickle=False)
    spec = oconv.psf_spectrum(synth.make_gaussian_psf(2 * crop.shape[0], psf_fwhm))
    with one_core():
        t0 = time.perf_counter()
        det, sm = mp_iuwt.mp_iuwt_decomposition(crop, scales, store_smoothed=True, core_count=workers)
        mp_iuwt.mp_iuwt_recomposition(det.astype(np.float32), 0, core_count=workers, smoothed_array=sm)
        oconv.fft_convolve(crop, spec, "linear")
        dt = time.perf_counter() - t0
    print(json.dumps({"seconds": dt}))


def c1_accuracy(ref, det_gpu, sm_gpu, conv_gpu, margin=512):
    """Relative L2 of the GPU's full-size c1 outputs against the float64 oracle's crop outputs, on
    the part of the crop at least `margin` pixels from its border (beyond the reach of the crop's
    own border handling: the IUWT at S=7 reaches 382 pixels, the FWHM-6 Gaussian PSF a few tens)."""
    import numpy as np
    c0, det, sm, out = ref
    m = det.shape[-1]
    a, b = margin, m - margin
    g = slice(c0 + a, c0 + b)

    def rel(x, y):
        x = np.asarray(x, np.float64)
        return float(np.linalg.norm(x - y) / max(np.linalg.norm(y), 1e-300))

    det_err = [rel(det_gpu[k][g, g], det[k][a:b, a:b]) for k in range(det.shape[0])]
    return {"vs": "float64 oracle on the central %dx%d crop, compared on its central %dx%d" % (m, m, b - a, b - a),
            "iuwt_detail_rel_l2_max": max(det_err), "iuwt_smoothed_rel_l2": rel(sm_gpu[g, g], sm[a:b, a:b]),
            "fft_convolve_rel_l2": rel(conv_gpu[g, g], out[a:b, a:b])}


MORESANE_KW = dict(loop_gain=0.2, sigma_level=4, major_loop_miter=100, tolerance=0.75, enforce_positivity=True,
                   conv_mode="linear", minor_loop_miter=30, accuracy=1e-6)


def one_core():
    """Context that keeps BLAS/OpenMP pools of the CPU legs at one thread (numpy's FFT, median and
    scipy's labelling are single-threaded already)."""
    try:
        from threadpoolctl import threadpool_limits
        return threadpool_limits(1)
    except ImportError:  # pragma: no cover - threadpoolctl ships with the image
        import contextlib
        return contextlib.nullcontext()


def bench_c0(cpu=True):
    """c0: one plain MORESANE run (256^2 dirty, 512^2 PSF, S=4, gain 0.2, sigma 4, 100 major, tol 0.75,
    positivity, linear) on the GPU, the median of five runs after one warm run, and -- as its CPU baseline -- the
    float64 oracle's run of the same deconvolution on one core, with the GPU result's error against it."""
    import numpy as np
    from pymoresane_amd import synth
    from pymoresane_amd.main import DataImage
    dirty, psf = synth.config_c0()
    kw = dict(MORESANE_KW, scale_count=4)
    DataImage(dirty, psf, None, 1.0, 1.0).moresane(**kw)  # warm: code objects, allocations
    times = []
    for _ in range(5):
        t0 = time.perf_counter()
        img = DataImage(dirty, psf, None, 1.0, 1.0)
        img.moresane(**kw)
        times.append(time.perf_counter() - t0)
    hip_s = float(np.median(times))
    majors = sum(1 for e in img.trace if e[0] == "update")
    rec = {"workload": "c0: 256x256 dirty, 512x512 PSF, moresane S=4, gain 0.2, sigma 4, 100 major, tol 0.75, "
                       "positivity, linear", "hip_s": round(hip_s, 4), "hip_s_min": round(min(times), 4),
           "runs": len(times), "major_iters": majors}
    if cpu:
        from oracle import moresane64 as omain
        ref = omain.OracleImage(dirty, psf)
        with one_core():
            t0 = time.perf_counter()
            ref.moresane(**kw)
            cpu_s = time.perf_counter() - t0

        def rel(a, b):
            a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
            return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))

        def decisions(trace):  # the discrete part of each trace event
            out = []
            for e in trace:
                e = tuple(e)
                out.append(e[:5] if e[0] == "scales" else e[:4] if e[0] == "minor" else
                           (e[0], e[1], bool(e[4])) if e[0] == "update" else e)
            return out

        rec.update({"cpu_baseline": {"seconds": round(cpu_s, 3), "cores": 1, "kind": "port",
                                     "sample": "float64 numpy oracle, the same deconvolution"},
                    "speedup_vs_cpu": round(cpu_s / hip_s, 1), "model_rel_l2": rel(img.model, ref.model),
                    "residual_rel_l2": rel(img.residual, ref.residual),
                    "trace_equal": decisions(img.trace) == decisions(ref.trace)})
    return rec


def pct(vals, q):
    import numpy as np
    return float(np.percentile(np.asarray(vals), q)) if len(vals) else None


def bench_batched(max_batch=32, workers=8):
    """96 ragged pairs (32x256^2, 32x512^2, 24x1024^2, 8x2048^2), S=5: plain (one job at a time) vs
    batched (each size class stepped in lockstep, up to max_batch jobs per batched launch, classes
    concurrent on `workers` streams)."""
    import numpy as np
    from pymoresane_amd import batch, synth
    sides = synth.config_c3_sides()
    log("c3: generating %d pairs" % len(sides))
    pairs = [synth.config_c3_pair(i, s) for i, s in enumerate(sides)]
    kw = dict(MORESANE_KW, scale_count=5)
    jobs = batch.DeviceJobs(pairs, [kw] * len(pairs))
    del pairs
    order = sorted(range(len(sides)), key=lambda i: -sides[i])  # largest first
    # plain: one job at a time on one stream
    t0 = time.perf_counter()
    plain_stats = batch.run_batch(jobs, n_workers=1, max_batch=1, order=order)
    t_plain = time.perf_counter() - t0
    ref = jobs.results()
    jobs.reset()
    t0 = time.perf_counter()
    stats = batch.run_batch(jobs, n_workers=workers, max_batch=max_batch, order=order)
    t_batch = time.perf_counter() - t0
    same = all(np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]) for a, b in zip(jobs.results(), ref))
    lat = [s[2] for s in stats]
    # per size class: the plain run's time on it (jobs back to back) and the span of its jobs in the
    # batched run (classes overlap there, so the spans do not add up to the whole)
    classes = {}
    for side in sorted(set(sides)):
        ids = [i for i, s in enumerate(sides) if s == side]
        t0c = min(stats[i][3] for i in ids)
        classes[str(side)] = {"jobs": len(ids), "plain_s": round(sum(plain_stats[i][2] for i in ids), 3),
                              "batched_span_s": round(max(stats[i][3] + stats[i][2] for i in ids) - t0c, 3)}
    log("c3: plain %.2fs batched %.2fs (x%.2f) identical=%s" % (t_plain, t_batch, t_plain / t_batch, same))
    return {"workload": "c3: 96 ragged pairs 32x256^2+32x512^2+24x1024^2+8x2048^2, PSF 2N, S=5",
            "plain_images_per_s": round(len(sides) / t_plain, 3),
            "batched_images_per_s": round(len(sides) / t_batch, 3),
            "speedup": round(t_plain / t_batch, 3), "max_batch": max_batch, "class_streams": workers,
            "identical_to_alone": bool(same),
            "batched_job_p50_s": round(pct(lat, 50), 4), "batched_job_p99_s": round(pct(lat, 99), 4),
            "classes": classes}


def bench_c3_class(side, max_batch=32, plain=False):
    """One size class of c3 alone (its jobs stepped in lockstep on one stream, or one at a time with
    plain=True): the run that a kernel trace of the batched form profiles."""
    from pymoresane_amd import batch, synth
    sides = synth.config_c3_sides()
    idx = [i for i, s in enumerate(sides) if s == side]
    if not idx:
        raise SystemExit("c3 has no %d^2 class (sides: %s)" % (side, sorted(set(sides))))
    pairs = [synth.config_c3_pair(i, side) for i in idx]
    jobs = batch.DeviceJobs(pairs, [dict(MORESANE_KW, scale_count=5)] * len(pairs))
    t0 = time.perf_counter()
    batch.run_batch(jobs, n_workers=1, max_batch=1 if plain else max_batch)
    dt = time.perf_counter() - t0
    return {"c3_class": side, "jobs": len(idx), "form": "plain" if plain else "batched",
            "max_batch": 1 if plain else max_batch, "seconds": round(dt, 4),
            "images_per_s": round(len(idx) / dt, 3)}


def bench_served(count=300, max_batch=32, rate=20.0, budget=200e9):
    """Poisson arrivals at `rate`/s, sides 512/1024/2048/4096 (50/30/15/5 %), S = log2(side) - 5."""
    import numpy as np
    from pymoresane_amd import batch, synth
    reqs = synth.config_c4_requests(count, rate=rate)
    log("c4: generating %d requests" % count)
    psf_cache = {}
    inputs = []
    for t_arr, side, seed in reqs:
        rng = np.random.default_rng(seed)
        if side not in psf_cache:
            psf_cache[side] = synth.make_psf(2 * side)
        sky = synth.make_sky(side, rng)
        dirty = (synth.linear_convolve(sky, psf_cache[side]) + rng.normal(0.0, 1e-3, (side, side))).astype(np.float32)
        inputs.append((t_arr, side, dirty))
    out = {}
    results = {}
    for label, nw in (("plain", 1), ("served", max_batch)):
        srv = batch.Server(max_batch=nw, budget_bytes=budget)
        handles = []
        t0 = time.perf_counter()
        for t_arr, side, dirty in inputs:
            wait = t_arr - (time.perf_counter() - t0)
            if wait > 0:
                time.sleep(wait)
            handles.append(srv.submit(dirty, psf_cache[side], scale_count=int(np.log2(side)) - 5, **MORESANE_KW))
        srv.drain()
        span = time.perf_counter() - t0
        lat = [h["req"].t_done - h["req"].t_submit for h in handles]
        ok = all(h["req"].status == 0 for h in handles)
        results[label] = [(h["model"], h["resid"]) for h in handles]
        srv.close()
        out[label] = {"images_per_s": round(len(handles) / span, 3), "p50_latency_s": round(pct(lat, 50), 4),
                      "p99_latency_s": round(pct(lat, 99), 4), "makespan_s": round(span, 3), "max_batch": nw,
                      "all_ok": ok}
        log("c4 %s: %.2f img/s p50 %.3fs p99 %.3fs" % (label, len(handles) / span, pct(lat, 50), pct(lat, 99)))
    same = all(np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
               for a, b in zip(results["plain"], results["served"]))
    out["identical_to_alone"] = bool(same)
    out["speedup_images_per_s"] = round(out["served"]["images_per_s"] / out["plain"]["images_per_s"], 3)
    out["workload"] = ("c4: %d Poisson arrivals at %.0f/s, sides 512/1024/2048/4096 (50/30/15/5%%), "
                       "S=log2(side)-5, HBM budget %.0f GB" % (count, rate, budget / 1e9))
    return out


C2_FIXTURE = os.path.join(ROOT, "tests", "golden", "c2_oracle.npz")


def bench_c2(n=2048, stop_scale=6, out_path=None):
    """c2, the precision form: 2048^2 dirty / 4096^2 PSF (2000 points + 50 blobs, noise 1e-4),
    moresane_by_scale(1..6), in the float32 bulk mode and in the float64 mode, each against the float64
    oracle's run of the same case (32 minutes on one core; its trace and model are the fixture
    tests/golden/c2_oracle.npz, the residual rebuilt from that model as dirty - H model)."""
    import numpy as np
    from scipy import signal
    from pymoresane_amd import synth
    from pymoresane_amd.main import DataImage
    dirty, psf = synth.config_c2(n=n)
    fx = np.load(C2_FIXTURE)
    kw = json.loads(str(fx["kw"]))
    model_o = np.zeros(tuple(fx["shape"]))
    model_o.flat[np.cumsum(fx["model_gaps"].astype(np.int64)) - 1] = fx["model_val"]
    resid_o = dirty.astype(np.float64) - signal.fftconvolve(model_o, psf.astype(np.float64))[n:2 * n, n:2 * n]
    trace_o = json.loads(str(fx["trace"]))

    def rel(a, b):
        a = np.asarray(a, np.float64)
        return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))

    def decisions(trace):
        return [tuple(v for v in e if not isinstance(v, float)) for e in trace]

    rec = {"workload": "c2 precision form: %dx%d dirty, %dx%d PSF, moresane_by_scale(1..%d), against the "
                       "float64 oracle run (%d decisions)" % (n, n, 2 * n, 2 * n, stop_scale, len(trace_o))}
    for prec in ("float32", "float64"):
        img = DataImage(dirty, psf, None, 1, 1)
        t0 = time.perf_counter()
        img.moresane_by_scale(start_scale=1, stop_scale=stop_scale, precision=prec, **kw)
        dt = time.perf_counter() - t0
        dg, do = decisions(img.trace), decisions([tuple(e) for e in trace_o])
        first = next((i for i, (a, b) in enumerate(zip(dg, do)) if a != b), None if len(dg) == len(do) else
                     min(len(dg), len(do)))
        rec[prec] = {"seconds": round(dt, 3), "decisions": len(img.trace), "trace_equal": first is None,
                     "first_divergence": first, "model_rel_l2": rel(img.model, model_o),
                     "residual_rel_l2": rel(img.residual, resid_o)}
        log("c2 %s: %.1f s, trace equal %s, model %.2e, residual %.2e" % (
            prec, dt, first is None, rec[prec]["model_rel_l2"], rec[prec]["residual_rel_l2"]))
        if out_path and prec == "float32":
            np.savez(out_path, model=img.model, residual=img.residual,
                     trace=np.array(json.dumps([list(t) for t in img.trace])))
    rec["float64_cost"] = round(rec["float64"]["seconds"] / rec["float32"]["seconds"], 2)
    return rec


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=4096)
    ap.add_argument("--scales", type=int, default=7)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched (c3) and served (c4) forms")
    ap.add_argument("--no-c0", action="store_true", help="skip the plain c0 deconvolution and its CPU baseline")
    ap.add_argument("--c3-class", type=int, metavar="SIDE",
                    help="run only the c3 size class SIDE (256/512/1024/2048) batched, print its line, exit")
    ap.add_argument("--c3-plain", action="store_true", help="with --c3-class: one job at a time")
    ap.add_argument("--cpu-mp-child", nargs=4, metavar=("CROP_NPY", "SCALES", "FWHM", "WORKERS"),
                    help=argparse.SUPPRESS)
    ap.add_argument("--c0-only", action="store_true", help="run only the c0 leg (with its CPU baseline unless "
                                                            "--no-cpu), print its line, exit")
    ap.add_argument("--serve-requests", type=int, default=300)
    ap.add_argument("--serve-rate", type=float, default=20.0, help="Poisson arrival rate of the served form, per s")
    ap.add_argument("--served-only", action="store_true", help="run only the served form (c4) and print its object")
    ap.add_argument("--max-batch", type=int, default=32, help="jobs per lockstep batch (batched/served forms)")
    ap.add_argument("--calibrate-runs", type=int, nargs="?", const=64, metavar="BYTES",
                    help="only run a copy that reads runs of BYTES (default 64) bytes, every other run of a "
                         "512 MiB buffer, 5 times: the HBM counter calibration for short runs")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a 256 MiB float32 copy kernel 5 times (HBM counter calibration)")
    ap.add_argument("--precision", nargs="?", const="", metavar="OUT_NPZ",
                    help="only run the 2048^2 by-scale precision case (float32 and float64 modes against the "
                         "oracle fixture); with OUT_NPZ also write the float32 model/residual/trace there")
    ap.add_argument("--no-c2", action="store_true", help="skip the c2 precision form")
    args = ap.parse_args()
    if args.cpu_mp_child:  # the mp baseline's child process: CPU only
        p_, s_, f_, w_ = args.cpu_mp_child
        _cpu_mp_child(p_, int(s_), float(f_), int(w_))
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the MORESANE hot path runs one image per device and this build "
              "has no multi-GPU path; run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import ctypes
    import numpy as np
    from pymoresane_amd import _lib, synth
    from pymoresane_amd._lib import DeviceArray, Event, check

    lib = _lib.lib()
    if args.calibrate_runs:
        # the HIP runtime's pitched copy: the first RUN bytes of every 2*RUN-byte row of a 512 MiB buffer
        # (one RUN-byte run per row) into a dense 256 MiB buffer, 5 times
        run = args.calibrate_runs
        if run < 4 or run % 4 or (256 << 20) % run:
            raise SystemExit("--calibrate-runs takes a multiple of 4 bytes that divides 256 MiB")
        hip = ctypes.CDLL("libamdhip64.so")
        hip.hipMemcpy2D.argtypes = [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p, ctypes.c_size_t,
                                    ctypes.c_size_t, ctypes.c_size_t, ctypes.c_int]
        hip.hipMemcpy2D.restype = ctypes.c_int
        rows = (256 << 20) // run
        a, b2 = DeviceArray((rows * run // 2,)), DeviceArray((rows * run // 4,))
        a.fill(1.0)
        check(lib.pms_device_synchronize())
        t0 = time.perf_counter()
        for _ in range(5):
            if hip.hipMemcpy2D(b2.ptr, run, a.ptr, 2 * run, run, rows, 3) != 0:  # 3: device to device
                raise RuntimeError("hipMemcpy2D failed")
        check(lib.pms_device_synchronize())
        ms = (time.perf_counter() - t0) * 1e3 / 5
        print(json.dumps({"run_bytes": run, "read_bytes": rows * run, "write_bytes": rows * run, "ms": round(ms, 4)}))
        return
    if args.calibrate_copy:
        count = 64 << 20
        a, b2 = DeviceArray((count,)), DeviceArray((count,))
        a.fill(1.0)
        e0, e1 = Event(), Event()
        e0.record()
        for _ in range(5):
            check(lib.pms_copy_f32(b2.ptr, a.ptr, count, None))
        e1.record()
        ms = e0.elapsed_ms(e1) / 5
        print(json.dumps({"copy_bytes_each_way": 4 * count, "ms": ms, "GBps_read_plus_write": gbs(8 * count, ms)}))
        return
    if args.precision is not None:
        print(json.dumps(bench_c2(out_path=args.precision or None)))
        return
    if args.c3_class:
        print(json.dumps(bench_c3_class(args.c3_class, args.max_batch, args.c3_plain)))
        return
    if args.served_only:
        print(json.dumps(bench_served(args.serve_requests, args.max_batch, rate=args.serve_rate)))
        return
    if args.c0_only:
        print(json.dumps(bench_c0(cpu=not args.no_cpu)))
        return
    n, S = args.n, args.scales
    p = 2 * n
    log("generating c1 inputs %dx%d, PSF %dx%d" % (n, n, p, p))
    img, psf = synth.config_c1(n=n)
    d_img = DeviceArray.from_host(img)
    d_psf = DeviceArray.from_host(psf)
    del psf
    det = DeviceArray((S, n, n))
    sm = DeviceArray((n, n))
    rec = DeviceArray((n, n))
    out = DeviceArray((n, n))
    spec = DeviceArray((lib.pms_fft_spectrum_bytes(p),), np.uint8)
    iw = DeviceArray((lib.pms_iuwt_work_bytes(n, n),), np.uint8)
    fw = DeviceArray((lib.pms_fft_work_bytes(n, p),), np.uint8)
    fw_psf = DeviceArray((lib.pms_fft_work_bytes(p, p),), np.uint8)
    stream = ctypes.c_void_p(0)
    check(lib.pms_stream_create(ctypes.byref(stream)))
    st = stream.value

    def spectrum():
        check(lib.pms_psf_spectrum(d_psf.ptr, p, p, spec.ptr, fw_psf.ptr, st))

    def decompose():
        check(lib.pms_iuwt_decompose(d_img.ptr, n, n, S, 0, det.ptr, sm.ptr, iw.ptr, st))

    def recompose():
        check(lib.pms_iuwt_recompose(det.ptr, S, n, n, 0, sm.ptr, rec.ptr, iw.ptr, st))

    def convolve():
        check(lib.pms_fft_convolve(d_img.ptr, n, p, spec.ptr, out.ptr, fw.ptr, st))

    def step():
        decompose()
        recompose()
        convolve()

    def timed(fn, k):
        e0, e1 = Event(), Event()
        e0.record(st)
        for _ in range(k):
            fn()
        e1.record(st)
        return e0.elapsed_ms(e1) / k

    spectrum()
    check(lib.pms_stream_synchronize(st))
    spec_ms = timed(spectrum, 3)
    for _ in range(args.warmup):
        step()
    check(lib.pms_stream_synchronize(st))
    log("warmup done; timing %d steps" % args.steps)

    # whole-step throughput: barrier (sync) on both sides of exactly K steps
    check(lib.pms_stream_synchronize(st))
    t0 = time.perf_counter()
    ms_step_ev = timed(step, args.steps)
    check(lib.pms_stream_synchronize(st))
    wall_ms = (time.perf_counter() - t0) * 1e3 / args.steps
    log("step %.3f ms (events), %.3f ms (wall)" % (ms_step_ev, wall_ms))

    # per-op averages on the same stream
    k_op = max(5, args.steps // 2)
    ms_dec = timed(decompose, k_op)
    ms_rec = timed(recompose, k_op)
    ms_conv = timed(convolve, k_op)
    check(lib.pms_stream_synchronize(st))

    plane = 4 * n * n
    b_iuwt = (S + 2) * plane
    # row-walk planes recompose detail scales two per launch, paired from the finest up
    # (at this size every pair qualifies)
    n_pairs = S // 2 if n * n > 256 * 256 and os.environ.get("PMS_IUWT_PAIR", "2") != "0" else 0
    b_conv = 8 * n * n + 3 * 8 * p * (p // 2 + 1)
    ops = {
        "iuwt_decompose": {"ms": round(ms_dec, 4), "alg_bytes": b_iuwt, "GBps": round(gbs(b_iuwt, ms_dec), 1),
                           "launches": S},
        "iuwt_recompose": {"ms": round(ms_rec, 4), "alg_bytes": b_iuwt, "GBps": round(gbs(b_iuwt, ms_rec), 1),
                           "launches": S - n_pairs},
        "fft_convolve": {"ms": round(ms_conv, 4), "alg_bytes": b_conv, "GBps": round(gbs(b_conv, ms_conv), 1),
                         "launches": 3},
        "psf_spectrum_once": {"ms": round(spec_ms, 4)},
    }
    for k, v in ops.items():
        if "GBps" in v:
            v["frac_of_8TBps"] = round(v["GBps"] / HBM_PEAK_GBS, 4)
    # Bytes the launches as structured must move at least (each launch reads its input planes and
    # writes its outputs once): a one-scale decomposition launch reads C_s and writes C_s+1 and w_s;
    # a two-scale recomposition launch reads R, w_a, w_b and writes R', a one-scale one 3 planes.
    # Against these, GBps shows how close the kernels run to the bandwidth; against alg_bytes, what
    # one launch per scale (or pair) costs relative to a fully fused transform.
    for k, nb in (("iuwt_decompose", 3 * S * plane),
                  ("iuwt_recompose", (4 * n_pairs + 3 * (S - 2 * n_pairs)) * plane)):
        ms_k = ops[k]["ms"]
        ops[k]["launch_bytes"] = nb
        ops[k]["launch_GBps"] = round(gbs(nb, ms_k), 1)
    # The per-major-iteration noise/extraction operations on the same S detail planes (not part of the
    # step): MAD thresholds (algorithmic: read the planes once), thresholding (read + write), source
    # extraction of the thresholded planes (read, write the extracted planes and the byte mask).
    thr = DeviceArray((S,), np.float64)
    tpl = DeviceArray((S, n, n))
    ext = DeviceArray((S, n, n))
    msk = DeviceArray((S, n, n), np.uint8)
    tw_ = DeviceArray((lib.pms_threshold_work_bytes(S, n, n),), np.uint8)
    xw_ = DeviceArray((lib.pms_extract_work_bytes(S, n, n),), np.uint8)

    def est():
        check(lib.pms_estimate_threshold(det.ptr, S, n, n, 0, 0, thr.ptr, tw_.ptr, st))

    def app():
        check(lib.pms_apply_threshold(det.ptr, S, n, n, thr.ptr, 4.0, tpl.ptr, tw_.ptr, st))

    def extract():
        check(lib.pms_source_extraction(tpl.ptr, S, n, n, 0.75, 0, ext.ptr, msk.ptr, None, xw_.ptr, st))

    est()
    app()
    extract()
    check(lib.pms_stream_synchronize(st))
    for name, fn, nb in (("estimate_threshold", est, S * plane), ("apply_threshold", app, 2 * S * plane),
                         ("source_extraction", extract, 2 * S * plane + S * n * n)):
        ms_k = timed(fn, k_op)
        ops[name] = {"ms": round(ms_k, 4), "alg_bytes": nb, "GBps": round(gbs(nb, ms_k), 1),
                     "frac_of_8TBps": round(gbs(nb, ms_k) / HBM_PEAK_GBS, 4), "step": False}
    check(lib.pms_stream_synchronize(st))
    for buf in (thr, tpl, ext, msk, tw_, xw_):
        buf.free()
    # the roofline line reports the step operation furthest below the HBM roofline (lowest fraction);
    # every step operation's fraction is listed beside it
    step_ops = ("iuwt_decompose", "iuwt_recompose", "fft_convolve")
    dom_name = min(step_ops, key=lambda k: ops[k]["frac_of_8TBps"])
    dom = ops[dom_name]

    # correctness of the measured step: perfect reconstruction and float64 spot check of the convolution
    recon_err = float(np.linalg.norm(rec.get().astype(np.float64) - img) / np.linalg.norm(img))
    log("recomposition rel err %.2e" % recon_err)

    value = 1000.0 / ms_step_ev
    res = {
        "metric": METRIC, "value": round(value, 3), "unit": "images/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms_step_ev, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "c1 kernel microbench: %dx%d IUWT S=%d decompose+recompose (store_smoothed) + "
                               "linear fft_convolve with %dx%d PSF spectrum" % (n, n, S, p, p),
                   "image": n, "scale_count": S, "psf": p, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "kernel": dom_name, "achieved": dom["GBps"], "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(dom["GBps"] / HBM_PEAK_GBS, 4),
                     "traffic": measured_traffic(dom_name, {"image": n, "scale_count": S, "psf": p}),
                     "alg_bytes": dom["alg_bytes"], "launches": dom["launches"],
                     "step_ops_frac": {k: ops[k]["frac_of_8TBps"] for k in step_ops}},
        "ops": ops,
        "wall_ms_per_step": round(wall_ms, 4),
        "recon_rel_err": recon_err,
    }
    gpu_out = None if args.no_cpu else (det.get(), sm.get(), out.get())
    check(lib.pms_stream_destroy(st))
    for buf in (d_img, d_psf, det, sm, rec, out, spec, iw, fw, fw_psf):
        buf.free()
    if not args.no_batch:
        res["batched"] = bench_batched(args.max_batch)
        res["served"] = bench_served(args.serve_requests, args.max_batch, rate=args.serve_rate)
    if not args.no_c0:
        log("c0 plain deconvolution%s" % ("" if args.no_cpu else " and its float64 oracle run (1 core)"))
        res["c0"] = bench_c0(cpu=not args.no_cpu)
    if not args.no_c2:
        log("c2 precision form, float32 and float64 modes")
        res["c2"] = bench_c2()
    if not args.no_cpu:
        log("cpu baseline (float64 oracle, 1 core, 2048^2 sample)")
        with one_core():
            res["cpu_baseline"], ref = cpu_baseline(img, S)
        log("cpu baseline, reference 'mp' mode (strip-parallel IUWT), same sample")
        c0_ = ref[0]
        res["cpu_baseline_mp"] = cpu_baseline_mp(img[c0_:c0_ + ref[1].shape[-1], c0_:c0_ + ref[1].shape[-1]], n, S)
        res["c1_error"] = c1_accuracy(ref, *gpu_out)
        del ref, gpu_out
    print(json.dumps(res))


if __name__ == "__main__":
    main()
