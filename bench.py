#!/usr/bin/env python3
"""Benchmark of the batched small complex FFT on one MI355X.

Prints ONE JSON line.  The headline workload is the FP64 forward transform at
N = 4096 over 16384 signals (1 GiB of complex128 in, 1 GiB out), inputs
N(0,1) re/im from Philox seed 7, run by the library's best FP64 path
(``fftoc_fft``).  A step is one launch over the whole batch; inputs are
resident in HBM before the timed region starts.  ``extra`` carries the same
measurement for the other configurations (N = 64 x 65536, the N = 8..512
sweep at 256 MiB per run, and the MFMA / radix paths side by side).

Roofline: algorithmic bytes per launch = 32 * N * batch (16 B read + 16 B
written per point) over the average launch duration = time between HIP events
recorded on the launch stream around the K timed launches, divided by K;
peak = 8 TB/s (MI355X HBM3E).
"""
from __future__ import annotations

import argparse
import json
import math
import os
import statistics
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBPS = 8000.0
FP64_PEAK_TFLOPS = 78.6
MALL_BYTES = 256 << 20
RAMP_MS = 200.0
FIRST_RAMP_MS = 1000.0  # the process's first workload starts from idle clocks
_ramped = False


def gen_signals(n: int, batch: int, seed: int, kind: str) -> np.ndarray:
    g = np.random.Generator(np.random.Philox(seed))
    x = np.empty((batch, n), dtype=np.complex128)
    v = x.view(np.float64)
    if kind in ("normal", "stress"):
        v[...] = g.standard_normal(v.shape)
        if kind == "stress":  # dynamic-range stress: 1% of the entries x 1e4
            x[g.random((batch, n)) < 0.01] *= 1e4
    else:
        v[...] = g.uniform(-1.0, 1.0, v.shape)
    return x


class Workload:
    def __init__(self, name, n, batch, seed, kind, path, flags=0):
        self.name, self.n, self.batch, self.seed, self.kind, self.path = name, n, batch, seed, kind, path
        self.flags = flags


def run_gpu(fo, wl: Workload, steps: int, warmup: int, host_x: np.ndarray | None = None):
    """Times `steps` launches of one path over the resident batch."""
    from fft_onchip_amd import hip

    n, batch = wl.n, wl.batch
    c64 = (wl.flags & fo.DATA_C64) != 0  # complex64 data (8 B per point)
    es = 8 if c64 else 16
    nbytes = n * batch * es
    x = gen_signals(n, batch, wl.seed, wl.kind) if host_x is None else host_x
    if c64:
        x = x.astype(np.complex64)
    # enough in/out pairs that the rotated footprint is >= 4x the 256 MB MALL
    pairs = max(1, min(8, math.ceil(4 * MALL_BYTES / (2 * nbytes))))
    ins = [hip.DeviceBuffer.from_array(x) for _ in range(pairs)]
    outs = [hip.DeviceBuffer(nbytes) for _ in range(pairs)]
    stream = hip.Stream()
    if wl.path.startswith("variant:"):  # one named kernel variant
        vidx = fo.variant_index(n, wl.path[len("variant:"):], c64=c64)

        def fn(x_, b_, n_, out=None, flags=0, stream=None):
            fo.fft_variant(x_, b_, n_, vidx, out=out, flags=flags, stream=stream)
    else:
        fn = {"best": fo.fft, "tensor": fo.tensor_fft, "legacy8": fo.legacy_fft,
              "legacy16": fo.legacy16_fft}[wl.path]
    # Clock ramp: MI355X needs ~100 ms of sustained load to reach its
    # clocks (longer from idle); launches continue for the ramp time before
    # the W warm-up steps.
    global _ramped
    ramp_ms = RAMP_MS if _ramped else FIRST_RAMP_MS
    _ramped = True
    ramp = 0
    t_ramp = time.perf_counter()
    while (time.perf_counter() - t_ramp) * 1e3 < ramp_ms:
        for _ in range(8):
            fn(ins[ramp % pairs], batch, n, out=outs[ramp % pairs], flags=wl.flags, stream=stream)
            ramp += 1
        stream.synchronize()
    for i in range(warmup):
        fn(ins[i % pairs], batch, n, out=outs[i % pairs], flags=wl.flags, stream=stream)
    stream.synchronize()
    t_start, t_end = hip.Event(), hip.Event()
    hip.synchronize()
    wall0 = time.perf_counter()
    t_start.record(stream)
    for i in range(steps):
        fn(ins[i % pairs], batch, n, out=outs[i % pairs], flags=wl.flags, stream=stream)
    t_end.record(stream)
    stream.synchronize()
    hip.synchronize()
    wall = time.perf_counter() - wall0
    total_ms = t_start.elapsed_ms(t_end)
    # separate pass with an event pair around each launch: spread only
    ev = [(hip.Event(), hip.Event()) for _ in range(min(steps, 10))]
    for i, (a, b) in enumerate(ev):
        a.record(stream)
        fn(ins[i % pairs], batch, n, out=outs[i % pairs], flags=wl.flags, stream=stream)
        b.record(stream)
    stream.synchronize()
    launch_ms = [a.elapsed_ms(b) for a, b in ev]
    result = outs[(steps - 1) % pairs]
    sample = result.download((min(batch, 256), n), dtype=np.complex64 if c64 else np.complex128)
    for b in ins + outs:
        b.free()
    # average launch duration = HIP events around the timed region / K (the
    # per-launch event pairs add ~2 us of their own to a 25 us launch)
    mean_launch_ms = total_ms / steps
    ffts_per_s = batch * steps / (total_ms * 1e-3)
    alg_bytes = 2.0 * es * n * batch
    achieved = alg_bytes / (mean_launch_ms * 1e-3) / 1e9
    flops = 5.0 * n * math.log2(n) * batch
    return {
        "workload": wl.name, "n": n, "batch": batch, "path": wl.path,
        "ffts_per_s": ffts_per_s, "ms_per_step": total_ms / steps,
        "launch_ms_mean": mean_launch_ms,
        "single_launch_ms_median": statistics.median(launch_ms),
        "single_launch_ms_min": min(launch_ms), "wall_s": wall,
        "GBps": achieved, "hbm_frac": achieved / HBM_PEAK_GBPS,
        "GFLOPps": flops / (mean_launch_ms * 1e-3) / 1e9,
        "rotation_pairs": pairs, "clock_ramp_launches": ramp, "clock_ramp_ms": ramp_ms,
    }, x, sample


# The reference's on-chip rows (scripts/results/configs.csv): microseconds per
# in-shared-memory pass over 32 FFTs per SM on a 108-SM GPU (3456 FFTs), best
# matrix-core functor per length, cuFFTDx at N = 4096 (no matrix-core row).
ONCHIP_REF = {8: (19, "MMA FPU=8 UPB=2", 0.2355),
              64: (195, "MMA FPU=1 UPB=32", 1.7070),
              128: (236, "MMA FPU=1 UPB=16", 3.8533),
              256: (261, "MMA FPU=1 UPB=8", 8.3804),
              512: (267, "MMA FPU=2 UPB=1", 35.714),
              4096: (274, "cuFFTDx", 253.27)}
ONCHIP_REF_FFTS = 32 * 108
ONCHIP_FAMILIES = {0: "radix (fftoc::LdsFft)", 1: "tensor Karatsuba (fftoc::LdsTensorFft)",
                   2: "tensor two-product (fftoc::LdsTensorFft<.., kTwoProduct>)",
                   3: "tensor Karatsuba, one matrix stage (fftoc::LdsTensorFft<.., kKaratsubaOneStage>)"}


def onchip_library(fo):
    """The example library's on-chip harness entry point and the device's
    FP64 twiddle table."""
    import ctypes

    lib = ctypes.CDLL(os.path.join(ROOT, "fft_onchip_amd", "libfftoc_fused_example.so"))
    run = lib.fftoc_example_onchip_fft_shape
    run.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int, ctypes.c_void_p, ctypes.c_int,
                    ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_void_p]
    run.restype = ctypes.c_int
    tw = ctypes.c_void_p()
    if fo.lib().fftoc_twiddle_table(ctypes.byref(tw), None) != 0:
        raise RuntimeError("fftoc_twiddle_table failed")
    return run, tw


def onchip_pass_us(run, tw, stream, xs: np.ndarray, fam: int, reps: int, shape: int = 0) -> float:
    """Microseconds of one in-LDS pass over the signals xs (batch, n):
    (t[R=1100] - t[R=100]) / 1000, medians of `reps` launches each.  shape 0:
    the example's workgroup shapes (up to 256 threads); 1: one-wave workgroups
    with the fewest signals a functor unit takes."""
    from fft_onchip_amd import hip

    batch, n = xs.shape
    buf = hip.DeviceBuffer.from_array(np.ascontiguousarray(xs))

    def timed(r):
        a, b = hip.Event(), hip.Event()
        a.record(stream)
        rc = run(buf.ptr, batch, n, tw, fam, shape, r, 1.0 / math.sqrt(n), stream.handle)
        b.record(stream)
        stream.synchronize()
        if rc != 0:
            raise RuntimeError(f"fftoc_example_onchip_fft n={n} family={fam}: {rc}")
        return a.elapsed_ms(b)

    timed(100)  # warm-up (clocks, code load)
    t100 = statistics.median(timed(100) for _ in range(reps))
    t1100 = statistics.median(timed(1100) for _ in range(reps))
    buf.free()
    return (t1100 - t100) * 1e3 / 1000.0


# The reference's saturation sweep (scripts/results/saturation.csv, from
# src/saturation.cu:15-62): microseconds per on-chip pass over k FFTs per SM
# on 108 SMs, matrix-core functor and cuFFTDx; N -> {k: (line, tensor us,
# cuFFTDx us)}.
SATURATION_REF = {
    8: {4: (2, 0.959488, 1.06394), 16: (5, 0.95744, 1.06189), 64: (17, 0.96256, 1.64454),
        256: (65, 1.41517, 4.82099)},
    64: {4: (67, 0.390144, 1.55443), 16: (70, 1.21037, 1.54931), 64: (82, 4.5865, 4.51891),
         256: (130, 18.0838, 17.4725)},
    128: {4: (132, 1.0537, 1.6087), 16: (135, 2.0265, 2.56614), 64: (147, 7.83667, 8.67533),
          256: (195, 31.0579, 33.6507)},
    256: {4: (197, 2.23232, 1.66605), 16: (200, 4.3479, 4.58957), 64: (212, 16.9667, 17.1561),
          256: (260, 67.542, 67.4499)},
    512: {4: (262, 5.29408, 2.77914), 16: (265, 18.3081, 9.05421), 64: (277, 71.3052, 34.902),
          256: (325, 283.205, 138.338)},
}


def run_saturation(fo, reps: int = 3):
    """The reference's saturation sweep (src/saturation.cu:15-62) on this
    GPU: on-chip rate of the radix, Karatsuba and two-product matrix-core
    functors at k FFTs per CU, k = 4, 16, 64, 256, each at the better of two
    workgroup shapes, beside the reference's rows at the same k per SM (its
    FFT/s = k * 108 / time)."""
    from fft_onchip_amd import hip

    run, tw = onchip_library(fo)
    cus = hip.cu_count()
    stream = hip.Stream()
    out = {"method": "on-chip pass time as extra.onchip; k FFTs per CU (k * CUs signals)",
           "cus": cus, "rows": {}}
    for n, ref in SATURATION_REF.items():
        g = np.random.Generator(np.random.Philox(41 + n))
        kmax = max(ref)
        xs = g.uniform(-1, 1, (kmax * cus, n)) + 1j * g.uniform(-1, 1, (kmax * cus, n))
        rows = {}
        for k, (line, t_tensor, t_dx) in ref.items():
            r = {"reference": {"row": f"scripts/results/saturation.csv:{line}",
                               "tensor_ffts_per_s": k * 108 / (t_tensor * 1e-6),
                               "cufftdx_ffts_per_s": k * 108 / (t_dx * 1e-6)}}
            for fam in ((0, 1) if n == 8 else (0, 1, 2)):  # two-product from N = 16
                # both workgroup shapes; at small k the one-wave shape spreads
                # the signals over every CU, as the reference's 2-FFT blocks do
                rates = {}
                for shape in (0, 1):
                    us = onchip_pass_us(run, tw, stream, xs[:k * cus], fam, reps, shape)
                    rates[shape] = k * cus / (us * 1e-6) if us > 0 else float("nan")
                best_shape = max(rates, key=lambda q: rates[q])
                r[f"family{fam}_ffts_per_s"] = rates[best_shape]
                r[f"family{fam}_shape"] = "one wave" if best_shape else "default"
            best_ref = max(r["reference"]["tensor_ffts_per_s"], r["reference"]["cufftdx_ffts_per_s"])
            r["best_vs_reference"] = round(max(v for q, v in r.items() if q.endswith("_ffts_per_s")
                                               and q.startswith("family")) / best_ref, 3)
            rows[str(k)] = r
        out["rows"][str(n)] = rows
        print(f"[bench] saturation n={n}: " + ", ".join(
            f"k={k} {v['best_vs_reference']}x" for k, v in rows.items()), file=sys.stderr, flush=True)
    return out


def run_per_cu(fo, rounds: int = 3, launches: int = 20):
    """Workgroups per CU as a measured dimension (the reference's blocks-per-
    SM sweep, src/saturation.cu:15-62, here over HBM-resident batches): the
    variant fftoc_fft picks for each workload, at its occupancy and at every
    smaller cap (FFTOC_WGS_PER_CU), back-to-back launches over buffers that
    rotate past the 256 MB last-level cache; median over rounds that visit
    the caps in turn.  `default_cap` is the cap fftoc_fft applies."""
    from fft_onchip_amd import hip

    lib = fo.lib()
    stream = hip.Stream()
    e0, e1 = hip.Event(), hip.Event()
    out = {}
    for n, batch in ((4096, 16384), (1024, 16384), (512, 32768), (256, 65536), (64, 65536)):
        nbytes = n * batch * 16
        pairs = max(1, min(8, math.ceil(4 * MALL_BYTES / (2 * nbytes))))
        x = gen_signals(n, batch, 3, "uniform")
        ins = [hip.DeviceBuffer.from_array(x) for _ in range(pairs)]
        outs = [hip.DeviceBuffer(nbytes) for _ in range(pairs)]
        v = lib.fftoc_get_variant(n, 0, batch)
        occ = lib.fftoc_variant_occupancy(n, 0, v)
        caps = list(range(0, min(occ, 16)))
        gbps = {c: [] for c in caps}
        for _ in range(rounds):
            for c in caps:
                fl = fo.make_flags(per_cu=c)
                for i in range(3):
                    fo.fft_variant(ins[i % pairs], batch, n, v, out=outs[i % pairs], flags=fl,
                                   stream=stream)
                e0.record(stream)
                for i in range(launches):
                    fo.fft_variant(ins[i % pairs], batch, n, v, out=outs[i % pairs], flags=fl,
                                   stream=stream)
                e1.record(stream)
                e1.synchronize()
                gbps[c].append(2 * nbytes * launches / (e0.elapsed_ms(e1) * 1e-3) / 1e9)
        for b in ins + outs:
            b.free()
        out[f"fp64_n{n}_b{batch}"] = {
            "variant": fo.variants(n)[v], "occupancy": occ,
            "default_cap": lib.fftoc_get_workgroups_per_cu(n, 0, batch),
            "GBps": {("occupancy" if c == 0 else str(c)): round(statistics.median(t), 1)
                     for c, t in gbps.items()}}
        print(f"[bench] per-CU n={n}: {out[f'fp64_n{n}_b{batch}']}", file=sys.stderr, flush=True)
    return out


def run_onchip(fo, reps: int = 3):
    """On-chip timing with the reference's method (includes/testing.cuh:94-127):
    one launch loads the batch into LDS, runs the in-LDS functor R times and
    stores once; R = 100 and R = 1100 are timed with HIP events on the launch
    stream and (t1100 - t100) / 1000 is the time of one pass over the batch,
    HBM traffic excluded.  Every functor family at N = 8, 64, 128, 256, 512,
    4096; batches of 32 FFTs per CU (the reference's shape), and one and four
    LDS-fulls of signals per CU (160 KiB = 10240 points each), each with the
    example's default workgroups and with one-wave workgroups; the best is
    reported as the saturated rate."""
    from fft_onchip_amd import hip

    run, tw = onchip_library(fo)
    cus = hip.cu_count()
    stream = hip.Stream()
    out = {"method": "(t[R=1100] - t[R=100]) / 1000 per pass; HIP events on the launch stream; "
                     f"median of {reps}", "cus": cus, "rows": {}}
    shapes = {"32_per_cu": lambda n: 32 * cus, "lds_x1": lambda n: max(1, 10240 * cus // n),
              "lds_x4": lambda n: max(1, 4 * 10240 * cus // n)}
    for n, (row, cfg, ref_us) in ONCHIP_REF.items():
        ref_rate = ONCHIP_REF_FFTS / (ref_us * 1e-6)
        entry = {"reference": {"row": f"scripts/results/configs.csv:{row}",
                               "config": cfg, "us_per_pass": ref_us, "ffts_per_pass": ONCHIP_REF_FFTS,
                               "ffts_per_s": ref_rate}}
        flops = 5.0 * n * math.log2(n)
        g = np.random.Generator(np.random.Philox(31 + n))
        big = max(fb(n) for fb in shapes.values())
        xs = g.uniform(-1, 1, (big, n)) + 1j * g.uniform(-1, 1, (big, n))
        for fam, fam_name in ONCHIP_FAMILIES.items():
            if fam == 2 and n == 8:
                continue  # the two-product contraction starts at N = 16
            if fam == 3 and n not in (256, 4096):
                continue  # one matrix stage of a 256-point core
            res = {"functor": fam_name}
            best = 0.0
            for shape, fb in shapes.items():
                batch = fb(n)
                # both workgroup shapes of the example harness: its default
                # (up to 256 threads) and one-wave workgroups
                for wg, key in ((0, shape), (1, shape + "/one_wave")):
                    us = onchip_pass_us(run, tw, stream, xs[:batch], fam, reps, wg)
                    rate = batch / (us * 1e-6) if us > 0 else float("nan")
                    res[key] = {"batch": batch, "us_per_pass": round(us, 4), "ffts_per_s": rate,
                                "GFLOPps": rate * flops / 1e9}
                    best = max(best, rate)
            res["ffts_per_s"] = best
            res["GFLOPps"] = best * flops / 1e9
            res["vs_reference"] = round(best / ref_rate, 3)
            entry[f"family{fam}"] = res
            print(f"[bench] onchip n={n} {fam_name}: {best:.4g} FFT/s ({best / ref_rate:.2f}x reference)",
                  file=sys.stderr, flush=True)
        out["rows"][str(n)] = entry
    return out


def run_precision(fo, steps: int, warmup: int):
    """C4: FP64 vs FP32 vs 3-term split-BF16 MFMA stage GEMMs on the
    dynamic-range stress inputs (seed 11), N = 64 x 65536 and 4096 x 16384.
    Errors here are against the FP64 GPU result of the same signals (itself
    within 1e-13 of the oracle); the CPU leg adds errors against the oracle."""
    out, samples = {}, {}
    for n, batch in ((64, 65536), (4096, 16384)):
        x = gen_signals(n, batch, 11, "stress")
        ref = None
        for prec, flag in (("fp64", 0), ("fp32", fo.PREC_FP32), ("bf16x3", fo.PREC_BF16X3)):
            wl = Workload(f"c4_n{n}_{prec}", n, batch, 11, "stress", "tensor", flag)
            r, _, sample = run_gpu(fo, wl, steps, warmup, host_x=x)
            if ref is None:
                ref = sample
            d = np.sqrt(np.sum(np.abs(sample - ref) ** 2, axis=-1) / np.sum(np.abs(ref) ** 2, axis=-1))
            entry = {k: (round(v, 6) if isinstance(v, float) else v) for k, v in r.items()}
            entry.update({"rel_l2_vs_fp64_max": float(d.max()), "rel_l2_vs_fp64_median": float(np.median(d)),
                          "max_abs_vs_fp64": float(np.max(np.abs(sample - ref))),
                          "checked_signals": int(sample.shape[0])})
            out[wl.name] = entry
            samples[wl.name] = (x[: sample.shape[0]], sample)
            print(f"[bench] {wl.name}: {r['ffts_per_s']:.4g} FFT/s, {r['GBps']:.0f} GB/s, "
                  f"rel L2 vs fp64 {d.max():.2e}", file=sys.stderr, flush=True)
    return out, samples


def precision_vs_oracle(samples, count: int = 64):
    """Per-result error of each C4 run against the long-double oracle (the
    CPU leg, where the oracle is the checker)."""
    import oracle
    res = {}
    for name, (x, got) in samples.items():
        ref = oracle.reference(x[:count])
        e = oracle.rel_l2(got[:count], ref, axis=-1)
        res[name] = {"rel_l2_max": float(e.max()), "rel_l2_median": float(np.median(e)),
                     "max_abs": float(oracle.max_abs(got[:count], ref)), "signals": count}
    return res


def c5_requests(count: int, seed: int, max_batch: int = 4096):
    """Request mix of the served configuration: N in 8..4096 with P ~ 1/log2 N,
    batch uniform 1..max_batch, FP64 or FP32 (complex64) data 50/50."""
    rng = np.random.default_rng(seed)
    sizes = np.array([1 << k for k in range(3, 13)])
    w = 1.0 / np.log2(sizes)
    ns = rng.choice(sizes, size=count, p=w / w.sum())
    batches = rng.integers(1, max_batch + 1, size=count)
    c64 = rng.integers(0, 2, size=count).astype(bool)
    return rng, ns, batches, c64


def run_batched(fo, steps: int, warmup: int, njobs: int = 64, max_batch: int = 4096):
    """Mixed-size plan (C5 job mix, batch 1..max_batch) as ONE persistent launch
    from a hipGraph, against the same jobs issued as one fftoc_fft call each
    (from this Python process, so each call also carries its host overhead)."""
    from fft_onchip_amd import hip
    rng, ns, batches, c64 = c5_requests(njobs, seed=13, max_batch=max_batch)
    bufs, jobs, nsig, nbytes = [], [], 0, 0
    for n, b, c in zip(ns, batches, c64):
        es = 8 if c else 16
        x = rng.uniform(-1, 1, 2 * int(n) * int(b)).astype(np.float32 if c else np.float64)
        src, dst = hip.DeviceBuffer.from_array(x), hip.DeviceBuffer(x.nbytes)
        bufs += [src, dst]
        jobs.append((src, dst, int(b), int(n), fo.DATA_C64 if c else 0))
        nsig += int(b)
        nbytes += 2 * int(n) * int(b) * es
    plan = fo.Plan(jobs)
    stream = hip.Stream()
    plan.autotune(reps=2, stream=stream)

    def timed(run):
        for _ in range(warmup):
            run()
        stream.synchronize()
        a, b = hip.Event(), hip.Event()
        a.record(stream)
        for _ in range(steps):
            run()
        b.record(stream)
        stream.synchronize()
        return a.elapsed_ms(b) / steps

    ms_graph = timed(lambda: plan.execute(stream=stream, graph=True))
    ms_plain = timed(lambda: [fo.fft(j[0], j[2], j[3], out=j[1], flags=j[4], stream=stream) for j in jobs])
    plan.close()
    for b in bufs:
        b.free()
    return {"jobs": njobs, "max_batch": max_batch, "signals": nsig, "bytes": nbytes,
            "graph_ms": round(ms_graph, 4), "graph_ffts_per_s": nsig / (ms_graph * 1e-3),
            "graph_GBps": round(nbytes / (ms_graph * 1e-3) / 1e9, 1),
            "per_call_ms": round(ms_plain, 4), "per_call_ffts_per_s": nsig / (ms_plain * 1e-3),
            "per_call_GBps": round(nbytes / (ms_plain * 1e-3) / 1e9, 1)}


def run_plan_single(fo, steps: int, warmup: int, n: int = 4096, batch: int = 16384, c64: bool = False):
    """One 1 GiB job (N = 4096 x 16384, complex128) as a plan -- autotuned
    (fftoc_plan_autotune picks the routine: a plan whose jobs share one
    routine runs the kernel specialized to it), one launch from a hipGraph --
    against the same transform as one fftoc_fft call, same buffers."""
    from fft_onchip_amd import hip
    es = 8 if c64 else 16
    flags = fo.DATA_C64 if c64 else 0
    x = gen_signals(n, batch, 7, "normal")
    if c64:
        x = x.astype(np.complex64)
    src = hip.DeviceBuffer.from_array(x)
    dst = hip.DeviceBuffer(x.nbytes)
    stream = hip.Stream()
    plan = fo.Plan([(src, dst, batch, n, flags)])
    # the candidates are timed at full clocks: launches for RAMP_MS first
    t0 = time.perf_counter()
    while (time.perf_counter() - t0) * 1e3 < RAMP_MS:
        for _ in range(8):
            plan.execute(stream=stream)
        stream.synchronize()
    plan.autotune(reps=10, stream=stream)

    def timed(run):
        for _ in range(warmup):
            run()
        stream.synchronize()
        a, b = hip.Event(), hip.Event()
        a.record(stream)
        for _ in range(steps):
            run()
        b.record(stream)
        stream.synchronize()
        return a.elapsed_ms(b) / steps

    ms_plan = timed(lambda: plan.execute(stream=stream, graph=True))
    ms_call = timed(lambda: fo.fft(src, batch, n, out=dst, flags=flags, stream=stream))
    routine = plan.routine(n, c64=c64)
    plan.close()
    src.free()
    dst.free()
    nbytes = 2 * n * batch * es
    return {"n": n, "batch": batch, "dtype": "c64" if c64 else "c128", "routine": routine,
            "plan_ms": round(ms_plan, 4), "plan_GBps": round(nbytes / (ms_plan * 1e-3) / 1e9, 1),
            "per_call_ms": round(ms_call, 4),
            "per_call_GBps": round(nbytes / (ms_call * 1e-3) / 1e9, 1),
            "plan_over_per_call": round(ms_call / ms_plan, 3)}


def run_served(fo, seconds: float = 1.0, rate: float = 20000.0, saturated: int = 4000):
    """Poisson request stream through the device-resident server (C5)."""
    from fft_onchip_amd import hip
    count = int(seconds * rate)
    rng, ns, batches, c64 = c5_requests(count + saturated, seed=13)
    lake_bytes = 1 << 30
    lake = hip.DeviceBuffer.from_array(rng.uniform(-1, 1, lake_bytes // 8))
    reqs = []
    for n, b, c in zip(ns, batches, c64):
        nb = int(n) * int(b) * (8 if c else 16)
        off = int(rng.integers(0, (lake_bytes - nb) // 256)) * 256
        reqs.append((lake.ptr + off, int(b), int(n), fo.DATA_C64 if c else 0))
    arrivals = np.cumsum(rng.exponential(1.0 / rate, count))
    out = {}
    with fo.Server(8 << 30, 2048) as srv:
        warm = [0.0] * 64
        srv.replay(reqs[:64], warm)
        lat, span = srv.replay(reqs[:count], arrivals.tolist())
        sig = sum(r[1] for r in reqs[:count])
        out["poisson"] = {"requests": count, "rate_per_s": rate, "span_s": round(span, 4),
                          "ffts_per_s": sig / span,
                          "p50_us": round(float(np.percentile(lat, 50)) * 1e6, 1),
                          "p99_us": round(float(np.percentile(lat, 99)) * 1e6, 1),
                          "mean_us": round(float(np.mean(lat)) * 1e6, 1)}
        # the loop's own round trip: 1-signal N=8 requests at 2 k/s, one at a time
        tiny = [(lake.ptr, 1, 8, 0)] * 1000
        lat0, _ = srv.replay(tiny, (np.arange(1000) * 5e-4).tolist())
        out["round_trip_n8_b1"] = {"requests": 1000,
                                   "p50_us": round(float(np.percentile(lat0, 50)) * 1e6, 1),
                                   "p99_us": round(float(np.percentile(lat0, 99)) * 1e6, 1)}
        sat = reqs[count:]
        lat2, span2 = srv.replay(sat, [0.0] * len(sat))
        sig2 = sum(r[1] for r in sat)
        nbytes2 = sum(2 * r[1] * r[2] * (8 if r[3] else 16) for r in sat)
        out["saturated"] = {"requests": len(sat), "span_s": round(span2, 4),
                            "ffts_per_s": sig2 / span2, "GBps": round(nbytes2 / span2 / 1e9, 1),
                            "p50_us": round(float(np.percentile(lat2, 50)) * 1e6, 1),
                            "p99_us": round(float(np.percentile(lat2, 99)) * 1e6, 1)}
    lake.free()
    return out


def run_rocfft(wl: Workload, steps: int, warmup: int, host_x: np.ndarray):
    """The vendor library on the same workload (the role cuFFTDx plays in the
    reference's tables): rocFFT's C API through ctypes, out-of-place double
    complex forward, same buffer rotation, clock ramp and timing as run_gpu.
    Returns (result, sample of the output for the accuracy check)."""
    import ctypes

    from fft_onchip_amd import hip
    try:
        rf = ctypes.CDLL("librocfft.so.0")
    except OSError as e:
        return {"error": f"librocfft not loadable: {e}"}, None
    vp, sz = ctypes.c_void_p, ctypes.c_size_t
    rf.rocfft_setup.argtypes = []
    rf.rocfft_plan_create.argtypes = [ctypes.POINTER(vp), ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                      sz, ctypes.POINTER(sz), sz, vp]
    rf.rocfft_plan_get_work_buffer_size.argtypes = [vp, ctypes.POINTER(sz)]
    rf.rocfft_execution_info_create.argtypes = [ctypes.POINTER(vp)]
    rf.rocfft_execution_info_set_work_buffer.argtypes = [vp, vp, sz]
    rf.rocfft_execution_info_set_stream.argtypes = [vp, vp]
    rf.rocfft_execute.argtypes = [vp, ctypes.POINTER(vp), ctypes.POINTER(vp), vp]
    rf.rocfft_plan_destroy.argtypes = [vp]
    rf.rocfft_execution_info_destroy.argtypes = [vp]
    if rf.rocfft_setup() != 0:
        return {"error": "rocfft_setup failed"}, None
    n, batch = wl.n, wl.batch
    plan, info = vp(), vp()
    lengths = (sz * 1)(n)
    # placement 1 = not in place, transform 0 = complex forward, precision 1 = double
    if rf.rocfft_plan_create(ctypes.byref(plan), 1, 0, 1, 1, lengths, batch, None) != 0:
        return {"error": "rocfft_plan_create failed"}, None
    wbytes = sz(0)
    rf.rocfft_plan_get_work_buffer_size(plan, ctypes.byref(wbytes))
    rf.rocfft_execution_info_create(ctypes.byref(info))
    work = hip.DeviceBuffer(wbytes.value) if wbytes.value else None
    if work is not None:
        rf.rocfft_execution_info_set_work_buffer(info, vp(work.ptr), wbytes)
    stream = hip.Stream()
    rf.rocfft_execution_info_set_stream(info, vp(stream.handle))
    nbytes = n * batch * 16
    pairs = max(1, min(8, math.ceil(4 * MALL_BYTES / (2 * nbytes))))
    ins = [hip.DeviceBuffer.from_array(host_x) for _ in range(pairs)]
    outs = [hip.DeviceBuffer(nbytes) for _ in range(pairs)]
    ptrs = [((vp * 1)(ins[i].ptr), (vp * 1)(outs[i].ptr)) for i in range(pairs)]

    def execute(i):
        ib, ob = ptrs[i % pairs]
        if rf.rocfft_execute(plan, ib, ob, info) != 0:
            raise RuntimeError("rocfft_execute failed")

    t_ramp, i = time.perf_counter(), 0
    while (time.perf_counter() - t_ramp) * 1e3 < RAMP_MS:
        for _ in range(8):
            execute(i)
            i += 1
        stream.synchronize()
    for i in range(warmup):
        execute(i)
    stream.synchronize()
    t0, t1 = hip.Event(), hip.Event()
    t0.record(stream)
    for i in range(steps):
        execute(i)
    t1.record(stream)
    stream.synchronize()
    ms = t0.elapsed_ms(t1) / steps
    # accuracy sample from a fresh copy of the input (rocFFT may overwrite it)
    ins[0].upload(host_x)
    execute(0)
    stream.synchronize()
    sample = outs[0].download((min(batch, 256), n))
    rf.rocfft_execution_info_destroy(info)
    rf.rocfft_plan_destroy(plan)
    for b in ins + outs + ([work] if work is not None else []):
        b.free()
    gbps = 32.0 * n * batch / (ms * 1e-3) / 1e9
    return {"workload": wl.name, "n": n, "batch": batch, "path": "rocfft",
            "ffts_per_s": batch / (ms * 1e-3), "ms_per_step": round(ms, 6), "GBps": round(gbps, 1),
            "hbm_frac": round(gbps / HBM_PEAK_GBPS, 4), "work_buffer_bytes": wbytes.value}, sample


def copy_bandwidth(nbytes: int = 1 << 30, reps: int = 10) -> float:
    """Device-to-device hipMemcpy of `nbytes` (read + write) in GB/s."""
    import ctypes

    from fft_onchip_amd import hip
    src, dst = hip.DeviceBuffer(nbytes), hip.DeviceBuffer(nbytes)
    stream = hip.Stream()
    h = hip.lib()
    for _ in range(2):
        hip.check(h.hipMemcpyAsync(ctypes.c_void_p(dst.ptr), ctypes.c_void_p(src.ptr), nbytes, 3,
                                   ctypes.c_void_p(stream.handle)))
    a, b = hip.Event(), hip.Event()
    a.record(stream)
    for _ in range(reps):
        hip.check(h.hipMemcpyAsync(ctypes.c_void_p(dst.ptr), ctypes.c_void_p(src.ptr), nbytes, 3,
                                   ctypes.c_void_p(stream.handle)))
    b.record(stream)
    stream.synchronize()
    ms = a.elapsed_ms(b) / reps
    src.free()
    dst.free()
    return 2.0 * nbytes / (ms * 1e-3) / 1e9


def copy_ceiling(nbytes: int = 1 << 30, launches: int = 20, ramp: int = 200) -> dict:
    """Plain copies of `nbytes` (read + write) moving 1, 4, 16 and 64 KiB per
    workgroup (fft_onchip_amd/libfftoc_copy.so): the streaming rate a kernel
    of that granularity can reach on this device.  A transform's workgroup
    owns whole signals (64 KiB at N = 4096 complex128), so its rate is
    compared with the copy at the same block size.  Median of 3 timed runs of
    `launches` launches, after `ramp` untimed ones."""
    import ctypes

    from fft_onchip_amd import hip
    lib = ctypes.CDLL(os.path.join(ROOT, "fft_onchip_amd", "libfftoc_copy.so"))
    fn = lib.fftoc_copy_ceiling
    fn.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, ctypes.c_int, ctypes.c_void_p]
    fn.restype = ctypes.c_int
    src, dst = hip.DeviceBuffer(nbytes), hip.DeviceBuffer(nbytes)
    stream = hip.Stream()
    hip.check(hip.lib().hipMemsetAsync(ctypes.c_void_p(src.ptr), 0, nbytes,
                                       ctypes.c_void_p(stream.handle)))
    out = {"method": "one-shot copy kernels, non-temporal 16 B per lane, 1 GiB in + 1 GiB out",
           "GBps": {}, "frac": {}}
    for kib in (1, 4, 16, 64):
        def launch():
            rc = fn(src.ptr, dst.ptr, nbytes, kib, stream.handle)
            if rc != 0:
                raise RuntimeError(f"fftoc_copy_ceiling({kib} KiB) returned {rc}")
        for _ in range(ramp):
            launch()
        times = []
        for _ in range(3):
            a, b = hip.Event(), hip.Event()
            a.record(stream)
            for _ in range(launches):
                launch()
            b.record(stream)
            stream.synchronize()
            times.append(a.elapsed_ms(b) / launches)
        gbs = 2.0 * nbytes / (sorted(times)[1] * 1e-3) / 1e9
        out["GBps"][f"{kib} KiB"] = round(gbs, 1)
        out["frac"][f"{kib} KiB"] = round(gbs / (HBM_PEAK_GBPS), 4)
    src.free()
    dst.free()
    return out


def cpu_baseline(x: np.ndarray, gpu_sample: np.ndarray, threads: int, budget_s: float = 12.0):
    """FP64 FFT of the same signals on the host cores (the oracle's Stockham
    transform, OpenMP over the batch); median of 5 timed runs."""
    import oracle  # the CPU leg: timed baseline and checker of the GPU sample

    batch, n = x.shape
    out = oracle.fft_f64(x, threads=threads)          # warm-up over the whole batch
    t0 = time.perf_counter()
    oracle.fft_f64(x, threads=threads)
    one = max(time.perf_counter() - t0, 1e-6)
    reps = max(1, int(budget_s / 5.0 / one))           # each timed run ~budget/5 s
    times = []
    for _ in range(5):
        t0 = time.perf_counter()
        for _ in range(reps):
            oracle.fft_f64(x, threads=threads)
        times.append(time.perf_counter() - t0)
    med = statistics.median(times)
    m = gpu_sample.shape[0]
    ref = oracle.reference(x[:m])
    err = float(np.max(oracle.rel_l2(gpu_sample, ref, axis=-1)))
    cpu_err = float(np.max(oracle.rel_l2(out[:m], ref, axis=-1)))
    count = batch * reps
    host = _host_info()
    return {
        "value": count / med, "unit": "FFTs/s", "cores": threads, "kind": "port",
        "sample": f"all {batch} N={n} signals of the workload, {reps}x per timed run; "
                  f"median of 5 runs ({sum(times):.1f} s of CPU work)",
        "GFLOPps": 5.0 * n * math.log2(n) * count / med / 1e9,
        "threads": threads, "threads_rule": host_threads()[1],
        "nproc": host["nproc"], "affinity_cpus": host["affinity_cpus"],
        "cpu_model": host["model"], "cpu_id": host["id"],
        "cpu_max_rel_l2_vs_oracle": cpu_err,
        "gpu_max_rel_l2_vs_oracle": err, "gpu_checked_signals": m,
    }


def host_threads():
    """Threads for the CPU baseline: every CPU this process may run on, up to
    the host-thread share allotted to one GPU (OMP_NUM_THREADS, which the GPU
    host sets to its per-GPU share), with the rule that chose the count."""
    affinity = len(os.sched_getaffinity(0))
    share = os.environ.get("OMP_NUM_THREADS", "")
    if share.isdigit() and 0 < int(share) < affinity:
        return int(share), (f"OMP_NUM_THREADS={share}: the host threads allotted to one GPU "
                            f"({affinity} CPUs in this process's affinity mask)")
    return affinity, f"all {affinity} CPUs in this process's affinity mask"


def _host_info() -> dict:
    """Host CPU as /proc/cpuinfo reports it: model string plus vendor, family,
    model and stepping numbers (a virtual machine may report a generic model
    string), and the CPU counts."""
    fields = {}
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if not line.strip():
                    break  # first processor's block only
                k, _, v = line.partition(":")
                fields[k.strip()] = v.strip()
    except OSError:
        pass
    ident = {k: fields.get(k) for k in ("vendor_id", "cpu family", "model", "stepping", "cpu MHz",
                                        "cache size")}
    ident["avx512f"] = "avx512f" in fields.get("flags", "").split()
    return {"model": fields.get("model name", "unknown"), "id": ident,
            "nproc": os.cpu_count(), "affinity_cpus": len(os.sched_getaffinity(0))}


TRAFFIC_JSON = os.path.join(ROOT, "profiles", "traffic.json")
COPY_KERNEL = "__amd_rocclr_copyBuffer"


HEADLINE = "fp64_c2c_fwd_n4096_b16384"


def run_power_evidence(fo, steps: int, warmup: int, x: np.ndarray) -> dict:
    """N = 4096 x 16384 tensor kernels on N(0,1) data and on zeros: the same
    instructions and HBM traffic, different switching activity.  A kernel that
    runs faster on zeros is bound by the power its matrix products draw (the
    clock drops), not by its memory system or issue."""
    zeros = np.zeros((16384, 4096), np.complex128)
    rows = {}
    for key, variant in (("1 matrix stage (default)", "tensor:16x16x16/VALU stages A and C, 4 waves"),
                         ("2 matrix stages", "tensor:16x16x16/VALU stage A, 4 waves"),
                         ("2 matrix stages, 2-product", "tensor:16x16x16/VALU stage A, 4 waves/2-product"),
                         ("3 matrix stages", "tensor:16x16x16/8 waves")):
        row = {"variant": variant}
        for data, hx in (("normal", x), ("zeros", zeros)):
            wl = Workload(f"power_{data}", 4096, 16384, 7, "normal", "variant:" + variant)
            r, _, _ = run_gpu(fo, wl, steps, warmup, host_x=hx)
            row[f"{data}_GBps"] = round(r["GBps"], 1)
        row["zeros_over_normal"] = round(row["zeros_GBps"] / row["normal_GBps"], 3)
        rows[key] = row
    return rows


def headline_variant(fo, wl) -> str:
    """Name of the kernel variant fftoc_fft runs for the workload."""
    lib = fo.lib()
    return lib.fftoc_variant_name(wl.n, 0, lib.fftoc_get_variant(wl.n, 0, wl.batch)).decode()


def load_traffic(workload: str, variant: str):
    """HBM bytes per launch of the headline kernel, from the committed
    rocprofv3 --pmc summary (profiles/traffic.json, written by --traffic from
    counter runs of `bench.py --quick`, which launches only that kernel and
    the calibration copy).  None unless the summary was taken of this
    workload run by this kernel variant (a stale summary is not reported)."""
    try:
        with open(TRAFFIC_JSON) as f:
            t = json.load(f)
        if t.get("workload") != workload or t.get("variant") != variant or len(t["kernels"]) != 1:
            return None
        (name, v), = t["kernels"].items()
        return dict(v, kernel=name, fetch_scale=t["fetch_scale"], write_scale=t["write_scale"])
    except (OSError, ValueError, KeyError):
        return None


def traffic_from_pmc(fetch_csv: str, write_csv: str, out_path: str = "",
                     workload: str = "", variant: str = "") -> dict:
    """Scales FETCH_SIZE / WRITE_SIZE (KiB per dispatch) by what the same
    counters read for the 1 GiB device-to-device copy this script runs
    (known: 1 GiB read, 1 GiB written) and writes the result as JSON
    (default profiles/traffic.json, the headline workload's)."""
    import csv
    import collections

    def mean_by_kernel(path, counter):
        acc = collections.defaultdict(list)
        for r in csv.DictReader(open(path)):
            if r["Counter_Name"] == counter:
                acc[r["Kernel_Name"]].append(float(r["Counter_Value"]))
        return {k: statistics.fmean(v) for k, v in acc.items()}

    fetch = mean_by_kernel(fetch_csv, "FETCH_SIZE")
    write = mean_by_kernel(write_csv, "WRITE_SIZE")
    copy_f = [v for k, v in fetch.items() if COPY_KERNEL in k]
    copy_w = [v for k, v in write.items() if COPY_KERNEL in k]
    if not copy_f or not copy_w:
        raise SystemExit("no copy kernel in the counter files")
    kib = (1 << 30) / 1024.0
    f_scale, w_scale = kib / max(copy_f), kib / max(copy_w)
    out = {"fetch_scale": f_scale, "write_scale": w_scale,
           "calibration": "1 GiB hipMemcpy device-to-device (read 1 GiB + write 1 GiB)",
           "workload": workload or HEADLINE, "variant": variant, "kernels": {}}
    for k in fetch:
        if COPY_KERNEL in k or k not in write:
            continue
        key = k.split("(HIP_vector_type")[0]
        out["kernels"][key] = {"fetch_bytes": fetch[k] * 1024 * f_scale,
                               "write_bytes": write[k] * 1024 * w_scale,
                               "bytes": (fetch[k] * f_scale + write[k] * w_scale) * 1024}
    with open(out_path or TRAFFIC_JSON, "w") as f:
        json.dump(out, f, indent=1)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--quick", action="store_true", help="headline workload only")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--onchip", action="store_true", help="with --quick: the on-chip leg too")
    ap.add_argument("--only", metavar="PATH:N[:PREC]",
                    help="time one workload (path best|tensor|legacy8|legacy16, length N at its "
                         "benchmark batch, precision fp64|fp32|bf16x3) and print its result; "
                         "for counter runs")
    ap.add_argument("--variant", metavar="NAME",
                    help="with --only: run this named kernel variant of the length")
    ap.add_argument("--traffic", nargs=2, metavar=("FETCH_CSV", "WRITE_CSV"),
                    help="write profiles/traffic.json from two rocprofv3 --pmc runs and exit")
    ap.add_argument("--traffic-out", default="", metavar="JSON",
                    help="with --traffic: write here instead (counter runs of other workloads)")
    ap.add_argument("--traffic-variant", default="", metavar="NAME",
                    help="with --traffic: the kernel variant the counter runs measured "
                         "(default: the headline workload's)")
    ap.add_argument("--traffic-workload", default="", metavar="NAME",
                    help="with --traffic: the workload the counter runs measured")
    args = ap.parse_args()
    if args.traffic:
        variant = args.traffic_variant
        if not variant and not args.traffic_workload:
            # the headline workload: the variant fftoc_fft picks for it (a
            # table lookup, no device needed)
            import fft_onchip_amd as fo
            variant = headline_variant(fo, Workload(HEADLINE, 4096, 16384, 7, "normal", "best"))
        print(json.dumps(traffic_from_pmc(*args.traffic, out_path=args.traffic_out,
                                          workload=args.traffic_workload, variant=variant),
                         indent=1))
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build measures one GPU; the batched FFT shards by contiguous "
              "signal ranges with no exchange, so N>1 is not run here.", file=sys.stderr)
        sys.exit(2)

    import fft_onchip_amd as fo
    from fft_onchip_amd import hip
    if hip.device_count() < 1:
        print("bench.py: no HIP device visible", file=sys.stderr)
        sys.exit(1)
    fo.init()

    if args.only:
        parts = args.only.split(":")
        path, n = parts[0], int(parts[1])
        prec = parts[2] if len(parts) > 2 else "fp64"
        flag = {"fp64": 0, "fp32": fo.PREC_FP32, "bf16x3": fo.PREC_BF16X3}[prec]
        batch = {4096: 16384, 64: 65536}.get(n, (1 << 24) // n)
        if args.variant:
            path = "variant:" + args.variant
        wl = Workload(f"only_{path}_n{n}_{prec}", n, batch, 7, "normal", path, flag)
        r, _, _ = run_gpu(fo, wl, args.steps, args.warmup)
        # the 1 GiB copies that calibrate FETCH_SIZE / WRITE_SIZE in counter runs
        r["hip_d2d_copy_GBps"] = copy_bandwidth()
        print(json.dumps({k: (round(v, 6) if isinstance(v, float) else v) for k, v in r.items()}))
        return

    head = Workload(HEADLINE, 4096, 16384, 7, "normal", "best")
    res, x, sample = run_gpu(fo, head, args.steps, args.warmup)
    print(f"[bench] {head.name}: {res['ffts_per_s']:.4g} FFT/s, {res['GBps']:.0f} GB/s", file=sys.stderr, flush=True)

    extra = {}
    if not args.quick:
        others = [Workload("fp64_n4096_b16384_tensor", 4096, 16384, 7, "normal", "tensor"),
                  # Karatsuba with two of the three DFT16 stages on the matrix cores
                  Workload("fp64_n4096_b16384_tensor_2mfma", 4096, 16384, 7, "normal",
                           "variant:tensor:16x16x16/VALU stage A, 4 waves"),
                  # the tensor path with all three DFT16 stages on the matrix cores
                  Workload("fp64_n4096_b16384_tensor_3mfma", 4096, 16384, 7, "normal",
                           "variant:tensor:16x16x16/8 waves"),
                  # the two-real-product DFT16 contraction (packed cosine/sine columns)
                  Workload("fp64_n4096_b16384_tensor_2product", 4096, 16384, 7, "normal",
                           "variant:tensor:16x16x16/VALU stage A, 4 waves/2-product"),
                  Workload("fp64_n4096_b16384_legacy16", 4096, 16384, 7, "normal", "legacy16"),
                  Workload("fp64_n4096_b16384_legacy8", 4096, 16384, 7, "normal", "legacy8")]
        for path in ("best", "tensor", "legacy8"):
            others.append(Workload(f"fp64_n64_b65536_{path}", 64, 65536, 0, "uniform", path))
        others.append(Workload("fp64_n64_b65536_tensor_2product", 64, 65536, 0, "uniform",
                               "variant:tensor:16x4/8 per wave/persistent/2-product"))
        for i, n in enumerate((8, 16, 128, 256, 512)):
            for path in ("best", "tensor"):
                others.append(Workload(f"fp64_n{n}_b{(1 << 24) // n}_{path}", n, (1 << 24) // n,
                                       1 + i, "uniform", path))
            if n in (8, 512):
                others.append(Workload(f"fp64_n{n}_b{(1 << 24) // n}_legacy8", n, (1 << 24) // n,
                                       1 + i, "uniform", "legacy8"))
            if n in (16, 256):
                others.append(Workload(f"fp64_n{n}_b{(1 << 24) // n}_legacy16", n, (1 << 24) // n,
                                       1 + i, "uniform", "legacy16"))
        # complex64 data (FFTOC_DATA_C64, FP32 arithmetic), the input bytes of
        # C3 and C1: 8 B per point in and out
        for n, b, seed in ((4096, 32768, 7), (64, 131072, 0)):
            others.append(Workload(f"c64_n{n}_b{b}_best", n, b, seed,
                                   "normal" if n == 4096 else "uniform", "best", fo.DATA_C64))
        # the lengths outside C2, at the same 2^24 points
        for n, seed in ((32, 21), (1024, 22), (2048, 23)):
            for path in ("best", "tensor") if n == 32 else ("best",):
                others.append(Workload(f"fp64_n{n}_b{(1 << 24) // n}_{path}", n, (1 << 24) // n,
                                       seed, "uniform", path))
        cache = {}
        for wl in others:
            key = (wl.n, wl.batch, wl.seed, wl.kind)
            if key not in cache:
                cache.clear()
                cache[key] = x if key == (head.n, head.batch, head.seed, head.kind) else \
                    gen_signals(wl.n, wl.batch, wl.seed, wl.kind)
            r, _, _ = run_gpu(fo, wl, args.steps, args.warmup, host_x=cache[key])
            extra[wl.name] = {k: (round(v, 6) if isinstance(v, float) else v) for k, v in r.items()}
            print(f"[bench] {wl.name}: {r['ffts_per_s']:.4g} FFT/s, {r['GBps']:.0f} GB/s", file=sys.stderr, flush=True)
    if not args.quick:
        extra["tensor_power_n4096"] = run_power_evidence(fo, args.steps, args.warmup, x)
        print(f"[bench] tensor power: {extra['tensor_power_n4096']}", file=sys.stderr, flush=True)
    extra["hip_d2d_copy_GBps"] = round(copy_bandwidth(), 1)
    extra["copy_ceiling"] = copy_ceiling()
    extra["copy_ceiling"]["headline_over_64KiB_copy"] = round(res["GBps"] / extra["copy_ceiling"]["GBps"]["64 KiB"], 3)
    print(f"[bench] copy ceiling: {extra['copy_ceiling']}", file=sys.stderr, flush=True)
    rocfft_samples = {}
    if not args.quick:
        extra["rocfft_x4"] = {}
        for wl in (Workload("fp64_n4096_b16384", 4096, 16384, 7, "normal", "rocfft"),
                   Workload("fp64_n64_b65536", 64, 65536, 0, "uniform", "rocfft")):
            hx = x if wl.n == head.n else gen_signals(wl.n, wl.batch, wl.seed, wl.kind)
            r, smp = run_rocfft(wl, args.steps, args.warmup, hx)
            ours = extra.get(f"fp64_n{wl.n}_b{wl.batch}_best") if wl.n != head.n else res
            if ours and "ffts_per_s" in r:
                r["fftoc_best_over_rocfft"] = round(ours["ffts_per_s"] / r["ffts_per_s"], 3)
            extra["rocfft_x4"][wl.name] = r
            if smp is not None:
                rocfft_samples[wl.name] = (hx[:smp.shape[0]], smp)
            print(f"[bench] rocfft {wl.name}: {r}", file=sys.stderr, flush=True)
    prec_samples = {}
    if not args.quick:
        extra["precision_c4"], prec_samples = run_precision(fo, args.steps, args.warmup)
        extra["batched_mixed_c5"] = run_batched(fo, args.steps, args.warmup)
        # many small jobs: where one launch saves the per-call overheads
        extra["batched_small_jobs"] = run_batched(fo, args.steps, args.warmup, njobs=1024,
                                                  max_batch=16)
        print(f"[bench] batched plan: {extra['batched_mixed_c5']}", file=sys.stderr, flush=True)
        extra["plan_single_n4096"] = run_plan_single(fo, args.steps, args.warmup)
        print(f"[bench] single-job plan: {extra['plan_single_n4096']}", file=sys.stderr, flush=True)
        extra["served_c5"] = run_served(fo)
        print(f"[bench] served: {extra['served_c5']}", file=sys.stderr, flush=True)
    if not args.quick or args.onchip:
        extra["onchip"] = run_onchip(fo)
        extra["saturation"] = run_saturation(fo)
    if not args.quick:
        extra["workgroups_per_cu"] = run_per_cu(fo)

    cpu = None
    if not args.no_cpu:
        threads, _ = host_threads()
        cpu = cpu_baseline(x, sample, threads)
        if prec_samples:
            extra["precision_c4_vs_oracle"] = precision_vs_oracle(prec_samples)
        for name, (xin, got) in rocfft_samples.items():
            import oracle  # the CPU leg checks the vendor library's output too
            ref = oracle.reference(xin)
            extra["rocfft_x4"][name]["max_rel_l2_vs_oracle"] = float(
                np.max(oracle.rel_l2(got, ref, axis=-1)))
        print(f"[bench] cpu baseline: {cpu['value']:.4g} FFT/s on {threads} threads", file=sys.stderr, flush=True)

    traffic = load_traffic(head.name, headline_variant(fo, head))
    roof = {"bound": "hbm", "achieved": round(res["GBps"], 2), "peak": HBM_PEAK_GBPS, "unit": "GB/s",
            "frac": round(res["hbm_frac"], 4),
            "traffic": round(traffic["bytes"]) if traffic else None,
            "traffic_detail": traffic,
            "kernel": "fftoc_fft N=4096 = " + headline_variant(fo, head),
            "algorithmic_bytes_per_launch": 32 * head.n * head.batch,
            "launch_ms_mean": round(res["launch_ms_mean"], 5)}
    line = {
        "metric": "FFTs/s (FP64 batched c2c forward FFT, N=4096)",
        "value": round(res["ffts_per_s"], 1),
        "unit": "FFTs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(res["ms_per_step"], 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (Philox seed 7, N(0,1) re/im), resident in HBM",
        "config": {"workload": head.name, "n": head.n, "batch": head.batch, "path": "fftoc_fft",
                   "layout": "interleaved complex128, signal-contiguous", "direction": "forward",
                   "GFLOPps": round(res["GFLOPps"], 1),
                   "clock_ramp": f"{res['clock_ramp_launches']} untimed launches ({res['clock_ramp_ms']:.0f} ms) "
                                 "before the warm-up steps"},
        "roofline": roof,
        "cpu_baseline": cpu,
        "extra": extra,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
