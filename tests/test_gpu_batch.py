"""GPU tests of complex64 (FP32) transforms and of mixed-size batch plans.

Tolerances: FP64 data <= 1e-13 relative L2 per signal; complex64 data
(FP32 arithmetic, FP32 twiddles) <= 2e-6 relative L2 per signal.
"""
import numpy as np
import pytest

import fft_onchip_amd as fo
import oracle

pytestmark = pytest.mark.gpu
TOL64 = 1e-13
TOL32 = 2e-6


def signals(n, batch, seed, c64=False):
    g = np.random.Generator(np.random.Philox(seed))
    x = g.uniform(-1, 1, (batch, n)) + 1j * g.uniform(-1, 1, (batch, n))
    return x.astype(np.complex64) if c64 else x


@pytest.fixture(scope="module", autouse=True)
def _device():
    assert fo.device_count() >= 1, "no HIP device visible"
    fo.init()


@pytest.mark.parametrize("n", fo.ALL_SIZES)
@pytest.mark.parametrize("path", ["best", "legacy8", "legacy16"])
def test_complex64_matches_oracle(n, path):
    if path == "legacy8" and n not in fo.LEGACY8_SIZES or path == "legacy16" and n not in fo.LEGACY16_SIZES:
        pytest.skip("size not in this family")
    x = signals(n, 257, seed=n, c64=True)
    ref = oracle.reference(x.astype(np.complex128))
    got = fo.fft_host(x, path=path)
    assert got.dtype == np.complex64
    assert np.max(oracle.rel_l2(got, ref, axis=-1)) <= TOL32
    refi = oracle.reference(x.astype(np.complex128), inverse=True)
    goti = fo.fft_host(x, path=path, inverse=True)
    assert np.max(oracle.rel_l2(goti, refi, axis=-1)) <= TOL32


@pytest.mark.parametrize("n", fo.ALL_SIZES)
def test_every_complex64_variant_matches_oracle(n):
    """Each complex64 kernel variant (the complex64 list of the length, as
    fftoc_variant_name enumerates it with FFTOC_DATA_C64), forward and
    inverse, on a batch that leaves a partial last workgroup; one past the
    last index is rejected."""
    lib = fo.lib()
    batch = max(3, min(515, 65536 // n)) | 1
    x = signals(n, batch, seed=700 + n, c64=True)
    refs = {False: oracle.reference(x.astype(np.complex128)),
            True: oracle.reference(x.astype(np.complex128), inverse=True)}
    src = fo.DeviceBuffer.from_array(x)
    dst = fo.DeviceBuffer(x.nbytes)
    names = fo.variants(n, c64=True)
    assert names, n
    for v, name in enumerate(names):
        for inv in (False, True):
            print(f"[c64 variant] n={n} v={v} {name} inverse={inv}", flush=True)
            flags = fo.DATA_C64 | (fo.INVERSE if inv else 0)
            assert lib.fftoc_fft_variant(src.ptr, dst.ptr, batch, n, flags, v, None) == 0
            fo.synchronize()
            got = dst.download(x.shape, np.complex64)
            assert np.max(oracle.rel_l2(got, refs[inv], axis=-1)) <= TOL32, (n, name, inv)
    assert lib.fftoc_fft_variant(src.ptr, dst.ptr, batch, n, fo.DATA_C64, len(names), None) == 13


def make_jobs(seed, c64_mix=True):
    rng = np.random.default_rng(seed)
    jobs = []
    for i, n in enumerate(list(fo.ALL_SIZES) * 2):
        batch = int(rng.integers(1, max(2, 40000 // n)))
        c64 = c64_mix and (i % 2 == 1)
        inv = bool(rng.integers(2))
        norm = inv and bool(rng.integers(2))
        x = signals(n, batch, seed=1000 * seed + i, c64=c64)
        flags = (fo.INVERSE if inv else 0) | (fo.NORMALIZE if norm else 0) | (fo.DATA_C64 if c64 else 0)
        jobs.append(dict(n=n, x=x, flags=flags, c64=c64, inv=inv, norm=norm))
    return jobs


def check_job(j, got):
    ref = oracle.reference(j["x"].astype(np.complex128), inverse=j["inv"])
    if j["norm"]:
        ref = oracle.DD(ref.hi / j["n"], ref.lo / j["n"])
    err = np.max(oracle.rel_l2(got, ref, axis=-1))
    assert err <= (TOL32 if j["c64"] else TOL64), (j["n"], j["c64"], j["inv"], err)


def run_plan(jobs, graph=False, replays=1, autotune=False):
    bufs = [(fo.DeviceBuffer.from_array(j["x"]), fo.DeviceBuffer(j["x"].nbytes)) for j in jobs]
    plan = fo.Plan([(i, o, j["x"].shape[0], j["n"], j["flags"]) for (i, o), j in zip(bufs, jobs)])
    if autotune:
        plan.autotune(reps=2)
    outs = []
    for r in range(replays):
        plan.execute(graph=graph)
        fo.synchronize()
        outs.append([o.download(j["x"].shape, dtype=j["x"].dtype) for (_, o), j in zip(bufs, jobs)])
    plan.close()
    return outs, bufs, plan


def test_plan_mixed_sizes_and_types_match_oracle():
    jobs = make_jobs(1)
    print(f"[plan] {len(jobs)} jobs", flush=True)
    (outs,), _, _ = run_plan(jobs)
    for j, got in zip(jobs, outs):
        check_job(j, got)


def test_plan_graph_replays_match_eager():
    """Replays of the captured graph, with eager work and new input data in
    between, each match the eager launch."""
    jobs = make_jobs(2)
    bufs = [(fo.DeviceBuffer.from_array(j["x"]), fo.DeviceBuffer(j["x"].nbytes)) for j in jobs]
    plan = fo.Plan([(i, o, j["x"].shape[0], j["n"], j["flags"]) for (i, o), j in zip(bufs, jobs)])
    plan.execute()
    fo.synchronize()
    eager = [o.download(j["x"].shape, dtype=j["x"].dtype) for (_, o), j in zip(bufs, jobs)]
    for j, got in zip(jobs, eager):
        check_job(j, got)
    for rep in range(4):
        print(f"[plan] graph replay {rep}", flush=True)
        # eager work on the device between replays
        fo.fft_host(signals(256, 64, seed=rep))
        if rep == 2:  # new contents in the same buffers
            for (i, _), j in zip(bufs, jobs):
                j["x"] = signals(j["n"], j["x"].shape[0], seed=99 + j["n"], c64=j["c64"])
                i.upload(j["x"])
        plan.execute(graph=True)
        fo.synchronize()
        for (_, o), j, e in zip(bufs, jobs, eager):
            got = o.download(j["x"].shape, dtype=j["x"].dtype)
            if rep < 2:
                assert np.array_equal(got, e)
            else:
                check_job(j, got)
    plan.close()


def test_plan_result_same_alone_and_in_mix():
    jobs = make_jobs(3)
    (mixed,), _, _ = run_plan(jobs)
    for k in (0, 5, 12, len(jobs) - 1):
        (alone,), _, _ = run_plan([jobs[k]])
        assert np.array_equal(alone[0], mixed[k]), (jobs[k]["n"], jobs[k]["c64"])


def test_plan_autotune_keeps_results_correct():
    jobs = make_jobs(4, c64_mix=True)
    (outs,), _, _ = run_plan(jobs, autotune=True)
    for j, got in zip(jobs, outs):
        check_job(j, got)


def test_plan_workgroups_per_cu_keeps_results_bit_for_bit():
    """The plan's workgroups per CU (set by hand, or by the autotuner's sweep
    after its routine choice) changes the grid only: mixed and one-routine
    plans give the same bits at every setting."""
    mixed = make_jobs(6)
    one = [dict(n=4096, x=signals(4096, 300, seed=61), flags=0, c64=False, inv=False, norm=False),
           dict(n=4096, x=signals(4096, 200, seed=62), flags=0, c64=False, inv=False, norm=False)]
    for jobs in (mixed, one):
        bufs = [(fo.DeviceBuffer.from_array(j["x"]), fo.DeviceBuffer(j["x"].nbytes)) for j in jobs]
        plan = fo.Plan([(i, o, j["x"].shape[0], j["n"], j["flags"]) for (i, o), j in zip(bufs, jobs)])
        assert plan.workgroups_per_cu == 0
        results = []
        for k in (0, 1, 3, 63):
            print(f"[plan per-cu] jobs={len(jobs)} k={k}", flush=True)
            plan.workgroups_per_cu = k
            assert plan.workgroups_per_cu == k
            for _, o in bufs:
                o.upload(np.zeros(o.nbytes // 8, np.float64))
            plan.execute()
            fo.synchronize()
            results.append([o.download(j["x"].shape, dtype=j["x"].dtype) for (_, o), j in zip(bufs, jobs)])
        for r in results[1:]:
            for a, b in zip(results[0], r):
                assert np.array_equal(a, b)
        with pytest.raises(fo.FftocError):
            plan.workgroups_per_cu = 64
        for j, got in zip(jobs, results[0]):
            check_job(j, got)
        plan.autotune(reps=2)  # may change the routines, hence the bits
        tuned = plan.workgroups_per_cu
        assert 0 <= tuned <= 16
        outs = []
        for k in (tuned, 0):
            plan.workgroups_per_cu = k
            plan.execute()
            fo.synchronize()
            outs.append([o.download(j["x"].shape, dtype=j["x"].dtype) for (_, o), j in zip(bufs, jobs)])
        for j, a, b in zip(jobs, outs[0], outs[1]):
            assert np.array_equal(a, b)
            check_job(j, a)
        plan.close()


@pytest.mark.parametrize("n", fo.ALL_SIZES)
def test_every_plan_routine_matches_oracle(n):
    """Each tile routine of length n, pinned with fftoc_plan_set_routine, for
    both data types: a forward job and a normalised inverse job of several
    work items each, the last tile partial."""
    routines = fo.batch_routines(n)
    assert routines, n
    batch = (1 << 20) // n + 3
    for c64 in (False, True):
        x = signals(n, batch, seed=700 + n, c64=c64)
        ref_f = oracle.reference(x.astype(np.complex128))
        ref_i = oracle.reference(x.astype(np.complex128), inverse=True)
        ref_i = oracle.DD(ref_i.hi / n, ref_i.lo / n)
        src = fo.DeviceBuffer.from_array(x)
        outs = (fo.DeviceBuffer(x.nbytes), fo.DeviceBuffer(x.nbytes))
        dt = fo.DATA_C64 if c64 else 0
        plan = fo.Plan([(src, outs[0], batch, n, dt),
                        (src, outs[1], batch, n, dt | fo.INVERSE | fo.NORMALIZE)])
        tol = TOL32 if c64 else TOL64
        for name in routines:
            print(f"[routine] n={n} c64={c64} {name}", flush=True)
            plan.set_routine(n, name, c64=c64)
            assert plan.routine(n, c64=c64) == name
            plan.execute()
            fo.synchronize()
            for o, ref in zip(outs, (ref_f, ref_i)):
                got = o.download(x.shape, dtype=x.dtype)
                assert np.max(oracle.rel_l2(got, ref, axis=-1)) <= tol, (n, c64, name)
        with pytest.raises(fo.FftocError):
            plan.set_routine(n, "no such routine", c64=c64)
        plan.close()


def test_plan_rejects_bad_jobs():
    a = fo.DeviceBuffer(64 * 16)
    with pytest.raises(fo.FftocError) as e:
        fo.Plan([(a, a, 1, 100, 0)])
    assert e.value.name == "INVALID_SIZE"
    with pytest.raises(fo.FftocError) as e:
        fo.Plan([(a, a, 1, 64, fo.PREC_FP32)])
    assert e.value.name == "UNSUPPORTED"
    empty = fo.Plan([(a, a, 0, 64, 0)])
    empty.execute()
    empty.execute(graph=True)
    fo.synchronize()
    empty.close()


def test_shutdown_waits_for_plans_and_servers():
    """fftoc_shutdown refuses while a plan or a server holds the device tables
    (their kernels read them), succeeds once they are gone, and the library
    re-initialises on the next call."""
    import gc
    gc.collect()
    x = np.zeros((4, 64), dtype=np.complex128)
    buf = fo.DeviceBuffer.from_array(x)
    plan = fo.Plan([(buf, buf, 4, 64, 0)])
    assert fo.lib().fftoc_shutdown() == 6  # FFTOC_ERROR_UNSUPPORTED
    plan.close()
    with fo.Server(1 << 20, 4):
        assert fo.lib().fftoc_shutdown() == 6
    assert fo.lib().fftoc_shutdown() == 0
    y = np.random.default_rng(1).standard_normal((8, 256)) + 0j
    got = fo.fft_host(y)
    assert np.max(oracle.rel_l2(got, oracle.reference(y), axis=-1)) <= 1e-13


def test_plan_jobs_with_more_items_than_the_grid_bit_for_bit():
    """Jobs of far more work items than the persistent grid has workgroups
    (about 64 KiB of input per item; 2048 and 1500 items against ~512
    workgroups) beside small jobs: each equals, bit for bit, the same job in a
    plan of its own, and its first and last 32 signals match the oracle."""
    rng = np.random.default_rng(61)
    specs = [(4096, 2048, False, False), (1024, 6000, True, True), (64, 1000, False, True),
             (8, 777, True, False), (512, 3001, False, False)]
    jobs = []
    for n, batch, c64, inv in specs:
        ft = np.float32 if c64 else np.float64
        x = rng.uniform(-1, 1, 2 * n * batch).astype(ft).view(np.complex64 if c64 else np.complex128)
        flags = (fo.INVERSE if inv else 0) | (fo.DATA_C64 if c64 else 0)
        jobs.append(dict(n=n, x=x.reshape(batch, n), flags=flags, c64=c64, inv=inv, norm=False))
    print(f"[plan] {len(jobs)} jobs, two larger than the grid", flush=True)
    (mixed,), _, _ = run_plan(jobs)
    for k in (0, 1):
        (alone,), _, _ = run_plan([jobs[k]])
        assert np.array_equal(alone[0], mixed[k]), (jobs[k]["n"], jobs[k]["c64"])
    for j, got in zip(jobs, mixed):
        b = j["x"].shape[0]
        for sl in (slice(0, 32), slice(b - 32, b)):
            check_job(dict(j, x=j["x"][sl]), got[sl])
