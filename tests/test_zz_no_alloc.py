"""GPU test of the allocation-free batch entry point (fftoc_batch_fft_noalloc)
against the plan entry points it is a variant of (fftoc_plan_create +
fftoc_plan_execute of the same jobs): bit for bit, and within the path's
tolerance of the oracle (FP64 data <= 1e-13 relative L2 per signal, complex64
<= 2e-6), on two calls in a row, the second with fewer and smaller jobs, so
the second runs on the scratch array the first left behind.  The entry point
is the first library call of the test (the device set-up goes through it)."""
import numpy as np
import pytest

import fft_onchip_amd as fo
import oracle

pytestmark = pytest.mark.gpu
TOL64, TOL32 = 1e-13, 2e-6

# (n, batch, complex64, inverse, normalise)
FIRST = [(4096, 700, False, False, False), (64, 20001, True, True, False),
         (256, 3001, False, True, True), (8, 50000, False, False, False),
         (1024, 900, True, False, False), (32, 9999, False, True, False),
         (128, 4097, True, False, True), (512, 2500, False, False, False),
         (16, 777, True, True, False), (2048, 301, False, False, False)]
SECOND = [(512, 33, False, True, False), (8, 1000, True, False, False),
          (4096, 5, False, False, True)]


def make_jobs(specs, seed):
    rng = np.random.default_rng(seed)
    jobs = []
    for n, batch, c64, inv, norm in specs:
        ft = np.float32 if c64 else np.float64
        x = rng.uniform(-1, 1, 2 * n * batch).astype(ft).view(np.complex64 if c64 else np.complex128)
        flags = (fo.INVERSE if inv else 0) | (fo.NORMALIZE if norm else 0) | (fo.DATA_C64 if c64 else 0)
        jobs.append(dict(n=n, x=x.reshape(batch, n), flags=flags, c64=c64, inv=inv, norm=norm,
                         src=fo.DeviceBuffer.from_array(x), a=fo.DeviceBuffer(x.nbytes),
                         b=fo.DeviceBuffer(x.nbytes)))
    return jobs


def run_both(jobs):
    fo.batch_fft_noalloc([(j["src"], j["a"], j["x"].shape[0], j["n"], j["flags"]) for j in jobs])
    fo.synchronize()
    got = [j["a"].download(j["x"].shape, j["x"].dtype) for j in jobs]
    plan = fo.Plan([(j["src"], j["b"], j["x"].shape[0], j["n"], j["flags"]) for j in jobs])
    plan.execute()
    fo.synchronize()
    ref = [j["b"].download(j["x"].shape, j["x"].dtype) for j in jobs]
    plan.close()
    return got, ref


def check_oracle(j, got, edge=32):
    b = j["x"].shape[0]
    for sl in (slice(0, min(edge, b)), slice(max(0, b - edge), b)):
        ref = oracle.reference(j["x"][sl].astype(np.complex128), inverse=j["inv"])
        if j["norm"]:
            ref = oracle.DD(ref.hi / j["n"], ref.lo / j["n"])
        err = np.max(oracle.rel_l2(got[sl], ref, axis=-1))
        assert err <= (TOL32 if j["c64"] else TOL64), (j["n"], b, j["c64"], err)


def test_noalloc_batch_matches_plan_two_calls_in_a_row():
    assert fo.device_count() >= 1, "no HIP device visible"
    for call, (specs, seed) in enumerate(((FIRST, 71), (SECOND, 72))):
        jobs = make_jobs(specs, seed)
        print(f"[noalloc] call {call}: {len(jobs)} jobs", flush=True)
        got, ref = run_both(jobs)
        for j, g, r in zip(jobs, got, ref):
            assert np.array_equal(g, r), (call, j["n"], j["x"].shape, j["c64"])
            check_oracle(j, g)


def test_noalloc_batch_limits():
    x = np.zeros((4, 64), np.complex128)
    buf = fo.DeviceBuffer.from_array(x)
    outs = [fo.DeviceBuffer(x.nbytes) for _ in range(fo.NOALLOC_MAX_JOBS + 1)]
    fo.batch_fft_noalloc([])                                    # nothing to do
    fo.batch_fft_noalloc([(buf, outs[0], 0, 64, 0)])             # empty job
    with pytest.raises(fo.FftocError) as e:                      # beyond the scratch array
        fo.batch_fft_noalloc([(buf, o, 4, 64, 0) for o in outs])
    assert e.value.name == "BUDGET"
    with pytest.raises(fo.FftocError) as e:
        fo.batch_fft_noalloc([(buf, outs[0], 4, 64, 0), (buf, outs[0], 4, 64, 0)])
    assert e.value.name == "OVERLAP"
    with pytest.raises(fo.FftocError) as e:
        fo.batch_fft_noalloc([(buf, outs[0], 4, 100, 0)])
    assert e.value.name == "INVALID_SIZE"
    fo.batch_fft_noalloc([(buf, o, 4, 64, 0) for o in outs[:fo.NOALLOC_MAX_JOBS]])
    fo.synchronize()
    for o in outs[:fo.NOALLOC_MAX_JOBS]:
        assert not np.any(o.download(x.shape))                  # FFT of zeros
