"""GPU tests of the device-resident serving loop (fftoc_server_*).

Tolerances as for single calls: FP64 data <= 1e-13 relative L2 per signal,
complex64 data <= 2e-6.  Every wait on the device is bounded.
"""
import time

import numpy as np
import pytest

import fft_onchip_amd as fo
from fft_onchip_amd import hip
import oracle

pytestmark = pytest.mark.gpu
TOL64, TOL32 = 1e-13, 2e-6


def signals(n, batch, seed, c64=False):
    g = np.random.Generator(np.random.Philox(seed))
    x = g.uniform(-1, 1, (batch, n)) + 1j * g.uniform(-1, 1, (batch, n))
    return x.astype(np.complex64) if c64 else x


@pytest.fixture(scope="module", autouse=True)
def _device():
    assert fo.device_count() >= 1, "no HIP device visible"
    fo.init()


def submit_retry(srv, buf, batch, n, flags, timeout=20.0):
    t0 = time.perf_counter()
    while True:
        rid = srv.submit(buf, batch, n, flags)
        if rid is not None:
            return rid
        assert time.perf_counter() - t0 < timeout, "admission stalled"


def make_requests(seed, count, max_batch=300):
    rng = np.random.default_rng(seed)
    sizes = list(fo.ALL_SIZES)
    reqs = []
    for i in range(count):
        n = sizes[i % len(sizes)]
        batch = int(rng.integers(1, max_batch + 1))
        c64 = bool(rng.integers(2))
        inv = bool(rng.integers(2))
        x = signals(n, batch, seed=10 * seed + i, c64=c64)
        flags = (fo.INVERSE if inv else 0) | (fo.DATA_C64 if c64 else 0)
        reqs.append(dict(n=n, x=x, flags=flags, c64=c64, inv=inv, buf=fo.DeviceBuffer.from_array(x)))
    return reqs


def run_requests(srv, reqs, all_at_once=True):
    outs = [None] * len(reqs)
    if all_at_once:
        ids = [submit_retry(srv, r["buf"], r["x"].shape[0], r["n"], r["flags"]) for r in reqs]
    else:
        ids = []
    for k, r in enumerate(reqs):
        if not all_at_once:
            ids.append(submit_retry(srv, r["buf"], r["x"].shape[0], r["n"], r["flags"]))
        ptr = srv.wait(ids[k])
        outs[k] = hip.download_ptr(ptr, r["x"].shape, r["x"].dtype)
        srv.release(ids[k])
    return outs


def check(r, got):
    ref = oracle.reference(r["x"].astype(np.complex128), inverse=r["inv"])
    err = np.max(oracle.rel_l2(got, ref, axis=-1))
    assert err <= (TOL32 if r["c64"] else TOL64), (r["n"], r["c64"], err)


def test_server_requests_match_oracle():
    reqs = make_requests(1, 30)
    with fo.Server(256 << 20, 64) as srv:
        print("[server] 30 requests", flush=True)
        outs = run_requests(srv, reqs)
    for r, got in zip(reqs, outs):
        check(r, got)


def test_server_reuses_slots_many_times():
    """Three ring slots for 120 requests, submitted as slots free up: every
    slot (and its completion counter) serves 40 requests in turn."""
    reqs = make_requests(5, 120, max_batch=64)
    with fo.Server(64 << 20, 3) as srv:
        print("[server] 120 requests through 3 slots", flush=True)
        pending, outs = [], [None] * len(reqs)
        for k, r in enumerate(reqs):
            pending.append((k, submit_retry(srv, r["buf"], r["x"].shape[0], r["n"], r["flags"])))
            if len(pending) == 3:
                j, rid = pending.pop(0)
                outs[j] = hip.download_ptr(srv.wait(rid), reqs[j]["x"].shape, reqs[j]["x"].dtype)
                srv.release(rid)
        for j, rid in pending:
            outs[j] = hip.download_ptr(srv.wait(rid), reqs[j]["x"].shape, reqs[j]["x"].dtype)
            srv.release(rid)
    for r, got in zip(reqs, outs):
        check(r, got)


def test_server_result_same_alone_and_in_mix():
    reqs = make_requests(2, 40)
    with fo.Server(512 << 20, 128) as srv:
        mixed = run_requests(srv, reqs, all_at_once=True)
        for k in (0, 3, 17, 39):
            alone = run_requests(srv, [reqs[k]])[0]
            assert np.array_equal(alone, mixed[k]), (reqs[k]["n"], reqs[k]["c64"])
    for r, got in zip(reqs, mixed):
        check(r, got)


def test_server_budget_and_slot_admission():
    x = signals(64, 256, seed=3)                    # 256 KiB of output per request
    buf = fo.DeviceBuffer.from_array(x)
    with fo.Server(1 << 20, 8) as srv:
        with pytest.raises(fo.FftocError) as e:     # can never fit
            srv.submit(buf, 1025, 64, 0)
        assert e.value.name == "BUDGET"
        ids = [srv.submit(buf, 256, 64, 0) for _ in range(4)]
        assert all(i is not None for i in ids)
        assert srv.submit(buf, 256, 64, 0) is None  # budget in use until released
        assert srv.stats() == (1 << 20, 4)
        srv.wait(ids[0])
        srv.release(ids[0])
        late = srv.submit(buf, 256, 64, 0)
        assert late is not None
        for i in ids[1:] + [late]:
            got = hip.download_ptr(srv.wait(i), x.shape)
            assert np.max(oracle.rel_l2(got, oracle.reference(x), axis=-1)) <= TOL64
            srv.release(i)
        small = fo.DeviceBuffer.from_array(signals(8, 1, seed=4))
        ids = [srv.submit(small, 1, 8, 0) for _ in range(8)]
        assert all(i is not None for i in ids)
        assert srv.submit(small, 1, 8, 0) is None   # ring full: slot of request 8 busy
        for i in ids:
            srv.wait(i)
            srv.release(i)
        assert srv.stats() == (0, 0)


def test_server_relaunches_after_idle_exit():
    x = signals(256, 16, seed=5)
    buf = fo.DeviceBuffer.from_array(x)
    with fo.Server(16 << 20, 16) as srv:
        srv.set_idle_timeout(0.05)                   # from the next launch on
        rid = srv.submit(buf, 16, 256, 0)
        srv.wait(rid)
        srv.release(rid)
        print("[server] idle 0.3 s", flush=True)
        time.sleep(0.3)                              # the device loop times out (20 ms)
        rid = srv.submit(buf, 16, 256, 0)
        srv.wait(rid)
        srv.release(rid)
        time.sleep(0.3)                              # and again (50 ms)
        rid = srv.submit(buf, 16, 256, 0)
        got = hip.download_ptr(srv.wait(rid), x.shape)
        srv.release(rid)
    assert np.max(oracle.rel_l2(got, oracle.reference(x), axis=-1)) <= TOL64


def test_server_replay_poisson_trace():
    rng = np.random.default_rng(13)
    lake = fo.DeviceBuffer.from_array(rng.uniform(-1, 1, (1 << 22,)))   # 32 MiB of input data
    sizes = np.array(fo.ALL_SIZES)
    w = 1.0 / np.log2(sizes)
    count = 400
    arrivals = np.cumsum(rng.exponential(1 / 20000.0, count))
    reqs = []
    for i in range(count):
        n = int(rng.choice(sizes, p=w / w.sum()))
        batch = int(rng.integers(1, 257))
        c64 = bool(rng.integers(2))
        es = 8 if c64 else 16
        nbytes = n * batch * es
        off = int(rng.integers(0, (lake.nbytes - nbytes) // 256)) * 256
        reqs.append((lake.ptr + off, batch, n, fo.DATA_C64 if c64 else 0))
    with fo.Server(1 << 30, 256) as srv:
        lat, span = srv.replay(reqs, arrivals.tolist())
    assert np.all(np.isfinite(lat)) and np.all(lat > 0)
    print(f"[server] replay {count} requests in {span:.3f} s: p50 {np.percentile(lat, 50)*1e6:.0f} us, "
          f"p99 {np.percentile(lat, 99)*1e6:.0f} us", flush=True)
