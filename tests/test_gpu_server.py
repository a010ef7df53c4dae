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


def test_server_slow_stream_every_request_completes():
    """About 3 s of 1-signal requests at ~100 per second with the default idle
    timeout.  Each CU's primary worker holds a ticket far ahead of the
    published items (request k's item is ticket k), so holders wait seconds
    for their item while the dispatcher, reset by every arrival, never times
    out.  Every request must complete within a host-side deadline: a worker
    that gave its ticket up on its own timer would leave that request's item
    unrun for good."""
    x = signals(8, 1, seed=31)
    buf = fo.DeviceBuffer.from_array(x)
    ref = oracle.reference(x)
    with fo.Server(1 << 20, 8) as srv:
        t0 = time.perf_counter()
        # no HIP call inside the stream (a synchronous copy can share the
        # loop's hardware queue and wait for the loop to time out, which would
        # restart the loop and its tickets every request): polls read host
        # memory only
        for k in range(300):
            rid = submit_retry(srv, buf, 1, 8, 0, timeout=5.0)
            srv.wait(rid, timeout_s=2.0)
            srv.release(rid)
            if k % 100 == 0:
                print(f"[server] slow stream: request {k} at {time.perf_counter() - t0:.2f} s",
                      flush=True)
            # pace at ~100 requests per second
            delay = t0 + 0.01 * (k + 1) - time.perf_counter()
            if delay > 0:
                time.sleep(delay)
        assert time.perf_counter() - t0 > 2.5
        rid = submit_retry(srv, buf, 1, 8, 0)
        got = hip.download_ptr(srv.wait(rid), x.shape)
        srv.release(rid)
    assert np.max(oracle.rel_l2(got, ref, axis=-1)) <= TOL64


# Requests of the load test, in submission order: (n, batch, complex64,
# inverse, normalise).  With nothing released, the output pool fills as they
# are submitted: the first two are sized with the light-load rule, every later
# one finds more than 256 MiB of outputs in flight and gets the deep-backlog
# item size (256 KiB - 4 MiB, growing with the backlog).  A, C and I have
# more work items (1024, 1024, ~560) than the loop has workers (511), so
# their items are counted in pairs.
DEEP_REQUESTS = [
    (4096, 4096, False, False, False),   # A 256 MiB
    (512, 32768, True, False, False),    # B 128 MiB
    (2048, 8192, False, True, False),    # C 256 MiB
    (8, 1 << 20, False, False, False),   # D 128 MiB
    (64, 200003, True, True, True),      # E
    (256, 50001, False, False, True),    # F
    (4096, 2001, True, False, False),    # G
    (16, 12345, False, True, False),     # H
    (1024, 20000, False, False, False),  # I 312 MiB
    (128, 70001, True, False, False),    # J
    (32, 99999, False, False, False),    # K
]


def _load_inputs(seed):
    rng = np.random.default_rng(seed)
    reqs = []
    for n, batch, c64, inv, norm in DEEP_REQUESTS:
        ft = np.float32 if c64 else np.float64
        x = rng.uniform(-1, 1, 2 * n * batch).astype(ft).view(np.complex64 if c64 else np.complex128)
        x = x.reshape(batch, n)
        flags = (fo.INVERSE if inv else 0) | (fo.NORMALIZE if norm else 0) | (fo.DATA_C64 if c64 else 0)
        reqs.append(dict(n=n, x=x, flags=flags, c64=c64, inv=inv, norm=norm,
                         buf=fo.DeviceBuffer.from_array(x)))
    return reqs


def _check_edges(r, got, edge=32):
    """First and last `edge` signals against the oracle."""
    b = r["x"].shape[0]
    for sl in (slice(0, min(edge, b)), slice(max(0, b - edge), b)):
        ref = oracle.reference(r["x"][sl].astype(np.complex128), inverse=r["inv"])
        if r["norm"]:
            ref = oracle.DD(ref.hi / r["n"], ref.lo / r["n"])
        err = np.max(oracle.rel_l2(got[sl], ref, axis=-1))
        assert err <= (TOL32 if r["c64"] else TOL64), (r["n"], r["x"].shape, r["c64"], err)


def test_server_deep_backlog_matches_alone_bit_for_bit():
    """Past the deep-backlog threshold (backlog-sized items) and with requests
    of more items than workers (paired completion counts), every request's
    output equals, bit for bit, the same request served alone on an idle
    server (light-load item sizes), and its first and last 32 signals match
    the oracle."""
    reqs = _load_inputs(41)
    total = sum(r["x"].nbytes for r in reqs)
    with fo.Server(4 << 30, 64) as srv:
        print(f"[server] deep backlog: {len(reqs)} requests, {total / 2**20:.0f} MiB", flush=True)
        ids = [submit_retry(srv, r["buf"], r["x"].shape[0], r["n"], r["flags"]) for r in reqs]
        used, inflight = srv.stats()
        assert used >= total and inflight == len(reqs)
        mixed = []
        for r, rid in zip(reqs, ids):
            mixed.append(hip.download_ptr(srv.wait(rid, timeout_s=30.0), r["x"].shape, r["x"].dtype))
        for rid in ids:
            srv.release(rid)
        assert srv.stats() == (0, 0)
        for k, r in enumerate(reqs):
            print(f"[server] alone: n={r['n']} batch={r['x'].shape[0]} c64={r['c64']}", flush=True)
            rid = submit_retry(srv, r["buf"], r["x"].shape[0], r["n"], r["flags"])
            alone = hip.download_ptr(srv.wait(rid, timeout_s=30.0), r["x"].shape, r["x"].dtype)
            srv.release(rid)
            assert np.array_equal(alone, mixed[k]), (r["n"], r["x"].shape, r["c64"])
    for r, got in zip(reqs, mixed):
        _check_edges(r, got)


def test_server_replay_refuses_a_stall_and_drains():
    """A slot held outside the trace makes admission impossible: the replay
    returns QUEUE_FULL after its stall bound (it used to spin forever), and the
    server still serves afterwards.  A budget that is not a multiple of the
    pool's 256-byte granule is rounded down, so a request in the gap is
    refused as over budget instead of waiting forever."""
    x = signals(64, 4, seed=51)
    buf = fo.DeviceBuffer.from_array(x)
    with fo.Server(1 << 20, 1) as srv:
        held = srv.submit(buf, 4, 64, 0)
        assert held is not None
        srv.wait(held)
        t0 = time.perf_counter()
        with pytest.raises(fo.FftocError) as e:
            srv.replay([(buf, 4, 64, 0)] * 3, [0.0, 0.0, 0.0])
        assert e.value.name == "QUEUE_FULL"
        assert time.perf_counter() - t0 < 10.0
        srv.release(held)
        lat, _ = srv.replay([(buf, 4, 64, 0)] * 3, [0.0, 0.0, 0.0])
        assert np.all(np.isfinite(lat))
        assert srv.stats() == (0, 0)
    with fo.Server(4096 + 100, 4) as srv:
        y = signals(8, 33, seed=52)                  # 4224 bytes: in the 100-byte gap
        ybuf = fo.DeviceBuffer.from_array(y)
        with pytest.raises(fo.FftocError) as e:
            srv.submit(ybuf, 33, 8, 0)
        assert e.value.name == "BUDGET"
        with pytest.raises(fo.FftocError) as e:
            srv.replay([(ybuf, 33, 8, 0)], [0.0])
        assert e.value.name == "BUDGET"
        rid = srv.submit(ybuf, 32, 8, 0)             # 4096 bytes fit
        got = hip.download_ptr(srv.wait(rid), (32, 8))
        srv.release(rid)
        assert np.max(oracle.rel_l2(got, oracle.reference(y[:32]), axis=-1)) <= TOL64
