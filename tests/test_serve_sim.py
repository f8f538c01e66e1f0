"""Served mode's admission, memory and latency bookkeeping on a fake clock
(host only, no GPU): omr_serve_simulate runs the server's own code for these
(omr_admission.cpp: the request arena, in-order admission, the records
fallback, the latency percentiles) against seeded Poisson arrivals and a
device modelled as finishing samples at a fixed rate in admission order.

Checked against what the rules fix: admission in submission order and never
before arrival, the live window, blocks that never overlap and stay inside
the arena (replayed against a granule-bitmap first-fit model), completion
times from the FIFO service recurrence done_i = max(done_{i-1}, admit_i) +
samples_i / rate, latencies done - arrival and their nearest-rank
percentiles, rejection of requests whose image exceeds the arena, the
records fallback and cancellation."""
import numpy as np
import pytest

import omr

MiB = 1 << 20


def image_bytes(w, h):
    a = lambda b: (b + 255) // 256 * 256  # noqa: E731
    return a(w * h * 24) + a(w * h * 12)


def record_bytes(samples):
    return (samples * 24 + 255) // 256 * 256


def poisson_requests(n, rate_per_ms, seed, max_px=1920 * 1080):
    """Seeded arrivals: exponential gaps, pixels log-uniform in 64^2..max_px
    (16:9 where possible), spp in {1, 4, 16, 64, 256}, 1x1 sub-pixels."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate_per_ms, n))
    px = np.exp(rng.uniform(np.log(64 * 64), np.log(max_px), n))
    out = []
    for i in range(n):
        h = max(1, int(np.sqrt(px[i] * 9 / 16)))
        w = max(1, int(px[i] / h))
        out.append(dict(arrival_ms=float(t[i]), w=w, h=h, ss=1, spp=int(rng.choice([1, 4, 16, 64, 256]))))
    return out


def samples(r):
    return r["w"] * r["h"] * r["ss"] * r["ss"] * r["spp"]


def nearest_rank(v, q):
    s = np.sort(v)
    return s[min(len(s) - 1, int(np.floor(q * (len(s) - 1) + 0.5)))]


class BitmapFirstFit:
    """First fit over 256-byte granules by scanning a bitmap for the lowest
    run of free granules that holds the request."""

    def __init__(self, size):
        self.free = np.ones(size // 256, dtype=bool)

    def take(self, nbytes):
        g = (nbytes + 255) // 256
        f = np.concatenate([[False], self.free, [False]]).astype(np.int8)
        starts = np.flatnonzero(np.diff(f) == 1)
        ends = np.flatnonzero(np.diff(f) == -1)
        for s, e in zip(starts, ends):
            if e - s >= g:
                assert self.free[s:s + g].all()
                self.free[s:s + g] = False
                return int(s) * 256
        return None

    def give(self, off, nbytes):
        g = (nbytes + 255) // 256
        assert not self.free[off // 256:off // 256 + g].any()
        self.free[off // 256:off // 256 + g] = True


RATE = 2.0e6  # samples per ms of the modelled device


@pytest.fixture(scope="module")
def stream():
    reqs = poisson_requests(400, rate_per_ms=0.08, seed=11)
    arena = 384 * MiB
    res, sm = omr.serve_simulate(reqs, arena, RATE, record_limit_bytes=96 * MiB, max_live_jobs=24)
    return reqs, arena, res, sm


def test_deterministic(stream):
    reqs, arena, res, sm = stream
    res2, sm2 = omr.serve_simulate(reqs, arena, RATE, record_limit_bytes=96 * MiB, max_live_jobs=24)
    assert res.tobytes() == res2.tobytes() and sm == sm2


def test_stream_exercises_queueing(stream):
    reqs, arena, res, sm = stream
    arrival = np.array([r["arrival_ms"] for r in reqs])
    done = res["state"] == 2
    assert done.all(), np.unique(res["state"])
    waited = res["admit_ms"] > arrival
    assert waited.mean() > 0.2  # the stream is heavy enough that requests queue
    assert (res["records"] == 0).any() and (res["records"] == 1).any()
    assert sm["peak_used_bytes"] > arena // 2


def test_admission_in_order_and_after_arrival(stream):
    reqs, _, res, _ = stream
    arrival = np.array([r["arrival_ms"] for r in reqs])
    assert (res["admit_ms"] >= arrival).all()
    assert (np.diff(res["admit_ms"]) >= 0).all()
    # admission happens only at events: an arrival, or a completion freeing room
    events = np.concatenate([arrival, res["done_ms"]])
    assert all(np.isin(a, events) for a in res["admit_ms"])


def test_fifo_service_recurrence(stream):
    reqs, _, res, _ = stream
    prev = 0.0
    for i, r in enumerate(reqs):
        expect = max(prev, res["admit_ms"][i]) + samples(r) / RATE
        assert res["done_ms"][i] == pytest.approx(expect, rel=1e-9, abs=1e-9)
        prev = res["done_ms"][i]


def test_latency_bookkeeping(stream):
    reqs, _, res, sm = stream
    arrival = np.array([r["arrival_ms"] for r in reqs])
    assert np.allclose(res["latency_ms"], res["done_ms"] - arrival, rtol=0, atol=1e-9)
    lat = res["latency_ms"]
    assert sm["lat_p50_ms"] == nearest_rank(lat, 0.50)
    assert sm["lat_p99_ms"] == nearest_rank(lat, 0.99)
    assert sm["lat_max_ms"] == lat.max()
    assert sm["jobs_done"] == len(reqs) and sm["makespan_ms"] == res["done_ms"].max()


def test_window_and_memory_never_exceeded(stream):
    reqs, arena, res, sm = stream
    ev = sorted([(t, 1, i) for i, t in enumerate(res["admit_ms"])] + [(t, 0, i) for i, t in enumerate(res["done_ms"])])
    live, used, peak_live, peak_used = set(), 0, 0, 0
    for t, kind, i in ev:
        if kind == 0:
            live.discard(i)
            used -= int(res["bytes"][i])
        else:
            live.add(i)
            used += int(res["bytes"][i])
            peak_live, peak_used = max(peak_live, len(live)), max(peak_used, used)
            assert len(live) <= 24 and used <= arena
    assert peak_live == sm["peak_live_jobs"] and peak_used == sm["peak_used_bytes"]
    assert ((res["off"] + res["bytes"]) <= arena).all()


def test_blocks_follow_first_fit(stream):
    """Replays takes (in admission order) and returns (before the takes at the
    same time, as completions precede the admissions they enable) on a
    granule bitmap: every block is where first fit puts it, so live blocks
    never overlap."""
    reqs, arena, res, _ = stream
    ev = sorted([(t, 1, i) for i, t in enumerate(res["admit_ms"])] + [(t, 0, i) for i, t in enumerate(res["done_ms"])])
    model = BitmapFirstFit(arena)
    for t, kind, i in ev:
        if kind == 0:
            model.give(int(res["off"][i]), int(res["bytes"][i]))
        else:
            assert model.take(int(res["bytes"][i])) == int(res["off"][i]), f"request {i}"


def test_records_planned_and_dropped():
    """Records when they fit the per-request limit and, with the image, the
    arena; a request whose records do not fit the free space at admission is
    admitted without them (the second here), while one admitted when a whole
    block is free again has them (the third, in the first one's block); one
    whose records exceed the limit never has them."""
    w, h, spp = 256, 256, 16
    img, rec = image_bytes(w, h), record_bytes(w * h * spp)
    arena = 2 * img + rec
    reqs = [dict(arrival_ms=0.0, w=w, h=h, ss=1, spp=spp), dict(arrival_ms=0.0, w=w, h=h, ss=1, spp=spp),
            dict(arrival_ms=0.0, w=w, h=h, ss=1, spp=spp)]
    res, _ = omr.serve_simulate(reqs, arena, RATE, record_limit_bytes=rec)
    assert list(res["records"]) == [1, 0, 1]
    assert list(res["bytes"]) == [img + rec, img, img + rec]
    assert res["admit_ms"][2] == res["done_ms"][0] and res["off"][2] == 0  # waits for the first block
    res, _ = omr.serve_simulate(reqs[:1], arena, RATE, record_limit_bytes=rec - 256)
    assert res["records"][0] == 0 and res["bytes"][0] == img


def test_waits_for_memory_and_keeps_order():
    """A request that does not fit waits for a block, and the smaller ones
    behind it wait with it (no overtaking)."""
    big = dict(w=1024, h=1024, ss=1, spp=4)
    small = dict(w=64, h=64, ss=1, spp=1)
    arena = image_bytes(1024, 1024) + 2 * image_bytes(64, 64)
    reqs = [dict(arrival_ms=0.0, **big), dict(arrival_ms=0.5, **big), dict(arrival_ms=0.6, **small),
            dict(arrival_ms=0.7, **small)]
    res, sm = omr.serve_simulate(reqs, arena, RATE, record_limit_bytes=0)
    assert (res["state"] == 2).all()
    assert res["admit_ms"][1] == res["done_ms"][0]
    assert res["admit_ms"][2] == res["admit_ms"][1] and res["admit_ms"][3] == res["admit_ms"][1]
    assert res["off"][1] == 0  # the freed block, first fit
    assert sm["peak_live_jobs"] == 3


def test_rejected_and_cancelled():
    """A request whose image exceeds the arena is rejected at submission and
    takes nothing; a cancelled one leaves the queue without a block when its
    turn comes; neither delays the others."""
    arena = 32 * MiB
    reqs = [dict(arrival_ms=0.0, w=512, h=512, ss=1, spp=4), dict(arrival_ms=0.1, w=4096, h=4096, ss=1, spp=1),
            dict(arrival_ms=0.2, w=512, h=512, ss=1, spp=4, cancel=1), dict(arrival_ms=0.3, w=512, h=512, ss=1, spp=4)]
    res, sm = omr.serve_simulate(reqs, arena, RATE, max_live_jobs=1)
    assert list(res["state"]) == [2, 5, 3, 2]
    assert res["admit_ms"][1] == -1 and res["admit_ms"][2] == -1 and res["bytes"][1] == 0
    assert res["admit_ms"][3] == res["done_ms"][0]
    assert sm["jobs_rejected"] == 1 and sm["jobs_cancelled"] == 1 and sm["jobs_done"] == 2


def test_window_of_one_is_sequential():
    reqs = poisson_requests(60, rate_per_ms=0.5, seed=3, max_px=256 * 256)
    res, sm = omr.serve_simulate(reqs, 1 << 34, RATE, max_live_jobs=1)
    arrival = np.array([r["arrival_ms"] for r in reqs])
    assert sm["peak_live_jobs"] == 1
    assert res["admit_ms"][0] == arrival[0]
    assert np.allclose(res["admit_ms"][1:], np.maximum(arrival[1:], res["done_ms"][:-1]), rtol=0, atol=0)


def test_invalid_arguments():
    ok = [dict(arrival_ms=1.0, w=8, h=8, ss=1, spp=1)]
    with pytest.raises(omr.OmrError):
        omr.serve_simulate(ok + [dict(arrival_ms=0.5, w=8, h=8, ss=1, spp=1)], MiB, RATE)  # arrivals out of order
    with pytest.raises(omr.OmrError):
        omr.serve_simulate(ok, MiB, 0.0)
    with pytest.raises(omr.OmrError):
        omr.serve_simulate(ok, MiB, RATE, max_live_jobs=0)
    with pytest.raises(omr.OmrError):
        omr.serve_simulate([dict(arrival_ms=0.0, w=8, h=8, ss=1, spp=0)], MiB, RATE)
    res, sm = omr.serve_simulate([], MiB, RATE)
    assert len(res) == 0 and sm["jobs_done"] == 0
