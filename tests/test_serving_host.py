"""Host logic of the serving engine: request stream and slot allocator."""
import numpy as np

from geoflood_amd.serving import SlotAllocator, make_requests


def test_request_stream_statistics():
    reqs = make_requests(4000, rate=8.0, seed=0)
    gaps = np.diff([0.0] + [r.arrival for r in reqs])
    assert abs(gaps.mean() - 1 / 8.0) < 0.01
    sizes = np.array([r.n for r in reqs])
    for n, w in zip((128, 256, 512, 1024), (0.4, 0.3, 0.2, 0.1)):
        assert abs((sizes == n).mean() - w) < 0.03
    man = np.array([r.manning for r in reqs])
    assert man.min() >= 0.02 and man.max() <= 0.06
    peak = np.array([r.peak_q for r in reqs])
    assert peak.min() >= 20 and peak.max() <= 2000
    assert abs(np.median(np.log(peak)) - 0.5 * (np.log(20) + np.log(2000))) < 0.15
    assert [r.seed for r in reqs] == list(range(4000))
    assert make_requests(10, 2.0, seed=3)[5] == make_requests(10, 2.0, seed=3)[5]


def test_slot_allocator_first_fit_and_merge():
    a = SlotAllocator(100)
    s1 = a.alloc(40)
    s2 = a.alloc(40)
    assert (s1, s2) == (0, 40)
    assert a.alloc(30) is None and a.largest() == 20
    a.release(s1, 40)
    assert a.alloc(30) == 0
    a.release(0, 30)
    a.release(s2, 40)
    assert a.free == [(0, 100)]


def test_result_record_is_json_line():
    """A finished request's record (the --jsonl output): arrival, admission,
    finish, latency and the step counters, serialisable as one JSON line."""
    import json
    from geoflood_amd.serving import Result
    r = Result(7, 256, 1.5, 1.75, 3.0, 120, 2, 1800.0, 2.5, 900, 1.0e5)
    line = json.dumps(r.record())
    back = json.loads(line)
    assert "\n" not in line
    assert back["rid"] == 7 and back["latency"] == 1.5 and back["cell_updates"] == 120 * 256 * 256
    assert back["admitted"] - back["arrival"] == 0.25 and back["status"] == 0
