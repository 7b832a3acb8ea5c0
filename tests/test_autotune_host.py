"""The step-variant cache (geoflood_amd.autotune.TuneCache) on the host: the
population buckets, a choice timed once per (patch size, method, bucket),
cached choices applied without timing, and the JSON file a later run reads.
A stand-in pool records what the cache asks of it (no device)."""
import json

from geoflood_amd.autotune import TuneCache


class _Pool:
    def __init__(self, mx, best=2, nvar=4):
        self.mx, self.best, self.nvar = mx, best, nvar
        self.tuned = 0
        self.variant = 0

    def autotune(self, reps=2):
        self.tuned += 1
        self.variant = self.best
        return self.best, [0.1 * (k + 1) for k in range(self.nvar)]

    def set_variant(self, v):
        self.variant = v


def test_buckets_are_rounded_log2_of_the_population():
    assert [TuneCache.bucket(a) for a in (0, 1, 1.4, 1.5, 3, 1000, 1448, 1449)] == [0, 0, 0, 1, 2, 10, 10, 11]
    assert TuneCache.key(16, "default", 590.5) == "16:default:9"


def test_a_bucket_is_timed_once_and_its_choice_reused(tmp_path):
    path = str(tmp_path / "variants.json")
    cache = TuneCache(path)
    p = _Pool(16, best=3)
    assert cache.choose(p, 600.0, levels={3: 86, 4: 3752}) == 3 and p.tuned == 1
    assert cache.choose(p, 500.0) == 3 and p.tuned == 1          # same bucket (9): no timing
    p.variant = 0
    assert cache.choose(p, 700.0) == 3 and p.variant == 3 and p.tuned == 1
    assert cache.choose(p, 2000.0) == 3 and p.tuned == 2         # bucket 11: timed
    assert [e["cached"] for e in cache.log] == [False, True, True, False]
    on_disk = json.load(open(path))
    assert set(on_disk) == {"16:default:9", "16:default:11"}
    assert on_disk["16:default:9"]["levels"] == {"3": 86, "4": 3752}
    assert on_disk["16:default:9"]["ms"] == [0.1, 0.2, 0.3, 0.4]


def test_a_later_run_reads_the_file_and_times_nothing(tmp_path):
    path = str(tmp_path / "variants.json")
    first = TuneCache(path)
    first.choose(_Pool(32, best=1), 1500.0)
    second = TuneCache(path)
    p = _Pool(32, best=2)
    assert second.choose(p, 1600.0) == 1 and p.tuned == 0 and p.variant == 1   # bucket 11 again
    # another patch size or method is another key
    assert second.choose(_Pool(16, best=2), 1600.0) == 2
    assert second.choose(p, 1600.0, method="order1") == 2 and p.tuned == 1


def test_no_path_keeps_choices_in_memory(tmp_path):
    cache = TuneCache()
    p = _Pool(8, best=3)
    assert cache.choose(p, 5000.0) == 3 and cache.choose(p, 5000.0) == 3 and p.tuned == 1
    assert not list(tmp_path.iterdir())
