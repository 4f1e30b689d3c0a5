"""LaunchTuner (seq_vp_3d_amd/batch.py) on the host: it times every candidate of a size class the same number of
times, keeps the fastest median once each has its timings, waits for unfinished events, and saves/loads its table."""
from seq_vp_3d_amd.batch import LaunchTuner


class FakeEvent:
    def __init__(self, t, done=True):
        self.t, self.done = t, done

    def query(self):
        return self.done

    def elapsed_time(self, end):
        return end.t - self.t


def test_tuner_times_each_candidate_then_keeps_the_fastest(tmp_path):
    cands = ((1, 0), (2, 0), (4, 512))
    cost = {(1, 0): 3., (2, 0): 1., (4, 512): 2.}
    tuner = LaunchTuner(cands, trials=2)
    seen = []
    for _ in range(6):
        cls, cand, timed = tuner.choose(5000)
        assert timed and cls == 12
        seen.append(cand)
        tuner.record(cls, cand, FakeEvent(0.), FakeEvent(cost[cand]))
    assert sorted(seen) == sorted(cands * 2)                  # round robin over the candidates
    cls, cand, timed = tuner.choose(5000)
    assert (cls, cand, timed) == (12, (2, 0), False)
    assert tuner.report()[12]["best"] == (2, 0)
    # another size class tunes on its own
    assert tuner.choose(40)[2]
    path = tmp_path / "tuned.json"
    tuner.save(path)
    again = LaunchTuner.load(path, candidates=cands)
    assert again.choose(5000) == (12, (2, 0), False)


def test_tuner_waits_for_unfinished_timings():
    cands = ((1, 0), (2, 0))
    tuner = LaunchTuner(cands, trials=1)
    end = FakeEvent(1., done=False)
    cls, cand, _ = tuner.choose(100)
    tuner.record(cls, cand, FakeEvent(0.), end)
    cls2, cand2, _ = tuner.choose(100)
    assert cand2 != cand                                       # the pending timing counts as taken
    tuner.record(cls2, cand2, FakeEvent(0.), FakeEvent(5.))
    assert tuner.choose(100)[2]                                # still tuning: one timing has not finished
    end.done = True
    assert tuner.choose(100) == (cls, cand, False)             # 1 ms beats 5 ms
