"""Host logic of the batched ensemble (medyan_amd.ensemble), no device: the
ensemble's cylinder counts, the pass-rule names, and the autotuner's
coordinate search (threads per block under the first pass rule, then the
pass rule, then the share exponent; each candidate timed once, the best of
each stage kept) over a stand-in for mdy_minimize_batch."""
import numpy as np
import pytest

from medyan_amd import ensemble, mechanics


def test_ensemble_sizes_log_uniform_multiples_of_ten():
    s = ensemble.ensemble_sizes(256)
    assert len(s) == 256 and np.all(s % 10 == 0)
    assert s.min() >= 500 and s.max() <= 20000
    assert np.array_equal(s, ensemble.ensemble_sizes(256))   # seeded
    # log-uniform: about half of the counts below the geometric midpoint
    mid = np.sqrt(500 * 20000)
    assert 0.35 < np.mean(s < mid) < 0.65


def test_pass_rule_names():
    assert ensemble.pass_rule(0) == "one pass"
    assert ensemble.pass_rule(-1) == "one pass"
    assert ensemble.pass_rule(500) == "passes capped at 500 x 4^k force calls"
    assert ensemble.pass_rule(-8) == "passes end when 1/8 of their networks have finished"


class _Ctx:
    def __init__(self, w):
        self.w = w


def test_autotune_coordinate_search(monkeypatch):
    # a synthetic cost surface: 128 threads best, then the completion rule -4,
    # then the exponent 1.25; every call must restore the coordinates first
    calls, restores = [], []

    def fake_batch(ctxs, params, d_coords, threads_per_block=64, weights=None, first_pass_calls=0):
        a = round(float(np.log(weights[1] / weights[0]) / np.log(2.0)), 6)
        calls.append((threads_per_block, first_pass_calls, a))
        t = {64: 10.0, 128: 8.0, 256: 9.0}[threads_per_block]
        t += {-8: 1.0, -4: 0.5, -2: 2.0, 250: 3.0, 500: 3.0, 1000: 3.0, 2000: 3.0}[first_pass_calls]
        t += {1.0: 0.2, 1.25: 0.1}[a]
        assert len(restores) == len(calls)
        return None, None, t

    monkeypatch.setattr(mechanics, "minimize_batch", fake_batch)
    monkeypatch.setattr(ensemble, "work", lambda c: c.w)
    ctxs = [_Ctx(1.0), _Ctx(2.0)]   # work ratio 2: weights[1] / weights[0] = 2^a
    best, times = ensemble.autotune(ctxs, None, None, lambda: restores.append(1))
    assert best["threads_per_block"] == 128 and best["first_pass_calls"] == -4
    assert np.allclose(best["weights"], np.array([1.0, 2.0]) ** 1.25)
    # 3 block sizes under -8, 6 more pass rules at 128, 1 more exponent: each candidate once
    assert len(calls) == 3 + 6 + 1 == len(set(calls)) == len(times)
    assert calls[:3] == [(64, -8, 1.0), (128, -8, 1.0), (256, -8, 1.0)]
    assert times[(128, -4, 1.25)] == pytest.approx(8.6)
