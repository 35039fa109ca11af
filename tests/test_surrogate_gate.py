"""The surrogate gate's calibration rule on the host, with a stub ensemble
whose predictions are fixed: the threshold is the largest ensemble std s such
that every calibration point with std <= s is within tol, accepted only with
at least min_points such points; ties at s count as above it; the threshold
is a std, never capped by tol."""
import numpy as np

from pauliprop import surrogate as sg


class Stub:
    def __init__(self, mean, sd):
        self.mean, self.sd = np.asarray(mean, float), np.asarray(sd, float)

    def predict(self, x):
        return self.mean, self.sd, None, 0.0


def gate(mean, sd, tol=0.1, min_points=1):
    g = sg.GatedSurrogate(Stub(mean, sd), exact=lambda t: np.full(len(t), 9.0), tol=tol, min_points=min_points)
    return g


TH = np.zeros((6, 20))


def test_threshold_is_the_last_std_before_the_first_miss():
    # std order: 0.01 ok, 0.02 ok, 0.03 ok, 0.5 miss, 0.6 ok, 0.7 ok
    sd = [0.03, 0.5, 0.01, 0.7, 0.02, 0.6]
    err = [0.0, 0.2, 0.05, 0.0, 0.09, 0.0]
    g = gate(err, sd)
    cal = g.calibrate(TH, np.zeros(6))
    assert cal["threshold"] == 0.03 and cal["within_tol_prefix"] == 3
    assert list(g.serves(np.array([0.03, 0.031, 0.0]))) == [True, False, True]


def test_threshold_is_a_std_not_capped_by_tol():
    g = gate([0.0] * 6, [2.0, 3.0, 4.0, 5.0, 6.0, 7.0], tol=1e-3)
    assert g.calibrate(TH, np.zeros(6))["threshold"] == 7.0


def test_ties_across_the_first_miss_count_as_above():
    sd = [0.1, 0.2, 0.2, 0.3, 0.4, 0.5]
    err = [0.0, 0.0, 0.5, 0.0, 0.0, 0.0]  # a miss tied at std 0.2 with a hit
    g = gate(err, sd)
    cal = g.calibrate(TH, np.zeros(6))
    assert cal["threshold"] == 0.1 and cal["within_tol_prefix"] == 1


def test_min_points_and_no_hit_serve_nothing():
    sd = [0.1, 0.2, 0.3, 0.4, 0.5, 0.6]
    err = [0.0, 0.0, 0.5, 0.0, 0.0, 0.0]
    g = gate(err, sd, min_points=3)  # only 2 points before the first miss
    assert g.calibrate(TH, np.zeros(6))["threshold"] < 0
    assert not g.serves(np.array([0.0, 0.1])).any()
    g = gate([1.0] * 6, sd)  # the most confident point already misses
    assert g.calibrate(TH, np.zeros(6))["threshold"] < 0
    vals, fall, _ = g.evaluate(TH)
    assert fall.all() and np.all(vals == 9.0)
