"""Host logic of the threshold scans (inovesa_amd/scan.py) with a stand-in
for the VFP runs: a run is unstable exactly above a known threshold."""
import numpy as np
import pytest

from inovesa_amd import scan


class FakeRunner:
    def __init__(self, thresholds):
        self.th = thresholds            # (gap, damping) -> factor
        self.runs = 0
        self.rounds = 0
        self.gpu_seconds = 0.0

    def run(self, points):
        self.rounds += 1
        self.runs += len(points)
        self.gpu_seconds += 1.0
        return [scan.RunResult(f, g, d, f > self.th[(g, d)], 0.0, 1.0)
                for f, g, d in points]


class FakeModel:
    def __init__(self, thresholds, error):
        self.th, self.err = thresholds, error

    def predict_threshold(self, g, d, lo, hi, band=False):
        f = self.th[(g, d)] * (1 + self.err)
        return (f, (f, f)) if band else f


TARGETS = [(0.02, 1500.0), (0.045, 3000.0), (0.06, 4500.0)]
TH = {TARGETS[0]: 1.2048, TARGETS[1]: 1.1395, TARGETS[2]: 0.7123}


def test_dense_scan_brackets_the_threshold():
    r = FakeRunner(TH)
    res, _ = scan.dense_scan(TARGETS, runner=r, rel_tol=0.02)
    for x in res:
        t = TH[(x.gap, x.damping)]
        assert x.lo <= t <= x.hi and x.width <= 0.02
        assert abs(x.threshold / t - 1) < 0.01
    assert r.rounds == 2 and r.runs == 2 * 16 * len(TARGETS)


@pytest.mark.parametrize("error", [0.0, 0.02, -0.025])
def test_surrogate_scan_confirms_good_predictions(error):
    """A prediction within the +-3% window is confirmed in one round."""
    r = FakeRunner(TH)
    res, _ = scan.surrogate_scan(FakeModel(TH, error), TARGETS, runner=r,
                                 rel_tol=0.02)
    assert r.rounds == 1
    for x in res:
        t = TH[(x.gap, x.damping)]
        assert x.confirmed and x.lo <= t <= x.hi and x.width <= 0.02 + 1e-9


@pytest.mark.parametrize("error,rounds", [(-0.035, 2), (0.15, 2), (-0.15, 2),
                                          (0.25, 2), (-0.6, 4), (1.0, 4)])
def test_surrogate_scan_falls_back(error, rounds):
    """A miss of up to 1.02^16 = 1.37 beyond the grid costs one expansion
    round; larger misses continue with the dense rounds."""
    r = FakeRunner(TH)
    res, _ = scan.surrogate_scan(FakeModel(TH, error), TARGETS, runner=r,
                                 rel_tol=0.02)
    assert r.rounds == rounds
    for x in res:
        t = TH[(x.gap, x.damping)]
        assert not x.confirmed and x.lo <= t <= x.hi and x.width <= 0.02 + 1e-9


class BandModel(FakeModel):
    """Predicts 8% high but reports a band of +-10% around it."""

    def predict_threshold(self, g, d, lo, hi, band=False):
        f = self.th[(g, d)] * 1.08
        return (f, (f / 1.1, f * 1.1)) if band else f


def test_surrogate_band_widens_the_grid():
    r = FakeRunner(TH)
    res, _ = scan.surrogate_scan(BandModel(TH, 0.0), TARGETS, runner=r,
                                 rel_tol=0.02)
    assert r.rounds == 1
    n_band = int(np.ceil(np.log(1.1 ** 2) / np.log(1.02))) + 1
    assert r.runs == len(TARGETS) * n_band
    for x in res:
        t = TH[(x.gap, x.damping)]
        assert x.confirmed and x.lo <= t <= x.hi and x.width <= 0.02 + 1e-9


def test_detector():
    s = scan.ScanSetup(steps_per_ts=10, window=2, tol=5e-3)
    flat = np.full(100, 2.0) + 1e-4 * np.sin(np.arange(100))
    burst = np.full(100, 2.0) * (1 + 0.2 * np.sin(np.arange(100) / 3.0))
    assert scan.detect(flat, s)[0] is False
    assert scan.detect(burst, s)[0] is True
    assert scan.detect(np.full(100, np.nan), s)[0] is True
