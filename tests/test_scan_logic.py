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


def test_mesh_mix_assigns_each_target_one_size():
    """ScanSetup(sizes=...): a target (gap, damping) always runs on the same
    mesh (dense scan, steered scan and training data agree), and the
    targets of a scan spread over every size."""
    from inovesa_amd.pipeline import scan_targets
    s = scan.ScanSetup(sizes=(128, 256, 384, 512))
    t = scan_targets(128)
    sizes = [s.size_of(g, d) for g, d in t]
    assert sizes == [s.size_of(g, d) for g, d in t]
    counts = {n: sizes.count(n) for n in (128, 256, 384, 512)}
    assert counts[128] >= 16 and counts[256] >= 16, counts
    assert scan.ScanSetup().size_of(0.03, 2000.0) == 128
    # only meshes on which the explicit FP update of the damping time
    # stays bounded (diffusion number below 1/2)
    from inovesa_amd.workloads import fp_diffusion_number
    for (g, d), n in zip(t, sizes):
        assert fp_diffusion_number(n, 2.0 / d) < 0.45
    assert {s.size_of(0.03, 9000.0 + k) for k in range(200)} == {
        128, 256, 384, 512}
