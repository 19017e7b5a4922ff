"""The learned surrogate of the threshold scan.

A multilayer perceptron 3 -> 256 -> 256 -> 256 -> 65 (GELU; BF16 weights,
FP32 accumulation on the matrix cores; trained and evaluated by the
library's own kernels, vfps_mlp_*) maps (bunch current, chamber gap,
damping time) to the 64-bin log CSR power spectrum of the run and the logit
of its instability verdict.  It is trained on runs of the batched VFP
solver; the scan (scan.surrogate_scan) asks it for the threshold current
and lets the VFP confirm or correct it.

Inputs are normalised to about [-1, 1]: log of the current in units of
the scaling-law threshold, the gap and the log of the damping time over
their training ranges.  Spectrum targets are standardised per bin.
"""
import json
import os

import numpy as np

from . import vfps as V
from .scan import Runner, ScanSetup

DIMS = (3, 256, 256, 256, 65)
# weights trained by `python -m inovesa_amd.pipeline` (default arguments:
# 128^2 runs), and by the same with --sizes 128,256,384,512 (the uneven mix
# of the batched scan)
DEFAULT_MODEL = os.path.join(os.path.dirname(os.path.abspath(__file__)),
                             "data", "surrogate.npz")
MIX_MODEL = os.path.join(os.path.dirname(os.path.abspath(__file__)),
                         "data", "surrogate_mix.npz")
MIX_SIZES = (128, 256, 384, 512)
FACTOR_RANGE = (0.3, 3.0)       # current / scaling-law threshold
FOCUS_RANGE = (0.75, 1.8)       # where the thresholds of the ranges below lie
GAP_RANGE = (0.016, 0.064)      # full chamber gap (m)
DAMPING_RANGE = (1000.0, 5000.0)  # damping time (time steps)


def features(factor, gap, damping):
    f = np.log(np.asarray(factor, np.float64))
    lf = np.log(FACTOR_RANGE)
    g = np.asarray(gap, np.float64)
    ld = np.log(DAMPING_RANGE)
    dm = np.log(np.asarray(damping, np.float64))
    f, g, dm = np.broadcast_arrays(f, g, dm)
    x = np.stack([2 * (f - lf[0]) / (lf[1] - lf[0]) - 1,
                  2 * (g - GAP_RANGE[0]) / (GAP_RANGE[1] - GAP_RANGE[0]) - 1,
                  2 * (dm - ld[0]) / (ld[1] - ld[0]) - 1], axis=-1)
    return x.astype(np.float32)


def sample_points(count, seed, focus=0.0):
    """Training points: current log-uniform, gap uniform, damping
    log-uniform over the ranges above; a fraction `focus` of the currents
    log-uniform over FOCUS_RANGE instead, where the verdict changes."""
    rng = np.random.Generator(np.random.Philox(seed))
    f = np.exp(rng.uniform(*np.log(FACTOR_RANGE), count))
    near = rng.random(count) < focus
    f[near] = np.exp(rng.uniform(*np.log(FOCUS_RANGE), int(near.sum())))
    g = rng.uniform(*GAP_RANGE, count)
    dmp = np.exp(rng.uniform(*np.log(DAMPING_RANGE), count))
    return [(float(a), float(b), float(c)) for a, b, c in zip(f, g, dmp)]


def generate(count, setup=None, seed=7, chunk=1024, dev=None, log=None,
             focus=0.0):
    """Run `count` random points on the batched solver (chunk instances per
    batch); returns (points, spectra [count][64], labels, fluct, gpu s)."""
    runner = Runner(setup or ScanSetup(), dev=dev, spectra=True)
    pts = sample_points(count, seed, focus)
    spec, lab, fl = [], [], []
    for k in range(0, count, chunk):
        res = runner.run(pts[k:k + chunk])
        spec += [r.spectrum for r in res]
        lab += [float(r.unstable) for r in res]
        fl += [r.fluct for r in res]
        if log:
            log(f"generated {min(k + chunk, count)}/{count} runs, "
                f"{runner.gpu_seconds:.1f} s of GPU time")
    return (np.array(pts), np.array(spec, np.float32),
            np.array(lab, np.float32), np.array(fl), runner.gpu_seconds)


class Surrogate:
    """The network with its input and output normalisation."""

    def __init__(self, seed=0, dev=None, max_batch=512):
        self.net = V.MLP(DIMS, max_batch=max_batch, seed=seed, dev=dev)
        self.mean = np.zeros(64, np.float32)
        self.std = np.ones(64, np.float32)

    def fit(self, points, spectra, labels, epochs=200, batch=256, lr=1e-3,
            label_weight=1.0, seed=0):
        pts = np.asarray(points, np.float64)
        x = features(pts[:, 0], pts[:, 1], pts[:, 2])
        spectra = np.asarray(spectra, np.float32)
        self.mean = spectra.mean(axis=0).astype(np.float32)
        self.std = (spectra.std(axis=0) + 1e-3).astype(np.float32)
        t = (spectra - self.mean) / self.std
        return self.net.train(x, t, labels, batch=batch, epochs=epochs, lr=lr,
                              label_weight=label_weight, seed=seed)

    def predict(self, factor, gap, damping):
        """(log spectrum [..][64], instability logit [..])."""
        x = features(factor, gap, damping).reshape(-1, 3)
        outs = []
        for k in range(0, len(x), self.net.max_batch):
            outs.append(self.net.forward(x[k:k + self.net.max_batch]))
        y = np.concatenate(outs)
        return y[:, :64] * self.std + self.mean, y[:, 64]

    def predict_threshold(self, gap, damping, lo=FACTOR_RANGE[0],
                          hi=FACTOR_RANGE[1], grid=512, band=False,
                          p_band=0.1):
        """Current factor where the predicted logit first turns positive
        (log-linear interpolation on a grid); with band=True also the
        factors where the verdict probability first passes p_band and
        1 - p_band around it."""
        f = np.geomspace(lo, hi, grid)
        _, logit = self.predict(f, gap, damping)

        def crossing(level):
            pos = np.nonzero(logit > level)[0]
            if len(pos) == 0:
                return float(hi)
            k = int(pos[0])
            if k == 0:
                return float(lo)
            a, b = logit[k - 1] - level, logit[k] - level
            t = a / (a - b)
            return float(np.exp(np.log(f[k - 1])
                                + t * (np.log(f[k]) - np.log(f[k - 1]))))

        f0 = crossing(0.0)
        if not band:
            return f0
        lev = float(np.log(p_band / (1 - p_band)))
        return f0, (min(crossing(lev), f0), max(crossing(-lev), f0))

    def save(self, path):
        np.savez_compressed(path, params=self.net.params(), mean=self.mean,
                            std=self.std, dims=np.array(DIMS))

    @classmethod
    def load(cls, path=DEFAULT_MODEL, dev=None):
        z = np.load(path)
        assert tuple(z["dims"]) == DIMS
        s = cls(dev=dev)
        s.net.set_params(z["params"])
        s.mean = z["mean"]
        s.std = z["std"]
        return s
