"""Instability threshold scans over the bunch current.

No reference counterpart: Inovesa runs one current per job.  Here a scan
runs many VFP instances at once in one batch (vfps_batch_*), each from the
seeded Gaussian, and classifies each run with a detector of the bursting
(microbunching) instability; the threshold current for a chamber gap and a
damping time is bracketed by the stable and unstable currents.

Detector: the relative fluctuation std(P)/mean(P) of the CSR power P over
the last `window` synchrotron periods of the run.  Below the threshold the
damped bunch settles and P is constant to a few 1e-4; above it the bunch
bursts and P fluctuates by 1e-2 and more; `tol` (5e-3) separates the two.

Dense scan (the plain method): rounds of `points` currents spaced
geometrically over the bracket, starting from [lo, hi] times the
scaling-law threshold Ith = Inorm (0.5 + 0.34 Pi) (src/main.cpp:383-390);
each round keeps the stable/unstable pair around the first unstable
current, until the bracket is narrower than rel_tol.

Surrogate-steered scan: the network (surrogate.py) predicts the threshold
f* and the band of currents where its verdict is uncertain; the VFP runs
one geometric grid over that band (at least f*/(1+w) .. f*(1+w)) whose
spacing already meets rel_tol; where the transition lies inside the grid
the scan is done in one round, otherwise the dense rounds continue from the
bracket the grid leaves.
"""
import time
from dataclasses import dataclass, field

import numpy as np

from . import vfps as V
from .workloads import derive, initial_mesh

C_LIGHT = 2.99792458e8


@dataclass
class ScanSetup:
    n: int = 128                 # mesh size of every run
    steps_per_ts: int = 500      # time steps per synchrotron period
    periods: int = 30            # synchrotron periods per run
    window: int = 5              # periods at the end seen by the detector
    tol: float = 5e-3            # relative CSR fluctuation marking bursting
    seed: int = 1                # initial-noise seed (the same for all runs)

    @property
    def steps(self):
        return self.steps_per_ts * self.periods


def scaling_law(gap, setup):
    """Scaling-law threshold current (A) for a chamber gap (m)."""
    return derive(grid_size=setup.n, steps_per_ts=setup.steps_per_ts,
                  vacuum_gap=gap)["Ith"]


def instance(factor, gap, damping_steps, setup):
    """One run: current factor * scaling-law threshold, parallel-plates CSR
    with the given full gap (m), damping time in time steps; rotation
    scheme with the cubic clamped RotationMap (desc, derived, init)."""
    base = derive(grid_size=setup.n, steps_per_ts=setup.steps_per_ts,
                  vacuum_gap=gap)
    d = derive(grid_size=setup.n, steps_per_ts=setup.steps_per_ts,
               vacuum_gap=gap, damping_time=damping_steps * base["dt"],
               bunch_current=factor * base["Ith"])
    desc = dict(n=setup.n, qmin=d["qmin"], qmax=d["qmax"], pmin=d["pmin"],
                pmax=d["pmax"], qscale=d["qscale"], pscale=d["pscale"],
                scheme=1, interp=4, rot_clamp=1, angle=d["angle"],
                rf_linear=1, rf_angle=d["angle"],
                revolutionpart=d["revolutionpart"], V_RF=d["V_eff"],
                f_RF=d["f_RF"], V0=d["V0"], slip=d["slip"],
                E0=d["ring"]["beam_energy"], use_fp=1, fptype=3, e1=d["e1"],
                deriv=4, use_wake=1, spacing_bins=d["spacing_bins"],
                wakescaling=d["wakescaling"], use_csr=1,
                fc=d["ring"]["cutoff_frequency"])
    return desc, d, initial_mesh(setup.n, seed=setup.seed)


def detect(csr, setup):
    """(unstable, fluctuation) of one CSR power series (last window)."""
    last = np.asarray(csr[-setup.window * setup.steps_per_ts:], np.float64)
    mean = last.mean()
    fl = float(last.std() / abs(mean)) if mean != 0 else float("inf")
    return (not np.isfinite(fl)) or fl > setup.tol, fl


@dataclass
class RunResult:
    factor: float
    gap: float
    damping: float
    unstable: bool
    fluct: float
    csr_mean: float
    profile: np.ndarray = field(repr=False, default=None)
    spectrum: np.ndarray = field(repr=False, default=None)


class Runner:
    """Runs sets of (factor, gap, damping) points as one batch and keeps
    the GPU wall time and the number of VFP runs it spent."""

    def __init__(self, setup=None, dev=None, spectra=False, snapshots=5,
                 tune=True):
        self.setup = setup or ScanSetup()
        self.dev = dev
        self.spectra = spectra
        self.snapshots = snapshots
        self.tune = tune
        self.gpu_seconds = 0.0
        self.runs = 0
        self.tuning = None

    def run(self, points):
        s = self.setup
        descs, inits, zws, zcs, ds = [], [], [], [], []
        for f, g, dmp in points:
            desc, d, init = instance(f, g, dmp, s)
            # wake and CSR impedances are the same parallel-plates table
            zw = V.ParallelPlatesCSR(d["padded_bins"], np.float32(d["f0"]),
                                     d["fmax"], g, dev=self.dev)
            zc = zw
            descs.append(desc)
            inits.append(init)
            zws.append(zw)
            zcs.append(zc)
            ds.append(d)
        b = V.Batch(descs, inits, zws, zcs, dev=self.dev)
        if self.tune:
            # tile shape per mesh size (results do not depend on it); the
            # tuning steps count as GPU time of the scan
            t0 = time.perf_counter()
            self.tuning = b.autotune(steps=8)
            self.gpu_seconds += time.perf_counter() - t0
        w = s.window * s.steps_per_ts
        b.record(s.steps)
        spectra = [np.zeros(64) for _ in points] if self.spectra else None
        t0 = time.perf_counter()
        if self.spectra:
            b.run(s.steps - w)
            chunk = w // self.snapshots
            for k in range(self.snapshots):
                b.run(chunk if k < self.snapshots - 1 else w - chunk * k)
                for i, p in enumerate(b.projections()):
                    spectra[i] += csr_spectrum(p, zcs[i], ds[i])
        else:
            b.run(s.steps)
        b.dev.sync()
        self.gpu_seconds += time.perf_counter() - t0
        self.runs += len(points)
        hist = b.csr_histories(0, s.steps)
        profs = b.projections()
        out = []
        for i, (f, g, dmp) in enumerate(points):
            un, fl = detect(hist[i], s)
            r = RunResult(f, g, dmp, un, fl,
                          float(np.mean(hist[i][-w:])), profs[i])
            if spectra is not None:
                r.spectrum = np.log10(spectra[i] / self.snapshots + 1e-30)
            out.append(r)
        return out


def csr_spectrum(profile, zc, d, bins=64):
    """CSR power spectrum of a bunch profile in `bins` log-spaced frequency
    bins: Re Z(f_k) |F_k|^2 (the terms of src/PS/ElectricField.cpp:133-184
    without the cutoff and unit factors), F the zero-padded transform."""
    nmax = d["padded_bins"]
    F = np.fft.rfft(np.asarray(profile, np.float64), nmax)
    z = zc.data()[:nmax // 2 + 1].real.astype(np.float64)
    s = z * np.abs(F) ** 2
    edges = np.unique(np.geomspace(1, nmax // 2 + 1, bins + 1).astype(int))
    out = np.zeros(bins)
    for k in range(bins):
        a = edges[min(k, len(edges) - 2)]
        b = max(edges[min(k + 1, len(edges) - 1)], a + 1)
        out[k] = s[a:b].mean()
    return np.maximum(out, 0.0)


@dataclass
class ScanResult:
    gap: float
    damping: float
    lo: float                    # highest stable factor seen
    hi: float                    # lowest unstable factor seen
    runs: int = 0
    surrogate: bool = False
    confirmed: bool = False      # the surrogate's bracket held
    predicted: float = 0.0       # the surrogate's threshold

    @property
    def threshold(self):
        return float(np.sqrt(self.lo * self.hi))

    @property
    def width(self):
        return self.hi / self.lo - 1.0


def _refine(runner, results, points=16, rel_tol=0.02, max_rounds=8):
    """Dense rounds on every unresolved scan point together."""
    for _ in range(max_rounds):
        todo = [r for r in results if r.width > rel_tol * (1 + 1e-9)]
        if not todo:
            break
        pts, owner = [], []
        for r in todo:
            fs = np.geomspace(r.lo, r.hi, points + 2)[1:-1]
            for f in fs:
                pts.append((float(f), r.gap, r.damping))
                owner.append(r)
        res = runner.run(pts)
        for r in todo:
            mine = [x for x, o in zip(res, owner) if o is r]
            r.runs += len(mine)
            for x in sorted(mine, key=lambda x: x.factor):
                if x.unstable:
                    r.hi = min(r.hi, x.factor)
                    break
                r.lo = max(r.lo, x.factor)
            # a stable run above an unstable one leaves the first unstable
            # as the upper end (the detector is monotone in practice)
    return results


def dense_scan(targets, setup=None, lo=0.3, hi=3.0, points=16, rel_tol=0.02,
               runner=None):
    """Plain threshold scan of every (gap, damping) in targets; returns
    (results, runner)."""
    runner = runner or Runner(setup)
    results = [ScanResult(g, dmp, lo, hi) for g, dmp in targets]
    _refine(runner, results, points, rel_tol)
    return results, runner


def surrogate_scan(model, targets, setup=None, lo=0.3, hi=3.0, points=16,
                   rel_tol=0.02, runner=None, window=0.04, near=None):
    """Threshold scan steered by the surrogate: `near` VFP runs per target
    spaced geometrically over [f*/(1+window), f*(1+window)] around the
    predicted threshold f* (by default as many as bring the bracket under
    rel_tol in one round); dense rounds continue only where the runs do not
    contain the transition."""
    runner = runner or Runner(setup)
    preds, wins = [], []
    for g, dmp in targets:
        f, (a, b) = model.predict_threshold(g, dmp, lo, hi, band=True)
        # the window: at least +-window around f*, widened to the band where
        # the network is unsure (verdict probability 0.1-0.9)
        a = max(lo, min(f / (1 + window), a))
        b = min(hi, max(f * (1 + window), b))
        preds.append(f)
        wins.append((a, b))
    pts, owner = [], []
    for k, ((g, dmp), (a, b)) in enumerate(zip(targets, wins)):
        m = near or int(np.ceil(np.log(b / a) / np.log1p(rel_tol))) + 1
        for x in np.geomspace(a, b, max(2, m)):
            pts.append((float(x), g, dmp))
            owner.append(k)
    res = runner.run(pts)
    results = []
    for k, (g, dmp) in enumerate(targets):
        mine = sorted((x for x, o in zip(res, owner) if o == k),
                      key=lambda x: x.factor)
        r = ScanResult(g, dmp, lo, hi, runs=len(mine), surrogate=True,
                       predicted=preds[k])
        for x in mine:
            if x.unstable:
                r.hi = x.factor
                break
            r.lo = x.factor
        # the transition lies between two of the proposed currents
        r.confirmed = (r.lo >= mine[0].factor and r.hi <= mine[-1].factor)
        results.append(r)
    # a missed target's grid is all stable (go up from its top) or all
    # unstable (go down from its bottom)
    _expand(runner, [(r, r.hi >= hi) for r in results if not r.confirmed],
            points, rel_tol)
    _refine(runner, results, points, rel_tol)
    return results, runner


def _expand(runner, missed, points=16, rel_tol=0.02):
    """One round for the targets whose transition lies outside the proposed
    grid, given as (result, upwards): `points` currents spaced by
    (1 + rel_tol) beyond the grid edge the transition lies past, so that a
    miss of up to (1 + rel_tol)^points is resolved in this round."""
    pts, owner = [], []
    for r, up in missed:
        k = np.arange(1, points + 1) * np.log1p(rel_tol)
        fs = r.lo * np.exp(k) if up else r.hi * np.exp(-k)
        for f in fs[(fs > r.lo) & (fs < r.hi)]:
            pts.append((float(f), r.gap, r.damping))
            owner.append(r)
    if not pts:
        return
    res = runner.run(pts)
    for r, _ in missed:
        mine = [x for x, o in zip(res, owner) if o is r]
        r.runs += len(mine)
        for x in sorted(mine, key=lambda x: x.factor):
            if x.unstable:
                r.hi = min(r.hi, x.factor)
                break
            r.lo = max(r.lo, x.factor)
