"""Instability threshold scans over the bunch current.

No reference counterpart: Inovesa runs one current per job.  Here a scan
runs many VFP instances at once in one batch (vfps_batch_*), each from the
seeded Gaussian, and classifies each run with a detector of the bursting
(microbunching) instability; the threshold current for a chamber gap and a
damping time is bracketed by the stable and unstable currents.

Detector: the relative fluctuation std(P)/mean(P) of the CSR power P over
the last `window` synchrotron periods of the run, against `tol`.  With the
kick-drift maps a damped bunch below the threshold holds P constant to
1e-4 and less, and a bursting one fluctuates by 3e-3 and more: tol = 1e-3.
With the rotation maps the two levels are a few 1e-4 and 1e-2 and more:
tol = 5e-3.

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
from .workloads import derive, fp_diffusion_number, initial_mesh, scan_instance

C_LIGHT = 2.99792458e8


@dataclass
class ScanSetup:
    n: int = 128                 # mesh size of every run (unless sizes)
    steps_per_ts: int = 500      # time steps per synchrotron period
    periods: int = 30            # synchrotron periods per run
    window: int = 5              # periods at the end seen by the detector
    tol: float = None            # relative CSR fluctuation marking bursting
                                 # (None: 1e-3 kick-drift, 5e-3 rotation)
    seed: int = 1                # initial-noise seed (the same for all runs)
    scheme: str = "kick_drift"   # maps of every run (workloads.scan_instance)
    sizes: tuple = None          # an uneven mix of mesh sizes: each (gap,
                                 # damping) target runs on one of them

    def __post_init__(self):
        if self.tol is None:
            self.tol = 1e-3 if self.scheme == "kick_drift" else 5e-3
        if self.sizes:
            self.sizes = tuple(int(v) for v in self.sizes)

    def size_of(self, gap, damping):
        """Mesh size of the runs of target (gap, damping): n, or one of
        `sizes` chosen by a hash of the target, so that every scan (and
        the training data) puts a target on the same mesh.  Only sizes on
        which the explicit Fokker-Planck update of this damping time stays
        bounded are drawn: its diffusion number e1/dp^2 (e1 = 2/damping
        steps) grows with the square of the mesh size, and above 1/2 the
        update diverges (kept below 0.45)."""
        if not self.sizes:
            return self.n
        ok = [m for m in self.sizes
              if fp_diffusion_number(m, 2.0 / float(damping)) < 0.45]
        ok = ok or [min(self.sizes)]
        key = (int(round(float(gap) * 1e7)) * 2654435761
               + int(round(float(damping) * 1e3)) * 40503) & 0xFFFFFFFF
        return ok[(key >> 7) % len(ok)]

    @property
    def steps(self):
        return self.steps_per_ts * self.periods


def scaling_law(gap, setup):
    """Scaling-law threshold current (A) for a chamber gap (m) (the same on
    every mesh: the ring does not depend on it)."""
    return derive(grid_size=setup.n, steps_per_ts=setup.steps_per_ts,
                  vacuum_gap=gap)["Ith"]


def instance(factor, gap, damping_steps, setup):
    """One run: current factor * scaling-law threshold, parallel-plates CSR
    with the given full gap (m), damping time in time steps, the maps of
    setup.scheme, on the target's mesh (desc, derived, init)."""
    n = setup.size_of(gap, damping_steps)
    desc, d = scan_instance(n, factor, gap, damping_steps,
                            setup.steps_per_ts, setup.scheme)
    return desc, d, initial_mesh(n, seed=setup.seed)


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


def batch_class(count):
    """Batch-size class of `count` runs for the tiling cache: the next power
    of two, at least 16."""
    c = 16
    while c < count:
        c *= 2
    return c


class Runner:
    """Runs sets of (factor, gap, damping) points as one batch and keeps
    the wall time it spent and the number of VFP runs.  That time
    (gpu_seconds) covers a round's whole cost on the device and its host
    side: building every run's impedance table and the batch (with the
    direct wake's Toeplitz tables), tuning, and the runs.  The first batch
    of each mesh size and batch-size class is autotuned; later batches of
    that class reuse the tiling."""

    def __init__(self, setup=None, dev=None, spectra=False, snapshots=5,
                 tune=True, tiles=None):
        self.setup = setup or ScanSetup()
        self.dev = dev
        self.spectra = spectra
        self.snapshots = snapshots
        self.tune = tune
        self.gpu_seconds = 0.0
        self.build_seconds = 0.0     # part of gpu_seconds building batches
        self.tune_seconds = 0.0      # part of gpu_seconds spent tuning
        self.runs = 0
        self.tuning = None
        # (mesh size, batch class) -> (T, threads, persistent); a dict
        # shared between runners makes the tiling a one-time cost (like
        # FFTW wisdom)
        self.tiles = {} if tiles is None else tiles
        self.rounds = []             # (runs, GPU seconds) of every batch

    def run(self, points):
        s = self.setup
        t_build = time.perf_counter()
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
        b.dev.sync()
        dt = time.perf_counter() - t_build
        self.gpu_seconds += dt
        self.build_seconds += dt
        if self.tune:
            # tile shape per mesh size and batch-size class (a few runs want
            # short tiles for latency, many runs tall ones for throughput;
            # results do not depend on it); the tuning steps count as GPU
            # time of the scan
            t0 = time.perf_counter()
            counts = {}
            for desc in descs:
                counts[desc["n"]] = counts.get(desc["n"], 0) + 1
            keys = {n: (n, batch_class(c)) for n, c in counts.items()}
            if all(k in self.tiles for k in keys.values()):
                for n, k in keys.items():
                    b.retile_size(n, *self.tiles[k])
            else:
                self.tuning = b.autotune(steps=8)
                for r in self.tuning:
                    if r["chosen"] and r["n"] in keys:
                        self.tiles[keys[r["n"]]] = (r["T"], r["threads"],
                                                    r["persistent"], r["strip"])
            dt = time.perf_counter() - t0
            self.gpu_seconds += dt
            self.tune_seconds += dt
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
        dt_run = time.perf_counter() - t0
        self.gpu_seconds += dt_run
        self.runs += len(points)
        self.rounds.append((len(points), dt_run))
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


def tune_tiles(setup, count, gap=0.04, damping=2500.0):
    """The tiling of setup's mesh size, autotuned once on a batch of `count`
    runs (currents spread over 0.5-2x the scaling law): a dict to pass to
    Runner(tiles=...), and the GPU seconds the tuning took."""
    tiles = {}
    runner = Runner(setup, tiles=tiles)
    pts = [(float(f), gap, damping) for f in np.geomspace(0.5, 2.0, count)]
    descs, inits, zs = [], [], []
    for f, g, dmp in pts:
        desc, d, init = instance(f, g, dmp, setup)
        zw = V.ParallelPlatesCSR(d["padded_bins"], np.float32(d["f0"]),
                                 d["fmax"], g, dev=runner.dev)
        descs.append(desc)
        inits.append(init)
        zs.append(zw)
    b = V.Batch(descs, inits, zs, zs, dev=runner.dev)
    t0 = time.perf_counter()
    for r in b.autotune(steps=8):
        if r["chosen"]:
            tiles[(r["n"], batch_class(count))] = (r["T"], r["threads"],
                                                   r["persistent"], r["strip"])
    return tiles, time.perf_counter() - t0


def dense_scan(targets, setup=None, lo=0.3, hi=3.0, points=16, rel_tol=0.02,
               runner=None):
    """Plain threshold scan of every (gap, damping) in targets; returns
    (results, runner)."""
    runner = runner or Runner(setup)
    results = [ScanResult(g, dmp, lo, hi) for g, dmp in targets]
    _refine(runner, results, points, rel_tol)
    return results, runner


def surrogate_scan(model, targets, setup=None, lo=0.3, hi=3.0, points=16,
                   rel_tol=0.02, runner=None, window=0.03, near=None):
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
