"""Benchmark of the per-step phase-space update on one MI355X.

Workload (the configuration the headline metric is quoted on): plain VFP,
one 512x512 FP32 mesh, kick-drift scheme (wake kick, sinusoidal RF kick,
linear drift, Fokker-Planck), parallel-plates CSR impedance with a 32 mm
gap, FFT length 4096, current 1.5x the scaling-law threshold, 2000 steps per
synchrotron period; the bunch profile and the CSR power are recorded on the
device every step.  One "step" is one time step of the mesh.

Prints one JSON line.  value = mesh-cell updates per second
(n^2 * steps / wall time of the timed steps, state resident on the GPU).
The x projection is summed in the reference's sequential order, bit for
bit (--projection wavefront: as a wavefront reduction; other_projection
reports the order not chosen).

Further keys of the line:
  roofline      the step kernel (the whole step in one launch; 8 bytes of
                mesh traffic per cell), timed with HIP events on its stream;
  fft_wake      the same workload with the wake and CSR power from the FFT
                field kernel;
  per_map       the same workload one kernel per map (the per-map path an
                unchanged reference loop drives);
  batched       64 independent meshes of 128^2-512^2 (kick-drift maps), one
                launch per step and size class;
  batched_rotation  the same 64 meshes on the rotation maps (wake kick,
                cubic clamped rotation, Fokker-Planck);
  production    16 meshes of 1024^2 with the large production maps and
                impedance;
  scan          the threshold scan with and without the trained network,
                targets on an uneven mix of 128^2-512^2 meshes;
  cpu_baseline  the oracle (the FP32 C restatement of the reference CPU
                path), compiled on the host with -O3 -march=native, on one
                host core: 50 warm-up steps, then the median of 5 runs.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("mesh-cell updates/s at 512²; batched+surrogate threshold scan "
          "time vs plain @2%")
HBM_PEAK_GBS = 8000.0


# tilings the autotuner chose on the MI355X for the batch lines (mesh size:
# rows per tile, workgroup size, persistent, tiles per strip), used by
# --only so that a profiler run times the same kernels without the tuner's
# launches
PROFILE_TILES = {
    "batched": {128: (2, 256, False, 0), 256: (4, 256, False, 0),
                384: (8, 512, False, 0), 512: (8, 512, False, 0)},
    "batched_rotation": {128: (2, 256, False, 0), 256: (4, 256, False, 0),
                         384: (8, 512, False, 0), 512: (8, 512, False, 0)},
    "production": (4, 512, False, 8),
}


def host_cpu():
    """nproc and the lscpu model name of the host that times the baseline."""
    import subprocess
    model = None
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True,
                             timeout=20).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                model = line.split(":", 1)[1].strip()
                break
    except (OSError, subprocess.SubprocessError):
        pass
    return os.cpu_count(), model


def cpu_baseline(desc, zw, zc, init, budget_s=15.0, warmup=50, runs=5):
    """The FP32 restatement of the reference CPU path, compiled on this host
    with the reference's release flags (-O3 -march=native), on one thread:
    `warmup` steps, then the median of `runs` timed runs of the same number
    of steps (about budget_s seconds in all), each continuing the state."""
    import numpy as np
    from oracle.build_oracle import build_native
    from oracle.oracle import Oracle
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from _helpers import oracle_cfg
    o = Oracle(path=build_native())
    sim = o.sim(oracle_cfg(desc, zw.n), init, zw.data(), zc.data())
    t0 = time.perf_counter()
    sim.run(warmup, want_csr=True)
    per = (time.perf_counter() - t0) / warmup
    steps = max(10, int(budget_s / runs / per))
    times = []
    for _ in range(runs):
        t0 = time.perf_counter()
        sim.run(steps, want_csr=True)
        times.append(time.perf_counter() - t0)
    dt = float(np.median(times))
    n = desc["n"]
    nproc, model = host_cpu()
    return {"value": n * n * steps / dt, "unit": "cell-updates/s", "cores": 1,
            "kind": "port",
            "sample": f"oracle FP32 restatement built with -O3 -march=native "
                      f"on this host, the same {n}x{n} kick-drift workload "
                      f"from the same initial state, one thread: {warmup} "
                      f"warm-up steps, then the median of {runs} runs of "
                      f"{steps} steps ({dt:.2f} s each); FFT: the "
                      f"restatement's own (FFTW is not available)",
            "nproc": nproc, "cpu_model": model,
            "run_seconds": [round(t, 3) for t in times]}


def launch_label(steps, no_graph, batch=False):
    """How a run of `steps` steps is launched: a single mesh's last graph
    also closes the run (last CSR power, error flag); a batch launches its
    close after the graphs."""
    if no_graph:
        return "host launches"
    full, rest = divmod(steps, 64)
    parts = []
    if full:
        parts.append(f"{full} replays of a 64-step hipGraph")
    if rest:
        parts.append(f"one {rest}-step hipGraph")
    return (" + ".join(parts) + " (captured before the timed region), "
            + ("then one finalize launch" if batch
               else "the last of them closing the run"))


def form_label(r):
    """Launch form of an autotuner row: per-tile, persistent, or strips of
    k tiles."""
    if r.get("strip"):
        return f"strips of {r['strip']} tiles"
    return "persistent" if r["persistent"] else "per-tile"


def batched_throughput(steps, log, scheme="kick_drift", tiles=None,
                       no_graph=False, wake_method=2):
    """64 independent meshes of uneven sizes (16 each of 128^2, 256^2,
    384^2, 512^2; currents log-spaced over 0.3-3x threshold; gaps 16-64 mm,
    each its own parallel-plates impedance) advanced together after the
    autotuner picked the tiles: whole-batch cell updates per second.
    scheme "rotation": the maps of the 256^2 rotation workload (wake kick,
    cubic clamped rotation, Fokker-Planck); "kick_drift": those of the
    512^2 workload.  wake_method 1: every instance with the FFT wake (one
    field-kernel launch per step over all of them), 2: the direct wake."""
    import numpy as np
    from inovesa_amd import Batch
    from inovesa_amd.workloads import batched_scan_config, impedances
    cfg = batched_scan_config(scheme=scheme)
    zs = [impedances(d) for _, d, _ in cfg]
    b = Batch([dict(c[0], no_graph=int(no_graph), wake_method=wake_method)
               for c in cfg],
              [c[2] for c in cfg], [z[0] for z in zs], [z[1] for z in zs])
    if tiles:
        for size, (T, thr, persistent, strip) in tiles.items():
            b.retile_size(size, T, thr, persistent=persistent, strip=strip)
        table = [dict(n=size, T=T, threads=thr, persistent=persistent,
                      strip=strip, us_per_step=0.0, chosen=True)
                 for size, (T, thr, persistent, strip) in tiles.items()]
    else:
        table = b.autotune(steps=10)
    cells = b.info()["cells"]
    b.record(steps + 128)
    b.run(128)
    b.prepare(steps)
    b.dev.sync()
    t0 = time.perf_counter()
    b.run(steps)
    b.dev.sync()
    dt = time.perf_counter() - t0
    chosen = {r["n"]: (r["T"], r["threads"], form_label(r),
                       round(r["us_per_step"], 2))
              for r in table if r["chosen"]}
    log(f"batched ({scheme}, wake {wake_method}): {len(cfg)} instances, "
        f"{cells} cells, {steps} steps in {dt:.2f} s")
    return {"value": cells * steps / dt, "unit": "cell-updates/s",
            "maps": scheme, "instances": len(cfg),
            "wake": "FFT field kernel" if wake_method == 1
                    else "direct (FP64 Toeplitz)",
            "cells_per_step": int(cells),
            "steps": steps, "ms_per_step": dt * 1e3 / steps,
            "roofline_frac": 8.0 * cells * steps / dt / (HBM_PEAK_GBS * 1e9),
            "traffic_per_step": (load_traffic("batch_ragged_64",
                                              "bytes_per_step")
                                 if scheme == "kick_drift" and wake_method == 2
                                 else None),
            "launch": launch_label(steps, no_graph, batch=True),
            "tiles_chosen": {str(k): v for k, v in chosen.items()}}


def production_throughput(steps, log, tiles=None, no_graph=False):
    """The large production workload: 16 independent 1024^2 meshes
    (second-order momentum compaction, parallel-plates CSR plus resistive
    wall, five-point Fokker-Planck, 4000 steps per period; currents
    0.8-1.5x threshold, gaps 24-56 mm) in one batch after the autotuner:
    whole-batch cell updates per second."""
    import numpy as np
    from inovesa_amd import Batch
    from inovesa_amd.workloads import (impedances, initial_mesh,
                                       production_config)
    descs, inits, zws, zcs = [], [], [], []
    for k, (f, g) in enumerate(zip(np.geomspace(0.8, 1.5, 16),
                                   np.linspace(0.024, 0.056, 16))):
        desc, d, rw = production_config(gap=float(g), current_factor=float(f))
        zw, zc = impedances(d, resistive_wall=rw)
        descs.append(dict(desc, no_graph=int(no_graph)))
        inits.append(initial_mesh(desc["n"], seed=100 + k))
        zws.append(zw)
        zcs.append(zc)
    b = Batch(descs, inits, zws, zcs)
    if tiles:
        T, thr, persistent, strip = tiles
        b.retile_size(1024, T, thr, persistent=persistent, strip=strip)
        table = [dict(n=1024, T=T, threads=thr, persistent=persistent,
                      strip=strip, us_per_step=0.0, chosen=True)]
    else:
        table = b.autotune(steps=6)
    cells = b.info()["cells"]
    b.record(steps + 64)
    b.run(64)
    b.prepare(steps)
    b.dev.sync()
    t0 = time.perf_counter()
    b.run(steps)
    b.dev.sync()
    dt = time.perf_counter() - t0
    chosen = [(r["T"], r["threads"], form_label(r),
               round(r["us_per_step"], 2)) for r in table if r["chosen"]]
    log(f"production: 16 x 1024^2, {steps} steps in {dt:.2f} s")
    return {"value": cells * steps / dt, "unit": "cell-updates/s",
            "instances": 16, "n": 1024, "cells_per_step": int(cells),
            "steps": steps, "ms_per_step": dt * 1e3 / steps,
            "roofline_frac": 8.0 * cells * steps / dt / (HBM_PEAK_GBS * 1e9),
            "traffic": load_traffic("batch_production_16x1024"),
            "launch": launch_label(steps, no_graph, batch=True),
            "tiles_chosen": chosen[0] if chosen else None}


def scan_ablation(targets, log):
    """Threshold scans of `targets` (gap, damping) pairs on the uneven mix
    of mesh sizes of the batched scan (each target on one of 128^2-512^2;
    only sizes on which the explicit Fokker-Planck update of its damping
    time stays bounded): the plain dense scan and the scan steered by the
    network trained on runs of that mix."""
    from collections import Counter
    from inovesa_amd.pipeline import compare, scan_targets
    from inovesa_amd.scan import ScanSetup
    from inovesa_amd.surrogate import MIX_MODEL, MIX_SIZES, Surrogate
    setup = ScanSetup(sizes=MIX_SIZES)
    tg = scan_targets(targets)
    res = compare(Surrogate.load(MIX_MODEL), tg, setup)
    log(f"scan: dense {res['dense_gpu_s']:.2f} s, surrogate "
        f"{res['surrogate_gpu_s']:.2f} s, ratio {res['gpu_time_ratio']:.3f}")
    out = {k: v for k, v in res.items() if k != "rows"}
    out["workload"] = ("kick-drift maps, PP-CSR, 500 steps/Ts x 30 periods; "
                       "targets on an uneven mix of mesh sizes")
    out["targets_per_size"] = {str(k): v for k, v in sorted(
        Counter(setup.size_of(g, d) for g, d in tg).items())}
    return out


def copy_calibration(mib):
    """A plain device copy of `mib` MiB (read + write), for scaling the HBM
    counters of a profiler run to bytes."""
    import torch
    a = torch.ones(mib * 262144, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(5):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_mib": mib, "bytes_read": mib * 1048576,
                      "bytes_written": mib * 1048576}), flush=True)


def load_traffic(name, key="bytes_per_launch"):
    """HBM bytes from the counter summary profiles/traffic_<name>.json
    (FETCH_SIZE and WRITE_SIZE scaled by a copy of known size)."""
    p = os.path.join(ROOT, "profiles", f"traffic_{name}.json")
    if os.path.exists(p):
        with open(p) as f:
            return json.load(f).get(key)
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=40000)
    ap.add_argument("--warmup", type=int, default=2000)
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-scan", action="store_true")
    ap.add_argument("--no-graph", action="store_true",
                    help="launch every step from the host instead of "
                         "replaying a hipGraph (for tools that do not "
                         "follow graph launches)")
    ap.add_argument("--projection", choices=("wavefront", "ordered"),
                    default="ordered",
                    help="x projection: wavefront reduction, or the "
                         "reference's sequential order (bit-exact)")
    ap.add_argument("--batch-steps", type=int, default=2000)
    ap.add_argument("--scan-targets", type=int, default=128)
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a device copy of MIB MiB and exit")
    ap.add_argument("--only", choices=("batched", "batched_rotation",
                                       "batched_fft_wake", "production"),
                    help="run only this batch line, with the tiling the "
                         "autotuner last chose (for profiler runs), and "
                         "print its JSON object")
    args = ap.parse_args()
    log = lambda m: print(m, file=sys.stderr, flush=True)
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.only:
        if args.only == "production":
            r = production_throughput(args.batch_steps, log,
                                      tiles=PROFILE_TILES["production"],
                                      no_graph=args.no_graph)
        else:
            fft = args.only == "batched_fft_wake"
            r = batched_throughput(args.batch_steps, log,
                                   scheme="rotation" if args.only ==
                                   "batched_rotation" else "kick_drift",
                                   tiles=PROFILE_TILES["batched" if fft
                                                       else args.only],
                                   no_graph=args.no_graph,
                                   wake_method=1 if fft else 2)
        print(json.dumps({args.only: r}), flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py: the step runs on one GPU; --gpus {args.gpus} is "
              "not supported (independent instances shard across GPUs by "
              "process: one process per GPU)", file=sys.stderr)
        sys.exit(2)

    from inovesa_amd import Step
    from inovesa_amd.workloads import impedances, initial_mesh, kick_drift_config

    n = args.n
    desc, d = kick_drift_config(n=n, steps_per_ts=2000, gap=0.032,
                                current_factor=1.5)
    zw, zc = impedances(d)
    init = initial_mesh(n, seed=1)
    # x projection in the reference's sequential order (bit-exact; the
    # wake amplifies a last-bit change of the profile some 300 times, so
    # only this order keeps the per-step wake within 1e-5) or as a
    # wavefront reduction
    fast = 1 if args.projection == "wavefront" else 0
    st = Step(dict(desc, fast_projection=fast, no_graph=int(args.no_graph)),
              init, zw, zc)
    st.record(args.warmup + args.steps, profiles=True)
    # the timed run's graphs are captured first, so that the warm-up steps
    # run right before the timed ones
    st.prepare(args.steps)
    st.prepare(args.warmup)
    st.run(args.warmup)
    st.dev.sync()
    t0 = time.perf_counter()
    st.run(args.steps)
    st.dev.sync()
    dt = time.perf_counter() - t0
    cnt_csr, _ = st.history()
    assert len(cnt_csr) == args.warmup + args.steps
    value = n * n * args.steps / dt

    cfg = st.config()
    _, ms_step_kernel, _ = st.timing(400)
    bytes_per_launch = 8.0 * n * n
    achieved = bytes_per_launch / (ms_step_kernel * 1e-3) / 1e9
    roofline = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                "traffic": load_traffic(f"step_kd_{n}"),
                "kernel": "step_kernel (direct wake + CSR power, wake kick, "
                          "RF kick, drift, Fokker-Planck, projection: the "
                          "whole step)",
                "kernel_ms": ms_step_kernel,
                "bytes_per_launch": bytes_per_launch}

    # the same workload with the other projection order
    other = Step(dict(desc, fast_projection=1 - fast,
                      no_graph=int(args.no_graph)), init, zw, zc)
    other.run(640)
    other.prepare(min(args.steps, 12800))
    other.dev.sync()
    ko = min(args.steps, 12800)
    t0 = time.perf_counter()
    other.run(ko)
    other.dev.sync()
    dto = time.perf_counter() - t0
    other_line = {"projection": "ordered" if fast else "wavefront",
                  "value": n * n * ko / dto, "steps": ko,
                  "ms_per_step": dto * 1e3 / ko}
    del other

    # the same workload with the FFT wake (field kernel + step kernel)
    fft = Step(dict(desc, wake_method=1, fast_projection=fast,
                    no_graph=int(args.no_graph)), init, zw, zc)
    fft.run(640)
    fft.prepare(min(args.steps, 6400))
    fft.dev.sync()
    kf = min(args.steps, 6400)
    t0 = time.perf_counter()
    fft.run(kf)
    fft.dev.sync()
    dtf = time.perf_counter() - t0
    ms_field, ms_maps_fft, _ = fft.timing(200)
    fft_line = {"value": n * n * kf / dtf, "steps": kf,
                "ms_per_step": dtf * 1e3 / kf, "field_kernel_ms": ms_field,
                "step_kernel_ms": ms_maps_fft}

    # the per-map path: one kernel per map and the FFT field, as an
    # unchanged reference loop calls it (SourceMap::apply per map)
    pm = Step(dict(desc, fused=0, no_graph=int(args.no_graph)), init, zw, zc)
    pm.run(64)
    kp = min(args.steps, 2000)
    pm.prepare(kp)
    pm.dev.sync()
    t0 = time.perf_counter()
    pm.run(kp)
    pm.dev.sync()
    dtp = time.perf_counter() - t0
    per_map_line = {"value": n * n * kp / dtp, "steps": kp,
                    "ms_per_step": dtp * 1e3 / kp,
                    "roofline_frac": 8.0 * n * n * kp / dtp
                    / (HBM_PEAK_GBS * 1e9),
                    "kernels": "field (FFT wake + CSR), wake kick, RF kick, "
                               "drift, Fokker-Planck, x projection"}
    del pm

    out = {"metric": METRIC, "value": value, "unit": "cell-updates/s",
           "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": dt * 1e3 / args.steps, "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f32",
           "data": "synthetic (seeded Gaussian + 1e-4 uniform noise)",
           "config": {"workload": f"plain VFP {n}x{n} kick-drift, PP-CSR "
                                  "gap 32 mm, nmax "
                                  f"{zw.n}, 1.5x scaling-law threshold, "
                                  "2000 steps/Ts, profile + CSR power "
                                  "recorded every step",
                      "n": n, "nmax": zw.n, "steps_per_ts": 2000,
                      "Ib_A": d["Ib"], "wake": "direct (FP64 Toeplitz)",
                      "x_projection": ("wavefront reduction" if fast else
                                       "sequential, reference order"),
                      "launch": launch_label(args.steps, args.no_graph),
                      "tile_rows": cfg["tile_rows"],
                      "threads": cfg["threads"]},
           "roofline": roofline, "fft_wake": fft_line,
           "other_projection": other_line, "per_map": per_map_line}
    if not args.no_batch:
        out["batched"] = batched_throughput(args.batch_steps, log)
        out["batched_rotation"] = batched_throughput(args.batch_steps, log,
                                                     scheme="rotation")
        # the same 64 instances with the FFT wake, on the direct line's
        # tiling (the batched radix FFT: one field-kernel workgroup per
        # instance and step)
        out["batched_fft_wake"] = batched_throughput(
            args.batch_steps, log, tiles=PROFILE_TILES["batched"],
            wake_method=1)
        out["production"] = production_throughput(1000, log)
    if not args.no_scan:
        out["scan"] = scan_ablation(args.scan_targets, log)
    if not args.no_cpu:
        cb = cpu_baseline(desc, zw, zc, init, args.cpu_budget)
        cb["ratio_gpu_over_cpu"] = value / cb["value"]
        out["cpu_baseline"] = cb
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
