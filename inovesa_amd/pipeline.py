"""Train the threshold surrogate on the batched solver and compare the
surrogate-steered scan with the plain dense scan.

    python -m inovesa_amd.pipeline --runs 8192 --focus 0.6 --out DIR
    python -m inovesa_amd.pipeline --sizes 128,256,384,512 ...   (a mix of
        mesh sizes: each (gap, damping) target on one of them)

Steps (all on the GPU through the library):
  1. `runs` VFP runs at random (current, gap, damping) points, batched,
     a fraction `focus` of them at currents near the thresholds;
     per run the 64-bin log CSR spectrum and the detector's verdict;
  2. the network trained on them with the library's forward, backward and
     Adam kernels;
  3. thresholds of `targets` (gap, damping) pairs by the dense scan and by
     the surrogate-steered scan, with the GPU time and VFP runs of each.
Writes DIR/surrogate.npz (weights and normalisation) and DIR/scan.json.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

from .scan import ScanSetup, dense_scan, surrogate_scan
from .surrogate import Surrogate, generate


def scan_targets(count, seed=11):
    rng = np.random.Generator(np.random.Philox(seed))
    g = rng.uniform(0.02, 0.06, count)
    d = np.exp(rng.uniform(np.log(1200.0), np.log(4500.0), count))
    return [(float(a), float(b)) for a, b in zip(g, d)]


def compare(model, targets, setup, rel_tol=0.02):
    """Dense and surrogate scans of the same targets: a JSON-able dict.
    A scan's time covers building its runs (impedance tables, batches),
    tuning the tiling of every batch-size class it meets once, and the
    runs themselves (*_build_s and *_tune_s are the first two parts;
    run_time_ratio compares the runs alone)."""
    t0 = time.perf_counter()
    dres, drun = dense_scan(targets, setup, rel_tol=rel_tol)
    t1 = time.perf_counter()
    sres, srun = surrogate_scan(model, targets, setup, rel_tol=rel_tol)
    t2 = time.perf_counter()
    rows = []
    for a, b in zip(dres, sres):
        rows.append(dict(gap=a.gap, damping=a.damping,
                         dense=a.threshold, dense_width=a.width,
                         predicted=b.predicted,
                         prediction_error=b.predicted / a.threshold - 1.0,
                         surrogate=b.threshold, surrogate_width=b.width,
                         confirmed=b.confirmed, dense_runs=a.runs,
                         surrogate_runs=b.runs,
                         rel_diff=abs(b.threshold / a.threshold - 1.0)))
    return dict(
        targets=len(targets), rel_tol=rel_tol,
        dense_gpu_s=drun.gpu_seconds, surrogate_gpu_s=srun.gpu_seconds,
        dense_build_s=drun.build_seconds,
        surrogate_build_s=srun.build_seconds,
        dense_tune_s=drun.tune_seconds, surrogate_tune_s=srun.tune_seconds,
        run_time_ratio=(srun.gpu_seconds - srun.tune_seconds
                        - srun.build_seconds)
        / max(drun.gpu_seconds - drun.tune_seconds - drun.build_seconds,
              1e-12),
        dense_rounds=drun.rounds, surrogate_rounds=srun.rounds,
        dense_wall_s=t1 - t0, surrogate_wall_s=t2 - t1,
        dense_runs=drun.runs, surrogate_runs=srun.runs,
        gpu_time_ratio=srun.gpu_seconds / max(drun.gpu_seconds, 1e-12),
        max_rel_diff=max(r["rel_diff"] for r in rows),
        max_prediction_error=max(abs(r["prediction_error"]) for r in rows),
        confirmed=sum(r["confirmed"] for r in rows), rows=rows)


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--runs", type=int, default=8192)
    ap.add_argument("--focus", type=float, default=0.6,
                    help="fraction of the runs near the threshold currents")
    ap.add_argument("--chunk", type=int, default=1024)
    ap.add_argument("--epochs", type=int, default=300)
    ap.add_argument("--targets", type=int, default=8)
    ap.add_argument("--out", default="surrogate_run",
                    help="directory for the training data, model and scan "
                         "summary")
    ap.add_argument("--model", default=None,
                    help="load a trained model instead of training")
    ap.add_argument("--sizes", default=None,
                    help="comma-separated mesh sizes: each (gap, damping) "
                         "target runs on one of them (default: every run "
                         "on 128^2)")
    ap.add_argument("--data", default=None,
                    help="train on the runs of a training_data.npz written "
                         "by an earlier call (labels recomputed from their "
                         "fluctuations with this setup's detector)")
    args = ap.parse_args(argv)
    os.makedirs(args.out, exist_ok=True)
    setup = ScanSetup(sizes=tuple(int(v) for v in args.sizes.split(","))
                      if args.sizes else None)
    log = lambda m: print(m, flush=True)
    summary = dict(setup=setup.__dict__)
    if args.model:
        model = Surrogate.load(args.model)
    else:
        if args.data:
            z = np.load(args.data)
            pts, spec, fl = z["points"], z["spectra"], z["fluct"]
            lab = ((~np.isfinite(fl)) | (fl > setup.tol)).astype(np.float32)
            gpu_s = 0.0
        else:
            pts, spec, lab, fl, gpu_s = generate(args.runs, setup,
                                                 chunk=args.chunk, log=log,
                                                 focus=args.focus)
        np.savez_compressed(os.path.join(args.out, "training_data.npz"),
                            points=pts, spectra=spec, labels=lab, fluct=fl)
        log(f"data: {len(lab)} runs, {int(lab.sum())} unstable, "
            f"{gpu_s:.1f} s GPU")
        # hold out a tenth to measure the fit
        n_val = max(1, len(lab) // 10)
        model = Surrogate(seed=1)
        t0 = time.perf_counter()
        losses = model.fit(pts[n_val:], spec[n_val:], lab[n_val:],
                           epochs=args.epochs)
        train_s = time.perf_counter() - t0
        vs, vl = model.predict(pts[:n_val, 0], pts[:n_val, 1], pts[:n_val, 2])
        acc = float(((vl > 0) == (lab[:n_val] > 0.5)).mean())
        rmse = float(np.sqrt(((vs - spec[:n_val]) ** 2).mean()))
        log(f"trained {args.epochs} epochs in {train_s:.1f} s: loss "
            f"{losses[0]:.4f} -> {losses[-1]:.4f}; held-out verdict accuracy "
            f"{acc:.3f}, log-spectrum RMSE {rmse:.3f}")
        model.save(os.path.join(args.out, "surrogate.npz"))
        summary.update(runs=len(lab), focus=args.focus, unstable=int(lab.sum()),
                       data_gpu_s=gpu_s, train_s=train_s,
                       epochs=args.epochs, loss_first=float(losses[0]),
                       loss_last=float(losses[-1]),
                       epoch_losses=[float(v) for v in losses],
                       heldout_accuracy=acc, heldout_logspec_rmse=rmse)
    res = compare(model, scan_targets(args.targets), setup)
    log(f"scan of {res['targets']} targets: dense {res['dense_runs']} runs "
        f"{res['dense_gpu_s']:.2f} s GPU; surrogate {res['surrogate_runs']} "
        f"runs {res['surrogate_gpu_s']:.2f} s GPU (ratio "
        f"{res['gpu_time_ratio']:.3f}); max threshold difference "
        f"{100 * res['max_rel_diff']:.2f}%, {res['confirmed']} confirmed")
    summary["scan"] = res
    with open(os.path.join(args.out, "scan.json"), "w") as f:
        json.dump(summary, f, indent=1)
    return summary


if __name__ == "__main__":
    main()
