"""Threshold scans on the GPU: the plain dense scan against the scaling law,
and the ablation of the trained surrogate (scan with the network against
the scan without it, on the same targets)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_dense_scan_against_scaling_law():
    """Ith = Inorm (0.5 + 0.34 Pi) (src/main.cpp:383-390) is a fit to
    VFP thresholds; the dense scan lands within a factor 0.8-1.5 of it and
    brackets the threshold to 2%."""
    from inovesa_amd.scan import ScanSetup, dense_scan
    res, runner = dense_scan([(0.032, 2000.0), (0.05, 3000.0)], ScanSetup())
    for r in res:
        print(f"\ngap {r.gap} damping {r.damping}: threshold "
              f"{r.threshold:.4f} x scaling law (bracket {r.lo:.4f}-"
              f"{r.hi:.4f}), {r.runs} runs")
        assert 0.8 < r.threshold < 1.5
        assert r.width <= 0.02 + 1e-9
    assert runner.runs == 64


def test_surrogate_ablation():
    """The trained network (inovesa_amd/data/surrogate.npz) against the
    method without it: the same thresholds within 2%, a fraction of the
    VFP runs and of the GPU time."""
    from inovesa_amd.pipeline import compare, scan_targets
    from inovesa_amd.scan import ScanSetup
    from inovesa_amd.surrogate import Surrogate
    model = Surrogate.load()
    res = compare(model, scan_targets(24, seed=23), ScanSetup())
    print(f"\n24 targets: dense {res['dense_runs']} runs "
          f"{res['dense_gpu_s']:.2f} s, surrogate {res['surrogate_runs']} runs "
          f"{res['surrogate_gpu_s']:.2f} s (ratio {res['gpu_time_ratio']:.3f}), "
          f"max threshold difference {100 * res['max_rel_diff']:.2f}%, "
          f"{res['confirmed']} confirmed")
    assert res["max_rel_diff"] <= 0.02
    assert res["surrogate_runs"] <= res["dense_runs"] / 5
    assert res["confirmed"] >= 22
    # a smaller batch than the 48-target benchmark: the per-step launch
    # floor weighs more, hence the margin over 1/5
    assert res["gpu_time_ratio"] <= 0.25


def test_surrogate_outputs():
    """The network's spectrum and verdict on its own inputs: finite,
    unstable well above and stable well below the scaling law."""
    from inovesa_amd.surrogate import Surrogate
    model = Surrogate.load()
    spec, logit = model.predict(np.array([0.4, 2.5]), 0.04, 2000.0)
    assert np.all(np.isfinite(spec)) and spec.shape == (2, 64)
    assert logit[0] < 0 < logit[1]
    f = model.predict_threshold(0.04, 2000.0)
    assert 0.8 < f < 1.5


def test_surrogate_ablation_on_mesh_mix():
    """The same ablation on the uneven mix of mesh sizes of the batched
    scan (each target on one of 128^2-512^2 where the explicit
    Fokker-Planck update of its damping time stays bounded), with the
    network trained on runs of that mix (inovesa_amd/data/
    surrogate_mix.npz)."""
    from collections import Counter
    from inovesa_amd.pipeline import compare, scan_targets
    from inovesa_amd.scan import ScanSetup
    from inovesa_amd.surrogate import MIX_MODEL, MIX_SIZES, Surrogate
    setup = ScanSetup(sizes=MIX_SIZES)
    targets = scan_targets(24, seed=29)
    sizes = Counter(setup.size_of(g, d) for g, d in targets)
    assert len(sizes) >= 2, sizes
    res = compare(Surrogate.load(MIX_MODEL), targets, setup)
    print(f"\n24 targets on {dict(sizes)}: dense {res['dense_runs']} runs "
          f"{res['dense_gpu_s']:.2f} s, surrogate {res['surrogate_runs']} runs "
          f"{res['surrogate_gpu_s']:.2f} s (ratio {res['gpu_time_ratio']:.3f}), "
          f"max threshold difference {100 * res['max_rel_diff']:.2f}%, "
          f"{res['confirmed']} confirmed")
    assert res["max_rel_diff"] <= 0.02
    assert res["surrogate_runs"] <= res["dense_runs"] / 5
    assert res["confirmed"] >= 22
    assert res["gpu_time_ratio"] <= 0.25
