"""Batched ragged step (many independent meshes of uneven sizes, one launch
per step) and its autotuner.  Pinned by the single-mesh step: every
instance of a batch must equal the same instance run alone, bit for bit,
whatever tiling the autotuner picks."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def V():
    import inovesa_amd as V
    return V


def _instances(V, sizes, per_size, kick_drift=False, seed=3,
               scheme="kick_drift"):
    from inovesa_amd.workloads import (batched_scan_config, impedances,
                                       kick_drift_config, initial_mesh)
    if scheme == "mixed":
        cfg = (batched_scan_config(sizes=sizes, per_size=per_size, seed=seed,
                                   scheme="rotation")
               + batched_scan_config(sizes=sizes, per_size=per_size,
                                     seed=seed + 1, scheme="kick_drift"))
    else:
        cfg = batched_scan_config(sizes=sizes, per_size=per_size, seed=seed,
                                  scheme=scheme)
    out = []
    for desc, d, init in cfg:
        zw, zc = impedances(d)
        out.append((desc, zw, zc, init))
    if kick_drift:
        desc, d = kick_drift_config(n=256, steps_per_ts=500)
        zw, zc = impedances(d)
        out.append((desc, zw, zc, initial_mesh(256, seed=11)))
    return out


def _batch(V, inst):
    return V.Batch([i[0] for i in inst], [i[3] for i in inst],
                   [i[1] for i in inst], [i[2] for i in inst])


def _single(V, i):
    desc, zw, zc, init = i
    return V.Step(dict(desc, wake_method=2), init, zw, zc)


@pytest.mark.parametrize("scheme", ["kick_drift", "rotation"])
def test_batch_equals_single_steps(V, scheme):
    """Each scheme alone, plus one instance of the other scheme's maps (a
    256^2 kick-drift mesh beside the rotation instances)."""
    inst = _instances(V, (128, 256, 384), 2, scheme == "rotation",
                      scheme=scheme)
    b = _batch(V, inst)
    info = b.info()
    assert info["instances"] == len(inst)
    # one launch per kernel variant, workgroup size and LDS size class:
    # at least one per variant, at most one per (variant, mesh size)
    variants = 2 if scheme == "rotation" else 1
    assert variants <= info["launch_groups"] <= 3 + (variants - 1)
    steps = 70                       # one graph replay and 6 single steps
    b.record(steps + 8)
    b.run(steps)
    for k, i in enumerate(inst):
        s = _single(V, i)
        s.record(steps + 8)
        s.run(steps)
        for x, y in zip(b.mesh(k), s.mesh()):
            assert np.array_equal(x, y), k
        cb, _ = b.history(k)
        cs, _ = s.history()
        assert len(cb) == steps and np.array_equal(cb, cs), k


@pytest.mark.parametrize("tile,threads", [(1, 128), (1, 256), (1, 512)])
def test_persistent_launch_equals_single_steps(V, tile, threads):
    """The persistent form (as many workgroups as stay resident, each
    walking several tiles of the ragged map) gives every instance's result
    bit for bit."""
    inst = _instances(V, (128, 256, 384, 512), 4)
    b = _batch(V, inst)
    for k in range(len(inst)):
        b.retile(k, tile, threads, persistent=True)
        assert b.config(k)["persistent"] == 1
    shapes = b.launches()
    print("\nlaunches (threads, tiles, workgroups, LDS):",
          [tuple(x.values()) for x in shapes])
    # at least one launch walks several tiles per workgroup
    assert any(x["workgroups"] < x["tiles"] for x in shapes)
    assert all(x["workgroups"] % 8 == 0 for x in shapes)
    steps = 66                       # one graph replay and 2 single steps
    b.record(steps + 4)
    b.run(steps)
    for k in (0, 5, 9, len(inst) - 1):
        s = _single(V, inst[k])
        s.record(steps + 4)
        s.run(steps)
        for x, y in zip(b.mesh(k), s.mesh()):
            assert np.array_equal(x, y), k
        cb, _ = b.history(k)
        cs, _ = s.history()
        assert np.array_equal(cb, cs), k


def test_production_batch_equals_single_steps(V):
    """Four instances of the 1024^2 production workload (currents and gaps
    apart) in one batch, autotuned: each equals its run alone, bit for
    bit."""
    from inovesa_amd.workloads import impedances, initial_mesh, production_config
    inst = []
    for k, (f, g) in enumerate([(0.8, 0.024), (1.0, 0.032), (1.2, 0.04),
                                (1.5, 0.056)]):
        desc, d, rw = production_config(gap=g, current_factor=f)
        zw, zc = impedances(d, resistive_wall=rw)
        inst.append((desc, zw, zc, initial_mesh(desc["n"], seed=40 + k)))
    b = _batch(V, inst)
    b.autotune(steps=4)
    steps = 66
    b.record(steps + 4)
    b.run(steps)
    for k in (0, 3):
        s = _single(V, inst[k])
        s.record(steps + 4)
        s.run(steps)
        for x, y in zip(b.mesh(k), s.mesh()):
            assert np.array_equal(x, y), k
        assert np.array_equal(b.history(k)[0], s.history()[0]), k


@pytest.mark.parametrize("tile,strip,threads", [(4, 8, 512), (2, 8, 256),
                                                (8, 8, 512), (4, 4, 128),
                                                (2, 32, 256)])
def test_strip_form_equals_single_steps(V, tile, strip, threads):
    """The strip form (a workgroup advances `strip` consecutive tiles of one
    instance, kicking every row once into a ring of kicked rows, and sums
    the strip's x projection one row per lane) gives every instance's
    result bit for bit: mesh, profile, wake offsets, CSR history.  196 rows
    end in a partial strip and a partial tile."""
    inst = _instances(V, (128, 196, 384, 512), 2)
    b = _batch(V, inst)
    for k in range(len(inst)):
        b.retile(k, tile, threads, strip=strip)
        cfg = b.config(k)
        assert (cfg["tile_rows"], cfg["threads"], cfg["strip"]) == (
            tile, threads, strip)
    steps = 70                       # one graph replay and 6 single steps
    b.record(steps + 8)
    b.run(steps)
    for k, i in enumerate(inst):
        s = _single(V, i)
        s.record(steps + 8)
        s.run(steps)
        for x, y in zip(b.mesh(k), s.mesh()):
            assert np.array_equal(x, y), (k, i[0]["n"])
        cb, pb = b.history(k)
        cs, ps = s.history()
        assert np.array_equal(cb, cs), k


def test_strip_form_beside_rotation_tiles(V):
    """Kick-drift instances in strips and rotation instances in tiles in
    one batch step (two launches): each equals its run alone."""
    inst = _instances(V, (128, 256), 2, scheme="mixed")
    b = _batch(V, inst)
    for k, i in enumerate(inst):
        if i[0]["scheme"] == 0:
            b.retile(k, 4, 256, strip=8)
    with pytest.raises(V.VfpsInvalid):
        b.retile([k for k, i in enumerate(inst) if i[0]["scheme"] == 1][0],
                 4, 256, strip=8)
    steps = 66
    b.record(steps + 4)
    b.run(steps)
    for k, i in enumerate(inst):
        s = _single(V, i)
        s.record(steps + 4)
        s.run(steps)
        for x, y in zip(b.mesh(k), s.mesh()):
            assert np.array_equal(x, y), k
        assert np.array_equal(b.history(k)[0], s.history()[0]), k


def test_strip_form_rejects_and_keeps_tiling(V):
    inst = _instances(V, (128,), 2)
    b = _batch(V, inst)
    before = b.config(0)
    for T, K in ((16, 2), (4, 32)):   # tiles above 8 rows, strips above 64
        with pytest.raises(V.VfpsInvalid):
            b.retile(0, T, 256, strip=K)
        assert b.config(0) == before
    b.retile_size(128, 4, 256, strip=8)
    assert all(b.config(k)["strip"] == 8 for k in range(len(inst)))
    b.retile_size(128, 2, 256)
    assert all(b.config(k)["strip"] == 0 for k in range(len(inst)))
    b.run(3)


def test_production_strip_form_equals_single_steps(V):
    """Two instances of the 1024^2 production workload (five-point FP
    stencil, second-order drift) in strips of 8 tiles of 4 rows."""
    from inovesa_amd.workloads import impedances, initial_mesh, production_config
    inst = []
    for k, (f, g) in enumerate([(0.8, 0.024), (1.5, 0.056)]):
        desc, d, rw = production_config(gap=g, current_factor=f)
        zw, zc = impedances(d, resistive_wall=rw)
        inst.append((desc, zw, zc, initial_mesh(desc["n"], seed=40 + k)))
    b = _batch(V, inst)
    b.retile_size(1024, 4, 512, strip=8)
    steps = 66
    b.record(steps + 4)
    b.run(steps)
    for k in range(len(inst)):
        s = _single(V, inst[k])
        s.record(steps + 4)
        s.run(steps)
        for x, y in zip(b.mesh(k), s.mesh()):
            assert np.array_equal(x, y), k
        assert np.array_equal(b.history(k)[0], s.history()[0]), k


def test_merged_launch_equals_split(V):
    """One launch for all size classes (every tile with the largest LDS)
    against one per LDS class: the same results bit for bit."""
    inst = _instances(V, (128, 256, 384), 2)
    a = _batch(V, inst)
    b = _batch(V, inst)
    b.set_merge(True)
    assert len(b.launches()) < len(a.launches())
    a.run(66)
    b.run(66)
    for k in range(len(inst)):
        for x, y in zip(a.mesh(k), b.mesh(k)):
            assert np.array_equal(x, y), k


def test_retile_rejects_and_keeps_tiling(V):
    inst = _instances(V, (128,), 2)
    b = _batch(V, inst)
    before = b.config(0)
    with pytest.raises(V.VfpsInvalid):
        b.retile(0, 1, 100)
    assert b.config(0) == before
    b.retile_size(128, 2, 256, persistent=True)
    for k in range(len(inst)):
        cfg = b.config(k)
        assert (cfg["tile_rows"], cfg["threads"], cfg["persistent"]) == (2, 256, 1)
    b.run(3)


def test_autotune_keeps_results_and_reports_table(V):
    inst = _instances(V, (128, 256, 384, 512), 1)
    a = _batch(V, inst)
    b = _batch(V, inst)
    a.run(10)
    b.run(10)
    table = a.autotune(steps=6)
    sizes = sorted({r["n"] for r in table})
    assert sizes == [128, 256, 384, 512]
    for n in sizes:
        rows = [r for r in table if r["n"] == n]
        assert len(rows) >= 3
        chosen = [r for r in rows if r["chosen"]]
        assert len(chosen) == 1
        assert all(r["persistent"] in (False, True) for r in rows)
        retimed = [r for r in rows if r["retimed"]]
        assert 1 <= len(retimed) <= 3
        fastest = min(retimed, key=lambda r: r["us_per_step"])
        if chosen[0] is not fastest:
            # a persistent choice that lost on the whole batch gives way to
            # the size's fastest per-tile tiling
            assert fastest["persistent"] and not chosen[0]["persistent"]
        k = [j for j, i in enumerate(inst) if i[0]["n"] == n][0]
        cfg = a.config(k)
        assert (cfg["tile_rows"], cfg["threads"], cfg["persistent"]) == (
            chosen[0]["T"], chosen[0]["threads"], int(chosen[0]["persistent"]))
    print("\nautotune table (n, T, threads, us/step, chosen):")
    for r in table:
        print(f"  {r['n']:4d} {r['T']:2d} {r['threads']:4d} "
              f"{r['us_per_step']:8.2f} {'*' if r['chosen'] else ''}")
    # tuning restored the state: both batches continue identically
    a.run(66)
    b.run(66)
    for k in range(len(inst)):
        for x, y in zip(a.mesh(k), b.mesh(k)):
            assert np.array_equal(x, y), k


def test_batch_timing_and_rejects(V):
    inst = _instances(V, (128,), 2)
    b = _batch(V, inst)
    ms = b.timing(8)
    assert 0 < ms < 10
    with pytest.raises((IndexError, V.VfpsInvalid)):
        b.mesh(5)
    bad = dict(inst[0][0], scheme=7)
    with pytest.raises(V.VfpsInvalid):
        V.Batch([bad], [inst[0][3]], [inst[0][1]], [inst[0][2]])


@pytest.mark.parametrize("persistent", [False, True])
def test_rotation_and_kick_drift_batch_equals_single_steps(V, persistent):
    """The batched scan workload on the rotation maps (cubic clamped
    RotationMap, src/SM/RotationMap.cpp:97-176) mixed with kick-drift
    instances over 128^2-512^2, so that both kernel variants run in one
    step, per-tile and persistent: every instance equals its run alone,
    bit for bit."""
    inst = _instances(V, (128, 256, 384, 512), 4, scheme="mixed")
    assert {i[0]["scheme"] for i in inst} == {0, 1}
    b = _batch(V, inst)
    if persistent:
        for k in range(len(inst)):
            b.retile(k, 1, 256, persistent=True)
        assert any(x["workgroups"] < x["tiles"] for x in b.launches())
    groups = b.info()["launch_groups"]
    assert groups >= 2
    steps = 66
    b.record(steps + 4)
    b.run(steps)
    # every mesh size of both schemes (sizes are interleaved per scheme)
    for k in list(range(4)) + list(range(16, 20)) + [15, 31]:
        i = inst[k]
        s = _single(V, i)
        s.record(steps + 4)
        s.run(steps)
        for x, y in zip(b.mesh(k), s.mesh()):
            assert np.array_equal(x, y), (k, i[0]["scheme"], i[0]["n"])
        assert np.array_equal(b.history(k)[0], s.history()[0]), k


def _mixed_wake_batch(V, sizes, per_size, fft_every=2):
    """Instances of the batched scan, every fft_every-th with the FFT wake,
    the rest with the direct wake."""
    inst = _instances(V, sizes, per_size)
    descs = [dict(i[0], wake_method=1 if k % fft_every == 0 else 2)
             for k, i in enumerate(inst)]
    b = V.Batch(descs, [i[3] for i in inst], [i[1] for i in inst],
                [i[2] for i in inst])
    return inst, descs, b


def _equals_single_runs(V, b, inst, descs, steps, ks):
    for k in ks:
        i, d = inst[k], descs[k]
        s = V.Step(d, i[3], i[1], i[2])
        s.record(steps + 8, profiles=True)
        s.run(steps)
        for x, y in zip(b.mesh(k), s.mesh()):
            assert np.array_equal(x, y), k
        cb, pb = b.history(k)
        cs, ps = s.history()
        assert len(cb) == steps, k
        assert np.array_equal(cb, cs) and np.array_equal(pb, ps), k


def test_fft_wake_batch_equals_single_steps(V):
    """FFT-wake instances (one field-kernel launch over all of them, a
    workgroup each with its own transform length, ahead of their step
    kernels) beside direct-wake instances in one batch, over 128^2-512^2:
    each equals its single-mesh run with the same wake form bit for bit,
    mesh, profile, wake, profile history and CSR power history."""
    inst, descs, b = _mixed_wake_batch(V, (128, 256, 384, 512), 2)
    assert {b.config(k)["wake_method"] for k in range(len(inst))} == {1, 2}
    # FFT and direct instances never share a step-kernel launch
    assert b.info()["launch_groups"] >= 2
    steps = 70                       # one graph replay and 6 single steps
    b.record(steps + 8, profiles=True)
    b.run(steps)
    _equals_single_runs(V, b, inst, descs, steps, range(len(inst)))


def test_fft_wake_batch_autotune_and_forms(V):
    """The autotuner times FFT-wake size classes in the per-tile form only
    (the resident-grid and strip forms are direct-wake forms), a request
    for them is refused with the tiling kept, and the tuned batch still
    equals the single runs bit for bit."""
    inst, descs, b = _mixed_wake_batch(V, (128, 256), 2, fft_every=1)
    table = b.autotune(steps=4)
    assert table and not any(r["persistent"] or r["strip"] for r in table)
    before = b.config(0)
    with pytest.raises(V.VfpsInvalid):
        b.retile(0, 2, 256, persistent=True)
    assert b.config(0) == before
    steps = 66
    b.record(steps + 8, profiles=True)
    b.run(steps)
    _equals_single_runs(V, b, inst, descs, steps, (0, len(inst) - 1))
