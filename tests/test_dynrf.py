"""DynamicRFKickMap's modulation queue (src/SM/DynamicRFKickMap.cpp:80-95)
and its per-step RF kick offsets (src/SM/RFKickMap.cpp:59-75), on the CPU:
the oracle is pinned by the C++ standard's fixed mt19937 output, closed
forms and the noise statistics; the library's host-side queue equals it."""
import math

import numpy as np
import pytest

ARGS = dict(revolutionpart=0.2022, phasespread=1e-3, amplspread=2e-3,
            modampl=5e-3, modtimeincrement=1.0 / 1700.0)


def test_mt19937_pinned(orc):
    # [rand.predef]: the 10000th invocation of a default-constructed
    # std::mt19937 produces 4123659995
    assert orc.mt19937_10000th() == 4123659995


def test_queue_closed_form_without_noise(orc):
    """No noise: phase = syncphase + modampl sin(float(2 pi inc) * i) in
    float, amplitude exactly 1."""
    steps, sync = 300, np.float32(0.0477)
    q = orc.dynrf_modulation(steps, 7, sync, 0.2022, 0.0, 0.0, 5e-3, 1 / 1700.0)
    dt = np.float32(2 * math.pi * (1 / 1700.0))
    i = np.arange(steps, dtype=np.float32)
    pm = np.float32(5e-3) * np.sin(dt * i).astype(np.float32)
    np.testing.assert_allclose(q[:, 0], sync + pm, rtol=0, atol=2e-9)
    assert np.all(q[:, 1] == 1.0)


def test_queue_noise_statistics(orc):
    """Phase and amplitude noise: zero mean, standard deviation
    spread / sqrt(revolutionpart), independent of each other."""
    steps = 40000
    q = orc.dynrf_modulation(steps, 11, 0.0, ARGS["revolutionpart"], 1e-3,
                             2e-3, 0.0, 0.0)
    sp = 1e-3 / math.sqrt(ARGS["revolutionpart"])
    sa = 2e-3 / math.sqrt(ARGS["revolutionpart"])
    ph, am = q[:, 0].astype(np.float64), q[:, 1].astype(np.float64) - 1.0
    assert abs(ph.mean()) < 5 * sp / math.sqrt(steps)
    assert abs(am.mean()) < 5 * sa / math.sqrt(steps)
    assert abs(ph.std() / sp - 1) < 0.02
    assert abs(am.std() / sa - 1) < 0.02
    assert abs(np.corrcoef(ph, am)[0, 1]) < 0.02
    # a different seed gives a different queue
    q2 = orc.dynrf_modulation(100, 12, 0.0, ARGS["revolutionpart"], 1e-3,
                              2e-3, 0.0, 0.0)
    assert not np.array_equal(q[:100], q2)


def test_library_queue_equals_oracle(orc):
    from inovesa_amd import dynrf_modulation
    for seed in (0, 5, 123456):
        a = dynrf_modulation(500, seed, np.float32(0.0477), **ARGS)
        b = orc.dynrf_modulation(500, seed, np.float32(0.0477), **ARGS)
        assert np.array_equal(a, b)


def test_offsets_at_synchronous_phase_equal_static_kick(orc):
    """_calcKick(_syncphase, 1) is the static RFKickMap's kick."""
    n = 256
    ax0, d0, zb0 = orc.ruler(n, -6.0, 6.0)
    ax1, d1, _ = orc.ruler(n, -6.0, 6.0)
    qs, ps = np.float32(2.0153e-3), np.float32(6.11e5)
    V, V0, f = 998.86e3, 47.65e3, 4.5e8
    sync = np.float32(math.asin(np.float32(V0) / np.float32(V)))
    a = orc.rfkick_offsets_at(ax0, 0.0494, V, f, V0, qs, d1, ps, sync, 1.0)
    b = orc.rfkick_offsets(ax0, 0.0494, V, f, V0, qs, d1, ps)
    assert np.array_equal(a, b)
    ang = np.float32(2 * math.pi / 1000)
    a = orc.rfkick_offsets_linear_at(n, zb0, ang, qs, f, d0, 0.0, 0.0, 1.0)
    b = orc.rfkick_offsets_linear(n, zb0, ang, qs, f, d0)
    assert np.array_equal(a, b)
    # a phase shift of the linear kick is a constant offset
    c = orc.rfkick_offsets_linear_at(n, zb0, ang, qs, f, d0, 0.0, 1e-3, 1.0)
    bl2phase = np.float32(np.float64(qs) / 2.99792458e8 * f * 2 * math.pi)
    shift = np.tan(ang) * np.float32(-1e-3) / bl2phase / d0
    np.testing.assert_allclose(c - b, shift, rtol=1e-4)
