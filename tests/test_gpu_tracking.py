"""Passive particle tracking on the GPU (SourceMap::applyTo,
src/SM/KickMap.cpp:280-304, src/SM/RotationMap.cpp:178-188,
src/SM/FokkerPlanckMap.cpp:214-268) against the oracle."""
import math

import numpy as np
import pytest

from _helpers import noisy_mesh

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def V():
    import inovesa_amd as V
    return V


def _cos_sin(angle):
    """cosf/sinf of -angle, as RotationMap's constructor computes them in
    float (src/SM/RotationMap.cpp:24-25)"""
    import ctypes
    m = ctypes.CDLL("libm.so.6")
    m.cosf.restype = m.sinf.restype = ctypes.c_float
    m.cosf.argtypes = m.sinf.argtypes = [ctypes.c_float]
    a = -float(np.float32(angle))
    return np.float32(m.cosf(a)), np.float32(m.sinf(a))


def _positions(n, k, seed):
    rng = np.random.default_rng(seed)
    return (1.0 + rng.random((k, 2)) * (n - 2.5)).astype(np.float32)


def test_map_apply_to_bitexact(V, orc):
    n = 128
    d = noisy_mesh(n, 13)
    a = V.PhaseSpace(n, -6.0, 6.0, 2.0e-3, -6.0, 6.0, 6.11e5, data=d)
    b = V.PhaseSpace(n, -6.0, 6.0, 2.0e-3, -6.0, 6.0, 6.11e5)
    pos = _positions(n, 300, 1)
    off = (np.random.default_rng(2).standard_normal(n) * 3).astype(np.float32)
    for axis in (0, 1):
        km = V.KickMap(a, b, 4, False, axis)
        km.swapOffset(off)
        assert np.array_equal(km.applyToAll(pos), orc.track_kick(axis, n,
                                                                 off, pos))
    ang = np.float32(2 * math.pi / 50)
    rm = V.RotationMap(a, b, ang, 4, True, n * n)
    _, _, zb0 = orc.ruler(n, -6.0, 6.0)
    cs, sn = _cos_sin(ang)
    assert np.array_equal(rm.applyToAll(pos),
                          orc.track_rotation(cs, sn, zb0, zb0, pos))
    ax1, d1, zb1 = orc.ruler(n, -6.0, 6.0)
    for mode in (0, 1, 2):
        fp = V.FokkerPlanckMap(a, b, 3, mode, 2e-2, 4)
        h, ip = orc.fp_build(n, 3, np.float32(2e-2), 4, ax1, d1, zb1)
        got = fp.applyToAll(pos)
        if mode == 0:
            assert np.array_equal(got, pos)
        else:
            assert np.array_equal(got, orc.track_fp(mode, h, ip, n, d[0],
                                                    pos)), mode
    ident = V.Identity(a, b)
    assert np.array_equal(ident.applyToAll(pos), pos)


def test_stochastic_tracking_statistics(V):
    """FPTracking::stochastic: y -= y e1 + N(0, sqrt(2 e1)/dp): over many
    particles the mean shift and the spread of one step."""
    n, e1, k = 128, 1e-3, 20000
    a = V.PhaseSpace(n, -6.0, 6.0, 2.0e-3, -6.0, 6.0, 6.11e5)
    b = V.PhaseSpace(n, -6.0, 6.0, 2.0e-3, -6.0, 6.0, 6.11e5)
    fp = V.FokkerPlanckMap(a, b, 3, 3, e1, 4)
    pos = np.tile(np.array([[60.0, 64.0]], np.float32), (k, 1))
    out = fp.applyToAll(pos, seed=5)
    dp = 12.0 / (n - 1)
    sigma = math.sqrt(2 * e1) / dp
    dy = out[:, 1].astype(np.float64) - 64.0
    assert np.all(out[:, 0] == 60.0)
    assert abs(dy.mean() + 64.0 * e1) < 5 * sigma / math.sqrt(k)
    assert abs(dy.std() / sigma - 1) < 0.03
    again = fp.applyToAll(pos, seed=5)    # the next call draws anew
    assert not np.array_equal(again, out)


@pytest.mark.parametrize("scheme", [0, 1])
def test_step_tracking_vs_oracle(V, orc, scheme):
    """Particles through every step of the fused step (wake kick, RF kick,
    drift, FP approximation1; or wake kick, rotation, FP), against the
    oracle's applyTo with the same offsets, bit for bit."""
    from inovesa_amd.workloads import (impedances, initial_mesh,
                                       kick_drift_config, rotation_config)
    n = 256
    if scheme == 0:
        desc, d = kick_drift_config(n=n, steps_per_ts=500)
    else:
        desc, d = rotation_config(n=n)
    zw, zc = impedances(d)
    st = V.Step(desc, initial_mesh(n, seed=3), zw, zc)
    pos = _positions(n, 64, 7)
    st.set_particles(pos, fptrack=1)
    ax0, d0, zb0 = orc.ruler(n, desc["qmin"], desc["qmax"])
    ax1, d1, zb1 = orc.ruler(n, desc["pmin"], desc["pmax"])
    h, ip = orc.fp_build(n, desc["fptype"], np.float32(desc["e1"]),
                         desc["deriv"], ax1, d1, zb1)
    slip = np.array(desc["slip"], np.float32)
    if scheme == 0:
        rf = orc.rfkick_offsets(ax0, np.float32(desc["revolutionpart"]),
                                np.float32(desc["V_RF"]),
                                np.float32(desc["f_RF"]),
                                np.float32(desc["V0"]),
                                np.float32(desc["qscale"]), d1,
                                np.float32(desc["pscale"]))
        dr = orc.drift_offsets(ax1, d0, slip, np.float32(desc["pscale"]),
                               np.float32(desc["E0"]))
    cs, sn = _cos_sin(desc["angle"])
    p = pos.copy()
    for k in range(6):
        st.run(1)
        wake = st.mesh()[2]            # the offsets this step kicked with
        p = orc.track_kick(1, n, wake, p)
        if scheme == 0:
            p = orc.track_kick(1, n, rf, p)
            p = orc.track_kick(0, n, dr, p)
        else:
            p = orc.track_rotation(cs, sn, zb0, zb1, p)
        p = orc.track_fp(1, h, ip, n, None, p)
        got = st.particles()
        assert np.array_equal(got, p), (k, np.abs(got - p).max())
    # tracking leaves the mesh update alone
    plain = V.Step(desc, initial_mesh(n, seed=3), zw, zc)
    plain.run(6)
    assert np.array_equal(plain.mesh()[0], st.mesh()[0])
    with pytest.raises(V.VfpsInvalid):
        st.set_particles(pos, fptrack=2)   # approximation2: per-map step only
