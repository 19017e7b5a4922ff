"""SourceMaps and PhaseSpace on the GPU against the oracle (bit-exact for
the maps; 1e-5 relative for reordered FP32 sums)."""
import math

import numpy as np
import pytest

from _helpers import noisy_mesh, rel_linf

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def V():
    import inovesa_amd as V
    return V


def mesh(V, n, data, qmin=-1.0, qmax=1.0, pmin=-1.0, pmax=1.0, nb=1):
    return V.PhaseSpace(n, qmin, qmax, 2.0, pmin, pmax, 4.0,
                        filling=[1.0 / nb] * nb if nb > 1 else [1.0],
                        data=data)


@pytest.mark.parametrize("case", ["kickmap_x", "kickmap_x_overflow",
                                  "kickmap_y"])
def test_reference_kick_vectors(V, orc, golden, case):
    v = golden("reference_vectors.json")[case]
    d = np.array(v["data_in"], np.float32).reshape(1, 4, 4)
    a, b = mesh(V, 4, d), mesh(V, 4, None)
    km = V.KickMap(a, b, v["it"], False, v["axis"])
    km.swapOffset(v["offset"])
    km.apply()
    out = b.getData()
    np.testing.assert_allclose(out.ravel(), v["data_out"], atol=1e-5)
    assert np.array_equal(out, orc.kick(v["axis"], d, v["offset"], v["it"]))
    np.testing.assert_array_equal(km.getForce(), v["offset"])


@pytest.mark.parametrize("case", ["rotation_linear_noclamp_map",
                                  "rotation_linear_offcenter",
                                  "rotation_quadratic_noclamp_nomap",
                                  "rotation_cubic_clamp_map"])
def test_reference_rotation_vectors(V, golden, case):
    v = golden("reference_vectors.json")[case]
    d = np.array(v["data_in"], np.float32).reshape(1, 4, 4)
    a = mesh(V, 4, d, v["qmin"], v["qmax"], v["pmin"], v["pmax"])
    b = mesh(V, 4, None, v["qmin"], v["qmax"], v["pmin"], v["pmax"])
    rm = V.RotationMap(a, b, np.float32(v["angle"]), v["it"], v["clamp"],
                       v["rotmapsize"])
    rm.apply()
    np.testing.assert_allclose(b.getData().ravel(), v["data_out"], atol=1e-5)


def test_rotation_constructor_throws(V, golden):
    a, b = mesh(V, 4, None), mesh(V, 4, None)
    for c in golden("reference_vectors.json")["rotation_throws"]:
        with pytest.raises(V.VfpsInvalid):
            V.RotationMap(a, b, np.float32(math.pi / 2), c["it"], c["clamp"],
                          c["rotmapsize"])


def _offsets(n, seed):
    rng = np.random.default_rng(seed)
    off = rng.uniform(-4, 4, n).astype(np.float32)
    special = [0, 1, -1, -0.5, -1.0, -1.5, n / 2 - 0.25, n / 2, -n / 2,
               -n / 2 - 0.75, -n / 2 - 1, 1e9, -1e9, 3.0, -2.0, 0.999999]
    off[: len(special)] = special
    off[len(special): len(special) + 8] = rng.uniform(-n, n, 8)
    return off


@pytest.mark.parametrize("axis", [0, 1])
@pytest.mark.parametrize("it", [1, 2, 3, 4])
def test_kick_random_bitexact(V, orc, axis, it):
    n = 512
    d = noisy_mesh(n, 10 + it, neg=True)
    off = _offsets(n, it)
    a, b = mesh(V, n, d, -6, 6, -6, 6), mesh(V, n, None, -6, 6, -6, 6)
    km = V.KickMap(a, b, it, False, axis)
    km.swapOffset(off)
    km.apply()
    assert np.array_equal(b.getData(), orc.kick(axis, d, off, it))


def test_drift_and_rf_maps_bitexact(V, orc):
    n = 512
    d = noisy_mesh(n, 3)
    qmin, qmax, pmin, pmax = -6.0, 6.0, -6.0, 6.0
    qscale, pscale, E0 = 2.0153e-3, 6.11e5, 1.3e9
    a = V.PhaseSpace(n, qmin, qmax, qscale, pmin, pmax, pscale, data=d)
    b = V.PhaseSpace(n, qmin, qmax, qscale, pmin, pmax, pscale)
    ax0, d0, zb0 = orc.ruler(n, qmin, qmax)
    ax1, d1, _ = orc.ruler(n, pmin, pmax)
    angle = np.float32(2 * math.pi / 2000)
    slip = np.array([angle, angle * 0.1, angle * 0.5], np.float32)
    dm = V.DriftMap(a, b, slip, E0, 4)
    ref_off = orc.drift_offsets(ax1, d0, slip, np.float32(pscale), E0)
    assert np.array_equal(dm.getForce()[:n], ref_off)
    dm.apply()
    assert np.array_equal(b.getData(), orc.kick(0, d, ref_off, 4))
    rf = V.RFKickMap(a, b, f_RF=4.5e8, revolutionpart=0.0494,
                     V_RF=998.86e3, V0=47.65e3, it=4)
    ref_off = orc.rfkick_offsets(ax0, 0.0494, 998.86e3, 4.5e8, 47.65e3,
                                 np.float32(qscale), d1, np.float32(pscale))
    assert np.array_equal(rf.getForce()[:n], ref_off)
    rf.apply()
    assert np.array_equal(b.getData(), orc.kick(1, d, ref_off, 4))
    lin = V.RFKickMap(a, b, angle=angle, f_RF=4.5e8, it=4)
    ref_off = orc.rfkick_offsets_linear(n, zb0, angle, np.float32(qscale),
                                        4.5e8, d0)
    assert np.array_equal(lin.getForce()[:n], ref_off)


@pytest.mark.parametrize("deriv", [3, 4, 5])
@pytest.mark.parametrize("fptype", [0, 1, 2, 3])
def test_fokkerplanck_bitexact(V, orc, deriv, fptype):
    n = 256
    d = noisy_mesh(n, 5 + deriv, neg=True)
    a, b = mesh(V, n, d, -6, 6, -6, 6), mesh(V, n, None, -6, 6, -6, 6)
    fp = V.FokkerPlanckMap(a, b, fptype, 0, 2e-3, deriv)
    fp.apply()
    ax1, d1, zb1 = orc.ruler(n, -6, 6)
    h, ip = orc.fp_build(n, fptype, 2e-3, deriv, ax1, d1, zb1)
    assert np.array_equal(b.getData(), orc.fp_apply(d, h, ip))


def test_fokkerplanck_unknown_derivation(V):
    a, b = mesh(V, 8, None), mesh(V, 8, None)
    with pytest.raises(V.VfpsInvalid):
        V.FokkerPlanckMap(a, b, 3, 0, 1e-3, 7)


@pytest.mark.parametrize("it,clamp", [(1, False), (2, False), (3, False),
                                      (4, False), (4, True)])
@pytest.mark.parametrize("angle", [2 * math.pi / 1000, 2 * math.pi / 97,
                                   math.pi / 2, 3.0])
def test_rotation_bitexact(V, orc, it, clamp, angle):
    n = 256
    d = noisy_mesh(n, 21, neg=True)
    lo, hi = -6.0, 6.0
    a, b = mesh(V, n, d, lo, hi, lo, hi), mesh(V, n, None, lo, hi, lo, hi)
    rms = n * n if clamp else 0
    rm = V.RotationMap(a, b, np.float32(angle), it, clamp, rms)
    rm.apply()
    ax, dd, zb = orc.ruler(n, lo, hi)
    ref = orc.rotation(d[0], ax, dd, zb, ax, dd, zb, np.float32(angle), it,
                       clamp, rms)
    assert np.array_equal(b.getData()[0], ref)


def test_rotation_offcentre_axes_bitexact(V, orc):
    n = 128
    d = noisy_mesh(n, 8)
    a = mesh(V, n, d, -7.0, 5.0, -5.5, 6.5)
    b = mesh(V, n, None, -7.0, 5.0, -5.5, 6.5)
    rm = V.RotationMap(a, b, np.float32(0.05), 4, True, n * n)
    rm.apply()
    a0, d0, z0 = orc.ruler(n, -7.0, 5.0)
    a1, d1, z1 = orc.ruler(n, -5.5, 6.5)
    ref = orc.rotation(d[0], a0, d0, z0, a1, d1, z1, np.float32(0.05), 4,
                       True, n * n)
    assert np.array_equal(b.getData()[0], ref)


def test_identity(V):
    d = noisy_mesh(64, 2)
    a, b = mesh(V, 64, d), mesh(V, 64, None)
    V.Identity(a, b).apply()
    assert np.array_equal(b.getData(), d)


def test_multibunch_maps_bitexact(V, orc):
    n, nb = 64, 3
    d = noisy_mesh(n, 4, nb=nb, neg=True)
    a, b = mesh(V, n, d, nb=nb), mesh(V, n, None, nb=nb)
    off = _offsets(n * nb, 5)
    for axis in (0, 1):
        km = V.KickMap(a, b, 4, False, axis)
        km.swapOffset(off)
        km.apply()
        h = orc.kick_update_sm(n, off, 4)
        assert np.array_equal(b.getData(), orc.kick_apply(axis, d, h, 4))
    fp = V.FokkerPlanckMap(a, b, 3, 0, 1e-3, 4)
    fp.apply()
    ax1, d1, zb1 = orc.ruler(n, -1, 1)
    h, ip = orc.fp_build(n, 3, 1e-3, 4, ax1, d1, zb1)
    assert np.array_equal(b.getData(), orc.fp_apply(d, h, ip))


def test_phasespace_gaussian_and_moments(V, orc):
    # test/TestPhaseSpace.cpp:38-107 on the device
    for filling in ([1.0], [0.75, 0.0, 0.15, 0.1, 0.0]):
        ps = V.PhaseSpace(32, -12, 12, 3, -12, 12, 4, filling=filling)
        fill, integral = ps.integrate()
        assert abs(integral - 1) < 1e-3
        np.testing.assert_allclose(fill, filling, rtol=1e-3, atol=1e-7)
        ref, px, py, ofill, _ = orc.phasespace_gaussian(
            len(filling), 32, -12, 12, -12, 12, filling)
        assert rel_linf(ps.getData(), ref) < 1e-5
        ps.updateYProjection()
        for axis in (0, 1):
            mean, var, rms = ps.variance(axis)
            for i, f in enumerate(filling):
                assert abs(mean[i]) < 1e-3
                assert (abs(rms[i] - 1) < 1e-3) if f > 0 else rms[i] == 0
    with pytest.raises(V.VfpsInvalid):
        V.PhaseSpace(32, -12, 12, 2, -12, 12, 4, filling=[0.75, 0.75])


def test_projection_integral_normalize(V, orc):
    n = 512
    d = noisy_mesh(n, 6) * np.float32(3.0)
    ps = V.PhaseSpace(n, -6, 6, 1, -6, 6, 1, data=d)
    ws = ps.simpsonWeights()
    ax, dd, _ = orc.ruler(n, -6, 6)
    assert np.array_equal(ws, orc.simpson_weights(n, dd))
    ps.updateXProjection()
    px = ps.getProjection(0)
    # both projections keep the reference's summation order: bit-exact
    assert np.array_equal(px, orc.project_x(d, ws))
    ps.updateYProjection()
    assert np.array_equal(ps.getProjection(1), orc.project_y(d, ws))
    fill, integral = ps.integrate()
    rf, ri = orc.integrate_f64(px, ws)
    assert abs(integral - ri) <= 1e-12 * abs(ri)
    ps.normalize()
    ps.updateXProjection()
    assert abs(ps.integrate()[1] - 1) < 1e-6


@pytest.mark.parametrize("linear", [False, True])
def test_dynamic_rf_kick_bitexact(V, orc, linear):
    """DynamicRFKickMap (src/SM/DynamicRFKickMap.cpp:97-111): before every
    apply the RF kick is recomputed from the next (phase, amplitude) of the
    queue; each step's offsets and kicked mesh equal the oracle's bit for
    bit, getPastModulation returns the applied entries once, and an apply
    past the queue fails."""
    n, steps, seed = 256, 6, 42
    d = noisy_mesh(n, 21)
    qscale, pscale = 2.0153e-3, 6.11e5
    a = V.PhaseSpace(n, -6.0, 6.0, qscale, -6.0, 6.0, pscale, data=d)
    b = V.PhaseSpace(n, -6.0, 6.0, qscale, -6.0, 6.0, pscale)
    ax0, d0, zb0 = orc.ruler(n, -6.0, 6.0)
    ax1, d1, _ = orc.ruler(n, -6.0, 6.0)
    revpart, f_RF, Vrf, V0 = 0.0494, 4.5e8, 998.86e3, 47.65e3
    noise = dict(phasespread=2e-2, amplspread=1e-2, modampl=5e-2,
                 modtimeincrement=1.0 / 17.0)
    ang = np.float32(2 * math.pi / 1000)
    if linear:
        m = V.DynamicRFKickMap(a, b, revpart, f_RF, steps=steps, seed=seed,
                               angle=ang, **noise)
        sync = np.float32(0.0)
    else:
        m = V.DynamicRFKickMap(a, b, revpart, f_RF, steps=steps, seed=seed,
                               V_RF=Vrf, V0=V0, **noise)
        sync = np.float32(math.asin(np.float32(V0) / np.float32(Vrf)))
    q = orc.dynrf_modulation(steps, seed, sync, revpart, **noise)
    for k in range(steps):
        m.apply()
        if linear:
            off = orc.rfkick_offsets_linear_at(n, zb0, ang, np.float32(qscale),
                                               f_RF, d0, sync, q[k, 0],
                                               q[k, 1])
        else:
            off = orc.rfkick_offsets_at(ax0, revpart, Vrf, f_RF, V0,
                                        np.float32(qscale), d1,
                                        np.float32(pscale), q[k, 0], q[k, 1])
        assert np.array_equal(m.getForce()[:n], off), k
        assert np.array_equal(b.getData(), orc.kick(1, d, off, 4)), k
    assert np.array_equal(m.getPastModulation(), q)
    assert m.getPastModulation().shape == (0, 2)
    with pytest.raises(V.VfpsInvalid):
        m.apply()
