"""The oracle against the reference's own unit-test vectors (CPU)."""
import numpy as np
import pytest


def _ruler(orc, n, a, b):
    return orc.ruler(n, a, b)


@pytest.mark.parametrize("case", ["kickmap_x", "kickmap_x_overflow",
                                  "kickmap_y"])
def test_kickmap_vectors(orc, golden, case):
    v = golden("reference_vectors.json")[case]
    d = np.array(v["data_in"], np.float32).reshape(1, 4, 4)
    out = orc.kick(v["axis"], d, v["offset"], v["it"])
    np.testing.assert_allclose(out.ravel(), v["data_out"], atol=1e-5)


@pytest.mark.parametrize("case", ["rotation_linear_noclamp_map",
                                  "rotation_linear_offcenter",
                                  "rotation_quadratic_noclamp_nomap",
                                  "rotation_cubic_clamp_map"])
def test_rotation_vectors(orc, golden, case):
    v = golden("reference_vectors.json")[case]
    a0, d0, z0 = _ruler(orc, 4, v["qmin"], v["qmax"])
    a1, d1, z1 = _ruler(orc, 4, v["pmin"], v["pmax"])
    d = np.array(v["data_in"], np.float32).reshape(4, 4)
    out = orc.rotation(d, a0, d0, z0, a1, d1, z1, np.float32(v["angle"]),
                       v["it"], v["clamp"], v["rotmapsize"])
    np.testing.assert_allclose(out.ravel(), v["data_out"], atol=1e-5)


def test_rotation_throws(orc, golden):
    a0, d0, z0 = _ruler(orc, 4, -1, 1)
    for c in golden("reference_vectors.json")["rotation_throws"]:
        with pytest.raises(ValueError):
            orc.rotation(np.zeros((4, 4), np.float32), a0, d0, z0, a0, d0, z0,
                         np.float32(np.pi / 2), c["it"], c["clamp"],
                         c["rotmapsize"])


def test_ruler_and_power_of_two(orc, golden):
    v = golden("reference_vectors.json")
    r = v["ruler"]
    data, delta, zb = orc.ruler(r["steps"], r["min"], r["max"])
    assert zb == r["zerobin"]
    assert data[-1] - data[0] == r["length"]
    for a, b in v["upper_power_of_two"]:
        assert orc.upper_power_of_two(a) == b


def test_phasespace_gaussian(orc, golden):
    g = golden("reference_vectors.json")["phasespace_gaussian"]
    n = g["n"]
    data, px, py, fill, integral = orc.phasespace_gaussian(
        1, n, g["min"], g["max"], g["min"], g["max"], [1.0])
    assert abs(integral - 1) < g["rel_tol"]
    ax, d, _ = orc.ruler(n, g["min"], g["max"])
    assert abs(d - g["delta"]) < 1e-3 * g["delta"]
    for proj in (px, py):
        mean, var, rms = orc.variance(proj, ax, d, np.ones(1, np.float32),
                                      fill)
        assert abs(mean[0]) < 1e-3
        assert abs(rms[0] - g["sigma"]) < g["rel_tol"]
    # test/TestPhaseSpace.cpp:109-116: filling must sum to one
    with pytest.raises(ValueError):
        orc.phasespace_gaussian(2, n, -12, 12, -12, 12, [0.75, 0.75])


def test_five_buckets(orc):
    # test/TestPhaseSpace.cpp:75-107
    buckets = [0.75, 0, 0.15, 0.1, 0]
    data, px, py, fill, integral = orc.phasespace_gaussian(
        5, 32, -12, 12, -12, 12, buckets)
    assert abs(integral - 1) < 1e-3
    np.testing.assert_allclose(fill, buckets, rtol=1e-3, atol=1e-7)
    ax, d, _ = orc.ruler(32, -12, 12)
    mean, var, rms = orc.variance(px, ax, d, np.array(buckets, np.float32),
                                  fill)
    for i, b in enumerate(buckets):
        if b > 0:
            assert abs(rms[i] - 1) < 1e-3
        else:
            assert rms[i] == 0


def test_const_and_sum_impedance(orc, golden):
    v = golden("reference_vectors.json")
    c = v["const_impedance"]
    z = orc.z_const(c["n"], complex(c["z"], 0))
    np.testing.assert_array_equal(z.real, c["expect"])
    np.testing.assert_array_equal(z.imag, 0)
