"""The oracle's restatement of SourceMap::applyTo (passive particle
tracking), pinned by closed forms on the CPU."""
import math

import numpy as np


def test_kick_moves_by_interpolated_offset_and_clamps(orc):
    n = 32
    off = np.linspace(-2.0, 3.0, n).astype(np.float32)
    pos = np.array([[10.0, 5.25], [20.0, 30.5], [1.5, 2.0], [30.0, 31.0]],
                   np.float32)
    out = orc.track_kick(0, n, off, pos)   # x moves by the offset at y
    exp = pos[0, 0] - (0.75 * off[5] + 0.25 * off[6])
    assert abs(out[0, 0] - exp) < 1e-5 and out[0, 1] == pos[0, 1]
    # y + 1 beyond the last offset: no kick, only the clamp to [1, n-1]
    assert out[3, 0] == 30.0
    out = orc.track_kick(1, n, np.full(n, 40.0, np.float32), pos)
    assert np.all(out[:, 1] == 1.0)        # clamped below
    out = orc.track_kick(1, n, np.full(n, -40.0, np.float32), pos)
    assert np.all(out[:, 1] == n - 1)      # clamped above


def test_rotation_returns_after_full_turn(orc):
    k = 40
    a = np.float32(2 * math.pi / k)
    cs, sn = np.float32(math.cos(-a)), np.float32(math.sin(-a))
    zb = np.float32(63.5)
    pos = np.array([[80.0, 63.5], [40.0, 70.0]], np.float32)
    p = pos.copy()
    for _ in range(k):
        p = orc.track_rotation(cs, sn, zb, zb, p)
    np.testing.assert_allclose(p, pos, atol=2e-3)
    r0 = np.hypot(pos[:, 0] - zb, pos[:, 1] - zb)
    q = orc.track_rotation(cs, sn, zb, zb, pos)
    np.testing.assert_allclose(np.hypot(q[:, 0] - zb, q[:, 1] - zb), r0,
                               rtol=1e-6)
    # a quarter turn of +x lands on the y axis
    c4, s4 = np.float32(math.cos(-math.pi / 2)), np.float32(math.sin(-math.pi / 2))
    q = orc.track_rotation(c4, s4, zb, zb, np.array([[zb + 10, zb]], np.float32))
    np.testing.assert_allclose(q[0], [zb, zb + 10], atol=1e-5)


def test_fp_approximations_damp_towards_centre(orc):
    n = 64
    ax, d, zb = orc.ruler(n, -6.0, 6.0)
    e1 = np.float32(1e-2)
    h, ip = orc.fp_build(n, 1, e1, 4, ax, d, zb)       # damping only
    pos = np.array([[10.0, 50.2], [10.0, 12.7]], np.float32)
    out = orc.track_fp(1, h, ip, n, None, pos)
    assert out[0, 1] < pos[0, 1] and out[1, 1] > pos[1, 1]
    # no damping, no diffusion: the stencil is the identity
    h0, ip0 = orc.fp_build(n, 0, e1, 4, ax, d, zb)
    assert np.array_equal(orc.track_fp(1, h0, ip0, n, None, pos), pos)
    # approximation2 on a uniform mesh: the stencil's mean displacement
    # sum_j w_j (index_j - y) / sum_j w_j, the opposite sign of
    # approximation1's sum_j (y - index_j) w_j (src/SM/FokkerPlanckMap.cpp
    # :230 and :255 differ in sign), divided by sum_j w_j = 1 + e1
    data = np.ones((n, n), np.float32)
    a2 = orc.track_fp(2, h, ip, n, data, pos)
    np.testing.assert_allclose(a2[:, 1] - pos[:, 1],
                               -(out[:, 1] - pos[:, 1]) / (1 + e1),
                               rtol=1e-4)
