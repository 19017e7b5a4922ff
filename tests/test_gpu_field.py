"""ElectricField, impedances and the FP64 Airy kernel on the GPU against the
oracle and the goldens (1e-5 relative L-inf for FFT results)."""
import math

import numpy as np
import pytest

from _helpers import noisy_mesh, rel_linf

pytestmark = pytest.mark.gpu

C_LIGHT = 2.99792458e8


@pytest.fixture(scope="module")
def V():
    import inovesa_amd as V
    return V


def test_airy_against_mpmath(V, golden):
    g = golden("airy_mpmath.json")
    xs = np.array([float(r[0]) for r in g["rows"]])
    ref = np.array([[float(v) for v in r[1:]] for r in g["rows"]])
    vals, ovf = V.airy(xs)
    assert not ovf.any()
    for i in range(len(xs)):
        for k in range(4):
            r = ref[i, k]
            if abs(r) < 1e-290:
                continue
            assert abs(vals[i, k] - r) <= 2e-13 * abs(r), (xs[i], k)
    _, ovf = V.airy([x for x, _ in g["overflow"]])
    assert list(ovf) == [e for _, e in g["overflow"]]


@pytest.mark.parametrize("n", [1024, 2048, 4096, 8192])
def test_freespace_rw_const_tables(V, orc, n):
    frev, fmax = np.float32(9e6), np.float32(6.35e12)
    z = V.FreeSpaceCSR(n, frev, fmax).data()
    ref = orc.z_freespace(n, frev, fmax)
    np.testing.assert_allclose(z, ref, rtol=4e-7, atol=0)
    L, s, xi, b = C_LIGHT / float(frev), 1.4e6, 0.0, 0.016
    z = V.ResistiveWall(n, frev, fmax, L, s, xi, b).data()
    np.testing.assert_allclose(z, orc.z_resistive_wall(n, frev, fmax, L, s,
                                                       xi, b),
                               rtol=4e-7, atol=0)
    z = V.ConstImpedance(n, fmax, 1e3 + 5j).data()
    np.testing.assert_array_equal(z, orc.z_const(n, 1e3 + 5j))
    z = V.CollimatorImpedance(n, fmax, 0.016, 0.004).data()
    np.testing.assert_allclose(z, orc.z_collimator(n, 0.016, 0.004),
                               rtol=1e-7)
    zs = V.FreeSpaceCSR(n, frev, fmax)
    zs += V.ConstImpedance(n, fmax, 2.0)
    np.testing.assert_allclose(zs.data(),
                               (orc.z_freespace(n, frev, fmax)
                                + orc.z_const(n, 2.0 + 0j)).astype(np.complex64),
                               rtol=4e-7)


def test_parallel_plates_golden(V, golden):
    g = golden("pp_mpmath.json")
    z = V.ParallelPlatesCSR(g["nfreqs"], g["f0"], g["f_max"], g["g"]).data()
    assert z[0] == 0 and np.all(z[g["nfreqs"] // 2 + 1:] == 0)
    for i, re, im in g["rows"]:
        ef = np.complex64(complex(re, im))
        assert abs(z[i].real - ef.real) <= 3e-7 * abs(ef.real), i
        assert abs(z[i].imag - ef.imag) <= 3e-7 * abs(ef.imag), i


@pytest.mark.parametrize("n,gap", [(512, 0.032), (256, 0.016), (1024, 0.064)])
def test_parallel_plates_config_tables(V, orc, n, gap):
    d = V.derive(grid_size=n, vacuum_gap=gap)
    nmax, f0, fmax = d["padded_bins"], np.float32(d["f0"]), d["fmax"]
    z = V.ParallelPlatesCSR(nmax, f0, fmax, gap).data()
    ref, terms = orc.z_parallel_plates(nmax, f0, np.float32(fmax), gap)
    assert terms > 0
    scale = np.abs(ref).max()
    assert np.abs(z - ref).max() <= 1e-6 * scale
    big = np.abs(ref) > 1e-3 * scale
    np.testing.assert_allclose(z[big], ref[big], rtol=1e-6)


def _field_fixture(V, filling, zval):
    n, spacing = 32, 256
    spaced = spacing * len(filling)
    ps = V.PhaseSpace(n, -12, 12, 3, -12, 12, 4, filling=filling)
    z = V.ConstImpedance(spaced, 1e9, zval)
    ws = V.ElectricField.scaling(1e-3, 1e-3, 3.0, ps.getDelta(1), 1e3, 1e9)
    f = V.ElectricField(ps, z, [0], spacing, 1e6, 0.1, ws)
    return n, ps, z, f


def test_padding_and_constant_wake(V, orc):
    # test/TestElectricField.cpp:110-168 (single bunch; the padded length
    # is a power of two, as RoundPadding makes it)
    n, ps, z, f = _field_fixture(V, [1.0], 1.0)
    f.padBunchProfiles()
    pad = f.getPaddedBunchProfiles()
    ax = ps.getAxis(0)
    exp = np.zeros(256)
    exp[:n] = orc.gaus(ax)
    np.testing.assert_allclose(pad, exp, rtol=1e-4 * 1e-2, atol=1e-9)
    wp = f.wakePotential()
    prof = ps.getProjection(0)
    assert np.abs(wp - prof * f.getWakeScaling()).max() < 1e-6


def test_forward_wake_causal(V):
    # test/TestElectricField.cpp:175-239
    n, ps, z, f = _field_fixture(V, [1.0], 0.0)
    z += V.FreeSpaceCSR(256, 1e6, 1e12)
    w0 = f.wakePotential()
    prof = ps.getProjection(0)
    c = n // 2
    dl = np.float32(0.1) * prof[0, c]
    prof[0, c] -= dl
    prof[0, c + 1] += dl
    prof[0, c + 2] += dl
    prof[0, c + 3] -= dl
    ps.setProjection(0, prof)
    w1 = f.wakePotential()
    assert np.abs(w0 - w1).max() > 2e-8
    assert np.abs(w0[0, :c] - w1[0, :c]).max() < 2e-8


@pytest.mark.parametrize("n,zkind", [(256, "fs"), (512, "pp"), (1024, "pprw"),
                                     (128, "const")])
def test_wake_potential_vs_oracle(V, orc, n, zkind):
    d = V.derive(grid_size=n, vacuum_gap=0.032 if zkind != "fs" else -1.0)
    nmax = d["padded_bins"]
    f0, fmax = np.float32(d["f0"]), np.float32(d["fmax"])
    if zkind == "fs":
        z = V.FreeSpaceCSR(nmax, f0, fmax)
    elif zkind == "const":
        z = V.ConstImpedance(nmax, fmax, 300 + 40j)
    else:
        z = V.ParallelPlatesCSR(nmax, f0, fmax, 0.032)
        if zkind == "pprw":
            z += V.ResistiveWall(nmax, np.float32(9e6), fmax,
                                 C_LIGHT / 9e6, 1.4e6, 0.0, 0.016)
    data = noisy_mesh(n, 31)
    ps = V.PhaseSpace(n, d["qmin"], d["qmax"], d["qscale"], d["pmin"],
                      d["pmax"], d["pscale"], data=data)
    f = V.ElectricField(ps, z, [0], d["spacing_bins"], 9e6,
                        d["revolutionpart"], d["wakescaling"])
    wp = f.wakePotential()
    prof = ps.getProjection(0)
    ref, pad = orc.wake_potential(prof, z.data(), [0], d["spacing_bins"],
                                  np.float32(d["wakescaling"]))
    assert rel_linf(wp, ref) <= 1e-5
    assert rel_linf(f.getPaddedBunchProfiles(), pad) == 0


@pytest.mark.parametrize("n,fc", [(512, 23e9), (256, 0.0), (1024, 5e11)])
def test_csr_vs_oracle(V, orc, n, fc):
    d = V.derive(grid_size=n, vacuum_gap=0.032)
    nmax = d["padded_bins"]
    z = V.ParallelPlatesCSR(nmax, np.float32(d["f0"]), d["fmax"], 0.032)
    ps = V.PhaseSpace(n, d["qmin"], d["qmax"], d["qscale"], d["pmin"],
                      d["pmax"], d["pscale"], data=noisy_mesh(n, 9))
    f = V.ElectricField(ps, z, [0], 0, 9e6, d["revolutionpart"])
    power, spec = f.updateCSR(np.float32(fc))
    prof = ps.getProjection(0)
    rspec, rpower = orc.update_csr(prof, z.data(), ps.getDelta(0),
                                   np.float32(d["qscale"]), np.float32(fc))
    assert rel_linf(spec, rspec) <= 1e-5
    # the power is the FP64 sum of the reference's float terms; the
    # reference sums them sequentially in float (error <= nmax*eps*sum)
    nmax = spec.shape[1]
    fdelta = np.float32(np.float32(1) / ps.getDelta(0)) / np.float32(nmax - 1)
    terms = (fdelta * rspec[0]).astype(np.float64)
    exact = terms.sum()
    assert abs(power[0] - exact) <= 1e-6 * abs(exact)
    assert abs(rpower[0] - exact) <= nmax * 2.0 ** -23 * np.abs(terms).sum()


def _multibunch_fixture(V, pattern, zval):
    """test/TestElectricField.cpp:22-69: n = 32, 8 bins of bunch spacing
    per mesh width (spacing 256 bins), one bucket per pattern entry (the
    FFT length is 256 per bucket: 768 for three), filled buckets numbered
    from the end of the pattern."""
    n, spacing = 32, 256
    spaced = spacing * len(pattern)
    buckets = [len(pattern) - 1 - i for i, f in enumerate(pattern) if f > 0]
    bunches = [f for f in pattern if f > 0]
    ps = V.PhaseSpace(n, -12, 12, 3, -12, 12, 4, filling=bunches)
    z = V.ConstImpedance(spaced, 1e9, zval)
    ws = V.ElectricField.scaling(1e-3, 1e-3, 3.0, ps.getDelta(1), 1e3, 1e9)
    f = V.ElectricField(ps, z, buckets, spacing, 1e6, 0.1, ws)
    return n, spacing, buckets, ps, z, f


PATTERNS = [[1.0], [0.75, 0.0, 0.25]]


@pytest.mark.parametrize("pattern", PATTERNS)
def test_multibunch_padding(V, orc, pattern):
    # test/TestElectricField.cpp:110-141
    n, spacing, buckets, ps, z, f = _multibunch_fixture(V, pattern, 1.0)
    f.padBunchProfiles()
    pad = f.getPaddedBunchProfiles()
    g = orc.gaus(ps.getAxis(0)).astype(np.float64)
    exp = np.zeros(spacing * len(pattern))
    for bucket, norm in enumerate(reversed(pattern)):
        exp[bucket * spacing: bucket * spacing + n] = norm * g
    np.testing.assert_allclose(pad, exp, rtol=1e-6, atol=1e-9)


@pytest.mark.parametrize("pattern", PATTERNS)
def test_multibunch_constant_wake(V, orc, pattern):
    """test/TestElectricField.cpp:147-168: with a constant impedance of 1
    the wake potential of every bunch is its profile times the scaling
    (within 1e-6), here with an FFT length of 768 (3 x 256) for the
    three-bucket pattern; and it equals the oracle's."""
    n, spacing, buckets, ps, z, f = _multibunch_fixture(V, pattern, 1.0)
    wp = f.wakePotential()
    prof = ps.getProjection(0)
    assert wp.shape == (len(buckets), n)
    assert np.abs(wp - prof * f.getWakeScaling()).max() < 1e-6
    ws = V.ElectricField.scaling(1e-3, 1e-3, 3.0, ps.getDelta(1), 1e3, 1e9)
    ref, pad = orc.wake_potential(prof, z.data(), buckets, spacing,
                                  np.float32(ws))
    assert rel_linf(wp, ref) <= 1e-5
    assert rel_linf(f.getPaddedBunchProfiles(), pad) == 0


@pytest.mark.parametrize("pattern", PATTERNS)
def test_multibunch_forward_wake_causal(V, pattern):
    """test/TestElectricField.cpp:175-239 on every bunch: free-space CSR
    over 256 or 768 bins; a change of the profile from x = n/2 on moves the
    wake only from there on."""
    n, spacing, buckets, ps, z, f = _multibunch_fixture(V, pattern, 0.0)
    z += V.FreeSpaceCSR(spacing * len(pattern), 1e6, 1e12)
    w0 = f.wakePotential()
    prof = ps.getProjection(0)
    c = n // 2
    for b in range(len(buckets)):
        dl = np.float32(0.1) * prof[b, c]
        prof[b, c] -= dl
        prof[b, c + 1] += dl
        prof[b, c + 2] += dl
        prof[b, c + 3] -= dl
    ps.setProjection(0, prof)
    w1 = f.wakePotential()
    for b in range(len(buckets)):
        assert np.abs(w0[b] - w1[b]).max() > 2e-8
        assert np.abs(w0[b, :c] - w1[b, :c]).max() < 2e-8


@pytest.mark.parametrize("nmax", [768, 1536, 3072])
def test_nonpower_of_two_wake_and_csr_vs_oracle(V, orc, nmax):
    """FFT lengths that are not powers of two (RoundPadding off: 3 x 2^k)
    through the direct DFT: wake potential and CSR spectrum of two bunches
    against the oracle."""
    n = 128 if nmax < 3072 else 384
    d = V.derive(grid_size=n, vacuum_gap=0.032)
    f0, fmax = np.float32(d["f0"]), np.float32(d["fmax"])
    z = V.ParallelPlatesCSR(nmax, f0, fmax, 0.032)
    z += V.ConstImpedance(nmax, fmax, 50 + 10j)
    data = np.concatenate([noisy_mesh(n, 41), noisy_mesh(n, 42)])
    ps = V.PhaseSpace(n, d["qmin"], d["qmax"], d["qscale"], d["pmin"],
                      d["pmax"], d["pscale"], filling=[0.6, 0.4], data=data)
    spacing = (nmax - n) // 2
    buckets = [2, 0]
    f = V.ElectricField(ps, z, buckets, spacing, 9e6, d["revolutionpart"],
                        d["wakescaling"])
    wp = f.wakePotential()
    prof = ps.getProjection(0)
    ref, pad = orc.wake_potential(prof, z.data(), buckets, spacing,
                                  np.float32(d["wakescaling"]))
    assert rel_linf(wp, ref) <= 1e-5
    power, spec = f.updateCSR(np.float32(23e9))
    rspec, rpower = orc.update_csr(prof, z.data(), ps.getDelta(0),
                                   np.float32(d["qscale"]), np.float32(23e9))
    assert rel_linf(spec, rspec) <= 1e-5
    np.testing.assert_allclose(power, rpower, rtol=2e-5)


@pytest.mark.parametrize("nmax", [4, 8, 16, 32, 64, 128, 256, 512, 1024, 2048,
                                  4096, 8192])
def test_every_fft_length_vs_oracle(V, orc, nmax):
    """The field kernel's transform at every power-of-two length it takes
    (complex length M = nmax/2 from 2 to 4096: radix-8 passes with the
    radix-4 or radix-2 pass each length needs), on a random complex
    impedance: the wake potential within 1e-5 relative L-inf of the
    oracle's, which numpy's FFT confirms (test_toeplitz pins the oracle's
    transform)."""
    n = min(nmax // 2 if nmax >= 8 else nmax, 512)
    d = V.derive(grid_size=n)
    rng = np.random.default_rng(nmax)
    table = (rng.standard_normal(nmax)
             + 1j * rng.standard_normal(nmax)).astype(np.complex64)
    z = V.Impedance(nmax, np.float32(d["fmax"]), table=table)
    ps = V.PhaseSpace(n, d["qmin"], d["qmax"], d["qscale"], d["pmin"],
                      d["pmax"], d["pscale"], data=noisy_mesh(n, nmax))
    f = V.ElectricField(ps, z, [0], 0, 9e6, d["revolutionpart"],
                        np.float32(1.0))
    wp = f.wakePotential()
    prof = ps.getProjection(0)
    ref, _ = orc.wake_potential(prof, z.data(), [0], 0, np.float32(1.0))
    assert rel_linf(wp, ref) <= 1e-5
    # numpy's transform of the reference's recipe: W_k = Z_k F_k below the
    # Nyquist bin (which stays zero), Im W_0 dropped, unnormalised C2R
    # times the scaling over nmax
    F = np.fft.rfft(np.pad(prof[0].astype(np.float64), (0, nmax - n)))
    W = np.zeros(nmax // 2 + 1, complex)
    W[:nmax // 2] = table[:nmax // 2].astype(complex) * F[:nmax // 2]
    W[0] = W[0].real
    exact = np.fft.irfft(W, nmax)[:n]
    assert rel_linf(wp[0], exact) <= 1e-5
