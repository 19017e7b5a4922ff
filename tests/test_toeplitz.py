"""The Toeplitz kernels of the direct wake and CSR forms (host code of the
library, no GPU): the direct sums must reproduce the zero-padded FFT
convolution of src/PS/ElectricField.cpp:215-241 and the CSR power of
src/PS/ElectricField.cpp:133-184."""
import numpy as np
import pytest

from _helpers import rel_linf


@pytest.fixture(scope="module")
def V():
    import inovesa_amd as V
    return V


def _fft_wake(z, p, nmax):
    """numpy FP64: C2R(Z * R2C(pad(p))) with the Nyquist bin zero and the
    imaginary part of bin 0 dropped, unnormalised."""
    pad = np.zeros(nmax)
    pad[:len(p)] = p
    F = np.fft.rfft(pad)
    W = z[:nmax // 2 + 1].astype(np.complex128) * F
    W[-1] = 0
    W[0] = W[0].real
    return np.fft.irfft(W, nmax)[:len(p)] * nmax


def _toeplitz_apply(k, p):
    n = len(p)
    idx = np.arange(n)[:, None] - np.arange(n)[None, :] + n - 1
    return k[idx] @ p


@pytest.mark.parametrize("nmax,n", [(16, 8), (64, 20), (256, 128),
                                    (1024, 128)])
def test_wake_kernel_is_the_fft_convolution(V, nmax, n):
    rng = np.random.default_rng(nmax + n)
    z = (rng.standard_normal(nmax)
         + 1j * rng.standard_normal(nmax)).astype(np.complex64)
    p = rng.random(n)
    g, _ = V.toeplitz_kernels(nmax, n, z=z)
    ref = _fft_wake(z, p, nmax)
    assert rel_linf(_toeplitz_apply(g, p), ref) <= 1e-13


@pytest.mark.parametrize("nmax,n", [(16, 8), (64, 20), (1024, 128)])
def test_csr_kernel_is_the_power_sum(V, nmax, n):
    rng = np.random.default_rng(7 * nmax + n)
    zc = (rng.standard_normal(nmax)
          + 1j * rng.standard_normal(nmax)).astype(np.complex64)
    rn = rng.random(nmax).astype(np.float32)
    fdelta = np.float32(0.37)
    p = rng.random(n)
    _, h = V.toeplitz_kernels(nmax, n, zc=zc, renorm=rn, fdelta=fdelta)
    pad = np.zeros(nmax)
    pad[:n] = p
    F = np.fft.rfft(pad)
    c = (np.float64(fdelta) * rn[:nmax // 2 + 1].astype(np.float64)
         * zc[:nmax // 2 + 1].real.astype(np.float64))
    power = (c * np.abs(F) ** 2).sum()
    assert abs(p @ _toeplitz_apply(h, p) - power) <= 1e-13 * abs(power)


def test_kernel_symmetries(V):
    # a real, even impedance gives an even g; h is always even (cosines)
    nmax, n = 128, 40
    rng = np.random.default_rng(3)
    z = rng.standard_normal(nmax).astype(np.complex64)
    g, h = V.toeplitz_kernels(nmax, n, z=z, zc=z, renorm=np.ones(nmax),
                              fdelta=1.0)
    assert np.allclose(g, g[::-1], rtol=0, atol=1e-12 * np.abs(g).max())
    assert np.allclose(h, h[::-1], rtol=0, atol=1e-12 * np.abs(h).max())
    # a constant impedance is the identity (test/TestElectricField.cpp:147)
    # up to the missing Nyquist bin: g[d] = Z (nmax delta_d - (-1)^d)
    g, _ = V.toeplitz_kernels(nmax, n, z=np.full(nmax, 2.5, np.complex64))
    d = np.arange(-(n - 1), n)
    e = -2.5 * (-1.0) ** d
    e[n - 1] += 2.5 * nmax
    np.testing.assert_allclose(g, e, rtol=0, atol=1e-12 * nmax)


@pytest.mark.parametrize("n,gap", [(128, 0.032), (256, 0.016)])
def test_direct_wake_matches_oracle(V, orc, n, gap):
    """The oracle keeps the reference's float roundings (F, W = Z F and the
    C2R output are float); the direct FP64 form sits within 1e-6 of it."""
    from _helpers import noisy_mesh
    d = V.derive(grid_size=n, vacuum_gap=gap)
    nmax = d["padded_bins"]
    z, _ = orc.z_parallel_plates(nmax, np.float32(d["f0"]),
                                 np.float32(d["fmax"]), gap)
    ax, dq, _ = orc.ruler(n, d["qmin"], d["qmax"])
    ws = orc.simpson_weights(n, dq)
    proj = orc.project_x(noisy_mesh(n, 5), ws)
    wp, _ = orc.wake_potential(proj, z, [0], 0, np.float32(d["wakescaling"]))
    g, _ = V.toeplitz_kernels(nmax, n, z=z)
    scale = np.float32(np.float32(d["wakescaling"]) / np.float32(nmax))
    direct = scale * _toeplitz_apply(g, proj[0].astype(np.float64)).astype(
        np.float32)
    assert rel_linf(direct, wp[0]) <= 1e-6


def test_bad_arguments(V):
    with pytest.raises(V.VfpsInvalid):
        V.toeplitz_kernels(6, 4, z=np.ones(6, np.complex64))
    with pytest.raises(V.VfpsInvalid):
        V.toeplitz_kernels(16, 4, zc=np.ones(16, np.complex64))
