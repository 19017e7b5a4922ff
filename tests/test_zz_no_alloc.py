"""vfps_wake_potential_noalloc (the wake potential on caller-owned device
buffers, allocating nothing) against vfps_field_wake_potential on the same
profiles and impedance, on consecutive calls of shrinking size."""
import numpy as np
import pytest

from _helpers import rel_linf

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def V():
    import inovesa_amd as V
    return V


def _field_wake(V, proj, ztab, buckets, spacing, ws):
    nb, n = proj.shape
    ps = V.PhaseSpace(n, -6.0, 6.0, 1e-3, -6.0, 6.0, 1e5,
                      filling=[1.0 / nb] * nb)
    ps.setProjection(0, proj)
    z = V.Impedance(len(ztab), 1e12, table=ztab)
    f = V.ElectricField(ps, z, buckets, spacing, wakescaling=ws)
    return f.wakePotential()


class DevBuf:
    """A device buffer of at least `size` floats: the data array of a mesh
    of the library (its device pointer), filled from the host."""

    def __init__(self, V, values, size=None):
        values = np.ascontiguousarray(values, np.float32).ravel()
        size = max(size or values.size, values.size, 4)
        self.m = int(np.ceil(np.sqrt(size)))
        self.ps = V.PhaseSpace(self.m, -1.0, 1.0, 1.0, -1.0, 1.0, 1.0)
        data = np.full(self.m * self.m, np.nan, np.float32)
        data[:values.size] = values
        self.ps.setData(data.reshape(1, self.m, self.m))
        self.ptr = int(self.ps.device_ptr())

    def read(self, count):
        return self.ps.getData().ravel()[:count]


def _noalloc_wake(V, proj, ztab, buckets, spacing, ws):
    nb, n = proj.shape
    pd = DevBuf(V, proj)
    zd = DevBuf(V, np.ascontiguousarray(ztab).view(np.float32))
    wd = DevBuf(V, np.full(nb * n, np.nan, np.float32))
    V.wake_potential_noalloc(pd.ptr, nb, n, zd.ptr, len(ztab), buckets,
                             spacing, ws, wd.ptr)
    V.Device.default().sync()
    return wd.read(nb * n).reshape(nb, n)


def _inputs(n, nmax, nb, seed):
    rng = np.random.default_rng(seed)
    x = np.linspace(-6.0, 6.0, n)
    proj = (np.exp(-0.5 * x * x)[None, :]
            * (1.0 + 0.1 * rng.random((nb, n)))).astype(np.float32)
    k = np.arange(nmax)
    z = ((k ** (1.0 / 3.0)) * (1.0 + 0.5j)
         * (1.0 + 0.05 * rng.standard_normal(nmax))).astype(np.complex64)
    z[nmax // 2 + 1:] = 0
    return proj, z


CASES = [  # n, nmax, buckets, spacing, wakescaling: shrinking inputs
    (512, 4096, [0], 0, np.float32(3.5e-4)),
    (64, 1024, [0, 2, 5], 100, np.float32(2.0e-3)),
    (32, 256, [1, 0], 40, np.float32(1.0)),
]


def test_noalloc_wake_equals_field_wake(V):
    for k, (n, nmax, buckets, spacing, ws) in enumerate(CASES):
        print(f"noalloc case {k}: n={n} nmax={nmax} buckets={buckets}",
              flush=True)
        proj, z = _inputs(n, nmax, len(buckets), seed=10 + k)
        ref = _field_wake(V, proj, z, buckets, spacing, ws)
        out = _noalloc_wake(V, proj, z, buckets, spacing, ws)
        assert np.all(np.isfinite(out))
        # the twiddle tables come from the device's cos/sin here and the
        # host's in the field object: last-bit differences of FP64 inputs
        err = rel_linf(out, ref)
        print(f"  rel-Linf {err:.2e}", flush=True)
        assert err <= 1e-6, (k, err)


def test_noalloc_rejects_beyond_scratch(V):
    """Sizes beyond the scratch are refused before any device work."""
    buf = DevBuf(V, np.zeros(2 * 16384, np.float32))
    p = buf.ptr
    with pytest.raises(V.VfpsInvalid, match="scratch"):
        V.wake_potential_noalloc(p, 1, 64, p, 16384, [0], 0, 1.0, p)
    with pytest.raises(V.VfpsInvalid):
        V.wake_potential_noalloc(p, 65, 1, p, 1024, [0] * 65, 0, 1.0, p)
    with pytest.raises(V.VfpsInvalid, match="fit"):
        V.wake_potential_noalloc(p, 1, 64, p, 1024, [1], 1000, 1.0, p)
