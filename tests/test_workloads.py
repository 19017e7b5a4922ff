"""The synthetic workloads: stability of the explicit Fokker-Planck update
(checked on the oracle, the CPU restatement of the reference)."""
import numpy as np
import pytest

from _helpers import oracle_cfg


def test_diffusion_number():
    from inovesa_amd.workloads import fp_diffusion_number, rotation_config
    assert fp_diffusion_number(256, 2e-3) == pytest.approx(0.9031, rel=1e-3)
    assert fp_diffusion_number(128, 2e-3) == pytest.approx(0.2240, rel=1e-3)
    for n in (128, 256, 384, 512):
        desc, _ = rotation_config(n=n)
        assert fp_diffusion_number(n, desc["e1"]) < 0.5


@pytest.mark.parametrize("damping,bounded", [(1000.0, False), (1.0e4, True)])
def test_reference_fp_update_bound(orc, damping, bounded):
    """A damping time of 1000 steps at 256^2 (diffusion number 0.90) makes
    the reference's explicit update diverge; 1e4 steps (0.09) does not."""
    from inovesa_amd.workloads import initial_mesh, rotation_config
    n = 256
    desc, d = rotation_config(n=n, damping_steps=damping)
    z = orc.z_freespace(d["padded_bins"], np.float32(d["f0"]),
                        np.float32(d["fmax"]))
    sim = orc.sim(oracle_cfg(desc, len(z)), initial_mesh(n, seed=1), z, z)
    sim.run(48)
    mesh, _, _ = sim.get()
    peak = float(np.abs(mesh).max())
    assert (peak < 0.2) == bounded, peak
