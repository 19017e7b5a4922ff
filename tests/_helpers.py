"""Shared test helpers (checker side)."""
import numpy as np

ORACLE_KEYS = ("n", "qmin", "qmax", "pmin", "pmax", "qscale", "pscale",
               "scheme", "interp", "rot_clamp", "angle", "rf_linear",
               "revolutionpart", "V_RF", "f_RF", "V0", "slip", "E0", "use_fp",
               "fptype", "e1", "deriv", "use_wake", "spacing_bins", "use_csr",
               "fc", "renormalize")


def oracle_cfg(desc, nmax):
    cfg = {k: desc[k] for k in ORACLE_KEYS if k in desc}
    if desc.get("scheme", 0) == 0 and desc.get("rf_linear"):
        cfg["angle"] = desc["rf_angle"]
    cfg["nmax"] = nmax
    cfg["wakescaling_arg"] = desc["wakescaling"]
    return cfg


def rel_linf(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def noisy_mesh(n, seed, nb=1, neg=False):
    rng = np.random.default_rng(seed)
    x = np.linspace(-3, 3, n)
    g = np.exp(-0.5 * (x[:, None] ** 2 + x[None, :] ** 2))
    d = g[None] + 1e-3 * rng.random((nb, n, n))
    if neg:
        d -= 2e-3 * rng.random((nb, n, n))
    return d.astype(np.float32)
