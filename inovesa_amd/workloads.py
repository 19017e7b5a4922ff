"""Synthetic workloads of the benchmark configurations.

The initial state of every workload is generated here, once, as FP32 bytes
that are handed unchanged to the GPU library and to any checker: a unit
Gaussian in normalised (z, delta) coordinates (src/PS/PhaseSpace.cpp:471-484
evaluated on the mesh axes of inc/PS/Ruler.hpp) plus seeded uniform noise in
[0, 1e-4), renormalised once to unit charge with the Simpson weights of
src/PS/PhaseSpace.cpp:486-501.
"""
import numpy as np

from .vfps import derive

ONE_DIV_ROOT_TWO_PI = np.float32(0.3989422804014327)


def ruler(n, vmin, vmax):
    """inc/PS/Ruler.hpp:31-51 in float32."""
    vmin, vmax = np.float32(vmin), np.float32(vmax)
    delta = (vmax - vmin) / np.float32(n - 1)
    data = vmin + np.arange(n, dtype=np.float32) * delta
    return data.astype(np.float32), np.float32(delta)


def simpson_weights(n, delta0):
    ws = np.empty(n, np.float32)
    h03 = np.float32(delta0) / np.float32(3)
    ws[0] = h03
    dc = np.float32(1)
    for x in range(1, n - 1):
        ws[x] = h03 * (np.float32(3) + dc)
        dc = -dc
    ws[n - 1] = h03
    return ws


def initial_mesh(n, seed=1, noise=1e-4, half=6.0):
    """Gaussian + U[0, noise) on an n x n mesh over +-half; unit charge."""
    ax, d = ruler(n, -half, half)
    g = (np.float64(ONE_DIV_ROOT_TWO_PI)
         * np.exp(-0.5 * ax.astype(np.float64) ** 2)).astype(np.float32)
    data = g[:, None] * g[None, :]
    rng = np.random.Generator(np.random.Philox(seed))
    data = (data + (rng.random((n, n)) * noise).astype(np.float32))
    data = data.astype(np.float32)
    ws = simpson_weights(n, d).astype(np.float64)
    integral = (data.astype(np.float64) @ ws) @ ws
    data = data * np.float32(1.0 / integral)
    return np.ascontiguousarray(data, np.float32)


def kick_drift_config(n=512, steps_per_ts=2000, gap=0.032, current_factor=1.5,
                      Ib=None, padding=8.0, damping_time=-1.0):
    """Plain VFP, kick-drift scheme (src/main.cpp:1034-1041): DriftMap with
    linear momentum compaction, sinusoidal RFKickMap, parallel-plates CSR
    shielded by `gap`, FokkerPlanckMap (cubic derivative)."""
    base = derive(grid_size=n, steps_per_ts=steps_per_ts, vacuum_gap=gap,
                  padding=padding, damping_time=damping_time)
    if Ib is None:
        Ib = current_factor * base["Ith"]
    d = derive(grid_size=n, steps_per_ts=steps_per_ts, vacuum_gap=gap,
               padding=padding, damping_time=damping_time, bunch_current=Ib)
    desc = dict(n=n, qmin=d["qmin"], qmax=d["qmax"], pmin=d["pmin"],
                pmax=d["pmax"], qscale=d["qscale"], pscale=d["pscale"],
                scheme=0, interp=4, rot_clamp=0, angle=d["angle"],
                rf_linear=0, rf_angle=d["angle"],
                revolutionpart=d["revolutionpart"], V_RF=d["V_eff"],
                f_RF=d["f_RF"], V0=d["V0"], slip=d["slip"],
                E0=d["ring"]["beam_energy"], use_fp=1, fptype=3, e1=d["e1"],
                deriv=4, use_wake=1, spacing_bins=d["spacing_bins"],
                wakescaling=d["wakescaling"], use_csr=1,
                fc=d["ring"]["cutoff_frequency"])
    return desc, d


def production_config(n=1024, steps_per_ts=4000, gap=0.032, alpha1=4e-4,
                      current_factor=1.2, Ib=None, padding=8.0,
                      resistive_wall=(1.4e6, 0.0)):
    """Large production workload: the kick-drift maps with second-order
    momentum compaction (alpha2 = 5 alpha1, three drift terms of
    src/main.cpp:663-664), a five-point Fokker-Planck stencil, 4000 steps
    per synchrotron period, and a wake impedance of parallel-plates CSR
    plus the resistive wall (conductivity, permeability parameter) of
    src/Z/ResistiveWall.cpp.  Returns (desc, derived, resistive_wall) for
    impedances(derived, resistive_wall=...)."""
    ring = dict(grid_size=n, steps_per_ts=steps_per_ts, vacuum_gap=gap,
                padding=padding, alpha1=alpha1, alpha2=5.0 * alpha1)
    base = derive(**ring)
    d = derive(**ring, bunch_current=Ib if Ib is not None
               else current_factor * base["Ith"])
    desc = dict(n=n, qmin=d["qmin"], qmax=d["qmax"], pmin=d["pmin"],
                pmax=d["pmax"], qscale=d["qscale"], pscale=d["pscale"],
                scheme=0, interp=4, rot_clamp=0, angle=d["angle"],
                rf_linear=0, rf_angle=d["angle"],
                revolutionpart=d["revolutionpart"], V_RF=d["V_eff"],
                f_RF=d["f_RF"], V0=d["V0"], slip=d["slip"],
                E0=d["ring"]["beam_energy"], use_fp=1, fptype=3, e1=d["e1"],
                deriv=5, use_wake=1, spacing_bins=d["spacing_bins"],
                wakescaling=d["wakescaling"], use_csr=1,
                fc=d["ring"]["cutoff_frequency"])
    return desc, d, resistive_wall


def fp_diffusion_number(n, e1, pq_size=12.0):
    """e1 / dp^2 of the FokkerPlanckMap diffusion stencil
    (src/SM/FokkerPlanckMap.cpp:38-118) on an n-point momentum axis of
    pq_size: the explicit update stays bounded only up to 1/2."""
    dp = pq_size / (n - 1)
    return float(e1) / (dp * dp)


def rotation_config(n=256, steps_per_ts=1000, damping_steps=1.0e4,
                    current_factor=0.5, Ib=None, padding=8.0):
    """Plain VFP, rotation scheme: WakePotentialMap, RotationMap (cubic,
    clamped, precomputed semantics), FokkerPlanckMap, free-space CSR.
    The damping time is damping_steps time steps: 1000 steps gives
    e1 = 2e-3, whose diffusion number e1/dp^2 exceeds 1/2 from n = 190 on
    (0.90 at n = 256), where the explicit Fokker-Planck update of the
    reference diverges; 1e4 steps keeps it at 0.09 (n = 256)."""
    base = derive(grid_size=n, steps_per_ts=steps_per_ts, vacuum_gap=-1.0,
                  padding=padding)
    dt = base["dt"]
    if Ib is None:
        Ib = current_factor * base["Ith"]
    d = derive(grid_size=n, steps_per_ts=steps_per_ts, vacuum_gap=-1.0,
               padding=padding, damping_time=damping_steps * dt,
               bunch_current=Ib)
    desc = dict(n=n, qmin=d["qmin"], qmax=d["qmax"], pmin=d["pmin"],
                pmax=d["pmax"], qscale=d["qscale"], pscale=d["pscale"],
                scheme=1, interp=4, rot_clamp=1, angle=d["angle"],
                rf_linear=1, rf_angle=d["angle"],
                revolutionpart=d["revolutionpart"], V_RF=d["V_eff"],
                f_RF=d["f_RF"], V0=d["V0"], slip=d["slip"],
                E0=d["ring"]["beam_energy"], use_fp=1, fptype=3, e1=d["e1"],
                deriv=4, use_wake=1, spacing_bins=d["spacing_bins"],
                wakescaling=d["wakescaling"], use_csr=1,
                fc=d["ring"]["cutoff_frequency"])
    return desc, d


def impedances(d, dev=None, resistive_wall=None):
    """Wake and CSR impedance tables of src/main.cpp:717-727 on the device:
    (wake, csr)."""
    from .vfps import makeImpedance
    ring = d["ring"]
    gap = ring["vacuum_gap"]
    s, xi = (resistive_wall or (0.0, 0.0))
    zw = makeImpedance(d["padded_bins"], d["fmax"], d["R_bend"],
                       ring["f_rev"], gap, True, s, xi, dev=dev)
    zc = makeImpedance(d["padded_bins"], d["fmax"], d["R_bend"],
                       ring["f_rev"], gap if gap > 0 else -1.0, dev=dev)
    return zw, zc


def scan_instance(n, factor, gap, damping_steps, steps_per_ts, scheme="kick_drift"):
    """One run of a scan: current factor x the scaling-law threshold,
    parallel-plates CSR with full gap `gap` (m), damping time damping_steps
    time steps.  scheme "kick_drift": the maps of the plain 512^2 workload
    (wake kick, sinusoidal RF kick, linear drift, cubic Fokker-Planck;
    src/main.cpp:1034-1041); "rotation": wake kick, cubic clamped
    RotationMap, Fokker-Planck.  Returns (desc, derived)."""
    base = derive(grid_size=n, steps_per_ts=steps_per_ts, vacuum_gap=gap)
    d = derive(grid_size=n, steps_per_ts=steps_per_ts, vacuum_gap=gap,
               damping_time=damping_steps * base["dt"],
               bunch_current=factor * base["Ith"])
    kd = scheme == "kick_drift"
    if not kd and scheme != "rotation":
        raise ValueError(f"scheme must be kick_drift or rotation: {scheme}")
    desc = dict(n=n, qmin=d["qmin"], qmax=d["qmax"], pmin=d["pmin"],
                pmax=d["pmax"], qscale=d["qscale"], pscale=d["pscale"],
                scheme=0 if kd else 1, interp=4, rot_clamp=0 if kd else 1,
                angle=d["angle"], rf_linear=0 if kd else 1,
                rf_angle=d["angle"], revolutionpart=d["revolutionpart"],
                V_RF=d["V_eff"], f_RF=d["f_RF"], V0=d["V0"], slip=d["slip"],
                E0=d["ring"]["beam_energy"], use_fp=1, fptype=3, e1=d["e1"],
                deriv=4, use_wake=1, spacing_bins=d["spacing_bins"],
                wakescaling=d["wakescaling"], use_csr=1,
                fc=d["ring"]["cutoff_frequency"])
    return desc, d


def batched_scan_config(sizes=(128, 256, 384, 512), per_size=16, seed=3,
                        current_range=(0.3, 3.0), gap_range=(0.016, 0.064),
                        steps_per_ts=2000, damping_steps=1.0e4,
                        scheme="kick_drift"):
    """Independent instances of uneven mesh sizes for a threshold scan:
    per_size instances of each size (sizes interleaved), currents
    log-spaced over current_range times each instance's scaling-law
    threshold, parallel-plates gaps drawn uniformly from gap_range (m) with
    the given seed, each instance with its own parallel-plates CSR
    impedance, the maps of `scheme` (see scan_instance) and a damping time
    of damping_steps steps (the default keeps the Fokker-Planck diffusion
    number below 1/2 up to n = 512; see rotation_config).  Returns a list
    of (desc, derived, init) per instance."""
    k = len(sizes) * per_size
    rng = np.random.Generator(np.random.Philox(seed))
    gaps = rng.uniform(gap_range[0], gap_range[1], k)
    factors = np.geomspace(current_range[0], current_range[1], k)
    out = []
    for i in range(k):
        n = sizes[i % len(sizes)]
        desc, d = scan_instance(n, float(factors[i]), float(gaps[i]),
                                damping_steps, steps_per_ts, scheme)
        out.append((desc, d, initial_mesh(n, seed=seed * 1000 + i)))
    return out
