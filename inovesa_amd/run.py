"""The reference program on the GPU: from the reference's options
(configuration file and command line, src/IO/ProgramOptions.cpp) through
the setup of src/main.cpp:163-840 to the time loop (:948-1053) and its
results file (:840-997, 1056-1100).

    python -m inovesa_amd.run -c run.cfg -o result.h5 [--name value ...]

One bunch with static RF runs as the fused step (one kernel per step,
replayed from captured graphs).  Several bunches, RF noise or modulation
(DynamicRFKickMap) and the mesh-reading particle tracking
(FPTrack = 2) run the reference's loop map by map with this library's map
kernels (the per-map path).
"""
import math
import os
import sys
import time

import numpy as np

from . import vfps as V
from .options import ProgramOptions

C_LIGHT = 2.99792458e8


def upper_power_of_two(v):
    """src/HelperFunctions.cpp:178-189"""
    v = int(v) - 1
    for s in (1, 2, 4, 8, 16, 32):
        v |= v >> s
    return v + 1


def read_impedance_file(path, nfreqs):
    """Impedance::readData (src/Z/Impedance.cpp:93-111): lines of `index
    real imag`, a repeated index read once; zero-padded (or cut) to nfreqs
    entries, the size the factory adds it to."""
    rows = []
    last = None
    with open(path) as f:
        for line in f:
            parts = line.split()
            if len(parts) < 3:
                continue
            idx = int(parts[0])
            if idx != last:
                rows.append(complex(np.float32(float(parts[1])),
                                    np.float32(float(parts[2]))))
            last = idx
    z = np.zeros(nfreqs, np.complex64)
    k = min(len(rows), nfreqs)
    z[:k] = rows[:k]
    return z


class Setup:
    """The derived quantities and objects of src/main.cpp:163-840."""

    def __init__(self, opts, dev=None, log=None):
        self.opts = o = opts
        self.log = log or (lambda m: None)
        self.dev = dev or V.Device.default()
        buckets, currents = o.bunches()
        if not currents:
            raise ValueError("BunchCurrent: no filled bucket")
        self.buckets, self.currents = buckets, currents
        self.nbuckets = len(o["BunchCurrent"])
        self.nb = len(currents)
        self.d = d = V.derive(**o.ring())
        self.n = n = o["GridSize"]
        self.Ib = d["Ib"]
        # filling pattern normalised (src/main.cpp:275-279)
        self.filling = [c / sum(currents) for c in currents]
        # src/main.cpp:304-321
        pq = np.float32(o["PhaseSpaceSize"])
        spacing_ps = (1.0 / d["f_RF"]) * C_LIGHT / d["qscale"] / float(pq)
        self.spacing_bins = int(round(n * spacing_ps))
        spaced = math.ceil(n * self.nbuckets * spacing_ps)
        if o["RoundPadding"]:
            spaced = upper_power_of_two(spaced)
        self.padded_bins = d["padded_bins"]
        self.spaced_bins = spaced
        self.wake_bins = spaced if self.nbuckets > 1 else self.padded_bins
        gap = o["VacuumGap"]
        table = (read_impedance_file(o["Impedance"], self.wake_bins)
                 if o["Impedance"] else None)
        # src/main.cpp:717-727
        self.zwake = V.makeImpedance(self.wake_bins, d["fmax"], d["R_bend"],
                                     d["ring"]["f_rev"], gap, o["UseCSR"],
                                     o["WallConductivity"],
                                     o["WallSusceptibility"],
                                     o["CollimatorRadius"], table, self.dev)
        self.zrad = V.makeImpedance(self.padded_bins, d["fmax"], d["R_bend"],
                                    d["ring"]["f_rev"],
                                    gap if gap > 0 else -1.0, dev=self.dev)
        self.rf_noise = o.rf_noise(d["dt"])
        ph, am, mod_a, mod_s = self.rf_noise
        # src/main.cpp:595-598: a dynamic RF map when any noise or the
        # modulation (both its amplitude and frequency) is set
        self.dynamic_rf = (ph > 0 or am > 0 or (mod_a > 0 and mod_s != 0))
        self.e1 = d["e1"]
        # the information the reference prints and saves (:371-378): its
        # shielding parameter and CSR strength are only set for gap == 0
        self.shield, self.S_csr = 0.0, 0.0
        if gap == 0:
            self.S_csr = self.Ib / d["Inorm"]

    # -- the initial distribution (src/main.cpp:466-535) -------------------
    def initial_mesh(self):
        o, d, n = self.opts, self.d, self.n
        path = o["InitialDistFile"]
        if path:
            ext = os.path.splitext(path)[1].lower()
            if ext in (".h5", ".hdf5"):
                from .hdf5 import readPhaseSpace
                mesh, _ = readPhaseSpace(path, o["InitialDistStep"])
                if mesh.shape[1] != n:
                    raise ValueError(f"{path}: mesh of {mesh.shape[1]}, "
                                     f"GridSize {n}")
            else:
                raise ValueError(f"{path}: only HDF5 result files are read "
                                 "as initial distributions")
            ps = self.phasespace(data=mesh[: self.nb])
        else:
            ps = self.phasespace(zoom=o["InitialDistZoom"])
        # an initial renormalisation (:528-532)
        if o["RenormalizeCharge"] >= 0:
            ps.updateXProjection()
            ps.integrate()
            ps.normalize()
        return ps

    def phasespace(self, data=None, zoom=1.0):
        d = self.d
        return V.PhaseSpace(self.n, d["qmin"], d["qmax"], d["qscale"],
                            d["pmin"], d["pmax"], d["pscale"], dev=self.dev,
                            charge=self.Ib / d["ring"]["f_rev"],
                            current=self.Ib, filling=self.filling, zoom=zoom,
                            data=data)

    def step_desc(self):
        """The fused step of the reference's loop: wake kick, RF kick,
        drift, Fokker-Planck (src/main.cpp:1034-1041)."""
        o, d = self.opts, self.d
        return dict(n=self.n, qmin=d["qmin"], qmax=d["qmax"], pmin=d["pmin"],
                    pmax=d["pmax"], qscale=d["qscale"], pscale=d["pscale"],
                    scheme=0, interp=o["InterpolationPoints"],
                    rot_clamp=int(o["InterpolateClamped"]),
                    angle=d["angle"], rf_linear=int(o["LinearRF"]),
                    rf_angle=d["angle"], revolutionpart=d["revolutionpart"],
                    V_RF=d["V_eff"], f_RF=d["f_RF"], V0=d["V0"],
                    slip=d["slip"], E0=o["BeamEnergy"],
                    use_fp=int(self.e1 > 0), fptype=o["FPType"], e1=self.e1,
                    deriv=o["derivation"],
                    use_wake=int(self.zwake is not None),
                    spacing_bins=self.spacing_bins,
                    wakescaling=d["wakescaling"], use_csr=1,
                    fc=o["CutoffFreq"],
                    renormalize=max(0, o["RenormalizeCharge"]))

    def tracking(self, ps):
        """Particle start positions (src/main.cpp:779-799): lines of q p,
        converted to grid coordinates x(q), y(p)."""
        path = self.opts["tracking"]
        if not path:
            return np.zeros((0, 2), np.float32)
        qp = np.loadtxt(path, dtype=np.float64, ndmin=2)
        d = self.d
        dq = (np.float32(d["qmax"]) - np.float32(d["qmin"])) / np.float32(self.n - 1)
        dp = (np.float32(d["pmax"]) - np.float32(d["pmin"])) / np.float32(self.n - 1)
        x = (qp[:, 0].astype(np.float32) - np.float32(d["qmin"])) / dq
        y = (qp[:, 1].astype(np.float32) - np.float32(d["pmin"])) / dp
        return np.stack([x, y], axis=1).astype(np.float32)


class FusedRun:
    """One bunch: the fused step, with the results file of StepWriter."""

    def __init__(self, setup):
        from .hdf5 import StepWriter
        self.s = s = setup
        ps = s.initial_mesh()
        init = ps.getData()[0]
        self.step = V.Step(s.step_desc(), init, s.zwake, s.zrad, dev=s.dev)
        self.laststep = s.d["laststep"]
        self.step.record(0)
        self.particles = s.tracking(ps)
        if len(self.particles):
            self.step.set_particles(self.particles, s.opts["FPTrack"])
        self.writer = None
        out = s.opts["output"]
        if out and os.path.splitext(out)[1].lower() in (".h5", ".hdf5"):
            self.writer = StepWriter(out, self.step, s.step_desc(), s.d,
                                     s.zwake, s.zrad, s.opts["outstep"],
                                     s.opts["SavePhaseSpace"], dev=s.dev)
            s.opts.save_h5(self.writer.file)
            self.writer.file.addParameterToGroup("/Info", "CSRStrength",
                                                 s.S_csr)

    def run(self, steps=None):
        steps = self.laststep if steps is None else steps
        if self.writer:
            self.writer.run(steps)
            self.writer.finish()
        else:
            self.step.run(steps)
        return self.step


class MapLoop:
    """The loop of src/main.cpp:948-1053 map by map (several bunches, a
    dynamic RF kick, tracking that reads the mesh): three phase spaces and
    the maps of :595-770 on the device, driven from the host."""

    def __init__(self, setup, seed=0):
        s = self.s = setup
        o, d = s.opts, s.d
        it, clamp = o["InterpolationPoints"], o["InterpolateClamped"]
        self.t1 = s.initial_mesh()
        self.t2 = self.t1.clone()
        self.t3 = self.t1.clone()
        self.laststep = d["laststep"]
        self.steps = d["steps"]
        self.rdtn = V.ElectricField(self.t1, s.zrad, s.buckets, 0,
                                    d["ring"]["f_rev"],
                                    np.float32(d["revolutionpart"]))
        self.wfield = None
        if s.zwake is not None:
            self.wfield = V.ElectricField(self.t1, s.zwake, s.buckets,
                                          s.spacing_bins, d["ring"]["f_rev"],
                                          np.float32(d["revolutionpart"]),
                                          d["wakescaling"])
            self.wm = V.WakePotentialMap(self.t1, self.t2, self.wfield, it,
                                         clamp)
        else:
            self.wm = V.Identity(self.t1, self.t2)
        ph, am, mod_a, mod_s = s.rf_noise
        self.drfm = None
        if s.dynamic_rf:
            kw = dict(phasespread=ph, amplspread=am, modampl=mod_a,
                      modtimeincrement=mod_s, steps=self.laststep + 1,
                      seed=seed, it=it, clamp=clamp)
            if o["LinearRF"]:
                self.rfm = V.DynamicRFKickMap(self.t2, self.t1,
                                              d["revolutionpart"], d["f_RF"],
                                              angle=d["angle"], **kw)
            else:
                self.rfm = V.DynamicRFKickMap(self.t2, self.t1,
                                              d["revolutionpart"], d["f_RF"],
                                              V_RF=d["V_eff"], V0=d["V0"],
                                              **kw)
            self.drfm = self.rfm
        elif o["LinearRF"]:
            self.rfm = V.RFKickMap(self.t2, self.t1, angle=d["angle"],
                                   f_RF=d["f_RF"], it=it, clamp=clamp)
        else:
            self.rfm = V.RFKickMap(self.t2, self.t1, f_RF=d["f_RF"], it=it,
                                   clamp=clamp,
                                   revolutionpart=d["revolutionpart"],
                                   V_RF=d["V_eff"], V0=d["V0"])
        self.drm = V.DriftMap(self.t1, self.t3, np.array(d["slip"],
                                                         np.float32),
                              o["BeamEnergy"], it, clamp)
        if s.e1 > 0:
            self.fpm = V.FokkerPlanckMap(self.t3, self.t1, o["FPType"],
                                         o["FPTrack"], s.e1, o["derivation"])
        else:
            self.fpm = V.Identity(self.t3, self.t1)
        self.particles = s.tracking(self.t1)
        self.seed = seed
        self.simulationstep = 0
        self.outstepnr = 0
        self.file = None
        out = o["output"]
        if out and os.path.splitext(out)[1].lower() in (".h5", ".hdf5"):
            self.file = self._open(out)

    def _open(self, path):
        from .hdf5 import HDF5File
        s, d = self.s, self.d_
        n = s.n
        ax = self.t1.getAxis(0)
        dq = self.t1.getDelta(0)
        faxis = np.float32(0) + np.arange(s.padded_bins, dtype=np.float32) * (
            (np.float32(1) / dq) / np.float32(s.padded_bins - 1))
        meter = float(np.float32(d["qscale"]))
        hertz = C_LIGHT / meter
        f_rev = float(np.float32(d["ring"]["f_rev"]))
        wph = 2 * 1.0 * s.Ib * s.Ib / f_rev
        f = HDF5File(path, s.nb, n, ax, meter, float(np.float32(d["pscale"])),
                     1.0 / d["f_s"], f_rev, s.Ib, s.Ib / f_rev,
                     buckets=s.buckets, maxn=s.padded_bins // 2,
                     axis_f=faxis[: s.padded_bins // 2], hertz=hertz,
                     volts=float(self.t1.getDelta(1)) * float(np.float32(d["pscale"]))
                     / float(np.float32(d["revolutionpart"])),
                     factor4WattPerHertz=wph, factor4Watts=wph * hertz,
                     impedance=s.zwake.data() if s.zwake is not None else None,
                     nparticles=len(self.particles))
        s.opts.save_h5(f)
        f.addParameterToGroup("/Info", "CSRStrength", s.S_csr)
        f.addParameterToGroup("/Info", "ShieldingParameter", s.shield)
        if self.wfield is not None:
            self.wfield.wakePotential()
            f.appendPadded(self.wfield.getPaddedBunchProfiles(),
                           self.wfield.getPaddedWakePotential())
        if s.opts["SavePhaseSpace"] == 0:
            f.append(0.0, f.PhaseSpace, mesh=self.t1.getData())
        return f

    @property
    def d_(self):
        return self.s.d

    def _output(self, final=False):
        f, t1 = self.file, self.t1
        h5save = self.s.opts["SavePhaseSpace"]
        at = (f.All if final or (h5save > 0 and
                                 self.outstepnr % h5save == 0) else f.Defaults)
        fill, _ = t1.integrate()
        mx, _, rx = t1.variance(0)
        t1.updateYProjection()
        my, _, ry = t1.variance(1)
        f.append(self.simulationstep / self.steps, at,
                 mesh=t1.getData() if at == f.All else None,
                 xproj=t1.getProjection(0), yproj=t1.getProjection(1),
                 length=rx, position=mx, spread=ry, average=my,
                 population=fill.astype(np.float32))
        power, spec = self.rdtn.updateCSR(np.float32(self.s.opts["CutoffFreq"]))
        f.append_field(power, spec[:, : f.maxn])
        if self.wfield is not None:
            f.append_wake(self.wm.getForce().reshape(self.s.nb, self.s.n))
        if len(self.particles):
            f.appendTracks(self._physical(self.particles))
        if self.drfm is not None:
            f.appendRFKicks(self.drfm.getPastModulation())
        if final and self.wfield is not None:
            f.appendPadded(self.wfield.getPaddedBunchProfiles(),
                           self.wfield.getPaddedWakePotential())
        self.outstepnr += 1

    def _physical(self, p):
        d = self.d_
        n = self.s.n
        dq = (np.float32(d["qmax"]) - np.float32(d["qmin"])) / np.float32(n - 1)
        dp = (np.float32(d["pmax"]) - np.float32(d["pmin"])) / np.float32(n - 1)
        return np.stack([np.float32(d["qmin"]) + p[:, 0] * dq,
                         np.float32(d["pmin"]) + p[:, 1] * dp], axis=1)

    def step(self):
        """One pass of the loop body (src/main.cpp:949-1049)."""
        o = self.s.opts
        ren = o["RenormalizeCharge"]
        if self.wfield is not None:
            self.wm.update()
        if ren > 0 and self.simulationstep % ren == 0:
            self.t1.integrateAndNormalize()
        else:
            self.t1.integrate()
        outstep = o["outstep"]
        if self.file and outstep > 0 and self.simulationstep % outstep == 0:
            self._output()
        p = self.particles
        for m in (self.wm, self.rfm, self.drm, self.fpm):
            m.apply()
            if len(p):
                p = m.applyToAll(p, seed=self.seed)
        self.particles = p
        self.t1.updateXProjection()
        self.simulationstep += 1

    def run(self, steps=None):
        steps = self.laststep if steps is None else steps
        for _ in range(steps):
            self.step()
        if self.file:
            if self.wfield is not None:
                self.wm.update()
            ren = self.s.opts["RenormalizeCharge"]
            if ren > 0 and self.simulationstep % ren == 0:
                self.t1.integrateAndNormalize()
            else:
                self.t1.integrate()
            self._output(final=True)
            self.file.close()
        return self


def simulation(opts, dev=None, log=None, seed=0):
    """The run the reference program makes of these options."""
    s = Setup(opts, dev, log)
    if (s.nb == 1 and not s.dynamic_rf and
            not (opts["tracking"] and opts["FPTrack"] == 2)):
        return FusedRun(s)
    return MapLoop(s, seed=seed)


def main(argv=None):
    argv = sys.argv[1:] if argv is None else argv
    opts = ProgramOptions().parse(argv)
    log = (lambda m: print(m, flush=True))
    sim = simulation(opts, log=log)
    kind = "fused step" if isinstance(sim, FusedRun) else "per-map loop"
    log(f"{kind}: {opts['GridSize']}^2 mesh, {sim.laststep} steps, "
        f"{len(opts.bunches()[1])} bunch(es)")
    if opts["output"]:
        opts.save(opts["output"] + ".cfg")
    t0 = time.perf_counter()
    sim.run()
    dt = time.perf_counter() - t0
    log(f"done: {sim.laststep} steps in {dt:.2f} s")
    return 0


if __name__ == "__main__":
    sys.exit(main())
