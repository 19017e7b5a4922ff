"""Writes the golden fixtures under tests/golden/.

reference_vectors.json
    Input and expected output meshes of the reference's own unit tests
    (test/TestKickMap.cpp:18-172, test/TestRotationMap.cpp:18-205,
    test/TestRuler.cpp:13-28, test/TestHelperFunctions.cpp:73-77,
    test/TestImpedance.cpp:34-86, test/TestConstImpedance.cpp:15-27,
    test/impedance.dat), transcribed as data.
airy_mpmath.json
    Ai, Ai', Bi, Bi' from mpmath (50 digits) on x in [0, 110].
fft_numpy.json
    Real FFTs from numpy.fft (pocketfft) of seeded vectors, n = 8..64.
pp_mpmath.json
    ParallelPlatesCSR entries computed with mpmath's Airy functions and the
    summation rule of src/Z/ParallelPlatesCSR.cpp:38-67 (terms stop where
    Bi or Bi' exceed the double range), for a small table.
"""
import json
import math
import os

import mpmath
import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def reference_vectors():
    v = {}
    # test/TestKickMap.cpp: axes -1..1 (4 points), linear interpolation
    v["kickmap_x"] = dict(
        axis=0, it=2, offset=[-1, 0.0, 1.0, 1.0],
        data_in=[0.0, 0.1, 0.0, 1.0, 0.0, 0.3, 0.7, 0.0,
                 0.1, 0.7, 0.3, 0.0, 0.9, 0.0, 0.0, 0.0],
        data_out=[0.0, 0.1, 0.7, 0.0, 0.0, 0.3, 0.3, 0.0,
                  0.0, 0.7, 0.0, 0.0, 0.1, 0.0, 0.0, 0.0])
    v["kickmap_x_overflow"] = dict(
        axis=0, it=2, offset=[-1, 0.0, 5.0, 5.0],
        data_in=v["kickmap_x"]["data_in"],
        data_out=[0.0, 0.1, 0.0, 0.0, 0.0, 0.3, 0.0, 0.0,
                  0.0, 0.7, 0.0, 0.0, 0.1, 0.0, 0.0, 0.0])
    v["kickmap_y"] = dict(
        axis=1, it=2, offset=[-1, 0.0, 1.0, 1.0],
        data_in=[0.0, 0.1, 0.0, 0.3, 0.0, 0.3, 0.7, 0.0,
                 0.1, 0.7, 0.3, 0.0, 0.5, 0.0, 0.6, 0.0],
        data_out=[0.0, 0.0, 0.1, 0.0, 0.0, 0.3, 0.7, 0.0,
                  0.7, 0.3, 0.0, 0.0, 0.0, 0.6, 0.0, 0.0])
    hp = float(np.float32(math.pi / 2))
    pi = float(np.float32(math.pi))
    v["rotation_linear_noclamp_map"] = dict(
        qmin=-1, qmax=1, pmin=-1, pmax=1, angle=hp, it=2, clamp=False,
        rotmapsize=0,
        data_in=[0.0, 0.0, 0.0, 0.0, 0.0, 0.3, 0.5, 0.0,
                 0.0, 0.7, 0.3, 0.0, 0.0, 0.0, 0.0, 0.0],
        data_out=[0.0, 0.0, 0.0, 0.0, 0.0, 0.5, 0.3, 0.0,
                  0.0, 0.3, 0.7, 0.0, 0.0, 0.0, 0.0, 0.0])
    v["rotation_linear_offcenter"] = dict(
        qmin=-2, qmax=1, pmin=-2, pmax=2, angle=pi, it=2, clamp=False,
        rotmapsize=0,
        data_in=[0.0, 0.0, 0.0, 0.0, 0.0, 0.1, 0.2, 0.0,
                 0.0, 0.3, 0.4, 0.0, 0.0, 0.0, 0.0, 0.0],
        data_out=[0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0,
                  0.0, 0.4, 0.3, 0.0, 0.0, 0.2, 0.1, 0.0])
    sym = [0.0, 0.0, 0.0, 0.0, 0.0, 0.3, 0.7, 0.0,
           0.0, 0.7, 0.3, 0.0, 0.0, 0.0, 0.0, 0.0]
    sym_rot = [0.0, 0.0, 0.0, 0.0, 0.0, 0.7, 0.3, 0.0,
               0.0, 0.3, 0.7, 0.0, 0.0, 0.0, 0.0, 0.0]
    v["rotation_quadratic_noclamp_nomap"] = dict(
        qmin=-1, qmax=1, pmin=-1, pmax=1, angle=hp, it=3, clamp=False,
        rotmapsize=0, data_in=sym, data_out=sym_rot)
    v["rotation_cubic_clamp_map"] = dict(
        qmin=-1, qmax=1, pmin=-1, pmax=1, angle=hp, it=4, clamp=True,
        rotmapsize=16, data_in=sym, data_out=sym_rot)
    v["rotation_throws"] = [dict(it=3, clamp=True, rotmapsize=0),
                            dict(it=4, clamp=True, rotmapsize=0)]
    v["ruler"] = dict(steps=15, min=-4, max=4, zerobin=7, length=8)
    v["upper_power_of_two"] = [[7, 8], [8, 8], [9, 16]]
    # test/TestImpedance.cpp:34-46 with test/impedance.dat
    v["impedance_file"] = dict(
        lines=[[1, -2, 2.5], [2, -1, 1.5], [3, 0, 0.5], [4, 1, -0.5],
               [5, 2, -1.5], [6, 3, -2.5]],
        expect=[[i - 2, 2.5 - i] for i in range(6)])
    # test/TestConstImpedance.cpp:15-27: n = 6, Z = 1e3
    v["const_impedance"] = dict(n=6, z=1e3,
                                expect=[1e3, 1e3, 1e3, 0, 0, 0])
    # test/TestImpedance.cpp:48-86: sums of two tables
    v["impedance_addition"] = dict(
        z1=[[i, -i + 4] for i in range(8)], z2=[[i - 2, 3] for i in range(8)])
    # test/TestPhaseSpace.cpp:38-107: Gaussian on -12..12, n = 32
    v["phasespace_gaussian"] = dict(n=32, min=-12, max=12, integral=1.0,
                                    delta=24 / 31.0, sigma=1.0, rel_tol=1e-3)
    return v


def airy_golden():
    mpmath.mp.dps = 50
    xs = sorted(set([0.0, 1e-9, 1e-6, 1e-3, 0.05, 0.1, 0.25, 0.3, 0.5, 0.7,
                     1.0, 1.3, 1.9, 2.0, 2.5, 3.0, 4.0, 5.0, 6.1, 7.5, 8.0,
                     9.9, 11.9, 11.99, 12.0, 12.01, 13.0, 15.0, 20.0, 30.0,
                     45.0, 60.0, 80.0, 100.0, 104.0, 104.2]))
    rows = []
    for x in xs:
        mx = mpmath.mpf(x)
        vals = [mpmath.airyai(mx), mpmath.airyai(mx, derivative=1),
                mpmath.airybi(mx), mpmath.airybi(mx, derivative=1)]
        rows.append([repr(x)] + [mpmath.nstr(v, 20) for v in vals])
    overflow = []
    for x in (104.25, 104.3, 105.0, 110.0):
        bip = mpmath.airybi(mpmath.mpf(x), derivative=1)
        overflow.append([x, bool(bip > mpmath.mpf(np.finfo(np.float64).max))])
    return dict(rows=rows, overflow=overflow)


def fft_golden():
    rng = np.random.default_rng(11)
    out = []
    for n in (8, 16, 32, 64):
        x = rng.standard_normal(n)
        f = np.fft.rfft(x)
        out.append(dict(n=n, x=x.tolist(), re=f.real.tolist(),
                        im=f.imag.tolist()))
    return out


def pp_golden():
    """ParallelPlatesCSR for nfreqs=64 with f0, f_max, g chosen so that the
    Airy sums reach the overflow cut."""
    mpmath.mp.dps = 40
    nfreqs = 64
    f0 = float(np.float32(9e6))
    f_max = float(np.float32(5e11))
    g = 0.02
    c = 2.99792458e8
    pi = mpmath.pi
    delta = float(np.float32(f_max / f0)) / (nfreqs - 1.0)
    r_bend = c / (2 * math.pi * f0)
    z0 = 1.0 / (8.854187817e-12 * c)
    dmax = mpmath.mpf(np.finfo(np.float64).max)
    rows = []
    for i in range(1, nfreqs // 2 + 1):
        n = i * delta
        m = n * (g / r_bend) ** 1.5
        maxp = int(2 * m * (r_bend / g) ** 1.5 * f0 * g / c)
        b = m ** (-4.0 / 3.0)
        zr = mpmath.mpf(0)
        zi = mpmath.mpf(0)
        for p in range(1, maxp + 1, 2):
            u = mpmath.mpf(math.pi ** 2 * p * p / 2 ** (2.0 / 3.0) * b)
            bi = mpmath.airybi(u)
            bip = mpmath.airybi(u, derivative=1)
            if bi > dmax or bip > dmax:
                break
            ai = mpmath.airyai(u)
            aip = mpmath.airyai(u, derivative=1)
            zr += aip * aip + u * ai * ai
            zi += -aip * bip - u * ai * bi
        sc = 4.0 * b * n * g / r_bend * math.pi ** 2 * 2 ** (1.0 / 3.0) * z0
        rows.append([i, float(zr * sc), float(zi * sc)])
    return dict(nfreqs=nfreqs, f0=f0, f_max=f_max, g=g, rows=rows)


def main():
    with open(os.path.join(HERE, "reference_vectors.json"), "w") as f:
        json.dump(reference_vectors(), f, indent=1)
    with open(os.path.join(HERE, "airy_mpmath.json"), "w") as f:
        json.dump(airy_golden(), f, indent=1)
    with open(os.path.join(HERE, "fft_numpy.json"), "w") as f:
        json.dump(fft_golden(), f)
    with open(os.path.join(HERE, "pp_mpmath.json"), "w") as f:
        json.dump(pp_golden(), f, indent=1)


if __name__ == "__main__":
    main()
