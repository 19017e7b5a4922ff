"""Build recipe for the CPU oracle (test infrastructure).

Compiles oracle/vfps_oracle.c twice with gcc:
  oracle/liboracle_f32.so  -- real = float  (parity target)
  oracle/liboracle_f64.so  -- real = double (FP64 drift oracle)
and oracle/dynrf_oracle.cpp with g++:
  oracle/liboracle_dynrf.so -- the DynamicRFKickMap modulation queue

-ffp-contract=off keeps every rounding where the reference source puts it.
-march=x86-64-v3 (not -march=native) so the library also runs on the host
CPU of another machine.

The reference itself (Inovesa) cannot be compiled here: its hot-path
translation units include fftw3.h and boost/multi_array.hpp, which this
image does not provide, so there is no oracle/_ref build.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "vfps_oracle.c")
FLAGS = ["-O3", "-march=x86-64-v3", "-ffp-contract=off", "-fPIC", "-shared",
         "-std=c11", "-Wall", "-Wno-unused-function"]


def build(verbose=False):
    outs = []
    for name, defs in (("liboracle_f32.so", []),
                       ("liboracle_f64.so", ["-DVFPS_ORACLE_DOUBLE"])):
        out = os.path.join(HERE, name)
        tmp = out + ".tmp"
        cmd = ["gcc"] + FLAGS + defs + [SRC, "-o", tmp, "-lm"]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        os.replace(tmp, out)
        outs.append(out)
    # the DynamicRFKickMap modulation queue (C++: the reference's generators
    # are the standard library's)
    out = os.path.join(HERE, "liboracle_dynrf.so")
    tmp = out + ".tmp"
    cmd = ["g++", "-O2", "-std=c++14", "-fPIC", "-shared", "-ffp-contract=off",
           "-Wall", os.path.join(HERE, "dynrf_oracle.cpp"), "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, out)
    outs.append(out)
    return outs


def build_native(outdir=None, verbose=False):
    """The FP32 restatement built with the reference's release flags
    (CMakeLists.txt:22,164: -O3 -march=native, the compiler's default
    contraction of a*b+c) on the machine that times it: the CPU baseline of
    the benchmark.  Not the parity target (that one keeps every rounding,
    -ffp-contract=off).  Returns the library path."""
    outdir = outdir or os.path.join(HERE, "_native")
    os.makedirs(outdir, exist_ok=True)
    out = os.path.join(outdir, "liboracle_f32_native.so")
    tmp = out + ".tmp%d" % os.getpid()
    cmd = ["gcc", "-O3", "-march=native", "-fPIC", "-shared", "-std=gnu11",
           "-w", SRC, "-o", tmp, "-lm"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, out)
    return out


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
