"""Build recipe for the CPU oracle (test infrastructure).

Compiles oracle/vfps_oracle.c twice with gcc:
  oracle/liboracle_f32.so  -- real = float  (parity target)
  oracle/liboracle_f64.so  -- real = double (FP64 drift oracle)

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
    return outs


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
