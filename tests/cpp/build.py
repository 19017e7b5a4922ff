"""Build the C++ test program of the reference-shaped interface
(include/inovesa_amd.hpp) against the in-tree libvfps.so with g++."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
SRC = os.path.join(HERE, "test_reference_api.cpp")
BIN = os.path.join(HERE, "test_reference_api.bin")


def command(out=BIN, syntax_only=False):
    cmd = ["g++", "-std=c++14", "-O2", "-Wall", "-Wextra",
           "-I", os.path.join(ROOT, "include"), SRC]
    if syntax_only:
        return cmd + ["-fsyntax-only"]
    lib = os.path.join(ROOT, "inovesa_amd", "lib")
    return cmd + ["-L", lib, "-lvfps", "-Wl,-rpath,$ORIGIN/../../inovesa_amd/lib",
                  "-Wl,-rpath-link,/opt/rocm/lib", "-o", out]


def build(out=BIN):
    subprocess.run(command(out), check=True)
    return out


if __name__ == "__main__":
    print(build())
