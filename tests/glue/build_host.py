"""Builds the test host around the reference-side binding.

tandav_amd/glue/tandav_glue.c + tests/glue/host.c -> tests/glue/_build/
libtandav_host{,_periodic}.so, linked against tandav_amd/libtandav_gpu.so
(found at run time through an $ORIGIN-relative rpath).  gcc, C99 + OpenMP,
the flags the reference is built with (-std=gnu99 -O2 -fopenmp
-ffp-contract=off).  The periodic variant also has GRAVITY_FORCETEST.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
OUT = os.path.join(HERE, "_build")
VARIANTS = {"libtandav_host.so": [],
            "libtandav_host_periodic.so": ["-DPERIODIC",
                                           "-DGRAVITY_FORCETEST"]}


def lib_path(periodic):
    name = "libtandav_host_periodic.so" if periodic else "libtandav_host.so"
    return os.path.join(OUT, name)


def build(verbose=False):
    os.makedirs(OUT, exist_ok=True)
    glue = os.path.join(ROOT, "tandav_amd", "glue", "tandav_glue.c")
    host = os.path.join(HERE, "host.c")
    libdir = os.path.join(ROOT, "tandav_amd")
    for name, defs in VARIANTS.items():
        cmd = ["gcc", "-std=gnu99", "-O2", "-fopenmp", "-ffp-contract=off",
               "-fPIC", "-shared", "-Wall", "-Wno-unknown-pragmas",
               "-D_GNU_SOURCE", '-DTANDAV_GLUE_ENV="host_env.h"',
               "-I", os.path.join(ROOT, "include"), "-I", HERE] + defs + [
               glue, host, "-L", libdir, "-ltandav_gpu",
               "-Wl,-rpath,$ORIGIN/../../../tandav_amd", "-Wl,-Bsymbolic",
               "-o", os.path.join(OUT, name)]
        if verbose:
            print(" ".join(cmd))
        subprocess.run(cmd, check=True)


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
