#!/usr/bin/env python3
"""Build the oracle (test infrastructure only).

  oracle/_ref/liboracle32.so  -- cress_oracle.c with REAL=float
  oracle/_ref/liboracle64.so  -- cress_oracle.c with REAL=double
  oracle/_ref/libcressref32.so -- the reference's own routines (build_ref.py),
                                  only where /root/reference is present;
  oracle/_ref/libcressrefmpi32.so -- likewise the routines that reach its
                                  halo-exchange layer, with the image's MPICH;
  oracle/_ref/libcressrefdrop32.so -- the same with the Fortran drop-in
                                  modules in place of four reference modules,
                                  their C calls answered by the oracle.

Plain gcc, -O2, no -march (x86-64 baseline has no FMA, and -ffp-contract=off
keeps it that way), so the FP32 build rounds like the reference's FP32 build.
OpenMP parallelises only loops whose iterations are independent; it does not
change any result.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "_ref")


def build_oracle():
    os.makedirs(OUT, exist_ok=True)
    src = os.path.join(HERE, "cress_oracle.c")
    deps = [src, os.path.join(HERE, "cress_oracle.h"), os.path.join(HERE, "cress_micro.c"),
            os.path.join(HERE, "cress_obc.c"),
            os.path.join(HERE, "cress_smooth.c"), os.path.join(HERE, "cress_turb.c")]
    libs = []
    for prec, defs in ((32, ["-DOR_FP32", "-DREAL=float"]), (64, ["-DREAL=double"])):
        lib = os.path.join(OUT, "liboracle%d.so" % prec)
        libs.append(lib)
        if os.path.exists(lib) and all(
                os.path.getmtime(lib) > os.path.getmtime(d) for d in deps if os.path.exists(d)):
            continue
        cmd = ["gcc", "-O2", "-fPIC", "-shared", "-ffp-contract=off", "-fno-fast-math",
               "-fopenmp", "-Wall", "-Wextra", "-Wno-unused-parameter"] + defs + [
               src, "-o", lib, "-lm"]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            sys.stderr.write(r.stdout + r.stderr)
            raise SystemExit("oracle build failed")
    return libs


def build(with_ref=True):
    libs = build_oracle()
    if with_ref:
        sys.path.insert(0, HERE)
        import build_ref
        build_ref.build(32)
        build_ref.build_mpi(32)
        build_ref.build_dropin()
    return libs


if __name__ == "__main__":
    build()
    print("oracle built in", OUT)
