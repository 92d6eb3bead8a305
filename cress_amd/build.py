"""Build cress_amd/_lib/libcresship.so: the HIP kernels and the C ABI of
include/cress_hip.h, compiled for gfx950 with hipcc.

Flags that matter for results:
  -ffp-contract=off   no a*b+c -> fma contraction, so every kernel rounds
                      like the reference's FP32 build;
  IEEE binary32 division and sqrt (hipcc's default for HIP);
  -Xclang -target-feature -Xclang -packed-fp32-ops
                      no packed FP32 instructions.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "_lib")
LIB = os.path.join(LIBDIR, "libcresship.so")
SOURCES = ["dyn.hip", "micro.hip", "obc.hip", "turb.hip", "mlp.hip", "diag.hip", "setup.hip", "capi.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLANG = os.environ.get("FLANG", "/opt/rocm/lib/llvm/bin/flang")
FSRC = os.path.join(HERE, "fortran")
FHOST = os.path.join(LIBDIR, "cress_run")
ARCH = os.environ.get("CRESS_OFFLOAD_ARCH", "gfx950")

FLAGS = ["--offload-arch=" + ARCH, "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-I" + os.path.join(ROOT, "include"), "-I" + CSRC, "-Wall"]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    err = "\n".join(l for l in (r.stdout + r.stderr).splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + err + "\n")
        raise SystemExit("HIP build failed")
    return err


def build(force=True, jobs=4):
    os.makedirs(LIBDIR, exist_ok=True)
    objdir = os.path.join(LIBDIR, "obj")
    os.makedirs(objdir, exist_ok=True)
    hdrs = [os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".h")]
    hdrs.append(os.path.join(ROOT, "include", "cress_hip.h"))
    newest_hdr = max(os.path.getmtime(h) for h in hdrs)
    objs, cmds = [], []
    for s in SOURCES:
        src = os.path.join(CSRC, s)
        obj = os.path.join(objdir, s.replace(".hip", ".o"))
        objs.append(obj)
        if (not force and os.path.exists(obj) and os.path.getmtime(obj) > os.path.getmtime(src)
                and os.path.getmtime(obj) > newest_hdr):
            continue
        cmds.append([HIPCC] + FLAGS + ["-c", src, "-o", obj])
    from concurrent.futures import ThreadPoolExecutor
    with ThreadPoolExecutor(jobs) as ex:
        list(ex.map(_run, cmds))
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs)
    os.replace(tmp, LIB)
    build_fortran_host()
    return LIB


def build_fortran_host():
    """The Fortran 90 host (fortran/cress_run.f90 over the ISO_C_BINDING
    module fortran/m_cress_hip.f90), linked against libcresship.so with an
    rpath to its own directory.  Skipped with a note when no flang exists."""
    if not os.path.exists(FLANG):
        sys.stderr.write("cress_amd: %s not found, Fortran host not built\n" % FLANG)
        return None
    moddir = os.path.join(LIBDIR, "mod")
    os.makedirs(moddir, exist_ok=True)
    objs = []
    for f in ("m_cress_hip.f90", "m_cress_files.f90", "cress_run.f90"):
        obj = os.path.join(moddir, f.replace(".f90", ".o"))
        _run([FLANG, "-O2", "-ffp-contract=off", "-module-dir", moddir, "-I", moddir, "-c",
              os.path.join(FSRC, f), "-o", obj])
        objs.append(obj)
    tmp = FHOST + ".tmp"
    _run([FLANG, "-o", tmp] + objs + ["-L" + LIBDIR, "-lcresship", "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, FHOST)
    return FHOST


if __name__ == "__main__":
    print(build())
