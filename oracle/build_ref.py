#!/usr/bin/env python3
"""Build oracle/_ref/libcressref32.so: the reference's own Fortran routines,
compiled from their sources where they lie under /root/reference, plus the
project-written bind(C) harness in oracle/ref_harness.f90.

Test infrastructure only.  Nothing in the product path links this library.

Recipe (mirrors what the reference's configure.csh/compile.csh would stage,
without running them -- they need csh and MPI):
  * sources: SrcOrig/*.f90, with SrcDevelop/0rig/UserModAdd/*.f90 taking
    precedence (configure.csh:218,288 copy SrcOrig then overlay UserModAdd);
  * only the routines the harness `use`s, plus their module closure, are
    compiled.  Every routine in that closure is MPI-free (m_defmpi is not
    reached), so no external library is linked;
  * flang -O2 -fPIC, FP32 (the reference's default REAL), no OpenMP.
Outputs go to oracle/_ref/ only (git-ignored).

build_mpi() makes oracle/_ref/libcressrefmpi32.so in the same way for the
routines whose module closure reaches the halo-exchange layer (the harness
oracle/ref_harness_mpi.f90: s_intgdrv, s_hevi, s_stepuv, s_stepwi,
s_steppi, s_steps, s_cloudphy, s_fallblk, s_phy2cnt, s_shift2nd).  Those
need an MPI library: the image's MPICH (/opt/conda: mpif.h, libmpifort,
libmpi).  Its Fortran module file is gfortran's and unreadable by flang, so
m_defmpi is compiled with the alternative the reference itself carries:
defmpi.f90 says `use mpi` with `include 'mpif.h'` commented out beside it
(SrcOrig/defmpi.f90:18,22); the staged copy under oracle/_ref/ comments
out the `use` line and restores the `include` line, placed after the
module's `implicit none` (flang does not accept the include's declarations
before it); nothing else changes.
"""
import os
import re
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

REF = os.environ.get("CRESS_REFERENCE", "/root/reference")
HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "_ref")
FLANG = os.environ.get("FLANG", "/opt/rocm/lib/llvm/bin/flang")


def solver_objects():
    """Object names of the solver's build list (SrcOrig/Make_solver)."""
    txt = open(os.path.join(REF, "SrcOrig", "Make_solver"), errors="replace").read()
    return set(re.findall(r"^\s*(\w+)\.o\s*\\?\s*$", txt, re.M))


def scan_sources(prefer_solver=False):
    """Source files by name, the module each defines and the modules each
    uses.  Some modules have two sources (inimpi_multi.f90 and
    inimpi_single.f90, ...); with prefer_solver the one in the solver's
    object list (the _multi, MPI variant) defines it, else the later name."""
    files = {}
    for d in (os.path.join(REF, "SrcOrig"),
              os.path.join(REF, "SrcDevelop", "0rig", "UserModAdd")):
        for f in sorted(os.listdir(d)):
            if f.endswith(".f90"):
                files[f[:-4]] = os.path.join(d, f)      # overlay wins
    objs = solver_objects() if prefer_solver else set()
    mod2file, uses = {}, {}
    for name, path in files.items():
        txt = open(path, errors="replace").read().lower()
        for m in re.findall(r"^\s*module\s+(m_\w+)", txt, re.M):
            if m in mod2file and prefer_solver and mod2file[m] in objs and name not in objs:
                continue
            mod2file[m] = name
        uses[name] = set(re.findall(r"^\s*use\s+(m_\w+)", txt, re.M))
    return files, mod2file, uses


def harness_modules(path):
    txt = open(path).read().lower()
    return sorted(set(re.findall(r"^\s*use\s+(m_\w+)", txt, re.M)))


def closure(roots, mod2file, uses):
    seen = set()
    stack = [mod2file[m] for m in roots if m in mod2file]
    while stack:
        f = stack.pop()
        if f in seen:
            continue
        seen.add(f)
        for m in uses.get(f, ()):
            g = mod2file.get(m)
            if g and g not in seen:
                stack.append(g)
    return seen


def levels(names, mod2file, uses):
    lvl = {}

    def depth(f):
        if f in lvl:
            return lvl[f]
        lvl[f] = 0
        deps = [mod2file[m] for m in uses.get(f, ()) if m in mod2file]
        lvl[f] = 1 + max((depth(g) for g in deps if g != f), default=-1)
        return lvl[f]

    for n in names:
        depth(n)
    out = {}
    for n in names:
        out.setdefault(lvl[n], []).append(n)
    return [sorted(out[k]) for k in sorted(out)]


def run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise SystemExit("oracle/_ref build failed")


def build(prec=32, jobs=8):
    if not os.path.isdir(os.path.join(REF, "SrcOrig")):
        print("reference sources absent; oracle/_ref not rebuilt")
        return None
    files, mod2file, uses = scan_sources()
    harness = os.path.join(HERE, "ref_harness.f90")
    names = closure(harness_modules(harness), mod2file, uses)
    if "defmpi" in names:
        raise SystemExit("harness closure reaches MPI; refusing to build")
    obj = os.path.join(OUT, "obj%d" % prec)
    os.makedirs(obj, exist_ok=True)
    flags = ["-O2", "-fPIC", "-ffp-contract=off", "-module-dir", obj, "-I", obj]
    if prec == 64:
        flags.append("-fdefault-real-8")
    objs = []
    for group in levels(names, mod2file, uses):
        cmds = []
        for n in group:
            o = os.path.join(obj, n + ".o")
            objs.append(o)
            src = files[n]
            if (os.path.exists(o) and os.path.getmtime(o) > os.path.getmtime(src)):
                continue
            cmds.append([FLANG, "-c"] + flags + [src, "-o", o])
        with ThreadPoolExecutor(jobs) as ex:
            list(ex.map(run, cmds))
    ho = os.path.join(obj, "ref_harness.o")
    run([FLANG, "-c"] + flags + [harness, "-o", ho])
    lib = os.path.join(OUT, "libcressref%d.so" % prec)
    run([FLANG, "-shared", "-fPIC", "-o", lib, ho] + objs)
    print("built", lib, "from", len(objs), "reference files")
    return lib


MPI_INC = os.environ.get("CRESS_MPI_INC", "/opt/conda/include")
MPI_LIB = os.environ.get("CRESS_MPI_LIB", "/opt/conda/lib")


def have_mpi():
    return (os.path.exists(os.path.join(MPI_INC, "mpif.h"))
            and os.path.exists(os.path.join(MPI_LIB, "libmpifort.so")))


def stage_defmpi(files, dst):
    """The reference's defmpi.f90 with its `use mpi` line commented out and
    its commented `include 'mpif.h'` line restored."""
    txt = open(files["defmpi"]).read()
    use = re.compile(r"^([ \t]*)use mpi[ \t]*$", re.M | re.I)
    inc = re.compile(r"^!OLD([ \t]*)include 'mpif.h'[ \t]*$", re.M | re.I)
    if len(use.findall(txt)) != 1 or len(inc.findall(txt)) != 1:
        raise SystemExit("defmpi.f90 does not have the expected use/include lines")
    txt = use.sub(r"!\1use mpi", txt)
    # The include's specification statements may not precede IMPLICIT NONE
    # (flang rejects that order), so the restored line moves below it.
    txt = inc.sub("!OLD  (moved below implicit none)", txt)
    imp = re.compile(r"^([ \t]*)implicit none[ \t]*$", re.M | re.I)
    if len(imp.findall(txt)) != 1:
        raise SystemExit("defmpi.f90 does not have one implicit none line")
    txt = imp.sub(r"\1implicit none\n\1include 'mpif.h'", txt)
    os.makedirs(os.path.dirname(dst), exist_ok=True)
    if not os.path.exists(dst) or open(dst).read() != txt:
        with open(dst, "w") as f:
            f.write(txt)
    return dst


def build_mpi(prec=32, jobs=8):
    if not os.path.isdir(os.path.join(REF, "SrcOrig")):
        print("reference sources absent; oracle/_ref (MPI part) not rebuilt")
        return None
    if not have_mpi():
        print("MPICH not found; oracle/_ref (MPI part) not built")
        return None
    files, mod2file, uses = scan_sources(prefer_solver=True)
    obj = os.path.join(OUT, "objmpi%d" % prec)
    os.makedirs(obj, exist_ok=True)
    files["defmpi"] = stage_defmpi(files, os.path.join(obj, "src", "defmpi.f90"))
    h1 = os.path.join(HERE, "ref_harness.f90")
    h2 = os.path.join(HERE, "ref_harness_mpi.f90")
    h3 = os.path.join(HERE, "ref_harness_files.f90")   # linked by build_dropin
    names = closure(harness_modules(h1) + harness_modules(h2) + harness_modules(h3), mod2file, uses)
    flags = ["-O2", "-fPIC", "-ffp-contract=off", "-module-dir", obj, "-I", obj, "-I", MPI_INC]
    if prec == 64:
        flags.append("-fdefault-real-8")
    objs = []
    for group in levels(names, mod2file, uses):
        cmds = []
        for n in group:
            o = os.path.join(obj, n + ".o")
            objs.append(o)
            src = files[n]
            if (os.path.exists(o) and os.path.getmtime(o) > os.path.getmtime(src)):
                continue
            cmds.append([FLANG, "-c"] + flags + [src, "-o", o])
        with ThreadPoolExecutor(jobs) as ex:
            list(ex.map(run, cmds))
    hos = []
    for h in (h1, h2):
        ho = os.path.join(obj, os.path.basename(h).replace(".f90", ".o"))
        run([FLANG, "-c"] + flags + [h, "-o", ho])
        hos.append(ho)
    lib = os.path.join(OUT, "libcressrefmpi%d.so" % prec)
    run([FLANG, "-shared", "-fPIC", "-o", lib] + hos + objs +
        ["-L" + MPI_LIB, "-Wl,-rpath," + MPI_LIB, "-lmpifort", "-lmpi"])
    print("built", lib, "from", len(objs), "reference files")
    return lib


# Drop-in modules run inside the reference's own solver on the CPU: the
# modules of cress_amd/fortran/dropin/ replace the reference's sources of
# the same modules, and oracle/dropin_cpu.c answers their tier-2 C calls
# with the oracle (liboracle32.so).  Built from the same objects as
# build_mpi(); the replaced reference objects are left out of the link.
DROPINS = ("gaussel", "advs", "advuvw", "advp", "turbdrv", "turbflx", "turbs", "forcetke")


def build_dropin(names=DROPINS, jobs=8):
    if not os.path.isdir(os.path.join(REF, "SrcOrig")) or not have_mpi():
        print("reference sources or MPICH absent; oracle/_ref (drop-in part) not built")
        return None
    base = build_mpi(32, jobs)
    if base is None:
        return None
    files, mod2file, uses = scan_sources(prefer_solver=True)
    objm = os.path.join(OUT, "objmpi32")
    objd = os.path.join(OUT, "objdrop32")
    os.makedirs(objd, exist_ok=True)
    repo = os.path.dirname(HERE)
    fdir = os.path.join(repo, "cress_amd", "fortran")
    flags = ["-O2", "-fPIC", "-ffp-contract=off", "-I", objm, "-I", objd, "-I", MPI_INC, "-module-dir", objd]
    run([FLANG, "-c"] + flags + [os.path.join(fdir, "m_cress_hip.f90"), "-o", os.path.join(objd, "m_cress_hip.o")])
    run([FLANG, "-c"] + flags + [os.path.join(fdir, "m_cress_files.f90"), "-o", os.path.join(objd, "m_cress_files.o")])
    dobjs = [os.path.join(objd, "m_cress_hip.o"), os.path.join(objd, "m_cress_files.o")]
    for n in names:
        o = os.path.join(objd, n + ".o")
        run([FLANG, "-c"] + flags + [os.path.join(fdir, "dropin", n + ".f90"), "-o", o])
        dobjs.append(o)
    shim = os.path.join(objd, "dropin_cpu.o")
    run(["gcc", "-c", "-O2", "-fPIC", "-ffp-contract=off", "-I", os.path.join(repo, "include"), "-I", HERE,
         os.path.join(HERE, "dropin_cpu.c"), "-o", shim])
    mock = os.path.join(objd, "ctx_cpu.o")
    run(["gcc", "-c", "-O2", "-fPIC", "-I", os.path.join(repo, "include"), os.path.join(HERE, "ctx_cpu.c"),
         "-o", mock])
    h1 = os.path.join(HERE, "ref_harness.f90")
    h2 = os.path.join(HERE, "ref_harness_mpi.f90")
    h3 = os.path.join(HERE, "ref_harness_files.f90")
    h3o = os.path.join(objd, "ref_harness_files.o")
    run([FLANG, "-c"] + flags + [h3, "-o", h3o])
    closure_names = closure(harness_modules(h1) + harness_modules(h2) + harness_modules(h3), mod2file, uses)
    replaced = {mod2file["m_" + n] for n in names}
    robjs = [os.path.join(objm, n + ".o") for g in levels(closure_names, mod2file, uses) for n in g
             if n not in replaced]
    hobjs = [os.path.join(objm, "ref_harness.o"), os.path.join(objm, "ref_harness_mpi.o")]
    lib = os.path.join(OUT, "libcressrefdrop32.so")
    run([FLANG, "-shared", "-fPIC", "-o", lib] + hobjs + [h3o] + robjs + dobjs + [shim, mock] +
        ["-L" + OUT, "-Wl,-rpath," + OUT, "-loracle32", "-lm",
         "-L" + MPI_LIB, "-Wl,-rpath," + MPI_LIB, "-lmpifort", "-lmpi"])
    print("built", lib, "with the drop-ins", ", ".join(names))
    return lib


# The reference's own solver program (solver.f90 and its module closure,
# the solver's build list) with OpenMP, FP32, one executable: the CPU
# baseline of the reference itself, timed in this container only (it never
# travels to a GPU machine).  Same staged defmpi.f90 as build_mpi().
def build_solver(jobs=8, opt="-O3"):
    if not os.path.isdir(os.path.join(REF, "SrcOrig")) or not have_mpi():
        print("reference sources or MPICH absent; oracle/_ref/solver32.exe not built")
        return None
    files, mod2file, uses = scan_sources(prefer_solver=True)
    obj = os.path.join(OUT, "objexe32")
    os.makedirs(obj, exist_ok=True)
    files["defmpi"] = stage_defmpi(files, os.path.join(obj, "src", "defmpi.f90"))
    names = closure([m for m in uses["solver"]], mod2file, uses) | {"solver"}
    flags = [opt, "-fopenmp", "-module-dir", obj, "-I", obj, "-I", MPI_INC]
    objs = []
    for group in levels(names, mod2file, uses):
        cmds = []
        for n in group:
            o = os.path.join(obj, n + ".o")
            objs.append(o)
            if os.path.exists(o) and os.path.getmtime(o) > os.path.getmtime(files[n]):
                continue
            cmds.append([FLANG, "-c"] + flags + [files[n], "-o", o])
        with ThreadPoolExecutor(jobs) as ex:
            list(ex.map(run, cmds))
    exe = os.path.join(OUT, "solver32.exe")
    run([FLANG, "-fopenmp", "-o", exe] + objs +
        ["-L" + MPI_LIB, "-Wl,-rpath," + MPI_LIB, "-lmpifort", "-lmpi"])
    print("built", exe, "from", len(objs), "reference files")
    return exe


if __name__ == "__main__":
    build(int(sys.argv[1]) if len(sys.argv) > 1 else 32)
    build_mpi(int(sys.argv[1]) if len(sys.argv) > 1 else 32)
    build_dropin()
