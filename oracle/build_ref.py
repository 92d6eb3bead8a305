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


def scan_sources():
    files = {}
    for d in (os.path.join(REF, "SrcOrig"),
              os.path.join(REF, "SrcDevelop", "0rig", "UserModAdd")):
        for f in sorted(os.listdir(d)):
            if f.endswith(".f90"):
                files[f[:-4]] = os.path.join(d, f)      # overlay wins
    mod2file, uses = {}, {}
    for name, path in files.items():
        txt = open(path, errors="replace").read().lower()
        for m in re.findall(r"^\s*module\s+(m_\w+)", txt, re.M):
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


if __name__ == "__main__":
    build(int(sys.argv[1]) if len(sys.argv) > 1 else 32)
