"""Times the reference's own solver program (oracle/_ref/solver32.exe,
build_ref.build_solver: SrcOrig with flang -O3 -fopenmp, FP32, MPICH) on
this machine's cores -- the CPU baseline of the reference itself.  It never
runs on a GPU machine (the executable and this script's run directory are
kept off the GPU runs).

The run starts from the reference's own idealised supercell form
(Form/3.5.1/user.conf.DELL with its sounding DELL.sounding.txt: a warm
bubble in a sheared sounding) with the grid, time step, boundaries and
options of the workload set as below, one rank (numpe=1, MPI singleton),
OpenMP threads as given (unbound).  The program has no timers, so the time per large
step is (wall(N2) - wall(N1)) / (N2 - N1) over two runs of N1 and N2 steps
with no dumps, restarts or monitors inside the forecast (dmpitv, resitv >
etime, mxnopt=0), which cancels the setup and the t=0 output.  The setup's
own wall time varies by up to 2x between runs on this machine, so each
length runs `--repeat` times, alternating, and the shortest of each counts.

    python oracle/time_solver.py --grid 1024 1024 80 --steps 2 12 --threads 8
"""
import argparse
import os
import re
import shutil
import subprocess
import time

HERE = os.path.dirname(os.path.abspath(__file__))
REF = os.environ.get("CRESS_REFERENCE", "/root/reference")
EXE = os.path.join(HERE, "_ref", "solver32.exe")
FORM = os.path.join(REF, "Form", "3.5.1")


def user_conf(nx, ny, nz, dx, dz, dtbig, dtsml, nsteps, cphopt, wbc):
    """The form's namelist with the workload's settings."""
    txt = open(os.path.join(FORM, "user.conf.DELL")).read()
    # groups of the form that this version's rdconf does not read (it reads
    # its groups in a fixed order from standard input)
    for grp in ("radation",):
        txt = re.sub(r"^\s*&%s\b.*?^\s*/\s*$" % grp, "", txt, count=1, flags=re.M | re.S)
    # and members of &cloudphy it does not declare
    for mem in ("flqcqi_opt", "ucqcst", "ucncst", "uiqcst", "uincst"):
        txt = re.sub(r"^\s*%s\s*=.*$\n" % mem, "", txt, count=1, flags=re.M)
    etime = nsteps * dtbig
    far = 1.0e3 * etime + 1.0e6
    sets = {
        "numpe": 1, "xgroup": 1, "ygroup": 1, "xsub": 1, "ysub": 1,
        "xdim": nx + 3, "ydim": ny + 3, "zdim": nz + 3,
        "dx": dx, "dy": dx, "dz": dz, "sthopt": 0,
        "stime": 0.0, "etime": etime,
        "wbc": wbc, "ebc": wbc, "sbc": wbc, "nbc": wbc, "bbc": 2, "tbc": 2,
        "movopt": 0, "pt0cx": 0.5 * nx * dx, "pt0cy": 0.5 * ny * dx,
        "dtbig": dtbig, "dtsml": dtsml, "advopt": 3, "smtopt": 0, "divopt": 1, "buyopt": 1,
        "coropt": 0, "cphopt": cphopt, "haiopt": 0, "tubopt": 0, "botbar": 0.75 * nz * dz,
        "dmpitv": far, "monitv": far, "resopt": 0, "resitv": far, "mxnopt": 0,
    }
    # no variables in the two dumps the run still writes (t=0 and the end),
    # so that their I/O does not blur the difference of the two runs
    txt, n = re.subn(r"^(\s*dmpvar\s*=\s*)'[^']*'", lambda m: m.group(1) + "'" + "x" * 16 + "'", txt, count=1,
                     flags=re.M)
    if n != 1:
        raise SystemExit("user.conf.DELL has no line for dmpvar")
    for k, v in sets.items():
        val = ("%.6e" % v) if isinstance(v, float) else str(v)
        pat = re.compile(r"^(\s*%s\s*=\s*)([^!\n]*?)(\s*(!.*)?)$" % k, re.M)
        txt, n = pat.subn(lambda m: m.group(1) + val + m.group(3), txt, count=1)
        if n != 1:
            raise SystemExit("user.conf.DELL has no line for %s" % k)
    return txt


def run(conf, rundir, threads):
    if os.path.isdir(rundir):
        shutil.rmtree(rundir)
    os.makedirs(rundir)
    shutil.copy(os.path.join(FORM, "DELL.sounding.txt"), rundir)
    # threads unbound: OMP_PROC_BIND=close with OMP_PLACES=cores runs 3x
    # slower in this container (measured at 128x128x40)
    env = {k: v for k, v in os.environ.items() if not k.startswith("OMP_")}
    env["OMP_NUM_THREADS"] = str(threads)
    t0 = time.perf_counter()
    with open(os.path.join(rundir, "stdout.txt"), "w") as out:
        r = subprocess.run([EXE], input=conf, text=True, cwd=rundir, env=env, stdout=out,
                           stderr=subprocess.STDOUT)
    el = time.perf_counter() - t0
    log = open(os.path.join(rundir, "stdout.txt"), errors="replace").read()
    if r.returncode != 0 or "stopped abnormally" in log:
        raise SystemExit("solver32.exe failed (%d); see %s/stdout.txt" % (r.returncode, rundir))
    return el


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--grid", type=int, nargs=3, default=[1024, 1024, 80])
    ap.add_argument("--dx", type=float, default=1000.0)
    ap.add_argument("--dz", type=float, default=250.0)
    ap.add_argument("--dt", type=float, nargs=2, default=[4.0, 1.0])
    ap.add_argument("--steps", type=int, nargs=2, default=[2, 12])
    ap.add_argument("--threads", type=int, default=8)
    ap.add_argument("--cphopt", type=int, default=2)
    ap.add_argument("--wbc", type=int, default=1)
    ap.add_argument("--repeat", type=int, default=2,
                    help="runs of each length, alternating; the shortest of each counts")
    a = ap.parse_args()
    nx, ny, nz = a.grid
    walls = {n: [] for n in a.steps}
    for r in range(a.repeat):
        for n in a.steps:
            conf = user_conf(nx, ny, nz, a.dx, a.dz, a.dt[0], a.dt[1], n, a.cphopt, a.wbc)
            walls[n].append(run(conf, os.path.join(HERE, "_ref", "run", "%dx%dx%d_%d" % (nx, ny, nz, n)), a.threads))
            print("steps %d: wall %.2f s" % (n, walls[n][-1]), flush=True)
    n1, n2 = a.steps
    per = (min(walls[n2]) - min(walls[n1])) / (n2 - n1)
    print("reference solver32.exe %dx%dx%d, %d threads: %.3f s per large step, %.3e cell-steps/s"
          % (nx, ny, nz, a.threads, per, nx * ny * nz / per))


if __name__ == "__main__":
    main()
