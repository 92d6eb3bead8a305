#!/usr/bin/env python3
"""Benchmark: CReSS large steps on one MI355X, in cell-steps/s.

A step is one large time step (s_intgdrv) of the whole model on the default
workload, the 1024x1024x80 warm-core vortex on an f-plane (dtbig=4 s, 8
acoustic substeps, cold-rain microphysics), with every field resident in HBM
before the timed region starts.  cells = nx*ny*nz interior cells.

Prints one JSON line (see the repository's benchmark contract) with
  roofline     -- the dominant kernel: algorithmic bytes per launch over its
                  average launch time (HIP events on the solver stream over
                  the timed region), against 8 TB/s;
  cpu_baseline -- the FP32 oracle (oracle/cress_oracle.c, OpenMP over the
                  host's threads) on a 128x128x80 sample of the same workload.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM = 8.0e12

# Algorithmic bytes per interior cell for one launch of each kernel family,
# as (fields, metrics): the distinct 3-D arrays the launch must read plus
# those it must write (4 B each; 2-D arrays and top-sponge-only reads left
# out), split into prognostic/work fields and base-state/metric fields.  A
# metric field uploaded with all columns equal (flat terrain, horizontally
# uniform base state -- every workload here) is read from a per-level table
# instead of HBM, so it adds no bytes; otherwise it counts as a field.
KERNEL_ARRAYS = {
    # u,v,w -> wc,rstxu,rstxv,rstxwc | rst8u,rst8v,rst8w,jcb8w
    "k_prep": (7, 4),
    # u,v,up,vp,rstxu,rstxv,rstxwc -> ufrc,vfrc | ubr,vbr,rst8u,rst8v,rst
    "k_force_uv": (9, 5),
    # w,wp,ptp,qv..qg(6),rstxu,rstxv,rstxwc -> wfrc | qvbr,ptbr,rst,rst8w
    "k_force_w": (13, 4),
    # u,v,wc,pp,ppp -> pfrc | jcb,jcb8u,jcb8v,jcb8w
    "k_force_p": (6, 4),
    # 8 scalars, rstxu,rstxv,rstxwc,w -> 8 forcings | ptbr,rbr
    "k_force_s": (20, 2),
    # uf,vf,wc,ppf,ufrc,vfrc -> ptil,uf,vf | jcb,rst8u,rst8v,rst8w
    "k_ac_ptuv": (9, 4),
    # uf,vf,ptil,pfrc,wfrc,wf,ppf -> wf,wc,ppf | jcb,jcb8u,jcb8v,jcb8w,rcsq,rst,rbr,rst8w
    "k_ac_col": (10, 8),
    # ptpp,ptfrc,qp(6),qvfrc,ncip,qf(5 forcings),ncif -> ptpf,qf(6),ncif | rst
    "k_steps": (24, 1),
    # 12 fields x 3 levels -> 12 present, 5 future (satadjst) | pbr,ptbr
    "k_asselin_sat": (53, 2),
    # past pp,ptp,q(6),nci, forcings ptfrc,qvfrc,5 q,nci -> future ptp,q(6),nci,
    #   urq,usq,ugq | jcb,pbr,ptbr,rbr,rst
    "k_micro_src": (28, 5),
    # urq,usq,ugq, past qc,qi, future qc,qr,qi,qs,qg,nci -> future qc,qr,qi,qs,qg,nci | rbr,rst
    "k_micro_fall": (17, 2),
}


def kernel_arrays(name, uniform=True):
    f, m = KERNEL_ARRAYS[name]
    return f if uniform else f + m


# Calibrated HBM bytes per launch (FETCH_SIZE/WRITE_SIZE passes of
# rocprofv3), one file per workload; see profiles/.
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "vortex_1024x1024x80_traffic.txt")


def measured_traffic(workload, kernel):
    """HBM bytes per launch of a kernel from the committed counter summary
    of the same workload, or None."""
    try:
        lines = open(TRAFFIC_FILE).read().splitlines()
    except OSError:
        return None
    if len(lines) < 2 or lines[1].strip() != workload:
        return None
    for ln in lines[2:]:
        f = ln.split()
        if len(f) == 3 and f[0] == kernel:
            return float(f[1]) + float(f[2])
    return None


def step_bytes_per_cell(nsstp, cphopt, uniform=True):
    """Whole-step algorithmic bytes per interior cell: the sum over one
    step's launches of their algorithmic bytes (KERNEL_ARRAYS)."""
    k = lambda n: 4 * kernel_arrays(n, uniform)
    forcing = k("k_prep") + k("k_force_uv") + k("k_force_w") + k("k_force_p") + k("k_force_s")
    acoustic = k("k_ac_ptuv") + k("k_ac_col")
    if cphopt == 2:
        return forcing + nsstp * acoustic + k("k_micro_src") + k("k_micro_fall") + k("k_asselin_sat")
    # dry dynamics: two scalars in k_force_s and k_steps, six fields in timeflt
    return (forcing - 4 * 12) + nsstp * acoustic + 4 * 6 + 4 * 24


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="vortex", choices=["vortex", "supercell", "bubble"])
    ap.add_argument("--nx", type=int, default=None)
    ap.add_argument("--ny", type=int, default=None)
    ap.add_argument("--nz", type=int, default=None)
    ap.add_argument("--cphopt", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-nxy", type=int, default=128)
    ap.add_argument("--ensemble", type=int, default=12,
                    help="members of the batched-ensemble measurement (0: none)")
    ap.add_argument("--ens-steps", type=int, default=6)
    ap.add_argument("--no-tune", action="store_true",
                    help="keep the default launch shapes instead of autotuning them on the workload")
    ap.add_argument("--launch", default=None,
                    help="explicit launch shapes, e.g. ptuv_by=16,col_by=2 (the rest default; no tuning)")
    ap.add_argument("--mlp", type=int, default=1, help="measure the learned microphysics (0: no)")
    ap.add_argument("--mlp-nxy", type=int, default=256)
    ap.add_argument("--mlp-steps", type=int, default=60,
                    help="large steps of the plain and the network run, from the cold start")
    ap.add_argument("--mlp-train", type=int, default=3000, help="Adam steps")
    return ap.parse_args()


MLP_FLOP = 2 * (9 * 128 + 128 * 128 + 128 * 128 + 128 * 7)   # forward flop per cell
PEAK_BF16 = 2.5e15                                           # dense BF16 matrix-core peak


def mlp_bench(args):
    """The learned microphysics (a 9-128-128-128-7 GELU network in place of
    srcblk's local sources) on the supercell (open sides, cold rain, dtbig
    3 s, 6 substeps) at mlp_nxy^2 x 64: training pairs recorded from the
    plain scheme's own steps, the network trained on the GPU (BF16 matrix
    cores, FP32 accumulation, Adam), then the same mlp_steps large steps
    from the cold start with srcblk and with the network, timed, and scored
    on the accumulated surface precipitation and the mixing ratios."""
    from cress_amd import cases, host, learned
    n = args.mlp_nxy
    case = cases.build("supercell", n, n, 64)
    cfg = case.cfg
    cells = (cfg.ni - 3) * (cfg.nj - 3) * (cfg.nk - 3)
    k = args.mlp_steps
    m = host.Mlp(capacity=1 << 25, max_batch=16384)
    t0 = time.perf_counter()
    npairs = learned.record(m, case, k)
    t_rec = time.perf_counter() - t0
    m.fit_norm(0.05)
    m.init(1)
    t0 = time.perf_counter()
    losses = learned.train(m, args.mlp_train, batch=16384, lr=2e-3)
    t_train = time.perf_counter() - t0
    print("bench: mlp %d pairs from %d plain steps (%.1f s), %d Adam steps in %.1f s, loss %.3g -> %.3g"
          % (npairs, k, t_rec, args.mlp_train, t_train, losses[:20].mean(), losses[-20:].mean()),
          file=sys.stderr, flush=True)
    L = host.lib()
    nk = L.cress_kernel_count()
    L.cress_kernel_name.restype = __import__("ctypes").c_char_p
    names = [L.cress_kernel_name(i).decode() for i in range(nk)]

    def run(mode):
        import ctypes as C
        s = host.Solver(cfg)
        case.load(s)
        if mode:
            s.set_mlp(m, 2)
        s.step(1, first=True)
        s.sync()
        t = time.perf_counter()
        s.step(k - 1)
        s.sync()
        el = time.perf_counter() - t
        # one more step with every kernel family timed
        L.cress_reset_timing(s.ctx)
        s.set_timing(True)
        s.step(1)
        s.sync()
        s.set_timing(False)
        tot = (C.c_double * nk)()
        cnt = (C.c_long * nk)()
        L.cress_kernel_times(s.ctx, tot, cnt, nk)
        kt = {names[i]: tot[i] / max(cnt[i], 1) for i in range(nk) if cnt[i]}
        net_cells = s.mlp_cells() if mode else 0
        return s, el / (k - 1), kt, net_cells

    s_plain, t_plain, kt_plain, _ = run(0)
    s_net, t_net, kt_net, ncells = run(1)
    sc = learned.score(s_net, s_plain)
    t_inf = kt_net.get("k_mlp_infer", float("nan")) * 1e-3
    achieved = ncells * MLP_FLOP / t_inf if t_inf > 0 else float("nan")
    for s in (s_plain, s_net):
        s.close()
    m.close()
    micro = lambda kt: sum(kt.get(x, 0.0) for x in ("k_micro_src", "k_mlp_select", "k_mlp_infer"))
    out = {"workload": "supercell %dx%dx64 (open sides, cold rain, nsstp %d)" % (n, n, cfg.nsstp),
           "steps": k, "pairs": npairs, "train_steps": args.mlp_train, "train_batch": 16384,
           "train_s": t_train, "loss_first": float(losses[:20].mean()), "loss_last": float(losses[-20:].mean()),
           "plain_ms_per_step": 1e3 * t_plain, "net_ms_per_step": 1e3 * t_net, "speedup": t_plain / t_net,
           "micro_sources_ms": {"plain": micro(kt_plain), "net": micro(kt_net)},
           "kernels_ms": {"plain": {k: kt_plain.get(k, 0.0) for k in ("k_micro_src", "k_micro_fall")},
                          "net": {k: kt_net.get(k, 0.0) for k in ("k_mlp_select", "k_mlp_infer", "k_micro_src",
                                                                  "k_micro_fall")}},
           "net_cells_per_step": ncells, "net_cell_frac": ncells / cells,
           "precip_plain": sc["precip_plain"], "precip_net": sc["precip"],
           "precip_rel_err": sc["precip_rel_err"],
           "q_rmse_rel": {q: sc["rmse_" + q] / max(sc["scale_" + q], 1e-30) for q in learned.QS},
           "mfma": {"kernel": "k_mlp_infer", "achieved": achieved / 1e12, "peak": PEAK_BF16 / 1e12,
                    "unit": "TFLOP/s", "frac": achieved / PEAK_BF16}}
    print("bench: mlp plain %.3f ms/step, network %.3f ms/step (%.2fx), precip err %.3g, infer %.1f TFLOP/s"
          % (1e3 * t_plain, 1e3 * t_net, t_plain / t_net, sc["precip_rel_err"], achieved / 1e12),
          file=sys.stderr, flush=True)
    return out


def ensemble_bench(args):
    """The batched ensemble against the same members stepped one after
    another: members drawn like the reference-style supercell ensemble (open
    sides, cold rain, dtbig 3 s, 6 substeps; nx=ny from 128..384, nz 48 or
    64, bubble 1-4 K, shear 15-35 m/s, member m seeded 1000+m), a cold-start
    step and one graph cycle of warm-up, then ens_steps large steps each
    way.  Steps are a multiple of the 6-substep members' 3-step cycle."""
    from cress_amd import cases, host
    rng = np.random.default_rng(7)
    solvers, cells = [], 0
    t0 = time.time()
    for m in range(args.ensemble):
        n = int(rng.choice([128, 192, 256, 320, 384]))
        nz = int(rng.choice([48, 64]))
        c = cases.build("supercell", n, n, nz, seed=1000 + m, amp=1.0 + 3.0 * rng.random(),
                        shear=15.0 + 20.0 * rng.random())
        s = host.Solver(c.cfg)
        c.load(s)
        del c
        if not args.no_tune:
            s.autotune()          # cached per grid size: members of one size are tuned once
        solvers.append(s)
        cells += n * n * nz
        print("bench: ensemble member %d %dx%dx%d" % (m, n, n, nz), file=sys.stderr, flush=True)
    print("bench: ensemble of %d members (%.1f M cells) built in %.1f s"
          % (args.ensemble, cells / 1e6, time.time() - t0), file=sys.stderr, flush=True)
    ens = host.Ensemble(solvers)
    ens.step(1, first=True)
    ens.step(6)
    ens.sync()
    k = args.ens_steps
    t0 = time.perf_counter()
    for s in solvers:
        s.step(k)
        s.sync()
    seq = time.perf_counter() - t0
    t0 = time.perf_counter()
    ens.step(k)
    ens.sync()
    bat = time.perf_counter() - t0
    ens.close()
    for s in solvers:
        s.close()
    out = {"members": args.ensemble, "cells": cells, "steps": k,
           "sequential_cell_steps_per_s": cells * k / seq, "batched_cell_steps_per_s": cells * k / bat,
           "speedup": seq / bat,
           "how": "members one after another with cress_large_step vs one hipGraph launch per "
                  "3-step cycle with one branch per member (cress_ens_step)"}
    print("bench: ensemble sequential %.1f ms/step, batched %.1f ms/step, speedup %.2f"
          % (1e3 * seq / k, 1e3 * bat / k, seq / bat), file=sys.stderr, flush=True)
    return out


def cpu_baseline(args, nz):
    """The FP32 oracle (C restatement, OpenMP over levels) on the host cores,
    on a cpu_nxy x cpu_nxy x nz sample of the same workload: about 10 s of
    steps after one warm-up step."""
    from cress_amd import cases
    from oracle import pyoracle as po
    nxy = args.cpu_nxy
    case = cases.build(args.workload, nxy, nxy, nz, cphopt=args.cphopt)
    m = po.Model(case, 32)
    threads = int(m.o.lib.or_threads())
    m.step(1, first=True)
    n, t0 = 0, time.perf_counter()
    while True:
        m.step(1)
        n += 1
        el = time.perf_counter() - t0
        if el > 10.0 or n >= 50:
            break
    return {"value": nxy * nxy * nz * n / el, "unit": "cell-steps/s", "cores": threads, "kind": "port",
            "sample": "FP32 oracle (C restatement, %d OpenMP threads), %s %dx%dx%d, %d large steps after 1 warm-up"
                      % (threads, args.workload, nxy, nxy, nz, n),
            "cpu_model": host_cpu_model(), "cpus_allowed": len(os.sched_getaffinity(0)),
            "omp": {k: os.environ.get(k, "unset") for k in ("OMP_NUM_THREADS", "OMP_PROC_BIND", "OMP_PLACES")}}


def host_cpu_model():
    """The host CPU's model name (first processor of /proc/cpuinfo)."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only; --gpus %d is not supported" % args.gpus,
              file=sys.stderr)
        sys.exit(2)
    from cress_amd import cases, host
    t0 = time.time()
    case = cases.build(args.workload, args.nx, args.ny, args.nz, cphopt=args.cphopt)
    cfg = case.cfg
    nx, ny, nz = cfg.ni - 3, cfg.nj - 3, cfg.nk - 3
    print("bench: %s %dx%dx%d nsstp=%d built in %.1f s" % (args.workload, nx, ny, nz, cfg.nsstp,
                                                          time.time() - t0), file=sys.stderr, flush=True)
    s = host.Solver(cfg)
    case.load(s)
    del case
    s.sync()
    how = "default" if args.no_tune else "autotuned"
    if args.launch:
        shapes = s.get_launch()
        for kv in args.launch.split(","):
            k, v = kv.split("=")
            if k not in shapes:
                raise SystemExit("bench: unknown launch shape %r (known: %s)" % (k, ", ".join(shapes)))
            shapes[k] = int(v)
        s.set_launch(shapes)
        shapes, how = s.get_launch(), "explicit"
    else:
        shapes = s.get_launch() if args.no_tune else s.autotune()
    print("bench: launch shapes %s (%s)" % (shapes, how), file=sys.stderr, flush=True)
    s.step(1, first=True)
    import ctypes as C
    L = host.lib()
    nk = L.cress_kernel_count()
    L.cress_kernel_name.restype = C.c_char_p
    names = [L.cress_kernel_name(i).decode() for i in range(nk)]

    def kernel_times():
        tot = (C.c_double * nk)()
        cnt = (C.c_long * nk)()
        L.cress_kernel_times(s.ctx, tot, cnt, nk)
        return {names[i]: (tot[i], cnt[i]) for i in range(nk) if cnt[i]}

    # the remaining warm-up steps time every kernel family (and the stages),
    # which names the dominant kernel; the timed steps then bracket only its
    # launches with HIP events, so the rest of the step runs uninstrumented
    L.cress_reset_timing(s.ctx)
    s.set_timing(True)
    s.step(max(args.warmup - 1, 1))
    s.sync()
    s.set_timing(False)
    kern = kernel_times()
    stages = s.stage_times()
    nwarm = max(args.warmup - 1, 1)
    dom = max((k for k in kern if k in KERNEL_ARRAYS), key=lambda k: kern[k][0])
    print("bench: warm-up done", file=sys.stderr, flush=True)
    L.cress_reset_timing(s.ctx)
    L.cress_set_timing(s.ctx, 2 + names.index(dom))
    s.sync()
    t0 = time.perf_counter()
    s.step(args.steps)
    s.sync()
    el = time.perf_counter() - t0
    s.set_timing(False)
    cells = nx * ny * nz
    value = cells * args.steps / el
    ms_per_step = 1e3 * el / args.steps
    kdom = kernel_times()[dom]
    avg_ms = kdom[0] / kdom[1]
    per_launch_bytes = kernel_arrays(dom) * 4.0 * cells
    achieved = per_launch_bytes / (avg_ms * 1e-3)
    step_b = step_bytes_per_cell(cfg.nsstp, cfg.cphopt)
    traffic = measured_traffic("workload %s %dx%dx%d nsstp=%d cphopt=%d" % (
        args.workload, nx, ny, nz, cfg.nsstp, cfg.cphopt), dom)
    print("bench: %.2f ms/step, whole-step %.2f TB/s algorithmic; %s %.3f ms per launch; "
          "kernels (instrumented warm-up): %s" % (
              ms_per_step, step_b * cells / (ms_per_step * 1e-3) / 1e12, dom, avg_ms,
              ", ".join("%s %.3f ms x%d" % (k, v[0] / max(v[1], 1), v[1]) for k, v in
                        sorted(kern.items(), key=lambda kv: -kv[1][0]))), file=sys.stderr, flush=True)
    s.close()
    ensemble = ensemble_bench(args) if args.ensemble > 0 else None
    mlp = mlp_bench(args) if args.mlp else None
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(args, nz)
    out = {
        "metric": "cell-steps/s (1 GPU), % HBM roofline; ensemble & MLP-microphysics vs plain",
        "value": value,
        "unit": "cell-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "%s %dx%dx%d" % (args.workload, nx, ny, nz), "nsstp": cfg.nsstp,
                   "dtbig": cfg.dtbig, "cphopt": cfg.cphopt, "advopt": cfg.advopt,
                   "lateral": ("open (option 7, gwave %g m/s)" % cfg.gwave) if cfg.wbc == 7 else "periodic",
                   "parallelism": "single GPU",
                   "launch_shapes": dict(shapes, how=how)},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved / 1e9, "peak": PEAK_HBM / 1e9,
                     "unit": "GB/s", "frac": achieved / PEAK_HBM, "traffic": traffic,
                     "step_frac": step_b * cells / (ms_per_step * 1e-3) / PEAK_HBM,
                     "step_bytes_per_cell": step_b},
        "stages_ms": {k: v / nwarm for k, v in stages.items()},
        "ensemble": ensemble,
        "mlp": mlp,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
