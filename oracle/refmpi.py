"""The reference's own MPI-reaching routines (s_intgdrv, s_hevi, s_stepuv,
s_stepwi, s_steppi, s_steps, s_cloudphy, s_fallblk, s_phy2cnt, s_shift2nd)
through oracle/_ref/libcressrefmpi32.so -- TEST INFRASTRUCTURE ONLY.

MPI can be initialised once per process and the reference allocates its
microphysics module arrays once per grid, so every job runs in a fresh
child process (spawned, not forked): `run(job, *args)` calls the module
function `job` there with a handle on the library, already configured for
the case's options and grid, and returns what it returns.  The library
exists only where the reference sources and MPICH were present at build
time (this container); nothing here runs on a GPU.
"""
import ctypes as C
import multiprocessing as mp
import os

import numpy as np

from oracle import pyoracle as po

LIB = os.path.join(po.REFDIR, "libcressrefmpi32.so")
# the same routines with the drop-in modules of cress_amd/fortran/dropin/ in
# place of the reference's m_gaussel, m_advs, m_advuvw, m_advp, answered by
# the oracle on the CPU (build_ref.py: build_dropin)
LIB_DROPIN = os.path.join(po.REFDIR, "libcressrefdrop32.so")

# base block of the harness (ref_harness_mpi.f90)
BASE = ["jcb", "jcb8u", "jcb8v", "jcb8w", "ubr", "vbr", "pbr", "ptbr", "qvbr", "rbr", "rst",
        "rst8u", "rst8v", "rst8w", "rcsq"]
# per-level block
LEV = ["u", "v", "w", "pp", "ptp", "qv", "qc", "qr", "qi", "qs", "qg", "ncc", "ncr", "nci", "ncs", "ncg"]


def have():
    return os.path.exists(LIB)


class Lib:
    """Handle on the library inside the child: configured for one case."""

    def __init__(self, cfg, path=LIB):
        self.lib = C.CDLL(path)
        self.cfg = cfg
        self.lib.ref_mpi_init()
        f = C.c_float
        I = C.c_int
        self.lib.ref_configure(I(cfg.wbc), I(cfg.ebc), I(cfg.sbc), I(cfg.nbc), I(cfg.bbc),
                               I(cfg.tbc), I(cfg.advopt), I(cfg.impopt), I(cfg.gwmopt),
                               I(cfg.buyopt), I(cfg.divopt), I(cfg.cphopt), I(cfg.haiopt),
                               I(cfg.vspopt), I(cfg.coropt), f(cfg.dx), f(cfg.dy), f(cfg.dz),
                               f(cfg.weicoe), f(cfg.filcoe), f(cfg.thresq),
                               C.c_char_p(bytes(cfg.vspvar).ljust(10, b"x")))
        self.lib.ref_mpi_setup(I(cfg.ni), I(cfg.nj), I(cfg.nk), I(cfg.sthopt), I(cfg.cphopt),
                               f(cfg.gwave), I(cfg.smtopt), f(cfg.smhcoe), f(cfg.smvcoe))

    def __getattr__(self, name):
        fn = getattr(self.lib, name)

        def call(*args):
            conv = []
            for a in args:
                if isinstance(a, np.ndarray):
                    conv.append(po.ptr(a))
                elif isinstance(a, float):
                    conv.append(C.c_float(a))
                else:
                    conv.append(a)
            return fn(*conv)
        return call


def _child(job, cfgbytes, args, path):
    cfg = po.Cfg.from_buffer_copy(cfgbytes)
    return job(Lib(cfg, path), *args)


def run(job, cfg, *args, lib=LIB):
    """job(lib, *args) in a fresh spawned process; cfg is an oracle Cfg;
    lib the library's path."""
    ctx = mp.get_context("spawn")
    with ctx.Pool(1) as pool:
        # a reference routine that stops the process (its error exit) would
        # leave apply() waiting for ever: bound the wait
        return pool.apply_async(_child, (job, bytes(cfg), args, lib)).get(timeout=600)


# ---------------------------------------------------------------------------
# Whole large steps (s_intgdrv)

def state_blocks(case):
    """The harness's base/level blocks, fc, rbct and precipitation arrays
    for a case's initial state (past = present = the initial fields)."""
    sc = case.cfg
    shp = (sc.ni + 2, sc.nj + 2, sc.nk)
    base = np.zeros(shp + (len(BASE),), np.float32, order="F")
    for n, b in enumerate(BASE):
        base[..., n] = case.base[b]
    lev = np.zeros(shp + (len(LEV), 3), np.float32, order="F")
    for lv in (0, 1):
        for k, a in case.init.items():
            lev[..., LEV.index(k), lv] = a
    fc = np.asfortranarray(case.fc.astype(np.float32))
    rbct = np.asfortranarray(case.rbct.astype(np.float32))
    prwtr = np.zeros((shp[0], shp[1], 2, 2), np.float32, order="F")
    price = np.zeros((shp[0], shp[1], 2, 3), np.float32, order="F")
    return base, fc, rbct, lev, prwtr, price


def _job_intgdrv(lib, sc_vals, blocks, nsteps):
    ni, nj, nk, ksp0, nsstp, dtbig, dtsml, moist = sc_vals
    base, fc, rbct, lev, prwtr, price = blocks
    out = []
    for s in range(nsteps):
        first = s == 0
        dtb = np.float32(0.5) * np.float32(dtbig) if first else np.float32(dtbig)
        dts = np.float32(0.5) * np.float32(dtsml) if first else np.float32(dtsml)
        ctime = 0 if first else int(round(1000 * (dtbig * s)))
        lib.ref_intgdrv(C.c_int(ni), C.c_int(nj), C.c_int(nk), C.c_int(ksp0[0]), C.c_int(ksp0[1]),
                        C.c_int(nsstp), C.c_int64(ctime), C.c_float(dtb), C.c_float(dts),
                        C.c_int(moist), base, fc, rbct, lev, prwtr, price)
        out.append((lev[..., 1].copy(), prwtr.copy(), price.copy()))
    return out


def intgdrv(case, nsteps, lib=LIB):
    """nsteps large steps of the reference's s_intgdrv from the case's
    initial state, cold start (the first step forward with half intervals).
    Returns per step (present level (ni+2,nj+2,nk,16) in LEV order, prwtr,
    price)."""
    sc = case.cfg
    cfg = po.cfg_from_solver_config(sc)
    moist = 1 if sc.cphopt != 0 or np.any(case.init.get("qv", 0.0)) else 0
    vals = (sc.ni, sc.nj, sc.nk, (sc.ksp0[0], sc.ksp0[1]), sc.nsstp, float(sc.dtbig),
            float(sc.dtsml), moist)
    return run(_job_intgdrv, cfg, vals, state_blocks(case), nsteps, lib=lib)


# ---------------------------------------------------------------------------
# Single routines

def _job_calls(lib, calls):
    """calls: [(entry point, [args])]; numpy arrays are passed by address
    and come back with what the routine wrote."""
    for name, args in calls:
        getattr(lib, name)(*args)
    return [args for _, args in calls]


def calls(cfg, seq, lib=LIB):
    """Run the harness entry points in seq = [(name, [args]), ...] one after
    another in one child process configured for cfg (an oracle Cfg);
    returns each call's argument list after the call."""
    return run(_job_calls, cfg, seq, lib=lib)


# ---------------------------------------------------------------------------
# Base-state setup (s_stretch, s_jacobian, s_intrpsnd, s_setbase, s_baserho,
# s_var8uvw, s_sndwave, s_vspdmp)

def setup_base(sc, snd):
    """The reference's FP32 base-state chain over flat terrain for a solver
    config sc and a sounding dict (cases.Case.sounding: arrays z, u, v, p,
    pt, qv and scalars zsfc, dzmin, layer1, layer2, botbar, vspbar).
    Returns (zsth (nk,), base block (ni+2,nj+2,nk,15) in BASE order,
    rbct (ni,nj,nk,2), ksp0 (2,))."""
    cfg = po.cfg_from_solver_config(sc)
    ni, nj, nk = sc.ni, sc.nj, sc.nk
    arrs = [np.ascontiguousarray(snd[k], np.float32) for k in ("z", "u", "v", "p", "pt", "qv")]
    zsth = np.zeros(nk, np.float32)
    base = np.zeros((ni + 2, nj + 2, nk, len(BASE)), np.float32, order="F")
    rbct = np.zeros((ni, nj, nk, 2), np.float32, order="F")
    ksp0 = np.zeros(2, np.int32)
    f = lambda k: float(snd.get(k, 0.0))
    args = [ni, nj, nk, sc.sthopt, float(sc.dz), f("zsfc"), f("dzmin"), f("layer1"), f("layer2"),
            len(arrs[0]) - 1] + arrs + [f("botbar"), f("vspbar"), zsth, base, rbct, ksp0]
    out = calls(cfg, [("ref_setup_base", args)])[0]
    return out[-4], out[-3], out[-2], out[-1]
