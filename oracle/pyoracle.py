"""ctypes access to the oracle libraries -- TEST INFRASTRUCTURE ONLY.

Loaded by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg to
check or time against the product; never imported by cress_amd.

  Oracle(prec)  -- the C restatement (liboracle32.so / liboracle64.so)
  Ref()         -- the reference's own FP32 routines (libcressref32.so),
                   present only where the reference was available at build time
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REFDIR = os.path.join(HERE, "_ref")

I = C.c_int
P = C.c_void_p


class Cfg(C.Structure):
    _fields_ = [("ni", I), ("nj", I), ("nk", I),
                ("wbc", I), ("ebc", I), ("sbc", I), ("nbc", I), ("bbc", I), ("tbc", I),
                ("advopt", I), ("impopt", I), ("gwmopt", I), ("buyopt", I), ("divopt", I),
                ("cphopt", I), ("haiopt", I), ("vspopt", I), ("coropt", I), ("sthopt", I),
                ("iwest", I), ("ieast", I), ("jsouth", I), ("jnorth", I),
                ("ksp0", I * 2), ("vspvar", C.c_char * 10),
                ("dx", C.c_double), ("dy", C.c_double), ("dz", C.c_double),
                ("weicoe", C.c_double), ("filcoe", C.c_double), ("thresq", C.c_double),
                ("gwave", C.c_double), ("smtopt", I), ("smhcoe", C.c_double),
                ("smvcoe", C.c_double)]


def make_cfg(ni, nj, nk, dx=1000.0, dy=1000.0, dz=500.0, advopt=3, divopt=1, buyopt=1,
             cphopt=2, vspopt=2, coropt=0, sthopt=1, bbc=2, tbc=2, ksp0=(2, 2),
             vspvar=b"oooxoooooo", weicoe=0.6, filcoe=0.1, thresq=1e-12, wbc=-1, gwave=0.0,
             smtopt=0, smhcoe=0.0, smvcoe=0.0):
    """wbc: -1 periodic or 7 open (radiation with the constant phase speed
    gwave) on all four sides."""
    c = Cfg()
    c.ni, c.nj, c.nk = ni, nj, nk
    c.wbc = c.ebc = c.sbc = c.nbc = wbc
    c.bbc, c.tbc = bbc, tbc
    c.advopt, c.impopt, c.gwmopt, c.buyopt, c.divopt = advopt, 1, 0, buyopt, divopt
    c.cphopt, c.haiopt, c.vspopt, c.coropt, c.sthopt = cphopt, 0, vspopt, coropt, sthopt
    # 4th-order reduction next to non-periodic sides (sethalo.f90:536-580)
    c.iwest = c.ieast = c.jsouth = c.jnorth = 0 if wbc == -1 else 1
    c.gwave = gwave
    c.ksp0[0], c.ksp0[1] = ksp0
    c.vspvar = vspvar
    c.dx, c.dy, c.dz = dx, dy, dz
    c.weicoe, c.filcoe, c.thresq = weicoe, filcoe, thresq
    c.smtopt, c.smhcoe, c.smvcoe = smtopt, smhcoe, smvcoe
    return c


def ptr(a):
    if a is None:
        return None
    if not a.flags["F_CONTIGUOUS"]:
        raise ValueError("arrays passed to the oracle must be Fortran-ordered")
    return a.ctypes.data_as(P)


class _Lib:
    def __init__(self, path):
        if not os.path.exists(path):
            raise FileNotFoundError(path)
        self.lib = C.CDLL(path)

    def __getattr__(self, name):
        fn = getattr(self.lib, name)

        def call(*args):
            conv = []
            for a in args:
                if isinstance(a, np.ndarray):
                    conv.append(ptr(a))
                elif isinstance(a, C.Structure):
                    conv.append(C.byref(a))
                elif isinstance(a, float):
                    conv.append(self._real(a))
                elif isinstance(a, (bytes,)):
                    conv.append(C.c_char_p(a))
                else:
                    conv.append(a)
            return fn(*conv)
        return call


class Oracle(_Lib):
    def __init__(self, prec=32):
        super().__init__(os.path.join(REFDIR, "liboracle%d.so" % prec))
        self.prec = prec
        self.dtype = np.float32 if prec == 32 else np.float64
        self._real = C.c_float if prec == 32 else C.c_double
        assert self.lib.or_real_size() == (4 if prec == 32 else 8)


class Ref(_Lib):
    """The reference's own FP32 routines via oracle/ref_harness.f90."""

    def __init__(self):
        super().__init__(os.path.join(REFDIR, "libcressref32.so"))
        self.dtype = np.float32
        self._real = C.c_float

    def configure(self, cfg):
        f = C.c_float
        self.lib.ref_configure(I(cfg.wbc), I(cfg.ebc), I(cfg.sbc), I(cfg.nbc), I(cfg.bbc),
                               I(cfg.tbc), I(cfg.advopt), I(cfg.impopt), I(cfg.gwmopt),
                               I(cfg.buyopt), I(cfg.divopt), I(cfg.cphopt), I(cfg.haiopt),
                               I(cfg.vspopt), I(cfg.coropt), f(cfg.dx), f(cfg.dy), f(cfg.dz),
                               f(cfg.weicoe), f(cfg.filcoe), f(cfg.thresq),
                               C.c_char_p(bytes(cfg.vspvar).ljust(10, b"x")))


def have_ref():
    return os.path.exists(os.path.join(REFDIR, "libcressref32.so"))


def field(cfg, dtype, n=1):
    """Zeroed Fortran-ordered real(0:ni+1,0:nj+1,1:nk[,n]) block."""
    shape = (cfg.ni + 2, cfg.nj + 2, cfg.nk) + ((n,) if n > 1 else ())
    return np.zeros(shape, dtype=dtype, order="F")


BASE = ["jcb", "jcb8u", "jcb8v", "jcb8w", "ubr", "vbr", "pbr", "ptbr", "qvbr", "rbr", "rst",
        "rst8u", "rst8v", "rst8w", "rcsq"]
PROG = ["u", "v", "w", "pp", "ptp", "q", "nci"]
QNAMES = ["qv", "qc", "qr", "qi", "qs", "qg"]
NSCR = 24


class State(C.Structure):
    _fields_ = ([(b, P) for b in BASE] + [("fc", P), ("rbct", P)] +
                [(p, P * 3) for p in PROG] + [("prwtr", P), ("price", P), ("scr", P), ("cp", P)])


def cfg_from_solver_config(sc):
    """Oracle Cfg equivalent of a cress_amd.host.Config."""
    c = make_cfg(sc.ni, sc.nj, sc.nk, dx=sc.dx, dy=sc.dy, dz=sc.dz, advopt=sc.advopt,
                 divopt=sc.divopt, buyopt=sc.buyopt, cphopt=sc.cphopt, vspopt=sc.vspopt,
                 coropt=sc.coropt, sthopt=sc.sthopt, bbc=sc.bbc, tbc=sc.tbc,
                 ksp0=(sc.ksp0[0], sc.ksp0[1]), vspvar=bytes(sc.vspvar), weicoe=sc.weicoe,
                 filcoe=sc.filcoe, thresq=sc.thresq, wbc=sc.wbc, gwave=sc.gwave,
                 smtopt=sc.smtopt, smhcoe=sc.smhcoe, smvcoe=sc.smvcoe)
    return c


class Model:
    """The oracle's large-step driver (or_large_step) on numpy arrays."""

    def __init__(self, case, prec=32):
        self.o = Oracle(prec)
        self.dt = self.o.dtype
        sc = case.cfg
        self.sc = sc
        self.cfg = cfg_from_solver_config(sc)
        shp = (sc.ni + 2, sc.nj + 2, sc.nk)
        self.shape = shp
        n = shp[0] * shp[1] * shp[2]
        self.n = n
        f = lambda *extra: np.zeros(shp + extra, dtype=self.dt, order="F")
        self.base = {k: np.asfortranarray(case.base[k].astype(self.dt)) for k in BASE}
        self.fc = np.asfortranarray(case.fc.astype(self.dt))
        self.rbct = np.asfortranarray(case.rbct.astype(self.dt))
        self.lv = {p: [f() if p != "q" else f(6) for _ in range(3)] for p in PROG}
        for lv in (0, 1):
            for k, a in case.init.items():
                if k in QNAMES:
                    self.lv["q"][lv][..., QNAMES.index(k)] = a
                else:
                    self.lv[k][lv][...] = a
        self.prwtr = np.zeros((shp[0], shp[1], 2, 2), self.dt, order="F")
        self.price = np.zeros((shp[0], shp[1], 2, 3), self.dt, order="F")
        self.scr = np.zeros(NSCR * n, self.dt)
        st = State()
        for k in BASE:
            setattr(st, k, self.base[k].ctypes.data)
        st.fc, st.rbct = self.fc.ctypes.data, self.rbct.ctypes.data
        for p in PROG:
            arr = getattr(st, p)
            for l in range(3):
                arr[l] = self.lv[p][l].ctypes.data
        st.prwtr, st.price, st.scr = self.prwtr.ctypes.data, self.price.ctypes.data, self.scr.ctypes.data
        # open-boundary phase speeds: zero until the first step's phasev
        # (allocslv.f90:2015 zeroes them; a cold start reads no restart)
        self.cp = None
        if sc.wbc == 7:
            self.o.lib.or_cp_size.restype = C.c_size_t
            self.cp = np.zeros(self.o.lib.or_cp_size(C.byref(self.cfg)), self.dt)
            st.cp = self.cp.ctypes.data
        self.state = st

    def step(self, nsteps=1, first=False):
        for s in range(nsteps):
            rc = self.o.lib.or_large_step(C.byref(self.cfg), C.byref(self.state),
                                          1 if (first and s == 0) else 0,
                                          C.c_double(self.sc.dtbig), C.c_double(self.sc.dtsml),
                                          self.sc.nsstp)
            if rc != 0:
                raise RuntimeError("or_large_step failed: %d" % rc)

    def get(self, name, level=1):
        """A prognostic at a time level (0 past, 1 present, 2 future), or the
        precipitation arrays 'prwtr' (rate/accumulation x cloud/rain) and
        'price' (rate/accumulation x ice/snow/graupel)."""
        if name in ("prwtr", "price"):
            return getattr(self, name)
        if name in QNAMES:
            return self.lv["q"][level][..., QNAMES.index(name)]
        return self.lv[name][level]
