"""Host-side mirror of the reference's step interface over the C ABI.

`Solver` plays the role of the reference's time-loop body
(slvdrv.f90:975-1452 -> s_intgdrv): it owns one device-resident context
(include/cress_hip.h, tier 1), takes the base state and the three time levels
once, and advances whole large steps on the GPU.  Arrays are the reference's
real(0:ni+1,0:nj+1,1:nk) blocks in Fortran order (numpy order="F").

The library is loaded from cress_amd/_lib/libcresship.so; there is no CPU
fallback.  Importing this module on a machine without that library raises.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_lib", "libcresship.so")

I = C.c_int
D = C.c_double

# field ids (include/cress_hip.h)
BASE_FIELDS = ["jcb", "jcb8u", "jcb8v", "jcb8w", "ubr", "vbr", "pbr", "ptbr", "qvbr", "rbr",
               "rst", "rst8u", "rst8v", "rst8w", "rcsq"]
FID = {name: i for i, name in enumerate(BASE_FIELDS)}
PROG_FIELDS = ["u", "v", "w", "pp", "ptp", "qv", "qc", "qr", "qi", "qs", "qg", "nci"]
FID.update({name: 32 + i for i, name in enumerate(PROG_FIELDS)})
FID.update({"fc": 64, "rbct": 65, "prwtr": 66, "price": 67})
# cress_diag's result order (include/cress_hip.h)
DIAG_NAMES = ["vapour", "liquid", "ice", "precip", "dry", "ke", "ie", "pe", "theta",
              "wmax", "wmin", "ptpmax", "ptpmin"]
NDIAG = len(DIAG_NAMES)
# cress_base_columns' order
SETUP_COLS = ["zph", "zph8s", "ubr", "vbr", "pbr", "ptbr", "qvbr", "rbr", "jcb", "jcb8w", "rst", "rst8w", "rcsq",
              "rbct", "ptvbr", "pibr"]
NSETUPCOL = len(SETUP_COLS)


class Sounding(C.Structure):
    """cress_sounding (include/cress_hip.h)."""
    _fields_ = [("nlev", C.c_int)] + [(k, C.POINTER(C.c_double)) for k in ("z", "u", "v", "p", "pt", "qv")] + \
               [(k, C.c_double) for k in ("zsfc", "dzmin", "layer1", "layer2", "botbar", "vspbar", "fc")]


# launch shapes (cress_autotune / cress_set_launch order)
LAUNCH_NAMES = ["ptuv_by", "ptuv_slots", "ptuv_order", "col_by", "col_lds", "fs_by", "fs_order", "ff_by",
                "ff_order"]
NLAUNCH = len(LAUNCH_NAMES)

ERR = {0: "ok", -1: "invalid argument", -2: "no usable GPU", -3: "HIP runtime error",
       -4: "device out of memory", -5: "field id or level out of range",
       -6: "option outside the implemented plain path", -7: "grid larger than a fixed-size buffer"}


class CressError(RuntimeError):
    def __init__(self, code, what):
        super().__init__("%s: %s (%d)" % (what, ERR.get(code, "error"), code))
        self.code = code


class Config(C.Structure):
    _fields_ = [("ni", I), ("nj", I), ("nk", I),
                ("wbc", I), ("ebc", I), ("sbc", I), ("nbc", I), ("bbc", I), ("tbc", I),
                ("advopt", I), ("impopt", I), ("gwmopt", I), ("buyopt", I), ("divopt", I),
                ("cphopt", I), ("haiopt", I), ("vspopt", I), ("coropt", I), ("sthopt", I),
                ("ksp0", I * 2), ("vspvar", C.c_char * 10), ("nsstp", I),
                ("dx", D), ("dy", D), ("dz", D), ("dtbig", D), ("dtsml", D),
                ("weicoe", D), ("filcoe", D), ("thresq", D), ("gwave", D),
                ("smtopt", I), ("smhcoe", D), ("smvcoe", D)]


_lib = None


def lib():
    """The loaded product library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError("cress_amd: %s not built (run __graft_entry__.build())" % LIB_PATH)
        L = C.CDLL(LIB_PATH)
        P = C.c_void_p
        L.cress_ctx_create.argtypes = [C.POINTER(Config), C.POINTER(P)]
        L.cress_ctx_destroy.argtypes = [P]
        L.cress_upload.argtypes = [P, I, I, P]
        L.cress_download.argtypes = [P, I, I, P]
        L.cress_field_size.argtypes = [P, I, C.POINTER(C.c_size_t)]
        L.cress_field_device_ptr.argtypes = [P, I, I, C.POINTER(P)]
        L.cress_field_pitch.argtypes = [P, C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]
        L.cress_large_step.argtypes = [P, I, I]
        L.cress_sync.argtypes = [P]
        L.cress_set_timing.argtypes = [P, I]
        L.cress_stage_times.argtypes = [P, C.POINTER(D), I]
        L.cress_device_count.argtypes = [C.POINTER(I)]
        L.cress_error_string.restype = C.c_char_p
        L.cress_gaussel.argtypes = [I] * 9 + [P] * 4
        L.cress_advs.argtypes = [C.POINTER(Config)] + [P] * 5
        L.cress_bcycle.argtypes = [C.POINTER(Config)] + [I] * 9 + [P]
        L.cress_advuvw.argtypes = [C.POINTER(Config)] + [P] * 9
        L.cress_advp.argtypes = [C.POINTER(Config)] + [P] * 8
        L.cress_hevi.argtypes = [C.POINTER(Config), C.c_float] + [P] * 19
        L.cress_steps.argtypes = [C.POINTER(Config), C.c_float] + [P] * 14
        L.cress_cloudphy.argtypes = [C.POINTER(Config), C.c_float] + [P] * 18
        L.cress_turbcoef.argtypes = [C.POINTER(Config), C.c_float, I] + [P] * 28
        L.cress_turbdrv.argtypes = [C.POINTER(Config), C.c_float, I] + [P] * 26
        L.cress_turbflx.argtypes = [C.POINTER(Config)] + [P] * 8
        L.cress_turbs.argtypes = [C.POINTER(Config)] + [P] * 6
        L.cress_kh8uv.argtypes = [C.POINTER(Config)] + [P] * 3
        L.cress_forcetke.argtypes = [C.POINTER(Config)] + [P] * 18
        L.cress_stepstke.argtypes = [C.POINTER(Config), C.c_float] + [P] * 3
        L.cress_bc8u.argtypes = [C.POINTER(Config), I, P]
        L.cress_bc8v.argtypes = [C.POINTER(Config), I, P]
        L.cress_bc8w.argtypes = [C.POINTER(Config), P]
        L.cress_ens_create.argtypes = [C.POINTER(P), I, C.POINTER(P)]
        L.cress_ens_step.argtypes = [P, I, I]
        L.cress_ens_sync.argtypes = [P]
        L.cress_ens_destroy.argtypes = [P]
        L.cress_mlp_create.argtypes = [C.c_long, I, C.POINTER(P)]
        L.cress_mlp_destroy.argtypes = [P]
        L.cress_mlp_init.argtypes = [P, C.c_ulonglong]
        L.cress_mlp_set.argtypes = [P, P, P]
        L.cress_mlp_get.argtypes = [P, P, P]
        L.cress_mlp_samples.argtypes = [P, C.POINTER(C.c_long)]
        L.cress_mlp_clear_samples.argtypes = [P]
        L.cress_mlp_put_samples.argtypes = [P, P, P, C.c_long]
        L.cress_mlp_get_samples.argtypes = [P, C.c_long, C.c_long, P, P]
        L.cress_mlp_fit_norm.argtypes = [P, C.c_float]
        L.cress_mlp_forward.argtypes = [P, P, C.c_long, P]
        L.cress_mlp_grad.argtypes = [P, P, P, I, P, C.POINTER(D)]
        L.cress_mlp_train.argtypes = [P, I, I, C.c_float, C.c_uint, P]
        L.cress_ctx_set_mlp.argtypes = [P, P, I]
        L.cress_ctx_mlp_cells.argtypes = [P, C.POINTER(C.c_long)]
        L.cress_mlp_nparam.argtypes = []
        L.cress_mlp_nnorm.argtypes = []
        L.cress_diag.argtypes = [P, I, C.POINTER(D)]
        L.cress_autotune.argtypes = [P, C.POINTER(I), I]
        L.cress_set_launch.argtypes = [P, C.POINTER(I), I]
        L.cress_get_launch.argtypes = [P, C.POINTER(I), I]
        L.cress_set_base.argtypes = [P, C.POINTER(Sounding)]
        L.cress_base_columns.argtypes = [P, C.POINTER(D), I]
        L.cress_gaussel_noalloc.argtypes = [I] * 9 + [P] * 4
        _lib = L
    return _lib


def check(rc, what):
    if rc != 0:
        raise CressError(rc, what)


def make_config(ni, nj, nk, dx, dy, dz, dtbig, dtsml, nsstp=None, advopt=3, divopt=1, buyopt=1,
                cphopt=2, vspopt=2, coropt=0, sthopt=1, bbc=2, tbc=2, ksp0=(2, 2),
                vspvar=b"oooxoooooo", weicoe=0.6, filcoe=0.1, thresq=1e-12, wbc=-1,
                gwave=0.0, smtopt=0, smhcoe=0.0, smvcoe=0.0):
    """wbc: -1 periodic or 7 open (radiation at the constant phase speed
    gwave, m/s) on all four sides."""
    c = Config()
    c.ni, c.nj, c.nk = ni, nj, nk
    c.wbc = c.ebc = c.sbc = c.nbc = wbc
    c.bbc, c.tbc = bbc, tbc
    c.advopt, c.impopt, c.gwmopt, c.buyopt, c.divopt = advopt, 1, 0, buyopt, divopt
    c.cphopt, c.haiopt, c.vspopt, c.coropt, c.sthopt = cphopt, 0, vspopt, coropt, sthopt
    c.ksp0[0], c.ksp0[1] = ksp0
    c.vspvar = vspvar
    if nsstp is None:
        # slvstep.f90:176: nsstp = 20*int(100*(dtbig+.001))/int(1000*(dtsml+.0001))
        nsstp = 20 * int(1e2 * (dtbig + 0.001)) // int(1e3 * (dtsml + 0.0001))
    c.nsstp = nsstp
    c.dx, c.dy, c.dz, c.dtbig, c.dtsml = dx, dy, dz, dtbig, dtsml
    c.weicoe, c.filcoe, c.thresq = weicoe, filcoe, thresq
    c.gwave = gwave
    c.smtopt, c.smhcoe, c.smvcoe = smtopt, smhcoe, smvcoe
    return c


class Solver:
    """One device-resident CReSS model on cuda:0 (HIP device 0)."""

    def __init__(self, cfg):
        self.cfg = cfg
        self.L = lib()
        self.ctx = C.c_void_p()
        check(self.L.cress_ctx_create(C.byref(cfg), C.byref(self.ctx)), "cress_ctx_create")
        self.shape = (cfg.ni + 2, cfg.nj + 2, cfg.nk)

    def close(self):
        if self.ctx:
            self.L.cress_ctx_destroy(self.ctx)
            self.ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def size(self, name):
        n = C.c_size_t()
        check(self.L.cress_field_size(self.ctx, FID[name], C.byref(n)), "cress_field_size")
        return n.value

    def upload(self, name, arr, level=0):
        a = np.asarray(arr, dtype=np.float32)
        a = np.asfortranarray(a)
        if a.size != self.size(name):
            raise ValueError("%s: expected %d floats, got %d" % (name, self.size(name), a.size))
        check(self.L.cress_upload(self.ctx, FID[name], level, a.ctypes.data), "upload " + name)

    def download(self, name, level=0):
        out = np.empty(self.size(name), dtype=np.float32)
        check(self.L.cress_download(self.ctx, FID[name], level, out.ctypes.data), "download " + name)
        if name in FID and FID[name] < 64:
            return out.reshape(self.shape, order="F")
        if name in ("prwtr", "price"):
            return out.reshape(self.shape[:2] + (2, 2 if name == "prwtr" else 3), order="F")
        return out

    def pitch(self):
        """(row, plane) float strides of the 3-D fields on the device."""
        r, p = C.c_size_t(), C.c_size_t()
        check(self.L.cress_field_pitch(self.ctx, C.byref(r), C.byref(p)), "cress_field_pitch")
        return r.value, p.value

    def device_ptr(self, name, level=0):
        p = C.c_void_p()
        check(self.L.cress_field_device_ptr(self.ctx, FID[name], level, C.byref(p)), "device_ptr")
        return p.value

    def step(self, nsteps=1, first=False):
        check(self.L.cress_large_step(self.ctx, nsteps, 1 if first else 0), "cress_large_step")

    def sync(self):
        check(self.L.cress_sync(self.ctx), "cress_sync")

    def set_timing(self, on=True):
        check(self.L.cress_set_timing(self.ctx, 1 if on else 0), "cress_set_timing")

    def set_mlp(self, mlp, mode):
        """mode 0: srcblk; 1: srcblk, recording training pairs into mlp;
        2: the network for the active cells inside its envelope."""
        check(self.L.cress_ctx_set_mlp(self.ctx, mlp.m if mlp is not None else None, mode), "cress_ctx_set_mlp")

    def diag(self, level=1):
        """FP64 mass and energy budgets of a level (cress_diag): a dict of
        vapour, liquid, ice, precip (kg), dry (kg), ke, ie, pe (J), theta
        (sum rst theta' dV), wmax, wmin, ptpmax, ptpmin."""
        buf = (D * NDIAG)()
        check(self.L.cress_diag(self.ctx, level, buf), "cress_diag")
        return dict(zip(DIAG_NAMES, list(buf)))

    def set_base(self, snd):
        """Base state and metrics from a 1-D sounding, in binary64 on the
        device (cress_set_base).  snd: dict with arrays z, u, v, p, pt, qv
        (levels 0..nlev) and scalars zsfc, dzmin, layer1, layer2, botbar,
        vspbar, fc."""
        arrs = {k: np.ascontiguousarray(snd[k], np.float64) for k in ("z", "u", "v", "p", "pt", "qv")}
        s = Sounding()
        s.nlev = len(arrs["z"]) - 1
        for k, a in arrs.items():
            setattr(s, k, a.ctypes.data_as(C.POINTER(D)))
        for k in ("zsfc", "dzmin", "layer1", "layer2", "botbar", "vspbar", "fc"):
            setattr(s, k, float(snd.get(k, 0.0)))
        check(self.L.cress_set_base(self.ctx, C.byref(s)), "cress_set_base")
        return s

    def base_columns(self):
        """cress_set_base's binary64 per-level columns: dict name -> (nk+1,)
        array (slot 0 unused)."""
        nk = self.cfg.nk
        out = np.empty((NSETUPCOL, nk + 1), np.float64)
        check(self.L.cress_base_columns(self.ctx, out.ctypes.data_as(C.POINTER(D)), NSETUPCOL), "cress_base_columns")
        return dict(zip(SETUP_COLS, out))

    def autotune(self):
        """Time every candidate launch shape on this grid and keep the
        fastest (cress_autotune); returns the chosen shapes."""
        buf = (C.c_int * NLAUNCH)()
        check(self.L.cress_autotune(self.ctx, buf, NLAUNCH), "cress_autotune")
        return dict(zip(LAUNCH_NAMES, list(buf)))

    def set_launch(self, shapes):
        """Set the launch shapes (a dict with LAUNCH_NAMES keys)."""
        buf = (C.c_int * NLAUNCH)(*[int(shapes[k]) for k in LAUNCH_NAMES])
        check(self.L.cress_set_launch(self.ctx, buf, NLAUNCH), "cress_set_launch")

    def get_launch(self):
        buf = (C.c_int * NLAUNCH)()
        check(self.L.cress_get_launch(self.ctx, buf, NLAUNCH), "cress_get_launch")
        return dict(zip(LAUNCH_NAMES, list(buf)))

    def mlp_cells(self):
        """Cells the network handled in the last large step (mode 2)."""
        c = C.c_long()
        check(self.L.cress_ctx_mlp_cells(self.ctx, C.byref(c)), "cress_ctx_mlp_cells")
        return c.value

    def stage_times(self):
        buf = (D * 8)()
        check(self.L.cress_stage_times(self.ctx, buf, 8), "cress_stage_times")
        return dict(zip(["force", "acoustic", "steps", "micro", "timeflt", "satadjst", "halo2",
                         "other"], list(buf)))


class Ensemble:
    """Solvers of any sizes stepped together: each cycle of large steps of
    every member is one launch of a captured hipGraph (cress_ens_*)."""

    def __init__(self, solvers):
        self.members = list(solvers)
        self.L = lib()
        arr = (C.c_void_p * len(self.members))(*[m.ctx for m in self.members])
        self.e = C.c_void_p()
        check(self.L.cress_ens_create(arr, len(self.members), C.byref(self.e)), "cress_ens_create")

    def step(self, nsteps=1, first=False):
        check(self.L.cress_ens_step(self.e, nsteps, 1 if first else 0), "cress_ens_step")

    def sync(self):
        check(self.L.cress_ens_sync(self.e), "cress_ens_sync")

    def close(self):
        if self.e:
            self.L.cress_ens_destroy(self.e)
            self.e = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Mlp:
    """The learned microphysics network (cress_mlp_*): parameters, training
    pairs and the training kernels on the GPU."""

    NIN, NOUT = 9, 7

    def __init__(self, capacity=1 << 20, max_batch=1 << 15):
        self.L = lib()
        self.m = C.c_void_p()
        check(self.L.cress_mlp_create(capacity, max_batch, C.byref(self.m)), "cress_mlp_create")
        self.nparam = self.L.cress_mlp_nparam()
        self.nnorm = self.L.cress_mlp_nnorm()

    def close(self):
        if self.m:
            self.L.cress_mlp_destroy(self.m)
            self.m = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def init(self, seed):
        check(self.L.cress_mlp_init(self.m, seed), "cress_mlp_init")

    def set(self, params=None, norm=None):
        p = None if params is None else np.ascontiguousarray(params, np.float32)
        n = None if norm is None else np.ascontiguousarray(norm, np.float32)
        check(self.L.cress_mlp_set(self.m, None if p is None else p.ctypes.data,
                                   None if n is None else n.ctypes.data), "cress_mlp_set")

    def get(self):
        p = np.empty(self.nparam, np.float32)
        n = np.empty(self.nnorm, np.float32)
        check(self.L.cress_mlp_get(self.m, p.ctypes.data, n.ctypes.data), "cress_mlp_get")
        return p, n

    def samples(self):
        c = C.c_long()
        check(self.L.cress_mlp_samples(self.m, C.byref(c)), "cress_mlp_samples")
        return c.value

    def clear_samples(self):
        check(self.L.cress_mlp_clear_samples(self.m), "cress_mlp_clear_samples")

    def put_samples(self, inp, tgt):
        a = np.ascontiguousarray(inp, np.float32)
        b = np.ascontiguousarray(tgt, np.float32)
        check(self.L.cress_mlp_put_samples(self.m, a.ctypes.data, b.ctypes.data, a.shape[0]),
              "cress_mlp_put_samples")

    def get_samples(self, first, n):
        a = np.empty((n, self.NIN), np.float32)
        b = np.empty((n, self.NOUT), np.float32)
        check(self.L.cress_mlp_get_samples(self.m, first, n, a.ctypes.data, b.ctypes.data),
              "cress_mlp_get_samples")
        return a, b

    def fit_norm(self, margin=0.05):
        check(self.L.cress_mlp_fit_norm(self.m, margin), "cress_mlp_fit_norm")

    def forward(self, inp):
        a = np.ascontiguousarray(inp, np.float32)
        out = np.empty((a.shape[0], self.NOUT), np.float32)
        check(self.L.cress_mlp_forward(self.m, a.ctypes.data, a.shape[0], out.ctypes.data), "cress_mlp_forward")
        return out

    def grad(self, inp, tgt):
        a = np.ascontiguousarray(inp, np.float32)
        b = np.ascontiguousarray(tgt, np.float32)
        g = np.empty(self.nparam, np.float32)
        loss = D()
        check(self.L.cress_mlp_grad(self.m, a.ctypes.data, b.ctypes.data, a.shape[0], g.ctypes.data,
                                    C.byref(loss)), "cress_mlp_grad")
        return g, loss.value

    def train(self, steps, batch, lr, seed=1):
        losses = np.empty(steps, np.float64)
        check(self.L.cress_mlp_train(self.m, steps, batch, lr, seed, losses.ctypes.data), "cress_mlp_train")
        return losses


def device_count():
    n = I(0)
    lib().cress_device_count(C.byref(n))
    return n.value
