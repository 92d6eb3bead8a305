"""Grid, base state and initial states of the benchmark cases (host, FP64).

This is the setup that feeds the large step (the reference's s_setgrid /
s_inivar: stretch.f90, jacobian.f90, var8uvw.f90, baserho.f90, sndwave.f90,
vspdmp.f90, getpt0.f90), evaluated once in float64 with numpy and handed to
the device in float32.  It is not on the timed path.

Cases (names follow the workload each one is):
  bubble    -- moist warm bubble, Weisman-Klemp sounding, no wind, 64x64x40,
               dx=dy=1 km, uniform dz=500 m, dtbig=5 s
  supercell -- WK sounding + quarter-circle hodograph, 256x256x64, dx=500 m,
               stretched dz 100 m -> 18 km top, dtbig=3 s, 6 substeps (run
               here with periodic sides, the only lateral BC implemented)
  vortex    -- 1024x1024x80 tropical sounding with a warm-core vortex on an
               f-plane (f=5e-5), stretched dz 50 m -> 20 km, dtbig=4 s,
               8 substeps (the bandwidth-stress workload)
Every case takes nx, ny, nz overrides so that the same physics can be run at
oracle-sized grids.
"""
import math

import numpy as np

from . import host

G, CP, CV, RD, RV, P0 = 9.8, 1004.0, 717.0, 287.0, 461.0, 1.0e5
EPSVA = 0.622
CC = 3.141592   # commath.f90:42


# --------------------------------------------------------------------------
# vertical grid: s_stretch (stretch.f90:174-400), sthopt 0 (uniform) / 1 (cubic)
# --------------------------------------------------------------------------
def stretch_levels(nk, dz, sthopt=0, dzmin=None, layer1=0.0, layer2=None, zsfc=0.0):
    """Physical heights zsth(1..nk) of the zeta (w) levels, zeta(k)=(k-2)*dz."""
    ztop2 = zsfc + (nk - 3) * dz
    if sthopt == 0 or dzmin is None or dzmin >= dz:
        return np.array([zsfc + (k - 2) * dz for k in range(1, nk + 1)], dtype=np.float64)
    layer2 = ztop2 if layer2 is None else layer2
    z1 = zsfc + max(min(ztop2 - zsfc, layer1), 0.0)
    z2 = z1 + max(min(ztop2 - z1, layer2), 0.0)
    z1 = max(zsfc, z1)
    z2 = min(ztop2, z2)
    klow = int((z1 - zsfc) / dzmin + 0.1)
    if z2 >= ztop2:
        dzmid = (ztop2 - zsfc - klow * dzmin) / (nk - 3 - klow)
        dzhigh = 2.0 * dzmid - dzmin
        kmid = 0
    else:
        a = 2.0 * (nk - 3 - klow)
        b = 2.0 * zsfc - ztop2 - z2 - (nk - 3 - 3 * klow) * dzmin
        c = dzmin * (z2 - zsfc - klow * dzmin)
        dzmid = 0.5 * (math.sqrt(b * b - 4.0 * a * c) - b) / a
        kmid = int((ztop2 - z2) / (2.0 * dzmid - dzmin) + 0.1)
        a = zsfc + (klow - kmid) * dzmin + (nk - 3 - klow + kmid) * dzmid
        if kmid != 0:
            dzhigh = (2.0 * dzmid - dzmin) + (ztop2 - a) / kmid
        else:
            dzmid = dzmid + (ztop2 - a) / (nk - 3 - klow - kmid)
            dzhigh = 2.0 * dzmid - dzmin
    a = dzmid - dzmin
    b = 1.0 / (nk - 4 - kmid - klow)
    dzs = np.zeros(nk + 1)
    for k in range(1, klow + 2):
        dzs[k] = dzmin
    for k in range(klow + 2, nk - 1 - kmid):
        t = 2.0 * b * (k - 2 - klow) - 1.0
        dzs[k] = dzmid + a * t * t * t
    for k in range(nk - 1 - kmid, nk - 1):
        dzs[k] = dzhigh
    z = np.zeros(nk + 1)
    z[2] = zsfc
    for k in range(3, nk - 1):
        z[k] = z[k - 1] + dzs[k - 1]
    z[1] = z[2] - dzs[1]
    z[nk - 1] = ztop2
    z[nk] = z[nk - 1] + dzs[nk - 2]
    return z[1:]


# --------------------------------------------------------------------------
# soundings
# --------------------------------------------------------------------------
def tetens_qvs(t, p):
    """Saturation mixing ratio over water, Tetens as in setblk.f90:333-339."""
    es = 610.78 * np.exp(17.269 * (t - 273.16) / (t - 35.86))
    return EPSVA * es / (p - es)


def wk_sounding(z, rh_sfc=0.95):
    """Weisman-Klemp (1982): theta and RH profiles."""
    zt, tht, ttr = 12000.0, 343.0, 213.0
    th = np.where(z <= zt, 300.0 + (tht - 300.0) * (np.maximum(z, 0) / zt) ** 1.25,
                  tht * np.exp(G * (z - zt) / (CP * ttr)))
    rh = np.where(z <= zt, 1.0 - 0.75 * (np.maximum(z, 0) / zt) ** 1.25, 0.25)
    rh = np.minimum(rh, rh_sfc)
    return th, rh


def tropical_sounding(z):
    """Moist tropical profile: theta_sfc=300 K with the WK lapse structure,
    RH = 0.9 exp(-z/3 km)."""
    th, _ = wk_sounding(z)
    rh = 0.9 * np.exp(-np.maximum(z, 0.0) / 3000.0)
    return th, rh


def hydrostatic(zs, th, rh, psfc=1.0e5):
    """Exner-function integration (FP64): returns p, qv, rho at heights zs."""
    n = len(zs)
    pi = np.empty(n)
    qv = np.zeros(n)
    order = np.argsort(zs)
    k0 = order[np.argmin(np.abs(zs[order]))]
    for it in range(8):
        piv = np.empty(n)
        pi_s = (psfc / P0) ** (RD / CP)
        # integrate from z=0 up and down with the mid-point rule
        zz = np.concatenate([[0.0], zs[order]])
        thv = th * (1.0 + 0.608 * qv)
        thv_s = np.interp(zz, zs[order], thv[order])
        pis = np.empty(len(zz))
        pis[0] = pi_s
        for m in range(1, len(zz)):
            pis[m] = pis[m - 1] - G * (zz[m] - zz[m - 1]) / (CP * 0.5 * (thv_s[m] + thv_s[m - 1]))
        piv[order] = pis[1:]
        pi = piv
        p = P0 * pi ** (CP / RD)
        t = th * pi
        qv = rh * tetens_qvs(t, p)
    p = P0 * pi ** (CP / RD)
    t = th * pi
    rho = p / (RD * t * (1.0 + 0.608 * qv))
    del k0
    return p, qv, rho


# --------------------------------------------------------------------------
# deterministic noise: splitmix64 counter hash
# --------------------------------------------------------------------------
def splitmix_uniform(seed, idx):
    x = (np.uint64(seed) << np.uint64(40)) + idx.astype(np.uint64)
    with np.errstate(over="ignore"):
        x = x + np.uint64(0x9E3779B97F4A7C15)
        x = (x ^ (x >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        x = (x ^ (x >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        x = x ^ (x >> np.uint64(31))
    return (x >> np.uint64(11)).astype(np.float64) * 2.0 ** -53


# --------------------------------------------------------------------------
# assembly
# --------------------------------------------------------------------------
class Case:
    def __init__(self, cfg, base, fc, rbct, init, zph, zs, name, snd=None):
        self.cfg, self.base, self.fc, self.rbct, self.init = cfg, base, fc, rbct, init
        self.zph, self.zs, self.name = zph, zs, name
        self.snd = snd

    def sounding(self, dzf=25.0):
        """The case's sounding as a fine 1-D profile (levels every dzf m from
        the ground to 1 km above the model top: z, u, v, p, pt, qv in FP64,
        hydrostatic), with the vertical-grid, sponge and Coriolis settings:
        the input of cress_set_base (Solver.set_base)."""
        sd = self.snd
        ztop = float(self.zph[-1]) + 1000.0
        z = np.arange(0.0, ztop + dzf, dzf)
        th, rh = sd["fn"](z)
        p, qv, _ = hydrostatic(z, th, rh)
        u, v = hodograph(z, sd["shear"])
        return {"z": z, "u": u, "v": v, "p": p, "pt": th, "qv": qv, "zsfc": 0.0,
                "dzmin": sd["dzmin"] or 0.0, "layer1": 0.0, "layer2": sd["layer2"],
                "botbar": sd["botbar"], "vspbar": 0.003333, "fc": sd["fcor"]}

    def write_state(self, path):
        """The base state, fc, rbct and the past and present levels in the
        stream format the Fortran host (cress_amd/fortran/cress_run.f90)
        reads: 8 int32 header ('CRSS', 1, ni, nj, nk, ksp0(2), nrec), then
        per record int32 field id, level, nfloat and the float32 data."""
        c = self.cfg
        recs = [(host.FID[k], 0, v) for k, v in self.base.items()]
        recs += [(host.FID["fc"], 0, self.fc), (host.FID["rbct"], 0, self.rbct)]
        for lv in (0, 1):
            recs += [(host.FID[k], lv, v) for k, v in self.init.items()]
        with open(path, "wb") as f:
            hdr = [0x53535243, 1, c.ni, c.nj, c.nk, c.ksp0[0], c.ksp0[1], len(recs)]
            f.write(np.asarray(hdr, np.int32).tobytes())
            for fid, lv, v in recs:
                a = np.asarray(v, np.float32).ravel(order="F")
                f.write(np.asarray([fid, lv, a.size], np.int32).tobytes())
                f.write(a.tobytes())

    def write_userconf(self, path, nsteps, stime=0.0, extra=()):
        """A user.conf with the reference's namelist groups for this case,
        the forecast from stime to etime = stime + nsteps * dtbig, and the
        extra namelist lines given (e.g. "&outfomat ... /")."""
        c = self.cfg
        vsp = bytes(c.vspvar).decode() if not isinstance(c.vspvar, str) else c.vspvar
        lines = [
            "&dimset numpe=1, xgroup=1, ygroup=1, xsub=1, ysub=1, xdim=%d, ydim=%d, zdim=%d /"
            % (c.ni, c.nj, c.nk),
            "&gridset dx=%r, dy=%r, dz=%r /" % (c.dx, c.dy, c.dz),
            "&gridsth sthopt=%d /" % c.sthopt,
            "&terrain trnopt=0 /",
            "&flength stime=%r, etime=%r /" % (stime, stime + nsteps * c.dtbig),
            "&boundry wbc=%d, ebc=%d, sbc=%d, nbc=%d, bbc=%d, tbc=%d, gwave=%r /"
            % tuple([1 if b == -1 else b for b in (c.wbc, c.ebc, c.sbc, c.nbc)] + [c.bbc, c.tbc, c.gwave]),
            "&gpvpram vspopt=%d, vspvar='%s' /" % (c.vspopt, vsp[:10]),
            "&integrat dtbig=%r, dtsml=%r, gwmopt=%d, impopt=%d, advopt=%d, weicoe=%r, filcoe=%r /"
            % (c.dtbig, c.dtsml, c.gwmopt, c.impopt, c.advopt, c.weicoe, c.filcoe),
            "&smoother smtopt=%d, smhcoe=%r, smvcoe=%r /" % (c.smtopt, c.smhcoe, c.smvcoe),
            "&mapfcter mfcopt=0 /",
            "&coriolis coropt=%d /" % c.coropt,
            "&buoyancy buyopt=%d /" % c.buyopt,
            "&ddamping divopt=%d /" % c.divopt,
            "&cloudphy cphopt=%d, haiopt=%d, thresq=%r /" % (c.cphopt, c.haiopt, c.thresq),
            "&turbulen tubopt=0 /",
        ] + list(extra)
        with open(path, "w") as f:
            f.write("\n".join(lines) + "\n")

    def load(self, solver):
        for k, v in self.base.items():
            solver.upload(k, v)
        solver.upload("fc", self.fc)
        solver.upload("rbct", self.rbct)
        for lv in (0, 1):
            for k, v in self.init.items():
                solver.upload(k, v, level=lv)


def _periodic_index(n_int, nfull, offset=2):
    """Map every array index 0..nfull-1 onto the interior period."""
    return offset + np.mod(np.arange(nfull) - offset, n_int)


def _hydrometeors(X, Y, xc, yc, zs, tk, p, qv, ip, jp, seed, ring, cellnoise=True):
    """Seed every cold-rain species so that all of the microphysics' branches
    are visited: a cos^2 blob (or, for the vortex, a Gaussian eyewall ring)
    with per-cell noise; cloud water above -40 C, rain above -10 C,
    ice below -5 C, snow below +2 C, graupel below +7 C, and 1% water
    supersaturation where the weight exceeds one half.  The ring radius is
    100 km or 0.6 of the half-width of smaller domains, its e-folding width
    0.3 of the radius."""
    if ring:
        # the ring radius and width shrink with small test domains
        r = np.sqrt((X - xc) ** 2 + (Y - yc) ** 2)
        r0 = min(100000.0, 0.6 * min(xc, yc))
        h = np.exp(-((r - r0) / (0.3 * r0)) ** 2)
    else:
        rr = np.sqrt(((X - xc) / (0.35 * max(2 * xc, 1.0))) ** 2 + ((Y - yc) / (0.35 * max(2 * yc, 1.0))) ** 2)
        h = np.where(rr < 1.0, np.cos(0.5 * CC * np.minimum(rr, 1.0)) ** 2, 0.0)
    nk = len(zs)
    I, J = np.meshgrid(ip - 2, jp - 2, indexing="ij")
    out = {s: np.zeros(X.shape + (nk,)) for s in ("qc", "qr", "qi", "qs", "qg", "nci")}
    qvn = np.broadcast_to(qv[None, None, :], X.shape + (nk,)).copy()
    for k in range(nk):
        if zs[k] > 14000.0:
            continue
        lin = (I + 4096 * (J + 4096 * k)).astype(np.uint64)
        un = splitmix_uniform(seed + 17, lin.ravel()).reshape(I.shape)
        b = h * (0.25 + 0.75 * un) if cellnoise else h
        t = tk[k]
        if t > 233.16:
            out["qc"][:, :, k] = 1.0e-3 * b
        if t > 263.16:
            out["qr"][:, :, k] = 2.0e-3 * b
        if t < 268.16:
            out["qi"][:, :, k] = 1.0e-4 * b
            out["nci"][:, :, k] = 1.0e-4 * b / 5.0e-12
        if t < 275.16:
            out["qs"][:, :, k] = 5.0e-4 * b
        if t < 280.16:
            out["qg"][:, :, k] = 8.0e-4 * b
        qvs = tetens_qvs(t, p[k])
        qvn[:, :, k] = np.where(b > 0.5, 1.01 * qvs, qvn[:, :, k])
    out["qv"] = qvn
    return out


def hodograph(z, shear):
    """Quarter-circle hodograph of radius 8 m/s below 2 km, then straight to
    U_s = shear at 6 km (Weisman-Klemp style), shifted to a storm-relative
    frame; zero wind when shear is None."""
    u = np.zeros(len(z))
    v = np.zeros(len(z))
    if shear is None:
        return u, v
    r, us = 8.0, shear
    for k, zk in enumerate(z):
        if zk <= 2000.0:
            a = 0.5 * math.pi * max(zk, 0.0) / 2000.0
            u[k], v[k] = r - r * math.cos(a), r * math.sin(a)
        elif zk <= 6000.0:
            u[k] = r + (us - r) * (zk - 2000.0) / 4000.0
            v[k] = r
        else:
            u[k], v[k] = us, r
    return u - 0.5 * us, v - 0.5 * r


def build(name="bubble", nx=None, ny=None, nz=None, seed=None, cphopt=2, dtbig=None,
          dtsml=None, nsstp=None, advopt=3, moist=None, noise=None, divopt=1, wbc=None,
          gwave=None, smtopt=0, smhcoe=0.0008333, smvcoe=0.0008333, amp=None, shear=None):
    """Return a Case with config, base state and the initial state.

    wbc: -1 periodic or 7 open (radiation at the constant phase speed gwave,
    m/s) on all four sides; the supercell defaults to open sides with
    gwave=30, the other cases to periodic.
    smtopt: 0 none, 2/3 the 4th-order smoothing with smhcoe, smvcoe (1/s;
    defaults are the reference's documented values).
    amp, shear override the case's bubble amplitude (K) and hodograph
    shear U_s (m/s; cases with a hodograph only).

    moist (default: cphopt == 2) seeds the hydrometeor species and a
    supersaturated core so that the microphysics is active from step one.
    noise overrides the case's theta' noise amplitude; noise=0 also drops the
    per-cell factor of the hydrometeor seeding, leaving a state with the
    case's mirror symmetries."""
    open_default, gwave_default = name == "supercell", 30.0
    if name == "bubble":
        nx, ny, nz = nx or 64, ny or 64, nz or 40
        dx = dy = 1000.0
        dz, sthopt, dzmin = 500.0, 0, None
        dtbig = dtbig or 5.0
        dtsml = dtsml or 2.0          # 6 substeps is not representable (chknlslv.f90:2384-2392)
        sounding, amp, rx, rz, zc = wk_sounding, 2.0, 10000.0, 1500.0, 1500.0
        noise0, seed = 0.1, 1 if seed is None else seed
        coropt, fcor, shear = 0, 0.0, None
        botbar = 15000.0
    elif name == "supercell":
        nx, ny, nz = nx or 256, ny or 256, nz or 64
        dx = dy = 500.0
        dz, sthopt, dzmin = 18000.0 / 64, 1, 100.0
        dtbig = dtbig or 3.0
        dtsml = dtsml or 1.0
        sounding, amp, rx, rz, zc = wk_sounding, 3.0, 10000.0, 1500.0, 1500.0
        noise0, seed = 0.1, 2 if seed is None else seed
        coropt, fcor, shear = 0, 0.0, 30.0
        botbar = 15000.0
    elif name == "vortex":
        nx, ny, nz = nx or 1024, ny or 1024, nz or 80
        dx = dy = 1000.0
        dz, sthopt, dzmin = 250.0, 1, 50.0
        dtbig = dtbig or 4.0
        dtsml = dtsml or 1.0
        sounding, amp, rx, rz, zc = tropical_sounding, 0.0, 0.0, 1.0, 0.0
        noise0, seed = 0.2, 3 if seed is None else seed
        coropt, fcor, shear = 1, 5.0e-5, None
        botbar = 15000.0
    else:
        raise ValueError(name)
    noise = noise0 if noise is None else float(noise)
    if amp is not None:
        amp = float(amp)
    if shear is not None and name == "supercell":
        shear = float(shear)
    ni, nj, nk = nx + 3, ny + 3, nz + 3
    if sthopt == 1:
        # keep the physical top of the configuration: ztop = nz*dz_comp
        dzmin = min(dzmin, dz)
    zw = stretch_levels(nk, dz, sthopt, dzmin)            # w levels k=1..nk
    zs = np.empty(nk)
    zs[:nk - 1] = 0.5 * (zw[:-1] + zw[1:])
    zs[nk - 1] = zw[nk - 1] + 0.5 * (zw[nk - 1] - zw[nk - 2])
    th, rh = sounding(zs)
    p, qv, rho = hydrostatic(zs, th, rh)
    ptop_k = nk - 1
    del ptop_k
    jcb = np.empty(nk)
    jcb[:nk - 1] = (zw[1:] - zw[:-1]) / dz
    jcb[nk - 1] = jcb[nk - 2]
    rst = jcb * rho

    def w8(a, bbc=2, tbc=2):
        o = np.empty(nk)
        o[1:nk - 1] = 0.5 * (a[0:nk - 2] + a[1:nk - 1])
        o[0] = o[2] if bbc == 2 else o[1]
        o[nk - 1] = o[nk - 3] if tbc == 2 else o[nk - 2]
        return o

    shape = (ni + 2, nj + 2, nk)

    def col(prof):
        return np.broadcast_to(prof[None, None, :], shape).astype(np.float32, order="F").copy(order="F")

    ubr_p = np.zeros(nk)
    vbr_p = np.zeros(nk)
    if shear is not None:
        # quarter-circle hodograph of radius 8 m/s below 2 km, then straight to
        # U_s at 6 km (Weisman-Klemp style); shifted to a storm-relative frame
        r, us = 8.0, shear
        for k in range(nk):
            z = zs[k]
            if z <= 2000.0:
                a = 0.5 * math.pi * max(z, 0.0) / 2000.0
                ubr_p[k], vbr_p[k] = r - r * math.cos(a), r * math.sin(a)
            elif z <= 6000.0:
                ubr_p[k] = r + (us - r) * (z - 2000.0) / 4000.0
                vbr_p[k] = r
            else:
                ubr_p[k], vbr_p[k] = us, r
        ubr_p -= 0.5 * us
        vbr_p -= 0.5 * r
    base = {
        "jcb": col(jcb), "jcb8u": col(jcb), "jcb8v": col(jcb), "jcb8w": col(w8(jcb)),
        "ubr": col(ubr_p), "vbr": col(vbr_p), "pbr": col(p), "ptbr": col(th), "qvbr": col(qv),
        "rbr": col(rho), "rst": col(rst), "rst8u": col(rst), "rst8v": col(rst), "rst8w": col(w8(rst)),
        "rcsq": col(CP / CV * p),
    }
    fc = np.full((ni + 2, nj + 2, 2), 0.25 * fcor, dtype=np.float32, order="F")
    # vspdmp.f90:222-285 (vspopt=2): sponge above botbar, vspbar=0.003333
    vspbar = 0.003333
    rb = np.zeros(nk)
    for k in range(2, nk):            # Fortran k = 2..nk-1 -> index k-1
        zk = zw[k - 1]
        if zk > botbar:
            rb[k - 1] = 0.5 * vspbar * (1.0 - math.cos(CC * (zk - botbar) / (zw[nk - 2] - botbar)))
    ksp0 = 2
    for k in range(3, nk + 1):
        if zw[k - 1] > botbar:
            ksp0 = k - 1
            break
    rbct = np.zeros((ni, nj, nk, 2), dtype=np.float32, order="F")
    rbct[:, :, :, 1] = rb[None, None, :]
    # initial perturbations on the periodic interior, wrapped into the halos
    ip = _periodic_index(ni - 3, ni + 2)
    jp = _periodic_index(nj - 3, nj + 2)
    xs = (ip - 2 + 0.5) * dx
    ys = (jp - 2 + 0.5) * dy
    X, Y = np.meshgrid(xs, ys, indexing="ij")
    xc, yc = 0.5 * nx * dx, 0.5 * ny * dy
    ptp = np.zeros(shape)
    if amp > 0.0:
        for k in range(nk):
            r2 = ((X - xc) / rx) ** 2 + ((Y - yc) / rx) ** 2 + ((zs[k] - zc) / rz) ** 2
            r = np.sqrt(r2)
            ptp[:, :, k] = np.where(r < 1.0, amp * np.cos(0.5 * CC * r) ** 2, 0.0)
    u = np.zeros(shape)
    v = np.zeros(shape)
    pp = np.zeros(shape)
    if name == "vortex":
        # axisymmetric warm-core vortex: Vmax=15 m/s at r=100 km, decaying
        # linearly to 0 at 15 km; pressure from gradient-wind balance and
        # theta' from hydrostatic balance of that pressure (FP64)
        vmax, rmax, ztop_v = 15.0, 100000.0, 15000.0
        xu = (ip - 2) * dx
        Xu, Yu = np.meshgrid(xu, ys, indexing="ij")
        yv = (jp - 2) * dy
        Xv, Yv = np.meshgrid(xs, yv, indexing="ij")

        def vt(rr):
            return vmax * (rr / rmax) * np.exp(0.5 * (1.0 - (rr / rmax) ** 2))

        rgrid = np.linspace(0.0, 0.75 * max(nx * dx, ny * dy), 4000)
        dr = rgrid[1] - rgrid[0]
        vg = vt(rgrid)
        dpdr = vg ** 2 / np.maximum(rgrid, 1.0) + fcor * vg      # per unit density
        phi = np.cumsum(dpdr) * dr
        phi -= phi[-1]
        ru = np.sqrt((Xu - xc) ** 2 + (Yu - yc) ** 2)
        rv = np.sqrt((Xv - xc) ** 2 + (Yv - yc) ** 2)
        rs = np.sqrt((X - xc) ** 2 + (Y - yc) ** 2)
        u2 = -vt(ru) * (Yu - yc) / np.maximum(ru, 1.0)
        v2 = vt(rv) * (Xv - xc) / np.maximum(rv, 1.0)
        p2 = np.interp(rs, rgrid, phi)
        for k in range(nk):
            fz = max(0.0, 1.0 - zs[k] / ztop_v)
            u[:, :, k] = u2 * fz
            v[:, :, k] = v2 * fz
            pp[:, :, k] = rho[k] * p2 * fz * fz
        # hydrostatic theta' consistent with dp'/dz = -rho g theta'/theta
        for k in range(1, nk - 1):
            dpdz = (pp[:, :, k + 1] - pp[:, :, k - 1]) / (zs[k + 1] - zs[k - 1])
            ptp[:, :, k] = -dpdz * th[k] / (rho[k] * G)
    if noise > 0.0:
        ii, jj, kk = np.meshgrid(np.arange(2, ni - 1), np.arange(2, nj - 1), np.arange(2, nk - 1),
                                 indexing="ij")
        lin = (ii - 2) + nx * (jj - 2) + nx * ny * (kk - 2)
        un = splitmix_uniform(seed, lin.ravel(order="F")).reshape(lin.shape, order="F")
        ptp[2:ni - 1, 2:nj - 1, 1:nk - 2] += noise * (2.0 * un - 1.0)
        # re-wrap the noise into the halos
        ptp = ptp[ip][:, jp]
    # vertical boundary rows as vbcs/vbcu/vbcp leave them
    for a in (u, v, ptp, pp):
        a[:, :, 0] = a[:, :, 1]
        a[:, :, nk - 2] = a[:, :, nk - 3]
    qvp = np.broadcast_to(qv[None, None, :], shape).copy()
    init = {"u": u, "v": v, "w": np.zeros(shape), "pp": pp, "ptp": ptp, "qv": qvp}
    if cphopt == 2:
        for s in ("qc", "qr", "qi", "qs", "qg", "nci"):
            init[s] = np.zeros(shape)
        if moist is None or moist:
            tk = th * (p / P0) ** (RD / CP)
            init.update(_hydrometeors(X, Y, xc, yc, zs, tk, p, qv, ip, jp, seed, name == "vortex",
                                      cellnoise=noise > 0.0))
            for s in ("qv", "qc", "qr", "qi", "qs", "qg", "nci"):
                a = init[s]
                a[:, :, 0] = a[:, :, 1]
                a[:, :, nk - 2] = a[:, :, nk - 3]
    init = {k: np.asfortranarray(a.astype(np.float32)) for k, a in init.items()}
    cfg = host.make_config(ni, nj, nk, dx, dy, dz, dtbig, dtsml, nsstp=nsstp, advopt=advopt,
                           divopt=divopt, cphopt=cphopt, vspopt=2, coropt=coropt, sthopt=sthopt,
                           ksp0=(ksp0, ksp0), vspvar=b"oooxoooooo",
                           wbc=(7 if open_default else -1) if wbc is None else wbc,
                           gwave=(gwave_default if gwave is None else gwave),
                           smtopt=smtopt, smhcoe=smhcoe if smtopt else 0.0,
                           smvcoe=smvcoe if smtopt else 0.0)
    snd = {"fn": sounding, "shear": shear, "dzmin": dzmin if sthopt == 1 else None,
           "layer2": (nk - 3) * dz, "botbar": botbar, "fcor": fcor}
    return Case(cfg, base, fc, rbct, init, zw, zs, name, snd)
