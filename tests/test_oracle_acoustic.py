"""What the mathematics fixes for the acoustic part of the step, on the
FP64 oracle: the Lamb wave of a resting isothermal atmosphere.

With temperature T, H = R T / g and gamma = cp/cv, the linearised
equations (w = 0, theta' = 0) have the exact solution
    p' = P exp(-z / (gamma H)) cos(k x) cos(omega t),
    u  = U exp(z (1 - 1/gamma) / H) sin(k x) sin(omega t),
    omega = c k,  c^2 = gamma R T,
a horizontally propagating sound wave in hydrostatic balance.  Started from
its t = 0 pressure field, with the base state set up by the reference's own
chain (tests/setup_ref.py, pinned against s_setbase, s_baserho,
s_sndwave) from an isothermal sounding, rigid top and bottom, periodic
sides, no divergence damping, sponge, smoothing, Coriolis or moisture, the
step must

* oscillate at c k to within the second-order C-grid's dispersion (the
  acoustic gradients and divergences are centred differences over dx:
  omega_grid = c (2/dx) sin(k dx / 2)) plus a small margin, and
* not amplify it (the forward-backward substeps with the off-centred
  vertical implicit weighting are neutral or damping).

A wrong sound speed (rcsq without cp/cv: c too low by 16%), a sign error in
the pressure gradient or the divergence (growth), or a lost factor of the
substep interval all fail these bars."""
import math

import numpy as np

from cress_amd import cases
from oracle import pyoracle as po
import setup_ref

T0 = 300.0


def isothermal_sounding(ztop):
    z = np.arange(0.0, ztop + 1000.0, 10.0)
    H = setup_ref.RD * T0 / setup_ref.G
    p = setup_ref.P0 * np.exp(-z / H)
    pt = T0 * (setup_ref.P0 / p) ** (setup_ref.RD / setup_ref.CP)
    zero = np.zeros_like(z)
    return {"z": z, "u": zero, "v": zero, "p": p, "pt": pt, "qv": zero, "zsfc": 0.0, "botbar": 1.0e9,
            "vspbar": 0.0}


def lamb_case(nx=24, nz=20, dx=1000.0):
    case = cases.build("bubble", nx, 4, nz, cphopt=0, dtbig=4.0, dtsml=1.0, noise=0.0, divopt=0, moist=False)
    c = case.cfg
    c.vspopt = 0
    c.coropt = 0
    assert c.sthopt == 0 and abs(c.dx - dx) < 1e-9
    snd = isothermal_sounding((c.nk - 3) * c.dz)
    col = setup_ref.columns(c.nk, c.dz, 0, c.bbc, c.tbc, snd)
    sh = (c.ni + 2, c.nj + 2, c.nk)
    field = {"jcb": "jcb", "jcb8u": "jcb", "jcb8v": "jcb", "jcb8w": "jcb8w", "ubr": "ubr", "vbr": "vbr",
             "pbr": "pbr", "ptbr": "ptbr", "qvbr": "qvbr", "rbr": "rbr", "rst": "rst", "rst8u": "rst",
             "rst8v": "rst", "rst8w": "rst8w", "rcsq": "rcsq"}
    for f, k in field.items():
        case.base[f] = np.asfortranarray(np.broadcast_to(col[k][1:][None, None, :], sh).astype(np.float32))
    # the t = 0 pressure field of the wave (scalar points), everything else 0
    gam = setup_ref.CP / setup_ref.CV
    H = setup_ref.RD * T0 / setup_ref.G
    nxp = c.ni - 3                                 # periodic length in cells
    k = 2.0 * math.pi / (nxp * c.dx)
    i = np.arange(c.ni + 2)
    z = col["zph8s"][1:]                           # scalar-point heights
    pp = 1.0 * np.cos(k * (i - 2) * c.dx)[:, None, None] * np.exp(-z / (gam * H))[None, None, :]
    init = {f: np.zeros(sh, np.float32, order="F") for f in ("u", "v", "w", "ptp", "qv")}
    init["pp"] = np.asfortranarray(np.broadcast_to(pp, sh).astype(np.float32))
    case.init = init
    case.rbct = np.zeros_like(case.rbct)
    return case, k, math.sqrt(gam * setup_ref.RD * T0)


def _fit_omega(t, y, lo, hi):
    best = None
    for om in np.linspace(lo, hi, 4001):
        a = np.stack([np.cos(om * t), np.sin(om * t), np.ones_like(t)], 1)
        coef, res, _, _ = np.linalg.lstsq(a, y, rcond=None)
        r = float(np.sum((a @ coef - y) ** 2))
        if best is None or r < best[1]:
            best = (om, r)
    return best[0]


def test_lamb_wave_frequency_and_amplitude():
    case, k, c = lamb_case()
    cfg = case.cfg
    m = po.Model(case, 64)
    i0, j0, k0 = 2, 3, cfg.nk // 2                # a crest of the pressure wave, mid-depth
    p0 = m.get("pp", 1)[i0, j0, k0]
    ts, ys = [0.0], [p0]
    m.step(1, first=True)
    ts.append(0.5 * cfg.dtbig)
    ys.append(m.get("pp", 1)[i0, j0, k0])
    nsteps = 70
    for n in range(1, nsteps):
        m.step(1)
        ts.append((n + 0.5) * cfg.dtbig)
        ys.append(m.get("pp", 1)[i0, j0, k0])
    t, y = np.array(ts), np.array(ys)
    om_exact = c * k
    om_grid = c * (2.0 / cfg.dx) * math.sin(0.5 * k * cfg.dx)
    om = _fit_omega(t[1:], y[1:], 0.7 * om_exact, 1.3 * om_exact)
    amp = np.abs(y).max()
    w = np.abs(m.get("w", 1)).max()
    print("Lamb wave: omega %.5f, c k %.5f, grid %.5f (%.2f%% / %.2f%%); max |p'| %.4f of %.4f; max |w| %.2e"
          % (om, om_exact, om_grid, 100 * (om / om_exact - 1), 100 * (om / om_grid - 1), amp, abs(p0), w))
    assert abs(om / om_grid - 1.0) < 0.01
    assert amp <= 1.001 * abs(p0)
    assert np.isfinite(y).all()
