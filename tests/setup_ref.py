"""FP64 numpy restatement of cress_set_base's per-level chain (setup.hip)
for the tests: s_stretch, s_jacobian, scalar heights, s_intrpsnd (vint13),
s_setbase + s_bcbase, s_baserho, s_var8uvw (w points) with bc8w, s_sndwave,
s_vspdmp (vspopt=2); the same operations in the same order, column index
k = 1..nk (slot 0 unused)."""
import math

import numpy as np

G, CP, CV, RD, P0, EPSAV, CC = 9.8, 1004.0, 717.0, 287.0, 1.e5, 1.6077, 3.141592


def vint13(zz, var, z):
    n = len(zz) - 1
    if zz[0] > z:
        dk = (z - zz[0]) / (zz[1] - zz[0])
        return (1.0 - dk) * var[0] + dk * var[1]
    if zz[n] <= z:
        dk = (z - zz[n - 1]) / (zz[n] - zz[n - 1])
        return (1.0 - dk) * var[n - 1] + dk * var[n]
    for kl in range(n):
        if zz[kl] <= z < zz[kl + 1]:
            dk = (z - zz[kl]) / (zz[kl + 1] - zz[kl])
            return (1.0 - dk) * var[kl] + dk * var[kl + 1]
    return var[n]


def stretch(nk, dz, sthopt, zsfc, dzmin, layer1, layer2):
    zph = np.zeros(nk + 1)
    ztop2 = zsfc + (nk - 3) * dz
    layer1 = max(min(ztop2 - zsfc, layer1), 0.0)
    z1 = zsfc + layer1
    layer2 = max(min(ztop2 - z1, layer2), 0.0)
    z2 = z1 + layer2
    if sthopt == 0:
        for k in range(1, nk + 1):
            zph[k] = zsfc + (k - 2) * dz
        return zph
    z1 = max(zsfc, z1)
    z2 = min(ztop2, z2)
    klow = int((z1 - zsfc) / dzmin + 0.1)
    if z2 >= ztop2:
        dzmid = (ztop2 - zsfc - klow * dzmin) / (nk - 3 - klow)
        dzhigh = 2.0 * dzmid - dzmin
        kmid = 0
    else:
        aa = 2.0 * (nk - 3 - klow)
        bb = 2.0 * zsfc - ztop2 - z2 - (nk - 3 - 3 * klow) * dzmin
        cc = dzmin * (z2 - zsfc - klow * dzmin)
        dzmid = 0.5 * (math.sqrt(bb * bb - 4.0 * aa * cc) - bb) / aa
        kmid = int((ztop2 - z2) / (2.0 * dzmid - dzmin) + 0.1)
        aa = zsfc + (klow - kmid) * dzmin + (nk - 3 - klow + kmid) * dzmid
        dzhigh = 0.0
        if kmid != 0:
            dzhigh = (2.0 * dzmid - dzmin) + (ztop2 - aa) / kmid
        elif nk - 3 - klow - kmid != 0:
            dzmid = dzmid + (ztop2 - aa) / (nk - 3 - klow - kmid)
            dzhigh = 2.0 * dzmid - dzmin
    sa = dzmid - dzmin
    sb = 1.0 / (nk - 4 - kmid - klow)

    def dzs(k):
        if k <= klow + 1:
            return dzmin
        if k <= nk - 2 - kmid:
            t = 2.0 * sb * (k - 2 - klow) - 1.0
            return dzmid + sa * t * t * t
        return dzhigh

    zph[2] = zsfc
    for k in range(3, nk - 1):
        zph[k] = zph[k - 1] + dzs(k - 1)
    zph[1] = zph[2] - dzs(1)
    zph[nk - 1] = ztop2
    zph[nk] = zph[nk - 1] + dzs(nk - 2)
    return zph


def columns(nk, dz, sthopt, bbc, tbc, snd):
    """The per-level columns as cress_base_columns returns them."""
    c = {k: np.zeros(nk + 1) for k in ("zph", "zph8s", "ubr", "vbr", "pbr", "ptbr", "qvbr", "rbr", "jcb",
                                       "jcb8w", "rst", "rst8w", "rcsq", "rbct", "ptvbr", "pibr")}
    zph = c["zph"]
    zph[:] = stretch(nk, dz, sthopt, snd["zsfc"], snd.get("dzmin", 0.0), snd.get("layer1", 0.0),
                     snd.get("layer2", 0.0))
    jcb, z8s = c["jcb"], c["zph8s"]
    for k in range(1, nk):
        jcb[k] = (zph[k + 1] - zph[k]) / dz
        z8s[k] = 0.5 * (zph[k + 1] + zph[k])
    jcb[nk], z8s[nk] = jcb[nk - 1], z8s[nk - 1]
    zz = [float(v) for v in snd["z"]]
    ptv, pib = c["ptvbr"], c["pibr"]
    for k in range(2, nk - 1):
        for name, key in (("ubr", "u"), ("vbr", "v"), ("pbr", "p"), ("ptbr", "pt"), ("qvbr", "qv")):
            c[name][k] = vint13(zz, [float(v) for v in snd[key]], z8s[k])
        ptv[k] = c["ptbr"][k] * (1.0 + EPSAV * c["qvbr"][k]) / (1.0 + c["qvbr"][k])
        pib[k] = math.exp((RD / CP) * math.log((1.0 / P0) * c["pbr"][k]))
        c["rbr"][k] = c["pbr"][k] / (RD * ptv[k] * pib[k])
    for name in ("ubr", "vbr", "ptbr", "qvbr", "ptvbr"):
        c[name][1] = c[name][2]
        c[name][nk - 1] = c[name][nk - 2]
    gdvcp2, cpdvrd = 2.0 * G / CP, CP / RD
    pib[1] = pib[2] + gdvcp2 * (z8s[2] - z8s[1]) / (ptv[1] + ptv[2])
    pib[nk - 1] = pib[nk - 2] - gdvcp2 * (z8s[nk - 1] - z8s[nk - 2]) / (ptv[nk - 2] + ptv[nk - 1])
    for k in (1, nk - 1):
        c["pbr"][k] = P0 * math.exp(cpdvrd * math.log(pib[k]))
        c["rbr"][k] = c["pbr"][k] / (RD * ptv[k] * pib[k])
    for name in ("ubr", "vbr", "pbr", "ptbr", "qvbr", "rbr"):
        c[name][nk] = c[name][nk - 1]
    rst = c["rst"]
    for k in range(1, nk):
        rst[k] = abs(jcb[k]) * c["rbr"][k]
    rst[nk] = rst[nk - 1]
    for k in range(2, nk):
        c["rst8w"][k] = 0.5 * (rst[k - 1] + rst[k])
        c["jcb8w"][k] = 0.5 * (jcb[k - 1] + jcb[k])
    for name in ("rst8w", "jcb8w"):
        v = c[name]
        v[1] = v[3] if bbc == 2 else v[2]
        v[nk] = v[nk - 2] if tbc == 2 else v[nk - 1]
    for k in range(1, nk + 1):
        c["rcsq"][k] = (CP / CV) * c["pbr"][k]
    cbar05 = 0.5 * snd["vspbar"]
    for k in range(2, nk):
        if zph[k] > snd["botbar"]:
            c["rbct"][k] = cbar05 * (1.0 - math.cos(CC * (zph[k] - snd["botbar"]) / (zph[nk - 1] - snd["botbar"])))
    return c


def ksp0(zph, botbar):
    """s_vspdmp's lowest sponge level (vspopt=2): one below the first level
    k >= 3 whose height exceeds botbar, else 2."""
    nk = len(zph) - 1
    for k in range(3, nk + 1):
        if zph[k] > botbar:
            return k - 1
    return 2
