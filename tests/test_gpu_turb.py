"""cress_turbcoef -- the coefficients of the 1.5-order TKE closure
(turbdrv.f90:245-276: s_bruntv, s_defomten, s_defomssq, s_eddyvis with
tubopt=2, isoopt=1, flat terrain without map factors, sfcopt=0) -- against
the oracle's or_turbcoef, itself pinned bitwise to the reference's routines
(tests/test_oracle_turb.py).  Outputs start from the same random contents,
so what the kernels leave alone is checked too.  The deformation tensor,
ssq and N^2 of dry or cloud-free air are bitwise; rkh, rkv and priv
(binary32 exp and log of their own on the GPU, glibc's in the oracle) are
within 1e-5 of each cell's value, N^2 with cloud physics within 1e-5 of
the field's largest (the moist N^2 is a difference that can cancel) --
with cloud physics, the eddy viscosity stage on the GPU's own N^2 and ssq.  Also the
rest of s_turbdrv, s_turbflx, s_turbs, s_kh8uv and s_forcetke through
their C-ABI entry points."""
import ctypes as C

import numpy as np
import pytest

from cress_amd import host
from oracle import pyoracle as po
from turb_inputs import IN, OUT, TKE_IN, TKE_OUT, inputs, tke_inputs

pytestmark = pytest.mark.gpu

NI, NJ, NK = 70, 19, 24
DTB = 6.0
EXACT = ["t11", "t22", "t33", "t12", "t13", "t23", "t31", "t32", "ssq"]
CLOSE = ["rkh", "rkv", "priv"]


def _ptr(a):
    assert a.flags["F_CONTIGUOUS"] and a.dtype == np.float32
    return a.ctypes.data_as(C.c_void_p)


def rel(a, r):
    a, r = np.asarray(a, np.float64), np.asarray(r, np.float64)
    return np.abs(a - r).max() / max(np.abs(r).max(), 1e-30)


def worst_elementwise(a, r, floor=1e-7):
    """The largest per-cell error relative to max(|r|, floor max|r|): small
    cells are held to their own size, not to the field's largest value."""
    a, r = np.asarray(a, np.float64), np.asarray(r, np.float64)
    scale = np.maximum(np.abs(r), floor * max(np.abs(r).max(), 1e-30))
    return (np.abs(a - r) / scale).max()


@pytest.mark.parametrize("bbc", [2, 3])
@pytest.mark.parametrize("wbc", [-1, 7])
@pytest.mark.parametrize("moist,cphopt", [(0, 0), (1, 0), (1, 2)])
def test_turbcoef_matches_oracle(moist, cphopt, wbc, bbc):
    d, outs = inputs(NI, NJ, NK, 31 + cphopt + (wbc > 0) + 3 * bbc + 7 * moist)
    gpu = {n: a.copy(order="F") for n, a in outs.items()}
    ora = {n: a.copy(order="F") for n, a in outs.items()}
    cfg = host.make_config(NI, NJ, NK, 700.0, 900.0, 350.0, DTB, 1.0, cphopt=cphopt, wbc=wbc, gwave=30.0)
    cfg.bbc = cfg.tbc = bbc
    cfg.thresq = 1.e-5
    print("cress_turbcoef moist=%d cphopt=%d wbc=%d bbc=%d %dx%dx%d" % (moist, cphopt, wbc, bbc, NI, NJ, NK),
          flush=True)
    host.check(host.lib().cress_turbcoef(C.byref(cfg), C.c_float(DTB), moist, *[_ptr(d[n]) for n in IN],
                                         *[_ptr(gpu[n]) for n in OUT]), "cress_turbcoef")
    c = po.make_cfg(NI, NJ, NK, dx=700.0, dy=900.0, dz=350.0, wbc=wbc, cphopt=cphopt, bbc=bbc, tbc=bbc,
                    thresq=1.e-5, gwave=30.0)
    z = lambda: np.zeros((NI + 2, NJ + 2, NK), np.float32, order="F")
    po.Oracle(32).or_turbcoef(c, DTB, moist, *[d[n] for n in IN], *[ora[n] for n in OUT], z(), z(), z(), z())
    for n in EXACT:
        np.testing.assert_array_equal(gpu[n], ora[n], err_msg=n)
    if moist and cphopt:
        assert rel(gpu["nsq8w"], ora["nsq8w"]) <= 1e-5
        # eddyvis sums N^2 of two levels, which may nearly cancel; the
        # mixing length then amplifies the ulps of cloudy N^2, as it does
        # in the reference.  The stage is checked on the GPU's own N^2 and
        # ssq.
        for n in CLOSE:
            ora[n] = outs[n].copy(order="F")
        po.Oracle(32).or_eddyvis(c, DTB, d["jcb"], d["rbr"], d["tkep"], gpu["ssq"], gpu["nsq8w"], ora["rkh"],
                                 ora["rkv"], ora["priv"])
    else:
        np.testing.assert_array_equal(gpu["nsq8w"], ora["nsq8w"])
    for n in CLOSE:
        e = worst_elementwise(gpu[n], ora[n])
        print("  %s: max rel %.2e, worst cell %.2e" % (n, rel(gpu[n], ora[n]), e))
        assert e <= 1e-5, (n, e)
        untouched = np.ones(gpu[n].shape, bool)
        untouched[1:NI, 1:NJ, 0:NK - 1] = False
        np.testing.assert_array_equal(gpu[n][untouched], outs[n][untouched], err_msg=n)
    assert not np.array_equal(gpu["rkv"], outs["rkv"])


def test_turbcoef_rejects_bad_arguments():
    cfg = host.make_config(NI, NJ, NK, 700.0, 900.0, 350.0, DTB, 1.0, cphopt=0)
    d, outs = inputs(NI, NJ, NK, 3)
    args = [_ptr(d[n]) for n in IN] + [_ptr(outs[n]) for n in OUT]
    assert host.lib().cress_turbcoef(C.byref(cfg), C.c_float(0.0), 1, *args) == -1
    cfg.wbc = 4
    assert host.lib().cress_turbcoef(C.byref(cfg), C.c_float(DTB), 1, *args) == -6


DRV_IN = ["jcb", "jcb8u", "jcb8v", "jcb8w", "pbr", "ptbr", "rbr", "up", "vp", "wp", "ppp", "ptpp", "qvp", "tkep",
          "qallp"]
DRV_OUT = ["ufrc", "vfrc", "wfrc", "rkh", "rkv", "priv", "ssq", "nsq8w", "t13", "t23", "t33"]


@pytest.mark.parametrize("wbc", [-1, 7])
@pytest.mark.parametrize("moist,cphopt", [(0, 0), (1, 0), (1, 2)])
def test_turbdrv_matches_oracle(moist, cphopt, wbc):
    """cress_turbdrv (all of s_turbdrv).  The coefficients as above; the
    stress tensor scaled by strsten and the velocity forcings of turbuvw
    bitwise equal to the oracle's strsten and turbuvw run on the oracle's
    tensor with the GPU's eddy viscosities; the forcings of the oracle's
    whole chain within 1e-5 where no cloud physics is on."""
    d, outs = inputs(NI, NJ, NK, 61 + cphopt + (wbc > 0) + 7 * moist)
    rng = np.random.default_rng(5)
    for n in ("ufrc", "vfrc", "wfrc"):
        outs[n] = np.asfortranarray(rng.uniform(-1.0, 1.0, (NI + 2, NJ + 2, NK)).astype(np.float32))
    gpu = {n: outs[n].copy(order="F") for n in DRV_OUT}
    cfg = host.make_config(NI, NJ, NK, 700.0, 900.0, 350.0, DTB, 1.0, cphopt=cphopt, wbc=wbc, gwave=30.0)
    cfg.thresq = 1.e-5
    print("cress_turbdrv moist=%d cphopt=%d wbc=%d %dx%dx%d" % (moist, cphopt, wbc, NI, NJ, NK), flush=True)
    host.check(host.lib().cress_turbdrv(C.byref(cfg), C.c_float(DTB), moist, *[_ptr(d[n]) for n in DRV_IN],
                                        *[_ptr(gpu[n]) for n in DRV_OUT]), "cress_turbdrv")
    c = po.make_cfg(NI, NJ, NK, dx=700.0, dy=900.0, dz=350.0, wbc=wbc, cphopt=cphopt, thresq=1.e-5, gwave=30.0)
    o = po.Oracle(32)
    z = lambda: np.zeros((NI + 2, NJ + 2, NK), np.float32, order="F")
    # stage check: the oracle's tensor, the GPU's eddy viscosities
    t = {n: outs[n].copy(order="F") for n in OUT}
    o.or_turbcoef(c, DTB, moist, *[d[n] for n in IN], *[t[n] for n in OUT], z(), z(), z(), z())
    np.testing.assert_array_equal(gpu["ssq"][1:NI, 1:NJ, 0:NK - 1], t["ssq"][1:NI, 1:NJ, 0:NK - 1])
    f = {n: outs[n].copy(order="F") for n in ("ufrc", "vfrc", "wfrc")}
    o.or_strsten(c, gpu["rkh"], gpu["rkv"], *[t[n] for n in ("t11", "t22", "t33", "t12", "t13", "t23", "t31", "t32")])
    o.or_turbuvw(c, d["jcb"], d["jcb8u"], d["jcb8v"],
                 *[t[n] for n in ("t11", "t22", "t33", "t12", "t13", "t23", "t31", "t32")], f["ufrc"], f["vfrc"],
                 f["wfrc"])
    for n in ("ufrc", "vfrc", "wfrc"):
        np.testing.assert_array_equal(gpu[n], f[n], err_msg=n)
    for n in ("t13", "t23", "t33"):
        np.testing.assert_array_equal(gpu[n], t[n], err_msg=n)
    if not (moist and cphopt):
        full = {n: outs[n].copy(order="F") for n in DRV_OUT}
        w = {n: outs[n].copy(order="F") for n in ("t11", "t22", "t12", "t31", "t32")}
        o.or_turbdrv(c, DTB, moist, *[d[n] for n in IN], *[full[n] for n in DRV_OUT],
                     *[w[n] for n in ("t11", "t22", "t12", "t31", "t32")], z(), z(), z(), z())
        # the forcings are divergences of stresses that carry the eddy
        # viscosity's last-bit differences; a sum that cancels has no
        # per-cell relative bound, so the bar is the field's largest value
        for n in ("ufrc", "vfrc", "wfrc"):
            assert rel(gpu[n], full[n]) <= 1e-5, (n, rel(gpu[n], full[n]))
    for n in ("ufrc", "vfrc", "wfrc"):
        assert not np.array_equal(gpu[n], outs[n])


@pytest.mark.parametrize("wbc", [-1, 7])
def test_turbflx_matches_oracle(wbc):
    """cress_turbflx against the oracle's or_turbflx (pinned to s_turbflx):
    bitwise over whole arrays."""
    rng = np.random.default_rng(17 + (wbc > 0))
    f = lambda lo, hi: np.asfortranarray(rng.uniform(lo, hi, (NI + 2, NJ + 2, NK)).astype(np.float32))
    jcb, s, kh, kv, kw = f(0.9, 1.1), f(290.0, 310.0), f(0.0, 50.0), f(0.0, 50.0), f(0.0, 20.0)
    h = [f(-1.0, 1.0) for _ in range(3)]
    gpu = [a.copy(order="F") for a in h]
    ora = [a.copy(order="F") for a in h]
    cfg = host.make_config(NI, NJ, NK, 700.0, 900.0, 350.0, DTB, 1.0, cphopt=0, wbc=wbc, gwave=30.0)
    print("cress_turbflx wbc=%d %dx%dx%d" % (wbc, NI, NJ, NK), flush=True)
    host.check(host.lib().cress_turbflx(C.byref(cfg), *[_ptr(a) for a in (jcb, s, kh, kv, kw)],
                                        *[_ptr(a) for a in gpu]), "cress_turbflx")
    c = po.make_cfg(NI, NJ, NK, dx=700.0, dy=900.0, dz=350.0, wbc=wbc, cphopt=0, gwave=30.0)
    po.Oracle(32).or_turbflx(c, jcb, s, kh, kv, kw, *ora)
    for a, b, n in zip(gpu, ora, ("h1", "h2", "h3")):
        np.testing.assert_array_equal(a, b, err_msg=n)
    assert not np.array_equal(gpu[2], h[2])


@pytest.mark.parametrize("wbc", [-1, 7])
def test_turbs_matches_oracle(wbc):
    """cress_turbs against the oracle's or_turbs (pinned to s_turbs):
    bitwise over the whole forcing array."""
    rng = np.random.default_rng(23 + (wbc > 0))
    f = lambda lo, hi: np.asfortranarray(rng.uniform(lo, hi, (NI + 2, NJ + 2, NK)).astype(np.float32))
    ju, jv, h1, h2, h3, sf = f(0.9, 1.1), f(0.9, 1.1), f(-5.0, 5.0), f(-5.0, 5.0), f(-5.0, 5.0), f(-1.0, 1.0)
    gpu, ora = sf.copy(order="F"), sf.copy(order="F")
    cfg = host.make_config(NI, NJ, NK, 700.0, 900.0, 350.0, DTB, 1.0, cphopt=0, wbc=wbc, gwave=30.0)
    print("cress_turbs wbc=%d %dx%dx%d" % (wbc, NI, NJ, NK), flush=True)
    host.check(host.lib().cress_turbs(C.byref(cfg), *[_ptr(a) for a in (ju, jv, h1, h2, h3, gpu)]), "cress_turbs")
    c = po.make_cfg(NI, NJ, NK, dx=700.0, dy=900.0, dz=350.0, wbc=wbc, cphopt=0, gwave=30.0)
    po.Oracle(32).or_turbs(c, ju, jv, h1, h2, h3, ora)
    np.testing.assert_array_equal(gpu, ora)
    assert not np.array_equal(gpu, sf)


def test_kh8uv_matches_oracle():
    """cress_kh8uv against the oracle's or_kh8uv (pinned to s_kh8uv)."""
    rng = np.random.default_rng(29)
    f = lambda: np.asfortranarray(rng.uniform(0.0, 50.0, (NI + 2, NJ + 2, NK)).astype(np.float32))
    kh, u0, v0 = f(), f(), f()
    gu, gv, ou, ov = u0.copy(order="F"), v0.copy(order="F"), u0.copy(order="F"), v0.copy(order="F")
    cfg = host.make_config(NI, NJ, NK, 700.0, 900.0, 350.0, DTB, 1.0, cphopt=0)
    print("cress_kh8uv %dx%dx%d" % (NI, NJ, NK), flush=True)
    host.check(host.lib().cress_kh8uv(C.byref(cfg), _ptr(kh), _ptr(gu), _ptr(gv)), "cress_kh8uv")
    c = po.make_cfg(NI, NJ, NK, dx=700.0, dy=900.0, dz=350.0, cphopt=0)
    po.Oracle(32).or_kh8uv(c, kh, ou, ov)
    np.testing.assert_array_equal(gu, ou)
    np.testing.assert_array_equal(gv, ov)


def dissipation(d, dx, dy, dz):
    """|The dissipation term| of disptke.f90 (isoopt=1) in binary64: the
    one term of the TKE forcing that takes exp and log."""
    f = {n: np.asarray(d[n], np.float64) for n in ("priv", "rst", "tkep", "jcb")}
    ln = (f["priv"] - 1.0) * np.cbrt(0.125 * dx * dy * dz * f["jcb"]) + 1e-20
    return np.abs((0.37 * f["priv"] - 0.18) * f["rst"] * f["tkep"] * np.sqrt(f["tkep"]) / ln)


def check_tkefrc(g, o, d, calm, dx=700.0, dy=900.0, dz=350.0):
    """Bitwise where the past TKE is zero; elsewhere the difference within
    1e-5 of the dissipation term plus the forcing itself: the term's last
    bits, and the roundings of the sums after it, are all that differ."""
    np.testing.assert_array_equal(g[calm], o[calm])
    bar = 1e-5 * (dissipation(d, dx, dy, dz) + np.abs(np.asarray(o, np.float64)))
    err = np.abs(np.asarray(g, np.float64) - np.asarray(o, np.float64))
    worst = (err / np.maximum(bar, 1e-30)).max()
    print("  tkefrc: %d of %d cells differ, worst at %.3f of the bar, worst cell rel %.2e"
          % ((g != o).sum(), g.size, worst, worst_elementwise(g, o)))
    assert (err <= bar).all(), worst


@pytest.mark.parametrize("advopt,vsp,ksp", [(3, 0, 8), (3, 1, 8), (2, 1, 2)])
@pytest.mark.parametrize("wbc", [-1, 7])
def test_forcetke_matches_oracle(wbc, advopt, vsp, ksp):
    """cress_forcetke against the oracle's or_forcetke (pinned bitwise to
    s_forcetke): rkh, rkv and rkv8s bitwise; tkefrc bitwise where the
    past TKE is zero (no dissipation), elsewhere within 1e-5 of the
    dissipation term plus the forcing (the term's length scale takes exp
    and log; a cell whose forcing nearly cancels has no relative bound of
    its own)."""
    d, outs = tke_inputs(NI, NJ, NK, 61 + (wbc > 0) + 2 * vsp + 4 * advopt + ksp)
    gpu = {n: a.copy(order="F") for n, a in outs.items()}
    ora = {n: a.copy(order="F") for n, a in outs.items()}
    vspvar = b"ooooooooo" + (b"o" if vsp else b"x")
    cfg = host.make_config(NI, NJ, NK, 700.0, 900.0, 350.0, DTB, 1.0, advopt=advopt, cphopt=0, wbc=wbc,
                           gwave=30.0, vspopt=2, ksp0=(2, ksp), vspvar=vspvar)
    print("cress_forcetke wbc=%d advopt=%d vsp=%d ksp0=%d %dx%dx%d" % (wbc, advopt, vsp, ksp, NI, NJ, NK), flush=True)
    host.check(host.lib().cress_forcetke(C.byref(cfg), *[_ptr(d[n]) for n in TKE_IN],
                                         *[_ptr(gpu[n]) for n in TKE_OUT]), "cress_forcetke")
    c = po.make_cfg(NI, NJ, NK, dx=700.0, dy=900.0, dz=350.0, wbc=wbc, advopt=advopt, cphopt=0, vspopt=2,
                    ksp0=(2, ksp), vspvar=vspvar, gwave=30.0)
    sh = (NI + 2, NJ + 2, NK)
    po.Oracle(32).or_forcetke(c, *[d[n] for n in TKE_IN], *[ora[n] for n in TKE_OUT],
                              *[np.zeros(sh, np.float32, order="F") for _ in range(4)])
    for n in ("rkh", "rkv", "rkv8s"):
        np.testing.assert_array_equal(gpu[n], ora[n], err_msg=n)
    g, o = gpu["tkefrc"], ora["tkefrc"]
    calm = np.zeros(sh, bool)
    calm[2:NI - 1, 2:NJ - 1, 1:NK - 2] = d["tkep"][2:NI - 1, 2:NJ - 1, 1:NK - 2] == 0
    assert calm.sum() > 100
    check_tkefrc(g, o, d, calm)
    assert not np.array_equal(g, outs["tkefrc"])


def test_forcetke_matches_oracle_large():
    """cress_forcetke on a 256x128x40 grid with open sides and the TKE
    sponge: many workgroups in x, y and z, ragged tiles; the same bars."""
    ni, nj, nk = 259, 131, 43
    d, outs = tke_inputs(ni, nj, nk, 77)
    gpu = {n: a.copy(order="F") for n, a in outs.items()}
    ora = {n: a.copy(order="F") for n, a in outs.items()}
    vspvar = b"oooooooooo"
    cfg = host.make_config(ni, nj, nk, 700.0, 900.0, 350.0, DTB, 1.0, advopt=3, cphopt=0, wbc=7, gwave=30.0,
                           vspopt=2, ksp0=(2, 30), vspvar=vspvar)
    print("cress_forcetke wbc=7 advopt=3 vsp=1 ksp0=30 %dx%dx%d" % (ni, nj, nk), flush=True)
    host.check(host.lib().cress_forcetke(C.byref(cfg), *[_ptr(d[n]) for n in TKE_IN],
                                         *[_ptr(gpu[n]) for n in TKE_OUT]), "cress_forcetke")
    c = po.make_cfg(ni, nj, nk, dx=700.0, dy=900.0, dz=350.0, wbc=7, advopt=3, cphopt=0, vspopt=2, ksp0=(2, 30),
                    vspvar=vspvar, gwave=30.0)
    sh = (ni + 2, nj + 2, nk)
    po.Oracle(32).or_forcetke(c, *[d[n] for n in TKE_IN], *[ora[n] for n in TKE_OUT],
                              *[np.zeros(sh, np.float32, order="F") for _ in range(4)])
    for n in ("rkh", "rkv", "rkv8s"):
        np.testing.assert_array_equal(gpu[n], ora[n], err_msg=n)
    g, o = gpu["tkefrc"], ora["tkefrc"]
    calm = np.zeros(sh, bool)
    calm[2:ni - 1, 2:nj - 1, 1:nk - 2] = d["tkep"][2:ni - 1, 2:nj - 1, 1:nk - 2] == 0
    check_tkefrc(g, o, d, calm)
