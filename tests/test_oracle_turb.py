"""The coefficients of the 1.5-order TKE closure (tubopt=2, isoopt=1, flat
terrain without map factors, sfcopt=0): the oracle's or_turbcoef
(oracle/cress_turb.c) against the reference's own s_bruntv, s_defomten,
s_defomssq and s_eddyvis called in turbdrv.f90's order, bitwise over whole
arrays -- dry, moist without and with cloud physics, periodic and open
sides, mirrored and copied bottom/top -- and against closed forms of the
deformation tensor for a linear shear flow."""
import numpy as np
import pytest

from oracle import pyoracle as po
from turb_inputs import IN, OUT, TKE_IN, TKE_OUT, inputs, tke_inputs

NI, NJ, NK = 13, 11, 14
DTB = 6.0

pytestmark = pytest.mark.skipif(not po.have_ref(), reason="oracle/_ref not built")


@pytest.fixture(scope="module")
def libs():
    return po.Oracle(32), po.Ref()


def cfg(wbc, cphopt, bbc):
    return po.make_cfg(NI, NJ, NK, dx=700.0, dy=900.0, dz=350.0, wbc=wbc, cphopt=cphopt, bbc=bbc, tbc=bbc,
                       thresq=1.e-5, gwave=30.0)


def fld():
    return np.zeros((NI + 2, NJ + 2, NK), np.float32, order="F")


@pytest.mark.parametrize("bbc", [2, 3])
@pytest.mark.parametrize("wbc", [-1, 7])
@pytest.mark.parametrize("moist,cphopt", [(0, 0), (1, 0), (1, 2)])
def test_turbcoef_matches_reference(libs, moist, cphopt, wbc, bbc):
    o, r = libs
    c = cfg(wbc, cphopt, bbc)
    r.configure(c)
    d, outs = inputs(NI, NJ, NK, 11 + cphopt + (wbc > 0) + 3 * bbc + 7 * moist)
    mine = {n: a.copy(order="F") for n, a in outs.items()}
    ref = {n: a.copy(order="F") for n, a in outs.items()}
    o.or_turbcoef(c, DTB, moist, *[d[n] for n in IN], *[mine[n] for n in OUT], *[fld() for _ in range(4)])
    r.ref_turbcoef(moist, DTB, NI, NJ, NK, *[d[n] for n in IN], *[ref[n] for n in OUT],
                   *[fld() for _ in range(7)])
    for n in OUT:
        np.testing.assert_array_equal(mine[n], ref[n], err_msg=n)
    I, J, Kk = slice(1, NI), slice(1, NJ), slice(0, NK - 1)
    nsq = mine["nsq8w"][I, J, Kk] + mine["nsq8w"][I, J, 1:NK]
    assert (nsq < 0).any() and (nsq > 0).any()
    assert (mine["rkv"][I, J, Kk] > 0).any()


def test_deformation_of_a_linear_shear(libs):
    """u = a z, v = b x, w = 0, unit Jacobians: s13 = s31 = a and s12 = b
    everywhere inside, s11 = s22 = s33 = 0, so ssq = a^2 + b^2 where the
    four averaged neighbours are all inside."""
    o, _ = libs
    c = cfg(-1, 0, 3)
    d, outs = inputs(NI, NJ, NK, 5)
    a, b = np.float32(2.0 ** -6), np.float32(2.0 ** -7)
    for n in ("jcb", "jcb8u", "jcb8v", "jcb8w"):
        d[n][...] = 1.0
    z = (np.arange(NK, dtype=np.float32) * np.float32(350.0))[None, None, :]
    x = (np.arange(NI + 2, dtype=np.float32) * np.float32(700.0))[:, None, None]
    d["up"][...] = a * z
    d["vp"][...] = b * x
    d["wp"][...] = 0.0
    o.or_turbcoef(c, DTB, 0, *[d[n] for n in IN], *[outs[n] for n in OUT], *[fld() for _ in range(4)])
    np.testing.assert_allclose(outs["t13"][2:NI - 1, 1:NJ, 1:NK - 1], a, rtol=1e-5)
    np.testing.assert_allclose(outs["t31"][2:NI - 1, 1:NJ, 1:NK - 1], a, rtol=1e-5)
    np.testing.assert_allclose(outs["t12"][1:NI + 1, 1:NJ + 1, 0:NK - 1], b, rtol=1e-5)
    for n in ("t11", "t22", "t33"):
        np.testing.assert_array_equal(outs[n][1:NI, 1:NJ, 0:NK - 1], 0.0)
    np.testing.assert_allclose(outs["ssq"][2:NI - 1, 1:NJ - 1, 1:NK - 2], a * a + b * b, rtol=1e-5)


DRV_OUT = ["ufrc", "vfrc", "wfrc", "rkh", "rkv", "priv", "ssq", "nsq8w", "t13", "t23", "t33", "t11", "t22", "t12",
           "t31", "t32"]
INTERIOR = ["rkh", "rkv", "priv", "ssq"]   # the reference's work arrays elsewhere


def drv_inputs(seed):
    d, outs = inputs(NI, NJ, NK, seed)
    rng = np.random.default_rng(seed + 1000)
    for n in ("ufrc", "vfrc", "wfrc"):
        outs[n] = np.asfortranarray(rng.uniform(-1.0, 1.0, (NI + 2, NJ + 2, NK)).astype(np.float32))
    return d, outs


@pytest.mark.parametrize("bbc", [2, 3])
@pytest.mark.parametrize("wbc", [-1, 7])
@pytest.mark.parametrize("moist,cphopt", [(0, 0), (1, 0), (1, 2)])
def test_turbdrv_matches_reference(libs, moist, cphopt, wbc, bbc):
    """The whole of s_turbdrv (coefficients, strsten, turbuvw) against the
    reference's own s_turbdrv: the velocity forcings and every array it
    returns, bitwise."""
    o, r = libs
    c = cfg(wbc, cphopt, bbc)
    r.configure(c)
    d, outs = drv_inputs(21 + cphopt + (wbc > 0) + 3 * bbc + 7 * moist)
    mine = {n: a.copy(order="F") for n, a in outs.items()}
    ref = {n: a.copy(order="F") for n, a in outs.items()}
    o.or_turbdrv(c, DTB, moist, *[d[n] for n in IN], *[mine[n] for n in DRV_OUT], *[fld() for _ in range(4)])
    r.ref_turbdrv(moist, DTB, NI, NJ, NK, *[d[n] for n in ("jcb", "jcb8u", "jcb8v", "jcb8w", "pbr", "ptbr", "rbr",
                                                          "up", "vp", "wp", "ppp", "ptpp", "qvp", "tkep", "qallp")],
                  *[ref[n] for n in DRV_OUT], *[fld() for _ in range(3)])
    for n in DRV_OUT:
        if n in INTERIOR:
            np.testing.assert_array_equal(mine[n][1:NI, 1:NJ, 0:NK - 1], ref[n][1:NI, 1:NJ, 0:NK - 1], err_msg=n)
        else:
            np.testing.assert_array_equal(mine[n], ref[n], err_msg=n)
    for n in ("ufrc", "vfrc", "wfrc"):
        assert not np.array_equal(mine[n], outs[n])


@pytest.mark.parametrize("wbc", [-1, 7])
def test_turbflx_matches_reference(libs, wbc):
    """s_turbflx (trnopt=0, sfcopt=0): the fluxes of a scalar, bitwise."""
    o, r = libs
    c = cfg(wbc, 0, 2)
    r.configure(c)
    rng = np.random.default_rng(3 + (wbc > 0))
    f = lambda lo, hi: np.asfortranarray(rng.uniform(lo, hi, (NI + 2, NJ + 2, NK)).astype(np.float32))
    jcb, s, kh, kv, kw = f(0.9, 1.1), f(290.0, 310.0), f(0.0, 50.0), f(0.0, 50.0), f(0.0, 20.0)
    h = [f(-1.0, 1.0) for _ in range(3)]
    mine = [a.copy(order="F") for a in h]
    ref = [a.copy(order="F") for a in h]
    o.or_turbflx(c, jcb, s, kh, kv, kw, *mine)
    r.ref_turbflx(NI, NJ, NK, jcb, s, kh, kv, kw, *ref, *[fld() for _ in range(4)])
    for a, b, n in zip(mine, ref, ("h1", "h2", "h3")):
        np.testing.assert_array_equal(a, b, err_msg=n)
        assert not np.array_equal(a, h[("h1", "h2", "h3").index(n)])


@pytest.mark.parametrize("wbc", [-1, 7])
def test_turbs_matches_reference(libs, wbc):
    """s_turbs (trnopt=0, mfcopt=0): the flux divergence added to the
    forcing, bitwise."""
    o, r = libs
    c = cfg(wbc, 0, 2)
    r.configure(c)
    rng = np.random.default_rng(8 + (wbc > 0))
    f = lambda lo, hi: np.asfortranarray(rng.uniform(lo, hi, (NI + 2, NJ + 2, NK)).astype(np.float32))
    ju, jv, h1, h2, h3, sf = f(0.9, 1.1), f(0.9, 1.1), f(-5.0, 5.0), f(-5.0, 5.0), f(-5.0, 5.0), f(-1.0, 1.0)
    mine, ref = sf.copy(order="F"), sf.copy(order="F")
    o.or_turbs(c, ju, jv, h1, h2, h3, mine)
    r.ref_turbs(NI, NJ, NK, ju, jv, h1, h2, h3, ref, *[fld() for _ in range(4)])
    np.testing.assert_array_equal(mine, ref)
    assert not np.array_equal(mine, sf)


def test_kh8uv_matches_reference(libs):
    """s_kh8uv (mfcopt=0), bitwise over whole arrays."""
    o, r = libs
    c = cfg(-1, 0, 2)
    r.configure(c)
    rng = np.random.default_rng(4)
    f = lambda: np.asfortranarray(rng.uniform(0.0, 50.0, (NI + 2, NJ + 2, NK)).astype(np.float32))
    kh, u0, v0 = f(), f(), f()
    mu, mv, ru, rv = u0.copy(order="F"), v0.copy(order="F"), u0.copy(order="F"), v0.copy(order="F")
    o.or_kh8uv(c, kh, mu, mv)
    r.ref_kh8uv(NI, NJ, NK, kh, ru, rv, fld())
    np.testing.assert_array_equal(mu, ru)
    np.testing.assert_array_equal(mv, rv)
    assert not np.array_equal(mu, u0)


@pytest.mark.parametrize("advopt", [2, 3])
@pytest.mark.parametrize("vsp", [0, 1])
@pytest.mark.parametrize("wbc", [-1, 7])
def test_forcetke_matches_reference(libs, wbc, vsp, advopt):
    """s_forcetke (tubopt=2, isoopt=1, trnopt=0, mfcopt=0, no smoothing or
    lateral sponge) against the reference's own: the TKE forcing and the
    diffusivities it leaves in rkh, rkv and rkv8s, bitwise over whole
    arrays, with the vertical sponge on TKE on and off."""
    o, r = libs
    c = po.make_cfg(NI, NJ, NK, dx=700.0, dy=900.0, dz=350.0, wbc=wbc, advopt=advopt, cphopt=0, vspopt=2 * vsp,
                    ksp0=(2, 8), vspvar=b"ooooooooo" + (b"o" if vsp else b"x"), gwave=30.0)
    r.configure(c)
    d, outs = tke_inputs(NI, NJ, NK, 31 + (wbc > 0) + 2 * vsp + 4 * advopt)
    mine = {n: a.copy(order="F") for n, a in outs.items()}
    ref = {n: a.copy(order="F") for n, a in outs.items()}
    ssq_ref = d["ssq"].copy(order="F")
    o.or_forcetke(c, *[d[n] for n in TKE_IN], *[mine[n] for n in TKE_OUT], *[fld() for _ in range(4)])
    r.ref_forcetke(8, NI, NJ, NK, *[d[n] for n in ("jcb", "jcb8u", "jcb8v", "jcb8w")], fld(),
                   *[d[n] for n in ("rst", "rstxu", "rstxv", "rstxwc", "tke", "tkep", "rbct", "priv", "nsq8w")],
                   ssq_ref, ref["rkh"], ref["rkv"], ref["tkefrc"], fld(), ref["rkv8s"], *[fld() for _ in range(3)],
                   np.zeros((NI, NJ), np.float32, order="F"))
    for n in TKE_OUT:
        np.testing.assert_array_equal(mine[n], ref[n], err_msg=n)
        assert not np.array_equal(mine[n], outs[n]), n
    # the sponge changes the forcing only from level ksp0(2)-1 up
    I, J = slice(2, NI - 1), slice(2, NJ - 1)
    if vsp:
        off = {n: a.copy(order="F") for n, a in outs.items()}
        c0 = po.make_cfg(NI, NJ, NK, dx=700.0, dy=900.0, dz=350.0, wbc=wbc, advopt=advopt, cphopt=0, vspopt=0,
                         ksp0=(2, 8), gwave=30.0)
        o.or_forcetke(c0, *[d[n] for n in TKE_IN], *[off[n] for n in TKE_OUT], *[fld() for _ in range(4)])
        diff = mine["tkefrc"][I, J, 1:NK - 2] != off["tkefrc"][I, J, 1:NK - 2]
        # Fortran level 7 = ksp0(2)-1 is index 5 of the slice from level 2
        assert not diff[:, :, :5].any() and diff[:, :, 5:].mean() > 0.6   # a fifth of the cells hold no TKE


def test_forcetke_dissipation_closed_form(libs):
    """With no wind, uniform TKE, S^2 = N^2 = 0 and unit metrics, only the
    dissipation is left: -(.37 priv - .18) e^1.5 / ((priv-1) (dx dy dz/8)^(1/3))
    (Deardorff's, as disptke.f90 writes it), and the diffusivities become
    rkh = priv rkv, rkv8w the mean of the two levels, rkv8s half of it."""
    o, _ = libs
    c = po.make_cfg(NI, NJ, NK, dx=800.0, dy=800.0, dz=400.0, wbc=-1, cphopt=0, vspopt=0)
    d, outs = tke_inputs(NI, NJ, NK, 5)
    one = np.ones_like(d["jcb"], order="F")
    for n in ("jcb", "jcb8u", "jcb8v", "jcb8w", "rst"):
        d[n] = one.copy(order="F")
    for n in ("rstxu", "rstxv", "rstxwc", "nsq8w", "ssq"):
        d[n] = np.zeros_like(one, order="F")
    d["tke"] = np.full_like(one, 1.5, order="F")
    d["tkep"] = d["tke"].copy(order="F")
    d["priv"] = np.full_like(one, 2.5, order="F")
    kv = 10.0 * np.arange(1, NK + 1, dtype=np.float32)[None, None, :] * one
    outs["rkv"] = np.asfortranarray(kv)
    outs["rkh"] = outs["rkv"].copy(order="F")
    o.or_forcetke(c, *[d[n] for n in TKE_IN], *[outs[n] for n in TKE_OUT], *[fld() for _ in range(4)])
    ln = 1.5 * (800.0 * 800.0 * 400.0 / 8.0) ** (1.0 / 3.0)
    want = -(0.37 * 2.5 - 0.18) * 1.5 ** 1.5 / ln
    np.testing.assert_allclose(outs["tkefrc"][2:NI - 1, 2:NJ - 1, 1:NK - 2], want, rtol=2e-6)
    k = np.arange(1, NK, dtype=np.float64)
    np.testing.assert_allclose(outs["rkh"][3, 4, 0:NK - 1], 2.5 * 10.0 * k, rtol=1e-6)
    np.testing.assert_allclose(outs["rkv"][3, 4, 1:NK - 1], 2.5 * 10.0 * (k[1:] - 0.5), rtol=1e-6)
    np.testing.assert_allclose(outs["rkv8s"][3, 4, 0:NK - 1], 0.5 * 2.5 * 10.0 * k, rtol=1e-6)
