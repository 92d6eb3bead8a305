"""Pin the oracle (oracle/cress_oracle.c, FP32 build) against the reference's
own compiled Fortran routines (oracle/_ref/libcressref32.so, built from the
reference sources by oracle/build_ref.py) on identical seeded inputs.

The restatement mirrors loop ranges and operation order, so agreement is
required to be bitwise over the whole array (halo points included: a wrong
loop bound shows up as a mismatch in a halo row).  Skipped where the
reference library was not built (no reference sources at build time).
"""
import numpy as np
import pytest

from oracle import pyoracle as po

pytestmark = pytest.mark.skipif(not po.have_ref(), reason="oracle/_ref not built")

NI, NJ, NK = 13, 11, 14


@pytest.fixture(scope="module")
def libs():
    return po.Oracle(32), po.Ref()


def cfg_for(advopt=3, **kw):
    return po.make_cfg(NI, NJ, NK, dx=700.0, dy=900.0, dz=350.0, advopt=advopt, **kw)


def rnd(cfg, rng, lo=-1.0, hi=1.0, n=1):
    a = po.field(cfg, np.float32, n)
    a[...] = rng.uniform(lo, hi, a.shape)
    return a


def same(a, b):
    """Bitwise equal except for the sign of zero: the reference forms exact
    zeros as (u1+u2)*j31 with j31=0 on flat terrain (vbcw.f90:207-318), whose
    sign follows u; -0 and +0 enter every later expression identically."""
    np.testing.assert_array_equal(a, b)


def run_both(libs, cfg, name_or, name_ref, or_args, ref_args):
    o, r = libs
    r.configure(cfg)
    getattr(o, name_or)(cfg, *or_args)
    getattr(r, name_ref)(*ref_args)


# wbc=7: open sides, where the 4th-order terms drop to 2nd order next to
# the boundary (iwest..jnorth = 1, sethalo.f90:536-580)
@pytest.mark.parametrize("wbc", [-1, 7])
@pytest.mark.parametrize("advopt", [1, 2, 3])
def test_advs(libs, advopt, wbc):
    rng = np.random.default_rng(10 + advopt)
    c = cfg_for(advopt, wbc=wbc)
    ux, vx, wx, s = (rnd(c, rng) for _ in range(4))
    outs = [[rnd(c, rng) for _ in range(5)]]
    outs.append([a.copy(order="F") for a in outs[0]])
    o, r = libs
    r.configure(c)
    o.or_advs(c, ux, vx, wx, s, *outs[0])
    r.ref_advs(NI, NJ, NK, ux, vx, wx, s, *outs[1])
    same(outs[0][0], outs[1][0])   # sfrc


@pytest.mark.parametrize("wbc", [-1, 7])
@pytest.mark.parametrize("advopt", [1, 2, 3])
def test_advuvw(libs, advopt, wbc):
    rng = np.random.default_rng(20 + advopt)
    c = cfg_for(advopt, wbc=wbc)
    ins = [rnd(c, rng) for _ in range(6)]
    a = [rnd(c, rng) for _ in range(8)]
    b = [x.copy(order="F") for x in a]
    o, r = libs
    r.configure(c)
    o.or_advuvw(c, *ins, *a)
    r.ref_advuvw(NI, NJ, NK, *ins, *b)
    for x, y in zip(a[:3], b[:3]):
        same(x, y)


@pytest.mark.parametrize("wbc", [-1, 7])
def test_rstuvwc(libs, wbc):
    rng = np.random.default_rng(40)
    c = cfg_for(wbc=wbc)
    r8u, r8v, r8w, u, v, wc = (rnd(c, rng) for _ in range(6))
    mf = rnd(c, rng)
    a = [rnd(c, rng) for _ in range(3)]
    b = [x.copy(order="F") for x in a]
    o, r = libs
    r.configure(c)
    o.or_rstuvwc(c, r8u, r8v, r8w, u, v, wc, *a)
    r.ref_rstuvwc(NI, NJ, NK, mf, mf, r8u, r8v, r8w, u, v, wc, *b)
    for x, y in zip(a, b):
        same(x, y)


def test_buoywb(libs):
    rng = np.random.default_rng(50)
    c = cfg_for()
    ptbr = rnd(c, rng, 290, 400)
    qvbr = rnd(c, rng, 0, 0.02)
    rst = rnd(c, rng, 0.2, 1.3)
    ptp = rnd(c, rng, -2, 2)
    qv = rnd(c, rng, 0, 0.02)
    qall = rnd(c, rng, 0, 0.01)
    a = [rnd(c, rng) for _ in range(3)]
    b = [x.copy(order="F") for x in a]
    o, r = libs
    r.configure(c)
    o.or_buoywb(c, ptbr, qvbr, rst, ptp, qv, qall, *a)
    r.ref_buoywb(NI, NJ, NK, ptbr, qvbr, rst, ptp, qv, qall, *b)
    same(a[0], b[0])


def test_coriuv(libs):
    rng = np.random.default_rng(60)
    c = cfg_for(coropt=1)
    fc = po.field(c, np.float32, 2)
    fc[...] = rng.uniform(1e-5, 3e-5, fc.shape)
    rst, u, v = (rnd(c, rng) for _ in range(3))
    a = [rnd(c, rng) for _ in range(3)]
    b = [x.copy(order="F") for x in a]
    o, r = libs
    r.configure(c)
    o.or_coriuv(c, fc[:, :, 0].copy(order="F"), rst, u, v, *a)
    r.ref_coriuv(NI, NJ, NK, fc, rst, u, v, *b)
    same(a[0], b[0])
    same(a[1], b[1])


def sponge(c, rng):
    rbct = np.zeros((c.ni, c.nj, c.nk, 2), np.float32, order="F")
    rbct[...] = rng.uniform(0, 0.003, rbct.shape)
    return rbct


def test_vspuvw(libs):
    rng = np.random.default_rng(70)
    c = cfg_for(ksp0=(5, 6))
    ins = [rnd(c, rng) for _ in range(8)]
    rbct = sponge(c, rng)
    dmy = rnd(c, rng)
    a = [rnd(c, rng) for _ in range(4)]
    b = [x.copy(order="F") for x in a]
    o, r = libs
    r.configure(c)
    o.or_vspuvw(c, *ins, rbct, *a)
    k0 = (np.array([5, 6], dtype=np.int32))
    r.ref_vspuvw(k0, NI, NJ, NK, *ins, rbct, dmy, *b)
    for x, y in zip(a[:3], b[:3]):
        same(x, y)


def forcing_inputs(c, rng):
    d = {}
    for k in ("jcb", "jcb8u", "jcb8v", "jcb8w", "rst"):
        d[k] = rnd(c, rng, 0.6, 1.4)
    d["rbr"] = rnd(c, rng, 0.3, 1.2)
    d["ptbr"] = rnd(c, rng, 290, 400)
    d["qvbr"] = rnd(c, rng, 0, 0.02)
    for k in ("u", "v", "wc", "w", "pp", "ppp", "ptp", "ptpp", "qv", "qvp", "rstxu", "rstxv",
              "rstxwc"):
        d[k] = rnd(c, rng)
    d["rbct"] = sponge(c, rng)
    return d


@pytest.mark.parametrize("wbc", [-1, 7])
def test_forcep(libs, wbc):
    rng = np.random.default_rng(80)
    c = cfg_for(ksp0=(4, 5), vspvar=b"oooooooooo", wbc=wbc)
    d = forcing_inputs(c, rng)
    dmy = rnd(c, rng, n=2)
    a = [rnd(c, rng) for _ in range(9)]
    b = [x.copy(order="F") for x in a]
    o, r = libs
    r.configure(c)
    o.or_advp(c, d["jcb8u"], d["jcb8v"], d["jcb8w"], d["u"], d["v"], d["wc"], d["pp"], a[0],
              *a[1:8])
    o.or_vsps(c, d["jcb"], d["ppp"], None, d["rbct"], a[0])
    r.ref_forcep(np.array([4, 5], np.int32), NI, NJ, NK, d["jcb"], d["jcb8u"], d["jcb8v"],
                 d["jcb8w"], d["u"], d["v"], d["wc"], d["pp"], d["ppp"], d["rbct"], dmy, b[0],
                 *b[1:9])
    same(a[0], b[0])


def test_forcept_forceqv(libs):
    rng = np.random.default_rng(90)
    c = cfg_for(ksp0=(4, 5), vspvar=b"oooooooooo")
    d = forcing_inputs(c, rng)
    dmy = rnd(c, rng, n=2)
    o, r = libs
    r.configure(c)
    a = [rnd(c, rng) for _ in range(8)]
    b = [x.copy(order="F") for x in a]
    o.or_advs(c, d["rstxu"], d["rstxv"], d["rstxwc"], d["ptp"], a[0], *a[1:5])
    o.or_vsps(c, d["rst"], d["ptpp"], None, d["rbct"], a[0])
    o.or_advbspt(c, d["ptbr"], d["rbr"], d["w"], a[0], a[1])
    r.ref_forcept(np.array([4, 5], np.int32), NI, NJ, NK, d["jcb"], d["ptbr"], d["rbr"], d["rst"],
                  d["rstxu"], d["rstxv"], d["rstxwc"], d["w"], d["ptp"], d["ptpp"], d["rbct"],
                  dmy, b[0], *b[1:8])
    same(a[0], b[0])
    a = [rnd(c, rng) for _ in range(7)]
    b = [x.copy(order="F") for x in a]
    o.or_advs(c, d["rstxu"], d["rstxv"], d["rstxwc"], d["qv"], a[0], *a[1:5])
    o.or_vsps(c, d["rst"], d["qvp"], d["qvbr"], d["rbct"], a[0])
    r.ref_forceqv(np.array([4, 5], np.int32), NI, NJ, NK, d["jcb"], d["qvbr"], d["rbr"],
                  d["rst"], d["rstxu"], d["rstxv"], d["rstxwc"], d["qv"], d["qvp"], d["rbct"],
                  dmy, b[0], *b[1:7])
    same(a[0], b[0])


def test_forceq(libs):
    rng = np.random.default_rng(95)
    c = cfg_for(ksp0=(4, 5), vspvar=b"oooooooooo")
    d = forcing_inputs(c, rng)
    qw, qwp = rnd(c, rng, 0, 1e-3, 2), rnd(c, rng, 0, 1e-3, 2)
    qi, qip = rnd(c, rng, 0, 1e-3, 3), rnd(c, rng, 0, 1e-3, 3)
    ni_, nip = rnd(c, rng, 0, 1e5), rnd(c, rng, 0, 1e5)
    dmy = rnd(c, rng, n=4)
    o, r = libs
    r.configure(c)
    fa = [rnd(c, rng, n=2), rnd(c, rng, n=3), rnd(c, rng)]
    fb = [x.copy(order="F") for x in fa]
    t = [rnd(c, rng) for _ in range(6)]
    for frc, q, qp in ((fa[0], qw, qwp), (fa[1], qi, qip)):
        for s in range(q.shape[3]):
            out = np.asfortranarray(frc[..., s])
            o.or_advs(c, d["rstxu"], d["rstxv"], d["rstxwc"], np.asfortranarray(q[..., s]), out,
                      *t[:4])
            o.or_vsps(c, d["rst"], np.asfortranarray(qp[..., s]), None, d["rbct"], out)
            frc[..., s] = out
    o.or_advs(c, d["rstxu"], d["rstxv"], d["rstxwc"], ni_, fa[2], *t[:4])
    o.or_vsps(c, d["rst"], nip, None, d["rbct"], fa[2])
    r.ref_forceq(np.array([4, 5], np.int32), NI, NJ, NK, d["jcb"], d["rbr"], d["rst"], d["rstxu"],
                 d["rstxv"], d["rstxwc"], qw, qwp, qi, qip, ni_, nip, d["rbct"], dmy,
                 fb[0], fb[1], fb[2], *t)
    for x, y in zip(fa, fb):
        same(x, y)


@pytest.mark.parametrize("divopt", [0, 1])
def test_pgrad(libs, divopt):
    rng = np.random.default_rng(100 + divopt)
    c = cfg_for(divopt=divopt)
    jcb, r8u, r8v, r8w = (rnd(c, rng, 0.6, 1.4) for _ in range(4))
    u, v, wc, pp = (rnd(c, rng) for _ in range(4))
    dmy = rnd(c, rng, n=4)
    a = [rnd(c, rng) for _ in range(6)]
    b = [x.copy(order="F") for x in a]
    o, r = libs
    r.configure(c)
    o.or_pgrad(c, 0.75, jcb, r8u, r8v, r8w, u, v, wc, pp, *a)
    r.ref_pgrad(po.C.c_float(0.75), NI, NJ, NK, jcb, r8u, r8v, r8w, u, v, wc, pp, *b, dmy)
    for x, y in zip(a[:3], b[:3]):
        same(x, y)


def test_acoustic_pieces(libs):
    rng = np.random.default_rng(110)
    c = cfg_for()
    o, r = libs
    r.configure(c)
    j8u, j8v, rcsq, jcb, rst, rbr, ptbr = (rnd(c, rng, 0.5, 1.5) for _ in range(7))
    u, v, w, pfrc, wfrc, pp, ptp = (rnd(c, rng) for _ in range(7))
    dmy = rnd(c, rng, n=4)
    # diverpih
    a = [rnd(c, rng) for _ in range(3)]
    b = [x.copy(order="F") for x in a] + [rnd(c, rng)]
    o.or_diverpih(c, j8u, j8v, rcsq, u, v, *a)
    r.ref_diverpih(NI, NJ, NK, j8u, j8v, rcsq, u, v, *b, dmy)
    same(a[0], b[0])
    # diverpiv
    a, b = rnd(c, rng), None
    b = a.copy(order="F")
    o.or_diverpiv(c, rcsq, w, a)
    r.ref_diverpiv(NI, NJ, NK, rcsq, w, b)
    same(a, b)
    # advbspi back / fore
    ph, pv = rnd(c, rng), rnd(c, rng)
    for fore in (0, 1):
        a = rnd(c, rng)
        b = a.copy(order="F")
        o.or_advbspi(c, fore, rst, w, pfrc, ph, pv, a)
        r.ref_advbspi(fore, NI, NJ, NK, rst, w, pfrc, ph, pv, b)
        same(a, b)
    # pgradiv
    fp = rnd(c, rng)
    a = [rnd(c, rng), rnd(c, rng)]
    b = [x.copy(order="F") for x in a]
    o.or_pgradiv(c, 0.5, jcb, wfrc, fp, *a)
    r.ref_pgradiv(po.C.c_float(0.5), NI, NJ, NK, jcb, wfrc, fp, *b)
    same(a[0], b[0])
    # buoywsi
    a = [rnd(c, rng), rnd(c, rng)]
    b = [x.copy(order="F") for x in a]
    o.or_buoywsi(c, 0.5, rbr, rst, rcsq, pp, fp, *a)
    r.ref_buoywsi(po.C.c_float(0.5), NI, NJ, NK, rbr, ptbr, rst, rcsq, pp, ptp, fp, *b)
    same(a[0], b[0])


def test_gaussel(libs):
    """Diagonally dominant systems: bitwise vs the reference and, in FP64 via
    numpy, close to a dense solve (the Thomas algorithm is exact up to
    rounding)."""
    rng = np.random.default_rng(120)
    c = cfg_for()
    rr, tt, ff = (rnd(c, rng) for _ in range(3))
    ss = rnd(c, rng, 3.0, 5.0)
    a = [x.copy(order="F") for x in (rr, ss, tt, ff)]
    b = [x.copy(order="F") for x in (rr, ss, tt, ff)]
    o, r = libs
    r.configure(c)
    o.or_gaussel(c, 2, NI - 2, 2, NJ - 2, 3, NK - 2, *a)
    r.ref_gaussel(2, NI - 2, 2, NJ - 2, 3, NK - 2, NI, NJ, NK, *b, rnd(c, rng))
    same(a[3], b[3])
    # closed form: solve one column densely in float64
    i, j = 4, 5
    ks = list(range(3, NK - 1))      # Fortran k = 3..nk-2 -> 0-based index k-1
    m = len(ks)
    A = np.zeros((m, m))
    for t, k in enumerate(ks):
        A[t, t] = ss[i, j, k - 1]
        if t > 0:
            A[t, t - 1] = rr[i, j, k - 1]
        if t < m - 1:
            A[t, t + 1] = tt[i, j, k - 1]
    x = np.linalg.solve(A, ff[i, j, [k - 1 for k in ks]].astype(np.float64))
    np.testing.assert_allclose(a[3][i, j, [k - 1 for k in ks]], x, rtol=2e-5, atol=1e-6)


def test_boundary_fills(libs):
    rng = np.random.default_rng(130)
    o, r = libs
    for bbc, tbc in ((2, 2), (3, 3), (2, 3)):
        c = cfg_for(bbc=bbc, tbc=tbc)
        r.configure(c)
        zeros = po.field(c, np.float32)
        for kind, name in ((0, "or_vbcs"), (1, "or_vbcu"), (2, "or_vbcv"), (3, "or_vbcp"),
                           (4, "or_vbcwc")):
            a = rnd(c, rng)
            b = a.copy(order="F")
            getattr(o, name)(c, a)
            r.ref_vbc(kind, NI, NJ, NK, b, zeros, zeros, zeros, zeros, zeros, rnd(c, rng),
                      rnd(c, rng))
            same(a, b)
        j8w, wc, uf, vf = rnd(c, rng), rnd(c, rng), rnd(c, rng), rnd(c, rng)
        a = rnd(c, rng)
        b = a.copy(order="F")
        o.or_vbcw(c, j8w, wc, a)
        r.ref_vbc(5, NI, NJ, NK, b, zeros, j8w, uf, vf, wc, rnd(c, rng), rnd(c, rng))
        same(a, b)
    c = cfg_for()
    r.configure(c)
    for args in ((2, 1, NI - 2, NI - 1, 2, 1, NJ - 2, NJ - 1), (3, 1, NI - 2, NI, 2, 1, NJ - 2, NJ - 1),
                 (4, 0, NI - 3, NI + 1, 3, 0, NJ - 3, NJ), (3, 0, NI - 3, NI, 4, 0, NJ - 3, NJ + 1)):
        a = rnd(c, rng)
        b = a.copy(order="F")
        o.or_bcycle(c, *args, NK, a)
        r.ref_bcycle(*args, NI, NJ, NK, b)
        same(a, b)


def test_totalqwi(libs):
    rng = np.random.default_rng(140)
    c = cfg_for()
    o, r = libs
    r.configure(c)
    qw, qi = rnd(c, rng, 0, 1, 2), rnd(c, rng, 0, 1, 3)
    a = rnd(c, rng)
    b = a.copy(order="F")
    o.or_totalqwi(c, qw, qi, a)
    r.ref_totalqwi(NI, NJ, NK, qw, qi, b)
    same(a, b)


def test_timeflt(libs):
    """or_timeflt (Asselin filter of all twelve prognostics over each field's
    own range) against the reference's s_timeflt, whole arrays."""
    rng = np.random.default_rng(140)
    c = cfg_for(cphopt=2)
    o, r = libs
    r.configure(c)
    names = ["u", "v", "w", "pp", "ptp", "q", "nci"]
    nsp = {"q": 6}
    lv = {m: [rnd(c, rng, n=nsp.get(m, 1)) for _ in range(3)] for m in names}
    ref = {m: [a.copy(order="F") for a in lv[m]] for m in names}
    st = po.State()
    for m in names:
        arr = getattr(st, m)
        for level in range(3):
            arr[level] = lv[m][level].ctypes.data
    o.lib.or_timeflt(po.C.byref(c), po.C.byref(st))
    n = lv["u"][0].size
    q = [a.reshape(-1, order="F") for a in ref["q"]]
    dmy = po.field(c, np.float32, 8)
    r.ref_timeflt(NI, NJ, NK, ref["u"][0], ref["u"][2], ref["v"][0], ref["v"][2], ref["w"][0], ref["w"][2],
                  ref["pp"][0], ref["pp"][2], ref["ptp"][0], ref["ptp"][2], q[0][:n], q[2][:n],
                  q[0][n:3 * n], q[2][n:3 * n], q[0][3 * n:], q[2][3 * n:], ref["nci"][0], ref["nci"][2],
                  ref["u"][1], ref["v"][1], ref["w"][1], ref["pp"][1], ref["ptp"][1], q[1][:n],
                  q[1][n:3 * n], q[1][3 * n:], ref["nci"][1], dmy)
    for m in names:
        same(lv[m][1], ref[m][1])
