"""Pin the oracle's cold-rain microphysics (oracle/cress_micro.c, FP32)
against the reference's own compiled routines on identical seeded states
that exercise every branch (warm/cold, sub/supersaturated, each species
present or absent).  The oracle calls the same C library exp/log/sqrt as the
reference's FP32 build on this host, so agreement is required bitwise
(signed zeros aside)."""
import numpy as np
import pytest

from oracle import pyoracle as po

pytestmark = pytest.mark.skipif(not po.have_ref(), reason="oracle/_ref not built")

NI, NJ, NK = 13, 11, 14


@pytest.fixture(scope="module")
def libs():
    o, r = po.Oracle(32), po.Ref()
    r.lib.ref_setref()
    return o, r


def cfg():
    return po.make_cfg(NI, NJ, NK, dx=700.0, dy=900.0, dz=350.0)


def state(seed):
    """A random but physically plausible past/future state (FP32)."""
    rng = np.random.default_rng(seed)
    c = cfg()
    f = lambda: po.field(c, np.float32)
    shp = f().shape
    T = rng.uniform(215.0, 305.0, shp)
    p = rng.uniform(2.0e4, 1.0e5, shp)
    pp = rng.uniform(-200.0, 200.0, shp)
    pbr = p - pp
    pi = (p / 1e5) ** (287.0 / 1004.0)
    ptp = rng.uniform(-2.0, 2.0, shp)
    ptbr = T / pi - ptp
    rbr = p / (287.0 * T)
    es = 610.78 * np.exp(17.269 * (T - 273.16) / (T - 35.86))
    qvs = 0.622 * es / (p - es)
    qv = qvs * rng.uniform(0.5, 1.3, shp)

    def hyd(lo, hi, frac=0.6):
        v = np.exp(rng.uniform(np.log(lo), np.log(hi), shp))
        return np.where(rng.uniform(0, 1, shp) < frac, v, 0.0)

    q = {"qv": qv, "qc": hyd(1e-7, 3e-3), "qr": hyd(1e-7, 5e-3), "qi": hyd(1e-8, 2e-3),
         "qs": hyd(1e-7, 4e-3), "qg": hyd(1e-7, 6e-3)}
    nci = q["qi"] / np.exp(rng.uniform(np.log(2e-12), np.log(1e-10), shp))
    F = lambda a: np.asfortranarray(a.astype(np.float32))
    d = {k: F(v) for k, v in q.items()}
    d.update(pbr=F(pbr), pp=F(pp), ptbr=F(ptbr), ptp=F(ptp), rbr=F(rbr), nci=F(nci))
    # future level: a perturbed copy
    for k in ("qv", "qc", "qr", "qi", "qs", "qg"):
        d[k + "f"] = F(np.maximum(d[k] * rng.uniform(0.5, 1.5, shp), 0.0))
    d["ptpf"] = F(ptp + rng.uniform(-0.5, 0.5, shp))
    d["ncif"] = F(nci * rng.uniform(0.5, 1.5, shp))
    return c, d


def same(a, b, what=""):
    np.testing.assert_array_equal(a, b, err_msg=what)


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_prologue_termblk_srcblk(libs, seed):
    o, r = libs
    c, d = state(seed)
    r.configure(c)
    n = (NI + 2) * (NJ + 2) * NK
    F = lambda k=1: po.field(c, np.float32, k)
    # prologue: getvdens, getexner, diagnw, diagnsg
    qw = np.asfortranarray(np.stack([d["qc"], d["qr"]], axis=-1))
    qi = np.asfortranarray(np.stack([d["qi"], d["qs"], d["qg"]], axis=-1))
    rbv_r, pi_r, p_r, nw_r, nic_r = F(), F(), F(), F(2), F(3)
    r.ref_micro_prologue(NI, NJ, NK, d["rbr"], d["pbr"], d["pp"], qw, qi, rbv_r, pi_r, p_r, nw_r, nic_r)
    rbv, pi, p = F(), F(), F()
    ncc, ncr, ncs, ncg = F(), F(), F(), F()
    o.or_getvdens(c, d["rbr"], rbv)
    o.or_getexner(c, d["pbr"], d["pp"], pi, p)
    o.or_diagnw(c, d["rbr"], d["qc"], d["qr"], ncc, ncr)
    o.or_diagnsg(c, d["rbr"], rbv, d["qs"], d["qg"], ncs, ncg)
    same(rbv, rbv_r, "rbv")
    same(pi, pi_r, "pi")
    same(p, p_r, "p")
    same(ncc, nw_r[..., 0], "ncc")
    same(ncr, nw_r[..., 1], "ncr")
    same(ncs, nic_r[..., 1], "ncs")
    same(ncg, nic_r[..., 2], "ncg")
    # termblk
    uo = [F() for _ in range(10)]
    ur = [F() for _ in range(10)] + [F()]
    o.or_termblk(c, rbv, d["qc"], d["qr"], d["qi"], d["qs"], d["qg"], ncr, ncs, ncg, *uo)
    r.ref_termblk(NI, NJ, NK, rbv, d["qc"], d["qr"], d["qi"], d["qs"], d["qg"], ncc, ncr, d["nci"],
                  ncs, ncg, *ur)
    for a, b, nm in zip(uo, ur, ["ucq", "urq", "uiq", "usq", "ugq", "ucn", "urn", "uin", "usn", "ugn"]):
        same(a, b, nm)
    # srcblk
    inp = np.zeros(22 * n, np.float32)
    fields = [d["ptbr"], d["rbr"], rbv, pi, p, d["ptp"], d["qv"], d["qc"], d["qr"], d["qi"], d["qs"],
              d["qg"], ncc, ncr, d["nci"], ncs, ncg, uo[1], uo[3], uo[4]]
    for k, a in enumerate(fields):
        inp[k * n:(k + 1) * n] = a.ravel(order="F")
    fut0 = np.zeros(8 * n, np.float32)
    for k, nm in enumerate(["ptpf", "qvf", "qcf", "qrf", "qif", "qsf", "qgf", "ncif"]):
        fut0[k * n:(k + 1) * n] = d[nm].ravel(order="F")
    fo, fr = fut0.copy(), fut0.copy()
    wo, wr = np.zeros(62 * n, np.float32), np.zeros(80 * n, np.float32)
    dtb = np.float32(6.0)
    o.or_srcblk_block(c, po.C.c_float(dtb), inp, fo, wo)
    r.ref_srcblk(NI, NJ, NK, po.C.c_float(dtb), po.C.c_float(1e-12), inp, fr, wr)
    names = ["ptpf", "qvf", "qcf", "qrf", "qif", "qsf", "qgf", "ncif"]
    for k, nm in enumerate(names):
        same(fo[k * n:(k + 1) * n], fr[k * n:(k + 1) * n], nm)
    # the process rates themselves (reference work-array order)
    rates = {20: "nuvi", 21: "nuci", 22: "clcr", 23: "clcs", 24: "clcg", 25: "clri", 26: "clrs",
             27: "clrg", 28: "clir", 29: "clis", 30: "clig", 31: "clsr", 32: "clsg", 33: "clrsg",
             40: "agin", 42: "vdvr", 43: "vdvi", 44: "vdvs", 45: "vdvg", 46: "cncr", 47: "cnis",
             48: "cnsg", 50: "spsi", 51: "spgi", 52: "mlic", 53: "mlsr", 54: "mlgr", 55: "frrg",
             57: "shsr", 58: "shgr", 59: "pgwet"}
    mask = np.zeros((NI + 2, NJ + 2, NK), bool)
    mask[1:NI, 1:NJ, 0:NK - 1] = True
    m = mask.ravel(order="F")
    for k, nm in rates.items():
        same(wo[k * n:(k + 1) * n][m], wr[k * n:(k + 1) * n][m], nm)
    # every branch was exercised: some but not all cells active per process
    for k in (21, 22, 26, 32, 43, 44, 47, 52, 55):
        v = wr[k * n:(k + 1) * n][m]
        assert np.count_nonzero(v) > 0, rates[k]


def test_sedimentation(libs):
    o, r = libs
    c, d = state(7)
    r.configure(c)
    F = lambda k=1: po.field(c, np.float32, k)
    rst = np.asfortranarray(d["rbr"] * np.float32(1.1))
    u = np.asfortranarray(np.abs(np.random.default_rng(3).normal(3.0, 2.0, d["qr"].shape)).astype(np.float32))
    for dtp in (0.7, 2.5):
        qa, qb = d["qrf"].copy(order="F"), d["qrf"].copy(order="F")
        pa = np.zeros((NI + 2, NJ + 2, 2), np.float32, order="F")
        pb = pa.copy(order="F")
        o.or_upwqp(c, po.C.c_float(dtp), d["rbr"], rst, u, qa, pa, F())
        r.ref_upwqp(po.C.c_float(dtp), NI, NJ, NK, d["rbr"], rst, u, qb, pb, F())
        same(qa, qb, "qr")
        same(pa, pb, "precip")
        na, nb = d["ncif"].copy(order="F"), d["ncif"].copy(order="F")
        o.or_upwnp(c, po.C.c_float(dtp), d["rbr"], rst, u, na, F())
        r.ref_upwnp(po.C.c_float(dtp), NI, NJ, NK, d["rbr"], rst, u, nb, F())
        same(na, nb, "nci")


def test_sedimentation_conserves_mass(libs):
    """Closed form: per column, sum_k rst*q*dz changes by exactly the surface
    flux times dtp (upwqp.f90:164-202) when nothing is clipped at zero."""
    o, _ = libs
    c, d = state(8)
    rng = np.random.default_rng(5)
    q = np.asfortranarray(rng.uniform(1e-4, 2e-3, d["qr"].shape).astype(np.float64))
    o64 = po.Oracle(64)
    rbr = d["rbr"].astype(np.float64, order="F")
    rst = rbr.copy(order="F")
    u = np.asfortranarray(np.full(q.shape, 2.0))
    pr = np.zeros((NI + 2, NJ + 2, 2), np.float64, order="F")
    q0 = q.copy(order="F")
    dtp = 1.0
    o64.or_upwqp(c, po.C.c_double(dtp), rbr, rst, u, q, pr, po.field(c, np.float64))
    dz = c.dz
    col = lambda a: (rst[1:NI, 1:NJ, 0:NK - 2] * a[1:NI, 1:NJ, 0:NK - 2]).sum(axis=2) * dz
    # flux through the top of level nk-2 enters from level nk-1: account for it
    inflow = (rbr * u * q0)[1:NI, 1:NJ, NK - 2] * dtp
    outflow = (rbr * u * q0)[1:NI, 1:NJ, 0] * dtp
    np.testing.assert_allclose(col(q) - col(q0), inflow - outflow, rtol=1e-10, atol=1e-14)
    np.testing.assert_allclose(pr[1:NI, 1:NJ, 0], ((rbr * u * q0)[1:NI, 1:NJ, 0] +
                                                 (rbr * u * q0)[1:NI, 1:NJ, 1]) * 0.5 / 1000.0)


@pytest.mark.parametrize("seed", [4, 5])
def test_satadjst(libs, seed):
    o, r = libs
    c, d = state(seed)
    r.configure(c)
    F = lambda k=1: po.field(c, np.float32, k)
    q = np.asfortranarray(np.stack([d[k] for k in ("qv", "qc", "qr", "qi", "qs", "qg")], -1))
    qa, qb = q.copy(order="F"), q.copy(order="F")
    pa, pb = d["ptp"].copy(order="F"), d["ptp"].copy(order="F")
    na = d["nci"].copy(order="F")
    nb = F(3)
    nb[..., 0] = d["nci"]
    o.or_satadjst_level(c, d["pbr"], d["ptbr"], d["pp"], pa, qa, na, F(), F())
    r.ref_satadjst(NI, NJ, NK, d["pbr"], d["ptbr"], F(), d["pp"], pb, qb, F(2), nb, F(), F())
    same(pa, pb, "ptp")
    same(qa, qb, "q")
    same(na, nb[..., 0], "nci")
    # invariant: total water qv+qc+qi unchanged up to rounding
    tw = lambda x: x[..., 0].astype(np.float64) + x[..., 1] + x[..., 3]
    np.testing.assert_allclose(tw(qa), tw(q), rtol=2e-6, atol=1e-9)


def test_adjstnci(libs):
    """or_adjstnci (cloud-ice number clipped to the mean-mass bounds of qi)
    against the reference's s_adjstnci, whole arrays."""
    rng = np.random.default_rng(77)
    c = cfg()
    o, r = libs
    f = lambda a: np.asfortranarray(a.astype(np.float32))
    shp = po.field(c, np.float32).shape
    qi = f(np.where(rng.uniform(0, 1, shp) < 0.7, np.exp(rng.uniform(np.log(1e-9), np.log(2e-3), shp)), 0.0))
    nci = f(qi * np.exp(rng.uniform(np.log(1e8), np.log(1e15), shp)))
    a, b = nci.copy(order="F"), nci.copy(order="F")
    o.or_adjstnci(c, qi, a)
    r.ref_adjstnci(NI, NJ, NK, qi, b)
    np.testing.assert_array_equal(a, b)
    # both bounds are active somewhere in this draw
    assert (a < nci).any() and (a > nci).any()
