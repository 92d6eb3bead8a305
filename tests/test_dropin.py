"""Drop-in boundary: the tier-2 per-routine entry points and the Fortran
modules that replace the reference's m_gaussel, m_bcycle and m_advs.

CPU: the replacement modules (cress_amd/fortran/dropin/) compile against
the reference's own module files (the ones oracle/_ref builds from the
reference sources), i.e. they fit the reference tree with the same module,
generic and dummy names.
GPU: cress_gaussel, cress_advs, cress_advuvw, cress_advp, cress_hevi,
cress_steps, cress_bcycle and cress_bc8u/v/w give bitwise the oracle's
s_gaussel, s_advs, s_advuvw, s_advp, s_hevi (one substep), s_steps,
s_bcycle and s_bc8u/v/w results, and cress_cloudphy the oracle's
s_cloudphy within 1e-5 (the microphysics bar) on random arrays (the oracle being pinned
bitwise against the reference's routines in test_oracle_ref.py and
test_oracle_obc.py).
"""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

from cress_amd import build as hb
from cress_amd import host
from oracle import pyoracle as po

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REFMOD = os.path.join(ROOT, "oracle", "_ref", "obj32")
DROPIN = os.path.join(ROOT, "cress_amd", "fortran", "dropin")
NI, NJ, NK = 13, 11, 14


@pytest.mark.skipif(not os.path.isdir(REFMOD) or not os.path.exists(hb.FLANG),
                    reason="reference modules or flang absent")
@pytest.mark.parametrize("name", ["gaussel", "bcycle", "advs", "advuvw", "advp", "hevi", "steps", "bc8u",
                                  "bc8v", "bc8w", "cloudphy", "turbdrv", "turbflx", "turbs"])
def test_dropin_module_compiles_in_reference_tree(name, tmp_path):
    moddir = os.path.join(hb.LIBDIR, "mod")
    if not os.path.exists(os.path.join(moddir, "m_cress_hip.mod")):
        pytest.skip("m_cress_hip not built")
    r = subprocess.run([hb.FLANG, "-c", "-ffp-contract=off", "-I", REFMOD, "-I", moddir, "-module-dir",
                        str(tmp_path), os.path.join(DROPIN, name + ".f90"), "-o",
                        str(tmp_path / (name + ".o"))], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr


def _rnd(rng, lo=-1.0, hi=1.0):
    return np.asfortranarray(rng.uniform(lo, hi, (NI + 2, NJ + 2, NK)).astype(np.float32))


def _ptr(a):
    assert a.flags["F_CONTIGUOUS"] and a.dtype == np.float32
    return a.ctypes.data_as(C.c_void_p)


@pytest.mark.gpu
def test_tier2_gaussel_matches_oracle():
    rng = np.random.default_rng(7)
    rr, tt, ff = _rnd(rng), _rnd(rng), _rnd(rng)
    ss = _rnd(rng, 3.0, 5.0)
    g = [x.copy(order="F") for x in (rr, ss, tt, ff)]
    o = [x.copy(order="F") for x in (rr, ss, tt, ff)]
    host.check(host.lib().cress_gaussel(2, NI - 2, 2, NJ - 2, 3, NK - 2, NI, NJ, NK, *[_ptr(a) for a in g]),
               "cress_gaussel")
    c = po.make_cfg(NI, NJ, NK, dx=700.0, dy=900.0, dz=350.0)
    po.Oracle(32).or_gaussel(c, 2, NI - 2, 2, NJ - 2, 3, NK - 2, *o)
    np.testing.assert_array_equal(g[3], o[3])
    np.testing.assert_array_equal(g[1], o[1])


@pytest.mark.gpu
@pytest.mark.parametrize("advopt", [1, 2, 3])
def test_tier2_advs_matches_oracle(advopt):
    rng = np.random.default_rng(20 + advopt)
    ux, vx, wx, s = (_rnd(rng) for _ in range(4))
    cfg = host.make_config(NI, NJ, NK, 700.0, 900.0, 350.0, 4.0, 1.0, advopt=advopt, cphopt=0, sthopt=0)
    g = _rnd(rng)
    o = [g.copy(order="F")] + [_rnd(rng) for _ in range(4)]
    host.check(host.lib().cress_advs(C.byref(cfg), _ptr(ux), _ptr(vx), _ptr(wx), _ptr(s), _ptr(g)),
               "cress_advs")
    c = po.make_cfg(NI, NJ, NK, dx=700.0, dy=900.0, dz=350.0, advopt=advopt)
    po.Oracle(32).or_advs(c, ux, vx, wx, s, *o)
    # the routine writes sfrc over the interior only
    np.testing.assert_array_equal(g[2:NI - 1, 2:NJ - 1, 1:NK - 2], o[0][2:NI - 1, 2:NJ - 1, 1:NK - 2])


@pytest.mark.gpu
def test_tier2_bcycle_matches_oracle():
    rng = np.random.default_rng(31)
    cfg = host.make_config(NI, NJ, NK, 700.0, 900.0, 350.0, 4.0, 1.0, cphopt=0)
    c = po.make_cfg(NI, NJ, NK, dx=700.0, dy=900.0, dz=350.0)
    ora = po.Oracle(32)
    for args in ((2, 1, NI - 2, NI - 1, 2, 1, NJ - 2, NJ - 1), (4, 0, NI - 3, NI + 1, 3, 0, NJ - 3, NJ)):
        a = _rnd(rng)
        b = a.copy(order="F")
        host.check(host.lib().cress_bcycle(C.byref(cfg), *args, NK, _ptr(a)), "cress_bcycle")
        ora.or_bcycle(c, *args, NK, b)
        np.testing.assert_array_equal(a, b)


@pytest.mark.gpu
@pytest.mark.parametrize("wbc", [-1, 7])
@pytest.mark.parametrize("advopt", [1, 2, 3])
def test_tier2_advuvw_matches_oracle(advopt, wbc):
    rng = np.random.default_rng(40 + advopt + (wbc > 0))
    rx, ry, rz, u, v, w = (_rnd(rng) for _ in range(6))
    cfg = host.make_config(NI, NJ, NK, 700.0, 900.0, 350.0, 4.0, 1.0, advopt=advopt, cphopt=0, sthopt=0,
                           wbc=wbc, gwave=30.0)
    g = [_rnd(rng) for _ in range(3)]
    o = [a.copy(order="F") for a in g]
    host.check(host.lib().cress_advuvw(C.byref(cfg), *[_ptr(a) for a in (rx, ry, rz, u, v, w)],
                                       *[_ptr(a) for a in g]), "cress_advuvw")
    c = po.make_cfg(NI, NJ, NK, dx=700.0, dy=900.0, dz=350.0, advopt=advopt, wbc=wbc, gwave=30.0)
    po.Oracle(32).or_advuvw(c, rx, ry, rz, u, v, w, *o, *[_rnd(rng) for _ in range(5)])
    for a, b, n in zip(g, o, "uvw"):
        np.testing.assert_array_equal(a, b, err_msg=n)


@pytest.mark.gpu
@pytest.mark.parametrize("opts", [(2, 3, 3, 2, 2, 3), (7, 7, 7, 7, 3, 2), (-1, -1, -1, -1, 2, 2)])
def test_tier2_bc8_matches_oracle(opts):
    rng = np.random.default_rng(50)
    cfg = host.make_config(NI, NJ, NK, 700.0, 900.0, 350.0, 4.0, 1.0, cphopt=0)
    c = po.make_cfg(NI, NJ, NK, dx=700.0, dy=900.0, dz=350.0)
    for x in (cfg, c):
        x.wbc, x.ebc, x.sbc, x.nbc, x.bbc, x.tbc = opts
    ora = po.Oracle(32)
    L = host.lib()
    for kind, kmax in ((0, NK), (0, NK - 3), (1, NK), (1, 2), (2, NK)):
        a = _rnd(rng)
        b = a.copy(order="F")
        if kind == 0:
            host.check(L.cress_bc8u(C.byref(cfg), kmax, _ptr(a)), "cress_bc8u")
        elif kind == 1:
            host.check(L.cress_bc8v(C.byref(cfg), kmax, _ptr(a)), "cress_bc8v")
        else:
            host.check(L.cress_bc8w(C.byref(cfg), _ptr(a)), "cress_bc8w")
        ora.or_bc8(c, kind, kmax, b)
        np.testing.assert_array_equal(a, b, err_msg="kind %d kmax %d" % (kind, kmax))


@pytest.mark.gpu
@pytest.mark.parametrize("wbc", [-1, 7])
@pytest.mark.parametrize("advopt", [1, 2, 3])
def test_tier2_advp_matches_oracle(advopt, wbc):
    rng = np.random.default_rng(60 + advopt + (wbc > 0))
    j8u, j8v, j8w = (_rnd(rng, 0.5, 1.5) for _ in range(3))
    u, v, wc, pp = (_rnd(rng) for _ in range(4))
    cfg = host.make_config(NI, NJ, NK, 700.0, 900.0, 350.0, 4.0, 1.0, advopt=advopt, cphopt=0, sthopt=0,
                           wbc=wbc, gwave=30.0)
    g = _rnd(rng)
    o = g.copy(order="F")
    host.check(host.lib().cress_advp(C.byref(cfg), *[_ptr(a) for a in (j8u, j8v, j8w, u, v, wc, pp, g)]),
               "cress_advp")
    c = po.make_cfg(NI, NJ, NK, dx=700.0, dy=900.0, dz=350.0, advopt=advopt, wbc=wbc, gwave=30.0)
    po.Oracle(32).or_advp(c, j8u, j8v, j8w, u, v, wc, pp, o, *[_rnd(rng) for _ in range(7)])
    np.testing.assert_array_equal(g, o)


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["bubble", "vortex"])
def test_tier2_hevi_matches_oracle(name):
    """One acoustic substep through cress_hevi equals the oracle's or_hevi
    (the substep of its small-step driver) bitwise over whole arrays, on a
    case's base state and initial fields with random forcings."""
    from cress_amd import cases
    case = cases.build(name, 12, 10, 16, cphopt=0)
    cfg = case.cfg
    rng = np.random.default_rng(70)
    names = ["jcb", "jcb8u", "jcb8v", "jcb8w", "rbr", "rst", "rst8u", "rst8v", "rst8w", "rcsq"]
    base = [np.asfortranarray(case.base[k].astype(np.float32)) for k in names]
    frc = [np.asfortranarray((1e-3 * rng.uniform(-1.0, 1.0, base[0].shape)).astype(np.float32))
           for _ in range(4)]
    st = [np.asfortranarray(case.init[k].astype(np.float32)) for k in ("u", "v", "w")]
    wc = np.asfortranarray((st[2] / base[3]).astype(np.float32))
    pp = np.asfortranarray(case.init["pp"].astype(np.float32))
    g = [a.copy(order="F") for a in (st[0], st[1], st[2], wc, pp)]
    o = [a.copy(order="F") for a in g]
    dts = float(np.float32(cfg.dtsml))
    host.check(host.lib().cress_hevi(C.byref(cfg), C.c_float(dts), *[_ptr(a) for a in base + frc + g]),
               "cress_hevi")
    m = po.Model(case, 32)
    tmp = [np.zeros_like(base[0], order="F") for _ in range(6)]
    m.o.lib.or_hevi(C.byref(m.cfg), m.o._real(dts), *[a.ctypes.data_as(C.c_void_p) for a in base + frc + o + tmp])
    for a, b, n in zip(g, o, ["uf", "vf", "wf", "wc", "ppf"]):
        np.testing.assert_array_equal(a, b, err_msg=n)
    assert not np.array_equal(g[0], st[0])


@pytest.mark.gpu
@pytest.mark.parametrize("cphopt", [0, 2])
def test_tier2_steps_matches_oracle(cphopt):
    """cress_steps (the scalar leapfrog, halo fills and vbcs of s_steps)
    equals the oracle's or_steps bitwise over whole arrays, with species
    forcings that drive some cells below zero (the max(0, .) clamp)."""
    from cress_amd import cases
    case = cases.build("bubble", 12, 10, 16, cphopt=cphopt)
    cfg = case.cfg
    rng = np.random.default_rng(80 + cphopt)
    shp = (NI, NJ, NK)
    shp = case.base["rst"].shape
    F = lambda a: np.asfortranarray(np.asarray(a, np.float32))
    R = lambda lo, hi, *extra: F(rng.uniform(lo, hi, shp + extra))
    rst = F(case.base["rst"])
    ptpp, qvp = F(case.init["ptp"]), F(case.init["qv"])
    ptfrc, qvfrc = R(-1e-3, 1e-3), R(-1e-6, 1e-6)
    qwp, qip, nip = R(0.0, 1e-3, 2), R(0.0, 1e-3, 3), R(0.0, 1e3)
    qwf, qif, nif = R(-1e-6, 1e-6, 2), R(-1e-6, 1e-6, 3), R(-1.0, 1.0)
    ptpf, qvf, dtd = R(-1.0, 1.0), R(0.0, 1e-2), R(0.0, 1.0)
    dtb = float(np.float32(cfg.dtbig))
    g = [a.copy(order="F") for a in (qwf, qif, nif, ptpf, qvf, dtd)]
    host.check(host.lib().cress_steps(C.byref(cfg), C.c_float(dtb), *[_ptr(a) for a in
                                      (rst, ptpp, ptfrc, qvp, qvfrc, qwp, qip, nip)],
                                      *[_ptr(a) for a in g]), "cress_steps")
    m = po.Model(case, 32)
    qp = F(np.concatenate([qvp[..., None], qwp, qip], axis=3))
    qf = F(np.concatenate([qvf[..., None], qwf, qif], axis=3))
    o_ptpf, o_nif, o_dtd = ptpf.copy(order="F"), nif.copy(order="F"), dtd.copy(order="F")
    vp = lambda a: a.ctypes.data_as(C.c_void_p)
    m.o.lib.or_steps_a(C.byref(m.cfg), m.o._real(dtb), vp(rst), vp(ptpp), vp(ptfrc), vp(qp), vp(qvfrc), vp(nip),
                       vp(o_ptpf), vp(qf), vp(o_nif), vp(o_dtd))
    np.testing.assert_array_equal(g[3], o_ptpf, err_msg="ptpf")
    np.testing.assert_array_equal(g[4], qf[..., 0], err_msg="qvf")
    np.testing.assert_array_equal(g[5], o_dtd, err_msg="dtdrst")
    if cphopt == 2:
        np.testing.assert_array_equal(g[0], qf[..., 1:3], err_msg="qwtrf")
        np.testing.assert_array_equal(g[1], qf[..., 3:6], err_msg="qicef")
        np.testing.assert_array_equal(g[2], o_nif, err_msg="nicef")
        assert (g[0] == 0.0).any() and (g[0] > 0.0).any()


@pytest.mark.gpu
def test_tier2_stepstke_matches_oracle():
    """cress_stepstke (the TKE lines of s_steps) equals the oracle's
    or_stepstke, pinned bitwise to the reference's s_steps with TKE,
    over the whole array, the clamp at zero acting on part of it."""
    from cress_amd import cases
    case = cases.build("bubble", 12, 10, 16, cphopt=0)
    cfg = case.cfg
    rng = np.random.default_rng(91)
    shp = case.base["rst"].shape
    F = lambda a: np.asfortranarray(np.asarray(a, np.float32))
    rst = F(case.base["rst"])
    tkep = F(np.maximum(rng.uniform(-0.5, 2.0, shp), 0.0))
    tkef = F(rng.uniform(-0.05, 0.05, shp))
    dtb = float(np.float32(cfg.dtbig))
    g = tkef.copy(order="F")
    print("cress_stepstke bubble %s" % (shp,), flush=True)
    host.check(host.lib().cress_stepstke(C.byref(cfg), C.c_float(dtb), _ptr(rst), _ptr(tkep), _ptr(g)),
               "cress_stepstke")
    m = po.Model(case, 32)
    o = tkef.copy(order="F")
    vp = lambda a: a.ctypes.data_as(C.c_void_p)
    m.o.lib.or_stepstke(C.byref(m.cfg), m.o._real(dtb), vp(rst), vp(tkep), vp(o))
    np.testing.assert_array_equal(g, o)
    assert (g == 0.0).any() and (g > 0.0).any()


@pytest.mark.gpu
def test_tier2_cloudphy_matches_oracle():
    """cress_cloudphy (the cold-rain microphysics on a given future level)
    against the oracle's or_cloudphy: every field within 1e-5 of the FP32
    oracle (max |difference| / max |value|), the bar of the full step's
    microphysics; and the scheme does change the state."""
    from cress_amd import cases
    case = cases.build("bubble", 12, 10, 24, cphopt=2)
    cfg = case.cfg
    rng = np.random.default_rng(90)
    F = lambda a: np.asfortranarray(np.asarray(a, np.float32))
    b = {k: F(case.base[k]) for k in ("jcb", "pbr", "ptbr", "rbr", "rst")}
    ini = case.init
    zeros = np.zeros(b["rst"].shape, np.float32)
    past = {k: F(ini[k]) if k in ini else F(zeros) for k in ("pp", "ptp", "qv", "qc", "qr", "qi", "qs", "qg", "nci")}
    fut = {k: F(past[k] * (1.0 + 0.01 * rng.uniform(-1.0, 1.0, zeros.shape))) for k in past if k != "pp"}
    qwp, qip = F(np.stack([past["qc"], past["qr"]], 3)), F(np.stack([past["qi"], past["qs"], past["qg"]], 3))
    g = {"ptpf": fut["ptp"].copy(order="F"), "qvf": fut["qv"].copy(order="F"),
         "qwtrf": F(np.stack([fut["qc"], fut["qr"]], 3)), "qicef": F(np.stack([fut["qi"], fut["qs"], fut["qg"]], 3)),
         "nicef": fut["nci"].copy(order="F"),
         "prwtr": np.zeros((cfg.ni + 2, cfg.nj + 2, 2, 2), np.float32, order="F"),
         "price": np.zeros((cfg.ni + 2, cfg.nj + 2, 2, 3), np.float32, order="F")}
    dtb = float(np.float32(cfg.dtbig))
    host.check(host.lib().cress_cloudphy(
        C.byref(cfg), C.c_float(dtb), *[_ptr(b[k]) for k in ("jcb", "pbr", "ptbr", "rbr", "rst")],
        _ptr(past["pp"]), _ptr(past["ptp"]), _ptr(past["qv"]), _ptr(qwp), _ptr(qip), _ptr(past["nci"]),
        *[_ptr(g[k]) for k in ("ptpf", "qvf", "qwtrf", "qicef", "nicef", "prwtr", "price")]), "cress_cloudphy")
    m = po.Model(case, 32)
    qp = F(np.stack([past[k] for k in ("qv", "qc", "qr", "qi", "qs", "qg")], 3))
    qf = F(np.stack([fut[k] for k in ("qv", "qc", "qr", "qi", "qs", "qg")], 3))
    optp, onci = fut["ptp"].copy(order="F"), fut["nci"].copy(order="F")
    opw = np.zeros_like(g["prwtr"], order="F")
    opi = np.zeros_like(g["price"], order="F")
    vp = lambda a: a.ctypes.data_as(C.c_void_p)
    rc = m.o.lib.or_cloudphy_a(C.byref(m.cfg), m.o._real(dtb), *[vp(b[k]) for k in ("jcb", "pbr", "ptbr", "rbr", "rst")],
                               vp(past["pp"]), vp(past["ptp"]), vp(qp), vp(past["nci"]), vp(optp), vp(qf), vp(onci),
                               vp(opw), vp(opi))
    assert rc == 0

    def rel(a, r):
        a, r = np.asarray(a, np.float64), np.asarray(r, np.float64)
        return np.abs(a - r).max() / max(np.abs(r).max(), 1e-30)

    pairs = [("ptpf", g["ptpf"], optp), ("qvf", g["qvf"], qf[..., 0]), ("qcf", g["qwtrf"][..., 0], qf[..., 1]),
             ("qrf", g["qwtrf"][..., 1], qf[..., 2]), ("qif", g["qicef"][..., 0], qf[..., 3]),
             ("qsf", g["qicef"][..., 1], qf[..., 4]), ("qgf", g["qicef"][..., 2], qf[..., 5]),
             ("ncif", g["nicef"], onci), ("prwtr", g["prwtr"], opw), ("price", g["price"], opi)]
    for name, a, r in pairs:
        assert rel(a, r) <= 1e-5, (name, rel(a, r))
    assert not np.array_equal(g["qwtrf"][..., 0], fut["qc"])
    assert np.abs(opw).max() > 0.0
