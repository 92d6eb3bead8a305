"""Open lateral boundaries (wbc=ebc=sbc=nbc=7: radiation at a constant phase
speed): pin the oracle's restatement (oracle/cress_obc.c) against the
reference's own compiled phvuvw, phvs, rbcu, rbcv, rbcw, rbcss, rbcpt,
rbcqv, rbcq, rbcs, bc4news and lbcu/lbcv/lbcw/lbcs (oracle/_ref), bitwise
over whole arrays, on seeded random inputs.

The phase speeds fed to the radiation routines are random per entry (not
the constant option 7 produces), so that a wrong side, a transposed (j,k)
or a wrong boundary row in the restatement changes the result.
"""
import numpy as np
import pytest

from oracle import pyoracle as po

pytestmark = pytest.mark.skipif(not po.have_ref(), reason="oracle/_ref not built")

NI, NJ, NK = 13, 11, 14


@pytest.fixture(scope="module")
def libs():
    return po.Oracle(32), po.Ref()


def ocfg(gwave=37.5, advopt=3):
    return po.make_cfg(NI, NJ, NK, dx=700.0, dy=900.0, dz=350.0, advopt=advopt, wbc=7,
                       gwave=gwave)


def rnd(shape, rng, lo=-1.0, hi=1.0):
    return np.asfortranarray(rng.uniform(lo, hi, shape).astype(np.float32))


def fld(rng, lo=-1.0, hi=1.0):
    return rnd((NI + 2, NJ + 2, NK), rng, lo, hi)


def cp_pair(rng, lo=-0.9, hi=0.9):
    return rnd((NJ, NK, 2), rng, lo, hi), rnd((NI, NK, 2), rng, lo, hi)


# gwave*dtb/dx above 1 exercises the clip to [-1, 1]
@pytest.mark.parametrize("gwave,dtb,dts", [(37.5, 1.5, 0.5), (37.5, 3.0, 1.0), (400.0, 3.0, 1.0)])
def test_phvuvw(libs, gwave, dtb, dts):
    o, r = libs
    rng = np.random.default_rng(int(gwave) + int(10 * dtb))
    c = ocfg(gwave)
    r.configure(c)
    a = [x for _ in range(3) for x in cp_pair(rng, 5.0, 6.0)]   # unset entries keep these
    b = [x.copy(order="F") for x in a]
    rmf = np.asfortranarray(np.ones((NI + 2, NJ + 2, 4), np.float32))
    rmf8 = np.asfortranarray(np.ones((NI + 2, NJ + 2, 3), np.float32))
    u = fld(rng)
    scr = [np.zeros(n, np.float32) for n in ((NJ + 2) * NK, (NI + 2) * NK, NJ + 2, NI + 2)]
    o.or_phvuvw(c, dtb, dts, *a)
    r.ref_phvuvw(float(gwave), float(dtb), float(dts), NI, NJ, NK, rmf, rmf8, rmf8,
                 u, *b, *scr)
    for x, y in zip(a, b):
        np.testing.assert_array_equal(x, y)
    assert (a[0][:NJ - 1, 1:NK - 2, 0] < 0).all() and (a[0][:NJ - 1, 1:NK - 2, 1] > 0).all()


@pytest.mark.parametrize("sml", [0, 1])
@pytest.mark.parametrize("gwave", [37.5, 400.0])
def test_phvs(libs, sml, gwave):
    o, r = libs
    rng = np.random.default_rng(7 + sml)
    c = ocfg(gwave)
    r.configure(c)
    a = list(cp_pair(rng, 5.0, 6.0))
    b = [x.copy(order="F") for x in a]
    rmf = np.asfortranarray(np.ones((NI + 2, NJ + 2, 4), np.float32))
    u = fld(rng)
    o.or_phvs(c, sml, 3.0, 1.0, *a)
    r.ref_phvs(float(gwave), sml, 3.0, 1.0, NI, NJ, NK, rmf, u, *b,
               np.zeros(NJ + 2, np.float32), np.zeros(NI + 2, np.float32))
    for x, y in zip(a, b):
        np.testing.assert_array_equal(x, y)


@pytest.mark.parametrize("kind,name", [(1, "or_rbcu"), (2, "or_rbcv"), (3, "or_rbcw"),
                                       (4, "or_rbcss")])
def test_rbc_inplace(libs, kind, name):
    o, r = libs
    rng = np.random.default_rng(30 + kind)
    c = ocfg()
    r.configure(c)
    cx, cy = cp_pair(rng)
    a = fld(rng)
    b = a.copy(order="F")
    getattr(o, name)(c, cx, cy, a)
    r.ref_rbcuvw(kind, 1.0, NI, NJ, NK, cx, cy, fld(rng), b)
    np.testing.assert_array_equal(a, b)
    assert not np.array_equal(a, fld(np.random.default_rng(30 + kind)))


@pytest.mark.parametrize("kind", [1, 2, 3, 4])
def test_rbc_leapfrog(libs, kind):
    """rbcpt (1), rbcqv (2), rbcq (3), rbcs (4): sf on the boundary from the
    past value there and the present value inside; qv and q clip at 0."""
    o, r = libs
    rng = np.random.default_rng(40 + kind)
    c = ocfg()
    r.configure(c)
    cx, cy = cp_pair(rng)
    s, sp = fld(rng), fld(rng)
    a = fld(rng)
    b = a.copy(order="F")
    o.or_rbclf(c, 1 if kind in (2, 3) else 0, cx, cy, s, sp, a)
    r.ref_rbclf(kind, 3.0, NI, NJ, NK, s, sp, cx, cy, fld(rng), b)
    np.testing.assert_array_equal(a, b)
    if kind in (2, 3):
        assert (a[1, 2:NJ - 1, 1:NK - 2] >= 0).all()


@pytest.mark.parametrize("box", [(1, NI, 1, NJ - 1), (1, NI - 1, 1, NJ), (1, NI - 1, 1, NJ - 1),
                                 (0, NI + 1, 0, NJ), (0, NI, 0, NJ + 1), (0, NI, 0, NJ)])
def test_bc4news(libs, box):
    o, r = libs
    rng = np.random.default_rng(sum(box))
    c = ocfg()
    r.configure(c)
    a = fld(rng)
    b = a.copy(order="F")
    o.or_bc4news(c, *box, NK, a)
    r.ref_bc4news(*box, NI, NJ, NK, b)
    np.testing.assert_array_equal(a, b)
    isw, ise, jss, jsn = box
    np.testing.assert_array_equal(a[isw, jss], np.float32(0.5) * (a[isw + 1, jss] + a[isw, jss + 1]))


@pytest.mark.parametrize("advopt", [1, 2, 3])
@pytest.mark.parametrize("kind", [1, 2, 3, 4])
def test_lbc_is_a_noop_for_option7(libs, kind, advopt):
    """lbcu/lbcv/lbcw/lbcs act only for wbc 2 and 3 when advopt<=3
    (lbcs.f90:153-190): the oracle omits them on open sides."""
    _, r = libs
    r.configure(ocfg(advopt=advopt))
    a = fld(np.random.default_rng(kind))
    b = a.copy(order="F")
    r.ref_lbc(kind, NI, NJ, NK, b)
    np.testing.assert_array_equal(a, b)


def test_periodic_domain_has_no_radiation(libs):
    o, _ = libs
    c = po.make_cfg(NI, NJ, NK, dx=700.0, dy=900.0, dz=350.0, gwave=37.5)
    rng = np.random.default_rng(3)
    cx, cy = cp_pair(rng)
    a = fld(rng)
    b = a.copy(order="F")
    o.or_rbcu(c, cx, cy, a)
    o.or_rbclf(c, 1, cx, cy, fld(rng), fld(rng), a)
    o.or_bc4news(c, 1, NI, 1, NJ - 1, NK, a)
    np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("opts", [(2, 3, 3, 2, 2, 3), (7, 7, 7, 7, 3, 2), (-1, -1, -1, -1, 2, 2),
                                  (3, 2, 2, 3, 3, 3)])
def test_bc8(libs, opts):
    """s_bc8u/v/w (the staggered-point boundary fills): the oracle's or_bc8
    bitwise equal to the reference's routines; and what the options fix:
    a rigid side (2) copies the point two in, options >= 3 the point one in,
    periodic sides (-1) are untouched."""
    o, r = libs
    c = po.make_cfg(NI, NJ, NK, dx=700.0, dy=900.0, dz=350.0)
    c.wbc, c.ebc, c.sbc, c.nbc, c.bbc, c.tbc = opts
    r.configure(c)
    rng = np.random.default_rng(sum(abs(x) for x in opts))
    for kind, kmax in ((0, NK), (0, NK - 3), (1, NK), (1, 2), (2, NK)):
        a = np.asfortranarray(rng.uniform(-1.0, 1.0, (NI + 2, NJ + 2, NK)).astype(np.float32))
        b = a.copy(order="F")
        a0 = a.copy(order="F")
        o.or_bc8(c, kind, kmax, a)
        r.ref_bc8(kind, NI, NJ, kmax, b)
        np.testing.assert_array_equal(a, b, err_msg="kind %d kmax %d" % (kind, kmax))
        lo, hi = (opts[0], opts[1]) if kind == 0 else (opts[2], opts[3]) if kind == 1 else (opts[4], opts[5])
        ax = kind
        n = (NI, NJ, NK)[ax]
        sl = lambda idx: tuple(idx if d == ax else slice(None) for d in range(3))
        kk = slice(0, kmax) if kind < 2 else slice(None)
        def plane(arr, idx):
            p = arr[sl(idx)]
            return p[..., kk] if kind < 2 else p
        # Fortran index m on axis ax is array index m (x, y) or m-1 (z)
        off = 0 if kind < 2 else -1
        if lo == 2:
            np.testing.assert_array_equal(plane(a, 1 + off), plane(a0, 3 + off))
        elif lo >= 3:
            np.testing.assert_array_equal(plane(a, 1 + off), plane(a0, 2 + off))
        else:
            np.testing.assert_array_equal(plane(a, 1 + off), plane(a0, 1 + off))
        if hi == 2:
            np.testing.assert_array_equal(plane(a, n + off), plane(a0, n - 2 + off))
        elif hi >= 3:
            np.testing.assert_array_equal(plane(a, n + off), plane(a0, n - 1 + off))
        # nothing else changes
        mask = np.ones(a.shape, bool)
        mask[sl(1 + off)] = False
        mask[sl(n + off)] = False
        np.testing.assert_array_equal(a[mask], a0[mask])
