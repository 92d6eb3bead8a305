"""4th-order numerical smoothing (smtopt=2 horizontal, 3 horizontal and
vertical): pin the oracle's restatement (oracle/cress_smooth.c) against the
reference's own s_smoo4s, s_smoo4qv, s_smoo4p and s_smoo4uvw, bitwise over
whole forcing arrays, on periodic and open domains (the 4th-order part drops
next to open sides, like the advection)."""
import numpy as np
import pytest

from oracle import pyoracle as po

pytestmark = pytest.mark.skipif(not po.have_ref(), reason="oracle/_ref not built")

NI, NJ, NK = 13, 11, 14
SMH, SMV = 0.0008333, 0.0011


@pytest.fixture(scope="module")
def libs():
    return po.Oracle(32), po.Ref()


def fld(rng, lo=-1.0, hi=1.0):
    return np.asfortranarray(rng.uniform(lo, hi, (NI + 2, NJ + 2, NK)).astype(np.float32))


def cfg(smtopt, wbc):
    return po.make_cfg(NI, NJ, NK, dx=700.0, dy=900.0, dz=350.0, wbc=wbc, gwave=30.0, smtopt=smtopt,
                       smhcoe=SMH, smvcoe=SMV)


@pytest.mark.parametrize("wbc", [-1, 7])
@pytest.mark.parametrize("smtopt", [2, 3])
@pytest.mark.parametrize("kind", [0, 1, 2])
def test_smoo4_scalar(libs, kind, smtopt, wbc):
    o, r = libs
    rng = np.random.default_rng(100 * kind + 10 * smtopt + (wbc > 0))
    c = cfg(smtopt, wbc)
    r.configure(c)
    s, rbr, ref = fld(rng), fld(rng, 0.5, 1.5), fld(rng)
    f = fld(rng)
    g = f.copy(order="F")
    o.or_smoo4s(c, kind, rbr, ref, s, f, *[fld(rng) for _ in range(4)])
    dm = [fld(rng) for _ in range(6)]
    r.ref_smoo4(kind, smtopt, SMH, SMV, NI, NJ, NK, s, dm[0], dm[1], g, dm[2], dm[3], rbr, ref,
                *[fld(rng) for _ in range(6)], *[fld(rng) for _ in range(5)])
    np.testing.assert_array_equal(f, g)
    assert not np.array_equal(f, fld(np.random.default_rng(100 * kind + 10 * smtopt + (wbc > 0))))


@pytest.mark.parametrize("wbc", [-1, 7])
@pytest.mark.parametrize("smtopt", [2, 3])
def test_smoo4uvw(libs, smtopt, wbc):
    o, r = libs
    rng = np.random.default_rng(7 + smtopt + (wbc > 0))
    c = cfg(smtopt, wbc)
    r.configure(c)
    jcb8 = [fld(rng, 0.5, 1.5) for _ in range(3)]
    ubr, vbr = fld(rng), fld(rng)
    rst8 = [fld(rng, 0.5, 1.5) for _ in range(3)]
    u, v, w = fld(rng), fld(rng), fld(rng)
    fo = [fld(rng) for _ in range(3)]
    fr = [x.copy(order="F") for x in fo]
    o.or_smoo4uvw(c, *jcb8, ubr, vbr, *rst8, u, v, w, *fo, *[fld(rng) for _ in range(4)])
    r.ref_smoo4(3, smtopt, SMH, SMV, NI, NJ, NK, u, v, w, *fr, *jcb8, ubr, vbr, *rst8,
                *[fld(rng) for _ in range(5)])
    for a, b in zip(fo, fr):
        np.testing.assert_array_equal(a, b)


def test_smoothing_damps_the_two_grid_wave():
    """Independent of the reference: on a periodic domain, the 4th-order
    horizontal smoothing of a checkerboard in x is -16*smhcoe*a (each
    second difference of (-1)^i is -4 times the field), and it leaves a
    field linear in x untouched (all second differences vanish)."""
    o = po.Oracle(64)
    c = cfg(2, -1)
    i = np.arange(NI + 2)[:, None, None]
    s = np.asfortranarray(np.broadcast_to((-1.0) ** i, (NI + 2, NJ + 2, NK)).astype(np.float64))
    rbr = np.asfortranarray(np.ones_like(s))
    f = np.asfortranarray(np.zeros_like(s))
    scr = [np.asfortranarray(np.zeros_like(s)) for _ in range(4)]
    o.or_smoo4s(c, 0, rbr, None, s, f, *scr)
    inner = (slice(2, NI - 1), slice(2, NJ - 1), slice(1, NK - 2))
    np.testing.assert_allclose(f[inner], -16.0 * SMH * s[inner], rtol=1e-12)
    lin = np.asfortranarray(np.broadcast_to(1.0 + 0.25 * i, s.shape).astype(np.float64))
    f[...] = 0.0
    o.or_smoo4s(c, 0, rbr, None, lin, f, *scr)
    assert np.abs(f[inner]).max() <= 1e-14
