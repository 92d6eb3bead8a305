"""The restart file of the Fortran host (cress_amd/fortran/m_cress_files.f90)
against the reference's own writer and reader, on the CPU.

oracle/ref_harness_files.f90 runs, in oracle/_ref/libcressrefdrop32.so:
the host's write_restart and read_restart on a host-memory stand-in for
the device context (oracle/ctx_cpu.c), and the reference's s_outres
(outres.f90:102) and s_rdres (rdres.f90:100, as inivar.f90:725 calls it for
iniopt=12) on the same fields.  For each case:
  * the two files are the same bytes (record order, record lengths, every
    value, the name <exprim>res<ctime>.pe0000.bin);
  * the reference's reader, given the host's file, returns every field it
    was written from, and on open sides phase speeds equal to the pair the
    host writes for each field;
  * the host's reader, given the reference's file, returns every field.
Fields are seeded per case, each level and field different, so a swapped
record or level fails.  CPU only."""
import os

import numpy as np
import pytest

from cress_amd import cases
from oracle import pyoracle as po
from oracle import refmpi

pytestmark = pytest.mark.skipif(not (refmpi.have() and os.path.exists(refmpi.LIB_DROPIN)),
                                reason="oracle/_ref drop-in part not built")

FIELDS = ["u", "v", "w", "pp", "ptp", "qv", "qc", "qr", "qi", "qs", "qg", "nci"]
CTIME = 120.0


def _state(case, seed):
    c = case.cfg
    rng = np.random.default_rng(seed)
    sh = (c.ni + 2, c.nj + 2, c.nk)
    base = np.zeros(sh + (5,), np.float32, order="F")
    for n, k in enumerate(("ubr", "vbr", "pbr", "ptbr", "qvbr")):
        base[..., n] = case.base[k]
    lev = np.zeros(sh + (12, 2), np.float32, order="F")
    for n, f in enumerate(FIELDS):
        a = case.init[f].astype(np.float32) if f in case.init else np.zeros(sh, np.float32)
        scale = max(float(np.abs(a).max()), 1e-3)
        for l in range(2):
            lev[..., n, l] = a + (0.01 * scale * rng.standard_normal(sh)).astype(np.float32)
    pw = np.asfortranarray(rng.uniform(0.0, 1e-3, (c.ni + 2, c.nj + 2, 2, 2)).astype(np.float32))
    pi = np.asfortranarray(rng.uniform(0.0, 1e-3, (c.ni + 2, c.nj + 2, 2, 3)).astype(np.float32))
    return base, lev, pw, pi


def _expected_fields(cphopt, moist):
    """The level fields a restart holds (outres.f90:886-1163)."""
    n = 5 + (1 if moist else 0)
    if moist and cphopt >= 1:
        n = 8
    if moist and cphopt == 2:
        n = 12
    return list(range(n))


@pytest.mark.parametrize("name,kw,moist", [
    ("bubble", dict(cphopt=2), 1),
    ("supercell", dict(cphopt=2, wbc=7, gwave=30.0), 1),
    ("vortex", dict(cphopt=0), 1),
    ("bubble", dict(cphopt=0, wbc=7, gwave=20.0), 0),
])
def test_restart_matches_reference(tmp_path, name, kw, moist):
    case = cases.build(name, 12, 10, 14, **kw)
    sc = case.cfg
    cfg = po.cfg_from_solver_config(sc)
    base, lev, pw, pi = _state(case, 31 + len(name) + sc.cphopt + (sc.wbc > 0))
    opn = 1 if sc.wbc == 7 else 0
    dA, dB = str(tmp_path / "ours") + "/", str(tmp_path / "ref") + "/"
    os.makedirs(dA)
    os.makedirs(dB)
    fname = "pinres%08d.pe0000.bin" % int(CTIME)
    fA, fB = dA + fname, dB + fname
    f = float
    common = [sc.ni, sc.nj, sc.nk]
    phys = [f(sc.dx), f(sc.dy), f(sc.gwave), f(sc.dtbig), f(sc.dtsml)]
    z = lambda a: np.zeros_like(a, order="F")
    seq = [
        ("pin_write_ours", common + [moist, sc.cphopt, opn] + phys + [fA.encode(), len(fA), base, lev, pw, pi]),
        ("ref_outres_pin", common + [moist, sc.cphopt] + phys + [dB.encode(), len(dB), f(CTIME), base, lev, pw, pi]),
        ("ref_rdres_pin", common + [sc.cphopt, opn] + phys +
         [dA.encode(), len(dA), f(CTIME), np.zeros(1, np.int32), np.zeros(1, np.int32), z(base), z(lev), z(pw),
          z(pi)]),
        ("pin_read_ours", common + [sc.cphopt, opn] + phys + [fB.encode(), len(fB), np.zeros(1, np.int32), z(base),
                                                              z(lev), z(pw), z(pi)]),
    ]
    out = refmpi.calls(cfg, seq, lib=refmpi.LIB_DROPIN)
    a, b = open(fA, "rb").read(), open(fB, "rb").read()
    assert len(a) == len(b), (len(a), len(b))
    assert a == b
    keep = _expected_fields(sc.cphopt, moist)
    # the reference's reader on the host's file
    rd = out[2]
    moist_r, same = int(rd[-6][0]), int(rd[-5][0])
    assert moist_r == moist and same == 1
    np.testing.assert_array_equal(rd[-4], base)
    for n in keep:
        np.testing.assert_array_equal(rd[-3][..., n, :], lev[..., n, :], err_msg=FIELDS[n])
    if moist and sc.cphopt >= 1:
        np.testing.assert_array_equal(rd[-2], pw)
    if moist and sc.cphopt == 2:
        np.testing.assert_array_equal(rd[-1], pi)
    # the host's reader on the reference's file
    ro = out[3]
    assert int(ro[-5][0]) == moist
    np.testing.assert_array_equal(ro[-4], base)
    for n in keep:
        np.testing.assert_array_equal(ro[-3][..., n, :], lev[..., n, :], err_msg=FIELDS[n])
    if moist and sc.cphopt >= 1:
        np.testing.assert_array_equal(ro[-2], pw)
    if moist and sc.cphopt == 2:
        np.testing.assert_array_equal(ro[-1], pi)


@pytest.mark.parametrize("name,kw,moist,dmpvar", [
    ("bubble", dict(cphopt=2), 1, "oooooooxxxxxxoxx"),
    ("supercell", dict(cphopt=2, wbc=7, gwave=30.0), 1, "---------xxxxoxx"),
    ("vortex", dict(cphopt=0), 1, "oo-o-oxxxxxxxxxx"),
    ("bubble", dict(cphopt=0), 0, "ooooooxxxxxxxxxx"),
])
def test_dump_matches_reference(tmp_path, name, kw, moist, dmpvar):
    """The host's dump (write_dump: dmpfmt=2, dmplev=1, dmpmon=0) of one
    time is the same bytes as the reference's s_opendmp + s_outdmp +
    s_closedmp on the same fields: base and perturbation ('o') or their sum
    ('-'), u, v, w averaged to the scalar points, the species, and the
    surface precipitation (dmpvar(14:14))."""
    case = cases.build(name, 12, 10, 14, **kw)
    sc = case.cfg
    cfg = po.cfg_from_solver_config(sc)
    base, lev, pw, pi = _state(case, 57 + len(name) + sc.cphopt)
    lv = np.asfortranarray(lev[..., 1])
    dA, dB = str(tmp_path / "ours") + "/", str(tmp_path / "ref") + "/"
    os.makedirs(dA)
    os.makedirs(dB)
    fname = "pindmp%08d.pe0000.bin" % int(CTIME)
    fA, fB = dA + fname, dB + fname
    dv = dmpvar.encode()
    seq = [
        ("pin_dump_ours", [sc.ni, sc.nj, sc.nk, moist, sc.cphopt, dv, fA.encode(), len(fA), base, lv, pw, pi]),
        ("ref_dump_pin", [sc.ni, sc.nj, sc.nk, moist, sc.cphopt, float(sc.dz), dv, dB.encode(), len(dB),
                          float(CTIME), base, lv, pw, pi]),
    ]
    refmpi.calls(cfg, seq, lib=refmpi.LIB_DROPIN)
    a, b = open(fA, "rb").read(), open(fB, "rb").read()
    rec = (sc.ni - 3) * (sc.nj - 3) * 4
    assert len(b) > 0 and len(a) == len(b), (len(a), len(b), rec)
    if a != b:
        ra = np.frombuffer(a, np.float32).reshape(-1, rec // 4)
        rb = np.frombuffer(b, np.float32).reshape(-1, rec // 4)
        bad = [r for r in range(ra.shape[0]) if not np.array_equal(ra[r], rb[r])]
        raise AssertionError("records differ: %s of %d" % (bad[:10], ra.shape[0]))
