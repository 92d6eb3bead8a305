"""The base-state setup in binary64 on the GPU (cress_set_base): its
per-level columns against the FP64 restatement of the same chain
(tests/setup_ref.py: s_stretch, s_jacobian, s_intrpsnd, s_setbase/s_bcbase,
s_baserho, s_var8uvw + bc8w, s_sndwave, s_vspdmp) to 1e-12; every base
field, the sponge and the Coriolis array equal to those columns rounded once
to binary32 at every cell; and a large step on that base state equal to the
oracle's on the same fields (bitwise dynamics).  The restatement itself is
pinned against the reference's FP32 routines in test_oracle_ref_setup.py."""
import numpy as np
import pytest

from cress_amd import cases, host
from oracle import pyoracle as po
import setup_ref

pytestmark = pytest.mark.gpu

FIELD_COL = {"jcb": "jcb", "jcb8u": "jcb", "jcb8v": "jcb", "jcb8w": "jcb8w", "ubr": "ubr", "vbr": "vbr",
             "pbr": "pbr", "ptbr": "ptbr", "qvbr": "qvbr", "rbr": "rbr", "rst": "rst", "rst8u": "rst",
             "rst8v": "rst", "rst8w": "rst8w", "rcsq": "rcsq"}


@pytest.mark.parametrize("name,dims", [("bubble", (16, 14, 30)), ("supercell", (20, 18, 24)),
                                       ("vortex", (16, 14, 24))])
def test_set_base_matches_the_fp64_chain(name, dims):
    case = cases.build(name, *dims, cphopt=0)
    cfg = case.cfg
    snd = case.sounding()
    s = host.Solver(cfg)
    s.set_base(snd)
    got = s.base_columns()
    ref = setup_ref.columns(cfg.nk, cfg.dz, cfg.sthopt, cfg.bbc, cfg.tbc, snd)
    for k, r in ref.items():
        g = got[k]
        err = np.abs(g[1:] - r[1:]).max() / max(np.abs(r[1:]).max(), 1e-300)
        assert err <= 1e-12, (k, err)
    shape = (cfg.ni + 2, cfg.nj + 2, cfg.nk)
    for f, c in FIELD_COL.items():
        want = np.broadcast_to(got[c][1:].astype(np.float32)[None, None, :], shape)
        np.testing.assert_array_equal(s.download(f), want, err_msg=f)
    rb = s.download("rbct").reshape((cfg.ni, cfg.nj, cfg.nk, 2), order="F")
    np.testing.assert_array_equal(rb[..., 1], np.broadcast_to(got["rbct"][1:].astype(np.float32), rb[..., 1].shape))
    assert not rb[..., 0].any()
    fc = s.download("fc")
    assert (fc == np.float32(0.25 * snd["fc"])).all()
    # a large step on this base state: the dynamics bitwise the oracle's on
    # the same fields
    base = {f: s.download(f) for f in FIELD_COL}
    case.base, case.rbct, case.fc = base, rb.copy(order="F"), fc.reshape((cfg.ni + 2, cfg.nj + 2, 2), order="F")
    for k, a in case.init.items():
        for lv in (0, 1):
            s.upload(k, a, lv)
    s.step(2, first=True)
    m = po.Model(case, 32)
    m.step(2, first=True)
    for f in ("u", "v", "w", "pp", "ptp", "qv"):
        a = s.download(f, 1)[2:cfg.ni - 1, 2:cfg.nj - 1, 1:cfg.nk - 2]
        b = m.get(f, 1)[2:cfg.ni - 1, 2:cfg.nj - 1, 1:cfg.nk - 2]
        np.testing.assert_array_equal(a, b, err_msg=f)
    s.close()


def test_set_base_rejects_a_non_increasing_sounding():
    case = cases.build("bubble", 12, 10, 12, cphopt=0)
    snd = case.sounding()
    snd["z"] = snd["z"].copy()
    snd["z"][5] = snd["z"][4]
    s = host.Solver(case.cfg)
    with pytest.raises(host.CressError):
        s.set_base(snd)
    s.close()
