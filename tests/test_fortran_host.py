"""The Fortran 90 host (cress_amd/fortran/cress_run.f90): reads the
reference's user.conf namelist groups and a state file, drives the device
context through the ISO_C_BINDING module, and writes the result.

CPU: the options it resolves from user.conf equal the ones the Python host
builds for the same case (periodic flags 1 -> -1, nsstp from dtbig/dtsml),
and the interoperable cress_config it passes is accepted by the library's
validation (which runs before any device call).
GPU: its output after three large steps is bitwise the Python host's.
"""
import os
import subprocess

import numpy as np
import pytest

from cress_amd import build as hb
from cress_amd import cases, host

EXE = hb.FHOST


def _need_exe():
    if not os.path.exists(EXE):
        pytest.skip("Fortran host not built (no flang)")


def _files(tmp_path, case, nsteps):
    conf, st, out = (str(tmp_path / n) for n in ("user.conf", "state.bin", "out.bin"))
    case.write_state(st)
    case.write_userconf(conf, nsteps)
    return conf, st, out


@pytest.mark.parametrize("name", ["vortex", "supercell"])
def test_fortran_host_resolves_options(tmp_path, name):
    _need_exe()
    case = cases.build(name, 16, 14, 20, cphopt=2)
    conf, st, out = _files(tmp_path, case, 4)
    r = subprocess.run([EXE, conf, st, out, "--check"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    kv = {}
    for line in r.stdout.splitlines():
        parts = line.split()
        if parts:
            kv[parts[0]] = parts[1:]
    c = case.cfg
    assert [int(x) for x in kv["dims"]] == [c.ni, c.nj, c.nk]
    assert [int(x) for x in kv["bcs"]] == [c.wbc, c.ebc, c.sbc, c.nbc, c.bbc, c.tbc]
    assert [int(x) for x in kv["opts"]] == [c.advopt, c.impopt, c.gwmopt, c.buyopt, c.divopt,
                                            c.cphopt, c.haiopt, c.vspopt, c.coropt, c.sthopt]
    assert [int(x) for x in kv["ksp0_nsstp"]] == [c.ksp0[0], c.ksp0[1], c.nsstp]
    assert kv["vspvar"][0] == bytes(c.vspvar).decode()
    reals = [float(x) for x in kv["reals"]]
    want = [c.dx, c.dy, c.dz, c.dtbig, c.dtsml, c.weicoe, c.filcoe, c.thresq, c.gwave, c.smhcoe,
            c.smvcoe]
    # user.conf values pass through default REAL, as in the reference
    np.testing.assert_array_equal(np.float32(reals), np.float32(want))
    assert int(kv["nsteps"][0]) == 4
    assert int(kv["smtopt"][0]) == c.smtopt


def test_fortran_host_rejects_multirank_and_bad_state(tmp_path):
    _need_exe()
    case = cases.build("bubble", 12, 10, 12, cphopt=0)
    conf, st, out = _files(tmp_path, case, 1)
    txt = open(conf).read().replace("numpe=1", "numpe=4")
    open(conf, "w").write(txt)
    r = subprocess.run([EXE, conf, st, out, "--check"], capture_output=True, text=True, timeout=60)
    assert r.returncode != 0 and "numpe" in r.stdout
    conf, st, out = _files(tmp_path, case, 1)
    with open(st, "r+b") as f:
        f.write(b"XXXX")
    r = subprocess.run([EXE, conf, st, out, "--check"], capture_output=True, text=True, timeout=60)
    assert r.returncode != 0 and "not a state file" in r.stdout


def test_fortran_host_config_passes_validation(tmp_path):
    """Without a GPU the library answers cress_ctx_create with 'no usable
    GPU' only after its config validation accepted the struct the Fortran
    side filled (an invalid struct gives 'invalid argument')."""
    _need_exe()
    if host.device_count() > 0:
        pytest.skip("a GPU is present; covered by the GPU test")
    case = cases.build("bubble", 12, 10, 12, cphopt=2)
    conf, st, out = _files(tmp_path, case, 1)
    r = subprocess.run([EXE, conf, st, out], capture_output=True, text=True, timeout=60)
    assert r.returncode != 0
    assert "no usable GPU" in r.stdout, r.stdout


def _read_out(path):
    raw = open(path, "rb").read()
    hdr = np.frombuffer(raw[:32], np.int32)
    recs, off = {}, 32
    for _ in range(hdr[7]):
        fid, lv, n = np.frombuffer(raw[off:off + 12], np.int32)
        off += 12
        recs[int(fid)] = np.frombuffer(raw[off:off + 4 * n], np.float32)
        off += 4 * n
    return recs


@pytest.mark.gpu
@pytest.mark.parametrize("wbc", [-1, 7])
def test_fortran_host_matches_python_host(tmp_path, wbc):
    _need_exe()
    case = cases.build("bubble", 12, 10, 24, cphopt=2, wbc=wbc, gwave=30.0)
    conf, st, out = _files(tmp_path, case, 3)
    r = subprocess.run([EXE, conf, st, out], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    recs = _read_out(out)
    s = host.Solver(case.cfg)
    case.load(s)
    s.step(1, first=True)
    s.step(2)
    for name in ("u", "v", "w", "pp", "ptp", "qv", "qc", "qr", "qi", "qs", "qg", "nci"):
        np.testing.assert_array_equal(recs[host.FID[name]], s.download(name, 1).ravel(order="F"),
                                      err_msg=name)
    for name in ("prwtr", "price"):
        np.testing.assert_array_equal(recs[host.FID[name]], s.download(name).ravel(order="F"),
                                      err_msg=name)
    s.close()


def test_fortran_host_reads_output_groups(tmp_path):
    """&runame, &drname, &initype and &outfomat (defname.f90:802-871) are
    read; dumps other than dmpfmt=2/dmplev=1/dmpmon=0 are refused before
    anything runs."""
    _need_exe()
    case = cases.build("bubble", 12, 10, 12, cphopt=2)
    conf, st, out = (str(tmp_path / n) for n in ("user.conf", "state.bin", "out.bin"))
    case.write_state(st)
    groups = ["&runame exprim='run1' /", "&drname crsdir='%s/' /" % tmp_path,
              "&initype iniopt=2 /",
              "&outfomat dmpfmt=2, dmplev=1, dmpmon=0, dmpvar='ooooooo-xxxxxox', dmpitv=60.0, "
              "resopt=1, resitv=120.0 /"]
    case.write_userconf(conf, 4, stime=30.0, extra=groups)
    r = subprocess.run([EXE, conf, st, out, "--check"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    kv = {ln.split()[0]: ln.split()[1:] for ln in r.stdout.splitlines() if ln.split()}
    assert kv["outfomat"] == ["T", "2", "1", "0", "1", "ooooooo-xxxxxox"]
    assert [float(x) for x in kv["intervals"]] == [60.0, 120.0, 30.0]
    assert kv["files"] == ["2", str(tmp_path) + "/", "run1"]
    assert int(kv["nsteps"][0]) == 4
    for bad, msg in (("dmplev=2", "dmplev"), ("dmpfmt=1", "dmpfmt"), ("dmpmon=1", "dmpmon")):
        g = groups[:3] + ["&outfomat %s, dmpitv=60.0 /" % bad]
        case.write_userconf(conf, 4, extra=g)
        r = subprocess.run([EXE, conf, st, out, "--check"], capture_output=True, text=True, timeout=60)
        assert r.returncode != 0 and msg in r.stdout, r.stdout
