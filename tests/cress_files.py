"""Readers of the reference's restart and dump files, for the tests.

Restart files are Fortran sequential unformatted files: each record is a
4-byte length, the payload and the length again.  Their record sequence is
restated here from outres.f90:886-1163 for the options the GPU path runs
(sfcopt=0, advopt<=3, cphopt 0 or 2, no aerosol, tracer or TKE); each
entry names the arrays of one record.

Dump files (dmpfmt=2, dmplev=1, dmpmon=0) are direct-access files of
(ni-3)*(nj-3) floats per record: one record per level k=2..nk-2 of each
3-D variable (outdmp3d.f90), u, v, w averaged to the scalar points, and
one record per 2-D precipitation array (outdmp2d.f90), in the order of
outdmp.f90:385-700 and 1506-1590.
"""
import numpy as np


def read_records(path):
    """The payloads of a sequential unformatted file, as bytes."""
    raw = open(path, "rb").read()
    out, off = [], 0
    while off < len(raw):
        n = int(np.frombuffer(raw[off:off + 4], np.int32)[0])
        out.append(raw[off + 4:off + 4 + n])
        n2 = int(np.frombuffer(raw[off + 4 + n:off + 8 + n], np.int32)[0])
        assert n2 == n, "record markers differ"
        off += 8 + n
    return out


def restart_layout(cfg, moist):
    """[(name, shape)] of the records of a restart file, outres.f90:886-1163.
    A record of several arrays lists them as name:a+b."""
    ni, nj, nk = cfg.ni, cfg.nj, cfg.nk
    f3 = (ni + 2, nj + 2, nk)
    cx, cy = (nj, nk, 2), (ni, nk, 2)
    op = cfg.wbc >= 4 or cfg.ebc >= 4 or cfg.sbc >= 4 or cfg.nbc >= 4
    L = [("fmois", None)]                                           # :886
    L += [(n, f3) for n in ("ubr", "vbr", "pbr", "ptbr", "qvbr")]   # :888-901
    L += [(n, f3) for n in ("u", "v", "w", "pp", "ptp")]            # :903-918 advopt<=3
    L += [(n, f3) for n in ("up", "vp", "wp", "ppp", "ptpp")]       # :920-935
    if op:                                                          # :937-967
        for v in ("u", "v", "w", "p", "pt"):
            L += [(v + "cpx", cx), (v + "cpy", cy)]
    if moist:
        L += [("qv", f3), ("qvp", f3)]                              # :971-981
        if op:
            L += [("qvcpx", cx), ("qvcpy", cy)]                     # :985-992
    c = cfg.cphopt
    if moist and c >= 1:
        L += [("qwtr:qc+qr", f3 + (2,))]                            # :996-1000
        L += [("qwtrp:qc+qr", f3 + (2,))]                           # :1012-1015
        if op:
            L += [("qwcpx", cx + (2,)), ("qwcpy", cy + (2,))]       # :1027-1035
    if moist and c == 2:
        L += [("qice:qi+qs+qg", f3 + (3,))]                         # :1051-1055
        L += [("nice:nci", f3)]                                     # :1059-1066
        L += [("qicep:qi+qs+qg", f3 + (3,))]                        # :1077-1081
        L += [("nicep:nci", f3)]                                    # :1085-1091
        if op:                                                      # :1102-1118
            L += [("qicpx", cx + (3,)), ("qicpy", cy + (3,)), ("nicpx", cx), ("nicpy", cy)]
    if moist and c >= 1:                                            # :1132-1140
        L += [("prwtr1", (ni + 2, nj + 2, 2)), ("prwtr2", (ni + 2, nj + 2, 2))]
    if moist and c == 2:                                            # :1154-1163
        L += [("price1", (ni + 2, nj + 2, 3)), ("price2", (ni + 2, nj + 2, 3))]
    return L


def read_restart(path, cfg, moist):
    recs = read_records(path)
    lay = restart_layout(cfg, moist)
    assert len(recs) == len(lay), (len(recs), len(lay))
    out = {}
    for (name, shape), r in zip(lay, recs):
        if shape is None:
            out[name] = r.decode()
            continue
        a = np.frombuffer(r, np.float32)
        assert a.size == int(np.prod(shape)), (name, a.size, shape)
        out[name.split(":")[0]] = a.reshape(shape, order="F")
    return out


def dump_layout(cfg, moist, dmpvar):
    """[(name, kind)] of the variables of a dump: kind 'oxx', 'xox', 'xxo',
    'xxx' for 3-D variables (nk-3 records each), 'xx' for 2-D ones."""
    L = []
    d = dmpvar.ljust(15, "x")
    if d[0] == "o":
        L.append(("ubar", "oxx"))
    if d[0] in "o-":
        L.append(("u", "oxx"))
    if d[1] == "o":
        L.append(("vbar", "xox"))
    if d[1] in "o-":
        L.append(("v", "xox"))
    if d[2] == "o":
        L.append(("w", "xxo"))
    for pos, bar, prt, tot in ((3, "pbar", "pp", "p"), (4, "ptbar", "ptp", "pt")):
        if d[pos] == "o":
            L += [(bar, "xxx"), (prt, "xxx")]
        elif d[pos] == "-":
            L.append((tot, "xxx"))
    if moist:
        if d[5] == "o":
            L.append(("qvbar", "xxx"))
        if d[5] in "o-":
            L.append(("qv", "xxx"))
        if cfg.cphopt >= 1 and d[6] == "o":
            L += [("qc", "xxx"), ("qr", "xxx")]
        if cfg.cphopt >= 2 and d[6] == "o":
            L += [("qi", "xxx"), ("qs", "xxx"), ("qg", "xxx")]
        if cfg.cphopt >= 1 and d[13] == "o":       # flqcqi_opt=1 (temparam.f90:54)
            L += [("pcr", "xx"), ("pca", "xx"), ("prr", "xx"), ("pra", "xx")]
        if cfg.cphopt >= 2 and d[13] == "o":
            L += [("pir", "xx"), ("pia", "xx"), ("psr", "xx"), ("psa", "xx"), ("pgr", "xx"),
                  ("pga", "xx")]
    return L


def read_dump(path, cfg, moist, dmpvar):
    ni, nj, nk = cfg.ni, cfg.nj, cfg.nk
    a = np.fromfile(path, np.float32)
    per = (ni - 3) * (nj - 3)
    assert a.size % per == 0
    recs = a.reshape(-1, per)
    out, r = {}, 0
    for name, kind in dump_layout(cfg, moist, dmpvar):
        if kind == "xx":
            out[name] = recs[r].reshape((ni - 3, nj - 3), order="F")
            r += 1
        else:
            out[name] = np.stack([recs[r + m].reshape((ni - 3, nj - 3), order="F") for m in range(nk - 3)],
                                 axis=2)
            r += nk - 3
    assert r == recs.shape[0], (r, recs.shape[0])
    return out


def to_scalar_points(a, kind, cfg):
    """What a dump holds for a full array a(0:ni+1,0:nj+1,1:nk): levels
    k=2..nk-2 over i=2..ni-2, j=2..nj-2, averaged as outdmp3d.f90 does."""
    ni, nj, nk = cfg.ni, cfg.nj, cfg.nk
    a = np.asarray(a, np.float32)
    I, J, K = slice(2, ni - 1), slice(2, nj - 1), slice(1, nk - 2)
    h = np.float32(0.5)
    if kind == "oxx":
        return h * (a[I, J, K] + a[3:ni, J, K])
    if kind == "xox":
        return h * (a[I, J, K] + a[I, 3:nj, K])
    if kind == "xxo":
        return h * (a[I, J, K] + a[I, J, 2:nk - 1])
    return a[I, J, K]
