"""FP64 numpy restatement of cress_diag (the context's mass and energy
budgets) for the tests: the same sums over the interior cells of fields
downloaded from the context or taken from the oracle."""
import numpy as np

G, CP, RD, EPS, P0, RHOW = 9.8, 1004.0, 287.0, 0.622, 1.e5, 1.e3
NDIAG = 13


def heights(cfg, jcb):
    """z of the scalar levels over flat terrain: z_w(2) = 0,
    z_w(k+1) = z_w(k) + dz jcb(k), z_s(k) = z_w(k) + dz jcb(k)/2."""
    nk = cfg.nk
    z = np.zeros(nk + 1)
    zw = 0.0
    for k in range(2, nk):
        jk = float(jcb[0, 0, k - 1])
        z[k] = zw + 0.5 * cfg.dz * jk
        zw += cfg.dz * jk
    return z


def diag(cfg, base, get, level, prwtr, price):
    """base: dict of base fields (host layout); get(name, level) gives a
    prognostic (host layout, Fortran order)."""
    ni, nj, nk = cfg.ni, cfg.nj, cfg.nk
    I, J, K = slice(2, ni - 1), slice(2, nj - 1), slice(1, nk - 2)   # 2..n-2 (0-based k)
    f = lambda a: np.asarray(a, dtype=np.float64)
    rst, jcb = f(base["rst"])[I, J, K], f(base["jcb"])[I, J, K]
    pbr, ptbr = f(base["pbr"])[I, J, K], f(base["ptbr"])[I, J, K]
    q = {s: f(get(s, level))[I, J, K] for s in ("qv", "qc", "qr", "qi", "qs", "qg")}
    u, v, w = f(get("u", level)), f(get("v", level)), f(get("w", level))
    pp, ptp = f(get("pp", level))[I, J, K], f(get("ptp", level))[I, J, K]
    dv = cfg.dx * cfg.dy * cfg.dz
    pr = pbr + pp
    th = ptbr + ptp
    t = th * (pr / P0) ** (RD / CP)
    rhod = pr * EPS / ((EPS + q["qv"]) * RD * t)
    md = jcb * rhod * dv
    uc = 0.5 * (u[I, J, K] + u[3:ni, J, K])
    vc = 0.5 * (v[I, J, K] + v[I, 3:nj, K])
    wc = 0.5 * (w[I, J, K] + w[I, J, 2:nk - 1])
    z = heights(cfg, f(base["jcb"]))[2:nk - 1][None, None, :]
    out = np.zeros(NDIAG)
    out[0] = (rst * q["qv"] * dv).sum()
    out[1] = (rst * (q["qc"] + q["qr"]) * dv).sum()
    out[2] = (rst * (q["qi"] + q["qs"] + q["qg"]) * dv).sum()
    out[3] = (f(prwtr)[I, J, 1, :].sum() + f(price)[I, J, 1, :].sum()) * RHOW * cfg.dx * cfg.dy
    out[4] = md.sum()
    out[5] = (0.5 * md * (uc ** 2 + vc ** 2 + wc ** 2)).sum()
    out[6] = (md * (CP - RD) * t).sum()
    out[7] = (md * G * z).sum()
    out[8] = (rst * ptp * dv).sum()
    ww = f(get("w", level))[I, J, K]
    out[9], out[10] = ww.max(), ww.min()
    out[11], out[12] = ptp.max(), ptp.min()
    return out
