"""Seeded inputs of the TKE-closure coefficient tests (1.5-order closure,
tubopt=2): a stably stratified base state with perturbations that give
both signs of N^2, cloudy and clear cells, and TKE from zero up."""
import numpy as np

IN = ["jcb", "jcb8u", "jcb8v", "jcb8w", "pbr", "ptbr", "rbr", "ppp", "ptpp", "qvp", "qallp", "up", "vp", "wp",
      "tkep"]
OUT = ["t11", "t22", "t33", "t12", "t13", "t23", "t31", "t32", "ssq", "nsq8w", "rkh", "rkv", "priv"]


def inputs(ni, nj, nk, seed):
    rng = np.random.default_rng(seed)
    sh = (ni + 2, nj + 2, nk)

    def f(lo, hi):
        return np.asfortranarray(rng.uniform(lo, hi, sh).astype(np.float32))

    z = np.arange(nk, dtype=np.float32)[None, None, :]
    d = {}
    for n in ("jcb", "jcb8u", "jcb8v", "jcb8w"):
        d[n] = f(0.9, 1.1)
    d["pbr"] = np.asfortranarray((1.0e5 * np.exp(-0.04 * z) + 0 * f(0, 1)).astype(np.float32))
    d["ptbr"] = np.asfortranarray((300.0 + 1.2 * z + f(-0.05, 0.05)).astype(np.float32))
    d["rbr"] = f(0.8, 1.2)
    d["ppp"] = f(-50.0, 50.0)
    d["ptpp"] = f(-1.5, 1.5)
    d["qvp"] = f(0.0, 0.015)
    q = f(-2.e-3, 2.e-3)
    d["qallp"] = np.asfortranarray(np.maximum(q, 0.0).astype(np.float32))
    for n in ("up", "vp", "wp"):
        d[n] = f(-8.0, 8.0)
    t = f(-0.5, 2.0)
    d["tkep"] = np.asfortranarray(np.maximum(t, 0.0).astype(np.float32))
    outs = {n: f(-1.0, 1.0) for n in OUT}
    return d, outs


TKE_IN = ["jcb", "jcb8u", "jcb8v", "jcb8w", "rst", "rstxu", "rstxv", "rstxwc", "tke", "tkep", "rbct", "priv",
          "nsq8w", "ssq"]
TKE_OUT = ["rkh", "rkv", "tkefrc", "rkv8s"]


def tke_inputs(ni, nj, nk, seed):
    """Inputs of s_forcetke at the magnitudes the closure gives them: the
    Prandtl-number factor priv in its range 1..3, eddy viscosities of tens
    of m^2/s, S^2 and N^2 of a sheared, partly unstable flow, TKE from zero
    up, and mass fluxes of 10 m/s winds.  The outputs start as noise so
    that the cells a routine leaves alone are checked too."""
    rng = np.random.default_rng(seed)
    sh = (ni + 2, nj + 2, nk)

    def f(lo, hi, shape=sh):
        return np.asfortranarray(rng.uniform(lo, hi, shape).astype(np.float32))

    d = {n: f(0.9, 1.1) for n in ("jcb", "jcb8u", "jcb8v", "jcb8w")}
    d["rst"] = f(0.8, 1.2)
    for n in ("rstxu", "rstxv", "rstxwc"):
        d[n] = f(-10.0, 10.0)
    for n in ("tke", "tkep"):
        d[n] = np.asfortranarray(np.maximum(f(-0.5, 2.0), 0.0).astype(np.float32))
    d["rbct"] = f(0.0, 0.01, (ni, nj, nk, 2))
    d["priv"] = f(1.2, 3.0)
    d["nsq8w"] = f(-1.e-4, 3.e-4)
    d["ssq"] = f(0.0, 2.e-4)
    outs = {"rkh": f(0.0, 40.0), "rkv": f(0.0, 40.0), "tkefrc": f(-1.0, 1.0), "rkv8s": f(-1.0, 1.0)}
    return d, outs
