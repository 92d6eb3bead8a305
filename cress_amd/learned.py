"""Host driver of the learned microphysics: record training pairs from the
plain scheme, fit the normalisation and envelope, train the network on the
GPU, and score a run that uses it against the plain scheme.

The network and its kernels are the library's (cress_mlp_*, cress_ctx_set_mlp
in include/cress_hip.h); this module only sequences them.  Scores follow the
plain scheme's own diagnostics: the accumulated surface precipitation
(prwtr/price index 2, upwqp.f90:199-202) summed over the domain, and the
RMSE of each mixing ratio over the interior.
"""
import numpy as np

from cress_amd import cases, host

QS = ["qv", "qc", "qr", "qi", "qs", "qg"]


def interior(a, cfg):
    return a[2:cfg.ni - 1, 2:cfg.nj - 1, 1:cfg.nk - 2]


def accumulated_precip(s):
    """Domain sum of the accumulated precipitation (rain + cloud water +
    ice + snow + graupel), kg m-2 over the interior columns."""
    cfg = s.cfg
    pw, pi = s.download("prwtr"), s.download("price")
    tot = pw[2:cfg.ni - 1, 2:cfg.nj - 1, 1, :].sum() + pi[2:cfg.ni - 1, 2:cfg.nj - 1, 1, :].sum()
    return float(tot)


def record(mlp, case, nsteps, first=True):
    """nsteps plain steps of the case, recording every active cell's pair."""
    s = host.Solver(case.cfg)
    case.load(s)
    s.set_mlp(mlp, 1)
    s.step(nsteps, first=first)
    s.sync()
    s.close()
    return mlp.samples()


def run(case, nsteps, mlp=None, mode=0):
    """A solver stepped nsteps from the case's cold start (mode 2: with the
    network); returns the open solver."""
    s = host.Solver(case.cfg)
    case.load(s)
    if mlp is not None and mode:
        s.set_mlp(mlp, mode)
    s.step(nsteps, first=True)
    s.sync()
    return s


def score(s_net, s_plain):
    """Relative error of the accumulated precipitation and the RMSE of each
    mixing ratio (over the interior) of a run against the plain one."""
    cfg = s_plain.cfg
    p0, p1 = accumulated_precip(s_plain), accumulated_precip(s_net)
    out = {"precip_plain": p0, "precip": p1,
           "precip_rel_err": abs(p1 - p0) / p0 if p0 > 0 else float("nan")}
    for q in QS:
        a = interior(s_net.download(q, 1), cfg).astype(np.float64)
        b = interior(s_plain.download(q, 1), cfg).astype(np.float64)
        out["rmse_" + q] = float(np.sqrt(np.mean((a - b) ** 2)))
        out["scale_" + q] = float(np.sqrt(np.mean(b ** 2)))
    return out


def train(mlp, steps, batch=16384, lr=2e-3, seed=1, decay_at=(0.6, 0.85)):
    """Adam with a step decay of the learning rate (x0.3 at each fraction
    of decay_at); returns the per-step losses."""
    losses = []
    bounds = [0] + [int(f * steps) for f in decay_at] + [steps]
    for i in range(len(bounds) - 1):
        n = bounds[i + 1] - bounds[i]
        if n > 0:
            losses.append(mlp.train(n, batch, lr * (0.3 ** i), seed + i))
    return np.concatenate(losses)
