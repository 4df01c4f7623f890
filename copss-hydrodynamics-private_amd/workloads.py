"""Synthetic workloads of the GGEM particle path (seeded, reproducible).

Each workload is a parameter dict for copss_ggem.Context plus a generator
of positions, forces and Brownian noise.  Sizes follow the configurations
the path is measured on:

  beads100   100 beads, a = 1, [0,40)^3 periodic, 20^3 elements
  beads20k   20,000 beads, [0,200)^3 periodic, 64^3 elements
  slit25k    400 chains x 64 beads in a 160x160x100 slit (walls at z = 0, 100),
             80x80x50 elements
  beads1m    1,000,000 beads, [0,500)^3 periodic, 128^3 elements
  batch256   256 independent periodic suspensions, N_k uniform in [500, 8000],
             5% volume fraction, ~2a elements, 16 noise vectors each
  ggem3      the reference's GGEM validation case: 3 beads in a 30^3
             wall-bounded box, 15^3 elements, alpha = 0.35

Common: mu = 1/(6 pi), alpha = 0.5/a (cutoff 4/alpha = 8a), random
sequential addition at minimum distance 2.2a (std::mt19937_64(1000 + id)),
N(0,1) forces (std::mt19937_64(2000 + id)).  xi defaults to the reference's
bead value sqrt(pi)/(3a); xi_mode="json" uses 1/(a sqrt(pi)).
"""
import math

import numpy as np

import copss_ggem as G

MU = 1.0 / (6.0 * math.pi)


def _cfg(box_min, box_max, periodic, n_elem, alpha=0.5, xi_mode="reference", **kw):
    cfg = dict(mu=MU, alpha=alpha, box_min=list(box_min), box_max=list(box_max),
               periodic=list(periodic), n_elem=list(n_elem), point_type=G.POLYMER_BEAD,
               precision=G.FP64, compat_flags=G.COMPAT_ELEM_HYSTERESIS)
    if xi_mode == "json":
        cfg["ksi"] = 1.0 / math.sqrt(math.pi)
    cfg.update(kw)
    return cfg


WORKLOADS = {
    "beads100": dict(id=1, n=100, cfg=lambda **k: _cfg([0, 0, 0], [40, 40, 40], [1, 1, 1], [20, 20, 20], **k),
                     cheb_order=30, nqp_1d=3),
    "beads20k": dict(id=2, n=20000, cfg=lambda **k: _cfg([0, 0, 0], [200, 200, 200], [1, 1, 1], [64, 64, 64], **k),
                     cheb_order=30, nqp_1d=3),
    "slit25k": dict(id=3, n=25600, cfg=lambda **k: _cfg([0, 0, 0], [160, 160, 100], [1, 1, 0], [80, 80, 50], **k),
                    cheb_order=30, nqp_1d=3),
    "beads1m": dict(id=5, n=1000000, cfg=lambda **k: _cfg([0, 0, 0], [500, 500, 500], [1, 1, 1], [128, 128, 128], **k),
                    cheb_order=40, nqp_1d=3),
}


def make_sample(name, shrink=4, **kw):
    """(cfg, positions, forces) of a periodic workload shrunk by `shrink` per
    axis at the same bead density and element size: box / shrink,
    elements / shrink, beads / shrink^3 (random sequential addition at 2.2a
    and N(0,1) forces with the workload's seeds).  Every bead sees the same
    stencil, neighbour density and Gaussian support as in the full
    workload, so per-bead costs carry over by the bead count."""
    w = WORKLOADS[name]
    full = w["cfg"](**kw)
    assert all(full["periodic"]), "samples are taken of periodic workloads"
    cfg = dict(full, box_max=[b / shrink for b in full["box_max"]],
               n_elem=[int(ne) // shrink for ne in full["n_elem"]])
    n = w["n"] // shrink ** 3
    x = G.rsa_positions(cfg, n, 2.2, 1000 + w["id"])
    f = G.normal(3 * n, 2000 + w["id"]).reshape(n, 3)
    return cfg, x, f


def ggem3_cfg(**kw):
    return _cfg([-15, -15, -15], [15, 15, 15], [0, 0, 0], [15, 15, 15], alpha=0.35, **kw)


def make_inputs(name, n=None, **kw):
    """(cfg, positions [n,3], forces [n,3]) of a named workload."""
    w = WORKLOADS[name]
    cfg = w["cfg"](**kw)
    n = w["n"] if n is None else n
    if name == "slit25k":
        x = polymer_chains(cfg, 400, 64, seed=1000 + w["id"]) if n == w["n"] else \
            G.rsa_positions(cfg, n, 2.2, 1000 + w["id"])
    else:
        x = G.rsa_positions(cfg, n, 2.2, 1000 + w["id"])
    f = G.normal(3 * n, 2000 + w["id"]).reshape(n, 3)
    return cfg, x, f


def polymer_chains(cfg, n_chains, n_beads, seed, bond=0.97, excl=0.9, wall_margin=1.2):
    """Random-walk chains (bond length `bond`), self-avoiding at `excl`
    against all earlier beads, kept `wall_margin` away from the walls
    (outside the WCA range 2^(1/6) of the slit-wall force) by rejection;
    chain starts uniform in the box.  numpy Generator(PCG64(seed))."""
    rng = np.random.default_rng(seed)
    lo = np.array(cfg["box_min"], float)
    hi = np.array(cfg["box_max"], float)
    L = hi - lo
    per = np.array(cfg["periodic"], bool)
    cell = 2.0
    dims = np.maximum((L // cell).astype(int), 1)
    grid = {}
    out = np.zeros((n_chains * n_beads, 3))

    def key(p):
        c = np.minimum(((p - lo) / L * dims).astype(int), dims - 1)
        return tuple(c)

    def ok(p):
        if np.any(~per & ((p < lo + wall_margin) | (p > hi - wall_margin))):
            return False
        c = np.array(key(p))
        for dk in (-1, 0, 1):
            for dj in (-1, 0, 1):
                for di in (-1, 0, 1):
                    q = c + (di, dj, dk)
                    q = np.where(per, q % dims, q)
                    if np.any(q < 0) or np.any(q >= dims):
                        continue
                    for j in grid.get(tuple(q), ()):
                        d = out[j] - p
                        d = np.where(per & (d <= -L / 2), d + L, d)
                        d = np.where(per & (d > L / 2), d - L, d)
                        if d @ d < excl * excl:
                            return False
        return True

    i = 0
    for _c in range(n_chains):
        while True:
            p = lo + rng.random(3) * L
            if ok(p):
                break
        for b in range(n_beads):
            if b > 0:
                for _t in range(1000):
                    v = rng.normal(size=3)
                    q = out[i - 1] + bond * v / np.linalg.norm(v)
                    q = np.where(per, lo + np.mod(q - lo, L), q)
                    if ok(q):
                        p = q
                        break
                else:
                    raise RuntimeError("chain growth failed")
            out[i] = p
            grid.setdefault(key(p), []).append(i)
            i += 1
    return out


def batch_cfgs(nsys=256, seed=4004, n_lo=500, n_hi=8000, phi=0.05, **kw):
    """Parameter dicts and bead counts of a batch of periodic suspensions:
    N_k uniform in [n_lo, n_hi] (numpy Generator(PCG64(seed))), cube side
    L_k = (N_k (4 pi / 3) / phi)^(1/3) for volume fraction phi of unit
    beads, floor(L_k / 2) elements per axis."""
    rng = np.random.default_rng(seed)
    counts = rng.integers(n_lo, n_hi + 1, size=nsys)
    cfgs = []
    for nk in counts:
        L = (nk * (4.0 * math.pi / 3.0) / phi) ** (1.0 / 3.0)
        ne = int(L // 2.0)
        cfgs.append(_cfg([0, 0, 0], [L, L, L], [1, 1, 1], [ne, ne, ne], **kw))
    return cfgs, [int(c) for c in counts]


def make_batch(nsys=256, nrhs=16, seed=4004, **kw):
    """(cfgs, positions per system, forces per system, noise [nrhs, N, 3]):
    random sequential addition at 2.2a (std::mt19937_64(10000 + k)), N(0,1)
    forces and noise (std::mt19937_64(seed + 1000), (seed + 2000))."""
    cfgs, counts = batch_cfgs(nsys, seed, **kw)
    xs = [G.rsa_positions(c, nk, 2.2, 10000 + k) for k, (c, nk) in enumerate(zip(cfgs, counts))]
    ntot = sum(counts)
    f = G.normal(3 * ntot, seed + 1000).reshape(ntot, 3)
    off = np.concatenate([[0], np.cumsum(counts)])
    fs = [f[off[k]:off[k + 1]] for k in range(nsys)]
    dw = G.normal(nrhs * 3 * ntot, seed + 2000).reshape(nrhs, ntot, 3) if nrhs else None
    return cfgs, xs, fs, dw
