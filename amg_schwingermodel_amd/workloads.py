"""Seeded synthetic workloads (the reference only reads gauge and rhs files).

uneven_batch(): independent gauge configurations of mixed lattice sizes, each
with its own coupling and bare mass, one Gaussian right-hand side each,
grouped by lattice size so that every group can be one multi-configuration
context (amgs_create_multi) and the groups one amgs_solve_multi call.
"""
import numpy as np

from . import default_params, gauge_metropolis, gaussian

# 3 levels: aggregation 8x8 then 4x4 blocks, 10 test vectors, 4x4 Schwarz blocks
UNEVEN_ROWS = [(0, 8, 8, 10, 4, 4), (1, 4, 4, 10, 4, 4)]


def uneven_batch(nconf=48, seed=20260514, sizes=(128, 256, 512), beta=(1.0, 4.0), m0=(-0.22, -0.10), sweeps=500,
                 log=None):
    """Members i = 0..nconf-1 with lattice size, beta and m0 drawn from a
    seeded generator (size uniformly from `sizes`, beta and m0 uniformly
    from their ranges); the gauge field of member i comes from `sweeps`
    Metropolis sweeps from a cold start with seed seed + 1 + i, its
    right-hand side from gaussian(.., seed + 1000 + i).  Returns a list of
    groups, one per lattice size present: dict(n, members, beta, m0, U
    [k, n*n, 2], b [k, n*n*2]).  The generator itself runs on all host
    threads, so the members are generated one after another; log (a
    callable) receives a line per member."""
    rng = np.random.default_rng(seed)
    size = rng.choice(np.asarray(sizes), size=nconf)
    betas = rng.uniform(beta[0], beta[1], size=nconf)
    m0s = rng.uniform(m0[0], m0[1], size=nconf)
    fields = []
    for i in range(nconf):
        U, plaq = gauge_metropolis(int(size[i]), int(size[i]), float(betas[i]), sweeps, seed + 1 + i)
        fields.append(U)
        if log:
            log(f"uneven batch member {i}: {size[i]}^2, beta {betas[i]:.3f}, m0 {m0s[i]:.4f}, plaquette {plaq:.4f}")
    groups = []
    for n in sizes:
        idx = [i for i in range(nconf) if size[i] == n]
        if not idx:
            continue
        groups.append(dict(n=int(n), members=idx, beta=betas[idx], m0=m0s[idx],
                           U=np.stack([fields[i] for i in idx]),
                           b=np.stack([gaussian(int(n) * int(n) * 2, seed + 1000 + i) for i in idx])))
    return groups


def uneven_params(n, **kw):
    """Parameters of one size group of uneven_batch (masses per member go to
    amgs_create_multi)."""
    return default_params(n, n, UNEVEN_ROWS, **kw)
