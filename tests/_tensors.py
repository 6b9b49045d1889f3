"""Seeded synthetic sparse tensors for the tests (unique coordinates)."""
import numpy as np


def uniform_tensor(dims, nnz, seed, low=0.0):
    """nnz unique uniform coordinates in dims; values in (low, 1]."""
    rng = np.random.default_rng(seed)
    dims = [int(d) for d in dims]
    total = 1
    for d in dims:
        total *= d
    assert nnz <= total
    rows = np.zeros((0, len(dims)), dtype=np.uint64)
    while rows.shape[0] < nnz:
        m = 2 * (nnz - rows.shape[0]) + 16
        draw = np.stack([rng.integers(0, d, size=m, dtype=np.uint64) for d in dims], axis=1)
        rows = np.unique(np.concatenate([rows, draw]), axis=0)
    rows = rows[rng.permutation(rows.shape[0])[:nnz]]
    cidx = np.ascontiguousarray(rows.T)
    vals = 1.0 - rng.random(nnz) * (1.0 - low)
    return cidx, vals


def zipf_tensor(dims, nnz, seed, s=1.1, values="lognormal"):
    """Unique coordinates with Zipf(s) indices over a seeded permutation of each mode."""
    rng = np.random.default_rng(seed)
    dims = [int(d) for d in dims]
    cdfs, perms = [], []
    for d in dims:
        p = np.arange(1, d + 1, dtype=np.float64) ** (-s)
        cdfs.append(np.cumsum(p / p.sum()))
        perms.append(rng.permutation(d).astype(np.uint64))
    rows = np.zeros((0, len(dims)), dtype=np.uint64)
    for _ in range(200):
        if rows.shape[0] >= nnz:
            break
        m = 2 * (nnz - rows.shape[0]) + 64
        draw = np.stack([perms[n][np.minimum(np.searchsorted(cdfs[n], rng.random(m)), dims[n] - 1)]
                         for n in range(len(dims))], axis=1)
        rows = np.unique(np.concatenate([rows, draw]), axis=0)
    rows = rows[rng.permutation(rows.shape[0])[:nnz]]
    cidx = np.ascontiguousarray(rows.T)
    k = rows.shape[0]
    if values == "lognormal":
        vals = rng.lognormal(0.0, 1.0, k)
    else:
        vals = 1.0 - rng.random(k)
    return cidx, vals


def mttkrp_numpy(cidx, vals, factors, mode):
    """Brute-force MTTKRP straight from the definition."""
    R = factors[0].shape[1]
    prod = np.repeat(vals[:, None], R, axis=1).astype(np.float64)
    for m, U in enumerate(factors):
        if m != mode:
            prod *= U[cidx[m].astype(np.int64)]
    out = np.zeros_like(factors[mode], dtype=np.float64)
    np.add.at(out, cidx[mode].astype(np.int64), prod)
    return out
