"""Seeded synthetic sparse tensors generated on the GPU (torch as plumbing).

Zipf(s) indices over a seeded random permutation of each mode's range
(heavy-hitter rows at random positions), duplicates rejected until the target
number of unique coordinates is reached (sparse_coo) or kept (sparse_draws),
values lognormal(0,1) or uniform (0,1].  Used by bench.py for the large workloads; the tests use the
numpy generators in tests/_tensors.py.
"""
from __future__ import annotations

import math

import torch


def _zipf_icdf_table(dim: int, s: float, device) -> torch.Tensor:
    k = torch.arange(1, dim + 1, dtype=torch.float64, device=device)
    p = k.pow(-s)
    c = torch.cumsum(p, 0)
    return c / c[-1]


def _draw_mode(cdf, perm, m, gen):
    u = torch.rand(m, dtype=torch.float64, device=cdf.device, generator=gen)
    r = torch.searchsorted(cdf, u).clamp_(max=cdf.numel() - 1)
    return perm[r]


def sparse_draws(dims, nnz, seed, s=1.1, values="lognormal", device="cuda"):
    """nnz Zipf(s) draws per mode over a seeded permutation of each mode's
    range, duplicate coordinates kept (in COO they add up, as a nonzero
    listed twice does in the reference): (coords, vals).  One pass of
    sampling, no deduplication -- the generator of the serving stream's
    jobs, which are made at submission time."""
    dims = [int(d) for d in dims]
    gen = torch.Generator(device=device)
    gen.manual_seed(int(seed))
    coords = []
    for d in dims:
        perm = torch.randperm(d, generator=gen, device=device)
        coords.append(_draw_mode(_zipf_icdf_table(d, s, device), perm, nnz, gen).contiguous())
    if values == "lognormal":
        vals = torch.exp(torch.randn(nnz, dtype=torch.float64, device=device, generator=gen))
    else:
        vals = 1.0 - torch.rand(nnz, dtype=torch.float64, device=device, generator=gen)
    return coords, vals


def sparse_coo(dims, nnz, seed, dist="zipf", s=1.1, values="lognormal", device="cuda"):
    """Returns (coords: list of int64 device tensors, vals: float64 device tensor)."""
    dims = [int(d) for d in dims]
    N = len(dims)
    gen = torch.Generator(device=device)
    gen.manual_seed(int(seed))
    cdfs, perms = [], []
    for d in dims:
        if dist == "zipf":
            cdfs.append(_zipf_icdf_table(d, s, device))
        else:
            cdfs.append(None)
        perms.append(torch.randperm(d, generator=gen, device=device))
    # linear key when the index space fits 63 bits, else a (hi, lo) pair
    total_bits = sum(max((d - 1).bit_length(), 1) for d in dims)
    split = N if total_bits <= 63 else None
    if split is None:
        acc = 0
        split = 0
        for n, d in enumerate(dims):
            b = max((d - 1).bit_length(), 1)
            if acc + b > 62:
                break
            acc += b
            split = n + 1

    def key_of(cols, lo, hi):
        k = torch.zeros_like(cols[lo])
        for n in range(lo, hi):
            k = k * dims[n] + cols[n]
        return k

    uniq = None
    rounds = 0
    yield_rate = 1.0  # new unique coordinates per draw in the last round
    while True:
        have = 0 if uniq is None else uniq.shape[0]
        if have >= nnz or rounds > 200:
            break
        # skewed indices collide often: size the round by the observed yield
        m = int(min((nnz - have) * 1.3 / max(yield_rate, 0.02), 2 * nnz)) + 1024
        cols = []
        for n, d in enumerate(dims):
            if dist == "zipf":
                cols.append(_draw_mode(cdfs[n], perms[n], m, gen))
            else:
                cols.append(torch.randint(0, d, (m,), generator=gen, device=device))
        if split == N:
            k = key_of(cols, 0, N).unsqueeze(1)
        else:
            k = torch.stack([key_of(cols, 0, split), key_of(cols, split, N)], 1)
        allk = k if uniq is None else torch.cat([uniq, k], 0)
        uniq = torch.unique(allk, dim=0) if allk.shape[1] > 1 else torch.unique(allk.squeeze(1)).unsqueeze(1)
        del allk, k, cols
        yield_rate = (uniq.shape[0] - have) / m
        rounds += 1
    # seeded random subset of exactly nnz coordinates, in a shuffled order
    sel = torch.randperm(uniq.shape[0], generator=gen, device=device)[:nnz]
    uniq = uniq[sel]
    coords = [None] * N
    if split == N:
        rem = uniq[:, 0].clone()
        for n in reversed(range(N)):
            coords[n] = (rem % dims[n]).contiguous()
            rem //= dims[n]
    else:
        rem = uniq[:, 1].clone()
        for n in reversed(range(split, N)):
            coords[n] = (rem % dims[n]).contiguous()
            rem //= dims[n]
        rem = uniq[:, 0].clone()
        for n in reversed(range(split)):
            coords[n] = (rem % dims[n]).contiguous()
            rem //= dims[n]
    k = uniq.shape[0]
    if values == "lognormal":
        vals = torch.exp(torch.randn(k, dtype=torch.float64, device=device, generator=gen))
    else:
        vals = 1.0 - torch.rand(k, dtype=torch.float64, device=device, generator=gen)
    return coords, vals


def poisson_kruskal_counts(dims, rank, expected_nnz, seed, shape=0.5, device="cuda", max_nnz=None):
    """Count tensor from a rank-`rank` Kruskal model: factor entries Gamma(shape,1)
    with columns normalized to sum 1; component r contributes Poisson(lambda_r)
    samples whose mode indices are drawn from U_n[:, r]; duplicates merged into
    counts (the scheme of src/poisson_generator.cpp:200-333).  lambda is scaled
    so that about expected_nnz distinct cells result (the draw count is grown
    until the distinct count reaches it)."""
    dims = [int(d) for d in dims]
    N = len(dims)
    gen = torch.Generator(device=device)
    gen.manual_seed(int(seed))
    U = []
    for d in dims:
        g = torch._standard_gamma(torch.full((d, rank), shape, dtype=torch.float64, device=device), generator=gen)
        U.append(g / g.sum(0, keepdim=True))
    w = torch.rand(rank, dtype=torch.float64, device=device, generator=gen) + 0.5
    w = w / w.sum()
    cdfs = [torch.cumsum(u, 0) for u in U]
    # split the modes into two groups whose index bits each fit 62 bits
    bits = [max((d - 1).bit_length(), 1) for d in dims]
    split, acc = 0, 0
    for n in range(N):
        if acc + bits[n] > 62:
            break
        acc += bits[n]
        split = n + 1

    def draw(total):
        m_r = torch.poisson(w * total, generator=gen).to(torch.int64)
        parts = []
        for r in range(rank):
            m = int(m_r[r].item())
            if m == 0:
                continue
            hi = torch.zeros(m, dtype=torch.int64, device=device)
            lo = torch.zeros(m, dtype=torch.int64, device=device)
            for n, d in enumerate(dims):
                u = torch.rand(m, dtype=torch.float64, device=device, generator=gen)
                idx = torch.searchsorted(cdfs[n][:, r].contiguous(), u * cdfs[n][-1, r]).clamp_(max=d - 1)
                if n < split:
                    hi = hi * d + idx
                else:
                    lo = lo * d + idx
            parts.append(torch.stack([hi, lo], 1))
        return torch.cat(parts, 0)

    total = float(expected_nnz)
    keys = draw(total)
    uk, counts = torch.unique(keys, dim=0, return_counts=True)
    for _ in range(6):
        if uk.shape[0] >= expected_nnz * 0.98:
            break
        extra = draw(total * (expected_nnz - uk.shape[0]) / max(uk.shape[0], 1) + 1024)
        keys = torch.cat([keys, extra], 0)
        uk, counts = torch.unique(keys, dim=0, return_counts=True)
    del keys
    if max_nnz is not None and uk.shape[0] > max_nnz:
        sel = torch.randperm(uk.shape[0], generator=gen, device=device)[:max_nnz]
        uk, counts = uk[sel], counts[sel]
    coords = [None] * N
    rem = uk[:, 1].clone()
    for n in reversed(range(split, N)):
        coords[n] = (rem % dims[n]).contiguous()
        rem //= dims[n]
    rem = uk[:, 0].clone()
    for n in reversed(range(split)):
        coords[n] = (rem % dims[n]).contiguous()
        rem //= dims[n]
    return coords, counts.to(torch.float64), U, w


def nnz_bits(dims):
    return sum(max(math.ceil(math.log2(d)) if d > 1 else 0, 0) for d in dims)
