"""ORACLE — test infrastructure only.

Pure-Python restatement of the reference's Poisson tensor generator
(src/poisson_generator.cpp, used by the command line's -d / -s options through
MakeSparseTensor, src/main.cpp:653-675) for a given OpenMP thread count T,
to check the command line's generator on small tensors.  Its MT19937-64 is
pinned by the known answer the C++ standard fixes for ``std::mt19937_64``
(the 10000th draw from the default seed 5489 is 9981545732273789042), its
``rand()`` is glibc's own (called through ctypes after ``srand(1)``, the
state a fresh process starts from).  Column sums run in row order (one
thread); see each function for the reference lines it follows.
"""
from __future__ import annotations

import ctypes
import math

MASK64 = (1 << 64) - 1
BATCH = 100000  # BSIZE, src/poisson_generator.cpp:13
SEED = 1234     # SEED, src/poisson_generator.cpp:14


class MT64:
    """MT19937-64 (src/rng.cpp:69-110: init_genrand64 / genrand64_int64)."""
    NN, MM = 312, 156
    A, UM, LM = 0xB5026F5AA96619E9, 0xFFFFFFFF80000000, 0x7FFFFFFF

    def __init__(self, seed):
        mt = [0] * self.NN
        mt[0] = seed & MASK64
        for i in range(1, self.NN):
            mt[i] = (6364136223846793005 * (mt[i - 1] ^ (mt[i - 1] >> 62)) + i) & MASK64
        self.mt, self.i = mt, self.NN

    def _twist(self):
        mt, NN, MM = self.mt, self.NN, self.MM
        for k in range(NN):
            x = (mt[k] & self.UM) | (mt[(k + 1) % NN] & self.LM)
            xa = x >> 1
            if x & 1:
                xa ^= self.A
            mt[k] = mt[(k + MM) % NN] ^ xa
        self.i = 0

    def __call__(self):
        if self.i >= self.NN:
            self._twist()
        x = self.mt[self.i]
        self.i += 1
        x ^= (x >> 29) & 0x5555555555555555
        x ^= (x << 17) & 0x71D67FFFEDA60000
        x ^= (x << 37) & 0xFFF7EEE000000000
        x ^= x >> 43
        return x & MASK64


class Streams:
    """One MT19937-64 per thread, seeded 1234 + t (:120-127); a batch of n
    uniform draws is split in T chunks of ceil(n / T), chunk t from stream t
    (UniformRandom, :20-34; RNGStreamUniform, include/rng_stream.hpp:21-28)."""

    def __init__(self, threads):
        self.s = [MT64(SEED + t) for t in range(threads)]

    def uniform(self, n, low, high):
        T = len(self.s)
        chunk = (n + T - 1) // T
        out = [0.0] * n
        for t in range(T):
            for i in range(t * chunk, min(n, (t + 1) * chunk)):
                u = (float(self.s[t]() >> 12) + 0.5) * (1.0 / 4503599627370496.0)  # include/rng.hpp:20
                out[i] = u * (high - low) + low
        return out


def binning(rv, bins):
    """Binning (:37-56): binary search over len(bins) edges."""
    start, end = 0, len(bins) - 1
    while start <= end:
        mid = (end + start) // 2
        if rv > bins[mid]:
            start = mid + 1
        elif rv == bins[mid]:
            return mid - 1 if mid == len(bins) - 1 else mid
        else:
            end = mid - 1
    return end


def poisson_tensor(dims, num_edges, rank, threads=1, return_model=False):
    """PoissonGeneratorRun (:336-376) -> (cidx: list of per-mode lists, vals),
    and the model the edges were drawn from, (factors, lambda) as float64
    arrays, when return_model is set."""
    libc = ctypes.CDLL("libc.so.6")
    libc.srand(1)
    rng = Streams(threads)
    N = len(dims)
    U = []
    for d in dims:  # GenerateRandomModel (:144-192)
        u = [0.1] * (d * rank)
        nbig = int(math.ceil(0.2 * d))
        for r in range(rank):
            perm = list(range(d))  # RandomPerm (:87-98)
            for i in range(nbig):
                j = libc.rand() % (d - i) + i
                perm[i], perm[j] = perm[j], perm[i]
            vals = rng.uniform(nbig, 1.0, 10.0 * rank)
            for i in range(nbig):
                u[perm[i] * rank + r] = vals[i]
        U.append(u)
    lam = rng.uniform(rank, 0.0, 1.0)
    for n, d in enumerate(dims):  # KruskalModelNormalize (src/kruskal_model.cpp:128-176)
        cn = [0.0] * rank
        for i in range(d):
            for j in range(rank):
                cn[j] += abs(U[n][i * rank + j])
        for i in range(d):
            for j in range(rank):
                U[n][i * rank + j] /= cn[j]
        for j in range(rank):
            lam[j] *= cn[j]
    tot = 0.0
    for j in range(rank):
        tot += lam[j]
    tot = 1.0 / tot
    lam = [x * tot for x in lam]
    bins = [0.0]  # GenerateEdges (:195-260)
    for i in range(rank):
        bins.append(min(lam[i] + bins[i], 1.0))
    rcount = [0] * (rank + 1)
    for b in range(0, num_edges, BATCH):  # ProbCount (:72-84)
        for x in rng.uniform(min(BATCH, num_edges - b), 0.0, 1.0):
            rcount[binning(x, bins)] += 1
    edges = []
    for r in range(rank):
        nc = rcount[r]
        comp = [[0] * N for _ in range(nc)]
        for n, d in enumerate(dims):
            cb = [0.0]
            for i in range(d):
                cb.append(min(U[n][i * rank + r] + cb[i], 1.0))
            for b in range(0, nc, BATCH):  # ProbSample (:59-70)
                for k, x in enumerate(rng.uniform(min(BATCH, nc - b), 0.0, 1.0)):
                    comp[b + k][n] = binning(x, cb)
        edges.extend(comp)
    # ConstructCIdx (:263-333): least-significant-mode-first stable sort, merge
    for n in range(N):
        edges.sort(key=lambda e: e[n])
    cidx, vals = [], []
    for e in edges:
        if cidx and cidx[-1] == e:
            vals[-1] += 1.0
        else:
            cidx.append(e)
            vals.append(1.0)
    out = ([[c[n] for c in cidx] for n in range(N)], vals)
    if return_model:
        import numpy as np
        model = ([np.array(U[n], np.float64).reshape(d, rank) for n, d in enumerate(dims)],
                 np.array(lam, np.float64))
        return out + (model,)
    return out
