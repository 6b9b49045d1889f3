"""GPU parity of alto_cpd_sessions_iterate: sessions of uneven tensors
iterated in lockstep (one batched MTTKRP per mode index, then each session's
solves, then all fits) end where each session's own iterate leaves it --
FP64 within 1e-10, FP32 within 1e-5 and BF16 within 1e-3 of the fit -- and
stop at the same iteration under a fit tolerance."""
import numpy as np
import pytest

from tests._tensors import uniform_tensor, zipf_tensor

pytestmark = pytest.mark.gpu

SPECS = [
    ([60, 70, 80], 8000, "uniform"),
    ([300, 200, 500], 40000, "zipf"),
    ([30, 20, 10, 40], 5000, "uniform"),
    ([40, 30, 20, 25, 15], 6000, "uniform"),
    ([500, 3000, 40], 20000, "zipf"),
]


def _A():
    import alto_amd
    return alto_amd


def _tensors(specs, seed0=0):
    A = _A()
    out = []
    for k, (dims, nnz, kind) in enumerate(specs):
        gen = zipf_tensor if kind == "zipf" else uniform_tensor
        cidx, vals = gen(dims, nnz, 200 + seed0 + k)
        out.append((dims, A.AltoTensor(dims, cidx, vals, nprtn=8)))
    return out


@pytest.mark.parametrize("prec,tol", [(0, 1e-10), (1, 1e-5), (2, 1e-3)])
def test_lockstep_equals_each_session_alone(prec, tol):
    A = _A()
    ts = _tensors(SPECS)
    R, iters = 16, 8
    alone, lock = [], []
    for k, (dims, g) in enumerate(ts):
        U, _ = A.kruskal_random_init(dims, R, 10 + k)
        alone.append(A.CpdSession(g, U, precision=prec))
        lock.append(A.CpdSession(g, U, precision=prec))
    for s in alone:
        s.iterate(iters)
    ran = A.cpd_sessions_iterate(lock, iters)
    assert ran == [iters] * len(ts)
    for (dims, _), a, b in zip(ts, alone, lock):
        Ua, la, fa, _ = a.get()
        Ub, lb, fb, _ = b.get()
        assert abs(fa - fb) <= tol, (dims, fa, fb)
        if prec == 0:
            np.testing.assert_allclose(lb, la, rtol=1e-8)
            for ua, ub in zip(Ua, Ub):
                np.testing.assert_allclose(ub, ua, rtol=1e-7, atol=1e-9)


def test_lockstep_convergence_and_continuation():
    """Under a fit tolerance each session stops where it would alone, the
    others go on; a later lockstep call continues the iteration count."""
    A = _A()
    ts = _tensors(SPECS[:4], seed0=5)
    R = 8
    alone, lock = [], []
    for k, (dims, g) in enumerate(ts):
        U, _ = A.kruskal_random_init(dims, R, 30 + k)
        alone.append(A.CpdSession(g, U))
        lock.append(A.CpdSession(g, U))
    want = []
    for s in alone:
        f, _ = s.iterate(60, 1e-4)
        want.append(len(f))
    ran = A.cpd_sessions_iterate(lock, 3, 1e-4)
    ran2 = A.cpd_sessions_iterate(lock, 57, 1e-4)
    assert [a + b for a, b in zip(ran, ran2)] == want
    for a, b in zip(alone, lock):
        assert abs(a.get()[2] - b.get()[2]) <= 1e-10


def test_lockstep_arguments():
    A = _A()
    ts = _tensors(SPECS[:2])
    U0, _ = A.kruskal_random_init(ts[0][0], 8, 1)
    U1, _ = A.kruskal_random_init(ts[1][0], 16, 1)
    s0 = A.CpdSession(ts[0][1], U0)
    s1 = A.CpdSession(ts[1][1], U1)
    assert A.cpd_sessions_iterate([], 5) == []
    with pytest.raises(A.AltoError):  # ranks differ
        A.cpd_sessions_iterate([s0, s1], 2)
    with pytest.raises(A.AltoError):  # one session twice
        A.cpd_sessions_iterate([s0, s0], 2)
    assert A.cpd_sessions_iterate([s0], 0) == [0]


def _compare(ts, alone, lock, tol=1e-10):
    for (dims, _), a, b in zip(ts, alone, lock):
        Ua, la, fa, _ = a.get()
        Ub, lb, fb, _ = b.get()
        assert abs(fa - fb) <= tol, (dims, fa, fb)
        np.testing.assert_allclose(lb, la, rtol=1e-8)
        for ua, ub in zip(Ua, Ub):
            np.testing.assert_allclose(ub, ua, rtol=1e-7, atol=1e-9)


def test_lockstep_single_session():
    """One session in lockstep (solves on the caller's stream, no helper
    streams) for several iterations equals its own iterate."""
    A = _A()
    ts = _tensors(SPECS[1:2], seed0=30)
    dims, g = ts[0]
    U, _ = A.kruskal_random_init(dims, 16, 3)
    alone, lock = [A.CpdSession(g, U)], [A.CpdSession(g, U)]
    alone[0].iterate(6)
    assert A.cpd_sessions_iterate(lock, 6) == [6]
    _compare(ts, alone, lock)


def test_lockstep_sessions_on_their_own_streams():
    """Sessions created on streams of their own, iterated in lockstep on a
    different stream, equal each session alone on its stream."""
    import torch
    A = _A()
    ts = _tensors(SPECS, seed0=40)
    streams = [torch.cuda.Stream() for _ in ts]
    lock_stream = torch.cuda.Stream()
    alone, lock = [], []
    for k, ((dims, g), st) in enumerate(zip(ts, streams)):
        U, _ = A.kruskal_random_init(dims, 16, 20 + k)
        alone.append(A.CpdSession(g, U, stream=st.cuda_stream))
        lock.append(A.CpdSession(g, U, stream=st.cuda_stream))
    for s in alone:
        s.iterate(5)
    assert A.cpd_sessions_iterate(lock, 5, stream=lock_stream.cuda_stream) == [5] * len(ts)
    torch.cuda.synchronize()
    _compare(ts, alone, lock)


def test_lockstep_with_hot_rows():
    """Rank-ordered streams with LDS-staged hot rows (ALTO_TUNE_HOT_ROWS) in
    lockstep equal each session alone."""
    A = _A()
    A.check(A.lib().alto_tune_set(7, 1), "tune_set")
    try:
        ts = _tensors(SPECS[:3] + SPECS[4:], seed0=50)
        alone, lock = [], []
        for k, (dims, g) in enumerate(ts):
            U, _ = A.kruskal_random_init(dims, 16, 30 + k)
            alone.append(A.CpdSession(g, U))
            lock.append(A.CpdSession(g, U))
        for s in alone:
            s.iterate(5)
        assert A.cpd_sessions_iterate(lock, 5) == [5] * len(ts)
        _compare(ts, alone, lock)
    finally:
        A.check(A.lib().alto_tune_set(7, 0), "tune_set")


DET, DENSE_BATCH, LOCK_GRAPHS = 18, 20, 21


def _bit_identical(ts, alone, lock):
    for (dims, _), a, b in zip(ts, alone, lock):
        Ua, la, fa, _ = a.get()
        Ub, lb, fb, _ = b.get()
        assert fa == fb, (dims, fa, fb)
        assert np.array_equal(la, lb), dims
        assert all(np.array_equal(ua, ub) for ua, ub in zip(Ua, Ub)), dims


def _det_lockstep(specs, R, prec, iters, seed0, dense_batch=1, eps=0.0, graphs=0):
    """Sessions alone and in lockstep with deterministic reductions: the
    batched MTTKRP sums in the same order as each tensor's own, and the
    batched dense launches run the single-session kernels' bodies over the
    same blocks, so the two agree bit for bit (graphs: repeated iterations
    replayed from a captured graph)."""
    A = _A()
    lib = A.lib()
    A.check(lib.alto_tune_set(DET, 1), "tune_set")
    A.check(lib.alto_tune_set(DENSE_BATCH, dense_batch), "tune_set")
    A.check(lib.alto_tune_set(LOCK_GRAPHS, graphs), "tune_set")
    try:
        ts = _tensors(specs, seed0=seed0)
        alone, lock = [], []
        for k, (dims, g) in enumerate(ts):
            U, _ = A.kruskal_random_init(dims, R, seed0 + k)
            alone.append(A.CpdSession(g, U, precision=prec))
            lock.append(A.CpdSession(g, U, precision=prec))
        want = [len(s.iterate(iters, eps)[0]) for s in alone]
        assert A.cpd_sessions_iterate(lock, iters, eps) == want
        _bit_identical(ts, alone, lock)
    finally:
        A.check(lib.alto_tune_set(DET, 0), "tune_set")
        A.check(lib.alto_tune_set(DENSE_BATCH, 1), "tune_set")
        A.check(lib.alto_tune_set(LOCK_GRAPHS, 0), "tune_set")


@pytest.mark.parametrize("R", [1, 7, 16, 33, 64, 65])
def test_lockstep_batched_dense_bit_identical_fp64(R):
    """Ranks through every lane shape of the batched dense kernels (1-64)
    and one above them (65: the per-session chain)."""
    _det_lockstep(SPECS, R, 0, 5, 60)


@pytest.mark.parametrize("prec", [1, 2])
@pytest.mark.parametrize("R", [8, 33])
def test_lockstep_batched_dense_bit_identical_reduced(prec, R):
    _det_lockstep(SPECS, R, prec, 5, 70)


def test_lockstep_per_session_chain_bit_identical():
    """The per-session chains on the helper streams (batching off)."""
    _det_lockstep(SPECS, 16, 0, 5, 80, dense_batch=0)


def test_lockstep_batched_dense_many_sessions_and_convergence():
    """20 sessions (two groups of batched launches) under a fit tolerance:
    each stops where it would alone, with the same values."""
    rng = np.random.default_rng(3)
    specs = []
    for k in range(20):
        nm = 3 + k % 3
        dims = [int(x) for x in rng.integers(10, 120, size=nm)]
        specs.append((dims, int(rng.integers(500, 6000)), "zipf" if k % 2 else "uniform"))
    _det_lockstep(specs, 12, 0, 40, 90, eps=1e-4)


@pytest.mark.parametrize("prec,R", [(0, 16), (0, 64), (1, 33), (2, 8)])
def test_lockstep_graph_replays_bit_identical(prec, R):
    """Iterations replayed from the captured graph (from the third on) equal
    each session's eager run bit for bit."""
    _det_lockstep(SPECS, R, prec, 8, 100, graphs=1)


def test_lockstep_graph_recaptured_as_sessions_converge():
    """Sessions that converge leave the set: the next iterations run eagerly
    once and are captured again; every session stops where it would alone
    (each stop depends on the fits of every replayed iteration) with the
    same values."""
    rng = np.random.default_rng(5)
    specs = []
    for k in range(18):
        dims = [int(x) for x in rng.integers(10, 90, size=3 + k % 3)]
        specs.append((dims, int(rng.integers(400, 4000)), "zipf" if k % 2 else "uniform"))
    _det_lockstep(specs, 10, 0, 50, 110, eps=1e-4, graphs=1)


@pytest.mark.parametrize("prec,tol", [(0, 1e-10), (1, 1e-5)])
def test_lockstep_graph_with_atomic_reductions(prec, tol):
    """Graph replays with the default (atomic) reductions stay within the
    stated tolerance of each session alone."""
    A = _A()
    lib = A.lib()
    A.check(lib.alto_tune_set(LOCK_GRAPHS, 1), "tune_set")
    try:
        ts = _tensors(SPECS, seed0=120)
        alone, lock = [], []
        for k, (dims, g) in enumerate(ts):
            U, _ = A.kruskal_random_init(dims, 16, 120 + k)
            alone.append(A.CpdSession(g, U, precision=prec))
            lock.append(A.CpdSession(g, U, precision=prec))
        for x in alone:
            x.iterate(10)
        assert A.cpd_sessions_iterate(lock, 10) == [10] * len(ts)
        for (dims, _), a, b in zip(ts, alone, lock):
            assert abs(a.get()[2] - b.get()[2]) <= tol, dims
    finally:
        A.check(lib.alto_tune_set(LOCK_GRAPHS, 0), "tune_set")


@pytest.mark.parametrize("dense64", [1, 0])
@pytest.mark.parametrize("prec,tol", [(1, 1e-5), (2, 1e-3)])
def test_reduced_precision_against_oracle(prec, tol, dense64):
    """FP32 / BF16 sessions, alone and in lockstep, with the Grams and row
    solves on FP64 (ALTO_TUNE_DENSE_F64, the default) or FP32 MFMA, against
    the oracle's FP64 CP-ALS from the same initial model: the fit after 8
    iterations within 1e-5 (FP32) and 1e-3 (BF16)."""
    import oracle
    A = _A()
    A.check(A.lib().alto_tune_set(24, dense64), "tune_set")
    try:
        specs = SPECS[:2] + SPECS[3:]
        data, alone, lock = [], [], []
        for k, (dims, nnz, kind) in enumerate(specs):
            gen = zipf_tensor if kind == "zipf" else uniform_tensor
            cidx, vals = gen(dims, nnz, 400 + k)
            g = A.AltoTensor(dims, cidx, vals, nprtn=8)
            U, lam = oracle.random_init(dims, 16, 40 + k)
            data.append((dims, cidx, vals, g))
            alone.append(A.CpdSession(g, U, precision=prec))
            lock.append(A.CpdSession(g, U, precision=prec))
        for x in alone:
            x.iterate(8)
        assert A.cpd_sessions_iterate(lock, 8) == [8] * len(specs)
        for k, ((dims, cidx, vals, _), a, b) in enumerate(zip(data, alone, lock)):
            o = oracle.OracleAlto(dims, cidx, vals, nprtn=8)
            Uo, lo = oracle.random_init(dims, 16, 40 + k)
            _, fo = o.cpd_als(Uo, lo, 8, 0.0)
            assert abs(a.get()[2] - fo[-1]) <= tol, (dims, a.get()[2], fo[-1])
            assert abs(b.get()[2] - fo[-1]) <= tol, (dims, b.get()[2], fo[-1])
    finally:
        A.check(A.lib().alto_tune_set(24, 1), "tune_set")
