"""GPU parity on the five benchmark workloads against the CPU oracle, each at
the largest size the oracle finishes in seconds:

- 1000^3 uniform, 1M nonzeros, rank 16 (the CPU configuration): 20 CP-ALS
  iterations and the mode-0 MTTKRP, full size;
- 12092 x 9184 x 28818, 76.9M Zipf(1.1) nonzeros, rank 32, FP64: the full
  tensor, key order and sample round trips, 5 CP-ALS fits within 1e-10;
- 2^20 x 2^18 x 2^16 x 1024 Poisson counts with 128-bit keys, rank 10 CP-APR:
  the same shape with 2M counts, 3 outer x 10 inner iterations;
- 50K x 50K x 10K x 2K x 512, Zipf(1.2), rank 64, 128-bit keys (66 index
  bits): the same shape with 3M nonzeros, per-mode FP32 / BF16 MTTKRP error
  against the FP64 oracle and 10-iteration CP-ALS fits;
- the serving stream: its first 30 jobs (3-5 modes, up to 81 index bits,
  ranks 8-64, up to 15M nonzeros), each job's result alone and in the mix.
"""
import numpy as np
import pytest

import oracle
from tests._tensors import uniform_tensor

pytestmark = pytest.mark.gpu


def _A():
    import alto_amd
    return alto_amd


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _device_tensor(dims, coords, vals, **kw):
    A = _A()
    return A.AltoTensor.from_device(dims, [c.data_ptr() for c in coords], vals.data_ptr(), int(vals.numel()), **kw)


def test_c1_cpu_config_full_size():
    """1000^3 uniform, 1M nonzeros, rank 16: 20 CP-ALS iterations and the
    mode-0 MTTKRP (the -p -t 0 timing loop) against the oracle."""
    A = _A()
    dims = [1000, 1000, 1000]
    cidx, vals = uniform_tensor(dims, 1_000_000, 1)
    g = A.AltoTensor(dims, cidx, vals, nprtn=8)
    o = oracle.OracleAlto(dims, cidx, vals, nprtn=8)
    rng = np.random.default_rng(1)
    U = [rng.random((d, 16)) for d in dims]
    assert _rel(g.mttkrp(0, U), o.mttkrp(0, U)) < 1e-13
    Ug, lg = oracle.random_init(dims, 16, 1)
    Uo, lo = oracle.random_init(dims, 16, 1)
    n, fit, _ = g.cpd_als(Ug, lg, 20, 0.0, A.PREC_F64)
    no, fo = o.cpd_als(Uo, lo, 20, 0.0)
    assert n == no == 20
    np.testing.assert_allclose(fit, fo, rtol=1e-10, atol=1e-12)


def test_c2_full_size_fp64():
    """The full 76.9M-nonzero nell-2-shape tensor: keys strictly increasing
    (sorted, unique), a sample of nonzeros found at their deposited keys with
    their values, and 5 FP64 CP-ALS fits within 1e-10 of the oracle run on
    the exported ALTO arrays."""
    import torch

    from alto_amd import gen

    A = _A()
    dims = [12092, 9184, 28818]
    coords, vals = gen.sparse_coo(dims, 76_900_000, 2, dist="zipf", s=1.1, values="lognormal")
    nnz = int(vals.numel())
    assert nnz == 76_900_000
    g = _device_tensor(dims, coords, vals, nprtn=16)
    st = g.export()
    keys = st["idx"][:, 0]
    assert np.all(keys[1:] > keys[:-1])
    masks, _, bits = oracle.setup_masks(dims, 1)
    assert bits == 43 and st["alto_bits"] == 43
    sel = torch.randint(0, nnz, (100_000,), device="cuda", generator=torch.Generator("cuda").manual_seed(2))
    c = [x[sel].cpu().numpy().astype(np.uint64) for x in coords]
    v = vals[sel].cpu().numpy()
    L = oracle.lib()
    want = np.zeros(len(v), np.uint64)
    for m in range(3):
        mk = int(masks[m][0])
        want |= np.array([L.or_pdep64(int(x), mk) for x in c[m]], np.uint64)
    pos = np.searchsorted(keys, want)
    assert np.array_equal(keys[pos], want)
    assert np.array_equal(st["vals"][pos], v)
    del coords, vals, sel
    o = oracle.OracleAlto.wrap(dims, 1, 16, st["masks"], st["idx"], st["vals"], st["prtn_ptr"], st["intervals"],
                               st["cr_masks"])
    del st, keys
    Ug, lg = oracle.random_init(dims, 32, 2)
    Uo, lo = oracle.random_init(dims, 32, 2)
    n, fit, _ = g.cpd_als(Ug, lg, 5, 0.0, A.PREC_F64)
    no, fo = o.cpd_als(Uo, lo, 5, 0.0)
    assert n == no == 5
    np.testing.assert_allclose(fit, fo, rtol=1e-10, atol=1e-12)
    np.testing.assert_allclose(lg, lo, rtol=1e-8)


def test_c3_shape_cp_apr_u128():
    """CP-APR on Poisson counts of the 2^20 x 2^18 x 2^16 x 1024 shape with a
    128-bit mask, rank 10: LL, fit, KKT and inner counts against the oracle
    (bench mode: every inner iteration runs)."""
    from alto_amd import gen

    A = _A()
    dims = [2 ** 20, 2 ** 18, 2 ** 16, 1024]
    coords, vals, _, _ = gen.poisson_kruskal_counts(dims, 10, 2_000_000, 3)
    nnz = int(vals.numel())
    assert 1_900_000 <= nnz <= 2_100_000
    g = _device_tensor(dims, coords, vals, nprtn=16, lit_bits=128, val_kind=A.VAL_U64)
    cidx = np.stack([c.cpu().numpy().astype(np.uint64) for c in coords])
    vh = vals.cpu().numpy()
    assert np.all(vh >= 1) and np.all(vh == np.round(vh))
    o = oracle.OracleAlto(dims, cidx, vh, W=2, nprtn=16, val_int=True)
    U, lam = oracle.random_init(dims, 10, 3)
    Uo, lo = oracle.random_init(dims, 10, 3)
    r = g.cp_apr(U, lam, 3, 10, epsilon=0.0, eps_div_zero=0.0, dobench=True)
    ro = o.cp_apr(Uo, lo, 3, 10, eps=0.0, eps_div_zero=0.0, dobench=True)
    assert r["outer_iters"] == ro["outer_iters"] == 3
    assert r["total_inner"] == ro["total_inner"] == 120
    np.testing.assert_allclose(r["loglik"], ro["loglik"], rtol=1e-9)
    np.testing.assert_allclose(r["ls_fit"], ro["ls_fit"], rtol=1e-9, atol=1e-11)
    np.testing.assert_allclose(r["kkt"], ro["kkt"], rtol=1e-6, atol=1e-12)
    np.testing.assert_allclose(lam, lo, rtol=1e-8)


def test_poisson_count_generator_follows_its_model():
    """The count-tensor generator: distinct cells, positive integer counts,
    about the requested number of cells, a normalised model, and per-mode
    count marginals equal to the model's expected marginals sum_r w_r U_n[:, r]
    (the draw scheme of src/poisson_generator.cpp:200-333)."""
    import torch

    from alto_amd import gen

    dims = [300, 200, 100, 50]
    coords, counts, U, w = gen.poisson_kruskal_counts(dims, 5, 200_000, 7)
    n = int(counts.numel())
    assert 0.98 * 200_000 <= n <= 1.05 * 200_000
    key = torch.zeros(n, dtype=torch.int64, device="cuda")
    for c, d in zip(coords, dims):
        assert int(c.min()) >= 0 and int(c.max()) < d
        key = key * d + c
    assert int(torch.unique(key).numel()) == n
    ch = counts.cpu().numpy()
    assert np.all(ch >= 1) and np.all(ch == np.round(ch))
    wh = w.cpu().numpy()
    assert abs(wh.sum() - 1) < 1e-12
    total = ch.sum()
    for m, d in enumerate(dims):
        Um = U[m].cpu().numpy()
        np.testing.assert_allclose(Um.sum(axis=0), 1.0, rtol=1e-12)
        emp = np.bincount(coords[m].cpu().numpy(), weights=ch, minlength=d) / total
        exp = Um @ wh
        # L1 distance against 3x its expectation under multinomial sampling
        # of `total` draws, sum_i sqrt(2 p_i (1 - p_i) / (pi total))
        noise = np.sqrt(2 * exp * (1 - exp) / (np.pi * total)).sum()
        assert np.abs(emp - exp).sum() < 3 * noise, (m, np.abs(emp - exp).sum(), noise)


C4_DIMS = [50000, 50000, 10000, 2000, 512]


@pytest.fixture(scope="module")
def c4_shape():
    """3M nonzeros of the 5-mode Zipf(1.2) workload (66 index bits)."""
    from alto_amd import gen

    coords, vals = gen.sparse_coo(C4_DIMS, 3_000_000, 4, dist="zipf", s=1.2, values="uniform")
    cidx = np.stack([c.cpu().numpy().astype(np.uint64) for c in coords])
    vh = vals.cpu().numpy()
    g = _device_tensor(C4_DIMS, coords, vals, nprtn=16)
    assert g.W == 2 and g.info().alto_bits == 66
    o = oracle.OracleAlto(C4_DIMS, cidx, vh, W=2, nprtn=16)
    return g, o


def test_c4_shape_mttkrp_precision(c4_shape):
    """Per-mode MTTKRP error against the FP64 oracle, rank 64: FP64 1e-12,
    FP32 (FP32 factors and arithmetic, FP64 row flushes) 1e-5, BF16 (factors
    stored in BF16, 2^-9 relative rounding each) 1e-2, normwise."""
    A = _A()
    g, o = c4_shape
    rng = np.random.default_rng(4)
    U = [rng.random((d, 64)) for d in C4_DIMS]
    for mode in range(5):
        want = o.mttkrp(mode, U)
        assert _rel(g.mttkrp(mode, U, A.PREC_F64), want) < 1e-12
        assert _rel(g.mttkrp(mode, U, A.PREC_F32), want) < 1e-5
        assert _rel(g.mttkrp(mode, U, A.PREC_BF16), want) < 1e-2


def test_c4_shape_cp_als_fit_delta(c4_shape):
    """10 CP-ALS iterations at rank 64: the FP64 fit trace within 1e-10 of
    the oracle; FP32 within 1e-4 and BF16 within 1e-3 of it."""
    A = _A()
    g, o = c4_shape
    Uo, lo = oracle.random_init(C4_DIMS, 64, 4)
    no, fo = o.cpd_als(Uo, lo, 10, 0.0)
    assert no == 10
    for prec, tol in ((A.PREC_F64, 1e-10), (A.PREC_F32, 1e-4), (A.PREC_BF16, 1e-3)):
        U, lam = oracle.random_init(C4_DIMS, 64, 4)
        n, fit, _ = g.cpd_als(U, lam, 10, 0.0, prec)
        assert n == 10
        if prec == A.PREC_F64:
            np.testing.assert_allclose(fit, fo, rtol=tol, atol=1e-12)
        else:
            assert abs(fit[-1] - fo[-1]) <= tol * max(1.0, abs(fo[-1])), (prec, fit[-1], fo[-1])


def test_c5_stream_jobs_alone_equal_in_mix():
    """The first 30 jobs of the seeded serving stream (5 with 65-81 index
    bits, 6 of rank 64, 10K-15M nonzeros), FP64 / FP32 / BF16: each job's
    fit and iteration count through the server with 8 slots equal those with
    one slot (FP64 1e-10, FP32 1e-5, BF16 1e-3)."""
    from alto_amd.serve import Server, job_stream, materialize

    jobs = materialize(job_stream(30, seed=5))
    bits = [sum(max((d - 1).bit_length(), 1) for d in j["dims"]) for j in jobs]
    assert sum(b > 64 for b in bits) >= 1 and sum(j["rank"] == 64 for j in jobs) >= 1
    for prec, tol in ((0, 1e-10), (1, 1e-5), (2, 1e-3)):
        res = {}
        for slots in (1, 8):
            s = Server(200 << 30, slots)
            ids = [s.submit(j["dims"], [c.data_ptr() for c in j["coords"]], j["vals"].data_ptr(), j["rank"], 20, 1e-5,
                            j["seed"] % (2 ** 31), prec, device=True, nnz=j["nnz"]) for j in jobs]
            res[slots] = [s.result(i)[0] for i in ids]
            s.close()
        for a, b in zip(res[1], res[8]):
            assert a.status == 2 and b.status == 2
            assert abs(a.fit - b.fit) <= tol * max(1.0, abs(a.fit)), (prec, a.fit, b.fit)
            if prec == 0:
                assert a.iters == b.iters
