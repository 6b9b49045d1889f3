"""Deterministic-order MTTKRP and Phi (ALTO_TUNE_DETERMINISTIC): the same
values as the oracle (FP64 within 1e-12; FP32 / BF16 within their bounds),
and bit-identical results from run to run -- MTTKRP in every precision,
Phi, and whole CP-ALS / CP-APR runs -- on tensors whose heavy rows span many
segments of the walk (the rows summed across segments by the second
launch), with 64- and 128-bit keys and 3 to 5 modes."""
import numpy as np
import pytest

import oracle
from tests._tensors import uniform_tensor, zipf_tensor

pytestmark = pytest.mark.gpu

DET = 18


def _A():
    import alto_amd
    return alto_amd


@pytest.fixture
def det():
    A = _A()
    A.check(A.lib().alto_tune_set(DET, 1), "tune_set")
    yield
    A.check(A.lib().alto_tune_set(DET, 0), "tune_set")


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


CASES = [([3, 2000, 3000], 200000, 1, "zipf"),   # mode 0: ~67K nonzeros per row, 65+ segments each
         ([60, 70, 80], 30000, 1, "uniform"),
         ([30, 20, 10, 40], 15000, 1, "zipf"),
         ([2 ** 14, 2 ** 13, 40, 2 ** 12, 2 ** 14], 20000, 2, "zipf"),
         ([500, 9, 700], 40000, 1, "zipf")]


def _tensor(dims, nnz, seed, kind):
    if kind == "zipf":
        return zipf_tensor(dims, nnz, seed, s=1.2)
    return uniform_tensor(dims, nnz, seed)


@pytest.mark.parametrize("dims,nnz,W,kind", CASES)
@pytest.mark.parametrize("R", [1, 8, 16, 33, 64, 100])
def test_det_mttkrp_fp64_matches_oracle_and_repeats(det, dims, nnz, W, kind, R):
    A = _A()
    cidx, vals = _tensor(dims, nnz, 40 + R, kind)
    g = A.AltoTensor(dims, cidx, vals, nprtn=8, lit_bits=64 * W)
    o = oracle.OracleAlto(dims, cidx, vals, W=W, nprtn=8)
    rng = np.random.default_rng(R)
    U = [rng.random((d, R)) for d in dims]
    for mode in range(len(dims)):
        a = g.mttkrp(mode, U)
        assert _rel(a, o.mttkrp(mode, U)) < 1e-12, mode
        assert np.array_equal(a, g.mttkrp(mode, U)), mode


@pytest.mark.parametrize("prec,tol", [(1, 1e-5), (2, 2e-2)])
@pytest.mark.parametrize("R", [8, 32, 64])
def test_det_mttkrp_reduced_precision_repeats(det, prec, tol, R):
    A = _A()
    dims = [3, 2000, 3000]
    cidx, vals = zipf_tensor(dims, 200000, 7, s=1.2)
    g = A.AltoTensor(dims, cidx, vals, nprtn=8)
    o = oracle.OracleAlto(dims, cidx, vals, nprtn=8)
    rng = np.random.default_rng(R)
    U = [rng.random((d, R)) for d in dims]
    for mode in range(3):
        a = g.mttkrp(mode, U, prec)
        assert _rel(a, o.mttkrp(mode, U)) < tol, mode
        assert np.array_equal(a, g.mttkrp(mode, U, prec)), mode


@pytest.mark.parametrize("R", [2, 10, 33])
def test_det_phi_matches_oracle_and_repeats(det, R):
    A = _A()
    for dims, nnz, W, kind in CASES[:4]:
        cidx, vals = _tensor(dims, nnz, 3, kind)
        g = A.AltoTensor(dims, cidx, vals, nprtn=8, lit_bits=64 * W)
        o = oracle.OracleAlto(dims, cidx, vals, W=W, nprtn=8)
        rng = np.random.default_rng(R)
        U = [rng.random((d, R)) + 0.05 for d in dims]
        for mode in range(len(dims)):
            a = g.phi(mode, U, 1e-10)
            want = o.phi(mode, U, 1e-10)
            np.testing.assert_allclose(a, want, rtol=1e-11, atol=1e-12 * max(np.abs(want).max(), 1.0))
            assert np.array_equal(a, g.phi(mode, U, 1e-10)), (dims, mode)


def test_det_batched_equals_single(det):
    A = _A()
    ts, Us, modes = [], [], []
    for k, (dims, nnz, W, kind) in enumerate(CASES[:3]):
        cidx, vals = _tensor(dims, nnz, 60 + k, kind)
        ts.append(A.AltoTensor(dims, cidx, vals, nprtn=4))
        rng = np.random.default_rng(k)
        Us.append([rng.random((d, 16)) for d in dims])
        modes.append(k % len(dims))
    single = [t.mttkrp(m, U) for t, m, U in zip(ts, modes, Us)]
    for a, b in zip(single, A.mttkrp_batched(ts, modes, Us)):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("prec", [0, 1, 2])
def test_det_cp_als_runs_are_bit_identical(det, prec):
    A = _A()
    dims = [3, 2000, 3000]
    cidx, vals = zipf_tensor(dims, 200000, 9, s=1.2)
    g = A.AltoTensor(dims, cidx, vals, nprtn=8)
    runs = []
    for _ in range(2):
        U, lam = oracle.random_init(dims, 16, 4)
        n, fit, _ = g.cpd_als(U, lam, 8, 0.0, prec)
        runs.append((fit, U, lam))
    assert np.array_equal(runs[0][0], runs[1][0])
    assert all(np.array_equal(a, b) for a, b in zip(runs[0][1], runs[1][1]))
    if prec == 0:
        o = oracle.OracleAlto(dims, cidx, vals, nprtn=8)
        U, lam = oracle.random_init(dims, 16, 4)
        _, fo = o.cpd_als(U, lam, 8, 0.0)
        np.testing.assert_allclose(runs[0][0], fo, rtol=1e-10, atol=1e-12)


def test_det_cp_apr_runs_are_bit_identical(det):
    A = _A()
    dims = [40, 300, 500]
    cidx, vals = zipf_tensor(dims, 60000, 10, s=1.2)
    vals = np.ceil(vals * 3)
    g = A.AltoTensor(dims, cidx, vals, nprtn=8)
    res = []
    for _ in range(2):
        U, lam = oracle.random_init(dims, 6, 5)
        r = g.cp_apr(U, lam, 4, 3, epsilon=1e-4)
        res.append((r["loglik"], r["kkt"], U, lam))
    assert res[0][0] == res[1][0] and res[0][1] == res[1][1]
    assert all(np.array_equal(a, b) for a, b in zip(res[0][2], res[1][2]))
    o = oracle.OracleAlto(dims, cidx, vals, nprtn=8)
    U, lam = oracle.random_init(dims, 6, 5)
    ro = o.cp_apr(U, lam, 4, 3, eps=1e-4)
    assert abs(res[0][0] - ro["loglik"]) <= 1e-9 * abs(ro["loglik"])
