"""GPU parity of CP-APR (cp_apr_alto) through the C ABI against the CPU
oracle: log-likelihood, least-squares fit, KKT and iteration counts, final
factors (FP64, relative 1e-9), in bench mode and with early termination."""
import numpy as np
import pytest

import oracle
from tests._tensors import uniform_tensor

pytestmark = pytest.mark.gpu


def _counts(dims, nnz, seed):
    cidx, _ = uniform_tensor(dims, nnz, seed)
    rng = np.random.default_rng(seed)
    vals = rng.integers(1, 9, size=nnz).astype(np.float64)
    return cidx, vals


@pytest.mark.parametrize("dims,nnz,R,W", [([30, 40, 20], 3000, 5, 1), ([20, 10, 15, 12], 2500, 10, 2),
                                          ([64, 50, 40, 30, 20], 6000, 8, 1), ([500, 400, 300], 900, 6, 1)])
@pytest.mark.parametrize("bench", [True, False])
def test_cp_apr_matches_oracle(dims, nnz, R, W, bench):
    import alto_amd as A

    cidx, vals = _counts(dims, nnz, 3)
    g = A.AltoTensor(dims, cidx, vals, nprtn=8, lit_bits=64 * W, val_kind=A.VAL_U64)
    U, lam = oracle.random_init(dims, R, 11)
    eps = 0.0 if bench else 1e-4
    r = g.cp_apr(U, lam, 6, 5, epsilon=eps, eps_div_zero=0.0 if bench else 1e-10, dobench=bench)
    o = oracle.OracleAlto(dims, cidx, vals, W=W, nprtn=8, val_int=True)
    Uo, lo = oracle.random_init(dims, R, 11)
    ro = o.cp_apr(Uo, lo, 6, 5, eps=eps, eps_div_zero=0.0 if bench else 1e-10, dobench=bench)
    assert r["outer_iters"] == ro["outer_iters"]
    assert r["total_inner"] == ro["total_inner"]
    assert r["inner_per_mode"][: len(dims)] == ro["inner_per_mode"][: len(dims)]
    np.testing.assert_allclose(r["loglik"], ro["loglik"], rtol=1e-9)
    np.testing.assert_allclose(r["ls_fit"], ro["ls_fit"], rtol=1e-9, atol=1e-11)
    np.testing.assert_allclose(r["kkt"], ro["kkt"], rtol=1e-7, atol=1e-12)
    np.testing.assert_allclose(lam, lo, rtol=1e-9)
    for a, b in zip(U, Uo):
        np.testing.assert_allclose(a, b, rtol=1e-8, atol=1e-12)


def test_cp_apr_pi_cache_equals_gathers():
    """CP-APR with the other modes' row products kept per nonzero across a
    mode's inner iterations (the default) equals the run that gathers them
    in every inner iteration (3- and 4-mode tensors, 64- and 128-bit keys,
    bench mode and early stops)."""
    import alto_amd as A
    for dims, nnz, W, bench in (([40, 300, 500], 40000, 1, False), ([30, 20, 10, 40], 15000, 2, True)):
        cidx, vals = uniform_tensor(dims, nnz, 21)
        vals = np.ceil(vals * 4)
        g = A.AltoTensor(dims, cidx, vals, nprtn=8, lit_bits=64 * W)
        out = []
        for cache in (1, 0):
            A.check(A.lib().alto_tune_set(19, cache), "tune_set")
            try:
                U, lam = oracle.random_init(dims, 6, 3)
                r = g.cp_apr(U, lam, 5, 4, epsilon=1e-4, dobench=bench)
            finally:
                A.check(A.lib().alto_tune_set(19, 1), "tune_set")
            out.append((r, U, lam))
        (r1, U1, l1), (r0, U0, l0) = out
        assert r1["outer_iters"] == r0["outer_iters"] and r1["total_inner"] == r0["total_inner"]
        assert abs(r1["loglik"] - r0["loglik"]) <= 1e-10 * abs(r0["loglik"])
        for a, b in zip(U1, U0):
            np.testing.assert_allclose(a, b, rtol=1e-9, atol=1e-12)


@pytest.mark.parametrize("R", [6, 7, 10, 16])
def test_cp_apr_phi16_with_pi_cache_matches_oracle(R):
    """The 16-byte-lane Phi (ALTO_TUNE_PHI_VARIANT = 2) with the product
    cache (store, then stream) against the oracle's CP-APR, 3- and 4-mode
    tensors, 64- and 128-bit keys; odd ranks fall back to the 8-byte-lane
    kernel."""
    import alto_amd as A
    A.check(A.lib().alto_tune_set(11, 2), "tune_set")
    try:
        for dims, nnz, W, bench in (([40, 300, 500], 40000, 1, True), ([30, 20, 10, 40], 15000, 2, False)):
            cidx, vals = _counts(dims, nnz, 31 + R)
            g = A.AltoTensor(dims, cidx, vals, nprtn=8, lit_bits=64 * W, val_kind=A.VAL_U64)
            o = oracle.OracleAlto(dims, cidx, vals, W=W, nprtn=8, val_int=True)
            U, lam = oracle.random_init(dims, R, 4)
            Uo, lo = oracle.random_init(dims, R, 4)
            eps = 0.0 if bench else 1e-4
            r = g.cp_apr(U, lam, 4, 5, epsilon=eps, eps_div_zero=0.0 if bench else 1e-10, dobench=bench)
            ro = o.cp_apr(Uo, lo, 4, 5, eps=eps, eps_div_zero=0.0 if bench else 1e-10, dobench=bench)
            assert r["outer_iters"] == ro["outer_iters"] and r["total_inner"] == ro["total_inner"]
            np.testing.assert_allclose(r["loglik"], ro["loglik"], rtol=1e-9)
            np.testing.assert_allclose(lam, lo, rtol=1e-9)
            for a, b in zip(U, Uo):
                np.testing.assert_allclose(a, b, rtol=1e-8, atol=1e-12)
    finally:
        A.check(A.lib().alto_tune_set(11, 1), "tune_set")
