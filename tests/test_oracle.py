"""Pins the CPU oracle before it is trusted as the checker.

Each oracle function is checked against something independent of its own
formula: hardware/brute-force bit operations, the reference's worked mask
example, the reference's own compiled create_alto / mttkrp_alto_par /
CalculatePhi_alto_par (oracle/_ref, built from /root/reference), numpy
brute force from the definitions, and scipy's LAPACK solvers.
"""
import ctypes
import os

import numpy as np
import pytest
import scipy.linalg

import oracle
from tests._tensors import mttkrp_numpy, uniform_tensor, zipf_tensor

L = oracle.lib()


def brute_pext(x, m):
    out, b = 0, 0
    for i in range(128):
        if (m >> i) & 1:
            if (x >> i) & 1:
                out |= 1 << b
            b += 1
    return out


def brute_pdep(x, m):
    out, b = 0, 0
    for i in range(128):
        if (m >> i) & 1:
            if (x >> b) & 1:
                out |= 1 << i
            b += 1
    return out


def test_pdep_pext_64_bruteforce():
    rng = np.random.default_rng(0)
    for _ in range(3000):
        x = int(rng.integers(0, 2**63, dtype=np.uint64)) * 2 + int(rng.integers(0, 2))
        m = int(rng.integers(0, 2**63, dtype=np.uint64)) * 2 + int(rng.integers(0, 2))
        if rng.random() < 0.3:
            m &= int(rng.integers(0, 2**63, dtype=np.uint64))
        assert L.or_pext64(x, m) == brute_pext(x, m)
        assert L.or_pdep64(x, m) == brute_pdep(x, m)
        assert L.or_pext64(L.or_pdep64(x, m), m) == x & ((1 << bin(m).count("1")) - 1)


def test_pdep_pext_128_split():
    """u128 forms split at bit 64 with popcount(low mask) (include/bitops.hpp:60-79)."""
    rng = np.random.default_rng(1)
    for _ in range(2000):
        m = int.from_bytes(rng.bytes(16), "little")
        if rng.random() < 0.5:
            m &= int.from_bytes(rng.bytes(16), "little")
        x = int(rng.integers(0, 2**63, dtype=np.uint64))
        mw = np.array([m & (2**64 - 1), m >> 64], np.uint64)
        out = np.zeros(2, np.uint64)
        L.or_pdep128(x, mw, out)
        dep = int(out[0]) | (int(out[1]) << 64)
        assert dep == brute_pdep(x, m) & (2**128 - 1)
        kw = np.array([dep & (2**64 - 1), dep >> 64], np.uint64)
        assert L.or_pext128(kw, mw) == brute_pext(dep, m) & (2**64 - 1)


def test_mask_worked_example():
    """dims (4,8,2): masks (0xA, 0x34, 0x1); (3,5,1) encodes to 0x2F."""
    masks, am, bits = oracle.setup_masks([4, 8, 2], 1)
    assert [int(x) for x in masks[:, 0]] == [0xA, 0x34, 0x1]
    assert bits == 6 and int(am[0]) == 0x3F
    key = 0
    for c, m in zip([3, 5, 1], masks[:, 0]):
        key |= L.or_pdep64(c, int(m))
    assert key == 0x2F


def test_mask_nell2_shape():
    """12092 x 9184 x 28818 (14+14+15 bits)."""
    masks, am, bits = oracle.setup_masks([12092, 9184, 28818], 1)
    assert bits == 43
    assert [int(x) for x in masks[:, 0]] == [0x9249249249, 0x12492492492, 0x64924924924]


@pytest.mark.parametrize("dims,W", [([1000, 1000, 1000], 1), ([2**20, 2**18, 2**16, 1024], 2),
                                    ([50000, 50000, 10000, 2000, 512], 2), ([3, 70000, 5], 1),
                                    ([2**40, 2**30], 2), ([7], 1)])
def test_mask_invariants(dims, W):
    masks, am, bits = oracle.setup_masks(dims, W)
    ints = [sum(int(w) << (64 * k) for k, w in enumerate(row)) for row in masks]
    tot = 0
    for i, (a, d) in enumerate(zip(ints, dims)):
        assert bin(a).count("1") == max(int(d - 1).bit_length(), 0)
        for b in ints[i + 1:]:
            assert a & b == 0
        tot |= a
    assert tot == (1 << bits) - 1


def _ref_cases():
    return [
        dict(dims=[40, 50, 60], nnz=3000, P=8, W=1, seed=1),
        dict(dims=[13, 1000, 7], nnz=2500, P=4, W=1, seed=2),
        dict(dims=[5000, 4000, 3000], nnz=4000, P=8, W=1, seed=3),   # low reuse: oidx + CR bits
        dict(dims=[30, 20, 10, 40], nnz=5000, P=6, W=2, seed=4),
        dict(dims=[2**14, 2**13, 2**13, 2**12, 2**12], nnz=6000, P=8, W=2, seed=5),  # 66 bits
        dict(dims=[100, 90, 80, 70, 60], nnz=4000, P=3, W=1, seed=6),
    ]


@pytest.mark.skipif(not oracle.ref_available(1), reason="reference not built (oracle/_ref)")
@pytest.mark.parametrize("case", _ref_cases())
def test_create_matches_reference(case):
    if not oracle.ref_available(case["W"]):
        pytest.skip("reference build missing")
    cidx, vals = uniform_tensor(case["dims"], case["nnz"], case["seed"])
    o = oracle.OracleAlto(case["dims"], cidx, vals, W=case["W"], nprtn=case["P"]).state()
    r = oracle.RefAlto(case["dims"], cidx, vals, W=case["W"], nprtn=case["P"]).state()
    np.testing.assert_array_equal(o["masks"], r["masks"])
    np.testing.assert_array_equal(o["idx"], r["idx"])
    np.testing.assert_array_equal(o["vals"], r["vals"])
    np.testing.assert_array_equal(o["prtn_ptr"], r["prtn_ptr"])
    np.testing.assert_array_equal(o["intervals"], r["intervals"])
    np.testing.assert_array_equal(o["cr_masks"], r["cr_masks"])
    assert o["is_oidx"] == r["is_oidx"]
    # oidx tie order is unspecified in the reference: compare the row sequence and per-row sets
    for n, oi in o["oidx"].items():
        ri = r["oidx"][n]
        keys = np.array([int(k[0]) if case["W"] == 1 else (int(k[0]) | (int(k[1]) << 64)) for k in o["idx"]])
        m = sum(int(w) << (64 * k) for k, w in enumerate(o["masks"][n]))
        rows = np.array([brute_pext(int(k), m) for k in keys])
        np.testing.assert_array_equal(rows[oi.astype(np.int64)], rows[ri.astype(np.int64)])
        np.testing.assert_array_equal(np.sort(oi), np.sort(ri))


@pytest.mark.parametrize("case", _ref_cases())
def test_delinearize_roundtrip(case):
    cidx, vals = uniform_tensor(case["dims"], case["nnz"], case["seed"])
    at = oracle.OracleAlto(case["dims"], cidx, vals, W=case["W"], nprtn=case["P"])
    st = at.state()
    coords = at.delinearize()
    # every input coordinate appears exactly once, in ascending key order
    a = {tuple(int(x) for x in col) for col in cidx.T}
    b = {tuple(int(x) for x in col) for col in coords.T}
    assert a == b
    keys = [sum(int(w) << (64 * k) for k, w in enumerate(row)) for row in st["idx"]]
    am = sum(int(w) << (64 * k) for k, w in enumerate(st["alto_mask"]))
    stripped = [k & am for k in keys]
    assert all(x < y for x, y in zip(stripped, stripped[1:]))


@pytest.mark.parametrize("case", _ref_cases())
def test_mttkrp_variants_bruteforce(case):
    cidx, vals = uniform_tensor(case["dims"], case["nnz"], case["seed"])
    at = oracle.OracleAlto(case["dims"], cidx, vals, W=case["W"], nprtn=case["P"])
    rng = np.random.default_rng(7)
    R = 10
    U = [rng.random((d, R)) for d in case["dims"]]
    for mode in range(len(case["dims"])):
        ref = mttkrp_numpy(cidx, vals, U, mode)
        for variant in range(6):
            out = at.mttkrp(mode, U, variant)
            np.testing.assert_allclose(out, ref, rtol=1e-12, atol=1e-12)
        np.testing.assert_allclose(oracle.mttkrp_coo(case["dims"], cidx, vals, U, mode), ref, rtol=1e-12, atol=1e-13)


@pytest.mark.skipif(not oracle.ref_available(1), reason="reference not built (oracle/_ref)")
@pytest.mark.parametrize("case", _ref_cases())
@pytest.mark.parametrize("R", [8, 16, 10])
def test_mttkrp_and_phi_match_reference(case, R):
    if not oracle.ref_available(case["W"]):
        pytest.skip("reference build missing")
    cidx, vals = uniform_tensor(case["dims"], case["nnz"], case["seed"])
    at = oracle.OracleAlto(case["dims"], cidx, vals, W=case["W"], nprtn=case["P"])
    rf = oracle.RefAlto(case["dims"], cidx, vals, W=case["W"], nprtn=case["P"])
    rng = np.random.default_rng(8)
    U = [rng.random((d, R)) + 0.1 for d in case["dims"]]
    for mode in range(len(case["dims"])):
        a = at.mttkrp(mode, U)
        np.testing.assert_allclose(a, rf.mttkrp(mode, U), rtol=1e-12, atol=1e-13)
        np.testing.assert_allclose(a, rf.mttkrp(mode, U, seq=True), rtol=1e-12, atol=1e-13)
        # Phi: rows without nonzeros are 0 in the oracle; the reference's op_sched
        # variant leaves them untouched (here: the zeroed buffer), so both agree.
        np.testing.assert_allclose(at.phi(mode, U), rf.phi(mode, U), rtol=1e-12, atol=1e-13)


def test_phi_bruteforce():
    dims = [30, 40, 20]
    cidx, vals = uniform_tensor(dims, 2000, 11)
    at = oracle.OracleAlto(dims, cidx, vals, nprtn=5)
    rng = np.random.default_rng(3)
    R = 6
    U = [rng.random((d, R)) for d in dims]
    for mode in range(3):
        pi = np.ones((len(vals), R))
        for m in range(3):
            if m != mode:
                pi *= U[m][cidx[m].astype(np.int64)]
        v = np.sum(U[mode][cidx[mode].astype(np.int64)] * pi, axis=1)
        w = vals / np.maximum(v, 1e-10)
        ref = np.zeros_like(U[mode])
        np.add.at(ref, cidx[mode].astype(np.int64), w[:, None] * pi)
        np.testing.assert_allclose(at.phi(mode, U), ref, rtol=1e-12)


def test_random_init_is_glibc_rand():
    libc = ctypes.CDLL("libc.so.6")
    U, lam = oracle.random_init([3, 4], 2, 1234)
    libc.srand(1234)
    seq = [libc.rand() / 2147483647 for _ in range(14)]
    np.testing.assert_array_equal(np.concatenate([U[0].ravel(), U[1].ravel()]), seq)
    np.testing.assert_array_equal(lam, [1.0, 1.0])


def _pi(grams, mode, U, dim):
    gp = oracle.ptrs(grams)
    U = np.ascontiguousarray(U).copy()
    fb = L.or_pseudo_inverse(len(grams), U.shape[1], gp, U, dim, mode)
    return U, fb


def test_gram_and_pseudo_inverse_cholesky():
    rng = np.random.default_rng(4)
    R = 12
    Us = [rng.random((d, R)) for d in (50, 60, 70)]
    grams = []
    for U in Us:
        G = np.zeros((R, R))
        L.or_gram(U, U.shape[0], R, G)
        np.testing.assert_allclose(G, U.T @ U, rtol=1e-13)
        grams.append(G)
    M = rng.random((60, R))
    out, fb = _pi(grams, 1, M, 60)
    assert fb == 0
    V = grams[0] * grams[2]
    ref = scipy.linalg.cho_solve(scipy.linalg.cho_factor(V, lower=True), M.T).T
    np.testing.assert_allclose(out, ref, rtol=1e-9)


def test_pseudo_inverse_fallback_min_norm():
    """A zero factor column makes V singular: Cholesky fails, the minimum-norm
    least-squares solution (LAPACK gelsy on the symmetric V) is returned."""
    rng = np.random.default_rng(5)
    R = 6
    Us = [rng.random((d, R)) for d in (20, 25, 30)]
    Us[0][:, 2] = 0.0
    grams = [U.T @ U for U in Us]
    M = rng.random((25, R))
    out, fb = _pi(grams, 1, M, 25)
    assert fb == 1
    V = grams[0] * grams[2]
    ref = scipy.linalg.lstsq(V, M.T, lapack_driver="gelsy")[0].T
    np.testing.assert_allclose(out, ref, rtol=1e-8, atol=1e-10)


def test_mat_norms():
    rng = np.random.default_rng(6)
    U = rng.normal(size=(40, 5))
    U[:, 3] = -np.abs(U[:, 3])  # all-negative column: max norm clamps to 1
    lam = np.zeros(5)
    a = U.copy()
    L.or_mat_norm(40, 5, a, lam, 0)
    np.testing.assert_allclose(lam, np.linalg.norm(U, axis=0), rtol=1e-14)
    np.testing.assert_allclose(a, U / np.linalg.norm(U, axis=0), rtol=1e-14)
    b = U.copy()
    L.or_mat_norm(40, 5, b, lam, 1)
    np.testing.assert_allclose(lam, np.maximum(1.0, U.max(axis=0)), rtol=0)
    assert lam[3] == 1.0


def test_fit_matches_dense_residual():
    dims = [6, 7, 8]
    cidx, vals = uniform_tensor(dims, 150, 9)
    rng = np.random.default_rng(10)
    R = 3
    U = [rng.random((d, R)) for d in dims]
    lam = rng.random(R) + 0.5
    X = np.zeros(dims)
    X[tuple(c.astype(np.int64) for c in cidx)] = vals
    Mk = np.einsum("r,ir,jr,kr->ijk", lam, *U)
    ref = 1.0 - np.linalg.norm(X - Mk) / np.linalg.norm(X)
    grams = [u.T @ u for u in U]
    Um = mttkrp_numpy(cidx, vals, U, 2)
    fit = L.or_cpd_fit(3, R, np.array(dims, np.uint64), oracle.ptrs(U), lam, oracle.ptrs(grams), Um,
                       float(np.sum(vals**2)))
    assert abs(fit - ref) < 1e-12


def test_cpd_als_solves_normal_equations_and_recovers_rank1():
    """Each ALS mode update solves its least-squares problem; on an exactly
    rank-1 tensor the fit goes to 1."""
    dims = [8, 9, 10]
    rng = np.random.default_rng(12)
    a, b, c = rng.random(8) + 0.5, rng.random(9) + 0.5, rng.random(10) + 0.5
    X = np.einsum("i,j,k->ijk", a, b, c)
    cidx = np.array(np.nonzero(X), dtype=np.uint64)
    vals = X[np.nonzero(X)]
    at = oracle.OracleAlto(dims, cidx, vals, nprtn=4)
    U, lam = oracle.random_init(dims, 1, 3)
    it, fit = at.cpd_als(U, lam, 30, 1e-12)
    # the fit goes through sqrt(normX + normU - 2 inner): at an exact fit the
    # rounding noise of that difference (~1e-16 normX) shows as ~1e-8
    assert fit[-1] > 1 - 1e-6
    Mk = lam[0] * np.einsum("i,j,k->ijk", U[0][:, 0], U[1][:, 0], U[2][:, 0])
    assert np.linalg.norm(Mk - X) < 1e-10 * np.linalg.norm(X)
    # one more mode-0 update from the current state must leave U0*diag(lam) unchanged
    dims2 = [6, 7, 5]
    cidx, vals = uniform_tensor(dims2, 120, 13)
    at = oracle.OracleAlto(dims2, cidx, vals, nprtn=2)
    U, lam = oracle.random_init(dims2, 3, 5)
    at.cpd_als(U, lam, 1, 0.0)
    X = np.zeros(dims2)
    X[tuple(x.astype(np.int64) for x in cidx)] = vals
    # normal equations for the last updated mode (2): X_(2) KR = U2 diag(lam) (KR^T KR)
    KR = np.einsum("ir,jr->ijr", U[0], U[1]).reshape(-1, 3)
    lhs = X.transpose(2, 0, 1).reshape(5, -1) @ KR
    rhs = (U[2] * lam) @ (KR.T @ KR)
    np.testing.assert_allclose(lhs, rhs, rtol=1e-8, atol=1e-10)


def test_cp_apr_loglik_and_monotone():
    """LL matches the dense Poisson log-likelihood; more outer iterations do not decrease it."""
    dims = [7, 6, 5]
    rng = np.random.default_rng(14)
    cidx, _ = uniform_tensor(dims, 80, 15)
    vals = rng.integers(1, 6, size=80).astype(np.float64)
    lls = []
    for iters in (1, 3, 6):
        at = oracle.OracleAlto(dims, cidx, vals, nprtn=2, val_int=True)
        U, lam = oracle.random_init(dims, 3, 7)
        res = at.cp_apr(U, lam, iters, 5, eps=0.0, dobench=True)
        M = np.einsum("r,ir,jr,kr->ijk", lam, *U)
        x = np.zeros(dims)
        x[tuple(c.astype(np.int64) for c in cidx)] = vals
        ll = np.sum(x[x > 0] * np.log(M[x > 0])) - M.sum()
        assert abs(res["loglik"] - ll) < 1e-8 * max(1.0, abs(ll))
        assert res["outer_iters"] == iters and res["total_inner"] == iters * 3 * 5
        lls.append(res["loglik"])
    assert lls[0] <= lls[1] + 1e-9 <= lls[2] + 2e-9


def test_zipf_generator_shape():
    cidx, vals = zipf_tensor([300, 200, 500], 5000, 3)
    assert cidx.shape == (3, 5000) and np.all(vals > 0)
    assert len({tuple(c) for c in cidx.T.tolist()}) == 5000


def test_poisson_oracle_mt19937_64_known_answer():
    """The oracle generator's MT19937-64 is the published one: the C++
    standard fixes the 10000th draw of std::mt19937_64 from its default seed
    5489 at 9981545732273789042; its uniform draws lie in (0, 1)."""
    from oracle.poisson_gen import MT64, Streams, binning
    g = MT64(5489)
    for _ in range(9999):
        g()
    assert g() == 9981545732273789042
    u = Streams(2).uniform(1001, 0.0, 1.0)
    assert all(0.0 < x < 1.0 for x in u)
    # binning: the last edge below the draw; a draw on an edge takes the bin below
    bins = [0.0, 0.25, 0.5, 1.0]
    assert [binning(x, bins) for x in (0.1, 0.25, 0.3, 0.75, 1.0)] == [0, 1, 1, 2, 2]


@pytest.mark.skipif(not oracle.ref_available(1), reason="reference harness not built")
@pytest.mark.parametrize("dims,nnz,W", [([300, 200, 100], 20000, 1), ([2**14, 2**13, 2**13, 2**12, 2**12], 8000, 2)])
def test_reference_wrap_equals_create(dims, nnz, W):
    """The reference's AltoTensor filled from an ALTO state (the path the
    bench uses to time the reference's MTTKRP on the GPU's export) gives the
    MTTKRPs of the one create_alto built."""
    if not oracle.ref_available(W):
        pytest.skip("reference harness not built")
    cidx, vals = uniform_tensor(dims, nnz, 13)
    r = oracle.RefAlto(dims, cidx, vals, W=W, nprtn=8)
    st = r.state()
    masks, cr = st["masks"], st["cr_masks"]
    alto_mask = np.bitwise_or.reduce(masks, axis=0)
    alto_cr = np.bitwise_or.reduce(cr, axis=0)
    w = oracle.RefAlto.wrap(dims, W, 8, masks, alto_mask, st["idx"], st["vals"], st["prtn_ptr"], st["intervals"], cr,
                            alto_cr)
    rng = np.random.default_rng(2)
    U = [rng.random((d, 16)) for d in dims]
    for mode in range(len(dims)):
        np.testing.assert_allclose(w.mttkrp(mode, U), r.mttkrp(mode, U), rtol=1e-13, atol=1e-13)


@pytest.mark.parametrize("dims,nnz,R", [([40, 50, 60], 6000, 8), ([30, 20, 10, 40], 8000, 16),
                                        ([100, 90, 80, 70, 60], 20000, 10)])
def test_cpd_als_matches_coo_cpd(dims, nnz, R):
    """The oracle's ALTO CP-ALS equals the reference's COO cpd restated with
    numpy / scipy (COO MTTKRP, LAPACK Cholesky solve): fit trace within 1e-10,
    factors and lambda within 1e-8."""
    cidx, vals = uniform_tensor(dims, nnz, 5)
    o = oracle.OracleAlto(dims, cidx, vals, nprtn=8)
    U1, l1 = oracle.random_init(dims, R, 3)
    U2, l2 = oracle.random_init(dims, R, 3)
    n1, f1 = o.cpd_als(U1, l1, 10, 0.0)
    n2, f2 = oracle.cpd_coo(dims, cidx, vals, U2, l2, 10, 0.0)
    assert n1 == n2 == 10
    np.testing.assert_allclose(f1, f2, rtol=1e-10, atol=1e-12)
    np.testing.assert_allclose(l1, l2, rtol=1e-8)
    for a, b in zip(U1, U2):
        np.testing.assert_allclose(a, b, rtol=1e-7, atol=1e-9 * np.abs(b).max())
