"""Pins the oracle's CP-APR driver (include/apr.hpp:1100-1290,
src/kruskal_model.cpp:128-222,355-378) by checks a plausible mistake fails.

The driver's steps are observed through the oracle's hook at real states
and compared with what the mathematics fixes:
- Phi against a brute-force sum from its definition, the KKT value against
  max |min(B, 1 - Phi)|, the multiplicative update against B * Phi;
- the kappa fix moves exactly the entries with B < kappa_tol and Phi > 1,
  by exactly kappa;
- mode normalisation leaves unit L1 column norms and the model unchanged,
  redistribution leaves lambda = 1 and the model unchanged;
- an exact low-rank count tensor started at its own factors is a fixed
  point (Phi = 1): one inner step per mode and the eps^2 outer stop at
  iteration 1;
- the trajectory equals an independent restatement of the reference's
  COO CP-APR (src/apr.cpp:239-488).
Oracle builds with a planted mistake (fmax for fmin, the kappa condition
reversed, 2-norm normalisation, the outer stop reversed) must fail them.
"""
import numpy as np
import pytest

import oracle

EPS_DIV_ZERO = 1e-10


def count_tensor(dims, R, seed, mass=400.0):
    """Poisson counts of a random rank-R Kruskal model, duplicates merged."""
    rng = np.random.default_rng(seed)
    U = [rng.gamma(0.5, 1.0, size=(d, R)) for d in dims]
    U = [u / u.sum(axis=0) for u in U]
    lam = rng.uniform(0.5, 1.5, R)
    lam *= mass / lam.sum()
    rows = []
    for r in range(R):
        k = rng.poisson(lam[r])
        rows.append(np.stack([rng.choice(d, size=k, p=U[n][:, r]) for n, d in enumerate(dims)], axis=1))
    rows = np.concatenate(rows)
    uniq, cnt = np.unique(rows, axis=0, return_counts=True)
    return np.ascontiguousarray(uniq.T.astype(np.uint64)), cnt.astype(np.float64)


def init_model(dims, R, seed):
    rng = np.random.default_rng(seed)
    return [rng.uniform(0.05, 1.0, size=(d, R)) for d in dims], np.ones(R)


def phi_bruteforce(cidx, vals, U, mode, eps=EPS_DIV_ZERO):
    """Phi_n[i,:] = sum over nonzeros x in row i of v_x / max(<B[i,:], pi_x>, eps) * pi_x."""
    R = U[0].shape[1]
    out = np.zeros_like(U[mode])
    for x in range(len(vals)):
        pi = np.ones(R)
        for m in range(len(U)):
            if m != mode:
                pi = pi * U[m][int(cidx[m][x])]
        i = int(cidx[mode][x])
        out[i] += vals[x] / max(float(U[mode][i] @ pi), eps) * pi
    return out


DIMS, RANK = (9, 11, 7), 3


@pytest.fixture(scope="module")
def counts():
    return count_tensor(DIMS, RANK, seed=21)


def run_trace(L, cidx, vals, kappa_tol=1e-10, kappa=1e-2, max_iters=4, max_inner=3, eps=1e-4, dobench=False,
              dims=DIMS, R=RANK, seed=5):
    at = oracle.OracleAlto(dims, cidx, vals, W=1, nprtn=2, L=L)
    U, lam = init_model(dims, R, seed)
    return at.cp_apr_trace(U, lam, max_iters, max_inner, eps=eps, eps_div_zero=EPS_DIV_ZERO, kappa_tol=kappa_tol,
                           kappa=kappa, dobench=dobench)


# ---- checks shared by the default oracle and the planted-mistake builds ----

def check_phi_kkt_mu(L, cidx, vals):
    res, ev = run_trace(L, cidx, vals, max_iters=3, max_inner=3, dobench=True)
    phis = [e for e in ev if e["ev"] == oracle.APR_EV_PHI]
    assert len(phis) == 3 * len(DIMS) * 3
    for e in phis:
        B = e["U"][e["mode"]]
        bf = phi_bruteforce(cidx, vals, e["U"], e["mode"])
        np.testing.assert_allclose(e["phi"], bf, rtol=1e-12, atol=1e-15)
        assert e["kkt"] == pytest.approx(float(np.abs(np.minimum(B, 1.0 - bf)).max()), rel=1e-12, abs=1e-15)
    for k, e in enumerate(ev):
        if e["ev"] == oracle.APR_EV_MU:
            p = ev[k - 1]
            assert p["ev"] == oracle.APR_EV_PHI and p["mode"] == e["mode"]
            np.testing.assert_array_equal(e["U"][e["mode"]], p["U"][p["mode"]] * p["phi"])
    outer = [e["kkt"] for e in ev if e["ev"] == oracle.APR_EV_OUTER]
    for it, k in enumerate(outer):
        per_mode = [e["kkt"] for e in ev if e["ev"] == oracle.APR_EV_NORM and e["iter"] == it]
        assert k == max(per_mode)
    assert res["kkt"] == outer[-1]


def check_kappa_fix(L, cidx, vals):
    kappa_tol, kappa = 0.08, 0.01
    _, ev = run_trace(L, cidx, vals, kappa_tol=kappa_tol, kappa=kappa, max_iters=4, max_inner=2)
    pairs = 0
    seen = np.zeros(3, dtype=bool)  # flagged / small B with Phi <= 1 / large B with Phi > 1
    for k, e in enumerate(ev):
        if e["ev"] != oracle.APR_EV_PRE_FIX:
            continue
        post = ev[k + 1]
        assert post["ev"] == oracle.APR_EV_POST_FIX
        n = e["mode"]
        pre_b, post_b, phi = e["U"][n], post["U"][n], e["phi"]
        flagged = (pre_b < kappa_tol) & (phi > 1.0)
        np.testing.assert_array_equal(post_b != pre_b, flagged)
        np.testing.assert_array_equal(post_b[flagged], pre_b[flagged] + kappa)
        seen |= [flagged.any(), ((pre_b < kappa_tol) & (phi <= 1.0)).any(), ((pre_b >= kappa_tol) & (phi > 1.0)).any()]
        pairs += 1
    assert pairs == 3 * len(DIMS)
    assert seen.all(), "the state does not discriminate the kappa condition"


def check_normalize_redistribute(L, cidx, vals):
    _, ev = run_trace(L, cidx, vals, max_iters=3, max_inner=2, eps=1e-4)
    start = ev[0]
    assert start["ev"] == oracle.APR_EV_START
    for U in start["U"]:
        np.testing.assert_allclose(np.abs(U).sum(axis=0), 1.0, rtol=1e-13)
    last_b = {}
    nred = 0
    for k, e in enumerate(ev):
        n = e["mode"]
        if e["ev"] == oracle.APR_EV_PHI:
            if e["inner"] == 0:  # RedistributeLambda: U_n * lambda, lambda = 1
                p = ev[k - 1]  # the model just before (kappa fix, previous mode's normalisation or start)
                assert p["ev"] in (oracle.APR_EV_POST_FIX, oracle.APR_EV_NORM, oracle.APR_EV_START)
                np.testing.assert_array_equal(e["lam"], 1.0)
                np.testing.assert_array_equal(e["U"][n], p["U"][n] * p["lam"])
                for m in range(len(DIMS)):
                    if m != n:
                        np.testing.assert_array_equal(e["U"][m], p["U"][m])
                nred += 1
            last_b[n] = e["U"][n]
        elif e["ev"] == oracle.APR_EV_MU:
            last_b[n] = e["U"][n]
        elif e["ev"] == oracle.APR_EV_NORM:  # NormalizeMode: unit L1 columns, model unchanged
            np.testing.assert_allclose(np.abs(e["U"][n]).sum(axis=0), 1.0, rtol=1e-13)
            np.testing.assert_allclose(e["U"][n] * e["lam"], last_b[n], rtol=1e-13)
    assert nred == 3 * len(DIMS)


def exact_count_model():
    a = np.array([[1, 2], [3, 1], [2, 2]], float)
    b = np.array([[2, 1], [1, 1], [1, 3], [2, 2]], float)
    c = np.array([[1, 1], [2, 1], [1, 2], [3, 1], [1, 1]], float)
    lam = np.array([1.0, 2.0])
    X = np.einsum("r,ir,jr,kr->ijk", lam, a, b, c)
    idx = np.argwhere(X > 0)
    return [a, b, c], lam, np.ascontiguousarray(idx.T.astype(np.uint64)), X[X > 0].astype(np.float64)


def check_fixed_point_stop(L):
    U, lam, cidx, vals = exact_count_model()
    dims = [u.shape[0] for u in U]
    at = oracle.OracleAlto(dims, cidx, vals, W=1, nprtn=2, L=L)
    res, ev = at.cp_apr_trace([u.copy() for u in U], lam.copy(), 10, 10, eps=1e-4, eps_div_zero=EPS_DIV_ZERO,
                              kappa_tol=1e-10, kappa=1e-2, dobench=False)
    for e in ev:
        if e["ev"] == oracle.APR_EV_PHI:
            np.testing.assert_allclose(e["phi"], 1.0, rtol=0, atol=1e-12)  # Phi = 1 at the fixed point
    assert res["outer_iters"] == 1
    assert res["total_inner"] == 3
    assert res["inner_per_mode"][:3] == [2, 2, 2]
    assert res["kkt"] < 1e-12
    # log-likelihood of the model that generated X, written out densely
    model = np.einsum("r,ir,jr,kr->ijk", lam, *U)
    assert res["loglik"] == pytest.approx(float((vals * np.log(model[tuple(cidx.astype(np.int64))])).sum()
                                                - model.sum()), rel=1e-12)


def check_coo_trajectory(L, cidx, vals):
    """dobench runs every inner step in both drivers, so they do the same arithmetic."""
    res, ev = run_trace(L, cidx, vals, max_iters=6, max_inner=4, eps=0.0, dobench=True, kappa_tol=0.05)
    U, lam = init_model(DIMS, RANK, 5)
    coo = oracle.cp_apr_coo(DIMS, cidx, vals, U, lam, 6, 4, eps=0.0, eps_div_zero=EPS_DIV_ZERO, kappa_tol=0.05,
                            kappa=1e-2)
    outer = np.array([e["kkt"] for e in ev if e["ev"] == oracle.APR_EV_OUTER])
    np.testing.assert_allclose(outer, coo["kkt"], rtol=1e-9)
    assert res["loglik"] == pytest.approx(coo["loglik"], rel=1e-11)
    assert res["ls_fit"] == pytest.approx(coo["ls_fit"], rel=1e-9, abs=1e-12)


# ---- the checks on the oracle as committed ---------------------------------

def test_apr_phi_kkt_and_update_match_definitions(counts):
    check_phi_kkt_mu(oracle.lib(), *counts)


def test_apr_kappa_fix_moves_exactly_flagged_entries(counts):
    check_kappa_fix(oracle.lib(), *counts)


def test_apr_normalize_and_redistribute_keep_the_model(counts):
    check_normalize_redistribute(oracle.lib(), *counts)


def test_l1_helpers_closed_forms():
    rng = np.random.default_rng(3)
    U = rng.normal(size=(13, 4))
    lam = rng.uniform(0.5, 2.0, 4)
    U1, lam1 = U.copy(), lam.copy()
    oracle.lib().or_normalize_mode_l1(13, 4, U1, lam1)
    np.testing.assert_allclose(np.abs(U1).sum(axis=0), 1.0, rtol=1e-14)
    np.testing.assert_allclose(lam1, lam * np.abs(U).sum(axis=0), rtol=1e-14)
    np.testing.assert_allclose(U1 * lam1, U * lam, rtol=1e-14)
    U2, lam2 = U.copy(), lam.copy()
    oracle.lib().or_redistribute_lambda(13, 4, U2, lam2)
    np.testing.assert_array_equal(lam2, 1.0)
    np.testing.assert_array_equal(U2, U * lam)


def test_apr_fixed_point_stops_after_one_inner_step():
    check_fixed_point_stop(oracle.lib())


def test_apr_trajectory_matches_coo_restatement(counts):
    check_coo_trajectory(oracle.lib(), *counts)


def test_apr_trajectories_agree_until_a_stop(counts):
    """With a KKT tolerance both drivers break their inner loops by the same
    rule; the outer KKT values agree for as long as both run."""
    cidx, vals = counts
    res, ev = run_trace(oracle.lib(), cidx, vals, max_iters=8, max_inner=5, eps=1e-3)
    U, lam = init_model(DIMS, RANK, 5)
    coo = oracle.cp_apr_coo(DIMS, cidx, vals, U, lam, 8, 5, eps=1e-3, eps_div_zero=EPS_DIV_ZERO)
    outer = np.array([e["kkt"] for e in ev if e["ev"] == oracle.APR_EV_OUTER])
    k = min(len(outer), len(coo["kkt"]))
    assert k >= 2
    np.testing.assert_allclose(outer[:k], coo["kkt"][:k], rtol=1e-9)


MUTANT_CHECKS = {
    1: ["phi_kkt_mu", "fixed_point"],      # fmax(B, 1 - Phi) in the KKT value
    2: ["kappa"],                          # kappa fix on Phi < 1
    3: ["normalize", "coo"],               # 2-norm mode normalisation
    4: ["fixed_point"],                    # outer stop reversed
}


@pytest.mark.parametrize("k", sorted(MUTANT_CHECKS))
def test_planted_mistakes_are_caught(k, counts, tmp_path):
    L = oracle.load(oracle.build_mutant(k, str(tmp_path / f"liboracle_mutant{k}.so")))
    cidx, vals = counts
    checks = {
        "phi_kkt_mu": lambda: check_phi_kkt_mu(L, cidx, vals),
        "kappa": lambda: check_kappa_fix(L, cidx, vals),
        "normalize": lambda: check_normalize_redistribute(L, cidx, vals),
        "fixed_point": lambda: check_fixed_point_stop(L),
        "coo": lambda: check_coo_trajectory(L, cidx, vals),
    }
    for name in MUTANT_CHECKS[k]:
        with pytest.raises(AssertionError):
            checks[name]()
