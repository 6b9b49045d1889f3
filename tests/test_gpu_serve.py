"""Serving: each job's result is the same alone and in a concurrent mix
(FP64 fit within 1e-10 relative, FP32 within 1e-5, BF16 fit delta within 1e-3:
the atomic summation order differs between runs), the
admission loop respects the HBM budget, and a failed job is isolated."""
import numpy as np
import pytest

from tests._tensors import zipf_tensor

pytestmark = pytest.mark.gpu


def _jobs(n, seed=0):
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n):
        nm = int(rng.integers(3, 6))
        dims = [int(rng.integers(40, 400)) for _ in range(nm)]
        nnz = int(rng.integers(2000, 30000))
        cidx, vals = zipf_tensor(dims, nnz, 100 + k)
        out.append(dict(dims=dims, cidx=cidx, vals=vals, rank=int(rng.choice([8, 16, 32])), seed=7 + k))
    return out


@pytest.mark.parametrize("lockstep_nnz", [0, 15000, 1 << 30])
@pytest.mark.parametrize("prec,tol", [(0, 1e-10), (1, 1e-5), (2, 1e-3)])
def test_alone_equals_in_mix(prec, tol, lockstep_nnz):
    """lockstep_nnz: none, part or all of the mix iterated in lockstep groups."""
    from alto_amd.serve import Server
    jobs = _jobs(12)
    alone = []
    s1 = Server(8 << 30, 1)
    for j in jobs:
        jid = s1.submit(j["dims"], list(j["cidx"]), j["vals"], j["rank"], 15, 1e-5, j["seed"], prec, want_factors=True)
        r, U, lam = s1.result(jid, True, j["dims"], j["rank"])
        assert r.status == 2
        alone.append((r.fit, r.iters, U))
    s1.close()
    s8 = Server(8 << 30, 8, lockstep_nnz=lockstep_nnz)
    ids = [s8.submit(j["dims"], list(j["cidx"]), j["vals"], j["rank"], 15, 1e-5, j["seed"], prec, want_factors=True)
           for j in jobs]
    for j, jid, (fit, iters, U) in zip(jobs, ids, alone):
        r, U2, lam2 = s8.result(jid, True, j["dims"], j["rank"])
        assert r.status == 2
        assert abs(r.fit - fit) <= tol * max(1.0, abs(fit))
        if prec == 0:
            assert r.iters == iters
            for a, b in zip(U, U2):
                np.testing.assert_allclose(a, b, rtol=1e-7, atol=1e-9)
    s8.close()


@pytest.mark.parametrize("lockstep_nnz", [0, 1 << 30])
def test_budget_admission_and_error_isolation(lockstep_nnz):
    from alto_amd.serve import Server
    jobs = _jobs(6, seed=1)
    # tiny budget: jobs are admitted one at a time
    s = Server(200 << 20, 4, lockstep_nnz=lockstep_nnz)
    ids = [s.submit(j["dims"], list(j["cidx"]), j["vals"], j["rank"], 5, 0.0, j["seed"]) for j in jobs]
    # a job whose coordinates exceed its dims fails alone
    bad = jobs[0]
    bad_c = [c.copy() for c in bad["cidx"]]
    bad_c[0][0] = bad["dims"][0] + 5
    bid = s.submit(bad["dims"], bad_c, bad["vals"], bad["rank"], 5, 0.0, 1)
    for jid in ids:
        assert s.result(jid)[0].status == 2
    assert s.result(bid)[0].status < 0
    st = s.stats()
    assert st.hbm_peak <= max(st.hbm_budget, max(s.result(i)[0].footprint for i in ids))
    # finished and failed jobs can both be released; a released id is gone
    from alto_amd import AltoError
    assert s.L.alto_server_release(s.h, bid) == 0
    assert s.L.alto_server_release(s.h, ids[0]) == 0
    with pytest.raises(AltoError):
        s.result(bid)
    s.close()


def test_sequential_runner_equals_server():
    """The one-job-at-a-time runner (device-seeded sessions on the calling
    thread) and the server with 4 slots give every job of a seeded stream
    the same fit (FP64, within 1e-10 relative)."""
    from alto_amd.serve import Server, run_sequential
    jobs = _jobs(10, seed=2)
    seq = run_sequential(jobs, 0, max_iters=15, epsilon=1e-5)
    s4 = Server(8 << 30, 4)
    ids = [s4.submit(j["dims"], list(j["cidx"]), j["vals"], j["rank"], 15, 1e-5, j["seed"], 0) for j in jobs]
    fits = []
    for jid in ids:
        r = s4.result(jid)[0]
        assert r.status == 2
        fits.append(r.fit)
    s4.close()
    assert seq["jobs"] == 10
    np.testing.assert_allclose(seq["fits"], fits, rtol=1e-10, atol=1e-12)


def test_lockstep_groups_with_mixed_limits():
    """Jobs of several ranks (64 joins a group, 72 does not), tolerances and
    iteration limits through the lockstep lane, with fewer lockstep slots
    than jobs: each stops where it stops alone, with the same fit."""
    from alto_amd.serve import Server
    rng = np.random.default_rng(21)
    jobs = []
    for k in range(20):
        dims = [int(rng.integers(30, 200)) for _ in range(int(rng.integers(3, 6)))]
        cidx, vals = zipf_tensor(dims, int(rng.integers(2000, 20000)), 300 + k)
        jobs.append(dict(dims=dims, cidx=cidx, vals=vals, rank=int(rng.choice([8, 16, 64, 72])), seed=40 + k,
                         iters=int(rng.integers(3, 25)), eps=float(rng.choice([0.0, 1e-4, 1e-6]))))
    alone = []
    s1 = Server(8 << 30, 1)
    for j in jobs:
        jid = s1.submit(j["dims"], list(j["cidx"]), j["vals"], j["rank"], j["iters"], j["eps"], j["seed"])
        r, _, _ = s1.result(jid, True)
        assert r.status == 2
        alone.append(r)
    s1.close()
    s = Server(8 << 30, 4, lockstep_nnz=1 << 30, lockstep_sessions=6)
    ids = [s.submit(j["dims"], list(j["cidx"]), j["vals"], j["rank"], j["iters"], j["eps"], j["seed"]) for j in jobs]
    for j, jid, a in zip(jobs, ids, alone):
        r, _, _ = s.result(jid, True)
        assert r.status == 2
        assert r.iters == a.iters, (j["rank"], j["iters"], j["eps"])
        assert abs(r.fit - a.fit) <= 1e-10 * max(1.0, abs(a.fit))
    s.close()
