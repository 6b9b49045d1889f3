"""GPU parity of the CP-ALS pieces and driver through the C ABI:
Gram on MFMA vs numpy, the pseudo-inverse (Cholesky and its minimum-norm
fallback) vs scipy/LAPACK, and whole CP-ALS runs vs the CPU oracle (FP64:
fit trace within 1e-10 relative; FP32 / BF16: final fit within 1e-4 / 1e-3
of the FP64 oracle)."""
import ctypes

import numpy as np
import pytest
import scipy.linalg

import oracle
from tests._tensors import uniform_tensor, zipf_tensor

pytestmark = pytest.mark.gpu


def _A():
    import alto_amd
    return alto_amd


@pytest.mark.parametrize("R", [1, 10, 16, 32, 64, 100])
@pytest.mark.parametrize("prec", [0, 1, 2])
def test_gram_mfma(R, prec):
    A = _A()
    rng = np.random.default_rng(R + prec)
    I = 1000 + 37 * R
    U = rng.random((I, R)) - 0.3
    if prec == 0:
        d, ref = A.DeviceArray.from_numpy(U), U
    elif prec == 1:
        d, ref = A.DeviceArray.from_numpy(U.astype(np.float32)), U.astype(np.float32).astype(np.float64)
    else:
        b = A.to_bf16_bits(U)
        d, ref = A.DeviceArray.from_numpy(b), A.from_bf16_bits(b).astype(np.float64)
    G = A.DeviceArray((R, R), np.float64)
    A.check(A.lib().alto_update_gram(d.ptr, I, R, prec, G.ptr, None), "gram")
    want = ref.T @ ref
    tol = 1e-13 if prec == 0 else 2e-6
    np.testing.assert_allclose(G.numpy(), want, rtol=tol, atol=tol * np.abs(want).max())


def _pinv(grams, mode, M):
    A = _A()
    R = M.shape[1]
    g = A.DeviceArray.from_numpy(np.ascontiguousarray(np.stack(grams)))
    d = A.DeviceArray.from_numpy(M)
    fb = ctypes.c_int(-1)
    A.check(A.lib().alto_pseudo_inverse(g.ptr, len(grams), mode, R, d.ptr, M.shape[0], ctypes.byref(fb), None), "pinv")
    return d.numpy(), fb.value


@pytest.mark.parametrize("R", [5, 16, 32, 64, 100])
def test_pseudo_inverse_cholesky(R):
    rng = np.random.default_rng(R)
    Us = [rng.random((d, R)) for d in (300, 200, 250)]
    grams = [u.T @ u for u in Us]
    M = rng.random((200, R))
    out, fb = _pinv(grams, 1, M)
    assert fb == 0
    V = grams[0] * grams[2]
    want = scipy.linalg.cho_solve(scipy.linalg.cho_factor(V, lower=True), M.T).T
    np.testing.assert_allclose(out, want, rtol=1e-8, atol=1e-10 * np.abs(want).max())


@pytest.mark.parametrize("R", [8, 40, 100])
def test_pseudo_inverse_fallback_matches_gelsy(R):
    """The elimination fails on the zero pivot and the same launch computes
    the minimum-norm pseudo-inverse (256- and 1024-thread launches)."""
    rng = np.random.default_rng(9)
    Us = [rng.random((d + R, R)) for d in (30, 40, 50)]
    Us[2][:, 5] = 0.0  # V has a zero row/column: Cholesky fails
    grams = [u.T @ u for u in Us]
    M = rng.random((40, R))
    out, fb = _pinv(grams, 1, M)
    assert fb == 1
    V = grams[0] * grams[2]
    want = scipy.linalg.lstsq(V, M.T, lapack_driver="gelsy")[0].T
    np.testing.assert_allclose(out, want, rtol=1e-8, atol=1e-10)


def _als(dims, cidx, vals, R, iters, prec=0, seed=3, eps=0.0, lit_bits=0):
    A = _A()
    g = A.AltoTensor(dims, cidx, vals, nprtn=8, lit_bits=lit_bits)
    U, lam = oracle.random_init(dims, R, seed)
    n, fit, tm = g.cpd_als(U, lam, iters, eps, prec)
    return n, fit, U, lam


@pytest.mark.parametrize("dims,nnz,R,W", [([40, 50, 60], 6000, 8, 1), ([30, 20, 10, 40], 8000, 16, 1),
                                          ([100, 90, 80, 70, 60], 20000, 10, 1), ([200, 300, 250], 30000, 32, 1),
                                          ([2**14, 2**13, 2**13, 2**12, 2**12], 20000, 16, 2),
                                          ([300, 400, 500], 20000, 64, 2)])
def test_cpd_als_fp64_matches_oracle(dims, nnz, R, W):
    """W = 2: 128-bit ALTO keys (66 index bits, and a forced 128-bit mask)."""
    cidx, vals = uniform_tensor(dims, nnz, 5)
    n, fit, U, lam = _als(dims, cidx, vals, R, 12, lit_bits=64 * W)
    o = oracle.OracleAlto(dims, cidx, vals, W=W, nprtn=8)
    Uo, lo = oracle.random_init(dims, R, 3)
    no, fo = o.cpd_als(Uo, lo, 12, 0.0)
    assert n == no == 12
    np.testing.assert_allclose(fit, fo, rtol=1e-10, atol=1e-12)
    np.testing.assert_allclose(lam, lo, rtol=1e-8)
    for a, b in zip(U, Uo):
        np.testing.assert_allclose(a, b, rtol=1e-7, atol=1e-9 * np.abs(b).max())


def test_cpd_als_convergence_stop_and_rank1_recovery():
    dims = [8, 9, 10]
    rng = np.random.default_rng(12)
    X = np.einsum("i,j,k->ijk", rng.random(8) + 0.5, rng.random(9) + 0.5, rng.random(10) + 0.5)
    cidx = np.array(np.nonzero(X), dtype=np.uint64)
    vals = X[np.nonzero(X)]
    # near fit = 1 the residual sqrt(normX + normU - 2 inner) is cancellation
    # noise of ~1e-8, so the stop test uses a tolerance well above it
    n, fit, U, lam = _als(dims, cidx, vals, 1, 50, eps=1e-6)
    o = oracle.OracleAlto(dims, cidx, vals, nprtn=8)
    Uo, lo = oracle.random_init(dims, 1, 3)
    no, fo = o.cpd_als(Uo, lo, 50, 1e-6)
    # the fit itself carries that ~1e-8 noise; the recovered model must
    # reproduce X
    assert fit[-1] > 1 - 1e-6
    Mk = lam[0] * np.einsum("i,j,k->ijk", U[0][:, 0], U[1][:, 0], U[2][:, 0])
    assert np.linalg.norm(Mk - X) < 1e-10 * np.linalg.norm(X)
    assert n == no < 50


@pytest.mark.parametrize("dims,nnz,R", [([300, 200, 400], 40000, 16), ([31, 45, 27], 3000, 7)])
@pytest.mark.parametrize("prec,tol", [(1, 1e-4), (2, 1e-3)])
def test_cpd_als_reduced_precision_fit(prec, tol, dims, nnz, R):
    """Reduced-precision CP-ALS against the FP64 oracle; rank 7 with odd
    dims leaves FP32 buffers that are not a multiple of 16 bytes."""
    cidx, vals = zipf_tensor(dims, nnz, 4, s=1.1)
    n, fit, U, lam = _als(dims, cidx, vals, R, 15, prec=prec)
    o = oracle.OracleAlto(dims, cidx, vals, nprtn=8)
    Uo, lo = oracle.random_init(dims, R, 3)
    no, fo = o.cpd_als(Uo, lo, 15, 0.0)
    assert abs(fit[-1] - fo[-1]) <= tol * max(1.0, abs(fo[-1]))


def test_session_split_iterations_equal_single_run():
    A = _A()
    dims = [60, 70, 80]
    cidx, vals = uniform_tensor(dims, 8000, 21)
    g = A.AltoTensor(dims, cidx, vals, nprtn=8)
    U, lam = A.kruskal_random_init(dims, 12, 4)
    Uo, lo = oracle.random_init(dims, 12, 4)
    for a, b in zip(U, Uo):
        np.testing.assert_array_equal(a, b)
    s = A.CpdSession(g, U)
    f1, _ = s.iterate(3)
    f2, _ = s.iterate(4)
    U1, l1, fit, _ = s.get()
    n, fit_all, U2, l2 = _als(dims, cidx, vals, 12, 7, seed=4)
    np.testing.assert_allclose(np.concatenate([f1, f2]), fit_all, rtol=1e-13)
    np.testing.assert_allclose(l1, l2, rtol=1e-12)


@pytest.mark.parametrize("prec,tol", [(0, 1e-10), (1, 1e-5)])
def test_graph_replay_matches_eager(prec, tol):
    """Iterations replayed from a captured HIP graph match eager launches; the
    graph is replayed several times with eager work and syncs in between."""
    A = _A()
    dims = [70, 60, 50, 40]
    cidx, vals = uniform_tensor(dims, 12000, 31)
    g = A.AltoTensor(dims, cidx, vals, nprtn=8)
    U, _ = A.kruskal_random_init(dims, 16, 5)
    eager = A.CpdSession(g, U, precision=prec)
    fe, _ = eager.iterate(7)
    try:
        A.check(A.lib().alto_tune_set(5, 1), "tune")
        s = A.CpdSession(g, U, precision=prec)
        f1, _ = s.iterate(2)
        g.mttkrp(0, U)  # eager work between replays
        f2, _ = s.iterate(3)
        A.synchronize()
        f3, _ = s.iterate(2)
    finally:
        A.lib().alto_tune_set(5, 0)
    np.testing.assert_allclose(np.concatenate([f1, f2, f3]), fe, rtol=tol, atol=tol)
    Ue, le, _, _ = eager.get()
    Ug, lg, _, _ = s.get()
    np.testing.assert_allclose(lg, le, rtol=1e3 * tol)


def test_cpd_fit_of_returned_model_equals_last_fit():
    """cpd_fit_alto on the model CP-ALS returned (alto_cpd_fit: the inner
    product from the model's own MTTKRP) equals the fit of the last
    iteration (computed inside the solve) and the oracle's cpd_fit_alto."""
    A = _A()
    dims = [60, 50, 40]
    cidx, vals = uniform_tensor(dims, 6000, 41)
    R = 8
    n, fits, U, lam = _als(dims, cidx, vals, R, 6)
    g = A.AltoTensor(dims, cidx, vals, nprtn=8)
    f = ctypes.c_double(0)
    A.check(A.lib().alto_cpd_fit(g.h, R, A._ptr_array([u.ctypes.data for u in U]), lam.ctypes.data, ctypes.byref(f)),
            "cpd_fit")
    np.testing.assert_allclose(f.value, fits[-1], rtol=1e-9, atol=1e-12)
    o = oracle.OracleAlto(dims, cidx, vals, nprtn=8)
    Uo, lo = oracle.random_init(dims, R, 3)
    no, fo = o.cpd_als(Uo, lo, 6, 0.0)
    np.testing.assert_allclose(f.value, fo[-1], rtol=1e-9, atol=1e-12)


def _bf16_round(x):
    """FP32 -> BF16 with round to nearest even, back to float64."""
    a = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    r = ((a + 0x7FFF + ((a >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def _stored_model_fit(A, g, R, U, lam):
    Ub = [np.ascontiguousarray(_bf16_round(u)) for u in U]
    f = ctypes.c_double(0)
    A.check(A.lib().alto_cpd_fit(g.h, R, A._ptr_array([u.ctypes.data for u in Ub]), lam.ctypes.data, ctypes.byref(f)),
            "cpd_fit")
    return f.value


@pytest.mark.parametrize("lockstep", [False, True])
@pytest.mark.parametrize("R", [8, 33, 64, 80])
def test_bf16_fit_is_the_fit_of_the_stored_model(R, lockstep):
    """A BF16 model's fit (the inner product taken from the stored BF16
    factors in the last mode's Gram pass; rank 80 through the separate
    inner-product kernel) equals the FP64 fit of the model it returns with
    its factors rounded to BF16 -- the model its MTTKRPs use -- within the
    FP32 accumulation of its MTTKRP (1e-5 relative)."""
    A = _A()
    rng = np.random.default_rng(R)
    cases = [([120, 90, 300], 30000, "zipf"), ([40, 30, 20, 25], 9000, "uniform"), ([2000, 50, 900], 20000, "zipf")]
    sessions, tensors = [], []
    for k, (dims, nnz, kind) in enumerate(cases):
        gen = zipf_tensor if kind == "zipf" else uniform_tensor
        cidx, vals = gen(dims, nnz, 70 + k)
        vals = rng.lognormal(size=vals.shape[0])
        g = A.AltoTensor(dims, cidx, vals, nprtn=8)
        U, _ = oracle.random_init(dims, R, 5 + k)
        tensors.append((dims, g))
        sessions.append(A.CpdSession(g, U, precision=2))
    if lockstep and R <= 64:
        A.cpd_sessions_iterate(sessions, 10)
    else:
        for x in sessions:
            x.iterate(10)
    for (dims, g), x in zip(tensors, sessions):
        U, lam, fit, _ = x.get()
        f = _stored_model_fit(A, g, R, U, lam)
        assert abs(f - fit) <= 1e-5 * max(1.0, abs(fit)), (dims, f, fit)


@pytest.mark.parametrize("job", [1, 15, 31])
def test_bf16_fit_on_serving_jobs_with_negative_residual(job):
    """Jobs of the seeded serving stream (duplicate coordinates kept,
    lognormal values) whose residual goes negative, so that the fit formula
    gives 55-2800 and amplifies any inconsistency: the BF16 fit equals the
    stored model's within 1e-5 relative (a fit whose inner product came
    from the unrounded factors was 1.4e-3 - 2.2e-3 off on these jobs) and
    the FP64 fit within 1e-3 relative."""
    import torch
    from alto_amd.serve import _generate, job_stream
    A = _A()
    j = dict(job_stream(job + 1, seed=5)[job])
    coords, vals, nnz = _generate(j)
    torch.cuda.synchronize()
    g = A.AltoTensor.from_device(j["dims"], [c.data_ptr() for c in coords], vals.data_ptr(), nnz, nprtn=1)
    R = j["rank"]
    fits = {}
    for prec in (0, 2):
        h = ctypes.c_void_p()
        A.check(A.lib().alto_cpd_session_create_seeded(g.h, R, prec, j["seed"] % (2 ** 31), None, ctypes.byref(h)),
                "session")
        tr = np.zeros(20)
        ran, conv = ctypes.c_int(0), ctypes.c_int(0)
        A.check(A.lib().alto_cpd_session_iterate(h, 20, 1e-5, tr.ctypes.data, ctypes.byref(ran), ctypes.byref(conv),
                                                 0), "iterate")
        U = [np.zeros((d, R)) for d in j["dims"]]
        lam = np.zeros(R)
        f = ctypes.c_double(0)
        A.check(A.lib().alto_cpd_session_get(h, A._ptr_array([u.ctypes.data for u in U]), lam.ctypes.data,
                                             ctypes.byref(f), None), "get")
        A.lib().alto_cpd_session_destroy(h)
        fits[prec] = (tr[ran.value - 1], U, lam)
    fb, Ub, lb = fits[2]
    assert abs(fb) > 10  # the regime where the residual is negative
    f_exact = _stored_model_fit(A, g, R, Ub, lb)
    assert abs(f_exact - fb) <= 1e-5 * abs(fb), (f_exact, fb)
    assert abs(fb - fits[0][0]) <= 1e-3 * abs(fits[0][0]), (fb, fits[0][0])
