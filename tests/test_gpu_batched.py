"""GPU parity of alto_mttkrp_batched: many MTTKRPs of uneven tensors in one
launch give what each job's own alto_mttkrp gives (the same kernel body over
the same segments; only the order of the floating-point atomic additions
differs) and what FP64 brute force over the coordinates gives (FP32 / BF16
within their storage rounding)."""
import numpy as np
import pytest

from tests._tensors import mttkrp_numpy, uniform_tensor, zipf_tensor

pytestmark = pytest.mark.gpu

# 2..5 modes, 64- and 128-bit keys, 1 .. 200K nonzeros, Zipf heavy rows
SPECS = [
    ([500, 3000], 20000, "zipf", 1),
    ([40, 50, 60], 3000, "uniform", 1),
    ([300, 200, 500], 200000, "zipf", 1),
    ([30, 20, 10, 40], 5000, "uniform", 2),
    ([2**14, 2**13, 2**13, 2**12, 2**12], 6000, "uniform", 2),
    ([100, 90, 80, 70, 60], 4000, "uniform", 1),
    ([3, 4, 5], 1, "uniform", 1),
    ([1, 7, 1], 7, "uniform", 1),
    ([2000, 1500, 3000], 60000, "zipf", 1),
]


def _alto():
    import alto_amd
    return alto_amd


def _jobs(specs, seed0=0):
    A = _alto()
    out = []
    for k, (dims, nnz, kind, W) in enumerate(specs):
        gen = zipf_tensor if kind == "zipf" else uniform_tensor
        cidx, vals = gen(dims, nnz, 100 + seed0 + k)
        g = A.AltoTensor(dims, cidx, vals, nprtn=8, lit_bits=64 * W)
        out.append((dims, cidx, vals, g))
    return out


@pytest.fixture
def tune_reset():
    A = _alto()
    yield A
    for param, value in [(4, -1), (3, 0), (2, 0), (1, 0), (7, 0), (8, 512), (9, 32768), (10, 32), (11, 1), (12, 1), (13, 1), (14, 3), (15, 30), (16, 4 << 20), (17, 0)]:
        A.check(A.lib().alto_tune_set(param, value), "tune_set")


def _tol(prec):
    return 1e-12 if prec == 0 else 1e-5


@pytest.mark.parametrize("prec", [0, 1, 2])
@pytest.mark.parametrize("R", [8, 16, 32, 64, 100, 128])
def test_batched_equals_single(prec, R):
    A = _alto()
    jobs = _jobs(SPECS)
    rng = np.random.default_rng(R)
    Us = [[rng.random((d, R)) for d in dims] for dims, _, _, _ in jobs]
    for mode_of in (lambda k, n: 0, lambda k, n: (k + 1) % n):
        modes = [mode_of(k, len(j[0])) for k, j in enumerate(jobs)]
        got = A.mttkrp_batched([j[3] for j in jobs], modes, Us, prec)
        for (dims, cidx, vals, g), m, U, x in zip(jobs, modes, Us, got):
            want = g.mttkrp(m, U, prec)
            scale = max(np.abs(want).max(), 1e-300)
            assert np.abs(x.astype(np.float64) - want).max() <= _tol(prec) * scale, (dims, m)
            # against brute force in FP64 over the coordinates (the reduced
            # precisions within their storage rounding)
            ref = mttkrp_numpy(cidx, vals, U, m)
            if prec == 0:
                np.testing.assert_allclose(x, ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max())
            else:
                tol = 1e-5 if prec == 1 else 2e-2
                assert np.abs(x.astype(np.float64) - ref).max() <= tol * max(np.abs(ref).max(), 1e-300), (dims, m)


def test_batched_more_jobs_than_one_launch_holds():
    """40 jobs of one lane shape: three launches of at most 16 jobs."""
    A = _alto()
    specs = [([60 + k, 50 + 2 * k, 40 + 3 * k], 500 + 997 * k, "zipf" if k % 3 else "uniform", 1) for k in range(40)]
    jobs = _jobs(specs, seed0=7)
    R = 16
    rng = np.random.default_rng(5)
    Us = [[rng.random((d, R)) for d in dims] for dims, _, _, _ in jobs]
    modes = [k % 3 for k in range(40)]
    got = A.mttkrp_batched([j[3] for j in jobs], modes, Us, 0)
    for (dims, cidx, vals, g), m, U, x in zip(jobs, modes, Us, got):
        ref = mttkrp_numpy(cidx, vals, U, m)
        np.testing.assert_allclose(x, ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max())


@pytest.mark.parametrize("knobs", [[(4, 0)], [(4, 1)], [(7, 1)], [(12, 0)], [(1, 100)], [(2, 10)]],
                         ids=["strided", "contig", "hot_rows", "wide_small", "seg100", "chunk10"])
def test_batched_with_tuning_knobs(tune_reset, knobs):
    """Lane mappings the batched kernel does not take send every job to its
    own launch; the results stay those of brute force."""
    A = tune_reset
    for p, v in knobs:
        A.check(A.lib().alto_tune_set(p, v), "tune_set")
    jobs = _jobs(SPECS[:6], seed0=3)
    R = 32
    rng = np.random.default_rng(9)
    Us = [[rng.random((d, R)) for d in dims] for dims, _, _, _ in jobs]
    for prec in (0, 1):
        modes = [len(j[0]) - 1 for j in jobs]
        got = A.mttkrp_batched([j[3] for j in jobs], modes, Us, prec)
        for (dims, cidx, vals, g), m, U, x in zip(jobs, modes, Us, got):
            Uc = U if prec == 0 else [u.astype(np.float32).astype(np.float64) for u in U]
            ref = mttkrp_numpy(cidx, vals, Uc, m)
            tol = 1e-12 if prec == 0 else 1e-4
            assert np.abs(x.astype(np.float64) - ref).max() <= tol * np.abs(ref).max(), (dims, m)


def test_batched_arguments():
    A = _alto()
    jobs = _jobs(SPECS[1:2])
    g = jobs[0][3]
    assert A.mttkrp_batched([], [], [], 0) == []
    U = [np.ones((d, 8)) for d in jobs[0][0]]
    with pytest.raises(A.AltoError):
        A.mttkrp_batched([g], [3], [U], 0)
    with pytest.raises(A.AltoError):
        A.mttkrp_batched([g], [0], [U], 7)
