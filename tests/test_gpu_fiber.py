"""Fiber-ordered mode streams: MTTKRP and CP-ALS with the long-fiber kernel
(the fiber-mode factor row read once per fiber) against the oracle, with the
layout forced onto small tensors (3-5 modes, 64- and 128-bit keys, every
lane shape from rank 1 to 100, fiber thresholds 2-16, chunk sizes that cut
fibers, one or two fiber modes per stream), and the split of each stream
checked against its definition."""
import numpy as np
import pytest

import oracle
from tests._tensors import uniform_tensor, zipf_tensor

pytestmark = pytest.mark.gpu

FIBERS, FIBER_MIN, FIBER_CHUNK, FIBER_MIN_NNZ = 13, 14, 15, 16


def _A():
    import alto_amd
    return alto_amd


@pytest.fixture
def fibers():
    """Forces the fiber layout onto every tensor; restores the defaults."""
    A = _A()

    def set_(min_len=8, chunk=30, tiers=2):
        for p, v in ((FIBERS, tiers), (FIBER_MIN, min_len), (FIBER_CHUNK, chunk), (FIBER_MIN_NNZ, 0)):
            A.check(A.lib().alto_tune_set(p, v), "tune_set")

    yield set_
    for p, v in ((FIBERS, 1), (FIBER_MIN, 3), (FIBER_CHUNK, 30), (FIBER_MIN_NNZ, 4 << 20), (17, 0)):
        A.check(A.lib().alto_tune_set(p, v), "tune_set")


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


CASES = [([60, 70, 80], 20000, 1), ([300, 40, 500], 30000, 1), ([30, 20, 10, 40], 15000, 1),
         ([64, 50, 40, 30, 20], 25000, 1), ([2 ** 14, 2 ** 13, 40, 2 ** 12, 2 ** 14], 20000, 2),
         ([500, 9, 700], 40000, 2)]


@pytest.mark.parametrize("tiers", [1, 2])
@pytest.mark.parametrize("dims,nnz,W", CASES)
@pytest.mark.parametrize("R", [1, 8, 16, 32, 64, 100])
def test_fiber_mttkrp_fp64_matches_oracle(fibers, dims, nnz, W, R, tiers):
    A = _A()
    fibers(min_len=4, tiers=tiers)
    cidx, vals = zipf_tensor(dims, nnz, 31 + R, s=1.2)
    g = A.AltoTensor(dims, cidx, vals, nprtn=8, lit_bits=64 * W)
    o = oracle.OracleAlto(dims, cidx, vals, W=W, nprtn=8)
    rng = np.random.default_rng(R)
    U = [rng.random((d, R)) for d in dims]
    for mode in range(len(dims)):
        assert _rel(g.mttkrp(mode, U), o.mttkrp(mode, U)) < 1e-12, mode


@pytest.mark.parametrize("tiers", [1, 2])
@pytest.mark.parametrize("R", [7, 8, 16, 32, 64])
@pytest.mark.parametrize("prec,tol", [(1, 1e-5), (2, 1e-2)])
def test_fiber_mttkrp_reduced_precision(fibers, R, prec, tol, tiers):
    """FP32 (2 or 4 columns per lane) and BF16 storage against the FP64 oracle."""
    A = _A()
    fibers(min_len=8, tiers=tiers)
    dims = [200, 150, 300]
    cidx, vals = zipf_tensor(dims, 60000, 5, s=1.2)
    g = A.AltoTensor(dims, cidx, vals, nprtn=8)
    o = oracle.OracleAlto(dims, cidx, vals, nprtn=8)
    rng = np.random.default_rng(3)
    U = [rng.random((d, R)) for d in dims]
    for mode in range(3):
        assert _rel(g.mttkrp(mode, U, prec), o.mttkrp(mode, U)) < tol, mode


@pytest.mark.parametrize("tiers", [1, 2])
@pytest.mark.parametrize("min_len,chunk,rule", [(2, 30, 0), (4, 10, 1), (16, 12, 0), (1 << 20, 30, 1)])
def test_fiber_thresholds_and_chunks(fibers, min_len, chunk, rule, tiers):
    """Any threshold (every fiber long, none long), chunks that split fibers
    and either fiber-mode rule give the same MTTKRP."""
    A = _A()
    fibers(min_len=min_len, chunk=chunk, tiers=tiers)
    A.check(A.lib().alto_tune_set(17, rule), "tune_set")
    dims = [120, 90, 150, 60]
    cidx, vals = zipf_tensor(dims, 50000, 8, s=1.1)
    g = A.AltoTensor(dims, cidx, vals, nprtn=8)
    o = oracle.OracleAlto(dims, cidx, vals, nprtn=8)
    rng = np.random.default_rng(4)
    U = [rng.random((d, 32)) for d in dims]
    for mode in range(4):
        assert _rel(g.mttkrp(mode, U), o.mttkrp(mode, U)) < 1e-12, mode


def test_fiber_cpd_als_matches_oracle(fibers):
    A = _A()
    fibers(min_len=4)
    dims = [200, 300, 250]
    cidx, vals = uniform_tensor(dims, 30000, 5)
    g = A.AltoTensor(dims, cidx, vals, nprtn=8)
    U, lam = oracle.random_init(dims, 32, 3)
    n, fit, _ = g.cpd_als(U, lam, 12, 0.0, A.PREC_F64)
    o = oracle.OracleAlto(dims, cidx, vals, nprtn=8)
    Uo, lo = oracle.random_init(dims, 32, 3)
    no, fo = o.cpd_als(Uo, lo, 12, 0.0)
    assert n == no == 12
    np.testing.assert_allclose(fit, fo, rtol=1e-10, atol=1e-12)


def test_fiber_stream_split_definition(fibers):
    """The streams are not exported, so the layout is checked through its
    effect: with a threshold above every fiber length all nonzeros take the
    short part, with threshold 2 every repeated (row, fiber row) pair the
    long one, and both equal the oracle (the long part alone is exercised
    with min_len 2 on a tensor whose fibers all repeat)."""
    A = _A()
    fibers(min_len=2)
    dims = [8, 5, 300]
    rng = np.random.default_rng(9)
    # every (row_0, row_1) pair repeated over 20 rows of mode 2: all fibers long
    i0, i1 = np.meshgrid(np.arange(8), np.arange(5), indexing="ij")
    rows = []
    for a, b in zip(i0.ravel(), i1.ravel()):
        for c in rng.choice(300, size=20, replace=False):
            rows.append((a, b, c))
    cidx = np.ascontiguousarray(np.array(rows, dtype=np.uint64).T)
    vals = rng.random(cidx.shape[1]) + 0.5
    g = A.AltoTensor(dims, cidx, vals, nprtn=4)
    o = oracle.OracleAlto(dims, cidx, vals, nprtn=4)
    U = [rng.random((d, 16)) for d in dims]
    for mode in range(3):
        assert _rel(g.mttkrp(mode, U), o.mttkrp(mode, U)) < 1e-12, mode


def test_fiber_second_mode_fibers(fibers):
    """A tensor whose mode-0 stream has no fiber of the first fiber mode
    (mode 1, the fewest rows: each (row_0, row_1) pair holds at most 3
    nonzeros) but fibers of 10 nonzeros in mode 2: with two fiber modes
    they all take the second fiber part; every mode equals the oracle with
    one and with two fiber modes."""
    A = _A()
    dims = [8, 40, 60]
    rng = np.random.default_rng(12)
    rows = []
    for a in range(8):
        for c in rng.choice(60, size=3, replace=False):
            for b in rng.choice(40, size=10, replace=False):
                rows.append((a, b, c))
    cidx = np.ascontiguousarray(np.array(rows, dtype=np.uint64).T)
    vals = rng.random(cidx.shape[1]) + 0.5
    o = oracle.OracleAlto(dims, cidx, vals, nprtn=4)
    U = [rng.random((d, 16)) for d in dims]
    for tiers in (1, 2):
        fibers(min_len=4, tiers=tiers)
        g = A.AltoTensor(dims, cidx, vals, nprtn=4)
        for mode in range(3):
            assert _rel(g.mttkrp(mode, U), o.mttkrp(mode, U)) < 1e-12, (tiers, mode)
