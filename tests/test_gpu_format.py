"""GPU parity: ALTO format construction and MTTKRP / Phi through the C ABI,
against the CPU oracle (itself pinned to the compiled reference) on the same
seeded inputs.  Integer/index state must be bit-exact; FP64 sums within
1e-12 relative (summation order differs), FP32 within 1e-4."""
import numpy as np
import pytest

import oracle
from tests._tensors import mttkrp_numpy, uniform_tensor, zipf_tensor

pytestmark = pytest.mark.gpu

CASES = [
    dict(dims=[40, 50, 60], nnz=3000, P=8, W=1, seed=1),
    dict(dims=[13, 1000, 7], nnz=2500, P=4, W=1, seed=2),
    dict(dims=[5000, 4000, 3000], nnz=4000, P=8, W=1, seed=3),  # low reuse: oidx + CR bits
    dict(dims=[30, 20, 10, 40], nnz=5000, P=6, W=2, seed=4),
    dict(dims=[2**14, 2**13, 2**13, 2**12, 2**12], nnz=6000, P=8, W=2, seed=5),  # 66 index bits
    dict(dims=[100, 90, 80, 70, 60], nnz=4000, P=3, W=1, seed=6),
    dict(dims=[1, 7, 1], nnz=7, P=2, W=1, seed=7),  # singleton modes (0 index bits)
    dict(dims=[3, 4, 5], nnz=1, P=1, W=1, seed=8),  # a single nonzero
    dict(dims=[5000, 4000, 3000], nnz=4000, P=1, W=1, seed=9),  # low reuse, one partition: CR masks, no CR bits
]


def _alto():
    import alto_amd
    return alto_amd


def _pair(case):
    A = _alto()
    cidx, vals = uniform_tensor(case["dims"], case["nnz"], case["seed"])
    g = A.AltoTensor(case["dims"], cidx, vals, nprtn=case["P"], lit_bits=64 * case["W"])
    o = oracle.OracleAlto(case["dims"], cidx, vals, W=case["W"], nprtn=case["P"])
    return cidx, vals, g, o


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"{len(c['dims'])}m_{c['nnz']}_W{c['W']}")
def test_create_bit_exact(case):
    cidx, vals, g, o = _pair(case)
    gs, os_ = g.export(), o.state()
    np.testing.assert_array_equal(gs["masks"], os_["masks"])
    np.testing.assert_array_equal(gs["idx"], os_["idx"])
    np.testing.assert_array_equal(gs["vals"], os_["vals"])
    np.testing.assert_array_equal(gs["prtn_ptr"], os_["prtn_ptr"])
    np.testing.assert_array_equal(gs["intervals"], os_["intervals"])
    np.testing.assert_array_equal(gs["cr_masks"], os_["cr_masks"])
    assert gs["is_oidx"] == os_["is_oidx"]
    np.testing.assert_array_equal(g.delinearize(), o.delinearize())
    for n, oi in os_["oidx"].items():
        np.testing.assert_array_equal(g.oidx(n), oi)  # both stable: identical


def test_pdep_pext_device_arrays():
    A = _alto()
    case = CASES[4]
    cidx, vals, g, o = _pair(case)
    N, nnz = len(case["dims"]), case["nnz"]
    coords = A.DeviceArray.from_numpy(np.ascontiguousarray(cidx))
    keys = A.DeviceArray((nnz, g.W), np.uint64)
    A.check(A.lib().alto_pdep_keys(g.h, coords.ptr, nnz, keys.ptr, None), "pdep")
    back = A.DeviceArray((N, nnz), np.uint64)
    A.check(A.lib().alto_pext_keys(g.h, keys.ptr, nnz, back.ptr, None), "pext")
    A.synchronize()
    k = keys.numpy()
    masks = o.state()["masks"]
    for i in range(0, nnz, 97):
        want = 0
        for n in range(N):
            m = int(masks[n][0]) | (int(masks[n][1]) << 64)
            out = np.zeros(2, np.uint64)
            oracle.lib().or_pdep128(int(cidx[n, i]), np.array([m & (2**64 - 1), m >> 64], np.uint64), out)
            want |= int(out[0]) | (int(out[1]) << 64)
        assert int(k[i, 0]) | (int(k[i, 1]) << 64) == want
    np.testing.assert_array_equal(back.numpy(), cidx)


def _rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


@pytest.mark.parametrize("case", CASES[:6], ids=lambda c: f"{len(c['dims'])}m_{c['nnz']}")
@pytest.mark.parametrize("R", [8, 10, 16, 32, 64, 100])
def test_mttkrp_fp64(case, R):
    cidx, vals, g, o = _pair(case)
    rng = np.random.default_rng(R)
    U = [rng.random((d, R)) for d in case["dims"]]
    for mode in range(len(case["dims"])):
        got = g.mttkrp(mode, U)
        want = o.mttkrp(mode, U)
        np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-12 * np.abs(want).max())


@pytest.mark.parametrize("R", [16, 32, 64])
def test_mttkrp_fp32_bf16_error(R):
    A = _alto()
    dims = [300, 200, 500]
    cidx, vals = zipf_tensor(dims, 20000, 3)
    g = A.AltoTensor(dims, cidx, vals, nprtn=8)
    rng = np.random.default_rng(1)
    U = [rng.random((d, R)) for d in dims]
    for mode in range(3):
        want = mttkrp_numpy(cidx, vals, U, mode)
        assert _rel(g.mttkrp(mode, U, A.PREC_F32).astype(np.float64), want) < 1e-4
        # BF16 storage of the factors: ~2^-8 relative per factor entry
        Ub = [A.from_bf16_bits(A.to_bf16_bits(u)).astype(np.float64) for u in U]
        want_b = mttkrp_numpy(cidx, vals, Ub, mode)
        assert _rel(g.mttkrp(mode, U, A.PREC_BF16).astype(np.float64), want_b) < 1e-4
        assert _rel(g.mttkrp(mode, U, A.PREC_BF16).astype(np.float64), want) < 2e-2


@pytest.mark.parametrize("case", CASES[:6], ids=lambda c: f"{len(c['dims'])}m_{c['nnz']}")
@pytest.mark.parametrize("R", [2, 7, 10, 16, 32, 64, 100])
@pytest.mark.parametrize("variant", [0, 1, 2])
def test_phi_fp64(case, R, variant):
    """The Phi kernels (runtime batch loop / compile-time batch with DPP
    broadcasts / 16-byte lanes with the model row reloaded only where the row
    changes; odd ranks take the second) match the oracle for 2- to 5-mode
    tensors and 64- / 128-bit keys (mode 0 of the second case has runs of
    about 190 nonzeros per row)."""
    import alto_amd as A
    cidx, vals, g, o = _pair(case)
    rng = np.random.default_rng(R + 1)
    U = [rng.random((d, R)) + 0.05 for d in case["dims"]]
    A.check(A.lib().alto_tune_set(11, variant), "tune_set")
    try:
        for mode in range(len(case["dims"])):
            got = g.phi(mode, U, 1e-10)
            want = o.phi(mode, U, 1e-10)
            np.testing.assert_allclose(got, want, rtol=1e-11, atol=1e-12 * max(np.abs(want).max(), 1.0))
    finally:
        A.lib().alto_tune_set(11, 1)


def test_mttkrp_accumulate_and_zipf_heavy_rows():
    """Zipf(1.1) heavy-hitter rows: one output row receives a large share of
    all contributions; the result still matches brute force; accumulate adds."""
    A = _alto()
    dims = [2000, 1500, 3000]
    cidx, vals = zipf_tensor(dims, 200000, 11)
    g = A.AltoTensor(dims, cidx, vals, nprtn=8)
    R = 32
    rng = np.random.default_rng(2)
    U = [rng.random((d, R)) for d in dims]
    dev = [A.DeviceArray.from_numpy(u) for u in U]
    for mode in range(3):
        want = mttkrp_numpy(cidx, vals, U, mode)
        out = A.DeviceArray.from_numpy(np.ones_like(want))
        g.mttkrp_device(mode, dev, out, A.PREC_F64, accumulate=True)
        A.synchronize()
        np.testing.assert_allclose(out.numpy(), want + 1.0, rtol=1e-11)


@pytest.fixture
def tune_reset():
    A = _alto()
    yield A
    for param, value in [(4, -1), (3, 0), (2, 0), (1, 0), (7, 0), (8, 512), (9, 32768), (10, 32), (11, 1), (12, 1), (13, 1), (14, 3), (15, 30), (16, 4 << 20), (17, 0)]:
        A.check(A.lib().alto_tune_set(param, value), "tune_set")


@pytest.mark.parametrize("hot", [(0, 512, 49152), (1, 512, 49152), (1, 1024, 0), (1, 256, 20000)],
                         ids=lambda h: f"hot{h[0]}_bt{h[1]}_lds{h[2]}")
@pytest.mark.parametrize("variant", [-1, 0, 1, 2])
@pytest.mark.parametrize("chunk_log2", [0, 10])
@pytest.mark.parametrize("dims", [[300, 40, 700, 25], [300, 40, 700], [500, 3000]], ids=lambda d: f"{len(d)}m")
def test_mttkrp_tuning_knobs_keep_results(tune_reset, hot, variant, chunk_log2, dims):
    """Rank-addressed streams with hot rows staged in LDS (on/off, workgroup
    size, LDS budget including none; streams grouped by hot pattern for 2- and
    3-mode tensors), the lane mapping and the chunk size change only the
    summation order: every combination matches the oracle in FP64, FP32 and
    BF16, with the tolerances of the tests above.  Rank 7 and 100 rows are not
    16-byte aligned in every precision (the 16-byte mapping then falls back)."""
    A = tune_reset
    A.check(A.lib().alto_tune_set(7, hot[0]), "tune_set")
    A.check(A.lib().alto_tune_set(8, hot[1]), "tune_set")
    A.check(A.lib().alto_tune_set(9, hot[2]), "tune_set")
    A.check(A.lib().alto_tune_set(4, variant), "tune_set")
    A.check(A.lib().alto_tune_set(2, chunk_log2), "tune_set")
    cidx, vals = zipf_tensor(dims, 30000, 17)
    g = A.AltoTensor(dims, cidx, vals, nprtn=8)
    o = oracle.OracleAlto(dims, cidx, vals, W=1, nprtn=8)
    for R in (4, 7, 8, 16, 32, 64, 100):
        rng = np.random.default_rng(R)
        U = [rng.random((d, R)) for d in dims]
        for mode in range(len(dims)):
            want = o.mttkrp(mode, U)
            np.testing.assert_allclose(g.mttkrp(mode, U), want, rtol=1e-12, atol=1e-12 * np.abs(want).max())
            assert _rel(g.mttkrp(mode, U, A.PREC_F32).astype(np.float64), want) < 1e-4
            Ub = [A.from_bf16_bits(A.to_bf16_bits(u)).astype(np.float64) for u in U]
            assert _rel(g.mttkrp(mode, U, A.PREC_BF16).astype(np.float64), mttkrp_numpy(cidx, vals, Ub, mode)) < 1e-4


@pytest.mark.parametrize("seg", [1, 16, 37, 64, 256, 4096, 100000])
@pytest.mark.parametrize("dims", [[300, 40, 700, 25], [500, 3000]], ids=lambda d: f"{len(d)}m")
def test_mttkrp_segment_lengths_keep_results(tune_reset, seg, dims):
    """Any segment length per lane group (one nonzero, not a batch multiple,
    longer than the tensor) changes only the summation order."""
    A = tune_reset
    A.check(A.lib().alto_tune_set(1, seg), "tune_set")
    cidx, vals = zipf_tensor(dims, 20000, 29)
    g = A.AltoTensor(dims, cidx, vals, nprtn=8)
    o = oracle.OracleAlto(dims, cidx, vals, W=1, nprtn=8)
    for R in (8, 16, 64):
        rng = np.random.default_rng(R)
        U = [rng.random((d, R)) for d in dims]
        for mode in range(len(dims)):
            want = o.mttkrp(mode, U)
            np.testing.assert_allclose(g.mttkrp(mode, U), want, rtol=1e-12, atol=1e-12 * np.abs(want).max())
            assert _rel(g.mttkrp(mode, U, A.PREC_F32).astype(np.float64), want) < 1e-4


@pytest.mark.parametrize("narrow_small", [0, 1])
@pytest.mark.parametrize("dims", [[300, 40, 700, 25], [500, 3000]], ids=lambda d: f"{len(d)}m")
def test_mttkrp_two_columns_per_lane_small_tensors(tune_reset, narrow_small, dims):
    """FP32 / BF16 MTTKRPs of every rank with two columns per lane (the
    small-tensor default) or the wider lanes match numpy on the stored
    factors."""
    A = tune_reset
    A.check(A.lib().alto_tune_set(12, narrow_small), "tune_set")
    cidx, vals = zipf_tensor(dims, 20000, 31)
    g = A.AltoTensor(dims, cidx, vals, nprtn=8)
    for R in (24, 32, 64, 100, 128):
        rng = np.random.default_rng(R)
        U = [rng.random((d, R)) for d in dims]
        for mode in range(len(dims)):
            U32 = [u.astype(np.float32).astype(np.float64) for u in U]
            assert _rel(g.mttkrp(mode, U, A.PREC_F32).astype(np.float64), mttkrp_numpy(cidx, vals, U32, mode)) < 1e-4
            Ub = [A.from_bf16_bits(A.to_bf16_bits(u)).astype(np.float64) for u in U]
            assert _rel(g.mttkrp(mode, U, A.PREC_BF16).astype(np.float64), mttkrp_numpy(cidx, vals, Ub, mode)) < 1e-4


@pytest.mark.parametrize("prec", [0, 1, 2])
def test_mttkrp_autotune_keeps_results(prec):
    """The autotuner times every lane mapping and segment length per mode,
    keeps the fastest on the tensor, and the MTTKRPs that then use it still
    match the oracle (FP64 1e-12; FP32 / BF16 1e-4 of the FP64 result on
    the same stored factors)."""
    A = _alto()
    dims = [300, 400, 2000]
    cidx, vals = zipf_tensor(dims, 60000, 23)
    g = A.AltoTensor(dims, cidx, vals, nprtn=8)
    o = oracle.OracleAlto(dims, cidx, vals, W=1, nprtn=8)
    R = 32
    res = g.mttkrp_autotune(R, prec, reps=2)
    assert [r["mode"] for r in res] == [0, 1, 2]
    for r in res:
        assert r["variant"] in (-1, 0, 1, 2) and r["seg"] in (0, 16, 32, 64, 128, 256, 512, 1024, 2048, 4096)
        assert r["ms"] <= r["default_ms"] + 1e-12
    rng = np.random.default_rng(4)
    U = [rng.random((d, R)) for d in dims]
    for mode in range(3):
        if prec == 0:
            want = o.mttkrp(mode, U)
            np.testing.assert_allclose(g.mttkrp(mode, U), want, rtol=1e-12, atol=1e-12 * np.abs(want).max())
        else:
            Us = U if prec == 1 else [A.from_bf16_bits(A.to_bf16_bits(u)).astype(np.float64) for u in U]
            Us = [u.astype(np.float32).astype(np.float64) for u in Us]
            assert _rel(g.mttkrp(mode, U, prec).astype(np.float64), mttkrp_numpy(cidx, vals, Us, mode)) < 1e-4
