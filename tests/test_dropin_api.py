"""The C++ drop-in API (include/alto/*.hpp) as a user of the reference
library calls it: the reference's own command line (src/main.cpp) built
unchanged against the forwarding headers and linked with libalto_hip (CPU:
it builds and starts; GPU: its check, bench, ALS and APR runs against the
oracle), and a program that calls every group of the API (COO tensors and
their files, Kruskal models and their normalisations, COO and ALTO MTTKRP,
the fit from its parts, COO and ALTO CP-ALS and CP-APR, the Poisson
generator) with each result checked against the oracle."""
import os
import subprocess

import numpy as np
import pytest

import oracle
from tests._tensors import uniform_tensor

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INC = os.path.join(ROOT, "include")
LIBDIR = os.path.join(ROOT, "alto_amd")
REF_MAIN = {64: os.path.join(ROOT, "oracle", "_ref", "ref_main64"),
            128: os.path.join(ROOT, "oracle", "_ref", "ref_main128")}
HAVE_REF_TREE = os.path.isfile("/root/reference/src/main.cpp")
HAVE_LIB = os.path.exists(os.path.join(LIBDIR, "libalto_hip.so"))


def _write_tns(path, dims, cidx, vals):
    with open(path, "w") as f:
        f.write("# synthetic\n")
        for i in range(len(vals)):
            f.write(" ".join(str(int(cidx[m, i]) + 1) for m in range(len(dims))) + f" {vals[i]:.17g}\n")


def _env(threads=4):
    return dict(os.environ, OMP_NUM_THREADS=str(threads))


# ------------------------------------------------------------------ CPU

@pytest.mark.parametrize("hdr", ["common", "sptensor", "kruskal_model", "alto", "cpd", "apr", "poisson_generator"])
def test_each_header_stands_alone(tmp_path, hdr):
    """Every drop-in header (and its reference-named forwarder) compiles on
    its own, with and without OpenMP, warning-free."""
    for inc, extra in ((f'"alto/{hdr}.hpp"', []), (f'"{hdr}.hpp"', ["-I" + os.path.join(INC, "alto", "ref")])):
        src = tmp_path / "one.cpp"
        src.write_text(f"#include {inc}\nint main() {{ return 0; }}\n")
        for omp in ([], ["-fopenmp"]):
            r = subprocess.run(["g++", "-std=c++17", "-Wall", "-Wextra", "-Werror", "-Wno-unused-function",
                                "-fsyntax-only", "-I" + INC, *extra, *omp, str(src)], capture_output=True, text=True)
            assert r.returncode == 0, r.stderr


@pytest.mark.skipif(not (HAVE_REF_TREE and HAVE_LIB), reason="needs the reference tree and libalto_hip.so")
def test_reference_main_builds_against_dropin():
    """src/main.cpp of the reference, compiled where it lies with only
    -I include/alto/ref, links against libalto_hip (64- and 128-bit keys)
    and starts: -v and -h touch no GPU."""
    subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "dropin"])
    for w, exe in REF_MAIN.items():
        assert os.path.exists(exe)
        r = subprocess.run(["ldd", exe], capture_output=True, text=True)
        assert "libalto_hip.so" in r.stdout, r.stdout
        r = subprocess.run([exe, "-v"], capture_output=True, text=True, timeout=30)
        assert r.returncode == 0 and "version 0.1.2" in r.stdout
        r = subprocess.run([exe, "-h"], capture_output=True, text=True, timeout=30)
        assert r.returncode == 0 and "-c or --check" in r.stderr
        r = subprocess.run([exe], capture_output=True, text=True, timeout=30)  # no input: usage, exit -1
        assert r.returncode != 0 and "No dims specified" in r.stderr


def test_kruskal_export_format(tmp_path):
    """ExportKruskalModel writes <path>.<n>.out (matrix, 2, dims rank, then
    the entries column by column, %.20lf) and <path>.lambda.out."""
    src = tmp_path / "k.cpp"
    src.write_text(r'''
#include "alto/kruskal_model.hpp"
int main() {
    IType dims[2] = {3, 2};
    KruskalModel* M; CreateKruskalModel(2, dims, 2, &M);
    for (int n = 0; n < 2; ++n) for (IType i = 0; i < dims[n] * 2; ++i) M->U[n][i] = 10 * n + i + 0.5;
    M->lambda[0] = 0.25; M->lambda[1] = 4;
    ExportKruskalModel(M, (char*)"m");
    DestroyKruskalModel(M);
    return 0;
}
''')
    exe = tmp_path / "k"
    r = subprocess.run(["g++", "-std=c++17", "-I" + INC, str(src), "-o", str(exe)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    assert subprocess.run([str(exe)], cwd=tmp_path).returncode == 0
    m0 = (tmp_path / "m.0.out").read_text().split()
    assert m0[:4] == ["matrix", "2", "3", "2"]
    assert [float(x) for x in m0[4:]] == [0.5, 2.5, 4.5, 1.5, 3.5, 5.5]
    lam = (tmp_path / "m.lambda.out").read_text().split()
    assert lam[:3] == ["vector", "1", "2"] and [float(x) for x in lam[3:]] == [0.25, 4.0]
    assert m0[4] == "0.50000000000000000000"


# ------------------------------------------------------------------ GPU

def _read_model(prefix, nmodes):
    U = []
    for n in range(nmodes):
        t = open(f"{prefix}.{n}.out").read().split()
        assert t[0] == "matrix" and t[1] == "2"
        I, R = int(t[2]), int(t[3])
        U.append(np.array([float(x) for x in t[4:]]).reshape(R, I).T.copy())
    t = open(f"{prefix}.lambda.out").read().split()
    return U, np.array([float(x) for x in t[3:]])


_API_PROG = r'''
#include <math.h>
#include <stdio.h>
#include "alto/alto.hpp"
#include "alto/apr.hpp"
#include "alto/cpd.hpp"
#include "alto/poisson_generator.hpp"
// tns rank seed: every group of the drop-in API on one tensor; models are
// written with ExportKruskalModel, scalars printed as "name value"
typedef AltoTensor<unsigned long long> AT64;
static KruskalModel* fresh(SparseTensor* X, IType R, unsigned seed) {
    KruskalModel* M; CreateKruskalModel(X->nmodes, X->dims, R, &M); KruskalModelRandomInit(M, seed); return M;
}
int main(int argc, char** argv) {
    SparseTensor* X; ImportSparseTensor(argv[1], TEXT_FORMAT, &X, 0);
    const IType R = atoll(argv[2]); const unsigned seed = atoi(argv[3]);
    const int N = X->nmodes;
    AT64* AT; create_alto(X, &AT, 8);
    printf("alto nnz %llu oidx %d\n", AT->nnz, (int)AT->is_oidx);
    // COO MTTKRP, mode after mode, each adding into U[mode] (later modes see the updated factors)
    KruskalModel* M = fresh(X, R, seed);
    for (int m = 0; m < N; ++m) mttkrp(X, M, (IType)m);
    ExportKruskalModel(M, (char*)"coo"); DestroyKruskalModel(M);
    // the same with the ALTO MTTKRP
    M = fresh(X, R, seed);
    for (int m = 0; m < N; ++m) mttkrp_alto_par(m, M->U, R, AT, nullptr, nullptr);
    ExportKruskalModel(M, (char*)"alto"); DestroyKruskalModel(M);
    // Kruskal utilities, chained
    M = fresh(X, R, seed);
    KruskalModelNormalize(M);               ExportKruskalModel(M, (char*)"l1");
    KruskalModelNormalizeMode(M, N - 1);    ExportKruskalModel(M, (char*)"l1m");
    KruskalModelNorm(M, 1, MAT_NORM_2, nullptr);   ExportKruskalModel(M, (char*)"n2");
    KruskalModelNorm(M, 0, MAT_NORM_MAX, nullptr); ExportKruskalModel(M, (char*)"nmax");
    RedistributeLambda(M, 2);               ExportKruskalModel(M, (char*)"redist");
    // fit from explicit parts: Grams by update_gram, last-mode MTTKRP, ||X||^2
    FType** grams = (FType**)AlignedMalloc(N * sizeof(FType*));
    for (int m = 0; m < N; ++m) { grams[m] = (FType*)AlignedMalloc(R * R * sizeof(FType)); update_gram(grams[m], M, m); }
    FType* Um = (FType*)AlignedMalloc(X->dims[N - 1] * R * sizeof(FType));
    KruskalModel* Z = fresh(X, R, seed);
    for (int m = 0; m < N; ++m) ParMemcpy(Z->U[m], M->U[m], X->dims[m] * R * sizeof(FType));
    ParMemset(Z->U[N - 1], 0, X->dims[N - 1] * R * sizeof(FType));
    mttkrp(X, Z, (IType)(N - 1));
    ParMemcpy(Um, Z->U[N - 1], X->dims[N - 1] * R * sizeof(FType));
    double nx = 0; for (IType i = 0; i < X->nnz; ++i) nx += X->vals[i] * X->vals[i];
    printf("fit_alto %.17g\n", cpd_fit_alto(AT, M, grams, Um, nx));
    printf("fit_coo %.17g\n", cpd_fit(X, M, grams, Um));
    for (int m = 0; m < N; ++m) AlignedFree(grams[m]);
    AlignedFree(grams); AlignedFree(Um); DestroyKruskalModel(Z); DestroyKruskalModel(M);
    // CP-ALS: COO and ALTO drivers
    M = fresh(X, R, seed); cpd(X, M, 5, 0.0);      ExportKruskalModel(M, (char*)"cpd"); DestroyKruskalModel(M);
    M = fresh(X, R, seed); cpd_alto(AT, M, 5, 0.0); ExportKruskalModel(M, (char*)"als"); DestroyKruskalModel(M);
    // CP-APR: COO drivers (COO stop rule) and the ALTO driver
    M = fresh(X, R, seed); cp_apr_mu_alto(X, M, 6, 3, 1e-4, 1e-10, 1e-10, 1e-2);
    ExportKruskalModel(M, (char*)"apr_coo"); DestroyKruskalModel(M);
    M = fresh(X, R, seed); cp_apr_mu_alto_otf(X, M, 6, 3, 1e-4, 1e-10, 1e-10, 1e-2);
    ExportKruskalModel(M, (char*)"apr_otf"); DestroyKruskalModel(M);
    M = fresh(X, R, seed); cp_apr_alto(AT, M, 6, 3, 1e-4, 1e-10, 1e-10, 1e-2, false);
    ExportKruskalModel(M, (char*)"apr_alto"); DestroyKruskalModel(M);
    destroy_alto(AT);
    // COO tensors: CreateSparseTensor from row-major coordinates, binary round trip
    IType* rm = (IType*)AlignedMalloc(X->nnz * N * sizeof(IType));
    for (IType i = 0; i < X->nnz; ++i) for (int m = 0; m < N; ++m) rm[i * N + m] = X->cidx[m][i];
    SparseTensor* Y; CreateSparseTensor(N, X->dims, X->nnz, rm, X->vals, &Y);
    ExportSparseTensor("y.bin", BINARY_FORMAT, Y);
    SparseTensor* Y2; ImportSparseTensor("y.bin", BINARY_FORMAT, &Y2, 0);
    bool same = Y2->nnz == X->nnz && Y2->nmodes == N;
    for (int m = 0; m < N && same; ++m) same = Y2->dims[m] == X->dims[m] && !memcmp(Y2->cidx[m], X->cidx[m], X->nnz * 8);
    same = same && !memcmp(Y2->vals, X->vals, X->nnz * 8);
    printf("sptensor_roundtrip %d\n", (int)same);
    AlignedFree(rm); DestroySparseTensor(Y); DestroySparseTensor(Y2); DestroySparseTensor(X);
    // the Poisson generator with its model; its permutations come from
    // rand(), which KruskalModelRandomInit reseeded: back to a fresh process
    srand(1);
    IType gd[3] = {30, 20, 10};
    PoissonGenerator* pg; CreatePoissonGenerator(3, gd, &pg);
    KruskalModel* G; SparseTensor* GX;
    PoissonGeneratorRun(pg, 300, 4, &G, &GX);
    DestroyPoissonGenerator(pg);
    ExportSparseTensor("gen.bin", BINARY_FORMAT, GX); ExportKruskalModel(G, (char*)"gen");
    DestroyKruskalModel(G); DestroySparseTensor(GX);
    return 0;
}
'''


@pytest.fixture(scope="module")
def api_run(tmp_path_factory):
    d = tmp_path_factory.mktemp("api")
    dims, R, seed = [40, 30, 20], 6, 7
    cidx, vals = uniform_tensor(dims, 3000, 11)
    vals = np.ceil(vals * 4)  # counts, for CP-APR
    _write_tns(d / "x.tns", dims, cidx, vals)
    src = d / "api.cpp"
    src.write_text(_API_PROG)
    exe = d / "api"
    r = subprocess.run(["g++", "-std=c++17", "-O1", "-I" + INC, str(src), "-o", str(exe), "-L" + LIBDIR,
                        "-lalto_hip", "-Wl,-rpath," + LIBDIR], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    print("dropin api program: dims", dims, "rank", R, flush=True)
    r = subprocess.run([str(exe), str(d / "x.tns"), str(R), str(seed)], cwd=d, env=_env(1), capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    scal = {}
    for line in r.stdout.splitlines():
        p = line.split()
        if len(p) >= 2 and p[0] in ("fit_alto", "fit_coo", "sptensor_roundtrip"):
            scal[p[0]] = float(p[1])
    return dict(dir=d, dims=dims, R=R, seed=seed, cidx=cidx, vals=vals, out=r.stdout, scal=scal)


def _init(ctx):
    return oracle.random_init(ctx["dims"], ctx["R"], ctx["seed"])


def _model(ctx, tag):
    return _read_model(str(ctx["dir"] / tag), len(ctx["dims"]))


def _close(a, b, rtol):
    np.testing.assert_allclose(a, b, rtol=rtol, atol=rtol * max(1e-300, float(np.abs(b).max())))


@pytest.mark.gpu
def test_api_mttkrp_coo_and_alto(api_run):
    """mttkrp (COO, device kernel over the coordinate lists) and
    mttkrp_alto_par, applied mode after mode and adding into U[mode], equal
    the oracle's COO MTTKRP applied the same way."""
    ctx = api_run
    U, _ = _init(ctx)
    for m in range(3):
        U[m] = U[m] + oracle.mttkrp_coo(ctx["dims"], ctx["cidx"], ctx["vals"], U, m)
    for tag in ("coo", "alto"):
        G, _ = _model(ctx, tag)
        for m in range(3):
            _close(G[m], U[m], 1e-12)


@pytest.mark.gpu
def test_api_kruskal_utilities(api_run):
    """KruskalModelNormalize / NormalizeMode (L1), KruskalModelNorm
    (MAT_NORM_2, MAT_NORM_MAX) and RedistributeLambda, chained, against
    their definitions: afterwards the model (lambda times the columns) is
    unchanged where the operation only rescales, the normalised columns
    have unit L1 / 2-norm or maximum at most 1, and lambda equals the norm
    that was divided out."""
    ctx = api_run
    U, lam = _init(ctx)
    l1U, l1l = _model(ctx, "l1")
    cn = [np.abs(u).sum(axis=0) for u in U]
    for m in range(3):
        _close(l1U[m], U[m] / cn[m], 1e-13)
    _close(l1l, lam * cn[0] * cn[1] * cn[2], 1e-13)
    l1mU, l1ml = _model(ctx, "l1m")  # mode 2 already has unit L1 columns
    _close(l1mU[2], l1U[2], 1e-13)
    _close(l1ml, l1l * np.abs(l1U[2]).sum(axis=0), 1e-13)
    n2U, n2l = _model(ctx, "n2")  # lambda replaced by mode 1's column 2-norms
    nrm = np.sqrt((l1mU[1] ** 2).sum(axis=0))
    _close(n2l, nrm, 1e-13)
    _close(n2U[1], l1mU[1] / nrm, 1e-13)
    _close(n2U[0], l1mU[0], 0)
    nmU, nml = _model(ctx, "nmax")  # lambda = max(1, column maximum) of mode 0
    mx = np.maximum(n2U[0].max(axis=0), 1.0)
    _close(nml, mx, 1e-15)
    _close(nmU[0], n2U[0] / mx, 1e-15)
    rdU, rdl = _model(ctx, "redist")  # mode 2 columns times lambda, lambda = 1
    _close(rdU[2], nmU[2] * nml, 1e-15)
    assert np.array_equal(rdl, np.ones(ctx["R"]))


@pytest.mark.gpu
def test_api_fit_from_parts_is_the_residual(api_run):
    """cpd_fit_alto and cpd_fit from explicit Grams / MTTKRP / ||X||^2 equal
    1 - ||X - M||_F / ||X||_F computed by densifying X and the model."""
    ctx = api_run
    U, lam = _model(ctx, "redist")
    dims = ctx["dims"]
    Xd = np.zeros(dims)
    Xd[tuple(ctx["cidx"].astype(np.int64))] = ctx["vals"]
    Md = np.einsum("r,ir,jr,kr->ijk", lam, U[0], U[1], U[2])
    want = 1 - np.linalg.norm(Xd - Md) / np.linalg.norm(Xd)
    assert abs(ctx["scal"]["fit_alto"] - want) < 1e-10
    assert abs(ctx["scal"]["fit_coo"] - want) < 1e-10


@pytest.mark.gpu
def test_api_cp_als_coo_and_alto(api_run):
    """cpd (COO CP-ALS on the device from the building blocks) equals the
    oracle's COO CP-ALS restatement; cpd_alto equals the oracle's ALTO
    CP-ALS (5 sweeps, epsilon 0)."""
    ctx = api_run
    U, lam = _init(ctx)
    n, fits = oracle.cpd_coo(ctx["dims"], ctx["cidx"], ctx["vals"], U, lam, 5, 0.0)
    G, gl = _model(ctx, "cpd")
    for m in range(3):
        _close(G[m], U[m], 1e-8)
    _close(gl, lam, 1e-8)
    lines = [l for l in ctx["out"].splitlines() if l.startswith("it: ") and "time" not in l]
    assert len(lines) == 5
    for l, f in zip(lines, fits):
        assert abs(float(l.split("\t")[1].split()[1]) - f) <= 1e-5 * max(1.0, abs(f))
    o = oracle.OracleAlto(ctx["dims"], ctx["cidx"], ctx["vals"], nprtn=8)
    U2, lam2 = _init(ctx)
    o.cpd_als(U2, lam2, 5, 0.0)
    G, gl = _model(ctx, "als")
    for m in range(3):
        _close(G[m], U2[m], 1e-8)
    _close(gl, lam2, 1e-8)


@pytest.mark.gpu
def test_api_cp_apr_coo_and_alto(api_run):
    """cp_apr_mu_alto and cp_apr_mu_alto_otf (COO stop rule) equal the
    oracle's COO CP-APR restatement; cp_apr_alto equals the oracle's ALTO
    driver (6 outer iterations of up to 3 inner, epsilon 1e-4)."""
    ctx = api_run
    U, lam = _init(ctx)
    res = oracle.cp_apr_coo(ctx["dims"], ctx["cidx"], ctx["vals"], U, lam, 6, 3, eps=1e-4)
    for tag in ("apr_coo", "apr_otf"):
        G, gl = _model(ctx, tag)
        for m in range(3):
            _close(G[m], U[m], 1e-8)
        _close(gl, lam, 1e-8)
    outs = [l for l in ctx["out"].splitlines() if l.startswith("  Total outer iterations")]
    assert len(outs) == 3 and outs[0] == outs[1] == f"  Total outer iterations = {res['outer_iters']}"
    o = oracle.OracleAlto(ctx["dims"], ctx["cidx"], ctx["vals"], nprtn=8)
    U2, lam2 = _init(ctx)
    o.cp_apr(U2, lam2, 6, 3, eps=1e-4)
    G, gl = _model(ctx, "apr_alto")
    for m in range(3):
        _close(G[m], U2[m], 1e-8)
    _close(gl, lam2, 1e-8)


@pytest.mark.gpu
def test_api_tensor_files_and_generator(api_run):
    """CreateSparseTensor (row-major coordinates) -> ExportSparseTensor ->
    ImportSparseTensor is the identity; PoissonGeneratorRun with one stream
    gives the oracle's tensor and model."""
    from oracle.poisson_gen import poisson_tensor
    from tests.test_dropin_cli import _read_bin

    ctx = api_run
    assert ctx["scal"]["sptensor_roundtrip"] == 1
    ci, vo, (Uo, lo) = poisson_tensor([30, 20, 10], 300, 4, 1, return_model=True)
    d, cidx, vals = _read_bin(ctx["dir"] / "gen.bin")
    assert d == [30, 20, 10] and np.array_equal(cidx, np.array(ci, np.uint64)) and np.array_equal(vals, vo)
    G, gl = _model(ctx, "gen")
    for m in range(3):
        _close(G[m], Uo[m], 1e-15)
    _close(gl, lo, 1e-15)


def _ref_main(w):
    exe = REF_MAIN[w]
    if not os.path.exists(exe):
        pytest.skip("the reference's main was not built against the drop-in (no reference tree at build time)")
    return exe


def _its(out):
    return [float(l.split("\t")[1].split()[1]) for l in out.splitlines() if l.startswith("it: ")]


@pytest.mark.gpu
@pytest.mark.parametrize("w", [64, 128])
def test_reference_main_check_and_als(tmp_path, w):
    """The reference's own main over libalto_hip: -c prints "correct" for
    every mode (COO MTTKRP vs ALTO MTTKRP on the device); an ALS run prints
    the oracle's fit trace; -p (bench) times the MTTKRPs."""
    exe = _ref_main(w)
    dims = [30, 40, 50]
    cidx, vals = uniform_tensor(dims, 4000, 3)
    tns = tmp_path / "x.tns"
    _write_tns(tns, dims, cidx, vals)
    print("reference main", w, "check", flush=True)
    r = subprocess.run([exe, "-c", "-i", str(tns), "-r", "8", "-x", "1"], env=_env(), capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stderr.count("Results of mttkrp_alto is correct") == 3, r.stderr
    print("reference main", w, "ALS", flush=True)
    # -l is needed: the reference's main leaves the model unset without it
    r = subprocess.run([exe, "-l", "ALS", "-i", str(tns), "-r", "8", "-m", "6", "-x", "5", "-e", "1e-12"],
                       env=_env(), capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    got = _its(r.stdout)
    o = oracle.OracleAlto(dims, cidx, vals, nprtn=4)
    U, lam = oracle.random_init(dims, 8, 5)
    n, fits = o.cpd_als(U, lam, 6, 1e-12)
    assert len(got) == n == 6
    # the "it:" lines print the fit with %g (6 significant digits)
    for g, f in zip(got, fits):
        assert abs(g - f) <= 5e-6 * max(1.0, abs(f)), (got, fits)
    print("reference main", w, "bench", flush=True)
    r = subprocess.run([exe, "-l", "ALS", "-p", "-i", str(tns), "-r", "8", "-m", "3", "-x", "1"], env=_env(),
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and "ALTO creation time:" in r.stdout and "ALTO runtime:" in r.stdout, r.stdout


@pytest.mark.gpu
def test_reference_main_apr_and_generator(tmp_path):
    """The reference's main: -l APR runs cp_apr_alto and reports the
    oracle's final log-likelihood and iteration counts; -d / -s generates
    the reference's Poisson tensor (one stream) and decomposes it."""
    exe = _ref_main(64)
    dims = [30, 40, 50]
    cidx, vals = uniform_tensor(dims, 4000, 3)
    vals = np.ceil(vals * 5)
    tns = tmp_path / "c.tns"
    _write_tns(tns, dims, cidx, vals)
    print("reference main APR", flush=True)
    r = subprocess.run([exe, "-l", "APR", "-i", str(tns), "-r", "4", "-m", "5", "-n", "3", "-x", "2", "-e", "1e-4"],
                       env=_env(), capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    o = oracle.OracleAlto(dims, cidx, vals, nprtn=4)
    U, lam = oracle.random_init(dims, 4, 2)
    res = o.cp_apr(U, lam, 5, 3, eps=1e-4)
    line = [l for l in r.stdout.splitlines() if "Final log-likelihood" in l][0]
    assert abs(float(line.split("=")[1]) - res["loglik"]) <= 1e-6 * abs(res["loglik"]), (line, res)
    assert f"Total outer iterations = {res['outer_iters']}" in r.stdout
    assert f"Total inner iterations = {res['total_inner']}" in r.stdout
    # (-l APR -p is not run: the reference's BenchmarkAlto allocates
    # factors[m] only for ALS and then frees them for every model,
    # src/main.cpp:410-416 and 507-510, so its APR bench frees pointers it
    # never set; cpd64 -l APR -p covers the APR bench on the GPU)
    print("reference main generator", flush=True)
    r = subprocess.run([exe, "-l", "ALS", "-d", "30,20,10", "-s", "0.05", "-r", "4", "-m", "3", "-x", "1", "-f", "1"],
                       cwd=tmp_path, env=_env(1), capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "Generating a sparse tensor: 30x20x10, rank = 4 ..." in r.stdout
    assert len(_its(r.stdout)) == 3
    from oracle.poisson_gen import poisson_tensor
    from tests.test_dropin_cli import _read_bin
    ci, vo = poisson_tensor([30, 20, 10], 300, 4, 1)
    d, c2, v2 = _read_bin(tmp_path / f"3D_{len(vo)}.bin")
    assert d == [30, 20, 10] and np.array_equal(c2, np.array(ci, np.uint64)) and np.array_equal(v2, vo)


@pytest.mark.gpu
@pytest.mark.parametrize("dims", [[30, 20, 10], [9, 300, 7, 11], [6, 5, 4, 7, 3]])
@pytest.mark.parametrize("R", [1, 5, 16, 33, 64, 100])
def test_coo_mttkrp_matches_oracle(dims, R):
    """alto_coo_mttkrp (host pointers) equals the oracle's COO MTTKRP for
    3-5 modes and ranks 1-100, overwriting or adding into the output;
    repeated coordinates add up."""
    import alto_amd as A

    cidx, vals = uniform_tensor(dims, 2000, R)
    cidx = np.concatenate([cidx, cidx[:, :300]], axis=1)  # 300 repeated coordinates
    vals = np.concatenate([vals, vals[:300] * 0.5])
    rng = np.random.default_rng(R)
    U = [rng.random((d, R)) for d in dims]
    for mode in range(len(dims)):
        want = oracle.mttkrp_coo(dims, cidx, vals, U, mode)
        _close(A.coo_mttkrp(dims, cidx, vals, mode, U), want, 1e-13)
        base = rng.random((dims[mode], R))
        _close(A.coo_mttkrp(dims, cidx, vals, mode, U, out=base.copy()), base + want, 1e-13)


@pytest.mark.gpu
def test_coo_mttkrp_device_pointers_and_empty():
    """alto_coo_mttkrp on device buffers (asynchronous form), with a
    nonzero whose coordinate is out of range skipped, and with no nonzeros
    (the output is zeroed or left as it was)."""
    import ctypes as C

    import alto_amd as A

    dims, R = [50, 40, 30], 24
    cidx, vals = uniform_tensor(dims, 5000, 4)
    rng = np.random.default_rng(4)
    U = [rng.random((d, R)) for d in dims]
    bad = cidx.copy()
    bad[1, 17] = dims[1] + 5  # skipped on the device
    keep = np.ones(cidx.shape[1], bool)
    keep[17] = False
    want = oracle.mttkrp_coo(dims, cidx[:, keep], vals[keep], U, 0)
    dc = [A.DeviceArray.from_numpy(bad[m]) for m in range(3)]
    dv = A.DeviceArray.from_numpy(vals)
    du = [A.DeviceArray.from_numpy(u) for u in U]
    out = A.DeviceArray.from_numpy(np.full((dims[0], R), 7.0))
    d = (C.c_uint64 * 3)(*dims)
    L = A.lib()
    ptrs = lambda xs: (C.c_void_p * len(xs))(*[x.ptr for x in xs])  # noqa: E731
    A.check(L.alto_coo_mttkrp(3, d, vals.size, ptrs(dc), dv.ptr, 0, R, ptrs(du), out.ptr, 0, A.PTR_DEVICE, None),
            "coo_mttkrp")
    A.synchronize()
    _close(out.numpy(), want, 1e-13)
    A.check(L.alto_coo_mttkrp(3, d, 0, ptrs(dc), dv.ptr, 0, R, ptrs(du), out.ptr, 1, A.PTR_DEVICE, None), "empty")
    A.synchronize()
    _close(out.numpy(), want, 1e-13)
    A.check(L.alto_coo_mttkrp(3, d, 0, ptrs(dc), dv.ptr, 0, R, ptrs(du), out.ptr, 0, A.PTR_DEVICE, None), "empty")
    A.synchronize()
    assert not out.numpy().any()
