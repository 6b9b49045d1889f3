"""The drop-in C++ header compiles against the C ABI (CPU), and the cpd64 /
cpd128 command line reproduces the reference's check / bench / run output
lines on a .tns file (GPU)."""
import os
import subprocess
import sys

import numpy as np
import pytest

from tests._tensors import uniform_tensor

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "alto_amd", "bin")


def test_dropin_header_compiles(tmp_path):
    src = tmp_path / "use.cpp"
    src.write_text('''
#include "alto/alto.hpp"
#include "alto/apr.hpp"
#include "alto/cpd.hpp"
int main() {
    IType dims[3] = {4, 8, 2};
    KruskalModel* M; CreateKruskalModel(3, dims, 2, &M);
    AltoTensor<unsigned long long>* A64 = nullptr; AltoTensor<unsigned __int128>* A128 = nullptr;
    if (A64) { cpd_alto(A64, M, 1, 1e-5); cp_apr_alto(A128, M, 1, 1, 0.0, 0.0, 1e-10, 1e-2, true);
               FType** of; create_da_mem(-1, 2, A64, &of); mttkrp_alto_par(0, M->U, 2, A64, nullptr, of);
               int modes[1] = {0}; FType** fs[1] = {M->U}; AltoTensor<unsigned long long>* ts[1] = {A64};
               mttkrp_alto_batched(1, modes, fs, 2, ts); }
    DestroyKruskalModel(M);
    return 0;
}
''')
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-I" + os.path.join(ROOT, "include"), str(src)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def _write_tns(path, dims, nnz, seed):
    cidx, vals = uniform_tensor(dims, nnz, seed)
    with open(path, "w") as f:
        f.write("# synthetic\n")
        for i in range(nnz):
            f.write(" ".join(str(int(cidx[m, i]) + 1) for m in range(len(dims))) + f" {vals[i]:.17g}\n")
    return cidx, vals


@pytest.mark.gpu
def test_cli_check_bench_and_run(tmp_path):
    tns = tmp_path / "x.tns"
    _write_tns(tns, [30, 40, 50], 4000, 3)
    for exe in ("cpd64", "cpd128"):
        r = subprocess.run([os.path.join(BIN, exe), "-c", "-i", str(tns), "-r", "8", "-x", "1"],
                           capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stderr
        assert r.stderr.count("Results of mttkrp_alto is correct") == 3, r.stderr
    r = subprocess.run([os.path.join(BIN, "cpd64"), "-i", str(tns), "-r", "8", "-m", "5", "-p", "-t", "0", "-x", "1"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and "ALTO creation time:" in r.stdout and "ALTO runtime:" in r.stdout, r.stdout
    r = subprocess.run([os.path.join(BIN, "cpd64"), "-i", str(tns), "-r", "8", "-m", "6", "-x", "1", "-e", "1e-12"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    its = [l for l in r.stdout.splitlines() if l.startswith("it: ")]
    assert len(its) == 6 and "Per iteration" in r.stdout
    # the fits printed (%g, 6 significant digits) are the oracle's CP-ALS fits
    import oracle
    cidx, vals = uniform_tensor([30, 40, 50], 4000, 3)
    o = oracle.OracleAlto([30, 40, 50], cidx, vals, nprtn=16)
    U, lam = oracle.random_init([30, 40, 50], 8, 1)
    _, fits = o.cpd_als(U, lam, 6, 1e-12)
    got = [float(l.split("\t")[1].split()[1]) for l in its]
    assert all(abs(g - f) <= 5e-6 * max(1.0, abs(f)) for g, f in zip(got, fits)), (got, fits)
    r = subprocess.run([os.path.join(BIN, "cpd64"), "-l", "APR", "-i", str(tns), "-r", "4", "-m", "3", "-n", "2",
                        "-x", "1"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and "Final log-likelihood" in r.stdout, r.stdout + r.stderr
    U, lam = oracle.random_init([30, 40, 50], 4, 1)
    res = o.cp_apr(U, lam, 3, 2, eps=1e-5)
    line = [l for l in r.stdout.splitlines() if "Final log-likelihood" in l][0]
    assert abs(float(line.split("=")[1]) - res["loglik"]) <= 1e-6 * abs(res["loglik"]), (line, res)
    assert f"Total inner iterations = {res['total_inner']}" in r.stdout
    # -f: the text input is also written as <N>D_<nnz>.bin; running from that
    # binary file gives the same fit trace
    runs = []
    for args in (["-i", str(tns), "-f", "1"], ["-b", "3D_4000.bin"]):
        r = subprocess.run([os.path.join(BIN, "cpd64")] + args + ["-r", "8", "-m", "4", "-x", "1"], cwd=tmp_path,
                           capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stderr
        runs.append([l.split("\t")[1] for l in r.stdout.splitlines() if l.startswith("it: ")])
    assert (tmp_path / "3D_4000.bin").exists()
    assert len(runs[0]) == 4 and runs[0] == runs[1]
    # -d / -s: the reference's Poisson generator (one thread), then CP-ALS
    from oracle.poisson_gen import poisson_tensor
    env = dict(os.environ, OMP_NUM_THREADS="1")
    r = subprocess.run([os.path.join(BIN, "cpd64"), "-d", "30,20,10", "-s", "0.05", "-r", "4", "-m", "3", "-x", "1",
                        "-f", "1"], cwd=tmp_path, env=env, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "Generating a sparse tensor: 30x20x10, rank = 4 ..." in r.stdout
    assert len([l for l in r.stdout.splitlines() if l.startswith("it: ")]) == 3
    ci, vo = poisson_tensor([30, 20, 10], 300, 4, 1)
    d, cidx, vals = _read_bin(tmp_path / f"3D_{len(vo)}.bin")
    assert d == [30, 20, 10] and np.array_equal(cidx, np.array(ci, np.uint64)) and np.array_equal(vals, vo)


_IO_PROG = r'''
#include "alto/sptensor.hpp"
// t/z: read .tns (one-/zero-based); b: read binary; w: .tns -> binary under
// the export name; x: binary -> .tns under the export name.  Prints the
// tensor (nmodes nnz dims, then one line per nonzero) or the written name.
int main(int argc, char** argv) {
    const char k = argv[1][0];
    SparseTensor* X;
    ImportSparseTensor(argv[2], (k == 'b' || k == 'x') ? BINARY_FORMAT : TEXT_FORMAT, &X, k == 'z');
    if (k == 'w' || k == 'x') {
        ExportSparseTensor(NULL, k == 'w' ? BINARY_FORMAT : TEXT_FORMAT, X);
        printf("%s\n", alto_io::export_name(X, k == 'w').c_str());
        DestroySparseTensor(X);
        return 0;
    }
    printf("%d %llu", X->nmodes, (unsigned long long)X->nnz);
    for (int m = 0; m < X->nmodes; ++m) printf(" %llu", (unsigned long long)X->dims[m]);
    printf("\n");
    for (IType i = 0; i < X->nnz; ++i) {
        for (int m = 0; m < X->nmodes; ++m) printf("%llu ", (unsigned long long)X->cidx[m][i]);
        printf("%.17g\n", X->vals[i]);
    }
    DestroySparseTensor(X);
    return 0;
}
'''


def _io_tool(tmp_path):
    src = tmp_path / "io.cpp"
    src.write_text(_IO_PROG)
    exe = tmp_path / "io"
    r = subprocess.run(["g++", "-std=c++17", "-O1", "-I" + os.path.join(ROOT, "include"),
                        str(src), "-o", str(exe)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return exe


def _io_read(exe, kind, path, cwd):
    r = subprocess.run([str(exe), kind, str(path)], capture_output=True, text=True, cwd=cwd)
    assert r.returncode == 0, r.stderr
    lines = r.stdout.splitlines()
    head = [int(x) for x in lines[0].split()]
    nm, nnz, dims = head[0], head[1], head[2:]
    rows = [l.split() for l in lines[1:]]
    cidx = np.array([[int(r[m]) for r in rows] for m in range(nm)], dtype=np.uint64)
    vals = np.array([float(r[nm]) for r in rows])
    return nm, nnz, dims, cidx, vals


def test_tensor_io_formats(tmp_path):
    """The command line's tensor files follow the reference (src/sptensor.cpp):
    .tns dims are the largest index (one-based) or largest + 1 (-z), '#' and
    empty lines are skipped; the binary file round-trips exactly, its 8-byte
    mode count is read through its low 32 bits, and exports are named
    <N>D_<nnz>.bin / .tns with "%g" text values."""
    import struct
    exe = _io_tool(tmp_path)
    tns = tmp_path / "a.tns"
    tns.write_text("# comment\n\n1 2 3 1.5\n4 1 2 0.125\n\n# another\n2 7 1 3.0000000000000004\n")
    nm, nnz, dims, cidx, vals = _io_read(exe, "t", tns, tmp_path)
    assert (nm, nnz, dims) == (3, 3, [4, 7, 3])
    np.testing.assert_array_equal(cidx, [[0, 3, 1], [1, 0, 6], [2, 1, 0]])
    np.testing.assert_array_equal(vals, [1.5, 0.125, 3.0000000000000004])
    _, _, dims_z, cidx_z, _ = _io_read(exe, "z", tns, tmp_path)
    assert dims_z == [5, 8, 4]
    np.testing.assert_array_equal(cidx_z, cidx + 1)
    # .tns -> binary under the reference's export name, read back exactly
    r = subprocess.run([str(exe), "w", str(tns)], capture_output=True, text=True, cwd=tmp_path)
    assert r.returncode == 0 and r.stdout.strip() == "3D_3.bin", r.stdout + r.stderr
    _, _, dims_b, cidx_b, vals_b = _io_read(exe, "b", tmp_path / "3D_3.bin", tmp_path)
    assert dims_b == dims and np.array_equal(cidx_b, cidx) and np.array_equal(vals_b, vals)
    # a binary file whose mode count carries garbage in its upper 4 bytes
    d = np.array([4, 7, 3], np.uint64)
    blob = struct.pack("<Q", (0xDEADBEEF << 32) | 3) + d.tobytes() + struct.pack("<Q", 3)
    blob += cidx.astype(np.uint64).tobytes() + vals.tobytes()
    (tmp_path / "g.bin").write_bytes(blob)
    nm, nnz, dims_g, cidx_g, vals_g = _io_read(exe, "b", tmp_path / "g.bin", tmp_path)
    assert (nm, nnz, dims_g) == (3, 3, [4, 7, 3]) and np.array_equal(cidx_g, cidx) and np.array_equal(vals_g, vals)
    # binary -> .tns: one-based indices, "%g" values
    r = subprocess.run([str(exe), "x", str(tmp_path / "g.bin")], capture_output=True, text=True, cwd=tmp_path)
    assert r.returncode == 0 and r.stdout.strip() == "3D_3.tns"
    assert (tmp_path / "3D_3.tns").read_text().splitlines() == ["1 2 3 1.5", "4 1 2 0.125", "2 7 1 3"]


_GEN_PROG = r'''
#include "alto/poisson_generator.hpp"
// edges rank threads dims...: the -d / -s generator, written as g.bin
int main(int argc, char** argv) {
    std::vector<IType> dims;
    for (int i = 4; i < argc; ++i) dims.push_back(atoll(argv[i]));
    SparseTensor* X = alto_gen::poisson_tensor((int)dims.size(), dims.data(), atoll(argv[1]), atoll(argv[2]),
                                               atoi(argv[3]));
    ExportSparseTensor("g.bin", BINARY_FORMAT, X);
    DestroySparseTensor(X);
    return 0;
}
'''


def _read_bin(path):
    import struct
    b = open(path, "rb").read()
    nm = struct.unpack("<Q", b[:8])[0] & 0xFFFFFFFF
    dims = np.frombuffer(b[8:8 + 8 * nm], np.uint64)
    nnz = struct.unpack("<Q", b[8 + 8 * nm:16 + 8 * nm])[0]
    off = 16 + 8 * nm
    cidx = np.frombuffer(b[off:off + 8 * nm * nnz], np.uint64).reshape(nm, nnz)
    vals = np.frombuffer(b[off + 8 * nm * nnz:], np.float64)
    return [int(d) for d in dims], cidx, vals


@pytest.mark.parametrize("dims,edges,rank,threads", [([30, 20, 10], 300, 4, 1), ([40, 35, 30, 25], 2000, 5, 3),
                                                     ([60, 50, 40], 120000, 3, 2)])
def test_poisson_generator_matches_oracle(tmp_path, dims, edges, rank, threads):
    """The -d / -s generator equals the oracle's restatement of the
    reference's (same MT19937-64 streams per thread, glibc rand(), batches of
    100000 draws) bit for bit; the counts add up to the draws, coordinates
    are in range and sorted with the last mode most significant."""
    from oracle.poisson_gen import poisson_tensor
    src = tmp_path / "g.cpp"
    src.write_text(_GEN_PROG)
    exe = tmp_path / "g"
    r = subprocess.run(["g++", "-std=c++17", "-O2", "-I" + os.path.join(ROOT, "include"),
                        str(src), "-o", str(exe)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([str(exe), str(edges), str(rank), str(threads)] + [str(d) for d in dims], cwd=tmp_path,
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert r.stdout.startswith("Generating a sparse tensor: " + "x".join(map(str, dims)) + f", rank = {rank} ...")
    d, cidx, vals = _read_bin(tmp_path / "g.bin")
    ci, vo = poisson_tensor(dims, edges, rank, threads)
    assert d == dims
    np.testing.assert_array_equal(cidx, np.array(ci, np.uint64))
    np.testing.assert_array_equal(vals, np.array(vo))
    assert vals.sum() == edges and (vals >= 1).all()
    assert all((cidx[m] < dims[m]).all() for m in range(len(dims)))
    key = [tuple(int(cidx[m, i]) for m in reversed(range(len(dims)))) for i in range(cidx.shape[1])]
    assert all(a < b for a, b in zip(key, key[1:]))


_BATCH_PROG = r'''
#include <math.h>
#include "alto/alto.hpp"
// three tensors of 3-4 modes (LCG coordinates, duplicates allowed), rank 8:
// mttkrp_alto_batched equals mttkrp_alto_par per tensor and brute force
static unsigned long long s = 12345;
static unsigned long long nxt() { s = s * 6364136223846793005ull + 1442695040888963407ull; return s >> 33; }
int main() {
    const int J = 3, R = 8;
    IType dimsv[J][4] = {{30, 40, 50, 0}, {7, 300, 20, 9}, {64, 64, 64, 0}};
    int nm[J] = {3, 4, 3}, modes[J] = {0, 2, 1};
    IType nnzv[J] = {4000, 3000, 20000};
    SparseTensor spt[J];
    AltoTensor<unsigned long long>* at[J];
    FType** A[J]; FType** B[J];
    for (int j = 0; j < J; ++j) {
        spt[j].nmodes = nm[j]; spt[j].dims = dimsv[j]; spt[j].nnz = nnzv[j];
        spt[j].cidx = new IType*[nm[j]]; spt[j].vals = new FType[nnzv[j]];
        for (int m = 0; m < nm[j]; ++m) {
            spt[j].cidx[m] = new IType[nnzv[j]];
            for (IType i = 0; i < nnzv[j]; ++i) spt[j].cidx[m][i] = nxt() % dimsv[j][m];
        }
        for (IType i = 0; i < nnzv[j]; ++i) spt[j].vals[i] = 0.5 + (nxt() % 1000) / 1000.0;
        create_alto(&spt[j], &at[j], 4);
        A[j] = new FType*[nm[j]]; B[j] = new FType*[nm[j]];
        for (int m = 0; m < nm[j]; ++m) {
            A[j][m] = new FType[dimsv[j][m] * R]; B[j][m] = new FType[dimsv[j][m] * R];
            for (IType k = 0; k < dimsv[j][m] * R; ++k) A[j][m][k] = B[j][m][k] = (nxt() % 1000) / 1000.0;
        }
    }
    FType* C0[J];  // the target factors before the calls
    for (int j = 0; j < J; ++j) {
        const IType n = dimsv[j][modes[j]] * R;
        C0[j] = new FType[n];
        for (IType k = 0; k < n; ++k) C0[j][k] = A[j][modes[j]][k];
    }
    for (int j = 0; j < J; ++j) mttkrp_alto_par(modes[j], A[j], R, at[j], nullptr, nullptr);
    mttkrp_alto_batched(J, modes, B, R, at);
    double worst = 0;
    for (int j = 0; j < J; ++j) {
        const int t = modes[j];
        const IType I = dimsv[j][t];
        // brute force over the coordinates, added to the initial values
        double* bf = new double[I * R]();
        for (IType i = 0; i < nnzv[j]; ++i)
            for (int r = 0; r < R; ++r) {
                double p = spt[j].vals[i];
                for (int m = 0; m < nm[j]; ++m)
                    if (m != t) p *= A[j][m][spt[j].cidx[m][i] * R + r];
                bf[spt[j].cidx[t][i] * R + r] += p;
            }
        double mx = 0;
        for (IType k = 0; k < I * R; ++k) mx = fmax(mx, fabs(bf[k]));
        for (IType k = 0; k < I * R; ++k) {
            worst = fmax(worst, fabs(A[j][t][k] - B[j][t][k]) / mx);
            worst = fmax(worst, fabs(B[j][t][k] - C0[j][k] - bf[k]) / mx);
        }
        delete[] bf;
    }
    printf("max rel diff batched vs per-tensor %.3e\n", worst);
    for (int j = 0; j < J; ++j) destroy_alto(at[j]);
    return worst < 1e-12 ? 0 : 1;
}
'''


@pytest.mark.gpu
def test_dropin_batched_mttkrp_matches_per_tensor(tmp_path):
    """The C++ drop-in mttkrp_alto_batched over three tensors gives what
    mttkrp_alto_par gives per tensor and what brute force gives (both add
    into the target factor)."""
    src = tmp_path / "batch.cpp"
    src.write_text(_BATCH_PROG)
    exe = tmp_path / "batch"
    libdir = os.path.join(ROOT, "alto_amd")
    r = subprocess.run(["g++", "-std=c++17", "-O1", "-I" + os.path.join(ROOT, "include"), str(src), "-o", str(exe),
                        "-L" + libdir, "-lalto_hip", "-Wl,-rpath," + libdir], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "max rel diff" in r.stdout
