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
int main() {
    IType dims[3] = {4, 8, 2};
    KruskalModel* M; CreateKruskalModel(3, dims, 2, &M);
    AltoTensor<unsigned long long>* A64 = nullptr; AltoTensor<unsigned __int128>* A128 = nullptr;
    if (A64) { cpd_alto(A64, M, 1, 1e-5); cp_apr_alto(A128, M, 1, 1, 0.0, 0.0, 1e-10, 1e-2, true);
               FType** of; create_da_mem(-1, 2, A64, &of); mttkrp_alto_par(0, M->U, 2, A64, nullptr, of); }
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
    r = subprocess.run([os.path.join(BIN, "cpd64"), "-l", "APR", "-i", str(tns), "-r", "4", "-m", "3", "-n", "2",
                        "-x", "1"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and "Final log-likelihood" in r.stdout, r.stdout + r.stderr
