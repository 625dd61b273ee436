"""A C++ program written against the reference's host API builds against
include/veig_compat.hpp with g++ and links libveig.so (the drop-in for a
C++ caller of the reference's host API), and the shim's file readers (utils.cpp:17-83:
readMatrix, readVector, readMatricesFromFile, loadPackages) read what the
reference's would.  No GPU needed: only host code runs."""
from __future__ import annotations

import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

REPO = Path(__file__).resolve().parent.parent
PKG = REPO / "vectorized-eigensolver_amd"

PROGRAM = r"""
#include "veig_compat.hpp"
#include <cstdio>
#include <vector>
int main(int argc, char **argv) {
  // argv[1]: 2 matrices of order 3 (18 doubles); argv[2]: work directory
  std::vector<double> a(18), m(9), v(3);
  readMatricesFromFile(3, 2, a.data(), argv[1]);
  readMatrix(3, argv[1], m.data());
  readVector(3, argv[1], v.data());
  for (double x : a) std::printf("%.17g ", x);
  std::printf("\n%.17g %.17g\n", m[8], v[2]);
  try {
    readMatricesFromFile(3, 3, a.data(), argv[1]);  // 27 doubles asked, 18 present
    std::printf("no error\n");
  } catch (const veig::Error &e) {
    std::printf("error %d\n", e.status);
  }
  std::vector<ComplexDouble> c(2 * 2);
  loadPackages(1, 2, 2, c.data());  // testcases/f_coeffs_1.csv in the working directory
  for (const ComplexDouble &z : c) std::printf("%.17g %.17g ", z.r, z.i);
  std::printf("\n");
  return 0;
}
"""


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_cpp_program_builds_and_reads_files(tmp_path):
    src = tmp_path / "app.cpp"
    src.write_text(PROGRAM)
    exe = tmp_path / "app"
    subprocess.run(["g++", "-std=c++17", f"-I{REPO / 'include'}", str(src), f"-L{PKG}", "-lveig",
                    f"-Wl,-rpath,{PKG}", "-o", str(exe)], check=True, capture_output=True)
    data = np.random.default_rng(5).standard_normal(18)
    data.tofile(tmp_path / "m.bin")
    (tmp_path / "testcases").mkdir()
    (tmp_path / "testcases" / "f_coeffs_1.csv").write_text("1,2,3,4\n5.5,-6,7e-3,8\n")
    out = subprocess.run([str(exe), str(tmp_path / "m.bin")], cwd=tmp_path, check=True,
                         capture_output=True, text=True).stdout.splitlines()
    assert np.array_equal(np.array(out[0].split(), float), data)
    m8, v2 = map(float, out[1].split())
    assert m8 == data[8] and v2 == data[2]
    assert out[2] == "error 1"
    assert np.array_equal(np.array(out[3].split(), float), [1, 2, 3, 4, 5.5, -6, 7e-3, 8])
