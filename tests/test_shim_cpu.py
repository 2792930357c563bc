"""The PdhgDevice shim (samaya_amd/csrc/shim/pdhg_core_hip.cpp), the drop-in that lets samaya's own
host loop call this library: compiled against samaya's own pdhg_core.hpp and SparseMatrix
(src/gpu/pdhg_core.hpp:24-79, include/samaya/sparse_matrix.hpp) with the host compiler, linked
with libsamaya_amd.so and run. Without a GPU, PdhgDevice::create must return null (the
reference's "no usable GPU" contract, src/gpu/pdhg_core.cu:246-247) rather than throw, so
samaya's pdlp.cpp:343-360 takes its CPU path. Needs the reference tree (skipped without it)."""
import os
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
REF = Path("/root/reference")
SHIM = ROOT / "samaya_amd" / "csrc" / "shim" / "pdhg_core_hip.cpp"

MAIN = r"""
#include <cstdio>
#include <vector>
#include "gpu/pdhg_core.hpp"
int main() {
  using namespace samaya;
  // min x0 + x1 s.t. x0 + 2 x1 = r0, x1 - x2 = r1 (2 x 3), its transpose, n + m = 5 entries.
  std::vector<Triplet> t = {{0, 0, 1.0}, {0, 1, 2.0}, {1, 1, 1.0}, {1, 2, -1.0}};
  const SparseMatrix A = SparseMatrix::from_triplets(2, 3, t);
  const SparseMatrix At = A.transpose();
  const std::vector<double> cost = {1, 1, 0, 0, 0}, lower = {0, 0, 0, 1, 0}, upper = {5, 5, 5, 1, 0};
  auto dev = gpu::PdhgDevice::create(A, At, cost, lower, upper);
  std::printf("%s\n", dev ? "device" : "null");
  return 0;
}
"""


@pytest.mark.skipif(not (REF / "src" / "gpu" / "pdhg_core.hpp").exists(),
                    reason="the reference tree (samaya's headers) is not present")
def test_shim_builds_against_the_reference_headers_and_reports_no_device(tmp_path):
    import samaya_amd as smy
    smy.load()
    lib_dir = ROOT / "samaya_amd" / "_lib"
    main = tmp_path / "main.cpp"
    main.write_text(MAIN)
    exe = tmp_path / "shim_main"
    cmd = ["g++", "-std=c++20", "-O1", f"-I{REF / 'src'}", f"-I{REF / 'include'}",
           f"-I{ROOT / 'include'}", str(main), str(SHIM), str(REF / "src" / "linalg" / "sparse_matrix.cpp"),
           f"-L{lib_dir}", "-lsamaya_amd", f"-Wl,-rpath,{lib_dir}", "-Wl,-rpath,/opt/rocm/lib",
           "-o", str(exe)]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    env = dict(os.environ)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True, env=env, timeout=120)
    import torch
    want = "device" if torch.cuda.is_available() else "null"
    assert out.stdout.strip() == want
