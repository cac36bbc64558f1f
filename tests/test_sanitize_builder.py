"""The CB builder, the packers and the scattered-entry split's host layout
(builder.cpp, packer.cpp, split_layout.cpp) compiled on the host with
AddressSanitizer and UndefinedBehaviorSanitizer, run over
malformed COO streams (out-of-range and negative indices, rows out of
order), edge shapes (nnz = 0, m = 0, n = 0, one entry in a partial block,
1000 duplicates of one entry, a 40,000-entry row, a full dense matrix with
zeros), the encodings' switch points (n = 4095/4096/65535/65536, nnz =
9999/10000) and a seeded fuzz of 240 small matrices over block sizes,
builders, value types, panels and tile shapes (tests/cpp/sanitize_builder.cpp).
Any out-of-bounds access or undefined behaviour aborts the run.  Host code
only: no HIP source is compiled."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "cb-sparse-dcu_amd", "csrc")


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_builder_and_packers_under_asan_ubsan(tmp_path):
    exe = tmp_path / "sanitize_builder"
    cmd = ["g++", "-std=c++17", "-g", "-O1", "-fopenmp", "-fno-omit-frame-pointer",
           "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
           f"-I{os.path.join(ROOT, 'include')}", f"-I{SRC}",
           os.path.join(ROOT, "tests", "cpp", "sanitize_builder.cpp"),
           os.path.join(SRC, "builder.cpp"), os.path.join(SRC, "packer.cpp"), os.path.join(SRC, "split_layout.cpp"),
           "-o", str(exe)]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=1", UBSAN_OPTIONS="print_stacktrace=1",
               OMP_NUM_THREADS="4")
    r = subprocess.run([str(exe)], capture_output=True, text=True, env=env, timeout=600)
    assert r.returncode == 0, (r.stdout[-3000:] + r.stderr[-6000:])
    assert "OK (0 failures)" in r.stdout
