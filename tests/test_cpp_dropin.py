"""The C++ drop-in header (include/cb_spmv.hpp): a reference-style caller
compiled with g++ against it and libcbspmv.so builds the same CB structure
as the library (CPU), and on the GPU cb_spmv returns the oracle's y."""
import json
import os
import subprocess

import numpy as np
import pytest

from cbspmv import LIB_PATH, Plan, read_mtx
from oracle import pyoracle

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
MTX = os.path.join(HERE, "golden", "1138_bus.mtx")


@pytest.fixture(scope="module")
def exe(tmp_path_factory):
    out = tmp_path_factory.mktemp("dropin") / "dropin"
    libdir = os.path.dirname(LIB_PATH)
    cmd = ["g++", "-O1", "-std=c++17", "-I", os.path.join(ROOT, "include"), os.path.join(HERE, "cpp", "dropin_main.cpp"),
           "-o", str(out), "-L", libdir, "-lcbspmv", f"-Wl,-rpath,{libdir}"]
    subprocess.run(cmd, check=True, capture_output=True)
    return str(out)


def _sum(a):
    return [int(a.size), int(a.astype(np.int64).sum())]


def test_cpp_caller_builds_same_structure(exe):
    out = subprocess.run([exe, MTX], check=True, capture_output=True, text=True).stdout
    got = json.loads(out)
    m, n, r, c, v = read_mtx(MTX)
    p = Plan(m, n, r, c, v)
    st = p.stats()
    assert (got["m"], got["n"], got["nnz"]) == (m, n, r.size)
    for k in ("gather_flag", "direct_coo_columns", "compact_coo_indices", "super_sparse_blocks", "analyzed_blocks",
              "gb_nnzb"):
        assert got[k] == st[k], k
    for name in ("gb_row_idx", "gb_col_idx", "nnz_per_blk", "type_per_blk", "segment_rows", "pattern_row_ptr",
                 "block_cols"):
        assert got[name] == _sum(p.export(name)), name
    p.close()


@pytest.mark.gpu
def test_cpp_caller_cb_spmv(exe, tmp_path, gpu):
    yfile = tmp_path / "xy.bin"
    res = subprocess.run([exe, MTX, "--run", str(yfile)], capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stderr
    assert "PASS" in res.stdout
    m, n, r, c, v = read_mtx(MTX)
    xy = np.fromfile(yfile, dtype=np.float64)
    x, y = xy[:n], xy[n:]
    assert pyoracle.rel_l2(y, pyoracle.spmv_seq(m, r, c, v, x)) <= 1e-12
