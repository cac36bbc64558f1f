"""The C-ABI boundary: libcbspmv.so loads without a GPU, exports every
function that include/*.h declares, and its host-only entry points behave as
documented (status codes, last-error text, option validation).  No call here
touches a device."""
import ctypes
import glob
import os
import re
import subprocess

import numpy as np
import pytest

import cbspmv
from cbspmv import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = sorted(glob.glob(os.path.join(ROOT, "include", "*.h")))


def declared_functions():
    names = set()
    for h in HEADERS:
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        text = re.sub(r"//[^\n]*", "", text)
        for m in re.finditer(r"\b(cbspmv_[a-z0-9_]+)\s*\(", text):
            names.add(m.group(1))
    return names


def exported_symbols(path):
    out = subprocess.run(["nm", "-D", "--defined-only", path], capture_output=True, text=True, check=True).stdout
    return {line.split()[-1] for line in out.splitlines() if " T " in line}


def test_headers_found():
    assert any(h.endswith("cbspmv.h") for h in HEADERS)
    assert len(declared_functions()) >= 15


def test_every_declared_symbol_is_exported():
    missing = declared_functions() - exported_symbols(cbspmv.LIB_PATH)
    assert not missing, f"declared in include/*.h but not exported: {sorted(missing)}"


def test_library_loads_and_reports_version():
    L = cbspmv.lib()
    L.cbspmv_version.restype = ctypes.c_char_p
    assert L.cbspmv_version().decode().startswith("cbspmv")


def test_no_exported_symbols_outside_the_abi():
    """Only cbspmv_* names leave the library (no oracle, no torch types)."""
    syms = exported_symbols(cbspmv.LIB_PATH)
    ours = {s for s in syms if s.startswith("cbspmv_")}
    assert ours == declared_functions()
    assert not any(s.startswith("oracle_") or s.startswith("ref_") for s in syms)


def test_default_opts_are_reference_makefile_build():
    o = _lib.Opts()
    cbspmv.lib().cbspmv_default_opts(ctypes.byref(o))
    assert (o.block_size, o.waves_per_block, o.row_owned, o.pattern_wave, o.force_coo64,
            o.dense_sixteenths) == (16, 4, 1, 1, 0, 15)


def _build(m, n, r, c, v, **kw):
    o = cbspmv.make_opts(**kw)
    h = ctypes.c_void_p()
    r = np.ascontiguousarray(r, np.int32)
    c = np.ascontiguousarray(c, np.int32)
    v = np.ascontiguousarray(v, np.float64)
    rc = cbspmv.lib().cbspmv_build(m, n, r.size, r.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                   c.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                   v.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), ctypes.byref(o),
                                   ctypes.byref(h))
    return rc, h


def _last_error():
    L = cbspmv.lib()
    L.cbspmv_last_error.restype = ctypes.c_char_p
    return L.cbspmv_last_error().decode()


@pytest.mark.parametrize("kw,code", [
    (dict(block_size=12), _lib.ERR_ARG),
    (dict(waves=0), _lib.ERR_ARG),
    (dict(waves=17), _lib.ERR_ARG),
    (dict(dense_sixteenths=0), _lib.ERR_ARG),
    (dict(tile_waves=17), _lib.ERR_ARG),
    (dict(tile_kib=200), _lib.ERR_ARG),
])
def test_option_validation(kw, code):
    rc, h = _build(4, 4, [0, 1], [0, 1], [1.0, 2.0], **kw)
    assert rc == code and not h.value
    assert _last_error()


def test_status_codes_for_bad_input():
    rc, _ = _build(4, 4, [1, 0], [0, 0], [1.0, 1.0])
    assert rc == _lib.ERR_ORDER
    rc, _ = _build(4, 4, [0, 4], [0, 0], [1.0, 1.0])
    assert rc == _lib.ERR_RANGE
    rc, _ = _build(4, 4, [0, 1], [0, 9], [1.0, 1.0])
    assert rc == _lib.ERR_RANGE
    rc, _ = _build(-1, 4, [], [], [])
    assert rc == _lib.ERR_ARG
    assert "negative" in _last_error()


def test_spmv_before_upload_is_state_error_without_gpu():
    """cbspmv_spmv on a plan never uploaded fails with ERR_STATE before any
    HIP call."""
    rc, h = _build(4, 4, [0, 1], [0, 1], [1.0, 2.0])
    assert rc == 0
    L = cbspmv.lib()
    rc = L.cbspmv_spmv(h, ctypes.c_void_p(0), ctypes.c_void_p(0), ctypes.c_void_p(0))
    assert rc in (_lib.ERR_STATE, _lib.ERR_ARG)
    L.cbspmv_destroy(h)


def test_export_unknown_name():
    rc, h = _build(4, 4, [0, 1], [0, 1], [1.0, 2.0])
    cnt = ctypes.c_size_t()
    assert cbspmv.lib().cbspmv_export(h, b"no_such_array", None, ctypes.byref(cnt)) == _lib.ERR_ARG
    cbspmv.lib().cbspmv_destroy(h)


def test_reader_error_codes(tmp_path):
    bad = tmp_path / "bad.mtx"
    bad.write_text("not a matrix market file\n")
    with pytest.raises(cbspmv.CbspmvError):
        cbspmv.read_mtx(str(bad))
    with pytest.raises(cbspmv.CbspmvError):
        cbspmv.read_mtx(str(tmp_path / "missing.mtx"))
    ok = tmp_path / "sym.mtx"
    ok.write_text("%%MatrixMarket matrix coordinate real symmetric\n3 3 3\n1 1 5.0\n2 1 7.0\n3 2 9.0\n")
    m, n, r, c, v = cbspmv.read_mtx(str(ok))
    assert (m, n) == (3, 3)
    # symmetric expansion; values forced to 1; row-major
    assert sorted(zip(r.tolist(), c.tolist())) == [(0, 0), (0, 1), (1, 0), (1, 2), (2, 1)]
    assert np.all(v == 1.0)
    assert np.all(np.diff(r) >= 0)


def test_report_helpers(tmp_path):
    """writeCSV (utils.h:256-268) and the report fields (utils.h:147-201)."""
    import io
    import json

    from cbspmv import GLOBAL_BLOCK, PROGRAM_METRICS, benchmark_report_dict, coo2block_gather, print_benchmark_report
    from cbspmv import writeCSV

    r = np.array([0, 0, 1, 2], np.int32)
    c = np.array([0, 3, 1, 2], np.int32)
    v = np.ones(4)
    gb = GLOBAL_BLOCK(3, 4, 4)
    coo2block_gather(r, c, v, gb)
    met = PROGRAM_METRICS(runtime=0.5, gflops=1.25)
    f = tmp_path / "r.csv"
    writeCSV(str(f), met, "tiny", 4)
    writeCSV(str(f), met, "tiny", 4)
    assert f.read_text().splitlines() == ["tiny,4,0.5,1.25"] * 2
    d = benchmark_report_dict("tiny", 3, 4, 4, gb, met, True)
    assert json.loads(json.dumps(d))["nonzeros"] == 4 and d["validation"] is True
    buf = io.StringIO()
    print_benchmark_report("tiny", 3, 4, 4, gb, met, True, file=buf)
    assert "CB-SpMV Benchmark" in buf.getvalue() and "PASS" in buf.getvalue()


def test_vendor_comparator_library_exports():
    """bench.py's rocSPARSE comparators (CSR and BSR) are built by build()
    beside the library and export both entry points."""
    path = os.path.join(ROOT, "cb-sparse-dcu_amd", "libvendor_csr.so")
    if not os.path.exists(path):
        pytest.skip("libvendor_csr.so not built")
    assert {"vendor_csr_spmv", "vendor_bsr_spmv"} <= exported_symbols(path)
