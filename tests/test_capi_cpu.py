"""C-ABI checks that need no GPU: the library loads, exports every symbol the
header declares, and keeps the reference's error conventions
(src/cfunction_interface.jl; test/runtests.jl:3244-3703)."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "draugr_hip.h")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(draugr_\w+|init_julia|shutdown_julia)\s*\(", text)))


def test_header_symbols_exported(dh):
    syms = declared_symbols()
    assert len(syms) >= 20
    L = ctypes.CDLL(dh.LIB_PATH)
    missing = [s for s in syms if not hasattr(L, s)]
    assert not missing, missing
    # every declared symbol has a ctypes signature in the host mirror
    from draugr_hip import _lib
    assert set(syms) == set(_lib.SIGNATURES)


def test_config_json_defaults_and_errors(dh):
    L = dh.lib()
    assert L.draugr_amg_config_from_json(None) > 0
    assert L.draugr_amg_config_from_json(b"{}") > 0
    # unknown keys ignored, string-encoded numbers accepted, names case-insensitive
    h = L.draugr_amg_config_from_json(b'{"foo": 1, "theta": "0.3", "max_levels": "7", "smoother": "JACOBI",'
                                      b' "cycle": "W", "coarse_solve_on_cpu": "true", "gpu_setup": true,'
                                      b' "kernel": "pipe", "graphs": false}')
    assert h > 0
    assert L.draugr_amg_config_free(h) == 0
    assert L.draugr_amg_config_free(h) == -1
    assert "Invalid config handle" in dh.last_error()
    # graphs: a boolean, a string-encoded boolean or "auto"
    for g in (b'true', b'false', b'"true"', b'"FALSE"', b'"auto"', b'"1"'):
        h = L.draugr_amg_config_from_json(b'{"graphs": ' + g + b'}')
        assert h > 0, (g, dh.last_error())
        assert L.draugr_amg_config_free(h) == 0
    for bad, msg in [(b'{"graphs": "sometimes"}', "Unknown graphs"),
                     (b'{"coarsening": "foo"}', "Unknown coarsening"),
                     (b'{"smoother": "bar"}', "Unknown smoother"),
                     (b'{"interpolation": "baz"}', "Unknown interpolation"),
                     (b'{"cycle": "x"}', "Unknown cycle"),
                     (b'{"strength": "weird"}', "Unknown strength"),
                     (b'{"precision": "fp8"}', "Unknown precision"),
                     (b'{"kernel": "csr"}', "Unknown kernel"),
                     (b'{"theta": "abc"}', "cannot parse"),
                     (b'{"theta": 0.5', "JSON parse error"),
                     (b'[1, 2]', "top level")]:
        assert L.draugr_amg_config_from_json(bad) == -1
        assert msg in dh.last_error(), (bad, dh.last_error())
    # a successful call clears the message
    assert L.draugr_amg_config_from_json(b"{}") > 0
    assert dh.last_error() == ""


def _csr_args(A):
    p = lambda a: a.ctypes.data_as(ctypes.c_void_p)  # noqa: E731
    return A.n, A.nnz, p(A.rowptr), p(A.colval), p(A.nzval)


def test_invalid_handles_and_index_base(dh):
    L = dh.lib()
    A = dh.problems.poisson1d(10)
    c = L.draugr_amg_config_from_json(b'{"coarsening": "aggregation", "smoother": "jacobi"}')
    x = np.zeros(10)
    xp = x.ctypes.data_as(ctypes.c_void_p)
    assert L.draugr_amg_solve(9999, 10, xp, xp, c, 1e-8, 10) == -1
    assert "Invalid handle" in dh.last_error()
    assert L.draugr_amg_cycle(9999, 10, xp, xp, c) == -1
    assert L.draugr_amg_free(9999) == -1
    assert "Invalid hierarchy handle" in dh.last_error()
    assert L.draugr_hip_setup_host(*_csr_args(A), c, 2, 1) == -1
    assert "index_base must be 0 or 1, got 2" in dh.last_error()
    assert L.draugr_hip_setup_host(*_csr_args(A), 4242, 0, 1) == -1
    assert "Invalid config handle" in dh.last_error()
    # 1-based input gives the same hierarchy as 0-based input
    h0 = L.draugr_hip_setup_host(*_csr_args(A), c, 0, 1)
    A1 = dh.problems.CSR(A.rowptr + 1, A.colval + 1, A.nzval)
    h1 = L.draugr_hip_setup_host(*_csr_args(A1), c, 1, 1)
    assert h0 > 0 and h1 > 0
    H0 = dh.AMGHierarchy(h0, None, 10, False)
    H1 = dh.AMGHierarchy(h1, None, 10, False)
    for l in range(H0.nlevels):
        a, b = H0.level(l), H1.level(l)
        for k in a:
            assert np.array_equal(a[k], b[k])
    # solving on a host-only hierarchy is a clean error
    assert L.draugr_amg_cycle(h0, 10, xp, xp, c) == -1
    assert "not uploaded" in dh.last_error()
    # malformed CSR
    bad = dh.problems.CSR(A.rowptr.copy(), A.colval.copy(), A.nzval)
    bad.colval[3] = 50
    assert L.draugr_hip_setup_host(*_csr_args(bad), c, 0, 1) == -1
    assert "column index out of range" in dh.last_error()
    H0.free()
    with pytest.raises(dh.DraugrError):
        H0.free() if False else dh._lib.check(L.draugr_amg_free(h0), "free")


def test_device_count_without_gpu_is_clean(dh):
    assert dh.device_count() >= 0


def test_runtime_lifecycle_entry_points(dh):
    """init_julia / shutdown_julia (c_api/draugr_amg.h:26-36) are exported:
    init_julia is a no-op; shutdown_julia releases every open handle, after
    which the old handles are invalid.  Run in a child process so that the
    shutdown does not touch this process's handles."""
    import subprocess
    import sys
    L = ctypes.CDLL(dh.LIB_PATH)
    assert hasattr(L, "init_julia") and hasattr(L, "shutdown_julia")
    code = r'''
import ctypes, sys
sys.path[:0] = [%r, %r]
import numpy as np
import draugr_hip as dh
L = dh.lib()
L.init_julia(0, None)
c = L.draugr_amg_config_from_json(b'{"coarsening": "aggregation", "smoother": "jacobi"}')
A = dh.problems.poisson2d(10)
h = L.draugr_hip_setup_host(A.n, A.nnz, A.rowptr.ctypes.data_as(ctypes.c_void_p),
                            A.colval.ctypes.data_as(ctypes.c_void_p), A.nzval.ctypes.data_as(ctypes.c_void_p), c, 0, 1)
assert c > 0 and h > 0
out = (ctypes.c_int64 * 8)()
assert L.draugr_hip_hierarchy_info(h, out, 8) > 0
L.shutdown_julia(0)
assert L.draugr_hip_hierarchy_info(h, out, 8) == -1 and "Invalid" in dh.last_error()
assert L.draugr_amg_config_free(c) == -1 and "Invalid config handle" in dh.last_error()
assert L.draugr_amg_free(h) == -1
c2 = L.draugr_amg_config_from_json(b"{}")
assert c2 > h  # handles are never reused
print("lifecycle ok")
''' % (ROOT, os.path.join(ROOT, "draugr.jl_amd"))
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and "lifecycle ok" in r.stdout, r.stdout + r.stderr
