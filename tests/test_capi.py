"""The C-ABI libraries load and export every declared symbol; host-side logic of the
device layout.  No GPU calls."""
import ctypes
import os
import re

import numpy as np
import pytest

import oracle
from cicy_geom_amd import dea, network

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(cicy_[a-z0-9_]+)\s*\(", text)))


def test_dea_header_symbols_exported():
    names = declared("cicy_dea.h")
    assert len(names) >= 16
    lib = ctypes.CDLL(os.path.join(ROOT, "cicy_geom_amd", "libcicy_dea.so"))
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert sorted(dea.EXPORTED_SYMBOLS) == names
    assert sorted(dea.ASSEMBLY_SYMBOLS) == declared("cicy_assembly.h")


def test_all_headers_exported():
    libs = {"cicy_dea.h": "libcicy_dea.so", "cicy_assembly.h": "libcicy_dea.so",
            "cicy_netio.h": "libcicy_net.so", "cicy_clusters.h": "libcicy_net.so"}
    headers = [h for h in os.listdir(os.path.join(ROOT, "include")) if h.endswith(".h")]
    assert sorted(headers) == sorted(libs)  # every C header is checked against its library
    for h in headers:
        lib = ctypes.CDLL(os.path.join(ROOT, "cicy_geom_amd", libs[h]))
        names = declared(h)
        assert names, h
        for n in names:
            assert hasattr(lib, n), (h, n)


@pytest.mark.parametrize("fmt", [dea.FORMAT_SSS, dea.FORMAT_BAND])
@pytest.mark.parametrize("rb", [128, 256, 512, 1024])
def test_device_layouts_on_host(rb, fmt):
    s = network.largest_percolating_system(network.generate(1), 0)
    v = np.random.default_rng(rb).standard_normal(s.n)
    got = dea.host_blocked_spmv(s.row_ptr, s.col_ind, s.values, s.diag, v, rb, fmt)
    want = oracle.spmv_sss(s.row_ptr, s.col_ind, s.values, s.diag, v)
    assert np.allclose(got, want, rtol=1e-13, atol=1e-15 * np.abs(want).max())


def test_blocked_layout_rejects_bad_input():
    rp = np.array([0, 0, 1, 2], np.int64)
    col = np.array([0, 2], np.int64)       # row 2 refers to column 2: not strictly lower
    with pytest.raises(dea.DEAError):
        dea.host_blocked_spmv(rp, col, np.array([-1.0, -1.0]), np.ones(3), np.ones(3))
    with pytest.raises(dea.DEAError):
        dea.host_blocked_spmv(np.array([0, 0, 1], np.int64), np.array([0], np.int64),
                              np.array([-1.0]), np.array([1.0, 0.0]), np.ones(2))


def test_context_needs_gpu_or_raises():
    """Without a device the context refuses to start: there is no CPU fallback."""
    try:
        import torch  # noqa: F401  (only to learn whether a GPU is visible)
        has_gpu = torch.cuda.is_available()
    except Exception:
        has_gpu = False
    if has_gpu:
        pytest.skip("a GPU is present")
    with pytest.raises(dea.DEAError):
        dea.Context(0)


@pytest.mark.parametrize("fmt", [dea.FORMAT_SSS, dea.FORMAT_BAND])
def test_layouts_on_unstructured_matrix(fmt):
    """A matrix with no sub-diagonal structure (random junction graph) in either layout."""
    rng = np.random.default_rng(4)
    n = 3000
    pairs = {}
    for a, b in rng.integers(0, n, (6000, 2)):
        if a != b:
            pairs[(max(a, b), min(a, b))] = -rng.uniform(0.1, 1)
    keys = sorted(pairs)
    rp = np.zeros(n + 1, np.int64)
    for r, _ in keys:
        rp[r + 1] += 1
    rp = np.cumsum(rp)
    col = np.array([c for _, c in keys], np.int64)
    val = np.array([pairs[k] for k in keys])
    diag = np.full(n, 10.0)
    v = rng.standard_normal(n)
    got = dea.host_blocked_spmv(rp, col, val, diag, v, 256, fmt)
    want = oracle.spmv_sss(rp, col, val, diag, v)
    assert np.allclose(got, want, rtol=1e-13, atol=1e-13)


def test_negative_row_ptr_at_thread_boundaries_is_rejected():
    """build_blocked checks rows on host threads (up to 16, when n >= 65536).  A row_ptr
    entry that is negative where a thread's range starts must be rejected before any entry
    is read: each thread checks row_ptr[r] >= 0 for its own rows, so no thread reads
    col_ind at a negative index while the neighbouring thread reports the error.  The
    entry is set at every start a thread count of 2..16 would use, with a magnitude that
    would fault on any read through it."""
    n = 4 * 65536
    rp = np.arange(n + 1, dtype=np.int64)
    rp[0] = 0
    rp[1] = 0                      # row 0 has no entry; row r > 0 holds (r, r-1)
    rp[2:] = np.arange(1, n, dtype=np.int64)
    nnz = int(rp[-1])
    col = np.arange(nnz, dtype=np.int64)   # entry j belongs to row j + 1: column j
    val = -np.ones(nnz)
    diag = np.full(n, 4.0)
    v = np.ones(n)
    assert np.allclose(dea.host_blocked_spmv(rp, col, val, diag, v), oracle.spmv_sss(
        rp, col, val, diag, v))
    bad = rp.copy()
    for nth in range(2, 17):
        for t in range(1, nth):
            bad[n * t // nth] = -(1 << 40)
    with pytest.raises(dea.DEAError):
        dea.host_blocked_spmv(bad, col, val, diag, v)
