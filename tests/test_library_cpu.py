"""The C-ABI library without a GPU: it loads, exports every entry point of include/xcquad.h, and answers the host-only
calls (estimated points, memory estimate, argument errors with the reference's Status codes)."""

import ctypes
import os
import re
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols(header="xcquad.h", prefix="xq_"):
    text = open(os.path.join(ROOT, "include", header)).read()
    return sorted(set(re.findall(r"\b(" + prefix + r"[a-z_0-9]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    from pdynamo3_amd import _lib
    names = declared_symbols()
    assert len(names) >= 10
    for name in names:
        assert hasattr(_lib.lib, name), name
    assert sorted(_lib.SIGNATURES) == names
    scf = declared_symbols("qcscf.h", "qs_")
    assert len(scf) >= 10
    for name in scf:
        assert hasattr(_lib.lib, name), name
    assert sorted(_lib.SCF_SIGNATURES) == scf


def test_estimated_points_match_reference_count():
    from pdynamo3_amd import integrator
    # DFTGrid_EstimatedPoints for water on the Medium grid: O 13,450 + 2 x H 14,318 (no weight cull).
    assert integrator.estimated_points(2, [8, 1, 1]) == 42086
    assert integrator.estimated_points(2, [8, 1, 1] * 16) == 16 * 42086


def test_argument_errors_return_status_codes():
    from pdynamo3_amd import _lib
    h = ctypes.c_void_p()
    z = np.array([8, 1, 1], dtype=np.int32)
    x = np.zeros(9)
    assert _lib.lib.xq_grid_build(7, 3, _lib.iptr(z), _lib.dptr(x), ctypes.byref(h)) == 3
    assert "invalid" in _lib.last_error()
    assert _lib.lib.xq_grid_build(2, 0, _lib.iptr(z), _lib.dptr(x), ctypes.byref(h)) == 3
    bad = np.array([8, 0, 1], dtype=np.int32)
    assert _lib.lib.xq_grid_build(2, 3, _lib.iptr(bad), _lib.dptr(x), ctypes.byref(h)) == 3
    nanx = x.copy()
    nanx[4] = np.nan
    assert _lib.lib.xq_grid_build(2, 3, _lib.iptr(z), _lib.dptr(nanx), ctypes.byref(h)) == 3
    e, n = ctypes.c_double(1.0), ctypes.c_double(1.0)
    ids = np.array([106, 131], dtype=np.int32)
    assert _lib.lib.xq_integrate(None, 2, _lib.iptr(ids), _lib.dptr(x), ctypes.byref(e), ctypes.byref(n),
                                 _lib.dptr(x), None) == 3
    # The unrestricted entry points and the gradients check their arguments before any device work.
    assert _lib.lib.xq_integrate_uks(None, 2, _lib.iptr(ids), _lib.dptr(x), _lib.dptr(x), ctypes.byref(e),
                                     ctypes.byref(n), _lib.dptr(x), _lib.dptr(x), None) == 3
    assert _lib.lib.xq_gradients(None, 2, _lib.iptr(ids), _lib.dptr(x), _lib.dptr(x)) == 3
    assert _lib.lib.xq_gradients_uks(None, 2, _lib.iptr(ids), _lib.dptr(x), _lib.dptr(x), _lib.dptr(x)) == 3
    assert "invalid" in _lib.last_error()
    # The SCF's dense steps refuse an empty basis or a missing handle before any device work.
    h = ctypes.c_void_p()
    assert _lib.lib.qs_dense_create(0, _lib.dptr(x), _lib.dptr(x), ctypes.byref(h)) == 3 and not h
    assert _lib.lib.qs_dense_aufbau(None, _lib.dptr(x), 1, _lib.dptr(x), None) == 3
    assert _lib.lib.qs_dense_commutator(None, _lib.dptr(x), _lib.dptr(x), _lib.dptr(x)) == 3


def test_memory_estimate_scales():
    from pdynamo3_amd import integrator
    small = integrator.estimate_bytes(42086, 19, 19.0, "fp64")
    big = integrator.estimate_bytes(16 * 42086, 304, 200.0, "fp64")
    mixed = integrator.estimate_bytes(16 * 42086, 304, 200.0, "mixed")
    assert 0 < small < big
    # The estimate is the peak: a set-up holds the FP64 rows (32 B per stored value) while it writes the operand planes
    # of its precision (48 B per value for FP64, 21 B for the digit planes).
    values = (16 * 42086 / 128) * 200.0 * 128
    assert 32 * values + 48 * values < big < 2 * (32 + 48) * values
    assert 32 * values + 21 * values < mixed < big


def test_batch_argument_errors_return_status_codes():
    from pdynamo3_amd import _lib
    b = ctypes.c_void_p()
    assert _lib.lib.xq_batch_create(0, ctypes.byref(b)) == 3
    assert _lib.lib.xq_batch_create(1, None) == 3
    one = np.zeros(1)
    ids = np.zeros(4, dtype=np.int32)
    n = np.ones(1, dtype=np.int32)
    assert _lib.lib.xq_integrate_batch(None, 1, None, _lib.iptr(n), _lib.iptr(ids), None, _lib.dptr(one), _lib.dptr(one),
                                       None, None) == 3
    assert "invalid" in _lib.last_error()


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True, text=True)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
