"""MFMA lane layouts the kernels assume.

v_mfma_f64_16x16x4f64 (SP2 purification, myqc_amd/csrc/sp2.hip): lane l feeds
A[l%16][l/16] and B[l/16][l%16] and receives D rows l/16 + 4v, column l%16.

v_mfma_f32_16x16x32_bf16 (split-bf16 GEMM, myqc_amd/csrc/mfma_gemm.hip): lane
l feeds A[l%16][8(l/16)+e] and B[8(l/16)+e][l%16], e = 0..7, and receives D
rows 4(l/16) + v, column l%16."""
import ctypes
import os

import numpy as np
import pytest

from conftest import ROOT

pytestmark = pytest.mark.gpu
LIB = os.path.join(ROOT, "tests", "gpu_probe", "libmfmaprobe.so")


def test_f64_mfma_output_layout():
    L = ctypes.CDLL(LIB)
    rng = np.random.default_rng(0)
    A = rng.standard_normal((16, 4))
    B = rng.standard_normal((4, 16))
    raw = np.zeros(256)
    dp = ctypes.POINTER(ctypes.c_double)
    assert L.mfma_probe(np.ascontiguousarray(A).ctypes.data_as(dp), np.ascontiguousarray(B).ctypes.data_as(dp),
                        raw.ctypes.data_as(dp)) == 0
    raw = raw.reshape(64, 4)
    Dref = A @ B
    D = np.zeros((16, 16))
    for l in range(64):
        for v in range(4):
            D[l // 16 + 4 * v, l % 16] = raw[l, v]
    assert np.max(np.abs(D - Dref)) < 1e-12


def test_bf16_mfma_layout():
    L = ctypes.CDLL(LIB)
    rng = np.random.default_rng(1)
    # small integers: exact in bf16, and every product sum exact in FP32
    A = rng.integers(-8, 9, (16, 32)).astype(np.float32)
    B = rng.integers(-8, 9, (32, 16)).astype(np.float32)
    raw = np.zeros(256, dtype=np.float32)
    fp = ctypes.POINTER(ctypes.c_float)
    assert L.mfma_probe_bf16(np.ascontiguousarray(A).ctypes.data_as(fp), np.ascontiguousarray(B).ctypes.data_as(fp),
                             raw.ctypes.data_as(fp)) == 0
    raw = raw.reshape(64, 4)
    D = np.zeros((16, 16), dtype=np.float64)
    for l in range(64):
        for v in range(4):
            D[4 * (l // 16) + v, l % 16] = raw[l, v]
    assert np.array_equal(D, A.astype(np.float64) @ B.astype(np.float64))
