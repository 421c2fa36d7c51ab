"""The v_mfma_f64_16x16x4f64 lane layout the SP2 purification kernel
(myqc_amd/csrc/sp2.hip) assumes: lane l feeds A[l%16][l/16] and B[l/16][l%16]
and receives D rows l/16 + 4v, column l%16."""
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
