"""Lane layout of the 16x16x4 MFMA builtins the assembly kernels rely on
(tests/host/mfma_probe.hip).  Lane l holds column l % 16 of D; its rows are
4*(l//16) + r for FP32 and l//16 + 4*r for FP64 (r = 0..3), which is how
gtk_kernels.hip (mfma_row) reads the results."""
import ctypes as C
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
LIB = os.path.join(os.path.dirname(os.path.abspath(__file__)), "host", "libmfma_probe.so")


@pytest.mark.parametrize("f64", [1, 0])
@pytest.mark.parametrize("chain", [1, 3])
def test_mfma_16x16x4_layout(f64, chain):
    print(f"case mfma layout f64={f64} chain={chain}", flush=True)
    L = C.CDLL(LIB)
    L.mfma_probe.argtypes = [C.c_int, C.c_int, np.ctypeslib.ndpointer(np.float64, flags="C")]
    out = np.zeros(256)
    assert L.mfma_probe(f64, chain, out) == 0
    got = out.reshape(64, 4)
    lane = np.arange(64)[:, None]
    r = np.arange(4)[None, :]
    i = (lane // 16 + 4 * r) if f64 else (4 * (lane // 16) + r)
    j = lane % 16
    expect = chain * (1000.0 * i + j)
    if not np.array_equal(got, expect):
        v = got / chain
        for l in range(64):
            print(l, [(int(x // 1000), int(x % 1000)) for x in v[l]])
    assert np.array_equal(got, expect)


@pytest.mark.parametrize("chain", [1, 3])
def test_mfma_16x16x32_f16_layout(chain):
    """FP16 inputs, FP32 accumulation: lane l supplies A[l%16][8(l//16)+i]
    and B[8(l//16)+i][l%16]; it holds D rows 4(l//16)+r of column l%16, as
    the FP32 16x16x4 product."""
    print(f"case mfma layout f16x32 chain={chain}", flush=True)
    L = C.CDLL(LIB)
    L.mfma_probe_f16x32.argtypes = [C.c_int, np.ctypeslib.ndpointer(np.float64, flags="C")]
    out = np.zeros(256)
    assert L.mfma_probe_f16x32(chain, out) == 0
    got = out.reshape(64, 4)
    lane = np.arange(64)[:, None]
    r = np.arange(4)[None, :]
    expect = chain * (1000.0 * (4 * (lane // 16) + r) + lane % 16)
    if not np.array_equal(got, expect):
        v = got / chain
        for l in range(64):
            print(l, [(int(x // 1000), int(x % 1000)) for x in v[l]])
    assert np.array_equal(got, expect)
