"""qm_sample_noalloc (scratch in a fixed per-GPU device array) against
qm_sample through the C ABI: the drawn indices are bit-identical on two
calls in a row, the second on a smaller state with fewer shots, so stale
scratch from the first call cannot leak into the second. Sizes past the
fixed scratch fail with a message naming the limit."""
import ctypes

import numpy as np
import pytest

import qmetal_amd as qm
from qmetal_amd import circuits as C

pytestmark = pytest.mark.gpu


def _abi():
    lib = ctypes.CDLL(qm.library_path())
    f = lib.qm_sample_noalloc
    f.restype = ctypes.c_int
    f.argtypes = [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_uint64,
                  ctypes.POINTER(ctypes.c_uint64)]
    g = lib.qm_sample
    g.restype = ctypes.c_int
    g.argtypes = f.argtypes
    return f, g


def _draw(fn, sim, shots, seed):
    out = np.zeros(max(shots, 1), dtype=np.uint64)
    rc = fn(sim._h, shots, seed, out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)))
    assert rc == 0, qm._err()
    return out[:shots]


def test_noalloc_sample_equals_sample_on_two_calls():
    noalloc, sample = _abi()
    cases = ((20, C.product_state_prep(20, seed=5) + C.qft(20), 200000, 99),
             (12, C.brickwork(12, 3, seed=4), 3000, 7))
    for n, gates, shots, seed in cases:
        with qm.Simulator(n) as s:
            s.run(C.to_circuit(n, gates))
            want = _draw(sample, s, shots, seed)
            got = _draw(noalloc, s, shots, seed)
            again = _draw(noalloc, s, shots, seed)
        print(f"n={n}: {shots} shots, {len(np.unique(got))} distinct indices, "
              f"equal to qm_sample: {np.array_equal(got, want)}")
        assert np.array_equal(got, want) and np.array_equal(again, want)


def test_noalloc_sample_limits():
    noalloc, _ = _abi()
    with qm.Simulator(4) as s:
        out = np.zeros(1, dtype=np.uint64)
        rc = noalloc(s._h, (1 << 20) + 1, 1, out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)))
        assert rc == -1 and "limit" in qm._err()
        assert noalloc(s._h, 0, 1, None) == 0
    with qm.Simulator(31) as s:
        out = np.zeros(4, dtype=np.uint64)
        rc = noalloc(s._h, 4, 1, out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)))
        assert rc == -1 and "30" in qm._err()
