"""The C ABI boundary, on the CPU: the library loads, exports every function
include/qmetal/qmetal.h declares, and its host-side validation and error
convention match the reference (src/capi.cpp): -1 plus a thread-local
message, NULL from constructors, 0 from getters on NULL handles."""
import ctypes
import os
import re

import pytest

import qmetal_amd as qm
from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "qmetal", "qmetal.h")

REFERENCE_27 = """qm_version qm_last_error qm_is_available qm_device_name qm_max_buffer_length
qm_recommended_working_set qm_max_qubits qm_circuit_create qm_circuit_destroy qm_circuit_add
qm_circuit_size qm_circuit_plan_size qm_create qm_destroy qm_reset qm_run qm_run_unfused
qm_synchronize qm_num_qubits qm_dim qm_dispatch_count qm_amplitudes qm_norm
qm_expectation_pauli qm_expectation qm_sample qm_energy_and_gradient""".split()


def declared():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"\b(qm_[a-z0-9_]+)\s*\(", text)))


def test_exports_every_declared_symbol():
    lib = ctypes.CDLL(qm.library_path())
    names = declared()
    assert set(REFERENCE_27) <= set(names)
    for name in names:
        assert hasattr(lib, name), name


def test_gate_codes_and_version():
    text = open(HEADER).read()
    for name, code in (("QM_X", 0), ("QM_H", 3), ("QM_RX", 8), ("QM_P", 11), ("QM_CX", 12),
                       ("QM_CP", 16), ("QM_CCX", 17), ("QM_U3", 18)):
        assert re.search(rf"\b{name} = {code}\b", text), name
    assert qm.version() == (0, 1, 0)
    assert qm.max_qubits() == 34


def test_header_compiles_as_c11(tmp_path):
    src = tmp_path / "t.c"
    src.write_text('#include "qmetal/qmetal.h"\nint main(void){return 0;}\n')
    rc = os.system(f"gcc -std=c11 -Wall -Wextra -Werror -I{ROOT}/include -c {src} -o {tmp_path}/t.o")
    assert rc == 0


def test_circuit_validation():
    c = qm.Circuit(2)
    with pytest.raises(qm.QMetalError, match="out of range"):
        c.x(5)
    with pytest.raises(qm.QMetalError, match="arity"):
        c.add(0, ())
    with pytest.raises(qm.QMetalError, match="arity"):
        c.add(17, (0, 1, 0, 1))
    with pytest.raises(qm.QMetalError, match="three parameters"):
        c.add(8, (0,), (1, 2, 3, 4))
    with pytest.raises(qm.QMetalError, match="unknown gate"):
        c.add(99, (0,))
    assert len(c) == 0
    c.h(0).cx(0, 1)
    assert len(c) == 2


def test_plan_size_reports_unsupported_gate():
    c = qm.Circuit(3)
    c.add(12, (0,))  # CX with one wire: accepted by add, rejected by lowering
    assert c.plan_size(True) == 0
    assert "one-qubit" in qm._err() or "unsupported" in qm._err()


def test_null_handles():
    lib = ctypes.CDLL(qm.library_path())
    lib.qm_circuit_size.restype = ctypes.c_size_t
    lib.qm_circuit_size.argtypes = [ctypes.c_void_p]
    assert lib.qm_circuit_size(None) == 0
    lib.qm_dim.restype = ctypes.c_uint64
    lib.qm_dim.argtypes = [ctypes.c_void_p]
    assert lib.qm_dim(None) == 0
    lib.qm_run.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
    assert lib.qm_run(None, None) == -1
    assert qm._err() == "null argument"
    lib.qm_destroy.argtypes = [ctypes.c_void_p]
    lib.qm_destroy(None)
    lib.qm_circuit_destroy.argtypes = [ctypes.c_void_p]
    lib.qm_circuit_destroy(None)
    lib.qm_batch_spread.restype = ctypes.c_uint32
    lib.qm_batch_spread.argtypes = [ctypes.c_void_p]
    assert lib.qm_batch_spread(None) == 0
    lib.qm_batch_set_spread.argtypes = [ctypes.c_void_p, ctypes.c_uint32]
    assert lib.qm_batch_set_spread(None, 14) == -1
    assert qm._err() == "null batch"


def test_oversized_simulator_raises():
    with pytest.raises(qm.QMetalError, match="num_qubits"):
        qm.Simulator(64)
    with pytest.raises(qm.QMetalError, match="num_qubits"):
        qm.Simulator(0)


def test_no_device_fails_closed():
    if qm.is_available():
        pytest.skip("a HIP device is present")
    assert qm.device_name() == "none"
    with pytest.raises(qm.QMetalError, match="no HIP device"):
        qm.Simulator(4)
