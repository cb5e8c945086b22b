"""State placement probe (qm_state_placement): for 33 qubits the
constructor times a value-preserving pass over the top seven wires against
one over wires 18..12 and allocates again while the first is slower, keeping
the fastest allocation. The state must still start as |0...0> and run
circuits exactly as an unprobed one."""
import os

import numpy as np
import pytest

import qmetal_amd as qm
from qmetal_amd import circuits as C

pytestmark = pytest.mark.gpu


def test_probed_state_starts_at_zero_and_runs():
    n = 33
    with qm.Simulator(n) as s:
        attempts, ratio = s.placement
        print(f"n={n}: {attempts} allocation(s), top/mid pass time ratio {ratio:.3f}")
        assert 1 <= attempts <= 4 and ratio > 0.0
        first = s.amplitudes_at([0, 1, (1 << n) - 1])
        assert first[0] == 1.0 and first[1] == 0.0 and first[2] == 0.0
        assert abs(s.norm() - 1.0) <= 1e-6
        s.run(C.to_circuit(n, C.ghz(n)))
        assert abs(s.expectation_pauli(z=1 | (1 << (n - 1))) - 1.0) <= 1e-5
        assert abs(s.norm() - 1.0) <= 1e-5


def test_probe_off_and_outside_the_range():
    old = os.environ.get("QMETAL_PLACEMENT_PROBE")
    os.environ["QMETAL_PLACEMENT_PROBE"] = "0"
    try:
        with qm.Simulator(33) as s:
            assert s.placement == (0, 0.0)
            assert s.amplitudes_at([0])[0] == 1.0
    finally:
        if old is None:
            del os.environ["QMETAL_PLACEMENT_PROBE"]
        else:
            os.environ["QMETAL_PLACEMENT_PROBE"] = old
    for n in (20, 32):
        with qm.Simulator(n) as s:
            assert s.placement == (0, 0.0)
