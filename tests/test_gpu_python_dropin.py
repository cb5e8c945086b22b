"""The reference's end-to-end Python check (tests/test_python.py) run
through `import qmetal`, the name reference callers use: handle lifetimes,
array marshalling and error propagation against closed forms. One documented
divergence: max_qubits() is 34 on MI355X (the reference reports 33, its M4
ceiling)."""
import math

import pytest

pytestmark = pytest.mark.gpu


def test_reference_python_checks_through_import_qmetal():
    import qmetal
    import qmetal_amd
    assert qmetal.Simulator is qmetal_amd.Simulator
    print(f"qmetal {'.'.join(map(str, qmetal.version()))} on {qmetal.device_name()}")
    assert qmetal.is_available()
    assert qmetal.max_qubits() == 34

    n = 6
    c = qmetal.Circuit(n)
    c.h(0)
    for i in range(n - 1):
        c.cx(i, i + 1)
    assert len(c) == n
    sim = qmetal.Simulator(n)
    sim.run(c)
    assert abs(sim.norm() - 1.0) <= 1e-5
    assert abs(sim.expectation("Z" + "I" * (n - 1))) <= 1e-5
    assert abs(sim.expectation("ZZ" + "I" * (n - 2)) - 1.0) <= 1e-5
    counts = sim.counts(4000, seed=7)
    assert set(counts) == {"0" * n, "1" * n}
    assert abs(counts["0" * n] / 4000 - 0.5) < 0.03
    amps = sim.amplitudes()
    assert abs(abs(amps[0]) - 1 / math.sqrt(2)) <= 1e-5
    assert abs(abs(amps[(1 << n) - 1]) - 1 / math.sqrt(2)) <= 1e-5

    qc = qmetal.Circuit(16)
    for j in range(15, -1, -1):
        qc.h(j)
        for k in range(j - 1, -1, -1):
            qc.cp(k, j, math.pi / (1 << (j - k)))
    assert qc.plan_size(fuse=False) == len(qc)
    assert len(qc) / qc.plan_size(fuse=True) >= 3.0

    for t in (0.0, 0.4, 1.3, -0.7):
        g = qmetal.Circuit(2)
        g.ry(0, t)
        h = qmetal.Hamiltonian().add_string(1.0, "ZI")
        e, grad = qmetal.Simulator(2).energy_and_gradient(g, h)
        assert abs(e - math.cos(t)) <= 1e-5
        assert abs(grad[0] + math.sin(t)) <= 1e-5

    with pytest.raises(qmetal.QMetalError):
        qmetal.Circuit(2).x(5)
    with pytest.raises(qmetal.QMetalError):
        qmetal.Simulator(64)
    bad = qmetal.Circuit(2)
    bad.cp(0, 1, 0.3)
    with pytest.raises(qmetal.QMetalError):
        qmetal.Simulator(2).energy_and_gradient(bad, qmetal.Hamiltonian().add_string(1.0, "ZI"))
