"""Pins the complex128 C oracle before anything is judged against it.

Closed forms follow the reference's oracle suite (tests/test_reference.cpp);
amplitudes and expectations are compared with the reference oracle itself,
built from its sources (golden fixtures); sampling is checked against a
brute-force inverse CDF; the adjoint against exact parameter shift.
"""
import math

import numpy as np
import pytest

from conftest import golden_amps, golden_gates, hexf, load_golden
from qmetal_amd import circuits as C
from oracle import oracle as O

TOL = 1e-12


def test_bit_convention():
    for q in range(3):
        s = O.run(3, [(C.X, (q,), ())])
        assert abs(abs(s[1 << q]) - 1) <= TOL


def test_single_qubit_identities():
    assert abs(abs(O.run(1, [(C.H, (0,), ())] * 2)[0]) - 1) <= TOL
    assert abs(abs(O.run(1, [(C.X, (0,), ())] * 2)[0]) - 1) <= TOL
    a = O.run(1, [(C.X, (0,), ()), (C.S, (0,), ()), (C.S, (0,), ())])
    b = O.run(1, [(C.X, (0,), ()), (C.Z, (0,), ())])
    assert np.abs(a - b).max() <= TOL
    assert abs(O.run(1, [(C.X, (0,), ())] + [(C.T, (0,), ())] * 8)[1].real - 1) <= TOL
    assert abs(O.run(1, [(C.RZ, (0,), (2 * math.pi,))])[0].real + 1) <= TOL
    assert abs(abs(O.run(1, [(C.RY, (0,), (math.pi,))])[1]) - 1) <= TOL
    assert abs(O.run(1, [(C.RX, (0,), (math.pi,))])[1].imag + 1) <= TOL
    s = O.run(1, [(C.X, (0,), ()), (C.S, (0,), ()), (C.SDG, (0,), ()), (C.T, (0,), ()),
                  (C.TDG, (0,), ())])
    assert abs(s[1].real - 1) <= TOL


def test_two_qubit_gates():
    r = 1 / math.sqrt(2)
    s = O.run(2, [(C.H, (0,), ()), (C.CX, (0, 1), ())])
    assert abs(s[0] - r) <= TOL and abs(s[3] - r) <= TOL and abs(s[1]) <= TOL and abs(s[2]) <= TOL
    # CX^2 = I, CX(0,1) != CX(1,0) on |01>, CZ and CP(pi) agree, SWAP moves |01> to |10>
    s = O.run(2, [(C.X, (0,), ()), (C.CX, (0, 1), ()), (C.CX, (0, 1), ())])
    assert abs(s[1] - 1) <= TOL
    assert abs(O.run(2, [(C.X, (0,), ()), (C.CX, (0, 1), ())])[3] - 1) <= TOL
    assert abs(O.run(2, [(C.X, (0,), ()), (C.CX, (1, 0), ())])[1] - 1) <= TOL
    prep = [(C.H, (0,), ()), (C.H, (1,), ())]
    a = O.run(2, prep + [(C.CZ, (0, 1), ())])
    b = O.run(2, prep + [(C.CP, (0, 1), (math.pi,))])
    assert np.abs(a - b).max() <= TOL and abs(a[3] + 0.5) <= TOL
    assert abs(O.run(2, [(C.X, (0,), ()), (C.SWAP, (0, 1), ())])[2] - 1) <= TOL
    s = O.run(3, [(C.X, (0,), ()), (C.X, (1,), ()), (C.CCX, (0, 1, 2), ())])
    assert abs(s[7] - 1) <= TOL
    s = O.run(3, [(C.X, (0,), ()), (C.CCX, (0, 1, 2), ())])
    assert abs(s[1] - 1) <= TOL


def test_ghz_and_qft_closed_forms():
    r = 1 / math.sqrt(2)
    for n in (2, 3, 8, 12):
        s = O.run(n, C.ghz(n))
        assert abs(s[0] - r) <= TOL and abs(s[-1] - r) <= TOL
        assert np.sum(np.abs(s[1:-1]) ** 2) <= TOL
    for n in (1, 2, 5, 10):
        s = O.run(n, C.qft(n))
        assert np.abs(s - 1 / math.sqrt(1 << n)).max() <= TOL
    s = O.run(2, [(C.X, (0,), ())] + C.qft(2))
    assert np.abs(s - np.array([0.5, -0.5, 0.5j, -0.5j])).max() <= TOL
    n, N, x = 3, 8, 3
    s = O.run(n, [(C.X, (0,), ()), (C.X, (1,), ())] + C.qft(n))

    def rev(k):
        return int(format(k, f"0{n}b")[::-1], 2)

    want = np.array([np.exp(2j * math.pi * x * k / N) / math.sqrt(N) for k in range(N)])
    got = np.array([s[rev(k)] for k in range(N)])
    assert np.abs(got - want).max() <= TOL


def test_pauli_closed_forms():
    plus_i = O.run(1, [(C.H, (0,), ()), (C.S, (0,), ())])
    assert abs(O.expect(plus_i, y=1) - 1) <= TOL
    assert abs(O.expect(plus_i, x=1)) <= TOL and abs(O.expect(plus_i, z=1)) <= TOL
    minus_i = O.run(1, [(C.H, (0,), ()), (C.SDG, (0,), ())])
    assert abs(O.expect(minus_i, y=1) + 1) <= TOL
    bell = O.run(2, [(C.H, (0,), ()), (C.CX, (0, 1), ())])
    assert abs(O.expect(bell, y=3) + 1) <= TOL
    assert abs(O.expect(bell, x=3) - 1) <= TOL
    assert abs(O.expect(bell, z=3) - 1) <= TOL


def dense_pauli(n, x, y, z):
    mats = {"I": np.eye(2), "X": np.array([[0, 1], [1, 0]]),
            "Y": np.array([[0, -1j], [1j, 0]]), "Z": np.diag([1, -1])}
    out = np.array([[1.0]])
    for q in range(n):  # qubit 0 is the least significant factor
        ch = "X" if x >> q & 1 else "Y" if y >> q & 1 else "Z" if z >> q & 1 else "I"
        out = np.kron(mats[ch], out)
    return out


def test_expectation_against_dense_matrix():
    n = 5
    s = O.run(n, C.random_circuit(n, 4, seed=77))
    rng = C.SplitMix64(9)
    for _ in range(40):
        x = y = z = 0
        for q in range(n):
            r = rng.next_u64() % 4
            x |= (r == 1) << q
            y |= (r == 2) << q
            z |= (r == 3) << q
        want = np.vdot(s, dense_pauli(n, x, y, z) @ s).real
        assert abs(O.expect(s, x, y, z) - want) <= 1e-12


def test_gate_matrices_unitary_and_u3():
    for kind in range(12):
        m = O.gate_matrix(kind, (0.731,))
        assert np.abs(m.conj().T @ m - np.eye(2)).max() <= 1e-15
    # U3 special cases reduce to textbook gates: U3(pi/2, 0, pi) = H,
    # U3(t, -pi/2, pi/2) = RX(t), U3(t, 0, 0) = RY(t), U3(0, 0, l) = P(l).
    assert np.abs(O.gate_matrix(C.U3, (math.pi / 2, 0, math.pi)) - O.gate_matrix(C.H)).max() < 1e-15
    t = 0.37
    assert np.abs(O.gate_matrix(C.U3, (t, -math.pi / 2, math.pi / 2)) - O.gate_matrix(C.RX, (t,))).max() < 1e-15
    assert np.abs(O.gate_matrix(C.U3, (t, 0, 0)) - O.gate_matrix(C.RY, (t,))).max() < 1e-15
    assert np.abs(O.gate_matrix(C.U3, (0, 0, t)) - O.gate_matrix(C.P, (t,))).max() < 1e-15


def test_unitarity_and_fail_closed():
    for gates, n in ((C.qaoa_ring(7, 6), 7), (C.tfim_trotter(7, 20), 7), (C.grover_proxy(7, 4), 7),
                     (C.hardware_efficient(12, 3), 12), (C.brickwork(12, 10), 12)):
        assert abs(O.norm(O.run(n, gates)) - 1) <= 1e-11
    with pytest.raises(ValueError):
        O.run(2, [(C.CX, (0, 0), ())])
    with pytest.raises(ValueError):
        O.run(2, [(C.X, (5,), ())])


@pytest.mark.parametrize("name", ["frozen_small.json.gz", "fuzz.json.gz"])
def test_oracle_matches_reference_oracle(name):
    """Amplitudes, norm and expectations equal the reference oracle's."""
    d = load_golden(name)
    items = d.items() if isinstance(d, dict) else enumerate(d)
    for key, case in items:
        n = case["n"]
        got = O.run(n, golden_gates(case))
        assert np.abs(got - golden_amps(case)).max() <= 1e-14, key
        assert abs(O.norm(got) - hexf(case["norm"])) <= 1e-14, key
        for q, e in enumerate(case["expz"]):
            assert abs(O.expect(got, z=1 << q) - hexf(e)) <= 1e-13, key
        x = sum(1 << q for q in range(n) if q % 3 == 0)
        y = sum(1 << q for q in range(n) if q % 3 == 1)
        z = sum(1 << q for q in range(n) if q % 3 == 2)
        assert abs(O.expect(got, x, y, z) - hexf(case["exp_mixed"])) <= 1e-13, key


def test_splitmix64_published_vector():
    # Reference outputs of splitmix64 seeded with 1234567 (Vigna's splitmix64.c).
    assert O.splitmix64_stream(1234567, 5) == [
        6457827717110365317, 3203168211198807973, 9817491932198370423,
        4593380528125082431, 16408922859458223821]
    # The Python generator used by the workloads is the same stream.
    rng = C.SplitMix64(42)
    assert [rng.next_u64() for _ in range(8)] == O.splitmix64_stream(42, 8)


def test_sampler_against_brute_force_inverse_cdf():
    """Each shot is the index the splitmix64 uniform selects on the exact
    cumulative distribution (two-level blocks change nothing but ties)."""
    for n, gates in ((6, C.random_circuit(6, 3, seed=5)), (11, C.qaoa_ring(11, 2)),
                     (12, C.brickwork(12, 3))):
        s = O.run(n, gates)
        pr = np.abs(s) ** 2
        cum = np.cumsum(pr)
        shots = 4000
        got = O.sample(s, shots, 2024)
        us = [(r >> 11) * 2.0 ** -53 * cum[-1] for r in O.splitmix64_stream(2024, shots)]
        want = np.searchsorted(cum, us, side="right")
        mism = np.nonzero(got != want)[0]
        for k in mism:  # only at a cumulative boundary, within round-off
            i = int(got[k])
            assert min(abs(us[k] - cum[i]), abs(us[k] - (cum[i - 1] if i else 0))) <= 1e-12
        assert len(mism) <= 2
    ghz = O.run(6, C.ghz(6))
    out = O.sample(ghz, 20000, 12345)
    assert set(out.tolist()) <= {0, 63}
    assert abs(np.mean(out == 0) - 0.5) < 0.02


def test_adjoint_matches_parameter_shift():
    n = 6
    h = [(1.0, 0, 0, 3 << q) for q in range(n - 1)] + [(0.4, 1 << q, 0, 0) for q in range(n)]
    h.append((0.25, 0, 1, 0))
    for gates in (C.hardware_efficient(n, 2, seed=0xC0FFEE), C.tfim_trotter(n, 3), C.qaoa_ring(n, 2)):
        e, g = O.adjoint(n, gates, h)
        assert abs(e - O.expect_h(O.run(n, gates), h)) <= 1e-12
        assert np.abs(g - O.param_shift(n, gates, h)).max() <= 1e-12
    for t in (0.0, 0.3, 1.1, 2.7, -0.9):
        e, g = O.adjoint(2, [(C.RY, (0,), (t,))], [(1.0, 0, 0, 1)])
        assert abs(e - math.cos(t)) <= 1e-14 and abs(g[0] + math.sin(t)) <= 1e-14
    with pytest.raises(ValueError):
        O.adjoint(2, [(C.H, (0,), ()), (C.CP, (0, 1), (0.4,))], [(1.0, 0, 0, 1)])


def test_cpu_baseline_legs_run_the_same_workload():
    """The bench's CPU baseline: the reference's own oracle (timed inside
    oracle/_ref/ref_dump) and the restatement run the same preparation and
    QFT; their first and last amplitudes agree to 1e-12."""
    import os
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    sys.path.insert(0, root)
    import bench
    if not os.access(bench.REF_DUMP, os.X_OK):
        pytest.skip("oracle/_ref not built (the reference tree is absent)")
    n = 10
    dt, scaled, rec = bench.ref_cpu_baseline(n, 12)
    assert rec["gates"] == n * (n + 1) // 2 and dt >= 0
    assert scaled == pytest.approx(dt * 4 * (12 * 13) / (n * (n + 1)))
    st = O.run(n, C.qft(n), state=O.run(n, C.product_state_prep(n, seed=2024)))
    assert abs(complex(*rec["first"]) - st[0]) < 1e-12
    assert abs(complex(*rec["last"]) - st[-1]) < 1e-12
    assert abs(rec["norm"] - 1) < 1e-12


def test_qft_product_state_closed_form_matches_oracle():
    """The full-size checks' closed forms, pinned on the complex128 oracle:
    QFT (no swap layer) of a seeded product state at every output index, the
    product state's Bloch vectors, and the inverse stream undoing QFT and
    brickwork circuits."""
    import closed_forms as CF
    for n in (1, 2, 5, 8):
        prep = C.product_state_prep(n, seed=300 + n)
        a, b = CF.product_factors(prep)
        st = O.run(n, prep)
        xs, ys, zs = CF.bloch(a, b)
        for q in range(n):
            assert abs(O.expect(st, x=1 << q) - xs[q]) < 1e-12
            assert abs(O.expect(st, y=1 << q) - ys[q]) < 1e-12
            assert abs(O.expect(st, z=1 << q) - zs[q]) < 1e-12
        out = O.run(n, C.qft(n), state=st)
        want = CF.qft_product_amplitudes(a, b, np.arange(1 << n))
        assert np.abs(out - want).max() < 1e-12, n
        back = O.run(n, CF.inverse_stream(C.qft(n)), state=out)
        assert np.abs(back - st).max() < 1e-12
    # A dropped bit reversal or a flipped phase sign fails the form.
    n = 5
    prep = C.product_state_prep(n, seed=305)
    a, b = CF.product_factors(prep)
    out = O.run(n, prep + C.qft(n))
    idx = np.arange(1 << n)
    assert np.abs(out - CF.qft_product_amplitudes(a, b, CF.bit_reverse(idx, n))).max() > 1e-3
    assert np.abs(out - CF.qft_product_amplitudes(np.conj(a), np.conj(b), idx)).max() > 1e-3
    g = C.brickwork(6, 4, seed=9)
    assert np.abs(O.run(6, CF.inverse_stream(g), state=O.run(6, g)) - O.zero_state(6)).max() < 1e-12


def test_adjoint_follows_num_params():
    """simulator.mm:822-828, :917: a gate carrying a parameter must be a
    rotation; a rotation added without one gets gradient 0, and a P added
    without one is accepted."""
    h = [(1.0, 0, 0, 1), (0.5, 2, 0, 0)]
    gates = [(C.H, (0,), ()), (C.RX, (0,), ()), (C.RY, (1,), (0.7,)), (C.P, (1,), ()),
             (C.CX, (1, 0), ()), (C.RZ, (0,), (0.3,))]
    e, g = O.adjoint(2, gates, h)
    assert g[1] == 0.0 and g[3] == 0.0
    assert abs(g[2]) > 1e-3
    want = O.param_shift(2, gates, h)
    assert np.abs(g - want).max() < 1e-12
    with pytest.raises(ValueError):
        O.adjoint(2, [(C.P, (1,), (0.2,))], h)
