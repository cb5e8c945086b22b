"""HIP path against the complex128 oracle (the reference's GPU suite,
tests/test_gpu.mm, restated): every gate on every wire, every 2q gate on every
ordered pair, CCX on every triple, the frozen set, drift with depth, launch
grouping invariance, fail-closed behaviour and dispatch accounting.

Tolerances: 1e-5 max amplitude difference for single gates (complex64 floor,
test_gpu.mm:50), 1e-4 for deep circuits and the norm (test_gpu.mm:155-156);
with depth, a bound that grows with the gate count G: amplitudes within
1e-6 + 1e-8 G and the norm within 1e-6 + 2e-8 G (TFIM n=12 measured
2-4e-9 G and 6e-9 G up to G = 4500).
"""
import math

import numpy as np
import pytest

import qmetal_amd as qm
from conftest import FUSION, golden_amps, golden_gates, load_golden
from qmetal_amd import FusionOptions
from qmetal_amd import circuits as C
from oracle import oracle as O

pytestmark = pytest.mark.gpu

TOL1 = 1e-5
TOL_DEEP = 1e-4


def tol_amp(gates: int) -> float:
    """Max amplitude difference allowed after `gates` complex64 gates."""
    return 1e-6 + 1e-8 * gates


def tol_norm(gates: int) -> float:
    """Norm drift allowed after `gates` complex64 gates."""
    return 1e-6 + 2e-8 * gates


def gpu_state(n, gates, options=None, fuse=True):
    c = C.to_circuit(n, gates)
    with qm.Simulator(n) as s:
        s.run(c, fuse=fuse, options=options)
        return s.amplitudes().astype(np.complex128)


def diff(n, gates, **kw):
    return np.abs(gpu_state(n, gates, **kw) - O.run(n, gates)).max()


def test_device_present():
    assert qm.is_available(), "no HIP device: GPU tests need an MI355X"
    print("device:", qm.device_name())


def test_single_qubit_gates_every_wire():
    n = 5
    spread = [(C.H, (i,), ()) for i in range(n)]
    for k in (C.X, C.Y, C.Z, C.H, C.S, C.SDG, C.T, C.TDG):
        for q in range(n):
            assert diff(n, spread + [(k, (q,), ())]) <= TOL1, (k, q)
    for k in (C.RX, C.RY, C.RZ, C.P):
        for q in range(n):
            assert diff(n, spread + [(k, (q,), (0.7853981633974483 + 0.3 * q,))]) <= TOL1, (k, q)
    for q in range(n):
        assert diff(n, spread + [(C.U3, (q,), (0.4 + q, -1.1, 2.3 * q))]) <= TOL1, q


def test_two_qubit_gates_every_pair():
    n = 4
    prep = [(C.H, (i,), ()) for i in range(n)] + [(C.RZ, (0,), (0.4,)), (C.RY, (2,), (0.9,))]
    for k in (C.CX, C.CY, C.CZ, C.SWAP):
        for a in range(n):
            for b in range(n):
                if a != b:
                    assert diff(n, prep + [(k, (a, b), ())]) <= TOL1, (k, a, b)
    for th in (0.1, math.pi / 4, math.pi / 2, math.pi, -math.pi / 3):
        assert diff(n, [(C.H, (i,), ()) for i in range(n)] + [(C.CP, (1, 3), (th,))]) <= TOL1


def test_ccx_every_triple():
    n = 4
    prep = [(C.H, (i,), ()) for i in range(n)] + [(C.RZ, (0,), (0.55,))]
    for a in range(n):
        for b in range(n):
            for t in range(n):
                if len({a, b, t}) == 3:
                    assert diff(n, prep + [(C.CCX, (a, b, t), ())]) <= TOL1, (a, b, t)


def test_wide_wires_use_fixed_bit_paths():
    """Controls and targets above bit 4 take the half-state (fixed-bit)
    paths; every placement is checked at n = 14."""
    n = 14
    prep = C.brickwork(n, 2, seed=3)
    for c_, t in ((0, 13), (13, 0), (5, 9), (9, 5), (4, 1), (12, 11), (3, 4)):
        assert diff(n, prep + [(C.CX, (c_, t), ())]) <= TOL1, (c_, t)
        assert diff(n, prep + [(C.CP, (c_, t), (0.77,))]) <= TOL1, (c_, t)
        assert diff(n, prep + [(C.SWAP, (c_, t), ())]) <= TOL1, (c_, t)
    for a, b, t in ((4, 9, 0), (0, 1, 13), (11, 12, 13), (2, 7, 5)):
        assert diff(n, prep + [(C.CCX, (a, b, t), ())]) <= TOL1, (a, b, t)
    for q in (0, 1, 5, 6, 12, 13):
        assert diff(n, prep + [(C.H, (q,), ())]) <= TOL1, q
    # groups of 2..4 wires in every position class
    for wires in ((0, 1), (0, 13), (5, 6, 7), (1, 2, 3, 4), (0, 6, 12, 13), (9, 10, 11, 12)):
        g = [(C.RY, (w,), (0.3 + w,)) for w in wires] + [(C.RX, (w,), (0.2 * w,)) for w in wires]
        assert diff(n, prep + g) <= TOL1, wires


FROZEN = {
    "ghz": lambda n: C.ghz(n), "qft": lambda n: C.qft(n),
    "qaoa_ring_p6": lambda n: C.qaoa_ring(n, 6), "tfim_trotter_20": lambda n: C.tfim_trotter(n, 20),
    "grover_proxy_4": lambda n: C.grover_proxy(n, 4), "random_d10": lambda n: C.random_circuit(n, 10),
    "hea_d3": lambda n: C.hardware_efficient(n, 3), "brickwork_d10": lambda n: C.brickwork(n, 10),
}


@pytest.mark.parametrize("name", sorted(FROZEN))
def test_frozen_set_against_oracle(name):
    for n in (6, 10, 14, 20):
        gates = FROZEN[name](n)
        want = O.run(n, gates)
        for opt in ("fused", "none", "blocked8") if n >= 8 else ("fused", "none"):
            c = C.to_circuit(n, gates)
            with qm.Simulator(n) as s:
                s.run(c, options=FusionOptions(*FUSION[opt]))
                got = s.amplitudes().astype(np.complex128)
                nerr = abs(s.norm() - 1.0)
            d = np.abs(got - want).max()
            print(f"{name:16s} n={n:2d} {opt:9s} gates={len(gates):5d} diff={d:.3e} norm_err={nerr:.3e}")
            assert d <= TOL_DEEP and nerr <= TOL_DEEP, (name, n, opt)


def test_reference_streams_against_reference_amplitudes():
    """The reference generators' own streams (golden fixtures) at n = 4, 7,
    11, under every FusionOptions, against the reference oracle's amplitudes."""
    for key, d in load_golden("frozen_small.json.gz").items():
        gates, want = golden_gates(d), golden_amps(d)
        # A block wider than the state is rejected, as by the reference
        # executor (src/metal/simulator.mm:468).
        for opt in [o for o in FUSION if FUSION[o][4] <= d["n"]]:
            got = gpu_state(d["n"], gates, options=FusionOptions(*FUSION[opt]))
            assert np.abs(got - want).max() <= TOL_DEEP, (key, opt)


def test_fuzz_streams_against_reference_amplitudes():
    for k, d in enumerate(load_golden("fuzz.json.gz")):
        gates, want = golden_gates(d), golden_amps(d)
        for opt in ("fused", "none", "diag_only", "win_only"):
            got = gpu_state(d["n"], gates, options=FusionOptions(*FUSION[opt]))
            assert np.abs(got - want).max() <= TOL1 * 4, (k, opt)


def test_drift_with_depth():
    n = 12
    for depth in (1, 5, 20, 50, 85, 90, 100):
        gates = C.tfim_trotter(n, depth)
        want = O.run(n, gates)
        c = C.to_circuit(n, gates)
        with qm.Simulator(n) as s:
            s.run(c)
            d = np.abs(s.amplitudes() - want).max()
            nerr = abs(s.norm() - 1)
        print(f"depth {depth:4d} gates {len(gates):5d} diff {d:.3e} norm_err {nerr:.3e}")
        assert d <= TOL_DEEP and nerr <= TOL_DEEP
        assert d <= tol_amp(len(gates)) and nerr <= tol_norm(len(gates)), depth


def test_launch_grouping_invariance():
    """Splitting a circuit into several qm_run calls at plan-op boundaries
    changes no arithmetic with one launch per plan op: the result is
    bit-identical (the reference's flush-cadence test, test_gpu.mm:189-222).
    With executor-level pass merging the cut also moves pass boundaries, so
    phases are rounded in another order: equal to 2e-6."""
    n = 12
    gates = C.tfim_trotter(n, 30)
    # cut after each RX layer: RX layers end a window, so plans concatenate
    layer = 3 * (n - 1) + n
    for merge, tol in ((False, 0.0), (True, 2e-6)):
        with qm.Simulator(n) as s:
            s.set_pass_merging(merge)
            s.run(C.to_circuit(n, gates))
            whole = s.amplitudes().astype(np.complex128)
        for cut in (1, 2, 7):
            with qm.Simulator(n) as s:
                s.set_pass_merging(merge)
                for a in range(0, 30, cut):
                    s.run(C.to_circuit(n, gates[a * layer:(a + cut) * layer]))
                got = s.amplitudes().astype(np.complex128)
            if merge:
                assert np.abs(got - whole).max() <= tol, cut
            else:
                assert np.array_equal(got, whole), cut


def test_qft_26_against_oracle():
    n = 26
    O.set_threads(16)
    try:
        prep = C.product_state_prep(n, seed=2024)
        want = O.run(n, prep + C.qft(n))
    finally:
        O.set_threads(1)
    with qm.Simulator(n) as s:
        s.run(C.to_circuit(n, prep))
        s.run(C.to_circuit(n, C.qft(n)))
        got = s.amplitudes().astype(np.complex128)
        nerr = abs(s.norm() - 1)
    d = np.abs(got - want).max()
    print(f"QFT26 on product state: max diff {d:.3e} (max |amp| {np.abs(want).max():.3e}) norm err {nerr:.3e}")
    assert d <= TOL1 and nerr <= TOL1


def test_fails_closed():
    with qm.Simulator(3) as s:
        c = qm.Circuit(3)
        c.cx(1, 1)
        with pytest.raises(qm.QMetalError, match="identical"):
            s.run(c)
        c = qm.Circuit(3)
        c.ccx(0, 1, 1)
        with pytest.raises(qm.QMetalError):
            s.run(c)
        c = qm.Circuit(4)
        with pytest.raises(qm.QMetalError, match="size mismatch"):
            s.run(c)
    with qm.Simulator(5) as s:
        # H on five wires lowers to Group(0..3) + Dense1q(4): a blockable run
        c = C.to_circuit(5, [(C.H, (q,), ()) for q in range(5)])
        with pytest.raises(qm.QMetalError, match="block_bits"):
            s.run(c, options=FusionOptions.blocked(13))
    with qm.Simulator(6) as s:
        with pytest.raises(qm.QMetalError, match="larger than the state"):
            s.run(C.to_circuit(6, [(C.H, (q,), ()) for q in range(5)]),
                  options=FusionOptions.blocked(8))
    with pytest.raises(qm.QMetalError):
        qm.Simulator(qm.max_qubits() + 1)


def test_blocked_plans_at_width():
    """Blocked1q (LDS-staged) plans where the block fits: b = 8, 10, 12."""
    for n, b in ((10, 8), (14, 10), (16, 12)):
        for gates in (C.tfim_trotter(n, 4), C.brickwork(n, 4), C.hardware_efficient(n, 3)):
            d = diff(n, gates, options=FusionOptions.blocked(b))
            assert d <= TOL_DEEP, (n, b, d)


def test_dispatch_count():
    """tests/test_gpu.mm:247-256: the constructor's reset counts one, a run
    one per plan op; the default (merged) executor keeps that count."""
    with qm.Simulator(8) as s:
        before = s.dispatch_count
        assert before == 1  # the constructor's reset
        c = C.to_circuit(8, C.ghz(8))
        s.run(c)
        s.synchronize()
        assert s.dispatch_count - before == len(c) == c.plan_size(True)
        q = C.to_circuit(8, C.qft(8))
        b = s.dispatch_count
        s.run(q)
        assert s.dispatch_count - b == q.plan_size(True)
        s.set_pass_merging(False)  # one launch per plan op: same count
        b = s.dispatch_count
        s.run(q)
        assert s.dispatch_count - b == q.plan_size(True)
