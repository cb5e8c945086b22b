"""Batched persistent launch against the one-circuit path and the oracle."""
import math

import numpy as np
import pytest

import qmetal_amd as qm
from qmetal_amd import circuits as C
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def mixed_instances():
    """Uneven widths, every gate kind the adjoint accepts, X/Y/Z terms."""
    out = []
    rng = C.SplitMix64(31)
    for k in range(40):
        n = 3 + k % 12
        g = C.hardware_efficient(n, 2, seed=k)
        g += [(C.H, (0,), ()), (C.CZ, (0, n - 1), ()), (C.S, (1,), ()), (C.T, (2 % n,), ())]
        if n >= 3:
            g += [(C.CCX, (0, 1, 2), ()), (C.SWAP, (0, n - 1), ()), (C.CY, (n - 1, 1), ())]
        g += [(C.RX, (q,), (rng.uniform(0, 6.28),)) for q in range(n)]
        h = C.tfim_ring_hamiltonian(n) + [(0.25, 0, 1, 0), (0.1, 0, 3 % (1 << n), 1 << (n - 1))]
        out.append((n, g, h))
    for n in (15, 16):  # wider than the LDS cut-off: HBM scratch path
        g = C.qaoa_ring(n, 2, [0.3, 0.5], [0.2, 0.9])
        out.append((n, g, C.ring_zz_hamiltonian(n)))
    return out


def plain(n, g, h):
    with qm.Simulator(n) as s:
        e, gr = s.energy_and_gradient(C.to_circuit(n, g), C.to_hamiltonian(h))
    return e, np.array(gr)


def test_batch_matches_plain_path_and_oracle():
    inst = mixed_instances()
    circs = [C.to_circuit(n, g) for n, g, _ in inst]
    hams = [C.to_hamiltonian(h) for _, _, h in inst]
    b = qm.Batch().prepare(circs, hams)
    for cfg in ((256, 1, 12), (512, 1, 13), (512, 2, 10)):
        try:
            b.set_config(*cfg)
        except qm.QMetalError:
            continue  # LDS cut-off beyond this device
        e, grads, ms = b.run()
        for i, (n, g, h) in enumerate(inst):
            e0, g0 = plain(n, g, h)
            assert abs(e[i] - e0) <= 1e-5, (cfg, i)
            assert np.abs(grads[i] - g0).max() <= 1e-5, (cfg, i)
    for i in (0, 7, 13, 29):
        n, g, h = inst[i]
        want = O.param_shift(n, g, h)
        assert np.abs(grads[i] - want).max() <= 1e-5, i
        assert abs(e[i] - O.expect_h(O.run(n, g), h)) <= 1e-5


def test_spread_instances_match_plain_path():
    """Every width split over many workgroups (one launch per program step
    for all spread instances), none, and the default (14 qubits and up):
    the same energies and gradients as the one-circuit path."""
    inst = mixed_instances()
    inst += [(n, C.qaoa_ring(n, 2, [0.7, 0.1], [0.4, 0.3]), C.ring_zz_hamiltonian(n))
             for n in (17, 18)]
    circs = [C.to_circuit(n, g) for n, g, _ in inst]
    hams = [C.to_hamiltonian(h) for _, _, h in inst]
    want = [plain(n, g, h) for n, g, h in inst]
    b = qm.Batch().prepare(circs, hams)
    results = {}
    for spread in (3, 0, 14, 9):
        b.set_spread(spread)
        e, grads, ms = b.run()
        results[spread] = (e, grads)
        worst = max(max(abs(e[i] - e0), np.abs(grads[i] - g0).max())
                    for i, (e0, g0) in enumerate(want))
        print(f"spread from {spread} qubits: {ms:.2f} ms, worst diff vs plain {worst:.2e}")
        assert worst <= 1e-5, spread
    e, grads, _ = b.run()  # a replay with the last setting: identical
    assert np.array_equal(e, results[9][0])
    assert all(np.array_equal(x, y) for x, y in zip(grads, results[9][1]))


def test_batch_graph_replays_are_identical():
    inst = C.qaoa_instances(64, lo=6, hi=14, seed=5)
    circs = [C.to_circuit(n, C.qaoa_ring(n, 6, ga, be)) for n, ga, be in inst]
    hams = [C.to_hamiltonian(C.ring_zz_hamiltonian(n)) for n, _, _ in inst]
    b = qm.Batch().prepare(circs, hams)
    cfg = b.autotune()
    print("autotuned:", cfg)
    e1, g1, _ = b.run()
    e2, g2, _ = b.run()
    assert np.array_equal(e1, e2)
    assert all(np.array_equal(x, y) for x, y in zip(g1, g2))
    for i in range(0, 64, 9):
        n, ga, be = inst[i]
        e0, g0 = plain(n, C.qaoa_ring(n, 6, ga, be), C.ring_zz_hamiltonian(n))
        assert abs(e1[i] - e0) <= 1e-5 and np.abs(g1[i] - g0).max() <= 1e-5


def test_autotuned_spread_batch_matches_plain_path():
    """Wide instances (14..18 qubits): autotune times spreading from 1, 3 or
    5 below the widest instance, or not at all, on the whole batch; whatever
    it picks, replays are identical and every instance matches the
    one-circuit path."""
    inst = C.qaoa_instances(24, p=3, lo=14, hi=18, seed=11)
    top = max(n for n, _, _ in inst)
    circs = [C.to_circuit(n, C.qaoa_ring(n, 3, ga, be)) for n, ga, be in inst]
    hams = [C.to_hamiltonian(C.ring_zz_hamiltonian(n)) for n, _, _ in inst]
    b = qm.Batch().prepare(circs, hams)
    cfg = b.autotune()
    print("autotuned:", cfg)
    assert cfg["spread_from"] == 0 or 14 <= cfg["spread_from"] <= top
    e1, g1, _ = b.run()
    e2, g2, _ = b.run()
    assert np.array_equal(e1, e2) and all(np.array_equal(x, y) for x, y in zip(g1, g2))
    for i, (n, ga, be) in enumerate(inst):
        e0, g0 = plain(n, C.qaoa_ring(n, 3, ga, be), C.ring_zz_hamiltonian(n))
        assert abs(e1[i] - e0) <= 1e-5 and np.abs(g1[i] - g0).max() <= 1e-5, i


def test_batch_fails_closed():
    b = qm.Batch()
    bad = C.to_circuit(3, [(C.H, (0,), ()), (C.CP, (0, 1), (0.4,))])
    with pytest.raises(qm.QMetalError, match="differentiable"):
        b.prepare([bad], [C.to_hamiltonian(C.ring_zz_hamiltonian(3))])
    wide = C.to_circuit(21, [(C.H, (0,), ())])
    with pytest.raises(qm.QMetalError, match="width"):
        b.prepare([wide], [qm.Hamiltonian()])


def test_bench_workload_against_fp64_adjoint():
    """The benchmarked workload (QAOA ring p=6, n ~ U{10..20}, the same
    splitmix64 draws), all 4096 instances in one autotuned batch: every
    instance against the one-circuit path (qm_energy_and_gradient), and the
    first instance of every width 10..20 against the complex128 adjoint on
    the oracle (itself pinned to parameter shift). Every gradient entry within
    1e-5 (absolute), the energy within 3e-6 relative to max(1, |E|) -- the
    state after ~150 complex64 gates carries ~1e-6 relative error, which the
    sum of n ring terms keeps (measured: gradients <= 3.6e-6 against the
    one-circuit path and <= 3.3e-6 against the oracle, energies 1.2e-6
    relative at |E| = 9.6)."""
    inst = C.qaoa_instances(4096)
    gates = [C.qaoa_ring(n, 6, ga, be) for n, ga, be in inst]
    terms = [C.ring_zz_hamiltonian(n) for n, _, _ in inst]
    circs = [C.to_circuit(n, g) for (n, _, _), g in zip(inst, gates)]
    hams = [C.to_hamiltonian(h) for h in terms]
    b = qm.Batch().prepare(circs, hams)
    cfg = b.autotune()
    e, grads, _ = b.run()
    sims = {n: qm.Simulator(n) for n in sorted({n for n, _, _ in inst})}
    worst_g, worst_e = 0.0, 0.0
    try:
        for i, (n, _, _) in enumerate(inst):
            e0, g0 = sims[n].energy_and_gradient(circs[i], hams[i])
            dg = float(np.abs(np.array(g0) - grads[i]).max())
            de = abs(e0 - e[i]) / max(1.0, abs(e0))
            assert dg <= 1e-5 and de <= 3e-6, (i, n, dg, de)
            worst_g, worst_e = max(worst_g, dg), max(worst_e, de)
    finally:
        for s in sims.values():
            s.close()
    print(f"all {len(inst)} instances vs the one-circuit path: gradients {worst_g:.2e}, "
          f"energies {worst_e:.2e} relative")
    first = {}
    for i, (n, _, _) in enumerate(inst):
        first.setdefault(n, i)
    assert sorted(first) == list(range(10, 21))
    O.set_threads(8)
    rows = []
    for n, i in sorted(first.items()):
        e0, g0 = O.adjoint(n, gates[i], terms[i])
        de, dg = abs(e[i] - e0), float(np.abs(grads[i] - g0).max())
        rows.append((n, i, e0, de, dg, float(np.abs(g0).max())))
        print(f"n={n} inst {i}: E {e0:.6f} |dE| {de:.2e} max|dgrad| {dg:.2e} "
              f"(max|grad| {rows[-1][5]:.3f})")
    O.set_threads(1)
    print(f"autotuned {cfg}")
    for n, i, e0, de, dg, gm in rows:
        assert dg <= 1e-5 and de <= 3e-6 * max(1.0, abs(e0)), (n, i, de, dg)


def test_batched_classes_run_the_merged_passes():
    """Spread instances of 14 qubits or more that share width, gate
    structure and Hamiltonian run as one batched program of the
    single-circuit merged passes per class (every pass one launch over the
    class). Three classes (QAOA p=3 at 14, 15, 16 qubits, several instances
    each with their own angles) plus one instance whose Hamiltonian differs
    (a class of its own): all against the one-circuit path, replays
    bit-identical, and the batched count as expected under each spread
    width."""
    rng = C.SplitMix64(77)
    inst = []
    for n in (14, 15, 16, 14, 15, 16, 16):
        ga = [rng.uniform(0, 3.14) for _ in range(3)]
        be = [rng.uniform(0, 1.57) for _ in range(3)]
        inst.append((n, C.qaoa_ring(n, 3, ga, be), C.ring_zz_hamiltonian(n)))
    n = 15
    odd_h = C.ring_zz_hamiltonian(n) + [(0.3, 1 << 4, 0, 0)]
    inst.append((n, C.qaoa_ring(n, 3, [0.2, 0.4, 0.6], [0.1, 0.3, 0.5]), odd_h))
    circs = [C.to_circuit(n, g) for n, g, _ in inst]
    hams = [C.to_hamiltonian(h) for _, _, h in inst]
    want = [plain(n, g, h) for n, g, h in inst]
    b = qm.Batch().prepare(circs, hams)
    for spread, batched in ((14, 8), (16, 3), (0, 0)):
        b.set_spread(spread)
        assert b.batched_instances() == batched, (spread, b.batched_instances())
        e1, g1, _ = b.run()
        e2, g2, _ = b.run()
        assert np.array_equal(e1, e2) and all(np.array_equal(x, y) for x, y in zip(g1, g2))
        worst = max(max(abs(e1[i] - e0), np.abs(g1[i] - g0).max())
                    for i, (e0, g0) in enumerate(want))
        print(f"spread from {spread}: {batched} batched instances, worst diff vs plain {worst:.2e}")
        assert worst <= 1e-5, spread


def test_batched_classes_with_terms_no_tile_holds():
    """Batched classes whose Hamiltonian has flip groups on high wires, Y
    terms and a string flipping every wire (no tile holds it: the batched
    generic co-state launches), several instances per class at 15 and 16
    qubits: against the one-circuit path, replays bit-identical."""
    rng = C.SplitMix64(78)
    inst = []
    for n in (15, 16, 15, 16, 15):
        ga = [rng.uniform(0, 3.14) for _ in range(2)]
        be = [rng.uniform(0, 1.57) for _ in range(2)]
        h = (C.tfim_ring_hamiltonian(n) + [(0.3, 0, 1 << q, 0) for q in range(0, n, 3)]
             + [(0.2, (1 << 9) | (1 << 12), 0, 1)] + [(0.4, (1 << n) - 1, 0, 0)])
        inst.append((n, C.qaoa_ring(n, 2, ga, be), h))
    circs = [C.to_circuit(n, g) for n, g, _ in inst]
    hams = [C.to_hamiltonian(h) for _, _, h in inst]
    want = [plain(n, g, h) for n, g, h in inst]
    b = qm.Batch().prepare(circs, hams)
    b.set_spread(14)
    assert b.batched_instances() == len(inst)
    e1, g1, _ = b.run()
    e2, g2, _ = b.run()
    assert np.array_equal(e1, e2) and all(np.array_equal(x, y) for x, y in zip(g1, g2))
    worst = max(max(abs(e1[i] - e0), np.abs(g1[i] - g0).max()) for i, (e0, g0) in enumerate(want))
    print(f"batched classes with wide terms: worst diff vs plain {worst:.2e}")
    assert worst <= 1e-5
