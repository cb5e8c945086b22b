"""FP64 reductions, sampling and the adjoint gradient on the GPU against the
complex128 oracle and closed forms (reference tests/test_gpu.mm:258-480 and
tests/test_python.py:34-99, restated)."""
import math
import os

import numpy as np
import pytest

import qmetal_amd as qm
from qmetal_amd import circuits as C
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def sim_with(n, gates):
    s = qm.Simulator(n)
    s.run(C.to_circuit(n, gates))
    return s


def test_expectation_closed_forms_on_ghz():
    n = 8
    with sim_with(n, C.ghz(n)) as s:
        assert abs(s.expectation_pauli(z=1)) <= 1e-5
        assert abs(s.expectation_pauli(z=0b11) - 1) <= 1e-5
        assert abs(s.expectation_pauli(x=1)) <= 1e-5
        assert abs(s.expectation_pauli(x=(1 << n) - 1) - 1) <= 1e-5
        assert abs(s.expectation_pauli() - 1) <= 1e-5
        assert abs(s.expectation("Z" + "I" * (n - 1))) <= 1e-5
        assert abs(s.expectation("ZZ" + "I" * (n - 2)) - 1) <= 1e-5
    with qm.Simulator(6) as s:
        for q in range(6):
            assert abs(s.expectation_pauli(z=1 << q) - 1) <= 1e-6


def test_y_sign_closed_forms():
    with sim_with(1, [(C.H, (0,), ()), (C.S, (0,), ())]) as s:
        assert abs(s.expectation("Y") - 1) <= 1e-6
        assert abs(s.expectation("X")) <= 1e-6
    with sim_with(1, [(C.H, (0,), ()), (C.SDG, (0,), ())]) as s:
        assert abs(s.expectation("Y") + 1) <= 1e-6
    with sim_with(2, [(C.H, (0,), ()), (C.CX, (0, 1), ())]) as s:
        assert abs(s.expectation("YY") + 1) <= 1e-6
        assert abs(s.expectation("XX") - 1) <= 1e-6
        assert abs(s.expectation("ZZ") - 1) <= 1e-6


def test_random_paulis_against_oracle():
    for n in (10, 17):
        gates = C.random_circuit(n, 4, seed=0xABC)
        ref = O.run(n, gates)
        rng = C.SplitMix64(7)
        with sim_with(n, gates) as s:
            for _ in range(24):
                x = y = z = 0
                for q in range(n):
                    r = rng.next_u64() % 4
                    x |= (r == 1) << q
                    y |= (r == 2) << q
                    z |= (r == 3) << q
                want = O.expect(ref, x, y, z)
                got = s.expectation_pauli(x, y, z)
                assert abs(want - got) <= 1e-5, (n, x, y, z, want, got)


def test_hamiltonian_sum_and_norm():
    n = 8
    gates = C.hardware_efficient(n, 2, seed=0xFEED)
    ref = O.run(n, gates)
    h = qm.Hamiltonian()
    terms = []
    for q in range(n - 1):
        terms += [(1.0, 0, 0, 3 << q), (0.3, 1 << q, 0, 0)]
    for t in terms:
        h.add(*t)
    with sim_with(n, gates) as s:
        assert abs(O.expect_h(ref, terms) - s.expectation(h)) <= 1e-4
        assert abs(s.norm() - 1.0) <= 1e-5


def _mixed_terms(n, seed):
    """Diagonal, single-flip, two-wire, Y-carrying (i^ny phases), flip-bit-0
    and many-high-wire terms: more than one tile's worth of diagonal terms
    and flip groups, plus flips that no tile can hold."""
    rng = C.SplitMix64(seed)
    terms = []
    for q in range(n):
        terms.append((1.0, 0, 0, (1 << q) | (1 << ((q + 1) % n))))
        terms.append((0.7, 1 << q, 0, 0))
        terms.append((-0.4, 0, 1 << q, 0))
    for q in range(0, n - 1, 2):
        terms.append((0.25, 1 << q, 1 << (q + 1), 0))
        terms.append((0.15, 0, 3 << q, 1 << ((q + 5) % n)))
    for _ in range(40):
        z = rng.next_u64() & ((1 << n) - 1)
        terms.append((rng.uniform(-1, 1), 0, 0, z))
    for _ in range(30):
        x = rng.next_u64() & ((1 << n) - 1)
        y = rng.next_u64() & ((1 << n) - 1) & ~x
        z = rng.next_u64() & ((1 << n) - 1) & ~(x | y)
        terms.append((rng.uniform(-1, 1), x, y, z))
    terms.append((0.35, 3, 0, 0))  # X0 X1: flip with amplitude bit 0 across lanes
    terms.append((0.3, 0, 0, 0))  # identity
    terms.append((0.7, 1 << (n // 2), 0, 0))  # a duplicate of a single-X term
    terms.append((-0.25, 0, 0, 3))  # a duplicate of a ZZ term
    terms.append((-0.45, 1 | (1 << (n - 1)), 0, 1 << 2))
    terms.append((0.55, 1 << (n - 2), 0, 1 << (n - 3)))
    top = ((1 << n) - 1) & ~0x7F
    terms.append((0.5, top, 0, 1))  # every wire of 7 and up (six or more from n = 13)
    terms.append((-0.3, 1, top & -top, 0))
    return terms


@pytest.fixture
def tiles_always(monkeypatch):
    """Simulators made under it take the tile passes for expectation(H)
    however small the state (by default, below 256 MiB of per-term reads one
    launch of per-term reductions is faster)."""
    monkeypatch.setenv("QMETAL_EXPECT_TILES", "1")


@pytest.mark.parametrize("n", [12, 14, 20])
def test_hamiltonian_tile_passes_match_per_term_and_oracle(n, tiles_always):
    """expectation(H) sums tile-held terms in read-only co-state tile passes;
    it equals the one-reduction-per-term path on the same state and the
    complex128 oracle on the oracle's state."""
    gates = C.random_circuit(n, 5, seed=0x5EED + n)
    ref = O.run(n, gates)
    terms = _mixed_terms(n, 0xC0FFEE + n)
    h = qm.Hamiltonian()
    for t in terms:
        h.add(*t)
    scale = sum(abs(t[0]) for t in terms)
    want = O.expect_h(ref, terms)
    with sim_with(n, gates) as s:
        got = s.expectation(h)
        s.set_pass_merging(False)
        per_term = s.expectation(h)
        s.set_pass_merging(True)
        assert s.expectation(h) == got  # deterministic
    assert abs(got - per_term) <= 1e-6 * scale, (n, got, per_term)
    assert abs(got - want) <= 1e-5 * scale, (n, got, want)


def test_hamiltonian_many_terms_per_flip_and_per_tile(tiles_always):
    """Groups and diagonal sums larger than one tile holds: 70 diagonal
    strings, 60 terms sharing one flip, 40 single-wire groups on wires 0..6
    repeated, against the per-term path and the oracle."""
    n = 14
    gates = C.random_circuit(n, 4, seed=0xD1A9)
    ref = O.run(n, gates)
    rng = C.SplitMix64(99)
    m = (1 << n) - 1
    terms = [(rng.uniform(-1, 1), 0, 0, rng.next_u64() & m) for _ in range(70)]
    flip = (1 << 2) | (1 << 9)
    for _ in range(60):
        y = rng.next_u64() & flip
        terms.append((rng.uniform(-1, 1), flip & ~y, y, rng.next_u64() & m & ~flip))
    for k in range(40):
        terms.append((rng.uniform(-1, 1), 1 << (k % 7), 0, rng.next_u64() & m & ~(1 << (k % 7))))
    h = qm.Hamiltonian()
    for t in terms:
        h.add(*t)
    scale = sum(abs(t[0]) for t in terms)
    with sim_with(n, gates) as s:
        got = s.expectation(h)
        s.set_pass_merging(False)
        per_term = s.expectation(h)
    assert abs(got - per_term) <= 1e-6 * scale, (got, per_term)
    assert abs(got - O.expect_h(ref, terms)) <= 1e-5 * scale


@pytest.mark.parametrize("n", [16, 20])
def test_jordan_wigner_hopping_terms(n, tiles_always):
    """Jordan-Wigner hopping and pairing strings X_a Z..Z X_b, Y_a Z..Z Y_b,
    X_a Z..Z Y_b (two flipped wires, Z on every wire between: signs over
    lanes, waves, registers and tile bits) plus number terms Z_a: through the
    tile passes against the complex128 oracle."""
    gates = C.random_circuit(n, 4, seed=0x4A + n)
    ref = O.run(n, gates)
    rng = C.SplitMix64(0x4A57 + n)
    terms = []
    for _ in range(60):
        a, b = sorted(rng.next_u64() % n for _ in range(2))
        if a == b:
            continue
        zs = ((1 << b) - 1) & ~((1 << (a + 1)) - 1)
        c = rng.uniform(-1, 1)
        kind = rng.next_u64() % 3
        if kind == 0:
            terms.append((c, (1 << a) | (1 << b), 0, zs))
        elif kind == 1:
            terms.append((c, 0, (1 << a) | (1 << b), zs))
        else:
            terms.append((c, 1 << a, 1 << b, zs))
    terms += [(rng.uniform(-1, 1), 0, 0, 1 << q) for q in range(n)]
    h = qm.Hamiltonian()
    for t in terms:
        h.add(*t)
    scale = sum(abs(t[0]) for t in terms)
    with sim_with(n, gates) as s:
        got = s.expectation(h)
        s.set_pass_merging(False)
        per_term = s.expectation(h)
    assert abs(got - per_term) <= 1e-6 * scale, (got, per_term)
    assert abs(got - O.expect_h(ref, terms)) <= 1e-5 * scale


def test_hamiltonian_single_term_and_diagonal_only(tiles_always):
    """One term keeps the FP64 reduction; a diagonal-only Hamiltonian (one
    tile) and an X-only one (flip groups only) match the per-term path."""
    n = 16
    gates = C.brickwork(n, 3, seed=11)
    with sim_with(n, gates) as s:
        for terms in ([(0.9, 0, 0, 1 << 9)],
                      [(1.0, 0, 0, 3 << q) for q in range(n - 1)],
                      [(0.5, 1 << q, 0, 0) for q in range(n)]):
            h = qm.Hamiltonian()
            for t in terms:
                h.add(*t)
            got = s.expectation(h)
            s.set_pass_merging(False)
            per_term = s.expectation(h)
            s.set_pass_merging(True)
            assert abs(got - per_term) <= 1e-6 * sum(abs(t[0]) for t in terms), (terms[0], got, per_term)


def test_hamiltonian_small_state_default_path():
    """Default path on a state below the tile threshold: all terms in one
    launch of per-term FP64 reductions, equal to the oracle, and to the tile
    passes on the same state to 1e-6 of sum |c|."""
    n = 13
    gates = C.random_circuit(n, 4, seed=0x13)
    ref = O.run(n, gates)
    terms = _mixed_terms(n, 0xABCD)
    terms += [(0.01 * k, 1 << (k % n), 0, 1 << ((k + 3) % n)) for k in range(150)]
    h = qm.Hamiltonian()
    for t in terms:
        h.add(*t)
    scale = sum(abs(t[0]) for t in terms)
    with sim_with(n, gates) as s:
        got = s.expectation(h)
    assert abs(got - O.expect_h(ref, terms)) <= 1e-5 * scale
    os.environ["QMETAL_EXPECT_TILES"] = "1"
    try:
        with sim_with(n, gates) as s:
            tiles = s.expectation(h)
    finally:
        del os.environ["QMETAL_EXPECT_TILES"]
    assert abs(got - tiles) <= 1e-6 * scale, (got, tiles)


def test_norm_is_fp64_accurate_at_scale():
    """The GPU norm of 2^24 complex64 amplitudes equals the FP64 norm of the
    same amplitudes to 1e-12 (an FP32 accumulation would be off by ~1e-7)."""
    n = 24
    with sim_with(n, C.brickwork(n, 2, seed=4)) as s:
        a = s.amplitudes()
        want = math.sqrt(float(np.sum(a.real.astype(np.float64) ** 2 + a.imag.astype(np.float64) ** 2)))
        assert abs(s.norm() - want) <= 1e-12
        z = s.expectation_pauli(z=1 << 5)
        zz = np.where((np.arange(1 << n) >> 5) & 1, -1.0, 1.0)
        want_z = float(np.sum(zz * (a.real.astype(np.float64) ** 2 + a.imag.astype(np.float64) ** 2)))
        assert abs(z - want_z) <= 1e-12


def test_sampling_ghz_poles_and_determinism():
    n = 6
    with sim_with(n, C.ghz(n)) as s:
        shots = s.sample(20000, 12345)
        assert set(shots) <= {0, (1 << n) - 1}
        frac = shots.count(0) / len(shots)
        assert abs(frac - 0.5) < 0.02
        counts = s.counts(4000, seed=7)
        assert set(counts) == {"0" * n, "1" * n}
        assert abs(counts["0" * n] / 4000 - 0.5) < 0.03
    with sim_with(5, C.qft(5)) as s:
        a, b, c = s.sample(500, 999), s.sample(500, 999), s.sample(500, 1000)
        assert a == b and a != c


def test_sampling_total_variation():
    n = 6
    gates = C.random_circuit(n, 3, seed=0x1234)
    probs = np.abs(O.run(n, gates)) ** 2
    with sim_with(n, gates) as s:
        draw = np.array(s.sample(200000, 2024))
    hist = np.bincount(draw, minlength=1 << n) / len(draw)
    tv = 0.5 * np.abs(hist - probs).sum()
    print(f"total variation {tv:.4f}")
    assert tv < 0.02


def test_sampling_matches_oracle_procedure_on_gpu_amplitudes():
    """Given the GPU's own amplitudes, the oracle's two-level procedure picks
    the same index for every shot except ones whose uniform lies on a CDF
    boundary: each differing shot's u = (r >> 11) 2^-53 total must sit within
    round-off of the cumulative probability at one of the two indices
    (the GPU sums blocks and prefixes in another order than the oracle)."""
    shots_n = 20000
    us = np.array([(r >> 11) * 2.0 ** -53 for r in O.splitmix64_stream(99, shots_n)])
    for n, gates in ((11, C.qaoa_ring(11, 2)), (14, C.brickwork(14, 4)), (20, C.qft(20))):
        with sim_with(n, gates) as s:
            shots = np.array(s.sample(shots_n, 99), dtype=np.uint64)
            amps = s.amplitudes().astype(np.complex128)
        want = O.sample(amps, shots_n, 99)
        mism = np.nonzero(shots != want)[0]
        cum = np.cumsum(np.abs(amps) ** 2)
        total = cum[-1]
        for k in mism:
            u = us[k] * total
            edges = []
            for i in (int(shots[k]), int(want[k])):
                edges += [cum[i], cum[i - 1] if i else 0.0]
            gap = min(abs(u - e) for e in edges)
            print(f"n={n}: shot {k} GPU {int(shots[k])} oracle {int(want[k])}, "
                  f"|u - CDF edge| = {gap:.2e}")
            assert gap <= 1e-12 * total, (n, k)
        print(f"n={n}: {len(mism)} of {shots_n} shots differ from the oracle procedure")
        assert len(mism) <= 2


def test_ghz20_invariants_and_1e5_shots():
    n = 20
    with sim_with(n, C.ghz(n)) as s:
        assert abs(s.norm() - 1) <= 1e-5
        for q in range(n):
            assert abs(s.expectation_pauli(z=1 << q)) <= 1e-5
        assert abs(s.expectation_pauli(z=1 | (1 << (n - 1))) - 1) <= 1e-5
        shots = set(s.sample(100000, 7))
        assert shots == {0, (1 << n) - 1}


@pytest.mark.parametrize("n", [33, 34])
def test_ghz_full_size_invariants(n):
    """GHZ at 64 and 128 GiB (one allocation), merged passes: norm 1,
    <Z_q> = 0, <Z_0 Z_(n-1)> = 1, <X...X> = 1 (the two poles' coherence), a
    100-term Hamiltonian's <H> = n + 2 through the tile passes, and shots only on the two poles, both of them, near half each (4000
    shots: 2000 +- 32 sd, the bound is 6 sd)."""
    top = (1 << n) - 1
    with sim_with(n, C.ghz(n)) as s:
        norm = s.norm()
        zq = max(abs(s.expectation_pauli(z=1 << q)) for q in (0, 1, 12, 13, n // 2, n - 1))
        zz = s.expectation_pauli(z=1 | (1 << (n - 1)))
        xx = s.expectation_pauli(x=top)
        # <H> through the read-only tile passes: every Z_i Z_j is 1 and every
        # string flipping fewer than n wires is 0 on GHZ; X..X (too wide for a
        # tile: one reduction) is 1.
        terms = [(1.0, 0, 0, (1 << q) | (1 << ((q + 1) % n))) for q in range(n)]
        terms += [(0.5, (1 << q) | (1 << ((q + 1) % n)), 0, 0) for q in range(n)]
        terms += [(-0.3, 0, 1 << q, 0) for q in range(n)] + [(2.0, top, 0, 0)]
        h = qm.Hamiltonian()
        for t in terms:
            h.add(*t)
        eh = s.expectation(h)
        shots = s.sample(4000, 2024)
    ones = sum(1 for v in shots if v == top)
    print(f"GHZ{n}: norm {norm:.8f}, max|<Z_q>| {zq:.2e}, <Z0 Z{n - 1}> {zz:.8f}, "
          f"<X..X> {xx:.8f}, <H> {eh:.8f} (n + 2 = {n + 2}), {ones} of 4000 shots on |1..1>")
    assert abs(norm - 1) <= 1e-5 and zq <= 1e-5
    assert abs(zz - 1) <= 1e-5 and abs(xx - 1) <= 1e-5
    assert abs(eh - (n + 2)) <= 1e-5 * sum(abs(t[0]) for t in terms)
    assert set(shots) == {0, top} and abs(ones - 2000) <= 190


def tfim6():
    n = 6
    h = [(1.0, 0, 0, 3 << q) for q in range(n - 1)] + [(0.4, 1 << q, 0, 0) for q in range(n)]
    h.append((0.25, 0, 1, 0))  # a Y term, so an i^ny sign error cannot hide
    return n, h


@pytest.mark.parametrize("name", ["hea_d2", "tfim_s3", "qaoa_p2"])
def test_gradient_against_parameter_shift(name):
    n, terms = tfim6()
    gates = {"hea_d2": C.hardware_efficient(n, 2, seed=0xC0FFEE),
             "tfim_s3": C.tfim_trotter(n, 3), "qaoa_p2": C.qaoa_ring(n, 2)}[name]
    c = C.to_circuit(n, gates)
    with qm.Simulator(n) as s:
        e, g = s.energy_and_gradient(c, C.to_hamiltonian(terms))
    assert abs(e - O.expect_h(O.run(n, gates), terms)) <= 1e-5
    want = O.param_shift(n, gates, terms)
    worst = float(np.abs(np.array(g) - want).max())
    print(f"{name}: max |adjoint - shift| = {worst:.3e}")
    assert worst <= 1e-5
    for k, (kind, _, _) in enumerate(gates):
        if kind not in (C.RX, C.RY, C.RZ):
            assert g[k] == 0.0


def test_gradient_ry_closed_form():
    for t in (0.0, 0.3, 1.1, 2.7, -0.9, 0.4, 1.3, -0.7):
        with qm.Simulator(2) as s:
            e, g = s.energy_and_gradient(C.to_circuit(2, [(C.RY, (0,), (t,))]),
                                         qm.Hamiltonian().add_string(1.0, "ZI"))
        assert abs(e - math.cos(t)) <= 1e-6 and abs(g[0] + math.sin(t)) <= 1e-6


def test_gradient_follows_num_params():
    """Through the C ABI, as simulator.mm:822-828, :917: RX added with no
    parameter gets gradient 0, a P without one is accepted, a P with one
    fails; the rest equals the complex128 adjoint."""
    n = 3
    gates = [(C.H, (0,), ()), (C.RX, (0,), ()), (C.RY, (1,), (0.7,)), (C.P, (1,), ()),
             (C.CX, (1, 2), ()), (C.RZ, (2,), (0.3,)), (C.RX, (1,), (1.1,))]
    terms = [(1.0, 0, 0, 1), (0.5, 2, 0, 0), (0.7, 0, 0, 6)]
    for merge in (True, False):
        with qm.Simulator(n) as s:
            s.set_pass_merging(merge)
            e, g = s.energy_and_gradient(C.to_circuit(n, gates), C.to_hamiltonian(terms))
        e_ref, g_ref = O.adjoint(n, gates, terms)
        assert g[1] == 0.0 and g[3] == 0.0
        assert abs(e - e_ref) <= 1e-6 and np.abs(np.array(g) - g_ref).max() <= 1e-6
    b = qm.Batch()
    b.prepare([C.to_circuit(n, gates)], [C.to_hamiltonian(terms)])
    eb, gb, _ = b.run()
    b.close()
    assert gb[0][1] == 0.0 and gb[0][3] == 0.0
    assert abs(eb[0] - e_ref) <= 1e-6 and np.abs(gb[0] - g_ref).max() <= 1e-6
    bad = C.to_circuit(n, [(C.P, (1,), (0.2,))])
    with qm.Simulator(n) as s:
        with pytest.raises(qm.QMetalError, match="differentiable"):
            s.energy_and_gradient(bad, C.to_hamiltonian(terms))


def test_gradient_fails_closed():
    bad = qm.Circuit(2)
    bad.h(0).cp(0, 1, 0.4)
    with qm.Simulator(2) as s:
        with pytest.raises(qm.QMetalError, match="differentiable"):
            s.energy_and_gradient(bad, qm.Hamiltonian().add_string(1.0, "ZI"))


def test_gradient_hea_n12_against_oracle_adjoint():
    """Wider circuit, wires past the fixed-bit threshold, TFIM ring: GPU
    adjoint vs the complex128 adjoint."""
    n = 12
    gates = C.hardware_efficient(n, 3, seed=0x4EA)
    terms = C.tfim_ring_hamiltonian(n)
    with qm.Simulator(n) as s:
        e, g = s.energy_and_gradient(C.to_circuit(n, gates), C.to_hamiltonian(terms))
    e_ref, g_ref = O.adjoint(n, gates, terms)
    assert abs(e - e_ref) <= 1e-5
    assert np.abs(np.array(g) - g_ref).max() <= 1e-5


def test_scale_sweep_invariants_every_width():
    """The reference's scale sweep (bench/scale.mm:30-90, results/scale_n33.txt)
    at every width from 20 to 34, one allocation each: GHZ norm 1, <Z_0> = 0,
    <Z_0 Z_(n-1)> = 1, 4096 shots only on the two poles; QFT|0...0> norm 1 and
    <Z_q> = 0 on every qubit (uniform image). Bounds 1e-5, tighter than the
    reference's 1e-3 / 1e-2."""
    for n in range(20, 35):
        top = (1 << n) - 1
        with sim_with(n, C.ghz(n)) as s:
            norm = s.norm()
            z0 = s.expectation_pauli(z=1)
            zz = s.expectation_pauli(z=1 | (1 << (n - 1)))
            shots = s.sample(4096, 7)
            s.reset()
            s.run(C.to_circuit(n, C.qft(n)))
            qnorm = s.norm()
            qz = max(abs(s.expectation_pauli(z=1 << q)) for q in range(n))
        print(f"n={n}: GHZ norm {norm:.8f} <Z0> {z0:+.1e} <Z0 Zn> {zz:.8f}; "
              f"QFT norm {qnorm:.8f} max|<Zq>| {qz:.1e}", flush=True)
        assert abs(norm - 1) <= 1e-5 and abs(z0) <= 1e-5 and abs(zz - 1) <= 1e-5
        assert set(shots) <= {0, top} and len(set(shots)) == 2
        assert abs(qnorm - 1) <= 1e-5 and qz <= 1e-5
