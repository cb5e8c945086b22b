"""Merged HBM passes on the GPU: several plan ops applied between one load and
one store of each amplitude. Checked against the complex128 oracle and against
the same plan run one launch per op; launch accounting follows the grouping.

Tolerances: 1e-5 max amplitude difference after one gate (complex64 floor),
1e-4 for deep circuits; merged vs op-by-op 2e-6 (same FP32 arithmetic, phases
summed in another FP64 order).
"""
import numpy as np
import pytest

import qmetal_amd as qm
from qmetal_amd import FusionOptions
from qmetal_amd import circuits as C
from oracle import oracle as O

pytestmark = pytest.mark.gpu

TOL1 = 1e-5
TOL_DEEP = 1e-4
TOL_SAME = 2e-6

ONE = (C.X, C.Y, C.Z, C.H, C.S, C.SDG, C.T, C.TDG)
ROT = (C.RX, C.RY, C.RZ, C.P)


def mixed_stream(n, count, seed):
    """Every gate kind on wires drawn uniformly (splitmix64), so controls and
    targets land on register, thread-index, fixed and fill bits alike."""
    r = C.SplitMix64(seed)
    gates = [(C.H, (q,), ()) for q in range(n)]

    def wires(k):
        out = []
        while len(out) < k:
            q = r.next_u64() % n
            if q not in out:
                out.append(q)
        return tuple(out)

    for _ in range(count):
        u = r.next_u64() % 16
        if u < 3:
            gates.append((ONE[r.next_u64() % len(ONE)], wires(1), ()))
        elif u < 6:
            gates.append((ROT[r.next_u64() % len(ROT)], wires(1), (r.uniform(-3.2, 3.2),)))
        elif u < 7:
            gates.append((C.U3, wires(1), tuple(r.uniform(-3.2, 3.2) for _ in range(3))))
        elif u < 10:
            gates.append(((C.CX, C.CY, C.CZ, C.SWAP)[r.next_u64() % 4], wires(2), ()))
        elif u < 13:
            gates.append((C.CP, wires(2), (r.uniform(-3.2, 3.2),)))
        elif u < 14:
            gates.append((C.CCX, wires(3), ()))
        else:  # a ZZ coupling, recognised as a diagonal run
            a, b = wires(2)
            gates += C.zz(a, b, r.uniform(-3.2, 3.2))
    return gates


def shapes(c, options=None):
    """Shape of every pass qm_run makes for `c` (after validating the
    schedule: every op once, overlapping ops in plan order)."""
    from test_schedule import check_schedule, pass_shape
    check_schedule(c, options)
    plan = c.plan(options)
    return [pass_shape(plan, e, c.num_qubits) for e in c.schedule(options)]


def run(n, gates, merge=True, options=None, prep=None):
    with qm.Simulator(n) as s:
        s.set_pass_merging(merge)
        if prep:
            s.run(C.to_circuit(n, prep))
        before = s.launch_count
        s.run(C.to_circuit(n, gates), options=options)
        s.synchronize()
        launches = s.launch_count - before
        return s.amplitudes().astype(np.complex128), s.norm(), launches


def test_mixed_streams_against_oracle():
    for n in (5, 6, 8, 9, 11, 13):
        for seed in range(6):
            gates = mixed_stream(n, 40 + 10 * seed, 1000 * n + seed)
            want = O.run(n, gates)
            got, norm, _ = run(n, gates)
            ref, _, _ = run(n, gates, merge=False)
            d, dm = np.abs(got - want).max(), np.abs(got - ref).max()
            assert d <= TOL_DEEP and abs(norm - 1) <= TOL_DEEP, (n, seed, d)
            assert dm <= TOL_SAME, (n, seed, dm)


def test_mixed_streams_wide_states_every_fusion_setting():
    """Every gate kind on any wires at 16 and 18 qubits, where the scheduler
    reorders across disjoint wires, splits groups and uses every pass shape,
    under every fusion setting: the schedule is validated on the host, the
    amplitudes against the oracle and against one launch per plan op."""
    fusion = {"fused": (1, 1, 1, 4, 0), "none": (0, 0, 0, 1, 0), "diag_only": (0, 1, 1, 1, 0),
              "win_only": (1, 0, 0, 4, 0), "blocked8": (1, 1, 1, 4, 8)}
    for n in (16, 18):
        for seed in range(2):
            gates = mixed_stream(n, 120, 4000 * n + seed)
            want = O.run(n, gates)
            c = C.to_circuit(n, gates)
            for opt, f in fusion.items():
                options = FusionOptions(*f)
                kinds = shapes(c, options)
                got, norm, launches = run(n, gates, options=options)
                ref, _, _ = run(n, gates, merge=False, options=options)
                d, dm = np.abs(got - want).max(), np.abs(got - ref).max()
                print(f"mixed n={n} seed={seed} {opt}: {launches} launches "
                      f"{ {k: kinds.count(k) for k in set(kinds)} } diff {d:.2e} vs op-by-op {dm:.2e}")
                assert launches == len(kinds)
                assert d <= TOL_DEEP and abs(norm - 1) <= TOL_DEEP, (n, seed, opt, d)
                assert dm <= TOL_SAME, (n, seed, opt, dm)


@pytest.mark.parametrize("name,n", [
    ("qft", 14), ("qft", 20), ("ghz", 20), ("qaoa", 16), ("tfim", 14), ("grover", 12),
    ("hea", 14), ("brickwork", 16), ("random", 12)])
def test_configurations_merged_vs_op_by_op(name, n):
    gates = {
        "qft": lambda: C.qft(n), "ghz": lambda: C.ghz(n), "qaoa": lambda: C.qaoa_ring(n, 6),
        "tfim": lambda: C.tfim_trotter(n, 10), "grover": lambda: C.grover_proxy(n, 4),
        "hea": lambda: C.hardware_efficient(n, 3), "brickwork": lambda: C.brickwork(n, 10),
        "random": lambda: C.random_circuit(n, 10),
    }[name]()
    prep = C.product_state_prep(n, seed=99)
    want = O.run(n, prep + gates)
    c = C.to_circuit(n, gates)
    for opt in ("fused", "none", "diag_only", "win_only"):
        options = FusionOptions(*{"fused": (1, 1, 1, 4, 0), "none": (0, 0, 0, 1, 0),
                                  "diag_only": (0, 1, 1, 1, 0),
                                  "win_only": (1, 0, 0, 4, 0)}[opt])
        got, norm, launches = run(n, gates, options=options, prep=prep)
        ref, _, ref_launches = run(n, gates, merge=False, options=options, prep=prep)
        d, dm = np.abs(got - want).max(), np.abs(got - ref).max()
        passes = c.schedule(options)
        print(f"{name:9s} n={n:2d} {opt:9s} plan={len(passes):4d} launches={launches:4d} "
              f"diff={d:.2e} vs op-by-op {dm:.2e}")
        assert d <= TOL_DEEP and abs(norm - 1) <= TOL_DEEP, (opt, d)
        assert dm <= TOL_SAME, (opt, dm)
        assert launches == len(passes)
        assert ref_launches == sum(1 for k, op in enumerate(c.plan(options))
                                   if not (op["kind"] == "DiagonalLayer" and not op["masks"]))


def test_dispatch_count_follows_plan_launch_count_follows_grouping():
    """dispatch_count adds one per plan op whether or not ops are merged (the
    reference's accounting, tests/test_gpu.mm:247-256); launch_count adds one
    per plan op with merging off and one per merged pass with it on."""
    c = C.to_circuit(8, C.ghz(8))
    with qm.Simulator(8) as s:
        s.set_pass_merging(False)
        before, lb = s.dispatch_count, s.launch_count
        s.run(c)
        s.synchronize()
        assert s.dispatch_count - before == len(c) == c.plan_size(True)
        assert s.launch_count - lb == len(c)
        s.set_pass_merging(True)
        before, lb = s.dispatch_count, s.launch_count
        s.run(c)
        s.synchronize()
        assert s.dispatch_count - before == len(c)
        assert s.launch_count - lb == len(c.schedule()) < len(c)


def test_qft_24_merged_vs_op_by_op():
    n = 24
    prep = C.product_state_prep(n, seed=2024)
    got, norm, launches = run(n, C.qft(n), prep=prep)
    ref, _, ref_launches = run(n, C.qft(n), merge=False, prep=prep)
    dm = np.abs(got - ref).max()
    print(f"QFT24: {launches} launches vs {ref_launches}; merged vs op-by-op {dm:.2e}")
    assert launches == 3 and ref_launches == 2 * n - 1
    assert dm <= TOL_SAME and abs(norm - 1) <= TOL1


@pytest.mark.parametrize("n,want", [(13, 1), (14, 2), (20, 2)])
def test_qft_four_wave_low_pass_against_oracle(n, want):
    """QFT wires 12..0 in one low pass over four-wave 2^13 tiles (after a
    wide pass on 19..13 at n=20, a register pass on 13 at n=14)."""
    prep = C.product_state_prep(n, seed=77 + n)
    got, norm, launches = run(n, C.qft(n), prep=prep)
    ref, _, _ = run(n, C.qft(n), merge=False, prep=prep)
    want_amps = O.run(n, prep + C.qft(n))
    d, dm = np.abs(got - want_amps).max(), np.abs(got - ref).max()
    print(f"QFT{n}: {launches} launches; diff vs oracle {d:.2e}, vs op-by-op {dm:.2e}")
    assert launches == want
    assert d <= TOL1 and dm <= TOL_SAME and abs(norm - 1) <= TOL1


@pytest.mark.parametrize("n", [33, 34])
def test_qft_basis_state_full_size_closed_form(n):
    """QFT of a basis state |k> at full size (64 / 128 GiB; two- and
    four-wave low passes): the reference's QFT (no swap layer) leaves qubit q
    in (|0> + e^(i phi_q)|1>)/sqrt(2) with phi_q = 2 pi k / 2^(q+1), so
    <X_q> = cos phi_q, <Y_q> = sin phi_q, <Z_q> = 0 (the closed form is
    pinned against the oracle by the next test), on every qubit."""
    k = 0x2F5A3C1B7 & ((1 << n) - 1)
    prep = [(C.X, (q,), ()) for q in range(n) if (k >> q) & 1]
    with qm.Simulator(n) as s:
        s.run(C.to_circuit(n, prep))
        s.run(C.to_circuit(n, C.qft(n)))
        norm = s.norm()
        worst = 0.0
        for q in range(n):
            phi = 2 * np.pi * ((k % (1 << (q + 1))) / (1 << (q + 1)))
            ex, ey, ez = (s.expectation_pauli(x=1 << q), s.expectation_pauli(y=1 << q),
                          s.expectation_pauli(z=1 << q))
            worst = max(worst, abs(ex - np.cos(phi)), abs(ey - np.sin(phi)), abs(ez))
    print(f"QFT{n}|k>: norm {norm:.8f}, worst Bloch-component error {worst:.2e}")
    assert abs(norm - 1) <= TOL1 and worst <= TOL1


def test_qft_basis_state_closed_form_matches_oracle():
    """The closed form used at full size, pinned against the complex128 oracle
    at n = 6 for three basis states (Bloch components of every qubit)."""
    n = 6
    idx = np.arange(1 << n)
    for k in (5, 37, 50):
        prep = [(C.X, (q,), ()) for q in range(n) if (k >> q) & 1]
        st = O.run(n, prep + C.qft(n))
        for q in range(n):
            a0 = st[((idx >> q) & 1) == 0]
            a1 = st[(idx ^ (1 << q))[((idx >> q) & 1) == 0]]
            ov = np.sum(np.conj(a0) * a1)
            phi = 2 * np.pi * k / (1 << (q + 1))
            assert abs(2 * ov.real - np.cos(phi)) < 1e-12 and abs(2 * ov.imag - np.sin(phi)) < 1e-12


def test_qft_zero_state_is_uniform_merged():
    n = 28
    with qm.Simulator(n) as s:
        s.run(C.to_circuit(n, C.qft(n)))
        zs = [abs(s.expectation_pauli(z=1 << q)) for q in (0, 1, 5, 13, 20, 27)]
        xs = [s.expectation_pauli(x=1 << q) for q in (0, 9, 27)]
        norm = s.norm()
    assert abs(norm - 1) <= TOL1
    assert max(zs) <= TOL1
    assert all(abs(x - 1) <= TOL1 for x in xs)


def adjoint_stream(n, count, seed):
    """Rotations mixed with every other differentiable-circuit gate kind,
    on wires drawn uniformly (controls land on local and base-index bits)."""
    r = C.SplitMix64(seed)

    def wires(k):
        out = []
        while len(out) < k:
            q = r.next_u64() % n
            if q not in out:
                out.append(q)
        return tuple(out)

    gates = [(C.H, (q,), ()) for q in range(n)]
    for _ in range(count):
        u = r.next_u64() % 10
        if u < 4:
            gates.append(((C.RX, C.RY, C.RZ)[r.next_u64() % 3], wires(1), (r.uniform(-3.2, 3.2),)))
        elif u < 5:
            gates.append((ONE[r.next_u64() % len(ONE)], wires(1), ()))
        elif u < 8:
            gates.append(((C.CX, C.CY, C.CZ)[r.next_u64() % 3], wires(2), ()))
        elif u < 9:
            gates.append((C.SWAP, wires(2), ()))
        else:
            gates.append((C.CCX, wires(3), ()))
    return gates


def test_merged_adjoint_against_oracle_and_per_gate_sweep():
    """The backward sweep in merged passes (several inverse gates and rotation
    overlaps per read of both states) against the complex128 adjoint and the
    one-launch-per-gate sweep."""
    for n in (6, 8, 11):
        terms = C.tfim_ring_hamiltonian(n) + [(0.3, 0, 1 << (n // 2), 0)]
        h = C.to_hamiltonian(terms)
        for seed in range(3):
            gates = adjoint_stream(n, 60, 77 * n + seed)
            c = C.to_circuit(n, gates)
            with qm.Simulator(n) as s:
                e_m, g_m = s.energy_and_gradient(c, h)
                s.set_pass_merging(False)
                e_p, g_p = s.energy_and_gradient(c, h)
            e_ref, g_ref = O.adjoint(n, gates, terms)
            d_ref = float(np.abs(np.array(g_m) - g_ref).max())
            d_per = float(np.abs(np.array(g_m) - np.array(g_p)).max())
            print(f"n={n} seed={seed}: |merged - oracle| {d_ref:.2e}, |merged - per-gate| {d_per:.2e}")
            assert abs(e_m - e_ref) <= 1e-5 and abs(e_m - e_p) <= 1e-5
            assert d_ref <= 1e-5 and d_per <= 1e-5
            for k, (kind, _, _) in enumerate(gates):
                if kind not in (C.RX, C.RY, C.RZ):
                    assert g_m[k] == 0.0


def tile_adjoint_stream(n, count, seed):
    """adjoint_stream's kinds plus CX(a,b) RZ_b CX(a,b) sandwiches (undone as
    one diagonal unit by the tile sweep), on neighbouring and random pairs,
    so their wires land on register, lane, wave, high-tile and base bits."""
    r = C.SplitMix64(seed)
    base = adjoint_stream(n, count, seed)
    out = []
    for g in base:
        out.append(g)
        if r.next_u64() % 3 == 0:
            a = r.next_u64() % n
            b = (a + 1) % n if r.next_u64() % 2 else (a + 1 + r.next_u64() % (n - 1)) % n
            out += C.zz(a, b, r.uniform(-3.2, 3.2))
    return out


@pytest.mark.parametrize("n", (12, 13, 16))
def test_tile_adjoint_against_oracle_and_per_gate_sweep(n):
    """Backward sweep in LDS tile passes (twelve wires per pass, segments of
    three register wires, diagonal units anywhere) against the complex128
    adjoint and the one-launch-per-gate sweep."""
    terms = C.tfim_ring_hamiltonian(n) + [(0.3, 0, 1 << (n // 2), 0), (0.2, 1 << (n - 1), 0, 1)]
    h = C.to_hamiltonian(terms)
    for seed in range(2):
        gates = tile_adjoint_stream(n, 90, 31 * n + seed)
        c = C.to_circuit(n, gates)
        with qm.Simulator(n) as s:
            e_m, g_m = s.energy_and_gradient(c, h)
            s.set_pass_merging(False)
            e_p, g_p = s.energy_and_gradient(c, h)
        e_ref, g_ref = O.adjoint(n, gates, terms)
        d_ref = float(np.abs(np.array(g_m) - g_ref).max())
        d_per = float(np.abs(np.array(g_m) - np.array(g_p)).max())
        print(f"n={n} seed={seed} gates={len(gates)}: |tile - oracle| {d_ref:.2e}, "
              f"|tile - per-gate| {d_per:.2e}")
        assert abs(e_m - e_ref) <= 1e-5 and abs(e_m - e_p) <= 1e-5
        assert d_ref <= 1e-5 and d_per <= 1e-5
        for k, (kind, _, _) in enumerate(gates):
            if kind not in (C.RX, C.RY, C.RZ):
                assert g_m[k] == 0.0


def random_hamiltonian(n, count, seed):
    """Random Pauli strings of weight 1..4 (X/Y/Z on random wires, so flips
    land on wire 0, lane, wave, register and base bits), 60 Z-only strings
    (more diagonal terms than one co-state tile takes), a flip over seven high
    wires (no tile holds it) and one over wires 0..6."""
    r = C.SplitMix64(seed)
    terms = []
    for _ in range(count):
        x = y = z = 0
        for _ in range(1 + r.next_u64() % 4):
            q = r.next_u64() % n
            k = r.next_u64() % 3
            x &= ~(1 << q)
            y &= ~(1 << q)
            z &= ~(1 << q)
            if k == 0:
                x |= 1 << q
            elif k == 1:
                y |= 1 << q
            else:
                z |= 1 << q
        terms.append((r.uniform(-1, 1), x, y, z))
    for _ in range(60):
        z = (1 << (r.next_u64() % n)) | (1 << (r.next_u64() % n))
        terms.append((r.uniform(-1, 1), 0, 0, z))
    terms.append((0.3, sum(1 << q for q in range(n - 7, n)), 0, 1))
    terms.append((-0.2, 0, 0b1111110, 1 << (n - 1)))
    return terms


@pytest.mark.parametrize("n", (12, 14, 17))
def test_costate_tiles_against_oracle(n):
    """lambda = H psi in LDS tile passes (flips served from LDS, diagonal
    terms by a per-thread Walsh-Hadamard transform, the energy fused into
    the last pass) for Hamiltonians of every flip shape, through the energy
    and the gradient, against the complex128 adjoint."""
    terms = random_hamiltonian(n, 90, 13 * n)
    gates = tile_adjoint_stream(n, 60, 5 * n)
    with qm.Simulator(n) as s:
        e, g = s.energy_and_gradient(C.to_circuit(n, gates), C.to_hamiltonian(terms))
    e_ref, g_ref = O.adjoint(n, gates, terms)
    d = float(np.abs(np.array(g) - g_ref).max())
    print(f"n={n} terms={len(terms)}: energy {e:.6f} (oracle {e_ref:.6f}), max |grad diff| {d:.2e}")
    assert abs(e - e_ref) <= 1e-5 * max(1.0, abs(e_ref))
    assert d <= 2e-5


@pytest.mark.parametrize("name,n", (("qaoa", 14), ("qaoa", 19), ("hea", 15), ("hea", 20)))
def test_tile_adjoint_structured_circuits(name, n):
    """QAOA ring (its ZZ sandwiches run as diagonal units) and HEA (CX ladder)
    through the tile sweep, against the complex128 adjoint."""
    r = C.SplitMix64(7 * n)
    if name == "qaoa":
        gates = C.qaoa_ring(n, 4, [r.uniform(0, 1) for _ in range(4)],
                            [r.uniform(0, 1) for _ in range(4)])
        terms = [(1.0, 0, 0, (1 << q) | (1 << ((q + 1) % n))) for q in range(n)]
    else:
        gates = C.hardware_efficient(n, 3, seed=0x51 + n)
        terms = C.tfim_ring_hamiltonian(n)
    with qm.Simulator(n) as s:
        e, g = s.energy_and_gradient(C.to_circuit(n, gates), C.to_hamiltonian(terms))
    e_ref, g_ref = O.adjoint(n, gates, terms)
    d = float(np.abs(np.array(g) - g_ref).max())
    print(f"{name} n={n}: energy {e:.6f} (oracle {e_ref:.6f}), max |grad diff| {d:.2e}")
    assert abs(e - e_ref) <= 1e-5 * max(1.0, abs(e_ref))
    assert d <= 1e-5


def step_gate(r, w, kind):
    """A chain step's 2x2: H ("h": the butterfly mode), RY ("real": a real
    matrix), U3 ("cplx")."""
    if kind == "h":
        return (C.H, (w,), ())
    if kind == "real":
        return (C.RY, (w,), (r.uniform(-3.2, 3.2),))
    return (C.U3, (w,), tuple(r.uniform(-3.2, 3.2) for _ in range(3)))


STEP_KINDS = ((1, "h"), (2, "cplx"), (3, "real"))


def wide_stream(n, rounds, seed, kind, steps=6):
    """Rounds of `steps` chain steps on distinct wires >= 7: a 2x2 (step_gate)
    followed by phase gates that all hold that wire (CP, CZ, P onto any other
    wire: pass wires, lane wires and workgroup-uniform wires); a CX separates
    the rounds."""
    r = C.SplitMix64(seed)
    gates = []
    for _ in range(rounds):
        ws = []
        while len(ws) < steps:
            q = 7 + r.next_u64() % (n - 7)
            if q not in ws:
                ws.append(q)
        for w in ws:
            gates.append(step_gate(r, w, kind))
            for j in range(1 + r.next_u64() % 4):
                k = r.next_u64() % n
                if k == w:
                    k = (w + 1 + j) % n
                # the first is a two-qubit phase, so windows never merge steps
                u = r.next_u64() % (2 if j == 0 else 3)
                if u == 0:
                    gates.append((C.CP, (k, w), (r.uniform(-3.2, 3.2),)))
                elif u == 1:
                    gates.append((C.CZ, (w, k), ()))
                else:
                    gates.append((C.P, (w,), (r.uniform(-3.2, 3.2),)))
        gates.append((C.CX, (0, n - 1), ()))
    return gates


@pytest.mark.parametrize("steps", [6, 7])
def test_wide_passes_against_oracle_and_op_by_op(steps):
    """Six- and seven-step chains over a workgroup tile (two or three steps
    after the LDS exchange), real and complex, with phase terms on every kind
    of bit."""
    for n in (16, 18):
        for seed, kind in STEP_KINDS:
            prep = C.product_state_prep(n, seed=seed)
            gates = wide_stream(n, 2, 100 * n + seed + 7 * steps, kind, steps)
            c = C.to_circuit(n, gates)
            plan = c.plan()
            wide = [e for e, sh in zip(c.schedule(), shapes(c)) if sh == "chain"
                    and sum(plan[k]["kind"] == "Dense1q" for k, _ in e) == steps]
            assert len(wide) == 2, "wide passes not formed"
            want = O.run(n, prep + gates)
            got, norm, launches = run(n, gates, prep=prep)
            ref, _, _ = run(n, gates, merge=False, prep=prep)
            d, dm = np.abs(got - want).max(), np.abs(got - ref).max()
            print(f"wide{steps} n={n} seed={seed} {kind}: {len(wide)} wide passes, {launches} launches, "
                  f"diff {d:.2e}, vs op-by-op {dm:.2e}")
            assert d <= TOL_DEEP and abs(norm - 1) <= TOL_DEEP
            assert dm <= TOL_SAME


def low_stream(n, rounds, seed, kind, top=11):
    """Rounds of 5..top chain steps on distinct wires 1..top, some closed by
    a step on wire 0, each step a 2x2 (step_gate) followed by
    phase gates that all hold its wire (onto register, lane, wave and
    workgroup-uniform wires); a CX separates the rounds. With top = 12 every
    other round holds a step on wire 12 (four-wave tiles). Rounds 1, 4, ...
    start with four wires drawn from 7..top in any order, rounds 2, 5, ...
    with the four top wires scrambled: the coalesced arrangement then puts
    those wires on its registers and serves the first four steps in place."""
    r = C.SplitMix64(seed)
    gates = []
    for i in range(rounds):
        count = 5 + r.next_u64() % (top - 4)
        ws = [12] if top == 12 and i % 2 == 1 else []
        if i % 3 == 1:
            ws = []
            while len(ws) < 4:
                q = 7 + r.next_u64() % (top - 6)
                if q not in ws:
                    ws.append(q)
        if i % 3 == 2:
            ws = [top, top - 2, top - 1, top - 3]
        while len(ws) < count:
            q = 1 + r.next_u64() % top
            if q not in ws:
                ws.append(q)
        if i % 2 == 0:
            ws.append(0)
        for w in ws:
            gates.append(step_gate(r, w, kind))
            for j in range(1 + r.next_u64() % 4):
                k = r.next_u64() % n
                if k == w:
                    k = (w + 1 + j) % n
                u = r.next_u64() % (2 if j == 0 else 3)
                if u == 0:
                    gates.append((C.CP, (k, w), (r.uniform(-3.2, 3.2),)))
                elif u == 1:
                    gates.append((C.CZ, (w, k), ()))
                else:
                    gates.append((C.P, (w,), (r.uniform(-3.2, 3.2),)))
        gates.append((C.CX, (n - 1, 0), ()))
    return gates


def test_low_passes_against_oracle_and_op_by_op():
    """Chains of 5..11 steps on wires 1..11 (and 0) over two-wave tiles, and
    of 5..12 steps on wires 1..12 over four-wave tiles (states of 13 qubits
    or more; from 24 qubits every low pass takes them), permuted through LDS,
    real and complex, with phase terms on every kind of bit."""
    O.set_threads(8)
    for n, top in ((12, 11), (14, 11), (17, 11), (13, 12), (14, 12), (17, 12), (24, 11),
                   (24, 12)):
        for seed, kind in STEP_KINDS:
            prep = C.product_state_prep(n, seed=seed)
            gates = low_stream(n, 4, 300 * n + 7 * top + seed, kind, top)
            c = C.to_circuit(n, gates)
            plan = c.plan()
            sched = c.schedule()
            low = [e for e, sh in zip(sched, shapes(c)) if sh == "chain"]
            assert len(low) >= 3, "low passes not formed"
            if top == 12:
                four = [e for e in low if any(plan[k]["kind"] == "Dense1q"
                                              and plan[k]["qubits"][0] == 12 for k, _ in e)]
                assert len(four) >= 2, "four-wave low passes not formed"
            want = O.run(n, prep + gates)
            got, norm, launches = run(n, gates, prep=prep)
            ref, _, _ = run(n, gates, merge=False, prep=prep)
            d, dm = np.abs(got - want).max(), np.abs(got - ref).max()
            print(f"low n={n} top={top} seed={seed} {kind}: {len(low)} low passes, "
                  f"{launches} launches, "
                  f"diff {d:.2e}, vs op-by-op {dm:.2e}")
            assert d <= TOL_DEEP and abs(norm - 1) <= TOL_DEEP
            assert dm <= TOL_SAME
    O.set_threads(1)


def tri_stream(n, rounds, seed, top, upward=False):
    """Rounds of H steps on wires top..1 (descending) then 0, each followed
    by random-angle CP / CZ / P gates whose other wire is lower than the
    step's or above the tile: every step's phase then ignores the wires of
    the earlier steps (the QFT's form, with arbitrary angles), which the low
    kernel runs with one complex product per amplitude and step. Round
    lengths vary so that the last arrangement is partly filled. `upward`:
    the third step of every round also takes a CP with the second step's
    wire (an earlier register wire of its arrangement), so the tables are no
    longer triangular and the low kernel must take its general form."""
    r = C.SplitMix64(seed)
    gates = []
    for i in range(rounds):
        count = top - (r.next_u64() % 4 if i % 2 else 0)
        for p, w in enumerate(list(range(top, top - count, -1)) + [0]):
            gates.append((C.H, (w,), ()))
            if upward and p == 2:
                gates.append((C.CP, (w + 1, w), (r.uniform(0.5, 3.0),)))
            # At least two two-qubit phase gates, so that the plan keeps
            # them as a phase layer after the step's Dense1q.
            for j in range(2 + r.next_u64() % 3):
                lower = w > 0 and r.next_u64() % 3 != 0
                k = r.next_u64() % w if lower else top + 1 + r.next_u64() % (n - top - 1)
                u = r.next_u64() % (2 if j < 2 else 3)
                if u == 0:
                    gates.append((C.CP, (k, w), (r.uniform(-3.2, 3.2),)))
                elif u == 1:
                    gates.append((C.CZ, (w, k), ()))
                else:
                    gates.append((C.P, (w,), (r.uniform(-3.2, 3.2),)))
        gates.append((C.CX, (n - 1, 0), ()))
    return gates


@pytest.mark.parametrize("n,top", [(14, 11), (14, 12), (17, 12), (24, 12)])
def test_triangular_low_chains_against_oracle_and_op_by_op(n, top):
    """Low passes whose phase tables are triangular (tri_stream) against the
    complex128 oracle and one launch per plan op, over two- and four-wave
    tiles."""
    prep = C.product_state_prep(n, seed=5 * n + top)
    gates = tri_stream(n, 3, 91 * n + top, top)
    c = C.to_circuit(n, gates)
    low = [sh for sh in shapes(c) if sh == "chain"]
    assert len(low) >= 3, "low passes not formed"
    want = O.run(n, prep + gates)
    got, norm, launches = run(n, gates, prep=prep)
    ref, _, _ = run(n, gates, merge=False, prep=prep)
    d, dm = np.abs(got - want).max(), np.abs(got - ref).max()
    print(f"triangular low n={n} top={top}: {len(low)} low passes, {launches} launches, "
          f"diff {d:.2e}, vs op-by-op {dm:.2e}")
    assert d <= TOL_DEEP and abs(norm - 1) <= TOL_DEEP
    assert dm <= TOL_SAME


@pytest.mark.parametrize("n,top", [(14, 11), (17, 12)])
def test_nearly_triangular_low_chains_take_the_general_form(n, top):
    """tri_stream with one upward coupling per round: the host's table check
    must reject the triangular form (a kernel that skipped the coupling
    would miss a phase of 0.5..3 rad on a quarter of the state)."""
    prep = C.product_state_prep(n, seed=7 * n + top)
    gates = tri_stream(n, 3, 53 * n + top, top, upward=True)
    c = C.to_circuit(n, gates)
    low = [sh for sh in shapes(c) if sh == "chain"]
    assert len(low) >= 3, "low passes not formed"
    want = O.run(n, prep + gates)
    got, norm, _ = run(n, gates, prep=prep)
    ref, _, _ = run(n, gates, merge=False, prep=prep)
    d, dm = np.abs(got - want).max(), np.abs(got - ref).max()
    print(f"nearly triangular low n={n} top={top}: {len(low)} low passes, "
          f"diff {d:.2e}, vs op-by-op {dm:.2e}")
    assert d <= TOL_DEEP and abs(norm - 1) <= TOL_DEEP
    assert dm <= TOL_SAME


def tile_stream(n, count, seed):
    """Every mergeable gate kind with targets on wires 0..11 and controls,
    swap partners and phase wires anywhere, in runs long enough to need
    several register arrangements (generic low tile passes)."""
    r = C.SplitMix64(seed)
    gates = [(C.H, (q,), ()) for q in range(n)]

    def low():
        return r.next_u64() % 12

    def other(avoid):
        while True:
            q = r.next_u64() % n
            if q not in avoid:
                return q

    for _ in range(count):
        u = r.next_u64() % 12
        t = low()
        if u < 3:
            gates.append((C.U3, (t,), tuple(r.uniform(-3.2, 3.2) for _ in range(3))))
        elif u < 4:
            gates.append((ONE[r.next_u64() % len(ONE)], (t,), ()))
        elif u < 6:
            gates.append(((C.CX, C.CY)[r.next_u64() % 2], (other((t,)), t), ()))
        elif u < 8:
            gates.append((C.CP, (other((t,)), t), (r.uniform(-3.2, 3.2),)))
        elif u < 9:
            a = low()
            b = low() if a else 1 + low() % 11
            if a != b:
                gates.append((C.SWAP, (a, b), ()))
        elif u < 10:
            c0 = other((t,))
            gates.append((C.CCX, (c0, other((t, c0)), t), ()))
        else:
            a, b = t, other((t,))
            gates += C.zz(a, b, r.uniform(-3.2, 3.2))
    return gates


def test_tile_passes_against_oracle_and_op_by_op():
    """Generic low tile passes (merged-pass ops in several register
    arrangements of a 2^12 tile, permuted through LDS) with controls and
    phase terms on register, thread and workgroup-uniform bits."""
    for n in (12, 13, 15, 17):
        for seed in range(3):
            gates = tile_stream(n, 60, 500 * n + seed)
            c = C.to_circuit(n, gates)
            plan = c.plan()
            tiles = [sh for sh in shapes(c) if sh == "tile"]
            assert tiles, "no tile pass formed"
            want = O.run(n, gates)
            got, norm, launches = run(n, gates)
            ref, _, _ = run(n, gates, merge=False)
            d, dm = np.abs(got - want).max(), np.abs(got - ref).max()
            print(f"tile n={n} seed={seed}: {len(tiles)} tile passes, {launches} launches "
                  f"(plan {len(plan)}), diff {d:.2e}, vs op-by-op {dm:.2e}")
            assert d <= TOL_DEEP and abs(norm - 1) <= TOL_DEEP
            assert dm <= TOL_SAME


def gwide_stream(n, rounds, seed):
    """Rounds of mergeable gates on seven wires >= 7: a run with targets on
    four of them, then a run with targets on the other three and one of the
    four, with controls, swap partners and phase wires anywhere; an H on wire
    3 separates the rounds."""
    r = C.SplitMix64(seed)
    gates = [(C.H, (q,), ()) for q in range(n)]

    def pick(ws):
        return ws[r.next_u64() % len(ws)]

    def other(avoid):
        while True:
            q = r.next_u64() % n
            if q not in avoid:
                return q

    def run(ws, count):
        for _ in range(count):
            u = r.next_u64() % 8
            t = pick(ws)
            if u < 2:
                gates.append((C.U3, (t,), tuple(r.uniform(-3.2, 3.2) for _ in range(3))))
            elif u < 3:
                gates.append((C.RY, (t,), (r.uniform(-3.2, 3.2),)))
            elif u < 5:
                gates.append((C.CX, (other((t,)), t), ()))
            elif u < 6:
                gates.append((C.CP, (other((t,)), t), (r.uniform(-3.2, 3.2),)))
            elif u < 7:
                c0 = other((t,))
                gates.append((C.CCX, (c0, other((t, c0)), t), ()))
            else:
                b = pick(ws)
                if b != t:
                    gates.append((C.SWAP, (t, b), ()))

    for _ in range(rounds):
        ws = []
        while len(ws) < 7:
            q = 7 + r.next_u64() % (n - 7)
            if q not in ws:
                ws.append(q)
        run(ws[:4], 10)
        run(ws[4:] + [ws[3]], 10)
        gates.append((C.H, (3,), ()))
    return gates


def test_gwide_passes_against_oracle_and_op_by_op():
    """Generic wide passes (merged-pass ops on a wide chain tile before and
    after its LDS exchange) with controls and phase terms on register, lane,
    wave and workgroup-uniform bits."""
    for n in (16, 18):
        for seed in range(3):
            gates = gwide_stream(n, 3, 700 * n + seed)
            c = C.to_circuit(n, gates)
            plan = c.plan()
            wide = [sh for sh in shapes(c) if sh == "gwide"]
            assert len(wide) >= 2, "no generic wide pass"
            want = O.run(n, gates)
            got, norm, launches = run(n, gates)
            ref, _, _ = run(n, gates, merge=False)
            d, dm = np.abs(got - want).max(), np.abs(got - ref).max()
            print(f"gwide n={n} seed={seed}: {len(wide)} generic wide passes, {launches} launches "
                  f"(plan {len(plan)}), diff {d:.2e}, vs op-by-op {dm:.2e}")
            assert d <= TOL_DEEP and abs(norm - 1) <= TOL_DEEP
            assert dm <= TOL_SAME
