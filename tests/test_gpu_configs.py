"""The benchmark configurations at their full sizes, on the GPU.

- QFT on the seeded product state at n = 30, 32, 33 (the headline workload)
  and 34: amplitudes at 4096 indices against the closed form
  N^{-1/2} prod_q (a_q + b_q exp(2 pi i 2^q rev(y) / N)) in FP64, then the
  inverse QFT and every qubit's <X>, <Y>, <Z> against the product state's
  closed form (forms pinned on the oracle in tests/test_oracle.py).
- Brickwork depth 10 (Haar U3 layers and CX bricks): n = 26 against the
  complex128 oracle over every amplitude; n = 32 merged passes against one
  launch per plan op (amplitudes at 65536 indices, every qubit's Bloch
  vector, norm) and the merged circuit undone op by op back to |0...0>.
- The hardware-efficient ansatz, depth 3, with the TFIM ring Hamiltonian
  through torch.autograd (tests/test_gpu.mm:389-480 on the structure of
  src/circuits.cpp:117-131): n = 20 against FP64 parameter shift on a
  sample of parameters and the complex128 adjoint on all; n = 26 and 32
  against central differences (h = 1e-3) of the GPU energy.

Tolerances: amplitude checks relative to the mean amplitude 2^{-n/2}
(1e-3 of it: a wrong phase on any qubit is an O(1) error); Bloch
components and norms 1e-5 (complex64 state, FP64 reductions); gradients 1e-5
against FP64 references, 2e-3 against difference quotients of complex64
energies.
"""
import numpy as np
import pytest

import closed_forms as CF
import qmetal_amd as qm
from qmetal_amd import circuits as C
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def random_indices(n, count, seed):
    rng = np.random.default_rng(seed)
    idx = rng.integers(0, 1 << n, size=count, dtype=np.uint64)
    idx[:2] = (0, (1 << n) - 1)
    return idx


def bloch_all(s, n):
    xs = np.array([s.expectation_pauli(x=1 << q) for q in range(n)])
    ys = np.array([s.expectation_pauli(y=1 << q) for q in range(n)])
    zs = np.array([s.expectation_pauli(z=1 << q) for q in range(n)])
    return xs, ys, zs


@pytest.mark.parametrize("n", [30, 32, 33, 34])
def test_qft_product_state_full_size(n):
    prep = C.product_state_prep(n, seed=2024)
    a, b = CF.product_factors(prep)
    idx = random_indices(n, 4096, seed=n)
    want = CF.qft_product_amplitudes(a, b, idx)
    scale = 2.0 ** (-n / 2)
    with qm.Simulator(n) as s:
        s.run(C.to_circuit(n, prep))
        qc = C.to_circuit(n, C.qft(n))
        launches = s.launch_count
        s.run(qc)
        launches = s.launch_count - launches
        got = s.amplitudes_at(idx).astype(np.complex128)
        norm_f = s.norm()
        s.run(C.to_circuit(n, CF.inverse_stream(C.qft(n))))
        xs, ys, zs = bloch_all(s, n)
        norm_b = s.norm()
    d_amp = float(np.abs(got - want).max())
    wx, wy, wz = CF.bloch(a, b)
    d_bloch = float(max(np.abs(xs - wx).max(), np.abs(ys - wy).max(), np.abs(zs - wz).max()))
    print(f"QFT{n} product state ({launches} launches): max |amp - closed form| "
          f"{d_amp:.2e} ({d_amp / scale:.2e} of 2^-n/2); after the inverse QFT worst Bloch "
          f"error {d_bloch:.2e} over {3 * n} components; norms {norm_f:.8f} {norm_b:.8f}")
    assert launches <= 4
    assert d_amp <= 1e-3 * scale
    assert d_bloch <= 1e-5
    assert abs(norm_f - 1) <= 1e-5 and abs(norm_b - 1) <= 1e-5


def test_brickwork_d10_n26_against_oracle():
    n = 26
    gates = C.brickwork(n, 10)
    assert len(gates) == 385
    c = C.to_circuit(n, gates)
    assert c.plan_size(True) == 195
    O.set_threads(16)
    try:
        want = O.run(n, gates)
    finally:
        O.set_threads(1)
    with qm.Simulator(n) as s:
        l0 = s.launch_count
        s.run(c)
        launches = s.launch_count - l0
        got = s.amplitudes().astype(np.complex128)
        norm = s.norm()
    d = float(np.abs(got - want).max())
    print(f"brickwork d10 n=26: {launches} launches for 195 plan ops; max diff vs oracle "
          f"{d:.2e}; norm {norm:.8f}")
    assert d <= 1e-6 + 1e-8 * len(gates)
    assert abs(norm - 1) <= 1e-5


def test_brickwork_d10_n32_merged_vs_op_by_op():
    n = 32
    gates = C.brickwork(n, 10)
    assert len(gates) == 475
    c = C.to_circuit(n, gates)
    assert c.plan_size(True) == 235
    idx = random_indices(n, 65536, seed=32)
    scale = 2.0 ** (-n / 2)
    with qm.Simulator(n) as m, qm.Simulator(n) as r:
        r.set_pass_merging(False)
        l0 = m.launch_count
        m.run(c)
        launches = m.launch_count - l0
        r.run(c)
        am, ar = m.amplitudes_at(idx), r.amplitudes_at(idx)
        bm, br = bloch_all(m, n), bloch_all(r, n)
        nm, nr = m.norm(), r.norm()
        # Undo the merged circuit one launch per plan op: back to |0...0>.
        m.set_pass_merging(False)
        m.run(C.to_circuit(n, CF.inverse_stream(gates)))
        z_back = np.array([m.expectation_pauli(z=1 << q) for q in range(n)])
        a0 = m.amplitudes_at(np.array([0], dtype=np.uint64))[0]
    d_amp = float(np.abs(am.astype(np.complex128) - ar).max())
    d_bloch = float(max(np.abs(x - y).max() for x, y in zip(bm, br)))
    print(f"brickwork d10 n=32: {launches} launches for 235 plan ops; merged vs op-by-op: "
          f"amplitudes {d_amp:.2e} ({d_amp / scale:.2e} of 2^-n/2), Bloch {d_bloch:.2e}, "
          f"norms {nm:.8f} {nr:.8f}; round trip min <Z_q> {z_back.min():.8f}, "
          f"|<0|psi>| {abs(a0):.8f}")
    assert d_amp <= 1e-3 * scale and d_bloch <= 1e-5
    assert abs(nm - 1) <= 1e-5 and abs(nr - 1) <= 1e-5
    assert z_back.min() >= 1 - 1e-5 and abs(abs(a0) - 1) <= 1e-5


def hea(n, seed):
    rng = C.SplitMix64(seed)
    p0 = [rng.uniform(0, 2 * np.pi) for _ in range(2 * n * 3)]
    return p0, C.hardware_efficient(n, 3, params=p0)


def test_torch_hea_d3_n20_against_parameter_shift():
    torch = pytest.importorskip("torch")
    from qmetal_amd.torch_adjoint import AdjointEnergy
    n = 20
    p0, gates = hea(n, 20)
    assert len(gates) == 177
    terms = C.tfim_ring_hamiltonian(n)
    with qm.Simulator(n) as sim:
        theta = torch.tensor(p0, dtype=torch.float64, device="cuda", requires_grad=True)
        e = AdjointEnergy.apply(theta, sim, C.to_circuit(n, gates), C.to_hamiltonian(terms))
        e.backward()
        got = theta.grad.cpu().numpy()
    rot = [k for k, (kind, _, _) in enumerate(gates) if kind in (C.RX, C.RY, C.RZ)]
    O.set_threads(16)
    try:
        e_ref, g_ref = O.adjoint(n, gates, terms)
        # Exact parameter shift on every 6th parameter (both kinds, every layer).
        sample = list(range(0, len(rot), 6)) + [len(rot) - 1]
        shift = {}
        for j in sample:
            k = rot[j]
            vals = []
            for dlt in (np.pi / 2, -np.pi / 2):
                g2 = list(gates)
                kind, q, p = g2[k]
                g2[k] = (kind, q, (p[0] + dlt,))
                vals.append(O.expect_h(O.run(n, g2), terms))
            shift[j] = 0.5 * (vals[0] - vals[1])
    finally:
        O.set_threads(1)
    d_adj = float(np.abs(got - g_ref[rot]).max())
    d_shift = max(abs(got[j] - v) for j, v in shift.items())
    d_ref = max(abs(g_ref[rot[j]] - v) for j, v in shift.items())
    print(f"torch HEA d3 n=20: |E - E_ref| {abs(e.item() - e_ref):.2e}; grad vs complex128 "
          f"adjoint {d_adj:.2e} (120 params), vs parameter shift {d_shift:.2e} "
          f"({len(shift)} params; oracle adjoint vs shift {d_ref:.1e})")
    assert abs(e.item() - e_ref) <= 1e-5 * max(1.0, abs(e_ref))
    assert d_adj <= 1e-5 and d_shift <= 1e-5 and d_ref <= 1e-10


@pytest.mark.parametrize("n", [26, 32])
def test_torch_hea_d3_central_differences(n):
    torch = pytest.importorskip("torch")
    from qmetal_amd.torch_adjoint import AdjointEnergy, energy_and_grad
    p0, gates = hea(n, n)
    circ = C.to_circuit(n, gates)
    ham = C.to_hamiltonian(C.tfim_ring_hamiltonian(n))
    h = 1e-3
    P = len(p0)
    worst = 0.0
    with qm.Simulator(n) as sim:
        theta = torch.tensor(p0, dtype=torch.float64, device="cuda", requires_grad=True)
        e = AdjointEnergy.apply(theta, sim, circ, ham)
        (g,) = torch.autograd.grad(e, theta)
        for j in (0, 1, P // 3, P // 2 + 1, P - 2, P - 1):
            tp, tm = theta.detach().clone(), theta.detach().clone()
            tp[j] += h
            tm[j] -= h
            ep, _ = energy_and_grad(sim, circ, ham, tp)
            em, _ = energy_and_grad(sim, circ, ham, tm)
            fd = (ep.item() - em.item()) / (2 * h)
            worst = max(worst, abs(fd - g[j].item()))
            assert abs(fd - g[j].item()) <= 2e-3, (j, fd, g[j].item())
    print(f"torch HEA d3 n={n}: E {e.item():.6f}; worst |central difference - adjoint| "
          f"{worst:.2e} over 6 parameters")


def test_tfim_expectation_n32_tile_passes_vs_per_term():
    """<H> of the TFIM ring (64 terms) on the HEA d3 n=32 state: the tile
    passes (every term in one read-only pass here) against one FP64
    reduction per term on the same state, and against the energy the
    adjoint gradient fuses into its co-state pass."""
    import time
    n = 32
    p0, gates = hea(n, n)
    terms = C.tfim_ring_hamiltonian(n)
    ham = C.to_hamiltonian(terms)
    scale = sum(abs(t[0]) for t in terms)
    with qm.Simulator(n) as s:
        s.run(C.to_circuit(n, gates))
        s.expectation(ham)
        t0 = time.perf_counter()
        got = s.expectation(ham)
        t1 = time.perf_counter()
        s.set_pass_merging(False)
        per_term = s.expectation(ham)
        t2 = time.perf_counter()
        s.set_pass_merging(True)
        energy, _ = s.energy_and_gradient(C.to_circuit(n, gates), ham)
    assert abs(got - per_term) <= 1e-6 * scale, (got, per_term)
    assert abs(got - energy) <= 1e-6 * scale, (got, energy)
    print(f"TFIM ring n=32 <H> = {got:.9f}: tile passes {1e3 * (t1 - t0):.1f} ms, "
          f"per term {1e3 * (t2 - t1):.1f} ms (|diff| {abs(got - per_term):.2e})")
