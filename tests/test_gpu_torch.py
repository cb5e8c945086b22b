"""torch.autograd.Function over the HIP adjoint: gradients against exact
parameter shift on the FP64 oracle and against central differences (h=1e-3)
of the GPU energy itself, on the hardware-efficient ansatz with the TFIM ring
Hamiltonian (sum Z_i Z_{i+1} + 0.7 sum X_i)."""
import numpy as np
import pytest

import qmetal_amd as qm
from qmetal_amd import circuits as C
from oracle import oracle as O

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def hea_template(n, depth, params):
    return C.hardware_efficient(n, depth, params=params)


def test_torch_adjoint_matches_parameter_shift_n8():
    from qmetal_amd.torch_adjoint import AdjointEnergy
    assert torch.cuda.is_available()
    n, depth = 8, 3
    P = 2 * n * depth
    rng = C.SplitMix64(11)
    p0 = [rng.uniform(0, 2 * np.pi) for _ in range(P)]
    gates = hea_template(n, depth, p0)
    terms = C.tfim_ring_hamiltonian(n)
    sim = qm.Simulator(n)
    theta = torch.tensor(p0, dtype=torch.float64, device="cuda", requires_grad=True)
    e = AdjointEnergy.apply(theta, sim, C.to_circuit(n, gates), C.to_hamiltonian(terms))
    e.backward()
    got = theta.grad.cpu().numpy()
    want_full = O.param_shift(n, gates, terms)
    want = np.array([want_full[k] for k, (kind, _, _) in enumerate(gates) if kind in (8, 9, 10)])
    assert abs(e.item() - O.expect_h(O.run(n, gates), terms)) <= 1e-5
    worst = np.abs(got - want).max()
    print(f"torch adjoint vs parameter shift: {worst:.3e}")
    assert worst <= 1e-5


def test_torch_adjoint_central_differences_n14():
    from qmetal_amd.torch_adjoint import AdjointEnergy, energy_and_grad
    n, depth = 14, 3
    P = 2 * n * depth
    rng = C.SplitMix64(12)
    p0 = [rng.uniform(0, 2 * np.pi) for _ in range(P)]
    circ = C.to_circuit(n, hea_template(n, depth, p0))
    ham = C.to_hamiltonian(C.tfim_ring_hamiltonian(n))
    sim = qm.Simulator(n)
    theta = torch.tensor(p0, dtype=torch.float64, device="cuda", requires_grad=True)
    e = AdjointEnergy.apply(theta, sim, circ, ham)
    (g,) = torch.autograd.grad(e, theta)
    h = 1e-3
    for j in range(0, P, 7):
        tp = theta.detach().clone()
        tm = theta.detach().clone()
        tp[j] += h
        tm[j] -= h
        ep, _ = energy_and_grad(sim, circ, ham, tp)
        em, _ = energy_and_grad(sim, circ, ham, tm)
        fd = (ep.item() - em.item()) / (2 * h)
        # complex64 energies: differences of ~1e-7-accurate values over 2h
        assert abs(fd - g[j].item()) <= 5e-4, (j, fd, g[j].item())


def test_torch_parameter_count_checked():
    from qmetal_amd.torch_adjoint import AdjointEnergy
    n = 4
    circ = C.to_circuit(n, hea_template(n, 1, [0.1] * 8))
    sim = qm.Simulator(n)
    bad = torch.zeros(5, dtype=torch.float64, device="cuda", requires_grad=True)
    with pytest.raises(qm.QMetalError, match="parameters"):
        AdjointEnergy.apply(bad, sim, circ, C.to_hamiltonian(C.ring_zz_hamiltonian(n)))


def test_energy_only_when_no_gradient_is_needed():
    """Without a gradient to compute (params not requiring one, or gradients
    disabled) the call runs the forward passes and the energy reductions only:
    the same energy as the full adjoint call, and measurably less work."""
    import time
    from qmetal_amd.torch_adjoint import AdjointEnergy, adjoint_energy, energy_and_grad
    n = 18
    rng = C.SplitMix64(5)
    p0 = [rng.uniform(0, 2 * np.pi) for _ in range(6 * n)]
    circ = C.to_circuit(n, hea_template(n, 3, p0))
    ham = C.to_hamiltonian(C.tfim_ring_hamiltonian(n))
    with qm.Simulator(n) as sim:
        theta = torch.tensor(p0, dtype=torch.float64, device="cuda", requires_grad=True)
        e_full, g_full = energy_and_grad(sim, circ, ham, theta)
        e_only, g_none = energy_and_grad(sim, circ, ham, theta, need_grad=False)
        assert g_none is None
        assert abs(e_full.item() - e_only.item()) <= 1e-6
        with torch.no_grad():
            e_ng = adjoint_energy(theta, sim, circ, ham)
        e_plain = AdjointEnergy.apply(theta.detach(), sim, circ, ham)
        assert abs(e_ng.item() - e_full.item()) <= 1e-6
        assert abs(e_plain.item() - e_full.item()) <= 1e-6
        e_g = adjoint_energy(theta, sim, circ, ham)
        (g,) = torch.autograd.grad(e_g, theta)
        assert torch.allclose(g, g_full, atol=1e-9)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(5):
            energy_and_grad(sim, circ, ham, theta)
        torch.cuda.synchronize()
        t_full = time.perf_counter() - t0
        t0 = time.perf_counter()
        for _ in range(5):
            energy_and_grad(sim, circ, ham, theta, need_grad=False)
        torch.cuda.synchronize()
        t_energy = time.perf_counter() - t0
    print(f"energy and gradient {t_full / 5 * 1e3:.2f} ms, energy only {t_energy / 5 * 1e3:.2f} ms")
    assert t_energy < t_full


def test_energy_only_program_under_graph_capture():
    """With no gradient asked, the device program measures the energy in
    read-only tile passes (no co-state is allocated): captured in a graph,
    replays with new angles equal eager calls bit for bit and the full
    adjoint call's energy to 1e-6."""
    from qmetal_amd.torch_adjoint import energy_and_grad
    n, depth = 15, 2
    P = 2 * n * depth
    rng = C.SplitMix64(33)
    circ = C.to_circuit(n, hea_template(n, depth, [0.1] * P))
    ham = C.to_hamiltonian(C.tfim_ring_hamiltonian(n))
    with qm.Simulator(n) as sim:
        theta = torch.tensor([rng.uniform(0, 6.2) for _ in range(P)], dtype=torch.float64,
                             device="cuda")
        side = torch.cuda.Stream()
        side.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(side):
            for _ in range(2):
                energy_and_grad(sim, circ, ham, theta, need_grad=False)
        torch.cuda.current_stream().wait_stream(side)
        torch.cuda.synchronize()
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(graph):
            e_g, _ = energy_and_grad(sim, circ, ham, theta, need_grad=False)
        for trial in range(3):
            theta.copy_(torch.tensor([rng.uniform(0, 6.2) for _ in range(P)],
                                     dtype=torch.float64, device="cuda"))
            graph.replay()
            torch.cuda.synchronize()
            e_r = e_g.clone()
            e_e, _ = energy_and_grad(sim, circ, ham, theta, need_grad=False)
            torch.cuda.synchronize()
            e_full, _ = energy_and_grad(sim, circ, ham, theta)
            torch.cuda.synchronize()
            print(f"replay {trial}: energy {e_r.item():.8f} eager {e_e.item():.8f} "
                  f"with gradient {e_full.item():.8f}")
            assert e_r.item() == e_e.item()
            assert abs(e_r.item() - e_full.item()) <= 1e-6


def test_energy_only_with_a_term_no_tile_holds():
    """A Hamiltonian with a string flipping every wire (too wide for a
    tile): the energy-only device call builds the co-state instead, and
    agrees with the full call and the host path."""
    from qmetal_amd.torch_adjoint import energy_and_grad
    n = 14
    rng = C.SplitMix64(44)
    p0 = [rng.uniform(0, 2 * np.pi) for _ in range(4 * n)]
    gates = hea_template(n, 2, p0)
    circ = C.to_circuit(n, gates)
    terms = C.tfim_ring_hamiltonian(n) + [(0.8, (1 << n) - 1, 0, 0)]
    ham = C.to_hamiltonian(terms)
    with qm.Simulator(n) as sim:
        theta = torch.tensor(p0, dtype=torch.float64, device="cuda")
        e_only, g_none = energy_and_grad(sim, circ, ham, theta, need_grad=False)
        e_full, _ = energy_and_grad(sim, circ, ham, theta)
        e_host, _ = sim.energy_and_gradient(circ, ham)
    assert g_none is None
    assert abs(e_only.item() - e_full.item()) <= 1e-6
    assert abs(e_only.item() - e_host) <= 1e-6


def test_host_path_after_device_calls_without_synchronising():
    """Device-resident calls return before their kernels (on the caller's
    stream) finish; host-path calls that follow at once, on the simulator's
    own stream, wait for them: gradient, expectation and norm right after
    device calls equal the complex128 oracle."""
    from qmetal_amd.torch_adjoint import energy_and_grad
    n = 14
    rng = C.SplitMix64(45)
    p0 = [rng.uniform(0, 2 * np.pi) for _ in range(4 * n)]
    gates = hea_template(n, 2, p0)
    circ = C.to_circuit(n, gates)
    terms = C.tfim_ring_hamiltonian(n) + [(0.8, (1 << n) - 1, 0, 0)]
    ham = C.to_hamiltonian(terms)
    want = O.expect_h(O.run(n, gates), terms)
    with qm.Simulator(n) as sim:
        theta = torch.tensor(p0, dtype=torch.float64, device="cuda")
        for _ in range(3):
            e_only, _ = energy_and_grad(sim, circ, ham, theta, need_grad=False)
            e_full, _ = energy_and_grad(sim, circ, ham, theta)
            e_host, _ = sim.energy_and_gradient(circ, ham)
            e_dev2, _ = energy_and_grad(sim, circ, ham, theta)
            sim.reset()
            sim.run(circ)
            e_exp = sim.expectation(ham)
            nrm = sim.norm()
            for v in (e_only.item(), e_full.item(), e_host, e_dev2.item(), e_exp):
                assert abs(v - want) <= 1e-5, (v, want)
            assert abs(nrm - 1) <= 1e-5


def test_device_program_recompiles_when_template_changes():
    """One simulator alternating between two circuit templates and two
    Hamiltonians (each change compiles a new program, the old one retired
    after its kernels), energy only and with gradients, without
    synchronising: every call equals the host path on the same angles."""
    from qmetal_amd.torch_adjoint import energy_and_grad
    n = 13
    rng = C.SplitMix64(46)
    pa = [rng.uniform(0, 2 * np.pi) for _ in range(4 * n)]
    ga = hea_template(n, 2, pa)
    pb = [rng.uniform(0, np.pi) for _ in range(4)]
    gb = C.qaoa_ring(n, 2, gammas=pb[:2], betas=pb[2:])
    qb = [p[0] for k, _, p in gb if k in (C.RX, C.RY, C.RZ)]
    hs = [C.to_hamiltonian(C.tfim_ring_hamiltonian(n)),
          C.to_hamiltonian(C.ring_zz_hamiltonian(n) + [(0.5, (1 << n) - 1, 0, 0)])]
    with qm.Simulator(n) as sim:
        for k in range(6):
            gates, params = (ga, pa) if k % 2 == 0 else (gb, qb)
            circ = C.to_circuit(n, gates)
            ham = hs[(k // 2) % 2]
            theta = torch.tensor(params, dtype=torch.float64, device="cuda")
            need = k % 3 != 1
            e_dev, g_dev = energy_and_grad(sim, circ, ham, theta, need_grad=need)
            e_host, g_host = sim.energy_and_gradient(circ, ham)
            rot = [j for j, (kind, _, _) in enumerate(gates) if kind in (C.RX, C.RY, C.RZ)]
            assert abs(e_dev.item() - e_host) <= 1e-5, (k, e_dev.item(), e_host)
            if need:
                assert np.abs(g_dev.cpu().numpy() - np.array(g_host)[rot]).max() <= 1e-5, k


def _with_angles(gates, theta):
    """The template with its RX/RY/RZ angles replaced, in circuit order."""
    it = iter(theta)
    return [(k, q, (next(it),)) if k in (C.RX, C.RY, C.RZ) else (k, q, p) for k, q, p in gates]


@pytest.mark.parametrize("name,n", (("hea", 13), ("qaoa", 15), ("mixed", 16)))
def test_device_program_matches_host_path_and_oracle(name, n):
    """From 12 qubits the device entry runs a program compiled once per
    circuit template (tile passes; rotation matrices written on the device
    from the parameter tensor): against the host path on the same angles and
    the complex128 adjoint, twice with different angles (the second call
    reuses the program)."""
    from qmetal_amd.torch_adjoint import energy_and_grad
    import test_gpu_passes as TP
    r = C.SplitMix64(40 + n)
    if name == "hea":
        gates = C.hardware_efficient(n, 3, seed=9)
        terms = C.tfim_ring_hamiltonian(n)
    elif name == "qaoa":
        gates = C.qaoa_ring(n, 3, [0.2] * 3, [0.3] * 3)
        terms = [(1.0, 0, 0, (1 << q) | (1 << ((q + 1) % n))) for q in range(n)]
    else:
        gates = TP.tile_adjoint_stream(n, 60, 77 + n)
        terms = TP.random_hamiltonian(n, 30, 5 + n)
    nrot = sum(1 for k, _, _ in gates if k in (C.RX, C.RY, C.RZ))
    circ = C.to_circuit(n, gates)
    ham = C.to_hamiltonian(terms)
    with qm.Simulator(n) as sim:
        for trial in range(2):
            th = [r.uniform(-3.2, 3.2) for _ in range(nrot)]
            theta = torch.tensor(th, dtype=torch.float64, device="cuda")
            e_d, g_d = energy_and_grad(sim, circ, ham, theta)
            torch.cuda.synchronize()
            concrete = _with_angles(gates, th)
            e_h, g_h = sim.energy_and_gradient(C.to_circuit(n, concrete), ham)
            where = [k for k, (kind, _, _) in enumerate(gates) if kind in (C.RX, C.RY, C.RZ)]
            g_h = np.array([g_h[k] for k in where])
            e_o, g_o = O.adjoint(n, concrete, terms)
            g_o = np.array([g_o[k] for k in where])
            gd = g_d.cpu().numpy()
            print(f"{name} n={n} trial {trial}: |dev - host| {np.abs(gd - g_h).max():.2e}, "
                  f"|dev - oracle| {np.abs(gd - g_o).max():.2e}")
            assert abs(e_d.item() - e_o) <= 1e-5 * max(1.0, abs(e_o))
            assert abs(e_d.item() - e_h) <= 1e-5 * max(1.0, abs(e_o))
            assert np.abs(gd - g_o).max() <= 2e-5
            assert np.abs(gd - g_h).max() <= 2e-5


def test_device_program_under_graph_capture():
    """The device entry launches only kernels on the caller's stream after its
    first call, so it can be captured in a graph: replays with new angles in
    the captured parameter tensor equal eager calls bit for bit (eager work and
    synchronisation between the replays)."""
    from qmetal_amd.torch_adjoint import energy_and_grad
    n, depth = 14, 2
    P = 2 * n * depth
    rng = C.SplitMix64(21)
    circ = C.to_circuit(n, hea_template(n, depth, [0.1] * P))
    ham = C.to_hamiltonian(C.tfim_ring_hamiltonian(n))
    with qm.Simulator(n) as sim:
        theta = torch.tensor([rng.uniform(0, 6.2) for _ in range(P)], dtype=torch.float64,
                             device="cuda")
        side = torch.cuda.Stream()
        side.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(side):
            for _ in range(2):
                energy_and_grad(sim, circ, ham, theta)
        torch.cuda.current_stream().wait_stream(side)
        torch.cuda.synchronize()
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(graph):
            e_g, g_g = energy_and_grad(sim, circ, ham, theta)
        for trial in range(4):
            theta.copy_(torch.tensor([rng.uniform(0, 6.2) for _ in range(P)],
                                     dtype=torch.float64, device="cuda"))
            graph.replay()
            torch.cuda.synchronize()
            e_r, g_r = e_g.clone(), g_g.clone()
            e_e, g_e = energy_and_grad(sim, circ, ham, theta)
            torch.cuda.synchronize()
            print(f"replay {trial}: energy {e_r.item():.8f} eager {e_e.item():.8f}")
            assert e_r.item() == e_e.item()
            assert torch.equal(g_r, g_e)
