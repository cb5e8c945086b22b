"""Plan builder parity with the reference builder (src/plan.cpp), on the CPU.

Every golden fixture holds a gate stream and the plans the reference builder
produced from it under six FusionOptions; the plan built here through the C
ABI (qm_plan_build) must be identical: op kinds, wires, FP64 matrices, term
masks, angles, global phases and block bits, compared with float equality.
Plans are also executed on the complex128 oracle and compared with the
circuit run gate by gate (the reference suite tests/test_fusion.cpp:38-44).
"""
import math

import numpy as np
import pytest

from conftest import FUSION, golden_amps, golden_gates, hexf, load_golden
from qmetal_amd import FusionOptions
from qmetal_amd import circuits as C
from oracle import oracle as O


def build(qm, d):
    c = qm.Circuit(d["n"])
    for kind, q, p in golden_gates(d):
        c.add(kind, q, p)
    return c


def mine(c, opt):
    return c.plan(FusionOptions(*FUSION[opt]))


def compare_plan(got, want, tag):
    assert len(got) == len(want), f"{tag}: {len(got)} ops vs reference {len(want)}"
    for k, (g, w) in enumerate(zip(got, want)):
        t = f"{tag} op {k}"
        assert g["kind_id"] == w["kind"], t
        assert g["qubits"] == w["q"], t
        m = [hexf(x) for x in w["m"]]
        gm = [v for z in g["matrix"] for v in (z.real, z.imag)]
        assert gm == m, f"{t}: matrix"
        assert g["group_qubits"] == w["gq"], f"{t}: group wires"
        wgm = [hexf(x) for x in w["gm"]]
        ggm = [v for z in g["group_matrices"] for v in (z.real, z.imag)]
        assert ggm == wgm, f"{t}: group matrices"
        assert g["masks"] == [int(x, 16) for x in w["masks"]], f"{t}: masks"
        assert g["angles"] == [hexf(x) for x in w["angles"]], f"{t}: angles"
        assert g["global_phase"] == hexf(w["global"]), f"{t}: global phase"
        assert g["block_bits"] == w["bb"], t


def _cases(name):
    d = load_golden(name)
    return list(d.items()) if isinstance(d, dict) else list(enumerate(d))


@pytest.mark.parametrize("name", ["frozen_small.json.gz", "structural.json.gz", "fuzz.json.gz"])
def test_plans_match_reference_all_options(qm, name):
    for key, d in _cases(name):
        c = build(qm, d)
        for opt, want in d["plans"].items():
            if isinstance(want, dict):  # the reference threw for this stream
                pytest.fail(f"{key}/{opt}: unexpected reference error {want}")
            compare_plan(mine(c, opt), want, f"{key}/{opt}")


@pytest.mark.parametrize("name", ["frozen_large.json.gz", "configs.json.gz"])
def test_plans_match_reference_large(qm, name):
    for key, d in _cases(name):
        c = build(qm, d)
        for opt, want in d["plans"].items():
            compare_plan(mine(c, opt), want, f"{key}/{opt}")
        for opt, size in d.get("plan_sizes", {}).items():
            assert len(mine(c, opt)) == size, f"{key}/{opt} size"


def test_plan_size_abi_matches_reference(qm):
    """qm_circuit_plan_size(fuse) equals the reference's fused/unfused sizes."""
    for name in ("frozen_large.json.gz", "configs.json.gz"):
        for key, d in _cases(name):
            c = build(qm, d)
            assert c.plan_size(True) == d["plan_sizes"]["fused"], key
            assert c.plan_size(False) == d["plan_sizes"]["none"] == len(c), key


# Published pass counts (results/fusion_reduction.txt:4-17,
# results/fusion_bench_n26.txt:5-11, results/ablation_n26_b10_extended.txt:6-13);
# random_d10 is excluded: its stream depends on the C++ library.
PUBLISHED = {
    ("ghz", 24): 24, ("ghz", 33): 33, ("qft", 24): 47, ("qft", 33): 65,
    ("qaoa_ring_p6", 24): 48, ("qaoa_ring_p6", 33): 69,
    ("tfim_trotter_20", 24): 140, ("tfim_trotter_20", 33): 200,
    ("grover_proxy_4", 24): 62, ("grover_proxy_4", 33): 89,
    ("hea_d3", 24): 87, ("hea_d3", 33): 123,
    ("ghz", 26): 26, ("qft", 26): 51, ("qaoa_ring_p6", 26): 55, ("tfim_trotter_20", 26): 160,
    ("grover_proxy_4", 26): 71, ("hea_d3", 26): 96, ("local_patch_w8", 26): 125,
}
PUBLISHED_BLOCKED10 = {
    ("ghz", 26): 26, ("qft", 26): 51, ("qaoa_ring_p6", 26): 48, ("tfim_trotter_20", 26): 140,
    ("grover_proxy_4", 26): 62, ("hea_d3", 26): 93, ("local_patch_w8", 26): 113,
}


def test_published_pass_counts(qm):
    large = load_golden("frozen_large.json.gz")
    for (name, n), want in PUBLISHED.items():
        c = build(qm, large[f"{name}_{n}"])
        assert c.plan_size(True) == want, (name, n)
    for (name, n), want in PUBLISHED_BLOCKED10.items():
        c = build(qm, large[f"{name}_{n}"])
        assert len(c.plan(FusionOptions.blocked(10))) == want, (name, n)


def test_structural_cases(qm):
    """tests/test_fusion.cpp:80-163, on the plans built here."""
    s = load_golden("structural.json.gz")
    p = build(qm, s["h4"]).plan()
    assert len(p) == 0  # H^4 collapses away
    p = build(qm, s["ry_small"]).plan()
    assert len(p) == 1  # RY(1e-6) twice survives
    p = build(qm, s["xhx"]).plan()
    assert len(p) == 1
    p = build(qm, s["cp_ladder"]).plan()
    assert len(p) == 1 and len(p[0]["masks"]) == 7
    p = build(qm, s["mixed_diag"]).plan()
    assert len(p) == 3 and len(p[0]["masks"]) == 3 and len(p[2]["masks"]) == 2
    p = build(qm, s["rz_phase"]).plan()
    assert len(p) == 2 and abs(p[1]["global_phase"] + 0.617) < 1e-12


def test_qft_config_plans(qm):
    """QFT(n) on its own is n Dense1q + (n-1) DiagonalLayer = 2n - 1 ops with
    ladders of 1..n-1 terms sharing the stage bit and no global phase."""
    for n in (26, 30, 32, 33, 34):
        c = C.to_circuit(n, C.qft(n))
        p = c.plan()
        assert len(p) == 2 * n - 1
        kinds = [op["kind"] for op in p]
        assert kinds.count("Dense1q") == n and kinds.count("DiagonalLayer") == n - 1
        for op in p:
            if op["kind"] == "DiagonalLayer":
                common = ~0
                for m in op["masks"]:
                    common &= m
                assert bin(common).count("1") >= 1 and op["global_phase"] == 0.0


def test_brickwork_u3_plan_counts(qm):
    """Brickwork with U3 lowers like the RY stand-in the reference can build:
    385 gates -> 195 passes at n=26, 475 -> 235 at n=32."""
    cfg = load_golden("configs.json.gz")
    for n, gates, passes in ((26, 385, 195), (32, 475, 235)):
        c = C.to_circuit(n, C.brickwork(n, 10))
        assert len(c) == gates
        got = c.plan()
        want = cfg[f"brickwork_ry_{n}"]["plans"]["fused"]
        assert len(got) == len(want) == passes
        assert [o["kind_id"] for o in got] == [w["kind"] for w in want]
        assert [o["group_qubits"] for o in got] == [w["gq"] for w in want]
        assert [o["qubits"] for o in got] == [w["q"] for w in want]


@pytest.mark.parametrize("name", ["frozen_small.json.gz", "fuzz.json.gz"])
def test_plans_execute_exactly_on_oracle(qm, name):
    """Each plan run on the complex128 oracle reproduces the reference
    oracle's gate-by-gate amplitudes to 1e-12."""
    cases = _cases(name)
    for key, d in cases[:: 3 if name == "fuzz.json.gz" else 1]:
        want = golden_amps(d)
        c = build(qm, d)
        for opt in ("fused", "none", "blocked8", "diag_only"):
            got = O.apply_plan(d["n"], mine(c, opt))
            assert np.abs(got - want).max() <= 1e-12, f"{key}/{opt}"


def test_u3_plans_execute_exactly_on_oracle(qm):
    for n in (5, 7):
        gates = C.brickwork(n, 6, seed=n) + C.random_circuit(n, 3, seed=n)
        want = O.run(n, gates)
        c = C.to_circuit(n, gates)
        for opt in ("fused", "none", "blocked8"):
            got = O.apply_plan(n, c.plan(FusionOptions(*FUSION[opt])))
            assert np.abs(got - want).max() <= 1e-12, opt


def test_unfused_is_one_to_one(qm):
    for gates in (C.qft(9), C.qaoa_ring(5, 2), C.tfim_trotter(5, 2), C.grover_proxy(5, 1)):
        c = C.to_circuit(9 if len(gates) == 45 else 5, gates)
        assert c.plan_size(False) == len(c)
