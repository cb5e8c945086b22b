"""The build's deterministic generators emit the reference's gate streams
(src/circuits.cpp:24-93): GHZ, QFT, the QAOA ring with its fixed angles, TFIM
Trotter steps and the Grover proxy, compared gate by gate (kind, wires and
parameters with float ==) against the streams the reference's own
generators wrote into the golden fixtures at n = 4, 7, 11, 24, 26, 33.
The hardware-efficient ansatz draws its angles from a <random> distribution
in the reference (src/circuits.cpp:117-131), so only its structure (kinds
and wires) is compared."""
import pytest

from conftest import golden_gates, load_golden
from qmetal_amd import circuits as C

DETERMINISTIC = {
    "ghz": C.ghz,
    "qft": C.qft,
    "qaoa_ring_p6": lambda n: C.qaoa_ring(n, 6),
    "tfim_trotter_20": lambda n: C.tfim_trotter(n, 20),
    "grover_proxy_4": lambda n: C.grover_proxy(n, 4),
}


def _streams():
    for fname in ("frozen_small.json.gz", "frozen_large.json.gz"):
        for key, d in load_golden(fname).items():
            yield key, d


@pytest.mark.parametrize("name", sorted(DETERMINISTIC))
def test_generator_equals_reference_stream(name):
    seen = 0
    for key, d in _streams():
        if key.rsplit("_", 1)[0] != name:
            continue
        n = d["n"]
        want = golden_gates(d)
        got = [(k, tuple(q), tuple(float(x) for x in p)) for k, q, p in DETERMINISTIC[name](n)]
        assert len(got) == len(want), (key, len(got), len(want))
        for i, (g, w) in enumerate(zip(got, want)):
            assert g == w, (key, i, g, w)
        seen += 1
    assert seen == 6, (name, seen)


def test_hardware_efficient_structure_equals_reference():
    seen = 0
    for key, d in _streams():
        if not key.startswith("hea_d3_"):
            continue
        n = d["n"]
        want = [(k, q) for k, q, _ in golden_gates(d)]
        got = [(k, tuple(q)) for k, q, _ in C.hardware_efficient(n, 3)]
        assert got == want, key
        seen += 1
    assert seen == 6


def test_config_streams_equal_generators():
    """The configuration circuits recorded with the plans (QFT at every
    benchmarked width, GHZ 20) are the generators' streams."""
    cfg = load_golden("configs.json.gz")
    for key, gen in (("qft_26", lambda: C.qft(26)), ("qft_30", lambda: C.qft(30)),
                     ("qft_32", lambda: C.qft(32)), ("qft_33", lambda: C.qft(33)),
                     ("qft_34", lambda: C.qft(34)), ("ghz_20", lambda: C.ghz(20))):
        want = golden_gates(cfg[key])
        got = [(k, tuple(q), tuple(float(x) for x in p)) for k, q, p in gen()]
        assert got == want, key
