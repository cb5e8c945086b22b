"""Executor-level pass merging: the grouping of plan ops into HBM passes
(host logic, no GPU). The plan itself is unchanged; these tests pin which
ops share a pass and the constraints every pass must meet."""
import pytest

from conftest import FUSION, golden_gates, load_golden
from qmetal_amd import FusionOptions
from qmetal_amd import circuits as C

MAX_REG = 4  # register wires beyond bit 0 in one pass


def moving_wires(op):
    """Wires (other than 0) a plan op needs in the thread's registers."""
    k, q = op["kind"], op["qubits"]
    if k == "Dense1q":
        w = {q[0]}
    elif k == "Dense1qGroup":
        w = set(op["group_qubits"])
    elif k == "Controlled1q":
        w = {q[1]}
    elif k == "Swap":
        w = {q[0], q[1]}
    elif k == "CCX":
        w = {q[2]}
    else:
        w = set()
    return w - {0}


def chain_pass(ops, n):
    """A chain pass: Dense1q steps on distinct wires, each followed at most by
    a phase layer whose masks all hold its wire; wide: six or seven steps on
    wires >= 7 (n >= 16); low: 5..11 steps on wires 1..11, then at most one
    on wire 0 (n >= 12)."""
    steps, last = [], None
    for o in ops:
        if o["kind"] == "Dense1q":
            last = o["qubits"][0]
            steps.append(last)
        elif o["kind"] == "DiagonalLayer" and last is not None:
            if not all((m >> last) & 1 for m in o["masks"]):
                return False
            last = None
        else:
            return False
    if len(set(steps)) != len(steps):
        return False
    if min(steps) >= 7 and len(steps) in (6, 7):
        return n >= 16
    body = steps[:-1] if steps[-1] == 0 else steps
    return n >= 12 and 0 not in body and 5 <= len(body) <= 11 and max(body) <= 11


def tile_pass(ops, n):
    """A generic low tile pass: mergeable ops whose moving wires lie in
    0..11 (n >= 12), at most 64 of them, that split in plan order into at
    most 8 segments of at most 4 register wires."""
    if n < 12 or len(ops) > 64:
        return False
    segs, cur = 0, None
    for o in ops:
        if o["kind"] == "Blocked1q":
            return False
        w = moving_wires(o)
        if any(q > 11 for q in w):
            return False
        if cur is None or len(cur | w) > MAX_REG:
            segs, cur = segs + 1, set(w)
        else:
            cur |= w
    return 2 <= segs <= 8


def gwide_pass(ops, n):
    """A generic wide pass: mergeable ops on wires >= 7 (n >= 16), at most
    64, that split in plan order into a run on at most four wires A and a run
    on at most three new wires plus at most 4 - max(2, new) wires of A."""
    if n < 16 or len(ops) > 64:
        return False
    ws = [moving_wires(o) for o in ops]
    if any(o["kind"] == "Blocked1q" for o in ops) or any(q < 7 for w in ws for q in w):
        return False
    for cut in range(1, len(ops)):
        a = set().union(*ws[:cut])
        if len(a) > 4:
            break
        rest = set().union(*ws[cut:])
        new, keep = rest - a, rest & a
        if len(new) <= 3 and len(keep) <= 4 - max(2, len(new)):
            return True
    return False


def check_grouping(circ, options=None):
    plan = circ.plan(options)
    passes = circ.passes(options)
    assert len(passes) == len(plan)
    seen = [p for p in passes if p >= 0]
    # passes are contiguous runs in plan order, numbered from 0
    assert seen == sorted(seen)
    if seen:
        assert sorted(set(seen)) == list(range(max(seen) + 1))
    for k, op in enumerate(plan):
        if op["kind"] == "DiagonalLayer" and not op["masks"]:
            assert passes[k] == -1
        else:
            assert passes[k] >= 0
    groups = {}
    for k, p in enumerate(passes):
        if p >= 0:
            groups.setdefault(p, []).append(plan[k])
    for ops in groups.values():
        if len(ops) > 1:
            wires = set().union(*(moving_wires(o) for o in ops))
            assert all(o["kind"] != "Blocked1q" for o in ops)
            if len(wires) > MAX_REG:
                n = circ.num_qubits
                assert chain_pass(ops, n) or tile_pass(ops, n) or gwide_pass(ops, n), ops
    return max(seen) + 1 if seen else 0


def test_qft_pass_counts():
    """QFT: H on wire j then the CP ladder onto j, for j = n-1 .. 0. From 16
    qubits seven (else six) H's and their ladders on wires >= 7 share a wide
    pass; from wire 11 (or lower) down to 0 the rest share one low pass; a
    wire left between them gets an ordinary pass."""
    for n, want in ((12, 1), (13, 2), (16, 2), (20, 3), (26, 3), (30, 4), (32, 4), (33, 4),
                    (34, 5)):
        c = C.to_circuit(n, C.qft(n))
        assert c.plan_size(True) == 2 * n - 1
        assert check_grouping(c) == want, n


def test_ghz_pass_count():
    # H(0), CX(0,1), ..., CX(18,19): targets 1..11 in one low tile pass (three
    # segments), then four new targets per register-tile pass
    assert check_grouping(C.to_circuit(20, C.ghz(20))) == 3
    # below 12 qubits there are no tile passes
    assert check_grouping(C.to_circuit(11, C.ghz(11))) == 3


def test_narrow_states_run_op_by_op():
    for n in (2, 3, 4):
        c = C.to_circuit(n, C.qft(n))
        assert c.passes() == list(range(c.plan_size(True)))


def test_blocked_ops_stay_alone():
    c = C.to_circuit(12, C.brickwork(12, 4))
    opts = FusionOptions(*FUSION["blocked8"])
    plan, passes = c.plan(opts), c.passes(opts)
    for k, op in enumerate(plan):
        if op["kind"] == "Blocked1q":
            assert passes.count(passes[k]) == 1
    check_grouping(c, opts)


@pytest.mark.parametrize("opt", sorted(FUSION))
def test_grouping_invariants_on_streams(opt):
    """Every fuzz stream (n = 3..7) and the configuration circuits under every
    fusion setting: contiguous passes, each within the register budget."""
    opts = FusionOptions(*FUSION[opt])
    for d in load_golden("fuzz.json.gz"):
        if FUSION[opt][4] <= d["n"]:
            check_grouping(C.to_circuit(d["n"], golden_gates(d)), opts)
    for n, gates in ((14, C.brickwork(14, 6)), (12, C.qaoa_ring(12, 6)),
                     (13, C.hardware_efficient(13, 3)), (16, C.tfim_trotter(16, 5)),
                     (10, C.grover_proxy(10, 4))):
        if FUSION[opt][4] <= n:
            check_grouping(C.to_circuit(n, gates), opts)


def test_merging_reduces_launches():
    for n, gates in ((20, C.brickwork(20, 10)), (16, C.qaoa_ring(16, 6)),
                     (18, C.hardware_efficient(18, 3))):
        c = C.to_circuit(n, gates)
        assert check_grouping(c) < c.plan_size(True)
