"""Executor-level scheduling: the grouping of plan ops into HBM passes (host
logic, no GPU). The plan itself is unchanged; these tests pin that every op
runs exactly once, that ops whose wires overlap keep their plan order, and
the constraints every pass shape must meet."""
import pytest

from conftest import FUSION, golden_gates, load_golden
from qmetal_amd import FusionOptions
from qmetal_amd import circuits as C

MAX_REG = 4  # register wires beyond bit 0 in one pass


def bits(m):
    return {q for q in range(64) if (m >> q) & 1}


def moving_wires(op, wires=0):
    """Wires (other than 0) a plan op -- or the group wires `wires` of a
    Dense1qGroup -- needs in the thread's registers."""
    k, q = op["kind"], op["qubits"]
    if k == "Dense1q":
        w = {q[0]}
    elif k == "Dense1qGroup":
        w = bits(wires) if wires else set(op["group_qubits"])
    elif k == "Controlled1q":
        w = {q[1]}
    elif k == "Swap":
        w = {q[0], q[1]}
    elif k == "CCX":
        w = {q[2]}
    elif k == "Blocked1q":
        w = set(op["group_qubits"])
    else:
        w = set()
    return w - {0}


def support(op, wires=0):
    """Wires a plan op (entry) reads or writes."""
    k, q = op["kind"], op["qubits"]
    if k == "DiagonalLayer":
        s = set()
        for m in op["masks"]:
            s |= bits(m)
        return s
    if k == "Dense1qGroup":
        return bits(wires) if wires else set(op["group_qubits"])
    if k == "Blocked1q":
        return set(range(64))
    if k == "Dense1q":
        return {q[0]}
    if k in ("Controlled1q", "Swap"):
        return {q[0], q[1]}
    return {q[0], q[1], q[2]}


def mergeable(op):
    if op["kind"] == "Blocked1q":
        return False
    if op["kind"] == "DiagonalLayer":
        return bool(op["masks"]) and all(1 <= bin(m).count("1") <= 2 for m in op["masks"])
    return True


def chain_pass(ops, n):
    """A chain pass: Dense1q steps on distinct wires, each followed at most by
    a phase layer whose masks all hold its wire; wide: six or seven steps on
    wires >= 7 (n >= 16); low: 5..12 steps on wires 1..11 (n >= 12) or 1..12
    (n >= 13), then at most one on wire 0."""
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
    if not steps or len(set(steps)) != len(steps):
        return False
    if min(steps) >= 7 and len(steps) in (6, 7):
        return n >= 16
    body = steps[:-1] if steps[-1] == 0 else steps
    top = 12 if n >= 13 else 11
    return n >= 12 and 0 not in body and 5 <= len(body) <= top and max(body) <= top


def tile_pass(ws, n):
    """A generic low tile pass over entries moving the wire sets `ws` (in
    execution order): moved wires in 1..11 (n >= 12), at most 64 entries, in
    2..8 segments of at most 4 register wires."""
    if n < 12 or len(ws) > 64 or any(q > 11 for w in ws for q in w):
        return False
    segs, cur = 0, None
    for w in ws:
        if cur is None or len(cur | w) > MAX_REG:
            segs, cur = segs + 1, set(w)
        else:
            cur |= w
    return 2 <= segs <= 8


def gwide_pass(ws, n):
    """A generic wide pass: moved wires >= 7 (n >= 16), at most 64 entries,
    that split in execution order into a run on at most four wires A and a
    run on at most three new wires plus at most 4 - max(2, new) wires of A."""
    if n < 16 or len(ws) > 64 or any(q < 7 for w in ws for q in w):
        return False
    for cut in range(1, len(ws) + 1):
        a = set().union(*ws[:cut])
        if len(a) > 4:
            break
        rest = set().union(set(), *ws[cut:])
        new, keep = rest - a, rest & a
        if len(new) <= 3 and len(keep) <= 4 - max(2, len(new)):
            return True
    return False


def pass_shape(plan, entries, n):
    """'single', 'register', 'chain', 'tile' or 'gwide' for a valid pass;
    None otherwise."""
    ops = [plan[k] for k, _ in entries]
    if len(entries) == 1 and entries[0][1] == 0:
        return "single"  # one whole op through its own kernel
    if not all(mergeable(o) for o in ops):
        return None
    ws = [moving_wires(plan[k], w) for k, w in entries]
    if len(set().union(*ws)) <= MAX_REG:
        return "register"
    if chain_pass(ops, n):
        return "chain"
    if tile_pass(ws, n):
        return "tile"
    if gwide_pass(ws, n):
        return "gwide"
    return None


def check_schedule(circ, options=None):
    """Validates qm_run's schedule for `circ`; returns its number of
    launches."""
    plan = circ.plan(options)
    sched = circ.schedule(options)
    n = circ.num_qubits
    # Every op that launches something runs exactly once (a Dense1qGroup may
    # be split over passes by wire).
    covered = {}
    for entries in sched:
        for k, w in entries:
            op = plan[k]
            if op["kind"] == "Dense1qGroup" and w:
                assert not (covered.get(k, 0) & w), (k, w)
                covered[k] = covered.get(k, 0) | w
            else:
                assert k not in covered, k
                covered[k] = -1
    for k, op in enumerate(plan):
        if op["kind"] == "DiagonalLayer" and not op["masks"]:
            assert k not in covered
        elif op["kind"] == "Dense1qGroup" and covered.get(k, 0) != -1:
            assert covered.get(k, 0) == sum(1 << q for q in op["group_qubits"]), k
        else:
            assert covered.get(k) == -1, k
    # Entries whose wires overlap keep plan order (phase layers commute).
    flat = [(k, w) for entries in sched for k, w in entries]
    sup = [support(plan[k], w) for k, w in flat]
    diag = [plan[k]["kind"] == "DiagonalLayer" for k, _ in flat]
    for i in range(len(flat)):
        for j in range(i + 1, len(flat)):
            if flat[i][0] > flat[j][0] and sup[i] & sup[j] and not (diag[i] and diag[j]):
                raise AssertionError(f"order: {flat[i]} before {flat[j]}")
    for entries in sched:
        assert pass_shape(plan, entries, n), entries
    return len(sched)


def test_qft_pass_counts():
    """QFT: H on wire j then the CP ladder onto j, for j = n-1 .. 0. From 16
    qubits seven (else six) H's and their ladders on wires >= 7 share a wide
    pass; from wire 12 (11 at 12 qubits, or lower) down to 0 the rest share
    one low pass (the longer of the two shapes where both start); a wire
    left between them gets an ordinary pass."""
    for n, want in ((12, 1), (13, 1), (14, 2), (16, 2), (19, 2), (20, 2), (26, 3), (30, 4),
                    (32, 4), (33, 4), (34, 4)):
        c = C.to_circuit(n, C.qft(n))
        assert c.plan_size(True) == 2 * n - 1
        assert check_schedule(c) == want, n


def test_ghz_pass_count():
    # H(0), CX(0,1), ..., CX(18,19): targets 1..11 in one low tile pass, the
    # rest in wide and register passes
    assert check_schedule(C.to_circuit(20, C.ghz(20))) <= 3
    # below 12 qubits only register-tile passes (four new targets each)
    assert check_schedule(C.to_circuit(11, C.ghz(11))) == 3


def test_narrow_states_run_op_by_op():
    for n in (2, 3, 4):
        c = C.to_circuit(n, C.qft(n))
        assert c.schedule() == [[(k, 0)] for k in range(c.plan_size(True))]


def test_blocked_ops_stay_alone():
    c = C.to_circuit(12, C.brickwork(12, 4))
    opts = FusionOptions(*FUSION["blocked8"])
    plan = c.plan(opts)
    for entries in c.schedule(opts):
        if any(plan[k]["kind"] == "Blocked1q" for k, _ in entries):
            assert len(entries) == 1
    check_schedule(c, opts)


@pytest.mark.parametrize("opt", sorted(FUSION))
def test_schedule_invariants_on_streams(opt):
    """Every fuzz stream (n = 3..7) and the configuration circuits under every
    fusion setting: every op once, overlapping ops in plan order, every pass
    within its shape."""
    opts = FusionOptions(*FUSION[opt])
    for d in load_golden("fuzz.json.gz"):
        if FUSION[opt][4] <= d["n"]:
            check_schedule(C.to_circuit(d["n"], golden_gates(d)), opts)
    for n, gates in ((14, C.brickwork(14, 6)), (12, C.qaoa_ring(12, 6)),
                     (13, C.hardware_efficient(13, 3)), (16, C.tfim_trotter(16, 5)),
                     (10, C.grover_proxy(10, 4)), (18, C.brickwork(18, 4)),
                     (20, C.random_circuit(20, 6)), (17, C.hardware_efficient(17, 2))):
        if FUSION[opt][4] <= n:
            check_schedule(C.to_circuit(n, gates), opts)


def test_low_chains_reaching_wire_12():
    """Chains of H / phase steps whose wires reach 12 (four-wave low tiles
    from 13 qubits; two-wave ones, or none, below): every pass within its
    shape, every op once, overlapping ops in plan order."""
    for n in (12, 13, 16):
        for seed in range(3):
            r = C.SplitMix64(77 * n + seed)
            gates = []
            for i in range(4):
                ws = [12 - k for k in range(4)] if i % 2 else []
                while len(ws) < 5 + r.next_u64() % 8:
                    q = 1 + r.next_u64() % 12
                    if q not in ws:
                        ws.append(q)
                ws = [w for w in ws if w < n] + [0]
                for w in ws:
                    gates.append((C.H, (w,), ()))
                    gates.append((C.CP, ((w + 1 + r.next_u64() % (n - 1)) % n, w),
                                  (r.uniform(-3, 3),)))
                gates.append((C.CX, (n - 1, 0), ()))
            check_schedule(C.to_circuit(n, gates))


def test_scheduling_reduces_launches():
    for n, gates in ((20, C.brickwork(20, 10)), (16, C.qaoa_ring(16, 6)),
                     (18, C.hardware_efficient(18, 3))):
        c = C.to_circuit(n, gates)
        assert check_schedule(c) < c.plan_size(True)


def test_brickwork_launch_counts():
    """Brickwork depth 10: reordering across the layers' disjoint wires and
    splitting the U3 groups by wire packs several layers per pass."""
    for n, most in ((26, 20), (32, 24)):
        c = C.to_circuit(n, C.brickwork(n, 10))
        assert check_schedule(c) <= most, n
