"""Write the golden fixtures under tests/golden/ from the reference's own CPU
core, compiled by oracle/Makefile into oracle/_ref/ref_dump.

    python tests/golden/make_golden.py

Each fixture holds gate streams and what the reference builder and oracle
make of them: plans under six FusionOptions (op kinds, wires, FP64 matrices,
masks, angles, global phases as hex floats) and, for n <= 14, complex128
amplitudes plus a few expectations. The streams are serialised from the same
run, so plan parity is checked on identical gate streams; the reference's
random_d10 stream depends on the C++ library's <random> distributions and is
only meaningful through these recorded streams.
"""
from __future__ import annotations

import gzip
import json
import math
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from qmetal_amd import circuits as C  # noqa: E402

REF = os.path.join(ROOT, "oracle", "_ref", "ref_dump")
FROZEN = ["ghz", "qft", "qaoa_ring_p6", "tfim_trotter_20", "grover_proxy_4", "random_d10",
          "hea_d3", "local_patch_w8"]


def ensure_ref():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "ref"], check=True)


def frozen(name, n, amps):
    out = subprocess.run([REF, "frozen", name, str(n)] + (["amps"] if amps else []),
                         capture_output=True, text=True, check=True).stdout
    return json.loads(out)


def stream(n, gates, amps=True):
    lines = [str(n)]
    for kind, q, p in gates:
        qq = list(q) + [0] * (3 - len(q))
        pp = list(p) + [0.0] * (3 - len(p))
        lines.append(f"{kind} {len(q)} {qq[0]} {qq[1]} {qq[2]} "
                     f"{float(pp[0]).hex()} {float(pp[1]).hex()} {float(pp[2]).hex()}")
    out = subprocess.run([REF, "stream"] + (["amps"] if amps else []), input="\n".join(lines) + "\n",
                         capture_output=True, text=True, check=True).stdout
    return json.loads(out)


def slim(d, keep=("fused", "blocked10"), sizes=("none", "diag_only", "win_only", "blocked8")):
    """Large cases: full dumps for `keep`, only op counts for the rest."""
    plans = d["plans"]
    d["plans"] = {k: plans[k] for k in keep}
    d["plan_sizes"] = {k: len(plans[k]) for k in plans}
    return d


def fuzz_streams(count, seed=0xF022):
    """Random streams aimed at the plan passes' corner cases: ZZ triples that
    match and near-miss, diagonal runs, cancelling pairs, small angles,
    repeated wires and every gate kind of the reference."""
    rng = C.SplitMix64(seed)
    out = []
    for _ in range(count):
        n = 3 + rng.next_u64() % 5
        length = 6 + rng.next_u64() % 40
        gates = []
        while len(gates) < length:
            r = rng.next_u64() % 16
            q = int(rng.next_u64() % n)
            q2 = int((q + 1 + rng.next_u64() % (n - 1)) % n)
            q3 = [w for w in range(n) if w not in (q, q2)][int(rng.next_u64() % (n - 2))]
            ang = rng.uniform(-2 * math.pi, 2 * math.pi)
            if r == 0:
                gates += C.zz(q, q2, ang)
            elif r == 1:  # near miss: RZ on the control
                gates += [(C.CX, (q, q2), ()), (C.RZ, (q,), (ang,)), (C.CX, (q, q2), ())]
            elif r == 2:  # cancelling pair
                k = [C.H, C.X, C.Y, C.Z][rng.next_u64() % 4]
                gates += [(k, (q,), ()), (k, (q,), ())]
            elif r == 3:
                gates += [(C.S, (q,), ()), (C.SDG, (q,), ()), (C.T, (q,), ()), (C.TDG, (q,), ())]
            elif r == 4:  # small and opposite angles
                a = 1e-7 * rng.uniform(-1, 1)
                gates += [(C.RY, (q,), (a,)), (C.RZ, (q,), (ang,)), (C.RZ, (q,), (-ang,))]
            elif r == 5:
                gates.append((C.CP, (q, q2), (ang,)))
            elif r == 6:
                gates.append((C.CZ, (q, q2), ()))
            elif r == 7:
                gates.append((C.P, (q,), (ang,)))
            elif r == 8:
                gates.append(([C.RX, C.RY, C.RZ][rng.next_u64() % 3], (q,), (ang,)))
            elif r == 9:
                gates.append(([C.H, C.X, C.Y, C.Z, C.S, C.SDG, C.T, C.TDG][rng.next_u64() % 8], (q,), ()))
            elif r == 10:
                gates.append(([C.CX, C.CY][rng.next_u64() % 2], (q, q2), ()))
            elif r == 11:
                gates.append((C.SWAP, (q, q2), ()))
            elif r == 12:
                gates.append((C.CCX, (q, q2, q3), ()))
            elif r == 13:  # a layer of 1q gates on many wires
                for w in range(n):
                    gates.append(([C.RX, C.RY, C.H][rng.next_u64() % 3], (w,), (ang * (w + 1),)))
            elif r == 14:  # a diagonal ladder
                for w in range(n):
                    if w != q:
                        gates.append((C.CP, (w, q), (math.pi / (1 << (abs(w - q) % 10)),)))
            else:
                gates.append((C.RZ, (q,), (ang,)))
        out.append((int(n), gates[:length + 3]))
    return out


def structural():
    """The hand-built cases of the reference's fusion suite
    (tests/test_fusion.cpp:80-163)."""
    cases = {
        "h4": (3, [(C.H, (1,), ())] * 4),
        "ry_small": (2, [(C.RY, (0,), (1e-6,)), (C.RY, (0,), (1e-6,))]),
        "xhx": (2, [(C.X, (0,), ()), (C.H, (1,), ()), (C.X, (0,), ())]),
        "cp_ladder": (8, [(C.CP, (k, 0), (math.pi / float(1 << k),)) for k in range(1, 8)]),
        "mixed_diag": (4, [(C.T, (0,), ()), (C.CZ, (0, 1), ()), (C.RZ, (2,), (0.3,)),
                           (C.H, (3,), ()), (C.S, (1,), ()), (C.CP, (2, 3), (0.9,))]),
        "rz_phase": (2, [(C.H, (0,), ()), (C.RZ, (0,), (1.234,)), (C.CZ, (0, 1), ())]),
    }
    return {k: stream(n, g) for k, (n, g) in cases.items()}


def configs():
    out = {}
    for n in (26, 30, 32, 33, 34):
        out[f"qft_{n}"] = slim(stream(n, C.qft(n), amps=False), keep=("fused",))
    out["ghz_20"] = slim(stream(20, C.ghz(20), amps=False), keep=("fused", "none"))
    for n in (20, 26, 32):
        out[f"hea_d3_{n}"] = slim(stream(n, C.hardware_efficient(n, 3, seed=0x4EA + n), amps=False),
                                  keep=("fused",))
    rng = C.SplitMix64(0x0A0A)
    for n in (10, 13, 16, 20):
        g = [rng.uniform(0, math.pi) for _ in range(6)]
        b = [rng.uniform(0, math.pi / 2) for _ in range(6)]
        out[f"qaoa_p6_{n}"] = slim(stream(n, C.qaoa_ring(n, 6, g, b), amps=n <= 13), keep=("fused",))
    # Brickwork uses the appended U3 gate, which the reference does not have;
    # its plan structure is pinned with RY standing in for each U3 (both are
    # non-diagonal one-qubit gates that never cancel).
    for n in (26, 32):
        bw = [(C.RY, q, (p[0] + 0.1,)) if k == C.U3 else (k, q, p) for k, q, p in C.brickwork(n, 10)]
        out[f"brickwork_ry_{n}"] = slim(stream(n, bw, amps=False), keep=("fused",))
    return out


def write(name, obj):
    path = os.path.join(HERE, name)
    with gzip.open(path, "wt") as f:
        json.dump(obj, f, separators=(",", ":"))
    print(f"{name}: {os.path.getsize(path) / 1024:.0f} KiB")


def main():
    ensure_ref()
    write("frozen_small.json.gz",
          {f"{nm}_{n}": frozen(nm, n, True) for nm in FROZEN for n in (4, 7, 11)})
    write("frozen_large.json.gz",
          {f"{nm}_{n}": slim(frozen(nm, n, False)) for nm in FROZEN for n in (24, 26, 33)})
    write("fuzz.json.gz", [stream(n, g) for n, g in fuzz_streams(400)])
    write("structural.json.gz", structural())
    write("configs.json.gz", configs())


if __name__ == "__main__":
    main()
