"""Headline benchmark: fused QFT on one MI355X, in place on a seeded product
state, reported as seconds per circuit and as measured HBM bandwidth
against the 8 TB/s peak, with the dominant kernel's roofline measured by HIP
events on the simulator's stream and the reference's complex128 CPU oracle
(built from its sources into oracle/_ref; the build's own restatement where
that is absent) timed beside it. Also in the line: the batched QAOA adjoint
(grads/s batched vs one circuit per launch, with its on-chip FLOP roofline)
and the HEA depth-3 adjoint at n=32 (s per gradient, kernel roofline).
"plan_equiv_*" divide the reference plan's per-op bytes by the time: merged
passes move each amplitude once for many plan ops, so these exceed the
physical peak by design and are not bandwidths.

    python bench.py [--steps K] [--warmup W] [--n 33] [--cpu-n 26] [--no-cpu]

One step is one qm_run of QFT(n) (n(n+1)/2 gates, a plan of 2n-1 fused ops,
which the executor merges into a few read-modify-write passes over HBM, 4 at
n=33; --no-merge launches one kernel per plan op) on the resident state. Inputs are synthetic: RY(theta_q), RZ(phi_q) on every wire with
theta, phi ~ U[0, 2 pi) from splitmix64(2024), applied before timing.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def qft_bytes(n: int) -> float:
    """Algorithmic bytes of the fused QFT: n Dense1q passes of 16 B/amplitude
    plus n-1 phase ladders that change only the half with the stage bit set."""
    return (24 * n - 8) * float(1 << n)


def cpu_baseline(n_cpu: int, n_gpu: int, threads: int):
    """The complex128 oracle (own restatement of the reference's CPU oracle),
    gate by gate, on the same workload at n_cpu."""
    from oracle import oracle as O
    from qmetal_amd import circuits as C
    O.set_threads(threads)
    prep = O.run(n_cpu, C.product_state_prep(n_cpu, seed=2024))
    gates = C.qft(n_cpu)
    t0 = time.perf_counter()
    O.run(n_cpu, gates, state=prep)
    dt = time.perf_counter() - t0
    O.set_threads(1)
    scale = float(1 << (n_gpu - n_cpu)) * (n_gpu * (n_gpu + 1) / 2) / (n_cpu * (n_cpu + 1) / 2)
    return dt, dt * scale


REF_DUMP = os.path.join(ROOT, "oracle", "_ref", "ref_dump")


def ref_cpu_baseline(n_cpu: int, n_gpu: int):
    """The reference's own complex128 oracle (Reference::run, one thread),
    compiled from its sources into oracle/_ref/ref_dump, on the same
    workload at n_cpu: the product-state preparation untimed, then QFT gate
    by gate, timed inside the binary. None where the binary is absent."""
    import subprocess
    from qmetal_amd import circuits as C
    if not os.access(REF_DUMP, os.X_OK):
        return None
    prep = C.product_state_prep(n_cpu, seed=2024)
    lines = [str(n_cpu)]
    for kind, q, p in prep + C.qft(n_cpu):
        qq = list(q) + [0] * (3 - len(q))
        pp = [float(v) for v in p] + [0.0] * (3 - len(p))
        lines.append(f"{kind} {len(q)} {qq[0]} {qq[1]} {qq[2]} "
                     f"{pp[0].hex()} {pp[1].hex()} {pp[2].hex()}")
    out = subprocess.run([REF_DUMP, "time", str(len(prep))], input="\n".join(lines) + "\n",
                         capture_output=True, text=True, check=True, timeout=600).stdout
    r = json.loads(out)
    if abs(r["norm"] - 1) > 1e-9 or r["gates"] != n_cpu * (n_cpu + 1) // 2:
        raise RuntimeError(f"reference oracle run looks wrong: {r}")
    dt = float(r["seconds"])
    scale = float(1 << (n_gpu - n_cpu)) * (n_gpu * (n_gpu + 1) / 2) / (n_cpu * (n_cpu + 1) / 2)
    return dt, dt * scale, r


def batched_qaoa(qm, C, count: int, repeats: int, plain_count: int):
    """4096 QAOA-ring p=6 instances (n ~ U{10..20}, gamma ~ U[0, pi],
    beta ~ U[0, pi/2], splitmix64): energy <sum Z_i Z_i+1> and the full
    adjoint gradient of every instance from one persistent graph launch,
    against the same instances solved one circuit per launch."""
    import numpy as np
    inst = C.qaoa_instances(count)
    t0 = time.perf_counter()
    circs = [C.to_circuit(n, C.qaoa_ring(n, 6, g, b)) for n, g, b in inst]
    hams = [C.to_hamiltonian(C.ring_zz_hamiltonian(n)) for n, _, _ in inst]
    t_build = time.perf_counter() - t0
    b = qm.Batch()
    t0 = time.perf_counter()
    b.prepare(circs, hams)
    t_prep = time.perf_counter() - t0
    t0 = time.perf_counter()
    cfg = b.autotune()
    t_tune = time.perf_counter() - t0
    e, g, _ = b.run()
    dev_ms, wall = [], []
    for _ in range(repeats):
        w0 = time.perf_counter()
        e, g, ms = b.run()
        wall.append(time.perf_counter() - w0)
        dev_ms.append(ms)
    batch_s = min(wall)
    # One circuit per launch: qm_energy_and_gradient per instance, one
    # simulator per width (allocated up front, outside the timing).
    sims = {n: qm.Simulator(n) for n in sorted({n for n, _, _ in inst})}
    k = min(plain_count, count)
    worst_g, worst_e = 0.0, 0.0
    p0 = time.perf_counter()
    for i in range(k):
        n = inst[i][0]
        e0, g0 = sims[n].energy_and_gradient(circs[i], hams[i])
        worst_g = max(worst_g, float(np.abs(np.array(g0) - g[i]).max()))
        worst_e = max(worst_e, abs(e0 - e[i]) / max(1.0, abs(e0)))
    plain_s = (time.perf_counter() - p0) * count / k
    # The same calls with one launch per gate in the backward sweep and one
    # per plan op forward (the reference executor's shape), on every 8th
    # instance of the sample.
    for s in sims.values():
        s.set_pass_merging(False)
    sub = range(0, k, 8)
    q0 = time.perf_counter()
    for i in sub:
        sims[inst[i][0]].energy_and_gradient(circs[i], hams[i])
    per_gate_s = (time.perf_counter() - q0) * count / len(sub)
    for s in sims.values():
        s.close()
    params = sum(12 * n for n, _, _ in inst)
    # On-chip roofline of the batch: the adjoint's algorithmic arithmetic per
    # instance, about 404 n 2^n flops of gate passes (forward and backward on
    # both states, FP32) and 96 n 2^n of overlaps (the products are FP32 in
    # the tile sweeps that run the wide instances, FP64 in the persistent
    # kernel; every sum across threads is FP64), over the device time. Peak:
    # the published FP32 rate (157.3 TFLOP/s, two FMAs per lane per cycle
    # with packed instructions, which this build leaves off) and the
    # unpacked VALU rate (78.6 TFLOP/s) beside it.
    f_gate = sum(404.0 * n * (1 << n) for n, _, _ in inst)
    f_ovl = sum(96.0 * n * (1 << n) for n, _, _ in inst)
    dev_s = min(dev_ms) * 1e-3
    ach = (f_gate + f_ovl) / dev_s / 1e12
    roof = {"bound": "valu", "unit": "TFLOP/s", "achieved": round(ach, 2), "peak": 157.3,
            "frac": round(ach / 157.3, 4), "peak_unpacked": 78.6,
            "frac_unpacked": round(ach / 78.6, 4), "flops_gate": f_gate,
            "flops_overlap": f_ovl}
    return {
        "workload": f"qaoa_ring_p6_x{count}_n10-20",
        "instances": count,
        "batched_s": round(batch_s, 4),
        "batched_device_ms": round(min(dev_ms), 3),
        "batched_grads_per_s": round(count / batch_s, 1),
        "batched_param_derivs_per_s": round(params / batch_s, 1),
        "plain_s": round(plain_s, 3),
        "plain_grads_per_s": round(count / plain_s, 1),
        "plain_sample": k,
        "speedup": round(plain_s / batch_s, 2),
        "plain_per_gate_s": round(per_gate_s, 3),
        "speedup_vs_per_gate": round(per_gate_s / batch_s, 2),
        "max_abs_grad_diff_vs_plain": worst_g,
        "max_rel_energy_diff_vs_plain": worst_e,
        "roofline": roof,
        "autotuned": cfg,
        "host_build_s": round(t_build, 2),
        "prepare_s": round(t_prep, 2),
        "autotune_s": round(t_tune, 2),
    }


def oracle_threads() -> int:
    """Host cores this process may use: its CPU affinity set, narrowed by
    OMP_NUM_THREADS where the environment sets it (a shared GPU box gives
    each GPU's process a share of the host)."""
    try:
        cores = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        cores = os.cpu_count() or 1
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        cores = min(cores, int(env))
    return max(1, cores)


def host_cpu() -> dict:
    """What the CPU baseline ran on: the model name (/proc/cpuinfo, as lscpu
    prints it), logical CPUs (nproc), the process's affinity set and
    OMP_NUM_THREADS."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        affinity = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        affinity = None
    return {"model": model, "nproc": os.cpu_count(), "affinity": affinity,
            "omp_num_threads": os.environ.get("OMP_NUM_THREADS")}


def ref_cpu_n30(n_gpu: int):
    """The reference's own oracle at its ceiling, n = 30 (src/reference.cpp:100-104,
    a 16 GiB complex128 state): QFT(30)'s first stage -- H on wire 29 and the
    29 CP gates onto it, gate by gate, one thread -- from |0...0> (the gate
    loops do the same work for any amplitudes). Scaled to the n_gpu circuit
    by amplitudes x gates. None where oracle/_ref is absent."""
    import subprocess
    from qmetal_amd import circuits as C
    if not os.access(REF_DUMP, os.X_OK):
        return None
    n = 30
    stage = C.qft(n)[:n]
    lines = [str(n)]
    for kind, q, p in stage:
        qq = list(q) + [0] * (3 - len(q))
        pp = [float(v) for v in p] + [0.0] * (3 - len(p))
        lines.append(f"{kind} {len(q)} {qq[0]} {qq[1]} {qq[2]} "
                     f"{pp[0].hex()} {pp[1].hex()} {pp[2].hex()}")
    out = subprocess.run([REF_DUMP, "time", "0"], input="\n".join(lines) + "\n",
                         capture_output=True, text=True, check=True, timeout=900).stdout
    r = json.loads(out)
    if abs(r["norm"] - 1) > 1e-9 or r["gates"] != n:
        raise RuntimeError(f"reference oracle run looks wrong: {r}")
    dt = float(r["seconds"])
    scale = float(1 << (n_gpu - n)) * (n_gpu * (n_gpu + 1) / 2) / n
    return dt, dt * scale


def scale_sweep(qm, C, nmin: int, nmax: int, emit):
    """The reference's scale bench (bench/scale.mm:30-90) on this build: for
    every n from nmin to nmax, GHZ (norm, <Z_0>, <Z_0 Z_(n-1)>, 4096 shots on
    the two poles) and QFT from |0...0> (norm, max |<Z_q>| over up to eight
    qubits), each with the allocation time and the run time, one JSON line
    per circuit and width; `ok` applies the reference's thresholds."""
    for kind in ("ghz", "qft"):
        for n in range(nmin, nmax + 1):
            gates = C.ghz(n) if kind == "ghz" else C.qft(n)
            c = C.to_circuit(n, gates)
            t0 = time.perf_counter()
            s = qm.Simulator(n)
            s.synchronize()
            t_alloc = time.perf_counter() - t0
            try:
                t0 = time.perf_counter()
                s.run(c)
                s.synchronize()
                t_run = time.perf_counter() - t0
                norm = s.norm()
                rec = {"config": f"scale_{kind}_n{n}", "gib": (8 << n) / 2**30,
                       "alloc_s": round(t_alloc, 3), "run_s": round(t_run, 4),
                       "gates": len(gates), "norm": norm}
                if kind == "ghz":
                    z0 = s.expectation_pauli(z=1)
                    zz = s.expectation_pauli(z=1 | (1 << (n - 1)))
                    shots = s.sample(4096, 7)
                    ones = sum(1 for v in shots if v == (1 << n) - 1)
                    zeros = sum(1 for v in shots if v == 0)
                    other = len(shots) - ones - zeros
                    rec.update({"z0": z0, "z0_zlast": zz, "zeros": zeros, "ones": ones,
                                "other": other,
                                "ok": abs(norm - 1) < 1e-3 and abs(z0) < 1e-3 and
                                abs(zz - 1) < 1e-3 and other == 0})
                else:
                    step = n // 8 if n > 8 else 1
                    wz = max(abs(s.expectation_pauli(z=1 << q)) for q in range(0, n, step))
                    rec.update({"max_abs_z_q": wz, "ok": abs(norm - 1) < 1e-2 and wz < 1e-2})
                emit(rec)
            finally:
                s.close()


def suite(qm, C, emit):
    """Every configuration of the measurement plan, one JSON object each:
    GHZ n=20, GHZ n=30..33, QFT scale sweep (merged passes vs one launch per plan op vs the
    unfused plan, alternated), brickwork,
    HEA adjoint gradients. Correctness is checked against the complex128
    oracle where it fits (n <= 26) and against invariants above."""
    import numpy as np
    from oracle import oracle as O
    th = oracle_threads()

    def best_of(fn, reps):
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            fn()
            ts.append(time.perf_counter() - t0)
        return min(ts)

    # --- GHZ n=20 ---------------------------------------------------------
    n = 20
    g = C.to_circuit(n, C.ghz(n))
    with qm.Simulator(n) as s:
        def run_ghz():
            s.reset()
            s.run(g)
            s.synchronize()
        t = best_of(run_ghz, 5)
        amps = s.amplitudes().astype(np.complex128)
        zq = max(abs(s.expectation_pauli(z=1 << q)) for q in range(n))
        zz = s.expectation_pauli(z=1 | (1 << (n - 1)))
        shots = s.sample(100000, 7)
        norm = s.norm()
    O.set_threads(1)
    t0 = time.perf_counter()
    ref = O.run(n, C.ghz(n))
    t_cpu = time.perf_counter() - t0
    emit({"config": "ghz_n20", "gpu_s": t, "cpu_oracle_s_1thread": t_cpu,
          "max_abs_diff_vs_oracle": float(np.abs(amps - ref).max()), "norm": norm,
          "max_abs_z_q": zq, "z0_z19": zz, "shots": 100000,
          "distinct_outcomes": sorted(set(shots)), "plan_ops": g.plan_size(True)})

    # --- GHZ scale sweep (the reference's published GHZ row, README.md:12-19:
    # 0.81 / 1.65 / 3.41 / 7.67 s at n = 30..33 on M4 Max): wall time per
    # circuit from |0...0>, merged passes vs one launch per plan op, with the
    # invariants <Z_q> = 0, <Z_0 Z_(n-1)> = 1, norm 1 and two-pole shots.
    for n in (30, 31, 32, 33):
        g = C.to_circuit(n, C.ghz(n))
        with qm.Simulator(n) as s:
            times = {"merged": [], "op_by_op": []}
            for _ in range(2):
                for arm in times:
                    s.set_pass_merging(arm == "merged")
                    s.reset()
                    s.synchronize()
                    t0 = time.perf_counter()
                    s.run(g)
                    s.synchronize()
                    times[arm].append(time.perf_counter() - t0)
            s.set_pass_merging(True)
            s.reset()
            before = s.launch_count
            s.run(g)
            launches = s.launch_count - before
            zq = max(abs(s.expectation_pauli(z=1 << q)) for q in (0, 1, n // 2, n - 1))
            zz = s.expectation_pauli(z=1 | (1 << (n - 1)))
            norm = s.norm()
            shots = s.sample(1000, 7)
        emit({"config": f"ghz_n{n}", "merged_s": min(times["merged"]),
              "op_by_op_s": min(times["op_by_op"]), "launches": launches,
              "plan_ops": g.plan_size(True), "norm": norm, "max_abs_z_q": zq, "z0_zlast": zz,
              "shots_on_the_two_poles": set(shots) <= {0, (1 << n) - 1},
              "reference_m4max_s": {30: 0.81, 31: 1.65, 32: 3.41, 33: 7.67}[n]})

    # --- QFT scale sweep: fused plan with merged passes, the same plan one
    # launch per op, and the unfused plan (merged), alternated -------------
    for n in (26, 30, 32, 33, 34):
        qft = C.to_circuit(n, C.qft(n))
        with qm.Simulator(n) as s:
            s.run(C.to_circuit(n, C.product_state_prep(n, seed=2024)))
            s.run(qft)
            s.synchronize()
            times = {"merged": [], "op_by_op": [], "unfused": []}
            for _ in range(2):
                for arm in times:
                    s.set_pass_merging(arm != "op_by_op")
                    t0 = time.perf_counter()
                    s.run(qft, fuse=arm != "unfused")
                    s.synchronize()
                    times[arm].append(time.perf_counter() - t0)
            s.set_pass_merging(True)
            s.reset()
            s.run(qft)
            norm0 = s.norm()
            z0 = max(abs(s.expectation_pauli(z=1 << q)) for q in range(n))
            diff = None
            if n == 26:
                s.reset()
                prep = C.product_state_prep(n, seed=2024)
                s.run(C.to_circuit(n, prep))
                s.run(qft)
                got = s.amplitudes().astype(np.complex128)
                O.set_threads(th)
                ref = O.run(n, prep + C.qft(n))
                O.set_threads(1)
                diff = float(np.abs(got - ref).max())
        tf, to, tu = min(times["merged"]), min(times["op_by_op"]), min(times["unfused"])
        emit({"config": f"qft_n{n}", "fused_s": tf, "fused_op_by_op_s": to, "unfused_s": tu,
              "speedup_vs_op_by_op": to / tf, "speedup_vs_unfused": tu / tf,
              "plan_ops_fused": qft.plan_size(True), "plan_ops_unfused": qft.plan_size(False),
              "launches_fused": max(qft.passes()) + 1,
              "plan_equiv_GBps": qft_bytes(n) / tf / 1e9,
              "plan_equiv_x_peak": qft_bytes(n) / tf / 1e9 / HBM_PEAK_GBS,
              "qft_zero_norm": norm0, "qft_zero_max_abs_z_q": z0,
              "max_abs_diff_vs_oracle_product_state": diff})

    # --- brickwork depth 10 (Haar U3 per wire, alternating CX layers) -----
    for n in (26, 32):
        gates = C.brickwork(n, 10)
        c = C.to_circuit(n, gates)
        with qm.Simulator(n) as s:
            s.run(c)
            s.synchronize()
            def run_bw():
                s.reset()
                s.run(c)
                s.synchronize()
            s.set_pass_merging(False)
            t_op = best_of(run_bw, 1)
            s.set_pass_merging(True)
            t = best_of(run_bw, 2)
            norm = s.norm()
            diff = None
            if n == 26:
                got = s.amplitudes().astype(np.complex128)
                O.set_threads(th)
                ref = O.run(n, gates)
                O.set_threads(1)
                diff = float(np.abs(got - ref).max())
        bytes_ = 0.0
        for op in c.plan():
            k = op["kind"]
            bytes_ += (16 if k in ("Dense1q", "Dense1qGroup", "Blocked1q", "DiagonalLayer")
                       else 8 if k in ("Controlled1q", "Swap") else 4) * float(1 << n)
        emit({"config": f"brickwork_n{n}_d10", "gates": len(c), "plan_ops": c.plan_size(True),
              "launches": max(c.passes()) + 1, "gpu_s": t, "op_by_op_s": t_op,
              "plan_equiv_GBps": bytes_ / t / 1e9, "norm": norm,
              "max_abs_diff_vs_oracle": diff})

    # --- HEA depth 3 adjoint, TFIM ring -----------------------------------
    for n in (20, 26, 32):
        P = 6 * n
        rng = C.SplitMix64(0x4EA + n)
        params = [rng.uniform(0, 2 * math.pi) for _ in range(P)]
        gates = C.hardware_efficient(n, 3, params=params)
        terms = C.tfim_ring_hamiltonian(n)
        c, h = C.to_circuit(n, gates), C.to_hamiltonian(terms)
        rot = [k for k, (kind, _, _) in enumerate(gates) if kind in (8, 9, 10)]
        with qm.Simulator(n) as s:
            s.energy_and_gradient(c, h)
            t0 = time.perf_counter()
            e, gr = s.energy_and_gradient(c, h)
            t = time.perf_counter() - t0
            gr = np.array(gr)
            # central differences (h = 1e-3) on a few parameters, on the GPU
            fd_err = 0.0
            for j in rot[:: max(1, len(rot) // 4)][:4]:
                vals = []
                for d in (1e-3, -1e-3):
                    g2 = list(gates)
                    kind, q, p = g2[j]
                    g2[j] = (kind, q, (p[0] + d,))
                    s.reset()
                    s.run(C.to_circuit(n, g2))
                    vals.append(s.expectation(h))
                fd_err = max(fd_err, abs((vals[0] - vals[1]) / 2e-3 - gr[j]))
            # <H> alone: read-only co-state tile passes against one FP64
            # reduction per term, on the same state
            s.expectation(h)
            t0 = time.perf_counter()
            e_tiles = s.expectation(h)
            t_tiles = time.perf_counter() - t0
            s.set_pass_merging(False)
            t0 = time.perf_counter()
            e_terms = s.expectation(h)
            t_terms = time.perf_counter() - t0
            s.set_pass_merging(True)
        rec = {"config": f"hea_d3_adjoint_n{n}", "gates": len(gates), "params": P,
               "energy": e, "gpu_s": t, "grads_per_s": 1.0 / t,
               "max_abs_err_vs_central_diff_h1e-3": fd_err,
               "expectation_s": t_tiles, "expectation_per_term_s": t_terms,
               "expectation_tiles_vs_per_term_abs_diff": abs(e_tiles - e_terms)}
        if n == 20:
            O.set_threads(th)
            t0 = time.perf_counter()
            want = O.param_shift(n, gates, terms)
            rec["cpu_param_shift_s"] = time.perf_counter() - t0
            e_ref = O.expect_h(O.run(n, gates), terms)
            O.set_threads(1)
            rec["max_abs_err_vs_fp64_param_shift"] = float(np.abs(gr - want).max())
            rec["energy_err_vs_fp64"] = abs(e - e_ref)
            rec["cpu_threads"] = th
        emit(rec)


def tfim_expectation(qm, C, n, reps):
    """<H> of the TFIM ring (2n terms) on the HEA depth-3 state of the
    adjoint leg: the default path (terms several to a read-only tile pass)
    against one FP64 reduction per term (pass merging off). Each tile pass
    reads the state once, so its algorithmic bytes are 8 * 2^n."""
    rng = C.SplitMix64(0x4EA + n)
    params = [rng.uniform(0, 2 * math.pi) for _ in range(6 * n)]
    gates = C.hardware_efficient(n, 3, params=params)
    terms = C.tfim_ring_hamiltonian(n)
    h = C.to_hamiltonian(terms)
    rec = {"workload": f"tfim_ring_expectation_n{n}", "terms": len(terms)}
    with qm.Simulator(n) as s:
        s.run(C.to_circuit(n, gates))
        for key, merge in (("tiles", True), ("per_term", False)):
            s.set_pass_merging(merge)
            s.expectation(h)
            before = s.launch_count
            e = s.expectation(h)
            launches = s.launch_count - before
            # reads of the state: one per tile pass, or one per term (the
            # per-term reductions of up to 128 terms share a launch)
            reads = launches if merge else len(terms)
            best = None
            for _ in range(reps):
                t0 = time.perf_counter()
                e = s.expectation(h)
                dt = time.perf_counter() - t0
                best = dt if best is None else min(best, dt)
            rec[key] = {"energy": e, "s": best, "launches": launches, "state_reads": reads,
                        "read_GBps": reads * 8.0 * (1 << n) / best / 1e9}
        s.set_pass_merging(True)
    rec["speedup"] = rec["per_term"]["s"] / rec["tiles"]["s"]
    rec["abs_diff"] = abs(rec["tiles"]["energy"] - rec["per_term"]["energy"])
    return rec


# Published ratios of the reference on its own device (Apple M4 Max, n=26):
# fusion speedup per circuit (results/fusion_bench_n26.txt:5-13) and the
# fusion x blocking grid at b=10 (results/ablation_n26_b10_extended.txt:6-15).
REF_FUSION_N26 = {"ghz": 1.01, "qft": 8.42, "qaoa_ring_p6": 6.06, "tfim_trotter_20": 5.16,
                  "grover_proxy_4": 5.84, "random_d10": 1.88, "hea_d3": 3.09}
REF_ABLATION_N26 = {"fusion_alone": 4.01, "blocking_alone": 1.37, "blocking_on_fusion": 1.02}

FROZEN = ("ghz", "qft", "qaoa_ring_p6", "tfim_trotter_20", "grover_proxy_4", "random_d10",
          "hea_d3", "local_patch_w8")


def frozen_streams(n: int):
    """The reference's own gate streams of its extended frozen benchmark set
    (src/circuits.cpp:158-190) at n, as recorded in the committed fixtures
    (tests/golden/frozen_large.json.gz: n = 24, 26, 33)."""
    import gzip
    with gzip.open(os.path.join(ROOT, "tests", "golden", "frozen_large.json.gz"), "rt") as f:
        d = json.load(f)
    out = {}
    for name in FROZEN:
        rec = d.get(f"{name}_{n}")
        if rec is None:
            raise SystemExit(f"bench.py: no recorded frozen stream {name} at n={n} (24, 26, 33)")
        out[name] = [(k, tuple(q[:nq]), tuple(float.fromhex(x) for x in p[:np_]))
                     for k, nq, np_, q, p in rec["gates"]]
    return out


def frozen(qm, C, n: int, reps: int, block_bits: int, emit):
    """The reference's fusion_bench and ablation protocols (bench/fusion_bench.mm,
    bench/ablation.mm) on this GPU: per frozen circuit, plans without fusion
    (base), fused (F), blocking only (B), fused + blocking (FB), each run one
    launch per plan op as the reference's executor does, plus the fused plan
    under this build's pass merging (M, the default qm_run). Arms alternate,
    each timed best of `reps` after a warm-up, twice; one JSON line per
    circuit and one with the geometric means beside the published ones."""
    streams = frozen_streams(n)
    arms = {"base": qm.FusionOptions(0, 0, 0, 1, 0), "F": qm.FusionOptions(1, 1, 1, 4, 0),
            "B": qm.FusionOptions(0, 0, 0, 1, block_bits),
            "FB": qm.FusionOptions(1, 1, 1, 4, block_bits)}
    geo = {"F/base": 0.0, "B/base": 0.0, "FB/F": 0.0, "M/F": 0.0, "M/base": 0.0}
    with qm.Simulator(n) as s:
        def time_arm(c, opts, merge):
            s.set_pass_merging(merge)
            best = math.inf
            for r in range(reps + 1):
                s.reset()
                s.synchronize()
                t0 = time.perf_counter()
                s.run(c, options=opts)
                s.synchronize()
                dt = time.perf_counter() - t0
                if r > 0:
                    best = min(best, dt)
            return best

        for name in FROZEN:
            c = C.to_circuit(n, streams[name])
            t = {}
            for rnd in range(2):
                for arm, opts in arms.items():
                    t[arm] = min(t.get(arm, math.inf), time_arm(c, opts, False))
                    print(f"# {name} round {rnd} {arm}: {t[arm]:.4f} s", file=sys.stderr, flush=True)
                t["M"] = min(t.get("M", math.inf), time_arm(c, arms["F"], True))
            s.set_pass_merging(True)
            rec = {"config": f"frozen_{name}_n{n}", "gates": len(c),
                   "plan_ops": {a: len(c.plan(o)) for a, o in arms.items()},
                   "launches_M": max(c.passes(arms["F"])) + 1,
                   "s": {a: round(v, 6) for a, v in t.items()},
                   "F/base": round(t["base"] / t["F"], 3), "B/base": round(t["base"] / t["B"], 3),
                   "FB/F": round(t["F"] / t["FB"], 3), "M/F": round(t["F"] / t["M"], 3),
                   "M/base": round(t["base"] / t["M"], 3)}
            if n == 26 and name in REF_FUSION_N26:
                rec["reference_F/base_m4max"] = REF_FUSION_N26[name]
            for k in geo:
                geo[k] += math.log(rec[k])
            emit(rec)
    k = len(FROZEN)
    emit({"config": f"frozen_geomean_n{n}", "block_bits": block_bits,
          **{k2: round(math.exp(v / k), 3) for k2, v in geo.items()},
          "reference_m4max_n26": REF_ABLATION_N26})


def device_program_time(qm, s, c, h, params, reps: int):
    """Best wall seconds of qm_energy_and_gradient_device (parameters, energy
    and gradient in device memory, the null stream; from 12 qubits the
    device-resident program, no host synchronisation inside the call), after
    one warm-up call that compiles it. Device memory comes from the HIP
    runtime the library is bound to (no torch in this process)."""
    import ctypes
    from ctypes import POINTER, byref, c_int, c_size_t, c_void_p
    hip = ctypes.CDLL("libamdhip64.so")
    eg = qm._decl("qm_energy_and_gradient_device", c_int, c_void_p, c_void_p,
                  POINTER(qm._Hamiltonian), c_void_p, c_size_t, c_void_p, c_void_p, c_void_p)
    P = len(params)
    bufs = []

    def alloc(nbytes):
        p = c_void_p()
        if hip.hipMalloc(byref(p), c_size_t(nbytes)) != 0:
            raise RuntimeError("hipMalloc failed")
        bufs.append(p)
        return p

    try:
        dp, de, dg = alloc(8 * P), alloc(8), alloc(8 * P)
        host = (ctypes.c_double * P)(*params)
        if hip.hipMemcpy(dp, host, c_size_t(8 * P), 1) != 0:  # host to device
            raise RuntimeError("hipMemcpy failed")
        hc = h._c()

        def call():
            qm._check(eg(s._h, c._h, byref(hc), dp, P, de, dg, None))
            hip.hipDeviceSynchronize()

        call()
        best = math.inf
        for _ in range(reps):
            t0 = time.perf_counter()
            call()
            best = min(best, time.perf_counter() - t0)
        return best
    finally:
        for p in bufs:
            hip.hipFree(p)


def hea_adjoint(qm, C, n: int, reps: int):
    """The hardware-efficient ansatz, depth 3, with the TFIM ring Hamiltonian
    (sum Z_i Z_i+1 + 0.7 sum X_i): wall seconds per adjoint gradient and
    per-kernel device times (events on the simulator's stream), best of
    `reps` after one warm-up."""
    P = 6 * n
    rng = C.SplitMix64(0x4EA + n)
    params = [rng.uniform(0, 2 * math.pi) for _ in range(P)]
    c = C.to_circuit(n, C.hardware_efficient(n, 3, params=params))
    h = C.to_hamiltonian(C.tfim_ring_hamiltonian(n))
    dev_best = None
    with qm.Simulator(n) as s:
        s.energy_and_gradient(c, h)
        best, timings = math.inf, None
        for _ in range(reps):
            s.set_timing(True)
            s.read_timing()
            t0 = time.perf_counter()
            e, _ = s.energy_and_gradient(c, h)
            dt = time.perf_counter() - t0
            t = s.read_timing()
            s.set_timing(False)
            if dt < best:
                best, timings = dt, t
        dev_best = device_program_time(qm, s, c, h, params, reps)
    kinds = {}
    for kind, ms, by in timings:
        k = kinds.setdefault(kind, [0, 0.0, 0.0])
        k[0] += 1
        k[1] += ms
        k[2] += by
    dom = max(kinds, key=lambda k: kinds[k][1])
    cnt, ms_tot, by_tot = kinds[dom]
    achieved = by_tot / (ms_tot * 1e-3) / 1e9
    A = float(1 << n)
    # The reference's per-gradient bytes (gate-path forward, co-state,
    # backward over both states): 3 B_f + 16 A with B_f = (120 n - 24) A.
    ref_bytes = (3 * (120 * n - 24) + 16) * A
    return {
        "workload": f"hea_d3_tfim_ring_n{n}", "gates": 9 * n - 3, "params": P, "energy": e,
        "s_per_gradient": round(best, 4), "grads_per_s": round(1 / best, 3),
        "device_program_s_per_gradient": round(dev_best, 4) if dev_best else None,
        "kernel_ms_per_gradient": round(sum(v[1] for v in kinds.values()), 2),
        "plan_equiv_GBps": round(ref_bytes / best / 1e9, 1),
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": load_traffic(f"hea_d3_tfim_ring_n{n}", dom),
                     "kernels": {k: {"launches": c_, "avg_ms": round(ms / c_, 3),
                                     "GBps": round(by / (ms * 1e-3) / 1e9, 1)}
                                 for k, (c_, ms, by) in sorted(kinds.items())}},
    }


def brickwork_bench(qm, C, n: int, reps: int):
    """Brickwork depth 10 (a Haar U3 on every wire, then CX bricks) from
    |0...0>: seconds per circuit (best of `reps` after a warm-up), launches,
    and per-kernel device times by event."""
    gates = C.brickwork(n, 10)
    c = C.to_circuit(n, gates)
    with qm.Simulator(n) as s:
        s.run(c)
        s.synchronize()
        best, timings = math.inf, None
        for _ in range(reps):
            s.reset()
            s.synchronize()
            s.set_timing(True)
            s.read_timing()
            t0 = time.perf_counter()
            s.run(c)
            s.synchronize()
            dt = time.perf_counter() - t0
            t = s.read_timing()
            s.set_timing(False)
            if dt < best:
                best, timings = dt, t
        norm = s.norm()
    kinds = {}
    for kind, ms, by in timings:
        k = kinds.setdefault(kind, [0, 0.0, 0.0])
        k[0] += 1
        k[1] += ms
        k[2] += by
    dom = max(kinds, key=lambda k: kinds[k][1])
    cnt, ms_tot, by_tot = kinds[dom]
    achieved = by_tot / (ms_tot * 1e-3) / 1e9
    total_bytes = sum(v[2] for v in kinds.values())
    return {"workload": f"brickwork_d10_n{n}", "gates": len(gates),
            "plan_ops": c.plan_size(True), "launches": len(timings), "s_per_circuit": best,
            "kernel_ms": round(sum(v[1] for v in kinds.values()), 3),
            "hbm_GBps": round(total_bytes / best / 1e9, 1), "norm": norm,
            "roofline": {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1),
                         "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": round(achieved / HBM_PEAK_GBS, 4),
                         "traffic": load_traffic(f"brickwork_d10_n{n}", dom),
                         "kernels": {k: {"launches": c_, "avg_ms": round(ms / c_, 3),
                                         "GBps": round(by / (ms * 1e-3) / 1e9, 1)}
                                     for k, (c_, ms, by) in sorted(kinds.items())}}}


def load_traffic(workload: str, kind: str):
    """HBM bytes per launch of the workload's `kind` kernel, measured with
    rocprofv3 PMC counters (FETCH_SIZE, WRITE_SIZE) and kept in
    profiles/traffic.json; None when not measured."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            entry = json.load(f).get(workload)
    except (OSError, ValueError):
        return None
    return entry.get(kind) if isinstance(entry, dict) else None


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=33)
    ap.add_argument("--cpu-n", type=int, default=26)
    ap.add_argument("--cpu-threads", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--unfused", action="store_true", help="also time the unfused plan")
    ap.add_argument("--no-merge", action="store_true",
                    help="one kernel launch per plan op (executor-level pass merging off)")
    ap.add_argument("--batch-count", type=int, default=4096,
                    help="batched QAOA instances (0 skips the batched measurement)")
    ap.add_argument("--batch-repeats", type=int, default=3)
    ap.add_argument("--plain-count", type=int, default=4096,
                    help="instances solved one circuit per launch for the comparison")
    ap.add_argument("--adjoint-n", type=int, default=32,
                    help="width of the HEA depth-3 adjoint leg (0 skips it)")
    ap.add_argument("--no-cpu-n30", action="store_true",
                    help="skip the reference oracle's n=30 arm of the CPU baseline")
    ap.add_argument("--suite", action="store_true",
                    help="measure every configuration (one JSON line each) instead of the headline")
    ap.add_argument("--scale", type=int, nargs=2, default=None, metavar=("NMIN", "NMAX"),
                    help="the reference's scale sweep: GHZ and QFT at every width with invariants")
    ap.add_argument("--frozen", type=int, default=0, metavar="N",
                    help="run the reference's fusion_bench / ablation protocols on its frozen "
                         "benchmark set at n=N (24, 26 or 33) instead of the headline")
    ap.add_argument("--frozen-reps", type=int, default=2)
    ap.add_argument("--adjoint-only", action="store_true",
                    help="only time the HEA depth-3 adjoint at n=--adjoint-n, per kernel")
    ap.add_argument("--expectation", type=int, default=0, metavar="N",
                    help="only time <H> of the TFIM ring on the HEA depth-3 state at n=N: "
                         "read-only tile passes against one reduction per term")
    ap.add_argument("--batch-only", action="store_true",
                    help="only run the batched QAOA measurement")
    ap.add_argument("--brickwork", type=int, default=0, metavar="N",
                    help="only time brickwork depth 10 at n=N, per kernel, instead of the headline")
    ap.add_argument("--block-bits", type=int, default=10)
    ap.add_argument("--copy-probe", type=int, default=0, metavar="N",
                    help="only run a torch device copy of 2^N complex64 (8 * 2^N bytes read "
                         "and written) three times: the known-size kernel that calibrates the "
                         "FETCH_SIZE / WRITE_SIZE counters")
    args = ap.parse_args()
    if args.copy_probe:
        import torch
        src = torch.ones(1 << args.copy_probe, dtype=torch.complex64, device="cuda")
        dst = torch.empty_like(src)
        best = None
        for _ in range(3):
            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            a.record()
            dst.copy_(src)
            b.record()
            b.synchronize()
            ms = a.elapsed_time(b)
            best = ms if best is None else min(best, ms)
        nbytes = 8 << args.copy_probe
        print(f"copy probe: 3 x {nbytes} bytes read and written; best {best:.3f} ms = "
              f"{2 * nbytes / (best * 1e-3) / 1e9:.1f} GB/s (read + write)", flush=True)
        return 0
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported",
              file=sys.stderr)
        return 2

    # No torch here: use the ROCm runtime libqmetal.so was linked against.
    os.environ.setdefault("QMETAL_HIP_RUNTIME", "system")
    import qmetal_amd as qm
    from qmetal_amd import circuits as C

    if args.scale:
        print(f"# device {qm.device_name()}  scale {args.scale[0]}..{args.scale[1]}",
              file=sys.stderr, flush=True)
        scale_sweep(qm, C, args.scale[0], args.scale[1],
                    lambda rec: print(json.dumps(rec), flush=True))
        return 0
    if args.suite:
        print(f"# device {qm.device_name()}  suite", file=sys.stderr, flush=True)
        suite(qm, C, lambda rec: print(json.dumps(rec), flush=True))
        return 0
    if args.adjoint_only:
        print(f"# device {qm.device_name()}  HEA d3 adjoint n={args.adjoint_n}", file=sys.stderr,
              flush=True)
        print(json.dumps(hea_adjoint(qm, C, args.adjoint_n, 2)), flush=True)
        return 0
    if args.expectation:
        print(f"# device {qm.device_name()}  TFIM ring <H> n={args.expectation}", file=sys.stderr,
              flush=True)
        print(json.dumps(tfim_expectation(qm, C, args.expectation, 3)), flush=True)
        return 0
    if args.batch_only:
        print(f"# device {qm.device_name()}  batched QAOA x{args.batch_count}", file=sys.stderr,
              flush=True)
        print(json.dumps(batched_qaoa(qm, C, args.batch_count, args.batch_repeats,
                                      args.plain_count)), flush=True)
        return 0
    if args.brickwork:
        print(f"# device {qm.device_name()}  brickwork d10 n={args.brickwork}", file=sys.stderr,
              flush=True)
        print(json.dumps(brickwork_bench(qm, C, args.brickwork, 3)), flush=True)
        return 0
    if args.frozen:
        print(f"# device {qm.device_name()}  frozen set n={args.frozen}", file=sys.stderr, flush=True)
        frozen(qm, C, args.frozen, args.frozen_reps, args.block_bits,
               lambda rec: print(json.dumps(rec), flush=True))
        return 0

    n = args.n
    workload = f"qft_n{n}_product_state"
    print(f"# device {qm.device_name()}  workload {workload}", file=sys.stderr, flush=True)
    sim = qm.Simulator(n)
    sim.set_pass_merging(not args.no_merge)
    place_attempts, place_ratio = sim.placement
    print(f"# state placement: {place_attempts} allocation(s), top/mid pass time ratio "
          f"{place_ratio:.3f}", file=sys.stderr, flush=True)
    sim.run(C.to_circuit(n, C.product_state_prep(n, seed=2024)))
    qft = C.to_circuit(n, C.qft(n))
    plan_ops = qft.plan_size(True)
    launches = plan_ops if args.no_merge else max(qft.passes()) + 1
    for _ in range(args.warmup):
        sim.run(qft)
    sim.synchronize()

    sim.set_timing(True)
    sim.read_timing()
    sim.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        sim.run(qft)
    sim.synchronize()
    t1 = time.perf_counter()
    timings = sim.read_timing()
    sim.set_timing(False)
    s_per = (t1 - t0) / args.steps
    alg = qft_bytes(n)
    norm_after = sim.norm()

    # Per-kernel roofline from the event timings.
    kinds = {}
    for kind, ms, by in timings:
        k = kinds.setdefault(kind, [0, 0.0, 0.0])
        k[0] += 1
        k[1] += ms
        k[2] += by
    dom = max(kinds, key=lambda k: kinds[k][1])
    cnt, ms_tot, by_tot = kinds[dom]
    achieved = (by_tot / cnt) / (ms_tot / cnt * 1e-3) / 1e9
    kern_ms = sum(v[1] for v in kinds.values()) / args.steps
    # Bytes one circuit moves through HBM: per launch the PMC-measured
    # traffic (FETCH_SIZE + WRITE_SIZE, profiles/traffic.json) where it was
    # measured for this workload, else the kernel's own pass bytes.
    hbm_step, measured = 0.0, []
    for k, (c_, ms_, by_) in kinds.items():
        tr = load_traffic(workload, k)
        if tr is not None:
            measured.append(k)
            hbm_step += tr * c_ / args.steps
        else:
            hbm_step += by_ / args.steps
    hbm_src = ("PMC counters per launch" if len(measured) == len(kinds) else
               f"PMC counters for {measured}, kernel pass bytes for the rest")
    print(f"# {args.steps} steps: {s_per:.4f} s/circuit, kernels {kern_ms:.1f} ms/circuit, "
          f"norm after {norm_after:.8f}", file=sys.stderr, flush=True)
    for k, (c, ms, by) in sorted(kinds.items()):
        print(f"#   {k:14s} {c:4d} launches  avg {ms / c:8.3f} ms  "
              f"{by / c / 1e9:7.2f} GB/launch  {by / (ms * 1e-3) / 1e9:7.1f} GB/s",
              file=sys.stderr, flush=True)

    # Invariants on |0...0>: QFT maps it to the uniform state.
    sim.reset()
    sim.run(qft)
    worst_z = max(abs(sim.expectation_pauli(z=1 << q)) for q in range(0, n, max(1, n // 8)))
    norm0 = sim.norm()
    print(f"# QFT|0>: norm {norm0:.8f}  max|<Z_q>| {worst_z:.2e}", file=sys.stderr, flush=True)

    # The same plan, one launch per plan op, for comparison (one circuit).
    op_by_op = None
    if not args.no_merge:
        sim.set_pass_merging(False)
        sim.run(qft)
        sim.synchronize()
        o0 = time.perf_counter()
        sim.run(qft)
        sim.synchronize()
        op_by_op = time.perf_counter() - o0
        sim.set_pass_merging(True)
        print(f"# one launch per plan op: {op_by_op:.3f} s/circuit", file=sys.stderr, flush=True)

    unfused = None
    if args.unfused:
        sim.run(qft, fuse=False)
        sim.synchronize()
        u0 = time.perf_counter()
        sim.run(qft, fuse=False)
        sim.synchronize()
        unfused = time.perf_counter() - u0
        print(f"# unfused: {unfused:.3f} s/circuit", file=sys.stderr, flush=True)
    sim.close()

    batched = None
    if args.batch_count:
        try:
            batched = batched_qaoa(qm, C, args.batch_count, args.batch_repeats, args.plain_count)
            print(f"# batched QAOA x{args.batch_count}: {batched['batched_s']:.3f} s vs plain "
                  f"{batched['plain_s']:.2f} s ({batched['speedup']}x; one launch per gate "
                  f"{batched['plain_per_gate_s']:.2f} s, {batched['speedup_vs_per_gate']}x), vs plain: "
                  f"gradients {batched['max_abs_grad_diff_vs_plain']:.2e} abs, energies "
                  f"{batched['max_rel_energy_diff_vs_plain']:.2e} rel, config {batched['autotuned']}",
                  file=sys.stderr, flush=True)
        except Exception as exc:  # report, do not lose the headline line
            batched = {"error": str(exc)}

    adjoint = None
    if args.adjoint_n:
        try:
            adjoint = hea_adjoint(qm, C, args.adjoint_n, 2)
            r = adjoint["roofline"]
            print(f"# HEA d3 adjoint n={args.adjoint_n}: {adjoint['s_per_gradient']:.3f} s per "
                  f"gradient; {r['kernel']} {r['achieved']:.0f} GB/s = {r['frac']:.3f} of peak",
                  file=sys.stderr, flush=True)
        except Exception as exc:  # report, do not lose the headline line
            adjoint = {"error": str(exc)}

    cpu = None
    if not args.no_cpu:
        sample = (f"QFT n={args.cpu_n} gate by gate ({args.cpu_n * (args.cpu_n + 1) // 2} "
                  f"gates) on the same seeded product state, complex128")
        ref, ref_error = None, None
        if args.cpu_threads == 1:
            try:
                ref = ref_cpu_baseline(args.cpu_n, n)
            except Exception as exc:  # report it and fall back to the restatement
                ref_error = f"reference oracle leg failed: {exc}"
                print(f"# {ref_error}", file=sys.stderr, flush=True)
        dt, extrap = cpu_baseline(args.cpu_n, n, args.cpu_threads)
        port = {"value": round(extrap, 1), "unit": "s/circuit", "cores": args.cpu_threads,
                "kind": "port", "sample": f"{sample} oracle restatement (oracle/qm_oracle.c): "
                                          f"{dt:.2f} s measured, scaled to n={n} by amplitude-gates",
                "sample_s": round(dt, 3)}
        if ref is not None:
            cpu = {"value": round(ref[1], 1), "unit": "s/circuit", "cores": 1, "kind": "reference",
                   "sample": f"{sample} reference oracle (src/reference.cpp, built from its "
                             f"sources into oracle/_ref): {ref[0]:.2f} s measured, scaled to "
                             f"n={n} by amplitude-gates",
                   "sample_s": round(ref[0], 3), "port": port}
            print(f"# cpu reference oracle n={args.cpu_n}: {ref[0]:.3f} s (scaled to n={n}: "
                  f"{ref[1]:.1f} s)", file=sys.stderr, flush=True)
        else:
            cpu = port
            if ref_error:
                cpu["reference_error"] = ref_error
        print(f"# cpu oracle n={args.cpu_n}: {dt:.3f} s (scaled to n={n}: {extrap:.1f} s)",
              file=sys.stderr, flush=True)
        # The threaded arm: the restatement over every host core this
        # process may use (the reference's oracle has no threading).
        th = oracle_threads()
        cpu["host"] = host_cpu()
        if th > 1 and args.cpu_threads == 1:
            dt_t, extrap_t = cpu_baseline(args.cpu_n, n, th)
            cpu["threaded"] = {"value": round(extrap_t, 1), "unit": "s/circuit", "cores": th,
                               "kind": "port", "sample_s": round(dt_t, 3),
                               "sample": f"the same sample on {th} OpenMP threads (all cores "
                                         f"the process may use)"}
            print(f"# cpu oracle n={args.cpu_n}, {th} threads: {dt_t:.3f} s "
                  f"(scaled to n={n}: {extrap_t:.1f} s)", file=sys.stderr, flush=True)
        if not args.no_cpu_n30:
            try:
                r30 = ref_cpu_n30(n)
            except Exception as exc:
                r30 = None
                cpu["n30_error"] = str(exc)
            if r30 is not None:
                cpu["n30"] = {"value": round(r30[1], 1), "unit": "s/circuit", "cores": 1,
                              "kind": "reference", "sample_s": round(r30[0], 3),
                              "sample": "reference oracle at its n=30 ceiling: QFT(30)'s first "
                                        "stage (H + 29 CP) from |0...0>, scaled to the n="
                                        f"{n} circuit by amplitudes x gates"}
                print(f"# cpu reference oracle n=30, 30 gates: {r30[0]:.3f} s (scaled to "
                      f"n={n}: {r30[1]:.1f} s)", file=sys.stderr, flush=True)

    out = {
        "metric": "s/circuit + HBM GB/s (% of peak) at n=30-34",
        "value": round(s_per, 5),
        "unit": "s/circuit",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(s_per * 1e3, 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "complex64",
        "data": "synthetic",
        "config": {"workload": workload, "n": n, "gates": n * (n + 1) // 2,
                   "plan_ops": plan_ops, "fused": True, "launches": launches,
                   "pass_merging": not args.no_merge},
        "state_placement": {"allocations": place_attempts, "top_mid_ratio": round(place_ratio, 4)},
        "hbm_GBps": round(hbm_step / s_per / 1e9, 1),
        "hbm_bytes_per_circuit": hbm_step,
        "hbm_bytes_source": hbm_src,
        "plan_equiv_GBps": round(alg / s_per / 1e9, 1),
        "plan_equiv_x_peak": round(alg / s_per / 1e9 / HBM_PEAK_GBS, 2),
        "invariants": {"norm_after_steps": norm_after, "qft_zero_norm": norm0,
                       "qft_zero_max_abs_z": worst_z},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": load_traffic(workload, dom),
                     "kernels": {k: {"launches": c // args.steps,
                                     "avg_ms": round(ms / c, 3),
                                     "GBps": round(by / (ms * 1e-3) / 1e9, 1)}
                                 for k, (c, ms, by) in sorted(kinds.items())}},
        "cpu_baseline": cpu,
    }
    if op_by_op is not None:
        out["op_by_op_s_per_circuit"] = round(op_by_op, 4)
    if unfused is not None:
        out["unfused_s_per_circuit"] = round(unfused, 4)
    if batched is not None:
        out["batched_adjoint"] = batched
    if adjoint is not None:
        out["adjoint"] = adjoint
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
