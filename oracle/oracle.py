"""ctypes wrapper of the complex128 C oracle (oracle/qm_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py, as the checker. The product never imports it.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from ctypes import POINTER, c_double, c_int, c_int32, c_size_t, c_uint8, c_uint32, c_uint64

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_build", "liboracle.so")


def build() -> str:
    subprocess.run(["make", "-s", "-C", HERE, "oracle"], check=True)
    return LIB


def _load():
    if not os.path.exists(LIB):
        build()
    return ctypes.CDLL(LIB)


_lib = _load()
_f64p, _u64p, _u32p = POINTER(c_double), POINTER(c_uint64), POINTER(c_uint32)


def _d(name, res, *args):
    f = getattr(_lib, name)
    f.restype = res
    f.argtypes = list(args)
    return f


_set_threads = _d("qo_set_threads", None, c_int)
_gate_matrix = _d("qo_gate_matrix", c_int, c_int, _f64p, _f64p)
_reset = _d("qo_reset", None, _f64p, c_uint32)
_apply_1q = _d("qo_apply_1q", None, _f64p, c_uint32, c_uint32, _f64p)
_apply_ctrl = _d("qo_apply_controlled", None, _f64p, c_uint32, c_uint32, c_uint32, _f64p)
_apply_swap = _d("qo_apply_swap", None, _f64p, c_uint32, c_uint32, c_uint32)
_apply_ccx = _d("qo_apply_ccx", None, _f64p, c_uint32, c_uint32, c_uint32, c_uint32)
_run = _d("qo_run", c_int, _f64p, c_uint32, c_size_t, POINTER(c_int32), POINTER(c_uint8),
          _u32p, _f64p)
_apply_group = _d("qo_apply_group", None, _f64p, c_uint32, c_size_t, _u32p, _f64p)
_apply_diag = _d("qo_apply_diag", None, _f64p, c_uint32, c_size_t, _u64p, _f64p, c_double)
_norm = _d("qo_norm", c_double, _f64p, c_uint32)
_expect = _d("qo_expect_pauli", c_double, _f64p, c_uint32, c_uint64, c_uint64, c_uint64)
_splitmix = _d("qo_splitmix64", c_uint64, _u64p)
_sample = _d("qo_sample", None, _f64p, c_uint32, c_size_t, c_uint64, _u64p)
_adjoint = _d("qo_adjoint", c_int, c_uint32, c_size_t, POINTER(c_int32), POINTER(c_uint8),
              POINTER(c_uint8), _u32p, _f64p, c_uint32, _f64p, _u64p, _u64p, _u64p, _f64p,
              _f64p, _f64p)


def _p(a, t):
    return a.ctypes.data_as(POINTER(t))


def set_threads(k: int) -> None:
    _set_threads(int(k))


def gate_matrix(kind: int, params=(0.0, 0.0, 0.0)) -> np.ndarray:
    p = np.zeros(3)
    p[:len(params)] = params
    m = np.zeros(8)
    if _gate_matrix(kind, _p(p, c_double), _p(m, c_double)) != 0:
        raise ValueError("not a one-qubit gate")
    return m.view(np.complex128).reshape(2, 2)


def zero_state(n: int) -> np.ndarray:
    s = np.zeros(1 << n, dtype=np.complex128)
    s[0] = 1.0
    return s


def _pack(gates):
    g = len(gates)
    kinds = np.zeros(g, dtype=np.int32)
    nq = np.zeros(g, dtype=np.uint8)
    qs = np.zeros(3 * g, dtype=np.uint32)
    ps = np.zeros(3 * g, dtype=np.float64)
    for k, (kind, q, p) in enumerate(gates):
        kinds[k] = kind
        nq[k] = len(q)
        qs[3 * k:3 * k + len(q)] = q
        ps[3 * k:3 * k + len(p)] = p
    return kinds, nq, qs, ps


def run(n: int, gates, state: np.ndarray | None = None) -> np.ndarray:
    """Apply a gate stream [(kind, qubits, params), ...] to |0...0> (or `state`)."""
    s = zero_state(n) if state is None else np.ascontiguousarray(state, dtype=np.complex128).copy()
    kinds, nq, qs, ps = _pack(gates)
    rc = _run(_p(s, c_double), n, len(gates), _p(kinds, c_int32), _p(nq, c_uint8), _p(qs, c_uint32),
              _p(ps, c_double))
    if rc != 0:
        raise ValueError(f"oracle rejected gate {-1 - rc}")
    return s


def apply_plan(n: int, plan, state: np.ndarray | None = None) -> np.ndarray:
    """Execute plan ops (dicts as qmetal_amd.Circuit.plan returns) in complex128."""
    s = zero_state(n) if state is None else np.ascontiguousarray(state, dtype=np.complex128).copy()
    sp = _p(s, c_double)
    for op in plan:
        k = op["kind"]
        if k == "Dense1q":
            m = np.array(op["matrix"], dtype=np.complex128)
            _apply_1q(sp, n, op["qubits"][0], _p(m, c_double))
        elif k == "Controlled1q":
            m = np.array(op["matrix"], dtype=np.complex128)
            _apply_ctrl(sp, n, op["qubits"][0], op["qubits"][1], _p(m, c_double))
        elif k in ("Dense1qGroup", "Blocked1q"):
            q = np.array(op["group_qubits"], dtype=np.uint32)
            m = np.array(op["group_matrices"], dtype=np.complex128)
            _apply_group(sp, n, len(q), _p(q, c_uint32), _p(m, c_double))
        elif k == "DiagonalLayer":
            mk = np.array(op["masks"], dtype=np.uint64)
            an = np.array(op["angles"], dtype=np.float64)
            _apply_diag(sp, n, len(mk), _p(mk, c_uint64), _p(an, c_double), op["global_phase"])
        elif k == "Swap":
            _apply_swap(sp, n, op["qubits"][0], op["qubits"][1])
        elif k == "CCX":
            _apply_ccx(sp, n, *op["qubits"])
        else:
            raise ValueError(k)
    return s


def norm(state: np.ndarray) -> float:
    n = int(np.log2(state.size))
    return float(_norm(_p(np.ascontiguousarray(state, dtype=np.complex128), c_double), n))


def expect(state: np.ndarray, x: int = 0, y: int = 0, z: int = 0) -> float:
    s = np.ascontiguousarray(state, dtype=np.complex128)
    n = int(np.log2(s.size))
    return float(_expect(_p(s, c_double), n, x, y, z))


def expect_h(state: np.ndarray, terms) -> float:
    return sum(c * expect(state, x, y, z) for c, x, y, z in terms)


def splitmix64_stream(seed: int, count: int) -> list[int]:
    st = c_uint64(seed)
    return [int(_splitmix(ctypes.byref(st))) for _ in range(count)]


def sample(state: np.ndarray, shots: int, seed: int) -> np.ndarray:
    s = np.ascontiguousarray(state, dtype=np.complex128)
    n = int(np.log2(s.size))
    out = np.zeros(shots, dtype=np.uint64)
    _sample(_p(s, c_double), n, shots, seed, _p(out, c_uint64))
    return out


def adjoint(n: int, gates, terms):
    """(energy, grad) from the complex128 adjoint sweep."""
    kinds, nq, qs, ps = _pack(gates)
    npar = np.array([len(p) for _, _, p in gates] or [0], dtype=np.uint8)
    t = len(terms)
    co = np.array([x[0] for x in terms], dtype=np.float64)
    xm = np.array([x[1] for x in terms], dtype=np.uint64)
    ym = np.array([x[2] for x in terms], dtype=np.uint64)
    zm = np.array([x[3] for x in terms], dtype=np.uint64)
    work = np.zeros(4 << n)
    e = np.zeros(1)
    g = np.zeros(max(len(gates), 1))
    rc = _adjoint(n, len(gates), _p(kinds, c_int32), _p(nq, c_uint8), _p(npar, c_uint8),
                  _p(qs, c_uint32),
                  _p(ps, c_double), t, _p(co, c_double), _p(xm, c_uint64), _p(ym, c_uint64),
                  _p(zm, c_uint64), _p(work, c_double), _p(e, c_double), _p(g, c_double))
    if rc != 0:
        raise ValueError("adjoint: non-differentiable parameterised gate")
    return float(e[0]), g[:len(gates)].copy()


def param_shift(n: int, gates, terms):
    """Exact parameter-shift gradient for RX/RY/RZ carrying a parameter:
    [E(t+pi/2) - E(t-pi/2)] / 2."""
    grad = np.zeros(len(gates))
    for k, (kind, q, p) in enumerate(gates):
        if kind not in (8, 9, 10) or not p:
            continue
        vals = []
        for d in (np.pi / 2, -np.pi / 2):
            g2 = list(gates)
            g2[k] = (kind, q, (p[0] + d,) + tuple(p[1:]))
            vals.append(expect_h(run(n, g2), terms))
        grad[k] = 0.5 * (vals[0] - vals[1])
    return grad
