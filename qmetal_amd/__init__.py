"""Python bindings for the MI355X statevector simulator.

A ctypes wrapper over the C ABI in include/qmetal/qmetal.h, with the classes
and methods of the reference's Python package (python/qmetal/__init__.py):
``Circuit``, ``Simulator``, ``Hamiltonian``, ``QMetalError``, ``version``,
``is_available``, ``device_name``, ``max_qubits``. ``import qmetal_amd as
qmetal`` is a drop-in for that path.

    import qmetal_amd as qmetal
    c = qmetal.Circuit(4)
    c.h(0)
    for i in range(3):
        c.cx(i, i + 1)
    sim = qmetal.Simulator(4)
    sim.run(c)
    print(sim.counts(1000))

Differences from the reference, all forced by discrete HBM or appended:
``amplitudes()`` returns a host copy (numpy complex64) refreshed per call, not
a zero-copy view; ``max_qubits()`` is 34; ``Circuit.u3`` and
``Simulator.run(..., options=...)`` / ``Circuit.plan(...)`` are appended.
"""

from __future__ import annotations

import ctypes
import os
from ctypes import (POINTER, Structure, byref, c_char_p, c_double, c_float,
                    c_int, c_size_t, c_uint32, c_uint64, c_void_p)

__all__ = ["Circuit", "Simulator", "Hamiltonian", "QMetalError", "FusionOptions", "Batch",
           "version", "is_available", "device_name", "max_qubits", "library_path"]


class QMetalError(RuntimeError):
    """Raised when the native library reports a failure."""


def _find_library() -> str:
    env = os.environ.get("QMETAL_LIBRARY")
    if env:
        return env
    here = os.path.dirname(os.path.abspath(__file__))
    path = os.path.join(here, "libqmetal.so")
    if os.path.exists(path):
        return path
    raise QMetalError("libqmetal.so not found. Build it with `python -m qmetal_amd.build`, "
                      "or point QMETAL_LIBRARY at it.")


def _share_hip_runtime():
    """One process holds one HIP runtime. PyTorch-ROCm resolves its HIP
    runtime from its own lib directory, so when torch is installed its
    runtime is loaded first and libqmetal.so binds to the same one (matching
    soname) whether torch is imported before or after this package; device
    pointers and streams are then shared. QMETAL_HIP_RUNTIME=system keeps the
    runtime libqmetal.so was linked against (ROCm's), =torch requires torch's.
    """
    mode = os.environ.get("QMETAL_HIP_RUNTIME", "auto")
    if mode == "system":
        return None
    import importlib.util
    try:
        spec = importlib.util.find_spec("torch")
    except (ImportError, ValueError):
        spec = None
    locs = list(spec.submodule_search_locations or []) if spec else []
    hip = os.path.join(locs[0], "lib", "libamdhip64.so") if locs else ""
    if not hip or not os.path.exists(hip):
        if mode == "torch":
            raise QMetalError("QMETAL_HIP_RUNTIME=torch but torch's HIP runtime was not found")
        return None
    return ctypes.CDLL(hip, mode=ctypes.RTLD_GLOBAL)


_hip_runtime = _share_hip_runtime()
_path = _find_library()
_lib = ctypes.CDLL(_path)


def library_path() -> str:
    return _path


def _decl(name, restype, *argtypes):
    fn = getattr(_lib, name)
    fn.restype = restype
    fn.argtypes = list(argtypes)
    return fn


_u64p, _f64p, _u32p = POINTER(c_uint64), POINTER(c_double), POINTER(c_uint32)


class _Hamiltonian(Structure):
    _fields_ = [("num_terms", c_uint32), ("coeffs", _f64p), ("x_masks", _u64p),
                ("y_masks", _u64p), ("z_masks", _u64p)]


class FusionOptions(Structure):
    """qm_fusion_options: the reference's FusionOptions (plan.h:76-94)."""
    _fields_ = [("collapse_1q_windows", c_int), ("fuse_diagonals", c_int),
                ("recognize_zz", c_int), ("group_1q", c_uint32), ("block_bits", c_uint32)]

    @classmethod
    def default(cls):
        return cls(1, 1, 1, 4, 0)

    @classmethod
    def none(cls):
        return cls(0, 0, 0, 1, 0)

    @classmethod
    def blocked(cls, bits: int):
        return cls(1, 1, 1, 4, int(bits))


class _PlanOp(Structure):
    _fields_ = [("kind", c_uint32), ("qubits", c_uint32 * 3), ("matrix", c_double * 8),
                ("global_phase", c_double), ("block_bits", c_uint32),
                ("num_terms", c_size_t), ("num_group", c_size_t)]


_version = _decl("qm_version", None, POINTER(c_int), POINTER(c_int), POINTER(c_int))
_last_error = _decl("qm_last_error", c_char_p)
_is_available = _decl("qm_is_available", c_int)
_device_name = _decl("qm_device_name", c_char_p)
_max_qubits = _decl("qm_max_qubits", c_uint32)
_max_buffer = _decl("qm_max_buffer_length", c_uint64)

_circuit_create = _decl("qm_circuit_create", c_void_p, c_uint32)
_circuit_destroy = _decl("qm_circuit_destroy", None, c_void_p)
_circuit_add = _decl("qm_circuit_add", c_int, c_void_p, c_int, _u32p, c_uint32, _f64p, c_uint32)
_circuit_size = _decl("qm_circuit_size", c_size_t, c_void_p)
_plan_size = _decl("qm_circuit_plan_size", c_size_t, c_void_p, c_int)

_create = _decl("qm_create", c_void_p, c_uint32)
_destroy = _decl("qm_destroy", None, c_void_p)
_reset = _decl("qm_reset", c_int, c_void_p)
_run = _decl("qm_run", c_int, c_void_p, c_void_p)
_run_unfused = _decl("qm_run_unfused", c_int, c_void_p, c_void_p)
_run_options = _decl("qm_run_options", c_int, c_void_p, c_void_p, POINTER(FusionOptions))
_sync = _decl("qm_synchronize", c_int, c_void_p)
_num_qubits = _decl("qm_num_qubits", c_uint32, c_void_p)
_dim = _decl("qm_dim", c_uint64, c_void_p)
_dispatch_count = _decl("qm_dispatch_count", c_uint64, c_void_p)
_launch_count = _decl("qm_launch_count", c_uint64, c_void_p)
_amplitudes = _decl("qm_amplitudes", POINTER(c_float), c_void_p)
_amplitudes_at = _decl("qm_amplitudes_at", c_int, c_void_p, _u64p, c_size_t, POINTER(c_float))
_norm = _decl("qm_norm", c_int, c_void_p, _f64p)
_expect_pauli = _decl("qm_expectation_pauli", c_int, c_void_p, c_uint64, c_uint64, c_uint64, _f64p)
_expect = _decl("qm_expectation", c_int, c_void_p, POINTER(_Hamiltonian), _f64p)
_sample = _decl("qm_sample", c_int, c_void_p, c_size_t, c_uint64, _u64p)
_energy_grad = _decl("qm_energy_and_gradient", c_int, c_void_p, c_void_p,
                     POINTER(_Hamiltonian), _f64p, _f64p)

_set_timing = _decl("qm_set_timing", c_int, c_void_p, c_int)
_read_timing = _decl("qm_read_timing", c_int, c_void_p, _u32p, _f64p, _f64p, c_size_t,
                     POINTER(c_size_t))

_batch_create = _decl("qm_batch_create", c_void_p)
_batch_destroy = _decl("qm_batch_destroy", None, c_void_p)
_batch_prepare = _decl("qm_batch_prepare", c_int, c_void_p, c_size_t, POINTER(c_void_p),
                       POINTER(_Hamiltonian))
_batch_autotune = _decl("qm_batch_autotune", c_int, c_void_p, _u32p, _u32p, _u32p)
_batch_set_config = _decl("qm_batch_set_config", c_int, c_void_p, c_uint32, c_uint32, c_uint32)
_batch_run = _decl("qm_batch_run", c_int, c_void_p, _f64p, _f64p, _f64p)
_batch_set_spread = _decl("qm_batch_set_spread", c_int, c_void_p, c_uint32)
_batch_spread = _decl("qm_batch_spread", c_uint32, c_void_p)
_batch_batched = _decl("qm_batch_batched_instances", c_int, c_void_p, POINTER(c_size_t))

_plan_build = _decl("qm_plan_build", c_void_p, c_void_p, POINTER(FusionOptions))
_plan_destroy = _decl("qm_plan_destroy", None, c_void_p)
_plan_nops = _decl("qm_plan_size", c_size_t, c_void_p)
_plan_get = _decl("qm_plan_get", c_int, c_void_p, c_size_t, POINTER(_PlanOp))
_plan_arrays = _decl("qm_plan_get_arrays", c_int, c_void_p, c_size_t, _u32p, _f64p, _u64p, _f64p)
_plan_passes = _decl("qm_plan_passes", c_int, c_void_p, _u32p, POINTER(c_size_t))
_plan_schedule = _decl("qm_plan_schedule", c_int, c_void_p, _u32p, POINTER(c_uint64), _u32p,
                       c_size_t, POINTER(c_size_t))
_set_pass_merging = _decl("qm_set_pass_merging", c_int, c_void_p, c_int)
_state_placement = _decl("qm_state_placement", c_int, c_void_p, _u32p, POINTER(c_double))


def _err() -> str:
    return _last_error().decode("utf-8", "replace")


def _check(rc: int) -> None:
    if rc != 0:
        raise QMetalError(_err())


# Gate codes: qm_gate_kind in qmetal.h.
_X, _Y, _Z, _H, _S, _SDG, _T, _TDG = range(8)
_RX, _RY, _RZ, _P = 8, 9, 10, 11
_CX, _CY, _CZ, _SWAP, _CP, _CCX = 12, 13, 14, 15, 16, 17
_U3 = 18

PLAN_KINDS = ("Dense1q", "Dense1qGroup", "Blocked1q", "Controlled1q", "DiagonalLayer",
              "Swap", "CCX")
# Timing records: plan-op kinds, then merged passes of several plan ops
# (generic pass kernel, chain kernel, wide and low chain kernels, generic
# low tile and generic wide kernels), then the adjoint gradient's kernels
# (merged backward passes over both states, co-state build, energy overlap,
# one-launch-per-rotation steps).
TIMING_KINDS = PLAN_KINDS + ("Pass", "ChainPass", "WidePass", "LowPass", "TilePass", "GWidePass",
                             "AdjPass", "CoState", "Overlap", "AdjRot")


def version() -> tuple[int, int, int]:
    a, b, c = c_int(), c_int(), c_int()
    _version(byref(a), byref(b), byref(c))
    return a.value, b.value, c.value


def is_available() -> bool:
    return bool(_is_available())


def device_name() -> str:
    return _device_name().decode()


def max_qubits() -> int:
    return int(_max_qubits())


class Hamiltonian:
    """A real-weighted sum of Pauli strings, given as bitmasks."""

    def __init__(self):
        self.terms: list[tuple[float, int, int, int]] = []

    def add(self, coeff: float, x_mask: int = 0, y_mask: int = 0,
            z_mask: int = 0) -> "Hamiltonian":
        self.terms.append((float(coeff), int(x_mask), int(y_mask), int(z_mask)))
        return self

    def add_string(self, coeff: float, s: str) -> "Hamiltonian":
        """`s` is "IXYZ"-style, leftmost character is qubit 0."""
        x = y = z = 0
        for q, ch in enumerate(s.upper()):
            bit = 1 << q
            if ch == "X":
                x |= bit
            elif ch == "Y":
                y |= bit
            elif ch == "Z":
                z |= bit
            elif ch != "I":
                raise ValueError(f"bad Pauli character {ch!r}")
        return self.add(coeff, x, y, z)

    def _c(self):
        n = len(self.terms)
        coeffs = (c_double * n)(*[t[0] for t in self.terms])
        xs = (c_uint64 * n)(*[t[1] for t in self.terms])
        ys = (c_uint64 * n)(*[t[2] for t in self.terms])
        zs = (c_uint64 * n)(*[t[3] for t in self.terms])
        h = _Hamiltonian(n, coeffs, xs, ys, zs)
        h._keep = (coeffs, xs, ys, zs)  # the arrays must outlive the struct
        return h


class Circuit:
    """A gate stream. `gates` mirrors what was added: (kind, qubits, params)."""

    def __init__(self, num_qubits: int):
        self.num_qubits = int(num_qubits)
        self.gates: list[tuple[int, tuple[int, ...], tuple[float, ...]]] = []
        self._h = _circuit_create(self.num_qubits)
        if not self._h:
            raise QMetalError(_err())

    def __del__(self, _free=_circuit_destroy):
        # The destroy function is bound here: at interpreter exit the module's
        # globals may already be gone when the last handles are collected.
        if getattr(self, "_h", None):
            _free(self._h)
            self._h = None

    def __len__(self) -> int:
        return int(_circuit_size(self._h))

    def add(self, kind: int, qubits, params=()):
        qubits = tuple(int(q) for q in qubits)
        params = tuple(float(p) for p in params)
        qs = (c_uint32 * len(qubits))(*qubits)
        ps = (c_double * len(params))(*params) if params else None
        _check(_circuit_add(self._h, int(kind), qs, len(qubits), ps, len(params)))
        self.gates.append((int(kind), qubits, params))
        return self

    _add = add

    # Single-qubit
    def x(self, q): return self._add(_X, (q,))
    def y(self, q): return self._add(_Y, (q,))
    def z(self, q): return self._add(_Z, (q,))
    def h(self, q): return self._add(_H, (q,))
    def s(self, q): return self._add(_S, (q,))
    def sdg(self, q): return self._add(_SDG, (q,))
    def t(self, q): return self._add(_T, (q,))
    def tdg(self, q): return self._add(_TDG, (q,))
    def rx(self, q, theta): return self._add(_RX, (q,), (theta,))
    def ry(self, q, theta): return self._add(_RY, (q,), (theta,))
    def rz(self, q, theta): return self._add(_RZ, (q,), (theta,))
    def p(self, q, theta): return self._add(_P, (q,), (theta,))
    def u3(self, q, theta, phi, lam): return self._add(_U3, (q,), (theta, phi, lam))

    # Multi-qubit
    def cx(self, c, t): return self._add(_CX, (c, t))
    def cy(self, c, t): return self._add(_CY, (c, t))
    def cz(self, a, b): return self._add(_CZ, (a, b))
    def swap(self, a, b): return self._add(_SWAP, (a, b))
    def cp(self, a, b, theta): return self._add(_CP, (a, b), (theta,))
    def ccx(self, c0, c1, t): return self._add(_CCX, (c0, c1, t))

    def plan_size(self, fuse: bool = True) -> int:
        """GPU passes this circuit lowers to -- what execution actually costs."""
        return int(_plan_size(self._h, 1 if fuse else 0))

    def schedule(self, options: FusionOptions | None = None) -> list[list[tuple[int, int]]]:
        """The passes qm_run makes for this circuit, in execution order: each
        a list of (plan op index, group wires mask -- 0 for the whole op)."""
        opts = options if options is not None else FusionOptions.default()
        p = _plan_build(self._h, byref(opts))
        if not p:
            raise QMetalError(_err())
        try:
            count = c_size_t()
            _check(_plan_schedule(p, None, None, None, 0, byref(count)))
            m = max(int(count.value), 1)
            ops, wires, pas = (c_uint32 * m)(), (c_uint64 * m)(), (c_uint32 * m)()
            _check(_plan_schedule(p, ops, wires, pas, m, byref(count)))
            out: list[list[tuple[int, int]]] = []
            for e in range(int(count.value)):
                if pas[e] == len(out):
                    out.append([])
                out[pas[e]].append((int(ops[e]), int(wires[e])))
            return out
        finally:
            _plan_destroy(p)

    def passes(self, options: FusionOptions | None = None) -> list[int]:
        """Index of the first pass that runs each plan op under executor-level
        pass merging (-1 for an op that launches nothing); max + 1 is the
        number of launches qm_run makes for this circuit."""
        opts = options if options is not None else FusionOptions.default()
        p = _plan_build(self._h, byref(opts))
        if not p:
            raise QMetalError(_err())
        try:
            n = int(_plan_nops(p))
            out = (c_uint32 * max(n, 1))()
            count = c_size_t()
            _check(_plan_passes(p, out, byref(count)))
            return [-1 if out[k] == 0xFFFFFFFF else int(out[k]) for k in range(n)]
        finally:
            _plan_destroy(p)

    def plan(self, options: FusionOptions | None = None) -> list[dict]:
        """The lowered plan as a list of dicts (kind name, wires, matrices,
        masks, angles, global phase), as qm_plan_build returns it."""
        opts = options if options is not None else FusionOptions.default()
        p = _plan_build(self._h, byref(opts))
        if not p:
            raise QMetalError(_err())
        try:
            out = []
            for k in range(int(_plan_nops(p))):
                op = _PlanOp()
                _check(_plan_get(p, k, byref(op)))
                ng, nt = int(op.num_group), int(op.num_terms)
                gq = (c_uint32 * max(ng, 1))()
                gm = (c_double * max(8 * ng, 1))()
                mk = (c_uint64 * max(nt, 1))()
                an = (c_double * max(nt, 1))()
                _check(_plan_arrays(p, k, gq, gm, mk, an))
                m = list(op.matrix)
                out.append({
                    "kind": PLAN_KINDS[op.kind],
                    "kind_id": int(op.kind),
                    "qubits": list(op.qubits),
                    "matrix": [complex(m[2 * i], m[2 * i + 1]) for i in range(4)],
                    "group_qubits": list(gq)[:ng],
                    "group_matrices": [complex(gm[2 * i], gm[2 * i + 1]) for i in range(4 * ng)],
                    "masks": list(mk)[:nt],
                    "angles": list(an)[:nt],
                    "global_phase": float(op.global_phase),
                    "block_bits": int(op.block_bits),
                })
            return out
        finally:
            _plan_destroy(p)


class Simulator:
    def __init__(self, num_qubits: int):
        self._h = _create(int(num_qubits))
        if not self._h:
            raise QMetalError(_err())

    def __del__(self, _free=_destroy):
        # The destroy function is bound here: at interpreter exit the module's
        # globals may already be gone when the last handles are collected.
        if getattr(self, "_h", None):
            _free(self._h)
            self._h = None

    def close(self):
        self.__del__()

    def __enter__(self): return self
    def __exit__(self, *exc): self.__del__()

    @property
    def num_qubits(self) -> int: return int(_num_qubits(self._h))
    @property
    def dim(self) -> int: return int(_dim(self._h))
    @property
    def dispatch_count(self) -> int: return int(_dispatch_count(self._h))
    @property
    def launch_count(self) -> int:
        """Kernels actually launched: a merged pass counts once, while
        dispatch_count adds one per plan op as the reference does."""
        return int(_launch_count(self._h))

    def reset(self): _check(_reset(self._h)); return self

    def run(self, circuit: Circuit, fuse: bool = True, options: FusionOptions | None = None):
        if options is not None:
            _check(_run_options(self._h, circuit._h, byref(options)))
        else:
            _check((_run if fuse else _run_unfused)(self._h, circuit._h))
        return self

    def synchronize(self): _check(_sync(self._h)); return self

    def amplitudes(self):
        """Host copy of the state as numpy complex64 (a list without numpy)."""
        p = _amplitudes(self._h)
        if not p:
            raise QMetalError(_err())
        n = self.dim
        try:
            import numpy as np
        except ImportError:
            return [complex(p[2 * i], p[2 * i + 1]) for i in range(n)]
        buf = np.ctypeslib.as_array(p, shape=(2 * n,))
        return buf.view(np.complex64).copy()

    def amplitudes_at(self, indices):
        """numpy complex64 amplitudes at the given basis indices, read on the
        device (for states too large to copy whole)."""
        import numpy as np
        idx = np.ascontiguousarray(indices, dtype=np.uint64).ravel()
        out = np.zeros(2 * max(idx.size, 1), dtype=np.float32)
        _check(_amplitudes_at(self._h, idx.ctypes.data_as(_u64p), idx.size,
                              out.ctypes.data_as(POINTER(c_float))))
        return out[:2 * idx.size].view(np.complex64).copy()

    def norm(self) -> float:
        out = c_double()
        _check(_norm(self._h, byref(out)))
        return out.value

    def expectation(self, observable) -> float:
        out = c_double()
        if isinstance(observable, Hamiltonian):
            h = observable._c()
        elif isinstance(observable, str):
            h = Hamiltonian().add_string(1.0, observable)._c()
        else:
            raise TypeError("observable must be a Hamiltonian or a Pauli string")
        _check(_expect(self._h, byref(h), byref(out)))
        return out.value

    def expectation_pauli(self, x: int = 0, y: int = 0, z: int = 0) -> float:
        out = c_double()
        _check(_expect_pauli(self._h, int(x), int(y), int(z), byref(out)))
        return out.value

    def sample(self, shots: int, seed: int = 0):
        buf = (c_uint64 * int(shots))()
        _check(_sample(self._h, int(shots), int(seed), buf))
        return list(buf)

    def counts(self, shots: int, seed: int = 0) -> dict[str, int]:
        """Bitstring counts, little-endian: qubit 0 is the rightmost character."""
        n = self.num_qubits
        out: dict[str, int] = {}
        for v in self.sample(shots, seed):
            key = format(v, f"0{n}b")
            out[key] = out.get(key, 0) + 1
        return dict(sorted(out.items(), key=lambda kv: -kv[1]))

    @property
    def placement(self) -> tuple[int, float]:
        """(allocations made for the state, time ratio of a pass over the top
        seven wires to one over wires 18..12 for the one kept); (0, 0.0) when
        the placement was not probed (states other than 33 qubits)."""
        a, r = c_uint32(), c_double()
        _check(_state_placement(self._h, byref(a), byref(r)))
        return int(a.value), float(r.value)

    def set_pass_merging(self, enable: bool = True):
        """Merge consecutive plan ops into shared HBM passes (default on);
        off gives one kernel launch per plan op."""
        _check(_set_pass_merging(self._h, 1 if enable else 0))
        return self

    def set_timing(self, enable: bool = True):
        """Record device time of every plan-op kernel (events on the
        simulator's stream)."""
        _check(_set_timing(self._h, 1 if enable else 0))
        return self

    def read_timing(self, cap: int = 1 << 16):
        """[(op kind name, device ms, algorithmic bytes)] since the last read."""
        kinds = (c_uint32 * cap)()
        ms = (c_double * cap)()
        by = (c_double * cap)()
        count = c_size_t()
        _check(_read_timing(self._h, kinds, ms, by, cap, byref(count)))
        return [(TIMING_KINDS[kinds[k]], ms[k], by[k]) for k in range(min(count.value, cap))]

    def energy_and_gradient(self, circuit: Circuit, hamiltonian: Hamiltonian):
        """Adjoint gradient: one entry per gate, zero where unparameterised."""
        h = hamiltonian._c()
        energy = c_double()
        grad = (c_double * max(len(circuit), 1))()
        _check(_energy_grad(self._h, circuit._h, byref(h), byref(energy), grad))
        return energy.value, list(grad)[:len(circuit)]


class Batch:
    """Many small circuits' energies and adjoint gradients in one persistent,
    graph-captured launch (qm_batch_*)."""

    def __init__(self):
        self._h = _batch_create()
        if not self._h:
            raise QMetalError(_err())
        self._sizes: list[int] = []

    def __del__(self, _free=_batch_destroy):
        # The destroy function is bound here: at interpreter exit the module's
        # globals may already be gone when the last handles are collected.
        if getattr(self, "_h", None):
            _free(self._h)
            self._h = None

    def close(self):
        self.__del__()

    def prepare(self, circuits, hamiltonians):
        n = len(circuits)
        cs = (c_void_p * n)(*[c._h for c in circuits])
        hs_py = [h._c() for h in hamiltonians]
        hs = (_Hamiltonian * n)(*hs_py)
        _check(_batch_prepare(self._h, n, cs, hs))
        self._sizes = [len(c) for c in circuits]
        return self

    def autotune(self):
        t, p, b = c_uint32(), c_uint32(), c_uint32()
        _check(_batch_autotune(self._h, byref(t), byref(p), byref(b)))
        return {"threads": t.value, "per_cu": p.value, "lds_bits": b.value,
                "spread_from": int(_batch_spread(self._h))}

    def set_config(self, threads: int, per_cu: int, lds_bits: int, spread_from=None):
        _check(_batch_set_config(self._h, threads, per_cu, lds_bits))
        if spread_from is not None:
            self.set_spread(spread_from)
        return self

    def set_spread(self, min_qubits: int):
        """Instances of min_qubits or more run spread (0: none)."""
        _check(_batch_set_spread(self._h, min_qubits))
        return self

    def batched_instances(self) -> int:
        """Spread instances that run as batched merged-pass programs (classes
        sharing width, gate structure and Hamiltonian) under the current
        spread width."""
        out = c_size_t()
        _check(_batch_batched(self._h, byref(out)))
        return int(out.value)

    def run(self):
        """(energies, per-instance gradient lists, device ms)."""
        import numpy as np
        e = np.zeros(max(len(self._sizes), 1))
        g = np.zeros(max(sum(self._sizes), 1))
        ms = c_double()
        _check(_batch_run(self._h, e.ctypes.data_as(_f64p), g.ctypes.data_as(_f64p), byref(ms)))
        grads, off = [], 0
        for k in self._sizes:
            grads.append(g[off:off + k].copy())
            off += k
        return e[:len(self._sizes)].copy(), grads, ms.value
