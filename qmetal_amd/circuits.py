"""Workload generators: gate streams as [(kind, qubits, params), ...].

Structure follows the reference's frozen benchmark set (src/circuits.cpp);
randomness comes from one portable stream, splitmix64 -> 53-bit uniforms
(the same generator the sampler uses, src/metal/simulator.mm:75-80), never
from <random> distributions, whose output differs across C++ libraries.
"""
from __future__ import annotations

import math

import numpy as np

X, Y, Z, H, S, SDG, T, TDG = range(8)
RX, RY, RZ, P = 8, 9, 10, 11
CX, CY, CZ, SWAP, CP, CCX = 12, 13, 14, 15, 16, 17
U3 = 18

_MASK = (1 << 64) - 1


class SplitMix64:
    def __init__(self, seed: int):
        self.state = seed & _MASK

    def next_u64(self) -> int:
        self.state = (self.state + 0x9E3779B97F4A7C15) & _MASK
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _MASK
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _MASK
        return z ^ (z >> 31)

    def uniform(self, lo: float = 0.0, hi: float = 1.0) -> float:
        return lo + (hi - lo) * (self.next_u64() >> 11) * (1.0 / 9007199254740992.0)

    def normal(self) -> float:
        # Box-Muller on two uniforms; 1 - u keeps the log argument in (0, 1].
        u1, u2 = self.uniform(), self.uniform()
        return math.sqrt(-2.0 * math.log(1.0 - u1)) * math.cos(2.0 * math.pi * u2)


def to_circuit(n: int, gates):
    from . import Circuit
    c = Circuit(n)
    for kind, q, p in gates:
        c.add(kind, q, p)
    return c


def ghz(n: int):
    """H on 0, then CX(i, i+1) (src/circuits.cpp:24-31)."""
    return [(H, (0,), ())] + [(CX, (i, i + 1), ()) for i in range(n - 1)]


def qft(n: int):
    """Descending stages, H then CP(k, j, pi / 2^(j-k)), no swap layer
    (src/circuits.cpp:33-48)."""
    out = []
    for j in range(n - 1, -1, -1):
        out.append((H, (j,), ()))
        for k in range(j - 1, -1, -1):
            out.append((CP, (k, j), (math.pi / float(1 << (j - k)),)))
    return out


def zz(a: int, b: int, theta: float):
    """exp(-i theta Z_a Z_b) as CX, RZ(2 theta), CX (src/circuits.cpp:12-16)."""
    return [(CX, (a, b), ()), (RZ, (b,), (2.0 * theta,)), (CX, (a, b), ())]


def qaoa_ring(n: int, p: int, gammas=None, betas=None):
    """H layer, then per layer ZZ(gamma) on ring bonds and RX(2 beta)
    (src/circuits.cpp:50-63). Default angles are the reference's fixed ones."""
    out = [(H, (i,), ()) for i in range(n)]
    for layer in range(p):
        g = gammas[layer] if gammas is not None else 0.3 + 0.11 * layer
        b = betas[layer] if betas is not None else 0.7 - 0.05 * layer
        for i in range(n):
            out += zz(i, (i + 1) % n, g)
        out += [(RX, (i,), (2.0 * b,)) for i in range(n)]
    return out


def tfim_trotter(n: int, steps: int):
    """src/circuits.cpp:65-75."""
    out = []
    for _ in range(steps):
        for i in range(n - 1):
            out += zz(i, i + 1, 0.1)
        out += [(RX, (i,), (2.0 * 0.8 * 0.1,)) for i in range(n)]
    return out


def grover_proxy(n: int, iters: int):
    """src/circuits.cpp:77-93."""
    out = [(H, (i,), ()) for i in range(n)]
    for _ in range(iters):
        out.append((CZ, (n - 2, n - 1), ()))
        out += [(H, (i,), ()) for i in range(n)]
        out += [(X, (i,), ()) for i in range(n)]
        out.append((CZ, (n - 2, n - 1), ()))
        out += [(X, (i,), ()) for i in range(n)]
        out += [(H, (i,), ()) for i in range(n)]
    return out


def hardware_efficient(n: int, depth: int, seed: int = 0x5EED, lo: float = 0.0,
                       hi: float = 2.0 * math.pi, params=None):
    """RY, RZ on every wire then a CX ladder, per layer (src/circuits.cpp:117-131).
    Angles ~ U[lo, hi) from splitmix64(seed), or `params` in gate order."""
    rng = SplitMix64(seed)
    out = []
    it = iter(params) if params is not None else None
    for _ in range(depth):
        for q in range(n):
            a = next(it) if it else rng.uniform(lo, hi)
            b = next(it) if it else rng.uniform(lo, hi)
            out.append((RY, (q,), (a,)))
            out.append((RZ, (q,), (b,)))
        out += [(CX, (i, i + 1), ()) for i in range(n - 1)]
    return out


def product_state_prep(n: int, seed: int = 2024):
    """RY(theta_q), RZ(phi_q) on every wire, theta, phi ~ U[0, 2 pi)."""
    rng = SplitMix64(seed)
    out = []
    for q in range(n):
        out.append((RY, (q,), (rng.uniform(0, 2 * math.pi),)))
        out.append((RZ, (q,), (rng.uniform(0, 2 * math.pi),)))
    return out


def haar_su2(rng: SplitMix64) -> np.ndarray:
    """Haar 2x2 unitary: QR of a complex Gaussian matrix with R's diagonal
    made positive."""
    a = np.array([[complex(rng.normal(), rng.normal()) for _ in range(2)] for _ in range(2)])
    q, r = np.linalg.qr(a)
    d = np.diag(r)
    return q * (d / np.abs(d))


def u3_params(u: np.ndarray):
    """(theta, phi, lambda) with u = e^{i alpha} U3(theta, phi, lambda)."""
    c, s = abs(u[0, 0]), abs(u[1, 0])
    theta = 2.0 * math.atan2(s, c)
    if s < 1e-15:  # diagonal: alpha from u00, phi = 0
        return (theta, 0.0, float(np.angle(u[1, 1]) - np.angle(u[0, 0])))
    alpha = float(np.angle(u[0, 0])) if c > 1e-15 else 0.0
    return (theta, float(np.angle(u[1, 0])) - alpha, float(np.angle(-u[0, 1])) - alpha)


def brickwork(n: int, depth: int, seed: int = 0xB1C):
    """Per layer a Haar U3 on every wire, then CX on disjoint neighbour pairs
    starting at offset (layer mod 2) -- the reference random circuit's layer
    shape with one dense gate per wire."""
    rng = SplitMix64(seed)
    out = []
    for d in range(depth):
        for q in range(n):
            out.append((U3, (q,), u3_params(haar_su2(rng))))
        for i in range(d % 2, n - 1, 2):
            out.append((CX, (i, i + 1), ()))
    return out


def random_circuit(n: int, depth: int, seed: int = 0x5EED):
    """Reference random_circuit shape (src/circuits.cpp:95-115) drawn from
    splitmix64: one of RX/RY/RZ/H/T per wire, then a shifted CX layer."""
    rng = SplitMix64(seed)
    pool = (RX, RY, RZ, H, T)
    out = []
    for d in range(depth):
        for q in range(n):
            k = pool[rng.next_u64() % 5]
            if k in (RX, RY, RZ):
                out.append((k, (q,), (rng.uniform(0, 2 * math.pi),)))
            else:
                out.append((k, (q,), ()))
        for i in range(d % 2, n - 1, 2):
            out.append((CX, (i, i + 1), ()))
    return out


def tfim_ring_hamiltonian(n: int, hx: float = 0.7):
    """sum Z_i Z_{i+1 mod n} + hx sum X_i as (coeff, x, y, z) terms."""
    terms = [(1.0, 0, 0, (1 << i) | (1 << ((i + 1) % n))) for i in range(n)]
    terms += [(hx, 1 << i, 0, 0) for i in range(n)]
    return terms


def ring_zz_hamiltonian(n: int):
    return [(1.0, 0, 0, (1 << i) | (1 << ((i + 1) % n))) for i in range(n)]


def to_hamiltonian(terms):
    from . import Hamiltonian
    h = Hamiltonian()
    for c, x, y, z in terms:
        h.add(c, x, y, z)
    return h


def qaoa_instances(count: int = 4096, p: int = 6, lo: int = 10, hi: int = 20,
                   seed: int = 0x0A0A):
    """Batched QAOA-ring workload: per instance n ~ U{lo..hi},
    gamma_k ~ U[0, pi], beta_k ~ U[0, pi/2], all from splitmix64(seed)."""
    rng = SplitMix64(seed)
    out = []
    for _ in range(count):
        n = lo + int(rng.next_u64() % (hi - lo + 1))
        gam = [rng.uniform(0, math.pi) for _ in range(p)]
        bet = [rng.uniform(0, math.pi / 2) for _ in range(p)]
        out.append((n, gam, bet))
    return out
