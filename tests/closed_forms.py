"""Closed forms the full-size GPU tests check against, pinned on the
complex128 oracle at small widths in tests/test_oracle.py.

Product state: the preparation RY(theta_q) then RZ(phi_q) on every wire
(qmetal_amd.circuits.product_state_prep) leaves qubit q in
a_q|0> + b_q|1> with a_q = cos(theta_q/2) e^{-i phi_q/2} and
b_q = sin(theta_q/2) e^{+i phi_q/2}.

QFT: the reference's circuit (src/circuits.cpp:33-48: stages j = n-1..0, H
then CP(k, j, pi/2^(j-k)), no swap layer) maps |x> to
N^{-1/2} sum_y exp(2 pi i x rev(y) / N) |y>, rev = n-bit reversal. On a
product state the sum over x factorises, so every output amplitude is
    N^{-1/2} prod_q (a_q + b_q exp(2 pi i 2^q rev(y) / N)).
"""
from __future__ import annotations

import math

import numpy as np

from qmetal_amd import circuits as C


def product_factors(prep):
    """(a, b) arrays for a product_state_prep stream."""
    n = len(prep) // 2
    a = np.zeros(n, dtype=np.complex128)
    b = np.zeros(n, dtype=np.complex128)
    for q in range(n):
        (k1, w1, (th,)), (k2, w2, (ph,)) = prep[2 * q], prep[2 * q + 1]
        assert k1 == C.RY and k2 == C.RZ and w1 == w2 == (q,)
        a[q] = math.cos(th / 2) * np.exp(-0.5j * ph)
        b[q] = math.sin(th / 2) * np.exp(0.5j * ph)
    return a, b


def bloch(a, b):
    """(<X_q>, <Y_q>, <Z_q>) of the product state, per qubit."""
    ov = np.conj(a) * b
    return 2 * ov.real, 2 * ov.imag, np.abs(a) ** 2 - np.abs(b) ** 2


def bit_reverse(y, n):
    y = np.asarray(y, dtype=np.uint64)
    r = np.zeros_like(y)
    for q in range(n):
        r |= ((y >> np.uint64(q)) & np.uint64(1)) << np.uint64(n - 1 - q)
    return r


def qft_product_amplitudes(a, b, idx):
    """FP64 amplitudes of QFT(prod_q (a_q|0> + b_q|1>)) at basis indices idx."""
    n = len(a)
    mask = (1 << n) - 1
    yr = [int(v) for v in bit_reverse(idx, n)]
    out = np.ones(len(yr), dtype=np.complex128)
    for q in range(n):
        turns = np.array([((v << q) & mask) / float(1 << n) for v in yr])
        out *= a[q] + b[q] * np.exp(2j * np.pi * turns)
    return out / math.sqrt(float(1 << n))


def inverse_stream(gates):
    """U^dagger of a stream of self-inverse gates, phase gates and U3, in
    reverse order (angles negated; U3(t, p, l)^dagger = U3(-t, -l, -p))."""
    out = []
    for kind, q, p in reversed(gates):
        if kind in (C.RX, C.RY, C.RZ, C.P, C.CP):
            out.append((kind, q, (-p[0],)))
        elif kind == C.U3:
            out.append((kind, q, (-p[0], -p[2], -p[1])))
        elif kind in (C.H, C.X, C.Y, C.Z, C.CX, C.CY, C.CZ, C.SWAP, C.CCX):
            out.append((kind, q, p))
        else:
            raise ValueError(f"no inverse for gate kind {kind}")
    return out
