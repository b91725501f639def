"""Seeded circuit generators for the dense state-vector workloads.

Every generator is a pure function of its seed, driven by a SplitMix64
stream (Steele, Lea & Flood 2014) -- the PRNG style of the reference's
corpus generators (scripts/gen_diff_corpus.py:64-84): next = mix(state +=
0x9E3779B97F4A7C15), unit() = top 53 bits / 2^53.

Random-layer circuit (20-qubit CPU-runnable case and the 30-qubit bandwidth
sweep): optionally an H on every qubit first, then `depth` layers; a layer
gives every qubit one of RX / RY / RZ (kind = next() % 3) with angle
2*pi*unit(), then CNOT(q, q+1) for even q, then for odd q.
"""
from __future__ import annotations

import math
from typing import List, Tuple

import numpy as np

from . import GATE_DTYPE, GateType

_M64 = (1 << 64) - 1


class SplitMix64:
    def __init__(self, seed: int) -> None:
        self.state = seed & _M64

    def next_u64(self) -> int:
        self.state = (self.state + 0x9E3779B97F4A7C15) & _M64
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _M64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _M64
        return z ^ (z >> 31)

    def unit(self) -> float:
        return (self.next_u64() >> 11) / float(1 << 53)

    def below(self, bound: int) -> int:
        return self.next_u64() % bound


def random_layer_circuit(n: int, depth: int, seed: int, h_layer: bool = False) -> np.ndarray:
    rng = SplitMix64(seed)
    rows: List[Tuple] = []
    if h_layer:
        rows += [(int(GateType.H), q, -1, -1, 0.0, 0.0, 0.0) for q in range(n)]
    kinds = (GateType.RX, GateType.RY, GateType.RZ)
    for _ in range(depth):
        for q in range(n):
            k = kinds[rng.below(3)]
            rows.append((int(k), q, -1, -1, 2.0 * math.pi * rng.unit(), 0.0, 0.0))
        for q in range(0, n - 1, 2):
            rows.append((int(GateType.CNOT), q + 1, q, -1, 0.0, 0.0, 0.0))
        for q in range(1, n - 1, 2):
            rows.append((int(GateType.CNOT), q + 1, q, -1, 0.0, 0.0, 0.0))
    return np.array(rows, dtype=GATE_DTYPE)


def su2_cz_circuit(n: int, depth: int, seed: int) -> np.ndarray:
    """Depth layers of a random SU(2) (U3 with Haar-distributed angles) on
    every qubit followed by a CZ brickwork (even pairs, then odd pairs)."""
    rng = SplitMix64(seed)
    rows: List[Tuple] = []
    for _ in range(depth):
        for q in range(n):
            theta = 2.0 * math.acos(math.sqrt(1.0 - rng.unit()))
            phi = 2.0 * math.pi * rng.unit()
            lam = 2.0 * math.pi * rng.unit()
            rows.append((int(GateType.U3), q, -1, -1, theta, phi, lam))
        for q in range(0, n - 1, 2):
            rows.append((int(GateType.CZ), q + 1, q, -1, 0.0, 0.0, 0.0))
        for q in range(1, n - 1, 2):
            rows.append((int(GateType.CZ), q + 1, q, -1, 0.0, 0.0, 0.0))
    return np.array(rows, dtype=GATE_DTYPE)


def algorithmic_bytes(gates: np.ndarray, n: int, amp_bytes: int = 16) -> int:
    """Bytes a gate list must move at minimum: dense 1q (and RZ, which
    rescales both halves) read + write every amplitude (2 D b); phase-on-one
    and CNOT/CY/SWAP touch half (2 (D/2) b); CZ, Toffoli and Fredkin a
    quarter (2 (D/4) b)."""
    D = 1 << n
    half = {GateType.Z, GateType.S, GateType.T, GateType.SDG, GateType.TDG, GateType.PHASE,
            GateType.CNOT, GateType.CY, GateType.SWAP, GateType.CY_TRUE, GateType.CRX,
            GateType.CRY, GateType.CRZ}
    quarter = {GateType.CZ, GateType.CPHASE, GateType.TOFFOLI, GateType.FREDKIN}
    total = 0
    for t in gates["type"]:
        t = GateType(int(t))
        if t == GateType.I:
            continue
        if t in quarter:
            total += 2 * (D // 4) * amp_bytes
        elif t in half:
            total += 2 * (D // 2) * amp_bytes
        else:
            total += 2 * D * amp_bytes
    return total
