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


def hea_batch(count: int, seed: int, n_range=(8, 18), layer_range=(4, 12), h_range=(0.5, 1.5)):
    """Hardware-efficient-ansatz VQE instances: n ~ U{n_range}, layers ~
    U{layer_range}, angles ~ U[0, 2 pi), transverse field h ~ U[h_range]."""
    rng = SplitMix64(seed)
    out = []
    for _ in range(count):
        n = n_range[0] + rng.below(n_range[1] - n_range[0] + 1)
        L = layer_range[0] + rng.below(layer_range[1] - layer_range[0] + 1)
        ang = np.array([2.0 * math.pi * rng.unit() for _ in range(L * n * 2)]).reshape(L, n, 2)
        h = h_range[0] + (h_range[1] - h_range[0]) * rng.unit()
        out.append((n, L, ang, h))
    return out


def hea_gate_list(n: int, L: int, ang: np.ndarray) -> np.ndarray:
    rows = []
    for l in range(L):
        for q in range(n):
            rows.append((int(GateType.RY), q, -1, -1, float(ang[l, q, 0]), 0.0, 0.0))
            rows.append((int(GateType.RZ), q, -1, -1, float(ang[l, q, 1]), 0.0, 0.0))
        for q in range(n - 1):
            rows.append((int(GateType.CZ), q + 1, q, -1, 0.0, 0.0, 0.0))
    return np.array(rows, dtype=GATE_DTYPE)


def serving_requests(count: int, seed: int, rate_hz: float, n_range=(10, 30), depth_range=(10, 50),
                     shots: int = 1024):
    """Poisson stream of random-layer circuits: n in n_range with P(n)
    proportional to 2^(-(n - n_min)/4), depth ~ U{depth_range}, precision
    complex128 / complex64 at 50/50, `shots` shots + <Z_q> each.  Returns
    dicts with arrival time (s), n, depth, precision, shot seed and gates."""
    rng = SplitMix64(seed)
    ns = list(range(n_range[0], n_range[1] + 1))
    w = np.array([2.0 ** (-(n - n_range[0]) / 4.0) for n in ns])
    cdf = np.cumsum(w / w.sum())
    t = 0.0
    out = []
    for i in range(count):
        t += -math.log(1.0 - rng.unit()) / rate_hz
        n = ns[int(np.searchsorted(cdf, rng.unit(), side="right").clip(0, len(ns) - 1))]
        depth = depth_range[0] + rng.below(depth_range[1] - depth_range[0] + 1)
        prec = int(rng.below(2))
        cseed = rng.next_u64()
        sseed = rng.next_u64() | 1
        out.append({"arrival": t, "n": n, "depth": depth, "precision": prec, "shot_seed": sseed,
                    "shots": shots, "gates": random_layer_circuit(n, depth, cseed)})
    return out
