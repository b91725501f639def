// fusion.h -- k-qubit gate fusion (k <= 5): plan, FP64 composition, apply.
#pragma once

#include <stdint.h>

#include <complex>
#include <vector>

#include "../../include/moonlab_hip_statevec.h"

namespace mlx {

using cd = std::complex<double>;

// A fused block: the sorted qubits it acts on and its 2^k x 2^k unitary,
// row-major, local basis bit j = qubits[j].
struct FusedBlock {
    std::vector<int> qubits;
    std::vector<cd>  mat;
    int              num_gates = 0;
    const moonlab_hip_gate_t *only_gate = nullptr;   // the member when there is exactly one
};

// Qubits a gate acts on (1..3); returns the count.
int gate_qubits(const moonlab_hip_gate_t &g, int q[3]);

// A gate every engine path can run on n qubits: a known code (0..24) and
// distinct qubits in range.
bool gate_valid(const moonlab_hip_gate_t &g, int n);

// Dense matrix of a gate on its own qubits (local bit j = q[j] from
// gate_qubits), row-major 2^nq x 2^nq.
void gate_dense(const moonlab_hip_gate_t &g, cd *m);

// 2x2 matrix (row-major) of a single-qubit gate; false for multi-qubit codes.
bool gate_matrix_1q(const moonlab_hip_gate_t &g, cd u[4]);

// Fusion into blocks of at most kmax qubits (a wider gate stands alone),
// several open at once over disjoint qubit sets; each block is composed in
// FP64 (left-multiplication in gate order).
std::vector<FusedBlock> plan_fusion(const moonlab_hip_gate_t *g, int ng, int kmax);

// Window planning (each block's qubit set chosen among windows of kmax
// neighbouring qubits seeded by `seeds` pending gates), and the plan with
// fewer blocks of the two; n bounds the windows.
constexpr int kFusionSeeds = 8;
std::vector<FusedBlock> plan_fusion_windows(const moonlab_hip_gate_t *g, int ng, int kmax, int seeds, int n);
std::vector<FusedBlock> plan_fusion_best(const moonlab_hip_gate_t *g, int ng, int kmax);

// Apply a gate list through fusion (returns a moonlab_hip_status_t value).
int apply_fused(moonlab_hip_state_t *s, const moonlab_hip_gate_t *g, int ng, int kmax);

}  // namespace mlx
