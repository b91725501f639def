// sv_internal.h -- host-side interface between the engine (C++) and the HIP
// kernel translation units.  Not installed; the public ABI is include/*.h.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "../../include/moonlab_hip_statevec.h"

namespace mlx {

// A device state vector as the launchers see it.
struct SvDev {
    void       *amps;    // 2^n complex (double2 or float2)
    int         n;
    int         prec;    // MOONLAB_HIP_C128 / MOONLAB_HIP_C64
    hipStream_t stream;
};

// Launch tuning for the streaming gate kernels: pairs (or elements) handled
// per thread and threads per block.  Set once per process by the autotuner
// or the environment (MOONLAB_HIP_EPT / MOONLAB_HIP_BLOCK).
struct GateTuning {
    int ept;     // 1, 2, 4 or 8
    int block;   // 64, 128, 256, 512 or 1024
    uint64_t nt_min_bytes;   // states at least this large use streaming loads/stores
    int blocked_tile_bits = 0;    // tile of the single-state blocked path: 11, 12 or 13 qubits (0: by precision)
    int blocked_threads = 512;    // workgroup size of its LDS op engine: 256 or 512
    int blocked_engine = 0;       // 0, 2: register phases (k_reg_pass), 16 values per thread; 1: LDS
                                  // ops (k_tile_pass); 3: register phases of 8 values per thread
    int blocked_lookahead = 8;    // seed ops per pass / phase of the blocked planner's lookahead (0: greedy)
};
GateTuning &gate_tuning();

// Record the text moonlab_hip_last_error() returns on this thread (returns st).
moonlab_hip_status_t set_last_error(moonlab_hip_status_t st, const char *what, hipError_t e = hipSuccess);

// Point a state's work at another stream (waits for the state's current
// stream first); returns the stream it used before.
hipStream_t state_bind_stream(moonlab_hip_state_t *s, hipStream_t st);

// Gate launchers (sv_gates.hip).  All return a hipError_t value.
hipError_t launch_gate(const SvDev &s, const moonlab_hip_gate_t &g);
hipError_t launch_matrix_1q(const SvDev &s, int q, const double m[8]);
hipError_t launch_matrix_2q(const SvDev &s, int q0, int q1, const double m[32]);
hipError_t launch_controlled_1q(const SvDev &s, uint64_t ctrl_mask, int target,
                                const double m[8]);
hipError_t launch_mcx(const SvDev &s, uint64_t ctrl_mask, int target);
hipError_t launch_mcz(const SvDev &s, uint64_t all_mask);
hipError_t launch_cswap(const SvDev &s, uint64_t ctrl_mask, int t1, int t2);
hipError_t launch_diag_1q(const SvDev &s, int q, const double d[4]);
hipError_t launch_phase_mask(const SvDev &s, uint64_t all_mask, double pr, double pi);
hipError_t launch_init_zero(const SvDev &s);
hipError_t launch_scale(const SvDev &s, double f);
hipError_t launch_convert(const void *src, int src_prec, void *dst, int dst_prec,
                          uint64_t count, hipStream_t st);

// Reductions and sampling (sv_reduce.hip).
struct ReduceWs {
    double  *partial;     // [blocks][n + 1]
    int      blocks;
    double  *out;         // [n + 1]: norm, then P(1) per qubit
    double  *tile_sums;   // [ntiles]
    double  *tile_incl;   // [ntiles] inclusive prefix of tile_sums
    uint64_t ntiles;
    int      tile_bits;
};
void       reduce_geometry(int n, int *tile_bits, uint64_t *ntiles, int *blocks);
hipError_t launch_reduce_z(const SvDev &s, const ReduceWs &w, bool want_tiles);
hipError_t launch_tile_scan(const ReduceWs &w, hipStream_t st);
// Collapsing measurements (sv_measure.hip): the FP64 probabilities of the 2^k
// outcomes of qubits[0..k-1] (outcome bit j = qubits[j], distinct, k <= 20)
// into host_out (synchronizes the stream); and the collapse pass (mode 0:
// zero the amplitudes whose masked bits differ from want, 1: also scale the
// rest by inv, 2: the basis state want with amplitude 1).
hipError_t launch_outcome_probs(const SvDev &s, const int *qubits, int k, double *host_out);
hipError_t launch_collapse(const SvDev &s, uint64_t mask, uint64_t want, double inv, int mode);
hipError_t launch_sample(const SvDev &s, const ReduceWs &w, const double *d_u, int nshots,
                         uint64_t *d_out);
hipError_t launch_probabilities(const SvDev &s, double *d_p);
// Pauli-pair sums over one qubit: out[0] = <X_q>, out[1] = <Y_q>.
hipError_t launch_xy(const SvDev &s, int q, double *d_partial, int blocks, double *d_out);
// <Z_mask> = sum_i |a_i|^2 (-1)^popcount(i & mask).
hipError_t launch_pauli(const SvDev &s, uint64_t x_mask, uint64_t z_mask, double *d_partial, int blocks,
                        double *d_out);
hipError_t launch_zmask(const SvDev &s, uint64_t mask, double *d_partial, int blocks,
                        double *d_out);
// <a|b> (both same n / precision): out[0] = Re, out[1] = Im.
hipError_t launch_inner(const SvDev &a, const SvDev &b, double *d_partial, int blocks,
                        double *d_out);

}  // namespace mlx
