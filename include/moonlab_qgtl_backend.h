/*
 * moonlab_qgtl_backend.h -- circuit-ingestion C ABI, executed on MI355X.
 *
 * Same names, types, enum values, status codes and ownership rules as the
 * reference's src/applications/moonlab_qgtl_backend.h (lines cited per entry),
 * so code written against the reference compiles and links unchanged.  The
 * difference is where `moonlab_qgtl_execute` runs: here every gate is a HIP
 * kernel on the GPU and the shot loop samples one device-side CDF instead of
 * cloning the state per shot.
 */
#ifndef MOONLAB_MI355X_QGTL_BACKEND_H
#define MOONLAB_MI355X_QGTL_BACKEND_H

#include <stddef.h>
#include <stdint.h>

#ifndef MOONLAB_API
#define MOONLAB_API __attribute__((visibility("default")))
#endif

#ifdef __cplusplus
extern "C" {
#endif

/* Gate tags (moonlab_qgtl_backend.h:57-72). */
typedef enum {
    MOONLAB_QGTL_GATE_I    = 0,
    MOONLAB_QGTL_GATE_X    = 1,
    MOONLAB_QGTL_GATE_Y    = 2,
    MOONLAB_QGTL_GATE_Z    = 3,
    MOONLAB_QGTL_GATE_H    = 4,
    MOONLAB_QGTL_GATE_S    = 5,
    MOONLAB_QGTL_GATE_T    = 6,
    MOONLAB_QGTL_GATE_RX   = 7,
    MOONLAB_QGTL_GATE_RY   = 8,
    MOONLAB_QGTL_GATE_RZ   = 9,
    MOONLAB_QGTL_GATE_CNOT = 10,
    MOONLAB_QGTL_GATE_CY   = 11,
    MOONLAB_QGTL_GATE_CZ   = 12,
    MOONLAB_QGTL_GATE_SWAP = 13
} moonlab_qgtl_gate_t;

/* Status codes (moonlab_qgtl_backend.h:77-81). */
#define MOONLAB_QGTL_OK          ( 0)
#define MOONLAB_QGTL_BAD_ARG     (-301)
#define MOONLAB_QGTL_OOM         (-302)
#define MOONLAB_QGTL_UNSUPPORTED (-303)
#define MOONLAB_QGTL_INTERNAL    (-304)

typedef struct moonlab_qgtl_circuit moonlab_qgtl_circuit_t;

/* Execution options (moonlab_qgtl_backend.h:90-94). */
typedef struct {
    int      num_shots;
    uint64_t rng_seed;
    int      return_probabilities;
} moonlab_qgtl_exec_options_t;

/* Results (moonlab_qgtl_backend.h:104-109). */
typedef struct {
    int       num_qubits;
    int       num_shots;
    uint64_t *outcomes;
    double   *probabilities;
} moonlab_qgtl_results_t;

/* Replaces moonlab_qgtl_circuit_create (moonlab_qgtl_backend.h:121-122):
 * 1..32 qubits, else NULL. */
MOONLAB_API moonlab_qgtl_circuit_t *moonlab_qgtl_circuit_create(int num_qubits);
/* Replaces moonlab_qgtl_circuit_free (moonlab_qgtl_backend.h:128). */
MOONLAB_API void moonlab_qgtl_circuit_free(moonlab_qgtl_circuit_t *c);
/* Replaces moonlab_qgtl_add_gate (moonlab_qgtl_backend.h:144-149): validates
 * target/control range, control != target for 2q tags, params for RX/RY/RZ. */
MOONLAB_API int moonlab_qgtl_add_gate(moonlab_qgtl_circuit_t *c, moonlab_qgtl_gate_t type,
                                      int target, int control, const double *params);
/* Replaces moonlab_qgtl_execute (moonlab_qgtl_backend.h:164-167). */
MOONLAB_API int moonlab_qgtl_execute(moonlab_qgtl_circuit_t *c,
                                     const moonlab_qgtl_exec_options_t *opts,
                                     moonlab_qgtl_results_t *out);
/* Replaces moonlab_qgtl_results_free (moonlab_qgtl_backend.h:172). */
MOONLAB_API void moonlab_qgtl_results_free(moonlab_qgtl_results_t *r);
/* Replace moonlab_qgtl_circuit_num_qubits / _num_gates (moonlab_qgtl_backend.h:182-188). */
MOONLAB_API int moonlab_qgtl_circuit_num_qubits(const moonlab_qgtl_circuit_t *c);
MOONLAB_API int moonlab_qgtl_circuit_num_gates(const moonlab_qgtl_circuit_t *c);
/* Replace the "# moonlab-circuit v1" text format entry points
 * (moonlab_qgtl_backend.h:226-273). */
MOONLAB_API int moonlab_qgtl_circuit_serialize(const moonlab_qgtl_circuit_t *c, char *buf,
                                               size_t buf_size, size_t *out_written);
MOONLAB_API moonlab_qgtl_circuit_t *moonlab_qgtl_circuit_deserialize(const char *buf,
                                                                     size_t buf_size,
                                                                     int *out_status);
MOONLAB_API int moonlab_qgtl_circuit_save(const moonlab_qgtl_circuit_t *c, const char *path);
MOONLAB_API moonlab_qgtl_circuit_t *moonlab_qgtl_circuit_load(const char *path, int *out_status);

#ifdef __cplusplus
}
#endif
#endif
