/* qmetal.h -- stable C ABI of the MI355X statevector simulator.
 *
 * Drop-in for the reference ABI (include/qmetal/qmetal.h of matiasrodlo/qmetal):
 * the same 27 entry points, types, gate codes and version macros, with the
 * same conventions. Everything after the "appended" marker is new and only
 * ever appended to.
 *
 * Convention: return 0 on success and -1 on failure, with qm_last_error()
 * holding the message for the calling thread. Every entry point catches;
 * nothing throws across the ABI. Handle constructors return NULL on failure.
 *
 * Index order is LSB-first: qubit q occupies bit q of the basis index.
 * Amplitudes are complex64 stored interleaved (re, im, re, im, ...), resident
 * in device memory (HBM) on the current HIP device.
 */

#ifndef QMETAL_H
#define QMETAL_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define QMETAL_VERSION_MAJOR 0
#define QMETAL_VERSION_MINOR 1
#define QMETAL_VERSION_PATCH 0

/* Gate codes. ABI: append only, never renumber (reference qmetal.h:28-36). */
typedef enum {
  QM_X = 0, QM_Y = 1, QM_Z = 2, QM_H = 3,
  QM_S = 4, QM_SDG = 5, QM_T = 6, QM_TDG = 7,
  QM_RX = 8, QM_RY = 9, QM_RZ = 10, QM_P = 11,
  QM_CX = 12, QM_CY = 13, QM_CZ = 14, QM_SWAP = 15,
  QM_CP = 16,
  QM_CCX = 17,
  /* appended: general one-qubit unitary U3(theta, phi, lambda) =
   * [[cos(t/2), -e^{i l} sin(t/2)], [e^{i p} sin(t/2), e^{i(p+l)} cos(t/2)]] */
  QM_U3 = 18
} qm_gate_kind;

typedef struct qm_simulator_s *qm_simulator;
typedef struct qm_circuit_s *qm_circuit;

/* A Hamiltonian as parallel arrays: one coefficient and three Pauli masks per
 * term (reference qmetal.h:41-50). */
typedef struct {
  uint32_t num_terms;
  const double *coeffs;
  const uint64_t *x_masks;
  const uint64_t *y_masks;
  const uint64_t *z_masks;
} qm_hamiltonian;

/* --- diagnostics (reference qmetal.h:54-60) -------------------------------- */

void qm_version(int *major, int *minor, int *patch);
const char *qm_last_error(void);          /* thread-local, valid until next call */
int qm_is_available(void);                /* 1 if a HIP device exists */
const char *qm_device_name(void);
uint64_t qm_max_buffer_length(void);      /* device memory in bytes */
uint64_t qm_recommended_working_set(void);/* free device memory in bytes */
uint32_t qm_max_qubits(void);             /* 34 on MI355X */

/* --- circuits (reference qmetal.h:64-73) ----------------------------------- */

qm_circuit qm_circuit_create(uint32_t num_qubits);
void qm_circuit_destroy(qm_circuit c);
int qm_circuit_add(qm_circuit c, qm_gate_kind kind, const uint32_t *qubits,
                   uint32_t num_qubits, const double *params,
                   uint32_t num_params);
size_t qm_circuit_size(qm_circuit c);

/* Number of GPU passes the circuit lowers to. With fuse = 0 this equals the
 * gate count; with fuse = 1 it is what execution will actually cost. */
size_t qm_circuit_plan_size(qm_circuit c, int fuse);

/* --- simulator (reference qmetal.h:77-107) --------------------------------- */

qm_simulator qm_create(uint32_t num_qubits);
void qm_destroy(qm_simulator s);

int qm_reset(qm_simulator s);
int qm_run(qm_simulator s, qm_circuit c);         /* fused */
int qm_run_unfused(qm_simulator s, qm_circuit c); /* ablation baseline */
int qm_synchronize(qm_simulator s);

uint32_t qm_num_qubits(qm_simulator s);
uint64_t qm_dim(qm_simulator s);
uint64_t qm_dispatch_count(qm_simulator s);

/* Host copy of the interleaved state, refreshed by each call, valid until the
 * next mutation or call. Length is 2 * qm_dim(). Returns NULL on error. */
const float *qm_amplitudes(qm_simulator s);

int qm_norm(qm_simulator s, double *out);
int qm_expectation_pauli(qm_simulator s, uint64_t x_mask, uint64_t y_mask,
                         uint64_t z_mask, double *out);
/* sum_k c_k <psi|P_k|psi>. On states where per-term reads would pass 256 MiB
 * (12 qubits and up), terms that fit a 12-wire tile are evaluated several per
 * read-only pass over the state; the others, and every term with pass
 * merging off, take one FP64 reduction each (many terms per launch). */
int qm_expectation(qm_simulator s, const qm_hamiltonian *h, double *out);

/* Draws `shots` basis-state indices without collapsing the state. `out` must
 * hold `shots` entries. Identical seeds give identical samples. */
int qm_sample(qm_simulator s, size_t shots, uint64_t seed, uint64_t *out);

/* Adjoint gradient. `out_grad` must hold qm_circuit_size(c) entries; gates
 * carrying no parameter get 0. Fails if a parameterised gate has no native
 * generator (only RX, RY, RZ do). */
int qm_energy_and_gradient(qm_simulator s, qm_circuit c,
                           const qm_hamiltonian *h, double *out_energy,
                           double *out_grad);

/* === appended ============================================================ */

/* Fusion switches (reference include/qmetal/plan.h:76-94). qm_run uses
 * {1, 1, 1, 4, 0}; qm_run_unfused uses {0, 0, 0, 1, 0}. block_bits >= 2
 * enables the LDS-staged Blocked1q pass (2..12). */
typedef struct {
  int collapse_1q_windows;
  int fuse_diagonals;
  int recognize_zz;
  uint32_t group_1q;
  uint32_t block_bits;
} qm_fusion_options;

int qm_run_options(qm_simulator s, qm_circuit c, const qm_fusion_options *o);

/* Read-only view of a built plan. Op kinds: 0 Dense1q, 1 Dense1qGroup,
 * 2 Blocked1q, 3 Controlled1q, 4 DiagonalLayer, 5 Swap, 6 CCX. */
typedef struct qm_plan_s *qm_plan;
typedef struct {
  uint32_t kind;
  uint32_t qubits[3];
  double matrix[8];       /* Dense1q / Controlled1q, row-major (re, im) */
  double global_phase;    /* DiagonalLayer */
  uint32_t block_bits;    /* Blocked1q */
  size_t num_terms;       /* DiagonalLayer */
  size_t num_group;       /* Dense1qGroup / Blocked1q wires */
} qm_plan_op;

qm_plan qm_plan_build(qm_circuit c, const qm_fusion_options *o);
void qm_plan_destroy(qm_plan p);
size_t qm_plan_size(qm_plan p);
int qm_plan_get(qm_plan p, size_t k, qm_plan_op *out);
/* Arrays of op k, each may be NULL: group_qubits[num_group],
 * group_matrices[8 * num_group], masks[num_terms], angles[num_terms]. */
int qm_plan_get_arrays(qm_plan p, size_t k, uint32_t *group_qubits,
                       double *group_matrices, uint64_t *masks,
                       double *angles);

/* Adjoint energy and gradient with device-memory arguments, for framework
 * tensors (ROCm PyTorch): `params` holds one double per RX/RY/RZ gate in
 * circuit order and replaces those gates' angles; `out_energy` (1 double) and
 * `out_grad` (num_params doubles, one per rotation) are device pointers;
 * `out_grad` may be NULL, and then only the energy is computed (forward
 * passes and the Hamiltonian's reductions, no backward sweep; when every
 * term fits a tile the energy is read off in read-only passes and no
 * co-state is allocated, so the first call with a gradient allocates it
 * and waits).
 * `stream` is the caller's hipStream_t (or NULL for the null stream): work
 * already queued there completes before the parameters are read, and work
 * queued there afterwards sees the results.
 * States of 12 qubits or more (with pass merging on and a non-empty
 * Hamiltonian): the first call for a circuit template (gate kinds and
 * wires; angles come from `params`) and Hamiltonian compiles a program and
 * waits; every later call only launches kernels on `stream` (the rotation
 * matrices are written on the device from `params`): no host
 * synchronisation, no allocation, so it can be captured in a graph. Device
 * calls on one simulator must be ordered among themselves (one stream, or
 * synchronised), as they share its states and program; the simulator's other
 * calls (qm_run, qm_expectation, ...) wait on the GPU for an uncaptured
 * call's kernels. Smaller states take the host path: the parameters are
 * copied to the host and the call waits for `stream` up to that point. */
int qm_energy_and_gradient_device(qm_simulator s, qm_circuit c,
                                  const qm_hamiltonian *h, const double *params,
                                  size_t num_params, double *out_energy,
                                  double *out_grad, void *stream);

/* Batched adjoint: many independent circuits (widths 1..20, each with its own
 * Hamiltonian) solved for energy and per-gate gradient in one persistent,
 * graph-captured launch. prepare compiles and uploads the programs (circuits
 * and Hamiltonians are borrowed for the call); run replays the launch and
 * copies out out_energy[count] and out_grad (qm_circuit_size(c_i) entries per
 * instance, back to back, 0 for gates without a parameter); out_ms gets the
 * device time of the launch. autotune times the launch configurations on the
 * prepared batch and keeps the fastest (workgroup size, persistent
 * workgroups per CU, widest instance kept in LDS). */
typedef struct qm_batch_s *qm_batch;
qm_batch qm_batch_create(void);
void qm_batch_destroy(qm_batch b);
int qm_batch_prepare(qm_batch b, size_t count, const qm_circuit *circuits,
                     const qm_hamiltonian *hams);
int qm_batch_autotune(qm_batch b, uint32_t *threads, uint32_t *per_cu,
                      uint32_t *lds_bits);
int qm_batch_set_config(qm_batch b, uint32_t threads, uint32_t per_cu,
                        uint32_t lds_bits);
int qm_batch_run(qm_batch b, double *out_energy, double *out_grad,
                 double *out_ms);
/* Instances of min_qubits or more (0 = none, the default; autotune picks it
 * by timing the whole batch) run spread: their states live in HBM side by
 * side and every program step is one launch over all of them, each instance
 * split over many workgroups; the narrower ones stay in the persistent
 * kernel. The same results to rounding either way (reductions in a fixed
 * order, so replays are identical). */
int qm_batch_set_spread(qm_batch b, uint32_t min_qubits);
/* The current spread width (0 = none; 0 for a NULL handle). */
uint32_t qm_batch_spread(qm_batch b);
/* Spread instances of 14 qubits or more that share a width, a gate
 * structure (kinds and wires in order) and a Hamiltonian run as one batched
 * program of the single-circuit merged passes (every pass one launch over all
 * of them) instead of the lockstep spread steps; *out gets how many instances
 * run that way under the current spread width. */
int qm_batch_batched_instances(qm_batch b, size_t *out);

/* Device timing of plan-op kernels, measured with events on the simulator's
 * own stream around each launch. qm_read_timing synchronises, reports every
 * op timed since the previous read (up to `cap` records: plan-op kind, device
 * milliseconds, algorithmic bytes = one read and one write of each amplitude
 * the op can change) and sets *count to the number timed. */
int qm_set_timing(qm_simulator s, int enable);
int qm_read_timing(qm_simulator s, uint32_t *kinds, double *ms, double *bytes,
                   size_t cap, size_t *count);

/* Executor-level pass merging below the plan (no counterpart in the
 * reference, whose executor encodes one dispatch per plan op,
 * src/metal/simulator.mm:436-590). Consecutive plan ops whose moving wires fit
 * one thread's register tile (bit 0 plus up to 4 wires) share one kernel and
 * one read-modify-write pass over the state; phase layers join any pass.
 * The plan, qm_circuit_plan_size and qm_dispatch_count are unchanged
 * (qm_dispatch_count still adds one per plan op, none for a diagonal layer
 * without terms, as src/metal/simulator.mm:450-595 and
 * tests/test_gpu.mm:247-256 count); qm_launch_count counts the kernels
 * actually launched, merged passes once. On by default;
 * qm_set_pass_merging(s, 0) restores one launch per plan op. Timing records of kind 7 to 12 are
 * merged passes (generic, chain-shaped, wide and low chain-shaped, generic
 * low tile and generic wide kernels); 13 to 16 are the adjoint gradient's
 * kernels (merged backward passes over both states, co-state build, energy
 * overlap, one-launch-per-rotation steps).
 * qm_plan_passes reports the grouping qm_run would use: pass_of_op
 * (qm_plan_size entries, may be NULL) gets the index of the first pass that
 * runs each op (UINT32_MAX for an op that launches nothing), *num_passes the
 * number of launches. Ops whose wires do not overlap may run out of plan
 * order, and a Dense1qGroup's 2x2s may run in different passes:
 * qm_plan_schedule lists every pass entry in execution order -- plan op,
 * group wires it applies (0: the whole op), pass index -- writing at most
 * `cap` entries (arrays may be NULL) and setting *count to their number. */
int qm_set_pass_merging(qm_simulator s, int enable);
uint64_t qm_launch_count(qm_simulator s);

/* Placement of the state (no counterpart in the reference). On MI355X a
 * 64 GiB state streams passes whose tiles span rows 4 MB or more apart at
 * 5.0 TB/s on some allocations and 5.9 TB/s on others, and which one an
 * allocation gets is up to the allocator. For 33 qubits the constructor
 * (and the co-state of the adjoint) therefore times one
 * value-preserving pass over the top seven wires against one over wires
 * 18..12, and allocates again while the first takes over 1.05 times the
 * second, up to 4 allocations, keeping the fastest (one extra state is held
 * at a time; QMETAL_PLACEMENT_PROBE=0 turns this off). *attempts gets the
 * number of allocations made for the state (0 when not probed), *ratio that
 * time ratio for the one kept. */
int qm_state_placement(qm_simulator s, uint32_t *attempts, double *ratio);
int qm_plan_passes(qm_plan p, uint32_t *pass_of_op, size_t *num_passes);
int qm_plan_schedule(qm_plan p, uint32_t *entry_op, uint64_t *entry_wires,
                     uint32_t *entry_pass, size_t cap, size_t *count);

/* Amplitudes at chosen basis indices, for states too large to copy whole
 * (qm_amplitudes copies all 2^n): out[2k], out[2k+1] get the (re, im) of
 * amplitude indices[k], k < count. Fails, writing nothing, if an index is
 * not below qm_dim(s). No counterpart in the reference, whose amplitudes are
 * a zero-copy view of unified memory (src/metal/simulator.mm:599-602). */
int qm_amplitudes_at(qm_simulator s, const uint64_t *indices, size_t count, float *out);

/* qm_sample without allocating: the same procedure, kernels and draws as
 * qm_sample (identical seeds give identical indices), with every piece of
 * device scratch -- block CDF, scan totals, drawn indices -- in a fixed
 * 16 MiB array per GPU that lives until the process ends. It makes no
 * hipMalloc, hipFree or pinned allocation on any call. No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * around HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc). Limits: states of up to 30 qubits and up to 2^20 shots per
 * call; past either it fails with a message naming the limit. All calls on
 * one GPU share its array: calls that overlap on one GPU (from several host
 * threads) race, calls made one after another do not (each waits for its
 * own work). `out` may be pageable or pinned host memory. */
int qm_sample_noalloc(qm_simulator s, size_t shots, uint64_t seed, uint64_t *out);

#ifdef __cplusplus
}
#endif

#endif /* QMETAL_H */
