/* qm_oracle.h -- complex128 CPU oracle for the statevector path.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker that the HIP path is
 * compared against. Only tests/, __graft_entry__.smoke() and bench.py's CPU
 * baseline leg may load it; the product (qmetal_amd, libqmetal.so) never does.
 *
 * It restates, in plain C, the semantics of the reference's complex128 oracle
 * and of the host-side algebra its executor uses:
 *
 *   qo_gate_matrix      src/reference.cpp:60-98   (+ U3, appended gate code 18)
 *   qo_apply_1q         src/reference.cpp:113-123
 *   qo_apply_controlled src/reference.cpp:125-136
 *   qo_apply_diag2q     src/reference.cpp:138-142
 *   qo_apply_swap       src/reference.cpp:144-151
 *   qo_apply_ccx        src/reference.cpp:153-159
 *   qo_apply_gate       src/reference.cpp:161-201
 *   qo_apply_group      src/reference.cpp:216-222 (Dense1qGroup / Blocked1q)
 *   qo_apply_diag       src/reference.cpp:229-240 (DiagonalLayer)
 *   qo_norm             src/reference.cpp:250-254
 *   qo_expect_pauli     src/reference.cpp:264-284
 *   qo_splitmix64       src/metal/simulator.mm:75-80
 *   qo_sample           src/metal/simulator.mm:677-763 (the two-level block
 *                       CDF procedure, evaluated in double on given amplitudes)
 *   qo_adjoint          src/metal/simulator.mm:814-963 (adjoint sweep, done in
 *                       complex128 here; pinned by parameter shift)
 *
 * Pinning (what checks each function; see tests/test_oracle.py):
 *   gate matrices, 1q/2q/3q kernels, run   closed forms of tests/test_reference.cpp
 *                                          plus golden amplitudes from the
 *                                          reference oracle built in oracle/_ref
 *   group / diag plan execution            plan-vs-gate equality at 1e-12 on the
 *                                          reference's golden plans
 *   expectation                            Y/X/Z closed forms, Bell YY=-1,
 *                                          brute-force dense-matrix <psi|P|psi>
 *   splitmix64                             published test vector of splitmix64
 *   sample                                 exact brute-force inverse-CDF search
 *                                          on tiny states
 *   adjoint                                exact parameter shift (RX/RY/RZ)
 *
 * Index order is LSB-first: qubit q has stride 1 << q. States are interleaved
 * (re, im) doubles of length 2 * 2^n.
 */
#ifndef QM_ORACLE_H
#define QM_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Gate codes equal qm_gate_kind (include/qmetal/qmetal.h). */
enum {
  QO_X = 0, QO_Y, QO_Z, QO_H, QO_S, QO_SDG, QO_T, QO_TDG,
  QO_RX, QO_RY, QO_RZ, QO_P,
  QO_CX, QO_CY, QO_CZ, QO_SWAP, QO_CP, QO_CCX,
  QO_U3
};

void qo_set_threads(int threads);
int qo_gate_matrix(int kind, const double *params, double *m8);

void qo_reset(double *state, uint32_t n);
void qo_apply_1q(double *state, uint32_t n, uint32_t q, const double *m8);
void qo_apply_controlled(double *state, uint32_t n, uint32_t c, uint32_t t,
                         const double *m8);
void qo_apply_diag2q(double *state, uint32_t n, uint32_t a, uint32_t b,
                     double ph_re, double ph_im);
void qo_apply_swap(double *state, uint32_t n, uint32_t a, uint32_t b);
void qo_apply_ccx(double *state, uint32_t n, uint32_t c0, uint32_t c1,
                  uint32_t t);

/* Returns 0, or -1 for an unsupported gate / bad wires (fail closed). */
int qo_apply_gate(double *state, uint32_t n, int kind, const uint32_t *qubits,
                  uint32_t nq, const double *params);
/* Runs `ngates` gates: kinds[g], nq[g], qubits[3g..3g+2], params[3g..3g+2].
 * Returns -1 - g on the first failing gate, else 0. */
int qo_run(double *state, uint32_t n, size_t ngates, const int32_t *kinds,
           const uint8_t *nq, const uint32_t *qubits, const double *params);

void qo_apply_group(double *state, uint32_t n, size_t count,
                    const uint32_t *qubits, const double *mats8);
void qo_apply_diag(double *state, uint32_t n, size_t count,
                   const uint64_t *masks, const double *angles,
                   double global_phase);

double qo_norm(const double *state, uint32_t n);
double qo_expect_pauli(const double *state, uint32_t n, uint64_t x,
                       uint64_t y, uint64_t z);

uint64_t qo_splitmix64(uint64_t *s);
/* Block CDF sampler over the probabilities |amp|^2 of a complex128 state. */
void qo_sample(const double *state, uint32_t n, size_t shots, uint64_t seed,
               uint64_t *out);

/* Adjoint energy and gradient in complex128. Hamiltonian as parallel arrays.
 * grad has ngates entries (0 where the gate carries no parameter). Returns 0,
 * or -1 if a parameterised gate is not RX/RY/RZ. `work` must hold 2 states. */
int qo_adjoint(uint32_t n, size_t ngates, const int32_t *kinds,
               const uint8_t *nq, const uint8_t *npar, const uint32_t *qubits,
               const double *params,
               uint32_t nterms, const double *coeffs, const uint64_t *xm,
               const uint64_t *ym, const uint64_t *zm, double *work,
               double *energy, double *grad);

#ifdef __cplusplus
}
#endif
#endif
