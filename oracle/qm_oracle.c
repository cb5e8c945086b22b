/* qm_oracle.c -- complex128 CPU oracle. Test infrastructure only; see the
 * header for what each function restates and what pins it. */
#include "qm_oracle.h"

#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef double _Complex cplx;

static int g_threads = 1;
void qo_set_threads(int threads) { g_threads = threads > 0 ? threads : 1; }

#define PAR _Pragma("omp parallel for if(g_threads > 1) num_threads(g_threads) schedule(static)")

static inline cplx ld(const double *s, size_t i) {
  return CMPLX(s[2 * i], s[2 * i + 1]);
}
static inline void st(double *s, size_t i, cplx v) {
  s[2 * i] = creal(v);
  s[2 * i + 1] = cimag(v);
}
static inline cplx polar1(double t) { return CMPLX(cos(t), sin(t)); }
static inline void put(double *m8, int k, cplx v) {
  m8[2 * k] = creal(v);
  m8[2 * k + 1] = cimag(v);
}

/* Row-major 2x2 of every 1-qubit gate (src/reference.cpp:60-98). RX/RY take
 * half angles, RZ = diag(e^{-it/2}, e^{it/2}), T = diag(1, e^{i pi/4}). U3 is
 * the appended general 1-qubit unitary
 *   [[cos(t/2), -e^{i l} sin(t/2)], [e^{i p} sin(t/2), e^{i(p+l)} cos(t/2)]]. */
int qo_gate_matrix(int kind, const double *p, double *m8) {
  const double r = 1.0 / sqrt(2.0);
  cplx m[4] = {1, 0, 0, 1};
  switch (kind) {
    case QO_X: m[0] = 0; m[1] = 1; m[2] = 1; m[3] = 0; break;
    case QO_Y: m[0] = 0; m[1] = CMPLX(0, -1); m[2] = CMPLX(0, 1); m[3] = 0; break;
    case QO_Z: m[3] = -1; break;
    case QO_H: m[0] = r; m[1] = r; m[2] = r; m[3] = -r; break;
    case QO_S: m[3] = CMPLX(0, 1); break;
    case QO_SDG: m[3] = CMPLX(0, -1); break;
    case QO_T: m[3] = polar1(M_PI / 4); break;
    case QO_TDG: m[3] = polar1(-M_PI / 4); break;
    case QO_RX: {
      double c = cos(p[0] / 2), s = sin(p[0] / 2);
      m[0] = c; m[1] = CMPLX(0, -s); m[2] = CMPLX(0, -s); m[3] = c;
      break;
    }
    case QO_RY: {
      double c = cos(p[0] / 2), s = sin(p[0] / 2);
      m[0] = c; m[1] = -s; m[2] = s; m[3] = c;
      break;
    }
    case QO_RZ: m[0] = polar1(-p[0] / 2); m[3] = polar1(p[0] / 2); break;
    case QO_P: m[3] = polar1(p[0]); break;
    case QO_U3: {
      double c = cos(p[0] / 2), s = sin(p[0] / 2);
      m[0] = c;
      m[1] = -polar1(p[2]) * s;
      m[2] = polar1(p[1]) * s;
      m[3] = polar1(p[1] + p[2]) * c;
      break;
    }
    default: return -1;
  }
  for (int k = 0; k < 4; k++) put(m8, k, m[k]);
  return 0;
}

static void unpack(const double *m8, cplx m[4]) {
  for (int k = 0; k < 4; k++) m[k] = CMPLX(m8[2 * k], m8[2 * k + 1]);
}

void qo_reset(double *state, uint32_t n) {
  memset(state, 0, sizeof(double) * 2 * ((size_t)1 << n));
  state[0] = 1.0;
}

/* Pairs (i, i | 2^q) enumerated by inserting a zero bit at q
 * (src/reference.cpp:113-123). */
void qo_apply_1q(double *s, uint32_t n, uint32_t q, const double *m8) {
  cplx m[4];
  unpack(m8, m);
  const size_t stride = (size_t)1 << q, pairs = ((size_t)1 << n) >> 1;
  PAR
  for (size_t t = 0; t < pairs; t++) {
    size_t i = ((t >> q) << (q + 1)) | (t & (stride - 1)), j = i | stride;
    cplx a = ld(s, i), b = ld(s, j);
    st(s, i, m[0] * a + m[1] * b);
    st(s, j, m[2] * a + m[3] * b);
  }
}

/* 2x2 on the target where the control bit is set, each pair visited from its
 * target-0 side (src/reference.cpp:125-136). */
void qo_apply_controlled(double *s, uint32_t n, uint32_t c, uint32_t t,
                         const double *m8) {
  cplx m[4];
  unpack(m8, m);
  const size_t cm = (size_t)1 << c, tm = (size_t)1 << t, dim = (size_t)1 << n;
  PAR
  for (size_t i = 0; i < dim; i++) {
    if (!(i & cm) || (i & tm)) continue;
    size_t j = i | tm;
    cplx a = ld(s, i), b = ld(s, j);
    st(s, i, m[0] * a + m[1] * b);
    st(s, j, m[2] * a + m[3] * b);
  }
}

void qo_apply_diag2q(double *s, uint32_t n, uint32_t a, uint32_t b,
                     double ph_re, double ph_im) {
  const size_t ma = (size_t)1 << a, mb = (size_t)1 << b, dim = (size_t)1 << n;
  const cplx ph = CMPLX(ph_re, ph_im);
  PAR
  for (size_t i = 0; i < dim; i++)
    if ((i & ma) && (i & mb)) st(s, i, ld(s, i) * ph);
}

void qo_apply_swap(double *s, uint32_t n, uint32_t a, uint32_t b) {
  if (a == b) return;
  const size_t ma = (size_t)1 << a, mb = (size_t)1 << b, dim = (size_t)1 << n;
  PAR
  for (size_t i = 0; i < dim; i++) {
    if ((i & ma) && !(i & mb)) {
      size_t j = (i ^ ma) | mb;
      cplx x = ld(s, i);
      st(s, i, ld(s, j));
      st(s, j, x);
    }
  }
}

void qo_apply_ccx(double *s, uint32_t n, uint32_t c0, uint32_t c1,
                  uint32_t t) {
  const size_t m0 = (size_t)1 << c0, m1 = (size_t)1 << c1, mt = (size_t)1 << t;
  const size_t dim = (size_t)1 << n;
  PAR
  for (size_t i = 0; i < dim; i++) {
    if ((i & m0) && (i & m1) && !(i & mt)) {
      cplx x = ld(s, i);
      st(s, i, ld(s, i | mt));
      st(s, i | mt, x);
    }
  }
}

/* Gate dispatch, failing closed (src/reference.cpp:161-201). */
int qo_apply_gate(double *s, uint32_t n, int kind, const uint32_t *q,
                  uint32_t nq, const double *params) {
  double m8[8];
  for (uint32_t k = 0; k < nq; k++)
    if (q[k] >= n) return -1;
  if (nq == 1) {
    if (qo_gate_matrix(kind, params, m8) != 0) return -1;
    qo_apply_1q(s, n, q[0], m8);
    return 0;
  }
  if (nq == 2) {
    if (q[0] == q[1]) return -1;
    switch (kind) {
      case QO_CX:
      case QO_CY:
        qo_gate_matrix(kind == QO_CX ? QO_X : QO_Y, NULL, m8);
        qo_apply_controlled(s, n, q[0], q[1], m8);
        return 0;
      case QO_CZ: qo_apply_diag2q(s, n, q[0], q[1], -1.0, 0.0); return 0;
      case QO_CP:
        qo_apply_diag2q(s, n, q[0], q[1], cos(params[0]), sin(params[0]));
        return 0;
      case QO_SWAP: qo_apply_swap(s, n, q[0], q[1]); return 0;
      default: return -1;
    }
  }
  if (nq == 3 && kind == QO_CCX) {
    if (q[0] == q[1] || q[0] == q[2] || q[1] == q[2]) return -1;
    qo_apply_ccx(s, n, q[0], q[1], q[2]);
    return 0;
  }
  return -1;
}

int qo_run(double *s, uint32_t n, size_t ngates, const int32_t *kinds,
           const uint8_t *nq, const uint32_t *qubits, const double *params) {
  for (size_t g = 0; g < ngates; g++)
    if (qo_apply_gate(s, n, kinds[g], qubits + 3 * g, nq[g], params + 3 * g))
      return -1 - (int)g;
  return 0;
}

/* Dense1qGroup / Blocked1q: the 2x2s applied in list order
 * (src/reference.cpp:216-222). */
void qo_apply_group(double *s, uint32_t n, size_t count, const uint32_t *qubits,
                    const double *mats8) {
  for (size_t k = 0; k < count; k++) qo_apply_1q(s, n, qubits[k], mats8 + 8 * k);
}

/* DiagonalLayer: phi(i) = global + sum of angles whose mask is fully set in i,
 * accumulated in double (src/reference.cpp:229-240). */
void qo_apply_diag(double *s, uint32_t n, size_t count, const uint64_t *masks,
                   const double *angles, double global_phase) {
  const size_t dim = (size_t)1 << n;
  PAR
  for (size_t i = 0; i < dim; i++) {
    double phi = global_phase;
    for (size_t t = 0; t < count; t++)
      if (((uint64_t)i & masks[t]) == masks[t]) phi += angles[t];
    st(s, i, ld(s, i) * polar1(phi));
  }
}

double qo_norm(const double *s, uint32_t n) {
  const size_t dim = (size_t)1 << n;
  double acc = 0.0;
  for (size_t i = 0; i < dim; i++) acc += s[2 * i] * s[2 * i] + s[2 * i + 1] * s[2 * i + 1];
  return sqrt(acc);
}

/* <P> with P|i> = c(i)|i ^ flip>, coefficient taken at the source index
 * (src/reference.cpp:264-284). */
double qo_expect_pauli(const double *s, uint32_t n, uint64_t x, uint64_t y,
                       uint64_t z) {
  const size_t dim = (size_t)1 << n;
  const uint64_t flip = x | y, sign = y | z;
  const unsigned ny = (unsigned)__builtin_popcountll(y);
  cplx acc = 0;
  for (size_t i = 0; i < dim; i++) {
    size_t j = i ^ flip;
    double sg = (__builtin_popcountll((uint64_t)j & sign) & 1) ? -1.0 : 1.0;
    acc += conj(ld(s, i)) * ld(s, j) * sg;
  }
  switch (ny & 3u) {
    case 0: return creal(acc);
    case 1: return -cimag(acc);
    case 2: return -creal(acc);
    default: return cimag(acc);
  }
}

/* splitmix64 step (src/metal/simulator.mm:75-80). */
uint64_t qo_splitmix64(uint64_t *st_) {
  uint64_t z = (*st_ += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

/* Two-level sampler (src/metal/simulator.mm:677-763): per-1024-amplitude
 * block sums, a double CDF over blocks, u = (r >> 11) * 2^-53 * total, binary
 * search for the first block whose CDF exceeds u, then a walk inside it that
 * clamps to the last non-zero amplitude when round-off leaves it short. */
void qo_sample(const double *s, uint32_t n, size_t shots, uint64_t seed,
               uint64_t *out) {
  const size_t dim = (size_t)1 << n, blk = 1024;
  const size_t nb = (dim + blk - 1) / blk;
  double *cdf = (double *)malloc(sizeof(double) * nb);
  double run = 0.0;
  for (size_t b = 0; b < nb; b++) {
    double acc = 0.0;
    size_t end = (b + 1) * blk < dim ? (b + 1) * blk : dim;
    for (size_t i = b * blk; i < end; i++)
      acc += s[2 * i] * s[2 * i] + s[2 * i + 1] * s[2 * i + 1];
    run += acc;
    cdf[b] = run;
  }
  const double total = run;
  uint64_t rng = seed;
  for (size_t k = 0; k < shots; k++) {
    if (total <= 0.0) { out[k] = 0; continue; }
    double u = (double)(qo_splitmix64(&rng) >> 11) * (1.0 / 9007199254740992.0) * total;
    size_t lo = 0, hi = nb - 1;
    while (lo < hi) {
      size_t mid = (lo + hi) >> 1;
      if (cdf[mid] > u) hi = mid; else lo = mid + 1;
    }
    double rem = u - (lo > 0 ? cdf[lo - 1] : 0.0);
    size_t start = lo * blk, end = start + blk < dim ? start + blk : dim;
    double acc = 0.0;
    int64_t last = -1;
    size_t chosen = end - 1;
    int found = 0;
    for (size_t i = start; i < end; i++) {
      double pr = s[2 * i] * s[2 * i] + s[2 * i + 1] * s[2 * i + 1];
      if (pr > 0.0) last = (int64_t)i;
      acc += pr;
      if (acc > rem) { chosen = i; found = 1; break; }
    }
    if (!found) chosen = last >= 0 ? (size_t)last : end - 1;
    out[k] = chosen;
  }
  free(cdf);
}

/* (P psi) accumulated into lam with weight coeff (c * i^ny folded in). */
static void pauli_accum(double *lam, const double *psi, uint32_t n, uint64_t x,
                        uint64_t y, uint64_t z, cplx coeff) {
  const size_t dim = (size_t)1 << n;
  const uint64_t flip = x | y, sign = y | z;
  for (size_t i = 0; i < dim; i++) {
    size_t j = i ^ flip;
    double sg = (__builtin_popcountll((uint64_t)j & sign) & 1) ? -1.0 : 1.0;
    st(lam, i, ld(lam, i) + coeff * sg * ld(psi, j));
  }
}

static cplx ipow(unsigned k) {
  switch (k & 3u) {
    case 0: return 1;
    case 1: return CMPLX(0, 1);
    case 2: return -1;
    default: return CMPLX(0, -1);
  }
}

/* U^dagger of a gate (src/metal/simulator.mm:786-799), U3 appended:
 * U3(t, p, l)^dagger = U3(-t, -l, -p). */
static void invert(int kind, const double *p, int *ik, double *ip) {
  *ik = kind;
  ip[0] = p[0]; ip[1] = p[1]; ip[2] = p[2];
  switch (kind) {
    case QO_RX: case QO_RY: case QO_RZ: case QO_P: case QO_CP: ip[0] = -p[0]; break;
    case QO_S: *ik = QO_SDG; break;
    case QO_SDG: *ik = QO_S; break;
    case QO_T: *ik = QO_TDG; break;
    case QO_TDG: *ik = QO_T; break;
    case QO_U3: ip[0] = -p[0]; ip[1] = -p[2]; ip[2] = -p[1]; break;
    default: break;
  }
}

/* Adjoint sweep (src/metal/simulator.mm:814-963) in complex128:
 * lambda = H psi, E = Re<psi|lambda>, then for k = G-1..0 take
 * grad[k] = Im<lambda|P_k|psi> before undoing gate k on both states. */
int qo_adjoint(uint32_t n, size_t ng, const int32_t *kinds, const uint8_t *nq,
               const uint8_t *npar, const uint32_t *qubits, const double *params,
               uint32_t nterms, const double *coeffs, const uint64_t *xm,
               const uint64_t *ym, const uint64_t *zm, double *work, double *energy,
               double *grad) {
  const size_t dim = (size_t)1 << n;
  /* simulator.mm:822-828: a gate carrying a parameter needs a generator. */
  for (size_t g = 0; g < ng; g++) {
    int k = kinds[g];
    if (npar[g] > 0 && !(k == QO_RX || k == QO_RY || k == QO_RZ)) return -1;
  }
  double *psi = work, *lam = work + 2 * dim;
  qo_reset(psi, n);
  if (qo_run(psi, n, ng, kinds, nq, qubits, params)) return -1;
  memset(lam, 0, sizeof(double) * 2 * dim);
  for (uint32_t t = 0; t < nterms; t++)
    pauli_accum(lam, psi, n, xm[t], ym[t], zm[t],
                coeffs[t] * ipow((unsigned)__builtin_popcountll(ym[t])));
  double e = 0.0;
  for (size_t i = 0; i < dim; i++) e += creal(conj(ld(psi, i)) * ld(lam, i));
  if (energy) *energy = e;
  for (size_t g = ng; g-- > 0;) {
    int k = kinds[g];
    grad[g] = 0.0;
    /* simulator.mm:917: only gates with num_params > 0 get a gradient. */
    if (npar[g] > 0 && (k == QO_RX || k == QO_RY || k == QO_RZ)) {
      uint64_t bit = 1ull << qubits[3 * g];
      uint64_t x = k == QO_RX ? bit : 0, y = k == QO_RY ? bit : 0,
               z = k == QO_RZ ? bit : 0;
      uint64_t flip = x | y, sign = y | z;
      cplx acc = 0;
      for (size_t i = 0; i < dim; i++) {
        size_t j = i ^ flip;
        double sg = (__builtin_popcountll((uint64_t)j & sign) & 1) ? -1.0 : 1.0;
        acc += conj(ld(lam, i)) * ld(psi, j) * sg;
      }
      grad[g] = cimag(acc * ipow((unsigned)__builtin_popcountll(y)));
    }
    int ik;
    double ip[3];
    invert(k, params + 3 * g, &ik, ip);
    qo_apply_gate(psi, n, ik, qubits + 3 * g, nq[g], ip);
    qo_apply_gate(lam, n, ik, qubits + 3 * g, nq[g], ip);
  }
  return 0;
}
