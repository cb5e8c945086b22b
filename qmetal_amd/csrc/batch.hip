// batch.hip -- persistent, graph-captured batched energy + adjoint gradient.
//
// Programs: each instance becomes a list of BOps:
//   INIT                        psi = |0...0>
//   forward plan ops            GROUP (K independent 2x2s), CTRL (2x2 under a
//                               control mask), SWAP, DIAG (FP64 phase terms)
//   ACCUM per Pauli flip        lam (+)= sum_t c_t i^ny_t s_t(j) psi_j
//   ENERGY                      E = Re<psi|lam>, FP64 block reduction
//   backward sweep              segments of the gate stream, last first:
//     ROTGROUP  up to 3 one-qubit gates on distinct wires: for each, the FP64
//               overlap <lam|P|psi> (rotations) then its inverse, on both
//               states held in registers
//     DIAGRUN   a run of diagonal blocks (diagonal gates and CX.RZ.CX
//               triples): the FP64 overlaps Im<lam|Z...Z|psi> of its
//               rotations, then one inverse phase pass on both states
//     CTRL/SWAP any other gate: its inverse on both states
//
// The semantics are those of the single-circuit adjoint (reference
// src/metal/simulator.mm:814-963): grad[k] = Im<lam_k|P_k|psi_k> with both
// states after gate k and every later gate undone. Gates that commute can be
// undone in any order without changing these overlaps -- <U'lam|P|U'psi> =
// <lam|P|psi> whenever U' commutes with P -- which is what lets a segment
// take all its overlaps at one point: one-qubit gates on distinct wires
// commute, diagonal blocks commute, and CX(a,b) Z_b CX(a,b) = Z_a Z_b gives
// the generator of the rotation inside a CX.RZ.CX triple. Results are per
// gate, 0 for gates without a parameter.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <numeric>
#include <string>

#include "batch.h"
#include "simulator.h"

namespace qmamd {
namespace {

void check(hipError_t e, const char *what) {
  if (e != hipSuccess) fail(std::string("batch: ") + what + ": " + hipGetErrorString(e));
}

enum : uint8_t {
  B_INIT = 0, B_GROUP, B_CTRL, B_SWAP, B_DIAG, B_ACCUM, B_ENERGY, B_ROTGROUP, B_DIAGRUN
};
constexpr uint8_t F_BOTH = 1, F_ASSIGN = 2;
constexpr uint32_t kNoAxis = 3;
constexpr int kOverlapChunk = 8;
// Wires per ROTGROUP: both states' 2^3 amplitudes and 6 FP64 accumulators
// stay in registers within the 128 VGPRs a 1024-thread workgroup allows.
constexpr size_t kRotGroup = 3;

struct BOp {
  uint8_t code, flags, k, pad;
  uint8_t w[4];
  uint32_t a, b;  // group: matrices offset; diag/accum/diagrun: term offset, count
  uint32_t c, d;  // rotgroup: aux offset; diagrun: overlap offset, count
  uint64_t mask;
  float m[8];
};

struct BInst {
  uint32_t n, op_begin, op_end, pad;
};

struct BTerm {
  uint64_t mask;
  double x, y;
};

struct BatchArgs {
  const BOp *ops;
  const BInst *insts;
  const uint32_t *order;
  const BTerm *terms;
  const float *mats;
  const uint32_t *aux;
  uint32_t count;
  uint32_t lds_bits;
  uint64_t slot;
  float2 *scratch;
  uint32_t *counter;
  double *energy;
  double *grad;
};

constexpr int kMaxThreads = 512;
constexpr int kWaves = kMaxThreads / 64;

__device__ __forceinline__ uint32_t ins0(uint32_t x, uint32_t p) {
  return ((x >> p) << (p + 1)) | (x & ((1u << p) - 1u));
}

__device__ __forceinline__ void rot2(const float *m, float2 &x, float2 &y) {
  const float2 u = x, v = y;
  x.x = m[0] * u.x - m[1] * u.y + m[2] * v.x - m[3] * v.y;
  x.y = m[0] * u.y + m[1] * u.x + m[2] * v.y + m[3] * v.x;
  y.x = m[4] * u.x - m[5] * u.y + m[6] * v.x - m[7] * v.y;
  y.y = m[4] * u.y + m[5] * u.x + m[6] * v.y + m[7] * v.x;
}

__device__ __forceinline__ double sgn(uint32_t j, uint64_t sign) {
  return (__popcll(uint64_t(j) & sign) & 1) ? -1.0 : 1.0;
}

// Workgroup sums of N doubles; totals valid in thread 0.
template <int N>
__device__ void block_sum(double (&v)[N], double (*red)[kWaves]) {
#pragma unroll
  for (int i = 0; i < N; i++)
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v[i] += __shfl_down(v[i], o, 64);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0)
#pragma unroll
    for (int i = 0; i < N; i++) red[i][w] = v[i];
  __syncthreads();
  if (threadIdx.x == 0)
#pragma unroll
    for (int i = 0; i < N; i++) {
      double s = 0.0;
      for (int k = 0; k < int(blockDim.x >> 6); k++) s += red[i][k];
      v[i] = s;
    }
  __syncthreads();
}

__device__ __forceinline__ void acc_cj(double &re, double &im, float2 x, float2 y, double s) {
  re += s * (double(x.x) * y.x + double(x.y) * y.y);
  im += s * (double(x.x) * y.y - double(x.y) * y.x);
}

// Every pass below handles several units per thread per iteration, loading
// all of them before any arithmetic, so that a workgroup keeps enough loads
// in flight to stream a state that lives in HBM. A pass covers chunk c of nc
// equal chunks of its units: the whole state (c = 0, nc = 1) when one
// workgroup runs the instance, a slice when several workgroups share it.

struct Range {
  uint32_t lo, hi;
};
__device__ __forceinline__ Range chunk_of(uint32_t units, uint32_t c, uint32_t nc) {
  return Range{uint32_t(uint64_t(units) * c / nc), uint32_t(uint64_t(units) * (c + 1) / nc)};
}

// The units of an op a workgroup takes, by local index v < count: the
// contiguous range [base, base + count) (CMap), or, in a high chunk, every
// unit whose unit-space bits [lo_pos, lo_pos + mbits) equal `base` (HMap: the
// bits of v below lo_pos as they are, the rest moved above the fixed ones).
struct CMap {
  uint32_t base, count;
  __device__ __forceinline__ uint32_t operator()(uint32_t v) const { return base + v; }
};
struct HMap {
  uint32_t base, count, lo_pos, mbits;
  __device__ __forceinline__ uint32_t operator()(uint32_t v) const {
    const uint32_t low = v & ((1u << lo_pos) - 1u);
    return low | (base << lo_pos) | ((v >> lo_pos) << (lo_pos + mbits));
  }
};

// Where a workgroup works: chunk c of nc contiguous ones, or (high) chunk c
// of the 2^mbits chunks that fix index bits [kHighLow, kHighLow + mbits).
constexpr uint32_t kHighLow = 6;
struct Geo {
  uint32_t c, nc, mbits;
  bool high;
};

// Unit map of an op with `units` units whose moved wires are w[0..k): the
// contiguous chunk (HIGH false) or the high chunk of the geometry.
template <bool HIGH>
__device__ __forceinline__ auto unit_map(const Geo &g, uint32_t units, const uint8_t *w, int k) {
  if constexpr (!HIGH) {
    const Range rg = chunk_of(units, g.c, g.nc);
    return CMap{rg.lo, rg.hi - rg.lo};
  } else {
    uint32_t kl = 0;  // moved wires below the fixed bits
    for (int j = 0; j < k; j++) kl += w[j] < kHighLow ? 1u : 0u;
    return HMap{g.c, units >> g.mbits, kHighLow - kl, g.mbits};
  }
}

template <int K, typename M>
__device__ void group(float2 *s, M m, const BOp &op, const float *mats) {
  constexpr int U = K >= 4 ? 1 : 16 >> K;  // units per thread per iteration
  for (uint32_t v0 = threadIdx.x; v0 < m.count; v0 += U * blockDim.x) {
    float2 x[U][1 << K];
    uint32_t idx[U][1 << K];
#pragma unroll
    for (int r = 0; r < U; r++) {
      const uint32_t v = v0 + r * blockDim.x;
      if (v >= m.count) continue;
      uint32_t base = m(v);
#pragma unroll
      for (int j = 0; j < K; j++) base = ins0(base, op.w[j]);
#pragma unroll
      for (int l = 0; l < (1 << K); l++) {
        uint32_t o = base;
#pragma unroll
        for (int j = 0; j < K; j++)
          if ((l >> j) & 1) o |= 1u << op.w[j];
        idx[r][l] = o;
        x[r][l] = s[o];
      }
    }
#pragma unroll
    for (int r = 0; r < U; r++) {
      if (v0 + r * blockDim.x >= m.count) continue;
#pragma unroll
      for (int j = 0; j < K; j++) {
        const float *mt = mats + op.a + 8 * j;
#pragma unroll
        for (int l = 0; l < (1 << K); l++)
          if (!((l >> j) & 1)) rot2(mt, x[r][l], x[r][l | (1 << j)]);
      }
#pragma unroll
      for (int l = 0; l < (1 << K); l++) s[idx[r][l]] = x[r][l];
    }
  }
}

template <bool HIGH>
__device__ void group_any(float2 *s, uint32_t dim, const Geo &g, const BOp &op,
                          const float *mats) {
  const auto m = unit_map<HIGH>(g, dim >> op.k, op.w, op.k);
  switch (op.k) {
    case 1: group<1>(s, m, op, mats); break;
    case 2: group<2>(s, m, op, mats); break;
    case 3: group<3>(s, m, op, mats); break;
    case 4: group<4>(s, m, op, mats); break;
    default: break;
  }
}

// Up to three one-qubit gates on distinct wires, on both states: for each
// wire, the overlap <lam|P|psi> of its rotation (if any), then its inverse.
// aux holds (axis, grad slot) per wire; axis 0/1/2 = X/Y/Z, kNoAxis = none.
// With `part` the workgroup's sums go to part[j] (one chunk's share) instead
// of the gradient.
template <int K, typename M>
__device__ void rotgroup(float2 *psi, float2 *lam, M m, const BOp &op, const float *mats,
                         const uint32_t *aux, double *out, double *part,
                         double (*red)[kWaves]) {
  constexpr int U = 8 >> K;  // units per thread per iteration
  double acc[2 * K];
#pragma unroll
  for (int i = 0; i < 2 * K; i++) acc[i] = 0.0;
  for (uint32_t v0 = threadIdx.x; v0 < m.count; v0 += U * blockDim.x) {
    float2 p[U][1 << K], l[U][1 << K];
    uint32_t idx[U][1 << K];
#pragma unroll
    for (int r = 0; r < U; r++) {
      const uint32_t v = v0 + r * blockDim.x;
      if (v >= m.count) continue;
      uint32_t base = m(v);
#pragma unroll
      for (int j = 0; j < K; j++) base = ins0(base, op.w[j]);
#pragma unroll
      for (int t = 0; t < (1 << K); t++) {
        uint32_t o = base;
#pragma unroll
        for (int j = 0; j < K; j++)
          if ((t >> j) & 1) o |= 1u << op.w[j];
        idx[r][t] = o;
        p[r][t] = psi[o];
        l[r][t] = lam[o];
      }
    }
#pragma unroll
    for (int r = 0; r < U; r++) {
      if (v0 + r * blockDim.x >= m.count) continue;
#pragma unroll
      for (int j = 0; j < K; j++) {
        const uint32_t axis = aux[op.c + 2 * j];
#pragma unroll
        for (int t = 0; t < (1 << K); t++) {
          if ((t >> j) & 1) continue;
          const int t1 = t | (1 << j);
          if (axis == 0) {         // X
            acc_cj(acc[2 * j], acc[2 * j + 1], l[r][t], p[r][t1], 1.0);
            acc_cj(acc[2 * j], acc[2 * j + 1], l[r][t1], p[r][t], 1.0);
          } else if (axis == 1) {  // Y, before its factor i
            acc_cj(acc[2 * j], acc[2 * j + 1], l[r][t], p[r][t1], -1.0);
            acc_cj(acc[2 * j], acc[2 * j + 1], l[r][t1], p[r][t], 1.0);
          } else if (axis == 2) {  // Z
            acc_cj(acc[2 * j], acc[2 * j + 1], l[r][t], p[r][t], 1.0);
            acc_cj(acc[2 * j], acc[2 * j + 1], l[r][t1], p[r][t1], -1.0);
          }
          const float *mt = mats + op.a + 8 * j;
          rot2(mt, p[r][t], p[r][t1]);
          rot2(mt, l[r][t], l[r][t1]);
        }
      }
#pragma unroll
      for (int t = 0; t < (1 << K); t++) {
        psi[idx[r][t]] = p[r][t];
        lam[idx[r][t]] = l[r][t];
      }
    }
  }
  block_sum<2 * K>(acc, red);
  if (threadIdx.x == 0)
#pragma unroll
    for (int j = 0; j < K; j++) {
      const uint32_t axis = aux[op.c + 2 * j];
      // dE/dt = Im(i^ny S): ny = 1 for Y.
      const double v = axis == 1 ? acc[2 * j] : acc[2 * j + 1];
      if (part) part[j] = v;
      else if (axis != kNoAxis) out[aux[op.c + 2 * j + 1]] = v;
    }
}

template <bool HIGH>
__device__ void rotgroup_any(float2 *psi, float2 *lam, uint32_t dim, const Geo &g, const BOp &op,
                             const float *mats, const uint32_t *aux, double *out, double *part,
                             double (*red)[kWaves]) {
  const auto m = unit_map<HIGH>(g, dim >> op.k, op.w, op.k);
  switch (op.k) {
    case 1: rotgroup<1>(psi, lam, m, op, mats, aux, out, part, red); break;
    case 2: rotgroup<2>(psi, lam, m, op, mats, aux, out, part, red); break;
    case 3: rotgroup<3>(psi, lam, m, op, mats, aux, out, part, red); break;
    default: break;
  }
}

constexpr int kElems = 8;  // elements per thread per iteration of the flat passes

template <typename M>
__device__ void ctrl(float2 *s, M m, const BOp &op) {
  const uint32_t t = op.w[0], bit = 1u << t, cm = uint32_t(op.mask);
  for (uint32_t v0 = threadIdx.x; v0 < m.count; v0 += kElems * blockDim.x) {
    float2 x[kElems], y[kElems];
    uint32_t idx[kElems];
    bool on[kElems];
#pragma unroll
    for (int r = 0; r < kElems; r++) {
      const uint32_t v = v0 + r * blockDim.x;
      const uint32_t i = ins0(m(v), t);
      idx[r] = i;
      on[r] = v < m.count && (i & cm) == cm;
      if (on[r]) {
        x[r] = s[i];
        y[r] = s[i | bit];
      }
    }
#pragma unroll
    for (int r = 0; r < kElems; r++) {
      if (!on[r]) continue;
      rot2(op.m, x[r], y[r]);
      s[idx[r]] = x[r];
      s[idx[r] | bit] = y[r];
    }
  }
}

template <typename M>
__device__ void swap2(float2 *s, M m, const BOp &op) {
  const uint32_t a = op.w[0], b = op.w[1], lo = a < b ? a : b, hi = a < b ? b : a;
  for (uint32_t v0 = threadIdx.x; v0 < m.count; v0 += kElems * blockDim.x) {
    float2 x[kElems], y[kElems];
    uint32_t ia[kElems], ib[kElems];
#pragma unroll
    for (int r = 0; r < kElems; r++) {
      const uint32_t v = v0 + r * blockDim.x;
      if (v >= m.count) continue;
      const uint32_t base = ins0(ins0(m(v), lo), hi);
      ia[r] = base | (1u << a);
      ib[r] = base | (1u << b);
      x[r] = s[ia[r]];
      y[r] = s[ib[r]];
    }
#pragma unroll
    for (int r = 0; r < kElems; r++) {
      if (v0 + r * blockDim.x >= m.count) continue;
      s[ia[r]] = y[r];
      s[ib[r]] = x[r];
    }
  }
}

__device__ __forceinline__ float2 cmulf(float2 v, float2 e) {
  return make_float2(v.x * e.x - v.y * e.y, v.x * e.y + v.y * e.x);
}

// Phase of a diagonal op, sum_t angle_t [mask_t within i] (turns), over a
// block of 2^13 aligned indices: the bits above the block are fixed, so every
// term folds into one of four tables over the block's low bits (A = bits
// 0..6, B = bits 7..12). Masks hold at most two bits (one- and two-qubit
// diagonal gates), so a term with bits in both A and B has exactly one in B:
//   phase(i) = c + a[i_A] + b[i_B] + sum over set bits r of i_B of x[r][i_A].
// The tables come from the block's coefficients: the summed angle of every
// low-bit mask of weight <= 2 over the terms whose high bits hold.
constexpr int kTabA = 7, kTabB = 6;
constexpr uint32_t kPhaseBlock = 1u << (kTabA + kTabB);
constexpr int kBits = kTabA + kTabB;
constexpr int kCoefs = 1 + kBits + kBits * (kBits - 1) / 2;  // masks of weight 0, 1, 2
struct PhaseTab {
  double a[1 << kTabA];
  double b[1 << kTabB];
  double x[kTabB][1 << kTabA];
  double c;
  double coef[kCoefs];  // [0] weight 0, [1 + q] bit q, then pairs q < r
};

// Coefficient slot of a low mask of weight <= 2 (pairs in row-major q < r).
__host__ __device__ constexpr int pair_slot(int q, int r) {
  return 1 + kBits + q * (2 * kBits - q - 1) / 2 + (r - q - 1);
}

// All threads: the tables of terms [ta, ta + nt) for the block holding index
// `base`; ends with a barrier.
__device__ void build_tab(PhaseTab &T, uint32_t base, const BTerm *terms, uint32_t ta,
                          uint32_t nt) {
  const uint64_t hi = base & ~(kPhaseBlock - 1u);
  // One thread per coefficient, terms in order (deterministic sums).
  for (int k = int(threadIdx.x); k < kCoefs; k += int(blockDim.x)) {
    uint32_t want = 0;
    if (k >= 1 && k <= kBits) {
      want = 1u << (k - 1);
    } else if (k > kBits) {
      int q = 0, rem = k - 1 - kBits;  // inverse of pair_slot
      while (rem >= kBits - 1 - q) {
        rem -= kBits - 1 - q;
        q++;
      }
      want = (1u << q) | (1u << (q + 1 + rem));
    }
    double sum = 0.0;
    for (uint32_t t = ta; t < ta + nt; t++) {
      const uint64_t m = terms[t].mask;
      const uint64_t mh = m & ~uint64_t(kPhaseBlock - 1u);
      if ((hi & mh) == mh && uint32_t(m & (kPhaseBlock - 1u)) == want) sum += terms[t].x;
    }
    T.coef[k] = sum;
  }
  __syncthreads();
  // Entry e: 0..127 a, 128..191 b, 192..959 x, 960 c.
  constexpr uint32_t kEntries = (1u << kTabA) + (1u << kTabB) + kTabB * (1u << kTabA) + 1;
  for (uint32_t e = threadIdx.x; e < kEntries; e += blockDim.x) {
    double sum = 0.0;
    if (e < 128) {  // bits and pairs within A
      for (int q = 0; q < kTabA; q++) {
        if (!((e >> q) & 1)) continue;
        sum += T.coef[1 + q];
        for (int r = q + 1; r < kTabA; r++)
          if ((e >> r) & 1) sum += T.coef[pair_slot(q, r)];
      }
      T.a[e] = sum;
    } else if (e < 192) {  // bits and pairs within B
      const uint32_t bv = e - 128;
      for (int q = 0; q < kTabB; q++) {
        if (!((bv >> q) & 1)) continue;
        sum += T.coef[1 + kTabA + q];
        for (int r = q + 1; r < kTabB; r++)
          if ((bv >> r) & 1) sum += T.coef[pair_slot(kTabA + q, kTabA + r)];
      }
      T.b[bv] = sum;
    } else if (e < 960) {  // pairs (q in A, bit r of B)
      const uint32_t r = (e - 192) >> kTabA, av = (e - 192) & 127u;
      for (int q = 0; q < kTabA; q++)
        if ((av >> q) & 1) sum += T.coef[pair_slot(q, kTabA + int(r))];
      T.x[r][av] = sum;
    } else {
      T.c = T.coef[0];
    }
  }
  __syncthreads();
}

__device__ __forceinline__ float2 phase_tab(const PhaseTab &T, uint32_t i) {
  const uint32_t a = i & 127u, b = (i >> kTabA) & 63u;
  double ph = T.c + T.a[a] + T.b[b];
#pragma unroll
  for (int r = 0; r < kTabB; r++)
    if ((b >> r) & 1) ph += T.x[r][a];
  const float f = float(ph - rint(ph));
  return make_float2(__builtin_amdgcn_cosf(f), __builtin_amdgcn_sinf(f));
}

// Block by block over the range: the block's tables, then its elements.
__device__ void diag(float2 *s, Range rg, const BOp &op, const BTerm *terms, PhaseTab &T) {
  for (uint32_t blk = rg.lo & ~(kPhaseBlock - 1u); blk < rg.hi; blk += kPhaseBlock) {
    build_tab(T, blk, terms, op.a, op.b);
    const uint32_t lo = blk > rg.lo ? blk : rg.lo;
    const uint32_t end = blk + kPhaseBlock < rg.hi ? blk + kPhaseBlock : rg.hi;
    for (uint32_t i0 = lo + threadIdx.x; i0 < end; i0 += kElems * blockDim.x) {
      float2 x[kElems];
#pragma unroll
      for (int r = 0; r < kElems; r++) {
        const uint32_t i = i0 + r * blockDim.x;
        if (i < end) x[r] = s[i];
      }
#pragma unroll
      for (int r = 0; r < kElems; r++) {
        const uint32_t i = i0 + r * blockDim.x;
        if (i < end) s[i] = cmulf(x[r], phase_tab(T, i));
      }
    }
    __syncthreads();  // the tables are rebuilt for the next block
  }
}

// Overlap sums of the masks [c0, c0 + kOverlapChunk) of the run from one
// thread's elements; block-summed into part[] or the gradient.
__device__ __forceinline__ void overlap_out(double (&acc)[kOverlapChunk], const BOp &op,
                                            const BTerm *terms, uint32_t c0, double *out,
                                            double *part, double (*red)[kWaves]) {
  block_sum<kOverlapChunk>(acc, red);
  if (threadIdx.x == 0)
    for (uint32_t r = 0; r < kOverlapChunk && c0 + r < op.d; r++) {
      if (part) part[c0 + r] = acc[r];
      else out[uint32_t(terms[op.c + c0 + r].x)] = acc[r];
    }
}

// Overlaps Im<lam|Z_m|psi> for each Z-string mask of the run (one read pass
// per kOverlapChunk masks), then the run's inverse phase on both states,
// block by block. With `part` the workgroup's sums go to part[r] for overlap
// r.
__device__ void diagrun(float2 *psi, float2 *lam, Range rg, const BOp &op, const BTerm *terms,
                        double *out, double *part, double (*red)[kWaves], PhaseTab &T) {
  const uint32_t dim = rg.hi;
  for (uint32_t c0 = 0; c0 < op.d; c0 += kOverlapChunk) {
    double acc[kOverlapChunk];
    uint64_t zm[kOverlapChunk];
#pragma unroll
    for (int r = 0; r < kOverlapChunk; r++) {
      acc[r] = 0.0;
      zm[r] = c0 + r < op.d ? terms[op.c + c0 + r].mask : 0;
    }
    for (uint32_t i0 = rg.lo + threadIdx.x; i0 < dim; i0 += kElems * blockDim.x) {
      float2 l[kElems], p[kElems];
#pragma unroll
      for (int e = 0; e < kElems; e++) {
        const uint32_t i = i0 + e * blockDim.x;
        if (i < dim) {
          l[e] = lam[i];
          p[e] = psi[i];
        }
      }
#pragma unroll
      for (int e = 0; e < kElems; e++) {
        const uint32_t i = i0 + e * blockDim.x;
        if (i >= dim) continue;
        const double w = double(l[e].x) * p[e].y - double(l[e].y) * p[e].x;  // Im(conj(l) p)
#pragma unroll
        for (int r = 0; r < kOverlapChunk; r++) acc[r] += sgn(i, zm[r]) * w;
      }
    }
    overlap_out(acc, op, terms, c0, out, part, red);
  }
  for (uint32_t blk = rg.lo & ~(kPhaseBlock - 1u); blk < rg.hi; blk += kPhaseBlock) {
    build_tab(T, blk, terms, op.a, op.b);
    const uint32_t lo = blk > rg.lo ? blk : rg.lo;
    const uint32_t end = blk + kPhaseBlock < rg.hi ? blk + kPhaseBlock : rg.hi;
    for (uint32_t i0 = lo + threadIdx.x; i0 < end; i0 += kElems * blockDim.x) {
      float2 l[kElems], p[kElems];
#pragma unroll
      for (int e = 0; e < kElems; e++) {
        const uint32_t i = i0 + e * blockDim.x;
        if (i < end) {
          l[e] = lam[i];
          p[e] = psi[i];
        }
      }
#pragma unroll
      for (int e = 0; e < kElems; e++) {
        const uint32_t i = i0 + e * blockDim.x;
        if (i >= end) continue;
        const float2 ph = phase_tab(T, i);
        psi[i] = cmulf(p[e], ph);
        lam[i] = cmulf(l[e], ph);
      }
    }
    __syncthreads();
  }
}

__device__ void accum(const float2 *psi, float2 *lam, Range rg, const BOp &op,
                      const BTerm *terms) {
  const uint32_t flip = uint32_t(op.mask), dim = rg.hi;
  for (uint32_t i0 = rg.lo + threadIdx.x; i0 < dim; i0 += kElems * blockDim.x) {
    float2 y[kElems], l[kElems];
#pragma unroll
    for (int e = 0; e < kElems; e++) {
      const uint32_t i = i0 + e * blockDim.x;
      if (i >= dim) continue;
      y[e] = psi[i ^ flip];
      if (!(op.flags & F_ASSIGN)) l[e] = lam[i];
    }
#pragma unroll
    for (int e = 0; e < kElems; e++) {
      const uint32_t i = i0 + e * blockDim.x;
      if (i >= dim) continue;
      const uint32_t j = i ^ flip;
      double cr = 0.0, ci = 0.0;
      for (uint32_t t = op.a; t < op.a + op.b; t++) {
        const double sg = sgn(j, terms[t].mask);
        cr += sg * terms[t].x;
        ci += sg * terms[t].y;
      }
      float2 v = make_float2(float(cr * y[e].x - ci * y[e].y), float(cr * y[e].y + ci * y[e].x));
      if (!(op.flags & F_ASSIGN)) {
        v.x += l[e].x;
        v.y += l[e].y;
      }
      lam[i] = v;
    }
  }
}

__device__ void energy(const float2 *psi, const float2 *lam, Range rg, const BOp &op,
                       double *out, double *part, double (*red)[kWaves]) {
  const uint32_t dim = rg.hi;
  double v[2] = {0.0, 0.0};
  for (uint32_t i0 = rg.lo + threadIdx.x; i0 < dim; i0 += kElems * blockDim.x) {
    float2 p[kElems], l[kElems];
#pragma unroll
    for (int e = 0; e < kElems; e++) {
      const uint32_t i = i0 + e * blockDim.x;
      if (i < dim) {
        p[e] = psi[i];
        l[e] = lam[i];
      }
    }
#pragma unroll
    for (int e = 0; e < kElems; e++)
      if (i0 + e * blockDim.x < dim) acc_cj(v[0], v[1], p[e], l[e], 1.0);
  }
  block_sum<2>(v, red);
  if (threadIdx.x == 0) {
    if (part) part[0] = v[0];
    else out[op.a] = v[0];
  }
}

// One program op on chunk c of nc of the instance's states; `part` (spread
// launches) takes the workgroup's reduction sums instead of the outputs.
// HIGH: a high-chunk step, which holds only groups, rotation groups,
// controlled 2x2s and swaps.
template <bool HIGH>
__device__ void run_op(const BatchArgs &A, const BOp &op, float2 *psi, float2 *lam, uint32_t dim,
                       const Geo &g, double *part, double (*red)[kWaves], PhaseTab &T) {
  const uint32_t c = g.c, nc = g.nc;  // flat ops: contiguous chunks only
  if constexpr (HIGH) {
    switch (op.code) {
      case B_GROUP:
        group_any<true>(psi, dim, g, op, A.mats);
        if (op.flags & F_BOTH) group_any<true>(lam, dim, g, op, A.mats);
        break;
      case B_CTRL:
        ctrl(psi, unit_map<true>(g, dim / 2, op.w, 1), op);
        if (op.flags & F_BOTH) ctrl(lam, unit_map<true>(g, dim / 2, op.w, 1), op);
        break;
      case B_SWAP:
        swap2(psi, unit_map<true>(g, dim / 4, op.w, 2), op);
        if (op.flags & F_BOTH) swap2(lam, unit_map<true>(g, dim / 4, op.w, 2), op);
        break;
      case B_ROTGROUP:
        rotgroup_any<true>(psi, lam, dim, g, op, A.mats, A.aux, A.grad, part, red);
        break;
      default:
        break;
    }
    return;
  }
  switch (op.code) {
    case B_INIT: {
      const Range rg = chunk_of(dim, c, nc);
      for (uint32_t i = rg.lo + threadIdx.x; i < rg.hi; i += blockDim.x)
        psi[i] = make_float2(i == 0 ? 1.0f : 0.0f, 0.0f);
      break;
    }
    case B_GROUP:
      group_any<false>(psi, dim, g, op, A.mats);
      if (op.flags & F_BOTH) group_any<false>(lam, dim, g, op, A.mats);
      break;
    case B_CTRL:
      ctrl(psi, unit_map<false>(g, dim / 2, op.w, 1), op);
      if (op.flags & F_BOTH) ctrl(lam, unit_map<false>(g, dim / 2, op.w, 1), op);
      break;
    case B_SWAP:
      swap2(psi, unit_map<false>(g, dim / 4, op.w, 2), op);
      if (op.flags & F_BOTH) swap2(lam, unit_map<false>(g, dim / 4, op.w, 2), op);
      break;
    case B_DIAG:
      diag(psi, chunk_of(dim, c, nc), op, A.terms, T);
      break;
    case B_ACCUM:
      accum(psi, lam, chunk_of(dim, c, nc), op, A.terms);
      break;
    case B_ENERGY:
      energy(psi, lam, chunk_of(dim, c, nc), op, A.energy, part, red);
      break;
    case B_ROTGROUP:
      rotgroup_any<false>(psi, lam, dim, g, op, A.mats, A.aux, A.grad, part, red);
      break;
    case B_DIAGRUN:
      diagrun(psi, lam, chunk_of(dim, c, nc), op, A.terms, A.grad, part, red, T);
      break;
    default:
      break;
  }
}

__global__ __launch_bounds__(kMaxThreads) void k_batch(BatchArgs A) {
  extern __shared__ float2 lds[];
  __shared__ uint32_t s_idx;
  __shared__ double red[kOverlapChunk][kWaves];
  __shared__ PhaseTab tab;
  for (;;) {
    __syncthreads();
    if (threadIdx.x == 0) s_idx = atomicAdd(A.counter, 1u);
    __syncthreads();
    const uint32_t idx = s_idx;
    if (idx >= A.count) return;
    const BInst in = A.insts[A.order[idx]];
    const uint32_t dim = 1u << in.n;
    float2 *psi, *lam;
    if (in.n <= A.lds_bits) {
      psi = lds;
      lam = lds + dim;
    } else {
      psi = A.scratch + uint64_t(blockIdx.x) * 2 * A.slot;
      lam = psi + A.slot;
    }
    for (uint32_t o = in.op_begin; o < in.op_end; o++) {
      run_op<false>(A, A.ops[o], psi, lam, dim, Geo{0, 1, 0, false}, nullptr, red, tab);
      __syncthreads();
    }
  }
}

// Reduction sums an op leaves (one chunk's share each).
__host__ __device__ inline uint32_t op_slots(const BOp &op) {
  return op.code == B_ROTGROUP ? op.k
         : op.code == B_DIAGRUN ? op.d
         : op.code == B_ENERGY  ? 1u
                                : 0u;
}

// Spread instances (states in HBM, several workgroups each) run in lockstep:
// one launch per step, workgroup w taking chunk w - wg_base[l] of local
// instance l = wg_inst[w]. A step is one op, or a run of consecutive ops
// whose every unit stays inside one chunk (ops [steps[step_off[l] + j],
// steps[step_off[l] + j + 1]) for step j), applied one after another on the
// workgroup's chunk. Reductions leave one row of `slots` partial sums per
// workgroup (the step's ops' sums side by side); k_spread_reduce adds each
// instance's rows in chunk order (deterministic) into the energies and
// gradients.
constexpr uint32_t kStepHigh = 0x80000000u;  // step boundary flag: high chunks
constexpr uint32_t kStepOps = 0x7FFFFFFFu;

struct SpreadArgs {
  const uint32_t *inst;      // local instance -> instance
  const uint32_t *wg_base;   // local instance -> first workgroup (count + 1 entries)
  const uint32_t *wg_inst;   // workgroup -> local instance
  const uint64_t *state;     // local instance -> offset of psi in scratch (lam follows)
  const uint32_t *step_off;  // local instance -> its first entry in steps
  const uint32_t *nsteps;    // local instance -> its number of steps
  const uint32_t *steps;     // op boundaries (nsteps + 1 entries per instance)
  float2 *scratch;
  double *part;
  uint32_t slots;
};

// One kernel per chunk kind (in one kernel the two loop nests share a
// register allocation and spill): the workgroups of instances whose step is
// of the other kind return at once.
template <bool HIGH>
__global__ __launch_bounds__(kMaxThreads) void k_spread(BatchArgs A, SpreadArgs S, uint32_t step) {
  __shared__ double red[kOverlapChunk][kWaves];
  __shared__ PhaseTab tab;
  const uint32_t l = S.wg_inst[blockIdx.x];
  if (step >= S.nsteps[l]) return;
  const BInst in = A.insts[S.inst[l]];
  const uint32_t c = blockIdx.x - S.wg_base[l], nc = S.wg_base[l + 1] - S.wg_base[l];
  const uint32_t dim = 1u << in.n;
  float2 *psi = S.scratch + S.state[l];
  double *row = S.part + uint64_t(blockIdx.x) * S.slots;
  const uint32_t s0 = S.steps[S.step_off[l] + step], s1 = S.steps[S.step_off[l] + step + 1];
  const uint32_t o0 = s0 & kStepOps, o1 = s1 & kStepOps;
  if (((s0 & kStepHigh) != 0) != HIGH) return;
  Geo g{c, nc, 0, false};
  if constexpr (HIGH) {
    // High chunks: 2^(n-14) of them, each 2^14 amplitudes; the other
    // workgroups of the instance only zero their partial sums.
    const uint32_t mbits = in.n > kHighLow + 8 ? in.n - kHighLow - 8 : 0;
    if (c >= (1u << mbits)) {
      uint32_t ns = 0;
      for (uint32_t o = o0; o < o1; o++) ns += op_slots(A.ops[o]);
      for (uint32_t r = threadIdx.x; r < ns; r += blockDim.x) row[r] = 0.0;
      return;
    }
    g = Geo{c, 1u << mbits, mbits, true};
  }
  for (uint32_t o = o0; o < o1; o++) {
    const BOp op = A.ops[o];
    if (o > o0) __syncthreads();  // the previous op's writes to the chunk
    run_op<HIGH>(A, op, psi, psi + dim, dim, g, row, red, tab);
    row += op_slots(op);
  }
}

__global__ __launch_bounds__(64) void k_spread_reduce(BatchArgs A, SpreadArgs S, uint32_t step) {
  const uint32_t l = blockIdx.x;
  if (step >= S.nsteps[l]) return;
  const uint32_t o0 = S.steps[S.step_off[l] + step] & kStepOps;
  const uint32_t o1 = S.steps[S.step_off[l] + step + 1] & kStepOps;
  const uint32_t w0 = S.wg_base[l], w1 = S.wg_base[l + 1];
  uint32_t off = 0;
  for (uint32_t o = o0; o < o1; o++) {
    const BOp op = A.ops[o];
    const uint32_t nslot = op_slots(op);
    for (uint32_t r = threadIdx.x; r < nslot; r += blockDim.x) {
      double v = 0.0;
      for (uint32_t w = w0; w < w1; w++) v += S.part[uint64_t(w) * S.slots + off + r];
      if (op.code == B_ENERGY) {
        A.energy[op.a] = v;
      } else if (op.code == B_DIAGRUN) {
        A.grad[uint32_t(A.terms[op.c + r].x)] = v;
      } else if (A.aux[op.c + 2 * r] != kNoAxis) {
        A.grad[A.aux[op.c + 2 * r + 1]] = v;
      }
    }
    off += nslot;
  }
}

__global__ void k_reset(uint32_t *counter) {
  if (threadIdx.x == 0) counter[0] = 0u;
}

// --- host-side compilation ---------------------------------------------------

constexpr double kTwoPi = 6.283185307179586476925286766559;
double turns(double rad) {
  const double t = rad / kTwoPi;
  return t - std::rint(t);
}

void narrow(const cd *m, float *out) {
  for (int k = 0; k < 4; k++) {
    out[2 * k] = float(m[k].real());
    out[2 * k + 1] = float(m[k].imag());
  }
}

Gate inverse_gate(const Gate &g) {
  Gate v = g;
  switch (g.kind) {
    case K_RX: case K_RY: case K_RZ: case K_P: case K_CP: v.p[0] = -g.p[0]; break;
    case K_S: v.kind = K_SDG; break;
    case K_SDG: v.kind = K_S; break;
    case K_T: v.kind = K_TDG; break;
    case K_TDG: v.kind = K_T; break;
    case K_U3: v.p[0] = -g.p[0]; v.p[1] = -g.p[2]; v.p[2] = -g.p[1]; break;
    default: break;
  }
  return v;
}

bool is_rot(uint8_t k) { return k == K_RX || k == K_RY || k == K_RZ; }

// CX(a,b) RZ_b(t) CX(a,b) at position i.
bool zz_triple(const Circuit &c, size_t i) {
  if (i + 2 >= c.gates.size()) return false;
  const Gate &x = c.gates[i], &r = c.gates[i + 1], &y = c.gates[i + 2];
  return x.kind == K_CX && y.kind == K_CX && r.kind == K_RZ && x.q[0] == y.q[0] &&
         x.q[1] == y.q[1] && r.q[0] == x.q[1];
}

struct Program {
  std::vector<BOp> ops;
  std::vector<BTerm> terms;
  std::vector<float> mats;
  std::vector<uint32_t> aux;

  BOp &add(uint8_t code, uint8_t flags = 0) {
    BOp o{};
    o.code = code;
    o.flags = flags;
    ops.push_back(o);
    return ops.back();
  }
  void group1(uint32_t q, const cd *m, uint8_t flags) {
    BOp &o = add(B_GROUP, flags);
    o.k = 1;
    o.w[0] = uint8_t(q);
    o.a = uint32_t(mats.size());
    float f[8];
    narrow(m, f);
    mats.insert(mats.end(), f, f + 8);
  }
  void ctrl(uint64_t cmask, uint32_t t, const cd *m, uint8_t flags) {
    BOp &o = add(B_CTRL, flags);
    o.w[0] = uint8_t(t);
    o.mask = cmask;
    narrow(m, o.m);
  }
  void swap(uint32_t a, uint32_t b, uint8_t flags) {
    BOp &o = add(B_SWAP, flags);
    o.w[0] = uint8_t(a);
    o.w[1] = uint8_t(b);
  }
  // A gate (or its inverse) on psi, or on both states.
  void gate(const Gate &g, uint8_t flags) {
    static const cd X[4] = {0.0, 1.0, 1.0, 0.0};
    if (g.nq == 1) {
      cd m[4];
      gate_matrix(g.kind, g.p, m);
      group1(g.q[0], m, flags);
    } else if (g.kind == K_CX || g.kind == K_CY) {
      cd m[4];
      gate_matrix(g.kind == K_CX ? K_X : K_Y, nullptr, m);
      ctrl(1ull << g.q[0], g.q[1], m, flags);
    } else if (g.kind == K_CZ || g.kind == K_CP) {
      const cd ph = g.kind == K_CZ ? cd(-1, 0) : cd(std::cos(g.p[0]), std::sin(g.p[0]));
      const cd m[4] = {1.0, 0.0, 0.0, ph};
      ctrl(1ull << g.q[0], g.q[1], m, flags);
    } else if (g.kind == K_SWAP) {
      swap(g.q[0], g.q[1], flags);
    } else if (g.kind == K_CCX) {
      ctrl((1ull << g.q[0]) | (1ull << g.q[1]), g.q[2], X, flags);
    } else {
      fail(std::string("batch: unsupported gate ") + kind_name(g.kind));
    }
  }
  void plan_op(const PlanOp &p) {
    static const cd X[4] = {0.0, 1.0, 1.0, 0.0};
    switch (p.kind) {
      case OpKind::Dense1q:
        group1(p.q[0], p.m, 0);
        break;
      case OpKind::Group: {
        BOp &o = add(B_GROUP);
        o.k = uint8_t(p.wires.size());
        o.a = uint32_t(mats.size());
        for (size_t j = 0; j < p.wires.size(); j++) {
          o.w[j] = uint8_t(p.wires[j]);
          float f[8];
          narrow(&p.mats[4 * j], f);
          mats.insert(mats.end(), f, f + 8);
        }
        break;
      }
      case OpKind::Blocked:
        for (size_t j = 0; j < p.wires.size(); j++) group1(p.wires[j], &p.mats[4 * j], 0);
        break;
      case OpKind::Controlled:
        ctrl(1ull << p.q[0], p.q[1], p.m, 0);
        break;
      case OpKind::Diagonal: {
        if (p.masks.empty()) break;
        BOp &o = add(B_DIAG);
        o.a = uint32_t(terms.size());
        terms.push_back(BTerm{0, turns(p.global), 0.0});
        for (size_t t = 0; t < p.masks.size(); t++)
          terms.push_back(BTerm{p.masks[t], turns(p.angles[t]), 0.0});
        o.b = uint32_t(terms.size()) - o.a;
        break;
      }
      case OpKind::Swap:
        swap(p.q[0], p.q[1], 0);
        break;
      case OpKind::CCX:
        ctrl((1ull << p.q[0]) | (1ull << p.q[1]), p.q[2], X, 0);
        break;
    }
  }

  // Backward segment: one-qubit gates [i, j) of the forward stream, any
  // wires. Each wire's gates are undone last-first; the wires' remaining
  // last gates form a layer on distinct wires, issued as ROTGROUPs of <= 3.
  void window_backward(const Circuit &c, size_t i, size_t j, uint32_t grad_base) {
    std::map<uint32_t, std::vector<size_t>> per_wire;
    for (size_t k = i; k < j; k++) per_wire[c.gates[k].q[0]].push_back(k);
    for (;;) {
      std::vector<size_t> layer;
      for (auto &kv : per_wire)
        if (!kv.second.empty()) {
          layer.push_back(kv.second.back());
          kv.second.pop_back();
        }
      if (layer.empty()) break;
      for (size_t b = 0; b < layer.size(); b += kRotGroup) {
        BOp &o = add(B_ROTGROUP, F_BOTH);
        const size_t take = std::min<size_t>(kRotGroup, layer.size() - b);
        o.k = uint8_t(take);
        o.a = uint32_t(mats.size());
        o.c = uint32_t(aux.size());
        for (size_t t = 0; t < take; t++) {
          const Gate &g = c.gates[layer[b + t]];
          const Gate inv = inverse_gate(g);
          o.w[t] = uint8_t(g.q[0]);
          cd m[4];
          gate_matrix(inv.kind, inv.p, m);
          float f[8];
          narrow(m, f);
          mats.insert(mats.end(), f, f + 8);
          aux.push_back(g.kind == K_RX ? 0u : g.kind == K_RY ? 1u : g.kind == K_RZ ? 2u : kNoAxis);
          aux.push_back(grad_base + uint32_t(layer[b + t]));
        }
      }
    }
  }

  // Backward segment: a run of diagonal blocks. Inverse phase terms (all
  // angles and the global phase negated) and one Z-string overlap per
  // rotation: Z_q for a diagonal RZ_q, Z_a Z_b for the RZ of a CX.RZ.CX.
  void diagrun_backward(const Circuit &c, const std::vector<std::pair<size_t, size_t>> &blocks,
                        uint32_t grad_base) {
    std::vector<uint64_t> masks;
    std::vector<double> angles;
    double global = 0.0;
    std::vector<std::pair<uint64_t, uint32_t>> overlaps;
    for (const auto &bl : blocks) {
      const size_t k = bl.first;
      if (bl.second == 3) {
        const uint64_t ma = 1ull << c.gates[k].q[0], mb = 1ull << c.gates[k].q[1];
        const double t = c.gates[k + 1].p[0];
        masks.insert(masks.end(), {ma, mb, ma | mb});
        angles.insert(angles.end(), {t, t, -2.0 * t});
        global += -t / 2.0;
        overlaps.push_back({ma | mb, grad_base + uint32_t(k + 1)});
      } else {
        global += diagonal_term(c.gates[k], masks, angles);
        if (c.gates[k].kind == K_RZ)
          overlaps.push_back({1ull << c.gates[k].q[0], grad_base + uint32_t(k)});
      }
    }
    BOp &o = add(B_DIAGRUN, F_BOTH);
    o.a = uint32_t(terms.size());
    terms.push_back(BTerm{0, turns(-global), 0.0});
    for (size_t t = 0; t < masks.size(); t++) terms.push_back(BTerm{masks[t], turns(-angles[t]), 0.0});
    o.b = uint32_t(terms.size()) - o.a;
    o.c = uint32_t(terms.size());
    for (const auto &ov : overlaps) terms.push_back(BTerm{ov.first, double(ov.second), 0.0});
    o.d = uint32_t(overlaps.size());
  }
};

void compile(const Circuit &c, const Hamiltonian &h, uint32_t energy_slot, uint32_t grad_base,
             Program &pg) {
  if (c.n < 1 || c.n > kBatchMaxQubits)
    fail("batch: circuit width must be 1.." + std::to_string(kBatchMaxQubits));
  check_circuit(c);
  for (const Gate &g : c.gates) {
    // Reference simulator.mm:822-828: only a gate that carries a parameter
    // must have a native generator.
    if (g.np > 0 && !is_rot(g.kind))
      fail(std::string("gradient: ") + kind_name(g.kind) +
           " carries a parameter but has no native generator; only rx, ry and rz are "
           "differentiable");
  }
  for (const Term &t : h)
    if (!t.p.valid(c.n)) fail("gradient: invalid Pauli term");
  const size_t op0 = pg.ops.size();
  pg.add(B_INIT);
  for (const PlanOp &p : build_plan(c, FuseOpts{}).ops) pg.plan_op(p);
  std::map<uint64_t, std::vector<const Term *>> by_flip;
  for (const Term &t : h) by_flip[t.p.x | t.p.y].push_back(&t);
  if (by_flip.empty()) by_flip[0];
  bool first = true;
  for (const auto &kv : by_flip) {
    BOp &o = pg.add(B_ACCUM, first ? F_ASSIGN : 0);
    first = false;
    o.mask = kv.first;
    o.a = uint32_t(pg.terms.size());
    for (const Term *t : kv.second) {
      const uint32_t ny = uint32_t(__builtin_popcountll(t->p.y));
      double re = t->c, im = 0.0;
      switch (ny & 3u) {
        case 1: re = 0.0; im = t->c; break;
        case 2: re = -t->c; break;
        case 3: re = 0.0; im = -t->c; break;
        default: break;
      }
      pg.terms.push_back(BTerm{t->p.y | t->p.z, re, im});
    }
    pg.ops.back().b = uint32_t(pg.terms.size()) - pg.ops.back().a;
  }
  pg.add(B_ENERGY).a = energy_slot;

  // Segment the forward stream: runs of diagonal blocks that contain a
  // multi-qubit block, windows of one-qubit gates, single other gates.
  struct Seg {
    int kind;  // 0 window [i, j), 1 diagonal run, 2 single gate i
    size_t i, j;
    std::vector<std::pair<size_t, size_t>> blocks;
  };
  std::vector<Seg> segs;
  const size_t G = c.gates.size();
  auto block_len = [&](size_t k) -> size_t {
    if (zz_triple(c, k)) return 3;
    return is_diagonal(c.gates[k]) ? 1 : 0;
  };
  for (size_t i = 0; i < G;) {
    const Gate &g = c.gates[i];
    if (zz_triple(c, i) || (g.nq == 2 && is_diagonal(g))) {
      Seg s{1, i, i, {}};
      size_t j = i;
      for (size_t l; j < G && (l = block_len(j)) != 0; j += l) s.blocks.push_back({j, l});
      s.j = j;
      segs.push_back(std::move(s));
      i = j;
    } else if (g.nq == 1) {
      size_t j = i;
      while (j < G && c.gates[j].nq == 1) j++;
      segs.push_back(Seg{0, i, j, {}});
      i = j;
    } else {
      segs.push_back(Seg{2, i, i + 1, {}});
      i++;
    }
  }
  for (size_t s = segs.size(); s-- > 0;) {
    const Seg &sg = segs[s];
    if (sg.kind == 0) pg.window_backward(c, sg.i, sg.j, grad_base);
    else if (sg.kind == 1) pg.diagrun_backward(c, sg.blocks, grad_base);
    else pg.gate(inverse_gate(c.gates[sg.i]), F_BOTH);
  }
  // The phase tables take masks of at most two bits (one- and two-qubit
  // diagonal gates are all that reach a phase op).
  for (size_t o = op0; o < pg.ops.size(); o++) {
    const BOp &op = pg.ops[o];
    if (op.code != B_DIAG && op.code != B_DIAGRUN) continue;
    for (uint32_t t = op.a; t < op.a + op.b; t++)
      if (__builtin_popcountll(pg.terms[t].mask) > 2) fail("batch: internal: phase mask weight");
  }
}

template <typename T>
void upload(T *&dst, const std::vector<T> &src) {
  if (dst) check(hipFree(dst), "hipFree");
  dst = nullptr;
  if (src.empty()) return;
  check(hipMalloc(reinterpret_cast<void **>(&dst), src.size() * sizeof(T)), "hipMalloc");
  check(hipMemcpy(dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice), "hipMemcpy");
}

}  // namespace

// Spread launches: each workgroup takes 2^kSpreadChunkBits amplitudes of one
// instance; the states of one group of spread instances live in HBM together
// (at most kSpreadBudget bytes or half the free device memory, groups run one
// after another).
constexpr uint32_t kSpreadChunkBits = 13;
// Widest instances first form batched classes (see Batch::Impl::Class).
constexpr uint32_t kClassMinQubits = 14;
// 256 threads: a spread workgroup (one wave per SIMD) fits beside a
// persistent one on the same CU.
constexpr int kSpreadThreads = 256;
constexpr uint64_t kSpreadBudget = uint64_t(16) << 30;
constexpr uint32_t kFuseOps = 24, kFuseSlots = 64;

struct SpreadGroup {
  uint32_t inst_off = 0;  // into the local-instance arrays (inst, state)
  uint32_t base_off = 0;  // into wg_base (count + 1 entries per group)
  uint32_t wg_off = 0;    // into wg_inst
  uint32_t count = 0;
  std::vector<uint32_t> step_insts;  // per step: instances still running (a prefix)
  std::vector<uint32_t> step_wgs;    // per step: their workgroups (a prefix)
  std::vector<uint8_t> step_reduce;  // per step: some running op reduces
  std::vector<uint8_t> step_kinds;   // per step: bit 0 contiguous, bit 1 high chunks
};

struct Batch::Impl {
  int device = 0;
  int cus = 0;
  size_t max_lds = 0;
  hipStream_t stream = nullptr;
  hipStream_t side = nullptr;  // spread launches, beside the persistent kernel
  hipEvent_t fork = nullptr, join = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  BOp *ops = nullptr;
  BInst *insts = nullptr;
  uint32_t *order = nullptr;
  uint32_t *sample_order = nullptr;
  uint32_t sample_count = 0;
  BTerm *terms = nullptr;
  float *mats = nullptr;
  uint32_t *aux = nullptr;
  float2 *scratch = nullptr;
  size_t scratch_wgs = 0;
  uint64_t slot = 0;
  uint32_t *counter = nullptr;
  double *energy = nullptr;
  double *grad = nullptr;
  uint32_t count = 0;
  uint32_t max_n = 0;
  size_t ngrad = 0;
  BatchConfig cfg;
  // Host copies the spread plan is built from.
  std::vector<BInst> h_insts;
  std::vector<uint32_t> h_order;  // every instance, largest cost first
  std::vector<BOp> h_ops;
  // Gradient slots of rotations added without a parameter: reported as 0,
  // as the reference does for gates with num_params == 0.
  std::vector<uint32_t> h_nograd;
  // Spread plan for cfg.spread_bits (built by build_spread).
  uint32_t built_spread = ~0u;
  // Classes of wide instances (at least kClassMinQubits) that share a width,
  // a gate structure and a Hamiltonian: when spread, a class runs as one
  // batched program of the single-circuit merged passes (every pass one
  // launch over the class, BatchedAdjoint) instead of the lockstep spread
  // steps; a class that does not compile keeps the spread steps.
  struct Class {
    uint32_t n = 0;
    std::vector<uint32_t> members;
    bool tried = false;
    std::unique_ptr<BatchedAdjoint> prog;
  };
  std::vector<Class> classes;
  std::vector<int32_t> class_of;  // per instance; -1: none
  std::vector<Circuit> h_circ;    // per instance in a class, else empty
  std::vector<Hamiltonian> h_ham;
  std::vector<uint32_t> h_gbase;
  std::vector<uint32_t> active;  // classes run batched under cfg.spread_bits
  std::vector<SpreadGroup> groups;
  uint32_t *sp_inst = nullptr, *sp_wg_base = nullptr, *sp_wg_inst = nullptr;
  uint32_t *sp_step_off = nullptr, *sp_nsteps = nullptr, *sp_steps = nullptr;
  uint64_t *sp_state = nullptr;
  float2 *sp_scratch = nullptr;
  double *sp_part = nullptr;
  uint32_t sp_slots = 1;
  uint32_t p_count = 0;  // instances run by the persistent kernel
  hipGraphExec_t exec = nullptr;
  bool captured = false;
  double ms = 0.0;

  void free_spread() {
    (void)hipFree(sp_step_off);
    (void)hipFree(sp_nsteps);
    (void)hipFree(sp_steps);
    sp_step_off = sp_nsteps = sp_steps = nullptr;
    (void)hipFree(sp_inst);
    (void)hipFree(sp_wg_base);
    (void)hipFree(sp_wg_inst);
    (void)hipFree(sp_state);
    (void)hipFree(sp_scratch);
    hipFree(sp_part);
    sp_inst = sp_wg_base = sp_wg_inst = nullptr;
    sp_state = nullptr;
    sp_scratch = nullptr;
    sp_part = nullptr;
    groups.clear();
    built_spread = ~0u;
  }

  // Instances with n >= cfg.spread_bits (none when 0) leave the persistent
  // queue: sorted by program length, cut into groups whose states fit the
  // budget, each instance given 2^(n - kSpreadChunkBits) workgroups (at
  // least one). The persistent kernel keeps the rest, in cost order, and its
  // autotuning sample is every 16th of them.
  // Compiles class k once (a simulator of its width prepares the passes).
  void compile_class(Class &c) {
    if (c.tried) return;
    c.tried = true;
    std::vector<const Circuit *> cs;
    std::vector<uint32_t> gb;
    for (uint32_t i : c.members) {
      cs.push_back(&h_circ[i]);
      gb.push_back(h_gbase[i]);
    }
    check(hipSetDevice(device), "hipSetDevice");
    Simulator sim(c.n);
    c.prog = sim.compile_batched(cs, h_ham[c.members[0]], c.members, gb);
  }

  void build_spread() {
    if (built_spread == cfg.spread_bits) return;
    check(hipStreamSynchronize(stream), "hipStreamSynchronize");
    free_spread();
    active.clear();
    for (size_t k = 0; k < classes.size(); k++) {
      Class &c = classes[k];
      if (!cfg.spread_bits || c.n < cfg.spread_bits) continue;
      compile_class(c);
      if (c.prog) active.push_back(uint32_t(k));
    }
    std::vector<uint32_t> sp, rest;
    for (uint32_t i : h_order) {
      const bool spread = cfg.spread_bits && h_insts[i].n >= cfg.spread_bits;
      if (spread && class_of[i] >= 0 && classes[size_t(class_of[i])].prog) continue;
      (spread ? sp : rest).push_back(i);
    }
    std::vector<uint32_t> sample;
    for (size_t k = 0; k < rest.size(); k += 16) sample.push_back(rest[k]);
    upload(order, rest);
    upload(sample_order, sample);
    p_count = uint32_t(rest.size());
    sample_count = uint32_t(sample.size());
    // Steps of each spread instance: chunk-local ops merge into runs.
    std::map<uint32_t, std::vector<uint32_t>> bounds;
    for (uint32_t i : sp) bounds[i] = spread_steps(h_insts[i]);
    auto nsteps = [&](uint32_t i) { return uint32_t(bounds[i].size() - 1); };
    std::stable_sort(sp.begin(), sp.end(),
                     [&](uint32_t x, uint32_t y) { return nsteps(x) > nsteps(y); });
    // States of one group: kSpreadBudget, or half the free device memory.
    size_t free_b = 0, total_b = 0;
    uint64_t budget = kSpreadBudget;
    if (hipMemGetInfo(&free_b, &total_b) == hipSuccess) budget = std::min<uint64_t>(budget, free_b / 2);
    (void)hipGetLastError();
    std::vector<uint32_t> inst, base, wg_inst, step_off, nst, steps;
    std::vector<uint64_t> state;
    uint64_t scratch_max = 0, wgs_max = 0;
    sp_slots = 1;
    for (size_t k = 0; k < sp.size();) {
      SpreadGroup g;
      g.inst_off = uint32_t(inst.size());
      g.base_off = uint32_t(base.size());
      g.wg_off = uint32_t(wg_inst.size());
      uint64_t elems = 0;
      uint32_t wgs = 0;
      while (k < sp.size()) {
        const BInst &in = h_insts[sp[k]];
        const uint64_t need = uint64_t(2) << in.n;
        if (g.count && (elems + need) * sizeof(float2) > budget) break;
        inst.push_back(sp[k]);
        state.push_back(elems);
        base.push_back(wgs);
        const uint32_t chunks = in.n > kSpreadChunkBits ? 1u << (in.n - kSpreadChunkBits) : 1u;
        for (uint32_t c = 0; c < chunks; c++) wg_inst.push_back(g.count);
        wgs += chunks;
        elems += need;
        const std::vector<uint32_t> &bd = bounds[sp[k]];
        step_off.push_back(uint32_t(steps.size()));
        nst.push_back(uint32_t(bd.size() - 1));
        steps.insert(steps.end(), bd.begin(), bd.end());
        for (size_t j = 0; j + 1 < bd.size(); j++) {
          uint32_t sl = 0;
          for (uint32_t o = bd[j] & kStepOps; o < (bd[j + 1] & kStepOps); o++)
            sl += op_slots(h_ops[o]);
          sp_slots = std::max(sp_slots, sl);
        }
        g.count++;
        k++;
      }
      base.push_back(wgs);
      const uint32_t gsteps = nst[g.inst_off];
      for (uint32_t j = 0; j < gsteps; j++) {
        uint32_t running = 0;
        bool reduce = false;
        uint8_t kinds = 0;
        for (uint32_t l = 0; l < g.count; l++) {
          const uint32_t li = g.inst_off + l;
          if (j >= nst[li]) break;  // sorted: the rest have fewer steps
          running++;
          kinds |= (steps[step_off[li] + j] & kStepHigh) ? 2 : 1;
          for (uint32_t o = steps[step_off[li] + j] & kStepOps;
               o < (steps[step_off[li] + j + 1] & kStepOps); o++)
            reduce = reduce || op_slots(h_ops[o]) > 0;
        }
        g.step_insts.push_back(running);
        g.step_wgs.push_back(base[g.base_off + running]);
        g.step_reduce.push_back(reduce);
        g.step_kinds.push_back(kinds);
      }
      scratch_max = std::max(scratch_max, elems);
      wgs_max = std::max<uint64_t>(wgs_max, wgs);
      groups.push_back(std::move(g));
    }
    if (!groups.empty()) {
      upload(sp_inst, inst);
      upload(sp_wg_base, base);
      upload(sp_wg_inst, wg_inst);
      upload(sp_state, state);
      upload(sp_step_off, step_off);
      upload(sp_nsteps, nst);
      upload(sp_steps, steps);
      check(hipMalloc(reinterpret_cast<void **>(&sp_scratch), scratch_max * sizeof(float2)),
            "hipMalloc(spread states)");
      check(hipMalloc(reinterpret_cast<void **>(&sp_part), wgs_max * sp_slots * sizeof(double)),
            "hipMalloc(spread partials)");
    }
    built_spread = cfg.spread_bits;
  }

  // Whether every unit of `op` on an n-qubit state stays inside one spread
  // chunk (2^kSpreadChunkBits aligned amplitudes): its moved wires lie below
  // the chunk bits (controls and phase masks may be anywhere).
  static bool chunk_local(const BOp &op, uint32_t n) {
    if (n <= kSpreadChunkBits) return true;
    switch (op.code) {
      case B_INIT:
      case B_DIAG:
      case B_DIAGRUN:
      case B_ENERGY:
        return true;
      case B_ACCUM:
        return (op.mask >> kSpreadChunkBits) == 0;
      case B_GROUP:
      case B_ROTGROUP:
        for (uint32_t j = 0; j < op.k; j++)
          if (op.w[j] >= kSpreadChunkBits) return false;
        return true;
      case B_CTRL:
        return op.w[0] < kSpreadChunkBits;
      case B_SWAP:
        return op.w[0] < kSpreadChunkBits && op.w[1] < kSpreadChunkBits;
      default:
        return false;
    }
  }

  // Whether `op` stays inside one high chunk of an n-qubit state (bits
  // [kHighLow, n - 8) fixed, 2^14 amplitudes): a group, rotation group,
  // controlled 2x2 or swap whose moved wires lie below kHighLow or among the
  // top n - 8 .. n - 1.
  static bool high_local(const BOp &op, uint32_t n) {
    if (n <= kSpreadChunkBits) return false;
    auto ok = [&](uint32_t w) { return w < kHighLow || w + 8 >= n; };
    switch (op.code) {
      case B_GROUP:
      case B_ROTGROUP:
        for (uint32_t j = 0; j < op.k; j++)
          if (!ok(op.w[j])) return false;
        return true;
      case B_CTRL:
        return ok(op.w[0]);
      case B_SWAP:
        return ok(op.w[0]) && ok(op.w[1]);
      default:
        return false;
    }
  }

  // Step boundaries of an instance's spread program: runs of ops that all
  // stay inside one contiguous chunk, or all inside one high chunk (flagged
  // kStepHigh on the run's first boundary), at most kFuseOps ops and
  // kFuseSlots reduction sums each; other ops alone. An op that fits both
  // kinds starts a high run only when the next op fits high chunks alone.
  std::vector<uint32_t> spread_steps(const BInst &in) const {
    std::vector<uint32_t> bd;
    uint32_t run = 0, slots = 0;
    int mode = 0;  // of the open run: 1 contiguous, 2 high, 0 a single op
    for (uint32_t o = in.op_begin; o < in.op_end; o++) {
      const bool lo = chunk_local(h_ops[o], in.n), hi = high_local(h_ops[o], in.n);
      const uint32_t sl = op_slots(h_ops[o]);
      const bool room = run < kFuseOps && slots + sl <= kFuseSlots;
      if (run > 0 && room && ((mode == 1 && lo) || (mode == 2 && hi))) {
        run++;
        slots += sl;
        continue;
      }
      int m = lo ? 1 : hi ? 2 : 0;
      if (lo && hi && o + 1 < in.op_end && !chunk_local(h_ops[o + 1], in.n) &&
          high_local(h_ops[o + 1], in.n))
        m = 2;
      bd.push_back(o | (m == 2 ? kStepHigh : 0u));
      mode = m;
      run = 1;
      slots = sl;
    }
    bd.push_back(in.op_end);
    return bd;
  }

  void free_all() {
    free_spread();
    classes.clear();
    class_of.clear();
    h_circ.clear();
    h_ham.clear();
    h_gbase.clear();
    active.clear();
    (void)hipFree(ops);
    (void)hipFree(insts);
    (void)hipFree(order);
    (void)hipFree(sample_order);
    (void)hipFree(terms);
    (void)hipFree(mats);
    (void)hipFree(aux);
    (void)hipFree(energy);
    (void)hipFree(grad);
    ops = nullptr;
    insts = nullptr;
    order = nullptr;
    sample_order = nullptr;
    terms = nullptr;
    mats = nullptr;
    aux = nullptr;
    energy = nullptr;
    grad = nullptr;
  }

  size_t lds_bytes(const BatchConfig &c) const { return size_t(2) << c.lds_bits << 3; }
  // LDS of the persistent kernel beyond the states: phase tables, reductions.
  static constexpr size_t kStaticLds = sizeof(PhaseTab) + sizeof(double) * kOverlapChunk * kWaves + 64;

  void ensure_scratch(uint32_t wgs) {
    uint32_t widest = 0;  // widest instance the persistent kernel runs
    for (const BInst &in : h_insts)
      if (!cfg.spread_bits || in.n < cfg.spread_bits) widest = std::max(widest, in.n);
    const uint64_t need_slot = widest > cfg.lds_bits ? (uint64_t(1) << widest) : 0;
    if (need_slot == 0) return;
    if (scratch && slot >= need_slot && scratch_wgs >= wgs) return;
    check(hipStreamSynchronize(stream), "hipStreamSynchronize");
    if (scratch) check(hipFree(scratch), "hipFree");
    scratch = nullptr;
    slot = need_slot;
    scratch_wgs = wgs;
    check(hipMalloc(reinterpret_cast<void **>(&scratch), 2 * slot * wgs * sizeof(float2)),
          "hipMalloc(scratch)");
  }

  BatchArgs args(bool sample) {
    BatchArgs A{};
    A.ops = ops;
    A.insts = insts;
    A.order = sample ? sample_order : order;
    A.terms = terms;
    A.mats = mats;
    A.aux = aux;
    A.count = sample ? sample_count : p_count;
    A.lds_bits = cfg.lds_bits;
    A.slot = slot;
    A.scratch = scratch;
    A.counter = counter;
    A.energy = energy;
    A.grad = grad;
    return A;
  }

  // The spread groups step by step (the full batch only), then the
  // persistent kernel over the rest. The spread plan and scratch are set up
  // beforehand (build_spread, ensure_scratch): nothing here synchronises, so
  // it can be captured.
  //
  // With both parts present the spread launches go to a second stream that
  // forks from and joins back into the first, so that they run beside the
  // persistent kernel (which keeps its LDS-resident instances compute-bound
  // while the spread steps stream from HBM).
  void enqueue(bool sample) {
    const uint32_t wgs = uint32_t(cus) * cfg.per_cu;
    ensure_scratch(wgs);
    const BatchArgs A = args(sample);
    const bool both = !sample && (!groups.empty() || !active.empty()) && A.count;
    hipStream_t ss = both ? side : stream;
    if (both) {
      check(hipEventRecord(fork, stream), "hipEventRecord");
      check(hipStreamWaitEvent(side, fork, 0), "hipStreamWaitEvent");
    }
    if (!sample)
      for (uint32_t k : active) classes[k].prog->enqueue(ss, energy, grad);
    if (!sample)
      for (const SpreadGroup &g : groups) {
        SpreadArgs S{};
        S.inst = sp_inst + g.inst_off;
        S.wg_base = sp_wg_base + g.base_off;
        S.wg_inst = sp_wg_inst + g.wg_off;
        S.state = sp_state + g.inst_off;
        S.step_off = sp_step_off + g.inst_off;
        S.nsteps = sp_nsteps + g.inst_off;
        S.steps = sp_steps;
        S.scratch = sp_scratch;
        S.part = sp_part;
        S.slots = sp_slots;
        for (uint32_t j = 0; j < g.step_wgs.size(); j++) {
          if (g.step_kinds[j] & 1) k_spread<false><<<g.step_wgs[j], kSpreadThreads, 0, ss>>>(A, S, j);
          if (g.step_kinds[j] & 2) k_spread<true><<<g.step_wgs[j], kSpreadThreads, 0, ss>>>(A, S, j);
          if (g.step_reduce[j]) k_spread_reduce<<<g.step_insts[j], 64, 0, ss>>>(A, S, j);
        }
      }
    if (!A.count) return;
    const size_t lds = lds_bytes(cfg);
    check(hipFuncSetAttribute(reinterpret_cast<const void *>(k_batch),
                              hipFuncAttributeMaxDynamicSharedMemorySize, int(lds)),
          "hipFuncSetAttribute");
    k_reset<<<1, 64, 0, stream>>>(counter);
    k_batch<<<wgs, cfg.threads, lds, stream>>>(A);
    if (both) {
      check(hipEventRecord(join, side), "hipEventRecord");
      check(hipStreamWaitEvent(stream, join, 0), "hipStreamWaitEvent");
    }
  }

  void capture() {
    if (exec) {
      check(hipGraphExecDestroy(exec), "hipGraphExecDestroy");
      exec = nullptr;
    }
    build_spread();
    ensure_scratch(uint32_t(cus) * cfg.per_cu);
    hipGraph_t graph;
    check(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
    enqueue(false);
    check(hipStreamEndCapture(stream, &graph), "hipStreamEndCapture");
    check(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0), "hipGraphInstantiate");
    check(hipGraphDestroy(graph), "hipGraphDestroy");
    captured = true;
  }

  double timed(bool sample) {
    if (sample) {
      build_spread();
      ensure_scratch(uint32_t(cus) * cfg.per_cu);
    } else if (!captured) {
      capture();
    }
    check(hipEventRecord(ev0, stream), "hipEventRecord");
    if (sample) enqueue(true);
    else check(hipGraphLaunch(exec, stream), "hipGraphLaunch");
    check(hipEventRecord(ev1, stream), "hipEventRecord");
    check(hipStreamSynchronize(stream), "GPU error");
    float t = 0.0f;
    check(hipEventElapsedTime(&t, ev0, ev1), "hipEventElapsedTime");
    return t;
  }
};

Batch::Batch() : impl_(new Impl) {
  Impl *b = impl_.get();
  check(hipGetDevice(&b->device), "hipGetDevice");
  hipDeviceProp_t prop;
  check(hipGetDeviceProperties(&prop, b->device), "hipGetDeviceProperties");
  b->cus = prop.multiProcessorCount;
  b->max_lds = prop.sharedMemPerBlock;
  int optin = 0;
  if (hipDeviceGetAttribute(&optin, hipDeviceAttributeSharedMemPerBlockOptin, b->device) ==
          hipSuccess &&
      size_t(optin) > b->max_lds)
    b->max_lds = size_t(optin);
  (void)hipGetLastError();
  b->cfg.lds_bits = b->max_lds >= (size_t(2) << 13 << 3) + Impl::kStaticLds ? 13 : 12;
  b->cfg.threads = 512;
  check(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking), "hipStreamCreate");
  check(hipStreamCreateWithFlags(&b->side, hipStreamNonBlocking), "hipStreamCreate");
  check(hipEventCreateWithFlags(&b->fork, hipEventDisableTiming), "hipEventCreate");
  check(hipEventCreateWithFlags(&b->join, hipEventDisableTiming), "hipEventCreate");
  check(hipEventCreate(&b->ev0), "hipEventCreate");
  check(hipEventCreate(&b->ev1), "hipEventCreate");
  check(hipMalloc(reinterpret_cast<void **>(&b->counter), 64), "hipMalloc(counter)");
}

Batch::~Batch() {
  Impl *b = impl_.get();
  if (!b) return;
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  b->free_all();
  (void)hipFree(b->scratch);
  (void)hipFree(b->counter);
  if (b->ev0) (void)hipEventDestroy(b->ev0);
  if (b->ev1) (void)hipEventDestroy(b->ev1);
  if (b->fork) (void)hipEventDestroy(b->fork);
  if (b->join) (void)hipEventDestroy(b->join);
  if (b->side) (void)hipStreamDestroy(b->side);
  if (b->stream) (void)hipStreamDestroy(b->stream);
}

void Batch::prepare(const std::vector<const Circuit *> &circuits,
                    const std::vector<const Hamiltonian *> &hams) {
  Impl *b = impl_.get();
  if (circuits.size() != hams.size()) fail("batch: circuits and Hamiltonians differ in count");
  check(hipSetDevice(b->device), "hipSetDevice");
  Program pg;
  std::vector<BInst> insts;
  std::vector<double> cost;
  uint32_t grad_base = 0;
  uint32_t max_n = 0;
  std::vector<uint32_t> nograd;
  for (size_t i = 0; i < circuits.size(); i++) {
    BInst in{};
    in.n = circuits[i]->n;
    in.op_begin = uint32_t(pg.ops.size());
    compile(*circuits[i], *hams[i], uint32_t(i), grad_base, pg);
    in.op_end = uint32_t(pg.ops.size());
    insts.push_back(in);
    cost.push_back(double(in.op_end - in.op_begin) * double(uint64_t(1) << in.n));
    for (size_t k = 0; k < circuits[i]->gates.size(); k++)
      if (circuits[i]->gates[k].np == 0 && is_rot(circuits[i]->gates[k].kind))
        nograd.push_back(grad_base + uint32_t(k));
    grad_base += uint32_t(circuits[i]->gates.size());
    max_n = std::max(max_n, in.n);
  }
  // Largest first: the persistent workgroups then drain a queue whose tail
  // is made of the cheapest instances.
  std::vector<uint32_t> order(insts.size());
  std::iota(order.begin(), order.end(), 0u);
  std::stable_sort(order.begin(), order.end(),
                   [&](uint32_t x, uint32_t y) { return cost[x] > cost[y]; });
  check(hipStreamSynchronize(b->stream), "hipStreamSynchronize");
  b->free_all();
  b->max_n = max_n;
  b->h_insts = insts;
  b->h_order = order;
  b->h_ops = pg.ops;
  b->h_nograd = nograd;
  // Classes of wide instances: same width, gate kinds and wires, and
  // Hamiltonian (keyed by their bytes).
  b->class_of.assign(circuits.size(), -1);
  b->h_circ.resize(circuits.size());
  b->h_ham.resize(circuits.size());
  b->h_gbase.assign(circuits.size(), 0);
  {
    std::map<std::string, int32_t> keys;
    uint32_t gb = 0;
    for (size_t i = 0; i < circuits.size(); i++) {
      const Circuit &c = *circuits[i];
      b->h_gbase[i] = gb;
      gb += uint32_t(c.gates.size());
      if (c.n < kClassMinQubits) continue;
      std::string key(reinterpret_cast<const char *>(&c.n), sizeof(c.n));
      for (const Gate &g : c.gates) {
        const uint8_t hdr[3] = {g.kind, g.nq, uint8_t(g.np > 0)};
        key.append(reinterpret_cast<const char *>(hdr), sizeof(hdr));
        key.append(reinterpret_cast<const char *>(g.q), sizeof(g.q));
      }
      key.push_back('|');
      for (const Term &t : *hams[i]) {
        key.append(reinterpret_cast<const char *>(&t.c), sizeof(t.c));
        key.append(reinterpret_cast<const char *>(&t.p), sizeof(t.p));
      }
      auto it = keys.find(key);
      if (it == keys.end()) {
        it = keys.emplace(key, int32_t(b->classes.size())).first;
        b->classes.emplace_back();
        b->classes.back().n = c.n;
      }
      b->class_of[i] = it->second;
      b->classes[size_t(it->second)].members.push_back(uint32_t(i));
      b->h_circ[i] = c;
      b->h_ham[i] = *hams[i];
    }
  }
  upload(b->ops, pg.ops);
  upload(b->insts, insts);
  upload(b->terms, pg.terms);
  upload(b->mats, pg.mats);
  upload(b->aux, pg.aux);
  b->count = uint32_t(insts.size());
  b->ngrad = grad_base;
  check(hipMalloc(reinterpret_cast<void **>(&b->energy), std::max<size_t>(1, b->count) * 8),
        "hipMalloc(energy)");
  check(hipMalloc(reinterpret_cast<void **>(&b->grad), std::max<size_t>(1, b->ngrad) * 8),
        "hipMalloc(grad)");
  // Gates without a parameter keep gradient 0: the kernel never writes them.
  check(hipMemset(b->grad, 0, std::max<size_t>(1, b->ngrad) * 8), "hipMemset(grad)");
  b->captured = false;
  b->build_spread();
}

void Batch::set_config(const BatchConfig &c) {
  Impl *b = impl_.get();
  if (c.threads < 64 || c.threads > uint32_t(kMaxThreads) || c.threads % 64)
    fail("batch: threads must be a multiple of 64 up to 512");
  if (c.per_cu < 1 || c.per_cu > 4) fail("batch: per_cu must be 1..4");
  if (c.lds_bits > 14 || b->lds_bytes(c) + Impl::kStaticLds > b->max_lds)
    fail("batch: LDS cut-off too large");
  const uint32_t spread = b->cfg.spread_bits;
  b->cfg = c;
  b->cfg.spread_bits = spread;
  b->captured = false;
}

void Batch::set_spread(uint32_t bits) {
  Impl *b = impl_.get();
  if (bits > kBatchMaxQubits) fail("batch: spread width must be 0..20");
  b->cfg.spread_bits = bits;
  b->captured = false;
}

BatchConfig Batch::config() const { return impl_->cfg; }

// Times each candidate twice on the sample (every 16th instance by cost) and
// keeps the fastest; the full batch is then captured with it.
BatchConfig Batch::autotune() {
  Impl *b = impl_.get();
  if (!b->count) return b->cfg;
  check(hipSetDevice(b->device), "hipSetDevice");
  // The persistent kernel is tuned on a sample of every instance (nothing
  // spread); the spread width is then picked by timing the whole batch.
  set_spread(0);
  BatchConfig best = b->cfg;
  double best_ms = 1e300;
  for (uint32_t bits : {11u, 12u, 13u}) {
    BatchConfig c;
    c.lds_bits = bits;
    if (b->lds_bytes(c) + Impl::kStaticLds > b->max_lds) continue;
    for (uint32_t per : {1u, 2u}) {
      // All persistent workgroups of a CU must fit its 160 KiB of LDS.
      if (per * (b->lds_bytes(c) + Impl::kStaticLds) > (size_t(160) << 10)) continue;
      for (uint32_t th : {256u, 512u}) {
        if (per * th > 2048) continue;
        c.per_cu = per;
        c.threads = th;
        set_config(c);
        b->timed(true);
        const double ms = b->timed(true);
        if (ms < best_ms) {
          best_ms = ms;
          best = c;
        }
      }
    }
  }
  set_config(best);
  uint32_t best_spread = 0;
  if (b->max_n >= kSpreadMinQubits) {
    double best_full = 1e300;
    for (uint32_t sp : {0u, b->max_n - 1, b->max_n - 3, b->max_n - 5}) {
      if (sp && sp < kSpreadMinQubits) continue;
      set_spread(sp);
      b->timed(false);
      const double ms = b->timed(false);
      if (ms < best_full) {
        best_full = ms;
        best_spread = sp;
      }
    }
  }
  set_spread(best_spread);
  return b->cfg;
}

void Batch::run(double *energies, double *grads) {
  Impl *b = impl_.get();
  check(hipSetDevice(b->device), "hipSetDevice");
  if (!b->count) return;
  b->ms = b->timed(false);
  if (energies)
    check(hipMemcpy(energies, b->energy, b->count * sizeof(double), hipMemcpyDeviceToHost),
          "hipMemcpy(energy)");
  if (grads && b->ngrad) {
    check(hipMemcpy(grads, b->grad, b->ngrad * sizeof(double), hipMemcpyDeviceToHost),
          "hipMemcpy(grad)");
    for (uint32_t k : b->h_nograd) grads[k] = 0.0;
  }
}

double Batch::last_ms() const { return impl_->ms; }

size_t Batch::batched_instances() {
  Impl *b = impl_.get();
  b->build_spread();
  size_t k = 0;
  for (uint32_t c : b->active) k += b->classes[c].members.size();
  return k;
}
size_t Batch::total_gates() const { return impl_->ngrad; }

}  // namespace qmamd
