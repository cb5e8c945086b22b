// batch.hip -- persistent, graph-captured batched energy + adjoint gradient.
//
// Programs: each instance becomes a list of BOps:
//   INIT                        psi = |0...0>
//   forward plan ops            GROUP (K independent 2x2s), CTRL (2x2 under a
//                               control mask), SWAP, DIAG (FP64 phase terms)
//   ACCUM per Pauli flip        lam (+)= sum_t c_t i^ny_t s_t(j) psi_j
//   ENERGY                      E = Re<psi|lam>, FP64 block reduction
//   backward, last gate first   ROT: FP64 <lam|P|psi> then the inverse
//                               rotation on both states; other gates: the
//                               inverse on both states
// The semantics are those of the single-circuit adjoint (reference
// src/metal/simulator.mm:814-963); results are per gate, 0 for gates without
// a parameter.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <numeric>

#include "batch.h"

namespace qmamd {
namespace {

void check(hipError_t e, const char *what) {
  if (e != hipSuccess) fail(std::string("batch: ") + what + ": " + hipGetErrorString(e));
}

enum : uint8_t { B_INIT = 0, B_GROUP, B_CTRL, B_SWAP, B_DIAG, B_ACCUM, B_ENERGY, B_ROT };
constexpr uint8_t F_BOTH = 1, F_ASSIGN = 2;

struct BOp {
  uint8_t code, flags, k, axis;
  uint8_t w[4];
  uint32_t a, b;
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
  uint32_t count;
  uint32_t lds_bits;
  uint64_t slot;
  float2 *scratch;
  uint32_t *counter;
  double *energy;
  double *grad;
};

constexpr int kMaxThreads = 1024;

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

// Workgroup sum of two doubles; the totals are valid in thread 0.
__device__ void block_sum2(double &a, double &b, double (*red)[kMaxThreads / 64]) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    a += __shfl_down(a, o, 64);
    b += __shfl_down(b, o, 64);
  }
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    red[0][w] = a;
    red[1][w] = b;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double s0 = 0.0, s1 = 0.0;
    for (int k = 0; k < int(blockDim.x >> 6); k++) {
      s0 += red[0][k];
      s1 += red[1][k];
    }
    a = s0;
    b = s1;
  }
}

template <int K>
__device__ void group(float2 *s, uint32_t dim, const BOp &op, const float *mats) {
  const uint32_t units = dim >> K;
  for (uint32_t u = threadIdx.x; u < units; u += blockDim.x) {
    uint32_t base = u;
#pragma unroll
    for (int j = 0; j < K; j++) base = ins0(base, op.w[j]);
    float2 x[1 << K];
    uint32_t idx[1 << K];
#pragma unroll
    for (int l = 0; l < (1 << K); l++) {
      uint32_t o = base;
#pragma unroll
      for (int j = 0; j < K; j++)
        if ((l >> j) & 1) o |= 1u << op.w[j];
      idx[l] = o;
      x[l] = s[o];
    }
#pragma unroll
    for (int j = 0; j < K; j++) {
      const float *m = mats + op.a + 8 * j;
#pragma unroll
      for (int l = 0; l < (1 << K); l++)
        if (!((l >> j) & 1)) rot2(m, x[l], x[l | (1 << j)]);
    }
#pragma unroll
    for (int l = 0; l < (1 << K); l++) s[idx[l]] = x[l];
  }
}

__device__ void group_any(float2 *s, uint32_t dim, const BOp &op, const float *mats) {
  switch (op.k) {
    case 1: group<1>(s, dim, op, mats); break;
    case 2: group<2>(s, dim, op, mats); break;
    case 3: group<3>(s, dim, op, mats); break;
    case 4: group<4>(s, dim, op, mats); break;
    default: break;
  }
}

__device__ void ctrl(float2 *s, uint32_t dim, const BOp &op) {
  const uint32_t t = op.w[0], bit = 1u << t, cm = uint32_t(op.mask);
  for (uint32_t u = threadIdx.x; u < dim / 2; u += blockDim.x) {
    const uint32_t i = ins0(u, t);
    if ((i & cm) != cm) continue;
    float2 x = s[i], y = s[i | bit];
    rot2(op.m, x, y);
    s[i] = x;
    s[i | bit] = y;
  }
}

__device__ void swap2(float2 *s, uint32_t dim, const BOp &op) {
  const uint32_t a = op.w[0], b = op.w[1], lo = a < b ? a : b, hi = a < b ? b : a;
  for (uint32_t u = threadIdx.x; u < dim / 4; u += blockDim.x) {
    const uint32_t base = ins0(ins0(u, lo), hi);
    const uint32_t i = base | (1u << a), j = base | (1u << b);
    const float2 x = s[i];
    s[i] = s[j];
    s[j] = x;
  }
}

__device__ void diag(float2 *s, uint32_t dim, const BOp &op, const BTerm *terms) {
  for (uint32_t i = threadIdx.x; i < dim; i += blockDim.x) {
    double ph = 0.0;
    for (uint32_t t = op.a; t < op.a + op.b; t++) {
      const uint64_t m = terms[t].mask;
      if ((uint64_t(i) & m) == m) ph += terms[t].x;
    }
    float sn, cs;
    sincospif(float(2.0 * (ph - rint(ph))), &sn, &cs);
    const float2 v = s[i];
    s[i] = make_float2(v.x * cs - v.y * sn, v.x * sn + v.y * cs);
  }
}

__device__ void accum(const float2 *psi, float2 *lam, uint32_t dim, const BOp &op,
                      const BTerm *terms) {
  const uint32_t flip = uint32_t(op.mask);
  for (uint32_t i = threadIdx.x; i < dim; i += blockDim.x) {
    const uint32_t j = i ^ flip;
    double cr = 0.0, ci = 0.0;
    for (uint32_t t = op.a; t < op.a + op.b; t++) {
      const double sg = sgn(j, terms[t].mask);
      cr += sg * terms[t].x;
      ci += sg * terms[t].y;
    }
    const float2 y = psi[j];
    float2 v = make_float2(float(cr * y.x - ci * y.y), float(cr * y.y + ci * y.x));
    if (!(op.flags & F_ASSIGN)) {
      const float2 l = lam[i];
      v.x += l.x;
      v.y += l.y;
    }
    lam[i] = v;
  }
}

__device__ __forceinline__ void acc_cj(double &re, double &im, float2 x, float2 y, double s) {
  re += s * (double(x.x) * y.x + double(x.y) * y.y);
  im += s * (double(x.x) * y.y - double(x.y) * y.x);
}

__device__ void energy(const float2 *psi, const float2 *lam, uint32_t dim, const BOp &op,
                       double *out, double (*red)[kMaxThreads / 64]) {
  double re = 0.0, im = 0.0;
  for (uint32_t i = threadIdx.x; i < dim; i += blockDim.x) acc_cj(re, im, psi[i], lam[i], 1.0);
  block_sum2(re, im, red);
  if (threadIdx.x == 0) out[op.a] = re;
}

__device__ void rot_grad(float2 *psi, float2 *lam, uint32_t dim, const BOp &op, double *out,
                         double (*red)[kMaxThreads / 64]) {
  const uint32_t q = op.w[0], bit = 1u << q;
  double re = 0.0, im = 0.0;
  for (uint32_t u = threadIdx.x; u < dim / 2; u += blockDim.x) {
    const uint32_t i0 = ins0(u, q), i1 = i0 | bit;
    float2 p0 = psi[i0], p1 = psi[i1], l0 = lam[i0], l1 = lam[i1];
    if (op.axis == 0) {
      acc_cj(re, im, l0, p1, 1.0);
      acc_cj(re, im, l1, p0, 1.0);
    } else if (op.axis == 1) {
      acc_cj(re, im, l0, p1, -1.0);
      acc_cj(re, im, l1, p0, 1.0);
    } else {
      acc_cj(re, im, l0, p0, 1.0);
      acc_cj(re, im, l1, p1, -1.0);
    }
    rot2(op.m, p0, p1);
    rot2(op.m, l0, l1);
    psi[i0] = p0;
    psi[i1] = p1;
    lam[i0] = l0;
    lam[i1] = l1;
  }
  block_sum2(re, im, red);
  // dE/dt = Im(i^ny S): ny = 1 for Y.
  if (threadIdx.x == 0) out[op.a] = op.axis == 1 ? re : im;
}

__global__ __launch_bounds__(kMaxThreads) void k_batch(BatchArgs A) {
  extern __shared__ float2 lds[];
  __shared__ uint32_t s_idx;
  __shared__ double red[2][kMaxThreads / 64];
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
      const BOp op = A.ops[o];
      switch (op.code) {
        case B_INIT:
          for (uint32_t i = threadIdx.x; i < dim; i += blockDim.x)
            psi[i] = make_float2(i == 0 ? 1.0f : 0.0f, 0.0f);
          break;
        case B_GROUP:
          group_any(psi, dim, op, A.mats);
          if (op.flags & F_BOTH) group_any(lam, dim, op, A.mats);
          break;
        case B_CTRL:
          ctrl(psi, dim, op);
          if (op.flags & F_BOTH) ctrl(lam, dim, op);
          break;
        case B_SWAP:
          swap2(psi, dim, op);
          if (op.flags & F_BOTH) swap2(lam, dim, op);
          break;
        case B_DIAG:
          diag(psi, dim, op, A.terms);
          break;
        case B_ACCUM:
          accum(psi, lam, dim, op, A.terms);
          break;
        case B_ENERGY:
          energy(psi, lam, dim, op, A.energy, red);
          break;
        case B_ROT:
          rot_grad(psi, lam, dim, op, A.grad, red);
          break;
        default:
          break;
      }
      __syncthreads();
    }
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

struct Program {
  std::vector<BOp> ops;
  std::vector<BTerm> terms;
  std::vector<float> mats;

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
};

void compile(const Circuit &c, const Hamiltonian &h, uint32_t energy_slot, uint32_t grad_base,
             Program &pg) {
  if (c.n < 1 || c.n > kBatchMaxQubits)
    fail("batch: circuit width must be 1.." + std::to_string(kBatchMaxQubits));
  check_circuit(c);
  for (const Gate &g : c.gates) {
    const bool param = g.np > 0 || g.kind == K_RX || g.kind == K_RY || g.kind == K_RZ ||
                       g.kind == K_P || g.kind == K_CP || g.kind == K_U3;
    if (param && !is_rot(g.kind))
      fail(std::string("gradient: ") + kind_name(g.kind) +
           " carries a parameter but has no native generator; only rx, ry and rz are "
           "differentiable");
  }
  for (const Term &t : h)
    if (!t.p.valid(c.n)) fail("gradient: invalid Pauli term");
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
    pg.ops[pg.ops.size() - 1].b = uint32_t(pg.terms.size()) - pg.ops.back().a;
  }
  pg.add(B_ENERGY).a = energy_slot;
  for (size_t k = c.gates.size(); k-- > 0;) {
    const Gate &g = c.gates[k];
    const Gate inv = inverse_gate(g);
    if (is_rot(g.kind)) {
      BOp &o = pg.add(B_ROT, F_BOTH);
      o.axis = g.kind == K_RX ? 0 : g.kind == K_RY ? 1 : 2;
      o.w[0] = uint8_t(g.q[0]);
      o.a = grad_base + uint32_t(k);
      cd m[4];
      gate_matrix(inv.kind, inv.p, m);
      narrow(m, o.m);
    } else {
      pg.gate(inv, F_BOTH);
    }
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

struct Batch::Impl {
  int device = 0;
  int cus = 0;
  size_t max_lds = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  BOp *ops = nullptr;
  BInst *insts = nullptr;
  uint32_t *order = nullptr;
  BTerm *terms = nullptr;
  float *mats = nullptr;
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
  hipGraphExec_t exec = nullptr;
  bool captured = false;
  double ms = 0.0;

  void free_all() {
    hipFree(ops);
    hipFree(insts);
    hipFree(order);
    hipFree(terms);
    hipFree(mats);
    hipFree(energy);
    hipFree(grad);
    ops = nullptr;
    insts = nullptr;
    order = nullptr;
    terms = nullptr;
    mats = nullptr;
    energy = nullptr;
    grad = nullptr;
  }

  size_t lds_bytes(const BatchConfig &c) const { return size_t(2) << c.lds_bits << 3; }

  void ensure_scratch(uint32_t wgs) {
    const uint64_t need_slot = max_n > cfg.lds_bits ? (uint64_t(1) << max_n) : 0;
    if (need_slot == 0) return;
    if (scratch && slot >= need_slot && scratch_wgs >= wgs) return;
    if (scratch) check(hipFree(scratch), "hipFree");
    scratch = nullptr;
    slot = need_slot;
    scratch_wgs = wgs;
    check(hipMalloc(reinterpret_cast<void **>(&scratch), 2 * slot * wgs * sizeof(float2)),
          "hipMalloc(scratch)");
  }

  void capture() {
    if (exec) {
      check(hipGraphExecDestroy(exec), "hipGraphExecDestroy");
      exec = nullptr;
    }
    const uint32_t wgs = uint32_t(cus) * cfg.per_cu;
    ensure_scratch(wgs);
    BatchArgs A{};
    A.ops = ops;
    A.insts = insts;
    A.order = order;
    A.terms = terms;
    A.mats = mats;
    A.count = count;
    A.lds_bits = cfg.lds_bits;
    A.slot = slot;
    A.scratch = scratch;
    A.counter = counter;
    A.energy = energy;
    A.grad = grad;
    const size_t lds = lds_bytes(cfg);
    check(hipFuncSetAttribute(reinterpret_cast<const void *>(k_batch),
                              hipFuncAttributeMaxDynamicSharedMemorySize, int(lds)),
          "hipFuncSetAttribute");
    hipGraph_t graph;
    check(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
    k_reset<<<1, 64, 0, stream>>>(counter);
    k_batch<<<wgs, cfg.threads, lds, stream>>>(A);
    check(hipStreamEndCapture(stream, &graph), "hipStreamEndCapture");
    check(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0), "hipGraphInstantiate");
    check(hipGraphDestroy(graph), "hipGraphDestroy");
    captured = true;
  }

  void launch() {
    if (!captured) capture();
    check(hipEventRecord(ev0, stream), "hipEventRecord");
    check(hipGraphLaunch(exec, stream), "hipGraphLaunch");
    check(hipEventRecord(ev1, stream), "hipEventRecord");
    check(hipStreamSynchronize(stream), "GPU error");
    float t = 0.0f;
    check(hipEventElapsedTime(&t, ev0, ev1), "hipEventElapsedTime");
    ms = t;
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
  b->cfg.lds_bits = b->max_lds >= (size_t(2) << 13 << 3) ? 13 : 12;
  check(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking), "hipStreamCreate");
  check(hipEventCreate(&b->ev0), "hipEventCreate");
  check(hipEventCreate(&b->ev1), "hipEventCreate");
  check(hipMalloc(reinterpret_cast<void **>(&b->counter), 64), "hipMalloc(counter)");
}

Batch::~Batch() {
  Impl *b = impl_.get();
  if (!b) return;
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  if (b->exec) hipGraphExecDestroy(b->exec);
  b->free_all();
  hipFree(b->scratch);
  hipFree(b->counter);
  if (b->ev0) hipEventDestroy(b->ev0);
  if (b->ev1) hipEventDestroy(b->ev1);
  if (b->stream) hipStreamDestroy(b->stream);
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
  b->max_n = 0;
  for (size_t i = 0; i < circuits.size(); i++) {
    BInst in{};
    in.n = circuits[i]->n;
    in.op_begin = uint32_t(pg.ops.size());
    compile(*circuits[i], *hams[i], uint32_t(i), grad_base, pg);
    in.op_end = uint32_t(pg.ops.size());
    insts.push_back(in);
    cost.push_back(double(in.op_end - in.op_begin) * double(uint64_t(1) << in.n));
    grad_base += uint32_t(circuits[i]->gates.size());
    b->max_n = std::max(b->max_n, in.n);
  }
  // Largest first: the persistent workgroups then drain a queue whose tail
  // is made of the cheapest instances.
  std::vector<uint32_t> order(insts.size());
  std::iota(order.begin(), order.end(), 0u);
  std::stable_sort(order.begin(), order.end(),
                   [&](uint32_t x, uint32_t y) { return cost[x] > cost[y]; });
  check(hipStreamSynchronize(b->stream), "hipStreamSynchronize");
  b->free_all();
  upload(b->ops, pg.ops);
  upload(b->insts, insts);
  upload(b->order, order);
  upload(b->terms, pg.terms);
  upload(b->mats, pg.mats);
  b->count = uint32_t(insts.size());
  b->ngrad = grad_base;
  check(hipMalloc(reinterpret_cast<void **>(&b->energy), std::max<size_t>(1, b->count) * 8),
        "hipMalloc(energy)");
  check(hipMalloc(reinterpret_cast<void **>(&b->grad), std::max<size_t>(1, b->ngrad) * 8),
        "hipMalloc(grad)");
  // Gates without a parameter keep gradient 0: the kernel never writes them.
  check(hipMemset(b->grad, 0, std::max<size_t>(1, b->ngrad) * 8), "hipMemset(grad)");
  b->captured = false;
}

void Batch::set_config(const BatchConfig &c) {
  Impl *b = impl_.get();
  if (c.threads < 64 || c.threads > uint32_t(kMaxThreads) || c.threads % 64)
    fail("batch: threads must be a multiple of 64 up to 1024");
  if (c.per_cu < 1 || c.per_cu > 4) fail("batch: per_cu must be 1..4");
  if (c.lds_bits > 14 || b->lds_bytes(c) > b->max_lds) fail("batch: LDS cut-off too large");
  b->cfg = c;
  b->captured = false;
}

BatchConfig Batch::config() const { return impl_->cfg; }

BatchConfig Batch::autotune() {
  Impl *b = impl_.get();
  if (!b->count) return b->cfg;
  BatchConfig best = b->cfg;
  double best_ms = 1e300;
  for (uint32_t bits : {10u, 11u, 12u, 13u}) {
    BatchConfig c;
    c.lds_bits = bits;
    if (b->lds_bytes(c) > b->max_lds) continue;
    for (uint32_t per : {1u, 2u, 4u}) {
      // All persistent workgroups of a CU must fit its 160 KiB of LDS.
      if (per * b->lds_bytes(c) > (size_t(160) << 10)) continue;
      for (uint32_t th : {256u, 512u, 1024u}) {
        c.per_cu = per;
        c.threads = th;
        set_config(c);
        b->launch();
        b->launch();
        if (b->ms < best_ms) {
          best_ms = b->ms;
          best = c;
        }
      }
    }
  }
  set_config(best);
  return best;
}

void Batch::run(double *energies, double *grads) {
  Impl *b = impl_.get();
  check(hipSetDevice(b->device), "hipSetDevice");
  if (!b->count) return;
  b->launch();
  if (energies)
    check(hipMemcpy(energies, b->energy, b->count * sizeof(double), hipMemcpyDeviceToHost),
          "hipMemcpy(energy)");
  if (grads && b->ngrad)
    check(hipMemcpy(grads, b->grad, b->ngrad * sizeof(double), hipMemcpyDeviceToHost),
          "hipMemcpy(grad)");
}

double Batch::last_ms() const { return impl_->ms; }
size_t Batch::total_gates() const { return impl_->ngrad; }

}  // namespace qmamd
