// plan.cpp -- lowering a gate stream to the Plan the executor runs.
//
// The passes reproduce the reference builder (src/plan.cpp:187-359) op for
// op: the same pass count, op kinds, wires, term order, FP64 matrices and
// global phases, because qm_circuit_plan_size and the dispatch count are part
// of the ABI's observable behaviour.
//
//   pass 1  mark maximal runs of diagonal blocks (a block is one diagonal
//           gate or a CX.RZ.CX triple on the target); keep runs of >= 2
//           blocks and lone triples                     (src/plan.cpp:193-221)
//   pass 2  emit: a marked run becomes one Diagonal op; a window of >= 2
//           unmarked 1q gates becomes per-wire FP64 products, exact
//           identities (atol 1e-12) dropped, survivors ascending and chunked
//           into groups of `group`; anything else is lowered one to one
//                                                       (src/plan.cpp:223-298)
//   pass 3  with block_bits >= 2, merge runs of >= 2 Dense1q/Group ops whose
//           wires are all below block_bits into one Blocked op
//                                                       (src/plan.cpp:305-357)

#include <algorithm>
#include <array>
#include <cmath>

#include "core.h"

namespace qmamd {
namespace {

// out = a * b for row-major 2x2s (b applied first). Reads both operands
// before writing, so `out` may alias `b`.
void mul2(const cd *a, const cd *b, cd *out) {
  const cd r0 = a[0] * b[0] + a[1] * b[2];
  const cd r1 = a[0] * b[1] + a[1] * b[3];
  const cd r2 = a[2] * b[0] + a[3] * b[2];
  const cd r3 = a[2] * b[1] + a[3] * b[3];
  out[0] = r0;
  out[1] = r1;
  out[2] = r2;
  out[3] = r3;
}

// Absolute tolerance only (src/plan.cpp:25-34): a genuine RZ(1e-6) sits 1e-6
// from the identity, algebraic cancellation lands within ~1e-15.
bool near_identity(const cd *m) {
  const double tol = 1e-12;
  return std::abs(m[0] - cd(1, 0)) <= tol && std::abs(m[1]) <= tol &&
         std::abs(m[2]) <= tol && std::abs(m[3] - cd(1, 0)) <= tol;
}

PlanOp dense(uint32_t q, const cd *m) {
  PlanOp op;
  op.kind = OpKind::Dense1q;
  op.q[0] = q;
  std::copy(m, m + 4, op.m);
  return op;
}

// One gate, no fusion (src/plan.cpp:45-90).
PlanOp lower_one(const Gate &g) {
  PlanOp op;
  if (g.nq == 1) {
    cd m[4];
    gate_matrix(g.kind, g.p, m);
    return dense(g.q[0], m);
  }
  if (g.nq == 2) {
    if (g.kind == K_CX || g.kind == K_CY) {
      op.kind = OpKind::Controlled;
      op.q[0] = g.q[0];
      op.q[1] = g.q[1];
      gate_matrix(g.kind == K_CX ? K_X : K_Y, nullptr, op.m);
      return op;
    }
    if (g.kind == K_CZ || g.kind == K_CP) {
      op.kind = OpKind::Diagonal;
      op.global = diagonal_term(g, op.masks, op.angles);
      return op;
    }
    if (g.kind == K_SWAP) {
      op.kind = OpKind::Swap;
      op.q[0] = g.q[0];
      op.q[1] = g.q[1];
      return op;
    }
  } else if (g.nq == 3 && g.kind == K_CCX) {
    op.kind = OpKind::CCX;
    op.q[0] = g.q[0];
    op.q[1] = g.q[1];
    op.q[2] = g.q[2];
    return op;
  }
  fail(std::string("build_plan: unsupported gate ") + kind_name(g.kind));
}

// CX(a,b) RZ_b(t) CX(a,b) at position i (src/plan.cpp:137-157).
bool zz_at(const Circuit &c, size_t i, uint32_t *a, uint32_t *b, double *t) {
  if (i + 2 >= c.gates.size()) return false;
  const Gate &x = c.gates[i], &r = c.gates[i + 1], &y = c.gates[i + 2];
  if (x.kind != K_CX || y.kind != K_CX || r.kind != K_RZ) return false;
  if (x.q[0] != y.q[0] || x.q[1] != y.q[1]) return false;
  if (r.q[0] != x.q[1]) return false;
  *a = x.q[0];
  *b = x.q[1];
  *t = r.p[0];
  return true;
}

size_t block_len(const Circuit &c, size_t i, bool zz) {
  uint32_t a, b;
  double t;
  if (zz && zz_at(c, i, &a, &b, &t)) return 3;
  return is_diagonal(c.gates[i]) ? 1 : 0;
}

// Terms of a ZZ triple: x_a ^ x_b = x_a + x_b - 2 x_a x_b, so (a, t), (b, t),
// (ab, -2t) and a global -t/2 (src/plan.cpp:169-183).
void add_block(const Circuit &c, size_t i, size_t len, PlanOp &op) {
  if (len == 3) {
    uint32_t a, b;
    double t;
    zz_at(c, i, &a, &b, &t);
    const uint64_t ma = 1ull << a, mb = 1ull << b;
    op.masks.push_back(ma);
    op.angles.push_back(t);
    op.masks.push_back(mb);
    op.angles.push_back(t);
    op.masks.push_back(ma | mb);
    op.angles.push_back(-2.0 * t);
    op.global += -t / 2.0;
    return;
  }
  op.global += diagonal_term(c.gates[i], op.masks, op.angles);
}

bool low_1q(const PlanOp &op, uint32_t b) {
  if (op.kind == OpKind::Dense1q) return op.q[0] < b;
  if (op.kind == OpKind::Group) {
    for (uint32_t w : op.wires)
      if (w >= b) return false;
    return true;
  }
  return false;
}

}  // namespace

bool is_diagonal(const Gate &g) {
  switch (g.kind) {
    case K_Z: case K_S: case K_SDG: case K_T: case K_TDG: case K_RZ: case K_P:
      return g.nq == 1;
    case K_CZ: case K_CP:
      return g.nq == 2;
    default:
      return false;
  }
}

// (mask, angle) of a diagonal gate; returns its global phase
// (src/plan.cpp:106-135). RZ = e^{-it/2} diag(1, e^{it}).
double diagonal_term(const Gate &g, std::vector<uint64_t> &masks,
                     std::vector<double> &angles) {
  const uint64_t b0 = 1ull << g.q[0];
  double angle = 0.0, global = 0.0;
  uint64_t mask = b0;
  switch (g.kind) {
    case K_Z: angle = M_PI; break;
    case K_S: angle = M_PI / 2; break;
    case K_SDG: angle = -M_PI / 2; break;
    case K_T: angle = M_PI / 4; break;
    case K_TDG: angle = -M_PI / 4; break;
    case K_P: angle = g.p[0]; break;
    case K_RZ: angle = g.p[0]; global = -g.p[0] / 2.0; break;
    case K_CZ: mask |= 1ull << g.q[1]; angle = M_PI; break;
    case K_CP: mask |= 1ull << g.q[1]; angle = g.p[0]; break;
    default: fail("diagonal_term: gate is not diagonal");
  }
  masks.push_back(mask);
  angles.push_back(angle);
  return global;
}

Plan build_plan(const Circuit &c, const FuseOpts &o) {
  Plan plan;
  plan.n = c.n;
  const size_t G = c.gates.size();

  // Pass 1: marking.
  std::vector<uint8_t> blen(G, 0), marked(G, 0);
  if (o.diagonals) {
    for (size_t i = 0; i < G;) {
      if (block_len(c, i, o.zz) == 0) {
        i++;
        continue;
      }
      size_t j = i, blocks = 0;
      for (size_t l; j < G && (l = block_len(c, j, o.zz)) != 0; j += l) {
        blen[j] = uint8_t(l);
        blocks++;
      }
      if (blocks >= 2 || blen[i] == 3) std::fill(marked.begin() + i, marked.begin() + j, 1);
      i = j;
    }
  }

  // Pass 2: emission.
  for (size_t i = 0; i < G;) {
    if (marked[i]) {
      PlanOp op;
      op.kind = OpKind::Diagonal;
      size_t j = i;
      while (j < G && marked[j]) {
        const size_t len = blen[j] ? blen[j] : 1;
        add_block(c, j, len, op);
        j += len;
      }
      plan.ops.push_back(std::move(op));
      i = j;
      continue;
    }
    if (o.windows && c.gates[i].nq == 1) {
      size_t j = i;
      while (j < G && c.gates[j].nq == 1 && !marked[j]) j++;
      if (j - i >= 2) {
        // Gates on distinct wires commute: each wire's subsequence collapses
        // to one product, later gates multiplied on the left.
        uint32_t top = 0;
        for (size_t k = i; k < j; k++) top = std::max(top, c.gates[k].q[0] + 1);
        std::vector<std::array<cd, 4>> prod(top);
        std::vector<uint8_t> seen(top, 0);
        for (size_t k = i; k < j; k++) {
          const Gate &g = c.gates[k];
          cd m[4];
          gate_matrix(g.kind, g.p, m);
          cd *acc = prod[g.q[0]].data();
          if (!seen[g.q[0]]) {
            seen[g.q[0]] = 1;
            std::copy(m, m + 4, acc);
          } else {
            mul2(m, acc, acc);
          }
        }
        std::vector<uint32_t> live;
        for (uint32_t w = 0; w < top; w++)
          if (seen[w] && !near_identity(prod[w].data())) live.push_back(w);
        const size_t K = std::max<size_t>(1, o.group);
        for (size_t b = 0; b < live.size(); b += K) {
          const size_t take = std::min(K, live.size() - b);
          if (take == 1) {
            plan.ops.push_back(dense(live[b], prod[live[b]].data()));
            continue;
          }
          PlanOp op;
          op.kind = OpKind::Group;
          for (size_t k = 0; k < take; k++) {
            const cd *m = prod[live[b + k]].data();
            op.wires.push_back(live[b + k]);
            op.mats.insert(op.mats.end(), m, m + 4);
          }
          plan.ops.push_back(std::move(op));
        }
        i = j;
        continue;
      }
    }
    plan.ops.push_back(lower_one(c.gates[i]));
    i++;
  }

  // Pass 3: blocking.
  if (o.block_bits >= 2) {
    Plan out;
    out.n = plan.n;
    for (size_t k = 0; k < plan.ops.size();) {
      size_t j = k;
      while (j < plan.ops.size() && low_1q(plan.ops[j], o.block_bits)) j++;
      if (j - k >= 2) {
        PlanOp op;
        op.kind = OpKind::Blocked;
        op.block_bits = o.block_bits;
        for (size_t m = k; m < j; m++) {
          const PlanOp &src = plan.ops[m];
          if (src.kind == OpKind::Dense1q) {
            op.wires.push_back(src.q[0]);
            op.mats.insert(op.mats.end(), src.m, src.m + 4);
          } else {
            op.wires.insert(op.wires.end(), src.wires.begin(), src.wires.end());
            op.mats.insert(op.mats.end(), src.mats.begin(), src.mats.end());
          }
        }
        out.ops.push_back(std::move(op));
        k = j;
        continue;
      }
      out.ops.push_back(std::move(plan.ops[k]));
      k++;
    }
    return out;
  }
  return plan;
}

}  // namespace qmamd
