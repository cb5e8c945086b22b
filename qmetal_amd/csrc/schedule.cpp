// schedule.cpp -- grouping of plan ops into HBM passes (see schedule.h).

#include "schedule.h"

#include <algorithm>
#include <cmath>

namespace qmamd {

namespace {

bool low_weight_diagonal(const PlanOp &op) {
  for (uint64_t m : op.masks) {
    const int w = __builtin_popcountll(m);
    if (w < 1 || w > 2) return false;
  }
  return true;
}

// Global phase in turns reduced to [-1/2, 1/2]; zero means amplitudes outside
// every mask are left alone.
bool zero_phase(double rad) {
  const double t = rad / 6.283185307179586476925286766559;
  return t - std::rint(t) == 0.0;
}

uint64_t high_bits(uint64_t m) { return m & ~((uint64_t(1) << kPassFixedMin) - 1); }

}  // namespace

bool pass_mergeable(const PlanOp &op) {
  switch (op.kind) {
    case OpKind::Dense1q:
    case OpKind::Group:
    case OpKind::Controlled:
    case OpKind::Swap:
    case OpKind::CCX:
      return true;
    case OpKind::Diagonal:
      return !op.masks.empty() && low_weight_diagonal(op);
    case OpKind::Blocked:
      return false;
  }
  return false;
}

std::vector<uint32_t> pass_wires(const PlanOp &op) {
  std::vector<uint32_t> w;
  auto add = [&](uint32_t q) {
    if (q != 0 && std::find(w.begin(), w.end(), q) == w.end()) w.push_back(q);
  };
  switch (op.kind) {
    case OpKind::Dense1q: add(op.q[0]); break;
    case OpKind::Group:
      for (uint32_t q : op.wires) add(q);
      break;
    case OpKind::Controlled: add(op.q[1]); break;
    case OpKind::Swap:
      add(op.q[0]);
      add(op.q[1]);
      break;
    case OpKind::CCX: add(op.q[2]); break;
    default: break;
  }
  return w;
}

uint64_t pass_fixed(const PlanOp &op) {
  switch (op.kind) {
    case OpKind::Controlled: return high_bits(uint64_t(1) << op.q[0]);
    case OpKind::CCX:
      return high_bits((uint64_t(1) << op.q[0]) | (uint64_t(1) << op.q[1]));
    case OpKind::Diagonal: {
      if (!zero_phase(op.global) || op.masks.empty()) return 0;
      uint64_t common = ~uint64_t(0);
      for (uint64_t m : op.masks) common &= m;
      return high_bits(common);
    }
    default: return 0;
  }
}

double pass_bytes(uint32_t n, uint64_t fixed) {
  return 16.0 * std::ldexp(1.0, int(n) - __builtin_popcountll(fixed));
}

double op_alone_bytes(uint32_t n, const PlanOp &op) {
  return pass_bytes(n, pass_fixed(op));
}

uint32_t wide_chain_at(const Plan &p, uint32_t k, std::vector<uint32_t> &wires) {
  wires.clear();
  if (p.n < kWideMinQubits) return 0;
  uint32_t j = k;
  while (j < p.ops.size() && wires.size() < kWideSteps) {
    const PlanOp &op = p.ops[j];
    if (op.kind != OpKind::Dense1q) break;
    const uint32_t w = op.q[0];
    if (w < kWideMinWire || std::find(wires.begin(), wires.end(), w) != wires.end()) break;
    wires.push_back(w);
    j++;
    if (j < p.ops.size() && p.ops[j].kind == OpKind::Diagonal && !p.ops[j].masks.empty()) {
      const PlanOp &d = p.ops[j];
      bool ok = zero_phase(d.global);
      for (uint64_t m : d.masks) {
        const int wt = __builtin_popcountll(m);
        if (wt < 1 || wt > 2 || !((m >> w) & 1)) ok = false;
      }
      if (!ok) break;
      j++;
    }
  }
  if (wires.size() < kWideSteps) {
    wires.clear();
    return 0;
  }
  return j - k;
}

std::vector<Pass> schedule_passes(const Plan &p) {
  std::vector<Pass> out;
  const uint32_t n = p.n;
  if (n < kPassMinQubits) {
    for (uint32_t k = 0; k < p.ops.size(); k++) {
      if (p.ops[k].kind == OpKind::Diagonal && p.ops[k].masks.empty()) continue;
      Pass s;
      s.ops.push_back(k);
      out.push_back(std::move(s));
    }
    return out;
  }
  Pass cur;
  auto flush = [&]() {
    if (cur.ops.empty()) return;
    // A pass of one op runs through that op's own kernel, which can skip
    // more (e.g. a diagonal's common low bits) and keeps its launch shape.
    if (cur.ops.size() == 1) {
      cur.merged = false;
      cur.reg.clear();
      cur.fixed = 0;
    }
    out.push_back(std::move(cur));
    cur = Pass{};
  };
  std::vector<uint32_t> wide;
  for (uint32_t k = 0; k < p.ops.size(); k++) {
    const PlanOp &op = p.ops[k];
    if (op.kind == OpKind::Diagonal && op.masks.empty()) continue;  // no dispatch
    if (const uint32_t used = wide_chain_at(p, k, wide)) {
      flush();
      Pass s;
      s.merged = s.wide = true;
      for (uint32_t j = k; j < k + used; j++) s.ops.push_back(j);
      s.reg = wide;
      out.push_back(std::move(s));
      k += used - 1;
      continue;
    }
    if (!pass_mergeable(op)) {
      flush();
      Pass s;
      s.ops.push_back(k);
      out.push_back(std::move(s));
      continue;
    }
    const std::vector<uint32_t> w = pass_wires(op);
    const uint64_t f = pass_fixed(op);
    bool join = !cur.ops.empty() && cur.ops.size() < kPassMaxOps;
    std::vector<uint32_t> reg = cur.reg;
    if (join) {
      for (uint32_t q : w)
        if (std::find(reg.begin(), reg.end(), q) == reg.end()) reg.push_back(q);
      const uint64_t fixed = cur.fixed & f;
      // Registers plus fixed bits plus bit 0 must leave thread bits.
      if (reg.size() > kPassMaxReg ||
          reg.size() + 1 + uint32_t(__builtin_popcountll(fixed)) + 2 > n)
        join = false;
      else if (pass_bytes(n, fixed) > pass_bytes(n, cur.fixed) + op_alone_bytes(n, op))
        join = false;
      if (join) {
        // No register wire may be a fixed bit.
        for (uint32_t q : reg)
          if ((fixed >> q) & 1) join = false;
      }
      if (join) {
        cur.reg = reg;
        cur.fixed = fixed;
        cur.ops.push_back(k);
        continue;
      }
    }
    flush();
    cur.merged = true;
    cur.ops.push_back(k);
    cur.reg = w;
    cur.fixed = f;
  }
  flush();
  return out;
}

}  // namespace qmamd
