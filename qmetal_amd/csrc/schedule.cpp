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

namespace {

// Chain steps from plan op `k`: an uncontrolled Dense1q on a distinct wire
// that `accept` takes, optionally followed by a phase layer (no global
// phase, masks of weight 1 or 2 that all hold the wire). Stops after
// `max_steps` steps or at the first op that does not fit; returns the number
// of plan ops taken and appends the step wires.
template <typename Accept>
uint32_t chain_steps(const Plan &p, uint32_t k, size_t max_steps, Accept accept,
                     std::vector<uint32_t> &wires) {
  uint32_t j = k;
  while (j < p.ops.size() && wires.size() < max_steps) {
    const PlanOp &op = p.ops[j];
    if (op.kind != OpKind::Dense1q) break;
    const uint32_t w = op.q[0];
    if (!accept(w) || std::find(wires.begin(), wires.end(), w) != wires.end()) break;
    uint32_t used = 1;
    if (j + 1 < p.ops.size() && p.ops[j + 1].kind == OpKind::Diagonal &&
        !p.ops[j + 1].masks.empty()) {
      const PlanOp &d = p.ops[j + 1];
      bool ok = zero_phase(d.global);
      for (uint64_t m : d.masks) {
        const int wt = __builtin_popcountll(m);
        if (wt < 1 || wt > 2 || !((m >> w) & 1)) ok = false;
      }
      // A layer that does not fit ends the chain after this step.
      if (ok) used = 2;
      else max_steps = wires.size() + 1;
    }
    wires.push_back(w);
    j += used;
  }
  return j - k;
}

}  // namespace

uint32_t wide_chain_at(const Plan &p, uint32_t k, std::vector<uint32_t> &wires) {
  wires.clear();
  if (p.n < kWideMinQubits) return 0;
  const uint32_t used =
      chain_steps(p, k, kWideMaxSteps, [](uint32_t w) { return w >= kWideMinWire; }, wires);
  if (wires.size() < kWideMinSteps) {
    wires.clear();
    return 0;
  }
  return used;
}

uint32_t low_chain_at(const Plan &p, uint32_t k, std::vector<uint32_t> &wires) {
  wires.clear();
  if (p.n < kLowMinQubits) return 0;
  uint32_t used = chain_steps(
      p, k, kLowMaxWire, [](uint32_t w) { return w >= 1 && w <= kLowMaxWire; }, wires);
  if (wires.size() < kLowMinSteps) {
    wires.clear();
    return 0;
  }
  // A step on wire 0 may close the pass, unless the chain already stopped at
  // a phase layer it could not take.
  const uint32_t j = k + used;
  const bool stopped = j < p.ops.size() && p.ops[j].kind == OpKind::Diagonal;
  if (!stopped) {
    std::vector<uint32_t> last;
    const uint32_t more = chain_steps(p, j, 1, [](uint32_t w) { return w == 0; }, last);
    if (!last.empty()) {
      wires.push_back(0);
      used += more;
    }
  }
  return used;
}

namespace {

// Register-tile grouping of plan ops [k0, k1), greedy in plan order: an op
// joins the open pass when its wires still fit and the merged pass moves no
// more bytes than the open pass and the op would separately.
void register_passes(const Plan &p, uint32_t k0, uint32_t k1, std::vector<Pass> &out) {
  const uint32_t n = p.n;
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
  for (uint32_t k = k0; k < k1; k++) {
    const PlanOp &op = p.ops[k];
    if (op.kind == OpKind::Diagonal && op.masks.empty()) continue;  // no dispatch
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
}

double passes_bytes(const Plan &p, const std::vector<Pass> &ps) {
  double b = 0.0;
  for (const Pass &s : ps)
    b += s.merged ? pass_bytes(p.n, s.fixed) : op_alone_bytes(p.n, p.ops[s.ops[0]]);
  return b;
}

}  // namespace

uint32_t tile_pass_at(const Plan &p, uint32_t k, Pass &pass) {
  pass = Pass{};
  if (p.n < kTileMinQubits) return 0;
  std::vector<uint32_t> cur, chain;
  uint32_t j = k, len = 0;
  while (j < p.ops.size() && j - k < kTileMaxOps) {
    const PlanOp &op = p.ops[j];
    if (op.kind == OpKind::Diagonal && op.masks.empty()) break;
    if (!pass_mergeable(op)) break;
    // A chain pass starting here runs on its own, faster kernel.
    if (j > k && (wide_chain_at(p, j, chain) || low_chain_at(p, j, chain))) break;
    const std::vector<uint32_t> w = pass_wires(op);
    bool inside = true;
    for (uint32_t q : w)
      if (q > kTileMaxWire) inside = false;
    if (!inside) break;
    std::vector<uint32_t> reg = cur;
    for (uint32_t q : w)
      if (std::find(reg.begin(), reg.end(), q) == reg.end()) reg.push_back(q);
    if (pass.seg_len.empty() || reg.size() > kPassMaxReg) {
      if (pass.seg_len.size() == kTileMaxSegs) break;
      if (!pass.seg_len.empty()) {
        pass.seg_len.back() = len;
        pass.seg_reg.back() = cur;
      }
      pass.seg_len.push_back(0);
      pass.seg_reg.emplace_back();
      cur = w;
      len = 0;
    } else {
      cur = reg;
    }
    pass.ops.push_back(j);
    len++;
    j++;
  }
  if (!pass.seg_len.empty()) {
    pass.seg_len.back() = len;
    pass.seg_reg.back() = cur;
  }
  if (pass.seg_len.size() < 2) {
    pass = Pass{};
    return 0;
  }
  std::vector<Pass> regular;
  register_passes(p, k, j, regular);
  if (pass_bytes(p.n, 0) >= passes_bytes(p, regular)) {
    pass = Pass{};
    return 0;
  }
  pass.merged = pass.tile = true;
  return j - k;
}

uint32_t gwide_pass_at(const Plan &p, uint32_t k, Pass &pass) {
  pass = Pass{};
  if (p.n < kWideMinQubits) return 0;
  std::vector<uint32_t> A, fresh, keep, chain;
  auto has = [](const std::vector<uint32_t> &v, uint32_t q) {
    return std::find(v.begin(), v.end(), q) != v.end();
  };
  uint32_t j = k, na = 0, nb = 0;
  bool second = false;
  while (j < p.ops.size() && j - k < kTileMaxOps) {
    const PlanOp &op = p.ops[j];
    if (op.kind == OpKind::Diagonal && op.masks.empty()) break;
    if (!pass_mergeable(op)) break;
    if (j > k && (wide_chain_at(p, j, chain) || low_chain_at(p, j, chain))) break;
    const std::vector<uint32_t> w = pass_wires(op);
    bool high = true;
    for (uint32_t q : w)
      if (q < kWideMinWire) high = false;
    if (!high) break;
    if (!second) {
      std::vector<uint32_t> a = A;
      for (uint32_t q : w)
        if (!has(a, q)) a.push_back(q);
      if (a.size() <= 4) {
        A = a;
        na++;
        j++;
        continue;
      }
      second = true;
    }
    std::vector<uint32_t> f = fresh, kp = keep;
    for (uint32_t q : w) {
      if (has(A, q)) {
        if (!has(kp, q)) kp.push_back(q);
      } else if (!has(f, q)) {
        f.push_back(q);
      }
    }
    const size_t wb = std::max<size_t>(2, f.size());
    if (f.size() > 3 || kp.size() > 4 - wb) break;
    fresh = f;
    keep = kp;
    nb++;
    j++;
  }
  if (na == 0 || nb == 0) return 0;
  const size_t wb = std::max<size_t>(2, fresh.size());
  // Register positions: the wires that swap out first (local bits 1..wb),
  // then the ones that stay; every other wire of the tile is a spare at or
  // above kWideMinWire.
  std::vector<uint32_t> stay = keep, out;
  for (uint32_t q : A)
    if (!has(stay, q)) (stay.size() < 4 - wb ? stay : out).push_back(q);
  std::vector<uint32_t> wave = fresh;
  uint32_t spare = kWideMinWire;
  auto next_spare = [&]() {
    while (has(A, spare) || has(fresh, spare) || has(out, spare) || has(stay, spare) ||
           has(wave, spare))
      spare++;
    return spare;
  };
  while (out.size() < wb) out.push_back(next_spare());
  while (stay.size() < 4 - wb) stay.push_back(next_spare());
  while (wave.size() < wb) wave.push_back(next_spare());
  if (spare >= p.n) return 0;
  std::vector<Pass> regular;
  register_passes(p, k, j, regular);
  if (pass_bytes(p.n, 0) >= passes_bytes(p, regular)) return 0;
  pass.merged = pass.gwide = true;
  pass.reg = out;
  pass.reg.insert(pass.reg.end(), stay.begin(), stay.end());
  pass.reg.insert(pass.reg.end(), wave.begin(), wave.end());
  for (uint32_t i = k; i < j; i++) pass.ops.push_back(i);
  pass.seg_len = {na, nb};
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
  // Runs of ops that no chain or tile pass starts in go to the register-tile
  // grouping as a whole.
  uint32_t run0 = 0;
  std::vector<uint32_t> steps;
  for (uint32_t k = 0; k < p.ops.size(); k++) {
    const PlanOp &op = p.ops[k];
    if (op.kind == OpKind::Diagonal && op.masks.empty()) continue;  // no dispatch
    uint32_t used = wide_chain_at(p, k, steps);
    const bool is_wide = used != 0;
    if (!used) used = low_chain_at(p, k, steps);
    if (used) {
      register_passes(p, run0, k, out);
      Pass s;
      s.merged = true;
      s.wide = is_wide;
      s.low = !is_wide;
      for (uint32_t j = k; j < k + used; j++) s.ops.push_back(j);
      s.reg = steps;
      out.push_back(std::move(s));
      k += used - 1;
      run0 = k + 1;
      continue;
    }
    // A generic low tile or wide pass from here: the one taking more ops.
    Pass t, g;
    const uint32_t tu = tile_pass_at(p, k, t);
    const uint32_t gu = gwide_pass_at(p, k, g);
    if (tu || gu) {
      register_passes(p, run0, k, out);
      out.push_back(std::move(gu > tu ? g : t));
      k += std::max(tu, gu) - 1;
      run0 = k + 1;
    }
  }
  register_passes(p, run0, uint32_t(p.ops.size()), out);
  return out;
}

}  // namespace qmamd
