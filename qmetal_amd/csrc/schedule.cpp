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
  const uint32_t top = p.n > kLowMaxWire ? kLowMaxWire : kLowTwoWaveWire;
  uint32_t used = chain_steps(p, k, top, [top](uint32_t w) { return w >= 1 && w <= top; }, wires);
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

uint64_t bit(uint32_t q) { return uint64_t(1) << q; }

// One schedulable unit: a plan op, or one wire's 2x2 of a Dense1qGroup (the
// group's 2x2s act on distinct wires and commute, so they may run in
// different passes).
struct Atom {
  uint32_t op;
  uint64_t wires;  // group wire of this atom (0: the whole op)
  uint64_t mov;    // wires it needs in registers (bit 0 excluded)
  uint64_t sup;    // wires it reads or writes (controls, phase mask bits)
  bool diag;       // a phase layer: commutes with every other phase layer
  bool ok;         // may join a merged pass
  std::vector<uint32_t> succ;
  uint32_t ndeps = 0;
};

// Atoms of plan ops [k0, k1) and their order constraints: two atoms whose
// wires overlap keep their plan order unless both are phase layers.
std::vector<Atom> make_atoms(const Plan &p, uint32_t k0, uint32_t k1) {
  std::vector<Atom> at;
  for (uint32_t k = k0; k < k1; k++) {
    const PlanOp &op = p.ops[k];
    auto add = [&](uint64_t wires, uint64_t mov, uint64_t sup, bool diag, bool ok) {
      Atom a;
      a.op = k;
      a.wires = wires;
      a.mov = mov & ~uint64_t(1);
      a.sup = sup;
      a.diag = diag;
      a.ok = ok;
      at.push_back(std::move(a));
    };
    switch (op.kind) {
      case OpKind::Dense1q: add(0, bit(op.q[0]), bit(op.q[0]), false, true); break;
      case OpKind::Group:
        for (uint32_t w : op.wires) add(bit(w), bit(w), bit(w), false, true);
        break;
      case OpKind::Controlled: add(0, bit(op.q[1]), bit(op.q[0]) | bit(op.q[1]), false, true); break;
      case OpKind::Swap:
        add(0, bit(op.q[0]) | bit(op.q[1]), bit(op.q[0]) | bit(op.q[1]), false, true);
        break;
      case OpKind::CCX:
        add(0, bit(op.q[2]), bit(op.q[0]) | bit(op.q[1]) | bit(op.q[2]), false, true);
        break;
      case OpKind::Diagonal: {
        if (op.masks.empty()) break;  // launches nothing
        uint64_t sup = 0;
        for (uint64_t m : op.masks) sup |= m;
        add(0, 0, sup, true, pass_mergeable(op));
        break;
      }
      case OpKind::Blocked: {
        uint64_t sup = 0;
        for (uint32_t w : op.wires) sup |= bit(w);
        add(0, sup, ~uint64_t(0), false, false);
        break;
      }
    }
  }
  // Per wire: the last atom that is not a phase layer, and the phase layers
  // after it.
  std::vector<int64_t> last(64, -1);
  std::vector<std::vector<uint32_t>> diags(64);
  std::vector<uint32_t> deps;
  for (uint32_t i = 0; i < at.size(); i++) {
    deps.clear();
    for (uint32_t q = 0; q < 64; q++) {
      if (!((at[i].sup >> q) & 1)) continue;
      if (last[q] >= 0) deps.push_back(uint32_t(last[q]));
      if (!at[i].diag) deps.insert(deps.end(), diags[q].begin(), diags[q].end());
    }
    std::sort(deps.begin(), deps.end());
    deps.erase(std::unique(deps.begin(), deps.end()), deps.end());
    for (uint32_t d : deps) at[d].succ.push_back(i);
    at[i].ndeps = uint32_t(deps.size());
    for (uint32_t q = 0; q < 64; q++) {
      if (!((at[i].sup >> q) & 1)) continue;
      if (at[i].diag) {
        diags[q].push_back(i);
      } else {
        last[q] = i;
        diags[q].clear();
      }
    }
  }
  return at;
}

enum class Kind { Reg, Tile, GWide };

// A pass being built: its atoms in execution order and the register state of
// its kind.
struct Draft {
  Kind kind;
  std::vector<uint32_t> atoms;
  // Reg / Tile: register wires of the current (tile: last) segment.
  uint64_t reg = 0;
  std::vector<uint32_t> seg_len;
  std::vector<uint64_t> seg_reg;
  // GWide: wires of the first arrangement, then the new wires and the kept
  // first-arrangement wires of the second.
  bool second = false;
  uint64_t a = 0, fresh = 0, keep = 0;
  uint32_t na = 0;
};

int popc(uint64_t m) { return __builtin_popcountll(m); }

// Cost of adding an atom moving `mov` to the draft (lower is better), or -1
// when it does not fit; `apply` commits it.
int fit(const Plan &p, Draft &d, uint64_t mov, bool apply) {
  switch (d.kind) {
    case Kind::Reg: {
      const uint64_t r = d.reg | mov;
      if (popc(r) > int(kPassMaxReg) || popc(r) + 3 > int(p.n)) return -1;
      const int c = popc(r) - popc(d.reg);
      if (apply) d.reg = r;
      return c;
    }
    case Kind::Tile: {
      if (mov >> (kTileMaxWire + 1)) return -1;
      const uint64_t r = d.reg | mov;
      if (!d.seg_len.empty() && popc(r) <= int(kPassMaxReg)) {
        const int c = popc(r) - popc(d.reg);
        if (apply) {
          d.reg = r;
          d.seg_len.back()++;
          d.seg_reg.back() = r;
        }
        return c;
      }
      if (d.seg_len.size() >= kTileMaxSegs || popc(mov) > int(kPassMaxReg)) return -1;
      if (apply) {
        d.reg = mov;
        d.seg_len.push_back(1);
        d.seg_reg.push_back(mov);
      }
      return 8;
    }
    case Kind::GWide: {
      if (mov & (bit(kWideMinWire) - 1)) return -1;
      if (!d.second) {
        const uint64_t a = d.a | mov;
        if (popc(a) <= 4) {
          const int c = popc(a) - popc(d.a);
          if (apply) {
            d.a = a;
            d.na++;
          }
          return c;
        }
      }
      const uint64_t f = d.fresh | (mov & ~d.a), k = d.keep | (mov & d.a);
      const int wb = std::max(2, popc(f));
      if (popc(f) > 3 || popc(k) > 4 - wb || popc(d.a | f) > 7) return -1;
      const int c = 5 + popc(f) - popc(d.fresh);
      if (apply) {
        d.second = true;
        d.fresh = f;
        d.keep = k;
      }
      return c;
    }
  }
  return -1;
}

// Greedy draft of one kind from the ready atoms: repeatedly take the
// fitting atom with the lowest cost (then the earliest), letting atoms whose
// order constraints it releases join later.
Draft draft(const Plan &p, const std::vector<Atom> &at, const std::vector<uint32_t> &ready,
            std::vector<uint32_t> &ndeps, Kind kind) {
  Draft d;
  d.kind = kind;
  std::vector<uint32_t> cand = ready, touched;
  while (d.atoms.size() < kTileMaxOps) {
    int best = -1, best_c = 0;
    size_t best_j = 0;
    for (size_t j = 0; j < cand.size(); j++) {
      const Atom &a = at[cand[j]];
      if (!a.ok) continue;
      const int c = fit(p, d, a.mov, false);
      if (c < 0) continue;
      if (best < 0 || c < best_c || (c == best_c && cand[j] < uint32_t(best))) {
        best = int(cand[j]);
        best_c = c;
        best_j = j;
      }
    }
    if (best < 0) break;
    fit(p, d, at[best].mov, true);
    d.atoms.push_back(uint32_t(best));
    cand.erase(cand.begin() + long(best_j));
    for (uint32_t s : at[best].succ) {
      touched.push_back(s);
      if (--ndeps[s] == 0) cand.push_back(s);
    }
  }
  for (uint32_t s : touched) ndeps[s]++;  // undo: only the chosen draft commits
  return d;
}

uint64_t fixed_of(const PlanOp &op, uint64_t wires) { return wires ? 0 : pass_fixed(op); }

// Schedules plan ops [k0, k1) (no chain passes among them) into passes.
void dag_passes(const Plan &p, uint32_t k0, uint32_t k1, std::vector<Pass> &out) {
  std::vector<Atom> at = make_atoms(p, k0, k1);
  std::vector<uint32_t> ndeps(at.size()), ready;
  for (uint32_t i = 0; i < at.size(); i++) {
    ndeps[i] = at[i].ndeps;
    if (!ndeps[i]) ready.push_back(i);
  }
  size_t done = 0;
  while (done < at.size()) {
    Draft best = draft(p, at, ready, ndeps, Kind::Reg);
    if (p.n >= kTileMinQubits) {
      Draft t = draft(p, at, ready, ndeps, Kind::Tile);
      if (t.seg_len.size() >= 2 && t.atoms.size() > best.atoms.size()) best = std::move(t);
    }
    if (p.n >= kWideMinQubits) {
      Draft g = draft(p, at, ready, ndeps, Kind::GWide);
      if (g.second && g.atoms.size() > best.atoms.size()) best = std::move(g);
    }
    Pass s;
    if (best.atoms.empty()) {
      // Only ops that cannot join a merged pass are ready: the earliest
      // runs alone.
      best.atoms = {*std::min_element(ready.begin(), ready.end())};
      best.kind = Kind::Reg;
    }
    for (uint32_t i : best.atoms) {
      s.ops.push_back(at[i].op);
      s.wmask.push_back(at[i].wires);
    }
    if (best.kind == Kind::Tile) {
      s.merged = s.tile = true;
      s.seg_len = best.seg_len;
      for (uint64_t r : best.seg_reg) {
        std::vector<uint32_t> w;
        for (uint32_t q = 1; q < 64; q++)
          if ((r >> q) & 1) w.push_back(q);
        s.seg_reg.push_back(w);
      }
    } else if (best.kind == Kind::GWide) {
      s.merged = s.gwide = true;
      s.seg_len = {best.na, uint32_t(best.atoms.size()) - best.na};
      const int wb = std::max(2, popc(best.fresh));
      // Register positions: wires that swap out first (local bits 1..wb),
      // then the ones that stay, then the wave wires; spares at or above
      // kWideMinWire fill up.
      std::vector<uint32_t> out_w, stay, wave;
      for (uint32_t q = 0; q < 64; q++) {
        if ((best.keep >> q) & 1) stay.push_back(q);
        if ((best.fresh >> q) & 1) wave.push_back(q);
      }
      for (uint32_t q = 0; q < 64; q++)
        if (((best.a & ~best.keep) >> q) & 1) (int(stay.size()) < 4 - wb ? stay : out_w).push_back(q);
      uint64_t used = best.a | best.fresh;
      uint32_t spare = kWideMinWire;
      auto next = [&]() {
        while ((used >> spare) & 1) spare++;
        used |= bit(spare);
        return spare;
      };
      while (int(out_w.size()) < wb) out_w.push_back(next());
      while (int(stay.size()) < 4 - wb) stay.push_back(next());
      while (int(wave.size()) < wb) wave.push_back(next());
      if (spare >= p.n) {
        // No spare wire: run the atoms as register passes instead.
        s = Pass{};
        best.kind = Kind::Reg;
        best.atoms.resize(1);
        s.ops.push_back(at[best.atoms[0]].op);
        s.wmask.push_back(at[best.atoms[0]].wires);
      } else {
        s.reg = out_w;
        s.reg.insert(s.reg.end(), stay.begin(), stay.end());
        s.reg.insert(s.reg.end(), wave.begin(), wave.end());
      }
    }
    if (best.kind == Kind::Reg) {
      const bool whole = s.ops.size() == 1 && s.wmask[0] == 0;
      if (whole) {
        s.merged = false;
      } else {
        s.merged = true;
        uint64_t fixed = ~uint64_t(0), reg = 0;
        for (uint32_t i : best.atoms) {
          fixed &= fixed_of(p.ops[at[i].op], at[i].wires);
          reg |= at[i].mov;
        }
        s.fixed = fixed & ~reg & ~uint64_t(1);
        // Registers, fixed bits and bit 0 must leave thread bits.
        while (popc(reg) + 1 + popc(s.fixed) + 2 > int(p.n)) s.fixed &= s.fixed - 1;
        for (uint32_t q = 1; q < 64; q++)
          if ((reg >> q) & 1) s.reg.push_back(q);
      }
    }
    // Commit.
    for (uint32_t i : best.atoms) {
      ready.erase(std::find(ready.begin(), ready.end(), i));
      for (uint32_t t : at[i].succ)
        if (--ndeps[t] == 0) ready.push_back(t);
      done++;
    }
    out.push_back(std::move(s));
  }
}

}  // namespace

std::vector<Pass> schedule_passes(const Plan &p) {
  std::vector<Pass> out;
  const uint32_t n = p.n;
  if (n < kPassMinQubits) {
    for (uint32_t k = 0; k < p.ops.size(); k++) {
      if (p.ops[k].kind == OpKind::Diagonal && p.ops[k].masks.empty()) continue;
      Pass s;
      s.ops.push_back(k);
      s.wmask.push_back(0);
      out.push_back(std::move(s));
    }
    return out;
  }
  // Chain passes are cut out in plan order; the ops between them are
  // scheduled as a dependency graph.
  uint32_t run0 = 0;
  std::vector<uint32_t> steps;
  for (uint32_t k = 0; k < p.ops.size(); k++) {
    const PlanOp &op = p.ops[k];
    if (op.kind == OpKind::Diagonal && op.masks.empty()) continue;  // no dispatch
    // Where both shapes start, the one that takes more plan ops.
    std::vector<uint32_t> low_steps;
    uint32_t used = wide_chain_at(p, k, steps);
    const uint32_t low_used = low_chain_at(p, k, low_steps);
    const bool is_wide = used != 0 && used >= low_used;
    if (!is_wide) {
      used = low_used;
      steps.swap(low_steps);
    }
    if (!used) continue;
    dag_passes(p, run0, k, out);
    Pass s;
    s.merged = true;
    s.wide = is_wide;
    s.low = !is_wide;
    for (uint32_t j = k; j < k + used; j++) {
      s.ops.push_back(j);
      s.wmask.push_back(0);
    }
    s.reg = steps;
    out.push_back(std::move(s));
    k += used - 1;
    run0 = k + 1;
  }
  dag_passes(p, run0, uint32_t(p.ops.size()), out);
  return out;
}

}  // namespace qmamd
