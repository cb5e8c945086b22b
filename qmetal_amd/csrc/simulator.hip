// simulator.hip -- host executor: HBM state, stream, per-op launch geometry,
// FP64 reductions, GPU-side sampling and the adjoint sweep.
//
// Host-side counterpart of the reference executor (src/metal/simulator.mm);
// each function cites the part of it whose behaviour it keeps.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <tuple>

#include "kernels.h"
#include "schedule.h"
#include "simulator.h"

namespace qmamd {
namespace {

void check(hipError_t e, const char *what) {
  if (e != hipSuccess)
    fail(std::string(what) + ": " + hipGetErrorString(e));
}

// Controls and diagonal masks at or above this bit are handled as fixed-one
// bits (the amplitudes with the bit clear are never read): a run of 2^4
// complex64 is 128 bytes, a whole cache line.
constexpr uint32_t kFixedMin = 4;

// Phase layers skip amplitudes down to 128-byte runs (bit 4), a whole line:
// skipping 64-byte halves saves nothing, HBM moves full lines.
constexpr uint32_t kDiagFixedMin = 4;

// Thread-index bits of one workgroup (256 threads): these land on the lowest
// free positions; every other bit of the base index is uniform across the
// workgroup for one iteration of the grid-stride loop.
constexpr uint32_t kLaneBits = 8;

// Amplitudes a thread owns by default: 2^4, i.e. eight 16-byte loads in
// flight per thread.
constexpr int kLocalBits = 4;

constexpr double kTwoPi = 6.283185307179586476925286766559;

// Timing record kinds of merged passes (plan-op kinds are 0..6): the
// generic pass kernel, the chain kernel, the wide and low chain kernels, the
// generic low tile and generic wide kernels.
constexpr uint32_t kTimingPass = 7;
constexpr uint32_t kTimingChain = 8;
constexpr uint32_t kTimingWide = 9;
constexpr uint32_t kTimingLow = 10;
constexpr uint32_t kTimingTile = 11;
constexpr uint32_t kTimingGWide = 12;
// The adjoint's kernels: merged backward passes over both states, the
// co-state build (zeroing and Pauli accumulation), the energy overlap, and
// the one-launch-per-rotation step.
constexpr uint32_t kTimingAdjPass = 13;
constexpr uint32_t kTimingCoState = 14;
constexpr uint32_t kTimingOverlap = 15;
constexpr uint32_t kTimingAdjRot = 16;

// State placement. A 64 GiB state (33 qubits) streams passes over tiles whose
// rows lie 4 MB or more apart at 5.0 TB/s on some allocations and at 5.9 TB/s
// on others (rows closer together: 5.6 and 5.9 TB/s); which one an allocation
// gets was not set by its size, alignment, allocation call or virtual
// address in any trial, while successive allocations in one process
// alternate between the two (128 GiB states streamed 6.0 TB/s at every
// stride; 16 and 32 GiB states were always of the slower kind). A 33-qubit
// state is therefore probed -- the time of one value-preserving pass with
// the top seven wires as steps over one with wires 18..12 -- and, while that
// ratio is above kPlaceSlow, another allocation is made and probed, up to
// kPlaceAttempts in all; the fastest is kept.
constexpr uint32_t kPlaceMinQubits = 33;
constexpr uint32_t kPlaceMaxQubits = 33;
constexpr int kPlaceAttempts = 4;
constexpr double kPlaceSlow = 1.05;

int current_device() {
  int d = 0;
  check(hipGetDevice(&d), "hipGetDevice");
  return d;
}

void narrow(const cd *m, float *out) {
  for (int k = 0; k < 4; k++) {
    out[2 * k] = float(m[k].real());
    out[2 * k + 1] = float(m[k].imag());
  }
}

// Angle in turns, reduced to [-0.5, 0.5].
double turns(double rad) {
  const double t = rad / kTwoPi;
  return t - std::rint(t);
}

// Builds a gather layout: `reg` are the physical positions of tile-local bits
// 1..k (bit 0 is always local), padded with the highest free positions up to
// `want` local bits in total.
struct LayoutBuilder {
  uint32_t n;
  std::vector<uint32_t> reg;
  uint64_t fixed = 0;

  int local_of(uint32_t pos) const {
    if (pos == 0) return 0;
    for (size_t k = 0; k < reg.size(); k++)
      if (reg[k] == pos) return int(k) + 1;
    return -1;
  }
  int add(uint32_t pos) {
    int l = local_of(pos);
    if (l >= 0) return l;
    reg.push_back(pos);
    return int(reg.size());
  }
  Layout finish(int want, int *nl_out) {
    const uint32_t nfixed = uint32_t(__builtin_popcountll(fixed));
    int nl = int(reg.size()) + 1;
    for (uint32_t p = n; p-- > 1 && nl < want;) {
      if (local_of(p) >= 0 || ((fixed >> p) & 1)) continue;
      if (uint32_t(nl) + nfixed >= n) break;
      reg.push_back(p);
      nl++;
    }
    Layout L{};
    std::vector<uint32_t> ins = reg;
    ins.push_back(0);
    for (uint32_t p = 0; p < 64; p++)
      if ((fixed >> p) & 1) ins.push_back(p);
    std::sort(ins.begin(), ins.end());
    if (ins.size() > sizeof(L.ins) || reg.size() > sizeof(L.rpos))
      fail("internal: layout too wide");
    L.nins = uint32_t(ins.size());
    for (size_t k = 0; k < ins.size(); k++) {
      L.ins[k] = uint8_t(ins[k]);
      L.hm[k] = ~((uint64_t(1) << ins[k]) - 1);
    }
    for (size_t k = 0; k < reg.size(); k++) L.rpos[k] = uint8_t(reg[k]);
    L.fixed = fixed;
    L.nthreads = (uint64_t(1) << n) >> ins.size();
    L.streaming = 1;
    for (uint32_t p : ins)
      if (p >= 1 && p <= 3) L.streaming = 0;
    if (n < 12) L.streaming = 0;
    *nl_out = nl;
    return L;
  }
};

Slot make_slot(int lbit, uint32_t cmask, const float *m8) {
  Slot s{};
  s.lbit = uint8_t(lbit);
  s.cmask = uint8_t(cmask);
  std::memcpy(s.m, m8, sizeof(s.m));
  return s;
}

const float kX[8] = {0, 0, 1, 0, 1, 0, 0, 0};

// Section kind of a 2x2 slot (slot_asm, sel = 8 * kind + local bit): an X
// whose controls are all tile-local (or absent) swaps amplitudes (5, two
// instructions per pair instead of 16); a diagonal matrix multiplies each
// amplitude (6, eight per pair: RZ, and the phases of CZ and CP, on the
// locally controlled pairs; a control on a base-index bit turns into the
// identity matrix per lane, which is diagonal too); otherwise complex over
// the locally controlled pairs (2), real (0), real on the diagonal and
// imaginary off it (7: RX and its products with real factors, eight per
// pair), or complex (1).
uint32_t slot_kind(const float *m8, bool real, bool local_ctrl, bool lane_ctrl) {
  if (!lane_ctrl && std::memcmp(m8, kX, sizeof(kX)) == 0) return 5u;
  if (m8[2] == 0.0f && m8[3] == 0.0f && m8[4] == 0.0f && m8[5] == 0.0f) return 6u;
  if (local_ctrl) return 2u;
  if (real) return 0u;
  if (m8[1] == 0.0f && m8[2] == 0.0f && m8[4] == 0.0f && m8[7] == 0.0f) return 7u;
  return 1u;
}

// U^dagger (reference simulator.mm:786-799); U3(t,p,l)^dagger = U3(-t,-l,-p).
Gate inverse(const Gate &g) {
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

bool is_rotation(uint8_t k) { return k == K_RX || k == K_RY || k == K_RZ; }
// A gate gets a gradient entry when it carries a parameter and has a native
// generator (reference simulator.mm:822-828, :917): RX/RY/RZ added with no
// parameter report 0, and a P/CP/U3 without one is not differentiated.
bool differentiable(const Gate &g) { return g.np > 0 && is_rotation(g.kind); }
}  // namespace

struct Simulator::Impl {
  uint32_t n = 0;
  uint64_t dim = 0;
  int device = 0;
  // dispatches: one per plan op run (merged or not), reset, reduction and
  // sample; launches: kernels actually launched.
  uint64_t dispatches = 0;
  uint64_t launches = 0;
  hipStream_t stream = nullptr;
  float2 *state = nullptr;
  float2 *costate = nullptr;

  // FP64 reduction partials and result slots.
  double *partials = nullptr;
  double *results = nullptr;
  size_t result_cap = 0;
  double *host_results = nullptr;  // pinned
  size_t host_result_cap = 0;

  // Per-plan parameter block (diagonal terms, blocked gate lists), staged in
  // pinned memory and copied once per run ahead of the plan's kernels.
  uint8_t *arena = nullptr;
  size_t arena_cap = 0;
  uint8_t *staging = nullptr;
  size_t staging_cap = 0;
  hipEvent_t staging_done = nullptr;
  bool staging_pending = false;

  // Per-block FP64 partials of the merged adjoint passes.
  double *adj_partials = nullptr;
  size_t adj_partials_cap = 0;

  // Sampling scratch.
  double *cdf = nullptr;
  size_t cdf_cap = 0;
  double *chunks = nullptr;
  size_t chunks_cap = 0;
  uint64_t *shots_dev = nullptr;
  size_t shots_cap = 0;

  std::vector<float> mirror;

  // Selected-amplitude reads: indices in, amplitudes out.
  uint64_t *pick_idx = nullptr;
  size_t pick_idx_cap = 0;
  float2 *pick_out = nullptr;
  size_t pick_out_cap = 0;

  // After a device-resident call (energy_and_gradient_device): its kernels
  // ran on the caller's stream, and dev_done marks their end there.
  hipEvent_t dev_done = nullptr;
  bool dev_pending = false;

  // Every entry point starts here: the device current and, after a
  // device-resident call, this stream waiting (on the GPU) for its work on
  // the caller's stream, which shares the state and co-state.
  void use() {
    check(hipSetDevice(device), "hipSetDevice");
    if (dev_pending) {
      check(hipStreamWaitEvent(stream, dev_done, 0), "hipStreamWaitEvent");
      dev_pending = false;
    }
  }

  void launched() { check(hipGetLastError(), "kernel launch"); }

  template <typename T>
  void grow(T *&p, size_t &cap, size_t need, const char *what) {
    if (need <= cap) return;
    check(hipStreamSynchronize(stream), "hipStreamSynchronize");
    if (p) check(hipFree(p), "hipFree");
    p = nullptr;
    cap = 0;
    size_t c = std::max<size_t>(need, 64);
    check(hipMalloc(reinterpret_cast<void **>(&p), c * sizeof(T)), what);
    cap = c;
  }

  double *result_slots(size_t count) {
    grow(results, result_cap, count, "hipMalloc(results)");
    return results;
  }

  // Copies `count` result doubles to the host and waits.
  const double *fetch_results(size_t count) {
    if (count > host_result_cap) {
      if (host_results) check(hipHostFree(host_results), "hipHostFree");
      host_results = nullptr;
      host_result_cap = 0;
      check(hipHostMalloc(reinterpret_cast<void **>(&host_results),
                          std::max<size_t>(count, 64) * sizeof(double)),
            "hipHostMalloc(results)");
      host_result_cap = std::max<size_t>(count, 64);
    }
    check(hipMemcpyAsync(host_results, results, count * sizeof(double),
                         hipMemcpyDeviceToHost, stream),
          "hipMemcpyAsync(results)");
    check(hipStreamSynchronize(stream), "GPU error");
    return host_results;
  }

  // Reserves `bytes` of host staging for the next plan upload.
  uint8_t *stage(size_t bytes) {
    if (staging_pending) {
      check(hipEventSynchronize(staging_done), "hipEventSynchronize");
      staging_pending = false;
    }
    if (bytes > staging_cap) {
      if (staging) check(hipHostFree(staging), "hipHostFree");
      staging = nullptr;
      staging_cap = 0;
      size_t c = std::max<size_t>(bytes, 1 << 16);
      check(hipHostMalloc(reinterpret_cast<void **>(&staging), c), "hipHostMalloc(staging)");
      staging_cap = c;
    }
    grow(arena, arena_cap, bytes, "hipMalloc(arena)");
    return staging;
  }

  void upload(size_t bytes) {
    if (!bytes) return;
    check(hipMemcpyAsync(arena, staging, bytes, hipMemcpyHostToDevice, stream),
          "hipMemcpyAsync(params)");
    check(hipEventRecord(staging_done, stream), "hipEventRecord");
    staging_pending = true;
  }

  void gather(LayoutBuilder &lb, const std::vector<Slot> &slots, float2 *second,
              int want = kLocalBits) {
    GatherArgs g{};
    g.a = state;
    g.b = second;
    int nl = 0;
    g.lay = lb.finish(std::max<int>(want, int(lb.reg.size()) + 1), &nl);
    if (slots.size() > 4) fail("internal: too many slots");
    g.nslots = uint32_t(slots.size());
    for (size_t k = 0; k < slots.size(); k++) g.slots[k] = slots[k];
    if (nl < 1 || nl > 5) fail("internal: gather width");
    launch_gather(stream, nl, g);
    launched();
  }

  // One 2x2 on `target`, with controls; controls at or above kFixedMin become
  // fixed-one bits so only their half of the state is touched.
  void controlled(const std::vector<uint32_t> &ctrls, uint32_t target,
                  const float *m8, float2 *second) {
    LayoutBuilder lb{n, {}};
    const int lt = lb.add(target);
    uint32_t cm = 0;
    for (uint32_t c : ctrls) {
      if (c >= kFixedMin) lb.fixed |= 1ull << c;
      else cm |= 1u << lb.add(c);
    }
    gather(lb, {make_slot(lt, cm, m8)}, second);
  }

  void swap(uint32_t a, uint32_t b, float2 *second) {
    LayoutBuilder lb{n, {}};
    const int la = lb.add(a), lbb = lb.add(b);
    gather(lb,
           {make_slot(lbb, 1u << la, kX), make_slot(la, 1u << lbb, kX),
            make_slot(lbb, 1u << la, kX)},
           second);
  }

  void group(const std::vector<uint32_t> &wires, const std::vector<cd> &mats,
             float2 *second) {
    LayoutBuilder lb{n, {}};
    std::vector<Slot> slots;
    for (size_t k = 0; k < wires.size(); k++) {
      float m8[8];
      narrow(&mats[4 * k], m8);
      slots.push_back(make_slot(lb.add(wires[k]), 0, m8));
    }
    gather(lb, slots, second);
  }

  size_t diag_bytes(const PlanOp &op) const {
    return op.masks.size() * std::max(sizeof(DiagTerm), sizeof(uint64_t) + sizeof(double));
  }

  // Positions of a workgroup's thread-index bits under layout L: the lowest
  // positions the layout leaves free. Every other bit of a thread's base index
  // is the same across the workgroup for one iteration of the grid-stride loop.
  uint64_t lane_bits(const Layout &L) const {
    uint64_t used = L.fixed | 1ull, lane = 0;
    for (uint32_t k = 0; k < L.nins; k++) used |= 1ull << L.ins[k];
    uint32_t got = 0;
    for (uint32_t p = 0; p < n && got < kLaneBits; p++)
      if (!((used >> p) & 1)) {
        lane |= 1ull << p;
        got++;
      }
    return lane;
  }

  // Phase-layer terms relative to a layout: weight-1 and weight-2 masks are
  // folded into const + sum beta_k x_k + sum gamma_kj x_k x_j over the
  // tile-local bits (`local_of`), bits in `fixed` count as 1, and the terms
  // that read base-index bits are written to `dt`, workgroup-uniform ones
  // first. Returns the number of terms written; angles in turns.
  template <int D, typename LocalOf>
  size_t fold_terms(const PlanOp &op, uint64_t fixed, uint64_t lane, LocalOf local_of,
                    double &constant, double (&beta0)[D], double (&gamma)[D][D],
                    DiagTerm *dt, uint32_t &nuniform) const {
    double c = op.global;
    std::map<std::tuple<uint32_t, uint64_t, uint32_t>, double> terms;
    for (size_t t = 0; t < op.masks.size(); t++) {
      const uint64_t m = op.masks[t] & ~fixed;
      const double a = op.angles[t];
      if (m == 0) {
        c += a;
        continue;
      }
      const uint32_t p0 = uint32_t(__builtin_ctzll(m));
      const int l0 = local_of(p0);
      if (__builtin_popcountll(m) == 1) {
        if (l0 >= 0) beta0[l0] += a;
        else terms[{0u, m, 0u}] += a;
        continue;
      }
      const uint32_t p1 = 63u - uint32_t(__builtin_clzll(m));
      const int l1 = local_of(p1);
      if (l0 >= 0 && l1 >= 0) gamma[std::min(l0, l1)][std::max(l0, l1)] += a;
      else if (l0 >= 0) terms[{2u, 1ull << p1, uint32_t(l0)}] += a;
      else if (l1 >= 0) terms[{2u, 1ull << p0, uint32_t(l1)}] += a;
      else terms[{1u, m, 0u}] += a;
    }
    constant = turns(c);
    for (int k = 0; k < D; k++) {
      beta0[k] = turns(beta0[k]);
      for (int j = 0; j < D; j++) gamma[k][j] = turns(gamma[k][j]);
    }
    size_t k = 0;
    for (int pass = 0; pass < 2; pass++) {
      for (const auto &kv : terms) {
        const bool uniform = (std::get<1>(kv.first) & lane) == 0;
        if (uniform != (pass == 0)) continue;
        dt[k].kind = std::get<0>(kv.first);
        dt[k].mask = std::get<1>(kv.first);
        dt[k].k = std::get<2>(kv.first);
        dt[k].ang = turns(kv.second);
        k++;
      }
      if (pass == 0) nuniform = uint32_t(k);
    }
    return k;
  }

  // --- merged passes ---------------------------------------------------------

  struct PassLaunch {
    int nl = 0;
    PassArgs a{};
    double bytes = 0.0;
    // Chain-shaped pass (straight-line kernel) instead of the generic one.
    bool chain = false;
    int chain_k = 0;
    bool chain_zero = false;
    bool chain_real = false;
    ChainArgs c{};
    // Wide chain pass (six steps over a workgroup tile).
    bool wide = false;
    WideArgs w{};
    // Low chain pass (steps on wires 1..11 and 0 over two-wave tiles).
    bool low = false;
    int low_wb = 1;  // low pass wave bits
    LowArgs l{};
    int mode = 0;  // wide / low step mode (launch_wide)
    // Generic low tile pass (segments of merged-pass ops).
    bool tile = false;
    TileArgs t{};
    // Generic wide pass (merged-pass ops around the wide exchange).
    bool gwide = false;
    GWideArgs g{};
  };

  // Local coefficients of a phase op of a merged pass (turns).
  struct PhaseCoef {
    double beta0[kPassMaxNL] = {};
    double gamma[kPassMaxNL][kPassMaxNL] = {};
  };

  // 2x2 slots of a pass entry (`wires`: the Dense1qGroup wires it applies,
  // 0 for all).
  static size_t pass_slots(const PlanOp &op, uint64_t wires = 0) {
    switch (op.kind) {
      case OpKind::Group: return wires ? size_t(__builtin_popcountll(wires)) : op.wires.size();
      case OpKind::Swap: return 3;
      default: return 1;
    }
  }

  size_t pass_param_bytes(const Plan &p, const Pass &ps) const {
    size_t b = 0;
    for (size_t e = 0; e < ps.ops.size(); e++) {
      const PlanOp &op = p.ops[ps.ops[e]];
      b += pass_slots(op, ps.wmask[e]) * sizeof(PassOp);
      if (op.kind == OpKind::Diagonal) b += op.masks.size() * sizeof(DiagTerm);
    }
    b += ps.seg_len.size() * sizeof(TileSeg);
    return (b + 255) & ~size_t(255);
  }

  // 2x2 slots and phase layers of pass entries [e0, e1).
  static size_t entry_slots(const Plan &p, const Pass &ps, size_t e0, size_t e1) {
    size_t n = 0;
    for (size_t e = e0; e < e1; e++) n += pass_slots(p.ops[ps.ops[e]], ps.wmask[e]);
    return n;
  }

  // PassOps of plan ops `ks` (in order; `wmask`: the Dense1qGroup wires each
  // entry applies, 0 for all) against a tile frame: `local_of`
  // maps a wire to its tile-local bit (0..nl-1) or -1, `fixed` bits are 1 for
  // every amplitude the pass touches, `var` are the base-index bits that
  // differ between the threads of a workgroup (their terms are per thread,
  // the rest workgroup-uniform). Writes ops to `po` and phase terms to
  // dt[nt..], advancing nt; `coef` gets each phase op's local coefficients.
  // Returns the number of ops written.
  template <typename LocalOf>
  size_t emit_ops(const Plan &p, const uint32_t *ks, const uint64_t *wmask, size_t count,
                  LocalOf local_of, uint64_t fixed, uint64_t var, int nl, PassOp *po,
                  DiagTerm *dt, size_t &nt, PhaseCoef *coef) const {
    size_t i = 0;
    auto local = [&](uint32_t q) {
      const int l = local_of(q);
      if (l < 0) fail("internal: pass wire not local");
      return uint32_t(l);
    };
    // A control: nothing if fixed (always 1 here), a tile-local bit, or a
    // bit of the thread's base index.
    auto control = [&](uint32_t c, PassOp &o) {
      if ((fixed >> c) & 1) return;
      const int l = local_of(c);
      if (l >= 0) o.cmask |= 1u << l;
      else o.gmask |= 1ull << c;
    };
    auto slot = [&](uint32_t target, const float *m8) -> PassOp & {
      PassOp &o = po[i++];
      o = PassOp{};
      const bool real = m8[1] == 0.0f && m8[3] == 0.0f && m8[5] == 0.0f && m8[7] == 0.0f;
      o.type = real ? kPassRealSlot : kPassSlot;
      o.lbit = local(target);
      std::memcpy(o.m, m8, sizeof(o.m));
      return o;
    };
    for (size_t kk = 0; kk < count; kk++) {
      const PlanOp &op = p.ops[ks[kk]];
      float m8[8];
      switch (op.kind) {
        case OpKind::Dense1q: {
          narrow(op.m, m8);
          PassOp &o = slot(op.q[0], m8);
          if (slot_log)
            slot_log->emplace_back(size_t(reinterpret_cast<uint8_t *>(&o) - staging), ks[kk]);
          break;
        }
        case OpKind::Group:
          for (size_t w = 0; w < op.wires.size(); w++) {
            if (wmask[kk] && !((wmask[kk] >> op.wires[w]) & 1)) continue;
            narrow(&op.mats[4 * w], m8);
            slot(op.wires[w], m8);
          }
          break;
        case OpKind::Controlled: {
          narrow(op.m, m8);
          PassOp &o = slot(op.q[1], m8);
          control(op.q[0], o);
          break;
        }
        case OpKind::Swap: {
          const uint32_t a = op.q[0], b = op.q[1];
          control(a, slot(b, kX));
          control(b, slot(a, kX));
          control(a, slot(b, kX));
          break;
        }
        case OpKind::CCX: {
          PassOp &o = slot(op.q[2], kX);
          control(op.q[0], o);
          control(op.q[1], o);
          break;
        }
        case OpKind::Diagonal: {
          PhaseCoef &pc = coef[i];
          pc = PhaseCoef{};
          PassOp &o = po[i++];
          o = PassOp{};
          o.type = kPassPhase;
          // With no global phase an amplitude lacking a bit common to every
          // mask keeps phase 0: only combinations holding those bits change.
          uint64_t common = ~0ull;
          for (uint64_t m : op.masks) common &= m;
          if (turns(op.global) == 0.0) {
            for (uint32_t q = 0; q < 64; q++)
              if (((common & ~fixed) >> q) & 1) control(q, o);
          }
          o.term0 = uint32_t(nt);
          double constant = 0.0;
          double(&beta0)[kPassMaxNL] = pc.beta0;
          double(&gamma)[kPassMaxNL][kPassMaxNL] = pc.gamma;
          o.nterms = uint32_t(fold_terms(op, fixed, var, local_of, constant, beta0, gamma,
                                         dt + nt, o.nuniform));
          // Which sums the base-index terms feed; the workgroup-uniform ones
          // are reduced over the wave only where they feed something.
          for (uint32_t u = 0; u < o.nterms; u++) {
            const DiagTerm &tm = dt[nt + u];
            const uint32_t bit = tm.kind == 2 ? 2u << tm.k : 1u;
            if (u < o.nuniform) o.umask |= bit;
            if (tm.kind == 2) o.bmask |= 1u << tm.k;
            else o.has_phi = 1;
          }
          // Everything that does not depend on the thread, per local
          // combination: exp(2 pi i (constant + sum beta0 + sum gamma)).
          for (uint32_t h = 0; h < uint32_t(kPassMaxNL); h++)
            if ((o.bmask >> h) & 1) o.beta0[h] = beta0[h];
          for (uint32_t l = 0; l < (1u << nl); l++) {
            double th = constant;
            for (int h = 0; h < nl; h++) {
              if (!((l >> h) & 1)) continue;
              if (!((o.bmask >> h) & 1)) th += beta0[h];
              for (int k = 0; k < h; k++)
                if ((l >> k) & 1) th += gamma[k][h];
            }
            th -= std::rint(th);
            o.table[2 * l] = float(std::cos(kTwoPi * th));
            o.table[2 * l + 1] = float(std::sin(kTwoPi * th));
          }
          nt += o.nterms;
          break;
        }
        case OpKind::Blocked:
          fail("internal: blocked op in a merged pass");
      }
    }
    // Code section of each slot: its local bit and kind (real over all
    // pairs, complex over all pairs, complex over the pairs whose local
    // controls are set); a phase op selects no section.
    for (size_t j = 0; j < i; j++) {
      PassOp &o = po[j];
      if (o.type == kPassPhase) {
        o.sel = ~0u;
        continue;
      }
      o.sel = o.lbit + 8u * slot_kind(o.m, o.type == kPassRealSlot, o.cmask != 0, o.gmask != 0);
      o.pm = 0;
      uint32_t pidx = 0;
      for (uint32_t l = 0; l < (1u << nl); l++) {
        if ((l >> o.lbit) & 1) continue;
        if ((l & o.cmask) == o.cmask) o.pm |= 1u << pidx;
        pidx++;
      }
    }
    return i;
  }

  // Lays out a merged pass: register wires first, fixed bits, fill bits at
  // the top; then one PassOp per 2x2 slot or phase layer, in plan order.
  void prep_pass(const Plan &p, const Pass &ps, size_t off, PassLaunch &pl) {
    LayoutBuilder lb{n, {}};
    lb.fixed = ps.fixed;
    for (uint32_t q : ps.reg) lb.add(q);
    const int want = lb.reg.size() + 1 <= 4 ? 4 : 5;
    pl.a.lay = lb.finish(want, &pl.nl);
    if (pl.nl != want) fail("internal: pass width");
    const uint64_t lane = lane_bits(pl.a.lay);
    const size_t nops = entry_slots(p, ps, 0, ps.ops.size());
    PassOp *po = reinterpret_cast<PassOp *>(staging + off);
    DiagTerm *dt = reinterpret_cast<DiagTerm *>(staging + off + nops * sizeof(PassOp));
    size_t nt = 0;
    std::vector<PhaseCoef> coef(nops);
    if (emit_ops(p, ps.ops.data(), ps.wmask.data(), ps.ops.size(),
                 [&](uint32_t q) { return lb.local_of(q); }, lb.fixed, lane, pl.nl, po, dt, nt,
                 coef.data()) != nops)
      fail("internal: pass op count");
    pl.a.a = state;
    pl.a.ops = reinterpret_cast<const PassOp *>(arena + off);
    pl.a.terms = reinterpret_cast<const DiagTerm *>(arena + off + nops * sizeof(PassOp));
    pl.a.nops = uint32_t(nops);
    pl.bytes = pass_bytes(n, ps.fixed);
    to_chain(po, dt, coef, nops, nt, off, pl);
  }

  // A pass of chain shape -- uncontrolled 2x2s on local bits 1, 2, ..., K in
  // order (optionally one more on bit 0 last), each followed by at most one
  // phase layer whose every mask holds that step's bit --
  // is rewritten in place as ChainSteps for the straight-line kernel.
  void to_chain(const PassOp *po, const DiagTerm *dt, const std::vector<PhaseCoef> &coef,
                size_t nops, size_t nt, size_t off, PassLaunch &pl) {
    // Batched programs keep the generic kernel: the chain kernel's real /
    // complex choice depends on each instance's matrices.
    if (batch_compile) return;
    struct Step {
      int slot, phase;
    };
    std::vector<Step> st;
    bool real = true;
    for (size_t j = 0; j < nops; j++) {
      const PassOp &o = po[j];
      if (o.type != kPassPhase) {
        if (o.cmask || o.gmask) return;
        if (o.type != kPassRealSlot) real = false;
        st.push_back(Step{int(j), -1});
        continue;
      }
      if (st.empty() || st.back().phase >= 0) return;
      const uint32_t b = po[st.back().slot].lbit;
      // Every mask must hold the step's bit (then the phase of a combination
      // with that bit set is exactly beta + its local linear terms, and 0
      // without it); other common bits need no special handling.
      if (o.gmask || o.has_phi || !((o.cmask >> b) & 1) || (o.bmask & ~(1u << b))) return;
      st.back().phase = int(j);
    }
    const int S = int(st.size());
    int K = 0;
    while (K < S && po[st[K].slot].lbit == uint32_t(K + 1)) K++;
    const bool zero = K < S;
    if (zero && (K != S - 1 || po[st[K].slot].lbit != 0)) return;
    if (!((pl.nl == 5 && K == 4) || (pl.nl == 4 && K >= 1 && K <= 3))) return;
    // Copy what is read before the region is overwritten.
    const std::vector<PassOp> ops(po, po + nops);
    const std::vector<DiagTerm> terms(dt, dt + nt);
    ChainStep *cs = reinterpret_cast<ChainStep *>(staging + off);
    DiagTerm *uni = reinterpret_cast<DiagTerm *>(staging + off + size_t(S) * sizeof(ChainStep));
    size_t nuni = 0;
    for (int i = 0; i < S; i++) {
      const PassOp &o = ops[st[i].phase >= 0 ? st[i].phase : 0];
      if (st[i].phase >= 0)
        for (uint32_t u = 0; u < o.nuniform; u++) {
          uni[nuni] = terms[o.term0 + u];
          uni[nuni++].k = uint32_t(i);
        }
    }
    DiagTerm *lanes = uni + nuni;
    size_t nlane = 0;
    for (int i = 0; i < S; i++) {
      const PassOp &o = ops[st[i].phase >= 0 ? st[i].phase : 0];
      if (st[i].phase >= 0)
        for (uint32_t u = o.nuniform; u < o.nterms; u++) {
          lanes[nlane] = terms[o.term0 + u];
          lanes[nlane++].k = uint32_t(i);
        }
    }
    for (int i = 0; i < S; i++) {
      ChainStep &c = cs[i];
      c = ChainStep{};
      std::memcpy(c.m, ops[st[i].slot].m, sizeof(c.m));
      const int b = int(ops[st[i].slot].lbit);
      // Other local bits, ascending: the lowest two form the low group.
      int other[kPassMaxNL], no = 0;
      for (int k = 0; k < pl.nl; k++)
        if (k != b) other[no++] = k;
      double g[kPassMaxNL] = {};
      if (st[i].phase >= 0) {
        const PhaseCoef &pc = coef[size_t(st[i].phase)];
        c.beta0 = pc.beta0[b];
        for (int j = 0; j < no; j++)
          g[j] = pc.gamma[std::min(b, other[j])][std::max(b, other[j])];
      }
      for (int cc = 1; cc < 4; cc++) {
        double tl = 0.0, th = 0.0;
        for (int j = 0; j < 2; j++)
          if ((cc >> j) & 1) {
            tl += g[j];
            if (2 + j < no) th += g[2 + j];
          }
        tl -= std::rint(tl);
        th -= std::rint(th);
        c.lo[2 * cc - 2] = float(std::cos(kTwoPi * tl));
        c.lo[2 * cc - 1] = float(std::sin(kTwoPi * tl));
        c.hi[2 * cc - 2] = float(std::cos(kTwoPi * th));
        c.hi[2 * cc - 1] = float(std::sin(kTwoPi * th));
      }
    }
    pl.chain = true;
    pl.chain_k = K;
    pl.chain_zero = zero;
    pl.chain_real = real;
    pl.c.a = state;
    pl.c.lay = pl.a.lay;
    pl.c.steps = reinterpret_cast<const ChainStep *>(arena + off);
    pl.c.uni = reinterpret_cast<const DiagTerm *>(arena + off + size_t(S) * sizeof(ChainStep));
    pl.c.lane = pl.c.uni + nuni;
    pl.c.nuni = uint32_t(nuni);
    pl.c.nlane = uint32_t(nlane);
  }

  // Phase tables of a chain step on local bit `b` of a five-bit register
  // tile: the other four local bits, ascending, form the low pair (lo) and
  // the high pair (hi); entry c - 1 is the factor for pair value c = 1..3.
  static void step_tables(ChainStep &c, int b, const double (&gamma)[kPassMaxNL][kPassMaxNL]) {
    int other[4], no = 0;
    for (int q = 0; q < 5; q++)
      if (q != b) other[no++] = q;
    for (int cc = 1; cc < 4; cc++) {
      double tl = 0.0, th = 0.0;
      for (int j = 0; j < 2; j++)
        if ((cc >> j) & 1) {
          tl += gamma[std::min(b, other[j])][std::max(b, other[j])];
          th += gamma[std::min(b, other[2 + j])][std::max(b, other[2 + j])];
        }
      tl -= std::rint(tl);
      th -= std::rint(th);
      c.lo[2 * cc - 2] = float(std::cos(kTwoPi * tl));
      c.lo[2 * cc - 1] = float(std::sin(kTwoPi * tl));
      c.hi[2 * cc - 2] = float(std::cos(kTwoPi * th));
      c.hi[2 * cc - 1] = float(std::sin(kTwoPi * th));
    }
  }

  // A chain step's Dense1q: its matrix, and unit phase tables until a phase
  // layer sets them.
  static void step_matrix(ChainStep &c, const PlanOp &op) {
    c = ChainStep{};
    float m8[8];
    narrow(op.m, m8);
    std::memcpy(c.m, m8, sizeof(m8));
    for (int cc = 0; cc < 6; cc++) c.lo[cc] = c.hi[cc] = (cc % 2 == 0) ? 1.0f : 0.0f;
  }

  // True when the phase tables of a low pass's steps are triangular: step i
  // (local bit 1 + i % 4, or 0 for the wire-0 step i == S) has factors that
  // do not change with the register bits below its own (the wires of its
  // arrangement's earlier steps), nor, for the wire-0 step, with any
  // register bit. Exact float comparisons: the low kernel's triangular form
  // then skips only factors that are exactly 1.
  static bool low_tables_triangular(const ChainStep *cs, int S, int T) {
    for (int i = 0; i < T; i++) {
      const int b = i < S ? 1 + i % 4 : 0;
      uint32_t ignored = 0;  // bits of the other-bit index c
      for (int j = 0; j < 4; j++) {
        const int q = b == 0 ? j + 1 : (j < b ? j : j + 1);
        if (b == 0 || (q >= 1 && q < b)) ignored |= 1u << j;
      }
      for (int g = 0; g < 2; g++) {
        const float *t = g ? cs[i].hi : cs[i].lo;
        const uint32_t gm = (ignored >> (2 * g)) & 3u;
        for (uint32_t cc = 1; cc < 4; cc++) {
          const uint32_t r = cc & ~gm;
          const float re = r ? t[2 * r - 2] : 1.0f, im = r ? t[2 * r - 1] : 0.0f;
          if (t[2 * cc - 2] != re || t[2 * cc - 1] != im) return false;
        }
      }
    }
    return true;
  }

  static bool steps_real(const ChainStep *cs, int count) {
    for (int i = 0; i < count; i++)
      if (cs[i].m[1] != 0.0f || cs[i].m[3] != 0.0f || cs[i].m[5] != 0.0f || cs[i].m[7] != 0.0f)
        return false;
    return true;
  }

  // Step mode of a wide or low pass (launch_wide): 2 when every step is
  // s_i [[1,1],[1,-1]] -- then the last step's matrix takes the product of
  // the earlier s_i and the others run unscaled -- else 1 when all are real,
  // else 0.
  static int steps_mode(ChainStep *cs, int count) {
    if (!steps_real(cs, count)) return 0;
    double scale = 1.0;
    for (int i = 0; i < count; i++) {
      const float *m = cs[i].m;
      if (m[0] == 0.0f || m[2] != m[0] || m[4] != m[0] || m[6] != -m[0]) return 1;
      if (i + 1 < count) scale *= double(m[0]);
    }
    for (int j = 0; j < 8; j += 2) cs[count - 1].m[j] = float(double(cs[count - 1].m[j]) * scale);
    return 2;
  }

  // Wide chain pass: the six steps' constants and base-index terms, folded
  // against the tile arrangement each step runs in (WideArgs).
  void prep_wide(const Plan &p, const Pass &ps, size_t off, PassLaunch &pl) {
    const int steps = int(ps.reg.size());
    if (steps < int(kWideMinSteps) || steps > int(kWideMaxSteps) || ps.fixed)
      fail("internal: wide pass shape");
    const int wb = steps - 4;  // wave bits
    const uint32_t *w = ps.reg.data();
    WideArgs &wa = pl.w;
    wa = WideArgs{};
    uint64_t tile = 1ull;
    for (int i = 0; i < steps; i++) {
      wa.pos[i] = uint8_t(w[i]);
      tile |= 1ull << w[i];
    }
    for (uint32_t q = 1, got = 0; got < 6; q++) {
      if ((tile >> q) & 1) fail("internal: wide pass lanes");
      wa.lanepos[got++] = uint8_t(q);
      tile |= 1ull << q;
    }
    // Tiles: insert zeros at the tile positions, lowest first.
    Layout &L = wa.lay;
    L = Layout{};
    for (uint32_t q = 0; q < 64; q++)
      if ((tile >> q) & 1) {
        L.ins[L.nins] = uint8_t(q);
        L.hm[L.nins] = ~((uint64_t(1) << q) - 1);
        L.nins++;
      }
    L.nthreads = dim >> L.nins;
    L.streaming = 1;
    // Arrangements: local bit -> wire (0 = physical bit 0), and the
    // thread-varying bits (lanes and the wave bits). The exchange swaps local
    // bits 1..wb with the wave bits.
    uint64_t lanes = 0;
    for (int j = 0; j < 6; j++) lanes |= 1ull << wa.lanepos[j];
    const uint32_t loc_a[5] = {0, w[0], w[1], w[2], w[3]};
    uint32_t loc_b[5] = {0, w[0], w[1], w[2], w[3]};
    uint64_t var_a = lanes, var_b = lanes;
    for (int j = 0; j < wb; j++) {
      loc_b[1 + j] = w[4 + j];
      var_a |= 1ull << w[4 + j];
      var_b |= 1ull << w[j];
    }
    ChainStep *cs = reinterpret_cast<ChainStep *>(staging + off);
    std::vector<DiagTerm> uni, lan;
    int step = -1;
    for (uint32_t k : ps.ops) {
      const PlanOp &op = p.ops[k];
      if (op.kind == OpKind::Dense1q) {
        step_matrix(cs[++step], op);
        continue;
      }
      // The step's phase layer.
      const bool second = step >= 4;
      const uint32_t *loc = second ? loc_b : loc_a;
      const int b = second ? step - 3 : step + 1;  // the step's local bit
      auto local_of = [&](uint32_t q) {
        for (int l = 0; l < 5; l++)
          if (loc[l] == q) return l;
        return -1;
      };
      double constant = 0.0, beta0[kPassMaxNL] = {}, gamma[kPassMaxNL][kPassMaxNL] = {};
      std::vector<DiagTerm> t(op.masks.size());
      uint32_t nuniform = 0;
      const size_t nt = fold_terms(op, 0, second ? var_b : var_a, local_of, constant, beta0, gamma,
                                   t.data(), nuniform);
      if (constant != 0.0) fail("internal: wide pass phase constant");
      for (size_t u = 0; u < nt; u++) {
        if (t[u].kind != 2 || t[u].k != uint32_t(b)) fail("internal: wide pass phase term");
        t[u].k = uint32_t(step);
        (u < nuniform ? uni : lan).push_back(t[u]);
      }
      cs[step].beta0 = beta0[b];
      step_tables(cs[step], b, gamma);
    }
    if (step != steps - 1) fail("internal: wide pass steps");
    pl.mode = steps_mode(cs, steps);
    DiagTerm *tu = reinterpret_cast<DiagTerm *>(staging + off + size_t(steps) * sizeof(ChainStep));
    std::copy(uni.begin(), uni.end(), tu);
    std::copy(lan.begin(), lan.end(), tu + uni.size());
    wa.a = state;
    wa.steps = reinterpret_cast<const ChainStep *>(arena + off);
    wa.uni = reinterpret_cast<const DiagTerm *>(arena + off + size_t(steps) * sizeof(ChainStep));
    wa.lane = wa.uni + uni.size();
    wa.nuni = uint32_t(uni.size());
    wa.nlane = uint32_t(lan.size());
    pl.wide = true;
    pl.chain_k = steps;
    pl.nl = 5;
    pl.bytes = pass_bytes(n, 0);
  }

  // Low chain pass (LowArgs): WB wave bits, tile bits 0..10+WB: WB = 2 when
  // a wire reaches 12 or the state has kLowFourWaveQubits qubits or more
  // (then a tile of 2^13 amplitudes carries wire 12 even without a step on
  // it: QFT n=33's low pass over wires 11..0 took 26.7 ms in two-wave tiles),
  // else 1. Arrangement 0 is the coalesced one: lanes on bits 1..6, the
  // first four steps' wires on the registers when they are all 7 or above
  // (so that those steps run in place), else the four wires above the wave
  // bits, and the wave bits on the remaining tile wires of 7 and up.
  // Arrangement 1 + a holds steps 4a..4a+3 on register bits, filled up with
  // the lowest unused wires, and lanes then the wave bits on the remaining
  // tile wires, ascending. The wire-0 step runs in the last.
  void prep_low(const Plan &p, const Pass &ps, size_t off, PassLaunch &pl) {
    const std::vector<uint32_t> &w = ps.reg;
    const bool zero = !w.empty() && w.back() == 0;
    const int S = int(w.size()) - (zero ? 1 : 0);
    if (S < int(kLowMinSteps) || S > int(kLowMaxWire) || ps.fixed || n < kLowMinQubits)
      fail("internal: low pass shape");
    const uint32_t wmax = *std::max_element(w.begin(), w.end());
    const int WB = wmax <= kLowTwoWaveWire && n < kLowFourWaveQubits ? 1 : 2;
    const uint32_t top = uint32_t(10 + WB);  // highest tile wire
    if (wmax > kLowMaxWire || top >= n) fail("internal: low pass tile");
    const int A = (S + 3) / 4;
    LowArgs &la = pl.l;
    la = LowArgs{};
    {
      uint32_t regs = 0;
      bool first = S >= 4;
      for (int i = 0; first && i < 4; i++) first = w[i] >= 7 && w[i] <= top;
      for (int k = 0; k < 4; k++) {
        la.regpos[0][k] = uint8_t(first ? w[k] : uint32_t(7 + WB + k));
        regs |= 1u << la.regpos[0][k];
      }
      for (int j = 0; j < 6; j++) la.lanepos[0][j] = uint8_t(1 + j);
      int nw = 0;
      for (uint32_t q = 7; q <= top; q++)
        if (!((regs >> q) & 1)) la.lanepos[0][6 + nw++] = uint8_t(q);
      if (nw != WB) fail("internal: low pass wave bits");
    }
    uint64_t var[kLowMaxArr] = {};
    for (int a = 0; a < A; a++) {
      uint32_t used = 1u;
      int nr = 0, nl = 0;
      for (int i = 4 * a; i < std::min(S, 4 * a + 4); i++) {
        if (w[i] < 1 || w[i] > top || ((used >> w[i]) & 1)) fail("internal: low pass wires");
        la.regpos[1 + a][nr++] = uint8_t(w[i]);
        used |= 1u << w[i];
      }
      for (uint32_t q = 1; nr < 4; q++)
        if (!((used >> q) & 1)) {
          la.regpos[1 + a][nr++] = uint8_t(q);
          used |= 1u << q;
        }
      for (uint32_t q = 1; q <= top; q++)
        if (!((used >> q) & 1)) {
          la.lanepos[1 + a][nl++] = uint8_t(q);
          var[1 + a] |= 1ull << q;
        }
      if (nl != 6 + WB) fail("internal: low pass lanes");
    }
    // When the first four steps' wires are arrangement 0's register bits,
    // arrangement 0 takes their order and serves the first steps in place.
    if (S >= 4) {
      uint32_t m0 = 0, m1 = 0;
      for (int k = 0; k < 4; k++) {
        m0 |= 1u << la.regpos[0][k];
        m1 |= 1u << la.regpos[1][k];
      }
      bool same_lanes = true;
      for (int j = 0; j < 6 + WB; j++) same_lanes = same_lanes && la.lanepos[0][j] == la.lanepos[1][j];
      if (m0 == m1 && same_lanes) {
        for (int k = 0; k < 4; k++) la.regpos[0][k] = la.regpos[1][k];
        la.first_in_place = 1;
      }
    }
    Layout &L = la.lay;
    L = Layout{};
    for (uint32_t q = 0; q <= top; q++) {
      L.ins[L.nins] = uint8_t(q);
      L.hm[L.nins] = ~((uint64_t(1) << q) - 1);
      L.nins++;
    }
    L.nthreads = dim >> L.nins;
    L.streaming = 1;
    ChainStep *cs = reinterpret_cast<ChainStep *>(staging + off);
    std::vector<DiagTerm> uni, lan;
    int step = -1;
    for (uint32_t k : ps.ops) {
      const PlanOp &op = p.ops[k];
      if (op.kind == OpKind::Dense1q) {
        step_matrix(cs[++step], op);
        continue;
      }
      if (step < 0) fail("internal: low pass order");
      const int a = 1 + (step < S ? step / 4 : A - 1);
      const int b = step < S ? 1 + step % 4 : 0;  // the step's local bit
      auto local_of = [&](uint32_t q) {
        if (q == 0) return 0;
        for (int l = 0; l < 4; l++)
          if (la.regpos[a][l] == q) return 1 + l;
        return -1;
      };
      double constant = 0.0, beta0[kPassMaxNL] = {}, gamma[kPassMaxNL][kPassMaxNL] = {};
      std::vector<DiagTerm> t(op.masks.size());
      uint32_t nuniform = 0;
      const size_t nt =
          fold_terms(op, 0, var[a], local_of, constant, beta0, gamma, t.data(), nuniform);
      if (constant != 0.0) fail("internal: low pass phase constant");
      for (size_t u = 0; u < nt; u++) {
        if (t[u].kind != 2 || t[u].k != uint32_t(b)) fail("internal: low pass phase term");
        t[u].k = uint32_t(step);
        (u < nuniform ? uni : lan).push_back(t[u]);
      }
      cs[step].beta0 = beta0[b];
      step_tables(cs[step], b, gamma);
    }
    const int T = S + (zero ? 1 : 0);
    if (step != T - 1) fail("internal: low pass steps");
    DiagTerm *tu = reinterpret_cast<DiagTerm *>(staging + off + size_t(T) * sizeof(ChainStep));
    std::copy(uni.begin(), uni.end(), tu);
    std::copy(lan.begin(), lan.end(), tu + uni.size());
    la.a = state;
    la.steps = reinterpret_cast<const ChainStep *>(arena + off);
    la.uni = reinterpret_cast<const DiagTerm *>(arena + off + size_t(T) * sizeof(ChainStep));
    la.lane = la.uni + uni.size();
    la.nuni = uint32_t(uni.size());
    la.nlane = uint32_t(lan.size());
    pl.low = true;
    pl.low_wb = WB;
    pl.chain_k = S;
    pl.chain_zero = zero;
    pl.mode = steps_mode(cs, T);
    if (pl.mode == 2 && low_tables_triangular(cs, S, T)) pl.mode = kLowModeTri;
    pl.nl = 5;
    pl.bytes = pass_bytes(n, 0);
  }

  // Generic low tile pass (TileArgs): per segment, its register wires
  // (filled up to four with the lowest unused wires of 1..11) and the thread
  // bits on the remaining seven, ascending; the ops emitted against that
  // arrangement. Region layout: segments, then ops, then phase terms.
  void prep_tile(const Plan &p, const Pass &ps, size_t off, PassLaunch &pl) {
    const size_t nseg = ps.seg_len.size();
    if (nseg < 2 || nseg != ps.seg_reg.size() || ps.fixed || n < kTileMinQubits)
      fail("internal: tile pass shape");
    const size_t nops = entry_slots(p, ps, 0, ps.ops.size());
    TileSeg *sg = reinterpret_cast<TileSeg *>(staging + off);
    const size_t ops_off = off + nseg * sizeof(TileSeg);
    PassOp *po = reinterpret_cast<PassOp *>(staging + ops_off);
    DiagTerm *dt = reinterpret_cast<DiagTerm *>(staging + ops_off + nops * sizeof(PassOp));
    size_t nt = 0, o = 0, k0 = 0;
    std::vector<PhaseCoef> coef(nops);
    for (size_t s = 0; s < nseg; s++) {
      TileSeg &g = sg[s];
      g = TileSeg{};
      uint32_t used = 1u, rpos[4];
      int nr = 0, nth = 0;
      for (uint32_t q : ps.seg_reg[s]) {
        if (q < 1 || q > kTileMaxWire || ((used >> q) & 1) || nr == 4)
          fail("internal: tile pass wires");
        rpos[nr++] = q;
        used |= 1u << q;
      }
      for (uint32_t q = 1; nr < 4; q++)
        if (!((used >> q) & 1)) {
          rpos[nr++] = q;
          used |= 1u << q;
        }
      uint64_t var = 0;
      for (uint32_t q = 1; q <= kTileMaxWire; q++)
        if (!((used >> q) & 1)) {
          g.tpos[nth] = uint8_t(q);
          g.tslot[nth] = tile_swizzle(1u << (q - 1));
          var |= 1ull << q;
          nth++;
        }
      if (nth != 7) fail("internal: tile pass thread bits");
      for (int k = 0; k < 4; k++) g.rslot[k] = tile_swizzle(1u << (rpos[k] - 1));
      auto local_of = [&](uint32_t q) {
        if (q == 0) return 0;
        for (int k = 0; k < 4; k++)
          if (rpos[k] == q) return k + 1;
        return -1;
      };
      g.op0 = uint32_t(o);
      o += emit_ops(p, ps.ops.data() + k0, ps.wmask.data() + k0, ps.seg_len[s], local_of, 0, var,
                    5, po + o, dt, nt, coef.data() + o);
      g.op1 = uint32_t(o);
      k0 += ps.seg_len[s];
    }
    if (o != nops || k0 != ps.ops.size()) fail("internal: tile pass op count");
    TileArgs &ta = pl.t;
    ta = TileArgs{};
    Layout &L = ta.lay;
    for (uint32_t q = 0; q <= kTileMaxWire; q++) {
      L.ins[L.nins] = uint8_t(q);
      L.hm[L.nins] = ~((uint64_t(1) << q) - 1);
      L.nins++;
    }
    L.nthreads = dim >> L.nins;
    L.streaming = 1;
    ta.a = state;
    ta.segs = reinterpret_cast<const TileSeg *>(arena + off);
    ta.nsegs = uint32_t(nseg);
    ta.ops = reinterpret_cast<const PassOp *>(arena + ops_off);
    ta.terms = reinterpret_cast<const DiagTerm *>(arena + ops_off + nops * sizeof(PassOp));
    pl.tile = true;
    pl.nl = 5;
    pl.bytes = pass_bytes(n, 0);
  }

  // Generic wide pass (GWideArgs): the tile of a wide chain pass on wires
  // ps.reg (four register wires, the first wb of which swap with the wb wave
  // wires after them); ops of the first segment emitted against the first
  // arrangement, the rest against the second.
  void prep_gwide(const Plan &p, const Pass &ps, size_t off, PassLaunch &pl) {
    const int steps = int(ps.reg.size()), wb = steps - 4;
    if (wb < 2 || wb > 3 || ps.fixed || ps.seg_len.size() != 2) fail("internal: wide pass shape");
    GWideArgs &ga = pl.g;
    ga = GWideArgs{};
    const uint32_t *w = ps.reg.data();
    uint64_t tile = 1ull;
    for (int i = 0; i < steps; i++) {
      if (w[i] < kWideMinWire || w[i] >= n || ((tile >> w[i]) & 1)) fail("internal: wide pass wires");
      ga.pos[i] = uint8_t(w[i]);
      tile |= 1ull << w[i];
    }
    uint64_t lanes = 0;
    for (uint32_t q = 1, got = 0; got < 6; q++) {
      if ((tile >> q) & 1) fail("internal: wide pass lanes");
      ga.lanepos[got++] = uint8_t(q);
      lanes |= 1ull << q;
    }
    tile |= lanes;
    Layout &L = ga.lay;
    for (uint32_t q = 0; q < 64; q++)
      if ((tile >> q) & 1) {
        L.ins[L.nins] = uint8_t(q);
        L.hm[L.nins] = ~((uint64_t(1) << q) - 1);
        L.nins++;
      }
    L.nthreads = dim >> L.nins;
    L.streaming = 1;
    uint32_t loc_a[5] = {0, w[0], w[1], w[2], w[3]}, loc_b[5] = {0, w[0], w[1], w[2], w[3]};
    uint64_t var_a = lanes, var_b = lanes;
    for (int j = 0; j < wb; j++) {
      loc_b[1 + j] = w[4 + j];
      var_a |= 1ull << w[4 + j];
      var_b |= 1ull << w[j];
    }
    const size_t nops = entry_slots(p, ps, 0, ps.ops.size());
    PassOp *po = reinterpret_cast<PassOp *>(staging + off);
    DiagTerm *dt = reinterpret_cast<DiagTerm *>(staging + off + nops * sizeof(PassOp));
    size_t nt = 0;
    std::vector<PhaseCoef> coef(nops);
    auto frame = [](const uint32_t(&loc)[5]) {
      return [&loc](uint32_t q) {
        for (int l = 0; l < 5; l++)
          if (loc[l] == q) return l;
        return -1;
      };
    };
    const size_t na = emit_ops(p, ps.ops.data(), ps.wmask.data(), ps.seg_len[0], frame(loc_a), 0,
                               var_a, 5, po, dt, nt, coef.data());
    const size_t nb = emit_ops(p, ps.ops.data() + ps.seg_len[0], ps.wmask.data() + ps.seg_len[0],
                               ps.seg_len[1], frame(loc_b), 0, var_b, 5, po + na, dt, nt,
                               coef.data() + na);
    if (na + nb != nops || ps.seg_len[0] + ps.seg_len[1] != ps.ops.size())
      fail("internal: wide pass op count");
    ga.a = state;
    ga.ops = reinterpret_cast<const PassOp *>(arena + off);
    ga.terms = reinterpret_cast<const DiagTerm *>(arena + off + nops * sizeof(PassOp));
    ga.na = uint32_t(na);
    ga.nops = uint32_t(nops);
    pl.gwide = true;
    pl.chain_k = steps;
    pl.nl = 5;
    pl.bytes = pass_bytes(n, 0);
  }

  void pass_launch(const PassLaunch &pl) {
    hipEvent_t a = nullptr;
    if (timing) {
      a = take_event();
      check(hipEventRecord(a, stream), "hipEventRecord");
    }
    if (pl.wide) launch_wide(stream, pl.chain_k, pl.mode, pl.w);
    else if (pl.low) launch_low(stream, pl.chain_k, pl.chain_zero, pl.mode, pl.low_wb, pl.l);
    else if (pl.tile) launch_tile(stream, pl.t);
    else if (pl.gwide) launch_gwide(stream, pl.chain_k - 4, pl.g);
    else if (pl.chain) launch_chain(stream, pl.nl, pl.chain_k, pl.chain_zero, pl.chain_real, pl.c);
    else launch_pass(stream, pl.nl, pl.a);
    launched();
    launches++;
    if (timing) {
      hipEvent_t b = take_event();
      check(hipEventRecord(b, stream), "hipEventRecord");
      const uint32_t kind = pl.wide    ? kTimingWide
                            : pl.low   ? kTimingLow
                            : pl.tile  ? kTimingTile
                            : pl.gwide ? kTimingGWide
                            : pl.chain ? kTimingChain
                                       : kTimingPass;
      timings.push_back(OpTiming{kind, pl.bytes, a, b});
    }
  }

  bool merge = true;
  // QMETAL_EXPECT_TILES=1: expectation(Hamiltonian) takes the tile passes
  // whenever they save a pass, however small the state.
  bool expect_tiles_always = [] {
    const char *env = getenv("QMETAL_EXPECT_TILES");
    return env && env[0] == '1';
  }();
  // Placement probe of the state: allocations made, time ratio of the one kept.
  uint32_t place_attempts = 0;
  double place_ratio = 0.0;

  // Time of a value-preserving pass with the top seven wires as steps over
  // one with wires 18..12 (after one warm-up pass), on this stream.
  double placement_ratio(float2 *p) {
    hipEvent_t e[3];
    for (hipEvent_t &x : e) check(hipEventCreate(&x), "hipEventCreate");
    launch_placement_probe(stream, p, n, 12);
    check(hipEventRecord(e[0], stream), "hipEventRecord");
    launch_placement_probe(stream, p, n, n - 7);
    check(hipEventRecord(e[1], stream), "hipEventRecord");
    launch_placement_probe(stream, p, n, 12);
    check(hipEventRecord(e[2], stream), "hipEventRecord");
    launched();
    check(hipEventSynchronize(e[2]), "placement probe");
    float top = 0.0f, mid = 0.0f;
    check(hipEventElapsedTime(&top, e[0], e[1]), "hipEventElapsedTime");
    check(hipEventElapsedTime(&mid, e[1], e[2]), "hipEventElapsedTime");
    for (hipEvent_t &x : e) (void)hipEventDestroy(x);
    return mid > 0.0f ? double(top) / double(mid) : 0.0;
  }

  // A state buffer of `bytes` (nullptr if it cannot be allocated), placed as
  // the constants above describe; QMETAL_PLACEMENT_PROBE=0 turns the probe
  // off. A candidate is allocated while the best so far is held, so a
  // failed candidate costs nothing.
  float2 *alloc_state(uint64_t bytes, uint32_t *attempts, double *ratio) {
    const char *env = getenv("QMETAL_PLACEMENT_PROBE");
    const bool probe = n >= kPlaceMinQubits && n <= kPlaceMaxQubits && !(env && env[0] == '0');
    *attempts = 0;
    *ratio = 0.0;
    float2 *best = nullptr;
    if (hipMalloc(reinterpret_cast<void **>(&best), bytes) != hipSuccess) {
      (void)hipGetLastError();
      return nullptr;
    }
    if (!probe) return best;
    *attempts = 1;
    *ratio = placement_ratio(best);
    while (*ratio > kPlaceSlow && *attempts < uint32_t(kPlaceAttempts)) {
      float2 *cand = nullptr;
      if (free_memory() < bytes ||
          hipMalloc(reinterpret_cast<void **>(&cand), bytes) != hipSuccess) {
        (void)hipGetLastError();
        break;
      }
      ++*attempts;
      const double r = placement_ratio(cand);
      if (r < *ratio) {
        check(hipFree(best), "hipFree");
        best = cand;
        *ratio = r;
      } else {
        check(hipFree(cand), "hipFree");
      }
    }
    return best;
  }

  // Set while compiling a batched program (compile_batched).
  bool batch_compile = false;
  // While set, emit_ops records every Dense1q slot it writes: the slot's
  // byte offset from `staging` and its plan op (compile_device_program).
  std::vector<std::pair<size_t, uint32_t>> *slot_log = nullptr;

  // Device-resident gradient program (energy_and_gradient_device): a
  // circuit template's forward and backward tile passes and co-state tiles,
  // compiled once; each call writes the rotations' matrices into the arena
  // from the device parameters (k_patch_angles) and launches everything on
  // the caller's stream, with no host synchronisation and no allocation.
  struct DevProg {
    std::vector<Gate> shape;  // the template (angles ignored)
    Hamiltonian h;
    uint8_t *arena = nullptr;
    AnglePatch *patch = nullptr;
    uint32_t npatch = 0;
    std::vector<AdjTileArgs> fwd, bwd;
    // The forward sweep as merged passes of the template's unfused plan
    // (when empty, `fwd` holds it as single-state tile passes).
    std::vector<PassLaunch> mfwd;
    std::vector<size_t> bwd_rot0;
    std::vector<CoTileArgs> co;
    std::vector<AccumArgs> co_generic;
    // The energy alone (no gradient asked): read-only tile passes, each
    // leaving its share in eslots, summed into eslot; empty when a term fits
    // no tile.
    std::vector<CoTileArgs> eco;
    double *eslots = nullptr;
    uint32_t nrot = 0;
    double *eslot = nullptr, *rslots = nullptr;
    uint32_t *rgate = nullptr, *inst = nullptr;
    ~DevProg() {
      (void)hipFree(arena);
      (void)hipFree(patch);
      (void)hipFree(eslots);
      (void)hipFree(eslot);
      (void)hipFree(rslots);
      (void)hipFree(rgate);
      (void)hipFree(inst);
    }
  };
  std::unique_ptr<DevProg> dev_prog;
  // Launch count after the last device-resident call: other work since then
  // on the simulator's stream must finish before the next one starts.
  uint64_t dev_launches = ~0ull;

  // --- adjoint sweep building blocks (energy_and_gradient, batched programs)

  // One merged backward pass: inverse gates (indices into the circuit, in
  // execution order) whose targets fit wire 0 plus `reg` (<= kAdjMaxReg wires), with
  // `nrot` rotations, the first at sweep index rot0.
  // A unit of a backward tile pass: gate g, or (len 3) the sandwich
  // CX(a,b) RZ_b CX(a,b), undone as the diagonal exp(i t/2 Z_a Z_b) whose
  // overlap <l|Z_a Z_b|psi> is the RZ's (CX Z_b CX = Z_a Z_b).
  struct AUnit {
    uint32_t g;
    uint32_t len;
  };
  struct APass {
    std::vector<size_t> gates;
    std::vector<uint32_t> reg;
    uint32_t nrot = 0;
    size_t rot0 = 0;
    // Tile pass (k_adj_tile) when set: units in sweep order, each segment's
    // register wires (<= 3) and the end of its units, the five high wires.
    bool tile = false;
    std::vector<AUnit> units;
    std::vector<std::vector<uint32_t>> seg_reg;
    std::vector<size_t> seg_end;
    uint32_t hpos[5] = {};
  };

  // Backward sweep in merged passes: inverse gates whose target wires fit one
  // thread's tile (wire 0 plus kAdjMaxReg register wires), at most kAdjMaxRot
  // rotations each, applied to both states in one read and one write of
  // each; overlaps reduced per pass. Gates are taken from a dependency graph
  // of the reversed circuit: two gates keep their reverse order when their
  // wires overlap, unless both are diagonal. A gate moved past others
  // commutes with them, so each rotation's overlap <l|P|psi> is the one the
  // gate-by-gate sweep would see (U'P U'^dagger = P). rot_index gets each
  // rotation's sweep index.
  static std::vector<APass> adjoint_schedule(const Circuit &c, std::vector<size_t> &rot_index) {
    if (c.n >= kAdjTileMinQubits) return adjoint_tile_schedule(c, rot_index);
    const size_t G = c.gates.size();
    auto targets = [](const Gate &g) {
      std::vector<uint32_t> t;
      if (g.nq == 1) t.push_back(g.q[0]);
      else if (g.kind == K_SWAP) t = {g.q[0], g.q[1]};
      else t.push_back(g.q[g.nq - 1]);
      return t;
    };
    // Units in sweep order (gate G-1 first) and their order constraints.
    std::vector<std::vector<uint32_t>> succ(G);
    std::vector<uint32_t> ndeps(G, 0);
    {
      std::vector<int64_t> last(64, -1);
      std::vector<std::vector<uint32_t>> diags(64);
      std::vector<uint32_t> deps;
      for (size_t u = 0; u < G; u++) {
        const Gate &g = c.gates[G - 1 - u];
        const bool dg = is_diagonal(g);
        deps.clear();
        for (uint32_t j = 0; j < g.nq; j++) {
          const uint32_t q = g.q[j];
          if (last[q] >= 0) deps.push_back(uint32_t(last[q]));
          if (!dg) deps.insert(deps.end(), diags[q].begin(), diags[q].end());
        }
        std::sort(deps.begin(), deps.end());
        deps.erase(std::unique(deps.begin(), deps.end()), deps.end());
        for (uint32_t d : deps) succ[d].push_back(uint32_t(u));
        ndeps[u] = uint32_t(deps.size());
        for (uint32_t j = 0; j < g.nq; j++) {
          const uint32_t q = g.q[j];
          if (dg) {
            diags[q].push_back(uint32_t(u));
          } else {
            last[q] = int64_t(u);
            diags[q].clear();
          }
        }
      }
    }
    std::vector<uint32_t> ready;
    for (uint32_t u = 0; u < G; u++)
      if (!ndeps[u]) ready.push_back(u);
    std::vector<APass> passes;
    rot_index.assign(G, 0);
    size_t sweep_rot = 0, done = 0;
    while (done < G) {
      APass cur;
      cur.rot0 = sweep_rot;
      for (;;) {
        // The ready gate adding the fewest register wires, then the earliest.
        int best = -1;
        size_t best_j = 0, best_add = 0;
        for (size_t j = 0; j < ready.size(); j++) {
          const Gate &g = c.gates[G - 1 - ready[j]];
          size_t add = 0;
          for (uint32_t q : targets(g))
            if (q != 0 && std::find(cur.reg.begin(), cur.reg.end(), q) == cur.reg.end()) add++;
          const uint32_t r = is_rotation(g.kind) ? 1u : 0u;
          if (cur.reg.size() + add > size_t(kAdjMaxReg) || cur.nrot + r > uint32_t(kAdjMaxRot))
            continue;
          if (best < 0 || add < best_add || (add == best_add && ready[j] < uint32_t(best))) {
            best = int(ready[j]);
            best_j = j;
            best_add = add;
          }
        }
        if (best < 0) break;
        const size_t k = G - 1 - size_t(best);
        const Gate &g = c.gates[k];
        for (uint32_t q : targets(g))
          if (q != 0 && std::find(cur.reg.begin(), cur.reg.end(), q) == cur.reg.end())
            cur.reg.push_back(q);
        if (is_rotation(g.kind)) {
          rot_index[k] = sweep_rot++;
          cur.nrot++;
        }
        cur.gates.push_back(k);
        ready.erase(ready.begin() + long(best_j));
        for (uint32_t t : succ[size_t(best)])
          if (--ndeps[t] == 0) ready.push_back(t);
        done++;
      }
      if (cur.gates.empty()) fail("internal: adjoint schedule");
      passes.push_back(std::move(cur));
    }
    return passes;
  }

  // Units of the tile sweep: every gate, except that CX(a,b) RZ_b CX(a,b)
  // (three consecutive gates) is one diagonal unit.
  static std::vector<AUnit> adjoint_units(const Circuit &c) {
    std::vector<AUnit> u;
    const size_t G = c.gates.size();
    for (size_t k = 0; k < G;) {
      const Gate &g = c.gates[k];
      if (g.kind == K_CX && k + 2 < G) {
        const Gate &r = c.gates[k + 1], &h = c.gates[k + 2];
        if (r.kind == K_RZ && r.q[0] == g.q[1] && h.kind == K_CX && h.q[0] == g.q[0] &&
            h.q[1] == g.q[1]) {
          u.push_back(AUnit{uint32_t(k), 3});
          k += 3;
          continue;
        }
      }
      u.push_back(AUnit{uint32_t(k), 1});
      k++;
    }
    return u;
  }
  static bool unit_diagonal(const Circuit &c, const AUnit &u) {
    return u.len == 3 || is_diagonal(c.gates[u.g]);
  }
  // Wires that must be tile-local when the unit is undone (none for a
  // diagonal unit, whose phase and Z-type overlap read index bits anywhere).
  static std::vector<uint32_t> unit_targets(const Circuit &c, const AUnit &u) {
    if (unit_diagonal(c, u)) return {};
    const Gate &g = c.gates[u.g];
    if (g.nq == 1) return {g.q[0]};
    if (g.kind == K_SWAP) return {g.q[0], g.q[1]};
    return {g.q[g.nq - 1]};
  }
  // Gate whose overlap the unit carries, or -1.
  static int64_t unit_rotation(const Circuit &c, const AUnit &u) {
    if (u.len == 3) return int64_t(u.g) + 1;
    return is_rotation(c.gates[u.g].kind) ? int64_t(u.g) : -1;
  }

  // Backward sweep in tile passes (k_adj_tile, states of kAdjTileMinQubits
  // or more): units from the dependency graph of the reversed circuit (two
  // units keep their order when their wires overlap, unless both are
  // diagonal), each pass a tile over wires 0..6 and five more, in segments of
  // three register wires. Greedily: the ready unit adding the fewest register
  // wires to the current segment (then the earliest) whose targets fit the
  // tile (wires >= 7 join the tile while fewer than five have); a new segment
  // when none fits; the pass ends when a new segment would not help, at
  // kAdjTileMaxSeg segments or kAdjTileMaxRot rotations.
  // forward: the same passes for the circuit itself in gate order (units
  // taken first-gate-first, no rotation limit: no overlaps).
  static std::vector<APass> adjoint_tile_schedule(const Circuit &c,
                                                  std::vector<size_t> &rot_index,
                                                  bool forward = false) {
    const std::vector<AUnit> units = adjoint_units(c);
    const size_t U = units.size();
    // Unit u of the sweep order.
    auto unit_at = [&](size_t u) -> const AUnit & { return units[forward ? u : U - 1 - u]; };
    std::vector<std::vector<uint32_t>> succ(U);
    std::vector<uint32_t> ndeps(U, 0);
    {
      std::vector<int64_t> last(64, -1);
      std::vector<std::vector<uint32_t>> diags(64);
      std::vector<uint32_t> deps;
      for (size_t u = 0; u < U; u++) {
        const AUnit &un = unit_at(u);
        const Gate &g = c.gates[un.g];
        const bool dg = unit_diagonal(c, un);
        const uint32_t nw = un.len == 3 ? 2u : g.nq;
        deps.clear();
        for (uint32_t j = 0; j < nw; j++) {
          const uint32_t q = g.q[j];
          if (last[q] >= 0) deps.push_back(uint32_t(last[q]));
          if (!dg) deps.insert(deps.end(), diags[q].begin(), diags[q].end());
        }
        std::sort(deps.begin(), deps.end());
        deps.erase(std::unique(deps.begin(), deps.end()), deps.end());
        for (uint32_t d : deps) succ[d].push_back(uint32_t(u));
        ndeps[u] = uint32_t(deps.size());
        for (uint32_t j = 0; j < nw; j++) {
          const uint32_t q = g.q[j];
          if (dg) {
            diags[q].push_back(uint32_t(u));
          } else {
            last[q] = int64_t(u);
            diags[q].clear();
          }
        }
      }
    }
    std::vector<std::vector<uint32_t>> tg(U);
    std::vector<uint32_t> isrot(U);
    for (size_t u = 0; u < U; u++) {
      tg[u] = unit_targets(c, unit_at(u));
      isrot[u] = !forward && unit_rotation(c, unit_at(u)) >= 0 ? 1u : 0u;
    }
    std::vector<uint32_t> ready;
    for (uint32_t u = 0; u < U; u++)
      if (!ndeps[u]) ready.push_back(u);
    std::vector<APass> passes;
    rot_index.assign(c.gates.size(), 0);
    size_t sweep_rot = 0, done = 0;
    const uint32_t n = c.n;
    while (done < U) {
      APass cur;
      cur.tile = true;
      cur.rot0 = sweep_rot;
      std::vector<uint32_t> hw;  // high wires taken so far
      cur.seg_reg.emplace_back();
      for (;;) {
        int best = -1;
        size_t best_j = 0, best_add = 0;
        for (size_t j = 0; j < ready.size(); j++) {
          const uint32_t u = ready[j];
          if (cur.nrot + isrot[u] > uint32_t(kAdjTileMaxRot)) continue;
          size_t newh = 0, add = 0;
          const std::vector<uint32_t> &sr = cur.seg_reg.back();
          for (size_t i = 0; i < tg[u].size(); i++) {
            const uint32_t q = tg[u][i];
            if (q >= 7 && std::find(hw.begin(), hw.end(), q) == hw.end() &&
                std::find(tg[u].begin(), tg[u].begin() + long(i), q) == tg[u].begin() + long(i))
              newh++;
            if (q != 0 && std::find(sr.begin(), sr.end(), q) == sr.end()) add++;
          }
          if (hw.size() + newh > 5 || sr.size() + add > 3) continue;
          if (best < 0 || add < best_add || (add == best_add && u < uint32_t(best))) {
            best = int(u);
            best_j = j;
            best_add = add;
          }
        }
        if (best < 0) {
          if (cur.seg_reg.size() < size_t(kAdjTileMaxSeg) && !cur.seg_reg.back().empty()) {
            cur.seg_end.push_back(cur.units.size());
            cur.seg_reg.emplace_back();
            continue;
          }
          break;
        }
        const AUnit &un = unit_at(size_t(best));
        for (uint32_t q : tg[best]) {
          if (q >= 7 && std::find(hw.begin(), hw.end(), q) == hw.end()) hw.push_back(q);
          std::vector<uint32_t> &sr = cur.seg_reg.back();
          if (q != 0 && std::find(sr.begin(), sr.end(), q) == sr.end()) sr.push_back(q);
        }
        const int64_t rg = forward ? -1 : unit_rotation(c, un);
        if (rg >= 0) {
          rot_index[size_t(rg)] = sweep_rot++;
          cur.nrot++;
        }
        cur.units.push_back(un);
        ready.erase(ready.begin() + long(best_j));
        for (uint32_t t : succ[size_t(best)])
          if (--ndeps[t] == 0) ready.push_back(t);
        done++;
      }
      if (cur.units.empty()) fail("internal: adjoint tile schedule");
      // The last segment ends with the pass (an empty trailing one is dropped).
      if (!cur.seg_end.empty() && cur.seg_end.back() == cur.units.size()) cur.seg_reg.pop_back();
      else cur.seg_end.push_back(cur.units.size());
      if (cur.seg_end.size() != cur.seg_reg.size()) fail("internal: adjoint tile segments");
      // Fill the tile to five high wires with the lowest unused ones.
      for (uint32_t q = 7; hw.size() < 5 && q < n; q++)
        if (std::find(hw.begin(), hw.end(), q) == hw.end()) hw.push_back(q);
      if (hw.size() != 5) fail("internal: adjoint tile wires");
      std::sort(hw.begin(), hw.end());
      for (int k = 0; k < 5; k++) cur.hpos[k] = hw[size_t(k)];
      passes.push_back(std::move(cur));
    }
    return passes;
  }

  // Staging bytes of a backward pass (256-aligned): AdjOps, or for a tile
  // pass its segments then at most three AdjOps per unit.
  static size_t adj_pass_bytes(const Circuit &c, const APass &ps) {
    if (ps.tile) {
      const size_t segb = (ps.seg_reg.size() * sizeof(AdjSeg) + 255) & ~size_t(255);
      return segb + ((3 * ps.units.size() * sizeof(AdjOp) + 255) & ~size_t(255));
    }
    size_t nops = 0;
    for (size_t k : ps.gates) nops += c.gates[k].kind == K_SWAP ? 3 : 1;
    return (nops * sizeof(AdjOp) + 255) & ~size_t(255);
  }

  // Segments and AdjOps of tile pass `ps` into host block `dst`, whose
  // device copy is `dev`; sets every field of `a` but the states. Returns
  // the number of AdjOps.
  // forward: the gates themselves, no overlaps (a forward tile pass). With
  // `patches`, every op whose matrix depends on a rotation's angle gets a
  // record (byte offset from `dst`, the rotation's gate index as `param`,
  // the recipe) and a section fixed by the gate kind, not by the angle.
  size_t emit_adj_tile(const Circuit &c, const APass &ps, uint8_t *dst, const uint8_t *dev,
                       AdjTileArgs &a, bool forward = false,
                       std::vector<AnglePatch> *patches = nullptr) const {
    const size_t nseg = ps.seg_reg.size();
    const size_t segb = (nseg * sizeof(AdjSeg) + 255) & ~size_t(255);
    AdjSeg *sg = reinterpret_cast<AdjSeg *>(dst);
    AdjOp *ops = reinterpret_cast<AdjOp *>(dst + segb);
    a = AdjTileArgs{};
    // Tile bit b: wire b (b <= 6), hpos[b - 7] (b >= 7).
    uint32_t tw[12];
    for (uint32_t b = 0; b < 7; b++) tw[b] = b;
    for (int k = 0; k < 5; k++) {
      tw[7 + k] = ps.hpos[k];
      a.hpos[k] = ps.hpos[k];
    }
    auto tile_bit = [&](uint32_t q) {
      for (uint32_t b = 0; b < 12; b++)
        if (tw[b] == q) return int(b);
      return -1;
    };
    size_t no = 0, u0 = 0;
    uint32_t rs = 0;
    const auto bit = [](uint32_t q) { return uint64_t(1) << q; };
    for (size_t si = 0; si < nseg; si++) {
      // Register tile bits: the segment's wires, then the lowest unused.
      uint32_t rb[3];
      int nr = 0;
      uint32_t used = 1u;
      for (uint32_t q : ps.seg_reg[si]) {
        const int b = tile_bit(q);
        if (b < 1 || nr == 3) fail("internal: adjoint tile register wire");
        rb[nr++] = uint32_t(b);
        used |= 1u << b;
      }
      for (uint32_t b = 1; nr < 3; b++)
        if (!((used >> b) & 1)) {
          rb[nr++] = b;
          used |= 1u << b;
        }
      AdjSeg &g = sg[si];
      g = AdjSeg{};
      int nt = 0;
      for (uint32_t b = 1; b < 12; b++)
        if (!((used >> b) & 1)) {
          g.tslot[nt] = tile_swizzle(1u << (b - 1));
          g.tpos[nt] = uint8_t(tw[b]);
          nt++;
        }
      if (nt != 8) fail("internal: adjoint tile thread bits");
      for (int k = 0; k < 3; k++) g.rslot[k] = tile_swizzle(1u << (rb[k] - 1));
      auto local_of = [&](uint32_t q) {
        if (q == 0) return 0;
        for (int k = 0; k < 3; k++)
          if (tw[rb[k]] == q) return k + 1;
        return -1;
      };
      // Angle-dependent op: record the patch and fix the section (6 for
      // diagonals, 7 for RX, 0 for RY).
      uint32_t pcode = ~0u, pgate = 0;
      auto op = [&](int lt, const cd *m, const cd *malt, uint32_t cm, uint64_t gmask,
                    uint64_t pmask) -> AdjOp & {
        if (lt < 0) fail("internal: adjoint tile target not local");
        if (no >= 3 * ps.units.size()) fail("internal: adjoint tile op count");
        AdjOp &o = ops[no++];
        o = AdjOp{};
        o.lbit = uint32_t(lt);
        o.gmask = gmask;
        o.pmask = pmask;
        narrow(m, o.m);
        narrow(malt ? malt : m, o.malt);
        bool real = true;
        for (int i = 1; i < 8; i += 2) real = real && o.m[i] == 0.0f && o.malt[i] == 0.0f;
        uint32_t kind = slot_kind(o.m, real, cm != 0, gmask != 0 || pmask != 0);
        if (patches && pcode != ~0u) {
          const uint32_t rk = (pcode >> 4) & 15u;
          kind = (pcode & 15u) != 0 || rk == 2 ? 6u : rk == 0 ? 7u : 0u;
          if (cm != 0 && kind != 6u) fail("internal: controlled rotation patch");
          patches->push_back(AnglePatch{
              uint32_t(reinterpret_cast<uint8_t *>(&o) - dst), pgate, pcode, 0});
        }
        o.sel = o.lbit + 8u * kind;
        uint32_t pidx = 0;
        for (uint32_t l = 0; l < 16u; l++) {
          if ((l >> o.lbit) & 1) continue;
          if ((l & cm) == cm) o.pm |= 1u << pidx;
          pidx++;
        }
        return o;
      };
      // The gradient reads the overlap's real part for RY, the imaginary
      // part otherwise (Im(i^ny S)); only that part is computed.
      auto overlap = [&](AdjOp &o, uint32_t axis, uint32_t lb, uint32_t lb2) {
        o.type = 1;
        o.comp = axis == 1 ? 0u : 1u;
        o.axis = axis;
        o.lbit = lb == 99 ? o.lbit : lb;
        o.lbit2 = lb2;
        o.rslot = rs++;
      };
      g.op0 = uint32_t(no);
      for (size_t ui = u0; ui < ps.seg_end[si]; ui++) {
        const AUnit &un = ps.units[ui];
        const Gate &gt = c.gates[un.g];
        const uint32_t inv_bit = forward ? 0u : 256u;
        pcode = ~0u;
        if (un.len == 3) {
          const uint32_t qa = gt.q[0], qb = gt.q[1];
          cd r[4];
          const double t = c.gates[un.g + 1].p[0];
          const double p[3] = {forward ? t : -t, 0.0, 0.0};
          gate_matrix(K_RZ, p, r);
          const cd d0 = r[0], d1 = r[3];
          const int la = local_of(qa), lb = local_of(qb);
          pgate = un.g + 1;
          auto ov = [&](AdjOp &o, uint32_t axis, uint32_t l1, uint32_t l2) {
            if (!forward) overlap(o, axis, l1, l2);
          };
          if (la < 0 && lb < 0) {
            const cd m[4] = {d0, 0.0, 0.0, d0}, ma[4] = {d1, 0.0, 0.0, d1};
            pcode = 1u | (2u << 4) | inv_bit;
            ov(op(0, m, ma, 0, 0, bit(qa) | bit(qb)), 3, 99, 0);
          } else if (la < 0 || lb < 0) {
            const int l = la < 0 ? lb : la;
            const uint32_t other = la < 0 ? qa : qb;
            const cd m[4] = {d0, 0.0, 0.0, d1}, ma[4] = {d1, 0.0, 0.0, d0};
            pcode = 2u | (2u << 4) | inv_bit;
            ov(op(l, m, ma, 0, 0, bit(other)), 2, 99, 0);
          } else {
            const cd m[4] = {d0, 0.0, 0.0, d1};
            pcode = 0u | (2u << 4) | inv_bit;
            ov(op(lb, m, nullptr, 0, 0, 0), 4, uint32_t(std::min(la, lb)),
               uint32_t(std::max(la, lb)));
            const cd m2[4] = {d1 * std::conj(d0), 0.0, 0.0, d0 * std::conj(d1)};
            pcode = 3u | (2u << 4) | inv_bit;
            op(lb, m2, nullptr, 1u << la, 0, 0);
          }
          pcode = ~0u;
          continue;
        }
        const Gate inv = forward ? gt : inverse(gt);
        cd m[4];
        if (gt.nq == 1 && is_diagonal(gt)) {
          gate_matrix(inv.kind, inv.p, m);
          const int l = local_of(gt.q[0]);
          const bool rz = gt.kind == K_RZ;
          pgate = un.g;
          if (l >= 0) {
            if (rz) pcode = 0u | (2u << 4) | inv_bit;
            AdjOp &o = op(l, m, nullptr, 0, 0, 0);
            if (rz && !forward) overlap(o, 2, 99, 0);
          } else {
            const cd mm[4] = {m[0], 0.0, 0.0, m[0]}, ma[4] = {m[3], 0.0, 0.0, m[3]};
            if (rz) pcode = 1u | (2u << 4) | inv_bit;
            AdjOp &o = op(0, mm, ma, 0, 0, bit(gt.q[0]));
            if (rz && !forward) overlap(o, 3, 99, 0);
          }
          pcode = ~0u;
          continue;
        }
        if (gt.kind == K_CZ || gt.kind == K_CP) {
          const cd ph = inv.kind == K_CZ ? cd(-1, 0) : cd(std::cos(inv.p[0]), std::sin(inv.p[0]));
          const uint32_t qa = gt.q[0], qb = gt.q[1];
          const int la = local_of(qa), lb = local_of(qb);
          const cd mm[4] = {1.0, 0.0, 0.0, ph};
          if (lb >= 0) op(lb, mm, nullptr, la >= 0 ? 1u << la : 0u, la >= 0 ? 0 : bit(qa), 0);
          else if (la >= 0) op(la, mm, nullptr, 0, bit(qb), 0);
          else {
            const cd md[4] = {ph, 0.0, 0.0, ph};
            op(0, md, nullptr, 0, bit(qa) | bit(qb), 0);
          }
          continue;
        }
        // Other gates: targets are tile-local; controls local or base bits.
        auto ctrl = [&](std::initializer_list<uint32_t> qs, uint32_t &cm, uint64_t &gm) {
          cm = 0;
          gm = 0;
          for (uint32_t q : qs) {
            const int l = local_of(q);
            if (l >= 0) cm |= 1u << l;
            else gm |= bit(q);
          }
        };
        uint32_t cm = 0;
        uint64_t gm = 0;
        if (gt.nq == 1) {
          gate_matrix(inv.kind, inv.p, m);
          if (is_rotation(gt.kind)) {
            pgate = un.g;
            pcode = 0u | ((gt.kind == K_RX ? 0u : 1u) << 4) | inv_bit;
          }
          AdjOp &o = op(local_of(gt.q[0]), m, nullptr, 0, 0, 0);
          pcode = ~0u;
          if (is_rotation(gt.kind) && !forward) overlap(o, gt.kind == K_RX ? 0 : 1, 99, 0);
          continue;
        }
        switch (gt.kind) {
          case K_CX:
          case K_CY:
            gate_matrix(gt.kind == K_CX ? K_X : K_Y, nullptr, m);
            ctrl({gt.q[0]}, cm, gm);
            op(local_of(gt.q[1]), m, nullptr, cm, gm, 0);
            break;
          case K_SWAP: {
            gate_matrix(K_X, nullptr, m);
            ctrl({gt.q[0]}, cm, gm);
            op(local_of(gt.q[1]), m, nullptr, cm, gm, 0);
            ctrl({gt.q[1]}, cm, gm);
            op(local_of(gt.q[0]), m, nullptr, cm, gm, 0);
            ctrl({gt.q[0]}, cm, gm);
            op(local_of(gt.q[1]), m, nullptr, cm, gm, 0);
            break;
          }
          case K_CCX:
            gate_matrix(K_X, nullptr, m);
            ctrl({gt.q[0], gt.q[1]}, cm, gm);
            op(local_of(gt.q[2]), m, nullptr, cm, gm, 0);
            break;
          default:
            fail(std::string("unsupported gate: ") + kind_name(gt.kind));
        }
      }
      g.op1 = uint32_t(no);
      u0 = ps.seg_end[si];
    }
    if (no > 3 * ps.units.size() || rs != ps.nrot) fail("internal: adjoint tile op count");
    Layout &L = a.lay;
    uint32_t ins[12];
    for (uint32_t b = 0; b < 12; b++) ins[b] = tw[b];
    std::sort(ins, ins + 12);
    for (uint32_t b = 0; b < 12; b++) {
      L.ins[L.nins] = uint8_t(ins[b]);
      L.hm[L.nins] = ~((uint64_t(1) << ins[b]) - 1);
      L.nins++;
    }
    L.nthreads = dim >> 12;
    L.streaming = 1;
    a.segs = reinterpret_cast<const AdjSeg *>(dev);
    a.ops = reinterpret_cast<const AdjOp *>(dev + segb);
    a.nsegs = uint32_t(nseg);
    a.nrot = ps.nrot;
    return no;
  }

  // AdjOps of backward pass `ps` into `ops`; sets a.lay, a.nops, a.nrot.
  void emit_adj_pass(const Circuit &c, const APass &ps, AdjOp *ops, AdjPassArgs &a) const {
    LayoutBuilder lb{n, {}};
    for (uint32_t q : ps.reg) lb.add(q);
    int nl = 0;
    a.lay = lb.finish(kAdjMaxReg + 1, &nl);
    if (nl != kAdjMaxReg + 1) fail("internal: adjoint pass width");
    a.nl = uint32_t(nl);
    uint32_t no = 0, rs = 0;
    auto slot = [&](uint32_t target, const float *m8, std::vector<uint32_t> ctrls) -> AdjOp & {
      AdjOp &o = ops[no++];
      o = AdjOp{};
      const int lt = lb.local_of(target);
      if (lt < 0) fail("internal: adjoint target not local");
      o.lbit = uint32_t(lt);
      uint32_t cm = 0;
      for (uint32_t q : ctrls) {
        const int l = lb.local_of(q);
        if (l >= 0) cm |= 1u << l;
        else o.gmask |= 1ull << q;
      }
      std::memcpy(o.m, m8, sizeof(o.m));
      const bool real = m8[1] == 0.0f && m8[3] == 0.0f && m8[5] == 0.0f && m8[7] == 0.0f;
      o.sel = o.lbit + 8u * slot_kind(m8, real, cm != 0, o.gmask != 0);
      uint32_t pidx = 0;
      for (uint32_t l = 0; l < (1u << nl); l++) {
        if ((l >> o.lbit) & 1) continue;
        if ((l & cm) == cm) o.pm |= 1u << pidx;
        pidx++;
      }
      return o;
    };
    for (size_t k : ps.gates) {
      const Gate &g = c.gates[k];
      const Gate inv = inverse(g);
      float m8[8];
      if (g.nq == 1) {
        cd m[4];
        gate_matrix(inv.kind, inv.p, m);
        narrow(m, m8);
        AdjOp &o = slot(g.q[0], m8, {});
        if (is_rotation(g.kind)) {
          o.type = 1;
          o.axis = g.kind == K_RX ? 0 : g.kind == K_RY ? 1 : 2;
          o.rslot = rs++;
        }
        continue;
      }
      switch (g.kind) {
        case K_CX:
        case K_CY: {
          cd m[4];
          gate_matrix(g.kind == K_CX ? K_X : K_Y, nullptr, m);
          narrow(m, m8);
          slot(g.q[1], m8, {g.q[0]});
          break;
        }
        case K_CZ:
        case K_CP: {
          const cd ph = inv.kind == K_CZ ? cd(-1, 0) : cd(std::cos(inv.p[0]), std::sin(inv.p[0]));
          const cd m[4] = {1.0, 0.0, 0.0, ph};
          narrow(m, m8);
          slot(g.q[1], m8, {g.q[0]});
          break;
        }
        case K_SWAP:
          slot(g.q[1], kX, {g.q[0]});
          slot(g.q[0], kX, {g.q[1]});
          slot(g.q[1], kX, {g.q[0]});
          break;
        case K_CCX:
          slot(g.q[2], kX, {g.q[0], g.q[1]});
          break;
        default:
          fail(std::string("unsupported gate: ") + kind_name(g.kind));
      }
    }
    a.nops = no;
    a.nrot = ps.nrot;
  }

  // Launches that build lambda = H psi (after zeroing it): terms grouped by
  // flip mask; flips of at most one bit share register-tile passes (bit 0
  // and the no-flip terms ride with the first tile, the other flipped wires
  // three to a tile), flips of two or more bits take one pass each. State
  // pointers and sizes are left for the caller.
  struct AccumLaunch {
    bool tile = false;
    AccumTileArgs t{};
    AccumArgs g{};
  };
  // Co-state in LDS tile passes (k_costate_tile, states of kAdjTileMinQubits
  // or more): terms grouped by flip; a flip whose bits >= 7 number at most
  // five is served inside a tile over wires 0..6 and five more (its partner
  // amplitudes read from LDS), flip groups packed first-fit into as few tiles
  // as their high wires allow; the diagonal terms (flip 0) go with the first
  // tile (in chunks of kCoTerms across tiles when there are more). Flips that
  // fit no tile keep k_pauli_accum passes, after the tiles. The first tile
  // writes lambda without reading it. States are left for the caller.
  // `spread` (read-only energy passes, where every tile costs the same read
  // and its arithmetic decides how well the read is hidden): the tiles are
  // set by the groups with high wires alone, and the diagonal terms and the
  // groups on wires 0..6 go one by one to the tile with the least work.
  struct CoPlan {
    std::vector<CoTileArgs> tiles;
    std::vector<AccumLaunch> generic;
  };
  static CoPlan costate_tile_plan(const Hamiltonian &h, uint32_t n, bool spread = false) {
    CoPlan out;
    if (n < kAdjTileMinQubits || h.empty()) return out;
    std::map<uint64_t, std::vector<size_t>> by_flip;
    for (size_t k = 0; k < h.size(); k++) by_flip[h[k].p.x | h[k].p.y].push_back(k);
    auto coef = [&](const Term &t, double &re, double &im) {  // c * i^ny
      const uint32_t ny = uint32_t(__builtin_popcountll(t.p.y));
      re = t.c;
      im = 0.0;
      switch (ny & 3u) {
        case 1: re = 0.0; im = t.c; break;
        case 2: re = -t.c; break;
        case 3: re = 0.0; im = -t.c; break;
        default: break;
      }
    };
    const uint64_t low = (uint64_t(1) << 7) - 1;
    struct Group {
      uint64_t flip;
      std::vector<size_t> terms;
    };
    struct Tile {
      uint64_t hw = 0;  // high wires taken
      std::vector<Group> groups;
      std::vector<size_t> diag;
      size_t nterms = 0;
    };
    std::vector<Tile> tiles;
    std::vector<size_t> diag = by_flip.count(0) ? by_flip[0] : std::vector<size_t>{};
    std::vector<Group> groups;
    for (const auto &kv : by_flip) {
      if (kv.first == 0) continue;
      if (__builtin_popcountll(kv.first & ~low) > 5) {
        const std::vector<size_t> &ts = kv.second;
        for (size_t b = 0; b < ts.size(); b += kAccumTerms) {
          AccumLaunch L;
          AccumArgs &a = L.g;
          a.flip = kv.first;
          for (size_t j = b; j < ts.size() && j < b + kAccumTerms; j++) {
            const Term &t = h[ts[j]];
            coef(t, a.cre[a.nterms], a.cim[a.nterms]);
            a.sign[a.nterms] = t.p.y | t.p.z;
            a.nterms++;
          }
          out.generic.push_back(L);
        }
        continue;
      }
      for (size_t b = 0; b < kv.second.size(); b += kCoTerms) {
        Group g;
        g.flip = kv.first;
        g.terms.assign(kv.second.begin() + long(b),
                       kv.second.begin() + long(std::min(kv.second.size(), b + kCoTerms)));
        groups.push_back(std::move(g));
      }
    }
    // Lowest high wire first, so neighbouring wires share tiles.
    std::stable_sort(groups.begin(), groups.end(), [&](const Group &a, const Group &b) {
      const uint64_t ha = a.flip & ~low, hb = b.flip & ~low;
      const int la = ha ? __builtin_ctzll(ha) : -1, lb2 = hb ? __builtin_ctzll(hb) : -1;
      return la < lb2;
    });
    auto first_fit = [&](Group &g) {
      const uint64_t gh = g.flip & ~low;
      Tile *dst = nullptr;
      for (Tile &t : tiles)
        if (__builtin_popcountll(t.hw | gh) <= 5 && t.groups.size() < size_t(kCoGroups) &&
            t.nterms + g.terms.size() <= size_t(kCoTerms)) {
          dst = &t;
          break;
        }
      if (!dst) {
        tiles.emplace_back();
        dst = &tiles.back();
      }
      dst->hw |= gh;
      dst->nterms += g.terms.size();
      dst->groups.push_back(std::move(g));
    };
    if (!spread) {
      for (size_t b = 0; b < diag.size() || (b == 0 && tiles.empty()); b += kCoTerms) {
        Tile t;
        t.diag.assign(diag.begin() + long(std::min(b, diag.size())),
                      diag.begin() + long(std::min(diag.size(), b + kCoTerms)));
        t.nterms = t.diag.size();
        tiles.push_back(std::move(t));
        if (diag.empty()) break;
      }
      for (Group &g : groups) first_fit(g);
    } else {
      std::vector<Group> lowg;
      for (Group &g : groups) {
        if (g.flip & ~low)
          first_fit(g);
        else
          lowg.push_back(std::move(g));
      }
      // Work per thread and tile, in VALU operations: a group's products
      // and per-term sums, a diagonal term's signed coefficient.
      std::vector<size_t> work(tiles.size(), 0);
      for (size_t i = 0; i < tiles.size(); i++)
        for (const Group &g : tiles[i].groups) work[i] += 40 + 20 * g.terms.size();
      auto least = [&](size_t groups_needed, size_t terms_needed) {
        size_t best = tiles.size();
        for (size_t i = 0; i < tiles.size(); i++)
          if (tiles[i].groups.size() + groups_needed <= size_t(kCoGroups) &&
              tiles[i].nterms + terms_needed <= size_t(kCoTerms) &&
              (best == tiles.size() || work[i] < work[best]))
            best = i;
        if (best == tiles.size()) {
          tiles.emplace_back();
          work.push_back(0);
        }
        return best;
      };
      for (Group &g : lowg) {
        const size_t i = least(1, g.terms.size());
        work[i] += 40 + 20 * g.terms.size();
        tiles[i].nterms += g.terms.size();
        tiles[i].groups.push_back(std::move(g));
      }
      for (size_t k : diag) {
        const size_t i = least(0, 1);
        work[i] += 8;
        tiles[i].nterms++;
        tiles[i].diag.push_back(k);
      }
      if (tiles.empty()) tiles.emplace_back();
    }
    if (tiles.size() == 1 && tiles[0].diag.empty() && tiles[0].groups.empty()) tiles.clear();
    for (size_t ti = 0; ti < tiles.size(); ti++) {
      Tile &t = tiles[ti];
      for (uint32_t q = 7; __builtin_popcountll(t.hw) < 5 && q < n; q++) t.hw |= uint64_t(1) << q;
      if (__builtin_popcountll(t.hw) != 5) fail("internal: co-state tile wires");
      CoTileArgs a{};
      uint32_t hp[5];
      {
        int k = 0;
        for (uint32_t q = 7; q < n; q++)
          if ((t.hw >> q) & 1) hp[k++] = q;
      }
      for (int k = 0; k < 5; k++) a.hpos[k] = hp[k];
      auto lmask = [&](uint64_t sgn) {
        uint32_t m = uint32_t(sgn & 1);
        for (int k = 0; k < 3; k++) m |= uint32_t((sgn >> hp[2 + k]) & 1) << (1 + k);
        return m;
      };
      auto slot_bits = [&](uint64_t f) {
        uint32_t sl = 0;
        for (uint32_t w = 1; w < 7; w++)
          if ((f >> w) & 1) sl ^= tile_swizzle(1u << (w - 1));
        for (int k = 0; k < 5; k++)
          if ((f >> hp[k]) & 1) sl ^= tile_swizzle(1u << (6 + k));
        return sl;
      };
      uint32_t cnt = 0;
      // Diagonal terms by local mask (real coefficients: no Y without a flip).
      std::vector<std::vector<size_t>> bym(16);
      for (size_t k : t.diag) bym[lmask(h[k].p.z)].push_back(k);
      for (uint32_t m = 0; m < 16; m++) {
        a.doff[m] = cnt;
        for (size_t k : bym[m]) {
          a.sign[cnt] = h[k].p.z;
          a.lmask[cnt] = m;
          a.cre[cnt] = h[k].c;
          a.cim[cnt] = 0.0;
          cnt++;
        }
      }
      a.doff[16] = cnt;
      for (size_t gi = 0; gi < t.groups.size(); gi++) {
        const Group &g = t.groups[gi];
        a.gflip[gi] = g.flip;
        a.gslot[gi] = slot_bits(g.flip);
        a.goff[gi] = cnt;
        uint32_t part = 0;
        for (size_t k : g.terms) {
          double re, im;
          coef(h[k], re, im);
          part |= (re != 0.0 ? 1u : 0u) | (im != 0.0 ? 2u : 0u);
        }
        if (g.terms.size() == 1 && part == 1u && lmask(h[g.terms[0]].p.y | h[g.terms[0]].p.z) == 0) {
          part |= kCoGroupPlain;
          // A flip of a single local bit: bit 0 or one of the register wires.
          const uint32_t F = lmask(g.flip);
          if (__builtin_popcountll(g.flip) == 1 && F != 0)
            part |= kCoGroupLocal | (uint32_t(__builtin_ctz(F)) << 4);
        }
        // Plain terms (real coefficient, no sign over the local bits) first:
        // the co-state pass adds them with two FMAs per amplitude.
        std::vector<size_t> order(g.terms);
        auto plain = [&](size_t k) {
          double re, im;
          coef(h[k], re, im);
          return im == 0.0 && lmask(h[k].p.y | h[k].p.z) == 0;
        };
        const size_t nplain = size_t(
            std::stable_partition(order.begin(), order.end(), plain) - order.begin());
        a.gpart[gi] = part | (uint32_t(nplain) << 8);
        for (size_t k : order) {
          const Term &tm = h[k];
          const uint64_t sg = tm.p.y | tm.p.z;
          double re, im;
          coef(tm, re, im);
          if (__builtin_popcountll(g.flip & sg) & 1) {
            re = -re;
            im = -im;
          }
          a.sign[cnt] = sg;
          a.lmask[cnt] = lmask(sg);
          a.cre[cnt] = re;
          a.cim[cnt] = im;
          cnt++;
        }
      }
      a.ngroups = uint32_t(t.groups.size());
      a.goff[a.ngroups] = cnt;
      if (cnt > uint32_t(kCoTerms)) fail("internal: co-state tile terms");
      a.first = ti == 0 ? 1u : 0u;
      Layout &L = a.lay;
      uint32_t ins[12];
      for (uint32_t b = 0; b < 7; b++) ins[b] = b;
      for (int k = 0; k < 5; k++) ins[7 + k] = hp[k];
      for (uint32_t b = 0; b < 12; b++) {
        L.ins[L.nins] = uint8_t(ins[b]);
        L.hm[L.nins] = ~((uint64_t(1) << ins[b]) - 1);
        L.nins++;
      }
      L.nthreads = (uint64_t(1) << n) >> 12;
      L.streaming = 1;
      out.tiles.push_back(a);
    }
    return out;
  }

  static std::vector<AccumLaunch> costate_launches(const Hamiltonian &h) {
    std::vector<AccumLaunch> out;
    std::map<uint64_t, std::vector<size_t>> by_flip;
    for (size_t k = 0; k < h.size(); k++) by_flip[h[k].p.x | h[k].p.y].push_back(k);
    auto coef = [&](const Term &t, double &re, double &im) {  // c * i^ny
      const uint32_t ny = uint32_t(__builtin_popcountll(t.p.y));
      re = t.c;
      im = 0.0;
      switch (ny & 3u) {
        case 1: re = 0.0; im = t.c; break;
        case 2: re = -t.c; break;
        case 3: re = 0.0; im = -t.c; break;
        default: break;
      }
    };
    std::vector<uint32_t> wires1;
    for (const auto &kv : by_flip)
      if (__builtin_popcountll(kv.first) == 1 && kv.first != 1)
        wires1.push_back(uint32_t(__builtin_ctzll(kv.first)));
    const bool has_tile = by_flip.count(0) || by_flip.count(1) || !wires1.empty();
    const size_t ntiles =
        has_tile ? std::max<size_t>(1, (wires1.size() + kAccumTileWires - 1) / kAccumTileWires) : 0;
    for (size_t g = 0; g < ntiles; g++) {
      const size_t w0 = g * kAccumTileWires;
      const uint32_t nw =
          uint32_t(std::min<size_t>(kAccumTileWires, wires1.size() - std::min(w0, wires1.size())));
      // Terms of this tile by slot: 0 flip bit 0, 1 + k wire k, nw + 1 none.
      std::vector<std::vector<size_t>> slot(nw + 2);
      if (g == 0) {
        if (by_flip.count(1)) slot[0] = by_flip[1];
        if (by_flip.count(0)) slot[nw + 1] = by_flip[0];
      }
      for (uint32_t k = 0; k < nw; k++) slot[1 + k] = by_flip[1ull << wires1[w0 + k]];
      // Launches of at most kAccumTerms terms, each slot's terms contiguous.
      std::vector<std::pair<uint32_t, size_t>> order;  // (slot, term)
      for (uint32_t sl = 0; sl < nw + 2; sl++)
        for (size_t t : slot[sl]) order.push_back({sl, t});
      for (size_t b = 0; b < order.size(); b += kAccumTerms) {
        AccumLaunch L;
        L.tile = true;
        AccumTileArgs &a = L.t;
        a.nw = nw;
        for (uint32_t k = 0; k < nw; k++) a.w[k] = uint8_t(wires1[w0 + k]);
        const size_t e = std::min(order.size(), b + kAccumTerms);
        uint32_t cnt = 0;
        for (uint32_t sl = 0; sl < nw + 2; sl++) {
          a.off[sl] = cnt;
          for (size_t j = b; j < e; j++)
            if (order[j].first == sl) {
              const Term &t = h[order[j].second];
              coef(t, a.cre[cnt], a.cim[cnt]);
              a.sign[cnt] = t.p.y | t.p.z;
              cnt++;
            }
        }
        a.off[nw + 2] = cnt;
        out.push_back(L);
      }
    }
    for (const auto &kv : by_flip) {
      if (__builtin_popcountll(kv.first) <= 1) continue;  // done on a tile above
      const std::vector<size_t> &ts = kv.second;
      for (size_t b = 0; b < ts.size(); b += kAccumTerms) {
        AccumLaunch L;
        AccumArgs &a = L.g;
        a.flip = kv.first;
        for (size_t j = b; j < ts.size() && j < b + kAccumTerms; j++) {
          const Term &t = h[ts[j]];
          coef(t, a.cre[a.nterms], a.cim[a.nterms]);
          a.sign[a.nterms] = t.p.y | t.p.z;
          a.nterms++;
        }
        out.push_back(L);
      }
    }
    return out;
  }

  // Prepared launch of a parameterised op whose parameters sit in the arena.
  struct Pending {
    int type = 0;  // 1 diagonal, 2 generic diagonal, 3 blocked
    int nl = 0;
    DiagArgs d{};
    DiagGenericArgs g{};
    BlockedArgs b{};
  };

  // Phase layer (reference kernels.metal:203-218 semantics, FP64 phase).
  // Terms are folded on the host into const + sum beta_k x_k +
  // sum gamma_kj x_k x_j over the tile-local bits; terms on other bits go to
  // the arena for the per-thread fold.
  void prep_diagonal(const PlanOp &op, size_t off, Pending &pl) {
    const size_t T = op.masks.size();
    bool low_weight = true;
    uint64_t common = ~0ull;
    for (uint64_t m : op.masks) {
      const int w = __builtin_popcountll(m);
      if (w < 1 || w > 2) low_weight = false;
      common &= m;
    }
    if (!low_weight) {
      uint64_t *mk = reinterpret_cast<uint64_t *>(staging + off);
      double *an = reinterpret_cast<double *>(staging + off + T * sizeof(uint64_t));
      for (size_t t = 0; t < T; t++) {
        mk[t] = op.masks[t];
        an[t] = turns(op.angles[t]);
      }
      pl.type = 2;
      pl.g.a = state;
      pl.g.dim = dim;
      pl.g.masks = reinterpret_cast<const uint64_t *>(arena + off);
      pl.g.angles = reinterpret_cast<const double *>(arena + off + T * sizeof(uint64_t));
      pl.g.nterms = uint32_t(T);
      pl.g.global = turns(op.global);
      return;
    }
    LayoutBuilder lb{n, {}};
    // With no global phase, amplitudes missing any common mask bit keep
    // phase 0 and are never touched.
    if (turns(op.global) == 0.0) lb.fixed = common & ~((1ull << kDiagFixedMin) - 1ull);
    DiagArgs &d = pl.d;
    d = DiagArgs{};
    d.a = state;
    d.lay = lb.finish(kLocalBits, &pl.nl);
    if (pl.nl > 4) fail("internal: diagonal width");
    d.nterms = uint32_t(fold_terms(
        op, lb.fixed, lane_bits(d.lay), [&](uint32_t q) { return lb.local_of(q); },
        d.constant, d.beta0, d.gamma, reinterpret_cast<DiagTerm *>(staging + off),
        d.nuniform));
    d.terms = reinterpret_cast<const DiagTerm *>(arena + off);
    pl.type = 1;
  }

  // Reference simulator.mm:464-499 (block_bits 2..12, wires below it).
  void prep_blocked(const PlanOp &op, size_t off, Pending &pl) {
    const uint32_t b = op.block_bits;
    if (b < 2 || b > 12) fail("block_bits must be 2..12");
    if (b > n) fail("block larger than the state");
    Slot *sl = reinterpret_cast<Slot *>(staging + off);
    for (size_t k = 0; k < op.wires.size(); k++) {
      if (op.wires[k] >= b) fail("blocked gate outside the block");
      float m8[8];
      narrow(&op.mats[4 * k], m8);
      sl[k] = make_slot(int(op.wires[k]), 0, m8);
    }
    pl.type = 3;
    pl.b.a = state;
    pl.b.nblocks = dim >> b;
    pl.b.bits = b;
    pl.b.count = uint32_t(op.wires.size());
    pl.b.gates = reinterpret_cast<const Slot *>(arena + off);
  }

  void launch_pending(const Pending &pl) {
    if (pl.type == 1) launch_diag(stream, pl.nl, pl.d);
    else if (pl.type == 2) launch_diag_generic(stream, pl.g);
    else if (pl.type == 3) launch_blocked(stream, pl.b);
    launched();
  }

  size_t param_bytes(const PlanOp &op) const {
    size_t b = 0;
    if (op.kind == OpKind::Diagonal) b = diag_bytes(op);
    if (op.kind == OpKind::Blocked) b = op.wires.size() * sizeof(Slot);
    return (b + 255) & ~size_t(255);
  }

  // Algorithmic bytes of a plan op: one read and one write of every
  // amplitude it can change (8 bytes each).
  double op_bytes(const PlanOp &p) const {
    const double A = double(dim);
    switch (p.kind) {
      case OpKind::Dense1q:
      case OpKind::Group:
      case OpKind::Blocked:
        return 16.0 * A;
      case OpKind::Controlled:
      case OpKind::Swap:
        return 8.0 * A;
      case OpKind::CCX:
        return 4.0 * A;
      case OpKind::Diagonal: {
        if (turns(p.global) != 0.0) return 16.0 * A;
        uint64_t common = ~0ull;
        for (uint64_t m : p.masks) common &= m;
        return 16.0 * A / double(uint64_t(1) << __builtin_popcountll(common));
      }
    }
    return 0.0;
  }

  // Optional per-op device timing: events recorded on this stream around
  // each plan-op launch.
  struct OpTiming {
    uint32_t kind;
    double bytes;
    hipEvent_t a, b;
  };
  bool timing = false;
  std::vector<OpTiming> timings;
  std::vector<hipEvent_t> event_pool;

  hipEvent_t take_event() {
    if (!event_pool.empty()) {
      hipEvent_t e = event_pool.back();
      event_pool.pop_back();
      return e;
    }
    hipEvent_t e;
    check(hipEventCreate(&e), "hipEventCreate");
    return e;
  }

  // Runs `launch` between two timing events when timing is on; `bytes` is
  // one read and one write of every amplitude the kernel changes plus one
  // read of every amplitude it only reads.
  template <typename F>
  void timed(uint32_t kind, double bytes, F &&launch) {
    hipEvent_t a = nullptr;
    if (timing) {
      a = take_event();
      check(hipEventRecord(a, stream), "hipEventRecord");
    }
    launch();
    if (timing) {
      hipEvent_t b = take_event();
      check(hipEventRecord(b, stream), "hipEventRecord");
      timings.push_back(OpTiming{kind, bytes, a, b});
    }
  }

  void op(const PlanOp &p, const Pending *pl) {
    if (p.kind == OpKind::Diagonal && p.masks.empty()) return;  // no dispatch
    hipEvent_t a = nullptr;
    if (timing) {
      a = take_event();
      check(hipEventRecord(a, stream), "hipEventRecord");
    }
    op_launch(p, pl);
    if (timing) {
      hipEvent_t b = take_event();
      check(hipEventRecord(b, stream), "hipEventRecord");
      timings.push_back(OpTiming{uint32_t(p.kind), op_bytes(p), a, b});
    }
  }

  void op_launch(const PlanOp &p, const Pending *pl) {
    switch (p.kind) {
      case OpKind::Dense1q: {
        float m8[8];
        narrow(p.m, m8);
        controlled({}, p.q[0], m8, nullptr);
        break;
      }
      case OpKind::Group:
        if (p.wires.size() < 2 || p.wires.size() > 4)
          fail("group size must be 2..4");
        group(p.wires, p.mats, nullptr);
        break;
      case OpKind::Blocked:
        launch_pending(*pl);
        break;
      case OpKind::Controlled: {
        float m8[8];
        narrow(p.m, m8);
        controlled({p.q[0]}, p.q[1], m8, nullptr);
        break;
      }
      case OpKind::Diagonal:
        launch_pending(*pl);
        break;
      case OpKind::Swap:
        swap(p.q[0], p.q[1], nullptr);
        break;
      case OpKind::CCX:
        controlled({p.q[0], p.q[1]}, p.q[2], kX, nullptr);
        break;
    }
    launches++;
  }

  // Gate path: one gate on the state (and, for the adjoint, the same gate on
  // the co-state in the same pass). Reference simulator.mm:348-434.
  void gate(const Gate &g, float2 *second) {
    for (int k = 0; k < g.nq; k++)
      if (g.q[k] >= n) fail("qubit index out of range");
    if (g.nq == 1) {
      cd m[4];
      gate_matrix(g.kind, g.p, m);
      float m8[8];
      narrow(m, m8);
      controlled({}, g.q[0], m8, second);
      return;
    }
    if (g.nq == 2) {
      if (g.q[0] == g.q[1]) fail("2-qubit gate on identical wires");
      switch (g.kind) {
        case K_CX:
        case K_CY: {
          cd m[4];
          gate_matrix(g.kind == K_CX ? K_X : K_Y, nullptr, m);
          float m8[8];
          narrow(m, m8);
          controlled({g.q[0]}, g.q[1], m8, second);
          return;
        }
        case K_CZ:
        case K_CP: {
          const cd ph = g.kind == K_CZ ? cd(-1, 0) : cd(std::cos(g.p[0]), std::sin(g.p[0]));
          const cd m[4] = {1.0, 0.0, 0.0, ph};
          float m8[8];
          narrow(m, m8);
          controlled({g.q[0]}, g.q[1], m8, second);
          return;
        }
        case K_SWAP:
          swap(g.q[0], g.q[1], second);
          return;
        default:
          break;
      }
    }
    if (g.nq == 3 && g.kind == K_CCX) {
      if (g.q[0] == g.q[1] || g.q[0] == g.q[2] || g.q[1] == g.q[2])
        fail("CCX on repeated wires");
      controlled({g.q[0], g.q[1]}, g.q[2], kX, second);
      return;
    }
    fail(std::string("unsupported gate: ") + kind_name(g.kind));
  }

  double pauli_sum(const float2 *bra, const float2 *ket, const Pauli &p,
                   double *slot) {
    launch_pauli(stream, bra, ket, dim, p.x | p.y, p.y | p.z, partials, slot);
    launched();
    return 0.0;
  }
};

// ---------------------------------------------------------------------------

bool Simulator::available() {
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  return count > 0;
}

std::string Simulator::device_name() {
  if (!available()) return "none";
  hipDeviceProp_t prop;
  check(hipGetDeviceProperties(&prop, current_device()), "hipGetDeviceProperties");
  return std::string(prop.name) + " (" + prop.gcnArchName + ")";
}

uint64_t Simulator::total_memory() {
  if (!available()) return 0;
  size_t fr = 0, tot = 0;
  check(hipMemGetInfo(&fr, &tot), "hipMemGetInfo");
  return tot;
}

uint64_t Simulator::free_memory() {
  if (!available()) return 0;
  size_t fr = 0, tot = 0;
  check(hipMemGetInfo(&fr, &tot), "hipMemGetInfo");
  return fr;
}

// Reference simulator.mm:272-326: validate n, preflight the allocation,
// allocate, write |0...0>.
Simulator::Simulator(uint32_t n) : impl_(new Impl) {
  if (n == 0 || n > kMaxQubits)
    fail("Simulator: num_qubits must be 1.." + std::to_string(kMaxQubits));
  if (!available()) fail("Simulator: no HIP device");
  Impl *s = impl_.get();
  s->n = n;
  s->dim = uint64_t(1) << n;
  s->device = current_device();
  const uint64_t bytes = s->dim * 8ull;
  if (const char *cap = getenv("QMETAL_ALLOC_CAP_BYTES")) {
    const uint64_t c = strtoull(cap, nullptr, 10);
    if (c && bytes > c)
      fail(std::to_string(n) + " qubits needs " + std::to_string(bytes) +
           " bytes, past QMETAL_ALLOC_CAP_BYTES=" + std::to_string(c));
  }
  const uint64_t freeb = free_memory();
  if (bytes > freeb)
    fail(std::to_string(n) + " qubits needs " + std::to_string(bytes >> 30) +
         " GiB, only " + std::to_string(freeb >> 30) + " GiB of device memory free");
  check(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking), "hipStreamCreate");
  check(hipEventCreateWithFlags(&s->staging_done, hipEventDisableTiming), "hipEventCreate");
  check(hipEventCreateWithFlags(&s->dev_done, hipEventDisableTiming), "hipEventCreate");
  s->state = s->alloc_state(bytes, &s->place_attempts, &s->place_ratio);
  if (!s->state)
    fail("Simulator: state allocation failed (" + std::to_string(bytes >> 30) + " GiB)");
  check(hipMalloc(reinterpret_cast<void **>(&s->partials), 2 * kRedBlocks * sizeof(double)),
        "hipMalloc(partials)");
  reset();
}

Simulator::~Simulator() {
  Impl *s = impl_.get();
  if (!s) return;
  if (s->stream) (void)hipStreamSynchronize(s->stream);
  if (s->dev_prog) {
    (void)hipDeviceSynchronize();  // its kernels ran on the callers' streams
    s->dev_prog.reset();
  }
  (void)hipFree(s->state);
  (void)hipFree(s->costate);
  (void)hipFree(s->partials);
  (void)hipFree(s->results);
  (void)hipFree(s->arena);
  (void)hipFree(s->cdf);
  (void)hipFree(s->adj_partials);
  (void)hipFree(s->chunks);
  (void)hipFree(s->shots_dev);
  (void)hipFree(s->pick_idx);
  (void)hipFree(s->pick_out);
  if (s->staging) (void)hipHostFree(s->staging);
  if (s->host_results) (void)hipHostFree(s->host_results);
  if (s->staging_done) (void)hipEventDestroy(s->staging_done);
  if (s->dev_done) (void)hipEventDestroy(s->dev_done);
  for (auto &t : s->timings) {
    (void)hipEventDestroy(t.a);
    (void)hipEventDestroy(t.b);
  }
  for (hipEvent_t e : s->event_pool) (void)hipEventDestroy(e);
  if (s->stream) (void)hipStreamDestroy(s->stream);
}

uint32_t Simulator::num_qubits() const { return impl_->n; }
uint64_t Simulator::dim() const { return impl_->dim; }
uint64_t Simulator::dispatch_count() const { return impl_->dispatches; }
uint64_t Simulator::launch_count() const { return impl_->launches; }
void *Simulator::device_state() { return impl_->state; }

void Simulator::reset() {
  Impl *s = impl_.get();
  s->use();
  launch_fill_basis(s->stream, s->state, s->dim);
  s->launched();
  s->dispatches++;
  s->launches++;
}

void Simulator::run(const Circuit &c, const FuseOpts &o) {
  if (c.n != impl_->n) fail("circuit/simulator size mismatch");
  check_circuit(c);
  run_plan(build_plan(c, o));
}

void Simulator::run_plan(const Plan &p) {
  Impl *s = impl_.get();
  if (p.n != s->n) fail("plan/simulator size mismatch");
  s->use();
  // Executor-level grouping below the plan: consecutive ops that fit one
  // register tile share one HBM pass (or each op runs alone).
  std::vector<Pass> passes;
  if (s->merge) {
    passes = schedule_passes(p);
  } else {
    for (uint32_t k = 0; k < p.ops.size(); k++) {
      Pass ps;
      ps.ops.push_back(k);
      ps.wmask.push_back(0);
      passes.push_back(std::move(ps));
    }
  }
  // Every parameterised op or pass writes its region of one staging block,
  // which is uploaded once, ahead of the plan's kernels on the same stream.
  std::vector<size_t> offs(passes.size());
  size_t total = 0;
  for (size_t k = 0; k < passes.size(); k++) {
    offs[k] = total;
    total += passes[k].merged ? s->pass_param_bytes(p, passes[k])
                              : s->param_bytes(p.ops[passes[k].ops[0]]);
  }
  std::vector<Impl::Pending> pend(passes.size());
  std::vector<Impl::PassLaunch> merged(passes.size());
  if (total) {
    s->stage(total);
    for (size_t k = 0; k < passes.size(); k++) {
      if (passes[k].wide) {
        s->prep_wide(p, passes[k], offs[k], merged[k]);
        continue;
      }
      if (passes[k].low) {
        s->prep_low(p, passes[k], offs[k], merged[k]);
        continue;
      }
      if (passes[k].tile) {
        s->prep_tile(p, passes[k], offs[k], merged[k]);
        continue;
      }
      if (passes[k].gwide) {
        s->prep_gwide(p, passes[k], offs[k], merged[k]);
        continue;
      }
      if (passes[k].merged) {
        s->prep_pass(p, passes[k], offs[k], merged[k]);
        continue;
      }
      const PlanOp &op = p.ops[passes[k].ops[0]];
      if (op.kind == OpKind::Diagonal && !op.masks.empty())
        s->prep_diagonal(op, offs[k], pend[k]);
      else if (op.kind == OpKind::Blocked)
        s->prep_blocked(op, offs[k], pend[k]);
    }
    s->upload(total);
  }
  for (size_t k = 0; k < passes.size(); k++) {
    if (passes[k].merged) s->pass_launch(merged[k]);
    else s->op(p.ops[passes[k].ops[0]], &pend[k]);
  }
  // The reference's accounting (simulator.mm:450-595): one dispatch per plan
  // op, none for a diagonal layer without terms, however the ops were merged.
  for (const PlanOp &op : p.ops)
    if (!(op.kind == OpKind::Diagonal && op.masks.empty())) s->dispatches++;
}

void Simulator::set_pass_merging(bool on) { impl_->merge = on; }
bool Simulator::pass_merging() const { return impl_->merge; }

void Simulator::placement(uint32_t *attempts, double *ratio) const {
  *attempts = impl_->place_attempts;
  *ratio = impl_->place_ratio;
}

void Simulator::set_timing(bool on) { impl_->timing = on; }

std::vector<OpTime> Simulator::read_timing() {
  Impl *s = impl_.get();
  s->use();
  check(hipStreamSynchronize(s->stream), "GPU error");
  std::vector<OpTime> out;
  for (const auto &t : s->timings) {
    float ms = 0.0f;
    check(hipEventElapsedTime(&ms, t.a, t.b), "hipEventElapsedTime");
    out.push_back(OpTime{t.kind, double(ms), t.bytes});
    s->event_pool.push_back(t.a);
    s->event_pool.push_back(t.b);
  }
  s->timings.clear();
  return out;
}

void Simulator::apply_gate(const Gate &g) {
  Impl *s = impl_.get();
  s->use();
  s->gate(g, nullptr);
  s->dispatches++;
  s->launches++;
}

void Simulator::synchronize() {
  Impl *s = impl_.get();
  s->use();
  check(hipStreamSynchronize(s->stream), "GPU error");
}

const float *Simulator::amplitudes() {
  Impl *s = impl_.get();
  s->use();
  s->mirror.resize(2 * s->dim);
  check(hipMemcpyAsync(s->mirror.data(), s->state, s->dim * 8, hipMemcpyDeviceToHost, s->stream),
        "hipMemcpyAsync(amplitudes)");
  check(hipStreamSynchronize(s->stream), "GPU error");
  return s->mirror.data();
}

// Amplitudes at chosen indices, for states too large to copy whole: the
// indices go to the device, one gather launch, the values come back.
void Simulator::amplitudes_at(const uint64_t *idx, size_t count, float *out) {
  Impl *s = impl_.get();
  for (size_t k = 0; k < count; k++)
    if (idx[k] >= s->dim)
      fail("amplitude index " + std::to_string(idx[k]) + " out of range for " +
           std::to_string(s->n) + " qubits");
  if (!count) return;
  s->use();
  s->grow(s->pick_idx, s->pick_idx_cap, count, "hipMalloc(pick indices)");
  s->grow(s->pick_out, s->pick_out_cap, count, "hipMalloc(pick values)");
  check(hipMemcpyAsync(s->pick_idx, idx, count * sizeof(uint64_t), hipMemcpyHostToDevice,
                       s->stream),
        "hipMemcpyAsync(pick indices)");
  launch_pick(s->stream, s->state, s->pick_idx, count, s->pick_out);
  s->launched();
  check(hipMemcpyAsync(out, s->pick_out, count * sizeof(float2), hipMemcpyDeviceToHost,
                       s->stream),
        "hipMemcpyAsync(pick values)");
  check(hipStreamSynchronize(s->stream), "GPU error");
}

// Reference simulator.mm:615-627, with FP64 accumulation on the GPU.
double Simulator::abs2sum() {
  Impl *s = impl_.get();
  s->use();
  double *slot = s->result_slots(1);
  launch_abs2(s->stream, s->state, s->dim, s->partials, slot);
  s->launched();
  s->dispatches++;
  s->launches++;
  return s->fetch_results(1)[0];
}

double Simulator::norm() { return std::sqrt(abs2sum()); }

namespace {
// Real part of i^ny * (re + i im) (reference simulator.mm:653-660).
double real_ipow(uint32_t ny, double re, double im) {
  switch (ny & 3u) {
    case 0: return re;
    case 1: return -im;
    case 2: return -re;
    default: return im;
  }
}
}  // namespace

double Simulator::expectation(const Pauli &p) {
  Hamiltonian h;
  h.push_back(Term{1.0, p});
  return expectation(h);
}

constexpr uint64_t kExpectTileMinBytes = uint64_t(256) << 20;

// sum_k c_k <psi|P_k|psi> (reference simulator.mm:663-667, each term as in
// simulator.mm:629-660): tile passes and per-term FP64 reductions (many
// terms to a launch), all enqueued before a single read-back.
double Simulator::expectation(const Hamiltonian &h) {
  Impl *s = impl_.get();
  for (const Term &t : h)
    if (!t.p.valid(s->n)) fail("expectation: invalid Pauli string for this width");
  if (h.empty()) return 0.0;
  s->use();
  // Terms whose flip holds at most five wires of 7 and up share read-only
  // co-state tile passes: each tile's terms act on the state in registers and
  // only <psi|H_t psi> leaves the pass, so a tile costs one read of the state
  // however many terms it holds. The other terms, narrow states and
  // Hamiltonians that would not save a pass take one reduction per term.
  const uint64_t low = (uint64_t(1) << 7) - 1;
  const bool tiles_ok = s->merge && s->n >= kAdjTileMinQubits && h.size() > 1;
  auto tileable = [&](const Term &t) {
    return tiles_ok && __builtin_popcountll((t.p.x | t.p.y) & ~low) <= 5;
  };
  Hamiltonian tiled;
  for (const Term &t : h)
    if (tileable(t)) tiled.push_back(t);
  // Below about 256 MiB of per-term reads the state stays in the caches and
  // one launch of per-term reductions beats the tile passes' launches
  // (TFIM ring: n = 19, 38 terms, 51 against 63 us; n = 20, 80 against 64 us).
  const double term_bytes = double(tiled.size()) * 8.0 * double(s->dim);
  Impl::CoPlan cp;
  if (!tiled.empty() && (term_bytes >= double(kExpectTileMinBytes) || s->expect_tiles_always))
    cp = Impl::costate_tile_plan(tiled, s->n, true);
  if (!cp.generic.empty()) fail("internal: expectation tile plan");
  if (cp.tiles.size() >= tiled.size()) cp.tiles.clear();
  std::vector<size_t> single;
  for (size_t k = 0; k < h.size(); k++)
    if (cp.tiles.empty() || !tileable(h[k])) single.push_back(k);
  const size_t ns = single.size(), nt = cp.tiles.size();
  double *slots = s->result_slots(2 * (ns + nt));
  // The one-reduction terms, up to kPauliTermsPerLaunch per launch; the
  // tiles; both leave their partials in the adjoint partials buffer, one
  // launch after the other.
  const size_t chunk = std::min<size_t>(ns, kPauliTermsPerLaunch);
  s->grow(s->adj_partials, s->adj_partials_cap,
          std::max(nt ? size_t(kAdjTileBlocks) * 2 : 0, chunk * pauli_terms_blocks(s->dim) * 2),
          "hipMalloc(adjoint partials)");
  for (size_t j0 = 0; j0 < ns; j0 += kPauliTermsPerLaunch) {
    const uint32_t count = uint32_t(std::min<size_t>(ns - j0, kPauliTermsPerLaunch));
    PauliTerms pt{};
    for (uint32_t j = 0; j < count; j++) {
      const Pauli &p = h[single[j0 + j]].p;
      pt.flip[j] = p.x | p.y;
      pt.sign[j] = p.y | p.z;
    }
    launch_pauli_terms(s->stream, s->state, s->dim, pt, count, s->adj_partials, slots + 2 * j0);
    s->launched();
    s->launches++;
  }
  for (size_t t = 0; t < nt; t++) {
    CoTileArgs &a = cp.tiles[t];
    a.psi = s->state;
    a.lam = nullptr;
    launch_costate_tile(s->stream, a, s->adj_partials, slots + 2 * (ns + t));
    s->launched();
    s->launches++;
  }
  // As the reference counts them: one reduction per term.
  s->dispatches += h.size();
  const double *r = s->fetch_results(2 * (ns + nt));
  double total = 0.0;
  for (size_t j = 0; j < ns; j++) {
    const Term &t = h[single[j]];
    total += t.c * real_ipow(uint32_t(__builtin_popcountll(t.p.y)), r[2 * j], r[2 * j + 1]);
  }
  for (size_t t = 0; t < nt; t++) total += r[2 * (ns + t)];
  return total;
}

// Reference simulator.mm:677-763, resolved on the GPU: FP64 block sums,
// FP64 prefix, then one thread per shot (splitmix64, binary search over
// blocks, in-block walk clamped to the last non-zero amplitude).
std::vector<uint64_t> Simulator::sample(size_t shots, uint64_t seed) {
  Impl *s = impl_.get();
  std::vector<uint64_t> out(shots);
  if (shots == 0) return out;
  s->use();
  const uint64_t nb = (s->dim + kSampleBlock - 1) / kSampleBlock;
  s->grow(s->cdf, s->cdf_cap, nb, "hipMalloc(cdf)");
  s->grow(s->chunks, s->chunks_cap, (nb + 1023) / 1024, "hipMalloc(chunks)");
  s->grow(s->shots_dev, s->shots_cap, shots, "hipMalloc(shots)");
  launch_block_sums(s->stream, s->state, s->dim, s->cdf);
  s->launched();
  launch_prefix(s->stream, s->cdf, nb, s->chunks);
  s->launched();
  launch_shots(s->stream, s->state, s->dim, s->cdf, nb, shots, seed, s->shots_dev);
  s->launched();
  s->dispatches++;
  s->launches++;
  check(hipMemcpyAsync(out.data(), s->shots_dev, shots * sizeof(uint64_t), hipMemcpyDeviceToHost,
                       s->stream),
        "hipMemcpyAsync(shots)");
  check(hipStreamSynchronize(s->stream), "GPU error");
  return out;
}

// --- allocation-free sampling ----------------------------------------------
//
// qm_sample_noalloc keeps its scratch in `scratch`, one copy per GPU for the
// life of the process: the 1024-amplitude block sums that become the CDF
// (states of up to kNoAllocMaxQubits qubits), the prefix scan's chunk totals,
// and the drawn indices (up to kNoAllocMaxShots). Nothing in it needs
// clearing: each kernel writes every element the next one reads. Every call
// on a GPU shares that GPU's copy, so calls that overlap on one GPU (from
// several threads) race; calls made one after another do not, because each
// waits for its own work before returning.
namespace {
constexpr uint32_t kNoAllocMaxQubits = 30;
constexpr uint64_t kNoAllocMaxShots = uint64_t(1) << 20;
constexpr uint64_t kNoAllocBlocks = (uint64_t(1) << kNoAllocMaxQubits) / kSampleBlock;
constexpr uint64_t kNoAllocChunks = (kNoAllocBlocks + 1023) / 1024;
constexpr uint64_t kScratchDoubles = kNoAllocBlocks + kNoAllocChunks + kNoAllocMaxShots;
constexpr int kMaxDevices = 64;

static __device__ __align__(256) double scratch[kScratchDoubles];

std::mutex g_scratch_mu;
double *g_scratch[kMaxDevices] = {};

// Device address of the current GPU's `scratch`. In host code the array's
// name is only a stand-in for the symbol, so the address is looked up once
// per device, with that device current, and used wherever the host hands the
// memory to a kernel or a copy.
double *scratch_here() {
  const int d = current_device();
  if (d < 0 || d >= kMaxDevices) fail("sample_noalloc: device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch[d]) {
    void *p = nullptr;
    check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
    if (!p) fail("sample_noalloc: no device address for the scratch array");
    g_scratch[d] = static_cast<double *>(p);
  }
  return g_scratch[d];
}
}  // namespace

// qm_sample's procedure and kernels with the scratch above instead of the
// simulator's grown buffers: no device or pinned allocation on any call.
void Simulator::sample_noalloc(size_t shots, uint64_t seed, uint64_t *out) {
  Impl *s = impl_.get();
  if (s->n > kNoAllocMaxQubits)
    fail("sample_noalloc: " + std::to_string(s->n) + " qubits is past the limit of " +
         std::to_string(kNoAllocMaxQubits) + " (its fixed scratch holds the CDF of 2^" +
         std::to_string(kNoAllocMaxQubits) + " amplitudes); use qm_sample");
  if (shots > kNoAllocMaxShots)
    fail("sample_noalloc: " + std::to_string(shots) + " shots is past the limit of " +
         std::to_string(kNoAllocMaxShots) + " per call; use qm_sample");
  if (shots == 0) return;
  s->use();
  double *base = scratch_here();
  double *cdf = base;
  double *chunks = base + kNoAllocBlocks;
  uint64_t *drawn = reinterpret_cast<uint64_t *>(base + kNoAllocBlocks + kNoAllocChunks);
  const uint64_t nb = (s->dim + kSampleBlock - 1) / kSampleBlock;
  launch_block_sums(s->stream, s->state, s->dim, cdf);
  s->launched();
  launch_prefix(s->stream, cdf, nb, chunks);
  s->launched();
  launch_shots(s->stream, s->state, s->dim, cdf, nb, shots, seed, drawn);
  s->launched();
  s->dispatches++;
  s->launches++;
  check(hipMemcpyAsync(out, drawn, shots * sizeof(uint64_t), hipMemcpyDeviceToHost, s->stream),
        "hipMemcpyAsync(shots)");
  check(hipStreamSynchronize(s->stream), "GPU error");
}



// Adjoint gradient (reference simulator.mm:814-963). The forward pass runs
// the fused plan (only the final state is needed); the co-state is built
// with one pass per distinct Pauli flip; E = Re<psi|lambda> is an FP64
// reduction; the backward sweep fuses, per rotation, the FP64 overlap
// <lambda|P|psi> with the undo of that rotation on both states, and undoes
// every other gate on both states in one gather pass. Overlaps stay on the
// device until one read-back at the end.
std::pair<double, std::vector<double>> Simulator::energy_and_gradient(
    const Circuit &c, const Hamiltonian &h) {
  Impl *s = impl_.get();
  if (c.n != s->n) fail("circuit/simulator size mismatch");
  for (const Gate &g : c.gates)
    if (g.np > 0 && !is_rotation(g.kind))
      fail(std::string("gradient: ") + kind_name(g.kind) +
           " carries a parameter but has no native generator; only rx, ry and rz are "
           "differentiable");
  check_circuit(c);
  for (const Term &t : h)
    if (!t.p.valid(s->n)) fail("gradient: invalid Pauli term");
  s->use();
  if (!s->costate) {
    const uint64_t bytes = s->dim * 8ull;
    if (bytes > free_memory()) fail("gradient: co-state exceeds free device memory");
    uint32_t tries = 0;
    double ratio = 0.0;
    s->costate = s->alloc_state(bytes, &tries, &ratio);
    if (!s->costate) fail("gradient: co-state allocation failed");
  }
  reset();
  run_plan(build_plan(c, FuseOpts{}));

  // lambda = H psi, terms grouped by flip mask.
  const double amp_bytes = 8.0 * double(s->dim);
  const size_t G = c.gates.size();
  size_t nrot_total = 0;
  for (const Gate &g : c.gates) nrot_total += is_rotation(g.kind) ? 1 : 0;
  // Slots: [0, 2G) per gate (one launch per gate), [2G, 2G + 2) energy,
  // then per rotation in sweep order (merged passes).
  double *slots = s->result_slots(2 * (G + 1) + 2 * nrot_total);
  s->grow(s->adj_partials, s->adj_partials_cap,
          std::max(size_t(kAdjMaxRot) * kAdjBlocks, size_t(kAdjTileMaxRot) * kAdjTileBlocks) * 2,
          "hipMalloc(adjoint partials)");
  Impl::CoPlan cp;
  if (s->merge) cp = Impl::costate_tile_plan(h, s->n);
  bool energy_done = false;
  if (!cp.tiles.empty()) {
    // Tiles first (the first writes lambda), then the flips no tile holds;
    // the energy rides with the last tile when it is the last co-state pass.
    for (size_t k = 0; k < cp.tiles.size(); k++) {
      CoTileArgs &a = cp.tiles[k];
      a.psi = s->state;
      a.lam = s->costate;
      const bool fuse = k + 1 == cp.tiles.size() && cp.generic.empty();
      s->timed(kTimingCoState, (a.first ? 2 : 3) * amp_bytes, [&] {
        launch_costate_tile(s->stream, a, s->adj_partials, fuse ? slots + 2 * G : nullptr);
      });
      energy_done = fuse;
      s->launched();
      s->dispatches++;
      s->launches++;
    }
    for (Impl::AccumLaunch &L : cp.generic) {
      L.g.lam = s->costate;
      L.g.psi = s->state;
      L.g.dim = s->dim;
      s->timed(kTimingCoState, 3 * amp_bytes, [&] { launch_pauli_accum(s->stream, L.g); });
      s->launched();
      s->dispatches++;
      s->launches++;
    }
  } else {
  s->timed(kTimingCoState, amp_bytes, [&] { launch_zero(s->stream, s->costate, s->dim); });
  s->launched();
  for (Impl::AccumLaunch &L : Impl::costate_launches(h)) {
    if (L.tile) {
      L.t.lam = s->costate;
      L.t.psi = s->state;
      L.t.dim = s->dim;
      s->timed(kTimingCoState, 3 * amp_bytes, [&] { launch_pauli_accum_tile(s->stream, L.t); });
    } else {
      L.g.lam = s->costate;
      L.g.psi = s->state;
      L.g.dim = s->dim;
      s->timed(kTimingCoState, 3 * amp_bytes, [&] { launch_pauli_accum(s->stream, L.g); });
    }
    s->launched();
    s->dispatches++;
    s->launches++;
  }
  }
  if (!energy_done) {
    // E = Re<psi|lambda>
    s->timed(kTimingOverlap, 2 * amp_bytes,
             [&] { s->pauli_sum(s->state, s->costate, Pauli{}, slots + 2 * G); });
    s->dispatches++;
    s->launches++;
  }

  if (s->merge && s->n >= 6) {
    // Backward sweep in merged passes (adjoint_schedule); the ops of every
    // pass in one staging block.
    std::vector<size_t> rot_index;
    const std::vector<Impl::APass> passes = Impl::adjoint_schedule(c, rot_index);
    std::vector<size_t> offs(passes.size());
    size_t total = 0;
    for (size_t i = 0; i < passes.size(); i++) {
      offs[i] = total;
      total += Impl::adj_pass_bytes(c, passes[i]);
    }
    s->stage(total);
    std::vector<AdjPassArgs> args(passes.size());
    std::vector<AdjTileArgs> targs(passes.size());
    for (size_t i = 0; i < passes.size(); i++) {
      if (passes[i].tile) {
        AdjTileArgs &t = targs[i];
        s->emit_adj_tile(c, passes[i], s->staging + offs[i], s->arena + offs[i], t);
        t.psi = s->state;
        t.lam = s->costate;
        continue;
      }
      AdjPassArgs &a = args[i];
      s->emit_adj_pass(c, passes[i], reinterpret_cast<AdjOp *>(s->staging + offs[i]), a);
      a.psi = s->state;
      a.lam = s->costate;
      a.ops = reinterpret_cast<const AdjOp *>(s->arena + offs[i]);
    }
    s->upload(total);
    double *rot_slots = slots + 2 * (G + 1);
    for (size_t i = 0; i < passes.size(); i++) {
      s->timed(kTimingAdjPass, 4 * amp_bytes, [&] {
        if (passes[i].tile)
          launch_adj_tile(s->stream, targs[i], s->adj_partials, rot_slots + 2 * passes[i].rot0);
        else
          launch_adj_pass(s->stream, args[i], s->adj_partials, rot_slots + 2 * passes[i].rot0);
      });
      s->launched();
      s->dispatches++;
      s->launches++;
    }
    const double *r = s->fetch_results(2 * (G + 1) + 2 * nrot_total);
    std::vector<double> grad(G, 0.0);
    for (size_t k = 0; k < G; k++) {
      const Gate &g = c.gates[k];
      if (!differentiable(g)) continue;
      const double *o = r + 2 * (G + 1) + 2 * rot_index[k];
      grad[k] = g.kind == K_RY ? o[0] : o[1];
    }
    return {r[2 * G], grad};
  }

  for (size_t k = G; k-- > 0;) {
    const Gate &g = c.gates[k];
    const Gate inv = inverse(g);
    if (is_rotation(g.kind)) {
      cd m[4];
      gate_matrix(inv.kind, inv.p, m);
      AdjArgs a{};
      a.psi = s->state;
      a.lam = s->costate;
      a.dim = s->dim;
      a.q = g.q[0];
      a.axis = g.kind == K_RX ? 0 : g.kind == K_RY ? 1 : 2;
      narrow(m, a.m);
      s->timed(kTimingAdjRot, 4 * amp_bytes,
               [&] { launch_adjoint_rot(s->stream, a, s->partials, slots + 2 * k); });
      s->launched();
    } else {
      s->gate(inv, s->costate);
    }
    s->dispatches++;
    s->launches++;
  }
  const double *r = s->fetch_results(2 * (G + 1));
  std::vector<double> grad(G, 0.0);
  for (size_t k = 0; k < G; k++) {
    const Gate &g = c.gates[k];
    if (!differentiable(g)) continue;
    // grad = Im(i^ny S): ny = 1 for RY, 0 for RX/RZ.
    grad[k] = g.kind == K_RY ? r[2 * k] : r[2 * k + 1];
  }
  return {r[2 * G], grad};
}

// Variant for framework tensors: parameters are read from device memory (one
// per RX/RY/RZ, circuit order), energy and gradient are written to device
// memory, and the caller's stream is ordered before and after this work
// through events. The parameters cross to the host (the plan's window
// products and phase tables are built there), so the call waits for the
// caller's stream up to that point and cannot be captured in a graph; the
// results go back asynchronously. Without out_grad only the energy is
// computed: the forward passes and one FP64 reduction per term, no co-state
// and no backward sweep.
namespace {
template <typename T>
T *device_vector(const std::vector<T> &v, const char *what) {
  T *p = nullptr;
  check(hipMalloc(reinterpret_cast<void **>(&p), std::max<size_t>(v.size(), 1) * sizeof(T)), what);
  if (!v.empty())
    check(hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice), what);
  return p;
}

bool same_shape(const std::vector<Gate> &a, const std::vector<Gate> &b) {
  if (a.size() != b.size()) return false;
  for (size_t k = 0; k < a.size(); k++)
    if (a[k].kind != b[k].kind || a[k].nq != b[k].nq ||
        std::memcmp(a[k].q, b[k].q, sizeof(a[k].q)) != 0 ||
        (!is_rotation(a[k].kind) && std::memcmp(a[k].p, b[k].p, sizeof(a[k].p)) != 0))
      return false;
  return true;
}
bool same_hamiltonian(const Hamiltonian &a, const Hamiltonian &b) {
  if (a.size() != b.size()) return false;
  for (size_t k = 0; k < a.size(); k++)
    if (a[k].c != b[k].c || a[k].p.x != b[k].p.x || a[k].p.y != b[k].p.y || a[k].p.z != b[k].p.z)
      return false;
  return true;
}
}  // namespace

// Compiles the device-resident program for template c (every rotation a
// parameter) and Hamiltonian h; false when the shape is not supported (then
// the host path serves the call).
bool Simulator::compile_device_program(const Circuit &c, const Hamiltonian &h) {
  Impl *s = impl_.get();
  if (!s->merge || s->n < kAdjTileMinQubits || h.empty()) return false;
  if (s->dev_prog && same_shape(s->dev_prog->shape, c.gates) &&
      same_hamiltonian(s->dev_prog->h, h))
    return true;
  s->dev_prog.reset();
  Circuit cc = c;
  std::vector<uint32_t> pidx(c.gates.size(), 0);
  uint32_t np = 0;
  for (size_t k = 0; k < cc.gates.size(); k++)
    if (is_rotation(cc.gates[k].kind)) {
      pidx[k] = np++;
      cc.gates[k].np = std::max<uint8_t>(cc.gates[k].np, 1);
    }
  Impl::CoPlan cp = Impl::costate_tile_plan(h, s->n);
  if (cp.tiles.empty()) return false;
  // Forward sweep: the template's plan without fusion (one op per gate, so
  // the ops and their matrix structure do not depend on the angles) in the
  // executor's merged passes. Every rotation is a Dense1q slot whose section
  // is fixed by its kind (RY real, RZ diagonal, RX real diagonal with an
  // imaginary off-diagonal, as at the placeholder angle 1) and whose 2x2 is
  // patched from the parameters. A schedule with chain-shaped passes keeps
  // the single-state tile passes.
  Circuit pc = cc;
  for (Gate &g : pc.gates)
    if (is_rotation(g.kind)) g.p[0] = 1.0;
  const Plan fplan = build_plan(pc, FuseOpts::none());
  std::vector<Pass> mpasses;
  bool merged_fwd = fplan.ops.size() == pc.gates.size();
  for (size_t k = 0; merged_fwd && k < pc.gates.size(); k++)
    if (is_rotation(pc.gates[k].kind))
      merged_fwd = fplan.ops[k].kind == OpKind::Dense1q && fplan.ops[k].q[0] == pc.gates[k].q[0];
  if (merged_fwd) {
    for (Pass &ps : schedule_passes(fplan)) {
      if (ps.wide || ps.low) {
        merged_fwd = false;
        break;
      }
      if (!ps.merged) {
        const PlanOp &op = fplan.ops[ps.ops[0]];
        if (op.kind == OpKind::Diagonal && op.masks.empty()) continue;  // no work
        if (!pass_mergeable(op)) {
          merged_fwd = false;
          break;
        }
        ps.merged = true;
        ps.reg = pass_wires(op);
        ps.fixed = pass_fixed(op);
        if (ps.reg.size() > kPassMaxReg) {
          merged_fwd = false;
          break;
        }
      }
      mpasses.push_back(std::move(ps));
    }
  }
  if (!merged_fwd) mpasses.clear();
  std::vector<size_t> rot_index, unused;
  const std::vector<Impl::APass> fwd =
      merged_fwd ? std::vector<Impl::APass>{} : Impl::adjoint_tile_schedule(cc, unused, true);
  const std::vector<Impl::APass> bwd = Impl::adjoint_tile_schedule(cc, rot_index, false);
  std::vector<size_t> moff(mpasses.size()), foff(fwd.size()), boff(bwd.size());
  size_t total = 0;
  for (size_t k = 0; k < mpasses.size(); k++) {
    moff[k] = total;
    total = (total + s->pass_param_bytes(fplan, mpasses[k]) + 255) & ~size_t(255);
  }
  for (size_t i = 0; i < fwd.size(); i++) {
    foff[i] = total;
    total += Impl::adj_pass_bytes(cc, fwd[i]);
  }
  for (size_t i = 0; i < bwd.size(); i++) {
    boff[i] = total;
    total += Impl::adj_pass_bytes(cc, bwd[i]);
  }
  std::unique_ptr<Impl::DevProg> d(new Impl::DevProg);
  d->shape = c.gates;
  d->h = h;
  check(hipMalloc(reinterpret_cast<void **>(&d->arena), std::max<size_t>(total, 256)),
        "hipMalloc(device program)");
  std::vector<uint8_t> host(std::max<size_t>(total, 256));
  std::vector<AnglePatch> patches;
  auto emit = [&](const Impl::APass &ps, size_t off, bool forward, AdjTileArgs &a) {
    const size_t before = patches.size();
    s->emit_adj_tile(cc, ps, host.data() + off, d->arena + off, a, forward, &patches);
    for (size_t k = before; k < patches.size(); k++) {
      patches[k].off += uint32_t(off);
      patches[k].param = pidx[patches[k].param];
    }
  };
  if (merged_fwd) {
    // Through the executor's own pass preparation, pointed at this
    // program's host block and arena; chain kernels stay off (their real /
    // complex choice would depend on the angles).
    std::vector<std::pair<size_t, uint32_t>> slots;
    struct Swap {
      Impl *s;
      uint8_t *st, *ar;
      ~Swap() {
        s->staging = st;
        s->arena = ar;
        s->batch_compile = false;
        s->slot_log = nullptr;
      }
    } swap{s, s->staging, s->arena};
    s->staging = host.data();
    s->arena = d->arena;
    s->batch_compile = true;
    s->slot_log = &slots;
    for (size_t k = 0; k < mpasses.size(); k++) {
      Impl::PassLaunch pl{};
      if (mpasses[k].tile) s->prep_tile(fplan, mpasses[k], moff[k], pl);
      else if (mpasses[k].gwide) s->prep_gwide(fplan, mpasses[k], moff[k], pl);
      else s->prep_pass(fplan, mpasses[k], moff[k], pl);
      if (pl.chain || pl.wide || pl.low) fail("internal: chain pass in a device program");
      d->mfwd.push_back(pl);
    }
    for (const auto &e : slots) {
      const Gate &g = cc.gates[e.second];
      if (!is_rotation(g.kind)) continue;
      const uint32_t rk = g.kind == K_RX ? 0u : g.kind == K_RY ? 1u : 2u;
      patches.push_back(AnglePatch{uint32_t(e.first + offsetof(PassOp, m)), pidx[e.second],
                                   0u | (rk << 4) | kPatchPassOp, 0});
    }
  }
  for (size_t i = 0; i < fwd.size(); i++) {
    AdjTileArgs a{};
    emit(fwd[i], foff[i], true, a);
    d->fwd.push_back(a);
  }
  for (size_t i = 0; i < bwd.size(); i++) {
    AdjTileArgs a{};
    emit(bwd[i], boff[i], false, a);
    d->bwd.push_back(a);
    d->bwd_rot0.push_back(bwd[i].rot0);
  }
  check(hipMemcpy(d->arena, host.data(), host.size(), hipMemcpyHostToDevice), "device program");
  d->npatch = uint32_t(patches.size());
  d->patch = device_vector(patches, "device program patches");
  for (size_t k = 0; k < cc.gates.size(); k++)
    if (is_rotation(cc.gates[k].kind)) d->nrot++;
  std::vector<uint32_t> rgate(std::max<uint32_t>(d->nrot, 1), 0);
  for (size_t k = 0; k < cc.gates.size(); k++)
    if (is_rotation(cc.gates[k].kind))
      rgate[rot_index[k]] = pidx[k] | (cc.gates[k].kind == K_RY ? 0x80000000u : 0u);
  d->rgate = device_vector(rgate, "device program rotation map");
  d->inst = device_vector(std::vector<uint32_t>{0u}, "device program instance");
  check(hipMalloc(reinterpret_cast<void **>(&d->eslot), 2 * sizeof(double)), "hipMalloc");
  check(hipMalloc(reinterpret_cast<void **>(&d->rslots),
                  std::max<size_t>(d->nrot, 1) * 2 * sizeof(double)),
        "hipMalloc");
  d->co = cp.tiles;
  for (const Impl::AccumLaunch &L : cp.generic) d->co_generic.push_back(L.g);
  if (s->merge && h.size() > 1) {
    Impl::CoPlan ep = Impl::costate_tile_plan(h, s->n, true);
    if (ep.generic.empty() && !ep.tiles.empty()) {
      d->eco = ep.tiles;
      check(hipMalloc(reinterpret_cast<void **>(&d->eslots), d->eco.size() * 2 * sizeof(double)),
            "hipMalloc");
    }
  }
  s->dev_prog = std::move(d);
  return true;
}

void Simulator::energy_and_gradient_device(const Circuit &c, const Hamiltonian &h,
                                           const double *params, size_t nparams,
                                           double *out_energy, double *out_grad,
                                           void *caller_stream) {
  Impl *s = impl_.get();
  s->use();
  hipStream_t cs = static_cast<hipStream_t>(caller_stream);
  {
    size_t nrot = 0;
    for (const Gate &g : c.gates) nrot += is_rotation(g.kind) ? 1 : 0;
    if (nrot != nparams)
      fail("gradient: " + std::to_string(nparams) + " parameters given, circuit has " +
           std::to_string(nrot) + " rotations");
    for (const Gate &g : c.gates)
      if (g.np > 0 && !is_rotation(g.kind))
        fail(std::string("gradient: ") + kind_name(g.kind) +
             " carries a parameter but has no native generator; only rx, ry and rz are "
             "differentiable");
    check_circuit(c);
    for (const Term &t : h)
      if (!t.p.valid(s->n)) fail("gradient: invalid Pauli term");
  }
  const bool fresh = !s->dev_prog || !same_shape(s->dev_prog->shape, c.gates) ||
                     !same_hamiltonian(s->dev_prog->h, h);
  if (fresh || s->launches != s->dev_launches) {
    // Other work on the simulator's stream (or the compile below) comes first.
    check(hipStreamSynchronize(s->stream), "GPU error");
  }
  if (c.n == s->n && compile_device_program(c, h)) {
    const bool energy_only = !out_grad && !s->dev_prog->eco.empty();
    if (!s->costate && !energy_only) {
      const uint64_t bytes = s->dim * 8ull;
      if (bytes > free_memory()) fail("gradient: co-state exceeds free device memory");
      uint32_t tries = 0;
      double ratio = 0.0;
      s->costate = s->alloc_state(bytes, &tries, &ratio);
      if (!s->costate) fail("gradient: co-state allocation failed");
    }
    s->grow(s->adj_partials, s->adj_partials_cap,
            std::max(size_t(kAdjMaxRot) * kAdjBlocks, size_t(kAdjTileMaxRot) * kAdjTileBlocks) *
                2,
            "hipMalloc(adjoint partials)");
    Impl::DevProg &d = *s->dev_prog;
    uint64_t n_launch = 0;
    launch_patch_angles(cs, d.arena, d.patch, d.npatch, params);
    launch_fill_basis(cs, s->state, s->dim);
    n_launch += 2;
    for (Impl::PassLaunch &pl : d.mfwd) {
      if (pl.tile) {
        pl.t.a = s->state;
        launch_tile(cs, pl.t);
      } else if (pl.gwide) {
        pl.g.a = s->state;
        launch_gwide(cs, pl.chain_k - 4, pl.g);
      } else {
        pl.a.a = s->state;
        launch_pass(cs, pl.nl, pl.a);
      }
      n_launch++;
    }
    for (AdjTileArgs &a : d.fwd) {
      a.psi = s->state;
      a.lam = s->state;
      launch_fwd_tile(cs, a);
      n_launch++;
    }
    if (energy_only) {
      for (size_t k = 0; k < d.eco.size(); k++) {
        CoTileArgs &a = d.eco[k];
        a.psi = s->state;
        a.lam = nullptr;
        launch_costate_tile(cs, a, s->adj_partials, d.eslots + 2 * k);
        n_launch++;
      }
      launch_sum_slots(cs, d.eslots, uint32_t(d.eco.size()), d.eslot);
      n_launch++;
    } else {
      bool fused = false;
      for (size_t k = 0; k < d.co.size(); k++) {
        CoTileArgs &a = d.co[k];
        a.psi = s->state;
        a.lam = s->costate;
        const bool fuse = k + 1 == d.co.size() && d.co_generic.empty();
        launch_costate_tile(cs, a, s->adj_partials, fuse ? d.eslot : nullptr);
        fused = fused || fuse;
        n_launch++;
      }
      for (AccumArgs &g : d.co_generic) {
        g.psi = s->state;
        g.lam = s->costate;
        g.dim = s->dim;
        launch_pauli_accum(cs, g);
        n_launch++;
      }
      if (!fused) {
        launch_pauli(cs, s->state, s->costate, s->dim, 0, 0, s->partials, d.eslot);
        n_launch++;
      }
    }
    if (out_grad) {
      for (size_t j = 0; j < d.bwd.size(); j++) {
        AdjTileArgs &a = d.bwd[j];
        a.psi = s->state;
        a.lam = s->costate;
        launch_adj_tile(cs, a, s->adj_partials, d.rslots + 2 * d.bwd_rot0[j]);
        n_launch++;
      }
    }
    launch_batched_results(cs, d.eslot, d.rslots, 1, out_grad ? d.nrot : 0, d.rgate, d.inst,
                           d.inst, out_energy, out_grad);
    check(hipGetLastError(), "device gradient launch");
    // Later work on this simulator's own stream waits for these kernels (a
    // captured call is ordered by whoever replays the graph).
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    check(hipStreamIsCapturing(cs, &cap), "hipStreamIsCapturing");
    if (cap == hipStreamCaptureStatusNone) {
      check(hipEventRecord(s->dev_done, cs), "hipEventRecord(device program)");
      s->dev_pending = true;
    }
    s->launches += n_launch + 1;
    s->dispatches += n_launch + 1;
    s->dev_launches = s->launches;
    return;
  }
  hipEvent_t ev;
  check(hipEventCreateWithFlags(&ev, hipEventDisableTiming), "hipEventCreate");
  struct EventGuard {
    hipEvent_t e;
    ~EventGuard() { (void)hipEventDestroy(e); }
  } guard{ev};
  check(hipEventRecord(ev, cs), "hipEventRecord(caller)");
  check(hipStreamWaitEvent(s->stream, ev, 0), "hipStreamWaitEvent");
  std::vector<size_t> where;
  for (size_t k = 0; k < c.gates.size(); k++)
    if (is_rotation(c.gates[k].kind)) where.push_back(k);
  if (where.size() != nparams)
    fail("gradient: " + std::to_string(nparams) + " parameters given, circuit has " +
         std::to_string(where.size()) + " rotations");
  std::vector<double> host(nparams + 1);
  if (nparams) {
    check(hipMemcpyAsync(host.data(), params, nparams * sizeof(double), hipMemcpyDeviceToHost,
                         s->stream),
          "hipMemcpyAsync(params)");
    check(hipStreamSynchronize(s->stream), "GPU error");
  }
  Circuit cc = c;
  for (size_t j = 0; j < nparams; j++) {
    cc.gates[where[j]].p[0] = host[j];
    cc.gates[where[j]].np = std::max<uint8_t>(cc.gates[where[j]].np, 1);
  }
  std::pair<double, std::vector<double>> eg;
  if (out_grad) {
    eg = energy_and_gradient(cc, h);
  } else {
    reset();
    run(cc, FuseOpts{});
    eg.first = expectation(h);
  }
  // Results leave through pinned staging so the copies stay asynchronous.
  uint8_t *st = s->stage((nparams + 1) * sizeof(double));
  double *hv = reinterpret_cast<double *>(st);
  hv[0] = eg.first;
  if (out_grad)
    for (size_t j = 0; j < nparams; j++) hv[1 + j] = eg.second[where[j]];
  check(hipMemcpyAsync(out_energy, hv, sizeof(double), hipMemcpyHostToDevice, s->stream),
        "hipMemcpyAsync(energy)");
  if (nparams && out_grad)
    check(hipMemcpyAsync(out_grad, hv + 1, nparams * sizeof(double), hipMemcpyHostToDevice,
                         s->stream),
          "hipMemcpyAsync(grad)");
  check(hipEventRecord(s->staging_done, s->stream), "hipEventRecord");
  s->staging_pending = true;
  check(hipEventRecord(ev, s->stream), "hipEventRecord(done)");
  check(hipStreamWaitEvent(cs, ev, 0), "hipStreamWaitEvent(caller)");
}

// --- batched adjoint programs ------------------------------------------------

namespace {
// One forward pass of a batched program: the generic register-tile, low tile
// or wide tile kernel with its (instance 0) arguments.
struct BatchedPass {
  int kind = 0;  // 0 register tile, 1 low tile, 2 wide tile
  int nl = 0, wb = 0;
  PassArgs a{};
  TileArgs t{};
  GWideArgs g{};
};
struct BatchedAccum {
  bool tile = false;
  bool cotile = false;  // k_costate_tile (c); `fuse`: it leaves the energies
  bool fuse = false;
  AccumTileArgs t{};
  AccumArgs g{};
  CoTileArgs c{};
};

bool same_layout(const Layout &x, const Layout &y) {
  if (x.nthreads != y.nthreads || x.fixed != y.fixed || x.nins != y.nins || x.streaming != y.streaming)
    return false;
  for (uint32_t k = 0; k < x.nins; k++)
    if (x.ins[k] != y.ins[k] || x.hm[k] != y.hm[k]) return false;
  return std::memcmp(x.rpos, y.rpos, sizeof(x.rpos)) == 0;
}

// Plans equal in everything the executor's pass structure reads: kinds,
// wires, masks, block bits and which phase layers have a zero global phase.
bool same_structure(const Plan &x, const Plan &y) {
  if (x.n != y.n || x.ops.size() != y.ops.size()) return false;
  for (size_t k = 0; k < x.ops.size(); k++) {
    const PlanOp &a = x.ops[k], &b = y.ops[k];
    if (a.kind != b.kind || std::memcmp(a.q, b.q, sizeof(a.q)) != 0 || a.wires != b.wires ||
        a.masks != b.masks || a.block_bits != b.block_bits ||
        (turns(a.global) == 0.0) != (turns(b.global) == 0.0))
      return false;
  }
  return true;
}

}  // namespace

struct BatchedAdjoint::Data {
  uint32_t n = 0, count = 0, nrot = 0;
  uint64_t dim = 0;
  float2 *psi = nullptr, *lam = nullptr;
  uint8_t *arena = nullptr;
  double *red = nullptr, *adj = nullptr, *eslots = nullptr, *rslots = nullptr;
  uint32_t *rgate = nullptr, *inst = nullptr, *gbase = nullptr;
  std::vector<BatchedPass> fwd;
  std::vector<BatchedAccum> costate;
  struct Bwd {
    bool tile = false;
    AdjPassArgs a{};
    AdjTileArgs t{};
    size_t nops = 0;
  };
  std::vector<Bwd> bwd;
  std::vector<size_t> bwd_rot0;

  ~Data() {
    (void)hipFree(psi);
    (void)hipFree(lam);
    (void)hipFree(arena);
    (void)hipFree(red);
    (void)hipFree(adj);
    (void)hipFree(eslots);
    (void)hipFree(rslots);
    (void)hipFree(rgate);
    (void)hipFree(inst);
    (void)hipFree(gbase);
  }
};

BatchedAdjoint::BatchedAdjoint() : d_(new Data) {}
BatchedAdjoint::~BatchedAdjoint() = default;
uint32_t BatchedAdjoint::count() const { return d_->count; }
uint32_t BatchedAdjoint::num_qubits() const { return d_->n; }

// The single-circuit adjoint (energy_and_gradient with merged passes) for
// every instance at once: |0...0>, the forward passes, lambda = H psi, E =
// Re<psi|lambda>, the merged backward passes with their overlaps, then the
// results scattered to the caller's arrays.
void BatchedAdjoint::enqueue(void *stream, double *energy_out, double *grad_out) const {
  const Data &d = *d_;
  hipStream_t st = static_cast<hipStream_t>(stream);
  launch_fill_basis_batched(st, d.psi, d.dim, d.count);
  for (const BatchedPass &b : d.fwd) {
    if (b.kind == 1) launch_tile_batched(st, b.t, d.count);
    else if (b.kind == 2) launch_gwide_batched(st, b.wb, b.g, d.count);
    else launch_pass_batched(st, b.nl, b.a, d.count);
  }
  const bool cotile = !d.costate.empty() && d.costate[0].cotile;
  if (!cotile) launch_zero_batched(st, d.lam, d.dim, d.count);
  bool fused = false;
  for (const BatchedAccum &b : d.costate) {
    if (b.cotile) {
      launch_costate_tile_batched(st, b.c, d.red, b.fuse ? d.eslots : nullptr, d.count);
      fused = fused || b.fuse;
    } else if (b.tile) {
      launch_pauli_accum_tile_batched(st, b.t, d.count);
    } else {
      launch_pauli_accum_batched(st, b.g, d.count);
    }
  }
  if (!fused) launch_pauli_batched(st, d.psi, d.lam, d.dim, 0, 0, d.red, d.eslots, d.count);
  for (size_t j = 0; j < d.bwd.size(); j++) {
    double *out = d.rslots + 2 * d.bwd_rot0[j];
    if (d.bwd[j].tile) launch_adj_tile_batched(st, d.bwd[j].t, d.adj, out, 2ull * d.nrot, d.count);
    else launch_adj_pass_batched(st, d.bwd[j].a, d.adj, out, 2ull * d.nrot, d.count);
  }
  launch_batched_results(st, d.eslots, d.rslots, d.count, d.nrot, d.rgate, d.inst, d.gbase,
                         energy_out, grad_out);
  check(hipGetLastError(), "batched adjoint launch");
}

std::unique_ptr<BatchedAdjoint> Simulator::compile_batched(
    const std::vector<const Circuit *> &circuits, const Hamiltonian &h,
    const std::vector<uint32_t> &inst, const std::vector<uint32_t> &gbase) {
  Impl *s = impl_.get();
  const size_t m = circuits.size();
  if (!m || inst.size() != m || gbase.size() != m) fail("compile_batched: argument sizes");
  if (s->n < kPassMinQubits) return nullptr;
  const Circuit &c0 = *circuits[0];
  for (const Circuit *c : circuits) {
    if (c->n != s->n || c->gates.size() != c0.gates.size()) return nullptr;
    for (size_t k = 0; k < c0.gates.size(); k++) {
      const Gate &a = c->gates[k], &b = c0.gates[k];
      if (a.kind != b.kind || a.nq != b.nq || (a.np > 0) != (b.np > 0) ||
          std::memcmp(a.q, b.q, sizeof(a.q)) != 0)
        return nullptr;
      if (a.np > 0 && !is_rotation(a.kind)) fail("compile_batched: non-differentiable parameter");
    }
    check_circuit(*c);
  }
  for (const Term &t : h)
    if (!t.p.valid(s->n)) fail("compile_batched: invalid Pauli term");
  std::vector<Plan> plans(m);
  for (size_t i = 0; i < m; i++) {
    plans[i] = build_plan(*circuits[i], FuseOpts{});
    if (!same_structure(plans[i], plans[0])) return nullptr;
  }
  // Forward passes: the schedule of the first plan; a single op becomes a
  // one-op register-tile pass, chain-shaped passes are not batched.
  std::vector<Pass> passes;
  for (Pass &ps : schedule_passes(plans[0])) {
    if (ps.wide || ps.low) return nullptr;
    if (!ps.merged) {
      const PlanOp &op = plans[0].ops[ps.ops[0]];
      if (op.kind == OpKind::Diagonal && op.masks.empty()) continue;  // no work
      if (!pass_mergeable(op)) return nullptr;
      for (size_t i = 1; i < m; i++)
        if (pass_fixed(plans[i].ops[ps.ops[0]]) != pass_fixed(op)) return nullptr;
      ps.merged = true;
      ps.reg = pass_wires(op);
      ps.fixed = pass_fixed(op);
      if (ps.reg.size() > kPassMaxReg) return nullptr;
    }
    passes.push_back(std::move(ps));
  }
  std::vector<size_t> rot_index;
  const std::vector<Impl::APass> apasses = Impl::adjoint_schedule(c0, rot_index);
  // Per-instance arena region: forward passes, then backward passes.
  std::vector<size_t> foff(passes.size()), boff(apasses.size());
  size_t stride = 0;
  for (size_t k = 0; k < passes.size(); k++) {
    foff[k] = stride;
    stride += s->pass_param_bytes(plans[0], passes[k]);
  }
  for (size_t j = 0; j < apasses.size(); j++) {
    boff[j] = stride;
    stride += Impl::adj_pass_bytes(c0, apasses[j]);
  }
  stride = (stride + 255) & ~size_t(255);
  const uint64_t state_bytes = uint64_t(m) * s->dim * sizeof(float2);
  if (2 * state_bytes + m * stride > free_memory() / 2) return nullptr;

  std::unique_ptr<BatchedAdjoint> out(new BatchedAdjoint);
  BatchedAdjoint::Data &d = *out->d_;
  d.n = s->n;
  d.dim = s->dim;
  d.count = uint32_t(m);
  check(hipMalloc(reinterpret_cast<void **>(&d.arena), m * stride), "hipMalloc(batched arena)");
  std::vector<uint8_t> host(m * stride);
  {
    // Compile through the executor's own pass preparation, pointed at this
    // program's host block and arena.
    struct Swap {
      Impl *s;
      uint8_t *st, *ar;
      ~Swap() {
        s->staging = st;
        s->arena = ar;
        s->batch_compile = false;
      }
    } swap{s, s->staging, s->arena};
    s->staging = host.data();
    s->arena = d.arena;
    s->batch_compile = true;
    std::vector<Impl::PassLaunch> tmpl(passes.size());
    for (size_t i = 0; i < m; i++) {
      for (size_t k = 0; k < passes.size(); k++) {
        Impl::PassLaunch pl{};
        const size_t off = i * stride + foff[k];
        if (passes[k].tile) s->prep_tile(plans[i], passes[k], off, pl);
        else if (passes[k].gwide) s->prep_gwide(plans[i], passes[k], off, pl);
        else s->prep_pass(plans[i], passes[k], off, pl);
        if (pl.chain || pl.wide || pl.low) return nullptr;
        if (i == 0) {
          tmpl[k] = pl;
          continue;
        }
        const Impl::PassLaunch &t = tmpl[k];
        bool same = pl.tile == t.tile && pl.gwide == t.gwide && pl.nl == t.nl &&
                    pl.chain_k == t.chain_k;
        if (same && pl.tile) same = same_layout(pl.t.lay, t.t.lay) && pl.t.nsegs == t.t.nsegs;
        else if (same && pl.gwide)
          same = same_layout(pl.g.lay, t.g.lay) && pl.g.na == t.g.na && pl.g.nops == t.g.nops &&
                 std::memcmp(pl.g.pos, t.g.pos, sizeof(pl.g.pos)) == 0 &&
                 std::memcmp(pl.g.lanepos, t.g.lanepos, sizeof(pl.g.lanepos)) == 0;
        else if (same) same = same_layout(pl.a.lay, t.a.lay) && pl.a.nops == t.a.nops;
        if (!same) return nullptr;
      }
      for (size_t j = 0; j < apasses.size(); j++) {
        BatchedAdjoint::Data::Bwd b;
        b.tile = apasses[j].tile;
        if (b.tile) {
          b.nops = s->emit_adj_tile(*circuits[i], apasses[j], host.data() + i * stride + boff[j],
                                    d.arena + boff[j], b.t);
        } else {
          s->emit_adj_pass(*circuits[i], apasses[j],
                           reinterpret_cast<AdjOp *>(host.data() + i * stride + boff[j]), b.a);
          b.a.ops = reinterpret_cast<const AdjOp *>(d.arena + boff[j]);
          b.nops = b.a.nops;
        }
        if (i == 0) {
          d.bwd.push_back(b);
          d.bwd_rot0.push_back(apasses[j].rot0);
          continue;
        }
        const BatchedAdjoint::Data::Bwd &t = d.bwd[j];
        const bool same = b.tile ? same_layout(b.t.lay, t.t.lay) && b.t.nsegs == t.t.nsegs &&
                                       b.t.nrot == t.t.nrot && b.nops == t.nops &&
                                       std::memcmp(b.t.hpos, t.t.hpos, sizeof(b.t.hpos)) == 0
                                 : same_layout(b.a.lay, t.a.lay) && b.nops == t.nops &&
                                       b.a.nrot == t.a.nrot;
        if (!same) return nullptr;
      }
    }
    for (size_t k = 0; k < passes.size(); k++) {
      BatchedPass b;
      const Impl::PassLaunch &t = tmpl[k];
      b.nl = t.nl;
      if (t.tile) {
        b.kind = 1;
        b.t = t.t;
      } else if (t.gwide) {
        b.kind = 2;
        b.wb = t.chain_k - 4;
        b.g = t.g;
      } else {
        b.a = t.a;
      }
      d.fwd.push_back(b);
    }
  }
  check(hipMemcpy(d.arena, host.data(), host.size(), hipMemcpyHostToDevice), "batched arena");
  check(hipMalloc(reinterpret_cast<void **>(&d.psi), state_bytes), "hipMalloc(batched states)");
  check(hipMalloc(reinterpret_cast<void **>(&d.lam), state_bytes), "hipMalloc(batched co-states)");
  for (BatchedPass &b : d.fwd) {
    b.a.a = b.t.a = b.g.a = d.psi;
    b.a.bstate = b.t.bstate = b.g.bstate = d.dim;
    b.a.barena = b.t.barena = b.g.barena = stride;
  }
  for (BatchedAdjoint::Data::Bwd &b : d.bwd) {
    b.a.psi = b.t.psi = d.psi;
    b.a.lam = b.t.lam = d.lam;
    b.a.bstate = b.t.bstate = d.dim;
    b.a.barena = b.t.barena = stride;
  }
  Impl::CoPlan cp = Impl::costate_tile_plan(h, d.n);
  for (size_t k = 0; k < cp.tiles.size(); k++) {
    BatchedAccum b;
    b.cotile = true;
    b.fuse = k + 1 == cp.tiles.size() && cp.generic.empty();
    b.c = cp.tiles[k];
    b.c.lam = d.lam;
    b.c.psi = d.psi;
    b.c.bstate = d.dim;
    d.costate.push_back(b);
  }
  for (Impl::AccumLaunch &L : cp.tiles.empty() ? Impl::costate_launches(h) : cp.generic) {
    BatchedAccum b;
    b.tile = L.tile;
    b.t = L.t;
    b.g = L.g;
    b.t.lam = b.g.lam = d.lam;
    b.t.psi = b.g.psi = d.psi;
    b.t.dim = b.g.dim = d.dim;
    b.t.bstate = b.g.bstate = d.dim;
    d.costate.push_back(b);
  }
  // Rotations in sweep order: their gate and which part of the overlap is
  // the gradient (Im(i^ny S): the real part for RY, the imaginary part else).
  std::vector<uint32_t> rgate;
  for (size_t k = 0; k < c0.gates.size(); k++)
    if (is_rotation(c0.gates[k].kind)) d.nrot++;
  rgate.assign(d.nrot, 0);
  for (size_t k = 0; k < c0.gates.size(); k++)
    if (is_rotation(c0.gates[k].kind))
      rgate[rot_index[k]] = uint32_t(k) | (c0.gates[k].kind == K_RY ? 0x80000000u : 0u);
  d.rgate = device_vector(rgate, "batched rotation map");
  d.inst = device_vector(inst, "batched instances");
  d.gbase = device_vector(gbase, "batched gradient bases");
  size_t rblocks = pauli_batched_blocks(d.dim);
  for (const BatchedAccum &b : d.costate)
    if (b.cotile) rblocks = std::max<size_t>(rblocks, costate_tile_blocks(b.c.lay));
  check(hipMalloc(reinterpret_cast<void **>(&d.red), size_t(m) * rblocks * 2 * sizeof(double)),
        "hipMalloc(batched partials)");
  size_t aparts = 1;
  for (const BatchedAdjoint::Data::Bwd &b : d.bwd)
    aparts = std::max(aparts, b.tile ? size_t(kAdjTileMaxRot) * adj_tile_batched_blocks(b.t.lay)
                                     : size_t(kAdjMaxRot) * adj_batched_blocks(b.a.lay));
  check(hipMalloc(reinterpret_cast<void **>(&d.adj), size_t(m) * aparts * 2 * sizeof(double)),
        "hipMalloc(batched adjoint partials)");
  check(hipMalloc(reinterpret_cast<void **>(&d.eslots), size_t(m) * 2 * sizeof(double)),
        "hipMalloc(batched energies)");
  check(hipMalloc(reinterpret_cast<void **>(&d.rslots),
                  size_t(m) * std::max<uint32_t>(d.nrot, 1) * 2 * sizeof(double)),
        "hipMalloc(batched overlaps)");
  return out;
}

}  // namespace qmamd
