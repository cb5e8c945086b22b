// kernels.hip -- hand-written gfx950 kernels for in-place complex64
// statevector passes. Wave64 throughout; every index is 64-bit so n = 34
// (2^34 amplitudes, 128 GiB) is addressable; launches cap the grid and loop.
//
// Kernel families
//   k_gather<NL>   2x2 slots (optionally controlled) applied to the 2^NL
//                  amplitudes a thread owns: Dense1q, Group, Controlled1q,
//                  Swap (three CX slots), CCX, and the gate-by-gate path.
//                  Bit 0 always lives inside one 16-byte float4 load, so every
//                  load and store is a full dwordx4; thread-index bits go to
//                  the lowest free positions so neighbouring lanes touch
//                  neighbouring 16-byte words.
//   k_diag<NL>     phase layer with the phase written as
//                  const + sum beta_k x_k + sum gamma_kj x_k x_j over the
//                  thread's local bits after the base index has been folded in.
//   k_blocked      contiguous 2^b tile staged in LDS, gates applied in order.
//   reductions     FP64 accumulation from the first add; per-block partials
//                  reduced by one block in a fixed order (deterministic).
//   sampling       FP64 block sums, FP64 prefix, per-shot resolution on GPU.

#include <hip/hip_runtime.h>

#include <algorithm>

#include "kernels.h"
#include "kernel_util.h"

namespace qmamd {
namespace {

template <int NL, bool NT>
__device__ __forceinline__ void gather_one(float2 *__restrict__ a, uint64_t ub,
                                           uint32_t lane_bytes,
                                           const uint64_t (&off)[1 << (NL - 1)],
                                           const GatherArgs &g) {
  float2 x[1 << NL];
#pragma unroll
  for (int r = 0; r < (1 << (NL - 1)); r++) {
    const float4 v = ld4<NT>(pair_ptr(a, ub | off[r], lane_bytes));
    x[2 * r] = make_float2(v.x, v.y);
    x[2 * r + 1] = make_float2(v.z, v.w);
  }
  for (uint32_t k = 0; k < g.nslots; k++) apply_any<NL>(x, g.slots[k]);
#pragma unroll
  for (int r = 0; r < (1 << (NL - 1)); r++)
    st4<NT>(pair_ptr(a, ub | off[r], lane_bytes),
            make_float4(x[2 * r].x, x[2 * r].y, x[2 * r + 1].x, x[2 * r + 1].y));
}

template <int NL, bool NT>
__global__ __launch_bounds__(kT) void k_gather(GatherArgs g) {
  uint64_t off[1 << (NL - 1)];
  offsets<NL>(g.lay, off);
  const TileAddr ta = tile_addr(g.lay);
  const uint64_t stride = uint64_t(gridDim.x) * kT;
  for (uint64_t t0 = uint64_t(blockIdx.x) * kT; t0 < g.lay.nthreads; t0 += stride) {
    if (t0 + threadIdx.x >= g.lay.nthreads) break;
    const uint64_t ub = scatter(t0, g.lay) | g.lay.fixed;
    gather_one<NL, NT>(g.a, ub, ta.lane_bytes, off, g);
    if (g.b) gather_one<NL, NT>(g.b, ub, ta.lane_bytes, off, g);
  }
}

template <int NL, bool NT>
__global__ __launch_bounds__(kT) void k_diag(DiagArgs d) {
  uint64_t off[1 << (NL - 1)];
  offsets<NL>(d.lay, off);
  const TileAddr ta = tile_addr(d.lay);
  const uint64_t stride = uint64_t(gridDim.x) * kT;
  const uint32_t lane = threadIdx.x & 63;
  for (uint64_t t0 = uint64_t(blockIdx.x) * kT; t0 < d.lay.nthreads; t0 += stride) {
    // Workgroup-uniform terms (they read only base-index bits that are the
    // same for every thread of this iteration): the lanes of each wave take
    // one term each and the wave sums the hits.
    const uint64_t ub = scatter(t0, d.lay) | d.lay.fixed;
    double phi = 0.0;
    double beta[4] = {0.0, 0.0, 0.0, 0.0};
    for (uint32_t k = lane; k < d.nuniform; k += 64) {
      const DiagTerm tm = d.terms[k];
      if ((ub & tm.mask) != tm.mask) continue;
      if (tm.kind == 2) {
#pragma unroll
        for (int j = 0; j < 4; j++)
          if (tm.k == uint32_t(j)) beta[j] += tm.ang;
      } else {
        phi += tm.ang;
      }
    }
    if (d.nuniform) {
      phi = wave_allsum(phi);
#pragma unroll
      for (int j = 0; j < NL; j++) beta[j] = wave_allsum(beta[j]);
    }
    phi += d.constant;
#pragma unroll
    for (int j = 0; j < 4; j++) beta[j] += d.beta0[j];
    if (t0 + threadIdx.x >= d.lay.nthreads) continue;
    const uint64_t base = ub | ta.lane_bits;
    for (uint32_t k = d.nuniform; k < d.nterms; k++) {
      const DiagTerm tm = d.terms[k];
      const bool hit = (base & tm.mask) == tm.mask;
      if (tm.kind == 2) {
#pragma unroll
        for (int j = 0; j < 4; j++)
          if (hit && tm.k == uint32_t(j)) beta[j] += tm.ang;
      } else if (hit) {
        phi += tm.ang;
      }
    }
    double ph[1 << NL];
    ph[0] = phi;
#pragma unroll
    for (int h = 0; h < NL; h++) {
#pragma unroll
      for (int l = 0; l < (1 << h); l++) {
        double inc = beta[h];
#pragma unroll
        for (int k = 0; k < h; k++)
          if ((l >> k) & 1) inc += d.gamma[k][h];
        ph[l | (1 << h)] = ph[l] + inc;
      }
    }
    // All loads are issued before any store so every 16-byte load of the
    // thread is in flight at once.
    float4 v[1 << (NL - 1)];
#pragma unroll
    for (int r = 0; r < (1 << (NL - 1)); r++) v[r] = ld4<NT>(pair_ptr(d.a, ub | off[r], ta.lane_bytes));
#pragma unroll
    for (int r = 0; r < (1 << (NL - 1)); r++) {
      float s0, c0, s1, c1;
      turn_sincos(ph[2 * r], s0, c0);
      turn_sincos(ph[2 * r + 1], s1, c1);
      const float2 a0 = cmul(make_float2(v[r].x, v[r].y), make_float2(c0, s0));
      const float2 a1 = cmul(make_float2(v[r].z, v[r].w), make_float2(c1, s1));
      v[r] = make_float4(a0.x, a0.y, a1.x, a1.y);
    }
#pragma unroll
    for (int r = 0; r < (1 << (NL - 1)); r++) st4<NT>(pair_ptr(d.a, ub | off[r], ta.lane_bytes), v[r]);
  }
}

__global__ __launch_bounds__(kT) void k_diag_generic(DiagGenericArgs d) {
  const uint64_t stride = uint64_t(gridDim.x) * kT;
  for (uint64_t i = uint64_t(blockIdx.x) * kT + threadIdx.x; i < d.dim;
       i += stride) {
    double ph = d.global;
    for (uint32_t k = 0; k < d.nterms; k++)
      if ((i & d.masks[k]) == d.masks[k]) ph += d.angles[k];
    float s, c;
    turn_sincos(ph, s, c);
    d.a[i] = cmul(d.a[i], make_float2(c, s));
  }
}

// Ops [o0, o1) of a merged pass applied to the thread's 2^NL amplitudes:
// `ub` is the workgroup's base index (the bits every thread shares), `base`
// the thread's (ub plus its thread-index bits).
template <int NL>
__device__ __forceinline__ void pass_ops(float2 (&x)[1 << NL], const_ptr<PassOp> ops,
                                         const_ptr<DiagTerm> terms, uint32_t o0, uint32_t o1,
                                         uint64_t ub, uint64_t base, uint32_t lane) {
  // Every op runs both statements that write the tile, the 2x2 slot and
  // the per-combination phase multiply, each of which skips itself inside
  // (a branch around either would be a second definition of the tile).
  // Each op's header is loaded as one block into scalar registers (otherwise
  // each field is loaded, and waited for, in its own branch), one op ahead of
  // its use so that the load overlaps the current op's arithmetic.
  struct Hdr {
    uint32_t type, cm, sel, pm;
    uint64_t gmask;
    float m[8];
  };
  auto load_hdr = [&](uint32_t o, Hdr &h) {
    h.type = ops[o].type;
    h.cm = ops[o].cmask;
    h.sel = ops[o].sel;
    h.pm = ops[o].pm;
    h.gmask = ops[o].gmask;
#pragma unroll
    for (int i = 0; i < 8; i++) h.m[i] = ops[o].m[i];
  };
  auto pin = [](Hdr &h) {
    asm volatile("" : "+s"(h.type), "+s"(h.cm), "+s"(h.sel), "+s"(h.pm), "+s"(h.gmask),
                 "+s"(h.m[0]), "+s"(h.m[1]), "+s"(h.m[2]), "+s"(h.m[3]), "+s"(h.m[4]),
                 "+s"(h.m[5]), "+s"(h.m[6]), "+s"(h.m[7]));
  };
  Hdr next{};
  if (o0 < o1) {
    load_hdr(o0, next);
    pin(next);
  }
  for (uint32_t o = o0; o < o1; o++) {
    Hdr h = next;
    load_hdr(o + 1 < o1 ? o + 1 : o, next);
    const uint32_t type = h.type, cm = h.cm, sel = h.sel, pm = h.pm;
    const uint64_t gmask = h.gmask;
    float m[8];
#pragma unroll
    for (int i = 0; i < 8; i++) m[i] = h.m[i];
    const bool is_phase = type == kPassPhase;
    const bool gon = (base & gmask) == gmask;
    {
      // A control on a base-index bit: the identity where this lane's
      // base lacks it. A phase op's section code matches nothing.
#pragma unroll
      for (int i = 0; i < 8; i++) m[i] = gon ? m[i] : (i == 0 || i == 6 ? 1.0f : 0.0f);
      slot_asm<NL>(x, sel, pm, m);
    }
    // The phase multiply, for a phase op only (a uniform branch around
    // in-place statements: the tile keeps its registers on both paths).
    if (is_phase) {
      const uint32_t bmask = ops[o].bmask;
      const bool has_phi = ops[o].has_phi != 0;
      float2 e[NL];
#pragma unroll
      for (int j = 0; j < NL; j++) e[j] = make_float2(1.0f, 0.0f);
      float2 ephi = make_float2(1.0f, 0.0f);
      // Workgroup-uniform terms: one term per lane, summed over the wave
      // (every lane takes part, active or not).
      const uint32_t nuniform = ops[o].nuniform, nterms = ops[o].nterms;
      const uint32_t umask = ops[o].umask;
      const const_ptr<DiagTerm> tl = terms + ops[o].term0;
      double phi = 0.0;
      double beta[NL];
#pragma unroll
      for (int j = 0; j < NL; j++) beta[j] = 0.0;
      for (uint32_t k0 = 0; k0 < nuniform; k0 += 64) {
        // Branch-free: lanes past the end read the last term and add zero.
        const uint32_t k = k0 + lane;
        const uint32_t kc = k < nuniform ? k : nuniform - 1;
        const uint64_t mask = tl[kc].mask;
        const double add = (k < nuniform && (ub & mask) == mask) ? double(tl[kc].ang) : 0.0;
        const uint32_t kind = tl[kc].kind, kk = tl[kc].k;
        phi += kind != 2 ? add : 0.0;
#pragma unroll
        for (int j = 0; j < NL; j++) beta[j] += (kind == 2 && kk == uint32_t(j)) ? add : 0.0;
      }
      if (umask & 1u) phi = wave_allsum(phi);
#pragma unroll
      for (int j = 0; j < NL; j++)
        if ((umask >> (j + 1)) & 1u) beta[j] = wave_allsum(beta[j]);
      for (uint32_t k = nuniform; k < nterms; k++) {
        const uint64_t mask = tl[k].mask;
        const double add = (base & mask) == mask ? double(tl[k].ang) : 0.0;
        const uint32_t kind = tl[k].kind, kk = tl[k].k;
        if (kind == 2) {
#pragma unroll
          for (int j = 0; j < NL; j++) beta[j] += kk == uint32_t(j) ? add : 0.0;
        } else {
          phi += add;
        }
      }
      // One FP64-reduced sin/cos per thread-dependent sum; each amplitude
      // then takes a table entry times the factors of its local bits.
#pragma unroll
      for (int j = 0; j < NL; j++) {
        if ((bmask >> j) & 1u) {
          float sn, cs;
          turn_sincos(beta[j] + ops[o].beta0[j], sn, cs);
          e[j] = make_float2(cs, sn);
        }
      }
      if (has_phi) {
        float sn, cs;
        turn_sincos(phi, sn, cs);
        ephi = make_float2(cs, sn);
      }
#pragma unroll
      for (int l = 0; l < (1 << NL); l++) {
        float2 f = make_float2(ops[o].table[2 * l], ops[o].table[2 * l + 1]);
        if (has_phi) f = cmul(f, ephi);
#pragma unroll
        for (int h = 0; h < NL; h++)
          if (((l >> h) & 1) && ((bmask >> h) & 1u)) f = cmul(f, e[h]);
        // A thread whose base lacks a bit every term needs keeps phase 0.
        if (gmask) f = gon ? f : make_float2(1.0f, 0.0f);
        // Combinations without the local bits every term needs are left
        // alone.
        cmul_inplace(x[l], f, __builtin_amdgcn_readfirstlane((uint32_t(l) & cm) != cm));
      }
    }
    pin(next);
  }
}

// Merged pass: the thread's 2^NL amplitudes are loaded once, every op of the
// pass is applied to them in plan order, and they are stored once. Ops are
// workgroup-uniform; their descriptors are read with scalar loads.
template <int NL, bool NT>
__global__ __launch_bounds__(kT) void k_pass(PassArgs p) {
  uint64_t off[1 << (NL - 1)];
  offsets<NL>(p.lay, off);
  const uint64_t inst = blockIdx.y;
  float2 *const A = p.a + inst * p.bstate;
  const const_ptr<PassOp> ops = as_const(bshift(p.ops, inst * p.barena));
  const const_ptr<DiagTerm> terms = as_const(bshift(p.terms, inst * p.barena));
  const TileAddr ta = tile_addr(p.lay);
  const uint64_t stride = uint64_t(gridDim.x) * kT;
  const uint32_t lane = threadIdx.x & 63;
  for (uint64_t t0 = uint64_t(blockIdx.x) * kT; t0 < p.lay.nthreads; t0 += stride) {
    const uint64_t ub = scatter(t0, p.lay) | p.lay.fixed;
    const bool active = t0 + threadIdx.x < p.lay.nthreads;
    const uint64_t base = ub | ta.lane_bits;
    // No lane-divergent control flow between the load and the store (it
    // would make the compiler keep a second copy of the amplitudes): lanes
    // past the end compute on a valid tile they never store, and a control on
    // a base-index bit picks the identity per lane instead of branching.
    float2 x[1 << NL];
    {
      // Lanes past the end read the workgroup's first tile (in bounds),
      // compute on it and drop it.
      const uint32_t lb = active ? ta.lane_bytes : 0u;
      float4 v[1 << (NL - 1)];
#pragma unroll
      for (int r = 0; r < (1 << (NL - 1)); r++) v[r] = ld4<NT>(pair_ptr(A, ub | off[r], lb));
#pragma unroll
      for (int r = 0; r < (1 << (NL - 1)); r++) {
        x[2 * r] = make_float2(v[r].x, v[r].y);
        x[2 * r + 1] = make_float2(v[r].z, v[r].w);
      }
    }
    pass_ops<NL>(x, ops, terms, 0, p.nops, ub, base, lane);
    if (active) {
#pragma unroll
      for (int r = 0; r < (1 << (NL - 1)); r++)
        st4<NT>(pair_ptr(A, ub | off[r], ta.lane_bytes),
                make_float4(x[2 * r].x, x[2 * r].y, x[2 * r + 1].x, x[2 * r + 1].y));
    }
  }
}

// Chain-shaped merged pass (ChainStep): every step's local bit, and which
// combinations its phase touches, are compile-time, so the step sequence is
// straight-line code. The loads of the workgroup-uniform terms are issued
// ahead of the amplitude loads so their latency overlaps.
template <int NL, bool NT, int K, bool Z, bool REAL>
__global__ __launch_bounds__(kT) void k_chain(ChainArgs p) {
  constexpr int S = K + (Z ? 1 : 0);
  static_assert(S <= NL && K < NL, "chain longer than the tile");
  uint64_t off[1 << (NL - 1)];
  offsets<NL>(p.lay, off);
  const const_ptr<ChainStep> steps = as_const(p.steps);
  const const_ptr<DiagTerm> uni = as_const(p.uni);
  const const_ptr<DiagTerm> lane_terms = as_const(p.lane);
  const TileAddr ta = tile_addr(p.lay);
  const uint64_t stride = uint64_t(gridDim.x) * kT;
  const uint32_t lane = threadIdx.x & 63;
  // Lane terms read only thread-index bits, the same for every tile.
  double bl[S];
#pragma unroll
  for (int i = 0; i < S; i++) bl[i] = steps[i].beta0;
  for (uint32_t k = 0; k < p.nlane; k++) {
    const uint64_t mask = lane_terms[k].mask;
    const double add = (ta.lane_bits & mask) == mask ? double(lane_terms[k].ang) : 0.0;
    const uint32_t st = lane_terms[k].k;
#pragma unroll
    for (int i = 0; i < S; i++)
      if (st == uint32_t(i)) bl[i] += add;
  }
  for (uint64_t t0 = uint64_t(blockIdx.x) * kT; t0 < p.lay.nthreads; t0 += stride) {
    const uint64_t ub = scatter(t0, p.lay) | p.lay.fixed;
    const bool active = t0 + threadIdx.x < p.lay.nthreads;
    // Workgroup-uniform terms, one per lane (branch-free: lanes past the end
    // read the last term and add zero).
    double bu[S];
#pragma unroll
    for (int i = 0; i < S; i++) bu[i] = 0.0;
    for (uint32_t k0 = 0; k0 < p.nuni; k0 += 64) {
      const uint32_t k = k0 + lane;
      const uint32_t kc = k < p.nuni ? k : p.nuni - 1;
      const uint64_t mask = uni[kc].mask;
      const double add = (k < p.nuni && (ub & mask) == mask) ? double(uni[kc].ang) : 0.0;
      const uint32_t st = uni[kc].k;
#pragma unroll
      for (int i = 0; i < S; i++) bu[i] += st == uint32_t(i) ? add : 0.0;
    }
    float2 x[1 << NL];
    {
      // Lanes past the end read the workgroup's first tile and never store.
      const uint32_t lb = active ? ta.lane_bytes : 0u;
#pragma unroll
      for (int r = 0; r < (1 << (NL - 1)); r++) {
        const float4 v = ld4<NT>(pair_ptr(p.a, ub | off[r], lb));
        x[2 * r] = make_float2(v.x, v.y);
        x[2 * r + 1] = make_float2(v.z, v.w);
      }
    }
    // Per-thread factor of each step: exp(2 pi i beta).
    float2 e[S];
#pragma unroll
    for (int i = 0; i < S; i++) {
      float sn, cs;
      turn_sincos(bl[i] + wave_allsum(bu[i]), sn, cs);
      e[i] = make_float2(cs, sn);
    }
#pragma unroll
    for (int i = 0; i < S; i++) {
      const ChainStep __attribute__((address_space(4))) &st = steps[opaque_step(i, x[0].x)];
      if (i < K) {
        if constexpr (K >= 1) if (i == 0) chain_step<NL, 1, REAL>(x, st, e[i]);
        if constexpr (K >= 2) if (i == 1) chain_step<NL, 2, REAL>(x, st, e[i]);
        if constexpr (K >= 3) if (i == 2) chain_step<NL, 3, REAL>(x, st, e[i]);
        if constexpr (K >= 4) if (i == 3) chain_step<NL, 4, REAL>(x, st, e[i]);
      } else {
        chain_step<NL, 0, REAL>(x, st, e[i]);
      }
    }
    if (active) {
#pragma unroll
      for (int r = 0; r < (1 << (NL - 1)); r++)
        st4<NT>(pair_ptr(p.a, ub | off[r], ta.lane_bytes),
                make_float4(x[2 * r].x, x[2 * r].y, x[2 * r + 1].x, x[2 * r + 1].y));
    }
  }
}

// Wide chain pass (WideArgs) with WB wave bits. One workgroup tile per
// iteration; the exchange between the two arrangements goes through LDS in
// 2^(4-WB) rounds, one per value of the local bits that stay (16 KB per round
// for WB = 2, 64 KB for WB = 3).
template <int WB, int MODE>
__global__ __launch_bounds__(64 << WB) void k_wide(WideArgs p) {
  constexpr int ML = MODE == 2 ? 1 : MODE;  // the last step's mode
  constexpr int W = 1 << WB;         // waves, and registers per exchange round
  constexpr int ROUNDS = 16 / W;
  __shared__ float4 xch[W][W][64];  // [writer's wave][slot][lane]
  // The wave index is wave-uniform: made scalar so that tile addresses are a
  // scalar base plus the lane's byte offset.
  const uint32_t lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const const_ptr<ChainStep> steps = as_const(p.steps);
  const const_ptr<DiagTerm> uni = as_const(p.uni);
  const const_ptr<DiagTerm> lane_terms = as_const(p.lane);
  uint64_t lanebits = 0;
#pragma unroll
  for (int j = 0; j < 6; j++)
    if ((lane >> j) & 1) lanebits |= 1ull << p.lanepos[j];
  // Wave bits: w5.. before the exchange, w1..wWB after.
  uint64_t wave_a = 0, wave_b = 0;
#pragma unroll
  for (int j = 0; j < WB; j++)
    if ((wave >> j) & 1) {
      wave_a |= 1ull << p.pos[4 + j];
      wave_b |= 1ull << p.pos[j];
    }
  const uint32_t lane_bytes = uint32_t(lanebits) * 8u;
  // Lane terms read only lane and wave bits, the same for every tile.
  double bl[4 + WB];
#pragma unroll
  for (int i = 0; i < 4 + WB; i++) bl[i] = steps[i].beta0;
  for (uint32_t k = 0; k < p.nlane; k++) {
    const uint64_t mask = lane_terms[k].mask;
    const uint32_t st = lane_terms[k].k;
    const uint64_t bs = lanebits | (st < 4 ? wave_a : wave_b);
    const double add = (bs & mask) == mask ? double(lane_terms[k].ang) : 0.0;
#pragma unroll
    for (int i = 0; i < 4 + WB; i++)
      if (st == uint32_t(i)) bl[i] += add;
  }
  for (uint64_t tile = blockIdx.x; tile < p.lay.nthreads; tile += gridDim.x) {
    const uint64_t ub = scatter(tile, p.lay) | p.lay.fixed;
    double bu[4 + WB];
#pragma unroll
    for (int i = 0; i < 4 + WB; i++) bu[i] = 0.0;
    for (uint32_t k0 = 0; k0 < p.nuni; k0 += 64) {
      const uint32_t k = k0 + lane;
      const uint32_t kc = k < p.nuni ? k : p.nuni - 1;
      const uint64_t mask = uni[kc].mask;
      const double add = (k < p.nuni && (ub & mask) == mask) ? double(uni[kc].ang) : 0.0;
      const uint32_t st = uni[kc].k;
#pragma unroll
      for (int i = 0; i < 4 + WB; i++) bu[i] += st == uint32_t(i) ? add : 0.0;
    }
    float2 x[32];
#pragma unroll
    for (int r = 0; r < 16; r++) {
      uint64_t o = ub | wave_a;
#pragma unroll
      for (int k = 0; k < 4; k++)
        if ((r >> k) & 1) o |= 1ull << p.pos[k];
      const float4 v = ld4<true>(pair_ptr(p.a, o, lane_bytes));
      x[2 * r] = make_float2(v.x, v.y);
      x[2 * r + 1] = make_float2(v.z, v.w);
    }
    float2 e[4 + WB];
#pragma unroll
    for (int i = 0; i < 4 + WB; i++) {
      float sn, cs;
      turn_sincos(bl[i] + wave_allsum(bu[i]), sn, cs);
      e[i] = make_float2(cs, sn);
    }
    chain_step<5, 1, MODE>(x, steps[opaque_step(0, x[0].x)], e[0]);
    chain_step<5, 2, MODE>(x, steps[opaque_step(1, x[0].x)], e[1]);
    chain_step<5, 3, MODE>(x, steps[opaque_step(2, x[0].x)], e[2]);
    chain_step<5, 4, MODE>(x, steps[opaque_step(3, x[0].x)], e[3]);
    // Exchange: local bits 1..WB (w1..) <-> wave bits (w5..). Register r
    // (local bits r + 1) holds x[2r], x[2r + 1]; its low WB bits are swapped
    // with the wave index, its high bits pick the round.
#pragma unroll
    for (int c = 0; c < ROUNDS; c++) {
#pragma unroll
      for (int a = 0; a < W; a++) {
        const int l = 2 * ((c << WB) | a);
        xch[wave][a][lane] = make_float4(x[l].x, x[l].y, x[l + 1].x, x[l + 1].y);
      }
      __syncthreads();
#pragma unroll
      for (int d = 0; d < W; d++) {
        const int l = 2 * ((c << WB) | d);
        const float4 v = xch[d][wave][lane];
        x[l] = make_float2(v.x, v.y);
        x[l + 1] = make_float2(v.z, v.w);
      }
      __syncthreads();
    }
    chain_step<5, 1, MODE>(x, steps[opaque_step(4, x[0].x)], e[4]);
    if constexpr (WB >= 3) {
      chain_step<5, 2, MODE>(x, steps[opaque_step(5, x[0].x)], e[5]);
      chain_step<5, 3, ML>(x, steps[opaque_step(6, x[0].x)], e[6]);
    } else {
      chain_step<5, 2, ML>(x, steps[opaque_step(5, x[0].x)], e[5]);
    }
#pragma unroll
    for (int r = 0; r < 16; r++) {
      // Local bits 1..WB now w5..; the rest still w(WB+1)..w4.
      uint64_t o = ub | wave_b;
#pragma unroll
      for (int k = 0; k < 4; k++)
        if ((r >> k) & 1) o |= 1ull << (k < WB ? p.pos[4 + k] : p.pos[k]);
      st4<true>(pair_ptr(p.a, o, lane_bytes),
                make_float4(x[2 * r].x, x[2 * r].y, x[2 * r + 1].x, x[2 * r + 1].y));
    }
  }
}

// Generic wide pass (GWideArgs): the wide chain kernel's tile, load, LDS
// exchange and store around two runs of merged-pass ops.
template <int WB>
__global__ __launch_bounds__(64 << WB) void k_gwide(GWideArgs p) {
  constexpr int W = 1 << WB;
  constexpr int ROUNDS = 16 / W;
  __shared__ float4 xch[W][W][64];  // [writer's wave][slot][lane]
  const uint32_t lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const uint64_t inst = blockIdx.y;
  float2 *const A = p.a + inst * p.bstate;
  const const_ptr<PassOp> ops = as_const(bshift(p.ops, inst * p.barena));
  const const_ptr<DiagTerm> terms = as_const(bshift(p.terms, inst * p.barena));
  uint64_t lanebits = 0;
#pragma unroll
  for (int j = 0; j < 6; j++)
    if ((lane >> j) & 1) lanebits |= 1ull << p.lanepos[j];
  uint64_t wave_a = 0, wave_b = 0;
#pragma unroll
  for (int j = 0; j < WB; j++)
    if ((wave >> j) & 1) {
      wave_a |= 1ull << p.pos[4 + j];
      wave_b |= 1ull << p.pos[j];
    }
  const uint32_t lane_bytes = uint32_t(lanebits) * 8u;
  for (uint64_t tile = blockIdx.x; tile < p.lay.nthreads; tile += gridDim.x) {
    const uint64_t ub = scatter(tile, p.lay) | p.lay.fixed;
    float2 x[32];
#pragma unroll
    for (int r = 0; r < 16; r++) {
      uint64_t o = ub | wave_a;
#pragma unroll
      for (int k = 0; k < 4; k++)
        if ((r >> k) & 1) o |= 1ull << p.pos[k];
      const float4 v = ld4<true>(pair_ptr(A, o, lane_bytes));
      x[2 * r] = make_float2(v.x, v.y);
      x[2 * r + 1] = make_float2(v.z, v.w);
    }
    pass_ops<5>(x, ops, terms, 0, p.na, ub, ub | lanebits | wave_a, lane);
#pragma unroll
    for (int c = 0; c < ROUNDS; c++) {
#pragma unroll
      for (int a = 0; a < W; a++) {
        const int l = 2 * ((c << WB) | a);
        xch[wave][a][lane] = make_float4(x[l].x, x[l].y, x[l + 1].x, x[l + 1].y);
      }
      __syncthreads();
#pragma unroll
      for (int d = 0; d < W; d++) {
        const int l = 2 * ((c << WB) | d);
        const float4 v = xch[d][wave][lane];
        x[l] = make_float2(v.x, v.y);
        x[l + 1] = make_float2(v.z, v.w);
      }
      __syncthreads();
    }
    pass_ops<5>(x, ops, terms, p.na, p.nops, ub, ub | lanebits | wave_b, lane);
#pragma unroll
    for (int r = 0; r < 16; r++) {
      uint64_t o = ub | wave_b;
#pragma unroll
      for (int k = 0; k < 4; k++)
        if ((r >> k) & 1) o |= 1ull << (k < WB ? p.pos[4 + k] : p.pos[k]);
      st4<true>(pair_ptr(A, o, lane_bytes),
                make_float4(x[2 * r].x, x[2 * r].y, x[2 * r + 1].x, x[2 * r + 1].y));
    }
  }
}

// Moves the tile between two arrangements given by their LDS slot bases:
// `tf` / `tt` the thread's own slot part (from / to), `rf` / `rt` the
// register bits' bases (workgroup-uniform).
__device__ __forceinline__ void tile_permute(float4 *lds, float2 (&x)[32], uint32_t tf,
                                             const uint32_t (&rf)[4], uint32_t tt,
                                             const uint32_t (&rt)[4]) {
  char *base = reinterpret_cast<char *>(lds);
  __syncthreads();  // earlier reads of the buffer are done
#pragma unroll
  for (int r = 0; r < 16; r++) {
    const uint32_t o = ((r & 1) ? rf[0] : 0u) ^ ((r & 2) ? rf[1] : 0u) ^ ((r & 4) ? rf[2] : 0u) ^
                       ((r & 8) ? rf[3] : 0u);
    *reinterpret_cast<float4 *>(base + ((tf ^ o) << 4)) =
        make_float4(x[2 * r].x, x[2 * r].y, x[2 * r + 1].x, x[2 * r + 1].y);
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < 16; r++) {
    const uint32_t o = ((r & 1) ? rt[0] : 0u) ^ ((r & 2) ? rt[1] : 0u) ^ ((r & 4) ? rt[2] : 0u) ^
                       ((r & 8) ? rt[3] : 0u);
    const float4 v = *reinterpret_cast<const float4 *>(base + ((tt ^ o) << 4));
    x[2 * r] = make_float2(v.x, v.y);
    x[2 * r + 1] = make_float2(v.z, v.w);
  }
}

// Generic low tile pass (TileArgs): two waves per 2^12-amplitude tile.
__global__ __launch_bounds__(128) void k_tile(TileArgs p) {
  __shared__ float4 lds[2048];
  const uint32_t lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const uint64_t inst = blockIdx.y;
  float2 *const A = p.a + inst * p.bstate;
  const const_ptr<TileSeg> segs = as_const(bshift(p.segs, inst * p.barena));
  const const_ptr<PassOp> ops = as_const(bshift(p.ops, inst * p.barena));
  const const_ptr<DiagTerm> terms = as_const(bshift(p.terms, inst * p.barena));
  const uint32_t lane_bytes = lane * 16u;
  const uint64_t wave_off = uint64_t(wave) << 7;
  // Arrangement 0 (loads and stores): thread bit j on tile bit j + 1,
  // register bit k on tile bit 8 + k.
  uint32_t t0 = 0;
#pragma unroll
  for (int j = 0; j < 6; j++)
    if ((lane >> j) & 1) t0 ^= tile_swizzle(1u << j);
  if (wave) t0 ^= tile_swizzle(1u << 6);
  uint32_t r0[4];
#pragma unroll
  for (int k = 0; k < 4; k++) r0[k] = tile_swizzle(1u << (7 + k));
  for (uint64_t tile = blockIdx.x; tile < p.lay.nthreads; tile += gridDim.x) {
    const uint64_t ub = scatter(tile, p.lay) | p.lay.fixed;
    float2 x[32];
#pragma unroll
    for (int r = 0; r < 16; r++) {
      const float4 v = ld4<true>(pair_ptr(A, ub | wave_off | (uint64_t(r) << 8), lane_bytes));
      x[2 * r] = make_float2(v.x, v.y);
      x[2 * r + 1] = make_float2(v.z, v.w);
    }
    uint32_t tf = t0, rf[4] = {r0[0], r0[1], r0[2], r0[3]};
    for (uint32_t s = 0; s < p.nsegs; s++) {
      const TileSeg __attribute__((address_space(4))) &sg = segs[s];
      uint32_t tt = 0;
      uint64_t base = ub;
#pragma unroll
      for (int j = 0; j < 7; j++) {
        const bool on = j < 6 ? ((lane >> j) & 1) != 0 : wave != 0;
        if (on) {
          tt ^= sg.tslot[j];
          base |= 1ull << sg.tpos[j];
        }
      }
      uint32_t rt[4];
#pragma unroll
      for (int k = 0; k < 4; k++) rt[k] = sg.rslot[k];
      tile_permute(lds, x, tf, rf, tt, rt);
      pass_ops<5>(x, ops, terms, sg.op0, sg.op1, ub, base, lane);
      tf = tt;
#pragma unroll
      for (int k = 0; k < 4; k++) rf[k] = rt[k];
    }
    tile_permute(lds, x, tf, rf, t0, r0);
#pragma unroll
    for (int r = 0; r < 16; r++)
      st4<true>(pair_ptr(A, ub | wave_off | (uint64_t(r) << 8), lane_bytes),
                make_float4(x[2 * r].x, x[2 * r].y, x[2 * r + 1].x, x[2 * r + 1].y));
  }
}

__global__ __launch_bounds__(kT) void k_blocked(BlockedArgs b) {
  extern __shared__ float2 tile[];
  const uint32_t block = 1u << b.bits, pairs = block >> 1;
  for (uint64_t blk = blockIdx.x; blk < b.nblocks; blk += gridDim.x) {
    float2 *g = b.a + blk * block;
    for (uint32_t i = threadIdx.x; i < block; i += kT) tile[i] = g[i];
    __syncthreads();
    for (uint32_t k = 0; k < b.count; k++) {
      const Slot s = b.gates[k];
      const uint32_t q = s.lbit, stride = 1u << q;
      for (uint32_t t = threadIdx.x; t < pairs; t += kT) {
        const uint32_t i = ((t >> q) << (q + 1)) | (t & (stride - 1));
        rot(s.m, tile[i], tile[i | stride]);
      }
      __syncthreads();
    }
    for (uint32_t i = threadIdx.x; i < block; i += kT) g[i] = tile[i];
    __syncthreads();
  }
}

__global__ __launch_bounds__(kT) void k_fill_basis(float2 *a, uint64_t dim) {
  const uint64_t stride = uint64_t(gridDim.x) * kT;
  float4 *v = reinterpret_cast<float4 *>(a + uint64_t(blockIdx.y) * dim);  // instance blockIdx.y
  for (uint64_t i = uint64_t(blockIdx.x) * kT + threadIdx.x; i < dim / 2; i += stride)
    v[i] = make_float4(i == 0 ? 1.0f : 0.0f, 0.0f, 0.0f, 0.0f);
}

__global__ __launch_bounds__(kT) void k_zero(float2 *a, uint64_t dim) {
  const uint64_t stride = uint64_t(gridDim.x) * kT;
  float4 *v = reinterpret_cast<float4 *>(a + uint64_t(blockIdx.y) * dim);  // instance blockIdx.y
  for (uint64_t i = uint64_t(blockIdx.x) * kT + threadIdx.x; i < dim / 2; i += stride)
    v[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

// Placement probe: every amplitude read and written back times `one` (1.0,
// unknown to the compiler, so the stores stay), over tiles shaped like a
// wide chain pass: lanes on amplitude bits 1..6, 16 registers on bits
// p0..p0+3, 8 waves on p0+4..p0+6 (`lay` enumerates the tiles).
__global__ __launch_bounds__(512) void k_placement_probe(float4 *a, Layout lay, uint32_t p0,
                                                         float one) {
  const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint64_t woff = uint64_t(wave) << (p0 + 4);
  for (uint64_t tile = blockIdx.x; tile < lay.nthreads; tile += gridDim.x) {
    const uint64_t ub = scatter(tile, lay) | woff;
    float4 v[16];
#pragma unroll
    for (int r = 0; r < 16; r++) v[r] = a[((ub | (uint64_t(r) << p0)) >> 1) + lane];
#pragma unroll
    for (int r = 0; r < 16; r++) {
      const float4 x = v[r];
      a[((ub | (uint64_t(r) << p0)) >> 1) + lane] =
          make_float4(x.x * one, x.y * one, x.z * one, x.w * one);
    }
  }
}

// out[k] = a[idx[k]]; every index was checked against the state size.
__global__ __launch_bounds__(kT) void k_pick(const float2 *a, const uint64_t *idx, uint64_t count,
                                             float2 *out) {
  const uint64_t stride = uint64_t(gridDim.x) * kT;
  for (uint64_t k = uint64_t(blockIdx.x) * kT + threadIdx.x; k < count; k += stride)
    out[k] = a[idx[k]];
}

// --- FP64 reductions --------------------------------------------------------

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  return v;
}

// Sum over the workgroup; thread 0 gets the total. Fixed order.
__device__ __forceinline__ double block_sum(double v, double *sh) {
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) sh[w] = v;
  __syncthreads();
  double t = 0.0;
  if (threadIdx.x == 0)
    for (int k = 0; k < kT / 64; k++) t += sh[k];
  __syncthreads();
  return t;
}

// The streaming reductions keep kRedUnroll independent 16-byte nontemporal
// loads in flight per thread (a wave's k-th loads cover 1 KB, the block's
// kT * kRedUnroll float4 one contiguous run per loop step).
constexpr int kRedUnroll = 4;

__global__ __launch_bounds__(kT) void k_abs2(const float2 *a, uint64_t dim, double *part) {
  __shared__ double sh[kT / 64];
  constexpr int U = kRedUnroll;
  const uint64_t n4 = dim / 2, step = uint64_t(gridDim.x) * kT * U;
  double acc = 0.0;
  for (uint64_t c = uint64_t(blockIdx.x) * kT * U + threadIdx.x; c < n4; c += step) {
    float4 x[U];
#pragma unroll
    for (int k = 0; k < U; k++) {
      const uint64_t i = c + uint64_t(k) * kT;
      x[k] = i < n4 ? ld4<true>(a + 2 * i) : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
#pragma unroll
    for (int k = 0; k < U; k++)
      acc += double(x[k].x) * x[k].x + double(x[k].y) * x[k].y + double(x[k].z) * x[k].z +
             double(x[k].w) * x[k].w;
  }
  const double t = block_sum(acc, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = t;
}

__device__ __forceinline__ double sgn(uint64_t j, uint64_t sign) {
  return (__popcll(j & sign) & 1) ? -1.0 : 1.0;
}

// acc += s(j) conj(x) y
__device__ __forceinline__ void acc_cj(double &re, double &im, float2 x, float2 y, double s) {
  re += s * (double(x.x) * y.x + double(x.y) * y.y);
  im += s * (double(x.x) * y.y - double(x.y) * y.x);
}

// sum_i s(i ^ flip) conj(bra_i) ket_{i ^ flip}. Each amplitude is read once:
// a thread takes an even index i with the flip's top bit clear together with
// its partners, and adds both directions of the pairing.
// This workgroup's part of sum_i s(i ^ flip) conj(bra_i) ket_{i ^ flip}
// (workgroup bx of gx). With bra == ket (an expectation) each float4 is
// loaded once. Two float4 of each state per thread and step: kRedBlocks
// workgroups stay resident at once (more registers would leave a tail of
// workgroups).
__device__ __forceinline__ void pauli_part(const float2 *bra, const float2 *ket, uint64_t dim,
                                           uint64_t flip, uint64_t sign, uint32_t bx,
                                           uint32_t gx, double &re, double &im) {
  const bool same = bra == ket;
  constexpr int U = 2;
  const float4 zero = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  if (flip <= 1) {
    const uint64_t n4 = dim / 2, step = uint64_t(gx) * kT * U;
    for (uint64_t c = uint64_t(bx) * kT * U + threadIdx.x; c < n4; c += step) {
      float4 b[U], k[U];
#pragma unroll
      for (int q = 0; q < U; q++) {
        const uint64_t u = c + uint64_t(q) * kT;
        b[q] = u < n4 ? ld4<true>(bra + 2 * u) : zero;
        k[q] = same ? b[q] : (u < n4 ? ld4<true>(ket + 2 * u) : zero);
      }
#pragma unroll
      for (int q = 0; q < U; q++) {
        const uint64_t i = 2 * (c + uint64_t(q) * kT);
        const float2 b0 = make_float2(b[q].x, b[q].y), b1 = make_float2(b[q].z, b[q].w);
        const float2 k0 = make_float2(k[q].x, k[q].y), k1 = make_float2(k[q].z, k[q].w);
        if (flip == 0) {
          acc_cj(re, im, b0, k0, sgn(i, sign));
          acc_cj(re, im, b1, k1, sgn(i + 1, sign));
        } else {
          // i -> j = i + 1, and i + 1 -> i
          acc_cj(re, im, b0, k1, sgn(i + 1, sign));
          acc_cj(re, im, b1, k0, sgn(i, sign));
        }
      }
    }
  } else {
    const uint32_t h = 63 - __clzll(flip);  // top bit of flip, >= 1
    const uint64_t fe = flip & ~1ull;
    const bool odd = flip & 1ull;
    constexpr int V = U;  // pairs (i, partner) per thread and step
    const uint64_t n8 = dim / 4, step = uint64_t(gx) * kT * V;
    for (uint64_t c = uint64_t(bx) * kT * V + threadIdx.x; c < n8; c += step) {
      float4 bi[V], ki[V], bp[V], kp[V];
      uint64_t iv[V];
#pragma unroll
      for (int q = 0; q < V; q++) {
        const uint64_t u = c + uint64_t(q) * kT;
        // i: even, bit h clear; its even partner base p has bit h set
        uint64_t t = u < n8 ? u : 0;
        t = ((t >> (h - 1)) << h) | (t & ((1ull << (h - 1)) - 1ull));
        const uint64_t i = t << 1, p = i ^ fe;
        iv[q] = i;
        bi[q] = u < n8 ? ld4<true>(bra + i) : zero;
        bp[q] = u < n8 ? ld4<true>(bra + p) : zero;
        ki[q] = same ? bi[q] : (u < n8 ? ld4<true>(ket + i) : zero);
        kp[q] = same ? bp[q] : (u < n8 ? ld4<true>(ket + p) : zero);
      }
#pragma unroll
      for (int q = 0; q < V; q++) {
        const uint64_t i = iv[q];
        const float2 bi0 = make_float2(bi[q].x, bi[q].y), bi1 = make_float2(bi[q].z, bi[q].w);
        const float2 ki0 = make_float2(ki[q].x, ki[q].y), ki1 = make_float2(ki[q].z, ki[q].w);
        const float2 bp0 = make_float2(bp[q].x, bp[q].y), bp1 = make_float2(bp[q].z, bp[q].w);
        const float2 kp0 = make_float2(kp[q].x, kp[q].y), kp1 = make_float2(kp[q].z, kp[q].w);
        // partner of i is p (+1 if flip is odd); of i+1 is p+1 (-1 if odd)
        const float2 kj0 = odd ? kp1 : kp0, kj1 = odd ? kp0 : kp1;
        const float2 bj0 = odd ? bp1 : bp0, bj1 = odd ? bp0 : bp1;
        const uint64_t j0 = i ^ flip, j1 = (i + 1) ^ flip;
        acc_cj(re, im, bi0, kj0, sgn(j0, sign));
        acc_cj(re, im, bi1, kj1, sgn(j1, sign));
        acc_cj(re, im, bj0, ki0, sgn(i, sign));
        acc_cj(re, im, bj1, ki1, sgn(i + 1, sign));
      }
    }
  }
}

__global__ __launch_bounds__(kT) void k_pauli(const float2 *bra, const float2 *ket, uint64_t dim,
                                              uint64_t flip, uint64_t sign, double *part) {
  __shared__ double sh[kT / 64];
  // Batched launches: instance blockIdx.y, states of `dim` amplitudes back to
  // back, partials per instance.
  bra += uint64_t(blockIdx.y) * dim;
  ket += uint64_t(blockIdx.y) * dim;
  part += size_t(blockIdx.y) * gridDim.x * 2;
  double re = 0.0, im = 0.0;
  pauli_part(bra, ket, dim, flip, sign, blockIdx.x, gridDim.x, re, im);
  const double tr = block_sum(re, sh);
  const double ti = block_sum(im, sh);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = tr;
    part[2 * blockIdx.x + 1] = ti;
  }
}

// <a|P_t|a> for the terms of PauliTerms (blockIdx.y = term t): one launch for
// many terms, partials of term t after t * gridDim.x.
__global__ __launch_bounds__(kT) void k_pauli_terms(const float2 *a, uint64_t dim, PauliTerms t,
                                                    double *part) {
  __shared__ double sh[kT / 64];
  part += size_t(blockIdx.y) * gridDim.x * 2;
  double re = 0.0, im = 0.0;
  pauli_part(a, a, dim, t.flip[blockIdx.y], t.sign[blockIdx.y], blockIdx.x, gridDim.x, re, im);
  const double tr = block_sum(re, sh);
  const double ti = block_sum(im, sh);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = tr;
    part[2 * blockIdx.x + 1] = ti;
  }
}

// Reduce `count` partials with component stride `comps` into out[0..comps)
// (block b: instance b of a batched reduction, partials and outputs back to
// back).
__global__ __launch_bounds__(kT) void k_finish(const double *part, int count, int comps,
                                               double *out) {
  __shared__ double sh[kT / 64];
  part += size_t(blockIdx.x) * count * comps;
  out += size_t(blockIdx.x) * comps;
  for (int c = 0; c < comps; c++) {
    double acc = 0.0;
    for (int k = threadIdx.x; k < count; k += kT) acc += part[comps * k + c];
    const double t = block_sum(acc, sh);
    if (threadIdx.x == 0) out[c] = t;
  }
}

__global__ __launch_bounds__(kT) void k_pauli_accum(AccumArgs p) {
  const uint64_t stride = uint64_t(gridDim.x) * kT;
  const uint64_t fe = p.flip & ~1ull;
  const bool odd = p.flip & 1ull;
  const float2 *psi = p.psi + uint64_t(blockIdx.y) * p.bstate;
  float2 *lam = p.lam + uint64_t(blockIdx.y) * p.bstate;
  for (uint64_t u = uint64_t(blockIdx.x) * kT + threadIdx.x; u < p.dim / 2; u += stride) {
    const uint64_t i = 2 * u, src = i ^ fe;
    const float4 s = *reinterpret_cast<const float4 *>(psi + src);
    float4 *l = reinterpret_cast<float4 *>(lam + i);
    float4 lv = *l;
    const float2 y0 = odd ? make_float2(s.z, s.w) : make_float2(s.x, s.y);
    const float2 y1 = odd ? make_float2(s.x, s.y) : make_float2(s.z, s.w);
    const uint64_t j0 = i ^ p.flip, j1 = (i + 1) ^ p.flip;
    double c0r = 0, c0i = 0, c1r = 0, c1i = 0;
    for (uint32_t k = 0; k < p.nterms; k++) {
      const double s0 = sgn(j0, p.sign[k]), s1 = sgn(j1, p.sign[k]);
      c0r += s0 * p.cre[k];
      c0i += s0 * p.cim[k];
      c1r += s1 * p.cre[k];
      c1i += s1 * p.cim[k];
    }
    lv.x += float(c0r * y0.x - c0i * y0.y);
    lv.y += float(c0r * y0.y + c0i * y0.x);
    lv.z += float(c1r * y1.x - c1i * y1.y);
    lv.w += float(c1r * y1.y + c1i * y1.x);
    *l = lv;
  }
}

// Co-state terms of several flips in one pass (AccumTileArgs). Thread tile:
// bit 0 inside a float4, register bit k on wire w[k]; the source of element e
// under flip slot s is element e ^ (1 << s), a compile-time index. Each
// element's coefficient per slot is summed in FP64 over that slot's terms
// (signs from the source index), then applied in FP32 as in k_pauli_accum.
template <int NW>
__global__ __launch_bounds__(kT) void k_pauli_accum_tile(AccumTileArgs p) {
  constexpr int E = 2 << NW;  // elements per thread
  const uint64_t units = p.dim >> (NW + 1);
  const uint64_t stride = uint64_t(gridDim.x) * kT;
  const float2 *psi = p.psi + uint64_t(blockIdx.y) * p.bstate;
  float2 *lam = p.lam + uint64_t(blockIdx.y) * p.bstate;
  uint64_t wbit[NW > 0 ? NW : 1];
#pragma unroll
  for (int k = 0; k < NW; k++) wbit[k] = 1ull << p.w[k];
  for (uint64_t u = uint64_t(blockIdx.x) * kT + threadIdx.x; u < units; u += stride) {
    uint64_t base = u << 1;
#pragma unroll
    for (int k = 0; k < NW; k++) {
      const uint64_t lo = base & (wbit[k] - 1);
      base = ((base - lo) << 1) | lo;
    }
    uint64_t idx[E / 2];
    float2 y[E], l[E];
#pragma unroll
    for (int r = 0; r < E / 2; r++) {
      uint64_t o = base;
#pragma unroll
      for (int k = 0; k < NW; k++)
        if ((r >> k) & 1) o |= wbit[k];
      idx[r] = o;
      const float4 v = *reinterpret_cast<const float4 *>(psi + o);
      const float4 w = *reinterpret_cast<const float4 *>(lam + o);
      y[2 * r] = make_float2(v.x, v.y);
      y[2 * r + 1] = make_float2(v.z, v.w);
      l[2 * r] = make_float2(w.x, w.y);
      l[2 * r + 1] = make_float2(w.z, w.w);
    }
#pragma unroll
    for (int e = 0; e < E; e++) {
      const uint64_t i = idx[e >> 1] | uint64_t(e & 1);
#pragma unroll
      for (int sl = 0; sl < NW + 2; sl++) {
        // slot sl: flip bit 0 (sl = 0), wire w[sl - 1], or none (sl = NW + 1)
        const uint64_t f = sl == 0 ? 1ull : sl <= NW ? wbit[sl - 1] : 0ull;
        const int src = sl == 0 ? (e ^ 1) : sl <= NW ? (e ^ (1 << sl)) : e;
        const uint64_t j = i ^ f;
        double cr = 0.0, ci = 0.0;
        for (uint32_t t = p.off[sl]; t < p.off[sl + 1]; t++) {
          const double sg = sgn(j, p.sign[t]);
          cr += sg * p.cre[t];
          ci += sg * p.cim[t];
        }
        if (p.off[sl] == p.off[sl + 1]) continue;
        const float2 ys = y[src];
        l[e].x += float(cr * ys.x - ci * ys.y);
        l[e].y += float(cr * ys.y + ci * ys.x);
      }
    }
#pragma unroll
    for (int r = 0; r < E / 2; r++)
      *reinterpret_cast<float4 *>(lam + idx[r]) =
          make_float4(l[2 * r].x, l[2 * r].y, l[2 * r + 1].x, l[2 * r + 1].y);
  }
}

// Fused adjoint step for a rotation on wire q: overlap <lam|P|psi> in FP64
// (P = X, Y or Z on q, taken on the pair (i0, i1 = i0 | 2^q)), then the
// inverse rotation on both states.
__device__ __forceinline__ void adj_pair(const AdjArgs &a, float2 &p0, float2 &p1,
                                         float2 &l0, float2 &l1, double &re, double &im) {
  if (a.axis == 0) {         // X: conj(l0) p1 + conj(l1) p0
    acc_cj(re, im, l0, p1, 1.0);
    acc_cj(re, im, l1, p0, 1.0);
  } else if (a.axis == 1) {  // Y (before i^ny): -conj(l0) p1 + conj(l1) p0
    acc_cj(re, im, l0, p1, -1.0);
    acc_cj(re, im, l1, p0, 1.0);
  } else {                   // Z: conj(l0) p0 - conj(l1) p1
    acc_cj(re, im, l0, p0, 1.0);
    acc_cj(re, im, l1, p1, -1.0);
  }
  rot(a.m, p0, p1);
  rot(a.m, l0, l1);
}

// <lam|P|psi> over the tile's pairs on local bit LB (P on that bit).
template <int LB, int N>
__device__ __forceinline__ void tile_overlap(const float2 (&x)[N], const float2 (&y)[N],
                                             uint32_t axis, double &re, double &im) {
#pragma unroll
  for (int l = 0; l < N; l++) {
    if (l & (1 << LB)) continue;
    const int h = l | (1 << LB);
    if (axis == 0) {         // X: conj(l0) p1 + conj(l1) p0
      acc_cj(re, im, y[l], x[h], 1.0);
      acc_cj(re, im, y[h], x[l], 1.0);
    } else if (axis == 1) {  // Y (before i^ny): -conj(l0) p1 + conj(l1) p0
      acc_cj(re, im, y[l], x[h], -1.0);
      acc_cj(re, im, y[h], x[l], 1.0);
    } else {                 // Z: conj(l0) p0 - conj(l1) p1
      acc_cj(re, im, y[l], x[l], 1.0);
      acc_cj(re, im, y[h], x[h], -1.0);
    }
  }
}

// Merged backward pass (AdjOp). Both tiles are written only by slot_asm
// statements that every op runs (a branch around a write would make the
// compiler hold second copies); the overlap only reads them. Wave sums of
// each overlap go to per-wave LDS slots in program order, then to per-block
// partials in a fixed order: results are run-to-run deterministic.
template <int NL, bool NT>
__global__ __launch_bounds__(kT) void k_adj_pass(AdjPassArgs a, double *part) {
  constexpr int N = 1 << NL;
  __shared__ double acc[kAdjMaxRot][kT / 64][2];
  for (int i = threadIdx.x; i < kAdjMaxRot * (kT / 64) * 2; i += kT) (&acc[0][0][0])[i] = 0.0;
  __syncthreads();
  uint64_t off[N / 2];
  offsets<NL>(a.lay, off);
  const uint64_t inst = blockIdx.y;
  float2 *const psi = a.psi + inst * a.bstate;
  float2 *const lam = a.lam + inst * a.bstate;
  const const_ptr<AdjOp> ops = as_const(bshift(a.ops, inst * a.barena));
  part += size_t(inst) * kAdjMaxRot * gridDim.x * 2;
  const TileAddr ta = tile_addr(a.lay);
  const uint64_t stride = uint64_t(gridDim.x) * kT;
  const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (uint64_t t0 = uint64_t(blockIdx.x) * kT; t0 < a.lay.nthreads; t0 += stride) {
    const uint64_t ub = scatter(t0, a.lay) | a.lay.fixed;
    const bool active = t0 + threadIdx.x < a.lay.nthreads;
    const uint64_t base = ub | ta.lane_bits;
    // Lanes past the end read the workgroup's first tile, add nothing to
    // the overlaps and never store.
    const uint32_t lb = active ? ta.lane_bytes : 0u;
    float2 x[N], y[N];
#pragma unroll
    for (int r = 0; r < N / 2; r++) {
      const float4 v = ld4<NT>(pair_ptr(psi, ub | off[r], lb));
      const float4 w = ld4<NT>(pair_ptr(lam, ub | off[r], lb));
      x[2 * r] = make_float2(v.x, v.y);
      x[2 * r + 1] = make_float2(v.z, v.w);
      y[2 * r] = make_float2(w.x, w.y);
      y[2 * r + 1] = make_float2(w.z, w.w);
    }
    for (uint32_t o = 0; o < a.nops; o++) {
      if (ops[o].type == 1) {
        double re = 0.0, im = 0.0;
        const uint32_t axis = ops[o].axis;
        switch (ops[o].lbit) {
          case 0: tile_overlap<0>(x, y, axis, re, im); break;
          case 1: tile_overlap<1>(x, y, axis, re, im); break;
          case 2: tile_overlap<2>(x, y, axis, re, im); break;
          case 3: tile_overlap<3>(x, y, axis, re, im); break;
          default: tile_overlap<NL - 1>(x, y, axis, re, im); break;
        }
        re = wave_allsum(active ? re : 0.0);
        im = wave_allsum(active ? im : 0.0);
        if (lane == 0) {
          const uint32_t r = ops[o].rslot;
          acc[r][wave][0] += re;
          acc[r][wave][1] += im;
        }
      }
      const uint64_t gmask = ops[o].gmask;
      const bool gon = (base & gmask) == gmask;
      float m[8];
#pragma unroll
      for (int i = 0; i < 8; i++) {
        const float mi = ops[o].m[i];
        m[i] = gon ? mi : (i == 0 || i == 6 ? 1.0f : 0.0f);
      }
      const uint32_t sel = __builtin_amdgcn_readfirstlane(ops[o].sel);
      const uint32_t pm = __builtin_amdgcn_readfirstlane(ops[o].pm);
      slot_asm<NL>(x, sel, pm, m);
      slot_asm<NL>(y, sel, pm, m);
    }
    if (active) {
#pragma unroll
      for (int r = 0; r < N / 2; r++) {
        st4<NT>(pair_ptr(psi, ub | off[r], ta.lane_bytes),
                make_float4(x[2 * r].x, x[2 * r].y, x[2 * r + 1].x, x[2 * r + 1].y));
        st4<NT>(pair_ptr(lam, ub | off[r], ta.lane_bytes),
                make_float4(y[2 * r].x, y[2 * r].y, y[2 * r + 1].x, y[2 * r + 1].y));
      }
    }
  }
  __syncthreads();
  if (threadIdx.x < a.nrot * 2) {
    const uint32_t r = threadIdx.x >> 1, c = threadIdx.x & 1;
    double t = 0.0;
    for (int w = 0; w < kT / 64; w++) t += acc[r][w][c];
    part[(size_t(r) * gridDim.x + blockIdx.x) * 2 + c] = t;
  }
}

// Block r sums rotation r's per-block partials in a fixed order (block row y:
// instance y of a batched pass, its partials kAdjMaxRot * count * 2 doubles
// apart and its outputs `ostride` doubles apart).
__global__ __launch_bounds__(kT) void k_finish_rot(const double *part, int count, double *out,
                                                   uint64_t ostride, int maxrot) {
  __shared__ double sh[kT / 64];
  const size_t r = blockIdx.x;
  part += size_t(blockIdx.y) * maxrot * count * 2;
  out += size_t(blockIdx.y) * ostride;
  for (int c = 0; c < 2; c++) {
    double acc = 0.0;
    for (int k = threadIdx.x; k < count; k += kT) acc += part[(r * count + k) * 2 + c];
    const double t = block_sum(acc, sh);
    if (threadIdx.x == 0) out[2 * r + c] = t;
  }
}

// FP64 conj(x) y accumulated with fused multiply-adds (NEG: subtracted).
template <bool NEG>
__device__ __forceinline__ void acc_cj_fma(double &re, double &im, float2 x, float2 y) {
  const double xr = x.x, xi = x.y, yr = y.x, yi = y.y;
  if (NEG) {
    re = fma(-xr, yr, re);
    re = fma(-xi, yi, re);
    im = fma(-xr, yi, im);
    im = fma(xi, yr, im);
  } else {
    re = fma(xr, yr, re);
    re = fma(xi, yi, re);
    im = fma(xr, yi, im);
    im = fma(-xi, yr, im);
  }
}

// a +/-= part COMP (0 real, 1 imaginary) of conj(x) y with FP32 fused
// multiply-adds (the negations are free source modifiers).
template <bool NEG, int COMP>
__device__ __forceinline__ void acc_part(float &a, float2 x, float2 y) {
  if (COMP == 0) {
    a = fmaf(NEG ? -x.x : x.x, y.x, a);
    a = fmaf(NEG ? -x.y : x.y, y.y, a);
  } else {
    a = fmaf(NEG ? -x.x : x.x, y.y, a);
    a = fmaf(NEG ? x.y : -x.y, y.x, a);
  }
}

// Part COMP of <lam|P|psi> over a 16-amplitude tile for P = X / Y / Z on
// local bit LB (axis 0..2; Y before the i^ny factor, as tile_overlap), into
// two FP32 partial sums (even and odd pairs).
template <int LB, int COMP>
__device__ __forceinline__ void tile_overlap16(const float2 (&x)[16], const float2 (&y)[16],
                                               uint32_t axis, float (&a)[2]) {
  int k = 0;
#pragma unroll
  for (int l = 0; l < 16; l++) {
    if (l & (1 << LB)) continue;
    const int h = l | (1 << LB);
    float &v = a[k++ & 1];
    if (axis == 0) {
      acc_part<false, COMP>(v, y[l], x[h]);
      acc_part<false, COMP>(v, y[h], x[l]);
    } else if (axis == 1) {
      acc_part<true, COMP>(v, y[l], x[h]);
      acc_part<false, COMP>(v, y[h], x[l]);
    } else {
      acc_part<false, COMP>(v, y[l], x[l]);
      acc_part<true, COMP>(v, y[h], x[h]);
    }
  }
}

// Part COMP of sum_l s(l) conj(y_l) x_l with s(l) = (-1)^(bit LA of l + bit
// LB of l); LA = LB = -1: s = 1.
template <int LA, int LB, int COMP>
__device__ __forceinline__ void tile_overlap_parity(const float2 (&x)[16], const float2 (&y)[16],
                                                    float (&a)[2]) {
#pragma unroll
  for (int l = 0; l < 16; l++) {
    const int par = (LA >= 0 ? (l >> (LA < 0 ? 0 : LA)) & 1 : 0) ^
                    (LB >= 0 ? (l >> (LB < 0 ? 0 : LB)) & 1 : 0);
    if (par) acc_part<true, COMP>(a[l & 1], y[l], x[l]);
    else acc_part<false, COMP>(a[l & 1], y[l], x[l]);
  }
}

// The thread's part of one overlap: FP32 over its 16 amplitudes (at most 32
// products in two sums, relative error ~ 16 * 2^-24 of each sum), FP64 from
// there on (lanes, waves, workgroups, the finish).
template <int COMP>
__device__ __forceinline__ double adj_op_overlap(const float2 (&x)[16], const float2 (&y)[16],
                                                 uint32_t axis, uint32_t lb, uint32_t lb2) {
  float a[2] = {0.0f, 0.0f};
  if (axis <= 2) {
    switch (lb) {
      case 0: tile_overlap16<0, COMP>(x, y, axis, a); break;
      case 1: tile_overlap16<1, COMP>(x, y, axis, a); break;
      case 2: tile_overlap16<2, COMP>(x, y, axis, a); break;
      default: tile_overlap16<3, COMP>(x, y, axis, a); break;
    }
  } else if (axis == 3) {
    tile_overlap_parity<-1, -1, COMP>(x, y, a);
  } else {
    switch (lb * 4 + lb2) {  // lb < lb2
      case 1: tile_overlap_parity<0, 1, COMP>(x, y, a); break;
      case 2: tile_overlap_parity<0, 2, COMP>(x, y, a); break;
      case 3: tile_overlap_parity<0, 3, COMP>(x, y, a); break;
      case 6: tile_overlap_parity<1, 2, COMP>(x, y, a); break;
      case 7: tile_overlap_parity<1, 3, COMP>(x, y, a); break;
      default: tile_overlap_parity<2, 3, COMP>(x, y, a); break;
    }
  }
  return double(a[0]) + double(a[1]);
}

// Moves 16 amplitudes per thread (8 slots) between two arrangements, as
// tile_permute does for 32.
__device__ __forceinline__ void tile_permute16(float4 *lds, float2 (&x)[16], uint32_t tf,
                                               const uint32_t (&rf)[3], uint32_t tt,
                                               const uint32_t (&rt)[3]) {
  char *base = reinterpret_cast<char *>(lds);
  __syncthreads();  // earlier reads of the buffer are done
#pragma unroll
  for (int r = 0; r < 8; r++) {
    const uint32_t o = ((r & 1) ? rf[0] : 0u) ^ ((r & 2) ? rf[1] : 0u) ^ ((r & 4) ? rf[2] : 0u);
    *reinterpret_cast<float4 *>(base + ((tf ^ o) << 4)) =
        make_float4(x[2 * r].x, x[2 * r].y, x[2 * r + 1].x, x[2 * r + 1].y);
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < 8; r++) {
    const uint32_t o = ((r & 1) ? rt[0] : 0u) ^ ((r & 2) ? rt[1] : 0u) ^ ((r & 4) ? rt[2] : 0u);
    const float4 v = *reinterpret_cast<const float4 *>(base + ((tt ^ o) << 4));
    x[2 * r] = make_float2(v.x, v.y);
    x[2 * r + 1] = make_float2(v.z, v.w);
  }
}

// Wavefront sum from the device library (a fixed sequence of cross-lane
// steps: the same order on every run).
extern "C" __device__ __attribute__((const)) double __ockl_wfred_add_f64(double);

// AdjOps [o0, o1) on the thread's 16 amplitudes of each state. Each op's
// fields are loaded as one block into scalar registers (a lane-dependent
// choice between two fields read from memory would otherwise become a
// per-lane load; loading the next op's block ahead spills scalar registers). Overlaps go to acc[rslot][wave][re, im]
// (the part `comp` only; wave sums, lane 0 adds them in program order).
template <bool ONE>
__device__ __forceinline__ void adj_tile_ops(float2 (&x)[16], float2 (&y)[16],
                                             const_ptr<AdjOp> ops, uint32_t o0, uint32_t o1,
                                             uint64_t base, uint32_t lane, uint32_t wave,
                                             double *acc) {
  struct Hdr {
    uint32_t type, lbit, sel, pm, axis, rslot, lbit2, comp;
    uint64_t gmask, pmask;
    float m[8], malt[8];
  };
  auto load_hdr = [&](uint32_t o, Hdr &h) {
    h.type = ops[o].type;
    h.lbit = ops[o].lbit;
    h.sel = ops[o].sel;
    h.pm = ops[o].pm;
    h.axis = ops[o].axis;
    h.rslot = ops[o].rslot;
    h.lbit2 = ops[o].lbit2;
    h.comp = ops[o].comp;
    h.gmask = ops[o].gmask;
    h.pmask = ops[o].pmask;
#pragma unroll
    for (int i = 0; i < 8; i++) {
      h.m[i] = ops[o].m[i];
      h.malt[i] = ops[o].malt[i];
    }
  };
  auto pin = [](Hdr &h) {
    asm volatile("" : "+s"(h.type), "+s"(h.lbit), "+s"(h.sel), "+s"(h.pm), "+s"(h.axis),
                 "+s"(h.rslot), "+s"(h.lbit2), "+s"(h.comp), "+s"(h.gmask), "+s"(h.pmask));
    asm volatile("" : "+s"(h.m[0]), "+s"(h.m[1]), "+s"(h.m[2]), "+s"(h.m[3]), "+s"(h.m[4]),
                 "+s"(h.m[5]), "+s"(h.m[6]), "+s"(h.m[7]));
    asm volatile("" : "+s"(h.malt[0]), "+s"(h.malt[1]), "+s"(h.malt[2]), "+s"(h.malt[3]),
                 "+s"(h.malt[4]), "+s"(h.malt[5]), "+s"(h.malt[6]), "+s"(h.malt[7]));
  };
  for (uint32_t o = o0; o < o1; o++) {
    Hdr h;
    load_hdr(o, h);
    pin(h);
    const bool odd = (__builtin_popcountll(base & h.pmask) & 1) != 0;
    if (!ONE && h.type == 1) {
      double v = h.comp ? adj_op_overlap<1>(x, y, h.axis, h.lbit, h.lbit2)
                        : adj_op_overlap<0>(x, y, h.axis, h.lbit, h.lbit2);
      v = __ockl_wfred_add_f64(odd ? -v : v);
      if (lane == 0) acc[(h.rslot * (kAdjTileThreads / 64) + wave) * 2 + h.comp] += v;
    }
    float m[8];
    if ((h.gmask | h.pmask) == 0) {
#pragma unroll
      for (int i = 0; i < 8; i++) m[i] = h.m[i];
    } else {
      const bool gon = (base & h.gmask) == h.gmask;
#pragma unroll
      for (int i = 0; i < 8; i++) {
        const float mi = odd ? h.malt[i] : h.m[i];
        m[i] = gon ? mi : (i == 0 || i == 6 ? 1.0f : 0.0f);
      }
    }
    slot_asm<4>(x, h.sel, h.pm, m);
    if (!ONE) slot_asm<4>(y, h.sel, h.pm, m);
  }
}

// Backward tile pass (AdjTileArgs). Every workgroup works on whole tiles, so
// all its threads take part in every op; overlaps are reduced as in
// k_adj_pass (wave sums into per-wave LDS slots in program order, per-block
// partials, a fixed-order finish).
// ONE: the forward sweep of a device-resident gradient program, on psi only.
template <bool ONE>
__global__ __launch_bounds__(kAdjTileThreads) __attribute__((amdgpu_waves_per_eu(4)))
void k_adj_tile(AdjTileArgs a, double *part) {
  constexpr int W = kAdjTileThreads / 64;
  __shared__ float4 lds[2048];
  __shared__ double acc[kAdjTileMaxRot][W][2];
  for (int i = threadIdx.x; i < kAdjTileMaxRot * W * 2; i += kAdjTileThreads)
    (&acc[0][0][0])[i] = 0.0;
  __syncthreads();
  const uint32_t lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const uint64_t inst = blockIdx.y;
  float2 *const psi = a.psi + inst * a.bstate;
  float2 *const lam = ONE ? psi : a.lam + inst * a.bstate;
  const const_ptr<AdjSeg> segs = as_const(bshift(a.segs, inst * a.barena));
  const const_ptr<AdjOp> ops = as_const(bshift(a.ops, inst * a.barena));
  if (!ONE) part += size_t(inst) * kAdjTileMaxRot * gridDim.x * 2;
  const uint32_t lane_bytes = lane * 16u;
  // Load arrangement: lane bit j on tile bit j + 1 (wire j + 1), wave bit w
  // on tile bit 7 + w (hpos[w]), register bit k on tile bit 9 + k
  // (hpos[2 + k]).
  const uint64_t wave_off = ((wave & 1u) ? 1ull << a.hpos[0] : 0ull) |
                            ((wave & 2u) ? 1ull << a.hpos[1] : 0ull);
  uint64_t reg_off[8];
#pragma unroll
  for (int r = 0; r < 8; r++)
    reg_off[r] = ((r & 1) ? 1ull << a.hpos[2] : 0ull) | ((r & 2) ? 1ull << a.hpos[3] : 0ull) |
                 ((r & 4) ? 1ull << a.hpos[4] : 0ull);
  uint32_t t0 = 0;
#pragma unroll
  for (int j = 0; j < 6; j++)
    if ((lane >> j) & 1) t0 ^= tile_swizzle(1u << j);
#pragma unroll
  for (int w = 0; w < 2; w++)
    if ((wave >> w) & 1) t0 ^= tile_swizzle(1u << (6 + w));
  uint32_t r0[3];
#pragma unroll
  for (int k = 0; k < 3; k++) r0[k] = tile_swizzle(1u << (8 + k));
  for (uint64_t tile = blockIdx.x; tile < a.lay.nthreads; tile += gridDim.x) {
    const uint64_t ub = scatter(tile, a.lay) | a.lay.fixed;
    float2 x[16], y[16];
#pragma unroll
    for (int r = 0; r < 8; r++) {
      const float4 v = ld4<true>(pair_ptr(psi, ub | wave_off | reg_off[r], lane_bytes));
      x[2 * r] = make_float2(v.x, v.y);
      x[2 * r + 1] = make_float2(v.z, v.w);
      if (!ONE) {
        const float4 w = ld4<true>(pair_ptr(lam, ub | wave_off | reg_off[r], lane_bytes));
        y[2 * r] = make_float2(w.x, w.y);
        y[2 * r + 1] = make_float2(w.z, w.w);
      } else {
        y[2 * r] = y[2 * r + 1] = make_float2(0.0f, 0.0f);
      }
    }
    uint32_t tf = t0, rf[3] = {r0[0], r0[1], r0[2]};
    for (uint32_t s = 0; s < a.nsegs; s++) {
      // The segment's fields as scalars (each loaded once, whatever the
      // lane), then combined per lane.
      uint32_t ts[8], tp[8], rt[3];
#pragma unroll
      for (int j = 0; j < 8; j++) {
        ts[j] = segs[s].tslot[j];
        tp[j] = segs[s].tpos[j];
      }
#pragma unroll
      for (int k = 0; k < 3; k++) rt[k] = segs[s].rslot[k];
      const uint32_t op0 = segs[s].op0, op1 = segs[s].op1;
      asm volatile("" : "+s"(ts[0]), "+s"(ts[1]), "+s"(ts[2]), "+s"(ts[3]), "+s"(ts[4]),
                   "+s"(ts[5]), "+s"(ts[6]), "+s"(ts[7]));
      asm volatile("" : "+s"(tp[0]), "+s"(tp[1]), "+s"(tp[2]), "+s"(tp[3]), "+s"(tp[4]),
                   "+s"(tp[5]), "+s"(tp[6]), "+s"(tp[7]), "+s"(rt[0]), "+s"(rt[1]), "+s"(rt[2]));
      uint32_t tt = 0;
      uint64_t base = ub;
#pragma unroll
      for (int j = 0; j < 8; j++) {
        const bool on = j < 6 ? ((lane >> j) & 1) != 0 : ((wave >> (j - 6)) & 1) != 0;
        tt ^= on ? ts[j] : 0u;
        base |= on ? 1ull << tp[j] : 0ull;
      }
      tile_permute16(lds, x, tf, rf, tt, rt);
      if (!ONE) tile_permute16(lds, y, tf, rf, tt, rt);
      adj_tile_ops<ONE>(x, y, ops, op0, op1, base, lane, wave, &acc[0][0][0]);
      tf = tt;
#pragma unroll
      for (int k = 0; k < 3; k++) rf[k] = rt[k];
    }
    tile_permute16(lds, x, tf, rf, t0, r0);
    if (!ONE) tile_permute16(lds, y, tf, rf, t0, r0);
#pragma unroll
    for (int r = 0; r < 8; r++) {
      st4<true>(pair_ptr(psi, ub | wave_off | reg_off[r], lane_bytes),
                make_float4(x[2 * r].x, x[2 * r].y, x[2 * r + 1].x, x[2 * r + 1].y));
      if (!ONE)
        st4<true>(pair_ptr(lam, ub | wave_off | reg_off[r], lane_bytes),
                  make_float4(y[2 * r].x, y[2 * r].y, y[2 * r + 1].x, y[2 * r + 1].y));
    }
  }
  if (ONE) return;
  __syncthreads();
  if (threadIdx.x < a.nrot * 2) {
    const uint32_t r = threadIdx.x >> 1, c = threadIdx.x & 1;
    double t = 0.0;
    for (int w = 0; w < W; w++) t += acc[r][w][c];
    part[(size_t(r) * gridDim.x + blockIdx.x) * 2 + c] = t;
  }
}

// Co-state tile pass (CoTileArgs).
__global__ __launch_bounds__(kAdjTileThreads) void k_costate_tile(CoTileArgs a, double *epart) {
  constexpr int W = kAdjTileThreads / 64;
  __shared__ float4 lds[2048];
  __shared__ double red[W][2];
  const uint32_t lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const uint64_t inst = blockIdx.y;
  const float2 *const psi = a.psi + inst * a.bstate;
  float2 *const lam = a.lam + inst * a.bstate;
  const uint32_t lane_bytes = lane * 16u;
  const uint64_t wave_off = ((wave & 1u) ? 1ull << a.hpos[0] : 0ull) |
                            ((wave & 2u) ? 1ull << a.hpos[1] : 0ull);
  uint64_t reg_off[8];
#pragma unroll
  for (int r = 0; r < 8; r++)
    reg_off[r] = ((r & 1) ? 1ull << a.hpos[2] : 0ull) | ((r & 2) ? 1ull << a.hpos[3] : 0ull) |
                 ((r & 4) ? 1ull << a.hpos[4] : 0ull);
  // The thread's slots (load arrangement of k_adj_tile) and its base index
  // (lane and wave bits; local bits are 0).
  uint32_t t0 = 0;
#pragma unroll
  for (int j = 0; j < 6; j++)
    if ((lane >> j) & 1) t0 ^= tile_swizzle(1u << j);
#pragma unroll
  for (int w = 0; w < 2; w++)
    if ((wave >> w) & 1) t0 ^= tile_swizzle(1u << (6 + w));
  uint32_t slot[8];
#pragma unroll
  for (int r = 0; r < 8; r++)
    slot[r] = t0 ^ ((r & 1) ? tile_swizzle(1u << 8) : 0u) ^ ((r & 2) ? tile_swizzle(1u << 9) : 0u) ^
              ((r & 4) ? tile_swizzle(1u << 10) : 0u);
  const uint64_t thr_bits = (uint64_t(lane) << 1) | wave_off;
  double ere = 0.0, eim = 0.0;
  for (uint64_t tile = blockIdx.x; tile < a.lay.nthreads; tile += gridDim.x) {
    const uint64_t ub = scatter(tile, a.lay) | a.lay.fixed;
    const uint64_t base = ub | thr_bits;
    float2 x[16], y[16];
#pragma unroll
    for (int r = 0; r < 8; r++) {
      const float4 v = ld4<true>(pair_ptr(psi, ub | wave_off | reg_off[r], lane_bytes));
      x[2 * r] = make_float2(v.x, v.y);
      x[2 * r + 1] = make_float2(v.z, v.w);
    }
    if (a.first) {
#pragma unroll
      for (int l = 0; l < 16; l++) y[l] = make_float2(0.0f, 0.0f);
    } else {
#pragma unroll
      for (int r = 0; r < 8; r++) {
        const float4 w = ld4<true>(pair_ptr(lam, ub | wave_off | reg_off[r], lane_bytes));
        y[2 * r] = make_float2(w.x, w.y);
        y[2 * r + 1] = make_float2(w.z, w.w);
      }
    }
    __syncthreads();  // the previous tile's reads of the buffer are done
#pragma unroll
    for (int r = 0; r < 8; r++)
      lds[slot[r]] = make_float4(x[2 * r].x, x[2 * r].y, x[2 * r + 1].x, x[2 * r + 1].y);
    __syncthreads();
    if (a.doff[16] > 0) {
      // Coefficients by local mask, then a Walsh-Hadamard transform: factor
      // of local combination l = sum_m w[m] (-1)^popcount(l & m).
      double w[16];
#pragma unroll
      for (int m = 0; m < 16; m++) {
        double acc = 0.0;
        for (uint32_t t = a.doff[m]; t < a.doff[m + 1]; t++)
          acc += (__popcll(base & a.sign[t]) & 1) ? -a.cre[t] : a.cre[t];
        w[m] = acc;
      }
#pragma unroll
      for (int b = 1; b < 16; b <<= 1)
#pragma unroll
        for (int l = 0; l < 16; l++)
          if (!(l & b)) {
            const double u = w[l], v = w[l | b];
            w[l] = u + v;
            w[l | b] = u - v;
          }
#pragma unroll
      for (int l = 0; l < 16; l++) {
        const float f = float(w[l]);
        y[l].x = fmaf(f, x[l].x, y[l].x);
        y[l].y = fmaf(f, x[l].y, y[l].y);
      }
    }
    for (uint32_t g = 0; g < a.ngroups; g++) {
      const uint32_t gs = a.gslot[g];
      const bool swap = (a.gflip[g] & 1ull) != 0;
      float2 p[16];
#pragma unroll
      for (int r = 0; r < 8; r++) {
        const float4 v = lds[slot[r] ^ gs];
        p[2 * r] = swap ? make_float2(v.z, v.w) : make_float2(v.x, v.y);
        p[2 * r + 1] = swap ? make_float2(v.x, v.y) : make_float2(v.z, v.w);
      }
      // The group's plain terms first (real coefficient, no sign over the
      // local bits): y += s c p.
      const uint32_t tp = a.goff[g] + ((a.gpart[g] >> 8) & 0xFFu);
      for (uint32_t t = a.goff[g]; t < tp; t++) {
        const float c = float(a.cre[t]);
        const float sr = (__popcll(base & a.sign[t]) & 1) ? -c : c;
#pragma unroll
        for (int l = 0; l < 16; l++) {
          y[l].x = fmaf(sr, p[l].x, y[l].x);
          y[l].y = fmaf(sr, p[l].y, y[l].y);
        }
      }
      for (uint32_t t = tp; t < a.goff[g + 1]; t++) {
        const bool neg = (__popcll(base & a.sign[t]) & 1) != 0;
        const float cr = float(a.cre[t]), ci = float(a.cim[t]);
        const float sr = neg ? -cr : cr, si = neg ? -ci : ci;
        const uint32_t lm = a.lmask[t];
#pragma unroll
        for (int l = 0; l < 16; l++) {
          const bool fl = (__popc(uint32_t(l) & lm) & 1) != 0;
          const float qr = fl ? -sr : sr, qi = fl ? -si : si;
          y[l].x = fmaf(qr, p[l].x, fmaf(-qi, p[l].y, y[l].x));
          y[l].y = fmaf(qr, p[l].y, fmaf(qi, p[l].x, y[l].y));
        }
      }
    }
    if (epart) {
#pragma unroll
      for (int l = 0; l < 16; l++) acc_cj_fma<false>(ere, eim, x[l], y[l]);
    }
#pragma unroll
    for (int r = 0; r < 8; r++)
      st4<true>(pair_ptr(lam, ub | wave_off | reg_off[r], lane_bytes),
                make_float4(y[2 * r].x, y[2 * r].y, y[2 * r + 1].x, y[2 * r + 1].y));
  }
  if (epart) {
    ere = __ockl_wfred_add_f64(ere);
    eim = __ockl_wfred_add_f64(eim);
    if (lane == 0) {
      red[wave][0] = ere;
      red[wave][1] = eim;
    }
    __syncthreads();
    if (threadIdx.x < 2) {
      double t = 0.0;
      for (int w = 0; w < W; w++) t += red[w][threadIdx.x];
      epart[(size_t(inst) * gridDim.x + blockIdx.x) * 2 + threadIdx.x] = t;
    }
  }
}

// Walsh-signed sum over a thread's 16 amplitudes, sum_l (-1)^popcount(l & m)
// v[l]: the sign factors over the bits of l, so the sum folds one bit at a
// time (15 FMAs, the four +-1 factors uniform).
__device__ __forceinline__ float walsh16(const float (&v)[16], uint32_t m) {
  const float s0 = __uint_as_float(0x3f800000u | ((m & 1u) << 31));
  const float s1 = __uint_as_float(0x3f800000u | ((m & 2u) << 30));
  const float s2 = __uint_as_float(0x3f800000u | ((m & 4u) << 29));
  const float s3 = __uint_as_float(0x3f800000u | ((m & 8u) << 28));
  float a[8], b[4];
#pragma unroll
  for (int k = 0; k < 8; k++) a[k] = fmaf(s0, v[2 * k + 1], v[2 * k]);
#pragma unroll
  for (int k = 0; k < 4; k++) b[k] = fmaf(s1, a[2 * k + 1], a[2 * k]);
  return fmaf(s3, fmaf(s2, b[3], b[2]), fmaf(s2, b[1], b[0]));
}

// Sign of term t at the thread's base index: the thread's own part of the
// parity (bit t of tpar, fixed for the whole pass) and the tile's (upar, the
// same for every thread of the block).
__device__ __forceinline__ float term_sign(float v, uint64_t tpar, uint32_t upar, uint32_t t) {
  return __uint_as_float(__float_as_uint(v) ^ ((uint32_t(tpar >> t) ^ upar) << 31));
}

// Plain groups (one term, real coefficient, no sign over the thread's own
// bits): the share before c and the base sign is sum_l Re(conj(x_l) p_l),
// one FMA chain over the partners.
template <bool SWAP>
__device__ __forceinline__ float energy_lds_plain(const float4 *lds, const uint32_t (&slot)[8],
                                                 uint32_t gs, const float2 (&x)[16]) {
  float acc0 = 0.0f, acc1 = 0.0f;
#pragma unroll
  for (int r = 0; r < 8; r++) {
    const float4 v = lds[slot[r] ^ gs];
    const float2 p0 = SWAP ? make_float2(v.z, v.w) : make_float2(v.x, v.y);
    const float2 p1 = SWAP ? make_float2(v.x, v.y) : make_float2(v.z, v.w);
    acc0 = fmaf(x[2 * r].x, p0.x, acc0);
    acc0 = fmaf(x[2 * r].y, p0.y, acc0);
    acc1 = fmaf(x[2 * r + 1].x, p1.x, acc1);
    acc1 = fmaf(x[2 * r + 1].y, p1.y, acc1);
  }
  return acc0 + acc1;
}

// A plain group whose flip is one of the thread's own local bits (bit B of
// l): both members of every pair (l, l ^ 2^B) sit in this thread, and a
// Hermitian term's two contributions conj(x_i) (P x)_i and conj(x_j) (P x)_j
// are complex conjugates, so twice the real part over the half with bit B
// clear is the whole. k indexes that half (l with bit B removed).
template <int B>
__device__ __forceinline__ int half_index(int k) {
  return ((k >> B) << (B + 1)) | (k & ((1 << B) - 1));
}
template <int B>
__device__ __forceinline__ float energy_local_plain(const float2 (&x)[16]) {
  float acc0 = 0.0f, acc1 = 0.0f;
#pragma unroll
  for (int k = 0; k < 8; k += 2) {
    const float2 u0 = x[half_index<B>(k)], v0 = x[half_index<B>(k) | (1 << B)];
    const float2 u1 = x[half_index<B>(k + 1)], v1 = x[half_index<B>(k + 1) | (1 << B)];
    acc0 = fmaf(u0.x, v0.x, acc0);
    acc0 = fmaf(u0.y, v0.y, acc0);
    acc1 = fmaf(u1.x, v1.x, acc1);
    acc1 = fmaf(u1.y, v1.y, acc1);
  }
  return 2.0f * (acc0 + acc1);
}

// One flip group's share of the energy: partner amplitudes from LDS (slot
// offset gs; SWAP when the flip holds amplitude bit 0), z_l = conj(x_l) p_l
// in the parts PART asks for (bit 0 real, bit 1 imaginary), then per term
// Re(c sum_l (-1)^popcount(l & lm) z_l) with its sign at the base index.
template <bool SWAP, int PART>
__device__ __forceinline__ float energy_lds_group(const float4 *lds, const uint32_t (&slot)[8],
                                                 uint32_t gs, const float2 (&x)[16],
                                                 const CoTileArgs &a, uint32_t t0, uint32_t t1,
                                                 uint64_t tpar, uint64_t ub) {
  float zr[16], zi[16];
#pragma unroll
  for (int r = 0; r < 8; r++) {
    const float4 v = lds[slot[r] ^ gs];
    const float2 p0 = SWAP ? make_float2(v.z, v.w) : make_float2(v.x, v.y);
    const float2 p1 = SWAP ? make_float2(v.x, v.y) : make_float2(v.z, v.w);
    const float2 u0 = x[2 * r], u1 = x[2 * r + 1];
    if (PART & 1) {
      zr[2 * r] = fmaf(u0.x, p0.x, u0.y * p0.y);
      zr[2 * r + 1] = fmaf(u1.x, p1.x, u1.y * p1.y);
    }
    if (PART & 2) {
      zi[2 * r] = fmaf(u0.x, p0.y, -(u0.y * p0.x));
      zi[2 * r + 1] = fmaf(u1.x, p1.y, -(u1.y * p1.x));
    }
  }
  float f = 0.0f;
  for (uint32_t t = t0; t < t1; t++) {
    const uint32_t lm = a.lmask[t];
    const float cr = float(a.cre[t]), ci = float(a.cim[t]);
    float v = 0.0f;
    if ((PART & 1) && cr != 0.0f) v = cr * walsh16(zr, lm);
    if ((PART & 2) && ci != 0.0f) v = fmaf(-ci, walsh16(zi, lm), v);
    f += term_sign(v, tpar, uint32_t(__popcll(ub & a.sign[t]) & 1), t);
  }
  return f;
}
template <bool SWAP>
__device__ __forceinline__ float energy_lds_pick(uint32_t part, const float4 *lds,
                                                const uint32_t (&slot)[8], uint32_t gs,
                                                const float2 (&x)[16], const CoTileArgs &a,
                                                uint32_t t0, uint32_t t1, uint64_t tpar,
                                                uint64_t ub) {
  if (part == 1) return energy_lds_group<SWAP, 1>(lds, slot, gs, x, a, t0, t1, tpar, ub);
  if (part == 2) return energy_lds_group<SWAP, 2>(lds, slot, gs, x, a, t0, t1, tpar, ub);
  return energy_lds_group<SWAP, 3>(lds, slot, gs, x, a, t0, t1, tpar, ub);
}

// Energy-only co-state tile pass: <psi|H_t psi> of one tile's terms H_t
// (CoTileArgs without lam), psi read once and nothing written. With x_l the
// thread's 16 amplitudes (local combination l) and f_l = sum_m w_m
// (-1)^popcount(l & m) the diagonal factor, the diagonal part is
// sum_l f_l |x_l|^2 = sum_m w_m Q_m, Q the Walsh-Hadamard transform of
// |x_l|^2; a flip term with coefficient c, base sign s and local mask lm
// gives s c sum_l (-1)^popcount(l & lm) conj(x_l) p_l (p_l its partner
// amplitude), whose real part is what the energy keeps. Per-tile sums run in
// FP32 over the thread's 16 amplitudes, FP64 from there on.
__global__ __launch_bounds__(kAdjTileThreads) void k_costate_energy(CoTileArgs a, double *epart) {
  constexpr int W = kAdjTileThreads / 64;
  __shared__ float4 lds[2048];
  __shared__ double red[W];
  const uint32_t lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const uint64_t inst = blockIdx.y;
  const float2 *const psi = a.psi + inst * a.bstate;
  const uint32_t lane_bytes = lane * 16u;
  const uint64_t wave_off = ((wave & 1u) ? 1ull << a.hpos[0] : 0ull) |
                            ((wave & 2u) ? 1ull << a.hpos[1] : 0ull);
  uint64_t reg_off[8];
#pragma unroll
  for (int r = 0; r < 8; r++)
    reg_off[r] = ((r & 1) ? 1ull << a.hpos[2] : 0ull) | ((r & 2) ? 1ull << a.hpos[3] : 0ull) |
                 ((r & 4) ? 1ull << a.hpos[4] : 0ull);
  uint32_t t0 = 0;
#pragma unroll
  for (int j = 0; j < 6; j++)
    if ((lane >> j) & 1) t0 ^= tile_swizzle(1u << j);
#pragma unroll
  for (int w = 0; w < 2; w++)
    if ((wave >> w) & 1) t0 ^= tile_swizzle(1u << (6 + w));
  uint32_t slot[8];
#pragma unroll
  for (int r = 0; r < 8; r++)
    slot[r] = t0 ^ ((r & 1) ? tile_swizzle(1u << 8) : 0u) ^ ((r & 2) ? tile_swizzle(1u << 9) : 0u) ^
              ((r & 4) ? tile_swizzle(1u << 10) : 0u);
  const uint64_t thr_bits = (uint64_t(lane) << 1) | wave_off;
  static_assert(kCoTerms <= 64, "term parities are kept in one 64-bit mask");
  const uint32_t nterms = a.goff[a.ngroups];
  uint64_t tpar = 0;
  for (uint32_t t = 0; t < nterms; t++)
    tpar |= uint64_t(__popcll(thr_bits & a.sign[t]) & 1) << t;
  bool shared = false;  // some group reads its partners from LDS
  for (uint32_t g = 0; g < a.ngroups; g++) shared |= !(a.gpart[g] & kCoGroupLocal);
  const bool diag = a.doff[16] > 0;
  double e = 0.0;
  for (uint64_t tile = blockIdx.x; tile < a.lay.nthreads; tile += gridDim.x) {
    const uint64_t ub = scatter(tile, a.lay) | a.lay.fixed;
    float2 x[16];
#pragma unroll
    for (int r = 0; r < 8; r++) {
      const float4 v = ld4<true>(pair_ptr(psi, ub | wave_off | reg_off[r], lane_bytes));
      x[2 * r] = make_float2(v.x, v.y);
      x[2 * r + 1] = make_float2(v.z, v.w);
    }
    if (shared) {
      __syncthreads();  // the previous tile's reads of the buffer are done
#pragma unroll
      for (int r = 0; r < 8; r++)
        lds[slot[r]] = make_float4(x[2 * r].x, x[2 * r].y, x[2 * r + 1].x, x[2 * r + 1].y);
      __syncthreads();
    }
    if (diag) {
      float q[16];
#pragma unroll
      for (int l = 0; l < 16; l++) q[l] = fmaf(x[l].x, x[l].x, x[l].y * x[l].y);
#pragma unroll
      for (int b = 1; b < 16; b <<= 1)
#pragma unroll
        for (int l = 0; l < 16; l++)
          if (!(l & b)) {
            const float u = q[l], v = q[l | b];
            q[l] = u + v;
            q[l | b] = u - v;
          }
      double d = 0.0;
#pragma unroll
      for (int m = 0; m < 16; m++) {
        double w = 0.0;
        for (uint32_t t = a.doff[m]; t < a.doff[m + 1]; t++) {
          const uint32_t neg = (uint32_t(tpar >> t) ^ uint32_t(__popcll(ub & a.sign[t]))) << 31;
          const double c = a.cre[t];
          w += __hiloint2double(__double2hiint(c) ^ int(neg), __double2loint(c));
        }
        d = fma(w, double(q[m]), d);
      }
      e += d;
    }
    float f = 0.0f;
    for (uint32_t g = 0; g < a.ngroups; g++) {
      const uint32_t gp = a.gpart[g], part = gp & 3u;
      if (!part) continue;
      const uint32_t g0 = a.goff[g], g1 = a.goff[g + 1];
      const bool swap = (a.gflip[g] & 1ull) != 0;
      if (gp & kCoGroupPlain) {
        float v;
        if (gp & kCoGroupLocal) {
          switch ((gp >> 4) & 3u) {
            case 0: v = energy_local_plain<0>(x); break;
            case 1: v = energy_local_plain<1>(x); break;
            case 2: v = energy_local_plain<2>(x); break;
            default: v = energy_local_plain<3>(x); break;
          }
        } else {
          v = swap ? energy_lds_plain<true>(lds, slot, a.gslot[g], x)
                   : energy_lds_plain<false>(lds, slot, a.gslot[g], x);
        }
        f += term_sign(float(a.cre[g0]) * v, tpar, uint32_t(__popcll(ub & a.sign[g0]) & 1), g0);
      } else if (swap) {
        f += energy_lds_pick<true>(part, lds, slot, a.gslot[g], x, a, g0, g1, tpar, ub);
      } else {
        f += energy_lds_pick<false>(part, lds, slot, a.gslot[g], x, a, g0, g1, tpar, ub);
      }
    }
    e += double(f);
  }
  e = __ockl_wfred_add_f64(e);
  if (lane == 0) red[wave] = e;
  __syncthreads();
  if (threadIdx.x < 2) {
    double t = 0.0;
    if (threadIdx.x == 0)
      for (int w = 0; w < W; w++) t += red[w];
    epart[(size_t(inst) * gridDim.x + blockIdx.x) * 2 + threadIdx.x] = t;
  }
}

__global__ __launch_bounds__(kT) void k_adjoint_rot(AdjArgs a, double *part) {
  __shared__ double sh[kT / 64];
  const uint64_t stride = uint64_t(gridDim.x) * kT;
  double re = 0.0, im = 0.0;
  if (a.q == 0) {
    for (uint64_t u = uint64_t(blockIdx.x) * kT + threadIdx.x; u < a.dim / 2; u += stride) {
      float4 *pp = reinterpret_cast<float4 *>(a.psi + 2 * u);
      float4 *lp = reinterpret_cast<float4 *>(a.lam + 2 * u);
      const float4 pv = *pp, lv = *lp;
      float2 p0 = make_float2(pv.x, pv.y), p1 = make_float2(pv.z, pv.w);
      float2 l0 = make_float2(lv.x, lv.y), l1 = make_float2(lv.z, lv.w);
      adj_pair(a, p0, p1, l0, l1, re, im);
      *pp = make_float4(p0.x, p0.y, p1.x, p1.y);
      *lp = make_float4(l0.x, l0.y, l1.x, l1.y);
    }
  } else {
    const uint32_t q = a.q;
    const uint64_t bit = 1ull << q;
    for (uint64_t u = uint64_t(blockIdx.x) * kT + threadIdx.x; u < a.dim / 4; u += stride) {
      const uint64_t t = ((u >> (q - 1)) << q) | (u & ((1ull << (q - 1)) - 1ull));
      const uint64_t i = t << 1;
      float4 *pa = reinterpret_cast<float4 *>(a.psi + i);
      float4 *pb = reinterpret_cast<float4 *>(a.psi + (i | bit));
      float4 *la = reinterpret_cast<float4 *>(a.lam + i);
      float4 *lb = reinterpret_cast<float4 *>(a.lam + (i | bit));
      const float4 pav = *pa, pbv = *pb, lav = *la, lbv = *lb;
      float2 p0 = make_float2(pav.x, pav.y), p1 = make_float2(pbv.x, pbv.y);
      float2 l0 = make_float2(lav.x, lav.y), l1 = make_float2(lbv.x, lbv.y);
      float2 p2 = make_float2(pav.z, pav.w), p3 = make_float2(pbv.z, pbv.w);
      float2 l2 = make_float2(lav.z, lav.w), l3 = make_float2(lbv.z, lbv.w);
      adj_pair(a, p0, p1, l0, l1, re, im);
      adj_pair(a, p2, p3, l2, l3, re, im);
      *pa = make_float4(p0.x, p0.y, p2.x, p2.y);
      *pb = make_float4(p1.x, p1.y, p3.x, p3.y);
      *la = make_float4(l0.x, l0.y, l2.x, l2.y);
      *lb = make_float4(l1.x, l1.y, l3.x, l3.y);
    }
  }
  const double tr = block_sum(re, sh);
  const double ti = block_sum(im, sh);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = tr;
    part[2 * blockIdx.x + 1] = ti;
  }
}

// --- sampling ---------------------------------------------------------------

// One wave per 1024-amplitude block: each lane sums 16 amplitudes in FP64.
__global__ __launch_bounds__(kT) void k_block_sums(const float2 *a, uint64_t dim, double *sums,
                                                   uint64_t nblocks) {
  const uint64_t waves = uint64_t(gridDim.x) * (kT / 64);
  const int lane = threadIdx.x & 63;
  for (uint64_t b = uint64_t(blockIdx.x) * (kT / 64) + (threadIdx.x >> 6); b < nblocks;
       b += waves) {
    // 16-byte loads, all eight of the lane's in flight (dim is even, so a
    // float4 is either wholly inside the state or wholly past its end).
    const uint64_t start = b * kSampleBlock;
    float4 x[kSampleBlock / 128];
#pragma unroll
    for (int k = 0; k < int(kSampleBlock / 128); k++) {
      const uint64_t i = start + uint64_t(k) * 128 + 2 * uint64_t(lane);
      x[k] = i < dim ? ld4<true>(a + i) : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
    double acc = 0.0;
#pragma unroll
    for (int k = 0; k < int(kSampleBlock / 128); k++)
      acc += double(x[k].x) * x[k].x + double(x[k].y) * x[k].y + double(x[k].z) * x[k].z +
             double(x[k].w) * x[k].w;
    acc = wave_sum(acc);
    if (lane == 0) sums[b] = acc;
  }
}

// Inclusive scan of each 1024-element chunk in place; chunk totals out.
__global__ __launch_bounds__(kT) void k_scan_chunks(double *v, uint64_t count, double *chunk) {
  __shared__ double sh[kT];
  const uint64_t base = uint64_t(blockIdx.x) * 1024;
  double x[4];
  double run = 0.0;
  for (int k = 0; k < 4; k++) {
    const uint64_t i = base + threadIdx.x * 4 + k;
    run += (i < count) ? v[i] : 0.0;
    x[k] = run;
  }
  sh[threadIdx.x] = run;
  __syncthreads();
  // Hillis-Steele over the 256 per-thread totals.
  for (int o = 1; o < kT; o <<= 1) {
    const double add = int(threadIdx.x) >= o ? sh[threadIdx.x - o] : 0.0;
    __syncthreads();
    sh[threadIdx.x] += add;
    __syncthreads();
  }
  const double before = threadIdx.x ? sh[threadIdx.x - 1] : 0.0;
  for (int k = 0; k < 4; k++) {
    const uint64_t i = base + threadIdx.x * 4 + k;
    if (i < count) v[i] = x[k] + before;
  }
  if (threadIdx.x == kT - 1) chunk[blockIdx.x] = sh[kT - 1];
}

// Exclusive prefix of the chunk totals by one workgroup, sequential per
// thread segment.
__global__ __launch_bounds__(kT) void k_scan_totals(double *chunk, uint64_t nchunks) {
  __shared__ double sh[kT];
  const uint64_t per = (nchunks + kT - 1) / kT;
  const uint64_t lo = threadIdx.x * per;
  double s = 0.0;
  for (uint64_t i = lo; i < lo + per && i < nchunks; i++) s += chunk[i];
  sh[threadIdx.x] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double run = 0.0;
    for (int k = 0; k < kT; k++) {
      const double t = sh[k];
      sh[k] = run;
      run += t;
    }
  }
  __syncthreads();
  double run = sh[threadIdx.x];
  for (uint64_t i = lo; i < lo + per && i < nchunks; i++) {
    const double t = chunk[i];
    chunk[i] = run;
    run += t;
  }
}

__global__ __launch_bounds__(kT) void k_add_offsets(double *v, uint64_t count, const double *chunk) {
  const uint64_t stride = uint64_t(gridDim.x) * kT;
  for (uint64_t i = uint64_t(blockIdx.x) * kT + threadIdx.x; i < count; i += stride)
    v[i] += chunk[i / 1024];
}

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// Shot k uses the (k+1)-th splitmix64 output from `seed`: the generator's
// state after k+1 steps is seed + (k+1) * golden, so shots are independent.
__global__ __launch_bounds__(kT) void k_shots(const float2 *a, uint64_t dim, const double *cdf,
                                              uint64_t nblocks, uint64_t shots, uint64_t seed,
                                              uint64_t *out) {
  const uint64_t stride = uint64_t(gridDim.x) * kT;
  const double total = cdf[nblocks - 1];
  for (uint64_t k = uint64_t(blockIdx.x) * kT + threadIdx.x; k < shots; k += stride) {
    if (!(total > 0.0)) {
      out[k] = 0;
      continue;
    }
    const uint64_t r = mix64(seed + (k + 1) * 0x9E3779B97F4A7C15ull);
    const double u = double(r >> 11) * (1.0 / 9007199254740992.0) * total;
    uint64_t lo = 0, hi = nblocks - 1;
    while (lo < hi) {
      const uint64_t mid = (lo + hi) >> 1;
      if (cdf[mid] > u) hi = mid; else lo = mid + 1;
    }
    const double rem = u - (lo > 0 ? cdf[lo - 1] : 0.0);
    const uint64_t start = lo * kSampleBlock;
    const uint64_t end = start + kSampleBlock < dim ? start + kSampleBlock : dim;
    double acc = 0.0;
    int64_t last = -1;
    uint64_t chosen = end - 1;
    bool found = false;
    for (uint64_t i = start; i < end; i++) {
      const float2 x = a[i];
      const double pr = double(x.x) * x.x + double(x.y) * x.y;
      if (pr > 0.0) last = int64_t(i);
      acc += pr;
      if (acc > rem) {
        chosen = i;
        found = true;
        break;
      }
    }
    if (!found) chosen = last >= 0 ? uint64_t(last) : end - 1;
    out[k] = chosen;
  }
}

// Batched adjoint results: one thread per (instance, rotation), plus the
// energy.
__global__ __launch_bounds__(kT) void k_batched_results(const double *eslots, const double *rslots,
                                                        uint32_t count, uint32_t nrot,
                                                        const uint32_t *rgate, const uint32_t *inst,
                                                        const uint32_t *gbase, double *energy_out,
                                                        double *grad_out) {
  const uint64_t per = uint64_t(nrot) + 1;
  const uint64_t stride = uint64_t(gridDim.x) * kT;
  for (uint64_t t = uint64_t(blockIdx.x) * kT + threadIdx.x; t < per * count; t += stride) {
    const uint32_t i = uint32_t(t / per), r = uint32_t(t % per);
    if (r == nrot) {
      energy_out[inst[i]] = eslots[2 * size_t(i)];
      continue;
    }
    const uint32_t g = rgate[r];
    const double *o = rslots + (size_t(i) * nrot + r) * 2;
    grad_out[size_t(gbase[i]) + (g & 0x7fffffffu)] = (g >> 31) ? o[0] : o[1];
  }
}

}  // namespace

// --- launchers --------------------------------------------------------------

void launch_fill_basis(hipStream_t s, float2 *a, uint64_t dim) {
  k_fill_basis<<<grid_for(dim / 2), kT, 0, s>>>(a, dim);
}

void launch_zero(hipStream_t s, float2 *a, uint64_t dim) {
  k_zero<<<grid_for(dim / 2), kT, 0, s>>>(a, dim);
}

void launch_placement_probe(hipStream_t s, float2 *a, uint32_t n, uint32_t p0) {
  if (n < 14 || p0 < 7 || p0 + 7 > n) return;
  Layout L{};
  const uint64_t tile = 0x7Full | (0x7Full << p0);
  for (uint32_t q = 0; q < 64; q++)
    if ((tile >> q) & 1) {
      L.ins[L.nins] = uint8_t(q);
      L.hm[L.nins] = ~((uint64_t(1) << q) - 1);
      L.nins++;
    }
  L.nthreads = (uint64_t(1) << n) >> L.nins;
  const unsigned grid = unsigned(std::min<uint64_t>(L.nthreads, kPassGrid));
  k_placement_probe<<<grid, 512, 0, s>>>(reinterpret_cast<float4 *>(a), L, p0, 1.0f);
}

void launch_pick(hipStream_t s, const float2 *a, const uint64_t *idx, uint64_t count,
                 float2 *out) {
  k_pick<<<grid_for(count), kT, 0, s>>>(a, idx, count, out);
}

template <bool NT>
void launch_gather_t(hipStream_t s, int nl, const GatherArgs &g) {
  const unsigned grid = grid_for(g.lay.nthreads);
  switch (nl) {
    case 1: k_gather<1, NT><<<grid, kT, 0, s>>>(g); break;
    case 2: k_gather<2, NT><<<grid, kT, 0, s>>>(g); break;
    case 3: k_gather<3, NT><<<grid, kT, 0, s>>>(g); break;
    case 4: k_gather<4, NT><<<grid, kT, 0, s>>>(g); break;
    case 5: k_gather<5, NT><<<grid, kT, 0, s>>>(g); break;
    default: break;
  }
}

void launch_gather(hipStream_t s, int nl, const GatherArgs &g) {
  if (g.lay.streaming) launch_gather_t<true>(s, nl, g);
  else launch_gather_t<false>(s, nl, g);
}



template <bool NT>
void launch_diag_t(hipStream_t s, int nl, const DiagArgs &d) {
  const unsigned grid = grid_for(d.lay.nthreads);
  switch (nl) {
    case 1: k_diag<1, NT><<<grid, kT, 0, s>>>(d); break;
    case 2: k_diag<2, NT><<<grid, kT, 0, s>>>(d); break;
    case 3: k_diag<3, NT><<<grid, kT, 0, s>>>(d); break;
    case 4: k_diag<4, NT><<<grid, kT, 0, s>>>(d); break;
    default: break;
  }
}

void launch_diag(hipStream_t s, int nl, const DiagArgs &d) {
  if (d.lay.streaming) launch_diag_t<true>(s, nl, d);
  else launch_diag_t<false>(s, nl, d);
}

template <bool NT>
void launch_pass_t(hipStream_t s, int nl, const PassArgs &p) {
  const unsigned grid = grid_for(p.lay.nthreads, kPassGrid);
  switch (nl) {
    case 4: k_pass<4, NT><<<grid, kT, 0, s>>>(p); break;
    case 5: k_pass<5, NT><<<grid, kT, 0, s>>>(p); break;
    default: break;
  }
}

void launch_pass(hipStream_t s, int nl, const PassArgs &p) {
  if (p.lay.streaming) launch_pass_t<true>(s, nl, p);
  else launch_pass_t<false>(s, nl, p);
}

template <int NL, int K, bool Z>
void launch_chain_k(hipStream_t s, bool nt, bool real, const ChainArgs &c) {
  const unsigned grid = grid_for(c.lay.nthreads, kPassGrid);
  if (nt && real) k_chain<NL, true, K, Z, true><<<grid, kT, 0, s>>>(c);
  else if (nt) k_chain<NL, true, K, Z, false><<<grid, kT, 0, s>>>(c);
  else if (real) k_chain<NL, false, K, Z, true><<<grid, kT, 0, s>>>(c);
  else k_chain<NL, false, K, Z, false><<<grid, kT, 0, s>>>(c);
}

void launch_chain(hipStream_t s, int nl, int k, bool zero, bool real, const ChainArgs &c) {
  const bool nt = c.lay.streaming != 0;
  if (nl == 5 && k == 4) {
    if (zero) launch_chain_k<5, 4, true>(s, nt, real, c);
    else launch_chain_k<5, 4, false>(s, nt, real, c);
  } else if (nl == 4 && k >= 1 && k <= 3) {
    switch (k * 2 + (zero ? 1 : 0)) {
      case 2: launch_chain_k<4, 1, false>(s, nt, real, c); break;
      case 3: launch_chain_k<4, 1, true>(s, nt, real, c); break;
      case 4: launch_chain_k<4, 2, false>(s, nt, real, c); break;
      case 5: launch_chain_k<4, 2, true>(s, nt, real, c); break;
      case 6: launch_chain_k<4, 3, false>(s, nt, real, c); break;
      case 7: launch_chain_k<4, 3, true>(s, nt, real, c); break;
      default: break;
    }
  }
}

template <int WB>
void launch_wide_wb(hipStream_t s, int mode, const WideArgs &w) {
  const unsigned grid = wide_grid(w.lay.nthreads);
  if (mode == 2) k_wide<WB, 2><<<grid, 64 << WB, 0, s>>>(w);
  else if (mode == 1) k_wide<WB, 1><<<grid, 64 << WB, 0, s>>>(w);
  else k_wide<WB, 0><<<grid, 64 << WB, 0, s>>>(w);
}

void launch_wide(hipStream_t s, int steps, int mode, const WideArgs &w) {
  if (steps == 7) launch_wide_wb<3>(s, mode, w);
  else launch_wide_wb<2>(s, mode, w);
}

// Workgroups of a low pass: enough to fill the chip many times over, few
// enough that the one spare prefetch per workgroup stays negligible.

void launch_low(hipStream_t s, int steps, bool zero, int mode, int wb, const LowArgs &l) {
  // The low chain kernels are built in low_kernels.hip, a pair of step
  // counts per object (they take most of the build).
  switch (steps) {
    case 5: case 6: launch_low_pair<5>(s, steps, zero, mode, wb, l); break;
    case 7: case 8: launch_low_pair<7>(s, steps, zero, mode, wb, l); break;
    case 9: case 10: launch_low_pair<9>(s, steps, zero, mode, wb, l); break;
    case 11: case 12: launch_low_pair<11>(s, steps, zero, mode, wb, l); break;
    default: break;
  }
}

void launch_gwide(hipStream_t s, int wb, const GWideArgs &g) {
  const unsigned grid = wide_grid(g.lay.nthreads);
  if (wb == 3) k_gwide<3><<<grid, 512, 0, s>>>(g);
  else k_gwide<2><<<grid, 256, 0, s>>>(g);
}

void launch_tile(hipStream_t s, const TileArgs &t) {
  const unsigned grid = unsigned(std::min<uint64_t>(t.lay.nthreads, kLowGrid));
  k_tile<<<grid, 128, 0, s>>>(t);
}

void launch_diag_generic(hipStream_t s, const DiagGenericArgs &d) {
  k_diag_generic<<<grid_for(d.dim), kT, 0, s>>>(d);
}

void launch_blocked(hipStream_t s, const BlockedArgs &b) {
  const uint64_t g = b.nblocks < kMaxGrid ? b.nblocks : kMaxGrid;
  k_blocked<<<unsigned(g), kT, sizeof(float2) << b.bits, s>>>(b);
}

// Workgroups of the streaming reductions: enough for one loop step per
// thread (k_abs2: kRedUnroll float4, k_pauli: two float4 of each state),
// at most kRedBlocks; narrow states reduce fewer partials.
void launch_abs2(hipStream_t s, const float2 *a, uint64_t dim, double *partials,
                 double *out) {
  const unsigned blocks = grid_for((dim / 2 + kRedUnroll - 1) / kRedUnroll, kRedBlocks);
  k_abs2<<<blocks, kT, 0, s>>>(a, dim, partials);
  k_finish<<<1, kT, 0, s>>>(partials, int(blocks), 1, out);
}

void launch_pauli(hipStream_t s, const float2 *bra, const float2 *ket,
                  uint64_t dim, uint64_t flip, uint64_t sign, double *partials,
                  double *out) {
  const unsigned blocks = grid_for((dim + 3) / 4, kRedBlocks);
  k_pauli<<<blocks, kT, 0, s>>>(bra, ket, dim, flip, sign, partials);
  k_finish<<<1, kT, 0, s>>>(partials, int(blocks), 2, out);
}

unsigned pauli_terms_blocks(uint64_t dim) { return grid_for((dim + 3) / 4, kRedBlocks); }

void launch_pauli_terms(hipStream_t s, const float2 *a, uint64_t dim, const PauliTerms &t,
                        uint32_t count, double *partials, double *out) {
  if (!count) return;
  const unsigned blocks = pauli_terms_blocks(dim);
  k_pauli_terms<<<dim3(blocks, count), kT, 0, s>>>(a, dim, t, partials);
  k_finish<<<count, kT, 0, s>>>(partials, int(blocks), 2, out);
}

void launch_pauli_accum(hipStream_t s, const AccumArgs &a) {
  k_pauli_accum<<<grid_for(a.dim / 2), kT, 0, s>>>(a);
}

void launch_pauli_accum_tile(hipStream_t s, const AccumTileArgs &a) {
  const unsigned grid = grid_for(a.dim >> (a.nw + 1));
  switch (a.nw) {
    case 0: k_pauli_accum_tile<0><<<grid, kT, 0, s>>>(a); break;
    case 1: k_pauli_accum_tile<1><<<grid, kT, 0, s>>>(a); break;
    case 2: k_pauli_accum_tile<2><<<grid, kT, 0, s>>>(a); break;
    case 3: k_pauli_accum_tile<3><<<grid, kT, 0, s>>>(a); break;
    default: break;
  }
}

void launch_adjoint_rot(hipStream_t s, const AdjArgs &a, double *partials,
                        double *out2) {
  k_adjoint_rot<<<kRedBlocks, kT, 0, s>>>(a, partials);
  k_finish<<<1, kT, 0, s>>>(partials, kRedBlocks, 2, out2);
}

template <int NL>
void adj_pass_nl(dim3 grid, hipStream_t s, const AdjPassArgs &a, double *partials) {
  if (a.lay.streaming) k_adj_pass<NL, true><<<grid, kT, 0, s>>>(a, partials);
  else k_adj_pass<NL, false><<<grid, kT, 0, s>>>(a, partials);
}

void launch_adj_pass(hipStream_t s, const AdjPassArgs &a, double *partials, double *out) {
  adj_pass_nl<4>(dim3(kAdjBlocks), s, a, partials);
  if (a.nrot) k_finish_rot<<<a.nrot, kT, 0, s>>>(partials, kAdjBlocks, out, 0, kAdjMaxRot);
}

void launch_adj_tile(hipStream_t s, const AdjTileArgs &a, double *partials, double *out) {
  const unsigned blocks = unsigned(std::min<uint64_t>(a.lay.nthreads, kAdjTileBlocks));
  k_adj_tile<false><<<blocks, kAdjTileThreads, 0, s>>>(a, partials);
  if (a.nrot) k_finish_rot<<<a.nrot, kT, 0, s>>>(partials, int(blocks), out, 0, kAdjTileMaxRot);
}

void launch_fwd_tile(hipStream_t s, const AdjTileArgs &a) {
  const unsigned blocks = unsigned(std::min<uint64_t>(a.lay.nthreads, kAdjTileBlocks));
  k_adj_tile<true><<<blocks, kAdjTileThreads, 0, s>>>(a, nullptr);
}

// One thread per patch (AnglePatch).
__global__ __launch_bounds__(kT) void k_patch_angles(uint8_t *arena, const AnglePatch *p,
                                                     uint32_t count, const double *params) {
  const uint32_t i = blockIdx.x * kT + threadIdx.x;
  if (i >= count) return;
  const AnglePatch q = p[i];
  const double t = params[q.param];
  const double a = (q.code & 256u) ? -t : t;
  double sn, cs;
  sincos(0.5 * a, &sn, &cs);
  // d0 = exp(-i a / 2), d1 = exp(i a / 2)
  const double d0r = cs, d0i = -sn, d1r = cs, d1i = sn;
  double m[8], ma[8];
  for (int k = 0; k < 8; k++) m[k] = 0.0;
  switch (q.code & 15u) {
    case 0: {
      const uint32_t kind = (q.code >> 4) & 15u;
      if (kind == 0) {         // RX: [[c, -i s], [-i s, c]]
        m[0] = cs; m[3] = -sn; m[5] = -sn; m[6] = cs;
      } else if (kind == 1) {  // RY: [[c, -s], [s, c]]
        m[0] = cs; m[2] = -sn; m[4] = sn; m[6] = cs;
      } else {                 // RZ: diag(d0, d1)
        m[0] = d0r; m[1] = d0i; m[6] = d1r; m[7] = d1i;
      }
      for (int k = 0; k < 8; k++) ma[k] = m[k];
      break;
    }
    case 1:
      for (int k = 0; k < 8; k++) ma[k] = 0.0;
      m[0] = m[6] = d0r; m[1] = m[7] = d0i;
      ma[0] = ma[6] = d1r; ma[1] = ma[7] = d1i;
      break;
    case 2:
      for (int k = 0; k < 8; k++) ma[k] = 0.0;
      m[0] = d0r; m[1] = d0i; m[6] = d1r; m[7] = d1i;
      ma[0] = d1r; ma[1] = d1i; ma[6] = d0r; ma[7] = d0i;
      break;
    default: {
      // diag(d1 conj(d0), d0 conj(d1)) = diag(exp(i a), exp(-i a))
      double s2, c2;
      sincos(a, &s2, &c2);
      m[0] = c2; m[1] = s2; m[6] = c2; m[7] = -s2;
      for (int k = 0; k < 8; k++) ma[k] = m[k];
      break;
    }
  }
  if (q.code & kPatchPassOp) {
    float *pm = reinterpret_cast<float *>(arena + q.off);
    for (int k = 0; k < 8; k++) pm[k] = float(m[k]);
    return;
  }
  AdjOp *o = reinterpret_cast<AdjOp *>(arena + q.off);
  for (int k = 0; k < 8; k++) {
    o->m[k] = float(m[k]);
    o->malt[k] = float(ma[k]);
  }
}

void launch_patch_angles(hipStream_t s, uint8_t *arena, const AnglePatch *p, uint32_t count,
                         const double *params) {
  if (count) k_patch_angles<<<(count + kT - 1) / kT, kT, 0, s>>>(arena, p, count, params);
}

unsigned costate_tile_blocks(const Layout &lay) {
  return unsigned(std::min<uint64_t>(lay.nthreads, kAdjTileBlocks));
}

void launch_costate_tile(hipStream_t s, const CoTileArgs &a, double *partials, double *energy) {
  const unsigned blocks = costate_tile_blocks(a.lay);
  if (!a.lam) {
    if (!energy) return;  // nothing to produce
    k_costate_energy<<<blocks, kAdjTileThreads, 0, s>>>(a, partials);
    k_finish<<<1, kT, 0, s>>>(partials, int(blocks), 2, energy);
    return;
  }
  k_costate_tile<<<blocks, kAdjTileThreads, 0, s>>>(a, energy ? partials : nullptr);
  if (energy) k_finish<<<1, kT, 0, s>>>(partials, int(blocks), 2, energy);
}

void launch_sum_slots(hipStream_t s, const double *slots, uint32_t count, double *out) {
  k_finish<<<1, kT, 0, s>>>(slots, int(count), 2, out);
}

void launch_costate_tile_batched(hipStream_t s, const CoTileArgs &a, double *partials,
                                 double *energy, uint32_t count) {
  const unsigned blocks = costate_tile_blocks(a.lay);
  k_costate_tile<<<dim3(blocks, count), kAdjTileThreads, 0, s>>>(a, energy ? partials : nullptr);
  if (energy) k_finish<<<count, kT, 0, s>>>(partials, int(blocks), 2, energy);
}

void launch_block_sums(hipStream_t s, const float2 *a, uint64_t dim, double *sums) {
  const uint64_t nb = (dim + kSampleBlock - 1) / kSampleBlock;
  k_block_sums<<<grid_for(nb * 64), kT, 0, s>>>(a, dim, sums, nb);
}

void launch_prefix(hipStream_t s, double *v, uint64_t count, double *chunk) {
  const uint64_t nchunks = (count + 1023) / 1024;
  k_scan_chunks<<<unsigned(nchunks), kT, 0, s>>>(v, count, chunk);
  if (nchunks > 1) {
    k_scan_totals<<<1, kT, 0, s>>>(chunk, nchunks);
    k_add_offsets<<<grid_for(count), kT, 0, s>>>(v, count, chunk);
  }
}

void launch_shots(hipStream_t s, const float2 *a, uint64_t dim,
                  const double *cdf, uint64_t nblocks, uint64_t shots,
                  uint64_t seed, uint64_t *out) {
  k_shots<<<grid_for(shots), kT, 0, s>>>(a, dim, cdf, nblocks, shots, seed, out);
}

// --- batched launches ------------------------------------------------------

void launch_pass_batched(hipStream_t s, int nl, const PassArgs &p, uint32_t count) {
  const dim3 grid(grid_for(p.lay.nthreads, kPassGrid), count);
  const bool nt = p.lay.streaming != 0;
  if (nl == 4) {
    if (nt) k_pass<4, true><<<grid, kT, 0, s>>>(p);
    else k_pass<4, false><<<grid, kT, 0, s>>>(p);
  } else if (nl == 5) {
    if (nt) k_pass<5, true><<<grid, kT, 0, s>>>(p);
    else k_pass<5, false><<<grid, kT, 0, s>>>(p);
  }
}

void launch_tile_batched(hipStream_t s, const TileArgs &t, uint32_t count) {
  const dim3 grid(unsigned(std::min<uint64_t>(t.lay.nthreads, kLowGrid)), count);
  k_tile<<<grid, 128, 0, s>>>(t);
}

void launch_gwide_batched(hipStream_t s, int wb, const GWideArgs &g, uint32_t count) {
  const dim3 grid(wide_grid(g.lay.nthreads), count);
  if (wb == 3) k_gwide<3><<<grid, 512, 0, s>>>(g);
  else k_gwide<2><<<grid, 256, 0, s>>>(g);
}

void launch_fill_basis_batched(hipStream_t s, float2 *a, uint64_t dim, uint32_t count) {
  k_fill_basis<<<dim3(grid_for(dim / 2), count), kT, 0, s>>>(a, dim);
}

void launch_zero_batched(hipStream_t s, float2 *a, uint64_t dim, uint32_t count) {
  k_zero<<<dim3(grid_for(dim / 2), count), kT, 0, s>>>(a, dim);
}

void launch_pauli_accum_tile_batched(hipStream_t s, const AccumTileArgs &a, uint32_t count) {
  const dim3 grid(grid_for(a.dim >> (a.nw + 1)), count);
  switch (a.nw) {
    case 0: k_pauli_accum_tile<0><<<grid, kT, 0, s>>>(a); break;
    case 1: k_pauli_accum_tile<1><<<grid, kT, 0, s>>>(a); break;
    case 2: k_pauli_accum_tile<2><<<grid, kT, 0, s>>>(a); break;
    case 3: k_pauli_accum_tile<3><<<grid, kT, 0, s>>>(a); break;
    default: break;
  }
}

void launch_pauli_accum_batched(hipStream_t s, const AccumArgs &a, uint32_t count) {
  k_pauli_accum<<<dim3(grid_for(a.dim / 2), count), kT, 0, s>>>(a);
}

unsigned pauli_batched_blocks(uint64_t dim) { return grid_for((dim + 3) / 4, kRedBlocks); }

unsigned adj_batched_blocks(const Layout &lay) { return grid_for(lay.nthreads, kAdjBlocks); }

unsigned adj_tile_batched_blocks(const Layout &lay) {
  return unsigned(std::min<uint64_t>(lay.nthreads, kAdjTileBlocks));
}

void launch_adj_tile_batched(hipStream_t s, const AdjTileArgs &a, double *partials, double *out,
                             uint64_t ostride, uint32_t count) {
  const unsigned blocks = adj_tile_batched_blocks(a.lay);
  k_adj_tile<false><<<dim3(blocks, count), kAdjTileThreads, 0, s>>>(a, partials);
  if (a.nrot)
    k_finish_rot<<<dim3(a.nrot, count), kT, 0, s>>>(partials, int(blocks), out, ostride,
                                                   kAdjTileMaxRot);
}

void launch_pauli_batched(hipStream_t s, const float2 *bra, const float2 *ket, uint64_t dim,
                          uint64_t flip, uint64_t sign, double *partials, double *out,
                          uint32_t count) {
  const unsigned blocks = pauli_batched_blocks(dim);
  k_pauli<<<dim3(blocks, count), kT, 0, s>>>(bra, ket, dim, flip, sign, partials);
  k_finish<<<count, kT, 0, s>>>(partials, int(blocks), 2, out);
}

void launch_adj_pass_batched(hipStream_t s, const AdjPassArgs &a, double *partials, double *out,
                             uint64_t ostride, uint32_t count) {
  const unsigned blocks = adj_batched_blocks(a.lay);
  const dim3 grid(blocks, count);
  adj_pass_nl<4>(grid, s, a, partials);
  if (a.nrot)
    k_finish_rot<<<dim3(a.nrot, count), kT, 0, s>>>(partials, int(blocks), out, ostride,
                                                   kAdjMaxRot);
}

void launch_batched_results(hipStream_t s, const double *eslots, const double *rslots,
                            uint32_t count, uint32_t nrot, const uint32_t *rgate,
                            const uint32_t *inst, const uint32_t *gbase, double *energy_out,
                            double *grad_out) {
  k_batched_results<<<grid_for(uint64_t(count) * (nrot + 1)), kT, 0, s>>>(
      eslots, rslots, count, nrot, rgate, inst, gbase, energy_out, grad_out);
}

}  // namespace qmamd
