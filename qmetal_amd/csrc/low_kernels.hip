// low_kernels.hip -- the low chain pass kernels (k_low) and their launchers.
// Built once per pair of step counts (QM_LOW_S0 = 5, 7, 9, 11; both counts of
// a pair in one object) so that the 128 kernel instances compile in parallel;
// without QM_LOW_S0 the file holds every pair.

#include <hip/hip_runtime.h>

#include <algorithm>

#include "kernel_util.h"
#include "kernels.h"

namespace qmamd {
namespace {

// LDS slot of tile index i (float4 units, up to 12 bits): the low four bits are
// xored with the higher ones so that lanes stepping through high bits still
// spread over the banks (a bijection: the high bits are kept; linear over
// GF(2)).
__device__ __forceinline__ uint32_t low_swz(uint32_t i) { return tile_swizzle(i); }

// Tile index (float4 units: tile bit p is index bit p - 1) of this thread's
// lane and wave bits in arrangement `a`.
template <int WB>
__device__ __forceinline__ uint32_t low_thread_index(const LowArgs &p, int a, uint32_t lane,
                                                     uint32_t wave) {
  uint32_t i = 0;
#pragma unroll
  for (int j = 0; j < 6; j++)
    if ((lane >> j) & 1) i |= 1u << (p.lanepos[a][j] - 1);
#pragma unroll
  for (int k = 0; k < WB; k++)
    if ((wave >> k) & 1) i |= 1u << (p.lanepos[a][6 + k] - 1);
  return i;
}

// Moves the tile from arrangement FROM to arrangement TO through LDS. The
// swizzle is linear and thread and register bits are disjoint, so a slot is
// the thread's swizzled index xor the xor of the swizzled register bits
// (workgroup-uniform). Byte offsets are kept opaque so that the 32 slot
// addresses are formed here (one xor each) instead of being hoisted out of
// the tile loop into registers.
template <int FROM, int TO, int WB>
__device__ __forceinline__ void low_permute(const LowArgs &p, float4 *lds, float2 (&x)[32],
                                            uint32_t lane, uint32_t wave) {
  uint32_t lf = low_swz(low_thread_index<WB>(p, FROM, lane, wave)) << 4;
  uint32_t lt = low_swz(low_thread_index<WB>(p, TO, lane, wave)) << 4;
  uint32_t cf[4], ct[4];
#pragma unroll
  for (int k = 0; k < 4; k++) {
    cf[k] = low_swz(1u << (p.regpos[FROM][k] - 1)) << 4;
    ct[k] = low_swz(1u << (p.regpos[TO][k] - 1)) << 4;
  }
  asm volatile("" : "+v"(lf), "+v"(lt) : "v"(x[0].x));
  asm volatile("" : "+s"(cf[0]), "+s"(cf[1]), "+s"(cf[2]), "+s"(cf[3]), "+s"(ct[0]), "+s"(ct[1]),
               "+s"(ct[2]), "+s"(ct[3]));
  char *base = reinterpret_cast<char *>(lds);
  __syncthreads();  // earlier reads of the buffer are done
#pragma unroll
  for (int r = 0; r < 16; r++) {
    const uint32_t o = ((r & 1) ? cf[0] : 0u) ^ ((r & 2) ? cf[1] : 0u) ^ ((r & 4) ? cf[2] : 0u) ^
                       ((r & 8) ? cf[3] : 0u);
    *reinterpret_cast<float4 *>(base + (lf ^ o)) =
        make_float4(x[2 * r].x, x[2 * r].y, x[2 * r + 1].x, x[2 * r + 1].y);
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < 16; r++) {
    const uint32_t o = ((r & 1) ? ct[0] : 0u) ^ ((r & 2) ? ct[1] : 0u) ^ ((r & 4) ? ct[2] : 0u) ^
                       ((r & 8) ? ct[3] : 0u);
    const float4 v = *reinterpret_cast<const float4 *>(base + (lt ^ o));
    x[2 * r] = make_float2(v.x, v.y);
    x[2 * r + 1] = make_float2(v.z, v.w);
  }
}

// Other-bit index c of a step on local bit BIT (the four other local bits,
// ascending) -> the c bits its phase ignores in a triangular chain: the
// register bits below BIT, which hold the wires of this arrangement's earlier
// steps, and for the wire-0 step (BIT 0) all four.
template <int BIT>
constexpr uint32_t tri_ignored() {
  uint32_t m = 0;
  for (int j = 0; j < 4; j++) {
    const int q = BIT == 0 ? j + 1 : (j < BIT ? j : j + 1);
    if (BIT == 0 || (q >= 1 && q < BIT)) m |= 1u << j;
  }
  return m;
}

// The phase of a chain step in a triangular chain (descending wires, each
// step's phase reading only lower wires, as in the QFT): the factor of
// combination c is e * hi * lo over the bits the phase reads, formed once per
// distinct value, then one complex product per amplitude (against two in
// chain_step). The host checks the step tables before choosing it.
template <int BIT>
__device__ __forceinline__ void chain_phase_tri(float2 (&x)[32],
                                                const ChainStep __attribute__((address_space(4))) &st,
                                                float2 e) {
  constexpr uint32_t TM = tri_ignored<BIT>();
  float2 ah[4], f[16];
  ah[0] = e;
#pragma unroll
  for (int h = 1; h < 4; h++)
    if (!((uint32_t(h) << 2) & TM)) ah[h] = cmul(e, make_float2(st.hi[2 * h - 2], st.hi[2 * h - 1]));
#pragma unroll
  for (int c = 0; c < 16; c++) {
    if (uint32_t(c) & TM) continue;
    const int lo = c & 3, hi = c >> 2;
    f[c] = lo ? cmul(ah[hi], make_float2(st.lo[2 * lo - 2], st.lo[2 * lo - 1])) : ah[hi];
  }
#pragma unroll
  for (int l = 0; l < 32; l++) {
    if (!(l & (1 << BIT))) continue;
    const uint32_t c = uint32_t((l & ((1 << BIT) - 1)) | ((l >> (BIT + 1)) << BIT));
    cmul_v(x[l], f[c & ~TM]);
  }
}

template <int BIT, int M, bool TRI>
__device__ __forceinline__ void low_step(float2 (&x)[32],
                                         const ChainStep __attribute__((address_space(4))) &st,
                                         float2 e) {
  if constexpr (TRI) {
    chain_mix<5, BIT, M>(x, st);
    chain_phase_tri<BIT>(x, st, e);
  } else {
    chain_step<5, BIT, M>(x, st, e);
  }
}

// Steps I.. of a low chain pass as straight-line code (template recursion,
// so that every step index is a compile-time constant however long the
// chain): step i runs in arrangement 1 + i / 4 on register bit 1 + i % 4,
// the wire-0 step (i == S) on bit 0 of the last arrangement.
template <int I, int S, int T, int MODE, int WB, bool TRI>
__device__ __forceinline__ void low_steps(const LowArgs &p, float4 *lds, float2 (&x)[32],
                                          const float2 (&e)[T], const_ptr<ChainStep> steps,
                                          uint32_t lane, uint32_t wave) {
  if constexpr (I < T) {
    if constexpr (I == 4 && S > 4) low_permute<1, 2, WB>(p, lds, x, lane, wave);
    if constexpr (I == 8 && S > 8) low_permute<2, 3, WB>(p, lds, x, lane, wave);
    const ChainStep __attribute__((address_space(4))) &st = steps[opaque_step(I, x[0].x)];
    constexpr int M = (MODE == 2 && I == T - 1) ? 1 : MODE;
    if constexpr (I >= S) low_step<0, M, TRI>(x, st, e[I]);
    else low_step<1 + I % 4, M, TRI>(x, st, e[I]);
    low_steps<I + 1, S, T, MODE, WB, TRI>(p, lds, x, e, steps, lane, wave);
  }
}

// Low chain pass (LowArgs): 2^WB waves per workgroup tile of 2^(11+WB)
// amplitudes, 2^(14+WB) bytes of LDS. The next tile's loads are issued
// before the current tile is worked on. TRI: triangular phase tables
// (chain_phase_tri).
template <int S, bool Z, int MODE, int WB, bool TRI>
__global__ __launch_bounds__(64 << WB) void k_low(LowArgs p) {
  constexpr int T = S + (Z ? 1 : 0);
  constexpr int A = (S + 3) / 4;  // step arrangements
  static_assert(S >= 1 && A + 1 <= kLowMaxArr, "low chain shape");
  __shared__ float4 lds[1024 << WB];
  const uint32_t lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const const_ptr<ChainStep> steps = as_const(p.steps);
  const const_ptr<DiagTerm> uni = as_const(p.uni);
  const const_ptr<DiagTerm> lane_terms = as_const(p.lane);
  // Each step's factor exp(2 pi i beta): lane terms read only thread bits,
  // which are the same for every tile, so without workgroup-uniform terms
  // the factors are computed once.
  auto lane_beta = [&](double (&b)[T]) {
#pragma unroll
    for (int i = 0; i < T; i++) b[i] = steps[i].beta0;
    for (uint32_t k = 0; k < p.nlane; k++) {
      const uint64_t mask = lane_terms[k].mask;
      const uint32_t st = lane_terms[k].k;
      const int a = 1 + (st < uint32_t(S) ? int(st) / 4 : A - 1);
      uint64_t bs = 0;
#pragma unroll
      for (int v = 0; v < WB; v++)
        if ((wave >> v) & 1) bs |= 1ull << p.lanepos[a][6 + v];
#pragma unroll
      for (int j = 0; j < 6; j++)
        if ((lane >> j) & 1) bs |= 1ull << p.lanepos[a][j];
      const double add = (bs & mask) == mask ? double(lane_terms[k].ang) : 0.0;
#pragma unroll
      for (int i = 0; i < T; i++)
        if (st == uint32_t(i)) b[i] += add;
    }
  };
  float2 e[T];
  {
    double b[T];
    lane_beta(b);
#pragma unroll
    for (int i = 0; i < T; i++) {
      float sn, cs;
      turn_sincos(b[i], sn, cs);
      e[i] = make_float2(cs, sn);
    }
  }
  // Arrangement 0: lanes on bits 1..6 (1 KB runs), wave bit k on bit
  // lanepos[0][6 + k], register bit k on bit regpos[0][k].
  uint64_t rbit[4];
#pragma unroll
  for (int k = 0; k < 4; k++) rbit[k] = 1ull << p.regpos[0][k];
  auto roff = [&](int r) {
    return ((r & 1) ? rbit[0] : 0ull) | ((r & 2) ? rbit[1] : 0ull) | ((r & 4) ? rbit[2] : 0ull) |
           ((r & 8) ? rbit[3] : 0ull);
  };
  const uint32_t lane_bytes = lane * 16u;
  uint64_t wave_off = 0;
#pragma unroll
  for (int k = 0; k < WB; k++)
    if ((wave >> k) & 1) wave_off |= 1ull << p.lanepos[0][6 + k];
  float4 y[16];
  if (blockIdx.x < p.lay.nthreads) {
    const uint64_t ub0 = scatter(blockIdx.x, p.lay) | p.lay.fixed | wave_off;
#pragma unroll
    for (int r = 0; r < 16; r++)
      y[r] = ld4<true>(pair_ptr(p.a, ub0 | roff(r), lane_bytes));
  }
  for (uint64_t tile = blockIdx.x; tile < p.lay.nthreads; tile += gridDim.x) {
    const uint64_t ub = scatter(tile, p.lay) | p.lay.fixed;
    float2 x[32];
#pragma unroll
    for (int r = 0; r < 16; r++) {
      x[2 * r] = make_float2(y[r].x, y[r].y);
      x[2 * r + 1] = make_float2(y[r].z, y[r].w);
    }
    if (tile + gridDim.x < p.lay.nthreads) {
      const uint64_t ubn = scatter(tile + gridDim.x, p.lay) | p.lay.fixed | wave_off;
#pragma unroll
      for (int r = 0; r < 16; r++)
        y[r] = ld4<true>(pair_ptr(p.a, ubn | roff(r), lane_bytes));
    }
    if (p.nuni) {
      // Workgroup-uniform terms, one per lane (branch-free: lanes past the
      // end read the last term and add zero).
      double bu[T];
#pragma unroll
      for (int i = 0; i < T; i++) bu[i] = 0.0;
      for (uint32_t k0 = 0; k0 < p.nuni; k0 += 64) {
        const uint32_t k = k0 + lane;
        const uint32_t kc = k < p.nuni ? k : p.nuni - 1;
        const uint64_t mask = uni[kc].mask;
        const double add = (k < p.nuni && (ub & mask) == mask) ? double(uni[kc].ang) : 0.0;
        const uint32_t st = uni[kc].k;
#pragma unroll
        for (int i = 0; i < T; i++) bu[i] += st == uint32_t(i) ? add : 0.0;
      }
      double b[T];
      lane_beta(b);
#pragma unroll
      for (int i = 0; i < T; i++) {
        float sn, cs;
        turn_sincos(b[i] + wave_allsum(bu[i]), sn, cs);
        e[i] = make_float2(cs, sn);
      }
    }
    if (!p.first_in_place) low_permute<0, 1, WB>(p, lds, x, lane, wave);
    low_steps<0, S, T, MODE, WB, TRI>(p, lds, x, e, steps, lane, wave);
    low_permute<A, 0, WB>(p, lds, x, lane, wave);
#pragma unroll
    for (int r = 0; r < 16; r++)
      st4<true>(pair_ptr(p.a, ub | wave_off | roff(r), lane_bytes),
                make_float4(x[2 * r].x, x[2 * r].y, x[2 * r + 1].x, x[2 * r + 1].y));
  }
}

}  // namespace

namespace {

template <int S, int WB>
void launch_low_s(hipStream_t s, bool zero, int mode, const LowArgs &l) {
  // Four tiles per workgroup (the set-up evaluates every step's lane phase
  // terms), at most kLowGrid workgroups' worth of waves.
  const unsigned grid = tile_grid(l.lay.nthreads, 4, kLowGrid >> (WB - 1));
  constexpr int threads = 64 << WB;
  if (zero) {
    if (mode == kLowModeTri) k_low<S, true, 2, WB, true><<<grid, threads, 0, s>>>(l);
    else if (mode == 2) k_low<S, true, 2, WB, false><<<grid, threads, 0, s>>>(l);
    else if (mode == 1) k_low<S, true, 1, WB, false><<<grid, threads, 0, s>>>(l);
    else k_low<S, true, 0, WB, false><<<grid, threads, 0, s>>>(l);
  } else {
    if (mode == kLowModeTri) k_low<S, false, 2, WB, true><<<grid, threads, 0, s>>>(l);
    else if (mode == 2) k_low<S, false, 2, WB, false><<<grid, threads, 0, s>>>(l);
    else if (mode == 1) k_low<S, false, 1, WB, false><<<grid, threads, 0, s>>>(l);
    else k_low<S, false, 0, WB, false><<<grid, threads, 0, s>>>(l);
  }
}

template <int S>
void launch_low_w(hipStream_t s, bool zero, int mode, int wb, const LowArgs &l) {
  if (wb == 2) launch_low_s<S, 2>(s, zero, mode, l);
  else launch_low_s<S, 1>(s, zero, mode, l);
}

}  // namespace

template <int S0>
void launch_low_pair(hipStream_t s, int steps, bool zero, int mode, int wb, const LowArgs &l) {
  if (steps == S0) launch_low_w<S0>(s, zero, mode, wb, l);
  else launch_low_w<S0 + 1>(s, zero, mode, wb, l);
}

#ifdef QM_LOW_S0
template void launch_low_pair<QM_LOW_S0>(hipStream_t, int, bool, int, int, const LowArgs &);
#else
template void launch_low_pair<5>(hipStream_t, int, bool, int, int, const LowArgs &);
template void launch_low_pair<7>(hipStream_t, int, bool, int, int, const LowArgs &);
template void launch_low_pair<9>(hipStream_t, int, bool, int, int, const LowArgs &);
template void launch_low_pair<11>(hipStream_t, int, bool, int, int, const LowArgs &);
#endif

}  // namespace qmamd
