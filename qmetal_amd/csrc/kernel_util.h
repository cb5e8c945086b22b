// kernel_util.h -- device helpers shared by the kernel translation units
// (kernels.hip, low_kernels.hip): launch geometry, index layout, 16-byte
// loads and stores, in-place complex arithmetic, phase evaluation and the
// chain step. Everything has internal linkage; each unit keeps its own copy.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>

#include "kernels.h"

namespace qmamd {
namespace {

constexpr int kT = 256;             // threads per workgroup
constexpr uint64_t kMaxGrid = 1u << 20;

// Workgroups of a merged pass: enough to fill the chip many times over;
// more would only repeat the per-workgroup set-up (the lane phase terms, the
// step constants) for fewer tiles each.
constexpr uint64_t kPassGrid = 1u << 16;

// Workgroups of a pass over `tiles` tiles: `per` tiles per workgroup, so
// that each workgroup's set-up is shared (QFT n=30: 3.57 -> 3.09 ms per wide
// pass at two), but one per workgroup while that leaves fewer than
// kFillGrid, and at most `cap`.
constexpr uint64_t kFillGrid = 1024;
inline unsigned tile_grid(uint64_t tiles, uint64_t per, uint64_t cap) {
  uint64_t g = tiles / per;
  const uint64_t fill = tiles < kFillGrid ? tiles : kFillGrid;
  if (g < fill) g = fill;
  if (g < 1) g = 1;
  if (g > cap) g = cap;
  return unsigned(g);
}
inline unsigned wide_grid(uint64_t tiles) { return tile_grid(tiles, 2, kPassGrid); }

inline unsigned grid_for(uint64_t work, uint64_t cap = kMaxGrid) {
  uint64_t g = (work + kT - 1) / kT;
  if (g < 1) g = 1;
  if (g > cap) g = cap;
  return unsigned(g);
}

// Thread index -> base index without the fixed bits: zero bits inserted at
// the layout's positions, lowest first.
__device__ __forceinline__ uint64_t scatter(uint64_t t, const Layout &L) {
  for (uint32_t k = 0; k < L.nins; k++) t += t & L.hm[k];
  return t;
}

// Base index of thread-index value t0 + tid with t0 a multiple of the
// workgroup size: scatter is a bit placement, so the workgroup-uniform part
// scatter(t0) | fixed and the per-thread part scatter(tid) are disjoint. The
// per-thread part stays below 2^24 (the workgroup's index bits land on the
// lowest free positions), so a thread addresses its amplitudes as a uniform
// 64-bit pointer plus one 32-bit byte offset.
struct TileAddr {
  uint32_t lane_bytes;  // 8 * scatter(tid)
  uint64_t lane_bits;   // scatter(tid)
};

__device__ __forceinline__ TileAddr tile_addr(const Layout &L) {
  TileAddr a;
  a.lane_bits = scatter(threadIdx.x, L);
  a.lane_bytes = uint32_t(a.lane_bits) * 8u;
  return a;
}

// Amplitude pair (ub | o) + the thread's part, as a 16-byte address.
__device__ __forceinline__ const float2 *pair_ptr(const float2 *a, uint64_t ubo,
                                                  uint32_t lane_bytes) {
  return reinterpret_cast<const float2 *>(reinterpret_cast<const char *>(a + ubo) + lane_bytes);
}
__device__ __forceinline__ float2 *pair_ptr(float2 *a, uint64_t ubo, uint32_t lane_bytes) {
  return reinterpret_cast<float2 *>(reinterpret_cast<char *>(a + ubo) + lane_bytes);
}

template <int NL>
__device__ __forceinline__ void offsets(const Layout &L, uint64_t (&off)[1 << (NL - 1)]) {
#pragma unroll
  for (int r = 0; r < (1 << (NL - 1)); r++) {
    uint64_t o = 0;
#pragma unroll
    for (int k = 0; k < NL - 1; k++)
      if ((r >> k) & 1) o |= 1ull << L.rpos[k];
    off[r] = o;
  }
}

__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// x' = m00 x + m01 y, y' = m10 x + m11 y
__device__ __forceinline__ void rot(const float *m, float2 &x, float2 &y) {
  const float2 u = x, v = y;
  x.x = m[0] * u.x - m[1] * u.y + m[2] * v.x - m[3] * v.y;
  x.y = m[0] * u.y + m[1] * u.x + m[2] * v.y + m[3] * v.x;
  y.x = m[4] * u.x - m[5] * u.y + m[6] * v.x - m[7] * v.y;
  y.y = m[4] * u.y + m[5] * u.x + m[6] * v.y + m[7] * v.x;
}

// a *= f in place (f per lane) unless `skip` (workgroup-uniform, in an SGPR)
// is set; the branch is inside the statement, so a is defined once either way.
// Complex products in place: (ax, ay) <- (fx ax - fy ay, fx ay + fy ax), the
// cross products into temporaries first, then one fused multiply-add per
// component written in place (no moves).
__device__ __forceinline__ void cmul_inplace(float2 &a, float2 f, uint32_t skip) {
  float t, u;
  asm("s_cmp_lg_u32 %[skip], 0\n\t"
      "s_cbranch_scc1 .Lqm%=\n\t"
      "v_mul_f32 %[t], %[fy], %[ax]\n\t"
      "v_mul_f32 %[u], %[fy], %[ay]\n\t"
      "v_fma_f32 %[ax], %[fx], %[ax], -%[u]\n\t"
      "v_fma_f32 %[ay], %[fx], %[ay], %[t]\n"
      ".Lqm%=:"
      : [ax] "+v"(a.x), [ay] "+v"(a.y), [t] "=&v"(t), [u] "=&v"(u)
      : [fx] "v"(f.x), [fy] "v"(f.y), [skip] "s"(skip)
      : "scc");
}

#include "slot_asm.inc"

// In-place updates with tied operands for straight-line code (the chain
// kernel): every amplitude keeps its registers, so the compiler cannot hold
// a step's results and its inputs at the same time. The matrix / factor
// operands are workgroup-uniform (SGPRs) unless the name ends in _v.
__device__ __forceinline__ void bfly_real_s(float m00, float m01, float m10, float m11,
                                            float2 &a, float2 &b) {
  float t0, t1, t2, t3;
  asm("v_mul_f32 %[t0], %[m10], %[ax]\n\t"
      "v_mul_f32 %[t1], %[m01], %[bx]\n\t"
      "v_mul_f32 %[t2], %[m10], %[ay]\n\t"
      "v_mul_f32 %[t3], %[m01], %[by]\n\t"
      "v_fma_f32 %[bx], %[m11], %[bx], %[t0]\n\t"
      "v_fma_f32 %[ax], %[m00], %[ax], %[t1]\n\t"
      "v_fma_f32 %[by], %[m11], %[by], %[t2]\n\t"
      "v_fma_f32 %[ay], %[m00], %[ay], %[t3]"
      : [ax] "+v"(a.x), [ay] "+v"(a.y), [bx] "+v"(b.x), [by] "+v"(b.y), [t0] "=&v"(t0),
        [t1] "=&v"(t1), [t2] "=&v"(t2), [t3] "=&v"(t3)
      : [m00] "s"(m00), [m01] "s"(m01), [m10] "s"(m10), [m11] "s"(m11));
}

__device__ __forceinline__ void bfly_cplx_s(const float *m, float2 &a, float2 &b) {
  float t0, t1, t2, t3;
  asm("v_mul_f32 %[t0], %[m0i], %[ay]\n\t"
      "v_fma_f32 %[t0], %[m1r], %[bx], -%[t0]\n\t"
      "v_fma_f32 %[t0], -%[m1i], %[by], %[t0]\n\t"
      "v_mul_f32 %[t1], %[m0i], %[ax]\n\t"
      "v_fmac_f32 %[t1], %[m1r], %[by]\n\t"
      "v_fmac_f32 %[t1], %[m1i], %[bx]\n\t"
      "v_mul_f32 %[t2], %[m2r], %[ax]\n\t"
      "v_fma_f32 %[t2], -%[m2i], %[ay], %[t2]\n\t"
      "v_fma_f32 %[t2], -%[m3i], %[by], %[t2]\n\t"
      "v_mul_f32 %[t3], %[m2r], %[ay]\n\t"
      "v_fmac_f32 %[t3], %[m2i], %[ax]\n\t"
      "v_fmac_f32 %[t3], %[m3i], %[bx]\n\t"
      "v_fma_f32 %[ax], %[m0r], %[ax], %[t0]\n\t"
      "v_fma_f32 %[ay], %[m0r], %[ay], %[t1]\n\t"
      "v_fma_f32 %[bx], %[m3r], %[bx], %[t2]\n\t"
      "v_fma_f32 %[by], %[m3r], %[by], %[t3]"
      : [ax] "+v"(a.x), [ay] "+v"(a.y), [bx] "+v"(b.x), [by] "+v"(b.y), [t0] "=&v"(t0),
        [t1] "=&v"(t1), [t2] "=&v"(t2), [t3] "=&v"(t3)
      : [m0r] "s"(m[0]), [m0i] "s"(m[1]), [m1r] "s"(m[2]), [m1i] "s"(m[3]), [m2r] "s"(m[4]),
        [m2i] "s"(m[5]), [m3r] "s"(m[6]), [m3i] "s"(m[7]));
}

// a *= f in place.
__device__ __forceinline__ void cmul_s(float2 &a, float fx, float fy) {
  float t, u;
  asm("v_mul_f32 %[t], %[fy], %[ax]\n\t"
      "v_mul_f32 %[u], %[fy], %[ay]\n\t"
      "v_fma_f32 %[ax], %[fx], %[ax], -%[u]\n\t"
      "v_fma_f32 %[ay], %[fx], %[ay], %[t]"
      : [ax] "+v"(a.x), [ay] "+v"(a.y), [t] "=&v"(t), [u] "=&v"(u)
      : [fx] "s"(fx), [fy] "s"(fy));
}

__device__ __forceinline__ void cmul_v(float2 &a, float2 f) {
  float t, u;
  asm("v_mul_f32 %[t], %[fy], %[ax]\n\t"
      "v_mul_f32 %[u], %[fy], %[ay]\n\t"
      "v_fma_f32 %[ax], %[fx], %[ax], -%[u]\n\t"
      "v_fma_f32 %[ay], %[fx], %[ay], %[t]"
      : [ax] "+v"(a.x), [ay] "+v"(a.y), [t] "=&v"(t), [u] "=&v"(u)
      : [fx] "v"(f.x), [fy] "v"(f.y));
}

template <int NL, int LB>
__device__ __forceinline__ void apply_slot(float2 (&x)[1 << NL], uint32_t cm, const float *m) {
#pragma unroll
  for (int l = 0; l < (1 << NL); l++) {
    if (l & (1 << LB)) continue;
    if ((uint32_t(l) & cm) != cm) continue;
    rot(m, x[l], x[l | (1 << LB)]);
  }
}

// 2x2 `m` on tile-local bit `lbit` wherever the tile-local bits `cm` are 1.
template <int NL>
__device__ __forceinline__ void apply_any(float2 (&x)[1 << NL], uint32_t lbit, uint32_t cm,
                                          const float *m) {
  switch (lbit) {
    case 0: apply_slot<NL, 0>(x, cm, m); break;
    case 1: if constexpr (NL > 1) apply_slot<NL, 1>(x, cm, m); break;
    case 2: if constexpr (NL > 2) apply_slot<NL, 2>(x, cm, m); break;
    case 3: if constexpr (NL > 3) apply_slot<NL, 3>(x, cm, m); break;
    case 4: if constexpr (NL > 4) apply_slot<NL, 4>(x, cm, m); break;
    default: break;
  }
}

template <int NL>
__device__ __forceinline__ void apply_any(float2 (&x)[1 << NL], const Slot &s) {
  apply_any<NL>(x, s.lbit, s.cmask, s.m);
}

typedef float v4f __attribute__((ext_vector_type(4)));

template <bool NT>
__device__ __forceinline__ float4 ld4(const float2 *p) {
  if constexpr (NT) {
    const v4f v = __builtin_nontemporal_load(reinterpret_cast<const v4f *>(p));
    return make_float4(v.x, v.y, v.z, v.w);
  }
  return *reinterpret_cast<const float4 *>(p);
}

template <bool NT>
__device__ __forceinline__ void st4(float2 *p, float4 v) {
  if constexpr (NT) {
    const v4f w = {v.x, v.y, v.z, v.w};
    __builtin_nontemporal_store(w, reinterpret_cast<v4f *>(p));
  } else {
    *reinterpret_cast<float4 *>(p) = v;
  }
}

// Workgroups of the low tile passes (two-wave workgroups): enough waves to
// fill the chip many times over.
constexpr uint64_t kLowGrid = 40960;

// sin and cos of 2 pi t for a phase t in turns: t is reduced to [-1/2, 1/2]
// in FP64, then the hardware transcendental unit, which takes its argument in
// turns, evaluates both (v_sin_f32 / v_cos_f32, no range reduction needed).
__device__ __forceinline__ void turn_sincos(double t, float &s, float &c) {
  const float r = float(t - rint(t));
  s = __builtin_amdgcn_sinf(r);
  c = __builtin_amdgcn_cosf(r);
}

// Sum over the wave; every lane gets the total (butterfly).
__device__ __forceinline__ double wave_allsum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Read-only kernel inputs seen through the constant address space: uniform
// loads through such a pointer become scalar loads into SGPRs (the compiler
// need not prove that the state stores of the kernel do not alias them).
// Instance blockIdx.y of a batched launch reads its own state and its own
// region of the op arena: the state pointer moves by `bstate` amplitudes and
// arena pointers by `barena` bytes per instance (both 0 for one instance).
template <typename T>
__device__ __forceinline__ T *bshift(T *p, uint64_t bytes) {
  return reinterpret_cast<T *>(reinterpret_cast<uintptr_t>(p) + bytes);
}

template <typename T>
using const_ptr = const __attribute__((address_space(4))) T *;

template <typename T>
__device__ __forceinline__ const_ptr<T> as_const(const T *p) {
  return (const_ptr<T>)p;
}

// The 2x2 of a chain step (m) on all pairs of tile-local bit BIT.
// MODE: 0 complex 2x2, 1 real 2x2, 2 the unscaled butterfly [[1,1],[1,-1]]
// (the scale of a run of such steps is carried by the run's last step,
// which uses mode 1).
template <int NL, int BIT, int MODE>
__device__ __forceinline__ void chain_mix(float2 (&x)[1 << NL],
                                          const ChainStep __attribute__((address_space(4))) &st) {
  if constexpr (MODE == 2) {
#pragma unroll
    for (int l = 0; l < (1 << NL); l++) {
      if (l & (1 << BIT)) continue;
      const float2 u = x[l], v = x[l | (1 << BIT)];
      x[l] = make_float2(u.x + v.x, u.y + v.y);
      x[l | (1 << BIT)] = make_float2(u.x - v.x, u.y - v.y);
    }
  } else {
    float m[8];
#pragma unroll
    for (int j = 0; j < 8; j++) m[j] = st.m[j];
#pragma unroll
    for (int l = 0; l < (1 << NL); l++) {
      if (l & (1 << BIT)) continue;
      if constexpr (MODE == 1) bfly_real_s(m[0], m[2], m[4], m[6], x[l], x[l | (1 << BIT)]);
      else bfly_cplx_s(m, x[l], x[l | (1 << BIT)]);
    }
  }
}

// One chain step on tile-local bit BIT: the 2x2, then the phase of the
// combinations holding BIT, e * lo[low two other bits] * hi[the other bits
// above them] (other local bits ascending).
template <int NL, int BIT, int MODE>
__device__ __forceinline__ void chain_step(float2 (&x)[1 << NL],
                                           const ChainStep __attribute__((address_space(4))) &st,
                                           float2 e) {
  chain_mix<NL, BIT, MODE>(x, st);
  float2 a[4];
  a[0] = e;
#pragma unroll
  for (int h = 1; h < 4; h++) a[h] = cmul(e, make_float2(st.hi[2 * h - 2], st.hi[2 * h - 1]));
#pragma unroll
  for (int l = 0; l < (1 << NL); l++) {
    if (!(l & (1 << BIT))) continue;
    const int c = (l & ((1 << BIT) - 1)) | ((l >> (BIT + 1)) << BIT);
    const int lo = c & 3, hi = c >> 2;
    if (lo) cmul_s(x[l], st.lo[2 * lo - 2], st.lo[2 * lo - 1]);
    cmul_v(x[l], a[hi]);
  }
}

// Step index i read through an opaque register after the tile's last update,
// so the compiler cannot hoist every step's scalar constants to the top (it
// would spill SGPRs).
__device__ __forceinline__ uint32_t opaque_step(int i, float dep) {
  uint32_t si = uint32_t(i);
  asm volatile("" : "+s"(si) : "v"(dep));
  return si;
}

}  // namespace
}  // namespace qmamd
