// kernels.h -- launch interface of the gfx950 statevector kernels.
//
// Every launcher enqueues on the given stream and returns immediately; none
// synchronises. Argument structs travel by value as kernel arguments.

#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace qmamd {

// One 2x2 applied inside a gather tile: `lbit` is the tile-local bit it acts
// on, `cmask` the tile-local bits that must all be 1 for it to apply
// (controls that live inside the tile). Matrix row-major, (re, im) pairs.
struct Slot {
  uint8_t lbit;
  uint8_t cmask;
  uint8_t pad[2];
  float m[8];
};

// Gather layout. Thread T owns the 2^NL amplitudes
//   deposit(T) | fixed | sum_k bit_k(r) << rpos[k] | e,   r < 2^(NL-1), e < 2,
// where deposit() inserts zero bits at the ascending positions ins[], which
// are {0} + rpos[] + the fixed-one bits; hm[k] = ~(2^ins[k] - 1) (inserting a
// zero at p turns t into t + (t & hm)).
struct Layout {
  uint64_t nthreads;
  uint64_t fixed;
  uint32_t nins;
  uint8_t ins[16];
  uint8_t rpos[8];
  uint64_t hm[16];
  // 1 when positions 1..3 are thread-index bits: then eight neighbouring lanes
  // cover one whole 128-byte line and the accesses can bypass the caches
  // (nontemporal); otherwise a line is split across load instructions and
  // must be merged in L2.
  uint32_t streaming;
};

struct GatherArgs {
  float2 *a;   // state
  float2 *b;   // optional second state receiving the same op (or nullptr)
  Layout lay;
  uint32_t nslots;
  Slot slots[4];
};

// Diagonal phase layer after host-side classification, in turns (1 = 2 pi).
// Tile-local bit k (k = 0 is physical bit 0, k >= 1 is rpos[k-1]) carries
// linear coefficient beta0[k] and pair coefficient gamma[k][j]; terms on
// non-local bits are read from `terms`:
//   kind 0: linear, phi += ang if bit p of the base index is set
//   kind 1: pair,   phi += ang if both bits of mask are set in the base
//   kind 2: cross,  beta[k] += ang if bit p of the base index is set
struct DiagTerm {
  uint64_t mask;
  double ang;
  uint32_t kind;
  uint32_t k;
};

struct DiagArgs {
  float2 *a;
  Layout lay;
  double constant;
  double beta0[4];
  double gamma[4][4];
  const DiagTerm *terms;  // [0, nuniform): uniform in a workgroup
  uint32_t nterms;
  uint32_t nuniform;
};

// General diagonal (masks of any weight), one amplitude pair per thread.
struct DiagGenericArgs {
  float2 *a;
  uint64_t dim;
  const uint64_t *masks;
  const double *angles;  // turns
  uint32_t nterms;
  double global;         // turns
};

struct BlockedArgs {
  float2 *a;
  uint64_t nblocks;
  uint32_t bits;
  uint32_t count;
  const Slot *gates;  // lbit = physical wire (< bits)
};

// One op of a merged pass (several plan ops applied to the 2^NL amplitudes a
// thread holds between one load and one store of them).
//   type 0, a 2x2 slot: applies to tile-local bit `lbit` where the tile-local
//     bits `cmask` are 1 and the thread's base index has every bit of `gmask`;
//     type 2 is the same with a real matrix (imaginary parts zero).
//   type 1, a phase layer: amplitude l (tile-local combination) is multiplied
//     by table[l] * exp(2 pi i (phi + sum_{k in l} beta_k)), where table holds
//     everything that does not depend on the thread (constant, local pair
//     terms, linear terms of local bits outside `bmask`) and phi / beta_k
//     (k in `bmask`, starting from beta0[k]) are FP64 sums, in turns, of the
//     terms[term0 .. term0 + nterms) that read the thread's base index: the
//     first `nuniform` read only bits that are the same across the workgroup
//     (`umask` bit 0 / bit 1+k: they feed phi / beta_k). `has_phi`: some term
//     feeds phi. Applied to local combinations containing `cmask`, in threads
//     whose base has `gmask` (the others keep phase 0).
constexpr int kPassMaxNL = 5;
enum PassOpType : uint32_t { kPassSlot = 0, kPassPhase = 1, kPassRealSlot = 2 };
struct PassOp {
  uint32_t type;
  uint32_t lbit;
  uint32_t cmask;
  uint32_t sel;   // slot: code section, 8 * kind + lbit (kinds: slot_asm.inc's generator)
  uint32_t pm;    // slot, masked: pair p (p-th local index with lbit clear) applies
  uint32_t term0;
  uint32_t nterms;
  uint32_t nuniform;
  uint32_t umask;
  uint32_t bmask;
  uint32_t has_phi;
  uint32_t pad[3];
  uint64_t gmask;
  float m[8];
  double beta0[kPassMaxNL];
  float table[2 << kPassMaxNL];  // (re, im) per local combination
};

struct PassArgs {
  float2 *a;
  Layout lay;
  const PassOp *ops;
  const DiagTerm *terms;
  uint32_t nops;
  uint64_t bstate;  // batched launches: per-instance state stride (amplitudes)
  uint64_t barena;  // and op-arena stride (bytes); 0 for one instance
};

// A merged pass of chain shape: step i applies an uncontrolled 2x2 to
// tile-local bit b_i (b_i = i + 1 for i < K, then bit 0 if the chain ends on
// wire 0), then multiplies every local combination l with bit b_i set by
// exp(2 pi i (beta + sum_{k in l, k != b_i} gamma_k)). Every mask of such a
// phase layer holds b_i, so its local part is linear in the other local bits,
// which split into a low group (the lowest two) and a high group (the rest,
// up to two): lo[c] and hi[c] (c = 1..3, the group's combination) are the
// products of exp(2 pi i gamma_k) over the group, rounded to float. beta
// (turns) = beta0 + the step's base-index terms that hold in the thread's
// base index. A step without a phase layer has unit factors and no terms.
// Real chains use m[0], m[2], m[4], m[6].
struct ChainStep {
  float m[8];
  float lo[6];
  float hi[6];
  double beta0;
};

// The base-index terms of all steps in two flat lists (DiagTerm.k = step):
// `uni` reads only workgroup-uniform bits, `lane` the rest.
struct ChainArgs {
  float2 *a;
  Layout lay;
  const ChainStep *steps;
  const DiagTerm *uni;
  const DiagTerm *lane;
  uint32_t nuni;
  uint32_t nlane;
};

// Wide chain pass: 4 + WB chain steps (WB = 2 or 3 wave bits) on wires
// w1..w(4+WB) (pos[0..]) over a workgroup tile of 2^(11+WB) amplitudes held
// by 2^WB waves. Each thread holds 32 amplitudes: bit 0 and tile-local bits
// 1..4 = w1..w4; lane bits 0..5 of the thread index sit on positions
// lanepos[0..5] and the WB wave bits on w5.. . Steps 0-3 act on local bits
// 1..4; an LDS exchange then swaps local bits 1..WB with the wave bits, and
// steps 4.. act on local bits 1..WB (now w5..) while the wave bits stand for
// w1..wWB. `lay` enumerates the tiles (nthreads = tiles, ins = the tile
// positions and the fixed bits). Steps 0-3 are folded against the first
// arrangement, the rest against the second; base-index terms as in ChainArgs
// (k = step).
constexpr int kWideArgSteps = 7;
struct WideArgs {
  float2 *a;
  Layout lay;
  uint8_t lanepos[6];
  uint8_t pos[kWideArgSteps];
  const ChainStep *steps;
  const DiagTerm *uni;
  const DiagTerm *lane;
  uint32_t nuni;
  uint32_t nlane;
};

// Low chain pass: S chain steps on distinct wires 1..10+WB, then optionally
// one on wire 0, over workgroup tiles of 2^(11+WB) contiguous amplitudes
// (bits 0..10+WB) held by 2^WB waves (WB = 1 or 2). The tile is loaded and
// stored coalesced (arrangement 0: lanes on bits 1..6, the wave bits on 7..,
// the register bits on the four above them) and permuted through LDS into
// arrangement 1 + a for steps 4a .. 4a + 3, whose wires are register bits
// (tile-local 1..4); the wire-0 step runs in the last arrangement. `lanepos`
// gives each arrangement's lane bits (0..5) and wave bits (6..), `regpos` its
// register bits. Steps are folded against their arrangement; base-index
// terms as in ChainArgs (k = step).
constexpr int kLowMaxArr = 4;
struct LowArgs {
  float2 *a;
  Layout lay;  // tiles: ins = the tile bits
  uint8_t lanepos[kLowMaxArr][8];
  uint8_t regpos[kLowMaxArr][4];
  const ChainStep *steps;
  const DiagTerm *uni;
  const DiagTerm *lane;
  uint32_t nuni;
  uint32_t nlane;
  // Arrangement 1 is arrangement 0 (the first four steps' wires are its
  // register bits, in step order): no permute before the first steps.
  uint32_t first_in_place;
};

// Generic low tile pass: merged-pass ops on wires 0..11 over workgroup tiles
// of 2^12 contiguous amplitudes held by two waves, in segments. Segment s
// runs ops [op0, op1) (PassOps with NL = 5) in its own arrangement of the
// tile: local bit 0 = wire 0, local bits 1..4 = the wires at rpos[0..3], the
// thread bits (lanes 0..5, then the wave bit) on the wires at tpos[0..6].
// `tslot` / `rslot` are the LDS slot bases (float4 units) of those bits. The
// tile is loaded and stored coalesced (lanes on 1..6, wave bit 7, registers
// 8..11) and permuted through LDS before each segment and after the last.
struct TileSeg {
  uint32_t op0, op1;
  uint32_t tslot[7];
  uint32_t rslot[4];
  uint8_t tpos[7];
  uint8_t pad;
};
struct TileArgs {
  float2 *a;
  Layout lay;  // tiles: ins = bits 0..11
  const TileSeg *segs;
  uint32_t nsegs;
  const PassOp *ops;
  const DiagTerm *terms;
  uint64_t bstate;  // batched launches, as in PassArgs
  uint64_t barena;
};

// Generic wide pass: merged-pass ops (NL = 5) over the workgroup tile of a
// wide chain pass (WideArgs: wires pos[0..3+WB], lanes on lanepos): ops
// [0, na) run in the first arrangement (local bits 1..4 = pos[0..3]), ops
// [na, nops) after the LDS exchange (local bits 1..WB = pos[4..3+WB], the
// rest still pos[WB..3]).
struct GWideArgs {
  float2 *a;
  Layout lay;
  uint8_t lanepos[6];
  uint8_t pos[kWideArgSteps];
  const PassOp *ops;
  const DiagTerm *terms;
  uint32_t na;
  uint32_t nops;
  uint64_t bstate;  // batched launches, as in PassArgs
  uint64_t barena;
};

// LDS slot (float4 index) of tile index i, i = (tile bits 1..11) >> 1: the
// low four bits are xored with the higher ones (linear over GF(2)).
constexpr uint32_t tile_swizzle(uint32_t i) { return i ^ (((i >> 4) ^ (i >> 8)) & 0xFu); }

void launch_fill_basis(hipStream_t s, float2 *a, uint64_t dim);
void launch_zero(hipStream_t s, float2 *a, uint64_t dim);
// Reads and writes back, unchanged, every amplitude of a 2^n state in tiles
// shaped like a wide chain pass whose seven step bits are p0..p0+6 (needs
// 7 <= p0, p0 + 7 <= n, n >= 14; otherwise launches nothing).
void launch_placement_probe(hipStream_t s, float2 *a, uint32_t n, uint32_t p0);
// out[k] = a[idx[k]] for k < count (indices below the state size).
void launch_pick(hipStream_t s, const float2 *a, const uint64_t *idx, uint64_t count,
                 float2 *out);
void launch_gather(hipStream_t s, int nl, const GatherArgs &g);
void launch_diag(hipStream_t s, int nl, const DiagArgs &d);
void launch_diag_generic(hipStream_t s, const DiagGenericArgs &d);
void launch_blocked(hipStream_t s, const BlockedArgs &b);
// nl in {4, 5}.
void launch_pass(hipStream_t s, int nl, const PassArgs &p);
// nl = 5 with k = 4, or nl = 4 with k in 1..3; `zero`: a last step on bit 0.
void launch_chain(hipStream_t s, int nl, int k, bool zero, bool real, const ChainArgs &c);
// Step modes of the wide and low chain kernels: 0 complex 2x2s, 1 real
// 2x2s, 2 every step a multiple of [[1,1],[1,-1]] with the product of the
// multiples folded into the last step's matrix.
// steps = 6 or 7.
void launch_wide(hipStream_t s, int steps, int mode, const WideArgs &w);
// s in 5..12 steps on wires 1..10+wb, `zero`: a last step on wire 0.
// Low passes take one more mode, kLowModeTri: mode 2 with triangular phase
// tables (each step's phase ignores the register bits of its arrangement's
// earlier steps, and the wire-0 step's ignores all four), which need one
// complex product per amplitude and step.
constexpr int kLowModeTri = 3;
void launch_low(hipStream_t s, int steps, bool zero, int mode, int wb, const LowArgs &l);
// The low chain kernels for steps S0 and S0 + 1 (S0 = 5, 7, 9, 11).
template <int S0>
void launch_low_pair(hipStream_t s, int steps, bool zero, int mode, int wb, const LowArgs &l);
void launch_tile(hipStream_t s, const TileArgs &t);
// wb = 2 or 3 wave bits.
void launch_gwide(hipStream_t s, int wb, const GWideArgs &g);

// FP64 reductions. Each writes per-block partials into `partials` (capacity
// kRedBlocks doubles or double2) and then reduces them in a fixed order into
// out[0] (and out[1]), so results are run-to-run deterministic.
constexpr int kRedBlocks = 2048;
constexpr int kRedThreads = 256;

// Terms of one k_pauli_terms launch: flip = x | y, sign = y | z of each.
constexpr uint32_t kPauliTermsPerLaunch = 128;
struct PauliTerms {
  uint64_t flip[kPauliTermsPerLaunch];
  uint64_t sign[kPauliTermsPerLaunch];
};
// out[2 t], out[2 t + 1] = sum_i s_t(i ^ flip_t) conj(a_i) a_{i ^ flip_t} for
// t < count (<= kPauliTermsPerLaunch) in one launch; `partials` holds
// count * pauli_terms_blocks(dim) * 2 doubles.
unsigned pauli_terms_blocks(uint64_t dim);
void launch_pauli_terms(hipStream_t s, const float2 *a, uint64_t dim, const PauliTerms &t,
                        uint32_t count, double *partials, double *out);

// out[0] = sum |a_i|^2
void launch_abs2(hipStream_t s, const float2 *a, uint64_t dim, double *partials,
                 double *out);
// out[0..1] = sum_i s(j) conj(bra_i) ket_j,  j = i ^ flip,
// s(j) = (-1)^popcount(j & sign). bra == ket gives the Pauli expectation sum.
void launch_pauli(hipStream_t s, const float2 *bra, const float2 *ket,
                  uint64_t dim, uint64_t flip, uint64_t sign, double *partials,
                  double *out);

// lam_i += sum_t coef_t * s_t(j) * psi_j over terms sharing one flip (j = i ^
// flip); coefs are complex (c_t * i^ny_t) in double. Terms: sign masks and
// coefficients, up to kAccumTerms per launch.
constexpr int kAccumTerms = 32;
struct AccumArgs {
  float2 *lam;
  const float2 *psi;
  uint64_t bstate;  // batched launches: per-instance state stride
  uint64_t dim;
  uint64_t flip;
  uint32_t nterms;
  uint64_t sign[kAccumTerms];
  double cre[kAccumTerms];
  double cim[kAccumTerms];
};
void launch_pauli_accum(hipStream_t s, const AccumArgs &a);

// lambda += sum_t c_t s_t(i ^ f_t) psi[i ^ f_t] for terms whose flip f_t is 0
// or one bit of the register tile {bit 0, w[0..nw)} (nw <= 3): a thread holds
// the 2^(nw+1) amplitudes of a tile of both states, so one pass serves every
// such flip. Terms are sorted by slot: slot 0 = flip bit 0, slot 1 + k =
// flip bit w[k], slot nw + 1 = no flip; slot s owns terms [off[s], off[s+1]).
constexpr int kAccumTileWires = 3;
struct AccumTileArgs {
  float2 *lam;
  const float2 *psi;
  uint64_t bstate;  // batched launches: per-instance state stride
  uint64_t dim;
  uint32_t nw;
  uint8_t w[kAccumTileWires];
  uint32_t off[kAccumTileWires + 3];
  uint64_t sign[kAccumTerms];
  double cre[kAccumTerms];
  double cim[kAccumTerms];
};
void launch_pauli_accum_tile(hipStream_t s, const AccumTileArgs &a);

// Co-state pass over an LDS tile (states of kAdjTileMinQubits or more): the
// tile is 2^12 amplitudes over wires 0..6 and hpos[0..4], loaded as by
// k_adj_tile (16 per thread: bit 0, register bits on hpos[2..4]) and also
// written to LDS. lam_i (+)= sum over terms of coef_t s_t(j) psi_j, j = i ^ f:
//  - diagonal terms (f = 0; real coefficients): sorted by their local mask m
//    (bit 0 -> 1, hpos[2 + k] -> 2 << k), [doff[m], doff[m + 1]); each thread
//    sums the coefficients by local mask in FP64, signed by the parity of
//    its base index over the term's mask, and a Walsh-Hadamard transform over
//    its four local bits gives each amplitude's factor;
//  - flip groups g < ngroups: flip gflip[g] (within the tile), terms
//    [goff[g], goff[g + 1]), partner amplitudes read from LDS at slot offset
//    gslot[g]; coefficients already carry the sign of the flip's own parity
//    (s_t(j) = s_t(i) s_t(f)); `lmask` is each term's local mask.
// `first`: lam is written without being read. With `epart`, the pass also
// leaves per-block FP64 partials of <psi|lam> (after its update) for
// k_finish (2 doubles per block, blocks of instance y after y * gridDim).
// With lam null the pass reads psi only: the partials then hold <psi|H_t psi>
// of this tile's terms H_t alone, and nothing is written to HBM.
constexpr int kCoTerms = 48;
constexpr int kCoGroups = 24;
struct CoTileArgs {
  float2 *lam;
  const float2 *psi;
  Layout lay;
  uint32_t hpos[5];
  uint32_t first;
  uint32_t ngroups;
  uint32_t doff[17];
  uint32_t goff[kCoGroups + 1];
  uint32_t gslot[kCoGroups];
  uint64_t gflip[kCoGroups];
  uint64_t sign[kCoTerms];
  uint32_t lmask[kCoTerms];
  double cre[kCoTerms];
  double cim[kCoTerms];
  uint64_t bstate;  // batched launches: per-instance state stride
  // Energy-only passes (lam null), per flip group: bit 0 / bit 1 = some term
  // has a real / imaginary coefficient (the pass needs the real / imaginary
  // parts of the products conj(x) p); kCoGroupPlain = one term, real
  // coefficient, local mask 0; kCoGroupLocal (plain groups only) = the flip
  // is one of the thread's own local bits, bit (gpart >> 4) & 3 of the local
  // index. Both passes: bits 8..15 = how many of the group's terms, placed
  // first, have a real coefficient and local mask 0.
  uint32_t gpart[kCoGroups];
};
constexpr uint32_t kCoGroupPlain = 4;
constexpr uint32_t kCoGroupLocal = 8;
// Single state; with `energy` (2 doubles) the fused <psi|lam> lands there.
void launch_costate_tile(hipStream_t s, const CoTileArgs &a, double *partials, double *energy);
// out[0..1] = sum over k < count of slots[2 k], slots[2 k + 1] (FP64, fixed
// order): the shares of several energy passes.
void launch_sum_slots(hipStream_t s, const double *slots, uint32_t count, double *out);
// Batched: instance y's energy to energy[2 y] (when non-null).
unsigned costate_tile_blocks(const Layout &lay);
void launch_costate_tile_batched(hipStream_t s, const CoTileArgs &a, double *partials,
                                 double *energy, uint32_t count);

// Adjoint step for a rotation exp(-i t P / 2), P in {X, Y, Z} on wire q:
// accumulates <lam|P|psi> (FP64, per-block partials reduced into
// out[2*slot], out[2*slot+1]) and then applies the inverse rotation to both
// states, in one pass over both.
struct AdjArgs {
  float2 *psi;
  float2 *lam;
  uint64_t dim;
  uint32_t q;
  uint32_t axis;  // 0 = X, 1 = Y, 2 = Z
  float m[8];     // inverse rotation, row-major
};
void launch_adjoint_rot(hipStream_t s, const AdjArgs &a, double *partials,
                        double *out2);

// Merged backward pass of the adjoint sweep: consecutive inverse gates whose
// wires fit one thread's tile of 2^nl amplitudes (bit 0 and up to
// kAdjMaxReg register wires), applied to psi and lam together.
// Each op applies a 2x2 to both states (slot_asm section `sel`, pair mask
// `pm`; the identity in threads whose base index lacks a bit of `gmask`);
// type 1 first accumulates the rotation's overlap <lam|P|psi>, P = X, Y, Z
// (`axis` 0, 1, 2) on tile-local bit `lbit`, in FP64 into slot `rslot`.
constexpr int kAdjMaxRot = 8;
// Register wires per backward pass. Four (a 32-amplitude tile of each state,
// 175 VGPRs, two waves per SIMD) cut HEA n=32 from 31 to 24 passes but made
// each pass 29% slower: measured no faster, so three.
constexpr int kAdjMaxReg = 3;
constexpr int kAdjBlocks = 4096;
// Tile passes (k_adj_tile) also take: axis 3 = Z on wires outside the
// thread's tile-local bits (the overlap of every amplitude, signed per
// thread), axis 4 = Z Z on tile-local bits lbit and lbit2; a thread whose base
// index has odd parity over `pmask` applies `malt` instead of `m` and negates
// its overlap (a diagonal phase or Z generator on a wire that is not
// tile-local). They compute only the overlap's part `comp` (0 real, 1
// imaginary: the one the gradient reads) and leave the other 0.
struct AdjOp {
  uint32_t type;
  uint32_t lbit;
  uint32_t sel;
  uint32_t pm;
  uint32_t axis;
  uint32_t rslot;
  uint32_t lbit2;
  uint32_t comp;
  uint64_t gmask;
  uint64_t pmask;
  float m[8];
  float malt[8];
};

struct AdjPassArgs {
  float2 *psi;
  float2 *lam;
  Layout lay;
  const AdjOp *ops;
  uint32_t nops;
  uint32_t nrot;
  uint32_t nl;  // tile: bit 0 plus nl - 1 register wires (4)
  uint64_t bstate;  // batched launches, as in PassArgs
  uint64_t barena;
};

// Grid of kAdjBlocks workgroups (four waves per SIMD at the kernel's
// register use, which a streaming pass over two states needs to keep enough
// loads in flight); per-block FP64 partials of each rotation go to
// `partials` (kAdjMaxRot * kAdjBlocks * 2 doubles) and are reduced in a
// fixed order into out[2 r], out[2 r + 1] (r < nrot).
void launch_adj_pass(hipStream_t s, const AdjPassArgs &a, double *partials, double *out);

// Backward tile pass: each workgroup (kAdjTileThreads threads) holds a tile of
// 2^12 amplitudes of both states, over wires 0..6 and the five wires hpos
// (loaded with lanes on wires 1..6, the two wave bits on hpos[0..1] and the
// three register bits on hpos[2..4]); per segment the tile is permuted through
// LDS so that the segment's three register wires are any three tile wires
// other than 0, and the segment's AdjOps run on the thread's 16 amplitudes of
// each state as in the register-tile pass. The slot of tile bit b (1..11) in
// LDS is tile_swizzle(1 << (b - 1)), xor-combined.
constexpr uint32_t kAdjTileMinQubits = 12;
constexpr int kAdjTileThreads = 256;
constexpr int kAdjTileMaxSeg = 8;
constexpr int kAdjTileMaxRot = 32;
constexpr int kAdjTileBlocks = 4096;
struct AdjSeg {
  uint32_t op0, op1;  // AdjOps [op0, op1) run in this arrangement
  uint32_t tslot[8];  // slot part of thread bit j (lane bits 0..5, wave bits 6..7)
  uint32_t rslot[3];  // slot part of register bit k (tile-local bit k + 1)
  uint8_t tpos[8];    // wire of thread bit j
  uint32_t pad;
};
struct AdjTileArgs {
  float2 *psi;
  float2 *lam;
  Layout lay;  // tiles: ins = wires 0..6 and hpos, ascending
  const AdjSeg *segs;
  const AdjOp *ops;
  uint32_t nsegs;
  uint32_t nrot;
  uint32_t hpos[5];
  uint64_t bstate;  // batched launches, as in PassArgs
  uint64_t barena;
};
// Per-block partials: kAdjTileMaxRot * kAdjTileBlocks * 2 doubles.
void launch_adj_tile(hipStream_t s, const AdjTileArgs &a, double *partials, double *out);
// The same tile passes on one state (a.psi; lam and overlaps unused): the
// forward sweep of a device-resident gradient program.
void launch_fwd_tile(hipStream_t s, const AdjTileArgs &a);

// Angles of a device-resident gradient program, written on the device: the
// AdjOp at byte offset `off` of the arena takes m / malt for angle a = +t or
// -t (bit 8 of `code`: the inverse), t = params[param], by recipe code & 15:
//   0: the rotation's own 2x2 (code >> 4 & 15: 0 RX, 1 RY, 2 RZ), malt = m;
//   with d0 = exp(-i a / 2), d1 = exp(i a / 2):
//   1: m = diag(d0, d0), malt = diag(d1, d1);   2: m = diag(d0, d1), malt =
//   diag(d1, d0);   3: m = malt = diag(d1 conj(d0), d0 conj(d1)).
struct AnglePatch {
  uint32_t off;
  uint32_t param;
  uint32_t code;
  uint32_t pad;
};
// Patch code flag: the target is the 8 floats of a PassOp's matrix (a
// merged-pass slot) rather than an AdjOp's m and malt.
constexpr uint32_t kPatchPassOp = 512;
void launch_patch_angles(hipStream_t s, uint8_t *arena, const AnglePatch *p, uint32_t count,
                         const double *params);

// Sampling: FP64 sums of |a|^2 over 1024-amplitude blocks, an inclusive FP64
// prefix over them, and per-shot resolution (splitmix64 stream, binary search
// over blocks, in-block walk clamped to the last non-zero amplitude).
constexpr uint64_t kSampleBlock = 1024;
void launch_block_sums(hipStream_t s, const float2 *a, uint64_t dim,
                       double *sums);
void launch_prefix(hipStream_t s, double *v, uint64_t count, double *chunk);
void launch_shots(hipStream_t s, const float2 *a, uint64_t dim,
                  const double *cdf, uint64_t nblocks, uint64_t shots,
                  uint64_t seed, uint64_t *out);

// Batched launches: `count` instances side by side (grid row y = instance
// y), each reading its state at the args' state stride and its ops at the
// arena stride; reductions and their outputs per instance back to back.
void launch_pass_batched(hipStream_t s, int nl, const PassArgs &p, uint32_t count);
void launch_tile_batched(hipStream_t s, const TileArgs &t, uint32_t count);
void launch_gwide_batched(hipStream_t s, int wb, const GWideArgs &g, uint32_t count);
void launch_fill_basis_batched(hipStream_t s, float2 *a, uint64_t dim, uint32_t count);
void launch_zero_batched(hipStream_t s, float2 *a, uint64_t dim, uint32_t count);
void launch_pauli_accum_tile_batched(hipStream_t s, const AccumTileArgs &a, uint32_t count);
void launch_pauli_accum_batched(hipStream_t s, const AccumArgs &a, uint32_t count);
// Workgroups per instance of the batched reductions below (partials hold
// count * blocks * 2 doubles for launch_pauli_batched, count * kAdjMaxRot *
// blocks * 2 for launch_adj_pass_batched).
unsigned pauli_batched_blocks(uint64_t dim);
unsigned adj_batched_blocks(const Layout &lay);
// Workgroups per instance of launch_adj_tile_batched (partials: count *
// kAdjTileMaxRot * blocks * 2 doubles).
unsigned adj_tile_batched_blocks(const Layout &lay);
void launch_adj_tile_batched(hipStream_t s, const AdjTileArgs &a, double *partials, double *out,
                             uint64_t ostride, uint32_t count);
// out[2 i], out[2 i + 1] = instance i's sum (as launch_pauli).
void launch_pauli_batched(hipStream_t s, const float2 *bra, const float2 *ket, uint64_t dim,
                          uint64_t flip, uint64_t sign, double *partials, double *out,
                          uint32_t count);
// Instance i's rotation overlaps to out + i * ostride + 2 r.
void launch_adj_pass_batched(hipStream_t s, const AdjPassArgs &a, double *partials, double *out,
                             uint64_t ostride, uint32_t count);
// Per instance i: energy_out[inst[i]] = eslots[2 i]; for every rotation r of
// the sweep, grad_out[gbase[i] + (rgate[r] & 0x7fffffff)] = the real part of
// its overlap (rslots[i * 2 nrot + 2 r]) when bit 31 of rgate[r] is set (RY),
// else the imaginary part.
void launch_batched_results(hipStream_t s, const double *eslots, const double *rslots,
                            uint32_t count, uint32_t nrot, const uint32_t *rgate,
                            const uint32_t *inst, const uint32_t *gbase, double *energy_out,
                            double *grad_out);

}  // namespace qmamd
