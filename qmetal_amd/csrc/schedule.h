// schedule.h -- executor-level scheduling below the Plan: consecutive plan
// ops are merged into one read-modify-write pass over HBM when the wires they
// move fit in one thread's register tile.
//
// A merged pass holds 2^NL amplitudes per thread: index bit 0 (inside one
// 16-byte load) plus up to kPassMaxReg register wires. Every dense 2x2 of the
// pass acts on one of those wires; controls on any other bit become
// predicates on the thread's base index; diagonal layers are elementwise and
// need no register wire at all. Bits that every op of the pass requires to be
// 1 (controls, common diagonal mask bits, at or above bit kPassFixedMin) are
// fixed: the pass never reads amplitudes with such a bit clear.
//
// The Plan itself (and qm_circuit_plan_size) is unchanged; only the number of
// kernel launches and HBM passes drops. Pure C++ so the grouping is tested on
// the CPU.

#pragma once

#include <cstdint>
#include <vector>

#include "core.h"

namespace qmamd {

constexpr uint32_t kPassMaxReg = 4;     // register wires beyond bit 0 (NL <= 5)
constexpr uint32_t kPassFixedMin = 4;   // 2^4 complex64 = one 128-byte line
constexpr uint32_t kPassMinQubits = 5;  // below this a plan runs op by op
constexpr uint32_t kPassMaxOps = 16;    // plan ops per merged pass

struct Pass {
  bool merged = false;         // false: ops[0] alone through its own kernel
  bool wide = false;           // merged: six chain steps over a workgroup tile
  bool low = false;            // merged: chain steps over a low 2^12 tile
  bool tile = false;           // merged: any ops on wires 0..11 over a low
                               // 2^12 tile, in segments
  bool gwide = false;          // merged: any ops on wires >= 7 over a wide
                               // chain pass's tile (reg: its wires), in two
                               // segments around the exchange
  std::vector<uint32_t> seg_len;               // tile: plan ops per segment
  std::vector<std::vector<uint32_t>> seg_reg;  // tile: register wires per segment
  std::vector<uint32_t> ops;   // plan-op indices, in execution order
  std::vector<uint64_t> wmask; // per entry of ops: for a Dense1qGroup the
                               // wires of it this pass applies (0: all)
  std::vector<uint32_t> reg;   // register wires (merged), first use order;
                               // wide and low: the step wires in order
  uint64_t fixed = 0;          // fixed-one bits (merged)
};

// Wide passes: six or seven chain steps -- each an uncontrolled Dense1q on a
// distinct wire at or above kWideMinWire, optionally followed by a phase
// layer whose every mask (weight 1 or 2) holds that wire, with no global
// phase -- over a workgroup tile of 2^(7+steps) amplitudes whose six lowest
// free positions are the lanes of a 1 KB run. States below kWideMinQubits
// use ordinary passes.
constexpr uint32_t kWideMinSteps = 6;
constexpr uint32_t kWideMaxSteps = 7;
constexpr uint32_t kWideMinWire = 7;
constexpr uint32_t kWideMinQubits = 16;

// Number of plan ops from `k` that form a wide pass (0 if none); `wires`
// gets the step wires.
uint32_t wide_chain_at(const Plan &p, uint32_t k, std::vector<uint32_t> &wires);

// Low passes: kLowMinSteps to kLowMaxWire chain steps (as for wide passes)
// on distinct wires 1..kLowMaxWire, optionally followed by one on wire 0,
// over tiles of contiguous amplitudes that the waves of a workgroup permute
// through LDS: 2^12 amplitudes (two waves) when the wires stay within
// 1..kLowTwoWaveWire, else 2^13 (four waves; states of 13 qubits or more).
// States below kLowMinQubits use ordinary passes.
constexpr uint32_t kLowMaxWire = 12;
constexpr uint32_t kLowTwoWaveWire = 11;
constexpr uint32_t kLowMinSteps = 5;
constexpr uint32_t kLowMinQubits = 12;
// From this many qubits a low pass always takes the four-wave tile.
constexpr uint32_t kLowFourWaveQubits = 24;

// Number of plan ops from `k` that form a low pass (0 if none); `wires` gets
// the step wires (a final 0 for a step on wire 0).
uint32_t low_chain_at(const Plan &p, uint32_t k, std::vector<uint32_t> &wires);

// Between chain passes the plan's ops are scheduled as a dependency graph:
// two ops keep their plan order when their wires overlap, unless both are
// phase layers (those commute), and a Dense1qGroup's 2x2s are separate
// units. Each pass is drafted greedily from the ops whose predecessors have
// run, in three shapes, and the one taking the most ops wins (ties: the
// register tile):
//   register tile: at most kPassMaxReg wires moved (bit 0 aside);
//   generic low tile: moved wires in 1..kTileMaxWire, in at most
//     kTileMaxSegs segments of kPassMaxReg register wires, over tiles of
//     2^(kTileMaxWire+1) contiguous amplitudes permuted through LDS between
//     segments (taken with two segments or more);
//   generic wide: moved wires at or above kWideMinWire, over the tile of a
//     wide chain pass: ops on at most four register wires, then, after the
//     exchange, on at most three new wires plus the register wires that stay
//     (taken when the second part is used).
// At most kTileMaxOps units per pass; an op that cannot join a merged pass,
// and a pass of one whole op, runs through the op's own kernel.
constexpr uint32_t kTileMaxWire = 11;
constexpr uint32_t kTileMinQubits = 12;
constexpr uint32_t kTileMaxSegs = 8;
constexpr uint32_t kTileMaxOps = 64;

// Whether plan op `op` can join a merged pass.
bool pass_mergeable(const PlanOp &op);
// Wires (other than 0) that must be register-local for `op`.
std::vector<uint32_t> pass_wires(const PlanOp &op);
// Bits `op` alone requires to be 1 (controls and common diagonal mask bits at
// or above kPassFixedMin; a diagonal with a non-zero global phase has none).
uint64_t pass_fixed(const PlanOp &op);
// HBM bytes of one pass over a 2^n state with fixed bits `fixed`.
double pass_bytes(uint32_t n, uint64_t fixed);
// HBM bytes `op` moves when it runs alone through its own kernel.
double op_alone_bytes(uint32_t n, const PlanOp &op);

// The passes of a plan, in execution order.
std::vector<Pass> schedule_passes(const Plan &p);

}  // namespace qmamd
