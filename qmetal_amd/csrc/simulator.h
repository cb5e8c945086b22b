// simulator.h -- the MI355X statevector executor behind the C ABI.
//
// One Simulator owns one HBM allocation of 2^n complex64 amplitudes on the
// current HIP device and one stream of default priority. Plan ops are
// enqueued asynchronously, consecutive ops sharing one kernel (one HBM pass)
// where their wires fit a register tile; reads, reductions and samples
// synchronise. No HIP types appear here so the C ABI compiles as plain C++.

#pragma once

#include <cstdint>
#include <memory>
#include <string>
#include <utility>
#include <vector>

#include "core.h"

namespace qmamd {

// 2^34 complex64 amplitudes is one 128 GiB allocation, within the 288 GB of
// HBM on one MI355X. The adjoint gradient needs a second state of the same
// size.
constexpr uint32_t kMaxQubits = 34;

// Device time of one launch and its algorithmic bytes: a plan op (kind
// 0..6) or a merged pass (kinds 7..12: the generic, chain, wide chain, low
// chain, generic low tile and generic wide kernels; bytes of one pass over
// the amplitudes it reads).
struct OpTime {
  uint32_t kind;
  double ms;
  double bytes;
};

// Adjoint energy and gradient of many circuits that share one width, one
// gate structure (kinds and wires in order; angles free) and one
// Hamiltonian, compiled into batched launches: every kernel of the merged
// single-circuit path is one launch over all instances (one grid row each),
// each instance reading its own state and its own region of a shared op
// arena. Built by Simulator::compile_batched; owns its device buffers.
class BatchedAdjoint {
 public:
  BatchedAdjoint();
  ~BatchedAdjoint();
  BatchedAdjoint(const BatchedAdjoint &) = delete;
  BatchedAdjoint &operator=(const BatchedAdjoint &) = delete;
  uint32_t count() const;
  uint32_t num_qubits() const;
  // Enqueues the whole program on `stream` (a hipStream_t): instance i's
  // energy to energy_out[inst[i]] and its rotations' gradients to
  // grad_out[gbase[i] + gate index] (inst, gbase as compiled). Nothing
  // synchronises or allocates, so a graph can capture it.
  void enqueue(void *stream, double *energy_out, double *grad_out) const;

 private:
  friend class Simulator;
  struct Data;
  std::unique_ptr<Data> d_;
};

class Simulator {
 public:
  explicit Simulator(uint32_t n);
  ~Simulator();
  Simulator(const Simulator &) = delete;
  Simulator &operator=(const Simulator &) = delete;

  void reset();
  void run(const Circuit &c, const FuseOpts &o);
  void run_plan(const Plan &p);
  // Merging of consecutive plan ops into shared HBM passes (on by default).
  // Off: one kernel launch per plan op.
  void set_pass_merging(bool on);
  bool pass_merging() const;
  // Placement probe of the state (33 qubits): allocations made (0 when
  // not probed) and the kept state's time ratio of a pass over the top seven
  // wires to one over wires 18..12.
  void placement(uint32_t *attempts, double *ratio) const;
  void apply_gate(const Gate &g);
  void synchronize();

  uint32_t num_qubits() const;
  uint64_t dim() const;
  // One per plan op run (merged or not; none for a diagonal layer without
  // terms), reset, reduction and sample, as the reference counts.
  uint64_t dispatch_count() const;
  // Kernels actually launched (merged passes count once).
  uint64_t launch_count() const;

  // Host copy of the interleaved complex64 state, refreshed on every call and
  // valid until the next call or mutation.
  const float *amplitudes();
  // Interleaved complex64 values of the amplitudes at idx[0..count) into
  // out[0..2 count); fails on an index past the state.
  void amplitudes_at(const uint64_t *idx, size_t count, float *out);

  double abs2sum();
  double norm();
  double expectation(const Pauli &p);
  double expectation(const Hamiltonian &h);
  std::vector<uint64_t> sample(size_t shots, uint64_t seed);
  // The same draws into out[0..shots) with no device allocation: scratch in
  // a fixed per-GPU array (states of up to 30 qubits, up to 2^20 shots).
  void sample_noalloc(size_t shots, uint64_t seed, uint64_t *out);

  std::pair<double, std::vector<double>> energy_and_gradient(
      const Circuit &c, const Hamiltonian &h);

  // Same, with parameters (one per RX/RY/RZ, circuit order) read from and
  // energy / gradient (one per parameter) written to device memory; ordered
  // after and before work on `caller_stream` (a hipStream_t, may be null).
  // The device-resident program for template c and Hamiltonian h (compiled
  // on first use, cached); false when the path does not apply (fewer than
  // 12 qubits, pass merging off, an empty Hamiltonian).
  bool compile_device_program(const Circuit &c, const Hamiltonian &h);
  void energy_and_gradient_device(const Circuit &c, const Hamiltonian &h,
                                  const double *params, size_t nparams,
                                  double *out_energy, double *out_grad,
                                  void *caller_stream);

  // Compiles `circuits` (this simulator's width, identical gate kinds and
  // wires, each differentiable as for energy_and_gradient) under `h` into a
  // batched program (BatchedAdjoint); instance i reports to inst[i] and
  // gbase[i]. Returns null when the circuits' plans differ in structure or
  // need a pass kind the batched kernels do not run (wide or low chain
  // passes, an op that cannot join a merged pass). Allocates the instances'
  // two states each.
  std::unique_ptr<BatchedAdjoint> compile_batched(const std::vector<const Circuit *> &circuits,
                                                  const Hamiltonian &h,
                                                  const std::vector<uint32_t> &inst,
                                                  const std::vector<uint32_t> &gbase);

  // Per-op timing with events on the simulator's stream (off by default).
  void set_timing(bool on);
  std::vector<OpTime> read_timing();

  // Device address of the state (for callers that share device memory).
  void *device_state();

  static bool available();
  static std::string device_name();
  static uint64_t total_memory();
  static uint64_t free_memory();

 private:
  struct Impl;
  std::unique_ptr<Impl> impl_;
};

}  // namespace qmamd
