// batch.h -- many small circuits' energies and adjoint gradients in one
// persistent launch (and, for wide instances, lockstep spread launches).
//
// Each instance is compiled on the host into a short program (forward plan,
// co-state build, energy, backward sweep). A persistent kernel of a few
// workgroups per CU pulls instances from an atomic queue, largest first, and
// runs each program to completion inside one workgroup: both states in LDS
// when they fit, otherwise in that workgroup's slot of an HBM scratch buffer.
// The launch (queue reset + persistent kernel) is captured once in a
// hipGraph and replayed; block size, workgroups per CU and the LDS cut-off are
// picked by timing the candidates on the batch itself.

#pragma once

#include <cstdint>
#include <memory>
#include <vector>

#include "core.h"

namespace qmamd {

struct BatchConfig {
  uint32_t threads = 512;   // workgroup size (64..512)
  uint32_t per_cu = 1;      // persistent workgroups per CU
  uint32_t lds_bits = 12;   // instances with n <= lds_bits run in LDS
  // Instances with n >= spread_bits (0: none) run spread: several
  // workgroups per instance, one launch per program step for all of them,
  // before the persistent kernel takes the rest. autotune picks it by timing
  // the whole batch (none, or from 1, 3 or 5 below the widest instance).
  uint32_t spread_bits = 0;
};

// Narrowest spread width autotune tries (states wider than any LDS cut-off).
constexpr uint32_t kSpreadMinQubits = 14;

class Batch {
 public:
  Batch();
  ~Batch();
  Batch(const Batch &) = delete;
  Batch &operator=(const Batch &) = delete;

  // Compile the instances and upload their programs. Widths 1..kBatchMaxQubits.
  void prepare(const std::vector<const Circuit *> &circuits,
               const std::vector<const Hamiltonian *> &hams);
  // Pick the launch configuration by timing each candidate (twice) on the
  // prepared batch; returns the chosen one.
  BatchConfig autotune();
  // threads, per_cu and lds_bits (spread_bits is kept).
  void set_config(const BatchConfig &c);
  // Instances with n >= bits run spread (0: none).
  void set_spread(uint32_t bits);
  BatchConfig config() const;
  // Instances that run as batched merged-pass programs (one per class of
  // spread instances sharing width, gate structure and Hamiltonian) under
  // the current spread width; the other spread instances take the lockstep
  // spread steps.
  size_t batched_instances();
  // One replay of the captured launch; results copied back.
  void run(double *energies, double *grads);
  // Device milliseconds of the last run (events around the graph).
  double last_ms() const;
  size_t total_gates() const;

 private:
  struct Impl;
  std::unique_ptr<Impl> impl_;
};

constexpr uint32_t kBatchMaxQubits = 20;

}  // namespace qmamd
