// core.h -- gate stream, Pauli algebra and the Plan IR of the MI355X
// statevector simulator. Pure C++17; no HIP types, so the plan builder is
// tested on the CPU before a kernel runs.
//
// Convention (reference include/qmetal/circuit.h:7-8): LSB-first, qubit q is
// bit q of the basis index, amplitude stride 1 << q.

#pragma once

#include <complex>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

namespace qmamd {

using cd = std::complex<double>;

// Gate codes are the C ABI's qm_gate_kind values (include/qmetal/qmetal.h);
// one numbering end to end. U3 = 18 is appended.
enum Kind : uint8_t {
  K_X = 0, K_Y, K_Z, K_H, K_S, K_SDG, K_T, K_TDG,
  K_RX, K_RY, K_RZ, K_P,
  K_CX, K_CY, K_CZ, K_SWAP, K_CP, K_CCX,
  K_U3,
  K_COUNT
};

const char *kind_name(uint8_t k);

// One gate record (reference include/qmetal/circuit.h:45-51): controlled
// gates list (control..., target).
struct Gate {
  uint8_t kind = 0;
  uint8_t nq = 0;
  uint8_t np = 0;
  uint32_t q[3] = {0, 0, 0};
  double p[3] = {0, 0, 0};
};

struct Circuit {
  uint32_t n = 0;
  std::vector<Gate> gates;
};

// Row-major FP64 2x2 of a one-qubit gate (reference src/reference.cpp:60-98).
// Throws for a kind that is not a one-qubit gate.
void gate_matrix(uint8_t kind, const double *p, cd m[4]);

// Pauli string as bitmasks (reference include/qmetal/pauli.h:25-40).
struct Pauli {
  uint64_t x = 0, y = 0, z = 0;
  bool valid(uint32_t n) const {
    const uint64_t all = n >= 64 ? ~0ull : ((1ull << n) - 1ull);
    if ((x | y | z) & ~all) return false;
    return !((x & y) | (x & z) | (y & z));
  }
};

struct Term {
  double c = 1.0;
  Pauli p;
};
using Hamiltonian = std::vector<Term>;

// Plan IR (reference include/qmetal/plan.h:27-68). OpKind values keep the
// reference's order so dumped plans compare by integer.
enum class OpKind : uint8_t {
  Dense1q = 0,
  Group = 1,       // K independent 2x2s in one pass, wires ascending
  Blocked = 2,     // a run of 2x2s on wires below block_bits
  Controlled = 3,  // 2x2 on q[1] where q[0] is set
  Diagonal = 4,    // amp_i *= exp(i phi(i)), phi = global + sum angle[t]*[mask_t in i]
  Swap = 5,
  CCX = 6,
};

struct PlanOp {
  OpKind kind = OpKind::Dense1q;
  uint32_t q[3] = {0, 0, 0};
  cd m[4] = {1, 0, 0, 1};
  std::vector<uint64_t> masks;
  std::vector<double> angles;
  double global = 0.0;
  std::vector<uint32_t> wires;  // Group / Blocked
  std::vector<cd> mats;         // 4 per wire, row-major
  uint32_t block_bits = 0;
};

struct Plan {
  uint32_t n = 0;
  std::vector<PlanOp> ops;
};

// Fusion switches (reference include/qmetal/plan.h:76-94).
struct FuseOpts {
  bool windows = true;    // per-wire 2x2 products inside a 1q window
  bool diagonals = true;  // runs of commuting diagonal gates
  bool zz = true;         // CX . RZ . CX on the target is a ZZ coupling
  uint32_t group = 4;     // wires per register group; 1 disables grouping
  uint32_t block_bits = 0;

  static FuseOpts none() { return FuseOpts{false, false, false, 1, 0}; }
  static FuseOpts blocked(uint32_t b) {
    FuseOpts o;
    o.block_bits = b;
    return o;
  }
};

bool is_diagonal(const Gate &g);
double diagonal_term(const Gate &g, std::vector<uint64_t> &masks,
                     std::vector<double> &angles);
Plan build_plan(const Circuit &c, const FuseOpts &o);

// Executor-side validation shared by every run path: wires in range, no
// repeated wire within a gate, known kind.
void check_circuit(const Circuit &c);

[[noreturn]] inline void fail(const std::string &what) {
  throw std::runtime_error(what);
}

}  // namespace qmamd
