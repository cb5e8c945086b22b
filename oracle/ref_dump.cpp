// ref_dump.cpp -- a driver around the reference's own CPU core (plan builder,
// complex128 oracle, Pauli algebra and circuit generators), compiled from the
// reference sources by oracle/Makefile into oracle/_ref/ref_dump.
//
// TEST INFRASTRUCTURE ONLY: it writes golden fixtures (gate streams, plans,
// amplitudes, expectations) that tests/golden/make_golden.py commits, and
// times the reference oracle for the CPU baseline leg of bench.py. Nothing
// in the product links or runs it.
//
// Usage:
//   ref_dump frozen <name> <n> [amps]   gate stream of a reference generator
//   ref_dump stream [amps]              gate stream read from stdin:
//                                       line 1: n; then per gate
//                                       "kind nq q0 q1 q2 p0 p1 p2"
//   ref_dump time <k>                   same stdin stream; the reference
//                                       oracle runs its first k gates
//                                       untimed, then the rest gate by gate
//                                       (Reference::run, one thread), and
//                                       prints {"n", "gates", "seconds",
//                                       "norm", "first", "last"} (the
//                                       first and last amplitudes)
// Otherwise prints one JSON object: the gates, the plans for six FusionOptions and,
// with "amps", the oracle's amplitudes (n <= 14) and a few expectations.

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "qmetal/circuits.h"
#include "qmetal/pauli.h"
#include "qmetal/plan.h"
#include "qmetal/reference.h"

using namespace qmetal;

static void hexd(double v) { printf("\"%a\"", v); }

static void dump_gates(const Circuit &c) {
  printf("\"gates\":[");
  for (size_t g = 0; g < c.ops.size(); g++) {
    const Gate &x = c.ops[g];
    printf("%s[%d,%d,%d,[%u,%u,%u],[", g ? "," : "", int(x.kind),
           int(x.num_qubits), int(x.num_params), x.qubits[0], x.qubits[1],
           x.qubits[2]);
    for (int k = 0; k < 3; k++) {
      if (k) printf(",");
      hexd(x.params[k]);
    }
    printf("]]");
  }
  printf("]");
}

static void dump_plan(const Plan &p) {
  printf("[");
  for (size_t k = 0; k < p.ops.size(); k++) {
    const PlanOp &op = p.ops[k];
    printf("%s{\"kind\":%d,\"q\":[%u,%u,%u],\"m\":[", k ? "," : "",
           int(op.kind), op.qubits[0], op.qubits[1], op.qubits[2]);
    for (int i = 0; i < 4; i++) {
      if (i) printf(",");
      hexd(op.matrix[i].real());
      printf(",");
      hexd(op.matrix[i].imag());
    }
    printf("],\"gq\":[");
    for (size_t i = 0; i < op.group_qubits.size(); i++)
      printf("%s%u", i ? "," : "", op.group_qubits[i]);
    printf("],\"gm\":[");
    for (size_t i = 0; i < op.group_matrices.size(); i++) {
      if (i) printf(",");
      hexd(op.group_matrices[i].real());
      printf(",");
      hexd(op.group_matrices[i].imag());
    }
    printf("],\"masks\":[");
    for (size_t i = 0; i < op.masks.size(); i++)
      printf("%s\"%llx\"", i ? "," : "", (unsigned long long)op.masks[i]);
    printf("],\"angles\":[");
    for (size_t i = 0; i < op.angles.size(); i++) {
      if (i) printf(",");
      hexd(op.angles[i]);
    }
    printf("],\"global\":");
    hexd(op.global_phase);
    printf(",\"bb\":%u}", op.block_bits);
  }
  printf("]");
}

static void dump_all(const Circuit &c, bool amps) {
  printf("{\"n\":%u,", c.num_qubits);
  dump_gates(c);
  struct Named { const char *name; FusionOptions o; };
  const Named opts[] = {
      {"fused", FusionOptions{}},
      {"none", FusionOptions::none()},
      {"diag_only", FusionOptions{false, true, true, 4, 0}},
      {"win_only", FusionOptions{true, false, false, 4, 0}},
      {"blocked8", FusionOptions::blocked(8)},
      {"blocked10", FusionOptions::blocked(10)},
  };
  printf(",\"plans\":{");
  bool first = true;
  for (const Named &nm : opts) {
    printf("%s\"%s\":", first ? "" : ",", nm.name);
    first = false;
    try {
      dump_plan(build_plan(c, nm.o));
    } catch (const std::exception &e) {
      printf("{\"error\":\"%s\"}", e.what());
    }
  }
  printf("}");
  if (amps && c.num_qubits <= 14) {
    try {
      Reference r(c.num_qubits);
      r.run(c);
      printf(",\"amps\":[");
      for (size_t i = 0; i < r.dim(); i++) {
        if (i) printf(",");
        hexd(r.state()[i].real());
        printf(",");
        hexd(r.state()[i].imag());
      }
      printf("],\"norm\":");
      hexd(r.norm());
      // Single-Z expectations and one all-X / one mixed-Y string.
      printf(",\"expz\":[");
      for (uint32_t q = 0; q < c.num_qubits; q++) {
        if (q) printf(",");
        hexd(r.expectation(PauliString::single('Z', q)));
      }
      printf("]");
      PauliString mixed;
      for (uint32_t q = 0; q < c.num_qubits; q++) {
        if (q % 3 == 0) mixed.x_mask |= 1ull << q;
        if (q % 3 == 1) mixed.y_mask |= 1ull << q;
        if (q % 3 == 2) mixed.z_mask |= 1ull << q;
      }
      printf(",\"exp_mixed\":");
      hexd(r.expectation(mixed));
    } catch (const std::exception &e) {
      printf(",\"amps_error\":\"%s\"", e.what());
    }
  }
  printf("}\n");
}

// One gate stream from stdin: line 1 n, then "kind nq q0 q1 q2 p0 p1 p2".
bool read_stream(Circuit &c) {
  char buf[512];
  if (!fgets(buf, sizeof buf, stdin)) return false;
  c.num_qubits = uint32_t(atoi(buf));
  while (fgets(buf, sizeof buf, stdin)) {
    if (buf[0] == '\n' || buf[0] == 0) continue;
    Gate g{};
    int kind, nq;
    unsigned q0, q1, q2;
    char p0[64], p1[64], p2[64];
    if (sscanf(buf, "%d %d %u %u %u %63s %63s %63s", &kind, &nq, &q0, &q1,
               &q2, p0, p1, p2) != 8)
      return false;
    g.kind = GateKind(kind);
    g.num_qubits = uint8_t(nq);
    g.qubits[0] = q0;
    g.qubits[1] = q1;
    g.qubits[2] = q2;
    g.params[0] = strtod(p0, nullptr);
    g.params[1] = strtod(p1, nullptr);
    g.params[2] = strtod(p2, nullptr);
    g.num_params = (kind >= 8 && kind <= 11) || kind == 16 ? 1 : 0;
    c.add(g);
  }
  return true;
}

int main(int argc, char **argv) {
  if (argc < 2) {
    fprintf(stderr, "usage: ref_dump frozen <name> <n> [amps] | stream [amps] | time <k>\n");
    return 2;
  }
  std::string mode = argv[1];
  if (mode == "frozen") {
    if (argc < 4) return 2;
    std::string name = argv[2];
    uint32_t n = uint32_t(atoi(argv[3]));
    bool amps = argc > 4 && std::string(argv[4]) == "amps";
    for (const auto &spec : extended_benchmarks()) {
      if (spec.name == name) {
        dump_all(spec.build(n), amps);
        return 0;
      }
    }
    fprintf(stderr, "unknown circuit %s\n", name.c_str());
    return 2;
  }
  if (mode == "stream" || mode == "time") {
    Circuit c;
    if (!read_stream(c)) return 3;
    if (mode == "stream") {
      dump_all(c, argc > 2 && std::string(argv[2]) == "amps");
      return 0;
    }
    const size_t k = argc > 2 ? size_t(strtoull(argv[2], nullptr, 10)) : 0;
    if (k > c.ops.size()) return 2;
    Circuit prep, body;
    prep.num_qubits = body.num_qubits = c.num_qubits;
    for (size_t i = 0; i < c.ops.size(); i++) (i < k ? prep : body).add(c.ops[i]);
    Reference r(c.num_qubits);
    r.run(prep);
    const auto t0 = std::chrono::steady_clock::now();
    r.run(body);
    const auto t1 = std::chrono::steady_clock::now();
    // The first and last amplitudes, so a caller can check the timed run
    // against another implementation of the same workload.
    const auto &a0 = r.state()[0], &a1 = r.state()[r.dim() - 1];
    printf("{\"n\":%u,\"gates\":%zu,\"seconds\":%.6f,\"norm\":%.17g,"
           "\"first\":[%.17g,%.17g],\"last\":[%.17g,%.17g]}\n",
           c.num_qubits, body.ops.size(), std::chrono::duration<double>(t1 - t0).count(),
           r.norm(), a0.real(), a0.imag(), a1.real(), a1.imag());
    return 0;
  }
  return 2;
}
