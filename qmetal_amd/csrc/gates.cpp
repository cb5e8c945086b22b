// gates.cpp -- gate names, FP64 one-qubit matrices and circuit validation.

#include <cmath>

#include "core.h"

namespace qmamd {

const char *kind_name(uint8_t k) {
  static const char *names[K_COUNT] = {
      "x", "y", "z", "h", "s", "sdg", "t", "tdg", "rx", "ry",
      "rz", "p", "cx", "cy", "cz", "swap", "cp", "ccx", "u3"};
  return k < K_COUNT ? names[k] : "?";
}

// (cos t, sin t), the unit phasor; std::polar(1.0, t) evaluates the same two
// libm calls, which keeps the plan's FP64 matrices bit-identical to the
// reference builder's.
static cd phasor(double t) { return cd(std::cos(t), std::sin(t)); }

// Matrices follow reference src/reference.cpp:60-98. U3(t, p, l) is the
// appended general unitary [[c, -e^{il} s], [e^{ip} s, e^{i(p+l)} c]] with
// c = cos(t/2), s = sin(t/2).
void gate_matrix(uint8_t kind, const double *p, cd m[4]) {
  const double r = 1.0 / std::sqrt(2.0);
  switch (kind) {
    case K_X: m[0] = 0.0; m[1] = 1.0; m[2] = 1.0; m[3] = 0.0; return;
    case K_Y: m[0] = 0.0; m[1] = cd(0, -1); m[2] = cd(0, 1); m[3] = 0.0; return;
    case K_Z: m[0] = 1.0; m[1] = 0.0; m[2] = 0.0; m[3] = -1.0; return;
    case K_H: m[0] = r; m[1] = r; m[2] = r; m[3] = -r; return;
    case K_S: m[0] = 1.0; m[1] = 0.0; m[2] = 0.0; m[3] = cd(0, 1); return;
    case K_SDG: m[0] = 1.0; m[1] = 0.0; m[2] = 0.0; m[3] = cd(0, -1); return;
    case K_T: m[0] = 1.0; m[1] = 0.0; m[2] = 0.0; m[3] = phasor(M_PI / 4); return;
    case K_TDG: m[0] = 1.0; m[1] = 0.0; m[2] = 0.0; m[3] = phasor(-M_PI / 4); return;
    case K_RX: {
      const double c = std::cos(p[0] / 2), s = std::sin(p[0] / 2);
      m[0] = c; m[1] = cd(0, -s); m[2] = cd(0, -s); m[3] = c;
      return;
    }
    case K_RY: {
      const double c = std::cos(p[0] / 2), s = std::sin(p[0] / 2);
      m[0] = c; m[1] = -s; m[2] = s; m[3] = c;
      return;
    }
    case K_RZ:
      m[0] = phasor(-p[0] / 2); m[1] = 0.0; m[2] = 0.0; m[3] = phasor(p[0] / 2);
      return;
    case K_P: m[0] = 1.0; m[1] = 0.0; m[2] = 0.0; m[3] = phasor(p[0]); return;
    case K_U3: {
      const double c = std::cos(p[0] / 2), s = std::sin(p[0] / 2);
      const cd el = phasor(p[2]), ep = phasor(p[1]), epl = phasor(p[1] + p[2]);
      m[0] = c;
      m[1] = cd(-el.real() * s, -el.imag() * s);
      m[2] = cd(ep.real() * s, ep.imag() * s);
      m[3] = cd(epl.real() * c, epl.imag() * c);
      return;
    }
    default:
      fail(std::string("gate matrix: ") + kind_name(kind) +
           " is not a one-qubit gate");
  }
}

void check_circuit(const Circuit &c) {
  for (const Gate &g : c.gates) {
    if (g.kind >= K_COUNT) fail("unknown gate code");
    if (g.nq < 1 || g.nq > 3) fail("gate arity must be 1..3");
    for (int k = 0; k < g.nq; k++)
      if (g.q[k] >= c.n) fail("qubit index out of range");
    if (g.nq >= 2 && g.q[0] == g.q[1])
      fail(std::string(kind_name(g.kind)) + " on identical wires");
    if (g.nq == 3 && (g.q[0] == g.q[2] || g.q[1] == g.q[2]))
      fail(std::string(kind_name(g.kind)) + " on repeated wires");
  }
}

}  // namespace qmamd
