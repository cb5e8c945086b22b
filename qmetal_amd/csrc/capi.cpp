// capi.cpp -- the C ABI over the executor (reference src/capi.cpp).
//
// Every entry point catches: an exception becomes -1 (or NULL / 0 for
// constructors and getters) plus a thread-local message.

#include "qmetal/qmetal.h"

#include <cstring>
#include <new>
#include <string>

#include "batch.h"
#include "core.h"
#include "schedule.h"
#include "simulator.h"

using namespace qmamd;

namespace {

thread_local std::string g_error;

void set_error(const char *what) { g_error = what ? what : "unknown error"; }

template <typename Fn>
int guard(Fn &&fn) {
  try {
    fn();
    return 0;
  } catch (const std::exception &e) {
    set_error(e.what());
    return -1;
  } catch (...) {
    set_error("unknown exception");
    return -1;
  }
}

FuseOpts to_opts(const qm_fusion_options *o) {
  if (!o) return FuseOpts{};
  FuseOpts f;
  f.windows = o->collapse_1q_windows != 0;
  f.diagonals = o->fuse_diagonals != 0;
  f.zz = o->recognize_zz != 0;
  f.group = o->group_1q;
  f.block_bits = o->block_bits;
  return f;
}

Hamiltonian to_hamiltonian(const qm_hamiltonian *h) {
  Hamiltonian out;
  if (!h) return out;
  if (h->num_terms && (!h->coeffs || !h->x_masks || !h->y_masks || !h->z_masks))
    fail("hamiltonian: null array with non-zero num_terms");
  for (uint32_t i = 0; i < h->num_terms; i++) {
    Term t;
    t.c = h->coeffs[i];
    t.p.x = h->x_masks[i];
    t.p.y = h->y_masks[i];
    t.p.z = h->z_masks[i];
    out.push_back(t);
  }
  return out;
}

}  // namespace

struct qm_circuit_s {
  Circuit c;
};
struct qm_simulator_s {
  Simulator *s;
};
struct qm_plan_s {
  Plan p;
};
struct qm_batch_s {
  Batch *b;
};

extern "C" {

void qm_version(int *major, int *minor, int *patch) {
  if (major) *major = QMETAL_VERSION_MAJOR;
  if (minor) *minor = QMETAL_VERSION_MINOR;
  if (patch) *patch = QMETAL_VERSION_PATCH;
}

const char *qm_last_error(void) { return g_error.c_str(); }

int qm_is_available(void) {
  try {
    return Simulator::available() ? 1 : 0;
  } catch (...) {
    return 0;
  }
}

const char *qm_device_name(void) {
  static thread_local std::string name;
  try {
    name = Simulator::device_name();
  } catch (...) {
    name = "none";
  }
  return name.c_str();
}

uint64_t qm_max_buffer_length(void) {
  try { return Simulator::total_memory(); } catch (...) { return 0; }
}

uint64_t qm_recommended_working_set(void) {
  try { return Simulator::free_memory(); } catch (...) { return 0; }
}

uint32_t qm_max_qubits(void) { return kMaxQubits; }

qm_circuit qm_circuit_create(uint32_t num_qubits) {
  auto *c = new (std::nothrow) qm_circuit_s;
  if (!c) {
    set_error("out of memory");
    return nullptr;
  }
  c->c.n = num_qubits;
  return c;
}

void qm_circuit_destroy(qm_circuit c) { delete c; }

// Validation as reference src/capi.cpp:137-158: arity 1..3, at most three
// parameters, non-NULL arrays, wires in range.
int qm_circuit_add(qm_circuit c, qm_gate_kind kind, const uint32_t *qubits,
                   uint32_t nq, const double *params, uint32_t np) {
  if (!c) {
    set_error("null circuit");
    return -1;
  }
  return guard([&] {
    if (nq == 0 || nq > 3) fail("gate arity must be 1..3");
    if (np > 3) fail("at most three parameters");
    if (!qubits) fail("null qubit array");
    if (np && !params) fail("null parameter array");
    if (int(kind) < 0 || int(kind) >= K_COUNT) fail("unknown gate code");
    Gate g;
    g.kind = uint8_t(kind);
    g.nq = uint8_t(nq);
    g.np = uint8_t(np);
    for (uint32_t i = 0; i < nq; i++) {
      if (qubits[i] >= c->c.n) fail("qubit index out of range");
      g.q[i] = qubits[i];
    }
    for (uint32_t i = 0; i < np; i++) g.p[i] = params[i];
    c->c.gates.push_back(g);
  });
}

size_t qm_circuit_size(qm_circuit c) { return c ? c->c.gates.size() : 0; }

size_t qm_circuit_plan_size(qm_circuit c, int fuse) {
  if (!c) return 0;
  try {
    return build_plan(c->c, fuse ? FuseOpts{} : FuseOpts::none()).ops.size();
  } catch (const std::exception &e) {
    set_error(e.what());
    return 0;
  }
}

qm_simulator qm_create(uint32_t num_qubits) {
  auto *h = new (std::nothrow) qm_simulator_s;
  if (!h) {
    set_error("out of memory");
    return nullptr;
  }
  h->s = nullptr;
  try {
    h->s = new Simulator(num_qubits);
  } catch (const std::exception &e) {
    set_error(e.what());
    delete h;
    return nullptr;
  }
  return h;
}

void qm_destroy(qm_simulator s) {
  if (!s) return;
  delete s->s;
  delete s;
}

int qm_reset(qm_simulator s) {
  if (!s) {
    set_error("null simulator");
    return -1;
  }
  return guard([&] { s->s->reset(); });
}

int qm_run(qm_simulator s, qm_circuit c) {
  if (!s || !c) {
    set_error("null argument");
    return -1;
  }
  return guard([&] { s->s->run(c->c, FuseOpts{}); });
}

int qm_run_unfused(qm_simulator s, qm_circuit c) {
  if (!s || !c) {
    set_error("null argument");
    return -1;
  }
  return guard([&] { s->s->run(c->c, FuseOpts::none()); });
}

int qm_run_options(qm_simulator s, qm_circuit c, const qm_fusion_options *o) {
  if (!s || !c) {
    set_error("null argument");
    return -1;
  }
  return guard([&] { s->s->run(c->c, to_opts(o)); });
}

int qm_synchronize(qm_simulator s) {
  if (!s) {
    set_error("null simulator");
    return -1;
  }
  return guard([&] { s->s->synchronize(); });
}

uint32_t qm_num_qubits(qm_simulator s) { return s ? s->s->num_qubits() : 0; }
uint64_t qm_dim(qm_simulator s) { return s ? s->s->dim() : 0; }
uint64_t qm_dispatch_count(qm_simulator s) { return s ? s->s->dispatch_count() : 0; }
uint64_t qm_launch_count(qm_simulator s) { return s ? s->s->launch_count() : 0; }

const float *qm_amplitudes(qm_simulator s) {
  if (!s) {
    set_error("null simulator");
    return nullptr;
  }
  try {
    return s->s->amplitudes();
  } catch (const std::exception &e) {
    set_error(e.what());
    return nullptr;
  }
}

int qm_norm(qm_simulator s, double *out) {
  if (!s || !out) {
    set_error("null argument");
    return -1;
  }
  return guard([&] { *out = s->s->norm(); });
}

int qm_expectation_pauli(qm_simulator s, uint64_t x, uint64_t y, uint64_t z,
                         double *out) {
  if (!s || !out) {
    set_error("null argument");
    return -1;
  }
  return guard([&] {
    Pauli p;
    p.x = x;
    p.y = y;
    p.z = z;
    *out = s->s->expectation(p);
  });
}

int qm_expectation(qm_simulator s, const qm_hamiltonian *h, double *out) {
  if (!s || !h || !out) {
    set_error("null argument");
    return -1;
  }
  return guard([&] { *out = s->s->expectation(to_hamiltonian(h)); });
}

int qm_sample(qm_simulator s, size_t shots, uint64_t seed, uint64_t *out) {
  if (!s || (shots && !out)) {
    set_error("null argument");
    return -1;
  }
  return guard([&] {
    auto v = s->s->sample(shots, seed);
    if (!v.empty()) std::memcpy(out, v.data(), v.size() * sizeof(uint64_t));
  });
}

int qm_energy_and_gradient(qm_simulator s, qm_circuit c, const qm_hamiltonian *h,
                           double *out_energy, double *out_grad) {
  if (!s || !c || !h || !out_grad) {
    set_error("null argument");
    return -1;
  }
  return guard([&] {
    auto eg = s->s->energy_and_gradient(c->c, to_hamiltonian(h));
    if (out_energy) *out_energy = eg.first;
    if (!eg.second.empty())
      std::memcpy(out_grad, eg.second.data(), eg.second.size() * sizeof(double));
  });
}

int qm_energy_and_gradient_device(qm_simulator s, qm_circuit c, const qm_hamiltonian *h,
                                  const double *params, size_t num_params,
                                  double *out_energy, double *out_grad, void *stream) {
  if (!s || !c || !h || !out_energy || (num_params && !params)) {
    set_error("null argument");
    return -1;
  }
  return guard([&] {
    s->s->energy_and_gradient_device(c->c, to_hamiltonian(h), params, num_params, out_energy,
                                     out_grad, stream);
  });
}

qm_batch qm_batch_create(void) {
  auto *h = new (std::nothrow) qm_batch_s;
  if (!h) {
    set_error("out of memory");
    return nullptr;
  }
  h->b = nullptr;
  try {
    if (!Simulator::available()) fail("batch: no HIP device");
    h->b = new Batch();
  } catch (const std::exception &e) {
    set_error(e.what());
    delete h;
    return nullptr;
  }
  return h;
}

void qm_batch_destroy(qm_batch b) {
  if (!b) return;
  delete b->b;
  delete b;
}

int qm_batch_prepare(qm_batch b, size_t count, const qm_circuit *circuits,
                     const qm_hamiltonian *hams) {
  if (!b || (count && (!circuits || !hams))) {
    set_error("null argument");
    return -1;
  }
  return guard([&] {
    std::vector<Hamiltonian> hs(count);
    std::vector<const Circuit *> cs(count);
    std::vector<const Hamiltonian *> hp(count);
    for (size_t i = 0; i < count; i++) {
      if (!circuits[i]) fail("null circuit in batch");
      cs[i] = &circuits[i]->c;
      hs[i] = to_hamiltonian(&hams[i]);
      hp[i] = &hs[i];
    }
    b->b->prepare(cs, hp);
  });
}

int qm_batch_autotune(qm_batch b, uint32_t *threads, uint32_t *per_cu, uint32_t *lds_bits) {
  if (!b) {
    set_error("null batch");
    return -1;
  }
  return guard([&] {
    BatchConfig c = b->b->autotune();
    if (threads) *threads = c.threads;
    if (per_cu) *per_cu = c.per_cu;
    if (lds_bits) *lds_bits = c.lds_bits;
  });
}

int qm_batch_set_config(qm_batch b, uint32_t threads, uint32_t per_cu, uint32_t lds_bits) {
  if (!b) {
    set_error("null batch");
    return -1;
  }
  return guard([&] {
    BatchConfig c;
    c.threads = threads;
    c.per_cu = per_cu;
    c.lds_bits = lds_bits;
    b->b->set_config(c);
  });
}

int qm_batch_set_spread(qm_batch b, uint32_t min_qubits) {
  if (!b) {
    set_error("null batch");
    return -1;
  }
  return guard([&] { b->b->set_spread(min_qubits); });
}

uint32_t qm_batch_spread(qm_batch b) { return b ? b->b->config().spread_bits : 0; }

int qm_batch_batched_instances(qm_batch b, size_t *out) {
  if (!b || !out) {
    set_error("null argument");
    return -1;
  }
  return guard([&] { *out = b->b->batched_instances(); });
}

int qm_batch_run(qm_batch b, double *out_energy, double *out_grad, double *out_ms) {
  if (!b) {
    set_error("null batch");
    return -1;
  }
  return guard([&] {
    b->b->run(out_energy, out_grad);
    if (out_ms) *out_ms = b->b->last_ms();
  });
}

int qm_set_timing(qm_simulator s, int enable) {
  if (!s) {
    set_error("null simulator");
    return -1;
  }
  return guard([&] { s->s->set_timing(enable != 0); });
}

int qm_read_timing(qm_simulator s, uint32_t *kinds, double *ms, double *bytes,
                   size_t cap, size_t *count) {
  if (!s || !count) {
    set_error("null argument");
    return -1;
  }
  return guard([&] {
    auto t = s->s->read_timing();
    *count = t.size();
    for (size_t k = 0; k < t.size() && k < cap; k++) {
      if (kinds) kinds[k] = t[k].kind;
      if (ms) ms[k] = t[k].ms;
      if (bytes) bytes[k] = t[k].bytes;
    }
  });
}

qm_plan qm_plan_build(qm_circuit c, const qm_fusion_options *o) {
  if (!c) {
    set_error("null circuit");
    return nullptr;
  }
  auto *p = new (std::nothrow) qm_plan_s;
  if (!p) {
    set_error("out of memory");
    return nullptr;
  }
  try {
    p->p = build_plan(c->c, to_opts(o));
  } catch (const std::exception &e) {
    set_error(e.what());
    delete p;
    return nullptr;
  }
  return p;
}

void qm_plan_destroy(qm_plan p) { delete p; }

size_t qm_plan_size(qm_plan p) { return p ? p->p.ops.size() : 0; }

int qm_plan_get(qm_plan p, size_t k, qm_plan_op *out) {
  if (!p || !out) {
    set_error("null argument");
    return -1;
  }
  return guard([&] {
    if (k >= p->p.ops.size()) fail("plan op index out of range");
    const PlanOp &op = p->p.ops[k];
    std::memset(out, 0, sizeof(*out));
    out->kind = uint32_t(op.kind);
    for (int i = 0; i < 3; i++) out->qubits[i] = op.q[i];
    for (int i = 0; i < 4; i++) {
      out->matrix[2 * i] = op.m[i].real();
      out->matrix[2 * i + 1] = op.m[i].imag();
    }
    out->global_phase = op.global;
    out->block_bits = op.block_bits;
    out->num_terms = op.masks.size();
    out->num_group = op.wires.size();
  });
}

int qm_plan_get_arrays(qm_plan p, size_t k, uint32_t *gq, double *gm,
                       uint64_t *masks, double *angles) {
  if (!p) {
    set_error("null plan");
    return -1;
  }
  return guard([&] {
    if (k >= p->p.ops.size()) fail("plan op index out of range");
    const PlanOp &op = p->p.ops[k];
    for (size_t i = 0; i < op.wires.size(); i++) {
      if (gq) gq[i] = op.wires[i];
      if (gm)
        for (int j = 0; j < 4; j++) {
          gm[8 * i + 2 * j] = op.mats[4 * i + j].real();
          gm[8 * i + 2 * j + 1] = op.mats[4 * i + j].imag();
        }
    }
    for (size_t i = 0; i < op.masks.size(); i++) {
      if (masks) masks[i] = op.masks[i];
      if (angles) angles[i] = op.angles[i];
    }
  });
}

int qm_plan_passes(qm_plan p, uint32_t *pass_of_op, size_t *num_passes) {
  if (!p || !num_passes) {
    set_error("null argument");
    return -1;
  }
  return guard([&] {
    const std::vector<Pass> passes = schedule_passes(p->p);
    if (pass_of_op)
      for (size_t k = 0; k < p->p.ops.size(); k++) pass_of_op[k] = UINT32_MAX;
    for (size_t i = 0; i < passes.size(); i++)
      for (uint32_t k : passes[i].ops)
        if (pass_of_op && pass_of_op[k] == UINT32_MAX) pass_of_op[k] = uint32_t(i);
    *num_passes = passes.size();
  });
}

int qm_plan_schedule(qm_plan p, uint32_t *entry_op, uint64_t *entry_wires,
                     uint32_t *entry_pass, size_t cap, size_t *count) {
  if (!p || !count) {
    set_error("null argument");
    return -1;
  }
  return guard([&] {
    const std::vector<Pass> passes = schedule_passes(p->p);
    size_t e = 0;
    for (size_t i = 0; i < passes.size(); i++)
      for (size_t j = 0; j < passes[i].ops.size(); j++, e++) {
        if (e >= cap) continue;
        if (entry_op) entry_op[e] = passes[i].ops[j];
        if (entry_wires) entry_wires[e] = passes[i].wmask[j];
        if (entry_pass) entry_pass[e] = uint32_t(i);
      }
    *count = e;
  });
}

int qm_amplitudes_at(qm_simulator s, const uint64_t *indices, size_t count, float *out) {
  if (!s || (count && (!indices || !out))) {
    set_error("null argument");
    return -1;
  }
  return guard([&] { s->s->amplitudes_at(indices, count, out); });
}

int qm_sample_noalloc(qm_simulator s, size_t shots, uint64_t seed, uint64_t *out) {
  if (!s || (shots && !out)) {
    set_error("null argument");
    return -1;
  }
  return guard([&] { s->s->sample_noalloc(shots, seed, out); });
}

int qm_state_placement(qm_simulator s, uint32_t *attempts, double *ratio) {
  if (!s || !attempts || !ratio) {
    set_error("null argument");
    return -1;
  }
  return guard([&] { s->s->placement(attempts, ratio); });
}

int qm_set_pass_merging(qm_simulator s, int enable) {
  if (!s) {
    set_error("null simulator");
    return -1;
  }
  return guard([&] { s->s->set_pass_merging(enable != 0); });
}

}  // extern "C"
