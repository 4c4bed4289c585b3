// CPU driver of the host setup (within_amd/csrc/plan.cpp) for AddressSanitizer
// and UndefinedBehaviorSanitizer builds: builds plans for small designs (2-4
// factors, weighted and unweighted, additive and multiplicative), lays them
// out for the device, serialises the preconditioner, checks the round trip,
// and feeds the deserialiser byte-flipped, truncated and extended copies,
// each of which must either load or be rejected with wa::Error.  Run by
// tests/test_sanitize.py; exits non-zero on a failed check (the sanitizers
// abort on their own findings).
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include "../../within_amd/csrc/plan.hpp"

namespace wa {
// This CPU build factors every dense core on the host (the GPU path is never
// taken: cfg.device is false), so no DeviceDense is ever made.
bool device_anchored_inverse(i64, const u32*, const u32*, const double*, const double*,
                             std::shared_ptr<DeviceDense>&) {
  return false;
}
DeviceDense::~DeviceDense() {}
void DeviceDense::ready() const {}
std::vector<double> DeviceDense::download() const { return {}; }
}  // namespace wa

using namespace wa;

// what a solver built from bytes does after deserialising (runtime.cpp,
// create_solver): every level the bytes cover must be observed, then the
// operator passes are built and the arrays laid out for the device
static void finish_from_bytes(Plan& q) {
  std::vector<char> seen(q.d.n_dofs, 0);
  for (int f = 0; f < q.d.Q; ++f)
    for (uint32_t c : q.d.codes[f]) seen[q.d.offset[f] + c] = 1;
  for (uint32_t g : q.pre.loc_gidx)
    if (!seen[g]) throw Error(9, "the preconditioner covers an unobserved level");
  if (q.precond == 2) q.gram = Csr(), q.has_gram = false;
  build_plan(q);
  (void)make_device_layout(q);
}

static int fail(const char* what) {
  std::fprintf(stderr, "plan_driver: %s\n", what);
  return 1;
}

int main() {
  std::mt19937_64 rng(7);
  int checked = 0, rejected = 0, loaded = 0;
  for (int t = 0; t < 6; ++t) {
    const int Q = 2 + t % 3;
    const i64 n = 400 + 300 * t;
    std::vector<uint32_t> codes((size_t)n * Q);
    for (int q = 0; q < Q; ++q) {
      const uint32_t L = (uint32_t)(q == 0 ? n / 8 : 6 + 5 * q);
      for (i64 i = 0; i < n; ++i) codes[(size_t)q * n + i] = (uint32_t)(rng() % L);
    }
    std::vector<double> w;
    if (t % 2) {
      w.resize(n);
      for (double& x : w) x = 0.5 + (double)(rng() % 1000) / 500.0;
    }
    Plan p;
    p.d = make_design(codes.data(), n, Q, n, w.empty() ? nullptr : w.data());
    p.precond = 1 + t % 2;
    p.explicit_op = t % 3 == 0;
    p.cfg.dense_core = t % 2 ? 8192 : 1;
    p.cfg.ac_split = 1 + t % 2;
    build_plan(p);
    const DeviceLayout lay = make_device_layout(p);
    if (lay.m <= 0 || lay.m > p.d.n_dofs) return fail("device layout size");
    const std::vector<uint8_t> bytes = serialize_precond(p);
    {
      Plan q;
      q.d = p.d;
      if (deserialize_precond(bytes.data(), (i64)bytes.size(), q) != p.d.n_dofs) return fail("round trip DOFs");
      if (serialize_precond(q) != bytes) return fail("round trip bytes");
      q.pre_given = true;
      q.explicit_op = p.explicit_op;
      finish_from_bytes(q);
      if (make_device_layout(q).m != lay.m) return fail("round trip layout");
    }
    for (int k = 0; k < 150; ++k) {
      std::vector<uint8_t> b = bytes;
      const int mode = k % 5;
      if (mode == 0) b.resize(rng() % b.size());
      else if (mode == 1) b.push_back((uint8_t)rng());
      else
        for (int f = 0; f < 1 + mode; ++f) b[rng() % b.size()] ^= (uint8_t)(1u << (rng() % 8));
      Plan q;
      q.d = p.d;
      try {
        deserialize_precond(b.data(), (i64)b.size(), q);
        q.explicit_op = p.explicit_op;
        finish_from_bytes(q);
        ++loaded;
      } catch (const Error&) {
        ++rejected;
      }
      ++checked;
    }
  }
  std::printf("plan_driver: %d corrupted copies, %d rejected, %d loaded\n", checked, rejected, loaded);
  return 0;
}
