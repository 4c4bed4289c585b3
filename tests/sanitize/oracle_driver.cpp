// CPU driver of the oracle (oracle/oracle.cpp) for AddressSanitizer and
// UndefinedBehaviorSanitizer builds: small designs through every solver
// configuration (CG / GMRES / LSMR, additive / multiplicative / none,
// implicit / explicit), checking convergence.  Run by tests/test_sanitize.py.
#include <cstdio>
#include <random>
#include <vector>

#include "../../oracle/oracle.hpp"

using namespace oracle;

int main() {
  std::mt19937_64 rng(3);
  int solves = 0;
  for (int t = 0; t < 4; ++t) {
    const int Q = 2 + t % 3;
    const i64 n = 300 + 200 * t;
    std::vector<uint32_t> codes((size_t)n * Q);
    for (int q = 0; q < Q; ++q)
      for (i64 i = 0; i < n; ++i) codes[(size_t)q * n + i] = (uint32_t)(rng() % (uint32_t)(q == 0 ? n / 6 : 5 + 3 * q));
    std::vector<double> w(n), y(n);
    for (i64 i = 0; i < n; ++i) {
      w[i] = 0.5 + (double)(rng() % 100) / 100.0;
      y[i] = (double)(rng() % 2001) / 1000.0 - 1.0;
    }
    for (int cfg = 0; cfg < 5; ++cfg) {
      Params p;
      p.krylov = cfg == 2 ? 1 : (cfg == 4 ? 2 : 0);
      p.precond = cfg == 2 ? 2 : (cfg == 3 ? 0 : 1);
      p.explicit_op = cfg == 1;
      p.tol = 1e-9;
      Solver s(make_design(codes.data(), n, Q, n, t % 2 ? w.data() : nullptr), p);
      SolveOut o = s.solve(y.data());
      if (!o.converged) {
        std::fprintf(stderr, "oracle_driver: design %d config %d did not converge\n", t, cfg);
        return 1;
      }
      ++solves;
    }
  }
  std::printf("oracle_driver: %d solves converged\n", solves);
  return 0;
}
