// Drop-in check: a caller of the reference's C++ interface, used the way the
// reference's own driver uses it (src/main.cpp:16-51, 76-166): InputParser and
// make_problem read the `.dat` file named by -i and the algorithm named by -a, the
// caller owns flows[m+n-1], the model object is copy-initialised from a temporary
// (under -std=c++14 this needs a move constructor), then execute(), create_flows()
// and the public statistics.  Built with g++ -std=c++14 against include/flock and
// libflock_amd.so.  Prints one RESULT line; exit status 0 only if the flow list it got
// back balances every supply and demand and its cost equals objVal.
#include "logger.h"
#include "instance.h"
#include "structs.h"
#include "algos.h"

#include <cstdio>
#include <cstdlib>

template <class Model>
static int run(ProblemInstance& problem, flowInformation* flows, const char* name) {
  Model model = Model(&problem, flows);
  model.execute();
  model.create_flows();
  const int m = problem.numSupplies, n = problem.numDemands;
  std::vector<double> rs(m, 0.0), cs(n, 0.0);
  double cost = 0.0;
  bool ok = problem.active_flows >= 1 && problem.active_flows <= m + n - 1;
  for (int k = 0; k < problem.active_flows && ok; ++k) {
    const flowInformation& f = flows[k];
    ok = f.source >= 0 && f.source < m && f.destination >= 0 && f.destination < n && f.qty > 0;
    if (!ok) break;
    rs[f.source] += f.qty;
    cs[f.destination] += f.qty;
    cost += (double)f.qty * (double)problem.costs[(size_t)f.source * n + f.destination];
  }
  for (int i = 0; i < m && ok; ++i) ok = rs[i] == problem.supplies[i];
  for (int j = 0; j < n && ok; ++j) ok = cs[j] == problem.demands[j];
  ok = ok && std::fabs(cost - model.objVal) <= 1e-6 * std::fabs(model.objVal) + 1e-6;
  std::printf("RESULT %s objVal %.10g scaled %lld scale %d iterations %d active_flows %d flows_ok %d "
              "uv %.1f r %.1f pivot %.1f cycle %.1f adjust %.1f\n",
              name, model.objVal, (long long)model.objective_scaled, model.cost_scale, model.totalIterations,
              problem.active_flows, ok ? 1 : 0, model.uv_time, model.reduced_cost_time, model.pivot_time,
              model.cycle_discovery_time, model.adjustment_time);
  return ok ? 0 : 3;
}

int main(int argc, char** argv) {
  InputParser input(argc, argv);
  ProblemInstance problem = ProblemInstance();
  make_problem(input, problem);
  flowInformation* flows =
      (flowInformation*)malloc((problem.numSupplies + problem.numDemands - 1) * sizeof(flowInformation));
  int rc = 2;
  try {
    if (problem.algo == ProblemInstance::my_algo::parallel_uv)
      rc = run<uvModel_parallel>(problem, flows, "parallel_uv");
    else if (problem.algo == ProblemInstance::my_algo::parallel_ss)
      rc = run<ssModel_parallel>(problem, flows, "parallel_ss");
    else if (problem.algo == ProblemInstance::my_algo::switch_hybrid)
      rc = run<switchModel_parallel>(problem, flows, "switch_hybrid");
    else
      add_log_msg("error", "this check runs parallel_uv, parallel_ss or switch_hybrid");
  } catch (const std::exception& e) {
    add_log_msg("error", e.what());
    rc = 1;
  }
  free(flows);
  return rc;
}
