// flock/uv_model_parallel.h -- drop-in for the reference's solver classes
// (src/parallel_simplex_method/parallel_uv_method/uv_model_parallel.h:19-88,
//  parallel_ss_method/ss_model_parallel.h:17-39, switch/switch_model_parallel.h:10-32).
// Same constructor, execute(), create_flows() and public statistics; the work runs on
// the MI355X through the C ABI in flock_amd.h, in exact integer arithmetic.
//
// Differences a caller can observe (documented, all on the side of exactness):
//  * float costs are converted to int32 by the smallest decimal scale 10^k (k <= 6)
//    that reproduces every float exactly; otherwise execute() throws
//    std::runtime_error (FLOCK_E_NONINTEGRAL) instead of rounding;
//  * zero-flow basic arcs are dropped from the flow list exactly as the reference's
//    retrieval does (round + remove_if, parallel_kernels.cu:89,174-176), and flows are
//    integral;
//  * errors throw std::runtime_error instead of exit() (parallel_structs.h:96-104).
#pragma once
#include "../flock_amd.h"
#include "structs.h"

class flockModelBase {
 public:
  ProblemInstance* data;
  flowInformation* optimal_flows;
  flowInformation* feasible_flows;

  void execute();
  void create_flows();

  // Model Statistics >> (times in microseconds, totalSolveTime in milliseconds).
  // The phase times are device time measured inside the solving workgroup:
  //   uv_time             tree build + initial duals (later duals are updated by each pivot)
  //   reduced_cost_time   full sweeps + candidate re-pricing
  //   pivot_time          cycle_discovery_time + adjustment_time
  //   cycle_discovery_time  cycle search + leaving arc (+ arc claims of multi-pivot rounds)
  //   resolve_time        0 (conflicts are resolved inside cycle discovery)
  //   adjustment_time     flow update + subtree re-rooting
  double uv_time, reduced_cost_time, pivot_time, cycle_discovery_time, resolve_time, adjustment_time;
  double objVal;
  int totalIterations;
  double totalSolveTime;

  // exact results of the last execute()
  int64_t objective_scaled;   // objective in units of 10^-cost_scale
  int cost_scale;
  int status;

  flock_options options;      // pricing rule, initial basis, iteration budget

  flockModelBase(ProblemInstance* problem, flowInformation* flows, int pricing);
  ~flockModelBase();
  // The flow buffer is owned: moving hands it over (the reference's callers
  // copy-initialise, `uvModel_parallel model = uvModel_parallel(&problem, flows);`,
  // main.cpp:78, which needs an accessible move constructor before C++17); copying is
  // not allowed.
  flockModelBase(flockModelBase&& o) noexcept;
  flockModelBase& operator=(flockModelBase&& o) noexcept;
  flockModelBase(const flockModelBase&) = delete;
  flockModelBase& operator=(const flockModelBase&) = delete;
};

/* Transportation simplex, UV duals + Dantzig entering rule (the reference's default). */
class uvModel_parallel : public flockModelBase {
 public:
  uvModel_parallel(ProblemInstance* problem, flowInformation* flows)
      : flockModelBase(problem, flows, FLOCK_PRICE_DANTZIG) {}
};

/* Stepping-stone multi-pivot: parallel cycle discovery + conflict resolution. */
class ssModel_parallel : public flockModelBase {
 public:
  ssModel_parallel(ProblemInstance* problem, flowInformation* flows)
      : flockModelBase(problem, flows, FLOCK_PRICE_MULTI) {}
};

/* Hybrid: multi-pivot rounds until one accepts fewer than 10 pivots, then one Dantzig
 * pivot per sweep to optimality (switch_model_parallel.cu:166-191, 292-307). */
class switchModel_parallel : public flockModelBase {
 public:
  switchModel_parallel(ProblemInstance* problem, flowInformation* flows)
      : flockModelBase(problem, flows, FLOCK_PRICE_SWITCH) {}
};
