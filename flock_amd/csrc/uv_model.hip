// uv_model.hip -- the reference's C++ instance/solver API over the C ABI.
//   ProblemInstance / flowInformation      <- src/structs.{h,cpp}
//   readProblemFile                        <- src/instance.cpp:80-179 (readSize/readCosts)
//   uvModel_parallel / ssModel_parallel /
//   switchModel_parallel                   <- uv_model_parallel.cu:20-65, 227-400 and the
//                                             ss / switch classes with the same shape
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <stdexcept>
#include <vector>

#include "../../include/flock/uv_model_parallel.h"

ProblemInstance::ProblemInstance()
    : read_mode(false), algo(parallel_uv), numDemands(0), numSupplies(0), demands(nullptr), supplies(nullptr),
      active_flows(0), readTime(0), preprocessTime(0), solveTime(0), postprocessTime(0), totalFlowCost(0),
      costs(nullptr) {}

void ProblemInstance::allocate_memory() {
  supplies = (int*)malloc(sizeof(int) * numSupplies);
  demands = (int*)malloc(sizeof(int) * numDemands);
  costs = (float*)malloc(sizeof(float) * (size_t)numSupplies * numDemands);
}

ProblemInstance::~ProblemInstance() {
  free(supplies);
  free(demands);
  free(costs);
}

std::ostream& operator<<(std::ostream& o, const flowInformation& x) {
  o << "Flow - " << x.qty << " units from source/rowNo: " << x.source << " to destination/colNo: " << x.destination;
  return o;
}

void readProblemFile(const std::string& filename, ProblemInstance& problem) {
  auto t0 = std::chrono::steady_clock::now();
  std::ifstream in(filename);
  if (!in) throw std::runtime_error("input file not found: " + filename);
  problem.filename = filename;
  in >> problem.numSupplies >> problem.numDemands;
  if (!in || problem.numSupplies < 1 || problem.numDemands < 1) throw std::runtime_error("bad header in " + filename);
  problem.allocate_memory();
  for (int i = 0; i < problem.numSupplies; ++i) in >> problem.supplies[i];
  for (int j = 0; j < problem.numDemands; ++j) in >> problem.demands[j];
  const size_t cells = (size_t)problem.numSupplies * problem.numDemands;
  for (size_t k = 0; k < cells; ++k) in >> problem.costs[k];
  if (!in) throw std::runtime_error("truncated file " + filename);
  problem.active_flows = problem.numSupplies + problem.numDemands - 1;
  problem.readTime += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

flockModelBase::flockModelBase(ProblemInstance* problem, flowInformation* flows, int pricing)
    : data(problem), optimal_flows(flows), feasible_flows(nullptr), uv_time(0), reduced_cost_time(0),
      pivot_time(0), cycle_discovery_time(0), resolve_time(0), adjustment_time(0), objVal(0),
      totalIterations(0), totalSolveTime(0), objective_scaled(0), cost_scale(0), status(0) {
  flock_default_options(&options);
  options.pricing = pricing;
  feasible_flows = (flowInformation*)malloc(sizeof(flowInformation) * (size_t)(data->numSupplies + data->numDemands - 1));
}

flockModelBase::~flockModelBase() { free(feasible_flows); }

void flockModelBase::execute() {
  const int m = data->numSupplies, n = data->numDemands, V = m + n;
  std::vector<int32_t> icost((size_t)m * n);
  int32_t scale = 0;
  int st = flock_scale_float_costs((int64_t)m * n, data->costs, icost.data(), &scale);
  if (st) throw std::runtime_error(std::string("flock: ") + flock_last_error());
  std::vector<flock_flow> flows(V - 1);
  flock_instance inst{m, n, icost.data(), data->supplies, data->demands};
  flock_result res;
  std::memset(&res, 0, sizeof res);
  res.flows = flows.data();
  auto t0 = std::chrono::steady_clock::now();
  st = flock_solve(&inst, &options, &res);
  const double wall_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  status = res.status;
  if (st != FLOCK_OK) throw std::runtime_error(std::string("flock: ") + flock_last_error());
  cost_scale = scale;
  objective_scaled = res.objective;
  objVal = (double)res.objective / std::pow(10.0, scale);
  // the reference counts loop passes, including the final optimality check
  totalIterations = (int)res.iterations + 1;
  totalSolveTime = wall_ms;
  for (int k = 0; k < res.n_flows; ++k)
    feasible_flows[k] = flowInformation{flows[k].source, flows[k].destination, (float)flows[k].qty};
  data->active_flows = res.n_flows;
  data->totalFlowCost = objVal;
  data->solveTime = totalSolveTime;
}

void flockModelBase::create_flows() {
  std::memcpy(optimal_flows, feasible_flows, sizeof(flowInformation) * (size_t)data->active_flows);
}
