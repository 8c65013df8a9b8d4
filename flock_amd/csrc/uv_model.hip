// uv_model.hip -- the reference's C++ instance/solver API over the C ABI.
//   ProblemInstance / flowInformation      <- src/structs.{h,cpp}
//   add_log_msg                            <- src/logger.cpp:3-6
//   InputParser helpers, make_problem,
//   readSize / readCosts                   <- src/instance.cpp:3-179
//   uvModel_parallel / ssModel_parallel /
//   switchModel_parallel                   <- uv_model_parallel.cu:20-65, 227-400 and the
//                                             ss / switch classes with the same shape
//   lpModel                                <- gurobi_lp_method/lp_model.h (declared only;
//                                             no LP backend here)
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <stdexcept>
#include <vector>

#include "../../include/flock/algos.h"
#include "../../include/flock/instance.h"

ProblemInstance::ProblemInstance()
    : read_mode(false), algo(parallel_uv), numDemands(0), numSupplies(0), demands(nullptr), supplies(nullptr),
      active_flows(0), readTime(0), preprocessTime(0), solveTime(0), postprocessTime(0), totalFlowCost(0),
      costs(nullptr) {}

ProblemInstance::ProblemInstance(ProblemInstance&& o) noexcept
    : read_mode(o.read_mode), filename(std::move(o.filename)), algo(o.algo), numDemands(o.numDemands),
      numSupplies(o.numSupplies), demands(o.demands), supplies(o.supplies), active_flows(o.active_flows),
      readTime(o.readTime), preprocessTime(o.preprocessTime), solveTime(o.solveTime),
      postprocessTime(o.postprocessTime), totalFlowCost(o.totalFlowCost), costs(o.costs) {
  o.demands = nullptr; o.supplies = nullptr; o.costs = nullptr;
}

ProblemInstance& ProblemInstance::operator=(ProblemInstance&& o) noexcept {
  if (this == &o) return *this;
  free(supplies); free(demands); free(costs);
  read_mode = o.read_mode; filename = std::move(o.filename); algo = o.algo;
  numDemands = o.numDemands; numSupplies = o.numSupplies; active_flows = o.active_flows;
  readTime = o.readTime; preprocessTime = o.preprocessTime; solveTime = o.solveTime;
  postprocessTime = o.postprocessTime; totalFlowCost = o.totalFlowCost;
  demands = o.demands; supplies = o.supplies; costs = o.costs;
  o.demands = nullptr; o.supplies = nullptr; o.costs = nullptr;
  return *this;
}

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

void add_log_msg(std::string level, std::string log_msg) {
  std::cout << "[" << level << "] => " << log_msg << std::endl;
}

// ---------------------------------------------------------------------------
// `.dat` input (instance.cpp:80-179): "m n", m supplies, n demands, m*n costs.
// ---------------------------------------------------------------------------
void readSize(ProblemInstance& problem) {
  std::ifstream in(problem.filename);
  if (!in) {
    std::cerr << "Error: input file not found: " << problem.filename << std::endl;
    throw FileNotFoundException();
  }
  in >> problem.numSupplies >> problem.numDemands;
}

void readCosts(ProblemInstance& problem) {
  std::ifstream in(problem.filename);
  if (!in) {
    std::cerr << "Error: input file not found: " << problem.filename << std::endl;
    throw FileNotFoundException();
  }
  int m = 0, n = 0;
  in >> m >> n;
  if (m != problem.numSupplies || n != problem.numDemands)
    throw std::runtime_error("size of " + problem.filename + " changed while reading");
  for (int i = 0; i < m; ++i) in >> problem.supplies[i];
  for (int j = 0; j < n; ++j) in >> problem.demands[j];
  const size_t cells = (size_t)m * n;
  for (size_t k = 0; k < cells; ++k) in >> problem.costs[k];
  if (!in) throw std::runtime_error("truncated or malformed file " + problem.filename);
}

void readProblemFile(const std::string& filename, ProblemInstance& problem) {
  auto t0 = std::chrono::steady_clock::now();
  problem.filename = filename;
  problem.read_mode = true;
  readSize(problem);
  if (problem.numSupplies < 1 || problem.numDemands < 1) throw std::runtime_error("bad header in " + filename);
  problem.allocate_memory();
  readCosts(problem);
  problem.active_flows = problem.numSupplies + problem.numDemands - 1;
  problem.readTime += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

void get_input_mode_and_filename(InputParser input, ProblemInstance& problem) {
  if (!input.cmdOptionExists("-i")) {
    problem.read_mode = false;
    add_log_msg("error", "Provide input file with argument -i");
    throw FileNotFoundException();
  }
  problem.read_mode = true;
  problem.filename = input.getCmdOption("-i");
  if (problem.filename.empty()) {
    add_log_msg("error", "No filename provided for argument -i");
    throw FileNotFoundException();
  }
  add_log_msg("info", "File Name = " + problem.filename);
}

void get_algorithm(InputParser input, ProblemInstance& problem) {
  std::string a = input.cmdOptionExists("-a") ? input.getCmdOption("-a") : std::string();
  if (a.empty()) a = "cpu_lp_solve";
  if (a == "cpu_lp_solve") problem.algo = ProblemInstance::cpu_lp_solve;
  else if (a == "parallel_uv") problem.algo = ProblemInstance::parallel_uv;
  else if (a == "parallel_ss") problem.algo = ProblemInstance::parallel_ss;
  else if (a == "switch_hybrid") problem.algo = ProblemInstance::switch_hybrid;
  else add_log_msg("error", "Invalid Algorithm!");
}

void make_problem(InputParser input, ProblemInstance& problem) {
  auto t0 = std::chrono::steady_clock::now();
  get_input_mode_and_filename(input, problem);
  if (!std::ifstream(problem.filename).good()) {
    add_log_msg("error", "File :" + problem.filename + " doesn't exist!");
    throw FileNotFoundException();
  }
  readSize(problem);
  if (problem.numSupplies < 1 || problem.numDemands < 1)
    throw std::runtime_error("bad header in " + problem.filename);
  problem.allocate_memory();
  readCosts(problem);
  get_algorithm(input, problem);
  problem.readTime += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  problem.active_flows = problem.numSupplies + problem.numDemands - 1;
}

// ---------------------------------------------------------------------------
// LP model: declared for callers that name it; there is no LP backend here.
// ---------------------------------------------------------------------------
lpModel::lpModel(ProblemInstance* problem, flowInformation* flows)
    : data(problem), optimal_flows(flows), objVal(0), totalIterations(0), totalSolveTime(0) {}
lpModel::~lpModel() {}
void lpModel::execute() {
  throw std::runtime_error("flock: the LP model (cpu_lp_solve) needs a commercial LP solver that this library "
                           "does not include; use parallel_uv, parallel_ss or switch_hybrid");
}
void lpModel::create_flows() {}

// ---------------------------------------------------------------------------
// Simplex models
// ---------------------------------------------------------------------------
flockModelBase::flockModelBase(ProblemInstance* problem, flowInformation* flows, int pricing)
    : data(problem), optimal_flows(flows), feasible_flows(nullptr), uv_time(0), reduced_cost_time(0),
      pivot_time(0), cycle_discovery_time(0), resolve_time(0), adjustment_time(0), objVal(0),
      totalIterations(0), totalSolveTime(0), objective_scaled(0), cost_scale(0), status(0) {
  flock_default_options(&options);
  options.pricing = pricing;
  feasible_flows = (flowInformation*)malloc(sizeof(flowInformation) * (size_t)(data->numSupplies + data->numDemands - 1));
}

flockModelBase::flockModelBase(flockModelBase&& o) noexcept
    : data(o.data), optimal_flows(o.optimal_flows), feasible_flows(o.feasible_flows), uv_time(o.uv_time),
      reduced_cost_time(o.reduced_cost_time), pivot_time(o.pivot_time), cycle_discovery_time(o.cycle_discovery_time),
      resolve_time(o.resolve_time), adjustment_time(o.adjustment_time), objVal(o.objVal),
      totalIterations(o.totalIterations), totalSolveTime(o.totalSolveTime), objective_scaled(o.objective_scaled),
      cost_scale(o.cost_scale), status(o.status), options(o.options) {
  o.feasible_flows = nullptr;
}

flockModelBase& flockModelBase::operator=(flockModelBase&& o) noexcept {
  if (this == &o) return *this;
  free(feasible_flows);
  data = o.data; optimal_flows = o.optimal_flows; feasible_flows = o.feasible_flows;
  uv_time = o.uv_time; reduced_cost_time = o.reduced_cost_time; pivot_time = o.pivot_time;
  cycle_discovery_time = o.cycle_discovery_time; resolve_time = o.resolve_time; adjustment_time = o.adjustment_time;
  objVal = o.objVal; totalIterations = o.totalIterations; totalSolveTime = o.totalSolveTime;
  objective_scaled = o.objective_scaled; cost_scale = o.cost_scale; status = o.status; options = o.options;
  o.feasible_flows = nullptr;
  return *this;
}

flockModelBase::~flockModelBase() { free(feasible_flows); }

void flockModelBase::execute() {
  const int m = data->numSupplies, n = data->numDemands, V = m + n;
  if (!feasible_flows) throw std::runtime_error("flock: model used after it was moved from");
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
  uv_time = res.phase_us[1];
  reduced_cost_time = res.phase_us[2] + res.phase_us[3];
  cycle_discovery_time = res.phase_us[4];
  resolve_time = 0;
  adjustment_time = res.phase_us[5];
  pivot_time = cycle_discovery_time + adjustment_time;
  for (int k = 0; k < res.n_flows; ++k)
    feasible_flows[k] = flowInformation{flows[k].source, flows[k].destination, (float)flows[k].qty};
  data->active_flows = res.n_flows;
  data->totalFlowCost = objVal;
  data->solveTime = totalSolveTime;
}

void flockModelBase::create_flows() {
  std::memcpy(optimal_flows, feasible_flows, sizeof(flowInformation) * (size_t)data->active_flows);
}
