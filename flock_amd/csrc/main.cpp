// main.cpp -- `flock` command line driver, the counterpart of the reference's
// src/main.cpp:16-167: `flock -i <file.dat> -a <parallel_uv|parallel_ss|switch_hybrid>`
// reads the instance, runs the chosen solver class and prints the reference's
// statistics lines.  Extra flags: -p <dantzig|candidates|multi|switch> overrides the pricing
// rule, -f prints the flows.
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <string>
#include <vector>

#include "flock/uv_model_parallel.h"

static void usage() {
  std::cerr << "usage: flock -i <file.dat> [-a parallel_uv|parallel_ss|switch_hybrid] "
               "[-p dantzig|candidates|multi|switch] [-f]\n";
}

int main(int argc, char** argv) {
  std::string file, algo = "parallel_uv", pricing;
  bool print_flows = false;
  for (int k = 1; k < argc; ++k) {
    std::string a = argv[k];
    if (a == "-i" && k + 1 < argc) file = argv[++k];
    else if (a == "-a" && k + 1 < argc) algo = argv[++k];
    else if (a == "-p" && k + 1 < argc) pricing = argv[++k];
    else if (a == "-f") print_flows = true;
    else { usage(); return 2; }
  }
  if (file.empty()) { usage(); return 2; }
  try {
    ProblemInstance problem;
    readProblemFile(file, problem);
    std::cout << "[info] => Problem Dimension => Supplies: " << problem.numSupplies
              << "\tDemands: " << problem.numDemands << std::endl;
    std::vector<flowInformation> flows(problem.numSupplies + problem.numDemands - 1);
    int p = FLOCK_PRICE_DANTZIG;
    if (algo == "parallel_ss") p = FLOCK_PRICE_MULTI;
    else if (algo == "switch_hybrid") p = FLOCK_PRICE_SWITCH;
    else if (algo != "parallel_uv") { std::cerr << "unknown algorithm " << algo << "\n"; return 2; }
    if (pricing == "dantzig") p = FLOCK_PRICE_DANTZIG;
    else if (pricing == "candidates") p = FLOCK_PRICE_CANDIDATES;
    else if (pricing == "multi") p = FLOCK_PRICE_MULTI;
    else if (pricing == "switch") p = FLOCK_PRICE_SWITCH;
    flockModelBase model(&problem, flows.data(), p);
    model.execute();
    model.create_flows();
    std::cout << "[info] => >>>> BASIC STATISTICS | Objective: " << model.objVal
              << " | Iterations: " << model.totalIterations << " | Time: " << model.totalSolveTime
              << " | MATRIX : " << problem.numSupplies << " X " << problem.numDemands << std::endl;
    std::cout << "[info] => >>>> EXACT | Objective: " << model.objective_scaled << " x 10^-" << model.cost_scale
              << " | Active flows: " << problem.active_flows << std::endl;
    if (print_flows)
      for (int k = 0; k < problem.active_flows; ++k) std::cout << flows[k] << "\n";
  } catch (const std::exception& e) {
    std::cerr << "[error] => " << e.what() << std::endl;
    return 1;
  }
  return 0;
}
