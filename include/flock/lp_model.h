// flock/lp_model.h -- the reference's LP model class (src/gurobi_lp_method/lp_model.h),
// declared so that callers which name it compile.  The LP path (a commercial solver)
// is outside this library: with MAKE_WITH_GUROBI == 0 the reference's main() never
// constructs it (main.cpp:55-74), and here execute() throws std::runtime_error
// instead of silently returning a zero objective (lp_model.cpp:188-242).
#pragma once
#include "structs.h"

class lpModel {
 public:
  ProblemInstance* data;
  flowInformation* optimal_flows;

  double objVal;
  int totalIterations;
  double totalSolveTime;

  lpModel(ProblemInstance* problem, flowInformation* flows);
  ~lpModel();
  void execute();
  void create_flows();
};
