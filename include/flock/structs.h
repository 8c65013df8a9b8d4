// flock/structs.h -- the reference's instance/structs API (src/structs.h:21-52),
// kept with the same names, fields and meaning so callers of the reference compile
// unchanged against this library.  Like the reference's header it pulls in the
// standard headers its callers rely on (src/structs.h:1-10: streams, strings,
// chrono, containers, <cmath> for round()).
#pragma once
#include <algorithm>
#include <chrono>
#include <cmath>
#include <fstream>
#include <iostream>
#include <iterator>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "parameters.h"

/* Container for the problem construct (structs.h:21-40). */
class ProblemInstance {
 public:
  bool read_mode;
  std::string filename;
  enum my_algo { cpu_lp_solve, parallel_uv, parallel_ss, switch_hybrid };
  my_algo algo;
  int numDemands, numSupplies, *demands, *supplies, active_flows;

  // Statistics >>
  double readTime, preprocessTime, solveTime, postprocessTime;
  double totalFlowCost;

  float* costs;  // row-major i*numDemands + j

  void allocate_memory();
  ProblemInstance();
  ~ProblemInstance();
  // The arrays are owned: moving hands them over, copying is not allowed (a copy
  // would free them twice).  `ProblemInstance p = ProblemInstance();` (main.cpp:24)
  // uses the move constructor.
  ProblemInstance(ProblemInstance&& o) noexcept;
  ProblemInstance& operator=(ProblemInstance&& o) noexcept;
  ProblemInstance(const ProblemInstance&) = delete;
  ProblemInstance& operator=(const ProblemInstance&) = delete;
};

/* Information for a flow variable (structs.h:45-52). */
struct flowInformation {
  int source, destination;
  float qty;
};

std::ostream& operator<<(std::ostream& o, const flowInformation& x);

/* `.dat` reader (instance.cpp:80-179): "m n", m supplies, n demands, m*n costs.
 * Fills problem.numSupplies/numDemands, allocates and reads the arrays, sets
 * active_flows = m+n-1.  Throws std::runtime_error if the file cannot be read.
 * The reference's own entry points (make_problem, readSize, readCosts) are in
 * instance.h. */
void readProblemFile(const std::string& filename, ProblemInstance& problem);
