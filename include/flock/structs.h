// flock/structs.h -- the reference's instance/structs API (src/structs.h:21-52),
// kept with the same names, fields and meaning so callers of the reference compile
// unchanged against this library.
#pragma once
#include <iostream>
#include <string>

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
};

/* Information for a flow variable (structs.h:45-52). */
struct flowInformation {
  int source, destination;
  float qty;
};

std::ostream& operator<<(std::ostream& o, const flowInformation& x);

/* `.dat` reader (instance.cpp:80-179): "m n", m supplies, n demands, m*n costs.
 * Fills problem.numSupplies/numDemands, allocates and reads the arrays, sets
 * active_flows = m+n-1.  Throws std::runtime_error if the file cannot be read. */
void readProblemFile(const std::string& filename, ProblemInstance& problem);
