// flock/instance.h -- the reference's command-line parser and `.dat` reader
// (src/instance.h:13-53, src/instance.cpp:3-179), same names and behaviour:
//   -i <file>  input file (required; FileNotFoundException if absent or unreadable)
//   -a <algo>  cpu_lp_solve (default) | parallel_uv | parallel_ss | switch_hybrid
// make_problem() reads the size, allocates, reads supplies, demands and the row-major
// costs, sets the algorithm and active_flows = m+n-1, and adds the elapsed time in
// milliseconds to problem.readTime.
#pragma once
#include <exception>
#include <string>
#include <vector>

#include "logger.h"
#include "structs.h"

struct FileNotFoundException : public std::exception {
  const char* what() const throw() { return "File not found!"; }
};

class InputParser {
 public:
  InputParser(int& argc, char** argv) {
    for (int i = 1; i < argc; ++i) args_.emplace_back(argv[i]);
  }
  // The argument after `option`, or an empty string.
  const std::string& getCmdOption(const std::string& option) const {
    static const std::string none;
    for (size_t k = 0; k + 1 < args_.size(); ++k)
      if (args_[k] == option) return args_[k + 1];
    return none;
  }
  bool cmdOptionExists(const std::string& option) const {
    for (const std::string& a : args_)
      if (a == option) return true;
    return false;
  }

 private:
  std::vector<std::string> args_;
};

void get_input_mode_and_filename(InputParser input, ProblemInstance& problem);
void get_algorithm(InputParser input, ProblemInstance& problem);
void make_problem(InputParser input, ProblemInstance& problem);
void readSize(ProblemInstance& problem);
void readCosts(ProblemInstance& problem);
