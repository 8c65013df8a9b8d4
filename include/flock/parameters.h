// flock/parameters.h -- build switches the reference's callers test
// (src/parameters.h:7-22).  This library has no LP backend, so MAKE_WITH_GUROBI is 0
// unless a caller defines it; lpModel (lp_model.h) then refuses to run.
#pragma once

#ifndef LOG_TO_FILE
#define LOG_TO_FILE 0
#endif

#ifndef LOG_LEVEL
#define LOG_LEVEL "debug"
#endif

#ifndef MAKE_WITH_GUROBI
#define MAKE_WITH_GUROBI 0
#endif
