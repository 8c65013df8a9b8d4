// flock/algos.h -- every solver class a caller of the reference can name
// (src/algos.h:1-4): the LP model and the three simplex models.
#pragma once
#include "lp_model.h"
#include "uv_model_parallel.h"
