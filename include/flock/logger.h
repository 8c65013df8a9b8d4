// flock/logger.h -- the reference's stdout logger (src/logger.h, logger.cpp:3-6):
// prints "[level] => message" and a newline.
#pragma once
#include <string>

#include "structs.h"

void add_log_msg(std::string level, std::string log_msg);
