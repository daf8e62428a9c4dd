// Host-only helpers of libdgh_device.so: the reference CLI configuration.
#pragma once
#include "../../include/dgh/dgh_device.h"

namespace dgh {
dgh_config default_config();                                  // config.hpp:7-18
void validate_config(const dgh_config &c);                    // config.cpp:6-22
dgh_config parse_config(int argc, const char *const *argv);   // config.cpp:23-122
}  // namespace dgh
