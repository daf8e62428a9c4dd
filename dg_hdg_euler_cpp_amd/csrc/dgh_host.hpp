// Host-only helpers of libdgh_device.so: the reference CLI configuration.
#pragma once
#include <vector>
#include "../../include/dgh/dgh_device.h"

namespace dgh {
dgh_config default_config();                                  // config.hpp:7-18
void validate_config(const dgh_config &c);                    // config.cpp:6-22
dgh_config parse_config(int argc, const char *const *argv);   // config.cpp:23-122
struct Basis;
// Gauss-Legendre rule of n points, ascending (nodal_basis.hpp:13-25)
void gauss_rule(int n, std::vector<double> &x, std::vector<double> &w);
// Lagrange basis of the nodes x evaluated at z
std::vector<double> lagrange_at(const std::vector<double> &x, double z);
// nodal_basis.hpp:34-65 tables for n nodes; the mass entries use hx, hy;
// Q > n adds the Q-point volume rule
Basis make_basis(int n, double hx, double hy, int Q = 0);
}  // namespace dgh
