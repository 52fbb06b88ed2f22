// Exact packing certificate (host code).
#pragma once
#include <string>
#include <vector>

namespace sis {

// 0 and L (50 decimals) when no two squares overlap in their interiors;
// 2 and the reason otherwise.
int certify_packing(const std::vector<std::string>& xs, const std::vector<std::string>& ys,
                    const std::vector<std::string>& as, std::string& L50, std::string& reason);

// Shortest decimal that reads back to the same double (the float transport
// a CSV reader applies; cuda/certifier.cpp:914-929 mirrors it).
std::string shortest_repr(double v);

// Dilation ladder of cuda/demo.py:25-40 (centres x 1.000002, 1.00001, 1.0001,
// %.17g), each candidate certified literally and after the float round trip.
int export_certify(int n, const double* q, double* dilation, std::string& L50, std::string& reason);

}  // namespace sis
