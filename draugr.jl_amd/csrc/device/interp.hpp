// Strength of connection and extended+i interpolation on the GPU at setup,
// identical to the host's interpolation_cf; src/strength.jl:24-61,
// src/prolongation.jl:467-690.
#pragma once

#include <vector>

#include "../host/types.hpp"

namespace draugr {

// P (host CSR) from a C/F splitting, then (truncate) filter_prolongation(P,
// ip.trunc_factor); the rows are computed in HBM.  Other interpolation types,
// norm_p != 1 and levels of fewer than min_rows rows (where one host thread
// beats the launches: coarse levels have few, long rows) run the host routines.
CSR interpolation_device(const CSR& A, const std::vector<int8_t>& cf, int32_t nc, const InterpSpec& ip, double theta,
                         bool truncate, int32_t min_rows);

}  // namespace draugr
