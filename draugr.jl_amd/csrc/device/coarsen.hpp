// PMIS sweeps of the C/F splitting on the GPU at setup, identical to the
// host's; src/coarsening.jl:202-283 (PMIS), 575-624 (HMIS's second phase).
#pragma once

#include "../host/types.hpp"

namespace draugr {

// Same contract as the host's pmis_sweeps (PmisFn in host/types.hpp).
void pmis_sweeps_device(std::vector<int8_t>& cf, int32_t n, const int64_t* off, const int32_t* adj,
                        const uint8_t* s, const std::vector<double>& measure, bool low_measure_fine);

}  // namespace draugr
