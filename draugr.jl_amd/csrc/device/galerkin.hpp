// Galerkin coarse operator on the GPU at setup (pattern, triple map, values
// and the transpose map of P), identical to the host's coarse_sparsity and
// build_transpose_map; src/galerkin.jl:11-183, src/prolongation.jl:767-797.
#pragma once

#include <memory>

#include "../host/types.hpp"
#include "dbuf.hpp"

namespace draugr {

// Galerkin triple map of one level resident in HBM (src/types.jl:413-418):
// offsets per coarse nonzero, then the indices into P, A, P of every triple.
struct DeviceRMap {
    int64_t ntriples = 0;
    DBuf<int64_t> off;
    DBuf<int32_t> pi, ai, pj;
};

// The device map of a host restriction map: the one built on the GPU, or a
// new upload of the host vectors.
std::shared_ptr<DeviceRMap> device_rmap(const RestrictionMap& R, hipStream_t st);

// Host CSR in, host CSR and transpose map out; the product runs in HBM.  The
// triple map (when rmap is not null) is left in HBM: rmap->device.
void galerkin_device(const CSR& A, const CSR& P, int32_t nc, CSR& Ac, RestrictionMap* rmap, TransposeMap& T);

}  // namespace draugr
