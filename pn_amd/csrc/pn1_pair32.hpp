// pn1_pair32.hpp -- launches of the mixed-form pair kernels (k_pair32,
// k_pair32q, k_ens_pair32, k_ens_pair32x), which live in a translation unit
// of their own (pn1_pair32.hip): it is compiled with LLVM's max-ILP machine
// scheduler, the rest of the library with the default one.
#pragma once
#include <hip/hip_runtime.h>

#include "pn1_common.hpp"
#include "pn1_ctx.hpp"
#include "pn1_ens.hpp"

namespace pn1host {

// Kernel form index: (loop variant & 3) * 4 + 2 * (cancellation sums) +
// (equal masses).
inline int pair32_form(int variant, bool ts, bool eqm)
{
    return (variant & 3) * 4 + (ts ? 2 : 0) + (eqm ? 1 : 0);
}

// Single cluster: i blocks ib0 .. ib0 + nib - 1 against every star, the j
// range split in S chunks (pn1_pair32.hip).
void launch_pair32(pn1_ctx* c, const double* in, const pn1dev::Phys& P, int ib0, int nib, int S,
                   int form);
// Ensemble: one persistent launch of grid g over the work queues.
void launch_ens_pair32(pn1_ens* e, const double* in, const pn1dev::Phys& P, dim3 g, int form);
// Resident workgroups per CU of the ensemble kernel of loop variant var
// (0..3), in its form with the cancellation sums (the one with the most
// registers).
hipError_t ens_pair32_occupancy(int var, int* per_cu);

// pn1_api.hip
bool quad_pairs(const pn1_ctx* c, int S);
int pair_grid(int slots, int nib);

}  // namespace pn1host
