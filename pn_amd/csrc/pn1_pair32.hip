// pn1_pair32.hip -- the mixed-form pair kernels and their launches, in a
// translation unit of their own so that they can be compiled with LLVM's
// max-ILP machine scheduler (-mllvm -amdgpu-sched-strategy=max-ilp, see
// pn_amd/build.py).  At n = 131073 on MI355X it shortened k_pair32 by ~0.9%
// at the same 8 waves per SIMD (k_pair32 holds itself to them); the FP64
// kernels stay in pn1_api.hip under the default scheduler, because max-ILP
// raises k_pair64 from 72 to 84 VGPRs (fewer waves, 1.4% longer).  The
// scheduler only orders instructions: the sums are the same bits.
#include <hip/hip_runtime.h>

#include "pn1_ens_kernels.hpp"
#include "pn1_kernels.hpp"
#include "pn1_pair32.hpp"

using namespace pn1dev;

namespace pn1host {
namespace {

template <int PP, bool TS, bool EQ>
void launch32(pn1_ctx* c, const double* in, const Phys& P, int ib0, int nib, int S)
{
    if (quad_pairs(c, S))
        hipLaunchKernelGGL((k_pair32q<PP, TS, EQ>), dim3(pair_grid(S / QSUB, nib)), dim3(WGQ), 0,
                           c->st, c->j32, c->j64, c->thr, in, (float*)c->part, P, S, nib, ib0,
                           c->tile);
    else
        hipLaunchKernelGGL((k_pair32<PP, TS, EQ>), dim3(pair_grid(S, nib)), dim3(WG), 0, c->st,
                           c->j32, c->j64, c->thr, in, (float*)c->part, P, S, nib, ib0, c->tile);
}

using Launch32 = void (*)(pn1_ctx*, const double*, const Phys&, int, int, int);
const Launch32 k_launch32[16] = {
    launch32<0, false, false>, launch32<0, false, true>, launch32<0, true, false>,
    launch32<0, true, true>,   launch32<1, false, false>, launch32<1, false, true>,
    launch32<1, true, false>,  launch32<1, true, true>,   launch32<2, false, false>,
    launch32<2, false, true>,  launch32<2, true, false>,  launch32<2, true, true>,
    launch32<3, false, false>, launch32<3, false, true>,  launch32<3, true, false>,
    launch32<3, true, true>};

template <int PP, bool TS, bool EQ>
void ens_launch32(pn1_ens* e, const double* in, const Phys& P, dim3 g)
{
    if constexpr ((PP & 1) != 0)
        hipLaunchKernelGGL((k_ens_pair32x<PP, TS, EQ>), g, dim3(WG), 0, e->st, e->j32, e->j64,
                           e->thr, in, (float*)e->part, e->items, e->counter, e->qoff, e->cmmax,
                           P, e->tile);
    else
        hipLaunchKernelGGL((k_ens_pair32<PP, TS, EQ>), g, dim3(WG), 0, e->st, e->j32, e->j64,
                           e->thr, in, (float*)e->part, e->items, e->counter, e->qoff, e->cmmax,
                           P, e->tile);
}

using EnsLaunch32 = void (*)(pn1_ens*, const double*, const Phys&, dim3);
const EnsLaunch32 k_ens_launch32[16] = {
    ens_launch32<0, false, false>, ens_launch32<0, false, true>, ens_launch32<0, true, false>,
    ens_launch32<0, true, true>,   ens_launch32<1, false, false>, ens_launch32<1, false, true>,
    ens_launch32<1, true, false>,  ens_launch32<1, true, true>,   ens_launch32<2, false, false>,
    ens_launch32<2, false, true>,  ens_launch32<2, true, false>,  ens_launch32<2, true, true>,
    ens_launch32<3, false, false>, ens_launch32<3, false, true>,  ens_launch32<3, true, false>,
    ens_launch32<3, true, true>};

}  // namespace

void launch_pair32(pn1_ctx* c, const double* in, const Phys& P, int ib0, int nib, int S, int form)
{
    k_launch32[form & 15](c, in, P, ib0, nib, S);
}

void launch_ens_pair32(pn1_ens* e, const double* in, const Phys& P, dim3 g, int form)
{
    k_ens_launch32[form & 15](e, in, P, g);
}

hipError_t ens_pair32_occupancy(int var, int* per_cu)
{
    switch (var & 3) {
    case 0: return hipOccupancyMaxActiveBlocksPerMultiprocessor(per_cu, k_ens_pair32<0, true, false>, WG, 0);
    case 1: return hipOccupancyMaxActiveBlocksPerMultiprocessor(per_cu, k_ens_pair32x<1, true, false>, WG, 0);
    case 2: return hipOccupancyMaxActiveBlocksPerMultiprocessor(per_cu, k_ens_pair32<2, true, false>, WG, 0);
    default: return hipOccupancyMaxActiveBlocksPerMultiprocessor(per_cu, k_ens_pair32x<3, true, false>, WG, 0);
    }
}

}  // namespace pn1host
