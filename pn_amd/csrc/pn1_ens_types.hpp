// pn1_ens_types.hpp -- work-item tables and the work queues of the
// ensemble kernels (shared by the force/energy kernels in
// pn1_ens_kernels.hpp and the adjoint kernels in pn1_adj.hip).
#pragma once
#include "pn1_common.hpp"

namespace pn1dev {

// A pair work item: one 256-star i-block of a cluster against one j chunk,
// self-contained so that a workgroup starts it with one scalar load.
//   off    the cluster's BH index; its stars are off + 1 .. off + nst
//   ib     i-block (local stars ib * 256 ..)
//   j0, j1 the chunk's particle range
//   c      the cluster's index
//   pslot  first partial slot of the chunk: poff[c] + s * nst
struct __align__(32) EItem {
    int off, nst, ib, j0, j1, c;
    long long pslot;
};

// Star block of the O(n) kernels: cluster and first local star index.
// Partial sums of cluster c start at slot poff[c] (a slot is one chunk of
// one star; slots of c: nchunks(c) x its star count) times the number of
// values a kernel keeps per slot (NACC, NACC + 3 or NACCE).
struct __align__(8) EBlk {
    int c, first;
};

// Ensemble launch constants: chunk length L (j-records per item).
struct EConst {
    int nitems, L, nclusters;
};

__device__ __forceinline__ int ens_nchunks(int nst, int L) { return (nst + L - 1) / L; }

// Work queues, one per XCD.  The dispatcher deals workgroup ids to the 8
// XCDs round-robin, so workgroup b runs on XCD b % 8.  Whole clusters are
// assigned to queues (balanced by cost on the host), so the j-records and
// lane state of a cluster are fetched into one XCD's L2.  Queue q holds
// items [qoff[q], qoff[q + 1]) with counter[q] its next index; a
// workgroup drains its own XCD's queue, then takes items from the others
// in turn, and leaves when all 8 are empty.  The four waves of a
// workgroup work on one item, so they read the same j-record stream
// (through the scalar cache) in step.
constexpr int NXCD = 8;

// The next work item for this workgroup, identical in every lane (and
// known to the compiler as uniform); -1 when every queue is empty.  `k`
// (thread 0's count of exhausted queues) starts at 0.
__device__ __forceinline__ int ens_take(unsigned int* counter, const int* __restrict__ qoff,
                                        int& k)
{
    __shared__ int s_item;
    __syncthreads();                         // every lane has read the last item
    if (threadIdx.x == 0) {
        int it = -1;
        while (k < NXCD) {
            const int q = ((int)blockIdx.x + k) & (NXCD - 1);
            const int idx = qoff[q] + (int)atomicAdd(counter + q, 1u);
            if (idx < qoff[q + 1]) {
                it = idx;
                break;
            }
            ++k;
        }
        s_item = it;
    }
    __syncthreads();
    return __builtin_amdgcn_readfirstlane(s_item);
}

}  // namespace pn1dev
