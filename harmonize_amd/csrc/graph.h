// Device kernels shared by the BFS batch and the serving loop: R-MAT edge
// generation, CSR row offsets from sorted sources, and the level/iterator
// reset (a kernel rather than a memset, so that it can be captured in a
// hipGraph). Internal linkage: each translation unit gets its own copy.
#pragma once

#include "rmat.h"
#include "bfs.h"

namespace hz {
namespace graph {
static __global__ void __launch_bounds__(256)
    rmat_edges_kernel(uint32_t g, uint32_t edges, uint32_t vertices, uint32_t scale, uint32_t k0, uint32_t k1,
                      uint32_t* src, uint32_t* dst) {
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < edges; e += stride) {
        uint32_t s, d;
        hz::rmat::edge(g, e, vertices, scale, k0, k1, s, d);
        src[e] = s;
        dst[e] = d;
    }
}

// row_ptr[v] = first position of source v in the sorted source list.
static __global__ void __launch_bounds__(256)
    row_ptr_kernel(const uint32_t* sorted_src, uint32_t edges, uint32_t vertices, int32_t* row_ptr) {
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v <= vertices; v += stride) {
        uint32_t lo = 0, hi = edges;
        while (lo < hi) {
            uint32_t mid = lo + ((hi - lo) >> 1);
            if (sorted_src[mid] < v) {
                lo = mid + 1u;
            } else {
                hi = mid;
            }
        }
        row_ptr[v] = (int32_t)lo;
    }
}

// Levels to BFS_UNREACHED and the job iterator to [0, jobs): a kernel, so
// that it can be captured in a graph.
static __global__ void __launch_bounds__(256)
    bfs_reset_kernel(int32_t* level, uint64_t n, util::iter::AtomicIter<unsigned int>* iter, unsigned int jobs) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) level[i] = BFS_UNREACHED;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        iter->value = 0u;
        iter->limit = jobs;
    }
}

inline unsigned int grid_for(uint64_t n, unsigned int block = 256, unsigned int cap = 8192) {
    uint64_t g = (n + block - 1) / block;
    if (g > cap) g = cap;
    return g == 0 ? 1u : (unsigned int)g;
}

}  // namespace graph
}  // namespace hz
