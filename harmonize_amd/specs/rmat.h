// R-MAT graph generator (Chakrabarti, Zhan, Faloutsos, "R-MAT: A Recursive
// Model for Graph Mining", SDM 2004) on Philox4x32-10, portable between
// gfx950 and the host so both build the same edge list.
//
// Edge e of graph g at scale s (2^s >= vertices) descends s levels of the
// adjacency matrix; at each level a 16-bit uniform picks the quadrant
// a = (0,0), b = (0,1), c = (1,0), d = (1,1) with probabilities
// a, b, c = 0.57, 0.19, 0.19 (d = 0.05), appending one bit to the row
// (source) and one to the column (target). Eight levels use one Philox
// draw: counter (e, g, draw, RMAT_STREAM), key = the batch seed. The
// 2^s-vertex ids are then mapped onto [0, vertices) by id * vertices >> s,
// which keeps the skew (vertex 0 stays the heaviest). Duplicate edges and
// self-loops are kept; the graph is directed.
#pragma once

#include <cstdint>

#include "philox.h"

namespace hz {
namespace rmat {

// Cumulative quadrant thresholds on a 16-bit uniform: round(65536 x 0.57),
// round(65536 x 0.76), round(65536 x 0.95).
constexpr uint32_t T_A = 37356u;
constexpr uint32_t T_AB = 49807u;
constexpr uint32_t T_ABC = 62259u;
constexpr uint32_t RMAT_STREAM = 0x524D4154u;  // "RMAT"
constexpr uint32_t SIZE_STREAM = 0x53495A45u;  // "SIZE"

// Smallest s with 2^s >= n (n >= 1).
HZ_PHILOX_HD uint32_t scale_of(uint32_t n) {
    uint32_t s = 0;
    while (s < 32u && (1ull << s) < (uint64_t)n) ++s;
    return s;
}

HZ_PHILOX_HD uint32_t quadrant(uint32_t u16) {
    return u16 < T_A ? 0u : (u16 < T_AB ? 1u : (u16 < T_ABC ? 2u : 3u));
}

// Source and target (in [0, vertices)) of edge e of graph g.
HZ_PHILOX_HD void edge(uint32_t g, uint32_t e, uint32_t vertices, uint32_t scale, uint32_t key0, uint32_t key1,
                       uint32_t& src, uint32_t& dst) {
    uint32_t u = 0, v = 0;
    for (uint32_t level = 0, draw = 0; level < scale; ++draw) {
        philox::Ctr r = philox::philox4x32_10(e, g, draw, RMAT_STREAM, key0, key1);
        for (uint32_t j = 0; j < 8u && level < scale; ++j, ++level) {
            uint32_t q = quadrant((r.v[j >> 1] >> ((j & 1u) * 16u)) & 0xFFFFu);
            u = (u << 1) | (q >> 1);
            v = (v << 1) | (q & 1u);
        }
    }
    src = (uint32_t)(((uint64_t)u * vertices) >> scale);
    dst = (uint32_t)(((uint64_t)v * vertices) >> scale);
}

// The 32 uniform bits that set the size of graph g in a batch.
HZ_PHILOX_HD uint32_t size_bits(uint32_t g, uint32_t key0, uint32_t key1) {
    return philox::philox4x32_10(g, 0u, 0u, SIZE_STREAM, key0, key1).v[0];
}

}  // namespace rmat
}  // namespace hz
