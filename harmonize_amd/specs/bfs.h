// Batched breadth-first search: many graphs (jobs) in one program. Each job
// is an int32 CSR graph with an int32 level array; the program computes
// every job's BFS levels from its source vertex.
//
// Label-correcting formulation, so that the result does not depend on the
// order in which promises run:
//  - BfsVisit(job, v, l) relaxes the out-edges of v at level l + 1, unless
//    v's level has meanwhile dropped below l (a better visit exists);
//  - relaxing edge v -> w at level n does atomicMin(level[w], n) and, when
//    that lowered the level, queues BfsVisit(job, w, n) one depth level
//    deeper than the visit that relaxed it (BfsEdges pieces stay at their
//    visit's depth), so a program with STACK_SIZE depth levels sees the BFS
//    level as the depth;
//  - an edge range longer than the split threshold is cut into at most
//    BFS_FANOUT pieces (none shorter than the threshold), all but the first
//    queued as BfsEdges(job, begin, end, n), until the rest is short: a
//    hub's edges spread over many lanes within a few rounds of splitting.
// Levels converge to the exact BFS distances (each final level is attained
// by a visit that relaxed along a shortest path); unreached vertices keep
// BFS_UNREACHED. The number of visits depends on the order, the levels do not.
// On the GPU the edge ranges left after splitting are relaxed by the whole
// wavefront together (bfs_relax_wave), so R-MAT's skewed degrees do not leave
// lanes idle; the host executors relax each range in its own lane.
#pragma once

#include <cstdint>

constexpr int32_t BFS_UNREACHED = 0x7FFFFFFF;
constexpr uint32_t BFS_DEFAULT_SPLIT = 64u;
// Valid split thresholds are [1, BFS_MAX_SPLIT]: a threshold of 0 would cut a
// one-edge range forever, and the wavefront's prefix of range lengths
// (64 lanes x split) must fit in 32 bits. The device clamps to this range;
// the entry points reject values outside it.
constexpr uint32_t BFS_MAX_SPLIT = 1u << 24;
constexpr int BFS_FANOUT = 16;
// Edges a lane relaxes per round of the wavefront's cooperative expansion.
#ifndef HZ_BFS_CHAINS
#define HZ_BFS_CHAINS 4
#endif

// One graph of the batch (32 bytes).
struct BfsJob {
    const int32_t* row_ptr;  // [vertices + 1], local edge offsets
    const int32_t* col;      // [edges], local vertex ids
    int32_t* level;          // [vertices]
    uint32_t vertices;
    uint32_t source;
};

struct BfsState {
    const BfsJob* jobs;
    util::iter::AtomicIter<unsigned int>* job_iter;  // jobs still to start
    unsigned int job_count;
    unsigned int split;  // edge ranges longer than this are split
};

struct BfsVisit;
struct BfsEdges;

template <typename PROGRAM>
__device__ inline void bfs_relax(PROGRAM& prog, unsigned int job, const BfsJob& j, int begin, int end, int next);

struct BfsVisit {
    using Type = void (*)(unsigned int job, int vertex, int level);

    template <typename PROGRAM>
    __device__ static void eval(PROGRAM prog, unsigned int job, int vertex, int level) {
        const BfsJob j = prog.device.jobs[job];
        if (hz_load(&j.level[vertex]) < level) return;
        bfs_relax(prog, job, j, j.row_ptr[vertex], j.row_ptr[vertex + 1], level + 1);
    }
};

struct BfsEdges {
    using Type = void (*)(unsigned int job, int begin, int end, int level);

    template <typename PROGRAM>
    __device__ static void eval(PROGRAM prog, unsigned int job, int begin, int end, int level) {
        const BfsJob j = prog.device.jobs[job];
        bfs_relax(prog, job, j, begin, end, level);
    }
};


#if defined(__HIP_DEVICE_COMPILE__)
// One lane's edge range, as the wavefront shares it in LDS (32 bytes).
struct BfsRange {
    const int32_t* col;  // first column of the range
    int32_t* level;      // the job's level array
    unsigned int start;  // exclusive prefix of the wavefront's range lengths
    unsigned int job;
    int next;            // level written to the range's targets
    unsigned int pad;
};

// Relaxes the edge ranges of all active lanes together: edge e of the
// concatenated ranges goes to lane (e mod active count), so every lane has
// work until the wavefront's edges run out, whatever the degree skew. Each
// active lane publishes its range at its rank among the active lanes (rows
// [0, n) of its wavefront's LDS block are exactly the live ones), and the
// prefix of range lengths comes from ballots over their bit slices, so
// inactive lanes are never read.
template <typename PROGRAM>
__device__ inline void bfs_relax_wave(PROGRAM& prog, unsigned int job, const BfsJob& j, int begin, int end,
                                      int next) {
    constexpr unsigned int WAVES = (unsigned int)(PROGRAM::WORK_GROUP_SIZE / 64);
    __shared__ BfsRange ranges[WAVES][64];
    BfsRange* row = ranges[hz_block_thread() / 64u];
    const uint64_t live = hz::wave::active();
    const unsigned int n = hz::wave::popc(live);
    const unsigned int me = hz::wave::rank(live);
    const unsigned int len = (unsigned int)(end - begin);
    unsigned int start = 0u, total = 0u;
    for (unsigned int b = 0; b < 32u && hz::wave::ballot((len >> b) != 0u) != 0ull; ++b) {
        const uint64_t m = hz::wave::ballot(((len >> b) & 1u) != 0u);
        start += hz::wave::rank(m) << b;
        total += hz::wave::popc(m) << b;
    }
    row[me] = BfsRange{j.col + begin, j.level, start, job, next, 0u};
    hz::wave::sync();
    // HZ_BFS_CHAINS edges per lane per round keep that many memory chains in
    // flight.
    constexpr int CH = HZ_BFS_CHAINS;
    for (unsigned int base = 0; base < total; base += (unsigned int)CH * n) {
        int w[CH], seen[CH], before[CH];
        const BfsRange* owner[CH];
        bool has[CH];
#pragma unroll
        for (int k = 0; k < CH; ++k) {
            const unsigned int e = base + (unsigned int)k * n + me;
            has[k] = e < total;
            owner[k] = row;
            w[k] = 0;
            if (has[k]) {
                // The last range starting at or before e holds it: ranges
                // of length 0 share their successor's start.
                unsigned int lo = 0u, hi = n - 1u;
                while (lo < hi) {
                    const unsigned int mid = (lo + hi + 1u) >> 1;
                    if (row[mid].start <= e) lo = mid; else hi = mid - 1u;
                }
                owner[k] = row + lo;
                w[k] = owner[k]->col[e - owner[k]->start];
            }
        }
#pragma unroll
        for (int k = 0; k < CH; ++k) seen[k] = has[k] ? hz_load(&owner[k]->level[w[k]]) : 0;
#pragma unroll
        for (int k = 0; k < CH; ++k) {
            const int nx = owner[k]->next;
            before[k] = has[k] && seen[k] > nx ? atomicMin(&owner[k]->level[w[k]], nx) : nx;
        }
#pragma unroll
        for (int k = 0; k < CH; ++k) {
            if (has[k] && before[k] > owner[k]->next) {
                prog.template async_call_cast<BfsVisit>(1, Promise<BfsVisit>(owner[k]->job, w[k], owner[k]->next));
            }
        }
    }
    hz::wave::sync();
}
#endif

template <typename PROGRAM>
__device__ inline void bfs_relax(PROGRAM& prog, unsigned int job, const BfsJob& j, int begin, int end, int next) {
    unsigned int split_limit = prog.device.split;
    if (split_limit < 1u) split_limit = 1u;
    if (split_limit > BFS_MAX_SPLIT) split_limit = BFS_MAX_SPLIT;
    const int split = (int)split_limit;
    while (end - begin > split) {
        int piece = (end - begin + BFS_FANOUT - 1) / BFS_FANOUT;
        if (piece < split) piece = split;
        for (long long b = (long long)begin + piece; b < end; b += piece) {
            prog.template async<BfsEdges>(job, (int)b, end - b > piece ? (int)b + piece : end, next);
        }
        end = begin + piece;
    }
#if defined(__HIP_DEVICE_COMPILE__)
    bfs_relax_wave(prog, job, j, begin, end, next);
#else
    // Four edges at a time: their column reads, level reads and atomics are
    // independent, so a lane keeps four memory chains in flight.
    int i = begin;
    for (; i + 4 <= end; i += 4) {
        int w[4], seen[4], before[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) w[k] = j.col[i + k];
#pragma unroll
        for (int k = 0; k < 4; ++k) seen[k] = hz_load(&j.level[w[k]]);
#pragma unroll
        for (int k = 0; k < 4; ++k) before[k] = seen[k] > next ? atomicMin(&j.level[w[k]], next) : next;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            if (before[k] > next) prog.template async_call_cast<BfsVisit>(1, Promise<BfsVisit>(job, w[k], next));
        }
    }
    for (; i < end; ++i) {
        int w = j.col[i];
        if (hz_load(&j.level[w]) > next) {
            int before = atomicMin(&j.level[w], next);
            if (before > next) prog.template async_call_cast<BfsVisit>(1, Promise<BfsVisit>(job, w, next));
        }
    }
#endif
}

// GROUPS: wavefront groups per workgroup (the launch's work-group size / 64);
// STASH: LDS links per group (fewer links, more resident wavefronts).
template <unsigned int GROUPS, unsigned int STASH = 16>
struct BfsSpecT {
    typedef OpUnion<BfsVisit, BfsEdges> OpSet;
    typedef BfsState DeviceState;

    static const size_t STASH_SIZE = STASH;
    static const size_t FRAME_SIZE = 8;
    static const size_t POOL_SIZE = 8;
    static const size_t GROUPS_PER_BLOCK = GROUPS;
    // A search grows from one vertex per graph: hand links to idle groups.
    static const size_t SHARE_WORK = 1;
    // Run the links that filled first first, closer to level order: fewer
    // vertices are visited again at a lower level (77 M -> 69 M promises at
    // the default configuration, 57.6 -> 50.7 ms; the photon walk and
    // Collatz are faster without it).
    static const size_t OLDEST_FIRST = 1;
    // Depth levels = BFS levels (clamped), run shallowest first: a group runs
    // its lowest-level links before deeper ones, as a level-synchronous
    // search would, so fewer vertices are visited again at a lower level.
    // Four levels took the 64-graph batch from 45.3 M to 37.2 M visits and
    // 42.1 to 38.4 ms at the tuned configuration (two levels: 39.6 ms). The
    // stash then holds at least 2 x 4 + 3 = 11 links. The event program
    // ignores the levels.
#ifndef HZ_BFS_LEVELS
#define HZ_BFS_LEVELS 4
#endif
    static const size_t STACK_SIZE = HZ_BFS_LEVELS;
    static const size_t SHALLOW_FIRST = 1;

    template <typename PROGRAM>
    __device__ static void initialize(PROGRAM) {}

    template <typename PROGRAM>
    __device__ static void finalize(PROGRAM) {}

    // Starts one job per thread: level[source] = 0 and a visit of the source.
    template <typename PROGRAM>
    __device__ static bool make_work(PROGRAM prog) {
        util::iter::Iter<unsigned int> it = prog.device.job_iter->leap(1u);
        unsigned int job;
        while (it.step(job)) {
            const BfsJob j = prog.device.jobs[job];
            if (j.source < j.vertices) {
                atomicMin(&j.level[j.source], 0);
                prog.template async<BfsVisit>(job, (int)j.source, 0);
            }
        }
        return !prog.device.job_iter->done();
    }
};

using BfsSpec = BfsSpecT<1>;
