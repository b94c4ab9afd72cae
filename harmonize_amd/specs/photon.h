// Branching optical-photon random walk in a 3-D slab of scattering medium.
//
// The workload is defined by its configuration, not by a reference example;
// its structure follows the reference's neutron example (one step operation
// that re-queues the particle and spawns extra particles through async calls,
// examples/cpp/example_neutron/program.cpp:11-44, logic.cpp:179-248). The
// physical choices the configuration leaves open are fixed here:
//  - slab 0 <= z <= thickness, unbounded in x and y; photons start at the
//    origin moving along +z with weight 1;
//  - one promise = one event: a flight of length -ln(u)/sigma_t, then either
//    escape through a face (weight tallied as reflected or transmitted) or a
//    collision; at a collision the photon deposits w - w*albedo in its depth
//    bin (implicit capture) and keeps weight w*albedo;
//  - weight window by depth: target weight 2^-k in region k = floor(z / 2)
//    (importance doubles every two mean free paths). With r = w / target,
//    r > 2 splits the photon into c = the largest power of two <= r
//    (c <= max_split) copies of weight w / c (exact: division by a power of
//    two); r < 0.1 plays Russian roulette: survive with probability r at
//    weight = target, else die;
//  - every survivor scatters isotropically (Marsaglia's 1972 disc method)
//    and is re-queued with event index + 1; copy 0 keeps the lineage, copy
//    j > 0 gets a new lineage hashed from (lineage, event, j);
//  - randomness is Philox4x32-10 with key (seed) and counter
//    (photon id, event index, lineage, stream). Stream 0 gives four words:
//    the flight, the roulette, and the first direction attempt of copy 0;
//    stream 1 + 256 j + a gives attempt a of the rejection loop for copy j
//    (for copy 0 after its first attempt);
//  - tallies: per depth bin, FP64 sums of deposited weight, spread over
//    `replicas` copies chosen by group to avoid same-address atomics;
//    per photon, the number of events of its whole family (u32);
//  - a lineage is cut after max_events (<= 65535) events, its weight tallied
//    as capped; event index and pending count are 16-bit, so a promise is
//    40 bytes.
// Weight bookkeeping is exact per event (deposit = w - w*albedo exactly in
// binary64, splits divide by powers of two), so launched weight equals
// absorbed + reflected + transmitted + lost - gained to rounding of the FP64
// sums.
//
// Written against the spec API only; compiled for gfx950 by the runtime and
// for the host by the oracle's executors. All arithmetic is IEEE add/mul and
// conversions (detmath.h), so host and GPU agree bit for bit.
#pragma once

#include <cstdint>

#include "detmath.h"
#include "philox.h"

// Slots after the depth bins in each tally replica.
enum PhotonTallySlot : unsigned int {
    PHOTON_REFLECTED = 0,
    PHOTON_TRANSMITTED = 1,
    PHOTON_ROULETTE_LOST = 2,
    PHOTON_ROULETTE_GAINED = 3,
    PHOTON_CAPPED = 4,
    PHOTON_EXTRA_SLOTS = 5,
};

struct PhotonState {
    unsigned int photon_count;
    unsigned int seed_lo;
    unsigned int seed_hi;
    unsigned int bins;         // depth bins over [0, thickness)
    unsigned int replicas;     // tally replicas, a power of two
    unsigned int max_split;    // a power of two
    unsigned int max_events;   // per lineage, <= 65535
    unsigned int regions;      // weight-window regions
    float thickness;
    float mean_free_path;      // 1 / sigma_t
    float albedo;
    float bin_scale;           // bins / thickness
    float region_scale;        // 1 / region width
    float split_above;         // 2
    float roulette_below;      // 0.1
    float pad;
    double* tally;             // [replicas][bins + PHOTON_EXTRA_SLOTS]
    unsigned int* events;      // [photon_count]
    util::iter::AtomicIter<unsigned int>* iterator;
};

struct Photon {
    float x, y, z;
    float ux, uy, uz;
    float w;
    unsigned int id;
    unsigned int lineage;
    unsigned short event;
    unsigned short pending;  // events of this lineage not yet added to events[id]
};

namespace photon_detail {

__host__ __device__ inline float float_of_bits(unsigned int b) {
    float f;
    std::memcpy(&f, &b, sizeof f);
    return f;
}

// 2^-k for 0 <= k < 126, built from its bit pattern.
__host__ __device__ inline float pow2_neg(unsigned int k) { return float_of_bits((127u - k) << 23); }

// 2^k for 0 <= k <= 127, built from its bit pattern (exact; a shift of a
// 32-bit integer would be undefined from k = 32 on).
__host__ __device__ inline float pow2(unsigned int k) { return float_of_bits((127u + k) << 23); }

// Outcome of the weight window at one collision.
enum WindowFate : unsigned int {
    WINDOW_KEEP = 0,      // inside the window: one copy at the same weight
    WINDOW_SURVIVE = 1,   // roulette won: one copy at the target weight
    WINDOW_DIE = 2,       // roulette lost
    WINDOW_SPLIT = 3,     // `copies` copies of weight w / copies
};

struct WindowResult {
    unsigned int fate;
    unsigned int copies;
    float w;  // weight of each copy (the weight lost, for WINDOW_DIE)
};

// Weight window at a collision in region `region` for a photon of weight w
// (after implicit capture). With target = 2^-region and r = w / target:
//  - r < roulette_below: roulette with u = (rbits >> 8) / 2^24 uniform in
//    [0, 1): survive with probability r at the target weight (so the expected
//    weight is w), else die;
//  - r > split_above: c = the largest power of two <= min(r, max_split)
//    copies of weight w / c (exact, c is a power of two);
//  - otherwise one copy of weight w.
__host__ __device__ inline WindowResult weight_window(const PhotonState& st, float w, unsigned int region,
                                                      unsigned int rbits) {
    WindowResult out{WINDOW_KEEP, 1u, w};
    float ratio = w * pow2(region);  // w / 2^-region, exact
    if (ratio < st.roulette_below) {
        float u2 = (float)(rbits >> 8) * (1.0f / 16777216.0f);  // [0, 1)
        if (u2 >= ratio) {
            out.fate = WINDOW_DIE;
            return out;
        }
        out.fate = WINDOW_SURVIVE;
        out.w = pow2_neg(region);
    } else if (ratio > st.split_above) {
        unsigned int k = 0;
        while ((2u << k) <= st.max_split && (float)(2u << k) <= ratio) ++k;
        out.fate = k > 0 ? WINDOW_SPLIT : WINDOW_KEEP;
        out.copies = 1u << k;
        out.w = w * pow2_neg(k);  // exact
    }
    return out;
}

// Integer hash "lowbias32" (C. Wellons, 2018) of (lineage, event, copy).
__host__ __device__ inline unsigned int child_lineage(unsigned int lineage, unsigned int event, unsigned int copy) {
    unsigned int h = lineage ^ (event * 0x9E3779B9u) ^ ((copy + 1u) * 0x85EBCA6Bu);
    h ^= h >> 16;
    h *= 0x7FEB352Du;
    h ^= h >> 15;
    h *= 0x846CA68Bu;
    h ^= h >> 16;
    return h;
}

// One attempt of Marsaglia's method: (v1, v2) uniform on the square; inside
// the unit disc it gives (2 v1 t, 2 v2 t, 1 - 2 s) with s = v1^2 + v2^2 and
// t = sqrt(1 - s).
__host__ __device__ inline bool marsaglia(unsigned int a, unsigned int b, float& ux, float& uy, float& uz) {
    double v1 = (double)(a >> 8) * (1.0 / 8388608.0) - 1.0;  // k / 2^23 - 1, exact
    double v2 = (double)(b >> 8) * (1.0 / 8388608.0) - 1.0;
    double s = v1 * v1 + v2 * v2;  // exact
    if (!(s > 0.0 && s < 1.0)) return false;
    double t = 2.0 * hz::det::det_sqrt(1.0 - s);
    ux = (float)(v1 * t);
    uy = (float)(v2 * t);
    uz = (float)(1.0 - 2.0 * s);
    return true;
}

// Isotropic unit direction for copy `copy` of photon p. `first` (two words,
// or null) is tried before the copy's own stream.
__host__ __device__ inline void isotropic(const PhotonState& st, const Photon& p, unsigned int copy,
                                          const unsigned int* first, float& ux, float& uy, float& uz) {
    if (first && marsaglia(first[0], first[1], ux, uy, uz)) return;
    for (unsigned int a = 0; a < 32; ++a) {
        hz::philox::Ctr q =
            hz::philox::philox4x32_10(p.id, p.event, p.lineage, 1u + (copy << 8) + a, st.seed_lo, st.seed_hi);
        if (marsaglia(q.v[0], q.v[1], ux, uy, uz)) return;
        if (marsaglia(q.v[2], q.v[3], ux, uy, uz)) return;
    }
    ux = 0.0f;  // reached with probability below 1e-40
    uy = 0.0f;
    uz = 1.0f;
}

// Group-level tally: with a persistent program type (groups live for the
// whole launch) deposits go to an FP64 tally in the group's LDS, flushed to
// the device tally once per launch in finalize; otherwise (and for more
// bins than fit) they go straight to the group's device replica. Device-wide
// FP64 atomics are the costliest part of an event otherwise.
constexpr unsigned int PHOTON_LDS_SLOTS = 1024u + PHOTON_EXTRA_SLOTS;

template <typename PROGRAM>
__device__ inline double* group_tally() {
    __shared__ double slots[PHOTON_LDS_SLOTS];
    return slots;
}

#ifndef HZ_PHOTON_GROUP_TALLY
#define HZ_PHOTON_GROUP_TALLY 1
#endif

template <typename PROGRAM>
__device__ inline bool uses_group_tally(const PhotonState& st) {
    if constexpr (PROGRAM::PERSISTENT && HZ_PHOTON_GROUP_TALLY) {
        return st.bins + PHOTON_EXTRA_SLOTS <= PHOTON_LDS_SLOTS;
    } else {
        return false;
    }
}

__device__ inline double* device_replica(const PhotonState& st) {
    return st.tally + (size_t)(hz_group_index() & (st.replicas - 1u)) * (st.bins + PHOTON_EXTRA_SLOTS);
}

template <typename PROGRAM>
__device__ inline void tally_add(const PhotonState& st, unsigned int slot, double value) {
    if constexpr (PROGRAM::PERSISTENT && HZ_PHOTON_GROUP_TALLY) {
        if (uses_group_tally<PROGRAM>(st)) {
            atomicAdd(&group_tally<PROGRAM>()[slot], value);
            return;
        }
    }
    atomicAdd(&device_replica(st)[slot], value);
}

}  // namespace photon_detail

namespace photon_detail {

// One event of photon p, as the configuration defines it. The caller
// supplies where its results go:
//  - tally(slot, value) books weight in a tally slot;
//  - end(q) ends the lineage of q (q.pending counts its events not yet added
//    to the photon's event total);
//  - split(copies) is told, before any copy is queued, that the photon splits
//    into `copies` lineages (copies > 1);
//  - emit(c) queues copy c.
template <typename TALLY, typename END, typename SPLIT, typename EMIT>
__host__ __device__ inline void event(const PhotonState& st, Photon p, TALLY&& tally, END&& end, SPLIT&& split,
                                      EMIT&& emit) {
    p.pending = (unsigned short)(p.pending + 1u);

    hz::philox::Ctr r = hz::philox::philox4x32_10(p.id, p.event, p.lineage, 0u, st.seed_lo, st.seed_hi);
    double u = (double)((r.v[0] >> 8) + 1u) * (1.0 / 16777216.0);  // (0, 1]
    float dist = (float)(-hz::det::det_log(u) * (double)st.mean_free_path);
    float nx = p.x + dist * p.ux;
    float ny = p.y + dist * p.uy;
    float nz = p.z + dist * p.uz;

    if (nz < 0.0f || nz > st.thickness) {
        tally(st.bins + (nz < 0.0f ? PHOTON_REFLECTED : PHOTON_TRANSMITTED), (double)p.w);
        end(p);
        return;
    }

    unsigned int bin = (unsigned int)(nz * st.bin_scale);
    if (bin >= st.bins) bin = st.bins - 1u;
    float w = p.w * st.albedo;
    tally(bin, (double)p.w - (double)w);  // exact difference

    unsigned int region = (unsigned int)(nz * st.region_scale);
    if (region >= st.regions) region = st.regions - 1u;
    WindowResult win = weight_window(st, w, region, r.v[1]);
    if (win.fate == WINDOW_DIE) {
        tally(st.bins + PHOTON_ROULETTE_LOST, (double)w);
        end(p);
        return;
    }
    if (win.fate == WINDOW_SURVIVE) {
        tally(st.bins + PHOTON_ROULETTE_GAINED, (double)win.w - (double)w);
    }
    const unsigned int copies = win.copies;
    w = win.w;

    if (p.event + 1u >= st.max_events) {
        tally(st.bins + PHOTON_CAPPED, (double)w * (double)copies);
        end(p);
        return;
    }

    if (copies > 1u) split(copies);
    for (unsigned int j = 0; j < copies; ++j) {
        Photon c;
        c.x = nx;
        c.y = ny;
        c.z = nz;
        c.w = w;
        c.id = p.id;
        c.event = (unsigned short)(p.event + 1u);
        isotropic(st, p, j, j == 0 ? &r.v[2] : nullptr, c.ux, c.uy, c.uz);
        c.lineage = (j == 0) ? p.lineage : child_lineage(p.lineage, p.event, j);
        c.pending = (j == 0) ? p.pending : (unsigned short)0;
        emit(c);
    }
}

}  // namespace photon_detail

struct PhotonStep {
    using Type = void (*)(Photon p);

    template <typename PROGRAM>
    __device__ static void eval(PROGRAM prog, Photon p) {
        const PhotonState& st = prog.device;
        photon_detail::event(
            st, p, [&](unsigned int slot, double value) { photon_detail::tally_add<PROGRAM>(st, slot, value); },
            [&](const Photon& q) { atomicAdd(&st.events[q.id], (unsigned int)q.pending); }, [](unsigned int) {},
            [&](const Photon& c) { prog.template async<PhotonStep>(c); });
    }
};

// Stash links per group, and photon ids each thread claims from the shared
// iterator at once. make_work launches one photon per thread per call from
// the thread's claimed block, so the shared counter (an atomic that every
// group hits) is touched once per HZ_PHOTON_CLAIM calls. Claims of 4 took
// the 1e7-photon walk from 5.62-5.77 ms (8) to 5.54-5.63 ms; claims of 2,
// to 7.0 ms.
#ifndef HZ_PHOTON_STASH
#define HZ_PHOTON_STASH 4
#endif
#ifndef HZ_PHOTON_CLAIM
#define HZ_PHOTON_CLAIM 4
#endif
// Below this many unclaimed photons a thread claims one at a time, so a
// small job (or the end of a large one) spreads its lineages over more
// wavefronts.
#ifndef HZ_PHOTON_SPREAD_BELOW
#define HZ_PHOTON_SPREAD_BELOW (1u << 17)
#endif
#ifndef HZ_PHOTON_BLOCK
#define HZ_PHOTON_BLOCK 4
#endif

// Photon ids a thread has claimed and not yet launched.
struct PhotonClaim {
    unsigned int next;
    unsigned int end;
};

namespace photon_detail {

template <typename PROGRAM>
__device__ inline void launch(PROGRAM& prog, unsigned int id) {
    Photon p;
    p.x = 0.0f;
    p.y = 0.0f;
    p.z = 0.0f;
    p.ux = 0.0f;
    p.uy = 0.0f;
    p.uz = 1.0f;
    p.w = 1.0f;
    p.id = id;
    p.event = 0u;
    p.lineage = 0u;
    p.pending = 0u;
    prog.template async<PhotonStep>(p);
}

}  // namespace photon_detail

struct PhotonSpec {
    typedef OpUnion<PhotonStep> OpSet;
    typedef PhotonState DeviceState;
    typedef PhotonClaim ThreadState;

    static const size_t STASH_SIZE = HZ_PHOTON_STASH;
    static const size_t FRAME_SIZE = 8;
    static const size_t POOL_SIZE = 8;
    // Four groups per workgroup share one LDS tally.
    static const size_t GROUPS_PER_BLOCK = HZ_PHOTON_BLOCK;
    // Link forwarding (one operation, so most links would forward) costs
    // the 1e7-photon walk 5.65 -> 6.05 ms, so it is off here.
    static const size_t FORWARD_LINKS = 0;

    template <typename PROGRAM>
    __device__ static void initialize(PROGRAM prog) {
        prog.thread.next = 0u;
        prog.thread.end = 0u;
        if (photon_detail::uses_group_tally<PROGRAM>(prog.device)) {
            double* t = photon_detail::group_tally<PROGRAM>();
            for (unsigned int i = hz_block_thread(); i < photon_detail::PHOTON_LDS_SLOTS; i += hz_block_threads()) {
                t[i] = 0.0;
            }
            __syncthreads();
        }
    }

    // Claimed ids not launched yet are queued now, so a pass that ends early
    // (cycle budget, halt) loses none of them; the group tally is added to
    // the group's device replica.
    template <typename PROGRAM>
    __device__ static void finalize(PROGRAM prog) {
        while (prog.thread.next < prog.thread.end) {
            photon_detail::launch(prog, prog.thread.next);
            prog.thread.next += 1u;
        }
        if (photon_detail::uses_group_tally<PROGRAM>(prog.device)) {
            __syncthreads();
            const double* t = photon_detail::group_tally<PROGRAM>();
            double* rep = photon_detail::device_replica(prog.device);
            const unsigned int slots = prog.device.bins + PHOTON_EXTRA_SLOTS;
            for (unsigned int i = hz_block_thread(); i < slots; i += hz_block_threads()) {
                if (t[i] != 0.0) atomicAdd(&rep[i], t[i]);
            }
        }
    }

    // Launches one photon per thread from the origin along +z.
    template <typename PROGRAM>
    __device__ static bool make_work(PROGRAM prog) {
        if (prog.thread.next >= prog.thread.end) {
            util::iter::AtomicIter<unsigned int>& photons = *prog.device.iterator;
            unsigned int at = hz_load(&photons.value);
            unsigned int left = at < photons.limit ? photons.limit - at : 0u;
            util::iter::Iter<unsigned int> it = photons.leap(left < HZ_PHOTON_SPREAD_BELOW ? 1u : HZ_PHOTON_CLAIM);
            prog.thread.next = it.value;
            prog.thread.end = it.limit;
        }
        if (prog.thread.next < prog.thread.end) {
            photon_detail::launch(prog, prog.thread.next);
            prog.thread.next += 1u;
            return true;
        }
        return !prog.device.iterator->done();
    }
};
