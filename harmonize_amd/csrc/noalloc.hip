// hz_collatz_run_noalloc: the Collatz AsyncProgram run of hz_collatz_run
// (kind HZ_ASYNC) with every runtime buffer -- the work-claim iterator, the
// run status, the link queues and their slot rings, the link arena and its
// metadata -- carved from one fixed device array instead of device memory
// allocated per call. For callers that cannot allocate device memory. The
// kernels are the same as hz_collatz_run's; only where their buffers live
// differs.
//
// The array, `scratch`, is a file-scope __device__ object: every GPU the
// library runs on holds its own copy, at its own address, until the process
// ends. Host code never passes the array's name to a kernel (on the host it
// is only a stand-in for the symbol); the device address comes from
// hipGetSymbolAddress, taken once per device with that device current, and
// every launch receives that address. Calls share their GPU's array, so two
// calls that overlap on one GPU race; each call clears what must start at
// zero (queues, status, iterator) with kernels before it runs.
#include <chrono>
#include <mutex>

#include "common.h"
#include "collatz.h"

using hz::capi::check;
using hz::capi::guarded;

namespace {

using CollatzAsync = AsyncProgram<CollatzSpec>;
using ClaimIter = util::iter::AtomicIter<unsigned int>;

constexpr size_t round256(size_t n) { return (n + 255u) & ~(size_t)255u; }

// Byte offsets of the buffers inside the array. Queue q's slot ring starts at
// OFF_QUEUE_SLOTS + 4 * q * capacity, capacity = the arena's link count
// rounded up to a power of two (at least 64), so at most HZ_NOALLOC_MAX_LINKS.
constexpr size_t MAX_LINKS = HZ_NOALLOC_MAX_LINKS;
constexpr size_t MAX_QUEUES = HZ_NOALLOC_MAX_QUEUES;
constexpr size_t OFF_ITER = 0;
constexpr size_t OFF_STATUS = round256(OFF_ITER + sizeof(ClaimIter));
constexpr size_t OFF_QUEUES = round256(OFF_STATUS + sizeof(hz::AsyncStatus));
constexpr size_t OFF_FREE_QUEUE = round256(OFF_QUEUES + MAX_QUEUES * sizeof(hz::LinkQueue));
constexpr size_t OFF_QUEUE_SLOTS = round256(OFF_FREE_QUEUE + sizeof(hz::LinkQueue));
constexpr size_t OFF_FREE_SLOTS = round256(OFF_QUEUE_SLOTS + MAX_QUEUES * MAX_LINKS * 4u);
constexpr size_t OFF_ARENA = round256(OFF_FREE_SLOTS + MAX_LINKS * 4u);
constexpr size_t OFF_LINK_META = round256(OFF_ARENA + MAX_LINKS * CollatzAsync::LINK_BYTES);
constexpr size_t OFF_LINK_AUX = round256(OFF_LINK_META + MAX_LINKS * 4u);
constexpr size_t SCRATCH_BYTES = round256(OFF_LINK_AUX + 2u * MAX_LINKS * 4u);

static_assert((MAX_LINKS & (MAX_LINKS - 1u)) == 0 && MAX_LINKS >= 64, "the link bound must be a power of two");
static_assert(SCRATCH_BYTES <= HZ_NOALLOC_SCRATCH_BYTES, "the documented bound on the array is too small");

}  // namespace

static __device__ __align__(256) unsigned char scratch[SCRATCH_BYTES];

namespace {

constexpr int MAX_DEVICES = 64;
std::mutex scratch_mutex;
unsigned char* scratch_address[MAX_DEVICES] = {};

// The device address of the current device's copy of `scratch`, looked up
// once per device with that device current.
unsigned char* scratch_on_current_device() {
    int dev = -1;
    check(hipGetDevice(&dev), "hipGetDevice");
    if (dev < 0 || dev >= MAX_DEVICES) throw std::runtime_error("hz_collatz_run_noalloc: device index out of range");
    std::lock_guard<std::mutex> lock(scratch_mutex);
    if (scratch_address[dev] == nullptr) {
        void* p = nullptr;
        check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
        if (p == nullptr) throw std::runtime_error("hipGetSymbolAddress(scratch) returned a null address");
        scratch_address[dev] = (unsigned char*)p;
    }
    return scratch_address[dev];
}

__global__ void __launch_bounds__(64) claim_iter_reset_kernel(ClaimIter* it, unsigned int limit) {
    if (blockIdx.x == 0 && threadIdx.x == 0) *it = ClaimIter(0u, limit);
}

unsigned int pow2_at_least(uint64_t n) {
    unsigned int p = 64;
    while (p < n) p <<= 1;
    return p;
}

}  // namespace

extern "C" int hz_collatz_run_noalloc(const uint64_t* d_starts, uint32_t count, uint32_t* d_steps,
                                      const hz_run_config* config, hz_run_stats* stats) {
    return guarded([&] {
        // Every argument is checked before the device is touched.
        if (d_steps == nullptr && count > 0) throw std::invalid_argument("hz_collatz_run_noalloc: null steps");
        hz_run_config cfg{};
        if (config) cfg = *config;
        if (stats) std::memset(stats, 0, sizeof(*stats));
        if (cfg.arena > MAX_LINKS) {
            throw std::invalid_argument("hz_collatz_run_noalloc: arena of " + std::to_string(cfg.arena) +
                                        " links exceeds the static maximum of " + std::to_string(MAX_LINKS));
        }
        if (cfg.queues > MAX_QUEUES) {
            throw std::invalid_argument("hz_collatz_run_noalloc: " + std::to_string(cfg.queues) +
                                        " queues exceed the static maximum of " + std::to_string(MAX_QUEUES));
        }
        if (cfg.coalesce > 64) throw std::invalid_argument("hz_collatz_run_noalloc: coalesce must be at most 64");
        if (count == 0) return;

        unsigned char* base = scratch_on_current_device();
        const unsigned int links = cfg.arena ? (unsigned int)cfg.arena : (unsigned int)MAX_LINKS;
        const unsigned int capacity = pow2_at_least(links);

        CollatzAsync::DeviceContext ctx;
        ctx.status = (hz::AsyncStatus*)(base + OFF_STATUS);
        ctx.queues = (hz::LinkQueue*)(base + OFF_QUEUES);
        ctx.queue_slots = (unsigned int*)(base + OFF_QUEUE_SLOTS);
        ctx.free_queue = (hz::LinkQueue*)(base + OFF_FREE_QUEUE);
        ctx.free_slots = (unsigned int*)(base + OFF_FREE_SLOTS);
        ctx.arena = (unsigned int*)(base + OFF_ARENA);
        ctx.link_meta = (unsigned int*)(base + OFF_LINK_META);
        ctx.link_aux = (unsigned int*)(base + OFF_LINK_AUX);
        ctx.links = links;
        ctx.queue_count = cfg.queues ? cfg.queues : 8u;
        ctx.queue_mask = capacity - 1u;
        ctx.coalesce = cfg.coalesce ? cfg.coalesce : 64u;

        ClaimIter* iter = (ClaimIter*)(base + OFF_ITER);
        CollatzState ds{(const unsigned long long*)d_starts, d_steps, iter};

        const uint32_t groups = cfg.groups ? cfg.groups : (uint32_t)hz::capi::resident_groups_of<CollatzAsync>();
        const uint32_t cycles = cfg.cycles ? cfg.cycles : (1u << 24);
        const unsigned int blocks = (unsigned int)blocks_for<CollatzAsync>(groups);

        auto t0 = std::chrono::steady_clock::now();
        // Clear the queues, the status and the claim iterator (kernels, no memset).
        size_t slots = (size_t)capacity * (ctx.queue_count + 1u);
        unsigned int reset_blocks = (unsigned int)std::min<size_t>(1024, (slots + 255) / 256);
        hipLaunchKernelGGL(_hz_async_reset_kernel<CollatzAsync>, dim3(reset_blocks), dim3(256), 0, nullptr, ctx);
        check(hipGetLastError(), "reset launch");
        hipLaunchKernelGGL(claim_iter_reset_kernel, dim3(1), dim3(64), 0, nullptr, iter, count);
        check(hipGetLastError(), "iterator reset launch");

        hz::AsyncStatus st{};
        uint64_t launches = 0, last_progress = ~0ull;
        while (true) {
            if (cfg.max_launches && launches >= cfg.max_launches) {
                throw std::runtime_error("AsyncProgram: launch bound reached before completion");
            }
            hipLaunchKernelGGL(_hz_exec_kernel<CollatzAsync>, dim3(blocks), dim3((unsigned)CollatzAsync::WORK_GROUP_SIZE),
                               0, nullptr, ctx, ds, cycles);
            check(hipGetLastError(), "exec launch");
            check(hipDeviceSynchronize(), "exec");
            launches++;
            check(hipMemcpy(&st, ctx.status, sizeof(st), hipMemcpyDeviceToHost), "status read");
            if (CollatzAsync::status_complete(st)) break;
            if (st.promises == last_progress) {
                throw std::runtime_error("AsyncProgram: a launch made no progress before completion");
            }
            last_progress = st.promises;
        }
        double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        if (stats) {
            stats->promises = st.promises;
            stats->launches = launches;
            stats->spills = st.spills;
            stats->fills = st.fills;
            stats->make_work_calls = st.make_work_calls;
            stats->groups = blocks * (uint32_t)(CollatzAsync::WORK_GROUP_SIZE / 64);
            stats->elapsed_ms = ms;
            for (int i = 0; i < 8; ++i) stats->phase_cycles[i] = st.phase_cycles[i];
            stats->links_run = st.links_run;
            stats->forwarded = st.forwarded;
        }
    });
}
