// Collatz programs of the C ABI: EventProgram and AsyncProgram over
// CollatzSpec, the one-thread-per-start comparison kernel (the analogue of
// the reference's examples/cpp/collaz/collaz_std.cpp:5-23), and the seeded
// start generator.
#include "common.h"
#include "collatz.h"
#include "philox.h"
#include "collatz_kernels.h"

using hz::collatz::collatz_starts_kernel;

using CollatzEvent = EventProgram<CollatzSpec>;
using CollatzAsync = AsyncProgram<CollatzSpec>;

static_assert(sizeof(hz_collatz_state) == sizeof(CollatzState), "C and C++ Collatz states must match");
static_assert(CollatzAsync::WORDS == 4, "Collatz promises are 16 bytes");

HZ_DEFINE_PROGRAM_ABI(collatz_event, CollatzEvent, CollatzState)
HZ_DEFINE_PROGRAM_ABI(collatz_async, CollatzAsync, CollatzState)

using hz::capi::check;
using hz::capi::guarded;

// One thread per start, no runtime: the lower bound any scheduler is compared to.
__global__ void __launch_bounds__(256)
    collatz_plain_kernel(const unsigned long long* starts, unsigned int count, unsigned int* steps) {
    for (unsigned int i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += gridDim.x * blockDim.x) {
        unsigned long long v = starts ? starts[i] : (unsigned long long)i;
        unsigned int s = 0;
        bool overflow = false;
        while (v > 1) {
            if ((v & 1ull) == 0) {
                v >>= 1;
            } else {
                if (v > 0x5555555555555554ull) {
                    overflow = true;
                    break;
                }
                v = 3ull * v + 1ull;
            }
            s++;
        }
        steps[i] = overflow ? COLLATZ_OVERFLOW : s;
    }
}

namespace {

uint32_t default_groups_async() {
    return (uint32_t)hz::capi::resident_groups_of<CollatzAsync>();
}

uint32_t default_groups_event() {
    // Enough wavefronts to fill the device several times over; they need not be resident.
    return (uint32_t)hz::capi::compute_units() * 32u;
}

}  // namespace

extern "C" int hz_collatz_run(int kind, const uint64_t* d_starts, uint32_t count, uint32_t* d_steps,
                              const hz_run_config* config, hz_run_stats* stats) {
    return guarded([&] {
        if (d_steps == nullptr && count > 0) throw std::invalid_argument("hz_collatz_run: null steps");
        hz_run_config cfg{};
        if (config) cfg = *config;
        if (stats) std::memset(stats, 0, sizeof(*stats));
        if (count == 0) return;

        if (kind == HZ_PLAIN) {
            hz::capi::EventTimer t;
            t.start();
            unsigned int blocks = (count + 255u) / 256u;
            if (blocks > 65535u * 4u) blocks = 65535u * 4u;
            hipLaunchKernelGGL(collatz_plain_kernel, dim3(blocks), dim3(256), 0, nullptr,
                               (const unsigned long long*)d_starts, count, d_steps);
            check(hipGetLastError(), "collatz_plain_kernel launch");
            double ms = t.stop();
            if (stats) {
                stats->launches = 1;
                stats->elapsed_ms = ms;
                stats->groups = blocks;
            }
            return;
        }

        hz::capi::DeviceIter iter(count);
        CollatzState ds{(const unsigned long long*)d_starts, d_steps, iter.ptr};

        if (kind == HZ_ASYNC) {
            uint32_t groups = cfg.groups ? cfg.groups : default_groups_async();
            uint64_t arena = cfg.arena ? cfg.arena : (1ull << 16);
            uint32_t cycles = cfg.cycles ? cfg.cycles : (1u << 24);
            CollatzAsync::Instance inst(arena, ds, cfg.queues, cfg.coalesce ? cfg.coalesce : 64u);
            init<CollatzAsync>(inst, groups);
            check(hipDeviceSynchronize(), "init");
            hz::capi::run_async<CollatzAsync>(inst, groups, cycles, cfg.max_launches, stats,
                                              (cfg.flags & HZ_RUN_AUDIT) != 0);
        } else if (kind == HZ_EVENT) {
            uint32_t groups = cfg.groups ? cfg.groups : default_groups_event();
            uint64_t io = cfg.arena;
            if (io == 0) {
                io = 1ull << 16;
                while (io < count && io < (1ull << 26)) io <<= 1;
            }
            uint32_t chunk = cfg.cycles ? cfg.cycles : 4u;
            CollatzEvent::Instance inst(io, ds);
            hz::capi::run_event<CollatzEvent>(inst, groups, chunk, cfg.max_launches, stats);
        } else {
            throw std::invalid_argument("hz_collatz_run: unknown program kind");
        }
    });
}

extern "C" int hz_collatz_make_starts(uint64_t seed, uint32_t bits, uint32_t count, uint64_t* d_out) {
    return guarded([&] {
        if (bits < 1 || bits > 63) throw std::invalid_argument("hz_collatz_make_starts: bits must be in [1, 63]");
        if (count == 0) return;
        if (d_out == nullptr) throw std::invalid_argument("hz_collatz_make_starts: null output");
        unsigned int blocks = (count + 255u) / 256u;
        if (blocks > 8192u) blocks = 8192u;
        hipLaunchKernelGGL(collatz_starts_kernel, dim3(blocks), dim3(256), 0, nullptr, (unsigned long long)seed, bits,
                           count, (unsigned long long*)d_out);
        check(hipGetLastError(), "collatz_starts_kernel launch");
        check(hipDeviceSynchronize(), "collatz_starts_kernel");
    });
}

extern "C" int hz_device_query(hz_device_facts* facts) {
    return guarded([&] {
        if (!facts) throw std::invalid_argument("hz_device_query: null facts");
        std::memset(facts, 0, sizeof(*facts));
        int dev = 0;
        check(hipGetDevice(&dev), "hipGetDevice");
        hipDeviceProp_t prop;
        check(hipGetDeviceProperties(&prop, dev), "hipGetDeviceProperties");
        facts->device = dev;
        facts->compute_units = prop.multiProcessorCount;
        facts->total_memory = (int64_t)prop.totalGlobalMem;
        std::strncpy(facts->name, prop.name, sizeof(facts->name) - 1);
        std::strncpy(facts->arch, prop.gcnArchName, sizeof(facts->arch) - 1);
        facts->max_groups_collatz_async = (int32_t)default_groups_async();
        facts->max_groups_collatz_event = (int32_t)hz::capi::resident_groups_of<CollatzEvent>();
    });
}
