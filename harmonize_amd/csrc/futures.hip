// Futures program of the C ABI: the value-returning reduction tree of
// specs/futures.h, under AsyncProgram (tree and readers) or EventProgram
// (tree only: it cannot hold promises on a barrier).
#include "common.h"
#include "futures.h"

using FuturesAsync = AsyncProgram<FuturesSpec>;
using FuturesEvent = EventProgram<FuturesSpec>;

using hz::capi::check;
using hz::capi::guarded;

namespace {

using TotalFuture = Future<unsigned long long, RemappingBarrier<FuturesOps>>;
using UnitFuture = Future<unsigned long long, UnitBarrier<FuturesOps>>;

__global__ void futures_setup_kernel(FuturesState st, unsigned int count) {
    for (unsigned int k = blockIdx.x * blockDim.x + threadIdx.x; k < count; k += gridDim.x * blockDim.x) {
        futures_setup_node(st, k);
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) *st.total = TotalFuture(1u);
}

// out[r] = forwarded[r].data if that future resolved, else FUTURES_EARLY.
__global__ void futures_collect_kernel(const UnitFuture* forwarded, unsigned int readers, unsigned long long* out) {
    for (unsigned int r = blockIdx.x * blockDim.x + threadIdx.x; r < readers; r += gridDim.x * blockDim.x) {
        out[r] = forwarded[r].barrier.semaphore.sem == 0u ? forwarded[r].data : FUTURES_EARLY;
    }
}

unsigned int grid_of(unsigned int n) {
    unsigned int blocks = (n + 255u) / 256u;
    return blocks == 0u ? 1u : (blocks > 4096u ? 4096u : blocks);
}

}  // namespace

extern "C" int hz_futures_run(int kind, const uint64_t* d_starts, uint32_t leaves, uint32_t readers,
                              uint64_t* d_echo, uint64_t* d_forwarded, uint64_t* total, const hz_run_config* config,
                              hz_run_stats* stats) {
    return guarded([&] {
        if (leaves == 0 || (leaves & (leaves - 1u)) != 0 || leaves > (1u << 26)) {
            throw std::invalid_argument("hz_futures_run: leaves must be a power of two in [1, 2^26]");
        }
        if (readers > (1u << 24)) throw std::invalid_argument("hz_futures_run: at most 2^24 readers");
        if (!d_starts || !total || (readers && (!d_echo || !d_forwarded))) {
            throw std::invalid_argument("hz_futures_run: null argument");
        }
        if (kind == HZ_EVENT && readers) {
            throw std::invalid_argument("hz_futures_run: readers await a RemappingBarrier, which EventProgram "
                                        "cannot hold promises on; use HZ_ASYNC");
        }
        if (kind != HZ_EVENT && kind != HZ_ASYNC) throw std::invalid_argument("hz_futures_run: unknown program kind");
        hz_run_config cfg{};
        if (config) cfg = *config;
        if (stats) std::memset(stats, 0, sizeof(*stats));

        util::host::DevBuf<unsigned long long> slot((size_t)leaves * 2u);
        util::host::DevBuf<UnitBarrier<FuturesOps>> join((size_t)leaves);
        util::host::DevBuf<TotalFuture> total_future((size_t)1);
        util::host::DevBuf<UnitFuture> forwarded((size_t)(readers ? readers : 1u));
        hz::capi::DeviceIter iter(leaves + readers);

        FuturesState st{};
        st.starts = (const unsigned long long*)d_starts;
        st.slot = slot.data();
        st.join = join.data();
        st.total = total_future.data();
        st.forwarded = forwarded.data();
        st.echo = (unsigned long long*)d_echo;
        st.iter = iter.ptr;
        st.leaves = leaves;
        st.readers = readers;

        unsigned int setup = leaves > readers ? leaves : readers;
        hipLaunchKernelGGL(futures_setup_kernel, dim3(grid_of(setup)), dim3(256), 0, nullptr, st, setup);
        check(hipGetLastError(), "futures_setup_kernel launch");
        check(hipDeviceSynchronize(), "futures_setup_kernel");

        if (kind == HZ_ASYNC) {
            uint32_t groups = cfg.groups ? cfg.groups : (uint32_t)hz::capi::resident_groups_of<FuturesAsync>();
            // Held Echo and ReturnOp promises occupy arena links until the total resolves.
            uint64_t arena = cfg.arena ? cfg.arena : (1ull << 16) + 2ull * readers / 64u + 64u;
            uint32_t cycles = cfg.cycles ? cfg.cycles : (1u << 24);
            FuturesAsync::Instance inst(arena, st, cfg.queues, cfg.coalesce ? cfg.coalesce : 64u);
            init<FuturesAsync>(inst, groups);
            check(hipDeviceSynchronize(), "init");
            hz::capi::run_async<FuturesAsync>(inst, groups, cycles, cfg.max_launches, stats,
                                              (cfg.flags & HZ_RUN_AUDIT) != 0, true);
        } else {
            uint32_t groups = cfg.groups ? cfg.groups : (uint32_t)hz::capi::resident_groups_of<FuturesEvent>();
            uint64_t io = cfg.arena;
            if (io == 0) {
                io = 1ull << 16;
                while (io < leaves) io <<= 1;
            }
            uint32_t chunk = cfg.cycles ? cfg.cycles : 4u;
            FuturesEvent::Instance inst(io, st);
            hz::capi::run_event<FuturesEvent>(inst, groups, chunk, cfg.max_launches, stats);
        }

        TotalFuture host_total;
        check(hipMemcpy(&host_total, total_future.data(), sizeof(TotalFuture), hipMemcpyDeviceToHost), "total");
        if (host_total.barrier.semaphore.sem != 0u) {
            throw std::runtime_error("hz_futures_run: the program completed without resolving the total");
        }
        *total = host_total.data;
        if (readers) {
            hipLaunchKernelGGL(futures_collect_kernel, dim3(grid_of(readers)), dim3(256), 0, nullptr,
                               (const UnitFuture*)forwarded.data(), readers, (unsigned long long*)d_forwarded);
            check(hipGetLastError(), "futures_collect_kernel launch");
            check(hipDeviceSynchronize(), "futures_collect_kernel");
        }
    });
}
