// Photon-walk programs of the C ABI: EventProgram and AsyncProgram over
// PhotonSpec.
#include "common.h"
#include "photon.h"

using PhotonEvent = EventProgram<PhotonSpec>;
using PhotonAsync = AsyncProgram<PhotonSpec>;

static_assert(sizeof(hz_photon_state) == sizeof(PhotonState), "C and C++ photon states must match");
static_assert(sizeof(Photon) == 40, "photon promises are 10 words");

HZ_DEFINE_PROGRAM_ABI(photon_event, PhotonEvent, PhotonState)
HZ_DEFINE_PROGRAM_ABI(photon_async, PhotonAsync, PhotonState)

using hz::capi::check;
using hz::capi::guarded;

__global__ void __launch_bounds__(256)
    detmath_kernel(const double* in, double* out_log, double* out_sqrt, uint32_t* out_philox, uint32_t n) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        out_log[i] = hz::det::det_log(in[i]);
        out_sqrt[i] = hz::det::det_sqrt(in[i]);
        hz::philox::Ctr r = hz::philox::philox4x32_10(i, i * 3u, ~i, 7u, 0x5EEDu, i ^ 0xABCDu);
        for (int k = 0; k < 4; ++k) out_philox[4 * (size_t)i + k] = r.v[k];
    }
}

extern "C" int hz_detmath_eval(const double* d_in, double* d_log, double* d_sqrt, uint32_t* d_philox, uint32_t n) {
    return guarded([&] {
        if (n == 0) return;
        unsigned blocks = (n + 255u) / 256u;
        if (blocks > 4096u) blocks = 4096u;
        hipLaunchKernelGGL(detmath_kernel, dim3(blocks), dim3(256), 0, nullptr, d_in, d_log, d_sqrt, d_philox, n);
        check(hipGetLastError(), "detmath_kernel launch");
        check(hipDeviceSynchronize(), "detmath_kernel");
    });
}

extern "C" void hz_photon_defaults(hz_photon_params* p) {
    if (!p) return;
    p->photons = 10000000u;
    p->bins = 1000u;
    p->seed = 0x5EEDull;
    p->thickness = 20.0f;
    p->sigma_t = 1.0f;
    p->albedo = 0.9f;
    p->split_above = 2.0f;
    p->roulette_below = 0.1f;
    p->region_width = 2.0f;
    p->regions = 10u;
    p->max_split = 8u;
    p->max_events = 65535u;
    p->replicas = 256u;
}

#include "photon_state.h"

using hz::photon_abi::make_state;

extern "C" int hz_photon_run(int kind, const hz_photon_params* params, double* d_tally, uint32_t* d_events,
                             const hz_run_config* config, hz_run_stats* stats) {
    return guarded([&] {
        if (!params) throw std::invalid_argument("hz_photon_run: null params");
        hz_photon_params p = *params;
        hz_run_config cfg{};
        if (config) cfg = *config;
        if (stats) std::memset(stats, 0, sizeof(*stats));
        if (p.photons == 0) return;
        if (!d_tally || !d_events) throw std::invalid_argument("hz_photon_run: null tally or events");
        hz::capi::DeviceIter iter(p.photons);
        PhotonState st = make_state(p, d_tally, d_events, iter.ptr);
        size_t tally_bytes = (size_t)p.replicas * (p.bins + HZ_PHOTON_EXTRA_SLOTS) * sizeof(double);
        check(hipMemset(d_tally, 0, tally_bytes), "zero tally");
        check(hipMemset(d_events, 0, (size_t)p.photons * sizeof(uint32_t)), "zero events");

        if (kind == HZ_ASYNC) {
            uint32_t groups = cfg.groups ? cfg.groups
                                         : (uint32_t)hz::capi::resident_groups_of<PhotonAsync>();
            uint64_t arena = cfg.arena ? cfg.arena : (1ull << 16);
            uint32_t cycles = cfg.cycles ? cfg.cycles : (1u << 24);
            PhotonAsync::Instance inst(arena, st, cfg.queues, cfg.coalesce ? cfg.coalesce : 64u);
            init<PhotonAsync>(inst, groups);
            check(hipDeviceSynchronize(), "init");
            hz::capi::run_async<PhotonAsync>(inst, groups, cycles, cfg.max_launches, stats,
                                             (cfg.flags & HZ_RUN_AUDIT) != 0);
        } else if (kind == HZ_EVENT) {
            uint32_t groups = cfg.groups ? cfg.groups : (uint32_t)hz::capi::compute_units() * 32u;
            uint64_t io = cfg.arena;
            if (io == 0) {
                // Splitting can grow a generation well past the source count.
                io = 1ull << 20;
                while (io < p.photons && io < (1ull << 24)) io <<= 1;
            }
            uint32_t chunk = cfg.cycles ? cfg.cycles : 4u;
            PhotonEvent::Instance inst(io, st);
            hz::capi::run_event<PhotonEvent>(inst, groups, chunk, cfg.max_launches, stats);
        } else {
            throw std::invalid_argument("hz_photon_run: kind must be HZ_EVENT or HZ_ASYNC");
        }
    });
}
