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

namespace {

bool pow2(uint32_t x) { return x != 0 && (x & (x - 1)) == 0; }

PhotonState make_state(const hz_photon_params& p, double* tally, uint32_t* events,
                       util::iter::AtomicIter<unsigned int>* iter) {
    if (p.bins == 0) throw std::invalid_argument("photon: bins must be positive");
    if (!pow2(p.replicas)) throw std::invalid_argument("photon: replicas must be a power of two");
    if (!pow2(p.max_split) || p.max_split > 1024) throw std::invalid_argument("photon: max_split must be a power of two <= 1024");
    if (p.regions == 0 || p.regions > 120) throw std::invalid_argument("photon: regions must be in [1, 120]");
    if (!(p.thickness > 0.0f) || !(p.sigma_t > 0.0f) || !(p.region_width > 0.0f)) {
        throw std::invalid_argument("photon: thickness, sigma_t and region_width must be positive");
    }
    if (!(p.albedo >= 0.0f && p.albedo <= 1.0f)) throw std::invalid_argument("photon: albedo must be in [0, 1]");
    if (p.max_events == 0 || p.max_events > 65535u) {
        throw std::invalid_argument("photon: max_events must be in [1, 65535]");
    }
    PhotonState st{};
    st.photon_count = p.photons;
    st.seed_lo = (uint32_t)p.seed;
    st.seed_hi = (uint32_t)(p.seed >> 32);
    st.bins = p.bins;
    st.replicas = p.replicas;
    st.max_split = p.max_split;
    st.max_events = p.max_events;
    st.regions = p.regions;
    st.thickness = p.thickness;
    st.mean_free_path = 1.0f / p.sigma_t;
    st.albedo = p.albedo;
    st.bin_scale = (float)p.bins / p.thickness;
    st.region_scale = 1.0f / p.region_width;
    st.split_above = p.split_above;
    st.roulette_below = p.roulette_below;
    st.tally = tally;
    st.events = events;
    st.iterator = iter;
    return st;
}

}  // namespace

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
            hz::capi::run_async<PhotonAsync>(inst, groups, cycles, cfg.max_launches, stats);
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
