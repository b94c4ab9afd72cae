// Validation of hz_photon_params and its translation into the photon
// specification's device state; shared by the photon entry points and the
// serving loop.
#pragma once

#include <stdexcept>

#include "harmonize_amd.h"
#include "photon.h"

namespace hz {
namespace photon_abi {

inline bool pow2(uint32_t x) { return x != 0 && (x & (x - 1)) == 0; }

inline PhotonState make_state(const hz_photon_params& p, double* tally, uint32_t* events,
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

}  // namespace photon_abi
}  // namespace hz
