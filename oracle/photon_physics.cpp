// ORACLE — test infrastructure only (see host_shim.h).
//
// Direct access to the photon walk's two sampling decisions, so that tests
// can check them against what the mathematics fixes rather than against a
// second copy of the same code:
//  - the isotropic direction sampler (photon_detail::isotropic, Marsaglia's
//    disc method), exactly as PhotonStep draws it: copy 0 tries the stream-0
//    words first, copies j > 0 use their own streams;
//  - the weight window (photon_detail::weight_window): roulette below the
//    window, splitting above it.
// This file compiles on its own with the specification headers, so a test
// can also build it against a mutated copy of photon.h and show that the
// checks catch the mutation.
#include <cstdint>

#include "host_program.h"
#include "photon.h"

extern "C" {

// Directions of copy `copy` of photons with ids [0, n) at (event, lineage),
// under key `seed`: out[3i..3i+2] = (ux, uy, uz) of photon i.
void oracle_photon_directions(uint64_t seed, uint32_t event, uint32_t lineage, uint32_t copy, uint32_t n,
                              float* out) {
    PhotonState st{};
    st.seed_lo = (uint32_t)seed;
    st.seed_hi = (uint32_t)(seed >> 32);
    for (uint32_t i = 0; i < n; ++i) {
        Photon p{};
        p.id = i;
        p.event = (unsigned short)event;
        p.lineage = lineage;
        hz::philox::Ctr r = hz::philox::philox4x32_10(p.id, p.event, p.lineage, 0u, st.seed_lo, st.seed_hi);
        photon_detail::isotropic(st, p, copy, copy == 0 ? &r.v[2] : nullptr, out[3 * i], out[3 * i + 1],
                                 out[3 * i + 2]);
    }
}

// Weight-window decisions for n collisions of a photon of weight w in region
// `region`; the roulette word of collision i is word 1 of the stream-0 draw
// of photon id i (as in PhotonStep) under key `seed`. fate[i] is a
// photon_detail::WindowFate, copies[i] the number of copies and w_out[i]
// their weight.
void oracle_photon_window(float w, uint32_t region, float split_above, float roulette_below, uint32_t max_split,
                          uint64_t seed, uint32_t n, uint32_t* fate, uint32_t* copies, float* w_out) {
    PhotonState st{};
    st.split_above = split_above;
    st.roulette_below = roulette_below;
    st.max_split = max_split;
    for (uint32_t i = 0; i < n; ++i) {
        hz::philox::Ctr r = hz::philox::philox4x32_10(i, 0u, 0u, 0u, (uint32_t)seed, (uint32_t)(seed >> 32));
        photon_detail::WindowResult res = photon_detail::weight_window(st, w, region, r.v[1]);
        fate[i] = res.fate;
        copies[i] = res.copies;
        w_out[i] = res.w;
    }
}

}  // extern "C"
