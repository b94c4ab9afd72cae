// Shared plumbing of the C-ABI translation units: error capture, the
// per-program entry-point generator, and launch sizing.
#pragma once

#include <harmonize.h>

#include <chrono>
#include <cstring>
#include <string>

#include "harmonize_amd.h"

namespace hz {
namespace capi {

void set_error(const std::string& message);
void clear_error();

// Runs `fn`, turning exceptions into a status code and a stored message.
template <typename F>
int guarded(F&& fn) {
    try {
        clear_error();
        fn();
        return HZ_OK;
    } catch (const std::invalid_argument& e) {
        set_error(e.what());
        return HZ_ERR_ARGUMENT;
    } catch (const std::exception& e) {
        set_error(e.what());
        return HZ_ERR_RUNTIME;
    }
}

inline void check(hipError_t e, const char* what) {
    if (e != hipSuccess) {
        throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
    }
}

int compute_units();

// Groups (64-lane wavefronts) that fit on the device at once for `kernel`,
// launched in workgroups of `block_threads` threads.
template <typename K>
int resident_groups(K kernel, int block_threads = 64) {
    int per_cu = 0;
    check(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, block_threads, 0), "occupancy query");
    if (per_cu < 1) per_cu = 1;
    return per_cu * compute_units() * (block_threads / 64);
}

// The same for a program type's exec kernel.
template <typename P>
int resident_groups_of() {
    return resident_groups(_hz_exec_kernel<P>, (int)P::WORK_GROUP_SIZE);
}

// Owns a device AtomicIter<unsigned int> set to [0, limit).
struct DeviceIter {
    util::iter::AtomicIter<unsigned int>* ptr = nullptr;
    explicit DeviceIter(unsigned int limit) {
        util::iter::AtomicIter<unsigned int> host(0u, limit);
        check(hipMalloc((void**)&ptr, sizeof(host)), "hipMalloc(iterator)");
        check(hipMemcpy(ptr, &host, sizeof(host), hipMemcpyHostToDevice), "hipMemcpy(iterator)");
    }
    ~DeviceIter() {
        if (ptr) (void)hipFree(ptr);
    }
    DeviceIter(const DeviceIter&) = delete;
    DeviceIter& operator=(const DeviceIter&) = delete;
};

// Times a span on the null stream with hipEvents.
struct EventTimer {
    hipEvent_t a{}, b{};
    EventTimer() {
        check(hipEventCreate(&a), "hipEventCreate");
        check(hipEventCreate(&b), "hipEventCreate");
    }
    ~EventTimer() {
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
    }
    void start() { check(hipEventRecord(a, nullptr), "hipEventRecord"); }
    double stop() {
        check(hipEventRecord(b, nullptr), "hipEventRecord");
        check(hipEventSynchronize(b), "hipEventSynchronize");
        float ms = 0;
        check(hipEventElapsedTime(&ms, a, b), "hipEventElapsedTime");
        return ms;
    }
};

// Runs an AsyncProgram instance to completion: exec launches until complete().
// With `audit`, the instance's queues and arena are audited after every
// launch (Instance::audit) and the number of audits is reported.
template <typename P>
void run_async(typename P::Instance& inst, uint32_t groups, uint32_t cycles, uint32_t max_launches,
               hz_run_stats* stats, bool audit = false, bool barriers = false) {
    EventTimer t;
    t.start();
    uint64_t launches = 0;
    uint64_t last_progress = ~0ull;
    uint32_t audits = 0;
    do {
        if (max_launches && launches >= max_launches) {
            throw std::runtime_error("AsyncProgram: launch bound reached before completion");
        }
        exec<P>(inst, groups, cycles);
        check(hipGetLastError(), "exec launch");
        check(hipDeviceSynchronize(), "exec");
        launches++;
        const bool done = inst.complete();  // raises on a runtime error first
        if (audit) {
            hz::AsyncCensus c = inst.audit(barriers);
            audits++;
            if (done && (c.queued_links != 0 || c.held_links != 0)) {
                throw std::runtime_error("AsyncProgram audit: complete with links still queued or held");
            }
        }
        if (done) break;
        // A launch that evaluated no promise signals lost accounting rather
        // than unfinished work: report it instead of looping.
        hz::AsyncStatus s = inst.read_status();
        uint64_t progress = s.promises;
        if (progress == last_progress) {
            throw std::runtime_error("AsyncProgram: a launch made no progress before completion");
        }
        last_progress = progress;
    } while (true);
    double ms = t.stop();
    if (stats) {
        hz::AsyncStatus s = inst.read_status();
        stats->promises = s.promises;
        stats->launches = launches;
        stats->spills = s.spills;
        stats->fills = s.fills;
        stats->make_work_calls = s.make_work_calls;
        stats->groups = (uint32_t)(blocks_for<P>(groups) * (P::WORK_GROUP_SIZE / 64));
        stats->elapsed_ms = ms;
        for (int i = 0; i < 8; ++i) stats->phase_cycles[i] = s.phase_cycles[i];
        stats->links_run = s.links_run;
        stats->forwarded = s.forwarded;
        stats->audits = audits;
        stats->arena_links = s.claim < inst.arena_size ? s.claim : inst.arena_size;
    }
}

// Runs an EventProgram instance generation by generation until complete().
template <typename P>
void run_event(typename P::Instance& inst, uint32_t groups, uint32_t chunk, uint32_t max_launches,
               hz_run_stats* stats) {
    EventTimer t;
    t.start();
    uint64_t launches = 0;
    if (max_launches == 0) max_launches = 1u << 24;
    do {
        if (max_launches && launches >= max_launches) {
            throw std::runtime_error("EventProgram: launch bound reached before completion");
        }
        exec<P>(inst, groups, chunk);
        check(hipGetLastError(), "exec launch");
        launches++;
    } while (!inst.complete());
    double ms = t.stop();
    if (stats) {
        auto s = inst.read_status();
        stats->promises = s.promises;
        stats->launches = launches;
        stats->spills = 0;
        stats->fills = 0;
        stats->make_work_calls = s.make_work_calls;
        stats->groups = groups;
        stats->elapsed_ms = ms;
    }
}

}  // namespace capi
}  // namespace hz

// Generates the reference-shaped per-program C entry points
// (harmonize/python/templates.py:3-184) for program type PROG whose device
// state is a plain struct STATE. The instance keeps a by-value copy of the
// state and refreshes it from the device copy before every launch, so the
// device copy behaves like the reference's state pointer.
#define HZ_DEFINE_PROGRAM_ABI(suffix, PROG, STATE)                                                      \
    namespace {                                                                                         \
    struct suffix##_handle {                                                                            \
        PROG::Instance inst;                                                                            \
        void* dev_state;                                                                                \
        suffix##_handle(STATE s, size_t n, void* d) : inst(n, s), dev_state(d) {}                       \
        void refresh() {                                                                                \
            hz::capi::check(hipMemcpy(&inst.device_state, dev_state, sizeof(STATE), hipMemcpyDeviceToHost), \
                            "state refresh");                                                           \
        }                                                                                               \
    };                                                                                                  \
    }                                                                                                   \
    extern "C" void* alloc_state_##suffix(void) {                                                       \
        void* p = nullptr;                                                                              \
        hz::capi::guarded([&] { hz::capi::check(hipMalloc(&p, sizeof(STATE)), "hipMalloc(state)"); });   \
        return p;                                                                                       \
    }                                                                                                   \
    extern "C" void free_state_##suffix(void* state) {                                                  \
        hz::capi::guarded([&] { hz::capi::check(hipFree(state), "hipFree(state)"); });                  \
    }                                                                                                   \
    extern "C" void load_state_device_##suffix(void* host_ptr, void* dev_ptr) {                         \
        hz::capi::guarded([&] {                                                                         \
            hz::capi::check(hipMemcpy(host_ptr, dev_ptr, sizeof(STATE), hipMemcpyDeviceToHost), "load_state"); \
        });                                                                                             \
    }                                                                                                   \
    extern "C" void store_state_device_##suffix(void* dev_ptr, void* host_ptr) {                        \
        hz::capi::guarded([&] {                                                                         \
            hz::capi::check(hipMemcpy(dev_ptr, host_ptr, sizeof(STATE), hipMemcpyHostToDevice), "store_state"); \
        });                                                                                             \
    }                                                                                                   \
    extern "C" void* alloc_program_##suffix(void* device_arg, size_t size) {                            \
        suffix##_handle* h = nullptr;                                                                   \
        hz::capi::guarded([&] {                                                                         \
            if (device_arg == nullptr) throw std::invalid_argument("alloc_program: null device state");  \
            STATE s;                                                                                    \
            hz::capi::check(hipMemcpy(&s, device_arg, sizeof(STATE), hipMemcpyDeviceToHost), "state read"); \
            h = new suffix##_handle(s, size, device_arg);                                               \
        });                                                                                             \
        return h;                                                                                       \
    }                                                                                                   \
    extern "C" void free_program_##suffix(void* instance_ptr) {                                         \
        hz::capi::guarded([&] { delete (suffix##_handle*)instance_ptr; });                              \
    }                                                                                                   \
    extern "C" void init_program_##suffix(void* instance_ptr, size_t grid_size) {                       \
        hz::capi::guarded([&] {                                                                         \
            auto h = (suffix##_handle*)instance_ptr;                                                    \
            init<PROG>(h->inst, grid_size);                                                             \
            hz::capi::check(hipDeviceSynchronize(), "init_program");                                    \
        });                                                                                             \
    }                                                                                                   \
    extern "C" void exec_program_##suffix(void* instance_ptr, size_t grid_size, size_t cycle_count) {   \
        hz::capi::guarded([&] {                                                                         \
            auto h = (suffix##_handle*)instance_ptr;                                                    \
            h->refresh();                                                                               \
            exec<PROG>(h->inst, grid_size, cycle_count);                                                \
            hz::capi::check(hipGetLastError(), "exec_program launch");                                  \
            hz::capi::check(hipDeviceSynchronize(), "exec_program");                                    \
        });                                                                                             \
    }                                                                                                   \
    extern "C" void init_program_on_##suffix(void* instance_ptr, size_t grid_size, void* stream) {       \
        hz::capi::guarded([&] {                                                                         \
            auto h = (suffix##_handle*)instance_ptr;                                                    \
            h->refresh();                                                                               \
            init<PROG>(h->inst, grid_size, (hipStream_t)stream);                                        \
            hz::capi::check(hipGetLastError(), "init_program launch");                                  \
        });                                                                                             \
    }                                                                                                   \
    extern "C" void exec_program_on_##suffix(void* instance_ptr, size_t grid_size, size_t cycle_count,  \
                                             void* stream) {                                            \
        hz::capi::guarded([&] {                                                                         \
            auto h = (suffix##_handle*)instance_ptr;                                                    \
            exec<PROG>(h->inst, grid_size, cycle_count, (hipStream_t)stream);                           \
            hz::capi::check(hipGetLastError(), "exec_program launch");                                  \
        });                                                                                             \
    }                                                                                                   \
    extern "C" int complete_##suffix(void* instance_ptr) {                                              \
        int result = -1;                                                                                \
        hz::capi::guarded([&] { result = ((suffix##_handle*)instance_ptr)->inst.complete() ? 1 : 0; });  \
        return result;                                                                                  \
    }                                                                                                   \
    extern "C" void clear_flags_##suffix(void* instance_ptr) {                                          \
        hz::capi::guarded([&] { ((suffix##_handle*)instance_ptr)->inst.clear_flags(); });               \
    }                                                                                                   \
    extern "C" void set_device_##suffix(int device) {                                                   \
        hz::capi::guarded([&] { hz::capi::check(hipSetDevice(device), "hipSetDevice"); });              \
    }
