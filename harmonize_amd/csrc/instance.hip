// Generic instance entry points of the C ABI: one handle type for every
// built-in specification and program kind, dispatching to the per-program
// entry points (alloc_program_<spec>_<kind> and friends) through a table, so
// that a caller can select the specification at run time.
#include <string>

#include "common.h"

using hz::capi::check;
using hz::capi::guarded;

namespace {

struct ProgramEntry {
    const char* name;
    size_t state_bytes;
    void* (*alloc_state)();
    void (*free_state)(void*);
    void* (*alloc_program)(void*, size_t);
    void (*free_program)(void*);
    void (*init_program)(void*, size_t);
    void (*exec_program)(void*, size_t, size_t);
    int (*complete)(void*);
    void (*clear_flags)(void*);
    void (*init_on)(void*, size_t, void*);
    void (*exec_on)(void*, size_t, size_t, void*);
};

#define HZ_ENTRY(suffix, STATE)                                                                           \
    ProgramEntry {                                                                                        \
        #suffix, sizeof(STATE), alloc_state_##suffix, free_state_##suffix, alloc_program_##suffix,        \
            free_program_##suffix, init_program_##suffix, exec_program_##suffix, complete_##suffix,        \
            clear_flags_##suffix, init_program_on_##suffix, exec_program_on_##suffix                      \
    }

// [spec][kind], kind = HZ_EVENT or HZ_ASYNC; a null name marks a pair that does not exist.
const ProgramEntry* entry_for(int spec, int kind) {
    static const ProgramEntry table[4][2] = {
        {HZ_ENTRY(collatz_event, hz_collatz_state), HZ_ENTRY(collatz_async, hz_collatz_state)},
        {HZ_ENTRY(photon_event, hz_photon_state), HZ_ENTRY(photon_async, hz_photon_state)},
        {HZ_ENTRY(bfs_event, hz_bfs_state), HZ_ENTRY(bfs_async, hz_bfs_state)},
        {ProgramEntry{nullptr, 0, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
                      nullptr},
         HZ_ENTRY(barrier_async, hz_barrier_state)},
    };
    if (spec < 0 || spec > 3 || (kind != HZ_EVENT && kind != HZ_ASYNC)) return nullptr;
    const ProgramEntry* e = &table[spec][kind];
    return e->name ? e : nullptr;
}

#undef HZ_ENTRY

// The per-program entry points report failures through hz_last_error().
void rethrow_last(const char* what) {
    const char* msg = hz_last_error();
    if (msg) throw std::runtime_error(std::string(what) + ": " + msg);
}

}  // namespace

struct hz_instance {
    const ProgramEntry* entry = nullptr;
    void* dev_state = nullptr;
    void* program = nullptr;
};

extern "C" int hz_instance_create(int kind, int spec, uint64_t size, const void* device_state, size_t state_bytes,
                                  hz_instance** out) {
    return guarded([&] {
        if (!out) throw std::invalid_argument("hz_instance_create: null output");
        *out = nullptr;
        const ProgramEntry* e = entry_for(spec, kind);
        if (!e) throw std::invalid_argument("hz_instance_create: no such specification and kind");
        if (!device_state || state_bytes != e->state_bytes) {
            throw std::invalid_argument("hz_instance_create: state must be the specification's state struct (" +
                                        std::to_string(e->state_bytes) + " bytes)");
        }
        if (size == 0) throw std::invalid_argument("hz_instance_create: size must be positive");
        if (spec == HZ_SPEC_BFS) {
            const hz_bfs_state* bs = static_cast<const hz_bfs_state*>(device_state);
            if (bs->split < 1u || bs->split > HZ_BFS_MAX_SPLIT) {
                throw std::invalid_argument("hz_instance_create: hz_bfs_state.split must be in [1, 2^24]");
            }
        }
        auto inst = new hz_instance();
        inst->entry = e;
        try {
            inst->dev_state = e->alloc_state();
            rethrow_last("alloc_state");
            if (!inst->dev_state) throw std::runtime_error("alloc_state failed");
            check(hipMemcpy(inst->dev_state, device_state, state_bytes, hipMemcpyHostToDevice), "state upload");
            inst->program = e->alloc_program(inst->dev_state, (size_t)size);
            rethrow_last("alloc_program");
            if (!inst->program) throw std::runtime_error("alloc_program failed");
        } catch (...) {
            if (inst->program) e->free_program(inst->program);
            if (inst->dev_state) e->free_state(inst->dev_state);
            delete inst;
            throw;
        }
        *out = inst;
    });
}

extern "C" int hz_instance_destroy(hz_instance* inst) {
    return guarded([&] {
        if (!inst) return;
        if (inst->program) inst->entry->free_program(inst->program);
        if (inst->dev_state) inst->entry->free_state(inst->dev_state);
        delete inst;
    });
}

namespace {
hz_instance* need(hz_instance* inst) {
    if (!inst) throw std::invalid_argument("hz_instance: null handle");
    return inst;
}
}  // namespace

extern "C" int hz_init(hz_instance* inst, uint32_t groups) {
    return guarded([&] {
        need(inst)->entry->init_program(inst->program, groups);
        rethrow_last("init_program");
    });
}

extern "C" int hz_exec(hz_instance* inst, uint32_t groups, uint32_t cycles) {
    return guarded([&] {
        if (groups == 0) throw std::invalid_argument("hz_exec: groups must be positive");
        need(inst)->entry->exec_program(inst->program, groups, cycles);
        rethrow_last("exec_program");
    });
}

extern "C" int hz_complete(hz_instance* inst, int* done) {
    return guarded([&] {
        if (!done) throw std::invalid_argument("hz_complete: null output");
        int r = need(inst)->entry->complete(inst->program);
        rethrow_last("complete");
        *done = r;
    });
}

extern "C" int hz_init_on(hz_instance* inst, uint32_t groups, void* stream) {
    return guarded([&] {
        need(inst)->entry->init_on(inst->program, groups, stream);
        rethrow_last("init_program_on");
    });
}

extern "C" int hz_exec_on(hz_instance* inst, uint32_t groups, uint32_t cycles, void* stream) {
    return guarded([&] {
        if (groups == 0) throw std::invalid_argument("hz_exec_on: groups must be positive");
        need(inst)->entry->exec_on(inst->program, groups, cycles, stream);
        rethrow_last("exec_program_on");
    });
}

extern "C" int hz_complete_on(hz_instance* inst, void* stream, int* done) {
    return guarded([&] {
        if (!done) throw std::invalid_argument("hz_complete_on: null output");
        need(inst);
        check(hipStreamSynchronize((hipStream_t)stream), "hipStreamSynchronize");
        int r = inst->entry->complete(inst->program);
        rethrow_last("complete");
        *done = r;
    });
}

extern "C" int hz_stream_create(void** stream) {
    return guarded([&] {
        if (!stream) throw std::invalid_argument("hz_stream_create: null output");
        hipStream_t s = nullptr;
        check(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreateWithFlags");
        *stream = s;
    });
}

extern "C" int hz_stream_destroy(void* stream) {
    return guarded([&] {
        if (stream) check(hipStreamDestroy((hipStream_t)stream), "hipStreamDestroy");
    });
}

extern "C" int hz_clear_flags(hz_instance* inst) {
    return guarded([&] {
        need(inst)->entry->clear_flags(inst->program);
        rethrow_last("clear_flags");
    });
}

extern "C" int hz_state_load(hz_instance* inst, uint64_t offset, uint64_t bytes, void* host) {
    return guarded([&] {
        need(inst);
        if (offset + bytes > inst->entry->state_bytes || (bytes && !host)) {
            throw std::invalid_argument("hz_state_load: range outside the state");
        }
        if (bytes) check(hipMemcpy(host, (char*)inst->dev_state + offset, bytes, hipMemcpyDeviceToHost), "state load");
    });
}

extern "C" int hz_state_store(hz_instance* inst, uint64_t offset, uint64_t bytes, const void* host) {
    return guarded([&] {
        need(inst);
        if (offset + bytes > inst->entry->state_bytes || (bytes && !host)) {
            throw std::invalid_argument("hz_state_store: range outside the state");
        }
        if (bytes) check(hipMemcpy((char*)inst->dev_state + offset, host, bytes, hipMemcpyHostToDevice), "state store");
    });
}

extern "C" int hz_set_device(int device) {
    return guarded([&] { check(hipSetDevice(device), "hipSetDevice"); });
}
