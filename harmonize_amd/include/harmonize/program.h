// Host launch API of the program types (reference program/mod.h:17-80):
// init<P>(instance, groups) and exec<P>(instance, groups, n), where n is the
// cycle budget of an AsyncProgram launch or the chunk size of an EventProgram
// generation. Both take an optional stream; the reference launches on the
// null stream and so do the defaults here.
#pragma once

#include "async.h"
#include "event.h"

template <typename ProgType>
__global__ void __launch_bounds__(ProgType::WORK_GROUP_SIZE)
    _hz_exec_kernel(typename ProgType::DeviceContext ctx, typename ProgType::DeviceState state, unsigned int n) {
    ProgType::kernel_exec(ctx, state, n);
}

template <typename ProgType>
__global__ void __launch_bounds__(256) _hz_async_reset_kernel(typename ProgType::DeviceContext ctx) {
    ProgType::kernel_reset(ctx);
}

template <typename ProgType>
__global__ void __launch_bounds__(64) _hz_event_flip_kernel(typename ProgType::DeviceContext ctx) {
    ProgType::flip(ctx);
}

template <typename ProgType>
__global__ void __launch_bounds__(256) _hz_event_reset_kernel(typename ProgType::DeviceContext ctx) {
    ProgType::kernel_reset(ctx);
}

template <typename PROGRAM_SPEC>
void EventProgram<PROGRAM_SPEC>::Instance::reset(hipStream_t stream) {
    hipLaunchKernelGGL(_hz_event_reset_kernel<EventProgram<PROGRAM_SPEC>>, dim3(16), dim3(256), 0, stream,
                       to_context());
    util::host::auto_throw(hipGetLastError());
}

template <typename PROGRAM_SPEC>
void AsyncProgram<PROGRAM_SPEC>::Instance::reset(hipStream_t stream) {
    size_t slots = (size_t)queue_capacity * ((size_t)queue_count * AsyncProgram<PROGRAM_SPEC>::LEVELS + 1u);
    size_t blocks = (slots + 255) / 256;
    if (blocks > 1024) blocks = 1024;
    if (blocks == 0) blocks = 1;
    hipLaunchKernelGGL(_hz_async_reset_kernel<AsyncProgram<PROGRAM_SPEC>>, dim3((unsigned)blocks), dim3(256), 0,
                       stream, to_context());
    util::host::auto_throw(hipGetLastError());
}

namespace hz {
template <typename P>
struct IsAsync : std::false_type {};
template <typename S>
struct IsAsync<AsyncProgram<S>> : std::true_type {};
}  // namespace hz

// Prepares an instance for execution. For an AsyncProgram this empties its
// queues and clears its status (reference AsyncProgram::init_program,
// async.h:2405-2516); for an EventProgram it does nothing, as in the reference.
template <typename ProgType>
__host__ void init(typename ProgType::Instance& instance, size_t group_count, hipStream_t stream = nullptr) {
    (void)group_count;
    if constexpr (hz::IsAsync<ProgType>::value) {
        instance.reset(stream);
    }
}

// Workgroups needed for `group_count` groups (rounded up to whole workgroups).
template <typename ProgType>
__host__ size_t blocks_for(size_t group_count) {
    const size_t per = ProgType::WORK_GROUP_SIZE / 64;
    return (group_count + per - 1) / per;
}

// Launches one execution pass of `group_count` groups (64-lane wavefronts),
// rounded up to whole workgroups of ProgType::WORK_GROUP_SIZE threads.
template <typename ProgType>
__host__ void exec(typename ProgType::Instance& instance, size_t group_count, size_t n, hipStream_t stream = nullptr) {
    if (group_count == 0) group_count = 1;
    auto ctx = instance.to_context();
    hipLaunchKernelGGL(_hz_exec_kernel<ProgType>, dim3((unsigned)blocks_for<ProgType>(group_count)),
                       dim3((unsigned)ProgType::WORK_GROUP_SIZE), 0, stream, ctx, instance.device_state,
                       (unsigned int)(n > 0xFFFFFFFFull ? 0xFFFFFFFFull : n));
    if constexpr (!hz::IsAsync<ProgType>::value) {
        hipLaunchKernelGGL(_hz_event_flip_kernel<ProgType>, dim3(1), dim3(64), 0, stream, ctx);
    }
}
