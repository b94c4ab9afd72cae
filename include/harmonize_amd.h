/*
 * C-ABI boundary of libharmonize_amd.so, the gfx950 asynchronous execution
 * runtime. Plain C: opaque handles, plain pointers and sizes, no exceptions
 * across the boundary. Device pointers are HIP device allocations on the
 * current device (harmonize_alloc_device_bytes or any hipMalloc).
 *
 * Two layers:
 *  1. The per-program entry points that the reference generates for its
 *     Python binding (harmonize/python/templates.py), with the same names,
 *     argument order and meaning, for every built-in specification
 *     <spec>_<kind> with <spec> in {collatz, photon, bfs} and <kind> in
 *     {event, async}. These keep the reference's `void` returns; failures
 *     are reported through hz_last_error() instead of C++ exceptions
 *     crossing the boundary as in the reference.
 *  2. hz_* workload drivers that run a whole program to completion and
 *     report statistics. They return 0 on success and a negative status on
 *     failure (message in hz_last_error()).
 */
#ifndef HARMONIZE_AMD_H
#define HARMONIZE_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- status */

#define HZ_OK 0
#define HZ_ERR_ARGUMENT (-1)
#define HZ_ERR_RUNTIME (-2)   /* HIP error, arena/queue exhaustion, overflow */
#define HZ_ERR_UNSUPPORTED (-3)

/* Message of the last failure on this thread, or NULL if the last call succeeded. */
const char* hz_last_error(void);

/* Program kinds. */
#define HZ_EVENT 0 /* EventProgram: one launch per generation */
#define HZ_ASYNC 1 /* AsyncProgram: persistent coalescing scheduler */
#define HZ_PLAIN 2 /* one thread per work item, no runtime (Collatz only) */

/* --------------------------------------- device memory helpers (layer 1)
 * Replace templates.py:307-367 (harmonize_alloc_device_bytes,
 * harmonize_alloc_managed_bytes, harmonize_free_device_bytes,
 * harmonize_memcpy_host_to_device, harmonize_memcpy_device_to_host; the
 * reference's device-to-host copy passes the host-to-device kind,
 * templates.py:365, which is not repeated here). NULL on failure. */
void* harmonize_alloc_device_bytes(size_t size);
void* harmonize_alloc_managed_bytes(size_t size);
void harmonize_free_device_bytes(void* ptr);
void harmonize_memcpy_host_to_device(void* dev_ptr, void* host_ptr, size_t byte_count);
void harmonize_memcpy_device_to_host(void* host_ptr, void* dev_ptr, size_t byte_count);

/* ------------------------------------------ device states of the built-ins */

/* Collatz step counts. make_work claims indices of `starts` (or uses the
 * index itself as the start when starts is NULL) from `iterator`, a device
 * pointer to a {uint32 value, uint32 limit} pair. steps[i] receives the
 * number of Collatz steps of start i, or HZ_COLLATZ_OVERFLOW if the
 * trajectory leaves 64 bits. */
#define HZ_COLLATZ_OVERFLOW 0xFFFFFFFFu
typedef struct hz_collatz_state {
    const uint64_t* starts;
    uint32_t* steps;
    void* iterator;
} hz_collatz_state;

/* Branching photon walk in a slab (FP32 state, FP64 tallies). The tally
 * holds `replicas` copies of (bins + HZ_PHOTON_EXTRA_SLOTS) doubles: depth
 * bins of deposited weight, then reflected, transmitted, roulette-lost,
 * roulette-gained and capped weight. events[photon] = events of its family. */
#define HZ_PHOTON_EXTRA_SLOTS 5
typedef struct hz_photon_params {
    uint32_t photons;
    uint32_t bins;
    uint64_t seed;
    float thickness;
    float sigma_t;
    float albedo;
    float split_above;    /* split when w / target > split_above */
    float roulette_below; /* roulette when w / target < roulette_below */
    float region_width;   /* weight-window region depth; target = 2^-region */
    uint32_t regions;
    uint32_t max_split;   /* power of two */
    uint32_t max_events;  /* per lineage, <= 65535 */
    uint32_t replicas;    /* tally replicas, power of two */
} hz_photon_params;

/* Device state of the photon programs (layout of PhotonState). */
typedef struct hz_photon_state {
    uint32_t photon_count, seed_lo, seed_hi, bins, replicas, max_split, max_events, regions;
    float thickness, mean_free_path, albedo, bin_scale, region_scale, split_above, roulette_below, pad;
    double* tally;
    uint32_t* events;
    void* iterator;
} hz_photon_state;

/* Remapping-barrier program (after the reference's barrier demo,
 * examples/cpp/barrier_demo/demo.cpp:26-44): Step(id) for ids
 * [0, dep + non_dep); a dependent id awaits barrier[id % barriers] with a
 * held Step and decrements that barrier's semaphore. Layout of
 * BarrierDemoState. */
/* Batched BFS. A job is one int32 CSR graph (row_ptr[vertices + 1] of local
 * edge offsets, col[edges] of local vertex ids) with an int32 level array;
 * unreached vertices keep HZ_BFS_UNREACHED. */
#define HZ_BFS_UNREACHED 0x7FFFFFFF
#define HZ_BFS_MAX_SPLIT (1u << 24) /* largest edge-range split threshold */
typedef struct hz_bfs_job {
    const int32_t* row_ptr;
    const int32_t* col;
    int32_t* level;
    uint32_t vertices;
    uint32_t source;
} hz_bfs_job;
typedef struct hz_bfs_state {
    const hz_bfs_job* jobs;
    void* job_iter; /* {uint32 next, uint32 limit}: jobs still to start */
    uint32_t job_count;
    uint32_t split; /* edge ranges longer than this are cut into up to 16
                       pieces of at least `split` edges; valid range
                       [1, 2^24] (hz_instance_create rejects others, the
                       device clamps to it) */
} hz_bfs_state;

typedef struct hz_barrier_state {
    uint32_t dep_count, non_dep_count, barrier_count, pad;
    void* source_id_iter;
    uint32_t* counter;
    uint32_t* hits;
    uint32_t* violations;
    void* barrier;
} hz_barrier_state;

/* ------------------------------------------- per-program entry points
 * For each <suffix> in collatz_event, collatz_async, photon_event,
 * photon_async, bfs_event, bfs_async, barrier_async (the barrier needs
 * AsyncProgram):
 *
 *   void* alloc_state_<suffix>(void);
 *       Device allocation sized for the program's device state
 *       (templates.py:73-83).
 *   void free_state_<suffix>(void* state);                    (:154-159)
 *   void load_state_device_<suffix>(void* host, void* dev);   (:85-104)
 *   void store_state_device_<suffix>(void* dev, void* host);  (:106-125)
 *       Copy the whole device state between host and device.
 *   void* alloc_program_<suffix>(void* dev_state, size_t size);
 *       New instance bound to a device state from alloc_state_<suffix>.
 *       size = io capacity per operation (event, :43-52) or arena links of
 *       64 promises (async, :54-63). NULL on failure.
 *   void free_program_<suffix>(void* instance);               (:65-71)
 *   void init_program_<suffix>(void* instance, size_t grid);  (:3-14)
 *   void exec_program_<suffix>(void* instance, size_t grid, size_t cycles);
 *       One launch of `grid` groups, then a device synchronize (:17-40).
 *       The device state is re-read from its device copy at every launch,
 *       so stores between launches take effect as in the reference.
 *   int complete_<suffix>(void* instance);                    (:161-168)
 *       1 when finished, 0 when not, -1 on error.
 *   void clear_flags_<suffix>(void* instance);                (:170-177)
 *   void set_device_<suffix>(int device);                     (:179-184)
 *
 * Additions with no reference counterpart (the reference launches on the
 * null stream and synchronizes after every launch):
 *   void init_program_on_<suffix>(void* instance, size_t grid, void* stream);
 *   void exec_program_on_<suffix>(void* instance, size_t grid, size_t cycles,
 *                                 void* stream);
 *       Enqueue the same work on a hipStream_t and return without waiting,
 *       so that instances on different streams overlap. The device state is
 *       read at init_program_on; the caller synchronizes the stream before
 *       complete_<suffix>.
 */
#define HZ_DECLARE_PROGRAM_ABI(suffix)                                              \
    void* alloc_state_##suffix(void);                                               \
    void free_state_##suffix(void* state);                                          \
    void load_state_device_##suffix(void* host_ptr, void* dev_ptr);                 \
    void store_state_device_##suffix(void* dev_ptr, void* host_ptr);                \
    void* alloc_program_##suffix(void* device_arg, size_t size);                    \
    void free_program_##suffix(void* instance_ptr);                                 \
    void init_program_##suffix(void* instance_ptr, size_t grid_size);               \
    void exec_program_##suffix(void* instance_ptr, size_t grid_size, size_t cycle_count); \
    int complete_##suffix(void* instance_ptr);                                      \
    void clear_flags_##suffix(void* instance_ptr);                                  \
    void set_device_##suffix(int device);                                           \
    void init_program_on_##suffix(void* instance_ptr, size_t grid_size, void* stream); \
    void exec_program_on_##suffix(void* instance_ptr, size_t grid_size, size_t cycle_count, void* stream);

HZ_DECLARE_PROGRAM_ABI(collatz_event)
HZ_DECLARE_PROGRAM_ABI(collatz_async)
HZ_DECLARE_PROGRAM_ABI(photon_event)
HZ_DECLARE_PROGRAM_ABI(photon_async)
HZ_DECLARE_PROGRAM_ABI(bfs_event)
HZ_DECLARE_PROGRAM_ABI(bfs_async)
HZ_DECLARE_PROGRAM_ABI(barrier_async)

/* ------------------------------- per-label state accessors (reference names)
 * The reference's binding (harmonize/python/runtime.py:1348-1371) binds, for
 * every state label of a program, load_state_<label>_<suffix>,
 * store_state_<label>_<suffix> and store_pointer_state_<label>_<suffix>
 * (templates.py:85-152). Labels are "device", "group" and "thread", and
 * "device_<field>" for a device state given as fields (runtime.py:562-595).
 * Besides the whole-state device accessors above, this library exports:
 *
 *   void load_state_device_<field>_<suffix>(void* host, void* dev_state);
 *   void store_state_device_<field>_<suffix>(void* dev_state, void* host);
 *       Copy one field (its size, at its offset) of the device state.
 *   void store_pointer_state_device_<field>_<suffix>(void* dev_state, void* ptr);
 *       Store the pointer value `ptr` into a pointer field; for a non-pointer
 *       field it records the reference's error, "Attempted to set
 *       indirection of non-array state field." (templates.py:146-150).
 *   void store_pointer_state_device_<suffix>(void* dev_state, void* ptr);
 *       The whole state is a record, not an array field: records that error.
 *   void load_state_group_<suffix>, store_state_group_<suffix>,
 *        load_state_thread_<suffix>, store_state_thread_<suffix>
 *       (same argument order as the device ones): the built-in programs keep
 *       no host-visible group or thread state, so these copy zero bytes;
 *       store_pointer_state_group_<suffix> and _thread_<suffix> record the
 *       error above.
 *
 * Fields (1 = pointer field): */
#define HZ_COLLATZ_STATE_FIELDS(X, suffix) X(suffix, starts, 1) X(suffix, steps, 1) X(suffix, iterator, 1)
#define HZ_PHOTON_STATE_FIELDS(X, suffix)                                                        \
    X(suffix, photon_count, 0) X(suffix, seed_lo, 0) X(suffix, seed_hi, 0) X(suffix, bins, 0)    \
    X(suffix, replicas, 0) X(suffix, max_split, 0) X(suffix, max_events, 0) X(suffix, regions, 0) \
    X(suffix, thickness, 0) X(suffix, mean_free_path, 0) X(suffix, albedo, 0) X(suffix, bin_scale, 0) \
    X(suffix, region_scale, 0) X(suffix, split_above, 0) X(suffix, roulette_below, 0)            \
    X(suffix, tally, 1) X(suffix, events, 1) X(suffix, iterator, 1)
#define HZ_BFS_STATE_FIELDS(X, suffix) X(suffix, jobs, 1) X(suffix, job_iter, 1) X(suffix, job_count, 0) X(suffix, split, 0)
#define HZ_BARRIER_STATE_FIELDS(X, suffix)                                                       \
    X(suffix, dep_count, 0) X(suffix, non_dep_count, 0) X(suffix, barrier_count, 0)             \
    X(suffix, source_id_iter, 1) X(suffix, counter, 1) X(suffix, hits, 1) X(suffix, violations, 1) X(suffix, barrier, 1)

#define HZ_DECLARE_STATE_FIELD(suffix, field, is_pointer)                           \
    void load_state_device_##field##_##suffix(void* host_ptr, void* dev_ptr);       \
    void store_state_device_##field##_##suffix(void* dev_ptr, void* host_ptr);      \
    void store_pointer_state_device_##field##_##suffix(void* dev_ptr, void* host_ptr);
#define HZ_DECLARE_STATE_LABELS(suffix, FIELDS)                                     \
    FIELDS(HZ_DECLARE_STATE_FIELD, suffix)                                          \
    void store_pointer_state_device_##suffix(void* dev_ptr, void* host_ptr);        \
    void load_state_group_##suffix(void* host_ptr, void* dev_ptr);                  \
    void store_state_group_##suffix(void* dev_ptr, void* host_ptr);                 \
    void store_pointer_state_group_##suffix(void* dev_ptr, void* host_ptr);         \
    void load_state_thread_##suffix(void* host_ptr, void* dev_ptr);                 \
    void store_state_thread_##suffix(void* dev_ptr, void* host_ptr);                \
    void store_pointer_state_thread_##suffix(void* dev_ptr, void* host_ptr);

HZ_DECLARE_STATE_LABELS(collatz_event, HZ_COLLATZ_STATE_FIELDS)
HZ_DECLARE_STATE_LABELS(collatz_async, HZ_COLLATZ_STATE_FIELDS)
HZ_DECLARE_STATE_LABELS(photon_event, HZ_PHOTON_STATE_FIELDS)
HZ_DECLARE_STATE_LABELS(photon_async, HZ_PHOTON_STATE_FIELDS)
HZ_DECLARE_STATE_LABELS(bfs_event, HZ_BFS_STATE_FIELDS)
HZ_DECLARE_STATE_LABELS(bfs_async, HZ_BFS_STATE_FIELDS)
HZ_DECLARE_STATE_LABELS(barrier_async, HZ_BARRIER_STATE_FIELDS)

/* The device's wall-clock rate in Hz (templates.py:268-285), for converting
 * wall-clock readings taken inside specifications; 0 on error. */
long long wall_clock_rate(void);

/* -------------------------------------------- generic instances (layer 1)
 * One handle for any built-in specification and kind, chosen at run time
 * (the boundary the per-program names above expand to):
 *   spec = HZ_SPEC_COLLATZ, HZ_SPEC_PHOTON, HZ_SPEC_BFS or HZ_SPEC_BARRIER
 *   (the barrier program exists only as HZ_ASYNC); kind = HZ_EVENT or HZ_ASYNC;
 *   size = io capacity per operation (event) or arena links (async);
 *   device_state = host copy of the spec's state struct (hz_collatz_state,
 *   hz_photon_state, hz_bfs_state, hz_barrier_state), state_bytes its size.
 * hz_exec launches once and synchronizes (the reference's convention);
 * hz_complete sets *done to 1 when the program has finished. The state lives
 * on the device; hz_state_load/store copy a byte range of it and take effect
 * from the next launch.
 * hz_init_on / hz_exec_on enqueue on `stream` (a hipStream_t; NULL = the
 * null stream) and return at once; hz_complete_on synchronizes that stream
 * and then reports like hz_complete. hz_stream_create/destroy make and free
 * non-blocking streams for callers without HIP headers. */
#define HZ_SPEC_COLLATZ 0
#define HZ_SPEC_PHOTON 1
#define HZ_SPEC_BFS 2
#define HZ_SPEC_BARRIER 3
typedef struct hz_instance hz_instance;
int hz_instance_create(int kind, int spec, uint64_t size, const void* device_state, size_t state_bytes,
                       hz_instance** out);
int hz_instance_destroy(hz_instance* instance);
int hz_init(hz_instance* instance, uint32_t groups);
int hz_exec(hz_instance* instance, uint32_t groups, uint32_t cycles);
int hz_complete(hz_instance* instance, int* done);
int hz_clear_flags(hz_instance* instance);
int hz_init_on(hz_instance* instance, uint32_t groups, void* stream);
int hz_exec_on(hz_instance* instance, uint32_t groups, uint32_t cycles, void* stream);
int hz_complete_on(hz_instance* instance, void* stream, int* done);
int hz_stream_create(void** stream);
int hz_stream_destroy(void* stream);
int hz_state_load(hz_instance* instance, uint64_t offset, uint64_t bytes, void* host);
int hz_state_store(hz_instance* instance, uint64_t offset, uint64_t bytes, const void* host);
int hz_set_device(int device);

/* ------------------------------------------------------ drivers (layer 2) */

typedef struct hz_run_config {
    uint32_t groups;   /* wavefront groups per launch; 0 = fill the device */
    uint32_t cycles;   /* async: cycle budget per launch (0 = 2^24);
                          event: chunk size in wavefronts of promises (0 = 4) */
    uint64_t arena;    /* async: arena links of 64 promises; event: queue
                          capacity per operation, in promises (0 = default) */
    uint32_t queues;   /* async: number of link queues (0 = 8) */
    uint32_t coalesce; /* async: minimum promises in an open link before it
                          runs while make_work can still add (0 = 64) */
    uint32_t max_launches; /* bound on the host launch loop (0 = none) */
    uint32_t block_groups; /* async: wavefront groups per workgroup where the
                              program offers variants (BFS: 1, 2 or 4; 0 = 1) */
    uint32_t flags;        /* HZ_RUN_GRAPH: launch the pass through a captured
                              hipGraph (BFS batches, HZ_ASYNC) */
    uint32_t spec_param;   /* program-specific knob (BFS: longest edge range a
                              promise relaxes before splitting it, at most
                              2^24; 0 = 64) */
    uint32_t stash;        /* async: LDS links per group where the program
                              offers variants (BFS: 8 or 16; 0 = 16) */
} hz_run_config;

#define HZ_RUN_GRAPH 1u
/* Audit the AsyncProgram's queues and arena on the host after every launch
 * (a debugging aid, the role of the reference's runtime_overview,
 * program/async.h:2766-2883): every group has left, ring counters agree with
 * their slots, no link id is queued twice or both queued and free, the units
 * count equals the queued links, and every promise written into a link is
 * evaluated, queued or held by a barrier. A violation fails the run with
 * HZ_ERR_RUNTIME naming it. */
#define HZ_RUN_AUDIT 2u

typedef struct hz_run_stats {
    uint64_t promises;        /* promises evaluated */
    uint64_t launches;        /* kernel launches of the exec loop */
    uint64_t spills;          /* async: links moved LDS -> HBM */
    uint64_t fills;           /* async: links moved HBM -> LDS */
    uint64_t make_work_calls; /* wavefront calls of make_work */
    uint32_t groups;          /* groups actually launched */
    uint32_t audits;          /* async, HZ_RUN_AUDIT: launches after which
                                 the queues and arena were audited */
    double elapsed_ms;        /* hipEvent time of the exec loop */
    uint64_t phase_cycles[8]; /* async, runtime built with HZ_PHASE_TIMING:
                                 shader cycles summed over groups per phase
                                 (cycle, spill, pull, make_work, pick, read,
                                 eval, idle); zero otherwise */
    uint64_t links_run;       /* async: links evaluated; promises / links_run
                                 is the mean number of busy lanes per
                                 evaluation (64 = perfect coalescing) */
    uint64_t arena_links;     /* async: arena links ever taken from the bump
                                 allocator (the high-water mark of the arena) */
    uint64_t forwarded;       /* async: promises run straight from registers by
                                 link forwarding, bypassing the LDS stash */
} hz_run_stats;

/* Runs the Collatz program over `count` starts (d_starts may be NULL: start
 * = index) and writes d_steps[count]. kind = HZ_EVENT, HZ_ASYNC or HZ_PLAIN. */
int hz_collatz_run(int kind, const uint64_t* d_starts, uint32_t count, uint32_t* d_steps,
                   const hz_run_config* config, hz_run_stats* stats);

/* Depth levels (the reference's non-flat work stack, mem/work.h:133-164,
 * and depth-aware async calls, program/async.h:1164-1181, 1627-1660): a tree
 * of promises where each Node at height d < height queues `fanout` children
 * one depth level deeper and each leaf counts itself for its root, so
 * d_leaves[r] = fanout^height for every root r. `order` picks the program:
 *   HZ_TREE_FLAT        one level (STACK_SIZE 0);
 *   HZ_TREE_DEEPEST     8 levels, deepest first (the reference's order);
 *   HZ_TREE_SHALLOWEST  8 levels, shallowest first.
 * *d_violations counts nodes whose prog.depth() differed from their height
 * clamped to the levels. stats->arena_links shows how many arena links the
 * order needed. kind = HZ_ASYNC, or HZ_EVENT with HZ_TREE_FLAT. */
#define HZ_TREE_FLAT 0
#define HZ_TREE_DEEPEST 1
#define HZ_TREE_SHALLOWEST 2
int hz_tree_run(int kind, int order, uint32_t roots, uint32_t fanout, uint32_t height, uint32_t* d_leaves,
                uint32_t* d_violations, const hz_run_config* config, hz_run_stats* stats);

/* Asynchronous argument loads (the reference's LazyLoad and ASYNC_LOADS,
 * async/promise.h:204-240, 284-303; program/async.h:2082-2144), on linked
 * lists: list l starts at node d_heads[l] (HZ_CHASE_END: empty) and follows
 * d_nodes[i].next until HZ_CHASE_END; d_sums[l] = sum of its nodes' values,
 * d_lengths[l] = its node count. Each step consumes its node as a LazyLoad
 * argument: lazy = 1 declares the parameter LazyLoad<node>, so AsyncProgram
 * loads a link's nodes with direct-to-LDS loads while the link before it
 * runs; lazy = 0 declares a plain node, read when the promise is built.
 * kind = HZ_ASYNC or HZ_EVENT (which reads lazy nodes when it evaluates). */
#define HZ_CHASE_END 0xFFFFFFFFu
typedef struct hz_chase_node {
    uint32_t next;
    uint32_t value;
} hz_chase_node;
int hz_chase_run(int kind, int lazy, uint32_t lists, const uint32_t* d_heads, const hz_chase_node* d_nodes,
                 uint64_t* d_sums, uint32_t* d_lengths, const hz_run_config* config, hz_run_stats* stats);

/* Allocation-free variant of hz_collatz_run(HZ_ASYNC, ...). It has no
 * counterpart in the reference: it is for callers that cannot allocate device
 * memory (code that sets up HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc). It runs the same kernels, but the runtime's own
 * buffers (claim iterator, status, link queues, link arena) live in one fixed
 * device array of at most HZ_NOALLOC_SCRATCH_BYTES per GPU, held until the
 * process ends; the call allocates no device memory. The caller passes the
 * inputs and outputs as for hz_collatz_run. Limits: config->arena at most
 * HZ_NOALLOC_MAX_LINKS links (0 = that maximum), config->queues at most
 * HZ_NOALLOC_MAX_QUEUES (0 = 8); larger values are refused with
 * HZ_ERR_ARGUMENT before the device is touched. Every call on one GPU shares
 * that GPU's array: calls that overlap on one GPU (several host threads)
 * race and must be serialized by the caller. elapsed_ms is host wall time. */
#define HZ_NOALLOC_MAX_LINKS 16384u
#define HZ_NOALLOC_MAX_QUEUES 16u
#define HZ_NOALLOC_SCRATCH_BYTES (18u << 20)
int hz_collatz_run_noalloc(const uint64_t* d_starts, uint32_t count, uint32_t* d_steps,
                           const hz_run_config* config, hz_run_stats* stats);

/* Fills d_out[count] with starts drawn uniformly from [1, 2^bits) by
 * Philox4x32-10 keyed by seed: start i = (((x1 << 32) | x0) mod (2^bits - 1)) + 1
 * where (x0, x1, x2, x3) = Philox4x32-10(counter = {i, 0, 0, 0},
 * key = {seed & 0xFFFFFFFF, seed >> 32}). 1 <= bits <= 63. */
int hz_collatz_make_starts(uint64_t seed, uint32_t bits, uint32_t count, uint64_t* d_out);

/* Runs the photon walk for photons [0, params->photons). d_tally must hold
 * params->replicas * (bins + HZ_PHOTON_EXTRA_SLOTS) doubles and d_events
 * params->photons counts; both are zeroed first. kind = HZ_EVENT or HZ_ASYNC. */
int hz_photon_run(int kind, const hz_photon_params* params, double* d_tally, uint32_t* d_events,
                  const hz_run_config* config, hz_run_stats* stats);

/* Fills a hz_photon_params with the workload's defaults: 1e7 photons,
 * 1000 bins, thickness 20, sigma_t 1, albedo 0.9, split above 2, roulette
 * below 0.1, region width 2 (10 regions), max_split 8, max_events 65535,
 * 256 replicas, seed 0x5EED. */
void hz_photon_defaults(hz_photon_params* params);

/* Diagnostic: evaluates the specifications' deterministic math on the GPU:
 * d_log[i] = det_log(d_in[i]), d_sqrt[i] = det_sqrt(d_in[i]) and
 * d_philox[4i..4i+3] = Philox4x32-10(counter {i, 3i, ~i, 7},
 * key {0x5EED, i ^ 0xABCD}), so tests can compare them with the host bit for bit. */
int hz_detmath_eval(const double* d_in, double* d_log, double* d_sqrt, uint32_t* d_philox, uint32_t n);

/* Runs the barrier program (AsyncProgram only, as in the reference). Writes
 * *d_counter = Step evaluations (expected 2 * dep + non_dep),
 * d_hits[i] = runs of the step held for dependent i (expected 1), and
 * *d_violations = held steps that ran before their barrier's release. */
int hz_barrier_run(uint32_t dep_count, uint32_t non_dep_count, uint32_t barrier_count, uint32_t* d_counter,
                   uint32_t* d_hits, uint32_t* d_violations, const hz_run_config* config, hz_run_stats* stats);

/* Runs the futures program: value-returning operations resolving a binary
 * tree of futures over Collatz step counts (no reference counterpart: the
 * reference declares Future, UnitBarrier, ReturnOp and Return,
 * async/future.h:7-60, sync/barrier.h:648-699, async/promise.h:326-434,
 * but leaves their resolution unfinished).
 *   leaves   power of two; leaf i computes steps(d_starts[i]) through a chain
 *            of promises that returns the count;
 *   *total   the sum of all leaves' counts, resolved through the tree;
 *   readers  (HZ_ASYNC only) reader r awaits the total on its
 *            RemappingBarrier future and writes d_echo[r] = total + r, and
 *            forwards the future itself: d_forwarded[r] = total. A reader
 *            that ran early, or a future left unresolved, shows as
 *            UINT64_MAX. */
int hz_futures_run(int kind, const uint64_t* d_starts, uint32_t leaves, uint32_t readers, uint64_t* d_echo,
                   uint64_t* d_forwarded, uint64_t* total, const hz_run_config* config, hz_run_stats* stats);

/* ------------------------------------------------------- batched BFS
 * A batch owns the CSR graphs of many jobs, their level arrays, and the
 * program instances that run them (kept between runs, so a run with the same
 * configuration reuses its instance and, with HZ_RUN_GRAPH, its hipGraph).
 * Graphs are laid out back to back: graph g's levels start at vertex_base[g],
 * its col at edge_base[g], its row_ptr at row_base[g] = vertex_base[g] + g. */
typedef struct hz_bfs_batch hz_bfs_batch;
typedef struct hz_bfs_info {
    uint32_t graphs;
    uint32_t reserved;
    uint64_t total_vertices;
    uint64_t total_edges;
    const int32_t* d_row_ptr;
    const int32_t* d_col;
    int32_t* d_level;
    double build_ms; /* device time of generation + CSR build */
} hz_bfs_info;

/* Vertex counts of an R-MAT batch: graph g has
 * round(2^(min_log2 + (max_log2 - min_log2) * u_g)) vertices, u_g uniform in
 * [0, 1) from Philox4x32-10 keyed by seed. Host-only computation. */
int hz_bfs_batch_sizes(uint64_t seed, uint32_t graphs, double min_log2, double max_log2, uint32_t* vertices);

/* Generates `graphs` R-MAT graphs (a, b, c = 0.57, 0.19, 0.19) of
 * vertices[g] vertices and edge_factor * vertices[g] directed edges on the
 * device, duplicates and self-loops kept, and builds their CSR with
 * neighbours in generation order. BFS source: vertex 0. */
int hz_bfs_batch_create(uint64_t seed, uint32_t graphs, const uint32_t* vertices, uint32_t edge_factor,
                        hz_bfs_batch** out);
/* A batch from host CSR arrays laid out as above (sources may be NULL: 0). */
int hz_bfs_batch_from_csr(uint32_t graphs, const uint32_t* vertices, const uint64_t* edges, const int32_t* row_ptr,
                          const int32_t* col, const uint32_t* sources, hz_bfs_batch** out);
int hz_bfs_batch_destroy(hz_bfs_batch* batch);
int hz_bfs_batch_info(const hz_bfs_batch* batch, hz_bfs_info* info);
/* Per-graph layout (any output may be NULL). */
int hz_bfs_batch_layout(const hz_bfs_batch* batch, uint32_t* vertices, uint64_t* edges, uint64_t* vertex_base,
                        uint64_t* edge_base, uint64_t* row_base);
/* BFS of every job of the batch in one program (kind HZ_ASYNC or HZ_EVENT);
 * levels land in info.d_level. */
int hz_bfs_batch_run(hz_bfs_batch* batch, int kind, const hz_run_config* config, hz_run_stats* stats);
/* Frees the program instances the batch keeps for reuse. */
int hz_bfs_batch_release(hz_bfs_batch* batch);

/* ------------------------------------------------------------ autotuning
 * Coordinate descent over hz_run_config: each round tries, axis by axis,
 * every listed value with the other fields fixed, keeping the fastest
 * (minimum over `reps` runs); configurations that fail to run the workload
 * are skipped. Axes: groups, queues, coalesce, arena (smallest within 3% of
 * the fastest) for every program; BFS adds block_groups, stash and the edge
 * split threshold and tunes the hipGraph-launched pass. The workload's
 * outputs hold the last trial's results. `log` (may be NULL) receives up to
 * log_capacity trials in order.
 *
 * Results are cached per program, device and workload class when a cache
 * file is set (hz_tune_cache_set_path, or the HZ_TUNE_CACHE environment
 * variable): a tuner call whose key is in the file returns the stored
 * configuration without running a trial (trials = 0, cached = 1), and a
 * tuning run stores its result. The key is
 * "<program>|<gcn arch>|<compute units>|<workload class>", the class being
 * the workload's size rounded to a power of two (Collatz starts, photons;
 * BFS graphs, vertices and edges). The file holds one line per key:
 * key, tab, the ten hz_run_config fields, tab, best_ms. */
typedef struct hz_tune_result {
    hz_run_config best;
    double best_ms;   /* time of the best configuration */
    double start_ms;  /* time of the starting configuration (-1: it failed) */
    uint32_t trials;  /* runs made */
    uint32_t failures;
    uint32_t cached;  /* 1: taken from the cache, no trial run */
    uint32_t reserved;
} hz_tune_result;
typedef struct hz_tune_trial {
    hz_run_config config;
    double ms; /* -1: the run failed */
} hz_tune_trial;

int hz_collatz_tune(const uint64_t* d_starts, uint32_t count, uint32_t* d_steps, uint32_t rounds, uint32_t reps,
                    hz_tune_result* result, hz_tune_trial* log, uint32_t log_capacity);
int hz_photon_tune(const hz_photon_params* params, double* d_tally, uint32_t* d_events, uint32_t rounds,
                   uint32_t reps, hz_tune_result* result, hz_tune_trial* log, uint32_t log_capacity);
int hz_bfs_batch_tune(hz_bfs_batch* batch, uint32_t rounds, uint32_t reps, hz_tune_result* result,
                      hz_tune_trial* log, uint32_t log_capacity);
/* Sets the tuning cache file (NULL or "" disables the cache). */
int hz_tune_cache_set_path(const char* path);
/* Looks up / stores one cache entry by key (host only). */
int hz_tune_cache_lookup(const char* key, hz_run_config* config, double* best_ms, int* found);
int hz_tune_cache_store(const char* key, const hz_run_config* config, double best_ms);

/* ------------------------------------------------------------- serving
 * A request server for Collatz, photon-walk and BFS jobs. Up to `slots` units
 * of work run at once, each on its own stream with the slot's share of the
 * device; the program instances of every slot are created with the server,
 * so admission allocates nothing. A Collatz or photon job of more than 2^18
 * items is cut into up to `slots` chunks of contiguous items; a BFS job runs
 * as one unit under the EventProgram. Job state lives in a device pool; a
 * job's state is allocated when its first unit is admitted, once a slot is
 * free and the state fits. Units are admitted in arrival order, or largest
 * job first (config.policy); the unit chosen waits for memory rather than
 * being overtaken. Jobs:
 *   HZ_JOB_COLLATZ  size starts from Philox4x32-10 keyed by seed, uniform in
 *                   [1, 2^32) (hz_collatz_make_starts with bits = 32);
 *                   output 0: u32 steps, output 1: u64 starts.
 *   HZ_JOB_PHOTON   the photon walk with hz_photon_defaults except photons =
 *                   size, seed = seed, replicas = 64; output 0: u32 events per
 *                   photon, output 1: the tally replicas (doubles), 64 per chunk.
 *   HZ_JOB_BFS      an R-MAT graph of size vertices and 16 x size edges
 *                   (graph 0 of hz_bfs_batch_create with this seed), built on
 *                   the device, then BFS from vertex 0; output 0: i32 levels,
 *                   1: row_ptr, 2: col.
 * Times are milliseconds on the server's steady clock. */
#define HZ_JOB_COLLATZ 0u
#define HZ_JOB_PHOTON 1u
#define HZ_JOB_BFS 2u
#define HZ_JOB_PENDING 0
#define HZ_JOB_RUNNING 1
#define HZ_JOB_DONE 2
#define HZ_JOB_FAILED (-1)

typedef struct hz_server hz_server;
typedef struct hz_server_config {
    uint64_t pool_bytes;  /* device memory for job state (0 = 8 GiB) */
    uint64_t arena_links; /* per slot and async program (0 = 2^16) */
    uint32_t slots;       /* concurrent jobs, 1..8 (0 = 3: a process has 4 hardware queues and the
                             default stream uses one) */
    uint32_t max_pending; /* submissions beyond this many waiting jobs fail (0 = no bound) */
    uint32_t keep_outputs; /* 1: a finished job keeps its state until hz_server_release */
    uint32_t policy;       /* admission order: HZ_SERVE_FIFO (0) or HZ_SERVE_LARGEST_FIRST */
} hz_server_config;

#define HZ_SERVE_FIFO 0u
#define HZ_SERVE_LARGEST_FIRST 1u
typedef struct hz_job {
    uint32_t type; /* HZ_JOB_* */
    uint32_t size; /* work items: starts, photons or vertices */
    uint64_t seed;
} hz_job;
typedef struct hz_job_result {
    uint64_t id;
    uint32_t type;
    uint32_t size;
    uint64_t seed;
    int32_t status; /* HZ_JOB_PENDING, _RUNNING, _DONE or _FAILED */
    uint32_t slot;
    uint64_t promises;
    uint32_t launches;
    uint32_t groups;
    double submit_ms; /* arrival */
    double start_ms;  /* admission (launch) */
    double end_ms;    /* completion seen by the server */
    uint64_t state_bytes;
    uint32_t chunks;  /* units the job was cut into (Collatz, photon) */
    uint32_t reserved;
} hz_job_result;
typedef struct hz_server_info_t {
    uint32_t slots;
    uint32_t pending;
    uint64_t jobs;
    uint64_t pool_bytes;
    uint64_t pool_used;
    uint64_t pool_largest_free;
    uint32_t groups_collatz; /* groups a job may use, per program */
    uint32_t groups_photon;
    uint32_t groups_bfs;
    uint32_t reserved;
} hz_server_info_t;

int hz_server_create(const hz_server_config* config, hz_server** out);
int hz_server_destroy(hz_server* server);
/* Queues a job (stamped with the current time) and makes progress. */
int hz_server_submit(hz_server* server, const hz_job* job, uint64_t* id);
/* Non-blocking: reaps finished passes, admits pending jobs. */
int hz_server_progress(hz_server* server, uint32_t* running, uint32_t* pending);
/* Makes progress until no job is pending or running. */
int hz_server_drain(hz_server* server, double timeout_ms);
int hz_server_job(hz_server* server, uint64_t id, hz_job_result* result);
/* Copies output `which` of a finished job (kept with keep_outputs); `bytes`
 * must equal the output's size. */
int hz_server_output(hz_server* server, uint64_t id, uint32_t which, void* host, uint64_t bytes);
int hz_server_release(hz_server* server, uint64_t id);
int hz_server_info(const hz_server* server, hz_server_info_t* info);
/* Open-loop replay: job i arrives arrival_ms[i] (sorted) after the call
 * starts; returns when all are finished. results[i] describes job i. */
int hz_server_run_trace(hz_server* server, const hz_job* jobs, const double* arrival_ms, uint32_t count,
                        double timeout_ms, hz_job_result* results);
/* Request/poll names: hz_serve_submit is hz_server_submit; hz_serve_poll
 * makes progress and returns (up to `capacity`) the ids of jobs that finished
 * (done or failed) since the previous poll. */
int hz_serve_submit(hz_server* server, const hz_job* job, uint64_t* id);
int hz_serve_poll(hz_server* server, uint64_t* finished_ids, uint32_t capacity, uint32_t* count);
/* Resident serving: Collatz or photon-walk jobs admitted into a running
 * persistent kernel.
 * The serving plan the reference's design calls for (a resident kernel fed by
 * a host doorbell; the reference itself has no serving loop): the host writes
 * each job's descriptor into a ring in host-coherent memory, a running
 * AsyncProgram launch installs it into its job table and claims the job's
 * items in priority order, and reports the job's completion in the ring; a
 * launch with no job for idle_us ends, and the next submission launches
 * again. A server runs jobs of one kind:
 *   HZ_JOB_COLLATZ  starts uniform in [1, 2^bits) from the seed (the serving
 *                   loop's generator); outputs 0 = step counts (u32[size]),
 *                   1 = starts (u64[size]);
 *   HZ_JOB_PHOTON   the photon walk with hz_photon_defaults except photons =
 *                   size and seed, 64 tally replicas; outputs 0 = events per
 *                   photon (u32[size]), 1 = the tally (f64[64 x (bins + 5)]).
 * At most 64 jobs are in flight; the rest wait in order. */
typedef struct hz_resident hz_resident;
typedef struct hz_resident_config {
    uint64_t pool_bytes;   /* device memory for job outputs (0 = 4 GiB): 12 bytes per Collatz
                              start; 4 per photon plus the job's tally (8 x 64 x (bins + 5)
                              bytes) per photon job */
    uint64_t arena_links;  /* link arena of the program (0 = 2^18) */
    uint32_t groups;       /* groups of a launch (0 = every group that fits on the device) */
    uint32_t policy;       /* claim order: HZ_RESIDENT_BY_ADMISSION (0) or HZ_RESIDENT_SMALLEST_FIRST
                              (Collatz only) */
    uint32_t idle_us;      /* a launch with no job this long ends (0 = 1000) */
    uint32_t poll_us;      /* the kernel's housekeeping period (0 = 2) */
    uint32_t life_ms;      /* a launch stops claiming after this long; the next one continues (0 = 10000) */
    uint32_t bits;         /* Collatz starts in [1, 2^bits), 1..40 (0 = 32) */
    uint32_t keep_outputs; /* 1: a finished job keeps its outputs until hz_resident_release */
    uint32_t kind;         /* HZ_JOB_COLLATZ (0) or HZ_JOB_PHOTON */
} hz_resident_config;
#define HZ_RESIDENT_BY_ADMISSION 0u
#define HZ_RESIDENT_SMALLEST_FIRST 1u
typedef struct hz_resident_info_t {
    uint64_t jobs;
    uint32_t pending;        /* waiting for a job-table entry or pool memory */
    uint32_t in_flight;      /* published and not finished */
    uint32_t launches;       /* kernel launches so far */
    uint32_t joined_running; /* jobs published while a launch was running */
    uint32_t groups;
    uint32_t running;
    uint64_t pool_bytes;
    uint64_t pool_used;
    /* Runtime counters summed over the launches that have ended. */
    uint64_t promises;
    uint64_t links_run;
    uint64_t make_work_calls;
    uint64_t idle_polls;
    uint64_t phase_cycles[8]; /* filled only in a build with HZ_PHASE_TIMING */
} hz_resident_info_t;
int hz_resident_create(const hz_resident_config* config, hz_resident** out);
/* Waits for the running launch to finish its jobs and end, then frees all. */
int hz_resident_destroy(hz_resident* server);
/* Queues a job and makes progress (publishes it if the table has room). */
int hz_resident_submit(hz_resident* server, const hz_job* job, uint64_t* id);
/* Makes progress and returns (up to `capacity`) the ids of jobs that
 * finished since the previous poll. */
int hz_resident_poll(hz_resident* server, uint64_t* finished_ids, uint32_t capacity, uint32_t* count);
/* Makes progress until every job has finished and the launch has ended. */
int hz_resident_drain(hz_resident* server, double timeout_ms);
/* result->launches is 0 for a job published while a launch was running. */
int hz_resident_job(hz_resident* server, uint64_t id, hz_job_result* result);
int hz_resident_output(hz_resident* server, uint64_t id, uint32_t which, void* host, uint64_t bytes);
int hz_resident_release(hz_resident* server, uint64_t id);
int hz_resident_info(const hz_resident* server, hz_resident_info_t* info);
/* Open-loop replay as hz_server_run_trace. */
int hz_resident_run_trace(hz_resident* server, const hz_job* jobs, const double* arrival_ms, uint32_t count,
                          double timeout_ms, hz_job_result* results);

/* Host-only check of the job-state pool allocator: ops[i] >= 0 allocates
 * that many bytes (out[i] = offset, or all-ones if it does not fit);
 * ops[i] = -k - 1 frees the k-th allocation (out[i] = largest free block). */
int hz_pool_test(uint64_t size, const int64_t* ops, uint32_t count, uint64_t* out);

/* Host-only check of the HZ_RUN_AUDIT audit itself: it must accept a
 * consistent snapshot of queues and arena and reject each of a set of planted
 * inconsistencies (a link queued twice, ring counters off, a stray slot, the
 * units count off, a lost promise, a leaked link, a bad operation, an empty
 * queued link, a group still checked in, an invalid id). Returns the number
 * of plants that went undetected (0 = all caught), or a negative status;
 * `report` receives a summary. */
int hz_audit_selftest(char* report, size_t report_bytes);

/* Device facts used to size launches. */
typedef struct hz_device_facts {
    int32_t device;
    int32_t compute_units;
    int32_t max_groups_collatz_async; /* resident async groups per device */
    int32_t max_groups_collatz_event;
    int64_t total_memory;
    char name[64];
    char arch[64];
} hz_device_facts;
int hz_device_query(hz_device_facts* facts);

int hz_device_synchronize(void);

/* Device-to-device copy of `bytes` bytes, synchronous (used to calibrate
 * HBM byte counters against a copy of known size). */
int hz_device_copy(void* dst, const void* src, size_t bytes);

#ifdef __cplusplus
}
#endif

#endif /* HARMONIZE_AMD_H */
