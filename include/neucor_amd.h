/*
 * neucor_amd.h -- C ABI of the MI355X NeuroCorrelation simulation engine.
 *
 * The engine replaces the internals of the reference's simulation core
 * (class NeuCor, src/NeuCor.h:36-138, src/NeuCor.cpp).  Every entry point
 * below names the reference interface it stands in for.  Plain pointers and
 * sizes only; no exceptions cross this boundary; every call returns an
 * nc_status and nc_last_error() describes the last failure on the thread.
 *
 * Objects:
 *   nc_net  -- a network on the host: neuron positions and, per neuron, the
 *              incoming synapses in ascending source id (the order of the
 *              std::map Neuron::inSynapses, src/NeuCor.h:212).
 *   nc_sim  -- a device-resident engine holding one or more networks.  The
 *              plain path is an nc_sim with one network (id 0); the batched
 *              and served forms add many networks of uneven size to one
 *              nc_sim and advance them together.
 *
 * Threading: one host thread per nc_sim at a time; each nc_sim owns one HIP
 * stream; read calls synchronise that stream.
 */
#ifndef NEUCOR_AMD_H
#define NEUCOR_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum nc_status {
  NC_OK = 0,
  NC_ERR_ARG = 1,       /* bad argument (std::invalid_argument-like)            */
  NC_ERR_RANGE = 2,     /* id out of range (std::out_of_range in the reference)  */
  NC_ERR_HIP = 3,       /* a HIP runtime call failed                             */
  NC_ERR_OOM = 4,       /* device or host allocation failed                      */
  NC_ERR_CAPACITY = 5,  /* the nc_sim has no room for the network (admission)    */
  NC_ERR_STATE = 6,     /* call not valid in the current state, or an unexpected */
                        /* host-side error                                       */
  NC_ERR_NO_DEVICE = 7  /* no usable GPU                                          */
} nc_status;

typedef struct nc_net nc_net;
typedef struct nc_sim nc_sim;

/* ---------------------------------------------------------------------------
 * Network construction
 * ------------------------------------------------------------------------ */

/* Distance wiring rule of the workloads (a cube of uniform-random positions;
 * every ordered pair closer than `radius` is a synapse, in both directions,
 * as Neuron::makeConnections does with radius 1.0, src/NeuCor.cpp:418-444).
 * delay steps = clamp(ceil(delay_max * dist / radius), delay_min, delay_max);
 * weight ~ U[w_min, w_max), negated with probability inhib_frac
 * (src/NeuCor.cpp:471-475 uses [0.2, 1.0) and 0.2).  All draws come from
 * Philox4x32-10 keyed by `seed`.  spatial_order = 1 numbers the neurons along
 * a Morton curve of radius-sized cells. */
typedef struct nc_wiring_spec {
  uint32_t n_neurons;
  uint32_t spatial_order;
  uint64_t seed;
  float side;
  float radius;
  uint32_t delay_min;
  uint32_t delay_max;
  float w_min;
  float w_max;
  float inhib_frac;
  uint32_t n_threads; /* host threads for the build (0 = hardware concurrency) */
} nc_wiring_spec;

/* Replaces NeuCor::NeuCor(int) + Neuron::makeConnections
 * (src/NeuCor.cpp:17-42, :418-444) for the workload wiring rule. */
nc_status nc_net_generate(const nc_wiring_spec* spec, nc_net** out);

/* The same network, bit for bit, built on GPU `device`: positions, a cell
 * grid and the rows of sources within the radius, sorted by source id, then
 * the per-synapse delays, weights and sign classes (spec->n_threads is
 * ignored).  Also replaces Neuron::makeConnections' all-pairs scan
 * (src/NeuCor.cpp:418-444). */
nc_status nc_net_generate_device(const nc_wiring_spec* spec, int32_t device, nc_net** out);

/* Replaces NeuCor::createNeuron / NeuCor::createSynapse
 * (src/NeuCor.cpp:154-195): an explicit network.  xyz may be NULL.
 * Synapse k goes from[k] -> to[k]; duplicates of an existing (from,to) pair
 * are ignored as createSynapse does (:189-191); delay in steps (>= 1);
 * inhibitory may be NULL (then the sign of w decides, as setWeight, :576-579). */
nc_status nc_net_from_edges(uint32_t n_neurons, const float* xyz,
                            uint64_t n_syn, const uint32_t* to,
                            const uint32_t* from, const float* w,
                            const uint8_t* delay_steps,
                            const uint8_t* inhibitory, nc_net** out);

nc_status nc_net_info(const nc_net* net, uint32_t* n_neurons, uint64_t* n_syn,
                      uint32_t* max_in_degree, uint32_t* max_delay);

/* Copy out the network: in_ptr has n+1 entries; rows in ascending source. */
nc_status nc_net_export(const nc_net* net, float* xyz, int64_t* in_ptr,
                        uint32_t* in_src, float* w, uint8_t* delay_steps,
                        uint8_t* inhibitory);

void nc_net_free(nc_net* net);

/* ---------------------------------------------------------------------------
 * Simulation parameters and input drive
 * ------------------------------------------------------------------------ */

/* NeuCor's public parameters (src/NeuCor.h:68-73, defaults :18-29).
 * dt is the step in ms (a power of two, 2^-8 .. 1): NeuCor::runSpeed with
 * runAll = true.  learning_rate and the factors are read live each step
 * (src/NeuCor.cpp:757-758); the trace decays are fixed at creation as in
 * the reference (const members, src/NeuCor.h:236, :289). */
typedef struct nc_params {
  float dt;
  float learning_rate;
  float pre_factor;
  float post_factor;
  float pre_trace_decay;
  float post_trace_decay;
} nc_params;

/* Device-generated drive (replaces InputFirer, src/NeuCor.cpp:305-345, and
 * the background firing of NeuCor::run, :604-607, for the workloads).
 * n_inputs neurons are chosen by seeded hash rank; the first n_correlated of
 * them (by rank) form n_groups groups that each share one Bernoulli spike
 * train, the rest fire independently.  p(s) = p0 * (1 + amp * sin(2 pi s /
 * period)).  Input fires bypass threshold and refractory (InputFirer::run
 * calls Neuron::fire directly, :326-331).  bg_prob forces a
 * threshold-bypassing but refractory-respecting fire per neuron per step
 * (Neuron::scheduleFire, :658-661, :683). */
typedef struct nc_drive_spec {
  uint32_t n_inputs;
  uint32_t n_groups;
  uint32_t n_correlated;
  uint32_t period;
  float p0;
  float amp_corr;
  float amp_ind;
  float bg_prob;
  uint64_t seed;
} nc_drive_spec;

/* The inputs nc_drive_spec selects on a network (ids ascending by rank) and
 * their group index (independent inputs get their own group). */
nc_status nc_net_select_inputs(const nc_net* net, const nc_drive_spec* drive,
                               uint32_t* ids, uint32_t* groups);

/* ---------------------------------------------------------------------------
 * Engine
 * ------------------------------------------------------------------------ */

/* Capacity of an nc_sim: total neurons, synapses and networks it may hold. */
typedef struct nc_sim_spec {
  int32_t device;
  uint64_t max_neurons;
  uint64_t max_synapses;
  uint32_t max_nets;
  uint32_t raster_capacity; /* spike records kept on device (0 = no raster) */
  uint64_t max_slots;       /* synapse slots incl. layout padding (0 = derive) */
} nc_sim_spec;

nc_status nc_sim_create(const nc_sim_spec* spec, nc_sim** out);

/* Exact device layout size of a network: neuron positions (a multiple of
 * 64) and synapse slots including padding, to size an nc_sim_spec. */
nc_status nc_net_device_size(const nc_net* net, uint64_t* positions, uint64_t* slots);

/* Builds and keeps the network's device image: its layout, slot records and
 * out-edges in network-local indices, so that every later nc_sim_add of it
 * only relocates and uploads them.  Optional: without it nc_sim_add builds
 * the image on each call and does not keep it.  Host work only (no device
 * needed); different networks may be prepared on different threads at once,
 * one network by one thread.  No counterpart in the reference, whose
 * neurons hold their synapses directly (src/NeuCor.h:212-213); for callers
 * that build networks off the thread that steps the simulation. */
nc_status nc_net_prepare(nc_net* net);
void nc_sim_destroy(nc_sim* sim);

/* Bytes a network of this size takes on the device, and the free HBM the
 * admission policy sees (hipMemGetInfo). */
nc_status nc_sim_footprint(uint64_t n_neurons, uint64_t n_syn, uint64_t* bytes);
nc_status nc_sim_free_bytes(const nc_sim* sim, uint64_t* free_bytes);

/* Upload a network; its state starts as NeuCor's constructors leave it
 * (V = -70, never fired, no spike in flight).  drive may be NULL.
 * steps_to_run = 0 means "run until told otherwise".  Fails with
 * NC_ERR_CAPACITY when the nc_sim has no room, NC_ERR_ARG for an in-degree
 * above 262140. */
nc_status nc_sim_add(nc_sim* sim, const nc_net* net, const nc_params* params,
                     const nc_drive_spec* drive, uint32_t steps_to_run,
                     int32_t* net_id);
nc_status nc_sim_remove(nc_sim* sim, int32_t net_id);

/* Replaces NeuCor::run() (src/NeuCor.cpp:583-617) with runAll = true and
 * runSpeed = dt: advance every active network by up to n_steps steps
 * (networks with fewer steps left stop early).  Launches kernels and copies
 * the per-call network windows and due host events to buffers sized by
 * nc_sim_create, nc_sim_add and nc_push_forced: it never allocates device
 * memory (under a graph launch shape it captures and instantiates a graph
 * when the shape or the set of networks changed). */
nc_status nc_sim_step(nc_sim* sim, uint32_t n_steps);

/* Steps each network has completed and has left (0 = finished when it was
 * added with a step budget; UINT32_MAX = unbounded). */
nc_status nc_sim_progress(nc_sim* sim, int32_t net_id, uint32_t* steps_done,
                          uint32_t* steps_left);

/* Replaces NeuCor::getTime() (src/NeuCor.cpp:225): steps_done * dt. */
nc_status nc_get_time(nc_sim* sim, int32_t net_id, float* time_ms);

/* Live parameters (NeuCor::learningRate, presynapticFactor,
 * postsynapticFactor: src/NeuCor.h:71-73). */
nc_status nc_set_params(nc_sim* sim, int32_t net_id, float learning_rate,
                        float pre_factor, float post_factor);

/* Host-scheduled forced fires for the next steps (replaces the events that
 * InputFirer::schedule, src/NeuCor.cpp:333-345, and the background loop,
 * :604-607, push onto the simulation queue).  Event k fires neuron ids[k]
 * at step offset step_off[k] (1 = the next step) of net_id; kind[k] is 0 for
 * an input fire (Neuron::fire, bypasses refractory) and 1 for a background
 * fire (scheduledFireTime, respects refractory).  Consumed by nc_sim_step.
 * Grows the device event buffer to hold every queued event when needed
 * (the step itself never allocates); on failure no event is queued. */
nc_status nc_push_forced(nc_sim* sim, int32_t net_id, uint32_t n_events,
                         const uint32_t* step_off, const uint32_t* ids,
                         const uint8_t* kind);

/* State read-out, by neuron id / by synapse in nc_net_export order.
 * Neuron::potential (src/NeuCor.cpp:456), Neuron::lastFire (src/NeuCor.h:226,
 * as a step; NC_NEVER_STEP = never), Neuron::firings (:645). */
#define NC_NEVER_STEP (-(1 << 30))
nc_status nc_read_potential(nc_sim* sim, int32_t net_id, float* out, uint32_t n);
nc_status nc_read_last_fire(nc_sim* sim, int32_t net_id, int32_t* out, uint32_t n);
nc_status nc_read_firings(nc_sim* sim, int32_t net_id, uint32_t* out, uint32_t n);

/* Synapse::getWeight / setWeight (src/NeuCor.cpp:572-579). */
nc_status nc_read_weights(nc_sim* sim, int32_t net_id, float* out, uint64_t n);
nc_status nc_write_weights(nc_sim* sim, int32_t net_id, const float* w, uint64_t n);

/* nc_read_weights without any device allocation.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * weights are gathered in export order into a static per-GPU scratch array
 * of 16,777,216 floats (64 MiB, held by every GPU the library runs on until
 * the process ends) and copied into out.  A network with more synapses than
 * that fails with NC_ERR_RANGE (nc_read_weights has no such limit).  Calls
 * that overlap on one GPU (from different threads) share the scratch and
 * race; calls on one nc_sim are serialised by its threading rule. */
nc_status nc_read_weights_noalloc(nc_sim* sim, int32_t net_id, float* out, uint64_t n);
/* Synapse::setWeight on the listed synapses only (indices in export order). */
nc_status nc_set_weights(nc_sim* sim, int32_t net_id, uint64_t n, const uint64_t* syn, const float* w);

/* Synapse spike state for the renderer's getPrePot/getPostPot
 * (src/NeuCor.cpp:551-567): arrival step of the spike in flight (0 = none,
 * Synapse::AP_fireTime == 0) and last arrival step (NC_NEVER_STEP = -inf). */
nc_status nc_read_synapse_state(nc_sim* sim, int32_t net_id, int32_t* arrival,
                                int32_t* last_arrival, float* depol, uint64_t n);

/* Spike raster (the renderer's raster plot reads Neuron::lastFire after each
 * run, src/NeuCor_Renderer.cpp:1858-1860).  Returns up to cap (step, id)
 * records logged since the last call, ordered by step then id. */
nc_status nc_read_raster(nc_sim* sim, int32_t net_id, uint32_t* steps,
                         uint32_t* ids, uint64_t cap, uint64_t* n_out);

/* NeuCor::resetActivities (src/NeuCor.cpp:233-235): firings := 0. */
nc_status nc_reset_activities(nc_sim* sim, int32_t net_id);

/* Event counters since creation, summed over all networks:
 * [0] steps*neurons, [1] synapse-steps (sum of S per step),
 * [2] active slot-steps, [3] arrivals, [4] fired neurons,
 * [5] in-degree of fired, [6] out-degree of fired, [7] accepted spikes,
 * [8] pending-list entries scanned (spikes whose delays exceed the
 * arrival ring), [9] mask words of every stepped neuron, ceil(in-degree /
 * 32) each, [10] active-mask words
 * written back, [11] fired-list records (per 64-neuron slice with fired
 * neurons, one per share of its out-edges given to a fan-out wave: up to 16,
 * by fired neurons times their network's mean out-degree in 64-edge groups,
 * summed over steps), [12] active-mask words read (those a
 * summary bit flags), [13] arrival-bucket entries taken, [14] slot records
 * read by the charge (active slots past the active-slot cache), [15] weights
 * written at arrival (pre-side STDP that changed the weight), [16]
 * active-slot cache records written (16 bytes each), [17] slot records read at arrival (the
 * target fired recently, or the spike came through a pending list), [18]
 * weights written by post-side STDP (those whose bits changed).
 * [0], [1] and [9] are kept by the host; the others are device counters that
 * advance only while counting is on (nc_set_counting).
 * Writes the first n (at most NC_NUM_COUNTERS) into out. */
#define NC_NUM_COUNTERS 19
nc_status nc_read_counters(nc_sim* sim, uint64_t* out, uint32_t n);

/* Checkpoint of one network's dynamic state.  No counterpart in the
 * reference, which keeps nothing resumable: for the served form's eviction
 * and resume, and for reproducible fixtures.  nc_sim_save writes potentials,
 * last fires, firing counts, every synapse's weight, latched depolarisation,
 * last arrival and arrival step, queued host events, the step count and the
 * plasticity parameters, in network-local order (independent of where the
 * network sits in the simulation), into buf; *size receives the bytes needed
 * (buf = NULL asks for the size only).  nc_sim_restore loads such a state
 * into a network just added (not yet stepped) from the same wiring at the
 * same dt, which then continues exactly as the saved one would have.
 * nc_sim_restore refuses (NC_ERR_ARG) a checkpoint whose last fires or last
 * arrivals are neither NC_NEVER_STEP nor in [1, steps done], whose arrivals
 * are neither NC_NEVER_STEP nor in [1, steps done + 255], or whose events
 * are of an unknown kind, before it writes anything. */
/* Byte layout of a checkpoint (little endian, 4-byte fields unless noted):
 *   header, NC_CKPT_HEADER_BYTES: u32 magic 'NCK1', u32 version 1, u32 n,
 *     u32 0, u64 synapses, u32 steps done (at NC_CKPT_STEPS_OFFSET),
 *     f32 learning rate, pre factor, post factor, dt, u32 queued events,
 *     u64 wiring digest;
 *   per neuron (id order): f32 potential, i32 last fire step, u32 firings;
 *   per synapse (nc_net_export order): f32 weight, f32 latched
 *     depolarisation, i32 last arrival step, i32 arrival step of the latched
 *     spike (NC_NEVER_STEP: none);
 *   per queued host event: u32 step offset, u32 neuron id, u32 kind.
 * A caller may rewrite state fields of a saved checkpoint (steps as
 * described at nc_sim_restore) but not the header's identity fields. */
#define NC_CKPT_HEADER_BYTES 56
#define NC_CKPT_STEPS_OFFSET 24
nc_status nc_sim_save(nc_sim* sim, int32_t net_id, void* buf, uint64_t cap, uint64_t* size);
nc_status nc_sim_restore(nc_sim* sim, int32_t net_id, const void* buf, uint64_t size);

/* Device event counters on (enable != 0) or off (the default: the step
 * kernels then carry no instrumentation). */
nc_status nc_set_counting(nc_sim* sim, int enable);

/* Launch shape of the step kernels.  No counterpart in the reference: the
 * knobs of the batched form, for callers that pack many networks into one
 * nc_sim.  waves_per_block: 1, 2, 4 or 8 slices (64 neurons each) per
 * workgroup; xcd_swizzle: each of the 8 XCDs takes one contiguous range of
 * slices; graph_steps: steps per captured hipGraph chunk (0 = direct
 * launches, at most 1024; a chunk is replayed while the set of networks is
 * unchanged);
 * stage_tables: copy the synapse-trace table of post-side STDP into LDS per
 * workgroup (pays when many synapses see STDP per step) instead of reading
 * it through the caches; fanout_waves: the most waves of the fan-out kernel, which share
 * the step's records of slices with fired neurons (below the cap, one
 * record per wave; never more waves than records can need); persistent:
 * the whole step (neuron phase, a grid-wide barrier, fan-out) in one launch
 * of a grid the device holds at once, its waves looping over the slices and
 * then sharing the fan-out (xcd_swizzle and fanout_waves do not apply; with
 * graph_steps the captured chunk is one such launch per step).  A barrier
 * that cannot complete (a grid not wholly resident, as when another
 * simulation's kernels hold part of the device) gives up after 2 s and the
 * next synchronising call fails with NC_ERR_STATE.  Results do not depend
 * on the launch shape.  Default {4, 0, 0, 0, 0, 0}; nc_sim_tune keeps
 * fanout_waves as it finds it. */
typedef struct nc_launch_cfg {
  uint32_t waves_per_block;
  uint32_t xcd_swizzle;
  uint32_t graph_steps;
  uint32_t stage_tables;
  uint32_t fanout_waves; /* waves of the fan-out kernel, 64..2^20 (0 = 16384) */
  uint32_t persistent;   /* 1: one launch per step with a grid barrier */
} nc_launch_cfg;
nc_status nc_sim_set_launch(nc_sim* sim, const nc_launch_cfg* cfg);
nc_status nc_sim_get_launch(const nc_sim* sim, nc_launch_cfg* cfg);

/* Autotuner of the launch shape over 38 shapes: {1, 2, 4, 8 slices per
 * workgroup} x {XCD order off/on} x {LDS tables off/on} x {no graph, graph
 * chunks of steps_per_trial steps}, and the persistent form with {2, 4, 8}
 * slices per workgroup x {no graph, graph chunks}, with the current
 * fanout_waves: as nc_sim_tune_list with 3 rounds. */
nc_status nc_sim_tune(nc_sim* sim, uint32_t steps_per_trial, nc_launch_cfg* chosen);

/* Autotuner over the caller's candidate shapes.  The simulation's state is
 * saved on the device first (the call allocates a copy of its buffers); every
 * trial starts from that state -- a warm-up trial of steps_per_trial steps,
 * which also captures the candidate's graph, the state put back, then the
 * timed trial of the same steps -- so that every candidate times the same
 * work.  `rounds` rounds visit the candidates in rotated order; each
 * candidate's median device time is its score, written to ms_out[i] (may be
 * NULL).  The fastest becomes the launch shape and is reported in *chosen.
 * Afterwards the simulation is exactly where it was before the call (no
 * simulated step is taken). */
nc_status nc_sim_tune_list(nc_sim* sim, const nc_launch_cfg* cands, uint32_t n_cands,
                           uint32_t steps_per_trial, uint32_t rounds, nc_launch_cfg* chosen,
                           float* ms_out);

/* Block until every step enqueued on the nc_sim's stream has finished. */
nc_status nc_sim_synchronize(nc_sim* sim);

/* Per-kernel timing for the kernels in kernel_mask (bit k: 0 drive, 1 neuron
 * step, 3 fan-out; 0 = off): start/stop events carried by the dispatch of
 * every `every`-th step (1 = all steps; the events cost the queue a little
 * time per timed launch).  nc_read_kernel_timing returns the timed launches
 * and their total milliseconds since the last nc_set_kernel_timing call. */
nc_status nc_set_kernel_timing(nc_sim* sim, int kernel_mask, uint32_t every);
nc_status nc_read_kernel_timing(nc_sim* sim, uint32_t kernel, uint64_t* launches,
                                double* total_ms);

/* Message of the last failed call on this thread ("" if none). */
const char* nc_last_error(void);

/* ---------------------------------------------------------------------------
 * Served form.  No counterpart in the reference, which runs one network per
 * process (src/main.cpp:56-60): a native request loop for callers that
 * submit many independent simulation jobs.  Jobs are built on builder
 * threads (on the GPU with device_wiring), admitted into the server's
 * simulations ("lanes", each with its own stream) while they have room,
 * stepped together (one kernel pair per step covers every job of a lane;
 * each job stops at its own step count), and retired at their step count.
 * A job's result is exactly what it gives alone: the step of one network
 * does not depend on the others.  With evict_after_s > 0, a job that has
 * waited that long for room evicts the running job with the most steps
 * left (nc_sim_save, then nc_sim_remove); the evicted job resumes later
 * from its checkpoint (nc_sim_add, nc_sim_restore) with the same result.
 * Every call is thread-safe; one internal thread drives the GPU.
 * ------------------------------------------------------------------------ */
typedef struct nc_server nc_server;

typedef struct nc_server_spec {
  int32_t device;
  uint32_t lanes;          /* simulations stepped in turn (0 = 2) */
  uint32_t chunk;          /* steps per lane turn (0 = 200) */
  uint32_t builders;       /* host threads building job networks (0 = 4) */
  uint32_t device_wiring;  /* 1: build job networks with nc_net_generate_device */
  uint32_t max_nets;       /* jobs per lane (0 = 2048) */
  uint64_t max_neurons;    /* capacity of each lane (0 = from hbm_fraction) */
  uint64_t max_synapses;
  float hbm_fraction;      /* with max_neurons = 0: the lanes together take this share
                              of the device's free HBM, for networks of syn_per_neuron
                              synapses per neuron (0 = 0.5) */
  float syn_per_neuron;    /* (0 = 60) */
  float evict_after_s;     /* 0 = never evict */
  uint32_t overflow_lanes; /* extra lanes created while a job waits for room and free
                              HBM (read when the job has waited evict_after_s, or 1 s
                              without eviction) holds one; destroyed when empty (0 = none) */
} nc_server_spec;

typedef struct nc_job {
  uint64_t id;             /* the caller's; returned in the result */
  nc_wiring_spec wiring;   /* n_threads is ignored */
  nc_drive_spec drive;
  nc_params params;
  uint32_t steps;          /* > 0 */
} nc_job;

typedef struct nc_job_result {
  uint64_t id;
  int32_t status;          /* NC_OK, or why the job did not run (NC_ERR_CAPACITY: larger than a lane) */
  uint32_t evictions;      /* times the job was evicted and resumed */
  uint32_t n_neurons;
  uint32_t steps;
  uint64_t n_synapses;
  uint64_t spikes;         /* sum of the final firing counts */
  uint64_t digest;         /* nc_digest of the final potentials, then weights (export
                              order), then firing counts, each in neuron-id order */
  double t_submit, t_built, t_admit, t_done;  /* seconds since nc_server_create */
} nc_job_result;

typedef struct nc_server_stats {
  uint64_t submitted, built, admitted, finished, evictions, chunks;
  uint32_t running, waiting, lanes;
  double gpu_wait_s, admit_s, retire_s, idle_s;
  double mean_running;     /* jobs per stepped chunk */
} nc_server_stats;

nc_status nc_server_create(const nc_server_spec* spec, nc_server** out);
/* Queues a job (its arrival time is now). */
nc_status nc_server_submit(nc_server* server, const nc_job* job);
/* Moves up to cap finished jobs' results into out, waiting up to wait_s
 * seconds for the first one. */
nc_status nc_server_poll(nc_server* server, nc_job_result* out, uint32_t cap, uint32_t* n_out, double wait_s);
nc_status nc_server_stats_get(nc_server* server, nc_server_stats* out);
/* Stops the loop and frees everything; unfinished jobs are dropped.  Fails
 * (and the server is still freed) if the loop had stopped on an error, whose
 * message nc_last_error() then holds. */
nc_status nc_server_destroy(nc_server* server);

/* 64-bit FNV-1a of n bytes continued from h (start: 14695981039346656037),
 * the digest of served results. */
uint64_t nc_digest(const void* data, uint64_t n, uint64_t h);

#ifdef __cplusplus
}
#endif
#endif /* NEUCOR_AMD_H */
