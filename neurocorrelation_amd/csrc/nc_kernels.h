// Device kernels of one simulation step (gfx950, wave64).
//
// One step s of every active network is, in order:
//   k_drive   host-scheduled forced fires -> bitmasks       (only when present)
//   k_neuron  one lane per neuron (one wave per 64-neuron slice):
//             charge_insynapses over the neuron's active slots
//             (depolarisation and arrival of the first NC_ACACHE from the
//             lane's active-slot cache, the rest from their slot records)
//             with the latch releases, charge_passive, threshold / forced
//             fire, AP waveform; then the wave applies post-side STDP to the
//             incoming synapses of each of its fired lanes, and takes up the
//             slice's arrivals at s -- its bucket of step s, then its pending
//             list -- each by the lane that owns the slot (pre-side STDP,
//             active bits, cache)
//   k_fanout  waves sharing the step's records of slices with fired lanes
//             (64 lists; a slice's out-edges split over up to NC_FSPLIT
//             records, one record per wave below a cap): Synapse::fire on
//             the fired lanes' outgoing synapses (source-order latch copy,
//             out->in slot map), appending accepted spikes to the target
//             slice's bucket of their arrival step, or, past NC_NB steps or
//             a full bucket, to its pending list of their arrival's window
// The kernel boundary orders the phases.  Every synapse slot has one writer
// per phase (its target's wave in k_neuron, its source's wave in k_fanout);
// the atomics are the appends to the buckets and pending lists (per target
// slice) and to the fired lists (per list), and the firing count.
#pragma once
#include <hip/hip_runtime.h>

#include "nc_common.h"

#define NC_DEG_MASK 0x00FFFFFFu
#define NC_DEG_INPUT 0x01000000u
#define NC_DEG_VALID 0x02000000u
#define NC_SLICE_NONE 0xFFFFFFFFu
#define NC_LDS_TAB_BYTES 32768  // a pre-trace table up to this size is staged in LDS
#define NC_WAP_K 8       // AP-table entries passed as kernel arguments
// Spikes in flight to a slice sit in one of two pending lists by arrival
// window: list (arrival >> NC_PEND_LOG_W) & 1, each at the slice's slot range
// of its own array (at most one spike in flight per slot).  An entry is
// (slot - first slot of the slice) << 8 | (arrival & 255): a list scanned at
// step s holds no arrival before s nor past s + 255 (delays are 1..255).
#define NC_PEND_LOG_W 4
// Spikes with delays up to NC_NB steps go to an arrival bucket instead: per
// slice, one bucket per arrival step modulo NC_NB, of up to NC_BCAP entries
// (slot - first slot of the slice | (delay - 1) << 24 | sign class << 31,
// weight at the fire, whose product with 10.4 is the latched
// depolarisation).  Fan-out takes
// the place with an atomic on the bucket's count; the step of arrival a reads
// bucket a % NC_NB of each slice -- exactly its arrivals, with what they
// contribute -- and empties it.  A spike that finds its bucket full, or
// whose delay exceeds NC_NB, goes to the pending list of its window.
#define NC_NB 64u
#define NC_BCAP 64u
// Slices with fired lanes go to one of NC_FLISTS lists (by wave index), so
// that the appends of a step spread over that many counters.
#define NC_FLISTS 64
// A slice with fired lanes puts its fan-out work -- the fired lanes'
// out-edges, flattened, in 64-edge groups -- into n records of its list:
// record i takes groups i, i + n, i + 2n, ...  When the fan-out grid has a
// wave per active slice or more, n = min(NC_FSPLIT, the groups the fired
// lanes' edges fill at the active networks' mean out-degree), so that a burst
// of fired lanes in one slice is shared by up to NC_FSPLIT fan-out waves;
// with fewer waves than slices every wave has records enough and n = 1.  A
// record is {mask lo, mask hi, step, slice | i << 24 | (n - 1) << 28}; when
// one lane fired (most fired slices at 1% activity) it is {first out-edge,
// out-degree, step, slice | NC_FSINGLE | i << 24 | (n - 1) << 28}, which
// spares fan-out the out-row load and the prefix sum.
#define NC_FSPLIT 16u
#define NC_FSLICE_BITS 23
#define NC_FSINGLE (1u << 23)
#define NC_FCNT_STRIDE 32  // words between two counters (one cache line each)
// In-slots are stored chunk-major: NC_CHUNK consecutive slots of a lane (one
// 128-byte line of 16-byte records), then the next lane's.
#define NC_CHUNK_LOG 3
#define NC_CHUNK (1u << NC_CHUNK_LOG)
#define NC_MAX_IN_DEGREE 262140u  // slots of a slice stay within 2^24 (64 lanes x 32768 chunks of 8)
#define NC_NCTR 19  // counter columns (device-side: 2-8, 10-18)
// Each lane keeps the depolarisation and arrival of its first NC_ACACHE
// active slots, in slot order, in one 16-byte record per position
// (acache[p] = {depol 0, depol 1, depol 2, arrival 0 & 255 | (arrival 1 &
// 255) << 8 | (arrival 2 & 255) << 16}): the charge reads those with one
// coalesced load and only further active slots from their scattered slot
// records, each a dependent read of several microseconds under load.  An
// active slot's arrival is one of the last R < 250 steps, so its low byte
// gives it back.  A slot's depolarisation and arrival do not change while it
// is active (its source cannot fire into it before the latch releases).
#define NC_ACACHE 3
// Slot records fetched together by the charge for active slots past the
// cache (one: fewer registers, a wave more per SIMD; 57 against 59 us per step
// of k_neuron at 10^6 neurons).
#define NC_RB 1
// Group spike trains are drawn for at most NC_GRING steps ahead, into a ring
// of that many rows (a call of more steps draws them chunk by chunk).
#define NC_GRING 1024u

// Per-network descriptor, rewritten by the host before each batch of steps.
struct nc_netdesc {
  int32_t base_step;  // steps completed before this batch
  int32_t remaining;  // steps this network runs in this batch
  float lr, pre_f, post_f;
  uint32_t drive_lo, drive_hi;  // drive seed
  uint32_t bg_thr;              // background threshold on a u32 draw (0 = off)
  uint32_t bg_lo, bg_hi;        // background seed
  int32_t whost;                // first step after the last host weight write
  uint32_t pad[5];
};

// One record per active slice (one wave).
struct nc_slice_rec {
  uint32_t q;         // slice index (positions 64q .. 64q+63)
  uint32_t slot_off;  // first synapse slot of the slice
  uint32_t net;       // network id
  uint32_t mask_off;  // first active-mask word of the slice
};

// One outgoing synapse, in its source's out-list.
struct nc_edge {
  uint32_t j;      // slot of the synapse at its target
  uint32_t tq_d;   // target slice << 8 | delay in steps
  int32_t arr;     // arrival step of the last accepted spike (NC_NEVER: none);
                   // the latch holds until arrival + release window
  uint32_t rel;    // the slot's place in its slice (slot - the slice's first slot)
};

struct nc_group {
  uint32_t tbl_off;  // first threshold entry
  uint32_t period;   // entries
  uint32_t local;    // group index inside its network (Philox counter word)
  uint32_t net;      // network id
};

struct nc_step_args {
  nc_slot* slots;
  uint32_t* amask;  // per slice: [mask word][lane] bits of active slots
  uint32_t* asum;   // per position: bit b set if mask words [b*G, b*G+G) may hold bits
                    // (G = 1 up to 32 words, else ceil(words/32))
  float* V;
  int32_t* last_fire;
  uint32_t* firings;
  const uint2* ndw;  // per position: x = in-degree | flags, y = input group
  const uint32_t* nid;
  const uint32_t* slice_off;
  const nc_slice_rec* active_slices;
  uint32_t n_active_slices;
  uint32_t act_identity;  // active record i is slice act_q0 + i (one contiguous run)
  uint32_t act_q0;
  const nc_netdesc* nets;
  const nc_group* groups;
  const uint32_t* thr;
  uint32_t* gfire;  // [step of the call & (NC_GRING - 1)][group words]: group spike trains
  uint32_t gwords;  // words per step in gfire
  uint32_t n_groups;
  uint32_t* forced_in;
  uint32_t* forced_bg;
  int32_t use_forced;
  uint4* fired;  // slices with fired lanes in NC_FLISTS lists (list b at b * fcap): mask lo/hi, step, slice
  uint32_t* fired_cnt;  // list lengths: [parity of the step in the call][list], NC_FCNT_STRIDE apart
  uint32_t fcap;        // entries per list
  uint32_t fsplit_od;   // mean out-degree for splitting fan-out records (0: no split)
  uint32_t* raster;           // (pos, step) pairs
  uint32_t* raster_cnt;
  uint32_t raster_cap;
  unsigned long long* ctr;  // [ctr_rows][NC_NCTR]: workgroup b adds into row b % ctr_rows
  uint32_t ctr_rows;
  const float* t_pre;
  const float* t_post;
  const float* w_ap;
  int32_t len_pre, len_post;
  int32_t tabs_in_lds;     // the pre-trace table is staged in (dynamic) LDS
  float w_ap_k[NC_WAP_K];  // the AP table when it has <= NC_WAP_K entries
  float dt, P, depol_k;
  double c0, e;
  int32_t d_ref, r_rel;
  int32_t chunk_i;            // step of the launch inside its chunk
  const uint32_t* chunk_off;  // device: steps of this call before the chunk
  uint32_t xcd_swizzle;       // blocks of one XCD take one contiguous range of slices
  uint32_t xcd_lists;         // fired lists by XCD: a wave of workgroup b uses list (b % 8) * 8 + its
                              // rank among the XCD's waves % 8, in k_neuron and k_fanout alike
  const uint2* out_row;     // per position: first out-edge, out-degree
  nc_edge* edges;           // out-edges, rows in ascending target id
  uint32_t* pend0;         // pending list 0 of every slice, at its slot range
  uint32_t* pend1;         // pending list 1
  uint2* pend_cnt;         // per slice: entries in pending list 0, in list 1
  uint32_t* bcnt;          // per slice: NC_NB arrival-bucket counts (beyond NC_BCAP: the rest overflowed)
  uint2* bent;             // per slice: NC_NB buckets of NC_BCAP entries {slot | delay | class, weight bits}
  uint4* acache;           // per position: (depol bits, arrival) of its first two active slots
  const uint32_t* ev_pos;  // host events of this step
  const uint8_t* ev_kind;
  uint32_t n_ev;
};

__device__ __forceinline__ float nc_lookup(const float* t, int32_t len, int32_t d) {
  return (d >= 0 && d < len) ? t[d] : 0.0f;
}

// fmax(fmin(w, hi), lo) as the reference evaluates it on x86-64
// (src/NeuCor.cpp:760-761); exact in float because the bounds are.
__device__ __forceinline__ float nc_clamp_w(float w, uint32_t inh) {
  const float hi = inh ? 0.0f : 1.0f;
  const float lo = inh ? -1.0f : 0.0f;
  float m = (w != w) ? hi : (w < hi ? w : hi);
  return m > lo ? m : lo;
}

// Synapse::synapticPlasticity (src/NeuCor.cpp:740-764) with both traces given.
__device__ __forceinline__ float nc_stdp(float w, uint32_t inh, float tr_s, float tr_t, float pre_f,
                                         float post_f, float lr) {
  float dw = pre_f * tr_s - post_f * tr_t;
  float wn = w + dw * lr;
  return nc_clamp_w(wn, inh);
}

// Entry i of a register array of NC_ACACHE cache entries (zero past the end),
// by selects rather than a scratch-indexed load.
__device__ __forceinline__ uint2 nc_pick(const uint2 (&c)[NC_ACACHE], uint32_t i) {
  uint2 v = make_uint2(0u, 0u);
#pragma unroll
  for (int e = 0; e < NC_ACACHE; ++e) {
    v.x = i == (uint32_t)e ? c[e].x : v.x;
    v.y = i == (uint32_t)e ? c[e].y : v.y;
  }
  return v;
}

// Inserts in-slot k (below 2^24) arriving at step arr, with depolarisation
// bits v, into the ascending list (nk, nv) of n entries, keeping the
// NC_ACACHE smallest slots; nk holds slot | (arr & 255) << 24.  Slot numbers
// are distinct; entries past n hold nothing, so whatever the shifts leave
// there is unused.
__device__ __forceinline__ void nc_push(uint32_t (&nk)[NC_ACACHE], uint32_t (&nv)[NC_ACACHE], uint32_t& n, uint32_t k,
                                        uint32_t arr, uint32_t v) {
  k |= arr << 24;
#pragma unroll
  for (int e = 0; e < NC_ACACHE; ++e) {
    const bool here = (uint32_t)e >= n || (k & 0xFFFFFFu) < (nk[e] & 0xFFFFFFu);
    const uint32_t tk = nk[e], tv = nv[e];
    nk[e] = here ? k : tk;
    nv[e] = here ? v : tv;
    k = here ? tk : k;
    v = here ? tv : v;
  }
  n = n < NC_ACACHE ? n + 1 : n;
}

__device__ __forceinline__ uint32_t nc_lane() {
  return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
}

// Wave-aggregated append of `pred` lanes to list[*cnt ...]; returns the slot
// (or ~0u) for each appending lane.
__device__ __forceinline__ uint32_t nc_wave_append(bool pred, uint32_t* cnt) {
  unsigned long long m = __ballot(pred);
  if (m == 0ull) return ~0u;
  uint32_t lane = nc_lane();
  uint32_t leader = (uint32_t)__builtin_ctzll(m);
  uint32_t base = 0;
  if (lane == leader) base = atomicAdd(cnt, (uint32_t)__builtin_popcountll(m));
  base = __shfl(base, (int)leader);
  unsigned long long below = (lane == 0) ? 0ull : (m & (~0ull >> (64 - lane)));
  return pred ? base + (uint32_t)__builtin_popcountll(below) : ~0u;
}

// Adds per-lane counts into the block's own counter row: a wave sum, then
// one fire-and-forget atomic from lane 0 (a row sees at most 4 writers).
__device__ __forceinline__ void nc_count32(unsigned long long* ctr, uint32_t row, int k, uint32_t v) {
  for (int o = 32; o > 0; o >>= 1) v += (uint32_t)__shfl_xor((int)v, o);
  if (nc_lane() == 0 && v) atomicAdd(&ctr[(size_t)row * NC_NCTR + k], (unsigned long long)v);
}

// Block index after the optional XCD remap: workgroups are dealt round-robin
// to the 8 XCDs, so block b runs on XCD b % 8; the remap gives each XCD one
// contiguous range of blocks (and of slices), which keeps a slice range and
// its spatially close fan-out targets in one L2.
__device__ __forceinline__ uint32_t nc_block_id(uint32_t swizzle) {
  const uint32_t b = blockIdx.x;
  if (!swizzle) return b;
  const uint32_t nb = gridDim.x, q = nb >> 3, r = nb & 7u, x = b & 7u, i = b >> 3;
  return x < r ? x * (q + 1) + i : r * (q + 1) + (x - r) * q + i;
}

// Advances the device-side chunk offset by the chunk's step count (the last
// node of a captured chunk); k_set_u32 clears it at the start of a call.
__global__ void k_advance(uint32_t* chunk_off, uint32_t steps) {
  if (threadIdx.x == 0) *chunk_off += steps;
}
__global__ void k_set_u32(uint32_t* p, uint32_t n, uint32_t v) {
  if (threadIdx.x < n) p[threadIdx.x] = v;
}
__global__ void k_zero_strided(uint32_t* p, uint32_t n, uint32_t stride) {
  if (threadIdx.x < n) p[threadIdx.x * stride] = 0u;
}

// Slot index of in-slot k of lane l in a slice whose slots start at slot_off
// (chunk-major: NC_CHUNK slots per lane per chunk).
NC_HD uint32_t nc_slot_of(uint32_t slot_off, uint32_t l, uint32_t k) {
  return slot_off + (k >> NC_CHUNK_LOG) * (64u * NC_CHUNK) + l * NC_CHUNK + (k & (NC_CHUNK - 1u));
}

// A network's weights in its export order (rows by target id, ascending
// source): thread i walks neuron i's row, at its position id_pos[pos0 + i]
// (network-local) and its first synapse id_row[pos0 + i]; the in-degree is
// the position's; slots are chunk-major inside the slice.
__global__ __launch_bounds__(256) void k_export_w(const nc_slot* slots, const uint32_t* slice_off, const uint2* ndw,
                                                  const uint32_t* id_pos, const uint32_t* id_row, uint32_t pos0,
                                                  uint32_t q0, uint32_t n, float* out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t p = id_pos[pos0 + i], row = id_row[pos0 + i];
  const uint32_t deg = ndw[pos0 + p].x & NC_DEG_MASK;
  const uint32_t so = slice_off[q0 + (p >> 6)], l = p & 63u;
  for (uint32_t k = 0; k < deg; ++k) out[row + k] = slots[nc_slot_of(so, l, k)].w;
}

// The weights of a run of slot records, packed 4 bytes per slot (host reads
// move a quarter of the records' bytes).
__global__ __launch_bounds__(256) void k_gather_w(const nc_slot* slots, uint64_t n, float* out) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) out[i] = slots[i].w;
}

// A network's out-edges from their uploaded network-local form (target slot,
// target slice << 8 | delay): target slot moved by the network's first slot,
// target slice by its first slice, the slot's active-mask word and bit from
// the target slice's first slot and first mask word (already in place), no
// spike latched.
__global__ __launch_bounds__(256) void k_place_edges(nc_edge* e, const uint2* jq, uint64_t n, uint32_t slot0,
                                                     uint32_t q0, const uint32_t* slice_off) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const uint2 x = jq[i];
    const uint32_t tq_d = x.y + (q0 << 8);
    const uint32_t j = x.x + slot0, tq = tq_d >> 8;
    e[i] = nc_edge{j, tq_d, NC_NEVER, j - slice_off[tq]};
  }
}

__global__ __launch_bounds__(256) void k_place_slots(nc_slot* s, const float* w, const uint32_t* inh, uint64_t n) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    s[i] = nc_slot{w[i], 0.0f, NC_NEVER, 2 * NC_NEVER + (int32_t)((inh[i >> 5] >> (i & 31)) & 1u)};
}

// Fresh neuron state: potential at the base level, never fired, no firings.
__global__ __launch_bounds__(256) void k_place_neurons(float* V, int32_t* last_fire, uint32_t* firings, uint64_t n) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    V[i] = NC_BASELEVEL;
    last_fire[i] = NC_NEVER;
    firings[i] = 0u;
  }
}

// Group spike trains of steps c0 .. c0 + gridDim.y - 1 of a call (at most
// NC_GRING steps): one Philox draw per (step, group), shared by every member
// of the group, into ring row (step & (NC_GRING - 1)).  blockIdx.y is the
// step, blockIdx.x * blockDim.x + threadIdx.x the group bit.
__global__ __launch_bounds__(256) void k_groups(nc_step_args a, uint32_t c0) {
  const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= a.gwords * 32u) return;  // whole waves: gwords * 32 is a multiple of 32, blocks of 256
  const uint32_t i = c0 + blockIdx.y;
  bool fire = false;
  if (g < a.n_groups) {
    const nc_group G = a.groups[g];
    const nc_netdesc* nd = &a.nets[G.net];
    if (i < (uint32_t)nd->remaining) {
      const uint32_t s = (uint32_t)(nd->base_step + (int32_t)i + 1);
      const uint64_t seed = ((uint64_t)nd->drive_hi << 32) | nd->drive_lo;
      const uint32_t u = nc_draw(seed, s, G.local, NC_STREAM_DRIVE).v[0];
      fire = u < a.thr[G.tbl_off + s % G.period];
    }
  }
  const unsigned long long m = __ballot(fire);
  const uint32_t lane = nc_lane();
  uint32_t* const row = a.gfire + (size_t)(i & (NC_GRING - 1u)) * a.gwords;
  if (lane == 0) row[g >> 5] = (uint32_t)m;
  if (lane == 32) row[g >> 5] = (uint32_t)(m >> 32);
}

// Host-scheduled forced fires of this step -> bitmasks read by k_neuron.
__global__ void k_drive(nc_step_args a) {
  uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= a.n_ev) return;
  uint32_t pos = a.ev_pos[t];
  uint32_t* m = a.ev_kind[t] ? a.forced_bg : a.forced_in;
  atomicOr(&m[pos >> 5], 1u << (pos & 31));
}

// Neuron::run (src/NeuCor.cpp:619-641) for the 64 neurons of active-slice
// record `wave` at step ci of the call (one wave of a workgroup of WPB waves;
// blk is the workgroup's counter row).
template <int WPB>
__device__ __forceinline__ void nc_neuron_slice(const nc_step_args& a, uint32_t blk, uint32_t wave, int32_t ci,
                                                const float* tpre) {
  const uint32_t lane = threadIdx.x & 63;
  const bool wave_in = wave < a.n_active_slices;
  const uint32_t widx = wave_in ? wave : 0u;
  const nc_slice_rec rec = a.active_slices[widx];
  // With one contiguous run of slices the per-neuron loads need not wait
  // for the slice record.
  const uint32_t q = a.act_identity ? a.act_q0 + widx : rec.q;
  const uint32_t pos = q * 64 + lane;
  uint2 nw = a.ndw[pos];
  float V = a.V[pos];
  const uint32_t V_in = __float_as_uint(V);
  int32_t lf = a.last_fire[pos];
  const uint2 pc = a.pend_cnt[q];
  uint32_t sm = a.asum[pos];
  // Loaded by every lane, with no branch between these loads and their uses,
  // so that the wave waits for each one alone (a load under a branch makes
  // the wait before the next use cover every load in flight): the active-slot
  // cache (read only while the walk is within it), the id (used by the
  // background draw), all NC_BCAP entry places of the arrival bucket of step
  // s (the first b_n hold this step's arrivals: spikes with delays up to NC_NB
  // that found room in it; the others wait in the pending list of s's arrival
  // window) and the group-spike-train word of an input's group (word 0 for
  // the others; gfire points at a zero word when there are no groups).
  const uint4 cache = a.acache[pos];
  const uint32_t my_nid = a.nid[pos];
  const nc_netdesc nd = a.nets[rec.net];
  const bool live = wave_in && ci < nd.remaining;
  const int32_t s = nd.base_step + ci + 1;
  const size_t bi = (size_t)q * NC_NB + ((uint32_t)s & (NC_NB - 1u));
  uint32_t b_n = a.bcnt[bi];
  const uint2 be = a.bent[bi * NC_BCAP + lane];
  const uint32_t gword = a.gfire[(size_t)((uint32_t)ci & (NC_GRING - 1u)) * a.gwords +
                                 ((nw.x & NC_DEG_INPUT) ? (nw.y >> 5) : 0u)];
  const uint32_t plist = ((uint32_t)s >> NC_PEND_LOG_W) & 1u;
  uint32_t pend_n = plist ? pc.y : pc.x;
  // The target trace at s of a lane that does not fire now (its last fire
  // stays lf), read with the mask words instead of after the fire: from a
  // clamped place (the table holds at least its entry 0) of the table in
  // global memory, whatever the launch shape, with no branch around the load
  // and its use after the fire. A lane that fires now has trace 0: entry 0
  // of every trace table, where powf(decay, 0) == 1 counts as no trace.
  const int32_t dlf0 = s - lf;
  const float tr_tab = a.t_post[min(max(dlf0, 0), a.len_post - 1)];
  if (!live) {
    nw = make_uint2(0u, 0u);
    pend_n = 0u;
    sm = 0u;
    b_n = 0u;
  }
  uint32_t* const pl = (plist ? a.pend1 : a.pend0) + rec.slot_off;
  const uint32_t dw = nw.x;
  const bool valid = (dw & NC_DEG_VALID) != 0;
  const uint32_t deg = dw & NC_DEG_MASK;
  const float lr = nd.lr, pre_f = nd.pre_f, post_f = nd.post_f;
  uint32_t n_active = 0, n_arr = 0, n_mwb = 0, n_mwr = 0, n_rec = 0, n_arec = 0, n_wset = 0, n_cw = 0;
  const uint32_t nmw = (deg + 31) >> 5;
  const uint32_t G = nmw <= 32 ? 1u : (nmw + 31) >> 5;  // mask words per summary bit
  uint32_t* const mrow = a.amask + rec.mask_off + lane;
  // The bucket's entries are taken up after the fire.
  const uint32_t b_take = b_n < NC_BCAP ? b_n : NC_BCAP;
  // The weight now of the bucket entry this lane holds, read at its slot
  // record with the mask words instead of after the fire, when the target
  // fired since the source did (as of the previous step) and no host write
  // came after the fire (the entry's sign class then still holds; after a
  // host write the record is read below, as for a target that fires now).
  // Nothing changes that weight before the entry is taken up, except
  // post-side STDP of a target that fires in this step, whose trace at s is
  // 0: its pre-side STDP then does not use the weight.  No branch: a lane
  // with no such entry reads the weight word of its own entry place, so that
  // e_w is the weight for every entry of the bucket.
  const uint32_t e_rel = be.x & 0xFFFFFFu;
  const uint32_t e_tl = (e_rel & (64u * NC_CHUNK - 1u)) >> NC_CHUNK_LOG;
  const int32_t e_sf = s - (int32_t)(((be.x >> 24) & 63u) + 1u);
  const int32_t e_lf = __shfl(lf, (int)e_tl);
  const bool e_stale = lane < b_take && e_lf > e_sf && e_sf >= nd.whost;
  const float* const e_wp =
      e_stale ? &a.slots[nc_slot_of(rec.slot_off, e_tl,
                                    ((e_rel >> (6 + NC_CHUNK_LOG)) << NC_CHUNK_LOG) | (e_rel & (NC_CHUNK - 1u)))]
                     .w
              : reinterpret_cast<const float*>(&a.bent[bi * NC_BCAP + lane].y);
  const float e_w = *e_wp;
  n_arec += e_stale ? 1u : 0u;
  // The lane's cache of (depolarisation, arrival) of its first NC_ACACHE
  // active slots in slot order (an arrival from its low byte: one of the
  // last R steps).
  const uint2 ac[NC_ACACHE] = {make_uint2(cache.x, (uint32_t)s - (((uint32_t)s - cache.w) & 255u)),
                               make_uint2(cache.y, (uint32_t)s - (((uint32_t)s - (cache.w >> 8)) & 255u)),
                               make_uint2(cache.z, (uint32_t)s - (((uint32_t)s - (cache.w >> 16)) & 255u))};

  // charge_insynapses (src/NeuCor.cpp:688-700): the active slots (spikes
  // that arrived at s-R .. s-1) in ascending source order.  A lane walks its
  // mask words four at a time; the i-th active slot's depolarisation and
  // arrival come from the cache while i < NC_ACACHE, else from its slot
  // record.  The latch is released at o = R (:697).  The kept slots, then the
  // arrivals of this step, form the next cache.
  // the next cache, ascending: slot | (arrival & 255) << 24, depolarisation bits
  uint32_t nk[NC_ACACHE] = {0u, 0u, 0u}, nv[NC_ACACHE] = {0u, 0u, 0u};
  uint32_t ncnt = 0, idx = 0;
  bool cache_dirty = false;
  uint32_t kw[4] = {0u, 0u, 0u, 0u};  // kept bits of mask words 0-3
  uint32_t sm_keep = 0;  // summary bits whose words hold bits after the releases and arrivals
  for (uint32_t w0 = 0; w0 < nmw; w0 += 4) {
    uint32_t mw[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const uint32_t w = w0 + u;
      const bool flagged = w < nmw && ((sm >> (G == 1u ? w : w / G)) & 1u);
      mw[u] = flagged ? mrow[w * 64] : 0u;
      n_mwr += flagged ? 1u : 0u;
    }
    // the active slots of these four words, ascending: bit p of lo (words
    // w0, w0+1) or of hi (w0+2, w0+3) is in-slot 32 w0 + p
    unsigned long long lo = ((unsigned long long)mw[1] << 32) | mw[0];
    unsigned long long hi = ((unsigned long long)mw[3] << 32) | mw[2];
    unsigned long long klo = lo, khi = hi;  // bits kept after the releases
    while (lo | hi) {
      // up to NC_RB slots at a time, applied in slot order
      uint32_t pp[NC_RB];
      uint32_t cnt = 0;
#pragma unroll
      for (int e = 0; e < NC_RB; ++e) {
        uint32_t p = 0;
        if (lo) {
          p = (uint32_t)__builtin_ctzll(lo);
          lo &= lo - 1;
          cnt = e + 1;
        } else if (hi) {
          p = 64u + (uint32_t)__builtin_ctzll(hi);
          hi &= hi - 1;
          cnt = e + 1;
        }
        pp[e] = p;
      }
      uint2 dv[NC_RB];  // (depol bits, arrival) of each
#pragma unroll
      for (int e = 0; e < NC_RB; ++e) {
        dv[e] = nc_pick(ac, idx + e);
        if ((uint32_t)e < cnt && idx + e >= NC_ACACHE) {
          const nc_slot r = a.slots[nc_slot_of(rec.slot_off, lane, 32u * w0 + pp[e])];
          dv[e] = make_uint2(__float_as_uint(r.depol), (uint32_t)(r.arr_inh >> 1));
          ++n_rec;
        }
      }
#pragma unroll
      for (int e = 0; e < NC_RB; ++e) {
        if ((uint32_t)e >= cnt) break;
        const int32_t o = s - (int32_t)dv[e].y;  // in [1, R]: the bit is cleared at o == R
        const float inc = a.dt * __uint_as_float(dv[e].x);
        V = (float)((double)V + ((double)inc * a.c0) * a.e);
        if (o >= a.r_rel) {
          if (pp[e] < 64u) klo &= ~(1ull << pp[e]);
          else khi &= ~(1ull << (pp[e] - 64u));
          cache_dirty = true;
        } else {
          nc_push(nk, nv, ncnt, 32u * w0 + pp[e], dv[e].y, dv[e].x);
        }
      }
      idx += cnt;
      n_active += cnt;
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const uint32_t w = w0 + u;
      if (w >= nmw) break;
      const uint32_t kept = (uint32_t)((u < 2 ? klo : khi) >> (32 * (u & 1)));
      if (w0 == 0) kw[u] = kept;
      if (kept != mw[u]) {
        mrow[w * 64] = kept;
        ++n_mwb;
      }
      if (kept) sm_keep |= 1u << (G == 1u ? w : w / G);
    }
  }

  // charge_passive (src/NeuCor.cpp:677-680).
  V = (V - NC_BASELEVEL) * a.P + NC_BASELEVEL;

  // Forced fires: group spike trains, background, host events.
  bool f_in = false, f_bg = false;
  if (valid && (dw & NC_DEG_INPUT)) f_in = ((gword >> (nw.y & 31u)) & 1u) != 0;
  if (valid && nd.bg_thr) {
    const uint64_t seed = ((uint64_t)nd.bg_hi << 32) | nd.bg_lo;
    const uint32_t u = nc_draw(seed, (uint32_t)s, my_nid, NC_STREAM_BG).v[0];
    f_bg = u < nd.bg_thr;
  }
  if (a.use_forced && live) {
    const uint32_t wi = a.forced_in[pos >> 5], wb = a.forced_bg[pos >> 5];
    f_in |= ((wi >> (pos & 31)) & 1u) != 0;
    f_bg |= ((wb >> (pos & 31)) & 1u) != 0;
    if ((pos & 31) == 0) {
      if (wi) a.forced_in[pos >> 5] = 0u;
      if (wb) a.forced_bg[pos >> 5] = 0u;
    }
  }

  // charge_thresholdCheck (src/NeuCor.cpp:682-686) + input fires (:326-331).
  const bool refr_ok = (s - lf) > a.d_ref;
  const bool fire = valid && ((((NC_THRESHOLD < V) || f_bg) && refr_ok) || f_in);
  const bool tr_in = (uint32_t)(s - lf) < (uint32_t)a.len_post;  // the last fire before this step's
  if (fire) lf = s;
  // AP waveform (src/NeuCor.cpp:706-714) from the host table.
  const int32_t dlf = s - lf;
  if (dlf <= a.d_ref) {
    if (a.d_ref < NC_WAP_K) {
      float wv = a.w_ap_k[0];
#pragma unroll
      for (int k = 1; k < NC_WAP_K; ++k) wv = (dlf == k) ? a.w_ap_k[k] : wv;
      V = wv;
    } else {
      V = a.w_ap[dlf];
    }
  }

  uint32_t n_post = 0;
  const unsigned long long fm = __ballot(fire);
  // the slice's fan-out records (lane 0), written after post-side STDP
  uint32_t f_at = 0u, f_rx = 0u, f_ry = 0u, f_tag = 0u, f_n = 0u, f_b = 0u;
  {
    // This step's lists get the fired slices' fan-out records; the other
    // parity's lists, which served the previous step's fan-out, are cleared
    // for the next.
    const uint32_t par = (uint32_t)ci & 1u;
    if (blk == 0 && threadIdx.x < NC_FLISTS)
      a.fired_cnt[((par ^ 1u) * NC_FLISTS + threadIdx.x) * NC_FCNT_STRIDE] = 0u;
    if (fm && lane == 0) {
      // the fired lanes' out-edges, in 64-edge groups, shared by n records
      // (written by one lane: a record per lane would hold registers at the
      // kernel's tightest point)
      const uint32_t nfl = (uint32_t)__builtin_popcountll(fm);
      const uint32_t n_frec = a.fsplit_od ? min(NC_FSPLIT, (nfl * a.fsplit_od + 63u) >> 6) : 1u;
      // waves running together use different lists; with XCD lists, those of
      // the workgroup's own XCD, which its fan-out waves read
      const uint32_t b = a.xcd_lists ? (blockIdx.x & 7u) * 8u + ((((blockIdx.x >> 3) * WPB) + (threadIdx.x >> 6)) & 7u)
                                     : wave % NC_FLISTS;
      uint32_t rx = (uint32_t)fm, ry = (uint32_t)(fm >> 32), tag = q;
      if (nfl == 1u) {
        // one fired lane: its out-edge range, loaded beside the list atomic
        const uint2 orow = a.out_row[q * 64u + (uint32_t)__builtin_ctzll(fm)];
        rx = orow.x;
        ry = orow.y;
        tag |= NC_FSINGLE;
      }
      // The places now, the records after post-side STDP: the atomic's
      // return (many waves append to each list) then overlaps the row passes'
      // reads instead of holding them back.
      f_at = atomicAdd(&a.fired_cnt[(par * NC_FLISTS + b) * NC_FCNT_STRIDE], n_frec);
      f_rx = rx;
      f_ry = ry;
      f_tag = tag | ((n_frec - 1u) << 28);
      f_n = n_frec;
      f_b = b;
      if (a.ctr) atomicAdd(&a.ctr[(size_t)blk * NC_NCTR + 11], (unsigned long long)n_frec);
    }
  }
  if (fm) {
    // Post-side STDP of each fired lane (the inSynapses loop of Neuron::fire,
    // src/NeuCor.cpp:651-653): target trace powf(decay, 0) == 1 -> 0, synapse
    // trace of lastSpikeArrival from before this step's arrivals: the
    // record's arrival if it came before s, else last_arr (the one before
    // it, or a last arrival whose latch a restore cleared).  The whole wave
    // works on one fired lane's row at a time.
    unsigned long long mm = fm;
    uint32_t n_pw = 0;  // weights written (a wave count)
    while (mm) {
      const uint32_t l = (uint32_t)__builtin_ctzll(mm);
      mm &= mm - 1;
      const uint32_t dl = (uint32_t)__shfl((int)deg, (int)l);
      n_post += (lane < dl) ? ((dl - lane + 63) >> 6) : 0u;
      for (uint32_t k = lane; k < dl; k += 128) {
        const bool two = k + 64 < dl;
        const uint32_t j0 = nc_slot_of(rec.slot_off, l, k);
        const uint32_t j1 = nc_slot_of(rec.slot_off, l, k + 64);
        // both records in one round (the first again when there is no second;
        // the empty asm keeps the compiler from waiting for the first before
        // it issues the second)
        nc_slot s0 = a.slots[j0];
        nc_slot s1 = a.slots[two ? j1 : j0];
        asm volatile("" : "+v"(s0.w), "+v"(s0.last_arr), "+v"(s0.arr_inh), "+v"(s1.w), "+v"(s1.last_arr),
                     "+v"(s1.arr_inh));
        const int32_t a0 = s0.arr_inh >> 1, a1 = s1.arr_inh >> 1;
        // the record's arrival if it is a step in [1, s-1], else last_arr
        const int32_t la0 = (uint32_t)(a0 - 1) < (uint32_t)(s - 1) ? a0 : s0.last_arr;
        const int32_t la1 = (uint32_t)(a1 - 1) < (uint32_t)(s - 1) ? a1 : s1.last_arr;
        // Both trace reads in one round, before either store (a read after a
        // store to memory it may alias waits for that store): from clamped
        // places (the table holds at least its entry 0), with no branch
        // around them; the empty asm keeps each read out of a branch that
        // would otherwise be formed around the selects below.
        const int32_t d0 = s - la0, d1 = s - la1;
        float tv0 = tpre[min(max(d0, 0), a.len_pre - 1)];
        float tv1 = tpre[min(max(d1, 0), a.len_pre - 1)];
        asm volatile("" : "+v"(tv0), "+v"(tv1));
        const float tp0 = (uint32_t)d0 < (uint32_t)a.len_pre ? tv0 : 0.0f;
        const float tp1 = (uint32_t)d1 < (uint32_t)a.len_pre ? tv1 : 0.0f;
        // a weight whose synapse saw no recent arrival keeps its bits (the
        // trace is below half an ulp of it): no store
        const float w0 = nc_stdp(s0.w, (uint32_t)s0.arr_inh & 1u, tp0, 0.0f, pre_f, post_f, lr);
        const bool ch0 = __float_as_uint(w0) != __float_as_uint(s0.w);
        bool ch1 = false;
        float w1 = 0.0f;
        if (two) {
          w1 = nc_stdp(s1.w, (uint32_t)s1.arr_inh & 1u, tp1, 0.0f, pre_f, post_f, lr);
          ch1 = __float_as_uint(w1) != __float_as_uint(s1.w);
        }
        if (ch0) a.slots[j0].w = w0;
        if (ch1) a.slots[j1].w = w1;
        // counted per wave (a scalar), not per lane
        n_pw += (uint32_t)__builtin_popcountll(__ballot(ch0)) + (uint32_t)__builtin_popcountll(__ballot(ch1));
      }
    }
    if (a.ctr && lane == 0 && n_pw) atomicAdd(&a.ctr[(size_t)blk * NC_NCTR + 18], (unsigned long long)n_pw);
    // the arrivals below read records the row passes may have written
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
  }
  if (lane == 0)
    for (uint32_t i = 0; i < f_n; ++i)
      a.fired[f_b * a.fcap + f_at + i] = make_uint4(f_rx, f_ry, (uint32_t)s, f_tag | (i << 24));

  // Spikes arriving at s (Synapse::run, :718-726): the bucket's entries,
  // which carry the weight at the fire (hence the latched depolarisation),
  // the delay and the sign class, then the pending list's arrivals, whose
  // state is read from the slot record.  Each one takes its pre-side STDP
  // now -- synapse trace powf(decay, 0) == 1 -> 0, target trace of lastFire
  // after this step; the record already holds the arrival as
  // lastSpikeArrival -- and joins the masks (contributing from s+1) and the
  // next cache.  The weight is the entry's while nothing can have changed it
  // since the fire (the target has not fired since, and no host write came
  // after it); else, and for pending-list spikes, the record is read.  An
  // entry belongs to the lane that owns its slot: six ballots on the bits of
  // the target lane give each lane the set of entries that are its own.
  const float tr_t = (!fire && tr_in) ? tr_tab : 0.0f;
  uint32_t arrw = 0;  // which of mask words 0-3 took arrivals
  // batch 0: the bucket's entries; batch b > 0: pending entries 64 (b - 1) ..
  const uint32_t n_batches = 1u + (pend_n + 63u) / 64u;
  uint32_t kept = 0;  // pending entries kept (compacted in place: write index <= read index)
  for (uint32_t bt = (b_take ? 0u : 1u); bt < n_batches; ++bt) {
    const bool known = bt == 0u;  // bucket entries carry the weight, delay and class
    uint32_t rel = be.x & 0xFFFFFFu, info = be.x >> 24, wbits = be.y, e = 0u, t = 0u;
    bool ev = lane < b_take;
    if (!known) {
      t = (bt - 1u) * 64u + lane;
      e = t < pend_n ? pl[t] : 0u;
      ev = t < pend_n && (e & 255u) == ((uint32_t)s & 255u);
      rel = e >> 8;
      info = 0u;
      wbits = 0u;
    }
    // entry (slot - first slot of the slice) -> the lane that owns the slot
    const uint32_t tl = (rel & (64u * NC_CHUNK - 1u)) >> NC_CHUNK_LOG;
    const uint32_t kk = ((rel >> (6 + NC_CHUNK_LOG)) << NC_CHUNK_LOG) | (rel & (NC_CHUNK - 1u));
    // The weight now, the sign class and the latched depolarisation of each
    // entry, found by the lane that holds the entry.  A bucket entry's weight
    // was read with the mask words, unless a host write came after its fire:
    // then, as for a pending-list spike, the slot record is read now, all in
    // one round (the sign class too, which the write may have changed).  A
    // target whose trace at s is 0 takes no pre-side STDP and does not use
    // the weight.
    const int32_t s_fire = s - (int32_t)((info & 63u) + 1u);
    uint32_t w_now = known ? __float_as_uint(e_w) : 0u, kinh = kk | ((info >> 7) << 31);
    uint32_t dkv = __float_as_uint(a.depol_k * __uint_as_float(wbits));  // Synapse::fire (:730-732), as fan-out made it
    if (ev && (!known || s_fire < nd.whost)) {
      const nc_slot r = a.slots[nc_slot_of(rec.slot_off, tl, kk)];
      ++n_arec;
      w_now = __float_as_uint(r.w);
      kinh = kk | (((uint32_t)r.arr_inh & 1u) << 31);
      if (!known) dkv = __float_as_uint(r.depol);
    }
    unsigned long long mine = __ballot(ev);
#pragma unroll
    for (int bb = 0; bb < 6; ++bb) {
      const unsigned long long m = __ballot(ev && ((tl >> bb) & 1u));
      mine &= ((lane >> bb) & 1u) ? m : ~m;
    }
    while (__ballot(mine != 0ull)) {
      const uint32_t src = mine ? (uint32_t)__builtin_ctzll(mine) : lane;
      const uint32_t ki = (uint32_t)__shfl((int)kinh, (int)src);
      const uint32_t wb = (uint32_t)__shfl((int)w_now, (int)src);
      const uint32_t dk = (uint32_t)__shfl((int)dkv, (int)src);
      if (mine) {
        const uint32_t k = ki & 0x7FFFFFFFu;
        mine &= mine - 1;
        const uint32_t w = k >> 5, bit = 1u << (k & 31u);
        if (w < 4u) {
#pragma unroll
          for (int u = 0; u < 4; ++u)
            if (w == (uint32_t)u) kw[u] |= bit;
          arrw |= 1u << w;
        } else {
          mrow[w * 64] |= bit;
          ++n_mwb;
        }
        sm_keep |= 1u << (G == 1u ? w : w / G);
        const uint32_t j = nc_slot_of(rec.slot_off, lane, k);
        if (tr_t != 0.0f) {
          const float wn = nc_stdp(__uint_as_float(wb), ki >> 31, 0.0f, tr_t, pre_f, post_f, lr);
          if (__float_as_uint(wn) != wb) {
            a.slots[j].w = wn;
            ++n_wset;
          }
        }
        nc_push(nk, nv, ncnt, k, (uint32_t)s, dk);
        cache_dirty = true;
        ++n_arr;
      }
    }
    if (!known) {
      const bool keep = t < pend_n && !ev;
      const unsigned long long km = __ballot(keep);
      const uint32_t to = kept + (uint32_t)__builtin_popcountll(km & ((1ull << lane) - 1ull));
      if (keep && to != t) pl[to] = e;
      kept += (uint32_t)__builtin_popcountll(km);
    }
  }
  if (pend_n && lane == 0) reinterpret_cast<uint32_t*>(&a.pend_cnt[q])[plist] = kept;
  // the bucket is emptied for its next use (step s + NC_NB, whose spikes the
  // fan-out after this kernel may start to append)
  if (b_n && lane == 0) a.bcnt[bi] = 0u;
#pragma unroll
  for (int u = 0; u < 4; ++u)
    if ((arrw >> u) & 1u) {
      mrow[u * 64] = kw[u];
      ++n_mwb;
    }
  if (cache_dirty) {
    // one 16-byte store; entries past ncnt are never read (the lane has
    // ncnt active slots)
    a.acache[pos] = make_uint4(nv[0], nv[1], nv[2], (nk[0] >> 24) | ((nk[1] >> 24) << 8) | ((nk[2] >> 24) << 16));
    ++n_cw;
  }
  if (live && sm_keep != sm) a.asum[pos] = sm_keep;
  // The neuron's own state is written last (nothing in this kernel reads it
  // back), so that no wait on the way above covers these stores.
  if (valid) {
    // a neuron at rest stays at the base level bit for bit: no store
    if (__float_as_uint(V) != V_in) a.V[pos] = V;
    if (fire) {
      a.last_fire[pos] = lf;
      // an atomic whose result is unused: the wave does not wait for the count
      __hip_atomic_fetch_add(&a.firings[pos], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
  }

  if (a.raster_cap) {
    const uint32_t rt = nc_wave_append(fire, a.raster_cnt);
    if (fire && rt < a.raster_cap) {
      a.raster[2ull * rt] = pos;
      a.raster[2ull * rt + 1] = (uint32_t)s;
    }
  }
  if (a.ctr) {
    nc_count32(a.ctr, blk, 2, n_active);
    nc_count32(a.ctr, blk, 3, n_arr);
    if (lane == 0 && fm) {
      atomicAdd(&a.ctr[(size_t)blk * NC_NCTR + 4], (unsigned long long)__builtin_popcountll(fm));
    }
    nc_count32(a.ctr, blk, 5, n_post);
    nc_count32(a.ctr, blk, 10, n_mwb);
    if (lane == 0 && pend_n) atomicAdd(&a.ctr[(size_t)blk * NC_NCTR + 8], (unsigned long long)pend_n);
    nc_count32(a.ctr, blk, 12, n_mwr);
    if (lane == 0 && b_take) atomicAdd(&a.ctr[(size_t)blk * NC_NCTR + 13], (unsigned long long)b_take);
    nc_count32(a.ctr, blk, 14, n_rec);
    nc_count32(a.ctr, blk, 15, n_wset);
    nc_count32(a.ctr, blk, 16, n_cw);
    nc_count32(a.ctr, blk, 17, n_arec);
  }
}

// The synapse-trace table of post-side STDP in LDS when the launch shape asks
// for it (once per workgroup), else through the caches.  (The target-trace
// table is read once per lane, with the prologue's loads, from global memory.)
template <int WPB>
__device__ __forceinline__ const float* nc_trace_tables(const nc_step_args& a) {
  extern __shared__ float sh_tab[];  // the pre-trace table
  if (!a.tabs_in_lds) return a.t_pre;
  for (int32_t i = threadIdx.x; i < a.len_pre; i += 64 * WPB) sh_tab[i] = a.t_pre[i];
  __syncthreads();
  return sh_tab;
}

// The neuron phase of one step for every active slice, one wave per slice,
// WPB waves per workgroup.
template <int WPB>
__global__ __launch_bounds__(64 * WPB) void k_neuron(nc_step_args a) {
  const uint32_t blk = nc_block_id(a.xcd_swizzle);
  const uint32_t wave = __builtin_amdgcn_readfirstlane(blk * WPB + (threadIdx.x >> 6));
  const float* tpre = nc_trace_tables<WPB>(a);
  nc_neuron_slice<WPB>(a, blk % a.ctr_rows, wave, (int32_t)*a.chunk_off + a.chunk_i, tpre);
}

// Synapse::fire (src/NeuCor.cpp:727-738) on every outgoing synapse of every
// fired neuron: each wave takes every (nwaves / 64)-th record of one of the
// step's 64 fired lists, 64 per load, and the records one by one (a record is
// a share of one fired slice's out-edges); one writer per synapse (its
// source), latch permitting.  Accepted spikes latch
// depolarisation and arrival step into the target slot and join the bucket of
// their arrival step at the target slice (or the slice's pending list).
__device__ __forceinline__ void nc_fire_edge(const nc_step_args& a, uint32_t e, const nc_edge& ed, int32_t s,
                                             uint32_t& n_acc) {
  // a spike still in flight or in its release window (:728)
  if (ed.arr + a.r_rel > s) return;
  const uint32_t d = ed.tq_d & 0xffu;
  const int32_t arr = s + (int32_t)d;
  const uint32_t tq = ed.tq_d >> 8;
  // the bucket place is taken before the slot is read, so that the two round
  // trips overlap
  size_t bi = 0;
  uint32_t at = NC_BCAP;
  if (d <= NC_NB) {
    bi = (size_t)tq * NC_NB + ((uint32_t)arr & (NC_NB - 1u));
    at = atomicAdd(&a.bcnt[bi], 1u);
  }
  // The record's previous arrival (it has arrived: the latch released) moves
  // to last_arr, which stands for lastSpikeArrival until the new one arrives.
  const nc_slot r = a.slots[ed.j];
  const int32_t prev = r.arr_inh >> 1;
  nc_slot n;
  n.w = r.w;
  n.depol = a.depol_k * r.w;
  n.last_arr = prev != NC_NEVER ? prev : r.last_arr;
  n.arr_inh = arr * 2 + (r.arr_inh & 1);
  a.slots[ed.j] = n;
  if (at < NC_BCAP) {
    a.bent[bi * NC_BCAP + at] =
        make_uint2(ed.rel | ((d - 1u) << 24) | ((uint32_t)(r.arr_inh & 1) << 31), __float_as_uint(r.w));
  } else {
    // the pending list of its window, at the target slice's slot range
    const uint32_t tbase = ed.j - ed.rel;
    const uint32_t plist = ((uint32_t)arr >> NC_PEND_LOG_W) & 1u;
    const uint32_t pt = atomicAdd(reinterpret_cast<uint32_t*>(&a.pend_cnt[tq]) + plist, 1u);
    (plist ? a.pend1 : a.pend0)[tbase + pt] = (ed.rel << 8) | ((uint32_t)arr & 255u);
  }
  a.edges[e].arr = arr;
  ++n_acc;
}

// The wave takes list lst, entries first, first + per, ... (64 per load), per
// = the waves per list.
__device__ __forceinline__ void nc_fanout_wave(const nc_step_args& a, uint32_t blk, uint32_t lst, uint32_t first,
                                               uint32_t per, int32_t ci) {
  const uint32_t lane = threadIdx.x & 63;
  uint32_t n_acc = 0, n_out = 0;
  const uint32_t nf = first < per ? __builtin_amdgcn_readfirstlane(
                                        a.fired_cnt[((((uint32_t)ci & 1u) * NC_FLISTS) + lst) * NC_FCNT_STRIDE])
                                  : 0u;
  const uint4* const flist = a.fired + lst * a.fcap;
  for (uint32_t fb = first; fb < nf; fb += 64 * per) {
    const uint32_t fi = fb + lane * per;
    const uint4 fl = fi < nf ? flist[fi] : make_uint4(0u, 0u, 0u, 0u);
    unsigned long long has = __ballot((fl.x | fl.y) != 0u);
    while (has) {
      const uint32_t fsrc = (uint32_t)__builtin_ctzll(has);
      has &= has - 1;
      const uint32_t fx = (uint32_t)__shfl((int)fl.x, (int)fsrc), fy = (uint32_t)__shfl((int)fl.y, (int)fsrc);
      const int32_t s = __shfl((int)fl.z, (int)fsrc);
      const uint32_t fw = (uint32_t)__shfl((int)fl.w, (int)fsrc);
      const uint32_t fq = fw & ((1u << NC_FSLICE_BITS) - 1u);
      const uint32_t item = (fw >> 24) & 15u, n = (fw >> 28) + 1u;
      if (fw & NC_FSINGLE) {
        // one fired lane: its out-edges are [fx, fx + fy); this record's
        // share is groups item, item + n, ...
        if (item == 0u) n_out += fy;
        for (uint32_t base = 64u * item; base < fy; base += 64u * n) {
          const uint32_t i = base + lane;
          if (i < fy) {
            const nc_edge ed = a.edges[fx + i];
            nc_fire_edge(a, fx + i, ed, s, n_acc);
          }
        }
        continue;
      }
      const unsigned long long m = ((unsigned long long)fy << 32) | fx;
      // The out-lists of all fired lanes, flattened: lane-parallel row loads,
      // a wave prefix sum of out-degrees, then this wave's 64-edge groups.
      uint2 row = make_uint2(0u, 0u);
      if ((m >> lane) & 1ull) row = a.out_row[fq * 64 + lane];
      uint32_t incl = row.y;
#pragma unroll
      for (int o = 1; o < 64; o <<= 1) {
        const uint32_t v = (uint32_t)__shfl_up((int)incl, o);
        if (lane >= (uint32_t)o) incl += v;
      }
      const uint32_t total = (uint32_t)__shfl((int)incl, 63);
      // this record's share: groups item, item + n, ... of the slice's n records
      if (item == 0u) n_out += total;
      for (uint32_t base = 64u * item; base < total; base += 64u * n) {
        const uint32_t i = base + lane;
        // source lane: the first lane whose inclusive prefix exceeds i
        uint32_t lo = 0;
#pragma unroll
        for (uint32_t b = 32; b >= 1; b >>= 1) {
          const uint32_t v = (uint32_t)__shfl((int)incl, (int)(lo + b - 1));
          if (v <= i) lo += b;
        }
        const uint32_t src = lo > 63 ? 63 : lo;
        const uint32_t rx = (uint32_t)__shfl((int)row.x, (int)src);
        const uint32_t ry = (uint32_t)__shfl((int)row.y, (int)src);
        const uint32_t ic = (uint32_t)__shfl((int)incl, (int)src);
        if (i < total) {
          const uint32_t ie = rx + (i - (ic - ry));
          const nc_edge ed = a.edges[ie];
          nc_fire_edge(a, ie, ed, s, n_acc);
        }
      }
    }
  }
  if (a.ctr) {
    if (lane == 0 && n_out) atomicAdd(&a.ctr[(size_t)blk * NC_NCTR + 6], (unsigned long long)n_out);
    nc_count32(a.ctr, blk, 7, n_acc);
  }
}

// The fan-out phase of one step: WPB waves per workgroup share the step's
// fired-slice records.
template <int WPB>
__global__ __launch_bounds__(64 * WPB) void k_fanout(nc_step_args a) {
  const uint32_t blk = nc_block_id(a.xcd_swizzle);
  const uint32_t per = gridDim.x * WPB / NC_FLISTS;
  uint32_t lst, first;
  if (a.xcd_lists) {
    // the lists of this workgroup's XCD (the grid holds the same number of
    // waves on every XCD, a multiple of 8)
    const uint32_t wx = (blockIdx.x >> 3) * WPB + (threadIdx.x >> 6);
    lst = (blockIdx.x & 7u) * 8u + (wx & 7u);
    first = wx >> 3;
  } else {
    const uint32_t wave = blk * WPB + (threadIdx.x >> 6);
    lst = wave % NC_FLISTS;
    first = wave / NC_FLISTS;
  }
  nc_fanout_wave(a, blk % a.ctr_rows, lst, first, per, (int32_t)*a.chunk_off + a.chunk_i);
}

// Grid-wide barrier of the persistent step kernel: each workgroup adds one
// to the counter and waits until it reaches `target` (wrap-safe, the counter
// runs on across the launches of a call).  The wait is bounded in time:
// past NC_BARRIER_TICKS of the 100 MHz real-time counter the workgroup flags
// the failure in *err and goes on, so that every launch ends; the host then
// reports an error (a grid whose workgroups were not all resident).
#define NC_BARRIER_TICKS 200000000ull  // 2 s
#define NC_BARRIER_SPINS (1u << 22)
__device__ __forceinline__ void nc_grid_barrier(uint32_t* bar, uint32_t target, uint32_t* err) {
  __syncthreads();
  if (threadIdx.x == 0) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    __hip_atomic_fetch_add(bar, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const uint64_t t0 = __builtin_amdgcn_s_memrealtime();
    uint32_t spins = 0;
    while ((int32_t)(__hip_atomic_load(bar, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) - target) < 0) {
      __builtin_amdgcn_s_sleep(2);
      // bounded by time and, should the clock not advance, by count
      if (__builtin_amdgcn_s_memrealtime() - t0 > NC_BARRIER_TICKS || ++spins > NC_BARRIER_SPINS) {
        __hip_atomic_store(err, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        break;
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  }
  __syncthreads();
}

// One whole step in one launch of a resident grid (the persistent form):
// every wave takes active slices w, w + waves, ... through the neuron phase,
// the grid meets at a barrier, then all waves share the fan-out.  The
// launch's end orders it before the next step.  bar[0] counts barrier
// arrivals in this call (zeroed at its start), bar[32] flags a timed-out
// wait.
template <int WPB>
__global__ __launch_bounds__(64 * WPB) void k_step(nc_step_args a, uint32_t* bar) {
  const uint32_t blk = blockIdx.x % a.ctr_rows;
  const uint32_t nwaves = gridDim.x * WPB;
  const uint32_t w0 = __builtin_amdgcn_readfirstlane(blockIdx.x * WPB + (threadIdx.x >> 6));
  const int32_t ci = (int32_t)*a.chunk_off + a.chunk_i;
  const float* tpre = nc_trace_tables<WPB>(a);
  // every wave of the grid runs the same number of slice passes
  for (uint32_t b = 0; b < a.n_active_slices; b += nwaves) nc_neuron_slice<WPB>(a, blk, b + w0, ci, tpre);
  nc_grid_barrier(bar, ((uint32_t)ci + 1u) * gridDim.x, bar + 32);
  nc_fanout_wave(a, blk, w0 % NC_FLISTS, w0 / NC_FLISTS, nwaves / NC_FLISTS, ci);
}
