// Device kernels of one simulation step (gfx950, wave64).
//
// One step s of every active network is, in order:
//   k_drive   host-scheduled forced fires -> bitmasks       (only when present)
//   k_neuron  one lane per neuron: charge_insynapses, charge_passive,
//             threshold / forced fire, deferred arrival STDP, AP waveform,
//             latch release; ballot-compacts the fired neurons
//   k_post    one wave per fired neuron: post-side STDP over its incoming
//             synapses (the inSynapses loop of Neuron::fire)
//   k_fanout  one wave per fired neuron: Synapse::fire on its outgoing
//             synapses through the out->in slot map, latch permitting
// Kernel boundaries order the phases; no atomics touch weights.
#pragma once
#include <hip/hip_runtime.h>

#include "nc_common.h"

#define NC_DEG_MASK 0x00FFFFFFu
#define NC_DEG_INPUT 0x01000000u
#define NC_DEG_VALID 0x02000000u
#define NC_SLICE_NONE 0xFFFFFFFFu
#define NC_NCTR 8
#define NC_CTR_ROWS 64

// Per-network descriptor, rewritten by the host before each batch of steps.
struct nc_netdesc {
  int32_t base_step;  // steps completed before this batch
  int32_t remaining;  // steps this network runs in this batch
  float lr, pre_f, post_f;
  uint32_t drive_lo, drive_hi;  // drive seed
  uint32_t bg_thr;              // background threshold on a u32 draw (0 = off)
  uint32_t bg_lo, bg_hi;        // background seed
  uint32_t pad[6];
};

struct nc_group {
  uint32_t tbl_off;  // first threshold entry
  uint32_t period;   // entries
  uint32_t local;    // group index inside its network (Philox counter word)
  uint32_t pad;
};

struct nc_step_args {
  uint8_t* stamp;
  nc_slot* slots;
  float* V;
  int32_t* last_fire;
  uint32_t* firings;
  const uint32_t* degw;
  const uint32_t* nid;
  const uint32_t* in_group;
  const uint32_t* slice_off;
  const uint32_t* slice_net;
  const uint32_t* active_slices;
  uint32_t n_active_slices;
  const nc_netdesc* nets;
  const nc_group* groups;
  const uint32_t* thr;
  uint32_t* forced_in;
  uint32_t* forced_bg;
  int32_t use_forced;
  uint32_t* fired;      // this step's list
  uint32_t* fired_cnt;  // this step's count
  uint32_t* fired_cnt_next;
  uint32_t* raster;  // (pos, step) pairs
  uint32_t* raster_cnt;
  uint32_t raster_cap;
  unsigned long long* ctr;  // [NC_CTR_ROWS][NC_NCTR]
  const float* t_pre;
  const float* t_post;
  const float* w_ap;
  int32_t len_pre, len_post;
  float dt, P, depol_k;
  double c0, e;
  int32_t d_ref, r_rel;
  int32_t chunk_i;
  const uint32_t* out_beg;
  const uint32_t* out_cnt;
  const uint32_t* out_slot;
  const uint8_t* out_delay;
  const uint32_t* ev_pos;  // host events of this step
  const uint8_t* ev_kind;
  uint32_t n_ev;
};

__device__ __forceinline__ float nc_lookup(const float* t, int32_t len, int32_t d) {
  return (d >= 0 && d < len) ? t[d] : 0.0f;
}

// fmax(fmin(w, hi), lo) as the reference evaluates it on x86-64
// (src/NeuCor.cpp:760-761); exact in float because the bounds are.
__device__ __forceinline__ float nc_clamp_w(float w, uint32_t flags) {
  const float hi = (flags & 1u) ? 0.0f : 1.0f;
  const float lo = (flags & 1u) ? -1.0f : 0.0f;
  float m = (w != w) ? hi : (w < hi ? w : hi);
  return m > lo ? m : lo;
}

// Synapse::synapticPlasticity (src/NeuCor.cpp:740-764) with both traces given.
__device__ __forceinline__ float nc_stdp(float w, uint32_t flags, float tr_s, float tr_t,
                                         float pre_f, float post_f, float lr) {
  float dw = pre_f * tr_s - post_f * tr_t;
  float wn = w + dw * lr;
  return nc_clamp_w(wn, flags);
}

__device__ __forceinline__ uint32_t nc_lane() {
  return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
}

__device__ __forceinline__ unsigned long long nc_wave_sum(unsigned long long v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
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

__device__ __forceinline__ void nc_count(unsigned long long* ctr, int k, unsigned long long v) {
  v = nc_wave_sum(v);
  if (nc_lane() == 0 && v) atomicAdd(&ctr[(blockIdx.x & (NC_CTR_ROWS - 1)) * NC_NCTR + k], v);
}

// Host-scheduled forced fires of this step -> bitmasks read by k_neuron.
__global__ void k_drive(nc_step_args a) {
  uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= a.n_ev) return;
  uint32_t pos = a.ev_pos[t];
  uint32_t* m = a.ev_kind[t] ? a.forced_bg : a.forced_in;
  atomicOr(&m[pos >> 5], 1u << (pos & 31));
}

// Neuron::run (src/NeuCor.cpp:619-641) for every neuron of the active slices.
__global__ __launch_bounds__(256) void k_neuron(nc_step_args a) {
  const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  if (wave >= a.n_active_slices) return;
  const uint32_t lane = threadIdx.x & 63;
  const uint32_t q = a.active_slices[wave];
  const uint32_t net = a.slice_net[q];
  const nc_netdesc* nd = &a.nets[net];
  const int32_t remaining = nd->remaining;
  if (a.chunk_i >= remaining) return;
  const int32_t s = nd->base_step + a.chunk_i + 1;
  const uint32_t pos = q * 64 + lane;
  const uint32_t dw = a.degw[pos];
  const bool valid = (dw & NC_DEG_VALID) != 0;
  const uint32_t deg = dw & NC_DEG_MASK;
  const float lr = nd->lr, pre_f = nd->pre_f, post_f = nd->post_f;

  float V = a.V[pos];
  int32_t lf = a.last_fire[pos];
  const int32_t smod = s % 255;
  const uint32_t sbase = a.slice_off[q] + lane * 4;
  const uint32_t nch = (deg + 3) >> 2;
  uint32_t n_active = 0, n_arr = 0;

  // charge_insynapses (src/NeuCor.cpp:688-700): ascending source order.
  for (uint32_t c0 = 0; c0 < nch; c0 += 4) {
    uint32_t wv[4];
#pragma unroll
    for (int u = 0; u < 4; ++u)
      wv[u] = (c0 + u < nch) ? *(const uint32_t*)(a.stamp + sbase + (c0 + u) * 256) : 0u;
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      uint32_t w4 = wv[u];
      if (w4 == 0u) continue;
      const uint32_t j0 = sbase + (c0 + u) * 256;
#pragma unroll
      for (int b = 0; b < 4; ++b) {
        const uint32_t code = (w4 >> (8 * b)) & 0xffu;
        if (code == 0u) continue;
        const int32_t o = nc_stamp_offset(smod, code);
        if (o == 0 || o > a.r_rel) continue;  // arriving now, or still in flight
        const uint32_t j = j0 + b;
        nc_slot sl = a.slots[j];
        const float inc = a.dt * sl.depol;
        V = (float)((double)V + ((double)inc * a.c0) * a.e);
        ++n_active;
        if (o == 1 && sl.last_arr != s - 1) {
          // The arrival of step s-1 (Synapse::run, src/NeuCor.cpp:718-726):
          // lastSpikeArrival = s-1, then STDP with the synapse trace at 1 -> 0
          // and the target trace of lastFire as it stood after step s-1.
          const int32_t ar = s - 1;
          const float tr_t = nc_lookup(a.t_post, a.len_post, ar - lf);
          const float wn = nc_stdp(sl.w, sl.flags, 0.0f, tr_t, pre_f, post_f, lr);
          a.slots[j].w = wn;
          a.slots[j].last_arr = ar;
          ++n_arr;
        }
        if (o >= a.r_rel) a.stamp[j] = 0;  // latch release (:697)
      }
    }
  }

  // charge_passive (src/NeuCor.cpp:677-680).
  V = (V - NC_BASELEVEL) * a.P + NC_BASELEVEL;

  // Forced fires: device drive, background, host events.
  bool f_in = false, f_bg = false;
  if (valid && (dw & NC_DEG_INPUT)) {
    const nc_group g = a.groups[a.in_group[pos]];
    const uint64_t seed = ((uint64_t)nd->drive_hi << 32) | nd->drive_lo;
    const uint32_t u = nc_draw(seed, (uint32_t)s, g.local, NC_STREAM_DRIVE).v[0];
    f_in = u < a.thr[g.tbl_off + (uint32_t)s % g.period];
  }
  if (valid && nd->bg_thr) {
    const uint64_t seed = ((uint64_t)nd->bg_hi << 32) | nd->bg_lo;
    const uint32_t u = nc_draw(seed, (uint32_t)s, a.nid[pos], NC_STREAM_BG).v[0];
    f_bg = u < nd->bg_thr;
  }
  if (a.use_forced) {
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
  uint32_t fcount = 0;
  if (fire) {
    lf = s;
    fcount = a.firings[pos] + 1u;
    a.firings[pos] = fcount;
  }
  // AP waveform (src/NeuCor.cpp:706-714) from the host table.
  const int32_t dlf = s - lf;
  if (dlf <= a.d_ref) V = a.w_ap[dlf];

  if (valid) {
    a.V[pos] = V;
    if (fire) a.last_fire[pos] = lf;
  }
  const uint32_t at = nc_wave_append(fire, a.fired_cnt);
  if (fire) a.fired[at] = pos;
  if (a.raster_cap) {
    const uint32_t rt = nc_wave_append(fire, a.raster_cnt);
    if (fire && rt < a.raster_cap) {
      a.raster[2ull * rt] = pos;
      a.raster[2ull * rt + 1] = (uint32_t)s;
    }
  }
  if (a.ctr) {
    nc_count(a.ctr, 0, valid ? 1ull : 0ull);
    nc_count(a.ctr, 1, valid ? deg : 0ull);
    nc_count(a.ctr, 2, n_active);
    nc_count(a.ctr, 3, n_arr);
    nc_count(a.ctr, 4, fire ? 1ull : 0ull);
    nc_count(a.ctr, 5, fire ? deg : 0ull);
  }
}

// Post-side STDP of every fired neuron over its incoming synapses: the
// inSynapses loop of Neuron::fire (src/NeuCor.cpp:651-653) with the target
// trace at powf(decay, 0) == 1 -> 0 and the synapse trace of lastSpikeArrival
// from before this step's arrivals.
__global__ __launch_bounds__(256) void k_post(nc_step_args a) {
  const uint32_t nf = *a.fired_cnt;
  const uint32_t lane = threadIdx.x & 63;
  const uint32_t nw = (gridDim.x * blockDim.x) >> 6;
  for (uint32_t f = (blockIdx.x * blockDim.x + threadIdx.x) >> 6; f < nf; f += nw) {
    const uint32_t pos = a.fired[f];
    const uint32_t q = pos >> 6, l = pos & 63;
    const nc_netdesc* nd = &a.nets[a.slice_net[q]];
    const int32_t s = nd->base_step + a.chunk_i + 1;
    const float lr = nd->lr, pre_f = nd->pre_f, post_f = nd->post_f;
    const uint32_t deg = a.degw[pos] & NC_DEG_MASK;
    const uint32_t base = a.slice_off[q] + l * 4;
    for (uint32_t k = lane; k < deg; k += 64) {
      const uint32_t j = base + (k >> 2) * 256 + (k & 3);
      const nc_slot sl = a.slots[j];
      const float tr_s = nc_lookup(a.t_pre, a.len_pre, s - sl.last_arr);
      a.slots[j].w = nc_stdp(sl.w, sl.flags, tr_s, 0.0f, pre_f, post_f, lr);
    }
  }
}

// Synapse::fire (src/NeuCor.cpp:727-738) on every outgoing synapse of every
// fired neuron: one writer per synapse (its source), latch permitting.
__global__ __launch_bounds__(256) void k_fanout(nc_step_args a) {
  const uint32_t nf = *a.fired_cnt;
  const uint32_t lane = threadIdx.x & 63;
  const uint32_t nw = (gridDim.x * blockDim.x) >> 6;
  unsigned long long n_out = 0, n_acc = 0;
  for (uint32_t f = (blockIdx.x * blockDim.x + threadIdx.x) >> 6; f < nf; f += nw) {
    const uint32_t pos = a.fired[f];
    const nc_netdesc* nd = &a.nets[a.slice_net[pos >> 6]];
    const int32_t s = nd->base_step + a.chunk_i + 1;
    const uint32_t beg = a.out_beg[pos], cnt = a.out_cnt[pos];
    n_out += (lane < cnt) ? ((cnt - lane + 63) >> 6) : 0;
    for (uint32_t k = lane; k < cnt; k += 64) {
      const uint32_t j = a.out_slot[beg + k];
      if (a.stamp[j] != 0) continue;  // a spike is already in flight (:728)
      const float w = a.slots[j].w;
      a.slots[j].depol = a.depol_k * w;
      a.stamp[j] = (uint8_t)nc_stamp_code(s + (int32_t)a.out_delay[beg + k]);
      ++n_acc;
    }
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) *a.fired_cnt_next = 0u;
  if (a.ctr) {
    nc_count(a.ctr, 6, n_out);
    nc_count(a.ctr, 7, n_acc);
  }
}

// Applies the STDP of arrivals at step s_done that k_neuron defers to the
// next step, so that host reads and parameter changes see eager state.
__global__ __launch_bounds__(256) void k_flush(nc_step_args a, uint32_t q0, uint32_t nq,
                                               int32_t s_done, float lr, float pre_f, float post_f) {
  const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  if (wave >= nq) return;
  const uint32_t lane = threadIdx.x & 63;
  const uint32_t q = q0 + wave;
  const uint32_t pos = q * 64 + lane;
  const uint32_t deg = a.degw[pos] & NC_DEG_MASK;
  const int32_t lf = a.last_fire[pos];
  const int32_t smod = s_done % 255;
  const uint32_t sbase = a.slice_off[q] + lane * 4;
  for (uint32_t k = 0; k < deg; ++k) {
    const uint32_t j = sbase + (k >> 2) * 256 + (k & 3);
    const uint32_t code = a.stamp[j];
    if (code == 0u || nc_stamp_offset(smod, code) != 0) continue;
    const nc_slot sl = a.slots[j];
    if (sl.last_arr == s_done) continue;
    const float tr_t = nc_lookup(a.t_post, a.len_post, s_done - lf);
    a.slots[j].w = nc_stdp(sl.w, sl.flags, 0.0f, tr_t, pre_f, post_f, lr);
    a.slots[j].last_arr = s_done;
  }
}
