// Host side of the MI355X engine: device-resident pools of networks, their
// SELL-C-64 layout, the step tables and the C ABI of include/neucor_amd.h.
//
// Layout of one network in HBM (all structure-of-arrays):
//   neurons are renumbered into "positions": windows of 1024 ids are sorted
//   by in-degree (SELL-C-sigma, C = 64, sigma = 1024) and cut into slices of
//   64 lanes, one wave per slice.  The incoming synapses of a slice are
//   stored chunk-major: chunk c of lane l holds the lane's slots 8c..8c+7 at
//   slot index slice_off + c*512 + l*8, so a fired lane's row is read by the
//   wave as 8 full 128-byte lines per 64 slots.  Slot order inside a lane
//   is ascending source id.
//   slots  : 16 B per slot {w, depol, last_arr, 2*arrival + inhibitory}
//   amask  : 1 bit per slot, set while the slot's spike is in its release
//            window; words [w][lane] per slice, so a wave reads 256
//            contiguous bytes per 32 slots of each lane
//   pend0/1: per slice, the spikes still in flight, in two lists by
//            arrival window ((slot in slice) << 8 | arrival & 255 each)
//   out-CSR: per position a row of slot indices, u8 delays, target slices
//            and the source-side copy of each synapse's latch, for the
//            fan-out.
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>
#include <math.h>
#include <string.h>

#include <algorithm>
#include <mutex>
#include <map>
#include <numeric>
#include <string>
#include <vector>

#include "nc_common.h"
#include "nc_internal.h"
#include "nc_kernels.h"

using nc::fail;

// Device scratch of nc_read_weights_noalloc: a file-scope device array, one
// copy per GPU, held until the process ends.  Host code never passes its name
// to a kernel (on the host the name is only a stand-in): nc_sim_create looks
// up each device's address once, with that device current, and the entry
// point uses that address.  Every call on one GPU shares the array.
#define NC_SCRATCH_FLOATS (1u << 24)  // 64 MiB: networks of up to 16,777,216 synapses
static __device__ __align__(256) float scratch[NC_SCRATCH_FLOATS];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
float* g_scratch_ptr[kMaxDevices] = {};  // device address of `scratch`, per device

// With `device` current: its copy's address, looked up once.
nc_status scratch_setup(int device) {
  if (device < 0 || device >= kMaxDevices) return fail(NC_ERR_RANGE, "device index beyond the scratch table");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[device]) return NC_OK;
  void* ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !ptr)
    return fail(NC_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
  g_scratch_ptr[device] = static_cast<float*>(ptr);
  return NC_OK;
}
}  // namespace

namespace {

#define NC_HIP(call)                                                              \
  do {                                                                            \
    hipError_t e_ = (call);                                                       \
    if (e_ != hipSuccess)                                                         \
      return fail(e_ == hipErrorOutOfMemory ? NC_ERR_OOM : NC_ERR_HIP,            \
                  std::string(#call) + ": " + hipGetErrorString(e_));             \
  } while (0)

constexpr uint32_t kSigma = 1024;

// First-fit allocator over [0, cap) in units.
struct RangeAlloc {
  uint64_t cap = 0;
  std::map<uint64_t, uint64_t> free_;  // offset -> length
  void init(uint64_t c) {
    cap = c;
    free_.clear();
    if (c) free_[0] = c;
  }
  bool alloc(uint64_t n, uint64_t* off) {
    if (n == 0) { *off = 0; return true; }
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second >= n) {
        *off = it->first;
        uint64_t rest = it->second - n, o = it->first + n;
        free_.erase(it);
        if (rest) free_[o] = rest;
        return true;
      }
    }
    return false;
  }
  void release(uint64_t off, uint64_t n) {
    if (n == 0) return;
    auto it = free_.emplace(off, n).first;
    auto nx = std::next(it);
    if (nx != free_.end() && it->first + it->second == nx->first) {
      it->second += nx->second;
      free_.erase(nx);
    }
    if (it != free_.begin()) {
      auto pv = std::prev(it);
      if (pv->first + pv->second == it->first) {
        pv->second += it->second;
        free_.erase(it);
      }
    }
  }
  uint64_t free_units() const {
    uint64_t f = 0;
    for (auto& kv : free_) f += kv.second;
    return f;
  }
  // More room at the end (merged with a free tail).
  void extend(uint64_t c) {
    if (c <= cap) return;
    const uint64_t old = cap;
    cap = c;
    release(old, c - old);
  }
};

struct Layout {
  uint32_t n = 0, nq = 0;
  std::vector<uint32_t> pos_of_id, id_of_pos;  // id_of_pos: nq*64 (~0u = padding)
  std::vector<uint32_t> slice_slots;           // slots per slice (multiple of 64 * NC_CHUNK)
  std::vector<uint32_t> slice_mwords;          // active-mask words per slice (multiple of 64)
  uint64_t nslots = 0, nmwords = 0;
};

void make_layout(const nc_net& net, Layout* L) {
  const uint32_t n = net.n;
  L->n = n;
  L->nq = (n + 63) / 64;
  L->pos_of_id.resize(n);
  L->id_of_pos.assign((size_t)L->nq * 64, ~0u);
  std::vector<uint32_t> ids(n);
  std::iota(ids.begin(), ids.end(), 0u);
  for (uint32_t w0 = 0; w0 < n; w0 += kSigma) {
    uint32_t w1 = std::min(n, w0 + kSigma);
    std::stable_sort(ids.begin() + w0, ids.begin() + w1, [&](uint32_t a, uint32_t b) {
      return (net.in_ptr[a + 1] - net.in_ptr[a]) > (net.in_ptr[b + 1] - net.in_ptr[b]);
    });
  }
  for (uint32_t p = 0; p < n; ++p) {
    L->id_of_pos[p] = ids[p];
    L->pos_of_id[ids[p]] = p;
  }
  L->slice_slots.resize(L->nq);
  L->nslots = 0;
  for (uint32_t q = 0; q < L->nq; ++q) {
    uint32_t mx = 0;
    for (uint32_t l = 0; l < 64; ++l) {
      uint32_t id = L->id_of_pos[q * 64 + l];
      if (id == ~0u) continue;
      uint32_t d = (uint32_t)(net.in_ptr[id + 1] - net.in_ptr[id]);
      mx = std::max(mx, d);
    }
    uint32_t nch = (mx + NC_CHUNK - 1) / NC_CHUNK;
    L->slice_slots[q] = nch * 64 * NC_CHUNK;
    L->nslots += L->slice_slots[q];
    L->slice_mwords.push_back(((mx + 31) / 32) * 64);
    L->nmwords += L->slice_mwords.back();
  }
}

}  // namespace

namespace nc {
// What a simulation keeps of a network's layout: the maps between the
// network's export order and its device positions, slots and out-edges, all
// network-local (a simulation adds its own first position, slot, edge).
struct NetMaps {
  std::vector<uint32_t> pos_of_id;    // local position per neuron id
  std::vector<uint32_t> id_of_pos;    // neuron id per local position (nq*64, ~0u: padding)
  std::vector<uint32_t> slot_of_syn;  // local slot per synapse, export order
  std::vector<uint32_t> edge_of_syn;  // local out-edge per synapse, export order
  std::vector<int64_t> in_ptr;        // incoming rows (export order)
  uint64_t digest = 0;                // identity of the wiring (checkpoints are checked against it)
};

// A network's device image in network-local coordinates: every array
// nc_sim_add uploads, with slot, slice and edge indices relative to the
// network's own first slot, slice and out-edge.  nc_sim_add relocates it
// (host offsets for the per-neuron rows, k_relocate_edges for the edges).
struct NetImage {
  uint32_t nq = 0;
  uint64_t S = 0, nslots = 0, nmwords = 0, mwords_step = 0;
  std::vector<uint32_t> slice_off, slice_moff;  // per slice: first slot, first mask word
  std::vector<float> slot_w;                    // nslots weights (0 on padding slots)
  std::vector<uint32_t> slot_inh;               // one sign bit per slot (1: inhibitory class)
  std::vector<uint2> edge_jq;                   // S out-edges by source position, rows in ascending target
                                                // id: (target slot, target slice << 8 | delay)
  std::vector<uint2> out_row;                   // per position: (first out-edge, count)
  std::vector<uint32_t> degw;                   // per position: in-degree | NC_DEG_VALID (0: padding)
  std::vector<uint32_t> nid;                    // per position: neuron id (~0u: padding)
  std::shared_ptr<const NetMaps> maps;
};
}  // namespace nc

namespace {
using nc::NetImage;
using nc::NetMaps;

// FNV-1a over 64-bit words (then the tail bytes).
void fnv_mix(uint64_t* h, const void* p, size_t n) {
  const unsigned char* c = (const unsigned char*)p;
  uint64_t x = *h;
  size_t i = 0;
  for (; i + 8 <= n; i += 8) {
    uint64_t w;
    memcpy(&w, c + i, 8);
    x = (x ^ w) * 1099511628211ull;
  }
  for (; i < n; ++i) x = (x ^ c[i]) * 1099511628211ull;
  *h = x;
}

std::shared_ptr<const NetImage> build_image(const nc_net& net) {
  Layout L;
  make_layout(net, &L);
  auto maps = std::make_shared<NetMaps>();
  auto img = std::make_shared<NetImage>();
  NetImage& I = *img;
  NetMaps& M = *maps;
  const uint64_t S = net.n_syn();
  I.nq = L.nq;
  I.S = S;
  I.nslots = L.nslots;
  I.nmwords = L.nmwords;
  const uint32_t npos = L.nq * 64;
  I.slice_off.resize(L.nq);
  I.slice_moff.resize(L.nq);
  {
    uint64_t o = 0, mo = 0;
    for (uint32_t q = 0; q < L.nq; ++q) {
      I.slice_off[q] = (uint32_t)o;
      I.slice_moff[q] = (uint32_t)mo;
      o += L.slice_slots[q];
      mo += L.slice_mwords[q];
    }
  }
  uint64_t h = 1469598103934665603ull;
  fnv_mix(&h, net.in_ptr.data(), net.in_ptr.size() * 8);
  fnv_mix(&h, net.in_src.data(), net.in_src.size() * 4);
  fnv_mix(&h, net.delay.data(), net.delay.size());
  fnv_mix(&h, net.inh.data(), net.inh.size());
  M.digest = h;
  M.in_ptr = net.in_ptr;
  // Slot records, in layout order.
  I.slot_w.assign(L.nslots, 0.0f);
  I.slot_inh.assign((L.nslots + 31) / 32, 0u);
  M.slot_of_syn.resize(S);
  M.edge_of_syn.assign(S, 0u);
  for (uint32_t id = 0; id < net.n; ++id) {
    const uint32_t pl = L.pos_of_id[id];
    const uint32_t q = pl / 64, l = pl % 64;
    const int64_t b = net.in_ptr[id], e = net.in_ptr[id + 1];
    for (int64_t k = 0; k < e - b; ++k) {
      const uint32_t j = nc_slot_of(I.slice_off[q], l, (uint32_t)k);
      M.slot_of_syn[b + k] = j;
      I.slot_w[j] = net.w[b + k];
      if (net.inh[b + k]) I.slot_inh[j >> 5] |= 1u << (j & 31);
    }
  }
  // Out-CSR by source position; rows in ascending target id.
  std::vector<uint32_t> outcnt(npos, 0), outbeg(npos, 0);
  for (uint64_t k = 0; k < S; ++k) outcnt[L.pos_of_id[net.in_src[k]]]++;
  {
    uint64_t o = 0;
    for (uint32_t pl = 0; pl < npos; ++pl) {
      outbeg[pl] = (uint32_t)o;
      o += outcnt[pl];
    }
  }
  I.edge_jq.resize(S);
  {
    std::vector<uint32_t> fillp(outbeg);
    for (uint32_t tg = 0; tg < net.n; ++tg) {
      const uint32_t tq = L.pos_of_id[tg] / 64;
      for (int64_t k = net.in_ptr[tg]; k < net.in_ptr[tg + 1]; ++k) {
        const uint32_t e = fillp[L.pos_of_id[net.in_src[k]]]++;
        I.edge_jq[e] = make_uint2(M.slot_of_syn[k], (tq << 8) | net.delay[k]);
        M.edge_of_syn[k] = e;
      }
    }
  }
  I.out_row.resize(npos);
  I.degw.assign(npos, 0u);
  I.nid.assign(npos, ~0u);
  for (uint32_t pl = 0; pl < npos; ++pl) {
    I.out_row[pl] = make_uint2(outbeg[pl], outcnt[pl]);
    const uint32_t idl = L.id_of_pos[pl];
    if (idl == ~0u) continue;
    const uint32_t dg = (uint32_t)(net.in_ptr[idl + 1] - net.in_ptr[idl]);
    I.degw[pl] = dg | NC_DEG_VALID;
    I.mwords_step += (dg + 31) / 32;
    I.nid[pl] = idl;
  }
  M.pos_of_id = std::move(L.pos_of_id);
  M.id_of_pos = std::move(L.id_of_pos);
  I.maps = maps;
  return img;
}

struct Tables {
  float dt = 0.f, pre_decay = 0.f, post_decay = 0.f;
  std::vector<float> t_pre, t_post, w_ap;
  int32_t d_ref = 0, r_rel = 0;
  float P = 0.f, depol_k = 0.f;
  double e = 0.0;
};

// Every transcendental the reference evaluates in a step, as a function of
// an integer step difference d, evaluated with this host's libm exactly as
// the reference writes it.
void build_tables(float dt, float pre_decay, float post_decay, Tables* T) {
  T->dt = dt;
  T->pre_decay = pre_decay;
  T->post_decay = post_decay;
  // Synapse trace powf(presynapticTraceDecay, t - lastSpikeArrival) and the
  // neuron trace powf(postsynapticTraceDecay, t - lastFire), with the
  // "== 1 -> 0" rule (src/NeuCor.cpp:741-745, :671-675).
  auto trace_table = [dt](float decay, std::vector<float>* t) {
    t->clear();
    for (int32_t d = 0;; ++d) {
      float x = (float)d * dt;
      float v = powf(decay, x);
      if (v == 1.0f) v = 0.0f;
      if (v == 0.0f && d > 0) break;
      t->push_back(v);
      if (d > 100000000) break;
    }
  };
  trace_table(pre_decay, &T->t_pre);
  trace_table(post_decay, &T->t_post);
  // Refractory / AP window: AP_cutoff < t - lastFire (src/NeuCor.cpp:684, :707).
  int32_t dref = 0;
  while (!(NC_AP_CUTOFF < (float)(dref + 1) * dt)) ++dref;
  T->d_ref = dref;
  // Latch release: the contribution at offset o is added, then the latch
  // clears when AP_cutoff < o*dt (src/NeuCor.cpp:692-697).
  T->r_rel = dref + 1;
  // AP waveform (src/NeuCor.cpp:706-714) with lastFire = 0, currentT = d*dt.
  T->w_ap.resize(dref + 1);
  for (int32_t d = 0; d <= dref; ++d) {
    const float lastFire = 0.0f;
    const float currentT = (float)d * dt;
    float currentAP =
        NC_AP_H * (exp(-powf((currentT - lastFire) - NC_AP_DSTART, 2.0) /
                       (2.0 * NC_AP_DEPOLW * NC_AP_DEPOLW)) -
                   exp(-powf((currentT - lastFire) - NC_AP_DSTART - NC_AP_DELTAPOL, 2.0) /
                       (2.0 * NC_AP_POLW * NC_AP_POLW)) *
                       NC_AP_DEPOLFAC) +
        NC_BASELEVEL + (NC_THRESHOLD - NC_BASELEVEL) * fmax(1.0 + lastFire - currentT, 0.0);
    T->w_ap[d] = currentAP;
  }
  T->P = powf(NC_RECHARGE, dt);         // powf(recharge, deltaT) (:678)
  T->e = exp(NC_CHARGE_R * (double)dt);  // exp(0.3702*deltaT) (:695)
  float dfac = NC_AP_DEPOLFAC;
  dfac *= NC_SYN_DEPOL_MUL;  // AP_depolFac *= 52.0 (:730)
  T->depol_k = dfac;
}

struct NetRec {
  bool live = false;
  uint32_t n = 0;
  uint64_t S = 0;
  uint32_t q0 = 0, nq = 0;
  uint64_t slot0 = 0, nslots = 0, out0 = 0, mask0 = 0, nmwords = 0;
  uint32_t act0 = 0;  // first record of this network in the active-slice list
  bool act_valid = false;    // the device active-slice list holds this network's records at act0
  uint64_t mwords_step = 0;  // active-mask words k_neuron reads per step
  std::shared_ptr<const NetMaps> maps;  // network-local layout maps
  uint32_t pos(uint32_t id) const { return q0 * 64 + maps->pos_of_id[id]; }
  float lr = 1.f, pre_f = 0.13f, post_f = 0.30f;
  uint32_t steps_done = 0, steps_left = UINT32_MAX;
  // first step after the last host weight write: spikes fired before it
  // carry a stale weight in their arrival-bucket entries
  uint32_t whost = 0;
  uint32_t grp0 = 0, ngrp = 0, thr0 = 0, nthr = 0;
  nc_drive_spec drive{};
  bool has_drive = false;
  std::vector<std::pair<uint32_t, std::pair<uint32_t, uint8_t>>> events;  // abs step, (pos, kind)
};

template <class T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
};

}  // namespace

struct nc_sim {
  int device = 0;
  hipStream_t stream = nullptr;
  uint64_t max_pos = 0, max_slots = 0, max_out = 0;
  uint32_t max_nets = 0, raster_cap = 0;
  RangeAlloc a_slices, a_slots, a_out, a_masks;
  uint64_t max_mwords = 0;
  std::vector<NetRec> nets;
  bool tables_set = false;
  Tables T;
  uint64_t gstep = 0;
  int kernel_timing = 0;  // bit mask of timed kernels
  uint32_t timing_every = 1;
  uint64_t timing_step = 0;
  uint64_t k_launch[4] = {0, 0, 0, 0};
  double k_ms[4] = {0, 0, 0, 0};
  std::vector<hipEvent_t> ev_pool;

  // device buffers
  nc_slot* slots = nullptr;
  uint32_t* amask = nullptr;
  uint32_t* asum = nullptr;
  float* V = nullptr;
  int32_t* last_fire = nullptr;
  uint32_t *firings = nullptr, *nid = nullptr;
  uint2* ndw = nullptr;
  uint2* out_row = nullptr;
  nc_edge* edges = nullptr;
  uint32_t *pend0 = nullptr, *pend1 = nullptr;
  uint2* pend_cnt = nullptr;
  uint32_t* bcnt = nullptr;        // per slice: NC_NB arrival-bucket counts
  uint2* bent = nullptr;           // per slice: NC_NB buckets of NC_BCAP entries
  uint4* acache = nullptr;         // active-slot cache: one 16-byte entry per position
  uint32_t *id_pos = nullptr, *id_row = nullptr;  // per network, at its positions: neuron id -> local position, first synapse
  uint32_t* slice_moff = nullptr;  // per slice: first active-mask word
  uint32_t *slice_off = nullptr, *slice_net = nullptr;
  nc_slice_rec* active = nullptr;
  uint32_t *forced_in = nullptr, *forced_bg = nullptr;
  uint4* fired = nullptr;
  uint32_t* fired_cnt = nullptr;

  uint64_t ctr_rows = 0;
  uint32_t *raster = nullptr, *raster_cnt = nullptr;
  unsigned long long* ctr = nullptr;
  bool counting = false;
  uint32_t* d_chunk_off = nullptr;  // steps of the current call before the running chunk
  nc_launch_cfg cfg{4, 0, 0, 0, 0, 0};
  int resident[9] = {0};  // persistent form: workgroups of 1..8 waves the device holds at once (0 = not known)
  // captured chunk of cfg.graph_steps steps and the launch it was captured for
  hipGraphExec_t graph = nullptr;
  nc_step_args graph_args{};
  uint32_t graph_key[5] = {0, 0, 0, 0, 0};
  nc_netdesc* d_nets = nullptr;
  float *t_pre = nullptr, *t_post = nullptr, *w_ap = nullptr;
  DevBuf<nc_group> d_groups;
  DevBuf<uint32_t> d_gfire;
  DevBuf<float> d_wbuf;       // packed weights of one network, for host reads
  DevBuf<uint32_t> d_stage;   // a network's compact upload, expanded in place by the placing kernels
  DevBuf<uint32_t> d_thr;
  DevBuf<uint32_t> d_ev_pos;
  DevBuf<uint8_t> d_ev_kind;

  // Drive group table and threshold tables: ranges per network, reused after
  // removal (freed groups point at the never-live network max_nets).
  std::vector<nc_group> groups;
  std::vector<uint32_t> thr;
  RangeAlloc a_groups, a_thr;
  std::vector<nc_slice_rec> active_host;
  std::vector<uint32_t> slice_off_host, slice_moff_host;
  bool active_dirty = true;
  std::vector<uint32_t> slice_net_host;
  uint64_t raster_read = 0;
  bool act_identity = false;  // active record i is slice act_q0 + i
  uint32_t act_q0 = 0;
  uint64_t host_neuron_steps = 0, host_synapse_steps = 0, host_mword_steps = 0;
  // Host staging of the per-call uploads; reused only after upload_done.
  std::vector<nc_netdesc> h_nds;
  std::vector<uint32_t> h_ev_pos;
  std::vector<uint8_t> h_ev_kind;
  hipEvent_t upload_done = nullptr;
  bool upload_pending = false;
};

namespace {

nc_status check_net(nc_sim* sim, int32_t id, NetRec** out) {
  if (!sim) return fail(NC_ERR_ARG, "null sim");
  if (id < 0 || (size_t)id >= sim->nets.size() || !sim->nets[id].live)
    return fail(NC_ERR_RANGE, "no such network id");
  *out = &sim->nets[id];
  return NC_OK;
}

template <class T>
nc_status grow(DevBuf<T>* b, size_t n, hipStream_t st) {
  if (n <= b->n && b->p) return NC_OK;
  size_t cap = std::max<size_t>(n, b->n * 2);
  if (cap < 64) cap = 64;
  NC_HIP(hipStreamSynchronize(st));
  if (b->p) NC_HIP(hipFree(b->p));
  b->p = nullptr;
  b->n = 0;
  NC_HIP(hipMalloc(&b->p, cap * sizeof(T)));
  b->n = cap;
  return NC_OK;
}

nc_step_args base_args(nc_sim* sim) {
  nc_step_args a;
  memset(&a, 0, sizeof a);  // padding too: launch arguments are compared bytewise
  a.slots = sim->slots;
  a.amask = sim->amask;
  a.asum = sim->asum;
  a.V = sim->V;
  a.last_fire = sim->last_fire;
  a.firings = sim->firings;
  a.ndw = sim->ndw;
  a.nid = sim->nid;
  a.slice_off = sim->slice_off;
  a.active_slices = sim->active;
  a.n_active_slices = (uint32_t)sim->active_host.size();
  a.act_identity = sim->act_identity ? 1u : 0u;
  a.act_q0 = sim->act_q0;
  a.nets = sim->d_nets;
  a.groups = sim->d_groups.p;
  a.thr = sim->d_thr.p;
  a.forced_in = sim->forced_in;
  a.forced_bg = sim->forced_bg;
  a.raster = sim->raster;
  a.raster_cnt = sim->raster_cnt;
  a.raster_cap = sim->raster_cap;
  a.ctr = sim->counting ? sim->ctr : nullptr;
  a.ctr_rows = (uint32_t)sim->ctr_rows;
  a.t_pre = sim->t_pre;
  a.t_post = sim->t_post;
  a.w_ap = sim->w_ap;
  a.len_pre = (int32_t)sim->T.t_pre.size();
  a.len_post = (int32_t)sim->T.t_post.size();
  a.tabs_in_lds = sim->cfg.stage_tables && (uint64_t)a.len_pre * 4 <= NC_LDS_TAB_BYTES ? 1 : 0;
  a.chunk_off = sim->d_chunk_off;
  a.xcd_swizzle = sim->cfg.xcd_swizzle;
  for (int k = 0; k < NC_WAP_K; ++k)
    a.w_ap_k[k] = k < (int)sim->T.w_ap.size() ? sim->T.w_ap[k] : 0.0f;
  a.dt = sim->T.dt;
  a.P = sim->T.P;
  a.depol_k = sim->T.depol_k;
  a.c0 = NC_CHARGE_K;
  a.e = sim->T.e;
  a.d_ref = sim->T.d_ref;
  a.r_rel = sim->T.r_rel;
  a.out_row = sim->out_row;
  a.edges = sim->edges;
  a.pend0 = sim->pend0;
  a.pend1 = sim->pend1;
  a.pend_cnt = sim->pend_cnt;
  a.bcnt = sim->bcnt;
  a.bent = sim->bent;
  a.acache = sim->acache;
  a.fired = sim->fired;
  a.fired_cnt = sim->fired_cnt;
  return a;
}

// The list of slices the step kernels visit.  When the live networks' slices
// span at most 5/4 of their number, the list is that whole span in slice
// order, holes included (records of the never-live network max_nets), so
// that record i is slice q0 + i and the kernels' per-neuron loads need not
// wait for the record; otherwise every slice of every live network,
// networks in id order (each network's slices are contiguous).
nc_status rebuild_active(nc_sim* sim) {
  if (!sim->active_dirty) return NC_OK;
  if (sim->upload_pending) {
    NC_HIP(hipEventSynchronize(sim->upload_done));
    sim->upload_pending = false;
  }
  NC_HIP(hipStreamSynchronize(sim->stream));
  sim->active_host.clear();
  uint32_t qmin = UINT32_MAX, qmax = 0;
  uint64_t nlive = 0;
  for (const NetRec& r : sim->nets) {
    if (!r.live || r.nq == 0) continue;
    qmin = std::min(qmin, r.q0);
    qmax = std::max(qmax, r.q0 + r.nq);
    nlive += r.nq;
  }
  if (nlive && (uint64_t)(qmax - qmin) * 4 <= nlive * 5) {
    for (uint32_t q = qmin; q < qmax; ++q) {
      const uint32_t id = sim->slice_net_host[q];
      if (id != NC_SLICE_NONE && sim->nets[id].live)
        sim->active_host.push_back(nc_slice_rec{q, sim->slice_off_host[q], id, sim->slice_moff_host[q]});
      else
        sim->active_host.push_back(nc_slice_rec{q, 0u, sim->max_nets, 0u});
    }
    for (NetRec& r : sim->nets) {
      if (!r.live) continue;
      r.act0 = r.nq ? r.q0 - qmin : 0u;
      r.act_valid = true;
    }
    sim->act_identity = true;
    sim->act_q0 = qmin;
    NC_HIP(hipMemcpy(sim->active, sim->active_host.data(), sim->active_host.size() * sizeof(nc_slice_rec),
                     hipMemcpyHostToDevice));
    sim->active_dirty = false;
    return NC_OK;
  }
  for (size_t i = 0; i < sim->nets.size(); ++i) {
    NetRec& r = sim->nets[i];
    if (!r.live) continue;
    r.act0 = (uint32_t)sim->active_host.size();
    r.act_valid = true;
    for (uint32_t q = r.q0; q < r.q0 + r.nq; ++q)
      sim->active_host.push_back(nc_slice_rec{q, sim->slice_off_host[q], (uint32_t)i, sim->slice_moff_host[q]});
  }
  sim->act_identity = true;
  sim->act_q0 = sim->active_host.empty() ? 0u : sim->active_host[0].q;
  for (size_t i = 0; i < sim->active_host.size(); ++i)
    if (sim->active_host[i].q != sim->act_q0 + (uint32_t)i) sim->act_identity = false;
  if (!sim->active_host.empty())
    NC_HIP(hipMemcpy(sim->active, sim->active_host.data(), sim->active_host.size() * sizeof(nc_slice_rec),
                     hipMemcpyHostToDevice));
  sim->active_dirty = false;
  return NC_OK;
}

// Device room for every host event queued on the simulation's networks, so
// that nc_sim_step, which uploads those due in its steps, never allocates.
nc_status reserve_events(nc_sim* sim) {
  size_t total = 0;
  for (const NetRec& r : sim->nets)
    if (r.live) total += r.events.size();
  if (total <= sim->d_ev_pos.n && total <= sim->d_ev_kind.n) return NC_OK;
  if (sim->upload_pending) {
    NC_HIP(hipEventSynchronize(sim->upload_done));
    sim->upload_pending = false;
  }
  nc_status st = grow(&sim->d_ev_pos, total, sim->stream);
  if (st == NC_OK) st = grow(&sim->d_ev_kind, total, sim->stream);
  return st;
}

nc_status check_barrier(nc_sim* sim);

// A network's weights in export order into dev (at least S floats), then the
// first n of them to the caller's out.
nc_status export_weights(nc_sim* sim, NetRec& r, float* dev, float* out, uint64_t n) {
  const uint32_t pos0 = r.q0 * 64;
  hipLaunchKernelGGL(k_export_w, dim3((r.n + 255) / 256), dim3(256), 0, sim->stream, sim->slots, sim->slice_off,
                     sim->ndw, sim->id_pos, sim->id_row, pos0, r.q0, r.n, dev);
  NC_HIP(hipGetLastError());
  NC_HIP(hipMemcpyAsync(out, dev, n * 4, hipMemcpyDeviceToHost, sim->stream));
  NC_HIP(hipStreamSynchronize(sim->stream));
  return check_barrier(sim);
}

hipEvent_t get_event(nc_sim* sim, size_t i) {
  while (sim->ev_pool.size() <= i) {
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    sim->ev_pool.push_back(e);
  }
  return sim->ev_pool[i];
}

}  // namespace

// ---------------------------------------------------------------------------

extern "C" nc_status nc_sim_footprint(uint64_t n_neurons, uint64_t n_syn, uint64_t* bytes) try {
  if (!bytes) return fail(NC_ERR_ARG, "null argument");
  // the buffers nc_sim_create allocates for one network of this size
  const uint64_t P = (n_neurons + 63) / 64 * 64 + 64, Q = P / 64;
  const uint64_t slots = (n_syn + n_syn / 4 + P * NC_CHUNK + 255) / 256 * 256;  // padding allowance
  const uint64_t mwords = slots / 32 + 64 * Q;
  *bytes = P * (4 + 12 + 8 + 4 + 8) + P / 4                // asum, V/lastFire/firings, ndw, nid, out_row, forced
           + P * sizeof(uint4)                             // active-slot cache
           + Q * (8 + 4 + 4 + sizeof(nc_slice_rec))       // pend_cnt, slice_off, slice_net, active
           + (Q + NC_FLISTS * 8) * NC_FSPLIT * 16          // fired lists
           + Q * NC_NB * (4ull + 8ull * NC_BCAP) + Q * 4     // arrival buckets, slice mask offsets
           + slots * (sizeof(nc_slot) + 8) + mwords * 4    // slot records, pending lists, active masks
           + (n_syn + 64) * sizeof(nc_edge)                // out-edges
           + (Q + 2) * NC_NCTR * 8;                        // counters
  return NC_OK;
}
NC_ABI_CATCH

extern "C" nc_status nc_net_device_size(const nc_net* net, uint64_t* positions, uint64_t* slots) try {
  if (!net) return fail(NC_ERR_ARG, "null network");
  if (net->image) {
    if (positions) *positions = (uint64_t)net->image->nq * 64;
    if (slots) *slots = net->image->nslots;
    return NC_OK;
  }
  Layout L;
  make_layout(*net, &L);
  if (positions) *positions = (uint64_t)L.nq * 64;
  if (slots) *slots = L.nslots;
  return NC_OK;
}
NC_ABI_CATCH

extern "C" nc_status nc_net_prepare(nc_net* net) {
  if (!net) return fail(NC_ERR_ARG, "null network");
  if (net->image) return NC_OK;
  try {
    net->image = build_image(*net);
  } catch (const std::bad_alloc&) {
    return fail(NC_ERR_OOM, "host allocation failed");
  }
  return NC_OK;
}

extern "C" nc_status nc_sim_create(const nc_sim_spec* spec, nc_sim** out) try {
  if (!spec || !out) return fail(NC_ERR_ARG, "null argument");
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(NC_ERR_NO_DEVICE, "no HIP device");
  if (spec->device < 0 || spec->device >= ndev) return fail(NC_ERR_RANGE, "device index out of range");
  if (spec->max_neurons == 0 || spec->max_nets == 0) return fail(NC_ERR_ARG, "zero capacity");
  if (spec->max_synapses >= (1ull << 32) - (1ull << 24)) return fail(NC_ERR_ARG, "max_synapses too large");
  // fired-list records carry the slice index in NC_FSLICE_BITS bits
  if ((uint64_t)spec->max_neurons / 64 + 2ull * spec->max_nets >= (1ull << NC_FSLICE_BITS))
    return fail(NC_ERR_ARG, "max_neurons too large");
  nc_sim* sim = new (std::nothrow) nc_sim();
  if (!sim) return fail(NC_ERR_OOM, "host allocation failed");
  // destroys the half-built simulation on a failure or an exception
  std::unique_ptr<nc_sim, void (*)(nc_sim*)> guard(sim, nc_sim_destroy);
  sim->device = spec->device;
  nc_status st = NC_OK;
  auto H = [&](hipError_t e, const char* what) {
    if (e != hipSuccess && st == NC_OK)
      st = fail(e == hipErrorOutOfMemory ? NC_ERR_OOM : NC_ERR_HIP,
                std::string(what) + ": " + hipGetErrorString(e));
  };
  H(hipSetDevice(spec->device), "hipSetDevice");
  if (st == NC_OK) st = scratch_setup(spec->device);  // once per device, with it current
  H(hipStreamCreateWithFlags(&sim->stream, hipStreamNonBlocking), "hipStreamCreate");
  H(hipEventCreateWithFlags(&sim->upload_done, hipEventDisableTiming), "hipEventCreate");
  sim->max_pos = (spec->max_neurons + 63) / 64 * 64 + 64 * (uint64_t)spec->max_nets;
  sim->max_slots = spec->max_slots
                       ? (spec->max_slots + 255) / 256 * 256
                       : (spec->max_synapses + spec->max_synapses / 4 + sim->max_pos * NC_CHUNK + 255) / 256 * 256;
  if (sim->max_slots >= (1ull << 32)) sim->max_slots = ((1ull << 32) - 256);
  sim->max_out = spec->max_synapses + 64;
  sim->max_nets = spec->max_nets;
  sim->raster_cap = spec->raster_capacity;
  const uint64_t P = sim->max_pos, Q = P / 64;
  sim->max_mwords = sim->max_slots / 32 + 64 * Q;
  if (st == NC_OK) {
    H(hipMalloc(&sim->amask, sim->max_mwords * 4), "hipMalloc amask");
    H(hipMalloc(&sim->asum, P * 4), "hipMalloc asum");
    H(hipMalloc(&sim->slots, sim->max_slots * sizeof(nc_slot)), "hipMalloc slots");
    H(hipMalloc(&sim->V, P * 4), "hipMalloc V");
    H(hipMalloc(&sim->last_fire, P * 4), "hipMalloc lastFire");
    H(hipMalloc(&sim->firings, P * 4), "hipMalloc firings");
    H(hipMalloc(&sim->ndw, P * sizeof(uint2)), "hipMalloc ndw");
    H(hipMalloc(&sim->nid, P * 4), "hipMalloc nid");
    H(hipMalloc(&sim->out_row, P * sizeof(uint2)), "hipMalloc out_row");
    H(hipMalloc(&sim->edges, sim->max_out * sizeof(nc_edge)), "hipMalloc edges");
    H(hipMalloc(&sim->pend0, sim->max_slots * 4), "hipMalloc pend0");
    H(hipMalloc(&sim->pend1, sim->max_slots * 4), "hipMalloc pend1");
    H(hipMalloc(&sim->pend_cnt, Q * sizeof(uint2)), "hipMalloc pend_cnt");
    H(hipMalloc(&sim->bcnt, Q * NC_NB * 4ull), "hipMalloc bcnt");
    H(hipMalloc(&sim->bent, Q * NC_NB * NC_BCAP * sizeof(uint2)), "hipMalloc bent");
    H(hipMalloc(&sim->acache, P * sizeof(uint4)), "hipMalloc acache");
    H(hipMalloc(&sim->slice_moff, Q * 4), "hipMalloc slice_moff");
    H(hipMalloc(&sim->id_pos, P * 4), "hipMalloc id_pos");
    H(hipMalloc(&sim->id_row, P * 4), "hipMalloc id_row");
    H(hipMalloc(&sim->slice_off, Q * 4), "hipMalloc slice_off");
    H(hipMalloc(&sim->slice_net, Q * 4), "hipMalloc slice_net");
    H(hipMalloc(&sim->active, Q * sizeof(nc_slice_rec)), "hipMalloc active");
    H(hipMalloc(&sim->forced_in, (P / 32) * 4), "hipMalloc forced_in");
    H(hipMalloc(&sim->forced_bg, (P / 32) * 4), "hipMalloc forced_bg");
    H(hipMalloc(&sim->fired, (Q + NC_FLISTS * 8ull) * NC_FSPLIT * sizeof(uint4)), "hipMalloc fired");
    H(hipMalloc(&sim->fired_cnt, 2ull * NC_FLISTS * NC_FCNT_STRIDE * 4), "hipMalloc fired_cnt");

    H(hipMalloc(&sim->raster_cnt, 64), "hipMalloc raster_cnt");
    if (sim->raster_cap) H(hipMalloc(&sim->raster, 8ull * sim->raster_cap), "hipMalloc raster");
    sim->ctr_rows = Q + 2;
    // word 0: chunk offset; word 32: grid-barrier arrivals of the call;
    // word 64: a grid barrier timed out (sticky until reported); word 128:
    // zero, the group-spike-train word read when there are no groups
    H(hipMalloc(&sim->d_chunk_off, 1024), "hipMalloc chunk_off");
    H(hipMalloc(&sim->ctr, sim->ctr_rows * NC_NCTR * 8), "hipMalloc ctr");
    // one more, never written: the (never live) network of hole records
    H(hipMalloc(&sim->d_nets, (sim->max_nets + 1ull) * sizeof(nc_netdesc)), "hipMalloc nets");
  }
  if (st == NC_OK) {
    H(hipMemsetAsync(sim->amask, 0, sim->max_mwords * 4, sim->stream), "memset");
    H(hipMemsetAsync(sim->asum, 0, P * 4, sim->stream), "memset");
    H(hipMemsetAsync(sim->ndw, 0, P * sizeof(uint2), sim->stream), "memset");
    H(hipMemsetAsync(sim->forced_in, 0, (P / 32) * 4, sim->stream), "memset");
    H(hipMemsetAsync(sim->forced_bg, 0, (P / 32) * 4, sim->stream), "memset");
    H(hipMemsetAsync(sim->fired, 0, (Q + NC_FLISTS * 8ull) * NC_FSPLIT * sizeof(uint4), sim->stream), "memset");
    H(hipMemsetAsync(sim->fired_cnt, 0, 2ull * NC_FLISTS * NC_FCNT_STRIDE * 4, sim->stream), "memset");
    H(hipMemsetAsync(sim->pend_cnt, 0, Q * sizeof(uint2), sim->stream), "memset");
    H(hipMemsetAsync(sim->bcnt, 0, Q * NC_NB * 4ull, sim->stream), "memset");
    H(hipMemsetAsync(sim->acache, 0, P * sizeof(uint4), sim->stream), "memset");

    H(hipMemsetAsync(sim->raster_cnt, 0, 64, sim->stream), "memset");
    H(hipMemsetAsync(sim->d_chunk_off, 0, 1024, sim->stream), "memset");
    H(hipMemsetAsync(sim->ctr, 0, sim->ctr_rows * NC_NCTR * 8, sim->stream), "memset");
    H(hipMemsetAsync(sim->d_nets, 0, (sim->max_nets + 1ull) * sizeof(nc_netdesc), sim->stream), "memset");
    std::vector<uint32_t> none(Q, NC_SLICE_NONE);
    sim->slice_net_host = none;
    sim->slice_off_host.assign(Q, 0u);
    sim->slice_moff_host.assign(Q, 0u);
    H(hipMemcpyAsync(sim->slice_net, none.data(), Q * 4, hipMemcpyHostToDevice, sim->stream), "memcpy");
    H(hipStreamSynchronize(sim->stream), "sync");
  }
  sim->a_slices.init(Q);
  sim->a_slots.init(sim->max_slots / 256);
  sim->a_out.init(sim->max_out);
  sim->a_masks.init(sim->max_mwords / 64);
  if (st != NC_OK) return st;
  *out = guard.release();
  return NC_OK;
}
NC_ABI_CATCH

extern "C" void nc_sim_destroy(nc_sim* sim) {
  if (!sim) return;
  (void)hipSetDevice(sim->device);
  if (sim->stream) (void)hipStreamSynchronize(sim->stream);
  if (sim->graph) (void)hipGraphExecDestroy(sim->graph);
  void* ptrs[] = {sim->d_chunk_off, sim->amask, sim->asum, sim->slots, sim->V, sim->last_fire, sim->firings, sim->ndw,
                  sim->nid, sim->out_row, sim->edges, sim->pend0, sim->pend1, sim->pend_cnt, sim->bcnt, sim->bent, sim->acache, sim->slice_moff, sim->id_pos, sim->id_row, sim->slice_off, sim->slice_net, sim->active, sim->forced_in,
                  sim->forced_bg, sim->fired, sim->fired_cnt, sim->raster, sim->raster_cnt,
                  sim->ctr, sim->d_nets, sim->t_pre, sim->t_post, sim->w_ap, sim->d_groups.p,
                  sim->d_thr.p, sim->d_ev_pos.p, sim->d_ev_kind.p, sim->d_gfire.p, sim->d_wbuf.p, sim->d_stage.p};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  for (hipEvent_t e : sim->ev_pool) (void)hipEventDestroy(e);
  if (sim->upload_done) (void)hipEventDestroy(sim->upload_done);
  if (sim->stream) (void)hipStreamDestroy(sim->stream);
  delete sim;
}

extern "C" nc_status nc_sim_free_bytes(const nc_sim* sim, uint64_t* free_bytes) try {
  if (!sim || !free_bytes) return fail(NC_ERR_ARG, "null argument");
  NC_HIP(hipSetDevice(sim->device));
  size_t fr = 0, tot = 0;
  NC_HIP(hipMemGetInfo(&fr, &tot));
  *free_bytes = fr;
  return NC_OK;
}
NC_ABI_CATCH

extern "C" nc_status nc_sim_add(nc_sim* sim, const nc_net* net, const nc_params* params,
                                const nc_drive_spec* drive, uint32_t steps_to_run,
                                int32_t* net_id) try {
  if (!sim || !net || !params || !net_id) return fail(NC_ERR_ARG, "null argument");
  NC_HIP(hipSetDevice(sim->device));
  // The drive tables below are host vectors a pending upload may still read.
  NC_HIP(hipStreamSynchronize(sim->stream));
  const nc_params p = *params;
  {
    // dt must be 2^-m, 0 <= m <= 8, so that float time is exact on the grid.
    int ex = 0;
    float fr = frexpf(p.dt, &ex);
    if (!(p.dt > 0.0f) || fr != 0.5f || ex > 1 || ex < -7)
      return fail(NC_ERR_ARG, "dt must be a power of two in [2^-8, 1] ms");
  }
  if (!sim->tables_set) {
    build_tables(p.dt, p.pre_trace_decay, p.post_trace_decay, &sim->T);
    const Tables& T = sim->T;
    NC_HIP(hipMalloc(&sim->t_pre, std::max<size_t>(1, T.t_pre.size()) * 4));
    NC_HIP(hipMalloc(&sim->t_post, std::max<size_t>(1, T.t_post.size()) * 4));
    NC_HIP(hipMalloc(&sim->w_ap, T.w_ap.size() * 4));
    NC_HIP(hipMemcpy(sim->t_pre, T.t_pre.data(), T.t_pre.size() * 4, hipMemcpyHostToDevice));
    NC_HIP(hipMemcpy(sim->t_post, T.t_post.data(), T.t_post.size() * 4, hipMemcpyHostToDevice));
    NC_HIP(hipMemcpy(sim->w_ap, T.w_ap.data(), T.w_ap.size() * 4, hipMemcpyHostToDevice));
    sim->tables_set = true;
  } else if (p.dt != sim->T.dt || p.pre_trace_decay != sim->T.pre_decay ||
             p.post_trace_decay != sim->T.post_decay) {
    return fail(NC_ERR_ARG, "all networks of one nc_sim share dt and the trace decays");
  }
  if ((uint32_t)sim->T.r_rel >= 250u) return fail(NC_ERR_ARG, "dt too fine: the release window must be < 250 steps");
  // a pending entry holds a slot's offset inside its slice in 24 bits
  if (net->max_in_degree > NC_MAX_IN_DEGREE) return fail(NC_ERR_ARG, "in-degree too large (at most 262140)");

  // Free network id.
  int32_t id = -1;
  for (size_t i = 0; i < sim->nets.size(); ++i)
    if (!sim->nets[i].live) { id = (int32_t)i; break; }
  if (id < 0) {
    if (sim->nets.size() >= sim->max_nets) return fail(NC_ERR_CAPACITY, "network table full");
    sim->nets.emplace_back();
    id = (int32_t)sim->nets.size() - 1;
  }

  std::shared_ptr<const NetImage> img = net->image;
  if (!img) img = build_image(*net);  // not kept: nc_net_prepare is what caches it
  const NetImage& I = *img;
  const NetMaps& M = *I.maps;
  const uint64_t S = I.S;
  uint64_t q0 = 0, sl0 = 0, o0 = 0;
  if (!sim->a_slices.alloc(I.nq, &q0)) return fail(NC_ERR_CAPACITY, "no room for the neurons");
  if (!sim->a_slots.alloc(I.nslots / 256, &sl0)) {
    sim->a_slices.release(q0, I.nq);
    return fail(NC_ERR_CAPACITY, "no room for the synapses");
  }
  if (!sim->a_out.alloc(S, &o0)) {
    sim->a_slices.release(q0, I.nq);
    sim->a_slots.release(sl0, I.nslots / 256);
    return fail(NC_ERR_CAPACITY, "no room for the out-synapse lists");
  }
  uint64_t mk0 = 0;
  if (!sim->a_masks.alloc(I.nmwords / 64, &mk0)) {
    sim->a_slices.release(q0, I.nq);
    sim->a_slots.release(sl0, I.nslots / 256);
    sim->a_out.release(o0, S);
    return fail(NC_ERR_CAPACITY, "no room for the active-slot masks");
  }
  sl0 *= 256;
  mk0 *= 64;
  auto release_ranges = [&]() {
    sim->a_slices.release(q0, I.nq);
    sim->a_slots.release(sl0 / 256, I.nslots / 256);
    sim->a_out.release(o0, S);
    sim->a_masks.release(mk0 / 64, I.nmwords / 64);
    sim->nets[id].live = false;
  };

  NetRec& r = sim->nets[id];
  r = NetRec();
  r.n = net->n;
  r.S = S;
  r.q0 = (uint32_t)q0;
  r.nq = I.nq;
  r.slot0 = sl0;
  r.nslots = I.nslots;
  r.out0 = o0;
  r.mask0 = mk0;
  r.nmwords = I.nmwords;
  r.lr = p.learning_rate;
  r.pre_f = p.pre_factor;
  r.post_f = p.post_factor;
  r.steps_left = steps_to_run ? steps_to_run : UINT32_MAX;
  r.maps = I.maps;
  r.mwords_step = I.mwords_step;

  const uint32_t npos = I.nq * 64;
  const uint32_t pos0 = (uint32_t)(q0 * 64);
  std::vector<uint32_t> slice_off(I.nq), slice_moff(I.nq);
  for (uint32_t q = 0; q < I.nq; ++q) {
    slice_off[q] = (uint32_t)sl0 + I.slice_off[q];
    slice_moff[q] = (uint32_t)mk0 + I.slice_moff[q];
  }
  std::vector<uint2> orow(npos);
  for (uint32_t pl = 0; pl < npos; ++pl) orow[pl] = make_uint2((uint32_t)o0 + I.out_row[pl].x, I.out_row[pl].y);
  // Drive.
  std::vector<uint2> ndwv(npos);
  for (uint32_t pl = 0; pl < npos; ++pl) ndwv[pl] = make_uint2(I.degw[pl], 0u);
  nc_netdesc nd{};
  nd.lr = r.lr;
  nd.pre_f = r.pre_f;
  nd.post_f = r.post_f;
  if (drive) {
    nc::InputSel sel;
    nc_status stt = nc::select_inputs(*net, *drive, &sel);
    if (stt != NC_OK) {
      release_ranges();
      return stt;
    }
    r.has_drive = true;
    r.drive = *drive;
    const uint32_t T = drive->period ? drive->period : 1;
    const uint32_t ncorr = drive->n_correlated > 0 ? drive->n_groups : 0;
    // Room for this network's groups and its two threshold tables (one per
    // amplitude: correlated groups, independent inputs).
    uint64_t g0 = 0, t0 = 0;
    const nc_group dead{0u, 1u, 0u, sim->max_nets};
    bool grew = false;
    auto room = [&grew](RangeAlloc& A, auto& host, uint64_t n, auto fill, uint64_t* off) {
      if (A.alloc(n, off)) return;
      const uint64_t c = std::max<uint64_t>({A.cap * 2, A.cap + n, 1024});
      A.extend(c);
      host.resize(c, fill);
      grew = true;
      A.alloc(n, off);
    };
    room(sim->a_groups, sim->groups, sel.n_groups_total, dead, &g0);
    room(sim->a_thr, sim->thr, 2ull * T, 0u, &t0);
    r.grp0 = (uint32_t)g0;
    r.ngrp = sel.n_groups_total;
    r.thr0 = (uint32_t)t0;
    r.nthr = 2 * T;
    for (uint32_t tb = 0; tb < 2; ++tb) {
      const double amp = tb == 0 ? drive->amp_corr : drive->amp_ind;
      for (uint32_t k = 0; k < T; ++k) {
        double pr = (double)drive->p0 * (1.0 + amp * sin(2.0 * M_PI * (double)k / (double)T));
        uint32_t th;
        if (!(pr > 0.0)) th = 0;
        else if (pr >= 1.0) th = 0xFFFFFFFFu;
        else th = (uint32_t)floor(pr * 4294967296.0);
        sim->thr[t0 + (uint64_t)tb * T + k] = th;
      }
    }
    for (uint32_t g = 0; g < sel.n_groups_total; ++g)
      sim->groups[g0 + g] = nc_group{(uint32_t)t0 + (g < ncorr ? 0u : T), T, g, (uint32_t)id};
    const size_t gw = (sim->groups.size() + 31) / 32;
    if (grew || sim->d_groups.n < sim->groups.size() || sim->d_thr.n < sim->thr.size() ||
        sim->d_gfire.n < (size_t)NC_GRING * gw) {
      // The device tables and the ring of group spike trains follow the host
      // tables' capacity, here: nc_sim_step never allocates.
      nc_status sg = grow(&sim->d_groups, sim->groups.size(), sim->stream);
      if (sg == NC_OK) sg = grow(&sim->d_thr, sim->thr.size(), sim->stream);
      if (sg == NC_OK) sg = grow(&sim->d_gfire, (size_t)NC_GRING * gw, sim->stream);
      if (sg != NC_OK) {
        for (uint32_t g = 0; g < r.ngrp; ++g) sim->groups[g0 + g] = dead;
        sim->a_groups.release(g0, r.ngrp);
        sim->a_thr.release(t0, r.nthr);
        release_ranges();
        return sg;
      }
      NC_HIP(hipMemcpyAsync(sim->d_groups.p, sim->groups.data(), sim->groups.size() * sizeof(nc_group),
                            hipMemcpyHostToDevice, sim->stream));
      NC_HIP(hipMemcpyAsync(sim->d_thr.p, sim->thr.data(), sim->thr.size() * 4, hipMemcpyHostToDevice,
                            sim->stream));
    } else if (r.ngrp) {
      NC_HIP(hipMemcpyAsync(sim->d_groups.p + g0, sim->groups.data() + g0, r.ngrp * sizeof(nc_group),
                            hipMemcpyHostToDevice, sim->stream));
      NC_HIP(hipMemcpyAsync(sim->d_thr.p + t0, sim->thr.data() + t0, r.nthr * 4ull, hipMemcpyHostToDevice,
                            sim->stream));
    }
    for (size_t k = 0; k < sel.ids.size(); ++k) {
      uint32_t pl = M.pos_of_id[sel.ids[k]];
      ndwv[pl].x |= NC_DEG_INPUT;
      ndwv[pl].y = r.grp0 + sel.group[k];
    }
    nd.drive_lo = (uint32_t)drive->seed;
    nd.drive_hi = (uint32_t)(drive->seed >> 32);
    if (drive->bg_prob > 0.0f) {
      double th = floor((double)drive->bg_prob * 4294967296.0);
      nd.bg_thr = th >= 4294967295.0 ? 0xFFFFFFFFu : (uint32_t)th;
      if (nd.bg_thr == 0) nd.bg_thr = 1;
      uint64_t bs = drive->seed ^ 0x9E3779B97F4A7C15ull;
      nd.bg_lo = (uint32_t)bs;
      nd.bg_hi = (uint32_t)(bs >> 32);
    }
  }
  std::vector<uint32_t> snet(I.nq, (uint32_t)id);
  hipStream_t st = sim->stream;
  // The compact image (4 bytes and a bit per slot, 8 bytes per edge) goes up
  // to a staging buffer; the placing kernels expand it into the records.
  const uint64_t n_inh = I.slot_inh.size(), o_jq = (I.nslots + n_inh + 1) & ~1ull;
  {
    nc_status sg = grow(&sim->d_stage, o_jq + 2 * S + 2, st);
    if (sg != NC_OK) {
      if (r.has_drive) {
        const nc_group dead{0u, 1u, 0u, sim->max_nets};
        for (uint32_t g = 0; g < r.ngrp; ++g) sim->groups[r.grp0 + g] = dead;
        if (r.ngrp)
          NC_HIP(hipMemcpy(sim->d_groups.p + r.grp0, sim->groups.data() + r.grp0, r.ngrp * sizeof(nc_group),
                           hipMemcpyHostToDevice));
        sim->a_groups.release(r.grp0, r.ngrp);
        sim->a_thr.release(r.thr0, r.nthr);
      }
      release_ranges();
      return sg;
    }
  }
  uint32_t* const stg = sim->d_stage.p;
  auto blocks = [](uint64_t n) { return dim3((uint32_t)std::min<uint64_t>((n + 255) / 256, 4096)); };
  NC_HIP(hipMemcpyAsync(sim->slice_off + q0, slice_off.data(), I.nq * 4, hipMemcpyHostToDevice, st));
  NC_HIP(hipMemcpyAsync(sim->slice_moff + q0, slice_moff.data(), I.nq * 4, hipMemcpyHostToDevice, st));
  if (I.nslots) {
    NC_HIP(hipMemcpyAsync(stg, I.slot_w.data(), I.nslots * 4, hipMemcpyHostToDevice, st));
    NC_HIP(hipMemcpyAsync(stg + I.nslots, I.slot_inh.data(), n_inh * 4, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_place_slots, blocks(I.nslots), dim3(256), 0, st, sim->slots + sl0, (const float*)stg,
                       (const uint32_t*)(stg + I.nslots), (uint64_t)I.nslots);
    NC_HIP(hipGetLastError());
  }
  if (S) {
    NC_HIP(hipMemcpyAsync(stg + o_jq, I.edge_jq.data(), S * sizeof(uint2), hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_place_edges, blocks(S), dim3(256), 0, st, sim->edges + o0, (const uint2*)(stg + o_jq), S,
                       (uint32_t)sl0, (uint32_t)q0, (const uint32_t*)sim->slice_off);
    NC_HIP(hipGetLastError());
  }
  hipLaunchKernelGGL(k_place_neurons, blocks(npos), dim3(256), 0, st, sim->V + pos0, sim->last_fire + pos0,
                     sim->firings + pos0, (uint64_t)npos);
  NC_HIP(hipGetLastError());
  NC_HIP(hipMemsetAsync(sim->amask + mk0, 0, I.nmwords * 4, st));
  NC_HIP(hipMemsetAsync(sim->asum + pos0, 0, npos * 4, st));
  NC_HIP(hipMemcpyAsync(sim->ndw + pos0, ndwv.data(), npos * sizeof(uint2), hipMemcpyHostToDevice, st));
  NC_HIP(hipMemcpyAsync(sim->nid + pos0, I.nid.data(), npos * 4, hipMemcpyHostToDevice, st));
  // the export maps: weights are read in export order on the device
  std::vector<uint32_t> idrow(net->n);
  for (uint32_t i = 0; i < net->n; ++i) idrow[i] = (uint32_t)M.in_ptr[i];
  if (net->n) {
    NC_HIP(hipMemcpyAsync(sim->id_pos + pos0, M.pos_of_id.data(), net->n * 4ull, hipMemcpyHostToDevice, st));
    NC_HIP(hipMemcpyAsync(sim->id_row + pos0, idrow.data(), net->n * 4ull, hipMemcpyHostToDevice, st));
  }
  NC_HIP(hipMemcpyAsync(sim->out_row + pos0, orow.data(), npos * sizeof(uint2), hipMemcpyHostToDevice, st));
  NC_HIP(hipMemsetAsync(sim->pend_cnt + q0, 0, I.nq * sizeof(uint2), st));
  // no spike of a network that held these slices before may arrive here
  NC_HIP(hipMemsetAsync(sim->bcnt + (size_t)q0 * NC_NB, 0, (size_t)I.nq * NC_NB * 4, st));
  NC_HIP(hipMemcpyAsync(sim->slice_net + q0, snet.data(), I.nq * 4, hipMemcpyHostToDevice, st));
  NC_HIP(hipMemcpyAsync(sim->d_nets + id, &nd, sizeof(nd), hipMemcpyHostToDevice, st));
  NC_HIP(hipStreamSynchronize(st));
  for (uint32_t q = 0; q < I.nq; ++q) {
    sim->slice_net_host[q0 + q] = (uint32_t)id;
    sim->slice_off_host[q0 + q] = slice_off[q];
    sim->slice_moff_host[q0 + q] = slice_moff[q];
  }
  r.live = true;
  sim->active_dirty = true;
  *net_id = id;
  return NC_OK;
}
NC_ABI_CATCH

extern "C" nc_status nc_sim_remove(nc_sim* sim, int32_t net_id) try {
  NetRec* r;
  nc_status stt = check_net(sim, net_id, &r);
  if (stt != NC_OK) return stt;
  NC_HIP(hipSetDevice(sim->device));
  NC_HIP(hipStreamSynchronize(sim->stream));
  std::vector<uint32_t> none(r->nq, NC_SLICE_NONE);
  NC_HIP(hipMemcpy(sim->slice_net + r->q0, none.data(), r->nq * 4, hipMemcpyHostToDevice));
  for (uint32_t q = 0; q < r->nq; ++q) sim->slice_net_host[r->q0 + q] = NC_SLICE_NONE;
  sim->a_slices.release(r->q0, r->nq);
  sim->a_slots.release(r->slot0 / 256, r->nslots / 256);
  sim->a_out.release(r->out0, r->S);
  sim->a_masks.release(r->mask0 / 64, r->nmwords / 64);
  if (r->has_drive) {
    // the groups go back to the table, pointing at the never-live network
    const nc_group dead{0u, 1u, 0u, sim->max_nets};
    for (uint32_t g = 0; g < r->ngrp; ++g) sim->groups[r->grp0 + g] = dead;
    if (r->ngrp)
      NC_HIP(hipMemcpy(sim->d_groups.p + r->grp0, sim->groups.data() + r->grp0, r->ngrp * sizeof(nc_group),
                       hipMemcpyHostToDevice));
    sim->a_groups.release(r->grp0, r->ngrp);
    sim->a_thr.release(r->thr0, r->nthr);
  }
  r->live = false;
  r->maps.reset();
  sim->active_dirty = true;
  return NC_OK;
}
NC_ABI_CATCH

extern "C" nc_status nc_sim_step(nc_sim* sim, uint32_t n_steps) try {
  if (!sim) return fail(NC_ERR_ARG, "null sim");
  if (n_steps == 0) return NC_OK;
  NC_HIP(hipSetDevice(sim->device));
  hipStream_t st = sim->stream;
  // Per-network window of this call.
  if (sim->upload_pending) {
    NC_HIP(hipEventSynchronize(sim->upload_done));
    sim->upload_pending = false;
  }
  std::vector<nc_netdesc>& nds = sim->h_nds;
  nds.assign(sim->nets.size(), nc_netdesc{});
  bool any_active = false;
  uint32_t max_run = 0;
  for (size_t i = 0; i < sim->nets.size(); ++i) {
    NetRec& r = sim->nets[i];
    nc_netdesc& nd = nds[i];
    nd = nc_netdesc{};
    if (!r.live) continue;
    uint32_t run = std::min<uint64_t>(n_steps, r.steps_left);
    if ((uint64_t)r.steps_done + run >= (uint64_t)NC_MAX_STEPS)
      return fail(NC_ERR_RANGE, "step count limit reached");
    nd.base_step = (int32_t)r.steps_done;
    nd.remaining = (int32_t)run;
    nd.whost = (int32_t)r.whost;
    nd.lr = r.lr;
    nd.pre_f = r.pre_f;
    nd.post_f = r.post_f;
    if (r.has_drive) {
      nd.drive_lo = (uint32_t)r.drive.seed;
      nd.drive_hi = (uint32_t)(r.drive.seed >> 32);
      if (r.drive.bg_prob > 0.0f) {
        double th = floor((double)r.drive.bg_prob * 4294967296.0);
        nd.bg_thr = th >= 4294967295.0 ? 0xFFFFFFFFu : (uint32_t)th;
        if (nd.bg_thr == 0) nd.bg_thr = 1;
        uint64_t bs = r.drive.seed ^ 0x9E3779B97F4A7C15ull;
        nd.bg_lo = (uint32_t)bs;
        nd.bg_hi = (uint32_t)(bs >> 32);
      }
    }
    if (run) any_active = true;
    max_run = std::max(max_run, run);
  }
  if (!any_active) return NC_OK;
  {
    nc_status s0 = rebuild_active(sim);
    if (s0 != NC_OK) return s0;
  }
  NC_HIP(hipMemcpyAsync(sim->d_nets, nds.data(), nds.size() * sizeof(nc_netdesc), hipMemcpyHostToDevice, st));

  // Host-scheduled events for steps [1, max_run] of this call, per step.
  std::vector<uint32_t>& ev_pos = sim->h_ev_pos;
  std::vector<uint8_t>& ev_kind = sim->h_ev_kind;
  ev_pos.clear();
  ev_kind.clear();
  std::vector<uint32_t> ev_ptr(max_run + 1, 0);
  {
    std::vector<std::vector<std::pair<uint32_t, uint8_t>>> per(max_run);
    for (size_t i = 0; i < sim->nets.size(); ++i) {
      NetRec& r = sim->nets[i];
      if (!r.live || r.events.empty()) continue;
      std::vector<std::pair<uint32_t, std::pair<uint32_t, uint8_t>>> keep;
      const uint32_t run = (uint32_t)nds[i].remaining;
      for (auto& ev : r.events) {
        uint32_t k = ev.first - r.steps_done;  // 1-based offset in this call
        if (k >= 1 && k <= run) per[k - 1].push_back(ev.second);
        else if (k > run) keep.push_back(ev);
      }
      r.events.swap(keep);
    }
    for (uint32_t k = 0; k < max_run; ++k) {
      ev_ptr[k + 1] = ev_ptr[k] + (uint32_t)per[k].size();
      for (auto& e : per[k]) {
        ev_pos.push_back(e.first);
        ev_kind.push_back(e.second);
      }
    }
  }
  if (!ev_pos.empty()) {
    // nc_push_forced and nc_sim_restore reserve room for every queued event
    if (sim->d_ev_pos.n < ev_pos.size() || sim->d_ev_kind.n < ev_kind.size())
      return fail(NC_ERR_STATE, "event buffer smaller than the queued events");
    NC_HIP(hipMemcpyAsync(sim->d_ev_pos.p, ev_pos.data(), ev_pos.size() * 4, hipMemcpyHostToDevice, st));
    NC_HIP(hipMemcpyAsync(sim->d_ev_kind.p, ev_kind.data(), ev_kind.size(), hipMemcpyHostToDevice, st));
  }
  NC_HIP(hipEventRecord(sim->upload_done, st));
  sim->upload_pending = true;

  nc_step_args a = base_args(sim);
  a.groups = sim->d_groups.p;
  a.thr = sim->d_thr.p;
  a.n_groups = (uint32_t)sim->groups.size();
  a.gwords = (a.n_groups + 31) / 32;
  // k_neuron reads a group-spike-train word for every neuron (word 0 for
  // neurons that are not inputs): a zero word stands in for an empty ring
  a.gfire = sim->d_gfire.p ? sim->d_gfire.p : sim->d_chunk_off + 128;
  // nc_sim_add sizes the drive tables and the group-train ring; nothing here allocates
  if (a.n_groups && (sim->d_groups.n < a.n_groups || sim->d_gfire.n < (size_t)NC_GRING * a.gwords))
    return fail(NC_ERR_STATE, "drive tables were not placed");
  const uint32_t wpb = sim->cfg.waves_per_block;
  const uint32_t grid = (a.n_active_slices + wpb - 1) / wpb;
  // A fired list holds at most fcap records (up to NC_FSPLIT per k_neuron
  // wave of its residue).  k_fanout gives the lists about one reader per
  // record, up to fanout_waves waves in all (default 16384, 64 per CU): then
  // each wave takes every (waves / 64)-th record of its list.  A fired
  // slice's out-edges are split over several records only when that grid has
  // a wave per active slice or more (else every wave has records enough):
  // about one record per 64 edges at the active networks' mean out-degree.
  // With XCD lists (the XCD-ordered remap, ordinary launches), list (x, j)
  // takes the k_neuron waves of workgroups b = x mod 8 whose rank on the XCD
  // is j mod 8.
  a.xcd_lists = (sim->cfg.xcd_swizzle && !sim->cfg.persistent) ? 1u : 0u;
  a.fcap = (a.xcd_lists ? ((grid + 7) / 8 * wpb + 7) / 8 : (grid * wpb + NC_FLISTS - 1) / NC_FLISTS) * NC_FSPLIT;
  uint64_t syn_active = 0, n_active = 0;
  for (size_t i = 0; i < sim->nets.size(); ++i)
    if (sim->nets[i].live && nds[i].remaining > 0) {
      syn_active += sim->nets[i].S;
      n_active += sim->nets[i].n;
    }
  const uint64_t wave_cap = std::min<uint64_t>(sim->cfg.fanout_waves ? sim->cfg.fanout_waves : 16384u,
                                               (uint64_t)a.fcap * NC_FLISTS);
  const bool split = wave_cap >= a.n_active_slices && n_active > 0;
  a.fsplit_od = split ? (uint32_t)std::min<uint64_t>(1u << 20, (syn_active + n_active - 1) / n_active) : 0u;
  const uint64_t rec_bound =
      (uint64_t)grid * wpb + (split ? (uint64_t)grid * wpb + syn_active / 64 : 0u);
  const uint32_t fan_waves =
      (uint32_t)std::min<uint64_t>(wave_cap, (rec_bound + NC_FLISTS - 1) / NC_FLISTS * NC_FLISTS);
  uint32_t fan_grid = std::max<uint32_t>((fan_waves + wpb - 1) / wpb, (NC_FLISTS + wpb - 1) / wpb);
  // XCD lists: the same number of waves on every XCD, 8 per list round
  if (a.xcd_lists) fan_grid = (fan_grid + NC_FLISTS / wpb - 1) / (NC_FLISTS / wpb) * (NC_FLISTS / wpb);
  const uint32_t lds = a.tabs_in_lds ? (uint32_t)a.len_pre * 4 : 0u;
  const void* fn_neuron = nullptr;
  const void* fn_fanout = nullptr;
  switch (wpb) {
    case 1: fn_neuron = (const void*)k_neuron<1>, fn_fanout = (const void*)k_fanout<1>; break;
    case 2: fn_neuron = (const void*)k_neuron<2>, fn_fanout = (const void*)k_fanout<2>; break;
    case 8: fn_neuron = (const void*)k_neuron<8>, fn_fanout = (const void*)k_fanout<8>; break;
    default: fn_neuron = (const void*)k_neuron<4>, fn_fanout = (const void*)k_fanout<4>; break;
  }
  // The persistent form: one launch per step of a grid the device holds at
  // once (every workgroup resident, so that the grid barrier can end).
  const void* fn_step = nullptr;
  uint32_t pgrid = 0;
  uint32_t* const bar = sim->d_chunk_off + 32;
  if (sim->cfg.persistent) {
    switch (wpb) {
      case 1: fn_step = (const void*)k_step<1>; break;
      case 2: fn_step = (const void*)k_step<2>; break;
      case 8: fn_step = (const void*)k_step<8>; break;
      default: fn_step = (const void*)k_step<4>; break;
    }
    int per_cu = 0, cus = 0;
    NC_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn_step, (int)(64 * wpb), lds));
    NC_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, sim->device));
    pgrid = std::min<uint32_t>((uint32_t)std::max(per_cu, 0) * (uint32_t)std::max(cus, 0), std::max(grid, fan_grid));
    if (pgrid * wpb < NC_FLISTS) return fail(NC_ERR_STATE, "the device holds too few workgroups for the persistent form");
    hipLaunchKernelGGL(k_set_u32, dim3(1), dim3(64), 0, st, bar, 1u, 0u);
    NC_HIP(hipGetLastError());
  }
  hipLaunchKernelGGL(k_set_u32, dim3(1), dim3(64), 0, st, sim->d_chunk_off, 1u, 0u);
  NC_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_zero_strided, dim3(1), dim3(2 * NC_FLISTS), 0, st, sim->fired_cnt, 2u * NC_FLISTS,
                     (uint32_t)NC_FCNT_STRIDE);
  NC_HIP(hipGetLastError());
  size_t nev_used = 0;
  std::vector<int> timed_kinds;
  auto launch = [&](int kind, const void* fn, uint32_t g, uint32_t block, uint32_t shmem) -> nc_status {
    if (g == 0) return NC_OK;
    uint32_t* bar_arg = bar;
    void* args[] = {&a, &bar_arg};  // k_step takes the barrier words, the others only a
    const bool timed = ((sim->kernel_timing >> kind) & 1) && (sim->timing_step % sim->timing_every) == 0;
    if (timed) {
      // start/stop events carried by the dispatch itself: they time the
      // kernel, not the queue around it
      hipEvent_t e0 = get_event(sim, nev_used++);
      hipEvent_t e1 = get_event(sim, nev_used++);
      if (!e0 || !e1) return fail(NC_ERR_HIP, "hipEventCreate failed");
      NC_HIP(hipExtLaunchKernel(fn, dim3(g), dim3(block), args, shmem, st, e0, e1, 0));
      sim->k_launch[kind]++;
      timed_kinds.push_back(kind);
    } else {
      NC_HIP(hipLaunchKernel(fn, dim3(g), dim3(block), args, shmem, st));
    }
    return NC_OK;
  };
  // Whole chunks as captured graphs when nothing per-step comes from the host.
  const uint32_t G =
      (sim->cfg.graph_steps && ev_pos.empty() && sim->kernel_timing == 0 && max_run >= sim->cfg.graph_steps)
          ? sim->cfg.graph_steps
          : 0u;
  if (G) {
    a.chunk_i = 0;
    const uint32_t key[5] = {G, wpb, grid, lds, sim->cfg.persistent ? pgrid | 0x80000000u : fan_grid};
    if (!sim->graph || memcmp(&sim->graph_args, &a, sizeof a) != 0 || memcmp(sim->graph_key, key, sizeof key) != 0) {
      if (sim->graph) {
        (void)hipGraphExecDestroy(sim->graph);
        sim->graph = nullptr;
      }
      NC_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
      nc_status cs = NC_OK;
      for (uint32_t c = 0; c < G && cs == NC_OK; ++c) {
        a.chunk_i = (int32_t)c;
        uint32_t* bar_arg = bar;
        void* args[] = {&a, &bar_arg};
        if (fn_step ? hipLaunchKernel(fn_step, dim3(pgrid), dim3(64 * wpb), args, lds, st) != hipSuccess
                    : (hipLaunchKernel(fn_neuron, dim3(grid), dim3(64 * wpb), args, lds, st) != hipSuccess ||
                       hipLaunchKernel(fn_fanout, dim3(fan_grid), dim3(64 * wpb), args, 0, st) != hipSuccess))
          cs = fail(NC_ERR_HIP, "kernel launch during capture failed");
      }
      hipLaunchKernelGGL(k_advance, dim3(1), dim3(64), 0, st, sim->d_chunk_off, G);
      hipGraph_t g = nullptr;
      const hipError_t ce = hipStreamEndCapture(st, &g);
      if (cs != NC_OK) {
        if (g) (void)hipGraphDestroy(g);
        return cs;
      }
      NC_HIP(ce);
      const hipError_t ie = hipGraphInstantiate(&sim->graph, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      NC_HIP(ie);
      a.chunk_i = 0;
      sim->graph_args = a;
      memcpy(sim->graph_key, key, sizeof key);
    }
  }
  // The call in sub-chunks of at most NC_GRING steps (whole graph chunks):
  // the group spike trains of a sub-chunk are drawn into the ring first.
  // The device chunk offset holds the steps launched before the running
  // graph chunk or run of direct launches.
  const uint32_t L = G ? (NC_GRING / G) * G : NC_GRING;
  for (uint32_t c0 = 0; c0 < max_run; c0 += L) {
    const uint32_t len = std::min(L, max_run - c0);
    if (a.n_groups) {
      hipLaunchKernelGGL(k_groups, dim3((a.gwords * 32 + 255) / 256, len), dim3(256), 0, st, a, c0);
      NC_HIP(hipGetLastError());
    }
    uint32_t j = 0;
    if (G)
      for (; len - j >= G; j += G) NC_HIP(hipGraphLaunch(sim->graph, st));
    for (uint32_t i = c0 + j; i < c0 + len; ++i) {
      a.chunk_i = (int32_t)(i - (c0 + j));
      a.n_ev = ev_ptr[i + 1] - ev_ptr[i];
      a.ev_pos = sim->d_ev_pos.p + ev_ptr[i];
      a.ev_kind = sim->d_ev_kind.p + ev_ptr[i];
      a.use_forced = ev_pos.empty() ? 0 : 1;
      nc_status s1 = NC_OK;
      if (a.n_ev) {
        void* args[] = {&a};
        hipError_t e = hipLaunchKernel((const void*)k_drive, dim3((a.n_ev + 255) / 256), dim3(256), args, 0, st);
        if (e != hipSuccess) s1 = fail(NC_ERR_HIP, hipGetErrorString(e));
      }
      if (fn_step) {
        if (s1 == NC_OK) s1 = launch(1, fn_step, pgrid, 64 * wpb, lds);
      } else {
        if (s1 == NC_OK) s1 = launch(1, fn_neuron, grid, 64 * wpb, lds);
        if (s1 == NC_OK) s1 = launch(3, fn_fanout, fan_grid, 64 * wpb, 0);
      }
      if (s1 != NC_OK) return s1;
      sim->timing_step++;
    }
    if (j < len && c0 + len < max_run) {
      hipLaunchKernelGGL(k_advance, dim3(1), dim3(64), 0, st, sim->d_chunk_off, len - j);
      NC_HIP(hipGetLastError());
    }
  }
  sim->gstep += max_run;
  if (!timed_kinds.empty()) {
    NC_HIP(hipEventSynchronize(sim->ev_pool[nev_used - 1]));
    for (size_t q = 0; q < timed_kinds.size(); ++q) {
      float ms = 0.f;
      NC_HIP(hipEventElapsedTime(&ms, sim->ev_pool[2 * q], sim->ev_pool[2 * q + 1]));
      sim->k_ms[timed_kinds[q]] += ms;
    }
  }
  for (size_t i = 0; i < sim->nets.size(); ++i) {
    NetRec& r = sim->nets[i];
    if (!r.live) continue;
    uint32_t run = (uint32_t)nds[i].remaining;
    sim->host_neuron_steps += (uint64_t)run * r.n;
    sim->host_synapse_steps += (uint64_t)run * r.S;
    sim->host_mword_steps += (uint64_t)run * r.mwords_step;
    r.steps_done += run;
    if (r.steps_left != UINT32_MAX) r.steps_left -= run;
  }
  return NC_OK;
}
NC_ABI_CATCH

extern "C" nc_status nc_sim_progress(nc_sim* sim, int32_t net_id, uint32_t* steps_done,
                                     uint32_t* steps_left) try {
  NetRec* r;
  nc_status st = check_net(sim, net_id, &r);
  if (st != NC_OK) return st;
  if (steps_done) *steps_done = r->steps_done;
  if (steps_left) *steps_left = r->steps_left;
  return NC_OK;
}
NC_ABI_CATCH

extern "C" nc_status nc_get_time(nc_sim* sim, int32_t net_id, float* time_ms) try {
  NetRec* r;
  nc_status st = check_net(sim, net_id, &r);
  if (st != NC_OK) return st;
  if (!time_ms) return fail(NC_ERR_ARG, "null argument");
  *time_ms = (float)r->steps_done * sim->T.dt;
  return NC_OK;
}
NC_ABI_CATCH

extern "C" nc_status nc_set_params(nc_sim* sim, int32_t net_id, float learning_rate,
                                   float pre_factor, float post_factor) try {
  NetRec* r;
  nc_status st = check_net(sim, net_id, &r);
  if (st != NC_OK) return st;
  NC_HIP(hipSetDevice(sim->device));
  r->lr = learning_rate;
  r->pre_f = pre_factor;
  r->post_f = post_factor;
  return NC_OK;
}
NC_ABI_CATCH

extern "C" nc_status nc_push_forced(nc_sim* sim, int32_t net_id, uint32_t n_events,
                                    const uint32_t* step_off, const uint32_t* ids,
                                    const uint8_t* kind) try {
  NetRec* r;
  nc_status st = check_net(sim, net_id, &r);
  if (st != NC_OK) return st;
  if (n_events && (!step_off || !ids || !kind)) return fail(NC_ERR_ARG, "null argument");
  for (uint32_t k = 0; k < n_events; ++k) {
    if (ids[k] >= r->n) return fail(NC_ERR_RANGE, "neuron id out of range");
    if (step_off[k] < 1) return fail(NC_ERR_ARG, "step offsets start at 1");
    if (kind[k] > 1) return fail(NC_ERR_ARG, "kind must be 0 (input) or 1 (background)");
  }
  for (uint32_t k = 0; k < n_events; ++k)
    if (step_off[k] >= (uint32_t)NC_MAX_STEPS - r->steps_done) return fail(NC_ERR_RANGE, "event step out of range");
  const size_t before = r->events.size();
  for (uint32_t k = 0; k < n_events; ++k)
    r->events.push_back({r->steps_done + step_off[k], {r->pos(ids[k]), kind[k]}});
  NC_HIP(hipSetDevice(sim->device));
  st = reserve_events(sim);
  if (st != NC_OK) r->events.resize(before);  // all or nothing
  return st;
}
NC_ABI_CATCH

namespace {
// After a synchronisation: did a grid barrier of the persistent form time out
// (a grid whose workgroups were not all resident) since the last check?
nc_status check_barrier(nc_sim* sim) {
  uint32_t flag = 0;
  NC_HIP(hipMemcpy(&flag, sim->d_chunk_off + 64, 4, hipMemcpyDeviceToHost));
  if (!flag) return NC_OK;
  NC_HIP(hipMemset(sim->d_chunk_off + 64, 0, 4));
  return fail(NC_ERR_STATE, "a grid barrier of the persistent launch timed out: the steps since the last check "
                            "are invalid");
}

template <class T>
nc_status read_pos_array(nc_sim* sim, NetRec& r, const T* dev, T* out, uint32_t n) {
  if (!out) return fail(NC_ERR_ARG, "null argument");
  if (n > r.n) return fail(NC_ERR_RANGE, "more neurons requested than the network has");
  std::vector<T> tmp((size_t)r.nq * 64);
  NC_HIP(hipMemcpyAsync(tmp.data(), dev + (size_t)r.q0 * 64, tmp.size() * sizeof(T), hipMemcpyDeviceToHost,
                        sim->stream));
  NC_HIP(hipStreamSynchronize(sim->stream));
  {
    nc_status cb = check_barrier(sim);
    if (cb != NC_OK) return cb;
  }
  for (uint32_t i = 0; i < n; ++i) out[i] = tmp[r.maps->pos_of_id[i]];
  return NC_OK;
}

nc_status read_slots(nc_sim* sim, NetRec& r, std::vector<nc_slot>* hs) {
  NC_HIP(hipSetDevice(sim->device));
  hs->resize(r.nslots);
  NC_HIP(hipMemcpyAsync(hs->data(), sim->slots + r.slot0, r.nslots * sizeof(nc_slot), hipMemcpyDeviceToHost,
                        sim->stream));
  NC_HIP(hipStreamSynchronize(sim->stream));
  return check_barrier(sim);
}
}  // namespace

extern "C" nc_status nc_read_potential(nc_sim* sim, int32_t net_id, float* out, uint32_t n) try {
  NetRec* r;
  nc_status st = check_net(sim, net_id, &r);
  if (st != NC_OK) return st;
  NC_HIP(hipSetDevice(sim->device));
  return read_pos_array(sim, *r, sim->V, out, n);
}
NC_ABI_CATCH

extern "C" nc_status nc_read_last_fire(nc_sim* sim, int32_t net_id, int32_t* out, uint32_t n) try {
  NetRec* r;
  nc_status st = check_net(sim, net_id, &r);
  if (st != NC_OK) return st;
  NC_HIP(hipSetDevice(sim->device));
  return read_pos_array(sim, *r, sim->last_fire, out, n);
}
NC_ABI_CATCH

extern "C" nc_status nc_read_firings(nc_sim* sim, int32_t net_id, uint32_t* out, uint32_t n) try {
  NetRec* r;
  nc_status st = check_net(sim, net_id, &r);
  if (st != NC_OK) return st;
  NC_HIP(hipSetDevice(sim->device));
  return read_pos_array(sim, *r, sim->firings, out, n);
}
NC_ABI_CATCH

extern "C" nc_status nc_read_weights(nc_sim* sim, int32_t net_id, float* out, uint64_t n) try {
  NetRec* r;
  nc_status st = check_net(sim, net_id, &r);
  if (st != NC_OK) return st;
  if (!out) return fail(NC_ERR_ARG, "null argument");
  if (n > r->S) return fail(NC_ERR_RANGE, "more synapses requested than the network has");
  NC_HIP(hipSetDevice(sim->device));
  if (r->S == 0 || n == 0) return NC_OK;
  st = grow(&sim->d_wbuf, r->S, sim->stream);
  if (st != NC_OK) return st;
  return export_weights(sim, *r, sim->d_wbuf.p, out, n);
}
NC_ABI_CATCH

extern "C" nc_status nc_read_weights_noalloc(nc_sim* sim, int32_t net_id, float* out, uint64_t n) try {
  NetRec* r;
  nc_status st = check_net(sim, net_id, &r);
  if (st != NC_OK) return st;
  if (!out && n) return fail(NC_ERR_ARG, "null argument");
  if (n > r->S) return fail(NC_ERR_RANGE, "more synapses requested than the network has");
  if (r->S > NC_SCRATCH_FLOATS)
    return fail(NC_ERR_RANGE, "the network has more synapses than the static scratch holds (16777216): "
                              "use nc_read_weights");
  float* const scr = sim->device < kMaxDevices ? g_scratch_ptr[sim->device] : nullptr;
  if (!scr) return fail(NC_ERR_STATE, "the static scratch of this device is not set up");
  NC_HIP(hipSetDevice(sim->device));
  if (r->S == 0 || n == 0) return NC_OK;
  // the export kernel writes every one of the network's S entries before
  // the copy reads them: nothing needs clearing first
  return export_weights(sim, *r, scr, out, n);
}
NC_ABI_CATCH

extern "C" nc_status nc_write_weights(nc_sim* sim, int32_t net_id, const float* w, uint64_t n) try {
  NetRec* r;
  nc_status st = check_net(sim, net_id, &r);
  if (st != NC_OK) return st;
  if (!w) return fail(NC_ERR_ARG, "null argument");
  if (n > r->S) return fail(NC_ERR_RANGE, "more synapses given than the network has");
  std::vector<nc_slot> hs;
  st = read_slots(sim, *r, &hs);
  if (st != NC_OK) return st;
  // Synapse::setWeight also re-derives the sign class (src/NeuCor.cpp:576-579).
  for (uint64_t k = 0; k < n; ++k) {
    nc_slot& x = hs[r->maps->slot_of_syn[k]];
    x.w = w[k];
    x.arr_inh = (x.arr_inh & ~1) | ((w[k] < 0.0f) ? 1 : 0);
  }
  r->whost = r->steps_done + 1;
  NC_HIP(hipMemcpyAsync(sim->slots + r->slot0, hs.data(), r->nslots * sizeof(nc_slot), hipMemcpyHostToDevice,
                        sim->stream));
  NC_HIP(hipStreamSynchronize(sim->stream));
  return NC_OK;
}
NC_ABI_CATCH

extern "C" nc_status nc_set_weights(nc_sim* sim, int32_t net_id, uint64_t n, const uint64_t* syn, const float* w) try {
  NetRec* r;
  nc_status st = check_net(sim, net_id, &r);
  if (st != NC_OK) return st;
  if (n && (!syn || !w)) return fail(NC_ERR_ARG, "null argument");
  for (uint64_t i = 0; i < n; ++i)
    if (syn[i] >= r->S) return fail(NC_ERR_RANGE, "synapse index out of range");
  std::vector<nc_slot> hs;
  st = read_slots(sim, *r, &hs);
  if (st != NC_OK) return st;
  for (uint64_t i = 0; i < n; ++i) {
    nc_slot& x = hs[r->maps->slot_of_syn[syn[i]]];
    x.w = w[i];
    x.arr_inh = (x.arr_inh & ~1) | ((w[i] < 0.0f) ? 1 : 0);
  }
  r->whost = r->steps_done + 1;
  NC_HIP(hipMemcpyAsync(sim->slots + r->slot0, hs.data(), r->nslots * sizeof(nc_slot), hipMemcpyHostToDevice,
                        sim->stream));
  NC_HIP(hipStreamSynchronize(sim->stream));
  return NC_OK;
}
NC_ABI_CATCH

extern "C" nc_status nc_read_synapse_state(nc_sim* sim, int32_t net_id, int32_t* arrival,
                                           int32_t* last_arrival, float* depol, uint64_t n) try {
  NetRec* r;
  nc_status st = check_net(sim, net_id, &r);
  if (st != NC_OK) return st;
  if (n > r->S) return fail(NC_ERR_RANGE, "more synapses requested than the network has");
  std::vector<nc_slot> hs;
  st = read_slots(sim, *r, &hs);
  if (st != NC_OK) return st;
  // Synapse::AP_fireTime: the latched spike's arrival until its latch
  // releases (R steps after arrival), then 0.
  const int32_t s = (int32_t)r->steps_done;
  for (uint64_t k = 0; k < n; ++k) {
    const uint64_t j = r->maps->slot_of_syn[k];
    const int32_t a = hs[j].arr_inh >> 1;
    if (arrival) arrival[k] = (a != NC_NEVER && s < a + sim->T.r_rel) ? a : 0;
    // lastSpikeArrival: the latched spike's arrival once it has arrived, else
    // last_arr (the one before it, which fan-out moved there, or a last
    // arrival restored without a latched spike)
    if (last_arrival) last_arrival[k] = (a != NC_NEVER && a <= s) ? a : hs[j].last_arr;
    if (depol) depol[k] = hs[j].depol;
  }
  return NC_OK;
}
NC_ABI_CATCH

extern "C" nc_status nc_read_raster(nc_sim* sim, int32_t net_id, uint32_t* steps, uint32_t* ids,
                                    uint64_t cap, uint64_t* n_out) try {
  NetRec* r;
  nc_status st = check_net(sim, net_id, &r);
  if (st != NC_OK) return st;
  if (!n_out) return fail(NC_ERR_ARG, "null argument");
  if (!sim->raster_cap) return fail(NC_ERR_STATE, "raster logging is off (raster_capacity = 0)");
  NC_HIP(hipSetDevice(sim->device));
  uint32_t cnt = 0;
  NC_HIP(hipMemcpyAsync(&cnt, sim->raster_cnt, 4, hipMemcpyDeviceToHost, sim->stream));
  NC_HIP(hipStreamSynchronize(sim->stream));
  if (cnt > sim->raster_cap) return fail(NC_ERR_CAPACITY, "raster log overflowed; read it more often");
  std::vector<uint32_t> buf(2ull * cnt);
  if (cnt) NC_HIP(hipMemcpy(buf.data(), sim->raster, 8ull * cnt, hipMemcpyDeviceToHost));
  // Keep other networks' records; hand back this network's, sorted.
  std::vector<std::pair<uint32_t, uint32_t>> mine, rest;
  const uint32_t p0 = r->q0 * 64, p1 = p0 + r->nq * 64;
  for (uint32_t k = 0; k < cnt; ++k) {
    uint32_t pos = buf[2 * k], s = buf[2 * k + 1];
    if (pos >= p0 && pos < p1) mine.push_back({s, r->maps->id_of_pos[pos - p0]});
    else rest.push_back({pos, s});
  }
  std::sort(mine.begin(), mine.end());
  if (mine.size() > cap) return fail(NC_ERR_CAPACITY, "output buffer too small for the raster");
  for (size_t k = 0; k < mine.size(); ++k) {
    if (steps) steps[k] = mine[k].first;
    if (ids) ids[k] = mine[k].second;
  }
  *n_out = mine.size();
  std::vector<uint32_t> back;
  for (auto& x : rest) {
    back.push_back(x.first);
    back.push_back(x.second);
  }
  uint32_t nrest = (uint32_t)rest.size();
  if (nrest) NC_HIP(hipMemcpy(sim->raster, back.data(), back.size() * 4, hipMemcpyHostToDevice));
  NC_HIP(hipMemcpy(sim->raster_cnt, &nrest, 4, hipMemcpyHostToDevice));
  return NC_OK;
}
NC_ABI_CATCH

// ---------------------------------------------------------------------------
// Checkpoints: a network's dynamic state in network-local order.
//   header  : magic 'NCK1', version 1, n, S, steps_done, lr, pre_f, post_f,
//             dt, wiring digest, queued host events
//   neurons : V f32, lastFire i32 (step), firings u32          (id order)
//   synapses: w f32, depol f32, lastArrival i32, arrival i32   (export order)
//   events  : step offset u32, id u32, kind u32                (queued forced fires)
// Active masks, their summaries, the pending lists and the fan-out latches
// are functions of (arrival, steps_done) and are rebuilt on restore.
namespace {
struct nc_ckpt_header {
  uint32_t magic, version, n, pad0;
  uint64_t n_syn;
  uint32_t steps_done;
  float lr, pre_f, post_f, dt;
  uint32_t n_events;
  uint64_t digest;
};
static_assert(sizeof(nc_ckpt_header) == NC_CKPT_HEADER_BYTES, "checkpoint header layout");
static_assert(offsetof(nc_ckpt_header, steps_done) == NC_CKPT_STEPS_OFFSET, "checkpoint header layout");
}  // namespace

extern "C" nc_status nc_sim_save(nc_sim* sim, int32_t net_id, void* buf, uint64_t cap, uint64_t* size) try {
  NetRec* r;
  nc_status st = check_net(sim, net_id, &r);
  if (st != NC_OK) return st;
  if (!size) return fail(NC_ERR_ARG, "null size");
  const uint64_t need = sizeof(nc_ckpt_header) + 12ull * r->n + 16ull * r->S + 12ull * r->events.size();
  *size = need;
  if (!buf) return NC_OK;
  if (cap < need) return fail(NC_ERR_RANGE, "checkpoint buffer too small");
  std::vector<nc_slot> hs;
  st = read_slots(sim, *r, &hs);
  if (st != NC_OK) return st;
  unsigned char* p = (unsigned char*)buf;
  nc_ckpt_header h{};
  h.magic = 0x314b434eu;  // "NCK1"
  h.version = 1;
  h.n = r->n;
  h.n_syn = r->S;
  h.steps_done = r->steps_done;
  h.lr = r->lr, h.pre_f = r->pre_f, h.post_f = r->post_f, h.dt = sim->T.dt;
  h.n_events = (uint32_t)r->events.size();
  h.digest = r->maps->digest;
  memcpy(p, &h, sizeof h);
  p += sizeof h;
  std::vector<float> V(r->n);
  std::vector<int32_t> lf(r->n);
  std::vector<uint32_t> fi(r->n);
  st = read_pos_array(sim, *r, sim->V, V.data(), r->n);
  if (st == NC_OK) st = read_pos_array(sim, *r, sim->last_fire, lf.data(), r->n);
  if (st == NC_OK) st = read_pos_array(sim, *r, sim->firings, fi.data(), r->n);
  if (st != NC_OK) return st;
  for (uint32_t i = 0; i < r->n; ++i) {
    memcpy(p, &V[i], 4), memcpy(p + 4, &lf[i], 4), memcpy(p + 8, &fi[i], 4);
    p += 12;
  }
  for (uint64_t k = 0; k < r->S; ++k) {
    const nc_slot& x = hs[r->maps->slot_of_syn[k]];
    const int32_t arr = x.arr_inh >> 1;
    const int32_t last = (arr != NC_NEVER && arr <= (int32_t)r->steps_done) ? arr : x.last_arr;  // lastSpikeArrival
    memcpy(p, &x.w, 4), memcpy(p + 4, &x.depol, 4), memcpy(p + 8, &last, 4), memcpy(p + 12, &arr, 4);
    p += 16;
  }
  for (auto& ev : r->events) {
    const uint32_t off = ev.first - r->steps_done, pos = ev.second.first - r->q0 * 64;
    const uint32_t id = r->maps->id_of_pos[pos], kind = ev.second.second;
    memcpy(p, &off, 4), memcpy(p + 4, &id, 4), memcpy(p + 8, &kind, 4);
    p += 12;
  }
  return NC_OK;
}
NC_ABI_CATCH

extern "C" nc_status nc_sim_restore(nc_sim* sim, int32_t net_id, const void* buf, uint64_t size) try {
  NetRec* r;
  nc_status st = check_net(sim, net_id, &r);
  if (st != NC_OK) return st;
  if (!buf || size < sizeof(nc_ckpt_header)) return fail(NC_ERR_ARG, "not a checkpoint");
  nc_ckpt_header h;
  memcpy(&h, buf, sizeof h);
  if (h.magic != 0x314b434eu || h.version != 1) return fail(NC_ERR_ARG, "not a checkpoint (magic/version)");
  if (h.n != r->n || h.n_syn != r->S || h.digest != r->maps->digest)
    return fail(NC_ERR_ARG, "checkpoint of a different network");
  if (h.dt != sim->T.dt) return fail(NC_ERR_ARG, "checkpoint taken at another dt");
  if (r->steps_done != 0) return fail(NC_ERR_STATE, "restore into a network that has not stepped yet");
  if (size != sizeof(nc_ckpt_header) + 12ull * h.n + 16ull * h.n_syn + 12ull * h.n_events)
    return fail(NC_ERR_ARG, "checkpoint size does not match its header");
  if (h.steps_done >= (uint32_t)NC_MAX_STEPS) return fail(NC_ERR_RANGE, "checkpoint step count out of range");
  const unsigned char* p = (const unsigned char*)buf + sizeof h;
  NC_HIP(hipSetDevice(sim->device));
  NC_HIP(hipStreamSynchronize(sim->stream));
  const int32_t s = (int32_t)h.steps_done;
  const int32_t R = sim->T.r_rel;
  const uint32_t npos = r->nq * 64, pos0 = r->q0 * 64;
  // neurons
  std::vector<float> V((size_t)npos, NC_BASELEVEL);
  std::vector<int32_t> lf((size_t)npos, NC_NEVER);
  std::vector<uint32_t> fi((size_t)npos, 0u), asum((size_t)npos, 0u);
  NC_HIP(hipMemcpy(V.data(), sim->V + pos0, npos * 4, hipMemcpyDeviceToHost));
  // Step values a checkpoint may hold: NC_NEVER, or a step of the run so far
  // (arrivals: up to the longest delay ahead); anything else would index the
  // step tables out of range.
  auto past = [s](int32_t v) { return v == NC_NEVER || (v >= 1 && v <= s); };
  for (uint32_t i = 0; i < h.n; ++i) {
    const uint32_t pl = r->maps->pos_of_id[i];
    memcpy(&V[pl], p, 4), memcpy(&lf[pl], p + 4, 4), memcpy(&fi[pl], p + 8, 4);
    if (!past(lf[pl])) return fail(NC_ERR_ARG, "checkpoint last fire out of range");
    p += 12;
  }
  // synapses, and from their arrivals the active masks, summaries, pending
  // lists and fan-out latches
  std::vector<nc_slot> hs(r->nslots);
  NC_HIP(hipMemcpy(hs.data(), sim->slots + r->slot0, r->nslots * sizeof(nc_slot), hipMemcpyDeviceToHost));
  std::vector<uint32_t> masks(r->nmwords, 0u);
  // the active-slot cache: each position's first NC_ACACHE active slots,
  // {depol 0, depol 1, depol 2, low bytes of their arrivals}
  static_assert(NC_ACACHE == 3, "three depolarisations and their arrivals' low bytes per 16-byte record");
  std::vector<uint4> cache((size_t)npos, make_uint4(0u, 0u, 0u, 0u));
  std::vector<uint8_t> ncache((size_t)npos, 0u);
  std::vector<std::vector<uint2>> pend(r->nq);
  std::vector<int32_t> edge_arr(r->S);
  for (uint32_t t = 0; t < h.n; ++t) {
    const uint32_t pl = r->maps->pos_of_id[t], q = pl / 64, lane = pl % 64;
    const uint32_t deg = (uint32_t)(r->maps->in_ptr[t + 1] - r->maps->in_ptr[t]);
    const uint32_t nmw = (deg + 31) / 32, G = nmw <= 32 ? 1u : (nmw + 31) / 32;
    for (uint32_t k = 0; k < deg; ++k) {
      const uint64_t syn = (uint64_t)r->maps->in_ptr[t] + k;
      const unsigned char* e = p + 16ull * syn;
      nc_slot& x = hs[r->maps->slot_of_syn[syn]];
      int32_t arr;
      memcpy(&x.w, e, 4), memcpy(&x.depol, e + 4, 4), memcpy(&x.last_arr, e + 8, 4), memcpy(&arr, e + 12, 4);
      if (!past(x.last_arr)) return fail(NC_ERR_ARG, "checkpoint last arrival out of range");
      if (arr != NC_NEVER && (arr < 1 || arr > s + 255)) return fail(NC_ERR_ARG, "checkpoint arrival out of range");
      x.arr_inh = arr * 2 + (x.arr_inh & 1);
      edge_arr[syn] = arr;
      if (arr == NC_NEVER) continue;
      if (arr <= s && s < arr + R) {
        const uint64_t word = sim->slice_moff_host[r->q0 + q] + (k >> 5) * 64 + lane - r->mask0;
        masks[word] |= 1u << (k & 31);
        asum[pl] |= 1u << ((k >> 5) / G);
        uint32_t dbits;
        memcpy(&dbits, &x.depol, 4);
        if (ncache[pl] < NC_ACACHE) {
          uint4& c = cache[pl];
          const uint32_t e = ncache[pl]++;
          (e == 0 ? c.x : e == 1 ? c.y : c.z) = dbits;
          c.w |= ((uint32_t)arr & 255u) << (8 * e);
        }
      } else if (arr > s) {
        pend[q].push_back(make_uint2((uint32_t)r->slot0 + r->maps->slot_of_syn[syn], (uint32_t)arr));
      }
    }
  }
  p += 16ull * h.n_syn;
  std::vector<std::pair<uint32_t, std::pair<uint32_t, uint8_t>>> events;
  for (uint32_t i = 0; i < h.n_events; ++i) {
    uint32_t off, id, kind;
    memcpy(&off, p, 4), memcpy(&id, p + 4, 4), memcpy(&kind, p + 8, 4);
    p += 12;
    if (id >= r->n || off < 1 || kind > 1 || off >= (uint32_t)NC_MAX_STEPS - h.steps_done)
      return fail(NC_ERR_ARG, "checkpoint event out of range");
    events.push_back({h.steps_done + off, {r->pos(id), (uint8_t)kind}});
  }
  r->events = events;
  st = reserve_events(sim);
  if (st != NC_OK) {
    r->events.clear();
    return st;
  }
  // upload
  hipStream_t sq = sim->stream;
  NC_HIP(hipMemcpyAsync(sim->V + pos0, V.data(), npos * 4, hipMemcpyHostToDevice, sq));
  NC_HIP(hipMemcpyAsync(sim->last_fire + pos0, lf.data(), npos * 4, hipMemcpyHostToDevice, sq));
  NC_HIP(hipMemcpyAsync(sim->firings + pos0, fi.data(), npos * 4, hipMemcpyHostToDevice, sq));
  NC_HIP(hipMemcpyAsync(sim->asum + pos0, asum.data(), npos * 4, hipMemcpyHostToDevice, sq));
  NC_HIP(hipMemcpyAsync(sim->slots + r->slot0, hs.data(), r->nslots * sizeof(nc_slot), hipMemcpyHostToDevice, sq));
  NC_HIP(hipMemcpyAsync(sim->amask + r->mask0, masks.data(), r->nmwords * 4, hipMemcpyHostToDevice, sq));
  NC_HIP(hipMemcpyAsync(sim->acache + pos0, cache.data(), npos * sizeof(uint4), hipMemcpyHostToDevice, sq));
  // each slice's spikes in flight into the pending list of their arrival
  // window, entries relative to the slice's first slot
  std::vector<uint2> cnt2(r->nq);
  std::vector<std::vector<uint32_t>> lists(2 * (size_t)r->nq);
  for (uint32_t q = 0; q < r->nq; ++q) {
    const uint32_t first = sim->slice_off_host[r->q0 + q];
    for (const uint2& e : pend[q])
      lists[2 * q + ((e.y >> NC_PEND_LOG_W) & 1u)].push_back(((e.x - first) << 8) | (e.y & 255u));
    cnt2[q] = make_uint2((uint32_t)lists[2 * q].size(), (uint32_t)lists[2 * q + 1].size());
    for (uint32_t k = 0; k < 2; ++k) {
      const std::vector<uint32_t>& l = lists[2 * q + k];
      if (!l.empty())
        NC_HIP(hipMemcpyAsync((k ? sim->pend1 : sim->pend0) + first, l.data(), l.size() * 4, hipMemcpyHostToDevice,
                              sq));
    }
  }
  NC_HIP(hipMemcpyAsync(sim->pend_cnt + r->q0, cnt2.data(), r->nq * sizeof(uint2), hipMemcpyHostToDevice, sq));
  std::vector<nc_edge> ed(r->S);
  if (r->S) NC_HIP(hipMemcpy(ed.data(), sim->edges + r->out0, r->S * sizeof(nc_edge), hipMemcpyDeviceToHost));
  for (uint64_t k = 0; k < r->S; ++k) ed[r->maps->edge_of_syn[k]].arr = edge_arr[k];
  if (r->S) NC_HIP(hipMemcpyAsync(sim->edges + r->out0, ed.data(), r->S * sizeof(nc_edge), hipMemcpyHostToDevice, sq));
  NC_HIP(hipStreamSynchronize(sq));
  r->steps_done = h.steps_done;
  r->lr = h.lr, r->pre_f = h.pre_f, r->post_f = h.post_f;
  return NC_OK;
}
NC_ABI_CATCH

extern "C" nc_status nc_reset_activities(nc_sim* sim, int32_t net_id) try {
  NetRec* r;
  nc_status st = check_net(sim, net_id, &r);
  if (st != NC_OK) return st;
  NC_HIP(hipSetDevice(sim->device));
  NC_HIP(hipMemsetAsync(sim->firings + (size_t)r->q0 * 64, 0, (size_t)r->nq * 64 * 4, sim->stream));
  NC_HIP(hipStreamSynchronize(sim->stream));
  return NC_OK;
}
NC_ABI_CATCH

extern "C" nc_status nc_read_counters(nc_sim* sim, uint64_t* out, uint32_t n) try {
  if (!sim || (!out && n)) return fail(NC_ERR_ARG, "null argument");
  if (n > NC_NUM_COUNTERS) return fail(NC_ERR_RANGE, "at most NC_NUM_COUNTERS counters");
  NC_HIP(hipSetDevice(sim->device));
  std::vector<unsigned long long> h(sim->ctr_rows * NC_NCTR);
  NC_HIP(hipMemcpyAsync(h.data(), sim->ctr, h.size() * 8, hipMemcpyDeviceToHost, sim->stream));
  NC_HIP(hipStreamSynchronize(sim->stream));
  uint64_t all[NC_NUM_COUNTERS] = {};
  for (int k = 0; k < NC_NCTR && k < NC_NUM_COUNTERS; ++k) {
    uint64_t v = 0;
    for (uint64_t r = 0; r < sim->ctr_rows; ++r) v += h[r * NC_NCTR + k];
    all[k] = v;
  }
  all[0] = sim->host_neuron_steps;
  all[1] = sim->host_synapse_steps;
  all[9] = sim->host_mword_steps;
  for (uint32_t k = 0; k < n; ++k) out[k] = all[k];
  return NC_OK;
}
NC_ABI_CATCH

namespace {
// The simulation's state at a step boundary, for the autotuner: a device copy
// of every buffer the step kernels write, and the host-side step counts and
// queued events.  Restoring it puts the simulation back exactly where it was.
struct StateCopy {
  struct Buf {
    void* dev;
    void* copy;
    size_t bytes;
  };
  std::vector<Buf> bufs;
  struct NetState {
    uint32_t steps_done, steps_left;
    std::vector<std::pair<uint32_t, std::pair<uint32_t, uint8_t>>> events;
  };
  std::vector<NetState> nets;
  uint64_t gstep = 0, timing_step = 0, hn = 0, hs = 0, hm = 0;
  ~StateCopy() {
    for (Buf& b : bufs)
      if (b.copy) (void)hipFree(b.copy);
  }
};

nc_status state_save(nc_sim* sim, StateCopy* c) {
  const uint64_t P = sim->max_pos, Q = P / 64;
  const std::pair<void*, size_t> regions[] = {
      {sim->slots, sim->max_slots * sizeof(nc_slot)},
      {sim->amask, sim->max_mwords * 4},
      {sim->asum, P * 4},
      {sim->V, P * 4},
      {sim->last_fire, P * 4},
      {sim->firings, P * 4},
      {sim->edges, sim->max_out * sizeof(nc_edge)},
      {sim->pend0, sim->max_slots * 4},
      {sim->pend1, sim->max_slots * 4},
      {sim->pend_cnt, Q * sizeof(uint2)},
      {sim->bcnt, Q * NC_NB * 4ull},
      {sim->bent, Q * NC_NB * NC_BCAP * sizeof(uint2)},
      {sim->acache, P * sizeof(uint4)},
      {sim->fired, (Q + NC_FLISTS * 8ull) * NC_FSPLIT * sizeof(uint4)},
      {sim->fired_cnt, 2ull * NC_FLISTS * NC_FCNT_STRIDE * 4},
      {sim->forced_in, (P / 32) * 4},
      {sim->forced_bg, (P / 32) * 4},
      {sim->raster_cnt, 64},
      {sim->ctr, sim->ctr_rows * NC_NCTR * 8},
  };
  for (const auto& r : regions) {
    StateCopy::Buf b{r.first, nullptr, r.second};
    NC_HIP(hipMalloc(&b.copy, b.bytes));
    c->bufs.push_back(b);
    NC_HIP(hipMemcpyAsync(b.copy, b.dev, b.bytes, hipMemcpyDeviceToDevice, sim->stream));
  }
  for (const NetRec& r : sim->nets) c->nets.push_back({r.steps_done, r.steps_left, r.events});
  c->gstep = sim->gstep, c->timing_step = sim->timing_step;
  c->hn = sim->host_neuron_steps, c->hs = sim->host_synapse_steps, c->hm = sim->host_mword_steps;
  NC_HIP(hipStreamSynchronize(sim->stream));
  return NC_OK;
}

nc_status state_restore(nc_sim* sim, const StateCopy& c) {
  for (const StateCopy::Buf& b : c.bufs)
    NC_HIP(hipMemcpyAsync(b.dev, b.copy, b.bytes, hipMemcpyDeviceToDevice, sim->stream));
  for (size_t i = 0; i < c.nets.size() && i < sim->nets.size(); ++i) {
    sim->nets[i].steps_done = c.nets[i].steps_done;
    sim->nets[i].steps_left = c.nets[i].steps_left;
    sim->nets[i].events = c.nets[i].events;
  }
  sim->gstep = c.gstep, sim->timing_step = c.timing_step;
  sim->host_neuron_steps = c.hn, sim->host_synapse_steps = c.hs, sim->host_mword_steps = c.hm;
  return NC_OK;
}

nc_status check_launch(const nc_launch_cfg& c) {
  const uint32_t w = c.waves_per_block;
  if (w != 1 && w != 2 && w != 4 && w != 8) return fail(NC_ERR_ARG, "waves_per_block must be 1, 2, 4 or 8");
  if (c.xcd_swizzle > 1) return fail(NC_ERR_ARG, "xcd_swizzle must be 0 or 1");
  if (c.graph_steps > NC_GRING) return fail(NC_ERR_ARG, "graph_steps must be at most 1024");
  if (c.stage_tables > 1) return fail(NC_ERR_ARG, "stage_tables must be 0 or 1");
  if (c.fanout_waves && (c.fanout_waves < 64 || c.fanout_waves > (1u << 20)))
    return fail(NC_ERR_ARG, "fanout_waves must be 0 or in [64, 2^20]");
  if (c.persistent > 1) return fail(NC_ERR_ARG, "persistent must be 0 or 1");
  return NC_OK;
}
}  // namespace

extern "C" nc_status nc_sim_set_launch(nc_sim* sim, const nc_launch_cfg* cfg) try {
  if (!sim || !cfg) return fail(NC_ERR_ARG, "null argument");
  nc_status st = check_launch(*cfg);
  if (st != NC_OK) return st;
  sim->cfg = *cfg;
  return NC_OK;
}
NC_ABI_CATCH

extern "C" nc_status nc_sim_get_launch(const nc_sim* sim, nc_launch_cfg* cfg) try {
  if (!sim || !cfg) return fail(NC_ERR_ARG, "null argument");
  *cfg = sim->cfg;
  return NC_OK;
}
NC_ABI_CATCH

extern "C" nc_status nc_sim_tune_list(nc_sim* sim, const nc_launch_cfg* cands, uint32_t n_cands,
                                      uint32_t steps_per_trial, uint32_t rounds, nc_launch_cfg* chosen,
                                      float* ms_out) try {
  if (!sim || !cands || n_cands == 0 || steps_per_trial == 0 || rounds == 0)
    return fail(NC_ERR_ARG, "null sim, no candidates, zero steps or zero rounds");
  for (uint32_t i = 0; i < n_cands; ++i) {
    nc_status cs = check_launch(cands[i]);
    if (cs != NC_OK) return cs;
  }
  NC_HIP(hipSetDevice(sim->device));
  if (sim->upload_pending) {
    NC_HIP(hipEventSynchronize(sim->upload_done));
    sim->upload_pending = false;
  }
  NC_HIP(hipStreamSynchronize(sim->stream));
  StateCopy saved;
  nc_status st = state_save(sim, &saved);
  if (st != NC_OK) return st;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  NC_HIP(hipEventCreate(&e0));
  if (hipEventCreate(&e1) != hipSuccess) {
    (void)hipEventDestroy(e0);
    return fail(NC_ERR_HIP, "hipEventCreate failed");
  }
  const nc_launch_cfg start = sim->cfg;
  const int timing = sim->kernel_timing;
  sim->kernel_timing = 0;  // events on the launches would keep graphs out
  std::vector<std::vector<float>> ms((size_t)n_cands);
  // Every trial starts from the saved state: a warm-up trial (which also
  // captures the candidate's graph), the state put back, then the timed
  // trial.  Rounds visit the candidates in rotated order.
  for (uint32_t rd = 0; rd < rounds && st == NC_OK; ++rd)
    for (uint32_t k = 0; k < n_cands && st == NC_OK; ++k) {
      const uint32_t i = (k + rd * (n_cands / rounds + 1)) % n_cands;
      sim->cfg = cands[i];
      st = state_restore(sim, saved);
      if (st == NC_OK) st = nc_sim_step(sim, steps_per_trial);
      if (st == NC_OK) st = state_restore(sim, saved);
      if (st != NC_OK) break;
      if (hipEventRecord(e0, sim->stream) != hipSuccess) {
        st = fail(NC_ERR_HIP, "hipEventRecord failed");
        break;
      }
      st = nc_sim_step(sim, steps_per_trial);
      if (st != NC_OK) break;
      float t = 0.0f;
      if (hipEventRecord(e1, sim->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
          hipEventElapsedTime(&t, e0, e1) != hipSuccess) {
        st = fail(NC_ERR_HIP, "event timing failed");
        break;
      }
      ms[i].push_back(t);
    }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  sim->kernel_timing = timing;
  nc_status rs = state_restore(sim, saved);
  if (rs == NC_OK && hipStreamSynchronize(sim->stream) != hipSuccess) rs = fail(NC_ERR_HIP, "synchronize failed");
  if (st == NC_OK) st = rs;
  if (st != NC_OK) {
    sim->cfg = start;
    return st;
  }
  uint32_t best = 0;
  float best_ms = -1.0f;
  for (uint32_t i = 0; i < n_cands; ++i) {
    std::vector<float> v = ms[i];
    std::sort(v.begin(), v.end());
    const float med = v.empty() ? 0.0f : v[v.size() / 2];
    if (ms_out) ms_out[i] = med;
    if (best_ms < 0.0f || med < best_ms) {
      best_ms = med;
      best = i;
    }
  }
  sim->cfg = cands[best];
  if (chosen) *chosen = sim->cfg;
  return NC_OK;
}
NC_ABI_CATCH

extern "C" nc_status nc_sim_tune(nc_sim* sim, uint32_t steps_per_trial, nc_launch_cfg* chosen) try {
  if (!sim || steps_per_trial == 0) return fail(NC_ERR_ARG, "null sim or zero steps");
  std::vector<nc_launch_cfg> cands;
  const uint32_t waves[4] = {1, 2, 4, 8};
  const uint32_t graphs[2] = {0, std::min<uint32_t>(steps_per_trial, NC_GRING)};
  for (uint32_t g : graphs)
    for (uint32_t w : waves)
      for (uint32_t x = 0; x < 4; ++x) cands.push_back(nc_launch_cfg{w, x & 1u, g, x >> 1, sim->cfg.fanout_waves, 0});
  // the persistent form (one launch per step): 2, 4 or 8 slices per workgroup
  for (uint32_t g : graphs)
    for (uint32_t w : {2u, 4u, 8u}) cands.push_back(nc_launch_cfg{w, 0, g, 0, sim->cfg.fanout_waves, 1});
  return nc_sim_tune_list(sim, cands.data(), (uint32_t)cands.size(), steps_per_trial, 3, chosen, nullptr);
}
NC_ABI_CATCH

extern "C" nc_status nc_set_counting(nc_sim* sim, int enable) try {
  if (!sim) return fail(NC_ERR_ARG, "null sim");
  sim->counting = enable != 0;
  return NC_OK;
}
NC_ABI_CATCH

extern "C" nc_status nc_sim_synchronize(nc_sim* sim) try {
  if (!sim) return fail(NC_ERR_ARG, "null sim");
  NC_HIP(hipSetDevice(sim->device));
  NC_HIP(hipStreamSynchronize(sim->stream));
  return check_barrier(sim);
}
NC_ABI_CATCH

extern "C" nc_status nc_set_kernel_timing(nc_sim* sim, int kernel_mask, uint32_t every) try {
  if (!sim) return fail(NC_ERR_ARG, "null sim");
  sim->kernel_timing = kernel_mask & 0xF;
  sim->timing_every = every ? every : 1;
  sim->timing_step = 0;
  for (int k = 0; k < 4; ++k) {
    sim->k_launch[k] = 0;
    sim->k_ms[k] = 0.0;
  }
  return NC_OK;
}
NC_ABI_CATCH

extern "C" nc_status nc_read_kernel_timing(nc_sim* sim, uint32_t kernel, uint64_t* launches,
                                           double* total_ms) try {
  if (!sim) return fail(NC_ERR_ARG, "null sim");
  if (kernel > 3) return fail(NC_ERR_RANGE, "kernel index out of range");
  if (launches) *launches = sim->k_launch[kernel];
  if (total_ms) *total_ms = sim->k_ms[kernel];
  return NC_OK;
}
NC_ABI_CATCH
