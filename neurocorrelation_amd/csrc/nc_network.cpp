// Host-side network construction: the workload wiring rule (spatial hash,
// O(N*K) instead of the reference's O(N^2) all-pairs scan in
// Neuron::makeConnections, src/NeuCor.cpp:418-444), explicit networks
// (NeuCor::createNeuron / createSynapse, src/NeuCor.cpp:154-195) and input
// selection.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <numeric>
#include <thread>

#include "nc_common.h"
#include "nc_internal.h"

namespace nc {

static thread_local std::string g_error;

void set_error(const std::string& msg) { g_error = msg; }

nc_status fail(nc_status code, const std::string& msg) {
  g_error = msg;
  return code;
}

void finish_net_stats(nc_net* net) {
  net->max_in_degree = 0;
  for (uint32_t i = 0; i < net->n; ++i) {
    uint64_t d = (uint64_t)(net->in_ptr[i + 1] - net->in_ptr[i]);
    if (d > net->max_in_degree) net->max_in_degree = (uint32_t)d;
  }
  net->max_delay = 0;
  for (uint8_t d : net->delay)
    if (d > net->max_delay) net->max_delay = d;
}

nc_status select_inputs(const nc_net& net, const nc_drive_spec& d, InputSel* out) {
  if (d.n_inputs > net.n) return fail(NC_ERR_ARG, "n_inputs exceeds the neuron count");
  if (d.n_correlated > d.n_inputs) return fail(NC_ERR_ARG, "n_correlated exceeds n_inputs");
  if (d.n_correlated > 0 && d.n_groups == 0)
    return fail(NC_ERR_ARG, "correlated inputs need n_groups > 0");
  if (d.n_inputs > 0 && d.period == 0) return fail(NC_ERR_ARG, "period must be > 0");
  std::vector<uint64_t> keyed(net.n);
  for (uint32_t i = 0; i < net.n; ++i) {
    uint32_t h = nc_draw(d.seed, i, 0u, NC_STREAM_INSEL).v[0];
    keyed[i] = ((uint64_t)h << 32) | i;
  }
  uint32_t k = d.n_inputs;
  if (k > 0 && k < net.n)
    std::nth_element(keyed.begin(), keyed.begin() + k, keyed.end());
  std::sort(keyed.begin(), keyed.begin() + k);
  out->ids.resize(k);
  out->group.resize(k);
  uint32_t ng = d.n_correlated > 0 ? d.n_groups : 0;
  for (uint32_t r = 0; r < k; ++r) {
    out->ids[r] = (uint32_t)(keyed[r] & 0xffffffffu);
    out->group[r] = r < d.n_correlated ? r % ng : ng + (r - d.n_correlated);
  }
  out->n_groups_total = ng + (k - d.n_correlated);
  return NC_OK;
}

}  // namespace nc

using nc::fail;

namespace nc {
nc_status check_wiring_spec(const nc_wiring_spec& s) {
  if (s.n_neurons == 0) return fail(NC_ERR_ARG, "n_neurons must be > 0");
  if (!(s.side > 0.0f) || !(s.radius > 0.0f)) return fail(NC_ERR_ARG, "side and radius must be > 0");
  if (s.delay_min < 1 || s.delay_max < s.delay_min || s.delay_max > 250)
    return fail(NC_ERR_ARG, "delays must satisfy 1 <= delay_min <= delay_max <= 250");
  if (!(s.w_max >= s.w_min)) return fail(NC_ERR_ARG, "w_max < w_min");
  return NC_OK;
}
}  // namespace nc

extern "C" nc_status nc_net_generate(const nc_wiring_spec* spec, nc_net** out) {
  if (!spec || !out) return fail(NC_ERR_ARG, "null argument");
  *out = nullptr;
  const nc_wiring_spec s = *spec;
  {
    nc_status st = nc::check_wiring_spec(s);
    if (st != NC_OK) return st;
  }
  const uint32_t n = s.n_neurons;

  nc_net* net = new (std::nothrow) nc_net();
  if (!net) return fail(NC_ERR_OOM, "host allocation failed");
  try {
    // Positions in draw order.
    std::vector<float> px(n), py(n), pz(n);
    for (uint32_t i = 0; i < n; ++i) {
      nc_u32x4 r = nc_draw(s.seed, i, 0u, NC_STREAM_POS);
      px[i] = nc_u01(r.v[0]) * s.side;
      py[i] = nc_u01(r.v[1]) * s.side;
      pz[i] = nc_u01(r.v[2]) * s.side;
    }
    const int ncell = nc_cells_per_axis(s.side, s.radius);
    std::vector<uint32_t> cx(n), cy(n), cz(n);
    for (uint32_t i = 0; i < n; ++i) {
      cx[i] = nc_cell_of(px[i], ncell, s.side);
      cy[i] = nc_cell_of(py[i], ncell, s.side);
      cz[i] = nc_cell_of(pz[i], ncell, s.side);
    }
    // Neuron ids: draw order, or Morton order of the cells (ties by draw index).
    std::vector<uint32_t> order(n);
    std::iota(order.begin(), order.end(), 0u);
    if (s.spatial_order) {
      std::vector<uint64_t> key(n);
      for (uint32_t i = 0; i < n; ++i) {
        uint64_t m = nc_part1by2(cx[i]) | (nc_part1by2(cy[i]) << 1) | (nc_part1by2(cz[i]) << 2);
        key[i] = (m << 32) | i;
      }
      std::sort(key.begin(), key.end());
      for (uint32_t i = 0; i < n; ++i) order[i] = (uint32_t)(key[i] & 0xffffffffu);
    }
    net->n = n;
    net->xyz.resize(3ull * n);
    std::vector<float> X(n), Y(n), Z(n);
    std::vector<uint32_t> cellid(n);
    for (uint32_t id = 0; id < n; ++id) {
      uint32_t d = order[id];
      X[id] = px[d]; Y[id] = py[d]; Z[id] = pz[d];
      net->xyz[3ull * id] = px[d];
      net->xyz[3ull * id + 1] = py[d];
      net->xyz[3ull * id + 2] = pz[d];
      cellid[id] = (cz[d] * (uint32_t)ncell + cy[d]) * (uint32_t)ncell + cx[d];
    }
    // Bucket ids by cell (ascending id inside a cell).
    const uint64_t ncells = (uint64_t)ncell * ncell * ncell;
    std::vector<uint32_t> cstart(ncells + 1, 0), cmem(n);
    for (uint32_t id = 0; id < n; ++id) cstart[cellid[id] + 1]++;
    for (uint64_t c = 0; c < ncells; ++c) cstart[c + 1] += cstart[c];
    {
      std::vector<uint32_t> fillp(cstart.begin(), cstart.end() - 1);
      for (uint32_t id = 0; id < n; ++id) cmem[fillp[cellid[id]]++] = id;
    }
    const float wspan = s.w_max - s.w_min;
    const float fdmax = (float)s.delay_max;

    unsigned nt = s.n_threads ? s.n_threads : std::thread::hardware_concurrency();
    if (nt == 0) nt = 1;
    if (nt > 64) nt = 64;
    if (n < 4096) nt = 1;
    struct Part {
      std::vector<uint32_t> deg, src;
      std::vector<float> w;
      std::vector<uint8_t> dly, inh;
    };
    std::vector<Part> parts(nt);
    auto work = [&](unsigned t) {
      uint32_t lo = (uint32_t)((uint64_t)n * t / nt), hi = (uint32_t)((uint64_t)n * (t + 1) / nt);
      Part& P = parts[t];
      P.deg.resize(hi - lo);
      std::vector<std::pair<uint32_t, float>> cand;
      for (uint32_t tg = lo; tg < hi; ++tg) {
        cand.clear();
        const uint32_t d = order[tg];
        const int x0 = (int)cx[d], y0 = (int)cy[d], z0 = (int)cz[d];
        for (int dz = -1; dz <= 1; ++dz) {
          int zz = z0 + dz;
          if (zz < 0 || zz >= ncell) continue;
          for (int dy = -1; dy <= 1; ++dy) {
            int yy = y0 + dy;
            if (yy < 0 || yy >= ncell) continue;
            for (int dx = -1; dx <= 1; ++dx) {
              int xx = x0 + dx;
              if (xx < 0 || xx >= ncell) continue;
              uint32_t c = ((uint32_t)zz * ncell + (uint32_t)yy) * ncell + (uint32_t)xx;
              for (uint32_t q = cstart[c]; q < cstart[c + 1]; ++q) {
                uint32_t src = cmem[q];
                if (src == tg) continue;
                float ddx = X[src] - X[tg], ddy = Y[src] - Y[tg], ddz = Z[src] - Z[tg];
                float d2 = ddx * ddx + ddy * ddy;
                d2 = d2 + ddz * ddz;
                float dist = sqrtf(d2);
                if (dist < s.radius) cand.push_back({src, dist});
              }
            }
          }
        }
        std::sort(cand.begin(), cand.end(),
                  [](const std::pair<uint32_t, float>& a, const std::pair<uint32_t, float>& b) {
                    return a.first < b.first;
                  });
        P.deg[tg - lo] = (uint32_t)cand.size();
        for (auto& c : cand) {
          float q = c.second / s.radius;
          float dl = ceilf(q * fdmax);
          uint32_t dd = dl < (float)s.delay_min ? s.delay_min : (dl > fdmax ? s.delay_max : (uint32_t)dl);
          nc_u32x4 r = nc_draw(s.seed, c.first, tg, NC_STREAM_WEIGHT);
          float w = s.w_min + nc_u01(r.v[0]) * wspan;
          uint8_t ih = nc_u01(r.v[1]) < s.inhib_frac ? 1 : 0;
          if (ih) w = -w;
          P.src.push_back(c.first);
          P.w.push_back(w);
          P.dly.push_back((uint8_t)dd);
          P.inh.push_back(ih);
        }
      }
    };
    if (nt == 1) {
      work(0);
    } else {
      std::vector<std::thread> th;
      for (unsigned t = 0; t < nt; ++t) th.emplace_back(work, t);
      for (auto& t : th) t.join();
    }
    uint64_t S = 0;
    for (auto& P : parts) S += P.src.size();
    net->in_ptr.resize(n + 1);
    net->in_src.resize(S);
    net->w.resize(S);
    net->delay.resize(S);
    net->inh.resize(S);
    uint64_t off = 0;
    uint32_t id = 0;
    net->in_ptr[0] = 0;
    for (auto& P : parts) {
      memcpy(net->in_src.data() + off, P.src.data(), P.src.size() * 4);
      memcpy(net->w.data() + off, P.w.data(), P.w.size() * 4);
      memcpy(net->delay.data() + off, P.dly.data(), P.dly.size());
      memcpy(net->inh.data() + off, P.inh.data(), P.inh.size());
      off += P.src.size();
      for (uint32_t dg : P.deg) {
        net->in_ptr[id + 1] = net->in_ptr[id] + dg;
        ++id;
      }
      std::vector<uint32_t>().swap(P.src);
      std::vector<float>().swap(P.w);
    }
    nc::finish_net_stats(net);
  } catch (const std::bad_alloc&) {
    delete net;
    return fail(NC_ERR_OOM, "host allocation failed while wiring");
  }
  *out = net;
  return NC_OK;
}

extern "C" nc_status nc_net_from_edges(uint32_t n, const float* xyz, uint64_t n_syn,
                                       const uint32_t* to, const uint32_t* from, const float* w,
                                       const uint8_t* delay_steps, const uint8_t* inhibitory,
                                       nc_net** out) {
  if (!out) return fail(NC_ERR_ARG, "null argument");
  *out = nullptr;
  if (n_syn > 0 && (!to || !from || !w || !delay_steps)) return fail(NC_ERR_ARG, "null edge array");
  for (uint64_t k = 0; k < n_syn; ++k) {
    if (to[k] >= n || from[k] >= n) return fail(NC_ERR_RANGE, "synapse endpoint out of range");
    if (to[k] == from[k]) return fail(NC_ERR_ARG, "self synapse");
    if (delay_steps[k] < 1 || delay_steps[k] > 250) return fail(NC_ERR_ARG, "delay must be in 1..250 steps");
  }
  nc_net* net = new (std::nothrow) nc_net();
  if (!net) return fail(NC_ERR_OOM, "host allocation failed");
  try {
    net->n = n;
    net->xyz.assign(3ull * n, 0.0f);
    if (xyz) memcpy(net->xyz.data(), xyz, 12ull * n);
    // Stable sort by (to, from); the first of duplicate pairs wins, as
    // NeuCor::createSynapse refuses an existing pair (src/NeuCor.cpp:189-191).
    std::vector<uint64_t> idx(n_syn);
    std::iota(idx.begin(), idx.end(), 0ull);
    std::stable_sort(idx.begin(), idx.end(), [&](uint64_t a, uint64_t b) {
      if (to[a] != to[b]) return to[a] < to[b];
      return from[a] < from[b];
    });
    net->in_ptr.assign(n + 1, 0);
    for (uint64_t q = 0; q < n_syn; ++q) {
      uint64_t k = idx[q];
      if (q > 0) {
        uint64_t p = idx[q - 1];
        if (to[p] == to[k] && from[p] == from[k]) continue;
      }
      net->in_src.push_back(from[k]);
      net->w.push_back(w[k]);
      net->delay.push_back(delay_steps[k]);
      net->inh.push_back(inhibitory ? (inhibitory[k] ? 1 : 0) : (w[k] < 0.0f ? 1 : 0));
      net->in_ptr[to[k] + 1]++;
    }
    for (uint32_t i = 0; i < n; ++i) net->in_ptr[i + 1] += net->in_ptr[i];
    nc::finish_net_stats(net);
  } catch (const std::bad_alloc&) {
    delete net;
    return fail(NC_ERR_OOM, "host allocation failed");
  }
  *out = net;
  return NC_OK;
}

extern "C" nc_status nc_net_info(const nc_net* net, uint32_t* n, uint64_t* n_syn,
                                 uint32_t* max_in_degree, uint32_t* max_delay) try {
  if (!net) return fail(NC_ERR_ARG, "null network");
  if (n) *n = net->n;
  if (n_syn) *n_syn = net->n_syn();
  if (max_in_degree) *max_in_degree = net->max_in_degree;
  if (max_delay) *max_delay = net->max_delay;
  return NC_OK;
}
NC_ABI_CATCH

extern "C" nc_status nc_net_export(const nc_net* net, float* xyz, int64_t* in_ptr, uint32_t* in_src,
                                   float* w, uint8_t* delay_steps, uint8_t* inhibitory) try {
  if (!net) return fail(NC_ERR_ARG, "null network");
  const uint64_t S = net->n_syn();
  if (xyz) memcpy(xyz, net->xyz.data(), 12ull * net->n);
  if (in_ptr) memcpy(in_ptr, net->in_ptr.data(), 8ull * (net->n + 1));
  if (in_src) memcpy(in_src, net->in_src.data(), 4 * S);
  if (w) memcpy(w, net->w.data(), 4 * S);
  if (delay_steps) memcpy(delay_steps, net->delay.data(), S);
  if (inhibitory) memcpy(inhibitory, net->inh.data(), S);
  return NC_OK;
}
NC_ABI_CATCH

extern "C" void nc_net_free(nc_net* net) { delete net; }

extern "C" nc_status nc_net_select_inputs(const nc_net* net, const nc_drive_spec* drive,
                                          uint32_t* ids, uint32_t* groups) try {
  if (!net || !drive) return fail(NC_ERR_ARG, "null argument");
  nc::InputSel sel;
  nc_status st = nc::select_inputs(*net, *drive, &sel);
  if (st != NC_OK) return st;
  if (ids) memcpy(ids, sel.ids.data(), 4 * sel.ids.size());
  if (groups) memcpy(groups, sel.group.data(), 4 * sel.group.size());
  return NC_OK;
}
NC_ABI_CATCH

extern "C" const char* nc_last_error(void) { return nc::g_error.c_str(); }
