// The workload wiring built on the GPU: the network nc_net_generate builds on
// the host, bit for bit, from the same counter-based draws.  It replaces the
// reference's all-pairs scan (Neuron::makeConnections, src/NeuCor.cpp:418-444)
// for the distance rule of the benchmark workloads:
//   1. positions from Philox draws (one per draw index), their cells, and
//      the ids: draw order or the Morton order of the cells (radix sort);
//   2. a cell grid (radix sort of (cell, id));
//   3. per target, the sources within the radius among the 27 cells around
//      its own: counted, prefix-summed into the incoming rows, written;
//   4. each row sorted by source id (radix sort of (target, source)), the
//      order of Neuron::inSynapses (src/NeuCor.h:212);
//   5. per synapse the delay from the distance and the weight and sign class
//      from a draw keyed by (source, target).
// Every float operation is written out with round-to-nearest intrinsics in
// the host builder's order, so no contraction or reassociation can differ.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <new>
#include <vector>

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include "nc_common.h"
#include "nc_internal.h"

using nc::fail;

namespace {

#define NW_HIP(x)                                                          \
  do {                                                                     \
    hipError_t e_ = (x);                                                   \
    if (e_ != hipSuccess) return fail(NC_ERR_HIP, hipGetErrorString(e_)); \
  } while (0)

// Device buffers released together, allocated and freed in the order of the
// build's own stream (hipMallocAsync / hipFreeAsync), so that a build on one
// host thread does not wait for, or hold up, work on other streams.
struct Bufs {
  hipStream_t st;
  std::vector<void*> p;
  ~Bufs() {
    for (void* x : p) (void)hipFreeAsync(x, st);
  }
  template <class T>
  hipError_t alloc(T** out, size_t n) {
    void* q = nullptr;
    hipError_t e = hipMallocAsync(&q, (n ? n : 1) * sizeof(T), st);
    if (e == hipSuccess) p.push_back(q);
    *out = (T*)q;
    return e;
  }
};

dim3 grid_for(uint64_t n) { return dim3((uint32_t)std::min<uint64_t>((n + 255) / 256, 65535)); }

// 1. positions in draw order, their cells, and the id sort keys
__global__ __launch_bounds__(256) void k_wire_pos(uint64_t seed, uint32_t n, float side, int ncell,
                                                  uint32_t spatial, float* px, float* py, float* pz,
                                                  uint32_t* cell, uint64_t* key) {
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const nc_u32x4 r = nc_draw(seed, i, 0u, NC_STREAM_POS);
    const float x = __fmul_rn(nc_u01(r.v[0]), side);
    const float y = __fmul_rn(nc_u01(r.v[1]), side);
    const float z = __fmul_rn(nc_u01(r.v[2]), side);
    px[i] = x, py[i] = y, pz[i] = z;
    const uint32_t cx = (uint32_t)nc_cell_of(x, ncell, side), cy = (uint32_t)nc_cell_of(y, ncell, side),
                   cz = (uint32_t)nc_cell_of(z, ncell, side);
    cell[i] = (cz * (uint32_t)ncell + cy) * (uint32_t)ncell + cx;
    const uint64_t m = nc_part1by2(cx) | (nc_part1by2(cy) << 1) | (nc_part1by2(cz) << 2);
    key[i] = spatial ? ((m << 32) | i) : (uint64_t)i;
  }
}

// Positions and cells by id; (cell, id) keys for the grid.
__global__ __launch_bounds__(256) void k_wire_place(uint32_t n, const uint64_t* order, const float* px,
                                                    const float* py, const float* pz, const uint32_t* cell_d,
                                                    float* X, float* Y, float* Z, float* xyz, uint32_t* cell,
                                                    uint64_t* ckey) {
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t id = blockIdx.x * blockDim.x + threadIdx.x; id < n; id += stride) {
    const uint32_t d = (uint32_t)(order[id] & 0xffffffffu);
    X[id] = px[d], Y[id] = py[d], Z[id] = pz[d];
    xyz[3ull * id] = px[d], xyz[3ull * id + 1] = py[d], xyz[3ull * id + 2] = pz[d];
    cell[id] = cell_d[d];
    ckey[id] = ((uint64_t)cell_d[d] << 32) | id;
  }
}

// 2. the grid: members of each cell in ascending id, cells as ranges
__global__ __launch_bounds__(256) void k_wire_cells(uint32_t n, const uint64_t* ckey, uint32_t* cmem,
                                                    uint32_t* cstart, uint32_t* cend) {
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += stride) {
    const uint64_t v = ckey[k];
    const uint32_t c = (uint32_t)(v >> 32);
    cmem[k] = (uint32_t)(v & 0xffffffffu);
    if (k == 0 || (uint32_t)(ckey[k - 1] >> 32) != c) cstart[c] = k;
    if (k + 1 == n || (uint32_t)(ckey[k + 1] >> 32) != c) cend[c] = k + 1;
  }
}

// IEEE square root (round to nearest even) of a float: the approximation
// is moved to the neighbour on the other side of a rounding midpoint, the
// test done exactly in double (a midpoint between adjacent floats has 25
// significant bits, so its square is exact).
__device__ __forceinline__ float sqrt_rn(float x) {
  float r = sqrtf(x);
  if (!(x > 0.0f) || !(r < INFINITY)) return r;
  const double xd = (double)x;
  const float up = nextafterf(r, INFINITY), dn = nextafterf(r, 0.0f);
  const double mu = 0.5 * ((double)r + (double)up), md = 0.5 * ((double)r + (double)dn);
  if (xd >= mu * mu) r = up;
  else if (xd < md * md) r = dn;
  return r;
}

// The distance test of the host builder, operation by operation.
__device__ __forceinline__ float wire_dist(const float* X, const float* Y, const float* Z, uint32_t src,
                                           uint32_t tg) {
  const float ddx = __fsub_rn(X[src], X[tg]), ddy = __fsub_rn(Y[src], Y[tg]), ddz = __fsub_rn(Z[src], Z[tg]);
  float d2 = __fadd_rn(__fmul_rn(ddx, ddx), __fmul_rn(ddy, ddy));
  d2 = __fadd_rn(d2, __fmul_rn(ddz, ddz));
  return sqrt_rn(d2);
}

// 3. per target: count (fill == false) or write (fill == true) the sources
// within the radius, in cell-visiting order
template <bool FILL>
__global__ __launch_bounds__(256) void k_wire_scan(uint32_t n, int ncell, float radius, const float* X,
                                                   const float* Y, const float* Z, const uint32_t* cell,
                                                   const uint32_t* cmem, const uint32_t* cstart,
                                                   const uint32_t* cend, uint64_t* deg, const uint64_t* in_ptr,
                                                   uint64_t* ekey, float* edist) {
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t tg = blockIdx.x * blockDim.x + threadIdx.x; tg < n; tg += stride) {
    const uint32_t c0 = cell[tg];
    const int x0 = (int)(c0 % (uint32_t)ncell), y0 = (int)((c0 / (uint32_t)ncell) % (uint32_t)ncell),
              z0 = (int)(c0 / ((uint32_t)ncell * (uint32_t)ncell));
    uint64_t cnt = 0, at = FILL ? in_ptr[tg] : 0;
    for (int dz = -1; dz <= 1; ++dz) {
      const int zz = z0 + dz;
      if (zz < 0 || zz >= ncell) continue;
      for (int dy = -1; dy <= 1; ++dy) {
        const int yy = y0 + dy;
        if (yy < 0 || yy >= ncell) continue;
        for (int dx = -1; dx <= 1; ++dx) {
          const int xx = x0 + dx;
          if (xx < 0 || xx >= ncell) continue;
          const uint32_t c = ((uint32_t)zz * ncell + (uint32_t)yy) * ncell + (uint32_t)xx;
          const uint32_t b = cstart[c], e = cend[c];
          for (uint32_t q = b; q < e; ++q) {
            const uint32_t src = cmem[q];
            if (src == tg) continue;
            const float dist = wire_dist(X, Y, Z, src, tg);
            if (!(dist < radius)) continue;
            if (FILL) {
              ekey[at] = ((uint64_t)tg << 32) | src;
              edist[at] = dist;
              ++at;
            } else {
              ++cnt;
            }
          }
        }
      }
    }
    if (!FILL) deg[tg] = cnt;
  }
}

// 5. per synapse, in row order: source, delay, weight, sign class
__global__ __launch_bounds__(256) void k_wire_attr(uint64_t S, const uint64_t* ekey, const float* edist,
                                                   uint64_t seed, float radius, float fdmax, uint32_t dmin,
                                                   uint32_t dmax, float w_min, float wspan, float inhib,
                                                   uint32_t* in_src, float* w, uint8_t* delay, uint8_t* inh) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; k < S; k += stride) {
    const uint32_t tg = (uint32_t)(ekey[k] >> 32), src = (uint32_t)(ekey[k] & 0xffffffffu);
    const float q = edist[k] / radius;  // IEEE division (HIP's default for float)
    const float dl = ceilf(__fmul_rn(q, fdmax));
    const uint32_t dd = dl < (float)dmin ? dmin : (dl > fdmax ? dmax : (uint32_t)dl);
    const nc_u32x4 r = nc_draw(seed, src, tg, NC_STREAM_WEIGHT);
    float wv = __fadd_rn(w_min, __fmul_rn(nc_u01(r.v[0]), wspan));
    const uint8_t ih = nc_u01(r.v[1]) < inhib ? 1 : 0;
    if (ih) wv = -wv;
    in_src[k] = src;
    w[k] = wv;
    delay[k] = (uint8_t)dd;
    inh[k] = ih;
  }
}

template <class F>
nc_status with_temp(Bufs& B, F&& f) {
  size_t bytes = 0;
  NW_HIP(f(nullptr, bytes));
  void* tmp = nullptr;
  NW_HIP(B.alloc((unsigned char**)&tmp, bytes));
  NW_HIP(f(tmp, bytes));
  return NC_OK;
}

uint32_t bits_for(uint64_t v) {
  uint32_t b = 1;
  while (b < 64 && (v >> b)) ++b;
  return b;
}

}  // namespace

extern "C" nc_status nc_net_generate_device(const nc_wiring_spec* spec, int32_t device, nc_net** out) {
  if (!spec || !out) return fail(NC_ERR_ARG, "null argument");
  *out = nullptr;
  const nc_wiring_spec s = *spec;
  {
    nc_status st = nc::check_wiring_spec(s);
    if (st != NC_OK) return st;
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(NC_ERR_NO_DEVICE, "no HIP device");
  if (device < 0 || device >= ndev) return fail(NC_ERR_RANGE, "device index out of range");
  NW_HIP(hipSetDevice(device));
  const uint32_t n = s.n_neurons;
  const int ncell = nc_cells_per_axis(s.side, s.radius);
  const uint64_t ncells = (uint64_t)ncell * ncell * ncell;
  hipStream_t st = nullptr;
  NW_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
  struct StreamGuard {
    hipStream_t s;
    ~StreamGuard() {
      (void)hipStreamSynchronize(s);  // the buffers' frees, queued on it
      (void)hipStreamDestroy(s);
    }
  } sg{st};
  Bufs B{st, {}};
  float *px, *py, *pz, *X, *Y, *Z, *xyz;
  uint32_t *cell_d, *cell, *cmem, *cstart, *cend;
  uint64_t *key, *key2, *ckey, *ckey2, *deg, *in_ptr;
  NW_HIP(B.alloc(&px, n));
  NW_HIP(B.alloc(&py, n));
  NW_HIP(B.alloc(&pz, n));
  NW_HIP(B.alloc(&X, n));
  NW_HIP(B.alloc(&Y, n));
  NW_HIP(B.alloc(&Z, n));
  NW_HIP(B.alloc(&xyz, 3ull * n));
  NW_HIP(B.alloc(&cell_d, n));
  NW_HIP(B.alloc(&cell, n));
  NW_HIP(B.alloc(&cmem, n));
  NW_HIP(B.alloc(&cstart, ncells));
  NW_HIP(B.alloc(&cend, ncells));
  NW_HIP(B.alloc(&key, n));
  NW_HIP(B.alloc(&key2, n));
  NW_HIP(B.alloc(&ckey, n));
  NW_HIP(B.alloc(&ckey2, n));
  NW_HIP(B.alloc(&deg, n));
  NW_HIP(B.alloc(&in_ptr, n + 1ull));
  const dim3 gn = grid_for(n);
  hipLaunchKernelGGL(k_wire_pos, gn, dim3(256), 0, st, s.seed, n, s.side, ncell, s.spatial_order, px, py, pz,
                     cell_d, key);
  NW_HIP(hipGetLastError());
  const uint64_t* order = key;
  if (s.spatial_order) {
    // Morton codes of 10-bit cells take 30 bits above the 32-bit index
    nc_status r = with_temp(B, [&](void* t, size_t& b) {
      return rocprim::radix_sort_keys(t, b, key, key2, (size_t)n, 0u, 62u, st);
    });
    if (r != NC_OK) return r;
    order = key2;
  }
  hipLaunchKernelGGL(k_wire_place, gn, dim3(256), 0, st, n, order, px, py, pz, cell_d, X, Y, Z, xyz, cell, ckey);
  NW_HIP(hipGetLastError());
  {
    const uint32_t hi = 32u + bits_for(ncells - 1);
    nc_status r = with_temp(B, [&](void* t, size_t& b) {
      return rocprim::radix_sort_keys(t, b, ckey, ckey2, (size_t)n, 0u, hi, st);
    });
    if (r != NC_OK) return r;
  }
  NW_HIP(hipMemsetAsync(cstart, 0, ncells * 4, st));
  NW_HIP(hipMemsetAsync(cend, 0, ncells * 4, st));
  hipLaunchKernelGGL(k_wire_cells, gn, dim3(256), 0, st, n, ckey2, cmem, cstart, cend);
  NW_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_wire_scan<false>, gn, dim3(256), 0, st, n, ncell, s.radius, X, Y, Z, cell, cmem, cstart,
                     cend, deg, (const uint64_t*)nullptr, (uint64_t*)nullptr, (float*)nullptr);
  NW_HIP(hipGetLastError());
  NW_HIP(hipMemsetAsync(in_ptr, 0, 8, st));
  {
    nc_status r = with_temp(B, [&](void* t, size_t& b) {
      return rocprim::inclusive_scan(t, b, deg, in_ptr + 1, (size_t)n, rocprim::plus<uint64_t>(), st);
    });
    if (r != NC_OK) return r;
  }
  uint64_t S = 0;
  NW_HIP(hipMemcpyAsync(&S, in_ptr + n, 8, hipMemcpyDeviceToHost, st));
  NW_HIP(hipStreamSynchronize(st));
  if (S >= (1ull << 32) - (1ull << 24)) return fail(NC_ERR_ARG, "too many synapses");
  uint64_t *ekey, *ekey2;
  float *edist, *edist2, *w;
  uint32_t* in_src;
  uint8_t *delay, *inh;
  NW_HIP(B.alloc(&ekey, S));
  NW_HIP(B.alloc(&ekey2, S));
  NW_HIP(B.alloc(&edist, S));
  NW_HIP(B.alloc(&edist2, S));
  hipLaunchKernelGGL(k_wire_scan<true>, gn, dim3(256), 0, st, n, ncell, s.radius, X, Y, Z, cell, cmem, cstart,
                     cend, deg, (const uint64_t*)in_ptr, ekey, edist);
  NW_HIP(hipGetLastError());
  if (S) {
    const uint32_t hi = 32u + bits_for(n > 1 ? n - 1 : 1);
    nc_status r = with_temp(B, [&](void* t, size_t& b) {
      return rocprim::radix_sort_pairs(t, b, ekey, ekey2, edist, edist2, (size_t)S, 0u, hi, st);
    });
    if (r != NC_OK) return r;
  }
  NW_HIP(B.alloc(&in_src, S));
  NW_HIP(B.alloc(&w, S));
  NW_HIP(B.alloc(&delay, S));
  NW_HIP(B.alloc(&inh, S));
  if (S) {
    hipLaunchKernelGGL(k_wire_attr, grid_for(S), dim3(256), 0, st, S, ekey2, edist2, s.seed, s.radius,
                       (float)s.delay_max, s.delay_min, s.delay_max, s.w_min, s.w_max - s.w_min, s.inhib_frac,
                       in_src, w, delay, inh);
    NW_HIP(hipGetLastError());
  }
  nc_net* net = new (std::nothrow) nc_net();
  if (!net) return fail(NC_ERR_OOM, "host allocation failed");
  try {
    net->n = n;
    net->xyz.resize(3ull * n);
    net->in_ptr.resize(n + 1ull);
    net->in_src.resize(S);
    net->w.resize(S);
    net->delay.resize(S);
    net->inh.resize(S);
  } catch (const std::bad_alloc&) {
    delete net;
    return fail(NC_ERR_OOM, "host allocation failed");
  }
  hipError_t e = hipMemcpyAsync(net->xyz.data(), xyz, 12ull * n, hipMemcpyDeviceToHost, st);
  if (e == hipSuccess)
    e = hipMemcpyAsync(net->in_ptr.data(), in_ptr, 8ull * (n + 1ull), hipMemcpyDeviceToHost, st);
  if (e == hipSuccess && S) e = hipMemcpyAsync(net->in_src.data(), in_src, 4ull * S, hipMemcpyDeviceToHost, st);
  if (e == hipSuccess && S) e = hipMemcpyAsync(net->w.data(), w, 4ull * S, hipMemcpyDeviceToHost, st);
  if (e == hipSuccess && S) e = hipMemcpyAsync(net->delay.data(), delay, S, hipMemcpyDeviceToHost, st);
  if (e == hipSuccess && S) e = hipMemcpyAsync(net->inh.data(), inh, S, hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e != hipSuccess) {
    delete net;
    return fail(NC_ERR_HIP, hipGetErrorString(e));
  }
  nc::finish_net_stats(net);
  *out = net;
  return NC_OK;
}
