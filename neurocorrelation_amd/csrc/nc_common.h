// Shared host/device definitions of the NeuroCorrelation MI355X engine.
//
// Model constants are the reference defaults (src/NeuCor.cpp:18-29 for the
// network-level STDP parameters, :376-385 for the neuron, :469-485 for the
// synapse).  Time is an integer step count s; simulated time is s*dt with
// dt = 2^-m ms, so every time difference the reference takes in float is an
// exact multiple of dt and every transcendental it evaluates becomes a table
// indexed by a step difference (tables are built on the host from its libm).
#pragma once
#include <math.h>
#include <stdint.h>

#if defined(__HIP__)
#include <hip/hip_runtime.h>
#define NC_HD __host__ __device__ __forceinline__
#else
#define NC_HD static inline
#endif

// ---- reference constants -------------------------------------------------
#define NC_BASELEVEL   (-70.0f)   // Neuron::baselevel       src/NeuCor.cpp:376
#define NC_THRESHOLD   (-55.0f)   // Neuron::threshold       src/NeuCor.cpp:377
#define NC_RECHARGE    (0.5f)     // Neuron::recharge        src/NeuCor.cpp:378
#define NC_AP_CUTOFF   (2.0f)     // Neuron::AP_cutoff       src/NeuCor.cpp:385
#define NC_AP_H        (100.0f)   // AP shape                src/NeuCor.cpp:384
#define NC_AP_DEPOLW   (0.3f)
#define NC_AP_POLW     (0.6f)
#define NC_AP_DELTAPOL (1.16f)
#define NC_AP_DEPOLFAC (0.2f)
#define NC_AP_DSTART   (1.0f)
#define NC_SYN_DEPOL_MUL (52.0)   // Synapse::fire           src/NeuCor.cpp:730
#define NC_AP_SPEED    (2.0f)     // Synapse::AP_speed       src/NeuCor.cpp:485
#define NC_CHARGE_K    (0.9943)   // charge_insynapses       src/NeuCor.cpp:695
#define NC_CHARGE_R    (0.3702)

// "never" for lastFire (NaN in the reference) and lastSpikeArrival (-inf):
// with s < 2^29 every difference s - NC_NEVER exceeds all table lengths and
// all refractory / AP windows, which is exactly how NaN and -inf behave there.
#define NC_NEVER (-(1 << 30))
#define NC_MAX_STEPS (1 << 29)

// ---- per-slot record (16 B, one per incoming synapse) --------------------
// A synapse carries at most one spike (src/NeuCor.cpp:728): its record holds
// the depolarisation and arrival step latched when the spike was accepted.
typedef struct __attribute__((aligned(16))) nc_slot {
  float   w;        // Synapse::weight
  float   depol;    // Synapse::AP_depolFac latched at fire time
  int32_t last_arr; // Synapse::lastSpikeArrival as a step (NC_NEVER = -inf)
  int32_t arr_inh;  // 2 * arrival step of the latched spike (NC_NEVER: none)
                    // + 1 if inhibitory (sign class of the initial weight)
} nc_slot;

// ---- Philox4x32-10 (Salmon, Moraes, Dror, Shaw: "Parallel random numbers:
// as easy as 1, 2, 3", SC'11).  Counter-based, so every draw is a pure
// function of (key, counter) and is identical on host and device, in any
// launch order or network packing.
typedef struct { uint32_t v[4]; } nc_u32x4;

NC_HD void nc_mulhilo32(uint32_t a, uint32_t b, uint32_t* hi, uint32_t* lo) {
  uint64_t p = (uint64_t)a * (uint64_t)b;
  *hi = (uint32_t)(p >> 32);
  *lo = (uint32_t)p;
}

NC_HD nc_u32x4 nc_philox4x32_10(nc_u32x4 ctr, uint32_t k0, uint32_t k1) {
  for (int r = 0; r < 10; ++r) {
    uint32_t hi0, lo0, hi1, lo1;
    nc_mulhilo32(0xD2511F53u, ctr.v[0], &hi0, &lo0);
    nc_mulhilo32(0xCD9E8D57u, ctr.v[2], &hi1, &lo1);
    nc_u32x4 o;
    o.v[0] = hi1 ^ ctr.v[1] ^ k0;
    o.v[1] = lo1;
    o.v[2] = hi0 ^ ctr.v[3] ^ k1;
    o.v[3] = lo0;
    ctr = o;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  return ctr;
}

// Draw streams (third counter word).
#define NC_STREAM_POS    1u
#define NC_STREAM_WEIGHT 2u
#define NC_STREAM_INSEL  3u
#define NC_STREAM_DRIVE  4u
#define NC_STREAM_BG     5u

NC_HD nc_u32x4 nc_draw(uint64_t seed, uint32_t a, uint32_t b, uint32_t stream) {
  nc_u32x4 c;
  c.v[0] = a; c.v[1] = b; c.v[2] = stream; c.v[3] = 0u;
  return nc_philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
}

// Uniform in [0,1) with 24 random bits: exact in float.
NC_HD float nc_u01(uint32_t u) { return (float)(u >> 8) * (1.0f / 16777216.0f); }

// ---- the workload wiring's cell grid (host and device builders alike) -----
// 10 bits of x spread to every third bit (Morton order of 3-D cells).
NC_HD uint32_t nc_part1by2(uint32_t x) {
  x &= 0x3ff;
  x = (x | (x << 16)) & 0x030000ff;
  x = (x | (x << 8)) & 0x0300f00f;
  x = (x | (x << 4)) & 0x030c30c3;
  x = (x | (x << 2)) & 0x09249249;
  return x;
}

// Cells per axis: the largest count whose cell side is at least the radius
// (so the 27 cells around a neuron's cell hold every neighbour), at most 1024.
NC_HD int nc_cells_per_axis(float side, float radius) {
  int nc = (int)floor((double)side / (double)radius);
  if (nc < 1) nc = 1;
  while (nc > 1 && (double)side / nc < (double)radius) --nc;
  if (nc > 1024) nc = 1024;
  return nc;
}

NC_HD int nc_cell_of(float x, int ncell, float side) {
  int c = (int)floor((double)x * (double)ncell / (double)side);
  if (c < 0) c = 0;
  if (c > ncell - 1) c = ncell - 1;
  return c;
}
