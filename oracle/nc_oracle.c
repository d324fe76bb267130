/*
 * nc_oracle.c -- CPU oracle for the NeuroCorrelation simulation step.
 *
 * TEST INFRASTRUCTURE ONLY.  This file restates the reference's algorithm
 * (src/NeuCor.cpp of Axelwickm/NeuroCorrelation) in plain C so that the
 * MI355X engine can be checked against it.  Only the tests, the smoke check
 * and the CPU-baseline leg of bench.py may load it; the product never does.
 *
 * What it restates, and how it is pinned (see the tests named in brackets):
 *   - the step: Neuron::run / fire / charge_* / AP, Synapse::run / fire /
 *     synapticPlasticity, evaluated in float simulated time with the same
 *     float/double expression shapes and direct glibc libm calls
 *     [pinned bit-exactly against the compiled reference on tie-free
 *      micro-networks, against closed forms, and against the essay's
 *      FEW_NEURONS outcome];
 *   - the clocked order of one step (all neurons run at t, then fired
 *     neurons fan out), which the reference's event queue reduces to when
 *     runAll = true, runSpeed = dt and every event lies on the step grid;
 *   - Philox4x32-10 [Random123 known-answer vectors];
 *   - the workload wiring rule, input selection and drive, which are not in
 *     the reference but are specified for the workloads [brute force over all
 *     pairs; closed-form degree and delay properties].
 * Compile with -ffp-contract=off (the reference's x86-64 build has no FMA).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define OR_EXPORT __attribute__((visibility("default")))

/* ---------------------------------------------------------------------------
 * Philox4x32-10 (Salmon et al., SC'11, "Parallel random numbers: as easy as
 * 1, 2, 3"): 10 rounds of two 32x32->64 multiplies with the Weyl key bump.
 * ------------------------------------------------------------------------ */
OR_EXPORT void or_philox4x32_10(const uint32_t in_ctr[4], const uint32_t in_key[2], uint32_t out[4]) {
  uint32_t c0 = in_ctr[0], c1 = in_ctr[1], c2 = in_ctr[2], c3 = in_ctr[3];
  uint32_t k0 = in_key[0], k1 = in_key[1];
  for (int r = 0; r < 10; ++r) {
    uint64_t p0 = (uint64_t)0xD2511F53u * c0;
    uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
    uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
    uint32_t n1 = (uint32_t)p1;
    uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
    uint32_t n3 = (uint32_t)p0;
    c0 = n0; c1 = n1; c2 = n2; c3 = n3;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

static void draw(uint64_t seed, uint32_t a, uint32_t b, uint32_t stream, uint32_t out[4]) {
  uint32_t ctr[4] = {a, b, stream, 0u};
  uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  or_philox4x32_10(ctr, key, out);
}
static float u01(uint32_t u) { return (float)(u >> 8) * (1.0f / 16777216.0f); }

enum { ST_POS = 1, ST_WEIGHT = 2, ST_INSEL = 3, ST_DRIVE = 4, ST_BG = 5 };

/* ---------------------------------------------------------------------------
 * Workload wiring rule (cube, radius, delays, weights), brute force.
 * ------------------------------------------------------------------------ */
typedef struct or_wiring {
  uint32_t n_neurons;
  uint32_t spatial_order;
  uint64_t seed;
  float side, radius;
  uint32_t delay_min, delay_max;
  float w_min, w_max, inhib_frac;
} or_wiring;

static int or_cells(float side, float radius) {
  int nc = (int)floor((double)side / (double)radius);
  if (nc < 1) nc = 1;
  while (nc > 1 && (double)side / nc < (double)radius) --nc;
  if (nc > 1024) nc = 1024;
  return nc;
}
static uint32_t or_cell(float x, int nc, float side) {
  int c = (int)floor((double)x * (double)nc / (double)side);
  if (c < 0) c = 0;
  if (c > nc - 1) c = nc - 1;
  return (uint32_t)c;
}
static uint64_t spread3(uint32_t v) { /* bit b of v -> bit 3b */
  uint64_t r = 0;
  for (int b = 0; b < 10; ++b) r |= (uint64_t)((v >> b) & 1u) << (3 * b);
  return r;
}
static int cmp_u64(const void* a, const void* b) {
  uint64_t x = *(const uint64_t*)a, y = *(const uint64_t*)b;
  return x < y ? -1 : (x > y);
}

/* Positions by neuron id (xyz, 3 per neuron). */
OR_EXPORT int or_positions(const or_wiring* s, float* xyz) {
  uint32_t n = s->n_neurons;
  float* p = (float*)malloc(12ull * n);
  uint64_t* key = (uint64_t*)malloc(8ull * n);
  if (!p || !key) { free(p); free(key); return -1; }
  int nc = or_cells(s->side, s->radius);
  for (uint32_t i = 0; i < n; ++i) {
    uint32_t r[4];
    draw(s->seed, i, 0u, ST_POS, r);
    p[3 * i] = u01(r[0]) * s->side;
    p[3 * i + 1] = u01(r[1]) * s->side;
    p[3 * i + 2] = u01(r[2]) * s->side;
    uint64_t m = 0;
    if (s->spatial_order) {
      m = spread3(or_cell(p[3 * i], nc, s->side)) | (spread3(or_cell(p[3 * i + 1], nc, s->side)) << 1) |
          (spread3(or_cell(p[3 * i + 2], nc, s->side)) << 2);
    }
    key[i] = (m << 32) | i;
  }
  qsort(key, n, 8, cmp_u64);
  for (uint32_t id = 0; id < n; ++id) {
    uint32_t d = (uint32_t)(key[id] & 0xffffffffu);
    memcpy(xyz + 3ull * id, p + 3ull * d, 12);
  }
  free(p);
  free(key);
  return 0;
}

/* In-rows of the given targets by brute force over every source.  Returns
 * the number of synapses written (or needed, when the arrays are NULL). */
OR_EXPORT int64_t or_rows(const or_wiring* s, const float* xyz, uint32_t n_rows, const uint32_t* rows,
                          int64_t* row_ptr, uint32_t* src, float* w, uint8_t* delay, uint8_t* inh) {
  const uint32_t n = s->n_neurons;
  const float span = s->w_max - s->w_min, fdmax = (float)s->delay_max;
  int64_t k = 0;
  if (row_ptr) row_ptr[0] = 0;
  for (uint32_t q = 0; q < n_rows; ++q) {
    uint32_t tg = rows[q];
    for (uint32_t sr = 0; sr < n; ++sr) {
      if (sr == tg) continue;
      float dx = xyz[3 * sr] - xyz[3 * tg], dy = xyz[3 * sr + 1] - xyz[3 * tg + 1],
            dz = xyz[3 * sr + 2] - xyz[3 * tg + 2];
      float d2 = dx * dx + dy * dy;
      d2 = d2 + dz * dz;
      float d = sqrtf(d2);
      if (!(d < s->radius)) continue;
      if (src) {
        float dl = ceilf((d / s->radius) * fdmax);
        uint32_t dd = dl < (float)s->delay_min ? s->delay_min : (dl > fdmax ? s->delay_max : (uint32_t)dl);
        uint32_t r[4];
        draw(s->seed, sr, tg, ST_WEIGHT, r);
        float ww = s->w_min + u01(r[0]) * span;
        uint8_t ih = u01(r[1]) < s->inhib_frac;
        src[k] = sr;
        w[k] = ih ? -ww : ww;
        delay[k] = (uint8_t)dd;
        inh[k] = ih;
      }
      ++k;
    }
    if (row_ptr) row_ptr[q + 1] = k;
  }
  return k;
}

/* Input selection: the n_inputs ids with the smallest (hash, id). */
OR_EXPORT int or_select_inputs(uint32_t n, uint64_t seed, uint32_t n_inputs, uint32_t n_correlated,
                               uint32_t n_groups, uint32_t* ids, uint32_t* groups) {
  uint64_t* key = (uint64_t*)malloc(8ull * n);
  if (!key) return -1;
  for (uint32_t i = 0; i < n; ++i) {
    uint32_t r[4];
    draw(seed, i, 0u, ST_INSEL, r);
    key[i] = ((uint64_t)r[0] << 32) | i;
  }
  qsort(key, n, 8, cmp_u64);
  uint32_t ng = n_correlated > 0 ? n_groups : 0;
  for (uint32_t r = 0; r < n_inputs; ++r) {
    ids[r] = (uint32_t)(key[r] & 0xffffffffu);
    groups[r] = r < n_correlated ? r % ng : ng + (r - n_correlated);
  }
  free(key);
  return 0;
}

/* ---------------------------------------------------------------------------
 * The simulation (src/NeuCor.cpp:583-764 in the clocked order).
 * ------------------------------------------------------------------------ */
typedef struct or_sim {
  uint32_t n;
  uint64_t S;
  int64_t* in_ptr;
  uint32_t* in_src;
  uint8_t* inh;
  float* length2;   /* fl(length * AP_speed) = delay in ms */
  float* w;         /* Synapse::weight */
  float* fire_t;    /* Synapse::AP_fireTime (0 = no spike in flight) */
  float* depol;     /* Synapse::AP_depolFac */
  float* last_arr;  /* Synapse::lastSpikeArrival */
  int32_t* released;/* step at which the latch last cleared (tie diagnostics) */
  int32_t* wchanged;/* step at which post-side STDP last changed w (tie diagnostics) */
  uint32_t* tgt;    /* target of each synapse */
  int64_t* out_ptr; /* per source: in-CSR indices of its out-synapses */
  uint64_t* out_syn;
  float* V;         /* potAct[2i] */
  float* last_fire; /* Neuron::lastFire */
  float* last_ran;  /* simulator::lastRan */
  uint32_t* firings;
  uint8_t* fired;   /* this step */
  uint8_t* f_in;
  uint8_t* f_bg;
  float t;          /* NeuCor::currentTime */
  float dt;
  float lr, pre_f, post_f, pre_decay, post_decay;
  int32_t step;
  /* drive */
  uint32_t n_inputs;
  uint32_t* in_ids;
  uint32_t* in_groups;
  uint32_t n_corr_groups, period;
  float p0, amp_corr, amp_ind;
  uint64_t drive_seed;
  uint32_t bg_thr;
  uint64_t bg_seed;
  /* host events: (absolute step, id, kind) */
  uint64_t n_ev, cap_ev;
  uint32_t* ev;
  /* diagnostics */
  uint64_t ties_latch, ties_weight, arrivals, fires, accepted;
  int64_t first_tie; /* first step with a queue-order dependent outcome (-1: none) */
} or_sim;

OR_EXPORT void or_sim_free(or_sim* o);

OR_EXPORT or_sim* or_sim_create(uint32_t n, uint64_t S, const int64_t* in_ptr, const uint32_t* in_src,
                                const float* w, const uint8_t* delay, const uint8_t* inh, float dt, float lr,
                                float pre_f, float post_f, float pre_decay, float post_decay) {
  or_sim* o = (or_sim*)calloc(1, sizeof(or_sim));
  if (!o) return NULL;
  o->n = n;
  o->S = S;
  o->in_ptr = (int64_t*)malloc(8ull * (n + 1));
  o->in_src = (uint32_t*)malloc(4ull * S + 4);
  o->inh = (uint8_t*)malloc(S + 1);
  o->length2 = (float*)malloc(4ull * S + 4);
  o->w = (float*)malloc(4ull * S + 4);
  o->fire_t = (float*)calloc(S + 1, 4);
  o->depol = (float*)calloc(S + 1, 4);
  o->last_arr = (float*)malloc(4ull * S + 4);
  o->released = (int32_t*)malloc(4ull * S + 4);
  o->wchanged = (int32_t*)malloc(4ull * S + 4);
  o->tgt = (uint32_t*)malloc(4ull * S + 4);
  o->out_ptr = (int64_t*)calloc(n + 1, 8);
  o->out_syn = (uint64_t*)malloc(8ull * S + 8);
  o->V = (float*)malloc(4ull * n);
  o->last_fire = (float*)malloc(4ull * n);
  o->last_ran = (float*)calloc(n, 4);
  o->firings = (uint32_t*)calloc(n, 4);
  o->fired = (uint8_t*)calloc(n, 1);
  o->f_in = (uint8_t*)calloc(n, 1);
  o->f_bg = (uint8_t*)calloc(n, 1);
  if (!o->in_ptr || !o->in_src || !o->inh || !o->length2 || !o->w || !o->fire_t || !o->depol ||
      !o->last_arr || !o->released || !o->wchanged || !o->tgt || !o->out_ptr || !o->out_syn || !o->V || !o->last_fire ||
      !o->last_ran || !o->firings || !o->fired || !o->f_in || !o->f_bg) {
    or_sim_free(o);
    return NULL;
  }
  memcpy(o->in_ptr, in_ptr, 8ull * (n + 1));
  memcpy(o->in_src, in_src, 4ull * S);
  memcpy(o->w, w, 4ull * S);
  for (uint64_t k = 0; k < S; ++k) {
    o->inh[k] = inh ? inh[k] : (w[k] < 0.0f);
    /* length = delay*dt/2, so length*AP_speed = delay*dt exactly */
    float length = (float)delay[k] * dt / 2.0f;
    o->length2[k] = length * 2.0f;
    o->last_arr[k] = -INFINITY; /* Synapse::Synapse, src/NeuCor.cpp:469 */
    o->released[k] = -1;
    o->wchanged[k] = -1;
  }
  for (uint32_t tg = 0; tg < n; ++tg)
    for (int64_t k = in_ptr[tg]; k < in_ptr[tg + 1]; ++k) o->tgt[k] = tg;
  /* Out lists: targets visited in ascending id, so each source's list is in
   * ascending target order (Neuron::outSynapses order). */
  for (uint64_t k = 0; k < S; ++k) o->out_ptr[in_src[k] + 1]++;
  for (uint32_t i = 0; i < n; ++i) o->out_ptr[i + 1] += o->out_ptr[i];
  {
    int64_t* fill = (int64_t*)malloc(8ull * (n + 1));
    memcpy(fill, o->out_ptr, 8ull * (n + 1));
    for (uint32_t tg = 0; tg < n; ++tg)
      for (int64_t k = in_ptr[tg]; k < in_ptr[tg + 1]; ++k) o->out_syn[fill[in_src[k]]++] = (uint64_t)k;
    free(fill);
  }
  for (uint32_t i = 0; i < n; ++i) {
    o->V[i] = -70.0f;       /* setPotential(baselevel), src/NeuCor.cpp:394 */
    o->last_fire[i] = NAN;  /* :392 */
  }
  o->t = 0.0f;
  o->dt = dt;
  o->lr = lr;
  o->pre_f = pre_f;
  o->post_f = post_f;
  o->pre_decay = pre_decay;
  o->post_decay = post_decay;
  o->first_tie = -1;
  return o;
}

OR_EXPORT void or_sim_free(or_sim* o) {
  if (!o) return;
  void* p[] = {o->in_ptr, o->in_src, o->inh, o->length2, o->w, o->fire_t, o->depol, o->last_arr,
               o->released, o->wchanged, o->tgt, o->out_ptr, o->out_syn, o->V, o->last_fire, o->last_ran, o->firings,
               o->fired, o->f_in, o->f_bg, o->in_ids, o->in_groups, o->ev};
  for (unsigned i = 0; i < sizeof(p) / sizeof(p[0]); ++i) free(p[i]);
  free(o);
}

OR_EXPORT void or_sim_set_params(or_sim* o, float lr, float pre_f, float post_f) {
  o->lr = lr;
  o->pre_f = pre_f;
  o->post_f = post_f;
}

OR_EXPORT int or_sim_set_drive(or_sim* o, uint32_t n_inputs, const uint32_t* ids, const uint32_t* groups,
                               uint32_t n_corr_groups, uint32_t period, float p0, float amp_corr,
                               float amp_ind, float bg_prob, uint64_t seed) {
  free(o->in_ids);
  free(o->in_groups);
  o->in_ids = (uint32_t*)malloc(4ull * n_inputs + 4);
  o->in_groups = (uint32_t*)malloc(4ull * n_inputs + 4);
  if (!o->in_ids || !o->in_groups) return -1;
  memcpy(o->in_ids, ids, 4ull * n_inputs);
  memcpy(o->in_groups, groups, 4ull * n_inputs);
  o->n_inputs = n_inputs;
  o->n_corr_groups = n_corr_groups;
  o->period = period ? period : 1;
  o->p0 = p0;
  o->amp_corr = amp_corr;
  o->amp_ind = amp_ind;
  o->drive_seed = seed;
  o->bg_thr = 0;
  if (bg_prob > 0.0f) {
    double th = floor((double)bg_prob * 4294967296.0);
    o->bg_thr = th >= 4294967295.0 ? 0xFFFFFFFFu : (uint32_t)th;
    if (o->bg_thr == 0) o->bg_thr = 1;
  }
  o->bg_seed = seed ^ 0x9E3779B97F4A7C15ull;
  return 0;
}

/* Forced fires at absolute steps (kind 0 = input fire, 1 = background). */
OR_EXPORT int or_sim_push_forced(or_sim* o, uint32_t n, const uint32_t* step_off, const uint32_t* ids,
                                 const uint8_t* kind) {
  if (o->n_ev + n > o->cap_ev) {
    uint64_t cap = (o->n_ev + n) * 2 + 16;
    uint32_t* e = (uint32_t*)realloc(o->ev, 12ull * cap);
    if (!e) return -1;
    o->ev = e;
    o->cap_ev = cap;
  }
  for (uint32_t k = 0; k < n; ++k) {
    o->ev[3 * o->n_ev] = (uint32_t)o->step + step_off[k];
    o->ev[3 * o->n_ev + 1] = ids[k];
    o->ev[3 * o->n_ev + 2] = kind[k];
    o->n_ev++;
  }
  return 0;
}

static uint32_t drive_thr(const or_sim* o, uint32_t g, int32_t s) {
  double amp = g < o->n_corr_groups ? o->amp_corr : o->amp_ind;
  uint32_t k = (uint32_t)s % o->period;
  double p = (double)o->p0 * (1.0 + amp * sin(2.0 * M_PI * (double)k / (double)o->period));
  if (!(p > 0.0)) return 0u;
  if (p >= 1.0) return 0xFFFFFFFFu;
  return (uint32_t)floor(p * 4294967296.0);
}

/* Neuron::getTrace (src/NeuCor.cpp:671-675). */
static float neuron_trace(const or_sim* o, uint32_t i) {
  float tr = powf(o->post_decay, o->t - o->last_fire[i]);
  if (tr == tr) return tr;
  return 0;
}

/* Synapse::synapticPlasticity (src/NeuCor.cpp:740-764); k = in-CSR index,
 * i = its target.  The bookkeeping of :747-749 and the dead branch of :752
 * do not affect the weight. */
static void plasticity(or_sim* o, uint64_t k, uint32_t i) {
  float traceS = powf(o->pre_decay, o->t - o->last_arr[k]);
  float traceT = neuron_trace(o, i);
  if (traceT == 1) traceT = 0;
  if (traceS == 1) traceS = 0;
  float weightChange = o->pre_f * traceS - o->post_f * traceT;
  float weight = o->w[k];
  weight += weightChange * o->lr;
  if (!o->inh[k]) weight = fmax(fmin(weight, 1.0), 0.0);
  else weight = fmax(fmin(weight, 0.0), -1.0);
  o->w[k] = weight;
}

/* Neuron::AP (src/NeuCor.cpp:706-714). */
static void action_potential(or_sim* o, uint32_t i, float currentT) {
  const float lastFire = o->last_fire[i];
  if (lastFire != lastFire || 2.0f < (float)currentT - lastFire) return;
  const float AP_h = 100.0f, AP_depolW = 0.3f, AP_polW = 0.6f, AP_deltaPol = 1.16f, AP_depolFac = 0.2f,
              AP_deltaStart = 1.0f, baselevel = -70.0f, threshold = -55.0f;
  float currentAP = AP_h * (exp(-powf(((float)currentT - lastFire) - AP_deltaStart, 2.0) /
                                (2.0 * AP_depolW * AP_depolW)) -
                            exp(-powf(((float)currentT - lastFire) - AP_deltaStart - AP_deltaPol, 2.0) /
                                (2.0 * AP_polW * AP_polW)) *
                                AP_depolFac) +
                    baselevel + (threshold - baselevel) * fmax(1.0 + lastFire - currentT, 0.0);
  o->V[i] = currentAP;
}

/* One step: NeuCor::run() with runAll = true and runSpeed = dt.  Fired ids
 * are appended to fired_ids (ascending), with their step in fired_steps. */
static void step_once(or_sim* o, uint32_t* fired_ids, uint32_t* fired_steps, uint64_t cap, uint64_t* nf) {
  const uint32_t n = o->n;
  const float targetTime = o->t + o->dt; /* src/NeuCor.cpp:609 */
  o->t = targetTime;
  o->step++;
  const int32_t s = o->step;
  const float currentT = o->t;
  memset(o->f_in, 0, n);
  memset(o->f_bg, 0, n);
  memset(o->fired, 0, n);
  /* device-style drive: group spike trains and background */
  for (uint32_t r = 0; r < o->n_inputs; ++r) {
    uint32_t g = o->in_groups[r], u[4];
    draw(o->drive_seed, (uint32_t)s, g, ST_DRIVE, u);
    if (u[0] < drive_thr(o, g, s)) o->f_in[o->in_ids[r]] = 1;
  }
  if (o->bg_thr) {
    for (uint32_t i = 0; i < n; ++i) {
      uint32_t u[4];
      draw(o->bg_seed, (uint32_t)s, i, ST_BG, u);
      if (u[0] < o->bg_thr) o->f_bg[i] = 1;
    }
  }
  /* host events of this step */
  {
    uint64_t keep = 0;
    for (uint64_t e = 0; e < o->n_ev; ++e) {
      uint32_t es = o->ev[3 * e], id = o->ev[3 * e + 1], kind = o->ev[3 * e + 2];
      if ((int32_t)es == s) {
        if (kind) o->f_bg[id] = 1;
        else o->f_in[id] = 1;
      } else if ((int32_t)es > s) {
        memmove(o->ev + 3 * keep, o->ev + 3 * e, 12);
        ++keep;
      }
    }
    o->n_ev = keep;
  }
  /* Phase 1: Neuron::run (src/NeuCor.cpp:619-641) on every neuron. */
  for (uint32_t i = 0; i < n; ++i) {
    float deltaT = currentT - o->last_ran[i];
    o->last_ran[i] = currentT;
    if (deltaT == 0) continue;
    /* charge_insynapses (:688-700), inSynapses in ascending source order */
    float newPot = o->V[i];
    for (int64_t k = o->in_ptr[i]; k < o->in_ptr[i + 1]; ++k) {
      float timeOffset = currentT - o->fire_t[k];
      if (timeOffset <= 0.0 || o->fire_t[k] == 0) continue;
      newPot += deltaT * o->depol[k] * 0.9943 * exp(0.3702 * deltaT);
      if (2.0f < timeOffset) {
        o->fire_t[k] = 0;
        o->released[k] = s;
      }
    }
    o->V[i] = newPot;
    /* charge_passive (:677-680) */
    {
      const float baselevel = -70.0f, recharge = 0.5f;
      float np = ((float)o->V[i] - baselevel) * powf(recharge, deltaT) + baselevel;
      o->V[i] = np;
    }
    /* charge_thresholdCheck (:682-686); vesicles stay > 0 (:391, :702-704).
     * Background fires stand in for scheduledFireTime == currentT; input
     * fires call Neuron::fire directly (:326-331). */
    int do_fire = 0;
    {
      const float threshold = -55.0f;
      const float lastFire = o->last_fire[i];
      if ((threshold < o->V[i] || o->f_bg[i]) && (lastFire != lastFire || 2.0f < (float)currentT - lastFire))
        do_fire = 1;
      if (o->f_in[i]) do_fire = 1;
    }
    if (do_fire) {
      /* Neuron::fire (:643-656): out-synapses fire in phase 2 below */
      o->last_fire[i] = currentT;
      o->firings[i]++;
      o->fired[i] = 1;
      o->fires++;
      for (int64_t k = o->in_ptr[i]; k < o->in_ptr[i + 1]; ++k) {
        float before = o->w[k];
        plasticity(o, (uint64_t)k, i);
        if (o->w[k] != before) o->wchanged[k] = s;
      }
    }
    /* Synapse::run (:718-726) for spikes arriving now: after the target ran */
    for (int64_t k = o->in_ptr[i]; k < o->in_ptr[i + 1]; ++k) {
      if (o->fire_t[k] != currentT) continue;
      o->last_arr[k] = currentT;
      plasticity(o, (uint64_t)k, i);
      o->arrivals++;
    }
    action_potential(o, i, currentT);
  }
  /* Phase 2: Synapse::fire (:727-738) on every out-synapse of fired neurons. */
  for (uint32_t i = 0; i < n; ++i) {
    if (!o->fired[i]) continue;
    if (fired_ids && *nf < cap) {
      fired_ids[*nf] = i;
      if (fired_steps) fired_steps[*nf] = (uint32_t)s;
    }
    if (nf) (*nf)++;
    for (int64_t e = o->out_ptr[i]; e < o->out_ptr[i + 1]; ++e) {
      uint64_t k = o->out_syn[e];
      if (o->fire_t[k] != 0) continue;
      /* Orders the reference leaves to its event-queue tie-break: this
       * fan-out against the latch release of the same step, and against the
       * target's post-side STDP of the same step. */
      if (o->released[k] == s) o->ties_latch++;
      if (o->wchanged[k] == s) o->ties_weight++;
      if ((o->released[k] == s || o->wchanged[k] == s) && o->first_tie < 0) o->first_tie = s;
      float AP_depolFac = 0.2f;
      AP_depolFac *= 52.0;
      AP_depolFac *= o->w[k];
      o->depol[k] = AP_depolFac;
      float AP_fireTime = o->length2[k];
      AP_fireTime += currentT;
      o->fire_t[k] = AP_fireTime;
      o->accepted++;
    }
  }
}

OR_EXPORT void or_sim_step(or_sim* o, uint32_t n_steps, uint32_t* fired_ids, uint32_t* fired_steps,
                           uint64_t cap, uint64_t* n_fired) {
  uint64_t nf = 0;
  for (uint32_t k = 0; k < n_steps; ++k) step_once(o, fired_ids, fired_steps, cap, &nf);
  if (n_fired) *n_fired = nf;
}

OR_EXPORT void or_sim_read(const or_sim* o, float* V, float* last_fire, uint32_t* firings, float* w,
                           float* fire_t, float* last_arr, float* depol, float* time) {
  if (V) memcpy(V, o->V, 4ull * o->n);
  if (last_fire) memcpy(last_fire, o->last_fire, 4ull * o->n);
  if (firings) memcpy(firings, o->firings, 4ull * o->n);
  if (w) memcpy(w, o->w, 4ull * o->S);
  if (fire_t) memcpy(fire_t, o->fire_t, 4ull * o->S);
  if (last_arr) memcpy(last_arr, o->last_arr, 4ull * o->S);
  if (depol) memcpy(depol, o->depol, 4ull * o->S);
  if (time) *time = o->t;
}

OR_EXPORT void or_sim_counters(const or_sim* o, uint64_t out[6]) {
  out[5] = (uint64_t)o->first_tie;
  out[0] = o->ties_latch;
  out[1] = o->ties_weight;
  out[2] = o->arrivals;
  out[3] = o->fires;
  out[4] = o->accepted;
}

OR_EXPORT void or_sim_reset_activities(or_sim* o) { memset(o->firings, 0, 4ull * o->n); }

/* Synapse::setWeight (src/NeuCor.cpp:576-579) on the first n synapses (in-CSR
 * order): the weight, and the sign class re-derived from it. */
OR_EXPORT void or_sim_set_weights(or_sim* o, const float* w, uint64_t n) {
  for (uint64_t k = 0; k < n && k < o->S; ++k) {
    o->w[k] = w[k];
    o->inh[k] = w[k] < 0.0f ? 1 : 0;
  }
}

/* Closed-form pieces, for table checks. */
OR_EXPORT float or_trace(float decay, float x) { return powf(decay, x); }
OR_EXPORT float or_ap_value(float dt_since_fire) {
  or_sim tmp;
  memset(&tmp, 0, sizeof(tmp));
  float V = 0.0f, lf = 0.0f;
  tmp.V = &V;
  tmp.last_fire = &lf;
  action_potential(&tmp, 0, dt_since_fire);
  return V;
}

/* The raw background draws of steps s0+1 .. s0+steps for neurons 0 .. n-1
 * (the per-step Bernoulli(bg_prob) of or_sim_step, before the refractory
 * check the simulation applies), as (step, id) pairs ordered by step then id;
 * returns their number (entries past cap are counted, not written).  For
 * scripting the same background fires into the reference's own run(). */
OR_EXPORT uint64_t or_bg_events(uint64_t seed, float bg_prob, uint32_t n, uint32_t s0, uint32_t steps,
                                uint32_t* ev_steps, uint32_t* ev_ids, uint64_t cap) {
  uint32_t thr = 0;
  if (bg_prob > 0.0f) {
    double th = floor((double)bg_prob * 4294967296.0);
    thr = th >= 4294967295.0 ? 0xFFFFFFFFu : (uint32_t)th;
    if (thr == 0) thr = 1;
  }
  const uint64_t bg_seed = seed ^ 0x9E3779B97F4A7C15ull;
  uint64_t k = 0;
  for (uint32_t s = s0 + 1; thr && s <= s0 + steps; ++s)
    for (uint32_t i = 0; i < n; ++i) {
      uint32_t u[4];
      draw(bg_seed, s, i, ST_BG, u);
      if (u[0] < thr) {
        if (k < cap) {
          ev_steps[k] = s;
          ev_ids[k] = i;
        }
        ++k;
      }
    }
  return k;
}
