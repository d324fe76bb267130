// ref_harness.cpp -- headless driver around the reference's simulation core.
//
// TEST INFRASTRUCTURE ONLY.  Built by oracle/Makefile together with the
// reference's own, unmodified src/NeuCor.cpp (compiled where it lies under
// the reference checkout) into oracle/_ref/ref_harness.  Used to pin the C
// oracle bit-exactly on tie-free micro-networks, for statistical comparison,
// and as the CPU baseline.  Never linked into the product.
//
// Access to protected/private members goes through the reference's own
// friend declaration `friend class NeuCor_Renderer` (src/NeuCor.h:98, :270):
// this file defines a class of that name with static accessors only.
//
// rand() is supplied here so that the reference's background-fire loop
// (src/NeuCor.cpp:604-607) is scripted: it fires exactly the neurons the
// scenario asks for, at offset runSpeed (i.e. on the next step of the grid),
// and nothing else.  In --bench mode every draw also calls glibc's rand() so
// that the loop keeps its real cost.
//
// Scenario file (little endian):
//   u32 magic 'NCS1', u32 n, u64 S, u32 steps, u32 mode_flags,
//   f32 dt, lr, pre_f, post_f
//   i64 in_ptr[n+1]; u32 in_src[S]; f32 w[S]; u8 delay_steps[S]
//   per step: u32 n_in, u32 ids[n_in]; u32 n_bg, u32 ids[n_bg]
// Trace output (mode 0): u32 magic 'NCO1', n, steps, pad;
//   per step: u32 n_fired, u32 ids[n_fired], f32 V[n]
//   then f32 w[S], fire_t[S], last_arr[S] (synapse order of the file),
//   f32 last_fire[n]
#include <dlfcn.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <deque>
#include <vector>

#include "NeuCor.h"

// ---- scripted rand() -------------------------------------------------------
static std::deque<int> g_script;
static bool g_bench = false;
static bool g_libc_rand = false;  // --wiring: the C library's own stream
typedef int (*rand_fn)(void);
static rand_fn g_real_rand = nullptr;

extern "C" int rand(void) {
  if (g_bench || g_libc_rand) {
    if (!g_real_rand) g_real_rand = (rand_fn)dlsym(RTLD_NEXT, "rand");
    const int v = g_real_rand();
    if (g_libc_rand) return v;
  }
  if (g_script.empty()) return 1;  // 1 % period != 0: no background fire
  int v = g_script.front();
  g_script.pop_front();
  return v;
}

class NeuCor_Renderer {
 public:
  static std::deque<Neuron>& neurons(NeuCor& b) { return b.neurons; }
  static Synapse* syn(NeuCor& b, size_t from, size_t to) { return b.getSynapse(from, to); }
  static void set_length(Synapse& s, float L) { s.length = L; }
  static float fire_t(const Synapse& s) { return s.AP_fireTime; }
  static float last_arr(const Synapse& s) { return s.lastSpikeArrival; }
  static float length(const Synapse& s) { return s.length; }
  static size_t target(const Synapse& s) { return s.tN; }
};

// --wiring N SEED: the network NeuCor(N) builds after srand(SEED), printed as
// "n N", "p x y z" per neuron and "s from to weight length" per synapse in
// creation order (floats as C99 hex).
static int dump_wiring(int n, unsigned seed) {
  g_libc_rand = true;
  srand(seed);
  NeuCor brain(n);
  std::deque<Neuron>& neu = NeuCor_Renderer::neurons(brain);
  printf("n %zu\n", brain.getNeuronCount());
  for (const auto& s : brain.getNeuronSnapshots()) printf("p %a %a %a\n", s.position.x, s.position.y, s.position.z);
  for (auto& nn : neu)
    for (auto& sy : nn.outSynapses)
      printf("s %zu %zu %a %a\n", nn.getID(), NeuCor_Renderer::target(sy), sy.getWeight(), NeuCor_Renderer::length(sy));
  return 0;
}

// --standard SEED MS RUN_SPEED RUN_ALL WALK_EVERY RATE SKIP: the STANDARD
// preset of src/main.cpp:80-124 headless: NeuCor(750) after srand(SEED);
// three inputs on a circle of radius 2 (radius 0.8) at rates drawn as
// randomRate does (:44-46); every WALK_EVERY-th frame the rates take the
// preset's random walk (clamped to [0, 75], inputs[1] = inputs[0],
// :100-106) before run(); frames of RUN_SPEED ms with runAll = RUN_ALL until
// MS ms.  With RATE > 0 every input fires at RATE Hz throughout instead.
// SKIP extra rand() draws after construction shift the C library's stream
// (the background fires of src/NeuCor.cpp:604-607 draw from it).  Every
// synapse's sign class is repaired first (setWeight(getWeight())): the
// reference's Synapse copy leaves it unset when the out-synapse vectors grow
// (src/NeuCor.cpp:487-524), and its value is then whatever the memory held.  Fires are counted by
// polling lastFire after every frame (a neuron firing twice inside one frame
// counts once).  Prints "frames", "fires", "neurons_fired", "mean_w" and
// "w_hist" (20 bins over [-1, 1]).
static int standard(unsigned seed, double ms, float run_speed, int run_all, int walk_every, float rate,
                    int skip) {
  g_libc_rand = true;
  srand(seed);
  NeuCor brain(750);
  for (auto& n : NeuCor_Renderer::neurons(brain))
    for (auto& sy : n.outSynapses) {
      // the sign class, lost by vector growth (src/NeuCor.cpp:487-524)
      sy.setWeight(sy.getWeight());
    }
  brain.runSpeed = run_speed;
  brain.runAll = run_all != 0;
  for (int k = 0; k < skip; ++k) (void)rand();
  std::vector<float> inputs(3);
  for (float& x : inputs) x = (float)rand() / (float)RAND_MAX * 75.0f;
  if (rate > 0)
    for (float& x : inputs) x = rate;
  std::vector<float> radius(3, 0.8f);
  std::vector<coord3> pos = {{cosf(0.0f) * 2.0f, sinf(0.0f) * 2.0f, 0.0f},
                             {cosf(2.0944f) * 2.0f, sinf(2.0944f) * 2.0f, 0.0f},
                             {cosf(4.1888f) * 2.0f, sinf(4.1888f) * 2.0f, 0.0f}};
  brain.setInputRateArray(inputs.data(), 3, pos.data(), radius.data());
  std::deque<Neuron>& neu = NeuCor_Renderer::neurons(brain);
  std::vector<float> last(neu.size(), NAN);
  std::vector<uint8_t> ever(neu.size(), 0);
  uint64_t frames = 0, fires = 0;
  while (brain.getTime() < ms) {
    if (rate <= 0 && frames % (uint64_t)walk_every == 0) {
      for (float& x : inputs) {
        x += ((float)rand() / (float)RAND_MAX - 0.5f) * 2.0f;
        x = std::min(std::max(x, 0.0f), 75.0f);
      }
      inputs[1] = inputs[0];
    }
    brain.run();
    ++frames;
    for (size_t i = 0; i < neu.size(); ++i) {
      const float lf = neu[i].lastFire;
      if (lf == lf && !(lf == last[i])) {
        ++fires;
        ever[i] = 1;
        last[i] = lf;
      }
    }
  }
  uint64_t hist[20] = {};
  double sum_w = 0;
  uint64_t ns = 0;
  for (auto& n : neu)
    for (auto& sy : n.outSynapses) {
      const float w = sy.getWeight();
      int b = (int)std::floor((w + 1.0f) * 10.0f);
      b = std::min(std::max(b, 0), 19);
      ++hist[b];
      sum_w += w;
      ++ns;
    }
  uint64_t k = 0;
  for (uint8_t e : ever) k += e;
  printf("time %.3f\nframes %llu\nfires %llu\nneurons_fired %llu\nsynapses %llu\nmean_w %.9g\nw_hist",
         brain.getTime(), (unsigned long long)frames, (unsigned long long)fires, (unsigned long long)k,
         (unsigned long long)ns, ns ? sum_w / (double)ns : 0.0);
  for (uint64_t h : hist) printf(" %llu", (unsigned long long)h);
  printf("\n");
  return 0;
}

// --gap RUN_ALL: neuron 2 receives synapses from neurons 0 and 1 (weight
// 0.2, delay 4 ms); neuron 0 is fired at 1 ms and neuron 1 at 40 ms; frames
// of 1 ms with runAll = RUN_ALL until 60 ms.  Prints "t V lastFire" of
// neuron 2 after every frame.  With runAll = false the neuron runs only at
// its events, and charge_insynapses (src/NeuCor.cpp:688-700) integrates a
// still-latched synapse over the whole time since its last run.
static int gap(int run_all) {
  NeuCor brain(0);
  brain.runSpeed = 1.0f;
  brain.runAll = run_all != 0;
  for (int i = 0; i < 3; ++i) brain.createNeuron(coord3{(float)i, 0.0f, 0.0f});
  std::deque<Neuron>& neu = NeuCor_Renderer::neurons(brain);
  brain.createSynapse(2, 0, 0.2f);
  brain.createSynapse(2, 1, 0.2f);
  for (auto& n : neu)
    for (auto& sy : n.outSynapses) {
      NeuCor_Renderer::set_length(sy, 2.0f);
      sy.setWeight(sy.getWeight());
    }
  for (int f = 0; f < 60; ++f) {
    brain.run();
    if (f == 0) neu[0].fire();
    if (f == 39) neu[1].fire();
    printf("%g %a %a\n", brain.getTime(), neu[2].potential(), neu[2].lastFire);
  }
  return 0;
}

// --psp DELAY: neuron 1 receives one synapse from neuron 0 (weight 0.2,
// delay DELAY ms = 2 * length); neuron 0 is fired at 1 ms; runAll, frames of
// 1 ms until 12 ms.  Prints "t V" of neuron 1 after every frame.  An
// arrival off the 1 ms grid runs the neuron at the arrival and again 2 ms
// later (Synapse::run, Neuron::transfer, src/NeuCor.cpp:663-666, :718-726),
// steps of fractional dt between the grid's.
static int psp(float delay) {
  NeuCor brain(0);
  brain.runSpeed = 1.0f;
  brain.runAll = true;
  brain.createNeuron(coord3{0.0f, 0.0f, 0.0f});
  brain.createNeuron(coord3{1.0f, 0.0f, 0.0f});
  std::deque<Neuron>& neu = NeuCor_Renderer::neurons(brain);
  brain.createSynapse(1, 0, 0.2f);
  for (auto& sy : neu[0].outSynapses) {
    NeuCor_Renderer::set_length(sy, delay / 2.0f);
    sy.setWeight(sy.getWeight());
  }
  for (int f = 0; f < 12; ++f) {
    brain.run();
    if (f == 0) neu[0].fire();
    printf("%g %a\n", brain.getTime(), neu[1].potential());
  }
  return 0;
}

struct Scenario {
  uint32_t n = 0, steps = 0, flags = 0;
  uint64_t S = 0;
  float dt = 1, lr = 1, pre_f = 0.13f, post_f = 0.30f;
  std::vector<int64_t> in_ptr;
  std::vector<uint32_t> in_src;
  std::vector<float> w;
  std::vector<uint8_t> delay;
  std::vector<std::vector<uint32_t>> in_ev, bg_ev;
};

static bool rd(FILE* f, void* p, size_t n) { return fread(p, 1, n, f) == n; }

static bool load(const char* path, Scenario* sc) {
  FILE* f = fopen(path, "rb");
  if (!f) return false;
  uint32_t magic = 0;
  bool ok = rd(f, &magic, 4) && magic == 0x3153434eu && rd(f, &sc->n, 4) && rd(f, &sc->S, 8) &&
            rd(f, &sc->steps, 4) && rd(f, &sc->flags, 4) && rd(f, &sc->dt, 4) && rd(f, &sc->lr, 4) &&
            rd(f, &sc->pre_f, 4) && rd(f, &sc->post_f, 4);
  if (ok) {
    sc->in_ptr.resize(sc->n + 1);
    sc->in_src.resize(sc->S);
    sc->w.resize(sc->S);
    sc->delay.resize(sc->S);
    ok = rd(f, sc->in_ptr.data(), 8 * (sc->n + 1)) && rd(f, sc->in_src.data(), 4 * sc->S) &&
         rd(f, sc->w.data(), 4 * sc->S) && rd(f, sc->delay.data(), sc->S);
  }
  sc->in_ev.resize(sc->steps);
  sc->bg_ev.resize(sc->steps);
  for (uint32_t s = 0; ok && s < sc->steps; ++s) {
    uint32_t k = 0;
    ok = rd(f, &k, 4);
    sc->in_ev[s].resize(k);
    ok = ok && rd(f, sc->in_ev[s].data(), 4 * k);
    ok = ok && rd(f, &k, 4);
    sc->bg_ev[s].resize(k);
    ok = ok && rd(f, sc->bg_ev[s].data(), 4 * k);
  }
  fclose(f);
  return ok;
}

int main(int argc, char** argv) {
  if (argc < 3) {
    fprintf(stderr, "usage: ref_harness [--bench] scenario.bin out.bin\n");
    return 2;
  }
  if (argc == 4 && strcmp(argv[1], "--wiring") == 0) return dump_wiring(atoi(argv[2]), (unsigned)strtoul(argv[3], 0, 10));
  if (argc == 3 && strcmp(argv[1], "--gap") == 0) return gap(atoi(argv[2]));
  if (argc == 9 && strcmp(argv[1], "--standard") == 0)
    return standard((unsigned)strtoul(argv[2], 0, 10), atof(argv[3]), (float)atof(argv[4]), atoi(argv[5]),
                    std::max(1, atoi(argv[6])), (float)atof(argv[7]), atoi(argv[8]));
  if (argc == 3 && strcmp(argv[1], "--psp") == 0) return psp((float)atof(argv[2]));
  int a = 1;
  uint32_t warmup = 0;  // --bench: untimed steps before the timed ones
  if (strcmp(argv[a], "--bench") == 0) {
    g_bench = true;
    ++a;
    if (a + 1 < argc && strcmp(argv[a], "--warmup") == 0) {
      warmup = (uint32_t)strtoul(argv[a + 1], 0, 10);
      a += 2;
    }
  }
  if (argc - a < 2) return 2;
  Scenario sc;
  if (!load(argv[a], &sc)) {
    fprintf(stderr, "cannot read scenario %s\n", argv[a]);
    return 1;
  }
  auto t_build0 = std::chrono::steady_clock::now();
  NeuCor brain(0);
  brain.runAll = true;
  brain.runSpeed = sc.dt;
  brain.learningRate = sc.lr;
  brain.presynapticFactor = sc.pre_f;
  brain.postsynapticFactor = sc.post_f;
  for (uint32_t i = 0; i < sc.n; ++i) brain.createNeuron(coord3{0.0f, 0.0f, (float)i});
  std::deque<Neuron>& neu = NeuCor_Renderer::neurons(brain);
  for (uint32_t tg = 0; tg < sc.n; ++tg) {
    for (int64_t k = sc.in_ptr[tg]; k < sc.in_ptr[tg + 1]; ++k) {
      uint32_t src = sc.in_src[k];
      brain.createSynapse(tg, src, sc.w[k]);
      // Synapse delay is length * AP_speed (src/NeuCor.cpp:733); AP_speed = 2.
      NeuCor_Renderer::set_length(neu[src].outSynapses.back(), (float)sc.delay[k] * sc.dt / 2.0f);
    }
  }
  // Repair the sign-class flag lost by vector growth (src/NeuCor.cpp:487-524).
  for (auto& n : neu)
    for (auto& s : n.outSynapses) s.setWeight(s.getWeight());
  double build_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_build0).count();

  FILE* out = fopen(argv[a + 1], "wb");
  if (!out) return 1;
  if (g_bench) {
    uint64_t fires = 0;
    if (warmup > sc.steps) warmup = sc.steps;
    // background fires scripted as in the trace mode below (the real
    // rand() is still drawn for every trial of the loop)
    auto one_step = [&](uint32_t s) {
      for (uint32_t id : sc.bg_ev[s]) {
        g_script.push_back(0);
        g_script.push_back((int)id);
        g_script.push_back(RAND_MAX);
      }
      brain.run();
      g_script.clear();
      for (uint32_t id : sc.in_ev[s]) neu[id].fire();
    };
    for (uint32_t s = 0; s < warmup; ++s) one_step(s);
    auto t0 = std::chrono::steady_clock::now();
    for (uint32_t s = warmup; s < sc.steps; ++s) one_step(s);
    double run_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    for (auto& n : neu)
      if (n.lastFire == n.lastFire) ++fires;
    fprintf(out, "{\"build_s\": %.6f, \"run_s\": %.6f, \"steps\": %u, \"warmup\": %u, \"neurons\": %u, "
                 "\"synapses\": %llu, \"neurons_fired_ever\": %llu}\n",
            build_s, run_s, sc.steps - warmup, warmup, sc.n, (unsigned long long)sc.S, (unsigned long long)fires);
    fclose(out);
    return 0;
  }
  uint32_t hdr[4] = {0x314f434eu, sc.n, sc.steps, 0u};
  fwrite(hdr, 4, 4, out);
  std::vector<float> V(sc.n);
  std::vector<uint32_t> fired;
  for (uint32_t s = 0; s < sc.steps; ++s) {
    // Background fires of step s+1 are drawn by the run() that advances to
    // it: hit (0), neuron (id), offset randomUnit() * runSpeed = runSpeed.
    for (uint32_t id : sc.bg_ev[s]) {
      g_script.push_back(0);
      g_script.push_back((int)id);
      g_script.push_back(RAND_MAX);
    }
    brain.run();
    g_script.clear();
    const float t = brain.getTime();
    for (uint32_t id : sc.in_ev[s]) neu[id].fire();
    for (auto& n : neu) n.run();  // what runAll does at the start of the next run()
    fired.clear();
    for (uint32_t i = 0; i < sc.n; ++i) {
      if (neu[i].lastFire == t) fired.push_back(i);
      V[i] = neu[i].potential();
    }
    uint32_t nf = (uint32_t)fired.size();
    fwrite(&nf, 4, 1, out);
    fwrite(fired.data(), 4, nf, out);
    fwrite(V.data(), 4, sc.n, out);
  }
  std::vector<float> w(sc.S), ft(sc.S), la(sc.S), lf(sc.n);
  for (uint32_t tg = 0; tg < sc.n; ++tg) {
    for (int64_t k = sc.in_ptr[tg]; k < sc.in_ptr[tg + 1]; ++k) {
      Synapse* sy = NeuCor_Renderer::syn(brain, sc.in_src[k], tg);
      w[k] = sy->getWeight();
      ft[k] = NeuCor_Renderer::fire_t(*sy);
      la[k] = NeuCor_Renderer::last_arr(*sy);
    }
  }
  for (uint32_t i = 0; i < sc.n; ++i) lf[i] = neu[i].lastFire;
  fwrite(w.data(), 4, sc.S, out);
  fwrite(ft.data(), 4, sc.S, out);
  fwrite(la.data(), 4, sc.S, out);
  fwrite(lf.data(), 4, sc.n, out);
  fclose(out);
  return 0;
}
