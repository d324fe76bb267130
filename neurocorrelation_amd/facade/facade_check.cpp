// Command-line driver of class NeuCor (this directory) for the tests.
//
//   facade_check wiring N SEED
//       srand(SEED); NeuCor(N); prints the network (no GPU needed):
//       "n N", then "p x y z" per neuron, then "s from to weight length"
//       per synapse in creation order (floats as C99 hex).
//   facade_check few RUNS
//       the FEW_NEURONS preset of src/main.cpp:162-189 (3 neurons, 0 -> 1 and
//       0 -> 2 at w = 0.5, inputs at 50 Hz with offsets +2 / -2 ms,
//       runAll, runSpeed 0.02 ms) for RUNS calls of run(); prints
//       "t TIME", "dt STEP_MS", "w FROM TO W" per synapse, "v ID V" per
//       neuron (hex) and "r R", the next value of the caller's rand().
//   facade_check mirror RUNS
//       NeuCor(120) with two inputs for RUNS runs of 1 ms; the Neuron /
//       Synapse mirror must match the snapshots, and a mirror setWeight must
//       reach the engine (prints a summary; exit status 1 on a mismatch).
//   facade_check standard SEED MS RUN_SPEED WALK_EVERY RATE SKIP
//       the STANDARD preset of src/main.cpp:80-124 through the facade, as
//       the reference harness's --standard mode runs it (runAll is true
//       here whatever the caller sets): NeuCor(750) after srand(SEED), SKIP
//       extra rand() draws, three inputs (randomRate, or RATE Hz if RATE >
//       0) with the preset's random walk every WALK_EVERY-th frame, frames of
//       RUN_SPEED ms until MS ms; fires counted by polling lastFire after
//       every frame.  Prints "fires", "neurons_fired", "mean_w", "w_hist".
//   facade_check grow RUNS
//       the mirror after RUNS runs, then after a createNeuron/createSynapse
//       and a run too short to step (the engine network rebuilt), then 50
//       more runs.
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <exception>
#include <vector>

#include "NeuCor.h"

static int wiring(int n, unsigned seed) {
  srand(seed);
  NeuCor brain(n);
  printf("n %zu\n", brain.getNeuronCount());
  for (const auto& s : brain.getNeuronSnapshots())
    printf("p %a %a %a\n", s.position.x, s.position.y, s.position.z);
  for (const auto& s : brain.getSynapseSnapshots())
    printf("s %zu %zu %a %a\n", s.fromID, s.toID, s.weight, s.from.getDist(s.to));
  return 0;
}

static void dump_mirror(NeuCor& brain);

static int few(long runs) {
  NeuCor brain(0);
  brain.runAll = true;
  brain.runSpeed = 0.02f;
  const coord3 pos[3] = {{0, 0, 0}, {0.3f, 0.3f, 0}, {0.3f, -0.3f, 0}};
  for (const coord3& p : pos) brain.createNeuron(p);
  brain.createSynapse(1, 0, 0.5f);
  brain.createSynapse(2, 0, 0.5f);
  float inputs[3] = {50.0f, 50.0f, 50.0f};
  coord3 ipos[3] = {pos[0], pos[1], pos[2]};
  float irad[3] = {0.1f, 0.1f, 0.1f};
  brain.setInputRateArray(inputs, 3, ipos, irad);
  brain.addInputOffset(1, 2.0f);
  brain.addInputOffset(2, -2.0f);
  for (long k = 0; k < runs; ++k) brain.run();
  printf("t %a\n", brain.getTime());
  printf("dt %a\n", brain.stepMs);
  for (const auto& s : brain.getSynapseSnapshots()) printf("w %zu %zu %a\n", s.fromID, s.toID, s.weight);
  for (const auto& s : brain.getNeuronSnapshots()) printf("v %zu %a\n", s.id, s.potential);
  printf("r %d\n", rand());  // the caller's rand() stream after the run
  dump_mirror(brain);
  return 0;
}

// The renderer's view after a run: every out-synapse in the Neuron mirror
// against the snapshots, then a setWeight that must reach the engine.
class NeuCor_Renderer {
 public:
  static int check(long runs) {
    srand(3);
    NeuCor brain(120);
    brain.runAll = true;
    brain.runSpeed = 1.0f;
    float inputs[2] = {40.0f, 25.0f};
    brain.setInputRateArray(inputs, 2);
    for (long k = 0; k < runs; ++k) brain.run();
    const auto snaps = brain.getSynapseSnapshots();
    std::size_t i = 0, bad = 0, latched = 0;
    for (auto& neu : brain.neurons)
      for (auto& syn : neu.outSynapses) {
        const auto& sn = snaps[i++];
        if (sn.fromID != syn.pN || sn.toID != syn.tN || sn.weight != syn.getWeight() ||
            sn.prePotential != syn.getPrePot() || sn.postPotential != syn.getPostPot())
          ++bad;
        if (syn.AP_fireTime != 0.0f) ++latched;
        if (brain.getSynapse(syn.pN, syn.tN) != &syn) ++bad;
        if (brain.getNeuron(syn.tN)->inSynapses.count(syn.pN) != 1) ++bad;
      }
    std::size_t fired = 0;
    for (auto& neu : brain.neurons) fired += neu.lastFire == neu.lastFire;
    Synapse* s0 = &brain.neurons[0].outSynapses[0];
    s0->setWeight(-0.25f);
    brain.run();
    const float after = brain.getSynapseSnapshots()[0].weight;
    printf("synapses %zu mismatches %zu latched %zu fired %zu weight_after_set %a\n", i, bad, latched, fired, after);
    return bad == 0 ? 0 : 1;
  }

  // What the renderer reads of each neuron and out-synapse
  // (src/NeuCor_Renderer.cpp:655-699, :840-866, :1349-1490):
  // "m FROM TO W AP_FIRETIME LASTARRIVAL" and "l ID LASTFIRE POTENTIAL" (hex).
  static float last_fire(NeuCor& brain, std::size_t i) { return brain.neurons[i].lastFire; }
  static void dump(NeuCor& brain) {
    for (auto& neu : brain.neurons) {
      for (auto& syn : neu.outSynapses)
        printf("m %zu %zu %a %a %a\n", syn.pN, syn.tN, syn.getWeight(), syn.AP_fireTime, syn.lastSpikeArrival);
      printf("l %zu %a %a\n", neu.getID(), neu.lastFire, neu.potential());
    }
  }
};

static void dump_mirror(NeuCor& brain) { NeuCor_Renderer::dump(brain); }

// NeuCor(120) run RUNS runs of 1 ms; the mirror; then a neuron and a synapse
// created (the engine network is rebuilt at the next run) and a run too
// short to reach the next grid step; the mirror again ("after" section);
// then 50 more runs of 1 ms.
static int grow(long runs) {
  srand(3);
  NeuCor brain(120);
  brain.runAll = true;
  brain.runSpeed = 1.0f;
  float inputs[2] = {40.0f, 25.0f};
  brain.setInputRateArray(inputs, 2);
  for (long k = 0; k < runs; ++k) brain.run();
  printf("before\n");
  dump_mirror(brain);
  brain.createNeuron(coord3{9.0f, 9.0f, 9.0f});
  brain.createSynapse(brain.getNeuronCount() - 1, 0, 0.5f);
  brain.runSpeed = 1.0f / 1024;
  brain.run();
  printf("after\n");
  dump_mirror(brain);
  brain.runSpeed = 1.0f;
  for (int k = 0; k < 50; ++k) brain.run();
  printf("t %a\n", brain.getTime());
  return 0;
}

static int standard(unsigned seed, double ms, float run_speed, int walk_every, float rate, int skip) {
  srand(seed);
  NeuCor brain(750);
  brain.runSpeed = run_speed;
  brain.runAll = false;  // the preset's setting; the engine steps every neuron every step
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
  const std::size_t n = brain.getNeuronCount();
  std::vector<float> last(n, NAN);
  std::vector<char> ever(n, 0);
  unsigned long long frames = 0, fires = 0;
  while (brain.getTime() < ms) {
    if (rate <= 0 && frames % (unsigned long long)walk_every == 0) {
      for (float& x : inputs) {
        x += ((float)rand() / (float)RAND_MAX - 0.5f) * 2.0f;
        x = std::min(std::max(x, 0.0f), 75.0f);
      }
      inputs[1] = inputs[0];
    }
    brain.run();
    ++frames;
    for (std::size_t i = 0; i < n; ++i) {
      const float lf = NeuCor_Renderer::last_fire(brain, i);
      if (lf == lf && !(lf == last[i])) {
        ++fires;
        ever[i] = 1;
        last[i] = lf;
      }
    }
  }
  unsigned long long hist[20] = {}, ns = 0, k = 0;
  double sum_w = 0;
  for (const auto& sn : brain.getSynapseSnapshots()) {
    int b = (int)floorf((sn.weight + 1.0f) * 10.0f);
    b = std::min(std::max(b, 0), 19);
    ++hist[b];
    sum_w += sn.weight;
    ++ns;
  }
  for (char e : ever) k += e;
  printf("time %.3f\nframes %llu\nfires %llu\nneurons_fired %llu\nsynapses %llu\nmean_w %.9g\nw_hist", brain.getTime(),
         frames, fires, k, ns, ns ? sum_w / (double)ns : 0.0);
  for (unsigned long long h : hist) printf(" %llu", h);
  printf("\n");
  return 0;
}

int main(int argc, char** argv) {
  try {
    if (argc == 4 && strcmp(argv[1], "wiring") == 0) return wiring(atoi(argv[2]), (unsigned)strtoul(argv[3], 0, 10));
    if (argc == 3 && strcmp(argv[1], "few") == 0) return few(atol(argv[2]));
    if (argc == 3 && strcmp(argv[1], "mirror") == 0) return NeuCor_Renderer::check(atol(argv[2]));
    if (argc == 3 && strcmp(argv[1], "grow") == 0) return grow(atol(argv[2]));
    if (argc == 8 && strcmp(argv[1], "standard") == 0)
      return standard((unsigned)strtoul(argv[2], 0, 10), atof(argv[3]), (float)atof(argv[4]),
                      std::max(1, atoi(argv[5])), (float)atof(argv[6]), atoi(argv[7]));
  } catch (const std::exception& e) {
    fprintf(stderr, "facade_check: %s\n", e.what());
    return 1;
  }
  fprintf(stderr,
          "usage: facade_check wiring N SEED | few RUNS | mirror RUNS | grow RUNS | "
          "standard SEED MS RUN_SPEED WALK_EVERY RATE SKIP\n");
  return 2;
}
