// Every access the reference's front end makes to the simulation core,
// written against this directory's NeuCor.h.
//
// The reference's renderer (src/NeuCor_Renderer.cpp) and presets
// (src/main.cpp) reach NeuCor through its public interface and, for the
// renderer, through `friend class NeuCor_Renderer` (src/NeuCor.h:98, :270).
// Their GUI libraries (OpenGL, GLFW, ImGui, glm) are not part of this build,
// so those files themselves are not compiled here; this file repeats each of
// their expressions on the simulation core -- same member names, same uses
// of the results' types -- in a class of the renderer's name, so that a
// façade that drops a member or changes a type fails to compile.  Each
// expression cites the line it mirrors.  Compiled by the tests with
// -fsyntax-only; `renderer_access_demo` also runs it on a small network.
#include <math.h>
#include <stdio.h>

#include <cstddef>
#include <deque>
#include <map>
#include <type_traits>
#include <utility>
#include <vector>

#include "NeuCor.h"

class NeuCor_Renderer {
 public:
  explicit NeuCor_Renderer(NeuCor* b) : brain(b) {}

  // The per-frame path: run with a wall-clock-scaled speed, then read back
  // what the renderer draws.  Returns a checksum of what was read.
  double frame(float deltaTime, bool realRunspeed) {
    double acc = 0.0;
    // src/NeuCor_Renderer.cpp:632-638
    if (realRunspeed) {
      float staticRunSpeed = brain->runSpeed;
      brain->runSpeed = staticRunSpeed * deltaTime;
      brain->run();
      brain->runSpeed = staticRunSpeed;
    } else {
      brain->run();
    }
    // :471-484 buffer sizes from the neuron and position counts
    acc += (double)(brain->positions.size() * 3 * sizeof(float)) + (double)(brain->neurons.size() * 5);
    // :655-699 every out-synapse of every neuron: endpoints, pulses, activity
    std::vector<coord3> connections;
    std::vector<float> synPot;
    synPot.reserve(brain->neurons.size() * 8.0);
    for (auto& neu : brain->neurons) {
      for (auto& syn : neu.outSynapses) {
        connections.push_back(brain->getNeuron(syn.pN)->position());
        connections.push_back(brain->getNeuron(syn.tN)->position());
        synPot.push_back(syn.getPrePot() + 0.03);
        synPot.push_back(syn.getPostPot() + 0.03);
        synPot.back() *= log(brain->getNeuron(syn.tN)->activity() + 1.f);
        synPot.push_back(log(brain->getNeuron(syn.pN)->activity() + 1.f));
      }
    }
    for (const coord3& c : connections) acc += c.x + c.y + c.z;
    for (float v : synPot) acc += v;
    // :773-779 raw GL uploads: positions as 3 floats, potential/activity interleaved
    unsigned neuronC = brain->positions.size();
    static_assert(std::is_same<decltype(brain->positions.data()), coord3*>::value, "positions");
    static_assert(std::is_same<decltype(brain->potAct.data()), float*>::value, "potAct");
    static_assert(sizeof(coord3) == 3 * sizeof(float), "coord3 is three packed floats");
    const float* pa = brain->potAct.data();
    for (unsigned i = 0; i < neuronC * 2 && i < brain->potAct.size(); ++i) acc += pa[i];
    // :840-853 a neuron's snapshot for the plot windows
    float brainTime = brain->getTime();
    if (!brain->neurons.empty()) {
      Neuron* neu = brain->getNeuron(0);
      float voltage = neu->potential();
      acc += voltage + brainTime + (double)neu->outSynapses.size();
      for (auto& syn : neu->outSynapses) acc += syn.getWeight();
      // :1349-1403 the neuron window
      coord3 pos3D = neu->position();
      acc += pos3D.x + neu->activity();
      float since = brain->getTime() - neu->lastFire;
      if (since == since) acc += since;
      // :1433-1451 incoming synapses by their (from, to) key
      for (auto& synM : neu->inSynapses) {
        Synapse* syn = brain->getSynapse(synM);
        acc += floor(50 + syn->getPrePot() * 180.0f);
        acc += (double)syn->pN + (double)syn->tN + syn->getWeight() + (syn->inhibitory ? 1 : 0);
      }
      // :916-922, :1256, :1276 positions of inputs and detectors
      for (const auto& input : brain->inputHandler)
        if (input.a.x == input.a.x && input.a.y == input.a.y && input.a.z == input.a.z) acc += input.a.z;
      // :1478-1483 outgoing synapses
      for (auto& syn : neu->outSynapses) acc += (double)syn.tN + (double)syn.pN + syn.getPostPot();
      // :1092-1097 closeness: walk the out-synapses' targets
      for (auto& outSyn : brain->getNeuron(0)->outSynapses) acc += (double)outSyn.tN;
      acc += (double)neu->getID();
    }
    // :908-920 scene points from positions, inputs and detectors
    for (const coord3& pos : brain->positions) acc += pos.z;
    for (const auto& input : brain->inputHandler) acc += input.a.x + input.radius;
    for (const auto& detector : brain->voltageDetectors) acc += detector.a.y + detector.radius;
    // :1248-1263 input list: enable toggles, last fire, rate
    int inputHandlerID = 0;
    for (auto& inFi : brain->inputHandler) {
      if (inFi.enabled) acc += 0.8 - 10.0f / ((float)((float)brain->getTime() - inFi.lastFire) * 10.0f + 1.0f);
      inFi.enabled = !inFi.enabled;  // the UI's checkbox writes it
      inFi.enabled = !inFi.enabled;
      acc += brain->inputArray[inputHandlerID];
      ++inputHandlerID;
    }
    // :1272-1274 detectors
    for (unsigned k = 0; k < brain->voltageDetectors.size(); ++k) acc += brain->getDetectorVoltage(k);
    // :1615-1619 activity reset
    std::vector<float> lastA;
    lastA.reserve(brain->neurons.size());
    for (auto& neu : brain->neurons) lastA.push_back(neu.activity());
    brain->resetActivities();
    // :962-968, :1664, :1696, :1712 sliders and keys write the parameters
    float* lr = &brain->learningRate;
    float* speed = &brain->runSpeed;
    bool* all = &brain->runAll;
    brain->runSpeed = powf(brain->runSpeed, 0.99);
    brain->runSpeed = powf(brain->runSpeed, 1.01);
    acc += *lr + *speed + (*all ? 1 : 0);
    // :582, :605 id checks
    acc += (double)(brain->neurons.size() > 0);
    return acc;
  }

 private:
  NeuCor* brain;
};

// The presets' calls on the public interface (src/main.cpp:80-210).
static void presets_compile(NeuCor& brain, bool read_back) {
  // the rate array is borrowed and re-read by every run() (src/NeuCor.cpp:46-48,
  // :600): it must outlive the calls, as the presets' session members do
  static std::vector<float> inputs = {10.0f, 20.0f, 30.0f};
  static std::vector<float> inputRadius = {0.8f, 0.8f, 0.8f};
  static std::vector<coord3> inputPositions = {{cosf(0.0f) * 2.0f, sinf(0.0f) * 2.0f, 0.0f},
                                               {cosf(2.0944f) * 2.0f, sinf(2.0944f) * 2.0f, 0.0f},
                                               {cosf(4.1888f) * 2.0f, sinf(4.1888f) * 2.0f, 0.0f}};
  // STANDARD (:80-124)
  brain.runSpeed = 4;
  brain.setInputRateArray(inputs.data(), inputs.size(), inputPositions.data(), inputRadius.data());
  if (read_back && 10000.0f < brain.getTime()) brain.learningRate = 0;
  // USER_INPUT (:126-159)
  brain.runAll = true;
  brain.runSpeed = 0.02f;
  brain.setInputRateArray(inputs.data(), inputs.size());
  brain.setDetectors(1);
  // FEW_NEURONS (:161-189)
  const std::vector<coord3> neuronPositions = {{0, 0, 0}, {0.3f, 0.3f, 0}, {0.3f, -0.3f, 0}};
  for (const coord3& p : neuronPositions) brain.createNeuron(p);
  std::size_t n = brain.getNeuronCount();
  brain.createSynapse(n - 2, n - 3, 0.5f);
  brain.createSynapse(n - 1, n - 3, 0.5f);
  brain.addInputOffset(1, 2.0f);
  brain.addInputOffset(2, -2.0f);
  // ONE_INPUT (:191-210)
  brain.runAll = false;
  brain.runSpeed = 4.0f;
  // the snapshot API (src/NeuCor.h:64-91)
  if (read_back) {
    auto ns = brain.getNeuronSnapshots();
    auto ss = brain.getSynapseSnapshots();
    auto is = brain.getInputSnapshots();
    auto dv = brain.getDetectorVoltages();
    printf("snapshots %zu neurons %zu synapses %zu inputs %zu detectors\n", ns.size(), ss.size(), is.size(),
           dv.size());
  }
}

#ifdef NC_RENDERER_ACCESS_MAIN
// Runs the renderer's accesses on NeuCor(60) for a few frames.
int main() {
  srand(3);
  NeuCor brain(60);
  presets_compile(brain, false);
  brain.runAll = true;
  brain.runSpeed = 1.0f;
  NeuCor_Renderer r(&brain);
  double acc = 0.0;
  for (int f = 0; f < 5; ++f) acc += r.frame(0.5f, f % 2 == 0);
  presets_compile(brain, true);
  printf("renderer access ok %d\n", acc == acc ? 1 : 0);
  return acc == acc ? 0 : 1;
}
#endif
