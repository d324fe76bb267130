// class NeuCor over the MI355X engine.
//
// The public interface of NeuroCorrelation's NeuCor (src/NeuCor.h:36-91),
// with the same member names, types and defaults, so that code written
// against the reference (its presets in src/main.cpp, tests, batch drivers)
// compiles and runs unchanged against this header.  The simulation itself
// runs on the GPU through the C ABI of include/neucor_amd.h; this class keeps
// the host-side objects the reference's callers touch: positions, the
// (potential, activity) array, input firers and voltage detectors.
//
// Clock: the engine steps on a grid of stepMs (a power of two, at most 1 ms,
// fixed at the first run()).  Simulated time is the reference's own float
// clock (currentTime += runSpeed, src/NeuCor.cpp:609-616); after each run()
// the engine has executed every grid step at or before it, so a caller with
// a fractional or wall-clock-scaled runSpeed (src/NeuCor_Renderer.cpp:633-636)
// sees the same simulated time as the reference.  Input and background fires
// land on the first grid step at or after their time.  Synaptic delays
// (2 * length ms, src/NeuCor.cpp:733) are rounded to whole steps; every
// neuron is updated every step (runAll is taken as true).
#pragma once

#include <math.h>

#include <cstddef>
#include <cstdint>
#include <deque>
#include <map>
#include <utility>
#include <vector>

struct nc_net;
struct nc_sim;

// src/NeuCor.h:13-24
struct coord3 {
  float x, y, z;
  float getDist(coord3 c) const { return sqrtf(powf(x - c.x, 2) + powf(y - c.y, 2) + powf(z - c.z, 2)); }
  void setNAN() { x = y = z = NAN; }
};

class NeuCor;

// One input (src/NeuCor.h:167-180): fires every neuron within `radius` of
// `a` at `frequency` Hz.
struct InputFirer {
  InputFirer(NeuCor* p, coord3 position = {NAN, NAN, NAN}, float radius = 1.0f);
  bool enabled;
  coord3 a;
  float radius;
  std::vector<std::size_t> near;
  float lastFire;  // the last scheduled fire time (ms)
};

// Mean potential of the neurons within `radius` of `a` (src/NeuCor.h:182-191).
struct VoltageDetector {
  VoltageDetector(NeuCor* p, coord3 position = {NAN, NAN, NAN}, float radius = 1.0f);
  NeuCor* parentNet;
  coord3 a;
  float radius;
  std::vector<std::size_t> near;
  float getVoltage();
};

// A synapse as the reference's renderer sees it (src/NeuCor.h:255-294): a
// host mirror of the engine's state, refreshed after every run().
// setWeight writes through to the engine at the next run().
class Synapse {
 public:
  Synapse(NeuCor* p, std::size_t parent, std::size_t target, float weight, float length);
  float getWeight() const { return weight; }
  void setWeight(float w);

 protected:
  friend class NeuCor;
  friend class NeuCor_Renderer;
  float getPrePot() const;   // the renderer's pulse at the source end
  float getPostPot() const;  // and at the target end
  std::size_t pN;            // parent (source) neuron id
  std::size_t tN;            // target neuron id
  float lastSpikeStart;      // when the latched spike was fired (ms)
  float lastSpikeArrival;    // when a spike last arrived (ms; -inf: never)
  float AP_fireTime;         // arrival time of the latched spike (ms; 0: none)
  float AP_speed = 2.0f;     // ms per unit of length
  bool inhibitory;

 private:
  NeuCor* parentNet;
  float length;
  float weight;
  bool written = false;  // setWeight since the last run()
};

// A neuron as the reference's renderer sees it (src/NeuCor.h:194-250).
class Neuron {
 public:
  Neuron(NeuCor* p, std::size_t id);
  std::vector<Synapse> outSynapses;               // owns its outgoing synapses
  std::map<std::size_t, std::size_t> inSynapses;  // {from, to} of its incoming ones
  coord3 position() const;
  float potential() const;
  float activity() const;
  std::size_t getID() const { return ownID; }
  float lastFire;  // simulation time of the last fire (ms; NaN: never)
  std::size_t pos, PA;

 private:
  NeuCor* parentNet;
  std::size_t ownID;
};

class NeuCor {
 public:
  struct NeuronSnapshot {
    std::size_t id;
    coord3 position;
    float potential;
    float activity;
  };
  struct SynapseSnapshot {
    std::size_t fromID;
    std::size_t toID;
    coord3 from;
    coord3 to;
    float weight;
    float prePotential;
    float postPotential;
    bool inhibitory;
  };
  struct InputSnapshot {
    std::size_t id;
    coord3 position;
    float radius;
    float frequency;
    bool enabled;
  };

  NeuCor(int n_neurons);
  ~NeuCor();
  NeuCor(const NeuCor&) = delete;
  NeuCor& operator=(const NeuCor&) = delete;

  void run();
  float runSpeed;
  bool runAll;
  float getTime() const;
  float learningRate;
  float presynapticTraceDecay, postsynapticTraceDecay;
  float presynapticFactor, postsynapticFactor;

  void setInputRateArray(float inputs[], unsigned inputCount, coord3 inputPositions[] = nullptr,
                         float inputRadius[] = nullptr);
  void addInputOffset(unsigned inputID, float t);

  void setDetectors(unsigned detectorNumber, coord3 detectorPositions[] = nullptr, float detectorRadius[] = nullptr);
  float getDetectorVoltage(unsigned ID);
  std::vector<float> getDetectorVoltages();

  void createNeuron(coord3 position);
  void createSynapse(std::size_t toID, std::size_t fromID, float weight);
  void makeConnections();
  std::size_t getNeuronCount() const;
  std::vector<NeuronSnapshot> getNeuronSnapshots() const;
  std::vector<SynapseSnapshot> getSynapseSnapshots() const;
  std::vector<InputSnapshot> getInputSnapshots() const;

  // No counterpart in the reference: the engine's step (ms, a power of two;
  // 0 = chosen at the first run() from runSpeed), HIP device, and whether
  // run() refreshes the Neuron/Synapse mirror (off: snapshots and potAct
  // stay current, the per-synapse mirror does not).
  float stepMs = 0.0f;
  int device = 0;
  bool mirror = true;

 protected:
  friend class NeuCor_Renderer;
  friend struct InputFirer;
  friend struct VoltageDetector;
  friend class Neuron;
  friend class Synapse;

  std::deque<Neuron> neurons;  // by id
  Neuron* getNeuron(std::size_t ID) { return &neurons.at(ID); }
  const Neuron* getNeuron(std::size_t ID) const { return &neurons.at(ID); }
  Synapse* getSynapse(std::size_t fromID, std::size_t toID);
  const Synapse* getSynapse(std::size_t fromID, std::size_t toID) const;
  Synapse* getSynapse(std::pair<std::size_t, std::size_t> ID) { return getSynapse(ID.first, ID.second); }
  const Synapse* getSynapse(std::pair<std::size_t, std::size_t> ID) const {
    return getSynapse(ID.first, ID.second);
  }

  std::vector<coord3> positions;  // by neuron id
  std::vector<float> potAct;      // potential, activity, potential, ... (by id)
  void resetActivities();

  float* inputArray;  // borrowed, re-read by every run() (src/NeuCor.cpp:46-48, :600)
  unsigned inputArraySize;
  std::vector<InputFirer> inputHandler;
  std::vector<VoltageDetector> voltageDetectors;

 private:
  void connectNeuron(std::size_t id);  // Neuron::makeConnections (src/NeuCor.cpp:418-444)
  void addSynapse(std::size_t from, std::size_t to, bool givenWeight, float weight);
  void upload();     // (re)builds the device network from the host objects
  void pullState();  // device potentials and firings -> potAct
  void pullWeights();
  void pullMirror();  // last fires and the synapse mirror

  float currentTime = 0.0f;
  uint64_t stepsDone = 0;  // grid steps simulated (time stepsDone * stepMs)
  unsigned totalGenNeurons = 0;
  bool weightsDirty = false;  // a mirror Synapse::setWeight since the last run()
  std::vector<float> activityStart;
  std::vector<uint64_t> exportIndex;     // out-synapse (flattened, by source) -> engine synapse order
  std::vector<uint8_t> delaySteps;       // by engine synapse order
  std::vector<std::size_t> uploadedRow;  // out-synapses per source when the engine network was built
  bool dirty = true;
  float setLr = NAN, setPre = NAN, setPost = NAN;
  nc_net* net = nullptr;
  nc_sim* sim = nullptr;
  int32_t netId = -1;
};
