// class NeuCor over the MI355X engine (see NeuCor.h).
//
// Construction (NeuCor(int), createNeuron, createSynapse, makeConnections)
// draws from the C library's rand() in the same order and with the same
// float/double expression shapes as the reference (src/NeuCor.cpp:11-42,
// :154-195, :418-444, :463-486), so NeuCor(n) after srand(seed) builds the
// reference's network: the same positions, synapses (in creation order),
// weights and lengths.  run() turns the reference's per-call work (input
// schedules, background draws, src/NeuCor.cpp:598-607) into forced fires on
// the engine's step grid and advances the engine by the whole steps covered.
#include "NeuCor.h"

#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <cassert>
#include <cmath>
#include <map>
#include <stdexcept>
#include <string>
#include <utility>

#include "../../include/neucor_amd.h"

namespace {

inline float randomUnit() { return static_cast<float>(rand()) / static_cast<float>(RAND_MAX); }

void check(nc_status st, const char* what) {
  if (st != NC_OK) {
    const char* msg = nc_last_error();
    throw std::runtime_error(std::string("NeuCor: ") + what + ": " + (msg ? msg : ""));
  }
}

// The HIP runtime draws from the C library's rand() while it sets a device
// up, and the reference's own draws (construction, input positions,
// background fires) and its callers' share that process-wide stream.  Every
// engine call runs with a scratch rand() state installed, so the caller's
// stream advances exactly as it would with the reference (setstate keeps the
// position of a swapped-out state in the state array itself).
class CallerRandStream {
 public:
  CallerRandStream() { prev_ = initstate(1u, scratch_, sizeof scratch_); }
  ~CallerRandStream() { setstate(prev_); }
  CallerRandStream(const CallerRandStream&) = delete;
  CallerRandStream& operator=(const CallerRandStream&) = delete;

 private:
  static char scratch_[128];
  char* prev_;
};
char CallerRandStream::scratch_[128];

// Synapse::getPrePot / getPostPot's display curve (src/NeuCor.cpp:528-532).
float renderPulse(float val) {
  val = fmin(fmax(val, 0.0), 0.7);
  if (val < 0.5) val = 8.0 * 1000.0 * powf(val / 5.0, 3.0);
  else val = 8.0 * powf(3.5 - 5 * val, 2.0);
  return val;
}

}  // namespace

// ---------------------------------------------------------------------------
// Inputs and detectors (src/NeuCor.cpp:305-366)

InputFirer::InputFirer(NeuCor* p, coord3 position, float r) : radius(r), lastFire(0.0f) {
  if (position.x == position.x) a = position;
  else a = {(randomUnit() - 0.5f) * 5.f, (randomUnit() - 0.5f) * 5.f, (randomUnit() - 0.5f) * 5.f};
  enabled = true;
  for (std::size_t i = 0; i < p->positions.size(); ++i)
    if (p->positions[i].getDist(a) < radius) near.push_back(i);
}

VoltageDetector::VoltageDetector(NeuCor* p, coord3 position, float r) : parentNet(p), radius(r) {
  if (position.x == position.x) a = position;
  else a = {(randomUnit() - 0.5f) * 3.f, (randomUnit() - 0.5f) * 3.f, (randomUnit() - 0.5f) * 3.f};
  for (std::size_t i = 0; i < p->positions.size(); ++i)
    if (p->positions[i].getDist(a) < radius) near.push_back(i);
}

float VoltageDetector::getVoltage() {
  float avgV = 0;
  for (std::size_t id : near) avgV += parentNet->potAct.at(id * 2);
  return avgV / near.size();
}

// ---------------------------------------------------------------------------
// The renderer's view: Neuron and Synapse mirrors (src/NeuCor.cpp:454-579)

Synapse::Synapse(NeuCor* p, std::size_t parent, std::size_t target, float w, float len)
    : pN(parent),
      tN(target),
      lastSpikeStart(0.0f),
      lastSpikeArrival(-INFINITY),
      AP_fireTime(0.0f),
      inhibitory(w < 0.0f),
      parentNet(p),
      length(len),
      weight(w) {}

void Synapse::setWeight(float w) {
  weight = w;
  inhibitory = weight < 0.0f;
  written = true;
  parentNet->weightsDirty = true;
}

float Synapse::getPrePot() const {
  if (AP_fireTime != 0.0f) {
    return renderPulse(float(parentNet->getTime() - lastSpikeStart) / (length * AP_speed)) * weight;
  }
  return 0.0f;
}

float Synapse::getPostPot() const {
  if (AP_fireTime != 0.0f && parentNet->getTime() < AP_fireTime) {
    return renderPulse(float(AP_fireTime - parentNet->getTime()) / (length * AP_speed)) * weight;
  }
  return 0.0f;
}

Neuron::Neuron(NeuCor* p, std::size_t id) : lastFire(NAN), pos(id), PA(2 * id), parentNet(p), ownID(id) {}

coord3 Neuron::position() const { return parentNet->positions.at(pos); }
float Neuron::potential() const { return parentNet->potAct.at(PA); }
float Neuron::activity() const { return parentNet->potAct.at(PA + 1); }

Synapse* NeuCor::getSynapse(std::size_t fromID, std::size_t toID) {
  for (Synapse& s : getNeuron(fromID)->outSynapses)
    if (s.tN == toID) return &s;
  return nullptr;
}

const Synapse* NeuCor::getSynapse(std::size_t fromID, std::size_t toID) const {
  for (const Synapse& s : getNeuron(fromID)->outSynapses)
    if (s.tN == toID) return &s;
  return nullptr;
}

// ---------------------------------------------------------------------------
// Construction

NeuCor::NeuCor(int n_neurons)
    : runSpeed(1.0f),
      runAll(false),
      learningRate(1.0f),
      presynapticTraceDecay(0.75f),
      postsynapticTraceDecay(0.65f),
      presynapticFactor(0.13f),
      postsynapticFactor(0.30f),
      inputArray(nullptr),
      inputArraySize(0) {
  totalGenNeurons = n_neurons;
  for (int n = 0; n < n_neurons; ++n) {
    coord3 d;
    d.setNAN();
    createNeuron(d);
  }
  totalGenNeurons = 0;
  for (int n = 0; n < n_neurons; ++n) connectNeuron((std::size_t)n);
}

NeuCor::~NeuCor() {
  CallerRandStream keep;
  if (sim) nc_sim_destroy(sim);
  if (net) nc_net_free(net);
}

void NeuCor::createNeuron(coord3 position) {
  // random position in a sphere holding 8 neurons per unit volume
  float spawnSize = 2.0;
  if (position.x != position.x) {
    if (totalGenNeurons != 0) spawnSize = powf(totalGenNeurons / (1.3333 * 3.1459 * 8), 0.33333) * 2.0;
    do {
      position.x = (randomUnit() - 0.5f) * spawnSize;
      position.y = (randomUnit() - 0.5f) * spawnSize;
      position.z = (randomUnit() - 0.5f) * spawnSize;
    } while (pow(position.x, 2) + pow(position.y, 2) + pow(position.z, 2) > pow(spawnSize / 2.0, 2.0));
  }
  positions.push_back(position);
  potAct.push_back(-70.0f);  // Neuron's baselevel (src/NeuCor.cpp:379, :394)
  potAct.push_back(0.0f);
  activityStart.push_back(currentTime);
  neurons.emplace_back(this, positions.size() - 1);
  dirty = true;
}

void NeuCor::addSynapse(std::size_t from, std::size_t to, bool givenWeight, float w) {
  // Synapse::Synapse draws its weight and sign even when createSynapse then
  // overwrites it (src/NeuCor.cpp:471-473, :193-194)
  float weight = randomUnit() * 0.8f + 0.2f;
  if (randomUnit() < 0.2f) weight = -weight;
  if (givenWeight) weight = w;
  const float length = positions.at(from).getDist(positions.at(to));
  neurons.at(from).outSynapses.emplace_back(this, from, to, weight, length);
  neurons.at(to).inSynapses.emplace(from, to);
  dirty = true;
}

void NeuCor::connectNeuron(std::size_t id) {
  const coord3 nPos = positions.at(id);
  for (std::size_t i = 0; i < positions.size(); ++i) {
    const float distance = nPos.getDist(positions[i]);
    if (distance < 1.0 && i != id) {
      bool allowed = true;
      for (const Synapse& s : neurons[id].outSynapses)
        if (s.tN == i) {
          allowed = false;
          break;
        }
      if (allowed) addSynapse(id, i, false, 0.0f);
    }
  }
}

void NeuCor::createSynapse(std::size_t toID, std::size_t fromID, float weight) {
  for (const Synapse& s : neurons.at(fromID).outSynapses)
    if (s.tN == toID) return;
  (void)positions.at(toID);
  addSynapse(fromID, toID, true, weight);
}

void NeuCor::makeConnections() {
  for (std::size_t n = 0; n < positions.size(); ++n) connectNeuron(n);
}

std::size_t NeuCor::getNeuronCount() const { return positions.size(); }

void NeuCor::setInputRateArray(float inputs[], unsigned inputCount, coord3 inputPositions[], float inputRadius[]) {
  inputArray = inputs;
  inputArraySize = inputCount;
  const int change = (int)inputArraySize - (int)inputHandler.size();
  for (int i = 0; i < change; ++i) {
    if (inputPositions != nullptr) inputHandler.emplace_back(this, inputPositions[i], inputRadius[i]);
    else inputHandler.emplace_back(this);
  }
  for (int i = 0; i < -change; ++i) inputHandler.pop_back();
}

void NeuCor::addInputOffset(unsigned inputID, float t) { inputHandler.at(inputID).lastFire += t; }

void NeuCor::setDetectors(unsigned detectorNumber, coord3 detectorPositions[], float detectorRadius[]) {
  for (unsigned i = 0; i < detectorNumber; ++i) {
    if (detectorPositions != nullptr) voltageDetectors.emplace_back(this, detectorPositions[i], detectorRadius[i]);
    else voltageDetectors.emplace_back(this);
  }
}

float NeuCor::getDetectorVoltage(unsigned ID) { return voltageDetectors.at(ID).getVoltage(); }

std::vector<float> NeuCor::getDetectorVoltages() {
  std::vector<float> v;
  for (auto& d : voltageDetectors) v.push_back(d.getVoltage());
  return v;
}

float NeuCor::getTime() const { return currentTime; }

// ---------------------------------------------------------------------------
// Device network

void NeuCor::pullWeights() {
  if (!sim) return;
  CallerRandStream keep;
  std::vector<float> w(delaySteps.size());
  if (!w.empty()) check(nc_read_weights(sim, netId, w.data(), w.size()), "nc_read_weights");
  // synapses created since the engine network was built keep their weights
  std::size_t f = 0;
  for (std::size_t i = 0; i < uploadedRow.size(); ++i)
    for (std::size_t k = 0; k < uploadedRow[i]; ++k) neurons[i].outSynapses[k].weight = w[exportIndex[f++]];
}

void NeuCor::upload() {
  pullWeights();  // a rebuild keeps the learned weights
  CallerRandStream keep;
  // A network changed after it has run (createNeuron / createSynapse, as the
  // reference allows at any time) keeps the dynamic state of its existing
  // neurons and synapses, as the reference's objects do: potentials, last
  // fires, firing counts, and every synapse's latched spike and last arrival.
  struct OldSyn {
    float depol;
    int32_t arrival, last_arr;
  };
  const std::size_t n_old = uploadedRow.size();
  std::vector<float> oldV;
  std::vector<int32_t> oldLf;
  std::vector<uint32_t> oldFi;
  std::map<std::pair<std::size_t, std::size_t>, OldSyn> oldSyn;
  const bool carry = sim && stepsDone > 0 && n_old > 0;
  const uint32_t oldSteps = (uint32_t)stepsDone;
  if (carry) {
    oldV.resize(n_old), oldLf.resize(n_old), oldFi.resize(n_old);
    check(nc_read_potential(sim, netId, oldV.data(), (uint32_t)n_old), "nc_read_potential");
    check(nc_read_last_fire(sim, netId, oldLf.data(), (uint32_t)n_old), "nc_read_last_fire");
    check(nc_read_firings(sim, netId, oldFi.data(), (uint32_t)n_old), "nc_read_firings");
    const std::size_t S_old = delaySteps.size();
    std::vector<int32_t> arr(S_old), la(S_old);
    std::vector<float> dp(S_old);
    if (S_old)
      check(nc_read_synapse_state(sim, netId, arr.data(), la.data(), dp.data(), S_old), "nc_read_synapse_state");
    std::size_t f = 0;
    for (std::size_t i = 0; i < n_old; ++i)
      for (std::size_t k = 0; k < uploadedRow[i]; ++k) {
        const uint64_t e = exportIndex[f++];
        // nc_read_synapse_state reports a released latch as 0; the checkpoint
        // takes NC_NEVER_STEP for "no spike" (a released spike is no longer
        // in flight or in its window, so the two behave alike)
        oldSyn[{i, neurons[i].outSynapses[k].tN}] = OldSyn{dp[e], arr[e] ? arr[e] : NC_NEVER_STEP, la[e]};
      }
  }
  if (sim) {
    nc_sim_destroy(sim);
    sim = nullptr;
  }
  if (net) {
    nc_net_free(net);
    net = nullptr;
  }
  dirty = false;
  const uint32_t n = (uint32_t)positions.size();
  if (n == 0) return;
  float maxLen = 0.0f;
  std::size_t S = 0;
  for (auto& nn : neurons) {
    S += nn.outSynapses.size();
    for (const Synapse& s : nn.outSynapses) maxLen = std::max(maxLen, s.length);
  }
  if (stepMs <= 0.0f) {
    // the largest power of two not above runSpeed, at most 1 ms, and coarse
    // enough that every delay fits the engine's 8-bit step count
    float dt = 1.0f;
    while (dt > runSpeed && dt > 1.0f / 1024) dt *= 0.5f;
    while (2.0f * maxLen / dt > 250.0f && dt < 1.0f) dt *= 2.0f;
    stepMs = dt;
  }
  std::vector<uint32_t> to, from;
  std::vector<float> w;
  std::vector<uint8_t> d, inh;
  to.reserve(S), from.reserve(S), w.reserve(S), d.reserve(S), inh.reserve(S);
  for (uint32_t i = 0; i < n; ++i)
    for (const Synapse& s : neurons[i].outSynapses) {
      to.push_back((uint32_t)s.tN);
      from.push_back(i);
      w.push_back(s.weight);
      // delay = 2 * length ms (AP_speed, src/NeuCor.cpp:485, :733), in whole steps
      const long k = lround(2.0 * s.length / stepMs);
      if (k > 250) throw std::runtime_error("NeuCor: a synaptic delay exceeds 250 steps; raise stepMs");
      d.push_back((uint8_t)std::max(1L, k));
      inh.push_back(s.weight < 0.0f ? 1 : 0);  // Synapse::setWeight (src/NeuCor.cpp:576-579)
    }
  std::vector<float> xyz(3 * (std::size_t)n);
  for (uint32_t i = 0; i < n; ++i) {
    xyz[3 * i] = positions[i].x;
    xyz[3 * i + 1] = positions[i].y;
    xyz[3 * i + 2] = positions[i].z;
  }
  check(nc_net_from_edges(n, xyz.data(), S, to.data(), from.data(), w.data(), d.data(), inh.data(), &net),
        "nc_net_from_edges");
  uint64_t npos = 0, nslots = 0;
  check(nc_net_device_size(net, &npos, &nslots), "nc_net_device_size");
  nc_sim_spec spec{};
  spec.device = device;
  spec.max_neurons = npos;
  spec.max_synapses = std::max<uint64_t>(S, 1);
  spec.max_nets = 1;
  spec.raster_capacity = 0;
  spec.max_slots = nslots;
  check(nc_sim_create(&spec, &sim), "nc_sim_create");
  nc_params p{stepMs, learningRate, presynapticFactor, postsynapticFactor, presynapticTraceDecay,
              postsynapticTraceDecay};
  check(nc_sim_add(sim, net, &p, nullptr, 0, &netId), "nc_sim_add");
  setLr = learningRate, setPre = presynapticFactor, setPost = postsynapticFactor;
  // engine synapse order: rows by target, ascending source
  std::vector<int64_t> in_ptr(n + 1);
  std::vector<uint32_t> in_src(S);
  check(nc_net_export(net, nullptr, in_ptr.data(), in_src.data(), nullptr, nullptr, nullptr), "nc_net_export");
  exportIndex.assign(S, 0);
  delaySteps.assign(S, 0);
  std::size_t f = 0;
  for (uint32_t i = 0; i < n; ++i)
    for (const Synapse& s : neurons[i].outSynapses) {
      const uint32_t* b = in_src.data() + in_ptr[s.tN];
      const uint32_t* e = in_src.data() + in_ptr[s.tN + 1];
      const uint64_t k = (uint64_t)(std::lower_bound(b, e, i) - in_src.data());
      exportIndex[f] = k;
      delaySteps[k] = d[f];
      ++f;
    }
  // a rebuilt engine starts at step 0 of the current time
  stepsDone = (uint64_t)std::floor((double)currentTime / stepMs);
  uploadedRow.resize(n);
  for (uint32_t i = 0; i < n; ++i) uploadedRow[i] = neurons[i].outSynapses.size();
  if (!carry) return;
  // The carried state, as a checkpoint of the new network (nc_sim_save of the
  // fresh network gives its header and sizes; the neuron and synapse records
  // are then overwritten): header of NC_CKPT_HEADER_BYTES, then per neuron
  // {V, last fire, firings}, per synapse in engine order {w, depol, last
  // arrival, arrival}, 4 bytes each.
  uint64_t size = 0;
  check(nc_sim_save(sim, netId, nullptr, 0, &size), "nc_sim_save");
  std::vector<unsigned char> buf(size);
  check(nc_sim_save(sim, netId, buf.data(), size, &size), "nc_sim_save");
  memcpy(buf.data() + NC_CKPT_STEPS_OFFSET, &oldSteps, 4);
  unsigned char* nr = buf.data() + NC_CKPT_HEADER_BYTES;
  for (std::size_t i = 0; i < n_old && i < n; ++i) {
    memcpy(nr + 12 * i, &oldV[i], 4);
    memcpy(nr + 12 * i + 4, &oldLf[i], 4);
    memcpy(nr + 12 * i + 8, &oldFi[i], 4);
  }
  unsigned char* sr = nr + 12ull * n;
  std::size_t f2 = 0;
  for (uint32_t i = 0; i < n; ++i)
    for (const Synapse& s : neurons[i].outSynapses) {
      const uint64_t k = exportIndex[f2++];
      auto it = oldSyn.find({i, s.tN});
      if (it == oldSyn.end()) continue;
      memcpy(sr + 16 * k + 4, &it->second.depol, 4);
      memcpy(sr + 16 * k + 8, &it->second.last_arr, 4);
      memcpy(sr + 16 * k + 12, &it->second.arrival, 4);
    }
  check(nc_sim_restore(sim, netId, buf.data(), size), "nc_sim_restore");
  stepsDone = oldSteps;
}

void NeuCor::pullMirror() {
  const std::size_t n = positions.size(), S = delaySteps.size();
  if (!sim || n == 0) return;
  CallerRandStream keep;
  std::vector<int32_t> lf(n), arrival(S), lastArr(S);
  std::vector<float> w(S), depol(S);
  check(nc_read_last_fire(sim, netId, lf.data(), (uint32_t)n), "nc_read_last_fire");
  if (S) {
    check(nc_read_weights(sim, netId, w.data(), S), "nc_read_weights");
    check(nc_read_synapse_state(sim, netId, arrival.data(), lastArr.data(), depol.data(), S),
          "nc_read_synapse_state");
  }
  for (std::size_t i = 0; i < n; ++i)
    neurons[i].lastFire = lf[i] == NC_NEVER_STEP ? NAN : (float)((double)lf[i] * stepMs);
  std::size_t f = 0;
  for (std::size_t i = 0; i < n; ++i)
    for (Synapse& s : neurons[i].outSynapses) {
      const uint64_t k = exportIndex[f++];
      s.weight = w[k];
      s.inhibitory = w[k] < 0.0f;
      s.lastSpikeArrival = lastArr[k] == NC_NEVER_STEP ? -INFINITY : (float)((double)lastArr[k] * stepMs);
      if (arrival[k] != 0) {
        s.AP_fireTime = (float)((double)arrival[k] * stepMs);
        s.lastSpikeStart = (float)((double)(arrival[k] - (int32_t)delaySteps[k]) * stepMs);
      } else {
        s.AP_fireTime = 0.0f;
      }
    }
}

void NeuCor::pullState() {
  const std::size_t n = positions.size();
  if (!sim || n == 0) return;
  CallerRandStream keep;
  std::vector<float> V(n);
  std::vector<uint32_t> fi(n);
  check(nc_read_potential(sim, netId, V.data(), (uint32_t)n), "nc_read_potential");
  check(nc_read_firings(sim, netId, fi.data(), (uint32_t)n), "nc_read_firings");
  for (std::size_t i = 0; i < n; ++i) {
    potAct[2 * i] = V[i];
    // Neuron::run's activity (src/NeuCor.cpp:640)
    potAct[2 * i + 1] = fi[i] / (((float)currentTime - activityStart[i]) / 10.0);
  }
}

void NeuCor::resetActivities() {
  if (sim) {
    CallerRandStream keep;
    check(nc_reset_activities(sim, netId), "nc_reset_activities");
  }
  for (std::size_t i = 0; i < positions.size(); ++i) {
    activityStart[i] = currentTime;
    potAct[2 * i + 1] = 0.0f;
  }
}

// ---------------------------------------------------------------------------
// NeuCor::run (src/NeuCor.cpp:583-617)

void NeuCor::run() {
  assert(0 <= runSpeed);
  if (runSpeed <= 0.0f) return;
  if (dirty) upload();
  const std::size_t n = positions.size();

  std::vector<uint32_t> evOff, evId;
  std::vector<uint8_t> evKind;
  // the first grid step at or after t (ms), as an offset from the next step
  auto stepAt = [&](double t) -> uint32_t {
    const double k = std::ceil(t / (double)stepMs);
    const double off = k - (double)stepsDone;
    return off < 1.0 ? 1u : (uint32_t)off;
  };

  // InputFirer::schedule (src/NeuCor.cpp:333-345) for every input
  for (std::size_t i = 0; i < inputHandler.size(); ++i) {
    const float frequency = inputArray != nullptr && i < inputArraySize ? inputArray[i] : 0.0f;
    InputFirer& in = inputHandler[i];
    if (frequency == 0 || !in.enabled) continue;
    if (in.lastFire != in.lastFire) in.lastFire = 0;
    const float currentT = currentTime;
    for (float fireTime = in.lastFire + 1000.0 / frequency; fireTime < currentT + runSpeed;
         fireTime += 1000.0 / frequency) {
      if (currentT < fireTime) {
        for (std::size_t id : in.near) {
          evOff.push_back(stepAt(fireTime));
          evId.push_back((uint32_t)id);
          evKind.push_back(0);
        }
        in.lastFire = fireTime;
      }
    }
  }
  // background fires (src/NeuCor.cpp:604-607)
  for (std::size_t i = 0; i < n; ++i) {
    const int period = std::max(1, static_cast<int>(600.0f / runSpeed));
    if (rand() % period == 0) {
      const std::size_t id = rand() % n;
      const float off = randomUnit() * runSpeed;
      evOff.push_back(stepAt((double)currentTime + off));
      evId.push_back((uint32_t)id);
      evKind.push_back(1);
    }
  }

  // The engine follows the reference's own float clock: after this call it
  // has run every grid step at or before currentTime + runSpeed.
  const float targetTime = currentTime + runSpeed;
  const uint64_t target = (uint64_t)std::floor((double)targetTime / stepMs);
  const uint64_t steps = target > stepsDone ? target - stepsDone : 0;
  if (sim) {
    CallerRandStream keep;
    if (weightsDirty) {
      // Synapse::setWeight on the mirror since the last run(), for exactly
      // the synapses written (the others keep their sign class)
      std::vector<uint64_t> idx;
      std::vector<float> wv;
      std::size_t f = 0;
      for (std::size_t i = 0; i < uploadedRow.size(); ++i)
        for (std::size_t k = 0; k < uploadedRow[i]; ++k, ++f) {
          Synapse& sy = neurons[i].outSynapses[k];
          if (!sy.written) continue;
          idx.push_back(exportIndex[f]);
          wv.push_back(sy.weight);
          sy.written = false;
        }
      if (!idx.empty()) check(nc_set_weights(sim, netId, idx.size(), idx.data(), wv.data()), "nc_set_weights");
      weightsDirty = false;
    }
    if (learningRate != setLr || presynapticFactor != setPre || postsynapticFactor != setPost) {
      check(nc_set_params(sim, netId, learningRate, presynapticFactor, postsynapticFactor), "nc_set_params");
      setLr = learningRate, setPre = presynapticFactor, setPost = postsynapticFactor;
    }
    if (!evId.empty())
      check(nc_push_forced(sim, netId, (uint32_t)evId.size(), evOff.data(), evId.data(), evKind.data()),
            "nc_push_forced");
    if (steps) check(nc_sim_step(sim, (uint32_t)steps), "nc_sim_step");
  }
  stepsDone += steps;
  currentTime = targetTime;
  pullState();
  if (mirror) pullMirror();
}

// ---------------------------------------------------------------------------
// Snapshots (src/NeuCor.cpp:98-152)

std::vector<NeuCor::NeuronSnapshot> NeuCor::getNeuronSnapshots() const {
  std::vector<NeuronSnapshot> v;
  v.reserve(positions.size());
  for (std::size_t i = 0; i < positions.size(); ++i) v.push_back({i, positions[i], potAct[2 * i], potAct[2 * i + 1]});
  return v;
}

std::vector<NeuCor::SynapseSnapshot> NeuCor::getSynapseSnapshots() const {
  const std::size_t S = delaySteps.size();
  std::vector<float> w(S), depol(S);
  std::vector<int32_t> arrival(S), lastArr(S);
  const bool live = sim && !dirty && S;
  if (live) {
    CallerRandStream keep;
    check(nc_read_weights(sim, netId, w.data(), S), "nc_read_weights");
    check(nc_read_synapse_state(sim, netId, arrival.data(), lastArr.data(), depol.data(), S),
          "nc_read_synapse_state");
  }
  std::vector<SynapseSnapshot> v;
  std::size_t f = 0;
  for (std::size_t i = 0; i < neurons.size(); ++i)
    for (const Synapse& s : neurons[i].outSynapses) {
      float weight = s.weight, pre = 0.0f, post = 0.0f;
      if (live) {
        const uint64_t k = exportIndex[f];
        weight = w[k];
        if (arrival[k] != 0) {
          // AP_fireTime = arrival, lastSpikeStart = arrival - delay, AP_speed = 2
          const float fireT = (float)((double)arrival[k] * stepMs);
          const float startT = (float)((double)(arrival[k] - (int32_t)delaySteps[k]) * stepMs);
          pre = renderPulse(float(currentTime - startT) / (s.length * 2.0f)) * weight;
          if (currentTime < fireT) post = renderPulse(float(fireT - currentTime) / (s.length * 2.0f)) * weight;
        }
      }
      v.push_back({i, s.tN, positions[i], positions[s.tN], weight, pre, post, weight < 0.0f});
      ++f;
    }
  return v;
}

std::vector<NeuCor::InputSnapshot> NeuCor::getInputSnapshots() const {
  std::vector<InputSnapshot> v;
  v.reserve(inputHandler.size());
  for (std::size_t i = 0; i < inputHandler.size(); ++i) {
    const InputFirer& in = inputHandler[i];
    v.push_back({i, in.a, in.radius, inputArray != nullptr && i < inputArraySize ? inputArray[i] : 0.0f, in.enabled});
  }
  return v;
}
