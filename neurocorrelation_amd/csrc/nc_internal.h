// Host-side internals shared by the network builder and the engine.
#pragma once
#include <stdint.h>
#include <exception>
#include <memory>
#include <new>
#include <string>
#include <vector>

#include "../../include/neucor_amd.h"

// A network on the host.  Incoming synapses of neuron i are
// [in_ptr[i], in_ptr[i+1]) with strictly ascending source id: the iteration
// order of Neuron::inSynapses (std::map keyed by source, src/NeuCor.h:212),
// which fixes the float summation order of charge_insynapses
// (src/NeuCor.cpp:688-700).
namespace nc {
struct NetImage;  // the engine's device image of a network (nc_net_prepare)
}

struct nc_net {
  uint32_t n = 0;
  std::vector<float> xyz;          // 3 per neuron
  std::vector<int64_t> in_ptr;     // n + 1
  std::vector<uint32_t> in_src;    // S
  std::vector<float> w;            // S
  std::vector<uint8_t> delay;      // S, steps >= 1
  std::vector<uint8_t> inh;        // S, 1 = inhibitory sign class
  uint32_t max_in_degree = 0;
  uint32_t max_delay = 0;
  // Cached device image, built by nc_net_prepare; a network never changes
  // after construction, so the image stays valid for its lifetime.
  std::shared_ptr<const nc::NetImage> image;
  uint64_t n_syn() const { return in_src.size(); }
};

namespace nc {
void set_error(const std::string& msg);
nc_status fail(nc_status code, const std::string& msg);

// Handlers of the function-try-blocks around the C entry points: no C++
// exception crosses the C ABI; a host allocation failure is NC_ERR_OOM.
#define NC_ABI_CATCH                                                              \
  catch (const std::bad_alloc&) {                                                 \
    return nc::fail(NC_ERR_OOM, "host allocation failed");                        \
  }                                                                               \
  catch (const std::exception& e) {                                               \
    return nc::fail(NC_ERR_STATE, std::string("unexpected host error: ") + e.what()); \
  }
void finish_net_stats(nc_net* net);
nc_status check_wiring_spec(const nc_wiring_spec& s);

// Input selection shared by the engine and nc_net_select_inputs.
struct InputSel {
  std::vector<uint32_t> ids;     // by rank
  std::vector<uint32_t> group;   // group index per input
  uint32_t n_groups_total = 0;   // correlated groups + independent singletons
};
nc_status select_inputs(const nc_net& net, const nc_drive_spec& d, InputSel* out);
}  // namespace nc
