// The served form's request loop (nc_server_*), over the engine's C ABI.
//
// Jobs queue as they are submitted; builder threads turn each job's wiring
// spec into a network with its device image (on the GPU when asked), and one
// loop thread owns the simulations ("lanes") and the GPU.  Each lane is an
// nc_sim holding many jobs' networks: one nc_sim_step call advances all of
// them, and a job stops at its own step count.  With two or more lanes the
// GPU runs one lane's chunk while the loop retires and admits the next's.
//
// Eviction: a job that has waited evict_after_s for room takes it from the
// running job with the most steps left (if that is more than the waiting
// job's own): the running job is checkpointed (nc_sim_save), removed, and
// queued again with its checkpoint and its built network, to resume with
// nc_sim_add + nc_sim_restore.  Results do not change: a checkpoint holds
// the whole state, and the drive depends only on the step number.
//
// Overflow lanes: a job that finds no room anywhere after waiting (and no
// victim) makes the loop read the device's free HBM; if a lane sized for it
// fits, one is created, and destroyed again once empty.
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <list>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "nc_internal.h"

namespace {

using Clock = std::chrono::steady_clock;

struct JobRec {
  nc_job job{};
  nc_job_result res{};
  nc_net* net = nullptr;
  uint32_t n = 0;
  uint64_t S = 0;
  uint32_t left = 0;          // steps still to run
  double t_ready = 0;         // joined the ready queue (built or evicted)
  std::vector<unsigned char> ckpt;  // evicted: its state
  int32_t nid = -1;
  bool no_room_empty = false;  // refused by an empty regular lane
  std::list<std::unique_ptr<JobRec>>::iterator self;  // its place in nc_server::jobs
  ~JobRec() {
    if (net) nc_net_free(net);
  }
};

struct Lane {
  nc_sim* sim = nullptr;
  bool overflow = false;
  std::map<int32_t, JobRec*> running;
};

}  // namespace

struct nc_server {
  nc_server_spec spec{};
  Clock::time_point t0;
  std::mutex mu;
  std::condition_variable cv_build;   // builders: a job to build, or stop
  std::condition_variable cv_loop;    // loop: a job built, or stop
  std::condition_variable cv_result;  // pollers: a result
  std::list<std::unique_ptr<JobRec>> jobs;  // every job not yet finished
  std::deque<JobRec*> to_build;
  std::deque<JobRec*> ready;          // built (or evicted), waiting for room, in that order
  std::vector<nc_job_result> results;
  std::vector<std::thread> builders;
  std::thread loop;
  bool stop = false;
  nc_status loop_status = NC_OK;
  std::string loop_error;
  nc_server_stats st{};
  double running_sum = 0;
  uint32_t n_lanes = 0;
  // owned by the loop thread (a lane keeps its address while others come and go)
  std::vector<std::unique_ptr<Lane>> lanes;

  double now() const { return std::chrono::duration<double>(Clock::now() - t0).count(); }
};

namespace {

// Publishes a job's result and drops the job (caller holds s->mu).
void complete_locked(nc_server* s, JobRec* j) {
  s->results.push_back(j->res);
  ++s->st.finished;
  s->cv_result.notify_all();
  s->jobs.erase(j->self);
}

nc_status lane_create(nc_server* s, uint64_t max_n, uint64_t max_s, bool overflow, Lane* out) {
  nc_sim_spec ss{};
  ss.device = s->spec.device;
  ss.max_neurons = max_n;
  ss.max_synapses = max_s;
  ss.max_nets = s->spec.max_nets;
  // layout padding: a quarter more slots, and up to a chunk of 8 per lane of every slice
  ss.max_slots = max_s + max_s / 4 + (max_n + 64ull * s->spec.max_nets) * 4;
  nc_status r = nc_sim_create(&ss, &out->sim);
  out->overflow = overflow;
  return r;
}

void build_job(nc_server* s, JobRec* j) {
  nc_wiring_spec w = j->job.wiring;
  w.n_threads = 1;
  nc_net* net = nullptr;
  nc_status r = s->spec.device_wiring ? nc_net_generate_device(&w, s->spec.device, &net) : nc_net_generate(&w, &net);
  if (r == NC_OK) r = nc_net_prepare(net);
  uint32_t n = 0;
  uint64_t S = 0;
  if (r == NC_OK) r = nc_net_info(net, &n, &S, nullptr, nullptr);
  std::lock_guard<std::mutex> g(s->mu);
  j->res.t_built = s->now();
  ++s->st.built;
  if (r != NC_OK) {
    if (net) nc_net_free(net);
    j->res.status = r;
    j->res.t_done = j->res.t_built;
    complete_locked(s, j);
    return;
  }
  j->net = net;
  j->n = n;
  j->S = S;
  j->res.n_neurons = n;
  j->res.n_synapses = S;
  j->t_ready = j->res.t_built;
  s->ready.push_back(j);
  s->cv_loop.notify_one();
}

void builder_main(nc_server* s) {
  for (;;) {
    JobRec* j = nullptr;
    {
      std::unique_lock<std::mutex> g(s->mu);
      s->cv_build.wait(g, [&] { return s->stop || !s->to_build.empty(); });
      if (s->stop) return;
      j = s->to_build.front();
      s->to_build.pop_front();
    }
    build_job(s, j);
  }
}

// The finished job's result: potentials, weights (export order) and firing
// counts hashed in that order.
nc_status finish_job(nc_server* s, Lane& L, JobRec* j) {
  std::vector<float> V(j->n);
  std::vector<float> w(j->S);
  std::vector<uint32_t> f(j->n);
  nc_status r = nc_read_potential(L.sim, j->nid, V.data(), j->n);
  if (r == NC_OK) r = nc_read_weights(L.sim, j->nid, w.data(), j->S);
  if (r == NC_OK) r = nc_read_firings(L.sim, j->nid, f.data(), j->n);
  if (r == NC_OK) r = nc_sim_remove(L.sim, j->nid);
  if (r != NC_OK) return r;
  uint64_t h = 14695981039346656037ull;
  h = nc_digest(V.data(), 4ull * j->n, h);
  h = nc_digest(w.data(), 4ull * j->S, h);
  h = nc_digest(f.data(), 4ull * j->n, h);
  uint64_t spikes = 0;
  for (uint32_t x : f) spikes += x;
  std::lock_guard<std::mutex> g(s->mu);
  j->res.status = NC_OK;
  j->res.spikes = spikes;
  j->res.digest = h;
  j->res.t_done = s->now();
  --s->st.running;
  complete_locked(s, j);
  return NC_OK;
}

// Admits j into lane L: NC_OK, NC_ERR_CAPACITY/NC_ERR_OOM (no room), or an error.
nc_status admit(nc_server* s, Lane& L, JobRec* j) {
  int32_t nid = -1;
  nc_status r = nc_sim_add(L.sim, j->net, &j->job.params, &j->job.drive, j->left, &nid);
  if (r != NC_OK) return r;
  if (!j->ckpt.empty()) {
    r = nc_sim_restore(L.sim, nid, j->ckpt.data(), j->ckpt.size());
    if (r != NC_OK) {
      (void)nc_sim_remove(L.sim, nid);
      return r;
    }
    j->ckpt.clear();
    j->ckpt.shrink_to_fit();
  }
  j->nid = nid;
  L.running[nid] = j;
  if (j->res.t_admit == 0) j->res.t_admit = s->now();
  return NC_OK;
}

// Checkpoints and removes lane L's job with the most steps left, if that is
// more than `than`; the job goes back to the ready queue.  Returns whether it
// evicted one.
bool evict_one(nc_server* s, Lane& L, uint32_t than, nc_status* err) {
  JobRec* v = nullptr;
  for (auto& kv : L.running) {
    uint32_t done = 0, left = 0;
    if ((*err = nc_sim_progress(L.sim, kv.first, &done, &left)) != NC_OK) return false;
    kv.second->left = left;
    if (left > than && (!v || left > v->left)) v = kv.second;
  }
  if (!v) return false;
  uint64_t need = 0;
  if ((*err = nc_sim_save(L.sim, v->nid, nullptr, 0, &need)) != NC_OK) return false;
  v->ckpt.resize(need);
  if ((*err = nc_sim_save(L.sim, v->nid, v->ckpt.data(), need, &need)) != NC_OK) return false;
  if ((*err = nc_sim_remove(L.sim, v->nid)) != NC_OK) return false;
  L.running.erase(v->nid);
  v->nid = -1;
  ++v->res.evictions;
  std::lock_guard<std::mutex> g(s->mu);
  ++s->st.evictions;
  --s->st.running;
  v->t_ready = s->now();
  s->ready.push_back(v);
  return true;
}

nc_status loop_body(nc_server* s) {
  const uint32_t chunk = s->spec.chunk;
  Lane* in_flight = nullptr;
  const double patience = s->spec.evict_after_s > 0 ? s->spec.evict_after_s : 1.0;
  for (;;) {
    {
      std::lock_guard<std::mutex> g(s->mu);
      if (s->stop) break;
    }
    bool served = false;
    for (size_t li = 0; li < s->lanes.size(); ++li) {
      Lane& L = *s->lanes[li];
      if (&L == in_flight) {
        if (s->lanes.size() > 1) continue;  // the other lanes are serviced while it runs
        double t = s->now();
        nc_status r = nc_sim_synchronize(L.sim);
        if (r != NC_OK) return r;
        in_flight = nullptr;
        s->st.gpu_wait_s += s->now() - t;
      }
      // retirement
      double t1 = s->now();
      for (auto it = L.running.begin(); it != L.running.end();) {
        uint32_t done = 0, left = 0;
        nc_status r = nc_sim_progress(L.sim, it->first, &done, &left);
        if (r != NC_OK) return r;
        JobRec* j = it->second;
        if (left != 0) {
          j->left = left;
          ++it;
          continue;
        }
        it = L.running.erase(it);
        r = finish_job(s, L, j);
        if (r != NC_OK) return r;
      }
      double t2 = s->now();
      s->st.retire_s += t2 - t1;
      // admission in ready order while the lane has room
      for (;;) {
        JobRec* j = nullptr;
        {
          std::lock_guard<std::mutex> g(s->mu);
          if (s->ready.empty()) break;
          j = s->ready.front();
        }
        nc_status r = admit(s, L, j);
        if (r == NC_OK) {
          std::lock_guard<std::mutex> g(s->mu);
          s->ready.pop_front();
          ++s->st.admitted;
          ++s->st.running;
          continue;
        }
        if (r != NC_ERR_CAPACITY && r != NC_ERR_OOM) return r;
        if (L.running.empty()) {
          // an empty lane has no room for it: no regular lane will; an
          // overflow lane sized for it may, if free HBM allows
          if (!L.overflow) j->no_room_empty = true;
          if (!s->spec.overflow_lanes) {
            std::lock_guard<std::mutex> g(s->mu);
            s->ready.pop_front();
            j->res.status = NC_ERR_CAPACITY;
            j->res.t_done = s->now();
            complete_locked(s, j);
            continue;
          }
          break;
        }
        if (s->spec.evict_after_s > 0 && s->now() - j->t_ready >= patience) {
          nc_status e = NC_OK;
          if (evict_one(s, L, j->left, &e)) continue;  // try again with the room it left
          if (e != NC_OK) return e;
        }
        break;
      }
      s->st.admit_s += s->now() - t2;
      // the GPU finishes the other lane's chunk, then takes this one
      double t3 = s->now();
      if (in_flight) {
        nc_status r = nc_sim_synchronize(in_flight->sim);
        if (r != NC_OK) return r;
        in_flight = nullptr;
      }
      s->st.gpu_wait_s += s->now() - t3;
      if (!L.running.empty()) {
        nc_status r = nc_sim_step(L.sim, chunk);
        if (r != NC_OK) return r;
        in_flight = &L;
        served = true;
        ++s->st.chunks;
        s->running_sum += (double)L.running.size();
      }
    }
    // overflow lanes: created for a job that waited without room, destroyed when empty
    if (s->spec.overflow_lanes) {
      for (size_t li = 0; li < s->lanes.size();) {
        if (s->lanes[li]->overflow && s->lanes[li]->running.empty() && s->lanes[li].get() != in_flight) {
          nc_sim_destroy(s->lanes[li]->sim);
          s->lanes.erase(s->lanes.begin() + li);
          std::lock_guard<std::mutex> g(s->mu);
          s->n_lanes = (uint32_t)s->lanes.size();
        } else {
          ++li;
        }
      }
      JobRec* j = nullptr;
      {
        std::lock_guard<std::mutex> g(s->mu);
        if (!s->ready.empty() && s->now() - s->ready.front()->t_ready >= patience) j = s->ready.front();
      }
      uint32_t n_over = 0;
      for (auto& L : s->lanes) n_over += L->overflow;
      if (j) {
        uint64_t free_b = 0, need = 0;
        const uint64_t on = std::max<uint64_t>(j->n, s->spec.max_neurons / 4);
        const uint64_t os = std::max<uint64_t>(j->S, s->spec.max_synapses / 4);
        nc_status r = nc_sim_free_bytes(s->lanes[0]->sim, &free_b);
        if (r == NC_OK) r = nc_sim_footprint(on, os, &need);
        if (r != NC_OK) return r;
        bool made = false;
        if (n_over < s->spec.overflow_lanes && need < free_b / 2) {
          std::unique_ptr<Lane> L(new Lane());
          made = lane_create(s, on, os, true, L.get()) == NC_OK;
          if (made) {
            s->lanes.push_back(std::move(L));
            std::lock_guard<std::mutex> g(s->mu);
            s->n_lanes = (uint32_t)s->lanes.size();
          }
        }
        if (!made && j->no_room_empty && n_over == 0) {
          // larger than a regular lane, and no HBM for a lane of its own
          std::lock_guard<std::mutex> g(s->mu);
          s->ready.pop_front();
          j->res.status = NC_ERR_CAPACITY;
          j->res.t_done = s->now();
          complete_locked(s, j);
        }
      }
    }
    if (!served && !in_flight) {
      double t = s->now();
      std::unique_lock<std::mutex> g(s->mu);
      s->cv_loop.wait_for(g, std::chrono::milliseconds(1));
      s->st.idle_s += s->now() - t;
    }
  }
  if (in_flight) return nc_sim_synchronize(in_flight->sim);
  return NC_OK;
}

void loop_main(nc_server* s) {
  nc_status r;
  try {
    r = loop_body(s);
  } catch (const std::bad_alloc&) {
    r = nc::fail(NC_ERR_OOM, "host allocation failed");
  } catch (const std::exception& e) {
    r = nc::fail(NC_ERR_STATE, std::string("unexpected host error: ") + e.what());
  }
  std::lock_guard<std::mutex> g(s->mu);
  if (r != NC_OK) {
    s->loop_status = r;
    const char* m = nc_last_error();
    s->loop_error = m ? m : "";
  }
  s->stop = true;
  s->cv_build.notify_all();
  s->cv_result.notify_all();
}

}  // namespace

extern "C" uint64_t nc_digest(const void* data, uint64_t n, uint64_t h) {
  const unsigned char* p = (const unsigned char*)data;
  for (uint64_t i = 0; i < n; ++i) {
    h ^= p[i];
    h *= 1099511628211ull;
  }
  return h;
}

extern "C" nc_status nc_server_create(const nc_server_spec* spec, nc_server** out) try {
  if (!spec || !out) return nc::fail(NC_ERR_ARG, "null argument");
  *out = nullptr;
  std::unique_ptr<nc_server> s(new nc_server());
  s->spec = *spec;
  nc_server_spec& p = s->spec;
  if (!p.lanes) p.lanes = 2;
  if (!p.chunk) p.chunk = 200;
  if (!p.builders) p.builders = 4;
  if (!p.max_nets) p.max_nets = 2048;
  if (p.hbm_fraction <= 0) p.hbm_fraction = 0.5f;
  if (p.syn_per_neuron <= 0) p.syn_per_neuron = 60.0f;
  if (p.lanes > 64 || p.builders > 256 || p.hbm_fraction > 0.95f || p.evict_after_s < 0)
    return nc::fail(NC_ERR_ARG, "server spec out of range");
  if (!p.max_neurons) {
    // the lanes' share of free HBM, read through a probe simulation
    Lane probe;
    nc_status r = lane_create(s.get(), 64, 64, false, &probe);
    if (r != NC_OK) return r;
    uint64_t free_b = 0, per = 0;
    r = nc_sim_free_bytes(probe.sim, &free_b);
    nc_sim_destroy(probe.sim);
    if (r == NC_OK) r = nc_sim_footprint(1000000, (uint64_t)(1e6 * p.syn_per_neuron), &per);
    if (r != NC_OK) return r;
    const double n = (double)free_b * p.hbm_fraction / p.lanes / ((double)per / 1e6);
    p.max_neurons = (uint64_t)n;
    p.max_synapses = (uint64_t)(n * p.syn_per_neuron);
  }
  if (!p.max_synapses) p.max_synapses = (uint64_t)(p.max_neurons * p.syn_per_neuron);
  for (uint32_t i = 0; i < p.lanes; ++i) {
    std::unique_ptr<Lane> L(new Lane());
    nc_status r = lane_create(s.get(), p.max_neurons, p.max_synapses, false, L.get());
    if (r != NC_OK) {
      for (auto& o : s->lanes) nc_sim_destroy(o->sim);
      return r;
    }
    s->lanes.push_back(std::move(L));
  }
  s->t0 = Clock::now();
  s->n_lanes = (uint32_t)s->lanes.size();
  nc_server* raw = s.get();
  for (uint32_t i = 0; i < p.builders; ++i) s->builders.emplace_back(builder_main, raw);
  s->loop = std::thread(loop_main, raw);
  *out = s.release();
  return NC_OK;
}
NC_ABI_CATCH

extern "C" nc_status nc_server_submit(nc_server* s, const nc_job* job) try {
  if (!s || !job) return nc::fail(NC_ERR_ARG, "null argument");
  if (!job->steps) return nc::fail(NC_ERR_ARG, "a job needs a step count");
  nc_status r = nc::check_wiring_spec(job->wiring);
  if (r != NC_OK) return r;
  std::unique_ptr<JobRec> j(new JobRec());
  j->job = *job;
  j->left = job->steps;
  j->res.id = job->id;
  j->res.steps = job->steps;
  std::lock_guard<std::mutex> g(s->mu);
  if (s->stop) return nc::fail(NC_ERR_STATE, "the server has stopped: " + s->loop_error);
  j->res.t_submit = s->now();
  JobRec* raw = j.get();
  s->jobs.push_back(std::move(j));
  raw->self = std::prev(s->jobs.end());
  s->to_build.push_back(raw);
  ++s->st.submitted;
  s->cv_build.notify_one();
  return NC_OK;
}
NC_ABI_CATCH

extern "C" nc_status nc_server_poll(nc_server* s, nc_job_result* out, uint32_t cap, uint32_t* n_out,
                                    double wait_s) try {
  if (!s || !n_out || (!out && cap)) return nc::fail(NC_ERR_ARG, "null argument");
  std::unique_lock<std::mutex> g(s->mu);
  if (s->results.empty() && wait_s > 0 && !s->stop)
    s->cv_result.wait_for(g, std::chrono::duration<double>(wait_s),
                          [&] { return !s->results.empty() || s->stop; });
  const uint32_t k = (uint32_t)std::min<size_t>(cap, s->results.size());
  std::copy(s->results.begin(), s->results.begin() + k, out);
  s->results.erase(s->results.begin(), s->results.begin() + k);
  *n_out = k;
  if (!k && s->loop_status != NC_OK) return nc::fail(s->loop_status, "server loop stopped: " + s->loop_error);
  return NC_OK;
}
NC_ABI_CATCH

extern "C" nc_status nc_server_stats_get(nc_server* s, nc_server_stats* out) try {
  if (!s || !out) return nc::fail(NC_ERR_ARG, "null argument");
  std::lock_guard<std::mutex> g(s->mu);
  *out = s->st;
  out->waiting = (uint32_t)(s->ready.size() + s->to_build.size());
  out->lanes = s->n_lanes;
  out->mean_running = s->st.chunks ? s->running_sum / (double)s->st.chunks : 0.0;
  return NC_OK;
}
NC_ABI_CATCH

extern "C" nc_status nc_server_destroy(nc_server* s) try {
  if (!s) return NC_OK;
  {
    std::lock_guard<std::mutex> g(s->mu);
    s->stop = true;
  }
  s->cv_build.notify_all();
  s->cv_loop.notify_all();
  if (s->loop.joinable()) s->loop.join();
  for (std::thread& t : s->builders) t.join();
  const nc_status r = s->loop_status;
  const std::string msg = s->loop_error;
  s->jobs.clear();
  for (auto& L : s->lanes) nc_sim_destroy(L->sim);
  delete s;
  if (r != NC_OK) return nc::fail(r, "server loop stopped: " + msg);
  return NC_OK;
}
NC_ABI_CATCH
