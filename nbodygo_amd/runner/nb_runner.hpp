// nb_runner.hpp -- host side of the drop-in boundary in C++: the reference's body collection,
// result queue and computation runner (cmd/body/body_collection.go, cmd/runner/resultqueue.go,
// cmd/runner/computation-runner.go), driving libnbgpu.so through include/nbody_gpu.h.
//
// The device holds the body state for the whole run: it is loaded once (nb_load) and every
// later change happens there -- deletes (nb_remove_bodies), body queries and mods (nb_get_body,
// nb_set_body), and Cycle's compaction and adds (nb_add_bodies + nb_cycle,
// body_collection.go:253-296).  The host collection keeps what the device does not (ids, names,
// classes, colours, intensities, frag steps) index for index with the device, compacted with the
// Exists flags each render returns; its physical fields are as of the last SyncToHost (pulled
// with nb_store_bodies).  Phase order per cycle is that of runOneComputation
// (computation-runner.go:267-326): control polls, force + detection + resolution + update on the
// device (nb_step, one cycle), renderables into a result queue, Cycle.
#pragma once

#include <atomic>
#include <condition_variable>
#include <cstdint>
#include <deque>
#include <functional>
#include <memory>
#include <mutex>
#include <optional>
#include <string>
#include <thread>
#include <vector>

#include "../../include/nbody_gpu.h"

namespace nbr {

// The fields of body.Body (cmd/body/body.go:34-52) that the per-cycle path and the collection
// lifecycle use.
struct Body {
  int64_t id = 0;
  double x = 0, y = 0, z = 0, vx = 0, vy = 0, vz = 0, mass = 0, radius = 0;
  int32_t collision_behavior = NB_CB_ELASTIC;
  int32_t color = 0;
  double frag_factor = 0, frag_step = 0;
  double r = 1.0;  // coefficient of restitution carried by the body
  bool exists = true, fragmenting = false, is_sun = false, pinned = false;
  double intensity = 0;
  std::string name, cls;

  void SetNotExists() {  // body.go:93-96
    mass = 0;
    exists = false;
  }
  void SetSun(double i) {  // body.go:101-104
    is_sun = true;
    intensity = i;
  }
  // "name=value" modifications (body.go:274-313); returns true like the reference.
  bool ApplyMods(const std::vector<std::string> &mods);
};

int64_t NextId();  // body.go:316-329
constexpr double kFragmentRadius = 1.0;  // fragcalc.go:81-82: Pow(x, 1/3) with 1/3 == 0

// renderable.go:9-40
struct Renderable {
  int64_t id = 0;
  bool exists = false;
  float x = 0, y = 0, z = 0;
  double radius = 0;
  bool is_sun = false;
  float intensity = 0;
  int32_t color = 0;
};

// resultqueue.go: a bounded FIFO of per-cycle renderable queues.
struct ResultQueue {
  uint64_t queue_num = 0;
  std::vector<Renderable> queue;
  void Add(const Renderable &r) { queue.push_back(r); }
};

class ResultQueueHolder {
 public:
  explicit ResultQueueHolder(int max_queues) : max_(max_queues) {}
  // A new queue, or nothing when max_queues are waiting (the runner then skips the cycle).
  std::optional<ResultQueue> NewResultQueue();
  void Add(ResultQueue q);
  std::optional<ResultQueue> Next();
  int MaxQueues();
  bool Resize(int max_queues);
  int Size();

 private:
  std::mutex m_;
  std::deque<ResultQueue> q_;
  int max_;
  uint64_t queue_num_ = 0;
};

// FromCsv (cmd/sim/fromcsv.go:46-112): at most body_count parseable rows; unparseable rows are
// skipped silently; a malformed CSV record (bad quoting) is an error, as log.Fatalf is there.
bool FromCsv(const std::string &path, int body_count, int32_t default_collision,
             int32_t default_color, std::vector<Body> *out, std::string *err);
bool WriteCsv(const std::string &path, const std::vector<Body> &bodies);
int32_t ParseCollisionBehavior(const std::string &s);  // globals.go: unknown -> elastic
int32_t ParseBodyColor(const std::string &s);          // globals.go: unknown -> random

class BodyCollection {
 public:
  explicit BodyCollection(std::vector<Body> bodies) : arr_(std::move(bodies)) {}
  std::vector<Body> &GetArray() { return arr_; }
  int Count();
  void IterateOnce(const std::function<void(Body &)> &f) {
    for (auto &b : arr_) f(b);
  }
  // Adds are queued and applied by Cycle, like the reference's add events.
  void EnqueueAdd(const Body &b);
  bool HasAdds() {
    std::lock_guard<std::mutex> g(lock_);
    return !adds_.empty();
  }
  std::vector<Body> TakeAdds();
  // The host half of nb_cycle: keeps the bodies whose keep flag is set (nullptr: the exists
  // field), then appends the device's fragments and then adds, each newest first with r = R.
  // Returns true when the array changed.
  bool Cycle(double R, const uint8_t *keep, const std::vector<Body> &adds,
             const std::vector<Body> &frags);
  int cycle() const { return cycle_; }

 private:
  std::vector<Body> arr_;
  std::mutex lock_;
  std::vector<Body> adds_;
  int cycle_ = 0;
};

// A one-slot mailbox, the runner-side half of the reference's buffered control channels
// (timeScaleChan, RChan, delChan): a later Put replaces a value not yet taken.
template <class T>
class Mailbox {
 public:
  void Put(T v) {
    std::lock_guard<std::mutex> g(m_);
    v_ = std::move(v);
  }
  std::optional<T> Poll() {
    std::lock_guard<std::mutex> g(m_);
    std::optional<T> r;
    r.swap(v_);
    return r;
  }

 private:
  std::mutex m_;
  std::optional<T> v_;
};

struct RunnerStats {
  uint64_t iterations = 0, computations = 0, skipped_no_queue = 0, loads = 0, compactions = 0;
  double step_ms = 0, wall_ms = 0, interactions = 0;
  int64_t pairs_total = 0, removed_total = 0;
};

class ComputationRunner {
 public:
  // workerCnt is kept for the reference's signature (computation-runner.go:85); the device
  // replaces the worker pool.  barnesHut selects nothing: the direct sum always runs (the
  // reference's octree never builds, octree.go:248-250).
  ComputationRunner(int worker_cnt, double time_scaling, bool barnes_hut, ResultQueueHolder *rqh,
                    BodyCollection *bc, const nb_params &params);
  ~ComputationRunner();
  ComputationRunner(const ComputationRunner &) = delete;
  ComputationRunner &operator=(const ComputationRunner &) = delete;

  bool ok() const { return ctx_ != nullptr; }
  const std::string &error() const { return err_; }

  ComputationRunner &SetMaxIterations(int n);
  ComputationRunner &Start();   // runs cycles on a thread until stopped or max iterations
  void Stop();                  // asks the thread to stop and joins it
  void Wait();                  // joins after max iterations
  bool running() const { return running_; }

  void SetWorkers(int n) { worker_cnt_ = n; }
  int WorkerCount() const { return worker_cnt_; }
  double TimeScaling() const { return time_scaling_; }
  void SetTimeScaling(double t) { time_scale_box_.Put(t); }
  double CoefficientOfRestitution() const { return R_; }
  void SetCoefficientOfRestitution(double R) { R_box_.Put(R); }
  void RemoveBodies(int n) { del_box_.Put(n); }  // -1 removes every body
  // Synchronous query / modification, answered by the runner between cycles (the reference's
  // GetBody / ModBody handshakes, body_collection.go:111-191).
  std::optional<Body> GetBody(int64_t id, const std::string &name);
  int ModBody(int64_t id, const std::string &name, const std::string &cls,
              const std::vector<std::string> &mods);  // 0 no match, 1 none, 2 some, 3 all
  void AddBody(const Body &b) { bc_->EnqueueAdd(b); }

  // One cycle on the calling thread (also what the runner thread calls); returns an NB_* code.
  int RunOneComputation();
  // Pulls the whole device state into the host collection (for output or inspection).
  int SyncToHost();
  // nb_body <-> Body for the fields the device holds
  static nb_body ToDevice(const Body &b);
  static void FromDevice(const nb_body &d, Body *b);
  RunnerStats stats() const;
  void PrintStats() const;  // computation-runner.go:61-71

 private:
  void run();
  int cycle();
  void processDeletes();
  void handleRequests();
  int pushIfDirty();

  int worker_cnt_;
  double time_scaling_;
  double R_ = 1.0;
  ResultQueueHolder *rqh_;
  BodyCollection *bc_;
  nb_ctx *ctx_ = nullptr;
  std::string err_;
  bool dirty_ = true;       // nothing loaded on the device yet
  bool events_ = false;     // nb_params.events: subsume and fragmentation on the device
  bool host_stale_ = false; // device advanced since the last pull
  int max_iteration_ = 0;
  std::atomic<bool> running_{false}, stop_{false};
  std::thread thread_;
  Mailbox<double> time_scale_box_, R_box_;
  Mailbox<int> del_box_;
  struct Request {
    int kind = 0;  // 1 get, 2 mod
    int64_t id = -1;
    std::string name, cls;
    std::vector<std::string> mods;
    std::optional<Body> body;
    int result = 0;
    bool done = false;
  };
  std::mutex req_m_;
  std::condition_variable req_cv_;
  std::deque<Request *> requests_;
  mutable std::mutex stats_m_;
  RunnerStats st_;
  std::vector<float> xyz_;
  std::vector<uint8_t> ex_;
};

}  // namespace nbr
