// Request loop with continuous batching (one GPU).
//
// Requests carry k load cases for a mesh given by its options.  Set-up meshes are
// cached on the device (factorisations, explicit local operators, coarse
// problems, block buffers) and admitted by their HBM footprint; idle meshes are
// evicted least recently used when a new one does not fit.  Every cached mesh owns
// a block of slots: queued load cases of any request for that mesh take free
// slots at the next chunk boundary, all running slots advance together, and a
// slot whose case has stopped is recovered and released.  Meshes run on their own
// streams, so the work of different meshes overlaps on the device.  A case's
// arithmetic is independent of the other slots, so its result is the same alone
// and in the mix.  A cold mesh is set up on a worker thread (its own stream) while
// the cached meshes keep iterating; one set-up runs at a time, in arrival order
// except that a mesh with less than an eighth of the set-up work of the earliest
// waiting one may go first.  Optionally, evicted meshes are kept in pinned host
// memory (their set-up state, see state.hpp): a later request restores the state
// over the host link instead of setting the mesh up again.
#pragma once
#include <cstdint>
#include <deque>
#include <future>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "engine.hpp"
#include "state.hpp"

namespace hfls {

struct ServeRequest {
  int64_t id = 0;
  std::string key;
  int k = 0;
  std::vector<double> F;               // k packed local loads (empty for seeded tractions)
  bool seeded = false;                 // face tractions of load cases case0 .. case0 + k - 1
  uint64_t seed = 0;
  int case0 = 0;
  double t_submit = 0, t_admit = -1, t_done = -1;  // ms since the server started
  int started = 0, remaining = 0;
  int64_t ng = 0;                      // 3 * nPoints
  std::shared_ptr<DVec<double>> dU;    // k x 3*nPoints displacements, device resident until fetched
  std::vector<int> it, status;
  std::vector<double> res;
  std::string error;                   // set-up failure of the request's mesh
};

struct ServeSlot {
  int64_t req = -1;
  int idx = 0;
};

struct ServeMesh {
  std::string key;
  Options opts;
  std::unique_ptr<Engine> e;
  size_t bytes = 0;
  int64_t np = 0, ng = 0;
  std::vector<ServeSlot> slot;
  std::deque<ServeSlot> queue;         // cases admitted, waiting for a slot
  bool swept = false;
  double last_used = 0;
  std::vector<int> waited;             // per slot: rounds since its case stopped (-1: not stopped)
  std::vector<int> it_seen;            // per slot: iterations of its case at the last sync
  int it_max = 0, it_cases = 0;        // most iterations of a case this mesh has finished; cases
  bool no_save = false;                // its copy to the host cache failed (too large): not retried
};

struct ServeStats {
  int64_t cases_done = 0, requests_done = 0, setups = 0, evictions = 0, joins = 0, finishes = 0;
  int64_t failed_setups = 0, cancelled = 0, steps = 0;
  double setup_ms = 0;
  double join_ms = 0, sweep_ms = 0, finish_ms = 0;  // host wall time of the step phases (device work included)
  // host-memory cache of evicted meshes: saves of evicted meshes, restores in place
  // of set-ups, saved meshes dropped to make room; the ms as setup_ms (admission to
  // cached, on the worker thread); save_ms is the background copies' wall time
  int64_t saves = 0, restores = 0, host_drops = 0, host_meshes = 0;
  double save_ms = 0, restore_ms = 0;
  size_t host_bytes = 0;
};

class Server {
 public:
  Server(int device, size_t budget_bytes, int slots, int precision, double eps, int max_iter, int chunk);
  ~Server();
  // packed local length and global DOF count of a mesh (host-side problem)
  void sizes(const Options& o, int64_t* npacked, int64_t* nglob, int64_t* bytes);
  // F: k packed local loads, or null for k seeded face tractions (seed, case0..)
  int64_t submit(const Options& o, int k, const double* F, uint64_t seed, int case0);
  bool step();                                   // one scheduling round; false when idle
  int state(int64_t id) const;                   // 0 queued, 1 running, 2 done, -1 unknown
  int request_k(int64_t id) const;
  // per-case status: 2 converged, 3 iteration limit, 6 non-finite residual, 7 the
  // mesh's set-up failed (the message is thrown after the outputs are written)
  void result(int64_t id, double* u, int* it, int* status, double* res, double* latency_ms, int* k);
  // drop a request: its queued cases are removed, its running cases leave their
  // slots (without recovery); the id is unknown afterwards
  void cancel(int64_t id);
  // pinned host memory for the set-up state of evicted meshes (0: off, the default);
  // least recently used saved meshes are dropped when a new one does not fit
  void set_host_cache(size_t bytes);
  const ServeStats& stats() const { return st_; }
  size_t bytes_in_use() const;  // cached meshes, a set-up in flight, results not yet fetched
  int meshes() const { return (int)mesh_.size(); }

 private:
  int device_, slots_, prec_, max_iter_, chunk_;
  double eps_;
  size_t budget_;
  int64_t next_id_ = 1;
  double t0_;
  std::map<int64_t, ServeRequest> req_;
  std::deque<int64_t> wait_;                     // submitted, mesh not admitted yet
  std::map<std::string, std::unique_ptr<ServeMesh>> mesh_;
  std::map<std::string, std::unique_ptr<Problem>> host_;  // host problems (sizes, loads)
  ServeStats st_;
  struct SetupOut {  // what the worker thread hands back
    std::unique_ptr<ServeMesh> m;
    bool restored = false;
  };
  struct PendingSetup {  // a mesh being set up (or restored) on the worker thread
    std::string key;
    size_t need = 0;
    double t0 = 0;
    std::future<SetupOut> fut;
  };
  // host-memory cache of set-up states: written by the saver thread, read by the
  // set-up worker, guarded by host_mu_ (an entry in use by a restore stays alive
  // through its shared pointer when the saver drops it)
  struct SavedMesh {
    std::shared_ptr<HostMesh> h;
    double last_used = 0;
  };
  struct PendingSave {  // a cached mesh being copied to host memory
    ServeMesh* m = nullptr;
    double t0 = 0;
    std::future<int64_t> fut;  // saved meshes dropped to make room
  };
  std::mutex host_mu_;
  size_t host_budget_ = 0, saved_bytes_ = 0;
  std::map<std::string, SavedMesh> saved_;
  std::unique_ptr<PendingSave> save_;
  int64_t host_trim(size_t budget);  // drop LRU saved meshes down to budget (host_mu_ held)
  void start_save();                 // copy a cached mesh without a saved copy, in the background
  bool finish_save(bool wait);       // collect a finished copy; true if one was collected
  std::unique_ptr<PendingSetup> pend_;
  size_t result_bytes_ = 0;                      // device bytes of results not yet fetched
  bool finish_setup(bool wait);                  // install a finished set-up; true if one was installed
  void fail_waiting(const std::string& key, const std::string& msg);  // requests of a failed set-up
  bool any_work() const;                         // a cached mesh has running or queued cases
  double now() const;
  const Problem& host_problem(const std::string& key, const Options& o);
  size_t estimate(const Problem& p) const;
  bool admit(ServeRequest& r);
};

std::string options_key(const Options& o);

}  // namespace hfls
