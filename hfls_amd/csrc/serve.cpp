// Request loop with continuous batching; see serve.hpp.
#include "serve.hpp"

#include <algorithm>
#include <chrono>
#include <cstring>
#include <stdexcept>

namespace hfls {

static int64_t round8l(int64_t x) { return (x + 7) / 8 * 8; }
static int64_t align32l(int64_t x) { return (x + 31) / 32 * 32; }

std::string options_key(const Options& o) {
  Options full = default_options();
  for (auto& kv : o) full[kv.first] = kv.second;
  std::string k;
  for (auto& kv : full) {
    if (kv.first == "device" || kv.first == "eps_iter" || kv.first == "max_iter" || kv.first == "vtkWithinIter" ||
        kv.first == "print_matrices" || kv.first == "precision" || kv.first == "nrhs" || kv.first == "refine_tol")
      continue;  // do not change the set-up mesh
    k += kv.first + "=" + kv.second + ";";
  }
  return k;
}

Server::Server(int device, size_t budget, int slots, int prec, double eps, int max_iter, int chunk)
    : device_(device), slots_(slots), prec_(prec), max_iter_(max_iter), chunk_(chunk), eps_(eps), budget_(budget) {
  if (slots < 1 || slots > KMAX) throw std::runtime_error("slots per mesh must be 1..64");
  if (prec != 32 && prec != 64) throw std::runtime_error("precision must be 32 or 64");
  if (chunk < 1) throw std::runtime_error("chunk must be positive");
  t0_ = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

Server::~Server() {
  if (pend_ && pend_->fut.valid()) pend_->fut.wait();
  if (save_ && save_->fut.valid()) save_->fut.wait();
}

double Server::now() const {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count() - t0_;
}

const Problem& Server::host_problem(const std::string& key, const Options& o) {
  auto it = host_.find(key);
  if (it != host_.end()) return *it->second;
  std::unique_ptr<Problem> p(new Problem(o));
  const Problem& ref = *p;
  host_[key] = std::move(p);
  return ref;
}

// device bytes a set-up mesh keeps (plain path + block buffers), from the host
// problem; the transient factorisation scratch is sized by the free memory
// relative set-up work of a mesh: the dense elimination of every subdomain
// matrix, sum of n_i^3
static double setup_work(const Problem& p) {
  double w = 0.0;
  for (const Subdomain& s : p.subs) w += (double)s.n * s.n * s.n;
  return w;
}

size_t Server::estimate(const Problem& p) const {
  double b = 0.0;
  const double nl = (double)p.nlam, ng = 3.0 * (double)p.nPoints;
  double np = 0;
  for (const Subdomain& s : p.subs) {
    const int64_t ldb = round8l(s.n_s);
    const int64_t T = (s.n_s + 63) / 64, tiles = T * (T + 1) / 2 * 4096;
    const int64_t TL = (s.n_r + 63) / 64;
    np += s.n;
    b += 8.0 * (double)(TL * (TL + 1) / 2 * 4096);                        // L_rr tiles
    b += 2.0 * 12.0 * 81.0 * (double)s.n_s;                               // K_br, K_rb (stencil bound)
    b += (8.0 + 8.0 + 4.0 + 4.0) * (double)tiles;                         // K+_bb, S (FP64, FP32)
    b += 4.0 * (double)tgemm_h16b_part_floats(s.n_s, KMAX);               // band GEMM partials
    b += 8.0 * (6.0 * s.n + 6.0 * s.n_s + 4.0 * s.n + 2.0 * s.n_s);       // R, R_b, vectors
    b += 8.0 * (2.0 * align32l(ldb * 36) + 54.0 * ldb);                   // C, Bc^T, [C|R_b|Q|Yh]
    b += 8.0 * (double)T * (T + 1) / 2 * 64;                              // tile-row partials
    b += (8.0 + 8.0 + 4.0) * 64.0 * 64.0 * T;                             // b-blocks
    b += 8.0 * 64.0 * (2.0 * ((s.n_r + 63) / 64 * 64) + 64.0 * 3);        // interior, cr/coef/v6 blocks
  }
  b += 8.0 * (10.0 * nl + 9.0 * 64.0 * nl);                               // dual vectors and row indices
  b += 8.0 * 64.0 * (3.0 * np + ng) + 8.0 * (ng + 4.0 * np);               // local blocks, U
  return (size_t)(1.1 * b);
}

size_t Server::bytes_in_use() const {
  size_t s = (pend_ ? pend_->need : 0) + result_bytes_;
  for (auto& kv : mesh_) s += kv.second->bytes;
  return s;
}

bool Server::any_work() const {
  for (auto& kv : mesh_) {
    if (!kv.second->queue.empty()) return true;
    for (auto& sl : kv.second->slot)
      if (sl.req >= 0) return true;
  }
  return false;
}

void Server::fail_waiting(const std::string& key, const std::string& msg) {
  const double t = now();
  for (auto it = wait_.begin(); it != wait_.end();) {
    ServeRequest& r = req_[*it];
    if (r.key != key) {
      ++it;
      continue;
    }
    r.error = msg;
    std::fill(r.status.begin(), r.status.end(), 7);
    r.remaining = 0;
    r.t_done = t;
    st_.requests_done++;
    it = wait_.erase(it);
  }
}

// with host_mu_ held
int64_t Server::host_trim(size_t budget) {
  int64_t dropped = 0;
  while (saved_bytes_ > budget && !saved_.empty()) {
    auto lru = saved_.begin();
    for (auto it = saved_.begin(); it != saved_.end(); ++it)
      if (it->second.last_used < lru->second.last_used) lru = it;
    saved_bytes_ -= lru->second.h->bytes;
    saved_.erase(lru);
    dropped++;
  }
  return dropped;
}

void Server::set_host_cache(size_t bytes) {
  finish_save(true);  // no copy in flight while the budget changes
  std::lock_guard<std::mutex> lk(host_mu_);
  host_budget_ = bytes;
  st_.host_drops += host_trim(bytes);
  st_.host_bytes = saved_bytes_;
  st_.host_meshes = (int64_t)saved_.size();
}

// Start copying the set-up state of a cached mesh that has no saved copy yet (the
// most recently used first) to pinned host memory, on a thread and a stream of its
// own: the state is immutable after set-up, so the copy overlaps the mesh's own
// iterations, and evicting the mesh later costs nothing.
void Server::start_save() {
  if (save_ || host_budget_ == 0) return;
  ServeMesh* pick = nullptr;
  {
    std::lock_guard<std::mutex> lk(host_mu_);
    for (auto& kv : mesh_) {
      ServeMesh* m = kv.second.get();
      if (m->no_save || saved_.count(m->key)) continue;
      if (!pick || m->last_used > pick->last_used) pick = m;
    }
  }
  if (!pick) return;
  save_.reset(new PendingSave);
  save_->m = pick;
  save_->t0 = now();
  const int dev = device_;
  const std::string key = pick->key;
  Engine* e = pick->e.get();
  save_->fut = std::async(std::launch::async, [this, dev, key, e]() -> int64_t {
    HC(hipSetDevice(dev));
    hipStream_t cs = nullptr;  // the engine's own streams belong to the server's thread
    HC(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
    int64_t dropped = 0;
    std::shared_ptr<HostMesh> h;
    try {
      const size_t n = saved_bytes(*e, cs);  // host-side counting pass
      {
        std::lock_guard<std::mutex> lk(host_mu_);
        if (n > host_budget_) throw std::runtime_error("saved mesh larger than the host cache");
        dropped = host_trim(host_budget_ - n);
      }
      h = save_mesh_host(*e, cs);
    } catch (...) {
      (void)hipStreamSynchronize(cs);
      (void)hipStreamDestroy(cs);
      throw;
    }
    HC(hipStreamDestroy(cs));
    std::lock_guard<std::mutex> lk(host_mu_);
    saved_bytes_ += h->bytes;
    SavedMesh& s = saved_[key];
    s.h = std::move(h);
    s.last_used = now();
    return dropped;
  });
}

bool Server::finish_save(bool wait) {
  if (!save_) return false;
  if (!wait && save_->fut.wait_for(std::chrono::seconds(0)) != std::future_status::ready) return false;
  try {
    st_.host_drops += save_->fut.get();
    st_.saves++;
    st_.save_ms += now() - save_->t0;
  } catch (const std::exception&) {
    save_->m->no_save = true;  // does not fit (or no pinned memory): not tried again
  }
  save_.reset();
  std::lock_guard<std::mutex> lk(host_mu_);
  st_.host_bytes = saved_bytes_;
  st_.host_meshes = (int64_t)saved_.size();
  return true;
}

bool Server::finish_setup(bool wait) {
  if (!pend_) return false;
  if (!wait && pend_->fut.wait_for(std::chrono::seconds(0)) != std::future_status::ready) return false;
  SetupOut out;
  try {
    out = pend_->fut.get();  // rethrows a failed set-up
  } catch (const std::exception& ex) {
    // the waiting requests of this mesh end with status 7; the server goes on
    const std::string key = pend_->key;
    st_.setup_ms += now() - pend_->t0;
    st_.failed_setups++;
    pend_.reset();
    fail_waiting(key, std::string("set-up of the mesh failed: ") + ex.what());
    return false;
  }
  std::unique_ptr<ServeMesh> m = std::move(out.m);
  const double ms = now() - pend_->t0;
  if (out.restored) {
    st_.restores++;
    st_.restore_ms += ms;
  } else {
    st_.setups++;
    st_.setup_ms += ms;
  }
  m->last_used = now();
  mesh_.emplace(pend_->key, std::move(m));
  pend_.reset();
  return true;
}

void Server::sizes(const Options& o, int64_t* npacked, int64_t* nglob, int64_t* bytes) {
  const Problem& p = host_problem(options_key(o), o);
  int64_t np = 0;
  for (const Subdomain& s : p.subs) np += s.n;
  if (npacked) *npacked = np;
  if (nglob) *nglob = 3 * (int64_t)p.nPoints;
  if (bytes) *bytes = (int64_t)estimate(p);
}

int64_t Server::submit(const Options& o, int k, const double* F, uint64_t seed, int case0) {
  if (k < 1) throw std::runtime_error("a request needs at least one load case");
  const std::string key = options_key(o);
  const Problem& p = host_problem(key, o);
  if (estimate(p) > budget_) throw std::runtime_error("mesh does not fit the HBM budget of the server");
  int64_t np = 0;
  for (const Subdomain& s : p.subs) np += s.n;
  ServeRequest r;
  r.id = next_id_++;
  r.key = key;
  r.k = k;
  if (F) {
    r.F.resize((size_t)k * np);
    std::memcpy(r.F.data(), F, r.F.size() * sizeof(double));
  } else {  // generated at the join: fields on the host, lumping on the device
    r.seeded = true;
    r.seed = seed;
    r.case0 = case0;
  }
  r.t_submit = now();
  r.remaining = k;
  r.ng = 3 * (int64_t)p.nPoints;
  r.it.assign(k, 0);
  r.status.assign(k, 0);
  r.res.assign(k, 0.0);
  const int64_t id = r.id;
  req_[id] = std::move(r);
  wait_.push_back(id);
  return id;
}

bool Server::admit(ServeRequest& r) {
  auto it = mesh_.find(r.key);
  if (it == mesh_.end()) {
    if (pend_) return false;  // one set-up at a time (possibly this very mesh's)
    const Problem& p = *host_[r.key];
    const size_t need = estimate(p);
    // evict idle meshes, least recently used first, until the new one fits (a mesh
    // being copied to the host cache is evicted once its copy is done)
    while (bytes_in_use() + need > budget_) {
      ServeMesh* lru = nullptr;
      for (auto& kv : mesh_) {
        ServeMesh* m = kv.second.get();
        bool idle = m->queue.empty();
        for (auto& s : m->slot) idle = idle && s.req < 0;
        if (idle && (!lru || m->last_used < lru->last_used)) lru = m;
      }
      if (!lru) return false;  // everything cached is busy: wait
      if (save_ && save_->m == lru) finish_save(true);
      mesh_.erase(lru->key);
      st_.evictions++;
    }
    // set the mesh up (or restore it from the host cache) on the worker thread; the
    // request is admitted once it is cached
    pend_.reset(new PendingSetup);
    pend_->key = r.key;
    pend_->need = need;
    pend_->t0 = now();
    const Options o = p.o;
    const int64_t ng = 3 * (int64_t)p.nPoints;
    const int dev = device_, slots = slots_, mi = max_iter_, prec = prec_;
    const double eps = eps_;
    const std::string key = r.key;
    pend_->fut = std::async(std::launch::async, [this, o, ng, dev, slots, mi, prec, eps, key, need]() {
      HC(hipSetDevice(dev));
      SetupOut out;
      std::unique_ptr<ServeMesh> m(new ServeMesh);
      m->key = key;
      m->opts = o;
      size_t f0 = 0, f1 = 0, tot = 0;
      HC(hipMemGetInfo(&f0, &tot));
      std::shared_ptr<HostMesh> h;
      {
        std::lock_guard<std::mutex> lk(host_mu_);
        auto sv = saved_.find(key);
        if (sv != saved_.end()) {
          h = sv->second.h;  // kept alive by this copy even if the saver drops the entry
          sv->second.last_used = now();
        }
      }
      if (h) {
        try {
          m->e = load_mesh_host(*h, dev);
          out.restored = true;
        } catch (const std::exception&) {
          m->e.reset();  // set up instead
        }
        h.reset();
      }
      if (!m->e) {
        m->e.reset(new Engine(o, dev));
        m->e->setup();
      }
      m->e->blk_begin(slots, eps, mi, prec);
      HC(hipStreamSynchronize(m->e->stream()));
      HC(hipMemGetInfo(&f1, &tot));
      m->bytes = std::max(need, f0 > f1 ? f0 - f1 : (size_t)0);
      m->np = m->e->packed_n();
      m->ng = ng;
      m->slot.assign(slots, ServeSlot());
      out.m = std::move(m);
      return out;
    });
    return false;
  }
  ServeMesh& m = *it->second;
  for (int c = 0; c < r.k; c++) m.queue.push_back(ServeSlot{r.id, c});
  m.last_used = now();
  r.t_admit = now();
  return true;
}

bool Server::step() {
  HC(hipSetDevice(device_));
  // 1. admission in arrival order; a request whose mesh does not fit yet waits
  //    without blocking the requests behind it.  A finished set-up is installed
  //    first; with nothing else to run, the loop waits for the running set-up.
  finish_setup(false);
  finish_save(false);
  start_save();
  auto admission = [&] {
    // requests for cached meshes, in arrival order
    for (auto it = wait_.begin(); it != wait_.end();) {
      if (mesh_.count(req_[*it].key) && admit(req_[*it])) it = wait_.erase(it);
      else ++it;
    }
    // one set-up at a time: the earliest waiting request's mesh, unless a later
    // request needs a mesh with less than an eighth of its set-up work (then the
    // smallest such goes first: it delays the large one little and saves its own
    // requests from waiting behind it)
    if (!pend_ && !wait_.empty()) {
      const double head = setup_work(*host_[req_[wait_.front()].key]);
      int64_t best = wait_.front();
      double bw = head;
      for (int64_t id : wait_) {
        const double w = setup_work(*host_[req_[id].key]);
        if (w < head / 8.0 && w < bw) {
          best = id;
          bw = w;
        }
      }
      admit(req_[best]);  // starts the set-up (or waits for room)
    }
  };
  admission();
  if (pend_ && !any_work()) {
    finish_setup(true);
    admission();
  }
  // 2. free slots take queued cases (all meshes, asynchronously).  A join runs the
  //    initialisation over the whole block, so cases join in batches: when at least
  //    a quarter of the slots (or every queued case) can start, or when the block
  //    has nothing else to run
  const int batch = std::max(1, slots_ / 4);
  st_.steps++;
  double tph = now();
  for (auto& kv : mesh_) {
    ServeMesh& m = *kv.second;
    if (m.queue.empty()) continue;
    int nfree = 0, nbusy = 0;
    for (auto& sl : m.slot) (sl.req < 0 ? nfree : nbusy)++;
    if (nfree == 0 || (nbusy > 0 && nfree < std::min<int>((int)m.queue.size(), batch))) continue;
    std::vector<int> cols, fcols;
    std::vector<double> F;
    std::map<uint64_t, std::pair<std::vector<int>, std::vector<int>>> tcols;  // seed -> slots, load cases
    for (int j = 0; j < slots_ && !m.queue.empty(); j++) {
      if (m.slot[j].req >= 0) continue;
      ServeSlot c = m.queue.front();
      m.queue.pop_front();
      m.slot[j] = c;
      ServeRequest& r = req_[c.req];
      cols.push_back(j);
      if (r.seeded) {
        auto& tc = tcols[r.seed];
        tc.first.push_back(j);
        tc.second.push_back(r.case0 + c.idx);
      } else {
        fcols.push_back(j);
        F.insert(F.end(), r.F.begin() + (size_t)c.idx * m.np, r.F.begin() + (size_t)(c.idx + 1) * m.np);
      }
      r.started++;
    }
    if (!cols.empty()) {
      if (!fcols.empty()) m.e->blk_load(fcols, F.data());
      for (auto& tc : tcols) m.e->blk_load_traction(tc.second.first, tc.first, tc.second.second);
      m.e->blk_init(cols);
      st_.joins += (int64_t)cols.size();
    }
  }
  // 3. sweeps for every mesh with work (different streams overlap): whole chunks until,
  //    by the most iterations a finished case of the mesh has taken, every running
  //    case will have stopped when cases wait for a slot (the first of them when
  //    none waits), at most 8 chunks; one chunk while the mesh has no finished case.
  //    Every round costs a whole-block initialisation for its joins and a recovery
  //    for its finishes (each about as long as 5 to 9 sweeps of a full block), so
  //    under load rounds that end when the block can change hands keep the slots
  //    full without paying those for a few cases at a time.
  st_.join_ms += now() - tph;
  tph = now();
  for (auto& kv : mesh_) {
    ServeMesh& m = *kv.second;
    m.swept = false;
    std::vector<int64_t> rem;
    if ((int)m.it_seen.size() != slots_) m.it_seen.assign(slots_, 0);
    for (int j = 0; j < slots_; j++)
      if (m.slot[j].req >= 0) {
        m.swept = true;  // it_seen < 0: stopped, waiting for its recovery
        if (m.it_cases > 0 && m.it_seen[j] >= 0) rem.push_back(std::max<int64_t>(1, m.it_max - m.it_seen[j]));
      }
    if (m.swept) {
      int nrep = 1;
      if (!rem.empty()) {
        std::sort(rem.begin(), rem.end());
        const size_t q = m.queue.empty() ? 0 : rem.size() - 1;
        nrep = (int)std::min<int64_t>(8, std::max<int64_t>(1, (rem[q] + chunk_ - 1) / chunk_));
      }
      for (int r = 0; r < nrep; r++) m.e->blk_sweeps(chunk_);
      m.last_used = now();
    }
  }
  // 4. stopped slots: recover, hand the displacements to their requests, release
  for (auto& kv : mesh_) {
    ServeMesh& m = *kv.second;
    if (!m.swept) continue;
    BScalars sh;
    m.e->blk_sync(&sh);
    st_.sweep_ms += now() - tph;  // the meshes' chunks run concurrently: waited for once
    tph = now();
    // stopped: converged, iteration limit, or non-finite (a bad case ends alone;
    // the other slots' arithmetic does not depend on it)
    std::vector<int> cols;
    int nrun = 0, maxw = 0;
    if ((int)m.waited.size() != slots_) m.waited.assign(slots_, -1);
    for (int j = 0; j < slots_; j++) {
      if (m.slot[j].req < 0) continue;
      m.it_seen[j] = sh.it[j];
      if (sh.done[j] == 2 || sh.done[j] == 3 || sh.done[j] == 6) {
        m.it_seen[j] = -1;
        cols.push_back(j);
        m.waited[j] = m.waited[j] < 0 ? 0 : m.waited[j] + 1;
        maxw = std::max(maxw, m.waited[j]);
      } else {
        nrun++;
      }
    }
    if (cols.empty()) continue;
    // the recovery also runs over the whole block: stopped cases are recovered
    // together once a quarter of the slots have stopped, when nothing else runs,
    // when queued cases need their slots, or after 4 rounds of waiting
    int nfree = 0;
    for (auto& sl : m.slot) nfree += sl.req < 0 ? 1 : 0;
    const bool need = !m.queue.empty() && nfree < std::min<int>((int)m.queue.size(), batch);
    if (!((int)cols.size() >= batch || nrun == 0 || need || maxw >= 4)) continue;
    for (int j : cols) m.waited[j] = -1;
    // recovery on the device, the displacements stay there (device to device into
    // the request's result buffer) until result() fetches them
    m.e->blk_recover(&sh);
    for (int j : cols) {
      ServeRequest& r = req_[m.slot[j].req];
      if (!r.dU) {
        r.dU = std::make_shared<DVec<double>>();
        r.dU->alloc((size_t)r.k * r.ng);
        result_bytes_ += r.dU->n * sizeof(double);
      }
      m.e->blk_copy_u(j, r.dU->p + (size_t)m.slot[j].idx * r.ng);
    }
    m.e->blk_release(cols);
    BScalars s2;
    m.e->blk_sync(&s2);
    const double t = now();
    for (size_t c = 0; c < cols.size(); c++) {
      const int j = cols[c];
      ServeRequest& r = req_[m.slot[j].req];
      const int i = m.slot[j].idx;
      r.it[i] = sh.it[j];
      r.status[i] = sh.done[j];
      if (sh.done[j] == 2) m.it_max = std::max(m.it_max, sh.it[j]), m.it_cases++;
      m.it_seen[j] = 0;
      r.res[i] = s2.res[j];
      if (--r.remaining == 0) {
        r.t_done = t;
        st_.requests_done++;
      }
      st_.cases_done++;
      m.slot[j] = ServeSlot();
    }
    st_.finishes += (int64_t)cols.size();
    st_.finish_ms += now() - tph;
    tph = now();
  }
  bool busy = !wait_.empty() || pend_ != nullptr;
  for (auto& kv : mesh_) {
    if (!kv.second->queue.empty()) busy = true;
    for (auto& s : kv.second->slot)
      if (s.req >= 0) busy = true;
  }
  return busy;
}

int Server::state(int64_t id) const {
  auto it = req_.find(id);
  if (it == req_.end()) return -1;
  if (it->second.t_done >= 0) return 2;
  if (it->second.t_admit >= 0) return 1;
  return 0;
}

int Server::request_k(int64_t id) const {
  auto it = req_.find(id);
  if (it == req_.end()) throw std::runtime_error("unknown request");
  return it->second.k;
}

void Server::result(int64_t id, double* u, int* it, int* status, double* res, double* latency_ms, int* k) {
  auto f = req_.find(id);
  if (f == req_.end()) throw std::runtime_error("unknown request");
  ServeRequest& r = f->second;
  if (r.t_done < 0) throw std::runtime_error("request not finished");
  if (k) *k = r.k;
  if (u) {
    if (r.dU) {
      HC(hipSetDevice(device_));
      HC(hipMemcpy(u, r.dU->p, r.dU->n * sizeof(double), hipMemcpyDeviceToHost));
    } else {
      std::memset(u, 0, (size_t)r.k * r.ng * sizeof(double));  // no case ran (failed set-up)
    }
  }
  if (r.dU) result_bytes_ -= r.dU->n * sizeof(double);
  for (int i = 0; i < r.k; i++) {
    if (it) it[i] = r.it[i];
    if (status) status[i] = r.status[i];
    if (res) res[i] = r.res[i];
  }
  if (latency_ms) {
    latency_ms[0] = r.t_done - r.t_submit;
    latency_ms[1] = r.t_admit - r.t_submit;
  }
  const std::string err = r.error;
  req_.erase(f);
  if (!err.empty()) throw std::runtime_error(err);
}

void Server::cancel(int64_t id) {
  auto f = req_.find(id);
  if (f == req_.end()) throw std::runtime_error("unknown request");
  for (auto it = wait_.begin(); it != wait_.end();)
    if (*it == id) it = wait_.erase(it);
    else ++it;
  auto mi = mesh_.find(f->second.key);
  if (mi != mesh_.end()) {
    ServeMesh& m = *mi->second;
    for (auto it = m.queue.begin(); it != m.queue.end();)
      if (it->req == id) it = m.queue.erase(it);
      else ++it;
    std::vector<int> cols;
    for (int j = 0; j < (int)m.slot.size(); j++)
      if (m.slot[j].req == id) {
        cols.push_back(j);
        m.slot[j] = ServeSlot();
      }
    if (!cols.empty()) {
      HC(hipSetDevice(device_));
      m.e->blk_release(cols);  // stream ordered after the mesh's last chunk
    }
  }
  if (f->second.t_done < 0) st_.cancelled++;
  if (f->second.dU) result_bytes_ -= f->second.dU->n * sizeof(double);
  req_.erase(f);
}

}  // namespace hfls
