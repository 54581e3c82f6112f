// Saved set-up state of a mesh (state.hpp): the engine's fields, the file and
// pinned-host-memory streams, and the save/restore entry points.
//
// The reference keeps nothing between runs: every Driver run assembles and
// factorises again (Driver.cpp:20-38); its closest artefact is the MatrixMarket
// dump of the cluster matrices (Matrix.cpp:556-632), which is for inspection and
// is never read back.
#include "state.hpp"

#include <algorithm>
#include <chrono>
#include <cstring>
#include <stdexcept>
#include <vector>

namespace hfls {

static const char kMagic[8] = {'H', 'F', 'L', 'S', 'M', 'E', 'S', 'H'};
static const char kEnd[8] = {'H', 'F', 'L', 'S', '_', 'E', 'N', 'D'};
static const uint32_t kVersion = 1;

// ------------------------------------------------------------------ streams

// pinned host memory: a counting pass (mem null), then the copy
class MemStream final : public StateStream {
 public:
  MemStream(char* mem, size_t cap, bool load) : mem_(mem), cap_(cap), load_(load) {}
  bool loading() const override { return load_; }
  void host(void* p, size_t n) override {
    take(n);
    if (load_) std::memcpy(p, mem_ + off_ - n, n);
    else if (mem_) std::memcpy(mem_ + off_ - n, p, n);
  }
  void device(void* d, size_t n, hipStream_t st) override {
    take(n);
    if (load_) HC(hipMemcpyAsync(d, mem_ + off_ - n, n, hipMemcpyHostToDevice, st));
    else if (mem_) HC(hipMemcpyAsync(mem_ + off_ - n, d, n, hipMemcpyDeviceToHost, st));
  }
  size_t used() const { return off_; }

 private:
  void take(size_t n) {
    if (n > cap_ - off_) throw std::runtime_error("saved mesh: read past the end of the saved state");
    off_ += n;
  }
  char* mem_;
  size_t cap_, off_ = 0;
  bool load_;
};

// a file, device bytes staged through two pinned 64 MB buffers (a load overlaps the
// read of one chunk with the host-to-device copy of the previous one); the buffers
// are allocated at the first device transfer, so a header is read without the GPU
class FileStream final : public StateStream {
 public:
  static constexpr size_t kChunk = (size_t)64 << 20;
  FileStream(const std::string& path, bool load) : load_(load) {
    f_ = std::fopen(path.c_str(), load ? "rb" : "wb");
    if (!f_) throw std::runtime_error("cannot open " + path + (load ? " for reading" : " for writing"));
  }
  ~FileStream() { release(); }
  bool loading() const override { return load_; }
  void host(void* p, size_t n) override {
    if (!n) return;
    if (load_) {
      if (std::fread(p, 1, n, f_) != n) throw std::runtime_error("saved mesh file is truncated");
    } else if (std::fwrite(p, 1, n, f_) != n) {
      throw std::runtime_error("writing the saved mesh file failed");
    }
  }
  void device(void* d, size_t n, hipStream_t st) override {
    for (int i = 0; i < 2; i++) {
      if (!stage_[i]) HC(hipHostMalloc((void**)&stage_[i], kChunk, hipHostMallocDefault));
      if (!ev_[i]) HC(hipEventCreateWithFlags(&ev_[i], hipEventDisableTiming));
    }
    char* dc = (char*)d;
    for (size_t o = 0, i = 0; o < n; o += kChunk, i++) {
      const size_t c = std::min(kChunk, n - o);
      const int b = (int)(i & 1);
      if (load_) {
        HC(hipEventSynchronize(ev_[b]));  // the copy that last read this buffer is done
        host(stage_[b], c);
        HC(hipMemcpyAsync(dc + o, stage_[b], c, hipMemcpyHostToDevice, st));
        HC(hipEventRecord(ev_[b], st));
      } else {
        HC(hipMemcpyAsync(stage_[0], dc + o, c, hipMemcpyDeviceToHost, st));
        HC(hipStreamSynchronize(st));
        host(stage_[0], c);
      }
    }
  }
  // wait for staged copies; a save flushes and closes the file, reporting errors
  void finish() {
    for (int i = 0; i < 2; i++)
      if (ev_[i]) HC(hipEventSynchronize(ev_[i]));
    if (f_ && !load_) {
      const bool bad = std::fflush(f_) != 0 || std::ferror(f_);
      const bool closed = std::fclose(f_) == 0;
      f_ = nullptr;
      if (bad || !closed) throw std::runtime_error("writing the saved mesh file failed");
    }
  }

 private:
  void release() {
    for (int i = 0; i < 2; i++) {
      if (ev_[i]) {
        (void)hipEventSynchronize(ev_[i]);
        (void)hipEventDestroy(ev_[i]);
        ev_[i] = nullptr;
      }
      if (stage_[i]) (void)hipHostFree(stage_[i]);
      stage_[i] = nullptr;
    }
    if (f_) std::fclose(f_);
    f_ = nullptr;
  }
  bool load_;
  FILE* f_ = nullptr;
  char* stage_[2] = {nullptr, nullptr};
  hipEvent_t ev_[2] = {nullptr, nullptr};
};

// ------------------------------------------------------------------ framing

static void io_str(StateStream& s, std::string& x) {
  uint32_t n = (uint32_t)x.size();
  s.host(&n, 4);
  if (s.loading()) {
    if (n > (1u << 20)) throw std::runtime_error("saved mesh: option string too long");
    x.assign(n, '\0');
  }
  if (n) s.host(&x[0], n);
}

static void io_header(StateStream& s, Options& o) {
  char m[8];
  std::memcpy(m, kMagic, 8);
  s.host(m, 8);
  if (s.loading() && std::memcmp(m, kMagic, 8) != 0) throw std::runtime_error("not a saved HTFETI mesh (bad magic)");
  uint32_t v = kVersion;
  s.host(&v, 4);
  if (s.loading() && v != kVersion)
    throw std::runtime_error("saved mesh format version " + std::to_string(v) + ", this library reads version " +
                             std::to_string(kVersion));
  uint32_t n = (uint32_t)o.size();
  s.host(&n, 4);
  if (s.loading()) {
    if (n > 4096) throw std::runtime_error("saved mesh: too many options");
    o.clear();
    for (uint32_t i = 0; i < n; i++) {
      std::string k, val;
      io_str(s, k);
      io_str(s, val);
      o[k] = val;
    }
  } else {
    for (auto& kv : o) {
      std::string k = kv.first, val = kv.second;
      io_str(s, k);
      io_str(s, val);
    }
  }
}

static void io_end(StateStream& s) {
  char m[8];
  std::memcpy(m, kEnd, 8);
  s.host(m, 8);
  if (s.loading() && std::memcmp(m, kEnd, 8) != 0) throw std::runtime_error("saved mesh: missing end marker");
}

// a field's frame: name, element size, element count (read and checked on a load)
static void frame(StateStream& s, const char* name, uint32_t elem, uint64_t& count) {
  char nm[16] = {0};
  std::strncpy(nm, name, 15);
  if (s.loading()) {
    char got[16];
    uint32_t e = 0;
    s.host(got, 16);
    s.host(&e, 4);
    s.host(&count, 8);
    if (std::memcmp(got, nm, 16) != 0 || e != elem)
      throw std::runtime_error(std::string("saved mesh: expected field ") + name);
  } else {
    s.host(nm, 16);
    s.host(&elem, 4);
    s.host(&count, 8);
  }
}

static void check_count(const char* name, uint64_t got, uint64_t want) {
  if (got != want)
    throw std::runtime_error(std::string("saved mesh: field ") + name + " has " + std::to_string(got) +
                             " entries, the mesh of its options needs " + std::to_string(want));
}

template <class T>
static void io_vec(StateStream& s, const char* name, std::vector<T>& v, uint64_t want) {
  uint64_t n = v.size();
  if (!s.loading()) check_count(name, n, want);
  frame(s, name, sizeof(T), n);
  if (s.loading()) {
    check_count(name, n, want);
    v.resize(n);
  }
  if (n) s.host(v.data(), n * sizeof(T));
}

template <class T>
static void io_dvec(StateStream& s, const char* name, DVec<T>& v, uint64_t want, hipStream_t st) {
  uint64_t n = v.n;
  if (!s.loading()) check_count(name, n, want);
  frame(s, name, sizeof(T), n);
  if (s.loading()) {
    check_count(name, n, want);
    if (v.n != n || !v.p) v.alloc(n);
  }
  if (n) s.device(v.p, n * sizeof(T), st);
}

// ------------------------------------------------------------------ engine state

void Engine::state_fields(StateStream& s, hipStream_t cs) {
  const int N = P->nsub, ncl = P->ncl;
  // the layout rebuilt from the options must be the layout the state was saved with
  int64_t fp[10] = {N, ncl, P->nsub_cl, P->nPoints, P->nlam, off_n[N], NB, NR, off_L[N], off_P[N]};
  int64_t got[10];
  std::memcpy(got, fp, sizeof fp);
  uint64_t c = 10;
  frame(s, "fingerprint", 8, c);
  check_count("fingerprint", c, 10);
  s.host(got, sizeof got);
  if (s.loading() && std::memcmp(got, fp, sizeof fp) != 0)
    throw std::runtime_error("saved mesh does not match the mesh its options describe");
  SetupInfo si = sinfo;
  si.restore_ms = 0.0;  // the record describes the set-up; a load times itself
  c = 1;
  frame(s, "setupinfo", sizeof(SetupInfo), c);
  check_count("setupinfo", c, 1);
  s.host(&si, sizeof si);
  // kernel detection and gluing results (host), then the constraint layout
  io_vec(s, "defect", h_def, N);
  io_vec(s, "fixpos", h_fix, 6 * (uint64_t)N);
  io_vec(s, "nlc", h_nlc, N);
  io_vec(s, "nlcl", h_nlcl, ncl);
  if (s.loading()) {
    for (int d = 0; d < N; d++) {
      if (h_def[d] != 6) throw std::runtime_error("saved mesh: kernel defect other than 6");
      for (int k = 0; k < 6; k++)
        if (h_fix[6 * d + k] < 0 || h_fix[6 * d + k] >= h_ns[d])
          throw std::runtime_error("saved mesh: fixing DOF outside its subdomain's interface");
    }
    coarse_layout();  // checks the counts, uploads the offsets
  }
  int64_t ngc = 0;
  for (int x : h_nlcl) ngc += x;
  io_vec(s, "lcmap", h_lcmap, off_lc[N]);
  io_vec(s, "gcsrc", h_gcsrc, std::max<int64_t>(2 * ngc, 1));
  if (s.loading()) {
    for (int d = 0; d < N; d++) {
      const int cl = h_cl[d];
      for (int64_t q = off_lc[d]; q < off_lc[d + 1]; q++)
        if (h_lcmap[q] < h_la_off[cl] || h_lcmap[q] >= h_la_off[cl] + h_nlcl[cl])
          throw std::runtime_error("saved mesh: constraint row outside its cluster");
    }
    for (size_t q = 0; q < h_gcsrc.size(); q++) {
      const int v = h_gcsrc[q];
      const bool first = (q % 2 == 0) && ngc > 0;
      if (v >= off_lc[N] || v < (first ? 0 : -1)) throw std::runtime_error("saved mesh: constraint source out of range");
    }
    d_fixpos.upload(h_fix, st);
    d_defect.upload(h_def, st);
    d_lcmap.upload(h_lcmap, st);
    d_gc_src.upload(h_gcsrc, st);
    sinfo = si;
  }
  // device products of the factorisation and of the coarse problems
  io_dvec(s, "kdiag", d_kdiag, off_n[N], cs);
  io_dvec(s, "f", d_f, off_n[N], cs);
  io_dvec(s, "Lt", d_Lt, std::max<int64_t>(off_L[N], 1), cs);
  io_dvec(s, "R", d_R, off_R[N], cs);
  io_dvec(s, "Rb", d_Rb, off_Rb[N], cs);
  io_dvec(s, "brv", d_brv, d_brv.n, cs);
  io_dvec(s, "rbv", d_rbv, d_rbv.n, cs);
  io_dvec(s, "BcT", d_BcT, std::max<int64_t>(off_C[N], 1), cs);
  io_dvec(s, "C", d_C, off_C[N], cs);
  io_dvec(s, "Kt", d_Kt, off_P[N], cs);
  io_dvec(s, "St", d_St, off_P[N], cs);
  io_dvec(s, "H", d_H, 36 * (uint64_t)N, cs);
  io_dvec(s, "Ainv", d_Ainv, h_Aoff[ncl], cs);
  io_dvec(s, "invGtG", d_invGtG, 36 * (uint64_t)ncl * ncl, cs);
}

void Engine::save_state(StateStream& s, hipStream_t copy) {
  if (s.loading()) throw std::logic_error("save_state on a loading stream");
  if (!d_Kt.p) throw std::runtime_error("the mesh is not set up");
  const hipStream_t cs = copy ? copy : st;
  state_fields(s, cs);
  HC(hipStreamSynchronize(cs));
}

void Engine::load_state(StateStream& s) {
  if (!s.loading()) throw std::logic_error("load_state on a saving stream");
  if (d_Kt.p) throw std::runtime_error("the mesh is already set up");
  const auto t0 = std::chrono::steady_clock::now();
  d_err.alloc(1);
  HC(hipMemsetAsync(d_err.p, 0, sizeof(int), st));
  build_meta();
  state_fields(s, st);
  alloc_work();
  HC(hipStreamSynchronize(st));
  sinfo.restore_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

// ------------------------------------------------------------------ entry points

HostMesh::~HostMesh() {
  if (mem) (void)hipHostFree(mem);
}

size_t saved_bytes(Engine& e, hipStream_t copy) {
  MemStream c(nullptr, SIZE_MAX, false);
  Options o = e.problem().o;
  io_header(c, o);
  e.save_state(c, copy);
  io_end(c);
  return c.used();
}

std::unique_ptr<HostMesh> save_mesh_host(Engine& e, hipStream_t copy) {
  const size_t n = saved_bytes(e, copy);
  std::unique_ptr<HostMesh> h(new HostMesh);
  HC(hipHostMalloc((void**)&h->mem, n, hipHostMallocDefault));
  h->bytes = n;
  h->o = e.problem().o;
  MemStream m(h->mem, n, false);
  Options o = h->o;
  io_header(m, o);
  e.save_state(m, copy);
  io_end(m);
  if (m.used() != n) throw std::logic_error("saved mesh: size differs from the counting pass");
  return h;
}

std::unique_ptr<Engine> load_mesh_host(const HostMesh& h, int device) {
  MemStream m(h.mem, h.bytes, true);
  Options o;
  io_header(m, o);
  std::unique_ptr<Engine> e(new Engine(o, device));
  try {
    e->load_state(m);
    io_end(m);
  } catch (...) {
    (void)hipStreamSynchronize(e->stream());  // copies into the engine's buffers end first
    throw;
  }
  return e;
}

void save_mesh_file(Engine& e, const std::string& path) {
  // written next to the target and renamed when complete: a reader never sees a
  // partial file under the target's name
  const std::string tmp = path + ".part";
  try {
    FileStream f(tmp, false);
    Options o = e.problem().o;
    io_header(f, o);
    e.save_state(f);
    io_end(f);
    f.finish();
  } catch (...) {
    std::remove(tmp.c_str());
    throw;
  }
  if (std::rename(tmp.c_str(), path.c_str()) != 0) {
    std::remove(tmp.c_str());
    throw std::runtime_error("cannot rename " + tmp + " to " + path);
  }
}

std::unique_ptr<Engine> load_mesh_file(const std::string& path, int device) {
  FileStream f(path, true);
  Options o;
  io_header(f, o);
  std::unique_ptr<Engine> e(new Engine(o, device));
  try {
    e->load_state(f);
    io_end(f);
    f.finish();
  } catch (...) {
    (void)hipStreamSynchronize(e->stream());  // staged copies into the engine's buffers end first
    throw;
  }
  return e;
}

}  // namespace hfls
