// Saved set-up state of a mesh: the on-disk mesh cache (one file per mesh, read
// back in another process) and the serving loop's cache of evicted meshes in
// pinned host memory.
//
// Layout of a saved mesh (file or host memory, little-endian, no padding):
//   "HFLSMESH", u32 format version, u32 option count, (u32 length, bytes) per key and
//   per value of the mesh options;
//   i64[10] fingerprint: nsub, ncl, nsub_cl, nPoints, nlam, packed local length, b
//   positions, interior rows, L_rr tile doubles, K+_bb tile doubles;
//   fields in a fixed order, each: char[16] name, u32 element size, u64 count, data;
//   "HFLS_END".
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <memory>
#include <string>

#include "engine.hpp"

namespace hfls {

// A sequential byte stream in one direction.  host() moves host bytes; device()
// moves device bytes between device memory and the stream's storage, enqueued on
// stream st (finished by the time the save or load that issued it returns).
class StateStream {
 public:
  virtual ~StateStream() = default;
  virtual bool loading() const = 0;
  virtual void host(void* p, size_t n) = 0;
  virtual void device(void* d, size_t n, hipStream_t st) = 0;
};

// A saved mesh in pinned (page-locked) host memory, restored at the host link's
// full copy rate.
struct HostMesh {
  Options o;
  size_t bytes = 0;
  char* mem = nullptr;
  HostMesh() = default;
  HostMesh(const HostMesh&) = delete;
  HostMesh& operator=(const HostMesh&) = delete;
  ~HostMesh();
};

// bytes a saved copy of the set-up mesh takes (copy: as for save_mesh_host)
size_t saved_bytes(Engine& e, hipStream_t copy = nullptr);
// save / restore through pinned host memory (throws if it cannot be allocated);
// copy: the stream of the save's device-to-host copies (null: the engine's)
std::unique_ptr<HostMesh> save_mesh_host(Engine& e, hipStream_t copy = nullptr);
std::unique_ptr<Engine> load_mesh_host(const HostMesh& h, int device);
// save / restore through a file
void save_mesh_file(Engine& e, const std::string& path);
std::unique_ptr<Engine> load_mesh_file(const std::string& path, int device);

}  // namespace hfls
