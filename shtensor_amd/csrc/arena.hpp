// HBM arena: one hipMalloc'd region carved by a first-fit allocator.
//
// Semantics of the reference's MemoryPool (src/Shmem/MemoryPool.h:126-354):
// first fit in address order, split when the remainder is worth a chunk,
// coalesce with both neighbours on free, in-place shrink and in-place grow
// into a free successor, otherwise move.  Differences, by design for a GPU
// arena: chunk headers live out of band on the host (device memory holds only
// payload, so a header can never be overwritten by a kernel), alignment is
// 256 bytes instead of alignof(max_align_t), and the backing store is device
// memory rather than an MPI window (MemoryPool.cpp:68).
#pragma once

#include <cstddef>
#include <cstdint>
#include <map>
#include <mutex>

namespace sht {

class Arena {
 public:
  static constexpr size_t kAlign = 256;

  Arena() = default;
  Arena(const Arena&) = delete;
  Arena& operator=(const Arena&) = delete;
  void init(void* base, size_t bytes, bool host = false);

  void* allocate(size_t bytes);               // throws Error(SHT_ERR_OOM, "Out of memory")
  void* reallocate(void* p, size_t bytes);    // may move (device copy)
  void release(void* p);
  size_t size_of(const void* p) const;
  bool owns(const void* p) const;

  size_t total() const { return total_; }
  size_t free_bytes() const;
  size_t largest_free() const;
  char* base() const { return base_; }

 private:
  struct Chunk {
    size_t size;
    bool free;
  };
  static size_t round_up(size_t b) { return (b + kAlign - 1) / kAlign * kAlign; }
  std::map<size_t, Chunk>::iterator find_used(const void* p);

  char* base_ = nullptr;
  bool host_ = false;
  size_t total_ = 0;
  std::map<size_t, Chunk> chunks_;  // offset -> chunk, covering [0, total_)
  mutable std::mutex mu_;
};

}  // namespace sht
