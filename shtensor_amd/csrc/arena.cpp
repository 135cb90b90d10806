#include "arena.hpp"

#include <cstring>

#include "common.hpp"

namespace sht {

void Arena::init(void* base, size_t bytes, bool host)
{
  std::lock_guard<std::mutex> lk(mu_);
  host_ = host;
  base_ = static_cast<char*>(base);
  total_ = bytes / kAlign * kAlign;
  chunks_.clear();
  if (total_) chunks_[0] = Chunk{total_, true};
}

void* Arena::allocate(size_t bytes)
{
  std::lock_guard<std::mutex> lk(mu_);
  const size_t need = round_up(bytes ? bytes : 1);
  for (auto it = chunks_.begin(); it != chunks_.end(); ++it) {
    Chunk& c = it->second;
    if (!c.free || c.size < need) continue;
    const size_t rest = c.size - need;
    if (rest >= kAlign) {
      chunks_[it->first + need] = Chunk{rest, true};
      c.size = need;
    }
    c.free = false;
    return base_ + it->first;
  }
  throw Error(SHT_ERR_OOM, "Out of memory");
}

std::map<size_t, Arena::Chunk>::iterator Arena::find_used(const void* p)
{
  const char* cp = static_cast<const char*>(p);
  SHT_REQUIRE(cp >= base_ && cp < base_ + total_, SHT_ERR_INVALID, "pointer not owned by the arena");
  auto it = chunks_.find(static_cast<size_t>(cp - base_));
  SHT_REQUIRE(it != chunks_.end() && !it->second.free, SHT_ERR_INVALID, "pointer is not a live arena allocation");
  return it;
}

void Arena::release(void* p)
{
  if (!p) return;
  std::lock_guard<std::mutex> lk(mu_);
  auto it = find_used(p);
  it->second.free = true;
  auto nx = std::next(it);
  if (nx != chunks_.end() && nx->second.free) {
    it->second.size += nx->second.size;
    chunks_.erase(nx);
  }
  if (it != chunks_.begin()) {
    auto pv = std::prev(it);
    if (pv->second.free) {
      pv->second.size += it->second.size;
      chunks_.erase(it);
    }
  }
}

void* Arena::reallocate(void* p, size_t bytes)
{
  if (!p) return allocate(bytes);
  const size_t need = round_up(bytes ? bytes : 1);
  {
    std::lock_guard<std::mutex> lk(mu_);
    auto it = find_used(p);
    Chunk& c = it->second;
    if (need == c.size) return p;
    auto nx = std::next(it);
    if (need < c.size) {
      // shrink in place; the tail joins a free successor or becomes one
      const size_t tail = c.size - need;
      c.size = need;
      if (nx != chunks_.end() && nx->second.free) {
        const size_t sz = nx->second.size + tail;
        chunks_.erase(nx);
        chunks_[it->first + need] = Chunk{sz, true};
      } else {
        chunks_[it->first + need] = Chunk{tail, true};
      }
      return p;
    }
    const size_t grow = need - c.size;
    if (nx != chunks_.end() && nx->second.free && nx->second.size >= grow) {
      const size_t left = nx->second.size - grow;
      chunks_.erase(nx);
      c.size = need;
      if (left) chunks_[it->first + need] = Chunk{left, true};
      return p;
    }
  }
  const size_t old = size_of(p);
  void* q = allocate(bytes);
  if (host_) std::memcpy(q, p, old);
  else SHT_HIP(hipMemcpy(q, p, old, hipMemcpyDeviceToDevice));
  release(p);
  return q;
}

size_t Arena::size_of(const void* p) const
{
  std::lock_guard<std::mutex> lk(mu_);
  const char* cp = static_cast<const char*>(p);
  SHT_REQUIRE(cp >= base_ && cp < base_ + total_, SHT_ERR_INVALID, "pointer not owned by the arena");
  auto it = chunks_.find(static_cast<size_t>(cp - base_));
  SHT_REQUIRE(it != chunks_.end() && !it->second.free, SHT_ERR_INVALID, "pointer is not a live arena allocation");
  return it->second.size;
}

bool Arena::owns(const void* p) const
{
  const char* cp = static_cast<const char*>(p);
  return cp >= base_ && cp < base_ + total_;
}

size_t Arena::free_bytes() const
{
  std::lock_guard<std::mutex> lk(mu_);
  size_t f = 0;
  for (const auto& kv : chunks_) if (kv.second.free) f += kv.second.size;
  return f;
}

size_t Arena::largest_free() const
{
  std::lock_guard<std::mutex> lk(mu_);
  size_t f = 0;
  for (const auto& kv : chunks_) if (kv.second.free && kv.second.size > f) f = kv.second.size;
  return f;
}

}  // namespace sht
