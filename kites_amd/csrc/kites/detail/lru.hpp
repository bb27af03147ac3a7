// lru.hpp -- a small least-recently-used map from byte-string keys to
// values, with a release callback for every value it drops.  Pure host
// C++ (no HIP): the launch cache keeps its captured hipGraphs in one, and
// the host-only sanitizer tests exercise it directly.
#ifndef KITES_DETAIL_LRU_HPP_
#define KITES_DETAIL_LRU_HPP_

#include <cstddef>
#include <cstdint>
#include <utility>
#include <vector>

namespace kites {
namespace detail {

template <class V>
class byte_lru {
 public:
  using key_type = std::vector<unsigned char>;

  explicit byte_lru(std::size_t capacity) : cap_(capacity ? capacity : 1) {}
  byte_lru(byte_lru const&) = delete;
  byte_lru& operator=(byte_lru const&) = delete;

  // The value stored under `key` (and marks it most recently used), or null.
  V* find(key_type const& key) {
    for (entry& e : ents_) {
      if (e.key == key) {
        e.used = ++tick_;
        return &e.value;
      }
    }
    return nullptr;
  }

  // Stores `value` under `key`, which must not be present.  When the map is
  // full, the least recently used entry is passed to `release` and dropped
  // first; the entry just inserted is the most recently used, so it is
  // never the one dropped by the next insert.  Returns the stored value.
  template <class Release>
  V& insert(key_type key, V value, Release&& release) {
    if (ents_.size() >= cap_) {
      std::size_t lru = 0;
      for (std::size_t i = 1; i < ents_.size(); ++i)
        if (ents_[i].used < ents_[lru].used) lru = i;
      release(ents_[lru].value);
      ents_.erase(ents_.begin() + static_cast<std::ptrdiff_t>(lru));
    }
    ents_.push_back(entry{std::move(key), std::move(value), ++tick_});
    return ents_.back().value;
  }

  // Passes every value to `release` and empties the map.
  template <class Release>
  void clear(Release&& release) {
    for (entry& e : ents_) release(e.value);
    ents_.clear();
  }

  std::size_t size() const { return ents_.size(); }
  std::size_t capacity() const { return cap_; }

 private:
  struct entry {
    key_type key;
    V value;
    std::uint64_t used;
  };
  std::size_t cap_;
  std::vector<entry> ents_;
  std::uint64_t tick_ = 0;
};

}  // namespace detail
}  // namespace kites

#endif  // KITES_DETAIL_LRU_HPP_
