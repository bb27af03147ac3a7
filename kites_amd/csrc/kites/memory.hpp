// memory.hpp -- RAII, move-only host-pinned and device buffers, bitmaps and
// copies: the thin runtime layer under the graph and engine.
//
// Interface follows the reference memory layer: uva_buffer
// (kites/memory/uva_buffer.cuh:14-52), host_pinned_buffer + host_bitmap
// (host_pinned_buffer.cuh:21-172), device_buffer + device_bitmap
// (device_buffer.cuh:19-160), copyMem / copyMemOnStream (mem_ops.cuh:16-88).
// Differences on MI355X: device bitmaps use 64-bit words so one word is one
// wave's ballot; host bitmaps keep the reference's 32-bit words (their byte
// image equals the 64-bit image on little-endian hosts, so uploads are plain
// byte copies).
#ifndef KITES_MEMORY_HPP_
#define KITES_MEMORY_HPP_

#include <algorithm>
#include <chrono>
#include <cstring>
#include <iostream>

#include "common.hpp"
#include "device.hpp"

namespace kites {

template <typename T>
class uva_buffer {
 protected:
  std::size_t nElems;
  T* ptr;

 public:
  uva_buffer() : nElems(0), ptr(nullptr) {}
  uva_buffer(std::size_t n, T* p) : nElems{n}, ptr{p} {}
  T* get_ptr() const { return ptr; }
  std::size_t size() const { return nElems; }
  std::size_t sizeInBytes() const { return nElems * sizeof(T); }
  virtual void free() = 0;
  virtual ~uva_buffer() {}
};

template <typename T>
class host_pinned_buffer : public uva_buffer<T> {
  void construct(std::size_t n) {
    KITES_HIP_CHECK(hipHostMalloc((void**)&this->ptr, n * sizeof(T),
                                  hipHostMallocPortable));
    this->nElems = n;
  }

 public:
  void free() override {
    if (this->nElems != 0) {
      this->nElems = 0;
      KITES_HIP_CHECK(hipHostFree(this->ptr));
      this->ptr = nullptr;
    }
  }
  host_pinned_buffer() : uva_buffer<T>() {}
  explicit host_pinned_buffer(std::size_t n) {
    if (n > 0) construct(n);
  }
  // Copy-construct from any buffer (host or device): blocking copy.
  host_pinned_buffer(uva_buffer<T> const& in) : host_pinned_buffer{in.size()} {
    if (in.size() != 0)
      KITES_HIP_CHECK(hipMemcpy(this->ptr, in.get_ptr(), in.sizeInBytes(),
                                hipMemcpyDefault));
  }
  host_pinned_buffer& operator=(host_pinned_buffer const&) = delete;
  host_pinned_buffer(host_pinned_buffer&& in) : uva_buffer<T>{in.size(), in.get_ptr()} {
    in.nElems = 0;
    in.ptr = nullptr;
  }
  host_pinned_buffer& operator=(host_pinned_buffer&& in) {
    if (this != &in) {
      free();
      this->nElems = in.nElems;
      this->ptr = in.ptr;
      in.nElems = 0;
      in.ptr = nullptr;
    }
    return *this;
  }
  ~host_pinned_buffer() override { free(); }
  virtual void alloc(std::size_t n) {
    if (this->nElems == 0 && n > 0) construct(n);
  }
  T& at(std::size_t i) {
    if (i >= this->nElems)
      throw std::runtime_error("The referred element does not exist in the buffer.");
    return this->ptr[i];
  }
  T& operator[](std::size_t i) { return this->ptr[i]; }
  T const& operator[](std::size_t i) const { return this->ptr[i]; }
  friend std::ostream& operator<<(std::ostream& os, host_pinned_buffer& b) {
    for (std::size_t i = 0; i < b.size(); ++i) os << b.at(i) << "\n";
    return os;
  }
};

// Host frontier bitmap: one bit per vertex (or vertex group), 32-bit words.
class host_bitmap : public host_pinned_buffer<unsigned int> {
  std::size_t nBits;

 public:
  host_bitmap() : nBits{0} {}
  explicit host_bitmap(std::size_t n)
      : host_pinned_buffer<unsigned int>{(n + 31) / 32}, nBits{n} {}
  void alloc(std::size_t n) override {
    if (n > 0) nBits = n;
    host_pinned_buffer<unsigned int>::alloc((n + 31) / 32);
  }
  void setAt(std::size_t pos) { this->ptr[pos >> 5] |= (1u << (pos & 31)); }
  bool getAt(std::size_t pos) const { return (this->ptr[pos >> 5] >> (pos & 31)) & 1u; }
  void reset() {
    if (this->ptr) std::memset(this->ptr, 0, this->sizeInBytes());
  }
  std::size_t getnBits() const { return nBits; }
  unsigned int count() const {
    unsigned int r = 0;
    for (std::size_t i = 0; i < this->size(); ++i)
      r += __builtin_popcount(this->get_ptr()[i]);
    return r;
  }
};

template <typename T>
class device_buffer : public uva_buffer<T> {
  int residing_device_id;

  void construct(std::size_t n) {
    KITES_HIP_CHECK(hipSetDevice(residing_device_id));
    KITES_HIP_CHECK(hipMalloc((void**)&this->ptr, n * sizeof(T)));
    this->nElems = n;
  }

 public:
  void free() override {
    if (this->nElems != 0) {
      this->nElems = 0;
      KITES_HIP_CHECK(hipFree(this->ptr));
      this->ptr = nullptr;
    }
  }
  device_buffer() : residing_device_id{0} {}
  device_buffer(std::size_t n, int dev) : residing_device_id{dev} {
    if (n > 0) construct(n);
  }
  // Copy-construct from any buffer: blocking copy.
  device_buffer(uva_buffer<T> const& in, int dev = 0) : device_buffer{in.size(), dev} {
    if (in.size() != 0)
      KITES_HIP_CHECK(hipMemcpy(this->ptr, in.get_ptr(), in.sizeInBytes(),
                                hipMemcpyDefault));
  }
  device_buffer& operator=(device_buffer const&) = delete;
  device_buffer(device_buffer&& in)
      : uva_buffer<T>{in.size(), in.get_ptr()}, residing_device_id{in.residing_device_id} {
    in.nElems = 0;
    in.ptr = nullptr;
  }
  device_buffer& operator=(device_buffer&& in) {
    if (this != &in) {
      free();
      this->nElems = in.nElems;
      this->ptr = in.ptr;
      residing_device_id = in.residing_device_id;
      in.nElems = 0;
      in.ptr = nullptr;
    }
    return *this;
  }
  ~device_buffer() override { free(); }
  virtual void alloc(std::size_t n, int dev = 0) {
    residing_device_id = dev;
    if (this->nElems == 0 && n > 0) construct(n);
  }
  int get_residing_device_id() const { return residing_device_id; }
  void reset() {
    KITES_HIP_CHECK(hipSetDevice(residing_device_id));
    if (this->nElems) KITES_HIP_CHECK(hipMemset((void*)this->ptr, 0, this->sizeInBytes()));
  }
  void reset(stream& ss) {
    KITES_HIP_CHECK(hipSetDevice(residing_device_id));
    if (this->nElems)
      KITES_HIP_CHECK(hipMemsetAsync((void*)this->ptr, 0, this->sizeInBytes(), ss.get()));
  }
};

// Device frontier bitmap: 64-bit words, bit i of the bitmap is bit (i & 63)
// of word (i >> 6).  One guard word past the last is always allocated.
class device_bitmap : public device_buffer<unsigned long long> {
  std::size_t nBits;

 public:
  device_bitmap() : nBits{0} {}
  device_bitmap(std::size_t n, int dev = 0)
      : device_buffer<unsigned long long>{(n + 63) / 64 + 1, dev}, nBits{n} {}
  void alloc(std::size_t n, int dev = 0) override {
    if (n > 0) nBits = n;
    device_buffer<unsigned long long>::alloc((n + 63) / 64 + 1, dev);
  }
  std::size_t getnBits() const { return nBits; }
};

// ---- copies (mem_ops.cuh:16-88) -------------------------------------------

template <class T, launch lMode = launch::sync>
double copyMem(uva_buffer<T>& dst, uva_buffer<T>& src, std::size_t length,
               std::size_t dstOffst = 0, std::size_t srcOffst = 0) {
  if (lMode == launch::sync) {
    auto t1 = std::chrono::steady_clock::now();
    KITES_HIP_CHECK(hipMemcpy(dst.get_ptr() + dstOffst, src.get_ptr() + srcOffst,
                              length * sizeof(T), hipMemcpyDefault));
    auto t2 = std::chrono::steady_clock::now();
    return std::chrono::duration<double, std::milli>(t2 - t1).count();
  }
  KITES_HIP_CHECK(hipMemcpyAsync(dst.get_ptr() + dstOffst, src.get_ptr() + srcOffst,
                                 length * sizeof(T), hipMemcpyDefault, 0));
  return 0;
}

template <class T, launch lMode = launch::sync>
double copyMem(uva_buffer<T>& dst, uva_buffer<T>& src, std::size_t dstOffst = 0,
               std::size_t srcOffst = 0) {
  return copyMem<T, lMode>(dst, src, std::min(src.size(), dst.size()), dstOffst,
                           srcOffst);
}

template <class T>
void copyMemOnStream(uva_buffer<T>& dst, uva_buffer<T>& src, std::size_t length,
                     stream& s, std::size_t dstOffst = 0, std::size_t srcOffst = 0) {
  KITES_HIP_CHECK(hipMemcpyAsync(dst.get_ptr() + dstOffst, src.get_ptr() + srcOffst,
                                 length * sizeof(T), hipMemcpyDefault, s.get()));
}

template <class T>
void copyMemOnStream(uva_buffer<T>& dst, uva_buffer<T>& src, stream& s,
                     std::size_t dstOffst = 0, std::size_t srcOffst = 0) {
  copyMemOnStream(dst, src, std::min(src.size(), dst.size()), s, dstOffst, srcOffst);
}

// Host (32-bit words) -> device (64-bit words) frontier upload: byte copy of
// the host image; the device tail past the host bytes is cleared.
inline void copyMemOnStream(device_bitmap& dst, host_bitmap& src, stream& s) {
  std::size_t db = dst.sizeInBytes(), sb = src.sizeInBytes();
  KITES_HIP_CHECK(hipMemsetAsync(dst.get_ptr(), 0, db, s.get()));
  if (sb) KITES_HIP_CHECK(hipMemcpyAsync(dst.get_ptr(), src.get_ptr(), std::min(db, sb),
                                         hipMemcpyHostToDevice, s.get()));
}

}  // namespace kites

#endif  // KITES_MEMORY_HPP_
