// multi_array<T> (src/data/multi_array.h:14-116): Fortran-order 3D array with
// a device buffer and a host mirror; move-only; explicit sync_to_host /
// sync_to_device.  Device memory comes from hipMalloc; assign_dev uses the
// library's fill kernel (coffee_fill).
#ifndef COFFEE_MULTI_ARRAY_H
#define COFFEE_MULTI_ARRAY_H

#include <hip/hip_runtime_api.h>

#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "coffee/grid.h"
#include "coffee_hip.h"

namespace Coffee {

inline void hip_check(hipError_t e, const char* what) {
  if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

inline void coffee_check(int rc, const char* what, const coffee_ctx* ctx = nullptr) {
  if (rc != 0)
    throw std::runtime_error(std::string(what) + ": " + coffee_last_error(ctx));
}

template <typename T>
class multi_array {
 public:
  multi_array() = default;
  explicit multi_array(int width, int height = 1, int depth = 1) { alloc(Extent{width, height, depth}); }
  explicit multi_array(const Extent& e) { alloc(e); }
  multi_array(const multi_array&) = delete;
  multi_array& operator=(const multi_array&) = delete;
  multi_array(multi_array&& o) noexcept { swap(o); }
  multi_array& operator=(multi_array&& o) noexcept {
    if (this != &o) {
      free_mem();
      swap(o);
    }
    return *this;
  }
  ~multi_array() { free_mem(); }

  T& operator()(int x, int y = 0, int z = 0) { return m_host[x + (y + (size_t)z * m_ext.y) * m_ext.x]; }
  const T& operator()(int x, int y = 0, int z = 0) const {
    return m_host[x + (y + (size_t)z * m_ext.y) * m_ext.x];
  }
  T& operator[](size_t n) { return m_host[n]; }
  const T& operator[](size_t n) const { return m_host[n]; }

  void copy_from(const multi_array& o) {
    if (o.size() != size()) throw std::range_error("multi_array::copy_from size mismatch");
    std::memcpy(m_host.data(), o.m_host.data(), size() * sizeof(T));
    hip_check(hipMemcpy(m_dev, o.m_dev, size() * sizeof(T), hipMemcpyDeviceToDevice),
              "multi_array::copy_from");
  }
  void assign(const T& v) { std::fill(m_host.begin(), m_host.end(), v); }
  void assign_dev(const T& v) {
    coffee_check(coffee_fill(m_dev, size(), (int)sizeof(T), (double)v, nullptr), "assign_dev");
  }
  void sync_to_device() {
    hip_check(hipMemcpy(m_dev, m_host.data(), size() * sizeof(T), hipMemcpyHostToDevice),
              "multi_array::sync_to_device");
  }
  void sync_to_host() {
    hip_check(hipMemcpy(m_host.data(), m_dev, size() * sizeof(T), hipMemcpyDeviceToHost),
              "multi_array::sync_to_host");
  }
  int width() const { return m_ext.x; }
  int height() const { return m_ext.y; }
  int depth() const { return m_ext.z; }
  size_t size() const { return m_ext.size() * (m_dev != nullptr); }
  const Extent& extent() const { return m_ext; }
  T* host_ptr() { return m_host.data(); }
  const T* host_ptr() const { return m_host.data(); }
  T* dev_ptr() { return m_dev; }
  const T* dev_ptr() const { return m_dev; }

 private:
  void alloc(const Extent& e) {
    m_ext = e;
    m_host.assign(e.size(), T(0));
    hip_check(hipMalloc(&m_dev, e.size() * sizeof(T)), "multi_array hipMalloc");
    hip_check(hipMemset(m_dev, 0, e.size() * sizeof(T)), "multi_array hipMemset");
  }
  void free_mem() {
    if (m_dev) (void)hipFree(m_dev);
    m_dev = nullptr;
    m_host.clear();
    m_ext = Extent{};
  }
  void swap(multi_array& o) {
    std::swap(m_ext, o.m_ext);
    std::swap(m_dev, o.m_dev);
    m_host.swap(o.m_host);
  }
  Extent m_ext{};
  T* m_dev = nullptr;
  std::vector<T> m_host;
};

}  // namespace Coffee

#endif
