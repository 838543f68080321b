// vector_field<T> (src/data/fields.h:11-132, fields_impl.hpp:8-30): three
// multi_arrays on one grid with a stagger per component.
#ifndef COFFEE_FIELDS_H
#define COFFEE_FIELDS_H

#include "coffee/multi_array.h"

namespace Coffee {

template <typename T>
class vector_field {
 public:
  vector_field() = default;
  explicit vector_field(const Grid& g) : m_grid(&g) {
    for (int i = 0; i < 3; i++) m_data[i] = multi_array<T>(g.extent());
    set_default_stagger();
  }
  vector_field(vector_field&&) = default;
  vector_field& operator=(vector_field&&) = default;

  // all components to zero (on the device, unlike the reference CPU build
  // where assign_dev is a no-op)
  void initialize() { assign(T(0)); }
  template <typename Func>
  void initialize(int n, const Func& f) {
    auto& a = m_data[n];
    for (int k = 0; k < m_grid->dims[2]; k++) {
      double x3 = m_grid->pos(2, k, m_stagger[n][2]);
      for (int j = 0; j < m_grid->dims[1]; j++) {
        double x2 = m_grid->pos(1, j, m_stagger[n][1]);
        for (int i = 0; i < m_grid->dims[0]; i++) {
          double x1 = m_grid->pos(0, i, m_stagger[n][0]);
          a(i, j, k) = f(x1, x2, x3);
        }
      }
    }
    a.sync_to_device();
  }
  template <typename Func>
  void initialize(const Func& f) {
    for (int n = 0; n < 3; n++)
      initialize(n, [&f, n](T x1, T x2, T x3) { return f(n, x1, x2, x3); });
  }
  void assign(T v) {
    for (auto& a : m_data) a.assign_dev(v);
  }
  void assign(T v, int n) { m_data[n].assign_dev(v); }
  void copy_from(const vector_field& o) {
    for (int i = 0; i < 3; i++) m_data[i].copy_from(o.m_data[i]);
  }
  void copy_stagger(const vector_field& o) {
    for (int i = 0; i < 3; i++) m_stagger[i] = o.m_stagger[i];
  }
  void sync_to_host() {
    for (auto& a : m_data) a.sync_to_host();
  }
  void sync_to_device() {
    for (auto& a : m_data) a.sync_to_device();
  }
  multi_array<T>& data(int n) { return m_data[n]; }
  const multi_array<T>& data(int n) const { return m_data[n]; }
  T* dev_ptr(int n) { return m_data[n].dev_ptr(); }
  T* host_ptr(int n) { return m_data[n].host_ptr(); }
  Stagger stagger(int n) const { return m_stagger[n]; }
  void set_stagger(int n, Stagger s) { m_stagger[n] = s; }
  void set_stagger(Stagger s[]) {
    for (int i = 0; i < 3; i++) m_stagger[i] = s[i];
  }
  const Grid& grid() const { return *m_grid; }

 private:
  void set_default_stagger() {
    m_stagger[0] = Stagger(0b110);
    m_stagger[1] = Stagger(0b101);
    m_stagger[2] = Stagger(0b011);
  }
  const Grid* m_grid = nullptr;
  multi_array<T> m_data[3];
  Stagger m_stagger[3];
};

}  // namespace Coffee

#endif
