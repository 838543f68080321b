// Grid geometry (src/data/grid.h:11-266): dims include guards; collocated
// node positions pos(i, n, 1) = lower + delta * (n - guard).
#ifndef COFFEE_GRID_H
#define COFFEE_GRID_H

#include "coffee/typedefs.h"

namespace Coffee {

struct Extent {
  int x = 1, y = 1, z = 1;
  int width() const { return x; }
  int height() const { return y; }
  int depth() const { return z; }
  size_t size() const { return (size_t)x * y * z; }
};

class Stagger {
  unsigned char stagger = 0;

 public:
  Stagger() = default;
  Stagger(unsigned char s) : stagger(s) {}
  int operator[](int i) const { return (stagger >> i) & 1UL; }
};

struct Grid {
  int dims[3] = {1, 1, 1};
  int guard[3] = {0, 0, 0};
  Scalar delta[3] = {1, 1, 1};
  Scalar inv_delta[3] = {1, 1, 1};
  Scalar lower[3] = {0, 0, 0};
  Scalar sizes[3] = {0, 0, 0};
  int offset[3] = {0, 0, 0};

  int reduced_dim(int i) const { return dims[i] - 2 * guard[i]; }
  Scalar pos(int i, int n, int stagger) const {
    return lower[i] + delta[i] * (n - guard[i] + (Scalar)(stagger * -0.5 + 0.5));
  }
  int get_idx(int c1, int c2 = 0, int c3 = 0) const {
    return c1 + c2 * dims[0] + c3 * dims[0] * dims[1];
  }
  int size() const { return dims[0] * dims[1] * dims[2]; }
  Extent extent() const { return Extent{dims[0], dims[1], dims[2]}; }
  int dim() const { return (dims[1] <= 1 && dims[2] <= 1) ? 1 : (dims[2] <= 1 ? 2 : 3); }
};

}  // namespace Coffee

#endif
