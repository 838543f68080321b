// Device side of the field output (include/coffee_io.h): the sampling kernel
// behind data_exporter::add_grid_output.  Only the sampled float array leaves
// the GPU.
#include <hip/hip_runtime.h>

#include <string>

#include "coffee_io.h"
#include "ctx.hpp"

using namespace coffee;

namespace {

struct SampleGeom {
  int ex[3];      // sampled extent
  int step[3];    // sampling stride per axis
  int guard[3];
  long sy, sz;    // array strides (Fortran order)
  int mi, mj, mk; // minus-side offsets from the stagger (-1 or 0)
};

// Interpolation to the cell centre as interpolation.h:9-34 (3D) and :36-50
// (2D) compute it with out = 0b111: the plus-side offset is 0 on every axis,
// the minus-side offset is -1 where the stagger bit is 0; pairs are averaged
// innermost along z, then y, then x, in the Scalar type.
template <typename T, bool TWO_D>
__global__ void __launch_bounds__(256) downsample_kernel(const T* __restrict__ f,
                                                          float* __restrict__ out, SampleGeom g) {
  const long n = (long)g.ex[0] * g.ex[1] * g.ex[2];
  const T h = T(0.5);
  for (long o = blockIdx.x * (long)blockDim.x + threadIdx.x; o < n;
       o += (long)gridDim.x * blockDim.x) {
    const int i = (int)(o % g.ex[0]);
    const int j = (int)((o / g.ex[0]) % g.ex[1]);
    const int k = (int)(o / ((long)g.ex[0] * g.ex[1]));
    const long c = (long)(i * g.step[0] + g.guard[0]) + (long)(j * g.step[1] + g.guard[1]) * g.sy +
                   (long)(k * g.step[2] + g.guard[2]) * g.sz;
    const long di = g.mi, dj = g.mj * g.sy;
    T v;
    if (TWO_D) {
      const T f1 = h * (f[c] + f[c + dj]);
      const T f0 = h * (f[c + di] + f[c + di + dj]);
      v = h * (f1 + f0);
    } else {
      const long dk = g.mk * g.sz;
      const T f11 = h * (f[c + dk] + f[c]);
      const T f10 = h * (f[c + dj + dk] + f[c + dj]);
      const T f01 = h * (f[c + di + dk] + f[c + di]);
      const T f00 = h * (f[c + di + dj + dk] + f[c + di + dj]);
      const T f1 = h * (f11 + f10);
      const T f0 = h * (f01 + f00);
      v = h * (f1 + f0);
    }
    out[o] = (float)v;
  }
}

}  // namespace

extern "C" int coffee_downsample(const coffee_grid* grid, const void* src, int scalar_bytes,
                                 int stagger, int d, float* dst, void* stream) {
  if (!grid || !src || !dst) return fail(nullptr, COFFEE_ERR_INVALID, "null argument");
  if (scalar_bytes != 4 && scalar_bytes != 8)
    return fail(nullptr, COFFEE_ERR_INVALID, "scalar_bytes must be 4 or 8");
  if (d < 1) return fail(nullptr, COFFEE_ERR_INVALID, "downsample factor must be >= 1");
  if (stagger < 0 || stagger > 7) return fail(nullptr, COFFEE_ERR_INVALID, "stagger must be 0..7");
  SampleGeom g{};
  for (int a = 0; a < 3; a++) {
    const int n = grid->dims[a] - 2 * grid->guard[a];
    if (grid->dims[a] < 1 || n < 1) return fail(nullptr, COFFEE_ERR_INVALID, "invalid grid dims/guard");
    g.ex[a] = n > d ? n / d : n;
    g.step[a] = n > d ? d : 1;
    g.guard[a] = grid->guard[a];
  }
  const bool two_d = grid->dims[2] == 1;
  if (grid->dims[1] == 1 && two_d)
    return fail(nullptr, COFFEE_ERR_INVALID, "1D grids have no field-output sampling");
  // the minus-side neighbour of the first sample must lie inside the array
  for (int a = 0; a < (two_d ? 2 : 3); a++)
    if (!((stagger >> a) & 1) && grid->guard[a] < 1)
      return fail(nullptr, COFFEE_ERR_INVALID, "a staggered axis needs a guard cell");
  g.sy = grid->dims[0];
  g.sz = (long)grid->dims[0] * grid->dims[1];
  g.mi = (stagger & 1) ? 0 : -1;
  g.mj = (stagger & 2) ? 0 : -1;
  g.mk = (stagger & 4) ? 0 : -1;
  const long n = (long)g.ex[0] * g.ex[1] * g.ex[2];
  const int blocks = (int)std::min<long>((n + 255) / 256, 4096);
  hipStream_t s = (hipStream_t)stream;
  if (scalar_bytes == 4) {
    if (two_d) downsample_kernel<float, true><<<blocks, 256, 0, s>>>((const float*)src, dst, g);
    else downsample_kernel<float, false><<<blocks, 256, 0, s>>>((const float*)src, dst, g);
  } else {
    if (two_d) downsample_kernel<double, true><<<blocks, 256, 0, s>>>((const double*)src, dst, g);
    else downsample_kernel<double, false><<<blocks, 256, 0, s>>>((const double*)src, dst, g);
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(nullptr, COFFEE_ERR_HIP, std::string("downsample: ") + hipGetErrorString(e));
  return COFFEE_OK;
}
