// Test-infrastructure harness around the reference's own CPU field-output
// sampling, multi_array<Scalar>::downsample (src/data/multi_array.cpp:56-73),
// which data_exporter::add_grid_output calls for every fld output array
// (src/utils/data_exporter.cpp:685-705).
//
// Built only by oracle/Makefile, from the reference sources where they lie,
// into oracle/_ref/.  It is a checker: it never ships in the product.
//
// usage: ref_downsample_fXX <in.bin> <out.bin>
// in.bin (little endian): int32 {scalar_bytes, dims[3], guard[3], stagger, d},
//   then dims[0]*dims[1]*dims[2] scalars in Fortran order (guards included).
// out.bin: int32 {ext[3]}, then ext[0]*ext[1]*ext[2] floats, where
//   ext[i] = N[i] / d if N[i] > d else N[i], N[i] = dims[i] - 2 guard[i]
//   (the tmp_grid_data extent, data_exporter.cpp:70-79).

#include <cstdio>
#include <vector>

#include "data/multi_array.h"
#include "data/stagger.h"
#include "data/typedefs.h"

using namespace Coffee;

int main(int argc, char* argv[]) {
  if (argc != 3) {
    fprintf(stderr, "usage: %s in.bin out.bin\n", argv[0]);
    return 2;
  }
  FILE* f = fopen(argv[1], "rb");
  if (!f) return 1;
  int hdr[9];
  if (fread(hdr, sizeof(int), 9, f) != 9) return 1;
  if (hdr[0] != (int)sizeof(Scalar)) {
    fprintf(stderr, "scalar size mismatch\n");
    return 1;
  }
  const int dims[3] = {hdr[1], hdr[2], hdr[3]};
  const int guard[3] = {hdr[4], hdr[5], hdr[6]};
  const int stagger = hdr[7], d = hdr[8];
  multi_array<Scalar> a(dims[0], dims[1], dims[2]);
  const size_t n = (size_t)dims[0] * dims[1] * dims[2];
  if (fread(a.host_ptr(), sizeof(Scalar), n, f) != n) return 1;
  fclose(f);
  int ext[3];
  for (int i = 0; i < 3; i++) {
    ext[i] = dims[i] - 2 * guard[i];
    if (ext[i] > d) ext[i] /= d;
  }
  multi_array<float> out(ext[0], ext[1], ext[2]);
  a.downsample(d, out, Index(guard[0], guard[1], guard[2]), Stagger(stagger));
  FILE* g = fopen(argv[2], "wb");
  if (!g) return 1;
  fwrite(ext, sizeof(int), 3, g);
  fwrite(out.host_ptr(), sizeof(float), (size_t)ext[0] * ext[1] * ext[2], g);
  fclose(g);
  return 0;
}
