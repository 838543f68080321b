// Scalar type of the host surface: Coffee::Scalar is float, or double when
// USE_DOUBLE is defined, as in the reference (src/data/typedefs.h:6-12).
#ifndef COFFEE_TYPEDEFS_H
#define COFFEE_TYPEDEFS_H

namespace Coffee {
#ifndef USE_DOUBLE
typedef float Scalar;
#define TINY ((float)(1e-7))
#else
typedef double Scalar;
#define TINY 1e-11
#endif
}  // namespace Coffee

#endif
