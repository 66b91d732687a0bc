// Scalar and index types of the Poisson API (reference include/definitions.h:22-25).
#ifndef POISSON_DEFINITIONS
#define POISSON_DEFINITIONS

#include <cstdint>

#ifndef MIN
#define MIN(a, b) (((a) < (b)) ? (a) : (b))
#endif

namespace Poisson {
using double_t = double;
using float_t = float;
using uint_t = long int;
using int_t = long int;
}  // namespace Poisson
#endif
