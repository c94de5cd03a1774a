// prop3d.hip -- 3-D CPML propagator (placeholder until the z-sweep kernel lands).
#include "prop.h"

namespace sssi {
Prop *make_prop3d(const double *, int64_t, int64_t, int64_t, int, int, double, double, double,
                  double) {
    fail(SSSI_ESTATE, "3-D propagator not built");
}
}  // namespace sssi
