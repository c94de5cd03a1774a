// tune.h -- layout autotuner state: on/off switch and the cache of chosen
// layouts per problem shape (in-process, optionally persisted to a text file
// named by SSSI_TUNE_CACHE).  The timing itself runs in the propagation
// handles, on their own grids (prop2d.hip, prop3d.hip).
#pragma once

#include <cstdint>

namespace sssi {

// Off by default; SSSI_AUTOTUNE=1 or sssi_set_autotune(1) turns it on.
bool autotune_enabled();
void set_autotune(bool on);

// 2-D: (a, b) = (tile height rz, march groups); 3-D: (a, b) = (ty, z chunk).
struct TuneKey {
    int dim;
    int64_t nz, nx, ny;
    int N, nB;
};
bool tune_lookup(const TuneKey &k, int &a, int &b);
void tune_store(const TuneKey &k, int a, int b);

}  // namespace sssi
