// Friends-of-friends halo finder on the device (the groups of fof.c:40-1180
// on one rank: every pair of particles closer than the linking length is
// linked, periodic in all three axes).
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "../../include/cola_halo_amd.h"

namespace cola {

struct FofInput {
  float* base = nullptr;   // snapshot records: x at base[i * stride + a], v at base[i * stride + 3 + a]
  long long stride = 8;    // floats per record
  long long np = 0;
  float boxsize = 0;
  float ll = 0;            // linking length
  int nfof_min = 32;       // fof.c:224
};

struct FofResult {
  std::vector<cola_halo> halos;  // ordered by the smallest record index of each group
  int* d_label = nullptr;        // per record: smallest record index of its group, or -1 below nfof_min
  long long np = 0;
};

// Runs the finder.  Wraps the record positions into [0, L) in place first,
// as move_particles2_min does before fof.c builds its tree.  Device memory
// of the search is allocated and freed inside; d_label is (re)allocated in
// *out and owned by it (free with fof_free).
int fof_run(const FofInput& in, FofResult* out, hipStream_t s);
void fof_free(FofResult* r);

// Number of cells per axis of the linking grid (cell side >= 1.01 ll, at most 1600).
int fof_cells_per_axis(float boxsize, float ll);

}  // namespace cola
