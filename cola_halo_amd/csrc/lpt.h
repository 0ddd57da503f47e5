// 2LPT initial-condition kernels (lpt.c:145-578).
#pragma once

#include <hip/hip_runtime.h>

#include "pm.h"

namespace cola {

// lpt.c:172-207 seed table of one RANLUX stream (host, sequential).
void lpt_seedtable(int nmesh, int seed, unsigned int* table);
// first n uniforms of the stream seeded with seed (host)
void ranlux_uniforms(unsigned long long seed, int n, double* out);

// Device copy of the power table (power.c:26-29) for the mode fill.
struct DevPower {
  const double* logk = nullptr;
  const double* logD = nullptr;
  int n = 0;
  double norm = 1.0;
};

// lpt.c:209-351: white-noise modes times sqrt(P(k)) into the three ZA
// displacement spectra [i][j][k <= n/2] (every element written).
int launch_mode_fill(int n, const unsigned int* d_seedtable, const DevPower& pw, double box, double fac,
                     float2* cd0, float2* cd1, float2* cd2, hipStream_t s);
// lpt.c:409-422: second-order source into g3 (padded real grids).
int launch_lpt_source(int n, const float* g0, const float* g1, const float* g2, float* g3, const float* g4,
                      const float* g5, hipStream_t s);

struct LptParticleFactors {
  float dx;          // lpt.c:508
  float Dplus;       // lpt.c:509
  double D2, D20;    // lpt.c:523-524
  double nmesh3_inv; // lpt.c:515
};
// lpt.c:532-561: lattice particles from the real displacement grids.
int launch_lpt_particles(int n, const float* const disp[3], const float* const disp2[3],
                         const LptParticleFactors& f, const DevParticles& p, hipStream_t s);

}  // namespace cola
