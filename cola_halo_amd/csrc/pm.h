// Particle-mesh force kernels (pm.c) and the COLA particle kernels (cola.c,
// move.c wrap) on structure-of-arrays particle state.
#pragma once

#include <hip/hip_runtime.h>

#include "cosmo.h"

namespace cola {

// Device particle state, one float array per component (lattice-ID order;
// id = index + 1, lpt.c:516,557).
struct DevParticles {
  long long np = 0;
  float* x[3] = {nullptr, nullptr, nullptr};
  float* v[3] = {nullptr, nullptr, nullptr};
  float* dx1[3] = {nullptr, nullptr, nullptr};
  float* dx2[3] = {nullptr, nullptr, nullptr};
  float* f[3] = {nullptr, nullptr, nullptr};
};

// CIC geometry of pm.c on one rank.
struct CicGeom {
  int ng = 0;
  float scale = 0.0f;    // pm.c:215  Ngrid / BoxSize
  float wpar = 0.0f;     // pm.c:216  (Ngrid / nc)^3
  float x_right = 0.0f;  // pm.c:422-423  ghost threshold L/Ng*Ng - L/Ng
  float shift = 0.0f;    // pm.c:434  -BoxSize
  long long row = 0;     // real row length in floats, >= 2(ng/2+1) (in-place r2c)
  // readout: consecutive workgroups per XCD (xcd_group_remap); a lattice
  // slab's rows split into one y band per XCD keeps each XCD's grid rows in
  // its own L2
  int readout_group = 8;
};
CicGeom make_cic_geom(int ng, float boxsize, float wpar, long long row);

// Tiling of the density grid for the binned deposit (power-of-two tiles).
struct TileGeom {
  int tx = 8, ty = 8, tz = 32;
  int lx = 3, ly = 3, lz = 5;
  int ntx = 0, nty = 0, ntz = 0;
  long long ntiles = 0;
};
bool make_tile_geom(int ng, int tx, int ty, int tz, TileGeom* t);

struct BinBuffers {
  int* counts = nullptr;       // ntiles, zero between steps
  int* offsets = nullptr;      // ntiles + 1
  int* block_sums = nullptr;   // scan scratch
  float* sorted = nullptr;     // 3 floats per entry
  long long capacity = 0;      // entries
  int* overflow = nullptr;     // set when an index fell outside the grid or capacity
  double* tile_sums = nullptr; // per-tile density sums of the last deposit
};

int bin_alloc(BinBuffers* b, const TileGeom& t, long long np);
void bin_free(BinBuffers* b);

// move.c:33-41
int launch_wrap(const DevParticles& p, float boxsize, hipStream_t s);
// Counting sort of particle positions into every tile their CIC stencil (or
// their ghost copy's stencil) touches.  wrap_first folds move.c's wrap into
// the counting pass.  Three launches: count, scan, scatter.
int launch_bin_count(const DevParticles& p, const CicGeom& g, const TileGeom& t, BinBuffers& b, bool wrap_first,
                     float boxsize, hipStream_t s);
int launch_bin_scan(const TileGeom& t, BinBuffers& b, hipStream_t s);
int launch_bin_scatter(const DevParticles& p, const CicGeom& g, const TileGeom& t, BinBuffers& b, hipStream_t s);
// Per-step values the step kernels read from device memory (one small
// host-to-device copy per step), so that a captured step graph replays with
// new values: the mass-check ring slot, the time and the kick / drift factors.
struct StepParams {
  int slot;
  float a;  // scale factor of the positions the force is computed at
  KickFactors k;
  DriftFactors d;
};

// LDS-accumulated deposit, one workgroup per tile, no global atomics:
// density = -1 + WPAR * sum of CIC weights (pm.c:209-279).  ring[prm->slot]
// (zeroed by launch_zero_mass) receives the double sum over the grid
// (pm.c:535-562).
int launch_deposit_tiles(const CicGeom& g, const TileGeom& t, const BinBuffers& b, float* density, double* ring,
                         const StepParams* prm, hipStream_t s);
int launch_zero_mass(double* ring, const StepParams* prm, hipStream_t s);
// pm.c:359-412 + 505-532: readout of the three force grids into p.f.
// fsum: null, or the three sums of a residual added to the force grids,
// whose mean is taken off every particle's force.  fi: null (gather from
// fx, fy, fz), or one interleaved grid fi[cell][3] of the same forces.
int launch_readout(const DevParticles& p, const CicGeom& g, const float* fx, const float* fy, const float* fz,
                   const double* fsum, const float* fi, hipStream_t s);
// Readout fused with cola.c:104-112 kick and cola.c:142-149 drift.
int launch_readout_kick_drift(const DevParticles& p, const CicGeom& g, const float* fx, const float* fy,
                              const float* fz, const StepParams* prm, const double* fsum, const float* fi,
                              hipStream_t s);
int launch_kick(const DevParticles& p, const KickFactors& k, hipStream_t s);
int launch_drift(const DevParticles& p, const DriftFactors& d, hipStream_t s);

// Snapshot records (cola.c:388-410) / non-COLA initial velocities (cola.c:321-331)
struct DevParticleMin {
  float x[3];
  float v[3];
  long long id;
};
int launch_snapshot(const DevParticles& p, const SnapshotFactors& f, DevParticleMin* out, hipStream_t s);
int launch_noncola_initial(const DevParticles& p, const SnapshotFactors& f, DevParticleMin* out, hipStream_t s);

int launch_zero_doubles(double* p, int n, hipStream_t s);

// ---- batched step kernels: one launch per kernel for many realizations ----
// What the step kernels of one realization read.
struct PmMember {
  DevParticles p;
  CicGeom g;
  TileGeom t;
  BinBuffers b;
  float* density;
  const float* f[3];       // force grids the readout gathers from
  double* ring;            // mass-check ring (ring[prm->slot] gets the grid sum)
  const StepParams* prm;   // this step's values of the realization
  float boxsize;
};

// Device tables of a batched step: the members and, per kernel, the first
// workgroup of every member (kernel k: start[k * (n + 1) + m], m <= n).
enum PmBatchKernel : int { kBatchBin = 0, kBatchScan, kBatchTile, kBatchSum, kBatchRead, kPmBatchKernels };
struct PmBatch {
  int n = 0;
  const PmMember* members = nullptr;
  const int* start = nullptr;
  int blocks[kPmBatchKernels] = {};
  size_t deposit_lds = 0;
  void* d_mem = nullptr;
  int init(const PmMember* members, int n);
  void destroy();
};
// The kernels of assign_kernels for every member, each one launch: mass
// slot cleared, wrap + tile counts, scan (3 launches), scatter, deposit,
// grid sums.  Same per-particle and per-tile arithmetic as the single
// launches: the density grids are the same bit for bit.
int launch_assign_batch(const PmBatch& b, hipStream_t s);
// The fused readout + kick + drift of every member (plain force grids).
int launch_readout_kick_drift_batch(const PmBatch& b, hipStream_t s);

}  // namespace cola
