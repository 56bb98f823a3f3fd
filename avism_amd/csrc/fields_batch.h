// fields_batch.h -- the field stage of a batch of independent patches, each an
// AVISM grid of its own size, in persistent launches (fields.hip) driven by
// batch.hip.
//
// Every per-patch array is reached through base pointers offset to the
// patch, so the kernel bodies are the single-grid ones: particle indices are
// local to the patch (0..np-1, bin-sorted within it), cell indices local to
// its inner grid.  Work items (chunks of particles, tiles of cells) of all
// patches form one flat range; item ranges of patch p are [off[p], off[p+1]).
#pragma once

#include "avism_internal.h"
#include "knn_search.h"

namespace avism {

struct PatchDev {
  Bins B;                    // the patch's bin-sorted particles and its own bin starts
  const float* rad;          // RADX, reference index i at rad[i - rlo]
  int32_t rlo, n, periodic, uniform;
  float dx, rx1, r0;         // cell size, RADX(1), first kNN search radius
  double scale;              // DDENS fixed point: 2^62 / (np * max mass) rounded down to a power of 2
  int64_t np;
  const float4* vel;         // (vx, vy, vz, m/rho), bin-sorted
  uint32_t* hbits;           // HPART bits, bin-sorted
  float *pdens, *hk, *norm;  // bin-sorted
  float *u, *v, *w;          // inner velocity fields
  unsigned long long* acc;   // DDENS cell sums
  const int8_t* smask;       // inner SMASK, or null
};

// the per-XCD work queues of one persistent launch: 8 counters, XCD x takes
// items [total * x / 8, total * (x + 1) / 8) in order
struct ItemQueue {
  const int64_t* off;  // npatch + 1 item offsets
  int32_t npatch;
  int64_t total;
  unsigned* cnt;       // 8 counters, zero before the launch
};

// tile shapes the autotuner picks from
constexpr int kVvelShapes = 3;   // 8x4x4, 4x8x4, 4x4x8 cells (one cell per lane, 128 lanes)
constexpr int kDdensShapes = 3;  // 8x8x4 widened by 4, 8x8x8 by 3, 16x8x4 by 4
const char* vvel_shape_name(int s);
const char* ddens_shape_name(int s);
int64_t vvel_shape_tiles(int s, int n);
int64_t ddens_shape_tiles(int s, int n);
int64_t ppart_chunks(int64_t np);
// kNN bins per axis for np particles on n^3 cells (fields.hip)
int32_t bins_per_axis(int32_t n, int64_t np);

// launches (k <= 64); nblocks: the persistent grid
hipError_t launch_batch_ppart(hipStream_t st, const PatchDev* P, ItemQueue q, int k, unsigned nblocks);
hipError_t launch_batch_vvel(hipStream_t st, const PatchDev* P, ItemQueue q, int k, int shape, unsigned nblocks);
hipError_t launch_batch_ddens(hipStream_t st, const PatchDev* P, ItemQueue q, int shape, unsigned nblocks,
                              unsigned long long* npairs);
// per patch p: cell sums -> U1 with the patch's scale (read on the device)
hipError_t launch_batch_finalize(hipStream_t st, const PatchDev* P, int p, int64_t ncell, float den, float rote,
                                 float* umass, double* partial);
// fills (graph-capturable, unlike memsets)
hipError_t launch_fill_u32(hipStream_t st, int64_t n, uint32_t* a, uint32_t v);
hipError_t launch_fill_u64(hipStream_t st, int64_t n, unsigned long long* a, unsigned long long v);
// the remaining single-grid kernels, on explicit arguments
hipError_t launch_velrec(hipStream_t st, int64_t np, const uint32_t* orig, const float4* pos, const float* pdens,
                         const float* vx, const float* vy, const float* vz, float4* vel);
hipError_t launch_fill_h(hipStream_t st, int64_t np, float* hpart, const float* hk);
hipError_t launch_reduce_mean(hipStream_t st, const double* partial, int64_t nparts, int64_t ncell, double* slices,
                              double* sum_out, float* mean_out);
hipError_t launch_div_finalize(hipStream_t st, int n, float dx, int periodic, const float* u, const float* v,
                               const float* w, const float* umass, const float* mean, float* u1co, float* div);
int64_t finalize_blocks(int64_t ncell);
hipError_t launch_convert_check(hipStream_t st, int64_t np, float* m, const float* x, const float* y, const float* z,
                                float half, int* flags);
hipError_t launch_convert_vel(hipStream_t st, int64_t np, float* vx, float* vy, float* vz);
hipError_t launch_gather_sorted(hipStream_t st, int64_t np, const uint32_t* perm, const float* x, const float* y,
                                const float* z, const float* m, float4* pos, uint32_t* orig);
hipError_t launch_scatter_orig(hipStream_t st, int64_t np, const uint32_t* orig, int64_t orig0, const float* a,
                               float* out);

}  // namespace avism
