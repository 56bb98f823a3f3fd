// refine.h -- windows of a refined void-in-void level (refine.hip drives them,
// the field kernels are in fields.hip).
//
// A refined level has 2x the parent level's cells per axis over the same box,
// but only cubic windows around the parent voids are computed.  Windows are
// packed into an atlas: a non-periodic grid of side A holding each window's
// cells at its own atlas origin, so the void stage runs on the atlas as on
// any grid.  Cell l (0-based, per axis) of window W is cell o + l of the
// level's grid (wrapped in a periodic box) and atlas cell a + l.
#pragma once

#include "avism_internal.h"
#include "knn_search.h"

namespace avism {

struct WinDev {
  int32_t ox, oy, oz, w;        // origin in the level's cells (0-based, unwrapped) and side
  int32_t ax, ay, az, parent;   // origin in the atlas, parent void id
  int32_t sox, soy, soz, sw;    // the parent MARCA source window: origin in parent-level cells, side
  int32_t sax, say, saz, swrap; // its origin in the parent MARCA grid; swrap: the source is the periodic box
  int32_t bx0, by0, bz0, pad0;  // DDENS: first bin of the window's particle range (unwrapped)
  int32_t bnx, bny, bnz, pad1;  // bins of that range per axis
};

// the level's whole grid, as the field kernels see it
struct VirtG {
  const float* rad;  // RADX of the level's grid, reference index i at rad[i - rlo]
  int32_t rlo, n, periodic;
  float dx, rx1;
  int32_t A;         // atlas side
};

// items per window of each kernel, for the host's item offsets
int64_t win_vvel_items(int32_t w, int32_t k);
int64_t win_ddens_items(const WinDev& wd);
int64_t win_cell_items(int32_t w);
int64_t win_dep_items(int32_t w);
constexpr int kWinBins = 4;  // DDENS: a block takes 4 x 4 x 4 bins

hipError_t launch_win_vvel(hipStream_t st, const Bins& B, int64_t np, const WinDev* W, const int64_t* off,
                           int32_t nwin, int64_t items, const VirtG& V, int32_t k, float r0, const float4* vel,
                           uint32_t* hbits, float* u, float* v, float* w);
// the particles that meet a window marked in need (np bytes; items: the
// windows' bin tiles, off) and listed in list (np ints, the length in count,
// sel_tmp the selection's scratch), the NORM of each into norm, then the
// deposits tile by tile in LDS (dep_items: the windows' cell tiles, dep_off),
// each tile gathering the particles within hmax of its cells
hipError_t launch_win_ddens(hipStream_t st, const VirtG& V, const Bins& B, int64_t np, const float* hpart,
                            double scale, unsigned long long* acc, const WinDev* W, const int64_t* off, int32_t nwin,
                            int64_t items, unsigned long long* npairs, float* norm, uint8_t* need,
                            int32_t* list, int32_t* count, DBuf& sel_tmp, float hmax, const int64_t* dep_off,
                            int64_t dep_items);
hipError_t launch_win_marcap(hipStream_t st, const WinDev* W, const int64_t* off, int32_t nwin, int64_t items,
                             int32_t A, const int32_t* pmarca, int32_t pside, int32_t* marcap);
hipError_t launch_max_bits(hipStream_t st, int64_t n, const uint32_t* a, uint32_t* out);
hipError_t launch_mass_sum(hipStream_t st, int64_t np, const float* m, double* partial, int32_t nparts);
hipError_t launch_ddens_finalize(hipStream_t st, int64_t ncell, const unsigned long long* acc, double inv_scale,
                                 float den, float rote, float* umass, double* partial);

}  // namespace avism
