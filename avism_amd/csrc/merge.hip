// merge.hip -- cube volume order, MERGE_VOID and post-processing on gfx950.
//
// Reference: avism.f90:1727-1973, voidfinding.f90:718-999 (MERGE_VOID),
// 1007-1161 (SIMPLE_CONNECTION_FAST), 1367-1568 (VOID_PERIODIC),
// shape.f90:140-225 (VOID_SHAPE).
//
// MERGE_VOID walks the cubes in descending volume (rank 0,1,...).  Its outcome
// is reproduced without the sequential walk:
//  * cube c is absorbed at step j*(c) = the lowest rank j < rank(c) whose cube,
//    grown by NBUFF=1, overlaps c grown by 1 (the ball query of 918-929 is a
//    superset of those cubes, since later cubes are never larger); with boxes
//    grown by 2 painted as atomicMin(rank), j* is the minimum over c's box.
//  * every cube paints its box once, at key 2*rank (if it is a root, i.e. has
//    no such j) or 2*j*+1 (absorbed); fill-if-zero painting means MARCA(x) is
//    the group of the cube with the smallest key over x.
//  * a root looks for a "close brother" (829-860): the nearest cell of its box
//    grown by 1 already painted before its own step (key < 2*rank), ties to the
//    first cell in the reference's loop order.  Its parent pointer is that
//    cell's painter (key>>1); an absorbed cube's is j*; a root without brother
//    is a major.  Pointer jumping gives every cube its group (= major).
//  * GXC/VOLNEW are the reference's FP32 running volume-weighted means, folded
//    per group in the reference's event order (step, then ball order by
//    distance, ties by id).
//  * void-in-void levels: only cubes of the same parent void (FATHER) absorb
//    each other, by either rule (voidfinding.f90:848-850, 939), so groups
//    never mix fathers.  j* then takes the lowest rank among same-father
//    cubes.  A cell whose +-2 boxes all come from one father gives that
//    father's minimum rank directly; a cube whose box holds a cell where
//    several fathers' boxes meet finds j* by enumerating the cubes near it
//    (bucket grid of cube centres).  The close-brother test skips cells
//    whose painter belongs to another father.
#include <climits>
#include <cstring>

#include <rocprim/rocprim.hpp>

#include "avism_internal.h"

namespace avism {

namespace {

struct ExtG {
  int n, low1, low2, ne, periodic;
  __device__ __forceinline__ size_t ext(int i, int j, int k) const {
    return (size_t)(i - low1) + (size_t)ne * ((size_t)(j - low1) + (size_t)ne * (k - low1));
  }
};

struct Cubes {
  const int32_t* ci;  // [9][nv]
  const float* cf;    // [7][nv]
  const int32_t* father;
  int nv;
  __device__ __forceinline__ int ini(int a, int v) const { return ci[a * nv + v]; }
  __device__ __forceinline__ int fin(int a, int v) const { return ci[(3 + a) * nv + v]; }
  __device__ __forceinline__ int ic(int a, int v) const { return ci[(6 + a) * nv + v]; }
};

inline unsigned nbk(int64_t n, int b = 256) { return (unsigned)((n + b - 1) / b); }

__global__ void k_fill(int64_t n, int32_t* a, int32_t v) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = v;
}

__global__ void k_vol_keys(int nv, Cubes C, float* vol) {
  int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v < nv) vol[v] = C.cf[6 * nv + v];
}

// INDICE (descending volume, avism.f90:1740-1745) and rank of each id
__global__ void k_indice(int nv, const int32_t* __restrict__ asc, int32_t* indice, int32_t* rank) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nv) return;
  int id = asc[nv - 1 - i];
  indice[i] = id;
  rank[id - 1] = i;
}

__global__ void k_check_father(int nv, Cubes C, int32_t* flag) {
  int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v < nv && C.father[v] != C.father[0]) atomicOr(flag, 1);
}

constexpr int32_t kFatherEmpty = INT_MIN, kFatherMixed = INT_MIN + 1;

// the father of the +-2 boxes over each cell: one father, or kFatherMixed
// where boxes of several fathers meet (a mixed cell never changes back)
__global__ void k_paint_father(int nv, ExtG G, Cubes C, int32_t* F) {
  int v = blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64;
  int lane = threadIdx.x % 64;
  if (v >= nv) return;
  const int f = C.father[v];
  int lo[3], hi[3];
  for (int a = 0; a < 3; ++a) {
    lo[a] = max(C.ini(a, v) - 2, G.low1);
    hi[a] = min(C.fin(a, v) + 2, G.low2);
  }
  int sx = hi[0] - lo[0] + 1, sy = hi[1] - lo[1] + 1, sz = hi[2] - lo[2] + 1;
  int vol = sx * sy * sz;
  for (int q = lane; q < vol; q += 64) {
    int a = q % sx, b = (q / sx) % sy, c = q / (sx * sy);
    int32_t* cell = &F[G.ext(lo[0] + a, lo[1] + b, lo[2] + c)];
    int old = atomicCAS(cell, kFatherEmpty, f);
    if (old != kFatherEmpty && old != f && old != kFatherMixed) atomicExch(cell, kFatherMixed);
  }
}

// box of cube v grown by g, clipped to the grid, painted with atomicMin(val)
__global__ void k_paint_box(int nv, ExtG G, Cubes C, const int32_t* __restrict__ indice,
                            const int32_t* __restrict__ keyval, int grow, int32_t* grid) {
  int r = blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64;  // rank
  int lane = threadIdx.x % 64;
  if (r >= nv) return;
  int v = indice[r] - 1;
  int lo[3], hi[3];
  for (int a = 0; a < 3; ++a) {
    lo[a] = max(C.ini(a, v) - grow, G.low1);
    hi[a] = min(C.fin(a, v) + grow, G.low2);
  }
  int sx = hi[0] - lo[0] + 1, sy = hi[1] - lo[1] + 1, sz = hi[2] - lo[2] + 1;
  int vol = sx * sy * sz;
  int key = keyval ? keyval[r] : r;
  for (int q = lane; q < vol; q += 64) {
    int a = q % sx, b = (q / sx) % sy, c = q / (sx * sy);
    atomicMin(&grid[G.ext(lo[0] + a, lo[1] + b, lo[2] + c)], key);
  }
}

// j*(c): min of the +-2 painting over c's own box (rank order).  With the
// father grid F: c's own +-2 box covers each of these cells, so a cell of one
// father is c's father's and P holds that father's minimum; a mixed cell
// leaves j* to k_jstar_enumerate (jstar = -2, counted in namb).
__global__ void k_jstar(int nv, ExtG G, Cubes C, const int32_t* __restrict__ indice,
                        const int32_t* __restrict__ P, const int32_t* __restrict__ F, int32_t* jstar,
                        int32_t* key, int32_t* namb) {
  int r = blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64;
  int lane = threadIdx.x % 64;
  if (r >= nv) return;
  int v = indice[r] - 1;
  int lo[3], sz[3];
  for (int a = 0; a < 3; ++a) {
    lo[a] = C.ini(a, v);
    sz[a] = C.fin(a, v) - lo[a] + 1;
  }
  int vol = sz[0] * sz[1] * sz[2];
  int m = 0x7fffffff;
  bool mixed = false;
  for (int q = lane; q < vol; q += 64) {
    int a = q % sz[0], b = (q / sz[0]) % sz[1], c = q / (sz[0] * sz[1]);
    size_t e = G.ext(lo[0] + a, lo[1] + b, lo[2] + c);
    if (F && F[e] == kFatherMixed) mixed = true;
    else m = min(m, P[e]);
  }
  for (int off = 32; off > 0; off >>= 1) m = min(m, __shfl_xor(m, off));
  mixed = __any(mixed);
  if (lane == 0) {
    if (mixed) {
      jstar[r] = -2;
      atomicAdd(namb, 1);
      return;
    }
    int js = m < r ? m : -1;
    jstar[r] = js;
    key[r] = js >= 0 ? 2 * js + 1 : 2 * r;
  }
}

// cube centres bucketed on a grid of kBkt^3 ext cells (key = bucket + 1)
constexpr int kBkt = 8;
__global__ void k_bucket_keys(int nv, ExtG G, Cubes C, const int32_t* __restrict__ indice, int nbx,
                              uint32_t* key, int32_t* val, int32_t* rmax) {
  int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nv) return;
  int v = indice[r] - 1;
  int bx = (C.ic(0, v) - G.low1) / kBkt, by = (C.ic(1, v) - G.low1) / kBkt, bz = (C.ic(2, v) - G.low1) / kBkt;
  key[r] = (uint32_t)(bx + nbx * (by + nbx * bz)) + 1u;
  val[r] = r;
  atomicMax(rmax, (C.fin(0, v) - C.ini(0, v)) / 2);
}

// j* of a cube left undecided by k_jstar: the lowest rank j < r of its father
// whose box grown by 1 overlaps its own grown by 1, over the cubes whose
// centres lie within r + rmax + 2 cells (lanes over buckets)
__global__ void k_jstar_enumerate(int nv, ExtG G, Cubes C, const int32_t* __restrict__ indice, int nbx,
                                  const int32_t* __restrict__ bstart, const int32_t* __restrict__ bend,
                                  const int32_t* __restrict__ branks, const int32_t* __restrict__ rmaxp,
                                  int32_t* jstar, int32_t* key) {
  int r = blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64;
  int lane = threadIdx.x % 64;
  if (r >= nv || jstar[r] != -2) return;
  const int v = indice[r] - 1;
  const int f = C.father[v];
  int lo[3], hi[3], b0[3], nb[3];
  const int reach = (C.fin(0, v) - C.ini(0, v)) / 2 + *rmaxp + 2;
  for (int a = 0; a < 3; ++a) {
    lo[a] = C.ini(a, v) - 1;
    hi[a] = C.fin(a, v) + 1;
    int c0 = max(C.ic(a, v) - reach, G.low1), c1 = min(C.ic(a, v) + reach, G.low2);
    b0[a] = (c0 - G.low1) / kBkt;
    nb[a] = (c1 - G.low1) / kBkt - b0[a] + 1;
  }
  const int ntot = nb[0] * nb[1] * nb[2];
  int m = 0x7fffffff;
  for (int q = lane; q < ntot; q += 64) {
    int bx = b0[0] + q % nb[0], by = b0[1] + (q / nb[0]) % nb[1], bz = b0[2] + q / (nb[0] * nb[1]);
    int b = bx + nbx * (by + nbx * bz);
    for (int i = bstart[b]; i < bend[b]; ++i) {
      int j = branks[i];
      if (j >= r || j >= m) continue;
      int w = indice[j] - 1;
      if (C.father[w] != f) continue;
      bool ov = true;
      for (int a = 0; a < 3; ++a) ov = ov && C.ini(a, w) - 1 <= hi[a] && lo[a] <= C.fin(a, w) + 1;
      if (ov) m = j;
    }
  }
  for (int off = 32; off > 0; off >>= 1) m = min(m, __shfl_xor(m, off));
  if (lane == 0) {
    int js = m < r ? m : -1;
    jstar[r] = js;
    key[r] = js >= 0 ? 2 * js + 1 : 2 * r;
  }
}

// parent pointer (rank space): absorber, close brother, or self (major)
__global__ void k_parent(int nv, ExtG G, Cubes C, const int32_t* __restrict__ indice,
                         const int32_t* __restrict__ jstar, const int32_t* __restrict__ K,
                         int check_father, int32_t* parent) {
  int r = blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64;
  int lane = threadIdx.x % 64;
  if (r >= nv) return;
  if (jstar[r] >= 0) {
    if (lane == 0) parent[r] = jstar[r];
    return;
  }
  int v = indice[r] - 1;
  int lo[3], sz[3], cc[3];
  for (int a = 0; a < 3; ++a) {
    lo[a] = max(C.ini(a, v) - 1, G.low1);
    int hi = min(C.fin(a, v) + 1, G.low2);
    sz[a] = hi - lo[a] + 1;
    cc[a] = C.ic(a, v);
  }
  int vol = sz[0] * sz[1] * sz[2];
  const int f = C.father[v];
  unsigned long long best = ~0ull;
  int bestkey = -1;
  for (int q = lane; q < vol; q += 64) {  // q follows the reference's loop order
    int a = q % sz[0], b = (q / sz[0]) % sz[1], c = q / (sz[0] * sz[1]);
    int x = lo[0] + a, y = lo[1] + b, z = lo[2] + c;
    int kk = K[G.ext(x, y, z)];
    // the painter's group has the painter's father (voidfinding.f90:848-850)
    if (kk < 2 * r && (!check_father || C.father[indice[kk >> 1] - 1] == f)) {
      long long d2 = (long long)(cc[0] - x) * (cc[0] - x) + (long long)(cc[1] - y) * (cc[1] - y) +
                     (long long)(cc[2] - z) * (cc[2] - z);
      unsigned long long kq = ((unsigned long long)d2 << 32) | (unsigned)q;
      if (kq < best) { best = kq; bestkey = kk; }
    }
  }
  for (int off = 32; off > 0; off >>= 1) {
    unsigned long long ob = __shfl_xor(best, off);
    int ok = __shfl_xor(bestkey, off);
    if (ob < best) { best = ob; bestkey = ok; }
  }
  if (lane == 0) parent[r] = (bestkey >= 0) ? (bestkey >> 1) : r;
}

__global__ void k_jump(int nv, const int32_t* __restrict__ in, int32_t* out, int32_t* changed) {
  int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nv) return;
  int p = in[r];
  int pp = in[p];
  out[r] = pp;
  if (pp != p) atomicOr(changed, 1);
}

// UVOID (1-based ids; -1 for majors) and the major count
__global__ void k_uvoid(int nv, const int32_t* __restrict__ indice, const int32_t* __restrict__ grp,
                        int32_t* uvoid, int32_t* nmajor) {
  int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nv) return;
  int id = indice[r];
  int g = grp[r];
  uvoid[id - 1] = (g == r) ? -1 : indice[g];
  if (g == r) atomicAdd(nmajor, 1);
}

__global__ void k_marca(int64_t ne3, const int32_t* __restrict__ K, const int32_t* __restrict__ indice,
                        const int32_t* __restrict__ grp, int32_t* marca) {
  int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= ne3) return;
  int k = K[e];
  marca[e] = (k == 0x7fffffff) ? 0 : indice[grp[k >> 1]];
}

// centres (voidfinding.f90:763-776): GXC = RADX(1)+(ICX-1)*DX, VOLNEW = cube volume
__global__ void k_centres(int nv, Cubes C, float dx, float rx1, float* gxc, float* gyc, float* gzc,
                          float* volnew) {
  int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= nv) return;
  gxc[v] = rx1 + (float)(C.ic(0, v) - 1) * dx;
  gyc[v] = rx1 + (float)(C.ic(1, v) - 1) * dx;
  gzc[v] = rx1 + (float)(C.ic(2, v) - 1) * dx;
  const float* cf = C.cf;
  int nv_ = C.nv;
  volnew[v] = ((cf[3 * nv_ + v] - cf[v]) * (cf[4 * nv_ + v] - cf[nv_ + v])) * (cf[5 * nv_ + v] - cf[2 * nv_ + v]);
}

__device__ __forceinline__ float kd_axis(float v, float L, int periodic) {
  if (periodic) {
    float a = fabsf(v), b = L - fabsf(v);
    v = a < b ? a : b;
  }
  return v;
}

// absorption events: primary key (group, step, phase), secondary (ball distance, id)
__global__ void k_events(int nv, const int32_t* __restrict__ indice, const int32_t* __restrict__ jstar,
                         const int32_t* __restrict__ grp, const float* __restrict__ gxc,
                         const float* __restrict__ gyc, const float* __restrict__ gzc, float L,
                         int periodic, uint64_t* k1, uint64_t* k2, int32_t* val) {
  int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nv) return;
  int id = indice[r];
  int g = grp[r];
  uint64_t step;
  uint32_t dbits = 0;
  if (g == r) {  // major: no event; sorts last
    k1[r] = ~0ull;
    k2[r] = 0;
    val[r] = r;
    return;
  }
  int js = jstar[r];
  if (js < 0) {
    step = 2ull * (uint64_t)r;  // brother absorption at its own step (phase 0)
  } else {
    step = 2ull * (uint64_t)js + 1;  // ball absorption at step j*
    int t = indice[js] - 1, a = id - 1;
    // kd-tree metric over cube centres (voidfinding.f90:780-788, kdtree.f90:1441-1465)
    float dx = kd_axis(gxc[a] - gxc[t], L, periodic);
    float dy = kd_axis(gyc[a] - gyc[t], L, periodic);
    float dz = kd_axis(gzc[a] - gzc[t], L, periodic);
    float d = dx * dx;
    d = d + dy * dy;
    d = d + dz * dz;
    dbits = __float_as_uint(d);
  }
  k1[r] = ((uint64_t)g << 40) | step;
  k2[r] = ((uint64_t)dbits << 32) | (uint32_t)(id - 1);
  val[r] = r;
}

__global__ void k_group_ranges(int nv, const uint64_t* __restrict__ k1, int32_t* gstart, int32_t* gend) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nv) return;
  uint64_t key = k1[i];
  if (key == ~0ull) return;
  int g = (int)(key >> 40);
  if (i == 0 || (k1[i - 1] >> 40) != (uint64_t)g || k1[i - 1] == ~0ull) gstart[g] = i;
  if (i == nv - 1 || (k1[i + 1] >> 40) != (uint64_t)g || k1[i + 1] == ~0ull) gend[g] = i + 1;
}

// sequential FP32 fold per group (voidfinding.f90:865-869, 961-966)
__global__ void k_fold(int nv, const int32_t* __restrict__ indice, const int32_t* __restrict__ grp,
                       const int32_t* __restrict__ gstart, const int32_t* __restrict__ gend,
                       const int32_t* __restrict__ ev, const float* __restrict__ cx,
                       const float* __restrict__ cy, const float* __restrict__ cz,
                       const float* __restrict__ cvol, float* gxc, float* gyc, float* gzc,
                       float* volnew) {
  int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nv || grp[r] != r) return;
  int v = indice[r] - 1;
  float X = cx[v], Y = cy[v], Z = cz[v], V = cvol[v];
  int s = gstart[r], e = gend[r];
  for (int i = s; i < e; ++i) {
    int a = indice[ev[i]] - 1;
    float va = cvol[a];
    float den = V + va;
    X = (X * V + cx[a] * va) / den;
    Y = (Y * V + cy[a] * va) / den;
    Z = (Z * V + cz[a] * va) / den;
    V = V + va;
  }
  gxc[v] = X;
  gyc[v] = Y;
  gzc[v] = Z;
  volnew[v] = V;
}

__global__ void k_gather_u64(int n, const uint64_t* __restrict__ src, const int32_t* __restrict__ idx,
                             uint64_t* dst) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[idx[i]];
}

// absorbed cubes keep their own centre and volume (the reference never updates them)
__global__ void k_copy_nonleaders(int nv, const int32_t* __restrict__ indice, const int32_t* __restrict__ grp,
                                  const float* __restrict__ cx, const float* __restrict__ cy,
                                  const float* __restrict__ cz, const float* __restrict__ cvol, float* gxc,
                                  float* gyc, float* gzc, float* volnew) {
  int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nv || grp[r] == r) return;
  int v = indice[r] - 1;
  gxc[v] = cx[v];
  gyc[v] = cy[v];
  gzc[v] = cz[v];
  volnew[v] = cvol[v];
}

// ---------------------------------------------------------- post-processing
// the id shared by a wave's nonzero ids (0 if it has none, -1 if they differ)
__device__ __forceinline__ int wave_uniform_id(int id) {
  int lead = id;
  for (int off = 32; off > 0; off >>= 1) lead = max(lead, __shfl_xor(lead, off));
  return __all(id == 0 || id == lead) ? lead : -1;
}

__device__ __forceinline__ double wave_sum(double s) {
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
  return s;
}

// cells per void over the ext grid (grid: x chunks, y, z); a wave whose void
// cells share one id adds once
__global__ void k_count_cells(ExtG G, const int32_t* __restrict__ marca, int32_t* ncell) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t e = (int64_t)x + (int64_t)G.ne * ((int64_t)blockIdx.y + (int64_t)G.ne * blockIdx.z);
  const int id = x < G.ne ? marca[e] : 0;
  const int u = wave_uniform_id(id);
  if (u == 0) return;
  if (u > 0) {
    const int cnt = __popcll(__ballot(id > 0));
    if (threadIdx.x % 64 == 0) atomicAdd(&ncell[u - 1], cnt);
  } else if (id > 0) {
    atomicAdd(&ncell[id - 1], 1);
  }
}

// NCELL_MIN / RMIN removals (avism.f90:1812-1827, 1845-1859): decided per main void;
// nremoved counts them (zeroed by the caller)
__global__ void k_mark_removed(int nv, const int32_t* __restrict__ uvoid, const int32_t* __restrict__ ncell,
                               int mode, float dx, double rmin3, int32_t* removed, float* volnew,
                               int32_t* nremoved) {
  int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= nv) return;
  removed[v] = 0;
  if (uvoid[v] >= 0) return;
  if (mode == 0) {
    if (ncell[v] < kNcellMin) removed[v] = 1;
  } else {
    float vol = (float)((double)ncell[v] * (double)((dx * dx) * dx));
    volnew[v] = vol;
    if ((double)vol < rmin3) removed[v] = 1;
  }
  if (removed[v]) atomicAdd(nremoved, 1);
}

__global__ void k_apply_removed_cubes(int nv, const int32_t* __restrict__ removed, int32_t* uvoid,
                                      int32_t* ncell, float* volnew) {
  int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= nv) return;
  int u = uvoid[v];
  if (removed[v]) {
    uvoid[v] = 0;
    ncell[v] = 0;
    volnew[v] = 0.0f;
  } else if (u > 0 && removed[u - 1]) {
    uvoid[v] = 0;
  }
}

// grid-stride over a fixed grid, so that a pass with nothing removed costs one small launch
__global__ void k_apply_removed_cells(int64_t ne3, const int32_t* __restrict__ removed,
                                      const int32_t* __restrict__ nremoved, int32_t* marca) {
  if (*nremoved == 0) return;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < ne3; e += (int64_t)gridDim.x * blockDim.x) {
    const int id = marca[e];
    if (id > 0 && removed[id - 1]) marca[e] = 0;
  }
}

// VOID_SHAPE (shape.f90:155-172): per-cell FP32 terms, FP64 sums
// (grid: x chunks, y, z over the ext grid)
__global__ void k_inertia(ExtG G, const float* __restrict__ rad, int rlo,
                          const int32_t* __restrict__ marca, const float* __restrict__ gxc,
                          const float* __restrict__ gyc, const float* __restrict__ gzc, double* inert) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t e = (int64_t)x + (int64_t)G.ne * ((int64_t)blockIdx.y + (int64_t)G.ne * blockIdx.z);
  int id = x < G.ne ? marca[e] : 0;
  if (__all(id == 0)) return;
  double t[6] = {0, 0, 0, 0, 0, 0};
  if (id > 0) {
    int I = x + G.low1, J = (int)blockIdx.y + G.low1, K = (int)blockIdx.z + G.low1;
    float rk[3] = {rad[I - rlo] - gxc[id - 1], rad[J - rlo] - gyc[id - 1], rad[K - rlo] - gzc[id - 1]};
    float rr = sqrtf((rk[0] * rk[0] + rk[1] * rk[1]) + rk[2] * rk[2]);
    float d = (1.0f * rr) * rr;
    t[0] = (double)(d - rk[0] * rk[0]);
    t[1] = (double)(d - rk[1] * rk[1]);
    t[2] = (double)(d - rk[2] * rk[2]);
    t[3] = (double)(0.0f * rr * rr - rk[0] * rk[1]);
    t[4] = (double)(0.0f * rr * rr - rk[0] * rk[2]);
    t[5] = (double)(0.0f * rr * rr - rk[1] * rk[2]);
  }
  // waves whose void cells share one id add once
  int lead = id;
  for (int off = 32; off > 0; off >>= 1) lead = max(lead, __shfl_xor(lead, off));
  bool uniform = __all(id == 0 || id == lead);
  if (uniform) {
    for (int c = 0; c < 6; ++c) {
      double s = t[c];
      for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
      if ((threadIdx.x % 64) == 0 && lead > 0) atomicAdd(&inert[(size_t)(lead - 1) * 6 + c], s);
    }
  } else if (id > 0) {
    for (int c = 0; c < 6; ++c) atomicAdd(&inert[(size_t)(id - 1) * 6 + c], t[c]);
  }
}

// JACOBI (shape.f90:46-134) in float with the reference's promotions
__device__ void jacobi3(float A[3][3], float D[3]) {
  float B[3], Z[3];
  for (int ip = 0; ip < 3; ip++) { B[ip] = A[ip][ip]; D[ip] = B[ip]; Z[ip] = 0.0f; }
  for (int i = 1; i <= 50; i++) {
    float sm = 0.0f;
    for (int ip = 0; ip < 2; ip++)
      for (int iq = ip + 1; iq < 3; iq++) sm = sm + fabsf(A[ip][iq]);
    if (sm == 0.0f) return;
    float tresh = (i < 4) ? (float)(0.2 * (double)(sm * sm)) : 0.0f;
    for (int ip = 0; ip < 2; ip++)
      for (int iq = ip + 1; iq < 3; iq++) {
        float g = (float)(100.0 * (double)fabsf(A[ip][iq]));
        if (i > 4 && fabsf(D[ip]) + g == fabsf(D[ip]) && fabsf(D[iq]) + g == fabsf(D[iq])) {
          A[ip][iq] = 0.0f;
        } else if (fabsf(A[ip][iq]) > tresh) {
          float h = D[iq] - D[ip], t;
          if (fabsf(h) + g == fabsf(h)) t = A[ip][iq] / h;
          else {
            float theta = (float)(0.5 * (double)h / (double)A[ip][iq]);
            t = (float)(1.0 / ((double)fabsf(theta) + (double)sqrtf((float)(1.0 + (double)(theta * theta)))));
            if (theta < 0.0f) t = -t;
          }
          float c = (float)(1.0 / (double)sqrtf((float)(1.0 + (double)(t * t))));
          float s = t * c;
          float tau = (float)((double)s / (1.0 + (double)c));
          h = t * A[ip][iq];
          Z[ip] = Z[ip] - h; Z[iq] = Z[iq] + h;
          D[ip] = D[ip] - h; D[iq] = D[iq] + h;
          A[ip][iq] = 0.0f;
          for (int j = 0; j < ip; j++) {
            float gg = A[j][ip], hh = A[j][iq];
            A[j][ip] = gg - s * (hh + gg * tau); A[j][iq] = hh + s * (gg - hh * tau);
          }
          for (int j = ip + 1; j < iq; j++) {
            float gg = A[ip][j], hh = A[j][iq];
            A[ip][j] = gg - s * (hh + gg * tau); A[j][iq] = hh + s * (gg - hh * tau);
          }
          for (int j = iq + 1; j < 3; j++) {
            float gg = A[ip][j], hh = A[iq][j];
            A[ip][j] = gg - s * (hh + gg * tau); A[iq][j] = hh + s * (gg - hh * tau);
          }
        }
      }
    for (int ip = 0; ip < 3; ip++) { B[ip] = B[ip] + Z[ip]; D[ip] = B[ip]; Z[ip] = 0.0f; }
  }
}

__global__ void k_shape(int nv, const int32_t* __restrict__ uvoid, const int32_t* __restrict__ ncell,
                        const double* __restrict__ inert, float dx, float* eps, float* ipr) {
  int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= nv) return;
  eps[v] = 0.0f;
  ipr[v] = 0.0f;
  if (uvoid[v] != -1) return;
  const double* T = inert + (size_t)v * 6;
  float nc = (float)ncell[v];
  float A[3][3];
  A[0][0] = (float)T[0] / nc; A[1][1] = (float)T[1] / nc; A[2][2] = (float)T[2] / nc;
  A[0][1] = A[1][0] = (float)T[3] / nc;
  A[0][2] = A[2][0] = (float)T[4] / nc;
  A[1][2] = A[2][1] = (float)T[5] / nc;
  float D[3] = {0, 0, 0};
  jacobi3(A, D);
  float ax[3] = {sqrtf(2.5f * fabsf(D[1] + D[2] - D[0])), sqrtf(2.5f * fabsf(D[2] + D[0] - D[1])),
                 sqrtf(2.5f * fabsf(D[0] + D[1] - D[2]))};
  for (int i = 0; i < 2; i++) {  // SORT (shape.f90:2-29)
    int kk = i;
    float pp = ax[i];
    for (int j = i + 1; j < 3; j++)
      if (ax[j] >= pp) { kk = j; pp = ax[j]; }
    if (kk != i) { ax[kk] = ax[i]; ax[i] = pp; }
  }
  float e = 1.0f - ax[2] / ax[0];
  float volm = ((nc * dx) * dx) * dx;
  float vell = (float)(((((double)(4.0f / 3.0f) * kPi) * (double)ax[0]) * (double)ax[1]) * (double)ax[2]);
  float ip = volm / vell;
  if (ip > 1.0f) { ip = -99.0f; e = -99.0f; }
  eps[v] = e;
  ipr[v] = ip;
}

__global__ void k_xc(int nv, Cubes C, const int32_t* __restrict__ uvoid, float dx, float rx1, float* xc,
                     float* yc, float* zc) {
  int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= nv) return;
  if (uvoid[v] >= 0) { xc[v] = yc[v] = zc[v] = 0.0f; return; }
  xc[v] = rx1 + (float)(C.ic(0, v) - 1) * dx;
  yc[v] = rx1 + (float)(C.ic(1, v) - 1) * dx;
  zc[v] = rx1 + (float)(C.ic(2, v) - 1) * dx;
}

// ---- SIMPLE_CONNECTION_FAST as connected components of the empty cells
__device__ __forceinline__ int cc_find(const int32_t* lab, int x) {
  int p = lab[x];
  while (p != x) {
    x = p;
    p = lab[x];
  }
  return x;
}

__device__ void cc_union(int32_t* lab, int a, int b) {
  for (;;) {
    a = cc_find(lab, a);
    b = cc_find(lab, b);
    if (a == b) return;
    if (a < b) { int t = a; a = b; b = t; }  // link the larger root under the smaller
    int old = atomicMin(&lab[a], b);
    if (old == a) return;
    a = old;
  }
}

// block-local components first: an 8^3 block labels its empty cells in LDS
// (label = the component's smallest cell, which is also its smallest global
// index), so the global union-find only joins components across block faces.
// A block whose cells are all empty (most of the grid) is one component with
// label 0; otherwise labels propagate to neighbours until none changes.
// isrep marks each component's smallest cell (its representative).
constexpr int CCB = 8;

__global__ void __launch_bounds__(256) k_cc_local(ExtG G, const int32_t* __restrict__ marca, int32_t* lab,
                                                 uint8_t* isrep) {
  __shared__ int16_t L[CCB * CCB * CCB];
  __shared__ int changed;
  const int nbx = (G.ne + CCB - 1) / CCB;
  const int bx = blockIdx.x % nbx, by = (blockIdx.x / nbx) % nbx, bz = blockIdx.x / (nbx * nbx);
  const int tid = threadIdx.x;
  int64_t eg[2];
  bool full = true;  // every in-grid cell of this thread is empty
  for (int r = 0; r < 2; ++r) {
    const int c = tid + 256 * r;
    const int gx = bx * CCB + c % CCB, gy = by * CCB + (c / CCB) % CCB, gz = bz * CCB + c / (CCB * CCB);
    const bool in = gx < G.ne && gy < G.ne && gz < G.ne;
    eg[r] = in ? (int64_t)gx + (int64_t)G.ne * ((int64_t)gy + (int64_t)G.ne * gz) : -1;
    const bool empty = in && marca[eg[r]] == 0;
    L[c] = empty ? (int16_t)c : (int16_t)-1;
    full = full && (empty || !in);
  }
  if (__syncthreads_and(full)) {
    for (int r = 0; r < 2; ++r) {
      const int c = tid + 256 * r;
      if (eg[r] >= 0) L[c] = 0;
    }
  } else {
    for (;;) {
      if (tid == 0) changed = 0;
      __syncthreads();
      bool ch = false;
      for (int r = 0; r < 2; ++r) {
        const int c = tid + 256 * r;
        int m = L[c];
        if (m < 0) continue;
        const int x = c % CCB, y = (c / CCB) % CCB, z = c / (CCB * CCB);
        int t;
        if (x > 0 && (t = L[c - 1]) >= 0) m = min(m, t);
        if (x < CCB - 1 && (t = L[c + 1]) >= 0) m = min(m, t);
        if (y > 0 && (t = L[c - CCB]) >= 0) m = min(m, t);
        if (y < CCB - 1 && (t = L[c + CCB]) >= 0) m = min(m, t);
        if (z > 0 && (t = L[c - CCB * CCB]) >= 0) m = min(m, t);
        if (z < CCB - 1 && (t = L[c + CCB * CCB]) >= 0) m = min(m, t);
        m = min(m, (int)L[m]);  // jump along the label chain
        if (m < L[c]) {
          L[c] = (int16_t)m;
          ch = true;
        }
      }
      if (ch) changed = 1;
      __syncthreads();
      const int any = changed;
      __syncthreads();
      if (!any) break;
    }
  }
  __syncthreads();
  for (int r = 0; r < 2; ++r) {
    const int c = tid + 256 * r;
    if (eg[r] < 0) continue;
    const int m = L[c];
    if (m < 0) {
      lab[eg[r]] = -1;
    } else {
      const int gx = bx * CCB + m % CCB, gy = by * CCB + (m / CCB) % CCB, gz = bz * CCB + m / (CCB * CCB);
      lab[eg[r]] = (int32_t)((int64_t)gx + (int64_t)G.ne * ((int64_t)gy + (int64_t)G.ne * gz));
    }
    isrep[eg[r]] = m == c ? 1 : 0;
  }
}

// joins across the low faces of the blocks: one wave per 8 x 8 block face
// (item = block * 3 + direction, four items per workgroup), lane (u, v) on the
// face.  A pair of local labels equal to its left or lower neighbour's on the
// face is already joined by that neighbour (the first lane of each distinct
// pair has neither), so a face whose two sides are one component each makes
// one union, not 64.
constexpr int kFaceW = 4;  // waves (faces) per workgroup
__global__ void __launch_bounds__(64 * kFaceW) k_cc_faces(ExtG G, int nbc, int32_t* lab) {
  const int64_t item = (int64_t)blockIdx.x * kFaceW + (threadIdx.x >> 6);
  if (item >= (int64_t)nbc * nbc * nbc * 3) return;
  const int dir = (int)(item % 3), b = (int)(item / 3);
  const int bx = b % nbc, by = (b / nbc) % nbc, bz = b / (nbc * nbc);
  const int l = threadIdx.x & 63, u = l & (CCB - 1), v = l / CCB;
  int i, j, k;
  int64_t step;
  if (dir == 0) {
    if (bx == 0) return;
    i = bx * CCB;
    j = by * CCB + u;
    k = bz * CCB + v;
    step = 1;
  } else if (dir == 1) {
    if (by == 0) return;
    i = bx * CCB + u;
    j = by * CCB;
    k = bz * CCB + v;
    step = G.ne;
  } else {
    if (bz == 0) return;
    i = bx * CCB + u;
    j = by * CCB + v;
    k = bz * CCB;
    step = (int64_t)G.ne * G.ne;
  }
  int la = -1, lb = -1;
  if (i < G.ne && j < G.ne && k < G.ne) {
    const int64_t e = (int64_t)i + (int64_t)G.ne * ((int64_t)j + (int64_t)G.ne * k);
    la = lab[e];
    lb = lab[e - step];
  }
  const int pl = l > 0 ? l - 1 : 0, ql = l >= CCB ? l - CCB : 0;
  const int pa = __shfl(la, pl), pb = __shfl(lb, pl), qa = __shfl(la, ql), qb = __shfl(lb, ql);
  const bool dup = (u > 0 && pa == la && pb == lb) || (v > 0 && qa == la && qb == lb);
  if (la >= 0 && lb >= 0 && !dup) cc_union(lab, la, lb);
}

// every representative points at its root; any empty cell's root is then
// lab[lab[cell]] (a cell's entry is its representative, a root its own)
__global__ void k_cc_flatten(int64_t ne3, const uint8_t* __restrict__ isrep, int32_t* lab) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e < ne3 && isrep[e]) lab[e] = cc_find(lab, (int)e);
}

// holes (empty cells not connected to the corner, voidfinding.f90:1040-1135)
// take the id of the first non-hole cell in +x (1140-1156)
// (grid: x chunks, y, z)
__global__ void k_fill_holes(ExtG G, const int32_t* __restrict__ lab, int32_t* marca, int32_t* nholes) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= G.ne) return;
  const int64_t e = (int64_t)i + (int64_t)G.ne * ((int64_t)blockIdx.y + (int64_t)G.ne * blockIdx.z);
  const int l0 = lab[0];
  const int outside = l0 >= 0 ? lab[l0] : -1;
  const int l = lab[e];
  if (l < 0 || lab[l] == outside) return;
  atomicAdd(nholes, 1);
  int64_t f = e;
  for (int lf; i < G.ne && (lf = lab[f]) >= 0 && lab[lf] != outside;) {
    ++i;
    ++f;
  }
  marca[e] = (i < G.ne) ? marca[f] : 0;
}

// ---- VOID_PERIODIC (voidfinding.f90:1367-1568), exact sequential semantics
__global__ void k_inner_presence(int64_t ne3, ExtG G, const int32_t* __restrict__ marca, int32_t* has) {
  int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= ne3) return;
  int id = marca[e];
  if (id <= 0) return;
  int I = (int)(e % G.ne) + G.low1, J = (int)((e / G.ne) % G.ne) + G.low1,
      K = (int)(e / ((int64_t)G.ne * G.ne)) + G.low1;
  if (I >= 1 && I <= G.n && J >= 1 && J <= G.n && K >= 1 && K <= G.n) has[id - 1] = 1;
}

__global__ void k_absent_removed(int nv, const int32_t* __restrict__ uvoid, const int32_t* __restrict__ has,
                                 int32_t* removed, int32_t* nremoved) {
  int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= nv) return;
  removed[v] = (uvoid[v] == -1 && !has[v]) ? 1 : 0;
  if (removed[v]) atomicAdd(nremoved, 1);
}

__global__ void k_label_ranges(int64_t ne3, const uint32_t* __restrict__ key, int32_t* lstart, int32_t* lend) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ne3) return;
  uint32_t k = key[i];
  if (k == 0xffffffffu) return;
  if (i == 0 || key[i - 1] != k) lstart[k - 1] = (int32_t)i;
  if (i == ne3 - 1 || key[i + 1] != k) lend[k - 1] = (int32_t)(i + 1);
}

__device__ __forceinline__ int resolve(const int32_t* lmap, int l) {
  while (l > 0 && lmap[l - 1] != l) l = lmap[l - 1];
  return l;
}

// the part of VOID_PERIODIC's per-void test that no earlier void can change
// (voidfinding.f90:1440-1480): a main void whose centre lies outside the box
// and whose image cell, inside the ext grid, belongs to a void.  flag[I] for
// INDICE rank I.
__global__ void k_vp_flag(int nv, ExtG G, const int32_t* __restrict__ indice, const int32_t* __restrict__ uvoid,
                          const int32_t* __restrict__ marca_base, const float* __restrict__ gxc,
                          const float* __restrict__ gyc, const float* __restrict__ gzc, float half, float dx,
                          float rx1, int32_t* flag, int4* info, int32_t* vlab) {
  const int I = blockIdx.x * blockDim.x + threadIdx.x;
  if (I >= nv) return;
  const int id = indice[I];
  int f = 0;
  if (uvoid[id - 1] == -1) {
    const float xc1 = gxc[id - 1], yc1 = gyc[id - 1], zc1 = gzc[id - 1];
    const int sx = xc1 < -half ? 1 : (xc1 > half ? -1 : 0);
    const int sy = yc1 < -half ? 1 : (yc1 > half ? -1 : 0);
    const int sz = zc1 < -half ? 1 : (zc1 > half ? -1 : 0);
    if (sx | sy | sz) {
      const int ax = (int)((xc1 - rx1) / dx) + 1 + sx * G.n, ay = (int)((yc1 - rx1) / dx) + 1 + sy * G.n,
                az = (int)((zc1 - rx1) / dx) + 1 + sz * G.n;
      if (ax >= G.low1 && ax <= G.low2 && ay >= G.low1 && ay <= G.low2 && az >= G.low1 && az <= G.low2) {
        const int b2 = marca_base[G.ext(ax, ay, az)];
        f = b2 > 0;
        // id, base label at the image of the centre, image shifts (+1 so they fit 2 bits)
        info[I] = make_int4(id, b2, (sx + 1) | ((sy + 1) << 2) | ((sz + 1) << 4), 0);
        vlab[id - 1] = f;
      }
    }
  }
  flag[I] = f;
}

__global__ void k_vp_list(int nv, const int32_t* __restrict__ flag, const int32_t* __restrict__ pos,
                          const int4* __restrict__ info, int4* list) {
  const int I = blockIdx.x * blockDim.x + threadIdx.x;
  if (I < nv && flag[I]) list[pos[I]] = info[I];
}

// the cells of the flagged voids, as (label, ext cell) pairs, compacted with
// one atomic per wave (their order is irrelevant: they are sorted by label next)
__global__ void __launch_bounds__(256) k_vp_cells(int64_t ne3, const int32_t* __restrict__ marca,
                                                 const int32_t* __restrict__ vlab, uint32_t* key, uint32_t* val,
                                                 unsigned* count) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int id = 0;
  bool sel = false;
  if (e < ne3) {
    id = marca[e];
    sel = id > 0 && vlab[id - 1];
  }
  const unsigned long long m = __ballot(sel);
  if (m == 0ull) return;
  const int lane = threadIdx.x & 63, leader = __builtin_ctzll(m);
  unsigned base = 0;
  if (lane == leader) base = atomicAdd(count, (unsigned)__popcll(m));
  base = __shfl(base, leader);
  if (sel) {
    const unsigned r = base + (unsigned)__popcll(m & ((1ull << lane) - 1ull));
    key[r] = (uint32_t)id;
    val[r] = (uint32_t)e;
  }
}

// per flagged void, its overlap test counted before any earlier void of the
// list changes labels: cnt2 = cells whose image lies in the ext grid, and the
// histogram of the base labels met at those images (up to VPT labels, else
// -1).  The ordered pass then needs no recount for a void unless labels were
// merged into it before its turn or its histogram overflowed: the cells
// labelled ind are still its own, and a label's count goes to ind2 when it
// resolves to ind2 (voidfinding.f90:1476-1480).
constexpr int VPT = 16, VPW = 2 + 2 * VPT;  // cnt2, nlab (or -1), labels, counts
__global__ void __launch_bounds__(256) k_vp_count(const int4* __restrict__ list, ExtG G,
                                                 const int32_t* __restrict__ marca_base,
                                                 const uint32_t* __restrict__ cells,
                                                 const int32_t* __restrict__ lstart, const int32_t* __restrict__ lend,
                                                 int32_t* pre) {
  __shared__ int s_c2, s_ovf, s_lab[VPT], s_num[VPT];
  const int L = blockIdx.x, tid = threadIdx.x;
  const int4 in = list[L];
  const int ind = in.x;
  const int sx = (in.z & 3) - 1, sy = ((in.z >> 2) & 3) - 1, sz = ((in.z >> 4) & 3) - 1;
  if (tid == 0) {
    s_c2 = 0;
    s_ovf = 0;
  }
  if (tid < VPT) {
    s_lab[tid] = 0;
    s_num[tid] = 0;
  }
  __syncthreads();
  int c2 = 0;
  const int qe = lend[ind - 1];
  for (int q = lstart[ind - 1] + tid; q < qe; q += blockDim.x) {
    const int cell = (int)cells[q];
    const int i = cell % G.ne + G.low1, j = (cell / G.ne) % G.ne + G.low1, k = cell / (G.ne * G.ne) + G.low1;
    const int a = i + sx * G.n, bb = j + sy * G.n, cc = k + sz * G.n;
    if (a >= G.low1 && a <= G.low2 && bb >= G.low1 && bb <= G.low2 && cc >= G.low1 && cc <= G.low2) {
      c2++;
      const int l = marca_base[G.ext(a, bb, cc)];
      if (l > 0) {  // histogram slot of l (a small open table in LDS)
        int t = 0;
        for (; t < VPT; ++t) {
          const int old = atomicCAS(&s_lab[t], 0, l);
          if (old == 0 || old == l) break;
        }
        if (t < VPT)
          atomicAdd(&s_num[t], 1);
        else
          s_ovf = 1;
      }
    }
  }
  atomicAdd(&s_c2, c2);
  __syncthreads();
  if (tid == 0) {
    int nt = 0;
    while (nt < VPT && s_lab[nt] != 0) ++nt;
    int32_t* o = pre + (size_t)L * VPW;
    o[0] = s_c2;
    o[1] = s_ovf ? -1 : nt;
    for (int t = 0; t < VPT; ++t) {
      o[2 + t] = s_lab[t];
      o[2 + VPT + t] = s_num[t];
    }
  }
}

// the label map for the ordered pass: in LDS as 16-bit entries when the
// labels fit (it is read in every resolve), else the global map itself
struct VpMap {
  uint16_t* s;  // LDS copy, or null
  int32_t* g;
  __device__ __forceinline__ int get(int l) const { return s ? (int)s[l - 1] : g[l - 1]; }
  __device__ __forceinline__ void set(int l, int v) const {
    if (s) s[l - 1] = (uint16_t)v;
    g[l - 1] = v;
  }
  __device__ __forceinline__ int resolve(int l) const {
    while (l > 0 && get(l) != l) l = get(l);
    return l;
  }
};

// CHANGE_BELONG(IND2, IND) when the overlap exceeds 0.4 (voidfinding.f90:
// 1476-1480, 1520-1566): labels of IND resolve to IND2 from now on, IND's
// member list is appended to IND2's, and IND2 is marked as having members.
__device__ __forceinline__ void vp_decide(int ind, int ind2, int cnt, int cnt2, const VpMap& M, int32_t* mnext,
                                          int32_t* mtail, int32_t* selfm, uint32_t* grown, int32_t* stats) {
  const float inter = (cnt2 == 0) ? 0.0f : (float)cnt / (float)cnt2;
  if (!(inter > 0.4f)) return;
  stats[0]++;
  if (ind2 == ind) {
    selfm[ind - 1] = 1;
    return;
  }
  M.set(ind, ind2);
  mnext[mtail[ind2 - 1] - 1] = ind;
  mtail[ind2 - 1] = mtail[ind - 1];
  grown[(ind2 - 1) >> 5] |= 1u << ((ind2 - 1) & 31);
}

constexpr int kVpCh = 32;

// one workgroup walks the flagged main voids in INDICE order (each id occurs
// once).  Thread 0 decides every void whose counts follow from its histogram;
// a void that gained members, or whose histogram overflowed, is counted again
// by the whole workgroup with the labels of now.  Dynamic LDS: the grown
// bitmap ((nv + 32) / 32 words; in grown_g instead when that is given, zeroed)
// and the 16-bit map ((nv + 1) entries) when use_lds.
__global__ void __launch_bounds__(1024) k_void_periodic(
    const int4* __restrict__ list, const int32_t* __restrict__ nlist, int nv, int use_lds, uint32_t* grown_g, ExtG G,
    const int32_t* __restrict__ marca_base, const uint32_t* __restrict__ cells,
    const int32_t* __restrict__ lstart, const int32_t* __restrict__ lend, const int32_t* __restrict__ pre,
    int32_t* lmap, int32_t* mnext, int32_t* mtail, int32_t* selfm, int32_t* stats) {
  extern __shared__ uint32_t dyn[];
  __shared__ int s_cnt, s_cnt2, s_go, s_ind, s_ind2, s_sx, s_sy, s_sz, s_next;
  // the next kVpCh entries of the list and their histograms, staged by the
  // whole workgroup so that thread 0's ordered pass reads LDS, not HBM
  __shared__ int4 s_list[kVpCh];
  __shared__ int32_t s_pre[kVpCh * VPW];
  const int tid = threadIdx.x;
  const int nl = *nlist;
  const int nw = (nv + 32) / 32;
  uint32_t* grown = grown_g ? grown_g : dyn;
  uint16_t* lm16 = use_lds ? (uint16_t*)(grown_g ? dyn : dyn + nw) : nullptr;
  if (!grown_g)
    for (int t = tid; t < nw; t += blockDim.x) grown[t] = 0u;
  if (use_lds)
    for (int t = tid; t < nv; t += blockDim.x) lm16[t] = (uint16_t)lmap[t];
  if (tid == 0) s_next = 0;
  __syncthreads();
  const VpMap M{lm16, lmap};
  for (;;) {
    const int base = s_next;  // set before the last barrier
    if (base >= nl) break;
    const int nch = min(kVpCh, nl - base);
    for (int e = tid; e < nch; e += blockDim.x) s_list[e] = list[base + e];
    for (int e = tid; e < nch * VPW; e += blockDim.x) s_pre[e] = pre[(size_t)base * VPW + e];
    __syncthreads();
    if (tid == 0) {
      int L = base;
      s_go = 0;
      for (; L < base + nch; ++L) {
        const int4 in = s_list[L - base];
        const int id = in.x;
        const int32_t* pr = s_pre + (L - base) * VPW;
        const int ind2 = M.resolve(in.y);
        const bool gr = (grown[(id - 1) >> 5] >> ((id - 1) & 31)) & 1u;
        const int nlab = pr[1];
        if (gr || nlab < 0) {  // count this one again (voidfinding.f90:1476-1480 with the labels of now)
          s_go = 1;
          s_ind = id;
          s_ind2 = ind2;
          s_sx = (in.z & 3) - 1;
          s_sy = ((in.z >> 2) & 3) - 1;
          s_sz = ((in.z >> 4) & 3) - 1;
          break;
        }
        int cnt = 0;
        for (int t = 0; t < nlab; ++t)
          if (M.resolve(pr[2 + t]) == ind2) cnt += pr[2 + VPT + t];
        vp_decide(id, ind2, cnt, pr[0], M, mnext, mtail, selfm, grown, stats);
      }
      s_next = L;
      s_cnt = 0;
      s_cnt2 = 0;
    }
    __syncthreads();
    if (!s_go) continue;  // the chunk is done: stage the next one
    {
      const int ind = s_ind, ind2 = s_ind2;
      int c1 = 0, c2 = 0;
      // cells labelled ind now: base labels of the member list of ind
      for (int b = ind; b > 0; b = mnext[b - 1]) {
        int s = lstart[b - 1], e = lend[b - 1];
        for (int q = s + tid; q < e; q += blockDim.x) {
          const int cell = (int)cells[q];
          const int i = cell % G.ne + G.low1, j = (cell / G.ne) % G.ne + G.low1, k = cell / (G.ne * G.ne) + G.low1;
          int a = i + s_sx * G.n, bb = j + s_sy * G.n, cc = k + s_sz * G.n;
          if (a >= G.low1 && a <= G.low2 && bb >= G.low1 && bb <= G.low2 && cc >= G.low1 && cc <= G.low2) {
            c2++;
            int l = marca_base[G.ext(a, bb, cc)];
            if (l > 0 && M.resolve(l) == ind2) c1++;
          }
        }
        if (b == mtail[ind - 1]) break;
      }
      atomicAdd(&s_cnt, c1);
      atomicAdd(&s_cnt2, c2);
      __syncthreads();
      if (tid == 0) {
        vp_decide(ind, ind2, s_cnt, s_cnt2, M, mnext, mtail, selfm, grown, stats);
        s_next = s_next + 1;
      }
    }
    __syncthreads();
  }
}

__global__ void k_apply_lmap(int nv, const int32_t* __restrict__ lmap, const int32_t* __restrict__ selfm,
                             int32_t* uvoid, float* volnew) {
  int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= nv) return;
  int u = uvoid[v];
  if (u == -1) {
    int r = resolve(lmap, v + 1);
    if (r != v + 1) { uvoid[v] = r; volnew[v] = 0.0f; }
    else if (selfm[v]) { uvoid[v] = v + 1; volnew[v] = 0.0f; }
  } else if (u > 0) {
    uvoid[v] = resolve(lmap, u);
  }
}

__global__ void k_relabel_cells(int64_t ne3, const int32_t* __restrict__ lmap, int32_t* marca) {
  int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= ne3) return;
  int id = marca[e];
  if (id > 0) marca[e] = resolve(lmap, id);
}

__global__ void k_init_lmap(int nv, int32_t* lmap, int32_t* mnext, int32_t* mtail, int32_t* selfm) {
  int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= nv) return;
  lmap[v] = v + 1;
  mnext[v] = 0;
  mtail[v] = v + 1;
  selfm[v] = 0;
}

// final properties over the inner box (avism.f90:1918-1945)
// (grid: x chunks, y, z over the inner grid); a wave whose void cells share
// one id adds once, and the total cell count takes one atomic per wave
__global__ void k_final_props(ExtG G, const int32_t* __restrict__ marca, const int32_t* __restrict__ uvoid,
                              const float* __restrict__ u1co, float mtot_scale_md, float dx, float rete3,
                              int32_t* ncell, double* umean, double* mtot, unsigned long long* ncells_total) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int I = x + 1, J = (int)blockIdx.y + 1, K = (int)blockIdx.z + 1;
  int id = 0;
  float u = 0.0f, m = 0.0f;
  if (x < G.n) {
    id = marca[G.ext(I, J, K)];
    if (id > 0 && uvoid[id - 1] != -1) id = 0;
    if (id > 0) {
      u = u1co[(size_t)(I - 1) + (size_t)G.n * ((size_t)(J - 1) + (size_t)G.n * (K - 1))];
      m = ((((u * mtot_scale_md) * dx) * dx) * dx) * rete3;
    }
  }
  const int w = wave_uniform_id(id);
  if (w == 0) return;
  const unsigned long long inv = __ballot(id > 0);
  const bool lane0 = threadIdx.x % 64 == 0;
  if (lane0) atomicAdd(ncells_total, (unsigned long long)__popcll(inv));
  if (w > 0) {
    const double su = wave_sum((double)u), sm = wave_sum((double)m);
    if (lane0) {
      atomicAdd(&ncell[w - 1], __popcll(inv));
      atomicAdd(&umean[w - 1], su);
      atomicAdd(&mtot[w - 1], sm);
    }
  } else if (id > 0) {
    atomicAdd(&ncell[id - 1], 1);
    atomicAdd(&umean[id - 1], (double)u);
    atomicAdd(&mtot[id - 1], (double)m);
  }
}

__global__ void k_finish_catalogue(int nv, const int32_t* __restrict__ ncell, const double* __restrict__ umean,
                                   const double* __restrict__ mtot, float* fum, float* fmt) {
  int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= nv) return;
  fum[v] = ncell[v] > 0 ? (float)umean[v] / (float)ncell[v] : (float)umean[v];
  fmt[v] = (float)mtot[v];
}

__global__ void k_req(int nv, const int32_t* __restrict__ uvoid, const float* __restrict__ volnew,
                      float* req, int32_t* nvoidt) {
  int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= nv) return;
  if (uvoid[v] != -1) { req[v] = 0.0f; return; }
  atomicAdd(nvoidt, 1);
  double volm = (double)volnew[v];  // avism.f90:2074
  req[v] = (float)pow((3.0 * volm) / (4.0 * kPi), (double)(1.0f / 3.0f));
}

}  // namespace

// vf layout on the device: 12 blocks of nv floats (gxc gyc gzc xc yc zc volnew umean mtot eps ip req)
int merge_and_post(avism_ctx* c, int32_t stop_after) {
  if (!c->grid_set || !c->fields_ready) AV_FAIL(c, AVISM_ERR_STATE, "fields and voids of this grid not computed");
  const int nv = c->nvoid;
  const GridGeom& g = c->g;
  ExtG G{g.n, g.low1, g.low2, g.next, g.periodic};
  const int64_t ne3 = (int64_t)g.next * g.next * g.next;
  c->nvoid2 = 0;
  c->nvoidt = 0;
  c->nperiodic = 0;
  c->volt_clean = 0.0;
  c->post_ready = false;
  if (!c->marca.ensure((size_t)ne3 * 4) || !c->vf.ensure((size_t)12 * (nv + 1) * 4) ||
      !c->indice.ensure((size_t)(nv + 1) * 4) || !c->uvoid.ensure((size_t)(nv + 1) * 4) ||
      !c->ncellv.ensure((size_t)(nv + 1) * 4))
    AV_FAIL(c, AVISM_ERR_HIP, "out of device memory (merge)");
  if (nv == 0) {
    AV_CHECK_HIP(c, hipMemsetAsync(c->marca.p, 0, (size_t)ne3 * 4, c->stream));
    AV_CHECK_HIP(c, hipStreamSynchronize(c->stream));
    c->post_ready = stop_after == 0;
    return AVISM_OK;
  }
  Cubes C{c->cube_i.as<int32_t>(), c->cube_f.as<float>(), c->cube_father.as<int32_t>(), nv};
  float* vf = c->vf.as<float>();
  float *GX = vf, *GY = vf + nv, *GZ = vf + 2 * nv, *XC = vf + 3 * nv, *YC = vf + 4 * nv, *ZC = vf + 5 * nv;
  float *VOLNEW = vf + 6 * nv, *UMEAN = vf + 7 * nv, *MTOT = vf + 8 * nv, *EPS = vf + 9 * nv,
        *IPR = vf + 10 * nv, *REQ = vf + 11 * nv;
  int32_t* indice = c->indice.as<int32_t>();
  int32_t* uvoid = c->uvoid.as<int32_t>();
  int32_t* ncell = c->ncellv.as<int32_t>();
  // scratch per cube (ranks)
  if (!c->rank_of.ensure((size_t)(nv + 1) * 4 * 12) || !c->paintP.ensure((size_t)ne3 * 4) ||
      !c->paintK.ensure((size_t)ne3 * 4) || !c->ev_key.ensure((size_t)(nv + 1) * 8) ||
      !c->ev_key2.ensure((size_t)(nv + 1) * 8) || !c->ev_val.ensure((size_t)(nv + 1) * 4) ||
      !c->ev_val2.ensure((size_t)(nv + 1) * 4) || !c->inert.ensure((size_t)(nv + 1) * 6 * 8 * 3))
    AV_FAIL(c, AVISM_ERR_HIP, "out of device memory (merge scratch)");
  int32_t* R = c->rank_of.as<int32_t>();
  int32_t *rank = R, *jstar = R + (nv + 1), *keyv = R + 2 * (nv + 1), *par = R + 3 * (nv + 1),
          *par2 = R + 4 * (nv + 1), *gstart = R + 5 * (nv + 1), *gend = R + 6 * (nv + 1),
          *removed = R + 7 * (nv + 1), *asc = R + 8 * (nv + 1), *lmap = R + 9 * (nv + 1),
          *mnext = R + 10 * (nv + 1), *mtail = R + 11 * (nv + 1);
  float* cvol = (float*)c->ev_val2.p;  // own cube volumes (reused below as key scratch)
  int32_t* small = (int32_t*)((char*)c->scalars.ensure(4096) + 1024);  // counters
  const unsigned W = 4;  // waves per block in wave-per-cube kernels
  // ---- volume order
  {
    StageTimer t(c, "volume_sort");
    float* volk = (float*)c->ev_key.p;
    k_vol_keys<<<nbk(nv), 256, 0, c->stream>>>(nv, C, volk);
    int rc = indexx_device(c, volk, nv, asc);
    if (rc) return rc;
    k_indice<<<nbk(nv), 256, 0, c->stream>>>(nv, asc, indice, rank);
    AV_CHECK_HIP(c, hipMemsetAsync(small, 0, 64, c->stream));
    k_check_father<<<nbk(nv), 256, 0, c->stream>>>(nv, C, small);
  }
  int32_t hf = 0;  // cubes of several parent voids
  AV_CHECK_HIP(c, hipMemcpyAsync(&hf, small, 4, hipMemcpyDeviceToHost, c->stream));
  AV_CHECK_HIP(c, hipStreamSynchronize(c->stream));
  int32_t* P = c->paintP.as<int32_t>();
  int32_t* K = c->paintK.as<int32_t>();
  int32_t* F = nullptr;
  if (hf) {
    if (!c->cover_acc.ensure((size_t)ne3 * 4)) AV_FAIL(c, AVISM_ERR_HIP, "out of device memory (merge)");
    F = c->cover_acc.as<int32_t>();
  }
  {
    StageTimer t(c, "merge");
    k_fill<<<nbk(ne3), 256, 0, c->stream>>>(ne3, P, 0x7fffffff);
    k_fill<<<nbk(ne3), 256, 0, c->stream>>>(ne3, K, 0x7fffffff);
    k_paint_box<<<nbk(nv, W), 64 * W, 0, c->stream>>>(nv, G, C, indice, nullptr, 2, P);
    if (F) {
      k_fill<<<nbk(ne3), 256, 0, c->stream>>>(ne3, F, kFatherEmpty);
      k_paint_father<<<nbk(nv, W), 64 * W, 0, c->stream>>>(nv, G, C, F);
    }
    k_jstar<<<nbk(nv, W), 64 * W, 0, c->stream>>>(nv, G, C, indice, P, F, jstar, keyv, small + 2);
    AV_CHECK_HIP(c, hipGetLastError());
    if (F) {
      int32_t namb = 0;
      AV_CHECK_HIP(c, hipMemcpyAsync(&namb, small + 2, 4, hipMemcpyDeviceToHost, c->stream));
      AV_CHECK_HIP(c, hipStreamSynchronize(c->stream));
      if (namb > 0) {
        const int nbx = (g.next + kBkt - 1) / kBkt;
        const int64_t nbt = (int64_t)nbx * nbx * nbx;
        if (!c->bkt.ensure((size_t)(nbt * 2 + 4 * (int64_t)(nv + 1)) * 4))
          AV_FAIL(c, AVISM_ERR_HIP, "out of device memory (merge buckets)");
        int32_t* bstart = c->bkt.as<int32_t>();
        int32_t* bend = bstart + nbt;
        uint32_t* bkey = (uint32_t*)(bend + nbt);
        uint32_t* bkey2 = bkey + (nv + 1);
        int32_t* bval = (int32_t*)(bkey2 + (nv + 1));
        int32_t* bval2 = bval + (nv + 1);
        AV_CHECK_HIP(c, hipMemsetAsync(small + 5, 0, 4, c->stream));
        k_bucket_keys<<<nbk(nv), 256, 0, c->stream>>>(nv, G, C, indice, nbx, bkey, bval, small + 5);
        int bits = 1;
        while (bits < 32 && ((uint64_t)1 << bits) <= (uint64_t)nbt + 1) ++bits;
        size_t tmp = 0;
        AV_CHECK_HIP(c, rocprim::radix_sort_pairs(nullptr, tmp, bkey, bkey2, bval, bval2, (size_t)nv, 0, bits,
                                                  c->stream));
        if (!c->sort_tmp.ensure(tmp + 256)) AV_FAIL(c, AVISM_ERR_HIP, "out of device memory");
        AV_CHECK_HIP(c, rocprim::radix_sort_pairs(c->sort_tmp.p, tmp, bkey, bkey2, bval, bval2, (size_t)nv, 0,
                                                  bits, c->stream));
        AV_CHECK_HIP(c, hipMemsetAsync(bstart, 0, (size_t)nbt * 8, c->stream));
        k_label_ranges<<<nbk(nv), 256, 0, c->stream>>>(nv, bkey2, bstart, bend);
        k_jstar_enumerate<<<nbk(nv, W), 64 * W, 0, c->stream>>>(nv, G, C, indice, nbx, bstart, bend, bval2,
                                                                small + 5, jstar, keyv);
        AV_CHECK_HIP(c, hipGetLastError());
      }
    }
    k_paint_box<<<nbk(nv, W), 64 * W, 0, c->stream>>>(nv, G, C, indice, keyv, 0, K);
    k_parent<<<nbk(nv, W), 64 * W, 0, c->stream>>>(nv, G, C, indice, jstar, K, hf ? 1 : 0, par);
    int32_t* a = par;
    int32_t* b = par2;
    for (int it = 0; it < 64; ++it) {
      AV_CHECK_HIP(c, hipMemsetAsync(small, 0, 4, c->stream));
      k_jump<<<nbk(nv), 256, 0, c->stream>>>(nv, a, b, small);
      int32_t ch = 0;
      AV_CHECK_HIP(c, hipMemcpyAsync(&ch, small, 4, hipMemcpyDeviceToHost, c->stream));
      AV_CHECK_HIP(c, hipStreamSynchronize(c->stream));
      std::swap(a, b);
      if (!ch) break;
    }
    int32_t* grp = a;
    AV_CHECK_HIP(c, hipMemsetAsync(small + 1, 0, 4, c->stream));
    k_uvoid<<<nbk(nv), 256, 0, c->stream>>>(nv, indice, grp, uvoid, small + 1);
    k_marca<<<nbk(ne3), 256, 0, c->stream>>>(ne3, K, indice, grp, c->marca.as<int32_t>());
    // GXC / VOLNEW fold in the reference's event order
    float *cx = XC, *cy = YC, *cz = ZC;  // original centres (XC.. rewritten later)
    k_centres<<<nbk(nv), 256, 0, c->stream>>>(nv, C, g.dx, g.rx1, cx, cy, cz, cvol);
    uint64_t* k1 = (uint64_t*)c->ev_key.p;    // by rank
    uint64_t* k2 = (uint64_t*)c->ev_key2.p;   // by rank
    uint64_t* kb = (uint64_t*)c->inert.p;     // sorted keys scratch
    uint64_t* kc = kb + (nv + 1);
    int32_t* ranks = c->ev_val.as<int32_t>();
    int32_t* rb = (int32_t*)(kc + (nv + 1));
    int32_t* rs = rb + (nv + 1);
    k_events<<<nbk(nv), 256, 0, c->stream>>>(nv, indice, jstar, grp, cx, cy, cz, g.lado0plus, g.periodic,
                                             k1, k2, ranks);
    size_t tmp = 0, tmp2 = 0;
    AV_CHECK_HIP(c, rocprim::radix_sort_pairs(nullptr, tmp, k2, kb, ranks, rb, (size_t)nv, 0, 64, c->stream));
    AV_CHECK_HIP(c, rocprim::radix_sort_pairs(nullptr, tmp2, kb, kc, rb, rs, (size_t)nv, 0, 64, c->stream));
    if (!c->sort_tmp.ensure(std::max(tmp, tmp2) + 256)) AV_FAIL(c, AVISM_ERR_HIP, "out of device memory");
    AV_CHECK_HIP(c, rocprim::radix_sort_pairs(c->sort_tmp.p, tmp, k2, kb, ranks, rb, (size_t)nv, 0, 64,
                                              c->stream));
    k_gather_u64<<<nbk(nv), 256, 0, c->stream>>>(nv, k1, rb, kb);
    AV_CHECK_HIP(c, rocprim::radix_sort_pairs(c->sort_tmp.p, tmp2, kb, kc, rb, rs, (size_t)nv, 0, 64,
                                              c->stream));
    AV_CHECK_HIP(c, hipMemsetAsync(gstart, 0, (size_t)(nv + 1) * 4, c->stream));
    AV_CHECK_HIP(c, hipMemsetAsync(gend, 0, (size_t)(nv + 1) * 4, c->stream));
    k_group_ranges<<<nbk(nv), 256, 0, c->stream>>>(nv, kc, gstart, gend);
    k_fold<<<nbk(nv), 256, 0, c->stream>>>(nv, indice, grp, gstart, gend, rs, cx, cy, cz, cvol, GX, GY, GZ,
                                           VOLNEW);
    k_copy_nonleaders<<<nbk(nv), 256, 0, c->stream>>>(nv, indice, grp, cx, cy, cz, cvol, GX, GY, GZ, VOLNEW);
    AV_CHECK_HIP(c, hipGetLastError());
  }
  int32_t hmaj = 0;
  AV_CHECK_HIP(c, hipMemcpyAsync(&hmaj, small + 1, 4, hipMemcpyDeviceToHost, c->stream));
  AV_CHECK_HIP(c, hipStreamSynchronize(c->stream));
  c->nvoid2 = hmaj;
  if (stop_after == 1) return AVISM_OK;

  int32_t* marca = c->marca.as<int32_t>();
  // ---- NCELL_MIN filter (avism.f90:1796-1827)
  {
    StageTimer t(c, "post_filters");
    AV_CHECK_HIP(c, hipMemsetAsync(ncell, 0, (size_t)(nv + 1) * 4, c->stream));
    k_count_cells<<<dim3((unsigned)((G.ne + 255) / 256), (unsigned)G.ne, (unsigned)G.ne), 256, 0, c->stream>>>(
        G, marca, ncell);
    AV_CHECK_HIP(c, hipMemsetAsync(small + 7, 0, 4, c->stream));
    k_mark_removed<<<nbk(nv), 256, 0, c->stream>>>(nv, uvoid, ncell, 0, g.dx, 0.0, removed, VOLNEW, small + 7);
    k_apply_removed_cubes<<<nbk(nv), 256, 0, c->stream>>>(nv, removed, uvoid, ncell, VOLNEW);
    k_apply_removed_cells<<<4096, 256, 0, c->stream>>>(ne3, removed, small + 7, marca);
  }
  // ---- VOID_SHAPE (avism.f90:1831)
  float* d_rad = (float*)((char*)c->scalars.p + 4096);
  {
    StageTimer t(c, "shape");
    AV_CHECK_HIP(c, hipMemcpyAsync(d_rad, g.rad.data(), g.rad.size() * 4, hipMemcpyHostToDevice, c->stream));
    double* inert = (double*)c->inert.p;
    AV_CHECK_HIP(c, hipMemsetAsync(inert, 0, (size_t)(nv + 1) * 6 * 8, c->stream));
    k_inertia<<<dim3((unsigned)((G.ne + 255) / 256), (unsigned)G.ne, (unsigned)G.ne), 256, 0, c->stream>>>(
        G, d_rad, g.rlo, marca, GX, GY, GZ, inert);
    k_shape<<<nbk(nv), 256, 0, c->stream>>>(nv, uvoid, ncell, inert, g.dx, EPS, IPR);
  }
  // ---- VOLNEW from NCELLV + RMIN filter (1834-1859), centres XC (1863-1870)
  double rmin3 = (double)(4.0f / 3.0f) * kPi * (double)((c->vp.rmin * c->vp.rmin) * c->vp.rmin);
  {
    StageTimer t(c, "post_filters");
    AV_CHECK_HIP(c, hipMemsetAsync(VOLNEW, 0, (size_t)nv * 4, c->stream));
    AV_CHECK_HIP(c, hipMemsetAsync(small + 7, 0, 4, c->stream));
    k_mark_removed<<<nbk(nv), 256, 0, c->stream>>>(nv, uvoid, ncell, 1, g.dx, rmin3, removed, VOLNEW, small + 7);
    k_apply_removed_cubes<<<nbk(nv), 256, 0, c->stream>>>(nv, removed, uvoid, ncell, VOLNEW);
    k_apply_removed_cells<<<4096, 256, 0, c->stream>>>(ne3, removed, small + 7, marca);
    k_xc<<<nbk(nv), 256, 0, c->stream>>>(nv, C, uvoid, g.dx, g.rx1, XC, YC, ZC);
  }
  if (stop_after == 2) {
    AV_CHECK_HIP(c, hipStreamSynchronize(c->stream));
    return AVISM_OK;
  }
  // ---- SIMPLE_CONNECTION_FAST (1877)
  {
    StageTimer t(c, "hole_fill");
    int32_t* lab = P;  // reuse
    const int nbc = (G.ne + CCB - 1) / CCB;
    uint8_t* isrep = c->paintK.as<uint8_t>();  // free after the merge
    k_cc_local<<<(unsigned)((int64_t)nbc * nbc * nbc), 256, 0, c->stream>>>(G, marca, lab, isrep);
    const dim3 g3((unsigned)((G.ne + 255) / 256), (unsigned)G.ne, (unsigned)G.ne);
    k_cc_faces<<<(unsigned)(((int64_t)nbc * nbc * nbc * 3 + kFaceW - 1) / kFaceW), 64 * kFaceW, 0, c->stream>>>(
        G, nbc, lab);
    k_cc_flatten<<<nbk(ne3), 256, 0, c->stream>>>(ne3, isrep, lab);
    AV_CHECK_HIP(c, hipMemsetAsync(small + 2, 0, 4, c->stream));
    k_fill_holes<<<g3, 256, 0, c->stream>>>(G, lab, marca, small + 2);
    AV_CHECK_HIP(c, hipGetLastError());
  }
  if (stop_after == 3) {
    AV_CHECK_HIP(c, hipStreamSynchronize(c->stream));
    return AVISM_OK;
  }
  // ---- VOID_PERIODIC (1892-1910)
  if (g.periodic) {
    StageTimer t(c, "void_periodic");
    int32_t* has = gstart;
    AV_CHECK_HIP(c, hipMemsetAsync(has, 0, (size_t)(nv + 1) * 4, c->stream));
    k_inner_presence<<<nbk(ne3), 256, 0, c->stream>>>(ne3, G, marca, has);
    AV_CHECK_HIP(c, hipMemsetAsync(small + 7, 0, 4, c->stream));
    k_absent_removed<<<nbk(nv), 256, 0, c->stream>>>(nv, uvoid, has, removed, small + 7);
    k_apply_removed_cubes<<<nbk(nv), 256, 0, c->stream>>>(nv, removed, uvoid, ncell, VOLNEW);
    k_apply_removed_cells<<<4096, 256, 0, c->stream>>>(ne3, removed, small + 7, marca);
    int32_t* selfm = removed;
    k_init_lmap<<<nbk(nv), 256, 0, c->stream>>>(nv, lmap, mnext, mtail, selfm);
    AV_CHECK_HIP(c, hipMemsetAsync(small + 3, 0, 8, c->stream));  // [3] merges, [4] cell pairs
    // the statically possible images, compacted in INDICE order, and a flag per label
    // per void (nv + 1 entries each): flag, scan position, label flag, counts
    // (VPW ints), then 16-byte aligned the per-rank info and the compacted list
    const size_t nv1 = (size_t)nv + 1;
    const size_t ints = (nv1 * (3 + VPW) + 3) & ~(size_t)3;
    if (!c->cellkey.ensure((size_t)ne3 * 4) || !c->cellkey2.ensure((size_t)ne3 * 4) ||
        !c->cellid.ensure((size_t)ne3 * 4) || !c->cellid2.ensure((size_t)ne3 * 4) ||
        !c->per_tmp.ensure(ints * 4 + nv1 * 32))
      AV_FAIL(c, AVISM_ERR_HIP, "out of device memory (periodic)");
    int32_t* vflag = c->per_tmp.as<int32_t>();
    int32_t* vpos = vflag + nv1;
    int32_t* vlab = vpos + nv1;
    int32_t* pre = vlab + nv1;
    int4* vinfo = (int4*)(vflag + ints);
    int4* vl = vinfo + nv1;
    AV_CHECK_HIP(c, hipMemsetAsync(vlab, 0, (size_t)(nv + 1) * 4, c->stream));
    k_vp_flag<<<nbk(nv), 256, 0, c->stream>>>(nv, G, indice, uvoid, marca, GX, GY, GZ, g.L0 / 2.0f, g.dx, g.rx1,
                                              vflag, vinfo, vlab);
    AV_CHECK_HIP(c, hipMemsetAsync(vflag + nv, 0, 4, c->stream));
    size_t tmp2 = 0;
    AV_CHECK_HIP(c, rocprim::exclusive_scan(nullptr, tmp2, vflag, vpos, 0, (size_t)nv + 1,
                                            rocprim::plus<int32_t>(), c->stream));
    if (!c->sort_tmp.ensure(tmp2)) AV_FAIL(c, AVISM_ERR_HIP, "out of device memory (periodic)");
    AV_CHECK_HIP(c, rocprim::exclusive_scan(c->sort_tmp.p, tmp2, vflag, vpos, 0, (size_t)nv + 1,
                                            rocprim::plus<int32_t>(), c->stream));
    k_vp_list<<<nbk(nv), 256, 0, c->stream>>>(nv, vflag, vpos, vinfo, vl);
    // cell lists of the flagged voids (at most ne3 pairs), sorted by label
    uint32_t* ka = c->cellkey.as<uint32_t>();
    uint32_t* kb2 = c->cellkey2.as<uint32_t>();
    uint32_t* va = c->cellid.as<uint32_t>();
    uint32_t* vb = c->cellid2.as<uint32_t>();
    k_vp_cells<<<nbk(ne3), 256, 0, c->stream>>>(ne3, marca, vlab, ka, va, (unsigned*)(small + 4));
    int32_t hcnt[2] = {0, 0};  // flagged voids, their cells
    AV_CHECK_HIP(c, hipMemcpyAsync(&hcnt[0], vpos + nv, 4, hipMemcpyDeviceToHost, c->stream));
    AV_CHECK_HIP(c, hipMemcpyAsync(&hcnt[1], small + 4, 4, hipMemcpyDeviceToHost, c->stream));
    AV_CHECK_HIP(c, hipStreamSynchronize(c->stream));
    const int nl = hcnt[0];
    const int64_t mcell = hcnt[1];
    if (nl > 0 && mcell > 0) {
      unsigned bits = 1;
      while ((1ll << bits) <= (int64_t)nv) bits++;
      size_t tmp = 0;
      AV_CHECK_HIP(c, rocprim::radix_sort_pairs(nullptr, tmp, ka, kb2, va, vb, (size_t)mcell, 0, bits, c->stream));
      if (!c->sort_tmp.ensure(tmp)) AV_FAIL(c, AVISM_ERR_HIP, "out of device memory (periodic sort)");
      AV_CHECK_HIP(c, rocprim::radix_sort_pairs(c->sort_tmp.p, tmp, ka, kb2, va, vb, (size_t)mcell, 0, bits,
                                                c->stream));
      int32_t* lstart = gend;
      int32_t* lend = par2;
      AV_CHECK_HIP(c, hipMemsetAsync(lstart, 0, (size_t)(nv + 1) * 4, c->stream));
      AV_CHECK_HIP(c, hipMemsetAsync(lend, 0, (size_t)(nv + 1) * 4, c->stream));
      k_label_ranges<<<nbk(mcell), 256, 0, c->stream>>>(mcell, kb2, lstart, lend);
      k_vp_count<<<(unsigned)nl, 256, 0, c->stream>>>(vl, G, marca, vb, lstart, lend, pre);
      // LDS (at most 56 KB): the grown bitmap, and the 16-bit label map below
      // 65535 labels; what does not fit stays in HBM
      const size_t words = (size_t)(nv + 32) / 32, map16 = (size_t)(nv + 1) * 2;
      const bool bits_lds = words * 4 <= 56 * 1024;
      const bool use_lds = nv < 65535 && (bits_lds ? words * 4 : 0) + map16 <= 56 * 1024;
      const size_t dyn = (bits_lds ? words * 4 : 0) + (use_lds ? map16 : 0);
      uint32_t* grown_g = nullptr;
      if (!bits_lds) {
        grown_g = (uint32_t*)c->cellkey.p;  // the pair keys are sorted into cellkey2 by now
        AV_CHECK_HIP(c, hipMemsetAsync(grown_g, 0, words * 4, c->stream));
      }
      k_void_periodic<<<1, 1024, dyn, c->stream>>>(vl, vpos + nv, nv, use_lds ? 1 : 0, grown_g, G, marca, vb,
                                                    lstart, lend, pre, lmap, mnext, mtail, selfm, small + 3);
    }
    k_apply_lmap<<<nbk(nv), 256, 0, c->stream>>>(nv, lmap, selfm, uvoid, VOLNEW);
    k_relabel_cells<<<nbk(ne3), 256, 0, c->stream>>>(ne3, lmap, marca);
    AV_CHECK_HIP(c, hipGetLastError());
  }
  // ---- final properties over the inner box (1918-1973)
  {
    StageTimer t(c, "final_props");
    double* um = (double*)c->inert.p;
    double* mt = um + (nv + 1);
    unsigned long long* tot = (unsigned long long*)((char*)c->scalars.p + 2048);
    AV_CHECK_HIP(c, hipMemsetAsync(ncell, 0, (size_t)(nv + 1) * 4, c->stream));
    AV_CHECK_HIP(c, hipMemsetAsync(um, 0, (size_t)(nv + 1) * 16, c->stream));
    AV_CHECK_HIP(c, hipMemsetAsync(tot, 0, 8, c->stream));
    float rete3 = (c->rete * c->rete) * c->rete;
    k_final_props<<<dim3((unsigned)((g.n + 255) / 256), (unsigned)g.n, (unsigned)g.n), 256, 0, c->stream>>>(
        G, marca, uvoid, c->u1co.as<float>(), c->meandens, g.dx, rete3, ncell, um, mt, tot);
    k_finish_catalogue<<<nbk(nv), 256, 0, c->stream>>>(nv, ncell, um, mt, UMEAN, MTOT);
    AV_CHECK_HIP(c, hipMemsetAsync(VOLNEW, 0, (size_t)nv * 4, c->stream));
    AV_CHECK_HIP(c, hipMemsetAsync(small + 7, 0, 4, c->stream));
    k_mark_removed<<<nbk(nv), 256, 0, c->stream>>>(nv, uvoid, ncell, 1, g.dx, rmin3, removed, VOLNEW, small + 7);
    k_apply_removed_cubes<<<nbk(nv), 256, 0, c->stream>>>(nv, removed, uvoid, ncell, VOLNEW);
    k_apply_removed_cells<<<4096, 256, 0, c->stream>>>(ne3, removed, small + 7, marca);
    AV_CHECK_HIP(c, hipMemsetAsync(small + 4, 0, 4, c->stream));
    k_req<<<nbk(nv), 256, 0, c->stream>>>(nv, uvoid, VOLNEW, REQ, small + 4);
    AV_CHECK_HIP(c, hipGetLastError());
    unsigned long long ht = 0;
    int32_t hn = 0, hp = 0;
    AV_CHECK_HIP(c, hipMemcpyAsync(&ht, tot, 8, hipMemcpyDeviceToHost, c->stream));
    AV_CHECK_HIP(c, hipMemcpyAsync(&hn, small + 4, 4, hipMemcpyDeviceToHost, c->stream));
    if (g.periodic) AV_CHECK_HIP(c, hipMemcpyAsync(&hp, small + 3, 4, hipMemcpyDeviceToHost, c->stream));
    AV_CHECK_HIP(c, hipStreamSynchronize(c->stream));
    c->nvoidt = hn;
    c->nperiodic = hp;
    c->volt_clean = (double)ht * (double)((g.dx * g.dx) * g.dx);
  }
  c->post_ready = true;
  return AVISM_OK;
}

namespace {
// MARCAP(X,Y,Z) = MARCA(INT((X-1)/2)+1, ...) on the grid of 2N (avism.f90:2148-2156);
// grid (x chunks, y, z) over 2N
__global__ void k_upsample_marca(ExtG G, const int32_t* __restrict__ marca, int32_t* marcap) {
  const int n2 = 2 * G.n;
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= n2) return;
  const int y = blockIdx.y, z = blockIdx.z;
  marcap[(size_t)x + (size_t)n2 * ((size_t)y + (size_t)n2 * z)] = marca[G.ext(x / 2 + 1, y / 2 + 1, z / 2 + 1)];
}

// SMASK of the finer grid from the coarser one (avism.f90:875-889, NDOWN = 2)
__global__ void k_upsample_mask(int n, const int8_t* __restrict__ m, int8_t* m2) {
  const int n2 = 2 * n;
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= n2) return;
  const int y = blockIdx.y, z = blockIdx.z;
  m2[(size_t)x + (size_t)n2 * ((size_t)y + (size_t)n2 * z)] =
      m[(size_t)(x / 2) + (size_t)n * ((size_t)(y / 2) + (size_t)n * (z / 2))];
}

// REQP(IND) = REQ(IND) for the main voids (avism.f90:2140-2144); 0 for the rest
__global__ void k_reqp(int nv, const int32_t* __restrict__ uvoid, const float* __restrict__ req, float* reqp) {
  int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v < nv) reqp[v] = uvoid[v] == -1 ? req[v] : 0.0f;
}
}  // namespace

// hand-off to the next void-in-void level (avism.f90:2123-2161, 742-759):
// MARCAP and REQP from this level's catalogue, then the grid of 2N cells per
// axis over the same box.  The particles stay loaded, so the next level can
// interpolate them at its own resolution.
int next_level(avism_ctx* c) {
  if (!c->post_ready) AV_FAIL(c, AVISM_ERR_STATE, "avism_merge_and_post must complete before the next level");
  const GridGeom g = c->g;
  const int n2 = 2 * g.n;
  GridGeom g2 = make_grid(n2, g.L0, g.periodic);
  if ((int64_t)g2.next * g2.next * g2.next >= (int64_t)0x7fffffff)
    AV_FAIL(c, AVISM_ERR_LIMIT, "extended grid of the next level above 2^31 cells is not supported by this build");
  const int nv = c->nvoid;
  if (!c->marcap.ensure((size_t)n2 * n2 * n2 * 4) || !c->reqp.ensure((size_t)(nv + 1) * 4))
    AV_FAIL(c, AVISM_ERR_HIP, "out of device memory (next level)");
  ExtG G{g.n, g.low1, g.low2, g.next, g.periodic};
  k_upsample_marca<<<dim3((unsigned)((n2 + 255) / 256), (unsigned)n2, (unsigned)n2), 256, 0, c->stream>>>(
      G, c->marca.as<int32_t>(), c->marcap.as<int32_t>());
  if (nv > 0)
    k_reqp<<<nbk(nv), 256, 0, c->stream>>>(nv, c->uvoid.as<int32_t>(), c->vf.as<float>() + 11 * (size_t)nv,
                                           c->reqp.as<float>());
  if (c->has_mask) {  // the survey mask at the finer resolution (the old one is no longer needed)
    DBuf m2;
    if (!m2.ensure((size_t)n2 * n2 * n2)) AV_FAIL(c, AVISM_ERR_HIP, "out of device memory (next level)");
    k_upsample_mask<<<dim3((unsigned)((n2 + 255) / 256), (unsigned)n2, (unsigned)n2), 256, 0, c->stream>>>(
        g.n, c->smask.as<int8_t>(), m2.as<int8_t>());
    AV_CHECK_HIP(c, hipGetLastError());
    AV_CHECK_HIP(c, hipStreamSynchronize(c->stream));
    std::swap(c->smask.p, m2.p);
    std::swap(c->smask.bytes, m2.bytes);
    c->mask_observed *= 8;  // each observed cell becomes 2 x 2 x 2
  }
  AV_CHECK_HIP(c, hipGetLastError());
  AV_CHECK_HIP(c, hipStreamSynchronize(c->stream));
  c->g = g2;
  if (!c->scalars.ensure(scalars_bytes(c))) AV_FAIL(c, AVISM_ERR_HIP, "out of device memory (scalars)");
  c->has_marcap = true;
  c->has_reqp = true;
  c->nreqp = nv;
  c->fields_ready = false;
  c->post_ready = false;
  c->nvoid = 0;
  c->ncand = 0;
  return AVISM_OK;
}

int get_catalogue(avism_ctx* c, int32_t* indice, int32_t* uvoid, int32_t* ncellv, float* vf,
                  avism_level_stats* st) {
  int nv = c->nvoid;
  if (nv > 0) {
    if (indice) AV_CHECK_HIP(c, hipMemcpy(indice, c->indice.p, (size_t)nv * 4, hipMemcpyDeviceToHost));
    if (uvoid) AV_CHECK_HIP(c, hipMemcpy(uvoid, c->uvoid.p, (size_t)nv * 4, hipMemcpyDeviceToHost));
    if (ncellv) AV_CHECK_HIP(c, hipMemcpy(ncellv, c->ncellv.p, (size_t)nv * 4, hipMemcpyDeviceToHost));
    if (vf) AV_CHECK_HIP(c, hipMemcpy(vf, c->vf.p, (size_t)12 * nv * 4, hipMemcpyDeviceToHost));
  }
  if (st) {
    st->ncand = c->ncand;
    st->ncubes = c->nvoid;
    st->nvoid2 = c->nvoid2;
    st->nvoidt = c->nvoidt;
    st->nperiodic = c->nperiodic;
    st->volt_clean = c->volt_clean;
    st->meandens = c->meandens;
    float l3 = (c->g.L0 * c->g.L0) * c->g.L0;
    st->ff = (float)(c->volt_clean / (double)l3);
  }
  return AVISM_OK;
}

int get_marca(avism_ctx* c, int32_t* marca, int32_t inner_only, int32_t where) {
  const GridGeom& g = c->g;
  const int64_t ne3 = (int64_t)g.next * g.next * g.next;
  if (!c->marca.p || !c->post_ready) AV_FAIL(c, AVISM_ERR_STATE, "no MARCA computed (run avism_merge_and_post)");
  hipMemcpyKind kind = where == AVISM_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
  if (!inner_only) {
    AV_CHECK_HIP(c, hipMemcpy(marca, c->marca.p, (size_t)ne3 * 4, kind));
    return AVISM_OK;
  }
  // inner box = the n^3 ext cells from (1,1,1): one strided 3-D copy
  const int n = g.n, off = 1 - g.low1;
  hipMemcpy3DParms p{};
  p.srcPtr = make_hipPitchedPtr(c->marca.p, (size_t)g.next * 4, (size_t)g.next, (size_t)g.next);
  p.srcPos = make_hipPos((size_t)off * 4, (size_t)off, (size_t)off);
  p.dstPtr = make_hipPitchedPtr(marca, (size_t)n * 4, (size_t)n, (size_t)n);
  p.dstPos = make_hipPos(0, 0, 0);
  p.extent = make_hipExtent((size_t)n * 4, (size_t)n, (size_t)n);
  p.kind = kind;
  AV_CHECK_HIP(c, hipMemcpy3D(&p));
  return AVISM_OK;
}

}  // namespace avism
