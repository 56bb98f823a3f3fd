// fields.hip -- particle -> grid stage on gfx950.
//
// Reference: particles.f90 (PPART_DENS 742-802, VVEL_INTERP_SPH_VW 806-930,
// KERNEL_FUNC 935-963, FILL_H_ZEROS 398-431, H_DISTANCE 276-351,
// DDENS_INTERP_SPH 434-618), divergence.f90:2-50 (DIVER_UNIFORM) and the mean
// density block avism.f90:1249-1250, 1371-1414.
//
// Design (MI355X):
//  * particles are binned into a uniform grid of cells (radix sort by bin, then
//    float4 records in bin order), so a row of bins is one contiguous range;
//    each kNN query (one lane) finds its k-th distance with a ring sweep of
//    rows and a 32-bit selection in LDS, then sweeps that exact ball to hand
//    its neighbours to the consumer (knn_search.h).
//  * distances, kernel weights and the DDENS normalisation are evaluated in the
//    reference's FP32 operation order (compiled without FMA contraction), so
//    neighbour sets, HPART and NORM are bit-identical to the reference; the
//    PPART and VVEL sums are FP64 as there.
//  * DDENS deposits go into int64 fixed-point cell sums through an LDS copy of
//    each block's tile (integer atomics: the sums do not depend on the order).
//    HPART uses integer atomicMax on the float bits (order-free).
#include <climits>
#include <cstring>
#include <algorithm>
#include <cmath>
#include <type_traits>

#include <rocprim/rocprim.hpp>

#include "avism_internal.h"
#include "fields_batch.h"
#include "knn_search.h"
#include "refine.h"

namespace avism {

// ------------------------------------------------------------------ host bits
DBuf::~DBuf() { release(); }
void DBuf::release() {
  if (p) (void)hipFree(p);
  p = nullptr;
  bytes = 0;
}
void* DBuf::ensure(size_t nbytes) {
  if (nbytes == 0) nbytes = 256;
  if (nbytes <= bytes) return p;
  release();
  if (hipMalloc(&p, nbytes) != hipSuccess) {
    p = nullptr;
    bytes = 0;
    return nullptr;
  }
  bytes = nbytes;
  return p;
}

GridGeom make_grid(int32_t n, float L0, int32_t periodic) {
  GridGeom g{};
  int32_t nplus = 0, low1 = 1, low2 = n;
  float lado0plus = L0;
  if (periodic) {  // avism.f90:789-799
    float sizing = 50.0f;
    float dxp = L0 / (float)n;
    if (L0 / 2.0f < sizing) sizing = L0 / 2.0f;
    int32_t a = n / 8, b = (int32_t)(sizing / dxp);
    nplus = a > b ? a : b;
    low1 = -nplus + 1;
    low2 = n + nplus;
    lado0plus = L0 + (2.0f * (float)nplus) * dxp;
  }
  g.n = n; g.periodic = periodic; g.nplus = nplus; g.low1 = low1; g.low2 = low2;
  g.next = low2 - low1 + 1; g.L0 = L0; g.lado0plus = lado0plus;
  g.rlo = std::min(low1, 0);
  g.rhi = std::max(low2, n + 1);
  g.rad.assign((size_t)(g.rhi - g.rlo + 1), 0.0f);
  auto R = [&](int i) -> float& { return g.rad[(size_t)(i - g.rlo)]; };
  // MALLA (grid.f90:3-76) with LADO = LADO0 - LADO0/N (avism.f90:817)
  float lado = L0 - (L0 / (float)n);
  float A = -lado / 2.0f, B = lado / 2.0f;
  float C = (B - A) / (float)(n - 1);
  R(1) = A;
  for (int i = 2; i <= n; i++) R(i) = R(1) + (float)(i - 1) * C;
  R(0) = R(1) - C;
  R(n + 1) = R(n) + C;
  g.dx = R(2) - R(1);
  g.rx1 = R(1);
  if (periodic) {  // avism.f90:837-861
    for (int i = low1; i <= low2; i++) {
      if (i < 1) R(i) = R(1) - (float)(1 - i) * g.dx;
      else if (i > n) R(i) = R(n) + (float)(i - n) * g.dx;
    }
  }
  return g;
}

void add_timing(avism_ctx* c, const char* name, hipEvent_t a, hipEvent_t b) {
  c->timings.push_back({name, a, b});
}

// ------------------------------------------------------------ device helpers
// M4 spline (particles.f90:943-950 and 525-533)
__device__ __forceinline__ float m4(float q) {
  if (q <= 1.0f) return 1.0f - (1.5f * (q * q)) * (1.0f - 0.5f * q);
  if (q <= 2.0f) {
    float t = 2.0f - q;
    return 0.25f * ((t * t) * t);
  }
  return 0.0f;
}

struct Kahan {
  float s = 0.0f, c = 0.0f;
  __device__ __forceinline__ void add(float x) {
    float y = x - c;
    float t = s + y;
    c = (t - s) - y;
    s = t;
  }
};

// ----------------------------------------------------------------- kernels
// unit conversion of the masses (reader.f90:457-461), the box check
// (avism.f90:573-575), the largest |mass| (for the DDENS fixed point), and
// whether every mass equals the first (m0raw: the first mass as given, copied
// aside before the conversion in place); grid-stride with one atomic per
// block, since same-address atomics serialise
__global__ void __launch_bounds__(256) k_convert_check(int64_t np, float* m, const float* x, const float* y,
                                                      const float* z, float half, int* bad, unsigned* mmax,
                                                      const float* m0raw, int* mixed) {
  unsigned mb = 0;
  bool out = false, diff = false;
  const float m0 = *m0raw / kUM;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < np; i += (int64_t)gridDim.x * blockDim.x) {
    const float mi = m[i] / kUM;
    m[i] = mi;
    const float a = x[i], b = y[i], c = z[i];
    out = out || (a < -half || a > half || b < -half || b > half || c < -half || c > half);
    mb = max(mb, __float_as_uint(fabsf(mi)));  // non-negative floats order as their bits
    diff = diff || __float_as_uint(mi) != __float_as_uint(m0);
  }
  for (int o = 32; o > 0; o >>= 1) mb = max(mb, (unsigned)__shfl_xor((int)mb, o));
  __shared__ unsigned wmax[4];
  __shared__ int wbad[4], wdiff[4];
  const bool anyout = __ballot(out) != 0, anydiff = __ballot(diff) != 0;
  if ((threadIdx.x & 63) == 0) {
    wmax[threadIdx.x >> 6] = mb;
    wbad[threadIdx.x >> 6] = anyout ? 1 : 0;
    wdiff[threadIdx.x >> 6] = anydiff ? 1 : 0;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const unsigned bm = max(max(wmax[0], wmax[1]), max(wmax[2], wmax[3]));
    if (bm) atomicMax(mmax, bm);
    if (wbad[0] | wbad[1] | wbad[2] | wbad[3]) atomicOr(bad, 1);
    if (wdiff[0] | wdiff[1] | wdiff[2] | wdiff[3]) atomicOr(mixed, 1);
  }
}

// velocities to code units (reader.f90:457-459), on the copy stream
__global__ void __launch_bounds__(256) k_convert_vel(int64_t np, float* vx, float* vy, float* vz) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < np; i += (int64_t)gridDim.x * blockDim.x) {
    vx[i] = vx[i] / kUV;
    vy[i] = vy[i] / kUV;
    vz[i] = vz[i] / kUV;
  }
}

__global__ void k_fill_u64(int64_t n, unsigned long long* a, unsigned long long v) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = v;
}

// Query order of PPART: the particles grouped by blocks of 4^3 bins, in bin
// order inside a block, so that a wave's 64 queries (about one block at one
// particle per bin) are neighbours in 3-D rather than a run of 64 bins along
// x; their searches then take similar paths (10% fewer PPART instructions).
// A block's particles are its 16 rows of 4 bins, each a contiguous range.
constexpr int kPB = 4;  // bins per block side
__device__ __forceinline__ void pblock_rows(const Bins& B, int64_t b, int nbb, int& x0, int& x1, int& y0, int& z0,
                                            int& ny, int& nz) {
  const int X = (int)(b % nbb), Y = (int)((b / nbb) % nbb), Z = (int)(b / ((int64_t)nbb * nbb));
  x0 = X * kPB;
  x1 = min(x0 + kPB, B.nb);
  y0 = Y * kPB;
  z0 = Z * kPB;
  ny = min(kPB, B.nb - y0);
  nz = min(kPB, B.nb - z0);
}
__global__ void k_pblock_count(Bins B, int nbb, int32_t* cnt) {
  const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= (int64_t)nbb * nbb * nbb) return;
  int x0, x1, y0, z0, ny, nz;
  pblock_rows(B, b, nbb, x0, x1, y0, z0, ny, nz);
  int c = 0;
#pragma unroll
  for (int r = 0; r < kPB * kPB; ++r) {  // the 16 rows' loads all in flight
    const int dz = r / kPB, dy = r % kPB;
    if (dy < ny && dz < nz) {
      const int64_t row = (int64_t)B.nb * ((int64_t)(y0 + dy) + (int64_t)B.nb * (z0 + dz));
      c += B.start[row + x1] - B.start[row + x0];
    }
  }
  cnt[b] = c;
}
// one wave per block: its rows' particles, in order, from off[b]
__global__ void __launch_bounds__(256) k_pblock_write(Bins B, int nbb, const int32_t* __restrict__ off,
                                                     uint32_t* perm) {
  const int64_t b = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (b >= (int64_t)nbb * nbb * nbb) return;
  const int lane = threadIdx.x & 63;
  int x0, x1, y0, z0, ny, nz;
  pblock_rows(B, b, nbb, x0, x1, y0, z0, ny, nz);
  // lanes 0..15 load row dz * 4 + dy's range, then every lane writes
  int s0 = 0, s1 = 0;
  if (lane < kPB * kPB) {
    const int dz = lane / kPB, dy = lane % kPB;
    if (dy < ny && dz < nz) {
      const int64_t row = (int64_t)B.nb * ((int64_t)(y0 + dy) + (int64_t)B.nb * (z0 + dz));
      s0 = B.start[row + x0];
      s1 = B.start[row + x1];
    }
  }
  int o = off[b];
  for (int r = 0; r < kPB * kPB; ++r) {
    const int a = __shfl(s0, r), e = __shfl(s1, r);
    for (int q = lane; q < e - a; q += 64) perm[o + q] = (uint32_t)(a + q);
    o += e - a;
  }
}

__global__ void k_bin_keys(int64_t np, const float* x, const float* y, const float* z, Bins B,
                           uint32_t* key, uint32_t* val) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  int b = B.bin_of(x[i]) + B.nb * (B.bin_of(y[i]) + B.nb * B.bin_of(z[i]));
  key[i] = (uint32_t)b;
  val[i] = (uint32_t)i;
}

// bin starts from the sorted keys: position i starts every bin after key[i-1]
// up to key[i] (bin nbins at np), so each of the nbins + 1 entries is written
// once.  A thread writes at most kRun of its run; the rest of a longer run
// (empty regions) goes to a list that k_bin_runs fills with whole blocks.
constexpr int kRun = 256;
__global__ void k_bin_starts(int64_t np, int64_t nbins, const uint32_t* __restrict__ key, int32_t* start,
                             int4* longrun, int32_t* nlong) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i > np) return;
  const int64_t lo = i == 0 ? 0 : (int64_t)key[i - 1] + 1, hi = i == np ? nbins : (int64_t)key[i];
  const int64_t end = min(hi, lo + kRun - 1);
  for (int64_t b = lo; b <= end; ++b) start[b] = (int32_t)i;
  if (end < hi) longrun[atomicAdd(nlong, 1)] = make_int4((int)(end + 1), (int)hi, (int)i, 0);
}
__global__ void k_bin_runs(const int4* __restrict__ longrun, const int32_t* __restrict__ nlong, int32_t* start) {
  const int n = *nlong;
  for (int e = blockIdx.x; e < n; e += gridDim.x) {
    const int4 r = longrun[e];
    for (int b = r.x + (int)threadIdx.x; b <= r.y; b += blockDim.x) start[b] = r.z;
  }
}

__global__ void k_gather_sorted(int64_t np, const uint32_t* perm, const float* x, const float* y,
                                const float* z, const float* m, float4* pos, uint32_t* orig) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  uint32_t p = perm[i];
  pos[i] = make_float4(x[p], y[p], z[p], m[p]);
  orig[i] = p;
}

// PPART_DENS (particles.f90:742-802): rho_p = (FP64 sum of the k nearest
// masses, self included) / (4/3 pi d_k^3); d_k is kept for FILL_H_ZEROS.
// One particle per lane, the queries taken in blocks of 4^3 bins (a wave's
// queries are neighbours); the masses come with the bin-sorted records.

// threads per block of the kNN kernels: a lane keeps its k selection slots
// in LDS (5K/4 words), so K = 128 and 256 run 64 lanes a block (40 and 80 KB)
constexpr int kq_of(int K) { return K <= 64 ? 128 : 64; }
// the selection keeps its slots' candidates in registers (k <= 64), so that
// the k nearest are known without a second sweep when no tie at the k-th
// distance was turned away
constexpr bool knn_idx(int K) { return K <= 64; }

// Workgroups are dispatched to the 8 XCDs round-robin (block b on XCD b % 8),
// and each XCD has its own L2.  Renumbering the blocks so that every XCD gets
// one contiguous range of tiles keeps neighbouring tiles, which read the same
// bins, on the same L2.
__device__ __forceinline__ int64_t xcd_block(int64_t b, int64_t nblocks) {
  constexpr int NX = 8;
  const int64_t q = nblocks / NX, r = nblocks % NX, x = b % NX, y = b / NX;
  return x * q + (x < r ? x : r) + y;
}

// uniform: every particle has the same mass, so the FP64 sum of the k
// nearest masses is k * m exactly (each partial sum j * m, j <= 64, needs at
// most 30 significant bits) and the neighbours themselves are not needed.
// Otherwise (IDX) the masses of the selection's slots are summed when they
// are the k nearest, else those of a second sweep.
template <int PER, int K, bool IDX>
__device__ __forceinline__ void ppart_one(const Bins& B, int64_t np, int k, float r0, int uniform, int64_t i,
                                          uint32_t* keys, uint32_t* gmx, float* pdens, float* hk) {
  const float4 pq = B.pos[i];
  Sel<kq_of(K), K / 4, IDX> h;
  h.init(keys, gmx, k);
  bool scanned;
  const float dk2 = knn_kth<PER>(B, np, pq.x, pq.y, pq.z, r0, k, h, scanned);
  double bas8 = 0.0;  // particles.f90:779-788
  if (uniform) {
    bas8 = (double)k * (double)pq.w;
  } else if (IDX && !scanned && h.rej == 0) {
    float m[K];
#pragma unroll
    for (int t = 0; t < K; ++t) m[t] = t < k ? B.pos[h.idr[t]].w : 0.0f;
#pragma unroll
    for (int t = 0; t < K; ++t)
      if (t < k) bas8 += (double)m[t];
  } else {
    knn_each<PER>(B, np, pq.x, pq.y, pq.z, k, dk2, scanned, h, PosLoad{B.pos},
                  [&](int, float, const float4& p) { bas8 += (double)p.w; });
  }
  const float dk = sqrtf(dk2);  // kdtree.f90:539
  hk[i] = dk;
  const double vol = (4.0 / 3.0 * kPi) * (double)((dk * dk) * dk);
  pdens[i] = (float)(bas8 / vol);  // particles.f90:790-794
}

template <int K>
__global__ void __launch_bounds__(kq_of(K)) k_ppart(Bins B, int64_t np, int k, float r0, int uniform,
                                                   float* pdens, float* hk, const uint32_t* __restrict__ perm) {
  constexpr int BQ = kq_of(K);
  __shared__ uint32_t keys_s[K * BQ], gmx_s[K / 4 * BQ];
  const int64_t t = xcd_block(blockIdx.x, gridDim.x) * BQ + threadIdx.x;
  if (t >= np) return;
  const int64_t i = perm ? (int64_t)perm[t] : t;
  if (B.periodic) {
    if (uniform)
      ppart_one<1, K, false>(B, np, k, r0, 1, i, keys_s + threadIdx.x, gmx_s + threadIdx.x, pdens, hk);
    else
      ppart_one<1, K, knn_idx(K)>(B, np, k, r0, 0, i, keys_s + threadIdx.x, gmx_s + threadIdx.x, pdens, hk);
  } else {
    if (uniform)
      ppart_one<0, K, false>(B, np, k, r0, 1, i, keys_s + threadIdx.x, gmx_s + threadIdx.x, pdens, hk);
    else
      ppart_one<0, K, knn_idx(K)>(B, np, k, r0, 0, i, keys_s + threadIdx.x, gmx_s + threadIdx.x, pdens, hk);
  }
}

// VVEL_INTERP_SPH_VW (particles.f90:806-930): per cell the k nearest (or,
// where d_k <= DX, every particle within DX), HPART = max distance at which a
// particle served a cell (atomic max on the float bits), and the FP64
// weighted velocity average.  One cell per lane; a block covers 8 x 4 x 4 cells.
struct VvelOut {
  const float4* __restrict__ vel;  // (vx, vy, vz, m/rho), bin-sorted
  uint32_t* hbits;                 // HPART bits, bin-sorted
  float *u, *v, *w;
  int want_vel;
};

// CHK: also where the neighbours come from a sweep (ties at d_k, ball mode),
// read HPART before the atomic and skip it when the value is not larger.  For
// grids finer than the particles, where each particle serves many cells; a
// stale read is never above the true maximum, so no update is lost.  (When
// the selection's slots are the neighbours, their reads are always made, all
// together, before the atomics.)
inline bool vvel_chk(int64_t ncell, int64_t np) { return ncell >= 4 * np; }

template <int PER, int K, bool CHK = false>
__device__ __forceinline__ void vvel_one(const Bins& B, int64_t np, int k, float r0, float dx, float qx, float qy,
                                         float qz, size_t c, uint32_t* keys, uint32_t* gmx, const VvelOut& O) {
  constexpr int BS = kq_of(K);
  constexpr bool IDX = knn_idx(K);
  Sel<BS, K / 4, IDX> h;
  h.init(keys, gmx, k);
  bool scanned;
  const float dk2 = knn_kth<PER>(B, np, qx, qy, qz, r0, k, h, scanned);
  const float dk = sqrtf(dk2);
  double sw = 0.0, sx = 0.0, sy = 0.0, sz = 0.0;
  // particles.f90:882-885; old: HPART as read before (0: not read).  The
  // atomic returns nothing, so the lane does not wait for it.
  auto serve_v = [&](int j, float d, float hh, uint32_t old) {
    const float4 vv = O.want_vel ? O.vel[j] : make_float4(0.f, 0.f, 0.f, 0.f);  // vx, vy, vz, m/rho
    const uint32_t hb = __float_as_uint(d);
    if (hb > old) atomicMax(&O.hbits[j], hb);
    if (O.want_vel) {
      const float qq = d / (hh / 2.0f);                     // particles.f90:941
      const float wgt = m4(qq) * vv.w;                      // 897-900
      const double w8 = (double)wgt;
      sw += w8;
      sx += w8 * (double)vv.x;
      sy += w8 * (double)vv.y;
      sz += w8 * (double)vv.z;
    }
  };
  // one neighbour found by a sweep (reading HPART first there costs a load
  // latency per neighbour, so only CHK does)
  auto serve = [&](int j, float d, float hh) { serve_v(j, d, hh, CHK ? O.hbits[j] : 0u); };
  if (dk > dx) {  // k neighbours (particles.f90:866-867)
    if (IDX && !scanned && h.rej == 0) {
      // the k neighbours are known: their HPART values are read first, all
      // loads in flight together, and an atomic is skipped where HPART
      // already holds at least this distance (a stale value is never above
      // the true maximum)
      uint32_t old[K];
#pragma unroll
      for (int t = 0; t < K; ++t) old[t] = t < k ? O.hbits[h.idr[t]] : 0u;
#pragma unroll
      for (int t = 0; t < K; ++t)
        if (t < k) serve_v((int)h.idr[t], sqrtf(u2f(keys[t * BS])), dk, old[t]);
    } else {
      knn_each<PER>(B, np, qx, qy, qz, k, dk2, scanned, h, PosLoad{B.pos},
                    [&](int j, float d2, const float4&) { serve(j, sqrtf(d2), dk); });
    }
  } else {  // ball of radius DX (particles.f90:868-878): every particle with d <= DX
    const float r2 = dx * dx;
    float maxd2 = 0.0f;
    ball_each<PER>(B, np, qx, qy, qz, r2, PosLoad{B.pos},
                   [&](int, float d2, const float4&) { maxd2 = fmaxf(maxd2, d2); });
    const float hh = sqrtf(maxd2);
    ball_each<PER>(B, np, qx, qy, qz, r2, PosLoad{B.pos},
                   [&](int j, float d2, const float4&) { serve(j, sqrtf(d2), hh); });
  }
  if (O.want_vel) {
    O.u[c] = (float)(sx / sw);
    O.v[c] = (float)(sy / sw);
    O.w[c] = (float)(sz / sw);
  }
}

template <int K, bool CHK>
__global__ void __launch_bounds__(kq_of(K)) k_vvel(Bins B, int64_t np, int n, const float* __restrict__ rad,
                                                  int rlo, float dx, int k, float r0,
                                                  const int8_t* __restrict__ smask, VvelOut O) {
  constexpr int BQ = kq_of(K), TZ = BQ / 32;  // tile of 8 x 4 x TZ cells
  __shared__ uint32_t keys_s[K * BQ], gmx_s[K / 4 * BQ];
  const int ntx = (n + 7) / 8, nty = (n + 3) / 4;
  const int tid = threadIdx.x;
  const int64_t bt = xcd_block(blockIdx.x, gridDim.x);
  const int ix = (int)(bt % ntx) * 8 + tid % 8, iy = (int)((bt / ntx) % nty) * 4 + (tid / 8) % 4,
            iz = (int)(bt / ((int64_t)ntx * nty)) * TZ + tid / 32;
  if (ix >= n || iy >= n || iz >= n) return;
  const size_t c = (size_t)ix + (size_t)n * ((size_t)iy + (size_t)n * iz);
  if (smask && smask[c] == 0) return;
  const float qx = rad[ix + 1 - rlo], qy = rad[iy + 1 - rlo], qz = rad[iz + 1 - rlo];
  if (B.periodic)
    vvel_one<1, K, CHK>(B, np, k, r0, dx, qx, qy, qz, c, keys_s + tid, gmx_s + tid, O);
  else
    vvel_one<0, K, CHK>(B, np, k, r0, dx, qx, qy, qz, c, keys_s + tid, gmx_s + tid, O);
}

// the velocity record of every particle, bin-sorted: (vx, vy, vz, m/rho), m/rho
// being the SPH weight factor of particles.f90:897-900 (one FP32 division as
// there)
__global__ void k_velrec(int64_t np, const uint32_t* __restrict__ orig, const float4* __restrict__ pos,
                         const float* __restrict__ pdens, const float* __restrict__ vx,
                         const float* __restrict__ vy, const float* __restrict__ vz, float4* vel) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  const uint32_t p = orig[i];
  vel[i] = make_float4(vx[p], vy[p], vz[p], pos[i].w / pdens[i]);
}

// FILL_H_ZEROS (particles.f90:398-431): h = distance to the k-th neighbour
// where no cell was served (bin-sorted)
__global__ void k_fill_h(int64_t np, float* hpart, const float* hk) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  if (!(hpart[i] > 0.0f)) hpart[i] = hk[i];
}

// ---------------------------------------------------------------- TSC
// DENS_INTERP_TSC (particles.f90:16-128) and VEL_INTERP_TSC (132-272): each
// particle spreads over the 3 x 3 x 3 cells around its nearest centre with the
// triangular-shaped-cloud weights, in the reference's FP32 operations.  The
// sums are int64 fixed point (order-free, so reproducible); the reference adds
// REAL*8 (mass) and REAL*4 under OMP ATOMIC (velocities), whose results depend
// on its thread schedule.  Non-periodic: cells outside the grid are dropped
// (the reference writes outside its arrays there).  The sums are FP64 atomic
// adds: weights reach 1e-20 and below in the tails, which a fixed point would
// drop; FP64 addition order changes only the last bits of the FP64 sums,
// below the FP32 rounding of the results.
struct TscGeom {
  const float* __restrict__ rad;  // RADX, reference index i at rad[i - rlo]
  int rlo, n, periodic;
  float dx;
};

// the periodic wrap of a coordinate (particles.f90:165-181, 187-194)
__device__ __forceinline__ float tsc_wrap(const TscGeom& T, float p) {
  if (T.periodic) {
    const float lo = T.rad[1 - T.rlo] - 0.5f * T.dx, hi = T.rad[T.n - T.rlo] + 0.5f * T.dx;
    const float leng = hi - lo;
    if (p < lo) p = p + leng;
    if (p >= hi) p = p - leng;
  }
  return p;
}

// the nearest centre I and the weights of I-1, I, I+1 (particles.f90:196-217)
__device__ __forceinline__ int tsc_w(const TscGeom& T, float p, float V[3]) {
  const int I = (int)(((p - T.rad[1 - T.rlo]) / T.dx) + 0.49999f) + 1;
  float bas = fabsf(T.rad[I - 1 - T.rlo] - p);
  float t = 1.5f - bas / T.dx;
  V[0] = 0.5f * (t * t);
  bas = fabsf(T.rad[I - T.rlo] - p);
  t = bas / T.dx;
  V[1] = 0.75f - t * t;
  bas = fabsf(T.rad[I + 1 - T.rlo] - p);
  t = 1.5f - bas / T.dx;
  V[2] = 0.5f * (t * t);
  return I;
}

// cell index after the periodic wrap (particles.f90:223-230); -1 outside
__device__ __forceinline__ int tsc_cell(const TscGeom& T, int i) {
  if (T.periodic) {
    if (i < 1) i += T.n;
    if (i > T.n) i -= T.n;
  }
  return (i < 1 || i > T.n) ? -1 : i;
}

// one thread per particle (input order): the mass deposits (mass != null)
// into mass[N^3], the velocity sums (acc != null) into acc[4][N^3]
__global__ void __launch_bounds__(256) k_tsc(TscGeom T, int64_t np, const float* __restrict__ x,
                                            const float* __restrict__ y, const float* __restrict__ z,
                                            const float* __restrict__ m, const float* __restrict__ vx,
                                            const float* __restrict__ vy, const float* __restrict__ vz,
                                            double* mass, double* acc) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np) return;
  float VX[3], VY[3], VZ[3];
  const int I = tsc_w(T, tsc_wrap(T, x[p]), VX);
  const int J = tsc_w(T, tsc_wrap(T, y[p]), VY);
  const int K = tsc_w(T, tsc_wrap(T, z[p]), VZ);
  const size_t ncell = (size_t)T.n * T.n * T.n;
  const float mp = mass ? m[p] : 0.0f;
  const float up = acc ? vx[p] : 0.0f, vp = acc ? vy[p] : 0.0f, wp = acc ? vz[p] : 0.0f;
  for (int a = -1; a <= 1; ++a) {
    const int k3 = tsc_cell(T, K + a);
    if (k3 < 0) continue;
    for (int b = -1; b <= 1; ++b) {
      const int j3 = tsc_cell(T, J + b);
      if (j3 < 0) continue;
      for (int cc = -1; cc <= 1; ++cc) {
        const int i3 = tsc_cell(T, I + cc);
        if (i3 < 0) continue;
        const size_t o = (size_t)(i3 - 1) + (size_t)T.n * ((size_t)(j3 - 1) + (size_t)T.n * (k3 - 1));
        if (mass) {
          const float inc = ((mp * VX[cc + 1]) * VY[b + 1]) * VZ[a + 1];  // particles.f90:100-118
          atomicAdd(&mass[o], (double)inc);
        }
        if (acc) {
          const float wt = (VX[cc + 1] * VY[b + 1]) * VZ[a + 1];  // particles.f90:232
          atomicAdd(&acc[o], (double)(up * wt));
          atomicAdd(&acc[ncell + o], (double)(vp * wt));
          atomicAdd(&acc[2 * ncell + o], (double)(wp * wt));
          atomicAdd(&acc[3 * ncell + o], (double)wt);
        }
      }
    }
  }
}

// U = sum(v * w) / WEI with WEI = sum(w), 1 where no particle reached the cell
// (particles.f90:255-267)
__global__ void k_tsc_vel_finalize(int64_t ncell, const double* __restrict__ acc, float* u, float* v, float* w) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncell) return;
  float wei = (float)acc[3 * ncell + c];
  if (wei == 0.0f) wei = 1.0f;
  u[c] = (float)acc[c] / wei;
  v[c] = (float)acc[ncell + c] / wei;
  w[c] = (float)acc[2 * ncell + c] / wei;
}

// FILL_H_ZEROS (particles.f90:398-431) when no PPART query ran (TSC
// velocities): h = the k-th neighbour distance of each particle with h = 0
template <int PER, int K>
__global__ void __launch_bounds__(kq_of(K)) k_fill_h_knn(Bins B, int64_t np, int k, float r0, float* hpart) {
  constexpr int BQ = kq_of(K);
  __shared__ uint32_t keys_s[K * BQ], gmx_s[K / 4 * BQ];
  const int64_t i = xcd_block(blockIdx.x, gridDim.x) * BQ + threadIdx.x;
  if (i >= np || hpart[i] > 0.0f) return;
  const float4 pq = B.pos[i];
  Sel<BQ, K / 4> h;
  h.init(keys_s + threadIdx.x, gmx_s + threadIdx.x, k);
  bool scanned;
  hpart[i] = sqrtf(knn_kth<PER>(B, np, pq.x, pq.y, pq.z, r0, k, h, scanned));
}

// DDENS_INTERP_SPH pass 1 (particles.f90:479-547): per particle the NORM, in
// the reference's loop order and FP32, plus the reach of its support in cells.
struct DdensGeom {
  const float* __restrict__ rad;  // RADX, reference index i at rad[i - rlo]
  int rlo, n;
  float dx, rx1;
};

__device__ __forceinline__ void ddens_setup(const DdensGeom& G, float x, float y, float z,
                                            float h, int& ix, int& iy, int& iz, float& inv_h,
                                            float& h2max, int& inc, int& reach) {
  ix = (int)((x - G.rx1) / G.dx) + 1;  // particles.f90:489-491
  iy = (int)((y - G.rx1) / G.dx) + 1;
  iz = (int)((z - G.rx1) / G.dx) + 1;
  inv_h = 2.0f / h;
  h2max = (4.0f * h) * h;
  inc = (int)((2.0f * h) / G.dx + 0.5f);
  // beyond (int)(h/dx)+2 cells from the home cell the distance exceeds h + dx/2,
  // where the M4 weight (q = 2d/h > 2) is exactly zero
  int rr = (int)(h / G.dx) + 2;
  reach = rr < inc ? rr : inc;
}

// ddens_w with the row's squared y and z separations given (dy2 = ty * ty,
// dz2 = tz * tz, as ddens_w forms them): the same FP32 operations
__device__ __forceinline__ float ddens_w_row(const DdensGeom& G, float x, float dy2, float dz2, float inv_h,
                                            float h2max, int i2, bool& ok) {
  float tx = x - G.rad[i2 - G.rlo];
  float dx2 = tx * tx;
  float d2 = (dx2 + dy2) + dz2;
  ok = !(d2 > h2max);
  float q = sqrtf(d2) * inv_h;
  return m4(q);
}

__device__ __forceinline__ float ddens_w(const DdensGeom& G, float x, float y, float z, float inv_h,
                                        float h2max, int i2, int j2, int k2, bool& ok) {
  float tz = z - G.rad[k2 - G.rlo];
  float dz2 = tz * tz;
  float ty = y - G.rad[j2 - G.rlo];
  float dy2 = ty * ty;
  float tx = x - G.rad[i2 - G.rlo];
  float dx2 = tx * tx;
  float d2 = (dx2 + dy2) + dz2;
  ok = !(d2 > h2max);
  float q = sqrtf(d2) * inv_h;
  return m4(q);
}

// DDENS_INTERP_SPH (particles.f90:434-618), one thread per particle: the
// particle's NORM over its cell box (first loop), then its deposits m*W/NORM
// (second loop) added into int64 fixed-point cell sums.  Integer atomics make
// the sums independent of the deposit order (bit-reproducible), and the fixed
// point (scale = 2^62 / (np * max mass), so no sum can overflow) resolves
// 1e-11 of a particle mass at 10^8 particles; the reference adds FP64.
//
// A block takes the particles binned in one tile of DTX x DTY x DTZ cells and
// adds their deposits into an LDS copy of the tile widened by DR cells (LDS
// integer atomics), then adds that copy to HBM once per cell: at one particle
// per cell a cell receives ~32 deposits but only a handful of tile flushes.
// A particle whose cell box reaches past the widened tile (large h, in voids)
// adds directly to HBM.  ddens_tile is one tile (bt, in the grid's tile
// order) for a whole block of 256 threads; DTY and DTZ must be multiples of
// the bin size (1, 2 or 4 cells).
template <int DTX, int DTY, int DTZ, int DR, int FM = 1>
struct DdensTile {
  static constexpr int DBX = DTX + 2 * DR, DBY = DTY + 2 * DR, DBZ = DTZ + 2 * DR;
  static constexpr int LDS = DBX * DBY * DBZ;
  // rows of bins of a tile, with up to FM bins per cell along an axis
  static constexpr int ROWS = (DTY * FM) * (DTZ * FM);
  static int64_t tiles(int n) {
    return (int64_t)((n + DTX - 1) / DTX) * ((n + DTY - 1) / DTY) * ((n + DTZ - 1) / DTZ);
  }
};

template <int DTX, int DTY, int DTZ, int DR, int FM>
__device__ __forceinline__ void ddens_tile(const DdensGeom& G, const Bins& B, const float* __restrict__ hpart,
                                           double scale, unsigned long long* acc, float* norm,
                                           unsigned long long* npairs, int64_t bt, unsigned long long* box,
                                           int* qrow, int* qbeg) {
  using T = DdensTile<DTX, DTY, DTZ, DR, FM>;
  constexpr int DBX = T::DBX, DBY = T::DBY, DBZ = T::DBZ;
  const int n = G.n;
  const int ntx = (n + DTX - 1) / DTX, nty = (n + DTY - 1) / DTY;
  const int tx0 = (int)(bt % ntx) * DTX, ty0 = (int)((bt / ntx) % nty) * DTY, tz0 = (int)(bt / ((int64_t)ntx * nty)) * DTZ;
  const int tid = threadIdx.x;
  for (int c = tid; c < DBX * DBY * DBZ; c += 256) box[c] = 0ull;
  // the tile's particles: one contiguous range per row of bins.  Cell i
  // (0-based) holds bins [i nb / n, (i + 1) nb / n): a bin is s^3 cells (nb =
  // n / s, s dividing the tile) or a cell f^3 bins (nb = n f, f <= FM).
  const int nb = B.nb;
  const int ib0 = (int)((int64_t)tx0 * nb / n), ib1 = (int)min((int64_t)(tx0 + DTX) * nb / n, (int64_t)nb);
  const int jb0 = (int)((int64_t)ty0 * nb / n), jb1 = (int)min((int64_t)(ty0 + DTY) * nb / n, (int64_t)nb);
  const int kb0 = (int)((int64_t)tz0 * nb / n), kb1 = (int)min((int64_t)(tz0 + DTZ) * nb / n, (int64_t)nb);
  const int nry = jb1 - jb0, nr = nry * (kb1 - kb0);  // <= T::ROWS
  for (int r = tid; r < nr; r += 256) {
    const int rowb = nb * ((jb0 + r % nry) + nb * (kb0 + r / nry));
    const int b0 = B.start[rowb + ib0];
    qbeg[r] = b0;
    qrow[r + 1] = B.start[rowb + ib1] - b0;
  }
  __syncthreads();
  if (tid == 0) {  // counts -> offsets
    qrow[0] = 0;
    for (int r = 0; r < nr; ++r) qrow[r + 1] += qrow[r];
  }
  __syncthreads();
  // widened tile in 1-based cells
  const int bx0 = tx0 + 1 - DR, by0 = ty0 + 1 - DR, bz0 = tz0 + 1 - DR;
  const int nq = qrow[nr];
  unsigned pairs = 0;  // (particle, cell) pairs within h: the work of the gather model
  for (int q = tid; q < nq; q += 256) {
    int r = 0, rh = nr;  // the row holding q: the last r with qrow[r] <= q
    while (rh - r > 1) {
      const int mid = (r + rh) >> 1;
      if (qrow[mid] <= q) r = mid;
      else rh = mid;
    }
    const int64_t i = qbeg[r] + (q - qrow[r]);
    const float4 p = B.pos[i];
    const float hh = hpart[i];
    int ix, iy, iz, inc, reach;
    float inv_h, h2max;
    ddens_setup(G, p.x, p.y, p.z, hh, ix, iy, iz, inv_h, h2max, inc, reach);
    // the reference's cell box, clipped to the grid ...
    int k0 = max(iz - reach, 1), k1 = min(iz + reach, n);
    int j0 = max(iy - reach, 1), j1 = min(iy + reach, n);
    const int i0 = max(ix - reach, 1), i1 = min(ix + reach, n);
    // ... and to the cells whose centre can lie within h (beyond it M4(2d/h)
    // is exactly 0, and adding +0 leaves the FP32 NORM unchanged).  In cell
    // units u = (x - RADX(1)) / DX + 1 the centre of cell i is at u = i, so the
    // centres within r of x are the integers in [u - r/DX, u + r/DX]; the
    // bounds keep a margin of 1e-5 h^2 and 1e-3 cells, far above the rounding
    // involved
    const float hm2 = hh * hh * 1.00001f;
    const float rdx = 1.0f / G.dx;
    const float ux = (p.x - G.rx1) * rdx + 1.0f, uy = (p.y - G.rx1) * rdx + 1.0f, uz = (p.z - G.rx1) * rdx + 1.0f;
    const float hc = hh * 1.00001f * rdx + 1e-3f;
    k0 = max(k0, (int)ceilf(uz - hc));
    k1 = min(k1, (int)floorf(uz + hc));
    j0 = max(j0, (int)ceilf(uy - hc));
    j1 = min(j1, (int)floorf(uy + hc));
    const int ia = max(i0, (int)ceilf(ux - hc));  // bounds every row's run below
    const int ib = min(i1, (int)floorf(ux + hc));
    const bool local = ia >= bx0 && ib < bx0 + DBX && j0 >= by0 && j1 < by0 + DBY && k0 >= bz0 && k1 < bz0 + DBZ;
    auto deposit = [&](int ci, int cj, int ck, float v) {
      const unsigned long long f = (unsigned long long)llrint((double)v * scale);
      if (local)
        atomicAdd(&box[(ci - bx0) + DBX * ((cj - by0) + DBY * (ck - bz0))], f);
      else
        atomicAdd(&acc[(size_t)(ci - 1) + (size_t)n * ((size_t)(cj - 1) + (size_t)n * (ck - 1))], f);
    };
    float nrm = 0.0f;
    for (int pass = 0; pass < 2; ++pass) {
      if (pass == 1) {
        norm[i] = nrm;
        if (nrm == 0.0f) {  // particles.f90:549-557: all the mass to the home cell
          if (ix >= 1 && ix <= n && iy >= 1 && iy <= n && iz >= 1 && iz <= n) {
            const bool home_local = ix >= bx0 && ix < bx0 + DBX && iy >= by0 && iy < by0 + DBY && iz >= bz0 &&
                                    iz < bz0 + DBZ;
            const unsigned long long f = (unsigned long long)llrint((double)p.w * scale);
            if (home_local)
              atomicAdd(&box[(ix - bx0) + DBX * ((iy - by0) + DBY * (iz - bz0))], f);
            else
              atomicAdd(&acc[(size_t)(ix - 1) + (size_t)n * ((size_t)(iy - 1) + (size_t)n * (iz - 1))], f);
          }
          break;
        }
      }
      for (int k2 = k0; k2 <= k1; ++k2) {
        const float tz = p.z - G.rad[k2 - G.rlo];
        for (int j2 = j0; j2 <= j1; ++j2) {
          const float ty = p.y - G.rad[j2 - G.rlo];
          const float rem = hm2 - (ty * ty + tz * tz);
          if (rem < 0.0f) continue;  // no centre of this row within h
          const float rc = bound_sqrt(rem) * rdx + 1e-3f;
          const int a = max(i0, (int)ceilf(ux - rc));
          const int b = min(i1, (int)floorf(ux + rc));
          const float dy2 = ty * ty, dz2 = tz * tz;
          for (int i2 = a; i2 <= b; ++i2) {
            bool ok;
            const float wgt = ddens_w_row(G, p.x, dy2, dz2, inv_h, h2max, i2, ok);
            if (!ok) continue;
            if (pass == 0) {
              ++pairs;
              nrm = nrm + wgt;  // particles.f90:520-547, in the reference's loop order
            } else {
              const float v = (wgt * p.w) / nrm;  // particles.f90:599-600
              if (v != 0.0f) deposit(i2, j2, k2, v);
            }
          }
        }
      }
    }
  }
  for (int o = 32; o > 0; o >>= 1) pairs += (unsigned)__shfl_xor((int)pairs, o);
  if ((tid & 63) == 0 && pairs) atomicAdd(npairs, (unsigned long long)pairs);
  __syncthreads();
  for (int c = tid; c < DBX * DBY * DBZ; c += 256) {
    const unsigned long long v = box[c];
    if (v == 0ull) continue;
    const int ci = bx0 + c % DBX, cj = by0 + (c / DBX) % DBY, ck = bz0 + c / (DBX * DBY);
    atomicAdd(&acc[(size_t)(ci - 1) + (size_t)n * ((size_t)(cj - 1) + (size_t)n * (ck - 1))], v);
  }
}

// the single-grid tile: 8 x 8 x 8 cells widened by 3 (timed against 8 x 8 x 4
// widened by 4 and 16 x 8 x 4 by 4 by the batch autotuner: fastest at one
// particle per cell)
constexpr int DTX = 8, DTY = 8, DTZ = 8, DR = 3;

// FM: the most bins per cell along an axis the launch allows (1, or 4 where
// dense particle sets split cells)
template <int FM>
__global__ void __launch_bounds__(256) k_ddens_scatter(DdensGeom G, Bins B, const float* __restrict__ hpart,
                                                      double scale, unsigned long long* acc, float* norm,
                                                      unsigned long long* npairs) {
  using T = DdensTile<DTX, DTY, DTZ, DR, FM>;
  __shared__ unsigned long long box[T::LDS];
  __shared__ int qrow[T::ROWS + 1], qbeg[T::ROWS];
  ddens_tile<DTX, DTY, DTZ, DR, FM>(G, B, hpart, scale, acc, norm, npairs, xcd_block(blockIdx.x, gridDim.x), box,
                                    qrow, qbeg);
}

// bins per axis for np particles on n^3 cells, about one particle per bin:
// bins of s^3 cells (s = 1, 2, 4, dividing n; sparse sets, and the finer
// grids of void-in-void levels) or cells of f^3 bins (f = 2, 4; dense sets),
// at most 1024 bins per axis
int32_t bins_per_axis(int32_t n, int64_t np) {
  const double ppc = (double)np / ((double)n * n * n);
  if (ppc >= 8.0) {
    int f = ppc >= 64.0 ? 4 : 2;
    while (f > 1 && (int64_t)n * f > 1024) f /= 2;
    return n * f;
  }
  int s = 1;
  while (s < 4 && n % (2 * s) == 0 && (double)(8 * s * s * s) * ppc <= 1.5) s *= 2;
  return n / s;
}

// cell sums -> U1 = mass/(dx^3 rete^3)/rote*rote (avism.f90:1249-1250, 1386)
// and the FP64 block sums of U1 (avism.f90:1395-1405), fixed reduction order.
// acc: int64 fixed-point cell masses (fp64 == 0) or FP64 ones (fp64 != 0).
// A thread takes kFinV consecutive cells (16-byte loads and stores), a block
// kFinB * kFinV cells.
constexpr int kFinB = 256, kFinV = 4;

__device__ __forceinline__ float fin_cell(unsigned long long a, double inv_scale, int fp64, float den, float rote) {
  const float mass = fp64 ? (float)__longlong_as_double((long long)a) : (float)((double)a * inv_scale);
  float u = mass / den;
  u = u / rote;
  return u * rote;
}

__global__ void __launch_bounds__(kFinB) k_ddens_finalize(int64_t ncell, const unsigned long long* __restrict__ acc,
                                                         double inv_scale, int fp64, float den, float rote,
                                                         float* umass, double* partial) {
  const int64_t c0 = ((int64_t)blockIdx.x * kFinB + threadIdx.x) * kFinV;
  double contrib = 0.0;
  if (c0 + kFinV <= ncell) {
    const ulonglong2 a01 = *reinterpret_cast<const ulonglong2*>(acc + c0);
    const ulonglong2 a23 = *reinterpret_cast<const ulonglong2*>(acc + c0 + 2);
    float4 u;
    u.x = fin_cell(a01.x, inv_scale, fp64, den, rote);
    u.y = fin_cell(a01.y, inv_scale, fp64, den, rote);
    u.z = fin_cell(a23.x, inv_scale, fp64, den, rote);
    u.w = fin_cell(a23.y, inv_scale, fp64, den, rote);
    *reinterpret_cast<float4*>(umass + c0) = u;
    contrib = (((double)u.x + (double)u.y) + (double)u.z) + (double)u.w;
  } else {
    for (int64_t c = c0; c < ncell; ++c) {
      const float u = fin_cell(acc[c], inv_scale, fp64, den, rote);
      umass[c] = u;
      contrib += (double)u;
    }
  }
  // block sum in a fixed order: waves by shuffles, then the four wave sums
  for (int off = 32; off > 0; off >>= 1) contrib += __shfl_xor(contrib, off);
  __shared__ double wsum[kFinB / 64];
  if (threadIdx.x % 64 == 0) wsum[threadIdx.x / 64] = contrib;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < kFinB / 64; ++w) t += wsum[w];
    partial[blockIdx.x] = t;
  }
}

// FP64 sum of the block partials in a fixed order (deterministic): each block
// of the first pass sums a contiguous slice, the single block of the second
// pass sums the slices and forms the mean
constexpr int kRedSlices = 512;
static_assert(kRedSlices == 512, "scalars_bytes reserves 512 slots for the slices");
__global__ void __launch_bounds__(256) k_reduce_mean(const double* partial, int64_t nparts, int64_t ncell,
                                                    double* slices, int final_pass, double* sum_out,
                                                    float* mean_out) {
  __shared__ double red[256];
  const int64_t per = final_pass ? nparts : (nparts + kRedSlices - 1) / kRedSlices;
  const int64_t b0 = final_pass ? 0 : (int64_t)blockIdx.x * per;
  const int64_t b1 = b0 + per < nparts ? b0 + per : nparts;
  double s = 0.0;
  for (int64_t i = b0 + threadIdx.x; i < b1; i += 256) s += partial[i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    if (!final_pass) {
      slices[blockIdx.x] = red[0];
    } else {
      *sum_out = red[0];
      *mean_out = (float)(red[0] / (double)ncell);  // avism.f90:1405
    }
  }
}

// DIVER_UNIFORM (divergence.f90:2-50) fused with U1CO = U1 / MEANDENS.  A
// block takes a 64 x 8 (x, y) column of cells and marches kDivZ planes in z:
// each cell's w at z-1, z, z+1 slides through registers (every w value is
// read once per column), the x and y neighbours come from the same and the
// block's adjacent rows (cache hits); boundary cells use the reference's
// periodic or extrapolated differences with direct loads.
constexpr int kDivX = 64, kDivY = 8, kDivZ = 16;

__global__ void __launch_bounds__(kDivX * kDivY) k_div_finalize(int n, float dx, int periodic,
                                                                const float* __restrict__ u,
                                                                const float* __restrict__ v,
                                                                const float* __restrict__ w,
                                                                const float* __restrict__ umass,
                                                                const float* __restrict__ mean, float* u1co,
                                                                float* div) {
  const int i = blockIdx.x * kDivX + (int)threadIdx.x % kDivX;  // 0-based
  const int j = blockIdx.y * kDivY + (int)threadIdx.x / kDivX;
  if (i >= n || j >= n) return;
  const int z0 = blockIdx.z * kDivZ, z1 = min(z0 + kDivZ, n);
  auto A = [&](const float* f, int a, int b, int cc) -> float {
    return f[(size_t)a + (size_t)n * ((size_t)b + (size_t)n * cc)];
  };
  const int N1 = n - 1;
  const float mn = umass ? *mean : 1.0f;
  float wm = z0 >= 1 ? A(w, i, j, z0 - 1) : 0.0f, wc = A(w, i, j, z0);
  for (int k = z0; k < z1; ++k) {
    const float wp = k + 1 < n ? A(w, i, j, k + 1) : 0.0f;
    float fx, fy, fz;
    if (i >= 1 && i <= n - 2) fx = (A(u, i + 1, j, k) - A(u, i - 1, j, k)) / 2.0f;
    else if (periodic) fx = (i == 0) ? (A(u, 1, j, k) - A(u, N1, j, k)) / 2.0f
                                     : (A(u, 0, j, k) - A(u, n - 2, j, k)) / 2.0f;
    else if (i == 0) fx = 2.0f * (A(u, 1, j, k) - A(u, 0, j, k)) - (A(u, 2, j, k) - A(u, 0, j, k)) / 2.0f;
    else fx = 2.0f * (A(u, N1, j, k) - A(u, n - 2, j, k)) - (A(u, N1, j, k) - A(u, n - 3, j, k)) / 2.0f;
    if (j >= 1 && j <= n - 2) fy = (A(v, i, j + 1, k) - A(v, i, j - 1, k)) / 2.0f;
    else if (periodic) fy = (j == 0) ? (A(v, i, 1, k) - A(v, i, N1, k)) / 2.0f
                                     : (A(v, i, 0, k) - A(v, i, n - 2, k)) / 2.0f;
    else if (j == 0) fy = 2.0f * (A(v, i, 1, k) - A(v, i, 0, k)) - (A(v, i, 2, k) - A(v, i, 0, k)) / 2.0f;
    else fy = 2.0f * (A(v, i, N1, k) - A(v, i, n - 2, k)) - (A(v, i, N1, k) - A(v, i, n - 3, k)) / 2.0f;
    if (k >= 1 && k <= n - 2) fz = (wp - wm) / 2.0f;
    else if (periodic) fz = (k == 0) ? (A(w, i, j, 1) - A(w, i, j, N1)) / 2.0f
                                     : (A(w, i, j, 0) - A(w, i, j, n - 2)) / 2.0f;
    else if (k == 0) fz = 2.0f * (A(w, i, j, 1) - A(w, i, j, 0)) - (A(w, i, j, 2) - A(w, i, j, 0)) / 2.0f;
    else fz = 2.0f * (A(w, i, j, N1) - A(w, i, j, n - 2)) - (A(w, i, j, N1) - A(w, i, j, n - 3)) / 2.0f;
    const size_t c = (size_t)i + (size_t)n * ((size_t)j + (size_t)n * k);
    div[c] = (fx / dx + fy / dx) + fz / dx;
    if (umass) u1co[c] = umass[c] / mn;
    wm = wc;
    wc = wp;
  }
}

static inline dim3 div_grid(int n) {
  return dim3((unsigned)((n + kDivX - 1) / kDivX), (unsigned)((n + kDivY - 1) / kDivY),
              (unsigned)((n + kDivZ - 1) / kDivZ));
}

// velocities not available (FLAG_DIV = -1, avism.f90:1259-1263, 1646-1650):
// U1CO = U1 / MEANDENS and DIVERCO = -(OMEGA0^0.6 RETE HTE) (U1CO - 1), the
// constant formed on the host in the same FP32 order
__global__ void k_div_linear(int64_t ncell, float lin, const float* __restrict__ umass,
                             const float* __restrict__ mean, float* u1co, float* div) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncell) return;
  const float u = umass[c] / (*mean);
  u1co[c] = u;
  div[c] = -(lin * (u - 1.0f));
}

__global__ void k_grid_input(int64_t ncell, const float* delta, const float* vx, const float* vy,
                             const float* vz, float* u1co, float* u, float* v, float* w) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ncell) return;
  u1co[i] = delta[i] + 1.0f;  // reader.f90:545
  u[i] = vx[i] / kUV;         // reader.f90:547-549
  v[i] = vy[i] / kUV;
  w[i] = vz[i] / kUV;
}

__global__ void k_scatter_orig(int64_t np, const uint32_t* orig, const float* a_sorted, float* a_out) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  a_out[orig[i]] = a_sorted[i];
}

// ------------------------------------------------------- batches of patches
// Persistent kernels over the work items of all patches of a batch (see
// fields_batch.h): a block takes items from its XCD's queue until the queue
// is empty, each item being the single-grid kernel's unit of work on one
// patch, so every patch's results are those of the single-grid kernels.

// the next item of this block's XCD queue, or -1 (every thread of the block
// calls it; items are handed out once each)
__device__ __forceinline__ int64_t next_item(const ItemQueue& q, int64_t* slot) {
  const int x = blockIdx.x & 7;
  const int64_t lo = q.total * x / 8, hi = q.total * (x + 1) / 8;
  if (threadIdx.x == 0) *slot = lo + (int64_t)atomicAdd(&q.cnt[x], 1u);
  __syncthreads();
  const int64_t it = *slot;
  __syncthreads();  // every thread has read it before thread 0 writes the next
  return it < hi ? it : -1;
}

// the patch of item it: the last p with off[p] <= it
__device__ __forceinline__ int item_patch(const ItemQueue& q, int64_t it) {
  int lo = 0, hi = q.npatch;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (q.off[mid] <= it) lo = mid;
    else hi = mid;
  }
  return lo;
}

constexpr int kBatchQ = 128;  // lanes of the batched kNN kernels (k <= 64)

template <int K>
__global__ void __launch_bounds__(kBatchQ) k_batch_ppart(const PatchDev* __restrict__ P, ItemQueue q, int k) {
  static_assert(kq_of(K) == kBatchQ, "batched kNN kernels take k <= 64");
  __shared__ uint32_t keys_s[K * kBatchQ], gmx_s[K / 4 * kBatchQ];
  __shared__ int64_t item_s;
  for (;;) {
    const int64_t it = next_item(q, &item_s);
    if (it < 0) break;
    const int p = item_patch(q, it);
    const PatchDev& D = P[p];
    const int64_t i = (it - q.off[p]) * kBatchQ + threadIdx.x;
    if (i >= D.np) continue;
    uint32_t *ks = keys_s + threadIdx.x, *gm = gmx_s + threadIdx.x;
    if (D.B.periodic) {
      if (D.uniform)
        ppart_one<1, K, false>(D.B, D.np, k, D.r0, 1, i, ks, gm, D.pdens, D.hk);
      else
        ppart_one<1, K, knn_idx(K)>(D.B, D.np, k, D.r0, 0, i, ks, gm, D.pdens, D.hk);
    } else {
      if (D.uniform)
        ppart_one<0, K, false>(D.B, D.np, k, D.r0, 1, i, ks, gm, D.pdens, D.hk);
      else
        ppart_one<0, K, knn_idx(K)>(D.B, D.np, k, D.r0, 0, i, ks, gm, D.pdens, D.hk);
    }
  }
}

template <int TX, int TY, int TZ>
struct VvelShape {
  static_assert(TX * TY * TZ == kBatchQ, "one cell per lane");
  static int64_t tiles(int n) {
    return (int64_t)((n + TX - 1) / TX) * ((n + TY - 1) / TY) * ((n + TZ - 1) / TZ);
  }
};

template <int K, int TX, int TY, int TZ>
__global__ void __launch_bounds__(kBatchQ) k_batch_vvel(const PatchDev* __restrict__ P, ItemQueue q, int k) {
  __shared__ uint32_t keys_s[K * kBatchQ], gmx_s[K / 4 * kBatchQ];
  __shared__ int64_t item_s;
  const int tid = threadIdx.x;
  for (;;) {
    const int64_t it = next_item(q, &item_s);
    if (it < 0) break;
    const int p = item_patch(q, it);
    const PatchDev& D = P[p];
    const int n = D.n;
    const int64_t bt = it - q.off[p];
    const int ntx = (n + TX - 1) / TX, nty = (n + TY - 1) / TY;
    const int ix = (int)(bt % ntx) * TX + tid % TX, iy = (int)((bt / ntx) % nty) * TY + (tid / TX) % TY,
              iz = (int)(bt / ((int64_t)ntx * nty)) * TZ + tid / (TX * TY);
    if (ix >= n || iy >= n || iz >= n) continue;
    const size_t c = (size_t)ix + (size_t)n * ((size_t)iy + (size_t)n * iz);
    if (D.smask && D.smask[c] == 0) {  // masked cells keep zero velocity (particles.f90:856)
      D.u[c] = 0.0f;
      D.v[c] = 0.0f;
      D.w[c] = 0.0f;
      continue;
    }
    const float qx = D.rad[ix + 1 - D.rlo], qy = D.rad[iy + 1 - D.rlo], qz = D.rad[iz + 1 - D.rlo];
    const VvelOut O{D.vel, D.hbits, D.u, D.v, D.w, 1};
    if (D.B.periodic)
      vvel_one<1, K>(D.B, D.np, k, D.r0, D.dx, qx, qy, qz, c, keys_s + tid, gmx_s + tid, O);
    else
      vvel_one<0, K>(D.B, D.np, k, D.r0, D.dx, qx, qy, qz, c, keys_s + tid, gmx_s + tid, O);
  }
}

template <int TX, int TY, int TZ, int R>
__global__ void __launch_bounds__(256) k_batch_ddens(const PatchDev* __restrict__ P, ItemQueue q,
                                                    unsigned long long* npairs) {
  using T = DdensTile<TX, TY, TZ, R, 4>;
  __shared__ unsigned long long box[T::LDS];
  __shared__ int qrow[T::ROWS + 1], qbeg[T::ROWS];
  __shared__ int64_t item_s;
  for (;;) {
    const int64_t it = next_item(q, &item_s);
    if (it < 0) break;
    const int p = item_patch(q, it);
    const PatchDev& D = P[p];
    const DdensGeom G{D.rad, D.rlo, D.n, D.dx, D.rx1};
    ddens_tile<TX, TY, TZ, R, 4>(G, D.B, reinterpret_cast<const float*>(D.hbits), D.scale, D.acc, D.norm, npairs,
                                 it - q.off[p], box, qrow, qbeg);
    __syncthreads();  // the tile's LDS is flushed before the next tile clears it
  }
}

__global__ void __launch_bounds__(kFinB) k_batch_finalize(const PatchDev* __restrict__ P, int p, int64_t ncell,
                                                         float den, float rote, float* umass, double* partial) {
  // the single-grid kernel's body with the patch's scale read here
  const double inv_scale = 1.0 / P[p].scale;
  const unsigned long long* acc = P[p].acc;
  const int64_t c0 = ((int64_t)blockIdx.x * kFinB + threadIdx.x) * kFinV;
  double contrib = 0.0;
  if (c0 + kFinV <= ncell) {
    const ulonglong2 a01 = *reinterpret_cast<const ulonglong2*>(acc + c0);
    const ulonglong2 a23 = *reinterpret_cast<const ulonglong2*>(acc + c0 + 2);
    float4 u;
    u.x = fin_cell(a01.x, inv_scale, 0, den, rote);
    u.y = fin_cell(a01.y, inv_scale, 0, den, rote);
    u.z = fin_cell(a23.x, inv_scale, 0, den, rote);
    u.w = fin_cell(a23.y, inv_scale, 0, den, rote);
    *reinterpret_cast<float4*>(umass + c0) = u;
    contrib = (((double)u.x + (double)u.y) + (double)u.z) + (double)u.w;
  } else {
    for (int64_t c = c0; c < ncell; ++c) {
      const float u = fin_cell(acc[c], inv_scale, 0, den, rote);
      umass[c] = u;
      contrib += (double)u;
    }
  }
  for (int off = 32; off > 0; off >>= 1) contrib += __shfl_xor(contrib, off);
  __shared__ double wsum[kFinB / 64];
  if (threadIdx.x % 64 == 0) wsum[threadIdx.x / 64] = contrib;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < kFinB / 64; ++w) t += wsum[w];
    partial[blockIdx.x] = t;
  }
}

__global__ void k_fill_u32(int64_t n, uint32_t* a, uint32_t v) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = v;
}

__global__ void k_scatter_orig_off(int64_t np, const uint32_t* orig, int64_t orig0, const float* a_sorted,
                                   float* a_out) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  a_out[(int64_t)orig[i] - orig0] = a_sorted[i];
}

const char* vvel_shape_name(int s) {
  static const char* names[kVvelShapes] = {"8x4x4", "4x8x4", "4x4x8"};
  return (s >= 0 && s < kVvelShapes) ? names[s] : "?";
}
const char* ddens_shape_name(int s) {
  static const char* names[kDdensShapes] = {"8x8x4+4", "8x8x8+3", "16x8x4+4"};
  return (s >= 0 && s < kDdensShapes) ? names[s] : "?";
}
int64_t vvel_shape_tiles(int s, int n) {
  switch (s) {
    case 0: return VvelShape<8, 4, 4>::tiles(n);
    case 1: return VvelShape<4, 8, 4>::tiles(n);
    default: return VvelShape<4, 4, 8>::tiles(n);
  }
}
int64_t ddens_shape_tiles(int s, int n) {
  switch (s) {
    case 0: return DdensTile<8, 8, 4, 4>::tiles(n);
    case 1: return DdensTile<8, 8, 8, 3>::tiles(n);
    default: return DdensTile<16, 8, 4, 4>::tiles(n);
  }
}
int64_t ppart_chunks(int64_t np) { return (np + kBatchQ - 1) / kBatchQ; }

hipError_t launch_batch_ppart(hipStream_t st, const PatchDev* P, ItemQueue q, int k, unsigned nblocks) {
  if (k <= 32) k_batch_ppart<32><<<nblocks, kBatchQ, 0, st>>>(P, q, k);
  else k_batch_ppart<64><<<nblocks, kBatchQ, 0, st>>>(P, q, k);
  return hipGetLastError();
}

template <int K>
static void vvel_shape_launch(hipStream_t st, const PatchDev* P, ItemQueue q, int k, int shape, unsigned nb) {
  switch (shape) {
    case 0: k_batch_vvel<K, 8, 4, 4><<<nb, kBatchQ, 0, st>>>(P, q, k); break;
    case 1: k_batch_vvel<K, 4, 8, 4><<<nb, kBatchQ, 0, st>>>(P, q, k); break;
    default: k_batch_vvel<K, 4, 4, 8><<<nb, kBatchQ, 0, st>>>(P, q, k); break;
  }
}

hipError_t launch_batch_vvel(hipStream_t st, const PatchDev* P, ItemQueue q, int k, int shape, unsigned nblocks) {
  if (k <= 32) vvel_shape_launch<32>(st, P, q, k, shape, nblocks);
  else vvel_shape_launch<64>(st, P, q, k, shape, nblocks);
  return hipGetLastError();
}

hipError_t launch_batch_ddens(hipStream_t st, const PatchDev* P, ItemQueue q, int shape, unsigned nblocks,
                              unsigned long long* npairs) {
  switch (shape) {
    case 0: k_batch_ddens<8, 8, 4, 4><<<nblocks, 256, 0, st>>>(P, q, npairs); break;
    case 1: k_batch_ddens<8, 8, 8, 3><<<nblocks, 256, 0, st>>>(P, q, npairs); break;
    default: k_batch_ddens<16, 8, 4, 4><<<nblocks, 256, 0, st>>>(P, q, npairs); break;
  }
  return hipGetLastError();
}

int64_t finalize_blocks(int64_t ncell) { return (ncell + (int64_t)kFinB * kFinV - 1) / ((int64_t)kFinB * kFinV); }

hipError_t launch_batch_finalize(hipStream_t st, const PatchDev* P, int p, int64_t ncell, float den, float rote,
                                 float* umass, double* partial) {
  k_batch_finalize<<<(unsigned)finalize_blocks(ncell), kFinB, 0, st>>>(P, p, ncell, den, rote, umass, partial);
  return hipGetLastError();
}

static inline unsigned nblk256(int64_t n) { return (unsigned)((n + 255) / 256); }

hipError_t launch_fill_u32(hipStream_t st, int64_t n, uint32_t* a, uint32_t v) {
  if (n > 0) k_fill_u32<<<nblk256(n), 256, 0, st>>>(n, a, v);
  return hipGetLastError();
}
hipError_t launch_fill_u64(hipStream_t st, int64_t n, unsigned long long* a, unsigned long long v) {
  if (n > 0) k_fill_u64<<<nblk256(n), 256, 0, st>>>(n, a, v);
  return hipGetLastError();
}
hipError_t launch_velrec(hipStream_t st, int64_t np, const uint32_t* orig, const float4* pos, const float* pdens,
                         const float* vx, const float* vy, const float* vz, float4* vel) {
  if (np > 0) k_velrec<<<nblk256(np), 256, 0, st>>>(np, orig, pos, pdens, vx, vy, vz, vel);
  return hipGetLastError();
}
hipError_t launch_fill_h(hipStream_t st, int64_t np, float* hpart, const float* hk) {
  if (np > 0) k_fill_h<<<nblk256(np), 256, 0, st>>>(np, hpart, hk);
  return hipGetLastError();
}
hipError_t launch_reduce_mean(hipStream_t st, const double* partial, int64_t nparts, int64_t ncell, double* slices,
                              double* sum_out, float* mean_out) {
  k_reduce_mean<<<kRedSlices, 256, 0, st>>>(partial, nparts, ncell, slices, 0, sum_out, mean_out);
  k_reduce_mean<<<1, 256, 0, st>>>(slices, kRedSlices, ncell, nullptr, 1, sum_out, mean_out);
  return hipGetLastError();
}
hipError_t launch_div_finalize(hipStream_t st, int n, float dx, int periodic, const float* u, const float* v,
                               const float* w, const float* umass, const float* mean, float* u1co, float* div) {
  k_div_finalize<<<div_grid(n), kDivX * kDivY, 0, st>>>(n, dx, periodic, u, v, w, umass, mean, u1co, div);
  return hipGetLastError();
}
hipError_t launch_convert_check(hipStream_t st, int64_t np, float* m, const float* x, const float* y, const float* z,
                                float half, int* flags) {
  // flags: [0] out of the box, [1] max |mass| bits, [2] the first mass as given, [3] masses differ
  hipError_t e = hipMemsetAsync(flags, 0, 16, st);
  if (e == hipSuccess) e = hipMemcpyAsync(flags + 2, m, 4, hipMemcpyDeviceToDevice, st);
  if (e != hipSuccess) return e;
  k_convert_check<<<(unsigned)std::min<int64_t>(nblk256(np), 2048), 256, 0, st>>>(
      np, m, x, y, z, half, flags, (unsigned*)flags + 1, (const float*)(flags + 2), flags + 3);
  return hipGetLastError();
}
hipError_t launch_convert_vel(hipStream_t st, int64_t np, float* vx, float* vy, float* vz) {
  if (np > 0) k_convert_vel<<<(unsigned)std::min<int64_t>(nblk256(np), 4096), 256, 0, st>>>(np, vx, vy, vz);
  return hipGetLastError();
}
hipError_t launch_gather_sorted(hipStream_t st, int64_t np, const uint32_t* perm, const float* x, const float* y,
                                const float* z, const float* m, float4* pos, uint32_t* orig) {
  if (np > 0) k_gather_sorted<<<nblk256(np), 256, 0, st>>>(np, perm, x, y, z, m, pos, orig);
  return hipGetLastError();
}
hipError_t launch_scatter_orig(hipStream_t st, int64_t np, const uint32_t* orig, int64_t orig0, const float* a,
                               float* out) {
  if (np > 0) k_scatter_orig_off<<<nblk256(np), 256, 0, st>>>(np, orig, orig0, a, out);
  return hipGetLastError();
}

// ---------------------------------------------------------------- host side
static inline unsigned nblk(int64_t n, int b = 256) { return (unsigned)((n + b - 1) / b); }

static Bins make_bins(avism_ctx* c) {
  Bins B;
  B.pos = c->spos4.as<float4>();
  B.orig = c->sorig.as<uint32_t>();
  B.start = c->bin_start.as<int32_t>();
  B.nb = c->nb;
  B.lo = -c->box_L0 / 2.0f;
  B.bsz = c->bsz;
  B.L = c->box_L0;
  B.eps = c->box_L0 * 2e-6f + c->bsz * 1e-5f;
  B.periodic = c->box_periodic;
  return B;
}

int finish_copies(avism_ctx* c) {
  if (!c->copy_pending) return AVISM_OK;
  AV_CHECK_HIP(c, hipStreamSynchronize(c->cstream));
  c->copy_pending = false;
  return AVISM_OK;
}

// Positions and masses go up on the compute stream, which binning and
// PPART_DENS need first; then velocities on the copy stream, so that with
// pinned host arrays their transfer overlaps binning and PPART_DENS (ev_vel
// orders their first reader).  Returns once positions and masses are checked.
int load_particles(avism_ctx* c, int64_t np, const float* x, const float* y, const float* z,
                   const float* vx, const float* vy, const float* vz, const float* m, int where) {
  if (!c->grid_set) AV_FAIL(c, AVISM_ERR_STATE, "avism_set_grid must be called first");
  if (refine_grid_is_atlas(c))
    AV_FAIL(c, AVISM_ERR_STATE, "the grid is an atlas of windows, not the particles' box: avism_set_grid first");
  if (np <= 0) AV_FAIL(c, AVISM_ERR_ARG, "np must be positive");
  if (np >= (int64_t)0x7fffffff) AV_FAIL(c, AVISM_ERR_LIMIT, "more than 2^31-1 particles per call");
  // earlier copies on the copy stream end before any buffer is reused
  AV_CHECK_HIP(c, hipStreamSynchronize(c->cstream));
  c->copy_pending = false;
  size_t b = (size_t)np * 4;
  DBuf* bufs[7] = {&c->px, &c->py, &c->pz, &c->pvx, &c->pvy, &c->pvz, &c->pm};
  const float* src[7] = {x, y, z, vx, vy, vz, m};
  hipMemcpyKind kind = where == AVISM_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
  for (int i = 0; i < 7; i++)
    if (!bufs[i]->ensure(b)) AV_FAIL(c, AVISM_ERR_HIP, "out of device memory (particles)");
  c->np = np;
  c->fields_ready = false;
  c->ppart_valid = false;
  refine_free(c);  // plans hold the previous particles' parent maps
  // [0] out-of-box flag, [1] max |mass| bits, [2] the first mass as given, [3] masses differ
  int* bad = (int*)c->scalars.p;
  {
    StageTimer t(c, "upload_pos_mass");
    for (int i : {0, 1, 2, 6}) AV_CHECK_HIP(c, hipMemcpyAsync(bufs[i]->p, src[i], b, kind, c->stream));
    // the velocities follow on the copy stream once positions and masses are
    // in (sharing the host link would delay what binning needs first), and
    // after all earlier work, which may still read the previous particles
    AV_CHECK_HIP(c, hipEventRecord(c->ev_mark, c->stream));
    AV_CHECK_HIP(c, hipStreamWaitEvent(c->cstream, c->ev_mark, 0));
    AV_CHECK_HIP(c, hipMemsetAsync(bad, 0, 16, c->stream));
    AV_CHECK_HIP(c, hipMemcpyAsync(bad + 2, c->pm.p, 4, hipMemcpyDeviceToDevice, c->stream));
    k_convert_check<<<(unsigned)std::min<int64_t>(nblk(np), 2048), 256, 0, c->stream>>>(
        np, c->pm.as<float>(), c->px.as<float>(), c->py.as<float>(), c->pz.as<float>(), c->g.L0 / 2.0f, bad,
        (unsigned*)bad + 1, (const float*)(bad + 2), bad + 3);
    AV_CHECK_HIP(c, hipGetLastError());
  }
  {
    StageTimer t(c, "upload_vel", c->cstream);
    for (int i = 3; i < 6; i++) {
      if (src[i])
        AV_CHECK_HIP(c, hipMemcpyAsync(bufs[i]->p, src[i], b, kind, c->cstream));
      else  // no velocities (READ_BINARY_PART_2 data): zero, as the reader leaves them
        AV_CHECK_HIP(c, hipMemsetAsync(bufs[i]->p, 0, b, c->cstream));
    }
    if (vx)
      k_convert_vel<<<(unsigned)std::min<int64_t>(nblk(np), 4096), 256, 0, c->cstream>>>(
          np, c->pvx.as<float>(), c->pvy.as<float>(), c->pvz.as<float>());
    AV_CHECK_HIP(c, hipGetLastError());
  }
  AV_CHECK_HIP(c, hipEventRecord(c->ev_vel, c->cstream));
  int hbad[4] = {0, 0, 0, 0};
  AV_CHECK_HIP(c, hipMemcpyAsync(hbad, bad, 16, hipMemcpyDeviceToHost, c->stream));
  AV_CHECK_HIP(c, hipStreamSynchronize(c->stream));
  memcpy(&c->mass_max, &hbad[1], 4);
  c->mass_uniform = hbad[3] == 0;
  if (!std::isfinite(c->mass_max)) AV_FAIL(c, AVISM_ERR_INPUT, "non-finite particle mass");
  if (hbad[0]) AV_FAIL(c, AVISM_ERR_INPUT, "Particles outside the bounding box! Check L0! Particles must be inside [-L0/2, L0/2] in each direction.");
  return AVISM_OK;
}

static int build_bins(avism_ctx* c) {
  StageTimer t(c, "bin");
  const int64_t np = c->np;
  int nb = bins_per_axis(c->g.n, np);
  c->nb = nb;
  c->box_L0 = c->g.L0;
  c->box_periodic = c->g.periodic;
  c->bsz = c->box_L0 / (float)nb;
  int64_t nbins = (int64_t)nb * nb * nb;
  if (!c->bin_start.ensure((size_t)(nbins + 1) * 4) || !c->sbin_key.ensure((size_t)np * 4) ||
      !c->sbin_val.ensure((size_t)np * 4) || !c->sbin_key2.ensure((size_t)np * 4) ||
      !c->sbin_val2.ensure((size_t)np * 4) || !c->spos4.ensure((size_t)np * 16) ||
      !c->vel4.ensure((size_t)np * 16) || !c->sorig.ensure((size_t)np * 4))
    AV_FAIL(c, AVISM_ERR_HIP, "out of device memory (bins)");
  Bins B = make_bins(c);
  k_bin_keys<<<nblk(np), 256, 0, c->stream>>>(np, c->px.as<float>(), c->py.as<float>(),
                                              c->pz.as<float>(), B, c->sbin_key.as<uint32_t>(),
                                              c->sbin_val.as<uint32_t>());
  unsigned bits = 1;
  while ((1ll << bits) < nbins) bits++;
  size_t tmp = 0;
  AV_CHECK_HIP(c, rocprim::radix_sort_pairs(nullptr, tmp, c->sbin_key.as<uint32_t>(),
                                            c->sbin_key2.as<uint32_t>(), c->sbin_val.as<uint32_t>(),
                                            c->sbin_val2.as<uint32_t>(), (size_t)np, 0, bits,
                                            c->stream));
  // the sort's scratch also holds, afterwards, the list of long runs of empty
  // bins (at most nbins / kRun + 2 entries) and its count
  const size_t lr_bytes = ((size_t)(nbins / kRun) + 2) * sizeof(int4);
  if (!c->sort_tmp.ensure(std::max(tmp, lr_bytes + 16))) AV_FAIL(c, AVISM_ERR_HIP, "out of device memory (sort)");
  AV_CHECK_HIP(c, rocprim::radix_sort_pairs(c->sort_tmp.p, tmp, c->sbin_key.as<uint32_t>(),
                                            c->sbin_key2.as<uint32_t>(), c->sbin_val.as<uint32_t>(),
                                            c->sbin_val2.as<uint32_t>(), (size_t)np, 0, bits,
                                            c->stream));
  int4* longrun = c->sort_tmp.as<int4>();
  int32_t* nlong = (int32_t*)((char*)c->sort_tmp.p + lr_bytes);
  AV_CHECK_HIP(c, hipMemsetAsync(nlong, 0, 4, c->stream));
  k_bin_starts<<<nblk(np + 1), 256, 0, c->stream>>>(np, nbins, c->sbin_key2.as<uint32_t>(),
                                                    c->bin_start.as<int32_t>(), longrun, nlong);
  k_bin_runs<<<1024, 256, 0, c->stream>>>(longrun, nlong, c->bin_start.as<int32_t>());
  k_gather_sorted<<<nblk(np), 256, 0, c->stream>>>(
      np, c->sbin_val2.as<uint32_t>(), c->px.as<float>(), c->py.as<float>(), c->pz.as<float>(),
      c->pm.as<float>(), c->spos4.as<float4>(), c->sorig.as<uint32_t>());
  AV_CHECK_HIP(c, hipGetLastError());
  AV_CHECK_HIP(c, hipStreamSynchronize(c->stream));
  return AVISM_OK;
}

// f(std::integral_constant<int, K>) for the smallest K in {32, 64, 128, 256}
// with k <= K (k <= 256 checked by the caller): the selection slots per lane
template <class F>
static void with_k(int k, F&& f) {
  if (k <= 32)
    f(std::integral_constant<int, 32>{});
  else if (k <= 64)
    f(std::integral_constant<int, 64>{});
  else if (k <= 128)
    f(std::integral_constant<int, 128>{});
  else
    f(std::integral_constant<int, 256>{});
}

// blocks of k_vvel: tiles of 8 x 4 x (lanes / 32) cells
static unsigned vvel_tiles(int n, int K) {
  const int tz = kq_of(K) / 32;
  return (unsigned)(((n + 7) / 8) * ((n + 3) / 4) * ((n + tz - 1) / tz));
}

int fields_from_particles(avism_ctx* c, int32_t k, int32_t vel_interp, int32_t dens_interp) {
  if (c->np <= 0) AV_FAIL(c, AVISM_ERR_STATE, "no particles loaded");
  if (refine_grid_is_atlas(c))
    AV_FAIL(c, AVISM_ERR_STATE, "the grid is an atlas of windows: its fields come from avism_refine_load");
  if (k <= 0) AV_FAIL(c, AVISM_ERR_INPUT, "KNEIGHBOURS must be positive");
  if (k > 256) AV_FAIL(c, AVISM_ERR_UNSUPPORTED, "KNEIGHBOURS above 256 is not supported by this build");
  if (k > c->np) AV_FAIL(c, AVISM_ERR_INPUT, "KNEIGHBOURS exceeds the number of particles");
  // FLAG_VEL_INTERP / FLAG_DENS_INTERP: 0 = TSC, 1 = SPH (avism.f90:1178-1179, 1220-1221)
  // vel_interp -1: FLAG_VEL_AVAILABLE = 0 (avism_particles_to_density)
  if ((vel_interp != -1 && vel_interp != 0 && vel_interp != 1) || (dens_interp != 0 && dens_interp != 1))
    AV_FAIL(c, AVISM_ERR_INPUT, "FLAG_VEL_INTERP and FLAG_DENS_INTERP must be 0 (TSC) or 1 (SPH)");
  const int64_t np = c->np;
  const int n = c->g.n;
  const int64_t ncell = (int64_t)n * n * n;
  if (int rc = finish_copies(c)) return rc;
  for (int64_t& v : c->counters) v = 0;
  c->growth_pending = false;
  c->nvoid = 0;  // cubes, MARCA and the catalogue of earlier fields are stale
  c->post_ready = false;
  c->ppart_valid = false;
  int rc = build_bins(c);
  if (rc) return rc;
  if (!c->pdens.ensure((size_t)np * 4) || !c->hk.ensure((size_t)np * 4) ||
      !c->hpart.ensure((size_t)np * 4) || !c->norm.ensure((size_t)np * 4) ||
      !c->fu.ensure((size_t)ncell * 4) || !c->fv.ensure((size_t)ncell * 4) ||
      !c->fw.ensure((size_t)ncell * 4) || !c->u_mass.ensure((size_t)ncell * 4) ||
      !c->u1co.ensure((size_t)ncell * 4) || !c->div.ensure((size_t)ncell * 4))
    AV_FAIL(c, AVISM_ERR_HIP, "out of device memory (fields)");
  float* d_rad = (float*)c->scalars.ensure(scalars_bytes(c));
  float* rad = d_rad + 1024;
  AV_CHECK_HIP(c, hipMemcpyAsync(rad, c->g.rad.data(), c->g.rad.size() * 4, hipMemcpyHostToDevice,
                                 c->stream));
  Bins B = make_bins(c);
  // SMASK: VVEL_INTERP_SPH_VW and H_DISTANCE skip masked cells (particles.f90:312, 856)
  const int8_t* smask = c->has_mask ? c->smask.as<int8_t>() : nullptr;
  // first search radius: 1.3 x the k-th neighbour distance at the mean density
  const double rho = (double)np / ((double)c->g.L0 * c->g.L0 * c->g.L0);
  const float r0 = (float)(1.3 * std::cbrt(3.0 * k / (4.0 * kPi * rho)));
  AV_CHECK_HIP(c, hipMemsetAsync(c->hpart.p, 0, (size_t)np * 4, c->stream));
  const TscGeom T{rad, c->g.rlo, n, c->g.periodic, c->g.dx};
  if (vel_interp == 1) {
    {
      StageTimer t(c, "ppart_dens");
      // the query order: blocks of 4^3 bins (counts, offsets, then positions)
      const int nbb = (c->nb + kPB - 1) / kPB;
      const int64_t nblocks = (int64_t)nbb * nbb * nbb;
      // (in the binning scratch, free here; with far more bins than particles
      // the counts do not fit there and the queries keep the bin order)
      const uint32_t* perm = nullptr;
      if (nblocks < np) {
        int32_t* bcnt = c->sbin_key.as<int32_t>();
        int32_t* boff = c->sbin_val.as<int32_t>();
        k_pblock_count<<<nblk(nblocks), 256, 0, c->stream>>>(B, nbb, bcnt);
        size_t tmp = 0;
        AV_CHECK_HIP(c, rocprim::exclusive_scan(nullptr, tmp, bcnt, boff, 0, (size_t)nblocks,
                                                rocprim::plus<int32_t>(), c->stream));
        if (!c->sort_tmp.ensure(tmp)) AV_FAIL(c, AVISM_ERR_HIP, "out of device memory (scan)");
        AV_CHECK_HIP(c, rocprim::exclusive_scan(c->sort_tmp.p, tmp, bcnt, boff, 0, (size_t)nblocks,
                                                rocprim::plus<int32_t>(), c->stream));
        k_pblock_write<<<nblk(nblocks, 4), 256, 0, c->stream>>>(B, nbb, boff, c->sbin_val2.as<uint32_t>());
        perm = c->sbin_val2.as<uint32_t>();
      }
      with_k(k, [&](auto KC) {
        constexpr int K = decltype(KC)::value;
        k_ppart<K><<<nblk(np, kq_of(K)), kq_of(K), 0, c->stream>>>(B, np, k, r0, c->mass_uniform ? 1 : 0,
                                                                   c->pdens.as<float>(), c->hk.as<float>(), perm);
      });
      AV_CHECK_HIP(c, hipGetLastError());
    }
    c->counters[1] = np;
    c->ppart_valid = true;
    AV_CHECK_HIP(c, hipStreamWaitEvent(c->stream, c->ev_vel, 0));  // velocities uploaded
    k_velrec<<<nblk(np), 256, 0, c->stream>>>(np, c->sorig.as<uint32_t>(), c->spos4.as<float4>(),
                                              c->pdens.as<float>(), c->pvx.as<float>(), c->pvy.as<float>(),
                                              c->pvz.as<float>(), c->vel4.as<float4>());
    {
      StageTimer t(c, "vvel_sph_vw");
      const VvelOut O{c->vel4.as<float4>(), c->hpart.as<uint32_t>(), c->fu.as<float>(), c->fv.as<float>(),
                      c->fw.as<float>(), 1};
      if (smask) {  // masked cells keep zero velocity (particles.f90:856)
        AV_CHECK_HIP(c, hipMemsetAsync(c->fu.p, 0, (size_t)ncell * 4, c->stream));
        AV_CHECK_HIP(c, hipMemsetAsync(c->fv.p, 0, (size_t)ncell * 4, c->stream));
        AV_CHECK_HIP(c, hipMemsetAsync(c->fw.p, 0, (size_t)ncell * 4, c->stream));
      }
      with_k(k, [&](auto KC) {
        constexpr int K = decltype(KC)::value;
        if (vvel_chk(ncell, np))
          k_vvel<K, true><<<vvel_tiles(n, K), kq_of(K), 0, c->stream>>>(B, np, n, rad, c->g.rlo, c->g.dx, k, r0,
                                                                        smask, O);
        else
          k_vvel<K, false><<<vvel_tiles(n, K), kq_of(K), 0, c->stream>>>(B, np, n, rad, c->g.rlo, c->g.dx, k, r0,
                                                                         smask, O);
      });
      c->counters[2] = smask ? c->mask_observed : ncell;
      AV_CHECK_HIP(c, hipGetLastError());
    }
    {
      StageTimer t(c, "fill_h");
      k_fill_h<<<nblk(np), 256, 0, c->stream>>>(np, c->hpart.as<float>(), c->hk.as<float>());
      AV_CHECK_HIP(c, hipGetLastError());
    }
  } else if (vel_interp == -1) {  // no velocities: U2..U4 stay 0, as PART_DENS
    AV_CHECK_HIP(c, hipMemsetAsync(c->pdens.p, 0, (size_t)np * 4, c->stream));
    AV_CHECK_HIP(c, hipMemsetAsync(c->fu.p, 0, (size_t)ncell * 4, c->stream));
    AV_CHECK_HIP(c, hipMemsetAsync(c->fv.p, 0, (size_t)ncell * 4, c->stream));
    AV_CHECK_HIP(c, hipMemsetAsync(c->fw.p, 0, (size_t)ncell * 4, c->stream));
  } else {  // VEL_INTERP_TSC; PART_DENS stays 0 as in the reference
    if (!c->tsc_acc.ensure((size_t)ncell * 32)) AV_FAIL(c, AVISM_ERR_HIP, "out of device memory (TSC)");
    AV_CHECK_HIP(c, hipMemsetAsync(c->pdens.p, 0, (size_t)np * 4, c->stream));
    StageTimer t(c, "vel_tsc");
    AV_CHECK_HIP(c, hipStreamWaitEvent(c->stream, c->ev_vel, 0));  // velocities uploaded
    k_fill_u64<<<nblk(ncell * 4), 256, 0, c->stream>>>(ncell * 4, c->tsc_acc.as<unsigned long long>(), 0ull);
    k_tsc<<<nblk(np), 256, 0, c->stream>>>(T, np, c->px.as<float>(), c->py.as<float>(), c->pz.as<float>(),
                                           c->pm.as<float>(), c->pvx.as<float>(), c->pvy.as<float>(),
                                           c->pvz.as<float>(), nullptr, c->tsc_acc.as<double>());
    k_tsc_vel_finalize<<<nblk(ncell), 256, 0, c->stream>>>(ncell, c->tsc_acc.as<double>(), c->fu.as<float>(),
                                                           c->fv.as<float>(), c->fw.as<float>());
    AV_CHECK_HIP(c, hipGetLastError());
  }
  if (dens_interp == 1 && vel_interp != 1) {  // H_DISTANCE (particles.f90:276-351), then FILL_H_ZEROS
    StageTimer t(c, "h_distance");
    c->counters[2] = smask ? c->mask_observed : ncell;
    const VvelOut O{nullptr, c->hpart.as<uint32_t>(), nullptr, nullptr, nullptr, 0};
    with_k(k, [&](auto KC) {
      constexpr int K = decltype(KC)::value, BQ = kq_of(K);
      if (vvel_chk(ncell, np))
        k_vvel<K, true><<<vvel_tiles(n, K), BQ, 0, c->stream>>>(B, np, n, rad, c->g.rlo, c->g.dx, k, r0, smask, O);
      else
        k_vvel<K, false><<<vvel_tiles(n, K), BQ, 0, c->stream>>>(B, np, n, rad, c->g.rlo, c->g.dx, k, r0, smask, O);
      if (c->g.periodic)
        k_fill_h_knn<1, K><<<nblk(np, BQ), BQ, 0, c->stream>>>(B, np, k, r0, c->hpart.as<float>());
      else
        k_fill_h_knn<0, K><<<nblk(np, BQ), BQ, 0, c->stream>>>(B, np, k, r0, c->hpart.as<float>());
    });
    AV_CHECK_HIP(c, hipGetLastError());
  }
  DdensGeom G{rad, c->g.rlo, n, c->g.dx, c->g.rx1};
  // cosmology constants (avism.f90:469-479, 599-606, 1249-1250)
  float den = ((c->g.dx * c->g.dx) * c->g.dx) * ((c->rete * c->rete) * c->rete);
  const unsigned grid = (unsigned)((ncell + (int64_t)kFinB * kFinV - 1) / ((int64_t)kFinB * kFinV));
  int64_t nparts = grid;
  if (!c->partial64.ensure((size_t)nparts * 8) || !c->mass_fix.ensure((size_t)ncell * 8))
    AV_FAIL(c, AVISM_ERR_HIP, "out of device memory");
  // fixed point: every cell sum stays below np * max mass * scale <= 2^62
  double bound = (double)np * (double)(c->mass_max > 0.0f ? c->mass_max : 1.0f);
  int e2 = 0;
  std::frexp(bound, &e2);  // bound < 2^e2
  const double scale = std::ldexp(1.0, 62 - e2);
  unsigned long long* dpairs = (unsigned long long*)((char*)c->scalars.p + 1024);
  if (dens_interp == 1) {
    StageTimer t(c, "ddens_scatter");
    k_fill_u64<<<nblk(ncell), 256, 0, c->stream>>>(ncell, c->mass_fix.as<unsigned long long>(), 0ull);
    const unsigned ntiles = (unsigned)(((n + DTX - 1) / DTX) * ((n + DTY - 1) / DTY) * ((n + DTZ - 1) / DTZ));
    AV_CHECK_HIP(c, hipMemsetAsync(dpairs, 0, 8, c->stream));
    if (vvel_chk(ncell, np) && c->nb <= n) {
      // a grid finer than the particles: the window path with one window, the
      // whole grid (NORM per particle, then the deposits tile by tile in LDS)
      if (!c->pmark.ensure((size_t)np)) AV_FAIL(c, AVISM_ERR_HIP, "out of device memory (DDENS marks)");
      uint32_t* dmax = (uint32_t*)((char*)c->scalars.p + 1792);
      AV_CHECK_HIP(c, launch_fill_u32(c->stream, 1, dmax, 0u));
      AV_CHECK_HIP(c, launch_max_bits(c->stream, np, c->hpart.as<uint32_t>(), dmax));
      uint32_t hbits = 0;
      AV_CHECK_HIP(c, hipMemcpyAsync(&hbits, dmax, 4, hipMemcpyDeviceToHost, c->stream));
      AV_CHECK_HIP(c, hipStreamSynchronize(c->stream));
      float hmax = 0.0f;
      memcpy(&hmax, &hbits, 4);
      struct {
        WinDev w;
        int64_t off[4];
      } h{};
      h.w.w = n;
      h.w.bnx = h.w.bny = h.w.bnz = c->nb;
      WinDev* dw = (WinDev*)((char*)c->scalars.p + 1536);
      const int64_t items = win_ddens_items(h.w), dep = win_dep_items(n);
      h.off[1] = items;
      h.off[3] = dep;
      AV_CHECK_HIP(c, hipMemcpy(dw, &h, sizeof(h), hipMemcpyHostToDevice));
      const int64_t* doff = (const int64_t*)(dw + 1);
      const VirtG V{rad, c->g.rlo, n, c->g.periodic, c->g.dx, c->g.rx1, n};
      AV_CHECK_HIP(c, launch_win_ddens(c->stream, V, B, np, c->hpart.as<float>(), scale,
                                       c->mass_fix.as<unsigned long long>(), dw, doff, 1, items, dpairs,
                                       c->norm.as<float>(), c->pmark.as<uint8_t>(), c->sbin_val.as<int32_t>(),
                                       (int32_t*)((char*)c->scalars.p + 1800), c->sort_tmp, hmax, doff + 2, dep));
    } else if (c->nb > n)  // cells split into bins
      k_ddens_scatter<4><<<ntiles, 256, 0, c->stream>>>(G, B, c->hpart.as<float>(), scale,
                                                        c->mass_fix.as<unsigned long long>(), c->norm.as<float>(),
                                                        dpairs);
    else
      k_ddens_scatter<1><<<ntiles, 256, 0, c->stream>>>(G, B, c->hpart.as<float>(), scale,
                                                        c->mass_fix.as<unsigned long long>(), c->norm.as<float>(),
                                                        dpairs);
    AV_CHECK_HIP(c, hipGetLastError());
  } else {  // DENS_INTERP_TSC: FP64 cell masses in mass_fix (zero bits are +0.0)
    StageTimer t(c, "dens_tsc");
    k_fill_u64<<<nblk(ncell), 256, 0, c->stream>>>(ncell, c->mass_fix.as<unsigned long long>(), 0ull);
    k_tsc<<<nblk(np), 256, 0, c->stream>>>(T, np, c->px.as<float>(), c->py.as<float>(), c->pz.as<float>(),
                                           c->pm.as<float>(), c->pvx.as<float>(), c->pvy.as<float>(),
                                           c->pvz.as<float>(), c->mass_fix.as<double>(), nullptr);
    AV_CHECK_HIP(c, hipGetLastError());
  }
  c->mass_inv_scale = 1.0 / scale;
  c->mass_fp64 = dens_interp == 1 ? 0 : 1;
  c->mass_n = n;
  {
    StageTimer t(c, "ddens_finalize");
    k_ddens_finalize<<<grid, kFinB, 0, c->stream>>>(ncell, c->mass_fix.as<unsigned long long>(), 1.0 / scale,
                                                  dens_interp == 1 ? 0 : 1, c->raw_mass ? 1.0f : den,
                                                  c->raw_mass ? 1.0f : c->rote, c->u_mass.as<float>(),
                                                  c->partial64.as<double>());
    AV_CHECK_HIP(c, hipGetLastError());
  }
  double* dsum = (double*)((char*)c->scalars.p + 256);
  float* dmean = (float*)((char*)c->scalars.p + 264);
  {
    StageTimer t(c, "mean_div");
    double* slices = (double*)((char*)c->scalars.p + scalars_slices_offset(c));
    k_reduce_mean<<<kRedSlices, 256, 0, c->stream>>>(c->partial64.as<double>(), nparts, ncell, slices, 0, dsum,
                                                     dmean);
    k_reduce_mean<<<1, 256, 0, c->stream>>>(slices, kRedSlices, ncell, nullptr, 1, dsum, dmean);
    if (vel_interp == -1) {
      const float lin = (powf(c->omega0, 0.6f) * c->rete) * c->hte;  // avism.f90:1648
      k_div_linear<<<nblk(ncell), 256, 0, c->stream>>>(ncell, lin, c->u_mass.as<float>(), dmean,
                                                      c->u1co.as<float>(), c->div.as<float>());
    } else {
      k_div_finalize<<<div_grid(n), kDivX * kDivY, 0, c->stream>>>(n, c->g.dx, c->g.periodic, c->fu.as<float>(),
                                                c->fv.as<float>(), c->fw.as<float>(),
                                                c->u_mass.as<float>(), dmean, c->u1co.as<float>(),
                                                c->div.as<float>());
    }
    AV_CHECK_HIP(c, hipGetLastError());
  }
  float hmean = 0;
  unsigned long long hpairs = 0;
  if (dens_interp == 1)
    AV_CHECK_HIP(c, hipMemcpyAsync(&hpairs, dpairs, 8, hipMemcpyDeviceToHost, c->stream));
  AV_CHECK_HIP(c, hipMemcpyAsync(&hmean, dmean, 4, hipMemcpyDeviceToHost, c->stream));
  AV_CHECK_HIP(c, hipMemcpyAsync(&c->meansum, dsum, 8, hipMemcpyDeviceToHost, c->stream));
  AV_CHECK_HIP(c, hipStreamSynchronize(c->stream));
  c->meandens_raw = hmean;
  c->counters[0] = (int64_t)hpairs;
  float ul3 = (kUL * kUL) * kUL;
  c->meandens = hmean * (kUM / ul3);  // avism.f90:1409
  c->fields_ready = true;
  return AVISM_OK;
}

// ---- survey mode (FLAG_SURVEY = 1, avism.f90:1415-1597)

static void swap_buf(DBuf& a, DBuf& b) {
  std::swap(a.p, b.p);
  std::swap(a.bytes, b.bytes);
}

void swap_particle_sets(avism_ctx* c) {
  c->ppart_valid = false;
  swap_buf(c->px, c->rpx); swap_buf(c->py, c->rpy); swap_buf(c->pz, c->rpz);
  swap_buf(c->pvx, c->rpvx); swap_buf(c->pvy, c->rpvy); swap_buf(c->pvz, c->rpvz); swap_buf(c->pm, c->rpm);
  swap_buf(c->spos4, c->rspos4); swap_buf(c->sorig, c->rsorig); swap_buf(c->vel4, c->rvel4);
  swap_buf(c->bin_start, c->rbin_start);
  swap_buf(c->pdens, c->rpdens); swap_buf(c->hk, c->rhk); swap_buf(c->hpart, c->rhpart); swap_buf(c->norm, c->rnorm);
  std::swap(c->np, c->nrand); std::swap(c->nb, c->rnb); std::swap(c->bsz, c->rbsz);
  std::swap(c->mass_max, c->rmass_max); std::swap(c->mass_uniform, c->rmass_uniform);
}

// per-block FP64 sums over the observed cells (SMASK > 0) of a and b, in a
// fixed order (grid-stride per thread, then a tree in LDS): partial[2 * block]
__global__ void __launch_bounds__(256) k_mask_sums(int64_t ncell, const float* __restrict__ a,
                                                   const float* __restrict__ b,
                                                   const int8_t* __restrict__ smask, double* partial) {
  __shared__ double sa[256], sb[256];
  double xa = 0.0, xb = 0.0;
  for (int64_t c = (int64_t)blockIdx.x * 256 + threadIdx.x; c < ncell; c += (int64_t)gridDim.x * 256) {
    if (smask[c] == 0) continue;
    xa += (double)a[c];
    if (b) xb += (double)b[c];
  }
  sa[threadIdx.x] = xa;
  sb[threadIdx.x] = xb;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) {
      sa[threadIdx.x] += sa[threadIdx.x + h];
      sb[threadIdx.x] += sb[threadIdx.x + h];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    partial[2 * blockIdx.x] = sa[0];
    partial[2 * blockIdx.x + 1] = sb[0];
  }
}

// out[0..1] = the sums of the nparts partial pairs, in block order
__global__ void k_sum_pairs(int nparts, const double* __restrict__ partial, double* out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double a = 0.0, b = 0.0;
  for (int i = 0; i < nparts; ++i) {
    a += partial[2 * i];
    b += partial[2 * i + 1];
  }
  out[0] = a;
  out[1] = b;
}

// D/(alpha R) per observed cell (avism.f90:1562-1581): alpha = DSUM/RSUM,
// the ratio clipped to [RATIOFLOOR, RATIOMAX] in FP64, stored in FP32;
// unobserved cells keep U1CO
__global__ void k_survey_ratio(int64_t ncell, const double* __restrict__ sums, const float* __restrict__ ugal,
                               const float* __restrict__ urand, const int8_t* __restrict__ smask, float* u1co) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncell) return;
  const float u = ugal[c];
  if (smask[c] == 0) {
    u1co[c] = u;
    return;
  }
  const double alpha = sums[0] / sums[1];
  const float r = urand[c];
  double b = r > 0.0f ? (double)u / (alpha * (double)r) : (double)1e-4f;
  b = fmax((double)1e-4f, b);
  b = fmin(1000.0, b);
  u1co[c] = (float)b;
}

// U1CO / MEANRATIO on observed cells (MEANRATIO = REAL*4 of the FP64 mean,
// avism.f90:1584-1588), then the linear-theory DIVERCO over every cell
__global__ void k_survey_final(int64_t ncell, const double* __restrict__ msum, double count, float lin,
                               const int8_t* __restrict__ smask, float* u1co, float* div) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncell) return;
  const float meanratio = (float)(msum[0] / count);
  float u = u1co[c];
  if (smask[c] != 0) u = u / meanratio;
  u1co[c] = u;
  div[c] = -(lin * (u - 1.0f));
}

int survey_density(avism_ctx* c, int32_t k, int32_t dens_interp) {
  if (!c->has_mask) AV_FAIL(c, AVISM_ERR_STATE, "survey mode needs the survey mask (avism_set_mask)");
  if (c->nrand <= 0) AV_FAIL(c, AVISM_ERR_STATE, "survey mode needs the randoms (avism_load_randoms)");
  if (c->np <= 0) AV_FAIL(c, AVISM_ERR_STATE, "no particles loaded");
  if (k <= 0) AV_FAIL(c, AVISM_ERR_INPUT, "KNEIGHBOURS must be positive");
  const int n = c->g.n;
  const int64_t ncell = (int64_t)n * n * n;
  // randoms: DENS_INTERP_TSC, or H_DISTANCE and FILL_H_ZEROS with
  // KNEIGHBOURS2 = MIN(8 KNEIGHBOURS, 256), then DDENS (avism.f90:1459-1500);
  // U1RAND is the cell mass itself
  const int k2 = std::min(8 * k, 256);
  swap_particle_sets(c);
  c->raw_mass = true;
  int rc = fields_from_particles(c, k2, -1, dens_interp);
  c->raw_mass = false;
  swap_particle_sets(c);
  if (rc) return rc;
  if (!c->u_rand.ensure((size_t)ncell * 4)) AV_FAIL(c, AVISM_ERR_HIP, "out of device memory (survey)");
  AV_CHECK_HIP(c, hipMemcpyAsync(c->u_rand.p, c->u_mass.p, (size_t)ncell * 4, hipMemcpyDeviceToDevice,
                                 c->stream));
  // galaxies: the density without velocities; u_mass is then U1CO * ROTE's
  // value before the mean (avism.f90:1386)
  rc = fields_from_particles(c, k, -1, dens_interp);
  if (rc) return rc;
  StageTimer t(c, "survey_ratio");
  const int nparts = (int)std::min<int64_t>(nblk(ncell), 1024);
  if (!c->partial64.ensure((size_t)nparts * 16 + 64)) AV_FAIL(c, AVISM_ERR_HIP, "out of device memory");
  double* part = c->partial64.as<double>();
  double* sums = part + 2 * nparts;  // DSUM, RSUM; then the ratio sum
  const int8_t* sm = c->smask.as<int8_t>();
  k_mask_sums<<<nparts, 256, 0, c->stream>>>(ncell, c->u_mass.as<float>(), c->u_rand.as<float>(), sm, part);
  k_sum_pairs<<<1, 64, 0, c->stream>>>(nparts, part, sums);
  k_survey_ratio<<<nblk(ncell), 256, 0, c->stream>>>(ncell, sums, c->u_mass.as<float>(), c->u_rand.as<float>(),
                                                      sm, c->u1co.as<float>());
  k_mask_sums<<<nparts, 256, 0, c->stream>>>(ncell, c->u1co.as<float>(), nullptr, sm, part);
  k_sum_pairs<<<1, 64, 0, c->stream>>>(nparts, part, sums + 2);
  AV_CHECK_HIP(c, hipGetLastError());
  // the observed-cell count, on the host from the mask
  int64_t nobs = 0;
  {
    std::vector<int8_t> hm((size_t)ncell);
    AV_CHECK_HIP(c, hipMemcpyAsync(hm.data(), sm, (size_t)ncell, hipMemcpyDeviceToHost, c->stream));
    AV_CHECK_HIP(c, hipStreamSynchronize(c->stream));
    for (int8_t v : hm) nobs += v > 0;
  }
  if (nobs == 0) AV_FAIL(c, AVISM_ERR_INPUT, "the survey mask has no observed cell");
  const float lin = (powf(c->omega0, 0.6f) * c->rete) * c->hte;  // avism.f90:1648
  k_survey_final<<<nblk(ncell), 256, 0, c->stream>>>(ncell, sums + 2, (double)nobs, lin, sm, c->u1co.as<float>(),
                                                      c->div.as<float>());
  AV_CHECK_HIP(c, hipGetLastError());
  AV_CHECK_HIP(c, hipStreamSynchronize(c->stream));
  const float ul3 = (kUL * kUL) * kUL;
  c->meandens = c->rote * (kUM / ul3);  // avism.f90:1590
  c->fields_ready = true;
  return AVISM_OK;
}

int fields_from_grid(avism_ctx* c, const float* delta, const float* vx, const float* vy,
                     const float* vz, int where) {
  if (!c->grid_set) AV_FAIL(c, AVISM_ERR_STATE, "avism_set_grid must be called first");
  if (int rc = finish_copies(c)) return rc;
  c->nvoid = 0;
  c->post_ready = false;
  const int n = c->g.n;
  const int64_t ncell = (int64_t)n * n * n;
  if (!c->fu.ensure((size_t)ncell * 4) || !c->fv.ensure((size_t)ncell * 4) ||
      !c->fw.ensure((size_t)ncell * 4) || !c->u_mass.ensure((size_t)ncell * 4) ||
      !c->u1co.ensure((size_t)ncell * 4) || !c->div.ensure((size_t)ncell * 4) ||
      !c->px.ensure((size_t)ncell * 4) || !c->py.ensure((size_t)ncell * 4) ||
      !c->pz.ensure((size_t)ncell * 4) || !c->pm.ensure((size_t)ncell * 4))
    AV_FAIL(c, AVISM_ERR_HIP, "out of device memory (grid input)");
  hipMemcpyKind kind = where == AVISM_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
  size_t b = (size_t)ncell * 4;
  AV_CHECK_HIP(c, hipMemcpyAsync(c->px.p, delta, b, kind, c->stream));
  AV_CHECK_HIP(c, hipMemcpyAsync(c->py.p, vx, b, kind, c->stream));
  AV_CHECK_HIP(c, hipMemcpyAsync(c->pz.p, vy, b, kind, c->stream));
  AV_CHECK_HIP(c, hipMemcpyAsync(c->pm.p, vz, b, kind, c->stream));
  k_grid_input<<<nblk(ncell), 256, 0, c->stream>>>(ncell, c->px.as<float>(), c->py.as<float>(),
                                                   c->pz.as<float>(), c->pm.as<float>(),
                                                   c->u1co.as<float>(), c->fu.as<float>(),
                                                   c->fv.as<float>(), c->fw.as<float>());
  k_div_finalize<<<div_grid(n), kDivX * kDivY, 0, c->stream>>>(n, c->g.dx, c->g.periodic, c->fu.as<float>(),
                                            c->fv.as<float>(), c->fw.as<float>(), nullptr, nullptr,
                                            nullptr, c->div.as<float>());
  AV_CHECK_HIP(c, hipGetLastError());
  AV_CHECK_HIP(c, hipStreamSynchronize(c->stream));
  // grid input: U1CO is 1+delta and MEANDENS the background (avism.f90:1599-1600)
  float ul3 = (kUL * kUL) * kUL;
  c->meandens = c->rote * (kUM / ul3);
  c->meandens_raw = 1.0f;
  c->np = 0;
  c->fields_ready = true;
  return AVISM_OK;
}

int get_particle_fields(avism_ctx* c, float* part_dens, float* hpart) {
  if (c->np <= 0 || !c->fields_ready) AV_FAIL(c, AVISM_ERR_STATE, "no particle fields computed");
  const int64_t np = c->np;
  float* tmp = (float*)c->sbin_key.ensure((size_t)np * 4);
  float* srcs[2] = {c->pdens.as<float>(), c->hpart.as<float>()};  // bin-sorted on the device
  float* dsts[2] = {part_dens, hpart};
  for (int i = 0; i < 2; i++) {
    if (!dsts[i]) continue;
    k_scatter_orig<<<nblk(np), 256, 0, c->stream>>>(np, c->sorig.as<uint32_t>(), srcs[i], tmp);
    AV_CHECK_HIP(c, hipMemcpyAsync(dsts[i], tmp, (size_t)np * 4, hipMemcpyDeviceToHost, c->stream));
    AV_CHECK_HIP(c, hipStreamSynchronize(c->stream));
  }
  return AVISM_OK;
}

// ------------------------------------------------ windows of a refined level
// The single-grid kernel bodies on the cells of cubic windows of a level's
// grid, written into an atlas (refine.h).  A window cell's query point,
// neighbours, weights, NORM and deposits are those the whole-grid kernels
// compute for the same cell of the level's grid.

__device__ __forceinline__ int wrapn(int v, int n) {
  v %= n;
  return v < 0 ? v + n : v;
}

// the window of item it: the last W with off[W] <= it
__device__ __forceinline__ int item_window(const int64_t* __restrict__ off, int nwin, int64_t it) {
  int lo = 0, hi = nwin;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (off[mid] <= it) lo = mid;
    else hi = mid;
  }
  return lo;
}

int64_t win_vvel_items(int32_t w, int32_t k) {
  const int tz = (k <= 64 ? 128 : 64) / 32;  // kq_of of the slot class of k, over 32
  return (int64_t)((w + 7) / 8) * ((w + 3) / 4) * ((w + tz - 1) / tz);
}
int64_t win_ddens_items(const WinDev& wd) {
  return (int64_t)((wd.bnx + kWinBins - 1) / kWinBins) * ((wd.bny + kWinBins - 1) / kWinBins) *
         ((wd.bnz + kWinBins - 1) / kWinBins);
}
int64_t win_cell_items(int32_t w) { return ((int64_t)w * w * w + 255) / 256; }

// VVEL_INTERP_SPH_VW on the windows: tiles of 8 x 4 x (lanes / 32) window cells
template <int K, bool CHK>
__global__ void __launch_bounds__(kq_of(K)) k_win_vvel(Bins B, int64_t np, const WinDev* __restrict__ W,
                                                      const int64_t* __restrict__ off, int nwin, VirtG V, int k,
                                                      float r0, VvelOut O) {
  constexpr int BQ = kq_of(K), TZ = BQ / 32;
  __shared__ uint32_t keys_s[K * BQ], gmx_s[K / 4 * BQ];
  const int64_t it = xcd_block(blockIdx.x, gridDim.x);
  const int wi = item_window(off, nwin, it);
  const WinDev wd = W[wi];
  const int64_t t = it - off[wi];
  const int w = wd.w, ntx = (w + 7) / 8, nty = (w + 3) / 4;
  const int tid = threadIdx.x;
  const int lx = (int)(t % ntx) * 8 + tid % 8, ly = (int)((t / ntx) % nty) * 4 + (tid / 8) % 4,
            lz = (int)(t / ((int64_t)ntx * nty)) * TZ + tid / 32;
  if (lx >= w || ly >= w || lz >= w) return;
  int gx = wd.ox + lx, gy = wd.oy + ly, gz = wd.oz + lz;
  if (V.periodic) {
    gx = wrapn(gx, V.n);
    gy = wrapn(gy, V.n);
    gz = wrapn(gz, V.n);
  }
  const float qx = V.rad[gx + 1 - V.rlo], qy = V.rad[gy + 1 - V.rlo], qz = V.rad[gz + 1 - V.rlo];
  const size_t c =
      (size_t)(wd.ax + lx) + (size_t)V.A * ((size_t)(wd.ay + ly) + (size_t)V.A * (size_t)(wd.az + lz));
  if (B.periodic)
    vvel_one<1, K, CHK>(B, np, k, r0, V.dx, qx, qy, qz, c, keys_s + tid, gmx_s + tid, O);
  else
    vvel_one<0, K, CHK>(B, np, k, r0, V.dx, qx, qy, qz, c, keys_s + tid, gmx_s + tid, O);
}

// cell c (1-based) of the level's grid -> its index in the window, or -1
__device__ __forceinline__ int win_local(int c, int o, int w, int n, int periodic) {
  int l = c - 1 - o;
  if (periodic) l = wrapn(l, n);
  return (l >= 0 && l < w) ? l : -1;
}
// some cell of lo..hi (1-based, within 1..n) lies in the window
__device__ __forceinline__ bool win_hits(int lo, int hi, int o, int w, int n, int periodic) {
  if (lo > hi) return false;
  if (!periodic) return max(lo - 1, o) < min(hi, o + w);
  const int l0 = wrapn(lo - 1 - o, n);
  return l0 < w || l0 + (hi - lo + 1) > n;
}

// A particle's cell box on the level's grid (1-based cells, the reference's
// box clipped to the grid and to the centres within h, as in ddens_tile) and
// its home cell.
struct NormBox {
  float4 p;
  float h, inv_h, h2max;
  int ix, iy, iz;          // home cell (may lie outside 1..n)
  int i0, i1, ia, ib, j0, j1, k0, k1;
};
__device__ __forceinline__ void norm_box(const DdensGeom& G, const Bins& B, const float* __restrict__ hpart,
                                         int64_t i, NormBox& q) {
  q.p = B.pos[i];
  q.h = hpart[i];
  int inc, reach;
  ddens_setup(G, q.p.x, q.p.y, q.p.z, q.h, q.ix, q.iy, q.iz, q.inv_h, q.h2max, inc, reach);
  const int n = G.n;
  q.k0 = max(q.iz - reach, 1), q.k1 = min(q.iz + reach, n);
  q.j0 = max(q.iy - reach, 1), q.j1 = min(q.iy + reach, n);
  q.i0 = max(q.ix - reach, 1), q.i1 = min(q.ix + reach, n);
  const float rdx = 1.0f / G.dx;
  const float ux = (q.p.x - G.rx1) * rdx + 1.0f, uy = (q.p.y - G.rx1) * rdx + 1.0f, uz = (q.p.z - G.rx1) * rdx + 1.0f;
  const float hc = q.h * 1.00001f * rdx + 1e-3f;
  q.k0 = max(q.k0, (int)ceilf(uz - hc));
  q.k1 = min(q.k1, (int)floorf(uz + hc));
  q.j0 = max(q.j0, (int)ceilf(uy - hc));
  q.j1 = min(q.j1, (int)floorf(uy + hc));
  q.ia = max(q.i0, (int)ceilf(ux - hc));
  q.ib = min(q.i1, (int)floorf(ux + hc));
}

// DDENS_INTERP_SPH into the windows, first part: which particles need a
// NORM.  A block takes the particles of 4^3 bins of a window's particle range
// (the bins within the largest DDENS reach of it) and marks need[i] for each
// whose cell box or home cell meets the window.  Windows overlap in their
// ranges, so this pass is cheap where the NORMs are not: k_norm_marked then
// computes each marked particle's NORM once.
__global__ void __launch_bounds__(256) k_win_mark(VirtG V, Bins B, const float* __restrict__ hpart,
                                                 const WinDev* __restrict__ W, const int64_t* __restrict__ off,
                                                 int nwin, uint8_t* need) {
  constexpr int TB = kWinBins, NR = 2 * TB * TB;  // rows of bins, each in at most two runs
  __shared__ int qrow[NR + 1], qbeg[NR];
  const int64_t it = xcd_block(blockIdx.x, gridDim.x);
  const int wi = item_window(off, nwin, it);
  const WinDev wd = W[wi];
  const int64_t t = it - off[wi];
  const int ntx = (wd.bnx + TB - 1) / TB, nty = (wd.bny + TB - 1) / TB;
  const int tx = (int)(t % ntx), ty = (int)((t / ntx) % nty), tz = (int)(t / ((int64_t)ntx * nty));
  const int bxa = wd.bx0 + tx * TB, bxb = wd.bx0 + min(tx * TB + TB, wd.bnx);
  const int bya = wd.by0 + ty * TB, byb = wd.by0 + min(ty * TB + TB, wd.bny);
  const int bza = wd.bz0 + tz * TB, bzb = wd.bz0 + min(tz * TB + TB, wd.bnz);
  const int nb = B.nb, ry = byb - bya, nrows = ry * (bzb - bza);
  const int tid = threadIdx.x;
  if (tid < NR) {
    const int r = tid >> 1, part = tid & 1;
    int beg = 0, cnt = 0;
    if (r < nrows) {
      const int by = wrapn(bya + r % ry, nb), bz = wrapn(bza + r / ry, nb);
      const int s = wrapn(bxa, nb), len = bxb - bxa;
      const int a = part == 0 ? s : 0;
      const int b = part == 0 ? min(s + len, nb) : max(s + len - nb, 0);
      if (b > a) {
        const int64_t rowb = (int64_t)nb * ((int64_t)by + (int64_t)nb * bz);
        beg = B.start[rowb + a];
        cnt = B.start[rowb + b] - beg;
      }
    }
    qbeg[tid] = beg;
    qrow[tid + 1] = cnt;
  }
  __syncthreads();
  if (tid == 0) {
    qrow[0] = 0;
    for (int r = 0; r < NR; ++r) qrow[r + 1] += qrow[r];
  }
  __syncthreads();
  const int n = V.n, per = V.periodic;
  const DdensGeom G{V.rad, V.rlo, n, V.dx, V.rx1};
  const int nq = qrow[NR];
  for (int q = tid; q < nq; q += 256) {
    int r = 0, rh = NR;
    while (rh - r > 1) {
      const int mid = (r + rh) >> 1;
      if (qrow[mid] <= q) r = mid;
      else rh = mid;
    }
    const int64_t i = qbeg[r] + (q - qrow[r]);
    NormBox b;
    norm_box(G, B, hpart, i, b);
    const bool box_in = win_hits(b.ia, b.ib, wd.ox, wd.w, n, per) && win_hits(b.j0, b.j1, wd.oy, wd.w, n, per) &&
                        win_hits(b.k0, b.k1, wd.oz, wd.w, n, per);
    const bool home_in = b.ix >= 1 && b.ix <= n && b.iy >= 1 && b.iy <= n && b.iz >= 1 && b.iz <= n &&
                         win_local(b.ix, wd.ox, wd.w, n, per) >= 0 && win_local(b.iy, wd.oy, wd.w, n, per) >= 0 &&
                         win_local(b.iz, wd.oz, wd.w, n, per) >= 0;
    if (box_in || home_in) need[i] = 1;
  }
}

// NORM of every listed particle (particles.f90:520-547, in the reference's
// loop order, as ddens_tile computes it), one thread each; the list keeps the
// bin order, so a wave's particles are neighbours
__global__ void __launch_bounds__(256) k_norm_marked(VirtG V, Bins B, const float* __restrict__ hpart,
                                                    const int32_t* __restrict__ list, const int32_t* __restrict__ count,
                                                    float* norm, unsigned long long* npairs) {
  const int nl = *count;
  unsigned pairs = 0;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < nl; t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = list[t];
    const DdensGeom G{V.rad, V.rlo, V.n, V.dx, V.rx1};
    NormBox b;
    norm_box(G, B, hpart, i, b);
    const float hm2 = b.h * b.h * 1.00001f;
    const float rdx = 1.0f / G.dx;
    const float ux = (b.p.x - G.rx1) * rdx + 1.0f;
    float nrm = 0.0f;
    for (int k2 = b.k0; k2 <= b.k1; ++k2) {
      const float tz2 = b.p.z - G.rad[k2 - G.rlo];
      for (int j2 = b.j0; j2 <= b.j1; ++j2) {
        const float ty2 = b.p.y - G.rad[j2 - G.rlo];
        const float rem = hm2 - (ty2 * ty2 + tz2 * tz2);
        if (rem < 0.0f) continue;
        const float rc = bound_sqrt(rem) * rdx + 1e-3f;
        const int a = max(b.i0, (int)ceilf(ux - rc));
        const int e = min(b.i1, (int)floorf(ux + rc));
        const float dy2 = ty2 * ty2, dz2 = tz2 * tz2;
        for (int i2 = a; i2 <= e; ++i2) {
          bool ok;
          const float wgt = ddens_w_row(G, b.p.x, dy2, dz2, b.inv_h, b.h2max, i2, ok);
          if (!ok) continue;
          ++pairs;
          nrm = nrm + wgt;
        }
      }
    }
    norm[i] = nrm;
  }
  for (int o = 32; o > 0; o >>= 1) pairs += (unsigned)__shfl_xor((int)pairs, o);
  if ((threadIdx.x & 63) == 0 && pairs) atomicAdd(npairs, (unsigned long long)pairs);
}

// The deposits of DDENS_INTERP_SPH into the windows (particles.f90:560-618),
// tile by tile: a block takes kDepX x kDepY x kDepZ cells of a window, keeps
// their int64 fixed-point sums in LDS, and gathers the particles whose cell
// box (clipped to the cells whose centre lies within h) meets the tile: those
// binned within the level's largest h of the tile's cell centres.  Each wave
// takes one such particle at a time and its lanes go through the particle's
// cells in the tile, adding m W / NORM (NORM from k_norm_marked) with LDS
// atomics; cells past h get M4 = 0 and add nothing.  The tile's sums then go
// to the atlas once per cell.  On grids finer than the particles each
// particle deposits into hundreds of cells, and the sums stay in LDS instead
// of costing one HBM atomic per (particle, cell) pair.  Integer sums: the
// result does not depend on the order.
constexpr int kDepX = 16, kDepY = 16, kDepZ = 8, kDepT = 256, kDepL = 768;

int64_t win_dep_items(int32_t w) {
  return (int64_t)((w + kDepX - 1) / kDepX) * ((w + kDepY - 1) / kDepY) * ((w + kDepZ - 1) / kDepZ);
}

// one axis of a tile: window-local cells [t0, t1) as up to two runs of level
// cells (0-based) g .. g + len - 1, the first at window-local cell l
struct CellRun {
  int g, len, l;
};
__device__ __forceinline__ int tile_runs(int o, int t0, int t1, int n, int per, CellRun* r) {
  if (t1 <= t0) return 0;
  if (!per) {  // the windows of a box that is not periodic lie inside the grid
    r[0] = CellRun{o + t0, t1 - t0, t0};
    return 1;
  }
  const int s = wrapn(o + t0, n), len = t1 - t0;
  if (s + len <= n) {
    r[0] = CellRun{s, len, t0};
    return 1;
  }
  r[0] = CellRun{s, n - s, t0};
  r[1] = CellRun{0, s + len - n, t0 + (n - s)};
  return 2;
}

// particle i's cells (1-based level cells, x a..b, y c..d, z e..f) as the
// deposit loop of ddens_tile bounds them: the reference's box clipped to the
// grid and to the centres within h
struct DepBox {
  float4 p;
  float h, inv_h, h2max, nrm;
  int a, b, c, d, e, f;
  int ix, iy, iz;  // home cell
};
__device__ __forceinline__ void dep_box(const DdensGeom& G, const Bins& B, const float* __restrict__ hpart,
                                        int64_t i, DepBox& q) {
  q.p = B.pos[i];
  const float hh = hpart[i];
  q.h = hh;
  int inc, reach;
  ddens_setup(G, q.p.x, q.p.y, q.p.z, hh, q.ix, q.iy, q.iz, q.inv_h, q.h2max, inc, reach);
  const int ix = q.ix, iy = q.iy, iz = q.iz;
  const int n = G.n;
  const float rdx = 1.0f / G.dx;
  const float ux = (q.p.x - G.rx1) * rdx + 1.0f, uy = (q.p.y - G.rx1) * rdx + 1.0f, uz = (q.p.z - G.rx1) * rdx + 1.0f;
  const float hc = hh * 1.00001f * rdx + 1e-3f;
  q.a = max(max(ix - reach, 1), (int)ceilf(ux - hc));
  q.b = min(min(ix + reach, n), (int)floorf(ux + hc));
  q.c = max(max(iy - reach, 1), (int)ceilf(uy - hc));
  q.d = min(min(iy + reach, n), (int)floorf(uy + hc));
  q.e = max(max(iz - reach, 1), (int)ceilf(uz - hc));
  q.f = min(min(iz + reach, n), (int)floorf(uz + hc));
}

__global__ void __launch_bounds__(kDepT) k_win_dep(VirtG V, Bins B, const float* __restrict__ hpart,
                                                 const float* __restrict__ norm, double scale, float hm,
                                                 int big, unsigned long long* acc, const WinDev* __restrict__ W,
                                                 const int64_t* __restrict__ off, int nwin) {
  __shared__ unsigned long long box[kDepZ][kDepY][kDepX];
  // list: particles with fewer than big cells in the tile from the front (s_ns
  // of them, a thread each), the others from the back (s_nb, a wave each)
  __shared__ int qcnt[kDepT + 1], qbeg[kDepT], list[kDepL], s_ns, s_nb;
  const int64_t it = xcd_block(blockIdx.x, gridDim.x);
  const int wi = item_window(off, nwin, it);
  const WinDev wd = W[wi];
  const int64_t t = it - off[wi];
  const int w = wd.w, ntx = (w + kDepX - 1) / kDepX, nty = (w + kDepY - 1) / kDepY;
  const int tx0 = (int)(t % ntx) * kDepX, ty0 = (int)((t / ntx) % nty) * kDepY,
            tz0 = (int)(t / ((int64_t)ntx * nty)) * kDepZ;
  const int tx1 = min(tx0 + kDepX, w), ty1 = min(ty0 + kDepY, w), tz1 = min(tz0 + kDepZ, w);
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  for (int c = tid; c < kDepX * kDepY * kDepZ; c += kDepT) (&box[0][0][0])[c] = 0ull;
  if (tid == 0) s_ns = s_nb = 0;
  const int n = V.n, per = V.periodic, nb = B.nb;
  const DdensGeom G{V.rad, V.rlo, n, V.dx, V.rx1};
  CellRun rx[2], ry[2], rz[2];
  const int nrx = tile_runs(wd.ox, tx0, tx1, n, per, rx), nry = tile_runs(wd.oy, ty0, ty1, n, per, ry),
            nrz = tile_runs(wd.oz, tz0, tz1, n, per, rz);
  // the bins of particles within hm of a run's cell centres (one bin of margin)
  auto bins_of = [&](const CellRun& R, int& b0, int& b1) {
    const float lo = G.rad[R.g + 1 - G.rlo] - hm, hi = G.rad[R.g + R.len - G.rlo] + hm;
    b0 = max((int)floorf((lo - B.lo) / B.bsz) - 1, 0);
    b1 = min((int)floorf((hi - B.lo) / B.bsz) + 1, nb - 1);
  };
  for (int sz = 0; sz < nrz; ++sz)
    for (int sy = 0; sy < nry; ++sy)
      for (int sx = 0; sx < nrx; ++sx) {
        const CellRun X = rx[sx], Y = ry[sy], Z = rz[sz];
        // the sub-tile in 1-based level cells, and its offset into box[][][]
        const int xa = X.g + 1, xb = X.g + X.len, ya = Y.g + 1, yb = Y.g + Y.len, za = Z.g + 1, zb = Z.g + Z.len;
        const int ox = X.l - tx0 - xa, oy = Y.l - ty0 - ya, oz = Z.l - tz0 - za;
        int bx0, bx1, by0, by1, bz0, bz1;
        bins_of(X, bx0, bx1);
        bins_of(Y, by0, by1);
        bins_of(Z, bz0, bz1);
        const int nyr = by1 - by0 + 1, nrows = nyr * (bz1 - bz0 + 1);
        // the listed particles' deposits
        auto deposit_thread = [&](int nl) {  // a particle per thread, rows clipped to the centres within h
          for (int e = tid; e < nl; e += kDepT) {
            DepBox q;
            dep_box(G, B, hpart, list[e], q);
            const float nrm = norm[list[e]];
            const int a0 = max(q.a, xa), b0 = min(q.b, xb), c0 = max(q.c, ya), d0 = min(q.d, yb);
            const int e0 = max(q.e, za), f0 = min(q.f, zb);
            const float hm2 = q.h * q.h * 1.00001f, rdx = 1.0f / G.dx;
            const float ux = (q.p.x - G.rx1) * rdx + 1.0f;
            for (int k2 = e0; k2 <= f0; ++k2) {
              const float tz2 = q.p.z - G.rad[k2 - G.rlo];
              for (int j2 = c0; j2 <= d0; ++j2) {
                const float ty2 = q.p.y - G.rad[j2 - G.rlo];
                const float rem = hm2 - (ty2 * ty2 + tz2 * tz2);
                if (rem < 0.0f) continue;
                const float rc = bound_sqrt(rem) * rdx + 1e-3f;
                const int a = max(a0, (int)ceilf(ux - rc)), b = min(b0, (int)floorf(ux + rc));
                const float dy2 = ty2 * ty2, dz2 = tz2 * tz2;
                for (int i2 = a; i2 <= b; ++i2) {
                  bool ok;
                  const float wgt = ddens_w_row(G, q.p.x, dy2, dz2, q.inv_h, q.h2max, i2, ok);
                  if (!ok) continue;
                  const float v = (wgt * q.p.w) / nrm;  // particles.f90:599-600
                  if (v != 0.0f)
                    atomicAdd(&box[k2 + oz][j2 + oy][i2 + ox], (unsigned long long)llrint((double)v * scale));
                }
              }
            }
          }
        };
        auto deposit_wave = [&](int nl) {  // a particle per wave, its cells across the lanes
          for (int e = wave; e < nl; e += kDepT / 64) {
            const int64_t i = list[kDepL - 1 - e];
            DepBox q;
            dep_box(G, B, hpart, i, q);
            q.nrm = norm[i];
            const int a = max(q.a, xa), b = min(q.b, xb), c = max(q.c, ya), d = min(q.d, yb);
            const int e0 = max(q.e, za), f = min(q.f, zb);
            const int mx = b - a + 1, my = d - c + 1, cells = mx * my * (f - e0 + 1);
            for (int u = lane; u < cells; u += 64) {
              const int i2 = a + u % mx, j2 = c + (u / mx) % my, k2 = e0 + u / (mx * my);
              bool ok;
              const float wgt = ddens_w(G, q.p.x, q.p.y, q.p.z, q.inv_h, q.h2max, i2, j2, k2, ok);
              if (!ok) continue;
              const float v = (wgt * q.p.w) / q.nrm;  // particles.f90:599-600
              if (v != 0.0f)
                atomicAdd(&box[k2 + oz][j2 + oy][i2 + ox], (unsigned long long)llrint((double)v * scale));
            }
          }
        };
        auto deposit = [&](int ns, int nbig) {
          deposit_thread(ns);
          deposit_wave(nbig);
        };
        for (int r0 = 0; r0 < nrows; r0 += kDepT) {
          __syncthreads();  // qcnt and qbeg of the previous rows are no longer read
          {
            const int r = r0 + tid;
            int beg = 0, cnt = 0;
            if (r < nrows) {
              const int64_t rowb = (int64_t)nb * ((int64_t)(by0 + r % nyr) + (int64_t)nb * (bz0 + r / nyr));
              beg = B.start[rowb + bx0];
              cnt = B.start[rowb + bx1 + 1] - beg;
            }
            qbeg[tid] = beg;
            qcnt[tid + 1] = cnt;
          }
          __syncthreads();
          if (tid == 0) {
            qcnt[0] = 0;
            for (int r = 0; r < kDepT; ++r) qcnt[r + 1] += qcnt[r];
          }
          __syncthreads();
          const int nq = qcnt[kDepT];
          for (int q0 = 0; q0 < nq; q0 += kDepT) {
            const int qi = q0 + tid;
            if (qi < nq) {
              int r = 0, rh = kDepT;
              while (rh - r > 1) {
                const int mid = (r + rh) >> 1;
                if (qcnt[mid] <= qi) r = mid;
                else rh = mid;
              }
              const int64_t i = qbeg[r] + (qi - qcnt[r]);
              DepBox q;
              dep_box(G, B, hpart, i, q);
              const int mx = min(q.b, xb) - max(q.a, xa) + 1, my = min(q.d, yb) - max(q.c, ya) + 1,
                        mz = min(q.f, zb) - max(q.e, za) + 1;
              const bool meets = mx > 0 && my > 0 && mz > 0;
              const bool home = q.ix >= xa && q.ix <= xb && q.iy >= ya && q.iy <= yb && q.iz >= za && q.iz <= zb;
              if (meets || home) {
                if (norm[i] != 0.0f) {
                  if (meets) {
                    if (mx * my * mz < big)
                      list[atomicAdd(&s_ns, 1)] = (int)i;
                    else
                      list[kDepL - 1 - atomicAdd(&s_nb, 1)] = (int)i;
                  }
                } else if (home) {  // particles.f90:549-557: NORM = 0, all the mass to the home cell
                  atomicAdd(&box[q.iz + oz][q.iy + oy][q.ix + ox], (unsigned long long)llrint((double)q.p.w * scale));
                }
              }
            }
            __syncthreads();
            const int ns = s_ns, nbig = s_nb;
            if (ns + nbig > kDepL - kDepT) {  // flush before the list can overflow
              deposit(ns, nbig);
              __syncthreads();
              if (tid == 0) s_ns = s_nb = 0;
            }
            __syncthreads();
          }
        }
        __syncthreads();
        const int ns = s_ns, nbig = s_nb;
        if (ns + nbig > 0) deposit(ns, nbig);
        __syncthreads();
        if (tid == 0) s_ns = s_nb = 0;
      }
  __syncthreads();
  for (int c = tid; c < kDepX * kDepY * kDepZ; c += kDepT) {
    const unsigned long long v = (&box[0][0][0])[c];
    if (v == 0ull) continue;
    const int lx = c % kDepX, ly = (c / kDepX) % kDepY, lz = c / (kDepX * kDepY);
    const size_t ai = (size_t)(wd.ax + tx0 + lx) +
                      (size_t)V.A * ((size_t)(wd.ay + ty0 + ly) + (size_t)V.A * (size_t)(wd.az + tz0 + lz));
    atomicAdd(&acc[ai], v);
  }
}

// MARCAP of the windows: the parent level's MARCA at cell floor((o + l) / 2),
// kept where it is the window's parent
__global__ void __launch_bounds__(256) k_win_marcap(const WinDev* __restrict__ W, const int64_t* __restrict__ off,
                                                   int nwin, int A, const int32_t* __restrict__ pmarca, int pside,
                                                   int32_t* marcap) {
  const int64_t it = blockIdx.x;
  const int wi = item_window(off, nwin, it);
  const WinDev wd = W[wi];
  const int w = wd.w;
  const int64_t l = (it - off[wi]) * 256 + threadIdx.x;
  if (l >= (int64_t)w * w * w) return;
  const int lx = (int)(l % w), ly = (int)((l / w) % w), lz = (int)(l / ((int64_t)w * w));
  int sx = ((wd.ox + lx) >> 1) - wd.sox, sy = ((wd.oy + ly) >> 1) - wd.soy, sz = ((wd.oz + lz) >> 1) - wd.soz;
  if (wd.swrap) {
    sx = wrapn(sx, wd.sw);
    sy = wrapn(sy, wd.sw);
    sz = wrapn(sz, wd.sw);
  }
  int32_t v = 0;
  if (sx >= 0 && sx < wd.sw && sy >= 0 && sy < wd.sw && sz >= 0 && sz < wd.sw) {
    const int32_t m = pmarca[(size_t)(wd.sax + sx) +
                             (size_t)pside * ((size_t)(wd.say + sy) + (size_t)pside * (size_t)(wd.saz + sz))];
    v = m == wd.parent ? m : 0;
  }
  marcap[(size_t)(wd.ax + lx) + (size_t)A * ((size_t)(wd.ay + ly) + (size_t)A * (size_t)(wd.az + lz))] = v;
}

// the largest of n float bits (non-negative floats: the largest value)
__global__ void __launch_bounds__(256) k_max_bits(int64_t n, const uint32_t* __restrict__ a, uint32_t* out) {
  uint32_t m = 0;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) m = max(m, a[i]);
  for (int o = 32; o > 0; o >>= 1) m = max(m, (uint32_t)__shfl_xor((int)m, o));
  if ((threadIdx.x & 63) == 0) atomicMax(out, m);
}

// FP64 sums of the particle masses over nparts contiguous slices (fixed order)
__global__ void __launch_bounds__(256) k_mass_sum(int64_t np, const float* __restrict__ m, double* partial) {
  __shared__ double red[256];
  const int64_t per = (np + gridDim.x - 1) / gridDim.x;
  const int64_t b0 = (int64_t)blockIdx.x * per, b1 = min(b0 + per, np);
  double s = 0.0;
  for (int64_t i = b0 + threadIdx.x; i < b1; i += 256) s += (double)m[i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

hipError_t launch_win_vvel(hipStream_t st, const Bins& B, int64_t np, const WinDev* W, const int64_t* off,
                           int32_t nwin, int64_t items, const VirtG& V, int32_t k, float r0, const float4* vel,
                           uint32_t* hbits, float* u, float* v, float* w) {
  if (items <= 0) return hipSuccess;
  const VvelOut O{vel, hbits, u, v, w, vel ? 1 : 0};
  const bool chk = vvel_chk((int64_t)V.n * V.n * V.n, np);
  with_k(k, [&](auto KC) {
    constexpr int K = decltype(KC)::value;
    if (chk)
      k_win_vvel<K, true><<<(unsigned)items, kq_of(K), 0, st>>>(B, np, W, off, nwin, V, k, r0, O);
    else
      k_win_vvel<K, false><<<(unsigned)items, kq_of(K), 0, st>>>(B, np, W, off, nwin, V, k, r0, O);
  });
  return hipGetLastError();
}

hipError_t launch_win_ddens(hipStream_t st, const VirtG& V, const Bins& B, int64_t np, const float* hpart,
                            double scale, unsigned long long* acc, const WinDev* W, const int64_t* off, int32_t nwin,
                            int64_t items, unsigned long long* npairs, float* norm, uint8_t* need,
                            int32_t* list, int32_t* count, DBuf& sel_tmp, float hmax, const int64_t* dep_off,
                            int64_t dep_items) {
  if (items <= 0) return hipSuccess;
  hipError_t e = hipMemsetAsync(need, 0, (size_t)np, st);
  if (e != hipSuccess) return e;
  k_win_mark<<<(unsigned)items, 256, 0, st>>>(V, B, hpart, W, off, nwin, need);
  // the marked particles' positions, in order (rocPRIM's flagged select)
  size_t tb = 0;
  e = rocprim::select(nullptr, tb, rocprim::counting_iterator<int32_t>(0), need, list, count, (size_t)np, st);
  if (e != hipSuccess) return e;
  if (!sel_tmp.ensure(tb)) return hipErrorOutOfMemory;
  e = rocprim::select(sel_tmp.p, tb, rocprim::counting_iterator<int32_t>(0), need, list, count, (size_t)np, st);
  if (e != hipSuccess) return e;
  k_norm_marked<<<8192, 256, 0, st>>>(V, B, hpart, list, count, norm, npairs);
  e = hipGetLastError();
  if (e != hipSuccess || dep_items <= 0) return e;
  const float hm = hmax * 1.00001f + V.dx;
  // a wave per particle pays off where particles reach thousands of cells
  const int big = (double)V.n * V.n * V.n >= 32.0 * (double)np ? 256 : INT_MAX;
  k_win_dep<<<(unsigned)dep_items, kDepT, 0, st>>>(V, B, hpart, norm, scale, hm, big, acc, W, dep_off, nwin);
  return hipGetLastError();
}

hipError_t launch_win_marcap(hipStream_t st, const WinDev* W, const int64_t* off, int32_t nwin, int64_t items,
                             int32_t A, const int32_t* pmarca, int32_t pside, int32_t* marcap) {
  if (items <= 0) return hipSuccess;
  k_win_marcap<<<(unsigned)items, 256, 0, st>>>(W, off, nwin, A, pmarca, pside, marcap);
  return hipGetLastError();
}

hipError_t launch_max_bits(hipStream_t st, int64_t n, const uint32_t* a, uint32_t* out) {
  k_max_bits<<<(unsigned)std::max<int64_t>(1, std::min<int64_t>(nblk(n), 4096)), 256, 0, st>>>(n, a, out);
  return hipGetLastError();
}

hipError_t launch_mass_sum(hipStream_t st, int64_t np, const float* m, double* partial, int32_t nparts) {
  k_mass_sum<<<(unsigned)nparts, 256, 0, st>>>(np, m, partial);
  return hipGetLastError();
}

hipError_t launch_ddens_finalize(hipStream_t st, int64_t ncell, const unsigned long long* acc, double inv_scale,
                                 float den, float rote, float* umass, double* partial) {
  const unsigned grid = (unsigned)((ncell + (int64_t)kFinB * kFinV - 1) / ((int64_t)kFinB * kFinV));
  k_ddens_finalize<<<grid, kFinB, 0, st>>>(ncell, acc, inv_scale, 0, den, rote, umass, partial);
  return hipGetLastError();
}

}  // namespace avism
