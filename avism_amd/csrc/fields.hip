// fields.hip -- particle -> grid stage on gfx950.
//
// Reference: particles.f90 (PPART_DENS 742-802, VVEL_INTERP_SPH_VW 806-930,
// KERNEL_FUNC 935-963, FILL_H_ZEROS 398-431, H_DISTANCE 276-351,
// DDENS_INTERP_SPH 434-618), divergence.f90:2-50 (DIVER_UNIFORM) and the mean
// density block avism.f90:1249-1250, 1371-1414.
//
// Design (MI355X):
//  * particles are binned into a uniform grid of cells (stable radix sort by
//    bin, then SoA float4 copies), so every kNN query walks Chebyshev shells of
//    bins nearest-first with bounding-box pruning; the k best (squared FP32
//    distance, particle index) pairs live in registers (one query per lane).
//  * distances, kernel weights and the DDENS normalisation are evaluated in the
//    reference's FP32 operation order (compiled without FMA contraction), so
//    neighbour sets, HPART and NORM are bit-identical to the reference; the
//    accumulations are FP32 with Kahan compensation instead of FP64.
//  * DDENS is a gather: each cell sums the contributions of the particles
//    whose support box covers it, in a fixed order -- deterministic, no float
//    atomics.  HPART uses integer atomicMax on the float bits (order-free).
//  * only the mean-density reduction runs in FP64.
#include <cstring>
#include <algorithm>
#include <cmath>

#include <rocprim/rocprim.hpp>

#include "avism_internal.h"

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
  float ms = 0.0f;
  (void)hipEventSynchronize(b);
  (void)hipEventElapsedTime(&ms, a, b);
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  c->timings.push_back({name, ms});
}

// ------------------------------------------------------------ device helpers
__device__ __forceinline__ float axis_d(float p, float q, float L, int periodic) {
  float d = p - q;
  if (periodic) {
    float a = fabsf(d), b = L - fabsf(d);
    d = a < b ? a : b;
  }
  return d;
}

// kdtree.f90:1441-1465 (squared distance, x, y, z accumulated in order)
__device__ __forceinline__ float dist2(float px, float py, float pz, float qx, float qy, float qz,
                                       float L, int periodic) {
  float dx = axis_d(px, qx, L, periodic);
  float dy = axis_d(py, qy, L, periodic);
  float dz = axis_d(pz, qz, L, periodic);
  float d = dx * dx;
  d = d + dy * dy;
  d = d + dz * dz;
  return d;
}

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

struct Bins {
  const float4* __restrict__ pos;      // sorted (x,y,z,m)
  const uint32_t* __restrict__ orig;   // sorted -> input index
  const int32_t* __restrict__ start;   // nb^3 + 1
  int32_t nb;
  float lo, bsz, L, eps;
  int periodic;
  __device__ __forceinline__ int bin_of(float v) const {
    int b = (int)floorf((v - lo) / bsz);
    return b < 0 ? 0 : (b >= nb ? nb - 1 : b);
  }
  __device__ __forceinline__ int wrap(int b) const {
    if (b < 0) b += nb;
    else if (b >= nb) b -= nb;
    return b;
  }
  // squared distance from q to the (unwrapped) bin box at index b, axis-wise
  __device__ __forceinline__ float box_d(float q, int b) const {
    float blo = lo + (float)b * bsz - eps, bhi = lo + (float)(b + 1) * bsz + eps;
    float d = 0.0f;
    if (q < blo) d = blo - q;
    else if (q > bhi) d = q - bhi;
    return d;
  }
  // lower bound of the distance from q (in bin b) to anything outside shells 0..s
  __device__ __forceinline__ float outside(float q, int b, int s) const {
    float dlo = q - (lo + (float)(b - s) * bsz) - eps;
    float dhi = (lo + (float)(b + s + 1) * bsz) - q - eps;
    if (!periodic) {
      if (b - s <= 0) dlo = INFINITY;
      if (b + s + 1 >= nb) dhi = INFINITY;
    }
    return dlo < dhi ? dlo : dhi;
  }
};

// k best (squared distance, input index) pairs in registers; K >= runtime k
template <int K>
struct TopK {
  float d[K];
  uint32_t id[K];   // sorted position
  int cnt;
  float wd;         // worst distance
  uint32_t wid;     // its sorted position
  int ws;           // its slot

  __device__ __forceinline__ void init() {
    cnt = 0;
    wd = INFINITY;
    wid = 0;
    ws = 0;
#pragma clang loop unroll(full)
    for (int t = 0; t < K; ++t) { d[t] = INFINITY; id[t] = 0; }
  }
  __device__ __forceinline__ void recompute(int k, const uint32_t* orig) {
    wd = -1.0f;
    ws = 0;
    wid = 0;
#pragma clang loop unroll(full)
    for (int t = 0; t < K; ++t) {
      if (t < k) {
        bool take = d[t] > wd;
        if (!take && d[t] == wd) take = orig[id[t]] > orig[wid];
        wd = take ? d[t] : wd;
        ws = take ? t : ws;
        wid = take ? id[t] : wid;
      }
    }
  }
  __device__ __forceinline__ void consider(float v, uint32_t j, int k, const uint32_t* orig) {
    if (cnt < k) {
#pragma clang loop unroll(full)
      for (int t = 0; t < K; ++t) {
        // value selects (not conditional stores) keep the arrays in VGPRs
        d[t] = (t == cnt) ? v : d[t];
        id[t] = (t == cnt) ? j : id[t];
      }
      ++cnt;
      if (cnt == k) recompute(k, orig);
      return;
    }
    if (v > wd) return;
    if (v == wd && orig[j] >= orig[wid]) return;
#pragma clang loop unroll(full)
    for (int t = 0; t < K; ++t) {
      d[t] = (t == ws) ? v : d[t];
      id[t] = (t == ws) ? j : id[t];
    }
    recompute(k, orig);
  }
};

// exact kNN of q into tk (k neighbours, or fewer if np < k)
template <int K>
__device__ __forceinline__ void knn_search(const Bins& B, float qx, float qy, float qz, int k, TopK<K>& tk) {
  tk.init();
  const int bx = B.bin_of(qx), by = B.bin_of(qy), bz = B.bin_of(qz);
  const int nb = B.nb;
  const int smax_p = (nb - 1) / 2;
  for (int s = 0;; ++s) {
    bool full_scan = B.periodic && s > smax_p;
    if (full_scan) break;  // handled below
    bool any = false;
    for (int dz = -s; dz <= s; ++dz) {
      int cz = bz + dz;
      if (!B.periodic && (cz < 0 || cz >= nb)) continue;
      float ddz = B.box_d(qz, cz);
      int wz = B.periodic ? B.wrap(cz) : cz;
      for (int dy = -s; dy <= s; ++dy) {
        int cy = by + dy;
        if (!B.periodic && (cy < 0 || cy >= nb)) continue;
        float ddy = B.box_d(qy, cy);
        int wy = B.periodic ? B.wrap(cy) : cy;
        bool edge = (dz == -s || dz == s || dy == -s || dy == s);
        int step = (edge || s == 0) ? 1 : 2 * s;
        for (int dxo = -s; dxo <= s; dxo += step) {
          int cx = bx + dxo;
          if (!B.periodic && (cx < 0 || cx >= nb)) continue;
          any = true;
          float ddx = B.box_d(qx, cx);
          float mind = ddx * ddx + ddy * ddy + ddz * ddz;
          if (tk.cnt >= k && mind > tk.wd) continue;
          int wx = B.periodic ? B.wrap(cx) : cx;
          int bin = wx + nb * (wy + nb * wz);
          int e = B.start[bin + 1];
          for (int j = B.start[bin]; j < e; ++j) {
            float4 p = B.pos[j];
            float d = dist2(p.x, p.y, p.z, qx, qy, qz, B.L, B.periodic);
            tk.consider(d, (uint32_t)j, k, B.orig);
          }
          if (s == 0) break;
        }
      }
    }
    if (tk.cnt >= k) {
      float b1 = B.outside(qx, bx, s), b2 = B.outside(qy, by, s), b3 = B.outside(qz, bz, s);
      float bd = fminf(b1, fminf(b2, b3));
      if (bd == INFINITY || (bd > 0.0f && tk.wd < bd * bd)) return;
    }
    if (!B.periodic && !any && s > 0) {
      int far = max(max(max(bx, nb - 1 - bx), max(by, nb - 1 - by)), max(bz, nb - 1 - bz));
      if (s > far) return;
    }
    if (B.periodic && s + 1 > smax_p) break;
  }
  // periodic box smaller than the search: scan every bin once
  tk.init();
  int nbins = nb * nb * nb;
  int e = B.start[nbins];
  for (int j = 0; j < e; ++j) {
    float4 p = B.pos[j];
    float d = dist2(p.x, p.y, p.z, qx, qy, qz, B.L, B.periodic);
    tk.consider(d, (uint32_t)j, k, B.orig);
  }
}

// visit every particle with d2 <= r2 (ball_search, kdtree.f90:1015-1072, 1126)
template <class F>
__device__ __forceinline__ void ball_visit(const Bins& B, float qx, float qy, float qz, float r, float r2, F&& f) {
  const int bx = B.bin_of(qx), by = B.bin_of(qy), bz = B.bin_of(qz);
  const int nb = B.nb;
  int s = (int)ceilf((r + B.eps) / B.bsz) + 1;
  if (B.periodic && 2 * s + 1 > nb) {
    int e = B.start[nb * nb * nb];
    for (int j = 0; j < e; ++j) {
      float4 p = B.pos[j];
      float d = dist2(p.x, p.y, p.z, qx, qy, qz, B.L, B.periodic);
      if (d <= r2) f(j, d);
    }
    return;
  }
  for (int dz = -s; dz <= s; ++dz) {
    int cz = bz + dz;
    if (!B.periodic && (cz < 0 || cz >= nb)) continue;
    float ddz = B.box_d(qz, cz);
    if (ddz > r) continue;
    int wz = B.periodic ? B.wrap(cz) : cz;
    for (int dy = -s; dy <= s; ++dy) {
      int cy = by + dy;
      if (!B.periodic && (cy < 0 || cy >= nb)) continue;
      float ddy = B.box_d(qy, cy);
      if (ddy > r) continue;
      int wy = B.periodic ? B.wrap(cy) : cy;
      for (int dxo = -s; dxo <= s; ++dxo) {
        int cx = bx + dxo;
        if (!B.periodic && (cx < 0 || cx >= nb)) continue;
        float ddx = B.box_d(qx, cx);
        if (ddx * ddx + ddy * ddy + ddz * ddz > r2 * 1.0001f + 1e-6f) continue;
        int wx = B.periodic ? B.wrap(cx) : cx;
        int bin = wx + nb * (wy + nb * wz);
        int e = B.start[bin + 1];
        for (int j = B.start[bin]; j < e; ++j) {
          float4 p = B.pos[j];
          float d = dist2(p.x, p.y, p.z, qx, qy, qz, B.L, B.periodic);
          if (d <= r2) f(j, d);
        }
      }
    }
  }
}

// ----------------------------------------------------------------- kernels
__global__ void k_convert_check(int64_t np, float* vx, float* vy, float* vz, float* m,
                                const float* x, const float* y, const float* z, float half,
                                int* bad) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  m[i] = m[i] / kUM;  // reader.f90:457-461
  vx[i] = vx[i] / kUV;
  vy[i] = vy[i] / kUV;
  vz[i] = vz[i] / kUV;
  float a = x[i], b = y[i], c = z[i];
  if (a < -half || a > half || b < -half || b > half || c < -half || c > half) atomicOr(bad, 1);
}

__global__ void k_bin_keys(int64_t np, const float* x, const float* y, const float* z, Bins B,
                           uint32_t* key, uint32_t* val, int32_t* count) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  int b = B.bin_of(x[i]) + B.nb * (B.bin_of(y[i]) + B.nb * B.bin_of(z[i]));
  key[i] = (uint32_t)b;
  val[i] = (uint32_t)i;
  atomicAdd(&count[b], 1);
}

__global__ void k_gather_sorted(int64_t np, const uint32_t* perm, const float* x, const float* y,
                                const float* z, const float* m, const float* vx, const float* vy,
                                const float* vz, float4* pos, float4* vel, uint32_t* orig) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  uint32_t p = perm[i];
  pos[i] = make_float4(x[p], y[p], z[p], m[p]);
  vel[i] = make_float4(vx[p], vy[p], vz[p], 0.0f);
  orig[i] = p;
}

// PPART_DENS (particles.f90:742-802): one query per particle, in bin order.
// hk = distance to the k-th neighbour, reused by FILL_H_ZEROS (same query).
template <int K>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1, 4))) k_ppart(Bins B, int64_t np, int k, int want_dens,
                                               float* pdens, float* hk) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  float4 q = B.pos[i];
  TopK<K> tk;
  knn_search<K>(B, q.x, q.y, q.z, k, tk);
  float h = sqrtf(tk.wd);
  hk[i] = h;
  if (!want_dens) return;
  Kahan ms;
#pragma clang loop unroll(full)
  for (int t = 0; t < K; ++t)
    if (t < tk.cnt) ms.add(B.pos[tk.id[t]].w);
  float vol = 4.18879020478639098f * ((h * h) * h);  // (4/3) pi h^3
  pdens[i] = ms.s / vol;
}

// VVEL_INTERP_SPH_VW (particles.f90:806-930) and H_DISTANCE (276-351).
// One cell per lane; a block covers an 8x8x4 tile of cells.
template <int K>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1, 4))) k_vvel(Bins B, int n, const float* __restrict__ rad,
                                              int rlo, float dx, int k, int want_vel,
                                              const float4* __restrict__ vel,
                                              const float* __restrict__ pdens,
                                              const int8_t* __restrict__ smask, uint32_t* hbits,
                                              float* u, float* v, float* w) {
  int ix = blockIdx.x * 8 + threadIdx.x % 8;
  int iy = blockIdx.y * 8 + (threadIdx.x / 8) % 8;
  int iz = blockIdx.z * 4 + threadIdx.x / 64;
  if (ix >= n || iy >= n || iz >= n) return;
  size_t c = (size_t)ix + (size_t)n * ((size_t)iy + (size_t)n * iz);
  if (smask && smask[c] == 0) return;
  float qx = rad[ix + 1 - rlo], qy = rad[iy + 1 - rlo], qz = rad[iz + 1 - rlo];
  TopK<K> tk;
  knn_search<K>(B, qx, qy, qz, k, tk);
  float dk = sqrtf(tk.wd);
  if (dk > dx) {  // k neighbours (particles.f90:866-867)
    float h = dk;
    Kahan sw, sx, sy, sz;
#pragma clang loop unroll(full)
    for (int t = 0; t < K; ++t) {
      if (t < tk.cnt) {
        uint32_t j = tk.id[t];
        float d = sqrtf(tk.d[t]);
        atomicMax(&hbits[j], __float_as_uint(d));  // particles.f90:882-885
        if (want_vel) {
          float q = d / (h / 2.0f);
          float wgt = m4(q) * (B.pos[j].w / pdens[j]);
          float4 vv = vel[j];
          sw.add(wgt);
          sx.add(wgt * vv.x);
          sy.add(wgt * vv.y);
          sz.add(wgt * vv.z);
        }
      }
    }
    if (want_vel) {
      u[c] = sx.s / sw.s;
      v[c] = sy.s / sw.s;
      w[c] = sz.s / sw.s;
    }
    return;
  }
  // ball of radius DX (particles.f90:868-878): every particle with d <= DX
  float r2 = dx * dx;
  float maxd2 = 0.0f;
  ball_visit(B, qx, qy, qz, dx, r2, [&](int j, float d2) { maxd2 = fmaxf(maxd2, d2); });
  float h = sqrtf(maxd2);
  Kahan sw, sx, sy, sz;
  ball_visit(B, qx, qy, qz, dx, r2, [&](int j, float d2) {
    float d = sqrtf(d2);
    atomicMax(&hbits[j], __float_as_uint(d));
    if (want_vel) {
      float q = d / (h / 2.0f);
      float wgt = m4(q) * (B.pos[j].w / pdens[j]);
      float4 vv = vel[j];
      sw.add(wgt);
      sx.add(wgt * vv.x);
      sy.add(wgt * vv.y);
      sz.add(wgt * vv.z);
    }
  });
  if (want_vel) {
    u[c] = sx.s / sw.s;
    v[c] = sy.s / sw.s;
    w[c] = sz.s / sw.s;
  }
}

// FILL_H_ZEROS (particles.f90:398-431): h = distance to the k-th neighbour
__global__ void k_fill_h(int64_t np, float* hpart, const float* hk) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  if (!(hpart[i] > 0.0f)) hpart[i] = hk[i];
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

__global__ void k_ddens_norm(int64_t np, DdensGeom G, const float4* __restrict__ pos,
                             const float* __restrict__ hpart, float* norm, int32_t* bin_reach,
                             Bins B, int32_t* rmax) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  float4 p = pos[i];
  int ix, iy, iz, inc, reach;
  float inv_h, h2max;
  ddens_setup(G, p.x, p.y, p.z, hpart[i], ix, iy, iz, inv_h, h2max, inc, reach);
  float nrm = 0.0f;
  for (int k2 = iz - reach; k2 <= iz + reach; ++k2) {
    if (k2 < 1 || k2 > G.n) continue;
    for (int j2 = iy - reach; j2 <= iy + reach; ++j2) {
      if (j2 < 1 || j2 > G.n) continue;
      for (int i2 = ix - reach; i2 <= ix + reach; ++i2) {
        if (i2 < 1 || i2 > G.n) continue;
        bool ok;
        float wgt = ddens_w(G, p.x, p.y, p.z, inv_h, h2max, i2, j2, k2, ok);
        if (ok) nrm = nrm + wgt;
      }
    }
  }
  norm[i] = nrm;
  // reach in bins (home cell and bin differ by at most one cell)
  int rb = reach + 1;
  int b = B.bin_of(p.x) + B.nb * (B.bin_of(p.y) + B.nb * B.bin_of(p.z));
  atomicMax(&bin_reach[b], rb);
  atomicMax(rmax, rb);
}

// DDENS pass 2: gather per cell.  Writes U1 = mass/(dx^3 rete^3)/rote*rote
// (avism.f90:1249-1250, 1386) and the FP64 block sums of it (avism.f90:1395-1405).
__global__ void __launch_bounds__(256) k_ddens_gather(DdensGeom G, Bins B,
                                                      const float* __restrict__ hpart,
                                                      const float* __restrict__ norm,
                                                      const int32_t* __restrict__ bin_reach,
                                                      int rsearch, float den, float rote,
                                                      float* umass, double* partial) {
  int ix = blockIdx.x * 8 + threadIdx.x % 8;
  int iy = blockIdx.y * 8 + (threadIdx.x / 8) % 8;
  int iz = blockIdx.z * 4 + threadIdx.x / 64;
  double contrib = 0.0;
  if (ix < G.n && iy < G.n && iz < G.n) {
    int ci = ix + 1, cj = iy + 1, ck = iz + 1;  // reference indices
    float qx = G.rad[ci - G.rlo], qy = G.rad[cj - G.rlo], qz = G.rad[ck - G.rlo];
    int bx = B.bin_of(qx), by = B.bin_of(qy), bz = B.bin_of(qz);
    Kahan acc;
    const int nb = B.nb;
    for (int dz = -rsearch; dz <= rsearch; ++dz) {
      int cz = bz + dz;
      if (cz < 0 || cz >= nb) continue;  // DDENS never wraps (particles.f90:505-512)
      for (int dy = -rsearch; dy <= rsearch; ++dy) {
        int cy = by + dy;
        if (cy < 0 || cy >= nb) continue;
        for (int dxo = -rsearch; dxo <= rsearch; ++dxo) {
          int cx = bx + dxo;
          if (cx < 0 || cx >= nb) continue;
          int bin = cx + nb * (cy + nb * cz);
          int rb = bin_reach[bin];
          int ad = max(abs(dxo), max(abs(dy), abs(dz)));
          if (ad > rb) continue;
          int e = B.start[bin + 1];
          for (int j = B.start[bin]; j < e; ++j) {
            float4 p = B.pos[j];
            int hx, hy, hz, inc, reach;
            float inv_h, h2max;
            ddens_setup(G, p.x, p.y, p.z, hpart[j], hx, hy, hz, inv_h, h2max, inc, reach);
            int ax = abs(ci - hx), ay = abs(cj - hy), az = abs(ck - hz);
            float nrm = norm[j];
            if (nrm == 0.0f) {  // particles.f90:549-557
              if (ax == 0 && ay == 0 && az == 0) acc.add(p.w);
              continue;
            }
            if (ax > reach || ay > reach || az > reach) continue;
            bool ok;
            float wgt = ddens_w(G, p.x, p.y, p.z, inv_h, h2max, ci, cj, ck, ok);
            if (!ok) continue;
            acc.add((wgt * p.w) / nrm);  // particles.f90:599-600
          }
        }
      }
    }
    float u = acc.s / den;
    u = u / rote;
    u = u * rote;
    size_t c = (size_t)ix + (size_t)G.n * ((size_t)iy + (size_t)G.n * iz);
    umass[c] = u;
    contrib = (double)u;
  }
  // block FP64 sum (fixed reduction order -> deterministic)
  __shared__ double red[256];
  red[threadIdx.x] = contrib;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    size_t b = blockIdx.x + (size_t)gridDim.x * (blockIdx.y + (size_t)gridDim.y * blockIdx.z);
    partial[b] = red[0];
  }
}

__global__ void k_reduce_mean(const double* partial, int64_t nparts, int64_t ncell, double* sum_out,
                              float* mean_out) {
  __shared__ double red[256];
  double s = 0.0;
  for (int64_t i = threadIdx.x; i < nparts; i += 256) s += partial[i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    *sum_out = red[0];
    *mean_out = (float)(red[0] / (double)ncell);  // avism.f90:1405
  }
}

// DIVER_UNIFORM (divergence.f90:2-50) fused with U1CO = U1 / MEANDENS
__global__ void k_div_finalize(int n, float dx, int periodic, const float* __restrict__ u,
                               const float* __restrict__ v, const float* __restrict__ w,
                               const float* __restrict__ umass, const float* __restrict__ mean,
                               float* u1co, float* div) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;  // 0-based
  int j = blockIdx.y, k = blockIdx.z;
  if (i >= n) return;
  auto A = [&](const float* f, int a, int b, int cc) -> float {
    return f[(size_t)a + (size_t)n * ((size_t)b + (size_t)n * cc)];
  };
  size_t c = (size_t)i + (size_t)n * ((size_t)j + (size_t)n * k);
  float fx, fy, fz;
  const int N1 = n - 1;
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
  if (k >= 1 && k <= n - 2) fz = (A(w, i, j, k + 1) - A(w, i, j, k - 1)) / 2.0f;
  else if (periodic) fz = (k == 0) ? (A(w, i, j, 1) - A(w, i, j, N1)) / 2.0f
                                   : (A(w, i, j, 0) - A(w, i, j, n - 2)) / 2.0f;
  else if (k == 0) fz = 2.0f * (A(w, i, j, 1) - A(w, i, j, 0)) - (A(w, i, j, 2) - A(w, i, j, 0)) / 2.0f;
  else fz = 2.0f * (A(w, i, j, N1) - A(w, i, j, n - 2)) - (A(w, i, j, N1) - A(w, i, j, n - 3)) / 2.0f;
  div[c] = (fx / dx + fy / dx) + fz / dx;
  if (umass) u1co[c] = umass[c] / (*mean);
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

// ---------------------------------------------------------------- host side
static inline unsigned nblk(int64_t n, int b = 256) { return (unsigned)((n + b - 1) / b); }

static Bins make_bins(avism_ctx* c) {
  Bins B;
  B.pos = c->spos4.as<float4>();
  B.orig = c->sorig.as<uint32_t>();
  B.start = c->bin_start.as<int32_t>();
  B.nb = c->nb;
  B.lo = -c->g.L0 / 2.0f;
  B.bsz = c->bsz;
  B.L = c->g.L0;
  B.eps = c->g.L0 * 2e-6f + c->bsz * 1e-5f;
  B.periodic = c->g.periodic;
  return B;
}

int load_particles(avism_ctx* c, int64_t np, const float* x, const float* y, const float* z,
                   const float* vx, const float* vy, const float* vz, const float* m, int where) {
  if (!c->grid_set) AV_FAIL(c, AVISM_ERR_STATE, "avism_set_grid must be called first");
  if (np <= 0) AV_FAIL(c, AVISM_ERR_ARG, "np must be positive");
  if (np >= (int64_t)0x7fffffff) AV_FAIL(c, AVISM_ERR_LIMIT, "more than 2^31-1 particles per call");
  size_t b = (size_t)np * 4;
  DBuf* bufs[7] = {&c->px, &c->py, &c->pz, &c->pvx, &c->pvy, &c->pvz, &c->pm};
  const float* src[7] = {x, y, z, vx, vy, vz, m};
  hipMemcpyKind kind = where == AVISM_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
  for (int i = 0; i < 7; i++) {
    if (!bufs[i]->ensure(b)) AV_FAIL(c, AVISM_ERR_HIP, "out of device memory (particles)");
    AV_CHECK_HIP(c, hipMemcpyAsync(bufs[i]->p, src[i], b, kind, c->stream));
  }
  c->np = np;
  c->fields_ready = false;
  int* bad = (int*)c->scalars.ensure(4096);
  AV_CHECK_HIP(c, hipMemsetAsync(bad, 0, 4, c->stream));
  k_convert_check<<<nblk(np), 256, 0, c->stream>>>(np, c->pvx.as<float>(), c->pvy.as<float>(),
                                                   c->pvz.as<float>(), c->pm.as<float>(),
                                                   c->px.as<float>(), c->py.as<float>(),
                                                   c->pz.as<float>(), c->g.L0 / 2.0f, bad);
  int hbad = 0;
  AV_CHECK_HIP(c, hipMemcpyAsync(&hbad, bad, 4, hipMemcpyDeviceToHost, c->stream));
  AV_CHECK_HIP(c, hipStreamSynchronize(c->stream));
  if (hbad) AV_FAIL(c, AVISM_ERR_INPUT, "Particles outside the bounding box! Check L0! Particles must be inside [-L0/2, L0/2] in each direction.");
  return AVISM_OK;
}

static int build_bins(avism_ctx* c) {
  StageTimer t(c, "bin");
  const int64_t np = c->np;
  int nb = c->g.n;  // one bin per grid cell
  if (nb < 1) nb = 1;
  c->nb = nb;
  c->bsz = c->g.L0 / (float)nb;
  int64_t nbins = (int64_t)nb * nb * nb;
  if (!c->bin_start.ensure((size_t)(nbins + 1) * 4) || !c->sbin_key.ensure((size_t)np * 4) ||
      !c->sbin_val.ensure((size_t)np * 4) || !c->sbin_key2.ensure((size_t)np * 4) ||
      !c->sbin_val2.ensure((size_t)np * 4) || !c->spos4.ensure((size_t)np * 16) ||
      !c->svel4.ensure((size_t)np * 16) || !c->sorig.ensure((size_t)np * 4) ||
      !c->bin_reach.ensure((size_t)nbins * 4))
    AV_FAIL(c, AVISM_ERR_HIP, "out of device memory (bins)");
  int32_t* count = c->bin_reach.as<int32_t>();  // reused as the histogram
  AV_CHECK_HIP(c, hipMemsetAsync(count, 0, (size_t)nbins * 4, c->stream));
  Bins B = make_bins(c);
  k_bin_keys<<<nblk(np), 256, 0, c->stream>>>(np, c->px.as<float>(), c->py.as<float>(),
                                              c->pz.as<float>(), B, c->sbin_key.as<uint32_t>(),
                                              c->sbin_val.as<uint32_t>(), count);
  unsigned bits = 1;
  while ((1ll << bits) < nbins) bits++;
  size_t tmp = 0;
  AV_CHECK_HIP(c, rocprim::radix_sort_pairs(nullptr, tmp, c->sbin_key.as<uint32_t>(),
                                            c->sbin_key2.as<uint32_t>(), c->sbin_val.as<uint32_t>(),
                                            c->sbin_val2.as<uint32_t>(), (size_t)np, 0, bits,
                                            c->stream));
  if (!c->sort_tmp.ensure(tmp)) AV_FAIL(c, AVISM_ERR_HIP, "out of device memory (sort)");
  AV_CHECK_HIP(c, rocprim::radix_sort_pairs(c->sort_tmp.p, tmp, c->sbin_key.as<uint32_t>(),
                                            c->sbin_key2.as<uint32_t>(), c->sbin_val.as<uint32_t>(),
                                            c->sbin_val2.as<uint32_t>(), (size_t)np, 0, bits,
                                            c->stream));
  size_t tmp2 = 0;
  AV_CHECK_HIP(c, rocprim::exclusive_scan(nullptr, tmp2, count, c->bin_start.as<int32_t>(), 0,
                                          (size_t)nbins + 1, rocprim::plus<int32_t>(), c->stream));
  if (!c->sort_tmp.ensure(std::max(tmp, tmp2))) AV_FAIL(c, AVISM_ERR_HIP, "out of device memory");
  // the scan reads nbins+1 inputs: make the extra one a zero
  AV_CHECK_HIP(c, hipMemsetAsync(c->sbin_key.p, 0, 4, c->stream));
  // exclusive scan of counts over nbins entries, final total appended
  AV_CHECK_HIP(c, rocprim::exclusive_scan(c->sort_tmp.p, tmp2, count, c->bin_start.as<int32_t>(), 0,
                                          (size_t)nbins, rocprim::plus<int32_t>(), c->stream));
  int32_t total = (int32_t)np;
  AV_CHECK_HIP(c, hipMemcpyAsync(c->bin_start.as<int32_t>() + nbins, &total, 4,
                                 hipMemcpyHostToDevice, c->stream));
  k_gather_sorted<<<nblk(np), 256, 0, c->stream>>>(
      np, c->sbin_val2.as<uint32_t>(), c->px.as<float>(), c->py.as<float>(), c->pz.as<float>(),
      c->pm.as<float>(), c->pvx.as<float>(), c->pvy.as<float>(), c->pvz.as<float>(),
      c->spos4.as<float4>(), c->svel4.as<float4>(), c->sorig.as<uint32_t>());
  AV_CHECK_HIP(c, hipGetLastError());
  AV_CHECK_HIP(c, hipStreamSynchronize(c->stream));
  return AVISM_OK;
}

template <int K>
static void launch_ppart(avism_ctx* c, Bins B, int k, int want_dens) {
  k_ppart<K><<<nblk(c->np), 256, 0, c->stream>>>(B, c->np, k, want_dens, c->pdens.as<float>(),
                                                 c->hk.as<float>());
}

template <int K>
static void launch_vvel(avism_ctx* c, Bins B, const float* rad, int k, int want_vel) {
  int n = c->g.n;
  dim3 grid((n + 7) / 8, (n + 7) / 8, (n + 3) / 4);
  k_vvel<K><<<grid, 256, 0, c->stream>>>(B, n, rad, c->g.rlo, c->g.dx, k, want_vel,
                                         c->svel4.as<float4>(), c->pdens.as<float>(), nullptr,
                                         c->hpart.as<uint32_t>(), c->fu.as<float>(),
                                         c->fv.as<float>(), c->fw.as<float>());
}

int fields_from_particles(avism_ctx* c, int32_t k, int32_t vel_interp, int32_t dens_interp) {
  if (c->np <= 0) AV_FAIL(c, AVISM_ERR_STATE, "no particles loaded");
  if (k <= 0) AV_FAIL(c, AVISM_ERR_INPUT, "KNEIGHBOURS must be positive");
  if (k > 64) AV_FAIL(c, AVISM_ERR_UNSUPPORTED, "KNEIGHBOURS above 64 is not supported by this build");
  if (k > c->np) AV_FAIL(c, AVISM_ERR_INPUT, "KNEIGHBOURS exceeds the number of particles");
  if (vel_interp != 1 || dens_interp != 1)
    AV_FAIL(c, AVISM_ERR_UNSUPPORTED, "only SPH interpolation (FLAG_VEL_INTERP=FLAG_DENS_INTERP=1) is built");
  const int64_t np = c->np;
  const int n = c->g.n;
  const int64_t ncell = (int64_t)n * n * n;
  int rc = build_bins(c);
  if (rc) return rc;
  if (!c->pdens.ensure((size_t)np * 4) || !c->hk.ensure((size_t)np * 4) ||
      !c->hpart.ensure((size_t)np * 4) || !c->norm.ensure((size_t)np * 4) ||
      !c->fu.ensure((size_t)ncell * 4) || !c->fv.ensure((size_t)ncell * 4) ||
      !c->fw.ensure((size_t)ncell * 4) || !c->u_mass.ensure((size_t)ncell * 4) ||
      !c->u1co.ensure((size_t)ncell * 4) || !c->div.ensure((size_t)ncell * 4))
    AV_FAIL(c, AVISM_ERR_HIP, "out of device memory (fields)");
  float* d_rad = (float*)c->scalars.ensure(4096 + c->g.rad.size() * 4);
  float* rad = d_rad + 1024;
  AV_CHECK_HIP(c, hipMemcpyAsync(rad, c->g.rad.data(), c->g.rad.size() * 4, hipMemcpyHostToDevice,
                                 c->stream));
  Bins B = make_bins(c);
  {
    StageTimer t(c, "ppart_dens");
    if (k <= 16) launch_ppart<16>(c, B, k, 1);
    else if (k <= 32) launch_ppart<32>(c, B, k, 1);
    else launch_ppart<64>(c, B, k, 1);
    AV_CHECK_HIP(c, hipGetLastError());
  }
  AV_CHECK_HIP(c, hipMemsetAsync(c->hpart.p, 0, (size_t)np * 4, c->stream));
  {
    StageTimer t(c, "vvel_sph_vw");
    if (k <= 16) launch_vvel<16>(c, B, rad, k, 1);
    else if (k <= 32) launch_vvel<32>(c, B, rad, k, 1);
    else launch_vvel<64>(c, B, rad, k, 1);
    AV_CHECK_HIP(c, hipGetLastError());
    k_fill_h<<<nblk(np), 256, 0, c->stream>>>(np, c->hpart.as<float>(), c->hk.as<float>());
  }
  DdensGeom G{rad, c->g.rlo, n, c->g.dx, c->g.rx1};
  int32_t* rmax = (int32_t*)c->scalars.p + 8;
  int64_t nbins = (int64_t)c->nb * c->nb * c->nb;
  {
    StageTimer t(c, "ddens_norm");
    AV_CHECK_HIP(c, hipMemsetAsync(c->bin_reach.p, 0, (size_t)nbins * 4, c->stream));
    AV_CHECK_HIP(c, hipMemsetAsync(rmax, 0, 4, c->stream));
    k_ddens_norm<<<nblk(np), 256, 0, c->stream>>>(np, G, c->spos4.as<float4>(), c->hpart.as<float>(),
                                                   c->norm.as<float>(), c->bin_reach.as<int32_t>(),
                                                   B, rmax);
    AV_CHECK_HIP(c, hipGetLastError());
  }
  int32_t hr = 0;
  AV_CHECK_HIP(c, hipMemcpyAsync(&hr, rmax, 4, hipMemcpyDeviceToHost, c->stream));
  AV_CHECK_HIP(c, hipStreamSynchronize(c->stream));
  // cosmology constants (avism.f90:469-479, 599-606, 1249-1250)
  float den = ((c->g.dx * c->g.dx) * c->g.dx) * ((c->rete * c->rete) * c->rete);
  dim3 grid((n + 7) / 8, (n + 7) / 8, (n + 3) / 4);
  int64_t nparts = (int64_t)grid.x * grid.y * grid.z;
  if (!c->partial64.ensure((size_t)nparts * 8)) AV_FAIL(c, AVISM_ERR_HIP, "out of device memory");
  {
    StageTimer t(c, "ddens_gather");
    k_ddens_gather<<<grid, 256, 0, c->stream>>>(G, B, c->hpart.as<float>(), c->norm.as<float>(),
                                                c->bin_reach.as<int32_t>(), hr, den, c->rote,
                                                c->u_mass.as<float>(), c->partial64.as<double>());
    AV_CHECK_HIP(c, hipGetLastError());
  }
  double* dsum = (double*)((char*)c->scalars.p + 256);
  float* dmean = (float*)((char*)c->scalars.p + 264);
  {
    StageTimer t(c, "mean_div");
    k_reduce_mean<<<1, 256, 0, c->stream>>>(c->partial64.as<double>(), nparts, ncell, dsum, dmean);
    dim3 g2((n + 255) / 256, n, n);
    k_div_finalize<<<g2, 256, 0, c->stream>>>(n, c->g.dx, c->g.periodic, c->fu.as<float>(),
                                              c->fv.as<float>(), c->fw.as<float>(),
                                              c->u_mass.as<float>(), dmean, c->u1co.as<float>(),
                                              c->div.as<float>());
    AV_CHECK_HIP(c, hipGetLastError());
  }
  float hmean = 0;
  AV_CHECK_HIP(c, hipMemcpyAsync(&hmean, dmean, 4, hipMemcpyDeviceToHost, c->stream));
  AV_CHECK_HIP(c, hipMemcpyAsync(&c->meansum, dsum, 8, hipMemcpyDeviceToHost, c->stream));
  AV_CHECK_HIP(c, hipStreamSynchronize(c->stream));
  c->meandens_raw = hmean;
  float ul3 = (kUL * kUL) * kUL;
  c->meandens = hmean * (kUM / ul3);  // avism.f90:1409
  c->fields_ready = true;
  return AVISM_OK;
}

int fields_from_grid(avism_ctx* c, const float* delta, const float* vx, const float* vy,
                     const float* vz, int where) {
  if (!c->grid_set) AV_FAIL(c, AVISM_ERR_STATE, "avism_set_grid must be called first");
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
  dim3 g2((n + 255) / 256, n, n);
  k_div_finalize<<<g2, 256, 0, c->stream>>>(n, c->g.dx, c->g.periodic, c->fu.as<float>(),
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
  float* srcs[2] = {c->pdens.as<float>(), c->hpart.as<float>()};
  float* dsts[2] = {part_dens, hpart};
  for (int i = 0; i < 2; i++) {
    if (!dsts[i]) continue;
    k_scatter_orig<<<nblk(np), 256, 0, c->stream>>>(np, c->sorig.as<uint32_t>(), srcs[i], tmp);
    AV_CHECK_HIP(c, hipMemcpyAsync(dsts[i], tmp, (size_t)np * 4, hipMemcpyDeviceToHost, c->stream));
    AV_CHECK_HIP(c, hipStreamSynchronize(c->stream));
  }
  return AVISM_OK;
}

}  // namespace avism
