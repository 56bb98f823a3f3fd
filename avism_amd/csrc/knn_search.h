// knn_search.h -- the exact k-nearest-neighbour search over bin-sorted
// particles (kdtree.f90:504-675 semantics), shared by the field kernels.
//
// Written to compile both as HIP device code (fields.hip) and as plain host
// C++, so that a host test can check it against brute force.
#pragma once

#include <math.h>
#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define AVK_DEV __device__ __forceinline__
#else
#include <algorithm>
#define AVK_DEV inline
struct float4 {
  float x, y, z, w;
};
inline float4 make_float4(float x, float y, float z, float w) { return float4{x, y, z, w}; }
using std::max;
using std::min;
#endif

namespace avism {

// per-axis separation, minimum image under periodic boundaries
AVK_DEV float axis_d(float p, float q, float L, int periodic) {
  float d = p - q;
  if (periodic) {  // the smaller of |d| and L - |d| (fminf: both are finite and non-negative)
    const float a = fabsf(d);
    d = fminf(a, L - a);
  }
  return d;
}

// kdtree.f90:1441-1465 (squared distance, x, y, z accumulated in order)
AVK_DEV float dist2(float px, float py, float pz, float qx, float qy, float qz, float L, int periodic) {
  float dx = axis_d(px, qx, L, periodic);
  float dy = axis_d(py, qy, L, periodic);
  float dz = axis_d(pz, qz, L, periodic);
  float d = dx * dx;
  d = d + dy * dy;
  d = d + dz * dz;
  return d;
}

// uniform bins over the box; particles sorted by bin, start[] = bin offsets
struct Bins {
  const float4* __restrict__ pos;      // sorted (x,y,z,m)
  const uint32_t* __restrict__ orig;   // sorted -> input index
  const int32_t* __restrict__ start;   // nb^3 + 1
  int32_t nb;
  float lo, bsz, L, eps;
  int periodic;
  AVK_DEV int bin_of(float v) const {
    int b = (int)floorf((v - lo) / bsz);
    return b < 0 ? 0 : (b >= nb ? nb - 1 : b);
  }
  AVK_DEV int wrap(int b) const {
    if (b < 0) b += nb;
    else if (b >= nb) b -= nb;
    return b;
  }
  // squared distance from q to the (unwrapped) bin box at index b, axis-wise
  AVK_DEV float box_d(float q, int b) const {
    float blo = lo + (float)b * bsz - eps, bhi = lo + (float)(b + 1) * bsz + eps;
    float d = 0.0f;
    if (q < blo) d = blo - q;
    else if (q > bhi) d = q - bhi;
    return d;
  }
  // lower bound of the distance from q (in bin b) to anything outside shells 0..s
  AVK_DEV float outside(float q, int b, int s) const {
    float dlo = q - (lo + (float)(b - s) * bsz) - eps;
    float dhi = (lo + (float)(b + s + 1) * bsz) - q - eps;
    if (!periodic) {
      if (b - s <= 0) dlo = INFINITY;
      if (b + s + 1 >= nb) dhi = INFINITY;
    }
    return dlo < dhi ? dlo : dhi;
  }
};

// ------------------------------------------------------------------ kNN
AVK_DEV uint32_t f2u(float f) {
  union {
    float f;
    uint32_t u;
  } c;
  c.f = f;
  return c.u;
}
AVK_DEV float u2f(uint32_t u) {
  union {
    uint32_t u;
    float f;
  } c;
  c.u = u;
  return c.f;
}

// The k best candidates of one query (one lane), as 64-bit keys: the squared
// distance (non-negative floats order as their bit patterns) over the
// particle's input index.  Keys are unique, and the smaller key is the closer
// particle, on equal distances the one with the lower input index: the
// deterministic form of kdtree.f90's strict-< admission.
//
// The 4*NG slots and the NG group maxima live in LDS (entry t of lane l at
// [t * BS + l], so every lane hits its own bank); only the worst key and its
// group stay in registers.  A candidate farther than the worst key costs one
// compare.  Replacing the worst key reads its group's 4 keys and the NG group
// maxima and writes two entries, with no data-dependent branches.  Slots at or
// beyond k hold key 0 and never become a maximum.
template <int BS, int NG>
struct Sel {
  uint64_t* keys;  // this lane's 4*NG slots in LDS
  uint64_t* gmx;   // this lane's NG group maxima in LDS
  uint64_t worst;  // largest key held once full (~0 while filling)
  int wg, k, cnt;

  AVK_DEV void init(uint64_t* ks, uint64_t* gm, int k_) {
    keys = ks;
    gmx = gm;
    k = k_;
    reset();
  }
  AVK_DEV void reset() {
    cnt = 0;
    worst = ~0ull;
    wg = 0;
#pragma unroll
    for (int t = 0; t < 4 * NG; ++t) keys[t * BS] = 0ull;
  }
  AVK_DEV bool full() const { return cnt >= k; }
  // squared distance of the worst key held (INFINITY while filling)
  AVK_DEV float wd() const { return cnt >= k ? u2f((uint32_t)(worst >> 32)) : INFINITY; }
  // the largest group maximum and its group
  AVK_DEV void find_worst() {
    uint64_t w = gmx[0];
    int g0 = 0;
#pragma unroll
    for (int g = 1; g < NG; ++g) {
      const uint64_t x = gmx[g * BS];
      const bool b = x > w;
      w = b ? x : w;
      g0 = b ? g : g0;
    }
    worst = w;
    wg = g0;
  }
  // candidate at squared distance v with input index o
  AVK_DEV void consider(float v, uint32_t o) {
    const uint32_t vb = f2u(v);
    if (vb > (uint32_t)(worst >> 32)) return;  // also false for every candidate while filling
    const uint64_t key = ((uint64_t)vb << 32) | o;
    if (cnt < k) {
      keys[cnt * BS] = key;
      if (++cnt == k) {  // full: the group maxima, then the worst
#pragma unroll
        for (int g = 0; g < NG; ++g) {
          uint64_t m = keys[(4 * g) * BS];
#pragma unroll
          for (int s = 1; s < 4; ++s) {
            const uint64_t x = keys[(4 * g + s) * BS];
            m = x > m ? x : m;
          }
          gmx[g * BS] = m;
        }
        find_worst();
      }
      return;
    }
    if (key >= worst) return;
    // the worst key sits in group wg: replace it and recompute that group's maximum
    uint64_t* grp = keys + (4 * wg) * BS;
    uint64_t x[4];
#pragma unroll
    for (int s = 0; s < 4; ++s) x[s] = grp[s * BS];
    int p = 0;
#pragma unroll
    for (int s = 1; s < 4; ++s) p = x[s] == worst ? s : p;
#pragma unroll
    for (int s = 0; s < 4; ++s) x[s] = s == p ? key : x[s];
    grp[p * BS] = key;
    uint64_t m = x[0];
#pragma unroll
    for (int s = 1; s < 4; ++s) m = x[s] > m ? x[s] : m;
    gmx[wg * BS] = m;
    find_worst();
  }
  AVK_DEV float dist(int t) const { return u2f((uint32_t)(keys[t * BS] >> 32)); }
  AVK_DEV uint32_t id(int t) const { return (uint32_t)keys[t * BS]; }
};

// exact kNN of q over the bin-sorted particles (k <= number of particles).
//
// The rows of bins (fixed y and z bin, a run of x bins) around q's bin are
// visited in square rings of increasing size in y-z, centre first.  A pass
// takes every particle within a radius R of q: a row's x run is clipped to the
// bins that can hold a particle within min(R, current k-th distance), and
// candidates farther than R are skipped, so a pass considers each particle
// within R exactly once.  If k particles were found the answer is exact (all
// particles within R were seen); otherwise the next pass takes the shell
// between R and 2R.  R starts at the caller's estimate of the k-th distance.
// Each particle in a row run is one contiguous range of the sorted particles
// (two under the periodic wrap); the rows are taken RB at a time so that their
// 4*RB bin-start loads are in flight together, and their candidates are then
// read in one flat pass, the next candidate's load in flight while the current
// one is considered.
constexpr int RB = 4;  // rows per batch (2 ranges each: pick8 below)

// the sorted-particle range of bins xa..xb (unwrapped, xb - xa < nb) of the row
// starting at rowb, as up to two [start index, end index) slots into B.start
AVK_DEV void row_slots(const Bins& B, int rowb, int xa, int xb, bool on, int& a0, int& b0,
                                          bool& v0, int& a1, int& b1, bool& v1) {
  const int nb = B.nb;
  v0 = v1 = false;
  if (!on) return;
  if (xa < 0) {  // periodic images on the low side: bins xa..min(xb,-1)
    a0 = rowb + xa + nb;
    b0 = rowb + min(xb, -1) + nb + 1;
    v0 = true;
    xa = 0;
  } else if (xb >= nb) {  // and on the high side: bins max(xa,nb)..xb
    a0 = rowb + max(xa, nb) - nb;
    b0 = rowb + xb - nb + 1;
    v0 = true;
    xb = nb - 1;
  }
  if (xa <= xb) {
    a1 = rowb + xa;
    b1 = rowb + xb + 1;
    v1 = true;
  }
}

// a[g] for a runtime g in 0..7 without dynamic register indexing
AVK_DEV int pick8(int a0, int a1, int a2, int a3, int a4, int a5, int a6, int a7, int g) {
  const int lo = g & 1 ? ((g & 2) ? a3 : a1) : ((g & 2) ? a2 : a0);
  const int hi = g & 1 ? ((g & 2) ? a7 : a5) : ((g & 2) ? a6 : a4);
  return (g & 4) ? hi : lo;
}

// rows of ring r >= 1 of the (2r+1)^2 square of rows, t in 0..8r-1: the
// two lines at dz = -r and dz = r, then the columns at dy = -r and dy = r
AVK_DEV void ring_row(int r, int t, int& dy, int& dz) {
  const int side = 2 * r + 1;
  if (t < side) {
    dy = t - r;
    dz = -r;
  } else if (t < 2 * side) {
    dy = t - side - r;
    dz = r;
  } else {
    const int u = t - 2 * side;
    dy = (u & 1) ? r : -r;
    dz = (u >> 1) - r + 1;
  }
}

// PER: B.periodic as a compile-time constant; r0: first search radius
template <int PER, class H>
AVK_DEV void knn_global(const Bins& B, int64_t np, float qx, float qy, float qz, float r0, H& h) {
  const int bx = B.bin_of(qx), by = B.bin_of(qy), bz = B.bin_of(qz);
  const int nb = B.nb;
  const float e2 = 2.0f * B.eps;
  const float ib = 1.0f / B.bsz;
  // rows and x runs reach at most rmax bins from q's bin (periodic: no bin twice)
  const int rmax = PER ? (nb - 1) / 2 : nb;
  float R2 = r0 * r0, lo2 = -1.0f;
  for (;;) {
    bool exhausted = true;
    // ring 1 carries the centre row as well
    for (int r = rmax < 1 ? 0 : 1; r <= rmax; ++r) {
      if (r > 1) {  // every row of ring r is farther than the bound in y or z
        const float lb = fminf(B.outside(qy, by, r - 1), B.outside(qz, bz, r - 1));
        if (lb == INFINITY || (lb > 0.0f && lb * lb > fminf(R2, h.wd()))) {
          exhausted = false;
          break;
        }
      }
      const int nrow = r == 0 ? 1 : (r == 1 ? 9 : 8 * r);
      for (int r0w = 0; r0w < nrow; r0w += RB) {
        const float T2 = fminf(R2, h.wd());
        int sa[2 * RB], sb[2 * RB];
        bool sv[2 * RB];
#pragma unroll
        for (int t = 0; t < RB; ++t) {
          const int row = r0w + t;
          int dy = 0, dz = 0;
          if (r > 1)
            ring_row(r, row, dy, dz);
          else if (r == 1 && row > 0)
            ring_row(1, row - 1, dy, dz);
          const int cy = by + dy, cz = bz + dz;
          bool on = row < nrow;
          if (!PER) on = on && cy >= 0 && cy < nb && cz >= 0 && cz < nb;
          int lo_b = bx - rmax, hi_b = bx + rmax;
          if (!PER) {
            lo_b = max(lo_b, 0);
            hi_b = min(hi_b, nb - 1);
          }
          const float ddy = B.box_d(qy, cy), ddz = B.box_d(qz, cz);
          const float rem = T2 - (ddy * ddy + ddz * ddz);
          if (rem < 0.0f) {
            on = false;
          } else if (rem < INFINITY) {  // bins whose box can hold d^2 <= T2
            const float rr = sqrtf(rem);
            lo_b = max(lo_b, (int)floorf((qx - rr - e2 - B.lo) * ib));
            hi_b = min(hi_b, (int)floorf((qx + rr + e2 - B.lo) * ib));
          }
          on = on && lo_b <= hi_b;
          const int wy = PER ? B.wrap(cy) : cy, wz = PER ? B.wrap(cz) : cz;
          row_slots(B, nb * (wy + nb * wz), lo_b, hi_b, on, sa[2 * t], sb[2 * t], sv[2 * t], sa[2 * t + 1],
                    sb[2 * t + 1], sv[2 * t + 1]);
        }
        int j0[2 * RB], j1[2 * RB];
#pragma unroll
        for (int g = 0; g < 2 * RB; ++g) {
          j0[g] = sv[g] ? B.start[sa[g]] : 0;
          j1[g] = sv[g] ? B.start[sb[g]] : 0;
        }
        // one flat pass over the batch's candidates: a lane moves on to its next
        // range as soon as one ends, so a wave pays for its longest lane's total
        // rather than for the longest lane of every range
        unsigned nz = 0;  // the batch's non-empty ranges
#pragma unroll
        for (int t = 0; t < 2 * RB; ++t) nz |= (j1[t] > j0[t] ? 1u : 0u) << t;
        int g = -1, j = 0, je = 0;
        auto adv = [&]() {  // on to the next non-empty range once this one is done
          if (j < je) return;
          const unsigned rest = nz & (~0u << (g + 1));
          if (rest == 0u) return;
          g = __builtin_ctz(rest);
          j = pick8(j0[0], j0[1], j0[2], j0[3], j0[4], j0[5], j0[6], j0[7], g);
          je = pick8(j1[0], j1[1], j1[2], j1[3], j1[4], j1[5], j1[6], j1[7], g);
        };
        adv();
        float4 p = make_float4(0.f, 0.f, 0.f, 0.f);
        uint32_t o = 0;
        if (j < je) {
          p = B.pos[j];
          o = B.orig[j];
        }
        while (j < je) {
          const float4 pc = p;
          const uint32_t oc = o;
          ++j;
          adv();
          if (j < je) {  // the next candidate's load is in flight while this one is considered
            p = B.pos[j];
            o = B.orig[j];
          }
          const float d2 = dist2(pc.x, pc.y, pc.z, qx, qy, qz, B.L, PER);
          if (d2 <= R2 && d2 > lo2) h.consider(d2, oc);
        }
      }
    }
    if (h.full()) {
      // exact unless the +-rmax clip (periodic boxes) cut into the ball of the
      // k-th distance: in x always, in y and z when the rings ran out
      float bd = B.outside(qx, bx, rmax);
      if (exhausted) bd = fminf(bd, fminf(B.outside(qy, by, rmax), B.outside(qz, bz, rmax)));
      if (bd == INFINITY || (bd > 0.0f && h.wd() < bd * bd)) return;
      break;
    }
    if (exhausted) break;  // the rings reached rmax without k particles
    lo2 = R2;
    R2 *= 4.0f;
  }
  // periodic box smaller than the search: every particle once
  h.reset();
  for (int64_t j = 0; j < np; ++j) {
    const float4 p = B.pos[j];
    h.consider(dist2(p.x, p.y, p.z, qx, qy, qz, B.L, PER), B.orig[j]);
  }
}

}  // namespace avism
