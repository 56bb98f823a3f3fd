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

// square root for search bounds only: the hardware approximation (within 1
// ulp), far inside the eps margins the bounds carry
AVK_DEV float bound_sqrt(float x) {
#if defined(__HIPCC__)
  return __builtin_amdgcn_sqrtf(x);
#else
  return sqrtf(x);
#endif
}

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
  // distance from q to the (unwrapped) bin box at index b along one axis
  AVK_DEV float box_d(float q, int b) const {
    const float blo = lo + (float)b * bsz - eps, bhi = lo + (float)(b + 1) * bsz + eps;
    return fmaxf(fmaxf(blo - q, q - bhi), 0.0f);  // at most one of the two is positive
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

// The k smallest squared distances seen by one query (one lane), as their
// bit patterns (non-negative floats order as their bits), with multiplicity.
// The 4*NG slots and the NG group maxima live in LDS (entry t of lane l at
// [t * BS + l], so every lane hits its own bank); only the worst value and its
// group stay in registers.  A value not below the worst costs one compare;
// replacing the worst reads its group's 4 slots and the NG group maxima and
// writes two entries, with no data-dependent branches.  Slots at or beyond k
// hold 0 and never become a maximum.  The same structure selects the m
// smallest input indices among tied candidates.
//
// IDX: each slot's candidate (its sorted position) is also kept, in
// registers (idr, written with a select per slot so that it never leaves
// them), and rej counts the candidates equal to the worst value that were
// turned away or evicted since the worst took that value.  A search that ends
// with rej == 0 holds every candidate at the k-th distance, so its slots are
// exactly the k neighbours.
template <int BS, int NG, bool IDX = false>
struct Sel {
  uint32_t* keys;  // this lane's 4*NG slots in LDS
  uint32_t* gmx;   // this lane's NG group maxima in LDS
  uint32_t worst;  // largest value held once full (~0 while filling)
  int wg, k, cnt;
  uint32_t rej;
  uint32_t idr[IDX ? 4 * NG : 1];

  AVK_DEV void init(uint32_t* ks, uint32_t* gm, int k_) {
    keys = ks;
    gmx = gm;
    reset(k_);
  }
  AVK_DEV void reset(int k_) {
    k = k_;
    cnt = 0;
    worst = ~0u;
    wg = 0;
    rej = 0;
#pragma unroll
    for (int t = 0; t < 4 * NG; ++t) keys[t * BS] = 0u;
  }
  // slot = 4 * g + s: one compare per group and per position, their products
  // formed on the scalar unit
  AVK_DEV void set_idx(int g, int s, uint32_t j) {
    if (IDX) {
      bool mg[NG], ms[4];
#pragma unroll
      for (int t = 0; t < NG; ++t) mg[t] = t == g;
#pragma unroll
      for (int t = 0; t < 4; ++t) ms[t] = t == s;
#pragma unroll
      for (int t = 0; t < 4 * NG; ++t) idr[t] = mg[t / 4] && ms[t % 4] ? j : idr[t];
    }
  }
  AVK_DEV bool full() const { return cnt >= k; }
  // the k-th smallest value as a float (INFINITY while filling)
  AVK_DEV float wd() const { return cnt >= k ? u2f(worst) : INFINITY; }
  AVK_DEV void find_worst() {
    uint32_t w = gmx[0];
    int g0 = 0;
#pragma unroll
    for (int g = 1; g < NG; ++g) {
      const uint32_t x = gmx[g * BS];
      const bool b = x > w;
      w = b ? x : w;
      g0 = b ? g : g0;
    }
    worst = w;
    wg = g0;
  }
  AVK_DEV void consider(uint32_t key, uint32_t j = 0) {
    if (key >= worst) {  // never true while filling
      if (IDX && key == worst) ++rej;
      return;
    }
    if (cnt < k) {
      keys[cnt * BS] = key;
      set_idx(cnt >> 2, cnt & 3, j);
      if (++cnt == k) {  // full: the group maxima, then the worst
#pragma unroll
        for (int g = 0; g < NG; ++g) {
          uint32_t m = keys[(4 * g) * BS];
#pragma unroll
          for (int s = 1; s < 4; ++s) {
            const uint32_t x = keys[(4 * g + s) * BS];
            m = x > m ? x : m;
          }
          gmx[g * BS] = m;
        }
        find_worst();
      }
      return;
    }
    // a copy of the worst value sits in group wg: replace it, recompute the group's maximum
    uint32_t* grp = keys + (4 * wg) * BS;
    uint32_t x[4];
#pragma unroll
    for (int s = 0; s < 4; ++s) x[s] = grp[s * BS];
    int p = 0;
#pragma unroll
    for (int s = 1; s < 4; ++s) p = x[s] == worst ? s : p;
#pragma unroll
    for (int s = 0; s < 4; ++s) x[s] = s == p ? key : x[s];
    grp[p * BS] = key;
    set_idx(wg, p, j);
    uint32_t m = x[0];
#pragma unroll
    for (int s = 1; s < 4; ++s) m = x[s] > m ? x[s] : m;
    gmx[wg * BS] = m;
    const uint32_t w0 = worst;
    find_worst();
    // the evicted entry held w0: still the worst, it is a tie turned away
    if (IDX) rej = worst == w0 ? rej + 1 : 0u;
  }
};

// Exact kNN of q over the bin-sorted particles (k <= number of particles), in
// two steps: the k-th smallest squared distance dk2 (knn_kth), then a sweep of
// the ball d^2 <= dk2 that hands the k nearest to the caller (knn_each).  The
// neighbours are those with d^2 < dk2 and, of those with d^2 == dk2, the ones
// with the lowest input indices: the deterministic form of kdtree.f90's
// strict-< admission.
//
// Sweeps visit the rows of bins (fixed y and z bin, a run of x bins) around
// q's bin in square rings of increasing size in y-z, centre first, each row
// clipped to the x bins that can hold a particle within the visitor's current
// bound.  Each row run is one contiguous range of the sorted particles (two
// under the periodic wrap); the rows are taken RB at a time so that their 4*RB
// bin-start loads are in flight together, and their candidates are read in
// one flat pass, the next candidate's load in flight while the current one is
// handled.
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

// rows and x runs reach at most rmax bins from q's bin (periodic: no bin twice)
template <int PER>
AVK_DEV int sweep_rmax(const Bins& B) {
  return PER ? (B.nb - 1) / 2 : B.nb;
}

// A candidate's record as a sweep loads it: by default the bin-sorted
// (x, y, z, m); a consumer that needs more per neighbour (velocities, HPART)
// supplies a loader whose record also carries those, so that they are loaded
// ahead with the position.
struct PosLoad {
  using Rec = float4;
  const float4* pos;
  AVK_DEV Rec load(int j) const { return pos[j]; }
  AVK_DEV static const float4& xyz(const Rec& r) { return r; }
};

// One sweep: every candidate in the visited rows goes to v.cand(d2, j, p) (j =
// sorted position, p = its record), rows pruned by v.bound() (re-read per
// batch, so it may shrink).  Returns true if the rings ran out at rmax, false
// if they stopped because every further row is beyond the bound.
template <int PER, class V>
AVK_DEV bool ring_sweep(const Bins& B, float qx, float qy, float qz, V& v) {
  const int bx = B.bin_of(qx), by = B.bin_of(qy), bz = B.bin_of(qz);
  const int nb = B.nb;
  const float e2 = 2.0f * B.eps;
  const float ib = 1.0f / B.bsz;
  const int rmax = sweep_rmax<PER>(B);
  // ring 1 carries the centre row as well
  for (int r = rmax < 1 ? 0 : 1; r <= rmax; ++r) {
    if (r > 1) {  // every row of ring r is beyond the bound in y or z
      const float lb = fminf(B.outside(qy, by, r - 1), B.outside(qz, bz, r - 1));
      if (lb == INFINITY || (lb > 0.0f && lb * lb > v.bound())) return false;
    }
    const int nrow = r == 0 ? 1 : (r == 1 ? 9 : 8 * r);
    for (int r0w = 0; r0w < nrow; r0w += RB) {
      const float T2 = v.bound();
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
          const float rr = bound_sqrt(rem);
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
      typename V::Rec p{};
      if (j < je) p = v.ld.load(j);
      // two candidates per trip: the loaded record alternates between two
      // register sets instead of being copied each candidate
#pragma unroll 2
      while (j < je) {
        const typename V::Rec pc = p;
        const int jc = j;
        ++j;
        adv();
        if (j < je) p = v.ld.load(j);  // the next candidate's load is in flight while this one is handled
        const float4& x = V::Load::xyz(pc);
        v.cand(dist2(x.x, x.y, x.z, qx, qy, qz, B.L, PER), jc, pc);
      }
    }
  }
  return true;
}

// after a sweep bounded by w: did it see every particle with d^2 <= w?  Only
// the +-rmax clip of periodic boxes can cut into that ball: in x always, in y
// and z when the rings ran out.
template <int PER>
AVK_DEV bool sweep_complete(const Bins& B, float qx, float qy, float qz, bool exhausted, float w) {
  const int rmax = sweep_rmax<PER>(B);
  float bd = B.outside(qx, B.bin_of(qx), rmax);
  if (exhausted) bd = fminf(bd, fminf(B.outside(qy, B.bin_of(qy), rmax), B.outside(qz, B.bin_of(qz), rmax)));
  return bd == INFINITY || (bd > 0.0f && w < bd * bd);
}

// the k-th smallest squared distance from q.  A pass takes every particle
// within a radius R of q, the rows clipped to min(R, current k-th distance):
// if k were found the result is exact, otherwise the next pass takes the
// shell between R and 2R.  R starts at r0, the caller's estimate of the k-th
// distance.  scanned: the ring sweeps could not cover the ball (a periodic box
// smaller than the search) and every particle was scanned instead.
template <int PER, class S>
struct KthVisitor {
  using Load = PosLoad;
  using Rec = PosLoad::Rec;
  PosLoad ld;
  S& sel;
  float R2, lo2;
  AVK_DEV float bound() const { return fminf(R2, sel.wd()); }
  AVK_DEV void cand(float d2, int j, const Rec&) {
    if (d2 <= R2 && d2 > lo2) sel.consider(f2u(d2), (uint32_t)j);
  }
};

template <int PER, class S>
AVK_DEV float knn_kth(const Bins& B, int64_t np, float qx, float qy, float qz, float r0, int k, S& sel,
                      bool& scanned) {
  sel.reset(k);
  scanned = false;
  KthVisitor<PER, S> v{PosLoad{B.pos}, sel, r0 * r0, -1.0f};
  for (;;) {
    const bool ex = ring_sweep<PER>(B, qx, qy, qz, v);
    if (sel.full()) {
      if (sweep_complete<PER>(B, qx, qy, qz, ex, sel.wd())) return sel.wd();
      break;
    }
    if (ex) break;  // the rings reached rmax without k particles
    v.lo2 = v.R2;
    v.R2 *= 4.0f;
  }
  scanned = true;
  sel.reset(k);
  for (int64_t j = 0; j < np; ++j) {
    const float4 p = B.pos[j];
    sel.consider(f2u(dist2(p.x, p.y, p.z, qx, qy, qz, B.L, PER)), (uint32_t)j);
  }
  return sel.wd();
}

// visits every particle with d^2 <= w once (a ring sweep, or all particles)
template <int PER, class V>
AVK_DEV void ball_sweep(const Bins& B, int64_t np, float qx, float qy, float qz, bool scanned, V& v) {
  if (!scanned) {
    ring_sweep<PER>(B, qx, qy, qz, v);
    return;
  }
  for (int64_t j = 0; j < np; ++j) {
    const typename V::Rec r = v.ld.load((int)j);
    const float4& p = V::Load::xyz(r);
    v.cand(dist2(p.x, p.y, p.z, qx, qy, qz, B.L, PER), (int)j, r);
  }
}

template <class L, class F>
struct LessVisitor {  // d^2 < dk2: neighbours; d^2 == dk2: counted
  using Load = L;
  using Rec = typename L::Rec;
  L ld;
  float dk2;
  F& f;
  int nlt, neq, jeq;
  Rec peq;
  AVK_DEV float bound() const { return dk2; }
  AVK_DEV void cand(float d2, int j, const Rec& p) {
    if (d2 < dk2) {
      ++nlt;
      f(j, d2, p);
    } else if (d2 == dk2) {
      ++neq;
      jeq = j;
      peq = p;
    }
  }
};
template <class S>
struct TieSelVisitor {  // the m lowest input indices among d^2 == dk2
  using Load = PosLoad;
  using Rec = PosLoad::Rec;
  PosLoad ld;
  float dk2;
  const uint32_t* orig;
  S& sel;
  AVK_DEV float bound() const { return dk2; }
  AVK_DEV void cand(float d2, int j, const Rec&) {
    if (d2 == dk2) sel.consider(orig[j]);
  }
};
template <class L, class F>
struct TieVisitor {  // d^2 == dk2 and input index <= omax
  using Load = L;
  using Rec = typename L::Rec;
  L ld;
  float dk2;
  const uint32_t* orig;
  uint32_t omax;
  F& f;
  AVK_DEV float bound() const { return dk2; }
  AVK_DEV void cand(float d2, int j, const Rec& p) {
    if (d2 == dk2 && orig[j] <= omax) f(j, d2, p);
  }
};

// f(j, d2, r) for each of the k nearest particles (j = sorted position, r = its
// record as ld loads it): those with d^2 < dk2 in sweep order, then the ones
// chosen among the ties at dk2.  dk2 and scanned come from knn_kth; sel is
// free for reuse.
template <int PER, class S, class L, class F>
AVK_DEV void knn_each(const Bins& B, int64_t np, float qx, float qy, float qz, int k, float dk2, bool scanned,
                      S& sel, const L& ld, F&& f) {
  LessVisitor<L, F> lv{ld, dk2, f, 0, 0, 0, typename L::Rec{}};
  ball_sweep<PER>(B, np, qx, qy, qz, scanned, lv);
  const int m = k - lv.nlt;  // ties to take, 1 <= m <= neq
  if (lv.neq == 1) {  // the usual case: the k-th neighbour alone at dk2
    f(lv.jeq, dk2, lv.peq);
    return;
  }
  uint32_t omax = ~0u;
  if (m < lv.neq) {
    sel.reset(m);
    TieSelVisitor<S> sv{PosLoad{B.pos}, dk2, B.orig, sel};
    ball_sweep<PER>(B, np, qx, qy, qz, scanned, sv);
    omax = sel.worst;
  }
  TieVisitor<L, F> tv{ld, dk2, B.orig, omax, f};
  ball_sweep<PER>(B, np, qx, qy, qz, scanned, tv);
}

template <class L, class F>
struct BallVisitor {  // every particle with d^2 <= r2
  using Load = L;
  using Rec = typename L::Rec;
  L ld;
  float r2;
  F& f;
  AVK_DEV float bound() const { return r2; }
  AVK_DEV void cand(float d2, int j, const Rec& p) {
    if (d2 <= r2) f(j, d2, p);
  }
};

// f(j, d2, r) for every particle with d^2 <= r2 (ball_search, kdtree.f90:
// 1015-1072, 1126), in sweep order
template <int PER, class L, class F>
AVK_DEV void ball_each(const Bins& B, int64_t np, float qx, float qy, float qz, float r2, const L& ld, F&& f) {
  BallVisitor<L, F> v{ld, r2, f};
  // a ring sweep covers the ball unless the +-rmax clip cuts into it
  ball_sweep<PER>(B, np, qx, qy, qz, !sweep_complete<PER>(B, qx, qy, qz, true, r2), v);
}

}  // namespace avism
