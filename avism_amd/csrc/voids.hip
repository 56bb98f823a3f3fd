// voids.hip -- MARK_ALL + VOIDFIND (voidfinding.f90:13-713) on gfx950.
//
//  1. k_mark: one pass over the extended grid computes FLAG_SUB, FLAGV and the
//     six directional "cannot expand" flags of voidfinding.f90:256-432 into one
//     byte per cell.  Ghost cells read the inner fields through the periodic
//     wrap (REAL_MAKE_PERIODIC, num.f90:89-133), so no extended copies of U1CO
//     or DIVERCO are materialised.
//  2. ordered compaction of the candidate centres (Fortran order, 161-176).
//  3. INDEXX on their divergence (sort.hip), reversed: descending order.
//  4. growth: every candidate's half-width depends only on the static flag
//     grid, so all candidates grow in parallel (one lane each, shell by shell;
//     candidates still growing after a few shells continue in a kernel where a
//     whole wave checks each shell's faces with ballots).  The reference's
//     edge-zone rules (FLAG2, voidfinding.f90:509-679) are applied exactly.
//  5. acceptance: candidate t is kept iff it grew to r >= 1 and no kept cube of
//     a lower t covers its centre (voidfinding.f90:457).  Resolved in rounds:
//     each round every undecided cube paints its cells with atomicMin of its
//     rank; a candidate whose centre carries its own rank has no undecided
//     predecessor over it and is accepted, one whose centre is covered by an
//     accepted lower rank is rejected.  The lowest undecided rank is always
//     decided, so the rounds terminate; ids are the ranks' prefix count.
#include <algorithm>
#include <cstring>

#include <rocprim/rocprim.hpp>

#include "avism_internal.h"

namespace avism {

namespace {

struct ExtGeom {
  int n, low1, low2, ne, periodic;
  __device__ __forceinline__ int wrap(int i) const {  // ext ref index -> inner 1..n
    if (i < 1) return i + n;
    if (i > n) return i - n;
    return i;
  }
  __device__ __forceinline__ size_t inner(int i, int j, int k) const {
    return (size_t)(wrap(i) - 1) + (size_t)n * ((size_t)(wrap(j) - 1) + (size_t)n * (wrap(k) - 1));
  }
  __device__ __forceinline__ size_t ext(int i, int j, int k) const {
    return (size_t)(i - low1) + (size_t)ne * ((size_t)(j - low1) + (size_t)ne * (k - low1));
  }
};

struct MarkParams {
  float dthr1, dthr2, grad, divthr;
  float two_dx;
};

// voidfinding.f90:256-432 (and MARK_ALL, 30-52) for every ext cell; ref
// order +X,-X,+Y,-Y,+Z,-Z.  A block takes a 64 x 8 (x, y) column of cells and
// walks kMarkZ planes of it in z, kMarkP planes per step: the planes' U1CO
// (halo 3) and DIVERCO (halo 1) are staged in LDS for the x and y stencils,
// and each cell's z line (U1CO at z-3..z+3, DIVERCO at z-1..z+1) slides
// through registers.  The next step's planes are loaded into registers while
// the current step is computed from LDS.  Ext cells read the inner fields
// through the periodic wrap; halo cells outside LOW1..LOW2 are clamped into
// it (the stencil never uses them).
constexpr int kMarkX = 64, kMarkY = 8, kMarkZ = 32, kMarkP = 2;
constexpr int kMarkT = kMarkX * kMarkY;                          // threads per block
constexpr int kMarkU = (kMarkY + 6) * (kMarkX + 6);              // U1CO tile entries per plane
constexpr int kMarkD = (kMarkY + 2) * (kMarkX + 2);              // DIVERCO tile entries per plane
constexpr int kMarkUP = (kMarkU + kMarkT - 1) / kMarkT;          // per thread
constexpr int kMarkDP = (kMarkD + kMarkT - 1) / kMarkT;
static_assert(kMarkZ % kMarkP == 0, "whole steps per block");

struct Line {  // one axis' stencil values at offsets -3..3 (div at -1..1)
  float u[7];
  float d[3];
};

__device__ __forceinline__ bool mark_dir(const MarkParams& P, const Line& A, int pos, bool plus, int L1,
                                         int L2) {
  bool flag = plus ? (pos >= L2 - kNsec) : (pos <= L1 + kNsec);
  if (flag) return true;
  const int s = plus ? 1 : -1;
  const int II = pos + s;
  const float un = A.u[3 + s], dn = A.d[1 + s];
  const int edge = plus ? L2 : L1;
  if (II == edge) {
    if (dn < P.divthr) flag = true;
    if (un >= P.dthr2) flag = true;
  } else {
    const float dd = (A.u[3 + 2 * s] - A.u[3]) / P.two_dx;
    if (dd >= P.grad) flag = true;
    const bool nextok = plus ? (II < L2 - 1) : (II > L1 + 1);
    if (flag && nextok) {
      const float dd2 = (A.u[3 + 3 * s] - un) / P.two_dx;
      if (dd2 < P.grad) flag = false;
    }
    if (dn < P.divthr) flag = true;
    if (un >= P.dthr2) flag = true;
  }
  return flag;
}

// MASK: SMASK given (inner int8; the periodic shell is masked, avism.f90:866-904)
template <int MASK>
__global__ void __launch_bounds__(kMarkT) k_mark(ExtGeom G, MarkParams P, const float* __restrict__ u1co,
                                                 const float* __restrict__ div,
                                                 const int32_t* __restrict__ marcap,
                                                 const int8_t* __restrict__ smask, uint8_t* flags) {
  __shared__ float su[kMarkP][kMarkY + 6][kMarkX + 6];
  __shared__ float sd[kMarkP][kMarkY + 2][kMarkX + 2];
  __shared__ int8_t sm[MASK ? kMarkP : 1][kMarkY + 2][kMarkX + 2];
  const int L1 = G.low1, L2 = G.low2, n = G.n, ne = G.ne;
  const int tx = threadIdx.x % kMarkX, ty = threadIdx.x / kMarkX;
  const int x0 = blockIdx.x * kMarkX, y0 = blockIdx.y * kMarkY, z0 = blockIdx.z * kMarkZ;
  const int I = x0 + tx + L1, J = y0 + ty + L1;
  const bool live = x0 + tx < ne && y0 + ty < ne;
  auto clampc = [&](int a) { return a < L1 ? L1 : (a > L2 ? L2 : a); };
  auto inner_xy = [&](int xe, int ye) {
    return (size_t)(G.wrap(clampc(xe)) - 1) + (size_t)n * (size_t)(G.wrap(clampc(ye)) - 1);
  };
  auto plane = [&](int k) { return (size_t)n * (size_t)n * (size_t)(G.wrap(clampc(k)) - 1); };
  const size_t cxy = inner_xy(I, J);
  // the (x, y) inner offsets of the tile entries this thread stages (fixed over z)
  size_t uo[kMarkUP], dof[kMarkDP];
#pragma unroll
  for (int m = 0; m < kMarkUP; ++m) {
    const int q = threadIdx.x + m * kMarkT;
    uo[m] = inner_xy(x0 + q % (kMarkX + 6) - 3 + L1, y0 + q / (kMarkX + 6) - 3 + L1);
  }
#pragma unroll
  for (int m = 0; m < kMarkDP; ++m) {
    const int q = threadIdx.x + m * kMarkT;
    dof[m] = inner_xy(x0 + q % (kMarkX + 2) - 1 + L1, y0 + q / (kMarkX + 2) - 1 + L1);
  }
  // SMASK of an ext cell: outside the inner box it is 0
  auto inner_or = [&](int xe, int ye) -> int64_t {
    return (xe < 1 || xe > n || ye < 1 || ye > n) ? -1 : (int64_t)(xe - 1) + (int64_t)n * (ye - 1);
  };
  auto mask_at = [&](int64_t xy, int ze) -> int8_t {
    return (xy < 0 || ze < 1 || ze > n) ? (int8_t)0 : smask[xy + (int64_t)n * n * (ze - 1)];
  };
  int64_t mo[MASK ? kMarkDP : 1];
  const int64_t mxy = MASK ? inner_or(I, J) : 0;
  if (MASK) {
#pragma unroll
    for (int m = 0; m < kMarkDP; ++m) {
      const int q = threadIdx.x + m * kMarkT;
      mo[m] = inner_or(x0 + q % (kMarkX + 2) - 1 + L1, y0 + q / (kMarkX + 2) - 1 + L1);
    }
  }
  const int z1 = min(z0 + kMarkZ, ne);
  // the z line of the step's planes: U1CO at z-3..z+kMarkP+2, DIVERCO (and
  // SMASK) at z-1..z+kMarkP
  float zu[kMarkP + 6], zd[kMarkP + 2];
  int8_t zm[MASK ? kMarkP + 2 : 1];
  if (MASK) {
#pragma unroll
    for (int o = 0; o < kMarkP + 2; ++o) zm[o] = mask_at(mxy, z0 + L1 + o - 1);
  }
#pragma unroll
  for (int o = 0; o < kMarkP + 6; ++o) zu[o] = u1co[cxy + plane(z0 + L1 + o - 3)];
#pragma unroll
  for (int o = 0; o < kMarkP + 2; ++o) zd[o] = div[cxy + plane(z0 + L1 + o - 1)];
  // the step's planes go through registers: the next step's global loads are
  // issued before the current step is computed, so their latency is hidden
  float pu[kMarkP][kMarkUP], pd[kMarkP][kMarkDP];
  int8_t pm[MASK ? kMarkP : 1][MASK ? kMarkDP : 1];
  float nzu[kMarkP], nzd[kMarkP];
  int8_t nzm[MASK ? kMarkP : 1];
  auto fetch = [&](int z) {
#pragma unroll
    for (int pl = 0; pl < kMarkP; ++pl) {
      const size_t pk = plane(z + pl + L1);
#pragma unroll
      for (int m = 0; m < kMarkUP; ++m) {
        const int q = threadIdx.x + m * kMarkT;
        pu[pl][m] = q < kMarkU ? u1co[uo[m] + pk] : 0.0f;
      }
#pragma unroll
      for (int m = 0; m < kMarkDP; ++m) {
        const int q = threadIdx.x + m * kMarkT;
        pd[pl][m] = q < kMarkD ? div[dof[m] + pk] : 0.0f;
        if (MASK) pm[pl][m] = q < kMarkD ? mask_at(mo[m], z + pl + L1) : (int8_t)0;
      }
    }
  };
  fetch(z0);
  for (int z = z0; z < z1; z += kMarkP) {
#pragma unroll
    for (int pl = 0; pl < kMarkP; ++pl) {
#pragma unroll
      for (int m = 0; m < kMarkUP; ++m) {
        const int q = threadIdx.x + m * kMarkT;
        if (q < kMarkU) (&su[pl][0][0])[q] = pu[pl][m];
      }
#pragma unroll
      for (int m = 0; m < kMarkDP; ++m) {
        const int q = threadIdx.x + m * kMarkT;
        if (q < kMarkD) (&sd[pl][0][0])[q] = pd[pl][m];
        if (MASK && q < kMarkD) (&sm[pl][0][0])[q] = pm[pl][m];
      }
    }
    const bool more = z + kMarkP < z1;
    if (more) {  // the next step's planes and z line, in flight while this step is computed
      fetch(z + kMarkP);
#pragma unroll
      for (int o = 0; o < kMarkP; ++o) {
        nzu[o] = u1co[cxy + plane(z + kMarkP + L1 + o + 3)];
        nzd[o] = div[cxy + plane(z + kMarkP + L1 + o + 1)];
        if (MASK) nzm[o] = mask_at(mxy, z + kMarkP + L1 + o + 1);
      }
    }
    __syncthreads();
#pragma unroll
    for (int pl = 0; pl < kMarkP; ++pl) {
      const int K = z + pl + L1;
      if (live && z + pl < z1) {
        Line X, Y, Z;
#pragma unroll
        for (int o = 0; o < 7; ++o) {
          X.u[o] = su[pl][ty + 3][tx + o];
          Y.u[o] = su[pl][ty + o][tx + 3];
          Z.u[o] = zu[pl + o];
        }
#pragma unroll
        for (int o = 0; o < 3; ++o) {
          X.d[o] = sd[pl][ty + 1][tx + o];
          Y.d[o] = sd[pl][ty + o][tx + 1];
          Z.d[o] = zd[pl + o];
        }
        const int mp = marcap ? marcap[cxy + plane(K)] : -1;
        uint8_t f = 0;
        // MARK_ALL (voidfinding.f90:30-52): nothing in a masked cell
        if ((!MASK || zm[pl + 1] != 0) && mp != 0) {
          f |= FL_SUB;
          if (Z.u[3] < P.dthr1 && Z.d[1] > 0.0f) f |= FL_V;
        }
        // directions in the reference's order; a masked neighbour ends the
        // cell (CYCLE, voidfinding.f90:283, 308, ...): later directions stay 0
        auto step = [&](const Line& A, int pos, bool plus, int8_t nmask, uint8_t bit) -> bool {
          const bool edge = plus ? (pos >= L2 - kNsec) : (pos <= L1 + kNsec);
          if (!edge && MASK && nmask == 0) return false;
          if (mark_dir(P, A, pos, plus, L1, L2)) f |= bit;
          return true;
        };
        const int8_t mxp = MASK ? sm[pl][ty + 1][tx + 2] : 1, mxm = MASK ? sm[pl][ty + 1][tx] : 1;
        const int8_t myp = MASK ? sm[pl][ty + 2][tx + 1] : 1, mym = MASK ? sm[pl][ty][tx + 1] : 1;
        const int8_t mzp = MASK ? zm[pl + 2] : 1, mzm = MASK ? zm[pl] : 1;
        (void)(step(X, I, true, mxp, FL_X2) && step(X, I, false, mxm, FL_X1) && step(Y, J, true, myp, FL_Y2) &&
               step(Y, J, false, mym, FL_Y1) && step(Z, K, true, mzp, FL_Z2) && step(Z, K, false, mzm, FL_Z1));
        flags[(size_t)(x0 + tx) + (size_t)ne * ((size_t)(y0 + ty) + (size_t)ne * (z + pl))] = f;
      }
    }
    // slide the z line kMarkP planes on
    if (more) {
#pragma unroll
      for (int o = 0; o < 6; ++o) zu[o] = zu[o + kMarkP];
#pragma unroll
      for (int o = 6; o < kMarkP + 6; ++o) zu[o] = nzu[o - 6];
#pragma unroll
      for (int o = 0; o < 2; ++o) zd[o] = zd[o + kMarkP];
#pragma unroll
      for (int o = 2; o < kMarkP + 2; ++o) zd[o] = nzd[o - 2];
      if (MASK) {
#pragma unroll
        for (int o = 0; o < 2; ++o) zm[o] = zm[o + kMarkP];
#pragma unroll
        for (int o = 2; o < kMarkP + 2; ++o) zm[o] = nzm[o - 2];
      }
    }
    __syncthreads();
  }
}

constexpr int kChunk = 16;  // cells per thread in the ordered compaction

__global__ void k_count_cand(int64_t ne3, const uint8_t* __restrict__ flags, int32_t* blk) {
  __shared__ int32_t red[256];
  int64_t base = ((int64_t)blockIdx.x * 256 + threadIdx.x) * kChunk;
  int cnt = 0;
  for (int t = 0; t < kChunk; ++t) {
    int64_t e = base + t;
    if (e < ne3 && (flags[e] & FL_V)) cnt++;
  }
  red[threadIdx.x] = cnt;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) blk[blockIdx.x] = red[0];
}

__global__ void k_write_cand(int64_t ne3, ExtGeom G, const uint8_t* __restrict__ flags,
                             const int32_t* __restrict__ blkoff, const float* __restrict__ div,
                             float* key, int32_t* lin) {
  __shared__ int32_t sc[256];
  int64_t base = ((int64_t)blockIdx.x * 256 + threadIdx.x) * kChunk;
  int cnt = 0;
  for (int t = 0; t < kChunk; ++t) {
    int64_t e = base + t;
    if (e < ne3 && (flags[e] & FL_V)) cnt++;
  }
  sc[threadIdx.x] = cnt;
  __syncthreads();
  // inclusive scan (Hillis-Steele) over the block
  for (int off = 1; off < 256; off <<= 1) {
    int v = (int)threadIdx.x >= off ? sc[threadIdx.x - off] : 0;
    __syncthreads();
    sc[threadIdx.x] += v;
    __syncthreads();
  }
  int out = blkoff[blockIdx.x] + sc[threadIdx.x] - cnt;
  for (int t = 0; t < kChunk; ++t) {
    int64_t e = base + t;
    if (e < ne3 && (flags[e] & FL_V)) {
      int I = (int)(e % G.ne) + G.low1;
      int J = (int)((e / G.ne) % G.ne) + G.low1;
      int K = (int)(e / ((int64_t)G.ne * G.ne)) + G.low1;
      key[out] = div[G.inner(I, J, K)];
      lin[out] = (int32_t)e;
      out++;
    }
  }
}

// ---------------------------------------------------------------- growth
// candidate status
enum : int8_t { ST_SKIP = 0, ST_VALID = 1, ST_INVALID = 2, ST_DEFER = 3, ST_ACCEPT = 4, ST_REJECT = 5 };

// blocked test for the face cell p in direction dir (voidfinding.f90:512-520 etc.)
__device__ __forceinline__ bool blocked(const ExtGeom& G, const uint8_t* __restrict__ flags,
                                        const int32_t* __restrict__ marcap, int indv, int dir,
                                        int I, int J, int K) {
  const uint8_t bits[6] = {FL_X2, FL_X1, FL_Y2, FL_Y1, FL_Z2, FL_Z1};
  int a = I, b = J, c = K;
  switch (dir) {
    case 0: a++; break;
    case 1: a--; break;
    case 2: b++; break;
    case 3: b--; break;
    case 4: c++; break;
    default: c--; break;
  }
  uint8_t fn = flags[G.ext(a, b, c)];
  if (!(fn & FL_SUB)) return true;
  if (marcap && marcap[G.inner(a, b, c)] != indv) return true;
  return (flags[G.ext(I, J, K)] & bits[dir]) != 0;
}

// is any face cell of the half-width-s cube around (ci,cj,ck) blocked?
// nchk gains the shell's face cells (counting each one as it is examined
// made the lane kernel a third slower)
__device__ bool shell_blocked(const ExtGeom& G, const uint8_t* __restrict__ flags,
                              const int32_t* __restrict__ marcap, int indv, int ci, int cj, int ck,
                              int s, int& nchk) {
  nchk += 6 * (2 * s + 1) * (2 * s + 1);
  for (int u = -s; u <= s; ++u)
    for (int v = -s; v <= s; ++v) {
      if (blocked(G, flags, marcap, indv, 0, ci + s, cj + u, ck + v)) return true;
      if (blocked(G, flags, marcap, indv, 1, ci - s, cj + u, ck + v)) return true;
      if (blocked(G, flags, marcap, indv, 2, ci + u, cj + s, ck + v)) return true;
      if (blocked(G, flags, marcap, indv, 3, ci + u, cj - s, ck + v)) return true;
      if (blocked(G, flags, marcap, indv, 4, ci + u, cj + v, ck + s)) return true;
      if (blocked(G, flags, marcap, indv, 5, ci + u, cj + v, ck - s)) return true;
    }
  return false;
}

// growth work per candidate (cw[t]): the face cells of the shells examined
// (a lane counts a whole shell; a wave the cells its lanes examined), bit 31
// set when a wave grew it; summed only when the counters are read
constexpr uint32_t kGrownByWave = 0x80000000u;

__device__ __forceinline__ void decode(const ExtGeom& G, int e, int& I, int& J, int& K) {
  I = e % G.ne + G.low1;
  J = (e / G.ne) % G.ne + G.low1;
  K = e / (G.ne * G.ne) + G.low1;
}

// distance (in shells) from the centre to the NSEC edge zone, and whether
// reaching it is a rejection (FLAG2: some axis reaches the zone on one side
// only) or, in the degenerate symmetric case, an acceptance at that size
__device__ __forceinline__ void zone_info(const ExtGeom& G, int ci, int cj, int ck, int& szone,
                                          bool& flag2) {
  const int lo = G.low1 + kNsec, hi = G.low2 - kNsec;
  int c3[3] = {ci, cj, ck};
  szone = 0x7fffffff;
  for (int a = 0; a < 3; ++a) {
    int d1 = c3[a] - lo, d2 = hi - c3[a];
    int m = d1 < d2 ? d1 : d2;
    if (m < szone) szone = m;
  }
  flag2 = false;
  for (int a = 0; a < 3; ++a) {
    bool e1 = c3[a] - szone <= lo, e2 = c3[a] + szone >= hi;
    if (e1 != e2) flag2 = true;
  }
}

// status/r for a candidate given the first blocked shell (or >= szone)
__device__ __forceinline__ void finish_growth(int sblock, int szone, bool zflag2, int8_t& st,
                                              int16_t& r) {
  if (sblock < szone) {
    r = (int16_t)sblock;
    st = sblock >= 1 ? ST_VALID : ST_INVALID;
  } else {
    r = (int16_t)szone;
    st = zflag2 ? ST_INVALID : ST_VALID;
  }
}

// ParentRadii: REQP of the parent level, ids 1..n (null: no parent-size test)
struct ParentRadii {
  const float* reqp;
  int n;
  float rmin_sub;
};

// one candidate per lane; returns the face cells examined
__device__ int grow_lane(int64_t t, const ExtGeom& G, const int32_t* __restrict__ order,
                         const int32_t* __restrict__ lin, const uint8_t* __restrict__ flags,
                         const int32_t* __restrict__ marcap, const ParentRadii& PR,
                         const int8_t* __restrict__ prop, int smax, int wave_from, int8_t* status, int16_t* rr) {
  int nchk = 0;
  int e = lin[order[t]];
  int I, J, K;
  decode(G, e, I, J, K);
  // centre within NSEC of the edge: skipped (voidfinding.f90:460-461)
  if (I <= G.low1 + kNsec || J <= G.low1 + kNsec || K <= G.low1 + kNsec || I >= G.low2 - kNsec ||
      J >= G.low2 - kNsec || K >= G.low2 - kNsec) {
    status[t] = ST_SKIP;
    rr[t] = -1;
    return 0;
  }
  int indv = marcap ? marcap[G.inner(I, J, K)] : -1;
  // FLAG1_GRID (voidfinding.f90:267-274, 491): a parent void smaller than
  // RMIN_SUB stops the cube at its centre cell, which FLAG2 then rejects (657-659)
  if (indv > 0 && PR.reqp && indv <= PR.n && PR.reqp[indv - 1] < PR.rmin_sub) {
    status[t] = ST_INVALID;
    rr[t] = 0;
    return 0;
  }
  int szone;
  bool zf2;
  zone_info(G, I, J, K, szone, zf2);
  int s0 = 0;
  if (prop) {
    // learned proposal r_hat.  A large one goes to the wave kernel, whose
    // lanes check the faces of shells 0, 1, ... together; a small one is
    // checked here, shells 0..r_hat in one sweep, and the first blocked shell
    // found is exact because every lower shell was checked
    int rh = prop[t];
    if (rh < 0) rh = 0;
    if (rh >= wave_from) {
      status[t] = ST_DEFER;
      rr[t] = 0;
      return 0;
    }
    if (rh > szone - 1) rh = szone - 1;
    int first = -1;
    for (int s = 0; s <= rh && first < 0; ++s)
      if (shell_blocked(G, flags, marcap, indv, I, J, K, s, nchk)) first = s;
    if (first >= 0) {
      int8_t st; int16_t r;
      finish_growth(first, szone, zf2, st, r);
      status[t] = st;
      rr[t] = r;
      return nchk;
    }
    s0 = rh + 1;  // under-prediction: keep growing
  }
  for (int s = s0; s < szone; ++s) {
    if (s >= smax) {  // continue in the wave kernel
      status[t] = ST_DEFER;
      rr[t] = (int16_t)s;
      return nchk;
    }
    if (shell_blocked(G, flags, marcap, indv, I, J, K, s, nchk)) {
      int8_t st; int16_t r;
      finish_growth(s, szone, zf2, st, r);
      status[t] = st;
      rr[t] = r;
      return nchk;
    }
  }
  int8_t st; int16_t r;
  finish_growth(szone, szone, zf2, st, r);
  status[t] = st;
  rr[t] = r;
  return nchk;
}

__global__ void k_grow_lane(int64_t ncand, ExtGeom G, const int32_t* __restrict__ order,
                            const int32_t* __restrict__ lin, const uint8_t* __restrict__ flags,
                            const int32_t* __restrict__ marcap, ParentRadii PR,
                            const int8_t* __restrict__ prop, int smax, int wave_from, int8_t* status, int16_t* rr,
                            uint32_t* cw) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < ncand) cw[t] = (uint32_t)grow_lane(t, G, order, lin, flags, marcap, PR, prop, smax, wave_from, status, rr);
}

// large cubes: one wave per candidate.  The face cells of shells rr[t],
// rr[t] + 1, ... are taken in that order, 64 at a time across the lanes (a
// chunk may end one shell and start the next); the first chunk with a blocked
// cell gives the first blocked shell, that of its lowest blocked lane, since
// every cell of every lower shell has been examined.  cw[t] gains the face
// cells examined here and the grown-by-a-wave bit.
__device__ __forceinline__ int shell_cells(int s) { return 6 * (2 * s + 1) * (2 * s + 1); }

__global__ void k_grow_wave(int64_t ncand, ExtGeom G, const int32_t* __restrict__ order,
                            const int32_t* __restrict__ lin, const uint8_t* __restrict__ flags,
                            const int32_t* __restrict__ marcap, int8_t* status, int16_t* rr, uint32_t* cw) {
  int64_t t = (int64_t)blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64;
  int lane = threadIdx.x % 64;
  if (t >= ncand) return;
  if (status[t] != ST_DEFER) return;
  int e = lin[order[t]];
  int I, J, K;
  decode(G, e, I, J, K);
  int indv = marcap ? marcap[G.inner(I, J, K)] : -1;
  int szone;
  bool zf2;
  zone_info(G, I, J, K, szone, zf2);
  int sblock = szone;
  int nchk = 0;
  // chunk start: shell sb, cell ob within it
  int sb = rr[t], ob = 0;
  while (sb < szone) {
    int s = sb, o = ob + lane;
    while (s < szone && o >= shell_cells(s)) {
      o -= shell_cells(s);
      ++s;
    }
    bool b = false;
    if (s < szone) {
      ++nchk;
      const int side = 2 * s + 1, nface = side * side;
      const int dir = o / nface, f = o % nface;
      const int u = f % side - s, v = f / side - s;
      int a = I, bb = J, cc = K;
      switch (dir) {
        case 0: a = I + s; bb = J + u; cc = K + v; break;
        case 1: a = I - s; bb = J + u; cc = K + v; break;
        case 2: a = I + u; bb = J + s; cc = K + v; break;
        case 3: a = I + u; bb = J - s; cc = K + v; break;
        case 4: a = I + u; bb = J + v; cc = K + s; break;
        default: a = I + u; bb = J + v; cc = K - s; break;
      }
      b = blocked(G, flags, marcap, indv, dir, a, bb, cc);
    }
    const unsigned long long m = __ballot(b);
    if (m) {
      sblock = __shfl(s, __ffsll((long long)m) - 1);
      break;
    }
    ob += 64;
    while (sb < szone && ob >= shell_cells(sb)) {
      ob -= shell_cells(sb);
      ++sb;
    }
  }
  for (int o = 32; o > 0; o >>= 1) nchk += __shfl_xor(nchk, o);
  if (lane == 0) {
    cw[t] = (cw[t] + (uint32_t)nchk) | kGrownByWave;
    int8_t st; int16_t r;
    finish_growth(sblock, szone, zf2, st, r);
    status[t] = st;
    rr[t] = r;
  }
}

// ------------------------------------------------------------ acceptance
__global__ void k_fill_u64(int64_t n, uint64_t* a, uint64_t v) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = v;
}
__global__ void k_fill_i32(int64_t n, int32_t* a, int32_t v) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = v;
}

// one wave per cube: paint its cells
__global__ void k_paint_und(int64_t ncand, ExtGeom G, const int32_t* __restrict__ order,
                            const int32_t* __restrict__ lin, const int8_t* __restrict__ status,
                            const int16_t* __restrict__ rr, uint64_t tag, uint64_t* und) {
  int64_t t = (int64_t)blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64;
  int lane = threadIdx.x % 64;
  if (t >= ncand || status[t] != ST_VALID) return;
  int I, J, K;
  decode(G, lin[order[t]], I, J, K);
  int r = rr[t], side = 2 * r + 1;
  int vol = side * side * side;
  uint64_t key = (tag << 32) | (uint64_t)(uint32_t)t;
  for (int q = lane; q < vol; q += 64) {
    int a = q % side - r, b = (q / side) % side - r, c = q / (side * side) - r;
    atomicMin((unsigned long long*)&und[G.ext(I + a, J + b, K + c)], (unsigned long long)key);
  }
}

__global__ void k_decide(int64_t ncand, ExtGeom G, const int32_t* __restrict__ order,
                         const int32_t* __restrict__ lin, int8_t* status,
                         const uint64_t* __restrict__ und, const int32_t* __restrict__ acc,
                         uint64_t tag, int32_t* nund) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ncand || status[t] != ST_VALID) return;
  int e = lin[order[t]];
  if (acc[e] < (int32_t)t) {  // centre inside an earlier kept cube (voidfinding.f90:457)
    status[t] = ST_REJECT;
    return;
  }
  uint64_t u = und[e];
  if ((u >> 32) == tag && (uint32_t)u == (uint32_t)t) {
    status[t] = ST_ACCEPT + 10;  // newly accepted (painted next)
    return;
  }
  atomicAdd(nund, 1);
}

__global__ void k_paint_acc(int64_t ncand, ExtGeom G, const int32_t* __restrict__ order,
                            const int32_t* __restrict__ lin, int8_t* status,
                            const int16_t* __restrict__ rr, int32_t* acc) {
  int64_t t = (int64_t)blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64;
  int lane = threadIdx.x % 64;
  if (t >= ncand || status[t] != ST_ACCEPT + 10) return;
  int I, J, K;
  decode(G, lin[order[t]], I, J, K);
  int r = rr[t], side = 2 * r + 1;
  int vol = side * side * side;
  for (int q = lane; q < vol; q += 64) {
    int a = q % side - r, b = (q / side) % side - r, c = q / (side * side) - r;
    atomicMin(&acc[G.ext(I + a, J + b, K + c)], (int32_t)t);
  }
  // all lanes have read the status before lane 0 rewrites it
  __builtin_amdgcn_wave_barrier();
  if (lane == 0) status[t] = ST_ACCEPT;
}

__global__ void k_accept_flags(int64_t ncand, const int8_t* __restrict__ status, int32_t* f) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < ncand) f[t] = status[t] == ST_ACCEPT ? 1 : 0;
}

// cube records (voidfinding.f90:466-496, 681-696); id = rank among accepted
__global__ void k_write_cubes(int64_t ncand, ExtGeom G, const int32_t* __restrict__ order,
                              const int32_t* __restrict__ lin, const int8_t* __restrict__ status,
                              const int16_t* __restrict__ rr, const int32_t* __restrict__ idx,
                              const int32_t* __restrict__ marcap, int32_t nv, float dx, float rx1,
                              int32_t* ci, float* cf, int32_t* father) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ncand || status[t] != ST_ACCEPT) return;
  int v = idx[t];  // 0-based cube index
  int I, J, K;
  decode(G, lin[order[t]], I, J, K);
  int r = rr[t];
  int ini[3] = {I - r, J - r, K - r}, fin[3] = {I + r, J + r, K + r};
  for (int a = 0; a < 3; ++a) {
    ci[a * nv + v] = ini[a];
    ci[(3 + a) * nv + v] = fin[a];
  }
  ci[6 * nv + v] = I;
  ci[7 * nv + v] = J;
  ci[8 * nv + v] = K;
  for (int a = 0; a < 3; ++a) {
    cf[a * nv + v] = (rx1 + (float)(ini[a] - 1) * dx) - 0.5f * dx;
    cf[(3 + a) * nv + v] = (rx1 + (float)(fin[a] - 1) * dx) + 0.5f * dx;
  }
  float A = (float)(2 * r + 1) * dx;
  cf[6 * nv + v] = (A * A) * A;
  father[v] = marcap ? marcap[G.inner(I, J, K)] : -1;
}

// largest id in the parent map (grid-stride; one atomic per wave)
__global__ void k_max_id(int64_t n, const int32_t* __restrict__ a, int32_t* out) {
  int m = 0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    m = max(m, a[i]);
  for (int off = 32; off > 0; off >>= 1) m = max(m, __shfl_xor(m, off));
  if (threadIdx.x % 64 == 0) atomicMax(out, m);
}

// processing order t -> compacted candidate (largest divergence first)
__global__ void k_reverse_order(int64_t n, const int32_t* __restrict__ indx, int32_t* ord) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) ord[t] = indx[n - 1 - t] - 1;
}

inline unsigned nbk(int64_t n, int b = 256) { return (unsigned)((n + b - 1) / b); }
}  // namespace

int void_find(avism_ctx* c, int32_t growth_mode) {
  if (!c->fields_ready) AV_FAIL(c, AVISM_ERR_STATE, "fields not computed");
  c->post_ready = false;
  const GridGeom& g = c->g;
  ExtGeom G{g.n, g.low1, g.low2, g.next, g.periodic};
  const int64_t ne3 = (int64_t)g.next * g.next * g.next;
  MarkParams P;
  P.dthr1 = c->vp.dens_thre + 1.0f;   // voidfinding.f90:41
  P.dthr2 = c->vp.dens_thre2 + 1.0f;  // voidfinding.f90:298
  P.grad = c->vp.grad_thre;
  P.divthr = c->vp.div_thre;
  P.two_dx = 2.0f * g.dx;  // exact; see voidfinding.f90:290
  const int32_t* marcap = c->has_marcap ? c->marcap.as<int32_t>() : nullptr;
  if (marcap && c->has_reqp) {  // every parent id must have its REQP
    const int64_t n3 = (int64_t)g.n * g.n * g.n;
    int32_t* mx = (int32_t*)((char*)c->scalars.ensure(4096) + 768);
    AV_CHECK_HIP(c, hipMemsetAsync(mx, 0, 4, c->stream));
    k_max_id<<<(unsigned)std::min<int64_t>(nbk(n3), 4096), 256, 0, c->stream>>>(n3, marcap, mx);
    AV_CHECK_HIP(c, hipGetLastError());
    int32_t h = 0;
    AV_CHECK_HIP(c, hipMemcpyAsync(&h, mx, 4, hipMemcpyDeviceToHost, c->stream));
    AV_CHECK_HIP(c, hipStreamSynchronize(c->stream));
    if (h > c->nreqp) AV_FAIL(c, AVISM_ERR_INPUT, "the parent map holds void ids above the number of parent radii");
  }
  if (!c->flags.ensure((size_t)ne3)) AV_FAIL(c, AVISM_ERR_HIP, "out of device memory (flags)");
  {
    StageTimer t(c, "mark_flags");
    const dim3 gm((unsigned)((G.ne + kMarkX - 1) / kMarkX), (unsigned)((G.ne + kMarkY - 1) / kMarkY),
                  (unsigned)((G.ne + kMarkZ - 1) / kMarkZ));
    if (c->has_mask)
      k_mark<1><<<gm, kMarkT, 0, c->stream>>>(G, P, c->u1co.as<float>(), c->div.as<float>(), marcap,
                                              c->smask.as<int8_t>(), c->flags.as<uint8_t>());
    else
      k_mark<0><<<gm, kMarkT, 0, c->stream>>>(G, P, c->u1co.as<float>(), c->div.as<float>(), marcap, nullptr,
                                              c->flags.as<uint8_t>());
    AV_CHECK_HIP(c, hipGetLastError());
  }
  // ordered compaction
  int64_t nblocks = (ne3 + 256 * kChunk - 1) / (256 * kChunk);
  if (!c->blk_cnt.ensure((size_t)(nblocks + 1) * 8)) AV_FAIL(c, AVISM_ERR_HIP, "out of device memory");
  int32_t* blk = c->blk_cnt.as<int32_t>();
  int32_t* blkoff = blk + nblocks + 1;
  {
    StageTimer t(c, "candidates");
    AV_CHECK_HIP(c, hipMemsetAsync(blk, 0, (size_t)(nblocks + 1) * 4, c->stream));
    k_count_cand<<<(unsigned)nblocks, 256, 0, c->stream>>>(ne3, c->flags.as<uint8_t>(), blk);
    size_t tmp = 0;
    AV_CHECK_HIP(c, rocprim::exclusive_scan(nullptr, tmp, blk, blkoff, 0, (size_t)nblocks + 1,
                                            rocprim::plus<int32_t>(), c->stream));
    if (!c->sort_tmp.ensure(tmp)) AV_FAIL(c, AVISM_ERR_HIP, "out of device memory");
    AV_CHECK_HIP(c, rocprim::exclusive_scan(c->sort_tmp.p, tmp, blk, blkoff, 0, (size_t)nblocks + 1,
                                            rocprim::plus<int32_t>(), c->stream));
    int32_t hn = 0;
    AV_CHECK_HIP(c, hipMemcpyAsync(&hn, blkoff + nblocks, 4, hipMemcpyDeviceToHost, c->stream));
    AV_CHECK_HIP(c, hipStreamSynchronize(c->stream));
    c->ncand = hn;
    if (!c->cand_key.ensure((size_t)(hn + 1) * 4) || !c->cand_lin.ensure((size_t)(hn + 1) * 4))
      AV_FAIL(c, AVISM_ERR_HIP, "out of device memory (candidates)");
    k_write_cand<<<(unsigned)nblocks, 256, 0, c->stream>>>(ne3, G, c->flags.as<uint8_t>(), blkoff,
                                                           c->div.as<float>(), c->cand_key.as<float>(),
                                                           c->cand_lin.as<int32_t>());
    AV_CHECK_HIP(c, hipGetLastError());
  }
  const int64_t nc = c->ncand;
  c->nvoid = 0;
  if (!c->sort_idx.ensure((size_t)(nc + 1) * 4) || !c->cand_r.ensure((size_t)(nc + 1) * 2) ||
      !c->cand_state.ensure((size_t)(nc + 1)) || !c->acc_rank.ensure((size_t)(nc + 2) * 4) ||
      !c->misc_i32.ensure((size_t)(nc + 2) * 4))
    AV_FAIL(c, AVISM_ERR_HIP, "out of device memory (growth)");
  if (nc == 0) return AVISM_OK;
  // INDEXX ascending, then reversed (voidfinding.f90:185-201)
  int rc = indexx_device(c, c->cand_key.as<float>(), nc, c->sort_idx.as<int32_t>());
  if (rc) return rc;
  int32_t* order = c->misc_i32.as<int32_t>();  // processing order -> compacted index
  k_reverse_order<<<nbk(nc), 256, 0, c->stream>>>(nc, c->sort_idx.as<int32_t>(), order);
  AV_CHECK_HIP(c, hipGetLastError());
  int8_t* status = c->cand_state.as<int8_t>();
  int16_t* rr = c->cand_r.as<int16_t>();
  // growth work per candidate (summed by avism_get_counters)
  if (!c->cand_work.ensure((size_t)(nc + 1) * 4)) AV_FAIL(c, AVISM_ERR_HIP, "out of device memory (growth)");
  uint32_t* cw = c->cand_work.as<uint32_t>();
  c->growth_pending = true;
  const int8_t* prop = nullptr;
  if (growth_mode == AVISM_GROW_LEARNED) {
    rc = cnn_propose(c);
    if (rc) return rc;
    prop = c->proposals.as<int8_t>();
  } else if (growth_mode == AVISM_GROW_PROPOSAL) {
    if (c->proposals.bytes < (size_t)nc) AV_FAIL(c, AVISM_ERR_STATE, "no proposals set for these candidates");
    prop = c->proposals.as<int8_t>();
  } else if (growth_mode != AVISM_GROW_PLAIN) {
    AV_FAIL(c, AVISM_ERR_ARG, "unknown growth mode");
  }
  {
    StageTimer t(c, prop ? "grow_proposal" : "grow");
    ParentRadii PR{c->has_reqp && marcap ? c->reqp.as<float>() : nullptr, c->nreqp, c->vp.rmin_sub};
    // plain: shells 0..3 in a lane, larger cubes continue in the wave kernel;
    // with proposals, a candidate proposed at 4 or more goes to the wave
    // kernel from shell 0 (routing from 2 up sent 59% of the candidates of a
    // 512^3 Zel'dovich box to waves and took 1.9 ms against 0.5)
    k_grow_lane<<<nbk(nc), 256, 0, c->stream>>>(nc, G, order, c->cand_lin.as<int32_t>(),
                                                c->flags.as<uint8_t>(), marcap, PR, prop, 4, 4, status, rr, cw);
    k_grow_wave<<<nbk(nc, 4), 256, 0, c->stream>>>(nc, G, order, c->cand_lin.as<int32_t>(),
                                                   c->flags.as<uint8_t>(), marcap, status, rr, cw);
    AV_CHECK_HIP(c, hipGetLastError());
  }
  // acceptance rounds
  if (!c->cover_und.ensure((size_t)ne3 * 8) || !c->cover_acc.ensure((size_t)ne3 * 4))
    AV_FAIL(c, AVISM_ERR_HIP, "out of device memory (acceptance)");
  uint64_t* und = c->cover_und.as<uint64_t>();
  int32_t* acc = c->cover_acc.as<int32_t>();
  int32_t* nund = (int32_t*)((char*)c->scalars.ensure(4096) + 512);
  int rounds = 0;
  {
    StageTimer t(c, "accept");
    k_fill_u64<<<nbk(ne3), 256, 0, c->stream>>>(ne3, und, ~0ull);
    k_fill_i32<<<nbk(ne3), 256, 0, c->stream>>>(ne3, acc, 0x7fffffff);
    for (;;) {
      ++rounds;
      uint64_t tag = 0xffffffffull - (uint64_t)rounds;  // later rounds win the atomicMin
      k_paint_und<<<nbk(nc, 4), 256, 0, c->stream>>>(nc, G, order, c->cand_lin.as<int32_t>(), status,
                                                     rr, tag, und);
      AV_CHECK_HIP(c, hipMemsetAsync(nund, 0, 4, c->stream));
      k_decide<<<nbk(nc), 256, 0, c->stream>>>(nc, G, order, c->cand_lin.as<int32_t>(), status, und,
                                               acc, tag, nund);
      k_paint_acc<<<nbk(nc, 4), 256, 0, c->stream>>>(nc, G, order, c->cand_lin.as<int32_t>(), status,
                                                     rr, acc);
      AV_CHECK_HIP(c, hipGetLastError());
      int32_t h = 0;
      AV_CHECK_HIP(c, hipMemcpyAsync(&h, nund, 4, hipMemcpyDeviceToHost, c->stream));
      AV_CHECK_HIP(c, hipStreamSynchronize(c->stream));
      if (h == 0) break;
      if (rounds > 100000) AV_FAIL(c, AVISM_ERR_LIMIT, "acceptance did not converge");
    }
  }
  c->growth_stats[0] = rounds;
  // ids and cube records
  int32_t* accf = c->acc_rank.as<int32_t>();
  int32_t* accx = accf + 0;  // reuse: flags then in-place scan into misc buffer
  k_accept_flags<<<nbk(nc), 256, 0, c->stream>>>(nc, status, accf);
  int32_t* idx = (int32_t*)c->cand_key.p;  // keys no longer needed
  size_t tmp = 0;
  AV_CHECK_HIP(c, rocprim::exclusive_scan(nullptr, tmp, accx, idx, 0, (size_t)nc + 1,
                                          rocprim::plus<int32_t>(), c->stream));
  if (!c->sort_tmp.ensure(tmp)) AV_FAIL(c, AVISM_ERR_HIP, "out of device memory");
  AV_CHECK_HIP(c, hipMemsetAsync(accf + nc, 0, 4, c->stream));
  AV_CHECK_HIP(c, rocprim::exclusive_scan(c->sort_tmp.p, tmp, accx, idx, 0, (size_t)nc + 1,
                                          rocprim::plus<int32_t>(), c->stream));
  int32_t nv = 0;
  AV_CHECK_HIP(c, hipMemcpyAsync(&nv, idx + nc, 4, hipMemcpyDeviceToHost, c->stream));
  AV_CHECK_HIP(c, hipStreamSynchronize(c->stream));
  if (nv > c->vp.nvoid_max) AV_FAIL(c, AVISM_ERR_LIMIT, "NVOID > NVOID_MAX!! Increase NVOID_MAX");
  c->nvoid = nv;
  if (!c->cube_i.ensure((size_t)9 * (nv + 1) * 4) || !c->cube_f.ensure((size_t)7 * (nv + 1) * 4) ||
      !c->cube_father.ensure((size_t)(nv + 1) * 4))
    AV_FAIL(c, AVISM_ERR_HIP, "out of device memory (cubes)");
  k_write_cubes<<<nbk(nc), 256, 0, c->stream>>>(nc, G, order, c->cand_lin.as<int32_t>(), status, rr, idx,
                                                marcap, nv, g.dx, g.rx1, c->cube_i.as<int32_t>(),
                                                c->cube_f.as<float>(), c->cube_father.as<int32_t>());
  AV_CHECK_HIP(c, hipGetLastError());
  AV_CHECK_HIP(c, hipStreamSynchronize(c->stream));
  return AVISM_OK;
}

namespace {
__global__ void k_sum_growth(int64_t n, const uint32_t* __restrict__ cw, unsigned long long* out) {
  unsigned long long cells = 0, waves = 0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const uint32_t v = cw[i];
    cells += v & ~kGrownByWave;
    waves += v >> 31;
  }
  for (int o = 32; o > 0; o >>= 1) {
    cells += __shfl_xor(cells, o);
    waves += __shfl_xor(waves, o);
  }
  if ((threadIdx.x & 63) == 0) {
    atomicAdd(out, cells);
    atomicAdd(out + 1, waves);
  }
}
}  // namespace

// counters[3], counters[4] from the per-candidate growth work of the last
// avism_find_voids (once per growth)
int growth_counters(avism_ctx* c) {
  if (!c->growth_pending) return AVISM_OK;
  c->growth_pending = false;
  c->counters[3] = c->counters[4] = 0;
  if (c->ncand <= 0) return AVISM_OK;
  unsigned long long* d = (unsigned long long*)((char*)c->scalars.ensure(4096) + 3584);
  AV_CHECK_HIP(c, hipMemsetAsync(d, 0, 16, c->stream));
  k_sum_growth<<<64, 256, 0, c->stream>>>(c->ncand, c->cand_work.as<uint32_t>(), d);
  AV_CHECK_HIP(c, hipGetLastError());
  unsigned long long h[2] = {0, 0};
  AV_CHECK_HIP(c, hipMemcpyAsync(h, d, 16, hipMemcpyDeviceToHost, c->stream));
  AV_CHECK_HIP(c, hipStreamSynchronize(c->stream));
  c->counters[3] = (int64_t)h[0];
  c->counters[4] = (int64_t)h[1];
  return AVISM_OK;
}

int set_proposals(avism_ctx* c, const int8_t* r, int64_t n, int where) {
  if (n != c->ncand) AV_FAIL(c, AVISM_ERR_ARG, "proposal count differs from the candidate count");
  if (!c->proposals.ensure((size_t)n + 1)) AV_FAIL(c, AVISM_ERR_HIP, "out of device memory");
  if (n > 0)
    AV_CHECK_HIP(c, hipMemcpy(c->proposals.p, r, (size_t)n,
                              where == AVISM_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice));
  return AVISM_OK;
}

int get_candidates(avism_ctx* c, int64_t* lin, int32_t* r, int64_t cap, int64_t* n) {
  if (n) *n = c->ncand;
  if (!lin && !r) return AVISM_OK;
  int64_t m = c->ncand < cap ? c->ncand : cap;
  if (m <= 0) return AVISM_OK;
  std::vector<int32_t> ord((size_t)m), l32((size_t)c->ncand);
  std::vector<int16_t> rr((size_t)m);
  std::vector<int8_t> st((size_t)m);
  AV_CHECK_HIP(c, hipMemcpy(ord.data(), c->misc_i32.p, (size_t)m * 4, hipMemcpyDeviceToHost));
  AV_CHECK_HIP(c, hipMemcpy(l32.data(), c->cand_lin.p, (size_t)c->ncand * 4, hipMemcpyDeviceToHost));
  AV_CHECK_HIP(c, hipMemcpy(rr.data(), c->cand_r.p, (size_t)m * 2, hipMemcpyDeviceToHost));
  AV_CHECK_HIP(c, hipMemcpy(st.data(), c->cand_state.p, (size_t)m, hipMemcpyDeviceToHost));
  for (int64_t t = 0; t < m; ++t) {
    if (lin) lin[t] = l32[(size_t)ord[(size_t)t]];
    if (r) r[t] = st[(size_t)t] == ST_SKIP ? -1 : rr[(size_t)t];
  }
  return AVISM_OK;
}

int get_cubes(avism_ctx* c, int32_t* ci, float* cf, int32_t* father) {
  int nv = c->nvoid;
  if (nv == 0) return AVISM_OK;
  if (ci) AV_CHECK_HIP(c, hipMemcpy(ci, c->cube_i.p, (size_t)9 * nv * 4, hipMemcpyDeviceToHost));
  if (cf) AV_CHECK_HIP(c, hipMemcpy(cf, c->cube_f.p, (size_t)7 * nv * 4, hipMemcpyDeviceToHost));
  if (father) AV_CHECK_HIP(c, hipMemcpy(father, c->cube_father.p, (size_t)nv * 4, hipMemcpyDeviceToHost));
  return AVISM_OK;
}

}  // namespace avism
