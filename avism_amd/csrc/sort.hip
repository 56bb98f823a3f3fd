// sort.hip -- INDEXX (num.f90:2-84) on the GPU, reproducing its permutation.
//
// The reference sorts candidate divergences (voidfinding.f90:185) and cube
// volumes (avism.f90:1738) with Numerical Recipes' indexx.  It is not stable,
// and its order of equal keys decides cube and void ids, so a generic GPU sort
// cannot be used.  Instead the quicksort is replayed level by level: all
// segments of one recursion depth are partitioned at once.
//
// Partition of segment [l, r] (1-based, r-l >= M=7), after the reference's
// median-of-three set-up (pivot a = arr(indx(l)), arr(indx(l+1)) <= a <= arr(indx(r))):
//   G = positions p in [l+2, r]   with arr >= a, ascending   (left-scan stops)
//   S = positions p in [l+1, r-1] with arr <= a, descending  (right-scan stops)
// The k-th exchange swaps G_k and S_k while G_k < S_k (both lists are in
// untouched territory until they cross), so K = #{k : G_k < S_k} exchanges
// happen, after which the scans stop at
//   i' = min(G_{K+1}, S_K),  j' = max(S_{K+1}, G_K)
// and, when i' == j' (an element equal to the pivot), one more empty round
// leaves i = p+1, j = p-1.  The pivot then moves to j and the children are
// [l, j-1] and [i, r].  G/S ranks come from prefix sums over the whole array,
// so each level is a handful of element-parallel kernels.  Segments of at most
// M+1 elements get the reference's insertion sort (num.f90:15-30).
#include <cstring>
#include <mutex>

#include <rocprim/rocprim.hpp>

#include "avism_internal.h"

namespace avism {

namespace {
constexpr int kM = 7;

struct SegTab {
  int32_t* l;
  int32_t* r;
  int32_t* pid;    // pivot id (1-based key index)
  float* piv;      // pivot key
  int32_t* kk;     // number of exchanges
  int32_t* iff;    // final i
  int32_t* jff;    // final j
  int32_t* cl;     // child segment ids (-1 none)
  int32_t* cr;
  int32_t* small;  // insertion-sorted at this level
};

__global__ void k_init_indx(int64_t n, int32_t* indx, int32_t* segid) {
  int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  indx[q] = (int32_t)(q + 1);
  segid[q] = 0;
}

// per segment: insertion sort (small) or median-of-three set-up (num.f90:33-48)
__device__ void seg_setup(int s, SegTab T, const float* __restrict__ arr, int32_t* indx) {
  int l = T.l[s], r = T.r[s];
  auto IX = [&](int p) -> int32_t& { return indx[p - 1]; };
  auto A = [&](int32_t id) -> float { return arr[id - 1]; };
  if (r - l < kM) {
    for (int j = l + 1; j <= r; ++j) {
      int32_t it = IX(j);
      float a = A(it);
      int i;
      for (i = j - 1; i >= l; --i) {
        if (A(IX(i)) <= a) break;
        IX(i + 1) = IX(i);
      }
      IX(i + 1) = it;
    }
    T.small[s] = 1;
    return;
  }
  T.small[s] = 0;
  int k = (l + r) / 2;
  int32_t t = IX(k); IX(k) = IX(l + 1); IX(l + 1) = t;
  if (A(IX(l + 1)) > A(IX(r))) { t = IX(l + 1); IX(l + 1) = IX(r); IX(r) = t; }
  if (A(IX(l)) > A(IX(r))) { t = IX(l); IX(l) = IX(r); IX(r) = t; }
  if (A(IX(l + 1)) > A(IX(l))) { t = IX(l + 1); IX(l + 1) = IX(l); IX(l) = t; }
  T.pid[s] = IX(l);
  T.piv[s] = A(IX(l));
}

__global__ void k_seg_setup(const int32_t* __restrict__ nsegp, SegTab T, const float* __restrict__ arr,
                            int32_t* indx) {
  const int nseg = *nsegp;
  for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < nseg; s += gridDim.x * blockDim.x) seg_setup(s, T, arr, indx);
}

// g/s flags per position
__global__ void k_flags(int64_t n, SegTab T, const int32_t* __restrict__ segid,
                        const float* __restrict__ arr, const int32_t* __restrict__ indx,
                        int32_t* gf, int32_t* sf) {
  int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q > n) return;
  int g = 0, sflag = 0;
  if (q < n) {
    int s = segid[q];
    if (s >= 0 && !T.small[s]) {
      int p = (int)q + 1, l = T.l[s], r = T.r[s];
      float a = T.piv[s];
      float v = arr[indx[q] - 1];
      g = (p >= l + 2 && p <= r && v >= a) ? 1 : 0;
      sflag = (p >= l + 1 && p <= r - 1 && v <= a) ? 1 : 0;
    }
  }
  gf[q] = g;
  sf[q] = sflag;
}

__global__ void k_scatter_gs(int64_t n, SegTab T, const int32_t* __restrict__ segid,
                             const int32_t* __restrict__ gf, const int32_t* __restrict__ sf,
                             const int32_t* __restrict__ gx, const int32_t* __restrict__ sx,
                             int32_t* gpos, int32_t* spos) {
  int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  int s = segid[q];
  if (s < 0 || T.small[s]) return;
  int p = (int)q + 1, l = T.l[s], r = T.r[s];
  int base = l - 1;
  if (gf[q]) {
    int rank = gx[q] - gx[l - 1] + 1;
    gpos[base + rank - 1] = p;
  }
  if (sf[q]) {
    int rank = sx[r - 1] - sx[q];
    spos[base + rank - 1] = p;
  }
}

__device__ void seg_partition(int s, SegTab T, const int32_t* __restrict__ gx, const int32_t* __restrict__ sx,
                              const int32_t* __restrict__ gpos, const int32_t* __restrict__ spos) {
  int l = T.l[s], r = T.r[s], base = l - 1;
  int ng = gx[r] - gx[l - 1];
  int ns = sx[r] - sx[l - 1];
  int m = ng < ns ? ng : ns;
  // K = number of leading k with G_k < S_k (monotone predicate)
  int lo = 0, hi = m;
  while (lo < hi) {
    int mid = (lo + hi + 1) / 2;
    if (gpos[base + mid - 1] < spos[base + mid - 1]) lo = mid;
    else hi = mid - 1;
  }
  int K = lo;
  const int BIG = 0x7fffffff;
  int gk1 = (K + 1 <= ng) ? gpos[base + K] : BIG;
  int sk = (K >= 1) ? spos[base + K - 1] : BIG;
  int sk1 = (K + 1 <= ns) ? spos[base + K] : -BIG;
  int gk = (K >= 1) ? gpos[base + K - 1] : -BIG;
  int ip = gk1 < sk ? gk1 : sk;
  int jp = sk1 > gk ? sk1 : gk;
  int i_f = ip, j_f = jp;
  if (jp >= ip) {  // equal element at p = ip: one more empty exchange
    i_f = ip + 1;
    j_f = ip - 1;
  }
  T.kk[s] = K;
  T.iff[s] = i_f;
  T.jff[s] = j_f;
}

__global__ void k_seg_partition(const int32_t* __restrict__ nsegp, SegTab T, const int32_t* __restrict__ gx,
                                const int32_t* __restrict__ sx, const int32_t* __restrict__ gpos,
                                const int32_t* __restrict__ spos) {
  const int nseg = *nsegp;
  for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < nseg; s += gridDim.x * blockDim.x) {
    if (!T.small[s]) seg_partition(s, T, gx, sx, gpos, spos);
  }
}

__global__ void k_exchange(int64_t n, SegTab T, const int32_t* __restrict__ segid,
                           const int32_t* __restrict__ gf, const int32_t* __restrict__ gx,
                           const int32_t* __restrict__ spos, int32_t* indx) {
  int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n || !gf[q]) return;
  int s = segid[q];
  int l = T.l[s];
  int rank = gx[q] - gx[l - 1] + 1;
  if (rank > T.kk[s]) return;
  int partner = spos[l - 1 + rank - 1];
  int32_t a = indx[q], b = indx[partner - 1];
  indx[q] = b;
  indx[partner - 1] = a;
}

__device__ void seg_finish(int s, SegTab T, SegTab N, int32_t* indx, int32_t* counter) {
  T.cl[s] = -1;
  T.cr[s] = -1;
  if (T.small[s]) return;
  int l = T.l[s], r = T.r[s], j = T.jff[s], i = T.iff[s];
  indx[l - 1] = indx[j - 1];  // num.f90:60-61
  indx[j - 1] = T.pid[s];
  if (j - 1 - l + 1 >= 2) {
    int id = atomicAdd(counter, 1);
    N.l[id] = l;
    N.r[id] = j - 1;
    T.cl[s] = id;
  }
  if (r - i + 1 >= 2) {
    int id = atomicAdd(counter, 1);
    N.l[id] = i;
    N.r[id] = r;
    T.cr[s] = id;
  }
}

__global__ void k_seg_finish(const int32_t* __restrict__ nsegp, SegTab T, SegTab N, int32_t* indx,
                             int32_t* counter) {
  const int nseg = *nsegp;
  for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < nseg; s += gridDim.x * blockDim.x)
    seg_finish(s, T, N, indx, counter);
}

__global__ void k_relabel(int64_t n, SegTab T, int32_t* segid) {
  int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  int s = segid[q];
  if (s < 0) return;
  if (T.small[s]) {
    segid[q] = -1;
    return;
  }
  int p = (int)q + 1;
  int l = T.l[s], r = T.r[s], j = T.jff[s], i = T.iff[s];
  if (p >= l && p <= j - 1) segid[q] = T.cl[s];
  else if (p >= i && p <= r) segid[q] = T.cr[s];
  else segid[q] = -1;
}

inline unsigned nb(int64_t n, int b = 256) { return (unsigned)((n + b - 1) / b); }

// the level counters and the first segment [1, n]
__global__ void k_seg_init(int64_t n, int32_t* l, int32_t* r, int32_t* counter) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  l[0] = 1;
  r[0] = (int32_t)n;
  counter[0] = 1;
  counter[1] = 0;
}

// scratch layout for n keys: segid, gf, sf, gx, sx, gpos, spos (n+1 each),
// two segment tables, the two level counters, then rocPRIM's scan state
size_t indexx_tables_bytes(int64_t n) {
  const size_t N1 = (size_t)n + 1, tab_entries = (size_t)n / 2 + 2;
  return ((7 * N1 * 4 + 2 * 10 * tab_entries * 4 + 64) + 255) & ~(size_t)255;
}

}  // namespace

// INDEXX of d_keys[0..n) into d_indx (1-based) on the context stream, with
// all scratch in [scratch, scratch + bytes): the tables, then tmpb bytes of
// scan state.  Allocates nothing.
static int indexx_core(avism_ctx* c, const float* d_keys, int64_t n, int32_t* d_indx, char* scratch, size_t bytes,
                       size_t tmpb) {
  const size_t N1 = (size_t)n + 1;
  const size_t tab_entries = (size_t)n / 2 + 2;
  if (indexx_tables_bytes(n) + tmpb > bytes) AV_FAIL(c, AVISM_ERR_LIMIT, "indexx scratch too small");
  int32_t* segid = (int32_t*)scratch;
  int32_t* gf = segid + N1;
  int32_t* sf = gf + N1;
  int32_t* gx = sf + N1;
  int32_t* sx = gx + N1;
  int32_t* gpos = sx + N1;
  int32_t* spos = gpos + N1;
  int32_t* tabs = spos + N1;
  void* tmp = scratch + indexx_tables_bytes(n);
  auto mk = [&](int32_t* b) {
    SegTab t;
    size_t e = tab_entries;
    t.l = b; t.r = b + e; t.pid = b + 2 * e; t.piv = (float*)(b + 3 * e); t.kk = b + 4 * e;
    t.iff = b + 5 * e; t.jff = b + 6 * e; t.cl = b + 7 * e; t.cr = b + 8 * e; t.small = b + 9 * e;
    return t;
  };
  SegTab A = mk(tabs), B = mk(tabs + 10 * tab_entries);
  int32_t* counter = (int32_t*)(tabs + 20 * tab_entries);
  k_init_indx<<<nb(n), 256, 0, c->stream>>>(n, d_indx, segid);
  k_seg_init<<<1, 64, 0, c->stream>>>(n, A.l, A.r, counter);
  // levels run in batches with the segment count kept on the device; the
  // host checks it between batches (a level with no segment does nothing)
  constexpr int kBatch = 4;
  const unsigned gseg = nb((int64_t)tab_entries) < 1024 ? nb((int64_t)tab_entries) : 1024;
  int levels = 0;
  for (;;) {
    for (int b = 0; b < kBatch; ++b, ++levels) {
      int32_t* cur = counter + (levels & 1);
      int32_t* nxt = counter + ((levels + 1) & 1);
      k_seg_setup<<<gseg, 256, 0, c->stream>>>(cur, A, d_keys, d_indx);
      k_flags<<<nb((int64_t)N1), 256, 0, c->stream>>>(n, A, segid, d_keys, d_indx, gf, sf);
      AV_CHECK_HIP(c, rocprim::exclusive_scan(tmp, tmpb, gf, gx, 0, N1, rocprim::plus<int32_t>(), c->stream));
      AV_CHECK_HIP(c, rocprim::exclusive_scan(tmp, tmpb, sf, sx, 0, N1, rocprim::plus<int32_t>(), c->stream));
      k_scatter_gs<<<nb(n), 256, 0, c->stream>>>(n, A, segid, gf, sf, gx, sx, gpos, spos);
      k_seg_partition<<<gseg, 256, 0, c->stream>>>(cur, A, gx, sx, gpos, spos);
      k_exchange<<<nb(n), 256, 0, c->stream>>>(n, A, segid, gf, gx, spos, d_indx);
      AV_CHECK_HIP(c, hipMemsetAsync(nxt, 0, 4, c->stream));
      k_seg_finish<<<gseg, 256, 0, c->stream>>>(cur, A, B, d_indx, nxt);
      k_relabel<<<nb(n), 256, 0, c->stream>>>(n, A, segid);
      AV_CHECK_HIP(c, hipGetLastError());
      std::swap(A, B);
    }
    int32_t left = 0;
    AV_CHECK_HIP(c, hipMemcpyAsync(&left, counter + (levels & 1), 4, hipMemcpyDeviceToHost, c->stream));
    AV_CHECK_HIP(c, hipStreamSynchronize(c->stream));
    if (left == 0) break;
    if (levels > 4096) AV_FAIL(c, AVISM_ERR_LIMIT, "indexx did not converge");
  }
  return AVISM_OK;
}

static size_t scan_state_bytes(int64_t n, hipStream_t s) {
  size_t tmpb = 0;
  if (rocprim::exclusive_scan(nullptr, tmpb, (int32_t*)nullptr, (int32_t*)nullptr, 0, (size_t)n + 1,
                              rocprim::plus<int32_t>(), s) != hipSuccess)
    return ~(size_t)0;
  return (tmpb + 255) & ~(size_t)255;
}

int indexx_device(avism_ctx* c, const float* d_keys, int64_t n, int32_t* d_indx) {
  StageTimer tm(c, "indexx");
  if (n <= 0) return AVISM_OK;
  const size_t tmpb = scan_state_bytes(n, c->stream);
  if (tmpb == ~(size_t)0) AV_FAIL(c, AVISM_ERR_HIP, "rocprim scan size query failed");
  const size_t need = indexx_tables_bytes(n) + tmpb;
  char* base = (char*)c->sort_misc.ensure(need);
  if (!base) AV_FAIL(c, AVISM_ERR_HIP, "out of device memory (indexx)");
  return indexx_core(c, d_keys, n, d_indx, base, need, tmpb);
}

// ---- INDEXX without allocation (callers that cannot allocate device memory)
// The scratch is a file-scope device array: every GPU holds its own copy for
// the life of the process, found by its per-device address (the array's name
// on the host is only a stand-in).  Calls that overlap on one GPU race on it.
constexpr int64_t kNoAllocMaxKeys = (int64_t)1 << 19;
constexpr size_t kNoAllocScratch = ((7 * ((size_t)kNoAllocMaxKeys + 1) * 4 +
                                     2 * 10 * ((size_t)kNoAllocMaxKeys / 2 + 2) * 4 + 64 + 255) & ~(size_t)255) +
                                   ((size_t)1 << 20);  // + scan state
static __device__ __align__(256) char scratch[kNoAllocScratch];
constexpr int kMaxDevices = 64;
static char* scratch_ptr[kMaxDevices];  // per device, set by noalloc_init
static std::mutex scratch_mu;

// once per device, with that device current (avism_create)
int noalloc_init(int device) {
  if (device < 0 || device >= kMaxDevices) return AVISM_ERR_ARG;
  std::lock_guard<std::mutex> lk(scratch_mu);
  if (scratch_ptr[device]) return AVISM_OK;
  void* p = nullptr;
  if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return AVISM_ERR_HIP;
  scratch_ptr[device] = (char*)p;
  return AVISM_OK;
}

int indexx_noalloc(avism_ctx* c, const float* d_keys, int64_t n, int32_t* d_indx) {
  if (n < 0) AV_FAIL(c, AVISM_ERR_ARG, "negative length");
  if (n > kNoAllocMaxKeys)
    AV_FAIL(c, AVISM_ERR_LIMIT, "avism_indexx_noalloc takes at most 524288 keys (its static scratch)");
  if (n == 0) return AVISM_OK;
  char* base = nullptr;
  if (c->device >= 0 && c->device < kMaxDevices) {
    std::lock_guard<std::mutex> lk(scratch_mu);
    base = scratch_ptr[c->device];
  }
  if (!base) AV_FAIL(c, AVISM_ERR_STATE, "static scratch of this device not set up (avism_create)");
  const size_t tmpb = scan_state_bytes(n, c->stream);
  if (tmpb == ~(size_t)0) AV_FAIL(c, AVISM_ERR_HIP, "rocprim scan size query failed");
  return indexx_core(c, d_keys, n, d_indx, base, kNoAllocScratch, tmpb);
}

}  // namespace avism
