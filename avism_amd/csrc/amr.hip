// amr.hip -- MASCLET AMR input (FLAG_DATA = 0): the gas of a base grid and
// its level-1 refinement patches onto the uniform grid of a level, and the
// level's U1CO / DIVERCO from the gas and the dark-matter particles.
//
// Reference:
//   READ_MASCLET (reader.f90:86-405): the arrays this takes; density contrasts
//     get +1 after reading (reader.f90:392-398);
//   level IR < 0: SMOOTH_2 (amr2uniform.f90:62-121); IR = 0: the base grid
//     (avism.f90:1077-1084); both then DIVER_UNIFORM (avism.f90:1129-1146);
//   level IR = 1: DIVER_FINA_GAS (divergence.f90:55-337) and VMESH
//     (amr2uniform.f90:130-248), U1GCO = UGR, DIVERGCO = DIVR (avism.f90:1100-1119);
//   U1CO from FLAG_DENS (avism.f90:1369-1372), x ROTE, FP64 mean, / MEANDENS
//     (1381-1409); DIVERCO from FLAG_DIV (1633-1649).
//
// Design: every output cell gathers its contributions in the reference's
// summation order, so no atomics and no order dependence:
//  * SMOOTH_2: coarse cell c sums the base cells that the reference's index
//    formula sends to c; the formula is separable, so the host tabulates the
//    base range of every coarse index per axis and the kernel walks it in
//    (z, y, x) order, the reference's loop order;
//  * VMESH: a fine cell visits the patches that overlap its 8^3-cell bucket
//    (a host-built list, ascending patch index), adds each covering patch's
//    density and DIVER (computed on the fly: patch differences inside, the base
//    grid's on and next to the patch border), divides by 1 + the count as the
//    reference does (its FLAGAMR starts at 1), and takes the base grid's values
//    where no patch refines its parent cell.
// Levels above 1 are not supported: the reference's VMESH indexes its
// base-grid flags with level-IR cells (amr2uniform.f90:201-204), which leaves
// that array beyond level 1.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "avism_internal.h"
#include "fields_batch.h"

namespace avism {

struct AmrState {
  int32_t nb = 0, npatch = 0;
  DBuf base;                    // delta, vx, vy, vz of the base grid (nb^3 each)
  DBuf pdata;                   // per patch delta, vx, vy, vz
  std::vector<int32_t> pint;    // 6 per patch: PATCHNX, NY, NZ, PATCHX, Y, Z
  std::vector<float> preal;     // 3 per patch: PATCHRX, RY, RZ
  std::vector<int64_t> poff;    // start of each patch's data in pdata
  DBuf dpatch, dbkt, dbktoff, dsmooth, dgas, dflagr, ddiv0, dtmp;
};

void amr_free(avism_ctx* c) {
  delete c->amr;
  c->amr = nullptr;
}

namespace {

inline unsigned nbk(int64_t n, int b = 256) { return (unsigned)((n + b - 1) / b); }

// a patch as the kernels see it
struct PatchAmr {
  int32_t n1, n2, n3, L1, L2, L3;  // sides, base cell of the first cells
  int32_t x1, y1, z1, pad;         // fine index of the cell before the first (IXR1)
  int64_t off;                     // its data in pdata
};

__device__ __forceinline__ float at3(const float* a, int n, int i, int j, int k) {  // 1-based
  return a[(size_t)(i - 1) + (size_t)n * ((size_t)(j - 1) + (size_t)n * (size_t)(k - 1))];
}

// U1G = delta + 1 on the base grid, and IR = 0 copies of the velocities
__global__ void k_amr_copy(int64_t n3, const float* __restrict__ base, float* g1, float* gu, float* gv, float* gw) {
  const int64_t c = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (c >= n3) return;
  g1[c] = base[c] + 1.0f;
  gu[c] = base[n3 + c];
  gv[c] = base[2 * n3 + c];
  gw[c] = base[3 * n3 + c];
}

// SMOOTH_2 (amr2uniform.f90:62-121) as a gather: rng[axis][c] = first and
// last base index (1-based) of coarse index c (1-based), tables of n + 1 pairs
__global__ void k_amr_smooth(int nb, int n, const float* __restrict__ base, const int2* __restrict__ rng, float wei,
                             float* g1, float* gu, float* gv, float* gw) {
  const int64_t c = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (c >= (int64_t)n * n * n) return;
  const int I = (int)(c % n) + 1, J = (int)((c / n) % n) + 1, K = (int)(c / ((int64_t)n * n)) + 1;
  const int2 rx = rng[I], ry = rng[(n + 1) + J], rz = rng[2 * (n + 1) + K];
  const size_t nb3 = (size_t)nb * nb * nb;
  float s1 = 0.0f, s2 = 0.0f, s3 = 0.0f, s4 = 0.0f;
  for (int kk = rz.x; kk <= rz.y; ++kk)
    for (int jj = ry.x; jj <= ry.y; ++jj)
      for (int ii = rx.x; ii <= rx.y; ++ii) {
        const size_t b = (size_t)(ii - 1) + (size_t)nb * ((size_t)(jj - 1) + (size_t)nb * (size_t)(kk - 1));
        s1 = s1 + (base[b] + 1.0f) * wei;
        s2 = s2 + base[nb3 + b] * wei;
        s3 = s3 + base[2 * nb3 + b] * wei;
        s4 = s4 + base[3 * nb3 + b] * wei;
      }
  g1[c] = s1;
  gu[c] = s2;
  gv[c] = s3;
  gw[c] = s4;
}

// DIVER0 (divergence.f90:140-173): differences over 2 DX0 inside, the edge
// planes copied in x, then y, then z, i.e. the value of the clamped cell
__device__ __forceinline__ float div_base(const float* __restrict__ base, int nb, float dx0, int a, int b, int c) {
  const size_t nb3 = (size_t)nb * nb * nb;
  const float *U2 = base + nb3, *U3 = base + 2 * nb3, *U4 = base + 3 * nb3;
  float b21 = at3(U2, nb, a + 1, b, c) - at3(U2, nb, a - 1, b, c);
  b21 = b21 / (2.0f * dx0);
  float b32 = at3(U3, nb, a, b + 1, c) - at3(U3, nb, a, b - 1, c);
  b32 = b32 / (2.0f * dx0);
  float b43 = at3(U4, nb, a, b, c + 1) - at3(U4, nb, a, b, c - 1);
  b43 = b43 / (2.0f * dx0);
  return (b21 + b32) + b43;
}

__global__ void k_amr_div0(int nb, float dx0, const float* __restrict__ base, float* div0) {
  const int64_t c = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (c >= (int64_t)nb * nb * nb) return;
  const int i = (int)(c % nb) + 1, j = (int)((c / nb) % nb) + 1, k = (int)(c / ((int64_t)nb * nb)) + 1;
  auto cl = [&](int v) { return v < 2 ? 2 : (v > nb - 1 ? nb - 1 : v); };
  div0[c] = div_base(base, nb, dx0, cl(i), cl(j), cl(k));
}

// FLAGR of VMESH: base cells with a refined child (writes of 1 only)
__global__ void k_amr_flagr(const PatchAmr* __restrict__ P, const int64_t* __restrict__ cells, int npatch, int nb,
                            uint8_t* flagr) {
  const int64_t it = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (it >= cells[npatch]) return;
  int lo = 0, hi = npatch;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (cells[mid] <= it) lo = mid;
    else hi = mid;
  }
  const PatchAmr p = P[lo];
  const int64_t l = it - cells[lo];
  const int ix = (int)(l % p.n1) + 1, iy = (int)((l / p.n1) % p.n2) + 1, iz = (int)(l / ((int64_t)p.n1 * p.n2)) + 1;
  const int X = p.x1 + ix, Y = p.y1 + iy, Z = p.z1 + iz;
  const int x0 = (int)((float)(X - 1) / 2.0f) + 1, y0 = (int)((float)(Y - 1) / 2.0f) + 1,
            z0 = (int)((float)(Z - 1) / 2.0f) + 1;
  flagr[(size_t)(x0 - 1) + (size_t)nb * ((size_t)(y0 - 1) + (size_t)nb * (size_t)(z0 - 1))] = 1;
}

// DIVER of patch cell (ix, iy, iz) after the border blending of
// divergence.f90:228-334
__device__ __forceinline__ float div_patch(const PatchAmr& p, const float* __restrict__ pdata,
                                           const float* __restrict__ base, int nb, float dx0, int ix, int iy, int iz) {
  const int n1 = p.n1, n2 = p.n2, n3 = p.n3;
  const bool border = ix == 1 || ix == n1 || iy == 1 || iy == n2 || iz == 1 || iz == n3;
  const bool near = ix == 2 || ix == n1 - 1 || iy == 2 || iy == n2 - 1 || iz == 2 || iz == n3 - 1;
  float dv = 0.0f, aux = 0.0f;
  if (!border) {
    const size_t pc = (size_t)n1 * n2 * n3;
    const float *P2 = pdata + p.off + pc, *P3 = P2 + pc, *P4 = P3 + pc;
    auto A = [&](const float* a, int i, int j, int k) {
      return a[(size_t)(i - 1) + (size_t)n1 * ((size_t)(j - 1) + (size_t)n2 * (size_t)(k - 1))];
    };
    const float dxr = dx0 / 2.0f;
    float b21 = A(P2, ix + 1, iy, iz) - A(P2, ix - 1, iy, iz);
    b21 = b21 / (2.0f * dxr);
    float b32 = A(P3, ix, iy + 1, iz) - A(P3, ix, iy - 1, iz);
    b32 = b32 / (2.0f * dxr);
    float b43 = A(P4, ix, iy, iz + 1) - A(P4, ix, iy, iz - 1);
    b43 = b43 / (2.0f * dxr);
    dv = (b21 + b32) + b43;
  }
  if (border || near)
    aux = div_base(base, nb, dx0, (ix + 1) / 2 + p.L1 - 1, (iy + 1) / 2 + p.L2 - 1, (iz + 1) / 2 + p.L3 - 1);
  if (border) return aux;
  if (near) return 0.5f * (aux + dv);
  return dv;
}

constexpr int kBkt = 8;  // VMESH buckets: 8^3 fine cells

// VMESH (amr2uniform.f90:130-248) for the gas density and divergence
__global__ void k_amr_vmesh(int n, int nb, float dx0, const PatchAmr* __restrict__ P, const int32_t* __restrict__ bkt,
                            const int32_t* __restrict__ bktoff, const float* __restrict__ pdata,
                            const float* __restrict__ base, const float* __restrict__ div0,
                            const uint8_t* __restrict__ flagr, float* g1, float* gd) {
  const int64_t c = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (c >= (int64_t)n * n * n) return;
  const int X = (int)(c % n) + 1, Y = (int)((c / n) % n) + 1, Z = (int)(c / ((int64_t)n * n)) + 1;
  const int x0 = (X - 1) / 2 + 1, y0 = (Y - 1) / 2 + 1, z0 = (Z - 1) / 2 + 1;
  const size_t b0 = (size_t)(x0 - 1) + (size_t)nb * ((size_t)(y0 - 1) + (size_t)nb * (size_t)(z0 - 1));
  if (!flagr[b0]) {  // the parent is not refined: the base grid's values
    g1[c] = base[b0] + 1.0f;
    gd[c] = div0[b0];
    return;
  }
  const int nbk1 = (n + kBkt - 1) / kBkt;
  const int bi = (X - 1) / kBkt + nbk1 * ((Y - 1) / kBkt + nbk1 * ((Z - 1) / kBkt));
  float su = 0.0f, sd = 0.0f;
  int cnt = 1;
  for (int q = bktoff[bi]; q < bktoff[bi + 1]; ++q) {
    const PatchAmr p = P[bkt[q]];
    const int ix = X - p.x1, iy = Y - p.y1, iz = Z - p.z1;
    if (ix < 1 || ix > p.n1 || iy < 1 || iy > p.n2 || iz < 1 || iz > p.n3) continue;
    ++cnt;
    const float d = pdata[p.off + (size_t)(ix - 1) + (size_t)p.n1 * ((size_t)(iy - 1) + (size_t)p.n2 * (size_t)(iz - 1))];
    su = su + (d + 1.0f);
    sd = sd + div_patch(p, pdata, base, nb, dx0, ix, iy, iz);
  }
  g1[c] = su / (float)cnt;
  gd[c] = sd / (float)cnt;
}

constexpr int kCmbB = 256, kCmbV = 4;  // the cell sums' fixed order: 4 cells a thread, 256 threads a block

// U1CO before the mean (avism.f90:1369-1383) and its FP64 block sums
__global__ void __launch_bounds__(kCmbB) k_amr_combine(int64_t ncell, const unsigned long long* __restrict__ mass,
                                                       double inv_scale, int fp64, float den, float rote,
                                                       const float* __restrict__ g1, int flag_dens, float* pre,
                                                       double* partial) {
  const int64_t c0 = ((int64_t)blockIdx.x * kCmbB + threadIdx.x) * kCmbV;
  double contrib = 0.0;
  for (int v = 0; v < kCmbV; ++v) {
    const int64_t c = c0 + v;
    if (c >= ncell) break;
    float dm = 0.0f;
    if (flag_dens <= 1) {  // U1DMCO = mass / (DX^3 RETE^3) / ROTE (avism.f90:1059-1060)
      const unsigned long long a = mass[c];
      const float m = fp64 ? (float)__longlong_as_double((long long)a) : (float)((double)a * inv_scale);
      dm = m / den;
      dm = dm / rote;
    }
    float u = flag_dens == 0 ? dm + g1[c] : (flag_dens == 1 ? dm : g1[c]);
    u = u * rote;
    pre[c] = u;
    contrib += (double)u;
  }
  for (int off = 32; off > 0; off >>= 1) contrib += __shfl_xor(contrib, off);
  __shared__ double wsum[kCmbB / 64];
  if (threadIdx.x % 64 == 0) wsum[threadIdx.x / 64] = contrib;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < kCmbB / 64; ++w) t += wsum[w];
    partial[blockIdx.x] = t;
  }
}

// U1CO = U1CO / MEANDENS and DIVERCO from FLAG_DIV (avism.f90:1409, 1633-1649);
// div holds DIVERDMCO on entry
__global__ void k_amr_final(int64_t ncell, const float* __restrict__ pre, const float* __restrict__ mean,
                            const float* __restrict__ gd, int flag_div, float lin, float* u1co, float* div) {
  const int64_t c = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (c >= ncell) return;
  const float u = pre[c] / (*mean);
  u1co[c] = u;
  const float dm = div[c];
  float d;
  if (flag_div == 2) d = gd[c];
  else if (flag_div == 1) d = dm;
  else if (flag_div == 0) d = fminf(gd[c], dm);
  else d = -(lin * (u - 1.0f));
  div[c] = d;
}

}  // namespace

int amr_set(avism_ctx* c, int32_t nb, const float* base, int32_t npatch, const int32_t* pint, const float* preal,
            const float* pdata, int where) {
  if (nb < 4) AV_FAIL(c, AVISM_ERR_ARG, "the base grid needs at least 4 cells per axis");
  if (npatch < 0) AV_FAIL(c, AVISM_ERR_ARG, "negative patch count");
  if (npatch > 0 && (!pint || !preal || !pdata)) AV_FAIL(c, AVISM_ERR_ARG, "null patch array");
  if ((int64_t)nb * nb * nb >= (int64_t)0x7fffffff) AV_FAIL(c, AVISM_ERR_LIMIT, "base grid above 2^31 cells");
  amr_free(c);
  AmrState* A = new AmrState();
  c->amr = A;
  A->nb = nb;
  A->npatch = npatch;
  const hipMemcpyKind kind = where == AVISM_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
  const size_t nb3 = (size_t)nb * nb * nb;
  if (!A->base.ensure(nb3 * 16)) AV_FAIL(c, AVISM_ERR_HIP, "out of device memory (AMR)");
  AV_CHECK_HIP(c, hipMemcpy(A->base.p, base, nb3 * 16, kind));
  A->pint.resize((size_t)npatch * 6);
  A->preal.resize((size_t)npatch * 3);
  if (npatch > 0) {
    AV_CHECK_HIP(c, hipMemcpy(A->pint.data(), pint, A->pint.size() * 4,
                              where == AVISM_DEVICE ? hipMemcpyDeviceToHost : hipMemcpyHostToHost));
    AV_CHECK_HIP(c, hipMemcpy(A->preal.data(), preal, A->preal.size() * 4,
                              where == AVISM_DEVICE ? hipMemcpyDeviceToHost : hipMemcpyHostToHost));
  }
  A->poff.assign((size_t)npatch + 1, 0);
  for (int32_t p = 0; p < npatch; ++p) {
    const int32_t* q = &A->pint[6 * (size_t)p];
    if (q[0] < 1 || q[1] < 1 || q[2] < 1) AV_FAIL(c, AVISM_ERR_INPUT, "a patch with no cells");
    A->poff[p + 1] = A->poff[p] + 4 * (int64_t)q[0] * q[1] * q[2];
  }
  if (!A->pdata.ensure((size_t)std::max<int64_t>(A->poff[npatch], 1) * 4))
    AV_FAIL(c, AVISM_ERR_HIP, "out of device memory (AMR)");
  if (npatch > 0) AV_CHECK_HIP(c, hipMemcpy(A->pdata.p, pdata, (size_t)A->poff[npatch] * 4, kind));
  return AVISM_OK;
}

int amr_fields(avism_ctx* c, int32_t ir, int32_t flag_dens, int32_t flag_div) {
  AmrState* A = c->amr;
  if (!A) AV_FAIL(c, AVISM_ERR_STATE, "no AMR gas loaded (avism_amr_set)");
  if (!c->grid_set) AV_FAIL(c, AVISM_ERR_STATE, "avism_set_grid must be called first");
  if (ir > 1) AV_FAIL(c, AVISM_ERR_UNSUPPORTED, "AMR levels above 1 are not supported (the reference's VMESH "
                                                  "indexes its base-grid flags past level 1)");
  if (ir < -8) AV_FAIL(c, AVISM_ERR_ARG, "level below -8");
  if (flag_dens < 0 || flag_dens > 2 || flag_div < -1 || flag_div > 2)
    AV_FAIL(c, AVISM_ERR_INPUT, "FLAG_DENS must be 0, 1 or 2 and FLAG_DIV -1, 0, 1 or 2");
  const int32_t nb = A->nb;
  const int32_t n = ir >= 0 ? nb << ir : nb >> (-ir);  // avism.f90:750
  if (ir < 0 && (nb % (1 << (-ir))) != 0) AV_FAIL(c, AVISM_ERR_INPUT, "base grid not divisible for that level");
  if (c->g.n != n) AV_FAIL(c, AVISM_ERR_INPUT, "the grid must have NHYX * 2^IR cells per axis");
  const bool need_dm = flag_dens <= 1 || flag_div == 0 || flag_div == 1;
  if (need_dm && (c->np <= 0 || c->mass_n != n || !c->fields_ready))
    AV_FAIL(c, AVISM_ERR_STATE, "FLAG_DENS or FLAG_DIV use the dark matter: avism_particles_to_grid on this grid first");
  if (int rc = finish_copies(c)) return rc;
  const int64_t n3 = (int64_t)n * n * n;
  const size_t nb3 = (size_t)nb * nb * nb;
  const GridGeom g0 = make_grid(nb, c->g.L0, 0);
  const GridGeom& gl = c->g;
  if (!A->dgas.ensure((size_t)n3 * 16) || !A->dtmp.ensure((size_t)n3 * 4) || !c->u1co.ensure((size_t)n3 * 4) ||
      !c->div.ensure((size_t)n3 * 4) || !c->fu.ensure((size_t)n3 * 4) || !c->fv.ensure((size_t)n3 * 4) ||
      !c->fw.ensure((size_t)n3 * 4) || !c->partial64.ensure((size_t)((n3 + kCmbB * kCmbV - 1) / (kCmbB * kCmbV)) * 8))
    AV_FAIL(c, AVISM_ERR_HIP, "out of device memory (AMR)");
  float* g1 = A->dgas.as<float>();
  float* gd = g1 + n3;
  float* gu = g1 + 2 * n3;
  float* gv = g1 + 3 * n3;
  float* gw = A->dtmp.as<float>();
  const float* dbase = A->base.as<float>();
  {
    StageTimer t(c, "amr_gas");
    if (ir == 0) {
      k_amr_copy<<<nbk(n3), 256, 0, c->stream>>>(n3, dbase, g1, gu, gv, gw);
    } else if (ir < 0) {  // SMOOTH_2: base ranges of every coarse index, per axis
      const float dx0 = g0.dx, dxl = gl.dx, rx1 = gl.radx(1);
      const float wei = ((dx0 * dx0) * dx0) / ((dxl * dxl) * dxl);
      std::vector<int32_t> lo(n + 1, INT32_MAX), hi(n + 1, INT32_MIN);
      for (int32_t ii = 1; ii <= nb; ++ii) {
        const int32_t I = (int32_t)(((g0.radx(ii) - rx1) / dxl) + 0.5f) + 1;
        if (I < 1 || I > n) AV_FAIL(c, AVISM_ERR_INPUT, "a base cell falls outside the coarse grid");
        lo[I] = std::min(lo[I], ii);
        hi[I] = std::max(hi[I], ii);
      }
      std::vector<int32_t> rng(6 * ((size_t)n + 1));
      for (int a = 0; a < 3; ++a)
        for (int32_t I = 0; I <= n; ++I) {
          const bool empty = lo[I] > hi[I];
          rng[2 * (a * ((size_t)n + 1) + I)] = empty ? 1 : lo[I];
          rng[2 * (a * ((size_t)n + 1) + I) + 1] = empty ? 0 : hi[I];
        }
      if (!A->dsmooth.ensure(rng.size() * 4)) AV_FAIL(c, AVISM_ERR_HIP, "out of device memory (AMR)");
      AV_CHECK_HIP(c, hipMemcpyAsync(A->dsmooth.p, rng.data(), rng.size() * 4, hipMemcpyHostToDevice, c->stream));
      k_amr_smooth<<<nbk(n3), 256, 0, c->stream>>>(nb, n, dbase, A->dsmooth.as<int2>(), wei, g1, gu, gv, gw);
      AV_CHECK_HIP(c, hipStreamSynchronize(c->stream));  // rng is freed on return
    }
    if (ir <= 0) {  // DIVER_UNIFORM of the gas velocities
      AV_CHECK_HIP(c, launch_div_finalize(c->stream, n, gl.dx, gl.periodic, gu, gv, gw, nullptr, nullptr, nullptr, gd));
    } else {  // DIVER_FINA_GAS + VMESH
      const float dx0 = g0.dx, dxl = gl.dx, rx1 = gl.radx(1);
      const int32_t np = A->npatch;
      std::vector<PatchAmr> pa(np);
      std::vector<int64_t> cells((size_t)np + 1, 0);
      const int nbk1 = (n + kBkt - 1) / kBkt;
      std::vector<std::vector<int32_t>> lists((size_t)nbk1 * nbk1 * nbk1);
      for (int32_t p = 0; p < np; ++p) {
        const int32_t* q = &A->pint[6 * (size_t)p];
        const float* r = &A->preal[3 * (size_t)p];
        PatchAmr& P = pa[p];
        P.n1 = q[0]; P.n2 = q[1]; P.n3 = q[2]; P.L1 = q[3]; P.L2 = q[4]; P.L3 = q[5];
        // IXR1 = INT(((PATCHRX - 0.5 DXR) - RX1) / DXR + 0.5) (amr2uniform.f90:178-180)
        P.x1 = (int32_t)((((r[0] - 0.5f * dxl) - rx1) / dxl) + 0.5f);
        P.y1 = (int32_t)((((r[1] - 0.5f * dxl) - rx1) / dxl) + 0.5f);
        P.z1 = (int32_t)((((r[2] - 0.5f * dxl) - rx1) / dxl) + 0.5f);
        P.off = A->poff[p];
        if (P.x1 < 0 || P.x1 + P.n1 > n || P.y1 < 0 || P.y1 + P.n2 > n || P.z1 < 0 || P.z1 + P.n3 > n)
          AV_FAIL(c, AVISM_ERR_INPUT, "a level-1 patch lies outside the level's grid");
        // the base cells the border stencil reads (divergence.f90:259-261) stay inside the base grid
        const int32_t lo[3] = {P.L1, P.L2, P.L3}, ns[3] = {P.n1, P.n2, P.n3};
        for (int a = 0; a < 3; ++a)
          if (lo[a] < 2 || (ns[a] + 1) / 2 + lo[a] - 1 > nb - 1)
            AV_FAIL(c, AVISM_ERR_INPUT, "a patch's parent cells touch the base grid's faces");
        cells[p + 1] = cells[p] + (int64_t)P.n1 * P.n2 * P.n3;
        for (int bz = P.z1 / kBkt; bz <= (P.z1 + P.n3 - 1) / kBkt; ++bz)
          for (int by = P.y1 / kBkt; by <= (P.y1 + P.n2 - 1) / kBkt; ++by)
            for (int bx = P.x1 / kBkt; bx <= (P.x1 + P.n1 - 1) / kBkt; ++bx)
              lists[(size_t)bx + (size_t)nbk1 * ((size_t)by + (size_t)nbk1 * bz)].push_back(p);
      }
      std::vector<int32_t> bktoff(lists.size() + 1, 0), bkt;
      for (size_t b = 0; b < lists.size(); ++b) {
        bktoff[b + 1] = bktoff[b] + (int32_t)lists[b].size();
        bkt.insert(bkt.end(), lists[b].begin(), lists[b].end());
      }
      if (bkt.empty()) bkt.push_back(0);
      if (!A->dpatch.ensure(sizeof(PatchAmr) * (size_t)std::max(np, 1) + 8 * ((size_t)np + 1)) ||
          !A->dbkt.ensure(bkt.size() * 4) || !A->dbktoff.ensure(bktoff.size() * 4) || !A->dflagr.ensure(nb3) ||
          !A->ddiv0.ensure(nb3 * 4))
        AV_FAIL(c, AVISM_ERR_HIP, "out of device memory (AMR)");
      PatchAmr* dpa = A->dpatch.as<PatchAmr>();
      int64_t* dcells = (int64_t*)((char*)A->dpatch.p + sizeof(PatchAmr) * (size_t)std::max(np, 1));
      if (np > 0) AV_CHECK_HIP(c, hipMemcpyAsync(dpa, pa.data(), sizeof(PatchAmr) * np, hipMemcpyHostToDevice, c->stream));
      AV_CHECK_HIP(c, hipMemcpyAsync(dcells, cells.data(), cells.size() * 8, hipMemcpyHostToDevice, c->stream));
      AV_CHECK_HIP(c, hipMemcpyAsync(A->dbkt.p, bkt.data(), bkt.size() * 4, hipMemcpyHostToDevice, c->stream));
      AV_CHECK_HIP(c, hipMemcpyAsync(A->dbktoff.p, bktoff.data(), bktoff.size() * 4, hipMemcpyHostToDevice, c->stream));
      AV_CHECK_HIP(c, hipMemsetAsync(A->dflagr.p, 0, nb3, c->stream));
      k_amr_div0<<<nbk((int64_t)nb3), 256, 0, c->stream>>>(nb, dx0, dbase, A->ddiv0.as<float>());
      if (cells[np] > 0)
        k_amr_flagr<<<nbk(cells[np]), 256, 0, c->stream>>>(dpa, dcells, np, nb, A->dflagr.as<uint8_t>());
      k_amr_vmesh<<<nbk(n3), 256, 0, c->stream>>>(n, nb, dx0, dpa, A->dbkt.as<int32_t>(), A->dbktoff.as<int32_t>(),
                                                  A->pdata.as<float>(), dbase, A->ddiv0.as<float>(),
                                                  A->dflagr.as<uint8_t>(), g1, gd);
      AV_CHECK_HIP(c, hipGetLastError());
      AV_CHECK_HIP(c, hipStreamSynchronize(c->stream));  // host tables are freed on return
    }
    AV_CHECK_HIP(c, hipGetLastError());
  }
  // the level's U1CO and DIVERCO
  const bool dm_div = ir <= 0 && need_dm;  // DIVERDMCO exists at levels <= 0 only (avism.f90:1127-1146)
  if (!dm_div) AV_CHECK_HIP(c, launch_fill_u32(c->stream, n3, c->div.as<uint32_t>(), 0u));
  const float den = ((gl.dx * gl.dx) * gl.dx) * ((c->rete * c->rete) * c->rete);
  double* dsum = (double*)((char*)c->scalars.p + 256);
  float* dmean = (float*)((char*)c->scalars.p + 264);
  double* slices = (double*)((char*)c->scalars.p + scalars_slices_offset(c));
  const int64_t nparts = (n3 + kCmbB * kCmbV - 1) / (kCmbB * kCmbV);
  {
    StageTimer t(c, "amr_combine");
    k_amr_combine<<<(unsigned)nparts, kCmbB, 0, c->stream>>>(
        n3, flag_dens <= 1 ? c->mass_fix.as<unsigned long long>() : nullptr, c->mass_inv_scale, c->mass_fp64, den,
        c->rote, g1, flag_dens, A->dtmp.as<float>(), c->partial64.as<double>());
    AV_CHECK_HIP(c, launch_reduce_mean(c->stream, c->partial64.as<double>(), nparts, n3, slices, dsum, dmean));
    const float lin = (powf(c->omega0, 0.6f) * c->rete) * c->hte;  // avism.f90:1648
    k_amr_final<<<nbk(n3), 256, 0, c->stream>>>(n3, A->dtmp.as<float>(), dmean, gd, flag_div, lin,
                                                c->u1co.as<float>(), c->div.as<float>());
    AV_CHECK_HIP(c, hipGetLastError());
  }
  if (!need_dm) {  // no dark matter: the velocity maps are zero
    AV_CHECK_HIP(c, launch_fill_u32(c->stream, n3, c->fu.as<uint32_t>(), 0u));
    AV_CHECK_HIP(c, launch_fill_u32(c->stream, n3, c->fv.as<uint32_t>(), 0u));
    AV_CHECK_HIP(c, launch_fill_u32(c->stream, n3, c->fw.as<uint32_t>(), 0u));
  }
  float hmean = 0;
  AV_CHECK_HIP(c, hipMemcpyAsync(&hmean, dmean, 4, hipMemcpyDeviceToHost, c->stream));
  AV_CHECK_HIP(c, hipMemcpyAsync(&c->meansum, dsum, 8, hipMemcpyDeviceToHost, c->stream));
  AV_CHECK_HIP(c, hipStreamSynchronize(c->stream));
  c->meandens_raw = hmean;
  const float ul3 = (kUL * kUL) * kUL;
  c->meandens = hmean * (kUM / ul3);  // avism.f90:1409
  c->nvoid = 0;
  c->post_ready = false;
  c->fields_ready = true;
  return AVISM_OK;
}

}  // namespace avism
