// batch.hip -- many independent AVISM patches of uneven size (each its own
// grid, particles and survey mask: one run of the reference per patch,
// avism.f90:1177-1973) through one field stage.
//
// Per batch:
//  * eager prefix: the upload, the unit conversion and box check per patch
//    (reader.f90:457-461, avism.f90:573-575), one radix sort of every particle
//    by (patch, bin), each patch's own bin starts, and a device table of
//    per-patch descriptors (fields_batch.h);
//  * the field stage -- PPART_DENS, VVEL_INTERP_SPH_VW, FILL_H_ZEROS,
//    DDENS_INTERP_SPH, the mean density and DIVER_UNIFORM of every patch -- as
//    persistent kernels over all patches' work items plus small per-patch
//    launches, captured once per layout into a hipGraph and replayed for the
//    next batches of that layout (the graph holds only this library's
//    kernels: no memset or library nodes);
//  * tile shapes and persistent grid sizes picked by an autotuner that times
//    the candidates on the layout's first batch and records its choices;
//  * the void stage (MARK_ALL .. post-processing) of each patch on its own
//    context and stream, the patches spread over host threads.
// Every patch's fields are those of the single-grid path on its particles.
#include <atomic>
#include <cmath>
#include <cstring>
#include <mutex>
#include <sstream>
#include <thread>

#include <rocprim/rocprim.hpp>

#include "fields_batch.h"

namespace avism {

struct BatchState {
  int32_t npatch = 0;
  std::vector<int32_t> n, periodic, nb;
  std::vector<float> L0, bsz;
  std::vector<int64_t> np, p0, c0, kbase, sbase, rad0, part0;
  std::vector<GridGeom> geom;
  int64_t np_tot = 0, ncell_tot = 0, nkeys = 0, nstart = 0, nparts = 0;
  DBuf x, y, z, vx, vy, vz, m, smask, key, val, key2, val2, sort_tmp, count, scan, start, pos4, orig, vel4;
  DBuf pdens, hk, hbits, norm, u, v, w, umass, u1co, div, acc, partial, rad, desc, flags, offs, qcnt, means,
      slices, pairs, p0dev, kbasedev;
  std::vector<PatchDev> hdesc;
  bool has_mask = false, loaded = false, fields_ready = false;
  std::vector<float> meandens, meandens_raw;
  // autotuner choices, per k
  int k_tuned = -1, vvel_shape = 0, ddens_shape = 0;
  unsigned ppart_blocks = 2048, vvel_blocks = 2048, ddens_blocks = 2048;
  std::string tuning_json;
  // the captured field stage
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int graph_k = -1;
  int64_t graph_launches = 0;
  hipEvent_t ev_done = nullptr;
  std::vector<avism_ctx*> sub;
  void drop_graph() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
    graph_k = -1;
  }
  ~BatchState() {
    drop_graph();
    if (ev_done) (void)hipEventDestroy(ev_done);
    for (avism_ctx* s : sub)
      if (s) avism_destroy(s);
  }
};

void batch_free(avism_ctx* c) {
  delete c->batch;
  c->batch = nullptr;
}

namespace {

inline unsigned nb256(int64_t n) { return (unsigned)((n + 255) / 256); }

// global particle i -> its patch (last p with p0[p] <= i)
__device__ __forceinline__ int patch_of(const int64_t* __restrict__ off, int npatch, int64_t i) {
  int lo = 0, hi = npatch;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (off[mid] <= i) lo = mid;
    else hi = mid;
  }
  return lo;
}

// sort key of each particle: kbase of its patch + its bin in the patch's bins
__global__ void k_batch_bin_keys(int64_t np, int npatch, const int64_t* __restrict__ p0,
                                 const int64_t* __restrict__ kbase, const PatchDev* __restrict__ P,
                                 const float* x, const float* y, const float* z, uint32_t* key, uint32_t* val,
                                 int32_t* count) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  const int p = patch_of(p0, npatch, i);
  const Bins& B = P[p].B;
  const int64_t b = kbase[p] + B.bin_of(x[i]) + (int64_t)B.nb * (B.bin_of(y[i]) + (int64_t)B.nb * B.bin_of(z[i]));
  key[i] = (uint32_t)b;
  val[i] = (uint32_t)i;
  atomicAdd(&count[b], 1);
}

// each patch's bin starts, relative to its first particle, with its end sentinel
__global__ void k_batch_bin_starts(int64_t nkeys, int npatch, const int64_t* __restrict__ kbase,
                                   const int64_t* __restrict__ p0, const int32_t* __restrict__ scan,
                                   int32_t* start) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= nkeys) return;
  const int p = patch_of(kbase, npatch, g);
  // patch p's starts sit at kbase[p] + p (one sentinel per earlier patch)
  start[g + p] = (int32_t)((int64_t)scan[g] - p0[p]);
  if (g + 1 == kbase[p + 1]) start[g + p + 1] = (int32_t)(p0[p + 1] - p0[p]);
}

}  // namespace

static int batch_check(avism_ctx* c, bool need_loaded = true) {
  if (!c->batch || c->batch->npatch == 0) AV_FAIL(c, AVISM_ERR_STATE, "no batch set (avism_batch_set)");
  if (need_loaded && !c->batch->loaded) AV_FAIL(c, AVISM_ERR_STATE, "no batch particles loaded (avism_batch_load)");
  return AVISM_OK;
}

int batch_set(avism_ctx* c, int32_t npatch, const int32_t* n, const float* L0, const int32_t* periodic,
              const int64_t* np) {
  if (npatch < 1 || npatch > 65536) AV_FAIL(c, AVISM_ERR_ARG, "npatch must be 1..65536");
  if (!n || !L0 || !np) AV_FAIL(c, AVISM_ERR_ARG, "null batch layout");
  if (int rc = finish_copies(c)) return rc;
  AV_CHECK_HIP(c, hipStreamSynchronize(c->stream));
  if (!c->batch) c->batch = new BatchState();
  BatchState& S = *c->batch;
  S.drop_graph();
  S.loaded = S.fields_ready = false;
  S.k_tuned = -1;
  S.npatch = 0;
  S.n.assign(n, n + npatch);
  S.L0.assign(L0, L0 + npatch);
  S.periodic.assign(npatch, 0);
  if (periodic) S.periodic.assign(periodic, periodic + npatch);
  S.np.assign(np, np + npatch);
  S.nb.resize(npatch);
  S.bsz.resize(npatch);
  S.p0.assign(npatch + 1, 0);
  S.c0.assign(npatch + 1, 0);
  S.kbase.assign(npatch + 1, 0);
  S.sbase.assign(npatch + 1, 0);
  S.rad0.assign(npatch + 1, 0);
  S.part0.assign(npatch + 1, 0);
  S.geom.clear();
  for (int p = 0; p < npatch; ++p) {
    if (S.n[p] < 4) AV_FAIL(c, AVISM_ERR_ARG, "patch grids need at least 4 cells per axis");
    if (!(S.L0[p] > 0.0f)) AV_FAIL(c, AVISM_ERR_ARG, "patch box sizes must be positive");
    if (S.periodic[p] != 0 && S.periodic[p] != 1) AV_FAIL(c, AVISM_ERR_ARG, "periodic must be 0 or 1");
    if (S.np[p] < 1) AV_FAIL(c, AVISM_ERR_ARG, "every patch needs particles");
    const int64_t n3 = (int64_t)S.n[p] * S.n[p] * S.n[p];
    S.nb[p] = bins_per_axis(S.n[p], S.np[p]);  // as the single grid
    S.bsz[p] = S.L0[p] / (float)S.nb[p];
    S.geom.push_back(make_grid(S.n[p], S.L0[p], S.periodic[p]));
    S.p0[p + 1] = S.p0[p] + S.np[p];
    S.c0[p + 1] = S.c0[p] + ((n3 + 3) & ~(int64_t)3);  // 16-byte aligned patch grids
    S.kbase[p + 1] = S.kbase[p] + (int64_t)S.nb[p] * S.nb[p] * S.nb[p];
    S.sbase[p + 1] = S.kbase[p + 1] + p + 1;
    S.rad0[p + 1] = S.rad0[p] + (int64_t)S.geom[p].rad.size();
    S.part0[p + 1] = S.part0[p] + finalize_blocks(n3);
  }
  S.np_tot = S.p0[npatch];
  S.ncell_tot = S.c0[npatch];
  S.nkeys = S.kbase[npatch];
  S.nstart = S.sbase[npatch];
  S.nparts = S.part0[npatch];
  if (S.np_tot >= (int64_t)0x7fffffff) AV_FAIL(c, AVISM_ERR_LIMIT, "a batch holds at most 2^31-1 particles");
  if (S.nkeys >= (int64_t)0x7fffffff) AV_FAIL(c, AVISM_ERR_LIMIT, "a batch holds at most 2^31-1 bins");
  const size_t P4 = (size_t)S.np_tot * 4, C4 = (size_t)S.ncell_tot * 4;
  bool ok = S.x.ensure(P4) && S.y.ensure(P4) && S.z.ensure(P4) && S.vx.ensure(P4) && S.vy.ensure(P4) &&
            S.vz.ensure(P4) && S.m.ensure(P4) && S.smask.ensure((size_t)S.ncell_tot) && S.key.ensure(P4) &&
            S.val.ensure(P4) && S.key2.ensure(P4) && S.val2.ensure(P4) && S.count.ensure((size_t)S.nkeys * 4 + 4) &&
            S.scan.ensure((size_t)S.nkeys * 4 + 4) && S.start.ensure((size_t)S.nstart * 4) &&
            S.pos4.ensure(P4 * 4) && S.orig.ensure(P4) && S.vel4.ensure(P4 * 4) && S.pdens.ensure(P4) &&
            S.hk.ensure(P4) && S.hbits.ensure(P4) && S.norm.ensure(P4) && S.u.ensure(C4) && S.v.ensure(C4) &&
            S.w.ensure(C4) && S.umass.ensure(C4) && S.u1co.ensure(C4) && S.div.ensure(C4) &&
            S.acc.ensure((size_t)S.ncell_tot * 8) && S.partial.ensure((size_t)S.nparts * 8) &&
            S.rad.ensure((size_t)S.rad0[npatch] * 4) && S.desc.ensure(sizeof(PatchDev) * npatch) &&
            S.flags.ensure((size_t)npatch * 16) && S.offs.ensure((size_t)3 * (npatch + 1) * 8) &&
            S.qcnt.ensure(3 * 8 * 4) && S.means.ensure((size_t)npatch * 16) &&
            S.slices.ensure((size_t)npatch * 512 * 8) && S.pairs.ensure(8) &&
            S.p0dev.ensure((size_t)(npatch + 1) * 8) && S.kbasedev.ensure((size_t)(npatch + 1) * 8);
  if (!ok) AV_FAIL(c, AVISM_ERR_HIP, "out of device memory (batch)");
  if (!S.ev_done) AV_CHECK_HIP(c, hipEventCreateWithFlags(&S.ev_done, hipEventDisableTiming));
  // RADX of every patch, the particle and bin offsets
  std::vector<float> hrad((size_t)S.rad0[npatch]);
  for (int p = 0; p < npatch; ++p)
    std::copy(S.geom[p].rad.begin(), S.geom[p].rad.end(), hrad.begin() + S.rad0[p]);
  AV_CHECK_HIP(c, hipMemcpy(S.rad.p, hrad.data(), hrad.size() * 4, hipMemcpyHostToDevice));
  AV_CHECK_HIP(c, hipMemcpy(S.p0dev.p, S.p0.data(), (size_t)(npatch + 1) * 8, hipMemcpyHostToDevice));
  AV_CHECK_HIP(c, hipMemcpy(S.kbasedev.p, S.kbase.data(), (size_t)(npatch + 1) * 8, hipMemcpyHostToDevice));
  // descriptors: everything but the data-dependent scale, uniform and r0
  S.hdesc.assign(npatch, PatchDev{});
  for (int p = 0; p < npatch; ++p) {
    PatchDev& D = S.hdesc[p];
    const int64_t a = S.p0[p], cc = S.c0[p];
    D.B.pos = S.pos4.as<float4>() + a;
    D.B.orig = S.orig.as<uint32_t>() + a;
    D.B.start = S.start.as<int32_t>() + S.sbase[p];
    D.B.nb = S.nb[p];
    D.B.lo = -S.L0[p] / 2.0f;
    D.B.bsz = S.bsz[p];
    D.B.L = S.L0[p];
    D.B.eps = S.L0[p] * 2e-6f + S.bsz[p] * 1e-5f;
    D.B.periodic = S.periodic[p];
    D.rad = S.rad.as<float>() + S.rad0[p];
    D.rlo = S.geom[p].rlo;
    D.n = S.n[p];
    D.periodic = S.periodic[p];
    D.dx = S.geom[p].dx;
    D.rx1 = S.geom[p].rx1;
    D.np = S.np[p];
    D.vel = S.vel4.as<float4>() + a;
    D.hbits = S.hbits.as<uint32_t>() + a;
    D.pdens = S.pdens.as<float>() + a;
    D.hk = S.hk.as<float>() + a;
    D.norm = S.norm.as<float>() + a;
    D.u = S.u.as<float>() + cc;
    D.v = S.v.as<float>() + cc;
    D.w = S.w.as<float>() + cc;
    D.acc = S.acc.as<unsigned long long>() + cc;
    D.smask = nullptr;
  }
  S.meandens.assign(npatch, 0.0f);
  S.meandens_raw.assign(npatch, 0.0f);
  S.npatch = npatch;
  return AVISM_OK;
}

static int upload_desc(avism_ctx* c) {
  BatchState& S = *c->batch;
  AV_CHECK_HIP(c, hipMemcpyAsync(S.desc.p, S.hdesc.data(), sizeof(PatchDev) * S.npatch, hipMemcpyHostToDevice,
                                 c->stream));
  AV_CHECK_HIP(c, hipStreamSynchronize(c->stream));  // the host table may change next
  return AVISM_OK;
}

int batch_load(avism_ctx* c, const float* x, const float* y, const float* z, const float* vx, const float* vy,
               const float* vz, const float* m, const int8_t* smask, int where) {
  if (int rc = batch_check(c, false)) return rc;
  BatchState& S = *c->batch;
  AV_CHECK_HIP(c, hipStreamSynchronize(c->stream));
  S.loaded = S.fields_ready = false;
  const size_t P4 = (size_t)S.np_tot * 4;
  const hipMemcpyKind kind = where == AVISM_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
  {
    StageTimer t(c, "batch_upload");
    DBuf* dst[7] = {&S.x, &S.y, &S.z, &S.vx, &S.vy, &S.vz, &S.m};
    const float* src[7] = {x, y, z, vx, vy, vz, m};
    for (int i = 0; i < 7; ++i) {
      if (src[i])
        AV_CHECK_HIP(c, hipMemcpyAsync(dst[i]->p, src[i], P4, kind, c->stream));
      else  // particles without velocities: zero
        AV_CHECK_HIP(c, hipMemsetAsync(dst[i]->p, 0, P4, c->stream));
    }
    S.has_mask = smask != nullptr;
    for (int p = 0; p < S.npatch; ++p) S.hdesc[p].smask = smask ? S.smask.as<int8_t>() + S.c0[p] : nullptr;
    if (smask) {  // the caller's masks are packed back to back: copy each to its aligned slot
      int64_t off = 0;
      for (int p = 0; p < S.npatch; ++p) {
        const int64_t n3 = (int64_t)S.n[p] * S.n[p] * S.n[p];
        AV_CHECK_HIP(c, hipMemcpyAsync(S.smask.as<int8_t>() + S.c0[p], smask + off, (size_t)n3, kind, c->stream));
        off += n3;
      }
    }
    for (int p = 0; p < S.npatch; ++p) {
      const int64_t a = S.p0[p];
      AV_CHECK_HIP(c, launch_convert_check(c->stream, S.np[p], S.m.as<float>() + a, S.x.as<float>() + a,
                                           S.y.as<float>() + a, S.z.as<float>() + a, S.L0[p] / 2.0f,
                                           S.flags.as<int>() + 4 * p));
    }
    AV_CHECK_HIP(c, launch_convert_vel(c->stream, S.np_tot, S.vx.as<float>(), S.vy.as<float>(), S.vz.as<float>()));
  }
  std::vector<int> hf((size_t)4 * S.npatch);
  AV_CHECK_HIP(c, hipMemcpyAsync(hf.data(), S.flags.p, hf.size() * 4, hipMemcpyDeviceToHost, c->stream));
  AV_CHECK_HIP(c, hipStreamSynchronize(c->stream));
  for (int p = 0; p < S.npatch; ++p) {
    float mmax;
    memcpy(&mmax, &hf[4 * p + 1], 4);
    if (hf[4 * p])
      AV_FAIL(c, AVISM_ERR_INPUT, "patch " + std::to_string(p) +
                                      ": Particles outside the bounding box! Check L0! Particles must be inside "
                                      "[-L0/2, L0/2] in each direction.");
    if (!std::isfinite(mmax)) AV_FAIL(c, AVISM_ERR_INPUT, "patch " + std::to_string(p) + ": non-finite particle mass");
    // DDENS fixed point: every cell sum stays below np * max mass * scale <= 2^62
    const double bound = (double)S.np[p] * (double)(mmax > 0.0f ? mmax : 1.0f);
    int e2 = 0;
    std::frexp(bound, &e2);
    S.hdesc[p].scale = std::ldexp(1.0, 62 - e2);
    S.hdesc[p].uniform = hf[4 * p + 3] == 0 ? 1 : 0;
  }
  if (int rc = upload_desc(c)) return rc;
  {
    StageTimer t(c, "batch_bin");
    AV_CHECK_HIP(c, hipMemsetAsync(S.count.p, 0, (size_t)S.nkeys * 4 + 4, c->stream));
    k_batch_bin_keys<<<nb256(S.np_tot), 256, 0, c->stream>>>(
        S.np_tot, S.npatch, S.p0dev.as<int64_t>(), S.kbasedev.as<int64_t>(), S.desc.as<PatchDev>(), S.x.as<float>(),
        S.y.as<float>(), S.z.as<float>(), S.key.as<uint32_t>(), S.val.as<uint32_t>(), S.count.as<int32_t>());
    AV_CHECK_HIP(c, hipGetLastError());
    unsigned bits = 1;
    while ((1ll << bits) < S.nkeys) bits++;
    size_t t1 = 0, t2 = 0;
    AV_CHECK_HIP(c, rocprim::radix_sort_pairs(nullptr, t1, S.key.as<uint32_t>(), S.key2.as<uint32_t>(),
                                              S.val.as<uint32_t>(), S.val2.as<uint32_t>(), (size_t)S.np_tot, 0, bits,
                                              c->stream));
    AV_CHECK_HIP(c, rocprim::exclusive_scan(nullptr, t2, S.count.as<int32_t>(), S.scan.as<int32_t>(), 0,
                                            (size_t)S.nkeys, rocprim::plus<int32_t>(), c->stream));
    if (!S.sort_tmp.ensure(std::max(t1, t2))) AV_FAIL(c, AVISM_ERR_HIP, "out of device memory (batch sort)");
    t1 = t2 = S.sort_tmp.bytes;
    AV_CHECK_HIP(c, rocprim::radix_sort_pairs(S.sort_tmp.p, t1, S.key.as<uint32_t>(), S.key2.as<uint32_t>(),
                                              S.val.as<uint32_t>(), S.val2.as<uint32_t>(), (size_t)S.np_tot, 0, bits,
                                              c->stream));
    AV_CHECK_HIP(c, rocprim::exclusive_scan(S.sort_tmp.p, t2, S.count.as<int32_t>(), S.scan.as<int32_t>(), 0,
                                            (size_t)S.nkeys, rocprim::plus<int32_t>(), c->stream));
    k_batch_bin_starts<<<nb256(S.nkeys), 256, 0, c->stream>>>(S.nkeys, S.npatch, S.kbasedev.as<int64_t>(),
                                                              S.p0dev.as<int64_t>(), S.scan.as<int32_t>(),
                                                              S.start.as<int32_t>());
    AV_CHECK_HIP(c, hipGetLastError());
    AV_CHECK_HIP(c, launch_gather_sorted(c->stream, S.np_tot, S.val2.as<uint32_t>(), S.x.as<float>(), S.y.as<float>(),
                                         S.z.as<float>(), S.m.as<float>(), S.pos4.as<float4>(),
                                         S.orig.as<uint32_t>()));
  }
  AV_CHECK_HIP(c, hipStreamSynchronize(c->stream));
  S.loaded = true;
  return AVISM_OK;
}

// item offsets of the three queues -- PPART chunks, VVEL tiles of shape vs,
// DDENS tiles of shape ds -- each npatch + 1 entries, into S.offs
static int upload_offsets(avism_ctx* c, int vs, int ds, int64_t tot[3]) {
  BatchState& S = *c->batch;
  const int np1 = S.npatch + 1;
  std::vector<int64_t> h((size_t)3 * np1, 0);
  for (int p = 0; p < S.npatch; ++p) {
    h[p + 1] = h[p] + ppart_chunks(S.np[p]);
    h[np1 + p + 1] = h[np1 + p] + vvel_shape_tiles(vs, S.n[p]);
    h[2 * np1 + p + 1] = h[2 * np1 + p] + ddens_shape_tiles(ds, S.n[p]);
  }
  for (int q = 0; q < 3; ++q) tot[q] = h[(size_t)q * np1 + S.npatch];
  AV_CHECK_HIP(c, hipMemcpyAsync(S.offs.p, h.data(), h.size() * 8, hipMemcpyHostToDevice, c->stream));
  AV_CHECK_HIP(c, hipStreamSynchronize(c->stream));
  return AVISM_OK;
}

struct Queues {
  ItemQueue pp, vv, dd;
};

static Queues queues(BatchState& S, const int64_t tot[3]) {
  const int np1 = S.npatch + 1;
  const int64_t* o = S.offs.as<int64_t>();
  unsigned* q = S.qcnt.as<unsigned>();
  return Queues{ItemQueue{o, S.npatch, tot[0], q}, ItemQueue{o + np1, S.npatch, tot[1], q + 8},
                ItemQueue{o + 2 * np1, S.npatch, tot[2], q + 16}};
}

#define BATCH_TRY(expr)                  \
  do {                                   \
    hipError_t _e = (expr);              \
    if (_e != hipSuccess) return _e;     \
  } while (0)

// the field stage's launch sequence: captured into the graph, or run
// eagerly; kernels only (the fills are kernels, not memsets)
static hipError_t field_sequence(avism_ctx* c, int k, const Queues& Q) {
  BatchState& S = *c->batch;
  const hipStream_t st = c->stream;
  const PatchDev* P = S.desc.as<PatchDev>();
  BATCH_TRY(launch_fill_u32(st, S.np_tot, S.hbits.as<uint32_t>(), 0u));
  BATCH_TRY(launch_fill_u32(st, 24, S.qcnt.as<uint32_t>(), 0u));
  BATCH_TRY(launch_batch_ppart(st, P, Q.pp, k, S.ppart_blocks));
  BATCH_TRY(launch_velrec(st, S.np_tot, S.orig.as<uint32_t>(), S.pos4.as<float4>(), S.pdens.as<float>(),
                          S.vx.as<float>(), S.vy.as<float>(), S.vz.as<float>(), S.vel4.as<float4>()));
  BATCH_TRY(launch_batch_vvel(st, P, Q.vv, k, S.vvel_shape, S.vvel_blocks));
  BATCH_TRY(launch_fill_h(st, S.np_tot, S.hbits.as<float>(), S.hk.as<float>()));
  BATCH_TRY(launch_fill_u64(st, S.ncell_tot, S.acc.as<unsigned long long>(), 0ull));
  BATCH_TRY(launch_batch_ddens(st, P, Q.dd, S.ddens_shape, S.ddens_blocks, S.pairs.as<unsigned long long>()));
  for (int p = 0; p < S.npatch; ++p) {
    const GridGeom& g = S.geom[p];
    const int64_t n3 = (int64_t)g.n * g.n * g.n, cc = S.c0[p];
    // cosmology constants (avism.f90:469-479, 599-606, 1249-1250)
    const float den = ((g.dx * g.dx) * g.dx) * ((c->rete * c->rete) * c->rete);
    double* dsum = (double*)((char*)S.means.p + 16 * p);
    float* dmean = (float*)((char*)S.means.p + 16 * p + 8);
    BATCH_TRY(launch_batch_finalize(st, P, p, n3, den, c->rote, S.umass.as<float>() + cc,
                                    S.partial.as<double>() + S.part0[p]));
    BATCH_TRY(launch_reduce_mean(st, S.partial.as<double>() + S.part0[p], S.part0[p + 1] - S.part0[p], n3,
                                 S.slices.as<double>() + (size_t)512 * p, dsum, dmean));
    BATCH_TRY(launch_div_finalize(st, g.n, g.dx, g.periodic, S.u.as<float>() + cc, S.v.as<float>() + cc,
                                  S.w.as<float>() + cc, S.umass.as<float>() + cc, dmean, S.u1co.as<float>() + cc,
                                  S.div.as<float>() + cc));
  }
  return hipSuccess;
}

// milliseconds of one launch of f on the context stream
template <class F>
static int time_launch(avism_ctx* c, F&& f, float& ms) {
  hipEvent_t a, b;
  AV_CHECK_HIP(c, hipEventCreate(&a));
  AV_CHECK_HIP(c, hipEventCreate(&b));
  AV_CHECK_HIP(c, hipEventRecord(a, c->stream));
  hipError_t e = f();
  AV_CHECK_HIP(c, hipEventRecord(b, c->stream));
  AV_CHECK_HIP(c, hipEventSynchronize(b));
  AV_CHECK_HIP(c, hipEventElapsedTime(&ms, a, b));
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  if (e != hipSuccess) AV_FAIL(c, AVISM_ERR_HIP, std::string("autotuner launch: ") + hipGetErrorString(e));
  return AVISM_OK;
}

// The autotuner: on this layout's batch, each candidate kernel configuration
// runs twice (the faster run counts: the first may load code), in the order of
// the field stage, so every kernel sees the inputs it will have; the fastest
// is kept.  The persistent grid sizes are multiples of the 256 CUs.
static int autotune(avism_ctx* c, int k) {
  BatchState& S = *c->batch;
  const hipStream_t st = c->stream;
  const PatchDev* P = S.desc.as<PatchDev>();
  const unsigned grids[3] = {1024, 2048, 4096};
  std::ostringstream js;
  js << "{\"k\": " << k << ", \"patches\": " << S.npatch << ", \"particles\": " << S.np_tot
     << ", \"cells\": " << S.ncell_tot;
  int64_t tot[3];
  if (int rc = upload_offsets(c, 0, 0, tot)) return rc;
  Queues Q = queues(S, tot);
  AV_CHECK_HIP(c, launch_fill_u32(st, S.np_tot, S.hbits.as<uint32_t>(), 0u));
  auto best_of = [&](auto&& launch, float& best) -> int {
    best = 1e30f;
    for (int rep = 0; rep < 2; ++rep) {
      float ms = 0;
      if (int rc = time_launch(c, launch, ms)) return rc;
      best = std::min(best, ms);
    }
    return AVISM_OK;
  };
  // PPART_DENS: grid size
  float bestp = 1e30f;
  js << ", \"ppart\": {";
  for (int gi = 0; gi < 3; ++gi) {
    float ms;
    if (int rc = best_of([&] {
          hipError_t e = launch_fill_u32(st, 24, S.qcnt.as<uint32_t>(), 0u);
          return e != hipSuccess ? e : launch_batch_ppart(st, P, Q.pp, k, grids[gi]);
        }, ms)) return rc;
    js << (gi ? ", " : "") << "\"" << grids[gi] << "\": " << ms;
    if (ms < bestp) { bestp = ms; S.ppart_blocks = grids[gi]; }
  }
  js << "}, \"ppart_blocks\": " << S.ppart_blocks;
  AV_CHECK_HIP(c, launch_velrec(st, S.np_tot, S.orig.as<uint32_t>(), S.pos4.as<float4>(), S.pdens.as<float>(),
                                S.vx.as<float>(), S.vy.as<float>(), S.vz.as<float>(), S.vel4.as<float4>()));
  // VVEL: tile shape x grid size
  float bestv = 1e30f;
  js << ", \"vvel\": {";
  for (int sh = 0; sh < kVvelShapes; ++sh) {
    if (int rc = upload_offsets(c, sh, 0, tot)) return rc;
    Q = queues(S, tot);
    for (int gi = 0; gi < 3; ++gi) {
      float ms;
      if (int rc = best_of([&] {
            hipError_t e = launch_fill_u32(st, 24, S.qcnt.as<uint32_t>(), 0u);
            return e != hipSuccess ? e : launch_batch_vvel(st, P, Q.vv, k, sh, grids[gi]);
          }, ms)) return rc;
      js << (sh || gi ? ", " : "") << "\"" << vvel_shape_name(sh) << "/" << grids[gi] << "\": " << ms;
      if (ms < bestv) { bestv = ms; S.vvel_shape = sh; S.vvel_blocks = grids[gi]; }
    }
  }
  js << "}, \"vvel_shape\": \"" << vvel_shape_name(S.vvel_shape) << "\", \"vvel_blocks\": " << S.vvel_blocks;
  AV_CHECK_HIP(c, launch_fill_h(st, S.np_tot, S.hbits.as<float>(), S.hk.as<float>()));
  // DDENS: tile shape x grid size
  float bestd = 1e30f;
  js << ", \"ddens\": {";
  for (int sh = 0; sh < kDdensShapes; ++sh) {
    if (int rc = upload_offsets(c, 0, sh, tot)) return rc;
    Q = queues(S, tot);
    for (int gi = 0; gi < 2; ++gi) {
      float ms;
      if (int rc = best_of([&] {
            hipError_t e = launch_fill_u32(st, 24, S.qcnt.as<uint32_t>(), 0u);
            if (e == hipSuccess) e = launch_fill_u64(st, S.ncell_tot, S.acc.as<unsigned long long>(), 0ull);
            return e != hipSuccess ? e : launch_batch_ddens(st, P, Q.dd, sh, grids[gi],
                                                            S.pairs.as<unsigned long long>());
          }, ms)) return rc;
      js << (sh || gi ? ", " : "") << "\"" << ddens_shape_name(sh) << "/" << grids[gi] << "\": " << ms;
      if (ms < bestd) { bestd = ms; S.ddens_shape = sh; S.ddens_blocks = grids[gi]; }
    }
  }
  js << "}, \"ddens_shape\": \"" << ddens_shape_name(S.ddens_shape) << "\", \"ddens_blocks\": " << S.ddens_blocks
     << "}";
  S.tuning_json = js.str();
  S.k_tuned = k;
  return AVISM_OK;
}

int batch_fields(avism_ctx* c, int32_t k) {
  if (int rc = batch_check(c)) return rc;
  BatchState& S = *c->batch;
  if (k <= 0) AV_FAIL(c, AVISM_ERR_INPUT, "KNEIGHBOURS must be positive");
  if (k > 64) AV_FAIL(c, AVISM_ERR_UNSUPPORTED, "batches take KNEIGHBOURS up to 64");
  for (int p = 0; p < S.npatch; ++p)
    if (k > S.np[p]) AV_FAIL(c, AVISM_ERR_INPUT, "KNEIGHBOURS exceeds the particles of patch " + std::to_string(p));
  // the first search radius of each patch: 1.3 x the k-th neighbour distance at its mean density
  for (int p = 0; p < S.npatch; ++p) {
    const double rho = (double)S.np[p] / ((double)S.L0[p] * S.L0[p] * S.L0[p]);
    S.hdesc[p].r0 = (float)(1.3 * std::cbrt(3.0 * k / (4.0 * kPi * rho)));
  }
  if (int rc = upload_desc(c)) return rc;
  S.fields_ready = false;
  if (S.k_tuned != k) {
    S.drop_graph();
    if (int rc = autotune(c, k)) return rc;
  }
  if (!S.exec || S.graph_k != k) {
    S.drop_graph();
    int64_t tot[3];
    if (int rc = upload_offsets(c, S.vvel_shape, S.ddens_shape, tot)) return rc;
    const Queues Q = queues(S, tot);
    AV_CHECK_HIP(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    const hipError_t e = field_sequence(c, k, Q);
    hipGraph_t g = nullptr;
    const hipError_t e2 = hipStreamEndCapture(c->stream, &g);
    if (e != hipSuccess || e2 != hipSuccess) {
      if (g) (void)hipGraphDestroy(g);
      AV_FAIL(c, AVISM_ERR_HIP, std::string("batch graph capture: ") + hipGetErrorString(e != hipSuccess ? e : e2));
    }
    S.graph = g;
    AV_CHECK_HIP(c, hipGraphInstantiate(&S.exec, S.graph, nullptr, nullptr, 0));
    S.graph_k = k;
  }
  {
    StageTimer t(c, "batch_fields");
    AV_CHECK_HIP(c, hipGraphLaunch(S.exec, c->stream));
  }
  ++S.graph_launches;
  AV_CHECK_HIP(c, hipEventRecord(S.ev_done, c->stream));
  std::vector<char> hm((size_t)16 * S.npatch);
  AV_CHECK_HIP(c, hipMemcpyAsync(hm.data(), S.means.p, hm.size(), hipMemcpyDeviceToHost, c->stream));
  AV_CHECK_HIP(c, hipStreamSynchronize(c->stream));
  const float ul3 = (kUL * kUL) * kUL;
  for (int p = 0; p < S.npatch; ++p) {
    float mr;
    memcpy(&mr, hm.data() + 16 * p + 8, 4);
    S.meandens_raw[p] = mr;
    S.meandens[p] = mr * (kUM / ul3);  // avism.f90:1409
  }
  S.fields_ready = true;
  return AVISM_OK;
}

// patch p's void stage on its own context (and stream)
static int patch_voids(avism_ctx* c, int p, int growth_mode, std::string& err) {
  BatchState& S = *c->batch;
  auto fail = [&](avism_ctx* s, int rc) {
    err = "patch " + std::to_string(p) + ": " + (s ? s->err : std::string("context"));
    return rc;
  };
  if (hipSetDevice(c->device) != hipSuccess) return fail(nullptr, AVISM_ERR_HIP);
  avism_ctx*& s = S.sub[p];
  if (!s) {
    int rc = avism_create(&s, c->device);
    if (rc) return fail(nullptr, rc);
  }
  int rc = avism_set_grid(s, S.n[p], S.L0[p], S.periodic[p]);
  if (rc) return fail(s, rc);
  avism_set_cosmology(s, c->ache, c->omega0, c->zeta);
  s->vp = c->vp;
  const int64_t n3 = (int64_t)S.n[p] * S.n[p] * S.n[p];
  const size_t b = (size_t)n3 * 4;
  if (!s->u1co.ensure(b) || !s->div.ensure(b) || !s->fu.ensure(b) || !s->fv.ensure(b) || !s->fw.ensure(b) ||
      (S.has_mask && !s->smask.ensure((size_t)n3)))
    return fail(s, AVISM_ERR_HIP);
  // the batch's fields, once its field stage is complete
  if (hipStreamWaitEvent(s->stream, S.ev_done, 0) != hipSuccess) return fail(s, AVISM_ERR_HIP);
  const int64_t cc = S.c0[p];
  void* dst[5] = {s->u1co.p, s->div.p, s->fu.p, s->fv.p, s->fw.p};
  const float* src[5] = {S.u1co.as<float>() + cc, S.div.as<float>() + cc, S.u.as<float>() + cc,
                         S.v.as<float>() + cc, S.w.as<float>() + cc};
  for (int i = 0; i < 5; ++i)
    if (hipMemcpyAsync(dst[i], src[i], b, hipMemcpyDeviceToDevice, s->stream) != hipSuccess)
      return fail(s, AVISM_ERR_HIP);
  if (S.has_mask) {
    if (hipMemcpyAsync(s->smask.p, S.smask.as<int8_t>() + cc, (size_t)n3, hipMemcpyDeviceToDevice, s->stream) !=
        hipSuccess)
      return fail(s, AVISM_ERR_HIP);
    s->has_mask = true;
  }
  if (c->cnn_ready) {
    if (!s->cnn_params.ensure((size_t)AVISM_CNN_NPARAM * 4) ||
        hipMemcpyAsync(s->cnn_params.p, c->cnn_params.p, (size_t)AVISM_CNN_NPARAM * 4, hipMemcpyDeviceToDevice,
                       s->stream) != hipSuccess)
      return fail(s, AVISM_ERR_HIP);
    s->cnn_ready = true;
  }
  s->meandens = S.meandens[p];
  s->meandens_raw = S.meandens_raw[p];
  s->np = 0;
  s->fields_ready = true;
  rc = void_find(s, growth_mode);
  if (rc) return fail(s, rc);
  rc = merge_and_post(s, 0);
  if (rc) return fail(s, rc);
  return AVISM_OK;
}

int batch_voids(avism_ctx* c, int32_t growth_mode, int32_t nthreads) {
  if (int rc = batch_check(c)) return rc;
  BatchState& S = *c->batch;
  if (!S.fields_ready) AV_FAIL(c, AVISM_ERR_STATE, "batch fields not computed (avism_batch_fields)");
  if (growth_mode == AVISM_GROW_PROPOSAL) AV_FAIL(c, AVISM_ERR_ARG, "batches grow plain or learned");
  if (growth_mode == AVISM_GROW_LEARNED && !c->cnn_ready)
    AV_FAIL(c, AVISM_ERR_STATE, "learned growth needs the network (avism_cnn_set_params)");
  S.sub.resize(S.npatch, nullptr);
  const int T = std::max(1, std::min<int>(nthreads > 0 ? nthreads : 8, S.npatch));
  std::atomic<int> next{0};
  std::mutex mu;
  int first_rc = AVISM_OK;
  std::string first_err;
  auto work = [&] {
    for (;;) {
      const int p = next.fetch_add(1);
      if (p >= S.npatch) return;
      std::string err;
      const int rc = patch_voids(c, p, growth_mode, err);
      if (rc) {
        std::lock_guard<std::mutex> lk(mu);
        if (!first_rc) {
          first_rc = rc;
          first_err = err;
        }
        next.store(S.npatch);  // stop handing out patches
      }
    }
  };
  std::vector<std::thread> pool;
  for (int t = 1; t < T; ++t) pool.emplace_back(work);
  work();
  for (auto& th : pool) th.join();
  if (first_rc) AV_FAIL(c, first_rc, first_err);
  return AVISM_OK;
}

int batch_patch(avism_ctx* c, int32_t p, avism_ctx** out) {
  if (int rc = batch_check(c)) return rc;
  BatchState& S = *c->batch;
  if (!out) AV_FAIL(c, AVISM_ERR_ARG, "null output");
  if (p < 0 || p >= S.npatch) AV_FAIL(c, AVISM_ERR_ARG, "patch index out of range");
  if ((int)S.sub.size() <= p || !S.sub[p]) AV_FAIL(c, AVISM_ERR_STATE, "the patch's void stage has not run");
  *out = S.sub[p];
  return AVISM_OK;
}

int batch_get_fields(avism_ctx* c, int32_t p, float* u1co, float* div, float* u, float* v, float* w,
                     float* part_dens, float* hpart, int where) {
  if (int rc = batch_check(c)) return rc;
  BatchState& S = *c->batch;
  if (!S.fields_ready) AV_FAIL(c, AVISM_ERR_STATE, "batch fields not computed");
  if (p < 0 || p >= S.npatch) AV_FAIL(c, AVISM_ERR_ARG, "patch index out of range");
  const int64_t n3 = (int64_t)S.n[p] * S.n[p] * S.n[p], cc = S.c0[p], a = S.p0[p];
  const hipMemcpyKind kind = where == AVISM_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
  float* dst[5] = {u1co, div, u, v, w};
  const float* src[5] = {S.u1co.as<float>() + cc, S.div.as<float>() + cc, S.u.as<float>() + cc,
                         S.v.as<float>() + cc, S.w.as<float>() + cc};
  for (int i = 0; i < 5; ++i)
    if (dst[i]) AV_CHECK_HIP(c, hipMemcpyAsync(dst[i], src[i], (size_t)n3 * 4, kind, c->stream));
  // per-particle fields in the patch's input order
  float* pdst[2] = {part_dens, hpart};
  const float* psrc[2] = {S.pdens.as<float>() + a, S.hbits.as<float>() + a};
  float* tmp = S.key.as<float>();  // the sort keys are no longer needed
  for (int i = 0; i < 2; ++i) {
    if (!pdst[i]) continue;
    AV_CHECK_HIP(c, launch_scatter_orig(c->stream, S.np[p], S.orig.as<uint32_t>() + a, a, psrc[i], tmp));
    AV_CHECK_HIP(c, hipMemcpyAsync(pdst[i], tmp, (size_t)S.np[p] * 4, kind, c->stream));
    AV_CHECK_HIP(c, hipStreamSynchronize(c->stream));
  }
  AV_CHECK_HIP(c, hipStreamSynchronize(c->stream));
  return AVISM_OK;
}

int batch_info(avism_ctx* c, char* tuning, int32_t len, int64_t* graph_launches, float* meandens) {
  if (int rc = batch_check(c, false)) return rc;
  BatchState& S = *c->batch;
  if (tuning && len > 0) {
    const size_t nb = std::min((size_t)len - 1, S.tuning_json.size());
    memcpy(tuning, S.tuning_json.data(), nb);
    tuning[nb] = 0;
  }
  if (graph_launches) *graph_launches = S.graph_launches;
  if (meandens)
    for (int p = 0; p < S.npatch; ++p) meandens[p] = S.meandens[p];
  return AVISM_OK;
}

}  // namespace avism
