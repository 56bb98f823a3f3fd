// capi.hip -- the extern "C" boundary declared in include/avism.h.
#include <cstring>

#include "avism_internal.h"

namespace avism {
int load_particles(avism_ctx* c, int64_t np, const float* x, const float* y, const float* z,
                   const float* vx, const float* vy, const float* vz, const float* m, int where);
int get_particle_fields(avism_ctx* c, float* part_dens, float* hpart);
int set_proposals(avism_ctx* c, const int8_t* r, int64_t n, int where);
int growth_counters(avism_ctx* c);
int get_candidates(avism_ctx* c, int64_t* lin, int32_t* r, int64_t cap, int64_t* n);
int get_cubes(avism_ctx* c, int32_t* ci, float* cf, int32_t* father);
int get_catalogue(avism_ctx* c, int32_t* indice, int32_t* uvoid, int32_t* ncellv, float* vf,
                  avism_level_stats* st);
int get_marca(avism_ctx* c, int32_t* marca, int32_t inner_only, int32_t where);
}  // namespace avism

using namespace avism;

#define CTX_GUARD(ctx)                      \
  if (!(ctx)) return AVISM_ERR_ARG;         \
  (ctx)->err.clear();                       \
  if (hipSetDevice((ctx)->device) != hipSuccess) { (ctx)->err = "hipSetDevice failed"; return AVISM_ERR_HIP; }

// Device pointers handed in may still be written by work the caller queued on
// other streams (the library's stream does not wait on them): wait for the
// whole device before reading or writing them.
#define DEVICE_ARGS(ctx, where)                                           \
  if ((where) == AVISM_DEVICE && hipDeviceSynchronize() != hipSuccess) { \
    (ctx)->err = "hipDeviceSynchronize failed";                          \
    return AVISM_ERR_HIP;                                                \
  }

extern "C" {

int avism_create(avism_ctx** out, int32_t device) {
  if (!out) return AVISM_ERR_ARG;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return AVISM_ERR_HIP;
  if (device < 0 || device >= ndev) return AVISM_ERR_ARG;
  if (hipSetDevice(device) != hipSuccess) return AVISM_ERR_HIP;
  avism_ctx* c = new avism_ctx();
  c->device = device;
  // both streams at the default priority
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
      hipStreamCreateWithFlags(&c->cstream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&c->ev_vel, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&c->ev_mark, hipEventDisableTiming) != hipSuccess) {
    if (c->stream) (void)hipStreamDestroy(c->stream);
    if (c->cstream) (void)hipStreamDestroy(c->cstream);
    delete c;
    return AVISM_ERR_HIP;
  }
  // the static scratch of the allocation-free entry points, on this device
  if (noalloc_init(device) != AVISM_OK) {
    (void)hipEventDestroy(c->ev_vel);
    (void)hipEventDestroy(c->ev_mark);
    (void)hipStreamDestroy(c->stream);
    (void)hipStreamDestroy(c->cstream);
    delete c;
    return AVISM_ERR_HIP;
  }
  c->vp = avism_void_params{-0.6f, 10.0f, 0.25f, 0.0f, 1.0f, 1.0f, 30000000};
  avism_set_cosmology(c, 0.678f, 0.31f, 0.0f);
  *out = c;
  return AVISM_OK;
}

static void drop_timings(avism_ctx* c) {
  for (auto& t : c->timings) {
    (void)hipEventDestroy(t.a);
    (void)hipEventDestroy(t.b);
  }
  c->timings.clear();
}

int avism_destroy(avism_ctx* c) {
  if (!c) return AVISM_OK;
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  (void)hipStreamSynchronize(c->cstream);
  batch_free(c);
  refine_free(c);
  amr_free(c);
  drop_timings(c);
  (void)hipEventDestroy(c->ev_vel);
  (void)hipEventDestroy(c->ev_mark);
  (void)hipStreamDestroy(c->cstream);
  (void)hipStreamDestroy(c->stream);
  delete c;
  return AVISM_OK;
}

const char* avism_last_error(avism_ctx* c) { return c ? c->err.c_str() : "null context"; }

int avism_set_timing(avism_ctx* c, int32_t on) {
  CTX_GUARD(c);
  c->timing_on = on != 0;
  drop_timings(c);
  return AVISM_OK;
}

int avism_get_timings(avism_ctx* c, char* names, int32_t names_len, float* ms, int32_t* count) {
  CTX_GUARD(c);
  if (!count) AV_FAIL(c, AVISM_ERR_ARG, "count must not be NULL");
  const int32_t cap = *count;
  const int32_t nwr = (int32_t)c->timings.size();
  std::string all;
  for (auto& t : c->timings) {
    all += t.name;
    all += ';';
  }
  *count = nwr;
  // too small a buffer: nothing is consumed, the caller retries with *count
  // entries and a names buffer of the size the message gives
  if (!ms || cap < nwr || !names || (size_t)names_len < all.size() + 1) {
    c->err = "timings buffer too small: " + std::to_string(nwr) + " entries, " +
             std::to_string(all.size() + 1) + " name bytes";
    return AVISM_ERR_LIMIT;
  }
  for (int32_t i = 0; i < nwr; ++i) {
    float tm = 0.0f;
    AV_CHECK_HIP(c, hipEventSynchronize(c->timings[i].b));
    AV_CHECK_HIP(c, hipEventElapsedTime(&tm, c->timings[i].a, c->timings[i].b));
    ms[i] = tm;
  }
  memcpy(names, all.data(), all.size());
  names[all.size()] = 0;
  drop_timings(c);
  return AVISM_OK;
}

int avism_set_grid(avism_ctx* c, int32_t n, float L0, int32_t periodic) {
  CTX_GUARD(c);
  if (n < 4) AV_FAIL(c, AVISM_ERR_ARG, "grid size must be at least 4 cells per axis");
  if (!(L0 > 0.0f)) AV_FAIL(c, AVISM_ERR_ARG, "box size must be positive");
  if (periodic != 0 && periodic != 1) AV_FAIL(c, AVISM_ERR_ARG, "periodic must be 0 or 1");
  if (int rc = finish_copies(c)) return rc;
  // the new geometry is validated and its scalars allocated before any state
  // changes: a rejected grid leaves the context unusable (AVISM_ERR_STATE on
  // every later stage call) rather than holding extents its buffers lack
  GridGeom g = make_grid(n, L0, periodic);
  refine_free(c);
  c->grid_set = false;
  c->fields_ready = false;
  c->post_ready = false;
  const int64_t next = g.next;
  if (next * next * next >= (int64_t)0x7fffffff)
    AV_FAIL(c, AVISM_ERR_LIMIT, "extended grid above 2^31 cells is not supported by this build");
  // small device scalars: counters and sums in the first 4 KiB, the cell-centre
  // table RADX from byte 4096 on (read by the field and shape kernels)
  if (!c->scalars.ensure(avism::scalars_bytes(g)))
    AV_FAIL(c, AVISM_ERR_HIP, "out of device memory (scalars)");
  c->g = std::move(g);
  c->grid_set = true;
  c->nrand = 0;  // survey randoms belong to the previous grid's box
  c->fields_ready = false;
  c->post_ready = false;
  c->np = 0;
  c->has_marcap = false;
  c->has_reqp = false;
  c->nreqp = 0;
  c->has_mask = false;
  return AVISM_OK;
}

int avism_get_grid(avism_ctx* c, avism_grid_info* out, float* rad_ext) {
  CTX_GUARD(c);
  if (!c->grid_set) AV_FAIL(c, AVISM_ERR_STATE, "no grid set");
  if (out) {
    out->n = c->g.n; out->periodic = c->g.periodic; out->nplus = c->g.nplus;
    out->low1 = c->g.low1; out->low2 = c->g.low2; out->next = c->g.next;
    out->L0 = c->g.L0; out->lado0plus = c->g.lado0plus; out->dx = c->g.dx; out->rx1 = c->g.rx1;
  }
  if (rad_ext)
    for (int i = c->g.low1; i <= c->g.low2; i++) rad_ext[i - c->g.low1] = c->g.radx(i);
  return AVISM_OK;
}

int avism_set_cosmology(avism_ctx* c, float ache, float omega0, float zeta) {
  if (!c) return AVISM_ERR_ARG;
  c->ache = ache; c->omega0 = omega0; c->zeta = zeta;
  // avism.f90:469-479 and 599-606
  float ache0 = (float)((double)ache * 3.66e-3);
  float rodo = (float)(((double)omega0 * 3.0) * (double)(ache0 * ache0));
  float re0 = 1.0f / 10.98f;
  if (zeta < 0.0f) zeta = 0.0f;
  float t = 1.0f + zeta;
  c->rote = rodo * ((t * t) * t);
  c->rete = re0 / t;
  float r = re0 / c->rete;
  float h = omega0 * (((r * r) * r) - 1.0f) + 1.0f;
  c->hte = ache * sqrtf(h);
  return AVISM_OK;
}

int avism_set_void_params(avism_ctx* c, const avism_void_params* p) {
  CTX_GUARD(c);
  if (!p) AV_FAIL(c, AVISM_ERR_ARG, "null params");
  if (p->nvoid_max <= 0) AV_FAIL(c, AVISM_ERR_ARG, "nvoid_max must be positive");
  c->vp = *p;
  return AVISM_OK;
}

int avism_set_parent_map(avism_ctx* c, const int32_t* marcap, int32_t where) {
  CTX_GUARD(c);
  DEVICE_ARGS(c, where);
  if (!c->grid_set) AV_FAIL(c, AVISM_ERR_STATE, "no grid set");
  if (!marcap) {
    c->has_marcap = false;
    return AVISM_OK;
  }
  size_t b = (size_t)c->g.n * c->g.n * c->g.n * 4;
  if (!c->marcap.ensure(b)) AV_FAIL(c, AVISM_ERR_HIP, "out of device memory");
  AV_CHECK_HIP(c, hipMemcpy(c->marcap.p, marcap, b,
                            where == AVISM_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice));
  c->has_marcap = true;
  return AVISM_OK;
}

int avism_set_parent_radii(avism_ctx* c, const float* reqp, int32_t n, int32_t where) {
  CTX_GUARD(c);
  DEVICE_ARGS(c, where);
  if (!c->grid_set) AV_FAIL(c, AVISM_ERR_STATE, "no grid set");
  if (!reqp) {
    c->has_reqp = false;
    c->nreqp = 0;
    return AVISM_OK;
  }
  if (n < 0) AV_FAIL(c, AVISM_ERR_ARG, "negative parent radii count");
  if (!c->reqp.ensure((size_t)(n + 1) * 4)) AV_FAIL(c, AVISM_ERR_HIP, "out of device memory");
  if (n > 0)
    AV_CHECK_HIP(c, hipMemcpy(c->reqp.p, reqp, (size_t)n * 4,
                              where == AVISM_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice));
  c->has_reqp = true;
  c->nreqp = n;
  return AVISM_OK;
}

int avism_set_mask(avism_ctx* c, const int8_t* smask, int32_t where) {
  CTX_GUARD(c);
  DEVICE_ARGS(c, where);
  if (!c->grid_set) AV_FAIL(c, AVISM_ERR_STATE, "no grid set");
  if (!smask) {
    c->has_mask = false;
    return AVISM_OK;
  }
  size_t b = (size_t)c->g.n * c->g.n * c->g.n;
  if (!c->smask.ensure(b)) AV_FAIL(c, AVISM_ERR_HIP, "out of device memory");
  AV_CHECK_HIP(c, hipMemcpy(c->smask.p, smask, b,
                            where == AVISM_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice));
  // observed cells, for the query counter
  int64_t nobs = 0;
  if (where == AVISM_DEVICE) {
    std::vector<int8_t> h(b);
    AV_CHECK_HIP(c, hipMemcpy(h.data(), c->smask.p, b, hipMemcpyDeviceToHost));
    for (int8_t v : h) nobs += v != 0;
  } else {
    for (size_t i = 0; i < b; ++i) nobs += smask[i] != 0;
  }
  c->mask_observed = nobs;
  c->has_mask = true;
  return AVISM_OK;
}

int avism_next_level(avism_ctx* c) {
  CTX_GUARD(c);
  if (refine_grid_is_atlas(c))
    AV_FAIL(c, AVISM_ERR_STATE, "the grid is an atlas of windows: use avism_refine_plan for its next level");
  if (int rc = finish_copies(c)) return rc;
  refine_free(c);
  return next_level(c);
}

int avism_load_particles(avism_ctx* c, int64_t np, const float* x, const float* y, const float* z,
                         const float* vx, const float* vy, const float* vz, const float* m,
                         int32_t where) {
  CTX_GUARD(c);
  DEVICE_ARGS(c, where);
  if (!x || !y || !z || !m) AV_FAIL(c, AVISM_ERR_ARG, "null particle array");
  if ((vx == nullptr) != (vy == nullptr) || (vx == nullptr) != (vz == nullptr))
    AV_FAIL(c, AVISM_ERR_ARG, "velocities: all three arrays or none");
  return load_particles(c, np, x, y, z, vx, vy, vz, m, where);
}

int avism_particles_to_grid(avism_ctx* c, int32_t k, int32_t vel_interp, int32_t dens_interp) {
  CTX_GUARD(c);
  return fields_from_particles(c, k, vel_interp, dens_interp);
}

int avism_particles_to_density(avism_ctx* c, int32_t k, int32_t dens_interp) {
  CTX_GUARD(c);
  return fields_from_particles(c, k, -1, dens_interp);
}

int avism_load_randoms(avism_ctx* c, int64_t nrand, const float* x, const float* y, const float* z,
                       const float* m, int32_t where) {
  CTX_GUARD(c);
  DEVICE_ARGS(c, where);
  if (!x || !y || !z || !m) AV_FAIL(c, AVISM_ERR_ARG, "null random array");
  swap_particle_sets(c);
  const int rc = load_particles(c, nrand, x, y, z, nullptr, nullptr, nullptr, m, where);
  swap_particle_sets(c);
  c->fields_ready = false;
  return rc;
}

int avism_survey_density(avism_ctx* c, int32_t k, int32_t dens_interp) {
  CTX_GUARD(c);
  return survey_density(c, k, dens_interp);
}

int avism_load_grid_fields(avism_ctx* c, const float* delta, const float* vx, const float* vy,
                           const float* vz, int32_t where) {
  CTX_GUARD(c);
  DEVICE_ARGS(c, where);
  if (!delta || !vx || !vy || !vz) AV_FAIL(c, AVISM_ERR_ARG, "null field array");
  return fields_from_grid(c, delta, vx, vy, vz, where);
}

int avism_load_void_fields(avism_ctx* c, const float* u1co, const float* div, float meandens,
                           int32_t where) {
  CTX_GUARD(c);
  DEVICE_ARGS(c, where);
  if (!c->grid_set) AV_FAIL(c, AVISM_ERR_STATE, "avism_set_grid must be called first");
  if (!u1co || !div) AV_FAIL(c, AVISM_ERR_ARG, "null field array");
  if (int rc = finish_copies(c)) return rc;
  c->nvoid = 0;
  c->post_ready = false;
  size_t b = (size_t)c->g.n * c->g.n * c->g.n * 4;
  if (!c->u1co.ensure(b) || !c->div.ensure(b) || !c->fu.ensure(b) || !c->fv.ensure(b) || !c->fw.ensure(b))
    AV_FAIL(c, AVISM_ERR_HIP, "out of device memory");
  hipMemcpyKind kind = where == AVISM_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
  AV_CHECK_HIP(c, hipMemcpy(c->u1co.p, u1co, b, kind));
  AV_CHECK_HIP(c, hipMemcpy(c->div.p, div, b, kind));
  AV_CHECK_HIP(c, hipMemset(c->fu.p, 0, b));
  AV_CHECK_HIP(c, hipMemset(c->fv.p, 0, b));
  AV_CHECK_HIP(c, hipMemset(c->fw.p, 0, b));
  c->meandens = meandens;
  c->np = 0;
  c->fields_ready = true;
  return AVISM_OK;
}

int avism_get_fields(avism_ctx* c, float* u1co, float* div, float* u, float* v, float* w,
                     int32_t where) {
  CTX_GUARD(c);
  DEVICE_ARGS(c, where);
  if (where != AVISM_HOST && where != AVISM_DEVICE && where != AVISM_HOST_ASYNC)
    AV_FAIL(c, AVISM_ERR_ARG, "where must be AVISM_HOST, AVISM_DEVICE or AVISM_HOST_ASYNC");
  if (!c->fields_ready) AV_FAIL(c, AVISM_ERR_STATE, "fields not computed");
  size_t b = (size_t)c->g.n * c->g.n * c->g.n * 4;
  hipMemcpyKind kind = where == AVISM_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
  float* dst[5] = {u1co, div, u, v, w};
  void* src[5] = {c->u1co.p, c->div.p, c->fu.p, c->fv.p, c->fw.p};
  if (where == AVISM_HOST_ASYNC) {  // on the copy stream, after the work queued so far
    AV_CHECK_HIP(c, hipEventRecord(c->ev_mark, c->stream));
    AV_CHECK_HIP(c, hipStreamWaitEvent(c->cstream, c->ev_mark, 0));
    StageTimer t(c, "download_fields", c->cstream);
    for (int i = 0; i < 5; i++)
      if (dst[i]) AV_CHECK_HIP(c, hipMemcpyAsync(dst[i], src[i], b, kind, c->cstream));
    c->copy_pending = true;
    return AVISM_OK;
  }
  for (int i = 0; i < 5; i++)
    if (dst[i]) AV_CHECK_HIP(c, hipMemcpyAsync(dst[i], src[i], b, kind, c->stream));
  AV_CHECK_HIP(c, hipStreamSynchronize(c->stream));
  return AVISM_OK;
}

int avism_sync(avism_ctx* c) {
  CTX_GUARD(c);
  AV_CHECK_HIP(c, hipStreamSynchronize(c->stream));
  return finish_copies(c);
}

int avism_get_counters(avism_ctx* c, int64_t* out, int32_t n) {
  CTX_GUARD(c);
  if (!out || n < 0) AV_FAIL(c, AVISM_ERR_ARG, "bad counter request");
  if (n > 3) {
    int rc = growth_counters(c);
    if (rc) return rc;
  }
  const int32_t m = n < 8 ? n : 8;
  for (int32_t i = 0; i < m; ++i) out[i] = c->counters[i];
  for (int32_t i = m; i < n; ++i) out[i] = 0;
  return AVISM_OK;
}

int avism_get_particle_fields(avism_ctx* c, float* part_dens, float* hpart) {
  CTX_GUARD(c);
  return get_particle_fields(c, part_dens, hpart);
}

int avism_find_voids(avism_ctx* c, int32_t growth_mode, int32_t* ncubes) {
  CTX_GUARD(c);
  int rc = void_find(c, growth_mode);
  if (ncubes) *ncubes = c->nvoid;
  return rc;
}

int avism_set_proposals(avism_ctx* c, const int8_t* r_hat, int64_t n, int32_t where) {
  CTX_GUARD(c);
  DEVICE_ARGS(c, where);
  return set_proposals(c, r_hat, n, where);
}

int avism_get_candidates(avism_ctx* c, int64_t* lin, int32_t* r, int64_t cap, int64_t* n) {
  CTX_GUARD(c);
  return get_candidates(c, lin, r, cap, n);
}

int avism_get_cubes(avism_ctx* c, int32_t* ci, float* cf, int32_t* father) {
  CTX_GUARD(c);
  return get_cubes(c, ci, cf, father);
}

int avism_indexx(avism_ctx* c, const float* keys, int64_t n, int32_t* indx, int32_t where) {
  CTX_GUARD(c);
  DEVICE_ARGS(c, where);
  if (n < 0) AV_FAIL(c, AVISM_ERR_ARG, "negative length");
  if (n == 0) return AVISM_OK;
  if (n >= (int64_t)0x7fffffff) AV_FAIL(c, AVISM_ERR_LIMIT, "INDEXX beyond 2^31-1 keys");
  const float* dk = keys;
  int32_t* di = indx;
  if (where != AVISM_DEVICE) {
    if (!c->sort_a.ensure((size_t)n * 4) || !c->sort_b.ensure((size_t)n * 4))
      AV_FAIL(c, AVISM_ERR_HIP, "out of device memory");
    AV_CHECK_HIP(c, hipMemcpyAsync(c->sort_a.p, keys, (size_t)n * 4, hipMemcpyHostToDevice, c->stream));
    dk = c->sort_a.as<float>();
    di = c->sort_b.as<int32_t>();
  }
  int rc = indexx_device(c, dk, n, di);
  if (rc) return rc;
  if (where != AVISM_DEVICE) {
    AV_CHECK_HIP(c, hipMemcpyAsync(indx, di, (size_t)n * 4, hipMemcpyDeviceToHost, c->stream));
    AV_CHECK_HIP(c, hipStreamSynchronize(c->stream));
  }
  return AVISM_OK;
}

int avism_indexx_noalloc(avism_ctx* c, const float* keys, int64_t n, int32_t* indx) {
  CTX_GUARD(c);
  DEVICE_ARGS(c, AVISM_DEVICE);
  if (n > 0 && (!keys || !indx)) AV_FAIL(c, AVISM_ERR_ARG, "null key or index array");
  return indexx_noalloc(c, keys, n, indx);
}

int avism_merge_and_post(avism_ctx* c, int32_t* nvoids) {
  CTX_GUARD(c);
  int rc = merge_and_post(c, 0);
  if (nvoids) *nvoids = c->nvoidt;
  return rc;
}

int avism_get_catalogue(avism_ctx* c, int32_t* indice, int32_t* uvoid, int32_t* ncellv, float* vf,
                        avism_level_stats* stats) {
  CTX_GUARD(c);
  return get_catalogue(c, indice, uvoid, ncellv, vf, stats);
}

int avism_get_marca(avism_ctx* c, int32_t* marca, int32_t inner_only, int32_t where) {
  CTX_GUARD(c);
  DEVICE_ARGS(c, where);
  return get_marca(c, marca, inner_only, where);
}

int avism_cnn_set_params(avism_ctx* c, const float* params, int32_t where) {
  CTX_GUARD(c);
  DEVICE_ARGS(c, where);
  if (!params) AV_FAIL(c, AVISM_ERR_ARG, "null parameter array");
  const size_t b = (size_t)AVISM_CNN_NPARAM * 4;
  if (!c->cnn_params.ensure(b)) AV_FAIL(c, AVISM_ERR_HIP, "out of device memory");
  AV_CHECK_HIP(c, hipMemcpy(c->cnn_params.p, params, b,
                            where == AVISM_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice));
  c->cnn_ready = true;
  return AVISM_OK;
}

int avism_cnn_propose(avism_ctx* c, int8_t* out, int64_t n, int32_t where) {
  CTX_GUARD(c);
  DEVICE_ARGS(c, where);
  if (n != c->ncand) AV_FAIL(c, AVISM_ERR_ARG, "proposal count differs from the candidate count");
  if (n > 0 && !out) AV_FAIL(c, AVISM_ERR_ARG, "null proposal array");
  if (c->ncand <= 0) return AVISM_OK;
  int rc = cnn_propose(c);
  if (rc) return rc;
  AV_CHECK_HIP(c, hipMemcpyAsync(out, c->proposals.p, (size_t)n,
                                 where == AVISM_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost,
                                 c->stream));
  AV_CHECK_HIP(c, hipStreamSynchronize(c->stream));
  return AVISM_OK;
}

int avism_cnn_patches(avism_ctx* c, const int64_t* ranks, int64_t n, uint16_t* patches) {
  CTX_GUARD(c);
  DEVICE_ARGS(c, AVISM_DEVICE);
  if (n < 0 || (n > 0 && (!ranks || !patches))) AV_FAIL(c, AVISM_ERR_ARG, "bad patch request");
  return cnn_patches(c, ranks, n, patches);
}

int avism_cnn_forward(avism_ctx* c, const float* params, int64_t n, const uint16_t* patches, float* logits,
                      uint16_t* cache) {
  CTX_GUARD(c);
  DEVICE_ARGS(c, AVISM_DEVICE);
  if (n < 0 || !params || (n > 0 && (!patches || !logits))) AV_FAIL(c, AVISM_ERR_ARG, "bad forward request");
  return cnn_forward(c, params, n, patches, logits, cache);
}

int avism_cnn_backward(avism_ctx* c, const float* params, int64_t n, const uint16_t* patches,
                       const uint16_t* cache, const float* logits, const int32_t* labels, float* grads,
                       float* loss) {
  CTX_GUARD(c);
  DEVICE_ARGS(c, AVISM_DEVICE);
  if (n < 0 || !params || !grads || !loss || (n > 0 && (!patches || !cache || !logits || !labels)))
    AV_FAIL(c, AVISM_ERR_ARG, "bad backward request");
  return cnn_backward(c, params, n, patches, cache, logits, labels, grads, loss);
}

int avism_cnn_adam(avism_ctx* c, float* params, const float* grads, float* m, float* v, float lr,
                   int32_t step) {
  CTX_GUARD(c);
  DEVICE_ARGS(c, AVISM_DEVICE);
  if (!params || !grads || !m || !v || step < 1) AV_FAIL(c, AVISM_ERR_ARG, "bad Adam request");
  return cnn_adam(c, params, grads, m, v, lr, step);
}

int avism_batch_set(avism_ctx* c, int32_t npatch, const int32_t* n, const float* L0, const int32_t* periodic,
                    const int64_t* np) {
  CTX_GUARD(c);
  return batch_set(c, npatch, n, L0, periodic, np);
}

int avism_batch_load(avism_ctx* c, const float* x, const float* y, const float* z, const float* vx,
                     const float* vy, const float* vz, const float* m, const int8_t* smask, int32_t where) {
  CTX_GUARD(c);
  DEVICE_ARGS(c, where);
  if (!x || !y || !z || !m) AV_FAIL(c, AVISM_ERR_ARG, "null particle array");
  if ((vx == nullptr) != (vy == nullptr) || (vx == nullptr) != (vz == nullptr))
    AV_FAIL(c, AVISM_ERR_ARG, "velocities: all three arrays or none");
  return batch_load(c, x, y, z, vx, vy, vz, m, smask, where);
}

int avism_batch_fields(avism_ctx* c, int32_t k) {
  CTX_GUARD(c);
  return batch_fields(c, k);
}

int avism_batch_voids(avism_ctx* c, int32_t growth_mode, int32_t nthreads) {
  CTX_GUARD(c);
  return batch_voids(c, growth_mode, nthreads);
}

int avism_batch_patch(avism_ctx* c, int32_t p, avism_ctx** patch_ctx) {
  CTX_GUARD(c);
  return batch_patch(c, p, patch_ctx);
}

int avism_batch_get_fields(avism_ctx* c, int32_t p, float* u1co, float* div, float* u, float* v, float* w,
                           float* part_dens, float* hpart, int32_t where) {
  CTX_GUARD(c);
  DEVICE_ARGS(c, where);
  return batch_get_fields(c, p, u1co, div, u, v, w, part_dens, hpart, where);
}

int avism_batch_info(avism_ctx* c, char* tuning, int32_t len, int64_t* graph_launches, float* meandens) {
  CTX_GUARD(c);
  return batch_info(c, tuning, len, graph_launches, meandens);
}

int avism_refine_plan(avism_ctx* c, int32_t margin, int64_t max_cells, int32_t* natlas) {
  CTX_GUARD(c);
  if (int rc = finish_copies(c)) return rc;
  return refine_plan(c, margin, max_cells, natlas);
}

int avism_refine_load(avism_ctx* c, int32_t atlas, int32_t k) {
  CTX_GUARD(c);
  return refine_load(c, atlas, k);
}

int avism_refine_windows(avism_ctx* c, int32_t* tab, int32_t cap, int32_t* n, avism_refine_info* info) {
  CTX_GUARD(c);
  return refine_windows(c, tab, cap, n, info);
}

int avism_refine_pop(avism_ctx* c) {
  CTX_GUARD(c);
  return refine_pop(c);
}

int avism_amr_set(avism_ctx* c, int32_t nhyx, const float* base, int32_t npatch, const int32_t* pint,
                  const float* preal, const float* pdata, int32_t where) {
  CTX_GUARD(c);
  DEVICE_ARGS(c, where);
  if (!base) AV_FAIL(c, AVISM_ERR_ARG, "null base grid");
  return amr_set(c, nhyx, base, npatch, pint, preal, pdata, where);
}

int avism_amr_fields(avism_ctx* c, int32_t ir, int32_t flag_dens, int32_t flag_div) {
  CTX_GUARD(c);
  if (refine_grid_is_atlas(c)) AV_FAIL(c, AVISM_ERR_STATE, "the grid is an atlas of windows");
  return amr_fields(c, ir, flag_dens, flag_div);
}

int avism_run_level(avism_ctx* c, int32_t k, int32_t vel_interp, int32_t dens_interp,
                    int32_t growth_mode, avism_level_stats* stats) {
  CTX_GUARD(c);
  int rc = fields_from_particles(c, k, vel_interp, dens_interp);
  if (rc) return rc;
  rc = void_find(c, growth_mode);
  if (rc) return rc;
  rc = merge_and_post(c, 0);
  if (rc) return rc;
  if (stats) return get_catalogue(c, nullptr, nullptr, nullptr, nullptr, stats);
  return AVISM_OK;
}

}  // extern "C"
