// refine.hip -- void-in-void levels computed only in windows around the
// parent voids (avism_refine_*, include/avism.h).
//
// Reference: the level loop and hand-off of avism.f90:742-759, 957-968,
// 2123-2161 (MARCAP = MARCA upsampled by 2 per axis, REQP = REQ of the main
// voids), with MARCAP gating in the void finder (voidfinding.f90:34-36,
// 267-274, 516-517, 939).  The reference computes every level on the whole
// box; here a level's whole grid is never materialised:
//  * planning (refine_plan): each qualifying main void V of the current level
//    (REQ >= RMIN_SUB) gets a cubic window: the bounding box of V's cells at
//    twice the resolution, grown by a margin.  Windows are packed side by side
//    (3-D shelf packing, largest first) into atlases of bounded size, which
//    bounds the memory of a level by its largest atlas;
//  * an atlas (refine_load) is a non-periodic grid holding every window's
//    cells at the window's atlas origin: its fields come from the window
//    kernels of fields.hip over all loaded particles, MARCAP is V on V's
//    cells of V's window and 0 elsewhere, REQP(V) = REQ(V).  Windows are
//    separated by their margins, whose cells hold no parent, so the void
//    stage (unchanged) treats each window's parent independently;
//  * plans nest: a plan made while an atlas is loaded refines that atlas's
//    voids (the next level), so levels are walked depth first and only one
//    parent MARCA per level is held.
#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#include <numeric>
#include <vector>

#include "avism_internal.h"
#include "fields_batch.h"
#include "refine.h"

namespace avism {

int get_marca(avism_ctx* c, int32_t* marca, int32_t inner_only, int32_t where);

struct WinHost {
  int32_t parent, w;
  int32_t o[3], a[3];   // origin in the level's cells (unwrapped), in the atlas
  int32_t so[3], sa[3]; // the parent MARCA source window: origin in parent-level cells, in the source grid
  int32_t sw, swrap;
  int32_t atlas;
};

struct RefPlan {
  int32_t n_level = 0, n_parent = 0, periodic = 0;
  int32_t pside = 0;              // side of the parent MARCA copy
  DBuf pmarca;                    // the box's inner MARCA, or the parent atlas's
  std::vector<float> reqp;        // REQP by parent id (index id - 1)
  std::vector<WinHost> win;       // grouped by atlas
  std::vector<int32_t> atlas_side, atlas_first;
  // HPART over the windows of every atlas (made once when there are several)
  DBuf hpart;
  int32_t hpart_k = 0;
};

struct RefineState {
  std::vector<RefPlan> plans;
  int32_t loaded = -1;            // atlas of plans.back() on the context's grid, -1 none
  bool grid_is_atlas = false;     // the context's grid is some atlas (loaded or stale)
  float level_dx = 0, level_rx1 = 0, meandens = 0;
  int64_t window_cells = 0;
  std::vector<WinHost> lwin;      // windows of the loaded atlas
  std::vector<int32_t> win_of_parent;  // loaded atlas: parent id -> window index (-1 none)
  DBuf dwin, doff, drad, dsum, dbbox, dmap;
};

void refine_free(avism_ctx* c) {
  delete c->refine;
  c->refine = nullptr;
}

bool refine_grid_is_atlas(const avism_ctx* c) { return c->refine && c->refine->grid_is_atlas; }

namespace {

inline unsigned nbk(int64_t n, int b = 256) { return (unsigned)((n + b - 1) / b); }

inline int64_t floordiv(int64_t a, int64_t b) {  // b > 0
  int64_t q = a / b;
  return (a % b != 0 && a < 0) ? q - 1 : q;
}

// per void: [0..2] min and [3..5] max of the cell coordinates; [6..11] the same
// for the coordinates shifted by n/2 (mod n) in a periodic box, so that a void
// across a face has a compact range in one of the two
struct BBoxArgs {
  const int32_t* marca;
  int32_t n, ne, off, periodic;  // cells per axis, ext side, inner offset in the ext grid
  // atlas frame: the voids' windows translate atlas cells to the level's cells
  const int32_t* father;         // cube -> parent id
  const int32_t* wmap;           // parent id -> window index
  const WinDev* win;
  int32_t atlas;
  int32_t nv;
};

__global__ void __launch_bounds__(256) k_void_bbox(BBoxArgs A, int* bb) {
  const int64_t n = A.n;
  const int64_t c = (int64_t)blockIdx.x * 256 + threadIdx.x;
  int v = 0, x = 0, y = 0, z = 0;
  if (c < n * n * n) {
    x = (int)(c % n);
    y = (int)((c / n) % n);
    z = (int)(c / (n * n));
    v = A.marca[(size_t)(x + A.off) + (size_t)A.ne * ((size_t)(y + A.off) + (size_t)A.ne * (size_t)(z + A.off))];
    if (v > A.nv) v = 0;
    if (v > 0 && A.atlas) {
      const int p = A.father[v - 1];
      const WinDev& wd = A.win[A.wmap[p]];
      x += wd.ox - wd.ax;
      y += wd.oy - wd.ay;
      z += wd.oz - wd.az;
    }
  }
  // one atomic per wave where the wave's cells belong to one void (the usual case)
  const int v0 = __shfl(v, 0);
  const bool same = __all(v == v0);
  int q[12];
  q[0] = x; q[1] = y; q[2] = z; q[3] = x; q[4] = y; q[5] = z;
  const int h = A.n / 2;
  const int sx = A.periodic && !A.atlas ? (x + h) % A.n : x, sy = A.periodic && !A.atlas ? (y + h) % A.n : y,
            sz = A.periodic && !A.atlas ? (z + h) % A.n : z;
  q[6] = sx; q[7] = sy; q[8] = sz; q[9] = sx; q[10] = sy; q[11] = sz;
  if (same) {
    if (v0 <= 0) return;
    for (int o = 32; o > 0; o >>= 1)
      for (int j = 0; j < 12; ++j) {
        const int t = __shfl_xor(q[j], o);
        q[j] = (j % 6) < 3 ? min(q[j], t) : max(q[j], t);
      }
    if ((threadIdx.x & 63) == 0)
      for (int j = 0; j < 12; ++j)
        (j % 6) < 3 ? atomicMin(&bb[12 * (v0 - 1) + j], q[j]) : atomicMax(&bb[12 * (v0 - 1) + j], q[j]);
    return;
  }
  if (v <= 0) return;
  for (int j = 0; j < 12; ++j)
    (j % 6) < 3 ? atomicMin(&bb[12 * (v - 1) + j], q[j]) : atomicMax(&bb[12 * (v - 1) + j], q[j]);
}

__global__ void k_init_bbox(int64_t n, int* bb) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) bb[i] = (i % 6) < 3 ? INT_MAX : INT_MIN;
}

// MEANDENS of the level: (total mass / den) / cells, from the slices' sums in
// a fixed order
__global__ void k_refine_mean(const double* partial, int nparts, double den, double ncell, double* sum_out,
                              float* mean_out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double s = 0.0;
  for (int i = 0; i < nparts; ++i) s += partial[i];
  *sum_out = s / den;
  *mean_out = (float)((s / den) / ncell);
}

// 3-D shelf packing of cubes of sides w (sorted, largest first) into atlases
// of side S: positions and atlas index per window; returns the atlas count
int shelf_pack(const std::vector<int32_t>& w, int32_t S, std::vector<int32_t>& pos, std::vector<int32_t>& atl,
               std::vector<int32_t>& side) {
  const size_t m = w.size();
  pos.assign(3 * m, 0);
  atl.assign(m, 0);
  side.clear();
  int32_t x = 0, y = 0, z = 0, rh = 0, lh = 0, ext = 0, a = 0;
  for (size_t i = 0; i < m; ++i) {
    const int32_t s = w[i];
    if (x + s > S) {  // next row
      y += rh;
      x = 0;
      rh = 0;
    }
    if (y + s > S) {  // next layer
      z += lh;
      y = 0;
      x = 0;
      rh = 0;
      lh = 0;
    }
    if (z + s > S) {  // next atlas
      side.push_back(std::max(ext, 4));
      ++a;
      x = y = z = rh = lh = ext = 0;
    }
    pos[3 * i] = x;
    pos[3 * i + 1] = y;
    pos[3 * i + 2] = z;
    atl[i] = a;
    x += s;
    rh = std::max(rh, s);
    lh = std::max(lh, s);
    ext = std::max(ext, std::max(x, std::max(y + s, z + s)));
  }
  if (m) side.push_back(std::max(ext, 4));
  return (int)side.size();
}

}  // namespace

int refine_plan(avism_ctx* c, int32_t margin, int64_t max_cells, int32_t* natlas) {
  if (natlas) *natlas = 0;
  if (margin < 4) AV_FAIL(c, AVISM_ERR_ARG, "the window margin must be at least 4 cells");
  if (max_cells == 0) max_cells = (int64_t)1 << 30;
  if (max_cells < 64 || max_cells > (int64_t)0x7fffffff)
    AV_FAIL(c, AVISM_ERR_ARG, "max_cells must lie in 64 .. 2^31-1");
  if (!c->post_ready) AV_FAIL(c, AVISM_ERR_STATE, "avism_merge_and_post must complete before avism_refine_plan");
  if (c->has_mask) AV_FAIL(c, AVISM_ERR_UNSUPPORTED, "refined windows do not support a survey mask");
  if (!c->refine) c->refine = new RefineState();
  RefineState& R = *c->refine;
  const bool from_atlas = R.grid_is_atlas;
  if (from_atlas && R.loaded < 0)
    AV_FAIL(c, AVISM_ERR_STATE, "the context holds an atlas of a dropped plan: load an atlas first");
  const GridGeom& g = c->g;
  const int32_t nv = c->nvoid;
  RefPlan P;
  P.periodic = c->box_periodic;
  P.n_parent = from_atlas ? R.plans.back().n_level : g.n;
  if ((int64_t)P.n_parent * 2 > (int64_t)1 << 28) AV_FAIL(c, AVISM_ERR_LIMIT, "level grid too fine");
  P.n_level = 2 * P.n_parent;
  if (nv > 0) {
    // the catalogue: main voids and their REQ (vf column 11)
    std::vector<int32_t> uvoid(nv);
    std::vector<float> req(nv);
    AV_CHECK_HIP(c, hipMemcpy(uvoid.data(), c->uvoid.p, (size_t)nv * 4, hipMemcpyDeviceToHost));
    AV_CHECK_HIP(c, hipMemcpy(req.data(), c->vf.as<float>() + 11 * (size_t)nv, (size_t)nv * 4,
                              hipMemcpyDeviceToHost));
    // bounding boxes of the voids' cells, in the level's cells of this grid
    if (!R.dbbox.ensure((size_t)nv * 12 * 4)) AV_FAIL(c, AVISM_ERR_HIP, "out of device memory (refine)");
    k_init_bbox<<<nbk((int64_t)nv * 12), 256, 0, c->stream>>>((int64_t)nv * 12, R.dbbox.as<int>());
    BBoxArgs A{};
    A.marca = c->marca.as<int32_t>();
    A.n = g.n;
    A.ne = g.next;
    A.off = 1 - g.low1;
    A.periodic = g.periodic;
    A.nv = nv;
    if (from_atlas) {
      A.atlas = 1;
      A.father = c->cube_father.as<int32_t>();
      A.wmap = R.dmap.as<int32_t>();
      A.win = R.dwin.as<WinDev>();
    }
    k_void_bbox<<<nbk((int64_t)g.n * g.n * g.n), 256, 0, c->stream>>>(A, R.dbbox.as<int>());
    AV_CHECK_HIP(c, hipGetLastError());
    std::vector<int> bb((size_t)nv * 12);
    AV_CHECK_HIP(c, hipMemcpyAsync(bb.data(), R.dbbox.p, bb.size() * 4, hipMemcpyDeviceToHost, c->stream));
    AV_CHECK_HIP(c, hipStreamSynchronize(c->stream));
    std::vector<int32_t> fathers;
    if (from_atlas) {
      fathers.resize(nv);
      AV_CHECK_HIP(c, hipMemcpy(fathers.data(), c->cube_father.p, (size_t)nv * 4, hipMemcpyDeviceToHost));
    }
    P.reqp.assign(nv, 0.0f);
    const int32_t nl = P.n_level;
    for (int32_t v = 0; v < nv; ++v) {
      if (uvoid[v] != -1) continue;
      P.reqp[v] = req[v];  // avism.f90:2140-2144
      const int* q = &bb[(size_t)v * 12];
      if (q[0] > q[3]) continue;             // no cells left (removed by the filters)
      if (!(req[v] >= c->vp.rmin_sub)) continue;  // FLAG1: no subvoid (voidfinding.f90:267-274)
      WinHost W{};
      W.parent = v + 1;
      int32_t lo[3], hi[3];
      for (int ax = 0; ax < 3; ++ax) {
        lo[ax] = q[ax];
        hi[ax] = q[3 + ax];
        if (!from_atlas && g.periodic && q[9 + ax] - q[6 + ax] < hi[ax] - lo[ax]) {
          lo[ax] = q[6 + ax] - g.n / 2;  // across the face: unwrapped range
          hi[ax] = q[9 + ax] - g.n / 2;
        }
        lo[ax] = 2 * lo[ax];
        hi[ax] = 2 * hi[ax] + 1;
      }
      int32_t span = 0;
      for (int ax = 0; ax < 3; ++ax) span = std::max(span, hi[ax] - lo[ax] + 1);
      W.w = span + 2 * margin;
      for (int ax = 0; ax < 3; ++ax) W.o[ax] = lo[ax] - margin - (W.w - (hi[ax] - lo[ax] + 1 + 2 * margin)) / 2;
      if (W.w >= nl) {
        W.w = nl;
        if (!P.periodic)
          for (int ax = 0; ax < 3; ++ax) W.o[ax] = 0;
      } else if (!P.periodic) {
        for (int ax = 0; ax < 3; ++ax) W.o[ax] = std::min(std::max(W.o[ax], 0), nl - W.w);
      }
      if (from_atlas) {  // V's cells lie in the window of V's own parent
        const int32_t fp = fathers[v];
        if (fp <= 0 || fp >= (int32_t)R.win_of_parent.size() || R.win_of_parent[fp] < 0)
          AV_FAIL(c, AVISM_ERR_STATE, "a void of the atlas has no window");
        const int32_t pw = R.win_of_parent[fp];
        const WinHost& S = R.lwin[pw];
        for (int ax = 0; ax < 3; ++ax) {
          W.so[ax] = S.o[ax];
          W.sa[ax] = S.a[ax];
        }
        W.sw = S.w;
        W.swrap = 0;
      } else {
        W.sw = g.n;
        W.swrap = g.periodic;
      }
      P.win.push_back(W);
    }
  }
  // pack the windows, largest first, into as few atlases as max_cells allows
  std::stable_sort(P.win.begin(), P.win.end(), [](const WinHost& a, const WinHost& b) { return a.w > b.w; });
  const int32_t amax = (int32_t)std::floor(std::cbrt((double)max_cells) + 1e-9);
  int64_t total = 0;
  std::vector<int32_t> ws;
  for (const WinHost& W : P.win) {
    if (W.w > amax)
      AV_FAIL(c, AVISM_ERR_LIMIT, "a window of " + std::to_string(W.w) + "^3 cells exceeds max_cells");
    ws.push_back(W.w);
    total += (int64_t)W.w * W.w * W.w;
  }
  if (!P.win.empty()) {
    int32_t S = std::max(ws[0], (int32_t)std::ceil(std::cbrt(1.25 * (double)total)));
    S = std::min(S, amax);
    std::vector<int32_t> pos, atl, side;
    for (;;) {
      const int na = shelf_pack(ws, S, pos, atl, side);
      if (na == 1 || S >= amax) break;
      S = std::min(amax, S + std::max(1, S / 10));
    }
    for (size_t i = 0; i < P.win.size(); ++i) {
      for (int ax = 0; ax < 3; ++ax) P.win[i].a[ax] = pos[3 * i + ax];
      P.win[i].atlas = atl[i];
    }
    P.atlas_side = side;
    P.atlas_first.assign(side.size() + 1, 0);
    for (size_t i = 0; i < P.win.size(); ++i) P.atlas_first[atl[i] + 1]++;
    for (size_t a = 0; a < side.size(); ++a) P.atlas_first[a + 1] += P.atlas_first[a];
  }
  // the parent map: the box's inner MARCA, or the loaded atlas's
  P.pside = g.n;
  if (!P.win.empty()) {
    if (!P.pmarca.ensure((size_t)g.n * g.n * g.n * 4)) AV_FAIL(c, AVISM_ERR_HIP, "out of device memory (refine)");
    AV_CHECK_HIP(c, hipStreamSynchronize(c->stream));
    if (int rc = get_marca(c, P.pmarca.as<int32_t>(), 1, AVISM_DEVICE)) return rc;
  }
  R.plans.push_back(std::move(P));
  R.loaded = -1;
  if (natlas) *natlas = (int32_t)R.plans.back().atlas_side.size();
  return AVISM_OK;
}

int refine_load(avism_ctx* c, int32_t atlas, int32_t k) {
  RefineState* Rp = c->refine;
  if (!Rp || Rp->plans.empty()) AV_FAIL(c, AVISM_ERR_STATE, "no refinement plan (avism_refine_plan)");
  RefineState& R = *Rp;
  RefPlan& P = R.plans.back();
  if (atlas < 0 || atlas >= (int32_t)P.atlas_side.size()) AV_FAIL(c, AVISM_ERR_ARG, "atlas index out of range");
  if (k <= 0 || k > 256) AV_FAIL(c, AVISM_ERR_ARG, "KNEIGHBOURS must lie in 1..256");
  if (c->np <= 0 || !c->ppart_valid)
    AV_FAIL(c, AVISM_ERR_STATE, "avism_particles_to_grid with SPH velocities must run on the loaded particles first");
  if (k > c->np) AV_FAIL(c, AVISM_ERR_INPUT, "KNEIGHBOURS exceeds the number of particles");
  if (int rc = finish_copies(c)) return rc;
  const int32_t A = P.atlas_side[atlas];
  const int64_t ncell = (int64_t)A * A * A;
  const int64_t np = c->np;
  // the level's whole grid (never stored) and the atlas grid
  const GridGeom gl = make_grid(P.n_level, c->box_L0, c->box_periodic);
  GridGeom ga = make_grid(A, (float)A * gl.dx, 0);
  // windows of this atlas
  const int32_t w0 = P.atlas_first[atlas], w1 = P.atlas_first[atlas + 1], nw = w1 - w0;
  std::vector<WinHost> lw(P.win.begin() + w0, P.win.begin() + w1);
  std::vector<WinDev> wd(nw);
  int64_t wcells = 0;
  for (int32_t i = 0; i < nw; ++i) {
    const WinHost& W = lw[i];
    WinDev& d = wd[i];
    memset(&d, 0, sizeof d);
    d.ox = W.o[0]; d.oy = W.o[1]; d.oz = W.o[2]; d.w = W.w;
    d.ax = W.a[0]; d.ay = W.a[1]; d.az = W.a[2]; d.parent = W.parent;
    d.sox = W.so[0]; d.soy = W.so[1]; d.soz = W.so[2]; d.sw = W.sw;
    d.sax = W.sa[0]; d.say = W.sa[1]; d.saz = W.sa[2]; d.swrap = W.swrap;
    wcells += (int64_t)W.w * W.w * W.w;
  }
  // buffers (the context's field and parent-map buffers, sized for the atlas)
  if (!c->fu.ensure((size_t)ncell * 4) || !c->fv.ensure((size_t)ncell * 4) || !c->fw.ensure((size_t)ncell * 4) ||
      !c->u_mass.ensure((size_t)ncell * 4) || !c->u1co.ensure((size_t)ncell * 4) ||
      !c->div.ensure((size_t)ncell * 4) || !c->mass_fix.ensure((size_t)ncell * 8) ||
      !c->marcap.ensure((size_t)ncell * 4) || !c->partial64.ensure((size_t)finalize_blocks(ncell) * 8) ||
      !c->hpart.ensure((size_t)np * 4) || !c->norm.ensure((size_t)np * 4) || !c->pmark.ensure((size_t)np) ||
      !R.dwin.ensure(sizeof(WinDev) * (size_t)std::max(nw, 1)) ||
      !R.drad.ensure(gl.rad.size() * 4) || !R.dsum.ensure(1024 * 8) ||
      !c->reqp.ensure((P.reqp.size() + 1) * 4))
    AV_FAIL(c, AVISM_ERR_HIP, "out of device memory (refine)");
  c->nvoid = 0;
  c->ncand = 0;
  c->post_ready = false;
  c->fields_ready = false;
  for (int64_t& v : c->counters) v = 0;
  c->growth_pending = false;
  AV_CHECK_HIP(c, hipMemcpyAsync(R.drad.p, gl.rad.data(), gl.rad.size() * 4, hipMemcpyHostToDevice, c->stream));
  AV_CHECK_HIP(c, hipMemcpyAsync(R.dwin.p, wd.data(), sizeof(WinDev) * (size_t)nw, hipMemcpyHostToDevice, c->stream));
  // item offsets: VVEL tiles, then (after the reach is known) DDENS bin tiles,
  // then MARCAP cell blocks, then DDENS deposit tiles, each nw + 1 entries
  std::vector<int64_t> off(4 * ((size_t)nw + 1), 0);
  for (int32_t i = 0; i < nw; ++i) {
    off[i + 1] = off[i] + win_vvel_items(lw[i].w, k);
    off[2 * (nw + 1) + i + 1] = off[2 * (nw + 1) + i] + win_cell_items(lw[i].w);
    off[3 * (nw + 1) + i + 1] = off[3 * (nw + 1) + i] + win_dep_items(lw[i].w);
  }
  if (!R.doff.ensure(off.size() * 8)) AV_FAIL(c, AVISM_ERR_HIP, "out of device memory (refine)");
  AV_CHECK_HIP(c, hipMemcpyAsync(R.doff.p, off.data(), off.size() * 8, hipMemcpyHostToDevice, c->stream));
  const int64_t* d_off = R.doff.as<int64_t>();
  // the particles' bins and PPART of the loaded particles
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
  const double rho = (double)np / ((double)c->box_L0 * c->box_L0 * c->box_L0);
  const float r0 = (float)(1.3 * std::cbrt(3.0 * k / (4.0 * kPi * rho)));
  const VirtG V{R.drad.as<float>(), gl.rlo, gl.n, gl.periodic, gl.dx, gl.rx1, A};
  uint32_t* dmax = (uint32_t*)((char*)c->scalars.p + 512);
  unsigned long long* dpairs = (unsigned long long*)((char*)c->scalars.p + 1024);
  const bool several = P.atlas_side.size() > 1;
  if (several && P.hpart_k != k) {
    // HPART is the largest distance at which a particle serves a cell of any
    // window of the level, so with several atlases a first pass takes it over
    // every window (VVEL's neighbour search without the velocity sums)
    StageTimer t(c, "refine_hpart");
    const int32_t nall = (int32_t)P.win.size();
    std::vector<WinDev> all(nall);
    std::vector<int64_t> aoff((size_t)nall + 1, 0);
    for (int32_t i = 0; i < nall; ++i) {
      memset(&all[i], 0, sizeof(WinDev));
      all[i].ox = P.win[i].o[0]; all[i].oy = P.win[i].o[1]; all[i].oz = P.win[i].o[2]; all[i].w = P.win[i].w;
      aoff[i + 1] = aoff[i] + win_vvel_items(P.win[i].w, k);
    }
    DBuf dall, daoff;
    if (!P.hpart.ensure((size_t)np * 4) || !dall.ensure(sizeof(WinDev) * (size_t)nall) ||
        !daoff.ensure(aoff.size() * 8))
      AV_FAIL(c, AVISM_ERR_HIP, "out of device memory (refine)");
    AV_CHECK_HIP(c, hipMemcpyAsync(dall.p, all.data(), sizeof(WinDev) * (size_t)nall, hipMemcpyHostToDevice, c->stream));
    AV_CHECK_HIP(c, hipMemcpyAsync(daoff.p, aoff.data(), aoff.size() * 8, hipMemcpyHostToDevice, c->stream));
    AV_CHECK_HIP(c, launch_fill_u32(c->stream, np, P.hpart.as<uint32_t>(), 0u));
    const VirtG V0{R.drad.as<float>(), gl.rlo, gl.n, gl.periodic, gl.dx, gl.rx1, 0};
    AV_CHECK_HIP(c, launch_win_vvel(c->stream, B, np, dall.as<WinDev>(), daoff.as<int64_t>(), nall, aoff[nall], V0, k,
                                    r0, nullptr, P.hpart.as<uint32_t>(), nullptr, nullptr, nullptr));
    AV_CHECK_HIP(c, launch_fill_h(c->stream, np, P.hpart.as<float>(), c->hk.as<float>()));
    AV_CHECK_HIP(c, hipStreamSynchronize(c->stream));  // before dall and daoff are freed
    P.hpart_k = k;
  }
  {
    StageTimer t(c, "refine_vvel");
    AV_CHECK_HIP(c, launch_fill_u32(c->stream, ncell, c->fu.as<uint32_t>(), 0u));
    AV_CHECK_HIP(c, launch_fill_u32(c->stream, ncell, c->fv.as<uint32_t>(), 0u));
    AV_CHECK_HIP(c, launch_fill_u32(c->stream, ncell, c->fw.as<uint32_t>(), 0u));
    if (several)  // the level's HPART; this atlas's cells are among those it covers
      AV_CHECK_HIP(c, hipMemcpyAsync(c->hpart.p, P.hpart.p, (size_t)np * 4, hipMemcpyDeviceToDevice, c->stream));
    else
      AV_CHECK_HIP(c, launch_fill_u32(c->stream, np, c->hpart.as<uint32_t>(), 0u));
    AV_CHECK_HIP(c, launch_win_vvel(c->stream, B, np, R.dwin.as<WinDev>(), d_off, nw, off[nw], V, k, r0,
                                    c->vel4.as<float4>(), c->hpart.as<uint32_t>(), c->fu.as<float>(),
                                    c->fv.as<float>(), c->fw.as<float>()));
    if (!several) AV_CHECK_HIP(c, launch_fill_h(c->stream, np, c->hpart.as<float>(), c->hk.as<float>()));
    AV_CHECK_HIP(c, launch_fill_u32(c->stream, 1, dmax, 0u));
    AV_CHECK_HIP(c, launch_max_bits(c->stream, np, c->hpart.as<uint32_t>(), dmax));
  }
  c->counters[2] = wcells;
  uint32_t hmax_bits = 0;
  AV_CHECK_HIP(c, hipMemcpyAsync(&hmax_bits, dmax, 4, hipMemcpyDeviceToHost, c->stream));
  AV_CHECK_HIP(c, hipStreamSynchronize(c->stream));
  float hmax = 0;
  memcpy(&hmax, &hmax_bits, 4);
  // each window's particle range: the bins of the cells within the largest
  // DDENS reach (int(h/dx) + 2 cells) of the window
  // (+2 cells: the home cell, (int)((x - RADX(1)) / DX) + 1, and the bin of a
  // particle can each sit a cell away from the cell holding it)
  const int64_t reach = (int64_t)(hmax / gl.dx) + 4, nb = c->nb, nl = gl.n;
  for (int32_t i = 0; i < nw; ++i) {
    int32_t* b0 = &wd[i].bx0;
    int32_t* bn = &wd[i].bnx;
    for (int ax = 0; ax < 3; ++ax) {
      const int64_t o = lw[i].o[ax], w = lw[i].w;
      int64_t lo = floordiv((o - reach) * nb, nl), hi = -floordiv(-(o + w + reach) * nb, nl);
      if (gl.periodic) {
        if (hi - lo >= nb) {
          lo = 0;
          hi = nb;
        }
      } else {
        lo = std::max<int64_t>(lo, 0);
        hi = std::min<int64_t>(hi, nb);
      }
      b0[ax] = (int32_t)lo;
      bn[ax] = (int32_t)std::max<int64_t>(hi - lo, 0);
    }
    off[(nw + 1) + i + 1] = off[(nw + 1) + i] + win_ddens_items(wd[i]);
  }
  AV_CHECK_HIP(c, hipMemcpyAsync(R.dwin.p, wd.data(), sizeof(WinDev) * (size_t)nw, hipMemcpyHostToDevice, c->stream));
  AV_CHECK_HIP(c, hipMemcpyAsync(R.doff.p, off.data(), off.size() * 8, hipMemcpyHostToDevice, c->stream));
  // DDENS with the whole-grid fixed point (every cell sum < np * max mass * scale <= 2^62)
  double bound = (double)np * (double)(c->mass_max > 0.0f ? c->mass_max : 1.0f);
  int e2 = 0;
  std::frexp(bound, &e2);
  const double scale = std::ldexp(1.0, 62 - e2);
  {
    StageTimer t(c, "refine_ddens");
    AV_CHECK_HIP(c, launch_fill_u64(c->stream, ncell, c->mass_fix.as<unsigned long long>(), 0ull));
    AV_CHECK_HIP(c, launch_fill_u64(c->stream, 1, dpairs, 0ull));
    AV_CHECK_HIP(c, launch_win_ddens(c->stream, V, B, np, c->hpart.as<float>(), scale,
                                     c->mass_fix.as<unsigned long long>(), R.dwin.as<WinDev>(), d_off + (nw + 1),
                                     nw, off[2 * (nw + 1) - 1], dpairs, c->norm.as<float>(),
                                     c->pmark.as<uint8_t>(), c->sbin_val.as<int32_t>(),
                                     (int32_t*)((char*)c->scalars.p + 1800), c->sort_tmp, hmax,
                                     d_off + 3 * (nw + 1), off[4 * (nw + 1) - 1]));
  }
  // U1, MEANDENS = total mass / the level's cells, U1CO and DIVERCO on the atlas
  const float den = ((gl.dx * gl.dx) * gl.dx) * ((c->rete * c->rete) * c->rete);
  double* dsum = (double*)((char*)c->scalars.p + 256);
  float* dmean = (float*)((char*)c->scalars.p + 264);
  {
    StageTimer t(c, "refine_finalize");
    AV_CHECK_HIP(c, launch_ddens_finalize(c->stream, ncell, c->mass_fix.as<unsigned long long>(), 1.0 / scale, den,
                                          c->rote, c->u_mass.as<float>(), c->partial64.as<double>()));
    AV_CHECK_HIP(c, launch_mass_sum(c->stream, np, c->pm.as<float>(), R.dsum.as<double>(), 1024));
    k_refine_mean<<<1, 64, 0, c->stream>>>(R.dsum.as<double>(), 1024, (double)den,
                                           (double)gl.n * (double)gl.n * (double)gl.n, dsum, dmean);
    AV_CHECK_HIP(c, hipGetLastError());
    AV_CHECK_HIP(c, launch_div_finalize(c->stream, A, gl.dx, 0, c->fu.as<float>(), c->fv.as<float>(),
                                        c->fw.as<float>(), c->u_mass.as<float>(), dmean, c->u1co.as<float>(),
                                        c->div.as<float>()));
    AV_CHECK_HIP(c, launch_fill_u32(c->stream, ncell, c->marcap.as<uint32_t>(), 0u));
    AV_CHECK_HIP(c, launch_win_marcap(c->stream, R.dwin.as<WinDev>(), d_off + 2 * (nw + 1), nw,
                                      off[3 * (nw + 1) - 1], A, P.pmarca.as<int32_t>(), P.pside,
                                      c->marcap.as<int32_t>()));
  }
  if (!P.reqp.empty())
    AV_CHECK_HIP(c, hipMemcpyAsync(c->reqp.p, P.reqp.data(), P.reqp.size() * 4, hipMemcpyHostToDevice, c->stream));
  float hmean = 0;
  unsigned long long hpairs = 0;
  AV_CHECK_HIP(c, hipMemcpyAsync(&hmean, dmean, 4, hipMemcpyDeviceToHost, c->stream));
  AV_CHECK_HIP(c, hipMemcpyAsync(&c->meansum, dsum, 8, hipMemcpyDeviceToHost, c->stream));
  AV_CHECK_HIP(c, hipMemcpyAsync(&hpairs, dpairs, 8, hipMemcpyDeviceToHost, c->stream));
  AV_CHECK_HIP(c, hipStreamSynchronize(c->stream));
  // the context's grid becomes the atlas
  c->g = std::move(ga);
  if (!c->scalars.ensure(scalars_bytes(c))) AV_FAIL(c, AVISM_ERR_HIP, "out of device memory (scalars)");
  AV_CHECK_HIP(c, hipMemcpy((char*)c->scalars.p + 4096, c->g.rad.data(), c->g.rad.size() * 4, hipMemcpyHostToDevice));
  c->counters[0] = (int64_t)hpairs;
  c->mass_n = 0;  // the cell masses are the atlas's, not a whole grid's
  c->has_marcap = true;
  c->has_reqp = true;
  c->nreqp = (int32_t)P.reqp.size();
  c->has_mask = false;
  c->meandens_raw = hmean;
  const float ul3 = (kUL * kUL) * kUL;
  c->meandens = hmean * (kUM / ul3);  // avism.f90:1409
  c->fields_ready = true;
  // the loaded atlas's windows, for translations and the next plan
  R.loaded = atlas;
  R.grid_is_atlas = true;
  R.level_dx = gl.dx;
  R.level_rx1 = gl.rx1;
  R.meandens = c->meandens;
  R.window_cells = wcells;
  R.lwin = std::move(lw);
  int32_t maxp = 0;
  for (const WinHost& W : R.lwin) maxp = std::max(maxp, W.parent);
  R.win_of_parent.assign((size_t)maxp + 1, -1);
  for (int32_t i = 0; i < nw; ++i) R.win_of_parent[R.lwin[i].parent] = i;
  if (!R.dmap.ensure(R.win_of_parent.size() * 4)) AV_FAIL(c, AVISM_ERR_HIP, "out of device memory (refine)");
  AV_CHECK_HIP(c, hipMemcpy(R.dmap.p, R.win_of_parent.data(), R.win_of_parent.size() * 4, hipMemcpyHostToDevice));
  return AVISM_OK;
}

int refine_windows(avism_ctx* c, int32_t* tab, int32_t cap, int32_t* n, avism_refine_info* info) {
  RefineState* R = c->refine;
  const bool loaded = R && R->loaded >= 0;
  const int32_t nw = loaded ? (int32_t)R->lwin.size() : 0;
  if (n) *n = nw;
  if (tab && loaded) {
    if (cap < nw) AV_FAIL(c, AVISM_ERR_LIMIT, "window table too small: " + std::to_string(nw) + " entries");
    for (int32_t i = 0; i < nw; ++i) {
      const WinHost& W = R->lwin[i];
      const int32_t e[8] = {W.parent, W.w, W.o[0], W.o[1], W.o[2], W.a[0], W.a[1], W.a[2]};
      memcpy(tab + 8 * (size_t)i, e, sizeof e);
    }
  }
  if (info) {
    memset(info, 0, sizeof *info);
    info->atlas = -1;
    if (R && !R->plans.empty()) {
      info->depth = (int32_t)R->plans.size();
      info->n_level = R->plans.back().n_level;
      info->natlas = (int32_t)R->plans.back().atlas_side.size();
    }
    if (loaded) {
      info->atlas = R->loaded;
      info->side = c->g.n;
      info->nwindows = nw;
      info->window_cells = R->window_cells;
      info->dx = R->level_dx;
      info->rx1_level = R->level_rx1;
      info->dx_atlas = c->g.dx;
      info->rx1_atlas = c->g.rx1;
      info->meandens = R->meandens;
    }
  }
  return AVISM_OK;
}

int refine_pop(avism_ctx* c) {
  if (!c->refine || c->refine->plans.empty()) AV_FAIL(c, AVISM_ERR_STATE, "no refinement plan to drop");
  c->refine->plans.pop_back();
  c->refine->loaded = -1;
  c->refine->lwin.clear();
  c->refine->win_of_parent.clear();
  return AVISM_OK;
}

}  // namespace avism
