// avism_internal.h -- shared types and helpers of the HIP implementation.
//
// Layout conventions (same as the reference, column-major, x fastest):
//   inner grid  : cell (i,j,k), 1..N      -> (i-1) + N*((j-1) + N*(k-1))
//   ext grid    : cell (I,J,K), LOW1..LOW2 -> (I-LOW1) + NE*((J-LOW1) + NE*(K-LOW1))
// Cube and void ids are 1-based, as in the reference.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/avism.h"

namespace avism {

constexpr float kUM = 9.1717E18f;      // commondata.f90:17
constexpr float kUV = 299792.458f;     // commondata.f90:16
constexpr float kUL = 10.98f;          // commondata.f90:15
constexpr int kNcellMin = 10;          // commondata.f90:10
constexpr int kNsec = 1;               // voidfinding.f90:101
constexpr int kNbuff = 1;              // voidfinding.f90:797
constexpr double kPi = 3.141592653589793238462643383279502884;

// flag byte of an extended cell (voidfinding.f90:256-432 and MARK_ALL 13-55)
enum : uint8_t {
  FL_X1 = 1, FL_X2 = 2, FL_Y1 = 4, FL_Y2 = 8, FL_Z1 = 16, FL_Z2 = 32,
  FL_SUB = 64,   // FLAG_SUB: may belong to a (sub)void
  FL_V = 128,    // FLAGV: candidate centre
};

struct GridGeom {
  int32_t n, periodic, nplus, low1, low2, next;
  float L0, lado0plus, dx, rx1;
  // RADX over reference indices rlo..rhi
  int32_t rlo, rhi;
  std::vector<float> rad;
  float radx(int i) const { return rad[i - rlo]; }
};

struct BatchState;   // a batch of patches (batch.hip)
struct RefineState;  // windows of refined void-in-void levels (refine.hip)
struct AmrState;     // MASCLET AMR gas (amr.hip)

GridGeom make_grid(int32_t n, float L0, int32_t periodic);

// Device buffer that only grows.
struct DBuf {
  void* p = nullptr;
  size_t bytes = 0;
  DBuf() = default;
  DBuf(const DBuf&) = delete;  // one owner per allocation
  DBuf& operator=(const DBuf&) = delete;
  DBuf(DBuf&& o) noexcept : p(o.p), bytes(o.bytes) {
    o.p = nullptr;
    o.bytes = 0;
  }
  DBuf& operator=(DBuf&& o) noexcept {
    if (this != &o) {
      release();
      p = o.p;
      bytes = o.bytes;
      o.p = nullptr;
      o.bytes = 0;
    }
    return *this;
  }
  ~DBuf();
  void* ensure(size_t nbytes);
  template <class T> T* as() const { return reinterpret_cast<T*>(p); }
  void release();
};

// a timed stage: two events on the context stream, read when the timings are
// fetched (so timing adds no synchronisation to the timed work)
struct Timing {
  std::string name;
  hipEvent_t a, b;
};

}  // namespace avism

// Context: owns all device memory and the stream.
struct avism_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  // host<->device copies that overlap the kernels run on cstream (default
  // priority, like stream): the particle velocities are uploaded there while
  // binning and PPART_DENS run (ev_vel orders their first reader), and
  // avism_get_fields(AVISM_HOST_ASYNC) downloads the maps there while the void
  // stage runs (copy_pending: wait for cstream before the fields change)
  hipStream_t cstream = nullptr;
  hipEvent_t ev_vel = nullptr, ev_mark = nullptr;
  bool copy_pending = false;
  int64_t counters[8] = {0};                   // work counters (avism_get_counters)
  std::string err;
  avism::GridGeom g{};
  bool grid_set = false;
  float ache = 0.678f, omega0 = 0.31f, zeta = 0.0f;
  float rote = 0, rete = 0, hte = 0;
  avism_void_params vp{};
  bool timing_on = false;
  std::vector<avism::Timing> timings;

  // particles (original order) and the binned copies
  int64_t np = 0;
  // the particles' box (the grid's box, except while an atlas of windows is
  // loaded, whose grid is not the box): bins and kNN wrap over it
  float box_L0 = 0;
  int32_t box_periodic = 0;
  bool ppart_valid = false;                    // pdens, hk and vel4 hold the loaded particles' PPART
  avism::DBuf px, py, pz, pvx, pvy, pvz, pm;   // converted units after upload
  avism::DBuf sbin_key, sbin_val, sbin_key2, sbin_val2, sort_tmp;
  avism::DBuf spos4, sorig;                    // sorted by bin
  avism::DBuf vel4;                            // (vx, vy, vz, m/rho), sorted by bin
  avism::DBuf bin_start;                       // nb^3 + 1
  int32_t nb = 0;
  float bsz = 0;
  avism::DBuf pdens, hk, hpart, norm;          // sorted order
  avism::DBuf pmark;                           // np bytes: particles marked for a NORM (fine-grid DDENS)
  avism::DBuf scalars;                         // small device scalars
  // inner fields
  avism::DBuf u_mass;                          // mass -> density (before /MEANDENS)
  avism::DBuf fu, fv, fw, u1co, div;
  avism::DBuf partial64;
  avism::DBuf marcap;                          // inner MARCAP (optional)
  bool has_marcap = false;
  avism::DBuf smask;                           // inner SMASK (int8), when has_mask
  bool has_mask = false;
  int64_t mask_observed = 0;                   // cells with SMASK != 0
  avism::DBuf reqp;                            // REQP of the parent level, ids 1..nreqp
  int32_t nreqp = 0;
  bool has_reqp = false;
  float mass_max = 0;                          // max particle mass (code units)
  bool mass_uniform = false;                   // every particle has the same mass
  avism::DBuf tsc_acc;                         // TSC velocity sums, FP64 [4][N^3]
  avism::DBuf mass_fix;                        // DDENS cell sums, int64 fixed point
  float meandens = 0;                          // MEANDENS in Msun/Mpc^3
  float meandens_raw = 0;                      // MEANDENS before unit scaling
  double meansum = 0;
  bool fields_ready = false;
  // survey randoms (FLAG_SURVEY = 1): a second particle set, swapped with the
  // first by swap_particle_sets; raw_mass: DDENS leaves cell masses unscaled
  int64_t nrand = 0;
  avism::DBuf rpx, rpy, rpz, rpvx, rpvy, rpvz, rpm, rspos4, rsorig, rvel4, rbin_start;
  avism::DBuf rpdens, rhk, rhpart, rnorm, u_rand;
  int32_t rnb = 0;
  float rbsz = 0, rmass_max = 0;
  bool rmass_uniform = false, raw_mass = false;

  // void stage
  avism::DBuf flags;                           // ext flag bytes
  avism::DBuf cand_key, cand_lin, cand_cnt, blk_cnt;
  avism::DBuf sort_idx, sort_a, sort_b, seg_a, seg_b, seg_small, sort_misc, sort_gpos, sort_spos;
  int64_t ncand = 0;
  avism::DBuf cand_r, cand_state, cand_flag2;  // growth results per sorted candidate
  avism::DBuf cand_work;                       // growth work per sorted candidate (counters)
  bool growth_pending = false;                 // cand_work not yet summed into counters
  avism::DBuf cover_acc, cover_und;            // ext grids for acceptance
  avism::DBuf acc_rank, misc_i32;
  avism::DBuf proposals;                       // half-width proposals (int8 per candidate)
  avism::DBuf cnn_params;                      // proposal network parameters (fp32)
  avism::DBuf cnn_work;                        // DIVERCO^2 partial sums and their total (FP64)
  bool cnn_ready = false;
  int32_t nvoid = 0;                           // accepted cubes
  avism::DBuf cube_i, cube_f, cube_father;     // [9][nvoid], [7][nvoid], [nvoid]
  // merge and post
  avism::DBuf indice, rank_of, paintP, paintK, jstar, pi_ptr, group, uvoid, marca;
  avism::DBuf ev_key, ev_val, ev_key2, ev_val2, grp_start, vf, ncellv, removed;
  avism::DBuf ccl, cellid, cellid2, cellkey, cellkey2, vstart, inert, per_tmp, bkt;
  int32_t nvoid2 = 0, nvoidt = 0, nperiodic = 0;
  bool post_ready = false;                     // MARCA and the catalogue are final
  double volt_clean = 0;
  int64_t growth_stats[8] = {0};
  avism::BatchState* batch = nullptr;          // avism_batch_* state
  avism::RefineState* refine = nullptr;        // avism_refine_* state
  avism::AmrState* amr = nullptr;              // avism_amr_* state
  // the DDENS / TSC cell masses in mass_fix: int64 fixed point with this
  // inverse scale, or FP64 (mass_fp64); mass_n: the grid they were made on
  double mass_inv_scale = 0;
  int32_t mass_fp64 = 0, mass_n = 0;
};

#define AV_CHECK_HIP(ctx, expr)                                                        \
  do {                                                                                 \
    hipError_t _e = (expr);                                                            \
    if (_e != hipSuccess) {                                                            \
      (ctx)->err = std::string(#expr) + ": " + hipGetErrorString(_e);                 \
      return AVISM_ERR_HIP;                                                            \
    }                                                                                  \
  } while (0)

#define AV_FAIL(ctx, code, msg) \
  do {                          \
    (ctx)->err = (msg);         \
    return (code);              \
  } while (0)

namespace avism {
// stage entry points (each returns an AVISM_* status)
int fields_from_particles(avism_ctx* c, int32_t k, int32_t vel_interp, int32_t dens_interp);
int fields_from_grid(avism_ctx* c, const float* delta, const float* vx, const float* vy,
                     const float* vz, int where);
int load_particles(avism_ctx* c, int64_t np, const float* x, const float* y, const float* z,
                   const float* vx, const float* vy, const float* vz, const float* m, int where);
void swap_particle_sets(avism_ctx* c);
int survey_density(avism_ctx* c, int32_t k, int32_t dens_interp);
int void_find(avism_ctx* c, int32_t growth_mode);
int merge_and_post(avism_ctx* c, int32_t stop_after);
int next_level(avism_ctx* c);
int indexx_device(avism_ctx* c, const float* d_keys, int64_t n, int32_t* d_indx);
int noalloc_init(int device);
// batches of independent patches (batch.hip)
void batch_free(avism_ctx* c);
int batch_set(avism_ctx* c, int32_t npatch, const int32_t* n, const float* L0, const int32_t* periodic,
              const int64_t* np);
int batch_load(avism_ctx* c, const float* x, const float* y, const float* z, const float* vx, const float* vy,
               const float* vz, const float* m, const int8_t* smask, int where);
int batch_fields(avism_ctx* c, int32_t k);
int batch_voids(avism_ctx* c, int32_t growth_mode, int32_t nthreads);
int batch_patch(avism_ctx* c, int32_t p, avism_ctx** out);
int batch_get_fields(avism_ctx* c, int32_t p, float* u1co, float* div, float* u, float* v, float* w,
                     float* part_dens, float* hpart, int where);
int batch_info(avism_ctx* c, char* tuning, int32_t len, int64_t* graph_launches, float* meandens);
int indexx_noalloc(avism_ctx* c, const float* d_keys, int64_t n, int32_t* d_indx);
// refined levels restricted to windows around the parent voids (refine.hip)
void refine_free(avism_ctx* c);
void amr_free(avism_ctx* c);
int amr_set(avism_ctx* c, int32_t nb, const float* base, int32_t npatch, const int32_t* pint, const float* preal,
            const float* pdata, int where);
int amr_fields(avism_ctx* c, int32_t ir, int32_t flag_dens, int32_t flag_div);
bool refine_grid_is_atlas(const avism_ctx* c);  // the grid is an atlas of windows
int refine_plan(avism_ctx* c, int32_t margin, int64_t max_cells, int32_t* natlas);
int refine_load(avism_ctx* c, int32_t atlas, int32_t k);
int refine_windows(avism_ctx* c, int32_t* tab, int32_t cap, int32_t* n, avism_refine_info* info);
int refine_pop(avism_ctx* c);
void add_timing(avism_ctx* c, const char* name, hipEvent_t a, hipEvent_t b);
int cnn_patches(avism_ctx* c, const int64_t* ranks, int64_t n, uint16_t* out);
int cnn_forward(avism_ctx* c, const float* params, int64_t n, const uint16_t* patches, float* logits,
                uint16_t* cache);
int cnn_backward(avism_ctx* c, const float* params, int64_t n, const uint16_t* patches, const uint16_t* cache,
                 const float* logits, const int32_t* labels, float* grads, float* loss);
int cnn_adam(avism_ctx* c, float* params, const float* grads, float* m, float* v, float lr, int32_t step);
int cnn_propose(avism_ctx* c);

// the small device scalars: counters and flags in the first 4 KB, RADX at
// 4096, then 8-byte slots for kRedSlices partial sums
inline size_t scalars_slices_offset(const GridGeom& g) { return (4096 + g.rad.size() * 4 + 127) & ~(size_t)63; }
inline size_t scalars_bytes(const GridGeom& g) { return scalars_slices_offset(g) + 512 * 8; }
inline size_t scalars_slices_offset(const avism_ctx* c) { return scalars_slices_offset(c->g); }
inline size_t scalars_bytes(const avism_ctx* c) { return scalars_bytes(c->g); }

// wait (on the host) for async map downloads before their source changes
int finish_copies(avism_ctx* c);

// scoped stage timer on the context stream (or another stream of the context)
struct StageTimer {
  avism_ctx* c;
  const char* name;
  hipStream_t s;
  hipEvent_t a = nullptr, b = nullptr;
  StageTimer(avism_ctx* c_, const char* n_, hipStream_t s_ = nullptr) : c(c_), name(n_), s(s_ ? s_ : c_->stream) {
    if (c->timing_on) {
      (void)hipEventCreate(&a);
      (void)hipEventCreate(&b);
      (void)hipEventRecord(a, s);
    }
  }
  ~StageTimer() {
    if (c->timing_on) {
      (void)hipEventRecord(b, s);
      add_timing(c, name, a, b);
    }
  }
};
}  // namespace avism
