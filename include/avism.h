/*
 * avism.h -- C ABI of libavism_hip.so, the MI355X (gfx950) implementation of
 * AVISM's grid-and-void path (oscarmonllor99/AVISM).
 *
 * The reference has no plugin or FFI layer: the path is a chain of Fortran
 * subroutines called from PROGRAM AVISM.  Each entry point below replaces one
 * link of that chain; the comment names the reference code it stands for.
 * A Fortran host binds these with ISO_C_BINDING (host/avism_host.f90), a
 * Python host with ctypes (avism_amd/_lib.py).
 *
 * Conventions
 *  - All calls return 0 (AVISM_OK) or a negative AVISM_ERR_* code; the message
 *    is available from avism_last_error().  Where the reference STOPs, the
 *    library returns AVISM_ERR_INPUT / AVISM_ERR_LIMIT with the same message.
 *  - Calls are synchronous and single-host-thread per context.  Pointer
 *    arguments are host memory unless an argument named `where` says
 *    AVISM_DEVICE (then they are device pointers on the context's GPU; the
 *    avism_cnn_* training calls always take device pointers).  A call given
 *    device pointers first waits for all work queued on the device, so data
 *    produced on any stream is complete; its device outputs are complete
 *    when it returns.
 *  - 3-D arrays are column-major (x fastest).  "inner" arrays cover cells 1..N
 *    of each axis; "ext" arrays cover the periodic-extended range LOW1..LOW2
 *    (avism.f90:780-813), LOW1=1, LOW2=N without periodic boundaries.
 *  - Cube and void ids are 1-based, as in the reference.
 */
#ifndef AVISM_H
#define AVISM_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define AVISM_OK 0
#define AVISM_ERR_ARG -1
#define AVISM_ERR_HIP -2
#define AVISM_ERR_INPUT -3
#define AVISM_ERR_LIMIT -4
#define AVISM_ERR_STATE -5
#define AVISM_ERR_UNSUPPORTED -6

#define AVISM_HOST 0
#define AVISM_DEVICE 1
/* avism_get_fields only: host pointers into pinned memory (hipHostMalloc); the
 * copies are queued behind the work already queued and the call returns at
 * once, so they overlap the void stage.  The data is complete after
 * avism_sync; a call that changes the fields first waits for the copies. */
#define AVISM_HOST_ASYNC 2

/* growth_mode of avism_find_voids */
#define AVISM_GROW_PLAIN 0     /* exact shell-by-shell growth of every candidate */
#define AVISM_GROW_PROPOSAL 1  /* start from proposals given by avism_set_proposals */
#define AVISM_GROW_LEARNED 2   /* proposals from the network of avism_cnn_set_params */

/* proposal network: parameter count, classes, bf16 activations cached per patch */
#define AVISM_CNN_NPARAM 19064
#define AVISM_CNN_NCLASS 8
#define AVISM_CNN_CACHE 7920
#define AVISM_CNN_PATCH 2187 /* 3 x 9^3 bf16 features */

typedef struct avism_ctx avism_ctx;

typedef struct {
    int32_t n, periodic, nplus, low1, low2, next;
    float L0, lado0plus, dx, rx1;
} avism_grid_info;

/* voids.dat void-finder block (config/voids.dat:20-33, avism.f90:201-213) */
typedef struct {
    float dens_thre, dens_thre2, grad_thre, div_thre, rmin, rmin_sub;
    int32_t nvoid_max;
} avism_void_params;

typedef struct {
    int64_t ncand;      /* NXYZ: candidate centres (voidfinding.f90:106-109) */
    int32_t ncubes;     /* NVOID after VOIDFIND */
    int32_t nvoid2;     /* voids after MERGE_VOID */
    int32_t nvoidt;     /* voids in the catalogue (avism.f90:1973) */
    int32_t nperiodic;  /* main voids VOID_PERIODIC relabelled (CHANGE_BELONG, voidfinding.f90:1476-1480) */
    double volt_clean;  /* total void volume (avism.f90:1928) */
    float meandens;     /* MEANDENS in Msun/Mpc^3 (avism.f90:1409) */
    float ff;           /* VOLT_CLEAN / L0^3 (catalogue header, avism.f90:2047) */
} avism_level_stats;

/* ---- context ---------------------------------------------------------- */
int avism_create(avism_ctx** ctx, int32_t device);
int avism_destroy(avism_ctx* ctx);
const char* avism_last_error(avism_ctx* ctx);
int avism_set_timing(avism_ctx* ctx, int32_t on);
/* names (';'-separated) and milliseconds of the stages timed since the last
 * call.  *count holds the capacity of ms on entry and the number of stages on
 * return.  If ms or names is too small (or NULL) nothing is consumed and the
 * call returns AVISM_ERR_LIMIT with *count set: call again with larger
 * buffers (avism_last_error gives the names length needed). */
int avism_get_timings(avism_ctx* ctx, char* names, int32_t names_len, float* ms, int32_t* count);

/* ---- set-up ------------------------------------------------------------ */
/* MALLA (grid.f90:3-76) + the PBC extension (avism.f90:780-865) */
int avism_set_grid(avism_ctx* ctx, int32_t n, float L0, int32_t periodic);
int avism_get_grid(avism_ctx* ctx, avism_grid_info* out, float* rad_ext /* next or NULL */);
/* cosmology of avism.f90:469-479, 599-606 (h, Omega_m, redshift) */
int avism_set_cosmology(avism_ctx* ctx, float ache, float omega0, float zeta);
int avism_set_void_params(avism_ctx* ctx, const avism_void_params* p);
/* MARCAP of the previous level (avism.f90:2123-2161); NULL = -1 everywhere.
 * Cells with MARCAP = 0 can hold no cube; a cube grows only over cells of its
 * centre's parent id, and only cubes of one parent merge (voidfinding.f90:34,
 * 487-601, 848-850, 939). */
int avism_set_parent_map(avism_ctx* ctx, const int32_t* marcap_inner, int32_t where);
/* REQP of the previous level, reqp[id-1] for parent ids 1..n (avism.f90:2138-2144):
 * a candidate whose parent has REQP < RMIN_SUB grows no cube
 * (voidfinding.f90:267-274, 491).  NULL: no parent-size test.  Fails with
 * AVISM_ERR_INPUT at avism_find_voids if the parent map holds ids above n. */
int avism_set_parent_radii(avism_ctx* ctx, const float* reqp, int32_t n, int32_t where);
/* SMASK on the inner grid (int8, 1 = observed; avism.f90:403-418, 866-904);
 * NULL = 1 everywhere.  The periodic shell counts as masked, as in survey
 * mode.  VVEL_INTERP_SPH_VW and H_DISTANCE skip masked cells
 * (particles.f90:312, 856), MARK_ALL marks nothing there and the flag stencil
 * ends a cell at a masked neighbour (voidfinding.f90:33, 283-300); DDENS and
 * the mean density do not read it (the reference's non-survey mean). */
int avism_set_mask(avism_ctx* ctx, const int8_t* smask_inner, int32_t where);
/* the void-in-void hand-off (avism.f90:2123-2161, 742-759), after
 * avism_merge_and_post: MARCAP = this level's MARCA upsampled by 2 per axis,
 * REQP = REQ of its main voids, the survey mask (if set) upsampled likewise,
 * and the grid becomes 2N cells per axis over the same box.  Loaded particles are kept, so avism_particles_to_grid then
 * interpolates them at the new resolution (the reference frees them inside
 * its level loop, avism.f90:1613-1629); grid fields for 2N can be loaded
 * instead. */
int avism_next_level(avism_ctx* ctx);

/* ---- field stage (avism.f90:1177-1257, 1371-1414) --------------------- */
/* READ_BINARY_PART_1 data (reader.f90:410-465): positions Mpc/h, velocities
 * km/s, masses Msun.  Units are converted as the reader does (m/UM, v/UV).
 * Fails with AVISM_ERR_INPUT for particles outside [-L0/2, L0/2]
 * (avism.f90:560-589).  vx, vy, vz may all be NULL (READ_BINARY_PART_2 data,
 * reader.f90:468-519: no velocities; they are taken as 0). */
int avism_load_particles(avism_ctx* ctx, int64_t np, const float* x, const float* y,
                         const float* z, const float* vx, const float* vy, const float* vz,
                         const float* m, int32_t where);
/* PPART_DENS, VVEL_INTERP_SPH_VW / VEL_INTERP_TSC, FILL_H_ZEROS, H_DISTANCE,
 * DDENS_INTERP_SPH / DENS_INTERP_TSC, mean density, DIVER_UNIFORM
 * (avism.f90:1177-1257).  vel_interp, dens_interp: FLAG_VEL_INTERP and
 * FLAG_DENS_INTERP of voids.dat, 1 = SPH, 0 = TSC. */
int avism_particles_to_grid(avism_ctx* ctx, int32_t k, int32_t vel_interp, int32_t dens_interp);
/* FLAG_VEL_AVAILABLE = 0 (avism.f90:1259-1300, 1646-1650): the density only,
 * by DENS_INTERP_TSC (dens_interp 0) or H_DISTANCE, FILL_H_ZEROS and
 * DDENS_INTERP_SPH (dens_interp 1); then the mean density, and DIVERCO from
 * linear theory, -OMEGA0^0.6 RETE HTE (U1CO - 1).  The velocity fields read
 * back as 0. */
int avism_particles_to_density(avism_ctx* ctx, int32_t k, int32_t dens_interp);
/* Survey randoms (input_data/rands.dat, avism.f90:1431-1441): positions
 * Mpc/h and masses Msun (m/UM applied), kept beside the particles. */
int avism_load_randoms(avism_ctx* ctx, int64_t nrand, const float* x, const float* y, const float* z,
                       const float* m, int32_t where);
/* FLAG_SURVEY = 1 (avism.f90:1415-1597, 1646-1650): the galaxies' density
 * as avism_particles_to_density makes it; the randoms' cell masses U1RAND by
 * DENS_INTERP_TSC, or H_DISTANCE and FILL_H_ZEROS with MIN(8k, 256)
 * neighbours and DDENS; on observed cells U1CO = D/(alpha R) clipped to
 * [1e-4, 1000] with alpha = sum D / sum R, then divided by its mean; MEANDENS
 * = the background density; DIVERCO from linear theory.  Needs
 * avism_set_mask and avism_load_randoms. */
int avism_survey_density(avism_ctx* ctx, int32_t k, int32_t dens_interp);
/* READ_BINARY_GRID (reader.f90:523-552): delta and velocities (km/s) on the
 * inner grid; then DIVER_UNIFORM.  Used for grid input (FLAG_DATA=2). */
int avism_load_grid_fields(avism_ctx* ctx, const float* delta, const float* vx, const float* vy,
                           const float* vz, int32_t where);
/* U1CO (1+delta) and DIVERCO already on the inner grid, as the reference holds
 * them after avism.f90:1409 and 1633-1646 (the state MARK_ALL starts from);
 * meandens in Msun/Mpc^3 is used for the void masses. */
int avism_load_void_fields(avism_ctx* ctx, const float* u1co, const float* div, float meandens,
                           int32_t where);
/* inner fields: U1CO (1+delta), DIVERCO, and the velocity components; NULL
 * skips.  where: AVISM_HOST, AVISM_DEVICE or AVISM_HOST_ASYNC.  Stands for the
 * map output's source arrays (avism.f90:2028-2040). */
int avism_get_fields(avism_ctx* ctx, float* u1co, float* div, float* u, float* v, float* w,
                     int32_t where);
/* waits for every copy and kernel the context has queued (no counterpart in
 * the reference, whose calls are all synchronous) */
int avism_sync(avism_ctx* ctx);
/* work counters of the last stages run, for throughput reports (no
 * counterpart in the reference): out[0] (particle, cell) pairs within h in
 * DDENS_INTERP_SPH (its cells with d <= h, particles.f90:520-547), out[1] kNN
 * queries of PPART_DENS, out[2] kNN queries of VVEL_INTERP_SPH_VW or
 * H_DISTANCE (unmasked cells); of the last avism_find_voids: out[3] face
 * cells of the shells the cube growth examined, out[4] candidates grown by
 * one wave each; the rest 0 */
int avism_get_counters(avism_ctx* ctx, int64_t* out, int32_t n);
/* per-particle PART_DENS and HPART in the input order; NULL skips */
int avism_get_particle_fields(avism_ctx* ctx, float* part_dens, float* hpart);

/* ---- void stage --------------------------------------------------------- */
/* MARK_ALL + VOIDFIND (voidfinding.f90:13-713): marking, flag stencil,
 * candidate sort (INDEXX order), cube growth and acceptance. */
int avism_find_voids(avism_ctx* ctx, int32_t growth_mode, int32_t* ncubes);
/* learned half-width proposal per candidate (sorted candidate order, int8);
 * used by AVISM_GROW_PROPOSAL.  n must equal the candidate count. */
int avism_set_proposals(avism_ctx* ctx, const int8_t* r_hat, int64_t n, int32_t where);
/* candidates in processing order: ext linear index, and half-width reached
 * (-1 for centres skipped at the edge zone); NULL skips */
int avism_get_candidates(avism_ctx* ctx, int64_t* lin, int32_t* r, int64_t cap, int64_t* n);
/* cubes: ci[9*ncubes] = inicio x,y,z, final x,y,z, ic x,y,z (blocks of ncubes);
 * cf[7*ncubes] = rini x,y,z, rfin x,y,z, vol; father[ncubes] */
int avism_get_cubes(avism_ctx* ctx, int32_t* ci, float* cf, int32_t* father);
/* INDEXX (num.f90:2-84) of a float array, reproducing the Numerical Recipes
 * permutation including its tie order; indx is 1-based. */
int avism_indexx(avism_ctx* ctx, const float* keys, int64_t n, int32_t* indx, int32_t where);
/* avism_indexx on device pointers that allocates no device memory: no
 * counterpart in the reference; for callers that cannot allocate (code under
 * graph capture, tight loops, hosts that forbid it).  Its scratch is a static
 * array of about 20 MiB held on each GPU for the life of the process, so n is
 * at most 524288 (AVISM_ERR_LIMIT beyond); calls that overlap on one GPU race
 * on that array.  Same permutation as avism_indexx. */
int avism_indexx_noalloc(avism_ctx* ctx, const float* keys, int64_t n, int32_t* indx);

/* ---- merge and post-processing (avism.f90:1727-1973) ------------------- */
int avism_merge_and_post(avism_ctx* ctx, int32_t* nvoids);
/* indice[ncubes] (descending volume), uvoid[ncubes], ncellv[ncubes],
 * vf[12*ncubes] = gxc,gyc,gzc, xc,yc,zc, volnew, umean, mtot, eps, ip, req */
int avism_get_catalogue(avism_ctx* ctx, int32_t* indice, int32_t* uvoid, int32_t* ncellv,
                        float* vf, avism_level_stats* stats);
/* MARCA on the ext grid, or on the inner grid (the map file, avism.f90:2036) */
int avism_get_marca(avism_ctx* ctx, int32_t* marca, int32_t inner_only, int32_t where);

/* ---- half-width proposal network (no counterpart in the reference) ------
 * A 3-D CNN proposes each candidate's half-width; AVISM_GROW_LEARNED growth
 * re-verifies every proposal with the exact criteria of voidfinding.f90:500-654,
 * so the catalogue equals AVISM_GROW_PLAIN's.  The training calls take device
 * pointers: params[AVISM_CNN_NPARAM] fp32, patches[n][3][9][9][9] bf16 (as
 * uint16), logits[n][8] fp32, cache[n][AVISM_CNN_CACHE] bf16, labels[n]. */
/* network used by AVISM_GROW_LEARNED (copied into the context) */
int avism_cnn_set_params(avism_ctx* ctx, const float* params, int32_t where);
/* the network's proposals (int8 classes, processing order) for the current
 * candidates, as AVISM_GROW_LEARNED computes them; after avism_find_voids and
 * before avism_merge_and_post; n = the candidate count */
int avism_cnn_propose(avism_ctx* ctx, int8_t* out, int64_t n, int32_t where);
/* feature patches of the current candidates at processing ranks ranks[n] */
int avism_cnn_patches(avism_ctx* ctx, const int64_t* ranks, int64_t n, uint16_t* patches);
/* forward: logits, and the activations the backward needs (cache may be NULL) */
int avism_cnn_forward(avism_ctx* ctx, const float* params, int64_t n, const uint16_t* patches,
                      float* logits, uint16_t* cache);
/* backward of the mean softmax cross-entropy: grads[AVISM_CNN_NPARAM], loss[1] */
int avism_cnn_backward(avism_ctx* ctx, const float* params, int64_t n, const uint16_t* patches,
                       const uint16_t* cache, const float* logits, const int32_t* labels,
                       float* grads, float* loss);
/* one Adam step (beta 0.9, 0.999, eps 1e-8) on params with moments m, v; step >= 1 */
int avism_cnn_adam(avism_ctx* ctx, float* params, const float* grads, float* m, float* v, float lr,
                   int32_t step);

/* ---- batches of independent patches ----------------------------------------
 * No single counterpart in the reference: each patch is one run of it
 * (avism.f90:1177-1973) on its own grid, particles and survey mask.  The
 * field stage of every patch (PPART_DENS, VVEL_INTERP_SPH_VW, FILL_H_ZEROS,
 * DDENS_INTERP_SPH, mean density, DIVER_UNIFORM; SPH only) runs in persistent
 * launches over all patches, with tile shapes picked by an autotuner on the
 * layout's first batch and the launch sequence captured in a graph that later
 * batches of the same layout replay.  The void stage (avism_find_voids +
 * avism_merge_and_post) of each patch runs on its own context. */
/* layout: npatch patches of n[p]^3 cells over a box of L0[p] (periodic[p], or
 * all non-periodic when NULL) holding np[p] particles */
int avism_batch_set(avism_ctx* ctx, int32_t npatch, const int32_t* n, const float* L0, const int32_t* periodic,
                    const int64_t* np);
/* particles of every patch back to back (patch p at offset np[0]+..+np[p-1]),
 * as avism_load_particles takes them; smask: the patches' inner SMASK grids back
 * to back (n[p]^3 bytes each) or NULL */
int avism_batch_load(avism_ctx* ctx, const float* x, const float* y, const float* z, const float* vx,
                     const float* vy, const float* vz, const float* m, const int8_t* smask, int32_t where);
/* the field stage of every patch, k <= 64 (AVISM_ERR_UNSUPPORTED above) */
int avism_batch_fields(avism_ctx* ctx, int32_t k);
/* the void stage of every patch (AVISM_GROW_PLAIN or AVISM_GROW_LEARNED with the
 * network of avism_cnn_set_params), the patches spread over nthreads host
 * threads (0: 8), with the context's cosmology and void parameters */
int avism_batch_voids(avism_ctx* ctx, int32_t growth_mode, int32_t nthreads);
/* patch p's context after avism_batch_voids: its catalogue, MARCA, cubes and
 * fields through the usual getters; owned by ctx (do not destroy) */
int avism_batch_patch(avism_ctx* ctx, int32_t p, avism_ctx** patch_ctx);
/* patch p's inner fields and, in its input order, PART_DENS and HPART; NULL skips */
int avism_batch_get_fields(avism_ctx* ctx, int32_t p, float* u1co, float* div, float* u, float* v, float* w,
                           float* part_dens, float* hpart, int32_t where);
/* the autotuner's record (JSON: candidates' milliseconds and the choices),
 * the number of graph launches so far, MEANDENS per patch; NULL skips */
int avism_batch_info(avism_ctx* ctx, char* tuning, int32_t len, int64_t* graph_launches, float* meandens);

/* ---- refined levels inside the parent voids --------------------------------
 * Void-in-void levels (avism.f90:742-759, 957-968, 2123-2161) computed only
 * around the parent voids, for boxes whose finer grids do not fit whole.
 * The reference runs every level on the whole box and cannot run more than
 * one level with particle input at all (its level loop frees the particles,
 * avism.f90:1613-1629); these semantics are this library's:
 *  - a level has 2x the parent level's cells per axis.  Each main void V of
 *    the parent level with REQ(V) >= RMIN_SUB (smaller ones can hold no cube,
 *    voidfinding.f90:267-274) gets one cubic window: the bounding box of V's
 *    cells at the finer resolution, grown by `margin` cells per side;
 *  - windows are packed side by side into atlases (non-periodic grids of at
 *    most max_cells cells) that are processed one at a time, so memory is
 *    bounded by the largest atlas, not by the level's whole grid;
 *  - in a window, MARCAP is V on V's cells and 0 elsewhere, REQP(V) =
 *    REQ(V): the void stage finds V's subvoids independently of other parents;
 *  - fields: VVEL_INTERP_SPH_VW queries the window's cells over all loaded
 *    particles (the box's periodic wrap); HPART is the largest distance at
 *    which a particle served any window cell of the level; DDENS_INTERP_SPH
 *    normalises each particle over its whole cell box on the level's grid
 *    (particles.f90:520-547) and keeps the deposits that fall in a window;
 *    MEANDENS is the total particle mass over the level's cells, which the
 *    whole-grid mean equals because DDENS conserves mass (particles.f90:549-604).
 * Call after avism_merge_and_post of the parent level (the whole box, or a
 * loaded atlas of the previous plan).  avism_particles_to_grid must have run
 * with SPH velocities on the loaded particles (its PPART_DENS is reused). */
typedef struct {
    int32_t depth;        /* plans on the stack (1: the first refined level) */
    int32_t n_level;      /* cells per axis of the level's whole grid */
    int32_t natlas;       /* atlases of the top plan */
    int32_t atlas;        /* atlas loaded on the context, -1 if none */
    int32_t side;         /* cells per axis of that atlas */
    int32_t nwindows;     /* its windows */
    int64_t window_cells; /* cells of its windows */
    float dx;             /* cell size of the level's grid */
    float rx1_level;      /* RADX(1) of the level's grid */
    float dx_atlas;       /* cell size of the atlas grid (MALLA over side cells of dx; within an ulp of dx) */
    float rx1_atlas;      /* RADX(1) of the atlas grid */
    float meandens;       /* MEANDENS of the level, Msun/Mpc^3 */
} avism_refine_info;
/* plan the next level's windows from the current level's catalogue and MARCA
 * (pushed on a stack of plans); margin >= 4 cells; max_cells <= 2^31-1 bounds
 * each atlas (0: 2^30); *natlas = the atlases to process (0: no parent
 * qualifies) */
int avism_refine_plan(avism_ctx* ctx, int32_t margin, int64_t max_cells, int32_t* natlas);
/* atlas a of the top plan becomes the context's grid (non-periodic, side^3
 * cells) with its fields, MARCAP and REQP; then avism_find_voids,
 * avism_merge_and_post and the getters work on it as on any grid (cube and
 * catalogue coordinates in the atlas frame: avism_refine_windows maps them
 * to the level's grid).  k: KNEIGHBOURS (SPH only). */
int avism_refine_load(avism_ctx* ctx, int32_t atlas, int32_t k);
/* the loaded atlas: tab[8*i .. 8*i+7] = parent id, side, origin x, y, z in
 * the level's cells (0-based; in a periodic box it may lie outside 0..n_level-1
 * and wraps), origin x, y, z in the atlas (0-based), for i < *n (cap: entries
 * tab holds; tab may be NULL); info may be NULL.  A position X in the atlas
 * frame is X - rx1_atlas - ax dx_atlas + rx1_level + ox dx in the level's. */
int avism_refine_windows(avism_ctx* ctx, int32_t* tab, int32_t cap, int32_t* n, avism_refine_info* info);
/* drop the top plan (its parent MARCA copy); the context's grid keeps the
 * last loaded atlas until another one is loaded */
int avism_refine_pop(avism_ctx* ctx);

/* ---- MASCLET AMR input (FLAG_DATA = 0) ------------------------------------
 * READ_MASCLET (reader.f90:86-405) reads a base grid of NHYX^3 cells and AMR
 * patches; the gas enters the void finder through a uniform grid per level
 * and the dark-matter particles through the SPH path.  avism_amr_set takes
 * the gas as the files hold it: base[4][nhyx^3] = density contrast, vx, vy,
 * vz (x fastest); npatch level-1 patches with pint[6*npatch] = PATCHNX,
 * PATCHNY, PATCHNZ, PATCHX, PATCHY, PATCHZ (1-based base cell of the first
 * cells), preal[3*npatch] = PATCHRX, PATCHRY, PATCHRZ, and pdata = per patch
 * density contrast, vx, vy, vz (nx*ny*nz each, x fastest), patches back to
 * back.  Fails with AVISM_ERR_INPUT for a patch with no cells. */
int avism_amr_set(avism_ctx* ctx, int32_t nhyx, const float* base, int32_t npatch, const int32_t* pint,
                  const float* preal, const float* pdata, int32_t where);
/* U1CO and DIVERCO of level ir on the context's grid (NHYX * 2^ir cells per
 * axis, avism.f90:750): the gas by SMOOTH_2 (ir < 0, amr2uniform.f90:62-121)
 * or the base grid (ir = 0) with DIVER_UNIFORM, or DIVER_FINA_GAS and VMESH
 * (ir = 1, divergence.f90:55-337, amr2uniform.f90:130-248; VMESH averages a
 * refined cell over 1 + its covering patches, as the reference's FLAGAMR
 * counts from 1); then U1CO by FLAG_DENS (0: dark matter + gas, 1: dark
 * matter, 2: gas) times ROTE over its FP64 mean, and DIVERCO by FLAG_DIV (0:
 * the smaller of gas and dark matter, 1: dark matter, 2: gas, -1: linear
 * theory), avism.f90:1369-1409, 1633-1649.  The dark matter is that of the
 * last avism_particles_to_grid on this grid (required unless FLAG_DENS = 2
 * and FLAG_DIV is 2 or -1); its divergence is 0 at ir = 1, as the reference
 * computes none there.  ir > 1: AVISM_ERR_UNSUPPORTED (the reference's VMESH
 * indexes its base-grid flags past level 1); AVISM_ERR_INPUT for a patch
 * outside the level's grid or whose parent cells touch the base grid's faces
 * (the reference would read outside its arrays). */
int avism_amr_fields(avism_ctx* ctx, int32_t ir, int32_t flag_dens, int32_t flag_div);

/* ---- whole level ---------------------------------------------------------- */
/* particles already loaded: fields, voids, merge and post in one call */
int avism_run_level(avism_ctx* ctx, int32_t k, int32_t vel_interp, int32_t dens_interp,
                    int32_t growth_mode, avism_level_stats* stats);

#ifdef __cplusplus
}
#endif
#endif
