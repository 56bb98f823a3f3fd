/*
 * avism_oracle.h -- CPU oracle for the AVISM grid-and-void path.
 *
 * TEST INFRASTRUCTURE ONLY. This library restates, in plain C, the
 * reference algorithm (oscarmonllor99/AVISM, Fortran) so the HIP product path
 * can be checked against it.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it.  The product never links it.
 *
 * Conventions (same as the reference):
 *   - 3-D arrays are column-major, x fastest: idx = i + nx*(j + ny*k).
 *   - Extended (ghost-shell) grids cover reference indices LOW1..LOW2 on each
 *     axis; element (I,J,K) lives at (I-LOW1) + NE*((J-LOW1) + NE*(K-LOW1)).
 *   - Cube/void ids are 1-based, as in the reference.
 *   - Floating point follows the reference's kinds: REAL*4 arithmetic is done
 *     in float in the reference's operation order (no FMA contraction: the
 *     file is compiled with -ffp-contract=off), REAL*8 accumulations in double.
 */
#ifndef AVISM_ORACLE_H
#define AVISM_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- grid geometry: MALLA (grid.f90:3-76) + PBC extension (avism.f90:780-865) */
typedef struct {
    int32_t n, periodic, nplus, low1, low2, next;
    float L0, lado0plus, dx, rx1;
    /* RADX over reference indices rlo..rhi (covers 0..n+1 and low1..low2) */
    int32_t rlo, rhi;
    float rad[1];  /* flexible; allocate with orc_grid_new */
} orc_grid;

orc_grid* orc_grid_new(int32_t n, float L0, int32_t periodic);
void orc_grid_free(orc_grid* g);
float orc_grid_rad(const orc_grid* g, int32_t i);
/* Copies header values and RADX(low1..low2) into caller arrays. */
void orc_grid_info(const orc_grid* g, int32_t* ints6, float* floats4, float* rad_ext);

/* ---- exact kNN (kdtree.f90:504-540 semantics, tie-break (d2, index)) ---- */
typedef struct orc_knn orc_knn;
orc_knn* orc_knn_new(int64_t np, const float* x, const float* y, const float* z,
                     float L0, int32_t periodic);
void orc_knn_free(orc_knn* t);
/* k nearest of (qx,qy,qz): indices (0-based) and squared distances, ascending. */
int32_t orc_knn_query(orc_knn* t, float qx, float qy, float qz, int32_t k,
                      int64_t* idx, float* d2);
/* all points with d2 <= r*r; returns count (<= cap) ascending by (d2, index) */
int64_t orc_ball_query(orc_knn* t, float qx, float qy, float qz, float r,
                       int64_t cap, int64_t* idx, float* d2);

/* ---- field stage ---- */
/* READ_BINARY_PART_1 unit conversion (reader.f90:457-461) */
void orc_convert_units(int64_t np, float* vx, float* vy, float* vz, float* m);
/* PPART_DENS (particles.f90:742-802); hk_out = distance to k-th neighbour */
void orc_ppart_dens(orc_knn* t, int64_t np, const float* x, const float* y,
                    const float* z, const float* m, int32_t k,
                    float* part_dens, float* hk_out);
/* VVEL_INTERP_SPH_VW (particles.f90:806-930) with KERNEL_FUNC (935-963).
 * smask: inner N^3 int8 or NULL (all ones).  hpart updated in place. */
void orc_vvel_vw(orc_knn* t, const orc_grid* g, int32_t k, const float* m,
                 const float* part_dens, const float* vx, const float* vy,
                 const float* vz, const int8_t* smask, float* hpart,
                 float* u, float* v, float* w);
/* H_DISTANCE (particles.f90:276-351) */
void orc_vvel_cells(orc_knn* t, const orc_grid* g, int32_t k, const float* m,
                    const float* part_dens, const float* vx, const float* vy,
                    const float* vz, int64_t ncells, const int64_t* cells,
                    float* u, float* v, float* w);
void orc_h_distance(orc_knn* t, const orc_grid* g, int32_t k, const int8_t* smask,
                    float* hpart);
/* FILL_H_ZEROS (particles.f90:398-431) */
void orc_fill_h_zeros(orc_knn* t, int64_t np, const float* x, const float* y,
                      const float* z, int32_t k, float* hpart);
/* DDENS_INTERP_SPH (particles.f90:434-618): mass per inner cell, FP64 sums */
void orc_ddens_sph(const orc_grid* g, int64_t np, const float* x, const float* y,
                   const float* z, const float* hpart, const float* m,
                   float* mass_out, double* mass64_out);
void orc_ddens_sph_par(const orc_grid* g, int64_t np, const float* x, const float* y,
                       const float* z, const float* hpart, const float* m,
                       float* mass_out, double* mass64_out);
/* DENS_INTERP_TSC / VEL_INTERP_TSC (particles.f90:16-128, 132-272) */
void orc_dens_tsc(const orc_grid* g, int64_t np, const float* x, const float* y,
                  const float* z, const float* m, float* mass_out);
void orc_vel_tsc(const orc_grid* g, int64_t np, const float* x, const float* y,
                 const float* z, const float* vx, const float* vy, const float* vz,
                 int32_t exact, float* u, float* v, float* w);
/* avism.f90:1249-1250, 1371-1414: mass -> U1CO (=1+delta); returns MEANDENS
 * (Msun/Mpc^3) and the raw FP64 sum through sum_out. */
float orc_normalise(const orc_grid* g, float ache, float omega0, float zeta,
                    const float* mass, float* u1co, double* sum_out);
/* DIVER_UNIFORM (divergence.f90:2-50) */
void orc_diver_uniform(int32_t n, float dx, const float* u, const float* v,
                       const float* w, int32_t periodic, float* div);

/* ---- void stage ---- */
typedef struct {
    float dens_thre, dens_thre2, grad_thre, div_thre, rmin, rmin_sub;
    int32_t nvoid_max;
} orc_void_params;

/* REAL_MAKE_PERIODIC / INT_MAKE_PERIODIC (num.f90:89-181): inner -> ext */
void orc_extend_f32(const orc_grid* g, const float* inner, float* ext);
void orc_extend_i32(const orc_grid* g, const int32_t* inner, int32_t* ext);
/* INDEXX (num.f90:2-84), Numerical Recipes indexx; indx is 1-based on output */
int32_t orc_indexx(int64_t n, const float* arr, int32_t* indx);

/* MARK_ALL + VOIDFIND (voidfinding.f90:13-55, 59-713).
 * u1co_ext, div_ext, marcap2_ext, smask_ext: extended grids.
 * reqp: parent radii (1-based ids -> reqp[id-1]) or NULL.
 * cubes: caller arrays of nvoid_max entries:
 *   ci[9*nvoid_max]: inicio x,y,z / final x,y,z / ic x,y,z  (blocks of nvoid_max)
 *   cf[7*nvoid_max]: rini x,y,z / rfin x,y,z / vol
 *   father[nvoid_max]
 * stats[4]: ncand (NXYZ), nflagv, nflag_sub, (unused)
 * cand_order (optional, NXYZ ints): candidate linear ext index in processing order
 * cand_r (optional, NXYZ ints): half-width reached per candidate (-1 skipped)
 * Returns NVOID (>=0) or a negative error code. */
int32_t orc_voidfind(const orc_grid* g, const orc_void_params* p,
                     const float* u1co_ext, const float* div_ext,
                     const int32_t* marcap2_ext, const int8_t* smask_ext,
                     const float* reqp, int32_t* ci, float* cf, int32_t* father,
                     int64_t* stats, int64_t ncand_cap, int64_t* cand_lin,
                     int32_t* cand_r);

/* Volume sort + MERGE_VOID + post-processing (avism.f90:1727-1973,
 * voidfinding.f90:718-1161, 1367-1568, shape.f90:140-225).
 * Inputs: the nvoid cubes from orc_voidfind (same array layout).
 * u1co_ext is read for UMEAN/MTOT (inner cells).
 * Outputs (caller-allocated):
 *   indice[nvoid] (1-based ids, descending volume), uvoid[nvoid],
 *   marca_ext[next^3],
 *   vf[13*nvoid]: gxc,gyc,gzc, xc,yc,zc, volnew, umean, mtot, eps, ip, req, (spare)
 *   ncellv[nvoid]
 *   stats_i[4]: nvoid2 (after merge), nvoidt (final), ncells_in_voids, (spare)
 *   stats_d[2]: volt_clean, ff
 * meandens, rete: MEANDENS (Msun/Mpc^3) and scale factor for MTOT. */
int32_t orc_merge_post(const orc_grid* g, const orc_void_params* p, int32_t nvoid,
                       const int32_t* ci, const float* cf, const int32_t* father,
                       const float* u1co_ext, float meandens, float rete,
                       int32_t* indice, int32_t* uvoid, int32_t* marca_ext,
                       float* vf, int32_t* ncellv, int32_t* stats_i, double* stats_d,
                       int32_t stop_after);

/* cosmology helpers (avism.f90:469-479, 599-606) */
void orc_cosmo(float ache, float omega0, float zeta, float* rote, float* rete,
               float* hte);

/* ---- MASCLET AMR gas (FLAG_DATA = 0) onto the uniform grid of level ir
 * (-3..1): base[4][nb^3] the base grid's gas density contrast and velocities as
 * the files hold them; npatch level-1 patches, pint[6*npatch] = PATCHNX, NY, NZ,
 * PATCHX, Y, Z (1-based base cell of the first cells), preal[3*npatch] =
 * PATCHRX, RY, RZ, pdata = per patch density contrast, vx, vy, vz (x fastest).
 * u1gco, divgco: n^3, n = nb * 2^ir.  0, or negative for an unsupported level
 * or a patch whose stencils leave the grids. */
int32_t orc_amr_gas(int32_t nb, float L0, int32_t periodic, int32_t ir, const float* base, int32_t npatch,
                    const int32_t* pint, const float* preal, const float* pdata, float* u1gco, float* divgco);

#ifdef __cplusplus
}
#endif
#endif
