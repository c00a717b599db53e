/*
 * oracle.h — FP64 OpenMP CPU oracle for the DEM–SPH per-timestep hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library restates the reference solver's
 * algorithms on the CPU so that tests and the benchmark's CPU-baseline leg
 * can check the HIP implementation against it.  Product code never links,
 * loads or calls it.
 *
 * The API mirrors include/dsph.h (same scene struct, same phase functions,
 * same read-back layout) with an `orc_` prefix.
 */
#ifndef DSPH_ORACLE_H
#define DSPH_ORACLE_H

#include <stdint.h>
#include "../include/dsph.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_ctx orc_ctx;

orc_ctx* orc_create(void);
void     orc_destroy(orc_ctx* o);
int      orc_upload_scene(orc_ctx* o, const dsph_scene* s);
int      orc_set_flags(orc_ctx* o, uint32_t flags);
int      orc_set_threads(orc_ctx* o, int n);
int      orc_initialize(orc_ctx* o, int compute_normals);
int      orc_neighbor_search(orc_ctx* o, int32_t i_step, int32_t gap);
int      orc_solid_before_contact(orc_ctx* o, double dt);
int      orc_fluid_integrate(orc_ctx* o, double dt, int32_t i_step, int32_t gap);
int      orc_contact_forces(orc_ctx* o, double dt);
int      orc_solid_after_contact(orc_ctx* o, double dt);
int      orc_step(orc_ctx* o, int32_t i_step, double dt, int32_t gap);
/* particles added after the run started (dsph_append_scene's semantics) */
int      orc_append_scene(orc_ctx* o, const dsph_scene* extra);
/* runtime mutation between the contact phase and solid_after_contact */
int      orc_add_external_force(orc_ctx* o, int32_t solid_index,
                                const double force[3]);
int      orc_add_global_damping(orc_ctx* o, int32_t solid_index, double c_d);

int orc_download_fluid(orc_ctx* o, double* pos, double* vel, double* acc,
                       double* drho, double* sumw0);
int orc_download_solid(orc_ctx* o, double* pos, double* vel, double* acc,
                       double* angvel, double* torque, double* normal);
int orc_download_clump(orc_ctx* o, double* centroid, double* vel, double* acc,
                       double* angvel, double* torque, double* orientation);
int orc_download_fluid_solid_force(orc_ctx* o, double* force, int64_t cap,
                                   int64_t* n_out);
int orc_download_contacts(orc_ctx* o, double* force, double* torque,
                          double* sliding, double* rolling, double* torsion,
                          int64_t cap, int64_t* n_out);
int orc_download_bonds(orc_ctx* o, double* normal_force, double* torsion,
                       double* shear, double* bending, int32_t* is_bonded,
                       int64_t cap, int64_t* n_out);
int orc_download_list(orc_ctx* o, int kind, int32_t* pointed,
                      int32_t* pointing, int64_t cap, int64_t* n_out,
                      int32_t* count);
/* transposed ("pointing") order of a list: index[] over pairs sorted stably
 * by pointing particle, and the per-pointing-particle [start,end) ranges   */
int orc_download_transpose(orc_ctx* o, int kind, int32_t* index, int64_t cap,
                           int32_t* range_start, int32_t* range_end);
int orc_grid_info(orc_ctx* o, int32_t* grid_size3, double* cell_size3,
                  int32_t* n_cells);
int orc_download_grid(orc_ctx* o, int species, int32_t* hash, int32_t* perm,
                      int32_t* cell_start, int32_t* cell_end);

/* --- single-formula entry points, for known-answer tests --------------- */
double orc_wendland(double r, double h);
void   orc_grad_wendland(const double rij[3], double h, double out[3]);
/* Hertz–Mindlin (model 0) or linear (model 1) contact between two spheres or
 * a sphere and a plane, one call of the reference's contact routine.
 * params: hertz (E,G,D,kr/ks,kt/ks,mu_s,mu_r,mu_t) or linear (11 values).
 * springs[9] in/out: sliding, rolling, torsion.  out[6]: force, torque.  */
void   orc_contact(int model, const double v_rel[3], const double w_rel[3],
                   const double n[3], double overlap, double m_eff,
                   double r_eff, double dt, const double* params,
                   double springs[9], double out[6]);
int    orc_parallel_bond(double state[8] /* Fn, Tt, Fs[3], Tb[3] */,
                         const double n_prev[3], const double n[3],
                         const double v_rel[3], const double w_a[3],
                         const double w_b[3], double r_a, double r_b,
                         double dt, const double params[6]);
void   orc_rotate_inverse_inertia(const double q[4], const double inv_i[6],
                                  double out[6]);
void   orc_quaternion_rotate(const double q[4], const double w[3], double dt,
                             double out[4]);
void   orc_rotate_vector(const double v[3], const double angle[3],
                         double out[3]);
int    orc_combined_index(int n_material, int a, int b);

/* ASCII VTU frames of host arrays as the reference's writers print them
 * (solverBase.cpp:69-333, vtu_ascii.cpp); 0, or -1 when the file cannot be
 * written.  Vectors are [3n] x,y,z. */
int orc_vtu_fluid_ascii(const char* path, double time, int32_t step, int64_t n,
                        const double* pos, const double* h, const double* vel,
                        const double* rho0, const double* drho, const double* c);
int orc_vtu_solid_ascii(const char* path, double time, int32_t step, int64_t n,
                        const double* pos, const double* radius,
                        const int32_t* cluster, const int32_t* clump,
                        const double* vel, const double* angvel,
                        const double* normal);
int orc_vtu_fluid_solid_ascii(const char* path, double time, int32_t step,
                              int64_t n, const int32_t* fluid,
                              const int32_t* solid, const double* fluid_pos,
                              const double* solid_pos, const double* force);

#ifdef __cplusplus
}
#endif
#endif
