/*
 * lbm_oracle.h -- FP64 CPU oracle of the AMR D3Q19 lattice-Boltzmann step.
 *
 * TEST INFRASTRUCTURE ONLY (never linked into the product library).
 *
 * The reference has no collide/stream arithmetic (the DSL emits an empty
 * body for lbm* kernels, src/parsing/SimulationBuilder.cpp:394-400); its
 * lattice constants come from src/parsing/registry/LatticeDescriptor.cpp:30-80
 * and the 7-entry fixture tests/unit/parsing/test_parsing.cpp:19-31.  The
 * scheme restated here is the build's own specification (documented in the
 * solver source): BGK, pull streaming, half-way bounce-back with Ladd's
 * moving-wall term, outflow copy, acoustic level scaling with time
 * sub-cycling and f_neq-rescaled coarse<->fine ghost fill.  Dense per-level
 * arrays keep this restatement independent of the GPU brick layout.
 */
#ifndef FL_LBM_ORACLE_H
#define FL_LBM_ORACLE_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { OR_BC_WALL = 0, OR_BC_OUTFLOW = 1, OR_BC_PERIODIC = 2 };

typedef struct {
    int nx0, ny0, nz0;     /* level-0 grid */
    int max_level;         /* finest level present */
    double tau0;           /* relaxation time at level 0 */
    int bc[6];             /* faces -x,+x,-y,+y,-z,+z */
    double bc_u[6][3];     /* wall velocity per face (lattice units) */
    double rho0;           /* wall density in the moving-wall term */
} or_lbm_params;

typedef struct or_lbm or_lbm;

/* Build from a leaf list (native coords per level, state 0 fluid 1 solid).
 * All fluid leaves and ghosts start at equilibrium(rho_init, u_init).
 * Returns NULL and sets *err if the mesh violates the coupling rules. */
or_lbm* or_lbm_create(const or_lbm_params* p, const int32_t* x, const int32_t* y,
                      const int32_t* z, const uint8_t* level, const uint8_t* state, size_t n,
                      double rho_init, const double u_init[3], int* err);
void or_lbm_destroy(or_lbm* s);
/* advance nsteps coarse (level-0) steps */
void or_lbm_step(or_lbm* s, int nsteps);
/* post-collision populations of the given leaves (19 per leaf, [leaf][q]) */
void or_lbm_get_f(const or_lbm* s, const int32_t* x, const int32_t* y, const int32_t* z,
                  const uint8_t* level, size_t n, double* f);
/* overwrite post-collision populations of given leaves (both buffers) */
void or_lbm_set_f(or_lbm* s, const int32_t* x, const int32_t* y, const int32_t* z,
                  const uint8_t* level, size_t n, const double* f);
/* density and velocity at the given leaves */
void or_lbm_moments(const or_lbm* s, const int32_t* x, const int32_t* y, const int32_t* z,
                    const uint8_t* level, size_t n, double* rho, double* u);
/* total mass (sum rho * 8^-level) and momentum over fluid leaves */
void or_lbm_totals(const or_lbm* s, double* mass, double mom[3]);
/* momentum-exchange force on solid leaves, in level-0 lattice units */
void or_lbm_force(const or_lbm* s, double F[3]);

/* Vorticity magnitude at the given leaves, level-0 units (regrid criterion) */
void or_lbm_vorticity(const or_lbm* s, const int32_t* x, const int32_t* y, const int32_t* z,
                      const uint8_t* level, size_t n, double* out);
/* A new solver on a new leaf list with populations carried over from `o`
 * (copy / split / merge rules in lbm_oracle.c). */
or_lbm* or_lbm_regrid(const or_lbm* o, const int32_t* x, const int32_t* y, const int32_t* z,
                      const uint8_t* level, const uint8_t* state, size_t n, int* err);

/* Tests only: multiply the c2f / f2c f_neq rescale factors (1.0 = the
 * scheme) so that a test can plant a plausible mistake and see it fail. */
void or_lbm_plant_coupling_error(or_lbm* s, double c2f_mult, double f2c_mult, double slope_mult);

/* D3Q19 constants as used by the oracle (for tests) */
void or_lbm_lattice(int* c /*19*3*/, double* w /*19*/, int* opp /*19*/);

#ifdef __cplusplus
}
#endif
#endif
