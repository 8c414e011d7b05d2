/*
 * lbm_oracle.c -- FP64 CPU oracle of the AMR D3Q19 step (test infrastructure
 * only; never linked into the product library).  See lbm_oracle.h.
 *
 * Scheme (the build's specification; the reference has no LBM arithmetic):
 *  - D3Q19, velocities ordered rest, +-x, +-y, +-z (the order of the lattice
 *    fixture in tests/unit/parsing/test_parsing.cpp:23-29), then the 12 edge
 *    directions in opposite pairs; opposites found as in
 *    LatticeDescriptor::computeOpposites (LatticeDescriptor.cpp:30-49).
 *  - Stored state: post-collision populations f*; pull streaming
 *    f_i(x) = f*_i(x - c_i); BGK f* = f - (f - feq)/tau.
 *  - Walls: half-way bounce-back f_i(x) = f*_opp(x) + 6 w_i rho0 (c_i.u_w)
 *    (Ladd).  Outflow faces are pressure outlets: the population of the
 *    position clamped into the domain, with that cell's density replaced by
 *    rho0 at its velocity, f_i = f*_i(s) + (rho0 - rho_s) w_i (1 + 3 c.u_s +
 *    4.5 (c.u_s)^2 - 1.5 u_s^2) (non-equilibrium extrapolation).
 *  - Level l: dx = 2^-l, dt = 2^-l, tau_l = 2^l (tau0 - 1/2) + 1/2.
 *  - Coarse step advance(l): f2c fill, capture pre-collision state of c2f
 *    sources (from the previous buffer), collide-stream, capture again, then
 *    two sub-steps of level l+1 each preceded by a c2f fill at alpha = 0, 1/2.
 *  - c2f: f_fine* = feq(rho_g,u_g) + (1-1/tau_f) (tau_f/(2 tau_c)) fneq_c, with
 *    rho_g, u_g linearly interpolated from the coarse cell and its coarse
 *    fluid neighbours.
 *  - f2c: average rho, j, fneq of the 8 children's pre-collision states;
 *    f_coarse* = feq + (1-1/tau_c) (2 tau_c/tau_f) avg fneq.
 */
#include "lbm_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* Populations and per-cell arithmetic are `real`: double for the oracle;
 * the Makefile also builds a float copy (liboracle_f32.so) that only the
 * benchmark's CPU baseline times.  The API stays double either way. */
#ifndef OR_REAL
#define OR_REAL double
#endif
typedef OR_REAL real;

#define Q 19
enum { T_NONE = 0, T_FLUID = 1, T_SOLID = 2, T_C2F = 3, T_F2C = 4, T_WALL = 5 };

static const int CX[Q] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0};
static const int CY[Q] = {0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0, 1, -1, 1, -1};
static const int CZ[Q] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1};
static real W[Q];
static int OPP[Q];
static int lattice_ready = 0;

static void init_lattice(void) {
    if (lattice_ready) return;
    for (int i = 0; i < Q; ++i) {
        int n = abs(CX[i]) + abs(CY[i]) + abs(CZ[i]);
        W[i] = n == 0 ? 1.0 / 3.0 : (n == 1 ? 1.0 / 18.0 : 1.0 / 36.0);
        for (int j = 0; j < Q; ++j)
            if (CX[i] == -CX[j] && CY[i] == -CY[j] && CZ[i] == -CZ[j]) { OPP[i] = j; break; }
    }
    lattice_ready = 1;
}

void or_lbm_lattice(int* c, double* w, int* opp) {
    init_lattice();
    for (int i = 0; i < Q; ++i) {
        c[3 * i] = CX[i]; c[3 * i + 1] = CY[i]; c[3 * i + 2] = CZ[i];
        w[i] = W[i]; opp[i] = OPP[i];
    }
}

typedef struct {
    int X, Y, Z;
    size_t N;
    uint8_t* type;
    int64_t* c2f_src;   /* coarse position for T_C2F */
    real* f[2];       /* [Q][N] */
    real* cap[2];     /* [Q][N] at capture cells */
    int cur;
    real tau;
    size_t nfluid, nc2f, nf2c, ncap;
    int64_t* fluid;     /* positions */
    int64_t* c2f;
    int64_t* f2c;
    int64_t* capl;
    uint8_t* is_cap;
} level_t;

struct or_lbm {
    or_lbm_params p;
    int L;              /* number of levels */
    level_t lv[24];
    real c2f_mult, f2c_mult, slope_mult; /* 1.0; tests plant mistakes through these */
};

static inline int64_t pidx(const level_t* v, int x, int y, int z) {
    return ((int64_t)z * v->Y + y) * v->X + x;
}
static inline void pxyz(const level_t* v, int64_t p, int* x, int* y, int* z) {
    *x = (int)(p % v->X);
    *y = (int)((p / v->X) % v->Y);
    *z = (int)(p / ((int64_t)v->X * v->Y));
}
static inline int regular(int t) { return t == T_FLUID || t == T_C2F || t == T_F2C; }

static inline real feq(real rho, real ux, real uy, real uz, int i) {
    real cu = CX[i] * ux + CY[i] * uy + CZ[i] * uz;
    real usq = ux * ux + uy * uy + uz * uz;
    return W[i] * rho * (1.0 + 3.0 * cu + 4.5 * cu * cu - 1.5 * usq);
}

static inline void moments(const real* f, real* rho, real* ux, real* uy, real* uz) {
    real r = 0, jx = 0, jy = 0, jz = 0;
    for (int i = 0; i < Q; ++i) {
        r += f[i];
        jx += CX[i] * f[i]; jy += CY[i] * f[i]; jz += CZ[i] * f[i];
    }
    *rho = r; *ux = jx / r; *uy = jy / r; *uz = jz / r;
}

/* Source of direction i for position (x,y,z) after periodic wrap.  Returns
 * 0 inside, else 1 and the wall face (or -1 for outflow) in *face. */
static int source(const or_lbm* s, const level_t* v, int x, int y, int z, int i, int* sx, int* sy,
                  int* sz, int* face) {
    int p[3] = {x - CX[i], y - CY[i], z - CZ[i]};
    int e[3] = {v->X, v->Y, v->Z};
    int out[3] = {0, 0, 0};
    int any = 0;
    for (int a = 0; a < 3; ++a) {
        int f = -1;
        if (p[a] < 0) f = 2 * a;
        else if (p[a] >= e[a]) f = 2 * a + 1;
        if (f < 0) continue;
        if (s->p.bc[f] == OR_BC_PERIODIC) {
            p[a] = (p[a] + e[a]) % e[a];
        } else {
            out[a] = f + 1;
            any = 1;
        }
    }
    *sx = p[0]; *sy = p[1]; *sz = p[2];
    if (!any) return 0;
    *face = -1;
    for (int a = 2; a >= 0; --a)
        if (out[a] && s->p.bc[out[a] - 1] == OR_BC_WALL) { *face = out[a] - 1; break; }
    if (*face < 0) {
        /* outflow: clamp the outside axes */
        for (int a = 0; a < 3; ++a)
            if (out[a]) p[a] = p[a] < 0 ? 0 : e[a] - 1;
        *sx = p[0]; *sy = p[1]; *sz = p[2];
    }
    return 1;
}

/* Coarse fluid neighbour of coarse position cp one cell along axis a in
 * direction d (+1/-1), across periodic faces; -1 when outside the domain or
 * not a fluid leaf. */
static int64_t coarse_nb(const or_lbm* s, const level_t* c, int64_t cp, int a, int d) {
    int q[3];
    pxyz(c, cp, &q[0], &q[1], &q[2]);
    const int e[3] = {c->X, c->Y, c->Z};
    q[a] += d;
    if (q[a] < 0 || q[a] >= e[a]) {
        if (s->p.bc[2 * a + (d > 0)] != OR_BC_PERIODIC) return -1;
        q[a] = (q[a] + e[a]) % e[a];
    }
    int64_t np = pidx(c, q[0], q[1], q[2]);
    return c->type[np] == T_FLUID ? np : -1;
}

static void gather(const or_lbm* s, const level_t* v, int64_t pos, const real* buf, real* out) {
    int x, y, z;
    pxyz(v, pos, &x, &y, &z);
    const size_t N = v->N;
    for (int i = 0; i < Q; ++i) {
        int sx, sy, sz, face;
        int o = source(s, v, x, y, z, i, &sx, &sy, &sz, &face);
        if (o && face >= 0) {
            const double* uw = s->p.bc_u[face];
            out[i] = buf[(size_t)OPP[i] * N + pos] +
                     6.0 * W[i] * s->p.rho0 * (CX[i] * uw[0] + CY[i] * uw[1] + CZ[i] * uw[2]);
            continue;
        }
        int64_t sp = pidx(v, sx, sy, sz);
        if (!regular(v->type[sp])) {
            out[i] = buf[(size_t)OPP[i] * N + pos];
        } else if (!o) {
            out[i] = buf[(size_t)i * N + sp];
        } else {
            /* pressure outlet: the clamped source's population with its
             * density set to rho0 at its own velocity (non-equilibrium
             * extrapolation; feq is linear in rho) */
            real fs[Q], rs, us, vs, ws;
            for (int j = 0; j < Q; ++j) fs[j] = buf[(size_t)j * N + sp];
            moments(fs, &rs, &us, &vs, &ws);
            out[i] = fs[i] + (s->p.rho0 - rs) * feq(1.0, us, vs, ws, i);
        }
    }
}

static void collide(const real* f, real tau, real* o) {
    real rho, ux, uy, uz;
    moments(f, &rho, &ux, &uy, &uz);
    for (int i = 0; i < Q; ++i) o[i] = f[i] - (f[i] - feq(rho, ux, uy, uz, i)) / tau;
}

static int classify(or_lbm* s, int l, int64_t p) {
    level_t* v = &s->lv[l];
    if (v->type[p] != T_NONE) return 0;
    int x, y, z;
    pxyz(v, p, &x, &y, &z);
    if (l > 0) {
        level_t* c = &s->lv[l - 1];
        int64_t cp = pidx(c, x >> 1, y >> 1, z >> 1);
        if (c->type[cp] == T_FLUID) {
            v->type[p] = T_C2F;
            v->c2f_src[p] = cp;
            c->is_cap[cp] = 1;
            return 0;
        }
        if (c->type[cp] == T_SOLID) { v->type[p] = T_WALL; return 0; }
    }
    if (l + 1 < s->L) {
        level_t* f = &s->lv[l + 1];
        int nf = 0, ns = 0;
        for (int q = 0; q < 8; ++q) {
            int t = f->type[pidx(f, 2 * x + (q & 1), 2 * y + ((q >> 1) & 1), 2 * z + ((q >> 2) & 1))];
            nf += t == T_FLUID;
            ns += t == T_SOLID;
        }
        if (nf == 8) { v->type[p] = T_F2C; return 0; }
        if (nf + ns == 8) { v->type[p] = T_WALL; return 0; }
    }
    return -1;
}

or_lbm* or_lbm_create(const or_lbm_params* p, const int32_t* x, const int32_t* y,
                      const int32_t* z, const uint8_t* level, const uint8_t* state, size_t n,
                      double rho_init, const double u_init[3], int* err) {
    init_lattice();
    *err = 0;
    or_lbm* s = (or_lbm*)calloc(1, sizeof(or_lbm));
    s->p = *p;
    s->L = p->max_level + 1;
    s->c2f_mult = 1.0;
    s->f2c_mult = 1.0;
    s->slope_mult = 1.0;
    for (int l = 0; l < s->L; ++l) {
        level_t* v = &s->lv[l];
        v->X = p->nx0 << l; v->Y = p->ny0 << l; v->Z = p->nz0 << l;
        v->N = (size_t)v->X * v->Y * v->Z;
        v->type = (uint8_t*)calloc(v->N, 1);
        v->is_cap = (uint8_t*)calloc(v->N, 1);
        v->c2f_src = (int64_t*)malloc(v->N * sizeof(int64_t));
        v->tau = ldexp(p->tau0 - 0.5, l) + 0.5;
    }
    for (size_t k = 0; k < n; ++k) {
        int l = level[k];
        if (l >= s->L) { *err = 1; or_lbm_destroy(s); return NULL; }
        level_t* v = &s->lv[l];
        if (x[k] < 0 || y[k] < 0 || z[k] < 0 || x[k] >= v->X || y[k] >= v->Y || z[k] >= v->Z) {
            *err = 2; or_lbm_destroy(s); return NULL;
        }
        int64_t q = pidx(v, x[k], y[k], z[k]);
        if (v->type[q] != T_NONE) { *err = 3; or_lbm_destroy(s); return NULL; }
        v->type[q] = state[k] == 0 ? T_FLUID : T_SOLID;
    }
    /* classify every source of every fluid leaf */
    for (int l = 0; l < s->L; ++l) {
        level_t* v = &s->lv[l];
        for (int64_t q = 0; q < (int64_t)v->N; ++q) {
            if (v->type[q] != T_FLUID) continue;
            int xx, yy, zz;
            pxyz(v, q, &xx, &yy, &zz);
            for (int i = 1; i < Q; ++i) {
                int sx, sy, sz, face;
                int o = source(s, v, xx, yy, zz, i, &sx, &sy, &sz, &face);
                if (o && face >= 0) continue;
                if (classify(s, l, pidx(v, sx, sy, sz)) != 0) {
                    *err = 4; or_lbm_destroy(s); return NULL;
                }
            }
        }
    }
    /* the coarse fluid neighbours of every c2f source are captured too: the
     * fill interpolates the coarse density and velocity linearly in space */
    for (int l = 1; l < s->L; ++l) {
        level_t* v = &s->lv[l];
        level_t* c = &s->lv[l - 1];
        for (size_t q = 0; q < v->N; ++q) {
            if (v->type[q] != T_C2F) continue;
            for (int a = 0; a < 3; ++a)
                for (int d = -1; d <= 1; d += 2) {
                    int64_t np = coarse_nb(s, c, v->c2f_src[q], a, d);
                    if (np >= 0) c->is_cap[np] = 1;
                }
        }
    }
    for (int l = 0; l < s->L; ++l) {
        level_t* v = &s->lv[l];
        size_t nf = 0, nc = 0, nd = 0, np = 0;
        for (size_t q = 0; q < v->N; ++q) {
            nf += v->type[q] == T_FLUID;
            nc += v->type[q] == T_C2F;
            nd += v->type[q] == T_F2C;
            np += v->is_cap[q];
        }
        v->fluid = (int64_t*)malloc((nf + 1) * sizeof(int64_t));
        v->c2f = (int64_t*)malloc((nc + 1) * sizeof(int64_t));
        v->f2c = (int64_t*)malloc((nd + 1) * sizeof(int64_t));
        v->capl = (int64_t*)malloc((np + 1) * sizeof(int64_t));
        v->nfluid = v->nc2f = v->nf2c = v->ncap = 0;
        for (size_t q = 0; q < v->N; ++q) {
            if (v->type[q] == T_FLUID) v->fluid[v->nfluid++] = (int64_t)q;
            if (v->type[q] == T_C2F) v->c2f[v->nc2f++] = (int64_t)q;
            if (v->type[q] == T_F2C) v->f2c[v->nf2c++] = (int64_t)q;
            if (v->is_cap[q]) v->capl[v->ncap++] = (int64_t)q;
        }
        for (int b = 0; b < 2; ++b) {
            v->f[b] = (real*)calloc((size_t)Q * v->N, sizeof(real));
            v->cap[b] = (real*)calloc((size_t)Q * v->N, sizeof(real));
        }
        for (size_t q = 0; q < v->N; ++q) {
            if (!regular(v->type[q])) continue;
            for (int i = 0; i < Q; ++i) {
                real e = feq(rho_init, u_init[0], u_init[1], u_init[2], i);
                v->f[0][(size_t)i * v->N + q] = e;
                v->f[1][(size_t)i * v->N + q] = e;
            }
        }
        v->cur = 0;
    }
    return s;
}

void or_lbm_destroy(or_lbm* s) {
    if (!s) return;
    for (int l = 0; l < s->L; ++l) {
        level_t* v = &s->lv[l];
        free(v->type); free(v->is_cap); free(v->c2f_src);
        free(v->f[0]); free(v->f[1]); free(v->cap[0]); free(v->cap[1]);
        free(v->fluid); free(v->c2f); free(v->f2c); free(v->capl);
    }
    free(s);
}

static void fill_f2c(or_lbm* s, int l) {
    level_t* v = &s->lv[l];
    level_t* f = &s->lv[l + 1];
    const real* fprev = f->f[1 - f->cur];
    real* dst = v->f[v->cur];
    const real fac = (1.0 - 1.0 / v->tau) * (2.0 * v->tau / f->tau) * s->f2c_mult;
#pragma omp parallel for schedule(static)
    for (int64_t k = 0; k < (int64_t)v->nf2c; ++k) {
        int64_t g = v->f2c[k];
        int x, y, z;
        pxyz(v, g, &x, &y, &z);
        real sr = 0, sjx = 0, sjy = 0, sjz = 0, sneq[Q];
        for (int i = 0; i < Q; ++i) sneq[i] = 0;
        for (int q = 0; q < 8; ++q) {
            int64_t cp = pidx(f, 2 * x + (q & 1), 2 * y + ((q >> 1) & 1), 2 * z + ((q >> 2) & 1));
            real fk[Q], r, ux, uy, uz;
            gather(s, f, cp, fprev, fk);
            moments(fk, &r, &ux, &uy, &uz);
            sr += r; sjx += r * ux; sjy += r * uy; sjz += r * uz;
            for (int i = 0; i < Q; ++i) sneq[i] += fk[i] - feq(r, ux, uy, uz, i);
        }
        real r = sr * 0.125;
        real ux = (sjx * 0.125) / r, uy = (sjy * 0.125) / r, uz = (sjz * 0.125) / r;
        for (int i = 0; i < Q; ++i)
            dst[(size_t)i * v->N + g] = feq(r, ux, uy, uz, i) + fac * (sneq[i] * 0.125);
    }
}

static void capture(or_lbm* s, int l, int slot) {
    level_t* v = &s->lv[l];
    const real* prev = v->f[1 - v->cur];
#pragma omp parallel for schedule(static)
    for (int64_t k = 0; k < (int64_t)v->ncap; ++k) {
        int64_t c = v->capl[k];
        real fk[Q];
        gather(s, v, c, prev, fk);
        for (int i = 0; i < Q; ++i) v->cap[slot][(size_t)i * v->N + c] = fk[i];
    }
}

static void cap_moments(const level_t* c, int64_t cp, real alpha, real* f, real* m) {
    for (int i = 0; i < Q; ++i)
        f[i] = (1.0 - alpha) * c->cap[0][(size_t)i * c->N + cp] + alpha * c->cap[1][(size_t)i * c->N + cp];
    moments(f, &m[0], &m[1], &m[2], &m[3]);
}

/* c2f: density and velocity interpolated linearly from the covering coarse
 * cell P to the fine cell centre (offset +-1/4 coarse cell per axis), with
 * per-axis slopes from P's coarse fluid neighbours (central difference, or
 * one-sided where only one side is a fluid leaf, or none); the
 * non-equilibrium part is P's, rescaled:
 *   f* = feq(rho_g, u_g) + (1 - 1/tau_f)(tau_f / (2 tau_c)) (f_P - feq(rho_P, u_P)) */
static void fill_c2f(or_lbm* s, int l, real alpha) {
    level_t* v = &s->lv[l];
    level_t* c = &s->lv[l - 1];
    real* dst = v->f[v->cur];
    const real fac = (1.0 - 1.0 / v->tau) * (v->tau / (2.0 * c->tau)) * s->c2f_mult;
#pragma omp parallel for schedule(static)
    for (int64_t k = 0; k < (int64_t)v->nc2f; ++k) {
        int64_t g = v->c2f[k];
        int64_t cp = v->c2f_src[g];
        int gc[3];
        pxyz(v, g, &gc[0], &gc[1], &gc[2]);
        real fp[Q], mp[4], fn[Q], mm[4], mq[4], phi[4];
        cap_moments(c, cp, alpha, fp, mp);
        for (int j = 0; j < 4; ++j) phi[j] = mp[j];
        for (int a = 0; a < 3; ++a) {
            const int64_t nm = coarse_nb(s, c, cp, a, -1), nq = coarse_nb(s, c, cp, a, +1);
            if (nm >= 0) cap_moments(c, nm, alpha, fn, mm);
            if (nq >= 0) cap_moments(c, nq, alpha, fn, mq);
            const real dl = ((gc[a] & 1) ? 0.25 : -0.25) * s->slope_mult;
            for (int j = 0; j < 4; ++j) {
                real gr = 0.0;
                if (nm >= 0 && nq >= 0) gr = 0.5 * (mq[j] - mm[j]);
                else if (nq >= 0) gr = mq[j] - mp[j];
                else if (nm >= 0) gr = mp[j] - mm[j];
                phi[j] += dl * gr;
            }
        }
        for (int i = 0; i < Q; ++i) {
            const real e = feq(phi[0], phi[1], phi[2], phi[3], i);
            dst[(size_t)i * v->N + g] = e + fac * (fp[i] - feq(mp[0], mp[1], mp[2], mp[3], i));
        }
    }
}

static void collide_stream(or_lbm* s, int l) {
    level_t* v = &s->lv[l];
    const real* src = v->f[v->cur];
    real* dst = v->f[1 - v->cur];
#pragma omp parallel for schedule(static)
    for (int64_t k = 0; k < (int64_t)v->nfluid; ++k) {
        int64_t p = v->fluid[k];
        real fk[Q], o[Q];
        gather(s, v, p, src, fk);
        collide(fk, v->tau, o);
        for (int i = 0; i < Q; ++i) dst[(size_t)i * v->N + p] = o[i];
    }
    v->cur = 1 - v->cur;
}

static void advance(or_lbm* s, int l) {
    int fine = l + 1 < s->L;
    if (fine) {
        fill_f2c(s, l);
        capture(s, l, 0);
    }
    collide_stream(s, l);
    if (fine) {
        capture(s, l, 1);
        for (int sub = 0; sub < 2; ++sub) {
            fill_c2f(s, l + 1, 0.5 * sub);
            advance(s, l + 1);
        }
    }
}

void or_lbm_step(or_lbm* s, int nsteps) {
    for (int t = 0; t < nsteps; ++t) advance(s, 0);
}

static int64_t leaf_pos(const or_lbm* s, int32_t x, int32_t y, int32_t z, int l) {
    if (l >= s->L) return -1;
    const level_t* v = &s->lv[l];
    if (x < 0 || y < 0 || z < 0 || x >= v->X || y >= v->Y || z >= v->Z) return -1;
    return pidx(v, x, y, z);
}

void or_lbm_get_f(const or_lbm* s, const int32_t* x, const int32_t* y, const int32_t* z,
                  const uint8_t* level, size_t n, double* f) {
    for (size_t k = 0; k < n; ++k) {
        int64_t p = leaf_pos(s, x[k], y[k], z[k], level[k]);
        const level_t* v = &s->lv[level[k]];
        for (int i = 0; i < Q; ++i)
            f[k * Q + i] = p < 0 ? NAN : v->f[v->cur][(size_t)i * v->N + p];
    }
}

void or_lbm_set_f(or_lbm* s, const int32_t* x, const int32_t* y, const int32_t* z,
                  const uint8_t* level, size_t n, const double* f) {
    for (size_t k = 0; k < n; ++k) {
        int64_t p = leaf_pos(s, x[k], y[k], z[k], level[k]);
        if (p < 0) continue;
        level_t* v = &s->lv[level[k]];
        for (int i = 0; i < Q; ++i) {
            v->f[0][(size_t)i * v->N + p] = f[k * Q + i];
            v->f[1][(size_t)i * v->N + p] = f[k * Q + i];
        }
    }
}

void or_lbm_moments(const or_lbm* s, const int32_t* x, const int32_t* y, const int32_t* z,
                    const uint8_t* level, size_t n, double* rho, double* u) {
    for (size_t k = 0; k < n; ++k) {
        int64_t p = leaf_pos(s, x[k], y[k], z[k], level[k]);
        if (p < 0) { rho[k] = NAN; u[3 * k] = u[3 * k + 1] = u[3 * k + 2] = NAN; continue; }
        const level_t* v = &s->lv[level[k]];
        real f[Q];
        for (int i = 0; i < Q; ++i) f[i] = v->f[v->cur][(size_t)i * v->N + p];
        real r, ux, uy, uz;
        moments(f, &r, &ux, &uy, &uz);
        rho[k] = r; u[3 * k] = ux; u[3 * k + 1] = uy; u[3 * k + 2] = uz;
    }
}

void or_lbm_totals(const or_lbm* s, double* mass, double mom[3]) {
    *mass = 0; mom[0] = mom[1] = mom[2] = 0;
    for (int l = 0; l < s->L; ++l) {
        const level_t* v = &s->lv[l];
        double vol = ldexp(1.0, -3 * l);
        for (size_t k = 0; k < v->nfluid; ++k) {
            int64_t p = v->fluid[k];
            real f[Q], r, ux, uy, uz;
            for (int i = 0; i < Q; ++i) f[i] = v->f[v->cur][(size_t)i * v->N + p];
            moments(f, &r, &ux, &uy, &uz);
            *mass += r * vol;
            mom[0] += r * ux * vol; mom[1] += r * uy * vol; mom[2] += r * uz * vol;
        }
    }
}

void or_lbm_force(const or_lbm* s, double F[3]) {
    F[0] = F[1] = F[2] = 0;
    for (int l = 0; l < s->L; ++l) {
        const level_t* v = &s->lv[l];
        real fx = 0, fy = 0, fz = 0;
        const real* b = v->f[v->cur];
        for (size_t k = 0; k < v->nfluid; ++k) {
            int64_t p = v->fluid[k];
            int x, y, z;
            pxyz(v, p, &x, &y, &z);
            for (int i = 1; i < Q; ++i) {
                int sx, sy, sz, face;
                if (source(s, v, x, y, z, i, &sx, &sy, &sz, &face)) continue;
                int t = v->type[pidx(v, sx, sy, sz)];
                if (t != T_SOLID && t != T_WALL) continue;
                real fo = b[(size_t)OPP[i] * v->N + p];
                fx += -2.0 * CX[i] * fo; fy += -2.0 * CY[i] * fo; fz += -2.0 * CZ[i] * fo;
            }
        }
        double sc = ldexp(1.0, -2 * l);
        F[0] += fx * sc; F[1] += fy * sc; F[2] += fz * sc;
    }
}

void or_lbm_plant_coupling_error(or_lbm* s, double c2f_mult, double f2c_mult, double slope_mult) {
    s->c2f_mult = c2f_mult;
    s->f2c_mult = f2c_mult;
    s->slope_mult = slope_mult;
}

/* ---- regrid support ------------------------------------------------------
 * Vorticity magnitude (level-0 units) at fluid leaves: central differences
 * over same-level fluid-leaf neighbours along each axis, one-sided when only
 * one side is a fluid leaf, zero when neither is; |w| * 2^level. */
static int leaf_u(const or_lbm* s, int l, int x, int y, int z, real* u) {
    const level_t* v = &s->lv[l];
    int p[3] = {x, y, z};
    int e[3] = {v->X, v->Y, v->Z};
    for (int a = 0; a < 3; ++a) {
        if (p[a] >= 0 && p[a] < e[a]) continue;
        if (s->p.bc[2 * a] != OR_BC_PERIODIC) return 0;
        p[a] = (p[a] + e[a]) % e[a];
    }
    int64_t q = pidx(v, p[0], p[1], p[2]);
    if (v->type[q] != T_FLUID) return 0;
    real f[Q], r;
    for (int i = 0; i < Q; ++i) f[i] = v->f[v->cur][(size_t)i * v->N + q];
    moments(f, &r, &u[0], &u[1], &u[2]);
    return 1;
}

void or_lbm_vorticity(const or_lbm* s, const int32_t* x, const int32_t* y, const int32_t* z,
                      const uint8_t* level, size_t n, double* out) {
    for (size_t k = 0; k < n; ++k) {
        int l = level[k];
        real u0[3], g[3][3];
        if (l >= s->L || !leaf_u(s, l, x[k], y[k], z[k], u0)) { out[k] = 0; continue; }
        for (int a = 0; a < 3; ++a) {
            int dp[3] = {0, 0, 0};
            dp[a] = 1;
            real up[3], um[3];
            int hp = leaf_u(s, l, x[k] + dp[0], y[k] + dp[1], z[k] + dp[2], up);
            int hm = leaf_u(s, l, x[k] - dp[0], y[k] - dp[1], z[k] - dp[2], um);
            for (int b = 0; b < 3; ++b) {
                if (hp && hm) g[a][b] = 0.5 * (up[b] - um[b]);
                else if (hp) g[a][b] = up[b] - u0[b];
                else if (hm) g[a][b] = u0[b] - um[b];
                else g[a][b] = 0.0;
            }
        }
        real wx = g[1][2] - g[2][1], wy = g[2][0] - g[0][2], wz = g[0][1] - g[1][0];
        out[k] = sqrt(wx * wx + wy * wy + wz * wz) * ldexp(1.0, l);
    }
}

/* New solver over a new leaf list with the populations carried over:
 * a leaf present before keeps its post-collision populations; a child of a
 * split parent gets feq(parent) + (1-1/tau_f)(tau_f/(2 tau_c)) fneq(parent);
 * a merged parent gets feq(avg) + (1-1/tau_c)(2 tau_c/tau_f) avg fneq of its
 * 8 children (pre-collision states gathered from the previous buffer, as in
 * the ghost fill); anything else starts at equilibrium(1, 0).  Both buffers
 * of the new solver hold the result; ghost slots restart at the
 * equilibrium of the leaves covering them. */
or_lbm* or_lbm_regrid(const or_lbm* o, const int32_t* x, const int32_t* y, const int32_t* z,
                      const uint8_t* level, const uint8_t* state, size_t n, int* err) {
    const double zero[3] = {0, 0, 0};
    or_lbm* s = or_lbm_create(&o->p, x, y, z, level, state, n, 1.0, zero, err);
    if (!s) return NULL;
    for (int l = 0; l < s->L; ++l) {
        level_t* v = &s->lv[l];
        for (size_t k = 0; k < v->nfluid; ++k) {
            int64_t p = v->fluid[k];
            int px, py, pz;
            pxyz(v, p, &px, &py, &pz);
            real f[Q];
            const level_t* ov = &o->lv[l];
            if (ov->type[p] == T_FLUID) {
                for (int i = 0; i < Q; ++i) f[i] = ov->f[ov->cur][(size_t)i * ov->N + p];
            } else if (l > 0 && o->lv[l - 1].type[pidx(&o->lv[l - 1], px >> 1, py >> 1, pz >> 1)] == T_FLUID) {
                const level_t* oc = &o->lv[l - 1];
                real fp[Q], r, ux, uy, uz;
                gather(o, oc, pidx(oc, px >> 1, py >> 1, pz >> 1), oc->f[1 - oc->cur], fp);
                moments(fp, &r, &ux, &uy, &uz);
                real fac = (1.0 - 1.0 / v->tau) * (v->tau / (2.0 * oc->tau));
                for (int i = 0; i < Q; ++i) {
                    real e = feq(r, ux, uy, uz, i);
                    f[i] = e + fac * (fp[i] - e);
                }
            } else {
                int all = l + 1 < o->L;
                for (int q = 0; q < 8 && all; ++q) {
                    const level_t* of = &o->lv[l + 1];
                    all = of->type[pidx(of, 2 * px + (q & 1), 2 * py + ((q >> 1) & 1), 2 * pz + ((q >> 2) & 1))] == T_FLUID;
                }
                if (all) {
                    const level_t* of = &o->lv[l + 1];
                    real sr = 0, sjx = 0, sjy = 0, sjz = 0, sneq[Q];
                    for (int i = 0; i < Q; ++i) sneq[i] = 0;
                    for (int q = 0; q < 8; ++q) {
                        real fk[Q], r, ux, uy, uz;
                        gather(o, of, pidx(of, 2 * px + (q & 1), 2 * py + ((q >> 1) & 1), 2 * pz + ((q >> 2) & 1)),
                               of->f[1 - of->cur], fk);
                        moments(fk, &r, &ux, &uy, &uz);
                        sr += r; sjx += r * ux; sjy += r * uy; sjz += r * uz;
                        for (int i = 0; i < Q; ++i) sneq[i] += fk[i] - feq(r, ux, uy, uz, i);
                    }
                    real r = sr * 0.125;
                    real ux = (sjx * 0.125) / r, uy = (sjy * 0.125) / r, uz = (sjz * 0.125) / r;
                    real fac = (1.0 - 1.0 / v->tau) * (2.0 * v->tau / of->tau);
                    for (int i = 0; i < Q; ++i) f[i] = feq(r, ux, uy, uz, i) + fac * (sneq[i] * 0.125);
                } else {
                    for (int i = 0; i < Q; ++i) f[i] = feq(1.0, 0, 0, 0, i);
                }
            }
            for (int i = 0; i < Q; ++i) {
                v->f[0][(size_t)i * v->N + p] = f[i];
                v->f[1][(size_t)i * v->N + p] = f[i];
            }
        }
    }
    /* ghosts restart at the equilibrium of what covers them: the coarse
     * leaf (c2f) or the average density and momentum of the 8 children
     * (f2c), in both buffers */
    for (int l = 0; l < s->L; ++l) {
        level_t* v = &s->lv[l];
        for (size_t k = 0; k < v->nc2f; ++k) {
            int64_t g = v->c2f[k];
            const level_t* c = &s->lv[l - 1];
            real f[Q], r, ux, uy, uz;
            for (int i = 0; i < Q; ++i) f[i] = c->f[0][(size_t)i * c->N + v->c2f_src[g]];
            moments(f, &r, &ux, &uy, &uz);
            for (int i = 0; i < Q; ++i) {
                real e = feq(r, ux, uy, uz, i);
                v->f[0][(size_t)i * v->N + g] = e;
                v->f[1][(size_t)i * v->N + g] = e;
            }
        }
        for (size_t k = 0; k < v->nf2c; ++k) {
            int64_t g = v->f2c[k];
            const level_t* fv = &s->lv[l + 1];
            int px, py, pz;
            pxyz(v, g, &px, &py, &pz);
            real sr = 0, sjx = 0, sjy = 0, sjz = 0;
            for (int q = 0; q < 8; ++q) {
                int64_t cp = pidx(fv, 2 * px + (q & 1), 2 * py + ((q >> 1) & 1), 2 * pz + ((q >> 2) & 1));
                real f[Q], r, ux, uy, uz;
                for (int i = 0; i < Q; ++i) f[i] = fv->f[0][(size_t)i * fv->N + cp];
                moments(f, &r, &ux, &uy, &uz);
                sr += r; sjx += r * ux; sjy += r * uy; sjz += r * uz;
            }
            real r = sr * 0.125;
            real ux = (sjx * 0.125) / r, uy = (sjy * 0.125) / r, uz = (sjz * 0.125) / r;
            for (int i = 0; i < Q; ++i) {
                real e = feq(r, ux, uy, uz, i);
                v->f[0][(size_t)i * v->N + g] = e;
                v->f[1][(size_t)i * v->N + g] = e;
            }
        }
    }
    return s;
}
