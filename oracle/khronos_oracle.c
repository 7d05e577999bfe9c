/*
 * khronos_oracle.c -- CPU restatement of the reference's time step.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP library:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it,
 * and never as the thing measured or shipped.
 *
 * It restates, on the reference's own array layout (column-major, x fastest,
 * one ghost layer per side, staggered component extents src/utils.jl:26-37,
 * src/Fields.jl:131-142), what the reference executes on its KernelAbstractions
 * CPU() backend in a single-chunk simulation (num_chunks = 1):
 *
 *   step order ................ src/Kernels/Kernels.jl:20-88
 *   curl (forward/backward) ... src/Kernels/Helpers.jl:283-298
 *   sigma cascade dispatch .... src/Kernels/Helpers.jl:30-270 (generic_curl!)
 *   W/material stage .......... src/Kernels/Helpers.jl:323-366
 *   curl kernel, update kernel  src/Kernels/ReferenceKernels.jl:276-313, :468-512
 *   source injection .......... src/Sources/Sources.jl:346-357
 *   Gaussian / CW waveform .... src/Sources/TimeSources.jl:61-64, :123-132
 *   DFT accumulation .......... src/Monitors/Monitors.jl:274-290, :361-379
 *   decimation gate ........... src/Kernels/Kernels.jl:465-497
 *   source shut-off ........... src/Kernels/Kernels.jl:27-35
 *   time of a step ............ src/Simulation.jl:22
 *   chi3 Kerr correction ...... src/Kernels/Dispersive.jl:127-173
 *   ADE poles, total P ........ src/Kernels/Dispersive.jl:25-117, :186-228
 *   P subtracted in the update  src/Kernels/Helpers.jl:330-333
 *   periodic / Bloch wrap ..... src/Chunking.jl:1725-1768 (single chunk),
 *                               :1825-1852, :2138-2214 (copy + phase)
 *   complex fields (Bloch) .... src/Fields.jl:140-159: every field array is
 *                               complex; held here as real and imaginary parts
 *                               (*i arrays), updated by the same real-coefficient
 *                               operations part by part
 *
 * Build twice: -DREAL=float (liboracle_f32.so) and -DREAL=double
 * (liboracle_f64.so), always with -ffp-contract=off (Julia does not contract
 * a*b+c into FMA).
 *
 * Auxiliary-array presence follows src/Fields.jl:266-305 for one chunk: with
 * PML boundaries every component has U and W; C exists only with a material
 * conductivity array.  A method combination the reference has no method for
 * (e.g. C without U and a scalar sigma_next) is rejected with an error.
 */
#include <math.h>
#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifndef REAL
#define REAL double
#endif

typedef REAL real;

#define NCOMP 6
enum { EX = 0, EY, EZ, HX, HY, HZ };

typedef struct {
    int comp, start[3], dims[3];
    real *amp; /* interleaved re, im; column-major dims */
    int kind;  /* 0 gaussian, 1 cw, 2 table */
    double p[4];
    double *table; /* 2*n entries, values per step of the half step that uses it */
    int64_t ntable;
} orc_source;

typedef struct {
    int comp, start[3], dims[3], nf, decim, acc_f64;
    double *freqs;
    void *acc; /* complex, (dims..., nf), column-major; float pairs or double pairs */
} orc_dft;

typedef struct {
    /* src/Susceptibility.jl:60-85, cast to the backend type (Dispersive.jl:214-219) */
    real g1i, g1, w, sw, dc;
    int drude;
    real *sigma;         /* (nx, ny, nz) oscillator strength, shared by x/y/z */
    real *P[3], *Pp[3];  /* (nx+2, ny+2, nz+2) */
    real *Pi[3], *Ppi[3];/* imaginary parts with complex fields (P complex: the
                            reference's real P cannot hold a complex E) */
} orc_pole;

typedef struct {
    int nx, ny, nz;
    real dx, dy, dz, dt;
    int dim[NCOMP][3];   /* raw extents incl. ghosts */
    int gdim[NCOMP][3];  /* geometry extents (component voxel count) */
    real *F[NCOMP];      /* E/H */
    real *T[NCOMP];      /* D/B */
    real *C[NCOMP], *U[NCOMP], *W[NCOMP], *S[NCOMP];
    real *sig[2][3];     /* [0]=B (H step), [1]=D (E step); length 2N+1 or NULL */
    real *msig[NCOMP];   /* material conductivity sigma_D / sigma_B (geometry layout) */
    real *minv[NCOMP];   /* per-voxel eps^-1 / mu^-1 (geometry layout) or NULL */
    real minv_s[2];      /* scalar mu^-1 (H), eps^-1 (E) */
    int has_bound;
    double cutoff;
    int sources_active;
    int64_t timestep;
    orc_source *src;
    int nsrc;
    orc_dft *dft;
    int ndft;
    int pdim[3];         /* (nx+2, ny+2, nz+2): polarization array extents */
    real *chi3;          /* (nx, ny, nz) or NULL */
    real *fPD[3];        /* total polarization, allocated with the first pole */
    orc_pole *pole;
    int npole;
    int cplx;            /* complex fields: imaginary parts below */
    real *Fi[NCOMP], *Ti[NCOMP], *Ci[NCOMP], *Ui[NCOMP], *Wi[NCOMP], *Si[NCOMP];
    real *fPDi[3];
    int periodic[3];     /* axis wraps (Periodic or Bloch on both sides) */
    /* the default 3D + PML chunk plan (27 chunks; src/Chunking.jl:622-738,
     * :907-923): a cell's chunk has PML along axis a when the cell is in
     * [1, pml_end[a]] or [pml_start[a], N]; that decides which auxiliary
     * arrays the chunk holds (:1090-1122) and zero-fills the profiles of its
     * other axes (:1355-1378) */
    int chunk27;
    int pml_end[3], pml_start[3];
    double ph_fwd[3][2]; /* lower interior -> upper ghost: exp(+i k L) */
    double ph_rev[3][2]; /* upper interior -> lower ghost: exp(-i k L) */
} orc_sim;

static void comp_count(int nx, int ny, int nz, int c, int out[3]) {
    /* src/utils.jl:26-37 */
    int n[3] = {nx, ny, nz < 1 ? 1 : nz};
    int a = c % 3;
    if (c < 3) {
        out[0] = n[0] + 1; out[1] = n[1] + 1; out[2] = n[2] + 1;
        out[a] = n[a];
    } else {
        out[0] = n[0]; out[1] = n[1]; out[2] = n[2];
        out[a] = n[a] + 1;
    }
}

static inline size_t rix(const int *d, int i, int j, int k) {
    /* 1-based raw index into a column-major array of extents d */
    return (size_t)(i - 1) + (size_t)d[0] * ((size_t)(j - 1) + (size_t)d[1] * (size_t)(k - 1));
}

static real *zalloc(size_t n) { return (real *)calloc(n ? n : 1, sizeof(real)); }

static size_t vol(const int *d) { return (size_t)d[0] * d[1] * d[2]; }

orc_sim *orc_create(int nx, int ny, int nz, double dx, double dy, double dz, double dt,
                    const real *sig_b0, const real *sig_b1, const real *sig_b2,
                    const real *sig_d0, const real *sig_d1, const real *sig_d2,
                    const real *minv0, const real *minv1, const real *minv2,
                    const real *minv3, const real *minv4, const real *minv5,
                    const real *msig0, const real *msig1, const real *msig2,
                    const real *msig3, const real *msig4, const real *msig5,
                    double eps_inv_s, double mu_inv_s, double cutoff) {
    orc_sim *s = (orc_sim *)calloc(1, sizeof(orc_sim));
    const real *sb[3] = {sig_b0, sig_b1, sig_b2};
    const real *sd[3] = {sig_d0, sig_d1, sig_d2};
    const real *mi[6] = {minv0, minv1, minv2, minv3, minv4, minv5};
    const real *ms[6] = {msig0, msig1, msig2, msig3, msig4, msig5};
    int N[3] = {nx, ny, nz};
    s->nx = nx; s->ny = ny; s->nz = nz;
    s->dx = (real)dx; s->dy = (real)dy; s->dz = (real)dz; s->dt = (real)dt;
    s->minv_s[0] = (real)mu_inv_s;
    s->minv_s[1] = (real)eps_inv_s;
    s->cutoff = cutoff;
    s->sources_active = 1;
    s->has_bound = (sb[0] != NULL);
    s->pdim[0] = nx + 2; s->pdim[1] = ny + 2; s->pdim[2] = (nz < 1 ? 1 : nz) + 2;
    for (int a = 0; a < 3; a++) {
        size_t n = 2 * (size_t)N[a] + 1;
        if (sb[a]) { s->sig[0][a] = zalloc(n); memcpy(s->sig[0][a], sb[a], n * sizeof(real)); }
        if (sd[a]) { s->sig[1][a] = zalloc(n); memcpy(s->sig[1][a], sd[a], n * sizeof(real)); }
    }
    for (int c = 0; c < NCOMP; c++) {
        comp_count(nx, ny, nz, c, s->gdim[c]);
        for (int a = 0; a < 3; a++) s->dim[c][a] = s->gdim[c][a] + 2;
        size_t n = vol(s->dim[c]);
        s->F[c] = zalloc(n);
        s->T[c] = zalloc(n);
        int e = c < 3;
        int own = c % 3, nxt = (own + 1) % 3, prv = (own + 2) % 3;
        real **sg = s->sig[e ? 1 : 0];
        /* src/Fields.jl:266-305 */
        if (sg[nxt]) s->U[c] = zalloc(n);
        if (sg[own]) s->W[c] = zalloc(n);
        if (ms[c]) {
            size_t g = vol(s->gdim[c]);
            s->msig[c] = zalloc(g);
            memcpy(s->msig[c], ms[c], g * sizeof(real));
            if (sg[nxt] || sg[prv]) s->C[c] = zalloc(n);
        }
        if (mi[c]) {
            size_t g = vol(s->gdim[c]);
            s->minv[c] = zalloc(g);
            memcpy(s->minv[c], mi[c], g * sizeof(real));
        }
    }
    return s;
}

void orc_destroy(orc_sim *s) {
    if (!s) return;
    for (int c = 0; c < NCOMP; c++) {
        free(s->F[c]); free(s->T[c]); free(s->C[c]); free(s->U[c]); free(s->W[c]);
        free(s->S[c]); free(s->msig[c]); free(s->minv[c]);
    }
    for (int e = 0; e < 2; e++)
        for (int a = 0; a < 3; a++) free(s->sig[e][a]);
    for (int i = 0; i < s->nsrc; i++) { free(s->src[i].amp); free(s->src[i].table); }
    for (int i = 0; i < s->ndft; i++) { free(s->dft[i].freqs); free(s->dft[i].acc); }
    for (int i = 0; i < s->npole; i++) {
        free(s->pole[i].sigma);
        for (int a = 0; a < 3; a++) { free(s->pole[i].P[a]); free(s->pole[i].Pp[a]); }
    }
    for (int a = 0; a < 3; a++) { free(s->fPD[a]); free(s->fPDi[a]); }
    for (int c = 0; c < NCOMP; c++) {
        free(s->Fi[c]); free(s->Ti[c]); free(s->Ci[c]); free(s->Ui[c]); free(s->Wi[c]); free(s->Si[c]);
    }
    for (int i = 0; i < s->npole; i++)
        for (int a = 0; a < 3; a++) { free(s->pole[i].Pi[a]); free(s->pole[i].Ppi[a]); }
    free(s->chi3); free(s->pole);
    free(s->src); free(s->dft); free(s);
}

/* kind 0: gaussian p = {fcen, width, peak_time, cutoff}; kind 1: cw p = {fcen};
 * kind 2: explicit table (orc_set_table). */
int orc_add_source(orc_sim *s, int comp, const int *start, const int *dims, const real *amp,
                   int kind, const double *p) {
    s->src = (orc_source *)realloc(s->src, (size_t)(s->nsrc + 1) * sizeof(orc_source));
    orc_source *o = &s->src[s->nsrc];
    memset(o, 0, sizeof(*o));
    o->comp = comp;
    for (int a = 0; a < 3; a++) { o->start[a] = start[a]; o->dims[a] = dims[a]; }
    size_t n = vol(dims);
    o->amp = zalloc(2 * n);
    memcpy(o->amp, amp, 2 * n * sizeof(real));
    o->kind = kind;
    if (p) for (int i = 0; i < 4; i++) o->p[i] = p[i];
    /* the reference allocates a full-size S array for every source component */
    if (!s->S[comp]) s->S[comp] = zalloc(vol(s->dim[comp]));
    if (s->cplx && !s->Si[comp]) s->Si[comp] = zalloc(vol(s->dim[comp]));
    return s->nsrc++;
}

int orc_set_table(orc_sim *s, int id, int64_t n, const double *values) {
    orc_source *o = &s->src[id];
    free(o->table);
    o->table = (double *)malloc((size_t)(2 * n) * sizeof(double));
    memcpy(o->table, values, (size_t)(2 * n) * sizeof(double));
    o->ntable = n;
    return 0;
}

int orc_add_dft(orc_sim *s, int comp, const int *start, const int *dims, int nf,
                const double *freqs, int decim, int acc_f64) {
    s->dft = (orc_dft *)realloc(s->dft, (size_t)(s->ndft + 1) * sizeof(orc_dft));
    orc_dft *d = &s->dft[s->ndft];
    memset(d, 0, sizeof(*d));
    d->comp = comp; d->nf = nf; d->decim = decim; d->acc_f64 = acc_f64;
    for (int a = 0; a < 3; a++) { d->start[a] = start[a]; d->dims[a] = dims[a]; }
    d->freqs = (double *)malloc((size_t)nf * sizeof(double));
    memcpy(d->freqs, freqs, (size_t)nf * sizeof(double));
    size_t n = vol(dims) * (size_t)nf * 2;
    d->acc = calloc(n, acc_f64 ? sizeof(double) : sizeof(real));
    return s->ndft++;
}

/* Complex fields and periodic axes; call right after orc_create.  fwd/rev
 * hold (re, im) of exp(+i k L) and exp(-i k L) per axis
 * (src/Chunking.jl:1745-1747). */
int orc_set_periodic(orc_sim *s, int cplx, const int *periodic, const double *fwd, const double *rev) {
    if (s->nsrc || s->npole) return 1;
    s->cplx = cplx;
    for (int a = 0; a < 3; a++) {
        s->periodic[a] = periodic[a];
        s->ph_fwd[a][0] = fwd[2 * a]; s->ph_fwd[a][1] = fwd[2 * a + 1];
        s->ph_rev[a][0] = rev[2 * a]; s->ph_rev[a][1] = rev[2 * a + 1];
    }
    if (cplx)
        for (int c = 0; c < NCOMP; c++) {
            size_t n = vol(s->dim[c]);
            s->Fi[c] = zalloc(n);
            s->Ti[c] = zalloc(n);
            if (s->C[c]) s->Ci[c] = zalloc(n);
            if (s->U[c]) s->Ui[c] = zalloc(n);
            if (s->W[c]) s->Wi[c] = zalloc(n);
        }
    return 0;
}

/* Replace a per-voxel inverse material array (geometry layout), as the
 * design update writes eps_inv (src/Adjoint/DesignRegion.jl:248-288). */
/* Switch to the default 27-chunk plan's arithmetic (quirk: components whose
 * chunk has no PML on their next axis get no U array, and the curl stage
 * then drops the previous axis's damping, src/Kernels/Helpers.jl:200-223).
 * pml_end[a]: last cell of the lower PML interval (0: none); pml_start[a]:
 * first cell of the upper one (N + 1: none).  3D with PML, no material
 * conductivity (its C arrays follow other rules).  Returns 0 or 1. */
int orc_set_chunk27(orc_sim *s, const int *pml_end, const int *pml_start) {
    if (!s->has_bound || s->nz < 1) return 1;
    for (int c = 0; c < NCOMP; c++)
        if (s->msig[c]) return 1;
    s->chunk27 = 1;
    for (int a = 0; a < 3; a++) {
        s->pml_end[a] = pml_end[a];
        s->pml_start[a] = pml_start[a];
    }
    return 0;
}

int orc_set_minv(orc_sim *s, int comp, const real *m) {
    if (!s->minv[comp]) return 1;
    memcpy(s->minv[comp], m, vol(s->gdim[comp]) * sizeof(real));
    return 0;
}

/* chi3 array (nx, ny, nz) painted on the cell centres (src/Geometry.jl:610-636) */
int orc_set_chi3(orc_sim *s, const real *chi3) {
    size_t n = (size_t)s->nx * s->ny * (s->nz < 1 ? 1 : s->nz);
    if (!s->chi3) s->chi3 = zalloc(n);
    memcpy(s->chi3, chi3, n * sizeof(real));
    return 0;
}

/* One Lorentz/Drude pole: c = {gamma1_inv, gamma1, omega0_dt_sq,
 * sigma_omega0_dt_sq, drude_coeff} as compute_ade_coefficients returns them
 * (src/Susceptibility.jl:74-85), sigma (nx, ny, nz).  The first pole also
 * allocates the total polarization fPD (src/Geometry.jl:1259-1267). */
int orc_add_pole(orc_sim *s, const double *c, int is_drude, const real *sigma) {
    s->pole = (orc_pole *)realloc(s->pole, (size_t)(s->npole + 1) * sizeof(orc_pole));
    orc_pole *q = &s->pole[s->npole];
    memset(q, 0, sizeof(*q));
    q->g1i = (real)c[0]; q->g1 = (real)c[1]; q->w = (real)c[2]; q->sw = (real)c[3]; q->dc = (real)c[4];
    q->drude = is_drude;
    size_t n = (size_t)s->nx * s->ny * (s->nz < 1 ? 1 : s->nz);
    q->sigma = zalloc(n);
    memcpy(q->sigma, sigma, n * sizeof(real));
    size_t np = vol(s->pdim);
    for (int a = 0; a < 3; a++) {
        q->P[a] = zalloc(np);
        q->Pp[a] = zalloc(np);
        if (!s->fPD[a]) s->fPD[a] = zalloc(np);
        if (s->cplx) {
            q->Pi[a] = zalloc(np);
            q->Ppi[a] = zalloc(np);
            if (!s->fPDi[a]) s->fPDi[a] = zalloc(np);
        }
    }
    return s->npole++;
}

/* ---------------------------------------------------------------- time */
static double round_time(const orc_sim *s) {
    /* Float64(timestep * dt): the product is formed in the backend precision */
    return (double)((real)s->timestep * s->dt);
}

static void waveform(const orc_sim *s, const orc_source *o, double t, real *re, real *im) {
    (void)s;
    if (o->kind == 0) {
        /* src/Sources/TimeSources.jl:123-132 */
        real fcen = (real)o->p[0], w = (real)o->p[1], peak = (real)o->p[2], cut = (real)o->p[3];
        real tt = (real)t - peak;
        if (tt > cut) { *re = 0; *im = 0; return; }
        real two_pi = (real)2 * (real)M_PI;
        real g = (real)exp((double)(((-tt) * tt) / (((real)2 * w) * w)));
        real b = ((-two_pi) * fcen) * tt;
        real c = (real)cos((double)b), sn = (real)sin((double)b);
        *re = g * c; *im = g * sn;
    } else if (o->kind == 1) {
        /* src/Sources/TimeSources.jl:61-64: exp(-i*2*pi*fcen*t) */
        real two_pi = (real)2 * (real)M_PI;
        real ph = (two_pi * (real)o->p[0]) * (real)t;
        *re = (real)cos((double)ph);
        *im = -(real)sin((double)ph);
    } else {
        int64_t n = s->timestep;
        if (n < o->ntable) { *re = (real)o->table[2 * n]; *im = (real)o->table[2 * n + 1]; }
        else { *re = 0; *im = 0; }
    }
}

static void step_sources(orc_sim *s, int magnetic, double t) {
    /* src/Sources/Sources.jl:330-357: S[off + i] += real(f(t) * A[i]) */
    for (int q = 0; q < s->nsrc; q++) {
        orc_source *o = &s->src[q];
        if ((o->comp >= 3) != magnetic) continue;
        real fr, fi;
        waveform(s, o, t, &fr, &fi);
        real *S = s->S[o->comp];
        const int *d = s->dim[o->comp];
        for (int iz = 1; iz <= o->dims[2]; iz++)
            for (int iy = 1; iy <= o->dims[1]; iy++)
                for (int ix = 1; ix <= o->dims[0]; ix++) {
                    size_t ai = (size_t)(ix - 1) + (size_t)o->dims[0] * ((size_t)(iy - 1) + (size_t)o->dims[1] * (size_t)(iz - 1));
                    real ar = o->amp[2 * ai], aim = o->amp[2 * ai + 1];
                    real v = (fr * ar) - (fi * aim);
                    S[rix(d, ix + o->start[0], iy + o->start[1], iz + o->start[2])] += v;
                }
    }
}

/* ------------------------------------------------------- curl + cascade */
/* update_field_from_curl (src/Kernels/Helpers.jl:30-33); has_old / has_sig
 * encode the Nothing dispatch. */
static inline real ufc(real A, real B, int has_old, real Bold, int has_sig, real sig) {
    if (has_sig) {
        if (has_old) return (((real)1 - sig) * A + B - Bold) / ((real)1 + sig);
        return (((real)1 - sig) * A + B) / ((real)1 + sig);
    }
    if (has_old) return A + B - Bold;
    return A + B;
}

static int g_error = 0;

/* generic_curl! (src/Kernels/Helpers.jl:39-270), method resolution spelled out.
 * Updates T (and C, U) in place; returns 1 when the reference has no method
 * for the combination, 0 otherwise. */
static inline int generic_curl(real K, real *C, real *U, real *T, size_t f,
                                int hsD, real sD, int hn, real sn, int hp, real sp) {
    if (C && U) {
        if (!hsD) { /* :72 */
            real Uo = U[f];
            U[f] = ufc(U[f], K, 0, 0, hn, sn);
            T[f] = ufc(T[f], U[f], 1, Uo, hp, sp);
            return 0;
        }
        /* :39 */
        if (sD == (real)0) {
            real Uo = U[f];
            U[f] = ufc(U[f], K, 0, 0, hn, sn);
            T[f] = ufc(T[f], U[f], 1, Uo, hp, sp);
        } else {
            real Co = C[f];
            C[f] = ufc(C[f], K, 0, 0, 1, sD);
            real Uo = U[f];
            U[f] = ufc(U[f], C[f], 1, Co, hn, sn);
            T[f] = ufc(T[f], U[f], 1, Uo, hp, sp);
        }
        return 0;
    }
    if (C && !U) {
        if (hn) return 1; /* no method (only the never-taken array sigma_next one) */
        if (!hsD) { T[f] = ufc(T[f], K, 0, 0, hp, sp); return 0; } /* :96 */
        if (sD == (real)0) { T[f] = ufc(T[f], K, 0, 0, hp, sp); return 0; } /* :80 */
        real Co = C[f];
        C[f] = ufc(C[f], K, 0, 0, 1, sD);
        T[f] = ufc(T[f], C[f], 1, Co, hp, sp);
        return 0;
    }
    if (!C && U) {
        if (hsD || !hn) return 1; /* no method */
        /* :113 */
        if (sn == (real)0) { T[f] = ufc(T[f], K, 0, 0, hp, sp); return 0; }
        real Uo = U[f];
        U[f] = ufc(U[f], K, 0, 0, 1, sn);
        T[f] = ufc(T[f], U[f], 1, Uo, hp, sp);
        return 0;
    }
    /* C and U both absent */
    if (!hsD) { T[f] = T[f] + K; return 0; } /* :188, :208, :242, :257 */
    if (!hn && !hp) return 1; /* :141 vs :225 ambiguous */
    T[f] = ufc(T[f], K, 0, 0, 1, sD); /* :225 */
    return 0;
}

/* sigma sample for cell i (1-based): sig[2i-1] (src/Kernels/Helpers.jl:277) */
static inline real sig_at(const real *sg, int i) { return sg[2 * i - 2]; }

static void half_step(orc_sim *s, int electric) {
    /* step_curl! then update_field! (src/Kernels/Kernels.jl:260-283) */
    int base = electric ? 0 : 3;     /* components being updated */
    int src = electric ? 3 : 0;      /* components differentiated */
    int idx = electric ? -1 : 1;     /* idx_curl */
    real **sg = s->sig[electric ? 1 : 0];
    real inv[3] = {(real)1 / s->dx, (real)1 / s->dy, (real)1 / s->dz};
    int N[3] = {s->nx, s->ny, s->nz};
    const real *A[3] = {s->F[src], s->F[src + 1], s->F[src + 2]};
    const int *ad[3] = {s->dim[src], s->dim[src + 1], s->dim[src + 2]};
    int err = 0;
#pragma omp parallel for schedule(static) reduction(| : err)
    for (int iz = 1; iz <= N[2]; iz++)
        for (int iy = 1; iy <= N[1]; iy++)
            for (int ix = 1; ix <= N[0]; ix++) {
                int fx = ix + 1, fy = iy + 1, fz = iz + 1;
                int gi[3] = {ix, iy, iz};
                for (int a = 0; a < 3; a++) {
                    int c = base + a;
                    int n1 = (a + 1) % 3, n2 = (a + 2) % 3; /* next, prev */
                    /* curl_a = d_n2(A_n1) - d_n1(A_n2)  (src/Kernels/Helpers.jl:293-298) */
                    int o1[3] = {0, 0, 0}, o2[3] = {0, 0, 0};
                    o1[n2] = idx;
                    o2[n1] = idx;
                    real t1 = inv[n2] * (A[n1][rix(ad[n1], fx + o1[0], fy + o1[1], fz + o1[2])] - A[n1][rix(ad[n1], fx, fy, fz)]);
                    real t2 = inv[n1] * (A[n2][rix(ad[n2], fx + o2[0], fy + o2[1], fz + o2[2])] - A[n2][rix(ad[n2], fx, fy, fz)]);
#ifdef ORC_PLANT_CURL_SIGN
                    /* planted mistake for the test suite: wrong sign of one term */
                    if (a == 2) t2 = -t2;
#endif
                    real K = s->dt * (t1 - t2);
                    size_t f = rix(s->dim[c], fx, fy, fz);
                    int hsD = s->msig[c] != NULL;
                    real sD = 0;
                    if (hsD) sD = (real)0.5 * (s->dt * s->msig[c][rix(s->gdim[c], ix, iy, iz)]);
#ifdef ORC_PLANT_COND_SCALE
                    /* planted mistake for the test suite: sigma~ without the 1/2 */
                    if (hsD) sD = s->dt * s->msig[c][rix(s->gdim[c], ix, iy, iz)];
#endif
                    int hn = sg[n1] != NULL, hp = sg[n2] != NULL;
                    real sn = hn ? sig_at(sg[n1], gi[n1]) : 0;
                    real sp = hp ? sig_at(sg[n2], gi[n2]) : 0;
                    real *Uc = s->U[c];
                    if (s->chunk27) {
                        /* a chunk with PML on some axis has profiles on every
                         * axis (zero-filled off its PML axes) and U only with
                         * PML on the next axis; a PML-free chunk has none */
                        int hpa[3];
                        for (int b = 0; b < 3; b++)
                            hpa[b] = gi[b] <= s->pml_end[b] || gi[b] >= s->pml_start[b];
                        const int pml_chunk = hpa[0] || hpa[1] || hpa[2];
                        hn = hp = pml_chunk;
                        sn = hpa[n1] ? sn : 0;
                        sp = hpa[n2] ? sp : 0;
                        if (!hpa[n1]) Uc = NULL;
                    }
                    err |= generic_curl(K, s->C[c], Uc, s->T[c], f, hsD, sD, hn, sn, hp, sp);
                }
            }
    if (err) g_error = 1;
#pragma omp parallel for schedule(static)
    for (int iz = 1; iz <= N[2]; iz++)
        for (int iy = 1; iy <= N[1]; iy++)
            for (int ix = 1; ix <= N[0]; ix++) {
                int gi[3] = {ix, iy, iz};
                for (int a = 0; a < 3; a++) {
                    int c = base + a;
                    size_t f = rix(s->dim[c], ix + 1, iy + 1, iz + 1);
                    real m = s->minv[c] ? s->minv[c][rix(s->gdim[c], ix, iy, iz)] : s->minv_s[electric];
                    /* update_field_generic (src/Kernels/Helpers.jl:323-366) */
                    real net = s->T[c][f];
                    real *S = s->sources_active ? s->S[c] : NULL;
                    if (S) net += S[f];
                    if (electric && s->fPD[a]) net -= s->fPD[a][rix(s->pdim, ix + 1, iy + 1, iz + 1)];
                    if (S) S[f] = 0;
                    int hs = sg[a] != NULL;
                    real so = hs ? sig_at(sg[a], gi[a]) : 0;
                    if (s->chunk27) {
                        /* W only with PML on the own axis; elsewhere the profile is zero-filled */
                        const int own = gi[a] <= s->pml_end[a] || gi[a] >= s->pml_start[a];
                        if (!own) so = 0;
                    }
                    if (s->W[c] && hs && so != (real)0) {
                        real Wo = s->W[c][f];
                        s->W[c][f] = m * net;
                        s->F[c][f] = s->F[c][f] + ((real)1 + so) * s->W[c][f] - ((real)1 - so) * Wo;
                    } else {
                        s->F[c][f] = m * net;
                    }
                }
            }
}

/* _chi3_correction_kernel! (src/Kernels/Dispersive.jl:137-157).  With complex
 * fields real(e * conj(e)) = re*re + im*im exactly, summed x, y, z in order. */
static void chi3_step(orc_sim *s) {
    int N[3] = {s->nx, s->ny, s->nz < 1 ? 1 : s->nz};
#pragma omp parallel for schedule(static)
    for (int iz = 1; iz <= N[2]; iz++)
        for (int iy = 1; iy <= N[1]; iy++)
            for (int ix = 1; ix <= N[0]; ix++) {
                real c3 = s->chi3[(size_t)(ix - 1) + (size_t)N[0] * ((size_t)(iy - 1) + (size_t)N[1] * (size_t)(iz - 1))];
                if (c3 == (real)0) continue;
                size_t f[3];
                real e[3], ei[3] = {0, 0, 0}, m[3];
                for (int a = 0; a < 3; a++) {
                    f[a] = rix(s->dim[a], ix + 1, iy + 1, iz + 1);
                    e[a] = s->F[a][f[a]];
                    if (s->cplx) ei[a] = s->Fi[a][f[a]];
                    m[a] = s->cplx ? e[a] * e[a] + ei[a] * ei[a] : e[a] * e[a];
                }
                real esq = m[0] + m[1] + m[2];
                real corr = (real)1 / ((real)1 + c3 * esq);
                for (int a = 0; a < 3; a++) {
                    s->F[a][f[a]] = e[a] * corr;
                    if (s->cplx) s->Fi[a][f[a]] = ei[a] * corr;
                }
            }
}

/* One ADE update of one part (real or imaginary) of one component. */
static inline void ade(const orc_pole *q, real sg, real e, real *P, real *Pp, size_t p) {
    real pc = P[p];
#ifdef ORC_PLANT_ADE_INDEX
    /* planted mistake for the test suite: P in place of P_prev */
    Pp[p] = pc;
#endif
    if (q->drude)
        P[p] = q->g1i * (((real)2 * pc - q->g1 * Pp[p]) + (q->dc * sg) * e);
    else
        P[p] = q->g1i * ((((real)2 - q->w) * pc - q->g1 * Pp[p]) + (q->sw * sg) * e);
    Pp[p] = pc;
}

/* step_polarization! (src/Kernels/Dispersive.jl:186-228): fPD <- 0, then per
 * pole the ADE update (:25-90) and fPD += P (:97-105). */
static void polarization_step(orc_sim *s) {
    int N[3] = {s->nx, s->ny, s->nz < 1 ? 1 : s->nz};
#pragma omp parallel for schedule(static)
    for (int iz = 1; iz <= N[2]; iz++)
        for (int iy = 1; iy <= N[1]; iy++)
            for (int ix = 1; ix <= N[0]; ix++) {
                size_t g = (size_t)(ix - 1) + (size_t)N[0] * ((size_t)(iy - 1) + (size_t)N[1] * (size_t)(iz - 1));
                size_t p = rix(s->pdim, ix + 1, iy + 1, iz + 1);
                for (int a = 0; a < 3; a++) {
                    s->fPD[a][p] = (real)0;
                    if (s->cplx) s->fPDi[a][p] = (real)0;
                }
                for (int k = 0; k < s->npole; k++) {
                    orc_pole *q = &s->pole[k];
                    real sg = q->sigma[g];
                    if (sg != (real)0) {
                        for (int a = 0; a < 3; a++) {
                            size_t f = rix(s->dim[a], ix + 1, iy + 1, iz + 1);
                            ade(q, sg, s->F[a][f], q->P[a], q->Pp[a], p);
                            if (s->cplx) ade(q, sg, s->Fi[a][f], q->Pi[a], q->Ppi[a], p);
                        }
                    }
                    for (int a = 0; a < 3; a++) {
                        s->fPD[a][p] += q->P[a][p];
                        if (s->cplx) s->fPDi[a][p] += q->Pi[a][p];
                    }
                }
            }
}

/* Exchange the real and imaginary array sets (the update is part-wise). */
static void swap_parts(orc_sim *s) {
    real *t;
#define ORC_SWAP(A, B) do { t = (A); (A) = (B); (B) = t; } while (0)
    for (int c = 0; c < NCOMP; c++) {
        ORC_SWAP(s->F[c], s->Fi[c]);
        ORC_SWAP(s->T[c], s->Ti[c]);
        ORC_SWAP(s->C[c], s->Ci[c]);
        ORC_SWAP(s->U[c], s->Ui[c]);
        ORC_SWAP(s->W[c], s->Wi[c]);
        ORC_SWAP(s->S[c], s->Si[c]);
    }
    for (int a = 0; a < 3; a++) ORC_SWAP(s->fPD[a], s->fPDi[a]);
#undef ORC_SWAP
}

/* Periodic / Bloch wrap of one field group after its half step
 * (_exchange_halos_local!, src/Chunking.jl:2138-2172, with the single-chunk
 * connections of :1755-1768): ghost 0 <- plane N * exp(-ikL), ghost N+1 <-
 * plane 1 * exp(+ikL), transverse indices 1..N.  The phase is applied in
 * complex float64 and rounded to the field type; a unit phase is a plain copy. */
static void wrap_group(orc_sim *s, int electric) {
    int N[3] = {s->nx, s->ny, s->nz < 1 ? 1 : s->nz};
    int base = electric ? 0 : 3;
    for (int ax = 0; ax < 3; ax++) {
        if (!s->periodic[ax]) continue;
        int t1 = (ax + 1) % 3, t2 = (ax + 2) % 3;
        for (int side = 0; side < 2; side++) {
            /* side 0: upper interior -> lower ghost (phase_rev); side 1: lower -> upper (phase_fwd) */
            int src_g = side == 0 ? N[ax] : 1;
            int dst_g = side == 0 ? 0 : N[ax] + 1;
            const double *ph = side == 0 ? s->ph_rev[ax] : s->ph_fwd[ax];
            int unit = ph[0] == 1.0 && ph[1] == 0.0;
            for (int c = base; c < base + 3; c++) {
                const int *d = s->dim[c];
                for (int j = 1; j <= N[t2]; j++)
                    for (int i = 1; i <= N[t1]; i++) {
                        int gs[3], gd[3];
                        gs[ax] = src_g; gd[ax] = dst_g;
                        gs[t1] = gd[t1] = i;
                        gs[t2] = gd[t2] = j;
                        size_t fs = rix(d, gs[0] + 1, gs[1] + 1, gs[2] + 1);
                        size_t fd = rix(d, gd[0] + 1, gd[1] + 1, gd[2] + 1);
                        if (!s->cplx) {
                            s->F[c][fd] = s->F[c][fs];
                        } else if (unit) {
                            s->F[c][fd] = s->F[c][fs];
                            s->Fi[c][fd] = s->Fi[c][fs];
                        } else {
                            double re = (double)s->F[c][fs], im = (double)s->Fi[c][fs];
                            s->F[c][fd] = (real)(re * ph[0] - im * ph[1]);
                            s->Fi[c][fd] = (real)(re * ph[1] + im * ph[0]);
                        }
                    }
            }
        }
    }
}

static void update_dfts(orc_sim *s, int magnetic, double t) {
    /* src/Kernels/Kernels.jl:465-497 gate, src/Monitors/Monitors.jl:361-379 */
    for (int q = 0; q < s->ndft; q++) {
        orc_dft *d = &s->dft[q];
        /* comp 6..11: D / B (the T arrays) at the positions of E / H */
        int base = d->comp % 6;
        if ((base >= 3) != magnetic) continue;
        if (d->decim > 1 && (s->timestep % d->decim) != 0) continue;
        /* time_fac = complex_backend_number(im * 2 * pi * time) */
        double theta = (double)(real)(6.283185307179586 * t);
        double dts = (double)s->dt;
        const real *F = d->comp >= 6 ? s->T[base] : s->F[base];
        const real *Fim = s->cplx ? (d->comp >= 6 ? s->Ti[base] : s->Fi[base]) : NULL;
        const int *fd = s->dim[base];
        size_t ncell = vol(d->dims);
        for (int k = 0; k < d->nf; k++) {
            double arg = d->freqs[k] * theta;
            double wr = dts * cos(arg), wi = dts * sin(arg);
#pragma omp parallel for schedule(static)
            for (int iz = 1; iz <= d->dims[2]; iz++)
                for (int iy = 1; iy <= d->dims[1]; iy++)
                    for (int ix = 1; ix <= d->dims[0]; ix++) {
                        size_t ci = (size_t)(ix - 1) + (size_t)d->dims[0] * ((size_t)(iy - 1) + (size_t)d->dims[1] * (size_t)(iz - 1));
                        size_t ai = ci + ncell * (size_t)k;
                        size_t fi = rix(fd, ix + d->start[0], iy + d->start[1], iz + d->start[2]);
                        double v = (double)F[fi];
                        /* (w_r + i w_i)(v + i u): complex fields add the u terms */
                        double re = wr * v, im = wi * v;
                        if (s->cplx) {
                            double u = (double)Fim[fi];
                            re = wr * v - wi * u;
                            im = wr * u + wi * v;
                        }
                        if (d->acc_f64) {
                            double *acc = (double *)d->acc;
                            acc[2 * ai] = acc[2 * ai] + re;
                            acc[2 * ai + 1] = acc[2 * ai + 1] + im;
                        } else {
                            real *acc = (real *)d->acc;
                            acc[2 * ai] = (real)((double)acc[2 * ai] + re);
                            acc[2 * ai + 1] = (real)((double)acc[2 * ai + 1] + im);
                        }
                    }
        }
    }
}

int orc_step(orc_sim *s, int64_t n) {
    g_error = 0;
    for (int64_t it = 0; it < n; it++) {
        double t = round_time(s);
        if (s->sources_active && t > s->cutoff) s->sources_active = 0;
        double th = t + (double)(s->dt / (real)2);
        if (s->sources_active) step_sources(s, 1, t);
        half_step(s, 0);
        if (s->cplx) { swap_parts(s); half_step(s, 0); swap_parts(s); }
        wrap_group(s, 0);
        update_dfts(s, 1, t);
        if (s->sources_active) step_sources(s, 0, th);
        half_step(s, 1);
        if (s->cplx) { swap_parts(s); half_step(s, 1); swap_parts(s); }
        /* the E halo exchange closes step_E_fused!, before chi3 and P (src/Kernels/Kernels.jl:76-83, :452) */
        wrap_group(s, 1);
        if (s->chi3) chi3_step(s);
        if (s->npole) polarization_step(s);
        update_dfts(s, 0, th);
        s->timestep++;
        if (g_error) return -1;
    }
    return 0;
}

int64_t orc_timestep(const orc_sim *s) { return s->timestep; }

/* Copy a full raw array (reference layout, ghosts included). which: 0 field,
 * 1 T (B/D), 2 U, 3 W, 4 C, 5 S; 6..11 the imaginary parts of the same. */
int orc_read(const orc_sim *s, int which, int comp, real *out) {
    real *src[12] = {s->F[comp], s->T[comp], s->U[comp], s->W[comp], s->C[comp], s->S[comp],
                     s->Fi[comp], s->Ti[comp], s->Ui[comp], s->Wi[comp], s->Ci[comp], s->Si[comp]};
    size_t n = vol(s->dim[comp]);
    if (!src[which]) { memset(out, 0, n * sizeof(real)); return 1; }
    memcpy(out, src[which], n * sizeof(real));
    return 0;
}

int orc_write(orc_sim *s, int which, int comp, const real *in) {
    real *dst[12] = {s->F[comp], s->T[comp], s->U[comp], s->W[comp], s->C[comp], s->S[comp],
                     s->Fi[comp], s->Ti[comp], s->Ui[comp], s->Wi[comp], s->Ci[comp], s->Si[comp]};
    if (!dst[which]) return 1;
    memcpy(dst[which], in, vol(s->dim[comp]) * sizeof(real));
    return 0;
}

/* Polarization arrays of a pole, (nx+2, ny+2, nz+2): which 0 P, 1 P_prev;
 * pole -1 reads the total fPD. */
int orc_read_pole(const orc_sim *s, int pole, int which, int a, real *out) {
    /* which 2 / 3: imaginary parts of P / P_prev (fPD with pole -1) */
    const real *src;
    if (pole < 0) src = which >= 2 ? s->fPDi[a] : s->fPD[a];
    else {
        const orc_pole *q = &s->pole[pole];
        const real *opts[4] = {q->P[a], q->Pp[a], q->Pi[a], q->Ppi[a]};
        src = opts[which];
    }
    if (!src) return 1;
    memcpy(out, src, vol(s->pdim) * sizeof(real));
    return 0;
}

int orc_write_pole(orc_sim *s, int pole, int which, int a, const real *in) {
    real *dst;
    if (pole < 0) dst = which >= 2 ? s->fPDi[a] : s->fPD[a];
    else {
        orc_pole *q = &s->pole[pole];
        real *opts[4] = {q->P[a], q->Pp[a], q->Pi[a], q->Ppi[a]};
        dst = opts[which];
    }
    if (!dst) return 1;
    memcpy(dst, in, vol(s->pdim) * sizeof(real));
    return 0;
}

void orc_dims(const orc_sim *s, int comp, int *out) {
    for (int a = 0; a < 3; a++) out[a] = s->dim[comp][a];
}

/* DFT accumulator as complex128 (dims..., nf) */
int orc_read_dft(const orc_sim *s, int id, double *out) {
    const orc_dft *d = &s->dft[id];
    size_t n = vol(d->dims) * (size_t)d->nf * 2;
    for (size_t i = 0; i < n; i++)
        out[i] = d->acc_f64 ? ((double *)d->acc)[i] : (double)((real *)d->acc)[i];
    return 0;
}

int orc_real_size(void) { return (int)sizeof(real); }

/* ------------------------------------------------------------------------
 * Near-to-far projection (double precision whatever REAL is).
 *
 *   green3d! ..................... src/Monitors/Near2Far.jl:40-94
 *   equivalent currents per axis . src/Monitors/Near2Far.jl:147-150, :192-195,
 *                                  :238-241 (= CPU path :334-357)
 *   plane slicing, normal average  src/Monitors/Near2Far.jl:409-474
 *
 * Complex numbers are (re, im) pairs combined with Julia's rules, in its
 * order: z*w = (a c - b d, a d + b c); real*z and z/real act on both parts;
 * real/z = real*inv(z), inv((0, y)) = (0, -1/y).  The unit direction p and
 * r_hat x p are formed in full, as in the reference.
 * ------------------------------------------------------------------------ */
typedef struct { double re, im; } orc_cx;

static inline orc_cx cx(double a, double b) { orc_cx z = {a, b}; return z; }
static inline orc_cx cmul(orc_cx a, orc_cx b) { return cx(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re); }
static inline orc_cx cscale(orc_cx a, double s) { return cx(a.re * s, a.im * s); }
static inline orc_cx cdivr(orc_cx a, double s) { return cx(a.re / s, a.im / s); }
static inline orc_cx cadd(orc_cx a, orc_cx b) { return cx(a.re + b.re, a.im + b.im); }

#ifdef ORC_PLANT_GREEN_TERM2
#define ORC_T2_COEF 2.0 /* planted: 2/(ikr) where the Green's function has 3/(ikr) */
#else
#define ORC_T2_COEF 3.0
#endif

void orc_green3d(double *EH, const double *x, double freq, double eps, double mu, const double *x0, int c0,
                 double f0re, double f0im) {
    const double rv[3] = {x[0] - x0[0], x[1] - x0[1], x[2] - x0[2]};
    const double r = sqrt((rv[0] * rv[0] + rv[1] * rv[1]) + rv[2] * rv[2]);
    if (!(r >= 1e-20)) return;
    const double rh[3] = {rv[0] / r, rv[1] / r, rv[2] / r};
    const double n = sqrt(eps * mu);
    const double k = 6.283185307179586 * freq * n; /* 2 pi */
    const double Z = sqrt(mu / eps);
    const double kr = k * r;
    const orc_cx inv_ikr = cx(0.0, -(1.0 / kr)); /* 1/(i k r) */
    const double ikr2 = -(kr * kr);
    const double ph = kr + 1.5707963267948966;    /* k r + pi/2 */
    const orc_cx expfac = cmul(cscale(cx(f0re, f0im), k * n / (12.566370614359172 * r)), cx(cos(ph), sin(ph)));
    const int pc = (c0 - 1) % 3;
    const double p[3] = {pc == 0 ? 1.0 : 0.0, pc == 1 ? 1.0 : 0.0, pc == 2 ? 1.0 : 0.0};
    const double pd = p[0] * rh[0] + p[1] * rh[1] + p[2] * rh[2];
    const double rxp[3] = {rh[1] * p[2] - rh[2] * p[1], rh[2] * p[0] - rh[0] * p[2], rh[0] * p[1] - rh[1] * p[0]};
    /* term1 = 1 - 1/ikr + 1/ikr2 */
    const orc_cx term1 = cx((1.0 - inv_ikr.re) + 1.0 / ikr2, -inv_ikr.im);
    /* term2 = (-1 + 3/ikr - 3/ikr2) * (p . r_hat) */
    const orc_cx t3ikr = cscale(inv_ikr, ORC_T2_COEF);
    const orc_cx term2 = cscale(cx((-1.0 + t3ikr.re) - 3.0 / ikr2, t3ikr.im), pd);
    /* term3 = 1 - 1/ikr */
    const orc_cx term3 = cx(1.0 - inv_ikr.re, -inv_ikr.im);
    orc_cx add[6];
    if (c0 <= 3) {
        const orc_cx ef = cdivr(expfac, eps);
        for (int j = 0; j < 3; ++j) add[j] = cmul(ef, cadd(cscale(term1, p[j]), cscale(term2, rh[j])));
        for (int j = 0; j < 3; ++j) add[3 + j] = cdivr(cscale(cmul(ef, term3), rxp[j]), Z);
    } else {
        const orc_cx ef = cdivr(expfac, mu);
        const orc_cx mef = cx(-ef.re, -ef.im);
        for (int j = 0; j < 3; ++j) add[j] = cscale(cscale(cmul(mef, term3), rxp[j]), Z);
        for (int j = 0; j < 3; ++j) add[3 + j] = cmul(ef, cadd(cscale(term1, p[j]), cscale(term2, rh[j])));
    }
    for (int j = 0; j < 6; ++j) {
        EH[2 * j] += add[j].re;
        EH[2 * j + 1] += add[j].im;
    }
}

/* Value of one tangential plane at (i1, i2) for frequency kf: the sample at
 * normal index 0, averaged with index 1 when the DFT spans two planes. */
static orc_cx n2f_plane(const double *a, const int *d, int ax, int t0, int t1, int i1, int i2, int kf) {
    int ii[3];
    ii[ax] = 0;
    ii[t0] = i1;
    ii[t1] = i2;
    const size_t nc = (size_t)d[0] * d[1] * d[2];
    const size_t lin = (size_t)ii[0] + (size_t)d[0] * ((size_t)ii[1] + (size_t)d[1] * ii[2]);
    const double *p0 = a + 2 * (lin + nc * kf);
    orc_cx v = cx(p0[0], p0[1]);
    if (d[ax] >= 2) {
        const size_t st = ax == 0 ? 1 : (ax == 1 ? (size_t)d[0] : (size_t)d[0] * d[1]);
        v = cx((v.re + p0[2 * st]) / 2, (v.im + p0[2 * st + 1]) / 2);
    }
    return v;
}

/* Far field out[((k * 6 + j) * n_obs + o)] (re, im) of one planar surface:
 * planes e1, e2, h1, h2 complex (d0, d1, d2, nf) x fastest, dims[4][3];
 * base[4][3] the position of sample (0, 0) of each; i1 runs along the lower
 * tangential axis.  Samples are summed in the reference's order (i1
 * fastest), one (observation point, frequency) per OpenMP iteration. */
int orc_near2far(const double *e1, const double *e2, const double *h1, const double *h2, const int *dims, int nf,
                 int normal_axis, double ns, const double *base, double d1, double d2, double dA, int n1, int n2,
                 const double *freqs, double eps, double mu, const double *obs, int n_obs, double *out) {
    if (normal_axis < 1 || normal_axis > 3 || n1 < 1 || n2 < 1 || nf < 1 || n_obs < 1) return 1;
    const int ax = normal_axis - 1;
    const int t0 = ax == 0 ? 1 : 0, t1 = ax == 2 ? 1 : 2;
    const double *pl[4] = {e1, e2, h1, h2};
    for (int q = 0; q < 4; ++q)
        if (n1 > dims[3 * q + t0] || n2 > dims[3 * q + t1] || dims[3 * q + ax] > 2) return 2;
    /* (plane, source component 1..6, sign) in the reference's call order */
    int qs[4], cs[4];
    double sg[4];
#ifdef ORC_PLANT_N2F_MSIGN
    const double msg = -1.0; /* planted: magnetic currents with the opposite sign */
#else
    const double msg = 1.0;
#endif
    if (ax == 2) {        /* Jx = ns Hy, Jy = -ns Hx, Mx = -ns Ey, My = ns Ex */
        qs[0] = 3; cs[0] = 1; sg[0] = ns;
        qs[1] = 2; cs[1] = 2; sg[1] = -ns;
        qs[2] = 1; cs[2] = 4; sg[2] = -ns * msg;
        qs[3] = 0; cs[3] = 5; sg[3] = ns * msg;
    } else if (ax == 0) { /* Jy = ns Hz, Jz = -ns Hy, My = -ns Ez, Mz = ns Ey */
        qs[0] = 3; cs[0] = 2; sg[0] = ns;
        qs[1] = 2; cs[1] = 3; sg[1] = -ns;
        qs[2] = 1; cs[2] = 5; sg[2] = -ns * msg;
        qs[3] = 0; cs[3] = 6; sg[3] = ns * msg;
    } else {              /* Jz = ns Hx, Jx = -ns Hz, Mz = -ns Ex, Mx = ns Ez */
        qs[0] = 2; cs[0] = 3; sg[0] = ns;
        qs[1] = 3; cs[1] = 1; sg[1] = -ns;
        qs[2] = 0; cs[2] = 6; sg[2] = -ns * msg;
        qs[3] = 1; cs[3] = 4; sg[3] = ns * msg;
    }
    const int64_t work = (int64_t)nf * n_obs;
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t w = 0; w < work; ++w) {
        const int kf = (int)(w / n_obs), o = (int)(w % n_obs);
        double EH[12] = {0};
        for (int i2 = 0; i2 < n2; ++i2)
            for (int i1 = 0; i1 < n1; ++i1) {
                for (int c = 0; c < 4; ++c) {
                    const int q = qs[c];
                    const orc_cx v = n2f_plane(pl[q], dims + 3 * q, ax, t0, t1, i1, i2, kf);
                    double x0[3];
                    x0[ax] = base[3 * q + ax];
                    x0[t0] = base[3 * q + t0] + (double)i1 * d1;
                    x0[t1] = base[3 * q + t1] + (double)i2 * d2;
                    const orc_cx f0 = cscale(cscale(v, sg[c]), dA);
                    orc_green3d(EH, obs + 3 * o, freqs[kf], eps, mu, x0, cs[c], f0.re, f0.im);
                }
            }
        for (int j = 0; j < 6; ++j) {
            out[2 * (((int64_t)kf * 6 + j) * n_obs + o)] = EH[2 * j];
            out[2 * (((int64_t)kf * 6 + j) * n_obs + o) + 1] = EH[2 * j + 1];
        }
    }
    return 0;
}
