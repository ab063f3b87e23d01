/*
 * casim_oracle.c -- CPU restatement of CaSim's dyad hot path (the oracle).
 *
 * TEST INFRASTRUCTURE ONLY: the checker for libcasim_dyad.so and the CPU
 * baseline timed by bench.py.  The product never links or calls this file.
 *
 * Built with -ffp-contract=off so every FP64 operation rounds separately, in
 * the order the reference's C++ writes it (the reference builds with MSVC on
 * x86-64, which does not contract a*b+c into an FMA by default).
 *
 * Parity pins (tests/): the RNG and the kinetic rates are compared with the
 * reference's own rngstream.cpp / ParallelRng.cpp / SheepCannell.cpp /
 * TwoState*.cpp compiled into oracle/_ref (and with an exact-integer Python
 * MRG32k3a); the reaction-diffusion restatement is pinned by closed forms
 * (exact Laplacian of quadratics, the scalar reaction recurrence, equilibrium
 * fixed point, Fourier-mode decay, mass balance of the injected flux, efflux
 * of linear profiles).  DyadRDCPU.cpp itself cannot be compiled here (it needs
 * nlohmann/json >= 3.6, absent from the image).
 */
#include "casim_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------- */
/* MRG32k3a, restated from rngstream.cpp (double arithmetic, exact integers). */

static const double m1 = 4294967087.0;
static const double m2 = 4294944443.0;
static const double norm_ = 1.0 / (4294967087.0 + 1.0);
static const double a12 = 1403580.0;
static const double a13n = 810728.0;
static const double a21 = 527612.0;
static const double a23n = 1370589.0;
static const double two17 = 131072.0;
static const double two53 = 9007199254740992.0;
static const double fact = 5.9604644775390625e-8; /* 2^-24 */

/* (a*s + c) mod m for a, s, c < 2^35 (rngstream.cpp:88-105). */
static double mult_mod_m(double a, double s, double c, double m) {
    double v = a * s + c;
    long a1;
    if (v >= two53 || v <= -two53) {
        a1 = (long)(a / two17);
        a -= a1 * two17;
        v = a1 * s;
        a1 = (long)(v / m);
        v -= a1 * m;
        v = v * two17 + a * s + c;
    }
    a1 = (long)(v / m);
    if ((v -= a1 * m) < 0.0) return v += m;
    return v;
}

/* v = A s mod m (rngstream.cpp:112-125). */
static void mat_vec_mod_m(const double A[9], const double s[3], double v[3], double m) {
    double x[3];
    for (int i = 0; i < 3; ++i) {
        x[i] = mult_mod_m(A[3 * i + 0], s[0], 0.0, m);
        x[i] = mult_mod_m(A[3 * i + 1], s[1], x[i], m);
        x[i] = mult_mod_m(A[3 * i + 2], s[2], x[i], m);
    }
    for (int i = 0; i < 3; ++i) v[i] = x[i];
}

/* C = A B mod m (rngstream.cpp:132-148). */
static void mat_mat_mod_m(const double A[9], const double B[9], double C[9], double m) {
    double V[3], W[9];
    for (int i = 0; i < 3; ++i) {
        for (int j = 0; j < 3; ++j) V[j] = B[3 * j + i];
        mat_vec_mod_m(A, V, V, m);
        for (int j = 0; j < 3; ++j) W[3 * j + i] = V[j];
    }
    memcpy(C, W, sizeof W);
}

static const double A1p0[9] = {0.0, 1.0, 0.0, 0.0, 0.0, 1.0, -810728.0, 1403580.0, 0.0};
static const double A2p0[9] = {0.0, 1.0, 0.0, 0.0, 0.0, 1.0, -1370589.0, 0.0, 527612.0};

/* B = A^(2^e) mod m by e squarings (MatTwoPowModM, rngstream.cpp:154-167).
 * The first squaring reduces the negative entries of A1p0/A2p0. */
static void mat_two_pow_mod_m(const double A[9], double B[9], double m, int e) {
    memcpy(B, A, 9 * sizeof(double));
    for (int i = 0; i < e; ++i) mat_mat_mod_m(B, B, B, m);
}

static double jump127_1[9], jump127_2[9], jump76_1[9], jump76_2[9];
static int jumps_ready = 0;

static void ensure_jumps(void) {
    if (jumps_ready) return;
#pragma omp critical(oracle_jumps)
    {
        if (!jumps_ready) {
            mat_two_pow_mod_m(A1p0, jump127_1, m1, 127);
            mat_two_pow_mod_m(A2p0, jump127_2, m2, 127);
            mat_two_pow_mod_m(A1p0, jump76_1, m1, 76);
            mat_two_pow_mod_m(A2p0, jump76_2, m2, 76);
            jumps_ready = 1;
        }
    }
}

void oracle_rng_jump_matrix(int which, int e, double out[9]) {
    mat_two_pow_mod_m(which == 1 ? A1p0 : A2p0, out, which == 1 ? m1 : m2, e);
}

/* CheckSeed (rngstream.cpp:200-234): 0 if legal. */
int oracle_rng_check_seed(const uint64_t seed[6]) {
    for (int i = 0; i < 3; ++i)
        if ((double)seed[i] >= m1 || seed[i] >= 4294967087ULL) return -1;
    for (int i = 3; i < 6; ++i)
        if ((double)seed[i] >= m2 || seed[i] >= 4294944443ULL) return -1;
    if (seed[0] == 0 && seed[1] == 0 && seed[2] == 0) return -1;
    if (seed[3] == 0 && seed[4] == 0 && seed[5] == 0) return -1;
    return 0;
}

/* States of streams first..first+n-1 after SetPackageSeed(seed): each
 * RngStream constructor takes nextSeed, then advances it by A^(2^127)
 * (rngstream.cpp:303-321). */
void oracle_rng_stream_states(const uint64_t seed[6], int first, int n, uint64_t *out) {
    ensure_jumps();
    double next[6];
    for (int i = 0; i < 6; ++i) next[i] = (double)seed[i];
    for (int k = 0; k < first + n; ++k) {
        if (k >= first)
            for (int i = 0; i < 6; ++i) out[6 * (k - first) + i] = (uint64_t)next[i];
        mat_vec_mod_m(jump127_1, next, next, m1);
        mat_vec_mod_m(jump127_2, next + 3, next + 3, m2);
    }
}

/* ResetNextSubstream from the stream start (rngstream.cpp:347-353). */
void oracle_rng_substream(uint64_t state[6]) {
    ensure_jumps();
    double s[6];
    for (int i = 0; i < 6; ++i) s[i] = (double)state[i];
    mat_vec_mod_m(jump76_1, s, s, m1);
    mat_vec_mod_m(jump76_2, s + 3, s + 3, m2);
    for (int i = 0; i < 6; ++i) state[i] = (uint64_t)s[i];
}

/* U01 (rngstream.cpp:242-265), anti = false. */
static double u01(double *Cg) {
    long k;
    double p1, p2;
    p1 = a12 * Cg[1] - a13n * Cg[0];
    k = (long)(p1 / m1);
    p1 -= k * m1;
    if (p1 < 0.0) p1 += m1;
    Cg[0] = Cg[1];
    Cg[1] = Cg[2];
    Cg[2] = p1;
    p2 = a21 * Cg[5] - a23n * Cg[3];
    k = (long)(p2 / m2);
    p2 -= k * m2;
    if (p2 < 0.0) p2 += m2;
    Cg[3] = Cg[4];
    Cg[4] = Cg[5];
    Cg[5] = p2;
    return (p1 > p2) ? (p1 - p2) * norm_ : (p1 - p2 + m1) * norm_;
}

/* RandU01 with incPrec = true (rngstream.cpp:306, :485-491) = U01d (:271-283). */
static double rand_u01(double *Cg) {
    double u = u01(Cg);
    u += u01(Cg) * fact;
    return (u < 1.0) ? u : (u - 1.0);
}

void oracle_rng_draws(uint64_t state[6], int n, double *out) {
    double Cg[6];
    for (int i = 0; i < 6; ++i) Cg[i] = (double)state[i];
    for (int i = 0; i < n; ++i) out[i] = rand_u01(Cg);
    for (int i = 0; i < 6; ++i) state[i] = (uint64_t)Cg[i];
}

/* ------------------------------------------------------------------------- */
/* Kinetic models. */

/* SheepCannell::_KOpen / _KClose (SheepCannell.cpp:12-18) through
 * TwoStateModel::GetRate (TwoStateModel.cpp:4-8); TwoStatePoModel::GetRate
 * (TwoStatePoModel.cpp:3-8) with TwoStateMWC::Po (TwoStateMWC.cpp:15-23) and
 * TwoStateaHTG::Po (TwoStateaHTG.cpp:18-29). */
double oracle_channel_rate(int model, const double *p, int state, double ca, double mg) {
    if (model == CASIM_MODEL_SHEEP_CANNELL) {
        double kopen0 = p[0], ncaopen = p[1], kopenmax = p[2], kclose0 = p[3], ncaclose = p[4];
        if (state) return kclose0 * pow(ca * 1e-3, ncaclose) * 1e-3;
        double v = kopen0 * pow(ca * 1e-3, ncaopen);
        return ((kopenmax < v) ? kopenmax : v) * 1e-3; /* std::min(v, kopenmax) */
    }
    if (model == CASIM_MODEL_TWO_STATE_MWC) {
        double fCa = p[0], KCa = p[1], fMg = p[2], KMg = p[3], KO0 = p[4], KMgI = p[5], kclose = p[6];
        if (state) return kclose;
        double inac = KMgI * KMgI / (KMgI * KMgI + mg * mg);
        double open = pow(ca + KCa * fCa * (1 + mg / (fMg * KMg)), 4);
        double close = KO0 * pow(fCa * (ca + KCa * (1 + mg / KMg)), 4);
        double P = inac * open / (open + close);
        return kclose * P / (1 - P);
    }
    if (model == CASIM_MODEL_TWO_STATE_AHTG) {
        double fCa = p[0], KCa = p[1], fMg = p[2], KMg = p[3], KO0 = p[4], KMgI = p[5];
        double KCL = p[6], KI = p[7], KOL = p[8], kclose = p[9];
        if (state) return kclose;
        double mult = KMgI * KMgI / (KMgI * KMgI + mg * mg);
        double up = (KCL * KI * (pow(ca, 4) * pow(fMg, 4) * pow(fCa, 4) * pow(KMg, 4) * KO0 +
                                 pow(ca * fMg * KMg + fCa * (mg + fMg * KMg) * KCa, 4) * KOL));
        double down =
            (4 * pow(ca, 3) * pow(fMg, 3) * fCa * pow(KMg, 3) * KCa * KCL * KI *
                 (mg + fMg * KMg + fMg * pow(fCa, 3) * (mg + KMg) * KO0) * KOL +
             6 * pow(ca, 2) * pow(fMg, 2) * pow(fCa, 2) * pow(KMg, 2) * pow(KCa, 2) * KCL * KI *
                 (pow(mg + fMg * KMg, 2) + pow(fMg, 2) * pow(fCa, 2) * pow(mg + KMg, 2) * KO0) * KOL +
             4 * ca * fMg * pow(fCa, 3) * KMg * pow(KCa, 3) * KCL * KI *
                 (pow(mg + fMg * KMg, 3) + pow(fMg, 3) * fCa * pow(mg + KMg, 3) * KO0) * KOL +
             pow(fCa, 4) * pow(KCa, 4) * KCL * KI *
                 (pow(mg + fMg * KMg, 4) + pow(fMg, 4) * pow(mg + KMg, 4) * KO0) * KOL +
             pow(ca, 4) * pow(fMg, 4) * pow(KMg, 4) *
                 (KCL * KI * KOL + pow(fCa, 4) * KO0 * (1 + KI + KCL * KI + KCL * KI * KOL)));
        double P = mult * up / down;
        return kclose * P / (1 - P);
    }
    return 0.0;
}

/* ------------------------------------------------------------------------- */
/* GaussElimination (DyadRDCPU.cpp:163-193) on an n x (n+1) system. */
static void gauss_elimination(int n, double *a /* n*(n+1) row-major */, double *ans) {
    int m = n;
    int where[16];
    for (int i = 0; i < m; ++i) where[i] = -1;
    for (int col = 0, row = 0; col < m && row < n; ++col) {
        int sel = row;
        for (int i = row; i < n; ++i)
            if (fabs(a[i * (m + 1) + col]) > fabs(a[sel * (m + 1) + col])) sel = i;
        if (fabs(a[sel * (m + 1) + col]) < 1e-10) continue;
        for (int i = col; i <= m; ++i) {
            double t = a[sel * (m + 1) + i];
            a[sel * (m + 1) + i] = a[row * (m + 1) + i];
            a[row * (m + 1) + i] = t;
        }
        where[col] = row;
        for (int i = 0; i < n; ++i)
            if (i != row) {
                double c = a[i * (m + 1) + col] / a[row * (m + 1) + col];
                for (int j = col; j <= m; ++j) a[i * (m + 1) + j] -= a[row * (m + 1) + j] * c;
            }
        ++row;
    }
    for (int i = 0; i < m; ++i) ans[i] = (where[i] != -1) ? a[where[i] * (m + 1) + m] / a[where[i] * (m + 1) + i] : 0.0;
}

/* The equilibrium system of one buffer (DyadRDCPU.cpp:53-81):
 * row t: sum_I kon_t C_t B_I + koff_t B_t = kon_t C_t Btot. */
void oracle_buffer_equilibrium(int m, const double *kon, const double *koff, const double *c,
                               double total, double *out) {
    double a[16 * 17];
    for (int t = 0; t < m; ++t) {
        for (int I = 0; I < m; ++I) {
            a[t * (m + 1) + I] = kon[t] * c[t];
            if (I == t) a[t * (m + 1) + I] += koff[t];
        }
        a[t * (m + 1) + m] = kon[t] * c[t] * total;
    }
    gauss_elimination(m, a, out);
}

/* ------------------------------------------------------------------------- */
/* The dyad. */

struct OracleDyad {
    int dim, nx, ny, nz, S;
    double dx, dy, dz, zext, R, V;
    int n_ions, n_buf, n_bound;
    double *ion_D, *ion_c0, *buf_total, *bound_D, *bound_kon, *bound_koff, *bound_init;
    int *buf_faces, *bound_ion, *bound_buf;
    short *binds; /* [ion * n_buf + buf] -> bound index or -1 */
    int *sl_binds;
    double *N1, *K1, *N2, *K2;
    int ca, mg;
    int n_types;
    int *type_model;
    double *type_params, *type_g;
    int *ch_off, *ch_node, *ch_type, *ch_state;
    int nch_max;
    uint64_t *rng0;
    double *rng; /* [S*6] current Cg */
    long NE;     /* nodes of all sites */
    long NS;     /* nodes per site */
    double *ions, *buffers, *buf_free, *evo_ions, *evo_buffers, *currents;
    double *mouths, *gradients, *total_sr_current;
    double *rim_ions, *rim_bound, *jsr;
    int have_update;
    int start_eq;
    double grad_mult, sx, sy;
    double *mult_z;
    /* extension: per-node rim values set by a cytosol box, [(f*S + s)*ncol + col]
       for the fields f (ions, then bound) with rim_mask[f] = 1; NULL otherwise */
    double *rim_nodes;
    int *rim_mask;
    int ncol;
};

static long idx3(const OracleDyad *o, int x, int y, int z, int s) {
    return (long)z + o->nz * ((long)y + o->ny * ((long)x + (long)o->nx * s));
}

/* Total concentration of unique buffer u at plane z (extension: face-only buffers). */
static double buf_total_at(const OracleDyad *o, int u, int z) {
    if (o->buf_faces && o->buf_faces[u] && o->dim == 3 && z != 0 && z != o->nz - 1) return 0.0;
    return o->buf_total[u];
}

static void *xcalloc(size_t n, size_t sz) {
    void *p = calloc(n ? n : 1, sz);
    return p;
}

OracleDyad *oracle_create(const CasimDyadConfig *c) {
    if (!c || (c->dim != 2 && c->dim != 3) || c->nx < 3 || c->ny < 3 || c->n_sites < 1) return NULL;
    if (c->dim == 3 && c->nz < 2) return NULL;
    OracleDyad *o = (OracleDyad *)xcalloc(1, sizeof *o);
    o->dim = c->dim;
    o->nx = c->nx;
    o->ny = c->ny;
    o->nz = c->dim == 3 ? c->nz : 1;
    o->S = c->n_sites;
    o->dx = c->dx;
    o->dy = c->dy;
    o->dz = c->dz;
    o->zext = c->z;
    o->R = c->radius;
    o->V = c->voltage;
    o->n_ions = c->n_ions;
    o->n_buf = c->n_buffers;
    o->n_bound = c->n_bound;
    int ni = o->n_ions, nb = o->n_buf, nB = o->n_bound;
    o->ion_D = (double *)xcalloc(ni, sizeof(double));
    o->ion_c0 = (double *)xcalloc(ni, sizeof(double));
    memcpy(o->ion_D, c->ion_D, ni * sizeof(double));
    memcpy(o->ion_c0, c->ion_c0, ni * sizeof(double));
    o->buf_total = (double *)xcalloc(nb, sizeof(double));
    memcpy(o->buf_total, c->buffer_total, nb * sizeof(double));
    o->buf_faces = (int *)xcalloc(nb, sizeof(int));
    if (c->buffer_faces_only)
        for (int u = 0; u < nb; ++u) o->buf_faces[u] = c->buffer_faces_only[u];
    o->bound_ion = (int *)xcalloc(nB, sizeof(int));
    o->bound_buf = (int *)xcalloc(nB, sizeof(int));
    o->bound_D = (double *)xcalloc(nB, sizeof(double));
    o->bound_kon = (double *)xcalloc(nB, sizeof(double));
    o->bound_koff = (double *)xcalloc(nB, sizeof(double));
    o->bound_init = (double *)xcalloc(nB, sizeof(double));
    o->binds = (short *)xcalloc((size_t)ni * nb, sizeof(short));
    for (int i = 0; i < ni * nb; ++i) o->binds[i] = -1;
    for (int b = 0; b < nB; ++b) {
        o->bound_ion[b] = c->bound_ion[b];
        o->bound_buf[b] = c->bound_buffer[b];
        o->bound_D[b] = c->bound_D[b];
        o->bound_kon[b] = c->bound_kon[b];
        o->bound_koff[b] = c->bound_koff[b];
        o->binds[c->bound_ion[b] * nb + c->bound_buffer[b]] = (short)b;
    }
    /* Equilibrium per buffer over the ions it binds, in ion order
     * (DyadRDCPU.cpp:53-81). */
    o->start_eq = c->start_equilibrium;
    if (c->start_equilibrium) {
        for (int u = 0; u < nb; ++u) {
            double kon[16], koff[16], cc[16], out[16];
            int bidx[16], m = 0;
            for (int i = 0; i < ni; ++i) {
                int b = o->binds[i * nb + u];
                if (b >= 0) {
                    kon[m] = o->bound_kon[b];
                    koff[m] = o->bound_koff[b];
                    cc[m] = o->ion_c0[i];
                    bidx[m] = b;
                    ++m;
                }
            }
            if (m > 0) {
                oracle_buffer_equilibrium(m, kon, koff, cc, o->buf_total[u], out);
                for (int t = 0; t < m; ++t) o->bound_init[bidx[t]] = out[t];
            }
        }
    }
    o->sl_binds = (int *)xcalloc(ni, sizeof(int));
    o->N1 = (double *)xcalloc(ni, sizeof(double));
    o->K1 = (double *)xcalloc(ni, sizeof(double));
    o->N2 = (double *)xcalloc(ni, sizeof(double));
    o->K2 = (double *)xcalloc(ni, sizeof(double));
    for (int i = 0; i < ni; ++i) {
        o->sl_binds[i] = c->sl_binds ? c->sl_binds[i] : 0;
        if (o->sl_binds[i]) {
            o->N1[i] = c->sl_N1[i];
            o->K1[i] = c->sl_K1[i];
            o->N2[i] = c->sl_N2[i];
            o->K2[i] = c->sl_K2[i];
        }
    }
    o->ca = c->calcium_index;
    o->mg = c->magnesium_index;
    o->n_types = c->n_channel_types;
    o->type_model = (int *)xcalloc(o->n_types, sizeof(int));
    o->type_params = (double *)xcalloc((size_t)o->n_types * CASIM_MAX_MODEL_PARAMS, sizeof(double));
    o->type_g = (double *)xcalloc((size_t)o->n_types * ni, sizeof(double));
    memcpy(o->type_model, c->type_model, o->n_types * sizeof(int));
    memcpy(o->type_params, c->type_params, (size_t)o->n_types * CASIM_MAX_MODEL_PARAMS * sizeof(double));
    memcpy(o->type_g, c->type_conductance, (size_t)o->n_types * ni * sizeof(double));
    int S = o->S;
    o->ch_off = (int *)xcalloc(S + 1, sizeof(int));
    memcpy(o->ch_off, c->site_channel_offset, (S + 1) * sizeof(int));
    int nch_tot = o->ch_off[S];
    o->ch_node = (int *)xcalloc(nch_tot, sizeof(int));
    o->ch_type = (int *)xcalloc(nch_tot, sizeof(int));
    o->ch_state = (int *)xcalloc(nch_tot, sizeof(int));
    o->nch_max = 0;
    for (int s = 0; s < S; ++s) {
        int n = o->ch_off[s + 1] - o->ch_off[s];
        if (n > o->nch_max) o->nch_max = n;
    }
    for (int k = 0; k < nch_tot; ++k) {
        /* Point::GetCoordsOnGrid (Point.cpp:17-23); z = 0; DyadRD2DCPU maps
         * with dz = 1e-5 (DyadRD2DCPU.cpp:119). */
        int gx = (int)((c->channel_x[k] + 0.5 * o->dx) / o->dx);
        int gy = (int)((c->channel_y[k] + 0.5 * o->dy) / o->dy);
        int gz = o->dim == 3 ? (int)((0.0 + 0.5 * o->dz) / o->dz) : 0;
        if (gx < 0 || gx >= o->nx || gy < 0 || gy >= o->ny) {
            oracle_destroy(o);
            return NULL;
        }
        o->ch_node[k] = gz + o->nz * (gy + o->ny * gx);
        o->ch_type[k] = c->channel_type[k];
    }
    o->rng0 = (uint64_t *)xcalloc((size_t)S * 6, sizeof(uint64_t));
    memcpy(o->rng0, c->rng_state, (size_t)S * 6 * sizeof(uint64_t));
    o->rng = (double *)xcalloc((size_t)S * 6, sizeof(double));
    o->NS = (long)o->nx * o->ny * o->nz;
    o->NE = o->NS * S;
    o->ions = (double *)xcalloc((size_t)o->NE * ni, sizeof(double));
    o->currents = (double *)xcalloc((size_t)o->NE * ni, sizeof(double));
    o->evo_ions = (double *)xcalloc((size_t)o->NE * ni, sizeof(double));
    o->buffers = (double *)xcalloc((size_t)o->NE * nB, sizeof(double));
    o->evo_buffers = (double *)xcalloc((size_t)o->NE * nB, sizeof(double));
    o->buf_free = (double *)xcalloc((size_t)o->NE * nb, sizeof(double));
    o->mouths = (double *)xcalloc((size_t)o->nch_max * S * (ni + nB + nb), sizeof(double));
    o->gradients = (double *)xcalloc((size_t)(ni + nB) * S, sizeof(double));
    o->total_sr_current = (double *)xcalloc((size_t)ni * S, sizeof(double));
    o->rim_ions = (double *)xcalloc((size_t)ni * S, sizeof(double));
    o->rim_bound = (double *)xcalloc((size_t)nB * S, sizeof(double));
    o->jsr = (double *)xcalloc((size_t)o->nch_max * S * ni, sizeof(double));
    /* grad_mult = 3/(4 pi R^3) (DyadRDCPU.cpp:146); sx, sy (:29-30). */
    o->grad_mult = 3.0 / (4 * acos(-1.0) * o->R * o->R * o->R);
    if (o->dim == 3) {
        o->sx = o->dy * o->dz / o->dx;
        o->sy = o->dx * o->dz / o->dy;
    } else {
        o->sx = o->dy * o->zext / o->dx;
        o->sy = o->dx * o->zext / o->dy;
    }
    o->mult_z = (double *)xcalloc(o->nz, sizeof(double));
    /* (i > 0 || i < nz - 1) ? 1 : 0.5 is 1 for every plane when nz >= 2 (DyadRDCPU.cpp:27-28). */
    for (int i = 0; i < o->nz; ++i) o->mult_z[i] = (i > 0 || i < o->nz - 1) ? 1 : 0.5;
    oracle_reset(o);
    return o;
}

void oracle_destroy(OracleDyad *o) {
    if (!o) return;
    void *ptrs[] = {o->ion_D, o->ion_c0, o->buf_total, o->bound_D, o->bound_kon, o->bound_koff,
                    o->bound_init, o->buf_faces, o->bound_ion, o->bound_buf, o->binds, o->sl_binds,
                    o->N1, o->K1, o->N2, o->K2, o->type_model, o->type_params, o->type_g, o->ch_off,
                    o->ch_node, o->ch_type, o->ch_state, o->rng0, o->rng, o->ions, o->buffers,
                    o->buf_free, o->evo_ions, o->evo_buffers, o->currents, o->mouths, o->gradients,
                    o->total_sr_current, o->rim_ions, o->rim_bound, o->jsr, o->mult_z, o->rim_nodes,
                    o->rim_mask};
    for (size_t i = 0; i < sizeof ptrs / sizeof ptrs[0]; ++i) free(ptrs[i]);
    free(o);
}

/* DyadRDCPU::Reset (DyadRDCPU.cpp:203-229), plus channels closed and each
 * site's stream rewound to its start. */
void oracle_reset(OracleDyad *o) {
    int ni = o->n_ions, nb = o->n_buf;
    for (int s = 0; s < o->S; ++s)
        for (int x = 0; x < o->nx; ++x)
            for (int y = 0; y < o->ny; ++y)
                for (int z = 0; z < o->nz; ++z) {
                    long id = idx3(o, x, y, z, s);
                    for (int j = 0; j < ni; ++j) {
                        long ii = id + j * o->NE;
                        o->ions[ii] = o->ion_c0[j];
                        o->currents[ii] = 0;
                        for (int i = 0; i < nb; ++i) {
                            long ib = id + i * o->NE;
                            if (j == 0) o->buf_free[ib] = buf_total_at(o, i, z);
                            int b = o->binds[j * nb + i];
                            if (b >= 0) {
                                double t = buf_total_at(o, i, z) == 0.0 ? 0.0 : o->bound_init[b];
                                o->buffers[id + b * o->NE] = t;
                                o->buf_free[ib] -= t;
                            }
                        }
                    }
                }
    for (int k = 0; k < o->ch_off[o->S]; ++k) o->ch_state[k] = 0;
    for (int i = 0; i < o->S * 6; ++i) o->rng[i] = (double)o->rng0[i];
    memset(o->gradients, 0, sizeof(double) * (o->n_ions + o->n_bound) * o->S);
}

/* DyadChannels::InitOpening (DyadChannels.cpp:83-95), passive channels:
 * SetMacroState(0, runif) for every channel, then SetMacroState(1, runif). */
void oracle_init_opening(OracleDyad *o, int site, int channel) {
    double *Cg = o->rng + 6 * site;
    int off = o->ch_off[site], n = o->ch_off[site + 1] - off;
    for (int k = 0; k < n; ++k) {
        (void)rand_u01(Cg);
        o->ch_state[off + k] = 0;
    }
    (void)rand_u01(Cg);
    if (channel >= 0 && channel < n) o->ch_state[off + channel] = 1;
}

/* DyadRDCPU::_SetIonsAndBuffersNearChannels (DyadRDCPU.cpp:371-390). */
static void set_mouths(OracleDyad *o) {
    int S = o->S, M = o->nch_max, ni = o->n_ions, nB = o->n_bound, nb = o->n_buf;
    for (int i = 0; i < S; ++i) {
        int off = o->ch_off[i], n = o->ch_off[i + 1] - off;
        for (int j = 0; j < n; ++j) {
            long node = o->ch_node[off + j] + o->NS * i;
            for (int k = 0; k < ni; ++k) o->mouths[j + M * (i + k * S)] = o->ions[node + k * o->NE];
            for (int k = 0; k < nB; ++k) o->mouths[j + M * (i + (k + ni) * S)] = o->buffers[node + k * o->NE];
            for (int k = 0; k < nb; ++k)
                o->mouths[j + M * (i + (k + ni + nB) * S)] = o->buf_free[node + k * o->NE];
        }
    }
}

/* DyadRDCPU::_SetBoundaries (DyadRDCPU.cpp:290-320): x rims, then y rims. */
/* Rim column of boundary node (x, y): y = 0 row, y = ny-1 row, x = 0 column,
 * x = nx-1 column (corners in the rows). */
static int rim_col(const OracleDyad *o, int x, int y) {
    if (y == 0) return x;
    if (y == o->ny - 1) return o->nx + x;
    if (x == 0) return 2 * o->nx + y - 1;
    return 2 * o->nx + o->ny - 2 + y - 1;
}

static void set_boundaries(OracleDyad *o) {
    int S = o->S, ni = o->n_ions, nB = o->n_bound;
    for (int pass = 0; pass < 2; ++pass) {
        int na = pass == 0 ? o->nx : o->ny;
        for (int a = 0; a < na; a += na - 1)
            for (int b = 0; b < (pass == 0 ? o->ny : o->nx); ++b)
                for (int z = 0; z < o->nz; ++z)
                    for (int s = 0; s < S; ++s) {
                        int x = pass == 0 ? a : b, y = pass == 0 ? b : a;
                        long id = idx3(o, x, y, z, s);
                        int col = o->rim_nodes ? rim_col(o, x, y) : 0;
                        for (int j = 0; j < ni; ++j)
                            o->ions[id + j * o->NE] = o->rim_nodes && o->rim_mask[j]
                                                          ? o->rim_nodes[((long)j * S + s) * o->ncol + col]
                                                          : o->rim_ions[s + j * S];
                        for (int i = 0; i < nB; ++i) {
                            /* extension: face-only buffers stay 0 off their planes */
                            double v = buf_total_at(o, o->bound_buf[i], z) == 0.0 ? 0.0 : o->rim_bound[s + i * S];
                            if (o->rim_nodes && o->rim_mask[ni + i])
                                v = o->rim_nodes[((long)(ni + i) * S + s) * o->ncol + col];
                            o->buffers[id + i * o->NE] = v;
                        }
                    }
    }
}

/* DyadRDCPU::_GetGradients (DyadRDCPU.cpp:322-369) / DyadRD2DCPU (:305-350). */
static void get_gradients(OracleDyad *o) {
    int S = o->S, ni = o->n_ions, nB = o->n_bound, nIB = ni + nB;
    int zmax = o->dim == 3 ? o->nz - 1 : 1;
    for (int j = 0; j < nIB * S; ++j) o->gradients[j] = 0;
    for (int x = 0; x < o->nx; x += o->nx - 1)
        for (int y = 1; y < o->ny - 1; ++y)
            for (int z = 0; z < zmax; ++z)
                for (int s = 0; s < S; ++s) {
                    long f1 = x == 0 ? idx3(o, x + 1, y, z, s) : idx3(o, x - 1, y, z, s);
                    long f2 = idx3(o, x, y, z, s);
                    long f3 = x == 0 ? idx3(o, x + 2, y, z, s) : idx3(o, x - 2, y, z, s);
                    for (int j = 0; j < ni; ++j) {
                        const double *u = o->ions + (long)j * o->NE;
                        double t = o->dim == 3 ? o->ion_D[j] * o->grad_mult * o->sx * o->mult_z[z]
                                               : o->ion_D[j] * o->grad_mult * o->sx;
                        o->gradients[j + nIB * s] += t * (2 * u[f1] - 1.5 * u[f2] - 0.5 * u[f3]);
                    }
                    for (int i = 0; i < nB; ++i) {
                        const double *u = o->buffers + (long)i * o->NE;
                        double t = o->dim == 3 ? o->bound_D[i] * o->grad_mult * o->sx * o->mult_z[z]
                                               : o->bound_D[i] * o->grad_mult * o->sx;
                        o->gradients[i + ni + nIB * s] += t * (2 * u[f1] - 1.5 * u[f2] - 0.5 * u[f3]);
                    }
                }
    for (int y = 0; y < o->ny; y += o->ny - 1)
        for (int x = 0; x < o->nx; ++x)
            for (int z = 0; z < zmax; ++z)
                for (int s = 0; s < S; ++s) {
                    long f1 = y == 0 ? idx3(o, x, y + 1, z, s) : idx3(o, x, y - 1, z, s);
                    long f2 = idx3(o, x, y, z, s);
                    long f3 = y == 0 ? idx3(o, x, y + 2, z, s) : idx3(o, x, y - 2, z, s);
                    for (int j = 0; j < ni; ++j) {
                        const double *u = o->ions + (long)j * o->NE;
                        double t = o->dim == 3 ? o->ion_D[j] * o->grad_mult * o->sy * o->mult_z[z]
                                               : o->ion_D[j] * o->grad_mult * o->sy;
                        o->gradients[j + nIB * s] += t * (2 * u[f1] - 1.5 * u[f2] - 0.5 * u[f3]);
                    }
                    for (int i = 0; i < nB; ++i) {
                        const double *u = o->buffers + (long)i * o->NE;
                        double t = o->dim == 3 ? o->bound_D[i] * o->grad_mult * o->sy * o->mult_z[z]
                                               : o->bound_D[i] * o->grad_mult * o->sy;
                        o->gradients[i + ni + nIB * s] += t * (2 * u[f1] - 1.5 * u[f2] - 0.5 * u[f3]);
                    }
                }
}

/* DyadRDCPU::_SetCurrents (DyadRDCPU.cpp:250-271) with PassiveIonChannel::Flux
 * (PassiveIonChannel.cpp:28-34): J = state * (g[k] * (jsr - mouth)).  Only
 * ion 0's total is zeroed per call (as the reference does). */
static void set_currents(OracleDyad *o) {
    int S = o->S, M = o->nch_max, ni = o->n_ions;
    for (int i = 0; i < S; ++i) {
        o->total_sr_current[i] = 0;
        int off = o->ch_off[i], n = o->ch_off[i + 1] - off;
        for (int j = 0; j < n; ++j) {
            long node = o->ch_node[off + j] + o->NS * i;
            const double *g = o->type_g + (long)o->ch_type[off + j] * ni;
            for (int k = 0; k < ni; ++k) {
                double cond = g[k] * (o->jsr[j + M * (i + k * S)] - o->ions[node + k * o->NE]);
                double J = o->ch_state[off + j] * cond;
                o->currents[node + k * o->NE] = J;
                o->total_sr_current[i + k * S] += J;
            }
        }
    }
}

/* DyadRDCPU::Update (DyadRDCPU.cpp:392-402). */
void oracle_update(OracleDyad *o, const double *rim_ions, const double *rim_bound, const double *jsr,
                   double voltage) {
    int S = o->S;
    o->V = voltage;
    memcpy(o->rim_ions, rim_ions, sizeof(double) * o->n_ions * S);
    memcpy(o->rim_bound, rim_bound, sizeof(double) * o->n_bound * S);
    memcpy(o->jsr, jsr, sizeof(double) * o->nch_max * S * o->n_ions);
    o->have_update = 1;
    set_mouths(o);
    set_boundaries(o);
    get_gradients(o);
    set_currents(o);
}

/* DyadRDCPU::_EvolutionOp (DyadRDCPU.cpp:404-449) / DyadRD2DCPU (:394-435). */
static void evolution_op(OracleDyad *o, int x, int y, int z, int s) {
    int ni = o->n_ions, nb = o->n_buf;
    long NE = o->NE;
    long c = idx3(o, x, y, z, s);
    long xl = idx3(o, x - 1, y, z, s), xr = idx3(o, x + 1, y, z, s);
    long yl = idx3(o, x, y - 1, z, s), yr = idx3(o, x, y + 1, z, s);
    long zl = c, zr = c;
    if (o->dim == 3) {
        zl = z == 0 ? idx3(o, x, y, z + 1, s) : idx3(o, x, y, z - 1, s);
        zr = z == o->nz - 1 ? idx3(o, x, y, z - 1, s) : idx3(o, x, y, z + 1, s);
    }
    double hx = o->dx * o->dx, hy = o->dy * o->dy, hz = o->dz * o->dz;
    for (int j = 0; j < ni; ++j) {
        const double *u = o->ions + j * NE;
        double vc = u[c];
        double lap = (u[xr] - 2 * vc + u[xl]) / hx + (u[yr] - 2 * vc + u[yl]) / hy;
        if (o->dim == 3) lap = lap + (u[zr] - 2 * vc + u[zl]) / hz;
        o->evo_ions[c + j * NE] = o->ion_D[j] * lap;
    }
    for (int i = 0; i < o->n_bound; ++i) {
        const double *u = o->buffers + i * NE;
        double vc = u[c];
        double lap = (u[xr] - 2 * vc + u[xl]) / hx + (u[yr] - 2 * vc + u[yl]) / hy;
        if (o->dim == 3) lap = lap + (u[zr] - 2 * vc + u[zl]) / hz;
        o->evo_buffers[c + i * NE] = o->bound_D[i] * lap;
    }
    for (int j = 0; j < ni; ++j) {
        long sidx = c + j * NE;
        double R = 0;
        for (int i = 0; i < nb; ++i) {
            int b = o->binds[j * nb + i];
            if (b >= 0) {
                long s2 = c + b * NE, s3 = c + i * NE;
                double R1 = o->bound_koff[b] * o->buffers[s2] - o->bound_kon[b] * o->ions[sidx] * o->buf_free[s3];
                o->evo_buffers[s2] += -R1;
                R += R1;
            }
        }
        o->evo_ions[sidx] += R;
    }
}

/* DyadRDCPU::_GetFreeBuffers (DyadRDCPU.cpp:451-470). */
static void get_free_buffers(OracleDyad *o, int x, int y, int z, int s) {
    int ni = o->n_ions, nb = o->n_buf;
    long c = idx3(o, x, y, z, s);
    for (int j = 0; j < ni; ++j) {
        if (j == 0)
            for (int i = 0; i < nb; ++i) o->buf_free[c + i * o->NE] = buf_total_at(o, i, z);
        for (int i = 0; i < nb; ++i) {
            int b = o->binds[j * nb + i];
            if (b >= 0) o->buf_free[c + i * o->NE] -= o->buffers[c + b * o->NE];
        }
    }
}

/* DyadRDCPU::_UpdateEvo (DyadRDCPU.cpp:472-492); SL factor at z = nz-1 in 3D,
 * at every node in 2D (DyadRD2DCPU.cpp:458-478). */
static void update_evo(OracleDyad *o, int x, int y, int z, int s, double dt) {
    long c = idx3(o, x, y, z, s);
    for (int j = 0; j < o->n_ions; ++j) {
        long i2 = c + j * o->NE;
        o->evo_ions[i2] += o->currents[i2];
        if ((o->dim == 2 || z == o->nz - 1) && o->sl_binds[j]) {
            double s1 = o->K1[j] + o->ions[i2];
            double s2 = o->K2[j] + o->ions[i2];
            o->evo_ions[i2] *= 1 / (1 + o->N1[j] * o->K1[j] / (s1 * s1) + o->N2[j] * o->K2[j] / (s2 * s2));
        }
        o->ions[i2] += o->evo_ions[i2] * dt;
    }
    for (int i = 0; i < o->n_bound; ++i) {
        long i2 = c + i * o->NE;
        o->buffers[i2] += o->evo_buffers[i2] * dt;
    }
}

/* DyadRDCPU::RunRD (DyadRDCPU.cpp:495-506). */
void oracle_run_rd(OracleDyad *o, double dt, int s) {
    for (int x = 1; x < o->nx - 1; ++x)
        for (int y = 1; y < o->ny - 1; ++y)
            for (int z = 0; z < o->nz; ++z) evolution_op(o, x, y, z, s);
    for (int x = 1; x < o->nx - 1; ++x)
        for (int y = 1; y < o->ny - 1; ++y)
            for (int z = 0; z < o->nz; ++z) {
                update_evo(o, x, y, z, s, dt);
                get_free_buffers(o, x, y, z, s);
            }
}

/* DyadChannels::RunMC (DyadChannels.cpp:97-130), passive channels: per
 * channel one draw u1; if u1 < rate * dt the channel toggles with one more
 * draw (MakeTransition(runif())). */
void oracle_run_mc(OracleDyad *o, double dt, int s) {
    double *Cg = o->rng + 6 * s;
    int off = o->ch_off[s], n = o->ch_off[s + 1] - off, M = o->nch_max, S = o->S;
    for (int i = 0; i < n; ++i) {
        double ca = o->ca >= 0 ? o->mouths[i + M * (s + o->ca * S)] : 0.0;
        double mg = o->mg >= 0 ? o->mouths[i + M * (s + o->mg * S)] : 0.0;
        int t = o->ch_type[off + i];
        double rate = oracle_channel_rate(o->type_model[t], o->type_params + (long)t * CASIM_MAX_MODEL_PARAMS,
                                          o->ch_state[off + i], ca, mg);
        if (rand_u01(Cg) < rate * dt) {
            (void)rand_u01(Cg);
            o->ch_state[off + i] = (o->ch_state[off + i] + 1) % 2;
        }
    }
}

/* The loop of RunSimulationsCPU with a constant rim (ParallelHybridSolver.cpp:
 * 408-437): RunRD and RunMC of every site in an OpenMP loop, then the serial
 * Update.  serial_s (may be NULL) accumulates the wall time of the serial part. */
void oracle_step_timed(OracleDyad *o, double dt, int k, int n_threads, double *serial_s) {
    for (int it = 0; it < k; ++it) {
#pragma omp parallel for num_threads(n_threads) schedule(static)
        for (int s = 0; s < o->S; ++s) {
            oracle_run_rd(o, dt, s);
            oracle_run_mc(o, dt, s);
        }
        double t0 = serial_s ? omp_get_wtime() : 0.0;
        set_mouths(o);
        set_boundaries(o);
        get_gradients(o);
        set_currents(o);
        if (serial_s) *serial_s += omp_get_wtime() - t0;
    }
}

void oracle_step(OracleDyad *o, double dt, int k, int n_threads) { oracle_step_timed(o, dt, k, n_threads, NULL); }

int oracle_max_channels(const OracleDyad *o) { return o->nch_max; }
int oracle_site_channels(const OracleDyad *o, int s) { return o->ch_off[s + 1] - o->ch_off[s]; }
int oracle_channel_node(const OracleDyad *o, int s, int ch) { return o->ch_node[o->ch_off[s] + ch]; }

/* DyadRDCPU::GetElementVolume (DyadRDCPU.cpp:195-197): 0.5 dx dy dz;
 * DyadRD2DCPU (DyadRD2DCPU.cpp:189-191): dx dy z. */
double oracle_element_volume(const OracleDyad *o) {
    return o->dim == 3 ? 0.5 * o->dx * o->dy * o->dz : o->dx * o->dy * o->zext;
}

void oracle_bound_init(const OracleDyad *o, double *out) { memcpy(out, o->bound_init, sizeof(double) * o->n_bound); }

void oracle_get_field(const OracleDyad *o, int f, double *out) {
    const double *src = f < o->n_ions ? o->ions + (long)f * o->NE : o->buffers + (long)(f - o->n_ions) * o->NE;
    memcpy(out, src, sizeof(double) * o->NE);
}

void oracle_set_field(OracleDyad *o, int f, const double *in) {
    double *dst = f < o->n_ions ? o->ions + (long)f * o->NE : o->buffers + (long)(f - o->n_ions) * o->NE;
    memcpy(dst, in, sizeof(double) * o->NE);
    /* keep buf_free consistent with the bound fields, as after a RunRD sweep */
    for (int s = 0; s < o->S; ++s)
        for (int x = 0; x < o->nx; ++x)
            for (int y = 0; y < o->ny; ++y)
                for (int z = 0; z < o->nz; ++z) get_free_buffers(o, x, y, z, s);
}

void oracle_get_free_buffer(const OracleDyad *o, int u, double *out) {
    memcpy(out, o->buf_free + (long)u * o->NE, sizeof(double) * o->NE);
}

void oracle_get_channel_states(const OracleDyad *o, int32_t *out) {
    int M = o->nch_max;
    for (int s = 0; s < o->S; ++s) {
        int off = o->ch_off[s], n = o->ch_off[s + 1] - off;
        for (int j = 0; j < M; ++j) out[s * M + j] = j < n ? o->ch_state[off + j] : -1;
    }
}

void oracle_set_channel_states(OracleDyad *o, const int32_t *in) {
    int M = o->nch_max;
    for (int s = 0; s < o->S; ++s) {
        int off = o->ch_off[s], n = o->ch_off[s + 1] - off;
        for (int j = 0; j < n; ++j) o->ch_state[off + j] = in[s * M + j];
    }
}

void oracle_get_mouths(const OracleDyad *o, double *out) {
    memcpy(out, o->mouths, sizeof(double) * o->nch_max * o->S * (o->n_ions + o->n_bound + o->n_buf));
}

void oracle_get_effluxes(const OracleDyad *o, double *out) {
    memcpy(out, o->gradients, sizeof(double) * (o->n_ions + o->n_bound) * o->S);
}

void oracle_get_total_sr_current(const OracleDyad *o, double *out) {
    memcpy(out, o->total_sr_current, sizeof(double) * o->n_ions * o->S);
}

void oracle_get_rng_states(const OracleDyad *o, uint64_t *out) {
    for (int i = 0; i < o->S * 6; ++i) out[i] = (uint64_t)o->rng[i];
}

void oracle_set_rng_states(OracleDyad *o, const uint64_t *in) {
    for (int i = 0; i < o->S * 6; ++i) o->rng[i] = (double)in[i];
}

/* ------------------------------------------------------------------------- */
/* Compartments around the dyad and the batch loop of ParallelHybridSolver.  */

struct OracleHybrid {
    OracleDyad *o;
    int S, ni, M, nB; /* nB: dyad bound fields */
    int ref_indexing; /* 1: SRPool's reads exactly as written (index mix-ups for S > 1) */
    /* SRPool (SRPool.cpp:4-124) */
    int jsr_const;
    double cq, Vjsr, Velem;
    double *nCQ, *KCQ, *tau, *jc0;
    double *jions, *jnew, *jnsr, *IjSR; /* [M*S*ni], [M*S*ni], [ni*S], [ni*S] */
    double *nsr;                        /* NSRConstPool output [site + ion*S] */
    /* CytosolPool (CytosolPool.cpp:3-365) */
    int cyto;
    int nbc, nBc; /* unique buffers (dyad first, then cytosol), bound fields */
    short *cbinds; /* [ni * nbc] -> bound index or -1 */
    int *cb_dyad;  /* [nBc] 1 if the buffer is a dyad buffer */
    double *ctot, *ckon, *ckoff, *conet_b, *cinit, *conet_i, *cbg;
    int *serca;
    double *jmax, *kup;
    double *cions, *cbuf, *cfree, *cevo_i, *cevo_b, *cbb; /* cbb: boundary buffers [site + b*S] */
    double *rim_ions, *rim_bound;                         /* constant rim when there is no pool */
    struct OracleBox *box;                                /* cytosol box, or NULL */
};

struct OracleBox;
static void box_to_rim(OracleHybrid *h);
static void box_link_rates(OracleHybrid *h);
static void box_after_step(OracleHybrid *h, double dt);
static void box_reset(struct OracleBox *b);
static void box_destroy(struct OracleBox *b);

OracleHybrid *oracle_hybrid_create(OracleDyad *o, const CasimPoolConfig *pc, const double *rim_ions,
                                   const double *rim_bound, int ref_indexing) {
    OracleHybrid *h = (OracleHybrid *)xcalloc(1, sizeof *h);
    int S = o->S, ni = o->n_ions, M = o->nch_max, nB = o->n_bound, nb = o->n_buf;
    h->o = o;
    h->S = S;
    h->ni = ni;
    h->M = M;
    h->nB = nB;
    h->ref_indexing = ref_indexing;
    h->jsr_const = pc->jsr_constant;
    h->cq = pc->jsr_cq_total;
    h->Vjsr = pc->jsr_volume;
    h->Velem = oracle_element_volume(o);
    h->nCQ = (double *)xcalloc(ni, sizeof(double));
    h->KCQ = (double *)xcalloc(ni, sizeof(double));
    h->tau = (double *)xcalloc(ni, sizeof(double));
    h->jc0 = (double *)xcalloc(ni, sizeof(double));
    for (int i = 0; i < ni; ++i) {
        h->nCQ[i] = pc->jsr_cq_sites[i];
        h->KCQ[i] = pc->jsr_k_cq[i];
        h->tau[i] = pc->jsr_tau_refill[i];
        h->jc0[i] = pc->jsr_c0[i];
    }
    h->jions = (double *)xcalloc((size_t)M * S * ni, sizeof(double));
    h->jnew = (double *)xcalloc((size_t)M * S * ni, sizeof(double));
    h->jnsr = (double *)xcalloc((size_t)ni * S, sizeof(double));
    h->IjSR = (double *)xcalloc((size_t)ni * S, sizeof(double));
    h->nsr = (double *)xcalloc((size_t)ni * S, sizeof(double));
    for (int i = 0; i < ni; ++i)
        for (int s = 0; s < S; ++s) {
            h->nsr[s + i * S] = pc->nsr_c[i];
            h->jnsr[s + i * S] = h->jc0[i]; /* SRPool's ctor starts ions_NSR at the jSR value (SRPool.cpp:67-72) */
        }
    h->cyto = pc->cytosol_pool;
    h->rim_ions = (double *)xcalloc((size_t)ni * S, sizeof(double));
    h->rim_bound = (double *)xcalloc((size_t)(nB ? nB : 1) * S, sizeof(double));
    if (rim_ions) memcpy(h->rim_ions, rim_ions, sizeof(double) * ni * S);
    if (rim_bound && nB) memcpy(h->rim_bound, rim_bound, sizeof(double) * nB * S);
    /* cytosol species: dyad buffers then cytosol buffers, bound ion-major */
    int nc = pc->cytosol_pool ? pc->n_cyto_buffers : 0;
    h->nbc = nb + nc;
    h->cbinds = (short *)xcalloc((size_t)ni * (h->nbc ? h->nbc : 1), sizeof(short));
    int cap = nB + nc * ni + 1;
    h->cb_dyad = (int *)xcalloc(cap, sizeof(int));
    h->ckon = (double *)xcalloc(cap, sizeof(double));
    h->ckoff = (double *)xcalloc(cap, sizeof(double));
    h->conet_b = (double *)xcalloc(cap, sizeof(double));
    h->cinit = (double *)xcalloc(cap, sizeof(double));
    h->ctot = (double *)xcalloc(h->nbc + 1, sizeof(double));
    h->conet_i = (double *)xcalloc(ni, sizeof(double));
    h->cbg = (double *)xcalloc(ni, sizeof(double));
    h->serca = (int *)xcalloc(ni, sizeof(int));
    h->jmax = (double *)xcalloc(ni, sizeof(double));
    h->kup = (double *)xcalloc(ni, sizeof(double));
    for (int u = 0; u < nb; ++u) h->ctot[u] = o->buf_total[u];
    for (int u = 0; u < nc; ++u) h->ctot[nb + u] = pc->cyto_buffer_total[u];
    int k = 0;
    double D[256];
    for (int j = 0; j < ni; ++j) {
        h->cbg[j] = o->ion_c0[j];
        if (pc->cytosol_pool) h->conet_i[j] = pc->cyto_leak_mult * o->ion_D[j];
        for (int u = 0; u < h->nbc; ++u) {
            h->cbinds[j * h->nbc + u] = -1;
            if (u < nb) {
                int b = o->binds[j * nb + u];
                if (b < 0) continue;
                h->ckon[k] = o->bound_kon[b];
                h->ckoff[k] = o->bound_koff[b];
                D[k] = o->bound_D[b];
                h->cb_dyad[k] = 1;
            } else {
                int c = u - nb;
                if (!pc->cyto_buffer_binds[c * ni + j]) continue;
                h->ckon[k] = pc->cyto_buffer_kon[c * ni + j];
                h->ckoff[k] = pc->cyto_buffer_koff[c * ni + j];
                D[k] = pc->cyto_buffer_D[c * ni + j];
                h->cb_dyad[k] = 0;
            }
            h->cbinds[j * h->nbc + u] = (short)k;
            if (pc->cytosol_pool) h->conet_b[k] = pc->cyto_leak_mult * D[k];
            ++k;
        }
        if (pc->cytosol_pool && pc->serca_binds && pc->serca_binds[j]) {
            h->serca[j] = 1;
            h->jmax[j] = pc->serca_jmax[j];
            h->kup[j] = pc->serca_kup[j];
        }
    }
    h->nBc = k;
    /* equilibrium of every cytosol buffer over the ions it binds (CytosolPool.cpp:42-70) */
    for (int u = 0; u < h->nbc; ++u) {
        double kon[16], koff[16], cc[16], sol[16];
        int idx[16], m = 0;
        for (int j = 0; j < ni; ++j) {
            int b = h->cbinds[j * h->nbc + u];
            if (b >= 0) {
                kon[m] = h->ckon[b];
                koff[m] = h->ckoff[b];
                cc[m] = h->cbg[j];
                idx[m++] = b;
            }
        }
        if (m && o->start_eq) { /* the cytosol follows the dyad's "Start" (CytosolPool.cpp:49,60) */
            oracle_buffer_equilibrium(m, kon, koff, cc, h->ctot[u], sol);
            for (int t = 0; t < m; ++t) h->cinit[idx[t]] = sol[t];
        }
    }
    h->cions = (double *)xcalloc((size_t)ni * S, sizeof(double));
    h->cevo_i = (double *)xcalloc((size_t)ni * S, sizeof(double));
    h->cbuf = (double *)xcalloc((size_t)(h->nBc + 1) * S, sizeof(double));
    h->cevo_b = (double *)xcalloc((size_t)(h->nBc + 1) * S, sizeof(double));
    h->cfree = (double *)xcalloc((size_t)(h->nbc + 1) * S, sizeof(double));
    h->cbb = (double *)xcalloc((size_t)(nB ? nB : 1) * S, sizeof(double));
    return h;
}

void oracle_hybrid_destroy(OracleHybrid *h) {
    if (!h) return;
    void *ptrs[] = {h->nCQ, h->KCQ, h->tau, h->jc0, h->jions, h->jnew, h->jnsr, h->IjSR, h->nsr, h->cbinds,
                    h->cb_dyad, h->ctot, h->ckon, h->ckoff, h->conet_b, h->cinit, h->conet_i, h->cbg, h->serca,
                    h->jmax, h->kup, h->cions, h->cbuf, h->cfree, h->cevo_i, h->cevo_b, h->cbb, h->rim_ions,
                    h->rim_bound};
    for (size_t i = 0; i < sizeof ptrs / sizeof ptrs[0]; ++i) free(ptrs[i]);
    box_destroy(h->box);
    free(h);
}

/* CytosolPool::_GetFreeBuffers (CytosolPool.cpp:270-288). */
static void cyto_free(OracleHybrid *h, int s) {
    int S = h->S;
    for (int j = 0; j < h->ni; ++j) {
        if (j == 0)
            for (int u = 0; u < h->nbc; ++u) h->cfree[s + u * S] = h->ctot[u];
        for (int u = 0; u < h->nbc; ++u) {
            int b = h->cbinds[j * h->nbc + u];
            if (b >= 0) h->cfree[s + u * S] -= h->cbuf[s + b * S];
        }
    }
}

/* ResetSystem (ParallelHybridSolver.cpp:149-157): NSR (no-op), SRPool::Reset
 * (SRPool.cpp:103-116), dyad, CytosolPool::Reset (CytosolPool.cpp:290-315). */
void oracle_hybrid_reset(OracleHybrid *h) {
    int S = h->S, M = h->M;
    for (int i = 0; i < h->ni; ++i) {
        for (int s = 0; s < S; ++s) {
            for (int c = 0; c < M; ++c) {
                h->jions[c + M * (s + i * S)] = h->jc0[i];
                h->jnew[c + M * (s + i * S)] = h->jc0[i];
            }
            h->IjSR[s + i * S] = 0;
        }
    }
    oracle_reset(h->o);
    if (h->box) box_reset(h->box);
    for (int s = 0; s < S; ++s) {
        int kd = 0;
        for (int j = 0; j < h->ni; ++j) {
            h->cions[s + j * S] = h->cbg[j];
            for (int u = 0; u < h->nbc; ++u) {
                if (j == 0) h->cfree[s + u * S] = h->ctot[u];
                int b = h->cbinds[j * h->nbc + u];
                if (b >= 0) {
                    h->cbuf[s + b * S] = h->cinit[b];
                    if (h->cb_dyad[b]) h->cbb[s + (kd++) * S] = h->cinit[b];
                    h->cfree[s + u * S] -= h->cinit[b];
                }
            }
        }
    }
}

/* CytosolPool::RunRD (CytosolPool.cpp:217-268), effluxes from the last dyad update. */
static void cyto_run(OracleHybrid *h, double dt, int s) {
    int S = h->S, ni = h->ni, nIB = ni + h->nB;
    const double *cur = h->o->gradients;
    int kd = 0, ks = 0;
    for (int j = 0; j < ni; ++j) {
        double react = 0;
        long str = s + j * S;
        for (int u = 0; u < h->nbc; ++u) {
            int b = h->cbinds[j * h->nbc + u];
            if (b < 0) continue;
            long s2 = s + b * S, s3 = s + u * S;
            double R1 = h->ckoff[b] * h->cbuf[s2] - h->ckon[b] * h->cions[str] * h->cfree[s3];
            h->cevo_b[s2] = -R1;
            react += R1;
            if (h->cb_dyad[b]) {
                h->cevo_b[s2] += cur[ni + kd + s * nIB];
                ++kd;
            }
        }
        h->cevo_i[str] = react + cur[j + s * nIB];
        if (h->serca[j]) {
            /* ions_bg indexed by the SERCA counter, as CytosolPool.cpp:246 does */
            double ions2 = h->cions[str] * h->cions[str];
            double K2 = h->kup[j] * h->kup[j];
            double b2 = h->cbg[ks] * h->cbg[ks];
            h->cevo_i[str] -= h->jmax[j] * (ions2 / (ions2 + K2) - b2 / (b2 + K2));
            ++ks;
        }
    }
    for (int j = 0; j < ni; ++j) {
        long str = s + j * S;
        h->cevo_i[str] -= h->conet_i[j] * (h->cions[str] - h->cbg[j]);
        h->cions[str] += h->cevo_i[str] * dt;
    }
    kd = 0;
    for (int b = 0; b < h->nBc; ++b) {
        long str = s + b * S;
        h->cevo_b[str] -= h->conet_b[b] * (h->cbuf[str] - h->cinit[b]);
        h->cbuf[str] += h->cevo_b[str] * dt;
        if (h->cb_dyad[b]) h->cbb[s + (kd++) * S] = h->cbuf[str];
    }
    cyto_free(h, s);
}

/* SRPool::Run (SRPool.cpp:70-89). */
static void jsr_run(OracleHybrid *h, double dt, int s) {
    if (h->jsr_const) return;
    int S = h->S, M = h->M, ni = h->ni;
    double beta = 1;
    for (int j = 0; j < ni; ++j) {
        double c = h->ref_indexing ? h->jions[s + j * S] : h->jions[0 + M * (s + j * S)];
        beta += h->nCQ[j] * h->KCQ[j] * h->cq / ((h->KCQ[j] + c) * (h->KCQ[j] + c));
    }
    for (int i = 0; i < ni; ++i) {
        long id = 0 + M * (s + i * S);
        double ns = h->ref_indexing ? h->jnsr[i + s * ni] : h->jnsr[s + i * S];
        double Iref = 1 / h->tau[i] * (ns - h->jions[id]);
        double newc = h->jions[id] + dt / beta * (Iref - h->IjSR[s + i * S] / h->Vjsr);
        for (int c = 0; c < M; ++c) h->jnew[c + M * (s + i * S)] = newc;
    }
}

/* ParallelHybridSolver::Update (ParallelHybridSolver.cpp:159-175). */
void oracle_hybrid_update(OracleHybrid *h) {
    OracleDyad *o = h->o;
    if (h->box) box_to_rim(h);
    if (h->cyto)
        oracle_update(o, h->cions, h->cbb, h->jnew, o->V);
    else
        oracle_update(o, h->rim_ions, h->rim_bound, h->jnew, o->V);
    if (h->box) box_link_rates(h);
    /* SRPool::Update (SRPool.cpp:55-68) */
    if (!h->jsr_const) {
        int S = h->S, M = h->M;
        for (int i = 0; i < h->ni; ++i)
            for (int s = 0; s < S; ++s) {
                h->IjSR[s + i * S] = o->total_sr_current[s + i * S] * h->Velem;
                h->jnsr[s + i * S] = h->nsr[s + i * S];
                for (int c = 0; c < M; ++c) h->jions[c + M * (s + i * S)] = h->jnew[c + M * (s + i * S)];
            }
    }
}

/* The loop body of RunSimulationsCPU (ParallelHybridSolver.cpp:420-430), k times. */
void oracle_hybrid_step(OracleHybrid *h, double dt, int k, int n_threads) {
    for (int it = 0; it < k; ++it) {
#pragma omp parallel for num_threads(n_threads) schedule(static)
        for (int s = 0; s < h->S; ++s) {
            if (h->cyto) cyto_run(h, dt, s);
            oracle_run_rd(h->o, dt, s);
            jsr_run(h, dt, s);
            oracle_run_mc(h->o, dt, s);
        }
        if (h->box) box_after_step(h, dt);
        oracle_hybrid_update(h);
    }
}

/* jSR concentration of every site and ion, [site + ion*S] (the value all of a site's channels see). */
void oracle_hybrid_get_jsr(const OracleHybrid *h, double *out) {
    for (int i = 0; i < h->ni; ++i)
        for (int s = 0; s < h->S; ++s) out[s + i * h->S] = h->jnew[0 + h->M * (s + i * h->S)];
}

/* Cytosol ions [site + ion*S] and bound fields [site + b*S] (b over the cytosol's bound fields). */
void oracle_hybrid_get_cytosol(const OracleHybrid *h, double *ions, double *bound) {
    memcpy(ions, h->cions, sizeof(double) * h->ni * h->S);
    if (bound) memcpy(bound, h->cbuf, sizeof(double) * h->nBc * h->S);
}

int oracle_hybrid_cyto_bound(const OracleHybrid *h) { return h->nBc; }

/* Overwrite the cytosol pool's ions [site + ion*S] and bound fields
 * [site + b*S] (either may be NULL); free buffers and the dyad-side boundary
 * values follow (CytosolPool::_GetFreeBuffers).  For tests. */
void oracle_hybrid_set_cytosol(OracleHybrid *h, const double *ions, const double *bound) {
    if (ions) memcpy(h->cions, ions, sizeof(double) * h->ni * h->S);
    if (bound) memcpy(h->cbuf, bound, sizeof(double) * h->nBc * h->S);
    for (int s = 0; s < h->S; ++s) {
        int kd = 0;
        for (int b = 0; b < h->nBc; ++b)
            if (h->cb_dyad[b]) h->cbb[s + (kd++) * h->S] = h->cbuf[s + b * h->S];
        cyto_free(h, s);
    }
}

/* ------------------------------------------------------------------------- */
/* Extension: a 3D cytosol box around each site, coupled at the dyad's rim.
 * Semantics as declared with CasimCytosolBoxConfig in casim_dyad.h: rim nodes
 * from the bilinear interpolation of the box's z = 0 face, rim-link effluxes
 * D (u_in - u_rim) A / d accumulated over `substeps` dyad steps and added to
 * the box with the same weights, an explicit box step of substeps*dt.  The
 * reference's coupled solver (DyadRD2DwCytosol.cu:119-241) informs the form
 * (reaction terms, SERCA relative to rest); the geometry is new, so this
 * restatement is pinned by conservation, a discrete eigenmode and closed
 * forms rather than by reference vectors. */

typedef struct OracleBox {
    int nx, ny, nz, NB, S, ni, nbb, nbu, nf, ne, far, M, count;
    double h, ih2, ifv;
    double D[16], kon[8], koff[8], total[8], jmax[4], K2[4], base[4];
    int bion[8], bbuf[8];
    int ex_box[16], ex_dyad[16];
    double ex_D[16];
    int ncol, nlink, dnz;
    int *col_x, *col_y, *col_q;
    double *col_w;
    int *lk_rx, *lk_ry, *lk_ix, *lk_iy, *lk_col;
    double *lk_fac, *zw;
    int *dep_row, *dep_ptr, *dep_link;
    double *dep_w;
    double *u, *un, *init, *lrate, *acc;
} OracleBox;

static void box_destroy(OracleBox *b) {
    if (!b) return;
    void *ptrs[] = {b->col_x, b->col_y, b->col_q, b->col_w, b->lk_rx, b->lk_ry, b->lk_ix, b->lk_iy, b->lk_col,
                    b->lk_fac, b->zw, b->dep_row, b->dep_ptr, b->dep_link, b->dep_w, b->u, b->un, b->init,
                    b->lrate, b->acc};
    for (size_t i = 0; i < sizeof ptrs / sizeof ptrs[0]; ++i) free(ptrs[i]);
    free(b);
}

static void box_reset(OracleBox *b) {
    long per = (long)b->S * b->NB;
    for (int f = 0; f < b->nf; ++f)
        for (long n = 0; n < per; ++n) b->u[f * per + n] = b->init[f];
    b->count = 0;
}

/* Bilinear support of (px, py) on the z = 0 face; returns -1 when it touches
 * the x or y faces. */
static int box_bilinear(const OracleBox *b, double px, double py, int q[4], double w[4]) {
    double gx = px / b->h, gy = py / b->h;
    int ax = (int)floor(gx), ay = (int)floor(gy);
    double fx = gx - ax, fy = gy - ay;
    if (!(ax >= 1 && ay >= 1 && ax + (fx > 0) <= b->nx - 2 && ay + (fy > 0) <= b->ny - 2)) return -1;
    int x1 = fx > 0 ? ax + 1 : ax, y1 = fy > 0 ? ay + 1 : ay;
    q[0] = ax + b->nx * ay;
    q[1] = x1 + b->nx * ay;
    q[2] = ax + b->nx * y1;
    q[3] = x1 + b->nx * y1;
    w[0] = (1 - fx) * (1 - fy);
    w[1] = fx * (1 - fy);
    w[2] = (1 - fx) * fy;
    w[3] = fx * fy;
    return 0;
}

int oracle_hybrid_attach_box(OracleHybrid *h, const CasimCytosolBoxConfig *c) {
    OracleDyad *o = h->o;
    if (!c || h->box || h->cyto || c->nx < 4 || c->ny < 4 || c->nz < 2 || c->substeps < 1 || !(c->h > 0)) return -1;
    if (c->n_bound < 0 || c->n_bound > 8 || c->n_buffers < 0 || c->n_buffers > 8 || o->n_ions > 4) return -1;
    OracleBox *b = (OracleBox *)xcalloc(1, sizeof *b);
    int ni = o->n_ions, nB = o->n_bound, S = o->S;
    b->nx = c->nx;
    b->ny = c->ny;
    b->nz = c->nz;
    b->NB = c->nx * c->ny * c->nz;
    b->S = S;
    b->ni = ni;
    b->nbb = c->n_bound;
    b->nbu = c->n_buffers;
    b->nf = ni + b->nbb;
    b->far = c->far_dirichlet != 0;
    b->M = c->substeps;
    b->h = c->h;
    b->ih2 = 1.0 / (c->h * c->h);
    b->ifv = 2.0 / (c->h * c->h * c->h);
    for (int j = 0; j < ni; ++j) {
        b->D[j] = c->ion_D[j];
        double jm = c->serca_jmax ? c->serca_jmax[j] : 0.0, K = c->serca_k ? c->serca_k[j] : 0.0;
        b->jmax[j] = jm;
        b->K2[j] = K * K;
        double c02 = o->ion_c0[j] * o->ion_c0[j];
        b->base[j] = jm != 0.0 ? c02 / (c02 + b->K2[j]) : 0.0;
    }
    for (int u = 0; u < b->nbu; ++u) b->total[u] = c->buffer_total[u];
    for (int k = 0; k < b->nbb; ++k) {
        b->bion[k] = c->bound_ion[k];
        b->bbuf[k] = c->bound_buffer[k];
        b->D[ni + k] = c->bound_D[k];
        b->kon[k] = c->bound_kon[k];
        b->koff[k] = c->bound_koff[k];
    }
    /* exchanged fields: every ion, then the dyad's mobile bound fields */
    for (int j = 0; j < ni; ++j) {
        b->ex_box[b->ne] = j;
        b->ex_dyad[b->ne] = j;
        b->ex_D[b->ne] = o->ion_D[j];
        ++b->ne;
    }
    for (int db = 0; db < nB; ++db) {
        int k = -1;
        for (int t = 0; t < b->nbb; ++t)
            if (c->bound_dyad[t] == db) k = t;
        int face = o->buf_faces && o->buf_faces[o->bound_buf[db]] && o->dim == 3;
        if (face || o->bound_D[db] == 0.0) continue;
        if (k < 0) {
            box_destroy(b);
            return -1;
        }
        b->ex_box[b->ne] = ni + k;
        b->ex_dyad[b->ne] = ni + db;
        b->ex_D[b->ne] = o->bound_D[db];
        ++b->ne;
    }
    /* rest state */
    b->init = (double *)xcalloc(b->nf, sizeof(double));
    for (int j = 0; j < ni; ++j) b->init[j] = o->ion_c0[j];
    if (o->start_eq)
        for (int u = 0; u < b->nbu; ++u) {
            double kon[16], koff[16], cc[16], sol[16];
            int idx[16], m = 0;
            for (int k = 0; k < b->nbb; ++k)
                if (b->bbuf[k] == u) {
                    kon[m] = b->kon[k];
                    koff[m] = b->koff[k];
                    cc[m] = o->ion_c0[b->bion[k]];
                    idx[m++] = k;
                }
            if (!m) continue;
            oracle_buffer_equilibrium(m, kon, koff, cc, b->total[u], sol);
            for (int t = 0; t < m; ++t) b->init[ni + idx[t]] = sol[t];
        }
    /* rim columns and links */
    int nx = o->nx, ny = o->ny;
    b->dnz = o->dim == 3 ? o->nz : 1;
    b->ncol = 2 * nx + 2 * (ny - 2);
    b->nlink = 2 * (nx - 2) + 2 * (ny - 2);
    b->col_x = (int *)xcalloc(b->ncol, sizeof(int));
    b->col_y = (int *)xcalloc(b->ncol, sizeof(int));
    b->col_q = (int *)xcalloc(4 * b->ncol, sizeof(int));
    b->col_w = (double *)xcalloc(4 * b->ncol, sizeof(double));
    int bad = 0;
    for (int y = 0; y < ny; ++y)
        for (int x = 0; x < nx; ++x) {
            if (!(x == 0 || x == nx - 1 || y == 0 || y == ny - 1)) continue;
            int k = rim_col(o, x, y);
            b->col_x[k] = x;
            b->col_y[k] = y;
            if (box_bilinear(b, c->origin_x + x * o->dx, c->origin_y + y * o->dy, b->col_q + 4 * k, b->col_w + 4 * k))
                bad = 1;
        }
    if (bad) {
        box_destroy(b);
        return -1;
    }
    b->lk_rx = (int *)xcalloc(b->nlink, sizeof(int));
    b->lk_ry = (int *)xcalloc(b->nlink, sizeof(int));
    b->lk_ix = (int *)xcalloc(b->nlink, sizeof(int));
    b->lk_iy = (int *)xcalloc(b->nlink, sizeof(int));
    b->lk_col = (int *)xcalloc(b->nlink, sizeof(int));
    b->lk_fac = (double *)xcalloc(b->nlink, sizeof(double));
    double thick = o->dim == 3 ? o->dz : o->zext;
    int l = 0;
    for (int side = 0; side < 2; ++side) /* x-rims: interior neighbour along x */
        for (int y = 1; y < ny - 1; ++y, ++l) {
            b->lk_rx[l] = side ? nx - 1 : 0;
            b->lk_ix[l] = side ? nx - 2 : 1;
            b->lk_ry[l] = b->lk_iy[l] = y;
            b->lk_fac[l] = o->dy * thick / o->dx;
        }
    for (int side = 0; side < 2; ++side) /* y-rims: interior neighbour along y */
        for (int x = 1; x < nx - 1; ++x, ++l) {
            b->lk_ry[l] = side ? ny - 1 : 0;
            b->lk_iy[l] = side ? ny - 2 : 1;
            b->lk_rx[l] = b->lk_ix[l] = x;
            b->lk_fac[l] = o->dx * thick / o->dy;
        }
    for (l = 0; l < b->nlink; ++l) b->lk_col[l] = rim_col(o, b->lk_rx[l], b->lk_ry[l]);
    b->zw = (double *)xcalloc(b->dnz, sizeof(double));
    for (int z = 0; z < b->dnz; ++z) b->zw[z] = (o->dim == 3 && (z == 0 || z == b->dnz - 1)) ? 0.5 : 1.0;
    /* deposit rows: each z = 0 node gathers its (link, weight) pairs in link order */
    int nfn = b->nx * b->ny;
    int *cnt = (int *)xcalloc(nfn, sizeof(int));
    for (l = 0; l < b->nlink; ++l)
        for (int q = 0; q < 4; ++q)
            if (b->col_w[4 * b->lk_col[l] + q] != 0.0) ++cnt[b->col_q[4 * b->lk_col[l] + q]];
    b->dep_row = (int *)xcalloc(nfn, sizeof(int));
    int rows = 0, ent = 0;
    for (int n = 0; n < nfn; ++n) {
        b->dep_row[n] = cnt[n] ? rows++ : -1;
        ent += cnt[n];
    }
    b->dep_ptr = (int *)xcalloc(rows + 1, sizeof(int));
    b->dep_link = (int *)xcalloc(ent, sizeof(int));
    b->dep_w = (double *)xcalloc(ent, sizeof(double));
    for (int n = 0, r = 0; n < nfn; ++n)
        if (cnt[n]) {
            b->dep_ptr[r + 1] = b->dep_ptr[r] + cnt[n];
            ++r;
        }
    int *fill = (int *)xcalloc(rows + 1, sizeof(int));
    for (l = 0; l < b->nlink; ++l)
        for (int q = 0; q < 4; ++q) {
            int k = b->lk_col[l];
            if (b->col_w[4 * k + q] == 0.0) continue;
            int r = b->dep_row[b->col_q[4 * k + q]];
            int at = b->dep_ptr[r] + fill[r]++;
            b->dep_link[at] = l;
            b->dep_w[at] = b->col_w[4 * k + q];
        }
    free(cnt);
    free(fill);
    long per = (long)S * b->NB;
    b->u = (double *)xcalloc((size_t)b->nf * per, sizeof(double));
    b->un = (double *)xcalloc((size_t)b->nf * per, sizeof(double));
    b->lrate = (double *)xcalloc((size_t)b->ne * S * b->nlink, sizeof(double));
    b->acc = (double *)xcalloc((size_t)b->ne * S * b->nlink, sizeof(double));
    box_reset(b);
    /* the dyad takes its rim nodes from the box from now on */
    o->ncol = b->ncol;
    o->rim_nodes = (double *)xcalloc((size_t)(ni + nB) * S * b->ncol, sizeof(double));
    o->rim_mask = (int *)xcalloc(ni + nB, sizeof(int));
    for (int e = 0; e < b->ne; ++e) o->rim_mask[b->ex_dyad[e]] = 1;
    h->box = b;
    return 0;
}

/* Rim values of every exchanged dyad field from the box's z = 0 face. */
static void box_to_rim(OracleHybrid *h) {
    OracleBox *b = h->box;
    OracleDyad *o = h->o;
    long per = (long)b->S * b->NB;
    for (int e = 0; e < b->ne; ++e)
        for (int s = 0; s < b->S; ++s) {
            const double *u = b->u + b->ex_box[e] * per + (long)s * b->NB;
            for (int k = 0; k < b->ncol; ++k) {
                const int *q = b->col_q + 4 * k;
                const double *w = b->col_w + 4 * k;
                o->rim_nodes[((long)b->ex_dyad[e] * b->S + s) * b->ncol + k] =
                    w[0] * u[q[0]] + w[1] * u[q[1]] + w[2] * u[q[2]] + w[3] * u[q[3]];
            }
        }
}

/* Rim-link effluxes D (u_in - u_rim) A / d of the dyad's current state. */
static void box_link_rates(OracleHybrid *h) {
    OracleBox *b = h->box;
    OracleDyad *o = h->o;
    for (int e = 0; e < b->ne; ++e) {
        int f = b->ex_dyad[e];
        const double *u = f < o->n_ions ? o->ions + (long)f * o->NE : o->buffers + (long)(f - o->n_ions) * o->NE;
        for (int s = 0; s < b->S; ++s)
            for (int l = 0; l < b->nlink; ++l) {
                double t = 0.0;
                for (int z = 0; z < b->dnz; ++z)
                    t += b->zw[z] * (u[idx3(o, b->lk_ix[l], b->lk_iy[l], z, s)] - u[idx3(o, b->lk_rx[l], b->lk_ry[l], z, s)]);
                b->lrate[((long)e * b->S + s) * b->nlink + l] = b->ex_D[e] * (b->lk_fac[l] * t);
            }
    }
}

/* One explicit box step of site s. */
static void box_step_site(OracleBox *b, double dt, int s) {
    long per = (long)b->S * b->NB, base = (long)s * b->NB;
    int sy = b->nx, sz = b->nx * b->ny;
    for (int n = 0; n < b->NB; ++n) {
        int x = n % b->nx, t = n / b->nx, y = t % b->ny, z = t / b->ny;
        if (b->far && (x == 0 || x == b->nx - 1 || y == 0 || y == b->ny - 1 || z == b->nz - 1)) {
            for (int f = 0; f < b->nf; ++f) b->un[f * per + base + n] = b->u[f * per + base + n];
            continue;
        }
        int xl = x > 0 ? n - 1 : n + 1, xr = x < b->nx - 1 ? n + 1 : n - 1;
        int yl = y > 0 ? n - sy : n + sy, yr = y < b->ny - 1 ? n + sy : n - sy;
        int zl = z > 0 ? n - sz : n + sz, zr = z < b->nz - 1 ? n + sz : n - sz;
        double v[16], evo[16], fr[8], dep[16];
        for (int f = 0; f < b->nf; ++f) {
            const double *u = b->u + f * per + base;
            double vc = u[n];
            v[f] = vc;
            double lap = (u[xr] - 2 * vc + u[xl]) * b->ih2 + (u[yr] - 2 * vc + u[yl]) * b->ih2 +
                         (u[zr] - 2 * vc + u[zl]) * b->ih2;
            evo[f] = b->D[f] * lap;
        }
        for (int u = 0; u < b->nbu; ++u) fr[u] = b->total[u];
        for (int k = 0; k < b->nbb; ++k) fr[b->bbuf[k]] -= v[b->ni + k];
        for (int j = 0; j < b->ni; ++j) {
            double R = 0.0;
            for (int k = 0; k < b->nbb; ++k) {
                if (b->bion[k] != j) continue;
                double R1 = b->koff[k] * v[b->ni + k] - b->kon[k] * v[j] * fr[b->bbuf[k]];
                evo[b->ni + k] += -R1;
                R += R1;
            }
            evo[j] += R;
            if (b->jmax[j] != 0.0) {
                double c2 = v[j] * v[j];
                evo[j] -= b->jmax[j] * (c2 / (c2 + b->K2[j]) - b->base[j]);
            }
        }
        for (int f = 0; f < b->nf; ++f) dep[f] = 0.0;
        if (z == 0 && b->dep_row[n] >= 0) {
            int r = b->dep_row[n];
            for (int e = 0; e < b->ne; ++e) {
                const double *acc = b->acc + ((long)e * b->S + s) * b->nlink;
                double sum = 0.0;
                for (int k = b->dep_ptr[r]; k < b->dep_ptr[r + 1]; ++k) sum += b->dep_w[k] * acc[b->dep_link[k]];
                dep[b->ex_box[e]] = sum;
            }
        }
        for (int f = 0; f < b->nf; ++f) b->un[f * per + base + n] = v[f] + evo[f] * dt + dep[f] * b->ifv;
    }
}

/* After each dyad step: accumulate dt times the rates of the state the step
 * started from; every M steps, one box step (all sites). */
static void box_after_step(OracleHybrid *h, double dt) {
    OracleBox *b = h->box;
    long n = (long)b->ne * b->S * b->nlink;
    for (long i = 0; i < n; ++i) b->acc[i] = b->count == 0 ? dt * b->lrate[i] : b->acc[i] + dt * b->lrate[i];
    if (++b->count < b->M) return;
    b->count = 0;
#pragma omp parallel for schedule(static)
    for (int s = 0; s < b->S; ++s) box_step_site(b, b->M * dt, s);
    double *t = b->u;
    b->u = b->un;
    b->un = t;
}

int oracle_hybrid_box_fields(const OracleHybrid *h) { return h->box ? h->box->nf : 0; }

/* One box field, [site][z][y][x]. */
void oracle_hybrid_get_box(const OracleHybrid *h, int f, double *out) {
    long per = (long)h->box->S * h->box->NB;
    memcpy(out, h->box->u + f * per, sizeof(double) * per);
}

void oracle_hybrid_set_box(OracleHybrid *h, int f, const double *in) {
    long per = (long)h->box->S * h->box->NB;
    memcpy(h->box->u + f * per, in, sizeof(double) * per);
}
