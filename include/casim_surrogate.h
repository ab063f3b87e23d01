/*
 * casim_surrogate.h -- learned mouth-[Ca2+] surrogate and the quasi-stationary
 * release-site path it drives (part of libcasim_dyad.so).
 *
 * The quasi-stationary path is CaSim's event-driven release site
 * (QuasyStationaryReleaseSite.cpp:89-133): between gating events the mouth
 * [Ca2+] of every RyR is the quasi-steady value for the current open set;
 * the reference obtains it by linear superposition of point-source solutions
 * (StationaryDiffusionModel.cpp:3-11, LinearNaraghiNeher.cpp:9-75).  Here an
 * ensemble of MLPs trained on the full reaction-diffusion solver predicts it,
 * with the ensemble spread as an error estimate.
 *
 * The network: features -> width -> ... (depth hidden layers, SiLU) -> 1,
 * BF16 weights and activations in the GEMMs, FP32 accumulation (MFMA), FP32
 * master weights and Adam state.  All entry points return 0 or a CASIM_E*
 * code (casim_dyad.h); casim_last_error() describes failures.
 */
#ifndef CASIM_SURROGATE_H
#define CASIM_SURROGATE_H

#include <stddef.h>
#include <stdint.h>

#include "casim_dyad.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct CasimMLP CasimMLP;

/* An ensemble of `members` MLPs with n_in inputs, `depth` hidden layers of
 * `width` units and one output, initialised from `seed` (He-uniform
 * weights, zero biases).  No reference counterpart. */
int casim_mlp_create(int32_t n_in, int32_t width, int32_t depth, int32_t members, uint64_t seed, CasimMLP **out);
int casim_mlp_destroy(CasimMLP *m);

/* Number of FP32 parameters of one member, and their layout: for each layer
 * l = 0..depth, W_l row-major [in_l x out_l] followed by b_l [out_l]. */
int64_t casim_mlp_param_count(const CasimMLP *m);
int casim_mlp_get_params(CasimMLP *m, int32_t member, float *out);
int casim_mlp_set_params(CasimMLP *m, int32_t member, const float *in);

/* Feature and target normalisation applied inside train / predict:
 * x' = (x - in_mean) / in_std, prediction = y' * out_std + out_mean. */
int casim_mlp_set_normalization(CasimMLP *m, const float *in_mean, const float *in_std, float out_mean,
                                float out_std);

/* Training set, copied to the device: x [rows x n_in], y [rows] (raw units;
 * normalised with the values above). */
int casim_mlp_set_data(CasimMLP *m, const float *x, const float *y, int64_t rows);

/* `epochs` passes of Adam (lr, beta1 0.9, beta2 0.999, eps 1e-8) over the
 * training set in shuffled minibatches of `batch` rows, every member with its
 * own shuffle.  loss_out[member] receives the mean squared error (normalised
 * units) of the last epoch. */
int casim_mlp_train(CasimMLP *m, int32_t epochs, int32_t batch, float lr, uint64_t seed, float *loss_out);

/* Gradient of the mean squared error (normalised units) of one member on
 * rows of x/y (host arrays), written in the parameter layout above; also the
 * loss.  For checking the device backward pass. */
int casim_mlp_gradient(CasimMLP *m, int32_t member, const float *x, const float *y, int64_t rows, float *grad,
                       float *loss);

/* Ensemble prediction for rows of x (host arrays): mean and relative spread
 * (standard deviation over members / mean) of exp(prediction), i.e. in
 * concentration units when the target is ln [Ca].  Either output may be NULL;
 * member_out [members x rows] receives every member's raw output. */
int casim_mlp_predict(CasimMLP *m, const float *x, int64_t rows, float *mean, float *spread, float *member_out);

/* Inference kernel choice (no counterpart in the reference, which has no
 * learned model): with on = 1 (the default) casim_mlp_predict and the
 * quasi-stationary path run every layer of a member in one fused kernel
 * (activations in LDS, BF16 MFMA, FP32 accumulation) where the shape allows
 * it (members <= 8, layers <= 8, inputs and width <= 256, width a multiple
 * of 64); 0 selects the per-layer GEMM and bias kernels.  The outputs are
 * the same bits either way. */
int casim_mlp_set_fused_inference(CasimMLP *m, int32_t on);

/*
 * Ensemble prediction that allocates nothing (no reference counterpart): for
 * callers that cannot allocate device memory, such as code under HIP graph
 * capture or a tight serving loop that forbids hipMalloc.  Same result as
 * casim_mlp_predict (mean and relative spread of exp(prediction)) for up to
 * CASIM_MLP_NOALLOC_MAX_ROWS rows, but on device arrays: x [rows x n_in]
 * FP32, mean and spread [rows] (either may be NULL), all device pointers,
 * work issued on `stream` (a hipStream_t; NULL: the handle's stream) and not
 * waited for.  The activations live in a fixed workspace of
 * CASIM_MLP_NOALLOC_BYTES bytes per GPU, held from the first
 * casim_mlp_create on that GPU until the process ends; calls on one GPU share
 * it, so calls that overlap on one GPU (two streams, two threads) race.  The
 * call allocates and frees nothing.  Refused with CASIM_E_RANGE: rows above
 * the maximum, layers wider than CASIM_MLP_NOALLOC_MAX_WIDTH, more than
 * CASIM_MLP_NOALLOC_MAX_MEMBERS members.
 */
#define CASIM_MLP_NOALLOC_MAX_ROWS 4096
#define CASIM_MLP_NOALLOC_MAX_WIDTH 256
#define CASIM_MLP_NOALLOC_MAX_MEMBERS 8
#define CASIM_MLP_NOALLOC_BYTES \
    ((size_t)CASIM_MLP_NOALLOC_MAX_ROWS * CASIM_MLP_NOALLOC_MAX_WIDTH * 8 + \
     (size_t)CASIM_MLP_NOALLOC_MAX_MEMBERS * CASIM_MLP_NOALLOC_MAX_ROWS * 4)
int casim_mlp_predict_noalloc(CasimMLP *m, const float *x, int64_t rows, float *mean, float *spread, void *stream);

/*
 * Quasi-stationary release sites driven by the surrogate
 * (QuasyStationaryReleaseSite::RunSimulation, QuasyStationaryReleaseSite.cpp:89-133):
 * per site, InitOpening (one draw per channel, then one to open init_channel),
 * then Gillespie events until no channel is open or t >= t_max, three draws
 * per event: the waiting time -ln(u)/total rate (ParallelRng::rexp), the
 * channel (cumulative rates, :113-122) and MakeTransition's draw.  The mouth
 * [Ca2+] of every RyR for the current open set is the ensemble mean of the
 * surrogate.  A site whose rate-weighted ensemble spread, sum over its RyRs of
 * (rate_i / total rate) x (relative spread of Ca_i), exceeds spread_threshold
 * waits for the caller to supply the full solver's values
 * (casim_qs_set_mouths) before its next event.
 */
typedef struct CasimQSConfig {
    int32_t n_sites;
    const int32_t *site_channel_offset; /* [n_sites + 1] */
    const double *channel_x;            /* [channels] nm */
    const double *channel_y;            /* [channels] nm */
    double half_extent;                 /* nm: the cleft spans [0, 2 half_extent] in x and y */
    double pitch;                       /* nm: length unit of the features (RyR lattice pitch) */
    int32_t model;                      /* CASIM_MODEL_* */
    const double *params;               /* [CASIM_MAX_MODEL_PARAMS] */
    double mg;                          /* uM: the cleft's constant [Mg2+] (MWC / aHTG models) */
    const double *jsr;                  /* [n_sites] uM */
    const uint64_t *rng_state;          /* [n_sites * 6] */
    const int32_t *init_channel;        /* [n_sites] */
    double t_max;                       /* ms */
    double spread_threshold;            /* relative ensemble spread that asks for the full solver */
    /* Depleting jSR (jsr_deplete != 0): SRPool's pool (SRPool.cpp:70-89),
     * dc/dt = ((nsr - c)/tau - I V_elem / V_jSR) / beta, beta = 1 + n K CQ / (K + c)^2,
     * advanced over each interval between events with that interval's open
     * set, I = sum over open RyRs of conductance x (c - c_mouth), c_mouth taken
     * proportional to c within the interval.  The surrogate sees the current c. */
    int32_t jsr_deplete;
    double jsr_volume;                  /* nm^3 */
    double element_volume;              /* nm^3: the dyad node volume of the currents */
    double conductance;                 /* (uM/ms)/uM per open RyR (CasimDyadConfig's type conductance) */
    double jsr_cq_total;                /* uM */
    double jsr_cq_sites;
    double jsr_k_cq;                    /* uM */
    double jsr_tau_refill;              /* ms */
    double nsr;                         /* uM */
} CasimQSConfig;

typedef struct CasimQS CasimQS;

int casim_qs_create(const CasimQSConfig *cfg, CasimQS **out);
int casim_qs_destroy(CasimQS *q);
/* InitOpening of every site, clocks to 0, statistics cleared. */
int casim_qs_reset(CasimQS *q);
/* Up to max_rounds rounds of one event per active site that does not wait
 * for the full solver.  Returns early when no site is active, when
 * wait_batch sites wait, or when every active site waits.  rounds, n_active
 * and n_waiting receive the rounds run and the counts after the last round
 * (any may be NULL).  Calls on one handle must not overlap.  With a
 * stationary model set (casim_qs_set_stationary_model) m may be NULL: the
 * mouth values come from that model and no site waits. */
int casim_qs_run(CasimQS *q, CasimMLP *m, int32_t max_rounds, int32_t wait_batch, int32_t *rounds,
                 int32_t *n_active, int32_t *n_waiting);
/* Sites waiting for the full solver: flags [S] (1 = waiting) and the
 * channel states of every site [S * M] (M = max channels, padding -1). */
int casim_qs_get_waiting(CasimQS *q, int32_t *flags, int32_t *states);
/* Mouth [Ca2+] of one waiting site's channels from the full solver; its next
 * event uses them. */
int casim_qs_set_mouths(CasimQS *q, int32_t site, const float *ca);
/* Per site: amplitude, TTP, TTE, Q = integral of N dt, Q/TTE, events
 * (IndChannelCREAnalyzer.cpp:89-137 columns from the event sequence). */
int casim_qs_get_stats(CasimQS *q, double *out);
/* Per site, the record cut at t_max: amplitude, time of the last event,
 * integral of N dt and N, all over [0, t_max] [S x 4]; what a fixed-step run
 * of length t_max reports (casim_dyad_get_release_stats). */
int casim_qs_get_window_stats(CasimQS *q, double *out);
/* Adds a 57th feature, ln of the linear-superposition estimate of each
 * mouth's [Ca2+] (the structure of StationaryDiffusionModel::GetConcentrations,
 * StationaryDiffusionModel.cpp:3-11): background + (jSR / jsr_ref) x sum over
 * open RyRs (itself included) of added[r^2], r^2 in squared lattice units,
 * interpolated linearly in r^2 and decaying as 1/r past the table's end.
 * n = 0 removes it (56 features).  The surrogate must take the count in use.
 * residual != 0: the ensemble predicts ln [Ca2+] minus that feature, and the
 * mouth value is exp(prediction + feature). */
int casim_qs_set_linear_feature(CasimQS *q, const float *added, int32_t n, float background, float jsr_ref,
                                int32_t residual);

/* The reference's own mouth model in place of the ensemble: the
 * StationaryDiffusionModel that QuasyStationaryReleaseSite builds
 * (QuasyStationaryReleaseSite.cpp:48-49, StationaryDiffusionModel.cpp:3-11)
 * with LinearNaraghiNeher's table (LinearNaraghiNeher.cpp:9-75).  The mouth
 * [Ca2+] of channel j is background + sum over open channels i, in index
 * order, of added(r_ij), r_ij the distance of the channel_x / channel_y
 * positions given at creation, added(r) = sum_m coef[m] exp(-r decay[m]) / r
 * for r > 1e-6 and ca_open otherwise; FP64.  coef and decay are the table's
 * eigenmodes with the unit factors folded in (coef = aa * eigv[n][m] * 1e-4,
 * decay = sqrt|mu_m| * 1e-8; the host computes them, casim_amd/stationary.py).
 * n_modes <= 8; 0 turns the model off.  Not with a depleting jSR (the model
 * does not follow the jSR, as in the reference): CASIM_E_INVALID. */
int casim_qs_set_stationary_model(CasimQS *q, int32_t n_modes, const double *coef, const double *decay,
                                  double background, double ca_open);

/* Time-histogram input of IndChannelCREAnalyzer::RunAnalysis
 * (IndChannelCREAnalyzer.cpp:69-117): at each event at time t, bins from the
 * first unwritten one up to floor(t / bin_ms) (at most max_bins - 1) take the
 * open count before the event; 255 marks bins never written.  Set before
 * casim_qs_reset (which clears the bins); max_bins = 0 turns it off. */
int casim_qs_record_open_counts(CasimQS *q, double bin_ms, int32_t max_bins);
/* The recorded counts, [max_bins][S] bytes. */
int casim_qs_get_open_counts(CasimQS *q, uint8_t *out);
/* Feature rows of every channel for the current open sets [channels x 56 or 57]. */
int casim_qs_get_features(CasimQS *q, float *out);
/* jSR of every site now and the lowest it reached over [0, t_max] [S]
 * (either may be NULL). */
int casim_qs_get_jsr(CasimQS *q, double *now, double *lowest);
/* Last ensemble mean and relative spread of every channel. */
int casim_qs_get_mouths(CasimQS *q, float *mean, float *spread);

/* Plain C = op(A) op(B) in BF16 with FP32 accumulation, on device pointers
 * (for tests of the MFMA kernel): A is M x K (K x M stored when trans_a),
 * B is K x N (N x K stored when trans_b), leading dimensions in elements. */
int casim_gemm_bf16(const void *a, int32_t lda, int32_t trans_a, const void *b, int32_t ldb, int32_t trans_b,
                    float *c, int32_t ldc, int32_t M, int32_t N, int32_t K, float beta, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* CASIM_SURROGATE_H */
