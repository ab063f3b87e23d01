/*
 * casim_oracle.h -- CPU restatement of CaSim's dyad hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP
 * product (libcasim_dyad.so) is compared against, and the CPU baseline that
 * bench.py times beside it.  Nothing in the product may link, load or call it.
 *
 * It restates, in plain C with the reference's operation order:
 *   rngstream.cpp / ParallelRng.cpp      MRG32k3a streams, RandU01 = U01d
 *   DyadRDCPU.cpp / DyadRD2DCPU.cpp      constructor, Reset, Update, RunRD, RunMC
 *   DyadChannels.cpp                     InitOpening, RunMC
 *   PassiveIonChannel.cpp                Flux
 *   SheepCannell.cpp, TwoStateModel.cpp  two-state rates and transitions
 *   TwoStatePoModel.cpp, TwoStateMWC.cpp, TwoStateaHTG.cpp
 *   SRPool.cpp, CytosolPool.cpp, NSRConstPool.cpp, ParallelHybridSolver.cpp:408-437
 *
 * Per-site RNG contract (the reference ties streams to OpenMP threads,
 * ParallelRng.cpp:33-35): site k draws from its own stream whose start state
 * is given in the configuration; a single-site reference run (NumBatch = 1,
 * NumCores = 1) seeded with that state draws the same sequence.
 */
#ifndef CASIM_ORACLE_H
#define CASIM_ORACLE_H

#include <stdint.h>
#include "../include/casim_dyad.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct OracleDyad OracleDyad;

/* RNG (rngstream.cpp), double arithmetic as in the reference. */
int oracle_rng_check_seed(const uint64_t seed[6]);
void oracle_rng_stream_states(const uint64_t seed[6], int first, int n, uint64_t *out);
void oracle_rng_draws(uint64_t state[6], int n, double *out);
void oracle_rng_substream(uint64_t state[6]); /* ResetNextSubstream from a stream start */
void oracle_rng_jump_matrix(int which, int e, double out[9]); /* A1 (which=1) / A2 (2) ^ (2^e) mod m */

/* Kinetic models: rate for a channel in `state` given mouth Ca and Mg. */
double oracle_channel_rate(int model, const double *params, int state, double ca, double mg);

/* Equilibrium bound concentrations of one buffer binding m ions:
 * kon[m], koff[m], c[m] (ion concentrations), total -> out[m].
 * GaussElimination as DyadRDCPU.cpp:163-193. */
void oracle_buffer_equilibrium(int m, const double *kon, const double *koff, const double *c,
                               double total, double *out);

/* Dyad (DyadRDCPU / DyadRD2DCPU) */
OracleDyad *oracle_create(const CasimDyadConfig *cfg);
void oracle_destroy(OracleDyad *o);
void oracle_reset(OracleDyad *o);
void oracle_init_opening(OracleDyad *o, int site, int channel);
void oracle_update(OracleDyad *o, const double *rim_ions, const double *rim_bound,
                   const double *jsr, double voltage);
void oracle_run_rd(OracleDyad *o, double dt, int site);
void oracle_run_mc(OracleDyad *o, double dt, int site);
/* k steps of {RunRD(site), RunMC(site) for every site (OpenMP, n_threads)}
 * followed by Update with the rim/jSR of the last oracle_update. */
void oracle_step(OracleDyad *o, double dt, int k, int n_threads);
void oracle_step_timed(OracleDyad *o, double dt, int k, int n_threads, double *serial_s);

int oracle_max_channels(const OracleDyad *o);
int oracle_site_channels(const OracleDyad *o, int site);
int oracle_channel_node(const OracleDyad *o, int site, int ch);
double oracle_element_volume(const OracleDyad *o);
void oracle_bound_init(const OracleDyad *o, double *out);
void oracle_get_field(const OracleDyad *o, int field, double *out);
void oracle_set_field(OracleDyad *o, int field, const double *in);
void oracle_get_free_buffer(const OracleDyad *o, int buffer, double *out);
void oracle_get_channel_states(const OracleDyad *o, int32_t *out);
void oracle_set_channel_states(OracleDyad *o, const int32_t *in);
void oracle_get_mouths(const OracleDyad *o, double *out);
void oracle_get_effluxes(const OracleDyad *o, double *out);
void oracle_get_total_sr_current(const OracleDyad *o, double *out);
void oracle_get_rng_states(const OracleDyad *o, uint64_t *out);
void oracle_set_rng_states(OracleDyad *o, const uint64_t *in);

/* Compartments around the dyad and the reference's batch loop:
 * SRPool, NSRConstPool, CytosolPool (or a constant rim) and
 * ParallelHybridSolver::Update / RunSimulationsCPU.  ref_indexing = 1 makes
 * SRPool::Run read exactly the elements the reference reads (SRPool.cpp:75,80),
 * which equal the per-site elements only for one site. */
typedef struct OracleHybrid OracleHybrid;
OracleHybrid *oracle_hybrid_create(OracleDyad *o, const CasimPoolConfig *pc, const double *rim_ions,
                                   const double *rim_bound, int ref_indexing);
void oracle_hybrid_destroy(OracleHybrid *h);
void oracle_hybrid_reset(OracleHybrid *h);
void oracle_hybrid_update(OracleHybrid *h);
void oracle_hybrid_step(OracleHybrid *h, double dt, int k, int n_threads);
void oracle_hybrid_get_jsr(const OracleHybrid *h, double *out);
void oracle_hybrid_get_cytosol(const OracleHybrid *h, double *ions, double *bound);
int oracle_hybrid_cyto_bound(const OracleHybrid *h);
void oracle_hybrid_set_cytosol(OracleHybrid *h, const double *ions, const double *bound);

/* Extension: the 3D cytosol box of CasimCytosolBoxConfig (casim_dyad.h),
 * the same coupling order as casim_dyad_step.  Returns 0, or -1 for a
 * configuration it refuses.  Box fields are [site][z][y][x]. */
int oracle_hybrid_attach_box(OracleHybrid *h, const CasimCytosolBoxConfig *c);
int oracle_hybrid_box_fields(const OracleHybrid *h);
void oracle_hybrid_get_box(const OracleHybrid *h, int f, double *out);
void oracle_hybrid_set_box(OracleHybrid *h, int f, const double *in);

#ifdef __cplusplus
}
#endif
#endif
