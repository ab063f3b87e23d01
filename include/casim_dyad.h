/*
 * casim_dyad.h -- C ABI of the MI355X dyadic-cleft reaction-diffusion library
 * (libcasim_dyad.so).
 *
 * The library replaces the dyad compartment of CaSim: the `DyadSeries`
 * implementations DyadRDCPU / DyadRD2DCPU (CPU) and DyadRDMP / DyadRDMP2D
 * (CUDA), together with the per-site stochastic RyR gating they drive through
 * DyadChannels -> KineticModel -> ParallelRng.  Every entry point below names
 * the reference method it stands in for (file:line in IaBogdi/CaSim, directory
 * CaSim/).
 *
 * Conventions
 *  - Units are the reference's: nm, ms, uM; D in nm^2/ms; kon in 1/(uM ms);
 *    koff in 1/ms; channel conductance g in (uM/ms)/uM.
 *  - "site" = one release unit (one cleft + its RyR cluster); the reference
 *    calls it a thread / batch slot (idx_batch).
 *  - Species order is the reference's: ions and buffers in sorted JSON-key
 *    order, bound fields ion-major over (ion, buffer) pairs
 *    (DyadRDCPU.cpp:84-99).
 *  - Host arrays use the reference layouts, noted per argument.  Grid fields
 *    are [site][x][y][z] with z fastest (DyadRDCPU.cpp:277-279).
 *  - Every function returns 0 on success or a negative CASIM_E* code; the
 *    message of the last failure on the calling thread is casim_last_error().
 *    (The reference has no error reporting; this is the ABI's own.)
 *  - Threading: one handle must not be used from two threads at once.  All
 *    device work of a handle is issued on the handle's own HIP stream.
 */
#ifndef CASIM_DYAD_H
#define CASIM_DYAD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CASIM_OK 0
#define CASIM_E_INVALID (-1)   /* bad argument or configuration */
#define CASIM_E_DEVICE (-2)    /* HIP runtime error */
#define CASIM_E_STATE (-3)     /* call out of protocol order */
#define CASIM_E_RANGE (-4)     /* index out of range */

/* Kinetic models of a passive RyR (PassiveIonChannel.cpp:10-24). */
#define CASIM_MODEL_SHEEP_CANNELL 1 /* SheepCannell.cpp:4-18; params: kopen0, ncaopen, kopenmax, kclose0, ncaclose */
#define CASIM_MODEL_TWO_STATE_MWC 2 /* TwoStateMWC.cpp:4-23; params: fCa, KCa, fMg, KMg, KO0, KMgI, kclose */
#define CASIM_MODEL_TWO_STATE_AHTG 3 /* TwoStateaHTG.cpp:4-29; params: fCa, KCa, fMg, KMg, KO0, KMgI, KCL, KI, KOL, kclose */
#define CASIM_MAX_MODEL_PARAMS 16

/* Arithmetic of the state grid. */
#define CASIM_FP64 0
#define CASIM_FP32 1

/*
 * One batch of sites.  Mirrors what the DyadRDCPU / DyadRD2DCPU constructors
 * read from the "Dyad" and "jSR" JSON objects (DyadRDCPU.cpp:8-161,
 * DyadChannels.cpp:17-61), flattened.  All pointers are read during
 * casim_dyad_create only.
 */
typedef struct CasimDyadConfig {
    int32_t dim;                 /* 3 = DyadRDCPU geometry, 2 = DyadRD2DCPU geometry */
    int32_t nx, ny, nz;          /* nodes: x/dx + 1 etc. (DyadRDCPU.cpp:18-20); nz = 1 when dim == 2 */
    double dx, dy, dz;           /* node spacing, nm (dz unused when dim == 2) */
    double z;                    /* JSON "z": cleft height, nm (2D: element thickness) */
    double radius;               /* JSON "Radius", nm (efflux scale, DyadRDCPU.cpp:146) */
    double voltage;              /* JSON "Voltage" */
    int32_t n_sites;

    int32_t n_ions;              /* sorted JSON keys of "Ions" */
    int32_t n_buffers;           /* sorted JSON keys of "Buffers" (unique buffers) */
    int32_t n_bound;             /* (ion, buffer) bound fields, ion-major */
    const double *ion_D;         /* [n_ions] */
    const double *ion_c0;        /* [n_ions] resting concentration ("Concentration") */
    const double *buffer_total;  /* [n_buffers] "Total Concentration" */
    const int32_t *buffer_faces_only; /* [n_buffers] or NULL.  Extension (no reference
                                    counterpart): 1 = the buffer exists only on the
                                    z = 0 and z = nz-1 planes and must have D = 0 */
    const int32_t *bound_ion;    /* [n_bound] ion index of the bound field */
    const int32_t *bound_buffer; /* [n_bound] buffer index of the bound field */
    const double *bound_D;       /* [n_bound] */
    const double *bound_kon;     /* [n_bound] */
    const double *bound_koff;    /* [n_bound] */
    int32_t start_equilibrium;   /* JSON "Start" == "Equilibrium" (DyadRDCPU.cpp:60-81) */
    const int32_t *sl_binds;     /* [n_ions] 1 if "Sarcolemma" lists the ion */
    const double *sl_N1, *sl_K1, *sl_N2, *sl_K2; /* [n_ions] */
    int32_t calcium_index;       /* ion index of "Calcium", -1 if absent */
    int32_t magnesium_index;     /* ion index of "Magnesium", -1 if absent */

    /* Channel types: the sorted keys of "Channels" (DyadChannels.cpp:30-58). */
    int32_t n_channel_types;
    const int32_t *type_model;        /* [n_channel_types] CASIM_MODEL_* */
    const double *type_params;        /* [n_channel_types * CASIM_MAX_MODEL_PARAMS] */
    const double *type_conductance;   /* [n_channel_types * n_ions]; g[k] in sorted key
                                         order of "Conductance" (DyadChannels.cpp:40-43) */

    /* Channels of every site, CSR over sites.  The reference gives every site
       the same list (DyadChannels.cpp:45-58); per-site lists are an extension
       for heterogeneous batches. */
    const int32_t *site_channel_offset; /* [n_sites + 1] */
    const double *channel_x;            /* [total channels] nm */
    const double *channel_y;            /* [total channels] nm */
    const int32_t *channel_type;        /* [total channels] */

    /* MRG32k3a state (Cg[0..5] of RngStream, rngstream.cpp:57) that each
       site's stream starts from; see casim_rng_stream_states. */
    const uint64_t *rng_state;          /* [n_sites * 6] */

    int32_t precision;           /* CASIM_FP64 or CASIM_FP32 state grid */
} CasimDyadConfig;

/*
 * Compartments around the dyad, for the full step of the reference's batch
 * loop (ParallelHybridSolver.cpp:159-175, 408-437).  Arrays over ions follow
 * the dyad's ion order.
 */
typedef struct CasimPoolConfig {
    /* jSR pool (SRPool.cpp:4-124) */
    int32_t jsr_constant;          /* "Constant": the jSR stays at jsr_c0 */
    double jsr_cq_total;           /* "Total CQ (uM)" */
    double jsr_volume;             /* "V jSR (nm3)" */
    const double *jsr_c0;          /* [n_ions] "Concentration" (0 for ions the jSR lacks) */
    const double *jsr_cq_sites;    /* [n_ions] "Binding sites of CQ" */
    const double *jsr_k_cq;        /* [n_ions] "K CQ (uM)" */
    const double *jsr_tau_refill;  /* [n_ions] "T refill (ms)" (1 for ions the jSR lacks) */
    /* network SR (NSRConstPool.cpp:3-29): constant concentrations */
    const double *nsr_c;           /* [n_ions] "Concentration (uM)" */
    /* cytosol.  0: the rim keeps the values of the last update; 1: the
       well-mixed pool of CytosolPool.cpp:3-365 supplies the rim and takes the
       dyad's effluxes */
    int32_t cytosol_pool;
    double cyto_leak_mult;         /* "Leak mult" */
    int32_t n_cyto_buffers;        /* buffers of the cytosol only ("Cytosol"."Buffers", sorted) */
    const double *cyto_buffer_total;  /* [n_cyto_buffers] */
    const double *cyto_buffer_D;   /* [n_cyto_buffers * n_ions] D of (buffer, ion), 0 if unbound */
    const double *cyto_buffer_kon; /* [n_cyto_buffers * n_ions], kon = 0 and koff = 0 if unbound */
    const double *cyto_buffer_koff;
    const int32_t *cyto_buffer_binds; /* [n_cyto_buffers * n_ions] 1 if the buffer binds the ion */
    const int32_t *serca_binds;    /* [n_ions] "SERCA" lists the ion */
    const double *serca_jmax;      /* [n_ions] */
    const double *serca_kup;       /* [n_ions] */
} CasimPoolConfig;

typedef struct CasimDyad CasimDyad;

/* Message of the last failed call on this thread ("" if none). */
const char *casim_last_error(void);

/* Library build information: "<version> gfx950". */
const char *casim_version(void);

/*
 * Initial MRG32k3a states of n consecutive streams after
 * RngStream::SetPackageSeed(seed): stream k starts at A^(k * 2^127) seed,
 * exactly as the k-th RngStream constructed after the seed is set
 * (rngstream.cpp:303-321 constructor, :357-364 SetPackageSeed; seed check
 * :200-234).  Host-only; no device needed.  out: [n * 6].
 */
int casim_rng_stream_states(const uint64_t seed[6], int32_t first_stream, int32_t n, uint64_t *out);

/* Advance one MRG32k3a state by n_draws RandU01() draws (2 generator steps
 * per draw, rngstream.cpp:271-283 with incPrec = true, :306).  Host-only. */
int casim_rng_advance(uint64_t state[6], int64_t n_draws);

/* n_draws RandU01() values of one stream, advancing state (rngstream.cpp:485-491). */
int casim_rng_draws(uint64_t state[6], int32_t n_draws, double *out);

/*
 * Constructor: DyadRDCPU::DyadRDCPU (DyadRDCPU.cpp:8-161) /
 * DyadRD2DCPU::DyadRD2DCPU.  Computes the equilibrium bound concentrations
 * (GaussElimination, DyadRDCPU.cpp:163-193), maps channel coordinates to nodes
 * (Point::GetCoordsOnGrid, Point.cpp:17-23), allocates the site batch on the
 * current HIP device and resets it.
 */
int casim_dyad_create(const CasimDyadConfig *cfg, CasimDyad **out);
int casim_dyad_destroy(CasimDyad *d);

/* DyadSeries::Reset -> DyadRDCPU::Reset (DyadRDCPU.cpp:203-229).  All nodes
 * to the resting/equilibrium state, currents to zero, every channel closed,
 * every site's RNG back to cfg->rng_state. */
int casim_dyad_reset(CasimDyad *d);

/* DyadSeries::InitOpening -> DyadChannels::InitOpening (DyadChannels.cpp:83-95):
 * one draw per channel for SetMacroState(0), then one draw to open `channel`.
 * site = -1 applies to every site with channel[site] from `channels`
 * (NULL: the same `channel` everywhere). */
int casim_dyad_init_opening(CasimDyad *d, int32_t site, int32_t channel, const int32_t *channels);

/*
 * DyadSeries::Update -> DyadRDCPU::Update (DyadRDCPU.cpp:392-402): snapshot
 * the ions/buffers at channel mouths (:371-390), write the Dirichlet rim from
 * the cytosol (:290-320), compute the rim effluxes (:322-369) and the channel
 * currents (:250-271, PassiveIonChannel::Flux).
 *   rim_ions   [site + ion * S]
 *   rim_bound  [site + bound * S]
 *   jsr        [ch + n_ch * (site + ion * S)]   (n_ch = casim_dyad_max_channels)
 *   voltage    the cytosol's V
 */
int casim_dyad_update(CasimDyad *d, const double *rim_ions, const double *rim_bound,
                      const double *jsr, double voltage);

/*
 * Attach the compartments around the dyad: the jSR pool (SRPool.cpp), the
 * constant network SR (NSRConstPool.cpp) and, with cytosol_pool = 1, the
 * well-mixed cytosol (CytosolPool.cpp), all held on the device per site.
 * From then on casim_dyad_update takes its jSR from the pool (jsr may be
 * NULL) and, with a cytosol pool, its rim too (rim arrays may be NULL), and
 * casim_dyad_step advances the pools with the dyad in the reference's step
 * order (ParallelHybridSolver.cpp:420-430: cytosol RunRD, dyad RunRD, jSR
 * Run, gating, then Update).  casim_dyad_reset resets the pools with the
 * dyad (ResetSystem, ParallelHybridSolver.cpp:149-157).  The protocol calls
 * run_rd / run_mc do not advance the pools.  Call once per handle.
 */
int casim_dyad_attach_pools(CasimDyad *d, const CasimPoolConfig *pc);

/* Pool state: jSR [site + ion*S] (the value every channel of a site sees,
 * SRPool::GetIons), cytosol ions [site + ion*S] and cytosol bound fields
 * [site + b*S], b over the cytosol's bound fields (ion-major over dyad
 * buffers then cytosol buffers; count from casim_dyad_cytosol_bound).
 * Any output may be NULL. */
int casim_dyad_get_pools(CasimDyad *d, double *jsr, double *cyto_ions, double *cyto_bound);
int32_t casim_dyad_cytosol_bound(const CasimDyad *d);

/* DyadSeries::RunRD -> DyadRDCPU::RunRD (DyadRDCPU.cpp:495-506) for every
 * site at once (the reference GPU form, DyadRDMP.cu:853-860).  site must be -1. */
int casim_dyad_run_rd(CasimDyad *d, double dt, int32_t site);

/* DyadSeries::RunMC -> DyadChannels::RunMC (DyadChannels.cpp:97-130) for
 * every site at once, on the mouth snapshot of the last update.  site must be -1. */
int casim_dyad_run_mc(CasimDyad *d, double dt, int32_t site);

/*
 * Fused path (no reference counterpart as a single call): k_steps repetitions
 * of the step the reference's batch loop performs with a constant rim and a
 * constant jSR (ParallelHybridSolver.cpp:408-437 with DyadRDCPU, per step
 * RunRD + RunMC for every site, then Update).  Equivalent to k_steps times
 * { run_rd(dt); run_mc(dt); update(rim, jsr) } with the rim and jSR of the
 * last casim_dyad_update call.  The launches are captured once into a HIP
 * graph per (step count, dt) and replayed; one graph holds at most 2048 steps
 * (with a cytosol box, the most whole box steps below that), so a longer call
 * replays several graphs, with the results of one call.
 */
int casim_dyad_step(CasimDyad *d, double dt, int32_t k_steps);

/* The same k_steps as casim_dyad_step, launched eagerly (no graph) with a
 * HIP event pair around every kernel on the handle's stream.  kernel_ms[0]
 * receives the summed duration of the state-sweep kernels, kernel_ms[1] that
 * of the channel kernels.  For measurement; results equal casim_dyad_step's. */
int casim_dyad_step_profiled(CasimDyad *d, double dt, int32_t k_steps, double kernel_ms[2]);

/* Block until the handle's stream is idle. */
int casim_dyad_synchronize(CasimDyad *d);

/* DyadSeries::GetEffluxes (DyadRDCPU.cpp:273-275): [f + (n_ions + n_bound) * site]. */
int casim_dyad_get_effluxes(CasimDyad *d, double *out);

/* DyadSeries::GetTotalSRCurrent (DyadRDCPU.cpp:242-244): [site + ion * S]. */
int casim_dyad_get_total_sr_current(CasimDyad *d, double *out);

/* One field of DyadSeries::GetConcentrations (DyadRDCPU.cpp:516-539):
 * field < n_ions: ion; else bound field (field - n_ions).  out: [site][x][y][z].
 * Face-only bound fields read 0 away from their two planes. */
int casim_dyad_get_field(CasimDyad *d, int32_t field, double *out);

/* Overwrite one field (same layout as casim_dyad_get_field); for tests and restarts. */
int casim_dyad_set_field(CasimDyad *d, int32_t field, const double *in);

/* DyadSeries::GetChannelsStates (DyadRDCPU.cpp:541-550): [site * n_ch + ch],
 * n_ch = casim_dyad_max_channels; padding entries read -1. */
int casim_dyad_get_channel_states(CasimDyad *d, int32_t *out);
int casim_dyad_set_channel_states(CasimDyad *d, const int32_t *in);

/* Mouth snapshot of the last update, DyadRDCPU::ions_near_channels
 * (DyadRDCPU.cpp:371-390): [ch + n_ch * (site + k * S)], k over ions, bound
 * fields, free buffers. */
int casim_dyad_get_mouths(CasimDyad *d, double *out);

/* Release statistics of each site since its InitOpening, the columns of
 * IndChannelCREAnalyzer's scenarios table (IndChannelCREAnalyzer.cpp:89-137)
 * accumulated per step instead of per event (N = open channels after the
 * gating of a step, in effect from that step's end):
 *   out[s*6 + 0] amplitude  = max N
 *   out[s*6 + 1] TTP        = first time N reached the max (ms)
 *   out[s*6 + 2] TTE        = time of the last change of N (ms)
 *   out[s*6 + 3] Q          = sum of N dt (ms)
 *   out[s*6 + 4] Q / TTE    (0 while TTE = 0)
 *   out[s*6 + 5] N now */
int casim_dyad_get_release_stats(CasimDyad *d, double *out);

/* RNG state of each site's stream (RngStream::GetState, rngstream.cpp:415-419): [site * 6]. */
int casim_dyad_get_rng_states(CasimDyad *d, uint64_t *out);
int casim_dyad_set_rng_states(CasimDyad *d, const uint64_t *in);

/* DyadSeries::GetElementVolume (DyadRDCPU.cpp:195-197, DyadRD2DCPU.cpp:189-191). */
double casim_dyad_element_volume(const CasimDyad *d);
/* Largest channel count over the batch's sites (DyadSeries::GetNumSRChannels
 * for a uniform batch, DyadRDCPU.cpp:199-201). */
int32_t casim_dyad_max_channels(const CasimDyad *d);
/* Channel count of one site. */
int32_t casim_dyad_site_channels(const CasimDyad *d, int32_t site);
/* Grid node (flat index x*ny*nz + y*nz + z) of one channel's mouth. */
int32_t casim_dyad_channel_node(const CasimDyad *d, int32_t site, int32_t ch);

/* Algorithmic HBM bytes of one site-step of the state sweep (read every
 * stored node once, write every updated node once). */
double casim_dyad_bytes_per_site_step(const CasimDyad *d);

/* Tuning of the state sweep.  planes_per_block >= 0 selects the
 * plane-marching kernel with that many interior x-planes per workgroup
 * (0 = a whole site per workgroup); -1 selects the thread-per-node kernel.
 * threads_per_block: multiple of 64 (0 = default).  prefetch: planes loaded
 * ahead of the one being updated, 1..3 (0 = 1).  Results do not depend on
 * the tuning. */
int casim_dyad_set_tuning(CasimDyad *d, int32_t planes_per_block, int32_t threads_per_block, int32_t prefetch);

/* Time-histogram input of IndChannelCREAnalyzer::RunAnalysis
 * (IndChannelCREAnalyzer.cpp:53-137, bin width dt = 0.01 ms there): with
 * steps_per_bin = B, bin 0 holds each site's open-channel count after
 * InitOpening and bin k >= 1 the count after step k*B - 1 (what the analyzer
 * assigns to bin k from per-step records), for k < max_bins; 255 marks bins
 * not reached.  max_bins = 0 turns recording off.  Reset clears the bins. */
int casim_dyad_record_open_counts(CasimDyad *d, int32_t steps_per_bin, int32_t max_bins);
/* The recorded counts, [max_bins][S] bytes. */
int casim_dyad_get_open_counts(CasimDyad *d, uint8_t *out);

/* Steps per state pass of casim_dyad_step: 2 (the default) advances pairs of
 * steps in one read and one write of the state, where the configuration
 * allows it (plane-marching kernel, no cytosol pool, a cytosol box only with
 * an even number of substeps); 1 forces one pass per
 * step.  Results do not depend on it.  casim_dyad_steps_per_pass returns the
 * value in effect (1 or 2), -1 for a null handle. */
int casim_dyad_set_steps_per_pass(CasimDyad *d, int32_t steps);
int32_t casim_dyad_steps_per_pass(const CasimDyad *d);

/* Split pipelining of casim_dyad_step (no counterpart in the reference, whose
 * batch loop steps every site in turn): with on = 1 (the default), pairs of
 * steps of a batch of two or more sites without a cytosol pool run as two
 * half-batches on two streams inside the captured graph, the second half one
 * channel call behind the first, so that one half's channel kernel (and a
 * cytosol box's rim effluxes) run beside the other half's state pass; the
 * halves rejoin for each box step.  Sites are independent, so the results are
 * the same bit for bit; 0 keeps one stream. */
int casim_dyad_set_split(CasimDyad *d, int32_t on);

/*
 * 3D cytosol box around each site, coupled at the dyad's rim.  Extension:
 * the reference's coupled solver (DyadRD2DwCytosol.cu:53-59, 119-241) puts
 * the cleft and the cytosol on one 2D grid of two spacings; here each site
 * gets a box of nx*ny*nz nodes at spacing h whose z = 0 face holds the
 * dyad: dyad node (i, j) sits at (origin_x + i*dx, origin_y + j*dy) on that
 * face.
 *   - rim: every rim node of the dyad (all z in 3D) takes, for each field
 *     that exchanges, the box's value bilinearly interpolated on the z = 0
 *     face at its position;
 *   - efflux: each rim link (rim node, its interior neighbour) carries
 *     D (u_in - u_rim) A / d out of the dyad per unit time (A: the link's
 *     face area, dy*z or dy*dz per z node with half weight on the two
 *     z-planes in 3D; d: the link length), exactly what the dyad's explicit
 *     step removes through that link; dt times it accumulates over the
 *     substeps and is added to the box nodes with the same bilinear weights,
 *     divided by the node volume (h^3/2 on the z = 0 face);
 *   - box step: explicit Euler every `substeps` dyad steps with
 *     dt_box = substeps * dt; 7-point diffusion with mirror (no-flux) faces,
 *     buffer reactions koff B - kon C (Btot - sum B), SERCA
 *     -Jmax (c^2/(c^2+K^2) - c0^2/(c0^2+K^2)) (CytosolPool.cpp:240-250 form);
 *     far_dirichlet = 1 holds the faces x = 0, nx-1, y = 0, ny-1 and
 *     z = nz-1 at rest instead.
 * The box starts at the dyad's resting ion concentrations with its buffers
 * at equilibrium (or empty when the dyad does not start at equilibrium).
 * Dyad and box together conserve every ion's total (free + bound) when
 * channels, SERCA and Dirichlet faces are off.
 */
typedef struct CasimCytosolBoxConfig {
    int32_t nx, ny, nz;             /* box nodes */
    double h;                       /* box node spacing, nm */
    double origin_x, origin_y;      /* dyad node (0, 0) on the box's z = 0 face, nm */
    int32_t substeps;               /* dyad steps per box step, >= 1 */
    int32_t far_dirichlet;          /* 1: outer faces held at rest; 0: no-flux */
    const double *ion_D;            /* [n_ions of the dyad] diffusion in the box */
    int32_t n_buffers;              /* unique box buffers */
    const double *buffer_total;     /* [n_buffers] */
    int32_t n_bound;                /* (ion, buffer) bound fields of the box */
    const int32_t *bound_ion;       /* [n_bound] */
    const int32_t *bound_buffer;    /* [n_bound] */
    const double *bound_D;          /* [n_bound] */
    const double *bound_kon;        /* [n_bound] */
    const double *bound_koff;       /* [n_bound] */
    const int32_t *bound_dyad;      /* [n_bound] dyad bound field it exchanges with at the rim, or -1;
                                       every dyad bound field with D > 0 needs one */
    const double *serca_jmax;       /* [n_ions] or NULL (0 = no uptake) */
    const double *serca_k;          /* [n_ions] or NULL */
} CasimCytosolBoxConfig;

/* Attach a cytosol box to every site (not with a cytosol pool; a jSR pool is
 * fine).  Afterwards casim_dyad_update takes NULL rims (the box sets them),
 * casim_dyad_step needs k_steps to be a multiple of substeps (two steps per
 * state pass need an even substeps), casim_dyad_run_rd is refused, and
 * casim_dyad_reset also resets the box.  Requires the bilinear support of
 * every rim node to lie off the box's x and y faces and nx <= 64. */
int casim_dyad_attach_cytosol_box(CasimDyad *d, const CasimCytosolBoxConfig *c);
/* Box fields: the dyad's ions, then the box's bound fields; 0 without a box. */
int32_t casim_dyad_cytosol_box_fields(const CasimDyad *d);
/* One box field, [site][z][y][x] (x fastest).  A set takes effect on the
 * dyad's rim at the next casim_dyad_update (call it before stepping). */
int casim_dyad_get_cytosol_box(CasimDyad *d, int32_t field, double *out);
int casim_dyad_set_cytosol_box(CasimDyad *d, int32_t field, const double *in);

/* HIP stream of the handle (hipStream_t), for callers that time with events. */
void *casim_dyad_stream(CasimDyad *d);

#ifdef __cplusplus
}
#endif

#endif /* CASIM_DYAD_H */
