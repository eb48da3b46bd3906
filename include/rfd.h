/*
 * rfd.h -- C ABI of the MI355X-native RockFit-DEM bonded-particle-model step.
 *
 * Two tiers:
 *   1. Reference-layout tier: one entry point per launch site of the
 *      reference's per-step pipeline, taking the reference's own device
 *      buffers (slot layout: K slots per particle, SoA-by-component particle
 *      arrays, AoS float3 slot state, 0xFFFFFFFF sentinel).  A host driver that
 *      used to launch the CUDA kernels calls these instead.
 *   2. Native tier: an opaque rfd_system that owns SoA particle and edge-list
 *      bond storage in HBM (polydisperse radii, per-bond strengths, FP64
 *      position/rotation accumulators) and runs whole steps, FIRE load steps
 *      and batched specimens on the device.
 *
 * Conventions
 *   - Every entry point returns int: 0 ok, <0 invalid argument / library
 *     error (RFD_E_*), >0 a hipError_t code.  Nothing aborts; the message of
 *     the last failure is rfd_last_error(ctx).
 *   - Tier-1 pointers are caller-owned device memory; counts are uint32_t.
 *   - Calls are asynchronous on the context's stream, except those that return
 *     host values (platen sums, FIRE power, flags), which synchronise.
 *   - Kernels never allocate caller memory; the context owns its scratch: a
 *     grow-only device workspace, allocated by the first call that needs more
 *     than it holds (cell binning: 4 B per cell and per particle; list
 *     construct: a 16 B position record per particle; list reconstruct: that
 *     record plus 12 B per particle and 8 B per cell of its search grid, whose
 *     cell edge is 0.505 sqrt(aacut)) and freed by rfd_ctx_destroy.  Such a
 *     first call synchronises the context's stream.
 *   - One context per host thread.
 */
#ifndef RFD_H
#define RFD_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RFD_SENTINEL 0xFFFFFFFFu

#define RFD_OK 0
#define RFD_E_ARG (-1)       /* null pointer, zero size, inconsistent sizes */
#define RFD_E_NODEV (-2)     /* no HIP device */
#define RFD_E_STATE (-3)     /* call out of order (e.g. system not built) */
#define RFD_E_CAPACITY (-4)  /* a fixed capacity was exceeded */

/* device-side condition flags, read (and cleared) by rfd_read_flags */
#define RFD_FLAG_CELL_RANGE 0x1u /* a particle lies outside the cell grid */
#define RFD_FLAG_ROW_FULL 0x2u   /* a neighbour row needed more than K slots */

typedef struct rfd_ctx rfd_ctx;

/* ------------------------------------------------------------ context */
/* stream: a hipStream_t to run on, or NULL for the device's null stream */
int rfd_ctx_create(rfd_ctx **out, int device, void *stream);
int rfd_ctx_destroy(rfd_ctx *ctx);
const char *rfd_last_error(const rfd_ctx *ctx);
/* wait for all work queued on the context's stream
 * (the reference's implicit legacy-stream ordering, md_results.cu:36) */
int rfd_sync(rfd_ctx *ctx);
/* synchronise, return the OR of RFD_FLAG_* raised since the last call, clear */
int rfd_read_flags(rfd_ctx *ctx, uint32_t *flags);
/* library build description, e.g. "rfd gfx950 ..." */
const char *rfd_version(void);

/* No counterpart in the reference: the measured HBM ceiling that roofline
 * figures are read against.  Copies `bytes` (a multiple of 16, both buffers
 * 16-byte aligned) from src to dst with 16-byte non-temporal loads and
 * stores, `blocks` workgroups of 256 threads (0: 4096).  Asynchronous. */
int rfd_stream_copy(rfd_ctx *ctx, void *dst, const void *src, size_t bytes, uint32_t blocks);

/* ========================= tier 1: reference layout ===================== */

/* CellDistribution (md_cell_distribution.cu:52-101): d_CalculateCellIndex
 * (md_calculate_cell.cu:28), CUB SortPairs (:68,:87), d_DetermineCellPointer
 * (md_calculate_cell.cu:55).  CI[N]; CIs[2N] = sorted cell keys | particle
 * ids (stable); pnC[2C] = start | count, zero for empty cells. */
int rfd_cell_distribution(rfd_ctx *ctx, const float *R, uint32_t N, float inv_a,
                          uint32_t Nx, uint32_t Ny, uint32_t Nz, uint32_t *CI,
                          uint32_t *CIs, uint32_t *pnC);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): rfd_cell_distribution with the same arguments and
 * results, whose workspace is a fixed 64 MiB array in the library's own
 * device memory (one per GPU, held until the process ends) instead of the
 * context's grow-only workspace; the call allocates nothing.  Limit:
 * N + C + ceil(C / 8192) <= 16,777,216 (C = Nx Ny Nz), else RFD_E_CAPACITY.
 * Every call on one GPU shares that array: calls that overlap on one GPU (two
 * streams, two threads) race and must be ordered by the caller. */
int rfd_cell_distribution_noalloc(rfd_ctx *ctx, const float *R, uint32_t N, float inv_a,
                                  uint32_t Nx, uint32_t Ny, uint32_t Nz, uint32_t *CI,
                                  uint32_t *CIs, uint32_t *pnC);

/* InteractionListConstruct (md_interaction_list_construction.cu:62-67) =
 * memset IL 0xFF + d_ConstructInteractionList (md_interaction_list_construct.cu:35) */
int rfd_il_construct(rfd_ctx *ctx, const float *R, uint32_t N,
                     const uint32_t *CIs, const uint32_t *pnC, uint32_t *IL,
                     uint32_t K, float a, float inv_a, float aacut, uint32_t Nx,
                     uint32_t Ny, uint32_t Nz);

/* InteractionListReConstructBPM (md_interaction_list_construction.cu:75-79) =
 * d_ReConstructInteractionListbpm (md_interaction_list_reconstructbpm.cu:80).
 * quirk_p0 != 0 keeps the reference's particle-0 free-slot rule (:32).  New
 * partners are inserted in the order of the 8-cell scan over the caller's
 * lists: search cell by search cell, each cell's members in the order CIs
 * holds them (any order, not only ascending id). */
int rfd_il_reconstruct_bpm(rfd_ctx *ctx, const float *R, uint32_t N,
                           const uint32_t *CIs, const uint32_t *pnC,
                           uint32_t *IL, uint8_t *ILtype, uint32_t K, float a,
                           float inv_a, float aacut, uint32_t Nx, uint32_t Ny,
                           uint32_t Nz, int quirk_p0);

/* d_ConstructBoundInteractions (md_boundinteractions_construct.cu:35),
 * launched at md_calculationuniaxialcompression.cu:133.  float3 arrays are
 * passed as float* (3 floats per slot). */
int rfd_bonds_construct(rfd_ctx *ctx, const float *R, const uint32_t *IL,
                        uint8_t *ILtype, float *iL, float *rij, float *Oijt,
                        float *Fijn, float *Fijt, float *Mijn, float *Mijt,
                        float *Mijadd, uint32_t N, uint32_t K, float bdd);

/* d_DeleteFarLinks (md_interaction_list_reconstructbpm.cu:141), launched at
 * md_calculationuniaxialcompression.cu:142,186 */
int rfd_delete_far_links(rfd_ctx *ctx, const float *R, uint32_t N,
                         uint32_t *IL, uint8_t *ILtype, float *rij,
                         float *Oijt, float *Fijn, float *Fijt, float *Mijn,
                         float *Mijt, float *Mijadd, uint32_t K,
                         float aa_farcut);

/* cudaMemset F, M (md_calculationuniaxialcompression.cu:171-172) */
int rfd_zero_fm(rfd_ctx *ctx, float *F, float *M, uint32_t N);

/* d_CalculateForcesDEM_22 (md_calculateforcesdem3.cu:424): bond force and
 * moments with the fused break check, F/M += per particle.
 * Launched at md_calculationuniaxialcompression.cu:173. */
int rfd_forces_bpm(rfd_ctx *ctx, const float *R, const float *V,
                   const float *W, const uint32_t *IL, uint8_t *ILtype,
                   const float *iL, float *rij, float *Oijt, float *Mijn,
                   float *Mijt, float *F, float *M, uint32_t N, uint32_t K,
                   float dt, float E, float G, float b_r, float sc, float tc);

/* d_CalculateForcesDEM_21 (md_calculateforcesdem3.cu:338): Hertz-Mindlin
 * contact, F/M += per particle.  Launched at
 * md_calculationuniaxialcompression.cu:175. */
int rfd_forces_contact(rfd_ctx *ctx, const float *R, const float *V,
                       const float *W, const uint32_t *IL,
                       const uint8_t *ILtype, float *Oijt, float *F, float *M,
                       uint32_t N, uint32_t K, float dt, float Estar,
                       float Gstar, float pA, float mu, float mur, float m,
                       float r);

/* The superseded per-slot force path, kept for A/B comparisons:
 * d_CalculateForcesDEM (md_calculateforcesdem.cu:328) for mode 0 (bond slots;
 * empty slots zeroed; contact slots untouched, as its contact branch is off),
 * d_CalculateForcesDEM_1 (:501) for mode 1 (contact slots within 2 rP get the
 * Hertz-Mindlin law with E, G = the contact moduli; the others and empty slots
 * are zeroed), d_CalculateForcesDEM_2 (:581) for mode 2 (bond slots only).
 * Every slot's loads go to Fijn, Fijt, Mijn, Mijt, Mijadd (and its state to
 * rij, Oijt); rfd_sumup_forces then forms F, M and rfd_check_break judges the
 * bonds.  float3 arrays are passed as float* (3 floats per slot). */
int rfd_forces_slots(rfd_ctx *ctx, int mode, const float *R, const float *V,
                     const float *W, const uint32_t *IL, const uint8_t *ILtype,
                     const float *iL, float *rij, float *Oijt, float *Fijn,
                     float *Fijt, float *Mijn, float *Mijt, float *Mijadd,
                     uint32_t N, uint32_t K, float dt, float E, float G,
                     float b_r, float pA, float mu, float mur, float mP,
                     float rP);

/* d_SumUpForcesDEM (md_sumupforcesdem.cu:19): F,M = row sums of slot terms */
int rfd_sumup_forces(rfd_ctx *ctx, const float *Fijn, const float *Fijt,
                     const float *Mijn, const float *Mijt, const float *Mijadd,
                     float *F, float *M, uint32_t N, uint32_t K);

/* d_CheckBreakConditionDEM (md_checkbreakconditiondem.cu:20), S = N*K slots */
int rfd_check_break(rfd_ctx *ctx, const float *rij, float *Oijt,
                    const float *Fijn, const float *Fijt, const float *Mijn,
                    const float *Mijt, uint8_t *ILtype, uint32_t S, float sc,
                    float tc, float b_r);

/* d_CalculateIncrementsDEM (md_calculateincrementsdem.cu:19) */
int rfd_increments(rfd_ctx *ctx, const float *F, float *V, float *R,
                   const float *M, float *W, uint32_t N, float dt_m, float dt,
                   float dt_I);

/* d_UniaxialCompression3_simple (md_uniaxialcompression3_simple.cu:19) plus
 * the host sum of its partials (SumForcesULoading, md_results.cu:36-47).
 * Synchronous: out_sums[0] = bottom reaction, out_sums[1] = top reaction,
 * summed in FP64 on the device. */
int rfd_platen_uniaxial(rfd_ctx *ctx, float *R, float *V, float *F, uint32_t N,
                        float cx, float cy, float RR, float Zt, float Zb,
                        float Ztr, float Zbr, float vt, float Zcut,
                        double out_sums[2]);

/* d_UniaxialCompression_simple (md_uniaxialcompression_simple.cu:29), the
 * superseded penalty platens: a particle inside the footprint (RR, below
 * Zcut) past z = cz + Hd2 or cz - Hd2 gets F.z += C times its penetration,
 * with Coulomb friction mu |F.z| against its in-plane velocity
 * (dd_FrictionForce :19).  out_sums = (bottom, top) reaction sums in FP64
 * (the reference adds FP32 block partials on the host); synchronises. */
int rfd_platen_uniaxial_penalty(rfd_ctx *ctx, const float *R, const float *V,
                                float *F, uint32_t N, float cx, float cy,
                                float cz, float RR, float Hd2, float C,
                                float mu, float Zcut, double out_sums[2]);

/* d_braziltest_simple2 (md_braziltest_simple2.cu:19) + partial sums */
int rfd_platen_brazil(rfd_ctx *ctx, float *R, float *V, float *F, uint32_t N,
                      float cx, float cz, float RR, float Yt, float Yb,
                      float Ytr, float Ybr, float vt, float Zcut,
                      double out_sums[2]);

/* d_ParallelepipedCutRestriction (md_parallelepipedrestriction.cu:73) */
int rfd_cut_restriction(rfd_ctx *ctx, float *R, float *V, float *F, uint32_t N,
                        const float center[3], const float sized2[3],
                        const float hidden_point[3]);

/* d_CalculateIncrementsFIRE (md_firerelaxation.cu:18) */
int rfd_fire_increments(rfd_ctx *ctx, const float *F, float *V, float *R,
                        uint32_t N, float dt_m, float dt, float alpha);
/* d_CalculateDecrementsHalfStepFIRE (md_firerelaxation.cu:58) */
int rfd_fire_halfstep(rfd_ctx *ctx, const float *V, float *R, uint32_t N,
                      float dt_d2);
/* d_FdotVEntire (md_firerelaxation.cu:372) + host sum (:267-272), in FP64.
 * out[0]=sum F.V (+M.W), out[1]=sum |F|^2 (+|M|^2), out[2]=sum |V|^2 (+|W|^2),
 * out[3]=max_i |F_i|^2.  M and W may be NULL (translation only). */
int rfd_fire_power(rfd_ctx *ctx, const float *F, const float *V,
                   const float *M, const float *W, uint32_t N, double out[4]);

/* The FIRE relaxation loop of the reference (CalculateGPUStepsContractRelaxFIRE,
 * md_firerelaxation.cu:201-304) over the three kernels above: per iteration
 * the caller's force callback fills F, P = F.V is summed (in FP64 here; the
 * reference sums FP32 block partials on the host in FP32), then
 *   P > 0:  ++n_positive, n_negative = 0; after n_delay positive iterations
 *           dt = min(dt dtgrow, dtmax) and alpha *= alphashrink;
 *   else:   n_positive = 0, ++n_negative (stop once above n_negative_max);
 *           after the first n_delay iterations dt = max(dt dtshrink, dtmin)
 *           and alpha = alpha0; R -= V dt/2 with the new dt, V = 0;
 * then d_CalculateIncrementsFIRE with dt_m (the potential's dt/m, fixed) and the
 * current dt and alpha, and the caller's restriction callback (the reference
 * calls d_CylinderRestrictionZ there).  dt and alpha are float, as in
 * firerelax_data (md_data_types.h:296-301).  The state persists in rfd_fire_ref
 * across calls; rfd_fire_ref_init sets SetFIREData's values
 * (md_firerelaxation.cu:72-84: dtmax 100 dt0, dtmin 1e-3 dt0, dtgrow 1.1,
 * dtshrink 0.5, alphashrink 0.99, n_negative_max 30000, n_delay 20) and the
 * initial alpha, which the reference leaves unset (0.25 is its alpha0; its
 * packing uses 0.1, md_createcylindersample.cu:266).  A callback returns 0 on
 * success; its non-zero return ends the call with that status. */
typedef struct rfd_fire_ref {
  float dt0, dtmax, dtmin, dtgrow, dtshrink, alpha0, alphashrink;
  uint32_t n_negative_max, n_delay;
  float dt, alpha;            /* current time step and mixing weight        */
  uint32_t n_positive, n_negative;
  uint32_t steps;             /* iterations taken (the loop counter)        */
  int stopped;                /* 1 once n_negative exceeded n_negative_max  */
  double power;               /* F.V of the last iteration                  */
} rfd_fire_ref;
typedef int (*rfd_callback)(void *user);
void rfd_fire_ref_init(rfd_fire_ref *f, float dt0, float alpha0);
int rfd_fire_relax_reference(rfd_ctx *ctx, const float *F, float *V, float *R, uint32_t N,
                             float dt_m, rfd_fire_ref *f, uint32_t iterations,
                             rfd_callback forces, rfd_callback restriction, void *user);

/* =========================== tier 2: native system ======================
 * Particles: R in FP64 (position accumulation), V, W, F, M in FP32, SoA by
 * component; per-particle radius (polydisperse).  Bonds: one record per
 * undirected pair (i<j) in SoA, grouped by owner i (CSR), with per-bond
 * radius, rest length and strengths; bond state n_prev, O, Mn, Mt (FP32,
 * seen from the owner) double-buffered so the partner end re-evaluates the
 * bond from the same old state without a race.  A broken bond record keeps
 * acting through the contact law (as a reference slot with ILtype 0 does).
 * Never-bonded neighbours are held in per-particle contact rows rebuilt every
 * `rebuild_every` steps within a skin.  Platens, integration, the cut box and
 * the loading control (SumForcesULoading + CompressU/B) run on the device, so
 * an explicit step needs no host synchronisation. */

typedef struct rfd_system rfd_system;

typedef struct rfd_system_params {
  double E, nu, rho;     /* elastic modulus, Poisson ratio, density        */
  double mu, mur, e;     /* sliding friction, rolling friction, restitution */
  double dt;             /* time step                                       */
  double farcut;         /* far-link distance as a multiple of r_i + r_j    */
  double skin;           /* contact-row skin (length)                       */
  uint32_t rebuild_every;/* contact-row rebuild period in steps             */
  uint32_t contact_capacity; /* contact slots per particle                  */
  int refmode;           /* 1: reference bending-moment factor
                            (md_calculateforcesdem3.cu:148) and rolling-moment
                            wipe (:330-335); 0: physical law                */
  int fp64;              /* 1: every quantity of the step in double (the GPU
                            precision oracle of the FP32 path), explicit and
                            FIRE load steps; no batched specimens, no warm-
                            start residual query                            */
} rfd_system_params;

/* loading control, device resident; mirrors compression_data
 * (md_data_types.h:188-294) and the fracture logic of md_results.cu:34-74 */
typedef struct rfd_loading {
  int mode;              /* 0 none, 1 uniaxial (z), 2 Brazilian (y)         */
  int cut;               /* 1: park particles leaving the cut box           */
  double cx, cy, cz;     /* platen axis centre                              */
  double RR;             /* squared radius of the platen footprint          */
  double zcut;           /* particles above zcut are ignored by the platens */
  double top, bottom, top_r, bottom_r, band;
  double v, v_load;      /* current and nominal platen speed                */
  double size, size0;    /* current and initial specimen size on load axis  */
  double inv_area, min_stress, zero_stress, max_strain;
  double stress, strain, fracture_stress, fracture_strain;
  double sum_bottom, sum_top;
  int loading, fracture;
  uint32_t steps;        /* explicit steps taken                            */
  uint32_t hist_len;     /* history records written                         */
  double box_c[3], box_h[3], hidden[3];  /* cut box and parking point       */
} rfd_loading;

int rfd_system_create(rfd_ctx *ctx, const rfd_system_params *p, uint32_t N,
                      rfd_system **out);
int rfd_system_destroy(rfd_system *sys);
/* host SoA positions R[3N] (FP64) and radii[N]; V = W = 0 */
int rfd_system_set_particles(rfd_system *sys, const double *R, const float *radius);
/* GPU neighbour search: pairs (i<j) whose surface gap |rij|-ri-rj < gap, in
 * (i, j) order, written to host pairs[2*cap]; *count = number found (which
 * may exceed cap: then call again with a larger buffer) */
int rfd_system_find_pairs(rfd_system *sys, double gap, uint32_t *pairs_host,
                          uint32_t cap, uint32_t *count);
/* bonds (host arrays, sorted by (bi, bj), bi < bj); rest lengths L0 (NULL:
 * the current distance, as d_ConstructBoundInteractions sets 1/L0 from the
 * current positions, md_boundinteractions_construct.cu:65-66); state zero;
 * all active */
int rfd_system_set_bonds(rfd_system *sys, uint32_t Nb, const uint32_t *bi,
                         const uint32_t *bj, const float *br, const float *sc,
                         const float *tc, const float *L0);
/* bond state, host arrays: state[12*Nb] = n, O, Mn, Mt per bond in i's frame
 * (bond-major), flag[Nb] = 1 bond, 0 broken (contact law), 2 removed.  The
 * device keeps O, Mt and the twist Mn . n; n is the current bond direction
 * (the force pass takes the previous direction from R - V dt), reported on get
 * and ignored on set, where Mn is projected on it. */
int rfd_system_set_bond_state(rfd_system *sys, const float *state, const uint8_t *flag);
int rfd_system_get_bond_state(rfd_system *sys, float *state, uint8_t *flag);
int rfd_system_set_velocities(rfd_system *sys, const float *V, const float *W);

/* ---- the reference slot layout in and out ----
 * Device buffers of the reference's layout (md_data_types.h:12-56,124-144:
 * R, V, W SoA float[3N]; IL, ILtype, 1d_iL per slot, rows of K slots;
 * rij, Oijt, Mijn, Mijt float3 per slot), synchronous.
 * Import replaces the system's particles, bonds and contacts (N must be the
 * system's): radius, bond radius and strengths uniform (potential_data); a
 * pair i < j is a bond when slot i -> j has ILtype 1, with rest length 1/iL
 * (md_calculateforcesdem3.cu:96) and that slot's rij, Oijt, Mijn, Mijt as its
 * state (as d_ConstructBoundInteractions / d_CalculateForcesDEM_22 keep it);
 * type-0 slots with a spring or within 2 radius + skin become contact rows
 * with that spring, then a row search completes them.
 * Export writes the state back as rows of K slots: per particle its intact
 * bonds in both directions (partners ascending; the higher end's slot holds
 * the mirror image -rij, -Oijt, -Mijn, -Mijt, as the reference's own second
 * slot does), then broken bonds and contact rows as type-0 slots with their
 * springs; rij is the direction at R - V dt; free slots hold 0xFFFFFFFF.
 * A row that needs more than K slots is an argument error. */
int rfd_system_import_reference_layout(rfd_system *sys, const float *R, const float *V,
                                       const float *W, const uint32_t *IL,
                                       const uint8_t *ILtype, const float *iL, const float *rij,
                                       const float *Oijt, const float *Mijn, const float *Mijt,
                                       uint32_t N, uint32_t K, float radius, float b_r, float sc,
                                       float tc);
int rfd_system_export_reference_layout(rfd_system *sys, float *R, float *V, float *W,
                                       uint32_t *IL, uint8_t *ILtype, float *iL, float *rij,
                                       float *Oijt, float *Mijn, float *Mijt, uint32_t K);
int rfd_system_set_loading(rfd_system *sys, const rfd_loading *L);
int rfd_system_get_loading(rfd_system *sys, rfd_loading *L);
/* host copies; any pointer may be NULL */
int rfd_system_get_state(rfd_system *sys, double *R, float *V, float *W, float *F,
                         float *M);
/* the same in double: the FP64 path's own values without rounding (the state
 * a step-by-step comparison with an FP64 restatement starts each step from),
 * the FP32 path's widened */
int rfd_system_get_state64(rfd_system *sys, double *R, double *V, double *W, double *F,
                           double *M);
/* one force evaluation (bonds, broken-bond contacts, contact rows) into F, M;
 * commits the bond / contact state like a step does */
int rfd_system_compute_forces(rfd_system *sys);
/* n explicit steps: [contact-row rebuild every rebuild_every steps], forces,
 * platen + integration + cut pass, loading control.  Asynchronous. */
int rfd_system_step(rfd_system *sys, uint32_t nsteps);
/* stress-strain history (strain, bottom stress, top stress) per step */
int rfd_system_get_history(rfd_system *sys, double *hist, uint32_t cap, uint32_t *n);
/* counts of bond records by flag: out[0]=bonds, out[1]=broken, out[2]=removed */
int rfd_system_bond_counts(rfd_system *sys, uint32_t out[3]);
/* number of particles and bond records */
int rfd_system_sizes(rfd_system *sys, uint32_t *N, uint32_t *Nb);
/* entries in the current contact rows (never-bonded pairs within reach plus
 * skin, as last rebuilt); synchronises */
int rfd_system_contact_entries(rfd_system *sys, uint64_t *entries);
/* HIP-event timing of the force kernel launched by rfd_system_step /
 * rfd_system_compute_forces, recorded on the context stream: enable (also
 * clears), then read the summed kernel time in ms and the launch count */
int rfd_system_timing(rfd_system *sys, int enable);
int rfd_system_timing_read(rfd_system *sys, double *kernel_ms, uint32_t *launches);
/* ---- seeded packer ----
 * Packs the system's particles (positions as given, e.g. seeded uniform
 * random points inside the container; radii as given are the final radii)
 * into a cylinder of radius `radius` about the axis through (cx, cy) along
 * z, between z0 and z1: the radii start at s0 times their final value and
 * grow by `growth` per stage; each iteration rebuilds the contact rows and
 * moves every particle by beta/2 of each of its overlaps plus `momentum`
 * times its last move (heavy-ball relaxation), then keeps its sphere inside
 * the container (a clamped particle forgets its last move).  A stage ends when the largest overlap,
 * relative to the smaller radius of the pair, is below tol, or after
 * stage_iters iterations (final_iters at full size).  Replaces the
 * soft-sphere packing of h_CreateCylinderSample
 * (md_createcylindersample.cu:136-343, d_CalculateForces md_calculateforces.cu:20,
 * d_CylinderRestrictionZ md_cylinderrestriction.cu:74).  The system must hold
 * no bonds (set an empty bond list); velocities and loads are zero after it,
 * and the contact rows are rebuilt by the next step.  Row overflow during the
 * packing is reported in the result (and raised in the context flags). */
typedef struct rfd_pack_params {
  double cx, cy, radius, z0, z1;  /* container                                 */
  double s0, growth;              /* initial radius scale, growth per stage    */
  double beta;                    /* relaxation factor (0, 1]                  */
  double momentum;                /* share of the last move kept, [0, 1)       */
  double tol;                     /* overlap / smaller radius at convergence   */
  uint32_t stage_iters, final_iters;
} rfd_pack_params;
typedef struct rfd_pack_result {
  uint32_t iterations, stages;
  int converged, row_overflow;
  double max_overlap;             /* relative to the smaller radius            */
  double seconds;
} rfd_pack_result;
int rfd_system_pack(rfd_system *sys, const rfd_pack_params *p, rfd_pack_result *out);
/* ---- batched specimens ----
 * S independent specimens held in one system, specimen k owning particles
 * [first[k], first[k+1]) (first[0] = 0, first[S] = N), each with its own
 * loading control L[k] (all of mode 1 or all of mode 2) and a history of up to
 * hist_cap records.  A step then runs one force launch over every specimen's
 * tiles, one platen/integration launch whose workgroups each work inside one
 * specimen, and one control workgroup per specimen.  This replaces running
 * h_CalculationBrazilTest (md_calculationbraziltest.cu:113-198) or
 * h_CalculationUniaxialCompression (md_calculationuniaxialcompression.cu:113-197)
 * once per specimen, as the calibration loop does
 * (mp_determineconjgradient.cu:84-196).  Specimens must not touch each other
 * (bonds and contacts are not split by specimen).  S = 0 returns to the single
 * loading control.  FP32 path only. */
int rfd_system_set_specimens(rfd_system *sys, uint32_t S, const uint32_t *first,
                             const rfd_loading *L, uint32_t hist_cap);
/* the S loading controls (S must equal the specimen count) */
int rfd_system_get_specimen_loading(rfd_system *sys, uint32_t S, rfd_loading *L);
/* history of specimen k, as rfd_system_get_history */
int rfd_system_get_specimen_history(rfd_system *sys, uint32_t k, double *hist, uint32_t cap,
                                    uint32_t *n);
/* cell grid of the contact search (set from the particle bounds by
 * rfd_system_set_particles): origin, cell edge (>= 2.5 rmax + skin) and cell
 * counts; the cell data of cell_data::CalculateParameters (md_data_types.h:42-55) */
int rfd_system_set_grid(rfd_system *sys, const double origin[3], double cell,
                        const uint32_t dims[3]);
int rfd_system_get_grid(rfd_system *sys, double origin[3], double *cell, uint32_t dims[3]);
/* as rfd_system_step, with the same results bit for bit, but every run of
 * 2 x rebuild_every steps that starts on a rebuild step is one replay of a HIP
 * graph captured on first use (and again after any change of particles, bonds,
 * loading, specimens, grid or launch shapes); other steps run as launches.
 * The replay runs on a private stream ordered after and before the context
 * stream's work.  Steps run as launches while force-kernel timing is on. */
int rfd_system_step_graph(rfd_system *sys, uint32_t nsteps);
/* ---- launch-shape autotuner ----
 * Shapes: force-pass tile of 32, 64 (default) or 128 particles per workgroup,
 * with or without single evaluation (physical mode: a bond whose ends share a
 * tile is evaluated once and loads both ends; used where the tile's half-bonds
 * fit the workgroup's LDS); contact-row search by 2 x 2 x 2 cell blocks with
 * the surrounding cells staged in LDS (0, default), or one thread per particle
 * in blocks of 64, 128 or 256.  Results do not depend on the shapes.
 * rfd_system_autotune times each shape (median of 5 trial force passes, which
 * write only F, M and the spare state buffer; row searches into the spare row
 * buffers) and applies the fastest; the choice is kept per (device
 * architecture, floor(log2 N), batched or not) in a process table, which
 * use_cache = 1 consults first and rfd_tune_cache_save/load write to and read
 * from a text file.  F and M hold the forces of the current configuration
 * afterwards; the trajectory is unchanged. */
typedef struct rfd_tune_result {
  int force_tile, nbr_block;    /* chosen shapes                              */
  int single_eval;              /* 1: in-tile bonds evaluated once            */
  int cached;                   /* 1: taken from the table, nothing timed     */
  int bucket;                   /* floor(log2 N)                              */
  double force_ms[4];           /* tiles 64, 32, 128 single evaluation; 64 every half */
  double nbr_ms[4];             /* thread blocks 128, 64, 256; cell blocks (0) */
} rfd_tune_result;
int rfd_system_autotune(rfd_system *sys, int use_cache, rfd_tune_result *out);
int rfd_system_set_tiles(rfd_system *sys, int force_tile, int nbr_block, int single_eval);
/* the shapes in use; single_eval reports whether the single-evaluation force
 * pass runs (physical mode, FP32 path, tile halves within its LDS) */
int rfd_system_get_tiles(rfd_system *sys, int *force_tile, int *nbr_block, int *single_eval);

/* ---- fused explicit steps ----
 * Replaces the separate d_CalculateIncrementsDEM / platen launches after the
 * force kernels (md_calculationuniaxialcompression.cu:173-184): with one
 * specimen in the FP32 physical mode and the single-evaluation force pass,
 * rfd_system_step runs the platens, integration and cut box of each particle
 * inside the force pass, pairs of steps alternating between two particle
 * buffers.  On by default; the results are the same bits with it off.  A
 * system of at most 256 force tiles also runs the loading control
 * (SumForcesULoading, md_results.cu:34) in the force pass, by its last
 * workgroup, so that each step is one launch (same bits).  An FP64 system
 * whose force pass has at most 256 workgroups runs each step as one launch
 * too; there the platen reactions are summed per workgroup of that launch, so
 * the reaction sums (and the history) may differ from the separate pass's in
 * their last bits, never the particle state.  Fusion also moves the FIRE
 * measure of rfd_system_fire_load_step (platens at rest, power, norms,
 * reactions) into the epilogue of the quasi-static force pass; there only the
 * order of the FP64 partial sums differs.
 * rfd_system_get_fusion reports whether steps currently run fused. */
int rfd_system_set_fusion(rfd_system *sys, int on);
int rfd_system_get_fusion(rfd_system *sys, int *fused);
/* Multi-step launches.  Replaces the per-step launch sequence of the host
 * loops (md_calculationuniaxialcompression.cu:169-189,
 * md_calculationbraziltest.cu:169-190) for small specimens and batches, where
 * launch latency bounds a step: with it on, rfd_system_step runs the fused
 * steps between two contact-row rebuilds as one cooperative launch with a
 * grid barrier per step and the loading control inside.  It applies when
 * steps run fused (rfd_system_get_fusion) and every force tile gets a
 * resident workgroup (a few tens of thousands of particles), and not under
 * graph capture; otherwise steps run as before.  Off by default; the results
 * are the same bits either way.  rfd_system_get_persistent reports the
 * setting and whether the current system qualifies. */
int rfd_system_set_persistent(rfd_system *sys, int on);
int rfd_system_get_persistent(rfd_system *sys, int *on, int *applies);
int rfd_tune_cache_save(const char *path);
int rfd_tune_cache_load(const char *path);
/* ---- quasi-static load steps: axial strain increment + FIRE relaxation ----
 * Standard FIRE (velocity mixing, adaptive dt, half-step back-off on uphill
 * steps) over translations and rotations, run on the device; the reference's
 * FIRE (md_firerelaxation.cu:18-84,263-304) is translation-only and packing
 * only.  Platen particles are held by the displacement boundary condition of
 * the explicit platens with zero platen speed. */
typedef struct rfd_fire_params {
  double ftol;          /* converged when max|F_i| and max|M_i|/r_i < ftol   */
  double dt0, dt_max;   /* initial / maximum FIRE time step                   */
  double alpha0, f_inc, f_dec, f_alpha;
  uint32_t n_delay;     /* positive steps before dt may grow                  */
  uint32_t max_iter;    /* iteration cap per load step                        */
  uint32_t check_every; /* host convergence poll period (iterations)          */
} rfd_fire_params;

typedef struct rfd_fire_result {
  uint32_t iterations;
  int converged;
  double fmax0, fmax;   /* residual max|F_i| before / after relaxation        */
  double mmax;          /* final max|M_i|/r_i                                 */
  double sum_bottom, sum_top; /* platen reactions after relaxation            */
  double strain;        /* specimen strain after the step                     */
  double seconds;       /* wall time of the load step                         */
  int warm_used;        /* 1: the predicted start was taken                   */
  double fmax0_plain, fmax0_pred; /* residuals of the two candidate starts
                                     (-1 when no prediction was given)        */
  uint32_t breaks;      /* bonds broken at equilibrium during the step        */
  uint32_t break_rounds;/* relax-and-break rounds after the first relaxation  */
} rfd_fire_result;

/* One quasi-static load step: move the top platen by dstrain * size0 (load
 * axis of the loading mode; dstrain < 0 compresses; particles held by a
 * platen follow it), optionally apply a predicted increment (warm start: dU,
 * dTh are device arrays of N x 3 floats, row-major, or NULL), carry the move
 * into the incremental springs with one force pass, then relax with FIRE
 * (from rest) to ftol with bonds held intact; then bonds are judged at the
 * relaxed state (break criterion of md_calculateforcesdem3.cu:20-36), and
 * while any breaks the step relaxes again.  A prediction is judged by FIRE's
 * residual: the start
 * it gives is taken only if its max|F| is not above the plain start's, else
 * the step runs from the plain start (fallback).  The displacement of the
 * step is kept on the device (rfd_system_fire_buffers). */
int rfd_system_fire_load_step(rfd_system *sys, double dstrain, const rfd_fire_params *fp,
                              const float *dU, const float *dTh, rfd_fire_result *out);
/* residual of the current configuration (one force pass + constraint):
 * out[0] = max|F_i|, out[1] = max|M_i|/r_i over free DOFs */
int rfd_system_residual(rfd_system *sys, double out[2]);
/* device arrays (N x 3 floats, row-major) of the last two load-step
 * displacements and rotations, newest first; valid after a load step */
int rfd_system_fire_buffers(rfd_system *sys, float **du_last, float **du_prev,
                            float **dth_last, float **dth_prev);
/* bond graph on the device, for callers that build per-bond features in HBM:
 * hoff[N+1] row offsets; half[h] = {j, b} (uint32 pairs: other end, bond);
 * bond[b] = {1/L0, br, sc, tc} (floats); state[b] = {O.xyz, mn, Mt.xyz, flag}
 * (7 floats + uint32 flag: 1 bond, 0 broken, 2 removed), the latest state;
 * *H = number of halves (2 Nb).  Rows list lower partners first, each row by
 * partner id. */
int rfd_system_bond_graph(rfd_system *sys, const uint32_t **hoff, const void **half,
                          const void **bond, const void **state, uint32_t *H);

/* device pointers of the particle arrays (for callers that keep data in HBM):
 * P is N records of 4 doubles {x, y, z, radius}; V, W, F, M are N records of
 * 4 floats {x, y, z, 0} */
int rfd_system_device_arrays(rfd_system *sys, double **P, float **V, float **W,
                             float **F, float **M);

/* ============ learned warm start: message-passing network kernels ========
 * No counterpart in the reference (the warm start is new); these are the
 * device kernels a per-particle message-passing network over the bond graph
 * needs for inference and training.  All arrays are FP32 device memory,
 * row-major with the given leading dimensions; dense products round their
 * operands to BF16 and accumulate in FP32 on the matrix cores; every
 * reduction has a fixed order (deterministic). */

/* mirror half of each half-bond (hm[h] = the half of the same bond in the
 * other end's row), H entries */
int rfd_system_half_mirror(rfd_system *sys, const uint32_t **hm);
/* network inputs at the start of a load step: X[N x 32] = last two step
 * displacements / su and rotations / sth (12 columns), then the utilisation of
 * the particle's intact bonds, E strain / tensile strength (largest, mean)
 * and the share of its bond records broken (3 columns; the rest 0);
 * E[H x 4] per half = (1e3 * bond strain if intact else 0, unit direction to
 * the other end) */
int rfd_system_mpnn_inputs(rfd_system *sys, double su, double sth, float *X, float *E);

/* Y[M x N] = X[M x K] B[K x N] + bias (+ Y if accumulate), B(k, c) = W[k*wsk + c*wsc]
 * (wsk = N, wsc = 1 for row-major W; wsk = 1, wsc = ldw for its transpose);
 * K a multiple of 32 up to 128, N a multiple of 16 up to 128; bias may be NULL */
int rfd_mm_bf16(rfd_ctx *ctx, const float *X, uint32_t ldx, const float *W, uint32_t wsk,
                uint32_t wsc, const float *bias, float *Y, uint32_t ldy, uint32_t M,
                uint32_t K, uint32_t N, int accumulate);
/* G[K x N] (+)= X[M x K]^T D[M x N]; K a multiple of 16 up to 128, N of 16 up to 128 */
int rfd_mm_tn_bf16(rfd_ctx *ctx, const float *X, uint32_t ldx, const float *D, uint32_t ldd,
                   float *G, uint32_t M, uint32_t K, uint32_t N, int accumulate);
/* g[c] (+)= sum over rows of D[m][c], c < ncols */
int rfd_colsum(rfd_ctx *ctx, const float *D, uint32_t ldd, float *g, uint32_t M,
               uint32_t ncols, int accumulate);
/* elementwise over M x ncols: op 0 out = relu(a); 1 out = b + relu(a);
 * 2 out = b > 0 ? a : 0; 3 out = a + b; 4 out = a; 5 out = alpha (a + b);
 * 6 out = alpha a - b; 7 out = alpha a (out may alias a or b) */
int rfd_mpnn_ewise(rfd_ctx *ctx, int op, uint32_t M, uint32_t ncols, float alpha,
                   const float *a, uint32_t lda, const float *b, uint32_t ldb, float *out,
                   uint32_t ldo);
/* message step over half-bond rows (hoff[N+1], partner of half h at
 * hj[h * hj_stride]): q_h = Z[j] + E[h] We + bm (hidden width 64, We 4 x 64),
 * A[i] = sum over row i of relu(q_h) (row stride lda); Q[H x 64] keeps q when
 * not NULL (training) */
int rfd_mpnn_edge_fwd(rfd_ctx *ctx, uint32_t N, const uint32_t *hoff, const uint32_t *hj,
                      uint32_t hj_stride, const float *E, const float *We, const float *bm,
                      const float *Z, float *Q, float *A, uint32_t lda);
/* its backward: dZ[N x 64], dWe[4 x 64] and dbm[64] (+= if accumulate) from dA */
int rfd_mpnn_edge_bwd(rfd_ctx *ctx, uint32_t N, const uint32_t *hoff, const uint32_t *hj,
                      uint32_t hj_stride, const uint32_t *hm, const float *E, const float *Q,
                      const float *dA, uint32_t ldda, float *dZ, float *dWe, float *dbm,
                      int accumulate);
/* mean of (Y - T)^2 over N x ncols (FP64 sum, synchronises); dY = its gradient
 * (may be NULL) */
int rfd_mse_loss(rfd_ctx *ctx, const float *Y, uint32_t ldy, const float *T, uint32_t ldt,
                 uint32_t N, uint32_t ncols, float *dY, uint32_t lddy, double *loss);
/* one Adam update of n parameters at step t >= 1 */
int rfd_adam(rfd_ctx *ctx, float *p, const float *g, float *m, float *v, uint32_t n, float lr,
             float b1, float b2, float eps, uint32_t t);

#ifdef __cplusplus
}
#endif
#endif
