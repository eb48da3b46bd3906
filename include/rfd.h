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
 *   - Kernels never allocate caller memory; the context owns its scratch.
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

/* ========================= tier 1: reference layout ===================== */

/* CellDistribution (md_cell_distribution.cu:52-101): d_CalculateCellIndex
 * (md_calculate_cell.cu:28), CUB SortPairs (:68,:87), d_DetermineCellPointer
 * (md_calculate_cell.cu:55).  CI[N]; CIs[2N] = sorted cell keys | particle
 * ids (stable); pnC[2C] = start | count, zero for empty cells. */
int rfd_cell_distribution(rfd_ctx *ctx, const float *R, uint32_t N, float inv_a,
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
 * quirk_p0 != 0 keeps the reference's particle-0 free-slot rule (:32). */
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

#ifdef __cplusplus
}
#endif
#endif
