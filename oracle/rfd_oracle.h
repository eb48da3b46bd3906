/*
 * rfd_oracle.h -- CPU oracle for the RockFit-DEM bonded-particle-model step.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is what the HIP path is compared
 * against.  Only tests/, __graft_entry__.smoke() and the cpu_baseline
 * leg of bench.py may load it; the product library (librfd.so) never links,
 * loads or calls it.
 *
 * Layout is the reference's "slot" layout (md_data_types.h:12-56,124-144):
 *   particle arrays SoA by component  X[i], X[i+N], X[i+2N]   (float)
 *   slot arrays     row-major, K slots per particle, AoS float3 per slot
 *   IL sentinel     0xFFFFFFFF
 * Arithmetic is FP32 exactly as the reference kernels write it, evaluated
 * left to right with no FMA contraction (the library is built with
 * -ffp-contract=off).  norm3df is computed in double and rounded once.
 */
#ifndef RFD_ORACLE_H
#define RFD_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_SENTINEL 0xFFFFFFFFu

/* status codes */
#define ORC_OK 0
#define ORC_ERR_CELL_RANGE (-10)   /* particle or neighbour cell outside the grid */
#define ORC_ERR_ROW_FULL (-11)     /* more than K links for one particle */

/* a3+a4: cell index per particle, stable sort of (cell, id), cell start/count */
int orc_cell_distribution(const float *R, uint32_t N, float inv_a, uint32_t Nx,
                          uint32_t Ny, uint32_t Nz, uint32_t *CI, uint32_t *CIs,
                          uint32_t *pnC);
/* a5: fresh neighbour rows (IL prefilled with the sentinel by this call) */
int orc_il_construct(const float *R, uint32_t N, const uint32_t *CIs,
                     const uint32_t *pnC, uint32_t *IL, uint32_t K, float a,
                     float inv_a, float aacut, uint32_t Nx, uint32_t Ny,
                     uint32_t Nz);
/* a7: insert new contacts into free slots; quirk_p0 reproduces the particle-0
 * free-slot rule of the reference */
int orc_il_reconstruct_bpm(const float *R, uint32_t N, const uint32_t *CIs,
                           const uint32_t *pnC, uint32_t *IL, uint8_t *ILtype,
                           uint32_t K, float a, float inv_a, float aacut,
                           uint32_t Nx, uint32_t Ny, uint32_t Nz, int quirk_p0);
/* a6 */
void orc_bonds_construct(const float *R, const uint32_t *IL, uint8_t *ILtype,
                         float *iL, float *rij, float *Oijt, float *Fijn,
                         float *Fijt, float *Mijn, float *Mijt, float *Mijadd,
                         uint32_t N, uint32_t K, float bdd);
/* a8 */
void orc_delete_far_links(const float *R, uint32_t N, uint32_t *IL,
                          uint8_t *ILtype, float *rij, float *Oijt, float *Fijn,
                          float *Fijt, float *Mijn, float *Mijt, float *Mijadd,
                          uint32_t K, float aa_farcut);
/* a9+a11+a12: bond force, moments, fused break check, F/M += per particle */
void orc_forces_bpm(const float *R, const float *V, const float *W,
                    const uint32_t *IL, uint8_t *ILtype, const float *iL,
                    float *rij, float *Oijt, float *Mijn, float *Mijt, float *F,
                    float *M, uint32_t N, uint32_t K, float dt, float E, float G,
                    float b_r, float sc, float tc);
/* a10: Hertz-Mindlin contact, F/M += per particle */
void orc_forces_contact(const float *R, const float *V, const float *W,
                        const uint32_t *IL, const uint8_t *ILtype, float *Oijt,
                        float *F, float *M, uint32_t N, uint32_t K, float dt,
                        float Estar, float Gstar, float pA, float mu, float mur,
                        float m, float r);
/* a11 stand-alone: F = sum(Fijn+Fijt), M = sum(Mijn+Mijt+Mijadd) per row */
void orc_sumup_forces(const float *Fijn, const float *Fijt, const float *Mijn,
                      const float *Mijt, const float *Mijadd, float *F, float *M,
                      uint32_t N, uint32_t K);
/* f4 superseded per-slot force path (md_calculateforcesdem.cu:328, :501,
 * :581 by mode 0, 1, 2) and penalty platens
 * (md_uniaxialcompression_simple.cu:29) */
void orc_forces_slots(int mode, const float *R, const float *V, const float *W,
                      const uint32_t *IL, const uint8_t *ILtype, const float *iL, float *rij,
                      float *Oijt, float *Fijn, float *Fijt, float *Mijn, float *Mijt,
                      float *Mijadd, uint32_t N, uint32_t K, float dt, float E, float G,
                      float b_r, float pA, float mu, float mur, float mP, float rP);
void orc_platen_uniaxial_penalty(const float *R, const float *V, float *F, uint32_t N,
                                 float cx, float cy, float cz, float RR, float Hd2, float C,
                                 float mu, float Zcut, double sums[2]);
/* a12 stand-alone per-slot break check */
void orc_check_break(const float *rij, float *Oijt, const float *Fijn,
                     const float *Fijt, const float *Mijn, const float *Mijt,
                     uint8_t *ILtype, uint32_t S, float sc, float tc, float b_r);
/* a13 */
void orc_increments(const float *F, float *V, float *R, const float *M, float *W,
                    uint32_t N, float dt_m, float dt, float dt_I);
/* a15 */
void orc_platen_uniaxial(float *R, float *V, float *F, uint32_t N, float cx,
                         float cy, float RR, float Zt, float Zb, float Ztr,
                         float Zbr, float vt, float Zcut, double sums[2]);
void orc_platen_brazil(float *R, float *V, float *F, uint32_t N, float cx,
                       float cz, float RR, float Yt, float Yb, float Ytr,
                       float Ybr, float vt, float Zcut, double sums[2]);
void orc_cut_restriction(float *R, float *V, float *F, uint32_t N,
                         const float center[3], const float sized2[3],
                         const float hp[3]);
/* a14 reference FIRE kernels */
void orc_fire_increments(const float *F, float *V, float *R, uint32_t N,
                         float dt_m, float dt, float alpha);
void orc_fire_halfstep(const float *V, float *R, uint32_t N, float dt_d2);
/* out[0]=sum F.V, out[1]=sum |F|^2, out[2]=sum |V|^2, out[3]=max |F|^2
 * (translation);  W/M may be NULL, else rotational terms are included too */
void orc_fire_power(const float *F, const float *V, const float *M,
                    const float *W, uint32_t N, double out[4]);

/* correctly rounded helpers shared by every function above (exported so that
 * tests can check them against the mathematics directly) */
float orc_norm3(float x, float y, float z);
float orc_rsqrt(float x);

/* ---------------- native (edge-list) physics, FP64 precision oracle --------
 * One record per undirected bond (i<j).  Per-particle radius r, mass m and
 * inertia I; per-bond radius br, 1/L0, tensile/shear strength sc/tc.
 * state[8*b] = O(3), mn, Mt(3), unused: shear spring, twist along n and
 * bending moment as seen from end i; the previous bond direction is taken
 * at R - V dt.
 * Forces are accumulated into F,M (double, SoA by component, zeroed here).
 * Reference mode (refmode=1) keeps the extra bending factor of
 * md_calculateforcesdem3.cu:148; refmode=0 applies Mt*pi*br^4 only.       */
void orc_native_bonds_f64(const double *R, const double *V, const double *W,
                          uint32_t N, const uint32_t *bi, const uint32_t *bj,
                          const double *iL, const double *br, const double *sc,
                          const double *tc, double *state, uint8_t *active,
                          uint32_t Nb, double dt, double E, double G,
                          int refmode, double *F, double *M);

/* Native-tier contact law per directed pair (i sees j), double precision:
 * pairs (pi[k], pj[k]) overlap below reach[k]; O[3k] is the pair's shear
 * spring (updated); F, M (zeroed here) receive i's load.  Per-particle radii
 * rad[N]; masses rho43pi r^3.  cap (may be NULL) receives per pair the
 * Coulomb test's relative margin |Ft|^2 / (mu fn)^2 - 1 (0 for a pair not in
 * contact): the law switches from the damped spring to the capped force at
 * 0, so a margin near 0 is decided by rounding. */
void orc_native_contacts_f64(const double *R, const double *V, const double *W,
                             const double *rad, uint32_t N, const uint32_t *pi,
                             const uint32_t *pj, const double *reach, double *O,
                             uint32_t P, double dt, double Estar, double Gstar,
                             double pA, double mu, double mur, double rho43pi,
                             double *F, double *M, double *cap);

#ifdef __cplusplus
}
#endif
/* threads of the per-particle loops (1, or the OpenMP team of the -fopenmp
 * build, librfd_oracle_omp.so) */
int orc_threads(void);

#endif
