/*
 * rfd_oracle.c -- CPU restatement of the RockFit-DEM BPM step.
 *
 * TEST INFRASTRUCTURE ONLY: what the HIP path is compared against (see
 * rfd_oracle.h).
 * Built with -O2 -ffp-contract=off -fno-fast-math so that every FP32
 * expression below rounds after each operation, left to right, exactly as
 * written.  Each function cites the reference file:line it restates.
 *
 * CUDA intrinsics are matched from IEEE semantics, not from CUDA's code:
 *   __fmul_rn(a,b)   -> a*b              (no contraction anyway)
 *   __frcp_rn(x)     -> 1.0f/x           (IEEE division, correctly rounded)
 *   __fsqrt_rn(x)    -> sqrtf(x)         (IEEE sqrt, correctly rounded)
 *   __frsqrt_rn(x)   -> (float)(1/sqrt((double)x))   (rounded once more)
 *   norm3df(x,y,z)   -> (float)sqrt(x*x+y*y+z*z in double)
 * Comparisons of a float against a double literal (e.g. rr > 1e-10) are
 * done in double, as C/C++ promotion does in the reference.
 */
#include "rfd_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define PI_F 3.14159265358979f /* MCf_pi, md_math_constants.h:87 */
#define TWO_THIRDS_F 0.6666666667f /* MCf_2d3, md_math_constants.h:63 */

float orc_norm3(float x, float y, float z) {
  double dx = x, dy = y, dz = z;
  return (float)sqrt(dx * dx + dy * dy + dz * dz);
}

float orc_rsqrt(float x) { return (float)(1.0 / sqrt((double)x)); }

/* ------------------------------------------------------------------ cells */

/* md_calculate_cell.cu:28-53: c = floor(x*ia) + floor(y*ia)*Nx + floor(z*ia)*Nx*Ny */
static int cell_of(const float *R, uint32_t N, uint32_t i, float inv_a,
                   uint32_t Nx, uint32_t Ny, uint32_t Nz, uint32_t *c,
                   uint32_t ix[3]) {
  for (int d = 0; d < 3; ++d) {
    float f = floorf(R[i + (size_t)d * N] * inv_a);
    if (!(f >= 0.0f)) return ORC_ERR_CELL_RANGE;
    ix[d] = (uint32_t)f;
  }
  if (ix[0] >= Nx || ix[1] >= Ny || ix[2] >= Nz) return ORC_ERR_CELL_RANGE;
  *c = ix[0] + ix[1] * Nx + ix[2] * Ny * Nx;
  return ORC_OK;
}

/* md_cell_distribution.cu:52-101 (+ md_calculate_cell.cu:13,55): stable sort
 * of (CI, i) keys -> CIs[0..N) keys, CIs[N..2N) ids; pnC zeroed then
 * pnC[c] = first sorted position, pnC[c+C] = count for each occupied cell.
 * A counting sort in index order is a stable sort, as CUB's radix sort is. */
int orc_cell_distribution(const float *R, uint32_t N, float inv_a, uint32_t Nx,
                          uint32_t Ny, uint32_t Nz, uint32_t *CI, uint32_t *CIs,
                          uint32_t *pnC) {
  uint32_t C = Nx * Ny * Nz;
  for (uint32_t i = 0; i < N; ++i) {
    uint32_t ix[3];
    int st = cell_of(R, N, i, inv_a, Nx, Ny, Nz, &CI[i], ix);
    if (st) return st;
  }
  memset(pnC, 0, sizeof(uint32_t) * 2 * (size_t)C);
  uint32_t *cnt = pnC + C;
  for (uint32_t i = 0; i < N; ++i) cnt[CI[i]]++;
  uint32_t *start = (uint32_t *)malloc(sizeof(uint32_t) * ((size_t)C + 1));
  uint32_t acc = 0;
  for (uint32_t c = 0; c < C; ++c) {
    start[c] = acc;
    acc += cnt[c];
  }
  for (uint32_t c = 0; c < C; ++c) pnC[c] = cnt[c] ? start[c] : 0u;
  for (uint32_t i = 0; i < N; ++i) {
    uint32_t p = start[CI[i]]++;
    CIs[p] = CI[i];
    CIs[N + p] = i;
  }
  free(start);
  return ORC_OK;
}

/* side sign of md_interaction_list_construct.cu:54-56 */
static int side(float x, uint32_t ix, float a) {
  float mid = ((float)ix + 0.5f) * a;
  return (int)copysignf(1.0f, x - mid);
}

/* the 8 cells searched, in the reference's order (construct.cu:59-85).
 * A neighbour cell outside the grid is reported as empty (UINT32_MAX): the
 * reference reads past its cell table there (e.g. for particles parked at the
 * hidden point in the top cell layer, md_data_types.h:116). */
static int neighbour_cells(const float *R, uint32_t N, uint32_t i, float a,
                           float inv_a, uint32_t Nx, uint32_t Ny, uint32_t Nz,
                           uint32_t out[8]) {
  uint32_t ix[3], c;
  int st = cell_of(R, N, i, inv_a, Nx, Ny, Nz, &c, ix);
  if (st) return st;
  int sg[3];
  sg[0] = side(R[i], ix[0], a);
  sg[1] = side(R[i + N], ix[1], a);
  sg[2] = side(R[i + 2 * (size_t)N], ix[2], a);
  const uint32_t dim[3] = {Nx, Ny, Nz};
  /* bit d of the mask selects a step along axis d */
  static const int masks[8] = {0, 1, 2, 4, 3, 5, 6, 7};
  for (int q = 0; q < 8; ++q) {
    int64_t p[3];
    int ok = 1;
    for (int d = 0; d < 3; ++d) {
      p[d] = (int64_t)ix[d] + ((masks[q] >> d) & 1 ? sg[d] : 0);
      if (p[d] < 0 || p[d] >= (int64_t)dim[d]) ok = 0;
    }
    out[q] = ok ? (uint32_t)(p[0] + p[1] * Nx + p[2] * (int64_t)Nx * Ny) : 0xFFFFFFFFu;
  }
  return ORC_OK;
}

static float dist2(const float *R, uint32_t N, uint32_t i, uint32_t j) {
  float dx = R[j] - R[i];
  float dy = R[j + N] - R[i + N];
  float dz = R[j + 2 * (size_t)N] - R[i + 2 * (size_t)N];
  return dx * dx + dy * dy + dz * dz;
}

/* md_interaction_list_construct.cu:13-98 + construction.cu:62-67 */
int orc_il_construct(const float *R, uint32_t N, const uint32_t *CIs,
                     const uint32_t *pnC, uint32_t *IL, uint32_t K, float a,
                     float inv_a, float aacut, uint32_t Nx, uint32_t Ny,
                     uint32_t Nz) {
  uint32_t C = Nx * Ny * Nz;
  int status = ORC_OK;
  int cell_status = ORC_OK;
  for (size_t s = 0; s < (size_t)N * K; ++s) IL[s] = ORC_SENTINEL;
#pragma omp parallel for schedule(static)
  for (uint32_t i = 0; i < N; ++i) {
    uint32_t cells[8];
    int st = neighbour_cells(R, N, i, a, inv_a, Nx, Ny, Nz, cells);
    if (st) {
#pragma omp atomic write
      cell_status = st;
      continue;
    }
    uint32_t nk = 0;
    for (int q = 0; q < 8; ++q) {
      uint32_t c = cells[q];
      if (c == 0xFFFFFFFFu) continue;
      uint32_t s0 = pnC[c], n = pnC[c + C];
      for (uint32_t t = 0; t < n; ++t) {
        uint32_t j = CIs[N + s0 + t];
        float rr = dist2(R, N, i, j);
        if (rr < aacut && j != i) {
          if (nk < K) {
            IL[(size_t)i * K + nk] = j;
          } else {
#pragma omp atomic write
            status = ORC_ERR_ROW_FULL;
          }
          ++nk;
        }
      }
    }
  }
  return cell_status ? cell_status : status;
}

/* md_interaction_list_reconstructbpm.cu:13-46, 80-139.  New contacts go to
 * the first free slot with ILtype=0; existing links are never removed.
 * quirk_p0: for particle 0 a free slot 0 only counts when no later slot is
 * free (the kdxfree==0 test of :32).  A full row is reported instead of
 * overwriting absolute slot 0 as the reference would. */
int orc_il_reconstruct_bpm(const float *R, uint32_t N, const uint32_t *CIs,
                           const uint32_t *pnC, uint32_t *IL, uint8_t *ILtype,
                           uint32_t K, float a, float inv_a, float aacut,
                           uint32_t Nx, uint32_t Ny, uint32_t Nz, int quirk_p0) {
  uint32_t C = Nx * Ny * Nz;
  int status = ORC_OK, cell_status = ORC_OK;
#pragma omp parallel for schedule(static)
  for (uint32_t i = 0; i < N; ++i) {
    uint32_t cells[8];
    int st = neighbour_cells(R, N, i, a, inv_a, Nx, Ny, Nz, cells);
    if (st) {
#pragma omp atomic write
      cell_status = st;
      continue;
    }
    uint32_t *row = IL + (size_t)i * K;
    for (int q = 0; q < 8; ++q) {
      uint32_t c = cells[q];
      if (c == 0xFFFFFFFFu) continue;
      uint32_t s0 = pnC[c], n = pnC[c + C];
      for (uint32_t t = 0; t < n; ++t) {
        uint32_t j = CIs[N + s0 + t];
        float rr = dist2(R, N, i, j);
        if (rr < aacut && (double)rr > 1e-10) {
          int found = 0;
          int64_t freeslot = -1;
          for (uint32_t k = 0; k < K; ++k) {
            if (row[k] == j) {
              found = 1;
              break;
            }
            if (row[k] == ORC_SENTINEL) {
              if (freeslot < 0)
                freeslot = k;
              else if (quirk_p0 && i == 0 && freeslot == 0)
                freeslot = k;
            }
          }
          if (!found) {
            if (freeslot < 0) {
#pragma omp atomic write
              status = ORC_ERR_ROW_FULL;
            } else {
              row[freeslot] = j;
              ILtype[(size_t)i * K + freeslot] = 0;
            }
          }
        }
      }
    }
  }
  return cell_status ? cell_status : status;
}

static void zero3(float *p, size_t s) { p[3 * s] = p[3 * s + 1] = p[3 * s + 2] = 0.0f; }

/* md_boundinteractions_construct.cu:35-106.  rm2 = x*x+y*y+z*z (no
 * contraction); rij is overwritten by 0*r at :78-80 (signed zeros kept). */
void orc_bonds_construct(const float *R, const uint32_t *IL, uint8_t *ILtype,
                         float *iL, float *rij, float *Oijt, float *Fijn,
                         float *Fijt, float *Mijn, float *Mijt, float *Mijadd,
                         uint32_t N, uint32_t K, float bdd) {
#pragma omp parallel for schedule(static)
  for (uint32_t i = 0; i < N; ++i) {
    for (uint32_t k = 0; k < K; ++k) {
      size_t s = (size_t)i * K + k;
      uint32_t j = IL[s];
      if (j == ORC_SENTINEL) continue;
      float rx = R[j] - R[i];
      float ry = R[j + N] - R[i + N];
      float rz = R[j + 2 * (size_t)N] - R[i + 2 * (size_t)N];
      float rm2 = rx * rx + ry * ry + rz * rz;
      if (rm2 < bdd && (double)rm2 > 1e-18) {
        iL[s] = orc_rsqrt(rm2);
        ILtype[s] = 1;
      } else {
        ILtype[s] = 0;
      }
      rij[3 * s] = 0.0f * rx;
      rij[3 * s + 1] = 0.0f * ry;
      rij[3 * s + 2] = 0.0f * rz;
      zero3(Oijt, s);
      zero3(Fijn, s);
      zero3(Fijt, s);
      zero3(Mijn, s);
      zero3(Mijt, s);
      zero3(Mijadd, s);
    }
  }
}

/* md_interaction_list_reconstructbpm.cu:141-186 (1/L0 is left as it is: the
 * reference's `_1d_iL = 0;` at :173 clears a local pointer, not the array) */
void orc_delete_far_links(const float *R, uint32_t N, uint32_t *IL,
                          uint8_t *ILtype, float *rij, float *Oijt, float *Fijn,
                          float *Fijt, float *Mijn, float *Mijt, float *Mijadd,
                          uint32_t K, float aa_farcut) {
#pragma omp parallel for schedule(static)
  for (uint32_t i = 0; i < N; ++i) {
    for (uint32_t k = 0; k < K; ++k) {
      size_t s = (size_t)i * K + k;
      uint32_t j = IL[s];
      if (j == ORC_SENTINEL) continue;
      float rr = dist2(R, N, i, j);
      if (rr > aa_farcut || (double)rr < 1e-10) {
        IL[s] = ORC_SENTINEL;
        ILtype[s] = 0;
        zero3(rij, s);
        zero3(Oijt, s);
        zero3(Fijn, s);
        zero3(Fijt, s);
        zero3(Mijn, s);
        zero3(Mijt, s);
        zero3(Mijadd, s);
      }
    }
  }
}

/* md_calculateforcesdem3.cu:424-483 with the dd_*_2 helpers at :20-197.
 * Accumulation order per particle: slot order; per slot Fn, then Ft into fs
 * and Mn, Mt, 0.5 r x Ft into ms. */
void orc_forces_bpm(const float *R, const float *V, const float *W,
                    const uint32_t *IL, uint8_t *ILtype, const float *iL,
                    float *rij, float *Oijt, float *Mijn, float *Mijt, float *F,
                    float *M, uint32_t N, uint32_t K, float dt, float E, float G,
                    float b_r, float sc, float tc) {
  const size_t N1 = N, N2 = 2 * (size_t)N;
#pragma omp parallel for schedule(static)
  for (uint32_t i = 0; i < N; ++i) {
    float fsx = 0, fsy = 0, fsz = 0, msx = 0, msy = 0, msz = 0;
    float Sc = 0, Mc = 0;
    for (uint32_t k = 0; k < K; ++k) {
      size_t s = (size_t)i * K + k;
      uint32_t j = IL[s];
      if (j == ORC_SENTINEL || ILtype[s] != 1) continue;
      /* dd_Calculate_rijm_nij_2 :38-52 */
      float rx = R[j] - R[i], ry = R[j + N1] - R[i + N1], rz = R[j + N2] - R[i + N2];
      float rm = orc_norm3(rx, ry, rz);
      float inv_rm = 1.0f / rm;
      float nx = rx * inv_rm, ny = ry * inv_rm, nz = rz * inv_rm;
      float il = iL[s];
      /* dd_Calculate_fijn_bpm_2 :91-103 */
      float cef = E * (rm - 1.0f / il) * il;
      Sc = cef;
      cef = cef * (b_r * b_r * PI_F);
      fsx = fsx + nx * cef;
      fsy = fsy + ny * cef;
      fsz = fsz + nz * cef;
      /* dd_Calculate_M123_bpm_2 :54-89 */
      float px0 = rij[3 * s], py0 = rij[3 * s + 1], pz0 = rij[3 * s + 2];
      float tx = py0 * nz - pz0 * ny;
      float ty = pz0 * nx - px0 * nz;
      float tz = px0 * ny - py0 * nx;
      float wsx = W[j] + W[i], wsy = W[j + N1] + W[i + N1], wsz = W[j + N2] + W[i + N2];
      float c2 = (wsx * nx + wsy * ny + wsz * nz) * dt * 0.5f;
      float phx = nx * c2, phy = ny * c2, phz = nz * c2;
      rij[3 * s] = nx;
      rij[3 * s + 1] = ny;
      rij[3 * s + 2] = nz;
      float M1x = 1.0f + tz * phz + ty * phy;
      float M1y = phz - tz - ty * phx;
      float M1z = -phy - tz * phx + ty;
      float M2x = tz - phz - tx * phy;
      float M2y = tz * phz + 1.0f + tx * phx;
      float M2z = -tz * phy + phx - tx;
      float M3x = -ty - tx * phz + phy;
      float M3y = -ty * phz + tx - phx;
      float M3z = ty * phy + tx * phx + 1.0f;
      /* dd_Calculate_mijn_bpm_2 :105-129 */
      float dwx = W[j] - W[i], dwy = W[j + N1] - W[i + N1], dwz = W[j + N2] - W[i + N2];
      float c1 = nx * dwx + ny * dwy + nz * dwz;
      c1 = c1 * (dt * 0.5f * G * il);
      float mnx0 = Mijn[3 * s], mny0 = Mijn[3 * s + 1], mnz0 = Mijn[3 * s + 2];
      float mnx = M1x * mnx0 + M1y * mny0 + M1z * mnz0 + nx * c1;
      float mny = M2x * mnx0 + M2y * mny0 + M2z * mnz0 + ny * c1;
      float mnz = M3x * mnx0 + M3y * mny0 + M3z * mnz0 + nz * c1;
      Mijn[3 * s] = mnx;
      Mijn[3 * s + 1] = mny;
      Mijn[3 * s + 2] = mnz;
      float j4 = PI_F * b_r * b_r * b_r * b_r;
      msx = msx + mnx * j4;
      msy = msy + mny * j4;
      msz = msz + mnz * j4;
      /* dd_Calculate_mijt_bpm_2 :131-152 */
      float c1t = nx * dwx + ny * dwy + nz * dwz;
      float c2t = dt * 0.25f * E * il;
      float mtx0 = Mijt[3 * s], mty0 = Mijt[3 * s + 1], mtz0 = Mijt[3 * s + 2];
      float mtx = M1x * mtx0 + M1y * mty0 + M1z * mtz0 + (dwx - nx * c1t) * c2t;
      float mty = M2x * mtx0 + M2y * mty0 + M2z * mtz0 + (dwy - ny * c1t) * c2t;
      float mtz = M3x * mtx0 + M3y * mty0 + M3z * mtz0 + (dwz - nz * c1t) * c2t;
      Mijt[3 * s] = mtx;
      Mijt[3 * s + 1] = mty;
      Mijt[3 * s + 2] = mtz;
      c2t = c2t * (PI_F * b_r * b_r * b_r * b_r);
      msx = msx + mtx * c2t;
      msy = msy + mty * c2t;
      msz = msz + mtz * c2t;
      /* dd_Calculate_fijt_mijunsym_bpm_2 :154-197 */
      float hx = 0.5f * (W[j] + W[i]), hy = 0.5f * (W[j + N1] + W[i + N1]),
            hz = 0.5f * (W[j + N2] + W[i + N2]);
      float vx = V[j] - V[i] - hy * rz + hz * ry;
      float vy = V[j + N1] - V[i + N1] - hz * rx + hx * rz;
      float vz = V[j + N2] - V[i + N2] - hx * ry + hy * rx;
      float cv = nx * vx + ny * vy + nz * vz;
      float ox0 = Oijt[3 * s], oy0 = Oijt[3 * s + 1], oz0 = Oijt[3 * s + 2];
      float ox = (M1x * ox0 + M1y * oy0 + M1z * oz0) + (vx - nx * cv) * dt;
      float oy = (M2x * ox0 + M2y * oy0 + M2z * oz0) + (vy - ny * cv) * dt;
      float oz = (M3x * ox0 + M3y * oy0 + M3z * oz0) + (vz - nz * cv) * dt;
      float ct = G * il;
      Mc = ct;
      ct = ct * (b_r * b_r * PI_F);
      fsx = fsx + ox * ct;
      fsy = fsy + oy * ct;
      fsz = fsz + oz * ct;
      Oijt[3 * s] = ox;
      Oijt[3 * s + 1] = oy;
      Oijt[3 * s + 2] = oz;
      msx = msx + 0.5f * (ry * oz - rz * oy) * ct;
      msy = msy + 0.5f * (rz * ox - rx * oz) * ct;
      msz = msz + 0.5f * (rx * oy - ry * ox) * ct;
      /* dd_CheckBreakConditionDEM_2 :20-36 */
      Mc = Mc * orc_norm3(ox, oy, oz);
      Sc = Sc + orc_norm3(mtx, mty, mtz) * 4.0f * b_r;
      Mc = Mc + orc_norm3(mnx, mny, mnz) * 2.0f * b_r;
      if (Sc > sc || Mc > tc) {
        ILtype[s] = 0;
        zero3(Oijt, s);
        zero3(Mijn, s);
        zero3(Mijt, s);
      }
    }
    F[i] += fsx;
    F[i + N1] += fsy;
    F[i + N2] += fsz;
    M[i] += msx;
    M[i + N1] += msy;
    M[i + N2] += msz;
  }
}

/* md_calculateforcesdem3.cu:338-422 with :199-336 */
void orc_forces_contact(const float *R, const float *V, const float *W,
                        const uint32_t *IL, const uint8_t *ILtype, float *Oijt,
                        float *F, float *M, uint32_t N, uint32_t K, float dt,
                        float Estar, float Gstar, float pA, float mu, float mur,
                        float m, float r) {
  const size_t N1 = N, N2 = 2 * (size_t)N;
#pragma omp parallel for schedule(static)
  for (uint32_t i = 0; i < N; ++i) {
    float fsx = 0, fsy = 0, fsz = 0, msx = 0, msy = 0, msz = 0;
    for (uint32_t k = 0; k < K; ++k) {
      size_t s = (size_t)i * K + k;
      uint32_t j = IL[s];
      if (j == ORC_SENTINEL) {
        zero3(Oijt, s);
        continue;
      }
      if (ILtype[s] != 0) continue;
      float rx = R[j] - R[i], ry = R[j + N1] - R[i + N1], rz = R[j + N2] - R[i + N2];
      float rm = orc_norm3(rx, ry, rz);
      float inv_rm = 1.0f / rm;
      float nx = rx * inv_rm, ny = ry * inv_rm, nz = rz * inv_rm;
      if (!(rm < 2.0f * r)) {
        zero3(Oijt, s);
        continue;
      }
      /* dd_Calculate_fijn_ppHM_2 :235-252 */
      float q = sqrtf(r * fabsf(r - 0.5f * rm));
      float Kn = 2.0f * Estar * q;
      float Kt = 8.0f * Gstar * q;
      float wsx = W[j] + W[i], wsy = W[j + N1] + W[i + N1], wsz = W[j + N2] + W[i + N2];
      float vx = V[j] - V[i] + r * (ny * wsz - nz * wsy);
      float vy = V[j + N1] - V[i + N1] + r * (nz * wsx - nx * wsz);
      float vz = V[j + N2] - V[i + N2] + r * (nx * wsy - ny * wsx);
      float vn = nx * vx + ny * vy + nz * vz;
      float fn = (rm - 2.0f * r) * Kn * TWO_THIRDS_F;
      float dmp = -1.8257f * pA * vn * sqrtf(Kn * 0.5f * m);
      fsx = fsx + nx * (fn + dmp);
      fsy = fsy + ny * (fn + dmp);
      fsz = fsz + nz * (fn + dmp);
      float vtx = vx - nx * vn, vty = vy - ny * vn, vtz = vz - nz * vn;
      /* dd_Calculate_fijtmijt_ppHM_2 :255-314 */
      float Ox = Oijt[3 * s], Oy = Oijt[3 * s + 1], Oz = Oijt[3 * s + 2];
      float c1 = nx * Ox + ny * Oy + nz * Oz;
      float ox = Ox - nx * c1, oy = Oy - ny * c1, oz = Oz - nz * c1;
      c1 = ox * ox + oy * oy + oz * oz;
      if ((double)c1 > 1e-18) {
        c1 = 1.0f / c1;
        c1 = c1 * (Ox * Ox + Oy * Oy + Oz * Oz);
        c1 = sqrtf(c1);
        ox = ox * c1;
        oy = oy * c1;
        oz = oz * c1;
      }
      ox = ox + vtx * dt;
      oy = oy + vty * dt;
      oz = oz + vtz * dt;
      float ftx = ox * Kt, fty = oy * Kt, ftz = oz * Kt;
      float ftm = ftx * ftx + fty * fty + ftz * ftz;
      if (ftm > mu * mu * fn * fn) {
        c1 = mu * fabsf(fn) * orc_rsqrt(ftm);
        ftx = ftx * c1;
        fty = fty * c1;
        ftz = ftz * c1;
        c1 = 1.0f / Kt;
        ox = ftx * c1;
        oy = fty * c1;
        oz = ftz * c1;
      } else {
        c1 = -1.8257f * pA * sqrtf(Kt * 0.5f * m);
        ftx = ftx + vtx * c1;
        fty = fty + vty * c1;
        ftz = ftz + vtz * c1;
      }
      Oijt[3 * s] = ox;
      Oijt[3 * s + 1] = oy;
      Oijt[3 * s + 2] = oz;
      fsx = fsx + ftx;
      fsy = fsy + fty;
      fsz = fsz + ftz;
      msx = msx + (ny * ftz - nz * fty) * r;
      msy = msy + (nz * ftx - nx * ftz) * r;
      msz = msz + (nx * fty - ny * ftx) * r;
      /* dd_Calculate_mijroll_ppHM_2 :316-336 (else-branch wipes ms) */
      float wix = W[i], wiy = W[i + N1], wiz = W[i + N2];
      float wim = wix * wix + wiy * wiy + wiz * wiz;
      if ((double)wim > 1e-18) {
        float cr = -mur * fabsf(fn) * orc_rsqrt(wim);
        msx = msx + wix * cr;
        msy = msy + wiy * cr;
        msz = msz + wiz * cr;
      } else {
        msx = 0;
        msy = 0;
        msz = 0;
      }
    }
    F[i] += fsx;
    F[i + N1] += fsy;
    F[i + N2] += fsz;
    M[i] += msx;
    M[i + N1] += msy;
    M[i + N2] += msz;
  }
}

/* md_sumupforcesdem.cu:19-60 */
void orc_sumup_forces(const float *Fijn, const float *Fijt, const float *Mijn,
                      const float *Mijt, const float *Mijadd, float *F, float *M,
                      uint32_t N, uint32_t K) {
  for (uint32_t i = 0; i < N; ++i) {
    float f[3] = {0, 0, 0}, m[3] = {0, 0, 0};
    for (uint32_t k = 0; k < K; ++k) {
      size_t s = (size_t)i * K + k;
      for (int d = 0; d < 3; ++d) {
        f[d] = f[d] + (Fijn[3 * s + d] + Fijt[3 * s + d]);
        m[d] = m[d] + (Mijn[3 * s + d] + Mijt[3 * s + d] + Mijadd[3 * s + d]);
      }
    }
    for (int d = 0; d < 3; ++d) {
      F[i + (size_t)d * N] = f[d];
      M[i + (size_t)d * N] = m[d];
    }
  }
}

/* ---- the superseded per-slot force path (md_calculateforcesdem.cu) ----
 * Each slot's law goes to the slot arrays; d_SumUpForcesDEM
 * (orc_sumup_forces) then adds them per particle and d_CheckBreakConditionDEM
 * (orc_check_break) judges the bonds.  mode 0: d_CalculateForcesDEM (:328):
 * bonds, empty slots zeroed, contact slots untouched (its contact branch is
 * switched off by `&& false`, :361); mode 1: d_CalculateForcesDEM_1 (:501):
 * contact slots, those beyond 2 rP and empty slots zeroed; mode 2:
 * d_CalculateForcesDEM_2 (:581): bonds only. */
static void zero_slot(float *Fijn, float *Fijt, float *Oijt, float *Mijn, float *Mijt,
                      float *Mijadd, size_t s) {
  zero3(Fijn, s);
  zero3(Fijt, s);
  zero3(Oijt, s);
  zero3(Mijn, s);
  zero3(Mijt, s);
  zero3(Mijadd, s);
}

/* the bond law of one slot: dd_Calculate_fijn_bpm (:73), dd_Calculate_M123_bpm
 * (:36), dd_Calculate_mijn_bpm (:112), dd_Calculate_mijt_bpm (:129),
 * dd_Calculate_fijt_mijunsym_bpm (:147), in the order d_CalculateForcesDEM_2
 * calls them */
static void slot_bond(const float *R, const float *V, const float *W, const float *iL,
                      float *rij, float *Oijt, float *Fijn, float *Fijt, float *Mijn,
                      float *Mijt, float *Mijadd, size_t N, uint32_t i, uint32_t j, size_t s,
                      float dt, float E, float G, float b_r) {
  const size_t N1 = N, N2 = 2 * N;
  /* dd_Calculate_rijm_nij (:20-34) */
  float rx = R[j] - R[i], ry = R[j + N1] - R[i + N1], rz = R[j + N2] - R[i + N2];
  float rm = orc_norm3(rx, ry, rz);
  float inv = 1.0f / rm;
  float nx = rx * inv, ny = ry * inv, nz = rz * inv;
  float il = iL[s];
  /* fijn: b_r^2 pi E (|r| - L0) / L0 along n */
  float cef = b_r * b_r * PI_F * E * (rm - 1.0f / il) * il;
  Fijn[3 * s] = nx * cef;
  Fijn[3 * s + 1] = ny * cef;
  Fijn[3 * s + 2] = nz * cef;
  /* M123: the first-order rotation from the previous direction and the spin */
  float px = rij[3 * s], py = rij[3 * s + 1], pz = rij[3 * s + 2];
  float tx = py * nz - pz * ny, ty = pz * nx - px * nz, tz = px * ny - py * nx;
  float c2 = ((W[j] + W[i]) * nx + (W[j + N1] + W[i + N1]) * ny + (W[j + N2] + W[i + N2]) * nz) *
             dt * 0.5f;
  float phx = nx * c2, phy = ny * c2, phz = nz * c2;
  rij[3 * s] = nx;
  rij[3 * s + 1] = ny;
  rij[3 * s + 2] = nz;
  float M1x = 1.0f + tz * phz + ty * phy, M1y = phz - tz - ty * phx, M1z = -phy - tz * phx + ty;
  float M2x = tz - phz - tx * phy, M2y = tz * phz + 1.0f + tx * phx, M2z = -tz * phy + phx - tx;
  float M3x = -ty - tx * phz + phy, M3y = -ty * phz + tx - phx, M3z = ty * phy + tx * phx + 1.0f;
  /* mijn: twist, dt pi b_r^4 G / (2 L0) per unit relative spin along n */
  float dwx = W[j] - W[i], dwy = W[j + N1] - W[i + N1], dwz = W[j + N2] - W[i + N2];
  float c1 = nx * dwx + ny * dwy + nz * dwz;
  float k2 = dt * PI_F * b_r * b_r * b_r * b_r * 0.5f * G * il;
  float mx = Mijn[3 * s], my = Mijn[3 * s + 1], mz = Mijn[3 * s + 2];
  Mijn[3 * s] = M1x * mx + M1y * my + M1z * mz + nx * c1 * k2;
  Mijn[3 * s + 1] = M2x * mx + M2y * my + M2z * mz + ny * c1 * k2;
  Mijn[3 * s + 2] = M3x * mx + M3y * my + M3z * mz + nz * c1 * k2;
  /* mijt: bending, dt pi b_r^4 E / (4 L0) per unit relative spin across n */
  k2 = dt * PI_F * b_r * b_r * b_r * b_r * 0.25f * E * il;
  mx = Mijt[3 * s];
  my = Mijt[3 * s + 1];
  mz = Mijt[3 * s + 2];
  Mijt[3 * s] = M1x * mx + M1y * my + M1z * mz + (dwx - nx * c1) * k2;
  Mijt[3 * s + 1] = M2x * mx + M2y * my + M2z * mz + (dwy - ny * c1) * k2;
  Mijt[3 * s + 2] = M3x * mx + M3y * my + M3z * mz + (dwz - nz * c1) * k2;
  /* fijt and the unsymmetric moment */
  float hx = 0.5f * (W[j] + W[i]), hy = 0.5f * (W[j + N1] + W[i + N1]),
        hz = 0.5f * (W[j + N2] + W[i + N2]);
  float vx = V[j] - V[i] - hy * rz + hz * ry;
  float vy = V[j + N1] - V[i + N1] - hz * rx + hx * rz;
  float vz = V[j + N2] - V[i + N2] - hx * ry + hy * rx;
  float cv = nx * vx + ny * vy + nz * vz;
  float Ox = Oijt[3 * s], Oy = Oijt[3 * s + 1], Oz = Oijt[3 * s + 2];
  float ox = (M1x * Ox + M1y * Oy + M1z * Oz) + (vx - nx * cv) * dt;
  float oy = (M2x * Ox + M2y * Oy + M2z * Oz) + (vy - ny * cv) * dt;
  float oz = (M3x * Ox + M3y * Oy + M3z * Oz) + (vz - nz * cv) * dt;
  float ct = G * b_r * b_r * PI_F * il;
  Fijt[3 * s] = ox * ct;
  Fijt[3 * s + 1] = oy * ct;
  Fijt[3 * s + 2] = oz * ct;
  Oijt[3 * s] = ox;
  Oijt[3 * s + 1] = oy;
  Oijt[3 * s + 2] = oz;
  Mijadd[3 * s] = 0.5f * (ry * oz - rz * oy) * ct;
  Mijadd[3 * s + 1] = 0.5f * (rz * ox - rx * oz) * ct;
  Mijadd[3 * s + 2] = 0.5f * (rx * oy - ry * ox) * ct;
}

/* the contact law of one slot within 2 rP: dd_Calculate_fijn_ppHM (:189),
 * dd_Calculate_fijtmijt_ppHM (:245), Mijn zeroed, dd_Calculate_mijroll_ppHM
 * (:306) */
static void slot_contact(const float *V, const float *W, float *Oijt, float *Fijn, float *Fijt,
                         float *Mijn, float *Mijt, float *Mijadd, size_t N, uint32_t i,
                         uint32_t j, size_t s, float nx, float ny, float nz, float rm, float dt,
                         float E, float G, float pA, float mu, float mur, float mP, float rP) {
  const size_t N1 = N, N2 = 2 * N;
  float q = sqrtf(rP * fabsf(rP - 0.5f * rm));
  float Kn = 2.0f * E * q, Kt = 8.0f * G * q;
  float vx = V[j] - V[i] + rP * (ny * (W[j + N2] + W[i + N2]) - nz * (W[j + N1] + W[i + N1]));
  float vy = V[j + N1] - V[i + N1] + rP * (nz * (W[j] + W[i]) - nx * (W[j + N2] + W[i + N2]));
  float vz = V[j + N2] - V[i + N2] + rP * (nx * (W[j + N1] + W[i + N1]) - ny * (W[j] + W[i]));
  float vn = nx * vx + ny * vy + nz * vz;
  float fn = (rm - 2.0f * rP) * Kn * TWO_THIRDS_F;
  float dmp = -1.8257f * pA * vn * sqrtf(Kn * 0.5f * mP);
  Fijn[3 * s] = nx * (fn + dmp);
  Fijn[3 * s + 1] = ny * (fn + dmp);
  Fijn[3 * s + 2] = nz * (fn + dmp);
  float vtx = vx - nx * vn, vty = vy - ny * vn, vtz = vz - nz * vn;
  float Ox = Oijt[3 * s], Oy = Oijt[3 * s + 1], Oz = Oijt[3 * s + 2];
  float c1 = nx * Ox + ny * Oy + nz * Oz;
  float ox = Ox - nx * c1, oy = Oy - ny * c1, oz = Oz - nz * c1;
  c1 = ox * ox + oy * oy + oz * oz;
  if ((double)c1 > 1e-18) {
    c1 = 1.0f / c1;
    c1 = c1 * (Ox * Ox + Oy * Oy + Oz * Oz);
    c1 = sqrtf(c1);
    ox = ox * c1;
    oy = oy * c1;
    oz = oz * c1;
  }
  ox = ox + vtx * dt;
  oy = oy + vty * dt;
  oz = oz + vtz * dt;
  float fx = ox * Kt, fy = oy * Kt, fz = oz * Kt;
  float ftm = fx * fx + fy * fy + fz * fz;
  if (ftm > mu * mu * fn * fn) {
    c1 = mu * fabsf(fn) * orc_rsqrt(ftm);
    fx = fx * c1;
    fy = fy * c1;
    fz = fz * c1;
    c1 = 1.0f / Kt;
    ox = fx * c1;
    oy = fy * c1;
    oz = fz * c1;
  } else {
    c1 = -1.8257f * pA * sqrtf(Kt * 0.5f * mP);
    fx = fx + vtx * c1;
    fy = fy + vty * c1;
    fz = fz + vtz * c1;
  }
  Oijt[3 * s] = ox;
  Oijt[3 * s + 1] = oy;
  Oijt[3 * s + 2] = oz;
  Fijt[3 * s] = fx;
  Fijt[3 * s + 1] = fy;
  Fijt[3 * s + 2] = fz;
  Mijt[3 * s] = (ny * fz - nz * fy) * rP;
  Mijt[3 * s + 1] = (nz * fx - nx * fz) * rP;
  Mijt[3 * s + 2] = (nx * fy - ny * fx) * rP;
  zero3(Mijn, s);
  float wx = W[i], wy = W[i + N1], wz = W[i + N2];
  float wim = wx * wx + wy * wy + wz * wz;
  if ((double)wim > 1e-18) {
    float cr = -mur * fabsf(fn) * orc_rsqrt(wim);
    Mijadd[3 * s] = wx * cr;
    Mijadd[3 * s + 1] = wy * cr;
    Mijadd[3 * s + 2] = wz * cr;
  } else {
    zero3(Mijadd, s);
  }
}

void orc_forces_slots(int mode, const float *R, const float *V, const float *W,
                      const uint32_t *IL, const uint8_t *ILtype, const float *iL, float *rij,
                      float *Oijt, float *Fijn, float *Fijt, float *Mijn, float *Mijt,
                      float *Mijadd, uint32_t N, uint32_t K, float dt, float E, float G,
                      float b_r, float pA, float mu, float mur, float mP, float rP) {
  const size_t N1 = N, N2 = 2 * (size_t)N;
  for (uint32_t i = 0; i < N; ++i) {
    for (uint32_t k = 0; k < K; ++k) {
      size_t s = (size_t)i * K + k;
      uint32_t j = IL[s];
      if (j == ORC_SENTINEL) {
        if (mode != 2) zero_slot(Fijn, Fijt, Oijt, Mijn, Mijt, Mijadd, s);
        continue;
      }
      uint8_t type = ILtype[s];
      if (type == 1 && mode != 1) {
        slot_bond(R, V, W, iL, rij, Oijt, Fijn, Fijt, Mijn, Mijt, Mijadd, N, i, j, s, dt, E, G,
                  b_r);
      } else if (type == 0 && mode == 1) {
        float rx = R[j] - R[i], ry = R[j + N1] - R[i + N1], rz = R[j + N2] - R[i + N2];
        float rm = orc_norm3(rx, ry, rz);
        float inv = 1.0f / rm;
        if (rm < 2.0f * rP)
          slot_contact(V, W, Oijt, Fijn, Fijt, Mijn, Mijt, Mijadd, N, i, j, s, rx * inv,
                       ry * inv, rz * inv, rm, dt, E, G, pA, mu, mur, mP, rP);
        else
          zero_slot(Fijn, Fijt, Oijt, Mijn, Mijt, Mijadd, s);
      }
    }
  }
}

/* md_uniaxialcompression_simple.cu:19-29 dd_FrictionForce, :29-81
 * d_UniaxialCompression_simple: penalty platens at z = c.z +- Hd2 with
 * stiffness C and Coulomb friction mu against the in-plane velocity; the
 * reaction sums (bottom, top) are accumulated in double */
static void penalty_friction(const float *V, float *F, size_t N, uint32_t i, float fz,
                             float mu) {
  float vv = V[i] * V[i] + V[i + N] * V[i + N];
  if ((double)vv > 1e-10) {
    float iv = orc_rsqrt(vv);
    F[i] -= fabsf(fz) * mu * V[i] * iv;
    F[i + N] -= fabsf(fz) * mu * V[i + N] * iv;
  }
}

void orc_platen_uniaxial_penalty(const float *R, const float *V, float *F, uint32_t N,
                                 float cx, float cy, float cz, float RR, float Hd2, float C,
                                 float mu, float Zcut, double sums[2]) {
  const size_t N1 = N, N2 = 2 * (size_t)N;
  double sfb = 0, sft = 0;
  for (uint32_t i = 0; i < N; ++i) {
    float rx = R[i] - cx, ry = R[i + N1] - cy, rz = R[i + N2] - cz;
    if (rx * rx + ry * ry < RR && R[i + N2] < Zcut) {
      if (rz - Hd2 > 0) {
        float fz = (Hd2 - rz) * C;
        penalty_friction(V, F, N, i, fz, mu);
        F[i + N2] += fz;
        sft -= fz;
      }
      if (rz + Hd2 < 0) {
        float fz = (Hd2 + rz) * C;
        penalty_friction(V, F, N, i, fz, mu);
        F[i + N2] -= fz;
        sfb += fz;
      }
    }
  }
  sums[0] = sfb;
  sums[1] = sft;
}

/* md_checkbreakconditiondem.cu:20-55 */
void orc_check_break(const float *rij, float *Oijt, const float *Fijn,
                     const float *Fijt, const float *Mijn, const float *Mijt,
                     uint8_t *ILtype, uint32_t S, float sc, float tc, float b_r) {
  for (uint32_t s = 0; s < S; ++s) {
    if (ILtype[s] != 1) continue;
    float inv_area = 1.0f / (b_r * b_r * PI_F);
    const float *f = Fijn + 3 * (size_t)s, *n = rij + 3 * (size_t)s;
    float fn = f[0] * n[0] + f[1] * n[1] + f[2] * n[2];
    float Sn = orc_norm3(f[0], f[1], f[2]) * inv_area;
    if (fn < 0) Sn = Sn * -1.0f;
    float inv_I = 1.0f / (PI_F * b_r * b_r * b_r * b_r * 0.25f);
    const float *mt = Mijt + 3 * (size_t)s, *mn = Mijn + 3 * (size_t)s,
                *ft = Fijt + 3 * (size_t)s;
    float St = orc_norm3(mt[0], mt[1], mt[2]) * b_r * inv_I;
    float MSn = orc_norm3(mn[0], mn[1], mn[2]) * 0.5f * b_r * inv_I;
    float MSt = orc_norm3(ft[0], ft[1], ft[2]) * inv_area;
    if (Sn + St > sc || MSn + MSt > tc) {
      ILtype[s] = 0;
      zero3(Oijt, s);
    }
  }
}

/* md_calculateincrementsdem.cu:19-49 */
void orc_increments(const float *F, float *V, float *R, const float *M, float *W,
                    uint32_t N, float dt_m, float dt, float dt_I) {
#pragma omp parallel for schedule(static)
  for (uint32_t i = 0; i < N; ++i) {
    for (int d = 0; d < 3; ++d) V[i + (size_t)d * N] += F[i + (size_t)d * N] * dt_m;
    for (int d = 0; d < 3; ++d) R[i + (size_t)d * N] += V[i + (size_t)d * N] * dt;
    for (int d = 0; d < 3; ++d) W[i + (size_t)d * N] += M[i + (size_t)d * N] * dt_I;
  }
}

/* md_uniaxialcompression3_simple.cu:19-76; the reaction sums (bottom, top)
 * are accumulated in double here */
void orc_platen_uniaxial(float *R, float *V, float *F, uint32_t N, float cx,
                         float cy, float RR, float Zt, float Zb, float Ztr,
                         float Zbr, float vt, float Zcut, double sums[2]) {
  const size_t N1 = N, N2 = 2 * (size_t)N;
  double sfb = 0, sft = 0;
  for (uint32_t i = 0; i < N; ++i) {
    float rx = R[i] - cx, ry = R[i + N1] - cy, rz = R[i + N2];
    if (rx * rx + ry * ry < RR && rz < Zcut) {
      if ((double)(rz - Ztr) > -1e-9) {
        float fz = F[i + N2];
        if (fz > 0) sft += fz;
        if ((double)(rz - Zt) > -1e-9) {
          R[i + N2] = Zt;
          V[i + N2] = vt;
          F[i + N2] = 0;
        }
      }
      if ((double)(rz - Zbr) < 1e-9) {
        float fz = F[i + N2];
        if (fz < 0) sfb += fz;
        if ((double)(rz - Zb) < 1e-9) {
          R[i + N2] = Zb;
          V[i + N2] = 0;
          F[i + N2] = 0;
        }
      }
    }
  }
  sums[0] = sfb;
  sums[1] = sft;
}

/* md_braziltest_simple2.cu:19-92 */
void orc_platen_brazil(float *R, float *V, float *F, uint32_t N, float cx,
                       float cz, float RR, float Yt, float Yb, float Ytr,
                       float Ybr, float vt, float Zcut, double sums[2]) {
  const size_t N1 = N, N2 = 2 * (size_t)N;
  double sfb = 0, sft = 0;
  for (uint32_t i = 0; i < N; ++i) {
    float rx = R[i] - cx, ry = R[i + N1], rz = R[i + N2] - cz;
    if (rx * rx + rz * rz < RR && rz < Zcut) {
      if ((double)(ry - Ytr) > -1e-9) {
        float fy = F[i + N1];
        if (fy > 0) {
          sft += fy;
          float vv = V[i] * V[i] + V[i + N2] * V[i + N2];
          if ((double)vv > 1e-9) {
            vv = orc_rsqrt(vv);
            F[i] = F[i] - 0.45f * fy * V[i] * vv;
            F[i + N2] = F[i + N2] - 0.45f * fy * V[i + N2] * vv;
          }
        }
        if ((double)(ry - Yt) > -1e-9) {
          R[i + N1] = Yt;
          V[i + N1] = vt;
          F[i + N1] = 0;
        }
      }
      if ((double)(ry - Ybr) < 1e-9) {
        float fy = F[i + N1];
        if (fy < 0) {
          sfb += fy;
          float vv = V[i] * V[i] + V[i + N2] * V[i + N2];
          if ((double)vv > 1e-9) {
            vv = orc_rsqrt(vv);
            F[i] = F[i] + 0.45f * fy * V[i] * vv;
            F[i + N2] = F[i + N2] + 0.45f * fy * V[i + N2] * vv;
          }
        }
        if ((double)(ry - Yb) < 1e-9) {
          R[i + N1] = Yb;
          V[i + N1] = 0;
          F[i + N1] = 0;
        }
      }
    }
  }
  sums[0] = sfb;
  sums[1] = sft;
}

/* md_parallelepipedrestriction.cu:73-158 */
void orc_cut_restriction(float *R, float *V, float *F, uint32_t N,
                         const float center[3], const float sized2[3],
                         const float hp[3]) {
#pragma omp parallel for schedule(static)
  for (uint32_t i = 0; i < N; ++i) {
    int out = 0;
    for (int d = 0; d < 3; ++d) {
      float dr = R[i + (size_t)d * N] - center[d];
      if (dr - sized2[d] > 0) out = 1;
      if (dr + sized2[d] < 0) out = 1;
    }
    if (out) {
      for (int d = 0; d < 3; ++d) {
        R[i + (size_t)d * N] = hp[d];
        V[i + (size_t)d * N] = 0.0f;
        F[i + (size_t)d * N] = 0.0f;
      }
    }
  }
}

/* md_firerelaxation.cu:18-56 (V is recomputed from F, the old V is dropped) */
void orc_fire_increments(const float *F, float *V, float *R, uint32_t N,
                         float dt_m, float dt, float alpha) {
  const size_t N1 = N, N2 = 2 * (size_t)N;
  for (uint32_t i = 0; i < N; ++i) {
    float vx = F[i] * dt_m, vy = F[i + N1] * dt_m, vz = F[i + N2] * dt_m;
    float vv = vx * vx + vy * vy + vz * vz;
    float f = F[i] * F[i] + F[i + N1] * F[i + N1] + F[i + N2] * F[i + N2];
    float vdf;
    if (f > 1e-12f)
      vdf = alpha * sqrtf(vv * (1.0f / f));
    else
      vdf = (float)((double)(alpha * sqrtf(vv)) * 1e6);
    vx = (1.0f - alpha) * vx + vdf * F[i];
    vy = (1.0f - alpha) * vy + vdf * F[i + N1];
    vz = (1.0f - alpha) * vz + vdf * F[i + N2];
    V[i] = vx;
    V[i + N1] = vy;
    V[i + N2] = vz;
    R[i] += vx * dt;
    R[i + N1] += vy * dt;
    R[i + N2] += vz * dt;
  }
}

/* md_firerelaxation.cu:58-70 */
void orc_fire_halfstep(const float *V, float *R, uint32_t N, float dt_d2) {
  for (size_t s = 0; s < 3 * (size_t)N; ++s) R[s] -= V[s] * dt_d2;
}

/* md_firerelaxation.cu:372-490 + :266-272, summed in double */
void orc_fire_power(const float *F, const float *V, const float *M,
                    const float *W, uint32_t N, double out[4]) {
  double p = 0, ff = 0, vv = 0, fmax = 0;
  for (uint32_t i = 0; i < N; ++i) {
    double f2 = 0;
    for (int d = 0; d < 3; ++d) {
      double f = F[i + (size_t)d * N], v = V[i + (size_t)d * N];
      p += f * v;
      f2 += f * f;
      vv += v * v;
    }
    ff += f2;
    if (f2 > fmax) fmax = f2;
    if (M && W) {
      for (int d = 0; d < 3; ++d) {
        double m = M[i + (size_t)d * N], w = W[i + (size_t)d * N];
        p += m * w;
        ff += m * m;
        vv += w * w;
      }
    }
  }
  out[0] = p;
  out[1] = ff;
  out[2] = vv;
  out[3] = fmax;
}

/* ------------------------------------------------------------- native f64 */
static void cross3(const double a[3], const double b[3], double o[3]) {
  o[0] = a[1] * b[2] - a[2] * b[1];
  o[1] = a[2] * b[0] - a[0] * b[2];
  o[2] = a[0] * b[1] - a[1] * b[0];
}
static double dot3(const double a[3], const double b[3]) {
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}
/* y = Mrot * x, with the first-order rotation matrix of
 * md_calculateforcesdem3.cu:80-88 built from t = n_prev x n and phi */
static void rot_apply(const double t[3], const double ph[3], const double x[3],
                      double y[3]) {
  double M1[3] = {1.0 + t[2] * ph[2] + t[1] * ph[1], ph[2] - t[2] - t[1] * ph[0],
                  -ph[1] - t[2] * ph[0] + t[1]};
  double M2[3] = {t[2] - ph[2] - t[0] * ph[1], t[2] * ph[2] + 1.0 + t[0] * ph[0],
                  -t[2] * ph[1] + ph[0] - t[0]};
  double M3[3] = {-t[1] - t[0] * ph[2] + ph[1], -t[1] * ph[2] + t[0] - ph[0],
                  t[1] * ph[1] + t[0] * ph[0] + 1.0};
  y[0] = dot3(M1, x);
  y[1] = dot3(M2, x);
  y[2] = dot3(M3, x);
}

/* The same bond law as orc_forces_bpm, per undirected bond, in double, with
 * per-bond radius and strengths; the j end receives the mirrored load
 * (-Fn-Ft, -Mn*J-Mt*c + 0.5 r x Ft) that the reference's j->i slot computes.
 * Two restatements of the reference's stored state (md_calculateforcesdem3.cu
 * :62-88, 105-130), exact when positions advance by R += V dt between
 * evaluations: the previous bond direction (the reference's rij) is the one
 * at R - V dt, and the twisting moment Mn, parallel to n, is kept as its
 * component mn = Mn . n, incremented by (n . dW) dt G/(2 L0). */
void orc_native_bonds_f64(const double *R, const double *V, const double *W,
                          uint32_t N, const uint32_t *bi, const uint32_t *bj,
                          const double *iL, const double *br, const double *sc,
                          const double *tc, double *state, uint8_t *active,
                          uint32_t Nb, double dt, double E, double G,
                          int refmode, double *F, double *M) {
  const double pi = 3.14159265358979323846;
  memset(F, 0, sizeof(double) * 3 * (size_t)N);
  memset(M, 0, sizeof(double) * 3 * (size_t)N);
  for (uint32_t b = 0; b < Nb; ++b) {
    if (!active[b]) continue;
    uint32_t i = bi[b], j = bj[b];
    double r[3], rp[3], Wi[3], Wj[3], Vi[3], Vj[3];
    for (int d = 0; d < 3; ++d) {
      r[d] = R[j + (size_t)d * N] - R[i + (size_t)d * N];
      Wi[d] = W[i + (size_t)d * N];
      Wj[d] = W[j + (size_t)d * N];
      Vi[d] = V[i + (size_t)d * N];
      Vj[d] = V[j + (size_t)d * N];
      rp[d] = r[d] - (Vj[d] - Vi[d]) * dt;
    }
    double L = sqrt(dot3(r, r)), n[3] = {r[0] / L, r[1] / L, r[2] / L};
    double Lp = sqrt(dot3(rp, rp)), np[3] = {rp[0] / Lp, rp[1] / Lp, rp[2] / Lp};
    double il = iL[b], rb = br[b], A = pi * rb * rb, J = pi * rb * rb * rb * rb;
    double *st = state + 8 * (size_t)b;
    double strain_e = E * (L - 1.0 / il) * il;
    double Fn[3] = {n[0] * strain_e * A, n[1] * strain_e * A, n[2] * strain_e * A};
    double t[3], ws[3] = {Wi[0] + Wj[0], Wi[1] + Wj[1], Wi[2] + Wj[2]};
    cross3(np, n, t);
    double c2 = dot3(ws, n) * dt * 0.5, ph[3] = {n[0] * c2, n[1] * c2, n[2] * c2};
    double dw[3] = {Wj[0] - Wi[0], Wj[1] - Wi[1], Wj[2] - Wi[2]};
    double dwn = dot3(n, dw);
    double mt[3], o[3];
    rot_apply(t, ph, st + 4, mt);
    rot_apply(t, ph, st, o);
    double mn = st[3] + dwn * dt * 0.5 * G * il, cbend = dt * 0.25 * E * il;
    double h[3] = {0.5 * ws[0], 0.5 * ws[1], 0.5 * ws[2]}, hxr[3];
    cross3(h, r, hxr);
    double v[3] = {Vj[0] - Vi[0] - hxr[0], Vj[1] - Vi[1] - hxr[1], Vj[2] - Vi[2] - hxr[2]};
    double vn = dot3(n, v);
    for (int d = 0; d < 3; ++d) {
      mt[d] += (dw[d] - n[d] * dwn) * cbend;
      o[d] += (v[d] - n[d] * vn) * dt;
    }
    double kt = G * il * A;
    double Ft[3] = {o[0] * kt, o[1] * kt, o[2] * kt}, rxFt[3];
    cross3(r, Ft, rxFt);
    double mtfac = refmode ? cbend * J : J;
    for (int d = 0; d < 3; ++d) {
      double fi = Fn[d] + Ft[d];
      double mi = n[d] * mn * J + mt[d] * mtfac + 0.5 * rxFt[d];
      double mj = -n[d] * mn * J - mt[d] * mtfac + 0.5 * rxFt[d];
      F[i + (size_t)d * N] += fi;
      F[j + (size_t)d * N] -= fi;
      M[i + (size_t)d * N] += mi;
      M[j + (size_t)d * N] += mj;
      st[d] = o[d];
      st[4 + d] = mt[d];
    }
    st[3] = mn;
    double Sc = strain_e + sqrt(dot3(mt, mt)) * 4.0 * rb;
    double Mc = G * il * sqrt(dot3(o, o)) + fabs(mn) * 2.0 * rb;
    if (Sc > sc[b] || Mc > tc[b]) {
      active[b] = 0;
      for (int d = 0; d < 8; ++d) st[d] = 0.0;
    }
  }
}

/* Contact law of the native tier per directed pair (i sees j), in double:
 * md_calculateforcesdem3.cu:199-336 (normal Hertz force with viscous damping
 * dd_Calculate_fijn_ppHM_2 :235-252, tangential spring with Coulomb cap
 * dd_Calculate_fijtmijt_ppHM_2 :255-314, rolling resistance
 * dd_Calculate_mijroll_ppHM_2 :316-336) restated for per-particle radii and
 * masses: R* = ri rj/(ri+rj), m* = mi mj/(mi+mj), lever arms ri and rj.  The
 * pair overlaps once its centre distance is below reach[k] (ri + rj; for a
 * broken bond of the physical mode min(ri + rj, L0)).  Not overlapping: the
 * spring O[3k] is cleared.  The rolling-moment wipe of the reference mode is
 * not part of it. */
void orc_native_contacts_f64(const double *R, const double *V, const double *W,
                             const double *rad, uint32_t N, const uint32_t *pi,
                             const uint32_t *pj, const double *reach, double *O,
                             uint32_t P, double dt, double Estar, double Gstar,
                             double pA, double mu, double mur, double rho43pi,
                             double *F, double *M, double *cap) {
  memset(F, 0, sizeof(double) * 3 * (size_t)N);
  memset(M, 0, sizeof(double) * 3 * (size_t)N);
  if (cap) memset(cap, 0, sizeof(double) * (size_t)P);
  for (uint32_t k = 0; k < P; ++k) {
    const uint32_t i = pi[k], j = pj[k];
    double r[3], Vi[3], Vj[3], Wi[3], Wj[3];
    for (int d = 0; d < 3; ++d) {
      r[d] = R[j + (size_t)d * N] - R[i + (size_t)d * N];
      Vi[d] = V[i + (size_t)d * N];
      Vj[d] = V[j + (size_t)d * N];
      Wi[d] = W[i + (size_t)d * N];
      Wj[d] = W[j + (size_t)d * N];
    }
    double *o = O + 3 * (size_t)k;
    const double L = sqrt(dot3(r, r));
    if (!(L < reach[k]) || !(L > 1e-5)) {
      o[0] = o[1] = o[2] = 0.0;
      continue;
    }
    const double n[3] = {r[0] / L, r[1] / L, r[2] / L};
    const double ri = rad[i], rj = rad[j];
    const double Rs = ri * rj / (ri + rj);
    const double mi = rho43pi * ri * ri * ri, mj = rho43pi * rj * rj * rj;
    const double ms = mi * mj / (mi + mj);
    const double q = sqrt(Rs * (reach[k] - L));
    const double Kn = 2.0 * Estar * q, Kt = 8.0 * Gstar * q;
    double arm[3] = {Wi[0] * ri + Wj[0] * rj, Wi[1] * ri + Wj[1] * rj,
                     Wi[2] * ri + Wj[2] * rj}, nxa[3];
    cross3(n, arm, nxa);
    double v[3] = {Vj[0] - Vi[0] + nxa[0], Vj[1] - Vi[1] + nxa[1], Vj[2] - Vi[2] + nxa[2]};
    const double vn = dot3(n, v);
    const double fn = (L - reach[k]) * Kn * (2.0 / 3.0);
    const double dmp = -1.8257 * pA * vn * sqrt(Kn * ms);
    double vt[3], ot[3], ft[3];
    const double on = dot3(n, o);
    for (int d = 0; d < 3; ++d) {
      vt[d] = v[d] - n[d] * vn;
      ot[d] = o[d] - n[d] * on;
    }
    /* the old spring turned into the tangent plane, length kept */
    const double oo = dot3(ot, ot);
    if (oo > 1e-18) {
      const double s = sqrt(dot3(o, o) / oo);
      for (int d = 0; d < 3; ++d) ot[d] *= s;
    }
    for (int d = 0; d < 3; ++d) {
      ot[d] += vt[d] * dt;
      ft[d] = ot[d] * Kt;
    }
    const double ftm = dot3(ft, ft);
    if (cap) cap[k] = ftm / (mu * mu * fn * fn) - 1.0;
    if (ftm > mu * mu * fn * fn) {
      const double s = mu * fabs(fn) / sqrt(ftm);
      for (int d = 0; d < 3; ++d) {
        ft[d] *= s;
        ot[d] = ft[d] / Kt;
      }
    } else {
      const double c = -1.8257 * pA * sqrt(Kt * ms);
      for (int d = 0; d < 3; ++d) ft[d] += vt[d] * c;
    }
    double nxf[3];
    cross3(n, ft, nxf);
    const double wi2 = dot3(Wi, Wi);
    for (int d = 0; d < 3; ++d) {
      o[d] = ot[d];
      F[i + (size_t)d * N] += n[d] * (fn + dmp) + ft[d];
      double m = nxf[d] * ri;
      if (wi2 > 1e-18) m += Wi[d] * (-mur * fabs(fn) / sqrt(wi2));
      M[i + (size_t)d * N] += m;
    }
  }
}

/* threads the per-particle loops run on: 1 in the plain build, the OpenMP
 * team size in the -fopenmp build (the loops are independent per particle,
 * so both builds give the same bits) */
int orc_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}
