// rfd_forces.hip -- bond and contact forces in the reference slot layout.
//
// The law of each slot (of each listed bond slot in the bond kernel) is
// evaluated by a thread of its own, which leaves the terms the reference adds
// to the particle's sums in LDS; one thread per particle then adds them in
// slot order, term by term, so that F += fs, M += ms are formed in the same
// order as md_calculateforcesdem3.cu:437-480 and :353-419 and are bit-exact
// against a sequential restatement.
#include "rfd_internal.h"

using namespace rfd;

namespace {

constexpr int kMaxK = 64;

struct f3 {
  float x, y, z;
};
__device__ __forceinline__ f3 ld3(const float *p, size_t s) {
  return f3{p[3 * s], p[3 * s + 1], p[3 * s + 2]};
}
__device__ __forceinline__ void st3(float *p, size_t s, float x, float y, float z) {
  p[3 * s] = x;
  p[3 * s + 1] = y;
  p[3 * s + 2] = z;
}

// A particle's F and M, read before the slot terms are formed so that the
// row sums at the end wait for LDS only.
struct RowFM {
  float fx, fy, fz, mx, my, mz;
};
__device__ __forceinline__ RowFM load_fm(const float *__restrict__ F,
                                         const float *__restrict__ M, uint32_t ip, size_t N1,
                                         size_t N2) {
  return RowFM{F[ip], F[ip + N1], F[ip + N2], M[ip], M[ip + N1], M[ip + N2]};
}

// d_CalculateForcesDEM_22 (md_calculateforcesdem3.cu:424-483).  A workgroup
// takes the rows of P particles (P K <= 512 slots), lists their live bond
// slots in slot order (a ballot per wave), evaluates the bond law of each
// listed slot on its own lane and keeps the five terms the reference adds to
// the particle's sums in LDS; then one thread per particle adds them in slot
// order, term by term, as the reference's row loop does (same bits as a
// sequential walk of the row).  Listing the live slots first keeps every lane
// of the law busy: about a third of a row's slots hold bonds.
constexpr int kSlotThreads = 256;
constexpr uint32_t kBpmSlots = 512;  // slots of a workgroup's rows
// particles per workgroup: as many rows as fit in kBpmSlots, at most one per
// thread (the row sums take one thread each)
__host__ __device__ inline uint32_t bpm_rows(uint32_t K) {
  return min(kBpmSlots / K, (uint32_t)kSlotThreads);
}

// Lists, in slot order, the row-local slots ls < nsl <= kBpmSlots of a tile
// for which keep(ls, val) holds, stores the bits of the val it sets beside
// each listed slot in s_val, and returns how many.  The keep() tests of all
// sweeps of kSlotThreads slots are made first, so their loads are in flight
// together; one ballot per wave and sweep, one barrier.  Every thread of the
// workgroup calls it.
constexpr uint32_t kListSweeps = kBpmSlots / kSlotThreads;
template <class Keep>
__device__ __forceinline__ uint32_t list_slots(uint32_t nsl, uint16_t *s_slot, float *s_val,
                                               uint32_t (*s_wave)[kSlotThreads / 64],
                                               Keep keep) {
  const uint32_t t = threadIdx.x, lane = t & 63, wid = t >> 6;
  bool kept[kListSweeps];
  uint32_t val[kListSweeps];
#pragma unroll
  for (uint32_t r = 0; r < kListSweeps; ++r) {
    const uint32_t ls = r * kSlotThreads + t;
    val[r] = 0;
    kept[r] = ls < nsl && keep(ls, val[r]);
  }
  uint64_t bal[kListSweeps];
#pragma unroll
  for (uint32_t r = 0; r < kListSweeps; ++r) {
    bal[r] = __ballot(kept[r]);
    if (lane == 0) s_wave[r][wid] = (uint32_t)__popcll(bal[r]);
  }
  __syncthreads();
  uint32_t n = 0;
#pragma unroll
  for (uint32_t r = 0; r < kListSweeps; ++r) {
    uint32_t before = 0, all = 0;
    for (uint32_t w = 0; w < kSlotThreads / 64; ++w) {
      before += w < wid ? s_wave[r][w] : 0u;
      all += s_wave[r][w];
    }
    if (kept[r]) {
      const uint32_t at = n + before + (uint32_t)__popcll(bal[r] & ((1ull << lane) - 1ull));
      s_slot[at] = (uint16_t)(r * kSlotThreads + t);
      s_val[at] = __uint_as_float(val[r]);
    }
    n += all;
  }
  __syncthreads();
  return n;
}

// The listed entries [q0, q1) that belong to row lp (slots lp K .. lp K + K - 1).
__device__ __forceinline__ void row_range(const uint16_t *s_slot, uint32_t n, uint32_t lp,
                                          uint32_t K, uint32_t &q0, uint32_t &q1) {
  uint32_t lo = 0, hi = n;
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (s_slot[mid] < lp * K) lo = mid + 1; else hi = mid;
  }
  q0 = lo;
  hi = n;
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (s_slot[mid] < (lp + 1) * K) lo = mid + 1; else hi = mid;
  }
  q1 = lo;
}

__global__ void __launch_bounds__(kSlotThreads)
k_forces_bpm(const float *__restrict__ R, const float *__restrict__ V,
             const float *__restrict__ W, const uint32_t *__restrict__ IL,
             uint8_t *__restrict__ ILtype, const float *__restrict__ iL,
             float *__restrict__ rij, float *__restrict__ Oijt, float *__restrict__ Mijn,
             float *__restrict__ Mijt, float *__restrict__ F, float *__restrict__ M,
             uint32_t N, uint32_t K, float dt, float E, float G, float b_r, float scrit,
             float tcrit, uint32_t nblk, uint32_t per) {
  // per listed slot: normal force, twisting and bending moments, tangential
  // force, moment of the tangential force (3 floats each)
  __shared__ float s_c[15][kBpmSlots];
  __shared__ uint16_t s_slot[kBpmSlots];  // row-local slot (lp K + k) of each listed slot
  __shared__ uint32_t s_wave[kListSweeps][kSlotThreads / 64];
  const uint32_t P = bpm_rows(K);
  const uint32_t t = threadIdx.x;
  const uint32_t blk = xcd_block(per);
  if (blk >= nblk) return;
  const uint32_t p0 = blk * P;
  const uint32_t np = min(P, N - p0);
  const uint32_t nsl = np * K;
  const size_t g0 = (size_t)p0 * K;
  const size_t N1 = N, N2 = 2 * (size_t)N;
  const float area_k = b_r * b_r * kPiF;          // b_r*b_r*MCf_pi
  const float j4 = kPiF * b_r * b_r * b_r * b_r;  // MCf_pi*b_r^4
  // the row's F and M, read now so that the row sums wait for LDS only
  const uint32_t ip = p0 + t;
  RowFM fm{};
  if (t < np) fm = load_fm(F, M, ip, N1, N2);
  // 1. the live bond slots, in slot order, and their partners (in the last
  // term's column: the one thread that evaluates a listed slot reads its
  // partner there before it stores the term)
  const uint32_t nlive = list_slots(nsl, s_slot, s_c[14], s_wave, [&](uint32_t ls, uint32_t &j) {
    j = IL[g0 + ls];
    const uint8_t ty = ILtype[g0 + ls];  // read with the id, not after it
    return j != kSentinel && ty == 1;
  });
  // 2. the law of each listed slot
  for (uint32_t q = t; q < nlive; q += kSlotThreads) {
    const uint32_t ls = s_slot[q];
    const uint32_t i = p0 + ls / K;
    const size_t s = g0 + ls;
    const uint32_t j = __float_as_uint(s_c[14][q]);
    const float Rix = R[i], Riy = R[i + N1], Riz = R[i + N2];
    const float Vix = V[i], Viy = V[i + N1], Viz = V[i + N2];
    const float Wix = W[i], Wiy = W[i + N1], Wiz = W[i + N2];
    // dd_Calculate_rijm_nij_2 (:38-52)
    const float rx = R[j] - Rix, ry = R[j + N1] - Riy, rz = R[j + N2] - Riz;
    const float rm = norm3(rx, ry, rz);
    const float inv_rm = 1.0f / rm;
    const float nx = rx * inv_rm, ny = ry * inv_rm, nz = rz * inv_rm;
    const float il = iL[s];
    // dd_Calculate_fijn_bpm_2 (:91-103)
    float cef = E * (rm - 1.0f / il) * il;
    float Sc = cef;
    cef = cef * area_k;
    s_c[0][q] = nx * cef;
    s_c[1][q] = ny * cef;
    s_c[2][q] = nz * cef;
    // dd_Calculate_M123_bpm_2 (:54-89)
    const f3 np_ = ld3(rij, s);
    const float tx = np_.y * nz - np_.z * ny;
    const float ty = np_.z * nx - np_.x * nz;
    const float tz = np_.x * ny - np_.y * nx;
    const float Wjx = W[j], Wjy = W[j + N1], Wjz = W[j + N2];
    const float c2 = ((Wjx + Wix) * nx + (Wjy + Wiy) * ny + (Wjz + Wiz) * nz) * dt * 0.5f;
    const float phx = nx * c2, phy = ny * c2, phz = nz * c2;
    st3(rij, s, nx, ny, nz);
    const float M1x = 1.0f + tz * phz + ty * phy;
    const float M1y = phz - tz - ty * phx;
    const float M1z = -phy - tz * phx + ty;
    const float M2x = tz - phz - tx * phy;
    const float M2y = tz * phz + 1.0f + tx * phx;
    const float M2z = -tz * phy + phx - tx;
    const float M3x = -ty - tx * phz + phy;
    const float M3y = -ty * phz + tx - phx;
    const float M3z = ty * phy + tx * phx + 1.0f;
    // dd_Calculate_mijn_bpm_2 (:105-129)
    const float dwx = Wjx - Wix, dwy = Wjy - Wiy, dwz = Wjz - Wiz;
    float c1 = nx * dwx + ny * dwy + nz * dwz;
    const float dotw = c1;
    c1 = c1 * (dt * 0.5f * G * il);
    const f3 mn0 = ld3(Mijn, s);
    const float mnx = M1x * mn0.x + M1y * mn0.y + M1z * mn0.z + nx * c1;
    const float mny = M2x * mn0.x + M2y * mn0.y + M2z * mn0.z + ny * c1;
    const float mnz = M3x * mn0.x + M3y * mn0.y + M3z * mn0.z + nz * c1;
    s_c[3][q] = mnx * j4;
    s_c[4][q] = mny * j4;
    s_c[5][q] = mnz * j4;
    // dd_Calculate_mijt_bpm_2 (:131-152)
    float c2t = dt * 0.25f * E * il;
    const f3 mt0 = ld3(Mijt, s);
    const float mtx = M1x * mt0.x + M1y * mt0.y + M1z * mt0.z + (dwx - nx * dotw) * c2t;
    const float mty = M2x * mt0.x + M2y * mt0.y + M2z * mt0.z + (dwy - ny * dotw) * c2t;
    const float mtz = M3x * mt0.x + M3y * mt0.y + M3z * mt0.z + (dwz - nz * dotw) * c2t;
    c2t = c2t * j4;
    s_c[6][q] = mtx * c2t;
    s_c[7][q] = mty * c2t;
    s_c[8][q] = mtz * c2t;
    // dd_Calculate_fijt_mijunsym_bpm_2 (:154-197)
    const float hx = 0.5f * (Wjx + Wix), hy = 0.5f * (Wjy + Wiy), hz = 0.5f * (Wjz + Wiz);
    const float vx = V[j] - Vix - hy * rz + hz * ry;
    const float vy = V[j + N1] - Viy - hz * rx + hx * rz;
    const float vz = V[j + N2] - Viz - hx * ry + hy * rx;
    const float cv = nx * vx + ny * vy + nz * vz;
    const f3 o0 = ld3(Oijt, s);
    const float ox = (M1x * o0.x + M1y * o0.y + M1z * o0.z) + (vx - nx * cv) * dt;
    const float oy = (M2x * o0.x + M2y * o0.y + M2z * o0.z) + (vy - ny * cv) * dt;
    const float oz = (M3x * o0.x + M3y * o0.y + M3z * o0.z) + (vz - nz * cv) * dt;
    float ct = G * il;
    float Mc = ct;
    ct = ct * area_k;
    s_c[9][q] = ox * ct;
    s_c[10][q] = oy * ct;
    s_c[11][q] = oz * ct;
    s_c[12][q] = 0.5f * (ry * oz - rz * oy) * ct;
    s_c[13][q] = 0.5f * (rz * ox - rx * oz) * ct;
    s_c[14][q] = 0.5f * (rx * oy - ry * ox) * ct;
    // dd_CheckBreakConditionDEM_2 (:20-36)
    Mc = Mc * norm3(ox, oy, oz);
    Sc = Sc + norm3(mtx, mty, mtz) * 4.0f * b_r;
    Mc = Mc + norm3(mnx, mny, mnz) * 2.0f * b_r;
    if (Sc > scrit || Mc > tcrit) {
      ILtype[s] = 0;
      st3(Oijt, s, 0.0f, 0.0f, 0.0f);
      st3(Mijn, s, 0.0f, 0.0f, 0.0f);
      st3(Mijt, s, 0.0f, 0.0f, 0.0f);
    } else {
      st3(Oijt, s, ox, oy, oz);
      st3(Mijn, s, mnx, mny, mnz);
      st3(Mijt, s, mtx, mty, mtz);
    }
  }
  __syncthreads();
  // 3. the row sums, slot by slot and term by term in the reference's order
  if (t >= np) return;
  uint32_t q0, q1;
  row_range(s_slot, nlive, t, K, q0, q1);
  float fsx = 0, fsy = 0, fsz = 0, msx = 0, msy = 0, msz = 0;
  for (uint32_t q = q0; q < q1; ++q) {
    fsx = fsx + s_c[0][q];
    fsy = fsy + s_c[1][q];
    fsz = fsz + s_c[2][q];
    msx = msx + s_c[3][q];
    msy = msy + s_c[4][q];
    msz = msz + s_c[5][q];
    msx = msx + s_c[6][q];
    msy = msy + s_c[7][q];
    msz = msz + s_c[8][q];
    fsx = fsx + s_c[9][q];
    fsy = fsy + s_c[10][q];
    fsz = fsz + s_c[11][q];
    msx = msx + s_c[12][q];
    msy = msy + s_c[13][q];
    msz = msz + s_c[14][q];
  }
  F[ip] = fm.fx + fsx;
  F[ip + N1] = fm.fy + fsy;
  F[ip + N2] = fm.fz + fsz;
  M[ip] = fm.mx + msx;
  M[ip + N1] = fm.my + msy;
  M[ip + N2] = fm.mz + msz;
}

// d_CalculateForcesDEM_21 (md_calculateforcesdem3.cu:338-422), slot-parallel
// like k_forces_bpm: the terms of each contact slot go to LDS, the row sums
// are formed in slot order, the rolling-moment wipe (:330-335) included.
struct ContactLaw {
  float dt, Estar, Gstar, pA, mu, mur, m, r;
};

// The law of a contact slot s (row particle i, partner j, centre distance
// rm < 2r along r = (rx, ry, rz)): stores the slot's new tangential spring in
// Oijt and the twelve terms of the row sums in column c of s_c (normal force,
// tangential force, its moment, rolling moment); returns 1, or 2 when the
// particle does not spin and the reference clears ms instead of adding a
// rolling moment.
template <uint32_t C>
__device__ __forceinline__ uint8_t contact_law(const float *__restrict__ V,
                                               const float *__restrict__ W,
                                               float *__restrict__ Oijt, uint32_t i,
                                               uint32_t j, size_t s, size_t N1, size_t N2,
                                               float rx, float ry, float rz, float rm,
                                               const ContactLaw &L, float (&s_c)[12][C],
                                               uint32_t c) {
  const float r = L.r;
  const float inv_rm = 1.0f / rm;
  const float nx = rx * inv_rm, ny = ry * inv_rm, nz = rz * inv_rm;
  const float Vix = V[i], Viy = V[i + N1], Viz = V[i + N2];
  const float Wix = W[i], Wiy = W[i + N1], Wiz = W[i + N2];
  // dd_Calculate_fijn_ppHM_2 (:235-252)
  const float q = sqrtf(r * fabsf(r - 0.5f * rm));
  const float Kn = 2.0f * L.Estar * q;
  const float Kt = 8.0f * L.Gstar * q;
  const float Wjx = W[j], Wjy = W[j + N1], Wjz = W[j + N2];
  const float wsx = Wjx + Wix, wsy = Wjy + Wiy, wsz = Wjz + Wiz;
  const float vx = V[j] - Vix + r * (ny * wsz - nz * wsy);
  const float vy = V[j + N1] - Viy + r * (nz * wsx - nx * wsz);
  const float vz = V[j + N2] - Viz + r * (nx * wsy - ny * wsx);
  const float vn = nx * vx + ny * vy + nz * vz;
  const float fn = (rm - 2.0f * r) * Kn * kTwoThirdsF;
  const float dmp = -1.8257f * L.pA * vn * sqrtf(Kn * 0.5f * L.m);
  s_c[0][c] = nx * (fn + dmp);
  s_c[1][c] = ny * (fn + dmp);
  s_c[2][c] = nz * (fn + dmp);
  const float vtx = vx - nx * vn, vty = vy - ny * vn, vtz = vz - nz * vn;
  // dd_Calculate_fijtmijt_ppHM_2 (:255-314)
  const f3 O = ld3(Oijt, s);
  float c1 = nx * O.x + ny * O.y + nz * O.z;
  float ox = O.x - nx * c1, oy = O.y - ny * c1, oz = O.z - nz * c1;
  c1 = ox * ox + oy * oy + oz * oz;
  if ((double)c1 > 1e-18) {
    c1 = 1.0f / c1;
    c1 = c1 * (O.x * O.x + O.y * O.y + O.z * O.z);
    c1 = sqrtf(c1);
    ox = ox * c1;
    oy = oy * c1;
    oz = oz * c1;
  }
  ox = ox + vtx * L.dt;
  oy = oy + vty * L.dt;
  oz = oz + vtz * L.dt;
  float ftx = ox * Kt, fty = oy * Kt, ftz = oz * Kt;
  const float ftm = ftx * ftx + fty * fty + ftz * ftz;
  if (ftm > L.mu * L.mu * fn * fn) {
    c1 = L.mu * fabsf(fn) * rsqrt_rn(ftm);
    ftx = ftx * c1;
    fty = fty * c1;
    ftz = ftz * c1;
    c1 = 1.0f / Kt;
    ox = ftx * c1;
    oy = fty * c1;
    oz = ftz * c1;
  } else {
    c1 = -1.8257f * L.pA * sqrtf(Kt * 0.5f * L.m);
    ftx = ftx + vtx * c1;
    fty = fty + vty * c1;
    ftz = ftz + vtz * c1;
  }
  st3(Oijt, s, ox, oy, oz);
  s_c[3][c] = ftx;
  s_c[4][c] = fty;
  s_c[5][c] = ftz;
  s_c[6][c] = (ny * ftz - nz * fty) * r;
  s_c[7][c] = (nz * ftx - nx * ftz) * r;
  s_c[8][c] = (nx * fty - ny * ftx) * r;
  // dd_Calculate_mijroll_ppHM_2 (:316-336): the else branch clears ms
  const float wim = Wix * Wix + Wiy * Wiy + Wiz * Wiz;
  if ((double)wim > 1e-18) {
    const float cr = -L.mur * fabsf(fn) * rsqrt_rn(wim);
    s_c[9][c] = Wix * cr;
    s_c[10][c] = Wiy * cr;
    s_c[11][c] = Wiz * cr;
    return 1;
  }
  return 2;
}

// Adds the terms of one live slot q (s_live value lv) to the row sums, as the
// reference's row loop does: fs += normal, fs += tangential, ms += moment,
// then ms += rolling moment or, for lv == 2, ms = 0.
template <uint32_t C>
__device__ __forceinline__ void add_contact_terms(const float (&s_c)[12][C], uint32_t q,
                                                  uint32_t lv, float (&f)[3], float (&mm)[3]) {
  f[0] = f[0] + s_c[0][q];
  f[1] = f[1] + s_c[1][q];
  f[2] = f[2] + s_c[2][q];
  f[0] = f[0] + s_c[3][q];
  f[1] = f[1] + s_c[4][q];
  f[2] = f[2] + s_c[5][q];
  mm[0] = mm[0] + s_c[6][q];
  mm[1] = mm[1] + s_c[7][q];
  mm[2] = mm[2] + s_c[8][q];
  if (lv == 1) {
    mm[0] = mm[0] + s_c[9][q];
    mm[1] = mm[1] + s_c[10][q];
    mm[2] = mm[2] + s_c[11][q];
  } else {
    mm[0] = 0;
    mm[1] = 0;
    mm[2] = 0;
  }
}

// Row t's sums of the workgroup (slots t K .. t K + K - 1 of s_c and s_live),
// in slot order, added to the particle's F and M read earlier (fm) and
// stored.  WORDS: K is a multiple of 4 and s_live is read four slots per
// 32-bit word, skipping words without a live slot.
template <bool WORDS, uint32_t C>
__device__ __forceinline__ void contact_row_sums(const float (&s_c)[12][C],
                                                 const uint8_t *s_live, uint32_t t, uint32_t K,
                                                 const RowFM &fm, uint32_t ip, size_t N1,
                                                 size_t N2, float *__restrict__ F,
                                                 float *__restrict__ M) {
  float f[3] = {0, 0, 0}, mm[3] = {0, 0, 0};
  if constexpr (WORDS) {
    const uint32_t *w4 = reinterpret_cast<const uint32_t *>(s_live);
    for (uint32_t w = t * K / 4; w < (t + 1) * K / 4; ++w) {
      uint32_t lw = w4[w];
      while (lw) {
        const uint32_t u = __builtin_ctz(lw) >> 3;
        add_contact_terms(s_c, 4 * w + u, (lw >> (8 * u)) & 0xFFu, f, mm);
        lw &= ~(0xFFu << (8 * u));
      }
    }
  } else {
    for (uint32_t q = t * K; q < (t + 1) * K; ++q)
      if (s_live[q]) add_contact_terms(s_c, q, s_live[q], f, mm);
  }
  F[ip] = fm.fx + f[0];
  F[ip + N1] = fm.fy + f[1];
  F[ip + N2] = fm.fz + f[2];
  M[ip] = fm.mx + mm[0];
  M[ip + N1] = fm.my + mm[1];
  M[ip + N2] = fm.mz + mm[2];
}

// One thread per slot, floor(256 / K) rows per workgroup: for row lengths
// that are not a multiple of 4.
__global__ void __launch_bounds__(kSlotThreads)
k_forces_contact(const float *__restrict__ R, const float *__restrict__ V,
                 const float *__restrict__ W, const uint32_t *__restrict__ IL,
                 const uint8_t *__restrict__ ILtype, float *__restrict__ Oijt,
                 float *__restrict__ F, float *__restrict__ M, uint32_t N, uint32_t K,
                 ContactLaw L, uint32_t nblk, uint32_t per) {
  __shared__ float s_c[12][kSlotThreads];
  __shared__ uint8_t s_live[kSlotThreads];  // 0 none, 1 contact, 2 contact that wipes ms
  const uint32_t P = kSlotThreads / K;
  const uint32_t t = threadIdx.x, lp = t / K, k = t % K;
  const uint32_t blk = xcd_block(per);
  if (blk >= nblk) return;
  const uint32_t p0 = blk * P;
  const uint32_t i = p0 + lp;
  const size_t N1 = N, N2 = 2 * (size_t)N;
  const uint32_t ip = p0 + t;
  const bool row = t < P && ip < N;
  RowFM fm{};
  if (row) fm = load_fm(F, M, ip, N1, N2);
  uint8_t live = 0;
  if (lp < P && i < N) {
    const size_t s = (size_t)i * K + k;
    const uint32_t j = IL[s];
    if (j == kSentinel) {
      st3(Oijt, s, 0.0f, 0.0f, 0.0f);
    } else if (ILtype[s] == 0) {
      const float rx = R[j] - R[i], ry = R[j + N1] - R[i + N1], rz = R[j + N2] - R[i + N2];
      const float rm = norm3(rx, ry, rz);
      if (!(rm < 2.0f * L.r))
        st3(Oijt, s, 0.0f, 0.0f, 0.0f);
      else
        live = contact_law(V, W, Oijt, i, j, s, N1, N2, rx, ry, rz, rm, L, s_c, t);
    }
  }
  s_live[t] = live;
  __syncthreads();
  if (!row) return;
  contact_row_sums<false>(s_c, s_live, t, K, fm, ip, N1, N2, F, M);
}

// Four consecutive slots per thread (K a multiple of 4, so that they share
// a row and their ids, types and springs are 16-, 4- and 16-byte aligned),
// 1024 slots per workgroup: one 16-byte id load and one 4-byte type load per
// four slots, the partners' positions gathered together, and the springs of
// four empty or apart slots cleared by three 16-byte stores.  Most slots are
// empty, bonds, or candidates that do not touch, so this is the common path.
// The touching slots of the workgroup are numbered in slot order (a scan of
// per-thread counts) and their terms kept in kContactCap LDS columns, so the
// workgroup holds 12 KB of terms instead of 48 KB; a workgroup with more
// touching slots than that runs the law and the row sums in rounds of
// kContactCap numbers, each round's terms added in slot order after the last.
constexpr uint32_t kContact4Slots = 4 * kSlotThreads;
constexpr uint32_t kContactCap = 256;
__global__ void __launch_bounds__(kSlotThreads)
k_forces_contact4(const float *__restrict__ R, const float *__restrict__ V,
                  const float *__restrict__ W, const uint32_t *__restrict__ IL,
                  const uint8_t *__restrict__ ILtype, float *__restrict__ Oijt,
                  float *__restrict__ F, float *__restrict__ M, uint32_t N, uint32_t K,
                  ContactLaw L, uint32_t nblk, uint32_t per) {
  __shared__ float s_c[12][kContactCap];
  // per thread, a 16-bit entry for each of its four slots: 0, or the slot's
  // number among the touching slots << 2 | the law's result (1 or 2)
  __shared__ uint64_t s_meta[kSlotThreads];
  __shared__ uint32_t s_wave[kSlotThreads / 64];
  const uint32_t P = kContact4Slots / K;
  const uint32_t t = threadIdx.x, lane = t & 63, wid = t >> 6;
  const uint32_t blk = xcd_block(per);
  if (blk >= nblk) return;
  const uint32_t p0 = blk * P;
  const uint32_t i = p0 + 4 * t / K;
  const size_t N1 = N, N2 = 2 * (size_t)N;
  const uint32_t ip = p0 + t;
  const bool row = t < P && ip < N;
  RowFM fm{};
  if (row) fm = load_fm(F, M, ip, N1, N2);
  const size_t s0 = (size_t)p0 * K + 4 * t;
  uint32_t js[4] = {0, 0, 0, 0};
  bool touch[4] = {false, false, false, false};
  float rx[4] = {}, ry[4] = {}, rz[4] = {};
  if (4 * t < P * K && i < N) {
    const uint4 jv = *reinterpret_cast<const uint4 *>(IL + s0);
    const uint32_t tv = *reinterpret_cast<const uint32_t *>(ILtype + s0);
    js[0] = jv.x;
    js[1] = jv.y;
    js[2] = jv.z;
    js[3] = jv.w;
    bool cand[4];
#pragma unroll
    for (int u = 0; u < 4; ++u)
      cand[u] = js[u] != kSentinel && ((tv >> (8 * u)) & 0xFFu) == 0;
    // a slot's spring is cleared unless it holds a bond or a touching contact
    bool clear[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) clear[u] = js[u] == kSentinel;
    if (cand[0] | cand[1] | cand[2] | cand[3]) {
      const float Rix = R[i], Riy = R[i + N1], Riz = R[i + N2];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        if (!cand[u]) continue;
        rx[u] = R[js[u]] - Rix;
        ry[u] = R[js[u] + N1] - Riy;
        rz[u] = R[js[u] + N2] - Riz;
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        if (!cand[u]) continue;
        touch[u] = norm3(rx[u], ry[u], rz[u]) < 2.0f * L.r;
        clear[u] = !touch[u];
      }
    }
    if (clear[0] & clear[1] & clear[2] & clear[3]) {
      float4 *o = reinterpret_cast<float4 *>(Oijt + 3 * s0);
      const float4 z = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
      o[0] = z;
      o[1] = z;
      o[2] = z;
    } else {
#pragma unroll
      for (int u = 0; u < 4; ++u)
        if (clear[u]) st3(Oijt, s0 + u, 0.0f, 0.0f, 0.0f);
    }
  }
  // number the touching slots in slot order: the count of this thread (0-4)
  // is summed over the lanes below by its three bits' ballots
  const uint32_t cnt = (uint32_t)touch[0] + touch[1] + touch[2] + touch[3];
  const uint64_t below = (1ull << lane) - 1ull;
  uint32_t before = 0, wtot = 0;
#pragma unroll
  for (int bit = 0; bit < 3; ++bit) {
    const uint64_t bb = __ballot((cnt >> bit) & 1u);
    before += (uint32_t)__popcll(bb & below) << bit;
    wtot += (uint32_t)__popcll(bb) << bit;
  }
  if (lane == 0) s_wave[wid] = wtot;
  __syncthreads();
  uint32_t base = before, total = 0;
  for (uint32_t w = 0; w < kSlotThreads / 64; ++w) {
    base += w < wid ? s_wave[w] : 0u;
    total += s_wave[w];
  }
  float f[3] = {0, 0, 0}, mm[3] = {0, 0, 0};
  uint64_t meta = 0;
  for (uint32_t w0 = 0; w0 < total; w0 += kContactCap) {
    // one copy of the law, its slot's values picked by selects (the arrays
    // stay in registers)
    auto pick = [](const auto (&a)[4], uint32_t u) {
      return u == 0 ? a[0] : u == 1 ? a[1] : u == 2 ? a[2] : a[3];
    };
    uint32_t num = base;
#pragma unroll 1
    for (uint32_t u = 0; u < 4; ++u) {
      const bool tu = u == 0 ? touch[0] : u == 1 ? touch[1] : u == 2 ? touch[2] : touch[3];
      if (!tu) continue;
      if (num >= w0 && num < w0 + kContactCap) {
        const float x = pick(rx, u), y = pick(ry, u), z = pick(rz, u);
        const uint32_t code = contact_law(V, W, Oijt, i, pick(js, u), s0 + u, N1, N2, x, y, z,
                                          norm3(x, y, z), L, s_c, num - w0);
        meta |= (uint64_t)((num << 2) | code) << (16 * u);
      }
      ++num;
    }
    s_meta[t] = meta;
    __syncthreads();
    if (row) {
      // the row's slots in order, four per entry of s_meta; only those
      // numbered in this round's window
      for (uint32_t q4 = t * K / 4; q4 < (t + 1) * K / 4; ++q4) {
        uint64_t m = s_meta[q4];
        while (m) {
          const uint32_t u = __builtin_ctzll(m) >> 4;
          const uint32_t v = (uint32_t)(m >> (16 * u)) & 0xFFFFu;
          m &= ~(0xFFFFull << (16 * u));
          const uint32_t num2 = v >> 2;
          if (num2 >= w0 && num2 < w0 + kContactCap)
            add_contact_terms(s_c, num2 - w0, v & 3u, f, mm);
        }
      }
    }
    __syncthreads();
  }
  if (!row) return;
  F[ip] = fm.fx + f[0];
  F[ip + N1] = fm.fy + f[1];
  F[ip + N2] = fm.fz + f[2];
  M[ip] = fm.mx + mm[0];
  M[ip + N1] = fm.my + mm[1];
  M[ip + N2] = fm.mz + mm[2];
}

// d_SumUpForcesDEM (md_sumupforcesdem.cu:19-60)
__global__ void k_sumup(const float *__restrict__ Fijn, const float *__restrict__ Fijt,
                        const float *__restrict__ Mijn, const float *__restrict__ Mijt,
                        const float *__restrict__ Mijadd, float *__restrict__ F,
                        float *__restrict__ M, uint32_t N, uint32_t K) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += grid_stride()) {
    float f[3] = {0, 0, 0}, mm[3] = {0, 0, 0};
    for (uint32_t k = 0; k < K; ++k) {
      const size_t s = (size_t)i * K + k;
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        f[d] = f[d] + (Fijn[3 * s + d] + Fijt[3 * s + d]);
        mm[d] = mm[d] + (Mijn[3 * s + d] + Mijt[3 * s + d] + Mijadd[3 * s + d]);
      }
    }
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      F[i + (size_t)d * N] = f[d];
      M[i + (size_t)d * N] = mm[d];
    }
  }
}

// d_CheckBreakConditionDEM (md_checkbreakconditiondem.cu:20-55)
__global__ void k_check_break(const float *__restrict__ rij, float *__restrict__ Oijt,
                              const float *__restrict__ Fijn, const float *__restrict__ Fijt,
                              const float *__restrict__ Mijn, const float *__restrict__ Mijt,
                              uint8_t *__restrict__ ILtype, uint32_t S, float scrit,
                              float tcrit, float b_r) {
  for (uint32_t s = blockIdx.x * blockDim.x + threadIdx.x; s < S; s += grid_stride()) {
    if (ILtype[s] != 1) continue;
    const float inv_area = 1.0f / (b_r * b_r * kPiF);
    const f3 f = ld3(Fijn, s), n = ld3(rij, s);
    const float fn = f.x * n.x + f.y * n.y + f.z * n.z;
    float Sn = norm3(f.x, f.y, f.z) * inv_area;
    if (fn < 0) Sn = Sn * -1.0f;
    const float inv_I = 1.0f / (kPiF * b_r * b_r * b_r * b_r * 0.25f);
    const f3 mt = ld3(Mijt, s), mn = ld3(Mijn, s), ft = ld3(Fijt, s);
    const float St = norm3(mt.x, mt.y, mt.z) * b_r * inv_I;
    const float MSn = norm3(mn.x, mn.y, mn.z) * 0.5f * b_r * inv_I;
    const float MSt = norm3(ft.x, ft.y, ft.z) * inv_area;
    if (Sn + St > scrit || MSn + MSt > tcrit) {
      ILtype[s] = 0;
      st3(Oijt, s, 0.0f, 0.0f, 0.0f);
    }
  }
}

// ---- the superseded per-slot force path (md_calculateforcesdem.cu) ----
// Every slot's law goes to the slot arrays (Fijn, Fijt, Oijt, Mijn, Mijt,
// Mijadd, and the bond direction rij); d_SumUpForcesDEM (k_sumup) adds them
// per particle and d_CheckBreakConditionDEM (k_check_break) judges the bonds.
// Slots are independent, so one thread per slot.  MODE 0: d_CalculateForcesDEM
// (:328-499): bonds, empty slots zeroed, contact slots left alone (its contact
// branch is switched off by `&& false`, :361); MODE 1: d_CalculateForcesDEM_1
// (:501-579): contact slots, those beyond 2 rP and empty slots zeroed;
// MODE 2: d_CalculateForcesDEM_2 (:581-628): bonds only.
struct SlotArrays {
  float *rij, *Oijt, *Fijn, *Fijt, *Mijn, *Mijt, *Mijadd;
};

__device__ __forceinline__ void slot_zero(const SlotArrays &a, size_t s) {
  st3(a.Fijn, s, 0.f, 0.f, 0.f);
  st3(a.Fijt, s, 0.f, 0.f, 0.f);
  st3(a.Oijt, s, 0.f, 0.f, 0.f);
  st3(a.Mijn, s, 0.f, 0.f, 0.f);
  st3(a.Mijt, s, 0.f, 0.f, 0.f);
  st3(a.Mijadd, s, 0.f, 0.f, 0.f);
}

template <int MODE>
__global__ void __launch_bounds__(256)
k_forces_slots(const float *__restrict__ R, const float *__restrict__ V,
               const float *__restrict__ W, const uint32_t *__restrict__ IL,
               const uint8_t *__restrict__ ILtype, const float *__restrict__ iL, SlotArrays a,
               uint32_t N, uint32_t K, float dt, float E, float G, float b_r, float pA, float mu,
               float mur, float mP, float rP) {
  const size_t S = (size_t)N * K, N1 = N, N2 = 2 * (size_t)N;
  for (size_t s = blockIdx.x * (size_t)blockDim.x + threadIdx.x; s < S;
       s += (size_t)blockDim.x * gridDim.x) {
    const uint32_t i = (uint32_t)(s / K), j = IL[s];
    if (j == kSentinel) {
      if (MODE != 2) slot_zero(a, s);
      continue;
    }
    const uint8_t type = ILtype[s];
    // dd_Calculate_rijm_nij (:20-34)
    const float rx = R[j] - R[i], ry = R[j + N1] - R[i + N1], rz = R[j + N2] - R[i + N2];
    const float rm = norm3(rx, ry, rz);
    const float inv = 1.0f / rm;
    const float nx = rx * inv, ny = ry * inv, nz = rz * inv;
    if (type == 1 && MODE != 1) {
      const float il = iL[s];
      // dd_Calculate_fijn_bpm (:73-79)
      const float cef = b_r * b_r * kPiF * E * (rm - 1.0f / il) * il;
      st3(a.Fijn, s, nx * cef, ny * cef, nz * cef);
      // dd_Calculate_M123_bpm (:36-71)
      const f3 p = ld3(a.rij, s);
      const float tx = p.y * nz - p.z * ny, ty = p.z * nx - p.x * nz, tz = p.x * ny - p.y * nx;
      const float c2 = ((W[j] + W[i]) * nx + (W[j + N1] + W[i + N1]) * ny +
                        (W[j + N2] + W[i + N2]) * nz) *
                       dt * 0.5f;
      const float phx = nx * c2, phy = ny * c2, phz = nz * c2;
      st3(a.rij, s, nx, ny, nz);
      const float M1x = 1.0f + tz * phz + ty * phy, M1y = phz - tz - ty * phx,
                  M1z = -phy - tz * phx + ty;
      const float M2x = tz - phz - tx * phy, M2y = tz * phz + 1.0f + tx * phx,
                  M2z = -tz * phy + phx - tx;
      const float M3x = -ty - tx * phz + phy, M3y = -ty * phz + tx - phx,
                  M3z = ty * phy + tx * phx + 1.0f;
      // dd_Calculate_mijn_bpm (:112-127)
      const float dwx = W[j] - W[i], dwy = W[j + N1] - W[i + N1], dwz = W[j + N2] - W[i + N2];
      const float c1 = nx * dwx + ny * dwy + nz * dwz;
      float k2 = dt * kPiF * b_r * b_r * b_r * b_r * 0.5f * G * il;
      f3 m = ld3(a.Mijn, s);
      st3(a.Mijn, s, M1x * m.x + M1y * m.y + M1z * m.z + nx * c1 * k2,
          M2x * m.x + M2y * m.y + M2z * m.z + ny * c1 * k2,
          M3x * m.x + M3y * m.y + M3z * m.z + nz * c1 * k2);
      // dd_Calculate_mijt_bpm (:129-145)
      k2 = dt * kPiF * b_r * b_r * b_r * b_r * 0.25f * E * il;
      m = ld3(a.Mijt, s);
      st3(a.Mijt, s, M1x * m.x + M1y * m.y + M1z * m.z + (dwx - nx * c1) * k2,
          M2x * m.x + M2y * m.y + M2z * m.z + (dwy - ny * c1) * k2,
          M3x * m.x + M3y * m.y + M3z * m.z + (dwz - nz * c1) * k2);
      // dd_Calculate_fijt_mijunsym_bpm (:147-187)
      const float hx = 0.5f * (W[j] + W[i]), hy = 0.5f * (W[j + N1] + W[i + N1]),
                  hz = 0.5f * (W[j + N2] + W[i + N2]);
      const float vx = V[j] - V[i] - hy * rz + hz * ry;
      const float vy = V[j + N1] - V[i + N1] - hz * rx + hx * rz;
      const float vz = V[j + N2] - V[i + N2] - hx * ry + hy * rx;
      const float cv = nx * vx + ny * vy + nz * vz;
      const f3 o0 = ld3(a.Oijt, s);
      const float ox = (M1x * o0.x + M1y * o0.y + M1z * o0.z) + (vx - nx * cv) * dt;
      const float oy = (M2x * o0.x + M2y * o0.y + M2z * o0.z) + (vy - ny * cv) * dt;
      const float oz = (M3x * o0.x + M3y * o0.y + M3z * o0.z) + (vz - nz * cv) * dt;
      const float ct = G * b_r * b_r * kPiF * il;
      st3(a.Fijt, s, ox * ct, oy * ct, oz * ct);
      st3(a.Oijt, s, ox, oy, oz);
      st3(a.Mijadd, s, 0.5f * (ry * oz - rz * oy) * ct, 0.5f * (rz * ox - rx * oz) * ct,
          0.5f * (rx * oy - ry * ox) * ct);
    } else if (type == 0 && MODE == 1) {
      if (!(rm < 2.0f * rP)) {
        slot_zero(a, s);
        continue;
      }
      // dd_Calculate_fijn_ppHM (:189-243)
      const float q = sqrtf(rP * fabsf(rP - 0.5f * rm));
      const float Kn = 2.0f * E * q, Kt = 8.0f * G * q;
      const float vx =
          V[j] - V[i] + rP * (ny * (W[j + N2] + W[i + N2]) - nz * (W[j + N1] + W[i + N1]));
      const float vy =
          V[j + N1] - V[i + N1] + rP * (nz * (W[j] + W[i]) - nx * (W[j + N2] + W[i + N2]));
      const float vz =
          V[j + N2] - V[i + N2] + rP * (nx * (W[j + N1] + W[i + N1]) - ny * (W[j] + W[i]));
      const float vn = nx * vx + ny * vy + nz * vz;
      const float fn = (rm - 2.0f * rP) * Kn * kTwoThirdsF;
      const float dmp = -1.8257f * pA * vn * sqrtf(Kn * 0.5f * mP);
      st3(a.Fijn, s, nx * (fn + dmp), ny * (fn + dmp), nz * (fn + dmp));
      const float vtx = vx - nx * vn, vty = vy - ny * vn, vtz = vz - nz * vn;
      // dd_Calculate_fijtmijt_ppHM (:245-304)
      const f3 O = ld3(a.Oijt, s);
      float c1 = nx * O.x + ny * O.y + nz * O.z;
      float ox = O.x - nx * c1, oy = O.y - ny * c1, oz = O.z - nz * c1;
      c1 = ox * ox + oy * oy + oz * oz;
      if ((double)c1 > 1e-18) {
        c1 = 1.0f / c1;
        c1 = c1 * (O.x * O.x + O.y * O.y + O.z * O.z);
        c1 = sqrtf(c1);
        ox = ox * c1;
        oy = oy * c1;
        oz = oz * c1;
      }
      ox = ox + vtx * dt;
      oy = oy + vty * dt;
      oz = oz + vtz * dt;
      float fx = ox * Kt, fy = oy * Kt, fz = oz * Kt;
      const float ftm = fx * fx + fy * fy + fz * fz;
      if (ftm > mu * mu * fn * fn) {
        c1 = mu * fabsf(fn) * rsqrt_rn(ftm);
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
      st3(a.Oijt, s, ox, oy, oz);
      st3(a.Fijt, s, fx, fy, fz);
      st3(a.Mijt, s, (ny * fz - nz * fy) * rP, (nz * fx - nx * fz) * rP, (nx * fy - ny * fx) * rP);
      st3(a.Mijn, s, 0.f, 0.f, 0.f);
      // dd_Calculate_mijroll_ppHM (:306-326)
      const float wx = W[i], wy = W[i + N1], wz = W[i + N2];
      const float wim = wx * wx + wy * wy + wz * wz;
      if ((double)wim > 1e-18) {
        const float cr = -mur * fabsf(fn) * rsqrt_rn(wim);
        st3(a.Mijadd, s, wx * cr, wy * cr, wz * cr);
      } else {
        st3(a.Mijadd, s, 0.f, 0.f, 0.f);
      }
    }
  }
}

}  // namespace

extern "C" {

int rfd_forces_bpm(rfd_ctx *ctx, const float *R, const float *V, const float *W,
                   const uint32_t *IL, uint8_t *ILtype, const float *iL, float *rij,
                   float *Oijt, float *Mijn, float *Mijt, float *F, float *M, uint32_t N,
                   uint32_t K, float dt, float E, float G, float b_r, float sc, float tc) {
  if (!ctx) return RFD_E_ARG;
  RFD_CHECK_ARG(ctx, R && V && W && IL && ILtype && iL && rij && Oijt && Mijn && Mijt && F &&
                         M && N > 0 && K > 0 && (uint64_t)N * K < 0xFFFFFFFFull,
                "rfd_forces_bpm: bad arguments");
  RFD_CHECK_ARG(ctx, K <= (uint32_t)kMaxK, "rfd_forces_bpm: K=%u > %d", K, kMaxK);
  const uint32_t P = bpm_rows(K);
  const unsigned blocks = (N + P - 1) / P;
  const uint32_t per = (blocks + kXcds - 1) / kXcds;
  k_forces_bpm<<<per * kXcds, kSlotThreads, 0, ctx->stream>>>(
      R, V, W, IL, ILtype, iL, rij, Oijt, Mijn, Mijt, F, M, N, K, dt, E, G, b_r, sc, tc, blocks,
      per);
  RFD_LAUNCHED(ctx, "rfd_forces_bpm");
  return RFD_OK;
}

int rfd_forces_contact(rfd_ctx *ctx, const float *R, const float *V, const float *W,
                       const uint32_t *IL, const uint8_t *ILtype, float *Oijt, float *F,
                       float *M, uint32_t N, uint32_t K, float dt, float Estar, float Gstar,
                       float pA, float mu, float mur, float m, float r) {
  if (!ctx) return RFD_E_ARG;
  RFD_CHECK_ARG(ctx, R && V && W && IL && ILtype && Oijt && F && M && N > 0 && K > 0 &&
                         (uint64_t)N * K < 0xFFFFFFFFull,
                "rfd_forces_contact: bad arguments");
  RFD_CHECK_ARG(ctx, K <= (uint32_t)kMaxK, "rfd_forces_contact: K=%u > %d", K, kMaxK);
  const ContactLaw L{dt, Estar, Gstar, pA, mu, mur, m, r};
  const bool by4 = K % 4 == 0 && reinterpret_cast<uintptr_t>(IL) % 16 == 0 &&
                   reinterpret_cast<uintptr_t>(Oijt) % 16 == 0 &&
                   reinterpret_cast<uintptr_t>(ILtype) % 4 == 0;
  const uint32_t P = (by4 ? kContact4Slots : kSlotThreads) / K;
  const unsigned blocks = (N + P - 1) / P;
  const uint32_t per = (blocks + kXcds - 1) / kXcds;
  if (by4)
    k_forces_contact4<<<per * kXcds, kSlotThreads, 0, ctx->stream>>>(R, V, W, IL, ILtype, Oijt,
                                                                      F, M, N, K, L, blocks, per);
  else
    k_forces_contact<<<per * kXcds, kSlotThreads, 0, ctx->stream>>>(R, V, W, IL, ILtype, Oijt, F,
                                                                     M, N, K, L, blocks, per);
  RFD_LAUNCHED(ctx, "rfd_forces_contact");
  return RFD_OK;
}

int rfd_forces_slots(rfd_ctx *ctx, int mode, const float *R, const float *V, const float *W,
                     const uint32_t *IL, const uint8_t *ILtype, const float *iL, float *rij,
                     float *Oijt, float *Fijn, float *Fijt, float *Mijn, float *Mijt,
                     float *Mijadd, uint32_t N, uint32_t K, float dt, float E, float G,
                     float b_r, float pA, float mu, float mur, float mP, float rP) {
  if (!ctx) return RFD_E_ARG;
  RFD_CHECK_ARG(ctx, R && V && W && IL && ILtype && iL && rij && Oijt && Fijn && Fijt && Mijn &&
                         Mijt && Mijadd && N > 0 && K > 0 &&
                         (uint64_t)N * K < 0xFFFFFFFFull,
                "rfd_forces_slots: bad arguments");
  RFD_CHECK_ARG(ctx, mode >= 0 && mode <= 2, "rfd_forces_slots: mode %d not 0, 1 or 2", mode);
  const SlotArrays a{rij, Oijt, Fijn, Fijt, Mijn, Mijt, Mijadd};
  const unsigned nb = rfd_blocks((size_t)N * K, 256, 65536);
  hipStream_t q = ctx->stream;
  if (mode == 0)
    k_forces_slots<0><<<nb, 256, 0, q>>>(R, V, W, IL, ILtype, iL, a, N, K, dt, E, G, b_r, pA, mu,
                                          mur, mP, rP);
  else if (mode == 1)
    k_forces_slots<1><<<nb, 256, 0, q>>>(R, V, W, IL, ILtype, iL, a, N, K, dt, E, G, b_r, pA, mu,
                                          mur, mP, rP);
  else
    k_forces_slots<2><<<nb, 256, 0, q>>>(R, V, W, IL, ILtype, iL, a, N, K, dt, E, G, b_r, pA, mu,
                                          mur, mP, rP);
  RFD_LAUNCHED(ctx, "rfd_forces_slots");
  return RFD_OK;
}

int rfd_sumup_forces(rfd_ctx *ctx, const float *Fijn, const float *Fijt, const float *Mijn,
                     const float *Mijt, const float *Mijadd, float *F, float *M, uint32_t N,
                     uint32_t K) {
  if (!ctx) return RFD_E_ARG;
  RFD_CHECK_ARG(ctx, Fijn && Fijt && Mijn && Mijt && Mijadd && F && M && N > 0 && K > 0,
                "rfd_sumup_forces: bad arguments");
  k_sumup<<<rfd_blocks(N, 256), 256, 0, ctx->stream>>>(Fijn, Fijt, Mijn, Mijt, Mijadd, F, M, N,
                                                        K);
  RFD_LAUNCHED(ctx, "rfd_sumup_forces");
  return RFD_OK;
}

int rfd_check_break(rfd_ctx *ctx, const float *rij, float *Oijt, const float *Fijn,
                    const float *Fijt, const float *Mijn, const float *Mijt, uint8_t *ILtype,
                    uint32_t S, float sc, float tc, float b_r) {
  if (!ctx) return RFD_E_ARG;
  RFD_CHECK_ARG(ctx, rij && Oijt && Fijn && Fijt && Mijn && Mijt && ILtype && S > 0,
                "rfd_check_break: bad arguments");
  k_check_break<<<rfd_blocks(S, 256), 256, 0, ctx->stream>>>(rij, Oijt, Fijn, Fijt, Mijn, Mijt,
                                                              ILtype, S, sc, tc, b_r);
  RFD_LAUNCHED(ctx, "rfd_check_break");
  return RFD_OK;
}

}  // extern "C"
