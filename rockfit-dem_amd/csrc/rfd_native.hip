// rfd_native.hip -- native tier: AoS particle records (FP64 position +
// radius), bonds stored as two mirrored half-bonds in per-particle rows,
// slot-major contact rows, fused platen + integration pass, device-side
// loading control.
//
// Memory layout (all in HBM, sized for one GPU's share of a specimen):
//   particle  P4 {x, y, z, r} (double, 32 B), V4 / W4 / F4 / M4 (float4),
//             inv2 {dt/m, dt/I} (float2)
//   half-bond rows hoff[N+1]; per half h: HR {j, b} (8 B) -- the other end
//             and the bond index
//   bonds     per bond b (sorted by (i, j), i < j): BR {1/L0, br, sc, tc} (16 B),
//             state BSt {O, mn, Mt, flag} (32 B, two buffers: read one, the
//             owner writes the other; flag 1 bond, 0 broken, 2 removed)
//   contacts  slot-major rows cj[k*N+i], cO[k*N+i] (float4), ccnt[i]
// A bond (i, j) appears in row i and in row j, as the reference holds one
// directed slot per end, but its constants and state are stored once.  Both
// halves evaluate the law, as exact mirror images (every operation is odd or
// even under i<->j and IEEE rounding is sign-symmetric), so their loads are
// exact opposites and they break together; the owner half (row i < j) writes
// the new state.  The state is kept in i's frame: the j half reads O and Mt
// negated.  The twisting moment is kept as its component mn along n, and the
// previous bond direction is recomputed from the positions one move back
// (R - V dt), so the record is 32 B instead of the 48 B of n_prev, O, Mn, Mt.
//
// Force pass: one workgroup per 64 particles; their half-bonds form one
// contiguous range, evaluated one per lane in chunks of 512; each lane writes
// its load to LDS and each particle's thread then adds its own halves in row
// order.  No atomics: the pass is deterministic.
//
// Bond law: md_calculateforcesdem3.cu:38-197 with per-bond radius; stretch in
// FP64 from FP64 positions, the rest in FP32, break criterion (:20-36) in
// FP64.  A bond that breaks also feels the contact law in the same pass,
// because the reference runs its contact kernel after the bond kernel on the
// just-freed slot (md_calculationuniaxialcompression.cu:173-176).  Broken
// records act through the Hertz-Mindlin law of :199-336 with R* = ri rj /
// (ri + rj) and m* = mi mj / (mi + mj), which is the reference's law for
// equal radii.
#include <algorithm>
#include <chrono>
#include <cstddef>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include <hip/hip_cooperative_groups.h>

#include "rfd_internal.h"

// The native tier is compared with its oracles at a stated tolerance, not bit
// for bit: products and sums may contract to FMA here (every contraction is
// sign-symmetric, so the two halves of a bond stay exact mirror images).
#pragma clang fp contract(fast)

using namespace rfd;

namespace {

// force-pass tile shapes (particles per workgroup, threads): the default and
// the alternatives the autotuner times; results do not depend on the choice
constexpr int kFP = 64;    // particles per force workgroup (default shape)
constexpr int kFT = 256;   // threads per force workgroup (default shape)
constexpr int kCH = 512;   // bond ends per LDS chunk
constexpr uint8_t kBond = 1, kBroken = 0, kRemoved = 2;
constexpr int kHistCap = 1 << 20;
constexpr uint32_t kMaxBreakRounds = 200;  // relax-and-break rounds per load step
// gfx950 places workgroup b on XCD b % 8; tiles are remapped so that each XCD
// (own L2) works through one contiguous run of Morton-ordered particle tiles,
// which keeps the neighbour gathers of adjacent tiles in the same L2
constexpr uint32_t kXcd = 8;
// force passes of at most this many tiles (one per CU) run the loading
// control themselves, and load each tile's first records ahead of its barrier
constexpr uint32_t kCtlInPassTiles = 256;
__device__ __forceinline__ uint32_t xcd_tile(uint32_t per) {
  return (blockIdx.x % kXcd) * per + blockIdx.x / kXcd;
}

struct alignas(32) P4 {
  double x, y, z, r;
};
struct alignas(16) BSt {
  float4 a;        // O (shear spring, i's frame), mn (twist along n)
  float mx, my, mz;  // Mt (bending, i's frame)
  uint32_t fl;     // kBond, kBroken, kRemoved
};
struct alignas(8) HR {
  uint32_t j, b;  // the other end, the bond
};
struct alignas(16) BR {
  float il, br, sc, tc;  // 1 / rest length, bond radius, tensile and shear strength
};

struct v3 {
  float x, y, z;
};
__device__ __forceinline__ v3 operator+(v3 a, v3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ v3 operator-(v3 a, v3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ v3 operator-(v3 a) { return {-a.x, -a.y, -a.z}; }
__device__ __forceinline__ v3 operator*(v3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ float dot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ v3 cross(v3 a, v3 b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ double dnorm(v3 a) {
  double x = a.x, y = a.y, z = a.z;
  return __builtin_sqrt(x * x + y * y + z * z);
}
__device__ __forceinline__ v3 xyz(float4 q) { return {q.x, q.y, q.z}; }
__device__ __forceinline__ float4 f4(v3 v) { return make_float4(v.x, v.y, v.z, 0.f); }
__device__ constexpr v3 kZero = {0.f, 0.f, 0.f};

struct Mat {
  float E, G, Estar, Gstar, pA, mu, mur, rho43pi, dt;
  double farcut;  // far-link distance as a multiple of ri + rj
  int refmode;
  int nobreak;    // quasi-static relaxation: bonds are judged at equilibrium only
};

// first-order rotation of md_calculateforcesdem3.cu:80-88 built from
// t = n_prev x n and phi
struct Rot {
  float m[9];
  __device__ __forceinline__ void build(v3 t, v3 ph) {
    m[0] = 1.0f + t.z * ph.z + t.y * ph.y;
    m[1] = ph.z - t.z - t.y * ph.x;
    m[2] = -ph.y - t.z * ph.x + t.y;
    m[3] = t.z - ph.z - t.x * ph.y;
    m[4] = t.z * ph.z + 1.0f + t.x * ph.x;
    m[5] = -t.z * ph.y + ph.x - t.x;
    m[6] = -t.y - t.x * ph.z + ph.y;
    m[7] = -t.y * ph.z + t.x - ph.x;
    m[8] = t.y * ph.y + t.x * ph.x + 1.0f;
  }
  __device__ __forceinline__ v3 apply(v3 x) const {
    return {m[0] * x.x + m[1] * x.y + m[2] * x.z, m[3] * x.x + m[4] * x.y + m[5] * x.z,
            m[6] * x.x + m[7] * x.y + m[8] * x.z};
  }
};

struct EndData {
  P4 P;
  v3 V, W;
};

// Hertz-Mindlin contact for the pair (i, j) (md_calculateforcesdem3.cu:199-336),
// n from i to j, overlap ri + rj - L; O is i's tangential spring (in/out).
// Returns the load on i and the contact moments on i and j.
// reach: the centre distance at which the pair starts to overlap (ri + rj,
// or less for a broken bond made while overlapping)
// dl is the overlap reach - L (> 0), formed by the caller from the FP64
// positions: for a pair that just touches it is far smaller than either
// length, and its FP32 difference would lose most of its digits
__device__ __forceinline__ void contact_law(const Mat &mt, v3 n, float dl, float ri, float rj,
                                            v3 Vi, v3 Vj, v3 Wi, v3 Wj, v3 &O, v3 &Fi,
                                            v3 &Mi, v3 &Mj) {
  const float Rs = ri * rj / (ri + rj);
  const float mi = mt.rho43pi * ri * ri * ri, mj = mt.rho43pi * rj * rj * rj;
  const float ms = mi * mj / (mi + mj);
  const float q = sqrtf(Rs * dl);
  const float Kn = 2.0f * mt.Estar * q;
  const float Kt = 8.0f * mt.Gstar * q;
  const v3 v = Vj - Vi + cross(n, Wi * ri + Wj * rj);
  const float vn = dot(n, v);
  const float fn = -dl * Kn * kTwoThirdsF;
  const float dmp = -1.8257f * mt.pA * vn * sqrtf(Kn * ms);
  Fi = n * (fn + dmp);
  const v3 vt = v - n * vn;
  v3 o = O - n * dot(n, O);  // rotate the old spring into the tangent plane
  const float oo = dot(o, o);
  if ((double)oo > 1e-18) o = o * sqrtf(dot(O, O) / oo);
  o = o + vt * mt.dt;
  v3 ft = o * Kt;
  const float ftm = dot(ft, ft);
  if (ftm > mt.mu * mt.mu * fn * fn) {  // Coulomb cap
    ft = ft * (mt.mu * fabsf(fn) * rsqrt_rn(ftm));
    o = ft * (1.0f / Kt);
  } else {
    ft = ft + vt * (-1.8257f * mt.pA * sqrtf(Kt * ms));
  }
  O = o;
  Fi = Fi + ft;
  const v3 nxf = cross(n, ft);
  Mi = nxf * ri;
  Mj = nxf * rj;
  const float wi2 = dot(Wi, Wi), wj2 = dot(Wj, Wj);
  if ((double)wi2 > 1e-18) Mi = Mi + Wi * (-mt.mur * fabsf(fn) * rsqrt_rn(wi2));
  if ((double)wj2 > 1e-18) Mj = Mj + Wj * (-mt.mur * fabsf(fn) * rsqrt_rn(wj2));
}

struct BondOut {
  v3 Fi;        // load on i (j gets -Fi)
  v3 Mbi, Mbj;  // bond moments on i and j
  v3 Mci, Mcj;  // contact moments on i and j
  bool hasc;    // the contact law acted (counts for the rolling-wipe rule)
  BSt st;       // new state, i's frame of the owner half
};

// evaluate bond (A, B) from its old state, A the lower-index end (the state
// is kept in its frame).  Both halves of a bond call this with the same
// orientation and take their own end's loads (B's force is -Fi), so the two
// ends, and any tiling of the pass, see the same bits.
__device__ __forceinline__ void eval_bond(const Mat &mt, const EndData &A, const EndData &B,
                                          const BR &c, const BSt &old, BondOut &o) {
  const double rx = B.P.x - A.P.x, ry = B.P.y - A.P.y, rz = B.P.z - A.P.z;
  const double r2 = rx * rx + ry * ry + rz * rz;
  o.Fi = o.Mbi = o.Mbj = o.Mci = o.Mcj = kZero;
  o.hasc = false;
  o.st = old;
  // far-link removal (d_DeleteFarLinks, md_interaction_list_reconstructbpm.cu:
  // 141-186): a record longer than farcut*(ri+rj), or of coincident ends, is
  // removed; both halves see the same distance and go together
  const double lim = mt.farcut * (A.P.r + B.P.r);
  if (r2 > lim * lim || !(r2 > 1e-10)) {
    o.st.fl = kRemoved;
    return;
  }
  const v3 r = {(float)rx, (float)ry, (float)rz};
  const float r2f = (float)r2;
  const float invL = __builtin_amdgcn_rsqf(r2f);
  const float L = r2f * invL;
  const v3 n = r * invL;
  const float ri = (float)A.P.r, rj = (float)B.P.r;
  v3 O = xyz(old.a);
  bool contact_now = old.fl == kBroken;
  if (old.fl == kBond) {
    // bond direction one move back: both ends at R - V dt
    const v3 rp = r - (B.V - A.V) * mt.dt;
    const float pp = dot(rp, rp);
    const v3 np = pp > 0.f ? rp * __builtin_amdgcn_rsqf(pp) : n;
    const v3 Mt0 = v3{old.mx, old.my, old.mz};
    const float il = c.il;
    const float A2 = kPiF * c.br * c.br, J = kPiF * c.br * c.br * c.br * c.br;
    // stretch without cancellation: eps = (L^2/L0^2 - 1) / (1 + L/L0), the
    // numerator from the FP64 squared length
    const double ild = il;
    const float eps = (float)(r2 * ild * ild - 1.0) * __builtin_amdgcn_rcpf(1.0f + L * il);
    const float sig32 = mt.E * eps;
    const v3 Fn = n * (sig32 * A2);
    const v3 ws = A.W + B.W;
    Rot M;
    M.build(cross(np, n), n * (dot(ws, n) * mt.dt * 0.5f));
    const v3 dw = B.W - A.W;
    const float dwn = dot(n, dw);
    const float mn = old.a.w + dwn * (mt.dt * 0.5f * mt.G * il);
    const v3 Mn = n * mn;
    const float cb = mt.dt * 0.25f * mt.E * il;
    const v3 Mt = M.apply(Mt0) + (dw - n * dwn) * cb;
    const v3 v = B.V - A.V - cross(ws * 0.5f, r);
    const float vn = dot(n, v);
    O = M.apply(O) + (v - n * vn) * mt.dt;
    const float kt = mt.G * il * A2;
    const v3 Ft = O * kt;
    const float mtf = mt.refmode ? cb * J : J;
    const v3 half = cross(r, Ft) * 0.5f;
    o.Fi = Fn + Ft;
    o.Mbi = Mn * J + Mt * mtf + half;
    o.Mbj = half - Mn * J - Mt * mtf;
    // break criterion (dd_CheckBreakConditionDEM_2, :20-36): evaluated in FP32,
    // and again in FP64 when FP32 lands within 1e-4 of a threshold, so the
    // decision at the threshold is the FP64 one
    const float Sc32 = sig32 + __builtin_amdgcn_sqrtf(dot(Mt, Mt)) * 4.0f * c.br;
    const float Mc32 = mt.G * il * __builtin_amdgcn_sqrtf(dot(O, O)) + fabsf(mn) * 2.0f * c.br;
    bool broken = Sc32 > c.sc || Mc32 > c.tc;
    if (fabsf(Sc32 - c.sc) <= 1e-4f * c.sc || fabsf(Mc32 - c.tc) <= 1e-4f * c.tc) {
      const double eps64 = __builtin_sqrt(r2) * ild - 1.0;
      const double Sc = (double)mt.E * eps64 + dnorm(Mt) * 4.0 * (double)c.br;
      const double Mc = (double)mt.G * ild * dnorm(O) + fabs((double)mn) * 2.0 * (double)c.br;
      broken = Sc > (double)c.sc || Mc > (double)c.tc;
    }
    if (mt.nobreak) broken = false;
    if (broken) {
      o.st.fl = kBroken;
      O = kZero;
      o.st.a.w = 0.f;
      o.st.mx = o.st.my = o.st.mz = 0.f;
      contact_now = true;  // the freed pair meets the contact law this pass
    } else {
      o.st.a = make_float4(O.x, O.y, O.z, mn);
      o.st.mx = Mt.x;
      o.st.my = Mt.y;
      o.st.mz = Mt.z;
      return;
    }
  }
  if (contact_now) {
    // a broken bond meets the contact law at its rest length when it was made
    // overlapping, so the overlap it was made with does not spring apart
    // (reference mode: at ri + rj, as md_calculateforcesdem3.cu does)
    const double reach = mt.refmode ? A.P.r + B.P.r : fmin(A.P.r + B.P.r, 1.0 / (double)c.il);
    const double L64 = __builtin_sqrt(r2);
    if (L64 < reach) {
      v3 Fc;
      contact_law(mt, n, (float)(reach - L64), ri, rj, A.V, B.V, A.W, B.W, O, Fc, o.Mci,
                  o.Mcj);
      o.Fi = o.Fi + Fc;
      o.hasc = true;
    } else {
      O = kZero;
    }
    o.st.a.x = O.x;
    o.st.a.y = O.y;
    o.st.a.z = O.z;
  }
}

struct SysDev {
  uint32_t N, Kc;
  const P4 *P;
  const float4 *V, *W;
  float4 *F, *M;
  const uint32_t *hoff;
  const HR *hr;
  const BR *brc;
  BSt *bst0, *bst1;   // the two state buffers
  const int *curp;    // device index of the buffer holding the latest state
  uint8_t *fl8; // per-bond flag copy, written by the owner half when it changes
  const uint32_t *cj, *ccnt;
  float4 *cO;
  const float *dtp;  // when set, the time step of the last move (FIRE) replaces mt.dt
  const int *done;   // when set and nonzero (relaxation finished), the pass does nothing
  int trial;         // trial pass: writes only F, M and st_out (nothing the next pass reads)
  const uint32_t *hm;  // mirror half of each half
  uint32_t *nbreak;  // when set, counts the bonds this pass breaks
};

__device__ __forceinline__ EndData load_end(const SysDev &s, uint32_t i) {
  EndData e;
  e.P = s.P[i];
  e.V = xyz(s.V[i]);
  e.W = xyz(s.W[i]);
  return e;
}

template <int FP, int FT>
__global__ void __launch_bounds__(FT) __attribute__((amdgpu_waves_per_eu(5, 8)))
k_nat_forces(SysDev s, Mat mt, uint32_t ntiles, uint32_t per) {
  // the row offsets of a tile (FP + 1 of them) are loaded one per thread
  static_assert(FP < FT, "a force workgroup needs more threads than particles");
  __shared__ float sF[3][kCH], sMb[3][kCH], sMc[3][kCH];
  __shared__ uint8_t sC[kCH];
  __shared__ uint32_t sOff[FP + 1];
  __shared__ float sAcc[3][3][FP];  // F, Mb, Mc running sums per particle
  __shared__ uint8_t sAnyc[FP];
  __shared__ P4 sP[FP];
  __shared__ float4 sV[FP], sW[FP];
  const uint32_t tile = xcd_tile(per);
  if (tile >= ntiles) return;  // padding tiles of the last XCD run
  if (s.done && *s.done) return;
  if (s.dtp) mt.dt = *s.dtp;
  // the pass reads the latest state and writes the other buffer (the index
  // flips on the device after the pass, only when the pass ran)
  BSt *const st_in = *s.curp ? s.bst1 : s.bst0;
  BSt *const st_out = *s.curp ? s.bst0 : s.bst1;
  const uint32_t tid = threadIdx.x;
  const uint32_t p0 = tile * FP;
  const uint32_t np = min((uint32_t)FP, s.N - p0);
  if (tid <= np) sOff[tid] = s.hoff[p0 + tid];
  if (tid < np) {
    sP[tid] = s.P[p0 + tid];
    sV[tid] = s.V[p0 + tid];
    sW[tid] = s.W[p0 + tid];
  }
  for (uint32_t q = tid; q < 9 * FP; q += FT) (&sAcc[0][0][0])[q] = 0.f;
  if (tid < FP) sAnyc[tid] = 0;
  __syncthreads();
  const uint32_t h0 = sOff[0], n_end = sOff[np] - h0;
  v3 F = kZero, Mb = kZero, Mc = kZero;
  bool anyc = false;
  for (uint32_t c0 = 0; c0 < n_end; c0 += kCH) {
    const uint32_t c1 = min(c0 + (uint32_t)kCH, n_end);
    for (uint32_t e = c0 + tid; e < c1; e += FT) {
      const uint32_t h = h0 + e;
      // row particle of half h: last row whose start is <= h
      uint32_t lo = 0, hi = np;
      while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (sOff[mid] <= h) lo = mid; else hi = mid;
      }
      const HR hr = s.hr[h];
      const uint32_t i = p0 + lo;
      const BSt old = st_in[hr.b];
      BondOut o;
      const bool own = i < hr.j;
      if (old.fl == kRemoved) {
        o.Fi = o.Mbi = o.Mbj = o.Mci = o.Mcj = kZero;
        o.hasc = false;
      } else {
        EndData me;
        me.P = sP[lo];
        me.V = xyz(sV[lo]);
        me.W = xyz(sW[lo]);
        const EndData other = load_end(s, hr.j);
        eval_bond(mt, own ? me : other, own ? other : me, s.brc[hr.b], old, o);
        if (own) {
          if (o.st.fl == kRemoved) {
            // both buffers: no later pass may read a stale live flag
            st_out[hr.b].fl = kRemoved;
            if (!s.trial) st_in[hr.b].fl = kRemoved;
          } else {
            st_out[hr.b] = o.st;
          }
          if (!s.trial && o.st.fl != old.fl) {
            s.fl8[hr.b] = (uint8_t)o.st.fl;
            if (s.nbreak && o.st.fl == kBroken) atomicAdd(s.nbreak, 1u);
          }
        }
      }
      const uint32_t k = e - c0;
      // this end's loads; the physical mode keeps bond and contact moments in
      // one sum (the reference mode needs the contact part for its wipe rule)
      const v3 Fm = own ? o.Fi : -o.Fi;
      v3 Mb = own ? o.Mbi : o.Mbj;
      v3 Mcx = own ? o.Mci : o.Mcj;
      if (!mt.refmode) {
        Mb = Mb + Mcx;
        Mcx = kZero;
      }
      sF[0][k] = Fm.x;
      sF[1][k] = Fm.y;
      sF[2][k] = Fm.z;
      sMb[0][k] = Mb.x;
      sMb[1][k] = Mb.y;
      sMb[2][k] = Mb.z;
      sMc[0][k] = Mcx.x;
      sMc[1][k] = Mcx.y;
      sMc[2][k] = Mcx.z;
      sC[k] = o.hasc ? 1 : 0;
    }
    __syncthreads();
    // per-particle sums of this chunk: task (component group g, particle p),
    // each adds its particle's entries in row order
    for (uint32_t task = tid; task < 4 * np; task += FT) {
      const uint32_t p = task % np, g = task / np;
      const uint32_t a0 = max(sOff[p] - h0, c0), a1 = min(sOff[p + 1] - h0, c1);
      if (g < 3) {
        float (*src)[kCH] = g == 0 ? sF : (g == 1 ? sMb : sMc);
        // one running sum over the whole row, whatever the chunking: the
        // result does not depend on the tile shape
        float x = sAcc[g][0][p], y = sAcc[g][1][p], z = sAcc[g][2][p];
        for (uint32_t e = a0; e < a1; ++e) {
          const uint32_t k = e - c0;
          x = x + src[0][k];
          y = y + src[1][k];
          z = z + src[2][k];
        }
        sAcc[g][0][p] = x;
        sAcc[g][1][p] = y;
        sAcc[g][2][p] = z;
      } else {
        uint8_t c = 0;
        for (uint32_t e = a0; e < a1; ++e) c |= sC[e - c0];
        sAnyc[p] |= c;
      }
    }
    __syncthreads();
  }
  if (tid < np) {
    F = v3{sAcc[0][0][tid], sAcc[0][1][tid], sAcc[0][2][tid]};
    Mb = v3{sAcc[1][0][tid], sAcc[1][1][tid], sAcc[1][2][tid]};
    Mc = v3{sAcc[2][0][tid], sAcc[2][1][tid], sAcc[2][2][tid]};
    anyc = sAnyc[tid] != 0;
  }
  if (tid >= np) return;
  // never-bonded contact rows of this particle (per-end spring, in place)
  const uint32_t i = p0 + tid, N = s.N;
  EndData me;
  me.P = sP[tid];
  me.V = xyz(sV[tid]);
  me.W = xyz(sW[tid]);
  const uint32_t nc = s.ccnt[i];
  for (uint32_t k = 0; k < nc; ++k) {
    const size_t sl = (size_t)k * N + i;
    const uint32_t j = s.cj[sl];
    const EndData other = load_end(s, j);
    const double rx = other.P.x - me.P.x, ry = other.P.y - me.P.y, rz = other.P.z - me.P.z;
    const double L64 = __builtin_sqrt(rx * rx + ry * ry + rz * rz);
    const float ri = (float)me.P.r, rj = (float)other.P.r;
    const double reach = me.P.r + other.P.r;
    if (L64 < reach && L64 > 1e-5) {
      const double inv = 1.0 / L64;
      const v3 n = {(float)(rx * inv), (float)(ry * inv), (float)(rz * inv)};
      v3 O = xyz(s.cO[sl]);
      v3 Fi, Mi, Mj;
      contact_law(mt, n, (float)(reach - L64), ri, rj, me.V, other.V, me.W, other.W, O, Fi,
                  Mi, Mj);
      F = F + Fi;
      Mc = Mc + Mi;
      anyc = true;
      if (!s.trial) s.cO[sl] = f4(O);
    } else if (!s.trial) {
      s.cO[sl] = make_float4(0.f, 0.f, 0.f, 0.f);
    }
  }
  // rolling-moment wipe of md_calculateforcesdem3.cu:330-335: with W_i ~ 0 the
  // reference clears the whole contact moment accumulator of the particle
  if (mt.refmode && anyc && !((double)dot(me.W, me.W) > 1e-18)) Mc = kZero;
  s.F[i] = f4(F);
  s.M[i] = f4(Mb + Mc);
}

// Largest half-bond count of one tile in the single-evaluation pass: the
// tile's per-half loads stay in LDS for the whole pass, in dynamic LDS sized
// at launch to the system's largest tile (systems with larger tiles run the
// chunked pass above)
template <int FP> struct OnceCap;
template <> struct OnceCap<32> { static constexpr uint32_t v = 1024; };
template <> struct OnceCap<64> { static constexpr uint32_t v = 2040; };
template <> struct OnceCap<128> { static constexpr uint32_t v = 2040; };

// Work record of the single-evaluation pass, one per evaluated half-bond of a
// tile, built once per tile shape (k_work_*): a = slot e (11 bits) | row
// particle lo << 11 (7 bits) | partner's local index << 18 (7 bits, in-tile
// records); b = bond; c = the mirror half's slot (in-tile) or the partner id.
struct WorkRec {
  uint32_t a, b, c;
};

struct WorkDev {
  const WorkRec *rec;
  const uint32_t *ofs;   // ntiles + 1 record offsets
  const uint32_t *n_in;  // in-tile records per tile (they come first)
};

// 1/sqrt for the platen friction: correctly rounded in FP32, plain in FP64
__device__ __forceinline__ float platen_rsqrt(float x) { return rsqrt_rn(x); }
__device__ __forceinline__ double platen_rsqrt(double x) { return 1.0 / __builtin_sqrt(x); }

// Platen reactions of the FP32 explicit steps are summed in fixed point:
// each particle's reaction is rounded to a multiple of 2^-40 force units
// (1e3 N x 2^-40, about 1e-9 N) and the sums are 64-bit integers, so that
// they do not depend on the order of summation (tile shape, fused or separate
// pass, workgroup schedule) and can be gathered with integer atomics.  Sums
// up to 2^22 force units (4e9 N) are represented exactly.
constexpr double kFxScale = 0x1p40, kFxInv = 0x1p-40;
__device__ __forceinline__ long long to_fx(double f) {
  const double x = fmin(fmax(f * kFxScale, -0x1p62), 0x1p62);  // NaN goes to -2^62
  return (long long)__builtin_rint(x);
}
__device__ __forceinline__ void acc_add(double &a, double f) { a += f; }
__device__ __forceinline__ void acc_add(long long &a, double f) { a += to_fx(f); }

// wave-wide integer sum, then one atomic per wave that has something to add
__device__ __forceinline__ void fx_wave_add(unsigned long long *acc, long long qb,
                                            long long qt) {
  for (int off = 32; off > 0; off >>= 1) {
    qb += __shfl_xor(qb, off, 64);
    qt += __shfl_xor(qt, off, 64);
  }
  if ((threadIdx.x & 63) == 0) {
    if (qb) atomicAdd(acc, (unsigned long long)qb);
    if (qt) atomicAdd(acc + 1, (unsigned long long)qt);
  }
}

// platen clamp + reaction sums (md_uniaxialcompression3_simple.cu:19-76,
// md_braziltest_simple2.cu:19-92).  MODE 1: load axis z, 2: load axis y.
// VT is the velocity/force vector type (FP32 v3 or FP64 d3); AT the
// accumulator (double, or fixed-point long long).  friction = false drops the
// Brazilian platens' sliding friction against the in-plane velocity, which
// quasi-static (FIRE) passes leave out like the contact law's
// velocity-dependent terms.
template <int MODE, class VT, class AT>
__device__ __forceinline__ void platen(const rfd_loading &L, P4 &R, VT &V, VT &F, AT &sb,
                                       AT &stp, bool friction = true) {
  using T = decltype(V.x);
  double ra, rb, h, zc;
  if (MODE == 1) {
    ra = R.x - L.cx;
    rb = R.y - L.cy;
    h = R.z;
    zc = R.z;
  } else {
    ra = R.x - L.cx;
    rb = R.z - L.cz;
    h = R.y;
    zc = R.z - L.cz;
  }
  if (!(ra * ra + rb * rb < L.RR && zc < L.zcut)) return;
  T &Fa = MODE == 1 ? F.z : F.y;
  T &Va = MODE == 1 ? V.z : V.y;
  double &Ra = MODE == 1 ? R.z : R.y;
  const T fr = (T)0.45f;
  if (h - L.top_r > -1e-9) {
    const T f = Fa;
    if (f > 0) {
      acc_add(stp, f);
      if (MODE == 2 && friction) {
        T vv = V.x * V.x + V.z * V.z;
        if ((double)vv > 1e-9) {
          vv = platen_rsqrt(vv);
          F.x = F.x - fr * f * V.x * vv;
          F.z = F.z - fr * f * V.z * vv;
        }
      }
    }
    if (h - L.top > -1e-9) {
      Ra = L.top;
      Va = (T)L.v;
      Fa = (T)0;
    }
  }
  if (h - L.bottom_r < 1e-9) {
    const T f = Fa;
    if (f < 0) {
      acc_add(sb, f);
      if (MODE == 2 && friction) {
        T vv = V.x * V.x + V.z * V.z;
        if ((double)vv > 1e-9) {
          vv = platen_rsqrt(vv);
          F.x = F.x + fr * f * V.x * vv;
          F.z = F.z + fr * f * V.z * vv;
        }
      }
    }
    if (h - L.bottom < 1e-9) {
      Ra = L.bottom;
      Va = (T)0;
      Fa = (T)0;
    }
  }
}

// platens, symplectic Euler with FP64 positions (md_calculateincrementsdem.cu:19-49),
// cut box (md_parallelepipedrestriction.cu:73-158) for one particle held in
// registers: F and M are its summed loads, iv its (1/m, 1/I)
template <int MODE>
__device__ __forceinline__ void advance_regs(const rfd_loading &L, P4 &R, v3 &V, v3 &W, v3 &F,
                                             const v3 &M, float2 iv, float dt, long long &qb,
                                             long long &qt) {
  if (MODE) platen<MODE>(L, R, V, F, qb, qt);
  V = V + F * iv.x;
  R.x += (double)V.x * dt;
  R.y += (double)V.y * dt;
  R.z += (double)V.z * dt;
  W = W + M * iv.y;
  if (L.cut) {
    const double dx = R.x - L.box_c[0], dy = R.y - L.box_c[1], dz = R.z - L.box_c[2];
    if (dx - L.box_h[0] > 0 || dx + L.box_h[0] < 0 || dy - L.box_h[1] > 0 ||
        dy + L.box_h[1] < 0 || dz - L.box_h[2] > 0 || dz + L.box_h[2] < 0) {
      R.x = L.hidden[0];
      R.y = L.hidden[1];
      R.z = L.hidden[2];
      V = kZero;
      F = kZero;
    }
  }
}

// The explicit step's particle update fused into the force pass: each tile
// advances its own particles once their loads are summed, writing positions,
// velocities and spins to the other buffers (other tiles still read the
// current ones), with the platen reactions summed in fixed point.
struct AdvDev {
  P4 *P;
  float4 *V, *W;
  const float2 *inv;
  const rfd_loading *ctl;    // per specimen
  unsigned long long *acc;   // per specimen: (bottom, top)
  float dt;
  int mode;  // loading mode: 0 none, 1 uniaxial (z), 2 Brazilian (y)
  // batched specimens (null for one): the specimen of particle 32 k, and the
  // specimens' first particles (S + 1); every specimen holds at least 128
  // particles, so a tile meets at most two
  const uint32_t *spec32, *spec_first;
  // ctl holds the loadings of specimens ctl_base, ctl_base + 1, ... (0: all)
  uint32_t ctl_base;
  // when set, the workgroup that finishes the pass last runs the loading
  // control of every specimen (the step's second launch otherwise): done
  // counts the finished workgroups and is zero between passes
  uint32_t *done;
  rfd_loading *ctl_w;   // the loadings it updates (ctl, writable)
  uint32_t nspec;       // specimens
  double *hist;         // history records, hist_cap per specimen
  uint32_t hist_cap;
  double dt64;
  int *curp;            // bond-state index to flip
};

// Force pass of the physical mode with every bond whose ends share a tile
// evaluated once: its lower-index half computes the law with the partner from
// LDS and writes both halves' loads; bonds leaving the tile are evaluated by
// each of their halves in the tile.  The two kinds run as separate loops over
// precomputed records, so no lane waits on a load another lane needs.
// Per-particle sums in row order, then the contact rows: the same bits as
// k_nat_forces.
// The FIRE measure fused into a quasi-static force pass: each tile's summed
// loads meet the platens at rest without friction (on copies of R and the
// FIRE velocity: the clamp of positions and velocities is applied by the
// update that follows, with the same decisions), and each wave writes its
// partial power, norms, maxima and platen sums (kFireVals values) to
// part[(tile * waves per tile + wave) * kFireVals].
struct MeasDev {
  const rfd_loading *ctl;
  const float4 *Vf;
  double *part;
  int mode;
};
constexpr int kMeasVals = 9;  // P, |F|^2, |V|^2, |M|^2, |W|^2, max|F|^2, max(|M|/r)^2, sb, st

// One tile of the pass below: `cur` is the index of the bond-state buffer
// holding the latest state.  Every thread of the workgroup calls it; threads
// return from it only after its last workgroup barrier.
template <int FP, int FT, bool ADV, bool MEAS = false, bool PIPE = false>
__device__ __forceinline__ void forces1_tile(const SysDev &s, const Mat &mt, const WorkDev &w,
                                             uint32_t tile, uint32_t cap, const AdvDev &a,
                                             int cur, const MeasDev *fm = nullptr) {
  static_assert(FP < FT && FP <= 128, "tile shape");
  static_assert(2 * FP <= FT, "one sum task per thread");
  // per-half loads: F (3 x cap floats) then M (3 x cap), cap >= 2 FP
  extern __shared__ float sDyn[];
  float *const sF0 = sDyn, *const sF1 = sDyn + cap, *const sF2 = sDyn + 2 * cap;
  float *const sM0 = sDyn + 3 * cap, *const sM1 = sDyn + 4 * cap, *const sM2 = sDyn + 5 * cap;
  __shared__ uint32_t sOff[FP + 1];
  __shared__ P4 sP[FP];
  __shared__ float4 sV[FP], sW[FP];
  // per-particle sums, kept in the slot memory once the slots are read
  float (*const sAcc)[3][FP] = reinterpret_cast<float (*)[3][FP]>(sDyn);
  BSt *const st_in = cur ? s.bst1 : s.bst0;
  BSt *const st_out = cur ? s.bst0 : s.bst1;
  const uint32_t tid = threadIdx.x;
  const uint32_t p0 = tile * FP;
  const uint32_t np = min((uint32_t)FP, s.N - p0);
  // owner side of a bond: state out, flag copy, break count
  auto commit = [&](uint32_t bnd, const BSt &old, const BondOut &o) {
    if (o.st.fl == kRemoved) {
      st_out[bnd].fl = kRemoved;
      if (!s.trial) st_in[bnd].fl = kRemoved;
    } else {
      st_out[bnd] = o.st;
    }
    if (!s.trial && o.st.fl != old.fl) {
      s.fl8[bnd] = (uint8_t)o.st.fl;
      if (s.nbreak && o.st.fl == kBroken) atomicAdd(s.nbreak, 1u);
    }
  };
  auto put = [&](uint32_t e, v3 F, v3 M) {
    sF0[e] = F.x;
    sF1[e] = F.y;
    sF2[e] = F.z;
    sM0[e] = M.x;
    sM1[e] = M.y;
    sM2[e] = M.z;
  };
  uint32_t h0;
  if constexpr (!PIPE) {
    if (tid <= np) sOff[tid] = s.hoff[p0 + tid];
    if (tid < np) {
      sP[tid] = s.P[p0 + tid];
      sV[tid] = s.V[p0 + tid];
      sW[tid] = s.W[p0 + tid];
    }
    __syncthreads();
    h0 = sOff[0];
    const uint32_t w0 = w.ofs[tile], w1 = w.ofs[tile + 1], wx = w0 + w.n_in[tile];
    // 1. bonds inside the tile, once each (the lower end owns them)
    for (uint32_t k = w0 + tid; k < wx; k += FT) {
      const WorkRec r = w.rec[k];
      const uint32_t e = r.a & 0x7FFu, lo = (r.a >> 11) & 0x7Fu, jl = (r.a >> 18) & 0x7Fu;
      const BSt old = st_in[r.b];
      BondOut o;
      if (old.fl == kRemoved) {
        o.Fi = o.Mbi = o.Mbj = o.Mci = o.Mcj = kZero;
      } else {
        EndData A, B;
        A.P = sP[lo];
        A.V = xyz(sV[lo]);
        A.W = xyz(sW[lo]);
        B.P = sP[jl];
        B.V = xyz(sV[jl]);
        B.W = xyz(sW[jl]);
        eval_bond(mt, A, B, s.brc[r.b], old, o);
        commit(r.b, old, o);
      }
      put(e, o.Fi, o.Mbi + o.Mci);
      put(r.c, -o.Fi, o.Mbj + o.Mcj);
    }
    // 2. halves whose partner is outside the tile, each evaluated by its end
    for (uint32_t k = wx + tid; k < w1; k += FT) {
      const WorkRec r = w.rec[k];
      const uint32_t e = r.a & 0x7FFu, lo = (r.a >> 11) & 0x7Fu, j = r.c;
      const bool own = p0 + lo < j;
      const BSt old = st_in[r.b];
      BondOut o;
      if (old.fl == kRemoved) {
        o.Fi = o.Mbi = o.Mbj = o.Mci = o.Mcj = kZero;
      } else {
        EndData me;
        me.P = sP[lo];
        me.V = xyz(sV[lo]);
        me.W = xyz(sW[lo]);
        const EndData other = load_end(s, j);
        eval_bond(mt, own ? me : other, own ? other : me, s.brc[r.b], old, o);
        if (own) commit(r.b, old, o);
      }
      if (own)
        put(e, o.Fi, o.Mbi + o.Mci);
      else
        put(e, -o.Fi, o.Mbj + o.Mcj);
    }
  } else {
    // Small systems, where a tile's chain of dependent loads bounds the pass
    // (one tile per CU): the records run as one list, bonds inside the tile
    // first, then the leaving halves; a thread takes records tid, tid + FT,
    // ...; its first two records and their bond data are loaded together with
    // the tile's particles, before the barrier that publishes those.  The
    // same arithmetic as the two loops above, on more registers.
    const uint32_t w0 = w.ofs[tile], w1 = w.ofs[tile + 1], wx = w0 + w.n_in[tile];
    if (tid <= np) sOff[tid] = s.hoff[p0 + tid];
    if (tid < np) {
      sP[tid] = s.P[p0 + tid];
      sV[tid] = s.V[p0 + tid];
      sW[tid] = s.W[p0 + tid];
    }
    struct Fetched {
      WorkRec r;
      BSt old;
      BR c;
      EndData x;  // the partner of a leaving half
    };
    auto fetch = [&](uint32_t k, Fetched &f) {
      f.r = w.rec[k];
      f.old = st_in[f.r.b];
      f.c = s.brc[f.r.b];
      if (k >= wx) f.x = load_end(s, f.r.c);
    };
    auto eval = [&](uint32_t k, const Fetched &f) {
      const WorkRec &r = f.r;
      const uint32_t e = r.a & 0x7FFu, lo = (r.a >> 11) & 0x7Fu;
      BondOut o;
      if (k < wx) {
        const uint32_t jl = (r.a >> 18) & 0x7Fu;
        if (f.old.fl == kRemoved) {
          o.Fi = o.Mbi = o.Mbj = o.Mci = o.Mcj = kZero;
        } else {
          EndData A, B;
          A.P = sP[lo];
          A.V = xyz(sV[lo]);
          A.W = xyz(sW[lo]);
          B.P = sP[jl];
          B.V = xyz(sV[jl]);
          B.W = xyz(sW[jl]);
          eval_bond(mt, A, B, f.c, f.old, o);
          commit(r.b, f.old, o);
        }
        put(e, o.Fi, o.Mbi + o.Mci);
        put(r.c, -o.Fi, o.Mbj + o.Mcj);
      } else {
        const bool own = p0 + lo < r.c;
        if (f.old.fl == kRemoved) {
          o.Fi = o.Mbi = o.Mbj = o.Mci = o.Mcj = kZero;
        } else {
          EndData me;
          me.P = sP[lo];
          me.V = xyz(sV[lo]);
          me.W = xyz(sW[lo]);
          eval_bond(mt, own ? me : f.x, own ? f.x : me, f.c, f.old, o);
          if (own) commit(r.b, f.old, o);
        }
        if (own)
          put(e, o.Fi, o.Mbi + o.Mci);
        else
          put(e, -o.Fi, o.Mbj + o.Mcj);
      }
    };
    const uint32_t k1 = w0 + tid, k2 = k1 + FT;
    Fetched f1, f2;
    if (k1 < w1) fetch(k1, f1);
    if (k2 < w1) fetch(k2, f2);
    __syncthreads();
    h0 = sOff[0];
    if (k1 < w1) eval(k1, f1);
    if (k2 < w1) eval(k2, f2);
    for (uint32_t k = k2 + FT; k < w1; k += FT) {
      Fetched f;
      fetch(k, f);
      eval(k, f);
    }
  }
  __syncthreads();
  // 3. per-particle sums in row order: task (group g, particle p), one per
  //    thread; the results go to the slot memory after every slot is read
  float x = 0.f, y = 0.f, z = 0.f;
  const uint32_t tp = tid % np, tg = tid / np;
  if (tid < 2 * np) {
    const uint32_t a0 = sOff[tp] - h0, a1 = sOff[tp + 1] - h0;
    const float *const src = sDyn + 3 * cap * tg;
    for (uint32_t e = a0; e < a1; ++e) {
      x = x + src[e];
      y = y + src[cap + e];
      z = z + src[2 * cap + e];
    }
  }
  __syncthreads();
  if (tid < 2 * np) {
    sAcc[tg][0][tp] = x;
    sAcc[tg][1][tp] = y;
    sAcc[tg][2][tp] = z;
  }
  __syncthreads();
  // the fused update and measure keep whole waves for their wave-wide sums
  if (tid >= ((ADV || MEAS) ? (FP + 63u) & ~63u : np)) return;
  const bool act = tid < np;
  const uint32_t i = p0 + tid, N = s.N;
  v3 F = kZero, M = kZero;
  EndData me;
  if (act) {
    // 4. never-bonded contact rows of this particle (per-end spring, in place)
    F = v3{sAcc[0][0][tid], sAcc[0][1][tid], sAcc[0][2][tid]};
    const v3 Mb = v3{sAcc[1][0][tid], sAcc[1][1][tid], sAcc[1][2][tid]};
    v3 Mc = kZero;
    me.P = sP[tid];
    me.V = xyz(sV[tid]);
    me.W = xyz(sW[tid]);
    const uint32_t nc = s.ccnt[i];
    for (uint32_t k = 0; k < nc; ++k) {
      const size_t sl = (size_t)k * N + i;
      const uint32_t j = s.cj[sl];
      const EndData other = load_end(s, j);
      const double rx = other.P.x - me.P.x, ry = other.P.y - me.P.y, rz = other.P.z - me.P.z;
      const double L64 = __builtin_sqrt(rx * rx + ry * ry + rz * rz);
      const float ri = (float)me.P.r, rj = (float)other.P.r;
      const double reach = me.P.r + other.P.r;
      if (L64 < reach && L64 > 1e-5) {
        const double inv = 1.0 / L64;
        const v3 n = {(float)(rx * inv), (float)(ry * inv), (float)(rz * inv)};
        v3 O = xyz(s.cO[sl]);
        v3 Fi, Mi, Mj;
        contact_law(mt, n, (float)(reach - L64), ri, rj, me.V, other.V, me.W, other.W, O, Fi,
                    Mi, Mj);
        F = F + Fi;
        Mc = Mc + Mi;
        if (!s.trial) s.cO[sl] = f4(O);
      } else if (!s.trial) {
        s.cO[sl] = make_float4(0.f, 0.f, 0.f, 0.f);
      }
    }
    M = Mb + Mc;
  }
  if (MEAS) {
    // k_fire_measure's arithmetic for this particle; the sums in wave order
    double v[kMeasVals] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    if (act) {
      rfd_loading L = *fm->ctl;
      L.v = 0.0;
      v3 Vf = xyz(fm->Vf[i]);
      P4 R = me.P;
      if (fm->mode == 1)
        platen<1>(L, R, Vf, F, v[7], v[8], false);
      else if (fm->mode == 2)
        platen<2>(L, R, Vf, F, v[7], v[8], false);
      s.F[i] = f4(F);
      s.M[i] = f4(M);
      const v3 W = me.W;
      v[0] = (double)dot(F, Vf) + (double)dot(M, W);
      const double f2 = dot(F, F), m2 = dot(M, M);
      v[1] = f2;
      v[2] = (double)dot(Vf, Vf);
      v[3] = m2;
      v[4] = (double)dot(W, W);
      v[5] = f2;
      v[6] = m2 / (R.r * R.r);
    }
#pragma unroll
    for (int q = 0; q < kMeasVals; ++q)
      v[q] = (q == 5 || q == 6) ? wave_max(v[q]) : wave_sum(v[q]);
    if ((tid & 63) == 0) {
      constexpr uint32_t wpt = (FP + 63) / 64;
      double *dst = fm->part + ((size_t)tile * wpt + (tid >> 6)) * kMeasVals;
#pragma unroll
      for (int q = 0; q < kMeasVals; ++q) dst[q] = v[q];
    }
    return;
  }
  if (!ADV) {
    s.F[i] = f4(F);
    s.M[i] = f4(M);
    return;
  }
  // 5. platens, integration and cut box of this particle (the separate
  //    particle pass's arithmetic, on the sums in registers)
  long long qb = 0, qt = 0;
  uint32_t sp0 = 0, sp = 0;
  if (a.spec32) {
    sp0 = a.spec32[p0 >> 5];
    sp = sp0 + (i >= a.spec_first[sp0 + 1] ? 1u : 0u);
  }
  if (act) {
    const rfd_loading L = a.ctl[sp - a.ctl_base];
    const float2 iv = a.inv[i];
    if (a.mode == 1)
      advance_regs<1>(L, me.P, me.V, me.W, F, M, iv, a.dt, qb, qt);
    else if (a.mode == 2)
      advance_regs<2>(L, me.P, me.V, me.W, F, M, iv, a.dt, qb, qt);
    else
      advance_regs<0>(L, me.P, me.V, me.W, F, M, iv, a.dt, qb, qt);
    a.P[i] = me.P;
    a.V[i] = f4(me.V);
    a.W[i] = f4(me.W);
    s.F[i] = f4(F);
    s.M[i] = f4(M);
  }
  if (a.mode) {
    if (!a.spec32) {
      fx_wave_add(a.acc, qb, qt);
    } else {
      // the tile's first specimen and the next one: the lanes' reactions
      // split between the two (the second adds nothing when no lane has it)
      const bool hi = sp != sp0;
      fx_wave_add(a.acc + 2 * sp0, hi ? 0 : qb, hi ? 0 : qt);
      fx_wave_add(a.acc + 2 * (sp0 + 1), hi ? qb : 0, hi ? qt : 0);
    }
  }
}

// The loading control in the force pass (explicit steps of small systems,
// where the control's own launch is a large part of a step): every workgroup,
// once its platen reactions have been added to the fixed-point sums, counts
// itself done, and the last one reads and clears the sums and runs the
// control of every specimen, as k_nat_control_fx does.  The sums and the
// counter are only ever read, added to or cleared with device-scope atomics,
// which the GPU performs at one point per address for all of its XCDs, and
// each workgroup's additions are complete (s_waitcnt vmcnt(0)) before it
// counts itself; so the last workgroup sees every addition without a fence
// that writes back an L2.  Nothing else the control reads is written during
// the pass (each specimen's loading is read by the tiles and written only
// here, after every tile is done), and it leaves the counter at zero.
__device__ __forceinline__ void control_update(rfd_loading &L, double s0, double s1,
                                               double *hist, uint32_t hist_cap, double dt);
__device__ __forceinline__ void last_tile_control(const AdvDev &a, uint32_t ntiles) {
  __shared__ uint32_t sLast;
  asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
  __syncthreads();
  if (threadIdx.x == 0) sLast = atomicAdd(a.done, 1u) == ntiles - 1 ? 1u : 0u;
  __syncthreads();
  if (!sLast) return;
  for (uint32_t sp = threadIdx.x; sp < a.nspec; sp += blockDim.x) {
    const double s0 = (double)(long long)atomicExch(a.acc + 2 * sp, 0ull) * kFxInv;
    const double s1 = (double)(long long)atomicExch(a.acc + 2 * sp + 1, 0ull) * kFxInv;
    rfd_loading L = a.ctl_w[sp];
    control_update(L, s0, s1, a.hist + 3 * (size_t)a.hist_cap * sp, a.hist_cap, a.dt64);
    a.ctl_w[sp] = L;
  }
  if (threadIdx.x == 0) {
    a.curp[0] = 1 - a.curp[0];
    atomicExch(a.done, 0u);
  }
}

template <int FP, int FT, bool ADV>
__global__ void __launch_bounds__(FT) __attribute__((amdgpu_waves_per_eu(5, 8)))
k_nat_forces1(SysDev s, Mat mt, WorkDev w, uint32_t ntiles, uint32_t per, uint32_t cap,
              AdvDev a) {
  const uint32_t tile = xcd_tile(per);
  if (tile >= ntiles) return;
  if (s.done && *s.done) return;
  if (s.dtp) mt.dt = *s.dtp;
  forces1_tile<FP, FT, ADV>(s, mt, w, tile, cap, a, *s.curp);
}

// The fused explicit step of a small system (at most kCtlInPassTiles tiles,
// one per CU): the pipelined tile, which needs more registers than one CU
// can give five workgroups, and the loading control by the last workgroup
template <int FP, int FT>
__global__ void __launch_bounds__(FT) __attribute__((amdgpu_waves_per_eu(2, 8)))
k_nat_forces1_small(SysDev s, Mat mt, WorkDev w, uint32_t ntiles, uint32_t per,
                    uint32_t cap, AdvDev a) {
  const uint32_t tile = xcd_tile(per);
  if (tile >= ntiles) return;
  forces1_tile<FP, FT, true, false, true>(s, mt, w, tile, cap, a, *s.curp);
  last_tile_control(a, ntiles);
}

// a quasi-static force pass with the FIRE measure in its epilogue
template <int FP, int FT>
__global__ void __launch_bounds__(FT) __attribute__((amdgpu_waves_per_eu(5, 8)))
k_nat_forces1_fire(SysDev s, Mat mt, WorkDev w, uint32_t ntiles, uint32_t per, uint32_t cap,
                   MeasDev fm) {
  const uint32_t tile = xcd_tile(per);
  if (tile >= ntiles) return;
  if (s.done && *s.done) return;
  if (s.dtp) mt.dt = *s.dtp;
  forces1_tile<FP, FT, false, true>(s, mt, w, tile, cap, AdvDev{}, *s.curp, &fm);
}

// work records of the single-evaluation pass: per tile, count the in-tile
// bonds (once, at their lower end) and the halves leaving the tile
template <int FP>
__global__ void __launch_bounds__(256) k_work_count(const uint32_t *__restrict__ hoff,
                                                    const HR *__restrict__ hr, uint32_t N,
                                                    uint32_t ntiles, uint32_t *cnt,
                                                    uint32_t *n_in) {
  const uint32_t t = blockIdx.x;
  if (t >= ntiles) return;
  __shared__ uint32_t s_in, s_x;
  if (threadIdx.x == 0) s_in = s_x = 0;
  __syncthreads();
  const uint32_t p0 = t * FP, np = min((uint32_t)FP, N - p0);
  const uint32_t h0 = hoff[p0], h1 = hoff[p0 + np];
  uint32_t a = 0, x = 0;
  for (uint32_t h = h0 + threadIdx.x; h < h1; h += blockDim.x) {
    // row particle of h
    uint32_t lo = 0, hi = np;
    while (hi - lo > 1) {
      const uint32_t mid = (lo + hi) >> 1;
      if (hoff[p0 + mid] <= h) lo = mid; else hi = mid;
    }
    const uint32_t jl = hr[h].j - p0;
    if (jl >= np) ++x;
    else if (lo < jl) ++a;
  }
  atomicAdd(&s_in, a);
  atomicAdd(&s_x, x);
  __syncthreads();
  if (threadIdx.x == 0) {
    cnt[t] = s_in + s_x;
    n_in[t] = s_in;
  }
}

// fills the records in row order (ordered wave-ballot compaction per class)
template <int FP>
__global__ void __launch_bounds__(256) k_work_fill(const uint32_t *__restrict__ hoff,
                                                   const HR *__restrict__ hr,
                                                   const uint32_t *__restrict__ hm, uint32_t N,
                                                   uint32_t ntiles, const uint32_t *ofs,
                                                   const uint32_t *n_in, WorkRec *rec) {
  const uint32_t t = blockIdx.x;
  if (t >= ntiles) return;
  __shared__ uint32_t s_cnt[2][4];
  const uint32_t tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  const uint32_t p0 = t * FP, np = min((uint32_t)FP, N - p0);
  const uint32_t h0 = hoff[p0], h1 = hoff[p0 + np];
  uint32_t base_in = ofs[t], base_x = ofs[t] + n_in[t];
  for (uint32_t b0 = h0; b0 < h1; b0 += 256) {
    const uint32_t h = b0 + tid;
    bool isin = false, isx = false;
    uint32_t lo = 0, jl = 0;
    HR r{0, 0};
    if (h < h1) {
      uint32_t hi = np;
      while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (hoff[p0 + mid] <= h) lo = mid; else hi = mid;
      }
      r = hr[h];
      jl = r.j - p0;
      isx = jl >= np;
      isin = !isx && lo < jl;
    }
    const uint64_t bi = __ballot(isin), bx = __ballot(isx);
    const uint64_t below = (1ull << lane) - 1ull;
    if (lane == 0) {
      s_cnt[0][wid] = (uint32_t)__popcll(bi);
      s_cnt[1][wid] = (uint32_t)__popcll(bx);
    }
    __syncthreads();
    uint32_t oi = 0, ox = 0, ti = 0, tx = 0;
    for (uint32_t q = 0; q < 4; ++q) {
      oi += q < wid ? s_cnt[0][q] : 0u;
      ox += q < wid ? s_cnt[1][q] : 0u;
      ti += s_cnt[0][q];
      tx += s_cnt[1][q];
    }
    const uint32_t e = h - h0;
    if (isin) {
      WorkRec w;
      w.a = e | (lo << 11) | (jl << 18);
      w.b = r.b;
      w.c = hm[h] - h0;
      rec[base_in + oi + (uint32_t)__popcll(bi & below)] = w;
    } else if (isx) {
      WorkRec w;
      w.a = e | (lo << 11);
      w.b = r.b;
      w.c = r.j;
      rec[base_x + ox + (uint32_t)__popcll(bx & below)] = w;
    }
    base_in += ti;
    base_x += tx;
    __syncthreads();
  }
}

// flips the state-buffer index after a force pass that ran
__global__ void k_flip_state(int *curp, const int *done) {
  if (threadIdx.x == 0 && blockIdx.x == 0 && !(done && *done)) curp[0] = 1 - curp[0];
}

// ---------------------------------------------------------- particle pass
struct PassDev {
  uint32_t N;
  P4 *P;
  float4 *V, *W, *F;
  const float4 *M;
  const float2 *inv;
  float dt;
  rfd_loading *ctl;
  unsigned long long *acc;  // fixed-point platen sums (bottom, top), per specimen
};

template <int MODE>
__device__ __forceinline__ void advance_one(const PassDev &p, const rfd_loading &L, uint32_t i,
                                            long long &qb, long long &qt) {
  P4 R = p.P[i];
  v3 V = xyz(p.V[i]), F = xyz(p.F[i]), W = xyz(p.W[i]);
  advance_regs<MODE>(L, R, V, W, F, xyz(p.M[i]), p.inv[i], p.dt, qb, qt);
  p.P[i] = R;
  p.V[i] = f4(V);
  p.W[i] = f4(W);
  p.F[i] = f4(F);
}

// workgroup sum of the two platen reactions, written as this block's partial
// pair (FP64 path)
__device__ __forceinline__ void write_partials(double *part, double sb, double stp) {
  __shared__ double s_part[8];
  sb = wave_sum(sb);
  stp = wave_sum(stp);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) {
    s_part[2 * wid] = sb;
    s_part[2 * wid + 1] = stp;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double a = 0, c = 0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) {
      a += s_part[2 * w];
      c += s_part[2 * w + 1];
    }
    part[2 * blockIdx.x] = a;
    part[2 * blockIdx.x + 1] = c;
  }
}

template <int MODE>
__global__ void __launch_bounds__(256) k_nat_pass(PassDev p) {
  const rfd_loading L = *p.ctl;
  const uint32_t N = p.N;
  long long qb = 0, qt = 0;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += grid_stride())
    advance_one<MODE>(p, L, i, qb, qt);
  if (MODE) fx_wave_add(p.acc, qb, qt);
}

// one step of the loading control from the reduced platen sums
// (SumForcesULoading md_results.cu:34-74 and CompressU md_data_types.h:223):
// history record, stress, loading/fracture detection, platen move
__device__ __forceinline__ void control_update(rfd_loading &L, double s0, double s1,
                                               double *hist, uint32_t hist_cap, double dt) {
  L.steps += 1;
  if (!L.mode) return;
  L.sum_bottom = s0;
  L.sum_top = s1;
  if (L.hist_len < hist_cap) {
    if (hist) {  // null: the record is counted but written by another workgroup
      hist[3 * (size_t)L.hist_len] = (L.size - L.size0) / L.size0;
      hist[3 * (size_t)L.hist_len + 1] = s0 * L.inv_area;
      hist[3 * (size_t)L.hist_len + 2] = s1 * L.inv_area;
    }
    L.hist_len += 1;
  }
  L.stress = 0.5 * (fabs(s0) + fabs(s1)) * L.inv_area;
  if (!L.loading && L.stress > L.min_stress) L.loading = 1;
  if (L.fracture_stress < L.stress) {
    L.fracture_stress = L.stress;
    L.fracture_strain = (L.size0 - L.size) / L.size0;
  }
  if (L.loading && L.stress < L.zero_stress) L.fracture = 1;
  L.v = L.stress < 1e-9 ? 10.0 * L.v_load : L.v_load;
  L.size += L.v * dt;
  L.top += L.v * dt;
  L.top_r = L.top - L.band;
  L.strain = (L.size - L.size0) / L.size0;
}

// reduce nb platen partial pairs in a fixed order within one workgroup
__device__ __forceinline__ void reduce_pairs(const double *part, int b0, int b1, double &s0,
                                             double &s1) {
  __shared__ double s_part[8];
  double a = 0, c = 0;
  for (int b = b0 + threadIdx.x; b < b1; b += blockDim.x) {
    a += part[2 * b];
    c += part[2 * b + 1];
  }
  a = wave_sum(a);
  c = wave_sum(c);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) {
    s_part[2 * wid] = a;
    s_part[2 * wid + 1] = c;
  }
  __syncthreads();
  s0 = s1 = 0;
  for (int w = 0; w < (int)(blockDim.x >> 6); ++w) {
    s0 += s_part[2 * w];
    s1 += s_part[2 * w + 1];
  }
}

// loading control of the FP64 path: the platen partial sums are reduced by
// one workgroup in a fixed order, then one thread applies the control and
// flips the bond-state index that the step's force pass left
__global__ void __launch_bounds__(256) k_nat_control(rfd_loading *ctl, const double *part,
                                                     int nblocks, double *hist, double dt,
                                                     int *curp) {
  double s0, s1;
  reduce_pairs(part, 0, nblocks, s0, s1);
  if (threadIdx.x != 0) return;
  curp[0] = 1 - curp[0];
  rfd_loading L = *ctl;
  control_update(L, s0, s1, hist, (uint32_t)kHistCap, dt);
  *ctl = L;
}

// loading control of the FP32 steps, thread k for specimen k: the fixed-point
// sums of the pass are read and cleared for the next step.  curp, when set,
// is the bond-state index that the step's force pass leaves to be flipped.
__global__ void __launch_bounds__(64) k_nat_control_fx(rfd_loading *ctls,
                                                       unsigned long long *acc, uint32_t S,
                                                       double *hist, uint32_t hist_cap,
                                                       double dt, int *curp) {
  const uint32_t sp = blockIdx.x * blockDim.x + threadIdx.x;
  if (sp == 0 && curp) curp[0] = 1 - curp[0];
  if (sp >= S) return;
  const double s0 = (double)(long long)acc[2 * sp] * kFxInv;
  const double s1 = (double)(long long)acc[2 * sp + 1] * kFxInv;
  acc[2 * sp] = 0ull;
  acc[2 * sp + 1] = 0ull;
  rfd_loading L = ctls[sp];
  control_update(L, s0, s1, hist + 3 * (size_t)hist_cap * sp, hist_cap, dt);
  ctls[sp] = L;
}

// ------------------------------------------------- multi-step launch
// Explicit FP32 steps of the fused pass in one cooperative launch, for
// specimens and batches small enough that every force tile has a resident
// workgroup: step j reads particle set j & 1 and writes the other, a grid
// barrier ends each step, and each workgroup then runs the loading control of
// the step for the (at most two) specimens its tile meets.  Those inputs are
// the same for every workgroup, so every copy of a specimen's loading holds
// the same bits; the workgroup whose tile holds the specimen's first particle
// writes its history record and, at the end, its loading.  The fixed-point
// platen sums rotate through three buffers: step j adds into buffer j % 3,
// which every workgroup reads after the barrier; the owner clears buffer
// (j + 2) % 3 then (read after the previous barrier, added into two steps
// on).  The buffers are zero on entry and on exit.  The arithmetic of a step
// is that of k_nat_forces1<.., true> followed by k_nat_control_fx.
struct StepsDev {
  P4 *P[2];
  float4 *V[2], *W[2];
  rfd_loading *ctl;             // per specimen
  unsigned long long *acc3;     // 3 x 2 x S
  double *hist;                 // per specimen, hist_cap records of 3
  uint32_t hist_cap, S, nsteps;
  int *curp;
  double dt;
};

template <int FP, int FT>
__global__ void __launch_bounds__(FT) __attribute__((amdgpu_waves_per_eu(4, 8)))
k_nat_steps(SysDev s, Mat mt, WorkDev w, uint32_t ntiles, uint32_t per, uint32_t cap,
            AdvDev a, StepsDev q) {
  __shared__ rfd_loading sL[2];
  __shared__ uint32_t sSp[2];
  __shared__ int sOwn[2];
  cooperative_groups::grid_group grid = cooperative_groups::this_grid();
  const uint32_t tile = xcd_tile(per);
  const bool live = tile < ntiles;
  const int cur0 = *q.curp;
  const uint32_t S = q.S;
  if (threadIdx.x == 0) {
    uint32_t sp0 = 0, sp1 = 0;
    int own0 = 0, own1 = 0;
    if (live) {
      const uint32_t p0 = tile * FP, pl = min(p0 + FP, s.N) - 1;
      if (a.spec32) {
        sp0 = a.spec32[p0 >> 5];
        sp1 = sp0 + 1 < S && pl >= a.spec_first[sp0 + 1] ? sp0 + 1 : sp0;
        own0 = a.spec_first[sp0] >= p0;
        own1 = sp1 != sp0;  // the second specimen starts inside this tile
      } else {
        own0 = tile == 0;
      }
    }
    sSp[0] = sp0;
    sSp[1] = sp1;
    sOwn[0] = own0;
    sOwn[1] = own1;
    sL[0] = q.ctl[sp0];
    sL[1] = q.ctl[sp1];
  }
  __syncthreads();
  AdvDev b = a;
  b.ctl = sL;
  b.ctl_base = sSp[0];
  for (uint32_t j = 0; j < q.nsteps; ++j) {
    const int in = j & 1;  // (selects, not indexing: no copy of q to scratch)
    s.P = in ? q.P[1] : q.P[0];
    s.V = in ? q.V[1] : q.V[0];
    s.W = in ? q.W[1] : q.W[0];
    b.P = in ? q.P[0] : q.P[1];
    b.V = in ? q.V[0] : q.V[1];
    b.W = in ? q.W[0] : q.W[1];
    b.acc = q.acc3 + 2 * (size_t)S * (j % 3);
    if (live) forces1_tile<FP, FT, true>(s, mt, w, tile, cap, b, cur0 ^ (int)in);
    grid.sync();
    if (live && threadIdx.x < 2 && (threadIdx.x == 0 || sSp[1] != sSp[0])) {
      const uint32_t k = threadIdx.x, sp = sSp[k];
      const bool own = sOwn[k] != 0;
      unsigned long long *acc = q.acc3 + 2 * (size_t)S * (j % 3) + 2 * sp;
      const double s0 = (double)(long long)acc[0] * kFxInv;
      const double s1 = (double)(long long)acc[1] * kFxInv;
      control_update(sL[k], s0, s1, own ? q.hist + 3 * (size_t)q.hist_cap * sp : nullptr,
                     q.hist_cap, q.dt);
      if (own) {
        unsigned long long *old = q.acc3 + 2 * (size_t)S * ((j + 2) % 3) + 2 * sp;
        old[0] = 0ull;
        old[1] = 0ull;
      }
    }
    __syncthreads();
  }
  grid.sync();
  // every workgroup has read the last step's sums: clear them, hand out the
  // loadings and the bond-state index
  if (live && threadIdx.x < 2 && (threadIdx.x == 0 || sSp[1] != sSp[0]) && sOwn[threadIdx.x]) {
    const uint32_t sp = sSp[threadIdx.x];
    unsigned long long *last = q.acc3 + 2 * (size_t)S * ((q.nsteps + 2) % 3) + 2 * sp;
    last[0] = 0ull;
    last[1] = 0ull;
    q.ctl[sp] = sL[threadIdx.x];
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) *q.curp = cur0 ^ (int)(q.nsteps & 1u);
}

// batched pass: workgroup b advances particles [blk_first[b], blk_end[b]) of
// specimen blk_spec[b] against that specimen's platens
template <int MODE>
__global__ void __launch_bounds__(256) k_nat_pass_batch(PassDev p,
                                                        const uint32_t *__restrict__ blk_spec,
                                                        const uint32_t *__restrict__ blk_first,
                                                        const uint32_t *__restrict__ blk_end,
                                                        const rfd_loading *__restrict__ ctls) {
  const uint32_t sp = blk_spec[blockIdx.x];
  const rfd_loading L = ctls[sp];
  const uint32_t e = blk_end[blockIdx.x];
  long long qb = 0, qt = 0;
  for (uint32_t i = blk_first[blockIdx.x] + threadIdx.x; i < e; i += blockDim.x)
    advance_one<MODE>(p, L, i, qb, qt);
  fx_wave_add(p.acc + 2 * sp, qb, qt);
}

// ----------------------------------------------------- neighbour search
struct GridDev {
  double ox, oy, oz, inv_cell;
  uint32_t nx, ny, nz;
};

__device__ __forceinline__ uint32_t cell_clamped(const GridDev &g, const P4 &R, int c[3]) {
  const double f[3] = {(R.x - g.ox) * g.inv_cell, (R.y - g.oy) * g.inv_cell,
                       (R.z - g.oz) * g.inv_cell};
  const uint32_t n[3] = {g.nx, g.ny, g.nz};
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    double q = floor(f[d]);
    if (!(q >= 0)) q = 0;
    if (q > (double)(n[d] - 1)) q = (double)(n[d] - 1);
    c[d] = (int)q;
  }
  return (uint32_t)c[0] + (uint32_t)c[1] * g.nx + (uint32_t)c[2] * g.nx * g.ny;
}

__global__ void k_nat_cell_index(const P4 *__restrict__ P, uint32_t N, GridDev g,
                                 uint32_t *__restrict__ CI, uint32_t *__restrict__ cnt) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += grid_stride()) {
    int c[3];
    const uint32_t ci = cell_clamped(g, P[i], c);
    CI[i] = ci;
    atomicAdd(&cnt[ci], 1u);
  }
}

// exclusive scan of cnt[0..C) by one workgroup of 1024 threads; start[C] = total
__global__ void __launch_bounds__(1024)
k_nat_scan(const uint32_t *__restrict__ cnt, uint32_t C, uint32_t *__restrict__ start,
           uint32_t *__restrict__ cursor) {
  __shared__ uint32_t s_sum[16];
  __shared__ uint32_t s_carry;
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  if (tid == 0) s_carry = 0;
  __syncthreads();
  constexpr int IT = 8;
  for (uint32_t base = 0; base < C; base += 1024 * IT) {
    uint32_t v[IT], local = 0;
    const uint32_t b0 = base + tid * IT;
#pragma unroll
    for (int k = 0; k < IT; ++k) {
      v[k] = (b0 + k < C) ? cnt[b0 + k] : 0u;
      local += v[k];
    }
    uint32_t inc = local;
    for (int off = 1; off < 64; off <<= 1) {
      const uint32_t t = __shfl_up(inc, off, 64);
      if (lane >= off) inc += t;
    }
    if (lane == 63) s_sum[wid] = inc;
    __syncthreads();
    uint32_t wo = 0;
    for (int w = 0; w < wid; ++w) wo += s_sum[w];
    uint32_t run = s_carry + wo + inc - local;
#pragma unroll
    for (int k = 0; k < IT; ++k) {
      if (b0 + k < C) {
        start[b0 + k] = run;
        if (cursor) cursor[b0 + k] = run;
      }
      run += v[k];
    }
    __syncthreads();
    if (tid == 1023) s_carry = run;
    __syncthreads();
  }
  if (tid == 0) start[C] = s_carry;
}

// three-phase exclusive scan for long arrays: per-chunk totals, a scan of the
// totals (k_nat_scan, one workgroup), then per-chunk scans with their offsets
constexpr uint32_t kScanChunk = 1024 * 8;

__global__ void __launch_bounds__(1024)
k_scan_totals(const uint32_t *__restrict__ in, uint32_t n, uint32_t *__restrict__ tot) {
  __shared__ uint32_t s_w[16];
  const uint32_t base = blockIdx.x * kScanChunk, tid = threadIdx.x;
  uint32_t v = 0;
  for (uint32_t k = 0; k < 8; ++k) {
    const uint32_t i = base + tid * 8 + k;
    v += i < n ? in[i] : 0u;
  }
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  if ((tid & 63) == 0) s_w[tid >> 6] = v;
  __syncthreads();
  if (tid == 0) {
    uint32_t a = 0;
    for (int w = 0; w < 16; ++w) a += s_w[w];
    tot[blockIdx.x] = a;
  }
}

__global__ void __launch_bounds__(1024)
k_scan_apply(const uint32_t *__restrict__ in, uint32_t n, const uint32_t *__restrict__ off,
             uint32_t nchunks, uint32_t *__restrict__ start, uint32_t *__restrict__ cursor) {
  __shared__ uint32_t s_sum[16];
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  const uint32_t b0 = blockIdx.x * kScanChunk + tid * 8;
  uint32_t v[8], local = 0;
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    v[k] = (b0 + k < n) ? in[b0 + k] : 0u;
    local += v[k];
  }
  uint32_t inc = local;
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t t = __shfl_up(inc, o, 64);
    if (lane >= o) inc += t;
  }
  if (lane == 63) s_sum[wid] = inc;
  __syncthreads();
  uint32_t wo = 0;
  for (int w = 0; w < wid; ++w) wo += s_sum[w];
  uint32_t run = off[blockIdx.x] + wo + inc - local;
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    if (b0 + k < n) {
      start[b0 + k] = run;
      if (cursor) cursor[b0 + k] = run;
    }
    run += v[k];
  }
  if (blockIdx.x == nchunks - 1 && tid == 0) start[n] = off[nchunks];
}

__global__ void k_nat_scatter(const uint32_t *__restrict__ CI, uint32_t N,
                              uint32_t *__restrict__ cursor, uint32_t *__restrict__ tmp) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += grid_stride())
    tmp[atomicAdd(&cursor[CI[i]], 1u)] = i;
}

// stable order inside each cell: each particle counts the members of its cell
// with a smaller id (one thread per particle) and writes itself at that rank,
// with its position relative to the grid origin (FP32) and radius in cell
// order, for the contact-row search
__global__ void k_nat_rank(const uint32_t *__restrict__ tmp, const uint32_t *__restrict__ start,
                           const uint32_t *__restrict__ CI, uint32_t N,
                           uint32_t *__restrict__ sorted, const P4 *__restrict__ P, GridDev g,
                           float4 *__restrict__ cpos) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += grid_stride()) {
    const uint32_t c = CI[i], s0 = start[c], n = start[c + 1] - s0;
    uint32_t rank = 0;
    for (uint32_t t = 0; t < n; ++t) rank += (tmp[s0 + t] < i) ? 1u : 0u;
    sorted[s0 + rank] = i;
    const P4 R = P[i];
    cpos[s0 + rank] = make_float4((float)(R.x - g.ox), (float)(R.y - g.oy), (float)(R.z - g.oz),
                                  (float)R.r);
  }
}

// is there a bond record (not removed) between a and b ?  (row a holds b)
__device__ __forceinline__ bool bonded(const uint32_t *__restrict__ hoff,
                                       const HR *__restrict__ hr, const uint8_t *__restrict__ fl8,
                                       uint32_t a, uint32_t b) {
  for (uint32_t q = hoff[a]; q < hoff[a + 1]; ++q) {
    const HR h = hr[q];
    if (h.j == b) return fl8[h.b] != kRemoved;
  }
  return false;
}

struct NbrDev {
  uint32_t N;
  const P4 *P;
  GridDev g;
  const uint32_t *start, *sorted;
  const float4 *cpos;  // cell-ordered positions relative to the grid origin, radius
  double extra;
  uint32_t *count;
  const uint32_t *pair_off;
  uint32_t *pairs;
  uint32_t Kc;
  const uint32_t *hoff;
  const HR *hr;
  const uint8_t *fl8;
  const uint32_t *cj_old, *ccnt_old;
  const float4 *cO_old;
  const double4 *cO64_old;  // FP64 path springs (carried over beside cO when set)
  double4 *cO64_new;
  uint32_t *cj_new, *ccnt_new;
  float4 *cO_new;
  uint32_t *flags;
};

// Per particle, the 27 surrounding cells in fixed order, members by id.
// MODE 0: count pairs j>i with |rij| < ri + rj + extra; MODE 1: write them;
// MODE 2: rebuild the contact row (never-bonded j within ri + rj + extra),
// carrying each spring O over from the previous row.
template <int MODE, int NB = 128>
__global__ void __launch_bounds__(NB) k_nat_neighbours(NbrDev s, uint32_t per) {
  // XCD-aware: each XCD scans one contiguous run of Morton-ordered particles,
  // so the cells around a run are fetched into one L2 rather than into all
  const uint32_t i = xcd_tile(per) * NB + threadIdx.x;
  if (i >= s.N) return;
  const uint32_t N = s.N;
  const P4 Ri = s.P[i];
  int c[3];
  cell_clamped(s.g, Ri, c);
  // MODE 2 tests candidates in FP32 relative coordinates from the cell-ordered
  // copy (contiguous per cell); its skin absorbs the rounding
  const float pix = (float)(Ri.x - s.g.ox), piy = (float)(Ri.y - s.g.oy),
              piz = (float)(Ri.z - s.g.oz), rif = (float)Ri.r, exf = (float)s.extra;
  uint32_t nfound = 0;
  uint32_t out = MODE == 1 ? s.pair_off[i] : 0;
  for (int dz = -1; dz <= 1; ++dz) {
    const int z = c[2] + dz;
    if (z < 0 || z >= (int)s.g.nz) continue;
    for (int dy = -1; dy <= 1; ++dy) {
      const int y = c[1] + dy;
      if (y < 0 || y >= (int)s.g.ny) continue;
      for (int dx = -1; dx <= 1; ++dx) {
        const int x = c[0] + dx;
        if (x < 0 || x >= (int)s.g.nx) continue;
        const uint32_t cc = (uint32_t)x + (uint32_t)y * s.g.nx + (uint32_t)z * s.g.nx * s.g.ny;
        const uint32_t e1 = s.start[cc + 1];
        for (uint32_t e = s.start[cc]; e < e1; ++e) {
          uint32_t j;
          if (MODE == 2) {
            const float4 q = s.cpos[e];
            const float dx = q.x - pix, dy2 = q.y - piy, dz2 = q.z - piz;
            const float d2 = dx * dx + dy2 * dy2 + dz2 * dz2;
            const float lim = rif + q.w + exf;
            if (!(d2 >= 1e-10f) || d2 >= lim * lim) continue;
            j = s.sorted[e];
            if (j == i) continue;
          } else {
            j = s.sorted[e];
            if (j <= i) continue;
            const P4 Rj = s.P[j];
            const double rx = Rj.x - Ri.x, ry = Rj.y - Ri.y, rz = Rj.z - Ri.z;
            const double d2 = rx * rx + ry * ry + rz * rz;
            if (d2 < 1e-10) continue;
            const double lim = Ri.r + Rj.r + s.extra;
            if (d2 >= lim * lim) continue;
          }
          if (MODE == 0) {
            ++nfound;
          } else if (MODE == 1) {
            s.pairs[2 * (size_t)out] = i;
            s.pairs[2 * (size_t)out + 1] = j;
            ++out;
          } else {
            if (bonded(s.hoff, s.hr, s.fl8, i, j)) continue;
            if (nfound < s.Kc) {
              const size_t sl = (size_t)nfound * N + i;
              s.cj_new[sl] = j;
              float4 O = make_float4(0.f, 0.f, 0.f, 0.f);
              double4 O64 = make_double4(0.0, 0.0, 0.0, 0.0);
              const uint32_t no = s.ccnt_old ? s.ccnt_old[i] : 0;
              for (uint32_t k = 0; k < no; ++k) {
                const size_t so = (size_t)k * N + i;
                if (s.cj_old[so] == j) {
                  O = s.cO_old[so];
                  if (s.cO64_new) O64 = s.cO64_old[so];
                  break;
                }
              }
              s.cO_new[sl] = O;
              if (s.cO64_new) s.cO64_new[sl] = O64;
            } else {
              atomicOr(s.flags, RFD_FLAG_ROW_FULL);
            }
            ++nfound;
          }
        }
      }
    }
  }
  if (MODE == 0) s.count[i] = nfound;
  if (MODE == 2) s.ccnt_new[i] = nfound < s.Kc ? nfound : s.Kc;
}

// contact-row entry for candidate j of particle i that passed the distance
// and bond tests: spring carried over from the previous row by partner id
__device__ __forceinline__ void row_entry(const NbrDev &s, uint32_t i, uint32_t j,
                                          uint32_t &nfound) {
  const uint32_t N = s.N;
  if (nfound < s.Kc) {
    const size_t sl = (size_t)nfound * N + i;
    s.cj_new[sl] = j;
    float4 O = make_float4(0.f, 0.f, 0.f, 0.f);
    double4 O64 = make_double4(0.0, 0.0, 0.0, 0.0);
    const uint32_t no = s.ccnt_old ? s.ccnt_old[i] : 0;
    for (uint32_t k = 0; k < no; ++k) {
      const size_t so = (size_t)k * N + i;
      if (s.cj_old[so] == j) {
        O = s.cO_old[so];
        if (s.cO64_new) O64 = s.cO64_old[so];
        break;
      }
    }
    s.cO_new[sl] = O;
    if (s.cO64_new) s.cO64_new[sl] = O64;
  } else {
    atomicOr(s.flags, RFD_FLAG_ROW_FULL);
  }
  ++nfound;
}

// Contact-row rebuild by cell blocks: a workgroup takes the particles of a
// 2 x 2 x 2 block of cells and copies the candidates of the surrounding
// 4 x 4 x 4 cells into LDS, one contiguous range per x-row of cells, in box
// order (z, y, x, then members by id).  Every lane then scans the whole box in
// that order: the candidates of its 27 neighbour cells come in the order the
// per-particle search visits them, and every other candidate lies more than a
// cell edge (> ri + rj + skin) away and fails the same FP32 test, so the rows
// are the same as k_nat_neighbours<2>'s.  The scan is uniform across lanes
// (broadcast LDS reads).  A block whose candidates exceed the LDS copy reads
// them from global memory in the same order.
constexpr int kRowCB = 2;                       // home block edge (cells)
constexpr int kRowHB = kRowCB + 2;              // halo block edge
constexpr uint32_t kRowHC = kRowHB * kRowHB * kRowHB;
constexpr uint32_t kRowCand = 384;              // candidates held in LDS
constexpr uint32_t kRowStage = 16;              // bonded partners held per particle
constexpr uint32_t kRowFound = 16;              // in-reach candidates listed per particle

__global__ void __launch_bounds__(64) k_nat_rows_cells(NbrDev s, uint32_t nblocks, uint32_t per,
                                                       uint32_t nbx, uint32_t nby) {
  __shared__ float4 sC[kRowCand];
  __shared__ uint32_t sJ[kRowCand];
  __shared__ uint32_t sS[kRowHC + 1], sG[kRowHC];
  __shared__ uint16_t sF[kRowFound][64];  // box index of each in-reach candidate
  __shared__ uint32_t sHome[kRowCB * kRowCB + 1], sHomeG[kRowCB * kRowCB];
  const uint32_t b = xcd_tile(per);
  if (b >= nblocks) return;
  const GridDev &g = s.g;
  const uint32_t tid = threadIdx.x;
  const int x0 = (int)(b % nbx) * kRowCB, y0 = (int)((b / nbx) % nby) * kRowCB,
            z0 = (int)(b / (nbx * nby)) * kRowCB;
  const int nx = (int)g.nx, ny = (int)g.ny, nz = (int)g.nz;
  auto gcell = [&](int x, int y, int z) {
    return (uint32_t)x + (uint32_t)y * g.nx + (uint32_t)z * g.nx * g.ny;
  };
  // home particles: one contiguous range per home x-row (y, z), cells x0..x1
  const int x1 = min(x0 + kRowCB - 1, nx - 1);
  {
    uint32_t hc = 0;
    if (tid < (uint32_t)(kRowCB * kRowCB)) {
      const int y = y0 + (int)tid % kRowCB, z = z0 + (int)tid / kRowCB;
      uint32_t a = 0;
      if (y < ny && z < nz) {
        a = s.start[gcell(x0, y, z)];
        hc = s.start[gcell(x1, y, z) + 1] - a;
      }
      sHomeG[tid] = a;
    }
    uint32_t incl = hc;
#pragma unroll
    for (int o = 1; o < kRowCB * kRowCB; o <<= 1) {
      const uint32_t v = __shfl_up(incl, o, 64);
      if ((int)tid >= o) incl += v;
    }
    if (tid < (uint32_t)(kRowCB * kRowCB)) sHome[tid] = incl - hc;
    if (tid == kRowCB * kRowCB - 1) sHome[kRowCB * kRowCB] = incl;
  }
  __syncthreads();
  const uint32_t nhome = sHome[kRowCB * kRowCB];
  if (nhome == 0) return;
  // halo cells [x0-1, x0+kRowCB] x ... clipped to the grid, box order (x fastest)
  const int hx0 = max(x0 - 1, 0), hy0 = max(y0 - 1, 0), hz0 = max(z0 - 1, 0);
  const int hx1 = min(x0 + kRowCB, nx - 1), hy1 = min(y0 + kRowCB, ny - 1),
            hz1 = min(z0 + kRowCB, nz - 1);
  const int hnx = hx1 - hx0 + 1, hny = hy1 - hy0 + 1, hnz = hz1 - hz0 + 1;
  const uint32_t nh = (uint32_t)(hnx * hny * hnz);
  uint32_t cnt = 0;
  if (tid < nh) {
    const int qx = (int)tid % hnx, qy = ((int)tid / hnx) % hny, qz = (int)tid / (hnx * hny);
    const uint32_t gc = gcell(hx0 + qx, hy0 + qy, hz0 + qz);
    const uint32_t a = s.start[gc];
    sG[tid] = a;
    cnt = s.start[gc + 1] - a;
  }
  // exclusive scan of the counts over the (at most 64) halo cells: one wave
  uint32_t incl = cnt;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t v = __shfl_up(incl, o, 64);
    if ((int)tid >= o) incl += v;
  }
  if (tid < nh) sS[tid] = incl - cnt;
  if (tid == nh - 1) sS[nh] = incl;
  __syncthreads();
  const uint32_t total = sS[nh];
  const bool staged = total <= kRowCand;
  if (staged) {
    for (uint32_t e = tid; e < total; e += 64) {
      uint32_t lo = 0, hi = nh;
      while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (sS[mid] <= e) lo = mid; else hi = mid;
      }
      const uint32_t src = sG[lo] + (e - sS[lo]);
      sC[e] = s.cpos[src];
      sJ[e] = s.sorted[src];
    }
  }
  __syncthreads();
  const float exf = (float)s.extra;
  for (uint32_t k0 = 0; k0 < nhome; k0 += 64) {
    const uint32_t k = k0 + tid;
    const bool live = k < nhome;
    uint32_t i = 0;
    float4 me = make_float4(0.f, 0.f, 0.f, 0.f);
    bool fits = true;
    uint32_t jb[kRowStage];  // bonded partners (kSentinel: none or removed)
#pragma unroll
    for (uint32_t t = 0; t < kRowStage; ++t) jb[t] = kSentinel;
    if (live) {
      uint32_t r = 0;
      while (r + 1 < kRowCB * kRowCB && sHome[r + 1] <= k) ++r;
      const uint32_t eg = sHomeG[r] + (k - sHome[r]);
      i = s.sorted[eg];
      me = s.cpos[eg];
      const uint32_t r0 = s.hoff[i], nb = s.hoff[i + 1] - r0;
      fits = nb <= kRowStage;
      if (fits) {
        HR hh[kRowStage];
#pragma unroll
        for (uint32_t t = 0; t < kRowStage; ++t)
          if (t < nb) hh[t] = s.hr[r0 + t];
#pragma unroll
        for (uint32_t t = 0; t < kRowStage; ++t)
          if (t < nb && s.fl8[hh[t].b] != kRemoved) jb[t] = hh[t].j;
      }
    }
    // in-reach test of one candidate (FP32, as the per-particle search)
    auto reach = [&](const float4 q4) {
      const float ddx = q4.x - me.x, ddy = q4.y - me.y, ddz = q4.z - me.z;
      const float d2 = ddx * ddx + ddy * ddy + ddz * ddz;
      const float lim = me.w + q4.w + exf;
      return d2 >= 1e-10f && d2 < lim * lim;
    };
    auto is_bonded = [&](uint32_t j) {
      if (!fits) return bonded(s.hoff, s.hr, s.fl8, i, j);
      bool bd = false;
#pragma unroll
      for (uint32_t t = 0; t < kRowStage; ++t) bd |= jb[t] == j;
      return bd;
    };
    uint32_t nfound = 0;
    if (staged) {
      // 1. in-reach candidates in scan order (box indices), per lane: the
      //    lane's 27 cells are 9 x-rows of three cells, each a contiguous
      //    range of the box copy, visited in (z, y) order as the per-particle
      //    search visits them
      uint32_t nf = 0;
      if (live) {
        // the home cell in box coordinates: the home x-row's first cell, or
        // the next one along x when the particle lies past the first's members
        uint32_t r = 0;
        while (r + 1 < kRowCB * kRowCB && sHome[r + 1] <= k) ++r;
        const int hy = y0 + (int)r % kRowCB - hy0, hz = z0 + (int)r / kRowCB - hz0;
        const int c0 = (x0 - hx0) + hy * hnx + hz * hnx * hny;
        const uint32_t eg = sHomeG[r] + (k - sHome[r]);
        const int hxb = (x0 - hx0) + (eg >= sG[c0] + (sS[c0 + 1] - sS[c0]) ? 1 : 0);
        const int xa = max(hxb - 1, 0), xb = min(hxb + 1, hnx - 1);
        for (int dz = -1; dz <= 1; ++dz) {
          const int z = hz + dz;
          if (z < 0 || z >= hnz) continue;
          for (int dy = -1; dy <= 1; ++dy) {
            const int y = hy + dy;
            if (y < 0 || y >= hny) continue;
            const int row = y * hnx + z * hnx * hny;
            const uint32_t e1 = sS[row + xb + 1];
            for (uint32_t e = sS[row + xa]; e < e1; ++e)
              if (reach(sC[e])) {
                if (nf < kRowFound) sF[nf][tid] = (uint16_t)e;
                ++nf;
              }
          }
        }
      }
      // 2. drop self and bonded partners, write the row in scan order
      if (live && nf <= kRowFound) {
        for (uint32_t f = 0; f < nf; ++f) {
          const uint32_t j = sJ[sF[f][tid]];
          if (j != i && !is_bonded(j)) row_entry(s, i, j, nfound);
        }
      } else if (live) {
        // more in-reach candidates than the list holds: scan again
        for (uint32_t q = 0; q < total; ++q)
          if (reach(sC[q])) {
            const uint32_t j = sJ[q];
            if (j != i && !is_bonded(j)) row_entry(s, i, j, nfound);
          }
      }
    } else if (live) {
      // candidates read from global memory in box order
      for (uint32_t q = 0; q < nh; ++q) {
        const uint32_t a = sG[q], n = sS[q + 1] - sS[q];
        for (uint32_t e = a; e < a + n; ++e)
          if (reach(s.cpos[e])) {
            const uint32_t j = s.sorted[e];
            if (j != i && !is_bonded(j)) row_entry(s, i, j, nfound);
          }
      }
    }
    if (live) s.ccnt_new[i] = nfound < s.Kc ? nfound : s.Kc;
  }
}

// rest length = current distance, set by the owner half
__global__ void k_set_rest_length(const P4 *__restrict__ P, uint32_t N,
                                  const uint32_t *__restrict__ hoff, const HR *__restrict__ hr,
                                  BR *__restrict__ brc) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += grid_stride()) {
    const P4 A = P[i];
    for (uint32_t h = hoff[i]; h < hoff[i + 1]; ++h) {
      const HR q = hr[h];
      if (q.j < i) continue;
      const P4 B = P[q.j];
      const double rx = B.x - A.x, ry = B.y - A.y, rz = B.z - A.z;
      brc[q.b].il = (float)(1.0 / __builtin_sqrt(rx * rx + ry * ry + rz * rz));
    }
  }
}

__global__ void k_fill_u32n(uint32_t *p, size_t n, uint32_t v) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)blockDim.x * gridDim.x)
    p[i] = v;
}

}  // namespace

// ------------------------------------------------------------------- FIRE
// Standard FIRE (Bitzek et al. 2006, with the half-step back-off of FIRE 2.0)
// over translations and rotations.  The reference's variant
// (md_firerelaxation.cu:18-84, 263-304) recomputes V from F alone and runs on
// translations of unbonded packings only.
struct FireDev {
  double dt, dt_last, alpha;
  double cV, cW;  // mixing coefficients alpha*|V|/|F|, alpha*|W|/|M|
  double fmax, mmax, fmax0, sum_bottom, sum_top;
  double dt_max, dt_min, alpha0, f_inc, f_dec, f_alpha, ftol;
  uint32_t n_pos, iter, max_iter, n_delay;
  int done, converged, backoff, first, measure_only;
  float dt_f;   // time step of the last move, read by the force pass
  double dt_m;  // the same for the FP64 force pass
};

constexpr int kFireVals = 9;  // P, |F|^2, |V|^2, |M|^2, |W|^2, max|F|^2, max(|M|/r)^2, sb, st

// FIRE keeps its own translational velocity Vf; V holds the rate of the last
// move (displacement / dt_f, back-off included), which is what the force pass
// reads for the incremental springs and the previous bond direction
struct FireParticles {
  uint32_t N;
  P4 *P;
  float4 *V, *Vf, *W, *F, *Th;
  const P4 *P0;
  const float4 *M;
  const float2 *inv;
  float dt_phys;
  rfd_loading *ctl;
  FireDev *fire;
  double *part;
};

// platen constraint with the platens at rest, reaction sums, power and norms
template <int MODE>
__global__ void __launch_bounds__(256) k_fire_measure(FireParticles p) {
  __shared__ double s_part[4 * kFireVals];
  if (p.fire->done) return;
  rfd_loading L = *p.ctl;
  L.v = 0.0;
  double v[kFireVals] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < p.N; i += grid_stride()) {
    P4 R = p.P[i];
    v3 V = xyz(p.Vf[i]), F = xyz(p.F[i]);
    if (MODE) {
      const v3 V0 = V, F0 = F;
      const double z0 = R.z, y0 = R.y;
      platen<MODE>(L, R, V, F, v[7], v[8], false);
      if (R.z != z0 || R.y != y0) p.P[i] = R;
      if (V.x != V0.x || V.y != V0.y || V.z != V0.z) p.Vf[i] = f4(V);
      if (F.x != F0.x || F.y != F0.y || F.z != F0.z) p.F[i] = f4(F);
    }
    const v3 W = xyz(p.W[i]), M = xyz(p.M[i]);
    v[0] += (double)dot(F, V) + (double)dot(M, W);
    const double f2 = dot(F, F), m2 = dot(M, M);
    v[1] += f2;
    v[2] += (double)dot(V, V);
    v[3] += m2;
    v[4] += (double)dot(W, W);
    v[5] = fmax(v[5], f2);
    v[6] = fmax(v[6], m2 / (R.r * R.r));
  }
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < kFireVals; ++q) v[q] = (q == 5 || q == 6) ? wave_max(v[q]) : wave_sum(v[q]);
  if (lane == 0)
    for (int q = 0; q < kFireVals; ++q) s_part[wid * kFireVals + q] = v[q];
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int q = 0; q < kFireVals; ++q) {
      double a = 0;
      for (int w = 0; w < (int)(blockDim.x >> 6); ++w)
        a = (q == 5 || q == 6) ? fmax(a, s_part[w * kFireVals + q]) : a + s_part[w * kFireVals + q];
      p.part[blockIdx.x * kFireVals + q] = a;
    }
  }
}

// reduce the partials in a fixed order, then one thread runs the FIRE logic;
// curp, when set, is the bond-state index the iteration's force pass left to
// be flipped
__global__ void __launch_bounds__(256) k_fire_control(FireDev *fire, const double *part,
                                                      int nblocks, rfd_loading *ctl,
                                                      int *curp) {
  __shared__ double s_part[4 * kFireVals];
  if (fire->done) return;
  double v[kFireVals] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (int b = threadIdx.x; b < nblocks; b += blockDim.x)
    for (int q = 0; q < kFireVals; ++q) {
      const double x = part[b * kFireVals + q];
      v[q] = (q == 5 || q == 6) ? fmax(v[q], x) : v[q] + x;
    }
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  for (int q = 0; q < kFireVals; ++q) v[q] = (q == 5 || q == 6) ? wave_max(v[q]) : wave_sum(v[q]);
  if (lane == 0)
    for (int q = 0; q < kFireVals; ++q) s_part[wid * kFireVals + q] = v[q];
  __syncthreads();
  if (threadIdx.x != 0) return;
  if (curp) *curp = 1 - *curp;
  for (int q = 0; q < kFireVals; ++q) {
    double a = 0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w)
      a = (q == 5 || q == 6) ? fmax(a, s_part[w * kFireVals + q]) : a + s_part[w * kFireVals + q];
    v[q] = a;
  }
  FireDev f = *fire;
  f.fmax = __builtin_sqrt(v[5]);
  f.mmax = __builtin_sqrt(v[6]);
  f.sum_bottom = v[7];
  f.sum_top = v[8];
  if (f.measure_only) {
    f.done = 1;
    *fire = f;
    return;
  }
  if (f.first) {
    f.fmax0 = f.fmax;
    f.first = 0;
  }
  f.converged = f.fmax < f.ftol && f.mmax < f.ftol;
  if (f.converged || f.iter >= f.max_iter) {
    f.done = 1;
    ctl->sum_bottom = v[7];
    ctl->sum_top = v[8];
    *fire = f;
    return;
  }
  f.iter += 1;
  f.dt_last = f.dt;
  if (v[2] == 0.0 && v[4] == 0.0) {
    // at rest (first iteration of a step): no power to judge, plain kick
    f.backoff = 0;
    f.cV = f.cW = 0.0;
  } else if (v[0] > 0.0) {
    f.backoff = 0;
    f.cV = v[1] > 0 ? f.alpha * __builtin_sqrt(v[2] / v[1]) : 0.0;
    f.cW = v[3] > 0 ? f.alpha * __builtin_sqrt(v[4] / v[3]) : 0.0;
    f.n_pos += 1;
    if (f.n_pos > f.n_delay) {
      f.dt = fmin(f.dt * f.f_inc, f.dt_max);
      f.alpha *= f.f_alpha;
    }
  } else {
    f.backoff = 1;
    f.n_pos = 0;
    f.dt = fmax(f.dt * f.f_dec, f.dt_min);
    f.alpha = f.alpha0;
  }
  f.dt_f = (float)f.dt;
  f.dt_m = f.dt;
  *fire = f;
}

static_assert(kMeasVals == kFireVals, "measure partials");

// second stage of the fused measure's reduction: workgroup b adds (or takes
// the maxima of) a contiguous run of the per-wave partials, in a fixed order,
// into the partial pair slot b that k_fire_control reduces
__global__ void __launch_bounds__(256) k_fire_reduce(const FireDev *fire,
                                                     const double *__restrict__ tpart,
                                                     uint32_t nparts, double *__restrict__ part) {
  __shared__ double s_part[4 * kFireVals];
  if (fire->done) return;
  const uint32_t chunk = (nparts + gridDim.x - 1) / gridDim.x;
  const uint32_t b0 = blockIdx.x * chunk, b1 = min(b0 + chunk, nparts);
  double v[kFireVals] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (uint32_t b = b0 + threadIdx.x; b < b1; b += blockDim.x)
    for (int q = 0; q < kFireVals; ++q) {
      const double x = tpart[(size_t)b * kFireVals + q];
      v[q] = (q == 5 || q == 6) ? fmax(v[q], x) : v[q] + x;
    }
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < kFireVals; ++q) v[q] = (q == 5 || q == 6) ? wave_max(v[q]) : wave_sum(v[q]);
  if (lane == 0)
    for (int q = 0; q < kFireVals; ++q) s_part[wid * kFireVals + q] = v[q];
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int q = 0; q < kFireVals; ++q) {
      double a = 0;
      for (int w = 0; w < (int)(blockDim.x >> 6); ++w)
        a = (q == 5 || q == 6) ? fmax(a, s_part[w * kFireVals + q]) : a + s_part[w * kFireVals + q];
      part[blockIdx.x * kFireVals + q] = a;
    }
  }
}

// velocity mixing (or back-off), semi-implicit Euler kick and move.  The
// platen clamp of positions and FIRE velocities (platens at rest) runs first:
// k_fire_measure has applied it already (it is idempotent), the measure fused
// into the force pass has not.
__global__ void __launch_bounds__(256) k_fire_update(FireParticles p) {
  // no contraction here: the rounding of the move does not depend on how the
  // compiler schedules this kernel (warm-started relaxations near equilibrium
  // take visibly different iteration counts from last-bit differences)
#pragma clang fp contract(off)
  const FireDev f = *p.fire;
  if (f.done) return;
  rfd_loading L = *p.ctl;
  L.v = 0.0;
  const float a1 = (float)(1.0 - f.alpha), cV = (float)f.cV, cW = (float)f.cW;
  const float scale = (float)(f.dt / p.dt_phys);  // inv2 holds dt_phys/m, dt_phys/I
  const float dtf = (float)f.dt;
  const double inv_dt = 1.0 / (double)dtf;  // the force pass multiplies by dt_f
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < p.N; i += grid_stride()) {
    P4 R0 = p.P[i];
    v3 V = xyz(p.Vf[i]), W = xyz(p.W[i]);
    if (L.mode) {
      v3 Fd = kZero;
      double d0 = 0, d1 = 0;
      if (L.mode == 1)
        platen<1>(L, R0, V, Fd, d0, d1, false);
      else
        platen<2>(L, R0, V, Fd, d0, d1, false);
    }
    P4 R = R0;
    const v3 F = xyz(p.F[i]), M = xyz(p.M[i]);
    if (f.backoff) {
      const double h = 0.5 * f.dt_last;
      R.x -= (double)V.x * h;
      R.y -= (double)V.y * h;
      R.z -= (double)V.z * h;
      V = kZero;
      W = kZero;
    } else {
      V = V * a1 + F * cV;
      W = W * a1 + M * cW;
    }
    const float2 iv = p.inv[i];
    V = V + F * (iv.x * scale);
    W = W + M * (iv.y * scale);
    R.x += (double)V.x * f.dt;
    R.y += (double)V.y * f.dt;
    R.z += (double)V.z * f.dt;
    p.P[i] = R;
    p.Vf[i] = f4(V);
    p.V[i] = make_float4((float)((R.x - R0.x) * inv_dt), (float)((R.y - R0.y) * inv_dt),
                         (float)((R.z - R0.z) * inv_dt), 0.f);
    p.W[i] = f4(W);
    p.Th[i] = f4(xyz(p.Th[i]) + W * dtf);
  }
}

// start of a load step: R = R0 + dU (warm start) or R0; particles held by a
// platen at the start follow it exactly (the top one moved by dtop), and any
// particle past a platen is snapped onto it (displacement boundary
// condition); V = (R - R0) / dt0 and W = dTh / dt0 so that one force pass run
// with dt0 carries the whole move into the incremental springs; FIRE starts at
// rest.
template <int MODE>
__global__ void k_fire_begin(FireParticles p, const float *__restrict__ dU,
                             const float *__restrict__ dTh, float dt0, double dtop) {
  rfd_loading L = *p.ctl;
  L.v = 0.0;
  const double inv_dt = 1.0 / (double)dt0;
  const float inv_dtf = 1.0f / dt0;
  double sb = 0, stp = 0;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < p.N; i += grid_stride()) {
    const P4 R0 = p.P[i];
    P4 R = R0;
    if (dU) {
      R.x += (double)dU[3 * (size_t)i];
      R.y += (double)dU[3 * (size_t)i + 1];
      R.z += (double)dU[3 * (size_t)i + 2];
    }
    v3 Vd = kZero, F = kZero;
    if (MODE) {
      double ra, rb, h0, zc;
      if (MODE == 1) {
        ra = R0.x - L.cx;
        rb = R0.y - L.cy;
        h0 = R0.z;
        zc = R0.z;
      } else {
        ra = R0.x - L.cx;
        rb = R0.z - L.cz;
        h0 = R0.y;
        zc = R0.z - L.cz;
      }
      if (ra * ra + rb * rb < L.RR && zc < L.zcut) {
        double &Ra = MODE == 1 ? R.z : R.y;
        if (h0 - (L.top - dtop) > -1e-9)
          Ra = L.top;
        else if (h0 - L.bottom < 1e-9)
          Ra = L.bottom;
      }
      platen<MODE>(L, R, Vd, F, sb, stp);
    }
    const v3 th = dTh ? v3{dTh[3 * (size_t)i], dTh[3 * (size_t)i + 1], dTh[3 * (size_t)i + 2]}
                      : kZero;
    p.P[i] = R;
    p.V[i] = make_float4((float)((R.x - R0.x) * inv_dt), (float)((R.y - R0.y) * inv_dt),
                         (float)((R.z - R0.z) * inv_dt), 0.f);
    p.W[i] = f4(th * inv_dtf);
    p.Th[i] = f4(th);
    p.Vf[i] = make_float4(0.f, 0.f, 0.f, 0.f);
  }
}

// displacement of the load step: du = R - R0, dth = Th (N x 3 floats)
__global__ void k_fire_step_disp(const P4 *__restrict__ P, const P4 *__restrict__ P0,
                                 const float4 *__restrict__ Th, uint32_t N,
                                 float *__restrict__ du, float *__restrict__ dth) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += grid_stride()) {
    const P4 a = P[i], b = P0[i];
    du[3 * i] = (float)(a.x - b.x);
    du[3 * i + 1] = (float)(a.y - b.y);
    du[3 * i + 2] = (float)(a.z - b.z);
    const float4 t = Th[i];
    dth[3 * i] = t.x;
    dth[3 * i + 1] = t.y;
    dth[3 * i + 2] = t.z;
  }
}

// ------------------------------------------------------------- packer
// Seeded packing by radius growth with overlap removal (the role of
// md_createcylindersample.cu:136-343, which relaxes a monodisperse soft-sphere
// gas for 200k viscous steps and then by FIRE): radii grow in stages and each
// iteration moves every particle by half of each of its overlaps (Jacobi, a
// relaxation factor beta), then keeps it inside the container.

// radii of the current stage: r = s r_final
__global__ void k_pack_scale(P4 *__restrict__ P, const float *__restrict__ rfin, double sc,
                             uint32_t N) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += grid_stride())
    P[i].r = sc * (double)rfin[i];
}

// displacement of each particle from its overlaps (contact rows), and the
// largest overlap relative to the smaller radius (as ordered float bits)
__global__ void k_pack_disp(const P4 *__restrict__ P, const uint32_t *__restrict__ ccnt,
                            const uint32_t *__restrict__ cj, uint32_t N, double beta,
                            float4 *__restrict__ D, uint32_t *__restrict__ maxov) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += grid_stride()) {
    const P4 A = P[i];
    double dx = 0.0, dy = 0.0, dz = 0.0;
    float mo = 0.f;
    const uint32_t nc = ccnt[i];
    for (uint32_t k = 0; k < nc; ++k) {
      const P4 B = P[cj[(size_t)k * N + i]];
      const double rx = B.x - A.x, ry = B.y - A.y, rz = B.z - A.z;
      const double L = __builtin_sqrt(rx * rx + ry * ry + rz * rz);
      const double ov = A.r + B.r - L;
      if (ov > 0.0 && L > 1e-12) {
        const double c = -0.5 * beta * ov / L;
        dx += c * rx;
        dy += c * ry;
        dz += c * rz;
        mo = fmaxf(mo, (float)(ov / fmin(A.r, B.r)));
      }
    }
    D[i] = make_float4((float)dx, (float)dy, (float)dz, 0.f);
    if (mo > 0.f) atomicMax(maxov, __float_as_uint(mo));
  }
}

// apply the moves, then keep each sphere inside the container: a cylinder of
// radius rc about the axis through (cx, cy) along z, between z0 and z1 (the
// clamp of d_CylinderRestrictionZ, md_cylinderrestriction.cu:74-118, applied
// to sphere surfaces).  The move is the displacement plus mom times the last
// move (heavy-ball relaxation); a clamp drops the last move.
__global__ void k_pack_apply(P4 *__restrict__ P, const float4 *__restrict__ D,
                             float4 *__restrict__ U, uint32_t N, double mom, double cx,
                             double cy, double rc, double z0, double z1) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += grid_stride()) {
    P4 A = P[i];
    const float4 d = D[i], u0 = U[i];
    const double ux = (double)d.x + mom * (double)u0.x, uy = (double)d.y + mom * (double)u0.y,
                 uz = (double)d.z + mom * (double)u0.z;
    const double x0 = A.x, y0 = A.y, z00 = A.z;
    A.x += ux;
    A.y += uy;
    A.z += uz;
    bool hit = false;
    const double ex = A.x - cx, ey = A.y - cy, lim = rc - A.r;
    const double rho2 = ex * ex + ey * ey;
    if (rho2 > lim * lim) {
      const double f = lim / __builtin_sqrt(rho2);
      A.x = cx + ex * f;
      A.y = cy + ey * f;
      hit = true;
    }
    if (A.z > z1 - A.r) {
      A.z = z1 - A.r;
      hit = true;
    }
    if (A.z < z0 + A.r) {
      A.z = z0 + A.r;
      hit = true;
    }
    P[i] = A;
    U[i] = hit ? make_float4(0.f, 0.f, 0.f, 0.f)
               : make_float4((float)(A.x - x0), (float)(A.y - y0), (float)(A.z - z00), 0.f);
  }
}

__global__ void k_zero_f4(float4 *__restrict__ a, float4 *__restrict__ b, uint32_t N) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += grid_stride()) {
    a[i] = make_float4(0.f, 0.f, 0.f, 0.f);
    if (b) b[i] = make_float4(0.f, 0.f, 0.f, 0.f);
  }
}

// inputs of the learned warm start: per particle the last two step
// displacements and rotations (scaled), per half the bond strain (x 1e3, zero
// unless the bond is intact) and the unit direction to the other end
__global__ void k_mpnn_inputs(const P4 *__restrict__ P, uint32_t N,
                              const uint32_t *__restrict__ hoff, const HR *__restrict__ hr,
                              const BR *__restrict__ brc, const uint8_t *__restrict__ fl8,
                              const float *__restrict__ du0, const float *__restrict__ du1,
                              const float *__restrict__ dth0, const float *__restrict__ dth1,
                              float isu, float isth, float E_mod, float *__restrict__ X,
                              float *__restrict__ E) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += grid_stride()) {
    float *x = X + 32 * (size_t)i;
    for (int d = 0; d < 3; ++d) {
      x[d] = du0[3 * (size_t)i + d] * isu;
      x[3 + d] = du1[3 * (size_t)i + d] * isu;
      x[6 + d] = dth0[3 * (size_t)i + d] * isth;
      x[9 + d] = dth1[3 * (size_t)i + d] * isth;
    }
    for (int d = 12; d < 32; ++d) x[d] = 0.f;
    const P4 A = P[i];
    // utilisation of the particle's bonds: the tensile stress of the stretch
    // over the bond's tensile strength (the first term of the break
    // criterion, md_calculateforcesdem3.cu:23), its largest and mean value,
    // and the share of its bond records already broken
    float umax = 0.f, usum = 0.f, nint = 0.f, nbrk = 0.f;
    for (uint32_t h = hoff[i]; h < hoff[i + 1]; ++h) {
      const HR q = hr[h];
      const P4 B = P[q.j];
      const double rx = B.x - A.x, ry = B.y - A.y, rz = B.z - A.z;
      const double L = __builtin_sqrt(rx * rx + ry * ry + rz * rz);
      const double il = L > 0 ? 1.0 / L : 0.0;
      const bool intact = fl8[q.b] == kBond;
      const double strain = L * (double)brc[q.b].il - 1.0;
      const float eps = intact ? (float)(strain * 1e3) : 0.f;
      if (intact) {
        const float u = (float)((double)E_mod * strain / (double)brc[q.b].sc);
        umax = fmaxf(umax, u);
        usum += u;
        nint += 1.f;
      } else {
        nbrk += 1.f;
      }
      *reinterpret_cast<float4 *>(E + 4 * (size_t)h) =
          make_float4(eps, (float)(rx * il), (float)(ry * il), (float)(rz * il));
    }
    x[12] = umax;
    x[13] = nint > 0.f ? usum / nint : 0.f;
    x[14] = nint + nbrk > 0.f ? nbrk / (nint + nbrk) : 0.f;
  }
}

// =============================================================== FP64 path
// The same native-tier step (bond law with break check, contact law, platens,
// integration, loading control) with every quantity in double: the GPU
// precision oracle of the FP32 path.  One thread per particle walks its
// half-bond row and contact row in order (deterministic, no LDS staging: this
// path is for checking, not for speed).  Contact rows, cells, positions and
// the loading control are shared with the FP32 path.
struct d3 {
  double x, y, z;
};
__device__ __forceinline__ d3 operator+(d3 a, d3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ d3 operator-(d3 a, d3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ d3 operator*(d3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ double dot(d3 a, d3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ d3 cross(d3 a, d3 b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ d3 xyz(double4 q) { return {q.x, q.y, q.z}; }
__device__ __forceinline__ double4 d4(d3 v) { return make_double4(v.x, v.y, v.z, 0.0); }
__device__ constexpr d3 kZero64 = {0.0, 0.0, 0.0};

struct alignas(16) BSt64 {
  double ox, oy, oz, mn, mx, my, mz;
  uint32_t fl, pad;
};

struct Mat64 {
  double E, G, Estar, Gstar, pA, mu, mur, rho43pi, dt, farcut;
  int refmode;
  int nobreak;  // quasi-static relaxation: bonds are judged after convergence
};

struct Rot64 {
  double m[9];
  __device__ __forceinline__ void build(d3 t, d3 ph) {
    m[0] = 1.0 + t.z * ph.z + t.y * ph.y;
    m[1] = ph.z - t.z - t.y * ph.x;
    m[2] = -ph.y - t.z * ph.x + t.y;
    m[3] = t.z - ph.z - t.x * ph.y;
    m[4] = t.z * ph.z + 1.0 + t.x * ph.x;
    m[5] = -t.z * ph.y + ph.x - t.x;
    m[6] = -t.y - t.x * ph.z + ph.y;
    m[7] = -t.y * ph.z + t.x - ph.x;
    m[8] = t.y * ph.y + t.x * ph.x + 1.0;
  }
  __device__ __forceinline__ d3 apply(d3 x) const {
    return {m[0] * x.x + m[1] * x.y + m[2] * x.z, m[3] * x.x + m[4] * x.y + m[5] * x.z,
            m[6] * x.x + m[7] * x.y + m[8] * x.z};
  }
};

__device__ __forceinline__ void contact_law64(const Mat64 &mt, d3 n, double L, double ri,
                                              double rj, double reach, d3 Vi, d3 Vj, d3 Wi,
                                              d3 Wj, d3 &O, d3 &Fi, d3 &Mi, d3 &Mj) {
  const double Rs = ri * rj / (ri + rj);
  const double mi = mt.rho43pi * ri * ri * ri, mj = mt.rho43pi * rj * rj * rj;
  const double ms = mi * mj / (mi + mj);
  const double q = __builtin_sqrt(Rs * (reach - L));
  const double Kn = 2.0 * mt.Estar * q, Kt = 8.0 * mt.Gstar * q;
  const d3 v = Vj - Vi + cross(n, Wi * ri + Wj * rj);
  const double vn = dot(n, v);
  const double fn = (L - reach) * Kn * (2.0 / 3.0);
  const double dmp = -1.8257 * mt.pA * vn * __builtin_sqrt(Kn * ms);
  Fi = n * (fn + dmp);
  const d3 vt = v - n * vn;
  d3 o = O - n * dot(n, O);
  const double oo = dot(o, o);
  if (oo > 1e-18) o = o * __builtin_sqrt(dot(O, O) / oo);
  o = o + vt * mt.dt;
  d3 ft = o * Kt;
  const double ftm = dot(ft, ft);
  if (ftm > mt.mu * mt.mu * fn * fn) {
    ft = ft * (mt.mu * fabs(fn) / __builtin_sqrt(ftm));
    o = ft * (1.0 / Kt);
  } else {
    ft = ft + vt * (-1.8257 * mt.pA * __builtin_sqrt(Kt * ms));
  }
  O = o;
  Fi = Fi + ft;
  const d3 nxf = cross(n, ft);
  Mi = nxf * ri;
  Mj = nxf * rj;
  const double wi2 = dot(Wi, Wi), wj2 = dot(Wj, Wj);
  if (wi2 > 1e-18) Mi = Mi + Wi * (-mt.mur * fabs(fn) / __builtin_sqrt(wi2));
  if (wj2 > 1e-18) Mj = Mj + Wj * (-mt.mur * fabs(fn) / __builtin_sqrt(wj2));
}

struct Sys64 {
  uint32_t N, Kc;
  const P4 *P;
  const double4 *V, *W;
  double4 *F, *M;
  const uint32_t *hoff;
  const HR *hr;
  const BR *brc;
  BSt64 *b0, *b1;
  const int *curp;
  uint8_t *fl8;
  const uint32_t *cj, *ccnt;
  double4 *cO;
  const double *dtp;  // when set, the time step of the last move (FIRE) replaces mt.dt
  const int *done;    // when set and nonzero (relaxation finished), the pass does nothing
  int trial;          // trial pass: writes only F, M and st_out
  uint32_t *nbreak;   // when set, counts the bonds this pass breaks
};

// Lanes per particle in the FP64 force pass: the lanes of a group evaluate a
// row's entries side by side, and every lane then adds the group's terms in
// lane order, which is row order.
constexpr uint32_t kG64 = 8;

__device__ __forceinline__ double grp_get(double v, uint32_t k) {
  return __shfl(v, (int)k, (int)kG64);
}
__device__ __forceinline__ d3 grp_get(d3 v, uint32_t k) {
  return {grp_get(v.x, k), grp_get(v.y, k), grp_get(v.z, k)};
}

// One bond half of row i in the FP64 pass (md_calculateforcesdem3.cu:424-483
// in FP64; the crossing of the far cut removes the bond): its load on i (dF,
// and the bond moments dMb), the contact terms (dMc) once the bond is broken
// and the pair touches (touch).  The owner half (i < j) writes the bond
// state.  Returns false when the entry adds nothing.
__device__ __forceinline__ bool f64_half(const Sys64 &s, const Mat64 &mt, BSt64 *st_in,
                                         BSt64 *st_out, uint32_t i, const P4 &A, const d3 &Vi,
                                         const d3 &Wi, uint32_t h, d3 &dF, d3 &dMb, d3 &dMc,
                                         bool &touch) {
  const HR q = s.hr[h];
  const BSt64 old = st_in[q.b];
  if (old.fl == kRemoved) return false;
  const bool own = i < q.j;
  const double sg = own ? 1.0 : -1.0;
  const P4 B = s.P[q.j];
  const d3 Vj = xyz(s.V[q.j]), Wj = xyz(s.W[q.j]);
  const BR c = s.brc[q.b];
  const d3 r = {B.x - A.x, B.y - A.y, B.z - A.z};
  const double r2 = dot(r, r);
  const double lim = mt.farcut * (A.r + B.r);
  BSt64 nw = old;
  if (r2 > lim * lim || !(r2 > 1e-10)) {
    if (own) {
      st_out[q.b].fl = kRemoved;
      if (!s.trial) {
        st_in[q.b].fl = kRemoved;
        s.fl8[q.b] = kRemoved;
      }
    }
    return false;
  }
  const double L = __builtin_sqrt(r2);
  const d3 n = r * (1.0 / L);
  d3 O = d3{old.ox, old.oy, old.oz} * sg;
  bool contact_now = old.fl == kBroken;
  bool any = false;
  if (old.fl == kBond) {
    const d3 rp = r - (Vj - Vi) * mt.dt;
    const double pp = dot(rp, rp);
    const d3 np = pp > 0.0 ? rp * (1.0 / __builtin_sqrt(pp)) : n;
    const d3 Mt0 = d3{old.mx, old.my, old.mz} * sg;
    const double il = (double)c.il, br = (double)c.br;
    const double A2 = 3.14159265358979323846 * br * br, J = A2 * br * br;
    const double eps = L * il - 1.0;
    const double sig = mt.E * eps;
    const d3 Fn = n * (sig * A2);
    const d3 ws = Wi + Wj;
    Rot64 M;
    M.build(cross(np, n), n * (dot(ws, n) * mt.dt * 0.5));
    const d3 dw = Wj - Wi;
    const double dwn = dot(n, dw);
    const double mn = old.mn + dwn * (mt.dt * 0.5 * mt.G * il);
    const double cb = mt.dt * 0.25 * mt.E * il;
    const d3 Mt = M.apply(Mt0) + (dw - n * dwn) * cb;
    const d3 v = Vj - Vi - cross(ws * 0.5, r);
    const double vn = dot(n, v);
    O = M.apply(O) + (v - n * vn) * mt.dt;
    const d3 Ft = O * (mt.G * il * A2);
    const double mtf = mt.refmode ? cb * J : J;
    dF = Fn + Ft;
    dMb = n * (mn * J) + Mt * mtf + cross(r, Ft) * 0.5;
    any = true;
    const double Sc = sig + __builtin_sqrt(dot(Mt, Mt)) * 4.0 * br;
    const double Mcr = mt.G * il * __builtin_sqrt(dot(O, O)) + fabs(mn) * 2.0 * br;
    if (!mt.nobreak && (Sc > (double)c.sc || Mcr > (double)c.tc)) {
      nw.fl = kBroken;
      O = kZero64;
      nw.mn = nw.mx = nw.my = nw.mz = 0.0;
      contact_now = true;
    } else {
      const d3 Os = O * sg, Ms = Mt * sg;
      nw.ox = Os.x;
      nw.oy = Os.y;
      nw.oz = Os.z;
      nw.mn = mn;
      nw.mx = Ms.x;
      nw.my = Ms.y;
      nw.mz = Ms.z;
    }
  }
  if (contact_now) {
    const double reach = mt.refmode ? A.r + B.r : fmin(A.r + B.r, 1.0 / (double)c.il);
    if (L < reach) {
      d3 Fc, Mci, Mcj;
      contact_law64(mt, n, L, A.r, B.r, reach, Vi, Vj, Wi, Wj, O, Fc, Mci, Mcj);
      dF = dF + Fc;
      dMc = Mci;
      touch = true;
      any = true;
    } else {
      O = kZero64;
    }
    const d3 Os = O * sg;
    nw.ox = Os.x;
    nw.oy = Os.y;
    nw.oz = Os.z;
  }
  if (own) {
    st_out[q.b] = nw;
    if (!s.trial && nw.fl != old.fl) {
      s.fl8[q.b] = (uint8_t)nw.fl;
      if (s.nbreak && nw.fl == kBroken) atomicAdd(s.nbreak, 1u);
    }
  }
  return any;
}

// Contact entry k of particle i's contact row (md_calculateforcesdem3.cu:
// 338-422 in FP64); writes the entry's spring.  Returns whether it touches.
__device__ __forceinline__ bool f64_contact(const Sys64 &s, const Mat64 &mt, uint32_t i,
                                            const P4 &A, const d3 &Vi, const d3 &Wi,
                                            uint32_t k, d3 &dF, d3 &dMc) {
  const size_t sl = (size_t)k * s.N + i;
  const uint32_t j = s.cj[sl];
  const P4 B = s.P[j];
  const d3 r = {B.x - A.x, B.y - A.y, B.z - A.z};
  const double L = __builtin_sqrt(dot(r, r));
  if (L < A.r + B.r && L > 1e-5) {
    d3 O = xyz(s.cO[sl]);
    d3 Mj;
    contact_law64(mt, r * (1.0 / L), L, A.r, B.r, A.r + B.r, Vi, xyz(s.V[j]), Wi,
                  xyz(s.W[j]), O, dF, dMc, Mj);
    if (!s.trial) s.cO[sl] = d4(O);
    return true;
  }
  if (!s.trial) s.cO[sl] = make_double4(0.0, 0.0, 0.0, 0.0);
  return false;
}

// FP64 force pass: kG64 lanes per particle take its bond halves, then its
// contact entries, kG64 at a time; each lane adds the group's terms in row
// order, so every lane holds the row sums and the first one writes them.
__global__ void __launch_bounds__(256) k_f64_forces(Sys64 s, Mat64 mt) {
  if (s.done && *s.done) return;
  if (s.dtp) mt.dt = *s.dtp;
  BSt64 *const st_in = *s.curp ? s.b1 : s.b0;
  BSt64 *const st_out = *s.curp ? s.b0 : s.b1;
  const uint32_t g = threadIdx.x % kG64;
  const uint32_t stride = grid_stride() / kG64;
  for (uint32_t i = (blockIdx.x * blockDim.x + threadIdx.x) / kG64; i < s.N; i += stride) {
    const P4 A = s.P[i];
    const d3 Vi = xyz(s.V[i]), Wi = xyz(s.W[i]);
    d3 F = kZero64, Mb = kZero64, Mc = kZero64;
    bool anyc = false;
    const uint32_t h1 = s.hoff[i + 1];
    for (uint32_t h = s.hoff[i]; h < h1; h += kG64) {
      d3 dF = kZero64, dMb = kZero64, dMc = kZero64;
      bool touch = false;
      const bool on = h + g < h1 &&
                      f64_half(s, mt, st_in, st_out, i, A, Vi, Wi, h + g, dF, dMb, dMc, touch);
      const int flags = (on ? 1 : 0) | (touch ? 2 : 0);
      for (uint32_t k = 0; k < kG64; ++k) {
        const int fk = __shfl(flags, (int)k, (int)kG64);
        const d3 a = grp_get(dF, k), b = grp_get(dMb, k), c = grp_get(dMc, k);
        if (fk & 1) {
          F = F + a;
          Mb = Mb + b;
          Mc = Mc + c;
        }
        anyc = anyc || (fk & 2);
      }
    }
    const uint32_t nc = s.ccnt[i];
    for (uint32_t k0 = 0; k0 < nc; k0 += kG64) {
      d3 dF = kZero64, dMc = kZero64;
      const bool on = k0 + g < nc && f64_contact(s, mt, i, A, Vi, Wi, k0 + g, dF, dMc);
      const int flag = on ? 1 : 0;
      for (uint32_t k = 0; k < kG64; ++k) {
        const int fk = __shfl(flag, (int)k, (int)kG64);
        const d3 a = grp_get(dF, k), c = grp_get(dMc, k);
        if (fk) {
          F = F + a;
          Mc = Mc + c;
          anyc = true;
        }
      }
    }
    if (mt.refmode && anyc && !(dot(Wi, Wi) > 1e-18)) Mc = kZero64;
    if (g == 0) {
      s.F[i] = d4(F);
      s.M[i] = d4(Mb + Mc);
    }
  }
}

// The FP64 explicit step of a small system in one launch: k_f64_forces with,
// once a particle's group has its sums, the first lane running k_f64_pass's
// arithmetic on them in registers (platens, symplectic Euler, cut box) and
// writing the next position, velocity and spin to the second buffers (other
// groups still read the current ones); then the loading control of
// k_nat_control by the last workgroup, as k_f64_pass_ctl does.  The platen
// partials are per workgroup of this launch, so the reaction sums can differ
// from the separate pass's in their last bits.
struct Adv64 {
  P4 *P;
  double4 *V, *W;     // next positions, velocities, spins
  rfd_loading *ctl;
  double *part;       // platen partials, two per workgroup
  uint32_t *done;     // workgroups done (zero between steps)
  double *hist;
  int *curp;          // bond-state index to flip
  double dt, rho43pi;
};

__device__ __forceinline__ double atomic_read_f64(double *p) {
  return __longlong_as_double(
      (long long)atomicAdd(reinterpret_cast<unsigned long long *>(p), 0ull));
}

// the partials of a workgroup, stored with atomics; true in the last
// workgroup to count itself done, whose threads then hold the fixed-order
// reduction of every workgroup's partials (k_nat_control's order) in s0, s1
__device__ __forceinline__ bool last_block_sums(double *part, uint32_t *done, double sb,
                                                double stp, double &s0, double &s1) {
  __shared__ double s_part[8];
  __shared__ uint32_t sLast;
  sb = wave_sum(sb);
  stp = wave_sum(stp);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) {
    s_part[2 * wid] = sb;
    s_part[2 * wid + 1] = stp;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double a = 0, c = 0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) {
      a += s_part[2 * w];
      c += s_part[2 * w + 1];
    }
    atomicExch(reinterpret_cast<unsigned long long *>(part + 2 * blockIdx.x),
               (unsigned long long)__double_as_longlong(a));
    atomicExch(reinterpret_cast<unsigned long long *>(part + 2 * blockIdx.x + 1),
               (unsigned long long)__double_as_longlong(c));
  }
  asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
  __syncthreads();
  if (threadIdx.x == 0) sLast = atomicAdd(done, 1u) == gridDim.x - 1 ? 1u : 0u;
  __syncthreads();
  if (!sLast) return false;
  double a = 0, c = 0;
  for (int b = threadIdx.x; b < (int)gridDim.x; b += blockDim.x) {
    a += atomic_read_f64(part + 2 * b);
    c += atomic_read_f64(part + 2 * b + 1);
  }
  a = wave_sum(a);
  c = wave_sum(c);
  __syncthreads();
  if (lane == 0) {
    s_part[2 * wid] = a;
    s_part[2 * wid + 1] = c;
  }
  __syncthreads();
  s0 = s1 = 0;
  for (int w = 0; w < (int)(blockDim.x >> 6); ++w) {
    s0 += s_part[2 * w];
    s1 += s_part[2 * w + 1];
  }
  return true;
}

template <int MODE>
__global__ void __launch_bounds__(256) k_f64_step(Sys64 s, Mat64 mt, Adv64 a) {
  BSt64 *const st_in = *s.curp ? s.b1 : s.b0;
  BSt64 *const st_out = *s.curp ? s.b0 : s.b1;
  const rfd_loading L = *a.ctl;
  double sb = 0.0, stp = 0.0;
  const uint32_t g = threadIdx.x % kG64;
  const uint32_t stride = grid_stride() / kG64;
  for (uint32_t i = (blockIdx.x * blockDim.x + threadIdx.x) / kG64; i < s.N; i += stride) {
    const P4 A = s.P[i];
    const d3 Vi = xyz(s.V[i]), Wi = xyz(s.W[i]);
    d3 F = kZero64, Mb = kZero64, Mc = kZero64;
    bool anyc = false;
    const uint32_t h1 = s.hoff[i + 1];
    for (uint32_t h = s.hoff[i]; h < h1; h += kG64) {
      d3 dF = kZero64, dMb = kZero64, dMc = kZero64;
      bool touch = false;
      const bool on = h + g < h1 &&
                      f64_half(s, mt, st_in, st_out, i, A, Vi, Wi, h + g, dF, dMb, dMc, touch);
      const int flags = (on ? 1 : 0) | (touch ? 2 : 0);
      for (uint32_t k = 0; k < kG64; ++k) {
        const int fk = __shfl(flags, (int)k, (int)kG64);
        const d3 x = grp_get(dF, k), y = grp_get(dMb, k), z = grp_get(dMc, k);
        if (fk & 1) {
          F = F + x;
          Mb = Mb + y;
          Mc = Mc + z;
        }
        anyc = anyc || (fk & 2);
      }
    }
    const uint32_t nc = s.ccnt[i];
    for (uint32_t k0 = 0; k0 < nc; k0 += kG64) {
      d3 dF = kZero64, dMc = kZero64;
      const bool on = k0 + g < nc && f64_contact(s, mt, i, A, Vi, Wi, k0 + g, dF, dMc);
      const int flag = on ? 1 : 0;
      for (uint32_t k = 0; k < kG64; ++k) {
        const int fk = __shfl(flag, (int)k, (int)kG64);
        const d3 x = grp_get(dF, k), z = grp_get(dMc, k);
        if (fk) {
          F = F + x;
          Mc = Mc + z;
          anyc = true;
        }
      }
    }
    if (mt.refmode && anyc && !(dot(Wi, Wi) > 1e-18)) Mc = kZero64;
    if (g == 0) {
      const d3 M = Mb + Mc;
      P4 R = A;
      d3 V = Vi;
      if (MODE) platen<MODE>(L, R, V, F, sb, stp);
      const double m = a.rho43pi * R.r * R.r * R.r, I = 0.4 * m * R.r * R.r;
      V = V + F * (a.dt / m);
      R.x += V.x * a.dt;
      R.y += V.y * a.dt;
      R.z += V.z * a.dt;
      const d3 W = Wi + M * (a.dt / I);
      if (L.cut) {
        const double dx = R.x - L.box_c[0], dy = R.y - L.box_c[1], dz = R.z - L.box_c[2];
        if (dx - L.box_h[0] > 0 || dx + L.box_h[0] < 0 || dy - L.box_h[1] > 0 ||
            dy + L.box_h[1] < 0 || dz - L.box_h[2] > 0 || dz + L.box_h[2] < 0) {
          R.x = L.hidden[0];
          R.y = L.hidden[1];
          R.z = L.hidden[2];
          V = kZero64;
          F = kZero64;
        }
      }
      a.P[i] = R;
      a.V[i] = d4(V);
      a.W[i] = d4(W);
      s.F[i] = d4(F);
      s.M[i] = d4(M);
    }
  }
  double s0, s1;
  if (!last_block_sums(a.part, a.done, sb, stp, s0, s1)) return;
  if (threadIdx.x != 0) return;
  a.curp[0] = 1 - a.curp[0];
  rfd_loading Lc = *a.ctl;
  control_update(Lc, s0, s1, a.hist, (uint32_t)kHistCap, a.dt);
  *a.ctl = Lc;
  atomicExch(a.done, 0u);
}

struct Pass64 {
  uint32_t N;
  P4 *P;
  double4 *V, *W, *F;
  const double4 *M;
  double dt, rho43pi;
  rfd_loading *ctl;
  double *part;
};

template <int MODE>
__global__ void __launch_bounds__(256) k_f64_pass(Pass64 p) {
  const rfd_loading L = *p.ctl;
  double sb = 0.0, stp = 0.0;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < p.N; i += grid_stride()) {
    P4 R = p.P[i];
    d3 V = xyz(p.V[i]), F = xyz(p.F[i]);
    if (MODE) platen<MODE>(L, R, V, F, sb, stp);
    const double m = p.rho43pi * R.r * R.r * R.r, I = 0.4 * m * R.r * R.r;
    V = V + F * (p.dt / m);
    R.x += V.x * p.dt;
    R.y += V.y * p.dt;
    R.z += V.z * p.dt;
    const d3 W = xyz(p.W[i]) + xyz(p.M[i]) * (p.dt / I);
    if (L.cut) {
      const double dx = R.x - L.box_c[0], dy = R.y - L.box_c[1], dz = R.z - L.box_c[2];
      if (dx - L.box_h[0] > 0 || dx + L.box_h[0] < 0 || dy - L.box_h[1] > 0 ||
          dy + L.box_h[1] < 0 || dz - L.box_h[2] > 0 || dz + L.box_h[2] < 0) {
        R.x = L.hidden[0];
        R.y = L.hidden[1];
        R.z = L.hidden[2];
        V = kZero64;
        F = kZero64;
      }
    }
    p.P[i] = R;
    p.V[i] = d4(V);
    p.W[i] = d4(W);
    p.F[i] = d4(F);
  }
  if (MODE == 0) return;
  write_partials(p.part, sb, stp);
}

// The FP64 particle pass of a small system with the loading control of
// k_nat_control run by its last workgroup (see last_tile_control): each
// workgroup's platen partials are written, and then read, with device-scope
// atomics, so that the last workgroup to count itself done sees every one; it
// reduces them in k_nat_control's fixed order (the same 256 threads), applies
// the control, flips the bond-state index and resets the counter.
template <int MODE>
__global__ void __launch_bounds__(256) k_f64_pass_ctl(Pass64 p, uint32_t *done, double *hist,
                                                      int *curp) {
  __shared__ double s_part[8];
  __shared__ uint32_t sLast;
  {
    const rfd_loading L = *p.ctl;
    double sb = 0.0, stp = 0.0;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < p.N; i += grid_stride()) {
      P4 R = p.P[i];
      d3 V = xyz(p.V[i]), F = xyz(p.F[i]);
      if (MODE) platen<MODE>(L, R, V, F, sb, stp);
      const double m = p.rho43pi * R.r * R.r * R.r, I = 0.4 * m * R.r * R.r;
      V = V + F * (p.dt / m);
      R.x += V.x * p.dt;
      R.y += V.y * p.dt;
      R.z += V.z * p.dt;
      const d3 W = xyz(p.W[i]) + xyz(p.M[i]) * (p.dt / I);
      if (L.cut) {
        const double dx = R.x - L.box_c[0], dy = R.y - L.box_c[1], dz = R.z - L.box_c[2];
        if (dx - L.box_h[0] > 0 || dx + L.box_h[0] < 0 || dy - L.box_h[1] > 0 ||
            dy + L.box_h[1] < 0 || dz - L.box_h[2] > 0 || dz + L.box_h[2] < 0) {
          R.x = L.hidden[0];
          R.y = L.hidden[1];
          R.z = L.hidden[2];
          V = kZero64;
          F = kZero64;
        }
      }
      p.P[i] = R;
      p.V[i] = d4(V);
      p.W[i] = d4(W);
      p.F[i] = d4(F);
    }
    if (MODE) {
      // write_partials' sums, stored with atomics
      sb = wave_sum(sb);
      stp = wave_sum(stp);
      const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
      if (lane == 0) {
        s_part[2 * wid] = sb;
        s_part[2 * wid + 1] = stp;
      }
      __syncthreads();
      if (threadIdx.x == 0) {
        double a = 0, c = 0;
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) {
          a += s_part[2 * w];
          c += s_part[2 * w + 1];
        }
        atomicExch(reinterpret_cast<unsigned long long *>(p.part + 2 * blockIdx.x),
                   (unsigned long long)__double_as_longlong(a));
        atomicExch(reinterpret_cast<unsigned long long *>(p.part + 2 * blockIdx.x + 1),
                   (unsigned long long)__double_as_longlong(c));
      }
    }
  }
  asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
  __syncthreads();
  if (threadIdx.x == 0) sLast = atomicAdd(done, 1u) == gridDim.x - 1 ? 1u : 0u;
  __syncthreads();
  if (!sLast) return;
  // reduce_pairs' order over the gridDim.x partial pairs
  double a = 0, c = 0;
  if (MODE) {
    for (int b = threadIdx.x; b < (int)gridDim.x; b += blockDim.x) {
      a += atomic_read_f64(p.part + 2 * b);
      c += atomic_read_f64(p.part + 2 * b + 1);
    }
  }
  a = wave_sum(a);
  c = wave_sum(c);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) {
    s_part[2 * wid] = a;
    s_part[2 * wid + 1] = c;
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  double s0 = 0, s1 = 0;
  for (int w = 0; w < (int)(blockDim.x >> 6); ++w) {
    s0 += s_part[2 * w];
    s1 += s_part[2 * w + 1];
  }
  curp[0] = 1 - curp[0];
  rfd_loading L = *p.ctl;
  control_update(L, s0, s1, hist, (uint32_t)kHistCap, p.dt);
  *p.ctl = L;
  atomicExch(done, 0u);
}

__global__ void k_f32_to_f64(const float *__restrict__ soa, uint32_t N,
                             double4 *__restrict__ out) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += grid_stride())
    out[i] = make_double4(soa[i], soa[i + (size_t)N], soa[i + 2 * (size_t)N], 0.0);
}

// ------------------------------------------------------ FP64 FIRE load steps
// The FP32 FIRE kernels (k_fire_begin / measure / update / step_disp) with
// velocities, spins, loads and the accumulated rotation in double, and 1/m,
// 1/I from the radius as the FP64 pass takes them.
struct FireParticles64 {
  uint32_t N;
  P4 *P;
  double4 *V, *Vf, *W, *F, *Th;
  const P4 *P0;
  const double4 *M;
  double rho43pi;
  rfd_loading *ctl;
  FireDev *fire;
  double *part;
};

template <int MODE>
__global__ void __launch_bounds__(256) k_fire64_measure(FireParticles64 p) {
  __shared__ double s_part[4 * kFireVals];
  if (p.fire->done) return;
  rfd_loading L = *p.ctl;
  L.v = 0.0;
  double v[kFireVals] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < p.N; i += grid_stride()) {
    P4 R = p.P[i];
    d3 V = xyz(p.Vf[i]), F = xyz(p.F[i]);
    if (MODE) {
      const d3 V0 = V, F0 = F;
      const double z0 = R.z, y0 = R.y;
      platen<MODE>(L, R, V, F, v[7], v[8], false);
      if (R.z != z0 || R.y != y0) p.P[i] = R;
      if (V.x != V0.x || V.y != V0.y || V.z != V0.z) p.Vf[i] = d4(V);
      if (F.x != F0.x || F.y != F0.y || F.z != F0.z) p.F[i] = d4(F);
    }
    const d3 W = xyz(p.W[i]), M = xyz(p.M[i]);
    v[0] += dot(F, V) + dot(M, W);
    const double f2 = dot(F, F), m2 = dot(M, M);
    v[1] += f2;
    v[2] += dot(V, V);
    v[3] += m2;
    v[4] += dot(W, W);
    v[5] = fmax(v[5], f2);
    v[6] = fmax(v[6], m2 / (R.r * R.r));
  }
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < kFireVals; ++q) v[q] = (q == 5 || q == 6) ? wave_max(v[q]) : wave_sum(v[q]);
  if (lane == 0)
    for (int q = 0; q < kFireVals; ++q) s_part[wid * kFireVals + q] = v[q];
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int q = 0; q < kFireVals; ++q) {
      double a = 0;
      for (int w = 0; w < (int)(blockDim.x >> 6); ++w)
        a = (q == 5 || q == 6) ? fmax(a, s_part[w * kFireVals + q]) : a + s_part[w * kFireVals + q];
      p.part[blockIdx.x * kFireVals + q] = a;
    }
  }
}

__global__ void __launch_bounds__(256) k_fire64_update(FireParticles64 p) {
  const FireDev f = *p.fire;
  if (f.done) return;
  const double a1 = 1.0 - f.alpha, inv_dt = 1.0 / f.dt;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < p.N; i += grid_stride()) {
    const P4 R0 = p.P[i];
    P4 R = R0;
    d3 V = xyz(p.Vf[i]), W = xyz(p.W[i]);
    const d3 F = xyz(p.F[i]), M = xyz(p.M[i]);
    if (f.backoff) {
      const double h = 0.5 * f.dt_last;
      R.x -= V.x * h;
      R.y -= V.y * h;
      R.z -= V.z * h;
      V = kZero64;
      W = kZero64;
    } else {
      V = V * a1 + F * f.cV;
      W = W * a1 + M * f.cW;
    }
    const double m = p.rho43pi * R.r * R.r * R.r, I = 0.4 * m * R.r * R.r;
    V = V + F * (f.dt / m);
    W = W + M * (f.dt / I);
    R.x += V.x * f.dt;
    R.y += V.y * f.dt;
    R.z += V.z * f.dt;
    p.P[i] = R;
    p.Vf[i] = d4(V);
    p.V[i] = make_double4((R.x - R0.x) * inv_dt, (R.y - R0.y) * inv_dt, (R.z - R0.z) * inv_dt,
                          0.0);
    p.W[i] = d4(W);
    p.Th[i] = d4(xyz(p.Th[i]) + W * f.dt);
  }
}

template <int MODE>
__global__ void k_fire64_begin(FireParticles64 p, const float *__restrict__ dU,
                               const float *__restrict__ dTh, double dt0, double dtop) {
  rfd_loading L = *p.ctl;
  L.v = 0.0;
  const double inv_dt = 1.0 / dt0;
  double sb = 0, stp = 0;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < p.N; i += grid_stride()) {
    const P4 R0 = p.P[i];
    P4 R = R0;
    if (dU) {
      R.x += (double)dU[3 * (size_t)i];
      R.y += (double)dU[3 * (size_t)i + 1];
      R.z += (double)dU[3 * (size_t)i + 2];
    }
    d3 Vd = kZero64, F = kZero64;
    if (MODE) {
      double ra, rb, h0, zc;
      if (MODE == 1) {
        ra = R0.x - L.cx;
        rb = R0.y - L.cy;
        h0 = R0.z;
        zc = R0.z;
      } else {
        ra = R0.x - L.cx;
        rb = R0.z - L.cz;
        h0 = R0.y;
        zc = R0.z - L.cz;
      }
      if (ra * ra + rb * rb < L.RR && zc < L.zcut) {
        double &Ra = MODE == 1 ? R.z : R.y;
        if (h0 - (L.top - dtop) > -1e-9)
          Ra = L.top;
        else if (h0 - L.bottom < 1e-9)
          Ra = L.bottom;
      }
      platen<MODE>(L, R, Vd, F, sb, stp);
    }
    const d3 th = dTh ? d3{dTh[3 * (size_t)i], dTh[3 * (size_t)i + 1], dTh[3 * (size_t)i + 2]}
                      : kZero64;
    p.P[i] = R;
    p.V[i] = make_double4((R.x - R0.x) * inv_dt, (R.y - R0.y) * inv_dt, (R.z - R0.z) * inv_dt,
                          0.0);
    p.W[i] = d4(th * inv_dt);
    p.Th[i] = d4(th);
    p.Vf[i] = make_double4(0.0, 0.0, 0.0, 0.0);
  }
}

__global__ void k_fire64_step_disp(const P4 *__restrict__ P, const P4 *__restrict__ P0,
                                   const double4 *__restrict__ Th, uint32_t N,
                                   float *__restrict__ du, float *__restrict__ dth) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += grid_stride()) {
    const P4 a = P[i], b = P0[i];
    du[3 * i] = (float)(a.x - b.x);
    du[3 * i + 1] = (float)(a.y - b.y);
    du[3 * i + 2] = (float)(a.z - b.z);
    const double4 t = Th[i];
    dth[3 * i] = (float)t.x;
    dth[3 * i + 1] = (float)t.y;
    dth[3 * i + 2] = (float)t.z;
  }
}

__global__ void k_zero_d4(double4 *__restrict__ a, double4 *__restrict__ b, uint32_t N) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += grid_stride()) {
    a[i] = make_double4(0.0, 0.0, 0.0, 0.0);
    if (b) b[i] = make_double4(0.0, 0.0, 0.0, 0.0);
  }
}

// ================================================================= host side
struct rfd_system {
  rfd_ctx *ctx = nullptr;
  rfd_system_params p{};
  uint32_t N = 0, Nb = 0;
  P4 *P = nullptr;
  float4 *V = nullptr, *W = nullptr, *F = nullptr, *M = nullptr;
  float2 *inv = nullptr;
  uint32_t *hoff = nullptr;
  HR *hr = nullptr;
  uint32_t *hm = nullptr;  // mirror half of each half (graph consumers)
  BR *brc = nullptr;
  uint32_t H = 0;  // half-bonds (2 per bond)
  std::vector<uint32_t> hbi, hbj;  // host copy of the bond ends
  BSt *bst[2] = {nullptr, nullptr};
  int *d_cur = nullptr;  // device index of the buffer holding the latest state
  uint32_t *scan_tot = nullptr, *scan_off = nullptr;  // chunk totals / offsets
  uint32_t scan_cap = 0;
  uint32_t *cj[2] = {nullptr, nullptr}, *ccnt[2] = {nullptr, nullptr};
  float4 *cO[2] = {nullptr, nullptr};
  int ccur = 0;
  GridDev g{};
  uint32_t C = 0;
  uint32_t *CI = nullptr, *cstart = nullptr, *ccursor = nullptr, *ctmp = nullptr,
           *csorted = nullptr, *ccnt_cells = nullptr;
  float4 *cpos = nullptr;    // cell-ordered relative positions (contact search)
  // FP64 path (params.fp64): velocities, loads, bond state and contact springs
  bool fp64 = false;
  double4 *V64 = nullptr, *W64 = nullptr, *F64 = nullptr, *M64 = nullptr;
  BSt64 *b64[2] = {nullptr, nullptr};
  double4 *cO64[2] = {nullptr, nullptr};
  uint8_t *fl8 = nullptr;    // per-bond flag, kept beside the state (contact search)
  rfd_loading *ctl = nullptr;
  double *part = nullptr;               // FP64 path: per-workgroup platen partials
  unsigned long long *acc = nullptr;    // FP32 path: fixed-point platen sums
  uint32_t *done_cnt = nullptr;         // workgroups done, for the control in the pass
  double *hist = nullptr;
  // explicit steps with the particle update fused into the force pass write
  // positions, velocities and spins to these and swap them with P, V, W; the
  // steps run in pairs, so that P, V, W hold the state between calls
  P4 *P2 = nullptr;
  float4 *V2 = nullptr, *W2 = nullptr;
  // the same for the FP64 path's one-launch steps (small systems)
  double4 *V64b = nullptr, *W64b = nullptr;
  double *part_step = nullptr;  // their platen partials, two per workgroup
  bool fuse = true;
  bool ctl_in_pass = true;  // small systems: the loading control in the force pass
  uint32_t nblocks_pass = 0;
  uint64_t step_count = 0;
  int mode = 0;  // host copy of the loading mode (selects the pass kernel)
  float rmax = 0.f;
  bool have_particles = false, have_bonds = false, rows_valid = false;
  // quasi-static load steps
  FireDev *fire = nullptr;
  float *fire_dt = nullptr;  // &fire->dt_f, handed to the force pass
  double *fire_part = nullptr;
  double *fire_tpart = nullptr;  // per-wave partials of the measure fused into the force pass
  uint32_t fire_tpart_n = 0;
  P4 *P0 = nullptr;          // positions at the start of the load step
  float4 *Vf = nullptr;      // FIRE velocity
  uint32_t *d_nbreak = nullptr;  // bonds broken by the last break-check pass
  bool forces_current = false;  // the load step's first force pass is done
  float4 *Th = nullptr;      // rotation accumulated during the load step
  double4 *Vf64 = nullptr, *Th64 = nullptr;  // the same two for the FP64 path
  double *fire_dt64 = nullptr;  // &fire->dt_m, handed to the FP64 force pass
  float *du[2] = {nullptr, nullptr}, *dth[2] = {nullptr, nullptr};
  int dcur = 0;              // index of the newest step displacement
  bool timing = false;  // force-kernel events on the context stream
  std::vector<hipEvent_t> ev;
  uint32_t ev_used = 0;
  // launch shapes (autotuned or set): force tile particles, neighbour block
  int force_fp = kFP;
  int nbr_block = 0;  // contact-row search by cell blocks
  uint32_t tile_halves[3] = {0, 0, 0};  // most half-bonds in a tile of 32, 64, 128
  bool once_off = false;  // force the chunked pass (comparison runs)
  // work records of the single-evaluation pass, for tiles of work_fp particles
  WorkRec *wrec = nullptr;
  uint32_t *wofs = nullptr, *wnin = nullptr, *wcnt = nullptr, *wcur = nullptr;
  int work_fp = 0;
  // batched specimens: S independent specimens in contiguous index ranges,
  // each with its own loading control and history
  uint32_t S = 0;
  std::vector<uint32_t> spec_first;   // host copy, S + 1 entries
  rfd_loading *ctls = nullptr;        // S controls
  double *hist_b = nullptr;           // S x hist_cap x 3
  uint32_t hist_cap_b = 0;
  uint32_t *blk_spec = nullptr, *blk_first = nullptr, *blk_end = nullptr;  // pass work list
  uint32_t nblk_b = 0;
  unsigned long long *acc_b = nullptr;  // S fixed-point platen sum pairs
  uint32_t *spec32 = nullptr;           // specimen of particle 32 k (fused batch steps)
  uint32_t *spec_first_d = nullptr;     // S + 1 first particles, on the device
  bool batch_fusable = false;           // every specimen holds at least 128 particles
  // multi-step launches (k_nat_steps): on request, when every force tile gets
  // a resident workgroup; their three fixed-point sum buffers per specimen
  bool persist = false;
  unsigned long long *acc3 = nullptr;
  uint32_t acc3_S = 0;
  // step graphs: 2 x rebuild_every steps captured once, replayed
  uint64_t gen = 1;                   // bumped by every change a captured step depends on
  hipStream_t gstream = nullptr;
  hipEvent_t gev[2] = {nullptr, nullptr};
  hipGraphExec_t gexec = nullptr;
  uint64_t g_gen = 0;
  int g_ccur = -1;
  uint32_t g_steps = 0;
};

namespace {

template <class T>
int dalloc(rfd_ctx *ctx, T **p, size_t n) {
  if (*p) {
    (void)hipFree(*p);  // hipFree waits for the device before releasing
    *p = nullptr;
  }
  if (n == 0) n = 1;
  RFD_HIP(ctx, hipMalloc((void **)p, sizeof(T) * n));
  return RFD_OK;
}

template <class T>
void dfree(T *&p) {
  if (p) (void)hipFree(p);
  p = nullptr;
}

Mat make_mat(const rfd_system_params &p) {
  Mat m{};
  const double G = 0.5 * p.E / (1.0 + p.nu);
  m.E = (float)p.E;
  m.G = (float)G;
  m.Estar = (float)(0.5 * p.E / (1.0 - p.nu * p.nu));
  m.Gstar = (float)(0.5 * G / (2.0 - p.nu));
  const double le = log(p.e);
  m.pA = (float)(le / sqrt(M_PI * M_PI + le * le));
  m.mu = (float)p.mu;
  m.mur = (float)p.mur;
  m.rho43pi = (float)(p.rho * 4.0 / 3.0 * M_PI);
  m.dt = (float)p.dt;
  m.farcut = p.farcut;
  m.refmode = p.refmode;
  return m;
}

SysDev make_dev(rfd_system *s) {
  SysDev d{};
  d.N = s->N;
  d.Kc = s->p.contact_capacity;
  d.P = s->P;
  d.V = s->V;
  d.W = s->W;
  d.F = s->F;
  d.M = s->M;
  d.hoff = s->hoff;
  d.hr = s->hr;
  d.hm = s->hm;
  d.brc = s->brc;
  d.bst0 = s->bst[0];
  d.bst1 = s->bst[1];
  d.curp = s->d_cur;
  d.fl8 = s->fl8;
  d.cj = s->cj[s->ccur];
  d.ccnt = s->ccnt[s->ccur];
  d.cO = s->cO[s->ccur];
  return d;
}

// exclusive scan of in[0..n) into start[0..n] (start[n] = total)
int scan_u32(rfd_system *s, const uint32_t *in, uint32_t n, uint32_t *start, uint32_t *cursor) {
  hipStream_t st = s->ctx->stream;
  const uint32_t nc = (n + kScanChunk - 1) / kScanChunk;
  if (nc + 1 > s->scan_cap) {
    int e = dalloc(s->ctx, &s->scan_tot, (size_t)nc + 1);
    if (!e) e = dalloc(s->ctx, &s->scan_off, (size_t)nc + 1);
    if (e) return e;
    s->scan_cap = nc + 1;
  }
  k_scan_totals<<<nc, 1024, 0, st>>>(in, n, s->scan_tot);
  k_nat_scan<<<1, 1024, 0, st>>>(s->scan_tot, nc, s->scan_off, nullptr);
  k_scan_apply<<<nc, 1024, 0, st>>>(in, n, s->scan_off, nc, start, cursor);
  RFD_HIP(s->ctx, hipGetLastError());
  return RFD_OK;
}

int bin_particles(rfd_system *s) {
  hipStream_t st = s->ctx->stream;
  k_fill_u32n<<<rfd_blocks(s->C, 256, 4096), 256, 0, st>>>(s->ccnt_cells, s->C, 0u);
  k_nat_cell_index<<<rfd_blocks(s->N, 256), 256, 0, st>>>(s->P, s->N, s->g, s->CI,
                                                          s->ccnt_cells);
  {
    int e = scan_u32(s, s->ccnt_cells, s->C, s->cstart, s->ccursor);
    if (e) return e;
  }
  k_nat_scatter<<<rfd_blocks(s->N, 256), 256, 0, st>>>(s->CI, s->N, s->ccursor, s->ctmp);
  k_nat_rank<<<rfd_blocks(s->N, 256), 256, 0, st>>>(s->ctmp, s->cstart, s->CI, s->N, s->csorted,
                                                   s->P, s->g, s->cpos);
  RFD_HIP(s->ctx, hipGetLastError());
  return RFD_OK;
}

NbrDev make_nbr(rfd_system *s, double extra) {
  NbrDev d{};
  d.N = s->N;
  d.P = s->P;
  d.g = s->g;
  d.start = s->cstart;
  d.sorted = s->csorted;
  d.cpos = s->cpos;
  d.extra = extra;
  d.Kc = s->p.contact_capacity;
  d.hoff = s->hoff;
  d.hr = s->hr;
  d.fl8 = s->fl8;
  d.flags = s->ctx->d_flags;
  return d;
}

// neighbour-search launch: ceil(N/nb) tiles spread over kXcd runs
uint32_t nbr_per(uint32_t N, uint32_t nb = 128) { return ((N + nb - 1) / nb + kXcd - 1) / kXcd; }
uint32_t nbr_grid(uint32_t N, uint32_t nb = 128) { return nbr_per(N, nb) * kXcd; }

// contact-row rebuild with the chosen shape (rows do not depend on it):
// 0 = by cell blocks with LDS staging, else threads per particle block
void launch_rows(rfd_system *s, const NbrDev &d, int nb) {
  hipStream_t q = s->ctx->stream;
  if (nb == 0) {
    const uint32_t nbx = (s->g.nx + kRowCB - 1) / kRowCB, nby = (s->g.ny + kRowCB - 1) / kRowCB,
                   nbz = (s->g.nz + kRowCB - 1) / kRowCB;
    const uint32_t nblocks = nbx * nby * nbz, per = (nblocks + kXcd - 1) / kXcd;
    k_nat_rows_cells<<<per * kXcd, 64, 0, q>>>(d, nblocks, per, nbx, nby);
  } else if (nb == 64)
    k_nat_neighbours<2, 64><<<nbr_grid(s->N, 64), 64, 0, q>>>(d, nbr_per(s->N, 64));
  else if (nb == 256)
    k_nat_neighbours<2, 256><<<nbr_grid(s->N, 256), 256, 0, q>>>(d, nbr_per(s->N, 256));
  else
    k_nat_neighbours<2, 128><<<nbr_grid(s->N), 128, 0, q>>>(d, nbr_per(s->N));
}

int rebuild_rows(rfd_system *s) {
  int st = bin_particles(s);
  if (st) return st;
  NbrDev d = make_nbr(s, s->p.skin);
  const int o = s->ccur, n = 1 - s->ccur;
  d.cj_old = s->rows_valid ? s->cj[o] : nullptr;
  d.ccnt_old = s->rows_valid ? s->ccnt[o] : nullptr;
  d.cO_old = s->cO[o];
  if (s->fp64) {
    d.cO64_old = s->cO64[o];
    d.cO64_new = s->cO64[n];
  }
  d.cj_new = s->cj[n];
  d.ccnt_new = s->ccnt[n];
  d.cO_new = s->cO[n];
  launch_rows(s, d, s->nbr_block);
  RFD_HIP(s->ctx, hipGetLastError());
  s->ccur = n;
  s->rows_valid = true;
  return RFD_OK;
}

// dtp set: a quasi-static (FIRE) pass -- the time step is that of the last
// move, and the contact law drops its velocity-dependent terms (normal and
// tangential damping, rolling resistance against the spin direction), so
// that the residual FIRE judges is the static out-of-balance force and moment
Mat64 make_mat64(const rfd_system_params &p) {
  Mat64 m{};
  const double G = 0.5 * p.E / (1.0 + p.nu);
  m.E = p.E;
  m.G = G;
  m.Estar = 0.5 * p.E / (1.0 - p.nu * p.nu);
  m.Gstar = 0.5 * G / (2.0 - p.nu);
  const double le = log(p.e);
  m.pA = le / sqrt(M_PI * M_PI + le * le);
  m.mu = p.mu;
  m.mur = p.mur;
  m.rho43pi = p.rho * 4.0 / 3.0 * M_PI;
  m.dt = p.dt;
  m.farcut = p.farcut;
  m.refmode = p.refmode;
  return m;
}

// FP64 force pass; dtp, trial, nobreak and done as in forces() (FIRE passes)
Sys64 make_sys64(rfd_system *s, const double *dtp, bool trial, const int *done) {
  Sys64 d{};
  d.N = s->N;
  d.Kc = s->p.contact_capacity;
  d.P = s->P;
  d.V = s->V64;
  d.W = s->W64;
  d.F = s->F64;
  d.M = s->M64;
  d.hoff = s->hoff;
  d.hr = s->hr;
  d.brc = s->brc;
  d.b0 = s->b64[0];
  d.b1 = s->b64[1];
  d.curp = s->d_cur;
  d.fl8 = s->fl8;
  d.cj = s->cj[s->ccur];
  d.ccnt = s->ccnt[s->ccur];
  d.cO = s->cO64[s->ccur];
  d.dtp = dtp;
  d.done = done;
  d.trial = trial ? 1 : 0;
  d.nbreak = s->d_nbreak;
  return d;
}

int forces64(rfd_system *s, const double *dtp = nullptr, bool trial = false,
             bool nobreak = false, const int *done = nullptr, bool flip = true) {
  Sys64 d = make_sys64(s, dtp, trial, done);
  Mat64 m = make_mat64(s->p);
  if (dtp) {
    m.pA = 0.0;
    m.mur = 0.0;
  }
  m.nobreak = nobreak ? 1 : 0;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (s->timing && s->ev_used + 2 <= s->ev.size()) {
    e0 = s->ev[s->ev_used];
    e1 = s->ev[s->ev_used + 1];
    s->ev_used += 2;
    RFD_HIP(s->ctx, hipEventRecord(e0, s->ctx->stream));
  }
  k_f64_forces<<<rfd_blocks((size_t)s->N * kG64, 256), 256, 0, s->ctx->stream>>>(d, m);
  RFD_HIP(s->ctx, hipGetLastError());
  if (e1) RFD_HIP(s->ctx, hipEventRecord(e1, s->ctx->stream));
  if (!trial && flip) k_flip_state<<<1, 64, 0, s->ctx->stream>>>(s->d_cur, done);
  RFD_HIP(s->ctx, hipGetLastError());
  return RFD_OK;
}

// one launch over ceil(N / FP) tiles, spread over the XCD runs
// (the single-evaluation pass in the physical mode when every tile's halves
// fit its LDS, else the chunked pass)
template <int FP, int FT>
void launch_forces(rfd_system *s, const SysDev &d, const Mat &m, const AdvDev *adv) {
  const uint32_t nb = (s->N + FP - 1) / FP;
  const uint32_t per = (nb + kXcd - 1) / kXcd;
  if (!m.refmode && !s->once_off && s->work_fp == FP) {
    WorkDev w{s->wrec, s->wofs, s->wnin};
    const int q = FP == 32 ? 0 : (FP == 64 ? 1 : 2);
    // slots for the system's largest tile, at least the per-particle sums
    const uint32_t cap = std::max((s->tile_halves[q] + 31u) & ~31u, 2u * FP);
    const size_t lds = sizeof(float) * 6 * cap;
    if (adv && adv->done)
      k_nat_forces1_small<FP, FT><<<per * kXcd, FT, lds, s->ctx->stream>>>(d, m, w, nb, per,
                                                                           cap, *adv);
    else if (adv)
      k_nat_forces1<FP, FT, true><<<per * kXcd, FT, lds, s->ctx->stream>>>(d, m, w, nb, per,
                                                                            cap, *adv);
    else
      k_nat_forces1<FP, FT, false><<<per * kXcd, FT, lds, s->ctx->stream>>>(d, m, w, nb, per,
                                                                             cap, AdvDev{});
  } else {
    k_nat_forces<FP, FT><<<per * kXcd, FT, 0, s->ctx->stream>>>(d, m, nb, per);
  }
}

// work records for tiles of FP particles (none when a tile's halves exceed the
// single-evaluation pass's LDS); allocates and synchronises, so it runs at
// set-up, never inside a captured step
template <int FP>
int build_work_fp(rfd_system *s) {
  rfd_ctx *ctx = s->ctx;
  hipStream_t q = ctx->stream;
  const uint32_t nt = (s->N + FP - 1) / FP;
  int st = dalloc(ctx, &s->wcnt, nt);
  if (!st) st = dalloc(ctx, &s->wnin, nt);
  if (!st) st = dalloc(ctx, &s->wofs, (size_t)nt + 1);
  if (!st) st = dalloc(ctx, &s->wcur, nt);
  if (st) return st;
  k_work_count<FP><<<nt, 256, 0, q>>>(s->hoff, s->hr, s->N, nt, s->wcnt, s->wnin);
  RFD_HIP(ctx, hipGetLastError());
  st = scan_u32(s, s->wcnt, nt, s->wofs, s->wcur);
  if (st) return st;
  uint32_t total = 0;
  RFD_HIP(ctx, hipMemcpyAsync(&total, s->wofs + nt, sizeof(uint32_t), hipMemcpyDeviceToHost, q));
  RFD_HIP(ctx, hipStreamSynchronize(q));
  st = dalloc(ctx, &s->wrec, total);
  if (st) return st;
  k_work_fill<FP><<<nt, 256, 0, q>>>(s->hoff, s->hr, s->hm, s->N, nt, s->wofs, s->wnin,
                                      s->wrec);
  RFD_HIP(ctx, hipGetLastError());
  s->work_fp = FP;
  return RFD_OK;
}

int ensure_work(rfd_system *s) {
  const int fp = s->force_fp;
  if (!s->have_bonds || s->p.refmode || s->fp64 || s->work_fp == fp) return RFD_OK;
  const int q = fp == 32 ? 0 : (fp == 64 ? 1 : 2);
  s->work_fp = 0;
  s->gen++;
  const uint32_t cap = fp == 32 ? OnceCap<32>::v : (fp == 64 ? OnceCap<64>::v : OnceCap<128>::v);
  if (s->tile_halves[q] > cap) return RFD_OK;  // the chunked pass serves this shape
  if (fp == 32) return build_work_fp<32>(s);
  if (fp == 128) return build_work_fp<128>(s);
  return build_work_fp<64>(s);
}

// flip = false leaves the flip of the bond-state index to the kernel that
// follows (the loading control of FP32 explicit steps, the FIRE control)
int forces(rfd_system *s, const float *dtp = nullptr, bool trial = false,
           bool nobreak = false, const AdvDev *adv = nullptr, bool flip = true) {
  if (s->fp64)
    return forces64(s, dtp ? s->fire_dt64 : nullptr, trial, nobreak,
                    dtp ? (const int *)((const char *)s->fire + offsetof(FireDev, done))
                        : nullptr,
                    flip);
  SysDev d = make_dev(s);
  d.dtp = dtp;
  d.trial = trial ? 1 : 0;
  d.nbreak = s->d_nbreak;
  // after convergence the remaining launches of a FIRE chunk must not apply
  // the last move to the springs again
  if (dtp) d.done = (const int *)((const char *)s->fire + offsetof(FireDev, done));
  Mat m = make_mat(s->p);
  if (dtp) {
    m.pA = 0.f;
    m.mur = 0.f;
  }
  m.nobreak = nobreak ? 1 : 0;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (s->timing && s->ev_used + 2 <= s->ev.size()) {
    e0 = s->ev[s->ev_used];
    e1 = s->ev[s->ev_used + 1];
    s->ev_used += 2;
    RFD_HIP(s->ctx, hipEventRecord(e0, s->ctx->stream));
  }
  if (s->force_fp == 32)
    launch_forces<32, 128>(s, d, m, adv);
  else if (s->force_fp == 128)
    launch_forces<128, 256>(s, d, m, adv);
  else
    launch_forces<kFP, kFT>(s, d, m, adv);
  RFD_HIP(s->ctx, hipGetLastError());
  // a trial pass leaves the state index alone: its output buffer is dropped
  if (!trial && flip) k_flip_state<<<1, 64, 0, s->ctx->stream>>>(s->d_cur, d.done);
  RFD_HIP(s->ctx, hipGetLastError());
  if (e1) RFD_HIP(s->ctx, hipEventRecord(e1, s->ctx->stream));
  return RFD_OK;
}

// loading control of an FP32 step, for the one specimen or every batched
// specimen (and the flip of the bond-state index the step's force pass left)
int control(rfd_system *s) {
  if (s->S)
    k_nat_control_fx<<<(s->S + 63) / 64, 64, 0, s->ctx->stream>>>(
        s->ctls, s->acc_b, s->S, s->hist_b, s->hist_cap_b, s->p.dt, s->d_cur);
  else
    k_nat_control_fx<<<1, 64, 0, s->ctx->stream>>>(s->ctl, s->acc, 1u, s->hist,
                                                   (uint32_t)kHistCap, s->p.dt, s->d_cur);
  RFD_HIP(s->ctx, hipGetLastError());
  return RFD_OK;
}

int pass_and_control(rfd_system *s) {
  hipStream_t st64 = s->ctx->stream;
  if (s->fp64) {
    Pass64 q{};
    q.N = s->N;
    q.P = s->P;
    q.V = s->V64;
    q.W = s->W64;
    q.F = s->F64;
    q.M = s->M64;
    q.dt = s->p.dt;
    q.rho43pi = s->p.rho * 4.0 / 3.0 * M_PI;
    q.ctl = s->ctl;
    q.part = s->part;
    if (s->ctl_in_pass && s->nblocks_pass <= kCtlInPassTiles) {
      // small system: the control in the pass (one launch less per step)
      if (s->mode == 1)
        k_f64_pass_ctl<1><<<s->nblocks_pass, 256, 0, st64>>>(q, s->done_cnt, s->hist, s->d_cur);
      else if (s->mode == 2)
        k_f64_pass_ctl<2><<<s->nblocks_pass, 256, 0, st64>>>(q, s->done_cnt, s->hist, s->d_cur);
      else
        k_f64_pass_ctl<0><<<s->nblocks_pass, 256, 0, st64>>>(q, s->done_cnt, s->hist, s->d_cur);
      RFD_HIP(s->ctx, hipGetLastError());
      return RFD_OK;
    }
    if (s->mode == 1)
      k_f64_pass<1><<<s->nblocks_pass, 256, 0, st64>>>(q);
    else if (s->mode == 2)
      k_f64_pass<2><<<s->nblocks_pass, 256, 0, st64>>>(q);
    else
      k_f64_pass<0><<<s->nblocks_pass, 256, 0, st64>>>(q);
    k_nat_control<<<1, 256, 0, st64>>>(s->ctl, s->part, (int)s->nblocks_pass, s->hist, s->p.dt,
                                       s->d_cur);
    RFD_HIP(s->ctx, hipGetLastError());
    return RFD_OK;
  }
  PassDev p{};
  p.N = s->N;
  p.P = s->P;
  p.V = s->V;
  p.W = s->W;
  p.F = s->F;
  p.M = s->M;
  p.inv = s->inv;
  p.dt = (float)s->p.dt;
  p.ctl = s->ctl;
  p.acc = s->acc;
  hipStream_t st = s->ctx->stream;
  if (s->S) {
    // batched specimens: one launch over every specimen's tiles, then one
    // control thread per specimen
    p.acc = s->acc_b;
    p.ctl = nullptr;
    if (s->mode == 1)
      k_nat_pass_batch<1><<<s->nblk_b, 256, 0, st>>>(p, s->blk_spec, s->blk_first, s->blk_end,
                                                     s->ctls);
    else
      k_nat_pass_batch<2><<<s->nblk_b, 256, 0, st>>>(p, s->blk_spec, s->blk_first, s->blk_end,
                                                     s->ctls);
    return control(s);
  }
  if (s->mode == 1)
    k_nat_pass<1><<<s->nblocks_pass, 256, 0, st>>>(p);
  else if (s->mode == 2)
    k_nat_pass<2><<<s->nblocks_pass, 256, 0, st>>>(p);
  else
    k_nat_pass<0><<<s->nblocks_pass, 256, 0, st>>>(p);
  return control(s);
}

bool valid(rfd_system *s) { return s && s->ctx; }

// explicit steps fuse the particle update into the force pass for one
// specimen in the FP32 physical mode, when the single-evaluation pass serves
// the current tile shape
bool can_fuse(const rfd_system *s) {
  return s->fuse && !s->fp64 && (!s->S || s->batch_fusable) && !s->p.refmode &&
         !s->once_off && s->P2 && s->work_fp == s->force_fp;
}

int step_plain(rfd_system *s) {
  if (!s->rows_valid || s->step_count % s->p.rebuild_every == 0) {
    int st = rebuild_rows(s);
    if (st) return st;
  }
  // the loading control flips the bond-state index
  int st = forces(s, nullptr, false, false, nullptr, false);
  if (!st) st = pass_and_control(s);
  if (st) return st;
  s->step_count++;
  return RFD_OK;
}

// one fused step: reads P, V, W, writes P2, V2, W2 and swaps the two sets
// (also after an error, so that a pair of calls always restores them)
// the loading control runs in the force pass when the pass has few enough
// tiles that every tile's workgroup is resident at once (one round of
// workgroups): then the wait of each workgroup for its done count does not
// hold back other tiles, and the control's launch is saved
bool control_in_pass(const rfd_system *s) {
  if (!s->ctl_in_pass) return false;
  const uint32_t fp = s->force_fp == 32 ? 32u : (s->force_fp == 128 ? 128u : (uint32_t)kFP);
  return (s->N + fp - 1) / fp <= kCtlInPassTiles;
}

int step_fused(rfd_system *s) {
  int st = RFD_OK;
  bool ctl_in_pass = false;
  if (!s->rows_valid || s->step_count % s->p.rebuild_every == 0) st = rebuild_rows(s);
  if (!st) {
    AdvDev a{s->P2, s->V2, s->W2, s->inv, s->ctl, s->acc, (float)s->p.dt, s->mode,
             nullptr, nullptr};
    a.ctl_w = s->ctl;
    a.nspec = 1;
    a.hist = s->hist;
    a.hist_cap = (uint32_t)kHistCap;
    if (s->S) {
      a.ctl = a.ctl_w = s->ctls;
      a.acc = s->acc_b;
      a.spec32 = s->spec32;
      a.spec_first = s->spec_first_d;
      a.nspec = s->S;
      a.hist = s->hist_b;
      a.hist_cap = s->hist_cap_b;
    }
    a.dt64 = s->p.dt;
    a.curp = s->d_cur;
    ctl_in_pass = control_in_pass(s);
    if (ctl_in_pass) a.done = s->done_cnt;
    st = forces(s, nullptr, false, false, &a, false);
  }
  if (!st && !ctl_in_pass) st = control(s);
  std::swap(s->P, s->P2);
  std::swap(s->V, s->V2);
  std::swap(s->W, s->W2);
  if (!st) s->step_count++;
  return st;
}

// ------------------------------------------------- multi-step launches
// the launch shape of k_nat_steps for the current tile shape
struct StepsShape {
  uint32_t nb, per, cap;
  size_t lds;
};

template <int FP>
StepsShape steps_shape(const rfd_system *s) {
  StepsShape h;
  h.nb = (s->N + FP - 1) / FP;
  h.per = (h.nb + kXcd - 1) / kXcd;
  const int q = FP == 32 ? 0 : (FP == 64 ? 1 : 2);
  h.cap = std::max((s->tile_halves[q] + 31u) & ~31u, 2u * FP);
  h.lds = sizeof(float) * 6 * h.cap;
  return h;
}

// every workgroup of a cooperative launch must be resident at once
template <int FP, int FT>
bool steps_fit_fp(const rfd_system *s) {
  const StepsShape h = steps_shape<FP>(s);
  int bpc = 0, dev = 0, cus = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&bpc, (const void *)k_nat_steps<FP, FT>, FT,
                                                   h.lds) != hipSuccess ||
      hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
    return false;
  return (uint64_t)bpc * (uint64_t)cus >= (uint64_t)h.per * kXcd;
}

bool steps_fit(const rfd_system *s) {
  if (!can_fuse(s)) return false;
  if (s->force_fp == 32) return steps_fit_fp<32, 128>(s);
  if (s->force_fp == 128) return steps_fit_fp<128, 256>(s);
  return steps_fit_fp<kFP, kFT>(s);
}

template <int FP, int FT>
int launch_steps(rfd_system *s, SysDev d, Mat m, AdvDev a, StepsDev q) {
  const StepsShape h = steps_shape<FP>(s);
  WorkDev w{s->wrec, s->wofs, s->wnin};
  uint32_t nb = h.nb, per = h.per, cap = h.cap;
  void *args[] = {&d, &m, &w, &nb, &per, &cap, &a, &q};
  RFD_HIP(s->ctx, hipLaunchCooperativeKernel((const void *)k_nat_steps<FP, FT>,
                                             dim3(h.per * kXcd), dim3(FT), args,
                                             (unsigned)h.lds, s->ctx->stream));
  return RFD_OK;
}

// n (even) fused steps in one launch; no contact-row rebuild falls inside
int steps_persistent(rfd_system *s, uint32_t n) {
  const uint32_t Ssp = s->S ? s->S : 1u;
  if (s->acc3_S != Ssp) {
    int st = dalloc(s->ctx, &s->acc3, 6 * (size_t)Ssp);
    if (st) return st;
    RFD_HIP(s->ctx, hipMemsetAsync(s->acc3, 0, sizeof(unsigned long long) * 6 * Ssp,
                                   s->ctx->stream));
    s->acc3_S = Ssp;
  }

  SysDev d = make_dev(s);
  Mat m = make_mat(s->p);
  AdvDev a{s->P2, s->V2, s->W2, s->inv, s->ctl, s->acc, (float)s->p.dt, s->mode,
           nullptr, nullptr, 0u};
  StepsDev q{};
  q.P[0] = s->P;
  q.P[1] = s->P2;
  q.V[0] = s->V;
  q.V[1] = s->V2;
  q.W[0] = s->W;
  q.W[1] = s->W2;
  q.ctl = s->ctl;
  q.acc3 = s->acc3;
  q.hist = s->hist;
  q.hist_cap = (uint32_t)kHistCap;
  q.S = Ssp;
  q.nsteps = n;
  q.curp = s->d_cur;
  q.dt = s->p.dt;

  if (s->S) {
    a.spec32 = s->spec32;
    a.spec_first = s->spec_first_d;
    q.ctl = s->ctls;
    q.hist = s->hist_b;
    q.hist_cap = s->hist_cap_b;
  }
  int st;
  if (s->force_fp == 32)
    st = launch_steps<32, 128>(s, d, m, a, q);
  else if (s->force_fp == 128)
    st = launch_steps<128, 256>(s, d, m, a, q);
  else
    st = launch_steps<kFP, kFT>(s, d, m, a, q);
  if (!st) s->step_count += n;
  return st;
}

bool capturing(const rfd_system *s) {
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  return hipStreamIsCapturing(s->ctx->stream, &cs) != hipSuccess ||
         cs != hipStreamCaptureStatusNone;
}

// n explicit steps; fused ones run in pairs, so that P, V, W hold the state
// when this returns (the particle update is the same arithmetic either way,
// and the two orders of steps give the same bits).  With multi-step launches
// on, the steps between two contact-row rebuilds run as one launch (not
// under graph capture).
// FP64 explicit steps in one launch each (k_f64_step): small systems whose
// force grid fits the one-round limit, one specimen
bool can_fuse64(const rfd_system *s) {
  return s->fuse && s->fp64 && !s->S && s->ctl_in_pass && s->P2 && s->V64b && s->W64b &&
         rfd_blocks((size_t)s->N * kG64, 256) <= kCtlInPassTiles;
}

int step_fused64(rfd_system *s) {
  if (!s->rows_valid || s->step_count % s->p.rebuild_every == 0) {
    const int st = rebuild_rows(s);
    if (st) return st;
  }
  const Sys64 d = make_sys64(s, nullptr, false, nullptr);
  const Mat64 m = make_mat64(s->p);
  Adv64 a{};
  a.P = s->P2;
  a.V = s->V64b;
  a.W = s->W64b;
  a.ctl = s->ctl;
  a.part = s->part_step;
  a.done = s->done_cnt;
  a.hist = s->hist;
  a.curp = s->d_cur;
  a.dt = s->p.dt;
  a.rho43pi = s->p.rho * 4.0 / 3.0 * M_PI;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (s->timing && s->ev_used + 2 <= s->ev.size()) {
    e0 = s->ev[s->ev_used];
    e1 = s->ev[s->ev_used + 1];
    s->ev_used += 2;
    RFD_HIP(s->ctx, hipEventRecord(e0, s->ctx->stream));
  }
  const unsigned grid = rfd_blocks((size_t)s->N * kG64, 256);
  if (s->mode == 1)
    k_f64_step<1><<<grid, 256, 0, s->ctx->stream>>>(d, m, a);
  else if (s->mode == 2)
    k_f64_step<2><<<grid, 256, 0, s->ctx->stream>>>(d, m, a);
  else
    k_f64_step<0><<<grid, 256, 0, s->ctx->stream>>>(d, m, a);
  RFD_HIP(s->ctx, hipGetLastError());
  if (e1) RFD_HIP(s->ctx, hipEventRecord(e1, s->ctx->stream));
  std::swap(s->P, s->P2);
  std::swap(s->V64, s->V64b);
  std::swap(s->W64, s->W64b);
  s->step_count++;
  return RFD_OK;
}

int run_steps(rfd_system *s, uint32_t n) {
  uint32_t k = 0;
  if (can_fuse64(s)) {
    // in pairs, so that P, V, W hold the state between calls
    for (; k + 2 <= n; k += 2) {
      const int st = step_fused64(s);
      const int st2 = st ? st : step_fused64(s);
      if (st2) return st2;
    }
  }
  if (can_fuse(s)) {
    bool multi = s->persist && n >= 2 && !capturing(s) && steps_fit(s);
    while (k + 2 <= n) {
      uint32_t seg = 0;
      if (multi) {
        if (!s->rows_valid || s->step_count % s->p.rebuild_every == 0) {
          const int st = rebuild_rows(s);
          if (st) return st;
        }
        const uint32_t left =
            s->p.rebuild_every - (uint32_t)(s->step_count % s->p.rebuild_every);
        seg = std::min(n - k, left) & ~1u;
      }
      if (seg >= 2) {
        const int st = steps_persistent(s, seg);
        if (st) return st;
        k += seg;
        continue;
      }
      const int st = step_fused(s);
      const int st2 = step_fused(s);
      if (st) return st;
      if (st2) return st2;
      k += 2;
    }
  }
  for (; k < n; ++k) {
    const int st = step_plain(s);
    if (st) return st;
  }
  return RFD_OK;
}

int fire_alloc(rfd_system *s) {
  if (s->fire) return RFD_OK;
  rfd_ctx *ctx = s->ctx;
  int st = dalloc(ctx, &s->fire, 1);
  if (!st) st = dalloc(ctx, &s->fire_part, (size_t)kFireVals * s->nblocks_pass);
  if (!st) st = dalloc(ctx, &s->P0, s->N);
  if (!st) st = dalloc(ctx, &s->Vf, s->N);
  if (!st) st = dalloc(ctx, &s->d_nbreak, 1);
  s->gen++;  // the force pass now counts breaks
  if (!st) st = dalloc(ctx, &s->Th, s->N);
  if (!st && s->fp64) {
    st = dalloc(ctx, &s->Vf64, s->N);
    if (!st) st = dalloc(ctx, &s->Th64, s->N);
  }
  for (int q = 0; q < 2 && !st; ++q) {
    st = dalloc(ctx, &s->du[q], 3 * (size_t)s->N);
    if (!st) st = dalloc(ctx, &s->dth[q], 3 * (size_t)s->N);
  }
  if (st) return st;
  for (int q = 0; q < 2; ++q) {
    RFD_HIP(ctx, hipMemset(s->du[q], 0, sizeof(float) * 3 * s->N));
    RFD_HIP(ctx, hipMemset(s->dth[q], 0, sizeof(float) * 3 * s->N));
  }
  s->fire_dt = (float *)((char *)s->fire + offsetof(FireDev, dt_f));
  s->fire_dt64 = (double *)((char *)s->fire + offsetof(FireDev, dt_m));
  return RFD_OK;
}


FireParticles fire_particles(rfd_system *s) {
  FireParticles p{};
  p.N = s->N;
  p.P = s->P;
  p.V = s->V;
  p.Vf = s->Vf;
  p.W = s->W;
  p.F = s->F;
  p.Th = s->Th;
  p.P0 = s->P0;
  p.M = s->M;
  p.inv = s->inv;
  p.dt_phys = (float)s->p.dt;
  p.ctl = s->ctl;
  p.fire = s->fire;
  p.part = s->fire_part;
  return p;
}

FireParticles64 fire_particles64(rfd_system *s) {
  FireParticles64 p{};
  p.N = s->N;
  p.P = s->P;
  p.V = s->V64;
  p.Vf = s->Vf64;
  p.W = s->W64;
  p.F = s->F64;
  p.Th = s->Th64;
  p.P0 = s->P0;
  p.M = s->M64;
  p.rho43pi = s->p.rho * 4.0 / 3.0 * M_PI;
  p.ctl = s->ctl;
  p.fire = s->fire;
  p.part = s->fire_part;
  return p;
}

// the FIRE measure of the current forces (platens at rest, power, norms,
// reaction sums) into the per-workgroup partials
void fire_measure(rfd_system *s, hipStream_t q) {
  if (s->fp64) {
    const FireParticles64 fp = fire_particles64(s);
    if (s->mode == 1)
      k_fire64_measure<1><<<s->nblocks_pass, 256, 0, q>>>(fp);
    else if (s->mode == 2)
      k_fire64_measure<2><<<s->nblocks_pass, 256, 0, q>>>(fp);
    else
      k_fire64_measure<0><<<s->nblocks_pass, 256, 0, q>>>(fp);
    return;
  }
  const FireParticles fp = fire_particles(s);
  if (s->mode == 1)
    k_fire_measure<1><<<s->nblocks_pass, 256, 0, q>>>(fp);
  else if (s->mode == 2)
    k_fire_measure<2><<<s->nblocks_pass, 256, 0, q>>>(fp);
  else
    k_fire_measure<0><<<s->nblocks_pass, 256, 0, q>>>(fp);
}

// velocities, spins and rotations of the load step back to rest (FIRE velocity too)
void fire_rest(rfd_system *s, hipStream_t q, bool fire_velocity) {
  if (s->fp64) {
    k_zero_d4<<<rfd_blocks(s->N, 256), 256, 0, q>>>(s->V64, s->W64, s->N);
    if (fire_velocity) k_zero_d4<<<rfd_blocks(s->N, 256), 256, 0, q>>>(s->Vf64, nullptr, s->N);
  } else {
    k_zero_f4<<<rfd_blocks(s->N, 256), 256, 0, q>>>(s->V, s->W, s->N);
    if (fire_velocity) k_zero_f4<<<rfd_blocks(s->N, 256), 256, 0, q>>>(s->Vf, nullptr, s->N);
  }
}

// the FIRE measure runs in the epilogue of the quasi-static force pass (FP32
// physical mode, one specimen, single-evaluation pass for the current tiles)
// unless fusion is switched off
bool can_measure_fused(const rfd_system *s) {
  return s->fuse && !s->fp64 && !s->S && !s->p.refmode && !s->once_off &&
         s->work_fp == s->force_fp;
}

template <int FP, int FT>
void launch_forces_fire(rfd_system *s, const SysDev &d, const Mat &m, const MeasDev &fm) {
  const uint32_t nb = (s->N + FP - 1) / FP;
  const uint32_t per = (nb + kXcd - 1) / kXcd;
  WorkDev w{s->wrec, s->wofs, s->wnin};
  const int q = FP == 32 ? 0 : (FP == 64 ? 1 : 2);
  const uint32_t cap = std::max((s->tile_halves[q] + 31u) & ~31u, 2u * FP);
  k_nat_forces1_fire<FP, FT><<<per * kXcd, FT, sizeof(float) * 6 * cap, s->ctx->stream>>>(
      d, m, w, nb, per, cap, fm);
}

// first-stage workgroups of the fused measure's reduction (the partial pairs
// k_fire_control then reduces)
uint32_t fire_reduce_blocks(const rfd_system *s) { return std::min(s->nblocks_pass, 64u); }

// quasi-static force pass with the measure fused in, then the two-stage
// reduction into fire_part (the FIRE control follows)
int forces_fire_measured(rfd_system *s) {
  const uint32_t wpt = s->force_fp == 128 ? 2u : 1u;
  const uint32_t nparts = (s->N + s->force_fp - 1) / s->force_fp * wpt;
  if (s->fire_tpart_n < nparts) {
    int st = dalloc(s->ctx, &s->fire_tpart, (size_t)kFireVals * nparts);
    if (st) return st;
    s->fire_tpart_n = nparts;
  }
  SysDev d = make_dev(s);
  d.dtp = s->fire_dt;
  d.nbreak = s->d_nbreak;
  d.done = (const int *)((const char *)s->fire + offsetof(FireDev, done));
  Mat m = make_mat(s->p);
  m.pA = 0.f;
  m.mur = 0.f;
  m.nobreak = 1;
  const MeasDev fm{s->ctl, s->Vf, s->fire_tpart, s->mode};
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (s->timing && s->ev_used + 2 <= s->ev.size()) {
    e0 = s->ev[s->ev_used];
    e1 = s->ev[s->ev_used + 1];
    s->ev_used += 2;
    RFD_HIP(s->ctx, hipEventRecord(e0, s->ctx->stream));
  }
  if (s->force_fp == 32)
    launch_forces_fire<32, 128>(s, d, m, fm);
  else if (s->force_fp == 128)
    launch_forces_fire<128, 256>(s, d, m, fm);
  else
    launch_forces_fire<kFP, kFT>(s, d, m, fm);
  RFD_HIP(s->ctx, hipGetLastError());
  if (e1) RFD_HIP(s->ctx, hipEventRecord(e1, s->ctx->stream));
  k_fire_reduce<<<fire_reduce_blocks(s), 256, 0, s->ctx->stream>>>(s->fire, s->fire_tpart,
                                                                   nparts, s->fire_part);
  RFD_HIP(s->ctx, hipGetLastError());
  return RFD_OK;
}

// one FIRE iteration: force pass at the current configuration, constraint +
// reductions, control, update
int fire_iteration(rfd_system *s) {
  int *flip = nullptr;
  uint32_t nparts = s->nblocks_pass;  // partial pairs for the control
  hipStream_t q = s->ctx->stream;
  if (s->forces_current) {
    s->forces_current = false;  // the load step's first pass already ran
    fire_measure(s, q);
  } else {
    if (!s->rows_valid || s->step_count % s->p.rebuild_every == 0) {
      int st = rebuild_rows(s);
      if (st) return st;
    }
    // the FIRE control flips the bond-state index after this pass
    if (can_measure_fused(s)) {
      int st = forces_fire_measured(s);
      if (st) return st;
      nparts = fire_reduce_blocks(s);
    } else {
      int st = forces(s, s->fire_dt, false, true, nullptr, false);
      if (st) return st;
      fire_measure(s, q);
    }
    flip = s->d_cur;
  }
  k_fire_control<<<1, 256, 0, q>>>(s->fire, s->fire_part, (int)nparts, s->ctl, flip);
  if (s->fp64)
    k_fire64_update<<<s->nblocks_pass, 256, 0, q>>>(fire_particles64(s));
  else
    k_fire_update<<<s->nblocks_pass, 256, 0, q>>>(fire_particles(s));
  RFD_HIP(s->ctx, hipGetLastError());
  s->step_count++;
  return RFD_OK;
}

// start of a load step from the positions saved in P0: the move (warm start
// dU, dTh and the platens) enters the springs through one force pass run with
// dt0 (FireDev f written fresh), FIRE then starts at rest.  A trial pass
// reports the residual it leaves and leaves the relaxation state as it was.
int begin_pass(rfd_system *s, const FireDev &f, const float *dU, const float *dTh, double dtop,
               bool trial, double *fmax_out) {
  rfd_ctx *ctx = s->ctx;
  hipStream_t q = ctx->stream;
  RFD_HIP(ctx, hipMemcpyAsync(s->fire, &f, sizeof(f), hipMemcpyHostToDevice, q));
  RFD_HIP(ctx, hipMemcpyAsync(s->P, s->P0, sizeof(P4) * s->N, hipMemcpyDeviceToDevice, q));
  if (s->fp64) {
    const FireParticles64 fpp = fire_particles64(s);
    if (s->mode == 1)
      k_fire64_begin<1><<<s->nblocks_pass, 256, 0, q>>>(fpp, dU, dTh, f.dt, dtop);
    else if (s->mode == 2)
      k_fire64_begin<2><<<s->nblocks_pass, 256, 0, q>>>(fpp, dU, dTh, f.dt, dtop);
    else
      k_fire64_begin<0><<<s->nblocks_pass, 256, 0, q>>>(fpp, dU, dTh, f.dt, dtop);
  } else {
    const FireParticles fpp = fire_particles(s);
    const float dt0 = (float)f.dt;
    if (s->mode == 1)
      k_fire_begin<1><<<s->nblocks_pass, 256, 0, q>>>(fpp, dU, dTh, dt0, dtop);
    else if (s->mode == 2)
      k_fire_begin<2><<<s->nblocks_pass, 256, 0, q>>>(fpp, dU, dTh, dt0, dtop);
    else
      k_fire_begin<0><<<s->nblocks_pass, 256, 0, q>>>(fpp, dU, dTh, dt0, dtop);
  }
  RFD_HIP(ctx, hipGetLastError());
  // a trial pass writes its state to the spare buffer and leaves the index
  int st = forces(s, s->fire_dt, trial, true);
  if (st) return st;
  fire_rest(s, q, false);
  RFD_HIP(ctx, hipGetLastError());
  if (trial) {
    FireDev fm = f;
    fm.measure_only = 1;
    RFD_HIP(ctx, hipMemcpyAsync(s->fire, &fm, sizeof(fm), hipMemcpyHostToDevice, q));
    fire_measure(s, q);
    k_fire_control<<<1, 256, 0, q>>>(s->fire, s->fire_part, (int)s->nblocks_pass, s->ctl,
                                    nullptr);
    RFD_HIP(ctx, hipGetLastError());
    FireDev r;
    RFD_HIP(ctx, hipMemcpyAsync(&r, s->fire, sizeof(r), hipMemcpyDeviceToHost, q));
    RFD_HIP(ctx, hipStreamSynchronize(q));
    *fmax_out = r.fmax;
  }
  return RFD_OK;
}

int upload_vec3(rfd_ctx *ctx, float4 *dst, const float *soa, uint32_t N) {
  std::vector<float4> h(N);
  for (uint32_t i = 0; i < N; ++i)
    h[i] = make_float4(soa[i], soa[i + (size_t)N], soa[i + 2 * (size_t)N], 0.f);
  RFD_HIP(ctx, hipMemcpy(dst, h.data(), sizeof(float4) * N, hipMemcpyHostToDevice));
  return RFD_OK;
}

int download_vec3(rfd_ctx *ctx, float *soa, const float4 *src, uint32_t N) {
  std::vector<float4> h(N);
  RFD_HIP(ctx, hipMemcpy(h.data(), src, sizeof(float4) * N, hipMemcpyDeviceToHost));
  for (uint32_t i = 0; i < N; ++i) {
    soa[i] = h[i].x;
    soa[i + (size_t)N] = h[i].y;
    soa[i + 2 * (size_t)N] = h[i].z;
  }
  return RFD_OK;
}

// launch shapes chosen by the autotuner, per (arch, log2 N bucket, batched)
struct TuneChoice {
  int first, second, third;  // force tile, row-search block, single evaluation
};
std::map<std::string, TuneChoice> &tune_table() {
  static std::map<std::string, TuneChoice> t;
  return t;
}
std::mutex &tune_mu() {
  static std::mutex m;
  return m;
}
std::string tune_key(const char *arch, int bucket, bool batched) {
  std::string a(arch);
  const size_t colon = a.find(':');  // drop feature suffixes (":sramecc+:xnack-")
  if (colon != std::string::npos) a.resize(colon);
  return a + "/n" + std::to_string(bucket) + (batched ? "/batch" : "/single");
}

}  // namespace

extern "C" {

int rfd_system_create(rfd_ctx *ctx, const rfd_system_params *p, uint32_t N, rfd_system **out) {
  if (!ctx) return RFD_E_ARG;
  RFD_CHECK_ARG(ctx, p && out && N > 0, "rfd_system_create: bad arguments");
  RFD_CHECK_ARG(ctx, p->dt > 0 && p->E > 0 && p->rho > 0 && p->e > 0 && p->e < 1,
                "rfd_system_create: bad material parameters");
  RFD_CHECK_ARG(ctx, p->contact_capacity > 0 && p->contact_capacity <= 64,
                "rfd_system_create: contact_capacity must be 1..64");
  *out = nullptr;
  rfd_system *s = new rfd_system();
  s->ctx = ctx;
  s->p = *p;
  if (s->p.rebuild_every == 0) s->p.rebuild_every = 20;
  if (s->p.farcut <= 0) s->p.farcut = 1.45;
  s->N = N;
  int st = 0;
  const size_t KN = (size_t)p->contact_capacity * N;
  if (!st) st = dalloc(ctx, &s->P, N);
  if (!st) st = dalloc(ctx, &s->V, N);
  if (!st) st = dalloc(ctx, &s->W, N);
  if (!st) st = dalloc(ctx, &s->F, N);
  if (!st) st = dalloc(ctx, &s->M, N);
  if (!st) st = dalloc(ctx, &s->inv, N);
  if (!st) st = dalloc(ctx, &s->hoff, (size_t)N + 1);
  for (int q = 0; q < 2 && !st; ++q) {
    st = dalloc(ctx, &s->cj[q], KN);
    if (!st) st = dalloc(ctx, &s->ccnt[q], N);
    if (!st) st = dalloc(ctx, &s->cO[q], KN);
  }
  if (!st) st = dalloc(ctx, &s->CI, N);
  if (!st) st = dalloc(ctx, &s->ctmp, N);
  if (!st) st = dalloc(ctx, &s->csorted, N);
  if (!st) st = dalloc(ctx, &s->cpos, N);
  if (!st) st = dalloc(ctx, &s->ctl, 1);
  if (!st) st = dalloc(ctx, &s->d_cur, 1);
  s->fp64 = p->fp64 != 0;
  if (s->fp64) {
    if (!st) st = dalloc(ctx, &s->V64, N);
    if (!st) st = dalloc(ctx, &s->W64, N);
    if (!st) st = dalloc(ctx, &s->F64, N);
    if (!st) st = dalloc(ctx, &s->M64, N);
    for (int q = 0; q < 2 && !st; ++q) {
      st = dalloc(ctx, &s->cO64[q], KN);
      if (!st && hipMemset(s->cO64[q], 0, sizeof(double4) * KN) != hipSuccess)
        st = rfd_set_error(ctx, RFD_E_STATE, "rfd_system_create: memset");
    }
  }
  s->nblocks_pass = rfd_blocks(N, 256, 2048);
  if (!st) st = dalloc(ctx, &s->part, 2 * (size_t)s->nblocks_pass);
  if (!st) st = dalloc(ctx, &s->acc, 2);
  if (!st) st = dalloc(ctx, &s->done_cnt, 1);
  if (!st) st = dalloc(ctx, &s->hist, 3 * (size_t)kHistCap);
  if (!st && !s->fp64) {
    st = dalloc(ctx, &s->P2, N);
    if (!st) st = dalloc(ctx, &s->V2, N);
    if (!st) st = dalloc(ctx, &s->W2, N);
  }
  if (!st && s->fp64 && rfd_blocks((size_t)N * kG64, 256) <= kCtlInPassTiles) {
    st = dalloc(ctx, &s->P2, N);
    if (!st) st = dalloc(ctx, &s->V64b, N);
    if (!st) st = dalloc(ctx, &s->W64b, N);
    if (!st) st = dalloc(ctx, &s->part_step, 2 * (size_t)kCtlInPassTiles);
  }
  if (!st) {
    rfd_loading L{};
    hipError_t e = hipMemcpy(s->ctl, &L, sizeof(L), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemset(s->acc, 0, 2 * sizeof(unsigned long long));
    if (e == hipSuccess) e = hipMemset(s->done_cnt, 0, sizeof(uint32_t));
    for (int q = 0; q < 2 && e == hipSuccess; ++q)
      e = hipMemset(s->ccnt[q], 0, sizeof(uint32_t) * N);
    if (e != hipSuccess) st = rfd_hip_status(ctx, e, "rfd_system_create");
  }
  if (st) {
    rfd_system_destroy(s);
    return st;
  }
  *out = s;
  return RFD_OK;
}

int rfd_system_destroy(rfd_system *s) {
  if (!s) return RFD_OK;
  if (s->ctx) (void)hipStreamSynchronize(s->ctx->stream);
  dfree(s->P);
  dfree(s->V);
  dfree(s->W);
  dfree(s->F);
  dfree(s->M);
  dfree(s->inv);
  dfree(s->hoff);
  dfree(s->hr);
  dfree(s->hm);
  dfree(s->brc);
  dfree(s->Vf);
  dfree(s->d_nbreak);
  for (int q = 0; q < 2; ++q) {
    dfree(s->bst[q]);
    dfree(s->cj[q]);
    dfree(s->ccnt[q]);
    dfree(s->cO[q]);
  }
  dfree(s->CI);
  dfree(s->cstart);
  dfree(s->ccursor);
  dfree(s->ctmp);
  dfree(s->csorted);
  dfree(s->cpos);
  dfree(s->fl8);
  dfree(s->ccnt_cells);
  dfree(s->ctl);
  dfree(s->d_cur);
  dfree(s->V64);
  dfree(s->W64);
  dfree(s->F64);
  dfree(s->M64);
  for (int q = 0; q < 2; ++q) {
    dfree(s->b64[q]);
    dfree(s->cO64[q]);
  }
  dfree(s->part);
  dfree(s->acc);
  dfree(s->done_cnt);
  dfree(s->Vf64);
  dfree(s->Th64);
  dfree(s->P2);
  dfree(s->V64b);
  dfree(s->W64b);
  dfree(s->part_step);
  dfree(s->V2);
  dfree(s->W2);
  dfree(s->hist);
  dfree(s->scan_tot);
  dfree(s->scan_off);
  dfree(s->fire);
  dfree(s->fire_part);
  dfree(s->fire_tpart);
  dfree(s->P0);
  dfree(s->Th);
  for (int q = 0; q < 2; ++q) {
    dfree(s->du[q]);
    dfree(s->dth[q]);
  }
  dfree(s->wrec);
  dfree(s->wofs);
  dfree(s->wnin);
  dfree(s->wcnt);
  dfree(s->wcur);
  dfree(s->ctls);
  dfree(s->hist_b);
  dfree(s->blk_spec);
  dfree(s->blk_first);
  dfree(s->blk_end);
  dfree(s->acc_b);
  dfree(s->acc3);
  dfree(s->spec32);
  dfree(s->spec_first_d);
  if (s->gexec) (void)hipGraphExecDestroy(s->gexec);
  for (hipEvent_t e : s->gev)
    if (e) (void)hipEventDestroy(e);
  if (s->gstream) (void)hipStreamDestroy(s->gstream);
  for (hipEvent_t e : s->ev) (void)hipEventDestroy(e);
  delete s;
  return RFD_OK;
}

int rfd_system_timing(rfd_system *s, int enable) {
  if (!valid(s)) return RFD_E_ARG;
  rfd_ctx *ctx = s->ctx;
  RFD_HIP(ctx, hipStreamSynchronize(ctx->stream));
  if (enable && s->ev.empty()) {
    s->ev.resize(2 * 4096);
    for (auto &e : s->ev) RFD_HIP(ctx, hipEventCreate(&e));
  }
  s->timing = enable != 0;
  s->ev_used = 0;
  return RFD_OK;
}

int rfd_system_timing_read(rfd_system *s, double *kernel_ms, uint32_t *launches) {
  if (!valid(s)) return RFD_E_ARG;
  rfd_ctx *ctx = s->ctx;
  RFD_CHECK_ARG(ctx, kernel_ms && launches, "rfd_system_timing_read: null output");
  RFD_HIP(ctx, hipStreamSynchronize(ctx->stream));
  double tot = 0;
  for (uint32_t k = 0; k + 1 < s->ev_used; k += 2) {
    float ms = 0;
    RFD_HIP(ctx, hipEventElapsedTime(&ms, s->ev[k], s->ev[k + 1]));
    tot += ms;
  }
  *kernel_ms = tot;
  *launches = s->ev_used / 2;
  return RFD_OK;
}

int rfd_system_set_particles(rfd_system *s, const double *R, const float *radius) {
  if (!valid(s)) return RFD_E_ARG;
  rfd_ctx *ctx = s->ctx;
  RFD_CHECK_ARG(ctx, R && radius, "rfd_system_set_particles: null input");
  const uint32_t N = s->N;
  std::vector<P4> hp(N);
  std::vector<float2> hinv(N);
  double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
  float rmax = 0.f;
  for (uint32_t i = 0; i < N; ++i) {
    const double r = radius[i];
    RFD_CHECK_ARG(ctx, r > 0, "rfd_system_set_particles: radius[%u] <= 0", i);
    const double m = s->p.rho * 4.0 / 3.0 * M_PI * r * r * r;
    hinv[i] = make_float2((float)(s->p.dt / m), (float)(s->p.dt / (0.4 * m * r * r)));
    rmax = fmaxf(rmax, radius[i]);
    hp[i].x = R[i];
    hp[i].y = R[i + (size_t)N];
    hp[i].z = R[i + 2 * (size_t)N];
    hp[i].r = r;
    for (int d = 0; d < 3; ++d) {
      lo[d] = fmin(lo[d], R[i + (size_t)d * N]);
      hi[d] = fmax(hi[d], R[i + (size_t)d * N]);
    }
  }
  s->rmax = rmax;
  // grid cell >= the largest search distance: contact rows (2 rmax + skin) and
  // bond creation with a surface gap up to 0.5 rmax; margin for drift
  const double cell = fmax(2.5 * rmax + fmax(s->p.skin, 0.0) + 1e-6, 1e-3);
  const double margin = cell + s->p.farcut * 2.0 * rmax;
  s->g.ox = lo[0] - margin;
  s->g.oy = lo[1] - margin;
  s->g.oz = lo[2] - margin;
  s->g.inv_cell = 1.0 / cell;
  s->g.nx = (uint32_t)((hi[0] - lo[0] + 2 * margin) / cell) + 1;
  s->g.ny = (uint32_t)((hi[1] - lo[1] + 2 * margin) / cell) + 1;
  s->g.nz = (uint32_t)((hi[2] - lo[2] + 2 * margin) / cell) + 1;
  const uint64_t C = (uint64_t)s->g.nx * s->g.ny * s->g.nz;
  RFD_CHECK_ARG(ctx, C < (1u << 30), "rfd_system_set_particles: grid too large (%llu cells)",
                (unsigned long long)C);
  s->C = (uint32_t)C;
  int st = dalloc(ctx, &s->cstart, (size_t)s->C + 1);
  if (!st) st = dalloc(ctx, &s->ccursor, s->C);
  if (!st) st = dalloc(ctx, &s->ccnt_cells, s->C);
  if (st) return st;
  RFD_HIP(ctx, hipMemcpy(s->P, hp.data(), sizeof(P4) * N, hipMemcpyHostToDevice));
  RFD_HIP(ctx, hipMemcpy(s->inv, hinv.data(), sizeof(float2) * N, hipMemcpyHostToDevice));
  RFD_HIP(ctx, hipMemset(s->V, 0, sizeof(float4) * N));
  RFD_HIP(ctx, hipMemset(s->W, 0, sizeof(float4) * N));
  RFD_HIP(ctx, hipMemset(s->F, 0, sizeof(float4) * N));
  RFD_HIP(ctx, hipMemset(s->M, 0, sizeof(float4) * N));
  if (s->fp64) {
    RFD_HIP(ctx, hipMemset(s->V64, 0, sizeof(double4) * N));
    RFD_HIP(ctx, hipMemset(s->W64, 0, sizeof(double4) * N));
    RFD_HIP(ctx, hipMemset(s->F64, 0, sizeof(double4) * N));
    RFD_HIP(ctx, hipMemset(s->M64, 0, sizeof(double4) * N));
  }
  for (int q = 0; q < 2; ++q) RFD_HIP(ctx, hipMemset(s->ccnt[q], 0, sizeof(uint32_t) * N));
  s->have_particles = true;
  s->rows_valid = false;
  s->gen++;
  return RFD_OK;
}

int rfd_system_find_pairs(rfd_system *s, double gap, uint32_t *pairs_host, uint32_t cap,
                          uint32_t *count) {
  if (!valid(s)) return RFD_E_ARG;
  rfd_ctx *ctx = s->ctx;
  RFD_CHECK_ARG(ctx, count && s->have_particles, "rfd_system_find_pairs: bad state");
  RFD_CHECK_ARG(ctx, 2.0 * s->rmax + gap <= 1.0 / s->g.inv_cell,
                "rfd_system_find_pairs: gap %g exceeds the grid's search distance", gap);
  int st = bin_particles(s);
  if (st) return st;
  uint32_t *cnt = nullptr, *off = nullptr, *pairs = nullptr;
  st = dalloc(ctx, &cnt, s->N);
  if (!st) st = dalloc(ctx, &off, (size_t)s->N + 1);
  if (st) {
    dfree(cnt);
    return st;
  }
  NbrDev d = make_nbr(s, gap);
  d.count = cnt;
  k_nat_neighbours<0><<<nbr_grid(s->N), 128, 0, ctx->stream>>>(d, nbr_per(s->N));
  st = scan_u32(s, cnt, s->N, off, nullptr);
  if (st) {
    dfree(cnt);
    dfree(off);
    return st;
  }
  uint32_t total = 0;
  hipError_t e = hipMemcpyAsync(&total, off + s->N, sizeof(uint32_t), hipMemcpyDeviceToHost,
                                ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) {
    dfree(cnt);
    dfree(off);
    return rfd_hip_status(ctx, e, "rfd_system_find_pairs");
  }
  *count = total;
  if (pairs_host && total <= cap && total > 0) {
    st = dalloc(ctx, &pairs, 2 * (size_t)total);
    if (!st) {
      d.pair_off = off;
      d.pairs = pairs;
      k_nat_neighbours<1><<<nbr_grid(s->N), 128, 0, ctx->stream>>>(d, nbr_per(s->N));
      e = hipMemcpyAsync(pairs_host, pairs, sizeof(uint32_t) * 2 * total,
                         hipMemcpyDeviceToHost, ctx->stream);
      if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
      if (e != hipSuccess) st = rfd_hip_status(ctx, e, "rfd_system_find_pairs copy");
    }
  }
  dfree(cnt);
  dfree(off);
  dfree(pairs);
  return st;
}

int rfd_system_set_bonds(rfd_system *s, uint32_t Nb, const uint32_t *bi, const uint32_t *bj,
                         const float *br, const float *sc, const float *tc, const float *L0) {
  if (!valid(s)) return RFD_E_ARG;
  rfd_ctx *ctx = s->ctx;
  RFD_CHECK_ARG(ctx, s->have_particles, "rfd_system_set_bonds: set particles first");
  RFD_CHECK_ARG(ctx, Nb == 0 || (bi && bj && br && sc && tc), "rfd_system_set_bonds: null input");
  RFD_CHECK_ARG(ctx, Nb < 0x7FFFFFFFu, "rfd_system_set_bonds: too many bonds");
  const uint32_t N = s->N;
  // rows: particle p holds one half per incident bond, sorted by partner id
  std::vector<uint32_t> hoff(N + 1, 0);
  for (uint32_t b = 0; b < Nb; ++b) {
    RFD_CHECK_ARG(ctx, bi[b] < bj[b] && bj[b] < N, "rfd_system_set_bonds: bond %u (%u,%u) invalid",
                  b, bi[b], bj[b]);
    RFD_CHECK_ARG(ctx, b == 0 || bi[b - 1] < bi[b] || (bi[b - 1] == bi[b] && bj[b - 1] < bj[b]),
                  "rfd_system_set_bonds: bonds not sorted by (i, j) at %u", b);
    hoff[bi[b] + 1]++;
    hoff[bj[b] + 1]++;
  }
  for (uint32_t i = 0; i < N; ++i) hoff[i + 1] += hoff[i];
  const uint32_t H = 2 * Nb;
  std::vector<HR> hhr(H);
  std::vector<BR> hbr(Nb);
  std::vector<uint32_t> hm;  // the other half of each half's bond
  {
    // lower partners first (bonds where p is j, in bond order = ascending i),
    // then higher partners (bonds where p is i, ascending j): rows sorted by id
    std::vector<uint32_t> cur(hoff.begin(), hoff.end() - 1), hmir(Nb);
    hm.assign(H, 0);
    for (uint32_t b = 0; b < Nb; ++b) {
      hmir[b] = cur[bj[b]]++;
      hhr[hmir[b]] = HR{bi[b], b};
    }
    for (uint32_t b = 0; b < Nb; ++b) {
      const uint32_t ho = cur[bi[b]]++;
      hhr[ho] = HR{bj[b], b};
      hm[ho] = hmir[b];
      hm[hmir[b]] = ho;
    }
    for (uint32_t b = 0; b < Nb; ++b) {
      if (L0) RFD_CHECK_ARG(ctx, L0[b] > 0.f, "rfd_system_set_bonds: L0[%u] <= 0", b);
      hbr[b] = BR{L0 ? (float)(1.0 / (double)L0[b]) : 0.f, br[b], sc[b], tc[b]};
    }
  }
  s->hbi.assign(bi, bi + Nb);
  s->hbj.assign(bj, bj + Nb);
  s->Nb = Nb;
  s->H = H;
  for (int q = 0; q < 3; ++q) {
    const uint32_t fp = 32u << q;
    uint32_t most = 0;
    for (uint32_t p = 0; p < N; p += fp)
      most = std::max(most, hoff[std::min(p + fp, N)] - hoff[p]);
    s->tile_halves[q] = most;
  }
  int st = dalloc(ctx, &s->hr, H);
  if (!st) st = dalloc(ctx, &s->hm, H);
  if (!st) st = dalloc(ctx, &s->brc, Nb);
  if (!st) st = dalloc(ctx, &s->bst[0], Nb);
  if (!st) st = dalloc(ctx, &s->bst[1], Nb);
  if (!st) st = dalloc(ctx, &s->fl8, Nb);
  if (s->fp64) {
    if (!st) st = dalloc(ctx, &s->b64[0], Nb);
    if (!st) st = dalloc(ctx, &s->b64[1], Nb);
  }
  if (st) return st;
  RFD_HIP(ctx, hipMemcpy(s->hoff, hoff.data(), sizeof(uint32_t) * (N + 1), hipMemcpyHostToDevice));
  if (H) {
    std::vector<BSt> h0(Nb);
    for (auto &x : h0) x = BSt{make_float4(0.f, 0.f, 0.f, 0.f), 0.f, 0.f, 0.f, kBond};
    RFD_HIP(ctx, hipMemcpy(s->hr, hhr.data(), sizeof(HR) * H, hipMemcpyHostToDevice));
    RFD_HIP(ctx, hipMemcpy(s->hm, hm.data(), sizeof(uint32_t) * H, hipMemcpyHostToDevice));
    RFD_HIP(ctx, hipMemcpy(s->brc, hbr.data(), sizeof(BR) * Nb, hipMemcpyHostToDevice));
    for (int q = 0; q < 2; ++q)
      RFD_HIP(ctx, hipMemcpy(s->bst[q], h0.data(), sizeof(BSt) * Nb, hipMemcpyHostToDevice));
    RFD_HIP(ctx, hipMemset(s->fl8, kBond, Nb));
    if (s->fp64) {
      std::vector<BSt64> z(Nb);
      for (auto &x : z) x = BSt64{0, 0, 0, 0, 0, 0, 0, kBond, 0};
      for (int q = 0; q < 2; ++q)
        RFD_HIP(ctx, hipMemcpy(s->b64[q], z.data(), sizeof(BSt64) * Nb, hipMemcpyHostToDevice));
    }
    if (!L0) {
      k_set_rest_length<<<rfd_blocks(N, 256), 256, 0, ctx->stream>>>(s->P, N, s->hoff, s->hr,
                                                                     s->brc);
      RFD_HIP(ctx, hipGetLastError());
    }
  }
  {
    const int zero = 0;
    RFD_HIP(ctx, hipMemcpy(s->d_cur, &zero, sizeof(int), hipMemcpyHostToDevice));
  }
  s->have_bonds = true;
  s->gen++;
  s->work_fp = 0;
  {
    int e = ensure_work(s);
    if (e) return e;
  }
  s->rows_valid = false;
  RFD_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return RFD_OK;
}

// bond-level state API, 12 floats per bond in i's frame: n, O, Mn, Mt.  The
// device keeps O, mn = Mn . n and Mt; n (the direction the next pass treats as
// the previous one) follows from the positions and is reported, not stored.
int bond_dirs(rfd_system *s, std::vector<double> &n) {
  const uint32_t N = s->N, Nb = s->Nb;
  std::vector<P4> hp(N);
  RFD_HIP(s->ctx, hipMemcpy(hp.data(), s->P, sizeof(P4) * N, hipMemcpyDeviceToHost));
  n.assign(3 * (size_t)Nb, 0.0);
  for (uint32_t b = 0; b < Nb; ++b) {
    const P4 &a = hp[s->hbi[b]], &c = hp[s->hbj[b]];
    const double r[3] = {c.x - a.x, c.y - a.y, c.z - a.z};
    const double L = sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
    if (L > 0)
      for (int d = 0; d < 3; ++d) n[3 * (size_t)b + d] = r[d] / L;
  }
  return RFD_OK;
}

int rfd_system_set_bond_state(rfd_system *s, const float *state, const uint8_t *flag) {
  if (!valid(s)) return RFD_E_ARG;
  rfd_ctx *ctx = s->ctx;
  RFD_CHECK_ARG(ctx, s->have_bonds && state && flag, "rfd_system_set_bond_state: bad state");
  RFD_HIP(ctx, hipStreamSynchronize(ctx->stream));
  const uint32_t Nb = s->Nb;
  std::vector<double> n;
  int st = bond_dirs(s, n);
  if (st) return st;
  std::vector<BSt> hs(Nb);
  for (uint32_t b = 0; b < Nb; ++b) {
    const float *q = state + 12 * (size_t)b;
    RFD_CHECK_ARG(ctx, flag[b] <= kRemoved, "rfd_system_set_bond_state: flag[%u] = %u", b,
                  flag[b]);
    const double mn = q[6] * n[3 * (size_t)b] + q[7] * n[3 * (size_t)b + 1] +
                      q[8] * n[3 * (size_t)b + 2];
    hs[b] = BSt{make_float4(q[3], q[4], q[5], (float)mn), q[9], q[10], q[11], flag[b]};
  }
  for (int k = 0; k < 2; ++k)
    RFD_HIP(ctx, hipMemcpy(s->bst[k], hs.data(), sizeof(BSt) * Nb, hipMemcpyHostToDevice));
  RFD_HIP(ctx, hipMemcpy(s->fl8, flag, Nb, hipMemcpyHostToDevice));
  if (s->fp64) {
    std::vector<BSt64> h64(Nb);
    for (uint32_t b = 0; b < Nb; ++b) {
      const BSt &x = hs[b];
      h64[b] = BSt64{x.a.x, x.a.y, x.a.z, x.a.w, x.mx, x.my, x.mz, x.fl, 0};
    }
    for (int k = 0; k < 2; ++k)
      RFD_HIP(ctx, hipMemcpy(s->b64[k], h64.data(), sizeof(BSt64) * Nb, hipMemcpyHostToDevice));
  }
  return RFD_OK;
}

int rfd_system_get_bond_state(rfd_system *s, float *state, uint8_t *flag) {
  if (!valid(s)) return RFD_E_ARG;
  rfd_ctx *ctx = s->ctx;
  RFD_CHECK_ARG(ctx, s->have_bonds, "rfd_system_get_bond_state: no bonds");
  RFD_HIP(ctx, hipStreamSynchronize(ctx->stream));
  const uint32_t Nb = s->Nb;
  std::vector<BSt> hs(Nb);
  int cur = 0;
  RFD_HIP(ctx, hipMemcpy(&cur, s->d_cur, sizeof(int), hipMemcpyDeviceToHost));
  if (s->fp64) {
    std::vector<BSt64> h64(Nb);
    RFD_HIP(ctx, hipMemcpy(h64.data(), s->b64[cur], sizeof(BSt64) * Nb, hipMemcpyDeviceToHost));
    for (uint32_t b = 0; b < Nb; ++b) {
      const BSt64 &x = h64[b];
      hs[b] = BSt{make_float4((float)x.ox, (float)x.oy, (float)x.oz, (float)x.mn), (float)x.mx,
                  (float)x.my, (float)x.mz, x.fl};
    }
  } else {
    RFD_HIP(ctx, hipMemcpy(hs.data(), s->bst[cur], sizeof(BSt) * Nb, hipMemcpyDeviceToHost));
  }
  if (state) {
    std::vector<double> n;
    int st = bond_dirs(s, n);
    if (st) return st;
    for (uint32_t b = 0; b < Nb; ++b) {
      float *q = state + 12 * (size_t)b;
      const BSt &x = hs[b];
      for (int d = 0; d < 3; ++d) {
        q[d] = (float)n[3 * (size_t)b + d];
        q[6 + d] = x.a.w * (float)n[3 * (size_t)b + d];
      }
      q[3] = x.a.x;
      q[4] = x.a.y;
      q[5] = x.a.z;
      q[9] = x.mx;
      q[10] = x.my;
      q[11] = x.mz;
    }
  }
  if (flag)
    for (uint32_t b = 0; b < Nb; ++b) flag[b] = (uint8_t)hs[b].fl;
  return RFD_OK;
}

int upload_vec3_64(rfd_ctx *ctx, double4 *dst, const float *soa, uint32_t N) {
  std::vector<double4> h(N);
  for (uint32_t i = 0; i < N; ++i)
    h[i] = make_double4(soa[i], soa[i + (size_t)N], soa[i + 2 * (size_t)N], 0.0);
  RFD_HIP(ctx, hipMemcpy(dst, h.data(), sizeof(double4) * N, hipMemcpyHostToDevice));
  return RFD_OK;
}

int download_vec3_64(rfd_ctx *ctx, float *soa, const double4 *src, uint32_t N) {
  std::vector<double4> h(N);
  RFD_HIP(ctx, hipMemcpy(h.data(), src, sizeof(double4) * N, hipMemcpyDeviceToHost));
  for (uint32_t i = 0; i < N; ++i) {
    soa[i] = (float)h[i].x;
    soa[i + (size_t)N] = (float)h[i].y;
    soa[i + 2 * (size_t)N] = (float)h[i].z;
  }
  return RFD_OK;
}

// ------------------------------------------- reference slot layout in/out
int rfd_system_import_reference_layout(rfd_system *s, const float *R, const float *V,
                                       const float *W, const uint32_t *IL,
                                       const uint8_t *ILtype, const float *iL, const float *rij,
                                       const float *Oijt, const float *Mijn, const float *Mijt,
                                       uint32_t N, uint32_t K, float radius, float b_r, float sc,
                                       float tc) {
  if (!valid(s)) return RFD_E_ARG;
  rfd_ctx *ctx = s->ctx;
  RFD_CHECK_ARG(ctx, R && V && W && IL && ILtype && iL && rij && Oijt && Mijn && Mijt,
                "rfd_system_import_reference_layout: null buffer");
  RFD_CHECK_ARG(ctx, N == s->N && K > 0,
                "rfd_system_import_reference_layout: N = %u, the system holds %u", N, s->N);
  RFD_CHECK_ARG(ctx, radius > 0 && b_r > 0, "rfd_system_import_reference_layout: radii <= 0");
  RFD_CHECK_ARG(ctx, s->S == 0,
                "rfd_system_import_reference_layout: the system holds batched specimens");
  const size_t S = (size_t)N * K, N3 = 3 * (size_t)N;
  RFD_HIP(ctx, hipStreamSynchronize(ctx->stream));
  std::vector<float> hR(N3), hV(N3), hW(N3), hiL(S), hn(3 * S), hO(3 * S), hMn(3 * S),
      hMt(3 * S);
  std::vector<uint32_t> hIL(S);
  std::vector<uint8_t> hty(S);
  RFD_HIP(ctx, hipMemcpy(hR.data(), R, sizeof(float) * N3, hipMemcpyDeviceToHost));
  RFD_HIP(ctx, hipMemcpy(hV.data(), V, sizeof(float) * N3, hipMemcpyDeviceToHost));
  RFD_HIP(ctx, hipMemcpy(hW.data(), W, sizeof(float) * N3, hipMemcpyDeviceToHost));
  RFD_HIP(ctx, hipMemcpy(hIL.data(), IL, sizeof(uint32_t) * S, hipMemcpyDeviceToHost));
  RFD_HIP(ctx, hipMemcpy(hty.data(), ILtype, S, hipMemcpyDeviceToHost));
  RFD_HIP(ctx, hipMemcpy(hiL.data(), iL, sizeof(float) * S, hipMemcpyDeviceToHost));
  RFD_HIP(ctx, hipMemcpy(hn.data(), rij, sizeof(float) * 3 * S, hipMemcpyDeviceToHost));
  RFD_HIP(ctx, hipMemcpy(hO.data(), Oijt, sizeof(float) * 3 * S, hipMemcpyDeviceToHost));
  RFD_HIP(ctx, hipMemcpy(hMn.data(), Mijn, sizeof(float) * 3 * S, hipMemcpyDeviceToHost));
  RFD_HIP(ctx, hipMemcpy(hMt.data(), Mijt, sizeof(float) * 3 * S, hipMemcpyDeviceToHost));
  // particles (FP32 positions are exact in FP64), uniform radius
  std::vector<double> R64(hR.begin(), hR.end());
  std::vector<float> rad(N, radius);
  int st = rfd_system_set_particles(s, R64.data(), rad.data());
  if (st) return st;
  // bonds: slot i -> j of type 1 with i < j, in (i, j) order; state of that slot
  std::vector<std::pair<uint64_t, size_t>> own;
  for (uint32_t i = 0; i < N; ++i)
    for (uint32_t k = 0; k < K; ++k) {
      const size_t sl = (size_t)i * K + k;
      const uint32_t j = hIL[sl];
      if (j == 0xFFFFFFFFu || hty[sl] != 1 || j <= i) continue;
      RFD_CHECK_ARG(ctx, j < N, "rfd_system_import_reference_layout: IL[%zu] = %u >= N", sl, j);
      own.emplace_back(((uint64_t)i << 32) | j, sl);
    }
  std::sort(own.begin(), own.end());
  const uint32_t Nb = (uint32_t)own.size();
  std::vector<uint32_t> bi(Nb), bj(Nb);
  std::vector<float> br(Nb, b_r), bsc(Nb, sc), btc(Nb, tc), L0(Nb), state(12 * (size_t)Nb);
  std::vector<uint8_t> flag(Nb, (uint8_t)kBond);
  for (uint32_t b = 0; b < Nb; ++b) {
    RFD_CHECK_ARG(ctx, b == 0 || own[b].first != own[b - 1].first,
                  "rfd_system_import_reference_layout: pair (%u,%u) bonded twice",
                  (uint32_t)(own[b].first >> 32), (uint32_t)own[b].first);
    const size_t sl = own[b].second;
    bi[b] = (uint32_t)(own[b].first >> 32);
    bj[b] = (uint32_t)own[b].first;
    // the bond law uses 1/iL as the rest length (md_calculateforcesdem3.cu:96)
    L0[b] = 1.0f / hiL[sl];
    float *q = state.data() + 12 * (size_t)b;
    for (int d = 0; d < 3; ++d) {
      q[d] = hn[3 * sl + d];
      q[3 + d] = hO[3 * sl + d];
      q[6 + d] = hMn[3 * sl + d];
      q[9 + d] = hMt[3 * sl + d];
    }
  }
  st = rfd_system_set_bonds(s, Nb, bi.data(), bj.data(), br.data(), bsc.data(), btc.data(),
                            L0.data());
  if (!st && Nb) st = rfd_system_set_bond_state(s, state.data(), flag.data());
  if (!st) st = rfd_system_set_velocities(s, hV.data(), hW.data());
  if (st) return st;
  // contact slots with a spring, or within ri + rj + skin, become the contact
  // rows (each end keeps its own spring, as each slot does), then one row
  // search completes the rows and carries the springs over
  const uint32_t Kc = s->p.contact_capacity;
  std::vector<uint32_t> cj((size_t)Kc * N, 0), ccnt(N, 0);
  std::vector<float4> cO((size_t)Kc * N, make_float4(0.f, 0.f, 0.f, 0.f));
  const double reach = 2.0 * radius + std::max(s->p.skin, 0.0);
  for (uint32_t i = 0; i < N; ++i)
    for (uint32_t k = 0; k < K; ++k) {
      const size_t sl = (size_t)i * K + k;
      const uint32_t j = hIL[sl];
      if (j == 0xFFFFFFFFu || hty[sl] != 0 || j >= N || j == i) continue;
      const bool spring = hO[3 * sl] != 0.f || hO[3 * sl + 1] != 0.f || hO[3 * sl + 2] != 0.f;
      double rr = 0.0;
      for (int d = 0; d < 3; ++d) {
        const double x = (double)hR[j + (size_t)d * N] - hR[i + (size_t)d * N];
        rr += x * x;
      }
      if (!spring && rr >= reach * reach) continue;
      RFD_CHECK_ARG(ctx, ccnt[i] < Kc,
                    "rfd_system_import_reference_layout: particle %u has more than %u contacts",
                    i, Kc);
      const size_t c = (size_t)ccnt[i]++ * N + i;
      cj[c] = j;
      cO[c] = make_float4(hO[3 * sl], hO[3 * sl + 1], hO[3 * sl + 2], 0.f);
    }
  const int o = s->ccur;
  RFD_HIP(ctx, hipMemcpy(s->cj[o], cj.data(), sizeof(uint32_t) * cj.size(),
                         hipMemcpyHostToDevice));
  RFD_HIP(ctx, hipMemcpy(s->ccnt[o], ccnt.data(), sizeof(uint32_t) * N, hipMemcpyHostToDevice));
  RFD_HIP(ctx, hipMemcpy(s->cO[o], cO.data(), sizeof(float4) * cO.size(),
                         hipMemcpyHostToDevice));
  if (s->fp64) {
    std::vector<double4> c64(cO.size());
    for (size_t c = 0; c < cO.size(); ++c) c64[c] = make_double4(cO[c].x, cO[c].y, cO[c].z, 0.0);
    RFD_HIP(ctx, hipMemcpy(s->cO64[o], c64.data(), sizeof(double4) * c64.size(),
                           hipMemcpyHostToDevice));
  }
  s->rows_valid = true;
  st = rebuild_rows(s);
  if (st) return st;
  s->gen++;
  RFD_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return RFD_OK;
}

int rfd_system_export_reference_layout(rfd_system *s, float *R, float *V, float *W,
                                       uint32_t *IL, uint8_t *ILtype, float *iL, float *rij,
                                       float *Oijt, float *Mijn, float *Mijt, uint32_t K) {
  if (!valid(s)) return RFD_E_ARG;
  rfd_ctx *ctx = s->ctx;
  RFD_CHECK_ARG(ctx, R && V && W && IL && ILtype && iL && rij && Oijt && Mijn && Mijt && K > 0,
                "rfd_system_export_reference_layout: bad arguments");
  RFD_CHECK_ARG(ctx, s->have_particles, "rfd_system_export_reference_layout: no particles");
  RFD_HIP(ctx, hipStreamSynchronize(ctx->stream));
  const uint32_t N = s->N, Nb = s->have_bonds ? s->Nb : 0;
  const size_t S = (size_t)N * K, N3 = 3 * (size_t)N;
  std::vector<P4> hp(N);
  RFD_HIP(ctx, hipMemcpy(hp.data(), s->P, sizeof(P4) * N, hipMemcpyDeviceToHost));
  std::vector<float> hR(N3), hV(N3), hW(N3);
  if (s->fp64) {
    int st = download_vec3_64(ctx, hV.data(), s->V64, N);
    if (!st) st = download_vec3_64(ctx, hW.data(), s->W64, N);
    if (st) return st;
  } else {
    int st = download_vec3(ctx, hV.data(), s->V, N);
    if (!st) st = download_vec3(ctx, hW.data(), s->W, N);
    if (st) return st;
  }
  for (uint32_t i = 0; i < N; ++i) {
    hR[i] = (float)hp[i].x;
    hR[i + (size_t)N] = (float)hp[i].y;
    hR[i + 2 * (size_t)N] = (float)hp[i].z;
  }
  std::vector<float> state(12 * (size_t)std::max(Nb, 1u));
  std::vector<uint8_t> flag(std::max(Nb, 1u));
  if (Nb) {
    int st = rfd_system_get_bond_state(s, state.data(), flag.data());
    if (st) return st;
  }
  std::vector<uint32_t> hIL(S, 0xFFFFFFFFu), cnt(N, 0);
  std::vector<uint8_t> hty(S, 0);
  std::vector<float> hiL(S, 0.f), hn(3 * S, 0.f), hO(3 * S, 0.f), hMn(3 * S, 0.f),
      hMt(3 * S, 0.f);
  uint32_t full = UINT32_MAX;
  // one slot of row i; sgn = -1 writes the mirror image (the other end's view)
  auto put = [&](uint32_t i, uint32_t j, uint8_t ty, float il, const float *n3, const float *o3,
                 const float *mn3, const float *mt3, float sgn) {
    if (cnt[i] >= K) {
      full = std::min(full, i);
      return;
    }
    const size_t sl = (size_t)i * K + cnt[i]++;
    hIL[sl] = j;
    hty[sl] = ty;
    hiL[sl] = il;
    for (int d = 0; d < 3; ++d) {
      hn[3 * sl + d] = n3 ? sgn * n3[d] : 0.f;
      hO[3 * sl + d] = o3 ? sgn * o3[d] : 0.f;
      hMn[3 * sl + d] = mn3 ? sgn * mn3[d] : 0.f;
      hMt[3 * sl + d] = mt3 ? sgn * mt3[d] : 0.f;
    }
  };
  std::vector<BR> hb(Nb);
  if (Nb) RFD_HIP(ctx, hipMemcpy(hb.data(), s->brc, sizeof(BR) * Nb, hipMemcpyDeviceToHost));
  const double dt = s->p.dt;
  // intact bonds first, both directions, partners ascending; then broken bonds
  for (int pass = 0; pass < 2; ++pass)
    for (uint32_t b = 0; b < Nb; ++b) {
      if (flag[b] != (pass == 0 ? kBond : kBroken)) continue;
      const uint32_t i = s->hbi[b], j = s->hbj[b];
      const float *q = state.data() + 12 * (size_t)b;
      // previous direction at R - V dt, as the force pass forms it
      const P4 &a = hp[i], &c = hp[j];
      double r[3] = {c.x - a.x, c.y - a.y, c.z - a.z};
      for (int d = 0; d < 3; ++d)
        r[d] -= ((double)hV[j + (size_t)d * N] - hV[i + (size_t)d * N]) * dt;
      const double L = sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
      const float n3[3] = {(float)(r[0] / L), (float)(r[1] / L), (float)(r[2] / L)};
      // twist: the stored component along the current direction q[0..2]
      const float mn = q[6] * q[0] + q[7] * q[1] + q[8] * q[2];
      const float mn3[3] = {mn * n3[0], mn * n3[1], mn * n3[2]};
      const float il = hb[b].il;
      if (pass == 0) {
        put(i, j, 1, il, n3, q + 3, mn3, q + 9, 1.f);
        put(j, i, 1, il, n3, q + 3, mn3, q + 9, -1.f);
      } else {
        put(i, j, 0, il, nullptr, q + 3, nullptr, nullptr, 1.f);
        put(j, i, 0, il, nullptr, q + 3, nullptr, nullptr, -1.f);
      }
    }
  // never-bonded contact rows with their per-end springs
  if (s->rows_valid) {
    const uint32_t Kc = s->p.contact_capacity;
    const int o = s->ccur;
    std::vector<uint32_t> cj((size_t)Kc * N), ccnt(N);
    std::vector<float4> cO((size_t)Kc * N);
    RFD_HIP(ctx, hipMemcpy(ccnt.data(), s->ccnt[o], sizeof(uint32_t) * N, hipMemcpyDeviceToHost));
    RFD_HIP(ctx, hipMemcpy(cj.data(), s->cj[o], sizeof(uint32_t) * cj.size(),
                           hipMemcpyDeviceToHost));
    if (s->fp64) {
      std::vector<double4> c64(cO.size());
      RFD_HIP(ctx, hipMemcpy(c64.data(), s->cO64[o], sizeof(double4) * c64.size(),
                             hipMemcpyDeviceToHost));
      for (size_t c = 0; c < c64.size(); ++c)
        cO[c] = make_float4((float)c64[c].x, (float)c64[c].y, (float)c64[c].z, 0.f);
    } else {
      RFD_HIP(ctx, hipMemcpy(cO.data(), s->cO[o], sizeof(float4) * cO.size(),
                             hipMemcpyDeviceToHost));
    }
    for (uint32_t i = 0; i < N; ++i)
      for (uint32_t k = 0; k < std::min(ccnt[i], Kc); ++k) {
        const size_t c = (size_t)k * N + i;
        const float o3[3] = {cO[c].x, cO[c].y, cO[c].z};
        put(i, cj[c], 0, 0.f, nullptr, o3, nullptr, nullptr, 1.f);
      }
  }
  RFD_CHECK_ARG(ctx, full == UINT32_MAX,
                "rfd_system_export_reference_layout: row of particle %u needs more than %u slots",
                full, K);
  RFD_HIP(ctx, hipMemcpy(R, hR.data(), sizeof(float) * N3, hipMemcpyHostToDevice));
  RFD_HIP(ctx, hipMemcpy(V, hV.data(), sizeof(float) * N3, hipMemcpyHostToDevice));
  RFD_HIP(ctx, hipMemcpy(W, hW.data(), sizeof(float) * N3, hipMemcpyHostToDevice));
  RFD_HIP(ctx, hipMemcpy(IL, hIL.data(), sizeof(uint32_t) * S, hipMemcpyHostToDevice));
  RFD_HIP(ctx, hipMemcpy(ILtype, hty.data(), S, hipMemcpyHostToDevice));
  RFD_HIP(ctx, hipMemcpy(iL, hiL.data(), sizeof(float) * S, hipMemcpyHostToDevice));
  RFD_HIP(ctx, hipMemcpy(rij, hn.data(), sizeof(float) * 3 * S, hipMemcpyHostToDevice));
  RFD_HIP(ctx, hipMemcpy(Oijt, hO.data(), sizeof(float) * 3 * S, hipMemcpyHostToDevice));
  RFD_HIP(ctx, hipMemcpy(Mijn, hMn.data(), sizeof(float) * 3 * S, hipMemcpyHostToDevice));
  RFD_HIP(ctx, hipMemcpy(Mijt, hMt.data(), sizeof(float) * 3 * S, hipMemcpyHostToDevice));
  return RFD_OK;
}

int rfd_system_set_velocities(rfd_system *s, const float *V, const float *W) {
  if (!valid(s)) return RFD_E_ARG;
  int st = RFD_OK;
  if (s->fp64) {
    if (V) st = upload_vec3_64(s->ctx, s->V64, V, s->N);
    if (!st && W) st = upload_vec3_64(s->ctx, s->W64, W, s->N);
    return st;
  }
  if (V) st = upload_vec3(s->ctx, s->V, V, s->N);
  if (!st && W) st = upload_vec3(s->ctx, s->W, W, s->N);
  return st;
}

int rfd_system_set_loading(rfd_system *s, const rfd_loading *L) {
  if (!valid(s)) return RFD_E_ARG;
  rfd_ctx *ctx = s->ctx;
  RFD_CHECK_ARG(ctx, L && L->mode >= 0 && L->mode <= 2, "rfd_system_set_loading: bad loading");
  RFD_CHECK_ARG(ctx, s->S == 0, "rfd_system_set_loading: the system holds batched specimens");
  RFD_HIP(ctx, hipStreamSynchronize(ctx->stream));
  s->gen++;
  RFD_HIP(ctx, hipMemcpy(s->ctl, L, sizeof(*L), hipMemcpyHostToDevice));
  s->mode = L->mode;
  return RFD_OK;
}

int rfd_system_get_loading(rfd_system *s, rfd_loading *L) {
  if (!valid(s)) return RFD_E_ARG;
  rfd_ctx *ctx = s->ctx;
  RFD_CHECK_ARG(ctx, L, "rfd_system_get_loading: null output");
  RFD_HIP(ctx, hipStreamSynchronize(ctx->stream));
  RFD_HIP(ctx, hipMemcpy(L, s->ctl, sizeof(*L), hipMemcpyDeviceToHost));
  return RFD_OK;
}

int rfd_system_get_state(rfd_system *s, double *R, float *V, float *W, float *F, float *M) {
  if (!valid(s)) return RFD_E_ARG;
  rfd_ctx *ctx = s->ctx;
  const uint32_t N = s->N;
  RFD_HIP(ctx, hipStreamSynchronize(ctx->stream));
  if (R) {
    std::vector<P4> h(N);
    RFD_HIP(ctx, hipMemcpy(h.data(), s->P, sizeof(P4) * N, hipMemcpyDeviceToHost));
    for (uint32_t i = 0; i < N; ++i) {
      R[i] = h[i].x;
      R[i + (size_t)N] = h[i].y;
      R[i + 2 * (size_t)N] = h[i].z;
    }
  }
  int st = RFD_OK;
  if (s->fp64) {
    if (V) st = download_vec3_64(ctx, V, s->V64, N);
    if (!st && W) st = download_vec3_64(ctx, W, s->W64, N);
    if (!st && F) st = download_vec3_64(ctx, F, s->F64, N);
    if (!st && M) st = download_vec3_64(ctx, M, s->M64, N);
    return st;
  }
  if (V) st = download_vec3(ctx, V, s->V, N);
  if (!st && W) st = download_vec3(ctx, W, s->W, N);
  if (!st && F) st = download_vec3(ctx, F, s->F, N);
  if (!st && M) st = download_vec3(ctx, M, s->M, N);
  return st;
}

// device vec3 records (float4 or double4) to host SoA doubles
extern "C++" {
template <class T4>
int download_soa64(rfd_ctx *ctx, double *soa, const T4 *src, uint32_t N) {
  std::vector<T4> h(N);
  RFD_HIP(ctx, hipMemcpy(h.data(), src, sizeof(T4) * N, hipMemcpyDeviceToHost));
  for (uint32_t i = 0; i < N; ++i) {
    soa[i] = (double)h[i].x;
    soa[i + (size_t)N] = (double)h[i].y;
    soa[i + 2 * (size_t)N] = (double)h[i].z;
  }
  return RFD_OK;
}
}

int rfd_system_get_state64(rfd_system *s, double *R, double *V, double *W, double *F,
                           double *M) {
  if (!valid(s)) return RFD_E_ARG;
  int st = rfd_system_get_state(s, R, nullptr, nullptr, nullptr, nullptr);
  if (st) return st;
  rfd_ctx *ctx = s->ctx;
  const uint32_t N = s->N;
  if (s->fp64) {
    if (V) st = download_soa64(ctx, V, s->V64, N);
    if (!st && W) st = download_soa64(ctx, W, s->W64, N);
    if (!st && F) st = download_soa64(ctx, F, s->F64, N);
    if (!st && M) st = download_soa64(ctx, M, s->M64, N);
    return st;
  }
  if (V) st = download_soa64(ctx, V, s->V, N);
  if (!st && W) st = download_soa64(ctx, W, s->W, N);
  if (!st && F) st = download_soa64(ctx, F, s->F, N);
  if (!st && M) st = download_soa64(ctx, M, s->M, N);
  return st;
}

int rfd_system_compute_forces(rfd_system *s) {
  if (!valid(s)) return RFD_E_ARG;
  rfd_ctx *ctx = s->ctx;
  RFD_CHECK_ARG(ctx, s->have_bonds, "rfd_system_compute_forces: set bonds first");
  if (!s->rows_valid) {
    int st = rebuild_rows(s);
    if (st) return st;
  }
  return forces(s);
}

int rfd_system_step(rfd_system *s, uint32_t nsteps) {
  if (!valid(s)) return RFD_E_ARG;
  rfd_ctx *ctx = s->ctx;
  RFD_CHECK_ARG(ctx, s->have_bonds, "rfd_system_step: set bonds first");
  return run_steps(s, nsteps);
}

int rfd_system_set_fusion(rfd_system *s, int on) {
  if (!valid(s)) return RFD_E_ARG;
  s->fuse = on != 0;
  s->gen++;
  return RFD_OK;
}

int rfd_system_get_fusion(rfd_system *s, int *fused) {
  if (!valid(s) || !fused) return RFD_E_ARG;
  *fused = can_fuse(s) || can_fuse64(s) ? 1 : 0;
  return RFD_OK;
}

int rfd_system_set_persistent(rfd_system *s, int on) {
  if (!valid(s)) return RFD_E_ARG;
  s->persist = on != 0;
  return RFD_OK;
}

int rfd_system_get_persistent(rfd_system *s, int *on, int *applies) {
  if (!valid(s) || !on || !applies) return RFD_E_ARG;
  *on = s->persist ? 1 : 0;
  *applies = steps_fit(s) ? 1 : 0;
  return RFD_OK;
}

// ------------------------------------------------------- batched specimens
int rfd_system_set_specimens(rfd_system *s, uint32_t S, const uint32_t *first,
                             const rfd_loading *L, uint32_t hist_cap) {
  if (!valid(s)) return RFD_E_ARG;
  rfd_ctx *ctx = s->ctx;
  RFD_HIP(ctx, hipStreamSynchronize(ctx->stream));
  s->gen++;
  if (S == 0) {
    s->S = 0;
    s->spec_first.clear();
    return RFD_OK;
  }
  RFD_CHECK_ARG(ctx, first && L && hist_cap > 0, "rfd_system_set_specimens: bad arguments");
  RFD_CHECK_ARG(ctx, !s->fp64, "rfd_system_set_specimens: batched specimens run the FP32 path");
  RFD_CHECK_ARG(ctx, first[0] == 0 && first[S] == s->N,
                "rfd_system_set_specimens: first[0] must be 0 and first[S] = N");
  const int mode = L[0].mode;
  RFD_CHECK_ARG(ctx, mode == 1 || mode == 2, "rfd_system_set_specimens: loading mode must be 1 or 2");
  std::vector<uint32_t> bs, bf, be;
  for (uint32_t k = 0; k < S; ++k) {
    RFD_CHECK_ARG(ctx, first[k + 1] > first[k], "rfd_system_set_specimens: specimen %u is empty", k);
    RFD_CHECK_ARG(ctx, L[k].mode == mode, "rfd_system_set_specimens: specimens mix loading modes");
    for (uint32_t a = first[k]; a < first[k + 1]; a += 256) {
      bs.push_back(k);
      bf.push_back(a);
      be.push_back(std::min(a + 256u, first[k + 1]));
    }
  }
  const uint32_t nblk = (uint32_t)bs.size();
  int st = dalloc(ctx, &s->ctls, S);
  if (!st) st = dalloc(ctx, &s->hist_b, 3 * (size_t)hist_cap * S);
  if (!st) st = dalloc(ctx, &s->blk_spec, nblk);
  if (!st) st = dalloc(ctx, &s->blk_first, nblk);
  if (!st) st = dalloc(ctx, &s->blk_end, nblk);
  if (!st) st = dalloc(ctx, &s->acc_b, 2 * (size_t)S);
  const size_t n32 = ((size_t)s->N + 31) / 32;
  if (!st) st = dalloc(ctx, &s->spec32, n32);
  if (!st) st = dalloc(ctx, &s->spec_first_d, (size_t)S + 1);
  if (st) return st;
  RFD_HIP(ctx, hipMemset(s->acc_b, 0, sizeof(unsigned long long) * 2 * S));
  {
    std::vector<uint32_t> sp32(n32);
    uint32_t k = 0, smallest = UINT32_MAX;
    for (size_t t = 0; t < n32; ++t) {
      while (first[k + 1] <= 32 * t) ++k;
      sp32[t] = k;
    }
    for (uint32_t q = 0; q < S; ++q) smallest = std::min(smallest, first[q + 1] - first[q]);
    RFD_HIP(ctx, hipMemcpy(s->spec32, sp32.data(), sizeof(uint32_t) * n32,
                           hipMemcpyHostToDevice));
    RFD_HIP(ctx, hipMemcpy(s->spec_first_d, first, sizeof(uint32_t) * (S + 1),
                           hipMemcpyHostToDevice));
    s->batch_fusable = smallest >= 128;
  }
  std::vector<rfd_loading> Lh(L, L + S);
  for (auto &x : Lh) {
    x.steps = 0;
    x.hist_len = 0;
  }
  RFD_HIP(ctx, hipMemcpy(s->ctls, Lh.data(), sizeof(rfd_loading) * S, hipMemcpyHostToDevice));
  RFD_HIP(ctx, hipMemcpy(s->blk_spec, bs.data(), sizeof(uint32_t) * nblk, hipMemcpyHostToDevice));
  RFD_HIP(ctx, hipMemcpy(s->blk_first, bf.data(), sizeof(uint32_t) * nblk, hipMemcpyHostToDevice));
  RFD_HIP(ctx, hipMemcpy(s->blk_end, be.data(), sizeof(uint32_t) * nblk, hipMemcpyHostToDevice));
  s->S = S;
  s->spec_first.assign(first, first + S + 1);
  s->hist_cap_b = hist_cap;
  s->nblk_b = nblk;
  s->mode = mode;
  return RFD_OK;
}

int rfd_system_get_specimen_loading(rfd_system *s, uint32_t S, rfd_loading *L) {
  if (!valid(s)) return RFD_E_ARG;
  rfd_ctx *ctx = s->ctx;
  RFD_CHECK_ARG(ctx, L && S == s->S && S > 0,
                "rfd_system_get_specimen_loading: S must equal the specimen count");
  RFD_HIP(ctx, hipStreamSynchronize(ctx->stream));
  RFD_HIP(ctx, hipMemcpy(L, s->ctls, sizeof(rfd_loading) * S, hipMemcpyDeviceToHost));
  return RFD_OK;
}

int rfd_system_get_specimen_history(rfd_system *s, uint32_t k, double *hist, uint32_t cap,
                                    uint32_t *n) {
  if (!valid(s)) return RFD_E_ARG;
  rfd_ctx *ctx = s->ctx;
  RFD_CHECK_ARG(ctx, n && k < s->S, "rfd_system_get_specimen_history: bad specimen");
  rfd_loading L;
  RFD_HIP(ctx, hipStreamSynchronize(ctx->stream));
  RFD_HIP(ctx, hipMemcpy(&L, s->ctls + k, sizeof(L), hipMemcpyDeviceToHost));
  *n = L.hist_len;
  const uint32_t m = std::min(std::min(L.hist_len, cap), s->hist_cap_b);
  if (hist && m)
    RFD_HIP(ctx, hipMemcpy(hist, s->hist_b + 3 * (size_t)s->hist_cap_b * k,
                           sizeof(double) * 3 * m, hipMemcpyDeviceToHost));
  return RFD_OK;
}

// ------------------------------------------------------------- cell grid
int rfd_system_set_grid(rfd_system *s, const double origin[3], double cell,
                        const uint32_t dims[3]) {
  if (!valid(s)) return RFD_E_ARG;
  rfd_ctx *ctx = s->ctx;
  RFD_CHECK_ARG(ctx, origin && dims && s->have_particles, "rfd_system_set_grid: bad arguments");
  const double need = 2.5 * s->rmax + fmax(s->p.skin, 0.0);
  RFD_CHECK_ARG(ctx, cell >= need, "rfd_system_set_grid: cell %g below 2.5 rmax + skin = %g",
                cell, need);
  RFD_CHECK_ARG(ctx, dims[0] && dims[1] && dims[2], "rfd_system_set_grid: empty grid");
  const uint64_t C = (uint64_t)dims[0] * dims[1] * dims[2];
  RFD_CHECK_ARG(ctx, C < (1u << 30), "rfd_system_set_grid: grid too large (%llu cells)",
                (unsigned long long)C);
  RFD_HIP(ctx, hipStreamSynchronize(ctx->stream));
  s->g.ox = origin[0];
  s->g.oy = origin[1];
  s->g.oz = origin[2];
  s->g.inv_cell = 1.0 / cell;
  s->g.nx = dims[0];
  s->g.ny = dims[1];
  s->g.nz = dims[2];
  s->C = (uint32_t)C;
  int st = dalloc(ctx, &s->cstart, (size_t)s->C + 1);
  if (!st) st = dalloc(ctx, &s->ccursor, s->C);
  if (!st) st = dalloc(ctx, &s->ccnt_cells, s->C);
  if (st) return st;
  s->rows_valid = false;
  s->gen++;
  return RFD_OK;
}

int rfd_system_get_grid(rfd_system *s, double origin[3], double *cell, uint32_t dims[3]) {
  if (!valid(s)) return RFD_E_ARG;
  if (origin) {
    origin[0] = s->g.ox;
    origin[1] = s->g.oy;
    origin[2] = s->g.oz;
  }
  if (cell) *cell = 1.0 / s->g.inv_cell;
  if (dims) {
    dims[0] = s->g.nx;
    dims[1] = s->g.ny;
    dims[2] = s->g.nz;
  }
  return RFD_OK;
}

// ------------------------------------------------------------ step graphs
int rfd_system_step_graph(rfd_system *s, uint32_t nsteps) {
  if (!valid(s)) return RFD_E_ARG;
  rfd_ctx *ctx = s->ctx;
  RFD_CHECK_ARG(ctx, s->have_bonds, "rfd_system_step_graph: set bonds first");
  const uint32_t R = s->p.rebuild_every, G = 2 * R;
  uint32_t k = 0;
  while (k < nsteps) {
    // a graph covers two rebuild periods from a rebuild step, so that the
    // contact-row buffers it reads and writes come back to where they started
    const bool whole = s->rows_valid && s->step_count % R == 0 && nsteps - k >= G && !s->timing;
    if (!whole) {
      const int st = step_plain(s);
      if (st) return st;
      ++k;
      continue;
    }
    if (!s->gexec || s->g_gen != s->gen || s->g_ccur != s->ccur || s->g_steps != G) {
      if (s->gexec) {
        RFD_HIP(ctx, hipGraphExecDestroy(s->gexec));
        s->gexec = nullptr;
      }
      if (!s->gstream) {
        RFD_HIP(ctx, hipStreamCreateWithFlags(&s->gstream, hipStreamNonBlocking));
        RFD_HIP(ctx, hipEventCreateWithFlags(&s->gev[0], hipEventDisableTiming));
        RFD_HIP(ctx, hipEventCreateWithFlags(&s->gev[1], hipEventDisableTiming));
      }
      // capture: the step code launches onto the context stream, pointed at
      // the private stream for the duration; nothing is executed
      const hipStream_t home = ctx->stream;
      const uint64_t sc0 = s->step_count;
      const int cc0 = s->ccur;
      RFD_HIP(ctx, hipStreamBeginCapture(s->gstream, hipStreamCaptureModeThreadLocal));
      ctx->stream = s->gstream;
      const int st = run_steps(s, G);
      ctx->stream = home;
      hipGraph_t graph = nullptr;
      const hipError_t ec = hipStreamEndCapture(s->gstream, &graph);
      s->step_count = sc0;
      s->ccur = cc0;
      if (st) {
        if (graph) (void)hipGraphDestroy(graph);
        return st;
      }
      RFD_HIP(ctx, ec);
      const hipError_t ei = hipGraphInstantiate(&s->gexec, graph, nullptr, nullptr, 0);
      (void)hipGraphDestroy(graph);
      RFD_HIP(ctx, ei);
      s->g_gen = s->gen;
      s->g_ccur = s->ccur;
      s->g_steps = G;
    }
    // the private stream waits for the context stream's earlier work, and the
    // context stream for the replay
    RFD_HIP(ctx, hipEventRecord(s->gev[0], ctx->stream));
    RFD_HIP(ctx, hipStreamWaitEvent(s->gstream, s->gev[0], 0));
    RFD_HIP(ctx, hipGraphLaunch(s->gexec, s->gstream));
    RFD_HIP(ctx, hipEventRecord(s->gev[1], s->gstream));
    RFD_HIP(ctx, hipStreamWaitEvent(ctx->stream, s->gev[1], 0));
    s->step_count += G;
    k += G;
  }
  return RFD_OK;
}

// -------------------------------------------------------------- autotuner
int rfd_system_set_tiles(rfd_system *s, int force_tile, int nbr_block, int single_eval) {
  if (!valid(s)) return RFD_E_ARG;
  rfd_ctx *ctx = s->ctx;
  RFD_CHECK_ARG(ctx, force_tile == 32 || force_tile == 64 || force_tile == 128,
                "rfd_system_set_tiles: force tile must be 32, 64 or 128");
  RFD_CHECK_ARG(ctx, nbr_block == 0 || nbr_block == 64 || nbr_block == 128 || nbr_block == 256,
                "rfd_system_set_tiles: row search must be 0 (cell blocks), 64, 128 or 256");
  s->force_fp = force_tile;
  s->nbr_block = nbr_block;
  s->once_off = !single_eval;
  s->gen++;
  return ensure_work(s);
}

int rfd_system_get_tiles(rfd_system *s, int *force_tile, int *nbr_block, int *single_eval) {
  if (!valid(s)) return RFD_E_ARG;
  if (force_tile) *force_tile = s->force_fp;
  if (nbr_block) *nbr_block = s->nbr_block;
  if (single_eval)
    *single_eval = !s->p.refmode && !s->fp64 && !s->once_off && s->work_fp == s->force_fp;
  return RFD_OK;
}

int rfd_system_autotune(rfd_system *s, int use_cache, rfd_tune_result *out) {
  if (!valid(s)) return RFD_E_ARG;
  rfd_ctx *ctx = s->ctx;
  RFD_CHECK_ARG(ctx, out && s->have_bonds, "rfd_system_autotune: bad arguments");
  RFD_CHECK_ARG(ctx, !s->fp64, "rfd_system_autotune: the FP64 path has one launch shape");
  *out = rfd_tune_result{};
  hipDeviceProp_t prop;
  RFD_HIP(ctx, hipGetDeviceProperties(&prop, ctx->device));
  int bucket = 0;
  while ((2u << bucket) <= s->N && bucket < 31) ++bucket;
  const std::string key = tune_key(prop.gcnArchName, bucket, s->S > 0);
  out->bucket = bucket;
  if (use_cache) {
    std::lock_guard<std::mutex> lk(tune_mu());
    auto it = tune_table().find(key);
    if (it != tune_table().end()) {
      out->force_tile = it->second.first;
      out->nbr_block = it->second.second;
      out->single_eval = it->second.third;
      out->cached = 1;
      return rfd_system_set_tiles(s, out->force_tile, out->nbr_block, out->single_eval);
    }
  }
  hipStream_t q = ctx->stream;
  if (!s->rows_valid) {
    int st = rebuild_rows(s);
    if (st) return st;
  }
  const bool timing = s->timing;
  s->timing = false;
  hipEvent_t e0, e1;
  RFD_HIP(ctx, hipEventCreate(&e0));
  RFD_HIP(ctx, hipEventCreate(&e1));
  // trial force passes write only F, M and the spare state buffer
  auto timed = [&](auto &&launch, double *ms) -> int {
    constexpr int kReps = 5;
    float t[kReps];
    int e = launch();  // warm-up
    for (int r = 0; r < kReps && !e; ++r) {
      RFD_HIP(ctx, hipEventRecord(e0, q));
      e = launch();
      RFD_HIP(ctx, hipEventRecord(e1, q));
      RFD_HIP(ctx, hipEventSynchronize(e1));
      RFD_HIP(ctx, hipEventElapsedTime(&t[r], e0, e1));
    }
    if (e) return e;
    std::sort(t, t + kReps);
    *ms = t[kReps / 2];
    return RFD_OK;
  };
  // force variants: tiles of 64, 32, 128 with single evaluation of in-tile
  // bonds (where the tile's halves fit), and 64 with every half evaluated
  static const int kTiles[4] = {64, 32, 128, 64};
  static const bool kOnceOff[4] = {false, false, false, true};
  static const int kBlocks[4] = {128, 64, 256, 0};
  const int fp0 = s->force_fp;
  const bool off0 = s->once_off;
  int st = RFD_OK, best_t = 0, best_b = 0;
  for (int v = 0; v < 4 && !st; ++v) {
    s->force_fp = kTiles[v];
    s->once_off = kOnceOff[v];
    st = ensure_work(s);
    if (!st) st = timed([&]() { return forces(s, nullptr, true, false); }, &out->force_ms[v]);
    if (!st && out->force_ms[v] < out->force_ms[best_t]) best_t = v;
  }
  s->force_fp = fp0;
  s->once_off = off0;
  if (!st) st = ensure_work(s);
  if (!st) st = bin_particles(s);
  // the row search writes into the spare row buffers, which the next rebuild
  // overwrites whole
  NbrDev d = make_nbr(s, s->p.skin);
  const int o = s->ccur, n = 1 - s->ccur;
  d.cj_old = s->cj[o];
  d.ccnt_old = s->ccnt[o];
  d.cO_old = s->cO[o];
  d.cj_new = s->cj[n];
  d.ccnt_new = s->ccnt[n];
  d.cO_new = s->cO[n];
  for (int v = 0; v < 4 && !st; ++v) {
    st = timed(
        [&]() {
          launch_rows(s, d, kBlocks[v]);
          return (int)hipGetLastError();
        },
        &out->nbr_ms[v]);
    if (!st && out->nbr_ms[v] < out->nbr_ms[best_b]) best_b = v;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  s->timing = timing;
  if (st) return st > 0 ? rfd_hip_status(ctx, (hipError_t)st, "rfd_system_autotune") : st;
  out->force_tile = kTiles[best_t];
  out->single_eval = kOnceOff[best_t] ? 0 : 1;
  out->nbr_block = kBlocks[best_b];
  {
    std::lock_guard<std::mutex> lk(tune_mu());
    tune_table()[key] = {out->force_tile, out->nbr_block, out->single_eval};
  }
  return rfd_system_set_tiles(s, out->force_tile, out->nbr_block, out->single_eval);
}

int rfd_tune_cache_save(const char *path) {
  if (!path) return RFD_E_ARG;
  FILE *f = fopen(path, "w");
  if (!f) return RFD_E_ARG;
  std::lock_guard<std::mutex> lk(tune_mu());
  for (const auto &kv : tune_table())
    fprintf(f, "%s %d %d %d\n", kv.first.c_str(), kv.second.first, kv.second.second,
            kv.second.third);
  fclose(f);
  return RFD_OK;
}

int rfd_tune_cache_load(const char *path) {
  if (!path) return RFD_E_ARG;
  FILE *f = fopen(path, "r");
  if (!f) return RFD_E_ARG;
  char k[160];
  int a = 0, b = 0, c = 0;
  std::lock_guard<std::mutex> lk(tune_mu());
  while (fscanf(f, "%159s %d %d %d", k, &a, &b, &c) == 4)
    if ((a == 32 || a == 64 || a == 128) && (b == 0 || b == 64 || b == 128 || b == 256) &&
        (c == 0 || c == 1))
      tune_table()[k] = {a, b, c};
  fclose(f);
  return RFD_OK;
}

int rfd_system_get_history(rfd_system *s, double *hist, uint32_t cap, uint32_t *n) {
  if (!valid(s)) return RFD_E_ARG;
  rfd_ctx *ctx = s->ctx;
  RFD_CHECK_ARG(ctx, n, "rfd_system_get_history: null count");
  rfd_loading L;
  int st = rfd_system_get_loading(s, &L);
  if (st) return st;
  *n = L.hist_len;
  const uint32_t m = L.hist_len < cap ? L.hist_len : cap;
  if (hist && m)
    RFD_HIP(ctx, hipMemcpy(hist, s->hist, sizeof(double) * 3 * m, hipMemcpyDeviceToHost));
  return RFD_OK;
}

int rfd_system_pack(rfd_system *s, const rfd_pack_params *pp, rfd_pack_result *out) {
  if (!valid(s)) return RFD_E_ARG;
  rfd_ctx *ctx = s->ctx;
  RFD_CHECK_ARG(ctx, pp && out, "rfd_system_pack: null argument");
  RFD_CHECK_ARG(ctx, s->have_bonds && s->Nb == 0 && !s->fp64,
                "rfd_system_pack: the packer runs on particles without bonds (set an empty "
                "bond list first)");
  RFD_CHECK_ARG(ctx, pp->s0 > 0 && pp->s0 <= 1 && pp->growth > 1 && pp->beta > 0 &&
                         pp->beta <= 1 && pp->momentum >= 0 && pp->momentum < 1 &&
                         pp->tol > 0 && pp->stage_iters > 0 &&
                         pp->final_iters > 0 && pp->radius > 0 && pp->z1 > pp->z0,
                "rfd_system_pack: bad parameters");
  *out = rfd_pack_result{};
  const auto t0 = std::chrono::steady_clock::now();
  hipStream_t q = ctx->stream;
  const uint32_t N = s->N;
  // final radii, kept while the stages scale them
  float *rfin = nullptr;
  uint32_t *d_max = nullptr;
  int st = dalloc(ctx, &rfin, N);
  if (!st) st = dalloc(ctx, &d_max, 1);
  if (st) return st;
  {
    std::vector<P4> hp(N);
    RFD_HIP(ctx, hipMemcpy(hp.data(), s->P, sizeof(P4) * N, hipMemcpyDeviceToHost));
    std::vector<float> hr(N);
    for (uint32_t i = 0; i < N; ++i) hr[i] = (float)hp[i].r;
    RFD_HIP(ctx, hipMemcpy(rfin, hr.data(), sizeof(float) * N, hipMemcpyHostToDevice));
  }
  const uint32_t nb = rfd_blocks(N, 256);
  k_zero_f4<<<nb, 256, 0, q>>>(s->V, nullptr, N);  // no last move yet
  double sc = pp->s0;
  uint32_t iters = 0, stages = 0;
  float mo = 0.f;
  auto relax = [&](uint32_t max_it) -> int {
    for (uint32_t k = 0; k < max_it; ++k) {
      int e = rebuild_rows(s);
      if (e) return e;
      k_fill_u32n<<<1, 64, 0, q>>>(d_max, 1, 0u);
      k_pack_disp<<<nb, 256, 0, q>>>(s->P, s->ccnt[s->ccur], s->cj[s->ccur], N, pp->beta, s->F,
                                     d_max);
      k_pack_apply<<<nb, 256, 0, q>>>(s->P, s->F, s->V, N, pp->momentum, pp->cx, pp->cy,
                                      pp->radius, pp->z0, pp->z1);
      RFD_HIP(ctx, hipGetLastError());
      ++iters;
      if (k % 8 == 7 || k + 1 == max_it) {
        uint32_t bits = 0;
        RFD_HIP(ctx, hipMemcpyAsync(&bits, d_max, sizeof(bits), hipMemcpyDeviceToHost, q));
        RFD_HIP(ctx, hipStreamSynchronize(q));
        memcpy(&mo, &bits, sizeof(mo));
        if (mo < pp->tol) break;
      }
    }
    return RFD_OK;
  };
  for (;;) {
    k_pack_scale<<<nb, 256, 0, q>>>(s->P, rfin, sc, N);
    RFD_HIP(ctx, hipGetLastError());
    const bool last = sc >= 1.0;
    st = relax(last ? pp->final_iters : pp->stage_iters);
    if (st) break;
    ++stages;
    if (last) break;
    sc = std::min(1.0, sc * pp->growth);
  }
  (void)hipStreamSynchronize(q);
  dfree(rfin);
  dfree(d_max);
  if (st) return st;
  // the packing leaves no motion and no loads behind
  k_zero_f4<<<nb, 256, 0, q>>>(s->V, s->W, N);
  k_zero_f4<<<nb, 256, 0, q>>>(s->F, s->M, N);
  RFD_HIP(ctx, hipGetLastError());
  RFD_HIP(ctx, hipStreamSynchronize(q));
  uint32_t fl = 0;
  RFD_HIP(ctx, hipMemcpy(&fl, ctx->d_flags, sizeof(fl), hipMemcpyDeviceToHost));
  out->iterations = iters;
  out->stages = stages;
  out->max_overlap = mo;
  out->converged = mo < pp->tol ? 1 : 0;
  out->row_overflow = (fl & RFD_FLAG_ROW_FULL) ? 1 : 0;
  out->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  s->rows_valid = false;
  s->gen++;
  return RFD_OK;
}

int rfd_system_bond_counts(rfd_system *s, uint32_t out[3]) {
  if (!valid(s)) return RFD_E_ARG;
  rfd_ctx *ctx = s->ctx;
  RFD_CHECK_ARG(ctx, out, "rfd_system_bond_counts: null output");
  std::vector<uint8_t> f(s->Nb);
  int st = rfd_system_get_bond_state(s, nullptr, f.data());
  if (st) return st;
  out[0] = out[1] = out[2] = 0;
  for (uint8_t v : f) out[v == kBond ? 0 : (v == kBroken ? 1 : 2)]++;
  return RFD_OK;
}

int rfd_system_contact_entries(rfd_system *s, uint64_t *entries) {
  if (!valid(s) || !entries) return RFD_E_ARG;
  rfd_ctx *ctx = s->ctx;
  *entries = 0;
  if (!s->rows_valid || !s->ccnt[s->ccur]) return RFD_OK;
  std::vector<uint32_t> c(s->N);
  RFD_HIP(ctx, hipMemcpyAsync(c.data(), s->ccnt[s->ccur], sizeof(uint32_t) * s->N,
                              hipMemcpyDeviceToHost, ctx->stream));
  RFD_HIP(ctx, hipStreamSynchronize(ctx->stream));
  uint64_t n = 0;
  for (uint32_t v : c) n += v;
  *entries = n;
  return RFD_OK;
}

int rfd_system_sizes(rfd_system *s, uint32_t *N, uint32_t *Nb) {
  if (!valid(s)) return RFD_E_ARG;
  if (N) *N = s->N;
  if (Nb) *Nb = s->Nb;
  return RFD_OK;
}

int rfd_system_device_arrays(rfd_system *s, double **P, float **V, float **W, float **F,
                             float **M) {
  if (!valid(s)) return RFD_E_ARG;
  if (P) *P = (double *)s->P;
  if (V) *V = (float *)s->V;
  if (W) *W = (float *)s->W;
  if (F) *F = (float *)s->F;
  if (M) *M = (float *)s->M;
  return RFD_OK;
}

int rfd_system_fire_load_step(rfd_system *s, double dstrain, const rfd_fire_params *fp,
                              const float *dU, const float *dTh, rfd_fire_result *out) {
  if (!valid(s)) return RFD_E_ARG;
  rfd_ctx *ctx = s->ctx;
  RFD_CHECK_ARG(ctx, fp && out && s->have_bonds, "rfd_system_fire_load_step: bad arguments");
  if (s->S)
    return rfd_set_error(ctx, RFD_E_STATE, "rfd_system_fire_load_step: batched specimens run "
                                           "explicit steps only");
  RFD_CHECK_ARG(ctx, fp->dt0 > 0 && fp->dt_max >= fp->dt0 && fp->ftol > 0 && fp->max_iter > 0,
                "rfd_system_fire_load_step: bad FIRE parameters");
  const auto t0 = std::chrono::steady_clock::now();
  int st = fire_alloc(s);
  if (st) return st;
  hipStream_t q = ctx->stream;
  // platen move (displacement control) on the host copy of the loading state
  rfd_loading L;
  st = rfd_system_get_loading(s, &L);
  if (st) return st;
  double dtop = 0.0;
  if (L.mode) {
    const double d = dstrain * L.size0;
    dtop = d;
    L.size += d;
    L.top += d;
    L.top_r = L.top - L.band;
    L.strain = (L.size - L.size0) / L.size0;
    L.v = 0.0;
    RFD_HIP(ctx, hipMemcpyAsync(s->ctl, &L, sizeof(L), hipMemcpyHostToDevice, q));
  }
  RFD_HIP(ctx, hipMemcpyAsync(s->P0, s->P, sizeof(P4) * s->N, hipMemcpyDeviceToDevice, q));
  FireDev f{};
  f.dt = f.dt_last = fp->dt0;
  f.alpha = f.alpha0 = fp->alpha0;
  f.dt_max = fp->dt_max;
  f.dt_min = 1e-3 * fp->dt0;
  f.f_inc = fp->f_inc;
  f.f_dec = fp->f_dec;
  f.f_alpha = fp->f_alpha;
  f.ftol = fp->ftol;
  f.n_delay = fp->n_delay;
  f.max_iter = fp->max_iter;
  f.first = 1;
  f.dt_f = (float)fp->dt0;
  f.dt_m = fp->dt0;
  if (!s->rows_valid || s->step_count % s->p.rebuild_every == 0) {
    st = rebuild_rows(s);
    if (st) return st;
  }
  // a prediction is judged by the residual it leaves against the plain start's
  // (two trial passes that change nothing the relaxation reads); the better
  // start is then taken for real
  double fmax_plain = -1.0, fmax_pred = -1.0;
  bool use_pred = false;
  if (dU || dTh) {
    st = begin_pass(s, f, nullptr, nullptr, dtop, true, &fmax_plain);
    if (!st) st = begin_pass(s, f, dU, dTh, dtop, true, &fmax_pred);
    if (st) return st;
    use_pred = fmax_pred <= fmax_plain;
  }
  st = begin_pass(s, f, use_pred ? dU : nullptr, use_pred ? dTh : nullptr, dtop, false, nullptr);
  if (st) return st;
  s->step_count++;
  s->forces_current = true;
  const uint32_t chunk = fp->check_every ? fp->check_every : 50;
  const FireDev f_start = f;
  auto relax = [&]() -> int {
    for (;;) {
      for (uint32_t k = 0; k < chunk; ++k) {
        int e = fire_iteration(s);
        if (e) return e;
      }
      RFD_HIP(ctx, hipMemcpyAsync(&f, s->fire, sizeof(f), hipMemcpyDeviceToHost, q));
      RFD_HIP(ctx, hipStreamSynchronize(q));
      if (f.done) return RFD_OK;
    }
  };
  st = relax();
  if (st) return st;
  const FireDev f_first = f;
  uint32_t iters = f.iter, breaks = 0, rounds = 0;
  // bonds are judged at equilibrium: a pass without increments breaks the
  // over-stressed ones, and the step relaxes again until none breaks (a
  // relaxation that did not converge is not judged)
  while (f.converged) {
    FireDev fc = f_start;
    fc.dt_f = 0.0f;
    fc.dt_m = 0.0;
    RFD_HIP(ctx, hipMemcpyAsync(s->fire, &fc, sizeof(fc), hipMemcpyHostToDevice, q));
    RFD_HIP(ctx, hipMemsetAsync(s->d_nbreak, 0, sizeof(uint32_t), q));
    st = forces(s, s->fire_dt, false, false);
    if (st) return st;
    uint32_t nb = 0;
    RFD_HIP(ctx, hipMemcpyAsync(&nb, s->d_nbreak, sizeof(nb), hipMemcpyDeviceToHost, q));
    RFD_HIP(ctx, hipStreamSynchronize(q));
    if (nb == 0 || rounds >= kMaxBreakRounds) break;
    breaks += nb;
    ++rounds;
    RFD_HIP(ctx, hipMemcpyAsync(s->fire, &f_start, sizeof(f_start), hipMemcpyHostToDevice, q));
    fire_rest(s, q, true);
    RFD_HIP(ctx, hipGetLastError());
    s->forces_current = true;  // the check pass left the forces of this state
    st = relax();
    if (st) return st;
    iters += f.iter;
  }
  s->forces_current = false;
  // the step's displacement and rotation, newest first; the relaxed state is
  // at rest
  s->dcur = 1 - s->dcur;
  if (s->fp64)
    k_fire64_step_disp<<<rfd_blocks(s->N, 256), 256, 0, q>>>(s->P, s->P0, s->Th64, s->N,
                                                              s->du[s->dcur], s->dth[s->dcur]);
  else
    k_fire_step_disp<<<rfd_blocks(s->N, 256), 256, 0, q>>>(s->P, s->P0, s->Th, s->N,
                                                            s->du[s->dcur], s->dth[s->dcur]);
  fire_rest(s, q, false);
  RFD_HIP(ctx, hipGetLastError());
  st = rfd_system_get_loading(s, &L);
  if (st) return st;
  out->iterations = iters;
  out->converged = f.converged;
  out->breaks = breaks;
  out->break_rounds = rounds;
  out->fmax0 = f_first.fmax0;
  out->fmax = f.fmax;
  out->mmax = f.mmax;
  out->sum_bottom = f.sum_bottom;
  out->sum_top = f.sum_top;
  out->strain = L.strain;
  out->seconds =
      std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  out->warm_used = use_pred ? 1 : 0;
  out->fmax0_plain = fmax_plain;
  out->fmax0_pred = fmax_pred;
  return RFD_OK;
}

int rfd_system_residual(rfd_system *s, double out[2]) {
  if (!valid(s)) return RFD_E_ARG;
  rfd_ctx *ctx = s->ctx;
  RFD_CHECK_ARG(ctx, out && s->have_bonds, "rfd_system_residual: bad arguments");
  if (s->fp64)
    return rfd_set_error(ctx, RFD_E_STATE, "rfd_system_residual: FP32 path only");
  int st = fire_alloc(s);
  if (st) return st;
  hipStream_t q = ctx->stream;
  FireDev f{};
  f.measure_only = 1;
  f.dt_f = 0.0f;  // no spring increments: the state is only re-expressed
  f.dt_m = 0.0;
  RFD_HIP(ctx, hipMemcpyAsync(s->fire, &f, sizeof(f), hipMemcpyHostToDevice, q));
  if (!s->rows_valid) {
    st = rebuild_rows(s);
    if (st) return st;
  }
  st = forces(s, s->fire_dt, false, true);
  if (st) return st;
  FireParticles fp = fire_particles(s);
  if (s->mode == 1)
    k_fire_measure<1><<<s->nblocks_pass, 256, 0, q>>>(fp);
  else if (s->mode == 2)
    k_fire_measure<2><<<s->nblocks_pass, 256, 0, q>>>(fp);
  else
    k_fire_measure<0><<<s->nblocks_pass, 256, 0, q>>>(fp);
  k_fire_control<<<1, 256, 0, q>>>(s->fire, s->fire_part, (int)s->nblocks_pass, s->ctl,
                                    nullptr);
  RFD_HIP(ctx, hipGetLastError());
  RFD_HIP(ctx, hipMemcpyAsync(&f, s->fire, sizeof(f), hipMemcpyDeviceToHost, q));
  RFD_HIP(ctx, hipStreamSynchronize(q));
  out[0] = f.fmax;
  out[1] = f.mmax;
  return RFD_OK;
}

int rfd_system_fire_buffers(rfd_system *s, float **du_last, float **du_prev, float **dth_last,
                            float **dth_prev) {
  if (!valid(s)) return RFD_E_ARG;
  int st = fire_alloc(s);
  if (st) return st;
  if (du_last) *du_last = s->du[s->dcur];
  if (du_prev) *du_prev = s->du[1 - s->dcur];
  if (dth_last) *dth_last = s->dth[s->dcur];
  if (dth_prev) *dth_prev = s->dth[1 - s->dcur];
  return RFD_OK;
}

int rfd_system_half_mirror(rfd_system *s, const uint32_t **hm) {
  if (!valid(s)) return RFD_E_ARG;
  RFD_CHECK_ARG(s->ctx, hm && s->have_bonds, "rfd_system_half_mirror: no bonds");
  *hm = s->hm;
  return RFD_OK;
}

int rfd_system_mpnn_inputs(rfd_system *s, double su, double sth, float *X, float *E) {
  if (!valid(s)) return RFD_E_ARG;
  rfd_ctx *ctx = s->ctx;
  RFD_CHECK_ARG(ctx, X && E && su > 0 && sth > 0 && s->have_bonds,
                "rfd_system_mpnn_inputs: bad arguments");
  int st = fire_alloc(s);
  if (st) return st;
  const int a = s->dcur, b = 1 - s->dcur;
  k_mpnn_inputs<<<rfd_blocks(s->N, 256), 256, 0, ctx->stream>>>(
      s->P, s->N, s->hoff, s->hr, s->brc, s->fl8, s->du[a], s->du[b], s->dth[a], s->dth[b],
      (float)(1.0 / su), (float)(1.0 / sth), (float)s->p.E, X, E);
  RFD_HIP(ctx, hipGetLastError());
  return RFD_OK;
}

int rfd_system_bond_graph(rfd_system *s, const uint32_t **hoff, const void **half,
                          const void **bond, const void **state, uint32_t *H) {
  if (!valid(s)) return RFD_E_ARG;
  if (hoff) *hoff = s->hoff;
  if (half) *half = s->hr;
  if (bond) *bond = s->brc;
  if (state) {
    int cur = 0;
    RFD_HIP(s->ctx, hipMemcpy(&cur, s->d_cur, sizeof(int), hipMemcpyDeviceToHost));
    *state = s->bst[cur];
  }
  if (H) *H = s->H;
  return RFD_OK;
}

}  // extern "C"

