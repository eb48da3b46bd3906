// rfd_lists.hip -- cell binning and the neighbour/bond slot lists
// (reference-layout tier).
//
// Cell binning replaces CUB's stable radix sort with a stable counting sort:
//   histogram (atomics) -> exclusive scan (one workgroup) -> scatter by atomic
//   cursor -> per-cell rank by id (one wave per cell).
// Ranking each cell's members by particle id restores index order inside a
// cell, which is exactly the order a stable sort of (cell, id) produces.
//
// List reconstruction finds its candidates in a second, finer binning (cell
// edge about half the search radius) and visits the ones in reach in the
// order of the reference's 8-cell scan, so it inserts the same partners into
// the same slots while testing about a quarter of the candidates.
#include <cmath>
#include <mutex>

#include "rfd_internal.h"

using namespace rfd;

namespace {

__global__ void k_fill_u32(uint32_t *p, size_t n, uint32_t v) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)blockDim.x * gridDim.x)
    p[i] = v;
}

// cell coordinates of particle i (md_calculate_cell.cu:39-47); false when the
// particle lies outside the grid
__device__ __forceinline__ bool cell_coords(const float *__restrict__ R, uint32_t N,
                                            uint32_t i, float inv_a, uint32_t Nx,
                                            uint32_t Ny, uint32_t Nz, uint32_t ix[3]) {
  float fx = floorf(R[i] * inv_a);
  float fy = floorf(R[i + N] * inv_a);
  float fz = floorf(R[i + 2 * (size_t)N] * inv_a);
  if (!(fx >= 0.0f) || !(fy >= 0.0f) || !(fz >= 0.0f)) return false;
  ix[0] = (uint32_t)fx;
  ix[1] = (uint32_t)fy;
  ix[2] = (uint32_t)fz;
  return ix[0] < Nx && ix[1] < Ny && ix[2] < Nz;
}

// Runs of equal keys in consecutive lanes of a wave (every lane active): the
// lane's offset in its run, the run's first lane, and at that lane the run's
// length.  Consecutive particles mostly share a cell, so a run takes one
// atomic where each lane took one.
struct WaveRun {
  uint32_t first, off, len;
};
__device__ __forceinline__ WaveRun wave_run(uint32_t key) {
  const uint32_t lane = threadIdx.x & 63;
  const uint32_t prev = __shfl_up(key, 1, 64);
  const uint64_t heads = __ballot(lane == 0 || prev != key);
  const uint64_t upto = heads & (lane == 63 ? ~0ull : ((2ull << lane) - 1ull));
  const uint32_t first = 63u - (uint32_t)__clzll(upto);
  const uint64_t after = heads & ~(lane == 63 ? ~0ull : ((2ull << lane) - 1ull));
  const uint32_t next = after ? (uint32_t)__ffsll((long long)after) - 1u : 64u;
  return WaveRun{first, lane - first, next - lane};
}

__global__ void k_cell_index(const float *__restrict__ R, uint32_t N, float inv_a,
                             uint32_t Nx, uint32_t Ny, uint32_t Nz,
                             uint32_t *__restrict__ CI, uint32_t *__restrict__ cnt,
                             uint32_t *__restrict__ flags) {
  // whole waves step together (the runs need every lane)
  const uint32_t lane = threadIdx.x & 63;
  for (uint32_t i0 = blockIdx.x * blockDim.x + (threadIdx.x & ~63u); i0 < N;
       i0 += grid_stride()) {
    const uint32_t i = i0 + lane;
    uint32_t c = kSentinel;
    if (i < N) {
      uint32_t ix[3];
      if (!cell_coords(R, N, i, inv_a, Nx, Ny, Nz, ix))
        atomicOr(flags, RFD_FLAG_CELL_RANGE);
      else
        c = ix[0] + ix[1] * Nx + ix[2] * Ny * Nx;
      CI[i] = c;
    }
    const WaveRun run = wave_run(c);
    if (run.off == 0 && c != kSentinel) atomicAdd(&cnt[c], run.len);
  }
}

// exclusive scan of the cell counts by one workgroup of 1024 threads;
// cursor[c] = start, pnC[c] = start for occupied cells, 0 for empty ones
constexpr int kScanThreads = 1024;
constexpr int kScanItems = 8;
constexpr uint32_t kScanChunk = kScanThreads * kScanItems;

// Exclusive scan of n words in chunks of kScanChunk, three launches:
// k_chunk_sums (one workgroup per chunk) -> k_scan_chunk_sums (one workgroup,
// carries over the chunk totals in place) -> k_scan_cells (one workgroup per
// chunk, starting from its chunk's carry). Integer sums: any order gives the
// same result.
__device__ __forceinline__ uint32_t block_sum(uint32_t v, uint32_t *s_sum) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  if (lane == 0) s_sum[wid] = v;
  __syncthreads();
  uint32_t t = 0;
  for (int w = 0; w < kScanThreads / 64; ++w) t += s_sum[w];
  return t;
}

// exclusive prefix of each thread's `local` inside the workgroup
__device__ __forceinline__ uint32_t block_prefix(uint32_t local, uint32_t *s_sum) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  uint32_t inc = local;
  for (int off = 1; off < 64; off <<= 1) {
    uint32_t t = __shfl_up(inc, off, 64);
    if (lane >= off) inc += t;
  }
  if (lane == 63) s_sum[wid] = inc;
  __syncthreads();
  uint32_t wave_off = 0;
  for (int w = 0; w < wid; ++w) wave_off += s_sum[w];
  return wave_off + inc - local;
}

__global__ void __launch_bounds__(kScanThreads)
k_chunk_sums(const uint32_t *__restrict__ cnt, uint32_t n, uint32_t *__restrict__ part) {
  __shared__ uint32_t s_sum[kScanThreads / 64];
  const uint32_t b0 = blockIdx.x * kScanChunk + threadIdx.x * kScanItems;
  uint32_t local = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) local += (b0 + k < n) ? cnt[b0 + k] : 0u;
  const uint32_t t = block_sum(local, s_sum);
  if (threadIdx.x == 0) part[blockIdx.x] = t;
}

// part[0..m) -> exclusive prefix sums, in place (one workgroup)
__global__ void __launch_bounds__(kScanThreads)
k_scan_chunk_sums(uint32_t *__restrict__ part, uint32_t m) {
  __shared__ uint32_t s_sum[kScanThreads / 64];
  __shared__ uint32_t s_carry;
  if (threadIdx.x == 0) s_carry = 0;
  __syncthreads();
  for (uint32_t base = 0; base < m; base += kScanThreads) {
    const uint32_t e = base + threadIdx.x;
    const uint32_t v = e < m ? part[e] : 0u;
    const uint32_t carry = s_carry;
    const uint32_t pre = block_prefix(v, s_sum);
    if (e < m) part[e] = carry + pre;
    __syncthreads();
    if (threadIdx.x == kScanThreads - 1) s_carry = carry + pre + v;
    __syncthreads();
  }
}

// counts in pnC[C..2C) -> starts in pnC[0..C) (0 for an empty cell, as the
// reference's cell table holds) and a copy of the starts as scatter cursors
__global__ void __launch_bounds__(kScanThreads)
k_scan_cells(uint32_t *__restrict__ pnC, uint32_t C, uint32_t *__restrict__ cursor,
             const uint32_t *__restrict__ part) {
  __shared__ uint32_t s_sum[kScanThreads / 64];
  const uint32_t *cnt = pnC + C;
  const uint32_t b0 = blockIdx.x * kScanChunk + threadIdx.x * kScanItems;
  uint32_t v[kScanItems], local = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    v[k] = (b0 + k < C) ? cnt[b0 + k] : 0u;
    local += v[k];
  }
  uint32_t run = (part ? part[blockIdx.x] : 0u) + block_prefix(local, s_sum);
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    if (b0 + k < C) {
      cursor[b0 + k] = run;
      pnC[b0 + k] = v[k] ? run : 0u;
    }
    run += v[k];
  }
}

__global__ void k_scatter(const uint32_t *__restrict__ CI, uint32_t N,
                          uint32_t *__restrict__ cursor, uint32_t *__restrict__ tmp) {
  const uint32_t lane = threadIdx.x & 63;
  for (uint32_t i0 = blockIdx.x * blockDim.x + (threadIdx.x & ~63u); i0 < N;
       i0 += grid_stride()) {
    const uint32_t i = i0 + lane;
    const uint32_t c = i < N ? CI[i] : kSentinel;
    const WaveRun run = wave_run(c);
    uint32_t p = 0;
    if (run.off == 0 && c != kSentinel) p = atomicAdd(&cursor[c], run.len);
    p = __shfl(p, (int)run.first, 64) + run.off;
    if (c != kSentinel) tmp[p] = i;
  }
}

// one wave per cell: rank of each member = number of members with smaller id
__global__ void k_rank_cells(const uint32_t *__restrict__ tmp,
                             const uint32_t *__restrict__ pnC, uint32_t C, uint32_t N,
                             uint32_t *__restrict__ CIs) {
  const uint32_t lane = threadIdx.x & 63;
  const uint32_t waves = blockDim.x >> 6;
  for (uint32_t c = blockIdx.x * waves + (threadIdx.x >> 6); c < C;
       c += gridDim.x * waves) {
    uint32_t n = pnC[C + c];
    if (n == 0) continue;
    uint32_t s0 = pnC[c];
    if (n <= 64) {
      // members read once, one per lane; each compared with the others'
      // ids read from their lanes
      const uint32_t id = lane < n ? tmp[s0 + lane] : 0u;
      uint32_t rank = 0;
      for (uint32_t t = 0; t < n; ++t) rank += (__shfl(id, (int)t, 64) < id) ? 1u : 0u;
      if (lane < n) {
        CIs[s0 + rank] = c;
        CIs[N + s0 + rank] = id;
      }
      continue;
    }
    for (uint32_t e = lane; e < n; e += 64) {
      uint32_t id = tmp[s0 + e];
      uint32_t rank = 0;
      for (uint32_t t = 0; t < n; ++t) rank += (tmp[s0 + t] < id) ? 1u : 0u;
      CIs[s0 + rank] = c;
      CIs[N + s0 + rank] = id;
    }
  }
}

// The 8 cells searched for particle i, in the reference's order
// (md_interaction_list_construct.cu:53-85): own cell, then the neighbours on
// the sides nearest i along x, y, z, xy, xz, yz, xyz.  A neighbour outside the
// grid is returned as kSentinel (searched as empty).
__device__ __forceinline__ bool search_cells(const float *__restrict__ R, uint32_t N,
                                             uint32_t i, float a, float inv_a,
                                             uint32_t Nx, uint32_t Ny, uint32_t Nz,
                                             uint32_t cells[8]) {
  uint32_t ix[3];
  if (!cell_coords(R, N, i, inv_a, Nx, Ny, Nz, ix)) return false;
  int sg[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    float mid = ((float)ix[d] + 0.5f) * a;
    sg[d] = (int)copysignf(1.0f, R[i + (size_t)d * N] - mid);
  }
  const uint32_t dim[3] = {Nx, Ny, Nz};
  const int masks[8] = {0, 1, 2, 4, 3, 5, 6, 7};
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    int64_t p[3];
    bool ok = true;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      p[d] = (int64_t)ix[d] + (((masks[q] >> d) & 1) ? sg[d] : 0);
      ok = ok && p[d] >= 0 && p[d] < (int64_t)dim[d];
    }
    cells[q] = ok ? (uint32_t)(p[0] + p[1] * Nx + p[2] * (int64_t)Nx * Ny) : kSentinel;
  }
  return true;
}

__device__ __forceinline__ float pair_r2(const float *__restrict__ R, uint32_t N,
                                         uint32_t i, uint32_t j) {
  float dx = R[j] - R[i];
  float dy = R[j + N] - R[i + N];
  float dz = R[j + 2 * (size_t)N] - R[i + 2 * (size_t)N];
  return dx * dx + dy * dy + dz * dz;
}

// Positions in cell order, one 16-byte record per cell-list entry: x, y, z and
// the particle id (as float bits). A candidate test then reads one record
// instead of the id and three position words in three arrays; the coordinates
// are copies, so distances round exactly as they do from R. Entries past the
// binned particles hold whatever the id slot held; they are never read.
__global__ void k_cell_positions(const float *__restrict__ R, uint32_t N,
                                 const uint32_t *__restrict__ CIs, float4 *__restrict__ Rc) {
  for (uint32_t s = blockIdx.x * blockDim.x + threadIdx.x; s < N; s += grid_stride()) {
    const uint32_t j = CIs[N + s];
    float4 v = make_float4(0.0f, 0.0f, 0.0f, __uint_as_float(kSentinel));
    if (j < N) v = make_float4(R[j], R[j + N], R[j + 2 * (size_t)N], __uint_as_float(j));
    Rc[s] = v;
  }
}

// |r_j - r_i|^2 in the same operation order as pair_r2, from a cell-order record
__device__ __forceinline__ float rec_r2(float4 p, float xi, float yi, float zi) {
  float dx = p.x - xi;
  float dy = p.y - yi;
  float dz = p.z - zi;
  return dx * dx + dy * dy + dz * dz;
}

__global__ void k_il_construct(const float *__restrict__ R, uint32_t N,
                               const float4 *__restrict__ Rc,
                               const uint32_t *__restrict__ pnC,
                               uint32_t *__restrict__ IL, uint32_t K, float a,
                               float inv_a, float aacut, uint32_t Nx, uint32_t Ny,
                               uint32_t Nz, uint32_t *__restrict__ flags) {
  const uint32_t C = Nx * Ny * Nz;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += grid_stride()) {
    uint32_t *row = IL + (size_t)i * K;
    uint32_t cells[8];
    uint32_t nk = 0;
    if (!search_cells(R, N, i, a, inv_a, Nx, Ny, Nz, cells)) {
      atomicOr(flags, RFD_FLAG_CELL_RANGE);
    } else {
      const float xi = R[i], yi = R[i + N], zi = R[i + 2 * (size_t)N];
      for (int q = 0; q < 8; ++q) {
        uint32_t c = cells[q];
        if (c == kSentinel) continue;
        uint32_t s0 = pnC[c], n = pnC[C + c];
        for (uint32_t t = 0; t < n; ++t) {
          const float4 pj = Rc[s0 + t];
          const uint32_t j = __float_as_uint(pj.w);
          float rr = rec_r2(pj, xi, yi, zi);
          if (rr < aacut && j != i) {
            if (nk < K) row[nk] = j;
            ++nk;
          }
        }
      }
      if (nk > K) atomicOr(flags, RFD_FLAG_ROW_FULL);
    }
    for (uint32_t k = nk; k < K; ++k) row[k] = kSentinel;
  }
}

// Rows of one workgroup are staged in LDS (coalesced in and out); each thread
// then tests every new candidate against its own row.
constexpr int kRecThreads = 128;
constexpr int kRecMaxK = 64;

// LDS row pitch in words: K rounded up to 16 bytes, then an odd multiple of 4
// words, so that the 16-byte reads of 16 consecutive lanes fall in distinct
// banks. Padding words hold the sentinel, which no candidate id equals.
__host__ __device__ constexpr uint32_t rec_pitch(uint32_t K) { return ((K + 3u) & ~3u) | 4u; }
__global__ void __launch_bounds__(kRecThreads)
k_il_reconstruct(const float *__restrict__ R, uint32_t N, const float4 *__restrict__ Rc,
                 const uint32_t *__restrict__ pnC, uint32_t *__restrict__ IL,
                 uint8_t *__restrict__ ILtype, uint32_t K, float a, float inv_a,
                 float aacut, uint32_t Nx, uint32_t Ny, uint32_t Nz, int quirk_p0,
                 uint32_t *__restrict__ flags) {
  // the block's rows, rec_pitch(K) words apart (sized by K at launch, so that
  // short rows leave room for more resident blocks)
  extern __shared__ uint4 s_row4[];
  uint32_t *s_row = reinterpret_cast<uint32_t *>(s_row4);
  const uint32_t C = Nx * Ny * Nz;
  const uint32_t p0 = blockIdx.x * kRecThreads;
  const uint32_t np = min((uint32_t)kRecThreads, N - p0);
  const uint32_t pitch = rec_pitch(K);
  const size_t g0 = (size_t)p0 * K;
  for (uint32_t e = threadIdx.x; e < np * pitch; e += kRecThreads) {
    const uint32_t r = e / pitch, k = e % pitch;
    s_row[e] = k < K ? IL[g0 + (size_t)r * K + k] : kSentinel;
  }
  __syncthreads();
  const uint32_t t = threadIdx.x;
  if (t < np) {
    const uint32_t i = p0 + t;
    uint32_t *row = s_row + t * pitch;
    const uint4 *row4 = s_row4 + t * (pitch / 4);
    uint32_t cells[8];
    if (!search_cells(R, N, i, a, inv_a, Nx, Ny, Nz, cells)) {
      atomicOr(flags, RFD_FLAG_CELL_RANGE);
    } else {
      const float xi = R[i], yi = R[i + N], zi = R[i + 2 * (size_t)N];
      for (int q = 0; q < 8; ++q) {
        uint32_t c = cells[q];
        if (c == kSentinel) continue;
        uint32_t s0 = pnC[c], n = pnC[C + c];
        for (uint32_t u = 0; u < n; ++u) {
          const float4 pj = Rc[s0 + u];
          const uint32_t j = __float_as_uint(pj.w);
          float rr = rec_r2(pj, xi, yi, zi);
          if (rr < aacut && (double)rr > 1e-10) {
            // membership over the whole row, without an early exit: the
            // reads are independent and issue back to back
            bool found = false;
            for (uint32_t w = 0; w < pitch / 4; ++w) {
              const uint4 v = row4[w];
              found |= (v.x == j) | (v.y == j) | (v.z == j) | (v.w == j);
            }
            if (!found) {
              // the first free slot (the second for particle 0 under the
              // particle-0 rule), searched only when the candidate is new
              int freeslot = -1;
              for (uint32_t k = 0; k < K; ++k) {
                if (row[k] == kSentinel) {
                  if (freeslot < 0) freeslot = (int)k;
                  else if (quirk_p0 && i == 0 && freeslot == 0) freeslot = (int)k;
                }
              }
              if (freeslot < 0) {
                atomicOr(flags, RFD_FLAG_ROW_FULL);
              } else {
                row[freeslot] = j;
                ILtype[(size_t)i * K + freeslot] = 0;
              }
            }
          }
        }
      }
    }
  }
  __syncthreads();
  for (uint32_t e = threadIdx.x; e < np * pitch; e += kRecThreads) {
    const uint32_t r = e / pitch, k = e % pitch;
    if (k < K) IL[g0 + (size_t)r * K + k] = s_row[e];
  }
}

__device__ __forceinline__ void zero3(float *p, size_t s) {
  p[3 * s] = 0.0f;
  p[3 * s + 1] = 0.0f;
  p[3 * s + 2] = 0.0f;
}

__global__ void k_bonds_construct(const float *__restrict__ R, const uint32_t *__restrict__ IL,
                                  uint8_t *__restrict__ ILtype, float *__restrict__ iL,
                                  float *__restrict__ rij, float *__restrict__ Oijt,
                                  float *__restrict__ Fijn, float *__restrict__ Fijt,
                                  float *__restrict__ Mijn, float *__restrict__ Mijt,
                                  float *__restrict__ Mijadd, uint32_t N, uint32_t K,
                                  float bdd) {
  const uint32_t S = N * K;
  for (uint32_t s = blockIdx.x * blockDim.x + threadIdx.x; s < S; s += grid_stride()) {
    uint32_t j = IL[s];
    if (j == kSentinel) continue;
    uint32_t i = s / K;
    float rx = R[j] - R[i];
    float ry = R[j + N] - R[i + N];
    float rz = R[j + 2 * (size_t)N] - R[i + 2 * (size_t)N];
    float rm2 = rx * rx + ry * ry + rz * rz;
    if (rm2 < bdd && (double)rm2 > 1e-18) {
      iL[s] = rsqrt_rn(rm2);
      ILtype[s] = 1;
    } else {
      ILtype[s] = 0;
    }
    rij[3 * (size_t)s] = 0.0f * rx;
    rij[3 * (size_t)s + 1] = 0.0f * ry;
    rij[3 * (size_t)s + 2] = 0.0f * rz;
    zero3(Oijt, s);
    zero3(Fijn, s);
    zero3(Fijt, s);
    zero3(Mijn, s);
    zero3(Mijt, s);
    zero3(Mijadd, s);
  }
}

__global__ void k_delete_far(const float *__restrict__ R, uint32_t N, uint32_t *__restrict__ IL,
                             uint8_t *__restrict__ ILtype, float *__restrict__ rij,
                             float *__restrict__ Oijt, float *__restrict__ Fijn,
                             float *__restrict__ Fijt, float *__restrict__ Mijn,
                             float *__restrict__ Mijt, float *__restrict__ Mijadd,
                             uint32_t K, float aa_farcut) {
  const uint32_t S = N * K;
  for (uint32_t s = blockIdx.x * blockDim.x + threadIdx.x; s < S; s += grid_stride()) {
    uint32_t j = IL[s];
    if (j == kSentinel) continue;
    uint32_t i = s / K;
    float rr = pair_r2(R, N, i, j);
    if (rr > aa_farcut || (double)rr < 1e-10) {
      IL[s] = kSentinel;
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

// The same pruning four slots per thread (K a multiple of 4, IL 16-byte
// aligned): one 16-byte load of the slot ids and one load of the particle's
// position serve four distance tests, whose gathers are then in flight
// together. Every slot is decided on its own, as above.
__global__ void k_delete_far4(const float *__restrict__ R, uint32_t N, uint32_t *__restrict__ IL,
                              uint8_t *__restrict__ ILtype, float *__restrict__ rij,
                              float *__restrict__ Oijt, float *__restrict__ Fijn,
                              float *__restrict__ Fijt, float *__restrict__ Mijn,
                              float *__restrict__ Mijt, float *__restrict__ Mijadd,
                              uint32_t K, float aa_farcut) {
  const uint32_t Q = N * (K / 4);
  const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= Q) return;
  const uint4 v = reinterpret_cast<const uint4 *>(IL)[g];
  if ((v.x & v.y & v.z & v.w) == kSentinel) return;
  const uint32_t s0 = 4 * g, i = s0 / K;
  const float xi = R[i], yi = R[i + N], zi = R[i + 2 * (size_t)N];
  const uint32_t js[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    const uint32_t j = js[u];
    if (j == kSentinel) continue;
    const float dx = R[j] - xi;
    const float dy = R[j + N] - yi;
    const float dz = R[j + 2 * (size_t)N] - zi;
    const float rr = dx * dx + dy * dy + dz * dz;
    if (rr > aa_farcut || (double)rr < 1e-10) {
      const uint32_t s = s0 + u;
      IL[s] = kSentinel;
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

bool slots_fit(uint32_t N, uint32_t K) { return (uint64_t)N * K < 0xFFFFFFFFull; }


// ---- fine-grid candidate search for the list reconstruction
// A second binning of the particles, with a cell edge of 1.01 / h times the
// search radius sqrt(aacut), puts every partner that can pass the FP32 range
// test in the (2h + 1)^3 fine cells around the particle: along an axis, a
// partner h + 1 fine cells away is more than h + 1 - 0.01 fine cells, over
// 1.0001 radii, away, beyond what the FP32 distance and the FP32 fine
// coordinates (below 40,000 cells) can round across. The grid starts h cells
// below the coarse grid and ends h cells beyond it, so every partner of a
// particle in the coarse grid is binned. Of the candidates that pass the range test, those that the
// caller's cell lists hold in one of the particle's 8 search cells are kept
// and visited in the order of the 8-cell search: by search cell, then by the
// candidate's place in that cell's list (CIs), whatever order the caller's
// lists hold. So the same partners are inserted, in the same order, as by the
// scan of the 8 cells, from about 100 candidates per particle instead of 8
// coarse cells' worth.
constexpr uint32_t kFineMaxDim = 40000;
constexpr int kFineHalf = 2;       // fine cells searched on each side
constexpr int kFineFound = 32;     // in-range partners held (sorted) per particle
constexpr uint32_t kFineIdBits = 29;  // key = search cell << 29 | place in CIs

struct FineGrid {
  float af, inv_af;     // fine cell edge and its inverse
  float off;            // h * af: coordinate x maps to cell floor((x + off) / af)
  uint32_t nx, ny, nz;  // cells per axis
  int h;                // cells searched on each side (edge = 1.01 radius / h)
};

__device__ __forceinline__ bool fine_coords(const FineGrid &f, float x, float y, float z,
                                            uint32_t c[3]) {
  const float fx = floorf((x + f.off) * f.inv_af);
  const float fy = floorf((y + f.off) * f.inv_af);
  const float fz = floorf((z + f.off) * f.inv_af);
  if (!(fx >= 0.0f && fx < (float)f.nx) || !(fy >= 0.0f && fy < (float)f.ny) ||
      !(fz >= 0.0f && fz < (float)f.nz))
    return false;
  c[0] = (uint32_t)fx;
  c[1] = (uint32_t)fy;
  c[2] = (uint32_t)fz;
  return true;
}

// the place of every particle in the caller's lists, j = CIs[N + placeOf[j]]
// (particles the lists do not hold keep the sentinel)
__global__ void k_place_of(const uint32_t *__restrict__ CIs, uint32_t N,
                           uint32_t *__restrict__ placeOf) {
  for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < N; p += grid_stride()) {
    const uint32_t j = CIs[N + p];
    if (j < N) placeOf[j] = p;
  }
}

__global__ void k_fine_index(const float *__restrict__ R, uint32_t N, FineGrid f,
                             uint32_t *__restrict__ fci, uint32_t *__restrict__ cnt) {
  const uint32_t lane = threadIdx.x & 63;
  for (uint32_t i0 = blockIdx.x * blockDim.x + (threadIdx.x & ~63u); i0 < N;
       i0 += grid_stride()) {
    const uint32_t i = i0 + lane;
    uint32_t c = kSentinel;
    if (i < N) {
      uint32_t c3[3];
      if (fine_coords(f, R[i], R[i + N], R[i + 2 * (size_t)N], c3))
        c = c3[0] + f.nx * (c3[1] + f.ny * c3[2]);
      fci[i] = c;
    }
    const WaveRun run = wave_run(c);
    if (run.off == 0 && c != kSentinel) atomicAdd(&cnt[c], run.len);
  }
}

// exclusive prefix sums of cnt[0..n) into excl[0..n], excl[n] = total (one
// workgroup per chunk, from the chunk carries of k_scan_chunk_sums)
__global__ void __launch_bounds__(kScanThreads)
k_scan_excl(const uint32_t *__restrict__ cnt, uint32_t n, uint32_t *__restrict__ excl,
            const uint32_t *__restrict__ part) {
  __shared__ uint32_t s_sum[kScanThreads / 64];
  const uint32_t b0 = blockIdx.x * kScanChunk + threadIdx.x * kScanItems;
  uint32_t v[kScanItems], local = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    v[k] = (b0 + k < n) ? cnt[b0 + k] : 0u;
    local += v[k];
  }
  uint32_t run = (part ? part[blockIdx.x] : 0u) + block_prefix(local, s_sum);
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    if (b0 + k < n) {
      excl[b0 + k] = run;
      if (b0 + k == n - 1) excl[n] = run + v[k];
    }
    run += v[k];
  }
}

// positions in fine-cell order, (x, y, z, bits of the particle's place in the
// caller's lists); cnt cleared beforehand
__global__ void k_fine_scatter(const float *__restrict__ R, uint32_t N,
                               const uint32_t *__restrict__ fci,
                               const uint32_t *__restrict__ placeOf,
                               const uint32_t *__restrict__ excl, uint32_t *__restrict__ cnt,
                               float4 *__restrict__ Rf) {
  const uint32_t lane = threadIdx.x & 63;
  for (uint32_t i0 = blockIdx.x * blockDim.x + (threadIdx.x & ~63u); i0 < N;
       i0 += grid_stride()) {
    const uint32_t i = i0 + lane;
    const uint32_t c = i < N ? fci[i] : kSentinel;
    const WaveRun run = wave_run(c);
    uint32_t p = 0;
    if (run.off == 0 && c != kSentinel) p = excl[c] + atomicAdd(&cnt[c], run.len);
    p = __shfl(p, (int)run.first, 64) + run.off;
    if (c != kSentinel)
      Rf[p] = make_float4(R[i], R[i + N], R[i + 2 * (size_t)N], __uint_as_float(placeOf[i]));
  }
}

__global__ void __launch_bounds__(kRecThreads)
k_il_reconstruct_fine(const float *__restrict__ R, uint32_t N, const uint32_t *__restrict__ CIs,
                      const uint32_t *__restrict__ pnC, const float4 *__restrict__ Rf,
                      const uint32_t *__restrict__ excl, FineGrid f,
                      uint32_t *__restrict__ IL, uint8_t *__restrict__ ILtype,
                      uint32_t K, float a, float inv_a, float aacut, uint32_t Nx, uint32_t Ny,
                      uint32_t Nz, int quirk_p0, uint32_t *__restrict__ flags,
                      uint32_t nblk, uint32_t per) {
  extern __shared__ uint4 s_row4[];
  uint32_t *s_row = reinterpret_cast<uint32_t *>(s_row4);
  const uint32_t C = Nx * Ny * Nz;
  const uint32_t blk = xcd_block(per);
  if (blk >= nblk) return;
  const uint32_t p0 = blk * kRecThreads;
  const uint32_t np = min((uint32_t)kRecThreads, N - p0);
  const uint32_t pitch = rec_pitch(K);
  const size_t g0 = (size_t)p0 * K;
  for (uint32_t e = threadIdx.x; e < np * pitch; e += kRecThreads) {
    const uint32_t r = e / pitch, k = e % pitch;
    s_row[e] = k < K ? IL[g0 + (size_t)r * K + k] : kSentinel;
  }
  __syncthreads();
  const uint32_t t = threadIdx.x;
  if (t < np) {
    const uint32_t i = p0 + t;
    uint32_t *row = s_row + t * pitch;
    const uint4 *row4 = s_row4 + t * (pitch / 4);
    // a partner that passed the range test: insert it unless the row holds it
    auto consider = [&](uint32_t j) {
      bool found = false;
      for (uint32_t w = 0; w < pitch / 4; ++w) {
        const uint4 v = row4[w];
        found |= (v.x == j) | (v.y == j) | (v.z == j) | (v.w == j);
      }
      if (found) return;
      int freeslot = -1;
      for (uint32_t k = 0; k < K; ++k) {
        if (row[k] == kSentinel) {
          if (freeslot < 0) freeslot = (int)k;
          else if (quirk_p0 && i == 0 && freeslot == 0) freeslot = (int)k;
        }
      }
      if (freeslot < 0) {
        atomicOr(flags, RFD_FLAG_ROW_FULL);
      } else {
        row[freeslot] = j;
        ILtype[(size_t)i * K + freeslot] = 0;
      }
    };
    uint32_t cells[8];
    if (!search_cells(R, N, i, a, inv_a, Nx, Ny, Nz, cells)) {
      atomicOr(flags, RFD_FLAG_CELL_RANGE);
    } else {
      const float xi = R[i], yi = R[i + N], zi = R[i + 2 * (size_t)N];
      // the search cells' ranges of places in the lists (empty when outside)
      uint32_t cfirst[8], ccount[8];
#pragma unroll
      for (int qq = 0; qq < 8; ++qq) {
        const bool in = cells[qq] != kSentinel;
        cfirst[qq] = in ? pnC[cells[qq]] : 0u;
        ccount[qq] = in ? pnC[C + cells[qq]] : 0u;
      }
      uint32_t key[kFineFound];  // ascending; unused entries hold 0xFFFFFFFF
#pragma unroll
      for (int u = 0; u < kFineFound; ++u) key[u] = 0xFFFFFFFFu;
      uint32_t nf = 0;
      uint32_t fc[3];
      if (fine_coords(f, xi, yi, zi, fc)) {
        // the grid is built with f.h == kFineHalf
        constexpr int kRows = 2 * kFineHalf + 1;
        const uint32_t hu = (uint32_t)kFineHalf;
        const uint32_t xlo = fc[0] > hu ? fc[0] - hu : 0;
        const uint32_t xhi = fc[0] + hu < f.nx ? fc[0] + hu : f.nx - 1;
        // a candidate within range whose place lies in one of the 8 search
        // cells' ranges goes into the sorted keys
        auto test = [&](const float4 &pj) {
          const float rr = rec_r2(pj, xi, yi, zi);
          if (rr < aacut && (double)rr > 1e-10) {
            const uint32_t pl = __float_as_uint(pj.w);
            uint32_t q = 8;
#pragma unroll
            for (uint32_t qq = 0; qq < 8; ++qq)
              if (pl - cfirst[qq] < ccount[qq]) q = qq;
            if (q < 8) {
              uint32_t x = (q << kFineIdBits) | pl;
#pragma unroll
              for (int u = 0; u < kFineFound; ++u) {
                const uint32_t lo2 = min(key[u], x), hi2 = max(key[u], x);
                key[u] = lo2;
                x = hi2;
              }
              ++nf;
            }
          }
        };
        for (int dz = -kFineHalf; dz <= kFineHalf; ++dz) {
          const int z = (int)fc[2] + dz;
          if (z < 0 || z >= (int)f.nz) continue;
          // the plane's x-row ranges, read together
          uint32_t rlo[kRows], rhi[kRows];
#pragma unroll
          for (int r = 0; r < kRows; ++r) {
            const int y = (int)fc[1] + r - kFineHalf;
            rlo[r] = rhi[r] = 0;
            if (y >= 0 && y < (int)f.ny) {
              const uint32_t base = f.nx * ((uint32_t)y + f.ny * (uint32_t)z);
              rlo[r] = excl[base + xlo];
              rhi[r] = excl[base + xhi + 1];
            }
          }
          // the plane's candidates as one run v = 0 .. tot - 1 over its rows,
          // kBatch at a time with their loads in flight together (v is
          // clamped to the run, so every load reads a candidate of a row)
          uint32_t cum[kRows + 1];
          cum[0] = 0;
#pragma unroll
          for (int r = 0; r < kRows; ++r) cum[r + 1] = cum[r] + (rhi[r] - rlo[r]);
          const uint32_t tot = cum[kRows];
          constexpr uint32_t kBatch = 8;
          for (uint32_t v = 0; v < tot; v += kBatch) {
            float4 pj[kBatch];
#pragma unroll
            for (uint32_t u = 0; u < kBatch; ++u) {
              const uint32_t vv = min(v + u, tot - 1);
              uint32_t e = rlo[0] + vv;
#pragma unroll
              for (int r = 1; r < kRows; ++r)
                if (vv >= cum[r]) e = rlo[r] + (vv - cum[r]);
              pj[u] = Rf[e];
            }
#pragma unroll
            for (uint32_t u = 0; u < kBatch; ++u)
              if (v + u < tot) test(pj[u]);
          }
        }
      } else {
        nf = kFineFound + 1;
      }
      if (nf <= (uint32_t)kFineFound) {
        // the partners' ids from their places in the lists, all loads in
        // flight at once, then inserted in key order
#pragma unroll
        for (int u = 0; u < kFineFound; ++u)
          if ((uint32_t)u < nf) key[u] = CIs[N + (key[u] & ((1u << kFineIdBits) - 1u))];
        for (uint32_t n = 0; n < nf; ++n) {
          const uint32_t j = key[0];
#pragma unroll
          for (int u = 0; u + 1 < kFineFound; ++u) key[u] = key[u + 1];
          consider(j);
        }
      } else {
        // more partners than the sorted list holds: the 8-cell scan itself
        for (int q = 0; q < 8; ++q) {
          const uint32_t c = cells[q];
          if (c == kSentinel) continue;
          const uint32_t s0 = pnC[c], n = pnC[C + c];
          for (uint32_t u = 0; u < n; ++u) {
            const uint32_t j = CIs[N + s0 + u];
            const float rr = pair_r2(R, N, i, j);
            if (rr < aacut && (double)rr > 1e-10) consider(j);
          }
        }
      }
    }
  }
  __syncthreads();
  for (uint32_t e = threadIdx.x; e < np * pitch; e += kRecThreads) {
    const uint32_t r = e / pitch, k = e % pitch;
    if (k < K) IL[g0 + (size_t)r * K + k] = s_row[e];
  }
}

// Workspace words of one cell binning: scatter cursors (C), the unranked cell
// lists (N) and the scan's chunk sums.
size_t cell_ws_words(uint32_t N, uint32_t C) {
  return (size_t)C + N + ((size_t)C + kScanChunk - 1) / kScanChunk;
}

// Fixed workspace of rfd_cell_distribution_noalloc: 64 MiB in the code
// object's own device memory, one copy per GPU, held until the process ends.
// Host code must not pass this name to a kernel (on the host it is only a
// stand-in); it uses the device address kept per device in g_fixed_ptr.
constexpr size_t kFixedWords = (size_t)16 << 20;
static __device__ __align__(256) uint32_t scratch[kFixedWords];

constexpr int kMaxDevices = 64;
std::mutex g_fixed_mu;
void *g_fixed_ptr[kMaxDevices] = {};

// Cell binning with the workspace `ws` (cell_ws_words(N, C) words).  Every
// workspace word a kernel reads was written earlier in the same call (the
// cursors by the scan, the lists by the scatter, the chunk sums by the first
// scan pass), so the workspace needs no clearing.
int cell_distribution_ws(rfd_ctx *ctx, uint32_t *ws, const float *R, uint32_t N, float inv_a,
                         uint32_t Nx, uint32_t Ny, uint32_t Nz, uint32_t *CI, uint32_t *CIs,
                         uint32_t *pnC) {
  const uint32_t C = Nx * Ny * Nz;
  const uint32_t chunks = (uint32_t)(((size_t)C + kScanChunk - 1) / kScanChunk);
  uint32_t *cursor = ws, *tmp = cursor + C, *part = tmp + N;
  hipStream_t s = ctx->stream;
  k_fill_u32<<<rfd_blocks(2 * (size_t)C, 256, 4096), 256, 0, s>>>(pnC, 2 * (size_t)C, 0u);
  k_cell_index<<<rfd_blocks(N, 256), 256, 0, s>>>(R, N, inv_a, Nx, Ny, Nz, CI, pnC + C,
                                                  ctx->d_flags);
  if (chunks > 1) {
    k_chunk_sums<<<chunks, kScanThreads, 0, s>>>(pnC + C, C, part);
    k_scan_chunk_sums<<<1, kScanThreads, 0, s>>>(part, chunks);
  }
  // one chunk: its carry is zero (no chunk sums to form)
  k_scan_cells<<<chunks, kScanThreads, 0, s>>>(pnC, C, cursor, chunks > 1 ? part : nullptr);
  k_scatter<<<rfd_blocks(N, 256), 256, 0, s>>>(CI, N, cursor, tmp);
  k_rank_cells<<<rfd_blocks(C, 4), 256, 0, s>>>(tmp, pnC, C, N, CIs);
  RFD_LAUNCHED(ctx, "rfd_cell_distribution");
  return RFD_OK;
}

}  // namespace

// Device address of the fixed workspace on ctx's device, looked up once per
// device (with that device current) and kept in g_fixed_ptr; called by
// rfd_ctx_create.
int rfd_fixed_scratch_attach(rfd_ctx *ctx) {
  if (ctx->device < 0 || ctx->device >= kMaxDevices)
    return rfd_set_error(ctx, RFD_E_ARG, "device %d beyond the fixed-workspace table",
                         ctx->device);
  std::lock_guard<std::mutex> lock(g_fixed_mu);
  if (!g_fixed_ptr[ctx->device]) {
    void *p = nullptr;
    RFD_HIP(ctx, hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_fixed_ptr[ctx->device] = p;
  }
  ctx->fixed_scratch = g_fixed_ptr[ctx->device];
  return RFD_OK;
}

extern "C" {

int rfd_cell_distribution(rfd_ctx *ctx, const float *R, uint32_t N, float inv_a,
                          uint32_t Nx, uint32_t Ny, uint32_t Nz, uint32_t *CI,
                          uint32_t *CIs, uint32_t *pnC) {
  if (!ctx) return RFD_E_ARG;
  RFD_CHECK_ARG(ctx, R && CI && CIs && pnC && N > 0, "rfd_cell_distribution: null buffer or N=0");
  RFD_CHECK_ARG(ctx, Nx && Ny && Nz && (uint64_t)Nx * Ny * Nz < 0x7FFFFFFFull,
                "rfd_cell_distribution: bad grid %u x %u x %u", Nx, Ny, Nz);
  void *ws = nullptr;
  int st = rfd_scratch(ctx, sizeof(uint32_t) * cell_ws_words(N, Nx * Ny * Nz), &ws);
  if (st) return st;
  return cell_distribution_ws(ctx, (uint32_t *)ws, R, N, inv_a, Nx, Ny, Nz, CI, CIs, pnC);
}

int rfd_cell_distribution_noalloc(rfd_ctx *ctx, const float *R, uint32_t N, float inv_a,
                                  uint32_t Nx, uint32_t Ny, uint32_t Nz, uint32_t *CI,
                                  uint32_t *CIs, uint32_t *pnC) {
  if (!ctx) return RFD_E_ARG;
  RFD_CHECK_ARG(ctx, R && CI && CIs && pnC && N > 0,
                "rfd_cell_distribution_noalloc: null buffer or N=0");
  RFD_CHECK_ARG(ctx, Nx && Ny && Nz && (uint64_t)Nx * Ny * Nz < 0x7FFFFFFFull,
                "rfd_cell_distribution_noalloc: bad grid %u x %u x %u", Nx, Ny, Nz);
  const size_t words = cell_ws_words(N, Nx * Ny * Nz);
  if (words > kFixedWords)
    return rfd_set_error(ctx, RFD_E_CAPACITY,
                         "rfd_cell_distribution_noalloc: N + cells = %zu words of workspace, "
                         "the fixed workspace holds %zu (N + cells + ceil(cells / 8192) <= "
                         "16777216)",
                         words, kFixedWords);
  if (!ctx->fixed_scratch)
    return rfd_set_error(ctx, RFD_E_STATE, "rfd_cell_distribution_noalloc: no fixed workspace");
  return cell_distribution_ws(ctx, (uint32_t *)ctx->fixed_scratch, R, N, inv_a, Nx, Ny, Nz, CI,
                              CIs, pnC);
}

int rfd_il_construct(rfd_ctx *ctx, const float *R, uint32_t N, const uint32_t *CIs,
                     const uint32_t *pnC, uint32_t *IL, uint32_t K, float a,
                     float inv_a, float aacut, uint32_t Nx, uint32_t Ny, uint32_t Nz) {
  if (!ctx) return RFD_E_ARG;
  RFD_CHECK_ARG(ctx, R && CIs && pnC && IL && N > 0 && K > 0 && slots_fit(N, K),
                "rfd_il_construct: bad arguments");
  void *ws = nullptr;
  int st = rfd_scratch(ctx, sizeof(float4) * (size_t)N, &ws);
  if (st) return st;
  float4 *Rc = (float4 *)ws;
  k_cell_positions<<<rfd_blocks(N, 256), 256, 0, ctx->stream>>>(R, N, CIs, Rc);
  k_il_construct<<<rfd_blocks(N, 128), 128, 0, ctx->stream>>>(
      R, N, Rc, pnC, IL, K, a, inv_a, aacut, Nx, Ny, Nz, ctx->d_flags);
  RFD_LAUNCHED(ctx, "rfd_il_construct");
  return RFD_OK;
}

int rfd_il_reconstruct_bpm(rfd_ctx *ctx, const float *R, uint32_t N,
                           const uint32_t *CIs, const uint32_t *pnC, uint32_t *IL,
                           uint8_t *ILtype, uint32_t K, float a, float inv_a,
                           float aacut, uint32_t Nx, uint32_t Ny, uint32_t Nz,
                           int quirk_p0) {
  if (!ctx) return RFD_E_ARG;
  RFD_CHECK_ARG(ctx, R && CIs && pnC && IL && ILtype && N > 0 && K > 0 && slots_fit(N, K),
                "rfd_il_reconstruct_bpm: bad arguments");
  RFD_CHECK_ARG(ctx, K <= (uint32_t)kRecMaxK, "rfd_il_reconstruct_bpm: K=%u > %d", K, kRecMaxK);
  RFD_CHECK_ARG(ctx, Nx && Ny && Nz && (uint64_t)Nx * Ny * Nz < 0x7FFFFFFFull,
                "rfd_il_reconstruct_bpm: bad grid %u x %u x %u", Nx, Ny, Nz);
  hipStream_t q = ctx->stream;
  const unsigned blocks = (N + kRecThreads - 1) / kRecThreads;
  const size_t lds = sizeof(uint32_t) * kRecThreads * rec_pitch(K);
  // the fine grid spans the coarse grid plus kFineHalf fine cells on each side
  FineGrid f{};
  bool fine = aacut > 0.0f && std::isfinite(aacut) && N < (1u << kFineIdBits) - 1u;
  if (fine) {
    f.h = kFineHalf;
    f.af = sqrtf(aacut) * 1.01f / (float)kFineHalf;
    f.inv_af = 1.0f / f.af;
    f.off = (float)kFineHalf * f.af;
    const double ext[3] = {(double)Nx * a, (double)Ny * a, (double)Nz * a};
    uint32_t dims[3];
    for (int d = 0; d < 3; ++d) {
      const double n = (ext[d] + 2.0 * kFineHalf * (double)f.af) / f.af + 2.0;
      fine = fine && n < (double)kFineMaxDim;
      dims[d] = fine ? (uint32_t)n : 0;
    }
    f.nx = dims[0];
    f.ny = dims[1];
    f.nz = dims[2];
    fine = fine && (uint64_t)f.nx * f.ny * f.nz < 0x7FFFFFFFull;
  }
  if (!fine) {
    void *ws = nullptr;
    int st = rfd_scratch(ctx, sizeof(float4) * (size_t)N, &ws);
    if (st) return st;
    float4 *Rc = (float4 *)ws;
    k_cell_positions<<<rfd_blocks(N, 256), 256, 0, q>>>(R, N, CIs, Rc);
    k_il_reconstruct<<<blocks, kRecThreads, lds, q>>>(R, N, Rc, pnC, IL, ILtype, K, a, inv_a,
                                                      aacut, Nx, Ny, Nz, quirk_p0,
                                                      ctx->d_flags);
    RFD_LAUNCHED(ctx, "rfd_il_reconstruct_bpm");
    return RFD_OK;
  }
  const uint32_t C = Nx * Ny * Nz;
  const uint32_t Cf = f.nx * f.ny * f.nz;
  const uint32_t chunks = (uint32_t)(((size_t)Cf + kScanChunk - 1) / kScanChunk);
  // workspace: fine-ordered positions (16 B each), then placeOf[N], fci[N],
  // cnt[Cf], excl[Cf + 1], part[chunks]
  const size_t words = 3 * (size_t)N + 2 * (size_t)Cf + 1 + chunks;
  void *ws = nullptr;
  int st = rfd_scratch(ctx, sizeof(float4) * (size_t)N + sizeof(uint32_t) * words, &ws);
  if (st) return st;
  float4 *Rf = (float4 *)ws;
  uint32_t *placeOf = (uint32_t *)(Rf + N), *fci = placeOf + N, *cnt = fci + N,
           *excl = cnt + Cf, *part = excl + Cf + 1;
  k_fill_u32<<<rfd_blocks(N, 256, 4096), 256, 0, q>>>(placeOf, N, kSentinel);
  k_fill_u32<<<rfd_blocks(Cf, 256, 4096), 256, 0, q>>>(cnt, Cf, 0u);
  k_place_of<<<rfd_blocks(N, 256), 256, 0, q>>>(CIs, N, placeOf);
  k_fine_index<<<rfd_blocks(N, 256), 256, 0, q>>>(R, N, f, fci, cnt);
  if (chunks > 1) {
    k_chunk_sums<<<chunks, kScanThreads, 0, q>>>(cnt, Cf, part);
    k_scan_chunk_sums<<<1, kScanThreads, 0, q>>>(part, chunks);
  }
  k_scan_excl<<<chunks, kScanThreads, 0, q>>>(cnt, Cf, excl, chunks > 1 ? part : nullptr);
  k_fill_u32<<<rfd_blocks(Cf, 256, 4096), 256, 0, q>>>(cnt, Cf, 0u);
  k_fine_scatter<<<rfd_blocks(N, 256), 256, 0, q>>>(R, N, fci, placeOf, excl, cnt, Rf);
  const uint32_t per = (blocks + kXcds - 1) / kXcds;
  k_il_reconstruct_fine<<<per * kXcds, kRecThreads, lds, q>>>(
      R, N, CIs, pnC, Rf, excl, f, IL, ILtype, K, a, inv_a, aacut, Nx, Ny, Nz, quirk_p0,
      ctx->d_flags, blocks, per);
  RFD_LAUNCHED(ctx, "rfd_il_reconstruct_bpm");
  return RFD_OK;
}

int rfd_bonds_construct(rfd_ctx *ctx, const float *R, const uint32_t *IL, uint8_t *ILtype,
                        float *iL, float *rij, float *Oijt, float *Fijn, float *Fijt,
                        float *Mijn, float *Mijt, float *Mijadd, uint32_t N, uint32_t K,
                        float bdd) {
  if (!ctx) return RFD_E_ARG;
  RFD_CHECK_ARG(ctx, R && IL && ILtype && iL && rij && Oijt && Fijn && Fijt && Mijn && Mijt &&
                         Mijadd && N > 0 && K > 0 && slots_fit(N, K),
                "rfd_bonds_construct: bad arguments");
  k_bonds_construct<<<rfd_blocks((size_t)N * K, 256), 256, 0, ctx->stream>>>(
      R, IL, ILtype, iL, rij, Oijt, Fijn, Fijt, Mijn, Mijt, Mijadd, N, K, bdd);
  RFD_LAUNCHED(ctx, "rfd_bonds_construct");
  return RFD_OK;
}

int rfd_delete_far_links(rfd_ctx *ctx, const float *R, uint32_t N, uint32_t *IL,
                         uint8_t *ILtype, float *rij, float *Oijt, float *Fijn,
                         float *Fijt, float *Mijn, float *Mijt, float *Mijadd,
                         uint32_t K, float aa_farcut) {
  if (!ctx) return RFD_E_ARG;
  RFD_CHECK_ARG(ctx, R && IL && ILtype && rij && Oijt && Fijn && Fijt && Mijn && Mijt &&
                         Mijadd && N > 0 && K > 0 && slots_fit(N, K),
                "rfd_delete_far_links: bad arguments");
  if (K % 4 == 0 && reinterpret_cast<uintptr_t>(IL) % 16 == 0) {
    const size_t quads = (size_t)N * (K / 4);
    k_delete_far4<<<(unsigned)((quads + 255) / 256), 256, 0, ctx->stream>>>(
        R, N, IL, ILtype, rij, Oijt, Fijn, Fijt, Mijn, Mijt, Mijadd, K, aa_farcut);
  } else {
    k_delete_far<<<rfd_blocks((size_t)N * K, 256), 256, 0, ctx->stream>>>(
        R, N, IL, ILtype, rij, Oijt, Fijn, Fijt, Mijn, Mijt, Mijadd, K, aa_farcut);
  }
  RFD_LAUNCHED(ctx, "rfd_delete_far_links");
  return RFD_OK;
}

}  // extern "C"
