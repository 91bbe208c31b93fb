// mpf_permute.hip -- marker rows between the input order and the bins' order in two passes
// of whole blocks instead of one pass of scattered 4-byte accesses.
//
// A binning's permutation: bins-order index i holds input marker j = order[i].  A plan cuts
// the bins-order indices into chunks of PERM_CHUNK and the input indices into buckets of
// PERM_BLOCK.  Between the two orders lies a staging order: bucket after bucket, within a
// bucket chunk after chunk.  Moving rows from the bins' order to the input order takes two
// passes that each read and write whole blocks:
//  pass 1, one workgroup per chunk: the chunk's values go to LDS at their slot q[i] in the
//   chunk's bucket-grouped order, and from LDS slot q to staging position dst[chunk, q], so
//   that the values of one (chunk, bucket) pair leave as one run of consecutive addresses
//   (16 values on average at 2^26 markers);
//  pass 2, one workgroup per bucket: the bucket's staged values go to LDS at lo[k], their
//   input index within the bucket, and leave in the input order, contiguous.
// The opposite direction (input order to bins order) runs the passes backwards.  A plan is
// made once per binning (its counts per (chunk, bucket), their scan down each bucket, the
// slots): bins kept across steps keep their plan.
#include "mpf_internal.h"

namespace mpf {

constexpr int PERM_SHIFT = 15;                               // buckets of 2^15 input indices
constexpr int64_t PERM_BLOCK = int64_t(1) << PERM_SHIFT;
constexpr int PERM_CSHIFT = 15;                              // chunks of 2^15 bins-order indices
constexpr int64_t PERM_CHUNK = int64_t(1) << PERM_CSHIFT;
constexpr int PERM_THREADS = 1024;
constexpr int PERM_PER = (int)(PERM_BLOCK / PERM_THREADS);   // indices per work-item and bucket
constexpr int PERM_CPER = (int)(PERM_CHUNK / PERM_THREADS);  // indices per work-item and chunk
constexpr int PERM_MAX_BLOCKS = 8192;                        // 2^28 markers: a 256 MiB counts matrix
constexpr uint32_t PERM_NONE = 0xFFFFFFFFu;

static int64_t perm_blocks(int64_t n) { return (n + PERM_BLOCK - 1) >> PERM_SHIFT; }
static int64_t perm_chunks(int64_t n) { return (n + PERM_CHUNK - 1) >> PERM_CSHIFT; }
static size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

// The plan in one caller-given buffer: q[n] (uint16), dst[chunks * PERM_CHUNK] (uint32),
// lo[n] (uint16).
struct PermPlan {
  uint16_t *q;
  uint32_t *dst;
  uint16_t *lo;
};

static size_t plan_bytes(int64_t n) {
  return align256(2 * (size_t)n) + align256(4 * (size_t)perm_chunks(n) * PERM_CHUNK) + align256(2 * (size_t)n);
}

static PermPlan plan_at(const void *p, int64_t n) {
  char *b = (char *)p;
  PermPlan pl;
  pl.q = (uint16_t *)b;
  pl.dst = (uint32_t *)(b + align256(2 * (size_t)n));
  pl.lo = (uint16_t *)(b + align256(2 * (size_t)n) + align256(4 * (size_t)perm_chunks(n) * PERM_CHUNK));
  return pl;
}

// The bucket of input index j (kept in range whatever j holds, so a plan made from a
// sequence that is not a permutation still writes inside its arrays).
__device__ __forceinline__ uint32_t bucket_of(int64_t j, int nb) {
  const uint64_t b = (uint64_t)j >> PERM_SHIFT;
  return b < (uint64_t)nb ? (uint32_t)b : (uint32_t)(nb - 1);
}

// Exclusive prefix sum of a[0 .. count) in LDS (count <= PERM_MAX_BLOCKS), by all
// PERM_THREADS work-items; ends with a barrier.
__device__ void perm_block_scan(uint32_t *a, int count, uint32_t *wave_sums) {
  constexpr int E = PERM_MAX_BLOCKS / PERM_THREADS;
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  uint32_t v[E], sum = 0;
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int k = t * E + e;
    v[e] = k < count ? a[k] : 0u;
    sum += v[e];
  }
  // inclusive scan of the per-work-item sums across the wavefront
  uint32_t incl = sum;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint32_t o = __shfl_up(incl, d, 64);
    if (lane >= d) incl += o;
  }
  if (lane == 63) wave_sums[w] = incl;
  __syncthreads();
  uint32_t before = 0;
  for (int u = 0; u < w; ++u) before += wave_sums[u];
  uint32_t run = before + incl - sum;
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int k = t * E + e;
    if (k < count) a[k] = run;
    run += v[e];
  }
  __syncthreads();
}

// Plan, step 1 (one workgroup per chunk): the chunk's counts per bucket into its row
// cnt[chunk * blocks + bucket], and every index's slot q[i] in the chunk's bucket-grouped
// order (bucket by bucket; within a bucket in the order the LDS atomics give).
__global__ void __launch_bounds__(PERM_THREADS) perm_count_kernel(const int64_t *__restrict__ order, int64_t n, int nb,
                                                                  uint16_t *__restrict__ q, uint32_t *__restrict__ cnt) {
  __shared__ uint32_t hist[PERM_MAX_BLOCKS];
  __shared__ uint32_t wave_sums[PERM_THREADS / 64];
  const int c = blockIdx.x, t = threadIdx.x;
  const int64_t i0 = (int64_t)c << PERM_CSHIFT;
  for (int b = t; b < nb; b += PERM_THREADS) hist[b] = 0u;
  __syncthreads();
  uint32_t bk[PERM_CPER], rank[PERM_CPER];
#pragma unroll
  for (int e = 0; e < PERM_CPER; ++e) {
    const int64_t i = i0 + e * PERM_THREADS + t;
    bk[e] = i < n ? bucket_of(order[i], nb) : PERM_NONE;
  }
#pragma unroll
  for (int e = 0; e < PERM_CPER; ++e) rank[e] = bk[e] != PERM_NONE ? atomicAdd(&hist[bk[e]], 1u) : 0u;
  __syncthreads();
  for (int b = t; b < nb; b += PERM_THREADS) cnt[(int64_t)c * nb + b] = hist[b];
  perm_block_scan(hist, nb, wave_sums);
#pragma unroll
  for (int e = 0; e < PERM_CPER; ++e) {
    const int64_t i = i0 + e * PERM_THREADS + t;
    if (bk[e] != PERM_NONE) q[i] = (uint16_t)(hist[bk[e]] + rank[e]);
  }
}

// Plan, step 2: every (chunk, bucket) pair's first staging position, base[chunk * blocks +
// bucket] = bucket * PERM_BLOCK + the bucket's counts in the chunks before (a bucket holds
// PERM_BLOCK input indices, the last one the rest): an exclusive scan down each column of
// the counts.  One workgroup per 64 buckets, lane = bucket; its 16 wavefronts take 16
// consecutive ranges of chunks, sum them, scan the 16 sums and walk their ranges again.
__global__ void __launch_bounds__(PERM_THREADS) perm_base_kernel(int nb, int nc, const uint32_t *__restrict__ cnt,
                                                                 uint32_t *__restrict__ base) {
  constexpr int WAVES = PERM_THREADS / 64;
  __shared__ uint32_t sums[WAVES][64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int b = blockIdx.x * 64 + lane;
  const int per = (nc + WAVES - 1) / WAVES, c0 = w * per, c1 = c0 + per < nc ? c0 + per : nc;
  uint32_t sum = 0;
  if (b < nb)
    for (int c = c0; c < c1; ++c) sum += cnt[(int64_t)c * nb + b];
  sums[w][lane] = sum;
  __syncthreads();
  uint32_t run = (uint32_t)b << PERM_SHIFT;
  for (int u = 0; u < w; ++u) run += sums[u][lane];
  if (b < nb)
    for (int c = c0; c < c1; ++c) {
      const int64_t at = (int64_t)c * nb + b;
      const uint32_t k = cnt[at];
      base[at] = run;
      run += k;
    }
}

// Plan, step 3 (one workgroup per chunk): the staging position of every slot,
// dst[chunk * PERM_CHUNK + q] = base[chunk, bucket] + (q - the bucket's first slot in the
// chunk), and lo in staging order (the input index of each staged entry within its
// bucket).  Each index's bucket and lo go to LDS by slot, so that dst is written in slot
// order (contiguous) and lo by runs of one (chunk, bucket) pair.  LDS: the chunk's 2^15
// slots (128 KiB) and the bucket offsets (32 KiB); the scan's wave sums borrow the slots.
__global__ void __launch_bounds__(PERM_THREADS) perm_dest_kernel(const int64_t *__restrict__ order, int64_t n, int nb,
                                                                 const uint16_t *__restrict__ q,
                                                                 const uint32_t *__restrict__ cnt,
                                                                 const uint32_t *__restrict__ base,
                                                                 uint32_t *__restrict__ dst,
                                                                 uint16_t *__restrict__ lo) {
  __shared__ uint32_t off[PERM_MAX_BLOCKS];
  __shared__ uint32_t slots[PERM_CHUNK];  // bucket | lo << 16 of the index in each slot
  const int c = blockIdx.x, t = threadIdx.x;
  const int64_t i0 = (int64_t)c << PERM_CSHIFT;
  const int64_t count = n - i0 < PERM_CHUNK ? n - i0 : PERM_CHUNK;
  for (int b = t; b < nb; b += PERM_THREADS) off[b] = cnt[(int64_t)c * nb + b];
  __syncthreads();
  perm_block_scan(off, nb, slots);
  // base - first slot, modulo 2^32: adding a slot of the bucket gives its staging position
  for (int b = t; b < nb; b += PERM_THREADS) off[b] = base[(int64_t)c * nb + b] - off[b];
#pragma unroll 8
  for (int e = 0; e < PERM_CPER; ++e) {
    const int64_t i = i0 + e * PERM_THREADS + t;
    if (i < n) {
      const int64_t j = order[i];
      slots[q[i]] = bucket_of(j, nb) | (uint32_t)(((uint64_t)j & (uint64_t)(PERM_BLOCK - 1)) << 16);
    }
  }
  __syncthreads();
#pragma unroll 8
  for (int e = 0; e < PERM_CPER; ++e) {
    const int k = e * PERM_THREADS + t;
    if (k < count) {
      const uint32_t sl = slots[k];
      const uint32_t d = off[sl & 0xFFFFu] + (uint32_t)k;
      dst[i0 + k] = d;
      if (d < (uint64_t)n) lo[d] = (uint16_t)(sl >> 16);
    }
  }
}

// The LDS of the apply passes: a block of float32, or half a block of float64 per round.
template <class T>
struct PermRounds {
  static constexpr int H = (int)(sizeof(T) / 4);            // rounds per block
  static constexpr int SPAN = (int)(PERM_BLOCK / H);          // LDS entries per round
};

// The chunk of workgroup b of a run pass: workgroups go round-robin over the eight XCDs,
// and each XCD takes a contiguous range of chunks, so that the chunks an XCD runs at once
// are neighbours and their runs of one bucket, adjacent in the staging order, meet in that
// XCD's L2 (partial lines merge there before they are written back).
__device__ __forceinline__ int xcd_chunk(int b, int nb) {
  const int x = b % 8, k = b / 8, per = nb / 8, extra = nb % 8;
  return x * per + (x < extra ? x : extra) + k;
}

// Pass 1, bins order -> staging (one workgroup per chunk, every row).  Each work-item keeps
// the slots of its 32 bins-order indices and the staging positions of its 32 slots.
template <class T>
__global__ void __launch_bounds__(PERM_THREADS) perm_bins_to_stage(const T *__restrict__ src, int64_t ss,
                                                                   int64_t rows, int64_t n,
                                                                   const uint16_t *__restrict__ q,
                                                                   const uint32_t *__restrict__ dst,
                                                                   T *__restrict__ stage) {
  using R = PermRounds<T>;
  __shared__ float words[PERM_CHUNK];
  T *s = reinterpret_cast<T *>(words);
  const int c = xcd_chunk(blockIdx.x, (int)gridDim.x), t = threadIdx.x;
  const int64_t i0 = (int64_t)c << PERM_CSHIFT;
  const int64_t count = n - i0 < PERM_CHUNK ? n - i0 : PERM_CHUNK;
  uint32_t qs[PERM_CPER], ds[PERM_CPER];
#pragma unroll
  for (int e = 0; e < PERM_CPER; ++e) {
    const int k = e * PERM_THREADS + t;
    qs[e] = k < count ? (uint32_t)q[i0 + k] : PERM_NONE;
    ds[e] = k < count ? dst[i0 + k] : PERM_NONE;
  }
  for (int64_t r = 0; r < rows; ++r) {
    const T *row = src + r * ss + i0;
    T *out = stage + r * n;
#pragma unroll
    for (int h = 0; h < R::H; ++h) {
      const uint32_t lo = (uint32_t)(h * R::SPAN), hi = lo + (uint32_t)R::SPAN;
#pragma unroll
      for (int e = 0; e < PERM_CPER; ++e)
        if (qs[e] >= lo && qs[e] < hi) s[qs[e] - lo] = row[e * PERM_THREADS + t];
      __syncthreads();
#pragma unroll
      for (int e = 0; e < PERM_CPER; ++e) {
        const uint32_t k = (uint32_t)(e * PERM_THREADS + t);
        if (k >= lo && k < hi && ds[e] < (uint64_t)n) out[ds[e]] = s[k - lo];
      }
      __syncthreads();
    }
  }
}

// Pass 2, staging -> input order (one workgroup per bucket, every row).
template <class T>
__global__ void __launch_bounds__(PERM_THREADS) perm_stage_to_input(const T *__restrict__ stage, int64_t rows,
                                                                    int64_t n, int blk0, const uint16_t *__restrict__ lo_idx,
                                                                    T *__restrict__ dst, int64_t ds_) {
  using R = PermRounds<T>;
  __shared__ float words[PERM_BLOCK];
  T *s = reinterpret_cast<T *>(words);
  const int b = blk0 + blockIdx.x, t = threadIdx.x;
  const int64_t k0 = (int64_t)b << PERM_SHIFT;
  const int64_t size = n - k0 < PERM_BLOCK ? n - k0 : PERM_BLOCK;
  uint32_t ls[PERM_PER];
#pragma unroll
  for (int e = 0; e < PERM_PER; ++e) {
    const int k = e * PERM_THREADS + t;
    ls[e] = k < size ? (uint32_t)lo_idx[k0 + k] : PERM_NONE;
  }
  for (int64_t r = 0; r < rows; ++r) {
    const T *in = stage + r * n + k0;
    T *out = dst + r * ds_ + k0;
#pragma unroll
    for (int h = 0; h < R::H; ++h) {
      const uint32_t lo = (uint32_t)(h * R::SPAN), hi = lo + (uint32_t)R::SPAN;
#pragma unroll
      for (int e = 0; e < PERM_PER; ++e)
        if (ls[e] >= lo && ls[e] < hi) s[ls[e] - lo] = in[e * PERM_THREADS + t];
      __syncthreads();
#pragma unroll
      for (int e = 0; e < PERM_PER; ++e) {
        const int64_t k = e * PERM_THREADS + t;
        if (k >= lo && k < hi && k < size) out[k] = s[k - lo];
      }
      __syncthreads();
    }
  }
}

// The opposite of pass 2, input order -> staging (one workgroup per bucket).
template <class T>
__global__ void __launch_bounds__(PERM_THREADS) perm_input_to_stage(const T *__restrict__ src, int64_t ss,
                                                                    int64_t rows, int64_t n, int blk0,
                                                                    const uint16_t *__restrict__ lo_idx,
                                                                    T *__restrict__ stage) {
  using R = PermRounds<T>;
  __shared__ float words[PERM_BLOCK];
  T *s = reinterpret_cast<T *>(words);
  const int b = blk0 + blockIdx.x, t = threadIdx.x;
  const int64_t k0 = (int64_t)b << PERM_SHIFT;
  const int64_t size = n - k0 < PERM_BLOCK ? n - k0 : PERM_BLOCK;
  uint32_t ls[PERM_PER];
#pragma unroll
  for (int e = 0; e < PERM_PER; ++e) {
    const int k = e * PERM_THREADS + t;
    ls[e] = k < size ? (uint32_t)lo_idx[k0 + k] : PERM_NONE;
  }
  for (int64_t r = 0; r < rows; ++r) {
    const T *in = src + r * ss + k0;
    T *out = stage + r * n + k0;
#pragma unroll
    for (int h = 0; h < R::H; ++h) {
      const uint32_t lo = (uint32_t)(h * R::SPAN), hi = lo + (uint32_t)R::SPAN;
#pragma unroll
      for (int e = 0; e < PERM_PER; ++e) {
        const int64_t k = e * PERM_THREADS + t;
        if (k >= lo && k < hi && k < size) s[k - lo] = in[k];
      }
      __syncthreads();
#pragma unroll
      for (int e = 0; e < PERM_PER; ++e)
        if (ls[e] >= lo && ls[e] < hi) out[e * PERM_THREADS + t] = s[ls[e] - lo];
      __syncthreads();
    }
  }
}

// The opposite of pass 1, staging -> bins order (one workgroup per chunk).
template <class T>
__global__ void __launch_bounds__(PERM_THREADS) perm_stage_to_bins(const T *__restrict__ stage, int64_t rows,
                                                                   int64_t n, const uint16_t *__restrict__ q,
                                                                   const uint32_t *__restrict__ dst,
                                                                   T *__restrict__ out_rows, int64_t os) {
  using R = PermRounds<T>;
  __shared__ float words[PERM_CHUNK];
  T *s = reinterpret_cast<T *>(words);
  const int c = xcd_chunk(blockIdx.x, (int)gridDim.x), t = threadIdx.x;
  const int64_t i0 = (int64_t)c << PERM_CSHIFT;
  const int64_t count = n - i0 < PERM_CHUNK ? n - i0 : PERM_CHUNK;
  uint32_t qs[PERM_CPER], ds[PERM_CPER];
#pragma unroll
  for (int e = 0; e < PERM_CPER; ++e) {
    const int k = e * PERM_THREADS + t;
    qs[e] = k < count ? (uint32_t)q[i0 + k] : PERM_NONE;
    ds[e] = k < count ? dst[i0 + k] : PERM_NONE;
  }
  for (int64_t r = 0; r < rows; ++r) {
    const T *in = stage + r * n;
    T *out = out_rows + r * os + i0;
#pragma unroll
    for (int h = 0; h < R::H; ++h) {
      const uint32_t lo = (uint32_t)(h * R::SPAN), hi = lo + (uint32_t)R::SPAN;
#pragma unroll
      for (int e = 0; e < PERM_CPER; ++e) {
        const uint32_t k = (uint32_t)(e * PERM_THREADS + t);
        if (k >= lo && k < hi && ds[e] < (uint64_t)n) s[k - lo] = in[ds[e]];
      }
      __syncthreads();
#pragma unroll
      for (int e = 0; e < PERM_CPER; ++e)
        if (qs[e] >= lo && qs[e] < hi) out[e * PERM_THREADS + t] = s[qs[e] - lo];
      __syncthreads();
    }
  }
}

// The float32 forms of the four passes for the whole blocks (the last block, if partial,
// takes the generic kernels): one workgroup per CU walks its blocks, and the values of the
// next (block, row) are loaded while the current one leaves LDS, so that a workgroup (alone
// on its CU with 128 KiB of LDS) always has loads in flight.  The 16-bit slots of a
// work-item's indices are kept two to a register; staging positions are clamped into the
// rows (a plan of a sequence that is not a permutation still stays inside them).
__device__ __forceinline__ void load_slots(const uint16_t *__restrict__ idx, uint32_t (&p)[PERM_PER / 2]) {
#pragma unroll
  for (int e = 0; e < PERM_PER; e += 2) {
    const int k = e * PERM_THREADS + threadIdx.x;
    p[e / 2] = (uint32_t)idx[k] | ((uint32_t)idx[k + PERM_THREADS] << 16);
  }
}

__device__ __forceinline__ uint32_t slot_at(const uint32_t (&p)[PERM_PER / 2], int e) {
  return (e & 1) ? p[e / 2] >> 16 : p[e / 2] & 0xFFFFu;
}

__device__ __forceinline__ void load_block(const float *__restrict__ row, float (&v)[PERM_PER]) {
#pragma unroll
  for (int e = 0; e < PERM_PER; ++e) v[e] = row[e * PERM_THREADS + threadIdx.x];
}

// The block and row a workgroup takes after (blk, r): the next row of the block, else row 0
// of its next block; false past its last.
__device__ __forceinline__ bool next_task(int blk, int64_t r, int64_t rows, int nblocks, int &nblk, int64_t &nr) {
  if (r + 1 < rows) {
    nblk = blk;
    nr = r + 1;
    return true;
  }
  nblk = blk + (int)gridDim.x;
  nr = 0;
  return nblk < nblocks;
}

// Pass 2, staging -> input order.
__global__ void __launch_bounds__(PERM_THREADS) perm_stage_to_input_f32(const float *__restrict__ stage, int64_t rows,
                                                                        int64_t n, int nfull,
                                                                        const uint16_t *__restrict__ lo_idx,
                                                                        float *__restrict__ dst, int64_t ds_) {
  __shared__ float s[PERM_BLOCK];
  const int t = threadIdx.x;
  float v[PERM_PER];
  int blk = blockIdx.x;
  if (blk < nfull) load_block(stage + ((int64_t)blk << PERM_SHIFT), v);
  for (; blk < nfull; blk += gridDim.x) {
    const int64_t k0 = (int64_t)blk << PERM_SHIFT;
    uint32_t ls[PERM_PER / 2];
    load_slots(lo_idx + k0, ls);
    for (int64_t r = 0; r < rows; ++r) {
#pragma unroll
      for (int e = 0; e < PERM_PER; ++e) s[slot_at(ls, e)] = v[e];
      __syncthreads();
      int nblk;
      int64_t nr;
      if (next_task(blk, r, rows, nfull, nblk, nr)) load_block(stage + nr * n + ((int64_t)nblk << PERM_SHIFT), v);
      float *out = dst + r * ds_ + k0;
#pragma unroll
      for (int e = 0; e < PERM_PER; ++e) out[e * PERM_THREADS + t] = s[e * PERM_THREADS + t];
      __syncthreads();
    }
  }
}

// The opposite of pass 2, input order -> staging.
__global__ void __launch_bounds__(PERM_THREADS) perm_input_to_stage_f32(const float *__restrict__ src, int64_t ss,
                                                                        int64_t rows, int64_t n, int nfull,
                                                                        const uint16_t *__restrict__ lo_idx,
                                                                        float *__restrict__ stage) {
  __shared__ float s[PERM_BLOCK];
  const int t = threadIdx.x;
  float v[PERM_PER];
  int blk = blockIdx.x;
  if (blk < nfull) load_block(src + ((int64_t)blk << PERM_SHIFT), v);
  for (; blk < nfull; blk += gridDim.x) {
    const int64_t k0 = (int64_t)blk << PERM_SHIFT;
    uint32_t ls[PERM_PER / 2];
    load_slots(lo_idx + k0, ls);
    for (int64_t r = 0; r < rows; ++r) {
#pragma unroll
      for (int e = 0; e < PERM_PER; ++e) s[e * PERM_THREADS + t] = v[e];
      __syncthreads();
      int nblk;
      int64_t nr;
      if (next_task(blk, r, rows, nfull, nblk, nr)) load_block(src + nr * ss + ((int64_t)nblk << PERM_SHIFT), v);
      float *out = stage + r * n + k0;
#pragma unroll
      for (int e = 0; e < PERM_PER; ++e) out[e * PERM_THREADS + t] = s[slot_at(ls, e)];
      __syncthreads();
    }
  }
}

// Workgroups of the float32 passes: one per CU (their LDS allows no more).
static int perm_grid(int nblocks) {
  static int cus[64] = {0};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return nblocks;
  if (!cus[dev] && hipDeviceGetAttribute(&cus[dev], hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
    return nblocks;
  return nblocks < cus[dev] ? nblocks : cus[dev];
}

bool perm_plan_ok(int64_t n) { return n >= 1 && perm_blocks(n) <= PERM_MAX_BLOCKS; }

int perm_plan_build(const int64_t *order, int64_t n, void *plan, mpf_context *ctx, hipStream_t stream) {
  if (!perm_plan_ok(n)) {
    set_error("a permutation plan takes 1 to 2^28 markers, got %lld", (long long)n);
    return MPF_ERR_ARGUMENT;
  }
  const int nb = (int)perm_blocks(n), nc = (int)perm_chunks(n);
  const size_t entries = (size_t)nc * nb;
  int rc = ensure(ctx->perm_build, 2 * align256(4 * entries), false, stream);
  if (rc) return rc;
  uint32_t *cnt = (uint32_t *)ctx->perm_build.ptr;
  uint32_t *base = (uint32_t *)((char *)ctx->perm_build.ptr + align256(4 * entries));
  const PermPlan pl = plan_at(plan, n);
  hipLaunchKernelGGL(perm_count_kernel, dim3(nc), dim3(PERM_THREADS), 0, stream, order, n, nb, pl.q, cnt);
  MPF_LAUNCH_CHECK();
  hipLaunchKernelGGL(perm_base_kernel, dim3((nb + 63) / 64), dim3(PERM_THREADS), 0, stream, nb, nc, cnt, base);
  MPF_LAUNCH_CHECK();
  hipLaunchKernelGGL(perm_dest_kernel, dim3(nc), dim3(PERM_THREADS), 0, stream, order, n, nb, pl.q, cnt, base, pl.dst,
                     pl.lo);
  MPF_LAUNCH_CHECK();
  return MPF_OK;
}

template <class T>
static int perm_apply_t(const PermPlan &pl, int64_t n, int direction, const T *src, int64_t ss, T *dst,
                        int64_t ds_, int64_t rows, mpf_context *ctx, hipStream_t stream) {
  const int nb = (int)perm_blocks(n);
  int rc = ensure(ctx->perm_stage, (size_t)rows * n * sizeof(T), false, stream);
  if (rc) return rc;
  T *stage = (T *)ctx->perm_stage.ptr;
  // the run passes: one workgroup per chunk; the bucket passes of float32 rows: the whole
  // buckets by the pipelined kernels and a partial last one by the generic kernel, float64
  // rows: every bucket by the generic kernel
  const int nfull = sizeof(T) == 4 ? (int)(n >> PERM_SHIFT) : 0;
  const dim3 block(PERM_THREADS);
  auto buckets = [&](bool to_input) -> int {
    if constexpr (sizeof(T) == 4) {
      if (nfull > 0) {
        if (to_input)
          hipLaunchKernelGGL(perm_stage_to_input_f32, dim3(perm_grid(nfull)), block, 0, stream, stage, rows, n, nfull,
                             pl.lo, dst, ds_);
        else
          hipLaunchKernelGGL(perm_input_to_stage_f32, dim3(perm_grid(nfull)), block, 0, stream, src, ss, rows, n,
                             nfull, pl.lo, stage);
        MPF_LAUNCH_CHECK();
      }
    }
    if (nb > nfull) {
      if (to_input)
        hipLaunchKernelGGL(perm_stage_to_input<T>, dim3(nb - nfull), block, 0, stream, stage, rows, n, nfull, pl.lo,
                           dst, ds_);
      else
        hipLaunchKernelGGL(perm_input_to_stage<T>, dim3(nb - nfull), block, 0, stream, src, ss, rows, n, nfull, pl.lo,
                           stage);
      MPF_LAUNCH_CHECK();
    }
    return MPF_OK;
  };
  const int nc = (int)perm_chunks(n);
  if (direction == 0) {
    hipLaunchKernelGGL(perm_bins_to_stage<T>, dim3(nc), block, 0, stream, src, ss, rows, n, pl.q, pl.dst, stage);
    MPF_LAUNCH_CHECK();
    return buckets(true);
  }
  rc = buckets(false);
  if (rc) return rc;
  hipLaunchKernelGGL(perm_stage_to_bins<T>, dim3(nc), block, 0, stream, stage, rows, n, pl.q, pl.dst, dst, ds_);
  MPF_LAUNCH_CHECK();
  return MPF_OK;
}

int perm_apply(const void *plan, int64_t n, int direction, const void *src, int64_t src_stride, void *dst,
               int64_t dst_stride, int64_t rows, int dtype, mpf_context *ctx, hipStream_t stream) {
  if (rows == 0 || n == 0) return MPF_OK;
  const PermPlan pl = plan_at(plan, n);
  if (dtype == F32)
    return perm_apply_t<float>(pl, n, direction, (const float *)src, src_stride, (float *)dst, dst_stride, rows, ctx,
                               stream);
  return perm_apply_t<double>(pl, n, direction, (const double *)src, src_stride, (double *)dst, dst_stride, rows, ctx,
                              stream);
}

// Rows in bins order back to the input order through a plan made in the context (the
// in-call route of mpf_gather and mpf_unsort_rows): from PERM_MIN markers, where the plan
// and the two passes cost less than scattered writes.
int unsort_planned(const void *sorted, void *out, int64_t out_stride, const int64_t *order, int64_t rows, int64_t n,
                   int dtype, mpf_context *ctx, hipStream_t stream) {
  int rc = ensure(ctx->perm_plan, plan_bytes(n), false, stream);
  if (rc) return rc;
  rc = perm_plan_build(order, n, ctx->perm_plan.ptr, ctx, stream);
  if (rc) return rc;
  return perm_apply(ctx->perm_plan.ptr, n, 0, sorted, n, out, out_stride, rows, dtype, ctx, stream);
}

}  // namespace mpf

using namespace mpf;

extern "C" int mpf_perm_plan_bytes(int64_t n, int64_t *bytes) {
  if (!bytes) {
    set_error("perm_plan_bytes: bytes is NULL");
    return MPF_ERR_ARGUMENT;
  }
  *bytes = perm_plan_ok(n) ? (int64_t)plan_bytes(n) : 0;
  return MPF_OK;
}

extern "C" int mpf_perm_plan_build(const int64_t *order, int64_t n, void *plan, mpf_context *ctx,
                                   mpf_stream stream_) {
  hipStream_t stream = (hipStream_t)stream_;
  if (!order || !plan) {
    set_error("perm_plan_build: order and plan must be device pointers");
    return MPF_ERR_ARGUMENT;
  }
  if (!ctx) ctx = default_context(stream);
  if (!ctx) return MPF_ERR_HIP;
  return perm_plan_build(order, n, plan, ctx, stream);
}

extern "C" int mpf_perm_rows(const void *plan, int64_t n, int direction, const void *src, int64_t src_stride,
                             void *dst, int64_t dst_stride, int64_t rows, int dtype, mpf_context *ctx,
                             mpf_stream stream_) {
  hipStream_t stream = (hipStream_t)stream_;
  if (rows < 0 || n < 0 || (direction != 0 && direction != 1) || (dtype != F32 && dtype != F64)) {
    set_error("perm_rows: need rows >= 0, n >= 0, direction 0 or 1 and dtype F32 or F64");
    return MPF_ERR_ARGUMENT;
  }
  if (rows == 0 || n == 0) return MPF_OK;
  if (!perm_plan_ok(n)) {
    set_error("a permutation plan takes 1 to 2^28 markers, got %lld", (long long)n);
    return MPF_ERR_ARGUMENT;
  }
  if (!plan || !src || !dst || src_stride < n || dst_stride < n) {
    set_error("perm_rows: plan, src and dst must be device pointers and the row strides at least n");
    return MPF_ERR_ARGUMENT;
  }
  if (!ctx) ctx = default_context(stream);
  if (!ctx) return MPF_ERR_HIP;
  return perm_apply(plan, n, direction, src, src_stride, dst, dst_stride, rows, dtype, ctx, stream);
}
