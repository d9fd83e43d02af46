// Cell-key counting sort of the MPM particles.
//
// Key: the particle's base cell (the truncating key of fsi/mpm3d.py:239,
// computed exactly as P2G computes it), shifted by one and clamped to
// [0, n+1] per axis, in 4x4x4-blocked order (blocks row-major, z fastest,
// an even number of blocks in z). The P2G work list cuts the particles of
// each pair of z-adjacent blocks into chunks of at most 256, so a chunk's
// stencil box is at most 6x6x10 nodes and its LDS tile stays small.
// Inactive particles take the last key.
//
// Sort: histogram -> exclusive scan -> scatter with an atomic cursor per key
// -> each key's segment ordered by particle index (ties keep index order),
// so the permutation equals a stable sort of the keys, bit for bit. Then one
// gather moves every particle array into the second buffer set.
#include <algorithm>

#include "common.h"

namespace fsi {
namespace {

constexpr int kScanItems = 4;                      // items per thread in a scan block
constexpr int kScanBlock = kBlock * kScanItems;    // 1024 counts per scan block
constexpr int kInsertionMax = 32;                  // longer key segments use heapsort

struct SortGeom {
  int nx, ny, nz;           // lattice cells
  int pbx, pby, pbz;        // 4-cell blocks per axis of the padded (n+2) box (pbz even)
  uint32_t nkeys;           // keys of cells; key nkeys = inactive particles
};

SortGeom make_geom(const fsi_mpm_params* p) {
  SortGeom g;
  g.nx = p->nx;
  g.ny = p->ny;
  g.nz = p->nz;
  g.pbx = (p->nx + 2 + 3) / 4;
  g.pby = (p->ny + 2 + 3) / 4;
  // an even number of z blocks: the block pairs of a P2G chunk group are
  // then always neighbours in z within one row
  g.pbz = ((p->nz + 2 + 3) / 4 + 1) & ~1;
  g.nkeys = (uint32_t)((int64_t)g.pbx * g.pby * g.pbz * 64);
  return g;
}

int64_t key_space(const SortGeom& g) { return (int64_t)g.pbx * g.pby * g.pbz * 64 + 1; }

struct SortWs {
  int32_t* counts;     // nkeys+1, then reused as the scatter cursor
  int32_t* starts;     // nkeys+1, exclusive scan of counts
  int32_t* blocksum;   // scan blocks
  uint32_t* keys;      // n, unsorted keys
  int32_t* perm;       // n, sorted position -> particle index
  int32_t* nchunk;     // nblocks+1: P2G chunks per cell block (+ the inactive bucket)
  int32_t* chunk_off;  // nblocks+1: exclusive scan of nchunk
  int64_t bytes;
};

int64_t align256(int64_t b) { return (b + 255) & ~int64_t(255); }

constexpr int64_t kGroupKeysHost = 128;  // keys per P2G chunk group (see k_chunk_counts)

// chunk groups for nk = nkeys + 1 keys: ceil(nkeys / group) cell groups and
// the inactive bucket
int64_t chunk_groups(int64_t nk) { return (nk - 1 + kGroupKeysHost - 1) / kGroupKeysHost + 1; }

SortWs carve(void* base, int64_t nk, int64_t n) {
  SortWs w;
  char* c = (char*)base;
  int64_t off = 0;
  const int64_t nblk = (nk + kScanBlock - 1) / kScanBlock;  // also covers the cell-block scan
  const int64_t ncb = chunk_groups(nk);                      // chunk groups + inactive bucket
  w.counts = (int32_t*)(c + off);
  off += align256(nk * 4);
  w.starts = (int32_t*)(c + off);
  off += align256(nk * 4);
  w.blocksum = (int32_t*)(c + off);
  off += align256(nblk * 4);
  w.keys = (uint32_t*)(c + off);
  off += align256(std::max<int64_t>(n, 1) * 4);
  w.perm = (int32_t*)(c + off);
  off += align256(std::max<int64_t>(n, 1) * 4);
  w.nchunk = (int32_t*)(c + off);
  off += align256(ncb * 4);
  w.chunk_off = (int32_t*)(c + off);
  off += align256(ncb * 4);
  w.bytes = off;
  return w;
}

__global__ __launch_bounds__(kBlock) void k_zero_i32(int64_t n, int32_t* __restrict__ a) {
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x)
    a[t] = 0;
}

// Same float arithmetic as locate() in the P2G pass: x * inv_dx - 0.5 in
// the particle precision without contraction, truncated toward zero.
template <typename R>
__device__ __forceinline__ int base_cell(R x, R inv_dx) {
#pragma clang fp contract(off)
  const R s = x * inv_dx - R(0.5);
  return (int)s;
}

template <typename R>
__global__ __launch_bounds__(kBlock) void k_keys(SortGeom g, int num, int cap, R inv_dx,
                                                 const R* __restrict__ x, const int32_t* __restrict__ active,
                                                 uint32_t* __restrict__ keys, int32_t* __restrict__ counts) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= num) return;
  uint32_t key = g.nkeys;
  if (active[p] != 0) {
    const int n3[3] = {g.nx, g.ny, g.nz};
    int c[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const int b = base_cell<R>(x[a * cap + p], inv_dx);
      c[a] = min(max(b, -1), n3[a]) + 1;  // 0 .. n+1
    }
    const uint32_t blk = ((uint32_t)(c[0] >> 2) * g.pby + (uint32_t)(c[1] >> 2)) * g.pbz + (uint32_t)(c[2] >> 2);
    key = blk * 64u + (uint32_t)((((c[0] & 3) * 4) + (c[1] & 3)) * 4 + (c[2] & 3));
  }
  keys[p] = key;
  atomicAdd(&counts[key], 1);
}

// ---- exclusive scan of counts[0..nk) into starts (three passes) --------

__device__ __forceinline__ int block_exclusive_scan(int v, int* total) {
  __shared__ int wsum[kBlock / 64];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  int inc = v;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const int t = __shfl_up(inc, off);
    if (lane >= off) inc += t;
  }
  if (lane == 63) wsum[wid] = inc;
  __syncthreads();
  int before = 0, all = 0;
#pragma unroll
  for (int w = 0; w < kBlock / 64; ++w) {
    before += w < wid ? wsum[w] : 0;
    all += wsum[w];
  }
  __syncthreads();
  *total = all;
  return before + inc - v;
}

__global__ __launch_bounds__(kBlock) void k_scan_blocksums(int64_t nk, const int32_t* __restrict__ counts,
                                                          int32_t* __restrict__ blocksum) {
  const int64_t base = (int64_t)blockIdx.x * kScanBlock + threadIdx.x * kScanItems;
  int s = 0;
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) s += base + i < nk ? counts[base + i] : 0;
  int total;
  block_exclusive_scan(s, &total);
  if (threadIdx.x == 0) blocksum[blockIdx.x] = total;
}

// One workgroup turns the block sums into exclusive block offsets, walking
// them in chunks of kBlock with a running carry.
__global__ __launch_bounds__(kBlock) void k_scan_offsets(int64_t nblk, int32_t* __restrict__ blocksum) {
  int carry = 0;
  for (int64_t c0 = 0; c0 < nblk; c0 += kBlock) {
    const int64_t i = c0 + threadIdx.x;
    const int v = i < nblk ? blocksum[i] : 0;
    int total;
    const int ex = block_exclusive_scan(v, &total);
    if (i < nblk) blocksum[i] = carry + ex;
    carry += total;
  }
}

__global__ __launch_bounds__(kBlock) void k_scan_apply(int64_t nk, const int32_t* __restrict__ counts,
                                                      const int32_t* __restrict__ blockoff,
                                                      int32_t* __restrict__ starts) {
  const int64_t base = (int64_t)blockIdx.x * kScanBlock + threadIdx.x * kScanItems;
  int v[kScanItems], s = 0;
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) {
    v[i] = base + i < nk ? counts[base + i] : 0;
    s += v[i];
  }
  int total;
  int run = block_exclusive_scan(s, &total) + blockoff[blockIdx.x];
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) {
    if (base + i < nk) starts[base + i] = run;
    run += v[i];
  }
}

__global__ __launch_bounds__(kBlock) void k_scatter(int num, const uint32_t* __restrict__ keys,
                                                   const int32_t* __restrict__ starts, int32_t* __restrict__ cursor,
                                                   int32_t* __restrict__ perm) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= num) return;
  const uint32_t k = keys[p];
  perm[starts[k] + atomicAdd(&cursor[k], 1)] = p;
}

__device__ void sift_down(int32_t* a, int root, int n) {
  for (;;) {
    int child = 2 * root + 1;
    if (child >= n) return;
    if (child + 1 < n && a[child] < a[child + 1]) ++child;
    if (!(a[root] < a[child])) return;
    const int32_t t = a[root];
    a[root] = a[child];
    a[child] = t;
    root = child;
  }
}

// Orders each key's segment of perm by particle index. A segment holds the
// particles of one cell (a handful at 8 per cell): insertion sort; long
// segments (crowded cells, the inactive bucket) use heapsort, so the cost is
// O(c log c) in every case.
__global__ __launch_bounds__(kBlock) void k_order_segments(int64_t nk, const int32_t* __restrict__ starts,
                                                          const int32_t* __restrict__ cursor,
                                                          int32_t* __restrict__ perm) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nk) return;
  const int c = cursor[k];
  if (c < 2) return;
  int32_t* seg = perm + starts[k];
  if (c <= kInsertionMax) {
    for (int i = 1; i < c; ++i) {
      const int32_t v = seg[i];
      int j = i - 1;
      while (j >= 0 && seg[j] > v) {
        seg[j + 1] = seg[j];
        --j;
      }
      seg[j + 1] = v;
    }
    return;
  }
  for (int r = c / 2 - 1; r >= 0; --r) sift_down(seg, r, c);
  for (int end = c - 1; end > 0; --end) {
    const int32_t t = seg[0];
    seg[0] = seg[end];
    seg[end] = t;
    sift_down(seg, 0, end);
  }
}

// P2G work list. Group g < ncb-1 holds the keys [g*kGroupKeys,
// min((g+1)*kGroupKeys, nkeys)): two 4x4x4 cell blocks adjacent in z, so a
// chunk's cells lie in a 4x4x8 box (stencil box 6x6x10 = 360 nodes, inside
// the 512-node P2G tile); group ncb-1 is the inactive bucket (key nkeys).
// Each group's particles [lo, hi) of the sorted order are cut into chunks of
// at most kBlock.
constexpr int64_t kGroupKeys = kGroupKeysHost;

__device__ __forceinline__ void block_range(int64_t b, int64_t ncb, int64_t nkeys, int num,
                                            const int32_t* __restrict__ starts, int& lo, int& hi) {
  if (b + 1 < ncb) {
    lo = starts[b * kGroupKeys];
    hi = starts[min((b + 1) * kGroupKeys, nkeys)];
  } else {
    lo = starts[nkeys];
    hi = num;
  }
}

__global__ __launch_bounds__(kBlock) void k_chunk_counts(int64_t ncb, int64_t nkeys, int num,
                                                        const int32_t* __restrict__ starts,
                                                        int32_t* __restrict__ nchunk) {
  const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= ncb) return;
  int lo, hi;
  block_range(b, ncb, nkeys, num, starts, lo, hi);
  nchunk[b] = (hi - lo + kBlock - 1) / kBlock;
}

__global__ __launch_bounds__(kBlock) void k_chunk_write(int64_t ncb, int64_t nkeys, int num,
                                                       const int32_t* __restrict__ starts,
                                                       const int32_t* __restrict__ nchunk,
                                                       const int32_t* __restrict__ chunk_off, int max_chunks,
                                                       int32_t* __restrict__ chunks, int32_t* __restrict__ total) {
  const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= ncb) return;
  int lo, hi;
  block_range(b, ncb, nkeys, num, starts, lo, hi);
  const int o = chunk_off[b];
  for (int c = 0; c < nchunk[b]; ++c) {
    if (o + c >= max_chunks) break;
    chunks[2 * (o + c)] = lo + c * kBlock;
    chunks[2 * (o + c) + 1] = min(kBlock, hi - lo - c * kBlock);
  }
  if (b == ncb - 1) *total = o + nchunk[b];
}

template <typename R>
__global__ __launch_bounds__(kBlock) void k_gather(int num, int cap, const int32_t* __restrict__ perm,
                                                   fsi_mpm_buffers src, fsi_mpm_buffers dst,
                                                   const int32_t* __restrict__ ids_src, int32_t* __restrict__ ids_dst,
                                                   const uint32_t* __restrict__ keys, uint32_t* __restrict__ keys_out,
                                                   int32_t* __restrict__ perm_out) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= num) return;
  const int s = perm[t];
  const int64_t c = cap;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    ((R*)dst.x)[a * c + t] = ((const R*)src.x)[a * c + s];
    ((R*)dst.v)[a * c + t] = ((const R*)src.v)[a * c + s];
    ((R*)dst.pforce)[a * c + t] = ((const R*)src.pforce)[a * c + s];
  }
#pragma unroll
  for (int a = 0; a < 9; ++a) {
    ((R*)dst.C)[a * c + t] = ((const R*)src.C)[a * c + s];
    ((R*)dst.F)[a * c + t] = ((const R*)src.F)[a * c + s];
  }
  ((R*)dst.mass)[t] = ((const R*)src.mass)[s];
  ((R*)dst.volume)[t] = ((const R*)src.volume)[s];
  if (src.material) {
    ((R*)dst.material)[t] = ((const R*)src.material)[s];
    ((R*)dst.material)[c + t] = ((const R*)src.material)[c + s];
  }
  ((int32_t*)dst.active)[t] = src.active[s];
  ids_dst[t] = ids_src[s];
  if (keys_out) keys_out[t] = keys[s];
  if (perm_out) perm_out[t] = s;
}

// exclusive scan in[0..n) -> out[0..n) (three passes; blocksum holds
// ceil(n / kScanBlock) ints)
int scan_exclusive(int64_t n, const int32_t* in, int32_t* out, int32_t* blocksum, hipStream_t st) {
  const int64_t nblk = (n + kScanBlock - 1) / kScanBlock;
  hipLaunchKernelGGL(k_scan_blocksums, dim3((unsigned)nblk), dim3(kBlock), 0, st, n, in, blocksum);
  FSI_TRY_LAUNCH("mpm_sort scan 1");
  hipLaunchKernelGGL(k_scan_offsets, dim3(1), dim3(kBlock), 0, st, nblk, blocksum);
  FSI_TRY_LAUNCH("mpm_sort scan 2");
  hipLaunchKernelGGL(k_scan_apply, dim3((unsigned)nblk), dim3(kBlock), 0, st, n, in, (const int32_t*)blocksum, out);
  FSI_TRY_LAUNCH("mpm_sort scan 3");
  return 0;
}

template <typename R>
int sort_t(const fsi_mpm_params* p, const fsi_mpm_buffers* src, const fsi_mpm_buffers* dst, const int32_t* ids_src,
           int32_t* ids_dst, void* workspace, uint32_t* keys_out, int32_t* perm_out, int32_t* chunks, int max_chunks,
           int32_t* chunk_count, hipStream_t st) {
  const SortGeom g = make_geom(p);
  const int64_t nk = key_space(g);
  const int n = p->num_particles;
  SortWs w = carve(workspace, nk, n);
  const unsigned zero_blocks = (unsigned)std::min<int64_t>(blocks_for(nk), 4096);
  hipLaunchKernelGGL(k_zero_i32, dim3(zero_blocks), dim3(kBlock), 0, st, nk, w.counts);
  FSI_TRY_LAUNCH("mpm_sort zero");
  if (n == 0) {
    if (chunks) {
      hipLaunchKernelGGL(k_zero_i32, dim3(1), dim3(kBlock), 0, st, (int64_t)1, chunk_count);
      FSI_TRY_LAUNCH("mpm_sort chunk count");
    }
    return 0;
  }
  hipLaunchKernelGGL(k_keys<R>, dim3(blocks_for(n)), dim3(kBlock), 0, st, g, n, p->capacity, (R)p->inv_dx,
                     (const R*)src->x, src->active, w.keys, w.counts);
  FSI_TRY_LAUNCH("mpm_sort keys");
  if (int rc = scan_exclusive(nk, w.counts, w.starts, w.blocksum, st)) return rc;
  // the counts become the scatter cursors
  hipLaunchKernelGGL(k_zero_i32, dim3(zero_blocks), dim3(kBlock), 0, st, nk, w.counts);
  FSI_TRY_LAUNCH("mpm_sort zero cursor");
  hipLaunchKernelGGL(k_scatter, dim3(blocks_for(n)), dim3(kBlock), 0, st, n, (const uint32_t*)w.keys,
                     (const int32_t*)w.starts, w.counts, w.perm);
  FSI_TRY_LAUNCH("mpm_sort scatter");
  hipLaunchKernelGGL(k_order_segments, dim3(blocks_for(nk)), dim3(kBlock), 0, st, nk, (const int32_t*)w.starts,
                     (const int32_t*)w.counts, w.perm);
  FSI_TRY_LAUNCH("mpm_sort order");
  hipLaunchKernelGGL(k_gather<R>, dim3(blocks_for(n)), dim3(kBlock), 0, st, n, p->capacity,
                     (const int32_t*)w.perm, *src, *dst, ids_src, ids_dst, (const uint32_t*)w.keys, keys_out,
                     perm_out);
  FSI_TRY_LAUNCH("mpm_sort gather");
  if (chunks) {
    const int64_t nkeys = nk - 1;
    const int64_t ncb = chunk_groups(nk);
    hipLaunchKernelGGL(k_chunk_counts, dim3(blocks_for(ncb)), dim3(kBlock), 0, st, ncb, nkeys, n,
                       (const int32_t*)w.starts, w.nchunk);
    FSI_TRY_LAUNCH("mpm_sort chunk counts");
    if (int rc = scan_exclusive(ncb, w.nchunk, w.chunk_off, w.blocksum, st)) return rc;
    hipLaunchKernelGGL(k_chunk_write, dim3(blocks_for(ncb)), dim3(kBlock), 0, st, ncb, nkeys, n,
                       (const int32_t*)w.starts, (const int32_t*)w.nchunk, (const int32_t*)w.chunk_off, max_chunks,
                       chunks, chunk_count);
    FSI_TRY_LAUNCH("mpm_sort chunk write");
  }
  return 0;
}

}  // namespace
}  // namespace fsi

extern "C" {

int fsi_mpm_sort_workspace(const fsi_mpm_params* p, int64_t* bytes) {
  using namespace fsi;
  FSI_REQUIRE(p && bytes, "mpm_sort_workspace: null argument");
  FSI_REQUIRE(p->nx > 0 && p->ny > 0 && p->nz > 0 && p->capacity >= 0, "mpm_sort_workspace: bad dimensions");
  const SortGeom g = make_geom(p);
  const int64_t nk = key_space(g);
  FSI_REQUIRE(nk < ((int64_t)1 << 31), "mpm_sort: %lld keys exceed the 31-bit key space", (long long)nk);
  *bytes = carve(nullptr, nk, p->capacity).bytes;
  return 0;
}

int fsi_mpm_sort(const fsi_mpm_params* p, const fsi_mpm_buffers* src, const fsi_mpm_buffers* dst,
                 const int32_t* ids_src, int32_t* ids_dst, void* workspace, int64_t workspace_bytes,
                 uint32_t* keys_out, int32_t* perm_out, int32_t* chunks, int32_t max_chunks,
                 int32_t* chunk_count, void* stream) {
  using namespace fsi;
  FSI_REQUIRE(p && src && dst && ids_src && ids_dst && workspace, "mpm_sort: null argument");
  FSI_REQUIRE(src->x && src->v && src->C && src->F && src->mass && src->volume && src->active && src->pforce,
              "mpm_sort: null source particle buffer");
  FSI_REQUIRE(dst->x && dst->v && dst->C && dst->F && dst->mass && dst->volume && dst->active && dst->pforce,
              "mpm_sort: null destination particle buffer");
  FSI_REQUIRE(src->x != dst->x && ids_src != ids_dst, "mpm_sort: source and destination must differ");
  FSI_REQUIRE((src->material == nullptr) == (dst->material == nullptr),
              "mpm_sort: a per-particle material plane needs a destination plane (and only then)");
  FSI_REQUIRE(p->num_particles >= 0 && p->num_particles <= p->capacity, "mpm_sort: bad particle count");
  int64_t need = 0;
  if (int rc = fsi_mpm_sort_workspace(p, &need)) return rc;
  FSI_REQUIRE(workspace_bytes >= need, "mpm_sort: workspace of %lld bytes, %lld needed", (long long)workspace_bytes,
              (long long)need);
  if (chunks) {
    FSI_REQUIRE(chunk_count, "mpm_sort: chunks without chunk_count");
    const int64_t ncb = chunk_groups(key_space(make_geom(p)));
    const int64_t bound = (int64_t)p->num_particles / kBlock + std::min<int64_t>(ncb, p->num_particles) + 1;
    FSI_REQUIRE(max_chunks >= bound, "mpm_sort: room for %d chunks, up to %lld needed", max_chunks,
                (long long)bound);
  }
  hipStream_t st = (hipStream_t)stream;
  if (p->real == FSI_REAL_F64)
    return sort_t<double>(p, src, dst, ids_src, ids_dst, workspace, keys_out, perm_out, chunks, max_chunks,
                          chunk_count, st);
  return sort_t<float>(p, src, dst, ids_src, ids_dst, workspace, keys_out, perm_out, chunks, max_chunks,
                       chunk_count, st);
}

}  // extern "C"
