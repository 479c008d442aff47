// Device helpers shared by the VoxelSG kernels (gfx950, wave64).
//
// Every floating-point expression here is written in the reference's operation
// order and the library is compiled with -ffp-contract=off, so products and sums
// round separately exactly as the reference's kernels state them; divisions and
// square roots are IEEE correctly rounded (HIP's default).
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <stdint.h>

namespace vsg {

constexpr int kBlock = 8;                              // SDF_BLOCK_SIZE (globalPramasConfig.h:8)
constexpr int kBucket = 10;                            // HASH_BUCKET_SIZE (globalPramasConfig.h:9)
constexpr int kVox = kBlock * kBlock * kBlock;         // 512 voxels per block
constexpr int kRecWords = 16;                          // bucket record: 16 x u64 = 128 B
constexpr int kKeyBits = 21;                           // per block coordinate in a packed key
constexpr int32_t kKeyMin = -(1 << (kKeyBits - 1));
constexpr int32_t kKeyMax = (1 << (kKeyBits - 1)) - 1;
constexpr uint32_t kNoClaim = 0xFFFFFFFFu;

// ---- views passed to kernels by value -------------------------------------------

// Hash table: bucket b is a 128-byte record at rec + 16*b holding 10 packed keys
// (u64 words 0..9) and 10 block pointers (i32, bytes 80..119).  cnt[b] is the
// number of used slots; slots fill in order and are never freed, so the used slots
// are always slots [0, cnt[b]).  claim[b] is the per-frame insertion claim word.
// blk[p] = (x, y, z, hash entry index) of the block with pointer p.
// occ is the block occupancy map: one 64-bit word per super-block of 4x4x4 blocks,
// hashed into occ_mask + 1 words; bit occ_bit(x, y, z) of the word of block
// (x, y, z)'s super-block is set once the block is allocated.  A hash collision
// only sets extra bits, so a clear bit proves the block absent.  hyp is the same one
// level up: one word per hyper-block of 4x4x4 super-blocks, bit hyp_bit of a
// super-block set once one of its blocks is allocated.
// idx is the block index the raycaster reads: an open-addressing table of
// idx_mask + 1 16-byte entries {packed key lo, hi, block pointer, 0} (linear
// probing, load factor <= 1/4, empty key all ones), filled when a block is
// inserted.  One probe usually answers a lookup in one memory round trip, where the
// bucket records take the count, the keys and the pointer.
struct TableView {
  uint64_t* rec;
  uint8_t* cnt;
  uint32_t* claim;
  int4* blk;
  uint64_t* occ;
  uint64_t* hyp;
  uint4* idx;
  uint32_t* pcnt;    // per block and 4^3-voxel cell: voxels with weight > 0 and SDF > 0
  uint32_t* idxpos;  // per block: its entry slot in the block index
  uint32_t num_buckets;
  uint32_t num_blocks;
  uint32_t occ_mask;
  uint32_t hyp_mask;
  uint32_t idx_mask;
};

constexpr uint32_t kIdxEmpty = 0xFFFFFFFFu;  // both key words of an empty index entry

// Per-frame counter set (two sets alternate between frames so that each frame's
// scan kernel can clear the next frame's set without an extra launch).
struct FrameCounters {
  uint32_t valid_pixels;
  uint32_t claim_pixels;
  uint32_t winners;
  uint32_t heap_fail;
  uint32_t out_of_range;
  uint32_t visible;
  uint32_t updated;
  uint32_t nonfinite;
  uint32_t zmin_key;  // ordered-uint encoding of the smallest / largest camera z
  uint32_t zmax_key;  // of the frame's valid points
  uint32_t claimed_buckets;
  uint32_t pad[5];
};

struct DevState {
  uint32_t n_alloc;       // live blocks (block pointers are 0..n_alloc-1)
  uint32_t n_alloc_prev;  // live blocks before the current frame's insertions
  uint32_t frames;
  uint32_t n_alloc_next;  // live blocks after them (made live by the compaction)
  FrameCounters set[2];
};

struct FrameArgs {
  float M[16];  // row-major camera->world
  float fx, fy, cx, cy;
  float trunc;
  float voxel;
  int32_t W, H;
  int32_t cset;     // counter set of this frame
  int32_t cull_ok;  // pose rotation orthonormal enough for conservative culling
  double Md[16];    // the same pose in double for the culling test
};

// ---- exact float helpers (mirroring the reference's expressions) -------------------

// float->int like the hardware conversion: saturating, NaN -> 0.
__device__ __forceinline__ int32_t cvt_i32(float x) {
  // v_cvt_i32_f32 truncates toward zero, saturates out-of-range values and maps NaN
  // to 0: the conversion wanted, in one instruction
  int32_t r;
  asm("v_cvt_i32_f32 %0, %1" : "=v"(r) : "v"(x));
  return r;
}

__device__ __forceinline__ float sgnf(float q) { return q > 0.0f ? 1.0f : (q < 0.0f ? -1.0f : 0.0f); }

// Allocator world->voxel (VoxelScene.cu:78-89): (int)(q + 0.5*sign(q)), q = p/v.
__device__ __forceinline__ int32_t alloc_voxel_axis(float p, float v) {
  float q = p / v;
  return cvt_i32(q + sgnf(q) * 0.5f);
}

// Raycaster world->voxel (RayCastRender.cu:17-25): floor(q + 0.5*sign(q)).
__device__ __forceinline__ int32_t ray_voxel_of_q(float q) {
  // q + 0.5 sign(q), with sign(+-0) = 0 (a NaN stays NaN and converts to 0)
  return cvt_i32(floorf(q == 0.0f ? q : q + copysignf(0.5f, q)));
}

// virtualVoxelPosToSDFBlock (VoxelScene.cu:91-95, RayCastRender.cu:28-34).
__device__ __forceinline__ int32_t voxel_to_block_axis(int32_t vi) {
  if (vi < 0) vi = (int32_t)((uint32_t)vi - (uint32_t)(kBlock - 1));
  return vi / kBlock;
}

// virtualVoxelPosToLocal (RayCastRender.cu:37-42).
__device__ __forceinline__ int32_t voxel_local_axis(int32_t vi) { return ((vi % kBlock) + kBlock) % kBlock; }

// hashBlockCoordinate (VoxelScene.cu:61-67): 32-bit wrap-around.
__device__ __forceinline__ uint32_t hash_bucket(int32_t x, int32_t y, int32_t z, uint32_t nb) {
  uint32_t h = ((uint32_t)x * 73856093u) ^ ((uint32_t)y * 19349663u) ^ ((uint32_t)z * 83492791u);
  return h % nb;
}

__device__ __forceinline__ bool key_in_range(int32_t x, int32_t y, int32_t z) {
  return x >= kKeyMin && x <= kKeyMax && y >= kKeyMin && y <= kKeyMax && z >= kKeyMin && z <= kKeyMax;
}

__device__ __forceinline__ uint64_t pack_key(int32_t x, int32_t y, int32_t z) {
  const uint64_t m = (1ull << kKeyBits) - 1ull;
  return (((uint64_t)(uint32_t)x & m) << (2 * kKeyBits)) | (((uint64_t)(uint32_t)y & m) << kKeyBits) |
         ((uint64_t)(uint32_t)z & m);
}

// Row-major rigid transform, each row summed left to right (VoxelScene.cu:545-549).
__device__ __forceinline__ float3 xform_point(const float* M, float x, float y, float z) {
  return make_float3(M[0] * x + M[1] * y + M[2] * z + M[3], M[4] * x + M[5] * y + M[6] * z + M[7],
                     M[8] * x + M[9] * y + M[10] * z + M[11]);
}

// Pointer of block (x,y,z), or -1.  Compares the used slots of the home bucket:
// keys are unique per bucket, so this is the reference's 10-slot scan
// (RayCastRender.cu:81-111).  The slot count and the ten keys are read together
// (then the matching pointer): two round trips to memory instead of three
// dependent ones.  Slots at or beyond the count are ignored whatever they hold.
__device__ __forceinline__ int32_t table_lookup(const TableView& t, int32_t x, int32_t y, int32_t z) {
  if (!key_in_range(x, y, z)) return -1;
  const uint32_t h = hash_bucket(x, y, z, t.num_buckets);
  const uint64_t* r = t.rec + (size_t)h * kRecWords;
  const uint4* r4 = reinterpret_cast<const uint4*>(r);
  uint4 w[5];
#pragma unroll
  for (int i = 0; i < 5; ++i) w[i] = r4[i];
  const int c = t.cnt[h];
  const uint64_t key = pack_key(x, y, z);
  const uint32_t klo = (uint32_t)key, khi = (uint32_t)(key >> 32);
  int hit = -1;
#pragma unroll
  for (int j = 0; j < kBucket; ++j) {
    const uint32_t lo = (j & 1) ? w[j >> 1].z : w[j >> 1].x, hi = (j & 1) ? w[j >> 1].w : w[j >> 1].y;
    if (j < c && lo == klo && hi == khi) hit = j;
  }
  if (hit < 0) return -1;
  return reinterpret_cast<const int32_t*>(r + kBucket)[hit];
}

// 32-bit hash of a block or super-block coordinate: odd multipliers per axis, then
// the MurmurHash3 32-bit finaliser (every output bit depends on every input bit), so
// spatially clustered keys spread over the table.  32-bit arithmetic keeps it cheap
// in the ray march, which hashes on every block lookup.
__device__ __forceinline__ uint32_t hash3(int32_t x, int32_t y, int32_t z) {
  uint32_t h = (uint32_t)x * 0x9E3779B1u + (uint32_t)y * 0x85EBCA77u + (uint32_t)z * 0xC2B2AE3Du;
  h ^= h >> 16;
  h *= 0x85EBCA6Bu;
  h ^= h >> 13;
  h *= 0xC2B2AE35u;
  h ^= h >> 16;
  return h;
}

// Occupancy map word of super-block (sx, sy, sz) and bit of block (x, y, z) in the
// word of its super-block (x >> 2, y >> 2, z >> 2).
__device__ __forceinline__ uint32_t occ_index(int32_t sx, int32_t sy, int32_t sz, uint32_t mask) {
  return hash3(sx, sy, sz) & mask;
}

__device__ __forceinline__ uint32_t occ_bit(int32_t x, int32_t y, int32_t z) {
  return (uint32_t)(x & 3) | ((uint32_t)(y & 3) << 2) | ((uint32_t)(z & 3) << 4);
}

// Hyper-block map word of hyper-block (hx, hy, hz) and bit of super-block (sx, sy, sz)
// in the word of its hyper-block (sx >> 2, sy >> 2, sz >> 2).
__device__ __forceinline__ uint32_t hyp_index(int32_t hx, int32_t hy, int32_t hz, uint32_t mask) {
  return hash3(hx, hy, hz) & mask;
}

__device__ __forceinline__ uint32_t hyp_bit(int32_t sx, int32_t sy, int32_t sz) { return occ_bit(sx, sy, sz); }

// Home slot of block (x, y, z) in the block index.
__device__ __forceinline__ uint32_t idx_home(int32_t x, int32_t y, int32_t z, uint32_t mask) {
  return hash3(x, y, z) & mask;
}

// Block index entry: {key lo, key hi, pointer, positive-cell mask}.  Bit c of the
// mask is set when cell c of the block (4^3 voxels, c = (x >> 2) | (y >> 2) << 1 |
// (z >> 2) << 2 in voxel coordinates within the block) holds a voxel with weight > 0
// and SDF > 0; the integration keeps it exact (per-cell counts in TableView::pcnt).

// Entry slot of block (x, y, z) in the block index, or -1.
__device__ __forceinline__ int32_t index_slot(const TableView& t, int32_t x, int32_t y, int32_t z) {
  if (!key_in_range(x, y, z)) return -1;
  const uint64_t key = pack_key(x, y, z);
  const uint32_t klo = (uint32_t)key, khi = (uint32_t)(key >> 32);
  uint32_t s = idx_home(x, y, z, t.idx_mask);
  for (uint32_t n = 0; n <= t.idx_mask; ++n) {
    const uint4 e = t.idx[s];
    if (e.x == klo && e.y == khi) return (int32_t)s;
    if (e.x == kIdxEmpty && e.y == kIdxEmpty) return -1;
    s = (s + 1u) & t.idx_mask;
  }
  return -1;
}

// Pointer of block (x, y, z) from the block index, or -1; *mask its positive-cell
// mask (0 when absent).
__device__ __forceinline__ int32_t index_lookup(const TableView& t, int32_t x, int32_t y, int32_t z,
                                                uint32_t* mask) {
  *mask = 0u;
  if (!key_in_range(x, y, z)) return -1;
  const uint64_t key = pack_key(x, y, z);
  const uint32_t klo = (uint32_t)key, khi = (uint32_t)(key >> 32);
  uint32_t s = idx_home(x, y, z, t.idx_mask);
  for (uint32_t n = 0; n <= t.idx_mask; ++n) {
    const uint4 e = t.idx[s];
    if (e.x == klo && e.y == khi) {
      *mask = e.w;
      return (int32_t)e.z;
    }
    if (e.x == kIdxEmpty && e.y == kIdxEmpty) return -1;
    s = (s + 1u) & t.idx_mask;
  }
  return -1;
}

__device__ __forceinline__ int32_t index_lookup(const TableView& t, int32_t x, int32_t y, int32_t z) {
  uint32_t m;
  return index_lookup(t, x, y, z, &m);
}

// ---- ordered float encoding for atomicMin/atomicMax ----------------------------------
__device__ __forceinline__ uint32_t float_to_ordered(float f) {
  uint32_t u = __float_as_uint(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

// ---- wave64 reductions (call with the whole wave converged) -------------------------
__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
__device__ __forceinline__ uint32_t wave_min(uint32_t v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    uint32_t w = __shfl_xor(v, o, 64);
    v = w < v ? w : v;
  }
  return v;
}
__device__ __forceinline__ uint32_t wave_max(uint32_t v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    uint32_t w = __shfl_xor(v, o, 64);
    v = w > v ? w : v;
  }
  return v;
}

// Inclusive prefix sum over the 64 lanes of a wave.
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    uint32_t w = __shfl_up(v, o, 64);
    if (lane >= o) v += w;
  }
  return v;
}

// Exclusive prefix sum over a workgroup of NT threads (NT multiple of 64); also
// returns the workgroup total.  Every thread of the workgroup must call it.
template <int NT> __device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t* total) {
  __shared__ uint32_t wsum[NT / 64];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  uint32_t inc = wave_incl_scan(v);
  if (lane == 63) wsum[wid] = inc;
  __syncthreads();
  uint32_t base = 0, tot = 0;
#pragma unroll
  for (int w = 0; w < NT / 64; ++w) {
    uint32_t s = wsum[w];
    if (w < wid) base += s;
    tot += s;
  }
  __syncthreads();
  *total = tot;
  return base + inc - v;
}

}  // namespace vsg
