// VoxelSG depth-fusion oracle: a CPU restatement of the reference algorithm.
//
// TEST INFRASTRUCTURE ONLY.  Nothing in the product path links, loads or calls
// this code; only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
// leg use it, and only as the checker (or as the timed CPU baseline).
//
// It restates, operation by operation, what the reference's CUDA kernels do
// (all citations are relative to the reference repository root):
//   hash ........................ VoxelSG/src/VoxelScene.cu:61-67, RayCastRender.cu:47-53
//   world->block (allocator) .... VoxelSG/src/VoxelScene.cu:73-96
//   bucket insert (allocBlock) .. VoxelSG/src/VoxelScene.cu:137-226
//   heap / table init ........... VoxelSG/src/VoxelScene.cu:250-284,1166-1236
//   band DDA .................... VoxelSG/src/VoxelScene.cu:295-481
//   per-pixel allocation ........ VoxelSG/src/VoxelScene.cu:487-633
//   compaction .................. VoxelSG/src/VoxelScene.cu:1275-1300
//   voxel integration ........... VoxelSG/src/VoxelScene.cu:918-1146
//   raycast ..................... VoxelSG/src/RayCastRender.cu:13-517
//   surface-point extraction .... VoxelSG/src/RayCastRender.cu:629-868, RayCastRender.cpp:307-364
//   depth map from points ....... VoxelSG/src/CustomDepthMapGenerator.cpp:81-222,550-587
//   depth-range rule ............ VoxelSG/src/voxelSG.cpp:594-608
//
// Conventions this oracle fixes where the reference leaves them open:
//   * Allocation runs as a sequential loop over pixels in row-major order and,
//     per pixel, over DDA steps in order.  In the reference the winner of a
//     contended bucket is whichever GPU thread takes the bucket lock first; a
//     sequential loop makes the first request in (pixel, step) order the
//     winner.  Heap pops happen in the same order, so block pointers are
//     deterministic too.
//   * No floating-point contraction: every a*b+c is a rounded product and a
//     rounded sum (the build compiles this file with -ffp-contract=off).
//   * float->int conversions saturate and map NaN to 0 (what the GPU's
//     conversion instruction does; C++'s own cast is undefined there).
//   * Surface points come out in hash-entry order (bucket, then slot), then voxel
//     order within the block: the order a sequential run of the reference's
//     one-thread-per-entry kernel produces (its GPU order is that of atomic
//     arrival).  At most max_points are kept: the first ones in that order.
//
// Templated on the arithmetic type R: R=float is the parity oracle (bit-exact
// with the GPU path), R=double is the precision oracle.  The voxel store can
// hold the SDF as FP32 or as FP16 (with FP32/FP64 arithmetic around it).
#pragma once

#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace vsgo {

constexpr int kBlock = 8;     // SDF_BLOCK_SIZE (VoxelSG/include/globalPramasConfig.h:8)
constexpr int kBucket = 10;   // HASH_BUCKET_SIZE (VoxelSG/include/globalPramasConfig.h:9)
constexpr int kVox = kBlock * kBlock * kBlock;

template <class R> struct V3 { R x, y, z; };
struct I3 {
  int32_t x, y, z;
  bool operator==(const I3& o) const { return x == o.x && y == o.y && z == o.z; }
};

// Saturating conversion with NaN -> 0, as the GPU's v_cvt_i32_f32 does.
template <class R> inline int32_t cvt_i32(R x) {
  if (x != x) return 0;
  if (x >= R(2147483648.0)) return INT32_MAX;
  if (x <= R(-2147483648.0)) return INT32_MIN;
  return (int32_t)x;
}

template <class R> inline R sgn(R q) { return q > R(0) ? R(1) : (q < R(0) ? R(-1) : R(0)); }

// ---- IEEE binary16 conversion (round to nearest even), used by the FP16 store.
inline uint16_t f32_to_f16(float f) {
  uint32_t x;
  std::memcpy(&x, &f, 4);
  uint32_t sign = (x >> 16) & 0x8000u;
  uint32_t ax = x & 0x7fffffffu;
  if (ax >= 0x7f800000u) {  // inf or NaN
    return (uint16_t)(sign | 0x7c00u | (ax > 0x7f800000u ? 0x200u | ((ax >> 13) & 0x3ffu) : 0u));
  }
  if (ax >= 0x477ff000u) return (uint16_t)(sign | 0x7c00u);  // rounds to >= 65520 -> inf
  if (ax < 0x38800000u) {                                     // result is subnormal or zero
    if (ax < 0x33000000u) return (uint16_t)sign;              // < 2^-25 rounds to 0
    // value = m * 2^(e-150) with the implicit bit in m; the half subnormal unit
    // is 2^-24, so the quotient is m >> (126 - e), rounded to nearest even.
    uint32_t e = ax >> 23;
    uint32_t m = (ax & 0x7fffffu) | 0x800000u;
    uint32_t shift = 126u - e;
    uint32_t q = m >> shift;
    uint32_t rem = m & ((1u << shift) - 1u);
    uint32_t half = 1u << (shift - 1u);
    if (rem > half || (rem == half && (q & 1u))) q++;
    return (uint16_t)(sign | q);
  }
  uint32_t e = (ax >> 23) - 112u;  // rebias 127 -> 15
  uint32_t m = ax & 0x7fffffu;
  uint32_t q = (e << 10) | (m >> 13);
  uint32_t rem = m & 0x1fffu;
  if (rem > 0x1000u || (rem == 0x1000u && (q & 1u))) q++;
  return (uint16_t)(sign | q);
}

inline float f16_to_f32(uint16_t h) {
  uint32_t sign = (uint32_t)(h & 0x8000u) << 16;
  uint32_t e = (h >> 10) & 0x1fu;
  uint32_t m = h & 0x3ffu;
  uint32_t x;
  if (e == 0) {
    if (m == 0) {
      x = sign;
    } else {  // subnormal: normalise
      int k = -1;
      do { k++; m <<= 1; } while (!(m & 0x400u));
      x = sign | ((uint32_t)(112 - k) << 23) | ((m & 0x3ffu) << 13);
    }
  } else if (e == 31) {
    x = sign | 0x7f800000u | (m << 13);
  } else {
    x = sign | ((e + 112u) << 23) | (m << 13);
  }
  float f;
  std::memcpy(&f, &x, 4);
  return f;
}

// ---- A2: bucket hash, 32-bit wrap-around (VoxelScene.cu:61-67 == RayCastRender.cu:47-53)
inline uint32_t hash_bucket(I3 b, uint32_t num_buckets) {
  uint32_t h = ((uint32_t)b.x * 73856093u) ^ ((uint32_t)b.y * 19349663u) ^ ((uint32_t)b.z * 83492791u);
  return h % num_buckets;
}

// ---- Q1: allocator world->voxel->block (VoxelScene.cu:73-96)
template <class R> inline int32_t alloc_voxel_axis(R p, R v) {
  R q = p / v;
  return cvt_i32(q + sgn(q) * R(0.5f));  // (int) truncates toward zero
}
inline int32_t voxel_to_block_axis(int32_t vi) {
  if (vi < 0) vi = (int32_t)((uint32_t)vi - (uint32_t)(kBlock - 1));
  return vi / kBlock;  // C division truncates toward zero
}
template <class R> inline I3 alloc_world_to_block(V3<R> p, R v) {
  return I3{voxel_to_block_axis(alloc_voxel_axis(p.x, v)), voxel_to_block_axis(alloc_voxel_axis(p.y, v)),
            voxel_to_block_axis(alloc_voxel_axis(p.z, v))};
}

// ---- Q2: raycaster world->voxel (RayCastRender.cu:17-25), block (:28-34), local (:37-42)
template <class R> inline int32_t ray_voxel_axis(R p, R v) {
  R q = p / v;
  return cvt_i32(std::floor(q + R(0.5f) * sgn(q)));
}
inline int32_t voxel_local_axis(int32_t vi) { return ((vi % kBlock) + kBlock) % kBlock; }

// rigid transform with a row-major 4x4 camera->world matrix, each row summed left to right
template <class R> inline V3<R> xform_point(const R* M, V3<R> p) {
  return V3<R>{M[0] * p.x + M[1] * p.y + M[2] * p.z + M[3], M[4] * p.x + M[5] * p.y + M[6] * p.z + M[7],
               M[8] * p.x + M[9] * p.y + M[10] * p.z + M[11]};
}

// ---- Voxel hash state (VoxelSG/include/CUDAHashRef.h:11-79, CUDAHashRef.cu:40-152)
template <class R> struct Scene {
  using Real = R;
  uint32_t num_buckets = 0;  // Params::hashSlotNum
  uint32_t num_blocks = 0;   // Params::SDFBlockNum
  R voxel = R(0);            // Params::voxelSize
  int store_fp16 = 0;        // 0: FP32 SDF, 1: FP16 SDF
  int directional = 0;       // directional TSDF extension (two direction slots per voxel)

  std::vector<I3> slot_pos;      // T = B*10 (HashSlot::pos)
  std::vector<int32_t> slot_ptr; // T (HashSlot::ptr, -1 = free)
  std::vector<uint32_t> heap;    // N+1
  uint32_t heap_counter = 0;
  std::vector<int32_t> mutex;    // B
  std::vector<uint8_t> heap_failed;  // B: bucket saw an insertion fail for lack of blocks this frame

  // voxel store, indexed ptr*512 + voxel, grown on demand (ptr < num_blocks)
  std::vector<R> sdf;
  std::vector<uint16_t> sdf_h;
  std::vector<uint8_t> weight;
  std::vector<uint8_t> color;  // 3 per voxel
  std::vector<uint8_t> flags;  // bit0 isUpdated, bit1 isZeroCrossing
  // directional extension, two slots per voxel (index 2*voxel + slot)
  std::vector<uint8_t> dir_id;  // 0 free, 1..6 = +x, -x, +y, -y, +z, -z
  std::vector<uint8_t> dir_w;
  std::vector<R> dtsdf;
  uint32_t frames = 0;

  // per-frame counters
  // c_heap_fail: buckets whose insertion failed for lack of blocks in the last frame
  uint32_t c_valid_pixels = 0, c_new_blocks = 0, c_heap_fail = 0, c_updated = 0, c_visible = 0;
  // wall-clock seconds of the last integrate's stages: allocation, compaction, integration
  double t_alloc = 0.0, t_compact = 0.0, t_integrate = 0.0;

  void init(uint32_t B, uint32_t N, R v, int fp16) {
    num_buckets = B;
    num_blocks = N;
    voxel = v;
    store_fp16 = fp16;
    reset();
  }

  // resetHashDataCUDA (VoxelScene.cu:1166-1236): heap[i] = N-1-i, counter = N-1,
  // table ptr=-1 / pos=0, mutexes 0; voxel memory zero (CUDAHashRef.cu:86-93).
  void reset() {
    slot_pos.assign((size_t)num_buckets * kBucket, I3{0, 0, 0});
    slot_ptr.assign((size_t)num_buckets * kBucket, -1);
    heap.assign((size_t)num_blocks + 1, 0u);
    for (uint32_t i = 0; i < num_blocks; ++i) heap[i] = num_blocks - i - 1;
    heap_counter = num_blocks - 1;
    mutex.assign(num_buckets, 0);
    heap_failed.assign(num_buckets, 0);
    sdf.clear();
    sdf_h.clear();
    weight.clear();
    color.clear();
    flags.clear();
    dir_id.clear();
    dir_w.clear();
    dtsdf.clear();
    frames = 0;
  }

  void ensure_block(int32_t ptr) {
    size_t need = ((size_t)ptr + 1) * kVox;
    if (weight.size() >= need) return;
    if (store_fp16) sdf_h.resize(need, 0); else sdf.resize(need, R(0));
    weight.resize(need, 0);
    color.resize(need * 3, 0);
    flags.resize(need, 0);
    if (directional) {
      dir_id.resize(2 * need, 0);
      dir_w.resize(2 * need, 0);
      dtsdf.resize(2 * need, R(0));
    }
  }

  R load_sdf(size_t i) const { return store_fp16 ? R(f16_to_f32(sdf_h[i])) : sdf[i]; }
  void store_sdf(size_t i, R s) {
    if (store_fp16) sdf_h[i] = f32_to_f16((float)s); else sdf[i] = s;
  }

  // ---- A6: allocBlock (VoxelScene.cu:137-226), executed sequentially
  int32_t alloc_block(I3 c) {
    uint32_t h = hash_bucket(c, num_buckets);
    size_t hp = (size_t)h * kBucket;
    long first_empty = -1;
    for (int j = 0; j < kBucket; ++j) {
      size_t i = hp + j;
      if (slot_pos[i] == c && slot_ptr[i] != -1) return slot_ptr[i];
      if (first_empty == -1 && slot_ptr[i] == -1) first_empty = (long)i;
    }
    if (first_empty != -1) {
      int prev = mutex[h];
      mutex[h] = 1;
      if (prev != 1) {
        uint32_t heap_index = heap_counter;  // atomicSub returns the old value
        heap_counter -= 1u;
        if (heap_index <= num_blocks && heap_index != 0xFFFFFFFFu) {
          uint32_t block_index = heap[heap_index];
          slot_pos[first_empty] = c;
          slot_ptr[first_empty] = (int32_t)block_index;
          ensure_block((int32_t)block_index);
          c_new_blocks++;
          return (int32_t)block_index;  // the bucket lock stays taken (VoxelScene.cu:201-203)
        }
        heap_counter += 1u;  // out of blocks: restore, release the lock
        mutex[h] = 0;
        if (!heap_failed[h]) {  // counted once per bucket and frame
          heap_failed[h] = 1;
          c_heap_fail++;
        }
        return -1;
      }
      return -1;
    }
    return -1;
  }

  size_t num_allocated() const { return (size_t)(num_blocks - 1u - heap_counter); }

  // ---- streaming (VoxelStreamingManager.cu:36-91,258-358,443-719) --------------
  // One streamed block: coordinate and voxels (weight widened to u16 as the GPU
  // store keeps it; colour rgb + flags; direction slots when the extension is on).
  struct Streamed {
    I3 c;
    std::vector<R> sdf;
    std::vector<uint16_t> weight;
    std::vector<uint8_t> rgbf;  // 4 per voxel
    std::vector<uint8_t> dir_id, dir_w;
    std::vector<R> dtsdf;
  };

  // Stream-out: live entries in table order (the compacted list, VoxelScene.cu:
  // 1275-1300) whose world position (SDFBlockToWorld: coordinate * 8 * voxel,
  // VoxelScene.cu:101-107) lies at distance >= radius from c
  // (VoxelStreamingManager.cu:55-66), at most cap: voxels copied out then zeroed,
  // slot freed (ptr -1, pos 0), block index pushed on the heap (:279-356).
  std::vector<Streamed> stream_out(V3<R> c, R radius, size_t cap) {
    std::vector<Streamed> out;
    for (size_t e = 0; e < slot_ptr.size() && out.size() < cap; ++e) {
      const int32_t ptr = slot_ptr[e];
      if (ptr == -1) continue;
      const I3 b = slot_pos[e];
      const R dx = R(b.x * kBlock) * voxel - c.x;
      const R dy = R(b.y * kBlock) * voxel - c.y;
      const R dz = R(b.z * kBlock) * voxel - c.z;
      if (!(std::sqrt(dx * dx + dy * dy + dz * dz) >= radius)) continue;
      ensure_block(ptr);
      Streamed sb;
      sb.c = b;
      for (int v = 0; v < kVox; ++v) {
        const size_t i = (size_t)ptr * kVox + v;
        sb.sdf.push_back(load_sdf(i));
        sb.weight.push_back(weight[i]);
        for (int ch = 0; ch < 3; ++ch) sb.rgbf.push_back(color[3 * i + ch]);
        sb.rgbf.push_back(flags[i]);
        store_sdf(i, R(0));
        weight[i] = 0;
        color[3 * i] = color[3 * i + 1] = color[3 * i + 2] = 0;
        flags[i] = 0;
        if (directional) {
          for (int k = 0; k < 2; ++k) {
            sb.dir_id.push_back(dir_id[2 * i + k]);
            sb.dir_w.push_back(dir_w[2 * i + k]);
            sb.dtsdf.push_back(dtsdf[2 * i + k]);
            dir_id[2 * i + k] = dir_w[2 * i + k] = 0;
            dtsdf[2 * i + k] = R(0);
          }
        }
      }
      slot_ptr[e] = -1;
      slot_pos[e] = I3{0, 0, 0};
      heap_counter += 1u;  // atomicAdd returns the old top; the block goes above it
      heap[heap_counter] = (uint32_t)ptr;
      out.push_back(std::move(sb));
    }
    return out;
  }

  // Stream-in of one block, inputs taken in order: 4 coordinate beyond the GPU
  // store's key range (+-2^20 blocks), 1 already allocated (nothing written), 2 home
  // bucket full (the reference probes on into other buckets,
  // VoxelStreamingManager.cu:531-596, where its lookups never look), 3 heap empty,
  // else 0: first free slot of the home bucket, block popped from the heap as
  // allocBlock pops it (VoxelScene.cu:183-214), voxels written.
  int stream_in(const Streamed& sb) {
    const I3 b = sb.c;
    const int32_t lim = 1 << 20;
    if (b.x < -lim || b.x >= lim || b.y < -lim || b.y >= lim || b.z < -lim || b.z >= lim) return 4;
    const uint32_t h = hash_bucket(b, num_buckets);
    const size_t hp = (size_t)h * kBucket;
    long first_empty = -1;
    for (int j = 0; j < kBucket; ++j) {
      const size_t i = hp + j;
      if (slot_ptr[i] != -1 && slot_pos[i] == b) return 1;
      if (first_empty == -1 && slot_ptr[i] == -1) first_empty = (long)i;
    }
    if (first_empty == -1) return 2;
    const uint32_t heap_index = heap_counter;
    if (!(heap_index <= num_blocks && heap_index != 0xFFFFFFFFu)) return 3;
    heap_counter -= 1u;
    const uint32_t ptr = heap[heap_index];
    slot_pos[first_empty] = b;
    slot_ptr[first_empty] = (int32_t)ptr;
    ensure_block((int32_t)ptr);
    for (int v = 0; v < kVox; ++v) {
      const size_t i = (size_t)ptr * kVox + v;
      store_sdf(i, sb.sdf[v]);
      weight[i] = (uint8_t)sb.weight[v];
      for (int ch = 0; ch < 3; ++ch) color[3 * i + ch] = sb.rgbf[4 * v + ch];
      flags[i] = sb.rgbf[4 * v + 3];
      if (directional && !sb.dir_id.empty()) {
        for (int k = 0; k < 2; ++k) {
          dir_id[2 * i + k] = sb.dir_id[2 * v + k];
          dir_w[2 * i + k] = sb.dir_w[2 * v + k];
          dtsdf[2 * i + k] = sb.dtsdf[2 * v + k];
        }
      }
    }
    return 0;
  }
};

// ---- A5: block DDA from block(rayStart) to block(rayEnd) (VoxelScene.cu:295-481).
// Calls visit(blockCoord, iteration) for every visited block, in order.
template <class R, class F> inline void dda_band(V3<R> a, V3<R> b, R v, F&& visit) {
  I3 cur = alloc_world_to_block(a, v);
  I3 end = alloc_world_to_block(b, v);
  if (cur == end) return;  // Q6: start == end allocates nothing (VoxelScene.cu:345-348)
  V3<R> d{b.x - a.x, b.y - a.y, b.z - a.z};
  R len = std::sqrt(d.x * d.x + d.y * d.y + d.z * d.z);
  if (len < R(1e-6f)) return;
  d.x /= len;
  d.y /= len;
  d.z /= len;
  V3<R> step{d.x > R(0) ? R(1) : R(-1), d.y > R(0) ? R(1) : R(-1), d.z > R(0) ? R(1) : R(-1)};
  R ext = R(kBlock) * v;
  V3<R> sc{step.x > R(0) ? R(1) : R(0), step.y > R(0) ? R(1) : R(0), step.z > R(0) ? R(1) : R(0)};
  R hv = R(0.5f) * v;
  V3<R> bnd{((R)cur.x + sc.x) * ext - hv, ((R)cur.y + sc.y) * ext - hv, ((R)cur.z + sc.z) * ext - hv};
  V3<R> tmax{(bnd.x - a.x) / d.x, (bnd.y - a.y) / d.y, (bnd.z - a.z) / d.z};
  V3<R> tdel{(step.x * ext) / d.x, (step.y * ext) / d.y, (step.z * ext) / d.z};
  if (d.x == R(0)) { tmax.x = R(1e6f); tdel.x = R(1e6f); }
  if (d.y == R(0)) { tmax.y = R(1e6f); tdel.y = R(1e6f); }
  if (d.z == R(0)) { tmax.z = R(1e6f); tdel.z = R(1e6f); }
  const int32_t sx = (int32_t)step.x, sy = (int32_t)step.y, sz = (int32_t)step.z;
  for (int iter = 0; iter < 1024; ++iter) {
    visit(cur, iter);
    if (cur == end) break;
    if (tmax.x < tmax.y && tmax.x < tmax.z) {
      tmax.x += tdel.x;
      cur.x += sx;
    } else if (tmax.y < tmax.z) {
      tmax.y += tdel.y;
      cur.y += sy;
    } else {
      tmax.z += tdel.z;
      cur.z += sz;
    }
  }
}

// Truncation band of one pixel in world space (VoxelScene.cu:527-605).  Returns
// false when the reference's kernel returns before walking the band.
template <class R> inline bool pixel_band(const R* M, V3<R> cam, V3<R> P, R trunc, V3<R>* lo, V3<R>* hi) {
  if (P.z <= R(1e-6f)) return false;
  V3<R> w = xform_point(M, P);
  V3<R> rd{w.x - cam.x, w.y - cam.y, w.z - cam.z};
  R len = std::sqrt(rd.x * rd.x + rd.y * rd.y + rd.z * rd.z);
  if (len < R(1e-6f)) return false;
  *lo = xform_point(M, V3<R>{P.x, P.y, P.z - trunc});
  *hi = xform_point(M, V3<R>{P.x, P.y, P.z + trunc});
  return true;
}

// ---- A4: allocation for one frame, sequential canonical order.
template <class R> inline void alloc_frame(Scene<R>& s, const float* pts, int W, int H, R trunc, const R* M) {
  // resetHashBucketMutexKernel before every allocation (VoxelScene.cu:1444-1449)
  std::fill(s.mutex.begin(), s.mutex.end(), 0);
  std::fill(s.heap_failed.begin(), s.heap_failed.end(), 0);
  s.c_new_blocks = 0;
  s.c_heap_fail = 0;
  s.c_valid_pixels = 0;
  V3<R> cam{M[3], M[7], M[11]};  // cameraPos = translation (VoxelScene.cpp:462-466)
  for (int pix = 0; pix < W * H; ++pix) {
    V3<R> P{(R)pts[3 * pix], (R)pts[3 * pix + 1], (R)pts[3 * pix + 2]};
    if (P.z <= R(1e-6f)) continue;
    s.c_valid_pixels++;
    V3<R> lo, hi;
    if (!pixel_band(M, cam, P, trunc, &lo, &hi)) continue;
    dda_band(lo, hi, s.voxel, [&](I3 c, int) { s.alloc_block(c); });
  }
}

// Conservative per-block test: can any voxel of block c be updated by this frame?
// False only when provably no voxel projects onto an in-image pixel or lies within
// the truncation band's depth slab.  zmin/zmax: camera-z range of valid pixel points.
template <class R>
inline bool block_maybe_visible(const R* M, I3 c, R v, int W, int H, R fx, R fy, R cx, R cy, R trunc, R zmin,
                                R zmax) {
  double lo_z = 1e300, hi_z = -1e300, lo_u = 1e300, hi_u = -1e300, lo_v = 1e300, hi_v = -1e300;
  bool all_front = true;
  for (int k = 0; k < 8; ++k) {
    double wx = ((double)c.x * kBlock + ((k & 1) ? 7 : 0)) * (double)v;
    double wy = ((double)c.y * kBlock + ((k & 2) ? 7 : 0)) * (double)v;
    double wz = ((double)c.z * kBlock + ((k & 4) ? 7 : 0)) * (double)v;
    double rx = wx - (double)M[3], ry = wy - (double)M[7], rz = wz - (double)M[11];
    double X = (double)M[0] * rx + (double)M[4] * ry + (double)M[8] * rz;
    double Y = (double)M[1] * rx + (double)M[5] * ry + (double)M[9] * rz;
    double Z = (double)M[2] * rx + (double)M[6] * ry + (double)M[10] * rz;
    lo_z = Z < lo_z ? Z : lo_z;
    hi_z = Z > hi_z ? Z : hi_z;
    if (Z > 1e-4) {
      double u = X * (double)fx / Z + (double)cx, vv = Y * (double)fy / Z + (double)cy;
      lo_u = u < lo_u ? u : lo_u;
      hi_u = u > hi_u ? u : hi_u;
      lo_v = vv < lo_v ? vv : lo_v;
      hi_v = vv > hi_v ? vv : hi_v;
    } else {
      all_front = false;
    }
  }
  double tz = (double)trunc * 1.001 + 1e-4;
  if (hi_z < (double)zmin - tz || lo_z > (double)zmax + tz) return false;
  if (all_front) {
    if (hi_u < -2.5 || lo_u > (double)W + 0.5 || hi_v < -2.5 || lo_v > (double)H + 0.5) return false;
  }
  return true;
}

// ---- A8: integrate one frame into the listed blocks (VoxelScene.cu:918-1146).
// The normal map is read by the reference but its value never reaches the store
// (VoxelScene.cu:1075-1094), so it is not an input here.
// ---- A9: directional TSDF extension ------------------------------------------------
//
// The compiled reference computes the world normal of every integrated voxel's pixel
// and drops it (VoxelScene.cu:1074-1094); the per-direction fields and knobs exist
// only as comments (CUDAHashRef.h:34-39: dir_id, dir_w, dtsdf for 2 slots;
// globalParamasConfig.cpp:101-111,416: 2 slots, cosine threshold 0.8, weight max 100).
// This extension is the build's own definition of that design, pinned by its own
// tests (no reference behaviour pins it):
//   * direction bin of a unit world normal n: the signed axis of its largest |n_a|
//     (ties to the lower axis), id = 1 + 2a + (n_a < 0): +x, -x, +y, -y, +z, -z;
//   * no update when that |n_a| (the cosine to the bin) is below 0.8 or n is zero;
//   * the slot whose id matches, else the first free slot (weight 0), else none;
//   * dtsdf <- (dtsdf * w + sdf) / (w + 1), w <- min(w + 1, 100), with the truncated
//     SDF of the main update.
template <class R> inline void dir_update(Scene<R>& s, size_t vi, V3<R> n, R sdfv) {
  const R ax = std::fabs(n.x), ay = std::fabs(n.y), az = std::fabs(n.z);
  int a = 0;
  R c = ax;
  if (ay > c) {
    a = 1;
    c = ay;
  }
  if (az > c) {
    a = 2;
    c = az;
  }
  if (!(c >= R(0.8f))) return;
  const R comp = a == 0 ? n.x : (a == 1 ? n.y : n.z);
  const uint8_t id = (uint8_t)(1 + 2 * a + (comp < R(0) ? 1 : 0));
  int slot = -1;
  for (int j = 1; j >= 0; --j)
    if (s.dir_id[2 * vi + j] == id) slot = j;
  if (slot < 0) {
    for (int j = 1; j >= 0; --j)
      if (s.dir_w[2 * vi + j] == 0) slot = j;
  }
  if (slot < 0) return;
  const size_t k = 2 * vi + (size_t)slot;
  const R ow = (R)s.dir_w[k];
  const R nw = ow + R(1);
  s.dtsdf[k] = (s.dtsdf[k] * ow + sdfv * R(1)) / nw;
  s.dir_w[k] = (uint8_t)std::fmin(R(100), nw);
  s.dir_id[k] = id;
}

// World normal of a pixel as the reference computes it (VoxelScene.cu:1074-1094).
template <class R> inline V3<R> pixel_world_normal(const float* nrm, int pix, const R* M) {
  V3<R> nc{(R)nrm[3 * pix], (R)nrm[3 * pix + 1], (R)nrm[3 * pix + 2]};
  R l = std::sqrt(nc.x * nc.x + nc.y * nc.y + nc.z * nc.z);
  if (l > R(0)) {
    nc.x /= l;
    nc.y /= l;
    nc.z /= l;
  }
  V3<R> n{M[0] * nc.x + M[1] * nc.y + M[2] * nc.z, M[4] * nc.x + M[5] * nc.y + M[6] * nc.z,
          M[8] * nc.x + M[9] * nc.y + M[10] * nc.z};
  l = std::sqrt(n.x * n.x + n.y * n.y + n.z * n.z);
  if (l > R(1e-6f)) {
    n.x /= l;
    n.y /= l;
    n.z /= l;
  }
  return n;
}

template <class R>
inline void integrate_block(Scene<R>& s, I3 bc, int32_t ptr, const float* pts, const uint8_t* rgb, int W, int H,
                            R trunc, const R* M, R fx, R fy, R cx, R cy, uint32_t* updated,
                            const float* nrm = nullptr) {
  const R v = s.voxel;
  V3<R> cam{M[3], M[7], M[11]};
  uint32_t upd = 0;
  for (int idx = 0; idx < kVox; ++idx) {
    int lx = idx % kBlock, ly = (idx / kBlock) % kBlock, lz = idx / (kBlock * kBlock);
    V3<R> vw{(R)(bc.x * kBlock + lx) * v, (R)(bc.y * kBlock + ly) * v, (R)(bc.z * kBlock + lz) * v};
    V3<R> rel{vw.x - M[3], vw.y - M[7], vw.z - M[11]};
    V3<R> vc{M[0] * rel.x + M[4] * rel.y + M[8] * rel.z, M[1] * rel.x + M[5] * rel.y + M[9] * rel.z,
             M[2] * rel.x + M[6] * rel.y + M[10] * rel.z};
    R sxf = vc.x * fx / vc.z + cx;
    R syf = vc.y * fy / vc.z + cy;
    int32_t px = cvt_i32(sxf + R(0.5f)), py = cvt_i32(syf + R(0.5f));
    if (!(px >= 0 && px < W && py >= 0 && py < H)) continue;
    int pix = py * W + px;
    V3<R> P{(R)pts[3 * pix], (R)pts[3 * pix + 1], (R)pts[3 * pix + 2]};
    if (!(P.x != R(0) || P.y != R(0) || P.z != R(0))) continue;
    V3<R> D = xform_point(M, P);
    V3<R> sd{vw.x - D.x, vw.y - D.y, vw.z - D.z};
    R sdfv = std::sqrt(sd.x * sd.x + sd.y * sd.y + sd.z * sd.z);
    V3<R> cd{D.x - cam.x, D.y - cam.y, D.z - cam.z};
    R cam_dist = std::sqrt(cd.x * cd.x + cd.y * cd.y + cd.z * cd.z);
    R vox_dist = std::sqrt((vw.x - cam.x) * (vw.x - cam.x) + (vw.y - cam.y) * (vw.y - cam.y) +
                           (vw.z - cam.z) * (vw.z - cam.z));
    if (vox_dist < cam_dist) sdfv = -sdfv;
    if (sdfv < -trunc || sdfv > trunc) continue;
    sdfv = std::fmax(-trunc, std::fmin(trunc, sdfv));
    size_t vi = (size_t)ptr * kVox + idx;
    R old_sdf = s.load_sdf(vi);
    R old_w = (R)s.weight[vi];
    R new_w = old_w + R(1);
    if (new_w > R(0)) {
      R nsdf = (old_sdf * old_w + sdfv * R(1)) / new_w;
      s.store_sdf(vi, nsdf);
      R stored = s.load_sdf(vi);  // isZeroCrossing compares against the stored value
      s.weight[vi] = (uint8_t)std::fmin(R(255), new_w);
      for (int ch = 0; ch < 3; ++ch) {
        R oc = (R)s.color[3 * vi + ch];
        R nc = (oc * old_w + (R)rgb[3 * pix + ch] * R(1)) / new_w;
        s.color[3 * vi + ch] = (uint8_t)nc;
      }
      uint8_t fl = s.flags[vi] | 1u;
      if ((old_sdf > R(0) && stored <= R(0)) || (old_sdf <= R(0) && stored > R(0))) fl |= 2u;
      s.flags[vi] = fl;
      if (s.directional && nrm) dir_update(s, vi, pixel_world_normal(nrm, pix, M), sdfv);
      upd++;
    }
  }
  *updated += upd;
}

// One call of VoxelScene::integrateFromScanData (VoxelScene.cpp:404-833): allocate,
// compact (every allocated block, VoxelScene.cu:1275-1300), integrate.
// visible_only != 0 integrates only blocks passing block_maybe_visible; the
// result must be identical (that is what the test of the culling checks).
template <class R>
inline void integrate_frame(Scene<R>& s, const float* pts, const uint8_t* rgb, int W, int H, R trunc, const R* M,
                            R fx, R fy, R cx, R cy, int visible_only, const float* nrm = nullptr) {
  using clock = std::chrono::steady_clock;
  const auto t0 = clock::now();
  alloc_frame(s, pts, W, H, trunc, M);
  const auto t1 = clock::now();
  // compaction: every live entry, in table order
  std::vector<std::pair<I3, int32_t>> list;
  for (size_t i = 0; i < s.slot_ptr.size(); ++i)
    if (s.slot_ptr[i] != -1) list.push_back({s.slot_pos[i], s.slot_ptr[i]});
  const auto t2 = clock::now();
  R zmin = R(1e30f), zmax = R(-1e30f);
  bool finite = true;
  if (visible_only) {
    for (int pix = 0; pix < W * H; ++pix) {
      R X = (R)pts[3 * pix], Y = (R)pts[3 * pix + 1], Z = (R)pts[3 * pix + 2];
      if (!(X != R(0) || Y != R(0) || Z != R(0))) continue;
      if (!std::isfinite(X) || !std::isfinite(Y) || !std::isfinite(Z)) finite = false;
      zmin = Z < zmin ? Z : zmin;
      zmax = Z > zmax ? Z : zmax;
    }
  }
  std::vector<char> keep(list.size(), 1);
  if (visible_only && finite) {
    for (size_t i = 0; i < list.size(); ++i)
      keep[i] = block_maybe_visible(M, list[i].first, s.voxel, W, H, fx, fy, cx, cy, trunc, zmin, zmax) ? 1 : 0;
  }
  uint32_t updated = 0, visible = 0;
#pragma omp parallel for schedule(dynamic, 16) reduction(+ : updated, visible)
  for (long i = 0; i < (long)list.size(); ++i) {
    if (!keep[i]) continue;
    visible++;
    integrate_block(s, list[i].first, list[i].second, pts, rgb, W, H, trunc, M, fx, fy, cx, cy, &updated, nrm);
  }
  s.c_updated = updated;
  s.c_visible = visible;
  s.frames++;
  const auto t3 = clock::now();
  s.t_alloc = std::chrono::duration<double>(t1 - t0).count();
  s.t_compact = std::chrono::duration<double>(t2 - t1).count();
  s.t_integrate = std::chrono::duration<double>(t3 - t2).count();
}

// ---- A13: getVoxel (RayCastRender.cu:55-121): weight and sdf of the voxel at p.
template <class R> struct VoxelRead {
  R sdf;
  uint8_t w;
  uint8_t c[3];
};

template <class R> inline VoxelRead<R> get_voxel(const Scene<R>& s, V3<R> p, R v) {
  int32_t vx = ray_voxel_axis(p.x, v), vy = ray_voxel_axis(p.y, v), vz = ray_voxel_axis(p.z, v);
  I3 b{voxel_to_block_axis(vx), voxel_to_block_axis(vy), voxel_to_block_axis(vz)};
  int lx = voxel_local_axis(vx), ly = voxel_local_axis(vy), lz = voxel_local_axis(vz);
  uint32_t h = hash_bucket(b, s.num_buckets);
  size_t hp = (size_t)h * kBucket;
  for (int j = 0; j < kBucket; ++j) {
    if (s.slot_ptr[hp + j] == -1) continue;
    if (s.slot_pos[hp + j] == b) {
      size_t vi = (size_t)s.slot_ptr[hp + j] * kVox + (size_t)(lx + ly * kBlock + lz * kBlock * kBlock);
      VoxelRead<R> r;
      r.sdf = s.load_sdf(vi);
      r.w = s.weight[vi];
      r.c[0] = s.color[3 * vi];
      r.c[1] = s.color[3 * vi + 1];
      r.c[2] = s.color[3 * vi + 2];
      return r;
    }
  }
  return VoxelRead<R>{R(0), 0, {0, 0, 0}};
}

template <class R> inline R frac01(R x) { return x - std::floor(x); }

// ---- A12: trilinearInterpolationSimpleFastFast (RayCastRender.cu:127-215)
template <class R> inline bool trilinear(const Scene<R>& s, V3<R> p, R v, R* dist, uint8_t col[3]) {
  const R off = v;
  V3<R> pd{p.x - off / R(2), p.y - off / R(2), p.z - off / R(2)};
  V3<R> vp{pd.x / v, pd.y / v, pd.z / v};
  V3<R> w{frac01(vp.x), frac01(vp.y), frac01(vp.z)};
  R d = R(0), cr = R(0), cg = R(0), cb = R(0), wsum = R(0);
  for (int i = 0; i < 8; ++i) {
    R ax = (i & 1) ? off : R(0), ay = (i & 2) ? off : R(0), az = (i & 4) ? off : R(0);
    VoxelRead<R> vox = get_voxel(s, V3<R>{pd.x + ax, pd.y + ay, pd.z + az}, v);
    if (vox.w == 0) continue;
    R wi = ((i & 1) ? w.x : (R(1) - w.x)) * ((i & 2) ? w.y : (R(1) - w.y)) * ((i & 4) ? w.z : (R(1) - w.z));
    d += wi * vox.sdf;
    cr += wi * (R)vox.c[0];
    cg += wi * (R)vox.c[1];
    cb += wi * (R)vox.c[2];
    wsum += wi;
  }
  if (wsum <= R(1e-6f)) return false;
  R inv = R(1) / wsum;
  *dist = d * inv;
  cr *= inv;
  cg *= inv;
  cb *= inv;
  col[0] = (uint8_t)cr;
  col[1] = (uint8_t)cg;
  col[2] = (uint8_t)cb;
  return true;
}

// ---- A14: three regula-falsi steps (RayCastRender.cu:220-271)
template <class R>
inline bool refine(const Scene<R>& s, V3<R> cam, V3<R> dir, R v, R d0, R r0, R d1, R r1, R* alpha, uint8_t col[3]) {
  R a = r0, ad = d0, b = r1, bd = d1, c = R(0);
  for (int i = 0; i < 3; ++i) {
    c = a + (ad / (ad - bd)) * (b - a);
    R cd;
    if (!trilinear(s, V3<R>{cam.x + c * dir.x, cam.y + c * dir.y, cam.z + c * dir.z}, v, &cd, col)) return false;
    if (ad * cd > R(0)) {
      a = c;
      ad = cd;
    } else {
      b = c;
      bd = cd;
    }
  }
  *alpha = c;
  return true;
}

// ---- A10/A11: one pixel of renderKernel + traverseCoarseGridSimpleSampleAll
// (RayCastRender.cu:276-517).  Also returns the bracket [a, b] (ray parameters and
// sdf values) at which the accepted crossing was refined; the adjoint restarts there.
template <class R> struct MarchResult {
  bool marched = false;  // depth range valid (RayCastRender.cu:461-462)
  bool hit = false;
  R alpha = R(0), d2r = R(1);
  V3<R> cd{R(0), R(0), R(0)}, wd{R(0), R(0), R(0)};
  R a = R(0), ad = R(0), b = R(0), bd = R(0);
  uint8_t rgb[3] = {0, 0, 0};
};

template <class R>
inline MarchResult<R> march_pixel(const Scene<R>& s, int x, int y, V3<R> cam, const R* M, R fx, R fy, R cx, R cy,
                                  R minD, R maxD, R v) {
  const R MINF = R(1e10f);
  MarchResult<R> res;
  V3<R> cd{((R)x - cx) / fx, ((R)y - cy) / fy, R(1)};
  R cl = std::sqrt(cd.x * cd.x + cd.y * cd.y + cd.z * cd.z);
  if (cl > R(1e-6f)) {
    cd.x /= cl;
    cd.y /= cl;
    cd.z /= cl;
  }
  V3<R> wd{M[0] * cd.x + M[1] * cd.y + M[2] * cd.z, M[4] * cd.x + M[5] * cd.y + M[6] * cd.z,
           M[8] * cd.x + M[9] * cd.y + M[10] * cd.z};
  R wl = std::sqrt(wd.x * wd.x + wd.y * wd.y + wd.z * wd.z);
  if (wl > R(1e-6f)) {
    wd.x /= wl;
    wd.y /= wl;
    wd.z /= wl;
  }
  res.cd = cd;
  res.wd = wd;
  if (minD == R(0) || minD == MINF) return res;
  if (maxD == R(0) || maxD == MINF) return res;
  res.marched = true;
  R last_sdf = R(0), last_alpha = R(0);
  unsigned last_w = 0;
  const R d2r = R(1) / cd.z;
  res.d2r = d2r;
  R ray = d2r * std::fmax(minD, minD);
  R ray_end = d2r * std::fmin(maxD, maxD);
  const R th = v * R(3);
  const R inc = v * R(0.5f);
  while (ray < ray_end) {
    V3<R> p{cam.x + ray * wd.x, cam.y + ray * wd.y, cam.z + ray * wd.z};
    R dist;
    uint8_t c1[3];
    if (trilinear(s, p, v, &dist, c1)) {
      if (last_w > 0 && last_sdf < R(0) && dist > R(0)) {
        R alpha = R(0);
        uint8_t c2[3] = {0, 0, 0};
        bool b = refine(s, cam, wd, v, last_sdf, last_alpha, dist, ray, &alpha, c2);
        if (b && std::fabs(last_sdf - dist) < th) {
          if (std::fabs(dist) < th) {
            res.hit = true;
            res.alpha = alpha;
            res.a = last_alpha;
            res.ad = last_sdf;
            res.b = ray;
            res.bd = dist;
            res.rgb[0] = c2[0];
            res.rgb[1] = c2[1];
            res.rgb[2] = c2[2];
            return res;
          }
        }
      }
      last_sdf = dist;
      last_alpha = ray;
      last_w = 1;
    } else {
      last_w = 0;
    }
    // rayCurrent += voxelSize * 0.5f.  When the step is absorbed by rounding the
    // reference's loop never ends; this restatement stops there (a miss).
    R next = ray + inc;
    if (next == ray) break;
    ray = next;
  }
  return res;
}

// Outputs: depth[W*H], pos[4*W*H], col[4*W*H], nrm[4*W*H] (normals stay 1e10).
// Returns 0 if the kernel would not run (no allocated blocks: outputs untouched).
template <class R>
inline int render(const Scene<R>& s, int W, int H, V3<R> cam, const R* M, R fx, R fy, R cx, R cy, R minD, R maxD,
                  R voxel_size, float* depth, float* pos, float* col, float* nrm, uint32_t* hits, uint32_t* misses,
                  const int32_t* pix_list = nullptr, long n_list = 0) {
  *hits = 0;
  *misses = 0;
  if (s.num_allocated() == 0) return 0;  // RayCastRender.cu:553-556
  const R MINF = R(1e10f);
  uint32_t nh = 0, nm = 0;
  // Pixels to trace: all (row-major), or the listed ones (others left untouched).
  const long nwork = pix_list ? n_list : (long)W * H;
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : nh, nm)
  for (long wi = 0; wi < nwork; ++wi) {
    const int pix = pix_list ? pix_list[wi] : (int)wi;
    depth[pix] = (float)MINF;
    for (int k = 0; k < 4; ++k) {
      pos[4 * pix + k] = (float)MINF;
      col[4 * pix + k] = (float)MINF;
      nrm[4 * pix + k] = (float)MINF;
    }
    MarchResult<R> m = march_pixel(s, pix % W, pix / W, cam, M, fx, fy, cx, cy, minD, maxD, voxel_size);
    if (!m.marched) continue;
    if (m.hit) {
      const R alpha = m.alpha;
      V3<R> iso{cam.x + alpha * m.wd.x, cam.y + alpha * m.wd.y, cam.z + alpha * m.wd.z};
      depth[pix] = (float)(alpha / m.d2r);
      pos[4 * pix] = (float)iso.x;
      pos[4 * pix + 1] = (float)iso.y;
      pos[4 * pix + 2] = (float)iso.z;
      pos[4 * pix + 3] = 1.0f;
      col[4 * pix] = (float)((R)m.rgb[0] / R(255));
      col[4 * pix + 1] = (float)((R)m.rgb[1] / R(255));
      col[4 * pix + 2] = (float)((R)m.rgb[2] / R(255));
      col[4 * pix + 3] = 1.0f;
      nh++;
    } else {
      nm++;
    }
  }
  *hits = nh;
  *misses = nm;
  return 1;
}

// ---- Raycast adjoint (no counterpart in the reference) -----------------------------
//
// Rendered depth = alpha * cd.z, alpha the third regula-falsi point.  Differentiated
// along the computation the forward performed (fixed bracket choices, fixed trilinear
// corner sets; the field's jumps at half-voxel planes are not differentiated), with
// respect to
//   * a camera-frame SE(3) perturbation T(xi) = T * exp(xi^), xi = (omega, upsilon):
//     a point at ray parameter r moves as dp = R (upsilon + omega x r*cd) + wd dr;
//   * the SDF of every voxel whose value entered alpha (up to 8 corners of each of the
//     four trilinear evaluations alpha depends on).
// Derivatives are carried in forward mode in double: 6 pose components and 4 seeds, one
// per evaluation (its value perturbed directly); d alpha / d sdf_j = sum_k seed_k *
// (w_j / sum w) of evaluation k.

struct Dual10 {
  double v;
  double d[10];
};

// Trilinear value (bit-identical to trilinear()) plus its gradient in world
// coordinates and the corners' partials d f / d sdf_j = w_j / sum(w).
template <class R> struct TriGrad {
  bool ok = false;
  R val = R(0);
  double g[3] = {0, 0, 0};
  int64_t vox[8];
  double dfds[8];
};

template <class R> inline TriGrad<R> trilinear_grad(const Scene<R>& s, V3<R> p, R v) {
  TriGrad<R> t;
  uint8_t col[3];
  R dist;
  t.ok = trilinear(s, p, v, &dist, col);
  for (int i = 0; i < 8; ++i) {
    t.vox[i] = -1;
    t.dfds[i] = 0.0;
  }
  if (!t.ok) return t;
  t.val = dist;
  const R off = v;
  V3<R> pd{p.x - off / R(2), p.y - off / R(2), p.z - off / R(2)};
  V3<R> q{pd.x / v, pd.y / v, pd.z / v};
  const double fr[3] = {(double)frac01(q.x), (double)frac01(q.y), (double)frac01(q.z)};
  double W = 0, N = 0, dW[3] = {0, 0, 0}, dN[3] = {0, 0, 0}, w[8];
  double sv[8];
  for (int i = 0; i < 8; ++i) {
    w[i] = 0;
    R ax = (i & 1) ? off : R(0), ay = (i & 2) ? off : R(0), az = (i & 4) ? off : R(0);
    V3<R> c{pd.x + ax, pd.y + ay, pd.z + az};
    int32_t vx = ray_voxel_axis(c.x, v), vy = ray_voxel_axis(c.y, v), vz = ray_voxel_axis(c.z, v);
    I3 b{voxel_to_block_axis(vx), voxel_to_block_axis(vy), voxel_to_block_axis(vz)};
    uint32_t h = hash_bucket(b, s.num_buckets);
    int64_t found = -1;
    for (int j = 0; j < kBucket; ++j) {
      size_t e = (size_t)h * kBucket + j;
      if (s.slot_ptr[e] != -1 && s.slot_pos[e] == b)
        found = (int64_t)s.slot_ptr[e] * kVox + voxel_local_axis(vx) + voxel_local_axis(vy) * kBlock +
                voxel_local_axis(vz) * kBlock * kBlock;
    }
    if (found < 0 || s.weight[(size_t)found] == 0) continue;
    double phi[3], dphi[3];
    for (int a = 0; a < 3; ++a) {
      bool hi = (i >> a) & 1;
      phi[a] = hi ? fr[a] : 1.0 - fr[a];
      dphi[a] = (hi ? 1.0 : -1.0) / (double)v;
    }
    w[i] = phi[0] * phi[1] * phi[2];
    sv[i] = (double)s.load_sdf((size_t)found);
    t.vox[i] = found;
    W += w[i];
    N += w[i] * sv[i];
    const double dwi[3] = {dphi[0] * phi[1] * phi[2], phi[0] * dphi[1] * phi[2], phi[0] * phi[1] * dphi[2]};
    for (int a = 0; a < 3; ++a) {
      dW[a] += dwi[a];
      dN[a] += dwi[a] * sv[i];
    }
  }
  const double f = N / W;
  for (int a = 0; a < 3; ++a) t.g[a] = (dN[a] - f * dW[a]) / W;
  for (int i = 0; i < 8; ++i)
    if (t.vox[i] >= 0) t.dfds[i] = w[i] / W;
  return t;
}

// Dual of one evaluation f(p(r)) at ray parameter r (dual: its derivative), seed k.
template <class R>
inline Dual10 eval_dual(const TriGrad<R>& tg, const Dual10& r, const double* cdd, const double* wdd,
                        const double* Rm, int seed) {
  // camera-frame point and the world gradient rotated into the camera frame
  const double q[3] = {r.v * cdd[0], r.v * cdd[1], r.v * cdd[2]};
  const double gc[3] = {Rm[0] * tg.g[0] + Rm[3] * tg.g[1] + Rm[6] * tg.g[2],
                        Rm[1] * tg.g[0] + Rm[4] * tg.g[1] + Rm[7] * tg.g[2],
                        Rm[2] * tg.g[0] + Rm[5] * tg.g[1] + Rm[8] * tg.g[2]};
  const double gw = tg.g[0] * wdd[0] + tg.g[1] * wdd[1] + tg.g[2] * wdd[2];  // d f / d r
  Dual10 e;
  e.v = (double)tg.val;
  // omega: q x gc ; upsilon: gc
  const double dom[3] = {q[1] * gc[2] - q[2] * gc[1], q[2] * gc[0] - q[0] * gc[2], q[0] * gc[1] - q[1] * gc[0]};
  for (int k = 0; k < 10; ++k) e.d[k] = gw * r.d[k];
  for (int k = 0; k < 3; ++k) {
    e.d[k] += dom[k];
    e.d[3 + k] += gc[k];
  }
  e.d[6 + seed] += 1.0;
  return e;
}

// d depth / d xi (6) and the SDF partials of one pixel.  Returns false without a hit.
template <class R>
inline bool adjoint_pixel(const Scene<R>& s, int x, int y, V3<R> cam, const R* M, R fx, R fy, R cx, R cy, R minD,
                          R maxD, R v, double J[6], int64_t vox[32], double dds[32], double* depth) {
  MarchResult<R> m = march_pixel(s, x, y, cam, M, fx, fy, cx, cy, minD, maxD, v);
  *depth = m.hit ? (double)(m.alpha / m.d2r) : 1e10;
  for (int k = 0; k < 6; ++k) J[k] = 0;
  for (int k = 0; k < 32; ++k) {
    vox[k] = -1;
    dds[k] = 0;
  }
  if (!m.hit) return false;
  const double Rm[9] = {(double)M[0], (double)M[1], (double)M[2], (double)M[4], (double)M[5],
                        (double)M[6], (double)M[8], (double)M[9], (double)M[10]};
  const double cdd[3] = {(double)m.cd.x, (double)m.cd.y, (double)m.cd.z};
  const double wdd[3] = {(double)m.wd.x, (double)m.wd.y, (double)m.wd.z};
  TriGrad<R> tg[4];
  auto pos = [&](R r) { return V3<R>{cam.x + r * m.wd.x, cam.y + r * m.wd.y, cam.z + r * m.wd.z}; };
  // bracket: constants a, b; their evaluations are seeds 0 and 1
  Dual10 a{}, b{};
  a.v = (double)m.a;
  b.v = (double)m.b;
  tg[0] = trilinear_grad(s, pos(m.a), v);
  tg[1] = trilinear_grad(s, pos(m.b), v);
  Dual10 ad = eval_dual(tg[0], a, cdd, wdd, Rm, 0);
  Dual10 bd = eval_dual(tg[1], b, cdd, wdd, Rm, 1);
  // float state that reproduces the forward's sample positions and branch choices
  R fa = m.a, fad = m.ad, fb = m.b, fbd = m.bd, fc = R(0);
  Dual10 c{};
  for (int it = 0; it < 3; ++it) {
    fc = fa + (fad / (fad - fbd)) * (fb - fa);
    const double den = ad.v - bd.v, u = ad.v / den;
    c.v = a.v + u * (b.v - a.v);
    for (int k = 0; k < 10; ++k) {
      const double du = (-ad.d[k] * bd.v + ad.v * bd.d[k]) / (den * den);
      c.d[k] = a.d[k] + du * (b.v - a.v) + u * (b.d[k] - a.d[k]);
    }
    if (it == 2) break;
    tg[2 + it] = trilinear_grad(s, pos(fc), v);
    Dual10 cdual = c;
    cdual.v = (double)fc;
    Dual10 cd = eval_dual(tg[2 + it], cdual, cdd, wdd, Rm, 2 + it);
    const R fcd = tg[2 + it].val;
    if (fad * fcd > R(0)) {
      fa = fc;
      fad = fcd;
      a = cdual;
      ad = cd;
    } else {
      fb = fc;
      fbd = fcd;
      b = cdual;
      bd = cd;
    }
  }
  for (int k = 0; k < 6; ++k) J[k] = c.d[k] * cdd[2];
  for (int e = 0; e < 4; ++e)
    for (int i = 0; i < 8; ++i) {
      vox[8 * e + i] = tg[e].vox[i];
      dds[8 * e + i] = c.d[6 + e] * tg[e].dfds[i] * cdd[2];
    }
  return true;
}

// ---- surface-point extraction (RayCastRender.cu:629-868 kernelExtractSurfacePoints and
// computeNormalFromSDF, RayCastRender.cpp:307-364 extractSurfacePoints)
template <class R> struct SurfacePoint {
  R p[3];
  R n[3];
  uint8_t c[3];
};

// computeNormalFromSDF (RayCastRender.cu:629-676): central differences of getVoxel at
// +-v/2 per axis (an axis whose two voxels are not both weighted contributes 0),
// normalised when the length exceeds 1e-6.
template <class R> inline V3<R> sdf_normal(const Scene<R>& s, V3<R> p, R v) {
  const R eps = v * R(0.5f);
  R g[3] = {R(0), R(0), R(0)};
  for (int a = 0; a < 3; ++a) {
    V3<R> pp = p, pm = p;
    if (a == 0) { pp.x = p.x + eps; pm.x = p.x - eps; }
    if (a == 1) { pp.y = p.y + eps; pm.y = p.y - eps; }
    if (a == 2) { pp.z = p.z + eps; pm.z = p.z - eps; }
    const VoxelRead<R> vp = get_voxel(s, pp, v), vm = get_voxel(s, pm, v);
    if (vp.w > 0 && vm.w > 0) g[a] = (vp.sdf - vm.sdf) / (R(2.0f) * eps);
  }
  const R len = std::sqrt(g[0] * g[0] + g[1] * g[1] + g[2] * g[2]);
  if (len > R(1e-6f)) {
    g[0] = g[0] / len;
    g[1] = g[1] / len;
    g[2] = g[2] / len;
  }
  return V3<R>{g[0], g[1], g[2]};
}

// kernelExtractSurfacePoints (RayCastRender.cu:684-868), entries in hash order.  A voxel
// with weight >= max(1, min_weight) and SDF in [min_sdf, max_sdf] emits a point: the
// mean of the zero crossings towards its 6 neighbours (a neighbour inside the block by
// index, one outside through getVoxel at the voxel centre +- v; weightless neighbours
// and equal signs contribute nothing; crossing at t = sdf / (sdf - nsdf) along the
// axis), or its centre when there is none and |sdf| < 0.001.  Returns the number of
// points found; keeps the first max_points.
template <class R>
inline uint64_t extract_surface_points(const Scene<R>& s, R min_sdf, R max_sdf, uint32_t min_weight,
                                       uint64_t max_points, std::vector<SurfacePoint<R>>& out) {
  const R v = s.voxel;
  static const int off[6][3] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};
  uint64_t total = 0;
  out.clear();
  for (size_t e = 0; e < s.slot_ptr.size(); ++e) {
    const int32_t ptr = s.slot_ptr[e];
    if (ptr == -1) continue;
    const I3 b = s.slot_pos[e];
    const V3<R> bw{R(b.x * kBlock) * v, R(b.y * kBlock) * v, R(b.z * kBlock) * v};
    for (int z = 0; z < kBlock; ++z)
      for (int y = 0; y < kBlock; ++y)
        for (int x = 0; x < kBlock; ++x) {
          const size_t vi = (size_t)ptr * kVox + (size_t)(x + y * kBlock + z * kBlock * kBlock);
          const uint32_t w = s.weight[vi];
          if (w == 0 || w < min_weight) continue;
          const R sd = s.load_sdf(vi);
          if (sd < min_sdf || sd > max_sdf) continue;
          const V3<R> c{bw.x + (R(x) + R(0.5f)) * v, bw.y + (R(y) + R(0.5f)) * v, bw.z + (R(z) + R(0.5f)) * v};
          V3<R> acc{R(0), R(0), R(0)};
          int n = 0;
          for (int d = 0; d < 6; ++d) {
            const int nx = x + off[d][0], ny = y + off[d][1], nz = z + off[d][2];
            const R dx = R(off[d][0]), dy = R(off[d][1]), dz = R(off[d][2]);
            R nsd;
            if (nx < 0 || nx >= kBlock || ny < 0 || ny >= kBlock || nz < 0 || nz >= kBlock) {
              const VoxelRead<R> nv = get_voxel(s, V3<R>{c.x + dx * v, c.y + dy * v, c.z + dz * v}, v);
              if (nv.w == 0) continue;
              nsd = nv.sdf;
            } else {
              const size_t ni = (size_t)ptr * kVox + (size_t)(nx + ny * kBlock + nz * kBlock * kBlock);
              if (s.weight[ni] == 0) continue;
              nsd = s.load_sdf(ni);
            }
            if ((sd < R(0) && nsd > R(0)) || (sd > R(0) && nsd < R(0))) {
              const R t = sd / (sd - nsd);
              acc.x = acc.x + (c.x + dx * v * t);
              acc.y = acc.y + (c.y + dy * v * t);
              acc.z = acc.z + (c.z + dz * v * t);
              n++;
            }
          }
          V3<R> iso;
          if (n == 0) {
            if (!(std::fabs(sd) < R(0.001f))) continue;
            iso = c;
          } else {
            iso = V3<R>{acc.x / R(n), acc.y / R(n), acc.z / R(n)};
          }
          const V3<R> nrm = sdf_normal(s, iso, v);
          if (total < max_points) {
            SurfacePoint<R> sp;
            sp.p[0] = iso.x; sp.p[1] = iso.y; sp.p[2] = iso.z;
            sp.n[0] = nrm.x; sp.n[1] = nrm.y; sp.n[2] = nrm.z;
            sp.c[0] = s.color[3 * vi]; sp.c[1] = s.color[3 * vi + 1]; sp.c[2] = s.color[3 * vi + 2];
            out.push_back(sp);
          }
          total++;
        }
  }
  return total;
}

// ---- marching cubes (MarchingCubes.cu:313-467: vertexInterp, computeNormal and
// marchingCubesKernel).  Blocks in hash-entry order; in each block the cubes in voxel
// order (x fastest); cube (x, y, z) spans voxels x..x+1, y..y+1, z..z+1.  As in the
// reference, cubes stay inside the block (x, y, z <= 6); with `cross`, cubes at 7 take
// their far corners from the +x / +y / +z neighbours and are skipped when one of those
// is not allocated.  A cube is skipped when a corner's weight is below min_weight (0:
// no test, as in the reference).  Corner c (bit 0: +x, bit 1: +y, bit 2: +z) sits at the
// voxel centre (8 block + i + 0.5) v, block and local index i taken in the block that
// holds the voxel.  Case bit c is set when sdf < iso; `table` (256 rows of 32) lists each
// case's triangles as edge triples up to -1.  The vertex on the edge from corner a to
// corner b (a < b, the edge numbering of the table) is vertexInterp(iso, pa, pb, sa, sb):
// the midpoint 0.5 (pa + pb) when |sb - sa| < 1e-6, else pa + mu (pb - pa) with
// mu = (iso - sa) / (sb - sa) clamped to [0, 1].  A triangle's normal (all three
// vertices) is (v1 - v0) x (v2 - v0) scaled by 1 / length, zero when the length is at
// most 1e-6.  Returns the number of triangles found and keeps the first max_tris:
// 9 vertex and 9 normal coordinates each.
template <class R>
inline uint64_t marching_cubes(const Scene<R>& s, R iso, R v, int cross, uint32_t min_weight,
                               const int8_t (*table)[32], uint64_t max_tris, std::vector<R>& verts,
                               std::vector<R>& nrms) {
  static const int edge[12][2] = {{0, 1}, {2, 3}, {4, 5}, {6, 7}, {0, 2}, {1, 3},
                                  {4, 6}, {5, 7}, {0, 4}, {1, 5}, {2, 6}, {3, 7}};
  auto find = [&](I3 b) -> int32_t {
    const size_t hp = (size_t)hash_bucket(b, s.num_buckets) * kBucket;
    for (int j = 0; j < kBucket; ++j)
      if (s.slot_ptr[hp + j] != -1 && s.slot_pos[hp + j] == b) return s.slot_ptr[hp + j];
    return -1;
  };
  auto centre = [&](int32_t bc, int i) { return R(bc * kBlock) * v + (R(i) + R(0.5f)) * v; };
  verts.clear();
  nrms.clear();
  uint64_t total = 0;
  const int last = cross ? kBlock : kBlock - 1;
  for (size_t e = 0; e < s.slot_ptr.size(); ++e) {
    if (s.slot_ptr[e] == -1) continue;
    const I3 b = s.slot_pos[e];
    int32_t nb[8];  // blocks at offsets (c & 1, c >> 1 & 1, c >> 2 & 1)
    for (int c = 0; c < 8; ++c)
      nb[c] = c == 0 ? s.slot_ptr[e] : (cross ? find(I3{b.x + (c & 1), b.y + ((c >> 1) & 1), b.z + (c >> 2)}) : -1);
    for (int z = 0; z < last; ++z)
      for (int y = 0; y < last; ++y)
        for (int x = 0; x < last; ++x) {
          R sd[8], P[8][3];
          bool ok = true;
          int k = 0;
          for (int c = 0; c < 8 && ok; ++c) {
            const int g[3] = {x + (c & 1), y + ((c >> 1) & 1), z + (c >> 2)};
            const int o = (g[0] >> 3) | ((g[1] >> 3) << 1) | ((g[2] >> 3) << 2);
            if (nb[o] < 0) {
              ok = false;
              break;
            }
            const size_t vi = (size_t)nb[o] * kVox + (size_t)((g[0] & 7) + (g[1] & 7) * kBlock + (g[2] & 7) * kBlock * kBlock);
            if (s.weight[vi] < min_weight) ok = false;
            sd[c] = s.load_sdf(vi);
            const int32_t bc[3] = {b.x, b.y, b.z};
            for (int d = 0; d < 3; ++d) P[c][d] = centre(bc[d] + (g[d] >> 3), g[d] & 7);
            if (sd[c] < iso) k |= 1 << c;
          }
          if (!ok) continue;
          R V[12][3];
          for (int ed = 0; ed < 12; ++ed) {
            const int a = edge[ed][0], bb = edge[ed][1];
            const R den = sd[bb] - sd[a];
            if (std::fabs(den) < R(1e-6f)) {
              for (int d = 0; d < 3; ++d) V[ed][d] = R(0.5f) * (P[a][d] + P[bb][d]);
            } else {
              R mu = (iso - sd[a]) / den;
              mu = std::min(std::max(mu, R(0)), R(1));
              for (int d = 0; d < 3; ++d) V[ed][d] = P[a][d] + mu * (P[bb][d] - P[a][d]);
            }
          }
          for (int t = 0; t < 10 && table[k][3 * t] != -1; ++t, ++total) {
            if (total >= max_tris) continue;
            const R* v0 = V[table[k][3 * t]];
            const R* v1 = V[table[k][3 * t + 1]];
            const R* v2 = V[table[k][3 * t + 2]];
            const R a[3] = {v1[0] - v0[0], v1[1] - v0[1], v1[2] - v0[2]};
            const R c[3] = {v2[0] - v0[0], v2[1] - v0[1], v2[2] - v0[2]};
            R n[3] = {a[1] * c[2] - a[2] * c[1], a[2] * c[0] - a[0] * c[2], a[0] * c[1] - a[1] * c[0]};
            const R len = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
            if (len > R(1e-6f)) {
              const R inv = R(1) / len;
              for (int d = 0; d < 3; ++d) n[d] *= inv;
            } else {
              n[0] = n[1] = n[2] = R(0);
            }
            for (const R* q : {v0, v1, v2})
              for (int d = 0; d < 3; ++d) {
                verts.push_back(q[d]);
                nrms.push_back(n[d]);
              }
          }
        }
  }
  return total;
}

// ---- depth map from a camera-frame point cloud: generateFromPointCloud
// (CustomDepthMapGenerator.cpp:81-222) with projectPoint (:550-578) and isPointInBounds
// (:580-587), the reference's own sequential loop (its CUDA path is a TODO that falls
// back to this loop, :603-611).  Points in array order: z <= 0 skipped; perspective
// u = x fx / z + cx, v = y fy / z + cy (orthographic u = x + cx, v = y + cy); in bounds
// when 0 <= u < W and 0 <= v < H; pixel (floor u, floor v); every later point
// overwrites the pixel's point and depth, and its normal / colour when it has one
// (i < n_normals / n_colors).  With no normals at all, finite differences of the depth
// map (:164-208).  Outputs W*H: pmap float3, depth, nmap float3, cmap uchar3 (zeros
// where no point lands).  Returns the pixels with depth > 0.
inline uint32_t depthmap_from_points(int W, int H, float fx, float fy, float cx, float cy, bool ortho,
                                     const float* pts, size_t n, const float* nrm, size_t n_nrm, const uint8_t* col,
                                     size_t n_col, float* pmap, float* depth, float* nmap, uint8_t* cmap) {
  const size_t npix = (size_t)W * H;
  for (size_t i = 0; i < 3 * npix; ++i) {
    pmap[i] = 0.0f;
    nmap[i] = 0.0f;
    cmap[i] = 0;
  }
  for (size_t i = 0; i < npix; ++i) depth[i] = 0.0f;
  for (size_t i = 0; i < n; ++i) {
    const float x = pts[3 * i], y = pts[3 * i + 1], z = pts[3 * i + 2];
    if (z <= 0.0f) continue;
    float u, v;
    if (ortho) {
      u = x + cx;
      v = y + cy;
    } else {
      u = (x * fx / z) + cx;
      v = (y * fy / z) + cy;
    }
    if (!(u >= 0.0f && u < (float)W && v >= 0.0f && v < (float)H)) continue;
    const int px = cvt_i32(std::floor(u)), py = cvt_i32(std::floor(v));
    if (px < 0 || py < 0 || px >= W || py >= H) continue;
    const size_t p = (size_t)py * W + px;
    pmap[3 * p] = x;
    pmap[3 * p + 1] = y;
    pmap[3 * p + 2] = z;
    depth[p] = z;
    if (i < n_nrm) {
      nmap[3 * p] = nrm[3 * i];
      nmap[3 * p + 1] = nrm[3 * i + 1];
      nmap[3 * p + 2] = nrm[3 * i + 2];
    }
    if (i < n_col) {
      cmap[3 * p] = col[3 * i];
      cmap[3 * p + 1] = col[3 * i + 1];
      cmap[3 * p + 2] = col[3 * i + 2];
    }
  }
  if (n_nrm == 0) {
    for (int y = 1; y < H - 1; ++y)
      for (int x = 1; x < W - 1; ++x) {
        const size_t p = (size_t)y * W + x;
        const float d = depth[p];
        if (!(d > 0.0f)) continue;
        const float c0 = ((float)x - cx) * d / fx, c1 = ((float)y - cy) * d / fy, c2 = d;
        const float dr = depth[p + 1];
        const float r0 = ((float)(x + 1) - cx) * dr / fx, r1 = ((float)y - cy) * dr / fy, r2 = dr;
        const float dd = depth[p + W];
        const float d0 = ((float)x - cx) * dd / fx, d1 = ((float)(y + 1) - cy) * dd / fy, d2 = dd;
        const float a0 = r0 - c0, a1 = r1 - c1, a2 = r2 - c2;  // right - centre
        const float b0 = d0 - c0, b1 = d1 - c1, b2 = d2 - c2;  // down - centre
        const float n0 = a1 * b2 - a2 * b1, n1 = a2 * b0 - a0 * b2, n2 = a0 * b1 - a1 * b0;
        const float len = std::sqrt(n0 * n0 + n1 * n1 + n2 * n2);
        if (len > 0.0f) {
          const float s = 1.0f / len;
          nmap[3 * p] = n0 * s;
          nmap[3 * p + 1] = n1 * s;
          nmap[3 * p + 2] = n2 * s;
        }
      }
  }
  uint32_t valid = 0;
  for (size_t p = 0; p < npix; ++p) valid += depth[p] > 0.0f ? 1u : 0u;
  return valid;
}

}  // namespace vsgo
