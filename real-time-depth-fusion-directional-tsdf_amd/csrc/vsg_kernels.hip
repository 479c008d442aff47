// VoxelSG depth-fusion kernels for gfx950 (wave64).
//
// Per frame, on one stream, with no host synchronisation:
//   k_alloc_claim    one thread per pixel: truncation band -> block DDA -> for every
//                    block absent from its bucket, atomicMin of the request key
//                    (pixel << 10 | DDA step) on the bucket's claim word
//   k_alloc_resolve  pixels that claimed re-walk their band and count the steps
//                    whose key won; workgroup sums
//   k_alloc_insert   every workgroup sums its predecessors' winners itself; winners
//                    insert their block into the first free slot of the bucket with
//                    pointer = blocks before + rank in key order; workgroup 0 sets
//                    the frame's counters and clears the next frame's
//   k_compact        makes the new block count live; one thread per live block:
//                    conservative "can this frame update it" test, wave ballot +
//                    prefix-sum stream compaction
//   k_integrate      one 512-thread workgroup per listed block, one voxel per
//                    thread, grid-stride over the list
// and for rendering:
//   k_render         one thread per pixel (8x8 tiles): ray march in half-voxel
//                    steps, trilinear samples, three regula-falsi steps
//
// The reference lets GPU scheduling pick the one insertion per bucket per frame
// (bucket lock taken and never released on success, VoxelScene.cu:176-217).
// Here the request with the smallest (pixel, step) key wins: the outcome of a
// sequential loop over the same requests, so the block set and the block
// pointers are deterministic.
#include "vsg_device.hpp"
#include "vsg_internal.hpp"
#include "vsg_div.hpp"
#include "vsg_mc_table.hpp"
#include "vsg_replay.hpp"

namespace vsg {

// ---------------------------------------------------------------------------------
// Band and DDA (VoxelScene.cu:295-481, 527-605)
// ---------------------------------------------------------------------------------

// Band ends of one pixel; false when the reference's kernel returns early.
__device__ __forceinline__ bool pixel_band(const FrameArgs& fa, float Px, float Py, float Pz, float3* lo,
                                           float3* hi) {
  if (Pz <= 1e-6f) return false;
  const float3 w = xform_point(fa.M, Px, Py, Pz);
  const float rx = w.x - fa.M[3], ry = w.y - fa.M[7], rz = w.z - fa.M[11];
  const float len = sqrtf(rx * rx + ry * ry + rz * rz);
  if (len < 1e-6f) return false;
  *lo = xform_point(fa.M, Px, Py, Pz - fa.trunc);
  *hi = xform_point(fa.M, Px, Py, Pz + fa.trunc);
  return true;
}

template <class F> __device__ __forceinline__ void dda_band(float3 a, float3 b, float v, F&& visit) {
  int32_t cx = voxel_to_block_axis(alloc_voxel_axis(a.x, v));
  int32_t cy = voxel_to_block_axis(alloc_voxel_axis(a.y, v));
  int32_t cz = voxel_to_block_axis(alloc_voxel_axis(a.z, v));
  const int32_t ex = voxel_to_block_axis(alloc_voxel_axis(b.x, v));
  const int32_t ey = voxel_to_block_axis(alloc_voxel_axis(b.y, v));
  const int32_t ez = voxel_to_block_axis(alloc_voxel_axis(b.z, v));
  if (cx == ex && cy == ey && cz == ez) return;
  float dx = b.x - a.x, dy = b.y - a.y, dz = b.z - a.z;
  const float len = sqrtf(dx * dx + dy * dy + dz * dz);
  if (len < 1e-6f) return;
  dx /= len;
  dy /= len;
  dz /= len;
  const float stx = dx > 0.0f ? 1.0f : -1.0f, sty = dy > 0.0f ? 1.0f : -1.0f, stz = dz > 0.0f ? 1.0f : -1.0f;
  const float ext = (float)kBlock * v;
  const float scx = stx > 0.0f ? 1.0f : 0.0f, scy = sty > 0.0f ? 1.0f : 0.0f, scz = stz > 0.0f ? 1.0f : 0.0f;
  const float hv = 0.5f * v;
  const float bx = ((float)cx + scx) * ext - hv, by = ((float)cy + scy) * ext - hv, bz = ((float)cz + scz) * ext - hv;
  float tmx = (bx - a.x) / dx, tmy = (by - a.y) / dy, tmz = (bz - a.z) / dz;
  float tdx = (stx * ext) / dx, tdy = (sty * ext) / dy, tdz = (stz * ext) / dz;
  if (dx == 0.0f) { tmx = 1e6f; tdx = 1e6f; }
  if (dy == 0.0f) { tmy = 1e6f; tdy = 1e6f; }
  if (dz == 0.0f) { tmz = 1e6f; tdz = 1e6f; }
  const int32_t sx = (int32_t)stx, sy = (int32_t)sty, sz = (int32_t)stz;
  for (int iter = 0; iter < 1024; ++iter) {
    visit(cx, cy, cz, iter);
    if (cx == ex && cy == ey && cz == ez) break;
    if (tmx < tmy && tmx < tmz) {
      tmx += tdx;
      cx += sx;
    } else if (tmy < tmz) {
      tmy += tdy;
      cy += sy;
    } else {
      tmz += tdz;
      cz += sz;
    }
  }
}

// ---------------------------------------------------------------------------------
// Allocation
// ---------------------------------------------------------------------------------

// Per-workgroup partial counters of k_alloc_claim (reduced by k_alloc_scan).
constexpr int kClaimPart = 8;  // valid, claimed, out of range, non-finite, zmin key, zmax key, -, -

__global__ __launch_bounds__(256) void k_alloc_claim(TableView t, FrameArgs fa, const float* __restrict__ pts,
                                                     uint32_t* __restrict__ pixflag, uint8_t* __restrict__ mark_bkt,
                                                     uint32_t* __restrict__ part) {
  const int pix = blockIdx.x * 256 + threadIdx.x;
  const int npix = fa.W * fa.H;
  uint32_t valid = 0, claimed = 0, oor = 0, nonfinite = 0;
  uint32_t zmin = 0xFFFFFFFFu, zmax = 0u;
  if (pix < npix) {
    const float Px = pts[3 * pix], Py = pts[3 * pix + 1], Pz = pts[3 * pix + 2];
    if (Px != 0.0f || Py != 0.0f || Pz != 0.0f) {  // integrate-side validity (VoxelScene.cu:1037-1043)
      if (!(isfinite(Px) && isfinite(Py) && isfinite(Pz))) {
        nonfinite = 1;
      } else {
        zmin = zmax = float_to_ordered(Pz);
      }
    }
    if (!(Pz <= 1e-6f)) valid = 1;
    float3 lo, hi;
    if (valid && pixel_band(fa, Px, Py, Pz, &lo, &hi)) {
      dda_band(lo, hi, fa.voxel, [&](int32_t x, int32_t y, int32_t z, int iter) {
        if (!key_in_range(x, y, z)) {
          oor++;
          return;
        }
        const uint32_t h = hash_bucket(x, y, z, t.num_buckets);
        const int c = t.cnt[h];
        if (mark_bkt && c > 0) mark_bkt[h] = 1;
        if (c >= kBucket) return;  // full: present (no-op) or refused (no-op)
        const uint64_t key = pack_key(x, y, z);
        const uint64_t* r = t.rec + (size_t)h * kRecWords;
        bool present = false;
        for (int j = 0; j < c; ++j) present |= (r[j] == key);
        if (present) return;
        atomicMin(&t.claim[h], ((uint32_t)pix << 10) | (uint32_t)iter);
        claimed = 1;
      });
    }
    pixflag[pix] = claimed;
  }
  // one write per workgroup: same-address atomics from every wave would serialise
  __shared__ uint32_t red[4][6];
  const int w = threadIdx.x >> 6;
  valid = wave_sum(valid);
  claimed = wave_sum(claimed);
  oor = wave_sum(oor);
  nonfinite = wave_sum(nonfinite);
  zmin = wave_min(zmin);
  zmax = wave_max(zmax);
  if ((threadIdx.x & 63) == 0) {
    red[w][0] = valid;
    red[w][1] = claimed;
    red[w][2] = oor;
    red[w][3] = nonfinite;
    red[w][4] = zmin;
    red[w][5] = zmax;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t* o = part + (size_t)blockIdx.x * kClaimPart;
    o[0] = red[0][0] + red[1][0] + red[2][0] + red[3][0];
    o[1] = red[0][1] + red[1][1] + red[2][1] + red[3][1];
    o[2] = red[0][2] + red[1][2] + red[2][2] + red[3][2];
    o[3] = red[0][3] + red[1][3] + red[2][3] + red[3][3];
    o[4] = min(min(red[0][4], red[1][4]), min(red[2][4], red[3][4]));
    o[5] = max(max(red[0][5], red[1][5]), max(red[2][5], red[3][5]));
  }
}

__global__ __launch_bounds__(256) void k_alloc_resolve(TableView t, FrameArgs fa, const float* __restrict__ pts,
                                                       uint32_t* __restrict__ pixcnt, uint32_t* __restrict__ wgsum) {
  const int pix = blockIdx.x * 256 + threadIdx.x;
  const int npix = fa.W * fa.H;
  uint32_t n = 0;
  if (pix < npix && pixcnt[pix] != 0u) {
    const float Px = pts[3 * pix], Py = pts[3 * pix + 1], Pz = pts[3 * pix + 2];
    float3 lo, hi;
    if (pixel_band(fa, Px, Py, Pz, &lo, &hi)) {
      dda_band(lo, hi, fa.voxel, [&](int32_t x, int32_t y, int32_t z, int iter) {
        if (!key_in_range(x, y, z)) return;
        const uint32_t h = hash_bucket(x, y, z, t.num_buckets);
        if (t.claim[h] == (((uint32_t)pix << 10) | (uint32_t)iter)) n++;
      });
    }
    pixcnt[pix] = n;
  }
  __shared__ uint32_t ws[4];
  const uint32_t s = wave_sum(n);
  if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) wgsum[blockIdx.x] = ws[0] + ws[1] + ws[2] + ws[3];
}

// The frame counters from the claim kernel's per-workgroup partials and the winner
// total, by workgroup 0 of k_alloc_insert (all of its threads): valid and claiming
// pixels, out-of-range and non-finite points, the depth range, winners, heap
// failures; the block count after the insertions (n_alloc_next, published by the
// compaction) and the next frame's counter set cleared.
__device__ __forceinline__ void frame_counters(DevState* st, int32_t cset, const uint32_t* __restrict__ part,
                                               uint32_t nwg, uint32_t prev, uint32_t winners, uint32_t ok) {
  uint32_t cv = 0, cc = 0, co = 0, cn = 0, zmn = 0xFFFFFFFFu, zmx = 0u;
  for (uint32_t k = threadIdx.x; k < nwg; k += 256u) {
    const uint32_t* p = part + (size_t)k * kClaimPart;
    cv += p[0];
    cc += p[1];
    co += p[2];
    cn += p[3];
    zmn = min(zmn, p[4]);
    zmx = max(zmx, p[5]);
  }
  __shared__ uint32_t red[4][6];
  cv = wave_sum(cv);
  cc = wave_sum(cc);
  co = wave_sum(co);
  cn = wave_sum(cn);
  zmn = wave_min(zmn);
  zmx = wave_max(zmx);
  if ((threadIdx.x & 63) == 0) {
    const int w = threadIdx.x >> 6;
    red[w][0] = cv;
    red[w][1] = cc;
    red[w][2] = co;
    red[w][3] = cn;
    red[w][4] = zmn;
    red[w][5] = zmx;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    FrameCounters& fc = st->set[cset];
    uint32_t a[6] = {0, 0, 0, 0, 0xFFFFFFFFu, 0u};
    for (int w = 0; w < 4; ++w) {
      for (int k = 0; k < 4; ++k) a[k] += red[w][k];
      a[4] = min(a[4], red[w][4]);
      a[5] = max(a[5], red[w][5]);
    }
    fc.valid_pixels = a[0];
    fc.claim_pixels = a[1];
    fc.out_of_range = a[2];
    fc.nonfinite = a[3];
    fc.zmin_key = a[4];
    fc.zmax_key = a[5];
    st->n_alloc_prev = prev;
    st->n_alloc_next = prev + ok;
    st->frames += 1u;
    fc.winners = winners;
    fc.claimed_buckets = winners;
    fc.heap_fail = winners - ok;
    FrameCounters& nx = st->set[cset ^ 1];
    nx.valid_pixels = 0;
    nx.claim_pixels = 0;
    nx.winners = 0;
    nx.heap_fail = 0;
    nx.out_of_range = 0;
    nx.visible = 0;
    nx.updated = 0;
    nx.nonfinite = 0;
    nx.zmin_key = 0xFFFFFFFFu;
    nx.zmax_key = 0u;
    nx.claimed_buckets = 0;
  }
}

// Each workgroup sums the resolve's per-workgroup winner counts itself (its
// predecessors' for its pointer offset, all for the frame's total): no separate scan
// pass.  The live block count is read here and changes only when the compaction
// publishes the new one.
__global__ __launch_bounds__(256) void k_alloc_insert(TableView t, DevState* st, FrameArgs fa,
                                                      const float* __restrict__ pts, const uint32_t* __restrict__ pixcnt,
                                                      const uint32_t* __restrict__ wgsum,
                                                      const uint32_t* __restrict__ part, uint32_t nwg) {
  uint32_t before = 0, all = 0;
  for (uint32_t k = threadIdx.x; k < nwg; k += 256u) {
    const uint32_t w = wgsum[k];
    all += w;
    if (k < blockIdx.x) before += w;
  }
  uint32_t off, total;
  (void)block_excl_scan<256>(before, &off);
  (void)block_excl_scan<256>(all, &total);
  const uint32_t prev = st->n_alloc;
  const uint32_t room = t.num_blocks - prev;
  if (blockIdx.x == 0u) frame_counters(st, fa.cset, part, nwg, prev, total, total < room ? total : room);
  const int pix = blockIdx.x * 256 + threadIdx.x;
  const int npix = fa.W * fa.H;
  const uint32_t n = pix < npix ? pixcnt[pix] : 0u;
  uint32_t tot;
  uint32_t rank = block_excl_scan<256>(n, &tot);
  if (n == 0u) return;
  rank += off;
  const float Px = pts[3 * pix], Py = pts[3 * pix + 1], Pz = pts[3 * pix + 2];
  float3 lo, hi;
  if (!pixel_band(fa, Px, Py, Pz, &lo, &hi)) return;
  dda_band(lo, hi, fa.voxel, [&](int32_t x, int32_t y, int32_t z, int iter) {
    if (!key_in_range(x, y, z)) return;
    const uint32_t h = hash_bucket(x, y, z, t.num_buckets);
    if (t.claim[h] != (((uint32_t)pix << 10) | (uint32_t)iter)) return;
    t.claim[h] = kNoClaim;  // this thread is the bucket's only winner
    const uint32_t p = prev + rank;
    rank++;
    if (p >= t.num_blocks) return;  // heap exhausted (VoxelScene.cu:204-214)
    const int slot = t.cnt[h];
    uint64_t* r = t.rec + (size_t)h * kRecWords;
    r[slot] = pack_key(x, y, z);
    reinterpret_cast<int32_t*>(r + kBucket)[slot] = (int32_t)p;
    t.cnt[h] = (uint8_t)(slot + 1);
    t.blk[p] = make_int4(x, y, z, (int32_t)(h * kBucket + slot));
    atomicOr(reinterpret_cast<unsigned long long*>(&t.occ[occ_index(x >> 2, y >> 2, z >> 2, t.occ_mask)]),
             1ull << occ_bit(x, y, z));
    atomicOr(reinterpret_cast<unsigned long long*>(&t.hyp[hyp_index(x >> 4, y >> 4, z >> 4, t.hyp_mask)]),
             1ull << hyp_bit(x >> 2, y >> 2, z >> 2));
    // block index: claim the first empty entry from the home slot (keys are unique,
    // so no entry can already hold this key)
    const uint64_t key = pack_key(x, y, z);
    uint32_t sl = idx_home(x, y, z, t.idx_mask);
    for (uint32_t n = 0; n <= t.idx_mask; ++n) {
      unsigned long long* kw = reinterpret_cast<unsigned long long*>(&t.idx[sl]);
      if (atomicCAS(kw, 0xFFFFFFFFFFFFFFFFull, (unsigned long long)key) == 0xFFFFFFFFFFFFFFFFull) {
        t.idx[sl].z = (uint32_t)p;
        t.idxpos[p] = sl;
        break;
      }
      sl = (sl + 1u) & t.idx_mask;
    }
  });
}

// ---------------------------------------------------------------------------------
// Visible-block compaction
// ---------------------------------------------------------------------------------

__device__ __forceinline__ float ordered_to_float(uint32_t k) {
  return __uint_as_float((k & 0x80000000u) ? (k & 0x7fffffffu) : ~k);
}

// False only when no voxel of block (x,y,z) can be updated by this frame: every
// updated voxel lies within trunc of a valid point along the camera z axis, and
// (for blocks fully in front of the camera) projects inside the image.  Margins
// cover float rounding of the per-voxel arithmetic many times over.
__device__ __forceinline__ bool block_maybe_visible(const FrameArgs& fa, int4 b, double zmin, double zmax) {
  const double* M = fa.Md;
  const double v = (double)fa.voxel;
  double lo_z = 1e300, hi_z = -1e300, lo_u = 1e300, hi_u = -1e300, lo_v = 1e300, hi_v = -1e300;
  bool all_front = true;
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const double wx = ((double)b.x * kBlock + ((k & 1) ? 7 : 0)) * v;
    const double wy = ((double)b.y * kBlock + ((k & 2) ? 7 : 0)) * v;
    const double wz = ((double)b.z * kBlock + ((k & 4) ? 7 : 0)) * v;
    const double rx = wx - M[3], ry = wy - M[7], rz = wz - M[11];
    const double X = M[0] * rx + M[4] * ry + M[8] * rz;
    const double Y = M[1] * rx + M[5] * ry + M[9] * rz;
    const double Z = M[2] * rx + M[6] * ry + M[10] * rz;
    lo_z = Z < lo_z ? Z : lo_z;
    hi_z = Z > hi_z ? Z : hi_z;
    if (Z > 1e-4) {
      const double u = X * (double)fa.fx / Z + (double)fa.cx, vv = Y * (double)fa.fy / Z + (double)fa.cy;
      lo_u = u < lo_u ? u : lo_u;
      hi_u = u > hi_u ? u : hi_u;
      lo_v = vv < lo_v ? vv : lo_v;
      hi_v = vv > hi_v ? vv : hi_v;
    } else {
      all_front = false;
    }
  }
  const double tz = (double)fa.trunc * 1.001 + 1e-4;
  if (hi_z < zmin - tz || lo_z > zmax + tz) return false;
  if (all_front) {
    if (hi_u < -2.5 || lo_u > (double)fa.W + 0.5 || hi_v < -2.5 || lo_v > (double)fa.H + 0.5) return false;
  }
  return true;
}

__global__ __launch_bounds__(256) void k_compact(TableView t, DevState* st, FrameArgs fa, uint32_t* __restrict__ vis) {
  // the frame's block count after the insertions becomes the live count here
  const uint32_t n = st->n_alloc_next;
  if (blockIdx.x == 0u && threadIdx.x == 0u) st->n_alloc = n;
  FrameCounters& fc = st->set[fa.cset];
  const bool any_valid = fc.zmin_key != 0xFFFFFFFFu || fc.nonfinite != 0u;
  const bool cull = fa.cull_ok != 0 && fc.nonfinite == 0u;
  const double zmin = any_valid && cull ? (double)ordered_to_float(fc.zmin_key) : 0.0;
  const double zmax = any_valid && cull ? (double)ordered_to_float(fc.zmax_key) : 0.0;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  __shared__ uint32_t wcnt[4];
  __shared__ uint32_t wbase;
  for (uint32_t base = blockIdx.x * 256u; base < n; base += gridDim.x * 256u) {
    const uint32_t i = base + threadIdx.x;
    bool keep = false;
    if (i < n && any_valid) keep = !cull || block_maybe_visible(fa, t.blk[i], zmin, zmax);
    // wave ballot + prefix, then one atomic per workgroup for the list offset
    const uint64_t m = __ballot(keep);
    if (lane == 0) wcnt[w] = (uint32_t)__popcll(m);
    __syncthreads();
    if (threadIdx.x == 0) {
      const uint32_t tot = wcnt[0] + wcnt[1] + wcnt[2] + wcnt[3];
      wbase = tot ? atomicAdd(&fc.visible, tot) : 0u;
    }
    __syncthreads();
    uint32_t off = wbase;
    for (int k = 0; k < w; ++k) off += wcnt[k];
    if (keep) vis[off + (uint32_t)__popcll(m & ((1ull << lane) - 1ull))] = i;
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------------
// Integration (VoxelScene.cu:918-1146)
// ---------------------------------------------------------------------------------

// FP32 store: uint2 {sdf bits, weight}; FP16 store: uint32 {half sdf | weight << 16}.
// Colour plane: uint32 {r | g << 8 | b << 16 | flags << 24}.
template <bool F16> struct Store;
template <> struct Store<false> {
  using word = uint2;
  __device__ static float sdf(word w) { return __uint_as_float(w.x); }
  __device__ static uint32_t weight(word w) { return w.y & 0xFFFFu; }
  __device__ static word make(float s, uint32_t wt) { return make_uint2(__float_as_uint(s), wt); }
};
template <> struct Store<true> {
  using word = uint32_t;
  __device__ static float sdf(word w) { return __half2float(__ushort_as_half((unsigned short)(w & 0xFFFFu))); }
  __device__ static uint32_t weight(word w) { return w >> 16; }
  __device__ static word make(float s, uint32_t wt) {
    return (uint32_t)__half_as_ushort(__float2half_rn(s)) | (wt << 16);
  }
};

// Directional extension (DIR): the voxel's two direction slots {id, weight} are the
// bytes of dmeta[v] (id0, w0, id1, w1) and their dtsdf values dval[v].  The bin of
// the pixel's unit world normal is the signed axis of its largest component; no
// update when that component is below 0.8 (a zero normal included); the slot is
// the one holding the bin, else the first free one (weight 0), else none; dtsdf
// becomes (dtsdf * w + sdf) / (w + 1) and w min(w + 1, 100).  The world normal is
// the reference's (VoxelScene.cu:1074-1094): the camera-frame normal normalised when
// its length is positive, rotated, normalised when its length exceeds 1e-6.
__device__ __forceinline__ void dir_update(uint32_t* __restrict__ dmeta, float2* __restrict__ dval, size_t vi,
                                           const float* __restrict__ nrm, int pix, const float* M, float sdf) {
  float ncx = nrm[3 * pix], ncy = nrm[3 * pix + 1], ncz = nrm[3 * pix + 2];
  float l = sqrtf(ncx * ncx + ncy * ncy + ncz * ncz);
  if (l > 0.0f) {
    ncx /= l;
    ncy /= l;
    ncz /= l;
  }
  float nx = M[0] * ncx + M[1] * ncy + M[2] * ncz;
  float ny = M[4] * ncx + M[5] * ncy + M[6] * ncz;
  float nz = M[8] * ncx + M[9] * ncy + M[10] * ncz;
  l = sqrtf(nx * nx + ny * ny + nz * nz);
  if (l > 1e-6f) {
    nx /= l;
    ny /= l;
    nz /= l;
  }
  const float ax = fabsf(nx), ay = fabsf(ny), az = fabsf(nz);
  int a = 0;
  float c = ax;
  if (ay > c) {
    a = 1;
    c = ay;
  }
  if (az > c) {
    a = 2;
    c = az;
  }
  if (!(c >= 0.8f)) return;
  const float comp = a == 0 ? nx : (a == 1 ? ny : nz);
  const uint32_t id = 1u + 2u * (uint32_t)a + (comp < 0.0f ? 1u : 0u);
  uint32_t meta = dmeta[vi];
  const uint32_t id0 = meta & 0xFFu, w0 = (meta >> 8) & 0xFFu, id1 = (meta >> 16) & 0xFFu, w1 = meta >> 24;
  int slot = id0 == id ? 0 : (id1 == id ? 1 : (w0 == 0u ? 0 : (w1 == 0u ? 1 : -1)));
  if (slot < 0) return;
  float2 d = dval[vi];
  const float ow = (float)(slot == 0 ? w0 : w1);
  const float nw = ow + 1.0f;
  const float nd = ((slot == 0 ? d.x : d.y) * ow + sdf * 1.0f) / nw;
  const uint32_t wq = (uint32_t)fminf(100.0f, nw);
  if (slot == 0) {
    d.x = nd;
    meta = (meta & 0xFFFF0000u) | id | (wq << 8);
  } else {
    d.y = nd;
    meta = (meta & 0x0000FFFFu) | (id << 16) | (wq << 24);
  }
  dval[vi] = d;
  dmeta[vi] = meta;
}

// Adds this wave's positive-voxel changes (delta in {-1, 0, 1} per voxel: +1 when
// the voxel became positive, weight > 0 and SDF > 0, -1 when it stopped being) to
// the block's per-cell counts.  A wave holds one z slice of the block: lanes
// ly * 8 + lx, cells (lx >> 2) | (ly >> 2) << 1 | (lz >> 2) << 2.
__device__ __forceinline__ void wave_cell_delta(int delta, int lz, uint32_t* counts) {
  const uint64_t plus = __ballot(delta > 0), minus = __ballot(delta < 0);
  if ((plus | minus) == 0ull) return;
  const int lane = threadIdx.x & 63;
  if (lane < 4) {
    const uint64_t m = (0x0F0F0F0Full << (4 * (lane & 1))) << (32 * (lane >> 1));
    const int d = __popcll(plus & m) - __popcll(minus & m);
    if (d != 0) atomicAdd(&counts[(lane & 1) | ((lane >> 1) << 1) | ((lz >> 2) << 2)], (uint32_t)d);
  }
}

// Positive-cell mask of a block from its per-cell counts.
__device__ __forceinline__ uint32_t cell_mask_of(const uint32_t* counts) {
  const uint4 a = reinterpret_cast<const uint4*>(counts)[0], b = reinterpret_cast<const uint4*>(counts)[1];
  return (a.x ? 1u : 0u) | (a.y ? 2u : 0u) | (a.z ? 4u : 0u) | (a.w ? 8u : 0u) | (b.x ? 16u : 0u) |
         (b.y ? 32u : 0u) | (b.z ? 64u : 0u) | (b.w ? 128u : 0u);
}

template <bool F16, bool DIR>
__global__ __launch_bounds__(512) void k_integrate(TableView t, DevState* st, FrameArgs fa,
                                                   const float* __restrict__ pts, const uint8_t* __restrict__ rgb,
                                                   const float* __restrict__ nrm, const uint32_t* __restrict__ vis,
                                                   void* __restrict__ vox_, uint32_t* __restrict__ col,
                                                   uint32_t* __restrict__ dmeta, float2* __restrict__ dval,
                                                   uint32_t* __restrict__ upd_part) {
  using S = Store<F16>;
  typename S::word* vox = reinterpret_cast<typename S::word*>(vox_);
  FrameCounters& fc = st->set[fa.cset];
  const uint32_t nvis = fc.visible;
  const int idx = threadIdx.x;
  const int lx = idx % kBlock, ly = (idx / kBlock) % kBlock, lz = idx / (kBlock * kBlock);
  const float* M = fa.M;
  const float v = fa.voxel, tr = fa.trunc;
  const float camx = M[3], camy = M[7], camz = M[11];
  uint32_t upd = 0;
  for (uint32_t i = blockIdx.x; i < nvis; i += gridDim.x) {
    const uint32_t ptr = vis[i];
    const int4 b = t.blk[ptr];
    int delta = 0;  // +1: the voxel became positive (weight > 0, SDF > 0); -1: it stopped being
    do {
      const float vwx = (float)(b.x * kBlock + lx) * v;
      const float vwy = (float)(b.y * kBlock + ly) * v;
      const float vwz = (float)(b.z * kBlock + lz) * v;
      const float rx = vwx - M[3], ry = vwy - M[7], rz = vwz - M[11];
      const float vcx = M[0] * rx + M[4] * ry + M[8] * rz;
      const float vcy = M[1] * rx + M[5] * ry + M[9] * rz;
      const float vcz = M[2] * rx + M[6] * ry + M[10] * rz;
      const float sxf = vcx * fa.fx / vcz + fa.cx;
      const float syf = vcy * fa.fy / vcz + fa.cy;
      const int32_t px = cvt_i32(sxf + 0.5f), py = cvt_i32(syf + 0.5f);
      if (!(px >= 0 && px < fa.W && py >= 0 && py < fa.H)) break;
      const int pix = py * fa.W + px;
      const float Px = pts[3 * pix], Py = pts[3 * pix + 1], Pz = pts[3 * pix + 2];
      if (!(Px != 0.0f || Py != 0.0f || Pz != 0.0f)) break;
      const float3 D = xform_point(M, Px, Py, Pz);
      const float sdx = vwx - D.x, sdy = vwy - D.y, sdz = vwz - D.z;
      float sdf = sqrtf(sdx * sdx + sdy * sdy + sdz * sdz);
      // the truncation test (sdf < -tr || sdf > tr) depends on |sdf| only: the same as
      // sdf > tr before the sign is applied (NaN passes both), so the two distances that
      // decide the sign are computed only for voxels within the band
      if (sdf > tr) break;
      const float cdx = D.x - camx, cdy = D.y - camy, cdz = D.z - camz;
      const float cam_dist = sqrtf(cdx * cdx + cdy * cdy + cdz * cdz);
      const float vox_dist =
          sqrtf((vwx - camx) * (vwx - camx) + (vwy - camy) * (vwy - camy) + (vwz - camz) * (vwz - camz));
      if (vox_dist < cam_dist) sdf = -sdf;
      sdf = fmaxf(-tr, fminf(tr, sdf));
      const size_t vi = (size_t)ptr * kVox + (size_t)idx;
      const typename S::word w = vox[vi];
      const uint32_t c = col[vi];
      const float old_sdf = S::sdf(w);
      const float old_w = (float)S::weight(w);
      const float new_w = old_w + 1.0f;
      if (new_w > 0.0f) {
        const float nsdf = (old_sdf * old_w + sdf * 1.0f) / new_w;
        const uint32_t nw = (uint32_t)fminf(255.0f, new_w);
        const typename S::word nwd = S::make(nsdf, nw);
        const float stored = S::sdf(nwd);
        const uint8_t* cp = rgb + 3 * pix;
        uint32_t nc = 0;
#pragma unroll
        for (int ch = 0; ch < 3; ++ch) {
          const float oc = (float)((c >> (8 * ch)) & 0xFFu);
          const float ncf = (oc * old_w + (float)cp[ch] * 1.0f) / new_w;
          nc |= ((uint32_t)ncf & 0xFFu) << (8 * ch);
        }
        uint32_t fl = (c >> 24) | 1u;
        if ((old_sdf > 0.0f && stored <= 0.0f) || (old_sdf <= 0.0f && stored > 0.0f)) fl |= 2u;
        vox[vi] = nwd;
        col[vi] = nc | (fl << 24);
        if (DIR) dir_update(dmeta, dval, vi, nrm, pix, M, sdf);
        upd++;
        const bool was_pos = S::weight(w) != 0u && old_sdf > 0.0f;
        const bool is_pos = stored > 0.0f;
        delta = (int)is_pos - (int)was_pos;
      }
    } while (false);
    wave_cell_delta(delta, lz, t.pcnt + (size_t)ptr * 8u);
  }
  // one write per workgroup (summed when the counters are read)
  __shared__ uint32_t ws[8];
  upd = wave_sum(upd);
  if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = upd;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t tot = 0;
    for (int k = 0; k < 8; ++k) tot += ws[k];
    upd_part[blockIdx.x] = tot;
  }
}

// Positive-cell masks of the frame's visible blocks (the only blocks whose counts
// can have changed) into their block index entries, after k_integrate.
__global__ __launch_bounds__(256) void k_update_masks(TableView t, const DevState* __restrict__ st, int32_t cset,
                                                      const uint32_t* __restrict__ vis) {
  const uint32_t nvis = st->set[cset].visible;
  for (uint32_t i = blockIdx.x * 256u + threadIdx.x; i < nvis; i += gridDim.x * 256u) {
    const uint32_t ptr = vis[i];
    t.idx[t.idxpos[ptr]].w = cell_mask_of(t.pcnt + (size_t)ptr * 8u);
  }
}

// Recomputes the per-cell positive-voxel counts and the positive-cell masks of live
// blocks [first, first + n) from their voxels (after the voxels were overwritten
// from the host).  One workgroup per block.
template <bool F16>
__global__ __launch_bounds__(512) void k_cell_masks(TableView t, const DevState* __restrict__ st, uint32_t first,
                                                    uint32_t n, const void* __restrict__ vox_) {
  using S = Store<F16>;
  const typename S::word* vox = reinterpret_cast<const typename S::word*>(vox_);
  const int idx = threadIdx.x;
  const int lz = idx / (kBlock * kBlock);
  __shared__ __align__(16) uint32_t cnt[8];
  const uint32_t live = st->n_alloc;  // blocks beyond it have no index entry
  n = first >= live ? 0u : (n < live - first ? n : live - first);
  for (uint32_t i = blockIdx.x; i < n; i += gridDim.x) {
    const uint32_t ptr = first + i;
    if (idx < 8) cnt[idx] = 0u;
    __syncthreads();
    const typename S::word w = vox[(size_t)ptr * kVox + (size_t)idx];
    wave_cell_delta(S::weight(w) != 0u && S::sdf(w) > 0.0f ? 1 : 0, lz, cnt);
    __syncthreads();
    if (idx < 8) t.pcnt[(size_t)ptr * 8u + (uint32_t)idx] = cnt[idx];
    if (idx == 0) t.idx[t.idxpos[ptr]].w = cell_mask_of(cnt);
    __syncthreads();
  }
}

// Positive-cell masks of blocks [first, first + n) from their block index entries
// (0 for a block without an entry).
__global__ __launch_bounds__(256) void k_export_cell_masks(TableView t, uint32_t first, uint32_t n,
                                                           uint8_t* __restrict__ out) {
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  if (i >= n) return;
  const int4 b = t.blk[first + i];
  const int32_t sl = index_slot(t, b.x, b.y, b.z);  // by key: also checks idxpos
  out[i] = sl >= 0 && t.idx[sl].z == first + i && t.idxpos[first + i] == (uint32_t)sl ? (uint8_t)t.idx[sl].w : 0u;
}

// ---------------------------------------------------------------------------------
// Raycast (RayCastRender.cu:127-517)
// ---------------------------------------------------------------------------------

// Block pointers and positive-cell masks of the 2x2x2 blocks around the ray's
// current sample, looked up on demand.  The window holds blocks [bx, bx + 1] x
// [by, by + 1] x [bz, bz + 1]; block (x, y, z) sits in slot (x & 1) | (y & 1) << 1 |
// (z & 1) << 2, so when the window moves by one block along an axis the blocks that
// stay keep their slots and only the slots of the blocks that leave are forgotten.
// The slots' pointers live in LDS (slot s of lane l at slots[64 s + l]: lanes read
// different slots without a select network); bit s of `known`: slot s holds its
// window block; byte s of pm: its positive-cell mask.  DIAG marks every bucket with
// live entries that the reference's lookup of a used block probes and every block
// found (byte-model counters).
template <bool DIAG> struct BlockCache {
  int32_t bx, by, bz;
  int32_t* slots;  // this lane's slot 0 in LDS (slot s at slots[64 s])
  uint64_t pm;
  uint32_t known;
  uint8_t* mark_bkt;
  uint8_t* mark_blk;
  uint32_t* stats;  // march statistics (evaluations, super-block skips, block skips, replayed steps, cell skips)
  __device__ void init(int32_t* lane_slots, uint8_t* mb, uint8_t* mk, uint32_t* st = nullptr) {
    bx = by = bz = INT32_MIN;
    known = 0u;
    pm = 0ull;
    slots = lane_slots;
    mark_bkt = mb;
    mark_blk = mk;
    stats = st;
  }
  __device__ __forceinline__ int32_t get(const TableView& t, int32_t x, int32_t y, int32_t z, uint32_t* mask) {
    const int32_t q = index_lookup(t, x, y, z, mask);
    if (DIAG) mark(t, x, y, z, q);
    return q;
  }
  // Diagnostics of a lookup of block (x, y, z) that found pointer q (-1: absent).
  __device__ __forceinline__ void mark(const TableView& t, int32_t x, int32_t y, int32_t z, int32_t q) {
    if (key_in_range(x, y, z)) {
      const uint32_t h = hash_bucket(x, y, z, t.num_buckets);
      if (t.cnt[h] > 0) mark_bkt[h] = 1;
    }
    if (q >= 0) mark_blk[q] = 1;
  }
  // The window along one axis after a request for blocks [c0, c0 + s]: unchanged
  // when they fit, else moved (by one block: the block that leaves frees its slots;
  // farther: nothing is kept).  lo: the slots whose bit of this axis is clear.
  __device__ __forceinline__ int32_t move_axis(int32_t base, int32_t c0, bool s, uint32_t lo) {
    const int32_t d = (int32_t)((uint32_t)c0 - (uint32_t)base);
    if (d == 0 || (d == 1 && !s)) return base;
    if (d == 1 || d == -1) {
      const int32_t gone = d == 1 ? base : base + 1;
      known &= (gone & 1) ? lo : (~lo & 0xFFu);
      return d == 1 ? base + 1 : base - 1;
    }
    known = 0u;
    return c0;
  }
  // Makes blocks {x0, x0 + sx} x {y0, y0 + sy} x {z0, z0 + sz} known.
  __device__ __forceinline__ void ensure(const TableView& t, int32_t x0, int32_t y0, int32_t z0, bool sx, bool sy,
                                         bool sz) {
    bx = move_axis(bx, x0, sx, 0x55u);
    by = move_axis(by, y0, sy, 0x33u);
    bz = move_axis(bz, z0, sz, 0x0Fu);
    const uint32_t ex = sx ? 0xFFu : ((x0 & 1) ? 0xAAu : 0x55u);
    const uint32_t ey = sy ? 0xFFu : ((y0 & 1) ? 0xCCu : 0x33u);
    const uint32_t ez = sz ? 0xFFu : ((z0 & 1) ? 0xF0u : 0x0Fu);
    uint32_t missing = ex & ey & ez & ~known;
    known |= missing;
    // Missing blocks are looked up two at a time: both first index probes are issued
    // before either is used (one memory round trip per pair); a probe that meets
    // another key continues along the probe sequence.
    while (missing) {
      const int i = __builtin_ctz(missing);
      missing &= missing - 1u;
      const int j = missing ? __builtin_ctz(missing) : i;
      missing &= missing - 1u;
      const uint32_t hi_ = idx_home(slot_x(i), slot_y(i), slot_z(i), t.idx_mask);
      const uint32_t hj = idx_home(slot_x(j), slot_y(j), slot_z(j), t.idx_mask);
      const uint4 ei = t.idx[hi_], ej = t.idx[hj];
      resolve(t, i, hi_, ei);
      if (j != i) resolve(t, j, hj, ej);
    }
  }
  // Block of slot i from the first probe e (at home slot s) of its index lookup.
  __device__ __forceinline__ void resolve(const TableView& t, int i, uint32_t s, uint4 e) {
    const int32_t x = slot_x(i), y = slot_y(i), z = slot_z(i);
    int32_t q = -1;
    uint32_t mk = 0u;
    if (key_in_range(x, y, z)) {
      const uint64_t key = pack_key(x, y, z);
      const uint32_t klo = (uint32_t)key, khi = (uint32_t)(key >> 32);
      for (uint32_t n = 0; n <= t.idx_mask; ++n) {
        if (e.x == klo && e.y == khi) {
          q = (int32_t)e.z;
          mk = e.w;
          break;
        }
        if (e.x == kIdxEmpty && e.y == kIdxEmpty) break;
        s = (s + 1u) & t.idx_mask;
        e = t.idx[s];
      }
    }
    if (DIAG) mark(t, x, y, z, q);
    slots[64 * i] = q;
    pm = (pm & ~(0xFFull << (8 * i))) | ((uint64_t)(mk & 0xFFu) << (8 * i));
  }
  // Coordinates of the window block with slot i's parities.
  __device__ __forceinline__ int32_t slot_x(int i) const { return bx + (int32_t)(((uint32_t)i ^ (uint32_t)bx) & 1u); }
  __device__ __forceinline__ int32_t slot_y(int i) const {
    return by + (int32_t)((((uint32_t)i >> 1) ^ (uint32_t)by) & 1u);
  }
  __device__ __forceinline__ int32_t slot_z(int i) const {
    return bz + (int32_t)((((uint32_t)i >> 2) ^ (uint32_t)bz) & 1u);
  }
  // Pointers of the eight corner blocks of a sample whose corners lie in blocks
  // {x0, x0 + sx} x {y0, y0 + sy} x {z0, z0 + sz}; corner i takes x from bit 0, y from
  // bit 1, z from bit 2.
  __device__ __forceinline__ void corners(const TableView& t, int32_t x0, int32_t y0, int32_t z0, bool sx, bool sy,
                                          bool sz, int32_t out[8]) {
    ensure(t, x0, y0, z0, sx, sy, sz);
    const uint32_t px = (uint32_t)x0 & 1u, py = ((uint32_t)y0 & 1u) << 1, pz = ((uint32_t)z0 & 1u) << 2;
    const uint32_t qx = sx ? 1u : 0u, qy = sy ? 2u : 0u, qz = sz ? 4u : 0u;
#pragma unroll
    for (int i = 0; i < 8; ++i) out[i] = slots[64 * ((px ^ (i & qx)) | (py ^ (i & qy)) | (pz ^ (i & qz)))];
  }
  // True when no cell of cells [cx0, cx1] x [cy0, cy1] x [cz0, cz1] (one or two per
  // axis, 4^3 voxels each) holds a voxel with weight > 0 and SDF > 0.
  __device__ __forceinline__ bool cells_free(const TableView& t, int32_t cx0, int32_t cx1, int32_t cy0, int32_t cy1,
                                             int32_t cz0, int32_t cz1) {
    ensure(t, cx0 >> 1, cy0 >> 1, cz0 >> 1, (cx1 >> 1) != (cx0 >> 1), (cy1 >> 1) != (cy0 >> 1),
           (cz1 >> 1) != (cz0 >> 1));
    // bit of cell c in pm: 8 * slot + cell in block; per axis the slot bit is bit 1
    // of c (the block's parity) and the cell in block bit 0 of c
    const uint32_t x0 = (((uint32_t)cx0 & 2u) << 2) | ((uint32_t)cx0 & 1u);
    const uint32_t x1 = (((uint32_t)cx1 & 2u) << 2) | ((uint32_t)cx1 & 1u);
    const uint32_t y0 = (((uint32_t)cy0 & 2u) << 3) | (((uint32_t)cy0 & 1u) << 1);
    const uint32_t y1 = (((uint32_t)cy1 & 2u) << 3) | (((uint32_t)cy1 & 1u) << 1);
    const uint32_t z0 = (((uint32_t)cz0 & 2u) << 4) | (((uint32_t)cz0 & 1u) << 2);
    const uint32_t z1 = (((uint32_t)cz1 & 2u) << 4) | (((uint32_t)cz1 & 1u) << 2);
    const uint64_t mx = (1ull << x0) | (1ull << x1);
    const uint64_t mxy = (mx << y0) | (mx << y1);
    const uint64_t m = (mxy << z0) | (mxy << z1);
    return (pm & m) == 0ull;
  }
};

// Sample status of a trilinear evaluation.
enum : int { kSampleInvalid = 1, kSampleValid = 2 };

// trilinearInterpolationSimpleFastFast (RayCastRender.cu:127-215).  The eight
// corners' voxel indices come from two quotients per axis (posDual/v and
// (posDual+v)/v): the reference computes the same quotient once per corner.
// Returns kSampleValid with *dist (and *rgb), kSampleInvalid when no corner voxel
// has a weight (the reference's sum of weights <= 1e-6).
// Corner voxels and weights of a sample.  REF: the reference's scheme (weights
// frac(p/v - 1/2), corners rounded from p - v/2 and p + v/2: RayCastRender.cu:139-176),
// which samples the stored field half a voxel off.  Otherwise standard trilinear
// interpolation of voxels at i*v (weights frac(p/v), corners floor(p/v) + {0,1}),
// offered for tracking; it is not the parity path.
template <bool REF>
__device__ __forceinline__ void corner_setup(float px, float py, float pz, float v, int32_t vx[2], int32_t vy[2],
                                             int32_t vz[2], float* wx, float* wy, float* wz) {
  // quotients by the voxel size, correctly rounded (as the reference's divisions)
  const float rv = 1.0f / v;
  if (REF) {
    const float off = v;
    const float pdx = px - off / 2.0f, pdy = py - off / 2.0f, pdz = pz - off / 2.0f;
    const float qx0 = div_fixed(pdx, v, rv), qy0 = div_fixed(pdy, v, rv), qz0 = div_fixed(pdz, v, rv);
    *wx = qx0 - floorf(qx0);
    *wy = qy0 - floorf(qy0);
    *wz = qz0 - floorf(qz0);
    const float qx1 = div_fixed(pdx + off, v, rv), qy1 = div_fixed(pdy + off, v, rv), qz1 = div_fixed(pdz + off, v, rv);
    vx[0] = ray_voxel_of_q(qx0);
    vx[1] = ray_voxel_of_q(qx1);
    vy[0] = ray_voxel_of_q(qy0);
    vy[1] = ray_voxel_of_q(qy1);
    vz[0] = ray_voxel_of_q(qz0);
    vz[1] = ray_voxel_of_q(qz1);
  } else {
    const float qx = div_fixed(px, v, rv), qy = div_fixed(py, v, rv), qz = div_fixed(pz, v, rv);
    const float fx = floorf(qx), fy = floorf(qy), fz = floorf(qz);
    *wx = qx - fx;
    *wy = qy - fy;
    *wz = qz - fz;
    vx[0] = cvt_i32(fx);
    vy[0] = cvt_i32(fy);
    vz[0] = cvt_i32(fz);
    vx[1] = vx[0] + 1;
    vy[1] = vy[0] + 1;
    vz[1] = vz[0] + 1;
  }
}

// Corners of one sample: the store index of each corner voxel (0 for a corner
// without an allocated block, flagged in `none`) and the interpolation weights.
struct Corners {
  uint32_t vi[8];
  uint32_t none;  // bit i: corner i has no allocated block
  float wx, wy, wz;
};

// Corner setup and block pointers of the sample at (px, py, pz): the block lookups
// (through the lane's block cache) happen here, the voxel reads in corner_gather.
template <bool REF, class Cache>
__device__ __forceinline__ void corner_prep(const TableView& t, Cache& bc, float px, float py, float pz, float v,
                                            Corners& cn) {
  int32_t vx[2], vy[2], vz[2];
  corner_setup<REF>(px, py, pz, v, vx, vy, vz, &cn.wx, &cn.wy, &cn.wz);
  // block (floor division by 8) and offset in the block per axis and corner
  const int32_t bx0 = vx[0] >> 3, bx1 = vx[1] >> 3, by0 = vy[0] >> 3, by1 = vy[1] >> 3;
  const int32_t bz0 = vz[0] >> 3, bz1 = vz[1] >> 3;
  const uint32_t ox[2] = {(uint32_t)vx[0] & 7u, (uint32_t)vx[1] & 7u};
  const uint32_t oy[2] = {((uint32_t)vy[0] & 7u) << 3, ((uint32_t)vy[1] & 7u) << 3};
  const uint32_t oz[2] = {((uint32_t)vz[0] & 7u) << 6, ((uint32_t)vz[1] & 7u) << 6};
  int32_t p[8];
  bc.corners(t, bx0, by0, bz0, bx1 != bx0, by1 != by0, bz1 != bz0, p);
  cn.none = 0u;
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    const bool ok = p[i] >= 0;
    cn.vi[i] = ok ? (uint32_t)p[i] * (uint32_t)kVox + ox[i & 1] + oy[(i >> 1) & 1] + oz[(i >> 2) & 1] : 0u;
    cn.none |= ok ? 0u : (1u << i);
  }
}

// The interpolation of trilinear() from prepared corners.  Corners without an
// allocated block or with weight 0 contribute a zero weight: adding +-0 leaves every
// sum unchanged (a sum started at +0 never becomes -0), so this equals skipping
// them.  Every corner is loaded (a corner without a block reads voxel 0 of the store
// and its values are replaced by zeros), so the gather has no branches.
template <bool F16, bool COLOR>
__device__ __forceinline__ int corner_gather(const void* __restrict__ vox_, const uint32_t* __restrict__ col,
                                             const Corners& cn, float* dist, uint32_t* rgb) {
  using S = Store<F16>;
  const typename S::word* vox = reinterpret_cast<const typename S::word*>(vox_);
  const float wx = cn.wx, wy = cn.wy, wz = cn.wz;
  const float ux = 1.0f - wx, uy = 1.0f - wy, uz = 1.0f - wz;
  float d = 0.0f, cr = 0.0f, cg = 0.0f, cbl = 0.0f, wsum = 0.0f;
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    const bool ok = !((cn.none >> i) & 1u);
    const uint32_t vi = cn.vi[i];
    const typename S::word w = vox[vi];
    const uint32_t c = COLOR ? (ok ? col[vi] : 0u) : 0u;
    const float wi = ((i & 1) ? wx : ux) * ((i & 2) ? wy : uy) * ((i & 4) ? wz : uz);
    const float wk = ok && S::weight(w) != 0u ? wi : 0.0f;
    d += wk * (ok ? S::sdf(w) : 0.0f);
    if (COLOR) {
      cr += wk * (float)(c & 0xFFu);
      cg += wk * (float)((c >> 8) & 0xFFu);
      cbl += wk * (float)((c >> 16) & 0xFFu);
    }
    wsum += wk;
    // with colours, gather the corners in two halves: the colour words would
    // otherwise double the registers held by loads in flight
    if (COLOR && i == 3) __builtin_amdgcn_sched_barrier(0);
  }
  if (wsum <= 1e-6f) return kSampleInvalid;
  const float inv = 1.0f / wsum;
  *dist = d * inv;
  if (COLOR) {
    cr *= inv;
    cg *= inv;
    cbl *= inv;
    *rgb = ((uint32_t)cr & 0xFFu) | (((uint32_t)cg & 0xFFu) << 8) | (((uint32_t)cbl & 0xFFu) << 16);
  }
  return kSampleValid;
}

// The distance of corner_gather<F16, false> from the eight corner words w, already read.
template <bool F16>
__device__ __forceinline__ int interp_words(const Corners& cn, const typename Store<F16>::word w[8], float* dist) {
  using S = Store<F16>;
  const float wx = cn.wx, wy = cn.wy, wz = cn.wz;
  const float ux = 1.0f - wx, uy = 1.0f - wy, uz = 1.0f - wz;
  float d = 0.0f, wsum = 0.0f;
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    const bool ok = !((cn.none >> i) & 1u);
    const float wi = ((i & 1) ? wx : ux) * ((i & 2) ? wy : uy) * ((i & 4) ? wz : uz);
    const float wk = ok && S::weight(w[i]) != 0u ? wi : 0.0f;
    d += wk * (ok ? S::sdf(w[i]) : 0.0f);
    wsum += wk;
  }
  if (wsum <= 1e-6f) return kSampleInvalid;
  *dist = d * (1.0f / wsum);
  return kSampleValid;
}

// Two samples' distances (corner_gather<F16, false> of each) with all sixteen corner
// reads issued before either interpolation.
template <bool F16>
__device__ __forceinline__ void corner_gather2(const void* __restrict__ vox_, const Corners& ca, const Corners& cb,
                                               int* sa, float* da, int* sb, float* db) {
  using S = Store<F16>;
  const typename S::word* vox = reinterpret_cast<const typename S::word*>(vox_);
  typename S::word wa[8], wb[8];
#pragma unroll
  for (int i = 0; i < 8; ++i) wa[i] = vox[ca.vi[i]];
#pragma unroll
  for (int i = 0; i < 8; ++i) wb[i] = vox[cb.vi[i]];
  *sa = interp_words<F16>(ca, wa, da);
  *sb = interp_words<F16>(cb, wb, db);
}

template <bool F16, bool COLOR, bool REF, class Cache>
__device__ __forceinline__ int trilinear(const TableView& t, Cache& bc, const void* __restrict__ vox_,
                                         const uint32_t* __restrict__ col, float px, float py, float pz, float v,
                                         float* dist, uint32_t* rgb) {
  Corners cn;
  corner_prep<REF>(t, bc, px, py, pz, v, cn);
  return corner_gather<F16, COLOR>(vox_, col, cn, dist, rgb);
}

// ---- exact empty-space skipping ---------------------------------------------------
//
// A sample is invalid whenever none of its eight corner voxels lies in an allocated
// block.  Per axis, the corner voxel indices of a sample at p lie in (Q - 2, Q + 1]
// with Q = p/v (reference sampling: r(q) and r(q + 1) with q = Q - 1/2 and r(q) in
// (q - 3/2, q + 1/2]; standard trilinear: floor(Q) and floor(Q) + 1).  While Q stays
// inside [8*b0 + 1, 8*b1 + 7) on every axis, all corners fall in blocks b0..b1; if
// none of those blocks is allocated, every such sample is invalid and the march only
// has to replay its step additions (rayCurrent += v/2) to land on the same sample
// positions as the reference.  The same holds for 32-voxel super-blocks s0..s1 and
// [32*s0 + 1, 32*s1 + 31).  Both tests read the block occupancy map (a hash collision
// only makes a block look allocated); boxes are shrunk by a margin covering the float
// rounding of p and Q.

struct RaySkip {
  float camx, camy, camz;
  float wdx, wdy, wdz;  // direction
  float v, inv_v;
  float eps;            // safety margin in voxels
};

// Parameter interval [t0, t1) of the ray inside the box [lo, hi] (voxel units,
// shrunk by the margin); t1 <= t0 when the ray misses it.
__device__ __forceinline__ void box_interval(const RaySkip& k, float lox, float hix, float loy, float hiy, float loz,
                                             float hiz, float* t0, float* t1) {
  float tn = -INFINITY, tf = INFINITY;
  const float lo[3] = {lox, loy, loz}, hi[3] = {hix, hiy, hiz};
  const float c[3] = {k.camx, k.camy, k.camz}, d[3] = {k.wdx, k.wdy, k.wdz};
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const float l = (lo[a] + k.eps) * k.v - c[a], h = (hi[a] - k.eps) * k.v - c[a];
    if (d[a] == 0.0f) {  // parallel to this slab: inside for every t, or never
      if (!(l <= 0.0f && 0.0f <= h)) { tf = -INFINITY; }
      continue;
    }
    // hardware reciprocal (1 ulp): its error is far inside the relative margin below
    const float id = __builtin_amdgcn_rcpf(d[a]);
    const float ta = l * id, tb = h * id;
    tn = fmaxf(tn, fminf(ta, tb));
    tf = fminf(tf, fmaxf(ta, tb));
  }
  // shrink by a relative margin against the rounding of the products above
  *t0 = tn + fabsf(tn) * 1e-5f + 1e-7f;
  *t1 = tf - fabsf(tf) * 1e-5f - 1e-7f;
}

// Occupancy of the blocks around a sample: the range [bx0, bx1] x [by0, by1] x
// [bz0, bz1] (one or two blocks per axis) and the super-blocks holding it.  The lane
// keeps the last occupancy word it read (consecutive samples mostly share it).
struct OccCache {
  uint32_t idx;        // map index of the cached word
  uint64_t word;
  uint32_t hidx;       // hyper-block map index of the cached hyper word
  uint64_t hword;
  int32_t rx, ry, rz;  // block range of the last test that found an allocated block:
  uint32_t rs;         // lower corner, bits 0-2 its extent minus one per axis, bit 3 set when valid
};

__device__ __forceinline__ void occ_cache_init(OccCache& oc) {
  oc.idx = 0xFFFFFFFFu;
  oc.word = 0ull;
  oc.hidx = 0xFFFFFFFFu;
  oc.hword = 0ull;
  oc.rs = 0u;
}

__device__ __forceinline__ uint64_t occ_word(const TableView& t, OccCache& oc, int32_t sx, int32_t sy, int32_t sz) {
  const uint32_t wi = occ_index(sx, sy, sz, t.occ_mask);
  if (wi != oc.idx) {
    oc.idx = wi;
    oc.word = t.occ[wi];
  }
  return oc.word;
}

// 0: no block of the super-blocks is allocated; 1: some block of the super-blocks
// but none of the range is; 2: a block of the range is allocated.
__device__ __forceinline__ int occ_test(const TableView& t, OccCache& oc, int32_t bx0, int32_t bx1, int32_t by0,
                                        int32_t by1, int32_t bz0, int32_t bz1) {
  // one word per super-block the range touches (one or two per axis); in each, the
  // bits x + 4y + 16z of the range's blocks (local coordinates in the super-block)
  uint64_t any_word = 0ull, any_bit = 0ull;
  const int32_t sx0 = bx0 >> 2, sy0 = by0 >> 2, sz0 = bz0 >> 2;
  const uint32_t ex = (bx1 >> 2) != sx0 ? 1u : 0u, ey = (by1 >> 2) != sy0 ? 1u : 0u, ez = (bz1 >> 2) != sz0 ? 1u : 0u;
  const uint32_t nsb = 1u << (ex + ey + ez);
  for (uint32_t k = 0; k < nsb; ++k) {
    // k enumerates the split axes' choices (x first)
    uint32_t b = k;
    const uint32_t cx = ex ? (b & 1u) : 0u;
    b >>= ex;
    const uint32_t cy = ey ? (b & 1u) : 0u;
    b >>= ey;
    const uint32_t cz = ez ? (b & 1u) : 0u;
    // blocks of the range in this super-block: both, or the one on its side of the split
    const int32_t xl = ex ? (cx ? bx1 : bx0) : bx0, xh = ex ? xl : bx1;
    const int32_t yl = ey ? (cy ? by1 : by0) : by0, yh = ey ? yl : by1;
    const int32_t zl = ez ? (cz ? bz1 : bz0) : bz0, zh = ez ? zl : bz1;
    const uint32_t mx = (1u << (xl & 3)) | (1u << (xh & 3));
    const uint32_t mxy = (mx << ((yl & 3) << 2)) | (mx << ((yh & 3) << 2));
    const uint64_t m = ((uint64_t)mxy << ((zl & 3) << 4)) | ((uint64_t)mxy << ((zh & 3) << 4));
    const uint64_t w = occ_word(t, oc, sx0 + (int32_t)cx, sy0 + (int32_t)cy, sz0 + (int32_t)cz);
    any_word |= w;
    any_bit |= w & m;
  }
  return any_bit ? 2 : (any_word ? 1 : 0);
}

// One ray of renderKernel (RayCastRender.cu:407-517) with the march of
// traverseCoarseGridSimpleSampleAll (RayCastRender.cu:276-398).  Also returns the
// bracket [a, b] at which the accepted crossing was refined (the adjoint restarts
// there).
struct RayHit {
  bool marched;  // depth range valid (RayCastRender.cu:461-462)
  bool found;
  bool deferred;  // step budget spent: the march continues from the state below
  float alpha, d2r;
  float cdx, cdy, cdz, wdx, wdy, wdz;
  float a, ad, b, bd;
  uint32_t rgb;
  float ray, last_sdf, last_alpha;  // state of a deferred march
  uint32_t last_w;
};

// Per-ray constants of the march.
struct RaySetup {
  float camx, camy, camz;
  float ray0, ray_end, inc, th, v;
  RaySkip sk;
};

// Rows 0-2 of the render's camera->world pose: the caller's device copy when the
// render was given one (RayCastRender.h:66-70 passes a device pointer), else the copy
// passed by value.  All lanes read the same words.
__device__ __forceinline__ float uniform_load(const float* p) {
  // the address is the same in every lane: keep the value in a scalar register
  return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(*p)));
}

// DEV: the render reads its pose from ra.Mdev and its camera position from
// ra.camdev (both device memory); otherwise from the by-value copies.  A compile-time
// choice, so that the by-value kernels carry no code or registers for the other.
template <bool DEV> __device__ __forceinline__ void render_pose(const RenderArgs& ra, float M[12]) {
  if (DEV) {
#pragma unroll
    for (int i = 0; i < 12; ++i) M[i] = uniform_load(ra.Mdev + i);
  } else {
#pragma unroll
    for (int i = 0; i < 12; ++i) M[i] = ra.M[i];
  }
}

// The render's camera position (device copy or by value, as render_pose).
template <bool DEV> __device__ __forceinline__ void render_cam(const RenderArgs& ra, float c[3]) {
  if (DEV) {
    c[0] = uniform_load(ra.camdev);
    c[1] = uniform_load(ra.camdev + 1);
    c[2] = uniform_load(ra.camdev + 2);
  } else {
    c[0] = ra.cam[0];
    c[1] = ra.cam[1];
    c[2] = ra.cam[2];
  }
}

// Ray of pixel (x, y) (RayCastRender.cu:407-470).  Returns false when the depth
// range is invalid (no march; the pixel keeps the miss values).
template <bool DEV>
__device__ __forceinline__ bool ray_setup(const RenderArgs& ra, uint32_t x, uint32_t y, RayHit& h, RaySetup& rs) {
  const float MINF = 1e10f;
  h.marched = false;
  h.found = false;
  h.deferred = false;
  h.alpha = 0.0f;
  h.a = h.ad = h.b = h.bd = 0.0f;
  h.rgb = 0;
  h.ray = h.last_sdf = h.last_alpha = 0.0f;
  h.last_w = 0;
  float cdx = ((float)x - ra.cx) / ra.fx, cdy = ((float)y - ra.cy) / ra.fy, cdz = 1.0f;
  const float cl = sqrtf(cdx * cdx + cdy * cdy + cdz * cdz);
  if (cl > 1e-6f) {
    cdx /= cl;
    cdy /= cl;
    cdz /= cl;
  }
  float M[12];
  render_pose<DEV>(ra, M);
  float wdx = M[0] * cdx + M[1] * cdy + M[2] * cdz;
  float wdy = M[4] * cdx + M[5] * cdy + M[6] * cdz;
  float wdz = M[8] * cdx + M[9] * cdy + M[10] * cdz;
  const float wl = sqrtf(wdx * wdx + wdy * wdy + wdz * wdz);
  if (wl > 1e-6f) {
    wdx /= wl;
    wdy /= wl;
    wdz /= wl;
  }
  h.cdx = cdx;
  h.cdy = cdy;
  h.cdz = cdz;
  h.wdx = wdx;
  h.wdy = wdy;
  h.wdz = wdz;
  h.d2r = 1.0f / cdz;
  const float minD = ra.minD, maxD = ra.maxD;
  if (minD == 0.0f || minD == MINF || maxD == 0.0f || maxD == MINF) return false;
  h.marched = true;
  rs.v = ra.voxel;
  float cam[3];
  render_cam<DEV>(ra, cam);
  rs.camx = cam[0];
  rs.camy = cam[1];
  rs.camz = cam[2];
  rs.ray0 = h.d2r * fmaxf(minD, minD);
  rs.ray_end = h.d2r * fminf(maxD, maxD);
  rs.th = rs.v * 3.0f;
  rs.inc = rs.v * 0.5f;
  RaySkip& sk = rs.sk;
  sk.camx = rs.camx;
  sk.camy = rs.camy;
  sk.camz = rs.camz;
  sk.wdx = wdx;
  sk.wdy = wdy;
  sk.wdz = wdz;
  sk.v = rs.v;
  sk.inv_v = 1.0f / rs.v;
  const float mag = fmaxf(fmaxf(fabsf(rs.camx), fabsf(rs.camy)), fabsf(rs.camz)) + fabsf(rs.ray_end);
  sk.eps = 0.02f + 8e-6f * mag * sk.inv_v;
  if (ra.win_obs) {  // projective window of the tracking renders
    const float o = ra.win_obs[y * (uint32_t)ra.W + x];
    if (!(o > 0.0f)) {
      h.marched = false;
      return false;
    }
    rs.ray0 = h.d2r * fmaxf(minD, o - ra.win);
    rs.ray_end = h.d2r * fminf(maxD, o + ra.win);
  }
  return true;
}

// Skip test at ray parameter `ray`: 0 when the super-blocks around the sample hold
// no allocated block, 1 when the blocks around it hold none, 2 otherwise; [t0, t1)
// is the parameter interval of the largest box known to hold no allocated block (for
// 0: the hyper-block, an aligned 2x2x2 group of super-blocks or the super-blocks; for
// 1: an aligned 2x2x2 group of blocks or the blocks; the blocks' box for 2).
__device__ __forceinline__ int skip_test(const TableView& t, OccCache& oc, const RaySkip& sk, float wdx, float wdy,
                                         float wdz, float ray, float* t0, float* t1) {
  const float mq = 1.125f + sk.eps;  // range margin around Q (voxels)
  const float qx = (sk.camx + ray * wdx) * sk.inv_v, qy = (sk.camy + ray * wdy) * sk.inv_v,
              qz = (sk.camz + ray * wdz) * sk.inv_v;
  const int32_t bx0 = cvt_i32(floorf((qx - mq) * 0.125f)), bx1 = cvt_i32(floorf((qx + mq) * 0.125f));
  const int32_t by0 = cvt_i32(floorf((qy - mq) * 0.125f)), by1 = cvt_i32(floorf((qy + mq) * 0.125f));
  const int32_t bz0 = cvt_i32(floorf((qz - mq) * 0.125f)), bz1 = cvt_i32(floorf((qz + mq) * 0.125f));
  *t0 = *t1 = 0.0f;
  // consecutive samples mostly share their block range: an allocated block in it stays
  const uint32_t ext = 8u | (uint32_t)(bx1 - bx0) | ((uint32_t)(by1 - by0) << 1) | ((uint32_t)(bz1 - bz0) << 2);
  if (ext == oc.rs && bx0 == oc.rx && by0 == oc.ry && bz0 == oc.rz) return 2;
  const int occ = occ_test(t, oc, bx0, bx1, by0, by1, bz0, bz1);
  oc.rx = bx0;
  oc.ry = by0;
  oc.rz = bz0;
  oc.rs = occ == 2 ? ext : 0u;
  if (occ == 2) return 2;
  // box levels in blocks: 1 (the range's blocks), 2 (an aligned 2x2x2 group of blocks
  // with no allocated block, read from the super-block's word), 4 (super-blocks), 8 (an
  // aligned 2x2x2 group of empty super-blocks, read from the hyper-block's word), 16
  // (an empty hyper-block).  Bits x + 4y + 16z of a word: a group at local origin
  // (gx, gy, gz) in {0, 2}^3 is 0x330033 << (gx + 4 gy + 16 gz).
  int32_t sh = occ == 0 ? 2 : 0;
  if (occ == 1) {
    if ((bx0 >> 1) == (bx1 >> 1) && (by0 >> 1) == (by1 >> 1) && (bz0 >> 1) == (bz1 >> 1)) {
      const uint64_t w = occ_word(t, oc, bx0 >> 2, by0 >> 2, bz0 >> 2);
      const uint32_t gsh = (uint32_t)(bx0 & 2) + 4u * (uint32_t)(by0 & 2) + 16u * (uint32_t)(bz0 & 2);
      if ((w & (0x0000000000330033ull << gsh)) == 0ull) sh = 1;
    }
  } else if ((bx0 >> 4) == (bx1 >> 4) && (by0 >> 4) == (by1 >> 4) && (bz0 >> 4) == (bz1 >> 4)) {
    const uint32_t hi_ = hyp_index(bx0 >> 4, by0 >> 4, bz0 >> 4, t.hyp_mask);
    if (hi_ != oc.hidx) {
      oc.hidx = hi_;
      oc.hword = t.hyp[hi_];
    }
    if (oc.hword == 0ull) {
      sh = 4;
    } else if ((bx0 >> 3) == (bx1 >> 3) && (by0 >> 3) == (by1 >> 3) && (bz0 >> 3) == (bz1 >> 3)) {
      const uint32_t gsh = (uint32_t)((bx0 >> 2) & 2) + 4u * (uint32_t)((by0 >> 2) & 2) + 16u * (uint32_t)((bz0 >> 2) & 2);
      if ((oc.hword & (0x0000000000330033ull << gsh)) == 0ull) sh = 3;
    }
  }
  const float g = (float)(8 << sh);
  const float lo = 1.0f, hi = g - 1.0f;
  box_interval(sk, g * (float)(bx0 >> sh) + lo, g * (float)(bx1 >> sh) + hi, g * (float)(by0 >> sh) + lo,
               g * (float)(by1 >> sh) + hi, g * (float)(bz0 >> sh) + lo, g * (float)(bz1 >> sh) + hi, t0, t1);
  return occ;
}

// Bits lo..hi (0 <= lo <= hi <= 3) of a 4-bit row, and the same rows spread to the
// y rows (bits 4y) and z planes (bits 16z) of a super-block's 64-bit occupancy word.
__device__ __forceinline__ uint32_t occ_row(uint32_t lo, uint32_t hi) { return (2u << hi) - (1u << lo); }
__device__ __forceinline__ uint32_t occ_rows_y(uint32_t lo, uint32_t hi) {
  return (0x1111u >> (4u * (3u - hi))) & (0x1111u << (4u * lo));
}
__device__ __forceinline__ uint64_t occ_planes_z(uint32_t lo, uint32_t hi) {
  return (0x0001000100010001ull >> (16u * (3u - hi))) & (0x0001000100010001ull << (16u * lo));
}

// Block range of the samples of a ray with parameter in [ta, tb]: no such sample
// takes a corner voxel from outside blocks [b0, b1] per axis.  Sample positions are
// cam + r * wd in float, monotone in r per axis, so every sample of the interval lies
// between the positions of its ends (computed by the same float operations); its
// corners lie within 1.125 + eps voxels of its position (as in skip_test).  The
// interval must span at most 29 voxels per axis: then a range covers at most 5
// blocks, hence at most two super-blocks per axis (`fits` is false otherwise, and
// for non-finite positions).
struct SegRange {
  int32_t b0[3], b1[3];
  bool fits;
};

__device__ __forceinline__ void segment_range(const RaySkip& sk, float ta, float tb, SegRange& sr) {
  const float mq = 1.125f + sk.eps;
  const float c[3] = {sk.camx, sk.camy, sk.camz}, d[3] = {sk.wdx, sk.wdy, sk.wdz};
  sr.fits = true;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const float qa = (c[a] + ta * d[a]) * sk.inv_v, qb = (c[a] + tb * d[a]) * sk.inv_v;
    sr.b0[a] = cvt_i32(floorf((fminf(qa, qb) - mq) * 0.125f));
    sr.b1[a] = cvt_i32(floorf((fmaxf(qa, qb) + mq) * 0.125f));
    sr.fits = sr.fits && __builtin_isfinite(qa) && __builtin_isfinite(qb) &&
              (uint32_t)sr.b1[a] - (uint32_t)sr.b0[a] <= 4u;
  }
}

// Super-block k (0..7) of a range: s0 + (k >> a & 1) per axis, its share [lo, hi] of
// the block range in local coordinates; false when the range stays in s0 on an axis
// where k takes the second super-block.
__device__ __forceinline__ bool segment_part(const SegRange& sr, int k, int32_t s[3], uint32_t lo[3],
                                             uint32_t hi[3]) {
  bool in = true;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const int32_t s0 = sr.b0[a] >> 2, s1 = sr.b1[a] >> 2;
    const bool second = (k >> a) & 1;
    s[a] = s0 + (second ? 1 : 0);
    in = in && (!second || s1 != s0);
    lo[a] = second ? 0u : ((uint32_t)sr.b0[a] & 3u);
    hi[a] = (second || s1 == s0) ? ((uint32_t)sr.b1[a] & 3u) : 3u;
  }
  return in;
}

// Reads the occupancy words of the (up to 8) super-blocks of a range.
__device__ __forceinline__ void segment_words(const TableView& t, const SegRange& sr, uint64_t w[8]) {
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    int32_t s[3];
    uint32_t lo[3], hi[3];
    (void)segment_part(sr, k, s, lo, hi);
    w[k] = t.occ[occ_index(s[0], s[1], s[2], t.occ_mask)];
  }
}

// True when the occupancy words w of a fitting range show no allocated block in it.
__device__ __forceinline__ bool segment_clear(const SegRange& sr, const uint64_t w[8]) {
  uint64_t hit = 0ull;
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    int32_t s[3];
    uint32_t lo[3], hi[3];
    const bool in = segment_part(sr, k, s, lo, hi);
    const uint32_t mxy = occ_row(lo[0], hi[0]) * occ_rows_y(lo[1], hi[1]);
    const uint64_t m = in ? (uint64_t)mxy * occ_planes_z(lo[2], hi[2]) : 0ull;
    hit |= w[k] & m;
  }
  return sr.fits && hit == 0ull;
}

// Positive-cell test inside allocated blocks.  The sample's corner voxels lie in the
// 4^3-voxel cells c0..c1 per axis (Q in [4*c0 + 1, 4*c1 + 3) keeps them there, as for
// blocks).  A sample can be the upper end of an accepted crossing only if its
// distance is positive, which needs a corner voxel with weight > 0 and SDF > 0 (the
// distance is a positively weighted mean of the weighted corners' SDFs).  Returns
// true with [t0, t1), the parameter interval of the cells' box, when no cell holds
// such a voxel: no sample in the box can complete a crossing, and only the box's
// last sample matters for what follows (it is the lower end of the next bracket).
// *last: key of the last cell range found to hold a positive voxel (consecutive
// samples mostly share it; a key collision only costs a skipped test).
// The cells around the sample at ray parameter `ray` (as in cell_skip) and their key.
struct CellRange {
  int32_t cx0, cx1, cy0, cy1, cz0, cz1;
  uint32_t key;
};

__device__ __forceinline__ void cell_range(const RaySkip& sk, float wdx, float wdy, float wdz, float ray,
                                           CellRange& cr) {
  const float mq = 1.125f + sk.eps;
  const float qx = (sk.camx + ray * wdx) * sk.inv_v, qy = (sk.camy + ray * wdy) * sk.inv_v,
              qz = (sk.camz + ray * wdz) * sk.inv_v;
  cr.cx0 = cvt_i32(floorf((qx - mq) * 0.25f));
  cr.cx1 = cvt_i32(floorf((qx + mq) * 0.25f));
  cr.cy0 = cvt_i32(floorf((qy - mq) * 0.25f));
  cr.cy1 = cvt_i32(floorf((qy + mq) * 0.25f));
  cr.cz0 = cvt_i32(floorf((qz - mq) * 0.25f));
  cr.cz1 = cvt_i32(floorf((qz + mq) * 0.25f));
  cr.key = ((uint32_t)cr.cx0 & 0x1FFu) | (((uint32_t)cr.cy0 & 0x1FFu) << 9) | (((uint32_t)cr.cz0 & 0x1FFu) << 18) |
           ((uint32_t)(cr.cx1 - cr.cx0) << 27) | ((uint32_t)(cr.cy1 - cr.cy0) << 28) |
           ((uint32_t)(cr.cz1 - cr.cz0) << 29) | 0x80000000u;
}

template <class Cache>
__device__ __forceinline__ bool cell_skip(const TableView& t, Cache& bc, const RaySkip& sk, const CellRange& cr,
                                          uint32_t* last, float* t0, float* t1) {
  const int32_t cx0 = cr.cx0, cx1 = cr.cx1, cy0 = cr.cy0, cy1 = cr.cy1, cz0 = cr.cz0, cz1 = cr.cz1;
  const uint32_t key = cr.key;
  if (key == *last) return false;
  if (!bc.cells_free(t, cx0, cx1, cy0, cy1, cz0, cz1)) {
    *last = key;
    return false;
  }
  box_interval(sk, 4.0f * (float)cx0 + 1.0f, 4.0f * (float)cx1 + 3.0f, 4.0f * (float)cy0 + 1.0f,
               4.0f * (float)cy1 + 3.0f, 4.0f * (float)cz0 + 1.0f, 4.0f * (float)cz1 + 3.0f, t0, t1);
  return true;
}

// The march of one ray per lane.  budget > 0: a lane that has taken that many steps
// (skips or evaluations) stops with h.deferred and its state, to be continued by
// the cooperative tail pass.
template <bool F16, bool DIAG, bool REF, bool DEV, class Cache>
__device__ __forceinline__ RayHit march_ray(const TableView& t, const RenderArgs& ra, const void* __restrict__ vox,
                                            const uint32_t* __restrict__ col, uint32_t x, uint32_t y, Cache& bc,
                                            uint32_t budget, float zskip = 0.0f) {
  RayHit h;
  RaySetup rs;
  if (!ray_setup<DEV>(ra, x, y, h, rs)) return h;
  const float v = rs.v;
  const float camx = rs.camx, camy = rs.camy, camz = rs.camz;
  const float wdx = h.wdx, wdy = h.wdy, wdz = h.wdz;
  float ray = rs.ray0;
  const float ray_end = rs.ray_end;
  const float th = rs.th;
  const float inc = rs.inc;
  float last_sdf = 0.0f, last_alpha = 0.0f;
  uint32_t last_w = 0;
  const RaySkip& sk = rs.sk;
  // One step per lane and iteration: a skip over samples the occupancy map proves
  // invalid, or the evaluation of one sample.  Lanes never wait for each other's
  // skips, so the wave runs max(steps) iterations rather than the sum of the lanes'
  // skip runs.  A sign change that passes the reference's thresholds parks the lane
  // with its bracket; once no lane of the wave is still marching, the parked lanes
  // refine together (three evaluations each), so the refinement runs once per wave
  // instead of once per lane.  A failed refinement resumes the march exactly where
  // the reference continues (RayCastRender.cu:372-391).
  bool stalled = false;
  uint32_t n_eval = 0, n_super = 0, n_block = 0, n_replay = 0, n_steps = 0, n_cell = 0;
  uint32_t cell_key = 0u;
  OccCache oc;
  occ_cache_init(oc);
  if (zskip > 0.0f) {
    // every sample at a camera depth below zskip is invalid (bundle_start): on to the
    // first sample at or beyond the parameter of that depth, taken slightly low
    const float tz = (zskip * h.d2r) * (1.0f - 1.0f / 1048576.0f);
    if (tz > ray) {
      const float from = ray;
      ray = replay_steps(ray, fminf(tz, ray_end), inc, &stalled);
      if (DIAG) n_replay += (uint32_t)((ray - from) / inc + 0.5f);
    }
  }
  bool marching = ray < ray_end && !stalled, parked = false;
  while (true) {
    if (!__any(marching || parked)) break;
    if (marching && budget != 0u && n_steps >= budget) {
      h.deferred = true;
      h.ray = ray;
      h.last_sdf = last_sdf;
      h.last_alpha = last_alpha;
      h.last_w = last_w;
      marching = false;
    }
    if (marching) {
      n_steps++;
      bool step = true;
      CellRange cr;
      cell_range(sk, wdx, wdy, wdz, ray, cr);
      // the cells of the last range found to hold a positive voxel (hence in allocated
      // blocks): no skip applies, the sample is evaluated (a key collision only costs an
      // evaluation the skips would have avoided)
      if (cr.key != cell_key) {
        float t0, t1;
        const int occ = skip_test(t, oc, sk, wdx, wdy, wdz, ray, &t0, &t1);
        const bool inside = t0 <= ray && ray < t1;  // false for a NaN interval
        if (occ != 2 && inside) {
          if (occ == 0) n_super++;
          else n_block++;
          last_w = 0u;
          const float from = ray;
          ray = replay_steps(ray, fminf(t1, ray_end), inc, &stalled);
          if (DIAG) n_replay += (uint32_t)((ray - from) / inc + 0.5f);
          if (stalled || !(ray < ray_end)) marching = false;
          step = false;
        } else if (occ == 2 && cell_skip(t, bc, sk, cr, &cell_key, &t0, &t1) && t0 <= ray && ray < t1) {
          n_cell++;
          if (!(t1 < ray_end)) {
            // every remaining sample lies in the box: no crossing can be accepted
            marching = false;
            step = false;
          } else {
            // on to the box's last sample, evaluated below
            float prev;
            const float from = ray;
            (void)replay_steps(ray, t1, inc, &stalled, &prev);
            if (stalled) {
              marching = false;
              step = false;
            } else {
              ray = prev;
              if (DIAG) n_replay += (uint32_t)((ray - from) / inc + 0.5f);
            }
          }
        }
      }
      if (step) {
        const float ppx = camx + ray * wdx, ppy = camy + ray * wdy, ppz = camz + ray * wdz;
        float dist;
        const int stat = trilinear<F16, false, REF>(t, bc, vox, col, ppx, ppy, ppz, v, &dist, nullptr);
        n_eval++;
        if (stat == kSampleValid) {
          // findIntersectionBisection's bracket and the acceptance thresholds
          // (RayCastRender.cu:372-385); the refinement itself comes later
          if (last_w > 0u && last_sdf < 0.0f && dist > 0.0f && fabsf(last_sdf - dist) < th && fabsf(dist) < th) {
            // the bracket's lower end; its upper end becomes last_alpha / last_sdf below
            h.a = last_alpha;
            h.ad = last_sdf;
            parked = true;
            marching = false;
          }
          last_sdf = dist;
          last_alpha = ray;
          last_w = 1u;
        } else {
          last_w = 0u;
        }
        if (marching) {
          const float next = ray + inc;
          if (next == ray) marching = false;  // step absorbed by rounding: the reference would never end
          ray = next;
          if (!(ray < ray_end)) marching = false;
        }
      }
    }
    if (!__any(marching)) {
      if (parked) {
        // findIntersectionBisection (RayCastRender.cu:229-271)
        h.b = last_alpha;
        h.bd = last_sdf;
        float a = h.a, ad = h.ad, bb = h.b, bd = h.bd, c = 0.0f;
        uint32_t rgb = 0;
        bool ok = true;
#pragma unroll
        for (int it = 0; it < 3; ++it) {
          c = a + (ad / (ad - bd)) * (bb - a);
          float cd;
          // the colour of the last point is the one returned
          const int stat = it < 2 ? trilinear<F16, false, REF>(t, bc, vox, col, camx + c * wdx, camy + c * wdy,
                                                               camz + c * wdz, v, &cd, &rgb)
                                  : trilinear<F16, true, REF>(t, bc, vox, col, camx + c * wdx, camy + c * wdy,
                                                              camz + c * wdz, v, &cd, &rgb);
          if (stat != kSampleValid) {
            ok = false;
            break;
          }
          if (ad * cd > 0.0f) {
            a = c;
            ad = cd;
          } else {
            bb = c;
            bd = cd;
          }
        }
        parked = false;
        if (ok) {
          h.found = true;
          h.alpha = c;
          h.rgb = rgb;
        } else {
          // the reference keeps marching after the rejected crossing
          const float next = h.b + inc;
          ray = next;
          marching = next != h.b && next < ray_end;
        }
      }
    }
  }
  if (DIAG && bc.stats) {
    atomicAdd(&bc.stats[0], n_eval);
    atomicAdd(&bc.stats[1], n_super);
    atomicAdd(&bc.stats[2], n_block);
    atomicAdd(&bc.stats[3], n_replay);
    atomicAdd(&bc.stats[4], n_cell);
  }
  return h;
}

// Bundle look-ahead of an 8x8 tile (one wave): lane k tests the slab of the tile's
// frustum between camera depths z + k dz and z + (k + 1) dz against the occupancy map.
// The samples of every ray of the tile at a camera depth below the first slab that can
// reach an allocated block are invalid (their corner voxels lie in no allocated block),
// so each ray may start at its first sample at or beyond that depth.  The points of
// the tile's rays at camera depth z in [za, zb] lie in the convex hull of the four
// corner rays' points at za and zb (the pixel direction (u, w, 1) is linear in the
// pixel, and the pose's rotation is orthonormal: checked to 1e-5, with a matching
// margin); the hull's bounding box with the corner margins of skip_test gives a block
// range tested like segment_clear.  Returns 0 when nothing is proven.
constexpr float kBundleSlab = 6.0f;  // slab thickness in voxels
constexpr int kBundleRounds = 4;     // look-ahead rounds (64 slabs each)

template <bool DEV>
__device__ __forceinline__ float bundle_start(const TableView& t, const RenderArgs& ra, uint32_t tx, uint32_t ty) {
  const float MINF = 1e10f;
  const float minD = ra.minD, maxD = ra.maxD;
  if (minD == 0.0f || minD == MINF || maxD == 0.0f || maxD == MINF || !(minD < maxD) || ra.win_obs) return 0.0f;
  float M[12];
  render_pose<DEV>(ra, M);
  float cam[3];
  render_cam<DEV>(ra, cam);
  // rotation orthonormal to 1e-5 (else the hull argument does not hold)
  float dev = 0.0f;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
#pragma unroll
    for (int j = i; j < 3; ++j) {
      const float d = M[i] * M[j] + M[4 + i] * M[4 + j] + M[8 + i] * M[8 + j] - (i == j ? 1.0f : 0.0f);
      dev = fmaxf(dev, fabsf(d));
    }
  }
  if (!(dev <= 1e-5f)) return 0.0f;
  const float v = ra.voxel, inv_v = 1.0f / v;
  const float u0 = ((float)(8u * tx) - ra.cx) / ra.fx, u1 = ((float)(8u * tx + 7u) - ra.cx) / ra.fx;
  const float w0 = ((float)(8u * ty) - ra.cy) / ra.fy, w1 = ((float)(8u * ty + 7u) - ra.cy) / ra.fy;
  // world direction per unit camera depth of the four corner rays
  float dir[4][3];
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    const float u = (c & 1) ? u1 : u0, w = (c & 2) ? w1 : w0;
#pragma unroll
    for (int a = 0; a < 3; ++a) dir[c][a] = M[4 * a] * u + M[4 * a + 1] * w + M[4 * a + 2];
  }
  const float dmax = fmaxf(fmaxf(fabsf(u0), fabsf(u1)), fmaxf(fabsf(w0), fabsf(w1))) + 1.0f;
  const float mag = fmaxf(fmaxf(fabsf(cam[0]), fabsf(cam[1])), fabsf(cam[2])) + 2.0f * maxD * dmax;
  const float eps = 0.02f + 4e-5f * mag * inv_v;
  const float mq = 1.125f + eps;
  const uint32_t lane = threadIdx.x & 63u;
  const float dz = kBundleSlab * v;
  float z = minD;
  for (int round = 0; round < kBundleRounds; ++round) {
    const float za = z + (float)lane * dz, zb = z + (float)(lane + 1u) * dz;
    SegRange sr;
    sr.fits = true;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      float lo = INFINITY, hi = -INFINITY;
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        const float pa = cam[a] + dir[c][a] * za, pb = cam[a] + dir[c][a] * zb;
        lo = fminf(lo, fminf(pa, pb));
        hi = fmaxf(hi, fmaxf(pa, pb));
      }
      const float qa = lo * inv_v, qb = hi * inv_v;
      sr.b0[a] = cvt_i32(floorf((qa - mq) * 0.125f));
      sr.b1[a] = cvt_i32(floorf((qb + mq) * 0.125f));
      sr.fits = sr.fits && __builtin_isfinite(qa) && __builtin_isfinite(qb) &&
                (uint32_t)sr.b1[a] - (uint32_t)sr.b0[a] <= 4u;
    }
    uint64_t ow[8];
    segment_words(t, sr, ow);
    const bool reach = za < maxD && !segment_clear(sr, ow);
    const uint64_t bm = __ballot(reach);
    const uint32_t f = bm != 0ull ? (uint32_t)__builtin_ctzll(bm) : 64u;
    if (f == 0u) break;
    z = z + (float)f * dz;
    if (!(z < maxD)) break;
  }
  return z > minD ? z : 0.0f;
}

template <bool F16, bool DIAG, bool REF, bool DEV>
__global__ __launch_bounds__(64, 5) void k_render(TableView t, const DevState* __restrict__ st, RenderArgs ra,
                                               const void* __restrict__ vox, const uint32_t* __restrict__ col,
                                               float* __restrict__ out_depth, float4* __restrict__ out_pos,
                                               float4* __restrict__ out_col, float4* __restrict__ out_nrm,
                                               uint32_t* __restrict__ rc) {
  // queue controls: this render uses set tail_par; the other set, the previous render's
  // (whose tail pass has finished), is cleared for the next render
  if (blockIdx.x == 0u && threadIdx.x < 16u) ra.tail_ctl[16u * (ra.tail_par ^ 1u) + threadIdx.x] = 0u;
  // Workgroups are dealt to the 8 XCDs in turn (workgroup b runs on XCD b % 8).  The
  // image is cut into super-tiles of kSuperTile x kSuperTile 8x8 tiles, dealt to the
  // XCDs in turn: the rays of a super-tile, which read the same blocks, share one L2,
  // and every XCD gets super-tiles from all over the image (balanced load).
  const uint32_t tiles_x = ((uint32_t)ra.W + 7u) / 8u, tiles_y = ((uint32_t)ra.H + 7u) / 8u;
  const uint32_t stx_n = (tiles_x + kSuperTile - 1u) / kSuperTile;
  const uint32_t j = blockIdx.x >> 3, within = j % (kSuperTile * kSuperTile);
  const uint32_t st_idx = (j / (kSuperTile * kSuperTile)) * 8u + (blockIdx.x & 7u);
  const uint32_t tx = (st_idx % stx_n) * kSuperTile + within % kSuperTile;
  const uint32_t ty = (st_idx / stx_n) * kSuperTile + within / kSuperTile;
  if (tx >= tiles_x || ty >= tiles_y) return;
  const uint32_t tile = ty * tiles_x + tx;
  const size_t wg = tile;
  if (st->n_alloc == 0u) {  // RayCastRender.cu:553-556: nothing to render, outputs untouched
    if (threadIdx.x == 0) {
      rc[2 * wg] = 0u;
      rc[2 * wg + 1] = 0u;
    }
    return;
  }
  // the whole wave: the tile's rays skip what the bundle look-ahead proves invalid
  const float zskip = bundle_start<DEV>(t, ra, tx, ty);
  const uint32_t x = tx * 8u + (threadIdx.x & 7u);
  const uint32_t y = ty * 8u + (threadIdx.x >> 3);
  uint32_t hit = 0, miss = 0;
  bool deferred = false;
  TailRay tr;
  if (x < (uint32_t)ra.W && y < (uint32_t)ra.H) {
    const float MINF = 1e10f;
    const int pix = (int)y * ra.W + (int)x;
    __shared__ int32_t slot_lds[8 * 64];
    BlockCache<DIAG> bc;
    bc.init(slot_lds + threadIdx.x, ra.mark_bkt, ra.mark_blk, ra.march_stats);
    const RayHit h = march_ray<F16, DIAG, REF, DEV>(t, ra, vox, col, x, y, bc, ra.budget, zskip);
    deferred = h.deferred;
    tr.pix = (uint32_t)pix;
    tr.ray = h.ray;
    tr.last_sdf = h.last_sdf;
    tr.last_alpha = h.last_alpha;
    tr.last_w = h.last_w;
    tr.pad[0] = tr.pad[1] = tr.pad[2] = 0u;
    if (!deferred) {
    float o_depth = MINF;
    float4 o_pos = make_float4(MINF, MINF, MINF, MINF), o_col = o_pos;
    if (h.found) {
      const float alpha = h.alpha;
      o_depth = alpha / h.d2r;
      float cam[3];
      render_cam<DEV>(ra, cam);
      o_pos = make_float4(cam[0] + alpha * h.wdx, cam[1] + alpha * h.wdy, cam[2] + alpha * h.wdz, 1.0f);
      o_col = make_float4((float)(h.rgb & 0xFFu) / 255.0f, (float)((h.rgb >> 8) & 0xFFu) / 255.0f,
                          (float)((h.rgb >> 16) & 0xFFu) / 255.0f, 1.0f);
    }
    hit = (h.marched && h.found) ? 1u : 0u;
    miss = (h.marched && !h.found) ? 1u : 0u;
    out_depth[pix] = o_depth;
    out_pos[pix] = o_pos;
    out_col[pix] = o_col;
    out_nrm[pix] = make_float4(MINF, MINF, MINF, MINF);
    }
  }
  // deferred rays: one queue reservation per wave
  const uint64_t dm = __ballot(deferred);
  if (dm != 0ull) {
    uint32_t base = 0;
    if (DIAG && threadIdx.x == 0 && ra.march_stats) atomicAdd(&ra.march_stats[5], (uint32_t)__popcll(dm));
    // the segment of this workgroup's XCD (the tail pass's workgroups on the same XCD
    // take them, where the march left the blocks in L2)
    const uint32_t xcd = blockIdx.x & 7u;
    if (threadIdx.x == 0) base = atomicAdd(&ra.tail_ctl[16u * ra.tail_par + xcd], (uint32_t)__popcll(dm));
    base = __shfl(base, 0, 64);
    if (deferred) {
      const uint32_t rank = (uint32_t)__popcll(dm & ((1ull << threadIdx.x) - 1ull));
      if (base + rank < ra.tail_seg)
        reinterpret_cast<TailRay*>(ra.tail)[(size_t)xcd * ra.tail_seg + base + rank] = tr;
    }
  }
  // one wave per workgroup: one write of its hit / miss counts (summed when read)
  hit = wave_sum(hit);
  miss = wave_sum(miss);
  if (threadIdx.x == 0) {
    rc[2 * wg] = hit;
    rc[2 * wg + 1] = miss;
  }
}

// Tail pass: the rays whose march ran out of steps, one ray per group of kTailGroup
// lanes.  A group looks ahead along the ray against the occupancy map and skips what
// it proves invalid; otherwise its lanes evaluate 2 kTailGroup consecutive samples at
// once (the positions the reference's additions reach) and the group resolves them in
// order: the first sign change that passes the thresholds is refined, and the march
// continues after the window's last sample.  Group g starts with queue entry g;
// afterwards the groups of a wave that need a ray take them with one atomic per wave
// (a single queue counter would otherwise serialise the fetches), until the queue is
// empty; then a wave's last rays move to wider groups (tail_drain).  The queue
// controls alternate between two sets by render; the march clears the one it does
// not use (no workgroup of the tail pass needs to know it is the last).
constexpr int kTailGroup = 8;
// Widest group of a wave's last rays (tail_drain).
constexpr int kTailDrainGroup = 32;
// Look-ahead segment of the tail pass, in voxels along the ray (at most 29: see
// segment_range).
constexpr float kLookSeg = 16.0f;

// March state of a ray in the tail pass, one per group of lanes (the same in every lane
// of the group).
struct TailState {
  RayHit h;
  RaySetup rs;
  uint32_t pix;
  float ray, last_sdf, last_alpha;
  uint32_t last_w;
};

// Diagnostic counts of the tail pass (lane gl == 0 of each group counts).
struct TailStats {
  uint32_t n_eval = 0, n_super = 0, n_block = 0, n_replay = 0, n_window = 0, n_tskip = 0;
};

// Bit i: predicate p of lane gbase + i of this lane's group of G lanes.
template <int G>
__device__ __forceinline__ uint64_t group_bits(bool p, int gbase) {
  const uint64_t b = __ballot(p);
  if constexpr (G == 64) {
    return b;
  } else {
    return (b >> gbase) & ((1ull << G) - 1ull);
  }
}

// One iteration of the tail pass on the ray of a group of G lanes: a skip over samples
// the occupancy map proves invalid (the look-ahead over G segments of kLookSeg voxels),
// or a window of 2 G consecutive samples (lane gl takes samples 2 gl and 2 gl + 1)
// resolved in order; the first sign change that passes the thresholds is refined.
// Returns true when the ray is finished (s.h.found: whether it hit).  Every lane of the
// group takes part.
template <bool F16, bool DIAG, bool REF, int G, class Cache>
__device__ __forceinline__ bool tail_iteration(const TableView& t, const RenderArgs& ra, const void* __restrict__ vox,
                                               const uint32_t* __restrict__ col, Cache& bc, TailState& s, int gl,
                                               int gbase, TailStats& ts) {
  RayHit& h = s.h;
  const RaySetup& rs = s.rs;
  float& ray = s.ray;
  float& last_sdf = s.last_sdf;
  float& last_alpha = s.last_alpha;
  uint32_t& last_w = s.last_w;
  bool done = false;
  // samples 2 gl and 2 gl + 1 of the window (2 G consecutive samples): 2 gl
  // and 2 gl + 1 more additions of the step (an addition that leaves the value
  // unchanged ends the ray there)
  bool stl0;
  const float r0 = advance_steps(ray, rs.inc, 2u * (uint32_t)gl, &stl0);
  bool in0 = !stl0;
  const float r1 = r0 + rs.inc;
  bool in1 = in0 && r1 != r0 && r1 < rs.ray_end;
  in0 = in0 && r0 < rs.ray_end;
  // Look-ahead: lane gl tests segment [T_gl, T_gl+1] of the ray, T_k = ray + k * S.
  // Its occupancy words are read together with the block lookups of the lane's
  // window samples (one memory round trip for both); the voxels are read only when
  // the window is evaluated.
  const float S = kLookSeg * rs.v;
  const float ta = ray + (float)gl * S, tb = ray + (float)(gl + 1) * S;
  SegRange sr;
  segment_range(rs.sk, ta, tb, sr);
  uint64_t ow[8];
  segment_words(t, sr, ow);
  Corners c0, c1;
  if (in0) corner_prep<REF>(t, bc, rs.camx + r0 * h.wdx, rs.camy + r0 * h.wdy, rs.camz + r0 * h.wdz, rs.v, c0);
  c1 = c0;
  if (in1) corner_prep<REF>(t, bc, rs.camx + r1 * h.wdx, rs.camy + r1 * h.wdy, rs.camz + r1 * h.wdz, rs.v, c1);
  {
    // the samples before the first segment that can reach an allocated block are
    // invalid: the ray moves on to its first sample at or beyond that segment
    const bool reach = ta < rs.ray_end && !segment_clear(sr, ow);
    const uint64_t gm = group_bits<G>(reach, gbase);
    const int f = gm != 0ull ? __builtin_ctzll(gm) : G;
    if (f > 0) {
      if (gl == 0) ts.n_tskip++;
      last_w = 0u;
      bool stl;
      const float from = ray;
      ray = replay_steps(ray, fminf(ray + (float)f * S, rs.ray_end), rs.inc, &stl);
      if (DIAG && gl == 0) ts.n_replay += (uint32_t)((ray - from) / rs.inc + 0.5f);
      return stl || !(ray < rs.ray_end);
    }
  }
  {
    if (gl == 0) ts.n_window++;
    float d0 = 0.0f, d1 = 0.0f;
    int st0 = kSampleInvalid, st1 = kSampleInvalid;
    if (in0) {
      corner_gather2<F16>(vox, c0, c1, &st0, &d0, &st1, &d1);
      ts.n_eval += in1 ? 2u : 1u;
    }
    const bool v0 = in0 && st0 == kSampleValid, v1 = in1 && st1 == kSampleValid;
    // the sample before sample 2 gl: lane gl - 1's second, or the carried state for the
    // first lane; before sample 2 gl + 1: the lane's first
    float p_sdf = __shfl_up(d1, 1, G), p_r = __shfl_up(r1, 1, G);
    int p_v = __shfl_up(v1 ? 1 : 0, 1, G);
    if (gl == 0) {
      p_sdf = last_sdf;
      p_r = last_alpha;
      p_v = last_w ? 1 : 0;
    }
    const bool cand0 = v0 && p_v && p_sdf < 0.0f && d0 > 0.0f && fabsf(p_sdf - d0) < rs.th && fabsf(d0) < rs.th;
    const bool cand1 = v1 && v0 && d0 < 0.0f && d1 > 0.0f && fabsf(d0 - d1) < rs.th && fabsf(d1) < rs.th;
    const uint64_t g0 = group_bits<G>(cand0, gbase), g1 = group_bits<G>(cand1, gbase);
    const uint64_t gi0 = group_bits<G>(in0, gbase), gi1 = group_bits<G>(in1, gbase);
    if ((g0 | g1) != 0ull) {
      // the first candidate in sample order: lane f's first sample, else its second
      const int f = __builtin_ctzll(g0 | g1);
      const bool first = (g0 >> f) & 1ull;
      const float a0 = first ? __shfl(p_r, f, G) : __shfl(r0, f, G);
      const float ad0 = first ? __shfl(p_sdf, f, G) : __shfl(d0, f, G);
      const float b0 = first ? __shfl(r0, f, G) : __shfl(r1, f, G);
      const float bd0 = first ? __shfl(d0, f, G) : __shfl(d1, f, G);
      // findIntersectionBisection (RayCastRender.cu:229-271), the same in every lane of the group
      float a = a0, ad = ad0, bb = b0, bd = bd0, c = 0.0f;
      uint32_t rgb = 0;
      bool ok = true;
#pragma unroll
      for (int it = 0; it < 3; ++it) {
        c = a + (ad / (ad - bd)) * (bb - a);
        float cd;
        const float qx = rs.camx + c * h.wdx, qy = rs.camy + c * h.wdy, qz = rs.camz + c * h.wdz;
        const int stat = it < 2 ? trilinear<F16, false, REF>(t, bc, vox, col, qx, qy, qz, rs.v, &cd, &rgb)
                                : trilinear<F16, true, REF>(t, bc, vox, col, qx, qy, qz, rs.v, &cd, &rgb);
        if (stat != kSampleValid) {
          ok = false;
          break;
        }
        if (ad * cd > 0.0f) {
          a = c;
          ad = cd;
        } else {
          bb = c;
          bd = cd;
        }
      }
      if (ok) {
        h.found = true;
        h.alpha = c;
        h.rgb = rgb;
        h.a = a0;
        h.ad = ad0;
        h.b = b0;
        h.bd = bd0;
        done = true;
      } else {
        last_sdf = bd0;
        last_alpha = b0;
        last_w = 1u;
        const float nx = b0 + rs.inc;
        if (nx == b0 || !(nx < rs.ray_end)) done = true;
        ray = nx;
      }
    } else {
      // no accepted crossing: continue after the window's last sample (lane L's second
      // when it has one, else its first)
      const int L = 63 - __builtin_clzll(gi0 | 1ull);
      const bool second = (gi1 >> L) & 1ull;
      const float rL = second ? __shfl(r1, L, G) : __shfl(r0, L, G);
      const float dL = second ? __shfl(d1, L, G) : __shfl(d0, L, G);
      const int vL = second ? __shfl(v1 ? 1 : 0, L, G) : __shfl(v0 ? 1 : 0, L, G);
      if (vL) {
        last_sdf = dL;
        last_alpha = rL;
        last_w = 1u;
      } else {
        last_w = 0u;
      }
      const float nx = rL + rs.inc;
      if (L < G - 1 || !second || nx == rL || !(nx < rs.ray_end)) done = true;
      ray = nx;
    }
  }
  return done;
}

// Outputs of a finished tail ray (called by lane gl == 0 of its group); counts the hit or
// miss.
template <bool DEV>
__device__ __forceinline__ void tail_write(const RenderArgs& ra, const TailState& s, float* __restrict__ out_depth,
                                           float4* __restrict__ out_pos, float4* __restrict__ out_col,
                                           float4* __restrict__ out_nrm, uint32_t& hit, uint32_t& miss) {
  const float MINF = 1e10f;
  const RayHit& h = s.h;
  float o_depth = MINF;
  float4 o_pos = make_float4(MINF, MINF, MINF, MINF), o_col = o_pos;
  if (h.found) {
    const float alpha = h.alpha;
    o_depth = alpha / h.d2r;
    float cam[3];
    render_cam<DEV>(ra, cam);
    o_pos = make_float4(cam[0] + alpha * h.wdx, cam[1] + alpha * h.wdy, cam[2] + alpha * h.wdz, 1.0f);
    o_col = make_float4((float)(h.rgb & 0xFFu) / 255.0f, (float)((h.rgb >> 8) & 0xFFu) / 255.0f,
                        (float)((h.rgb >> 16) & 0xFFu) / 255.0f, 1.0f);
    hit++;
  } else {
    miss++;
  }
  out_depth[s.pix] = o_depth;
  out_pos[s.pix] = o_pos;
  out_col[s.pix] = o_col;
  out_nrm[s.pix] = make_float4(MINF, MINF, MINF, MINF);
}

// The end of a wave's tail work, when no group can take another ray: the rays still
// running (one per group of G lanes, `have` in its lanes) are finished; once they fit
// in groups of 2 G lanes (up to kTailDrainGroup), they move to such groups.
template <bool F16, bool DIAG, bool REF, bool DEV, int G, class Cache>
__device__ __forceinline__ void tail_regroup(const TableView& t, const RenderArgs& ra, const void* __restrict__ vox,
                                             const uint32_t* __restrict__ col, Cache& bc, const TailState& s_in,
                                             uint64_t act, TailStats& ts, float* __restrict__ out_depth,
                                             float4* __restrict__ out_pos, float4* __restrict__ out_col,
                                             float4* __restrict__ out_nrm, uint32_t& hit, uint32_t& miss);

template <bool F16, bool DIAG, bool REF, bool DEV, int G, class Cache>
__device__ __forceinline__ void tail_drain(const TableView& t, const RenderArgs& ra, const void* __restrict__ vox,
                                           const uint32_t* __restrict__ col, Cache& bc, TailState& s, bool have,
                                           TailStats& ts, float* __restrict__ out_depth, float4* __restrict__ out_pos,
                                           float4* __restrict__ out_col, float4* __restrict__ out_nrm, uint32_t& hit,
                                           uint32_t& miss) {
  const int lane = threadIdx.x;
  const int gl = lane & (G - 1), gbase = lane & ~(G - 1);
  while (true) {
    const uint64_t act = __ballot(have && gl == 0);
    if (act == 0ull) return;
    if constexpr (2 * G <= kTailDrainGroup) {
      if (__popcll(act) <= 64 / (2 * G)) {
        tail_regroup<F16, DIAG, REF, DEV, 2 * G>(t, ra, vox, col, bc, s, act, ts, out_depth, out_pos, out_col, out_nrm,
                                                  hit, miss);
        return;
      }
    }
    if (have && tail_iteration<F16, DIAG, REF, G>(t, ra, vox, col, bc, s, gl, gbase, ts)) {
      if (gl == 0) tail_write<DEV>(ra, s, out_depth, out_pos, out_col, out_nrm, hit, miss);
      have = false;
    }
  }
}

// Moves the rays of the groups in `act` (bit l: lane l leads a group with a ray) to
// groups of G lanes, the k-th ray to lanes [k G, (k + 1) G), and finishes them.
template <bool F16, bool DIAG, bool REF, bool DEV, int G, class Cache>
__device__ __forceinline__ void tail_regroup(const TableView& t, const RenderArgs& ra, const void* __restrict__ vox,
                                             const uint32_t* __restrict__ col, Cache& bc, const TailState& s_in,
                                             uint64_t act, TailStats& ts, float* __restrict__ out_depth,
                                             float4* __restrict__ out_pos, float4* __restrict__ out_col,
                                             float4* __restrict__ out_nrm, uint32_t& hit, uint32_t& miss) {
  const int lane = threadIdx.x;
  uint64_t m = act;
  for (int i = 0; i < lane / G; ++i) m &= m - 1ull;
  const bool have = m != 0ull;
  const int src = have ? __builtin_ctzll(m) : 0;
  TailState s;
  s.pix = __shfl(s_in.pix, src, 64);
  s.ray = __shfl(s_in.ray, src, 64);
  s.last_sdf = __shfl(s_in.last_sdf, src, 64);
  s.last_alpha = __shfl(s_in.last_alpha, src, 64);
  s.last_w = __shfl(s_in.last_w, src, 64);
  if (have) (void)ray_setup<DEV>(ra, s.pix % (uint32_t)ra.W, s.pix / (uint32_t)ra.W, s.h, s.rs);
  tail_drain<F16, DIAG, REF, DEV, G>(t, ra, vox, col, bc, s, have, ts, out_depth, out_pos, out_col, out_nrm, hit,
                                     miss);
}

template <bool F16, bool DIAG, bool REF, bool DEV>
__global__ __launch_bounds__(64, 3) void k_render_tail(TableView t, RenderArgs ra, const void* __restrict__ vox,
                                                    const uint32_t* __restrict__ col, float* __restrict__ out_depth,
                                                    float4* __restrict__ out_pos, float4* __restrict__ out_col,
                                                    float4* __restrict__ out_nrm, uint32_t* __restrict__ rc) {
  constexpr int G = kTailGroup;
  const int lane = threadIdx.x;
  const int gl = lane & (G - 1);
  const int gbase = lane & ~(G - 1);
  // workgroup b runs on XCD b % 8 and takes segment b % 8 of the queue (the rays the
  // march deferred on that XCD)
  const uint32_t xcd = blockIdx.x & 7u;
  uint32_t* const ctl = ra.tail_ctl + 16u * ra.tail_par;
  const uint32_t n = min(ctl[xcd], ra.tail_seg);
  const TailRay* __restrict__ queue = reinterpret_cast<const TailRay*>(ra.tail) + (size_t)xcd * ra.tail_seg;
  constexpr uint32_t kGroups = 64u / (uint32_t)G;         // groups per wave
  const uint32_t n_static = (gridDim.x >> 3) * kGroups;   // entries handed out without atomics
  uint32_t next_static = (blockIdx.x >> 3) * kGroups + (uint32_t)(gbase / G);
  uint32_t hit = 0, miss = 0;
  TailStats ts;
  __shared__ int32_t slot_lds[8 * 64];
  BlockCache<DIAG> bc;
  bc.init(slot_lds + threadIdx.x, ra.mark_bkt, ra.mark_blk, ra.march_stats);
  TailState s;
  s.pix = 0;
  s.ray = s.last_sdf = s.last_alpha = 0.0f;
  s.last_w = 0;
  bool have = false, exhausted = false;
  while (true) {
    // queue entries for the groups that need one: the first statically, then one
    // reservation per wave
    const bool want = !have && !exhausted;
    uint32_t e = 0xFFFFFFFFu;
    if (next_static != 0xFFFFFFFFu) {
      if (want) {
        e = next_static;
        next_static = 0xFFFFFFFFu;
      }
    } else {
      const uint64_t need = __ballot(want && gl == 0);
      if (need != 0ull) {
        uint32_t base = 0;
        if (lane == 0) base = atomicAdd(&ctl[8 + xcd], (uint32_t)__popcll(need));
        base = __shfl(base, 0, 64);
        if (want) e = n_static + base + (uint32_t)__popcll(need & ((1ull << gbase) - 1ull));
      }
    }
    if (want) {
      if (e < n) {
        const TailRay q = queue[e];
        s.pix = q.pix;
        (void)ray_setup<DEV>(ra, s.pix % (uint32_t)ra.W, s.pix / (uint32_t)ra.W, s.h, s.rs);
        s.ray = q.ray;
        s.last_sdf = q.last_sdf;
        s.last_alpha = q.last_alpha;
        s.last_w = q.last_w;
        have = true;
      } else {
        exhausted = true;
      }
    }
    if (!__any(have)) break;
    // After the fetch every group holds a ray or has found the queue empty (for good:
    // entries are only taken).  Once the rays left fit in groups twice as wide, the
    // wave regroups (tail_drain): a wave's last, long rays get more lanes instead of
    // running many short windows while its other lanes idle (groups of 16 lanes once
    // at most 4 rays are left, of 32 once at most 2 are: 0.495 ms against 0.497 with 16
    // lanes at most, 0.496 with 64 and 0.533 without the drain).
    const uint64_t act = __ballot(have && gl == 0);
    if (__popcll(act) <= 64 / (2 * G)) {
      tail_regroup<F16, DIAG, REF, DEV, 2 * G>(t, ra, vox, col, bc, s, act, ts, out_depth, out_pos, out_col, out_nrm,
                                                hit, miss);
      break;
    }
    if (!have) continue;
    if (tail_iteration<F16, DIAG, REF, G>(t, ra, vox, col, bc, s, gl, gbase, ts)) {
      if (gl == 0) tail_write<DEV>(ra, s, out_depth, out_pos, out_col, out_nrm, hit, miss);
      have = false;
    }
  }
  if (DIAG && ra.march_stats) {
    const uint32_t e = wave_sum(ts.n_eval), su = wave_sum(ts.n_super), bl = wave_sum(ts.n_block),
                   re = wave_sum(ts.n_replay);
    const uint32_t wi = wave_sum(ts.n_window), tsk = wave_sum(ts.n_tskip);
    if (lane == 0) {
      atomicAdd(&ra.march_stats[0], e);
      atomicAdd(&ra.march_stats[1], su);
      atomicAdd(&ra.march_stats[2], bl);
      atomicAdd(&ra.march_stats[3], re);
      atomicAdd(&ra.march_stats[6], wi);
      atomicAdd(&ra.march_stats[7], tsk);
    }
  }
  hit = wave_sum(hit);
  miss = wave_sum(miss);
  if (lane == 0) {
    rc[2 * blockIdx.x] = hit;
    rc[2 * blockIdx.x + 1] = miss;
  }
}

// ---------------------------------------------------------------------------------
// Raycast adjoint (no counterpart in the reference)
// ---------------------------------------------------------------------------------
//
// Rendered depth = alpha * cd.z with alpha the third regula-falsi point.  Recomputes
// the march (the same code as k_render, so the same bracket), then differentiates the
// computation the forward performed (fixed bracket choices and trilinear corner sets)
// in forward mode with 10 components: a camera-frame SE(3) perturbation
// T(xi) = T exp(xi^), xi = (omega, upsilon), under which a point at ray parameter r
// moves by R (upsilon + omega x r cd) + wd dr, and one seed per trilinear evaluation
// alpha depends on (the two bracket ends and the first two refinement points), whose
// product with that evaluation's d f / d sdf_j = w_j / sum(w) gives d alpha / d sdf_j.

struct TriGradF {
  float g[3];        // gradient of the sampled field in world coordinates
  int64_t vox[8];    // corner voxel indices (ptr*512 + voxel), -1 when not used
  float dfds[8];     // d f / d sdf of each used corner
};

// Gradient part of trilinear() at p (whose value the forward computed); the corner
// set and weights are recomputed exactly as the forward forms them.
template <bool F16, bool REF, class Cache>
__device__ __forceinline__ void trilinear_grad(const TableView& t, Cache& bc, const void* __restrict__ vox_, float px,
                                               float py, float pz, float v, float f, TriGradF* out) {
  using S = Store<F16>;
  const typename S::word* vox = reinterpret_cast<const typename S::word*>(vox_);
  int32_t vx[2], vy[2], vz[2];
  float fr[3];
  corner_setup<REF>(px, py, pz, v, vx, vy, vz, &fr[0], &fr[1], &fr[2]);
  const float iv = 1.0f / v;
  float W = 0.0f, dW[3] = {0.0f, 0.0f, 0.0f}, dN[3] = {0.0f, 0.0f, 0.0f}, w[8];
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    w[i] = 0.0f;
    out->vox[i] = -1;
    out->dfds[i] = 0.0f;
    const int32_t ix = vx[i & 1], iy = vy[(i >> 1) & 1], iz = vz[(i >> 2) & 1];
    uint32_t mk;
    const int32_t p = bc.get(t, voxel_to_block_axis(ix), voxel_to_block_axis(iy), voxel_to_block_axis(iz), &mk);
    if (p < 0) continue;
    const int64_t vi = (int64_t)p * kVox +
                       (voxel_local_axis(ix) + voxel_local_axis(iy) * kBlock + voxel_local_axis(iz) * kBlock * kBlock);
    const typename S::word wd = vox[vi];
    if (S::weight(wd) == 0u) continue;
    float phi[3], dphi[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const bool hi = (i >> a) & 1;
      phi[a] = hi ? fr[a] : 1.0f - fr[a];
      dphi[a] = hi ? iv : -iv;
    }
    w[i] = phi[0] * phi[1] * phi[2];
    const float sv = S::sdf(wd);
    out->vox[i] = vi;
    W += w[i];
    const float dwi[3] = {dphi[0] * phi[1] * phi[2], phi[0] * dphi[1] * phi[2], phi[0] * phi[1] * dphi[2]};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      dW[a] += dwi[a];
      dN[a] += dwi[a] * sv;
    }
  }
  const float invW = W > 0.0f ? 1.0f / W : 0.0f;
#pragma unroll
  for (int a = 0; a < 3; ++a) out->g[a] = (dN[a] - f * dW[a]) * invW;
#pragma unroll
  for (int i = 0; i < 8; ++i) out->dfds[i] = out->vox[i] >= 0 ? w[i] * invW : 0.0f;
}

struct Dual10 {
  float v;
  float d[10];
};

// Dual of an evaluation f(p(r)): r carries its derivative, tg the field gradient.
__device__ __forceinline__ Dual10 eval_dual(const TriGradF& tg, float fval, const Dual10& r, const RayHit& h,
                                            const float* M, int seed) {
  const float q[3] = {r.v * h.cdx, r.v * h.cdy, r.v * h.cdz};
  const float gc[3] = {M[0] * tg.g[0] + M[4] * tg.g[1] + M[8] * tg.g[2],
                       M[1] * tg.g[0] + M[5] * tg.g[1] + M[9] * tg.g[2],
                       M[2] * tg.g[0] + M[6] * tg.g[1] + M[10] * tg.g[2]};
  const float gw = tg.g[0] * h.wdx + tg.g[1] * h.wdy + tg.g[2] * h.wdz;
  Dual10 e;
  e.v = fval;
#pragma unroll
  for (int k = 0; k < 10; ++k) e.d[k] = gw * r.d[k];
  e.d[0] += q[1] * gc[2] - q[2] * gc[1];
  e.d[1] += q[2] * gc[0] - q[0] * gc[2];
  e.d[2] += q[0] * gc[1] - q[1] * gc[0];
  e.d[3] += gc[0];
  e.d[4] += gc[1];
  e.d[5] += gc[2];
  e.d[6 + seed] += 1.0f;
  return e;
}

// Pose Jacobian of the depth of a ray that crossed the surface at parameter tc, through
// the implicit function f(p(t)) = 0 with the field's gradient taken by central
// differences over +-1 voxel at the crossing: dt = -(g . dp) / (g . w).  For the
// camera-frame perturbation exp(xi^) the surface point moves by -omega x q - upsilon in
// the camera frame (q = t cd), so dt/domega = t (g_c x cd) / (g_c . cd) and dt/dupsilon =
// -g_c / (g_c . cd), g_c the gradient in the camera frame; depth = t cd_z.  0 where a
// sample is unweighted or the ray meets the surface at |cos| < 0.2.
template <bool F16, bool REF, bool DEV, class Cache>
__device__ __forceinline__ void smoothed_pose_jacobian(const TableView& t, Cache& bc, const void* __restrict__ vox,
                                                       const uint32_t* __restrict__ col, const RenderArgs& ra,
                                                       const RayHit& h, float tc, float J[6]) {
  const float v = ra.voxel;
  float cam[3];
  render_cam<DEV>(ra, cam);
  const float qx = cam[0] + tc * h.wdx, qy = cam[1] + tc * h.wdy, qz = cam[2] + tc * h.wdz;
  float gw[3];
  bool ok = true;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    float fp_ = 0.0f, fm_ = 0.0f;
    uint32_t rgb;
    const float ex = a == 0 ? v : 0.0f, ey = a == 1 ? v : 0.0f, ez = a == 2 ? v : 0.0f;
    ok = ok && trilinear<F16, false, REF>(t, bc, vox, col, qx + ex, qy + ey, qz + ez, v, &fp_, &rgb) == kSampleValid;
    ok = ok && trilinear<F16, false, REF>(t, bc, vox, col, qx - ex, qy - ey, qz - ez, v, &fm_, &rgb) == kSampleValid;
    gw[a] = ok ? (fp_ - fm_) : 0.0f;
  }
  float M[12];
  render_pose<DEV>(ra, M);
  const float gcx = M[0] * gw[0] + M[4] * gw[1] + M[8] * gw[2];
  const float gcy = M[1] * gw[0] + M[5] * gw[1] + M[9] * gw[2];
  const float gcz = M[2] * gw[0] + M[6] * gw[1] + M[10] * gw[2];
  const float gl = sqrtf(gcx * gcx + gcy * gcy + gcz * gcz);
  const float den = gcx * h.cdx + gcy * h.cdy + gcz * h.cdz;
  if (ok && gl > 0.0f && fabsf(den) >= 0.2f * gl) {
    const float s_ = h.cdz / den;
    J[0] = s_ * tc * (gcy * h.cdz - gcz * h.cdy);
    J[1] = s_ * tc * (gcz * h.cdx - gcx * h.cdz);
    J[2] = s_ * tc * (gcx * h.cdy - gcy * h.cdx);
    J[3] = -s_ * gcx;
    J[4] = -s_ * gcy;
    J[5] = -s_ * gcz;
  } else {
#pragma unroll
    for (int k = 0; k < 6; ++k) J[k] = 0.0f;
  }
}

// The smoothed pose Jacobian of every pixel of the last render, from its rendered depth
// (no second march): t = depth * d2r.
template <bool F16, bool REF, bool DEV>
__global__ __launch_bounds__(64) void k_gn_jacobian(TableView t, const DevState* __restrict__ st, RenderArgs ra,
                                                    const void* __restrict__ vox, const uint32_t* __restrict__ col,
                                                    const float* __restrict__ depth, float* __restrict__ jac) {
  const uint32_t x = blockIdx.x * 8u + (threadIdx.x & 7u);
  const uint32_t y = blockIdx.y * 8u + (threadIdx.x >> 3);
  if (x >= (uint32_t)ra.W || y >= (uint32_t)ra.H) return;
  const int pix = (int)y * ra.W + (int)x;
  float J[6] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
  const float d = depth[pix];
  if (st->n_alloc != 0u && d > 0.0f && d < 1e9f) {
    __shared__ int32_t slot_lds[8 * 64];
    BlockCache<false> bc;
    bc.init(slot_lds + threadIdx.x, nullptr, nullptr);
    RayHit h;
    RaySetup rs;
    (void)ray_setup<DEV>(ra, x, y, h, rs);
    smoothed_pose_jacobian<F16, REF, DEV>(t, bc, vox, col, ra, h, d * h.d2r, J);
  }
#pragma unroll
  for (int k = 0; k < 6; ++k) jac[6 * (size_t)pix + k] = J[k];
}

template <bool F16, bool REF, bool DEV>
__global__ __launch_bounds__(64) void k_render_adjoint(TableView t, const DevState* __restrict__ st, RenderArgs ra,
                                                       const void* __restrict__ vox, const uint32_t* __restrict__ col,
                                                       float* __restrict__ jac, const float* __restrict__ grad_depth,
                                                       float* __restrict__ grad_sdf) {
  const uint32_t x = blockIdx.x * 8u + (threadIdx.x & 7u);
  const uint32_t y = blockIdx.y * 8u + (threadIdx.x >> 3);
  if (x >= (uint32_t)ra.W || y >= (uint32_t)ra.H) return;
  const int pix = (int)y * ra.W + (int)x;
  float J[6] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
  if (st->n_alloc != 0u) {
    __shared__ int32_t slot_lds[8 * 64];
    BlockCache<false> bc;
    bc.init(slot_lds + threadIdx.x, nullptr, nullptr);
    const RayHit h = march_ray<F16, false, REF, DEV>(t, ra, vox, col, x, y, bc, 0u);
    if (h.found) {
      const float v = ra.voxel;
      float cam[3], M[12];
      render_cam<DEV>(ra, cam);
      render_pose<DEV>(ra, M);
      const float camx = cam[0], camy = cam[1], camz = cam[2];
      auto px = [&](float r) { return camx + r * h.wdx; };
      auto py = [&](float r) { return camy + r * h.wdy; };
      auto pz = [&](float r) { return camz + r * h.wdz; };
      TriGradF tg[4];
      Dual10 a, b, ad, bd, c;
#pragma unroll
      for (int k = 0; k < 10; ++k) a.d[k] = b.d[k] = c.d[k] = 0.0f;
      a.v = h.a;
      b.v = h.b;
      trilinear_grad<F16, REF>(t, bc, vox, px(h.a), py(h.a), pz(h.a), v, h.ad, &tg[0]);
      trilinear_grad<F16, REF>(t, bc, vox, px(h.b), py(h.b), pz(h.b), v, h.bd, &tg[1]);
      ad = eval_dual(tg[0], h.ad, a, h, M, 0);
      bd = eval_dual(tg[1], h.bd, b, h, M, 1);
      float fa = h.a, fad = h.ad, fb = h.b, fbd = h.bd;
      for (int it = 0; it < 3; ++it) {
        const float fc = fa + (fad / (fad - fbd)) * (fb - fa);  // the forward's point
        const float den = ad.v - bd.v, u = ad.v / den, den2 = den * den;
        c.v = fc;
#pragma unroll
        for (int k = 0; k < 10; ++k) {
          const float du = (bd.d[k] * ad.v - ad.d[k] * bd.v) / den2;
          c.d[k] = a.d[k] + du * (b.v - a.v) + u * (b.d[k] - a.d[k]);
        }
        if (it == 2) break;
        float fcd;
        uint32_t rgb;
        trilinear<F16, false, REF>(t, bc, vox, col, px(fc), py(fc), pz(fc), v, &fcd, &rgb);
        trilinear_grad<F16, REF>(t, bc, vox, px(fc), py(fc), pz(fc), v, fcd, &tg[2 + it]);
        const Dual10 cd = eval_dual(tg[2 + it], fcd, c, h, M, 2 + it);
        if (fad * fcd > 0.0f) {
          fa = fc;
          fad = fcd;
          a = c;
          ad = cd;
        } else {
          fb = fc;
          fbd = fcd;
          b = c;
          bd = cd;
        }
      }
#pragma unroll
      for (int k = 0; k < 6; ++k) J[k] = c.d[k] * h.cdz;
      if (grad_sdf && grad_depth) {
        const float g = grad_depth[pix] * h.cdz;
        if (g != 0.0f) {
          for (int e = 0; e < 4; ++e) {
            const float ge = g * c.d[6 + e];
            for (int i = 0; i < 8; ++i)
              if (tg[e].vox[i] >= 0) atomicAdd(&grad_sdf[tg[e].vox[i]], ge * tg[e].dfds[i]);
          }
        }
      }
    }
  }
  if (jac) {
#pragma unroll
    for (int k = 0; k < 6; ++k) jac[6 * (size_t)pix + k] = J[k];
  }
}

// ---------------------------------------------------------------------------------
// Gauss-Newton pose refinement, FP64 (no counterpart in the reference)
// ---------------------------------------------------------------------------------

constexpr int kGnTerms = 29;  // 21 (upper triangle of H) + 6 (g) + cost + count

__device__ __forceinline__ double shfl_xor_d(double v, int o) {
  int lo = __double2loint(v), hi = __double2hiint(v);
  lo = __shfl_xor(lo, o, 64);
  hi = __shfl_xor(hi, o, 64);
  return __hiloint2double(hi, lo);
}

// Per-workgroup partial sums of J J^T, J r, r^2 and the count over pixels with a
// rendered and an observed depth and |r| < max_res (r = rendered - observed), away
// from depth edges: a pixel whose four neighbours are not all rendered within
// max(2 cm, 2 % of its depth) of it is left out, since a small pose change can move
// its ray onto another surface, which the Jacobian does not describe.
// Residuals r = rendered - observed over pixels with both depths, |r| < max_res and no
// depth edge around them; weight (a / sigma(o))^2 with the sensor's noise model
// sigma(o) = a + b (o - z0)^2 when noise.x > 0 (noise = {a, b, z0}), else 1.
__global__ __launch_bounds__(256) void k_gn_accum(const float* __restrict__ depth, const float* __restrict__ obs,
                                                  const float* __restrict__ jac, int W, int H, float max_res,
                                                  float3 noise, double* __restrict__ partials) {
  const int npix = W * H;
  const int pix = blockIdx.x * 256 + threadIdx.x;
  double acc[kGnTerms];
#pragma unroll
  for (int k = 0; k < kGnTerms; ++k) acc[k] = 0.0;
  if (pix < npix) {
    const int x = pix % W, y = pix / W;
    const float d = depth[pix], o = obs[pix];
    const float r = d - o;
    bool use = d < 1e9f && o > 0.0f && fabsf(r) < max_res && x > 0 && y > 0 && x + 1 < W && y + 1 < H;
    if (use) {
      const float tol = fmaxf(0.02f, 0.02f * d);
      const float nb[4] = {depth[pix - 1], depth[pix + 1], depth[pix - W], depth[pix + W]};
#pragma unroll
      for (int k = 0; k < 4; ++k) use = use && nb[k] < 1e9f && fabsf(nb[k] - d) <= tol;
    }
    if (use) {
      double w = 1.0;
      if (noise.x > 0.0f) {
        const double dz = (double)o - (double)noise.z;
        const double sg = (double)noise.x + (double)noise.y * dz * dz;
        w = ((double)noise.x / sg) * ((double)noise.x / sg);
      }
      double Jd[6];
#pragma unroll
      for (int k = 0; k < 6; ++k) Jd[k] = (double)jac[6 * (size_t)pix + k];
      int n = 0;
#pragma unroll
      for (int i = 0; i < 6; ++i)
#pragma unroll
        for (int j = i; j < 6; ++j) acc[n++] = w * Jd[i] * Jd[j];
#pragma unroll
      for (int i = 0; i < 6; ++i) acc[21 + i] = w * Jd[i] * (double)r;
      acc[27] = w * (double)r * (double)r;
      acc[28] = 1.0;
    }
  }
  __shared__ double ws[4][kGnTerms];
#pragma unroll
  for (int k = 0; k < kGnTerms; ++k) {
    double v = acc[k];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += shfl_xor_d(v, o);
    if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6][k] = v;
  }
  __syncthreads();
  if (threadIdx.x < kGnTerms)
    partials[(size_t)blockIdx.x * kGnTerms + threadIdx.x] =
        ws[0][threadIdx.x] + ws[1][threadIdx.x] + ws[2][threadIdx.x] + ws[3][threadIdx.x];
}

struct Pose64 {
  double T[16];
};

// Sums the partials in a fixed order, solves (H + lambda diag(H)) delta = -g by
// Cholesky and composes pose * exp(delta^), all in FP64.  out: new pose (16),
// delta (6), cost, count, solved flag.
// The sum: lane l of the 256 adds rows l, l + 256, ... in turn, then the lanes' sums
// are added by a fixed butterfly within each wave and the four waves' sums in wave
// order, so the result does not depend on timing.
// pose_in (may be null: the pose passed by value) may be `out` itself (the previous
// iteration's result): lane 0 reads it before writing.  posef (may be null) receives the
// new pose as 16 floats and the camera position as 3 more, the form a render reads.
constexpr int kGnSolveThreads = 256;
__global__ __launch_bounds__(kGnSolveThreads) void k_gn_solve(const double* __restrict__ partials, int nwg,
                                                              double lambda, Pose64 pose, const double* pose_in,
                                                              double* out, float* __restrict__ posef) {
  __shared__ double wsum[kGnSolveThreads / 64][kGnTerms];
  __shared__ double sum[kGnTerms];
  {
    double acc[kGnTerms];
#pragma unroll
    for (int k = 0; k < kGnTerms; ++k) acc[k] = 0.0;
    for (int b = threadIdx.x; b < nwg; b += kGnSolveThreads) {
      const double* row = partials + (size_t)b * kGnTerms;
#pragma unroll
      for (int k = 0; k < kGnTerms; ++k) acc[k] += row[k];
    }
#pragma unroll
    for (int k = 0; k < kGnTerms; ++k) {
      double v = acc[k];
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) v += shfl_xor_d(v, o);
      if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6][k] = v;
    }
  }
  __syncthreads();
  if (threadIdx.x < kGnTerms) {
    double v = wsum[0][threadIdx.x];
#pragma unroll
    for (int w = 1; w < kGnSolveThreads / 64; ++w) v += wsum[w][threadIdx.x];
    sum[threadIdx.x] = v;
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  double P[16];
  for (int i = 0; i < 16; ++i) P[i] = pose_in ? pose_in[i] : pose.T[i];
  double H[36], g[6];
  int n = 0;
  for (int i = 0; i < 6; ++i)
    for (int j = i; j < 6; ++j) {
      H[6 * i + j] = H[6 * j + i] = sum[n++];
    }
  for (int i = 0; i < 6; ++i) g[i] = sum[21 + i];
  for (int i = 0; i < 6; ++i) H[7 * i] += lambda * H[7 * i] + 1e-12;
  // Cholesky H = L L^T
  double L[36];
  for (int k = 0; k < 36; ++k) L[k] = 0.0;
  bool ok = sum[28] >= 6.0;
  for (int j = 0; j < 6 && ok; ++j) {
    double d = H[7 * j];
    for (int k = 0; k < j; ++k) d -= L[6 * j + k] * L[6 * j + k];
    if (!(d > 0.0)) {
      ok = false;
      break;
    }
    L[7 * j] = sqrt(d);
    for (int i = j + 1; i < 6; ++i) {
      double s = H[6 * i + j];
      for (int k = 0; k < j; ++k) s -= L[6 * i + k] * L[6 * j + k];
      L[6 * i + j] = s / L[7 * j];
    }
  }
  double delta[6] = {0, 0, 0, 0, 0, 0};
  if (ok) {
    double z[6];
    for (int i = 0; i < 6; ++i) {
      double s = -g[i];
      for (int k = 0; k < i; ++k) s -= L[6 * i + k] * z[k];
      z[i] = s / L[7 * i];
    }
    for (int i = 5; i >= 0; --i) {
      double s = z[i];
      for (int k = i + 1; k < 6; ++k) s -= L[6 * k + i] * delta[k];
      delta[i] = s / L[7 * i];
    }
  }
  // exp(delta^): rotation by Rodrigues, translation V upsilon
  const double wx = delta[0], wy = delta[1], wz = delta[2];
  const double th = sqrt(wx * wx + wy * wy + wz * wz);
  double A, Bc, Cc;
  if (th < 1e-8) {
    A = 1.0 - th * th / 6.0;
    Bc = 0.5 - th * th / 24.0;
    Cc = 1.0 / 6.0 - th * th / 120.0;
  } else {
    A = sin(th) / th;
    Bc = (1.0 - cos(th)) / (th * th);
    Cc = (th - sin(th)) / (th * th * th);
  }
  const double K[9] = {0.0, -wz, wy, wz, 0.0, -wx, -wy, wx, 0.0};
  double K2[9];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double s = 0.0;
      for (int k = 0; k < 3; ++k) s += K[3 * i + k] * K[3 * k + j];
      K2[3 * i + j] = s;
    }
  double Rd[9], V[9];
  for (int i = 0; i < 9; ++i) {
    const double I = (i % 4 == 0) ? 1.0 : 0.0;
    Rd[i] = I + A * K[i] + Bc * K2[i];
    V[i] = I + Bc * K[i] + Cc * K2[i];
  }
  double td[3];
  for (int i = 0; i < 3; ++i) td[i] = V[3 * i] * delta[3] + V[3 * i + 1] * delta[4] + V[3 * i + 2] * delta[5];
  // pose * [Rd td; 0 1]
  double Tn[16];
  for (int i = 0; i < 3; ++i) {
    for (int j = 0; j < 3; ++j)
      Tn[4 * i + j] = P[4 * i] * Rd[j] + P[4 * i + 1] * Rd[3 + j] + P[4 * i + 2] * Rd[6 + j];
    Tn[4 * i + 3] = P[4 * i] * td[0] + P[4 * i + 1] * td[1] + P[4 * i + 2] * td[2] + P[4 * i + 3];
  }
  Tn[12] = Tn[13] = Tn[14] = 0.0;
  Tn[15] = 1.0;
  for (int i = 0; i < 16; ++i) out[i] = Tn[i];
  for (int i = 0; i < 6; ++i) out[16 + i] = delta[i];
  out[22] = sum[27];
  out[23] = sum[28];
  out[24] = ok ? 1.0 : 0.0;
  if (posef) {
    for (int i = 0; i < 16; ++i) posef[i] = (float)Tn[i];
    posef[16] = (float)Tn[3];
    posef[17] = (float)Tn[7];
    posef[18] = (float)Tn[11];
  }
}

// ---------------------------------------------------------------------------------
// Reset
// ---------------------------------------------------------------------------------

// Zeroes the voxels of the blocks that are live.  The reference's reset
// (VoxelScene.cu:1166-1236) leaves voxel memory as it was, so a block handed out
// again after a reset would start from stale values; zeroing them keeps a reset
// scene equal to a new one.  Runs before k_reset_table, which clears n_alloc.
__global__ __launch_bounds__(256) void k_reset_voxels(const DevState* st, void* vox, uint32_t word_bytes,
                                                      uint32_t* col, uint32_t* dmeta, float2* dval, uint32_t* pcnt) {
  const size_t tid = (size_t)blockIdx.x * 256 + threadIdx.x;
  const size_t stride = (size_t)gridDim.x * 256;
  const size_t nvox = (size_t)st->n_alloc * kVox;
  for (size_t i = tid; i < (size_t)st->n_alloc * 8u; i += stride) pcnt[i] = 0;
  for (size_t i = tid; i < nvox; i += stride) {
    if (word_bytes == 8) reinterpret_cast<uint2*>(vox)[i] = make_uint2(0u, 0u);
    else reinterpret_cast<uint32_t*>(vox)[i] = 0u;
    col[i] = 0u;
    if (dmeta) {
      dmeta[i] = 0u;
      dval[i] = make_float2(0.0f, 0.0f);
    }
  }
}

// Empties the table (no used slot, no claim) and initialises the scene state and
// both counter sets (VoxelScene.cu:250-284: heap full, table free, mutexes open).
__global__ __launch_bounds__(256) void k_reset_table(TableView t, DevState* st) {
  const size_t tid = (size_t)blockIdx.x * 256 + threadIdx.x;
  const size_t stride = (size_t)gridDim.x * 256;
  for (size_t b = tid; b < t.num_buckets; b += stride) {
    t.cnt[b] = 0;
    t.claim[b] = kNoClaim;
  }
  for (size_t w = tid; w <= t.occ_mask; w += stride) t.occ[w] = 0ull;
  for (size_t w = tid; w <= t.hyp_mask; w += stride) t.hyp[w] = 0ull;
  for (size_t w = tid; w <= t.idx_mask; w += stride) t.idx[w] = make_uint4(kIdxEmpty, kIdxEmpty, 0u, 0u);
  if (tid == 0) {
    st->n_alloc = 0;
    st->n_alloc_prev = 0;
    st->n_alloc_next = 0;
    st->frames = 0;
    for (int s = 0; s < 2; ++s) {
      FrameCounters& f = st->set[s];
      f.valid_pixels = f.claim_pixels = f.winners = f.heap_fail = f.out_of_range = 0;
      f.visible = f.updated = f.nonfinite = f.claimed_buckets = 0;
      f.zmin_key = 0xFFFFFFFFu;
      f.zmax_key = 0u;
    }
  }
}

// ---------------------------------------------------------------------------------
// Surface-point extraction (RayCastRender.cu:629-907, RayCastRender.cpp:307-364)
// ---------------------------------------------------------------------------------

// Device-wide exclusive scan of n values (u8 or u32) into out, in chunks of
// kScanChunk per 1024-thread workgroup: chunk sums, a scan of the sums by one
// workgroup (1024 sums per pass), then the chunks again.
constexpr uint32_t kScanChunk = kScanChunkSize;

template <class T>
__global__ __launch_bounds__(1024) void k_scan_sums(const T* __restrict__ in, uint32_t n, uint32_t* __restrict__ sums) {
  const uint32_t base = blockIdx.x * kScanChunk;
  uint32_t v = 0;
#pragma unroll
  for (uint32_t k = 0; k < kScanChunk / 1024u; ++k) {
    const uint32_t i = base + k * 1024u + threadIdx.x;
    v += i < n ? (uint32_t)in[i] : 0u;
  }
  uint32_t total;
  (void)block_excl_scan<1024>(v, &total);
  if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

__global__ __launch_bounds__(1024) void k_scan_top(uint32_t* __restrict__ sums, uint32_t nchunks) {
  uint32_t carry = 0;
  for (uint32_t base = 0; base < nchunks; base += 1024u) {
    const uint32_t i = base + threadIdx.x;
    const uint32_t v = i < nchunks ? sums[i] : 0u;
    uint32_t total;
    const uint32_t ex = block_excl_scan<1024>(v, &total);
    if (i < nchunks) sums[i] = ex + carry;
    carry += total;
  }
}

template <class T>
__global__ __launch_bounds__(1024) void k_scan_apply(const T* __restrict__ in, uint32_t n,
                                                     const uint32_t* __restrict__ sums, uint32_t* __restrict__ out) {
  // thread t holds elements base + 4t .. base + 4t + 3 (consecutive, in order)
  const uint32_t base = blockIdx.x * kScanChunk + threadIdx.x * (kScanChunk / 1024u);
  uint32_t v[kScanChunk / 1024u], run = 0;
#pragma unroll
  for (uint32_t k = 0; k < kScanChunk / 1024u; ++k) {
    const uint32_t i = base + k;
    v[k] = i < n ? (uint32_t)in[i] : 0u;
    run += v[k];
  }
  uint32_t total;
  uint32_t ex = block_excl_scan<1024>(run, &total) + sums[blockIdx.x];
#pragma unroll
  for (uint32_t k = 0; k < kScanChunk / 1024u; ++k) {
    const uint32_t i = base + k;
    if (i < n) out[i] = ex;
    ex += v[k];
  }
}

// Every live block at its rank in hash-entry order (bucket, then slot): bucket b's
// entries start at bucket_start[b] (the exclusive scan of the bucket counts).  oblk[k]
// = (block coordinates, pointer) of rank k.
__global__ __launch_bounds__(256) void k_ext_order(TableView t, const DevState* __restrict__ st,
                                                   const uint32_t* __restrict__ bucket_start, int4* __restrict__ oblk) {
  const uint32_t p = blockIdx.x * 256u + threadIdx.x;
  if (p >= st->n_alloc) return;
  const int4 b = t.blk[p];
  const uint32_t e = (uint32_t)b.w;
  oblk[bucket_start[e / kBucket] + e % kBucket] = make_int4(b.x, b.y, b.z, (int32_t)p);
}

// Pointers of the 3x3x3 blocks around every live block (-1 where none is allocated),
// kExtNb words per block: nbp[kExtNb * k + i] for neighbour i = (dx + 1) + 3 (dy + 1)
// + 9 (dz + 1) of rank k.
constexpr uint32_t kExtNb = kExtNbWords;
__global__ __launch_bounds__(256) void k_ext_neighbours(TableView t, const DevState* __restrict__ st,
                                                        const int4* __restrict__ oblk, int32_t* __restrict__ nbp) {
  const uint32_t g = blockIdx.x * 256u + threadIdx.x;
  const uint32_t k = g / kExtNb, i = g % kExtNb;
  if (k >= st->n_alloc || i >= 27u) return;
  const int4 b = oblk[k];
  const int dx = (int)(i % 3u) - 1, dy = (int)((i / 3u) % 3u) - 1, dz = (int)(i / 9u) - 1;
  nbp[(size_t)k * kExtNb + i] = (dx | dy | dz) == 0 ? b.w : index_lookup(t, b.x + dx, b.y + dy, b.z + dz);
}

struct ExtractArgs {
  float voxel;
  float min_sdf, max_sdf;
  uint32_t min_weight;
  uint32_t max_points;
};

// One block of the extraction in LDS (one wave per block): its 512 voxel words, the
// pointers of the 3x3x3 blocks around it (-1 where none is allocated), and per axis the
// voxel getVoxel reads at (corner + (i + 0.5) v + o v) for local coordinate i and
// offset o in {-1, 0, +1}: (block offset + 1) * 8 + local coordinate, kExtFar when it
// lies beyond the neighbouring blocks.  getVoxel's rounding of one coordinate depends
// on that coordinate alone, so a neighbour's voxel follows from three table entries.
constexpr uint8_t kExtFar = 255;
template <bool F16> struct ExtBlock {
  typename Store<F16>::word sv[kVox];
  int32_t nbp[27];
  uint8_t ax[3][kBlock][3];
  uint16_t list[kVox];  // compacted voxel indices
  uint64_t em[8];       // emitting voxels of the block, one bit per voxel
};

// getVoxel (RayCastRender.cu:55-121) at world position p: weight and SDF, zero when the
// block is not allocated.  Blocks next to the staged one come from its pointer table,
// the staged block from LDS, others from the block index.
template <bool F16>
__device__ __forceinline__ uint32_t ext_get_voxel(const TableView& t, const void* __restrict__ vox_,
                                                  const ExtBlock<F16>& eb, int4 b, float px, float py, float pz,
                                                  float v, float* sdf) {
  using S = Store<F16>;
  const int32_t vx = ray_voxel_of_q(px / v), vy = ray_voxel_of_q(py / v), vz = ray_voxel_of_q(pz / v);
  const int32_t bx = voxel_to_block_axis(vx), by = voxel_to_block_axis(vy), bz = voxel_to_block_axis(vz);
  const int l = voxel_local_axis(vx) + voxel_local_axis(vy) * kBlock + voxel_local_axis(vz) * kBlock * kBlock;
  const int32_t dx = bx - b.x, dy = by - b.y, dz = bz - b.z;
  typename S::word w;
  if ((dx | dy | dz) == 0) {
    w = eb.sv[l];
  } else {
    const bool near = dx >= -1 && dx <= 1 && dy >= -1 && dy <= 1 && dz >= -1 && dz <= 1;
    const int32_t q = near ? eb.nbp[(dx + 1) + 3 * (dy + 1) + 9 * (dz + 1)] : index_lookup(t, bx, by, bz);
    if (q < 0) {
      *sdf = 0.0f;
      return 0u;
    }
    w = reinterpret_cast<const typename S::word*>(vox_)[(size_t)q * kVox + (size_t)l];
  }
  *sdf = S::sdf(w);
  return S::weight(w);
}

// Voxel centre coordinate along one axis (block coordinate bc, local i).
__device__ __forceinline__ float ext_centre(int32_t bc, int i, float v) {
  return (float)(bc * kBlock) * v + ((float)i + 0.5f) * v;
}

// Fills eb.ax for block b (lanes of one wave; the caller synchronises).
template <bool F16> __device__ __forceinline__ void ext_axis_table(ExtBlock<F16>& eb, int4 b, float v) {
  for (uint32_t j = threadIdx.x; j < 72u; j += 64u) {
    const int a = (int)(j / 24u), i = (int)((j / 3u) % 8u), o = (int)(j % 3u) - 1;
    const int32_t bc = a == 0 ? b.x : (a == 1 ? b.y : b.z);
    const float p = ext_centre(bc, i, v) + (float)o * v;
    const int32_t vi = ray_voxel_of_q(p / v);
    const int32_t db = voxel_to_block_axis(vi) - bc;
    eb.ax[a][i][o + 1] = (db >= -1 && db <= 1) ? (uint8_t)((db + 1) * kBlock + voxel_local_axis(vi)) : kExtFar;
  }
}

// The neighbour of voxel (x, y, z) at offset (ox, oy, oz) (one axis +-1) as the
// reference reads it (RayCastRender.cu:766-800): by index inside the block, through
// getVoxel at the voxel centre + offset * v outside it.  Returns its weight.
template <bool F16>
__device__ __forceinline__ uint32_t ext_neighbour(const TableView& t, const void* __restrict__ vox_,
                                                  const ExtBlock<F16>& eb, int4 b, float v, int x, int y, int z,
                                                  int ox, int oy, int oz, float* nsd) {
  using S = Store<F16>;
  const int nx = x + ox, ny = y + oy, nz = z + oz;
  const bool inb = nx >= 0 && nx < kBlock && ny >= 0 && ny < kBlock && nz >= 0 && nz < kBlock;
  // per axis (block offset + 1) * 8 + local: the index itself inside the block, the
  // getVoxel table outside it
  const uint32_t ex = inb ? (uint32_t)(kBlock + nx) : eb.ax[0][x][ox + 1];
  const uint32_t ey = inb ? (uint32_t)(kBlock + ny) : eb.ax[1][y][oy + 1];
  const uint32_t ez = inb ? (uint32_t)(kBlock + nz) : eb.ax[2][z][oz + 1];
  if (ex == kExtFar || ey == kExtFar || ez == kExtFar)
    return ext_get_voxel<F16>(t, vox_, eb, b, ext_centre(b.x, x, v) + (float)ox * v,
                              ext_centre(b.y, y, v) + (float)oy * v, ext_centre(b.z, z, v) + (float)oz * v, v, nsd);
  const int nbi = (int)(ex / kBlock) + 3 * (int)(ey / kBlock) + 9 * (int)(ez / kBlock);  // 13: this block
  const int l = (int)(ex % kBlock) + (int)(ey % kBlock) * kBlock + (int)(ez % kBlock) * kBlock * kBlock;
  typename S::word w = eb.sv[l];
  if (nbi != 13) {
    const int32_t q = eb.nbp[nbi];
    if (q < 0) {
      *nsd = 0.0f;
      return 0u;
    }
    w = reinterpret_cast<const typename S::word*>(vox_)[(size_t)q * kVox + (size_t)l];
  }
  *nsd = S::sdf(w);
  return S::weight(w);
}

__device__ __forceinline__ void ext_dir(int d, int* ox, int* oy, int* oz) {
  *ox = d == 0 ? 1 : (d == 1 ? -1 : 0);
  *oy = d == 2 ? 1 : (d == 3 ? -1 : 0);
  *oz = d == 4 ? 1 : (d == 5 ? -1 : 0);
}

// Voxel passes the weight and SDF thresholds (RayCastRender.cu:738-752).
template <bool F16> __device__ __forceinline__ bool ext_in_band(typename Store<F16>::word w, const ExtractArgs& a) {
  using S = Store<F16>;
  const uint32_t wt = S::weight(w);
  const float sd = S::sdf(w);
  return wt != 0u && wt >= a.min_weight && !(sd < a.min_sdf) && !(sd > a.max_sdf);
}

// Whether in-band voxel l of the staged block emits a point: a weighted neighbour of
// the opposite sign, or |sdf| < 0.001 (RayCastRender.cu:760-834).
template <bool F16>
__device__ __forceinline__ bool ext_emits(const TableView& t, const void* __restrict__ vox_, const ExtBlock<F16>& eb,
                                          int4 b, float v, int l) {
  using S = Store<F16>;
  const int x = l % kBlock, y = (l / kBlock) % kBlock, z = l / (kBlock * kBlock);
  const float sd = S::sdf(eb.sv[l]);
  for (int d = 0; d < 6; ++d) {
    int ox, oy, oz;
    ext_dir(d, &ox, &oy, &oz);
    float nsd;
    if (ext_neighbour<F16>(t, vox_, eb, b, v, x, y, z, ox, oy, oz, &nsd) == 0u) continue;
    if ((sd < 0.0f && nsd > 0.0f) || (sd > 0.0f && nsd < 0.0f)) return true;
  }
  return fabsf(sd) < 0.001f;
}

// The point of emitting voxel l of the staged block: the mean of its zero crossings,
// else its centre (RayCastRender.cu:760-840).
template <bool F16>
__device__ __forceinline__ float3 ext_iso(const TableView& t, const void* __restrict__ vox_, const ExtBlock<F16>& eb,
                                          int4 b, float v, int l) {
  using S = Store<F16>;
  const int x = l % kBlock, y = (l / kBlock) % kBlock, z = l / (kBlock * kBlock);
  const float sd = S::sdf(eb.sv[l]);
  const float cx = ext_centre(b.x, x, v), cy = ext_centre(b.y, y, v), cz = ext_centre(b.z, z, v);
  float ax = 0.0f, ay = 0.0f, az = 0.0f;
  int n = 0;
  for (int d = 0; d < 6; ++d) {
    int ox, oy, oz;
    ext_dir(d, &ox, &oy, &oz);
    float nsd;
    if (ext_neighbour<F16>(t, vox_, eb, b, v, x, y, z, ox, oy, oz, &nsd) == 0u) continue;
    if ((sd < 0.0f && nsd > 0.0f) || (sd > 0.0f && nsd < 0.0f)) {
      const float tt = sd / (sd - nsd);
      ax = ax + (cx + (float)ox * v * tt);
      ay = ay + (cy + (float)oy * v * tt);
      az = az + (cz + (float)oz * v * tt);
      n++;
    }
  }
  if (n == 0) return make_float3(cx, cy, cz);
  return make_float3(ax / (float)n, ay / (float)n, az / (float)n);
}

// The voxel getVoxel reads at voxel indices (vx, vy, vz) (block index, local index),
// from the staged block, its neighbours' pointers or the block index.
template <bool F16>
__device__ __forceinline__ uint32_t ext_read_voxel(const TableView& t, const void* __restrict__ vox_,
                                                   const ExtBlock<F16>& eb, int4 b, int32_t vx, int32_t vy,
                                                   int32_t vz, float* sdf) {
  using S = Store<F16>;
  const int32_t bx = voxel_to_block_axis(vx), by = voxel_to_block_axis(vy), bz = voxel_to_block_axis(vz);
  const int l = voxel_local_axis(vx) + voxel_local_axis(vy) * kBlock + voxel_local_axis(vz) * kBlock * kBlock;
  const int32_t dx = bx - b.x, dy = by - b.y, dz = bz - b.z;
  typename S::word w = eb.sv[l];
  if ((dx | dy | dz) != 0) {
    const bool near = dx >= -1 && dx <= 1 && dy >= -1 && dy <= 1 && dz >= -1 && dz <= 1;
    const int32_t q = near ? eb.nbp[(dx + 1) + 3 * (dy + 1) + 9 * (dz + 1)] : index_lookup(t, bx, by, bz);
    if (q < 0) {
      *sdf = 0.0f;
      return 0u;
    }
    w = reinterpret_cast<const typename S::word*>(vox_)[(size_t)q * kVox + (size_t)l];
  }
  *sdf = S::sdf(w);
  return S::weight(w);
}

// computeNormalFromSDF (RayCastRender.cu:629-676): central differences of getVoxel at
// +-v/2 per axis (0 for an axis whose two voxels are not both weighted), normalised
// when the length exceeds 1e-6.  getVoxel's voxel index of a coordinate depends on
// that coordinate alone: three per axis (p - eps, p, p + eps) serve all six reads.
template <bool F16>
__device__ __forceinline__ float3 ext_normal(const TableView& t, const void* __restrict__ vox_,
                                             const ExtBlock<F16>& eb, int4 b, float3 p, float v) {
  const float eps = v * 0.5f;
  const int32_t x0 = ray_voxel_of_q(p.x / v), xm = ray_voxel_of_q((p.x - eps) / v), xp = ray_voxel_of_q((p.x + eps) / v);
  const int32_t y0 = ray_voxel_of_q(p.y / v), ym = ray_voxel_of_q((p.y - eps) / v), yp = ray_voxel_of_q((p.y + eps) / v);
  const int32_t z0 = ray_voxel_of_q(p.z / v), zm = ray_voxel_of_q((p.z - eps) / v), zp = ray_voxel_of_q((p.z + eps) / v);
  float g[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll
  for (int ax_ = 0; ax_ < 3; ++ax_) {
    float sp, sm;
    const uint32_t wp = ext_read_voxel<F16>(t, vox_, eb, b, ax_ == 0 ? xp : x0, ax_ == 1 ? yp : y0, ax_ == 2 ? zp : z0,
                                            &sp);
    const uint32_t wm = ext_read_voxel<F16>(t, vox_, eb, b, ax_ == 0 ? xm : x0, ax_ == 1 ? ym : y0, ax_ == 2 ? zm : z0,
                                            &sm);
    if (wp > 0u && wm > 0u) g[ax_] = (sp - sm) / (2.0f * eps);
  }
  const float len = sqrtf(g[0] * g[0] + g[1] * g[1] + g[2] * g[2]);
  if (len > 1e-6f) {
    g[0] = g[0] / len;
    g[1] = g[1] / len;
    g[2] = g[2] / len;
  }
  return make_float3(g[0], g[1], g[2]);
}

// The extraction's walk over the live blocks in hash order, one wave per block: lane
// L holds voxels L + 64 z (z = 0..7).  Workgroup g takes ranks g, g + G, ...; the next
// block's voxel words and neighbour pointers are loaded while the current block is
// processed, and the list entry of the block after that one.
template <bool F16> struct ExtWalk {
  using W = typename Store<F16>::word;
  const W* vw;
  const int4* oblk;
  const int32_t* nbp;
  uint32_t nb;
  int4 b;       // current block (coordinates, pointer)
  int4 bn;      // next block's list entry
  W w[8];       // this lane's voxel words of the current block
  int32_t nq;   // neighbour pointer (lanes < 27)

  __device__ void fetch(uint32_t k, int4 bk) {
    const size_t base = (size_t)bk.w * kVox + threadIdx.x;
#pragma unroll
    for (int z = 0; z < 8; ++z) w[z] = vw[base + 64u * (uint32_t)z];
    nq = threadIdx.x < 27u ? nbp[(size_t)k * kExtNb + threadIdx.x] : -1;
  }
  __device__ void stage(ExtBlock<F16>& eb, float v) const {
#pragma unroll
    for (int z = 0; z < 8; ++z) eb.sv[threadIdx.x + 64u * (uint32_t)z] = w[z];
    if (threadIdx.x < 27u) eb.nbp[threadIdx.x] = nq;
    ext_axis_table<F16>(eb, b, v);
  }
};

// Points per block (bcount) and the emitting voxels (ebits, 8 words per block).
template <bool F16>
__global__ __launch_bounds__(64) void k_ext_count(TableView t, const DevState* __restrict__ st,
                                                  const int4* __restrict__ oblk, const int32_t* __restrict__ nbp,
                                                  const void* __restrict__ vox, ExtractArgs a,
                                                  uint32_t* __restrict__ bcount, uint64_t* __restrict__ ebits) {
  using S = Store<F16>;
  __shared__ ExtBlock<F16> eb;
  ExtWalk<F16> wk;
  wk.vw = reinterpret_cast<const typename S::word*>(vox);
  wk.oblk = oblk;
  wk.nbp = nbp;
  wk.nb = st->n_alloc;
  const uint32_t G = gridDim.x, lane = threadIdx.x;
  const uint64_t below = (1ull << lane) - 1ull;
  uint32_t k = blockIdx.x;
  if (k < wk.nb) {
    wk.b = oblk[k];
    wk.fetch(k, wk.b);
    wk.bn = k + G < wk.nb ? oblk[k + G] : make_int4(0, 0, 0, 0);
  }
  for (; k < wk.nb; k += G) {
    // in-band voxels of the current block, in voxel order, before its words are staged
    uint32_t nband = 0;
#pragma unroll
    for (int z = 0; z < 8; ++z) {
      const bool in = ext_in_band<F16>(wk.w[z], a);
      const uint64_t m = __ballot(in);
      if (in) eb.list[nband + (uint32_t)__popcll(m & below)] = (uint16_t)(lane + 64u * (uint32_t)z);
      nband += (uint32_t)__popcll(m);
    }
    wk.stage(eb, a.voxel);
    if (lane < 8u) eb.em[lane] = 0ull;
    __syncthreads();
    const int4 b = wk.b;
    if (k + G < wk.nb) wk.fetch(k + G, wk.bn);
    wk.b = wk.bn;
    wk.bn = k + 2u * G < wk.nb ? oblk[k + 2u * G] : make_int4(0, 0, 0, 0);
    uint32_t c = 0;
    for (uint32_t j0 = 0; j0 < nband; j0 += 64u) {
      const uint32_t j = j0 + lane;
      if (j < nband) {
        const int l = eb.list[j];
        if (ext_emits<F16>(t, vox, eb, b, a.voxel, l)) {
          c++;
          atomicOr(reinterpret_cast<unsigned long long*>(&eb.em[l >> 6]), 1ull << (l & 63));
        }
      }
    }
    c = wave_sum(c);
    __syncthreads();
    if (lane == 0) bcount[k] = c;
    if (lane < 8u) ebits[(size_t)k * 8u + lane] = eb.em[lane];
    __syncthreads();  // eb is read above until every lane is done
  }
  // no points beyond the live blocks (the scan runs over all block slots)
  for (uint32_t j = wk.nb + blockIdx.x * 64u + lane; j < t.num_blocks; j += G * 64u) bcount[j] = 0u;
}

// The points: block k's start pstart[k], then voxel order; the first max_points kept.
template <bool F16>
__global__ __launch_bounds__(64) void k_ext_write(TableView t, const DevState* __restrict__ st,
                                                  const int4* __restrict__ oblk, const int32_t* __restrict__ nbp,
                                                  const void* __restrict__ vox, const uint32_t* __restrict__ col,
                                                  ExtractArgs a, const uint32_t* __restrict__ bcount,
                                                  const uint64_t* __restrict__ ebits,
                                                  const uint32_t* __restrict__ pstart, float4* __restrict__ out_pos,
                                                  float4* __restrict__ out_col, float4* __restrict__ out_nrm) {
  using S = Store<F16>;
  __shared__ ExtBlock<F16> eb;
  ExtWalk<F16> wk;
  wk.vw = reinterpret_cast<const typename S::word*>(vox);
  wk.oblk = oblk;
  wk.nbp = nbp;
  wk.nb = st->n_alloc;
  const uint32_t G = gridDim.x, lane = threadIdx.x;
  const uint64_t below = (1ull << lane) - 1ull;
  __shared__ float c255[256];  // c / 255 for the 8-bit colour channels
  for (uint32_t i = lane; i < 256u; i += 64u) c255[i] = (float)i / 255.0f;
  uint32_t k = blockIdx.x, start = 0, cnt = 0, sn = 0, cn = 0;
  uint64_t em = 0, emn = 0;  // lane z < 8: emit word z of the current / next block
  if (k < wk.nb) {
    wk.b = oblk[k];
    start = pstart[k];
    cnt = bcount[k];
    if (lane < 8u) em = ebits[(size_t)k * 8u + lane];
    if (cnt != 0u) wk.fetch(k, wk.b);
    if (k + G < wk.nb) {
      wk.bn = oblk[k + G];
      sn = pstart[k + G];
      cn = bcount[k + G];
      if (lane < 8u) emn = ebits[(size_t)(k + G) * 8u + lane];
    }
  }
  for (; k < wk.nb; k += G) {
    if (start >= a.max_points) break;  // starts do not decrease with k
    const int4 b = wk.b;
    const uint32_t s0 = start, c0 = cnt;
    if (c0 != 0u) {
      // emitting voxels in voxel order
      uint32_t n = 0;
#pragma unroll
      for (int z = 0; z < 8; ++z) {
        const uint64_t m = __shfl(em, z, 64);
        if ((m >> lane) & 1ull) eb.list[n + (uint32_t)__popcll(m & below)] = (uint16_t)(lane + 64u * (uint32_t)z);
        n += (uint32_t)__popcll(m);
      }
      wk.stage(eb, a.voxel);
      __syncthreads();
    }
    if (k + G < wk.nb && cn != 0u) wk.fetch(k + G, wk.bn);
    wk.b = wk.bn;
    start = sn;
    cnt = cn;
    em = emn;
    const uint32_t k2 = k + 2u * G;
    if (k2 < wk.nb) {
      wk.bn = oblk[k2];
      sn = pstart[k2];
      cn = bcount[k2];
      if (lane < 8u) emn = ebits[(size_t)k2 * 8u + lane];
    }
    if (c0 == 0u) continue;
    for (uint32_t j0 = 0; j0 < c0; j0 += 64u) {
      const uint32_t j = j0 + lane;
      const uint64_t idx = (uint64_t)s0 + j;
      if (j < c0 && idx < a.max_points) {
        const int l = eb.list[j];
        const uint32_t c = col[(size_t)b.w * kVox + (size_t)l];
        const float3 iso = ext_iso<F16>(t, vox, eb, b, a.voxel, l);
        const float3 nrm = ext_normal<F16>(t, vox, eb, b, iso, a.voxel);
        out_pos[idx] = make_float4(iso.x, iso.y, iso.z, 1.0f);
        out_col[idx] = make_float4(c255[c & 0xFFu], c255[(c >> 8) & 0xFFu], c255[(c >> 16) & 0xFFu], 1.0f);
        out_nrm[idx] = make_float4(nrm.x, nrm.y, nrm.z, 0.0f);
      }
    }
    __syncthreads();  // eb is read above until every lane is done
  }
}

// ---------------------------------------------------------------------------------
// Marching cubes (MarchingCubes.cu:313-467, runMarchingCubesCUDA :511-571)
// ---------------------------------------------------------------------------------

// One block's cube corners in LDS (one wave per block): SDF and usability of the voxels
// at local (x, y, z) in 0..8 (x fastest, 9^3), the ninth layer from the +x / +y / +z
// neighbours; a corner is usable when its block is allocated and its weight is at
// least min_weight; and the triangle count of every case.
constexpr int kMcG = kBlock + 1;
constexpr int kMcGrid = kMcG * kMcG * kMcG;
struct McBlock {
  float sd[kMcGrid];
  uint8_t ok[kMcGrid];
  uint8_t ntri[256];
};

// Stages block k's corners: its own voxels coalesced (lane L, voxel L + 64 z), the
// ninth layer (cross-block cubes only) through the neighbour pointers.
template <bool F16>
__device__ __forceinline__ void mc_stage(McBlock& g, const typename Store<F16>::word* __restrict__ vw,
                                         const int32_t* __restrict__ nbp, uint32_t k, int4 b, const McArgs& a) {
  using S = Store<F16>;
  const uint32_t lane = threadIdx.x;
  const size_t base = (size_t)b.w * kVox + lane;
  typename S::word w[8];
#pragma unroll
  for (int z = 0; z < 8; ++z) w[z] = vw[base + 64u * (uint32_t)z];
#pragma unroll
  for (int z = 0; z < 8; ++z) {
    const uint32_t l = lane + 64u * (uint32_t)z;
    const uint32_t i = (l & 7u) + kMcG * ((l >> 3) & 7u) + kMcG * kMcG * (l >> 6);
    g.sd[i] = S::sdf(w[z]);
    g.ok[i] = S::weight(w[z]) >= a.min_weight ? 1 : 0;
  }
  if (a.cross) {
    // the 217 points of the ninth layer: x = 8 (81), then y = 8 with x < 8 (72), then
    // z = 8 with x, y < 8 (64)
    for (uint32_t j = lane; j < (uint32_t)(kMcGrid - kVox); j += 64u) {
      uint32_t x, y, z;
      if (j < 81u) {
        x = 8u, y = j % 9u, z = j / 9u;
      } else if (j < 153u) {
        x = (j - 81u) % 8u, y = 8u, z = (j - 81u) / 8u;
      } else {
        x = (j - 153u) % 8u, y = (j - 153u) / 8u, z = 8u;
      }
      const uint32_t i = x + kMcG * y + kMcG * kMcG * z;
      const int32_t q = nbp[(size_t)k * kExtNb + 13u + (x >> 3) + 3u * (y >> 3) + 9u * (z >> 3)];
      if (q < 0) {
        g.sd[i] = 0.0f;
        g.ok[i] = 0;
        continue;
      }
      const typename S::word v = vw[(size_t)q * kVox + (x & 7u) + kBlock * (y & 7u) + kBlock * kBlock * (z & 7u)];
      g.sd[i] = S::sdf(v);
      g.ok[i] = S::weight(v) >= a.min_weight ? 1 : 0;
    }
  }
}

// Case of cube l = x + 8 y + 64 z of the staged block (bit c: corner c = (c & 1,
// c >> 1 & 1, c >> 2 & 1) has sdf < iso), or -1 when the cube is not marched: beyond
// local 6 without cross-block cubes, or a corner not usable.
__device__ __forceinline__ int mc_case(const McBlock& g, uint32_t l, const McArgs& a) {
  const uint32_t x = l & 7u, y = (l >> 3) & 7u, z = l >> 6;
  if (!a.cross && (x == 7u || y == 7u || z == 7u)) return -1;
  const uint32_t i0 = x + kMcG * y + kMcG * kMcG * z;
  int k = 0;
  bool ok = true;
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    const uint32_t i = i0 + (uint32_t)(c & 1) + kMcG * (uint32_t)((c >> 1) & 1) + kMcG * kMcG * (uint32_t)(c >> 2);
    ok = ok && g.ok[i] != 0;
    k |= (g.sd[i] < a.iso ? 1 : 0) << c;
  }
  return ok ? k : -1;
}

// Triangles per block (bcount; zero for the block slots beyond the live blocks).
template <bool F16>
__global__ __launch_bounds__(64) void k_mc_count(const DevState* __restrict__ st, uint32_t num_blocks,
                                                 const int4* __restrict__ oblk, const int32_t* __restrict__ nbp,
                                                 const void* __restrict__ vox, const int8_t* __restrict__ table,
                                                 McArgs a, uint32_t* __restrict__ bcount) {
  __shared__ McBlock g;
  const uint32_t lane = threadIdx.x, G = gridDim.x, nb = st->n_alloc;
  for (uint32_t j = lane; j < 256u; j += 64u) {
    int n = 0;
    while (n < 10 && table[j * kMcRow + 3u * (uint32_t)n] >= 0) ++n;
    g.ntri[j] = (uint8_t)n;
  }
  const auto* vw = reinterpret_cast<const typename Store<F16>::word*>(vox);
  for (uint32_t k = blockIdx.x; k < nb; k += G) {
    __syncthreads();  // g is read below until every lane is done
    mc_stage<F16>(g, vw, nbp, k, oblk[k], a);
    __syncthreads();
    uint32_t c = 0;
#pragma unroll
    for (int z = 0; z < 8; ++z) {
      const int cs = mc_case(g, lane + 64u * (uint32_t)z, a);
      c += cs >= 0 ? g.ntri[cs] : 0u;
    }
    c = wave_sum(c);
    if (lane == 0) bcount[k] = c;
  }
  for (uint32_t j = nb + blockIdx.x * 64u + lane; j < num_blocks; j += G * 64u) bcount[j] = 0u;
}

// Corner position along one axis: voxel centre (8 block + i + 0.5) v of local grid
// coordinate gi (0..8) of block coordinate bc, taken in the block that holds the voxel,
// so that neighbouring blocks compute the same position.
__device__ __forceinline__ float mc_coord(int32_t bc, uint32_t gi, float v) {
  return ext_centre(gi == 8u ? bc + 1 : bc, (int)(gi & 7u), v);
}

// vertexInterp (MarchingCubes.cu:313-327) on edge e of cube (x, y, z): from its lower
// corner a to its upper corner b (the edge's axis e >> 2; corners a = the two other
// coordinates' bits e & 3 with a zero inserted at the axis, b = a + that axis' bit).
__device__ __forceinline__ float3 mc_vertex(const McBlock& g, int4 b, uint32_t x, uint32_t y, uint32_t z, int e,
                                            const McArgs& a) {
  const int axis = e >> 2, j = e & 3;
  const int ca = ((j >> axis) << (axis + 1)) | (j & ((1 << axis) - 1));
  const int cb = ca | (1 << axis);
  const uint32_t ax = x + (uint32_t)(ca & 1), ay = y + (uint32_t)((ca >> 1) & 1), az = z + (uint32_t)(ca >> 2);
  const uint32_t bx = x + (uint32_t)(cb & 1), by = y + (uint32_t)((cb >> 1) & 1), bz = z + (uint32_t)(cb >> 2);
  const float sa = g.sd[ax + kMcG * ay + kMcG * kMcG * az], sb = g.sd[bx + kMcG * by + kMcG * kMcG * bz];
  const float p1x = mc_coord(b.x, ax, a.voxel), p1y = mc_coord(b.y, ay, a.voxel), p1z = mc_coord(b.z, az, a.voxel);
  const float p2x = mc_coord(b.x, bx, a.voxel), p2y = mc_coord(b.y, by, a.voxel), p2z = mc_coord(b.z, bz, a.voxel);
  const float den = sb - sa;
  if (fabsf(den) < 1e-6f) return make_float3(0.5f * (p1x + p2x), 0.5f * (p1y + p2y), 0.5f * (p1z + p2z));
  float mu = (a.iso - sa) / den;
  mu = fminf(fmaxf(mu, 0.0f), 1.0f);
  return make_float3(p1x + mu * (p2x - p1x), p1y + mu * (p2y - p1y), p1z + mu * (p2z - p1z));
}

// The triangles: block k's start pstart[k], then cube order, then table order; three
// vertices and three copies of the face normal (computeNormal, MarchingCubes.cu:329-348)
// per triangle; the first max_tris kept.  The block's triangles are spread evenly over
// the lanes (lane L takes triangles L, L + 64, ... of the block, finding its cube by a
// binary search of the cubes' inclusive prefix sums), so neighbouring lanes write
// neighbouring triangles.
template <bool F16>
__global__ __launch_bounds__(64) void k_mc_write(const DevState* __restrict__ st, const int4* __restrict__ oblk,
                                                 const int32_t* __restrict__ nbp, const void* __restrict__ vox,
                                                 const int8_t* __restrict__ table, McArgs a,
                                                 const uint32_t* __restrict__ bcount,
                                                 const uint32_t* __restrict__ pstart, float3* __restrict__ out_v,
                                                 float3* __restrict__ out_n) {
  __shared__ McBlock g;
  __shared__ uint32_t incl[kVox];  // triangles of cubes 0..l of the block
  __shared__ uint8_t cases[kVox];
  const uint32_t lane = threadIdx.x, G = gridDim.x, nb = st->n_alloc;
  for (uint32_t j = lane; j < 256u; j += 64u) {
    int n = 0;
    while (n < 10 && table[j * kMcRow + 3u * (uint32_t)n] >= 0) ++n;
    g.ntri[j] = (uint8_t)n;
  }
  const auto* vw = reinterpret_cast<const typename Store<F16>::word*>(vox);
  for (uint32_t k = blockIdx.x; k < nb; k += G) {
    const uint32_t start = pstart[k];
    if (start >= a.max_tris) break;  // starts do not decrease with k
    if (bcount[k] == 0u) continue;
    const int4 b = oblk[k];
    __syncthreads();  // g, incl and cases are read below until every lane is done
    mc_stage<F16>(g, vw, nbp, k, b, a);
    __syncthreads();
    uint32_t run = 0;
#pragma unroll
    for (int z = 0; z < 8; ++z) {
      const uint32_t l = lane + 64u * (uint32_t)z;
      const int cs = mc_case(g, l, a);
      const uint32_t c = cs >= 0 ? g.ntri[cs] : 0u;
      const uint32_t inc = wave_incl_scan(c);
      incl[l] = run + inc;
      cases[l] = (uint8_t)(cs >= 0 ? cs : 0);
      run += __shfl(inc, 63, 64);
    }
    __syncthreads();
    const uint32_t total = run;
    for (uint32_t j = lane; j < total; j += 64u) {
      const uint32_t o = start + j;
      if (o >= a.max_tris) break;
      // cube of triangle j: the first l with incl[l] > j
      uint32_t lo = 0, hi = kVox - 1u;
      while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (incl[mid] > j) hi = mid;
        else lo = mid + 1u;
      }
      const uint32_t l = lo, cs = cases[l];
      const uint32_t t = j - (incl[l] - g.ntri[cs]);
      const uint32_t x = l & 7u, y = (l >> 3) & 7u, zz = l >> 6;
      const int8_t* row = table + (size_t)cs * kMcRow + 3u * t;
      const float3 v0 = mc_vertex(g, b, x, y, zz, row[0], a);
      const float3 v1 = mc_vertex(g, b, x, y, zz, row[1], a);
      const float3 v2 = mc_vertex(g, b, x, y, zz, row[2], a);
      const float ax = v1.x - v0.x, ay = v1.y - v0.y, az = v1.z - v0.z;
      const float bx = v2.x - v0.x, by = v2.y - v0.y, bz = v2.z - v0.z;
      float nx = ay * bz - az * by, ny = az * bx - ax * bz, nz = ax * by - ay * bx;
      const float len = sqrtf(nx * nx + ny * ny + nz * nz);
      if (len > 1e-6f) {
        const float inv = 1.0f / len;
        nx *= inv;
        ny *= inv;
        nz *= inv;
      } else {
        nx = ny = nz = 0.0f;
      }
      const float3 n = make_float3(nx, ny, nz);
      out_v[3u * o] = v0;
      out_v[3u * o + 1u] = v1;
      out_v[3u * o + 2u] = v2;
      out_n[3u * o] = n;
      out_n[3u * o + 1u] = n;
      out_n[3u * o + 2u] = n;
    }
  }
}

// ---------------------------------------------------------------------------------
// Depth maps from point clouds (CustomDepthMapGenerator.cpp:81-222)
// ---------------------------------------------------------------------------------

// Pixel of camera-frame point (x, y, z), or -1 (CustomDepthMapGenerator.cpp:110-140 with
// projectPoint :550-578 and isPointInBounds :580-587): points with z <= 0 are skipped;
// perspective u = x fx / z + cx (orthographic u = x + cx); in bounds when 0 <= u < W and
// 0 <= v < H; pixel (floor u, floor v).
__device__ __forceinline__ int32_t dg_pixel(const DepthGenArgs& g, float x, float y, float z) {
  if (z <= 0.0f) return -1;
  float u, v;
  if (g.orthographic) {
    u = x + g.cx;
    v = y + g.cy;
  } else {
    u = (x * g.fx / z) + g.cx;
    v = (y * g.fy / z) + g.cy;
  }
  if (!(u >= 0.0f && u < (float)g.W && v >= 0.0f && v < (float)g.H)) return -1;
  const int32_t px = cvt_i32(floorf(u)), py = cvt_i32(floorf(v));
  if (px < 0 || py < 0 || px >= g.W || py >= g.H) return -1;
  return py * g.W + px;
}

__global__ __launch_bounds__(256) void k_dg_clear(uint32_t* __restrict__ win, uint32_t n) {
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  if (i < n) win[i] = 0u;
}

// Last writer in point order: the largest point index (+1) per pixel; per pixel also
// the largest index below n_normals / n_colors when those arrays are shorter.
__global__ __launch_bounds__(256) void k_dg_claim(DepthGenArgs g, const float* __restrict__ pts,
                                                  uint32_t* __restrict__ win, uint32_t* __restrict__ win_n,
                                                  uint32_t* __restrict__ win_c) {
  for (uint32_t i = blockIdx.x * 256u + threadIdx.x; i < g.n; i += gridDim.x * 256u) {
    const int32_t p = dg_pixel(g, pts[3 * (size_t)i], pts[3 * (size_t)i + 1], pts[3 * (size_t)i + 2]);
    if (p < 0) continue;
    atomicMax(&win[p], i + 1u);
    if (win_n && i < g.n_normals) atomicMax(&win_n[p], i + 1u);
    if (win_c && i < g.n_colors) atomicMax(&win_c[p], i + 1u);
  }
}

// Every point that won its pixel writes it: the point, its z, and its normal / colour
// when it also won the pixel's normal / colour (CustomDepthMapGenerator.cpp:143-161).
// The points are read in order; the writes scatter.
__global__ __launch_bounds__(256) void k_dg_scatter(DepthGenArgs g, const float* __restrict__ pts,
                                                    const float* __restrict__ nrm, const uint8_t* __restrict__ rgb,
                                                    const uint32_t* __restrict__ win,
                                                    const uint32_t* __restrict__ win_n,
                                                    const uint32_t* __restrict__ win_c, float* __restrict__ out_pts,
                                                    float* __restrict__ out_depth, float* __restrict__ out_nrm,
                                                    uint8_t* __restrict__ out_rgb) {
  for (uint32_t i = blockIdx.x * 256u + threadIdx.x; i < g.n; i += gridDim.x * 256u) {
    const float x = pts[3 * (size_t)i], y = pts[3 * (size_t)i + 1], z = pts[3 * (size_t)i + 2];
    const int32_t p = dg_pixel(g, x, y, z);
    if (p < 0) continue;
    const size_t q = (size_t)p;
    if (win[p] == i + 1u) {
      out_pts[3 * q] = x;
      out_pts[3 * q + 1] = y;
      out_pts[3 * q + 2] = z;
      if (out_depth) out_depth[p] = z;
    }
    if (i < g.n_normals && (win_n ? win_n[p] : win[p]) == i + 1u) {
      out_nrm[3 * q] = nrm[3 * (size_t)i];
      out_nrm[3 * q + 1] = nrm[3 * (size_t)i + 1];
      out_nrm[3 * q + 2] = nrm[3 * (size_t)i + 2];
    }
    if (i < g.n_colors && (win_c ? win_c[p] : win[p]) == i + 1u) {
      out_rgb[3 * q] = rgb[3 * (size_t)i];
      out_rgb[3 * q + 1] = rgb[3 * (size_t)i + 1];
      out_rgb[3 * q + 2] = rgb[3 * (size_t)i + 2];
    }
  }
}

// Zeros where no point (no normal, no colour) landed (CustomDepthMapGenerator.cpp:103-107).
__global__ __launch_bounds__(256) void k_dg_fill(DepthGenArgs g, const uint32_t* __restrict__ win,
                                                 const uint32_t* __restrict__ win_n,
                                                 const uint32_t* __restrict__ win_c, float* __restrict__ out_pts,
                                                 float* __restrict__ out_depth, float* __restrict__ out_nrm,
                                                 uint8_t* __restrict__ out_rgb) {
  const uint32_t p = blockIdx.x * 256u + threadIdx.x;
  const uint32_t npix = (uint32_t)g.W * (uint32_t)g.H;
  if (p < npix) {
    const size_t q = p;
    const uint32_t w = win[p];
    if (w == 0u) {
      out_pts[3 * q] = 0.0f;
      out_pts[3 * q + 1] = 0.0f;
      out_pts[3 * q + 2] = 0.0f;
      if (out_depth) out_depth[p] = 0.0f;
    }
    const uint32_t wn = g.n_normals == 0u ? 0u : (win_n ? win_n[p] : w);
    if (wn == 0u) {
      out_nrm[3 * q] = 0.0f;
      out_nrm[3 * q + 1] = 0.0f;
      out_nrm[3 * q + 2] = 0.0f;
    }
    const uint32_t wc = g.n_colors == 0u ? 0u : (win_c ? win_c[p] : w);
    if (wc == 0u) {
      out_rgb[3 * q] = 0;
      out_rgb[3 * q + 1] = 0;
      out_rgb[3 * q + 2] = 0;
    }
  }
}

// Pixels with depth > 0 (CustomDepthMapGenerator.cpp:211-216): those with a winner
// (every winner has z > 0).  *valid starts at 0; one addition per workgroup.
__global__ __launch_bounds__(1024) void k_dg_count(const uint32_t* __restrict__ win, uint32_t npix,
                                                   uint32_t* __restrict__ valid) {
  uint32_t c = 0;
  for (uint32_t p = blockIdx.x * 1024u + threadIdx.x; p < npix; p += gridDim.x * 1024u) c += win[p] != 0u ? 1u : 0u;
  uint32_t total;
  (void)block_excl_scan<1024>(c, &total);
  if (threadIdx.x == 0 && total != 0u) atomicAdd(valid, total);
}

// Finite-difference normals when the cloud has none (CustomDepthMapGenerator.cpp:164-208):
// interior pixels with depth > 0 get normalize((right - centre) x (down - centre)) of
// the back-projected depths (the neighbours' depths as they are, 0 included), scaled by
// 1 / length when the length is positive; other pixels keep 0.
__global__ __launch_bounds__(256) void k_dg_fd_normals(DepthGenArgs g, const float* __restrict__ pmap,
                                                       float* __restrict__ out_nrm) {
  const uint32_t p = blockIdx.x * 256u + threadIdx.x;
  const int32_t W = g.W, H = g.H;
  if (p >= (uint32_t)W * (uint32_t)H) return;
  const int32_t x = (int32_t)(p % (uint32_t)W), y = (int32_t)(p / (uint32_t)W);
  if (x < 1 || y < 1 || x >= W - 1 || y >= H - 1) return;
  const float dc = pmap[3 * (size_t)p + 2];
  if (!(dc > 0.0f)) return;
  const float dr = pmap[3 * (size_t)(p + 1u) + 2], dd = pmap[3 * (size_t)(p + (uint32_t)W) + 2];
  const float fx = g.fx, fy = g.fy, cx = g.cx, cy = g.cy;
  const float ccx = ((float)x - cx) * dc / fx, ccy = ((float)y - cy) * dc / fy, ccz = dc;
  const float rx = ((float)(x + 1) - cx) * dr / fx, ry = ((float)y - cy) * dr / fy, rz = dr;
  const float ddx = ((float)x - cx) * dd / fx, ddy = ((float)(y + 1) - cy) * dd / fy, ddz = dd;
  const float ax = rx - ccx, ay = ry - ccy, az = rz - ccz;
  const float bx = ddx - ccx, by = ddy - ccy, bz = ddz - ccz;
  const float nx = ay * bz - az * by, ny = az * bx - ax * bz, nz = ax * by - ay * bx;
  const float len = sqrtf(nx * nx + ny * ny + nz * nz);
  if (len > 0.0f) {
    const float inv = 1.0f / len;
    out_nrm[3 * (size_t)p] = nx * inv;
    out_nrm[3 * (size_t)p + 1] = ny * inv;
    out_nrm[3 * (size_t)p + 2] = nz * inv;
  }
}

__global__ __launch_bounds__(256) void k_convert(const float* __restrict__ in, int32_t* __restrict__ out, uint32_t n) {
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  if (i < n) out[i] = cvt_i32(in[i]);
}

// ---------------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------------

void launch_reset_voxels(hipStream_t s, DevState* st, void* vox, uint32_t word_bytes, uint32_t* col,
                         uint32_t* dmeta, float2* dval, uint32_t* pcnt) {
  hipLaunchKernelGGL(k_reset_voxels, dim3(2048), dim3(256), 0, s, st, vox, word_bytes, col, dmeta, dval, pcnt);
}

void launch_cell_masks(hipStream_t s, bool fp16, const TableView& t, const DevState* st, uint32_t first, uint32_t n,
                       const void* vox) {
  if (n == 0u) return;
  const uint32_t grid = n < 2048u ? n : 2048u;
  if (fp16) hipLaunchKernelGGL(k_cell_masks<true>, dim3(grid), dim3(512), 0, s, t, st, first, n, vox);
  else hipLaunchKernelGGL(k_cell_masks<false>, dim3(grid), dim3(512), 0, s, t, st, first, n, vox);
}

void launch_export_cell_masks(hipStream_t s, const TableView& t, uint32_t first, uint32_t n, uint8_t* out) {
  if (n == 0u) return;
  hipLaunchKernelGGL(k_export_cell_masks, dim3((n + 255u) / 256u), dim3(256), 0, s, t, first, n, out);
}

void launch_reset_table(hipStream_t s, const TableView& t, DevState* st) {
  hipLaunchKernelGGL(k_reset_table, dim3(1024), dim3(256), 0, s, t, st);
}

void launch_alloc_claim(hipStream_t s, const TableView& t, const FrameArgs& fa, const float* pts, uint32_t* pixcnt,
                        uint8_t* mark_bkt, uint32_t* part) {
  const uint32_t nwg = (uint32_t)((fa.W * fa.H + 255) / 256);
  hipLaunchKernelGGL(k_alloc_claim, dim3(nwg), dim3(256), 0, s, t, fa, pts, pixcnt, mark_bkt, part);
}

void launch_alloc_resolve(hipStream_t s, const TableView& t, const FrameArgs& fa, const float* pts, uint32_t* pixcnt,
                          uint32_t* wgsum) {
  const uint32_t nwg = (uint32_t)((fa.W * fa.H + 255) / 256);
  hipLaunchKernelGGL(k_alloc_resolve, dim3(nwg), dim3(256), 0, s, t, fa, pts, pixcnt, wgsum);
}

void launch_alloc_insert(hipStream_t s, const TableView& t, DevState* st, const FrameArgs& fa, const float* pts,
                         const uint32_t* pixcnt, const uint32_t* wgsum, const uint32_t* part) {
  const uint32_t nwg = (uint32_t)((fa.W * fa.H + 255) / 256);
  hipLaunchKernelGGL(k_alloc_insert, dim3(nwg), dim3(256), 0, s, t, st, fa, pts, pixcnt, wgsum, part, nwg);
}

void launch_compact(hipStream_t s, const TableView& t, DevState* st, const FrameArgs& fa, uint32_t* vis,
                    uint32_t grid) {
  hipLaunchKernelGGL(k_compact, dim3(grid), dim3(256), 0, s, t, st, fa, vis);
}

void launch_integrate(hipStream_t s, bool fp16, const TableView& t, DevState* st, const FrameArgs& fa,
                      const float* pts, const uint8_t* rgb, const float* nrm, const uint32_t* vis, void* vox,
                      uint32_t* col, uint32_t* dmeta, float2* dval, uint32_t grid, uint32_t* upd_part) {
#define VSG_INTEGRATE(F, D)                                                                                       \
  hipLaunchKernelGGL((k_integrate<F, D>), dim3(grid), dim3(512), 0, s, t, st, fa, pts, rgb, nrm, vis, vox, col, dmeta, \
                     dval, upd_part)
  const bool dir = dmeta != nullptr && nrm != nullptr;
  if (fp16) {
    if (dir) VSG_INTEGRATE(true, true); else VSG_INTEGRATE(true, false);
  } else {
    if (dir) VSG_INTEGRATE(false, true); else VSG_INTEGRATE(false, false);
  }
#undef VSG_INTEGRATE
  hipLaunchKernelGGL(k_update_masks, dim3(256), dim3(256), 0, s, t, st, fa.cset, vis);
}

void launch_gn_jacobian(hipStream_t s, bool fp16, const TableView& t, const DevState* st, const RenderArgs& ra,
                        const void* vox, const uint32_t* col, const float* depth, float* jac) {
  const dim3 grid((ra.W + 7) / 8, (ra.H + 7) / 8);
  const bool ref = ra.mode == 0;
  const bool dev = ra.Mdev != nullptr;
#define VSG_GNJ(F, R)                                                                                      \
  do {                                                                                                     \
    if (dev)                                                                                               \
      hipLaunchKernelGGL((k_gn_jacobian<F, R, true>), grid, dim3(64), 0, s, t, st, ra, vox, col, depth, jac); \
    else                                                                                                   \
      hipLaunchKernelGGL((k_gn_jacobian<F, R, false>), grid, dim3(64), 0, s, t, st, ra, vox, col, depth, jac); \
  } while (0)
  if (fp16) {
    if (ref) VSG_GNJ(true, true); else VSG_GNJ(true, false);
  } else {
    if (ref) VSG_GNJ(false, true); else VSG_GNJ(false, false);
  }
#undef VSG_GNJ
}

void launch_render_adjoint(hipStream_t s, bool fp16, const TableView& t, const DevState* st, const RenderArgs& ra,
                           const void* vox, const uint32_t* col, float* jac, const float* grad_depth,
                           float* grad_sdf) {
  const dim3 grid((ra.W + 7) / 8, (ra.H + 7) / 8);
  const bool ref = ra.mode == 0;
  const bool dev = ra.Mdev != nullptr;
#define VSG_ADJ(F, R)                                                                                          \
  do {                                                                                                         \
    if (dev)                                                                                                   \
      hipLaunchKernelGGL((k_render_adjoint<F, R, true>), grid, dim3(64), 0, s, t, st, ra, vox, col, jac, grad_depth, \
                         grad_sdf);                                                                            \
    else                                                                                                       \
      hipLaunchKernelGGL((k_render_adjoint<F, R, false>), grid, dim3(64), 0, s, t, st, ra, vox, col, jac,        \
                         grad_depth, grad_sdf);                                                                \
  } while (0)
  if (fp16) {
    if (ref) VSG_ADJ(true, true); else VSG_ADJ(true, false);
  } else {
    if (ref) VSG_ADJ(false, true); else VSG_ADJ(false, false);
  }
#undef VSG_ADJ
}

uint32_t gn_partial_count(int npix) { return (uint32_t)((npix + 255) / 256); }

void launch_gn(hipStream_t s, const float* depth, const float* obs, const float* jac, int W, int H, float max_res,
               const float noise[3], double lambda, const double pose[16], const double* pose_in, double* partials,
               double* out, float* posef) {
  const uint32_t nwg = gn_partial_count(W * H);
  const float3 nz = make_float3(noise[0], noise[1], noise[2]);
  hipLaunchKernelGGL(k_gn_accum, dim3(nwg), dim3(256), 0, s, depth, obs, jac, W, H, max_res, nz, partials);
  Pose64 p;
  for (int i = 0; i < 16; ++i) p.T[i] = pose ? pose[i] : 0.0;
  hipLaunchKernelGGL(k_gn_solve, dim3(1), dim3(kGnSolveThreads), 0, s, partials, (int)nwg, lambda, p, pose_in, out,
                     posef);
}

uint32_t render_grid(int32_t W, int32_t H) {
  const uint32_t tiles_x = (uint32_t)((W + 7) / 8), tiles_y = (uint32_t)((H + 7) / 8);
  const uint32_t nst = ((tiles_x + kSuperTile - 1u) / kSuperTile) * ((tiles_y + kSuperTile - 1u) / kSuperTile);
  return 8u * ((nst + 7u) / 8u) * kSuperTile * kSuperTile;
}

uint32_t tail_workgroups(int32_t W, int32_t H) {
  // three waves per SIMD (the tail kernel's register budget): 256 CUs x 4 SIMDs x 3
  const uint64_t n = ((uint64_t)W * (uint64_t)H + 63u) / 64u;
  const uint32_t c = (uint32_t)(n > 3072u ? 3072u : n);
  return c < 8u ? 8u : (c + 7u) / 8u * 8u;
}

uint32_t tail_segment(int32_t W, int32_t H) {
  const uint32_t tiles_x = (uint32_t)((W + 7) / 8), tiles_y = (uint32_t)((H + 7) / 8);
  const uint32_t nst = ((tiles_x + kSuperTile - 1u) / kSuperTile) * ((tiles_y + kSuperTile - 1u) / kSuperTile);
  return (nst + 7u) / 8u * kSuperTile * kSuperTile * 64u;
}

template <class T> static void launch_scan(hipStream_t s, const T* in, uint32_t n, uint32_t* sums, uint32_t* out) {
  if (n == 0u) return;
  const uint32_t nchunks = (n + kScanChunk - 1u) / kScanChunk;
  hipLaunchKernelGGL(k_scan_sums<T>, dim3(nchunks), dim3(1024), 0, s, in, n, sums);
  hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(1024), 0, s, sums, nchunks);
  hipLaunchKernelGGL(k_scan_apply<T>, dim3(nchunks), dim3(1024), 0, s, in, n, sums, out);
}

// Workgroups of `threads` lanes of `kernel` the GPU holds at once (at least 1).
template <class K> static uint32_t resident_workgroups(K kernel, int threads) {
  int dev = 0, cus = 0, per = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
      hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kernel, threads, 0) != hipSuccess || cus * per < 1) {
    (void)hipGetLastError();
    return 1024u;
  }
  return (uint32_t)(cus * per);
}

void launch_extract(hipStream_t s, bool fp16, const TableView& t, const DevState* st, const void* vox,
                    const uint32_t* col, float voxel, float min_sdf, float max_sdf, uint32_t min_weight,
                    uint32_t max_points, const ExtractScratch& x, float4* pos, float4* colr, float4* nrm) {
  // blocks in hash-entry order, their neighbour pointers
  launch_scan<uint8_t>(s, t.cnt, t.num_buckets, x.sums, x.bucket_start);
  hipLaunchKernelGGL(k_ext_order, dim3((t.num_blocks + 255u) / 256u), dim3(256), 0, s, t, st, x.bucket_start, x.oblk);
  const uint64_t nthreads = (uint64_t)t.num_blocks * kExtNb;
  hipLaunchKernelGGL(k_ext_neighbours, dim3((uint32_t)((nthreads + 255u) / 256u)), dim3(256), 0, s, t, st, x.oblk,
                     x.nbp);
  const ExtractArgs a{voxel, min_sdf, max_sdf, min_weight, max_points};
  // one wave per block, as many waves as the GPU holds at once (the walk is grid-stride)
  auto grid_of = [&](uint32_t resident) { return dim3(t.num_blocks < resident ? t.num_blocks : resident); };
  if (fp16)
    hipLaunchKernelGGL(k_ext_count<true>, grid_of(resident_workgroups(k_ext_count<true>, 64)), dim3(64), 0, s, t, st,
                       x.oblk, x.nbp, vox, a, x.bcount, x.ebits);
  else
    hipLaunchKernelGGL(k_ext_count<false>, grid_of(resident_workgroups(k_ext_count<false>, 64)), dim3(64), 0, s, t,
                       st, x.oblk, x.nbp, vox, a, x.bcount, x.ebits);
  // block starts over all num_blocks counts (those beyond the live blocks are cleared)
  launch_scan<uint32_t>(s, x.bcount, t.num_blocks, x.sums, x.pstart);
  if (fp16)
    hipLaunchKernelGGL(k_ext_write<true>, grid_of(resident_workgroups(k_ext_write<true>, 64)), dim3(64), 0, s, t, st,
                       x.oblk, x.nbp, vox, col, a, x.bcount, x.ebits, x.pstart, pos, colr, nrm);
  else
    hipLaunchKernelGGL(k_ext_write<false>, grid_of(resident_workgroups(k_ext_write<false>, 64)), dim3(64), 0, s, t,
                       st, x.oblk, x.nbp, vox, col, a, x.bcount, x.ebits, x.pstart, pos, colr, nrm);
}

void launch_marching_cubes(hipStream_t s, bool fp16, const TableView& t, const DevState* st, const void* vox,
                           const int8_t* table, const McArgs& a, const ExtractScratch& x, float3* verts,
                           float3* nrms) {
  launch_scan<uint8_t>(s, t.cnt, t.num_buckets, x.sums, x.bucket_start);
  hipLaunchKernelGGL(k_ext_order, dim3((t.num_blocks + 255u) / 256u), dim3(256), 0, s, t, st, x.bucket_start, x.oblk);
  if (a.cross) {
    const uint64_t nthreads = (uint64_t)t.num_blocks * kExtNb;
    hipLaunchKernelGGL(k_ext_neighbours, dim3((uint32_t)((nthreads + 255u) / 256u)), dim3(256), 0, s, t, st, x.oblk,
                       x.nbp);
  }
  auto grid_of = [&](uint32_t resident) { return dim3(t.num_blocks < resident ? t.num_blocks : resident); };
  if (fp16)
    hipLaunchKernelGGL(k_mc_count<true>, grid_of(resident_workgroups(k_mc_count<true>, 64)), dim3(64), 0, s, st,
                       t.num_blocks, x.oblk, x.nbp, vox, table, a, x.bcount);
  else
    hipLaunchKernelGGL(k_mc_count<false>, grid_of(resident_workgroups(k_mc_count<false>, 64)), dim3(64), 0, s, st,
                       t.num_blocks, x.oblk, x.nbp, vox, table, a, x.bcount);
  launch_scan<uint32_t>(s, x.bcount, t.num_blocks, x.sums, x.pstart);
  if (fp16)
    hipLaunchKernelGGL(k_mc_write<true>, grid_of(resident_workgroups(k_mc_write<true>, 64)), dim3(64), 0, s, st,
                       x.oblk, x.nbp, vox, table, a, x.bcount, x.pstart, verts, nrms);
  else
    hipLaunchKernelGGL(k_mc_write<false>, grid_of(resident_workgroups(k_mc_write<false>, 64)), dim3(64), 0, s, st,
                       x.oblk, x.nbp, vox, table, a, x.bcount, x.pstart, verts, nrms);
}

void launch_depthgen_count(hipStream_t s, const uint32_t* win, uint32_t npix, uint32_t* valid) {
  hipLaunchKernelGGL(k_dg_clear, dim3(1), dim3(256), 0, s, valid, 1u);
  const uint32_t wgs = (npix + 1023u) / 1024u;
  hipLaunchKernelGGL(k_dg_count, dim3(wgs < 64u ? wgs : 64u), dim3(1024), 0, s, win, npix, valid);
}

void launch_depthgen(hipStream_t s, const DepthGenArgs& g, const float* pts, const float* nrm, const uint8_t* rgb,
                     uint32_t* win, uint32_t* win_n, uint32_t* win_c, float* out_pts, float* out_depth,
                     float* out_nrm, uint8_t* out_rgb) {
  const uint32_t npix = (uint32_t)g.W * (uint32_t)g.H;
  const dim3 pgrid((npix + 255u) / 256u);
  hipLaunchKernelGGL(k_dg_clear, pgrid, dim3(256), 0, s, win, npix);
  if (win_n) hipLaunchKernelGGL(k_dg_clear, pgrid, dim3(256), 0, s, win_n, npix);
  if (win_c) hipLaunchKernelGGL(k_dg_clear, pgrid, dim3(256), 0, s, win_c, npix);
  if (g.n != 0u) {
    const uint32_t blocks = (g.n + 255u) / 256u;
    hipLaunchKernelGGL(k_dg_claim, dim3(blocks < 65536u ? blocks : 65536u), dim3(256), 0, s, g, pts, win, win_n, win_c);
  }
  if (g.n != 0u) {
    const uint32_t blocks = (g.n + 255u) / 256u;
    hipLaunchKernelGGL(k_dg_scatter, dim3(blocks < 65536u ? blocks : 65536u), dim3(256), 0, s, g, pts, nrm, rgb, win,
                       win_n, win_c, out_pts, out_depth, out_nrm, out_rgb);
  }
  hipLaunchKernelGGL(k_dg_fill, pgrid, dim3(256), 0, s, g, win, win_n, win_c, out_pts, out_depth, out_nrm, out_rgb);
  if (g.n_normals == 0u) hipLaunchKernelGGL(k_dg_fd_normals, pgrid, dim3(256), 0, s, g, out_pts, out_nrm);
}

void launch_convert(hipStream_t s, const float* in, int32_t* out, uint32_t n) {
  if (n) hipLaunchKernelGGL(k_convert, dim3((n + 255u) / 256u), dim3(256), 0, s, in, out, n);
}

void launch_render(hipStream_t s, bool fp16, const TableView& t, const DevState* st, const RenderArgs& ra,
                   const void* vox, const uint32_t* col, float* depth, float4* pos, float4* colr, float4* nrm,
                   uint32_t* rc) {
  const uint32_t tiles_x = (uint32_t)((ra.W + 7) / 8), tiles_y = (uint32_t)((ra.H + 7) / 8);
  const uint32_t ntiles = tiles_x * tiles_y;
  const dim3 grid(render_grid(ra.W, ra.H));
  const bool diag = ra.mark_blk != nullptr;
  const bool ref = ra.mode == 0;
  uint32_t* rc_tail = rc + 2 * (size_t)ntiles;
  const dim3 tgrid(tail_workgroups(ra.W, ra.H));
  // device poses (ra.Mdev) are instantiated without diagnostics only: the caller reads a
  // device pose back into the by-value copy when diagnostics are on
  const bool dev = ra.Mdev != nullptr && !diag;
#define VSG_RENDER_DEV(F, D, R, DV)                                                                             \
  do {                                                                                                          \
    hipLaunchKernelGGL((k_render<F, D, R, DV>), grid, dim3(64), 0, s, t, st, ra, vox, col, depth, pos, colr,    \
                       nrm, rc);                                                                                \
    if (ra.budget != 0u)                                                                                        \
      hipLaunchKernelGGL((k_render_tail<F, D, R, DV>), tgrid, dim3(64), 0, s, t, ra, vox, col, depth, pos, colr,  \
                         nrm, rc_tail);                                                                         \
  } while (0)
#define VSG_RENDER(F, D, R)                       \
  do {                                            \
    if (D) VSG_RENDER_DEV(F, true, R, false);     \
    else if (dev) VSG_RENDER_DEV(F, false, R, true); \
    else VSG_RENDER_DEV(F, false, R, false);      \
  } while (0)
  if (fp16) {
    if (diag) {
      if (ref) VSG_RENDER(true, true, true); else VSG_RENDER(true, true, false);
    } else {
      if (ref) VSG_RENDER(true, false, true); else VSG_RENDER(true, false, false);
    }
  } else {
    if (diag) {
      if (ref) VSG_RENDER(false, true, true); else VSG_RENDER(false, true, false);
    } else {
      if (ref) VSG_RENDER(false, false, true); else VSG_RENDER(false, false, false);
    }
  }
#undef VSG_RENDER
#undef VSG_RENDER_DEV
}

__global__ void k_set_cam(float* __restrict__ cam, float x, float y, float z) {
  cam[0] = x;
  cam[1] = y;
  cam[2] = z;
}

void launch_set_cam(hipStream_t s, float* cam, const float c[3]) {
  hipLaunchKernelGGL(k_set_cam, dim3(1), dim3(1), 0, s, cam, c[0], c[1], c[2]);
}

// ---------------------------------------------------------------------------------
// Streaming (VoxelStreamingManager.cu:36-91,258-358,443-719)
// ---------------------------------------------------------------------------------
//
// Live block pointers stay dense ([0, n_alloc)) through streaming, so every other
// kernel keeps its invariants.  Stream-out removes a set of blocks: their images are
// gathered, the kept blocks at pointers >= n_alloc - m move into the holes the
// removed ones leave below that bound (sources and destinations are disjoint), the
// bucket records drop the removed entries (keeping slot order) and take the moved
// pointers, the freed tail is zeroed, and the block index and occupancy maps are
// rebuilt from the remaining blocks.  Stream-in inserts blocks in rounds of one
// insertion per bucket, the smallest input index first, which is what a sequential
// loop over the inputs produces; pointers follow input order.

// Copies `bytes` (a multiple of 16) from src to dst with the workgroup's 256 lanes.
__device__ __forceinline__ void wg_copy(void* __restrict__ dst, const void* __restrict__ src, uint32_t bytes) {
  uint4* d = reinterpret_cast<uint4*>(dst);
  const uint4* s = reinterpret_cast<const uint4*>(src);
  for (uint32_t i = threadIdx.x; i < bytes / 16u; i += 256u) d[i] = s[i];
}

// Copies the planes of block p into image `img` (or back when to_image is false).
__device__ __forceinline__ void image_copy(const BlockPlanes& bp, uint32_t p, uint8_t* img, bool to_image) {
  const uint32_t vb = kVox * bp.word_bytes;
  uint8_t* planes[4] = {bp.vox + (size_t)p * vb, reinterpret_cast<uint8_t*>(bp.col + (size_t)p * kVox), nullptr,
                        nullptr};
  if (bp.dmeta) {
    planes[2] = reinterpret_cast<uint8_t*>(bp.dmeta + (size_t)p * kVox);
    planes[3] = reinterpret_cast<uint8_t*>(bp.dval + (size_t)p * kVox);
  }
  const uint32_t bytes[4] = {vb, kVox * 4u, kVox * 4u, kVox * 8u};
  const int nplanes = bp.dmeta ? 4 : 2;
  uint32_t off = 0;
  for (int k = 0; k < nplanes; ++k) {
    if (to_image) wg_copy(img + off, planes[k], bytes[k]);
    else wg_copy(planes[k], img + off, bytes[k]);
    off += bytes[k];
  }
}

// sel[p] = 1 when block p lies at distance >= radius from the centre, its world
// position being coordinate * 8 * voxel (SDFBlockToWorld, VoxelScene.cu:101-107;
// test as in streamOutFindBlocksKernel_OPTIMIZED, VoxelStreamingManager.cu:55-66).
__global__ __launch_bounds__(256) void k_so_select(const int4* __restrict__ blk, uint32_t n, float cx, float cy,
                                                   float cz, float radius, float voxel, uint8_t* __restrict__ sel) {
  const uint32_t p = blockIdx.x * 256u + threadIdx.x;
  if (p >= n) return;
  const int4 b = blk[p];
  const float dx = (float)(b.x * kBlock) * voxel - cx;
  const float dy = (float)(b.y * kBlock) * voxel - cy;
  const float dz = (float)(b.z * kBlock) * voxel - cz;
  sel[p] = sqrtf(dx * dx + dy * dy + dz * dz) >= radius ? 1 : 0;
}

// The first m selected blocks (pointer order) are removed: out[r] = pointer of the
// r-th; remap[p] = p for kept blocks below k = n - m, ~0 for removed ones (movers'
// entries are written by k_so_move).
__global__ __launch_bounds__(256) void k_so_list(uint32_t n, uint32_t k, uint32_t m, const uint8_t* __restrict__ sel,
                                                 const uint32_t* __restrict__ rank, uint32_t* __restrict__ out,
                                                 uint32_t* __restrict__ remap) {
  const uint32_t p = blockIdx.x * 256u + threadIdx.x;
  if (p >= n) return;
  const bool removed = sel[p] && rank[p] < m;
  if (removed) out[rank[p]] = p;
  if (removed) remap[p] = 0xFFFFFFFFu;
  else if (p < k) remap[p] = p;
}

// One workgroup per removed block: its coordinates and image.
__global__ __launch_bounds__(256) void k_so_gather(BlockPlanes bp, const int4* __restrict__ blk,
                                                   const uint32_t* __restrict__ out, int4* __restrict__ coords,
                                                   uint8_t* __restrict__ images, uint32_t image_bytes) {
  const uint32_t r = blockIdx.x, p = out[r];
  if (threadIdx.x == 0) coords[r] = blk[p];
  image_copy(bp, p, images + (size_t)r * image_bytes, true);
}

// One workgroup per pointer q in [k, n): a kept block there moves to the hole of
// its rank among the movers (the holes are the removed pointers below k, which are
// out[0 .. holes) in increasing order).
__global__ __launch_bounds__(256) void k_so_move(BlockPlanes bp, uint32_t* __restrict__ pcnt, uint32_t n, uint32_t k,
                                                 uint32_t m, const uint8_t* __restrict__ sel,
                                                 const uint32_t* __restrict__ rank, const uint32_t* __restrict__ out,
                                                 uint32_t* __restrict__ remap) {
  const uint32_t q = k + blockIdx.x;
  if (q >= n) return;
  const uint32_t rq = rank[q] < m ? rank[q] : m;  // removed blocks below q
  if (sel[q] && rank[q] < m) return;             // removed itself
  const uint32_t rk = rank[k] < m ? rank[k] : m;  // removed blocks below k (the holes)
  const uint32_t dst = out[(q - k) - (rq - rk)];
  if (threadIdx.x == 0) remap[q] = dst;
  if (threadIdx.x < 8) pcnt[(size_t)dst * 8u + threadIdx.x] = pcnt[(size_t)q * 8u + threadIdx.x];
  const uint32_t vb = kVox * bp.word_bytes;
  wg_copy(bp.vox + (size_t)dst * vb, bp.vox + (size_t)q * vb, vb);
  wg_copy(bp.col + (size_t)dst * kVox, bp.col + (size_t)q * kVox, kVox * 4u);
  if (bp.dmeta) {
    wg_copy(bp.dmeta + (size_t)dst * kVox, bp.dmeta + (size_t)q * kVox, kVox * 4u);
    wg_copy(bp.dval + (size_t)dst * kVox, bp.dval + (size_t)q * kVox, kVox * 8u);
  }
}

// Zeroes the planes and cell counts of blocks [k, n) (free again).
__global__ __launch_bounds__(256) void k_so_clear(BlockPlanes bp, uint32_t* __restrict__ pcnt, uint32_t k, uint32_t n) {
  const size_t tid = (size_t)blockIdx.x * 256 + threadIdx.x, stride = (size_t)gridDim.x * 256;
  const size_t v0 = (size_t)k * kVox, v1 = (size_t)n * kVox;
  for (size_t i = (size_t)k * 8u + tid; i < (size_t)n * 8u; i += stride) pcnt[i] = 0u;
  for (size_t i = v0 + tid; i < v1; i += stride) {
    if (bp.word_bytes == 8) reinterpret_cast<uint2*>(bp.vox)[i] = make_uint2(0u, 0u);
    else reinterpret_cast<uint32_t*>(bp.vox)[i] = 0u;
    bp.col[i] = 0u;
    if (bp.dmeta) {
      bp.dmeta[i] = 0u;
      bp.dval[i] = make_float2(0.0f, 0.0f);
    }
  }
}

// Per bucket: drops removed entries (the kept ones keep their order), renames moved
// pointers and rewrites blk[] of every kept entry with its new slot.
__global__ __launch_bounds__(256) void k_so_buckets(TableView t, const uint32_t* __restrict__ remap) {
  for (uint32_t h = blockIdx.x * 256u + threadIdx.x; h < t.num_buckets; h += gridDim.x * 256u) {
    const int c = t.cnt[h];
    if (c == 0) continue;
    uint64_t* r = t.rec + (size_t)h * kRecWords;
    int32_t* ptr = reinterpret_cast<int32_t*>(r + kBucket);
    int j2 = 0;
    for (int j = 0; j < c; ++j) {
      const uint64_t key = r[j];
      const int32_t p = ptr[j];
      const uint32_t np = remap[p];
      if (np == 0xFFFFFFFFu) continue;
      const int4 b = t.blk[p];
      r[j2] = key;
      ptr[j2] = (int32_t)np;
      t.blk[np] = make_int4(b.x, b.y, b.z, (int32_t)(h * kBucket + j2));
      ++j2;
    }
    t.cnt[h] = (uint8_t)j2;
  }
}

// Empties the block index and the occupancy maps.
__global__ __launch_bounds__(256) void k_clear_maps(TableView t) {
  const size_t tid = (size_t)blockIdx.x * 256 + threadIdx.x, stride = (size_t)gridDim.x * 256;
  for (size_t w = tid; w <= t.occ_mask; w += stride) t.occ[w] = 0ull;
  for (size_t w = tid; w <= t.hyp_mask; w += stride) t.hyp[w] = 0ull;
  for (size_t w = tid; w <= t.idx_mask; w += stride) t.idx[w] = make_uint4(kIdxEmpty, kIdxEmpty, 0u, 0u);
}

// Occupancy bits and a block index entry (pointer p, positive-cell mask `mask`) for
// block b: the same writes the allocation's insertion makes.
__device__ __forceinline__ void index_block(const TableView& t, int4 b, uint32_t p, uint32_t mask) {
  atomicOr(reinterpret_cast<unsigned long long*>(&t.occ[occ_index(b.x >> 2, b.y >> 2, b.z >> 2, t.occ_mask)]),
           1ull << occ_bit(b.x, b.y, b.z));
  atomicOr(reinterpret_cast<unsigned long long*>(&t.hyp[hyp_index(b.x >> 4, b.y >> 4, b.z >> 4, t.hyp_mask)]),
           1ull << hyp_bit(b.x >> 2, b.y >> 2, b.z >> 2));
  const uint64_t key = pack_key(b.x, b.y, b.z);
  uint32_t sl = idx_home(b.x, b.y, b.z, t.idx_mask);
  for (uint32_t n = 0; n <= t.idx_mask; ++n) {
    unsigned long long* kw = reinterpret_cast<unsigned long long*>(&t.idx[sl]);
    if (atomicCAS(kw, 0xFFFFFFFFFFFFFFFFull, (unsigned long long)key) == 0xFFFFFFFFFFFFFFFFull) {
      t.idx[sl].z = p;
      t.idx[sl].w = mask;
      t.idxpos[p] = sl;
      break;
    }
    sl = (sl + 1u) & t.idx_mask;
  }
}

// Re-indexes live blocks [0, n) (masks from their cell counts).
__global__ __launch_bounds__(256) void k_reindex(TableView t, uint32_t n) {
  const uint32_t p = blockIdx.x * 256u + threadIdx.x;
  if (p >= n) return;
  index_block(t, t.blk[p], p, cell_mask_of(t.pcnt + (size_t)p * 8u));
}

// Sets the live block count.
__global__ void k_set_alloc(DevState* st, uint32_t n) {
  st->n_alloc = n;
  st->n_alloc_prev = n;
  st->n_alloc_next = n;
}

// Stream-in status codes (per input block).
constexpr uint8_t kSiPending = 0xFF;

// Inputs out of the key range fail (4); the others start pending.
__global__ __launch_bounds__(256) void k_si_init(const int4* __restrict__ coords, uint32_t n,
                                                 uint8_t* __restrict__ status) {
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  if (i >= n) return;
  const int4 c = coords[i];
  status[i] = key_in_range(c.x, c.y, c.z) ? kSiPending : 4;
}

// One claim round: a pending input whose block is live is done (1), one whose home
// bucket is full fails (2); with no free block left (nospace) the others fail (3),
// else they bid for their bucket with their index.
__global__ __launch_bounds__(256) void k_si_claim(TableView t, const int4* __restrict__ coords, uint32_t n,
                                                  uint8_t* __restrict__ status, int32_t nospace) {
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  if (i >= n || status[i] != kSiPending) return;
  const int4 c = coords[i];
  if (table_lookup(t, c.x, c.y, c.z) >= 0) {
    status[i] = 1;
    return;
  }
  const uint32_t h = hash_bucket(c.x, c.y, c.z, t.num_buckets);
  if (t.cnt[h] >= kBucket) {
    status[i] = 2;
    return;
  }
  if (nospace) {
    status[i] = 3;
    return;
  }
  atomicMin(&t.claim[h], i);
}

// The round's winners take the next slot of their bucket (pointer set later);
// entry[i] = bucket * 10 + slot.
__global__ __launch_bounds__(256) void k_si_insert(TableView t, const int4* __restrict__ coords, uint32_t n,
                                                   uint8_t* __restrict__ status, uint32_t* __restrict__ entry) {
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  if (i >= n || status[i] != kSiPending) return;
  const int4 c = coords[i];
  const uint32_t h = hash_bucket(c.x, c.y, c.z, t.num_buckets);
  if (t.claim[h] != i) return;
  t.claim[h] = kNoClaim;
  const int slot = t.cnt[h];
  uint64_t* r = t.rec + (size_t)h * kRecWords;
  r[slot] = pack_key(c.x, c.y, c.z);
  reinterpret_cast<int32_t*>(r + kBucket)[slot] = 0;
  t.cnt[h] = (uint8_t)(slot + 1);
  entry[i] = h * kBucket + (uint32_t)slot;
  status[i] = 0;
}

// ok[i] = 1 for inserted inputs; inputs still pending after the last round fail (2).
__global__ __launch_bounds__(256) void k_si_flags(uint32_t n, uint8_t* __restrict__ status, uint8_t* __restrict__ ok) {
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  if (i >= n) return;
  if (status[i] == kSiPending) status[i] = 2;
  ok[i] = status[i] == 0 ? 1 : 0;
}

// One workgroup per input: an inserted block takes pointer n0 + its rank, its record
// pointer, blk entry, occupancy bits, index entry and planes (from its image).
__global__ __launch_bounds__(256) void k_si_place(TableView t, BlockPlanes bp, const int4* __restrict__ coords,
                                                  const uint8_t* __restrict__ ok, const uint32_t* __restrict__ rank,
                                                  const uint32_t* __restrict__ entry, uint32_t n0,
                                                  const uint8_t* __restrict__ images, uint32_t image_bytes) {
  const uint32_t i = blockIdx.x;
  if (!ok[i]) return;
  const uint32_t p = n0 + rank[i];
  if (threadIdx.x == 0) {
    const int4 c = coords[i];
    const uint32_t e = entry[i];
    reinterpret_cast<int32_t*>(t.rec + (size_t)(e / kBucket) * kRecWords + kBucket)[e % kBucket] = (int32_t)p;
    t.blk[p] = make_int4(c.x, c.y, c.z, (int32_t)e);
    index_block(t, c, p, 0u);
  }
  image_copy(bp, p, const_cast<uint8_t*>(images) + (size_t)i * image_bytes, false);
}

uint32_t block_image_bytes(const BlockPlanes& bp) { return kVox * (bp.word_bytes + 4u + (bp.dmeta ? 12u : 0u)); }

uint32_t stream_scan_words(uint32_t n) { return extract_sums_words(n); }

uint32_t stream_out_select(hipStream_t s, const TableView& t, uint32_t n, const float c[3], float radius, float voxel,
                           uint8_t* sel, uint32_t* rank, uint32_t* sums) {
  if (n == 0u) return 0u;
  hipLaunchKernelGGL(k_so_select, dim3((n + 255u) / 256u), dim3(256), 0, s, t.blk, n, c[0], c[1], c[2], radius, voxel,
                     sel);
  launch_scan(s, sel, n, sums, rank);
  uint32_t last_rank = 0;
  uint8_t last_sel = 0;
  if (hipMemcpyAsync(&last_rank, rank + (n - 1u), 4, hipMemcpyDeviceToHost, s) != hipSuccess) return 0xFFFFFFFFu;
  if (hipMemcpyAsync(&last_sel, sel + (n - 1u), 1, hipMemcpyDeviceToHost, s) != hipSuccess) return 0xFFFFFFFFu;
  if (hipStreamSynchronize(s) != hipSuccess) return 0xFFFFFFFFu;
  return last_rank + last_sel;
}

void stream_out_remove(hipStream_t s, const TableView& t, DevState* st, const BlockPlanes& bp, uint32_t n, uint32_t m,
                       const uint8_t* sel, const uint32_t* rank, uint32_t* out, uint32_t* remap, int4* coords,
                       uint8_t* images) {
  if (m == 0u) return;
  const uint32_t k = n - m;
  const uint32_t g = (n + 255u) / 256u;
  hipLaunchKernelGGL(k_so_list, dim3(g), dim3(256), 0, s, n, k, m, sel, rank, out, remap);
  hipLaunchKernelGGL(k_so_gather, dim3(m), dim3(256), 0, s, bp, t.blk, out, coords, images, block_image_bytes(bp));
  hipLaunchKernelGGL(k_so_move, dim3(m), dim3(256), 0, s, bp, t.pcnt, n, k, m, sel, rank, out, remap);
  hipLaunchKernelGGL(k_so_clear, dim3(m < 2048u ? m : 2048u), dim3(256), 0, s, bp, t.pcnt, k, n);
  const uint32_t gb = t.num_buckets / 256u + 1u;
  hipLaunchKernelGGL(k_so_buckets, dim3(gb < 4096u ? gb : 4096u), dim3(256), 0, s, t, remap);
  hipLaunchKernelGGL(k_clear_maps, dim3(2048), dim3(256), 0, s, t);
  hipLaunchKernelGGL(k_set_alloc, dim3(1), dim3(1), 0, s, st, k);
  if (k > 0u) hipLaunchKernelGGL(k_reindex, dim3((k + 255u) / 256u), dim3(256), 0, s, t, k);
}

uint32_t stream_in_insert(hipStream_t s, const TableView& t, uint32_t n, const int4* coords, uint8_t* status,
                          uint32_t* entry, uint8_t* ok, uint32_t* rank, uint32_t* sums, bool nospace) {
  if (n == 0u) return 0u;
  const uint32_t g = (n + 255u) / 256u;
  hipLaunchKernelGGL(k_si_init, dim3(g), dim3(256), 0, s, coords, n, status);
  // each round inserts one input per bucket; after kBucket + 1 rounds every input is
  // inserted, live, or facing a full bucket
  for (int round = 0; round <= kBucket; ++round) {
    hipLaunchKernelGGL(k_si_claim, dim3(g), dim3(256), 0, s, t, coords, n, status, (int32_t)nospace);
    if (nospace) break;
    hipLaunchKernelGGL(k_si_insert, dim3(g), dim3(256), 0, s, t, coords, n, status, entry);
  }
  hipLaunchKernelGGL(k_si_flags, dim3(g), dim3(256), 0, s, n, status, ok);
  launch_scan(s, ok, n, sums, rank);
  uint32_t last_rank = 0;
  uint8_t last_ok = 0;
  if (hipMemcpyAsync(&last_rank, rank + (n - 1u), 4, hipMemcpyDeviceToHost, s) != hipSuccess) return 0xFFFFFFFFu;
  if (hipMemcpyAsync(&last_ok, ok + (n - 1u), 1, hipMemcpyDeviceToHost, s) != hipSuccess) return 0xFFFFFFFFu;
  if (hipStreamSynchronize(s) != hipSuccess) return 0xFFFFFFFFu;
  return last_rank + last_ok;
}

void stream_in_place(hipStream_t s, bool fp16, const TableView& t, DevState* st, const BlockPlanes& bp, uint32_t n,
                     uint32_t n0, uint32_t inserted, const int4* coords, const uint8_t* ok, const uint32_t* rank,
                     const uint32_t* entry, const uint8_t* images) {
  if (inserted == 0u) return;
  hipLaunchKernelGGL(k_si_place, dim3(n), dim3(256), 0, s, t, bp, coords, ok, rank, entry, n0, images,
                     block_image_bytes(bp));
  hipLaunchKernelGGL(k_set_alloc, dim3(1), dim3(1), 0, s, st, n0 + inserted);
  launch_cell_masks(s, fp16, t, st, n0, inserted, bp.vox);
}

}  // namespace vsg
