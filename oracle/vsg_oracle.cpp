// C API over the VoxelSG oracle (vsg_oracle.hpp) for the Python test harness.
//
// TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and
// bench.py's cpu_baseline leg as the checker / CPU baseline, never by the
// product path.  Poses and intrinsics cross this API as double; the FP32
// oracle rounds them to float (exact for values that came from float).
#include <algorithm>
#include <cstdint>
#include <limits>
#include <utility>
#include <vector>

#include <omp.h>

#include "vsg_oracle.hpp"

using namespace vsgo;

namespace {

struct Handle {
  int fp64 = 0;
  Scene<float> sf;
  Scene<double> sd;
};

template <class R> void to_pose(const double* T, R* M) {
  for (int i = 0; i < 16; ++i) M[i] = (R)T[i];
}

template <class R>
void do_integrate(Scene<R>& s, const float* pts, const uint8_t* rgb, const float* nrm, int W, int H, double trunc,
                  const double* T, double fx, double fy, double cx, double cy, int visible_only, uint32_t* ctr) {
  R M[16];
  to_pose(T, M);
  integrate_frame(s, pts, rgb, W, H, (R)trunc, M, (R)fx, (R)fy, (R)cx, (R)cy, visible_only, nrm);
  if (ctr) {
    ctr[0] = s.c_valid_pixels;
    ctr[1] = s.c_new_blocks;
    ctr[2] = s.c_heap_fail;
    ctr[3] = (uint32_t)s.num_allocated();
    ctr[4] = s.c_visible;
    ctr[5] = s.c_updated;
    ctr[6] = s.heap_counter;
    ctr[7] = s.frames;
  }
}

template <class R>
int do_render(const Scene<R>& s, int W, int H, const double* cam3, const double* T, double fx, double fy, double cx,
              double cy, double minD, double maxD, double voxel, float* depth, float* pos, float* col, float* nrm,
              uint32_t* hm, const int32_t* pix_list, long n_list) {
  R M[16];
  to_pose(T, M);
  V3<R> cam{(R)cam3[0], (R)cam3[1], (R)cam3[2]};
  return render(s, W, H, cam, M, (R)fx, (R)fy, (R)cx, (R)cy, (R)minD, (R)maxD, (R)voxel, depth, pos, col, nrm, &hm[0],
                &hm[1], pix_list, n_list);
}

template <class R> uint64_t do_export_blocks(const Scene<R>& s, int32_t* coords, int32_t* ptrs, int32_t* slots,
                                             uint64_t cap) {
  std::vector<std::pair<int32_t, size_t>> live;
  for (size_t i = 0; i < s.slot_ptr.size(); ++i)
    if (s.slot_ptr[i] != -1) live.push_back({s.slot_ptr[i], i});
  std::sort(live.begin(), live.end());
  uint64_t n = std::min<uint64_t>(cap, live.size());
  for (uint64_t k = 0; k < n; ++k) {
    const I3& p = s.slot_pos[live[k].second];
    if (coords) {
      coords[3 * k] = p.x;
      coords[3 * k + 1] = p.y;
      coords[3 * k + 2] = p.z;
    }
    if (ptrs) ptrs[k] = live[k].first;
    if (slots) slots[k] = (int32_t)live[k].second;
  }
  return live.size();
}

template <class R> void do_export_voxels(const Scene<R>& s, uint64_t nblocks, double* sdf, uint8_t* w, uint8_t* c3,
                                         uint8_t* fl) {
  size_t have = s.weight.size();
  for (size_t i = 0; i < (size_t)nblocks * kVox; ++i) {
    bool ok = i < have;
    if (sdf) sdf[i] = ok ? (double)s.load_sdf(i) : 0.0;
    if (w) w[i] = ok ? s.weight[i] : 0;
    if (c3) {
      for (int ch = 0; ch < 3; ++ch) c3[3 * i + ch] = ok ? s.color[3 * i + ch] : 0;
    }
    if (fl) fl[i] = ok ? s.flags[i] : 0;
  }
}

}  // namespace

extern "C" {

void* vsgo_create(uint32_t num_buckets, uint32_t num_blocks, double voxel, int store_fp16, int fp64) {
  Handle* h = new Handle();
  h->fp64 = fp64;
  if (fp64) h->sd.init(num_buckets, num_blocks, voxel, store_fp16);
  else h->sf.init(num_buckets, num_blocks, (float)voxel, store_fp16);
  return h;
}

void vsgo_destroy(void* p) { delete (Handle*)p; }

void vsgo_reset(void* p) {
  Handle* h = (Handle*)p;
  if (h->fp64) h->sd.reset(); else h->sf.reset();
}

// counters: valid_pixels, new_blocks, heap_fail, num_allocated, visible, updated, heap_counter, frames.
// nrm (camera-frame normals, 3 floats per pixel) feeds the directional extension; may be null.
void vsgo_integrate(void* p, const float* pts, const uint8_t* rgb, const float* nrm, int W, int H, double trunc,
                    const double* T16, double fx, double fy, double cx, double cy, int visible_only,
                    uint32_t* counters) {
  Handle* h = (Handle*)p;
  if (h->fp64) do_integrate(h->sd, pts, rgb, nrm, W, H, trunc, T16, fx, fy, cx, cy, visible_only, counters);
  else do_integrate(h->sf, pts, rgb, nrm, W, H, trunc, T16, fx, fy, cx, cy, visible_only, counters);
}

// Seconds of the last integrate's allocation, compaction and integration (out3).
void vsgo_stage_times(void* p, double* out3) {
  Handle* h = (Handle*)p;
  if (h->fp64) {
    out3[0] = h->sd.t_alloc;
    out3[1] = h->sd.t_compact;
    out3[2] = h->sd.t_integrate;
  } else {
    out3[0] = h->sf.t_alloc;
    out3[1] = h->sf.t_compact;
    out3[2] = h->sf.t_integrate;
  }
}

// OpenMP threads of the parallel loops (integration over blocks, raycast over pixels);
// returns the number in effect.
int vsgo_set_threads(int n) {
  if (n > 0) omp_set_num_threads(n);
  return omp_get_max_threads();
}

// Directional extension on (before the first frame) or off.
void vsgo_set_directional(void* p, int on) {
  Handle* h = (Handle*)p;
  h->sd.directional = on;
  h->sf.directional = on;
}

// Direction slots of the first nblocks blocks (pointer order): 2 per voxel each.
void vsgo_export_directional(void* p, uint64_t nblocks, uint8_t* ids, uint8_t* ws, double* dts) {
  Handle* h = (Handle*)p;
  const size_t n = (size_t)nblocks * kVox * 2;
  for (size_t i = 0; i < n; ++i) {
    if (h->fp64) {
      const bool have = i < h->sd.dir_id.size();
      ids[i] = have ? h->sd.dir_id[i] : 0;
      ws[i] = have ? h->sd.dir_w[i] : 0;
      dts[i] = have ? h->sd.dtsdf[i] : 0.0;
    } else {
      const bool have = i < h->sf.dir_id.size();
      ids[i] = have ? h->sf.dir_id[i] : 0;
      ws[i] = have ? h->sf.dir_w[i] : 0;
      dts[i] = have ? (double)h->sf.dtsdf[i] : 0.0;
    }
  }
}

// Allocation step only (for allocation-level tests).
void vsgo_alloc(void* p, const float* pts, int W, int H, double trunc, const double* T16) {
  Handle* h = (Handle*)p;
  if (h->fp64) {
    double M[16];
    to_pose(T16, M);
    alloc_frame(h->sd, pts, W, H, trunc, M);
  } else {
    float M[16];
    to_pose(T16, M);
    alloc_frame(h->sf, pts, W, H, (float)trunc, M);
  }
}

// pix_list (optional, n_list entries): trace only these pixel indices.
int vsgo_render(void* p, int W, int H, const double* cam3, const double* T16, double fx, double fy, double cx,
                double cy, double minD, double maxD, double voxel, float* depth, float* pos, float* col, float* nrm,
                uint32_t* hits_misses, const int32_t* pix_list, long n_list) {
  Handle* h = (Handle*)p;
  if (h->fp64)
    return do_render(h->sd, W, H, cam3, T16, fx, fy, cx, cy, minD, maxD, voxel, depth, pos, col, nrm, hits_misses,
                     pix_list, n_list);
  return do_render(h->sf, W, H, cam3, T16, fx, fy, cx, cy, minD, maxD, voxel, depth, pos, col, nrm, hits_misses,
                   pix_list, n_list);
}

uint64_t vsgo_num_allocated(void* p) {
  Handle* h = (Handle*)p;
  return h->fp64 ? h->sd.num_allocated() : h->sf.num_allocated();
}

uint32_t vsgo_heap_counter(void* p) {
  Handle* h = (Handle*)p;
  return h->fp64 ? h->sd.heap_counter : h->sf.heap_counter;
}

uint64_t vsgo_export_blocks(void* p, int32_t* coords, int32_t* ptrs, int32_t* slots, uint64_t cap) {
  Handle* h = (Handle*)p;
  return h->fp64 ? do_export_blocks(h->sd, coords, ptrs, slots, cap) : do_export_blocks(h->sf, coords, ptrs, slots, cap);
}

void vsgo_export_voxels(void* p, uint64_t nblocks, double* sdf, uint8_t* weight, uint8_t* color3, uint8_t* flags) {
  Handle* h = (Handle*)p;
  if (h->fp64) do_export_voxels(h->sd, nblocks, sdf, weight, color3, flags);
  else do_export_voxels(h->sf, nblocks, sdf, weight, color3, flags);
}

// Surface points of the current store (positions, normals: 3 doubles each; colours:
// 3 bytes), the first `cap` in hash-entry order; returns the number found.
uint64_t vsgo_extract_surface_points(void* p, double min_sdf, double max_sdf, uint32_t min_weight, uint64_t cap,
                                     double* pos3, double* nrm3, uint8_t* col3) {
  Handle* h = (Handle*)p;
  auto run = [&](auto& sc, auto tag) -> uint64_t {
    using R = decltype(tag);
    std::vector<SurfacePoint<R>> pts;
    const uint64_t total = extract_surface_points(sc, (R)min_sdf, (R)max_sdf, min_weight, cap, pts);
    for (size_t i = 0; i < pts.size(); ++i) {
      for (int k = 0; k < 3; ++k) {
        if (pos3) pos3[3 * i + k] = (double)pts[i].p[k];
        if (nrm3) nrm3[3 * i + k] = (double)pts[i].n[k];
        if (col3) col3[3 * i + k] = pts[i].c[k];
      }
    }
    return total;
  };
  return h->fp64 ? run(h->sd, 0.0) : run(h->sf, 0.0f);
}

// Marching cubes over the scene with the given table (int8 [256][32]); vertices and
// normals: 9 values per triangle for the first cap triangles.  Returns the number found.
uint64_t vsgo_marching_cubes(void* p, double iso, double voxel, int32_t cross, uint32_t min_weight,
                             const int8_t* table, uint64_t cap, double* verts9, double* nrms9) {
  Handle* h = (Handle*)p;
  const int8_t(*tab)[32] = reinterpret_cast<const int8_t(*)[32]>(table);
  auto run = [&](auto& sc, auto tag) -> uint64_t {
    using R = decltype(tag);
    std::vector<R> vv, nn;
    const uint64_t total = marching_cubes(sc, (R)iso, (R)voxel, cross, min_weight, tab, cap, vv, nn);
    for (size_t i = 0; i < vv.size(); ++i) {
      if (verts9) verts9[i] = (double)vv[i];
      if (nrms9) nrms9[i] = (double)nn[i];
    }
    return total;
  };
  return h->fp64 ? run(h->sd, 0.0) : run(h->sf, 0.0f);
}

// ---- unit-level entry points (float arithmetic, as the reference kernels)
uint32_t vsgo_hash(int32_t x, int32_t y, int32_t z, uint32_t num_buckets) {
  return hash_bucket(I3{x, y, z}, num_buckets);
}

void vsgo_alloc_world_to_block(const float* p3, float v, int32_t* voxel3, int32_t* block3) {
  V3<float> p{p3[0], p3[1], p3[2]};
  voxel3[0] = alloc_voxel_axis(p.x, v);
  voxel3[1] = alloc_voxel_axis(p.y, v);
  voxel3[2] = alloc_voxel_axis(p.z, v);
  I3 b = alloc_world_to_block(p, v);
  block3[0] = b.x;
  block3[1] = b.y;
  block3[2] = b.z;
}

void vsgo_ray_world_to_voxel(const float* p3, float v, int32_t* voxel3, int32_t* block3, int32_t* local3) {
  int32_t vv[3] = {ray_voxel_axis(p3[0], v), ray_voxel_axis(p3[1], v), ray_voxel_axis(p3[2], v)};
  for (int k = 0; k < 3; ++k) {
    voxel3[k] = vv[k];
    block3[k] = voxel_to_block_axis(vv[k]);
    local3[k] = voxel_local_axis(vv[k]);
  }
}

int vsgo_dda(const float* a3, const float* b3, float v, int32_t* out, int cap) {
  int n = 0;
  dda_band(V3<float>{a3[0], a3[1], a3[2]}, V3<float>{b3[0], b3[1], b3[2]}, v, [&](I3 c, int) {
    if (n < cap) {
      out[3 * n] = c.x;
      out[3 * n + 1] = c.y;
      out[3 * n + 2] = c.z;
    }
    n++;
  });
  return n;
}

uint16_t vsgo_f32_to_f16(float f) { return f32_to_f16(f); }
float vsgo_f16_to_f32(uint16_t h) { return f16_to_f32(h); }

// Trilinear sample of the current store at world point p3 (float oracle only).
int vsgo_trilinear(void* p, const float* p3, float v, float* dist, uint8_t* col3) {
  Handle* h = (Handle*)p;
  if (h->fp64) {
    double d = 0;
    int ok = trilinear(h->sd, V3<double>{p3[0], p3[1], p3[2]}, (double)v, &d, col3) ? 1 : 0;
    *dist = (float)d;
    return ok;
  }
  return trilinear(h->sf, V3<float>{p3[0], p3[1], p3[2]}, v, dist, col3) ? 1 : 0;
}

// Accepted crossing of each listed pixel's march: the bracket's ray parameters
// (a, b) and the refined alpha, or NaN without a hit.  Used to tell pixels whose
// depth differs between arithmetics because a different sample pair was accepted
// from pixels that accepted the same pair.
void vsgo_march_brackets(void* p, int W, const double* cam3, const double* T16, double fx, double fy, double cx,
                         double cy, double minD, double maxD, double voxel, const int32_t* pix, long n, double* out3) {
  Handle* h = (Handle*)p;
  auto run = [&](auto& s, auto tag) {
    using R = decltype(tag);
    R M[16];
    to_pose(T16, M);
    V3<R> cam{(R)cam3[0], (R)cam3[1], (R)cam3[2]};
#pragma omp parallel for schedule(dynamic, 64)
    for (long k = 0; k < n; ++k) {
      MarchResult<R> m = march_pixel(s, pix[k] % W, pix[k] / W, cam, M, (R)fx, (R)fy, (R)cx, (R)cy, (R)minD, (R)maxD,
                                     (R)voxel);
      const double nan = std::numeric_limits<double>::quiet_NaN();
      out3[3 * k] = m.hit ? (double)m.a : nan;
      out3[3 * k + 1] = m.hit ? (double)m.b : nan;
      out3[3 * k + 2] = m.hit ? (double)m.alpha : nan;
    }
  };
  if (h->fp64) run(h->sd, 0.0);
  else run(h->sf, 0.0f);
}

// Raycast adjoint for listed pixels: J (6 per pixel, d depth / d xi for T * exp(xi^)),
// depth (in the oracle's arithmetic; 1e10 without a hit),
// and, when grad_depth is given, grad_sdf[ptr*512 + voxel] += grad_depth[k] * d depth / d sdf.
// hit[k] = 1 where the pixel has a hit.  grad_sdf covers num_allocated() blocks.
void vsgo_render_adjoint(void* p, int W, int H, const double* cam3, const double* T16, double fx, double fy, double cx,
                         double cy, double minD, double maxD, double voxel, const int32_t* pix, long n, double* J,
                         uint8_t* hit, double* depth, const double* grad_depth, double* grad_sdf) {
  Handle* h = (Handle*)p;
  (void)H;
  auto run = [&](auto& s, auto tag) {
    using R = decltype(tag);
    R M[16];
    to_pose(T16, M);
    V3<R> cam{(R)cam3[0], (R)cam3[1], (R)cam3[2]};
    std::vector<int64_t> vx((size_t)n * 32);
    std::vector<double> dd((size_t)n * 32);
#pragma omp parallel for schedule(dynamic, 16)
    for (long k = 0; k < n; ++k) {
      double dk = 0;
      bool ok = adjoint_pixel(s, pix[k] % W, pix[k] / W, cam, M, (R)fx, (R)fy, (R)cx, (R)cy, (R)minD, (R)maxD,
                              (R)voxel, &J[6 * k], &vx[32 * k], &dd[32 * k], &dk);
      if (depth) depth[k] = dk;
      if (hit) hit[k] = ok ? 1 : 0;
    }
    if (grad_depth && grad_sdf) {
      for (long k = 0; k < n; ++k)
        for (int j = 0; j < 32; ++j)
          if (vx[32 * k + j] >= 0) grad_sdf[vx[32 * k + j]] += grad_depth[k] * dd[32 * k + j];
    }
  };
  if (h->fp64) run(h->sd, 0.0);
  else run(h->sf, 0.0f);
}

// Reads / overwrites the stored SDF of one voxel (global index ptr*512 + voxel).
double vsgo_get_sdf(void* p, int64_t idx) {
  Handle* h = (Handle*)p;
  return h->fp64 ? h->sd.load_sdf((size_t)idx) : (double)h->sf.load_sdf((size_t)idx);
}
void vsgo_set_sdf(void* p, int64_t idx, double value) {
  Handle* h = (Handle*)p;
  if (h->fp64) h->sd.store_sdf((size_t)idx, value);
  else h->sf.store_sdf((size_t)idx, (float)value);
}

// Overwrites SDF, weight and (rgb3 != null) colour of n voxels (global indices
// ptr*512 + voxel) of allocated blocks: designed fields for the raycaster's
// refinement and acceptance checks.
void vsgo_set_voxels(void* p, uint64_t n, const int64_t* idx, const double* sdf, const uint8_t* weight,
                     const uint8_t* rgb3) {
  Handle* h = (Handle*)p;
  for (uint64_t i = 0; i < n; ++i) {
    const size_t k = (size_t)idx[i];
    uint8_t* col;
    if (h->fp64) {
      h->sd.store_sdf(k, sdf[i]);
      h->sd.weight[k] = weight[i];
      col = &h->sd.color[3 * k];
    } else {
      h->sf.store_sdf(k, (float)sdf[i]);
      h->sf.weight[k] = weight[i];
      col = &h->sf.color[3 * k];
    }
    if (rgb3)
      for (int c = 0; c < 3; ++c) col[c] = rgb3[3 * i + c];
  }
}

// Driver's per-frame depth range (VoxelSG/src/voxelSG.cpp:594-608), float arithmetic.
void vsgo_depth_range(const float* depth, long n, long stride, float* out2) {
  float mn = 1e10f, mx = -1e10f;
  for (long i = 0; i < n; ++i) {
    float d = depth[i * stride];
    if (d > 0.0f) {
      mn = std::fmin(mn, d);
      mx = std::fmax(mx, d);
    }
  }
  float margin = (mx - mn) * 0.1f;
  mn = std::fmax(0.01f, mn - margin);
  mx += margin;
  out2[0] = mn;
  out2[1] = mx;
}

uint32_t vsgo_depthmap_from_points(int W, int H, float fx, float fy, float cx, float cy, int ortho, const float* pts,
                                   uint64_t n, const float* nrm, uint64_t n_nrm, const uint8_t* col, uint64_t n_col,
                                   float* pmap, float* depth, float* nmap, uint8_t* cmap) {
  return depthmap_from_points(W, H, fx, fy, cx, cy, ortho != 0, pts, (size_t)n, nrm, (size_t)n_nrm, col, (size_t)n_col,
                              pmap, depth, nmap, cmap);
}

// Stream-out of the blocks at distance >= radius from c3 (at most cap, table order):
// returns how many; coords 3 per block, sdf / weight 512 per block, rgbf 4 per voxel,
// direction slots 2 per voxel (any output but coords may be null).
uint64_t vsgo_stream_out(void* p, const double* c3, double radius, uint64_t cap, int32_t* coords, double* sdf,
                         uint16_t* weight, uint8_t* rgbf, uint8_t* dir_id, uint8_t* dir_w, double* dtsdf) {
  Handle* h = (Handle*)p;
  auto run = [&](auto& sc) -> uint64_t {
    using R = typename std::decay_t<decltype(sc)>::Real;
    auto out = sc.stream_out(V3<R>{(R)c3[0], (R)c3[1], (R)c3[2]}, (R)radius, (size_t)cap);
    for (size_t k = 0; k < out.size(); ++k) {
      const auto& b = out[k];
      coords[3 * k] = b.c.x;
      coords[3 * k + 1] = b.c.y;
      coords[3 * k + 2] = b.c.z;
      for (int v = 0; v < kVox; ++v) {
        const size_t i = k * kVox + v;
        if (sdf) sdf[i] = (double)b.sdf[v];
        if (weight) weight[i] = b.weight[v];
        if (rgbf)
          for (int ch = 0; ch < 4; ++ch) rgbf[4 * i + ch] = b.rgbf[4 * v + ch];
        if (!b.dir_id.empty())
          for (int s = 0; s < 2; ++s) {
            if (dir_id) dir_id[2 * i + s] = b.dir_id[2 * v + s];
            if (dir_w) dir_w[2 * i + s] = b.dir_w[2 * v + s];
            if (dtsdf) dtsdf[2 * i + s] = (double)b.dtsdf[2 * v + s];
          }
      }
    }
    return out.size();
  };
  return h->fp64 ? run(h->sd) : run(h->sf);
}

// Stream-in of n blocks in order (layout of vsgo_stream_out; direction slots may be
// null); status per block (0 inserted, 1 already allocated, 2 bucket full, 3 heap
// empty, 4 out of the key range).
void vsgo_stream_in(void* p, uint64_t n, const int32_t* coords, const double* sdf, const uint16_t* weight,
                    const uint8_t* rgbf, const uint8_t* dir_id, const uint8_t* dir_w, const double* dtsdf,
                    uint8_t* status) {
  Handle* h = (Handle*)p;
  auto run = [&](auto& sc) {
    using S = std::decay_t<decltype(sc)>;
    using R = typename S::Real;
    for (uint64_t k = 0; k < n; ++k) {
      typename S::Streamed b;
      b.c = I3{coords[3 * k], coords[3 * k + 1], coords[3 * k + 2]};
      for (int v = 0; v < kVox; ++v) {
        const size_t i = k * kVox + v;
        b.sdf.push_back((R)sdf[i]);
        b.weight.push_back(weight[i]);
        for (int ch = 0; ch < 4; ++ch) b.rgbf.push_back(rgbf[4 * i + ch]);
        if (dir_id && dir_w && dtsdf)
          for (int s = 0; s < 2; ++s) {
            b.dir_id.push_back(dir_id[2 * i + s]);
            b.dir_w.push_back(dir_w[2 * i + s]);
            b.dtsdf.push_back((R)dtsdf[2 * i + s]);
          }
      }
      status[k] = (uint8_t)sc.stream_in(b);
    }
  };
  if (h->fp64) run(h->sd); else run(h->sf);
}

}  // extern "C"
