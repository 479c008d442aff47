// C ABI of the VoxelSG depth-fusion library (declared in include/vsg.h).
//
// Owns the device memory of scenes and renderers, validates arguments, and queues
// the kernels of vsg_kernels.hip on the object's stream.  No call synchronises
// except the ones vsg.h documents as synchronising.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "vsg.h"
#include "vsg_device.hpp"
#include "vsg_internal.hpp"
#include "vsg_mc_table.hpp"

using namespace vsg;

namespace {

thread_local std::string g_err;

vsg_status fail(vsg_status st, const std::string& msg) {
  g_err = msg;
  return st;
}

vsg_status hip_fail(hipError_t e, const char* what) {
  return fail(VSG_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

#define VSG_HIP(call)                                  \
  do {                                                 \
    hipError_t e_ = (call);                            \
    if (e_ != hipSuccess) return hip_fail(e_, #call);  \
  } while (0)

// ---- kernel timing ---------------------------------------------------------------
struct TimingRec {
  int kid;
  hipEvent_t a, b;
};

struct Timing {
  std::mutex mu;
  bool on = false;
  std::vector<TimingRec> pending;
  std::vector<hipEvent_t> pool;
  double ms[VSG_K_COUNT] = {};
  uint32_t n[VSG_K_COUNT] = {};

  hipEvent_t get() {
    if (!pool.empty()) {
      hipEvent_t e = pool.back();
      pool.pop_back();
      return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
  }
};
Timing g_timing;

// Brackets one launch with events when timing is on.
struct TimedLaunch {
  hipStream_t s;
  int kid;
  hipEvent_t a = nullptr, b = nullptr;
  bool on;
  TimedLaunch(hipStream_t s_, int kid_) : s(s_), kid(kid_) {
    std::lock_guard<std::mutex> g(g_timing.mu);
    on = g_timing.on;
    if (on) {
      a = g_timing.get();
      b = g_timing.get();
      if (!a || !b) on = false;
      else (void)hipEventRecord(a, s);
    }
  }
  ~TimedLaunch() {
    if (!on) return;
    (void)hipEventRecord(b, s);
    std::lock_guard<std::mutex> g(g_timing.mu);
    g_timing.pending.push_back(TimingRec{kid, a, b});
  }
};

vsg_status check_launch(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return hip_fail(e, what);
  return VSG_OK;
}

}  // namespace

// ---- objects ---------------------------------------------------------------------

struct vsg_scene {
  vsg_params p{};
  hipStream_t stream = nullptr;
  bool own_stream = false;
  uint32_t word_bytes = 8;
  TableView t{};
  void* vox = nullptr;
  uint32_t* col = nullptr;
  uint32_t* vis = nullptr;
  DevState* st = nullptr;
  uint32_t* pixcnt = nullptr;
  uint32_t* wgsum = nullptr;
  uint32_t* claim_part = nullptr;  // per-workgroup claim counters (8 per workgroup)
  uint32_t* upd_part = nullptr;    // per-workgroup updated-voxel counts of the last integrate
  size_t pix_cap = 0;
  uint32_t frames = 0;  // integrate calls since reset (host mirror)
  bool diag = false;
  uint8_t* mark_abkt = nullptr;  // diagnostics marks (B, B, N bytes)
  uint8_t* mark_rbkt = nullptr;
  uint8_t* mark_blk = nullptr;
  uint32_t* march_stats = nullptr;  // kMarchStats counters
  uint32_t* dmeta = nullptr;        // directional extension: slot ids and weights per voxel
  float2* dval = nullptr;           // directional extension: slot dtsdf per voxel
  uint8_t* stream_arena = nullptr;  // device scratch of vsg_stream_out / vsg_stream_in, grown on demand
  size_t stream_arena_bytes = 0;
};

struct vsg_depthgen {
  int32_t W = 0, H = 0;
  int32_t method = VSG_PROJ_PERSPECTIVE;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  uint32_t* win = nullptr;    // last point per pixel (3 * W * H words: points, normals, colours)
  uint32_t* valid = nullptr;  // pixels with depth > 0 of the last call (counted on request)
};

struct vsg_render {
  int32_t W = 0, H = 0;
  hipStream_t stream = nullptr;  // used when has_stream, else the scene's stream
  bool has_stream = false;
  hipStream_t last = nullptr;    // stream of the last render (valid when has_last)
  bool has_last = false;
  float* depth = nullptr;
  float4* pos = nullptr;
  float4* col = nullptr;
  float4* nrm = nullptr;
  uint32_t* rc = nullptr;  // {hits, misses} per workgroup of the last render (march, then tail pass)
  size_t rc_words = 0;
  size_t rc_march_words = 0;     // the march's slice of rc (the tail pass's follows it)
  uint32_t last_budget = 0;      // step budget of the last render (0: the tail pass did not run)
  void* tail = nullptr;          // tail queue: 8 segments (one per XCD) of tail_seg TailRays
  uint32_t tail_seg = 0;
  uint32_t* tail_ctl = nullptr;  // kTailCtlWords queue controls of the tail pass
  uint32_t budget = 0;           // march steps per lane before the tail pass takes the ray over
  uint32_t calls = 0;
  hipEvent_t ev = nullptr;
  int32_t sampling = VSG_SAMPLE_REFERENCE;
  RenderArgs last_args{};  // arguments of the last render (the adjoint's forward)
  bool has_args = false;
  float* jac = nullptr;       // GN scratch: Jacobian per pixel
  double* partials = nullptr; // GN scratch: per-workgroup sums
  double* gn_out = nullptr;   // GN result (25 doubles)
  float* gn_posef = nullptr;  // GN pose for the next render: 16 floats + camera position (3)
  float* cam_buf = nullptr;   // camera position of a render with a device pose (3 floats)
  int32_t gn_jacobian = VSG_GN_SMOOTHED;  // vsg_gn_jacobian of vsg_pose_gn_step
  bool gn_window = false;                 // vsg_render_set_gn_window
  float gn_noise[3] = {0.0f, 0.0f, 0.0f};  // depth noise model {a, b, z0} of the residual weights (a = 0: none)
  // surface-point extraction: outputs of capacity ext_cap, scratch sized for a scene of
  // ext_buckets buckets and ext_blocks blocks
  uint32_t max_points = 10000000u;
  uint32_t ext_cap = 0;
  float4* ext_pos = nullptr;
  float4* ext_col = nullptr;
  float4* ext_nrm = nullptr;
  uint32_t ext_count = 0;
  ExtractScratch ext{};
  uint32_t ext_buckets = 0, ext_blocks = 0;
  uint32_t* ext_total = nullptr;  // pinned host: pstart and bcount of the last block slot
};

namespace {

void free_scene(vsg_scene* s) {
  if (!s) return;
  (void)hipStreamSynchronize(s->stream);
  (void)hipFree(s->t.rec);
  (void)hipFree(s->t.cnt);
  (void)hipFree(s->t.claim);
  (void)hipFree(s->t.blk);
  (void)hipFree(s->t.occ);
  (void)hipFree(s->t.hyp);
  (void)hipFree(s->t.idx);
  (void)hipFree(s->t.pcnt);
  (void)hipFree(s->t.idxpos);
  (void)hipFree(s->dmeta);
  (void)hipFree(s->dval);
  (void)hipFree(s->vox);
  (void)hipFree(s->col);
  (void)hipFree(s->vis);
  (void)hipFree(s->st);
  (void)hipFree(s->pixcnt);
  (void)hipFree(s->wgsum);
  (void)hipFree(s->claim_part);
  (void)hipFree(s->upd_part);
  (void)hipFree(s->mark_abkt);
  (void)hipFree(s->mark_rbkt);
  (void)hipFree(s->mark_blk);
  (void)hipFree(s->march_stats);
  (void)hipFree(s->stream_arena);
  if (s->own_stream && s->stream) (void)hipStreamDestroy(s->stream);
  delete s;
}

template <class T> vsg_status dmalloc(T** p, size_t bytes, const char* what) {
  hipError_t e = hipMalloc((void**)p, bytes);
  if (e != hipSuccess) {
    *p = nullptr;
    (void)hipGetLastError();
    return fail(VSG_ERR_OOM, std::string("device allocation of ") + what + " (" + std::to_string(bytes) +
                                 " bytes) failed: " + hipGetErrorString(e));
  }
  return VSG_OK;
}

void free_extract_outputs(vsg_render* r) {
  (void)hipFree(r->ext_pos);
  (void)hipFree(r->ext_col);
  (void)hipFree(r->ext_nrm);
  r->ext_pos = r->ext_col = r->ext_nrm = nullptr;
  r->ext_cap = 0;
  r->ext_count = 0;
}

void free_extract_scratch(vsg_render* r) {
  (void)hipFree(r->ext.bucket_start);
  (void)hipFree(r->ext.oblk);
  (void)hipFree(r->ext.nbp);
  (void)hipFree(r->ext.ebits);
  (void)hipFree(r->ext.bcount);
  (void)hipFree(r->ext.pstart);
  (void)hipFree(r->ext.sums);
  r->ext = ExtractScratch{};
  r->ext_buckets = r->ext_blocks = 0;
}

vsg_status queue_reset(vsg_scene* s) {
  {
    TimedLaunch tl(s->stream, VSG_K_RESET);
    launch_reset_voxels(s->stream, s->st, s->vox, s->word_bytes, s->col, s->dmeta, s->dval, s->t.pcnt);
  }
  vsg_status r = check_launch("reset voxels");
  if (r != VSG_OK) return r;
  {
    TimedLaunch tl(s->stream, VSG_K_RESET);
    launch_reset_table(s->stream, s->t, s->st);
  }
  r = check_launch("reset table");
  if (r != VSG_OK) return r;
  s->frames = 0;
  return VSG_OK;
}

// Conservative culling needs the pose's rotation to be a rotation: R^T R = I.
bool pose_orthonormal(const float* M) {
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double d = 0;
      for (int k = 0; k < 3; ++k) d += (double)M[4 * k + i] * (double)M[4 * k + j];
      if (std::fabs(d - (i == j ? 1.0 : 0.0)) > 1e-5) return false;
    }
  return true;
}

bool all_finite(const float* v, int n) {
  for (int i = 0; i < n; ++i)
    if (!std::isfinite(v[i])) return false;
  return true;
}

}  // namespace

extern "C" {

int32_t vsg_version(void) { return 100; }

const char* vsg_last_error(void) { return g_err.c_str(); }

vsg_status vsg_scene_create(const vsg_params* params, vsg_scene** out) {
  if (!params || !out) return fail(VSG_ERR_INVALID_ARG, "vsg_scene_create: null argument");
  *out = nullptr;
  const vsg_params& p = *params;
  if (p.hash_slot_num <= 0 || p.hash_slot_num > (1 << 28))
    return fail(VSG_ERR_INVALID_ARG, "vsg_scene_create: hash_slot_num must be in [1, 2^28]");
  if (p.sdf_block_num <= 0 || p.sdf_block_num > (1 << 26))
    return fail(VSG_ERR_INVALID_ARG, "vsg_scene_create: sdf_block_num must be in [1, 2^26]");
  if (!(p.voxel_size > 0.0f) || !std::isfinite(p.voxel_size))
    return fail(VSG_ERR_INVALID_ARG, "vsg_scene_create: voxel_size must be a positive finite number");
  if (p.store != VSG_STORE_FP32 && p.store != VSG_STORE_FP16)
    return fail(VSG_ERR_INVALID_ARG, "vsg_scene_create: store must be VSG_STORE_FP32 or VSG_STORE_FP16");
  if ((p.flags & ~VSG_SCENE_DIRECTIONAL) != 0u) return fail(VSG_ERR_INVALID_ARG, "vsg_scene_create: unknown flags");
  for (int i = 0; i < 3; ++i)
    if (p.reserved[i] != 0) return fail(VSG_ERR_INVALID_ARG, "vsg_scene_create: reserved fields must be zero");

  vsg_scene* s = new vsg_scene();
  s->p = p;
  s->word_bytes = p.store == VSG_STORE_FP16 ? 4u : 8u;
  const size_t B = (size_t)p.hash_slot_num, N = (size_t)p.sdf_block_num;
  s->t.num_buckets = (uint32_t)B;
  s->t.num_blocks = (uint32_t)N;
  hipError_t e = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete s;
    return hip_fail(e, "hipStreamCreateWithFlags");
  }
  s->own_stream = true;
  vsg_status r = VSG_OK;
  if (r == VSG_OK) r = dmalloc(&s->t.rec, B * kRecWords * sizeof(uint64_t), "hash table");
  if (r == VSG_OK) r = dmalloc(&s->t.cnt, B, "bucket counts");
  if (r == VSG_OK) r = dmalloc(&s->t.claim, B * sizeof(uint32_t), "bucket claims");
  if (r == VSG_OK) r = dmalloc(&s->t.blk, N * sizeof(int4), "block list");
  // occupancy map: one word per 4x4x4-block super-block hashed into a power of two
  // of at least 2^16 words and at least one word per block slot, so that fewer
  // super-blocks than words are ever occupied
  size_t occ_words = 1u << 16;
  while (occ_words < N) occ_words <<= 1;
  s->t.occ_mask = (uint32_t)(occ_words - 1);
  if (r == VSG_OK) r = dmalloc(&s->t.occ, occ_words * sizeof(uint64_t), "block occupancy map");
  // hyper-block map: one word per 4x4x4 super-blocks, at least 2^12 words
  size_t hyp_words = 1u << 12;
  while (hyp_words < occ_words / 16) hyp_words <<= 1;
  s->t.hyp_mask = (uint32_t)(hyp_words - 1);
  if (r == VSG_OK) r = dmalloc(&s->t.hyp, hyp_words * sizeof(uint64_t), "hyper-block occupancy map");
  // block index: a power of two of at least 4 entries per block slot
  size_t idx_entries = 1u << 16;
  while (idx_entries < 4 * N) idx_entries <<= 1;
  s->t.idx_mask = (uint32_t)(idx_entries - 1);
  if (r == VSG_OK) r = dmalloc(&s->t.idx, idx_entries * sizeof(uint4), "block index");
  if (r == VSG_OK) r = dmalloc(&s->t.pcnt, N * 8 * sizeof(uint32_t), "positive-voxel counts per cell");
  if (r == VSG_OK) r = dmalloc(&s->t.idxpos, N * sizeof(uint32_t), "block index slots");
  if (r == VSG_OK) r = dmalloc(&s->vox, N * kVox * s->word_bytes, "voxel store");
  if (r == VSG_OK) r = dmalloc(&s->col, N * kVox * sizeof(uint32_t), "colour store");
  if (r == VSG_OK) r = dmalloc(&s->vis, N * sizeof(uint32_t), "visible list");
  if (r == VSG_OK && (p.flags & VSG_SCENE_DIRECTIONAL)) {
    r = dmalloc(&s->dmeta, N * kVox * sizeof(uint32_t), "direction slots");
    if (r == VSG_OK) r = dmalloc(&s->dval, N * kVox * sizeof(float2), "direction slot values");
  }
  if (r == VSG_OK) r = dmalloc(&s->st, sizeof(DevState), "scene state");
  if (r == VSG_OK) r = dmalloc(&s->upd_part, kIntegrateGrid * sizeof(uint32_t), "integrate counters");
  if (r != VSG_OK) {
    free_scene(s);
    return r;
  }
  // new scene: voxel memory zero (CUDAHashRef.cu:86-93), then reset
  e = hipMemsetAsync(s->vox, 0, N * kVox * s->word_bytes, s->stream);
  if (e == hipSuccess) e = hipMemsetAsync(s->col, 0, N * kVox * sizeof(uint32_t), s->stream);
  if (e == hipSuccess && s->dmeta) e = hipMemsetAsync(s->dmeta, 0, N * kVox * sizeof(uint32_t), s->stream);
  if (e == hipSuccess && s->dval) e = hipMemsetAsync(s->dval, 0, N * kVox * sizeof(float2), s->stream);
  if (e == hipSuccess) e = hipMemsetAsync(s->st, 0, sizeof(DevState), s->stream);
  if (e == hipSuccess) e = hipMemsetAsync(s->t.blk, 0, N * sizeof(int4), s->stream);
  if (e == hipSuccess) e = hipMemsetAsync(s->t.pcnt, 0, N * 8 * sizeof(uint32_t), s->stream);
  if (e == hipSuccess) e = hipMemsetAsync(s->t.idxpos, 0, N * sizeof(uint32_t), s->stream);
  if (e == hipSuccess) e = hipMemsetAsync(s->upd_part, 0, kIntegrateGrid * sizeof(uint32_t), s->stream);
  if (e != hipSuccess) {
    free_scene(s);
    return hip_fail(e, "hipMemsetAsync");
  }
  r = queue_reset(s);
  if (r == VSG_OK) {
    e = hipStreamSynchronize(s->stream);
    if (e != hipSuccess) r = hip_fail(e, "hipStreamSynchronize");
  }
  if (r != VSG_OK) {
    free_scene(s);
    return r;
  }
  *out = s;
  return VSG_OK;
}

vsg_status vsg_scene_destroy(vsg_scene* scene) {
  if (!scene) return fail(VSG_ERR_INVALID_ARG, "vsg_scene_destroy: null scene");
  free_scene(scene);
  return VSG_OK;
}

vsg_status vsg_scene_reset(vsg_scene* scene) {
  if (!scene) return fail(VSG_ERR_INVALID_ARG, "vsg_scene_reset: null scene");
  return queue_reset(scene);
}

vsg_status vsg_scene_set_stream(vsg_scene* scene, void* hip_stream) {
  if (!scene) return fail(VSG_ERR_INVALID_ARG, "vsg_scene_set_stream: null scene");
  VSG_HIP(hipStreamSynchronize(scene->stream));
  if (scene->own_stream) VSG_HIP(hipStreamDestroy(scene->stream));
  scene->stream = (hipStream_t)hip_stream;
  scene->own_stream = false;
  return VSG_OK;
}

vsg_status vsg_scene_get_stream(const vsg_scene* scene, void** hip_stream) {
  if (!scene || !hip_stream) return fail(VSG_ERR_INVALID_ARG, "vsg_scene_get_stream: null argument");
  *hip_stream = (void*)scene->stream;
  return VSG_OK;
}

namespace {

vsg_status check_frame(const char* who, vsg_scene* s, const vsg_frame* f, float truncation, const float pose[16],
                       const vsg_intrinsics* k) {
  char msg[160];
  auto bad = [&](const char* what) {
    std::snprintf(msg, sizeof(msg), "%s: %s", who, what);
    return fail(VSG_ERR_INVALID_ARG, msg);
  };
  if (!s || !f || !pose || !k) return bad("null argument");
  if (!f->points || !f->colors) return bad("null point or colour map");
  if (f->width <= 0 || f->height <= 0) return bad("empty frame");
  const size_t npix = (size_t)f->width * (size_t)f->height;
  if (npix > (1u << 22)) return bad("frames are limited to 2^22 pixels");  // keys are (pixel << 10 | DDA step)
  if (!std::isfinite(truncation)) return bad("truncation must be finite");
  if (!all_finite(pose, 16)) return bad("pose must be finite");
  const float kk[4] = {k->fx, k->fy, k->cx, k->cy};
  if (!all_finite(kk, 4)) return bad("intrinsics must be finite");
  return VSG_OK;
}

// The frame's kernels with the given per-pixel and per-workgroup scratch (every word
// of [0, npix) and [0, nwg) is written before it is read).
vsg_status integrate_frame(vsg_scene* s, const vsg_frame* f, float truncation, const float pose[16],
                           const vsg_intrinsics* k, uint32_t* pixcnt, uint32_t* wgsum, uint32_t* claim_part) {
  FrameArgs fa;
  std::memset(&fa, 0, sizeof(fa));
  for (int i = 0; i < 16; ++i) {
    fa.M[i] = pose[i];
    fa.Md[i] = (double)pose[i];
  }
  fa.fx = k->fx;
  fa.fy = k->fy;
  fa.cx = k->cx;
  fa.cy = k->cy;
  fa.trunc = truncation;
  fa.voxel = s->p.voxel_size;
  fa.W = f->width;
  fa.H = f->height;
  fa.cset = (int32_t)(s->frames & 1u);
  fa.cull_ok = pose_orthonormal(pose) ? 1 : 0;

  const float* pts = (const float*)f->points;
  const uint8_t* rgb = (const uint8_t*)f->colors;
  const bool fp16 = s->p.store == VSG_STORE_FP16;
  vsg_status r;
  {
    TimedLaunch tl(s->stream, VSG_K_ALLOC_CLAIM);
    launch_alloc_claim(s->stream, s->t, fa, pts, pixcnt, s->diag ? s->mark_abkt : nullptr, claim_part);
  }
  if ((r = check_launch("alloc claim")) != VSG_OK) return r;
  {
    TimedLaunch tl(s->stream, VSG_K_ALLOC_RESOLVE);
    launch_alloc_resolve(s->stream, s->t, fa, pts, pixcnt, wgsum);
  }
  if ((r = check_launch("alloc resolve")) != VSG_OK) return r;
  {
    TimedLaunch tl(s->stream, VSG_K_ALLOC_INSERT);
    launch_alloc_insert(s->stream, s->t, s->st, fa, pts, pixcnt, wgsum, claim_part);
  }
  if ((r = check_launch("alloc insert")) != VSG_OK) return r;
  {
    TimedLaunch tl(s->stream, VSG_K_COMPACT);
    launch_compact(s->stream, s->t, s->st, fa, s->vis, 1024);
  }
  if ((r = check_launch("compact")) != VSG_OK) return r;
  {
    TimedLaunch tl(s->stream, VSG_K_INTEGRATE);
    launch_integrate(s->stream, fp16, s->t, s->st, fa, pts, rgb, (const float*)f->normals, s->vis, s->vox, s->col,
                     s->dmeta, s->dval, kIntegrateGrid, s->upd_part);
  }
  if ((r = check_launch("integrate")) != VSG_OK) return r;
  s->frames++;
  return VSG_OK;
}

}  // namespace

vsg_status vsg_integrate(vsg_scene* s, const vsg_frame* f, float truncation, const float pose[16],
                         const vsg_intrinsics* k) {
  vsg_status r = check_frame("vsg_integrate", s, f, truncation, pose, k);
  if (r != VSG_OK) return r;
  const size_t npix = (size_t)f->width * (size_t)f->height;
  const uint32_t nwg = (uint32_t)((npix + 255) / 256);
  if (npix > s->pix_cap) {
    VSG_HIP(hipStreamSynchronize(s->stream));
    (void)hipFree(s->pixcnt);
    (void)hipFree(s->wgsum);
    (void)hipFree(s->claim_part);
    s->pixcnt = nullptr;
    s->wgsum = nullptr;
    s->claim_part = nullptr;
    s->pix_cap = 0;
    r = dmalloc(&s->pixcnt, npix * sizeof(uint32_t), "pixel scratch");
    if (r == VSG_OK) r = dmalloc(&s->wgsum, (size_t)nwg * sizeof(uint32_t), "workgroup scratch");
    if (r == VSG_OK) r = dmalloc(&s->claim_part, (size_t)nwg * 8 * sizeof(uint32_t), "workgroup counters");
    if (r != VSG_OK) return r;
    s->pix_cap = npix;
  }
  return integrate_frame(s, f, truncation, pose, k, s->pixcnt, s->wgsum, s->claim_part);
}

// ---- integration without allocation ---------------------------------------------------
//
// The scratch of vsg_integrate_noalloc: per-pixel counts, then per-workgroup sums, then
// per-workgroup claim counters, for frames of up to kNoAllocPixels pixels.  One copy per
// GPU (each device holds its own, for the life of the process); its device address is
// looked up once per device with that device current.
constexpr uint32_t kNoAllocPixels = 1u << 20;
constexpr uint32_t kNoAllocGroups = kNoAllocPixels / 256u;
constexpr uint32_t kNoAllocWords = kNoAllocPixels + kNoAllocGroups + 8u * kNoAllocGroups;
static __device__ __align__(256) uint32_t scratch[kNoAllocWords];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
uint32_t* g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on the current device (looked up on first use).
vsg_status scratch_on_current_device(uint32_t** out) {
  int dev = 0;
  VSG_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return fail(VSG_ERR_INVALID_ARG, "vsg_integrate_noalloc: device index too large");
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      (void)hipGetLastError();
      return fail(VSG_ERR_HIP, "vsg_integrate_noalloc: hipGetSymbolAddress(scratch) failed");
    }
    g_scratch_ptr[dev] = (uint32_t*)p;
  }
  *out = g_scratch_ptr[dev];
  return VSG_OK;
}

}  // namespace

vsg_status vsg_integrate_noalloc(vsg_scene* s, const vsg_frame* f, float truncation, const float pose[16],
                                 const vsg_intrinsics* k) {
  vsg_status r = check_frame("vsg_integrate_noalloc", s, f, truncation, pose, k);
  if (r != VSG_OK) return r;
  if ((size_t)f->width * (size_t)f->height > kNoAllocPixels)
    return fail(VSG_ERR_INVALID_ARG, "vsg_integrate_noalloc: frames are limited to 2^20 pixels (its static scratch)");
  uint32_t* base = nullptr;
  if ((r = scratch_on_current_device(&base)) != VSG_OK) return r;
  return integrate_frame(s, f, truncation, pose, k, base, base + kNoAllocPixels, base + kNoAllocPixels + kNoAllocGroups);
}

vsg_status vsg_scene_counters(vsg_scene* s, vsg_counters* out) {
  if (!s || !out) return fail(VSG_ERR_INVALID_ARG, "vsg_scene_counters: null argument");
  VSG_HIP(hipStreamSynchronize(s->stream));
  DevState h;
  VSG_HIP(hipMemcpy(&h, s->st, sizeof(h), hipMemcpyDeviceToHost));
  std::memset(out, 0, sizeof(*out));
  out->num_allocated = h.n_alloc;
  out->frames = h.frames;
  if (s->frames > 0) {
    const FrameCounters& f = h.set[(s->frames - 1u) & 1u];
    out->valid_pixels = f.valid_pixels;
    out->new_blocks = h.n_alloc - h.n_alloc_prev;
    out->heap_fail = f.heap_fail;
    out->visible_blocks = f.visible;
    std::vector<uint32_t> up(kIntegrateGrid);
    VSG_HIP(hipMemcpy(up.data(), s->upd_part, up.size() * sizeof(uint32_t), hipMemcpyDeviceToHost));
    uint32_t updated = 0;
    for (uint32_t u : up) updated += u;
    out->updated_voxels = updated;
    out->claimed_buckets = f.claimed_buckets;
    out->out_of_range = f.out_of_range;
  }
  return VSG_OK;
}

vsg_status vsg_heap_free_count(vsg_scene* s, uint32_t* out) {
  if (!s || !out) return fail(VSG_ERR_INVALID_ARG, "vsg_heap_free_count: null argument");
  VSG_HIP(hipStreamSynchronize(s->stream));
  uint32_t n = 0;
  VSG_HIP(hipMemcpy(&n, &s->st->n_alloc, sizeof(n), hipMemcpyDeviceToHost));
  *out = s->t.num_blocks - n;
  return VSG_OK;
}

vsg_status vsg_num_integrated_frames(const vsg_scene* s, uint32_t* out) {
  if (!s || !out) return fail(VSG_ERR_INVALID_ARG, "vsg_num_integrated_frames: null argument");
  *out = s->frames;
  return VSG_OK;
}

vsg_status vsg_scene_hash_data(const vsg_scene* s, vsg_hash_data* out) {
  if (!s || !out) return fail(VSG_ERR_INVALID_ARG, "vsg_scene_hash_data: null argument");
  out->buckets = s->t.rec;
  out->bucket_used = s->t.cnt;
  out->blocks = (int32_t*)s->t.blk;
  out->voxels = s->vox;
  out->colors = s->col;
  out->block_index = s->t.idx;
  out->num_allocated_device = &s->st->n_alloc;
  out->num_buckets = s->t.num_buckets;
  out->num_blocks = s->t.num_blocks;
  out->voxel_bytes = s->word_bytes;
  out->index_entries = s->t.idx_mask + 1u;
  out->store = s->p.store;
  return VSG_OK;
}

vsg_status vsg_scene_params(const vsg_scene* s, vsg_params* out) {
  if (!s || !out) return fail(VSG_ERR_INVALID_ARG, "vsg_scene_params: null argument");
  *out = s->p;
  return VSG_OK;
}

vsg_status vsg_export_blocks(vsg_scene* s, int32_t* coords, int32_t* slots, uint32_t capacity, uint32_t* count) {
  if (!s || !count) return fail(VSG_ERR_INVALID_ARG, "vsg_export_blocks: null argument");
  VSG_HIP(hipStreamSynchronize(s->stream));
  uint32_t n = 0;
  VSG_HIP(hipMemcpy(&n, &s->st->n_alloc, sizeof(n), hipMemcpyDeviceToHost));
  *count = n;
  const uint32_t m = n < capacity ? n : capacity;
  if (m == 0 || (!coords && !slots)) return VSG_OK;
  std::vector<int4> h(m);
  VSG_HIP(hipMemcpy(h.data(), s->t.blk, m * sizeof(int4), hipMemcpyDeviceToHost));
  for (uint32_t i = 0; i < m; ++i) {
    if (coords) {
      coords[3 * i] = h[i].x;
      coords[3 * i + 1] = h[i].y;
      coords[3 * i + 2] = h[i].z;
    }
    if (slots) slots[i] = h[i].w;
  }
  return VSG_OK;
}

vsg_status vsg_export_directional(vsg_scene* s, uint32_t first, uint32_t nblocks, uint8_t* ids, uint8_t* weights,
                                  float* dtsdf) {
  if (!s) return fail(VSG_ERR_INVALID_ARG, "vsg_export_directional: null scene");
  if (!s->dmeta) return fail(VSG_ERR_INVALID_ARG, "vsg_export_directional: scene created without VSG_SCENE_DIRECTIONAL");
  if ((uint64_t)first + nblocks > s->t.num_blocks)
    return fail(VSG_ERR_INVALID_ARG, "vsg_export_directional: block range beyond the store");
  VSG_HIP(hipStreamSynchronize(s->stream));
  const size_t nv = (size_t)nblocks * kVox, off = (size_t)first * kVox;
  if (nv == 0) return VSG_OK;
  std::vector<uint32_t> meta(nv);
  std::vector<float2> val(nv);
  VSG_HIP(hipMemcpy(meta.data(), s->dmeta + off, nv * sizeof(uint32_t), hipMemcpyDeviceToHost));
  VSG_HIP(hipMemcpy(val.data(), s->dval + off, nv * sizeof(float2), hipMemcpyDeviceToHost));
  for (size_t i = 0; i < nv; ++i) {
    if (ids) {
      ids[2 * i] = (uint8_t)(meta[i] & 0xFFu);
      ids[2 * i + 1] = (uint8_t)((meta[i] >> 16) & 0xFFu);
    }
    if (weights) {
      weights[2 * i] = (uint8_t)((meta[i] >> 8) & 0xFFu);
      weights[2 * i + 1] = (uint8_t)(meta[i] >> 24);
    }
    if (dtsdf) {
      dtsdf[2 * i] = val[i].x;
      dtsdf[2 * i + 1] = val[i].y;
    }
  }
  return VSG_OK;
}

vsg_status vsg_export_voxels(vsg_scene* s, uint32_t first, uint32_t nblocks, float* sdf, uint16_t* weight,
                             uint8_t* rgbf) {
  if (!s) return fail(VSG_ERR_INVALID_ARG, "vsg_export_voxels: null scene");
  if ((uint64_t)first + nblocks > s->t.num_blocks)
    return fail(VSG_ERR_INVALID_ARG, "vsg_export_voxels: block range beyond the store");
  VSG_HIP(hipStreamSynchronize(s->stream));
  const size_t nv = (size_t)nblocks * kVox, off = (size_t)first * kVox;
  if (nv == 0) return VSG_OK;
  if (sdf || weight) {
    std::vector<uint8_t> w(nv * s->word_bytes);
    VSG_HIP(hipMemcpy(w.data(), (uint8_t*)s->vox + off * s->word_bytes, w.size(), hipMemcpyDeviceToHost));
    for (size_t i = 0; i < nv; ++i) {
      if (s->word_bytes == 8) {
        uint32_t a, b;
        std::memcpy(&a, &w[8 * i], 4);
        std::memcpy(&b, &w[8 * i + 4], 4);
        if (sdf) std::memcpy(&sdf[i], &a, 4);
        if (weight) weight[i] = (uint16_t)(b & 0xFFFFu);
      } else {
        uint32_t a;
        std::memcpy(&a, &w[4 * i], 4);
        if (sdf) {
          // binary16 -> binary32 (exact)
          uint32_t h = a & 0xFFFFu, sign = (h & 0x8000u) << 16, e = (h >> 10) & 0x1Fu, m = h & 0x3FFu, x;
          if (e == 0) {
            if (m == 0) {
              x = sign;
            } else {
              int kk = -1;
              do {
                kk++;
                m <<= 1;
              } while (!(m & 0x400u));
              x = sign | ((uint32_t)(112 - kk) << 23) | ((m & 0x3FFu) << 13);
            }
          } else if (e == 31) {
            x = sign | 0x7F800000u | (m << 13);
          } else {
            x = sign | ((e + 112u) << 23) | (m << 13);
          }
          std::memcpy(&sdf[i], &x, 4);
        }
        if (weight) weight[i] = (uint16_t)(a >> 16);
      }
    }
  }
  if (rgbf) VSG_HIP(hipMemcpy(rgbf, s->col + off, nv * sizeof(uint32_t), hipMemcpyDeviceToHost));
  return VSG_OK;
}

// ---- renderer ----------------------------------------------------------------------

vsg_status vsg_render_create(int32_t width, int32_t height, vsg_render** out) {
  if (!out) return fail(VSG_ERR_INVALID_ARG, "vsg_render_create: null argument");
  *out = nullptr;
  if (width <= 0 || height <= 0 || (int64_t)width * height > (1 << 26))
    return fail(VSG_ERR_INVALID_ARG, "vsg_render_create: width and height must be positive, at most 2^26 pixels");
  vsg_render* r = new vsg_render();
  r->W = width;
  r->H = height;
  const size_t n = (size_t)width * height;
  vsg_status st = dmalloc(&r->depth, n * sizeof(float), "depth output");
  if (st == VSG_OK) st = dmalloc(&r->pos, n * sizeof(float4), "position output");
  if (st == VSG_OK) st = dmalloc(&r->col, n * sizeof(float4), "colour output");
  if (st == VSG_OK) st = dmalloc(&r->nrm, n * sizeof(float4), "normal output");
  const size_t nwg = (size_t)((width + 7) / 8) * ((height + 7) / 8);
  r->rc_march_words = 2 * nwg;
  r->rc_words = 2 * (nwg + tail_workgroups(width, height));
  if (st == VSG_OK) st = dmalloc(&r->rc, r->rc_words * sizeof(uint32_t), "render counters");
  if (st == VSG_OK && hipMemset(r->rc, 0, r->rc_words * sizeof(uint32_t)) != hipSuccess)
    st = fail(VSG_ERR_HIP, "hipMemset");
  r->tail_seg = tail_segment(width, height);
  if (st == VSG_OK) st = dmalloc(&r->tail, (size_t)8 * r->tail_seg * sizeof(TailRay), "tail queue");
  if (st == VSG_OK) st = dmalloc(&r->tail_ctl, kTailCtlWords * sizeof(uint32_t), "tail queue controls");
  if (st == VSG_OK) st = dmalloc(&r->cam_buf, 4 * sizeof(float), "camera position");
  if (st == VSG_OK && hipMemset(r->tail_ctl, 0, kTailCtlWords * sizeof(uint32_t)) != hipSuccess)
    st = fail(VSG_ERR_HIP, "hipMemset");
  r->budget = kDefaultStepBudget;
  if (st == VSG_OK && hipEventCreateWithFlags(&r->ev, hipEventDisableTiming) != hipSuccess)
    st = fail(VSG_ERR_HIP, "hipEventCreate");
  if (st != VSG_OK) {
    vsg_render_destroy(r);
    return st;
  }
  *out = r;
  return VSG_OK;
}

vsg_status vsg_render_destroy(vsg_render* r) {
  if (!r) return fail(VSG_ERR_INVALID_ARG, "vsg_render_destroy: null renderer");
  if (r->has_last) (void)hipStreamSynchronize(r->last);
  (void)hipFree(r->depth);
  (void)hipFree(r->pos);
  (void)hipFree(r->col);
  (void)hipFree(r->nrm);
  (void)hipFree(r->rc);
  (void)hipFree(r->tail);
  (void)hipFree(r->tail_ctl);
  (void)hipFree(r->jac);
  (void)hipFree(r->partials);
  (void)hipFree(r->gn_out);
  (void)hipFree(r->gn_posef);
  (void)hipFree(r->cam_buf);
  free_extract_outputs(r);
  free_extract_scratch(r);
  if (r->ext_total) (void)hipHostFree(r->ext_total);
  if (r->ev) (void)hipEventDestroy(r->ev);
  delete r;
  return VSG_OK;
}

vsg_status vsg_render_set_sampling(vsg_render* r, int32_t mode) {
  if (!r) return fail(VSG_ERR_INVALID_ARG, "vsg_render_set_sampling: null renderer");
  if (mode != VSG_SAMPLE_REFERENCE && mode != VSG_SAMPLE_TRILINEAR)
    return fail(VSG_ERR_INVALID_ARG, "vsg_render_set_sampling: unknown mode");
  r->sampling = mode;
  return VSG_OK;
}

vsg_status vsg_render_set_step_budget(vsg_render* r, uint32_t steps) {
  if (!r) return fail(VSG_ERR_INVALID_ARG, "vsg_render_set_step_budget: null renderer");
  // the queue controls stay consistent: every render that queues rays also drains them
  r->budget = steps;
  return VSG_OK;
}

vsg_status vsg_render_set_stream(vsg_render* r, void* hip_stream) {
  if (!r) return fail(VSG_ERR_INVALID_ARG, "vsg_render_set_stream: null renderer");
  if (r->has_last) VSG_HIP(hipStreamSynchronize(r->last));
  r->stream = (hipStream_t)hip_stream;
  r->has_stream = true;
  return VSG_OK;
}

namespace {

// Queues one render with the pose fields of `ra` already set (by value or device
// pointers); fills in the rest from the renderer and the scene.
vsg_status render_launch(vsg_render* r, vsg_scene* s, RenderArgs& ra, const vsg_intrinsics* k, float min_depth,
                         float max_depth, float voxel_size) {
  const float kk[4] = {k->fx, k->fy, k->cx, k->cy};
  if (!all_finite(kk, 4)) return fail(VSG_ERR_INVALID_ARG, "vsg_render_frame: intrinsics must be finite");
  if (std::isnan(min_depth) || std::isnan(max_depth) || std::isinf(min_depth) || std::isinf(max_depth))
    return fail(VSG_ERR_INVALID_ARG, "vsg_render_frame: depth range must be finite");
  ra.fx = k->fx;
  ra.fy = k->fy;
  ra.cx = k->cx;
  ra.cy = k->cy;
  ra.minD = min_depth;
  ra.maxD = max_depth;
  ra.voxel = voxel_size > 0.0f ? voxel_size : s->p.voxel_size;
  ra.W = r->W;
  ra.H = r->H;
  ra.mark_bkt = s->diag ? s->mark_rbkt : nullptr;
  ra.mark_blk = s->diag ? s->mark_blk : nullptr;
  ra.march_stats = s->diag ? s->march_stats : nullptr;
  ra.mode = r->sampling;
  ra.budget = r->budget;
  ra.tail = r->tail;
  ra.tail_ctl = r->tail_ctl;
  ra.tail_seg = r->tail_seg;
  ra.tail_par = r->calls & 1u;
  hipStream_t strm = r->has_stream ? r->stream : s->stream;
  if (strm != s->stream) {  // order after the scene's pending work
    VSG_HIP(hipEventRecord(r->ev, s->stream));
    VSG_HIP(hipStreamWaitEvent(strm, r->ev, 0));
  }
  if (ra.Mdev && !ra.camdev) {  // device pose, camera position by value: stage it on the stream
    launch_set_cam(strm, r->cam_buf, ra.cam);
    ra.camdev = r->cam_buf;
  }
  if (ra.Mdev && s->diag) {  // the diagnostics kernels take the pose by value: read it back
    VSG_HIP(hipStreamSynchronize(strm));
    VSG_HIP(hipMemcpy(ra.M, ra.Mdev, 16 * sizeof(float), hipMemcpyDeviceToHost));
    VSG_HIP(hipMemcpy(ra.cam, ra.camdev, 3 * sizeof(float), hipMemcpyDeviceToHost));
    ra.Mdev = ra.camdev = nullptr;
  }
  {
    TimedLaunch tl(strm, VSG_K_RAYCAST);
    launch_render(strm, s->p.store == VSG_STORE_FP16, s->t, s->st, ra, s->vox, s->col, r->depth, r->pos, r->col,
                  r->nrm, r->rc);
  }
  vsg_status st = check_launch("render");
  if (st != VSG_OK) return st;
  if (strm != s->stream) {  // the scene's next integrate waits for this render
    VSG_HIP(hipEventRecord(r->ev, strm));
    VSG_HIP(hipStreamWaitEvent(s->stream, r->ev, 0));
  }
  r->last = strm;
  r->has_last = true;
  r->calls++;
  r->last_budget = ra.budget;
  r->last_args = ra;
  r->has_args = true;
  return VSG_OK;
}

}  // namespace

vsg_status vsg_render_frame(vsg_render* r, vsg_scene* s, const float cam_pos[3], const float pose[16],
                            const vsg_intrinsics* k, float min_depth, float max_depth, float voxel_size) {
  if (!r || !s || !cam_pos || !pose || !k) return fail(VSG_ERR_INVALID_ARG, "vsg_render_frame: null argument");
  if (!all_finite(cam_pos, 3) || !all_finite(pose, 16))
    return fail(VSG_ERR_INVALID_ARG, "vsg_render_frame: camera position and pose must be finite");
  RenderArgs ra;
  std::memset(&ra, 0, sizeof(ra));
  for (int i = 0; i < 16; ++i) ra.M[i] = pose[i];
  for (int i = 0; i < 3; ++i) ra.cam[i] = cam_pos[i];
  return render_launch(r, s, ra, k, min_depth, max_depth, voxel_size);
}

vsg_status vsg_render_frame_device_pose(vsg_render* r, vsg_scene* s, const float cam_pos[3], const float* pose_device,
                                        const vsg_intrinsics* k, float min_depth, float max_depth, float voxel_size) {
  if (!r || !s || !cam_pos || !pose_device || !k)
    return fail(VSG_ERR_INVALID_ARG, "vsg_render_frame_device_pose: null argument");
  if (!all_finite(cam_pos, 3))
    return fail(VSG_ERR_INVALID_ARG, "vsg_render_frame_device_pose: camera position must be finite");
  RenderArgs ra;
  std::memset(&ra, 0, sizeof(ra));
  ra.Mdev = pose_device;
  for (int i = 0; i < 3; ++i) ra.cam[i] = cam_pos[i];
  return render_launch(r, s, ra, k, min_depth, max_depth, voxel_size);
}

vsg_status vsg_render_buffers(const vsg_render* r, float** depth, float** positions, float** colors, float** normals) {
  if (!r) return fail(VSG_ERR_INVALID_ARG, "vsg_render_buffers: null renderer");
  if (depth) *depth = r->depth;
  if (positions) *positions = (float*)r->pos;
  if (colors) *colors = (float*)r->col;
  if (normals) *normals = (float*)r->nrm;
  return VSG_OK;
}

vsg_status vsg_render_download(vsg_render* r, float* depth, float* positions, float* colors, float* normals) {
  if (!r) return fail(VSG_ERR_INVALID_ARG, "vsg_render_download: null renderer");
  if (r->has_last) VSG_HIP(hipStreamSynchronize(r->last));
  const size_t n = (size_t)r->W * r->H;
  if (depth) VSG_HIP(hipMemcpy(depth, r->depth, n * sizeof(float), hipMemcpyDeviceToHost));
  if (positions) VSG_HIP(hipMemcpy(positions, r->pos, n * sizeof(float4), hipMemcpyDeviceToHost));
  if (colors) VSG_HIP(hipMemcpy(colors, r->col, n * sizeof(float4), hipMemcpyDeviceToHost));
  if (normals) VSG_HIP(hipMemcpy(normals, r->nrm, n * sizeof(float4), hipMemcpyDeviceToHost));
  return VSG_OK;
}

vsg_status vsg_render_counters(vsg_render* r, uint32_t* hits, uint32_t* misses) {
  if (!r || !hits || !misses) return fail(VSG_ERR_INVALID_ARG, "vsg_render_counters: null argument");
  *hits = *misses = 0;
  if (r->calls == 0) return VSG_OK;
  if (r->has_last) VSG_HIP(hipStreamSynchronize(r->last));
  // without the tail pass its slice still holds an earlier render's counts
  const size_t words = r->last_budget == 0 ? r->rc_march_words : r->rc_words;
  std::vector<uint32_t> h(words);
  VSG_HIP(hipMemcpy(h.data(), r->rc, h.size() * sizeof(uint32_t), hipMemcpyDeviceToHost));
  for (size_t i = 0; i < words / 2; ++i) {
    *hits += h[2 * i];
    *misses += h[2 * i + 1];
  }
  return VSG_OK;
}

// ---- surface-point extraction ----------------------------------------------------------

vsg_status vsg_render_set_max_points(vsg_render* r, uint32_t max_points) {
  if (!r) return fail(VSG_ERR_INVALID_ARG, "vsg_render_set_max_points: null renderer");
  r->max_points = max_points;
  return VSG_OK;
}

vsg_status vsg_extract_surface_points(vsg_render* r, vsg_scene* s, float min_sdf, float max_sdf, uint32_t min_weight,
                                      uint32_t* stored, uint64_t* total) {
  if (!r || !s) return fail(VSG_ERR_INVALID_ARG, "vsg_extract_surface_points: null argument");
  if (std::isnan(min_sdf) || std::isnan(max_sdf))
    return fail(VSG_ERR_INVALID_ARG, "vsg_extract_surface_points: SDF range must not be NaN");
  const uint32_t B = s->t.num_buckets, N = s->t.num_blocks;
  // point counts and starts are 32-bit: at most 2^32 - 1 voxels
  if ((uint64_t)N * kVox > 0xFFFFFFFFull)
    return fail(VSG_ERR_INVALID_ARG, "vsg_extract_surface_points: scenes of more than 2^23 blocks are not supported");
  hipStream_t strm = r->has_stream ? r->stream : s->stream;
  if (r->has_last) VSG_HIP(hipStreamSynchronize(r->last));  // outputs and scratch are reused
  if (r->ext_cap != r->max_points || !r->ext_pos) {
    free_extract_outputs(r);
    const size_t cap = r->max_points ? r->max_points : 1u;
    vsg_status st = dmalloc(&r->ext_pos, cap * sizeof(float4), "extracted positions");
    if (st == VSG_OK) st = dmalloc(&r->ext_col, cap * sizeof(float4), "extracted colours");
    if (st == VSG_OK) st = dmalloc(&r->ext_nrm, cap * sizeof(float4), "extracted normals");
    if (st != VSG_OK) {
      free_extract_outputs(r);
      return st;
    }
    r->ext_cap = r->max_points;
  }
  if (r->ext_buckets != B || r->ext_blocks != N) {
    free_extract_scratch(r);
    const uint32_t sums = extract_sums_words(B > N ? B : N);
    vsg_status st = dmalloc(&r->ext.bucket_start, (size_t)B * sizeof(uint32_t), "extraction bucket starts");
    if (st == VSG_OK) st = dmalloc(&r->ext.oblk, (size_t)N * sizeof(int4), "extraction block order");
    if (st == VSG_OK)
      st = dmalloc(&r->ext.nbp, (size_t)N * kExtNbWords * sizeof(int32_t), "extraction neighbour pointers");
    if (st == VSG_OK) st = dmalloc(&r->ext.ebits, (size_t)N * 8 * sizeof(uint64_t), "extraction emit bits");
    if (st == VSG_OK) st = dmalloc(&r->ext.bcount, (size_t)N * sizeof(uint32_t), "extraction block counts");
    if (st == VSG_OK) st = dmalloc(&r->ext.pstart, (size_t)N * sizeof(uint32_t), "extraction block starts");
    if (st == VSG_OK) st = dmalloc(&r->ext.sums, (size_t)sums * sizeof(uint32_t), "extraction scan sums");
    if (st != VSG_OK) {
      free_extract_scratch(r);
      return st;
    }
    r->ext_buckets = B;
    r->ext_blocks = N;
  }
  if (!r->ext_total && hipHostMalloc((void**)&r->ext_total, 2 * sizeof(uint32_t), hipHostMallocDefault) != hipSuccess) {
    r->ext_total = nullptr;
    return fail(VSG_ERR_OOM, "vsg_extract_surface_points: pinned host allocation failed");
  }
  if (strm != s->stream) {  // order after the scene's pending work
    VSG_HIP(hipEventRecord(r->ev, s->stream));
    VSG_HIP(hipStreamWaitEvent(strm, r->ev, 0));
  }
  {
    TimedLaunch tl(strm, VSG_K_EXTRACT);
    launch_extract(strm, s->p.store == VSG_STORE_FP16, s->t, s->st, s->vox, s->col, s->p.voxel_size, min_sdf, max_sdf,
                   min_weight, r->ext_cap, r->ext, r->ext_pos, r->ext_col, r->ext_nrm);
  }
  vsg_status st = check_launch("extract surface points");
  if (st != VSG_OK) return st;
  VSG_HIP(hipMemcpyAsync(&r->ext_total[0], r->ext.pstart + (N - 1u), sizeof(uint32_t), hipMemcpyDeviceToHost, strm));
  VSG_HIP(hipMemcpyAsync(&r->ext_total[1], r->ext.bcount + (N - 1u), sizeof(uint32_t), hipMemcpyDeviceToHost, strm));
  VSG_HIP(hipStreamSynchronize(strm));
  const uint64_t found = (uint64_t)r->ext_total[0] + (uint64_t)r->ext_total[1];
  r->ext_count = (uint32_t)(found < r->ext_cap ? found : r->ext_cap);
  r->last = strm;
  r->has_last = true;
  if (stored) *stored = r->ext_count;
  if (total) *total = found;
  return VSG_OK;
}

vsg_status vsg_extracted_points(const vsg_render* r, float** positions, float** colors, float** normals,
                                uint32_t* count) {
  if (!r) return fail(VSG_ERR_INVALID_ARG, "vsg_extracted_points: null renderer");
  if (positions) *positions = (float*)r->ext_pos;
  if (colors) *colors = (float*)r->ext_col;
  if (normals) *normals = (float*)r->ext_nrm;
  if (count) *count = r->ext_count;
  return VSG_OK;
}

vsg_status vsg_extracted_points_download(vsg_render* r, float* positions, float* colors, float* normals,
                                         uint32_t capacity, uint32_t* count) {
  if (!r) return fail(VSG_ERR_INVALID_ARG, "vsg_extracted_points_download: null renderer");
  if (r->has_last) VSG_HIP(hipStreamSynchronize(r->last));
  const size_t n = r->ext_count < capacity ? r->ext_count : capacity;
  if (n && positions) VSG_HIP(hipMemcpy(positions, r->ext_pos, n * sizeof(float4), hipMemcpyDeviceToHost));
  if (n && colors) VSG_HIP(hipMemcpy(colors, r->ext_col, n * sizeof(float4), hipMemcpyDeviceToHost));
  if (n && normals) VSG_HIP(hipMemcpy(normals, r->ext_nrm, n * sizeof(float4), hipMemcpyDeviceToHost));
  if (count) *count = (uint32_t)n;
  return VSG_OK;
}

// ---- adjoint and pose refinement ---------------------------------------------------------

vsg_status vsg_render_adjoint(vsg_render* r, vsg_scene* s, float* jac, const float* grad_depth, float* grad_sdf) {
  if (!r || !s) return fail(VSG_ERR_INVALID_ARG, "vsg_render_adjoint: null argument");
  if (!r->has_args) return fail(VSG_ERR_NOT_READY, "vsg_render_adjoint: no render to differentiate");
  if ((grad_depth == nullptr) != (grad_sdf == nullptr))
    return fail(VSG_ERR_INVALID_ARG, "vsg_render_adjoint: grad_depth and grad_sdf go together");
  RenderArgs ra = r->last_args;
  ra.mark_bkt = ra.mark_blk = nullptr;
  {
    TimedLaunch tl(r->last, VSG_K_ADJOINT);
    launch_render_adjoint(r->last, s->p.store == VSG_STORE_FP16, s->t, s->st, ra, s->vox, s->col, jac, grad_depth,
                          grad_sdf);
  }
  return check_launch("render adjoint");
}

vsg_status vsg_render_set_gn_noise(vsg_render* r, float a, float b, float z0) {
  if (!r) return fail(VSG_ERR_INVALID_ARG, "vsg_render_set_gn_noise: null renderer");
  if (!std::isfinite(a) || !std::isfinite(b) || !std::isfinite(z0) || a < 0.0f || b < 0.0f)
    return fail(VSG_ERR_INVALID_ARG, "vsg_render_set_gn_noise: a and b must be finite and >= 0");
  r->gn_noise[0] = a;
  r->gn_noise[1] = b;
  r->gn_noise[2] = z0;
  return VSG_OK;
}

vsg_status vsg_render_set_gn_jacobian(vsg_render* r, int32_t mode) {
  if (!r) return fail(VSG_ERR_INVALID_ARG, "vsg_render_set_gn_jacobian: null renderer");
  if (mode != VSG_GN_SMOOTHED && mode != VSG_GN_EXACT)
    return fail(VSG_ERR_INVALID_ARG, "vsg_render_set_gn_jacobian: unknown mode");
  r->gn_jacobian = mode;
  return VSG_OK;
}

vsg_status vsg_render_set_gn_window(vsg_render* r, int32_t on) {
  if (!r) return fail(VSG_ERR_INVALID_ARG, "vsg_render_set_gn_window: null renderer");
  r->gn_window = on != 0;
  return VSG_OK;
}

namespace {

vsg_status gn_scratch(vsg_render* r) {
  if (r->jac) return VSG_OK;
  const int npix = r->W * r->H;
  vsg_status st = dmalloc(&r->jac, (size_t)npix * 6 * sizeof(float), "Jacobian scratch");
  if (st == VSG_OK) st = dmalloc(&r->partials, (size_t)gn_partial_count(npix) * 29 * sizeof(double), "GN scratch");
  if (st == VSG_OK) st = dmalloc(&r->gn_out, 32 * sizeof(double), "GN result");
  if (st == VSG_OK) st = dmalloc(&r->gn_posef, 20 * sizeof(float), "GN pose");
  return st;
}

// The Jacobian of the last render (the mode vsg_render_set_gn_jacobian chose).
vsg_status gn_jacobian(vsg_render* r, vsg_scene* s) {
  RenderArgs ra = r->last_args;
  ra.mark_bkt = ra.mark_blk = nullptr;
  {
    TimedLaunch tl(r->last, VSG_K_ADJOINT);
    if (r->gn_jacobian == VSG_GN_SMOOTHED)
      launch_gn_jacobian(r->last, s->p.store == VSG_STORE_FP16, s->t, s->st, ra, s->vox, s->col, r->depth, r->jac);
    else
      launch_render_adjoint(r->last, s->p.store == VSG_STORE_FP16, s->t, s->st, ra, s->vox, s->col, r->jac, nullptr,
                            nullptr);
  }
  return check_launch("render adjoint");
}

void gn_stats(const double out[25], double stats[8]) {
  stats[0] = out[22];
  stats[1] = out[23];
  stats[2] = out[24];
  stats[3] = std::sqrt(out[16] * out[16] + out[17] * out[17] + out[18] * out[18]);
  stats[4] = std::sqrt(out[19] * out[19] + out[20] * out[20] + out[21] * out[21]);
  stats[5] = stats[6] = stats[7] = 0.0;
}

}  // namespace

vsg_status vsg_pose_gn_step(vsg_render* r, vsg_scene* s, const float* observed_depth, const vsg_intrinsics* k,
                            float min_depth, float max_depth, float max_residual, double lambda, double pose[16],
                            double stats[8]) {
  if (!r || !s || !observed_depth || !k || !pose) return fail(VSG_ERR_INVALID_ARG, "vsg_pose_gn_step: null argument");
  for (int i = 0; i < 16; ++i)
    if (!std::isfinite(pose[i])) return fail(VSG_ERR_INVALID_ARG, "vsg_pose_gn_step: pose must be finite");
  vsg_status st = gn_scratch(r);
  if (st != VSG_OK) return st;
  RenderArgs ra;
  std::memset(&ra, 0, sizeof(ra));
  for (int i = 0; i < 16; ++i) ra.M[i] = (float)pose[i];
  ra.cam[0] = ra.M[3];
  ra.cam[1] = ra.M[7];
  ra.cam[2] = ra.M[11];
  if (!all_finite(ra.M, 16)) return fail(VSG_ERR_INVALID_ARG, "vsg_pose_gn_step: pose must be finite in float");
  if (r->gn_window) {
    ra.win_obs = observed_depth;
    ra.win = max_residual;
  }
  st = render_launch(r, s, ra, k, min_depth, max_depth, 0.0f);
  if (st == VSG_OK) st = gn_jacobian(r, s);
  if (st != VSG_OK) return st;
  {
    TimedLaunch tl(r->last, VSG_K_GN);
    launch_gn(r->last, r->depth, observed_depth, r->jac, r->W, r->H, max_residual, r->gn_noise, lambda, pose,
              nullptr, r->partials, r->gn_out, nullptr);
  }
  st = check_launch("gauss-newton");
  if (st != VSG_OK) return st;
  double out[25];
  VSG_HIP(hipStreamSynchronize(r->last));
  VSG_HIP(hipMemcpy(out, r->gn_out, sizeof(out), hipMemcpyDeviceToHost));
  for (int i = 0; i < 16; ++i) pose[i] = out[i];
  if (stats) gn_stats(out, stats);
  return VSG_OK;
}

vsg_status vsg_pose_gn_refine(vsg_render* r, vsg_scene* s, const float* observed_depth, const vsg_intrinsics* k,
                              float min_depth, float max_depth, uint32_t iterations, const float* max_residuals,
                              double lambda, double pose[16], double stats[8]) {
  if (!r || !s || !observed_depth || !k || !pose || (iterations && !max_residuals))
    return fail(VSG_ERR_INVALID_ARG, "vsg_pose_gn_refine: null argument");
  for (int i = 0; i < 16; ++i)
    if (!std::isfinite(pose[i])) return fail(VSG_ERR_INVALID_ARG, "vsg_pose_gn_refine: pose must be finite");
  for (uint32_t i = 0; i < iterations; ++i)
    if (!std::isfinite(max_residuals[i]))
      return fail(VSG_ERR_INVALID_ARG, "vsg_pose_gn_refine: residual windows must be finite");
  if (iterations == 0) {
    if (stats)
      for (int i = 0; i < 8; ++i) stats[i] = 0.0;
    return VSG_OK;
  }
  vsg_status st = gn_scratch(r);
  if (st != VSG_OK) return st;
  for (uint32_t it = 0; it < iterations; ++it) {
    RenderArgs ra;
    std::memset(&ra, 0, sizeof(ra));
    if (it == 0) {  // the host pose, rounded to float as vsg_pose_gn_step does
      for (int i = 0; i < 16; ++i) ra.M[i] = (float)pose[i];
      ra.cam[0] = ra.M[3];
      ra.cam[1] = ra.M[7];
      ra.cam[2] = ra.M[11];
    } else {  // the previous solve's pose, left on the device
      ra.Mdev = r->gn_posef;
      ra.camdev = r->gn_posef + 16;
    }
    if (r->gn_window) {
      ra.win_obs = observed_depth;
      ra.win = max_residuals[it];
    }
    st = render_launch(r, s, ra, k, min_depth, max_depth, 0.0f);
    if (st == VSG_OK) st = gn_jacobian(r, s);
    if (st != VSG_OK) return st;
    {
      TimedLaunch tl(r->last, VSG_K_GN);
      launch_gn(r->last, r->depth, observed_depth, r->jac, r->W, r->H, max_residuals[it], r->gn_noise, lambda, pose,
                it == 0 ? nullptr : r->gn_out, r->partials, r->gn_out, r->gn_posef);
    }
    st = check_launch("gauss-newton");
    if (st != VSG_OK) return st;
  }
  double out[25];
  VSG_HIP(hipStreamSynchronize(r->last));
  VSG_HIP(hipMemcpy(out, r->gn_out, sizeof(out), hipMemcpyDeviceToHost));
  for (int i = 0; i < 16; ++i) pose[i] = out[i];
  if (stats) gn_stats(out, stats);
  // the last render's pose lives in gn_posef, which the next iteration would overwrite:
  // an adjoint of this render must not read it through a device pointer
  if (iterations > 1) r->has_args = false;
  return VSG_OK;
}

vsg_status vsg_import_voxels(vsg_scene* s, uint32_t first, uint32_t nblocks, const float* sdf, const uint16_t* weight,
                             const uint8_t* rgbf) {
  if (!s || !sdf || !weight) return fail(VSG_ERR_INVALID_ARG, "vsg_import_voxels: null argument");
  if ((uint64_t)first + nblocks > s->t.num_blocks)
    return fail(VSG_ERR_INVALID_ARG, "vsg_import_voxels: block range beyond the store");
  VSG_HIP(hipStreamSynchronize(s->stream));
  const size_t nv = (size_t)nblocks * kVox, off = (size_t)first * kVox;
  if (nv == 0) return VSG_OK;
  std::vector<uint8_t> w(nv * s->word_bytes);
  for (size_t i = 0; i < nv; ++i) {
    if (s->word_bytes == 8) {
      uint32_t a, b = weight[i];
      std::memcpy(&a, &sdf[i], 4);
      std::memcpy(&w[8 * i], &a, 4);
      std::memcpy(&w[8 * i + 4], &b, 4);
    } else {
      // binary32 -> binary16, round to nearest even
      uint32_t x;
      std::memcpy(&x, &sdf[i], 4);
      const uint32_t sign = (x >> 16) & 0x8000u, ax = x & 0x7FFFFFFFu;
      uint32_t h;
      if (ax >= 0x7F800000u) h = sign | 0x7C00u | (ax > 0x7F800000u ? 0x200u : 0u);
      else if (ax >= 0x477FF000u) h = sign | 0x7C00u;
      else if (ax < 0x38800000u) {
        if (ax < 0x33000000u) h = sign;
        else {
          const uint32_t e = ax >> 23, m = (ax & 0x7FFFFFu) | 0x800000u, sh = 126u - e;
          uint32_t q = m >> sh;
          const uint32_t rem = m & ((1u << sh) - 1u), half = 1u << (sh - 1u);
          if (rem > half || (rem == half && (q & 1u))) q++;
          h = sign | q;
        }
      } else {
        uint32_t q = ((ax >> 23) - 112u) << 10 | ((ax & 0x7FFFFFu) >> 13);
        const uint32_t rem = ax & 0x1FFFu;
        if (rem > 0x1000u || (rem == 0x1000u && (q & 1u))) q++;
        h = sign | q;
      }
      const uint32_t word = (h & 0xFFFFu) | ((uint32_t)weight[i] << 16);
      std::memcpy(&w[4 * i], &word, 4);
    }
  }
  VSG_HIP(hipMemcpy((uint8_t*)s->vox + off * s->word_bytes, w.data(), w.size(), hipMemcpyHostToDevice));
  if (rgbf) VSG_HIP(hipMemcpy(s->col + off, rgbf, nv * sizeof(uint32_t), hipMemcpyHostToDevice));
  // the raycaster's positive-cell masks follow the new voxels
  launch_cell_masks(s->stream, s->p.store == VSG_STORE_FP16, s->t, s->st, first, nblocks, s->vox);
  vsg_status st = check_launch("cell masks");
  if (st != VSG_OK) return st;
  VSG_HIP(hipStreamSynchronize(s->stream));
  return VSG_OK;
}

vsg_status vsg_export_cell_masks(vsg_scene* s, uint32_t first, uint32_t nblocks, uint8_t* masks, uint8_t* counts) {
  if (!s) return fail(VSG_ERR_INVALID_ARG, "vsg_export_cell_masks: null scene");
  if ((uint64_t)first + nblocks > s->t.num_blocks)
    return fail(VSG_ERR_INVALID_ARG, "vsg_export_cell_masks: block range beyond the store");
  if (nblocks == 0) return VSG_OK;
  if (masks) {
    uint8_t* d = nullptr;
    vsg_status st = dmalloc(&d, nblocks, "mask export buffer");
    if (st != VSG_OK) return st;
    launch_export_cell_masks(s->stream, s->t, first, nblocks, d);
    st = check_launch("export cell masks");
    hipError_t e = hipStreamSynchronize(s->stream);
    if (st == VSG_OK && e == hipSuccess) e = hipMemcpy(masks, d, nblocks, hipMemcpyDeviceToHost);
    (void)hipFree(d);
    if (st != VSG_OK) return st;
    if (e != hipSuccess) return hip_fail(e, "vsg_export_cell_masks");
  }
  if (counts) {
    VSG_HIP(hipStreamSynchronize(s->stream));
    std::vector<uint32_t> c((size_t)nblocks * 8);
    VSG_HIP(hipMemcpy(c.data(), s->t.pcnt + (size_t)first * 8, c.size() * sizeof(uint32_t), hipMemcpyDeviceToHost));
    for (size_t i = 0; i < c.size(); ++i) counts[i] = (uint8_t)c[i];
  }
  return VSG_OK;
}

// ---- diagnostics ----------------------------------------------------------------------

vsg_status vsg_diag_enable(vsg_scene* s, int32_t on) {
  if (!s) return fail(VSG_ERR_INVALID_ARG, "vsg_diag_enable: null scene");
  VSG_HIP(hipStreamSynchronize(s->stream));
  if (on && !s->mark_blk) {
    vsg_status r = dmalloc(&s->mark_abkt, s->t.num_buckets, "diagnostic marks");
    if (r == VSG_OK) r = dmalloc(&s->mark_rbkt, s->t.num_buckets, "diagnostic marks");
    if (r == VSG_OK) r = dmalloc(&s->mark_blk, s->t.num_blocks, "diagnostic marks");
    if (r == VSG_OK) r = dmalloc(&s->march_stats, kMarchStats * sizeof(uint32_t), "diagnostic counters");
    if (r != VSG_OK) return r;
  }
  if (on) {
    VSG_HIP(hipMemset(s->march_stats, 0, kMarchStats * sizeof(uint32_t)));
    VSG_HIP(hipMemset(s->mark_abkt, 0, s->t.num_buckets));
    VSG_HIP(hipMemset(s->mark_rbkt, 0, s->t.num_buckets));
    VSG_HIP(hipMemset(s->mark_blk, 0, s->t.num_blocks));
  }
  s->diag = on != 0;
  return VSG_OK;
}

vsg_status vsg_diag_convert(const float* in, int32_t* out, uint32_t n) {
  if (!in || !out) return fail(VSG_ERR_INVALID_ARG, "vsg_diag_convert: null argument");
  if (n == 0) return VSG_OK;
  float* din = nullptr;
  int32_t* dout = nullptr;
  vsg_status st = dmalloc(&din, (size_t)n * sizeof(float), "conversion input");
  if (st == VSG_OK) st = dmalloc(&dout, (size_t)n * sizeof(int32_t), "conversion output");
  hipError_t e = hipSuccess;
  if (st == VSG_OK) e = hipMemcpy(din, in, (size_t)n * sizeof(float), hipMemcpyHostToDevice);
  if (st == VSG_OK && e == hipSuccess) {
    launch_convert(nullptr, din, dout, n);
    st = check_launch("convert");
    if (st == VSG_OK) e = hipMemcpy(out, dout, (size_t)n * sizeof(int32_t), hipMemcpyDeviceToHost);
  }
  (void)hipFree(din);
  (void)hipFree(dout);
  if (st != VSG_OK) return st;
  if (e != hipSuccess) return hip_fail(e, "vsg_diag_convert");
  return VSG_OK;
}

vsg_status vsg_diag_read(vsg_scene* s, vsg_diag* out, int32_t clear) {
  if (!s || !out) return fail(VSG_ERR_INVALID_ARG, "vsg_diag_read: null argument");
  std::memset(out, 0, sizeof(*out));
  if (!s->mark_blk) return VSG_OK;
  VSG_HIP(hipDeviceSynchronize());
  std::vector<uint8_t> a(s->t.num_buckets), r(s->t.num_buckets), b(s->t.num_blocks);
  VSG_HIP(hipMemcpy(a.data(), s->mark_abkt, a.size(), hipMemcpyDeviceToHost));
  VSG_HIP(hipMemcpy(r.data(), s->mark_rbkt, r.size(), hipMemcpyDeviceToHost));
  VSG_HIP(hipMemcpy(b.data(), s->mark_blk, b.size(), hipMemcpyDeviceToHost));
  for (size_t i = 0; i < a.size(); ++i) {
    out->alloc_buckets += a[i] ? 1u : 0u;
    out->ray_buckets += r[i] ? 1u : 0u;
    out->any_buckets += (a[i] | r[i]) ? 1u : 0u;
  }
  for (size_t i = 0; i < b.size(); ++i) out->ray_blocks += b[i] ? 1u : 0u;
  VSG_HIP(hipMemcpy(out->march, s->march_stats, kMarchStats * sizeof(uint32_t), hipMemcpyDeviceToHost));
  if (clear) {
    VSG_HIP(hipMemset(s->march_stats, 0, kMarchStats * sizeof(uint32_t)));
    VSG_HIP(hipMemset(s->mark_abkt, 0, s->t.num_buckets));
    VSG_HIP(hipMemset(s->mark_rbkt, 0, s->t.num_buckets));
    VSG_HIP(hipMemset(s->mark_blk, 0, s->t.num_blocks));
  }
  return VSG_OK;
}

// ---- timing -------------------------------------------------------------------------

vsg_status vsg_timing_enable(int32_t on) {
  std::lock_guard<std::mutex> g(g_timing.mu);
  g_timing.on = on != 0;
  return VSG_OK;
}

vsg_status vsg_timing_reset(void) {
  std::vector<TimingRec> pend;
  {
    std::lock_guard<std::mutex> g(g_timing.mu);
    pend.swap(g_timing.pending);
  }
  for (auto& t : pend) (void)hipEventSynchronize(t.b);
  std::lock_guard<std::mutex> g(g_timing.mu);
  for (auto& t : pend) {
    g_timing.pool.push_back(t.a);
    g_timing.pool.push_back(t.b);
  }
  for (int i = 0; i < VSG_K_COUNT; ++i) {
    g_timing.ms[i] = 0;
    g_timing.n[i] = 0;
  }
  return VSG_OK;
}

vsg_status vsg_timing_read(double ms[VSG_K_COUNT], uint32_t launches[VSG_K_COUNT]) {
  std::vector<TimingRec> pend;
  {
    std::lock_guard<std::mutex> g(g_timing.mu);
    pend.swap(g_timing.pending);
  }
  std::vector<float> el(pend.size(), 0.0f);
  for (size_t i = 0; i < pend.size(); ++i) {
    VSG_HIP(hipEventSynchronize(pend[i].b));
    VSG_HIP(hipEventElapsedTime(&el[i], pend[i].a, pend[i].b));
  }
  std::lock_guard<std::mutex> g(g_timing.mu);
  for (size_t i = 0; i < pend.size(); ++i) {
    g_timing.ms[pend[i].kid] += el[i];
    g_timing.n[pend[i].kid] += 1;
    g_timing.pool.push_back(pend[i].a);
    g_timing.pool.push_back(pend[i].b);
  }
  for (int i = 0; i < VSG_K_COUNT; ++i) {
    if (ms) ms[i] = g_timing.ms[i];
    if (launches) launches[i] = g_timing.n[i];
  }
  return VSG_OK;
}

// ---- depth maps from point clouds ------------------------------------------------------

vsg_status vsg_depthgen_create(int32_t width, int32_t height, vsg_depthgen** out) {
  if (!out) return fail(VSG_ERR_INVALID_ARG, "vsg_depthgen_create: null argument");
  *out = nullptr;
  if (width <= 0 || height <= 0 || (int64_t)width * height > (1 << 26))
    return fail(VSG_ERR_INVALID_ARG, "vsg_depthgen_create: width and height must be positive, at most 2^26 pixels");
  vsg_depthgen* g = new vsg_depthgen();
  g->W = width;
  g->H = height;
  hipError_t e = hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete g;
    return hip_fail(e, "hipStreamCreateWithFlags");
  }
  g->own_stream = true;
  const size_t n = (size_t)width * height;
  vsg_status st = dmalloc(&g->win, 3 * n * sizeof(uint32_t), "depth-map winners");
  if (st == VSG_OK) st = dmalloc(&g->valid, sizeof(uint32_t), "depth-map counter");
  if (st == VSG_OK && hipMemset(g->valid, 0, sizeof(uint32_t)) != hipSuccess) st = fail(VSG_ERR_HIP, "hipMemset");
  if (st == VSG_OK && hipMemset(g->win, 0, 3 * n * sizeof(uint32_t)) != hipSuccess) st = fail(VSG_ERR_HIP, "hipMemset");
  if (st != VSG_OK) {
    vsg_depthgen_destroy(g);
    return st;
  }
  *out = g;
  return VSG_OK;
}

vsg_status vsg_depthgen_destroy(vsg_depthgen* g) {
  if (!g) return fail(VSG_ERR_INVALID_ARG, "vsg_depthgen_destroy: null generator");
  if (g->stream) (void)hipStreamSynchronize(g->stream);
  (void)hipFree(g->win);
  (void)hipFree(g->valid);
  if (g->own_stream && g->stream) (void)hipStreamDestroy(g->stream);
  delete g;
  return VSG_OK;
}

vsg_status vsg_depthgen_set_stream(vsg_depthgen* g, void* hip_stream) {
  if (!g) return fail(VSG_ERR_INVALID_ARG, "vsg_depthgen_set_stream: null generator");
  VSG_HIP(hipStreamSynchronize(g->stream));
  if (g->own_stream) VSG_HIP(hipStreamDestroy(g->stream));
  g->stream = (hipStream_t)hip_stream;
  g->own_stream = false;
  return VSG_OK;
}

vsg_status vsg_depthgen_set_projection(vsg_depthgen* g, int32_t method) {
  if (!g) return fail(VSG_ERR_INVALID_ARG, "vsg_depthgen_set_projection: null generator");
  if (method != VSG_PROJ_PERSPECTIVE && method != VSG_PROJ_ORTHOGRAPHIC && method != VSG_PROJ_CUSTOM_CAMERA)
    return fail(VSG_ERR_INVALID_ARG, "vsg_depthgen_set_projection: unknown method");
  g->method = method;
  return VSG_OK;
}

vsg_status vsg_depthgen_from_points(vsg_depthgen* g, const vsg_intrinsics* k, const float* points, uint32_t n,
                                    const float* normals, uint32_t n_normals, const uint8_t* colors,
                                    uint32_t n_colors, float* pointmap, float* depth, float* normalmap,
                                    uint8_t* colormap) {
  if (!g || !k || !pointmap || !normalmap || !colormap)
    return fail(VSG_ERR_INVALID_ARG, "vsg_depthgen_from_points: null argument");
  if (n != 0u && !points) return fail(VSG_ERR_INVALID_ARG, "vsg_depthgen_from_points: null points");
  if (n == 0xFFFFFFFFu) return fail(VSG_ERR_INVALID_ARG, "vsg_depthgen_from_points: at most 2^32 - 2 points");
  if ((n_normals != 0u && !normals) || (n_colors != 0u && !colors))
    return fail(VSG_ERR_INVALID_ARG, "vsg_depthgen_from_points: null normals or colours with a nonzero count");
  const float kk[4] = {k->fx, k->fy, k->cx, k->cy};
  if (!all_finite(kk, 4)) return fail(VSG_ERR_INVALID_ARG, "vsg_depthgen_from_points: intrinsics must be finite");
  DepthGenArgs a;
  a.fx = k->fx;
  a.fy = k->fy;
  a.cx = k->cx;
  a.cy = k->cy;
  a.W = g->W;
  a.H = g->H;
  a.orthographic = g->method == VSG_PROJ_ORTHOGRAPHIC ? 1 : 0;
  a.n = n;
  a.n_normals = n_normals < n ? n_normals : n;
  a.n_colors = n_colors < n ? n_colors : n;
  const size_t npix = (size_t)g->W * g->H;
  uint32_t* win_n = a.n_normals < n && a.n_normals != 0u ? g->win + npix : nullptr;
  uint32_t* win_c = a.n_colors < n && a.n_colors != 0u ? g->win + 2 * npix : nullptr;
  {
    TimedLaunch tl(g->stream, VSG_K_DEPTHGEN);
    launch_depthgen(g->stream, a, points, normals, colors, g->win, win_n, win_c, pointmap, depth, normalmap, colormap);
  }
  return check_launch("depth map from points");
}

vsg_status vsg_depthgen_valid_pixels(vsg_depthgen* g, uint32_t* count) {
  if (!g || !count) return fail(VSG_ERR_INVALID_ARG, "vsg_depthgen_valid_pixels: null argument");
  launch_depthgen_count(g->stream, g->win, (uint32_t)((size_t)g->W * g->H), g->valid);
  vsg_status st = check_launch("depth map pixel count");
  if (st != VSG_OK) return st;
  VSG_HIP(hipStreamSynchronize(g->stream));
  VSG_HIP(hipMemcpy(count, g->valid, sizeof(uint32_t), hipMemcpyDeviceToHost));
  return VSG_OK;
}

// ---- meshing ---------------------------------------------------------------------------

vsg_status vsg_mc_table(int8_t* out) {
  if (!out) return fail(VSG_ERR_INVALID_ARG, "vsg_mc_table: null argument");
  int8_t tri[256][kMcRow];
  if (build_mc_table(tri) < 0) return fail(VSG_ERR_NOT_READY, "vsg_mc_table: table generation failed");
  std::memcpy(out, tri, sizeof(tri));
  return VSG_OK;
}

struct vsg_mesh {
  uint32_t cap = 0;         // maxTriangles
  float3* verts = nullptr;  // 3 * cap
  float3* nrms = nullptr;   // 3 * cap
  int8_t* table = nullptr;  // [256][kMcRow]
  int most = 0;             // triangles of the largest case
  ExtractScratch x{};       // scan scratch for a scene of x_buckets buckets, x_blocks blocks
  uint32_t x_buckets = 0, x_blocks = 0;
  uint32_t* found = nullptr;  // pinned host: pstart and bcount of the last block slot
  uint32_t count = 0;         // triangles stored by the last run
  hipStream_t last = nullptr;
  bool has_last = false;
};

static void free_mesh_scratch(vsg_mesh* m) {
  (void)hipFree(m->x.bucket_start);
  (void)hipFree(m->x.oblk);
  (void)hipFree(m->x.nbp);
  (void)hipFree(m->x.bcount);
  (void)hipFree(m->x.pstart);
  (void)hipFree(m->x.sums);
  m->x = ExtractScratch{};
  m->x_buckets = m->x_blocks = 0;
}

vsg_status vsg_mesh_create(uint32_t max_triangles, vsg_mesh** out) {
  if (!out) return fail(VSG_ERR_INVALID_ARG, "vsg_mesh_create: null argument");
  *out = nullptr;
  if (max_triangles == 0u || max_triangles > (1u << 28))
    return fail(VSG_ERR_INVALID_ARG, "vsg_mesh_create: max_triangles must be in [1, 2^28]");
  int8_t tri[256][kMcRow];
  const int most = build_mc_table(tri);
  if (most < 0) return fail(VSG_ERR_NOT_READY, "vsg_mesh_create: table generation failed");
  vsg_mesh* m = new vsg_mesh();
  m->cap = max_triangles;
  m->most = most;
  vsg_status st = dmalloc(&m->verts, (size_t)max_triangles * 3 * sizeof(float3), "mesh vertices");
  if (st == VSG_OK) st = dmalloc(&m->nrms, (size_t)max_triangles * 3 * sizeof(float3), "mesh normals");
  if (st == VSG_OK) st = dmalloc(&m->table, sizeof(tri), "triangle table");
  if (st == VSG_OK && hipMemcpy(m->table, tri, sizeof(tri), hipMemcpyHostToDevice) != hipSuccess)
    st = fail(VSG_ERR_HIP, "hipMemcpy");
  if (st == VSG_OK && hipHostMalloc((void**)&m->found, 2 * sizeof(uint32_t), hipHostMallocDefault) != hipSuccess) {
    m->found = nullptr;
    st = fail(VSG_ERR_OOM, "vsg_mesh_create: pinned host allocation failed");
  }
  if (st != VSG_OK) {
    vsg_mesh_destroy(m);
    return st;
  }
  *out = m;
  return VSG_OK;
}

vsg_status vsg_mesh_destroy(vsg_mesh* m) {
  if (!m) return fail(VSG_ERR_INVALID_ARG, "vsg_mesh_destroy: null mesh");
  if (m->has_last) (void)hipStreamSynchronize(m->last);
  (void)hipFree(m->verts);
  (void)hipFree(m->nrms);
  (void)hipFree(m->table);
  free_mesh_scratch(m);
  if (m->found) (void)hipHostFree(m->found);
  delete m;
  return VSG_OK;
}

vsg_status vsg_marching_cubes(vsg_mesh* m, vsg_scene* s, float iso, float voxel_size, int32_t cross_block,
                              uint32_t min_weight, uint32_t* stored, uint64_t* total) {
  if (!m || !s) return fail(VSG_ERR_INVALID_ARG, "vsg_marching_cubes: null argument");
  if (std::isnan(iso) || std::isnan(voxel_size))
    return fail(VSG_ERR_INVALID_ARG, "vsg_marching_cubes: iso value and voxel size must not be NaN");
  const uint32_t B = s->t.num_buckets, N = s->t.num_blocks;
  hipStream_t strm = s->stream;
  if (m->has_last) VSG_HIP(hipStreamSynchronize(m->last));  // outputs and scratch are reused
  VSG_HIP(hipStreamSynchronize(strm));
  uint32_t live = 0;
  VSG_HIP(hipMemcpy(&live, &s->st->n_alloc, sizeof(live), hipMemcpyDeviceToHost));
  // triangle counts and starts are 32-bit
  if ((uint64_t)live * kVox * (uint64_t)m->most > 0xFFFFFFFFull)
    return fail(VSG_ERR_INVALID_ARG, "vsg_marching_cubes: the scene's blocks could hold 2^32 or more triangles");
  if (m->x_buckets != B || m->x_blocks != N) {
    free_mesh_scratch(m);
    const uint32_t sums = extract_sums_words(B > N ? B : N);
    vsg_status st = dmalloc(&m->x.bucket_start, (size_t)B * sizeof(uint32_t), "mesh bucket starts");
    if (st == VSG_OK) st = dmalloc(&m->x.oblk, (size_t)N * sizeof(int4), "mesh block order");
    if (st == VSG_OK) st = dmalloc(&m->x.nbp, (size_t)N * kExtNbWords * sizeof(int32_t), "mesh neighbour pointers");
    if (st == VSG_OK) st = dmalloc(&m->x.bcount, (size_t)N * sizeof(uint32_t), "mesh block counts");
    if (st == VSG_OK) st = dmalloc(&m->x.pstart, (size_t)N * sizeof(uint32_t), "mesh block starts");
    if (st == VSG_OK) st = dmalloc(&m->x.sums, (size_t)sums * sizeof(uint32_t), "mesh scan sums");
    if (st != VSG_OK) {
      free_mesh_scratch(m);
      return st;
    }
    m->x_buckets = B;
    m->x_blocks = N;
  }
  const McArgs a{voxel_size > 0.0f ? voxel_size : s->p.voxel_size, iso, min_weight, cross_block ? 1u : 0u, m->cap};
  {
    TimedLaunch tl(strm, VSG_K_MESH);
    launch_marching_cubes(strm, s->p.store == VSG_STORE_FP16, s->t, s->st, s->vox, m->table, a, m->x, m->verts,
                          m->nrms);
  }
  vsg_status st = check_launch("marching cubes");
  if (st != VSG_OK) return st;
  VSG_HIP(hipMemcpyAsync(&m->found[0], m->x.pstart + (N - 1u), sizeof(uint32_t), hipMemcpyDeviceToHost, strm));
  VSG_HIP(hipMemcpyAsync(&m->found[1], m->x.bcount + (N - 1u), sizeof(uint32_t), hipMemcpyDeviceToHost, strm));
  VSG_HIP(hipStreamSynchronize(strm));
  const uint64_t found = (uint64_t)m->found[0] + (uint64_t)m->found[1];
  m->count = (uint32_t)(found < m->cap ? found : m->cap);
  m->last = strm;
  m->has_last = true;
  if (stored) *stored = m->count;
  if (total) *total = found;
  return VSG_OK;
}

vsg_status vsg_mesh_buffers(const vsg_mesh* m, float** vertices, float** normals, uint32_t* count) {
  if (!m) return fail(VSG_ERR_INVALID_ARG, "vsg_mesh_buffers: null mesh");
  if (vertices) *vertices = (float*)m->verts;
  if (normals) *normals = (float*)m->nrms;
  if (count) *count = m->count;
  return VSG_OK;
}

vsg_status vsg_mesh_download(vsg_mesh* m, float* vertices, float* normals, uint32_t capacity, uint32_t* count) {
  if (!m) return fail(VSG_ERR_INVALID_ARG, "vsg_mesh_download: null mesh");
  if (m->has_last) VSG_HIP(hipStreamSynchronize(m->last));
  const size_t n = m->count < capacity ? m->count : capacity;
  if (n && vertices) VSG_HIP(hipMemcpy(vertices, m->verts, n * 3 * sizeof(float3), hipMemcpyDeviceToHost));
  if (n && normals) VSG_HIP(hipMemcpy(normals, m->nrms, n * 3 * sizeof(float3), hipMemcpyDeviceToHost));
  if (count) *count = (uint32_t)n;
  return VSG_OK;
}

}  // extern "C"

// ---- streaming (VoxelStreamingManager.cu / .cpp) ----------------------------------

namespace {

BlockPlanes planes_of(vsg_scene* s) {
  return BlockPlanes{(uint8_t*)s->vox, s->word_bytes, s->col, s->dmeta, s->dval};
}

// The device scratch of a streaming call: consecutive 256-byte aligned pieces of the
// scene's stream arena.  The arena is kept between calls and replaced by a larger one
// only when a call needs more than it holds (both entry points synchronise the scene's
// stream first, so no kernel still uses the old one).
struct ArenaPlan {
  size_t bytes = 0;
  size_t add(size_t n) {
    const size_t at = bytes;
    bytes += (n + 255) & ~(size_t)255;
    return at;
  }
};

vsg_status stream_arena(vsg_scene* s, size_t bytes, uint8_t** base) {
  if (bytes > s->stream_arena_bytes) {
    (void)hipFree(s->stream_arena);
    s->stream_arena = nullptr;
    s->stream_arena_bytes = 0;
    vsg_status st = dmalloc(&s->stream_arena, bytes, "streaming scratch");
    if (st != VSG_OK) return st;
    s->stream_arena_bytes = bytes;
  }
  *base = s->stream_arena;
  return VSG_OK;
}

}  // namespace

extern "C" {

vsg_status vsg_block_image_bytes(const vsg_scene* s, uint32_t* out) {
  if (!s || !out) return fail(VSG_ERR_INVALID_ARG, "vsg_block_image_bytes: null argument");
  *out = (uint32_t)kVox * (s->word_bytes + 4u + (s->dmeta ? 12u : 0u));
  return VSG_OK;
}

vsg_status vsg_stream_out(vsg_scene* s, const float center[3], float radius, uint32_t capacity, int32_t* coords,
                          void* images, uint32_t* count) {
  if (!s || !center || !count) return fail(VSG_ERR_INVALID_ARG, "vsg_stream_out: null argument");
  if (capacity > 0 && (!coords || !images))
    return fail(VSG_ERR_INVALID_ARG, "vsg_stream_out: coords and images are required when capacity > 0");
  *count = 0;
  VSG_HIP(hipStreamSynchronize(s->stream));
  uint32_t n = 0;
  VSG_HIP(hipMemcpy(&n, &s->st->n_alloc, sizeof(n), hipMemcpyDeviceToHost));
  if (n == 0 || capacity == 0) return VSG_OK;
  const BlockPlanes bp = planes_of(s);
  const size_t ib = block_image_bytes(bp);
  // pieces for at most min(n, capacity) blocks leaving the scene
  const size_t mmax = capacity < n ? capacity : n;
  ArenaPlan plan;
  const size_t o_sel = plan.add(n), o_rank = plan.add((size_t)n * 4),
               o_sums = plan.add((size_t)stream_scan_words(n) * 4), o_out = plan.add(mmax * 4),
               o_remap = plan.add((size_t)n * 4), o_coords = plan.add(mmax * sizeof(int4)), o_img = plan.add(mmax * ib);
  uint8_t* base = nullptr;
  vsg_status st;
  if ((st = stream_arena(s, plan.bytes, &base)) != VSG_OK) return st;
  uint8_t* sel = base + o_sel;
  uint32_t* rank = (uint32_t*)(base + o_rank);
  uint32_t* sums = (uint32_t*)(base + o_sums);
  const uint32_t selected = stream_out_select(s->stream, s->t, n, center, radius, s->p.voxel_size, sel, rank, sums);
  if ((st = check_launch("stream-out selection")) != VSG_OK) return st;
  if (selected == 0xFFFFFFFFu) return fail(VSG_ERR_HIP, "vsg_stream_out: selection failed");
  const uint32_t m = selected < capacity ? selected : capacity;
  if (m == 0) return VSG_OK;
  uint32_t* out = (uint32_t*)(base + o_out);
  uint32_t* remap = (uint32_t*)(base + o_remap);
  int4* dcoords = (int4*)(base + o_coords);
  uint8_t* dimg = base + o_img;
  stream_out_remove(s->stream, s->t, s->st, bp, n, m, sel, rank, out, remap, dcoords, dimg);
  if ((st = check_launch("stream-out")) != VSG_OK) return st;
  VSG_HIP(hipStreamSynchronize(s->stream));
  std::vector<int4> hc(m);
  VSG_HIP(hipMemcpy(hc.data(), dcoords, (size_t)m * sizeof(int4), hipMemcpyDeviceToHost));
  VSG_HIP(hipMemcpy(images, dimg, (size_t)m * ib, hipMemcpyDeviceToHost));
  for (uint32_t i = 0; i < m; ++i) {
    coords[3 * i] = hc[i].x;
    coords[3 * i + 1] = hc[i].y;
    coords[3 * i + 2] = hc[i].z;
  }
  *count = m;
  return VSG_OK;
}

vsg_status vsg_stream_in(vsg_scene* s, uint32_t n, const int32_t* coords, const void* images, uint8_t* status,
                         uint32_t* inserted) {
  if (!s) return fail(VSG_ERR_INVALID_ARG, "vsg_stream_in: null scene");
  if (n > 0 && (!coords || !images)) return fail(VSG_ERR_INVALID_ARG, "vsg_stream_in: null coords or images");
  if (inserted) *inserted = 0;
  if (n == 0) return VSG_OK;
  VSG_HIP(hipStreamSynchronize(s->stream));
  const BlockPlanes bp = planes_of(s);
  const size_t ib = block_image_bytes(bp);
  const bool fp16 = s->p.store == VSG_STORE_FP16;
  ArenaPlan plan;
  const size_t o_coords = plan.add((size_t)n * sizeof(int4)), o_img = plan.add((size_t)n * ib),
               o_status = plan.add(n), o_ok = plan.add(n), o_entry = plan.add((size_t)n * 4),
               o_rank = plan.add((size_t)n * 4), o_sums = plan.add((size_t)stream_scan_words(n) * 4);
  uint8_t* base = nullptr;
  vsg_status st;
  if ((st = stream_arena(s, plan.bytes, &base)) != VSG_OK) return st;
  int4* dcoords = (int4*)(base + o_coords);
  uint8_t* dimg = base + o_img;
  uint8_t* dstatus = base + o_status;
  uint8_t* ok = base + o_ok;
  uint32_t* entry = (uint32_t*)(base + o_entry);
  uint32_t* rank = (uint32_t*)(base + o_rank);
  uint32_t* sums = (uint32_t*)(base + o_sums);
  std::vector<int4> hc(n);
  for (uint32_t i = 0; i < n; ++i) hc[i] = make_int4(coords[3 * i], coords[3 * i + 1], coords[3 * i + 2], 0);
  VSG_HIP(hipMemcpy(dcoords, hc.data(), (size_t)n * sizeof(int4), hipMemcpyHostToDevice));
  VSG_HIP(hipMemcpy(dimg, images, (size_t)n * ib, hipMemcpyHostToDevice));
  uint32_t total = 0;
  // Batches no larger than the free blocks, so every insertion of a batch gets a
  // block; once none is free the rest are classified without inserting (3 = no
  // block), as a sequential loop over the inputs would.
  uint32_t first = 0;
  while (first < n) {
    uint32_t n0 = 0;
    VSG_HIP(hipMemcpy(&n0, &s->st->n_alloc, sizeof(n0), hipMemcpyDeviceToHost));
    const uint32_t free_blocks = s->t.num_blocks - n0;
    const bool nospace = free_blocks == 0;
    const uint32_t len = nospace ? n - first : (n - first < free_blocks ? n - first : free_blocks);
    const uint32_t got = stream_in_insert(s->stream, s->t, len, dcoords + first, dstatus + first, entry + first,
                                          ok + first, rank + first, sums, nospace);
    if ((st = check_launch("stream-in insertion")) != VSG_OK) return st;
    if (got == 0xFFFFFFFFu) return fail(VSG_ERR_HIP, "vsg_stream_in: insertion failed");
    stream_in_place(s->stream, fp16, s->t, s->st, bp, len, n0, got, dcoords + first, ok + first, rank + first,
                    entry + first, dimg + (size_t)first * ib);
    if ((st = check_launch("stream-in placement")) != VSG_OK) return st;
    VSG_HIP(hipStreamSynchronize(s->stream));
    total += got;
    first += len;
  }
  if (status) VSG_HIP(hipMemcpy(status, dstatus, n, hipMemcpyDeviceToHost));
  if (inserted) *inserted = total;
  return VSG_OK;
}

}  // extern "C"
