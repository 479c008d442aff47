// Internal interface between the C ABI (vsg_capi.hip) and the kernels (vsg_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "vsg_device.hpp"

namespace vsg {

struct RenderArgs {
  float M[16];
  float cam[3];
  const float* Mdev;    // when non-null: the pose is the 16 floats here (device memory), not M
  const float* camdev;  // when non-null: the camera position is the 3 floats here, not cam
  float fx, fy, cx, cy;
  float minD, maxD;
  float voxel;
  int32_t W, H;
  uint8_t* mark_bkt;  // diagnostics (null when off): buckets with live entries probed
  uint8_t* mark_blk;  // diagnostics (null when off): blocks a trilinear corner touched
  uint32_t* march_stats;  // diagnostics (null when off): evaluations, super/block skips, replayed steps, cell
                          // skips, deferred rays, tail windows, tail skips
  int32_t mode;           // vsg_sampling: 0 the reference's trilinear sampling, 1 standard trilinear
  uint32_t budget;        // steps per lane before a ray moves to the cooperative tail pass (0: none)
  void* tail;             // tail queue (8 segments of tail_seg TailRays, one per XCD) when budget > 0
  uint32_t* tail_ctl;     // queue controls (kTailCtlWords): set tail_par for this render
  uint32_t tail_seg;      // entries per segment of the tail queue
  uint32_t tail_par;      // queue control set of this render (render count parity)
  const float* win_obs;   // when non-null: observed depth per pixel; a ray marches only over depths within
  float win;              // win of it (and within [minD, maxD]); a pixel without one (<= 0) is not marched
};

// A ray whose march continues in the tail pass: its pixel and march state.
struct TailRay {
  uint32_t pix;
  float ray, last_sdf, last_alpha;
  uint32_t last_w;
  uint32_t pad[3];
};

// Queue control words of a renderer's tail pass, two sets of 16 used by alternate
// renders: per XCD x the count of rays the march deferred into segment x (word
// 16 p + x) and the next entry handed out dynamically (word 16 p + 8 + x).  A render's
// march clears the other set (the previous render's, whose tail pass has finished).
constexpr uint32_t kTailCtlWords = 32;

// Counters of the march statistics (diagnostics).
constexpr uint32_t kMarchStats = 8;

// Raycast workgroups go to the XCDs in super-tiles of kSuperTile x kSuperTile 8x8 tiles.
constexpr uint32_t kSuperTile = 4;

// Default march steps per lane before a ray moves to the tail pass.
constexpr uint32_t kDefaultStepBudget = 28;

// Workgroups of the tail pass for a W x H render (a multiple of 8).
uint32_t tail_workgroups(int32_t W, int32_t H);

// Entries of one XCD's segment of the tail queue for a W x H render: the pixels of the
// super-tiles dealt to one XCD, at most.
uint32_t tail_segment(int32_t W, int32_t H);

// Workgroups of the march (and of the fused raycast) for a W x H render.
uint32_t render_grid(int32_t W, int32_t H);

void launch_reset_voxels(hipStream_t s, DevState* st, void* vox, uint32_t word_bytes, uint32_t* col,
                         uint32_t* dmeta, float2* dval, uint32_t* pcnt);
// Recomputes the positive-cell counts and masks of live blocks [first, first + n).
void launch_cell_masks(hipStream_t s, bool fp16, const TableView& t, const DevState* st, uint32_t first, uint32_t n,
                       const void* vox);
// Positive-cell masks of blocks [first, first + n) into out (device, n bytes).
void launch_export_cell_masks(hipStream_t s, const TableView& t, uint32_t first, uint32_t n, uint8_t* out);
void launch_reset_table(hipStream_t s, const TableView& t, DevState* st);
// Workgroups of k_integrate (grid-stride over the list): 4096 measured 60.9 us per room frame
// against 62.5 at 2048, 66.4 at 1024 and 62.6 at 8192.
constexpr uint32_t kIntegrateGrid = 4096;
void launch_alloc_claim(hipStream_t s, const TableView& t, const FrameArgs& fa, const float* pts, uint32_t* pixcnt,
                        uint8_t* mark_bkt, uint32_t* part);
void launch_alloc_resolve(hipStream_t s, const TableView& t, const FrameArgs& fa, const float* pts, uint32_t* pixcnt,
                          uint32_t* wgsum);
void launch_alloc_insert(hipStream_t s, const TableView& t, DevState* st, const FrameArgs& fa, const float* pts,
                         const uint32_t* pixcnt, const uint32_t* wgsum, const uint32_t* part);
void launch_compact(hipStream_t s, const TableView& t, DevState* st, const FrameArgs& fa, uint32_t* vis,
                    uint32_t grid);
void launch_integrate(hipStream_t s, bool fp16, const TableView& t, DevState* st, const FrameArgs& fa,
                      const float* pts, const uint8_t* rgb, const float* nrm, const uint32_t* vis, void* vox,
                      uint32_t* col, uint32_t* dmeta, float2* dval, uint32_t grid, uint32_t* upd_part);
// Scratch of the surface-point extraction (device buffers): bucket_start[num_buckets],
// oblk, bcount, pstart[num_blocks], nbp[kExtNbWords * num_blocks], ebits[8 * num_blocks],
// sums[extract_sums_words(max(num_buckets, num_blocks))].
struct ExtractScratch {
  uint32_t* bucket_start;
  int4* oblk;
  int32_t* nbp;
  uint64_t* ebits;
  uint32_t* bcount;
  uint32_t* pstart;
  uint32_t* sums;
};
constexpr uint32_t kExtNbWords = 32;
constexpr uint32_t kScanChunkSize = 4096;
inline uint32_t extract_sums_words(uint32_t n) { return (n + kScanChunkSize - 1u) / kScanChunkSize + 1u; }

// Surface points of the scene (RayCastRender.cu:629-907) in hash-entry order, then
// voxel order; the first max_points written to pos / colr / nrm.  The number found is
// pstart[num_blocks - 1] + bcount[num_blocks - 1] once the work is done.
void launch_extract(hipStream_t s, bool fp16, const TableView& t, const DevState* st, const void* vox,
                    const uint32_t* col, float voxel, float min_sdf, float max_sdf, uint32_t min_weight,
                    uint32_t max_points, const ExtractScratch& x, float4* pos, float4* colr, float4* nrm);

// Marching cubes (MarchingCubes.cu:352-467) over the scene's live blocks in hash-entry
// order, then cube order, then table order: 3 vertices and 3 normals per triangle, the
// first max_tris kept; the number found is pstart[num_blocks - 1] + bcount[num_blocks
// - 1] once the work is done.  table: int8 [256][kMcRow] on the device.  The scratch's
// ebits is unused; nbp only with cross-block cubes.
struct McArgs {
  float voxel, iso;
  uint32_t min_weight;
  uint32_t cross;  // 1: cubes at local 7 take corners from the +x / +y / +z neighbours
  uint32_t max_tris;
};
void launch_marching_cubes(hipStream_t s, bool fp16, const TableView& t, const DevState* st, const void* vox,
                           const int8_t* table, const McArgs& a, const ExtractScratch& x, float3* verts,
                           float3* nrms);

// Depth map from a camera-frame point cloud (CustomDepthMapGenerator.cpp:81-222).
struct DepthGenArgs {
  float fx, fy, cx, cy;
  int32_t W, H;
  int32_t orthographic;  // ProjectionMethod::ORTHOGRAPHIC, else perspective
  uint32_t n;            // points
  uint32_t n_normals;    // normals given for points [0, n_normals)
  uint32_t n_colors;     // colours given for points [0, n_colors)
};
// win: W*H words; win_n / win_c: W*H words when n_normals / n_colors < n, else null.
// out_depth may be null.
void launch_depthgen(hipStream_t s, const DepthGenArgs& g, const float* pts, const float* nrm, const uint8_t* rgb,
                     uint32_t* win, uint32_t* win_n, uint32_t* win_c, float* out_pts, float* out_depth,
                     float* out_nrm, uint8_t* out_rgb);
// *valid = pixels of the last launch_depthgen with a point (its win array).
void launch_depthgen_count(hipStream_t s, const uint32_t* win, uint32_t npix, uint32_t* valid);

// Per-block arrays of the store, in the order a block image holds them: voxel words
// (kVox * word_bytes), colours (kVox * 4 B), and with the directional extension the
// slot metadata (kVox * 4 B) and slot values (kVox * 8 B).
struct BlockPlanes {
  uint8_t* vox;
  uint32_t word_bytes;
  uint32_t* col;
  uint32_t* dmeta;  // null without the directional extension
  float2* dval;
};
uint32_t block_image_bytes(const BlockPlanes& bp);
// Words of the scan partials for n elements.
uint32_t stream_scan_words(uint32_t n);
// Marks the live blocks [0, n) at distance >= radius from c (sel: n bytes), ranks
// them in pointer order (rank: n words) and returns how many there are (synchronises
// the stream; ~0 on a HIP error).
uint32_t stream_out_select(hipStream_t s, const TableView& t, uint32_t n, const float c[3], float radius, float voxel,
                           uint8_t* sel, uint32_t* rank, uint32_t* sums);
// Removes the first m selected blocks: their coordinates into coords[0, m) and
// images into images (m images); out: m words, remap: n words of scratch.
void stream_out_remove(hipStream_t s, const TableView& t, DevState* st, const BlockPlanes& bp, uint32_t n, uint32_t m,
                       const uint8_t* sel, const uint32_t* rank, uint32_t* out, uint32_t* remap, int4* coords,
                       uint8_t* images);
// Inserts the n input blocks into their buckets (status per input: 0 inserted, 1
// already live, 2 bucket full, 3 no free block (nospace), 4 coordinate out of
// range) and returns how many were inserted (synchronises; ~0 on a HIP error).
// entry, rank: n words; ok: n bytes.
uint32_t stream_in_insert(hipStream_t s, const TableView& t, uint32_t n, const int4* coords, uint8_t* status,
                          uint32_t* entry, uint8_t* ok, uint32_t* rank, uint32_t* sums, bool nospace);
// Gives the inserted blocks pointers n0 + rank, index entries and their planes from
// the images, and recomputes their positive-cell counts and masks.
void stream_in_place(hipStream_t s, bool fp16, const TableView& t, DevState* st, const BlockPlanes& bp, uint32_t n,
                     uint32_t n0, uint32_t inserted, const int4* coords, const uint8_t* ok, const uint32_t* rank,
                     const uint32_t* entry, const uint8_t* images);

// out[i] = cvt_i32(in[i]) for i < n (device buffers).
void launch_convert(hipStream_t s, const float* in, int32_t* out, uint32_t n);
void launch_render_adjoint(hipStream_t s, bool fp16, const TableView& t, const DevState* st, const RenderArgs& ra,
                           const void* vox, const uint32_t* col, float* jac, const float* grad_depth,
                           float* grad_sdf);
// The pose Jacobian of each pixel of the render `ra` describes (depth: its output)
// through the implicit function at the crossing, with the field gradient by central
// differences over +-1 voxel (the Gauss-Newton's default).
void launch_gn_jacobian(hipStream_t s, bool fp16, const TableView& t, const DevState* st, const RenderArgs& ra,
                        const void* vox, const uint32_t* col, const float* depth, float* jac);
uint32_t gn_partial_count(int npix);
// out: new pose (16 doubles), delta (6), cost, count, solved flag (25 doubles)
// noise = {a, b, z0}: residual weights (a / (a + b (obs - z0)^2))^2 when a > 0.
// One Gauss-Newton accumulation and solve.  The pose comes from pose_in (device, may be
// the previous result `out`) or, when pose_in is null, from pose (host, by value); posef
// (may be null) receives the new pose in the form a render reads (16 + 3 floats).
void launch_gn(hipStream_t s, const float* depth, const float* obs, const float* jac, int W, int H, float max_res,
               const float noise[3], double lambda, const double pose[16], const double* pose_in, double* partials,
               double* out, float* posef);
// Writes the 3 floats of c to cam (device) on stream s.
void launch_set_cam(hipStream_t s, float* cam, const float c[3]);
void launch_render(hipStream_t s, bool fp16, const TableView& t, const DevState* st, const RenderArgs& ra,
                   const void* vox, const uint32_t* col, float* depth, float4* pos, float4* colr, float4* nrm,
                   uint32_t* rc);

}  // namespace vsg
