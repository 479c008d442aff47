/*
 * vsg.h -- C ABI of the MI355X (gfx950) VoxelSG depth-fusion library.
 *
 * The library replaces the hot path of the reference's
 *   VoxelScene::integrateFromScanData   (VoxelSG/include/VoxelScene.h:89-91)
 *   RayCastRender::initialize / render  (VoxelSG/include/RayCastRender.h:56,66-70)
 * and the launch wrappers under them (VoxelSG/include/VoxelScene.h:13-42,
 * VoxelSG/src/RayCastRender.cpp:28-45).  Those wrappers are extern "C" in the
 * reference but take C++ references to CUDAHashRef/Params; this header is a true
 * C ABI: plain pointers, sizes and POD structs, usable from C, C++, ctypes or cgo.
 * include/vsg_voxelscene.hpp keeps the reference's C++ class API on top of it.
 *
 * Conventions (identical to the reference):
 *   - pose: row-major float[16] camera->world [R|t; 0 0 0 1]; camera position = t.
 *   - pointmap: float3 per pixel in camera coordinates (x right, y down, z ahead),
 *     row-major W*H; a pixel is invalid when its point is (0,0,0).
 *   - colormap: uchar3 per pixel; normalmap: float3 per pixel (not used by the
 *     update, as in the reference).
 *   - frame buffers are DEVICE pointers owned by the caller; the scene owns the
 *     hash table and voxel store; a renderer owns its output buffers.
 *
 * Every call is asynchronous on the object's HIP stream (default: a stream the
 * library creates) except the ones documented as synchronising.  One scene per
 * host thread; calls on one scene must not overlap.
 */
#ifndef VSG_H
#define VSG_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum vsg_status {
  VSG_OK = 0,
  VSG_ERR_INVALID_ARG = 1, /* null pointer, bad size, or parameters out of range   */
  VSG_ERR_OOM = 2,         /* device allocation failed                              */
  VSG_ERR_HIP = 3,         /* a HIP runtime call or kernel launch failed            */
  VSG_ERR_NOT_READY = 4    /* object not initialised (RayCastRender.cpp:128-131)    */
} vsg_status;

typedef enum vsg_store {
  VSG_STORE_FP32 = 0, /* f32 SDF + u16 weight per voxel (8 B) + uchar4 colour     */
  VSG_STORE_FP16 = 1  /* f16 SDF + u16 weight per voxel (4 B) + uchar4 colour     */
} vsg_store;

/* Params (VoxelSG/include/CUDAHashData.h:9-42).  Block size 8 and bucket size 10
 * are fixed (VoxelSG/include/globalPramasConfig.h:8-9). */
typedef struct vsg_params {
  int32_t hash_slot_num; /* buckets B (Params::hashSlotNum)                        */
  int32_t sdf_block_num; /* voxel blocks N (Params::SDFBlockNum)                   */
  float voxel_size;      /* metres (Params::voxelSize)                             */
  int32_t store;         /* vsg_store                                              */
  uint32_t flags;        /* VSG_SCENE_* bits                                       */
  int32_t reserved[3];   /* must be zero                                           */
} vsg_params;

/* Directional TSDF extension: two direction slots per voxel {id, weight, dtsdf}
 * updated beside the reference's TSDF.  The compiled reference computes the
 * rotated pixel normal and drops it (VoxelScene.cu:1074-1094); its per-direction
 * fields and knobs exist only as comments (CUDAHashRef.h:34-39,
 * globalParamasConfig.cpp:101-111: 2 slots, cosine 0.8, weight max 100).  The
 * update rule is this library's definition of that design (see
 * vsg_export_directional); the reference's own fields are unchanged by it. */
#define VSG_SCENE_DIRECTIONAL 1u

typedef struct vsg_intrinsics {
  float fx, fy, cx, cy;
} vsg_intrinsics;

typedef struct vsg_frame {
  const void* points;  /* device float3[W*H]                                     */
  const void* colors;  /* device uchar3[W*H]                                     */
  const void* normals; /* device float3[W*H], may be NULL                        */
  int32_t width, height;
} vsg_frame;

/* Counters of the last integrate / render call (reference prints these:
 * VoxelSG/src/VoxelScene.cpp:487-532,708-714, VoxelSG/src/RayCastRender.cu:599-616). */
typedef struct vsg_counters {
  uint32_t valid_pixels;   /* pixels with z > 1e-6 (allocation kernel's counter)  */
  uint32_t new_blocks;     /* blocks inserted by the last integrate               */
  uint32_t heap_fail;      /* bucket winners that found the block heap empty      */
  uint32_t num_allocated;  /* live blocks after the last integrate                */
  uint32_t visible_blocks; /* blocks the last integrate visited                   */
  uint32_t updated_voxels; /* voxels the last integrate updated                   */
  uint32_t claimed_buckets;/* buckets with at least one insertion request         */
  uint32_t out_of_range;   /* requests for blocks beyond +-2^20 block indices     */
  uint32_t frames;         /* integrate calls since reset                         */
  uint32_t reserved[7];
} vsg_counters;

typedef struct vsg_scene vsg_scene;
typedef struct vsg_render vsg_render;

/* ---- library ------------------------------------------------------------------ */

/* Library version (major*10000 + minor*100 + patch). */
int32_t vsg_version(void);

/* Last error message of this host thread (static storage, never NULL). */
const char* vsg_last_error(void);

/* ---- scene: VoxelScene (VoxelSG/include/VoxelScene.h:52-209) -------------------- */

/* VoxelScene::VoxelScene(const Params&) + allocate() + reset()
 * (VoxelSG/src/VoxelScene.cpp:67-84,97-127,276-302).  Synchronising. */
vsg_status vsg_scene_create(const vsg_params* params, vsg_scene** out);

/* VoxelScene::~VoxelScene (VoxelSG/src/VoxelScene.cpp:89-92).  Synchronising. */
vsg_status vsg_scene_destroy(vsg_scene* scene);

/* VoxelScene::reset (VoxelSG/src/VoxelScene.cpp:276-302): empty hash table, full
 * block heap, zero voxels, frame count 0.  Asynchronous. */
vsg_status vsg_scene_reset(vsg_scene* scene);

/* Stream the scene's work is queued on (hipStream_t).  The reference uses the
 * legacy default stream and synchronises after every kernel; this library never
 * synchronises inside integrate/render. */
vsg_status vsg_scene_set_stream(vsg_scene* scene, void* hip_stream);
vsg_status vsg_scene_get_stream(const vsg_scene* scene, void** hip_stream);

/* VoxelScene::integrateFromScanData (VoxelSG/include/VoxelScene.h:89-91,
 * VoxelSG/src/VoxelScene.cpp:404-833): allocate the blocks along every valid
 * pixel's truncation band, list the blocks the frame can reach, integrate the
 * frame into them.  pose: HOST float[16] row-major camera->world (the reference
 * takes a cv::Mat and uploads it, VoxelScene.cpp:431-453).  Asynchronous. */
vsg_status vsg_integrate(vsg_scene* scene, const vsg_frame* frame, float truncation, const float pose[16],
                         const vsg_intrinsics* intrinsics);

/* vsg_integrate without any allocation: no hipMalloc, hipFree or allocating library
 * call, whatever the frame size, for callers that cannot allocate (code under HIP
 * graph capture, tight serving loops, harnesses that forbid it).  No counterpart in
 * the reference.  Same results as vsg_integrate.  Its per-pixel and per-workgroup
 * scratch is a static device array of the library (4.14 MiB, held by every GPU the
 * library runs on until the process ends), so frames are limited to 2^20 pixels
 * (VSG_ERR_INVALID_ARG beyond).  All calls on one GPU share that array: calls that
 * overlap on one GPU (scenes on different streams) race; serialise them.
 * Asynchronous. */
vsg_status vsg_integrate_noalloc(vsg_scene* scene, const vsg_frame* frame, float truncation, const float pose[16],
                                 const vsg_intrinsics* intrinsics);

/* Counters of the last integrate (synchronises the scene's stream). */
vsg_status vsg_scene_counters(vsg_scene* scene, vsg_counters* out);

/* VoxelScene::getHeapFreeCount (VoxelSG/src/VoxelScene.cpp:370-375): heap counter
 * + 1, i.e. N - allocated blocks.  Synchronising. */
vsg_status vsg_heap_free_count(vsg_scene* scene, uint32_t* out);

/* VoxelScene::getNumIntegratedFrames (VoxelSG/include/VoxelScene.h:104-106). */
vsg_status vsg_num_integrated_frames(const vsg_scene* scene, uint32_t* out);

/* Params the scene was created with (VoxelScene::getParams, VoxelScene.h:133-153). */
vsg_status vsg_scene_params(const vsg_scene* scene, vsg_params* out);

/* The scene's device arrays (VoxelScene::getHashData, VoxelScene.h:125-136, which
 * returns the reference's CUDAHashRef of device pointers).  The layout is this
 * library's, not the reference's HashSlot / VoxelData arrays:
 *   buckets      num_buckets records of 128 bytes: 10 uint64 keys (the block
 *                coordinates as 21-bit two's-complement fields, x << 42 | y << 21 | z),
 *                then 10 int32 block pointers; slots [0, bucket_used[b]) are live;
 *   bucket_used  num_buckets bytes;
 *   blocks       int32[4] per block pointer p < num_allocated: (x, y, z, bucket*10 + slot);
 *   voxels       num_blocks * 512 words of voxel_bytes, voxel order x fastest:
 *                FP32 store {float sdf, uint16 weight, uint16 0}, FP16 store
 *                {binary16 sdf | uint16 weight << 16};
 *   colors       num_blocks * 512 words {r, g, b, flags} (flags bit0 isUpdated, bit1
 *                isZeroCrossing);
 *   block_index  index_entries entries {uint32 key lo, key hi, pointer, positive-cell
 *                mask}, open addressing, all-ones key = empty;
 *   num_allocated_device  device uint32: live blocks (pointers 0 .. n-1).
 * The pointers stay valid until vsg_scene_destroy; writing through them bypasses the
 * raycaster's occupancy and cell summaries (use vsg_import_voxels to overwrite voxels). */
typedef struct vsg_hash_data {
  void* buckets;
  uint8_t* bucket_used;
  int32_t* blocks;
  void* voxels;
  uint32_t* colors;
  void* block_index;
  uint32_t* num_allocated_device;
  uint32_t num_buckets, num_blocks, voxel_bytes, index_entries;
  int32_t store; /* vsg_store */
} vsg_hash_data;
vsg_status vsg_scene_hash_data(const vsg_scene* scene, vsg_hash_data* out);

/* Allocated blocks in pointer order (block k has pointer k), copied to HOST
 * memory: coords[3*k..3*k+2] = block coordinate, slots[k] = hash entry index
 * (bucket*10 + slot).  The reference downloads its compacted list the same way
 * (VoxelSG/src/VoxelScene.cpp:540-560).  *count receives the number of
 * allocated blocks; at most `capacity` are written.  Synchronising. */
vsg_status vsg_export_blocks(vsg_scene* scene, int32_t* coords, int32_t* slots, uint32_t capacity,
                             uint32_t* count);

/* Voxels of blocks [first, first+nblocks) to HOST memory, 512 per block in the
 * reference's order (x fastest, VoxelScene.cu:955-957): sdf as float (FP16
 * store widened), weight, colour rgb + flags byte (bit0 isUpdated, bit1
 * isZeroCrossing, CUDAHashRef.h:23-45).  Any output may be NULL.  Synchronising. */
vsg_status vsg_export_voxels(vsg_scene* scene, uint32_t first, uint32_t nblocks, float* sdf, uint16_t* weight,
                             uint8_t* rgbf);

/* Direction slots of the voxels of blocks [first, first+nblocks) to HOST memory,
 * two per voxel (index 2*voxel + slot), voxels in vsg_export_voxels' order: id 0
 * free, 1..6 = +x, -x, +y, -y, +z, -z (world axes); weight <= 100; dtsdf the
 * running mean of the truncated SDF of the updates given to the slot.  An update
 * goes to the voxel's slot of the direction bin of the pixel's world normal (the
 * signed axis of its largest component) when that component is at least 0.8, else
 * to the first free slot, else nowhere.  VSG_ERR_INVALID_ARG when the scene was
 * created without VSG_SCENE_DIRECTIONAL.  Any output may be NULL.  Synchronising.
 * No counterpart in the compiled reference (extension of CUDAHashRef.h:34-39). */
vsg_status vsg_export_directional(vsg_scene* scene, uint32_t first, uint32_t nblocks, uint8_t* ids,
                                  uint8_t* weights, float* dtsdf);

/* ---- renderer: RayCastRender (VoxelSG/include/RayCastRender.h:20-123) ----------- */

/* RayCastRender::initialize(w, h) (VoxelSG/src/RayCastRender.cpp:88-118):
 * allocates depth (float), position, colour and normal (float4) outputs. */
vsg_status vsg_render_create(int32_t width, int32_t height, vsg_render** out);
vsg_status vsg_render_destroy(vsg_render* render);
vsg_status vsg_render_set_stream(vsg_render* render, void* hip_stream);

/* Trilinear sampling of the ray march.  VSG_SAMPLE_REFERENCE (default) is the
 * reference's scheme (RayCastRender.cu:139-176): weights frac(p/v - 1/2) with
 * corners rounded from p -+ v/2, which returns the stored field half a voxel off and
 * jumps at half-voxel planes.  VSG_SAMPLE_TRILINEAR interpolates the voxels at i*v
 * with weights frac(p/v): a continuous field for pose tracking.  Only the reference
 * scheme reproduces the reference's renders.  No counterpart in the reference. */
typedef enum vsg_sampling { VSG_SAMPLE_REFERENCE = 0, VSG_SAMPLE_TRILINEAR = 1 } vsg_sampling;
vsg_status vsg_render_set_sampling(vsg_render* render, int32_t mode);

/* Steps (skips or sample evaluations) a ray marches with one lane before the
 * cooperative tail pass continues it with several lanes; 0 marches every ray with
 * one lane to the end.  Results do not depend on it (default 28).  No counterpart
 * in the reference: the tail pass is this library's load balancing. */
vsg_status vsg_render_set_step_budget(vsg_render* render, uint32_t steps);

/* RayCastRender::render (VoxelSG/src/RayCastRender.cpp:123-183,
 * VoxelSG/src/RayCastRender.cu:407-620): one ray per pixel, marched in half-voxel
 * steps over [min_depth, max_depth] (z-depth), zero crossing refined by three
 * regula-falsi steps.  cam_pos and pose are HOST arrays (the reference passes the
 * pose as a device pointer: vsg_render_frame_device_pose).  voxel_size <= 0 uses the
 * scene's.  Work is queued on the renderer's stream after the scene's pending
 * work.  When the scene holds no block the outputs are left untouched
 * (RayCastRender.cu:553-556).  Asynchronous. */
vsg_status vsg_render_frame(vsg_render* render, vsg_scene* scene, const float cam_pos[3], const float pose[16],
                            const vsg_intrinsics* intrinsics, float min_depth, float max_depth, float voxel_size);

/* vsg_render_frame with the pose where the reference's render takes it
 * (RayCastRender.h:66-70, `float* d_cameraTransform`): pose_device is a DEVICE
 * pointer to 16 floats, row-major camera->world, that the kernels read when they run
 * (the reference copies nothing either).  cam_pos stays a host array (a float3 by
 * value in the reference).  The pose values must still be in place when the render
 * runs on the stream, and when a later vsg_render_adjoint of this render runs (it
 * reads them again).  No host synchronisation.  Asynchronous. */
vsg_status vsg_render_frame_device_pose(vsg_render* render, vsg_scene* scene, const float cam_pos[3],
                                        const float* pose_device, const vsg_intrinsics* intrinsics, float min_depth,
                                        float max_depth, float voxel_size);

/* Device output buffers (RayCastRender::getOutputDepth/Color/Normal/Positions,
 * RayCastRender.h:110-113): depth is float[W*H] (z-depth, 1e10 = no hit);
 * positions, colours and normals are float4[W*H]. */
vsg_status vsg_render_buffers(const vsg_render* render, float** depth, float** positions, float** colors,
                              float** normals);

/* downloadDepthFloat / downloadResults (RayCastRender.cpp:188-221) to HOST
 * memory; any output may be NULL.  Synchronising. */
vsg_status vsg_render_download(vsg_render* render, float* depth, float* positions, float* colors, float* normals);

/* Hit / miss counts of the last render (RayCastRender.cu:599-616).  Synchronising. */
vsg_status vsg_render_counters(vsg_render* render, uint32_t* hits, uint32_t* misses);

/* ---- meshing: marching cubes (VoxelScene::saveToMesh, VoxelScene.cpp:838-845, a stub
 *      in the reference; its unbuilt MarchingCubes.cu:352-467 sketches the kernel) ---- */

/* The triangle table (int8 [256][32]): case k (bit c set when corner c = (c & 1,
 * c >> 1 & 1, c >> 2 & 1) has sdf < iso) lists its triangles as edge triples, then -1.
 * Edge e joins corners {0,1},{2,3},{4,5},{6,7} (along x), {0,2},{1,3},{4,6},{5,7}
 * (along y), {0,4},{1,5},{2,6},{3,7} (along z).  Generated from the cube's geometry
 * (no copy of a published table): faces whose four edges change sign cut off each
 * inside corner; triangle normals (v1 - v0) x (v2 - v0) point to the outside corners.
 * Host-only (needs no GPU). */
vsg_status vsg_mc_table(int8_t* out);

typedef struct vsg_mesh vsg_mesh; /* MarchingCubesOutput (VoxelSG/include/MarchingCubes.h:19-29) */

/* allocateMarchingCubesOutput (VoxelSG/src/MarchingCubes.cu:474-489): room for
 * max_triangles triangles (3 float3 vertices and 3 float3 normals each, device memory),
 * the triangle table on the device and the scan scratch (sized at the first call on a
 * scene).  max_triangles must be positive, at most 2^28. */
vsg_status vsg_mesh_create(uint32_t max_triangles, vsg_mesh** out);

/* freeMarchingCubesOutput (MarchingCubes.cu:491-501). */
vsg_status vsg_mesh_destroy(vsg_mesh* mesh);

/* runMarchingCubesCUDA (MarchingCubes.cu:511-571, kernel :352-467) over every live
 * block of the scene, in hash-entry order (bucket, slot), then cube order (x fastest),
 * then table order (the reference walks the compactified list of the last frame and
 * appends in atomicAdd order).  Cube (x, y, z) spans voxels x..x+1, y..y+1, z..z+1 at
 * voxel centres (8 block + i + 0.5) voxel_size; cross_block = 0 keeps the reference's
 * cubes inside one block (x, y, z <= 6: gaps along block borders); cross_block = 1 adds
 * the cubes at 7, whose far corners come from the +x / +y / +z neighbours (skipped
 * where one is not allocated), for a closed surface.  Cubes with a corner of weight
 * below min_weight are skipped (0: none, as in the reference).  Case bit c when
 * sdf < iso (table of vsg_mc_table); vertex on the edge from corner a to b (a < b):
 * the midpoint when |sdf_b - sdf_a| < 1e-6, else pa + mu (pb - pa), mu = (iso - sdf_a)
 * / (sdf_b - sdf_a) clamped to [0, 1] (vertexInterp :313-327), so a vertex shared by
 * neighbouring cubes is the same float3 in both; the normal of all three vertices is
 * the unit (v1 - v0) x (v2 - v0), zero when its length is at most 1e-6 (computeNormal
 * :329-348).  voxel_size <= 0: the scene's (MarchingCubesParams::voxelSize, whose
 * default 1.0 the reference uses as given).  *total (may be NULL) = triangles found;
 * *stored (may be NULL) = min(total, max_triangles) written.  Enqueued on the scene's
 * stream; synchronising (the count is read back, as copyMarchingCubesTriangleCount
 * :503-509 does).  VSG_ERR_INVALID_ARG for a NaN iso or voxel size, or when the scene's
 * live blocks could hold 2^32 or more triangles (more than about 800,000 blocks). */
vsg_status vsg_marching_cubes(vsg_mesh* mesh, vsg_scene* scene, float iso, float voxel_size, int32_t cross_block,
                              uint32_t min_weight, uint32_t* stored, uint64_t* total);

/* Device buffers of the last run (MarchingCubesOutput::d_vertices / d_normals):
 * float3 [3 * max_triangles] each, triangle i at 3 i .. 3 i + 2; *count = triangles
 * stored.  Any output may be NULL. */
vsg_status vsg_mesh_buffers(const vsg_mesh* mesh, float** vertices, float** normals, uint32_t* count);

/* The first min(capacity, stored) triangles of the last run to HOST arrays of 9 floats
 * per triangle (vertices / normals, either may be NULL); *count = triangles copied.
 * Synchronising. */
vsg_status vsg_mesh_download(vsg_mesh* mesh, float* vertices, float* normals, uint32_t capacity, uint32_t* count);

/* ---- surface-point extraction ------------------------------------------------------- */

/* Capacity of the extraction output (m_maxExtractedPoints, RayCastRender.cpp:62:
 * default 10,000,000).  Takes effect at the next extraction (the buffers are
 * reallocated then). */
vsg_status vsg_render_set_max_points(vsg_render* render, uint32_t max_points);

/* RayCastRender::extractSurfacePoints (VoxelSG/src/RayCastRender.cpp:307-364,
 * kernelExtractSurfacePoints VoxelSG/src/RayCastRender.cu:684-868): every voxel with
 * weight >= max(1, min_weight) and SDF in [min_sdf, max_sdf] emits a point (the mean of
 * its zero crossings towards the 6 neighbours, or its centre when it has none and
 * |sdf| < 0.001), a normal (central differences of getVoxel, computeNormalFromSDF
 * :629-676) and its colour / 255.  Points are in hash-entry order (bucket, slot), then
 * voxel order (x fastest): the reference appends them in atomicAdd order, which no
 * run repeats.  *total (may be NULL) = points found; *stored (may be NULL) =
 * min(total, max points) written to the device buffers of vsg_extracted_points.  The
 * reference returns its atomic counter, which can exceed its buffer.  Synchronising
 * (the count is read back, as in the reference). */
vsg_status vsg_extract_surface_points(vsg_render* render, vsg_scene* scene, float min_sdf, float max_sdf,
                                      uint32_t min_weight, uint32_t* stored, uint64_t* total);

/* Device buffers of the last extraction (m_d_extractedPositions / Colors / Normals,
 * RayCastRender.h:33-36): float4[max points] each (position w = 1, colour a = 1,
 * normal w = 0); *count = points stored.  Any output may be NULL. */
vsg_status vsg_extracted_points(const vsg_render* render, float** positions, float** colors, float** normals,
                                uint32_t* count);

/* The first min(capacity, stored) points of the last extraction to HOST float4
 * arrays (any may be NULL); *count = points copied.  Synchronising. */
vsg_status vsg_extracted_points_download(vsg_render* render, float* positions, float* colors, float* normals,
                                         uint32_t capacity, uint32_t* count);

/* ---- depth maps from point clouds: CustomDepthMapGenerator
 *      (VoxelSG/include/CustomDepthMapGenerator.h, VoxelSG/src/CustomDepthMapGenerator.cpp) -- */

typedef struct vsg_depthgen vsg_depthgen;

/* CustomDepthMapGenerator::ProjectionMethod (CustomDepthMapGenerator.h:30-34); CUSTOM_CAMERA
 * projects as PERSPECTIVE, as in the reference (CustomDepthMapGenerator.cpp:564-573). */
typedef enum vsg_projection {
  VSG_PROJ_PERSPECTIVE = 0,
  VSG_PROJ_ORTHOGRAPHIC = 1,
  VSG_PROJ_CUSTOM_CAMERA = 2
} vsg_projection;

/* A generator for width x height maps (the reference's CameraParams width/height).
 * Its work runs on a stream of its own unless vsg_depthgen_set_stream gives one (give
 * the scene's, vsg_scene_get_stream, to integrate the maps without a host round trip). */
vsg_status vsg_depthgen_create(int32_t width, int32_t height, vsg_depthgen** out);
vsg_status vsg_depthgen_destroy(vsg_depthgen* gen);
vsg_status vsg_depthgen_set_stream(vsg_depthgen* gen, void* hip_stream);
/* setProjectionMethod (CustomDepthMapGenerator.cpp:76-79). */
vsg_status vsg_depthgen_set_projection(vsg_depthgen* gen, int32_t method);

/* generateFromPointCloud (VoxelSG/src/CustomDepthMapGenerator.cpp:81-222) on DEVICE
 * arrays: camera-frame points float3[n]; normals float3[n_normals] and colours
 * uchar3[n_colors] for the first points (NULL with a count of 0).  Each point with
 * z > 0 projecting inside the image lands on pixel (floor u, floor v); the last such
 * point in array order wins the pixel (the reference's loop order).  Outputs (device,
 * W*H): pointmap float3 (the point), depth float (its z; may be NULL), normalmap float3
 * (the normal of the last point with one that lands there), colormap uchar3 (likewise);
 * zeros where no point lands.  With n_normals == 0 the normals are the reference's
 * finite differences of the depth map (:164-208).  The outputs are vsg_frame's
 * inputs.  Asynchronous. */
vsg_status vsg_depthgen_from_points(vsg_depthgen* gen, const vsg_intrinsics* intrinsics, const float* points,
                                    uint32_t n, const float* normals, uint32_t n_normals, const uint8_t* colors,
                                    uint32_t n_colors, float* pointmap, float* depth, float* normalmap,
                                    uint8_t* colormap);

/* Pixels with depth > 0 of the last call (the reference prints it,
 * CustomDepthMapGenerator.cpp:211-219).  Synchronising. */
vsg_status vsg_depthgen_valid_pixels(vsg_depthgen* gen, uint32_t* count);

/* ---- differentiable raycast and pose refinement (no counterpart in the reference;
 *      the reference's renderer is forward only) ---------------------------------------- */

/* Adjoint of the last vsg_render_frame of this renderer.  Recomputes every pixel's
 * march and differentiates the rendered depth through the three regula-falsi steps
 * and the trilinear samples (the computation the forward performed; the sampled
 * field's jumps at half-voxel planes are not differentiated):
 *   jac        (device float[6*W*H] or NULL): d depth / d xi for the camera-frame
 *              perturbation pose * exp(xi^), xi = (omega, upsilon); 0 without a hit;
 *   grad_depth (device float[W*H]) with grad_sdf (device float[N*512]):
 *              grad_sdf[ptr*512 + voxel] += sum_pixels grad_depth * d depth / d sdf.
 * Asynchronous. */
vsg_status vsg_render_adjoint(vsg_render* render, vsg_scene* scene, float* jac, const float* grad_depth,
                              float* grad_sdf);

/* Jacobian of vsg_pose_gn_step.  VSG_GN_SMOOTHED (default): the rendered crossing
 * differentiated through the implicit function f(p(t)) = 0 with the field's gradient
 * taken by central differences over +-1 voxel at the crossing (pixels where a sample is
 * unweighted or the ray meets the surface at |cos| < 0.2 are left out).  VSG_GN_EXACT:
 * the derivative vsg_render_adjoint returns, of the forward's own computation; its
 * local trilinear gradient carries the noise of the fused field, which biases the
 * normal equations towards short steps. */
typedef enum vsg_gn_jacobian { VSG_GN_SMOOTHED = 0, VSG_GN_EXACT = 1 } vsg_gn_jacobian;
vsg_status vsg_render_set_gn_jacobian(vsg_render* render, int32_t mode);

/* Residual weights of vsg_pose_gn_step from the depth sensor's noise model
 * sigma(z) = a + b (z - z0)^2 of the observed depth z: each pixel counts
 * (a / sigma(z))^2.  a = 0 (default): all pixels count 1.  No counterpart in the
 * reference. */
vsg_status vsg_render_set_gn_noise(vsg_render* render, float a, float b, float z0);

/* Projective window of the renders of vsg_pose_gn_step / vsg_pose_gn_refine (on != 0;
 * off by default): each ray marches only over depths within the step's max_residual of
 * its pixel's observed depth (and within [min_depth, max_depth]), and pixels without an
 * observed depth are not marched.  A crossing outside the window would fail the
 * residual test anyway; what changes is that a nearer surface outside the window (an
 * occluder the full render would return, leaving the pixel out) is not seen, so the
 * pixel can count with the surface inside the window.  The march then covers a few
 * centimetres per ray instead of the whole depth range.  No counterpart in the
 * reference. */
vsg_status vsg_render_set_gn_window(vsg_render* render, int32_t on);

/* One Gauss-Newton step of frame-to-model pose refinement: renders at pose (camera
 * position = its translation), takes the Jacobian of vsg_render_set_gn_jacobian's mode,
 * accumulates in FP64
 * H = sum J J^T and g = sum J r over pixels with a rendered and an observed depth
 * (device float[W*H], 0 = invalid) and |r| < max_residual, r = rendered - observed;
 * solves (H + lambda diag H) delta = -g by FP64 Cholesky and composes
 * pose <- pose * exp(delta^) in FP64 on the device.  pose: HOST double[16] row-major
 * camera->world, updated in place.  stats (HOST double[8], may be NULL): cost before
 * the step (sum r^2), residual count, solved (1/0), |omega| (rad), |upsilon| (m).
 * Synchronising (the new pose is read back). */
vsg_status vsg_pose_gn_step(vsg_render* render, vsg_scene* scene, const float* observed_depth,
                            const vsg_intrinsics* intrinsics, float min_depth, float max_depth, float max_residual,
                            double lambda, double pose[16], double stats[8]);

/* `iterations` Gauss-Newton steps as vsg_pose_gn_step, step i with residual window
 * max_residuals[i] (HOST float[iterations]), with the pose kept on the device between
 * steps: every step's render reads the previous solve's pose from device memory, so
 * the steps queue without a host round trip and the pose is read back once at the end
 * (same results as as many vsg_pose_gn_step calls).  stats as vsg_pose_gn_step's, of
 * the last step.  The last render's pose is not kept, so vsg_render_adjoint reports
 * VSG_ERR_NOT_READY after a refinement of more than one step.  Synchronising.  No
 * counterpart in the reference. */
vsg_status vsg_pose_gn_refine(vsg_render* render, vsg_scene* scene, const float* observed_depth,
                              const vsg_intrinsics* intrinsics, float min_depth, float max_depth, uint32_t iterations,
                              const float* max_residuals, double lambda, double pose[16], double stats[8]);

/* Overwrites the voxels of blocks [first, first+nblocks) from HOST memory (the
 * layout vsg_export_voxels writes; rgbf may be NULL to keep colours and flags).
 * Synchronising. */
vsg_status vsg_import_voxels(vsg_scene* scene, uint32_t first, uint32_t nblocks, const float* sdf,
                             const uint16_t* weight, const uint8_t* rgbf);

/* The raycaster's positive-cell summary of blocks [first, first+nblocks) to HOST
 * memory: bit c of masks[b] is set when cell c of the block (4x4x4 voxels, c =
 * (x >> 2) | (y >> 2) << 1 | (z >> 2) << 2 for voxel (x, y, z) of the block) holds a
 * voxel with weight > 0 and SDF > 0; counts[8*b + c] is the number of such voxels.
 * The march skips samples whose corner voxels lie in cells without such a voxel up
 * to the last one before a cell that has one.  Either output may be NULL.
 * Synchronising.  No counterpart in the reference. */
vsg_status vsg_export_cell_masks(vsg_scene* scene, uint32_t first, uint32_t nblocks, uint8_t* masks,
                                 uint8_t* counts);

/* ---- streaming: VoxelStreamingManager (VoxelSG/include/VoxelStreamingManager.h:120-301) -- */

/* Bytes of one block image (the unit vsg_stream_out writes and vsg_stream_in reads):
 * the block's 512 voxel words as the store holds them (FP32 store: {f32 sdf, u32
 * weight}; FP16 store: {f16 sdf | u16 weight << 16}), then 512 colour words {r, g, b,
 * flags}, then with VSG_SCENE_DIRECTIONAL 512 slot words {id0, w0, id1, w1} and 512
 * float pairs (dtsdf of the two slots); voxels in vsg_export_voxels' order.  Replaces
 * the reference's SDFBlock (VoxelStreamingManager.h:33-42), 512 x 12-byte VoxelData. */
vsg_status vsg_block_image_bytes(const vsg_scene* scene, uint32_t* out);

/* Stream-out (VoxelStreamingManager::streamOutFindBlocksOnGPU + streamOutCopyBlocksCUDA,
 * VoxelSG/src/VoxelStreamingManager.cpp:303-500, VoxelStreamingManager.cu:36-91,
 * 258-441): every allocated block whose world position (coordinate * 8 * voxel size,
 * SDFBlockToWorld) lies at a distance >= radius from center is removed from the scene,
 * at most `capacity` of them (the first in the scene's block order): its coordinate
 * goes to coords[3*i..3*i+2] and its image to images + i * vsg_block_image_bytes
 * (HOST memory), its hash entry is freed and its block returns to the free store
 * with zeroed voxels.  *count = blocks removed; when it equals capacity more may
 * remain.  Synchronising.  The device scratch of vsg_stream_out and vsg_stream_in is
 * held by the scene between calls (as large as the largest call needed: for
 * vsg_stream_out about min(capacity, allocated blocks) images plus 13 bytes per
 * allocated block) and freed with it. */
vsg_status vsg_stream_out(vsg_scene* scene, const float center[3], float radius, uint32_t capacity,
                          int32_t* coords, void* images, uint32_t* count);

/* Stream-in (VoxelStreamingManager::streamInInsertToHashTable + streamInHashAllocCUDA /
 * streamInHashInitCUDA, VoxelSG/src/VoxelStreamingManager.cpp:566-833,
 * VoxelStreamingManager.cu:443-719): inserts n blocks (HOST coords and images in
 * vsg_stream_out's layout) in input order, as a sequential loop would: status[i] = 0
 * inserted, 1 already allocated (the image is not applied), 2 home bucket full (the
 * reference then probes other buckets, where its own lookups never look), 3 no free
 * block, 4 coordinate beyond the key range.  *inserted = blocks inserted.  status and
 * inserted may be NULL.  Synchronising. */
vsg_status vsg_stream_in(vsg_scene* scene, uint32_t n, const int32_t* coords, const void* images, uint8_t* status,
                         uint32_t* inserted);

/* ---- diagnostics (byte-model counters; no counterpart in the reference) ------------ */

typedef struct vsg_diag {
  uint32_t alloc_buckets; /* distinct buckets with live entries the allocation probed */
  uint32_t ray_buckets;   /* distinct buckets with live entries the raycast probed    */
  uint32_t any_buckets;   /* union of the two                                         */
  uint32_t ray_blocks;    /* distinct live blocks a raycast trilinear corner touched  */
  uint32_t march[8];      /* ray march: samples evaluated, super-block skips, block   */
                          /* skips, samples skipped by replaying the step additions,  */
                          /* positive-cell skips, rays deferred to the tail pass,     */
                          /* tail-pass sample windows, tail-pass skips                */
} vsg_diag;

/* When on, integrate and render run instrumented kernel variants that mark the
 * buckets and blocks they touch (the production kernels are not instrumented).
 * Turning it on allocates B + B + N bytes of marks and clears them. */
vsg_status vsg_diag_enable(vsg_scene* scene, int32_t on);

/* Counts the marks recorded since the last clear (synchronising); clear != 0
 * clears them afterwards. */
vsg_status vsg_diag_read(vsg_scene* scene, vsg_diag* out, int32_t clear);

/* The kernels' float -> int32 conversion applied on the device to n HOST floats
 * (truncation toward zero, out-of-range values saturated, NaN -> 0: the rule every
 * index computation of the path follows, the oracle's too), for verification.
 * Synchronising.  No counterpart in the reference. */
vsg_status vsg_diag_convert(const float* in, int32_t* out, uint32_t n);

/* ---- kernel timing ---------------------------------------------------------- */

/* Kernel ids for vsg_timing_read.  VSG_K_ALLOC_SCAN is no longer launched (the
 * insertion sums the resolve pass's workgroup counts itself): its slot reads zero
 * launches and keeps the other ids' values. */
enum {
  VSG_K_ALLOC_CLAIM = 0,
  VSG_K_ALLOC_RESOLVE = 1,
  VSG_K_ALLOC_SCAN = 2,
  VSG_K_ALLOC_INSERT = 3,
  VSG_K_COMPACT = 4,
  VSG_K_INTEGRATE = 5,
  VSG_K_RAYCAST = 6,
  VSG_K_RESET = 7,
  VSG_K_ADJOINT = 8,
  VSG_K_GN = 9,
  VSG_K_EXTRACT = 10,
  VSG_K_DEPTHGEN = 11,
  VSG_K_MESH = 12,
  VSG_K_COUNT = 13
};

/* When enabled, every kernel launch is bracketed by HIP events on the stream it
 * runs on (no host synchronisation).  vsg_timing_read synchronises, folds the
 * recorded events into per-kernel totals and returns, per kernel id, the total
 * milliseconds and the launch count since the last vsg_timing_reset. */
vsg_status vsg_timing_enable(int32_t on);
vsg_status vsg_timing_reset(void);
vsg_status vsg_timing_read(double ms[VSG_K_COUNT], uint32_t launches[VSG_K_COUNT]);

#ifdef __cplusplus
}
#endif

#endif /* VSG_H */
