// vsg_voxelscene.hpp -- the reference's C++ API (VoxelScene, RayCastRender, Params)
// over the C ABI of vsg.h, header-only.
//
// Same class names, method names, argument order and meaning as
//   VoxelSG/include/VoxelScene.h:52-209
//   VoxelSG/include/RayCastRender.h:20-123
//   VoxelSG/include/CUDAHashData.h:9-42
// so host code written against the reference compiles against this header and
// libvsg.so.  Differences, all at the edges of the API:
//   * the pose of integrateFromScanData is a host float[16] row-major
//     camera->world; the cv::Mat overload exists when VSG_HAVE_OPENCV is defined
//     (OpenCV is used by the reference only as a pose container);
//   * Params is filled directly instead of from GlobalParamsConfig;
//   * getHashData() returns vsg_hash_data, this library's device arrays (their layout
//     is documented in vsg.h), in place of the reference's CUDAHashRef;
//   * RayCastRender::render takes the device pose pointer and reads it in the kernels,
//     as the reference does; no host copy, no synchronisation;
//   * errors print and return, as the reference does (VoxelScene.cpp:420-428).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <unordered_map>
#include <vector>

#include "vsg.h"

#ifdef VSG_HAVE_OPENCV
#include <opencv2/core.hpp>
#endif

#define SDF_BLOCK_SIZE 8    // VoxelSG/include/globalPramasConfig.h:8
#define HASH_BUCKET_SIZE 10 // VoxelSG/include/globalPramasConfig.h:9

// Params (VoxelSG/include/CUDAHashData.h:9-42), field order kept.
struct Params {
  Params() = default;
  Params(int hashSlotNum_, int SDFBlockNum_, float voxelSize_, float truncation_ = 0.02f)
      : hashSlotNum(hashSlotNum_),
        SDFBlockNum(SDFBlockNum_),
        SDFBlockSize(SDF_BLOCK_SIZE),
        totalHashSize(hashSlotNum_ * HASH_BUCKET_SIZE),
        slotSize(HASH_BUCKET_SIZE),
        totalBlockSize(SDFBlockNum_ * SDF_BLOCK_SIZE * SDF_BLOCK_SIZE * SDF_BLOCK_SIZE),
        voxelSize(voxelSize_),
        truncationScale(0.0f),
        truncation(truncation_) {}

  int hashSlotNum = 0;
  int SDFBlockNum = 0;
  int SDFBlockSize = SDF_BLOCK_SIZE;
  int totalHashSize = 0;
  int slotSize = HASH_BUCKET_SIZE;
  int totalBlockSize = 0;
  float voxelSize = 0.0f;
  float truncationScale = 0.0f;
  float truncation = 0.0f;
};

namespace vsg_detail {
inline bool report(vsg_status st, const char* where) {
  if (st == VSG_OK) return true;
  std::fprintf(stderr, "%s: %s\n", where, vsg_last_error());
  return false;
}

// Binary little-endian PLY of a triangle soup (9 floats per triangle): vertices with
// equal bits merged in first-use order, faces as (uchar 3, int indices).
inline bool write_mesh_ply(const std::string& path, const float* tri, uint32_t n) {
  struct Key {
    uint32_t b[3];
    bool operator==(const Key& o) const { return b[0] == o.b[0] && b[1] == o.b[1] && b[2] == o.b[2]; }
  };
  struct Hash {
    size_t operator()(const Key& k) const {
      return (size_t)k.b[0] * 0x9E3779B97F4A7C15ull ^ (size_t)k.b[1] * 0xC2B2AE3D27D4EB4Full ^ (size_t)k.b[2];
    }
  };
  std::unordered_map<Key, int32_t, Hash> index;
  index.reserve((size_t)n * 2);
  std::vector<float> verts;
  std::vector<int32_t> faces((size_t)n * 3);
  for (size_t i = 0; i < (size_t)n * 3; ++i) {
    Key k;
    std::memcpy(k.b, tri + 3 * i, 12);
    auto it = index.find(k);
    if (it == index.end()) {
      it = index.emplace(k, (int32_t)(verts.size() / 3)).first;
      verts.insert(verts.end(), tri + 3 * i, tri + 3 * i + 3);
    }
    faces[i] = it->second;
  }
  std::FILE* f = std::fopen(path.c_str(), "wb");
  if (!f) {
    std::fprintf(stderr, "saveToMesh: cannot open %s\n", path.c_str());
    return false;
  }
  std::fprintf(f,
               "ply\nformat binary_little_endian 1.0\ncomment Generated by VoxelScene::saveToMesh\n"
               "element vertex %zu\nproperty float x\nproperty float y\nproperty float z\n"
               "element face %u\nproperty list uchar int vertex_indices\nend_header\n",
               verts.size() / 3, n);
  std::fwrite(verts.data(), sizeof(float), verts.size(), f);
  for (uint32_t t = 0; t < n; ++t) {
    const unsigned char three = 3;
    std::fwrite(&three, 1, 1, f);
    std::fwrite(&faces[3 * (size_t)t], sizeof(int32_t), 3, f);
  }
  return std::fclose(f) == 0;
}
}  // namespace vsg_detail

// VoxelScene (VoxelSG/include/VoxelScene.h:52-209)
class VoxelScene {
 public:
  // directional: also keep the two direction slots per voxel (vsg_export_directional)
  explicit VoxelScene(const Params& params, vsg_store store = VSG_STORE_FP32, bool directional = false)
      : m_params(params) {
    vsg_params p{};
    p.hash_slot_num = params.hashSlotNum;
    p.sdf_block_num = params.SDFBlockNum;
    p.voxel_size = params.voxelSize;
    p.store = store;
    p.flags = directional ? VSG_SCENE_DIRECTIONAL : 0u;
    m_isAllocated = vsg_detail::report(vsg_scene_create(&p, &m_scene), "VoxelScene");
  }
  ~VoxelScene() {
    if (m_scene) vsg_scene_destroy(m_scene);
  }
  VoxelScene(const VoxelScene&) = delete;
  VoxelScene& operator=(const VoxelScene&) = delete;

  void reset() {
    if (!m_isAllocated) {
      std::fprintf(stderr, "VoxelScene: Not allocated yet!\n");
      return;
    }
    vsg_detail::report(vsg_scene_reset(m_scene), "VoxelScene::reset");
  }

  // VoxelScene::integrateFromScanData (VoxelSG/include/VoxelScene.h:89-91): device
  // pointmap / colour map / normal map, host pose.
  void integrateFromScanData(const float3* depthmap, const uchar3* colormap, const float3* normalmap, int width,
                             int height, float truncationDistance, const float cameraTransform[16], float fx, float fy,
                             float cx, float cy) {
    if (depthmap == nullptr || colormap == nullptr || normalmap == nullptr) {
      std::fprintf(stderr, "VoxelScene: Invalid pointers (nullptr)!\n");
      return;
    }
    if (!m_isAllocated) {
      std::fprintf(stderr, "VoxelScene: Not allocated!\n");
      return;
    }
    vsg_frame f{depthmap, colormap, normalmap, width, height};
    vsg_intrinsics k{fx, fy, cx, cy};
    vsg_detail::report(vsg_integrate(m_scene, &f, truncationDistance, cameraTransform, &k),
                       "VoxelScene::integrateFromScanData");
  }

#ifdef VSG_HAVE_OPENCV
  void integrateFromScanData(const float3* depthmap, const uchar3* colormap, const float3* normalmap, int width,
                             int height, float truncationDistance, const cv::Mat& cameraTransform, float fx, float fy,
                             float cx, float cy) {
    float T[16];
    for (int i = 0; i < 4; ++i)
      for (int j = 0; j < 4; ++j) T[4 * i + j] = cameraTransform.at<float>(i, j);
    integrateFromScanData(depthmap, colormap, normalmap, width, height, truncationDistance, T, fx, fy, cx, cy);
  }
#endif

  // Compaction runs inside every integrate call.
  void compactifyHashEntries() {}

  unsigned int getHeapFreeCount() const {
    uint32_t n = 0;
    vsg_detail::report(vsg_heap_free_count(m_scene, &n), "VoxelScene::getHeapFreeCount");
    return n;
  }
  unsigned int getNumIntegratedFrames() const {
    uint32_t n = 0;
    vsg_num_integrated_frames(m_scene, &n);
    return n;
  }
  bool isAllocated() const { return m_isAllocated; }

  // VoxelScene::getHashData (VoxelSG/include/VoxelScene.h:125-136): the scene's device
  // arrays, in this library's layout (vsg_hash_data in vsg.h).
  vsg_hash_data getHashData() const {
    vsg_hash_data h{};
    vsg_detail::report(vsg_scene_hash_data(m_scene, &h), "VoxelScene::getHashData");
    return h;
  }
  Params& getParams() { return m_params; }
  const Params& getParams() const { return m_params; }

  void printStats() const {
    std::printf("\n=== VoxelScene Stats ===\nHash Slots: %d\nBucket Size: %d\nTotal Hash Entries: %d\n",
                m_params.hashSlotNum, m_params.slotSize, m_params.totalHashSize);
    std::printf("SDF Blocks: %d\nBlock Size: %dx%dx%d\nTotal Voxels: %d\nIntegrated Frames: %u\n",
                m_params.SDFBlockNum, m_params.SDFBlockSize, m_params.SDFBlockSize, m_params.SDFBlockSize,
                m_params.totalBlockSize, getNumIntegratedFrames());
    if (m_isAllocated) {
      unsigned int freeCount = getHeapFreeCount();
      std::printf("Free Blocks: %u / %d (%g%%)\n", freeCount, m_params.SDFBlockNum,
                  freeCount * 100.0 / m_params.SDFBlockNum);
    }
    std::printf("========================\n\n");
  }

  // VoxelScene::saveToMesh (a stub in the reference, VoxelSG/src/VoxelScene.cpp:838-845):
  // marching cubes over the whole scene (cubes across block borders included, cubes with
  // a corner of weight below minWeight left out), written as a binary little-endian PLY:
  // vertices shared by neighbouring triangles merged (equal float3, first-use order),
  // one face per triangle.  False when the scene has no triangles or on an error.
  bool saveToMesh(const std::string& filename, float isoValue = 0.0f, unsigned minWeight = 1) {
    if (!m_isAllocated) return false;
    uint32_t cap = 1u << 20, stored = 0;
    uint64_t total = 0;
    vsg_mesh* mesh = nullptr;
    for (int pass = 0; pass < 2; ++pass) {
      if (!vsg_detail::report(vsg_mesh_create(cap, &mesh), "VoxelScene::saveToMesh")) return false;
      if (!vsg_detail::report(vsg_marching_cubes(mesh, m_scene, isoValue, 0.0f, 1, minWeight, &stored, &total),
                              "VoxelScene::saveToMesh")) {
        vsg_mesh_destroy(mesh);
        return false;
      }
      if (total <= cap || total > (1u << 28)) break;
      vsg_mesh_destroy(mesh);
      mesh = nullptr;
      cap = (uint32_t)total;
    }
    if (stored == 0) {
      vsg_mesh_destroy(mesh);
      std::fprintf(stderr, "VoxelScene::saveToMesh: no triangles\n");
      return false;
    }
    std::vector<float> v((size_t)stored * 9);
    uint32_t got = 0;
    const bool ok = vsg_detail::report(vsg_mesh_download(mesh, v.data(), nullptr, stored, &got),
                                       "VoxelScene::saveToMesh");
    vsg_mesh_destroy(mesh);
    return ok && vsg_detail::write_mesh_ply(filename, v.data(), got);
  }

  vsg_scene* handle() const { return m_scene; }

 private:
  Params m_params;
  vsg_scene* m_scene = nullptr;
  bool m_isAllocated = false;
};

// VoxelStreamingManager (VoxelSG/include/VoxelStreamingManager.h:120-301): blocks far
// from a sphere centre move from the scene to a host chunk grid (gridDimensions
// chunks of voxelExtents metres from minGridPos, in chunk units) and back when the
// sphere reaches them.  A block belongs to the chunk of its world position
// (coordinate * 8 * voxel size) rounded half away from zero (worldToChunks,
// VoxelStreamingManager.cpp:905-922); blocks outside the grid are dropped, as the
// reference drops them (updateChunkGrid, :875-879), and counted.  The four stages
// keep the reference's names and order; they run synchronously on the caller's
// thread, so multiThreaded and streamingEnabled are accepted and ignored (the
// reference's worker thread and Windows events have no counterpart).
//
// Two departures from the reference's chunk lists (chunkData, VoxelStreamingManager.h:
// 78-118): a chunk holds one image per block coordinate, the newest (a block streamed
// out, re-created by integration and streamed out again replaces its older image, where
// the reference appends a second copy); and a chunk larger than what one stream-in may
// still collect is collected in parts over several calls, where the reference returns
// at that chunk (VoxelStreamingManager.cpp:1080-1085) and never collects it or any
// chunk after it.
class VoxelStreamingManager {
 public:
  VoxelStreamingManager(VoxelScene& scene, const Params& params, const float3& voxelExtents, const int3& gridDimensions,
                        const int3& minGridPos, unsigned int initialChunkListSize = 1000, bool streamingEnabled = false)
      : m_scene(&scene), m_params(params) {
    initialize(voxelExtents, gridDimensions, minGridPos, initialChunkListSize, streamingEnabled);
  }

  void initialize(const float3& voxelExtents, const int3& gridDimensions, const int3& minGridPos,
                  unsigned int initialChunkListSize, bool /*streamingEnabled*/) {
    m_voxelExtents = voxelExtents;
    m_gridDimensions = gridDimensions;
    m_minGridPos = minGridPos;
    m_maxGridPos = make_int3(minGridPos.x + gridDimensions.x, minGridPos.y + gridDimensions.y,
                             minGridPos.z + gridDimensions.z);
    m_initialChunkListSize = initialChunkListSize;
    m_grid.assign((size_t)gridDimensions.x * gridDimensions.y * gridDimensions.z, Chunk{});
    vsg_block_image_bytes(m_scene->handle(), &m_imageBytes);
  }

  // Removes the blocks at distance >= radius from sphereCenter from the scene (at
  // most m_maxNumberOfSDFBlocksIntegrateFromGlobalHash per call) and holds them for
  // streamOutCopyToChunkGrid.  Returns how many.
  unsigned int streamOutFindBlocksOnGPU(const float3& sphereCenter, float radius, bool /*useParts*/ = false,
                                        bool /*multiThreaded*/ = false) {
    const uint32_t cap = m_maxNumberOfSDFBlocksIntegrateFromGlobalHash;
    m_outCoords.resize((size_t)cap * 3);
    m_outImages.resize((size_t)cap * m_imageBytes);
    const float c[3] = {sphereCenter.x, sphereCenter.y, sphereCenter.z};
    uint32_t n = 0;
    vsg_detail::report(vsg_stream_out(m_scene->handle(), c, radius, cap, m_outCoords.data(), m_outImages.data(), &n),
                       "VoxelStreamingManager::streamOutFindBlocksOnGPU");
    nStreamdOutBlocks_ = n;
    return n;
  }

  void streamOutCopyToChunkGrid(bool /*multiThreaded*/ = false) {
    updateChunkGrid(m_outCoords.data(), m_outImages.data(), nStreamdOutBlocks_);
    nStreamdOutBlocks_ = 0;
  }

  // Collects the blocks of the streamed-out chunks the sphere touches (the chunk box
  // around the centre's chunk, isChunkInSphere) and removes them from the grid
  // (streamInFindChunk, VoxelStreamingManager.cpp:993-1130), at most
  // m_maxNumberOfSDFBlocksIntegrateFromGlobalHash: a chunk that does not fit whole
  // gives its first blocks and keeps the rest for the next call.  With useParts, one
  // chunk (or the part that fits) per call.  Returns how many.
  unsigned int streamInCopyToGPUBuffer(const float3& sphereCenter, float radius, bool useParts = false,
                                       bool /*multiThreaded*/ = false) {
    m_inCoords.clear();
    m_inImages.clear();
    const int3 cam = worldToChunks(sphereCenter);
    const int3 rad = make_int3((int)std::ceil(radius / m_voxelExtents.x), (int)std::ceil(radius / m_voxelExtents.y),
                               (int)std::ceil(radius / m_voxelExtents.z));
    const int3 lo = make_int3(std::max(cam.x - rad.x, m_minGridPos.x), std::max(cam.y - rad.y, m_minGridPos.y),
                              std::max(cam.z - rad.z, m_minGridPos.z));
    const int3 hi = make_int3(std::min(cam.x + rad.x, m_maxGridPos.x - 1), std::min(cam.y + rad.y, m_maxGridPos.y - 1),
                              std::min(cam.z + rad.z, m_maxGridPos.z - 1));
    for (int x = lo.x; x <= hi.x; ++x)
      for (int y = lo.y; y <= hi.y; ++y)
        for (int z = lo.z; z <= hi.z; ++z) {
          const int3 ch = make_int3(x, y, z);
          Chunk& c = m_grid[linearizeChunkPos(ch)];
          if (c.coords.empty() || !isChunkInSphere(ch, sphereCenter, radius)) continue;
          const size_t room = m_maxNumberOfSDFBlocksIntegrateFromGlobalHash - nStreamdInBlocks();
          const size_t take = std::min(c.coords.size() / 3, room);
          m_inCoords.insert(m_inCoords.end(), c.coords.begin(), c.coords.begin() + 3 * take);
          m_inImages.insert(m_inImages.end(), c.images.begin(), c.images.begin() + take * m_imageBytes);
          c.dropFront(take, m_imageBytes);
          if (useParts || nStreamdInBlocks() == m_maxNumberOfSDFBlocksIntegrateFromGlobalHash)
            return nStreamdInBlocks();
        }
    return nStreamdInBlocks();
  }

  // Inserts the collected blocks; blocks that find their bucket full or no free block
  // go back to the chunk grid.  Returns how many were inserted.
  unsigned int streamInInsertToHashTable(bool /*multiThreaded*/ = false) {
    const uint32_t n = nStreamdInBlocks();
    if (n == 0) return 0;
    std::vector<uint8_t> status(n);
    uint32_t inserted = 0;
    vsg_detail::report(vsg_stream_in(m_scene->handle(), n, m_inCoords.data(), m_inImages.data(), status.data(),
                                     &inserted),
                       "VoxelStreamingManager::streamInInsertToHashTable");
    std::vector<int32_t> bc;
    std::vector<uint8_t> bi;
    for (uint32_t i = 0; i < n; ++i)
      if (status[i] == 2 || status[i] == 3) {
        bc.insert(bc.end(), &m_inCoords[3 * (size_t)i], &m_inCoords[3 * (size_t)i] + 3);
        bi.insert(bi.end(), &m_inImages[(size_t)i * m_imageBytes], &m_inImages[(size_t)(i + 1) * m_imageBytes]);
      }
    if (!bc.empty()) updateChunkGrid(bc.data(), bi.data(), (uint32_t)(bc.size() / 3));
    m_inCoords.clear();
    m_inImages.clear();
    return inserted;
  }

  // One-call forms (VoxelStreamingManager.h: streamOut / streamIn).
  void streamOut(VoxelScene& /*scene*/, const float3& cameraPos, float radius) {
    for (;;) {
      const unsigned int n = streamOutFindBlocksOnGPU(cameraPos, radius);
      streamOutCopyToChunkGrid();
      if (n < m_maxNumberOfSDFBlocksIntegrateFromGlobalHash) return;
    }
  }
  // As many collect + insert passes as the cap needs: a pass that fills the cap and
  // inserts something is followed by another.
  void streamIn(VoxelScene& /*scene*/, const float3& cameraPos, float radius) {
    for (;;) {
      const unsigned int n = streamInCopyToGPUBuffer(cameraPos, radius);
      const unsigned int inserted = streamInInsertToHashTable();
      if (n < m_maxNumberOfSDFBlocksIntegrateFromGlobalHash || inserted == 0) return;
    }
  }

  size_t numStreamedBlocks() const {
    size_t n = 0;
    for (const Chunk& c : m_grid) n += c.coords.size() / 3;
    return n;
  }
  size_t droppedOutsideGrid() const { return m_dropped; }

  unsigned int m_maxNumberOfSDFBlocksIntegrateFromGlobalHash = 100000;  // VoxelStreamingManager.cpp:113

 private:
  // One image per block coordinate, in first-stream-out order; `where` maps a block's
  // packed coordinate to its position in coords/images.
  struct Chunk {
    std::vector<int32_t> coords;  // 3 per block
    std::vector<uint8_t> images;  // vsg_block_image_bytes per block
    std::unordered_map<uint64_t, uint32_t> where;

    // 21 bits per axis, as the scene's own keys: injective over the coordinates a
    // scene can hold ([-2^20, 2^20) per axis).
    static uint64_t key(const int32_t* c) {
      return ((uint64_t)(c[0] & 0x1FFFFF) << 42) | ((uint64_t)(c[1] & 0x1FFFFF) << 21) | (uint64_t)(c[2] & 0x1FFFFF);
    }
    void put(const int32_t* c, const uint8_t* image, size_t bytes) {
      const auto it = where.find(key(c));
      if (it != where.end()) {
        std::copy(image, image + bytes, images.begin() + (size_t)it->second * bytes);
        return;
      }
      where.emplace(key(c), (uint32_t)(coords.size() / 3));
      coords.insert(coords.end(), c, c + 3);
      images.insert(images.end(), image, image + bytes);
    }
    void dropFront(size_t k, size_t bytes) {
      if (k == coords.size() / 3) {
        *this = Chunk{};
        return;
      }
      coords.erase(coords.begin(), coords.begin() + 3 * k);
      images.erase(images.begin(), images.begin() + k * bytes);
      where.clear();
      for (size_t i = 0; i < coords.size() / 3; ++i) where.emplace(key(&coords[3 * i]), (uint32_t)i);
    }
  };

  unsigned int nStreamdInBlocks() const { return (unsigned int)(m_inCoords.size() / 3); }

  void updateChunkGrid(const int32_t* coords, const uint8_t* images, uint32_t n) {
    for (uint32_t i = 0; i < n; ++i) {
      const float v = m_params.voxelSize;
      const float3 w = make_float3((float)(coords[3 * i] * SDF_BLOCK_SIZE) * v,
                                   (float)(coords[3 * i + 1] * SDF_BLOCK_SIZE) * v,
                                   (float)(coords[3 * i + 2] * SDF_BLOCK_SIZE) * v);
      const int3 ch = worldToChunks(w);
      if (!isValidChunk(ch)) {
        ++m_dropped;
        continue;
      }
      Chunk& c = m_grid[linearizeChunkPos(ch)];
      if (c.coords.empty()) {
        c.coords.reserve((size_t)m_initialChunkListSize * 3);
      }
      c.put(coords + 3 * (size_t)i, images + (size_t)i * m_imageBytes, m_imageBytes);
    }
  }

  int3 worldToChunks(const float3& p) const {
    const float x = p.x / m_voxelExtents.x, y = p.y / m_voxelExtents.y, z = p.z / m_voxelExtents.z;
    auto sgn = [](float q) { return q > 0.0f ? 1.0f : (q < 0.0f ? -1.0f : 0.0f); };
    return make_int3((int)(x + sgn(x) * 0.5f), (int)(y + sgn(y) * 0.5f), (int)(z + sgn(z) * 0.5f));
  }
  bool isValidChunk(const int3& c) const {
    return c.x >= m_minGridPos.x && c.y >= m_minGridPos.y && c.z >= m_minGridPos.z && c.x < m_maxGridPos.x &&
           c.y < m_maxGridPos.y && c.z < m_maxGridPos.z;
  }
  size_t linearizeChunkPos(const int3& c) const {
    return (size_t)(c.z - m_minGridPos.z) * m_gridDimensions.x * m_gridDimensions.y +
           (size_t)(c.y - m_minGridPos.y) * m_gridDimensions.x + (size_t)(c.x - m_minGridPos.x);
  }
  bool isChunkInSphere(const int3& c, const float3& center, float radius) const {
    const float ext = std::max(std::max(m_voxelExtents.x, m_voxelExtents.y), m_voxelExtents.z);
    const float chunkRadius = 0.5f * ext * std::sqrt(3.0f);
    const float dx = c.x * m_voxelExtents.x - center.x, dy = c.y * m_voxelExtents.y - center.y,
                dz = c.z * m_voxelExtents.z - center.z;
    return std::sqrt(dx * dx + dy * dy + dz * dz) - chunkRadius <= radius;
  }

  VoxelScene* m_scene;
  Params m_params;
  float3 m_voxelExtents{};
  int3 m_gridDimensions{}, m_minGridPos{}, m_maxGridPos{};
  unsigned int m_initialChunkListSize = 1000;
  uint32_t m_imageBytes = 0;
  std::vector<Chunk> m_grid;
  std::vector<int32_t> m_outCoords, m_inCoords;
  std::vector<uint8_t> m_outImages, m_inImages;
  uint32_t nStreamdOutBlocks_ = 0;
  size_t m_dropped = 0;
};

// RayCastRender (VoxelSG/include/RayCastRender.h:20-123)
class RayCastRender {
 public:
  RayCastRender() = default;
  ~RayCastRender() {
    if (m_render) vsg_render_destroy(m_render);
    if (m_extractOnly) vsg_render_destroy(m_extractOnly);
  }
  RayCastRender(const RayCastRender&) = delete;
  RayCastRender& operator=(const RayCastRender&) = delete;

  bool initialize(int width, int height) {
    if (m_isAllocated) {
      std::fprintf(stderr, "RayCastRender: Already initialized!\n");
      return false;
    }
    if (!vsg_detail::report(vsg_render_create(width, height, &m_render), "RayCastRender: Failed to allocate GPU memory!"))
      return false;
    vsg_render_set_max_points(m_render, (uint32_t)m_maxExtractedPoints);
    m_width = width;
    m_height = height;
    m_isAllocated = true;
    return true;
  }

  // RayCastRender::render (VoxelSG/include/RayCastRender.h:66-70): the pose is a
  // DEVICE float[16] row-major camera->world, as the reference's driver passes it.
  // The kernels read the 16 floats when they run, so the buffer must hold the pose
  // until then (the reference's driver keeps it for the whole run).
  void render(const VoxelScene* voxelScene, float3 cameraPos, float* cameraTransform, float fx, float fy, float cx,
              float cy, float minDepth, float maxDepth) {
    if (!m_isAllocated) {
      std::fprintf(stderr, "RayCastRender: Not initialized!\n");
      return;
    }
    if (!voxelScene || !voxelScene->isAllocated()) {
      std::fprintf(stderr, "RayCastRender: Invalid voxel scene!\n");
      return;
    }
    const float cam[3] = {cameraPos.x, cameraPos.y, cameraPos.z};
    vsg_intrinsics k{fx, fy, cx, cy};
    vsg_detail::report(vsg_render_frame_device_pose(m_render, voxelScene->handle(), cam, cameraTransform, &k, minDepth,
                                                    maxDepth, voxelScene->getParams().voxelSize),
                       "RayCastRender::render");
  }

  // Same, with the pose in host memory (no copy).
  void renderHostPose(const VoxelScene* voxelScene, float3 cameraPos, const float cameraTransform[16], float fx,
                      float fy, float cx, float cy, float minDepth, float maxDepth) {
    if (!m_isAllocated) {
      std::fprintf(stderr, "RayCastRender: Not initialized!\n");
      return;
    }
    if (!voxelScene || !voxelScene->isAllocated()) {
      std::fprintf(stderr, "RayCastRender: Invalid voxel scene!\n");
      return;
    }
    const float cam[3] = {cameraPos.x, cameraPos.y, cameraPos.z};
    vsg_intrinsics k{fx, fy, cx, cy};
    vsg_detail::report(vsg_render_frame(m_render, voxelScene->handle(), cam, cameraTransform, &k, minDepth, maxDepth,
                                        voxelScene->getParams().voxelSize),
                       "RayCastRender::render");
  }

  // RayCastRender::downloadResults (RayCastRender.cpp:188-214): depth replicated
  // into xyz with w = 1, colours and normals as rendered.
  void downloadResults(float4* h_outputDepth, float4* h_outputColor, float4* h_outputNormal) {
    if (!m_isAllocated) {
      std::fprintf(stderr, "RayCastRender: Not initialized!\n");
      return;
    }
    std::vector<float> d((size_t)m_width * m_height);
    vsg_detail::report(vsg_render_download(m_render, d.data(), nullptr, (float*)h_outputColor, (float*)h_outputNormal),
                       "RayCastRender::downloadResults");
    for (size_t i = 0; i < d.size(); ++i) h_outputDepth[i] = make_float4(d[i], d[i], d[i], 1.0f);
  }

  void downloadDepthFloat(float* h_outputDepthFloat) {
    if (!m_isAllocated || h_outputDepthFloat == nullptr) return;
    vsg_render_download(m_render, h_outputDepthFloat, nullptr, nullptr, nullptr);
  }

  // RayCastRender::savePointCloudPLY (RayCastRender.cpp:221-302): ASCII PLY of the
  // hit pixels' world positions and colours.
  bool savePointCloudPLY(const std::string& filePath) {
    if (!m_isAllocated) {
      std::fprintf(stderr, "RayCastRender: Not initialized!\n");
      return false;
    }
    const size_t n = (size_t)m_width * m_height;
    std::vector<float> d(n);
    std::vector<float4> p(n), c(n);
    if (!vsg_detail::report(vsg_render_download(m_render, d.data(), (float*)p.data(), (float*)c.data(), nullptr),
                            "RayCastRender::savePointCloudPLY"))
      return false;
    size_t cnt = 0;
    for (size_t i = 0; i < n; ++i) cnt += d[i] < 1e9f ? 1 : 0;
    FILE* fp = std::fopen(filePath.c_str(), "w");
    if (!fp) return false;
    std::fprintf(fp, "ply\nformat ascii 1.0\nelement vertex %zu\n", cnt);
    std::fprintf(fp, "property float x\nproperty float y\nproperty float z\n");
    std::fprintf(fp, "property uchar red\nproperty uchar green\nproperty uchar blue\nend_header\n");
    for (size_t i = 0; i < n; ++i) {
      if (!(d[i] < 1e9f)) continue;
      auto u8 = [](float v) { return (int)(v < 0.f ? 0.f : (v > 1.f ? 255.f : v * 255.f)); };
      std::fprintf(fp, "%g %g %g %d %d %d\n", p[i].x, p[i].y, p[i].z, u8(c[i].x), u8(c[i].y), u8(c[i].z));
    }
    std::fclose(fp);
    return true;
  }

  // RayCastRender::extractSurfacePoints (RayCastRender.cpp:307-364): points in
  // hash-entry order, then voxel order.  Returns the number found (the reference's
  // counter); the first getMaxExtractedPoints() are kept.  Works without initialize(),
  // as in the reference.
  int extractSurfacePoints(const VoxelScene* voxelScene, float minSDF = -0.01f, float maxSDF = 0.01f,
                           unsigned int minWeight = 1) {
    if (!voxelScene || !voxelScene->isAllocated()) {
      std::fprintf(stderr, "RayCastRender::extractSurfacePoints: Invalid voxelScene!\n");
      return 0;
    }
    vsg_render* r = extractor();
    if (!r) return 0;
    uint32_t stored = 0;
    uint64_t total = 0;
    if (!vsg_detail::report(vsg_extract_surface_points(r, voxelScene->handle(), minSDF, maxSDF, minWeight, &stored,
                                                       &total),
                            "RayCastRender::extractSurfacePoints"))
      return 0;
    m_numExtractedPoints = (int)stored;
    return total > 0x7FFFFFFFull ? 0x7FFFFFFF : (int)total;
  }

  // RayCastRender::saveExtractedSurfacePointsPLY (RayCastRender.cpp:370-446): ASCII
  // PLY of the extracted points with finite positions (x y z nx ny nz red green blue).
  bool saveExtractedSurfacePointsPLY(const std::string& filePath) {
    vsg_render* r = m_isAllocated ? m_render : m_extractOnly;
    if (!r || m_numExtractedPoints == 0) {
      std::fprintf(stderr, "RayCastRender::saveExtractedSurfacePointsPLY: No points to save!\n");
      return false;
    }
    const size_t n = (size_t)m_numExtractedPoints;
    std::vector<float4> p(n), c(n), nr(n);
    uint32_t got = 0;
    if (!vsg_detail::report(vsg_extracted_points_download(r, (float*)p.data(), (float*)c.data(), (float*)nr.data(),
                                                          (uint32_t)n, &got),
                            "RayCastRender::saveExtractedSurfacePointsPLY"))
      return false;
    auto finite3 = [](const float4& v) { return std::isfinite(v.x) && std::isfinite(v.y) && std::isfinite(v.z); };
    size_t valid = 0;
    for (size_t i = 0; i < got; ++i) valid += finite3(p[i]) ? 1 : 0;
    FILE* fp = std::fopen(filePath.c_str(), "w");
    if (!fp) {
      std::fprintf(stderr, "RayCastRender::saveExtractedSurfacePointsPLY: Failed to open file: %s\n",
                   filePath.c_str());
      return false;
    }
    std::fprintf(fp, "ply\nformat ascii 1.0\ncomment Generated by RayCastRender::extractSurfacePoints\n");
    std::fprintf(fp, "element vertex %zu\n", valid);
    std::fprintf(fp, "property float x\nproperty float y\nproperty float z\n");
    std::fprintf(fp, "property float nx\nproperty float ny\nproperty float nz\n");
    std::fprintf(fp, "property uchar red\nproperty uchar green\nproperty uchar blue\nend_header\n");
    auto chan = [](float v) -> int {
      if (!std::isfinite(v)) return 255;
      v = v < 0.0f ? 0.0f : (v > 1.0f ? 1.0f : v);
      return (int)(v * 255.0f + 0.5f);
    };
    for (size_t i = 0; i < got; ++i) {
      if (!finite3(p[i])) continue;
      std::fprintf(fp, "%g %g %g %g %g %g %d %d %d\n", p[i].x, p[i].y, p[i].z, nr[i].x, nr[i].y, nr[i].z,
                   chan(c[i].x), chan(c[i].y), chan(c[i].z));
    }
    std::fclose(fp);
    return true;
  }

  // Capacity of the extraction (m_maxExtractedPoints, RayCastRender.cpp:62; 10M).
  void setMaxExtractedPoints(int maxPoints) {
    m_maxExtractedPoints = maxPoints < 0 ? 0 : maxPoints;
    if (m_render) vsg_render_set_max_points(m_render, (uint32_t)m_maxExtractedPoints);
    if (m_extractOnly) vsg_render_set_max_points(m_extractOnly, (uint32_t)m_maxExtractedPoints);
  }
  int getMaxExtractedPoints() const { return m_maxExtractedPoints; }
  int getNumExtractedPoints() const { return m_numExtractedPoints; }

  // Device buffers of the extracted points (float4 each).
  float4* getExtractedPositions() const { return (float4*)ext(0); }
  float4* getExtractedColors() const { return (float4*)ext(1); }
  float4* getExtractedNormals() const { return (float4*)ext(2); }

  float4* getOutputDepth() const { return (float4*)buf(0); }  // a float[W*H], as in the reference
  float4* getOutputColor() const { return (float4*)buf(2); }
  float4* getOutputNormal() const { return (float4*)buf(3); }
  float4* getOutputPositions() const { return (float4*)buf(1); }
  int getWidth() const { return m_width; }
  int getHeight() const { return m_height; }
  bool isAllocated() const { return m_isAllocated; }

  // Hit / miss counts of the last render (printed by the reference, RayCastRender.cu:599-616).
  void getCounters(unsigned int* hits, unsigned int* misses) const { vsg_render_counters(m_render, hits, misses); }

 private:
  float* buf(int i) const {
    float* b[4] = {nullptr, nullptr, nullptr, nullptr};
    if (m_render) vsg_render_buffers(m_render, &b[0], &b[1], &b[2], &b[3]);
    return b[i];
  }
  // the renderer extractions run on: the initialised one, else a 1x1 one of its own
  vsg_render* extractor() {
    if (m_isAllocated) return m_render;
    if (!m_extractOnly) {
      if (!vsg_detail::report(vsg_render_create(1, 1, &m_extractOnly), "RayCastRender::extractSurfacePoints"))
        return nullptr;
      vsg_render_set_max_points(m_extractOnly, (uint32_t)m_maxExtractedPoints);
    }
    return m_extractOnly;
  }
  float* ext(int i) const {
    float* b[3] = {nullptr, nullptr, nullptr};
    vsg_render* r = m_isAllocated ? m_render : m_extractOnly;
    if (r) vsg_extracted_points(r, &b[0], &b[1], &b[2], nullptr);
    return b[i];
  }
  vsg_render* m_render = nullptr;
  vsg_render* m_extractOnly = nullptr;
  int m_width = 0, m_height = 0;
  bool m_isAllocated = false;
  int m_maxExtractedPoints = 10000000;
  int m_numExtractedPoints = 0;
};

// CustomDepthMapGenerator (VoxelSG/include/CustomDepthMapGenerator.h:15-144): depth maps
// from camera-frame point clouds on the GPU (the reference's own generator runs on the
// CPU; its CUDA path is a TODO, CustomDepthMapGenerator.cpp:593-611).  Points, normals
// and colours are float3 / float3 / uchar3 arrays (cv::Point3f and cv::Vec3b have the
// same layout; the cv overloads take them when VSG_HAVE_OPENCV is defined).
class CustomDepthMapGenerator {
 public:
  struct CameraParams {  // CustomDepthMapGenerator.h:18-27, same defaults
    float fx = 400.0f, fy = 400.0f, cx = 200.0f, cy = 240.0f;
    float nearPlane = 0.1f, farPlane = 10.0f;
    int width = 400, height = 480;
  };
  enum class ProjectionMethod { PERSPECTIVE, ORTHOGRAPHIC, CUSTOM_CAMERA };
  struct GeneratedDepthMap {  // CustomDepthMapGenerator.h:37-45 (host arrays, W*H each)
    std::vector<float> depthmap;
    std::vector<float3> pointmap;
    std::vector<float3> normalmap;
    std::vector<uchar3> colormap;
    int width = 0, height = 0;
  };

  CustomDepthMapGenerator() = default;
  ~CustomDepthMapGenerator() { release(); }
  CustomDepthMapGenerator(const CustomDepthMapGenerator&) = delete;
  CustomDepthMapGenerator& operator=(const CustomDepthMapGenerator&) = delete;

  void setCameraParams(const CameraParams& params) { m_cam = params; }
  void setCameraParams(float fx, float fy, float cx, float cy, float nearPlane, float farPlane, int width, int height) {
    m_cam.fx = fx;
    m_cam.fy = fy;
    m_cam.cx = cx;
    m_cam.cy = cy;
    m_cam.nearPlane = nearPlane;
    m_cam.farPlane = farPlane;
    m_cam.width = width;
    m_cam.height = height;
  }
  const CameraParams& getCameraParams() const { return m_cam; }
  void setProjectionMethod(ProjectionMethod method) { m_method = method; }
  ProjectionMethod getProjectionMethod() const { return m_method; }
  // Queue the generator's work on this stream (e.g. a VoxelScene's, to integrate the
  // device maps without a host round trip).
  void setStream(hipStream_t stream) {
    m_stream = stream;
    m_haveStream = true;
    if (m_gen) vsg_depthgen_set_stream(m_gen, (void*)stream);
  }

  // generateFromPointCloud on DEVICE arrays into DEVICE maps (W*H each; depth may be
  // null).  Asynchronous.  Returns false on error.
  bool generateDevice(const float3* d_points, size_t n, const float3* d_normals, size_t nNormals,
                      const uchar3* d_colors, size_t nColors, float3* d_pointmap, float* d_depth, float3* d_normalmap,
                      uchar3* d_colormap) {
    vsg_depthgen* g = gen();
    if (!g) return false;
    vsg_intrinsics k{m_cam.fx, m_cam.fy, m_cam.cx, m_cam.cy};
    return vsg_detail::report(
        vsg_depthgen_from_points(g, &k, (const float*)d_points, (uint32_t)n, (const float*)d_normals,
                                 (uint32_t)nNormals, (const uint8_t*)d_colors, (uint32_t)nColors, (float*)d_pointmap,
                                 d_depth, (float*)d_normalmap, (uint8_t*)d_colormap),
        "CustomDepthMapGenerator::generateFromPointCloud");
  }

  // CustomDepthMapGenerator::generateFromPointCloud (CustomDepthMapGenerator.cpp:81-222)
  // on host arrays.
  GeneratedDepthMap generateFromPointCloud(const std::vector<float3>& points,
                                           const std::vector<float3>& normals = std::vector<float3>(),
                                           const std::vector<uchar3>& colors = std::vector<uchar3>()) {
    return generateHost(points.data(), points.size(), normals.data(), normals.size(), colors.data(), colors.size());
  }

#ifdef VSG_HAVE_OPENCV
  GeneratedDepthMap generateFromPointCloud(const std::vector<cv::Point3f>& points,
                                           const std::vector<cv::Point3f>& normals = std::vector<cv::Point3f>(),
                                           const std::vector<cv::Vec3b>& colors = std::vector<cv::Vec3b>()) {
    static_assert(sizeof(cv::Point3f) == sizeof(float3) && sizeof(cv::Vec3b) == sizeof(uchar3), "layout");
    return generateHost((const float3*)points.data(), points.size(), (const float3*)normals.data(), normals.size(),
                        (const uchar3*)colors.data(), colors.size());
  }
#endif

  // Valid pixels of the last generation (depth > 0).  Synchronising.
  int validPixels() {
    uint32_t c = 0;
    if (m_gen) vsg_depthgen_valid_pixels(m_gen, &c);
    return (int)c;
  }

 private:
  GeneratedDepthMap generateHost(const float3* pts, size_t n, const float3* nrm, size_t nn, const uchar3* col,
                                 size_t nc) {
    GeneratedDepthMap r;
    const size_t npix = (size_t)m_cam.width * m_cam.height;
    vsg_depthgen* g = gen();
    if (!g || n >= 0xFFFFFFFFull) return r;
    void *dp = nullptr, *dn = nullptr, *dc = nullptr, *out = nullptr;
    const size_t bytes_out = npix * (3 * sizeof(float) + sizeof(float) + 3 * sizeof(float) + 3);
    bool ok = hipMalloc(&out, bytes_out) == hipSuccess;
    if (ok && n) ok = hipMalloc(&dp, n * sizeof(float3)) == hipSuccess;
    if (ok && nn) ok = hipMalloc(&dn, nn * sizeof(float3)) == hipSuccess;
    if (ok && nc) ok = hipMalloc(&dc, nc * sizeof(uchar3)) == hipSuccess;
    if (ok && n) ok = hipMemcpy(dp, pts, n * sizeof(float3), hipMemcpyHostToDevice) == hipSuccess;
    if (ok && nn) ok = hipMemcpy(dn, nrm, nn * sizeof(float3), hipMemcpyHostToDevice) == hipSuccess;
    if (ok && nc) ok = hipMemcpy(dc, col, nc * sizeof(uchar3), hipMemcpyHostToDevice) == hipSuccess;
    float3* opts = (float3*)out;
    float* odep = (float*)(opts + npix);
    float3* onrm = (float3*)(odep + npix);
    uchar3* ocol = (uchar3*)(onrm + npix);
    if (ok) ok = generateDevice((const float3*)dp, n, (const float3*)dn, nn, (const uchar3*)dc, nc, opts, odep, onrm, ocol);
    uint32_t valid = 0;
    if (ok) ok = vsg_depthgen_valid_pixels(m_gen, &valid) == VSG_OK;  // synchronises the generator's stream
    if (ok) {
      r.width = m_cam.width;
      r.height = m_cam.height;
      r.pointmap.resize(npix);
      r.depthmap.resize(npix);
      r.normalmap.resize(npix);
      r.colormap.resize(npix);
      ok = hipMemcpy(r.pointmap.data(), opts, npix * sizeof(float3), hipMemcpyDeviceToHost) == hipSuccess &&
           hipMemcpy(r.depthmap.data(), odep, npix * sizeof(float), hipMemcpyDeviceToHost) == hipSuccess &&
           hipMemcpy(r.normalmap.data(), onrm, npix * sizeof(float3), hipMemcpyDeviceToHost) == hipSuccess &&
           hipMemcpy(r.colormap.data(), ocol, npix * sizeof(uchar3), hipMemcpyDeviceToHost) == hipSuccess;
      if (!ok) r = GeneratedDepthMap();
    }
    if (!ok) std::fprintf(stderr, "CustomDepthMapGenerator: GPU generation failed\n");
    (void)hipFree(dp);
    (void)hipFree(dn);
    (void)hipFree(dc);
    (void)hipFree(out);
    return r;
  }
  vsg_depthgen* gen() {
    if (m_gen && (m_genW != m_cam.width || m_genH != m_cam.height)) release();
    if (!m_gen) {
      if (!vsg_detail::report(vsg_depthgen_create(m_cam.width, m_cam.height, &m_gen), "CustomDepthMapGenerator"))
        return nullptr;
      m_genW = m_cam.width;
      m_genH = m_cam.height;
      if (m_haveStream) vsg_depthgen_set_stream(m_gen, (void*)m_stream);
    }
    vsg_depthgen_set_projection(m_gen, (int32_t)m_method);
    return m_gen;
  }
  void release() {
    if (m_gen) vsg_depthgen_destroy(m_gen);
    m_gen = nullptr;
  }
  CameraParams m_cam;
  ProjectionMethod m_method = ProjectionMethod::PERSPECTIVE;
  vsg_depthgen* m_gen = nullptr;
  int m_genW = 0, m_genH = 0;
  hipStream_t m_stream = nullptr;
  bool m_haveStream = false;
};

// MarchingCubes (VoxelSG/include/MarchingCubes.h:7-59).  The output's device buffers
// are the library's; d_triangleCount is not kept on the device (the count of the last
// run is triangleCount), and the run also takes the cross-block and weight options.
struct MarchingCubesParams {  // MarchingCubes.h:7-16, same defaults
  float isoValue = 0.0f;
  float voxelSize = 1.0f;  // used when positive, else Params::voxelSize
  int maxTriangles = 0;
};

struct MarchingCubesOutput {  // MarchingCubes.h:18-29
  float3* d_vertices = nullptr;
  float3* d_normals = nullptr;
  int* d_triangleCount = nullptr;
  int maxTriangles = 0;
  long long triangleCount = 0;  // triangles found by the last run
  vsg_mesh* mesh = nullptr;
};

// MarchingCubes.cu:491-501
inline void freeMarchingCubesOutput(MarchingCubesOutput& output) {
  if (output.mesh) vsg_mesh_destroy(output.mesh);
  output = MarchingCubesOutput{};
}

// MarchingCubes.cu:474-489
inline void allocateMarchingCubesOutput(MarchingCubesOutput& output, int maxTriangles) {
  if (maxTriangles <= 0) {
    std::printf("allocateMarchingCubesOutput: maxTriangles must be positive.\n");
    return;
  }
  freeMarchingCubesOutput(output);
  if (!vsg_detail::report(vsg_mesh_create((uint32_t)maxTriangles, &output.mesh), "allocateMarchingCubesOutput")) return;
  float *v = nullptr, *n = nullptr;
  vsg_mesh_buffers(output.mesh, &v, &n, nullptr);
  output.d_vertices = reinterpret_cast<float3*>(v);
  output.d_normals = reinterpret_cast<float3*>(n);
  output.maxTriangles = maxTriangles;
}

// MarchingCubes.cu:511-571 over every live block of the scene, in hash-entry, cube and
// table order; crossBlock adds the cubes across block borders (the reference's stay
// inside one block), minWeight > 0 leaves out cubes with a lighter corner.
inline void runMarchingCubesCUDA(VoxelScene& scene, const Params& params, const MarchingCubesParams& paramsMC,
                                 MarchingCubesOutput& output, bool crossBlock = false, unsigned minWeight = 0) {
  if (!output.mesh) {
    std::printf("runMarchingCubesCUDA: Output buffers not allocated.\n");
    return;
  }
  const float voxel = paramsMC.voxelSize > 0.0f ? paramsMC.voxelSize : params.voxelSize;
  uint64_t total = 0;
  if (vsg_detail::report(vsg_marching_cubes(output.mesh, scene.handle(), paramsMC.isoValue, voxel, crossBlock ? 1 : 0,
                                            minWeight, nullptr, &total),
                         "runMarchingCubesCUDA"))
    output.triangleCount = (long long)total;
}

// MarchingCubes.cu:503-509: the triangles found (may exceed maxTriangles, as the
// reference's counter does).
inline long long copyMarchingCubesTriangleCount(const MarchingCubesOutput& output) { return output.triangleCount; }
