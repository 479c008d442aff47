// Host driver written against the reference's C++ API (include/vsg_voxelscene.hpp),
// following the per-frame order of VoxelSG/src/voxelSG.cpp:392-690: upload the
// frame, VoxelScene::integrateFromScanData, create a RayCastRender, upload the pose,
// compute the depth range from the host depth map (voxelSG.cpp:594-608), render,
// download.  Frames come from the synthetic generator (libvsg_synth.so).
//
// After the last frame, RayCastRender::extractSurfacePoints on a renderer that was
// never initialised (as the reference allows) and, when a path is given,
// saveExtractedSurfacePointsPLY.
//
// With a fourth argument "gen", every frame goes through CustomDepthMapGenerator first: the
// valid pixels' points (with normals and colours, in pixel order) form a point cloud whose
// depth maps, generated on the device on the scene's stream, are integrated.
//
// Then runMarchingCubesCUDA with the reference's cubes (inside blocks, no weight test)
// and, when a mesh path is given, VoxelScene::saveToMesh.
//
// With a sixth argument "stream:<radius>[:<cap>]", a VoxelStreamingManager (1 m chunks,
// 64^3 chunks from -32; at most <cap> blocks per stream-out or stream-in pass, 100000 when
// not given) runs before every integration from the second frame on: stream-out of the
// blocks at distance >= radius from the camera, then stream-in of the chunks within
// radius (the order of VoxelSG/src/voxelSG.cpp:503-552).
//
// usage: vsg_shim_example <workload 0|1> <frames> [points.ply|-] [gen|-] [mesh.ply|-]
//                         [stream:<radius>[:<cap>]]
//        vsg_shim_example bench <frames> <passes>
//
// The bench form times the host code path of the drop-in classes on the 640x480 room
// sequence as bench.py times the Python one: the frames' points, colours, normals and
// poses resident in device memory before the timed region, then `passes` passes over
// the sequence from an empty scene (VoxelScene::reset), each frame
// integrateFromScanData then RayCastRender::render (pose as a device pointer, one
// renderer for all frames) at the frame's depth range; one untimed pass first.  Prints
// the frames per second and milliseconds per frame.
// prints: frames, allocated blocks, hits, misses, a checksum of the last depth map,
// the number of surface points found, the number of triangles and the number of
// blocks held in the streaming chunk grid.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "vsg_voxelscene.hpp"

extern "C" {
void* vsgs_create(int cfg, uint64_t seed, int nframes);
void vsgs_destroy(void* p);
void vsgs_info(void* p, int* whn, double* intr);
void vsgs_pose(void* p, int frame, double* T16);
void vsgs_render(void* p, int frame, const double* T16, int noise, float* pts, uint8_t* rgb, float* nrm, float* depth);
void vsgs_depth_range(const float* depth, long n, float* out2);
}

static int bench_main(int nframes, int passes) {
  void* sc = vsgs_create(1, 0, nframes);
  int whn[3];
  double intr[4];
  vsgs_info(sc, whn, intr);
  const int W = whn[0], H = whn[1];
  const float fx = (float)intr[0], fy = (float)intr[1], cx = (float)intr[2], cy = (float)intr[3];
  const size_t n = (size_t)W * H;
  Params params(1 << 20, 1 << 19, 0.004f);
  const float truncation = 0.02f;
  VoxelScene scene(params);
  if (!scene.isAllocated()) return 1;
  float3 *d_pts = nullptr, *d_nrm = nullptr;
  uchar3* d_rgb = nullptr;
  float* d_T = nullptr;
  if (hipMalloc(&d_pts, 12 * n * nframes) != hipSuccess || hipMalloc(&d_nrm, 12 * n * nframes) != hipSuccess ||
      hipMalloc(&d_rgb, 3 * n * nframes) != hipSuccess || hipMalloc(&d_T, 64 * (size_t)nframes) != hipSuccess)
    return 2;
  std::vector<float> pts(3 * n), nrm(3 * n), depth(n), T(16 * (size_t)nframes), cams(3 * (size_t)nframes),
      ranges(2 * (size_t)nframes);
  std::vector<uint8_t> rgb(3 * n);
  for (int i = 0; i < nframes; ++i) {
    double Td[16];
    vsgs_pose(sc, i, Td);
    vsgs_render(sc, i, Td, 1, pts.data(), rgb.data(), nrm.data(), depth.data());
    for (int k = 0; k < 16; ++k) T[16 * (size_t)i + k] = (float)Td[k];
    cams[3 * (size_t)i] = T[16 * (size_t)i + 3];
    cams[3 * (size_t)i + 1] = T[16 * (size_t)i + 7];
    cams[3 * (size_t)i + 2] = T[16 * (size_t)i + 11];
    vsgs_depth_range(depth.data(), (long)n, &ranges[2 * (size_t)i]);
    if (hipMemcpy(d_pts + n * i, pts.data(), 12 * n, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(d_rgb + n * i, rgb.data(), 3 * n, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(d_nrm + n * i, nrm.data(), 12 * n, hipMemcpyHostToDevice) != hipSuccess)
      return 3;
  }
  if (hipMemcpy(d_T, T.data(), 64 * (size_t)nframes, hipMemcpyHostToDevice) != hipSuccess) return 3;
  RayCastRender renderer;
  if (!renderer.initialize(W, H)) return 4;
  auto pass = [&]() {
    scene.reset();
    for (int i = 0; i < nframes; ++i) {
      const float* Ti = &T[16 * (size_t)i];
      scene.integrateFromScanData(d_pts + n * i, d_rgb + n * i, d_nrm + n * i, W, H, truncation, Ti, fx, fy, cx, cy);
      const float3 cam = make_float3(cams[3 * (size_t)i], cams[3 * (size_t)i + 1], cams[3 * (size_t)i + 2]);
      renderer.render(&scene, cam, d_T + 16 * (size_t)i, fx, fy, cx, cy, ranges[2 * (size_t)i],
                      ranges[2 * (size_t)i + 1]);
    }
  };
  pass();
  if (hipDeviceSynchronize() != hipSuccess) return 5;
  const auto t0 = std::chrono::steady_clock::now();
  for (int p = 0; p < passes; ++p) pass();
  if (hipDeviceSynchronize() != hipSuccess) return 5;
  const double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  const double frames = (double)nframes * passes;
  std::printf("bench frames %d passes %d fps %.3f ms_per_frame %.5f\n", nframes, passes, frames / s, s / frames * 1e3);
  (void)hipFree(d_pts);
  (void)hipFree(d_nrm);
  (void)hipFree(d_rgb);
  (void)hipFree(d_T);
  vsgs_destroy(sc);
  return 0;
}

int main(int argc, char** argv) {
  if (argc > 1 && std::string(argv[1]) == "bench")
    return bench_main(argc > 2 ? std::atoi(argv[2]) : 300, argc > 3 ? std::atoi(argv[3]) : 3);
  const int cfg = argc > 1 ? std::atoi(argv[1]) : 0;
  const int nframes = argc > 2 ? std::atoi(argv[2]) : 5;
  void* sc = vsgs_create(cfg, 0, 0);
  int whn[3];
  double intr[4];
  vsgs_info(sc, whn, intr);
  const int W = whn[0], H = whn[1];
  const float fx = (float)intr[0], fy = (float)intr[1], cx = (float)intr[2], cy = (float)intr[3];
  const size_t n = (size_t)W * H;

  Params params = cfg == 0 ? Params(200000, 1 << 17, 0.004f) : Params(1 << 20, 1 << 19, 0.004f);
  const float truncation = 0.02f;
  VoxelScene scene(params);
  if (!scene.isAllocated()) return 1;

  std::vector<float> pts(3 * n), nrm(3 * n), depth(n);
  std::vector<uint8_t> rgb(3 * n);
  float3 *d_pts = nullptr, *d_nrm = nullptr;
  uchar3* d_rgb = nullptr;
  float* d_T = nullptr;
  if (hipMalloc(&d_pts, 12 * n) != hipSuccess || hipMalloc(&d_nrm, 12 * n) != hipSuccess ||
      hipMalloc(&d_rgb, 3 * n) != hipSuccess || hipMalloc(&d_T, 64) != hipSuccess)
    return 2;

  const bool viaGenerator = argc > 4 && std::string(argv[4]) == "gen";
  CustomDepthMapGenerator generator;
  generator.setCameraParams(fx, fy, cx, cy, 0.1f, 10.0f, W, H);
  void* sceneStream = nullptr;
  vsg_scene_get_stream(scene.handle(), &sceneStream);
  generator.setStream((hipStream_t)sceneStream);
  float3 *c_pts = nullptr, *c_nrm = nullptr, *m_pts = nullptr, *m_nrm = nullptr;
  uchar3 *c_rgb = nullptr, *m_rgb = nullptr;
  if (viaGenerator && (hipMalloc(&c_pts, 12 * n) != hipSuccess || hipMalloc(&c_nrm, 12 * n) != hipSuccess ||
                       hipMalloc(&c_rgb, 3 * n) != hipSuccess || hipMalloc(&m_pts, 12 * n) != hipSuccess ||
                       hipMalloc(&m_nrm, 12 * n) != hipSuccess || hipMalloc(&m_rgb, 3 * n) != hipSuccess))
    return 2;
  std::vector<float> cp, cn;
  std::vector<uint8_t> cc;

  const std::string streamArg = argc > 6 ? argv[6] : "";
  const bool streaming = streamArg.rfind("stream:", 0) == 0;
  char* capArg = nullptr;
  const float streamRadius = streaming ? std::strtof(streamArg.c_str() + 7, &capArg) : 0.0f;
  VoxelStreamingManager streamer(scene, params, make_float3(1.0f, 1.0f, 1.0f), make_int3(64, 64, 64),
                                 make_int3(-32, -32, -32));
  if (streaming && capArg && *capArg == ':') {
    const long cap = std::strtol(capArg + 1, nullptr, 10);
    if (cap <= 0) {
      std::fprintf(stderr, "stream cap must be positive\n");
      return 2;
    }
    streamer.m_maxNumberOfSDFBlocksIntegrateFromGlobalHash = (unsigned int)cap;
  }

  unsigned int hits = 0, misses = 0;
  double checksum = 0.0;
  for (int i = 0; i < nframes; ++i) {
    double Td[16];
    vsgs_pose(sc, i, Td);
    vsgs_render(sc, i, Td, 1, pts.data(), rgb.data(), nrm.data(), depth.data());
    float T[16];
    for (int k = 0; k < 16; ++k) T[k] = (float)Td[k];
    if (hipMemcpy(d_pts, pts.data(), 12 * n, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(d_rgb, rgb.data(), 3 * n, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(d_nrm, nrm.data(), 12 * n, hipMemcpyHostToDevice) != hipSuccess)
      return 3;
    if (streaming && i > 0) {
      const float3 cam = make_float3(T[3], T[7], T[11]);
      streamer.streamOut(scene, cam, streamRadius);
      streamer.streamIn(scene, cam, streamRadius);
    }
    if (viaGenerator) {
      cp.clear();
      cn.clear();
      cc.clear();
      for (size_t k = 0; k < n; ++k) {
        if (!(depth[k] > 0.0f)) continue;
        cp.insert(cp.end(), &pts[3 * k], &pts[3 * k] + 3);
        cn.insert(cn.end(), &nrm[3 * k], &nrm[3 * k] + 3);
        cc.insert(cc.end(), &rgb[3 * k], &rgb[3 * k] + 3);
      }
      const size_t m = cp.size() / 3;
      if (hipMemcpy(c_pts, cp.data(), 12 * m, hipMemcpyHostToDevice) != hipSuccess ||
          hipMemcpy(c_nrm, cn.data(), 12 * m, hipMemcpyHostToDevice) != hipSuccess ||
          hipMemcpy(c_rgb, cc.data(), 3 * m, hipMemcpyHostToDevice) != hipSuccess)
        return 3;
      if (!generator.generateDevice(c_pts, m, c_nrm, m, c_rgb, m, m_pts, nullptr, m_nrm, m_rgb)) return 7;
      scene.integrateFromScanData(m_pts, m_rgb, m_nrm, W, H, truncation, T, fx, fy, cx, cy);
    } else {
      scene.integrateFromScanData(d_pts, d_rgb, d_nrm, W, H, truncation, T, fx, fy, cx, cy);
    }

    RayCastRender renderer;  // created per frame, as the driver does (voxelSG.cpp:578-579)
    if (!renderer.initialize(W, H)) return 4;
    float3 camPos = make_float3(T[3], T[7], T[11]);
    if (hipMemcpy(d_T, T, 64, hipMemcpyHostToDevice) != hipSuccess) return 5;
    float range[2];
    vsgs_depth_range(depth.data(), (long)n, range);
    renderer.render(&scene, camPos, d_T, fx, fy, cx, cy, range[0], range[1]);
    std::vector<float> out(n);
    renderer.downloadDepthFloat(out.data());
    renderer.getCounters(&hits, &misses);
    checksum = 0.0;
    for (size_t k = 0; k < n; ++k)
      if (out[k] < 1e9f) checksum += (double)out[k] * (double)((k % 97) + 1);
  }
  RayCastRender extractor;
  const int points = extractor.extractSurfacePoints(&scene);
  if (argc > 3 && std::string(argv[3]) != "-" && !extractor.saveExtractedSurfacePointsPLY(argv[3])) return 6;
  MarchingCubesOutput mc;
  allocateMarchingCubesOutput(mc, 1 << 22);
  MarchingCubesParams mcp;
  mcp.voxelSize = 0.0f;  // the scene's
  runMarchingCubesCUDA(scene, params, mcp, mc);
  const long long triangles = copyMarchingCubesTriangleCount(mc);
  freeMarchingCubesOutput(mc);
  if (argc > 5 && std::string(argv[5]) != "-" && !scene.saveToMesh(argv[5])) return 7;
  std::printf("frames %d blocks %u hits %u misses %u checksum %.9e points %d triangles %lld streamed %zu\n",
              scene.getNumIntegratedFrames(), (unsigned)(params.SDFBlockNum - scene.getHeapFreeCount()), hits, misses,
              checksum, points, triangles, streamer.numStreamedBlocks());
  (void)hipFree(d_pts);
  (void)hipFree(d_nrm);
  (void)hipFree(d_rgb);
  (void)hipFree(d_T);
  (void)hipFree(c_pts);
  (void)hipFree(c_nrm);
  (void)hipFree(c_rgb);
  (void)hipFree(m_pts);
  (void)hipFree(m_nrm);
  (void)hipFree(m_rgb);
  vsgs_destroy(sc);
  return 0;
}
