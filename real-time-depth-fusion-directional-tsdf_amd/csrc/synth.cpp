// Synthetic depth frames for the VoxelSG fusion path: analytic renders of seeded
// scenes along seeded camera trajectories, in the input formats
// VoxelScene::integrateFromScanData takes (VoxelSG/include/VoxelScene.h:89-91):
//   pointmap  float3 per pixel, camera frame (x right, y down, z forward),
//             pixel (u,v) back-projected at integer coordinates:
//             P = ((u-cx)*z/fx, (v-cy)*z/fy, z); invalid pixels are (0,0,0);
//   colormap  uchar3 per pixel, one seeded colour per object;
//   normalmap float3 per pixel, finite differences of the depth map exactly as
//             VoxelSG/src/CustomDepthMapGenerator.cpp:164-208 computes them
//             (normalize((right-centre) x (down-centre)), zero on the border);
//   pose      row-major float[16] camera->world.
//
// Scenes (sizes as the workloads define them):
//   0  sphere r=0.3 m centred 1 m ahead of frame 0 on a ground plane tangent to
//      its bottom; 30 cameras on a 20 deg arc (radius 1 m about the centre,
//      looking at it); 320x240, f=262.5; z noise N(0, 1 mm).
//   1  4x4x3 m box room (inward faces) with 10 spheres (r 0.1-0.5 m) and 10
//      axis-aligned boxes (edges 0.2-1 m); smooth random-walk trajectory
//      (<= 5 cm, <= 3 deg per frame); 640x480, f=525; noise
//      sigma(z) = 0.0012 + 0.0019 (z-0.4)^2.
//   4  8x8x3 m room, 30 spheres + 30 boxes, 1024x768, f=840, noise as 1.
// Depths outside [0.2, 4] m are invalid.  Seeds: scene 1, trajectory 2,
// noise 3, colour 4 (added to the caller's base seed).
//
// Random numbers are counter based (splitmix64 finaliser), so a frame's bytes do
// not depend on thread count or on which frames were rendered before it.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

constexpr double kPi = 3.14159265358979323846;

inline uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

struct Rng {  // sequential splitmix64 stream
  uint64_t s;
  explicit Rng(uint64_t seed) : s(seed) {}
  uint64_t next() { s += 0x9E3779B97F4A7C15ull; return mix64(s); }
  double uni() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }  // [0,1)
  double uni(double a, double b) { return a + (b - a) * uni(); }
  double gauss() {
    double u1 = 1.0 - uni(), u2 = uni();
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * kPi * u2);
  }
};

// N(0,1) for (seed, frame, pixel), independent of evaluation order.
inline double gauss_at(uint64_t seed, uint64_t frame, uint64_t pixel) {
  uint64_t k = mix64(seed ^ mix64(frame * 0x9E3779B97F4A7C15ull + 0x632BE59BD9B4E019ull) ^ (pixel * 0xD1B54A32D192ED03ull));
  uint64_t a = mix64(k), b = mix64(k ^ 0xA0761D6478BD642Full);
  double u1 = 1.0 - (double)(a >> 11) * (1.0 / 9007199254740992.0);
  double u2 = (double)(b >> 11) * (1.0 / 9007199254740992.0);
  return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * kPi * u2);
}

enum PrimType { kSphere = 0, kBox = 1, kPlane = 2, kRoom = 3 };

struct Prim {
  int type;
  double c[3] = {0, 0, 0};  // sphere centre
  double r = 0;             // sphere radius
  double lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};  // box / room
  double n[3] = {0, 0, 0}, d = 0;               // plane n.x = d
  uint8_t rgb[3] = {0, 0, 0};
  uint8_t face_rgb[6][3] = {};  // room: one colour per face
};

struct Synth {
  int cfg = 0, W = 0, H = 0, nframes = 0;
  double fx = 0, fy = 0, cx = 0, cy = 0;
  double zlo = 0.2, zhi = 4.0;
  double sig_a = 0.001, sig_b = 0.0;  // sigma(z) = a + b (z - 0.4)^2
  uint64_t seed = 0;
  std::vector<Prim> prims;
  std::vector<std::array<double, 16>> poses;
};

void rot_ypr(double yaw, double pitch, double roll, double* R) {
  // R = Ry(yaw) * Rx(pitch) * Rz(roll)
  double cy_ = std::cos(yaw), sy = std::sin(yaw), cp = std::cos(pitch), sp = std::sin(pitch), cr = std::cos(roll),
         sr = std::sin(roll);
  double Ry[9] = {cy_, 0, sy, 0, 1, 0, -sy, 0, cy_};
  double Rx[9] = {1, 0, 0, 0, cp, -sp, 0, sp, cp};
  double Rz[9] = {cr, -sr, 0, sr, cr, 0, 0, 0, 1};
  double T[9];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double s = 0;
      for (int k = 0; k < 3; ++k) s += Ry[3 * i + k] * Rx[3 * k + j];
      T[3 * i + j] = s;
    }
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double s = 0;
      for (int k = 0; k < 3; ++k) s += T[3 * i + k] * Rz[3 * k + j];
      R[3 * i + j] = s;
    }
}

void set_pose(std::array<double, 16>& P, const double* R, const double* t) {
  for (int i = 0; i < 3; ++i) {
    for (int j = 0; j < 3; ++j) P[4 * i + j] = R[3 * i + j];
    P[4 * i + 3] = t[i];
  }
  P[12] = P[13] = P[14] = 0;
  P[15] = 1;
}

void colour(Rng& rng, uint8_t* c) {
  for (int k = 0; k < 3; ++k) c[k] = (uint8_t)(40 + (int)(rng.uni() * 200.0));
}

void build_room_scene(Synth& s, double hx, double hy, double hz, int nsph, int nbox, double kx, double ky, double kz) {
  Rng srng(s.seed + 1), crng(s.seed + 4);
  Prim room;
  room.type = kRoom;
  room.lo[0] = -hx; room.lo[1] = -hy; room.lo[2] = -hz;
  room.hi[0] = hx; room.hi[1] = hy; room.hi[2] = hz;
  for (int f = 0; f < 6; ++f) colour(crng, room.face_rgb[f]);
  s.prims.push_back(room);
  const double keep = 0.3;  // clearance around the camera region
  auto box_clear = [&](const double* lo, const double* hi) {
    return hi[0] < -kx - keep || lo[0] > kx + keep || hi[1] < -ky - keep || lo[1] > ky + keep || hi[2] < -kz - keep ||
           lo[2] > kz + keep;
  };
  for (int i = 0; i < nsph + nbox; ++i) {
    Prim p;
    for (int attempt = 0; attempt < 10000; ++attempt) {
      if (i < nsph) {
        p.type = kSphere;
        p.r = srng.uni(0.1, 0.5);
        p.c[0] = srng.uni(-hx + p.r, hx - p.r);
        p.c[1] = srng.uni(-hy + p.r, hy - p.r);
        p.c[2] = srng.uni(-hz + p.r, hz - p.r);
        double lo[3] = {p.c[0] - p.r, p.c[1] - p.r, p.c[2] - p.r}, hi[3] = {p.c[0] + p.r, p.c[1] + p.r, p.c[2] + p.r};
        if (box_clear(lo, hi)) break;
      } else {
        p.type = kBox;
        double e[3] = {srng.uni(0.2, 1.0), srng.uni(0.2, 1.0), srng.uni(0.2, 1.0)};
        double h[3] = {hx, hy, hz};
        for (int k = 0; k < 3; ++k) {
          double c = srng.uni(-h[k] + e[k] / 2, h[k] - e[k] / 2);
          p.lo[k] = c - e[k] / 2;
          p.hi[k] = c + e[k] / 2;
        }
        if (box_clear(p.lo, p.hi)) break;
      }
    }
    colour(crng, p.rgb);
    s.prims.push_back(p);
  }
}

void build_random_walk(Synth& s, double kx, double ky, double kz) {
  Rng rng(s.seed + 2);
  const double deg = kPi / 180.0;
  double p[3] = {0, 0, 0}, v[3] = {0, 0, 0};
  double yaw = 0, pitch = 0, roll = 0, wy = 0, wp = 0, wr = 0;
  const double lim[3] = {kx, ky, kz};
  s.poses.resize(s.nframes);
  for (int f = 0; f < s.nframes; ++f) {
    double R[9];
    rot_ypr(yaw, pitch, roll, R);
    set_pose(s.poses[f], R, p);
    // translation: damped random velocity, <= 5 cm per frame, reflected at the region bounds
    for (int k = 0; k < 3; ++k) v[k] = 0.9 * v[k] + 0.012 * rng.gauss();
    double vn = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    if (vn > 0.05) for (int k = 0; k < 3; ++k) v[k] *= 0.05 / vn;
    for (int k = 0; k < 3; ++k) {
      p[k] += v[k];
      if (p[k] > lim[k]) { p[k] = 2 * lim[k] - p[k]; v[k] = -v[k]; }
      if (p[k] < -lim[k]) { p[k] = -2 * lim[k] - p[k]; v[k] = -v[k]; }
    }
    // rotation: damped random rates, |rate| <= 3 deg per frame, bounded pitch and roll
    wy = 0.9 * wy + 0.6 * deg * rng.gauss();
    wp = 0.9 * wp + 0.3 * deg * rng.gauss();
    wr = 0.9 * wr + 0.1 * deg * rng.gauss();
    if (pitch > 20 * deg && wp > 0) wp = -wp;
    if (pitch < -20 * deg && wp < 0) wp = -wp;
    if (roll > 5 * deg && wr > 0) wr = -wr;
    if (roll < -5 * deg && wr < 0) wr = -wr;
    double wn = std::sqrt(wy * wy + wp * wp + wr * wr);
    if (wn > 2.9 * deg) { wy *= 2.9 * deg / wn; wp *= 2.9 * deg / wn; wr *= 2.9 * deg / wn; }
    yaw += wy;
    pitch += wp;
    roll += wr;
  }
}

void build(Synth& s, int cfg, uint64_t seed, int nframes) {
  s.cfg = cfg;
  s.seed = seed;
  if (cfg == 0) {
    s.W = 320; s.H = 240; s.fx = s.fy = 262.5; s.cx = 159.5; s.cy = 119.5;
    s.nframes = nframes > 0 ? nframes : 30;
    s.sig_a = 0.001; s.sig_b = 0.0;
    Rng crng(seed + 4);
    Prim sph;
    sph.type = kSphere;
    sph.c[0] = 0; sph.c[1] = 0; sph.c[2] = 1.0; sph.r = 0.3;
    colour(crng, sph.rgb);
    Prim ground;  // y points down: the ground plane y = +0.3 touches the sphere's bottom
    ground.type = kPlane;
    ground.n[0] = 0; ground.n[1] = 1; ground.n[2] = 0; ground.d = 0.3;
    colour(crng, ground.rgb);
    s.prims = {sph, ground};
    s.poses.resize(s.nframes);
    for (int f = 0; f < s.nframes; ++f) {
      double th = (-10.0 + 20.0 * (s.nframes > 1 ? (double)f / (s.nframes - 1) : 0.0)) * kPi / 180.0;
      double R[9];
      rot_ypr(th, 0, 0, R);
      double t[3] = {sph.c[0] - R[2], sph.c[1] - R[5], sph.c[2] - R[8]};  // centre - forward * 1 m
      set_pose(s.poses[f], R, t);
    }
  } else if (cfg == 4) {
    s.W = 1024; s.H = 768; s.fx = s.fy = 840; s.cx = 511.5; s.cy = 383.5;
    s.nframes = nframes > 0 ? nframes : 1000;
    s.sig_a = 0.0012; s.sig_b = 0.0019;
    build_room_scene(s, 4.0, 1.5, 4.0, 30, 30, 1.5, 0.4, 1.5);
    build_random_walk(s, 1.5, 0.4, 1.5);
  } else {  // 1 (and the workloads derived from its frames)
    s.W = 640; s.H = 480; s.fx = s.fy = 525; s.cx = 319.5; s.cy = 239.5;
    s.nframes = nframes > 0 ? nframes : 300;
    s.sig_a = 0.0012; s.sig_b = 0.0019;
    build_room_scene(s, 2.0, 1.5, 2.0, 10, 10, 0.5, 0.3, 0.5);
    build_random_walk(s, 0.5, 0.3, 0.5);
  }
}

// Nearest positive hit of o + t d.  Returns t (inf if none) and the colour.
double trace(const Synth& s, const double* o, const double* d, uint8_t* rgb) {
  double best = INFINITY;
  const double eps = 1e-9;
  for (const Prim& p : s.prims) {
    double t = INFINITY;
    const uint8_t* c = p.rgb;
    if (p.type == kSphere) {
      double oc[3] = {o[0] - p.c[0], o[1] - p.c[1], o[2] - p.c[2]};
      double A = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
      double B = d[0] * oc[0] + d[1] * oc[1] + d[2] * oc[2];
      double C = oc[0] * oc[0] + oc[1] * oc[1] + oc[2] * oc[2] - p.r * p.r;
      double disc = B * B - A * C;
      if (disc >= 0) {
        double sq = std::sqrt(disc);
        double t0 = (-B - sq) / A, t1 = (-B + sq) / A;
        t = t0 > eps ? t0 : (t1 > eps ? t1 : INFINITY);
      }
    } else if (p.type == kPlane) {
      double nd = p.n[0] * d[0] + p.n[1] * d[1] + p.n[2] * d[2];
      if (std::fabs(nd) > 1e-15) {
        double tt = (p.d - (p.n[0] * o[0] + p.n[1] * o[1] + p.n[2] * o[2])) / nd;
        if (tt > eps) t = tt;
      }
    } else {  // box (entry) or room (exit)
      double tn = -INFINITY, tf = INFINITY;
      int an = -1, af = -1;
      bool miss = false;
      for (int k = 0; k < 3; ++k) {
        if (std::fabs(d[k]) < 1e-15) {
          if (o[k] < p.lo[k] || o[k] > p.hi[k]) miss = true;
          continue;
        }
        double t1 = (p.lo[k] - o[k]) / d[k], t2 = (p.hi[k] - o[k]) / d[k];
        int f1 = 2 * k, f2 = 2 * k + 1;
        if (t1 > t2) { std::swap(t1, t2); std::swap(f1, f2); }
        if (t1 > tn) { tn = t1; an = f1; }
        if (t2 < tf) { tf = t2; af = f2; }
      }
      if (!miss && tn <= tf) {
        if (p.type == kBox) {
          if (tn > eps) t = tn;
        } else if (tf > eps) {
          t = tf;
          c = p.face_rgb[af < 0 ? 0 : af];
        }
      }
      (void)an;
    }
    if (t < best) {
      best = t;
      rgb[0] = c[0]; rgb[1] = c[1]; rgb[2] = c[2];
    }
  }
  return best;
}

}  // namespace

extern "C" {

void* vsgs_create(int cfg, uint64_t seed, int nframes) {
  Synth* s = new Synth();
  build(*s, cfg, seed, nframes);
  return s;
}

void vsgs_destroy(void* p) { delete (Synth*)p; }

// out: W, H, nframes; intr: fx, fy, cx, cy
void vsgs_info(void* p, int* whn, double* intr) {
  Synth* s = (Synth*)p;
  whn[0] = s->W; whn[1] = s->H; whn[2] = s->nframes;
  intr[0] = s->fx; intr[1] = s->fy; intr[2] = s->cx; intr[3] = s->cy;
}

void vsgs_pose(void* p, int frame, double* T16) {
  Synth* s = (Synth*)p;
  const auto& P = s->poses[(size_t)frame % s->poses.size()];
  for (int i = 0; i < 16; ++i) T16[i] = P[i];
}

// Renders frame `frame` (noise stream index) at pose T16 (row-major camera->world;
// NULL: the trajectory pose).  depth (optional) receives the z value per pixel
// (0 when invalid).  noise=0 renders without depth noise.
void vsgs_render(void* p, int frame, const double* T16, int noise, float* pts, uint8_t* rgb, float* nrm, float* depth) {
  Synth* s = (Synth*)p;
  double T[16];
  if (T16) std::memcpy(T, T16, sizeof(T));
  else vsgs_pose(p, frame, T);
  const int W = s->W, H = s->H;
  std::vector<float> z((size_t)W * H, 0.0f);
  const float fx = (float)s->fx, fy = (float)s->fy, cx = (float)s->cx, cy = (float)s->cy;
#pragma omp parallel for schedule(static)
  for (int y = 0; y < H; ++y) {
    for (int x = 0; x < W; ++x) {
      size_t i = (size_t)y * W + x;
      double a = ((double)x - s->cx) / s->fx, b = ((double)y - s->cy) / s->fy;
      double d[3] = {T[0] * a + T[1] * b + T[2], T[4] * a + T[5] * b + T[6], T[8] * a + T[9] * b + T[10]};
      double o[3] = {T[3], T[7], T[11]};
      uint8_t c[3] = {0, 0, 0};
      double t = trace(*s, o, d, c);  // camera-space direction (a, b, 1): z of the hit = t
      double zz = t;
      if (noise && std::isfinite(zz)) {
        double sig = s->sig_a + s->sig_b * (zz - 0.4) * (zz - 0.4);
        zz += sig * gauss_at(s->seed + 3, (uint64_t)frame, i);
      }
      float zf = (std::isfinite(zz) && zz >= s->zlo && zz <= s->zhi) ? (float)zz : 0.0f;
      z[i] = zf;
      if (zf > 0.0f) {
        pts[3 * i] = ((float)x - cx) * zf / fx;
        pts[3 * i + 1] = ((float)y - cy) * zf / fy;
        pts[3 * i + 2] = zf;
        rgb[3 * i] = c[0]; rgb[3 * i + 1] = c[1]; rgb[3 * i + 2] = c[2];
      } else {
        pts[3 * i] = pts[3 * i + 1] = pts[3 * i + 2] = 0.0f;
        rgb[3 * i] = rgb[3 * i + 1] = rgb[3 * i + 2] = 0;
      }
      nrm[3 * i] = nrm[3 * i + 1] = nrm[3 * i + 2] = 0.0f;
      if (depth) depth[i] = zf;
    }
  }
  // finite-difference normals (CustomDepthMapGenerator.cpp:164-208), float arithmetic
#pragma omp parallel for schedule(static)
  for (int y = 1; y < H - 1; ++y) {
    for (int x = 1; x < W - 1; ++x) {
      size_t i = (size_t)y * W + x;
      float dc = z[i];
      if (!(dc > 0.0f)) continue;
      float cxp = ((float)x - cx) * dc / fx, cyp = ((float)y - cy) * dc / fy, czp = dc;
      float dr = z[i + 1];
      float rx = ((float)(x + 1) - cx) * dr / fx, ry = ((float)y - cy) * dr / fy, rz = dr;
      float dd = z[i + W];
      float dx = ((float)x - cx) * dd / fx, dy = ((float)(y + 1) - cy) * dd / fy, dz = dd;
      float ax = rx - cxp, ay = ry - cyp, az = rz - czp;
      float bx = dx - cxp, by = dy - cyp, bz = dz - czp;
      float nx = ay * bz - az * by, ny = az * bx - ax * bz, nz = ax * by - ay * bx;
      float len = std::sqrt(nx * nx + ny * ny + nz * nz);
      if (len > 0.0f) {
        float inv = 1.0f / len;
        nrm[3 * i] = nx * inv;
        nrm[3 * i + 1] = ny * inv;
        nrm[3 * i + 2] = nz * inv;
      }
    }
  }
}

// Per-frame depth range of the driver (VoxelSG/src/voxelSG.cpp:594-608): min/max
// of the valid depths, widened by 10 % of the span, min clamped to 0.01 m.
void vsgs_depth_range(const float* depth, long n, float* out2) {
  float mn = 1e10f, mx = -1e10f;
  for (long i = 0; i < n; ++i) {
    float d = depth[i];
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

}  // extern "C"
