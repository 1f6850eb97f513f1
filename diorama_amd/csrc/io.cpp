// What sits either side of the training path, over the C ABI: a COLMAP sparse
// reconstruction and its cameras (src/colmap.ts:18-129, src/sceneCameras.ts:53-67,
// src/loadDataset.ts:80-86), the starting Gaussians from its points
// (src/loadDataset.ts:87-149, the kernels in knn.hip) and the 3DGS .ply writer
// and loader (src/plyWriter.ts:14-43, src/plyLoader.ts:21-110). The .ply header is text handled here; the body is
// packed and unpacked by ply.hip on the device, so a trainer's Gaussians go to a
// file straight from HBM and a loaded scene lands in HBM in the renderer's layout.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "runtime.hpp"

using dio::Error;
using dio::alloc;
using dio::guard;
using dio::invalid;
using dio::need;
using dio::release;

namespace {

// The writer's column names in order (src/plyWriter.ts:8-12).
std::vector<std::string> ply_properties() {
  std::vector<std::string> names = {"x", "y", "z", "nx", "ny", "nz", "f_dc_0", "f_dc_1", "f_dc_2"};
  for (uint32_t k = 0; k < dio::kShRest; k++) names.push_back("f_rest_" + std::to_string(k));
  for (const char* n : {"opacity", "scale_0", "scale_1", "scale_2", "rot_0", "rot_1", "rot_2", "rot_3"}) names.push_back(n);
  return names;
}

std::string ply_header(uint32_t count) {
  std::string h = "ply\nformat binary_little_endian 1.0\nelement vertex " + std::to_string(count) + "\n";
  for (const auto& name : ply_properties()) h += "property float " + name + "\n";
  return h + "end_header\n";
}

bool is_space(char c) { return c == ' ' || c == '\t' || c == '\n' || c == '\r' || c == '\v' || c == '\f'; }

// A header line trimmed and split on runs of whitespace.
std::vector<std::string> words_of(const std::string& line) {
  std::vector<std::string> words;
  size_t i = 0;
  while (i < line.size()) {
    while (i < line.size() && is_space(line[i])) i++;
    size_t j = i;
    while (j < line.size() && !is_space(line[j])) j++;
    if (j > i) words.push_back(line.substr(i, j - i));
    i = j;
  }
  return words;
}

struct PlyLayout {
  uint32_t count = 0;
  uint32_t degree = 0;
  size_t body_offset = 0;
  dio::PlyColumns cols{};
};

// The loader's header checks, with its messages (src/plyLoader.ts:22-75).
PlyLayout parse_header(const uint8_t* bytes, size_t size) {
  static const char kMarker[] = "end_header\n";
  const size_t marker_len = sizeof(kMarker) - 1;
  const size_t limit = std::min<size_t>(size, 64 * 1024);
  size_t marker_at = SIZE_MAX;
  for (size_t i = 0; i + marker_len <= limit; i++) {
    if (std::memcmp(bytes + i, kMarker, marker_len) == 0) {
      marker_at = i;
      break;
    }
  }
  if (marker_at == SIZE_MAX) invalid("No .ply header found.");
  PlyLayout out;
  out.body_offset = marker_at + marker_len;
  const std::string header(reinterpret_cast<const char*>(bytes), out.body_offset);
  if (header.compare(0, 4, "ply\n") != 0) invalid("Not a .ply file.");
  // /^format binary_little_endian 1\.0$/m: some line, ended by \n or \r, is exactly this.
  bool little = false;
  for (size_t i = 0; i <= header.size();) {
    size_t j = header.find_first_of("\r\n", i);
    if (j == std::string::npos) j = header.size();
    if (header.compare(i, j - i, "format binary_little_endian 1.0") == 0) little = true;
    i = j + 1;
  }
  if (!little) invalid("Only binary little-endian .ply files are supported.");

  std::vector<std::string> elements, properties;
  bool have_count = false;
  double count = 0;
  for (size_t i = 0; i < header.size();) {
    size_t j = header.find('\n', i);
    if (j == std::string::npos) j = header.size();
    const auto words = words_of(header.substr(i, j - i));
    i = j + 1;
    if (words.empty()) continue;
    if (words[0] == "element") {
      elements.push_back(words.size() >= 2 ? words[1] : "");
      if (elements.back() == "vertex") {
        const std::string n = words.size() >= 3 ? words[2] : "";
        if (n.empty() || n.find_first_not_of("0123456789") != std::string::npos)
          invalid("The .ply vertex count \"" + n + "\" is not a whole number.");
        count = std::stod(n);
        have_count = true;
      }
    } else if (words[0] == "property" && !elements.empty() && elements.back() == "vertex") {
      const std::string type = words.size() >= 2 ? words[1] : "";
      const std::string name = words.size() >= 3 ? words[2] : "";
      if (type != "float") invalid("Property " + name + " is " + type + "; only float properties are supported.");
      properties.push_back(name);
    }
  }
  if (elements.empty() || elements[0] != "vertex") invalid("Expected the .ply to start with its vertex element.");
  if (!have_count || count > 4294967295.0) invalid("The .ply vertex count is out of range.");
  out.count = static_cast<uint32_t>(count);

  auto column = [&](const std::string& name) -> uint32_t {
    for (size_t k = 0; k < properties.size(); k++)
      if (properties[k] == name) return static_cast<uint32_t>(k);
    invalid("The .ply has no \"" + name + "\" property. Is it a 3D Gaussian Splatting file?");
  };
  dio::PlyColumns& c = out.cols;
  for (int k = 0; k < 3; k++) c.position[k] = column(std::string(1, static_cast<char>('x' + k)));
  for (int k = 0; k < 3; k++) c.dc[k] = column("f_dc_" + std::to_string(k));
  c.opacity = column("opacity");
  for (int k = 0; k < 3; k++) c.scale[k] = column("scale_" + std::to_string(k));
  for (int k = 0; k < 4; k++) c.rotation[k] = column("rot_" + std::to_string(k));
  uint32_t rest = 0;
  for (const auto& p : properties)
    if (p.compare(0, 7, "f_rest_") == 0) rest++;
  // degree d holds (d + 1)^2 - 1 coefficients per channel, d <= 3 (src/plyLoader.ts:60-65)
  uint32_t degree = 4;
  for (uint32_t d = 0; d <= 3; d++)
    if (3 * ((d + 1) * (d + 1) - 1) == rest) degree = d;
  if (degree > 3) invalid("Unexpected number of spherical-harmonic coefficients (" + std::to_string(rest) + ").");
  out.degree = degree;
  c.per_channel = rest / 3;
  for (uint32_t k = 0; k < rest; k++) c.rest[k] = column("f_rest_" + std::to_string(k));
  if (properties.size() > dio::kPlyLdsFloats)
    invalid("The .ply has " + std::to_string(properties.size()) + " properties per vertex; at most " +
            std::to_string(dio::kPlyLdsFloats) + " are supported.");
  c.stride = static_cast<uint32_t>(properties.size());
  const double body = static_cast<double>(out.count) * c.stride * 4;
  if (static_cast<double>(out.body_offset) + body > static_cast<double>(size))
    invalid("The .ply file is shorter than its header says.");
  return out;
}

}  // namespace

extern "C" int dio_ply_write(const dio_buf* params, const dio_buf* sh, uint32_t count, void* out, size_t capacity,
                             size_t* length) {
  return guard([&] {
    need(length, "length");
    const std::string header = ply_header(count);
    const size_t body = static_cast<size_t>(count) * dio::kPlyColumns * sizeof(float);
    *length = header.size() + body;
    if (!out) return;
    need(params, "params");
    need(sh, "sh");
    dio_ctx* ctx = params->ctx;
    if (params->bytes < static_cast<size_t>(count) * dio::kFloatsPerGaussian * 4 ||
        sh->bytes < static_cast<size_t>(count) * dio::kShRest * 4)
      invalid("the parameter buffers hold fewer than " + std::to_string(count) + " Gaussians");
    if (capacity < *length)
      invalid("the output holds " + std::to_string(capacity) + " bytes; the .ply needs " + std::to_string(*length));
    std::memcpy(out, header.data(), header.size());
    if (count == 0) return;
    dio_buf* staged = alloc(ctx, body);
    try {
      dio::ply_pack(ctx->stream, static_cast<const float*>(params->ptr), static_cast<const float*>(sh->ptr), count,
                    static_cast<float*>(staged->ptr));
      DIO_HIP(hipMemcpyAsync(static_cast<uint8_t*>(out) + header.size(), staged->ptr, body, hipMemcpyDeviceToHost,
                             ctx->stream));
      dio::sync(ctx);
    } catch (...) {
      release(staged);
      throw;
    }
    release(staged);
  });
}

extern "C" int dio_ply_parse(dio_ctx* ctx, const void* file, size_t bytes, dio_buf** gaussians, dio_buf** sh_rest,
                             uint32_t* count, uint32_t* sh_degree) {
  return guard([&] {
    need(file, "file");
    // The header is checked before anything touches the device.
    const PlyLayout layout = parse_header(static_cast<const uint8_t*>(file), bytes);
    need(ctx, "ctx");
    need(gaussians, "gaussians");
    need(sh_rest, "sh_rest");
    const size_t n = layout.count;
    dio_buf* staged = nullptr;
    dio_buf* g = nullptr;
    dio_buf* s = nullptr;
    try {
      g = alloc(ctx, n * dio::kFloatsPerGaussian * 4);
      s = alloc(ctx, n * dio::kShRest * 4);
      if (n) {
        const size_t body = n * layout.cols.stride * 4;
        staged = alloc(ctx, body);
        DIO_HIP(hipMemcpyAsync(staged->ptr, static_cast<const uint8_t*>(file) + layout.body_offset, body,
                               hipMemcpyHostToDevice, ctx->stream));
        dio::ply_unpack(ctx->stream, static_cast<const float*>(staged->ptr), layout.count, layout.cols,
                        static_cast<float*>(g->ptr), static_cast<float*>(s->ptr));
        dio::sync(ctx);
        release(staged);
      }
    } catch (...) {
      release(staged);
      release(g);
      release(s);
      throw;
    }
    *gaussians = g;
    *sh_rest = s;
    if (count) *count = layout.count;
    if (sh_degree) *sh_degree = layout.degree;
  });
}

namespace {

void check_points(const dio_buf* positions, uint32_t count) {
  need(positions, "positions");
  if (positions->bytes < static_cast<size_t>(count) * 12)
    invalid("positions hold fewer than " + std::to_string(count) + " points (3 floats each)");
}

// The mean squared 3-nearest distances of `count` points into `out`, with a
// workspace of its own for the call.
void neighbor_distances(dio_ctx* ctx, const dio_buf* positions, uint32_t count, float* out) {
  if (count == 0) return;
  if (count >= (1u << 30)) invalid("at most 2^30 points");
  dio_buf* ws = alloc(ctx, dio::neighbors_workspace_bytes(count), true);  // the sort's part starts zeroed
  try {
    dio::mean_neighbor_distance2(ctx->stream, static_cast<const float*>(positions->ptr), count, out, ws->ptr);
    dio::sync(ctx);
  } catch (...) {
    release(ws);
    throw;
  }
  release(ws);
}

}  // namespace

extern "C" int dio_mean_neighbor_distance2(dio_ctx* ctx, const dio_buf* positions, uint32_t count,
                                           dio_buf* distance2) {
  return guard([&] {
    need(ctx, "ctx");
    check_points(positions, count);
    need(distance2, "distance2");
    if (distance2->bytes < static_cast<size_t>(count) * 4) invalid("distance2 holds fewer than count floats");
    neighbor_distances(ctx, positions, count, static_cast<float*>(distance2->ptr));
  });
}

extern "C" int dio_initial_gaussians(dio_ctx* ctx, const dio_buf* positions, const dio_buf* colors, uint32_t count,
                                     dio_buf* params, dio_buf* sh) {
  return guard([&] {
    need(ctx, "ctx");
    check_points(positions, count);
    need(colors, "colors");
    need(params, "params");
    need(sh, "sh");
    if (colors->bytes < static_cast<size_t>(count) * 3) invalid("colors hold fewer than count rgb triples");
    if (params->bytes < static_cast<size_t>(count) * dio::kFloatsPerGaussian * 4 ||
        sh->bytes < static_cast<size_t>(count) * dio::kShRest * 4)
      invalid("params / sh hold fewer than count Gaussians");
    if (count == 0) return;
    dio_buf* d2 = alloc(ctx, static_cast<size_t>(count) * 4);
    try {
      neighbor_distances(ctx, positions, count, static_cast<float*>(d2->ptr));
      dio::initial_gaussians(ctx->stream, static_cast<const float*>(positions->ptr),
                             static_cast<const uint8_t*>(colors->ptr), static_cast<const float*>(d2->ptr), count,
                             static_cast<float*>(params->ptr), static_cast<float*>(sh->ptr));
      dio::sync(ctx);
    } catch (...) {
      release(d2);
      throw;
    }
    release(d2);
  });
}

// ================================================================= COLMAP

namespace {

// Little-endian reads over one COLMAP file, failing on a short file.
struct ColmapReader {
  const uint8_t* p;
  size_t size, at = 0;
  const char* what;
  void need_bytes(size_t n) {
    if (at + n > size || at + n < at) invalid(std::string("The COLMAP file ") + what + " is shorter than its contents say.");
  }
  template <class T>
  T read() {
    need_bytes(sizeof(T));
    T v;
    std::memcpy(&v, p + at, sizeof(T));
    at += sizeof(T);
    return v;
  }
  void skip(uint64_t n) {
    if (n > size) invalid(std::string("The COLMAP file ") + what + " is shorter than its contents say.");
    need_bytes(static_cast<size_t>(n));
    at += static_cast<size_t>(n);
  }
  std::string text() {
    const size_t start = at;
    while (true) {
      need_bytes(1);
      if (p[at++] == 0) break;
    }
    return std::string(reinterpret_cast<const char*>(p + start), at - 1 - start);
  }
};

struct Intrinsics {
  uint64_t width, height;
  double fx, fy;
};

}  // namespace

extern "C" int dio_colmap_parse(const void* cameras_bin, size_t cameras_bytes, const void* images_bin,
                                size_t images_bytes, const void* points_bin, size_t points_bytes,
                                dio_colmap_camera* cameras, uint32_t camera_capacity, uint32_t* camera_count,
                                float* positions, uint8_t* colors, uint32_t point_capacity, uint32_t* point_count) {
  return guard([&] {
    need(cameras_bin, "cameras_bin");
    need(images_bin, "images_bin");
    need(points_bin, "points_bin");
    need(camera_count, "camera_count");
    need(point_count, "point_count");
    // cameras.bin: the intrinsics of undistorted pinhole models (src/colmap.ts:72-94)
    ColmapReader rc{static_cast<const uint8_t*>(cameras_bin), cameras_bytes, 0, "cameras.bin"};
    std::vector<std::pair<int32_t, Intrinsics>> intrinsics;
    for (uint64_t n = rc.read<uint64_t>(); n > 0; n--) {
      const int32_t id = rc.read<int32_t>(), model = rc.read<int32_t>();
      Intrinsics k{rc.read<uint64_t>(), rc.read<uint64_t>(), 0, 0};
      if (model == 0) {  // SIMPLE_PINHOLE: f, cx, cy
        k.fx = k.fy = rc.read<double>();
        rc.skip(16);
      } else if (model == 1) {  // PINHOLE: fx, fy, cx, cy
        k.fx = rc.read<double>();
        k.fy = rc.read<double>();
        rc.skip(16);
      } else {
        invalid("Camera " + std::to_string(id) + " uses COLMAP model " + std::to_string(model) +
                "; only undistorted pinhole cameras are supported.");
      }
      intrinsics.emplace_back(id, k);
    }
    // images.bin: each photo's pose, world to camera x_c = R x_w + t (src/colmap.ts:96-109)
    ColmapReader ri{static_cast<const uint8_t*>(images_bin), images_bytes, 0, "images.bin"};
    std::vector<dio_colmap_camera> photos;
    for (uint64_t n = ri.read<uint64_t>(); n > 0; n--) {
      ri.read<int32_t>();  // image id
      double q[4], t[3];
      for (double& v : q) v = ri.read<double>();  // (w, x, y, z)
      for (double& v : t) v = ri.read<double>();
      const int32_t camera_id = ri.read<int32_t>();
      const std::string name = ri.text();
      const uint64_t points2d = ri.read<uint64_t>();
      if (points2d > ri.size / 24) invalid("The COLMAP file images.bin is shorter than its contents say.");
      ri.skip(points2d * 24);
      const Intrinsics* k = nullptr;
      for (const auto& c : intrinsics)
        if (c.first == camera_id) k = &c.second;
      if (!k) invalid("Photo " + name + " uses camera " + std::to_string(camera_id) + ", which isn't in cameras.bin.");
      if (name.size() >= sizeof(dio_colmap_camera::name))
        invalid("Photo name " + name + " is longer than " + std::to_string(sizeof(dio_colmap_camera::name) - 1) +
                " bytes.");
      dio_colmap_camera c{};
      std::memcpy(c.name, name.data(), name.size());
      c.width = static_cast<uint32_t>(k->width);
      c.height = static_cast<uint32_t>(k->height);
      c.fx = k->fx;
      c.fy = k->fy;
      // R of the (w, x, y, z) quaternion; its rows are the camera's axes in world
      // space (x right, y down, z forward) and the centre is -R^T t.
      const double w = q[0], x = q[1], y = q[2], z = q[3];
      const double r[3][3] = {{1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)},
                              {2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)},
                              {2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)}};
      for (int col = 0; col < 3; col++) {
        c.position[col] = -(r[0][col] * t[0] + r[1][col] * t[1] + r[2][col] * t[2]);
        c.right[col] = r[0][col];
        c.down[col] = r[1][col];
        c.forward[col] = r[2][col];
      }
      photos.push_back(c);
    }
    // sorted by file name, as the reference (src/colmap.ts:31)
    std::stable_sort(photos.begin(), photos.end(), [](const dio_colmap_camera& a, const dio_colmap_camera& b) {
      return std::strcmp(a.name, b.name) < 0;
    });
    // points3D.bin: xyz as float32 and rgb (src/colmap.ts:111-124)
    ColmapReader rp{static_cast<const uint8_t*>(points_bin), points_bytes, 0, "points3D.bin"};
    const uint64_t points = rp.read<uint64_t>();
    if (points > rp.size / 43) invalid("The COLMAP file points3D.bin is shorter than its contents say.");
    *camera_count = static_cast<uint32_t>(photos.size());
    *point_count = static_cast<uint32_t>(points);
    if (!cameras && !positions && !colors) return;  // counts only
    if (cameras) {
      if (camera_capacity < photos.size()) invalid("cameras holds fewer than the reconstruction's photos");
      std::copy(photos.begin(), photos.end(), cameras);
    }
    if ((positions || colors) && point_capacity < points) invalid("positions / colors hold fewer than its points");
    for (uint64_t i = 0; i < points; i++) {
      rp.read<uint64_t>();  // point id
      for (int k = 0; k < 3; k++) {
        const float v = static_cast<float>(rp.read<double>());
        if (positions) positions[3 * i + k] = v;
      }
      for (int k = 0; k < 3; k++) {
        const uint8_t v = rp.read<uint8_t>();
        if (colors) colors[3 * i + k] = v;
      }
      rp.read<double>();  // reprojection error
      const uint64_t track = rp.read<uint64_t>();
      if (track > rp.size / 8) invalid("The COLMAP file points3D.bin is shorter than its contents say.");
      rp.skip(track * 8);
    }
  });
}

extern "C" int dio_colmap_training_camera(const dio_colmap_camera* camera, uint32_t width, uint32_t height,
                                          dio_camera* out) {
  return guard([&] {
    need(camera, "camera");
    need(out, "out");
    if (width == 0 || height == 0) invalid("the viewport is empty");
    if (!(camera->fy > 0) || camera->height == 0) invalid("the camera has no focal length");
    const double* p = camera->position;
    const double up[3] = {-camera->down[0], -camera->down[1], -camera->down[2]};
    const double back[3] = {-camera->forward[0], -camera->forward[1], -camera->forward[2]};
    auto dot = [p](const double* a) { return a[0] * p[0] + a[1] * p[1] + a[2] * p[2]; };
    // cameraView (src/sceneCameras.ts:54-67): x right, y up, looking down -z; column-major
    const double view[16] = {camera->right[0], up[0], back[0], 0,
                             camera->right[1], up[1], back[1], 0,
                             camera->right[2], up[2], back[2], 0,
                             -dot(camera->right), -dot(up), -dot(back), 1};
    // perspective(verticalFov, aspect, 0.01, 1000) (src/loadDataset.ts:80-86, src/mat4.ts:39-48),
    // WebGPU's clip depth 0..1
    const double fov = 2 * std::atan(camera->height / (2 * camera->fy));
    const double f = 1 / std::tan(fov / 2), aspect = static_cast<double>(width) / height;
    const double near = 0.01, far = 1000, depth = 1 / (near - far);
    const double proj[16] = {f / aspect, 0, 0, 0, 0, f, 0, 0, 0, 0, far * depth, -1, 0, 0, near * far * depth, 0};
    for (int k = 0; k < 16; k++) {
      out->view[k] = static_cast<float>(view[k]);
      out->proj[k] = static_cast<float>(proj[k]);
    }
    out->viewport[0] = static_cast<float>(width);
    out->viewport[1] = static_cast<float>(height);
  });
}
