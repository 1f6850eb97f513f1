// What sits either side of the training path, over the C ABI: the starting
// Gaussians from structure-from-motion points (src/loadDataset.ts:87-149, the
// kernels in knn.hip) and the 3DGS .ply writer and loader (src/plyWriter.ts:14-43,
// src/plyLoader.ts:21-110). The .ply header is text handled here; the body is
// packed and unpacked by ply.hip on the device, so a trainer's Gaussians go to a
// file straight from HBM and a loaded scene lands in HBM in the renderer's layout.
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
  dio_buf* ws = alloc(ctx, dio::neighbors_workspace_bytes(count));
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
