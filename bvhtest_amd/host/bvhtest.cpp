// bvhtest -- JSON command host of the MI355X BVH pipeline.
//
// Keeps the reference's command interface (src/main.cpp:59-244,
// lib/base/Config.cpp:26-66, lib/base/Command.cpp:35-69): a config
// {baseConfig{name, verbose, basePath, input[]}, commands[{cmd, options}]}
// is run once per input; each command checks its prerequisites (logged
// only), runs setup -> runImpl (timed) -> teardown, and the first result
// other than OK / WARNING ends that input.  The config file is rewritten
// with every default filled in.  GPU commands call the C ABI of
// libbvhtest_amd.so (include/bvhtest_amd.h).
//
// Commands: load, genScene (synthetic meshes), export, camImport,
// copyToGPU, LBVH, bittner13GPU, CUDATracer, copyToHost, validate, status,
// writeData, BVHExport, BVHImport.
#include "bvhtest_amd.h"

#include <json.hpp>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <functional>
#include <iostream>
#include <map>
#include <memory>
#include <regex>
#include <string>
#include <vector>

namespace fs = std::filesystem;
using json = nlohmann::json;

namespace {

// ErrorCode -- lib/base/Command.hpp:33-42
enum class ErrorCode { OK, WARNING, IO_ERROR, PARSE_ERROR, BVH_ERROR, GL_ERROR, CUDA_ERROR, GENERIC_ERROR };
const char *toStr(ErrorCode e) {
  static const char *n[] = {"OK", "WARNING", "IO_ERROR", "PARSE_ERROR", "BVH_ERROR", "GL_ERROR", "CUDA_ERROR",
                            "GENERIC_ERROR"};
  return n[(int)e];
}

// CommandType -- lib/base/State.hpp:32-43
enum : uint64_t {
  IMPORT = 1u << 0,
  BVH_BUILD = 1u << 4,
  BVH_OPT1 = 1u << 5,
  BVH_OPT2 = 1u << 6,
  CUDA_DONE = 1u << 7,
  CUDA_INIT = 1u << 8,
  RAY_TRACE = 1u << 16,
  EXPORT = 1u << 24,
  SUMMARY = 1u << 28,
};

bool gVerbose = false;
#define LOG_INFO(...)                         \
  do {                                        \
    if (gVerbose) {                           \
      std::fprintf(stdout, "[info] " __VA_ARGS__); \
      std::fprintf(stdout, "\n");             \
    }                                         \
  } while (0)
#define LOG_WARN(...)                          \
  do {                                         \
    std::fprintf(stderr, "[warning] " __VA_ARGS__); \
    std::fprintf(stderr, "\n");                \
  } while (0)
#define LOG_ERR(...)                         \
  do {                                       \
    std::fprintf(stderr, "[error] " __VA_ARGS__); \
    std::fprintf(stderr, "\n");              \
  } while (0)

struct Camera {  // lib/misc/Camera.hpp:27-36, CameraBase.hpp:28-31
  bvht_vec3 pos = {0.0f, 0.0f, 2.0f}, lookAt = {0.0f, 0.0f, 0.0f}, up = {0.0f, 1.0f, 0.0f};
  float fov = 45.0f;
  uint32_t width = 2560, height = 1440;
};

struct State {  // lib/base/State.hpp:46-99
  uint64_t commandsRun = 0;
  std::string basePath, input, name;
  struct CMD {
    std::string name;
    uint64_t type;
    double ms;
  };
  std::vector<CMD> commands;
  struct OptStep {
    uint32_t step;
    float sah;
    double ms;
  };
  std::vector<OptStep> optSteps;
  uint32_t numSkipped = 0, numTotal = 0;
  std::vector<bvht_vec3> vert, norm;
  std::vector<bvht_Triangle> faces;
  std::vector<Camera> cameras;
  struct Image {
    std::string name;
    uint32_t width = 0, height = 0;
    std::vector<bvht_CUDAPixel> pixels;
    double ns = 0;
  };
  std::vector<Image> work;
  std::vector<bvht_BVHNode> bvh;  // host BVH
  uint32_t root = 0;
  uint16_t maxLevel = 0;
  bvht_MeshRaw rawMesh = {};
  bvht_CUDAMemoryBVHPointer dbvh = {};
};

std::string libError() {
  std::string e = bvht_getLastError();
  bvht_clearError();
  return e;
}

// ---------------------------------------------------------------------------
// host BVH utilities (lib/base/BVH.cpp:23-76, lib/misc/Validate.cpp:42-263)
// ---------------------------------------------------------------------------

float boxSA(const bvht_AABB &b) {
  float dx = b.minMax[1].x - b.minMax[0].x, dy = b.minMax[1].y - b.minMax[0].y, dz = b.minMax[1].z - b.minMax[0].z;
  return 2.0f * (dx * dy + dx * dz + dy * dz);
}

// BVH::fixLevels -- bit-stack DFS from the root
uint16_t fixLevels(std::vector<bvht_BVHNode> &b, uint32_t root) {
  if (b.empty()) return 0;
  unsigned __int128 stack = 0;
  uint16_t level = 0, maxLevel = 0;
  uint32_t node = root;
  while (true) {
    b[node].level = level;
    if (b[node].isLeafFlag == 0) {
      stack = (stack << 1) | 1;
      node = b[node].left;
      level++;
      maxLevel = std::max(level, maxLevel);
      continue;
    }
    while ((stack & 1) == 0) {
      if (stack == 0) return maxLevel;
      node = b[node].parent;
      stack >>= 1;
      level--;
    }
    node = b[b[node].parent].right;
    stack ^= 1;
  }
}

// BVH::calcSAH (f64 accumulation)
double calcSAH(const std::vector<bvht_BVHNode> &b, uint32_t root, float cInner = 1.0f, float cLeaf = 1.0f) {
  if (b.empty()) return -1.0;
  double s = 0.0;
  for (size_t i = 0; i < b.size(); ++i) {
    if (i == root) continue;
    float c = b[i].surfaceArea;
    c *= b[i].isLeafFlag ? cLeaf : cInner;
    s += c;
  }
  return s / b[root].surfaceArea;
}

// Adler-32 (RFC 1950, section 8.2), the checksum of the BVH dump format.
uint32_t adler32(const uint8_t *d, size_t n) {
  uint32_t a = 1, b = 0;
  while (n > 0) {
    size_t k = n < 3800 ? n : 3800;  // keep b below 2^32 before the modulo
    n -= k;
    for (size_t i = 0; i < k; ++i) {
      a += d[i];
      b += a;
    }
    d += k;
    a %= 65521u;
    b %= 65521u;
  }
  return (b << 16) | a;
}

fs::path inputDir(const State &s) {
  fs::path p = fs::path(s.basePath) / s.input;
  if (fs::is_directory(p)) return p;
  if (fs::is_regular_file(p)) return p.parent_path();
  return fs::path();
}

bvht_vec3 vecFromJSON(const json &j, bvht_vec3 d) {
  d.x = j.value("x", d.x);
  d.y = j.value("y", d.y);
  d.z = j.value("z", d.z);
  return d;
}
json vecToJSON(const bvht_vec3 &v) { return json{{"x", v.x}, {"y", v.y}, {"z", v.z}}; }

// ---------------------------------------------------------------------------
// commands
// ---------------------------------------------------------------------------

class Command {
 public:
  virtual ~Command() = default;
  virtual std::string name() const = 0;
  virtual std::string desc() const = 0;
  virtual uint64_t type() const = 0;
  virtual uint64_t required() const { return 0; }
  virtual void fromJSON(const json &) {}
  virtual json toJSON() const { return json::object(); }
  virtual ErrorCode setup(State &) { return ErrorCode::OK; }
  virtual ErrorCode runImpl(State &) = 0;
  virtual void teardown(State &) {}

  // Command::run -- lib/base/Command.cpp:35-69
  ErrorCode run(State &s) {
    LOG_INFO("Running: command %-16s -- %s", name().c_str(), desc().c_str());
    if ((s.commandsRun & required()) != required())
      LOG_ERR("Command %s has unmet requirements (required 0x%llx, executed 0x%llx)", name().c_str(),
              (unsigned long long)required(), (unsigned long long)s.commandsRun);
    ErrorCode r = setup(s);
    if (r != ErrorCode::OK) {
      LOG_ERR("Setup function of command %s returned %s", name().c_str(), toStr(r));
      return r;
    }
    auto t0 = std::chrono::high_resolution_clock::now();
    r = runImpl(s);
    auto t1 = std::chrono::high_resolution_clock::now();
    teardown(s);
    if (r == ErrorCode::WARNING) LOG_WARN("Command %s returned a warning", name().c_str());
    if (r != ErrorCode::OK && r != ErrorCode::WARNING) {
      LOG_ERR("Command %s returned %s", name().c_str(), toStr(r));
      return r;
    }
    s.commandsRun |= type();
    s.commands.push_back({name(), type(), std::chrono::duration<double, std::milli>(t1 - t0).count()});
    return ErrorCode::OK;
  }
};

// load -- lib/io/Load.cpp:41-187 (mesh format versions 1 and 3; version 2
// is LZO-compressed and not supported here)
class Load : public Command {
 public:
  std::string name() const override { return "load"; }
  std::string desc() const override { return "Load a mesh (data.json + raw binary)"; }
  uint64_t type() const override { return IMPORT; }
  ErrorCode runImpl(State &s) override {
    fs::path p = fs::absolute(s.basePath) / s.input;
    if (fs::is_directory(p)) p /= "data.json";
    if (!fs::exists(p) || !fs::is_regular_file(p)) {
      LOG_ERR("File %s does not exist", p.string().c_str());
      return ErrorCode::IO_ERROR;
    }
    std::ifstream ctl(p);
    json cfg;
    ctl >> cfg;
    uint32_t vers = cfg.at("version").get<uint32_t>();
    fs::path bin = p.parent_path() / cfg.at("bin").get<std::string>();
    if (!fs::exists(bin) || !fs::is_regular_file(bin)) {
      LOG_ERR("File %s does not exist", bin.string().c_str());
      return ErrorCode::IO_ERROR;
    }
    if (vers != 1 && vers != 3) {
      LOG_ERR("File format version is %u but 1 or 3 is required", vers);
      return ErrorCode::PARSE_ERROR;
    }
    size_t nv = cfg.at("mesh").at("vert").get<size_t>();
    size_t nn = cfg.at("mesh").at("normals").get<size_t>();
    size_t nf = cfg.at("mesh").at("faces").get<size_t>();
    std::ifstream b(bin, std::ios::binary);
    s.vert.resize(nv);
    s.norm.resize(nn);
    s.faces.resize(nf);
    b.read(reinterpret_cast<char *>(s.vert.data()), nv * 12);
    b.read(reinterpret_cast<char *>(s.norm.data()), nn * 12);
    b.read(reinterpret_cast<char *>(s.faces.data()), nf * 12);
    if (!b) {
      LOG_ERR("File %s is truncated", bin.string().c_str());
      return ErrorCode::IO_ERROR;
    }
    for (auto &f : s.faces)
      if (f.v1 >= nv || f.v2 >= nv || f.v3 >= nv) {
        LOG_ERR("Face references a vertex beyond %zu", nv);
        return ErrorCode::PARSE_ERROR;
      }
    return ErrorCode::OK;
  }
};

// genScene (extension): the synthetic scenes of the library, as an import
class GenScene : public Command {
  std::string vKind = "heightfield";
  uint32_t vGX = 256, vGZ = 128, vSeed = 1, vSpheres = 2048, vSoup = 500000, vTris = 100000;
  float vAmp = 0.2f;

 public:
  std::string name() const override { return "genScene"; }
  std::string desc() const override { return "Generate a synthetic mesh (heightfield | spheresSoup | soup)"; }
  uint64_t type() const override { return IMPORT; }
  void fromJSON(const json &j) override {
    vKind = j.value("type", vKind);
    vGX = j.value("gx", vGX);
    vGZ = j.value("gz", vGZ);
    vSeed = j.value("seed", vSeed);
    vAmp = j.value("amplitude", vAmp);
    vSpheres = j.value("numSpheres", vSpheres);
    vSoup = j.value("numSoup", vSoup);
    vTris = j.value("numTris", vTris);
  }
  json toJSON() const override {
    return json{{"type", vKind}, {"gx", vGX}, {"gz", vGZ}, {"seed", vSeed}, {"amplitude", vAmp},
                {"numSpheres", vSpheres}, {"numSoup", vSoup}, {"numTris", vTris}};
  }
  ErrorCode runImpl(State &s) override {
    uint32_t nv = 0, nf = 0;
    auto fill = [&](auto gen) {
      gen(nullptr, nullptr, &nv, &nf);
      s.vert.resize(nv);
      s.faces.resize(nf);
      gen(s.vert.data(), s.faces.data(), &nv, &nf);
    };
    if (vKind == "heightfield")
      fill([&](bvht_vec3 *v, bvht_Triangle *f, uint32_t *a, uint32_t *b) { bvht_sceneHeightfield(vGX, vGZ, vSeed, vAmp, v, f, a, b); });
    else if (vKind == "spheresSoup")
      fill([&](bvht_vec3 *v, bvht_Triangle *f, uint32_t *a, uint32_t *b) { bvht_sceneSpheresSoup(vSpheres, vSoup, vSeed, v, f, a, b); });
    else if (vKind == "soup")
      fill([&](bvht_vec3 *v, bvht_Triangle *f, uint32_t *a, uint32_t *b) { bvht_sceneSoup(vTris, vSeed, 0.01f, v, f, a, b); });
    else {
      LOG_ERR("Unknown scene type %s", vKind.c_str());
      return ErrorCode::PARSE_ERROR;
    }
    s.norm.clear();
    return ErrorCode::OK;
  }
};

// export -- lib/io/ExportMesh.cpp:53-124 (mesh format version 3)
class ExportMesh : public Command {
  std::string vOut = "data";

 public:
  std::string name() const override { return "export"; }
  std::string desc() const override { return "Export the mesh (format version 3)"; }
  uint64_t type() const override { return EXPORT; }
  uint64_t required() const override { return IMPORT; }
  void fromJSON(const json &j) override { vOut = j.value("outDir", vOut); }
  json toJSON() const override { return json{{"outDir", vOut}}; }
  ErrorCode runImpl(State &s) override {
    fs::path dir = fs::path(s.basePath) / vOut;
    fs::create_directories(dir);
    json ctl = {{"version", 3}, {"bin", "data.bin"},
                {"mesh", {{"vert", s.vert.size()}, {"normals", s.norm.size()}, {"faces", s.faces.size()}}}};
    std::ofstream(dir / "data.json") << ctl.dump(2);
    std::ofstream b(dir / "data.bin", std::ios::binary);
    b.write(reinterpret_cast<const char *>(s.vert.data()), s.vert.size() * 12);
    b.write(reinterpret_cast<const char *>(s.norm.data()), s.norm.size() * 12);
    b.write(reinterpret_cast<const char *>(s.faces.data()), s.faces.size() * 12);
    return b ? ErrorCode::OK : ErrorCode::IO_ERROR;
  }
};

// camImport -- lib/io/CameraImport.cpp:42-108 (cameras.json version 1)
class CamImport : public Command {
  bool vAbort = true;

 public:
  std::string name() const override { return "camImport"; }
  std::string desc() const override { return "Import cameras (cameras.json)"; }
  uint64_t type() const override { return IMPORT; }
  void fromJSON(const json &j) override { vAbort = j.value("abortOnError", vAbort); }
  json toJSON() const override { return json{{"abortOnError", vAbort}}; }
  ErrorCode runImpl(State &s) override {
    ErrorCode fail = vAbort ? ErrorCode::IO_ERROR : ErrorCode::OK;
    fs::path dir = inputDir(s);
    if (dir.empty()) {
      LOG_ERR("Invalid base path %s", (fs::path(s.basePath) / s.input).string().c_str());
      return fail;
    }
    fs::path p = fs::absolute(dir) / "cameras.json";
    if (!fs::is_regular_file(p)) {
      LOG_ERR("Path '%s' is not a regular file", p.string().c_str());
      return fail;
    }
    std::ifstream f(p);
    json d;
    f >> d;
    if (d.at("version").get<uint32_t>() != 1) {
      LOG_ERR("Incompatible camera file version");
      return fail;
    }
    for (const json &c : d.at("cameras")) {
      if (c.at("type").get<std::string>() != "PERSPECTIVE") {
        LOG_WARN("Currently only PERSPECTIVE type cameras supported");
        continue;
      }
      const json &cam = c.at("cam");
      Camera k;
      json cc = cam.value("camera", json::object()), res = cam.value("resolution", json::object());
      k.pos = vecFromJSON(cc.value("pos", json::object()), k.pos);
      k.lookAt = vecFromJSON(cc.value("lookAt", json::object()), k.lookAt);
      k.up = vecFromJSON(cc.value("up", json::object()), k.up);
      k.fov = cc.value("fov", k.fov);
      k.width = res.value("width", k.width);
      k.height = res.value("height", k.height);
      s.cameras.push_back(k);
    }
    return ErrorCode::OK;
  }
};

// copyToGPU -- lib/cuda/CopyToGPU.cpp:30-34
class CopyToGPU : public Command {
 public:
  std::string name() const override { return "copyToGPU"; }
  std::string desc() const override { return "Copy the mesh (and a host BVH) to the GPU"; }
  uint64_t type() const override { return CUDA_INIT; }
  uint64_t required() const override { return IMPORT; }
  ErrorCode runImpl(State &s) override {
    if (!copyBVHToGPU(s.bvh.data(), (uint32_t)s.bvh.size(), s.root, s.maxLevel, &s.dbvh)) {
      LOG_ERR("%s", libError().c_str());
      return ErrorCode::CUDA_ERROR;
    }
    if (!copyMeshToGPU(s.vert.data(), (uint32_t)s.vert.size(), s.norm.data(), (uint32_t)s.norm.size(), s.faces.data(),
                       (uint32_t)s.faces.size(), &s.rawMesh)) {
      LOG_ERR("%s", libError().c_str());
      return ErrorCode::CUDA_ERROR;
    }
    return ErrorCode::OK;
  }
};

// copyToHost -- lib/cuda/CopyToHost.cpp:30-36
class CopyToHost : public Command {
  bool vFixLevels = true;

 public:
  std::string name() const override { return "copyToHost"; }
  std::string desc() const override { return "Copy the BVH and mesh back to the host"; }
  uint64_t type() const override { return CUDA_DONE; }
  uint64_t required() const override { return CUDA_INIT; }
  void fromJSON(const json &j) override { vFixLevels = j.value("fixLevels", vFixLevels); }
  json toJSON() const override { return json{{"fixLevels", vFixLevels}}; }
  ErrorCode runImpl(State &s) override {
    if (s.dbvh.nodes) {
      s.bvh.resize(s.dbvh.numNodes);
      uint32_t n = 0;
      if (!copyBVHToHost(&s.dbvh, s.bvh.data(), (uint32_t)s.bvh.size(), &n, &s.root, &s.maxLevel)) {
        LOG_ERR("%s", libError().c_str());
        return ErrorCode::CUDA_ERROR;
      }
      s.bvh.resize(n);
    }
    if (s.rawMesh.faces) {
      s.vert.resize(s.rawMesh.numVert);
      s.norm.resize(s.rawMesh.numNorm);
      s.faces.resize(s.rawMesh.numFaces);
      if (!copyMeshToHost(&s.rawMesh, s.vert.data(), s.norm.data(), s.faces.data())) {
        LOG_ERR("%s", libError().c_str());
        return ErrorCode::CUDA_ERROR;
      }
    }
    resetCUDA();
    if (vFixLevels) s.maxLevel = fixLevels(s.bvh, s.root);
    return ErrorCode::OK;
  }
};

// LBVH -- lib/builder/LBVH.cpp:28-55 (+ options mortonBits / topology)
class LBVHCmd : public Command {
  float vCostInner = 1.2f, vCostLeaf = 1.0f;  // BuilderBase.cpp:27-32 (unused by LBVH)
  uint32_t vMortonBits = 30, vTopology = 0;
  bvht_LBVH_WorkingMemory vMem = {};

 public:
  std::string name() const override { return "LBVH"; }
  std::string desc() const override { return "LBVH GPU builder (gfx950)"; }
  uint64_t type() const override { return BVH_BUILD; }
  uint64_t required() const override { return IMPORT | CUDA_INIT; }
  void fromJSON(const json &j) override {
    vCostInner = j.value("costInner", vCostInner);
    vCostLeaf = j.value("costLeaf", vCostLeaf);
    vMortonBits = j.value("mortonBits", vMortonBits);
    vTopology = j.value("topology", vTopology);
  }
  json toJSON() const override {
    return json{{"costInner", vCostInner}, {"costLeaf", vCostLeaf}, {"mortonBits", vMortonBits}, {"topology", vTopology}};
  }
  ErrorCode setup(State &s) override {
    bvht_LBVH_Options o = bvht_LBVH_defaultOptions();
    o.mortonBits = vMortonBits;
    o.topology = vTopology;
    vMem = bvht_LBVH_allocateWorkingMemoryEx(&s.rawMesh, &o);
    if (!vMem.lRes) {
      LOG_ERR("%s", libError().c_str());
      return ErrorCode::CUDA_ERROR;
    }
    bvht_freeBVH(&s.dbvh);  // copyToGPU of an empty host BVH leaves nothing here
    if (!LBVH_allocateBVH(&s.dbvh, &s.rawMesh)) {
      LOG_ERR("%s", libError().c_str());
      LBVH_freeWorkingMemory(&vMem);
      return ErrorCode::CUDA_ERROR;
    }
    return ErrorCode::OK;
  }
  ErrorCode runImpl(State &s) override {
    bool ok = bvht_LBVH_build(&vMem, &s.rawMesh, &s.dbvh);
    LBVH_doCUDASyc();
    if (!ok || *bvht_getLastError()) {
      LOG_ERR("%s", libError().c_str());
      return ErrorCode::CUDA_ERROR;
    }
    s.root = 0;
    return ErrorCode::OK;
  }
  void teardown(State &) override { LBVH_freeWorkingMemory(&vMem); }
};

// Reads a little-endian fp32 .npy vector (format 1.x/2.x, C order).
bool loadNpyF32(const std::string &path, std::vector<float> &out) {
  std::ifstream in(path, std::ios::binary);
  if (!in) return false;
  char magic[6];
  uint8_t ver[2];
  if (!in.read(magic, 6) || std::memcmp(magic, "\x93NUMPY", 6) != 0 || !in.read(reinterpret_cast<char *>(ver), 2))
    return false;
  uint32_t hlen = 0;
  if (ver[0] == 1) {
    uint16_t h16 = 0;
    if (!in.read(reinterpret_cast<char *>(&h16), 2)) return false;
    hlen = h16;
  } else if (!in.read(reinterpret_cast<char *>(&hlen), 4)) {
    return false;
  }
  std::string hdr(hlen, ' ');
  if (!in.read(&hdr[0], hlen)) return false;
  if (hdr.find("'<f4'") == std::string::npos || hdr.find("'fortran_order': False") == std::string::npos) return false;
  std::vector<char> data((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
  if (data.size() % 4) return false;
  out.resize(data.size() / 4);
  std::memcpy(out.data(), data.data(), data.size());
  return true;
}

// bittner13GPU -- lib/builder/Bittner13GPU.cpp:46-133 with the
// OptimizerBase benchmark hooks (lib/builder/OptimizerBase.cpp:28-57).
// Extension options: "priority" ("plain": the reference heuristic;
// "learned": the MLP in "priorityModel", a .npy of bvht_mlpNumParams() fp32
// values).  A learned step that does not lower the unit SAH hands the next
// step to the plain heuristic.
class Bittner13GPUCmd : public Command {
  uint32_t vMaxNumStepps = 500, vNumChunks = 32, vBlock = 128;
  float vBatchPercent = 1.0f;
  bool vSort = true, vOffsetAccess = false, vAltFindNode = false, vAltFixTree = false, vAltSort = false,
       vLocalPatchCPY = false, vCalcSAH = false, vDoSync = false;
  std::string vPriority = "plain", vModelPath;
  bvht_PriorityModel *vModel = nullptr;
  bvht_GPUWorkingMemory vMem = {};
  uint32_t chunkSize(const State &s) const {
    uint32_t n = (uint32_t)((vBatchPercent / 100.0f) * (float)s.dbvh.numNodes);
    return n / vNumChunks;
  }

 public:
  std::string name() const override { return "bittner13GPU"; }
  std::string desc() const override { return "GPU parallel BVH optimizer based on Bittner et al. 2013"; }
  uint64_t type() const override { return BVH_OPT1; }
  uint64_t required() const override { return BVH_BUILD | CUDA_INIT; }
  void fromJSON(const json &j) override {
    vCalcSAH = j.value("base_calcSAH", vCalcSAH);
    vDoSync = j.value("base_doSync", vDoSync);
    vMaxNumStepps = j.value("maxNumStepps", vMaxNumStepps);
    vNumChunks = j.value("numChunks", vNumChunks);
    vBlock = j.value("CUDABlockSize", vBlock);
    vBatchPercent = j.value("batchPercent", vBatchPercent);
    vSort = j.value("sort", vSort);
    vOffsetAccess = j.value("offsetAccess", vOffsetAccess);
    vAltFindNode = j.value("altFindNode", vAltFindNode);
    vAltFixTree = j.value("altFixTree", vAltFixTree);
    vAltSort = j.value("altSort", vAltSort);
    vLocalPatchCPY = j.value("localPatchCPY", vLocalPatchCPY);
    vPriority = j.value("priority", vPriority);
    vModelPath = j.value("priorityModel", vModelPath);
    if (vBatchPercent <= 0.01f) vBatchPercent = 0.01f;
    if (vBatchPercent >= 75.0f) vBatchPercent = 75.0f;
  }
  json toJSON() const override {
    return json{{"base_calcSAH", vCalcSAH}, {"base_doSync", vDoSync},     {"maxNumStepps", vMaxNumStepps},
                {"numChunks", vNumChunks},  {"CUDABlockSize", vBlock},    {"batchPercent", vBatchPercent},
                {"sort", vSort},            {"offsetAccess", vOffsetAccess}, {"altFindNode", vAltFindNode},
                {"altFixTree", vAltFixTree}, {"altSort", vAltSort},       {"localPatchCPY", vLocalPatchCPY},
                {"priority", vPriority},    {"priorityModel", vModelPath}};
  }
  ErrorCode setup(State &s) override {
    if (vPriority != "plain" && vPriority != "learned") {
      LOG_ERR("bittner13GPU: priority must be \"plain\" or \"learned\", not \"%s\"", vPriority.c_str());
      return ErrorCode::PARSE_ERROR;
    }
    vMem = allocateMemory(&s.dbvh, chunkSize(s), s.rawMesh.numFaces);
    if (!vMem.result) {
      LOG_ERR("%s", libError().c_str());
      return ErrorCode::CUDA_ERROR;
    }
    initData(&vMem, &s.dbvh, vBlock);
    if (*bvht_getLastError()) {
      LOG_ERR("%s", libError().c_str());
      return ErrorCode::CUDA_ERROR;
    }
    if (vPriority == "learned") {
      std::vector<float> params;
      if (!loadNpyF32(vModelPath, params) || params.size() != bvht_mlpNumParams()) {
        LOG_ERR("bittner13GPU: cannot read %u fp32 parameters from \"%s\"", bvht_mlpNumParams(), vModelPath.c_str());
        return ErrorCode::IO_ERROR;
      }
      vModel = bvht_mlpCreate(0);
      if (!vModel || !bvht_mlpSetParams(vModel, params.data()) || !bvht_bn13_setModel(&vMem, vModel)) {
        LOG_ERR("%s", libError().c_str());
        return ErrorCode::CUDA_ERROR;
      }
    }
    return ErrorCode::OK;
  }
  ErrorCode runImpl(State &s) override {
    bvht_AlgoCFG c;
    std::memset(&c, 0, sizeof(c));
    c.numChunks = vNumChunks;
    c.chunkSize = chunkSize(s);
    c.blockSize = vBlock;
    c.offsetAccess = vOffsetAccess;
    c.altFindNode = vAltFindNode;
    c.altFixTree = vAltFixTree;
    c.altSort = vAltSort;
    c.sort = vSort;
    c.localPatchCPY = vLocalPatchCPY;
    const bool learned = vPriority == "learned";
    bool fallbackNext = false;
    float prevSAH = learned ? CUDAcalcSAH(&s.dbvh) : 0.0f;
    s.numSkipped = 0;
    s.numTotal = 0;
    s.optSteps.push_back({0, vCalcSAH ? CUDAcalcSAH(&s.dbvh) : 0.0f, 0.0});
    if (vDoSync) doCudaDevSync();
    for (uint32_t i = 0; i < vMaxNumStepps; ++i) {
      s.optSteps.push_back({(uint32_t)s.optSteps.size(), 0.0f, 0.0});
      if (vDoSync) doCudaDevSync();
      c.priority = (learned && !fallbackNext) ? 1u : 0u;
      auto t0 = std::chrono::system_clock::now();
      bn13_doAlgorithmStep(&vMem, &s.dbvh, c);
      if (vDoSync) doCudaDevSync();
      auto t1 = std::chrono::system_clock::now();
      s.optSteps.back().ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
      if (vCalcSAH || learned) {
        const float sah = CUDAcalcSAH(&s.dbvh);
        if (vCalcSAH) s.optSteps.back().sah = sah;
        fallbackNext = learned && c.priority == 1 && sah >= prevSAH;
        prevSAH = sah;
      }
      if (*bvht_getLastError()) {
        LOG_ERR("%s", libError().c_str());
        return ErrorCode::CUDA_ERROR;
      }
    }
    doCudaDevSync();
    return ErrorCode::OK;
  }
  void teardown(State &s) override {
    uint32_t k = chunkSize(s);
    s.numSkipped = calcNumSkipped(&vMem);
    s.numTotal = k * vNumChunks * vMaxNumStepps;
    LOG_INFO("Skipped %u of %u", s.numSkipped, s.numTotal);
    freeMemory(&vMem);
    bvht_mlpFree(vModel);
    vModel = nullptr;
  }
};

// CUDATracer -- lib/tracer/CUDATracer.cpp:43-150
class CUDATracerCmd : public Command {
  bvht_vec3 vLight = {2.0f, 2.0f, 2.0f};  // TracerBase.hpp default
  bool vBundles = false;
  bvht_Ray *vRays = nullptr;
  std::vector<bvht_CUDAPixel *> vImages;
  uint32_t vW = 0, vH = 0;

 public:
  std::string name() const override { return "CUDATracer"; }
  std::string desc() const override { return "GPU ray tracer (gfx950)"; }
  uint64_t type() const override { return RAY_TRACE; }
  uint64_t required() const override { return BVH_BUILD | CUDA_INIT; }
  void fromJSON(const json &j) override {
    json t = j;
    if (t.count("lightLocation") == 0) t["lightLocation"] = json::object();
    vLight = vecFromJSON(t["lightLocation"], vLight);
    vBundles = j.value("rayBundles", vBundles);
  }
  json toJSON() const override { return json{{"lightLocation", vecToJSON(vLight)}, {"rayBundles", vBundles}}; }
  ErrorCode setup(State &s) override {
    if (s.cameras.empty()) return ErrorCode::WARNING;
    vW = s.cameras[0].width;
    vH = s.cameras[0].height;
    if (!allocateRays(&vRays, vW * vH)) return ErrorCode::CUDA_ERROR;
    for (size_t i = 0; i < s.cameras.size(); ++i) {
      bvht_CUDAPixel *img = nullptr;
      if (!allocateImage(&img, vW, vH)) return ErrorCode::CUDA_ERROR;
      vImages.push_back(img);
    }
    return ErrorCode::OK;
  }
  ErrorCode runImpl(State &s) override {
    size_t off = s.work.size();
    s.work.resize(off + s.cameras.size());
    for (size_t i = 0; i < s.cameras.size(); ++i) {
      const Camera &c = s.cameras[i];
      auto t0 = std::chrono::high_resolution_clock::now();
      generateRays(vRays, vW, vH, c.pos, c.lookAt, c.up, c.fov);
      tracerImage(vRays, vImages[i], s.dbvh.nodes, s.root, s.rawMesh, vLight, vW, vH, vBundles);
      tracerDoCudaSync();
      auto t1 = std::chrono::high_resolution_clock::now();
      s.work[off + i].ns = std::chrono::duration<double, std::nano>(t1 - t0).count();
      if (*bvht_getLastError()) {
        LOG_ERR("%s", libError().c_str());
        return ErrorCode::CUDA_ERROR;
      }
    }
    return ErrorCode::OK;
  }
  void teardown(State &s) override {
    for (size_t i = 0; i < vImages.size(); ++i) {
      State::Image &img = s.work[s.work.size() - vImages.size() + i];
      img.name = s.name + "_CUDA_cam_" + std::to_string(i);
      img.width = vW;
      img.height = vH;
      img.pixels.resize((size_t)vW * vH);
      copyImageToHost(img.pixels.data(), vImages[i], vW, vH);
      freeImage(vImages[i]);
    }
    vImages.clear();
    freeRays(&vRays);
  }
};

// validate -- lib/misc/Validate.cpp:267-277
class ValidateCmd : public Command {
 public:
  std::string name() const override { return "validate"; }
  std::string desc() const override { return "Validate the BVH (tree, boxes, surface areas, triangles)"; }
  uint64_t type() const override { return SUMMARY; }
  uint64_t required() const override { return BVH_BUILD; }
  ErrorCode runImpl(State &s) override {
    auto &b = s.bvh;
    if (b.empty()) return ErrorCode::OK;
    uint32_t treeErr = 0, boxErr = 0, strictErr = 0, saErr = 0, triErr = 0;
    // checkTree (:42-124): bit-stack DFS, levels, parent/child links, orphans
    std::vector<uint8_t> seen(b.size(), 0);
    unsigned __int128 stack = 0;
    uint32_t node = s.root, level = 0, mlev = 0;
    if (b[node].parent != node) treeErr++;
    for (uint64_t guard = 0; guard < 4 * b.size() + 16; ++guard) {
      seen[node] = 1;
      if (b[node].level != level) treeErr++;
      if (node != s.root) {
        const auto &p = b[b[node].parent];
        if (b[node].isLeft ? p.left != node : p.right != node) treeErr++;
      }
      if (!b[node].isLeafFlag) {
        if (b[node].left == b[node].right || b[node].left == node || b[node].right == node) treeErr++;
        stack = (stack << 1) | 1;
        uint32_t last = node;
        node = b[node].left;
        if (node >= b.size() || b[node].parent != last) {
          treeErr++;
          break;
        }
        mlev = std::max(mlev, ++level);
        continue;
      }
      bool end = false;
      while ((stack & 1) == 0) {
        if (stack == 0) {
          end = true;
          break;
        }
        node = b[node].parent;
        stack >>= 1;
        level--;
      }
      if (end) break;
      uint32_t last = b[node].parent;
      node = b[last].right;
      if (b[node].parent != last) treeErr++;
      stack ^= 1;
    }
    for (uint8_t x : seen) treeErr += x ? 0 : 1;
    if (mlev != s.maxLevel) treeErr++;
    std::vector<uint8_t> tri(s.faces.size(), 0);
    for (size_t i = 0; i < b.size(); ++i) {
      const auto &n = b[i];
      if (!n.isLeafFlag) {
        const bvht_AABB &x = n.bbox, &l = b[n.left].bbox, &r = b[n.right].bbox;
        bool ok = true, strict = true;
        for (int a = 0; a < 3; ++a) {
          float xm = (&x.minMax[0].x)[a], xM = (&x.minMax[1].x)[a];
          float lm = (&l.minMax[0].x)[a], lM = (&l.minMax[1].x)[a], rm = (&r.minMax[0].x)[a], rM = (&r.minMax[1].x)[a];
          ok = ok && xm <= lm && xm <= rm && xM >= lM && xM >= rM;
          strict = strict && xm == std::min(lm, rm) && xM == std::max(lM, rM);
        }
        boxErr += ok ? 0 : 1;
        strictErr += strict ? 0 : 1;
      } else {
        bool ok = true;
        for (uint32_t j = n.left; j < n.left + n.right; ++j) {
          if (j >= s.faces.size()) {
            ok = false;
            continue;
          }
          tri[j] = 1;
          const bvht_Triangle &f = s.faces[j];
          for (uint32_t vi : {f.v1, f.v2, f.v3}) {
            const bvht_vec3 &v = s.vert[vi];
            ok = ok && n.bbox.minMax[0].x <= v.x && n.bbox.minMax[0].y <= v.y && n.bbox.minMax[0].z <= v.z &&
                 n.bbox.minMax[1].x >= v.x && n.bbox.minMax[1].y >= v.y && n.bbox.minMax[1].z >= v.z;
          }
        }
        triErr += ok ? 0 : 1;
      }
      if (std::fabs(n.surfaceArea - boxSA(n.bbox)) >= 0.000001f) saErr++;
    }
    for (uint8_t x : tri) triErr += x ? 0 : 1;
    if (treeErr) LOG_ERR("%u BVH tree errors", treeErr);
    if (boxErr) LOG_ERR("%u invalid bounding boxes", boxErr);
    if (strictErr) LOG_ERR("%u invalid bounding boxes (strict)", strictErr);
    if (saErr) LOG_ERR("%u wrong pre-calculated surface areas", saErr);
    if (triErr) LOG_ERR("%u BVH triangle errors", triErr);
    return (treeErr | boxErr | strictErr | saErr | triErr) == 0 ? ErrorCode::OK : ErrorCode::WARNING;
  }
};

// status -- lib/base/StatusDump.cpp:31-56
class StatusCmd : public Command {
 public:
  std::string name() const override { return "status"; }
  std::string desc() const override { return "Print BVH, mesh, image and command statistics"; }
  uint64_t type() const override { return SUMMARY; }
  ErrorCode runImpl(State &s) override {
    std::printf("Status:\n  -- BVH:\n    - Nodes:       %zu\n    - Tree height: %u\n    - SAH:         %.6f\n",
                s.bvh.size(), (unsigned)s.maxLevel, s.bvh.empty() ? -1.0 : calcSAH(s.bvh, s.root));
    std::printf("  -- Mesh info:\n    - Vertices: %zu\n    - Faces:    %zu\n", s.vert.size(), s.faces.size());
    std::printf("  -- Num Images: %zu\n", s.work.size());
    for (auto &w : s.work) std::printf("    - Image %-16s -- Time: %12.0fns\n", w.name.c_str(), w.ns);
    std::printf("  -- Command times:\n");
    for (auto &c : s.commands) std::printf("    - %-16s -- %10.3fms\n", c.name.c_str(), c.ms);
    return ErrorCode::OK;
  }
};

// writeData -- lib/io/WriteData.cpp:37-84
class WriteData : public Command {
  std::string vOut = "%d/%i/rData_%n.json";

 public:
  std::string name() const override { return "writeData"; }
  std::string desc() const override { return "Write optimizer data as JSON"; }
  uint64_t type() const override { return SUMMARY; }
  void fromJSON(const json &j) override { vOut = j.value("outFile", vOut); }
  json toJSON() const override { return json{{"outFile", vOut}}; }
  ErrorCode runImpl(State &s) override {
    if (!fs::is_directory(s.basePath)) return ErrorCode::IO_ERROR;
    std::string f = std::regex_replace(vOut, std::regex("%d"), s.basePath);
    f = std::regex_replace(f, std::regex("%i"), s.input);
    f = std::regex_replace(f, std::regex("%n"), s.name);
    double total = 0;
    for (auto &c : s.commands)
      if (c.type == BVH_OPT1 || c.type == BVH_OPT2) total = c.ms;
    json steps = json::array();
    for (auto &o : s.optSteps) steps.push_back({{"step", o.step}, {"sah", o.sah}, {"duration", (int64_t)o.ms}});
    json d = {{"totalTime", (int64_t)total}, {"optStepps", steps}, {"numSkipped", s.numSkipped},
              {"numTotal", s.numTotal}, {"numFaces", s.faces.size()}, {"numVert", s.vert.size()},
              {"numBVHNodes", s.bvh.size()}};
    std::ofstream o(f, std::ios::trunc);
    if (!o) return ErrorCode::IO_ERROR;
    o << d.dump(2);
    return ErrorCode::OK;
  }
};

// BVHExport / BVHImport -- lib/io/BVHExport.cpp:44-117, BVHImport.cpp:42-141
// (format version 9: raw BVHNode[] + JSON control with an Adler-32 checksum)
class BVHExportCmd : public Command {
  std::string vName = "genericBVH";

 public:
  std::string name() const override { return "BVHExport"; }
  std::string desc() const override { return "Export the raw BVH (format version 9)"; }
  uint64_t type() const override { return EXPORT; }
  uint64_t required() const override { return BVH_BUILD; }
  void fromJSON(const json &j) override { vName = j.value("name", vName); }
  json toJSON() const override { return json{{"name", vName}}; }
  ErrorCode runImpl(State &s) override {
    fs::path dir = inputDir(s);
    if (dir.empty()) return ErrorCode::IO_ERROR;
    const size_t bytes = s.bvh.size() * sizeof(bvht_BVHNode);
    json ctl = {{"version", 9}, {"bin", vName + "_bvh.bin"}, {"BVHSize", s.bvh.size()}, {"treeHeight", s.maxLevel},
                {"rawChecksum", adler32(reinterpret_cast<const uint8_t *>(s.bvh.data()), bytes)}};
    std::ofstream c(fs::absolute(dir) / (vName + "_bvh.json"), std::ios::trunc);
    std::ofstream b(fs::absolute(dir) / (vName + "_bvh.bin"), std::ios::trunc | std::ios::binary);
    if (!c || !b) return ErrorCode::IO_ERROR;
    c << ctl.dump(2);
    b.write(reinterpret_cast<const char *>(s.bvh.data()), bytes);
    return ErrorCode::OK;
  }
};

class BVHImportCmd : public Command {
  std::string vName = "genericBVH";

 public:
  std::string name() const override { return "BVHImport"; }
  std::string desc() const override { return "Import a raw BVH (format version 9)"; }
  uint64_t type() const override { return BVH_BUILD; }
  uint64_t required() const override { return IMPORT; }
  void fromJSON(const json &j) override { vName = j.value("name", vName); }
  json toJSON() const override { return json{{"name", vName}}; }
  ErrorCode runImpl(State &s) override {
    fs::path dir = inputDir(s);
    fs::path cp = fs::absolute(dir) / (vName + "_bvh.json");
    if (dir.empty() || !fs::is_regular_file(cp)) {
      LOG_ERR("Invalid BVH control file %s", cp.string().c_str());
      return ErrorCode::IO_ERROR;
    }
    std::ifstream c(cp);
    json ctl;
    c >> ctl;
    if (ctl.at("version").get<uint32_t>() != 9) {
      LOG_ERR("Incompatible BVH file version");
      return ErrorCode::IO_ERROR;
    }
    fs::path bp = dir / ctl.at("bin").get<std::string>();
    uint32_t n = ctl.at("BVHSize").get<uint32_t>();
    s.maxLevel = ctl.at("treeHeight").get<uint16_t>();
    uint32_t sum = ctl.at("rawChecksum").get<uint32_t>();
    std::ifstream b(bp, std::ios::binary);
    if (!b) {
      LOG_ERR("Invalid BVH binary file %s", bp.string().c_str());
      return ErrorCode::IO_ERROR;
    }
    s.bvh.resize(n);
    b.read(reinterpret_cast<char *>(s.bvh.data()), (std::streamsize)n * sizeof(bvht_BVHNode));
    if (!b || adler32(reinterpret_cast<const uint8_t *>(s.bvh.data()), n * sizeof(bvht_BVHNode)) != sum) {
      LOG_ERR("BVH binary file %s: corrupt data", bp.string().c_str());
      return ErrorCode::IO_ERROR;
    }
    s.root = 0;
    for (uint32_t i = 0; i < n; ++i)
      if (s.bvh[i].parent == i) {
        s.root = i;
        break;
      }
    return ErrorCode::OK;
  }
};

const std::vector<std::string> kCommands = {"load",         "genScene",   "export",   "camImport", "copyToGPU",
                                            "LBVH",         "bittner13GPU", "CUDATracer", "copyToHost", "validate",
                                            "status",       "writeData",  "BVHExport", "BVHImport"};

std::shared_ptr<Command> fromString(const std::string &s) {
  static const std::map<std::string, std::function<std::shared_ptr<Command>()>> reg = {
      {"load", [] { return std::make_shared<Load>(); }},
      {"genScene", [] { return std::make_shared<GenScene>(); }},
      {"export", [] { return std::make_shared<ExportMesh>(); }},
      {"camImport", [] { return std::make_shared<CamImport>(); }},
      {"copyToGPU", [] { return std::make_shared<CopyToGPU>(); }},
      {"LBVH", [] { return std::make_shared<LBVHCmd>(); }},
      {"bittner13GPU", [] { return std::make_shared<Bittner13GPUCmd>(); }},
      {"CUDATracer", [] { return std::make_shared<CUDATracerCmd>(); }},
      {"copyToHost", [] { return std::make_shared<CopyToHost>(); }},
      {"validate", [] { return std::make_shared<ValidateCmd>(); }},
      {"status", [] { return std::make_shared<StatusCmd>(); }},
      {"writeData", [] { return std::make_shared<WriteData>(); }},
      {"BVHExport", [] { return std::make_shared<BVHExportCmd>(); }},
      {"BVHImport", [] { return std::make_shared<BVHImportCmd>(); }},
  };
  auto it = reg.find(s);
  return it == reg.end() ? nullptr : it->second();
}

struct Config {  // lib/base/Config.cpp:26-66
  std::string name = "BVHTest", basePath = ".";
  bool verbose = true;
  std::vector<std::string> input;
  std::vector<std::shared_ptr<Command>> cmds;
  void fromJSON(const json &j) {
    if (j.count("baseConfig")) {
      const json &b = j["baseConfig"];
      name = b.value("name", name);
      verbose = b.value("verbose", verbose);
      basePath = b.value("basePath", basePath);
      input = b.value("input", input);
    }
    if (!j.count("commands") || !j["commands"].is_array()) return;
    for (const json &c : j["commands"]) {
      std::string n = c.at("cmd").get<std::string>();
      auto cmd = fromString(n);
      if (!cmd) {
        LOG_WARN("Command '%s' not found", n.c_str());
        continue;
      }
      cmd->fromJSON(c.count("options") ? c.at("options") : json::object());
      cmds.push_back(cmd);
    }
  }
  json toJSON() const {
    json j = {{"baseConfig", {{"name", name}, {"verbose", verbose}, {"basePath", basePath}, {"input", input}}},
              {"commands", json::array()}};
    for (auto &c : cmds) {
      json o = c->toJSON();
      if (o.empty())
        j["commands"].push_back(json{{"cmd", c->name()}});
      else
        j["commands"].push_back(json{{"cmd", c->name()}, {"options", o}});
    }
    return j;
  }
};

int usage() {
  std::cout << "Options:\n"
               "  help              -- display this help message\n"
               "  run CFG           -- run bvhtest with configuration file CFG\n"
               "  list              -- list all commands and exit\n"
               "  generate FILE CMDs -- generate config file with commands CMDs (use - for stdout)\n";
  return 0;
}

// run -- src/main.cpp:153-210
bool run(const std::string &file) {
  try {
    json j;
    {
      std::ifstream f(file);
      if (!f) {
        LOG_ERR("Cannot open %s", file.c_str());
        return false;
      }
      f >> j;
    }
    Config cfg;
    cfg.fromJSON(j);
    {
      std::ofstream f(file, std::ios::trunc);
      f << cfg.toJSON().dump(2);
    }
    gVerbose = cfg.verbose;
    uint32_t k = 0;
    for (auto &in : cfg.input) {
      LOG_INFO("Processing %s (%u of %zu)", in.c_str(), ++k, cfg.input.size());
      State s;
      s.input = in;
      s.basePath = cfg.basePath;
      s.name = cfg.name;
      for (auto &c : cfg.cmds) {
        ErrorCode r = c->run(s);
        if (r != ErrorCode::OK && r != ErrorCode::WARNING) break;
      }
      bvht_freeBVH(&s.dbvh);
      bvht_freeMesh(&s.rawMesh);
    }
  } catch (json::exception &e) {
    LOG_ERR("JSON exception: %s", e.what());
    return false;
  }
  return true;
}

}  // namespace

int main(int argc, char **argv) {
  std::vector<std::string> args(argv + 1, argv + argc);
  if (args.empty()) return usage();
  for (size_t i = 0; i < args.size(); ++i) {
    if (args[i] == "list") {
      std::cout << "Commands:\n";
      for (auto &n : kCommands) {
        auto c = fromString(n);
        std::printf("  - %-16s -- %s\n", n.c_str(), c->desc().c_str());
      }
      return 0;
    }
    if (args[i] == "generate") {
      if (i + 1 >= args.size()) {
        LOG_ERR("generate needs a file name");
        return 1;
      }
      Config cfg;
      for (size_t k = i + 2; k < args.size(); ++k) {
        auto c = fromString(args[k]);
        if (!c) {
          LOG_ERR("Command '%s' not found", args[k].c_str());
          return 2;
        }
        cfg.cmds.push_back(c);
      }
      std::string out = cfg.toJSON().dump(2);
      if (args[i + 1] == "-")
        std::cout << out << "\n";
      else
        std::ofstream(args[i + 1]) << out;
      return 0;
    }
    if (args[i] == "run") {
      if (++i >= args.size()) {
        LOG_ERR("Argument required for 'run'");
        return 1;
      }
      if (!run(args[i])) return 2;
      continue;
    }
    return usage();
  }
  return 0;
}
