// Mufflon's scene-file loader (src/mffloader) over the core C API of the HIP
// path tracer: the JSON scene description (doc/FileFormat.md, parsing in
// json_loader.cpp) and the binary geometry file (.mff, binary.cpp).
//
// The loader only talks to the core through include/mufflon_core.h, as the
// reference's plugin does: a scene loaded here is the world a front end
// would have built with the same world_* / scenario_* calls. LoDs are read on
// demand: the core calls loader_load_lod (registered with
// mufflon_set_lod_loader) when a scenario load needs one that is not loaded.
#include <zlib.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <map>
#include <memory>
#include <mutex>
#include <set>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "json_dom.hpp"
#include "mff_loader.h"
#include "mufflon_core.h"

namespace fs = std::filesystem;
using mffjson::Value;

namespace {

// ------------------------------------------------------------------ logging
thread_local std::string gError;
std::atomic<int> gLogLevel{LOG_WARNING};

void log_at(LogLevel level, const std::string& msg) {
    if ((int)level < gLogLevel.load()) return;
    static const char* kNames[] = {"PEDANTIC", "INFO", "WARNING", "ERROR", "FATAL"};
    std::fprintf(stderr, "[mffloader %s] %s\n", kNames[std::min<int>(level, 4)], msg.c_str());
}
void warn(const std::string& msg) { log_at(LOG_WARNING, msg); }
void set_error(const std::string& msg) {
    gError = msg;
    log_at(LOG_ERROR, msg);
}

struct LoaderError : std::runtime_error {
    using std::runtime_error::runtime_error;
};
struct Aborted {};

// ----------------------------------------------------------- file versions
// json_loader.hpp:30-82
struct FileVersion {
    unsigned major = 0, minor = 0;
    bool operator<(const FileVersion& o) const { return major < o.major || (major == o.major && minor < o.minor); }
    bool operator>=(const FileVersion& o) const { return !(*this < o); }
};
constexpr FileVersion kCurrentVersion{1, 7};
constexpr FileVersion kWorldToInstVersion{1, 4};      // instance matrices are world-to-instance
constexpr FileVersion kAbsoluteNearFarVersion{1, 5};  // camera planes default to absolute values
constexpr FileVersion kUniqueMatIndicesVersion{1, 7};  // objects list their unique material indices
constexpr float kDefaultNearFactor = 1.e-4f, kDefaultFarFactor = 2.f;
constexpr float kDefaultNear = 0.01f, kDefaultFar = 500.f;
constexpr float kPi = 3.14159265358979323846f;

// --------------------------------------------------------- binary format
// FileFormat.md "The binary file": little-endian sections tagged by four
// ASCII characters read as one u32.
constexpr uint32_t tag(char a, char b, char c, char d) {
    return (uint32_t)(uint8_t)a | ((uint32_t)(uint8_t)b << 8) | ((uint32_t)(uint8_t)c << 16) |
           ((uint32_t)(uint8_t)d << 24);
}
constexpr uint32_t kMatsTag = tag('M', 'a', 't', 's');
constexpr uint32_t kBoneTag = tag('B', 'o', 'n', 'e');
constexpr uint32_t kObjsTag = tag('O', 'b', 'j', 's');
constexpr uint32_t kObjTag = tag('O', 'b', 'j', '_');
constexpr uint32_t kInstTag = tag('I', 'n', 's', 't');
constexpr uint32_t kLodTag = tag('L', 'O', 'D', '_');
constexpr uint32_t kAttrTag = tag('A', 't', 't', 'r');
constexpr uint32_t kFlagDeflate = 1, kFlagCompressedNormals = 2;
constexpr uint32_t kNoFrame = 0xFFFFFFFFu;
// object record after its name: flags, keyframe, previous animation object, box
constexpr uint64_t kObjectInfoBytes = 9 * 4;

class BinaryFile {
public:
    explicit BinaryFile(const fs::path& path) : path_(path.string()) {
        f_ = std::fopen(path_.c_str(), "rb");
        if (!f_) throw LoaderError("Failed to open binary file '" + path_ + "'");
        std::fseek(f_, 0, SEEK_END);
        size_ = (uint64_t)std::ftell(f_);
        std::fseek(f_, 0, SEEK_SET);
    }
    ~BinaryFile() {
        if (f_) std::fclose(f_);
    }
    BinaryFile(const BinaryFile&) = delete;
    BinaryFile& operator=(const BinaryFile&) = delete;

    void seek(uint64_t pos) {
        if (pos > size_ || fseeko(f_, (off_t)pos, SEEK_SET) != 0)
            throw LoaderError("Seek past the end of binary file '" + path_ + "' (" + std::to_string(pos) + ")");
    }
    void skip(uint64_t bytes) { seek(tell() + bytes); }
    uint64_t tell() const { return (uint64_t)ftello(f_); }
    void bytes(void* out, uint64_t n) {
        if (n && std::fread(out, 1, n, f_) != n)
            throw LoaderError("Unexpected end of binary file '" + path_ + "'");
    }
    template <class T>
    T read() {
        T v;
        bytes(&v, sizeof(T));
        return v;
    }
    std::string string() {
        const uint32_t n = read<uint32_t>();
        if (n > size_ - tell()) throw LoaderError("String longer than binary file '" + path_ + "'");
        std::string s(n, '\0');
        bytes(&s[0], n);
        return s;
    }
    void expect(uint32_t magic, const std::string& what) {
        if (read<uint32_t>() != magic) throw LoaderError("Invalid " + what + " magic constant");
    }
    // DATA' block: u32 compressed bytes, u32 decompressed bytes, zlib stream
    // (binary.cpp:424-440)
    std::vector<uint8_t> inflate_block() {
        const uint32_t packed = read<uint32_t>();
        const uint32_t unpacked = read<uint32_t>();
        if (packed > size_ - tell()) throw LoaderError("Compressed block longer than binary file '" + path_ + "'");
        std::vector<uint8_t> in(packed), out(unpacked);
        bytes(in.data(), packed);
        uLongf got = unpacked;
        const int rc = uncompress(out.data(), &got, in.data(), packed);
        if (rc != Z_OK) throw LoaderError("Failed to deflate stream (error code " + std::to_string(rc) + ")");
        if (got != unpacked) throw LoaderError("Mismatch between expected and actual decompressed byte count");
        return out;
    }
    const std::string& path() const { return path_; }

private:
    std::string path_;
    FILE* f_ = nullptr;
    uint64_t size_ = 0;
};

// Octahedral normal decoding of COMPRESSED_NORMALS (FileFormat.md "Normal
// Compression"), then normalization.
Vec3 unpack_normal(uint32_t packed) {
    const float u = (float)(int16_t)(packed & 0xFFFFu) / 32767.0f;
    const float v = (float)(int16_t)(packed >> 16) / 32767.0f;
    const float z = 1.0f - std::fabs(u) - std::fabs(v);
    float x, y;
    if (z >= 0.0f) {
        x = u;
        y = v;
    } else {
        x = (1.0f - std::fabs(v)) * (u >= 0.0f ? 1.0f : -1.0f);
        y = (1.0f - std::fabs(u)) * (v >= 0.0f ? 1.0f : -1.0f);
    }
    const float len = std::sqrt(x * x + y * y + z * z);
    return Vec3{x / len, y / len, z / len};
}

struct ObjectRecord {
    std::string name;
    uint64_t offset = 0;
    ObjectHdl hdl = nullptr;
    float box[6] = {0, 0, 0, 0, 0, 0};
    uint32_t lodCount = 0;
};

// Everything a later loader_load_lod needs to find an object's LoDs.
struct FileState {
    fs::path binPath;
    FileVersion version;
    uint32_t flags = 0;
    std::vector<ObjectRecord> objects;
    std::map<const void*, uint32_t> objectIndex;  // object handle -> index in the file
};

struct LoaderInstance {
    MufflonInstanceHdl core = nullptr;
    std::mutex mutex;  // guards `file` and `status`
    FileState file;
    std::string status;
    std::atomic<bool> abort{false};
    std::atomic<bool> loading{false};
    std::string statusCopy;  // what loader_get_loading_status returned last

    void stage(const std::string& s) {
        std::lock_guard<std::mutex> lock(mutex);
        status = s;
    }
    void check_abort() const {
        if (abort.load()) throw Aborted{};
    }
};

// ---------------------------------------------------------- one LoD
// BinaryLoader::read_lod (binary.cpp:672-744). Returns the level loaded.
uint32_t read_lod(MufflonInstanceHdl core, const FileState& fs_, uint32_t objIndex, uint32_t lod, bool asReduced) {
    if (objIndex >= fs_.objects.size())
        throw LoaderError("Object index out of bounds (" + std::to_string(objIndex) + " >= " +
                          std::to_string(fs_.objects.size()) + ")");
    const ObjectRecord& obj = fs_.objects[objIndex];
    BinaryFile f(fs_.binPath);
    f.seek(obj.offset + 4);
    const uint32_t nameLen = f.read<uint32_t>();
    f.skip(nameLen + kObjectInfoBytes);
    const uint32_t lods = f.read<uint32_t>();
    if (lods == 0) throw LoaderError("Object '" + obj.name + "' has no LoDs");
    const uint32_t level = std::min(lod, lods - 1u);
    f.skip(8ull * level);
    f.seek(f.read<uint64_t>());
    if (f.read<uint32_t>() != kLodTag)
        throw LoaderError("Invalid LoD magic constant (object '" + obj.name + "', LoD " + std::to_string(level) + ")");
    const uint32_t numTris = f.read<uint32_t>(), numQuads = f.read<uint32_t>(), numSpheres = f.read<uint32_t>();
    const uint32_t numVerts = f.read<uint32_t>();
    (void)f.read<uint32_t>();  // edges
    const uint32_t vertAttribs = f.read<uint32_t>(), faceAttribs = f.read<uint32_t>();
    const uint32_t sphereAttribs = f.read<uint32_t>();

    const bool deflate = (fs_.flags & kFlagDeflate) != 0;
    const bool packedNormals = (fs_.flags & kFlagCompressedNormals) != 0;
    const size_t V = numVerts, T = numTris, Q = numQuads, S = numSpheres, F = T + Q;
    const size_t vertexBytes = V * (12 + (packedNormals ? 4 : 12) + 8);
    std::vector<uint8_t> vertexData, triData, quadData, matData, sphereData, sphereMatData;
    auto skip_attributes = [&](uint32_t count) {
        f.expect(kAttrTag, "attribute");
        if (deflate) {
            (void)f.inflate_block();
            return;
        }
        for (uint32_t i = 0; i < count; ++i) {  // binary.cpp:245-251
            (void)f.string();
            (void)f.string();
            (void)f.read<uint32_t>();
            (void)f.read<uint32_t>();
            f.skip(f.read<uint64_t>());
        }
    };
    auto block = [&](size_t bytes, std::vector<uint8_t>& out, const char* what) {
        if (deflate) {
            out = f.inflate_block();
            if (out.size() != bytes)
                throw LoaderError(std::string("Decompressed ") + what + " block of object '" + obj.name +
                                  "' has the wrong size");
        } else {
            out.resize(bytes);
            f.bytes(out.data(), bytes);
        }
    };
    // order of binary.cpp:711-733: vertices, vertex attributes, triangles,
    // quads, material indices (triangles, then quads), face attributes,
    // spheres (center and radius, then material indices: one stream when
    // deflated, binary.cpp:539-567), sphere attributes
    if (V > 0) block(vertexBytes, vertexData, "vertex");
    if (V > 0 && vertAttribs > 0) skip_attributes(vertAttribs);
    if (T > 0) block(T * 12, triData, "triangle");
    if (Q > 0) block(Q * 16, quadData, "quad");
    if (F > 0) block(F * 2, matData, "material index");
    if (F > 0 && faceAttribs > 0) skip_attributes(faceAttribs);
    if (S > 0) {
        if (deflate) {
            block(S * 18, sphereData, "sphere");
            sphereMatData.assign(sphereData.begin() + 16 * S, sphereData.end());
            sphereData.resize(16 * S);
        } else {
            block(S * 16, sphereData, "sphere");
            block(S * 2, sphereMatData, "sphere material index");
        }
        if (sphereAttribs > 0) skip_attributes(sphereAttribs);
    }

    const LodHdl lodHdl = object_add_lod(obj.hdl, level, asReduced);
    if (!lodHdl) throw LoaderError(std::string("Failed to add LoD: ") + core_get_dll_error());
    if (!polygon_reserve(lodHdl, V, T, Q)) throw LoaderError("Failed to reserve LoD polygon memory");
    if (!spheres_reserve(lodHdl, S)) throw LoaderError("Failed to reserve LoD sphere memory");
    if (V > 0) {
        const char* base = (const char*)vertexData.data();
        BulkLoader points{BULK_ARRAY, {}}, normals{BULK_ARRAY, {}}, uvs{BULK_ARRAY, {}};
        points.descriptor.bytes = base;
        normals.descriptor.bytes = base + 12 * V;
        uvs.descriptor.bytes = base + V * (12 + (packedNormals ? 4 : 12));
        AABB box{{obj.box[0], obj.box[1], obj.box[2]}, {obj.box[3], obj.box[4], obj.box[5]}};
        size_t pr = 0, nr = 0, ur = 0;
        if (polygon_add_vertex_bulk(lodHdl, V, &points, packedNormals ? nullptr : &normals, &uvs, &box, &pr,
                                    packedNormals ? nullptr : &nr, &ur) < 0)
            throw LoaderError("Failed to bulk-read vertices for object '" + obj.name + "'");
        if (pr != V || ur != V || (!packedNormals && nr != V))
            throw LoaderError("Not all vertices were fully read for object '" + obj.name + "'");
        if (packedNormals) {
            const uint32_t* packed = (const uint32_t*)(base + 12 * V);
            for (size_t v = 0; v < V; ++v)
                if (!polygon_set_vertex_normal(lodHdl, (VertexHdl)v, unpack_normal(packed[v])))
                    throw LoaderError("Failed to set normal for object '" + obj.name + "'");
        }
    }
    if (T > 0) {
        const UVec3* tris = (const UVec3*)triData.data();
        for (size_t t = 0; t < T; ++t)
            if (polygon_add_triangle(lodHdl, tris[t]) < 0)
                throw LoaderError("Failed to add triangle to object '" + obj.name + "': " + core_get_dll_error());
    }
    if (Q > 0) {
        const UVec4* quads = (const UVec4*)quadData.data();
        for (size_t q = 0; q < Q; ++q)
            if (polygon_add_quad(lodHdl, quads[q]) < 0)
                throw LoaderError("Failed to add quad to object '" + obj.name + "': " + core_get_dll_error());
    }
    if (F > 0) {
        BulkLoader mats{BULK_ARRAY, {}};
        mats.descriptor.bytes = (const char*)matData.data();
        if (polygon_set_material_idx_bulk(lodHdl, 0, F, &mats) != F)
            throw LoaderError("Failed to set face material for object '" + obj.name + "'");
    }
    if (S > 0) {
        BulkLoader spheres{BULK_ARRAY, {}}, mats{BULK_ARRAY, {}};
        spheres.descriptor.bytes = (const char*)sphereData.data();
        mats.descriptor.bytes = (const char*)sphereMatData.data();
        AABB box{{obj.box[0], obj.box[1], obj.box[2]}, {obj.box[3], obj.box[4], obj.box[5]}};
        size_t read = 0;
        if (spheres_add_sphere_bulk(lodHdl, S, &spheres, &box, &read) < 0 || read != S)
            throw LoaderError("Failed to add spheres to object '" + obj.name + "'");
        if (spheres_set_material_idx_bulk(lodHdl, 0, S, &mats) != S)
            throw LoaderError("Failed to set sphere materials for object '" + obj.name + "'");
    }
    return level;
}

// --------------------------------------------------------------- JSON
// A path of keys for messages, as ParserState names them (json_helper.cpp:36-78).
class Context {
public:
    void push(const std::string& k) { keys_.push_back(k); }
    void pop() { keys_.pop_back(); }
    std::string where() const {
        std::string s;
        for (const auto& k : keys_) s += "\"" + k + "\":";
        if (!s.empty()) s.pop_back();
        return s.empty() ? "\"Root\"" : s;
    }
    struct Scope {
        Context& c;
        Scope(Context& c, const std::string& k) : c(c) { c.push(k); }
        ~Scope() { c.pop(); }
    };

private:
    std::vector<std::string> keys_;
};

class JsonScene {
public:
    JsonScene(LoaderInstance& inst, const fs::path& path) : inst_(inst), core_(inst.core), path_(path) {}

    void load();

private:
    LoaderInstance& inst_;
    MufflonInstanceHdl core_;
    fs::path path_;
    Context ctx_;
    FileVersion version_ = kCurrentVersion;
    bool absoluteNearFar_ = true;
    float box_[6] = {1e30f, 1e30f, 1e30f, -1e30f, -1e30f, -1e30f};
    std::vector<std::string> binMaterials_;
    std::map<std::string, InstanceHdl> instanceByName_;
    std::map<std::string, LightHdl> lights_;
    std::map<std::string, MaterialHdl> materials_;
    std::map<std::string, std::string> unsupported_;  // camera / light / material name -> reason
    FileState file_;

    // ---- typed reads (json_helper.cpp:180-293)
    [[noreturn]] void type_error(const std::string& key, const char* expected) const {
        throw LoaderError("Key " + ctx_.where() + (ctx_.where() == "\"Root\"" ? ":" : ":") + "\"" + key +
                          "\" has invalid type (expected " + expected + ")");
    }
    const Value& get(const Value& parent, const char* key) const {
        const Value* v = parent.find(key);
        if (!v) throw LoaderError("Key " + ctx_.where() + " is missing key \"" + key + "\"");
        return *v;
    }
    float as_float(const Value& v, const char* key) const {
        if (!v.is_number()) type_error(key, "float");
        return (float)v.number;
    }
    uint32_t as_uint(const Value& v, const char* key) const {
        if (!v.is_uint()) type_error(key, "uint");
        return (uint32_t)v.number;
    }
    int32_t as_int(const Value& v, const char* key) const {
        if (!v.is_int()) type_error(key, "int");
        return (int32_t)v.number;
    }
    bool as_bool(const Value& v, const char* key) const {
        if (!v.is_bool()) type_error(key, "bool");
        return v.boolean;
    }
    const std::string& as_string(const Value& v, const char* key) const {
        if (!v.is_string()) type_error(key, "string");
        return v.str;
    }
    void as_floats(const Value& v, const char* key, float* out, size_t n) const {
        if (!v.is_array() || v.items.size() != n) type_error(key, n == 2 ? "Vec2 array" : "Vec3 array");
        for (size_t i = 0; i < n; ++i) out[i] = as_float(v.items[i], key);
    }
    Vec3 as_vec3(const Value& v, const char* key) const {
        float f[3];
        as_floats(v, key, f, 3);
        return Vec3{f[0], f[1], f[2]};
    }
    // a list of Vec3 ([[x,y,z], ...]); a single [x,y,z] counts as one entry
    // where the reference accepts it (read_opt_array_array, json_loader.cpp:38-52)
    std::vector<Vec3> vec3_list(const Value& v, const char* key, bool allowSingle) const {
        std::vector<Vec3> out;
        if (!v.is_array()) type_error(key, "array");
        if (allowSingle && (v.items.empty() || !v.items[0].is_array())) {
            out.push_back(as_vec3(v, key));
            return out;
        }
        for (const Value& e : v.items) out.push_back(as_vec3(e, key));
        return out;
    }
    std::vector<float> float_list(const Value& v, const char* key) const {
        std::vector<float> out;
        if (v.is_number()) {
            out.push_back(as_float(v, key));
            return out;
        }
        if (!v.is_array()) type_error(key, "float array");
        for (const Value& e : v.items) out.push_back(as_float(e, key));
        return out;
    }
    template <class T>
    static std::vector<T> broadcast(std::vector<T> v, size_t n) {
        if (v.size() == 1 && n != 1) v.assign(n, v.front());
        return v;
    }

    TextureHdl texture(const std::string& name, TextureSampling sampling) {
        fs::path p = fs::u8path(name);
        if (!p.is_absolute()) p = path_.parent_path() / p;
        if (!fs::exists(p)) throw LoaderError("Cannot find texture file '" + p.u8string() + "'");
        p = fs::canonical(p);
        TextureHdl t = world_add_texture(core_, p.u8string().c_str(), sampling, MIPMAP_NONE, nullptr, nullptr);
        if (!t) throw LoaderError("Failed to load texture '" + name + "': " + core_get_dll_error());
        return t;
    }
    TextureHdl value_texture(const float* v, uint32_t n) {
        TextureHdl t = world_add_texture_value(core_, v, n, SAMPLING_NEAREST);
        if (!t) throw LoaderError(std::string("Failed to add a texture value: ") + core_get_dll_error());
        return t;
    }

    void load_binary(uint32_t globalLod, bool deinstance, bool worldToInst, bool noDefaultInstances);
    void load_cameras(const Value& cameras);
    void load_lights(const Value& lights);
    void load_materials(const Value& materials);
    bool material(const Value& m, MaterialParams& p, std::string& unsupported);
    void load_scenarios(const Value& scenarios, const std::string& defaultScenario);
    Value resolve_scenario(const Value& scenarios, const std::string& name, int depth) const;
    static void replace_keys(const Value& from, Value& into);
};

// ------------------------------------------------------------- binary part
// BinaryLoader::load_file (binary.cpp:1126-1241) without the LoD data, which
// loader_load_lod reads on demand.
void JsonScene::load_binary(uint32_t globalLod, bool deinstance, bool worldToInst, bool noDefaultInstances) {
    (void)globalLod;
    inst_.stage("Loading binary file");
    BinaryFile f(file_.binPath);
    f.expect(kMatsTag, "materials header");
    const uint64_t next = f.read<uint64_t>();
    const uint32_t numMaterials = f.read<uint32_t>();
    for (uint32_t i = 0; i < numMaterials; ++i) binMaterials_.push_back(f.string());
    f.seek(next);
    uint32_t magic = f.read<uint32_t>();
    if (magic == kBoneTag) {
        const uint64_t objectStart = f.read<uint64_t>();
        warn("Bone animation data in '" + f.path() + "' is not used by the HIP path tracer");
        f.seek(objectStart);
        magic = f.read<uint32_t>();
    }
    if (magic != kObjsTag) throw LoaderError("Invalid objects header magic constant");
    const uint64_t instanceStart = f.read<uint64_t>();
    file_.flags = f.read<uint32_t>();
    std::vector<uint64_t> jump(f.read<uint32_t>());
    for (auto& j : jump) j = f.read<uint64_t>();
    {
        const uint64_t here = f.tell();
        f.seek(instanceStart + 4);
        const uint32_t instanceCount = f.read<uint32_t>();
        f.seek(here);
        world_reserve_objects_instances(core_, (uint32_t)jump.size(), (uint32_t)(instanceCount + jump.size()));
    }
    // objects (BinaryLoader::read_object, binary.cpp:746-767)
    for (size_t i = 0; i < jump.size(); ++i) {
        inst_.check_abort();
        if (i % std::max<size_t>(1, jump.size() / 200) == 0)
            inst_.stage("Reading object definition " + std::to_string(i) + " / " + std::to_string(jump.size()));
        f.seek(jump[i]);
        if (f.read<uint32_t>() != kObjTag)
            throw LoaderError("Invalid object magic constant (object " + std::to_string(i) + ")");
        ObjectRecord o;
        o.offset = jump[i];
        o.name = f.string();
        const uint32_t flags = f.read<uint32_t>();
        (void)f.read<uint32_t>();  // keyframe
        (void)f.read<uint32_t>();  // previous animation object
        f.bytes(o.box, sizeof(o.box));
        o.lodCount = f.read<uint32_t>();
        o.hdl = world_create_object(core_, o.name.c_str(), (ObjectFlags)flags);
        if (!o.hdl) throw LoaderError("Failed to create object '" + o.name + "': " + core_get_dll_error());
        object_allocate_lod_slots(o.hdl, o.lodCount);
        file_.objectIndex[o.hdl] = (uint32_t)file_.objects.size();
        file_.objects.push_back(std::move(o));
    }
    // instances (BinaryLoader::read_instances, binary.cpp:780-882)
    inst_.stage("Loading instances");
    f.seek(instanceStart);
    if (f.read<uint32_t>() != kInstTag) throw LoaderError("Invalid instance magic constant");
    const uint32_t numInstances = f.read<uint32_t>();
    std::vector<uint8_t> hasInstance(file_.objects.size(), 0);
    auto grow = [&](const float* b) {
        for (int k = 0; k < 3; ++k) {
            box_[k] = std::min(box_[k], b[k]);
            box_[3 + k] = std::max(box_[3 + k], b[3 + k]);
        }
    };
    for (uint32_t i = 0; i < numInstances; ++i) {
        inst_.check_abort();
        const std::string name = f.string();
        const uint32_t objId = f.read<uint32_t>();
        const uint32_t keyframe = f.read<uint32_t>();
        (void)f.read<uint32_t>();  // previous animation instance
        Mat3x4 m;
        f.bytes(&m, sizeof(m));
        if (objId >= file_.objects.size())
            throw LoaderError("Instance '" + name + "' references object " + std::to_string(objId) + " of " +
                              std::to_string(file_.objects.size()));
        const ObjectRecord& o = file_.objects[objId];
        InstanceHdl ih = world_create_instance(core_, o.hdl, keyframe);
        if (!ih) throw LoaderError("Failed to create instance for object ID " + std::to_string(objId));
        instanceByName_.emplace(name, ih);
        if (!instance_set_transformation_matrix(core_, ih, &m, worldToInst ? 1u : 0u))
            throw LoaderError("Failed to set transformation matrix for instance of object ID " +
                              std::to_string(objId));
        Mat3x4 w;
        if (!instance_get_transformation_matrix(core_, ih, &w))
            throw LoaderError("Failed to read back the instance-to-world matrix of object ID " +
                              std::to_string(objId));
        // the object box under the affine map: centre mapped, half extents
        // through |A| (the box of the eight transformed corners)
        const float* M = &w.v[0];
        float tb[6];
        for (int r = 0; r < 3; ++r) {
            float lo = M[4 * r + 3], hi = M[4 * r + 3];
            for (int c = 0; c < 3; ++c) {
                const float a = M[4 * r + c] * o.box[c], b = M[4 * r + c] * o.box[3 + c];
                lo += std::min(a, b);
                hi += std::max(a, b);
            }
            tb[r] = lo;
            tb[3 + r] = hi;
        }
        grow(tb);
        hasInstance[objId] = 1;
    }
    if (!noDefaultInstances) {
        inst_.stage("Creating default instances");
        for (size_t i = 0; i < file_.objects.size(); ++i) {
            if (hasInstance[i]) continue;
            InstanceHdl ih = world_create_instance(core_, file_.objects[i].hdl, kNoFrame);
            if (!ih) throw LoaderError("Failed to create instance for object ID " + std::to_string(i));
            instanceByName_.emplace(file_.objects[i].name + "###defaultInstance", ih);
            grow(file_.objects[i].box);
        }
    }
    if (deinstance)
        warn("\"deinstance\" is ignored: instances keep their transformation (same geometry up to rounding)");
}

// --------------------------------------------------------------- cameras
// JsonLoader::load_cameras (json_loader.cpp:453-537)
void JsonScene::load_cameras(const Value& cameras) {
    inst_.stage("Parsing cameras");
    Context::Scope s0(ctx_, "cameras");
    if (!cameras.is_object()) type_error("cameras", "object");
    const float diag = std::sqrt((box_[3] - box_[0]) * (box_[3] - box_[0]) + (box_[4] - box_[1]) * (box_[4] - box_[1]) +
                                 (box_[5] - box_[2]) * (box_[5] - box_[2]));
    for (const auto& [name, cam] : cameras.members) {
        inst_.check_abort();
        Context::Scope s1(ctx_, name);
        if (!cam.is_object()) type_error(name, "object");
        const Value* nearV = cam.find("near");
        const Value* farV = cam.find("far");
        const float nearPlane = nearV ? as_float(*nearV, "near")
                                      : (absoluteNearFar_ ? kDefaultNear : kDefaultNearFactor * diag);
        const float farPlane = farV ? as_float(*farV, "far") : (absoluteNearFar_ ? kDefaultFar : kDefaultFarFactor * diag);
        const std::string& type = as_string(get(cam, "type"), "type");
        std::vector<Vec3> path = vec3_list(get(cam, "path"), "path", false);
        std::vector<Vec3> dir = vec3_list(get(cam, "viewDir"), "viewDir", false);
        std::vector<Vec3> up;
        if (const Value* u = cam.find("up")) up = vec3_list(*u, "up", false);
        else up.push_back(Vec3{0, 1, 0});
        if (dir.size() == 1 && (path.size() != 1 || up.size() != 1)) dir.assign(path.size(), dir.front());
        if (up.size() == 1 && (path.size() != 1 || dir.size() != 1)) up.assign(path.size(), up.front());
        if (path.size() == 1 && (dir.size() != 1 || up.size() != 1)) path.assign(dir.size(), path.front());
        if (path.size() != dir.size()) throw LoaderError("Mismatched camera path size (view direction)");
        if (path.size() != up.size()) throw LoaderError("Mismatched camera path size (up direction)");
        if (type == "pinhole") {
            const Value* fovV = cam.find("fov");
            const float degrees = fovV ? as_float(*fovV, "fov") : 25.f;
            const float radians = degrees * kPi / 180.0f;  // Degrees -> Radians (util/degrad.hpp:25)
            if (!world_add_pinhole_camera(core_, name.c_str(), path.data(), dir.data(), up.data(), (uint32_t)path.size(),
                                          nearPlane, farPlane, radians))
                throw LoaderError(std::string("Failed to add pinhole camera: ") + core_get_dll_error());
        } else if (type == "focus" || type == "ortho") {
            warn("camera '" + name + "' of type " + type + " is not supported by the HIP path tracer");
            unsupported_["camera:" + name] = type + " camera";
        } else {
            warn("camera object '" + name + "' has unknown type '" + type + "'");
        }
    }
}

// ---------------------------------------------------------------- lights
// JsonLoader::load_lights (json_loader.cpp:539-780)
void JsonScene::load_lights(const Value& lights) {
    inst_.stage("Parsing lights");
    Context::Scope s0(ctx_, "lights");
    if (!lights.is_object()) type_error("lights", "object");
    for (const auto& [name, light] : lights.members) {
        inst_.check_abort();
        Context::Scope s1(ctx_, name);
        if (!light.is_object()) type_error(name, "object");
        const std::string& type = as_string(get(light, "type"), "type");
        if (type == "point") {
            std::vector<Vec3> pos;
            if (const Value* p = light.find("position")) pos = vec3_list(*p, "position", true);
            std::vector<float> scales{1.f};
            if (const Value* sc = light.find("scale")) scales = float_list(*sc, "scale");
            std::vector<Vec3> intensity;
            if (const Value* iv = light.find("intensity")) {
                intensity = vec3_list(*iv, "intensity", false);
            } else if (const Value* fv = light.find("flux")) {
                intensity = vec3_list(*fv, "flux", false);
                const float sphere = 4.0f * kPi;  // flux [W] -> intensity [W/sr]
                for (Vec3& v : intensity) v = Vec3{v.x / sphere, v.y / sphere, v.z / sphere};
            } else {
                (void)get(light, "temperature");
                throw LoaderError("point light '" + name +
                                  "': black-body temperatures are not supported by this loader (give an intensity)");
            }
            const size_t n = std::max(pos.size(), std::max(intensity.size(), scales.size()));
            intensity = broadcast(intensity, n);
            pos = broadcast(pos, n);
            scales = broadcast(scales, n);
            if (pos.size() != intensity.size()) throw LoaderError("Mismatched light path size (intensities)");
            if (pos.size() != scales.size()) throw LoaderError("Mismatched light path size (scales)");
            const LightHdl h = world_add_light(core_, name.c_str(), LIGHT_POINT, (uint32_t)pos.size());
            if (h.type != LIGHT_POINT) throw LoaderError(std::string("Failed to add point light: ") + core_get_dll_error());
            for (uint32_t i = 0; i < (uint32_t)pos.size(); ++i) {
                world_set_point_light_position(core_, h, pos[i], i);
                const Vec3 v{intensity[i].x * scales[i], intensity[i].y * scales[i], intensity[i].z * scales[i]};
                world_set_point_light_intensity(core_, h, v, i);
            }
            lights_[name] = h;
        } else if (type == "envmap") {
            Vec3 color{1.f, 1.f, 1.f};
            if (const Value* sc = light.find("scale")) {
                if (sc->is_array()) color = as_vec3(*sc, "scale");
                else {
                    const float f = as_float(*sc, "scale");
                    color = Vec3{f, f, f};
                }
            }
            if (const Value* map = light.find("map")) {
                TextureHdl t = texture(as_string(*map, "map"), SAMPLING_NEAREST);
                const LightHdl h = world_add_background_light(core_, name.c_str(), BACKGROUND_ENVMAP);
                if (h.type != LIGHT_ENVMAP)
                    throw LoaderError(std::string("Failed to add environment light: ") + core_get_dll_error());
                world_set_env_light_map(core_, h, t);
                world_set_env_light_scale(core_, h, color);
                lights_[name] = h;
            } else {
                const LightHdl h = world_add_background_light(core_, name.c_str(), BACKGROUND_MONOCHROME);
                if (h.type != LIGHT_ENVMAP)
                    throw LoaderError(std::string("Failed to add monochromatic environment light: ") +
                                      core_get_dll_error());
                world_set_env_light_color(core_, h, color);
                lights_[name] = h;
            }
        } else if (type == "spot" || type == "directional" || type == "sky" || type == "goniometric") {
            warn("light '" + name + "' of type " + type + " is not supported by the HIP path tracer");
            unsupported_["light:" + name] = type + " light";
        } else {
            warn("light object '" + name + "' has unknown type '" + type + "'");
        }
    }
}

// ------------------------------------------------------------- materials
// JsonLoader::load_material (json_loader.cpp:202-429). Returns false (with
// the reason) for a material the device cannot render.
bool JsonScene::material(const Value& m, MaterialParams& p, std::string& unsupported) {
    std::memset(&p, 0, sizeof(p));
    if (!m.is_object()) type_error("material", "object");
    if (const Value* outer = m.find("outerMedium")) {
        Context::Scope s(ctx_, "outerMedium");
        p.outerMedium.absorption = as_vec3(get(*outer, "absorption"), "absorption");
        const Value& ior = get(*outer, "ior");
        if (ior.is_array()) {
            float f[2];
            as_floats(ior, "ior", f, 2);
            p.outerMedium.refractionIndex = Vec2{f[0], f[1]};
        } else {
            p.outerMedium.refractionIndex = Vec2{as_float(ior, "ior"), 0.f};
        }
    } else {
        p.outerMedium.absorption = Vec3{0.f, 0.f, 0.f};
        p.outerMedium.refractionIndex = Vec2{1.f, 0.f};
    }
    const std::string& type = as_string(get(m, "type"), "type");
    auto ndf_of = [&](const Value& v) {
        const std::string& n = as_string(get(v, "ndf"), "ndf");
        if (n == "BS") return NDF_BECKMANN;
        if (n == "GGX") return NDF_GGX;
        if (n == "Cos") return NDF_COSINE;
        throw LoaderError("Unknown normal distribution function '" + n + "'");
    };
    auto shadowing_of = [&](const Value& v) {
        const Value* s = v.find("shadowingModel");
        const std::string n = s ? as_string(*s, "shadowingModel") : "vcavity";
        if (n == "smith") return SHADOWING_SMITH;
        if (n == "vcavity") return SHADOWING_VCAVITY;
        throw LoaderError("Unknown shadowing model '" + n + "'");
    };
    auto albedo_of = [&](const Value& v, TextureSampling sampling) {
        const Value& a = get(v, "albedo");
        if (a.is_array()) {
            const Vec3 c = as_vec3(a, "albedo");
            const float f[3] = {c.x, c.y, c.z};
            return value_texture(f, 3);
        }
        if (a.is_string()) return texture(a.str, sampling);
        throw LoaderError("Invalid type for albedo.");
    };
    auto roughness_of = [&](const Value& v) {
        const Value& r = get(v, "roughness");
        if (r.is_array()) {
            float f[2];
            as_floats(r, "roughness", f, 2);
            return value_texture(f, 2);
        }
        if (r.is_number()) {
            const float a = as_float(r, "roughness");
            return value_texture(&a, 1);
        }
        if (r.is_string()) return texture(r.str, SAMPLING_LINEAR);
        throw LoaderError("Invalid type for roughness.");
    };
    if (m.find("alpha")) {
        unsupported = "alpha-textured material";
        return false;
    }
    if (m.find("displacement")) {
        unsupported = "displacement-mapped material";
        return false;
    }
    if (type == "lambert") {
        p.innerType = MATERIAL_LAMBERT;
        p.inner.lambert.albedo = albedo_of(m, SAMPLING_LINEAR);
    } else if (type == "torrance") {
        p.innerType = MATERIAL_TORRANCE;
        p.inner.torrance.shadowingModel = shadowing_of(m);
        p.inner.torrance.ndf = ndf_of(m);
        p.inner.torrance.roughness = roughness_of(m);
        p.inner.torrance.albedo = albedo_of(m, SAMPLING_LINEAR);
    } else if (type == "microfacet" || type == "walter") {
        // the same parameters, loaded as a Walter pack (json_loader.cpp:287-292)
        p.innerType = type == "walter" ? MATERIAL_WALTER : MATERIAL_MICROFACET;
        p.inner.walter.shadowingModel = shadowing_of(m);
        p.inner.walter.ndf = ndf_of(m);
        p.inner.walter.absorption = as_vec3(get(m, "absorption"), "absorption");
        p.inner.walter.roughness = roughness_of(m);
        p.inner.walter.refractionIndex = as_float(get(m, "ior"), "ior");
    } else if (type == "emissive") {
        p.innerType = MATERIAL_EMISSIVE;
        p.inner.emissive.scale = Vec3{1.f, 1.f, 1.f};
        if (const Value* sc = m.find("scale")) p.inner.emissive.scale = as_vec3(*sc, "scale");
        if (const Value* r = m.find("radiance")) {
            if (r->is_array()) {
                const Vec3 c = as_vec3(*r, "radiance");
                const float f[3] = {c.x, c.y, c.z};
                p.inner.emissive.radiance = value_texture(f, 3);
            } else if (r->is_string()) {
                p.inner.emissive.radiance = texture(r->str, SAMPLING_LINEAR);
            } else {
                throw LoaderError("Invalid type for radiance");
            }
        } else {
            (void)get(m, "temperature");
            unsupported = "black-body emissive material";
            return false;
        }
    } else if (type == "orennayar" || type == "blend" || type == "fresnel" || type == "glass" ||
               type == "opaque") {
        unsupported = type + " material";
        return false;
    } else {
        throw LoaderError("Unknown material type '" + type + "'");
    }
    return true;
}

void JsonScene::load_materials(const Value& materials) {
    inst_.stage("Parsing materials");
    Context::Scope s0(ctx_, "materials");
    if (!materials.is_object()) type_error("materials", "object");
    for (const auto& [name, m] : materials.members) {
        inst_.check_abort();
        Context::Scope s1(ctx_, name);
        MaterialParams p;
        std::string why;
        if (!material(m, p, why)) {
            warn("material '" + name + "' (" + why + ") is not supported by the HIP path tracer");
            unsupported_["material:" + name] = why;
            continue;
        }
        MaterialHdl h = world_add_material(core_, name.c_str(), &p);
        if (!h) throw LoaderError("Failed to add material to world: " + std::string(core_get_dll_error()));
        materials_[name] = h;
    }
}

// ------------------------------------------------------------- scenarios
// JsonLoader::selective_replace_keys (json_loader.cpp:1012-1029): members of
// `from` override `into`; objects merge recursively.
void JsonScene::replace_keys(const Value& from, Value& into) {
    for (const auto& [k, v] : from.members) {
        Value* t = into.find(k.c_str());
        if (t) {
            if (v.is_object() && t->is_object()) replace_keys(v, *t);
            else *t = v;
        } else {
            into.members.emplace_back(k, v);
        }
    }
}

// JsonLoader::load_scenario (json_loader.cpp:990-1010): a scenario with its
// parentScenario chain applied.
Value JsonScene::resolve_scenario(const Value& scenarios, const std::string& name, int depth) const {
    if (depth < 1) throw LoaderError("Loop in scenario inheritance");
    const Value* sc = scenarios.find(name.c_str());
    if (!sc) throw LoaderError("Failed to find parent scenario: " + name);
    if (!sc->is_object()) type_error(name, "object");
    if (const Value* parent = sc->find("parentScenario")) {
        Value base = resolve_scenario(scenarios, as_string(*parent, "parentScenario"), depth - 1);
        replace_keys(*sc, base);
        return base;
    }
    return *sc;
}

// JsonLoader::load_scenarios (json_loader.cpp:808-988)
void JsonScene::load_scenarios(const Value& scenarios, const std::string& defaultScenario) {
    inst_.stage("Parsing scenarios");
    Context::Scope s0(ctx_, "scenarios");
    world_reserve_scenarios(core_, (uint32_t)scenarios.members.size());
    for (const auto& member : scenarios.members) {
        inst_.check_abort();
        const std::string& name = member.first;
        Context::Scope s1(ctx_, name);
        const Value sc = resolve_scenario(scenarios, name, (int)scenarios.members.size());
        const std::string& camera = as_string(get(sc, "camera"), "camera");
        const Value& res = get(sc, "resolution");
        if (!res.is_array() || res.items.size() != 2) type_error("resolution", "IVec2 array");
        const int32_t width = as_int(res.items[0], "resolution"), height = as_int(res.items[1], "resolution");
        const Value* lod = sc.find("lod");
        const uint32_t globalLod = lod ? as_uint(*lod, "lod") : 0u;

        // what the device cannot render keeps the scenario out
        std::string blocker;
        if (unsupported_.count("camera:" + camera)) blocker = "uses the " + unsupported_["camera:" + camera] + " '" + camera + "'";
        if (const Value* ls = sc.find("lights"))
            for (const Value& l : ls->items)
                if (l.is_string() && unsupported_.count("light:" + l.str))
                    blocker = "uses the " + unsupported_["light:" + l.str] + " '" + l.str + "'";
        const Value& assign = get(sc, "materialAssignments");
        if (!assign.is_object()) type_error("materialAssignments", "object");
        for (const std::string& bin : binMaterials_) {
            const Value* mv = assign.find(bin.c_str());
            if (mv && mv->is_string() && unsupported_.count("material:" + mv->str))
                blocker = "assigns the " + unsupported_["material:" + mv->str] + " '" + mv->str + "'";
        }
        if (!blocker.empty()) {
            if (name == defaultScenario)
                throw LoaderError("The default scenario '" + name + "' " + blocker +
                                  ", which the HIP path tracer does not support");
            warn("scenario '" + name + "' " + blocker + ", which the HIP path tracer does not support; skipped");
            continue;
        }

        CameraHdl cam = world_get_camera(core_, camera.c_str());
        if (!cam) throw LoaderError("Camera '" + camera + "' does not exist");
        ScenarioHdl h = world_create_scenario(core_, name.c_str());
        if (!h) throw LoaderError("Failed to create scenario");
        if (!scenario_set_camera(core_, h, cam)) throw LoaderError("Failed to set camera '" + camera + "'");
        if (!scenario_set_resolution(h, (uint32_t)width, (uint32_t)height))
            throw LoaderError("Failed to set resolution '" + std::to_string(width) + "x" + std::to_string(height) + "'");
        if (!scenario_set_global_lod_level(h, globalLod)) throw LoaderError("Failed to set LoD " + std::to_string(globalLod));
        if (const Value* ls = sc.find("lights")) {
            Context::Scope s2(ctx_, "lights");
            if (!ls->is_array()) type_error("lights", "array");
            for (const Value& l : ls->items) {
                const std::string& ln = as_string(l, "lights");
                auto it = lights_.find(ln);
                if (it == lights_.end()) {
                    warn("Unknown light source '" + ln + "' will be ignored");
                    continue;
                }
                if (!scenario_add_light(core_, h, it->second))
                    throw LoaderError("Failed to add light '" + ln + "': " + core_get_dll_error());
            }
        }
        if (const Value* op = sc.find("objectProperties")) {
            Context::Scope s2(ctx_, "objectProperties");
            if (!op->is_object()) type_error("objectProperties", "object");
            scenario_reserve_custom_object_properties(h, op->members.size());
            for (const auto& [oname, props] : op->members) {
                Context::Scope s3(ctx_, oname);
                if (!props.is_object()) type_error(oname, "object");
                ObjectHdl oh = world_get_object(core_, oname.c_str());
                if (!oh) throw LoaderError("Failed to find object '" + oname + "'");
                if (const Value* l = props.find("lod"))
                    if (!scenario_set_object_lod(h, oh, as_uint(*l, "lod")))
                        throw LoaderError("Failed to set LoD level of object '" + oname + "'");
                if (const Value* mk = props.find("mask"))
                    if (as_bool(*mk, "mask") && !scenario_mask_object(h, oh))
                        throw LoaderError("Failed to set mask for object '" + oname + "'");
                if (props.find("tessellation"))
                    warn("tessellation of object '" + oname + "' is ignored (the HIP path tracer renders the "
                         "stored triangles)");
            }
        }
        if (const Value* ip = sc.find("instanceProperties")) {
            Context::Scope s2(ctx_, "instanceProperties");
            if (!ip->is_object()) type_error("instanceProperties", "object");
            scenario_reserve_custom_instance_properties(h, ip->members.size());
            for (const auto& [iname, props] : ip->members) {
                Context::Scope s3(ctx_, iname);
                if (!props.is_object()) type_error(iname, "object");
                auto it = instanceByName_.find(iname);
                if (it == instanceByName_.end()) continue;
                if (const Value* l = props.find("lod"))
                    if (!scenario_set_instance_lod(h, it->second, as_uint(*l, "lod")))
                        throw LoaderError("Failed to set LoD level of instance '" + iname + "'");
                if (const Value* mk = props.find("mask"))
                    if (as_bool(*mk, "mask") && !scenario_mask_instance(h, it->second))
                        throw LoaderError("Failed to set mask of instance '" + iname + "'");
            }
        }
        {
            Context::Scope s2(ctx_, "materialAssignments");
            scenario_reserve_material_slots(h, binMaterials_.size());
            for (const std::string& bin : binMaterials_) {
                inst_.check_abort();
                const std::string& mname = as_string(get(assign, bin.c_str()), bin.c_str());
                const MatIdx slot = scenario_declare_material_slot(h, bin.c_str(), bin.size());
                if (slot == (MatIdx)0xFFFF) throw LoaderError("Failed to declare material slot");
                auto it = materials_.find(mname);
                if (it == materials_.end()) throw LoaderError("Unknown material name '" + mname + "' in association");
                if (!scenario_assign_material(h, slot, it->second))
                    throw LoaderError("Failed to associate material '" + mname + "'");
            }
        }
        const char* msg = "";
        if (!world_finalize_scenario(core_, h, &msg))
            throw LoaderError("Scenario '" + name + "' did not pass sanity check: " + std::string(msg ? msg : ""));
    }
}

// ---------------------------------------------------------------- driver
// JsonLoader::load_file (json_loader.cpp:1031-1185)
void JsonScene::load() {
    inst_.stage("Loading JSON");
    std::string text;
    {
        std::ifstream in(path_, std::ios::binary);
        if (!in) throw LoaderError("Failed to open scene file '" + path_.u8string() + "'");
        std::ostringstream ss;
        ss << in.rdbuf();
        text = ss.str();
    }
    Value doc;
    try {
        doc = mffjson::parse(text);
    } catch (const mffjson::ParseError& e) {
        throw LoaderError(e.what());
    }
    if (!doc.is_object()) throw LoaderError("Document root is not an object");
    if (const Value* v = doc.find("version")) {
        const std::string& s = as_string(*v, "version");
        unsigned a = 0, b = 0;
        std::sscanf(s.c_str(), "%u.%u", &a, &b);
        version_ = FileVersion{a, b};
        if (kCurrentVersion < version_)
            warn("Scene file: version mismatch (" + s + "(file) vs 1.7(current))");
    } else {
        warn("Scene file: no version specified (current one assumed)");
    }
    const bool worldToInst = version_ >= kWorldToInstVersion;
    absoluteNearFar_ = version_ >= kAbsoluteNearFarVersion;
    fs::path bin = fs::u8path(as_string(get(doc, "binary"), "binary"));
    if (bin.empty()) throw LoaderError("Scene file has an empty binary file path");
    if (bin.is_relative()) bin = path_.parent_path() / bin;
    if (!fs::exists(bin)) throw LoaderError("Binary file '" + bin.u8string() + "' does not exist");
    file_.binPath = fs::canonical(bin);
    file_.version = version_;
    float initTess = 0.f;
    if (const Value* t = doc.find("initTessellationLevel")) initTess = as_float(*t, "initTessellationLevel");
    if (initTess > 0.f) warn("\"initTessellationLevel\" is ignored (no tessellation on the HIP path tracer)");

    const Value& lights = get(doc, "lights");
    const Value& scenarios = get(doc, "scenarios");
    const Value& cameras = get(doc, "cameras");
    const Value& materials = get(doc, "materials");
    if (!scenarios.is_object()) type_error("scenarios", "object");
    if (scenarios.members.empty()) throw LoaderError("Scene file has no scenarios");
    std::string defaultScenario;
    if (const Value* d = doc.find("defaultScenario")) defaultScenario = as_string(*d, "defaultScenario");
    if (defaultScenario.empty()) defaultScenario = scenarios.members.front().first;
    uint32_t defaultLod = 0;
    {
        Context::Scope s(ctx_, "scenarios");
        const Value& def = get(scenarios, defaultScenario.c_str());
        if (const Value* l = def.find("lod")) defaultLod = as_uint(*l, "lod");
    }
    bool deinstance = false, noDefaultInstances = false;
    if (const Value* v = doc.find("deinstance")) deinstance = as_bool(*v, "deinstance");
    if (const Value* v = doc.find("noDefaultInstances")) noDefaultInstances = as_bool(*v, "noDefaultInstances");

    load_binary(defaultLod, deinstance, worldToInst, noDefaultInstances);
    {  // later LoD requests read from this file
        std::lock_guard<std::mutex> lock(inst_.mutex);
        inst_.file = file_;
    }
    load_cameras(cameras);
    load_lights(lights);
    load_materials(materials);
    inst_.stage("Checking world sanity");
    const char* msg = "";
    if (!world_finalize(core_, Vec3{box_[0], box_[1], box_[2]}, Vec3{box_[3], box_[4], box_[5]}, &msg))
        throw LoaderError("World did not pass sanity check: " + std::string(msg ? msg : ""));
    load_scenarios(scenarios, defaultScenario);
    ScenarioHdl def = world_find_scenario(core_, defaultScenario.c_str());
    if (!def) throw LoaderError("Cannot find the default scenario '" + defaultScenario + "'");
    inst_.stage("Loading initial scenario");
    gError.clear();  // a failing on-demand LoD read (loader_load_lod) leaves its reason here
    if (!world_load_scenario(core_, def)) {
        std::string why = core_get_dll_error();
        if (!gError.empty()) why += " (" + gError + ")";
        throw LoaderError("Cannot load the default scenario '" + defaultScenario + "': " + why);
    }
}

// C callbacks for the core (mufflon_set_lod_loader)
Boolean lod_callback(void* user, ObjectHdl obj, uint32_t lod, Boolean asReduced) {
    return loader_load_lod(user, obj, lod, asReduced);
}
Boolean mat_callback(void* user, uint32_t objId, uint16_t* buf, uint32_t* read) {
    return loader_load_object_material_indices(user, objId, buf, read);
}
Boolean meta_callback(void* user, LodMetadata* data, size_t* read) {
    return loader_load_lod_metadatas(user, data, read);
}

LoaderInstance* instance_of(MufflonLoaderInstanceHdl h) {
    if (!h) throw LoaderError("Invalid loader instance handle (nullptr)");
    return static_cast<LoaderInstance*>(h);
}

template <class R, class F>
R guarded(R fail, F&& f) {
    try {
        return f();
    } catch (const std::exception& e) {
        set_error(e.what());
    } catch (...) {
        set_error("unknown error");
    }
    return fail;
}

}  // namespace

extern "C" {

const char* loader_get_dll_error(void) { return gError.c_str(); }

Boolean loader_set_log_level(LogLevel level) {
    if ((int)level < (int)LOG_PEDANTIC || (int)level > (int)LOG_FATAL_ERROR) {
        set_error("Invalid log level");
        return 0;
    }
    gLogLevel.store((int)level);
    return 1;
}

MufflonLoaderInstanceHdl loader_initialize(MufflonInstanceHdl mffInstHdl) {
    return guarded((MufflonLoaderInstanceHdl) nullptr, [&]() -> MufflonLoaderInstanceHdl {
        if (!mffInstHdl) throw LoaderError("Invalid mufflon instance (nullptr)");
        auto* l = new LoaderInstance;
        l->core = mffInstHdl;
        return l;
    });
}

void loader_destroy(MufflonLoaderInstanceHdl hdl) { delete static_cast<LoaderInstance*>(hdl); }

LoaderStatus loader_load_json(MufflonLoaderInstanceHdl hdl, const char* path) {
    return guarded(LOADER_ERROR, [&]() -> LoaderStatus {
        LoaderInstance& l = *instance_of(hdl);
        if (!path) throw LoaderError("Invalid scene file path (nullptr)");
        const fs::path p = fs::u8path(path);
        if (!fs::exists(p)) throw LoaderError("File '" + p.u8string() + "' does not exist");
        if (fs::is_directory(p)) throw LoaderError("Path '" + p.u8string() + "' is a directory, not a file");
        if (p.extension() != ".json")
            warn("Scene file does not end with '.json'; attempting to parse it anyway");
        world_clear_all(l.core);
        {
            std::lock_guard<std::mutex> lock(l.mutex);
            l.file = FileState{};
        }
        if (!mufflon_set_lod_loader(l.core, lod_callback, mat_callback, meta_callback, hdl))
            throw LoaderError(std::string("Failed to register the LoD loader: ") + core_get_dll_error());
        l.abort.store(false);
        l.loading.store(true);
        struct Done {
            LoaderInstance& l;
            ~Done() { l.loading.store(false); }
        } done{l};
        try {
            JsonScene(l, p).load();
        } catch (const Aborted&) {
            l.stage("Aborted");
            return LOADER_ABORT;
        }
        l.stage("Done");
        return LOADER_SUCCESS;
    });
}

LoaderStatus loader_save_scene(MufflonLoaderInstanceHdl hdl, const char* path) {
    (void)hdl;
    (void)path;
    set_error("loader_save_scene is not provided by this loader");
    return LOADER_ERROR;
}

Boolean loader_load_lod(MufflonLoaderInstanceHdl hdl, ObjectHdl obj, uint32_t lod, Boolean asReduced) {
    return guarded((Boolean)0, [&]() -> Boolean {
        LoaderInstance& l = *instance_of(hdl);
        if (!obj) throw LoaderError("Invalid object handle");
        FileState f;
        {
            std::lock_guard<std::mutex> lock(l.mutex);
            f = l.file;
        }
        if (f.binPath.empty()) throw LoaderError("No scene file loaded");
        uint32_t index = 0;
        auto it = f.objectIndex.find(obj);
        if (it != f.objectIndex.end()) index = it->second;
        else if (!object_get_id(obj, &index)) throw LoaderError(core_get_dll_error());
        (void)read_lod(l.core, f, index, lod, asReduced != 0);
        return 1;
    });
}

Boolean loader_load_object_material_indices(MufflonLoaderInstanceHdl hdl, const uint32_t objId,
                                            uint16_t* indexBuffer, uint32_t* readIndices) {
    return guarded((Boolean)0, [&]() -> Boolean {
        LoaderInstance& l = *instance_of(hdl);
        if (!indexBuffer) throw LoaderError("Invalid material index buffer (nullptr)");
        FileState f;
        {
            std::lock_guard<std::mutex> lock(l.mutex);
            f = l.file;
        }
        if (f.binPath.empty()) throw LoaderError("No scene file loaded");
        if (f.version < kUniqueMatIndicesVersion) {
            if (readIndices) *readIndices = 0;
            return 1;
        }
        if (objId >= f.objects.size())
            throw LoaderError("Object index out of bounds (" + std::to_string(objId) + " >= " +
                              std::to_string(f.objects.size()) + ")");
        BinaryFile b(f.binPath);
        b.seek(f.objects[objId].offset + 4);
        b.skip(b.read<uint32_t>() + kObjectInfoBytes);
        b.skip(8ull * b.read<uint32_t>());
        const uint32_t n = b.read<uint32_t>();
        b.bytes(indexBuffer, 2ull * n);
        if (readIndices) *readIndices = n;
        return 1;
    });
}

Boolean loader_load_lod_metadatas(MufflonLoaderInstanceHdl hdl, LodMetadata* data, size_t* read) {
    return guarded((Boolean)0, [&]() -> Boolean {
        LoaderInstance& l = *instance_of(hdl);
        if (!data) throw LoaderError("Invalid LoD metadata buffer (nullptr)");
        FileState f;
        {
            std::lock_guard<std::mutex> lock(l.mutex);
            f = l.file;
        }
        if (f.binPath.empty()) throw LoaderError("No scene file loaded");
        BinaryFile b(f.binPath);
        size_t k = 0;
        for (const ObjectRecord& o : f.objects) {
            b.seek(o.offset + 4);
            b.skip(b.read<uint32_t>() + kObjectInfoBytes);
            std::vector<uint64_t> lods(b.read<uint32_t>());
            for (auto& x : lods) x = b.read<uint64_t>();
            for (uint64_t off : lods) {
                b.seek(off);
                b.expect(kLodTag, "LoD");
                LodMetadata& m = data[k++];
                m.triangles = b.read<uint32_t>();
                m.quads = b.read<uint32_t>();
                m.spheres = b.read<uint32_t>();
                m.vertices = b.read<uint32_t>();
                m.edges = b.read<uint32_t>();
            }
        }
        if (read) *read = k;
        return 1;
    });
}

Boolean loader_abort(MufflonLoaderInstanceHdl hdl) {
    return guarded((Boolean)0, [&]() -> Boolean {
        LoaderInstance& l = *instance_of(hdl);
        if (!l.loading.load()) throw LoaderError("No loader running that could be aborted");
        l.abort.store(true);
        return 1;
    });
}

const char* loader_get_loading_status(MufflonLoaderInstanceHdl hdl) {
    return guarded((const char*)nullptr, [&]() -> const char* {
        LoaderInstance& l = *instance_of(hdl);
        std::lock_guard<std::mutex> lock(l.mutex);
        l.statusCopy = l.status;
        return l.statusCopy.c_str();
    });
}

void loader_profiling_enable(void) {}
void loader_profiling_disable(void) {}

}  // extern "C"
