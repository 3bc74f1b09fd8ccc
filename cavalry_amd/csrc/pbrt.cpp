// PBRT scene front end on the host: lexer, parser, shape builders and PLY
// reader, producing the triangle soup + mesh table that cav_scene_create
// takes, the camera and the render settings (cav_pbrt_*,
// cav_scene_create_from_pbrt).
//
// Reference: SceneLoading/Lexing.cu (tokens, recognised keywords),
// SceneLoading/Parsing.cu (scene-wide options, subsections, Include,
// ReverseOrientation, AreaLightSource binding, named materials), the factories
// TriangleMesh::createFromParams / createFromPLY (Core/TriangleMesh.cu:52-183),
// ShapeMesh sphere / disk (ShapeMesh/Sphere.cu, Disk.cu), the material,
// light, camera, sampler and integrator factories (Materials/*.h,
// Lights/LightObject.h:89-125, Cameras/PerspectiveCamera.cu:22-26,
// Samplers/SamplerObject.h:157-168, Integrators/CreateIntegrator.cu:10-25) and
// the CLI overrides of main.cu:25-97.
//
// Differences, each reported through cav_render_setup::warnings and the
// message list: rough glass is rendered as specular glass; texture files are
// PNG or TGA (8-bit RGBA; the reference also reads EXR and, through its image
// loader, other formats); the "random" sampler renders with Halton. Lights
// follow the reference's list (Scene::lightsHost): area lights and
// "infinite" environment maps in directive order, and any other LightSource
// / AreaLightSource type as the black environment light the reference's
// factory returns for it (Lights/LightObject.h:124-126).
// Unlike the reference, which exits, every error returns CAV_ERR_PARSE with a
// message.
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <unordered_set>
#include <vector>

#include "../../include/cavalry.h"
#include "capi_internal.h"
#include "image_read.h"
#include "mat4.h"

using namespace cav::m4;

namespace {

struct ParseError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

// ------------------------------------------------------------------ lexing
enum class Tok { Keyword, String, Num, LBracket, RBracket };

struct Token {
    Tok type;
    std::string text;  // keyword or string contents
    float num = 0.f;
};

const std::unordered_set<std::string>& keywords() {
    static const std::unordered_set<std::string> k = {
        "ReverseOrientation", "Material", "MakeNamedMedium", "MediumInterface", "NamedMedium",
        "Include", "PixelFilter", "Transform", "NamedMaterial", "TransformBegin", "TransformEnd",
        "AreaLightSource", "MakeNamedMaterial", "Texture", "Scale", "Rotate", "Translate",
        "LightSource", "LookAt", "Camera", "Film", "WorldBegin", "WorldEnd", "AttributeBegin",
        "AttributeEnd", "Sampler", "Shape", "Integrator"};
    return k;
}

bool is_letter(char c) { return ('a' <= c && c <= 'z') || ('A' <= c && c <= 'Z'); }
bool is_digit(char c) { return '0' <= c && c <= '9'; }

std::string read_text(const std::string& path) {
    std::ifstream f(path, std::ios::binary);
    if (!f) throw ParseError("cannot open " + path);
    std::stringstream ss;
    ss << f.rdbuf();
    return ss.str();
}

// SceneLoading/Lexing.cu:8-101. A number is a run of digits, '-', '.', 'e'
// (also 'E' and '+' here, so exponents like 1e+05 lex); a leading '.' reads
// as "0.". Carriage returns count as white space.
std::vector<Token> lex(const std::string& path) {
    const std::string in = read_text(path);
    std::vector<Token> out;
    size_t pos = 0;
    while (pos < in.size()) {
        const char c = in[pos];
        if (c == ' ' || c == '\n' || c == '\t' || c == '\r') {
            ++pos;
        } else if (c == '#') {
            while (pos < in.size() && in[pos] != '\n') ++pos;
            ++pos;
        } else if (c == '[') {
            out.push_back({Tok::LBracket, "["});
            ++pos;
        } else if (c == ']') {
            out.push_back({Tok::RBracket, "]"});
            ++pos;
        } else if (is_letter(c)) {
            std::string w;
            while (pos < in.size() && is_letter(in[pos])) w += in[pos++];
            if (!keywords().count(w)) throw ParseError("Unrecognized Keyword: " + w + " in " + path);
            out.push_back({Tok::Keyword, w});
        } else if (c == '"') {
            ++pos;
            std::string s;
            while (pos < in.size() && in[pos] != '"') s += in[pos++];
            if (pos == in.size()) throw ParseError("Lexing Error: Didn't find ending quotation in " + path);
            ++pos;
            out.push_back({Tok::String, s});
        } else if (is_digit(c) || c == '-' || c == '.') {
            std::string s;
            while (pos < in.size() && (is_digit(in[pos]) || in[pos] == '-' || in[pos] == '.' || in[pos] == 'e' ||
                                       in[pos] == 'E' || in[pos] == '+'))
                s += in[pos++];
            if (s[0] == '.') s = "0" + s;
            Token t{Tok::Num, s};
            try {
                size_t used = 0;
                t.num = std::stof(s, &used);
                if (used != s.size()) throw std::invalid_argument(s);
            } catch (const std::exception&) {
                throw ParseError("Lexing Error: bad number '" + s + "' in " + path);
            }
            out.push_back(t);
        } else {
            throw ParseError(std::string("unrecognized character '") + c + "' in " + path);
        }
    }
    return out;
}

// ------------------------------------------------------------------ params
struct Params {
    std::unordered_map<std::string, std::string> strings;
    std::unordered_map<std::string, std::vector<std::string>> string_lists;
    std::unordered_map<std::string, float> nums;
    std::unordered_map<std::string, std::vector<float>> num_lists;

    // Core/Parameters.h:16-70: a one-element list also answers the scalar
    // queries, with its first element
    bool has_num(const std::string& k) const { return nums.count(k) || num_lists.count(k); }
    float num(const std::string& k) const {
        if (nums.count(k)) return nums.at(k);
        if (num_lists.count(k) && !num_lists.at(k).empty()) return num_lists.at(k)[0];
        throw ParseError("Params Num field not found: " + k);
    }
    bool has_string(const std::string& k) const { return strings.count(k) || string_lists.count(k); }
    std::string str(const std::string& k) const {
        if (strings.count(k)) return strings.at(k);
        if (string_lists.count(k) && !string_lists.at(k).empty()) return string_lists.at(k)[0];
        throw ParseError("Params String field not found: " + k);
    }
    bool has_list(const std::string& k) const { return num_lists.count(k) > 0; }
    const std::vector<float>& list(const std::string& k) const {
        if (!num_lists.count(k)) throw ParseError("Params NumList field not found: " + k);
        return num_lists.at(k);
    }
};

struct ObjectDef {
    std::string keyword, name;
    Params params;
    bool defined = false;
};

struct Buf {
    std::vector<Token> toks;
    size_t i = 0;
    const Token& peek(size_t off = 0) const {
        if (i + off >= toks.size()) throw ParseError("unexpected end of scene file");
        return toks[i + off];
    }
    const Token& pop(Tok t, const char* what) {
        const Token& k = peek();
        if (k.type != t)
            throw ParseError(std::string("Parsing Error: ") + what + " expected at token " + std::to_string(i) +
                             ": " + k.text);
        ++i;
        return k;
    }
};

std::vector<std::string> split_spaces(const std::string& s) {
    std::vector<std::string> r;
    size_t a = 0, b;
    while ((b = s.find(' ', a)) != std::string::npos) {
        r.push_back(s.substr(a, b - a));
        a = b + 1;
    }
    r.push_back(s.substr(a));
    return r;
}

std::vector<float> read_num_list(Buf& b) {
    std::vector<float> r;
    b.pop(Tok::LBracket, "[");
    while (b.peek().type == Tok::Num) r.push_back(b.pop(Tok::Num, "number").num);
    b.pop(Tok::RBracket, "]");
    return r;
}

std::vector<std::string> read_string_list(Buf& b) {
    std::vector<std::string> r;
    b.pop(Tok::LBracket, "[");
    while (b.peek().type == Tok::String) r.push_back(b.pop(Tok::String, "string").text);
    b.pop(Tok::RBracket, "]");
    return r;
}

// Parsing.cu:66-110: Keyword "name" then "type field" value pairs; the field
// is the last word of its declaration, the object name the first word.
ObjectDef read_object(Buf& b) {
    ObjectDef d;
    d.keyword = b.pop(Tok::Keyword, "keyword").text;
    d.name = split_spaces(b.pop(Tok::String, "object name").text)[0];
    while (b.i < b.toks.size() && b.peek().type == Tok::String) {
        const std::vector<std::string> w = split_spaces(b.pop(Tok::String, "field").text);
        const std::string field = w.back();
        const Token& v = b.peek();
        if (v.type == Tok::String) {
            d.params.strings[field] = b.pop(Tok::String, "string").text;
        } else if (v.type == Tok::Num) {
            d.params.nums[field] = b.pop(Tok::Num, "number").num;
        } else if (v.type == Tok::LBracket) {
            const Token& n = b.peek(1);
            if (n.type == Tok::Num || n.type == Tok::RBracket) d.params.num_lists[field] = read_num_list(b);
            else if (n.type == Tok::String) d.params.string_lists[field] = read_string_list(b);
            else throw ParseError("Unaccepted List Element at token " + std::to_string(b.i + 1));
        } else {
            throw ParseError("Read Object failed at token " + std::to_string(b.i) + ": " + v.text);
        }
    }
    d.defined = true;
    return d;
}

// skips an unsupported directive and its arguments
void skip_object(Buf& b) {
    b.pop(Tok::Keyword, "keyword");
    while (b.i < b.toks.size() && b.peek().type != Tok::Keyword) ++b.i;
}

// Parsing.cu:119-169. Returns false when the next keyword is not a transform.
bool read_transform(Buf& b, M4& t) {
    const Token& k = b.peek();
    if (k.type != Tok::Keyword) throw ParseError("Parsing Error: Keyword expected at token " + std::to_string(b.i));
    auto num = [&]() { return b.pop(Tok::Num, "number").num; };
    if (k.text == "Translate") {
        ++b.i;
        const float x = num(), y = num(), z = num();
        t = translate(t, {x, y, z});
        return true;
    }
    if (k.text == "Rotate") {
        ++b.i;
        const float a = num(), x = num(), y = num(), z = num();
        t = rotate(t, radians(a), {x, y, z});
        return true;
    }
    if (k.text == "Scale") {
        ++b.i;
        const float x = num(), y = num(), z = num();
        t = scale(t, {x, y, z});
        return true;
    }
    if (k.text == "Transform") {
        ++b.i;
        const std::vector<float> v = read_num_list(b);
        if (v.size() != 16) throw ParseError("Transform needs 16 numbers");
        M4 m;
        for (int j = 0; j < 16; ++j) m.m[j] = v[j];
        t = mul(m, t);  // pre-multiplied, as the reference does
        return true;
    }
    return false;
}

// ------------------------------------------------------------------ scene
struct Mesh {
    std::vector<V3> pos;
    std::vector<V3> nrm;     // per vertex, empty when the mesh has none
    std::vector<float> uv;   // 2 per vertex, empty when the mesh has none
    std::vector<int> idx;    // 3 per triangle
    float sphere[4] = {0, 0, 0, 0};  // centre, radius of a sphere shape (radius 0: none)
};

// normalize(normalMatrix * n) for each given normal
std::vector<V3> transform_normals(const std::vector<V3>& n, const M4& t) {
    float r[3][3];
    if (!normal_matrix(t, r)) throw ParseError("singular transform for normals");
    std::vector<V3> out;
    out.reserve(n.size());
    for (const V3& v : n) out.push_back(normalize(mul3(r, v)));
    return out;
}

struct Material {
    cav_mesh m{};  // material fields only
};

struct Builder {
    std::vector<float> soup;
    std::vector<float> normals;  // per triangle, 9 floats (zeros for meshes without normals)
    std::vector<uint8_t> mesh_normals;
    bool any_normals = false;
    std::vector<float> uv;       // per triangle, 6 floats (zeros for meshes without them)
    bool any_uv = false;
    std::vector<cav::Image> images;                    // textures, in definition order
    std::unordered_map<std::string, int> texture_ids;  // Texture name -> index into images
    std::vector<cav_mesh> meshes;
    std::vector<std::string> messages;
    int warnings = 0;
    std::unordered_map<std::string, Material> named;
    // pending AreaLightSource (binds to the next shape, across subsections)
    bool next_is_light = false;
    float light_L[3] = {1, 1, 1};
    int pending_slot = -1;               // its place in `order`, filled with the mesh id
    std::vector<cav_env_light> envs;     // environment-type lights
    std::vector<int32_t> order;          // the light list: mesh id, or -1 - k for envs[k]
    int env_background = -1;             // the last "infinite" light (Parsing.cu:339-341)

    void warn(const std::string& s) {
        ++warnings;
        messages.push_back(s);
    }
};

Material default_matte() {  // Parsing.cu:290-292: MatteMaterial(1, 1, 1)
    Material r;
    r.m.material = CAV_MAT_MATTE;
    r.m.texture = -1;
    for (int c = 0; c < 3; ++c) r.m.kd[c] = 1.f;
    r.m.ior = 1.5f;
    return r;
}

// GGX::roughnessToAlpha (BSDFs/MicrofacetDistribution.h:57-62)
float roughness_to_alpha(float r) {
    r = std::fmax(r, 1e-3f);
    const float x = std::log(r);
    return 1.62142f + 0.819955f * x + 0.1734f * x * x + 0.0171201f * x * x * x + 0.000640711f * x * x * x * x;
}

void rgb_param(const Params& p, const char* k, float* out, Builder& bd, const std::string& what) {
    if (p.has_list(k)) {
        const auto& v = p.list(k);
        if (v.size() < 3) throw ParseError(what + ": " + k + " needs 3 values");
        for (int c = 0; c < 3; ++c) out[c] = v[c];
    } else if (p.strings.count(k)) {
        bd.warn(what + ": texture '" + p.strings.at(k) + "' for " + k + " is not supported; using the constant");
    }
}

// Kd of matte and plastic: a colour, or the name of a texture defined
// before (MatteMaterial.h:41-52, Plastic.h:77-84)
void kd_param(const Params& p, cav_mesh& m, Builder& bd, const std::string& what) {
    if (p.has_list("Kd")) {
        const auto& v = p.list("Kd");
        if (v.size() < 3) throw ParseError(what + ": Kd needs 3 values");
        for (int c = 0; c < 3; ++c) m.kd[c] = v[c];
    } else if (p.strings.count("Kd")) {
        const std::string& name = p.strings.at("Kd");
        auto it = bd.texture_ids.find(name);
        if (it == bd.texture_ids.end()) throw ParseError(what + ": texture not found: " + name);
        m.texture = it->second;
    }
}

// u/v roughness of metal and substrate with GGX::createFromRoughness
// (BSDFs/MicrofacetDistribution.h:147-154): 0.1 each unless given,
// "roughness" sets both, remapped unless "remaproughness" is "false"
void ggx_roughness(const Params& p, cav_mesh& m) {
    float u = 0.1f, v = 0.1f;
    bool remap = true;
    if (p.has_num("uroughness")) u = p.num("uroughness");
    if (p.has_num("vroughness")) v = p.num("vroughness");
    if (p.has_num("roughness")) u = v = p.num("roughness");
    if (p.has_string("remaproughness")) remap = p.str("remaproughness") == "true";
    m.roughness = remap ? roughness_to_alpha(u) : u;
    m.roughness_v = remap ? roughness_to_alpha(v) : v;
}

// MaterialObject::createFromObjectDefinition (Materials/MaterialObject.h:82-104)
Material make_material(const Params& p, Builder& bd) {
    const std::string type = p.str("type");
    Material r = default_matte();
    cav_mesh& m = r.m;
    if (type == "matte") {  // MatteMaterial.h:41-52
        kd_param(p, m, bd, "matte");
    } else if (type == "plastic") {  // Plastic.h:58-85
        m.material = CAV_MAT_PLASTIC;
        float rough = 0.1f;
        bool remap = true;
        if (p.has_num("roughness")) rough = p.num("roughness");
        if (p.has_string("remaproughness")) remap = p.str("remaproughness") == "true";
        for (int c = 0; c < 3; ++c) m.kd[c] = m.ks[c] = 0.25f;
        kd_param(p, m, bd, "plastic");
        rgb_param(p, "Ks", m.ks, bd, "plastic");
        m.roughness = remap ? roughness_to_alpha(rough) : rough;
    } else if (type == "mirror") {  // Mirror.h:32-38
        m.material = CAV_MAT_MIRROR;
        const float kr = p.has_num("Kr") ? p.num("Kr") : 0.9f;
        for (int c = 0; c < 3; ++c) m.kr[c] = kr;
    } else if (type == "glass") {  // Glass.h:44-82
        m.material = CAV_MAT_GLASS;
        for (int c = 0; c < 3; ++c) m.kr[c] = m.kt[c] = 1.f;
        rgb_param(p, "Kr", m.kr, bd, "glass");
        rgb_param(p, "Kt", m.kt, bd, "glass");
        m.ior = 1.5f;
        if (p.has_num("eta")) m.ior = p.num("eta");
        else if (p.has_num("index")) m.ior = p.num("index");
        float ur = 0.f, vr = 0.f;
        if (p.has_num("uroughness")) ur = p.num("uroughness");
        if (p.has_num("vroughness")) vr = p.num("vroughness");
        if (p.has_num("roughness")) ur = vr = p.num("roughness");
        if (ur != 0.f || vr != 0.f) bd.warn("glass: rough glass is rendered as specular glass");
    } else if (type == "metal") {  // Metal.h:36-64
        m.material = CAV_MAT_METAL;
        float f0[3] = {0.955f, 0.638f, 0.538f};  // copper
        if (p.has_list("eta") && p.has_list("k")) {
            const auto& eta = p.list("eta");
            const auto& k = p.list("k");
            if (eta.size() < 3 || k.size() < 3) throw ParseError("metal: eta and k need 3 values");
            for (int c = 0; c < 3; ++c) {  // Fresnel::createFromIOR(eta, k): |((n - 1) / (n + 1))^2|
                const std::complex<float> n(eta[c], k[c]), one(1.f, 0.f);
                const std::complex<float> r = (n - one) / (n + one);
                f0[c] = std::abs(r * r);
            }
        }
        for (int c = 0; c < 3; ++c) m.ks[c] = f0[c];
        ggx_roughness(p, m);
    } else if (type == "substrate") {  // Substrate.h:50-83
        m.material = CAV_MAT_SUBSTRATE;
        for (int c = 0; c < 3; ++c) m.kd[c] = m.ks[c] = 0.5f;
        kd_param(p, m, bd, "substrate");
        rgb_param(p, "Ks", m.ks, bd, "substrate");
        ggx_roughness(p, m);
    } else {
        bd.warn("unsupported material type: " + type + "; using matte");
        rgb_param(p, "Kd", m.kd, bd, type);
    }
    return r;
}

// ---------------------------------------------------------------- PLY
// Minimal reader for the element / property layouts of pbrt scenes: ASCII or
// binary little-endian, scalar properties of any PLY type, one list property
// on faces. TriangleMesh::createFromPLY keeps the first three indices of
// every face.
struct PlyProp {
    std::string name, type, count_type;  // count_type empty unless a list
};
struct PlyElem {
    std::string name;
    int64_t count = 0;
    std::vector<PlyProp> props;
};

int ply_size(const std::string& t) {
    if (t == "char" || t == "uchar" || t == "int8" || t == "uint8") return 1;
    if (t == "short" || t == "ushort" || t == "int16" || t == "uint16") return 2;
    if (t == "int" || t == "uint" || t == "float" || t == "int32" || t == "uint32" || t == "float32") return 4;
    if (t == "double" || t == "float64") return 8;
    throw ParseError("PLY: unknown type " + t);
}

double ply_bin(const std::string& t, const unsigned char* p) {
    if (t == "char" || t == "int8") return (double)*(const int8_t*)p;
    if (t == "uchar" || t == "uint8") return (double)*p;
    int16_t s; uint16_t us; int32_t i; uint32_t u; float f; double d;
    if (t == "short" || t == "int16") { std::memcpy(&s, p, 2); return s; }
    if (t == "ushort" || t == "uint16") { std::memcpy(&us, p, 2); return us; }
    if (t == "int" || t == "int32") { std::memcpy(&i, p, 4); return i; }
    if (t == "uint" || t == "uint32") { std::memcpy(&u, p, 4); return u; }
    if (t == "float" || t == "float32") { std::memcpy(&f, p, 4); return f; }
    std::memcpy(&d, p, 8);
    return d;
}

Mesh read_ply(const std::string& path, const M4& t) {
    const std::string data = read_text(path);
    size_t pos = 0;
    auto line = [&]() {
        const size_t e = data.find('\n', pos);
        if (e == std::string::npos) throw ParseError("PLY: truncated header in " + path);
        std::string l = data.substr(pos, e - pos);
        if (!l.empty() && l.back() == '\r') l.pop_back();
        pos = e + 1;
        return l;
    };
    if (line() != "ply") throw ParseError("PLY: bad magic in " + path);
    std::string format;
    std::vector<PlyElem> elems;
    for (;;) {
        std::istringstream ls(line());
        std::string w;
        ls >> w;
        if (w == "format") ls >> format;
        else if (w == "element") {
            PlyElem e;
            ls >> e.name >> e.count;
            elems.push_back(e);
        } else if (w == "property") {
            if (elems.empty()) throw ParseError("PLY: property before element in " + path);
            PlyProp p;
            std::string ty;
            ls >> ty;
            if (ty == "list") ls >> p.count_type >> p.type >> p.name;
            else { p.type = ty; ls >> p.name; }
            elems.back().props.push_back(p);
        } else if (w == "end_header") break;
    }
    const bool ascii = format == "ascii";
    if (!ascii && format != "binary_little_endian") throw ParseError("PLY: unsupported format " + format);
    Mesh m;
    std::istringstream as(ascii ? data.substr(pos) : std::string());
    const unsigned char* bp = (const unsigned char*)data.data() + pos;
    const unsigned char* be = (const unsigned char*)data.data() + data.size();
    auto next = [&](const std::string& ty) -> double {
        if (ascii) {
            double v;
            if (!(as >> v)) throw ParseError("PLY: truncated body in " + path);
            return v;
        }
        const int n = ply_size(ty);
        if (bp + n > be) throw ParseError("PLY: truncated body in " + path);
        const double v = ply_bin(ty, bp);
        bp += n;
        return v;
    };
    for (const PlyElem& e : elems) {
        int ix = -1, iy = -1, iz = -1, ivi = -1, inx = -1, iny = -1, inz = -1, iu = -1, iv = -1;
        for (size_t k = 0; k < e.props.size(); ++k) {
            if (e.props[k].name == "x") ix = (int)k;
            if (e.props[k].name == "y") iy = (int)k;
            if (e.props[k].name == "z") iz = (int)k;
            if (e.props[k].name == "nx") inx = (int)k;
            if (e.props[k].name == "ny") iny = (int)k;
            if (e.props[k].name == "nz") inz = (int)k;
            if (e.props[k].name == "u") iu = (int)k;
            if (e.props[k].name == "v") iv = (int)k;
            if (e.props[k].name == "vertex_indices" || e.props[k].name == "vertex_index") ivi = (int)k;
        }
        std::vector<double> vals(e.props.size());
        for (int64_t r = 0; r < e.count; ++r) {
            for (size_t k = 0; k < e.props.size(); ++k) {
                const PlyProp& p = e.props[k];
                if (p.count_type.empty()) {
                    vals[k] = next(p.type);
                } else {
                    const int64_t n = (int64_t)next(p.count_type);
                    std::vector<int> li((size_t)std::max<int64_t>(n, 0));
                    for (int64_t q = 0; q < n; ++q) li[(size_t)q] = (int)next(p.type);
                    if (e.name == "face" && (int)k == ivi) {
                        if (n < 3) throw ParseError("PLY: face with fewer than 3 vertices in " + path);
                        m.idx.insert(m.idx.end(), li.begin(), li.begin() + 3);
                    }
                }
            }
            if (e.name == "vertex") {
                if (ix < 0 || iy < 0 || iz < 0) throw ParseError("PLY: vertex without x, y, z in " + path);
                m.pos.push_back(apply(t, {(float)vals[ix], (float)vals[iy], (float)vals[iz]}));
                if (inx >= 0 && iny >= 0 && inz >= 0)
                    m.nrm.push_back({(float)vals[inx], (float)vals[iny], (float)vals[inz]});
                if (iu >= 0 && iv >= 0) m.uv.insert(m.uv.end(), {(float)vals[iu], (float)vals[iv]});
            }
        }
    }
    if (!m.nrm.empty()) m.nrm = transform_normals(m.nrm, t);
    return m;
}

// ShapeMesh/Sphere.cu: octahedron, four midpoint subdivisions, scaled by the
// radius then transformed
Mesh sphere_mesh(float radius, const M4& t) {
    std::vector<V3> p = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};
    std::vector<int> cur = {0, 2, 4, 0, 2, 5, 0, 3, 4, 0, 3, 5, 1, 2, 4, 1, 2, 5, 1, 3, 4, 1, 3, 5};
    auto mid = [&](int a, int b) {
        const V3 s{(p[a].x + p[b].x) * 0.5f, (p[a].y + p[b].y) * 0.5f, (p[a].z + p[b].z) * 0.5f};
        p.push_back(normalize(s));
        return (int)p.size() - 1;
    };
    for (int it = 0; it < 4; ++it) {
        std::vector<int> nxt;
        for (size_t k = 0; k < cur.size(); k += 3) {
            const int i0 = cur[k], i1 = cur[k + 1], i2 = cur[k + 2];
            const int i01 = mid(i0, i1), i02 = mid(i0, i2), i12 = mid(i1, i2);
            nxt.insert(nxt.end(), {i01, i02, i12, i0, i01, i02, i1, i01, i12, i2, i02, i12});
        }
        cur.swap(nxt);
    }
    Mesh m;
    for (const V3& v : p) m.pos.push_back(apply(t, {v.x * radius, v.y * radius, v.z * radius}));
    m.nrm = transform_normals(p, t);  // the unit-sphere points are their own normals
    m.idx = cur;
    // shapeParams (ShapeMesh/Sphere.cu:86-88): the transformed centre, and the
    // distance from it to the transformed point (radius, 0, 0)
    const V3 c = apply(t, {0, 0, 0}), e = apply(t, {radius, 0, 0});
    const V3 dc{c.x - e.x, c.y - e.y, c.z - e.z};
    m.sphere[0] = c.x; m.sphere[1] = c.y; m.sphere[2] = c.z;
    m.sphere[3] = std::sqrt(dc.x * dc.x + dc.y * dc.y + dc.z * dc.z);
    return m;
}

// ShapeMesh/Disk.cu: 360 segments around the centre at the given height
Mesh disk_mesh(float radius, float height, const M4& t) {
    const int seg = 360;
    Mesh m;
    for (int s = 0; s < seg; ++s) {
        const float th = (float)(2.f * M_PI * (double)((float)s / (float)seg));
        m.pos.push_back(apply(t, {std::sin(th) * radius, std::cos(th) * radius, height}));
    }
    m.pos.push_back(apply(t, {0, 0, height}));
    for (int s = 0; s < seg; ++s) m.idx.insert(m.idx.end(), {(s + 1) % seg, s, seg});
    return m;
}

std::string join(const std::string& base, const std::string& rel) {
    if (!rel.empty() && rel[0] == '/') return rel;
    if (base.empty()) return rel;
    return base + "/" + rel;
}

std::string dir_of(const std::string& path) {
    const size_t s = path.find_last_of('/');
    return s == std::string::npos ? std::string() : path.substr(0, s);
}

// TriangleMesh::createFromObjectDefinition (Core/TriangleMesh.cu:111-183)
Mesh make_shape(const ObjectDef& d, const M4& t, const std::string& base) {
    if (d.name == "trianglemesh" || d.name == "plymesh") {
        if (d.params.has_string("filename")) return read_ply(join(base, d.params.str("filename")), t);
        if (d.params.has_list("P") && d.params.has_list("indices")) {
            const auto& P = d.params.list("P");
            const auto& I = d.params.list("indices");
            Mesh m;
            for (size_t k = 0; k + 2 < P.size(); k += 3) m.pos.push_back(apply(t, {P[k], P[k + 1], P[k + 2]}));
            for (size_t k = 0; k + 2 < I.size(); k += 3)
                m.idx.insert(m.idx.end(), {(int)I[k], (int)I[k + 1], (int)I[k + 2]});
            if (d.params.has_list("N")) {  // Core/TriangleMesh.cu:152-166
                const auto& N = d.params.list("N");
                if (N.size() != P.size()) throw ParseError("trianglemesh: N must give one normal per vertex");
                std::vector<V3> n;
                for (size_t k = 0; k + 2 < N.size(); k += 3) n.push_back({N[k], N[k + 1], N[k + 2]});
                m.nrm = transform_normals(n, t);
            }
            if (d.params.has_list("uv")) {  // Core/TriangleMesh.cu:171-178
                m.uv = d.params.list("uv");
                if (m.uv.size() != 2 * m.pos.size())
                    throw ParseError("trianglemesh: uv must give two numbers per vertex");
            }
            return m;
        }
        throw ParseError("No valid inputs for triangle mesh");
    }
    if (d.name == "sphere") return sphere_mesh(d.params.num("radius"), t);
    if (d.name == "disk") return disk_mesh(d.params.num("radius"), d.params.has_num("height") ? d.params.num("height") : 0.f, t);
    throw ParseError("ShapeMesh type not supported: " + d.name);
}

struct Setup {
    ObjectDef camera, film, integrator, sampler, filter;
    M4 view = identity();
};

// Parsing.cu:172-268
void parse_options(Buf& b, Setup& su, const cav_pbrt_overrides* ov, Builder& bd) {
    su.filter.defined = true;
    su.filter.name = "triangle";
    for (;;) {
        const Token& k = b.peek();
        if (k.type != Tok::Keyword)
            throw ParseError("Parsing Error: Keyword expected at token " + std::to_string(b.i) + ": " + k.text);
        if (k.text == "WorldBegin") break;
        if (k.text == "LookAt") {
            ++b.i;
            float v[9];
            for (float& x : v) x = b.pop(Tok::Num, "number").num;
            V3 eye = apply(su.view, {v[0], v[1], v[2]});
            V3 center = apply(su.view, {v[3], v[4], v[5]});
            V3 up = apply3(su.view, {v[6], v[7], v[8]});
            su.view = mul(look_at_lh(eye, center, up), su.view);
        } else if (k.text == "Camera") {
            su.camera = read_object(b);
        } else if (k.text == "Film") {
            su.film = read_object(b);
            if (ov && ov->output && ov->output[0]) su.film.params.strings["filename"] = ov->output;
        } else if (k.text == "Sampler") {
            su.sampler = read_object(b);
            if (ov && ov->spp > 0) su.sampler.params.nums["pixelsamples"] = (float)ov->spp;
        } else if (read_transform(b, su.view)) {
        } else if (k.text == "Integrator") {
            su.integrator = read_object(b);
            if (ov && ov->integrator && ov->integrator[0]) su.integrator.name = ov->integrator;
        } else if (k.text == "PixelFilter") {
            su.filter = read_object(b);
        } else {
            bd.warn("ignored scene-wide directive " + k.text);
            skip_object(b);
        }
    }
    if (!(su.camera.defined && su.film.defined && su.integrator.defined && su.sampler.defined))
        throw ParseError("incomplete scene-wide options: requires Camera, Film, Integrator, Sampler, PixelFilter");
}

struct State {
    M4 t;
    std::shared_ptr<Material> mat;  // null: default matte
};

void add_shape(Builder& bd, const Mesh& m, const Material& mat, bool reverse) {
    cav_mesh c = mat.m;
    const int64_t first = (int64_t)bd.soup.size() / 9;
    const int64_t nv = (int64_t)m.pos.size();
    int ntris = 0;
    const bool has_n = !m.nrm.empty();
    if (has_n && (int64_t)m.nrm.size() != nv) throw ParseError("normals and positions differ in count");
    const bool has_uv = !m.uv.empty();
    if (has_uv && (int64_t)m.uv.size() != 2 * nv) throw ParseError("texture coordinates and positions differ in count");
    for (size_t k = 0; k + 2 < m.idx.size(); k += 3) {
        for (int j = 0; j < 3; ++j) {
            const int vi = m.idx[k + j];
            if (vi < 0 || vi >= nv) throw ParseError("triangle index out of range");
            bd.soup.insert(bd.soup.end(), {m.pos[vi].x, m.pos[vi].y, m.pos[vi].z});
            if (has_n) bd.normals.insert(bd.normals.end(), {m.nrm[vi].x, m.nrm[vi].y, m.nrm[vi].z});
            else bd.normals.insert(bd.normals.end(), {0.f, 0.f, 0.f});
            if (has_uv) bd.uv.insert(bd.uv.end(), {m.uv[2 * vi], m.uv[2 * vi + 1]});
            else bd.uv.insert(bd.uv.end(), {0.f, 0.f});
        }
        ++ntris;
    }
    if (first + ntris > INT32_MAX) throw ParseError("more than 2^31 triangles");
    c.first_tri = (int32_t)first;
    c.num_tris = ntris;
    for (int q = 0; q < 4; ++q) c.sphere[q] = m.sphere[q];
    c.reverse_orientation = reverse ? 1 : 0;
    c.is_light = 0;
    const bool light = bd.next_is_light;
    if (light) {
        c.is_light = 1;
        for (int q = 0; q < 3; ++q) c.emission[q] = bd.light_L[q];
        bd.next_is_light = false;
    } else {
        for (int q = 0; q < 3; ++q) c.emission[q] = 0.f;
    }
    if (ntris == 0) {
        bd.warn("shape without triangles skipped");
        return;
    }
    if (light) {
        bd.order[(size_t)bd.pending_slot] = (int32_t)bd.meshes.size();
        bd.pending_slot = -1;
    }
    bd.meshes.push_back(c);
    bd.mesh_normals.push_back(has_n ? 1 : 0);
    bd.any_normals = bd.any_normals || has_n;
    bd.any_uv = bd.any_uv || has_uv;
}

void add_env_light(Builder& bd, const cav_env_light& e) {
    bd.order.push_back(-1 - (int32_t)bd.envs.size());
    bd.envs.push_back(e);
}

// Parsing.cu:272-388: one XBegin ... XEnd section
void parse_section(Buf& b, State st, const std::string& base, Builder& bd) {
    const Token& begin = b.pop(Tok::Keyword, "XBegin");
    const std::string& w = begin.text;
    if (w.size() < 5 || w.compare(w.size() - 5, 5, "Begin") != 0)
        throw ParseError("Parsing Error: XXXBegin expected at token " + std::to_string(b.i - 1) + ": " + w);
    const std::string name = w.substr(0, w.size() - 5);
    bool reverse = false;
    for (;;) {
        const Token& k = b.peek();
        if (k.type != Tok::Keyword)
            throw ParseError("Parsing Error: Keyword expected at token " + std::to_string(b.i) + ": " + k.text);
        const std::string& kw = k.text;
        if (kw.size() >= 3 && kw.compare(kw.size() - 3, 3, "End") == 0) {
            if (kw.substr(0, kw.size() - 3) != name)
                throw ParseError("Parsing Error: Mismatching subsection names " + name + " / " + kw);
            break;
        } else if (kw.size() >= 5 && kw.compare(kw.size() - 5, 5, "Begin") == 0) {
            parse_section(b, st, base, bd);
        } else if (kw == "Shape") {
            const ObjectDef d = read_object(b);
            const Mesh m = make_shape(d, st.t, base);
            add_shape(bd, m, st.mat ? *st.mat : default_matte(), reverse);
        } else if (kw == "AreaLightSource") {
            const ObjectDef d = read_object(b);
            float L[3] = {1, 1, 1};
            const char* key = d.params.has_list("L") ? "L" : (d.params.has_list("l") ? "l" : nullptr);
            if (key) {
                const auto& v = d.params.list(key);
                for (int q = 0; q < 3; ++q) L[q] = v.size() >= 3 ? v[q] : (v.empty() ? 1.f : v[0]);
            }
            if (d.name == "diffuse") {  // Parsing.cu:320-333: listed now, bound to the next shape
                if (bd.pending_slot >= 0) {
                    bd.warn("AreaLightSource without a shape dropped");
                    bd.order.erase(bd.order.begin() + bd.pending_slot);
                }
                bd.next_is_light = true;
                for (int q = 0; q < 3; ++q) bd.light_L[q] = L[q];
                bd.pending_slot = (int)bd.order.size();
                bd.order.push_back(INT32_MIN);
            } else {
                bd.warn("area light '" + d.name + "' counts as a black light, as in the reference");
                add_env_light(bd, cav_env_light{{0.f, 0.f, 0.f}, -1, {1, 0, 0, 0, 1, 0, 0, 0, 1}});
            }
        } else if (kw == "LightSource") {  // Parsing.cu:335-342, Lights/LightObject.h:89-126
            const ObjectDef d = read_object(b);
            cav_env_light e{{1.f, 1.f, 1.f}, -1, {1, 0, 0, 0, 1, 0, 0, 0, 1}};
            const char* key = d.params.has_list("L") ? "L" : (d.params.has_list("l") ? "l" : nullptr);
            if (d.name == "infinite" && d.params.has_list("scale")) key = "scale";
            if (key) {
                const auto& v = d.params.list(key);
                for (int q = 0; q < 3; ++q) e.color[q] = v.size() >= 3 ? v[q] : (v.empty() ? 1.f : v[0]);
            }
            if (d.name == "infinite") {
                if (d.params.has_string("mapname")) {
                    cav::Image img;
                    std::string err;
                    if (!cav::read_texture_file(join(base, d.params.str("mapname")), img, err)) throw ParseError(err);
                    e.texture = (int32_t)bd.images.size();
                    bd.images.push_back(std::move(img));
                }
                // glm::mat3(inverse(lightToWorld)), as a row-major 3x3
                float inv[16];
                if (!invert(st.t.m, inv)) throw ParseError("infinite light: singular transform");
                for (int r = 0; r < 3; ++r)
                    for (int cc = 0; cc < 3; ++cc) e.world_to_light[r * 3 + cc] = inv[cc * 4 + r];
                bd.env_background = (int)bd.envs.size();
            } else {
                bd.warn("light source '" + d.name + "' counts as a black light, as in the reference");
                e.color[0] = e.color[1] = e.color[2] = 0.f;
            }
            add_env_light(bd, e);
        } else if (read_transform(b, st.t)) {
        } else if (kw == "MakeNamedMaterial") {
            const ObjectDef d = read_object(b);
            bd.named[d.name] = make_material(d.params, bd);
        } else if (kw == "Material") {
            ObjectDef d = read_object(b);
            d.params.strings["type"] = d.name;
            st.mat = std::make_shared<Material>(make_material(d.params, bd));
        } else if (kw == "Texture") {  // Parsing.cu:355-358, Core/Texture.cu:12-27
            const ObjectDef d = read_object(b);
            cav::Image img;
            if (d.params.has_string("filename")) {
                std::string err;
                if (!cav::read_texture_file(join(base, d.params.str("filename")), img, err)) throw ParseError(err);
            } else {  // a texture without a file is one transparent black texel
                img.w = img.h = 1;
                img.rgba.assign(4, 0);
            }
            bd.texture_ids[d.name] = (int)bd.images.size();
            bd.images.push_back(std::move(img));
        } else if (kw == "NamedMaterial") {
            ++b.i;
            const std::string n = b.pop(Tok::String, "material name").text;
            if (!bd.named.count(n)) throw ParseError("Named material not found: " + n);
            st.mat = std::make_shared<Material>(bd.named.at(n));
        } else if (kw == "Include") {
            ++b.i;
            const std::string file = b.pop(Tok::String, "file name").text;
            std::vector<Token> inc = lex(join(base, file));
            b.toks.insert(b.toks.begin() + (std::ptrdiff_t)b.i, inc.begin(), inc.end());
        } else if (kw == "ReverseOrientation") {
            ++b.i;
            reverse = true;
        } else {
            bd.warn("ignored directive " + kw);
            skip_object(b);
        }
    }
    b.pop(Tok::Keyword, "XEnd");
}

}  // namespace

struct cav_pbrt {
    std::vector<float> soup;
    std::vector<float> normals;        // empty when no mesh has vertex normals
    std::vector<uint8_t> mesh_normals;
    std::vector<float> uv;             // empty when no mesh has texture coordinates
    std::vector<cav::Image> images;
    std::vector<cav_texture> textures; // views of images
    std::vector<cav_env_light> envs;
    std::vector<int32_t> order;        // the light list (cav_scene_desc::light_order)
    int env_background = -1;
    std::vector<cav_mesh> meshes;
    cav_render_setup setup{};
    std::string messages;
};

namespace {

int integrator_id(const std::string& n, int& depth, const Params& p, Builder& bd) {
    // Integrators/CreateIntegrator.cu:10-25
    if (n == "path" || n == "rlpath") {
        depth = (int)p.num("maxdepth");
        return n == "path" ? CAV_INTEGRATOR_PATH : CAV_INTEGRATOR_RLPATH;
    }
    if (n == "directlighting") {
        depth = 1;
        return CAV_INTEGRATOR_DIRECT;
    }
    bd.warn("unrecognized integrator name. using pathtracer as default.");
    depth = 8;
    return CAV_INTEGRATOR_PATH;
}

void copy_err(char* err, int err_len, const std::string& s) {
    if (err && err_len > 0) {
        std::snprintf(err, (size_t)err_len, "%s", s.c_str());
    }
}

}  // namespace

extern "C" {

int cav_pbrt_parse(const char* path, const cav_pbrt_overrides* ov, cav_pbrt** out, char* err, int err_len) {
    if (!path || !out) return CAV_ERR_INVALID;
    *out = nullptr;
    try {
        Buf b;
        b.toks = lex(path);
        Builder bd;
        Setup su;
        parse_options(b, su, ov, bd);
        State st{identity(), nullptr};
        parse_section(b, st, dir_of(path), bd);
        if (bd.meshes.empty()) throw ParseError("scene has no shapes");
        if (bd.pending_slot >= 0) {
            bd.warn("AreaLightSource without a shape dropped");
            bd.order.erase(bd.order.begin() + bd.pending_slot);
        }
        std::unique_ptr<cav_pbrt> r(new cav_pbrt());
        cav_render_setup& s = r->setup;
        s.width = (int)su.film.params.num("xresolution");
        s.height = (int)su.film.params.num("yresolution");
        if (s.width < 1 || s.height < 1) throw ParseError("film resolution must be positive");
        float c2w[16];
        if (!invert(su.view.m, c2w)) throw ParseError("camera transform is singular");
        // Cameras/CameraObject.h:91-93: only the perspective camera exists
        if (su.camera.name != "perspective") bd.warn("camera '" + su.camera.name + "' rendered as perspective");
        if (cav_camera_from_matrix(c2w, su.camera.params.num("fov"), s.width, s.height, &s.camera) != CAV_OK)
            throw ParseError("bad camera");
        s.spp = (int)su.sampler.params.num("pixelsamples");
        if (s.spp < 1) throw ParseError("pixelsamples must be positive");
        if (su.sampler.name != "halton") bd.warn("sampler '" + su.sampler.name + "' rendered with halton");
        s.integrator = integrator_id(su.integrator.name, s.max_depth, su.integrator.params, bd);
        // Scene::prepareForRender (Core/Scene.cu:141-146)
        s.bvh_opt_rounds = ov ? ov->bvh_opt_rounds : -1;
        if (s.bvh_opt_rounds < 0) s.bvh_opt_rounds = s.spp >= 64 ? 10 : 3;
        s.morton_bits = (ov && ov->morton_bits) ? ov->morton_bits : 63;
        const std::string fn =
            su.film.params.has_string("filename") ? su.film.params.str("filename") : std::string("output.png");
        std::snprintf(s.output, sizeof s.output, "%s", fn.c_str());
        s.warnings = bd.warnings;
        s.ntris = (int64_t)bd.soup.size() / 9;
        s.nmeshes = (int)bd.meshes.size();
        s.nlights = (int32_t)bd.order.size();
        r->soup.swap(bd.soup);
        r->meshes.swap(bd.meshes);
        if (bd.any_normals) {
            r->normals.swap(bd.normals);
            r->mesh_normals.swap(bd.mesh_normals);
        }
        if (bd.any_uv) r->uv.swap(bd.uv);
        r->envs.swap(bd.envs);
        r->order.swap(bd.order);
        r->env_background = bd.env_background;
        r->images.swap(bd.images);
        for (const cav::Image& im : r->images) r->textures.push_back(cav_texture{im.rgba.data(), im.w, im.h});
        for (const std::string& m : bd.messages) r->messages += m + "\n";
        *out = r.release();
        return CAV_OK;
    } catch (const std::exception& e) {
        copy_err(err, err_len, e.what());
        return CAV_ERR_PARSE;
    }
}

int cav_pbrt_get(const cav_pbrt* p, const float** tris, const cav_mesh** meshes, cav_render_setup* setup,
                 const char** messages) {
    if (!p) return CAV_ERR_INVALID;
    if (tris) *tris = p->soup.data();
    if (meshes) *meshes = p->meshes.data();
    if (setup) *setup = p->setup;
    if (messages) *messages = p->messages.c_str();
    return CAV_OK;
}

int cav_pbrt_get_normals(const cav_pbrt* p, const float** tri_normals, const uint8_t** mesh_normals) {
    if (!p) return CAV_ERR_INVALID;
    if (tri_normals) *tri_normals = p->normals.empty() ? nullptr : p->normals.data();
    if (mesh_normals) *mesh_normals = p->normals.empty() ? nullptr : p->mesh_normals.data();
    return CAV_OK;
}

int cav_pbrt_get_textures(const cav_pbrt* p, const float** tri_uv, const cav_texture** textures,
                          int32_t* ntextures) {
    if (!p) return CAV_ERR_INVALID;
    if (tri_uv) *tri_uv = p->uv.empty() ? nullptr : p->uv.data();
    if (textures) *textures = p->textures.empty() ? nullptr : p->textures.data();
    if (ntextures) *ntextures = (int32_t)p->textures.size();
    return CAV_OK;
}

int cav_pbrt_get_lights(const cav_pbrt* p, const cav_env_light** env_lights, int32_t* nenv,
                        const int32_t** light_order, int32_t* nlight_order, int32_t* env_background) {
    if (!p) return CAV_ERR_INVALID;
    if (env_lights) *env_lights = p->envs.empty() ? nullptr : p->envs.data();
    if (nenv) *nenv = (int32_t)p->envs.size();
    if (light_order) *light_order = p->order.data();
    if (nlight_order) *nlight_order = (int32_t)p->order.size();
    if (env_background) *env_background = p->env_background;
    return CAV_OK;
}

void cav_pbrt_destroy(cav_pbrt* p) { delete p; }

int cav_scene_create_from_pbrt(cav_ctx* ctx, const char* path, const cav_pbrt_overrides* ov, cav_scene** out,
                               cav_render_setup* setup) {
    if (!ctx || !path || !out) return CAV_ERR_INVALID;
    *out = nullptr;
    cav_pbrt* p = nullptr;
    char err[1024] = {0};
    int rc = cav_pbrt_parse(path, ov, &p, err, sizeof err);
    if (rc != CAV_OK) return cav::fail(ctx, rc, err);
    const cav_render_setup& s = p->setup;
    cav_scene_desc d{};
    d.tris = p->soup.data();
    d.ntris = (int32_t)s.ntris;
    d.meshes = p->meshes.data();
    d.nmeshes = s.nmeshes;
    d.tri_normals = p->normals.empty() ? nullptr : p->normals.data();
    d.mesh_normals = p->normals.empty() ? nullptr : p->mesh_normals.data();
    d.bvh_opt_rounds = s.bvh_opt_rounds;
    d.morton_bits = s.morton_bits;
    d.tri_uv = p->uv.empty() ? nullptr : p->uv.data();
    d.textures = p->textures.empty() ? nullptr : p->textures.data();
    d.ntextures = (int32_t)p->textures.size();
    d.env_lights = p->envs.empty() ? nullptr : p->envs.data();
    d.nenv = (int32_t)p->envs.size();
    d.light_order = p->order.data();
    d.nlight_order = (int32_t)p->order.size();
    d.env_background = p->env_background;
    rc = cav_scene_create_desc(ctx, &d, out);
    if (rc == CAV_OK && setup) *setup = s;
    cav_pbrt_destroy(p);
    return rc;
}

}  // extern "C"
