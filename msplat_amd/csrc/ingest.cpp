// ingest.cpp — see ingest.h.
#include "ingest.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <limits>
#include <random>
#include <sstream>
#include <stdexcept>
#include <unordered_map>

#include "image_io.h"
#include "json.h"

namespace fs = std::filesystem;

namespace msplat_hip {

namespace {

[[noreturn]] void bad(const std::string& what) { throw std::runtime_error(what); }

std::string read_text(const fs::path& p) {
    std::ifstream f(p, std::ios::binary);
    if (!f) bad("cannot open " + p.string());
    std::stringstream ss;
    ss << f.rdbuf();
    return ss.str();
}

// A little-endian binary reader that throws at the end of the data.
struct BinReader {
    std::vector<uint8_t> b;
    size_t p = 0;
    explicit BinReader(const fs::path& path) {
        std::ifstream f(path, std::ios::binary);
        if (!f) bad("cannot open " + path.string());
        b.assign(std::istreambuf_iterator<char>(f), std::istreambuf_iterator<char>());
    }
    void need(size_t n) const {
        if (p + n > b.size()) bad("truncated binary file");
    }
    template <class T> T get() {
        need(sizeof(T));
        T v;
        std::memcpy(&v, b.data() + p, sizeof(T));
        p += sizeof(T);
        return v;
    }
    void skip(uint64_t n) {
        if (n > b.size() - p) bad("truncated binary file");
        p += (size_t)n;
    }
    std::string cstr() {
        std::string s;
        while (true) {
            need(1);
            const char c = (char)b[p++];
            if (c == '\0') break;
            s += c;
        }
        return s;
    }
};

// ---------------------------------------------------------------------------------- COLMAP
struct ColmapCam {
    int model = 0, width = 0, height = 0;
    float fx = 0, fy = 0, cx = 0, cy = 0, k1 = 0, k2 = 0, p1 = 0, p2 = 0;
};

std::unordered_map<uint32_t, ColmapCam> colmap_cameras(const fs::path& path) {
    BinReader r(path);
    const uint64_t n = r.get<uint64_t>();
    std::unordered_map<uint32_t, ColmapCam> out;
    for (uint64_t i = 0; i < n; i++) {
        const uint32_t id = r.get<uint32_t>();
        ColmapCam c;
        c.model = (int)r.get<uint32_t>();
        c.width = (int)r.get<uint64_t>();
        c.height = (int)r.get<uint64_t>();
        auto d = [&]() { return (float)r.get<double>(); };
        switch (c.model) {  // SIMPLE_PINHOLE, PINHOLE, SIMPLE_RADIAL, RADIAL, OPENCV
            case 0: c.fx = c.fy = d(); c.cx = d(); c.cy = d(); break;
            case 1: c.fx = d(); c.fy = d(); c.cx = d(); c.cy = d(); break;
            case 2: c.fx = c.fy = d(); c.cx = d(); c.cy = d(); c.k1 = d(); break;
            case 3: c.fx = c.fy = d(); c.cx = d(); c.cy = d(); c.k1 = d(); c.k2 = d(); break;
            case 4: c.fx = d(); c.fy = d(); c.cx = d(); c.cy = d(); c.k1 = d(); c.k2 = d(); c.p1 = d(); c.p2 = d(); break;
            default: bad("Unsupported COLMAP camera model: " + std::to_string(c.model));
        }
        out[id] = c;
    }
    return out;
}

struct ColmapImage {
    uint32_t cam = 0;
    double q[4] = {1, 0, 0, 0}, t[3] = {0, 0, 0};
    std::string name;
};

std::vector<ColmapImage> colmap_images(const fs::path& path) {
    BinReader r(path);
    const uint64_t n = r.get<uint64_t>();
    std::vector<ColmapImage> out;
    for (uint64_t i = 0; i < n; i++) {
        ColmapImage im;
        (void)r.get<uint32_t>();  // image id
        for (double& v : im.q) v = r.get<double>();
        for (double& v : im.t) v = r.get<double>();
        im.cam = r.get<uint32_t>();
        im.name = r.cstr();
        r.skip(r.get<uint64_t>() * 24);  // 2-D points: x, y, point id
        out.push_back(im);
    }
    return out;
}

// world-to-camera (quaternion w, x, y, z and translation) -> camera-to-world, camera axes
// flipped to OpenGL (y up, z backwards), as load_colmap.cpp:10-18, 89-112 compute it
void colmap_c2w(const double q[4], const double t[3], float out[16]) {
    const double nrm = std::sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
    const double w = q[0] / nrm, x = q[1] / nrm, y = q[2] / nrm, z = q[3] / nrm;
    const float R[9] = {(float)(1 - 2 * (y * y + z * z)), (float)(2 * (x * y - w * z)), (float)(2 * (x * z + w * y)),
                        (float)(2 * (x * y + w * z)), (float)(1 - 2 * (x * x + z * z)), (float)(2 * (y * z - w * x)),
                        (float)(2 * (x * z - w * y)), (float)(2 * (y * z + w * x)), (float)(1 - 2 * (x * x + y * y))};
    const float Rt[9] = {R[0], R[3], R[6], R[1], R[4], R[7], R[2], R[5], R[8]};
    const float tf[3] = {(float)t[0], (float)t[1], (float)t[2]};
    float C[3];
    for (int i = 0; i < 3; i++) C[i] = -(Rt[3 * i] * tf[0] + Rt[3 * i + 1] * tf[1] + Rt[3 * i + 2] * tf[2]);
    for (int i = 0; i < 3; i++) {
        out[4 * i + 0] = Rt[3 * i];
        out[4 * i + 1] = -Rt[3 * i + 1];
        out[4 * i + 2] = -Rt[3 * i + 2];
        out[4 * i + 3] = C[i];
    }
    out[12] = 0.f; out[13] = 0.f; out[14] = 0.f; out[15] = 1.f;
}

// an existing file among path + ("", .png, .jpg, .jpeg, .JPG) (load_nerfstudio.cpp:11-16)
std::string with_image_ext(const std::string& p) {
    if (fs::exists(p)) return p;
    for (const char* e : {".png", ".jpg", ".jpeg", ".JPG"})
        if (fs::exists(p + e)) return p + e;
    return p;
}

// PLY property types and sizes
int ply_size(const std::string& t) {
    if (t == "char" || t == "int8" || t == "uchar" || t == "uint8") return 1;
    if (t == "short" || t == "int16" || t == "ushort" || t == "uint16") return 2;
    if (t == "int" || t == "int32" || t == "uint" || t == "uint32" || t == "float" || t == "float32") return 4;
    if (t == "double" || t == "float64") return 8;
    return 0;
}

double ply_value(const std::string& t, const uint8_t* p) {
    if (t == "float" || t == "float32") { float v; std::memcpy(&v, p, 4); return v; }
    if (t == "double" || t == "float64") { double v; std::memcpy(&v, p, 8); return v; }
    if (t == "uchar" || t == "uint8") return p[0];
    if (t == "char" || t == "int8") return (int8_t)p[0];
    if (t == "ushort" || t == "uint16") { uint16_t v; std::memcpy(&v, p, 2); return v; }
    if (t == "short" || t == "int16") { int16_t v; std::memcpy(&v, p, 2); return v; }
    if (t == "uint" || t == "uint32") { uint32_t v; std::memcpy(&v, p, 4); return v; }
    if (t == "int" || t == "int32") { int32_t v; std::memcpy(&v, p, 4); return v; }
    return 0.0;
}

bool ply_is_float(const std::string& t) { return t == "float" || t == "float32" || t == "double" || t == "float64"; }

// a binary colour sample to 8 bits as load_ply.cpp:107-113 reads it: 8-bit as is, 16-bit
// its high byte, 32-bit float times 255 (truncated), any other type grey (128)
uint8_t ply_color(const std::string& t, double v) {
    if (t == "uchar" || t == "uint8") return (uint8_t)v;
    if (t == "ushort" || t == "uint16") return (uint8_t)((uint32_t)v >> 8);
    if (t == "float" || t == "float32") return (uint8_t)((float)v * 255.0f);
    return 128;
}

}  // namespace

ScenePoints read_ply_points(const std::string& path) {
    std::ifstream f(path, std::ios::binary);
    if (!f) bad("Cannot open PLY file: " + path);
    std::string line;
    std::getline(f, line);
    if (line.find("ply") == std::string::npos) bad("Not a PLY file: " + path);
    bool binary = false, ascii = false, in_vertex = false;
    int64_t nv = 0;
    struct Prop { std::string name, type; int off; };
    std::vector<Prop> props;
    int stride = 0;
    while (std::getline(f, line)) {
        if (!line.empty() && line.back() == '\r') line.pop_back();
        if (line == "end_header") break;
        std::istringstream ss(line);
        std::string tok;
        ss >> tok;
        if (tok == "format") {
            std::string fmt;
            ss >> fmt;
            binary = fmt == "binary_little_endian";
            ascii = fmt == "ascii";
        } else if (tok == "element") {
            std::string name;
            ss >> name;
            in_vertex = name == "vertex";
            if (in_vertex) ss >> nv;
        } else if (tok == "property" && in_vertex) {
            std::string type, name;
            ss >> type >> name;
            if (type == "list") bad("PLY: list properties on vertices are not supported");
            props.push_back({name, type, stride});
            stride += ply_size(type);
        }
    }
    if (nv <= 0) bad("PLY has no vertices");
    if (!binary && !ascii) bad("Unsupported PLY format (need ascii or binary_little_endian)");
    auto find = [&](const char* n) -> int {
        for (size_t i = 0; i < props.size(); i++)
            if (props[i].name == n) return (int)i;
        return -1;
    };
    const int ix = find("x"), iy = find("y"), iz = find("z"), ir = find("red"), ig = find("green"), ib = find("blue");
    if (ix < 0 || iy < 0 || iz < 0) bad("PLY missing x/y/z properties");
    const bool color = ir >= 0 && ig >= 0 && ib >= 0;
    ScenePoints pts;
    pts.xyz.resize((size_t)nv * 3);
    pts.rgb.assign((size_t)nv * 3, 128);  // grey without colours
    if (binary) {
        std::vector<uint8_t> rec((size_t)stride);
        for (int64_t i = 0; i < nv; i++) {
            f.read(reinterpret_cast<char*>(rec.data()), stride);
            if (!f) bad("PLY: truncated vertex data");
            auto val = [&](int k) { return ply_value(props[k].type, rec.data() + props[k].off); };
            pts.xyz[3 * i] = (float)val(ix);
            pts.xyz[3 * i + 1] = (float)val(iy);
            pts.xyz[3 * i + 2] = (float)val(iz);
            if (color) {
                pts.rgb[3 * i] = ply_color(props[ir].type, val(ir));
                pts.rgb[3 * i + 1] = ply_color(props[ig].type, val(ig));
                pts.rgb[3 * i + 2] = ply_color(props[ib].type, val(ib));
            }
        }
    } else {
        for (int64_t i = 0; i < nv; i++) {
            if (!std::getline(f, line)) bad("PLY: truncated vertex data");
            std::istringstream ss(line);
            std::vector<std::string> tk;
            std::string t;
            while (ss >> t) tk.push_back(t);
            if (tk.size() < props.size()) bad("PLY: short vertex line");
            pts.xyz[3 * i] = std::stof(tk[ix]);
            pts.xyz[3 * i + 1] = std::stof(tk[iy]);
            pts.xyz[3 * i + 2] = std::stof(tk[iz]);
            if (color) {
                const int cs[3] = {ir, ig, ib};
                for (int c = 0; c < 3; c++) {
                    const float v = std::stof(tk[cs[c]]);
                    pts.rgb[3 * i + c] = ply_is_float(props[cs[c]].type) ? (uint8_t)(v * 255.0f) : (uint8_t)v;
                }
            }
        }
    }
    return pts;
}

ScenePoints read_colmap_points(const std::string& path) {
    BinReader r(path);
    const uint64_t n = r.get<uint64_t>();
    ScenePoints pts;
    pts.xyz.resize((size_t)n * 3);
    pts.rgb.resize((size_t)n * 3);
    for (uint64_t i = 0; i < n; i++) {
        (void)r.get<uint64_t>();  // point id
        for (int k = 0; k < 3; k++) pts.xyz[3 * i + k] = (float)r.get<double>();
        for (int k = 0; k < 3; k++) pts.rgb[3 * i + k] = r.get<uint8_t>();
        (void)r.get<double>();  // reprojection error
        r.skip(r.get<uint64_t>() * 8);  // track: image id, 2-D point index
    }
    return pts;
}

void auto_scale_and_center(Scene& s) {
    if (s.cams.empty()) return;
    float mean[3] = {0.f, 0.f, 0.f};
    for (const auto& c : s.cams)
        for (int k = 0; k < 3; k++) mean[k] += c.c2w[4 * k + 3];
    const int n = (int)s.cams.size();
    for (int k = 0; k < 3; k++) mean[k] /= (float)n;
    for (int k = 0; k < 3; k++) s.translation[k] = mean[k];
    for (auto& c : s.cams)
        for (int k = 0; k < 3; k++) c.c2w[4 * k + 3] -= mean[k];
    float mx = 0.f;
    for (const auto& c : s.cams)
        for (int k = 0; k < 3; k++) mx = std::max(mx, std::abs(c.c2w[4 * k + 3]));
    s.scale = mx > 0.f ? 1.0f / mx : 1.0f;
    for (auto& c : s.cams)
        for (int k = 0; k < 3; k++) c.c2w[4 * k + 3] *= s.scale;
    for (int64_t i = 0; i < s.points.count(); i++)
        for (int k = 0; k < 3; k++) s.points.xyz[3 * i + k] = (s.points.xyz[3 * i + k] - mean[k]) * s.scale;
}

Scene load_colmap(const std::string& root_s, const std::string& colmap_images_dir) {
    const fs::path root(root_s);
    const fs::path sparse = fs::exists(root / "cameras.bin") ? root : root / "sparse" / "0";
    const std::string image_dir = !colmap_images_dir.empty() ? colmap_images_dir
                                  : fs::exists(root / "images") ? (root / "images").string()
                                                                : root.string();
    const auto cams = colmap_cameras(sparse / "cameras.bin");
    auto images = colmap_images(sparse / "images.bin");
    std::sort(images.begin(), images.end(), [](const ColmapImage& a, const ColmapImage& b) { return a.name < b.name; });
    Scene s;
    for (const auto& im : images) {
        const auto it = cams.find(im.cam);
        if (it == cams.end()) continue;
        const ColmapCam& cc = it->second;
        SceneCamera c;
        c.width = cc.width; c.height = cc.height;
        c.fx = cc.fx; c.fy = cc.fy; c.cx = cc.cx; c.cy = cc.cy;
        c.k1 = cc.k1; c.k2 = cc.k2; c.p1 = cc.p1; c.p2 = cc.p2;
        c.path = image_dir + "/" + im.name;
        c.relative = im.name;
        colmap_c2w(im.q, im.t, c.c2w);
        s.cams.push_back(c);
    }
    if (fs::exists(sparse / "points3D.bin")) s.points = read_colmap_points((sparse / "points3D.bin").string());
    else if (fs::exists(sparse / "points3D.ply")) s.points = read_ply_points((sparse / "points3D.ply").string());
    auto_scale_and_center(s);
    return s;
}

Scene load_nerfstudio(const std::string& root_s) {
    const fs::path root(root_s);
    const Json j = Json::parse(read_text(root / "transforms.json"));
    const int gw = (int)j.value("w", 0.0), gh = (int)j.value("h", 0.0);
    const float gfx = (float)j.value("fl_x", 0.0), gfy = (float)j.value("fl_y", 0.0);
    const float gcx = (float)j.value("cx", 0.0), gcy = (float)j.value("cy", 0.0);
    const float gk1 = (float)j.value("k1", 0.0), gk2 = (float)j.value("k2", 0.0), gk3 = (float)j.value("k3", 0.0);
    const float gp1 = (float)j.value("p1", 0.0), gp2 = (float)j.value("p2", 0.0);
    Scene s;
    const Json& frames = j.at("frames");
    for (size_t i = 0; i < frames.size(); i++) {
        const Json& fr = frames[i];
        SceneCamera c;
        c.width = (int)fr.value("w", (double)gw);
        c.height = (int)fr.value("h", (double)gh);
        c.fx = (float)fr.value("fl_x", (double)gfx); c.fy = (float)fr.value("fl_y", (double)gfy);
        c.cx = (float)fr.value("cx", (double)gcx); c.cy = (float)fr.value("cy", (double)gcy);
        c.k1 = (float)fr.value("k1", (double)gk1); c.k2 = (float)fr.value("k2", (double)gk2);
        c.k3 = (float)fr.value("k3", (double)gk3);
        c.p1 = (float)fr.value("p1", (double)gp1); c.p2 = (float)fr.value("p2", (double)gp2);
        const std::string fp = fr.at("file_path").str();
        c.path = (!fp.empty() && (fp[0] == '/' || fp[0] == '.')) ? with_image_ext(fp) : with_image_ext((root / fp).string());
        const Json& tm = fr.at("transform_matrix");
        for (int r = 0; r < 4; r++)
            for (int k = 0; k < 4; k++) c.c2w[4 * r + k] = (float)tm[r][k].number();
        s.cams.push_back(c);
    }
    std::sort(s.cams.begin(), s.cams.end(), [](const SceneCamera& a, const SceneCamera& b) { return a.path < b.path; });
    if (j.contains("ply_file_path")) {
        std::string p = j.at("ply_file_path").str();
        if (!p.empty() && p[0] != '/') p = (root / p).string();
        if (fs::exists(p)) s.points = read_ply_points(p);
    }
    if (s.points.count() == 0)
        for (const char* p : {"sparse/0/points3D.ply", "points3D.ply"})
            if (fs::exists(root / p)) {
                s.points = read_ply_points((root / p).string());
                break;
            }
    auto_scale_and_center(s);
    return s;
}

Scene load_polycam(const std::string& root_s) {
    const fs::path root(root_s);
    const fs::path kf = root / "keyframes" / "corrected_cameras", imgs = root / "keyframes" / "corrected_images";
    Scene s;
    if (fs::exists(kf)) {
        std::vector<fs::path> files;
        for (const auto& e : fs::directory_iterator(kf))
            if (e.path().extension() == ".json") files.push_back(e.path());
        std::sort(files.begin(), files.end());
        for (const auto& jp : files) {
            const Json j = Json::parse(read_text(jp));
            SceneCamera c;
            c.width = (int)j.value("width", 0.0);
            c.height = (int)j.value("height", 0.0);
            c.fx = (float)j.value("fx", 0.0); c.fy = (float)j.value("fy", 0.0);
            c.cx = (float)j.value("cx", (double)((float)c.width / 2.0f));
            c.cy = (float)j.value("cy", (double)((float)c.height / 2.0f));
            float R[9], T[3];
            for (int r = 0; r < 3; r++)
                for (int k = 0; k < 3; k++) R[3 * r + k] = (float)j.value("R_" + std::to_string(r) + std::to_string(k), 0.0);
            // translation keys as load_polycam.cpp:49-51 reads them
            T[0] = (float)j.value("t_20", 0.0);
            T[1] = (float)j.value("t_21", 0.0);
            T[2] = (float)j.value("t_22", 0.0);
            const float sgn[3] = {1.f, -1.f, -1.f};  // rows 1 and 2 negated (OpenGL y / z)
            for (int r = 0; r < 3; r++) {
                for (int k = 0; k < 3; k++) c.c2w[4 * r + k] = sgn[r] * R[3 * r + k];
                c.c2w[4 * r + 3] = sgn[r] * T[r];
            }
            c.c2w[12] = 0.f; c.c2w[13] = 0.f; c.c2w[14] = 0.f; c.c2w[15] = 1.f;
            const std::string stem = jp.stem().string();
            c.path = (imgs / (stem + ".png")).string();
            for (const char* e : {".png", ".jpg", ".jpeg", ".JPG"})
                if (fs::exists(imgs / (stem + e))) {
                    c.path = (imgs / (stem + e)).string();
                    break;
                }
            s.cams.push_back(c);
        }
    } else {
        const Json j = Json::parse(read_text(root / "cameras.json"));
        const Json& frames = j.contains("frames") ? j.at("frames") : j;
        for (size_t i = 0; i < frames.size(); i++) {
            const Json& fr = frames[i];
            SceneCamera c;
            c.width = (int)fr.value("width", 0.0);
            c.height = (int)fr.value("height", 0.0);
            c.fx = (float)fr.value("fx", 0.0); c.fy = (float)fr.value("fy", 0.0);
            c.cx = (float)fr.value("cx", (double)((float)c.width / 2.0f));
            c.cy = (float)fr.value("cy", (double)((float)c.height / 2.0f));
            if (fr.contains("transform_matrix")) {
                const Json& tm = fr.at("transform_matrix");
                for (int r = 0; r < 4; r++)
                    for (int k = 0; k < 4; k++) c.c2w[4 * r + k] = (float)tm[r][k].number();
            }
            c.path = fr.value("file_path", std::string());
            if (!c.path.empty() && c.path[0] != '/') c.path = (root / c.path).string();
            s.cams.push_back(c);
        }
    }
    for (const char* p : {"keyframes/point_cloud.ply", "point_cloud.ply", "sparse.ply"})
        if (fs::exists(root / p)) {
            s.points = read_ply_points((root / p).string());
            break;
        }
    auto_scale_and_center(s);
    return s;
}

Scene load_scene(const std::string& root_s, const std::string& colmap_images_dir) {
    const fs::path root(root_s);
    if (fs::exists(root / "transforms.json")) return load_nerfstudio(root_s);
    if (fs::exists(root / "cameras.bin") || fs::exists(root / "sparse" / "0" / "cameras.bin"))
        return load_colmap(root_s, colmap_images_dir);
    if (fs::exists(root / "keyframes" / "corrected_cameras") || fs::exists(root / "cameras.json"))
        return load_polycam(root_s);
    bad("Unrecognized dataset format in: " + root_s +
        " (supported: COLMAP (cameras.bin), Nerfstudio (transforms.json), Polycam (keyframes/))");
}

std::vector<uint8_t> load_camera_image(SceneCamera& cam, float downscale) {
    HostImage img = read_image_rgb(cam.path);
    if (cam.width > 0 && cam.height > 0 && (img.width != cam.width || img.height != cam.height)) {
        const float sx = (float)img.width / (float)cam.width, sy = (float)img.height / (float)cam.height;
        cam.fx *= sx; cam.fy *= sy; cam.cx *= sx; cam.cy *= sy;
        cam.width = img.width; cam.height = img.height;
    } else if (cam.width <= 0 || cam.height <= 0) {
        cam.width = img.width; cam.height = img.height;
    }
    if (downscale > 1.0f) {
        const int nw = (int)((float)cam.width / downscale), nh = (int)((float)cam.height / downscale);
        if (nw <= 0 || nh <= 0) bad("downscale factor leaves an empty image: " + cam.path);
        img = resize_area(img, nw, nh);
        const float sx = (float)nw / (float)cam.width, sy = (float)nh / (float)cam.height;
        cam.fx *= sx; cam.fy *= sy; cam.cx *= sx; cam.cy *= sy;
        cam.width = nw; cam.height = nh;
    }
    if (cam.distorted()) {
        Undistorted u = undistort_image(img, cam.fx, cam.fy, cam.cx, cam.cy, cam.k1, cam.k2, cam.p1, cam.p2, cam.k3);
        img = std::move(u.image);
        cam.fx = u.fx; cam.fy = u.fy; cam.cx = u.cx; cam.cy = u.cy;
        cam.width = img.width; cam.height = img.height;
        cam.k1 = cam.k2 = cam.k3 = cam.p1 = cam.p2 = 0.f;
    }
    return quantize_rgb(img);
}

bool load_camera_mask(const SceneCamera& cam, const std::string& dir, std::vector<uint8_t>& mask, int64_t& kept) {
    kept = 0;
    if (dir.empty() || cam.relative.empty() || cam.width <= 0 || cam.height <= 0) return false;
    const fs::path appended = fs::path(dir) / (cam.relative + ".png");
    const fs::path swapped = (fs::path(dir) / cam.relative).replace_extension(".png");
    std::string path;
    if (fs::exists(appended)) path = appended.string();
    else if (fs::exists(swapped)) path = swapped.string();
    else return false;
    HostImage m = read_image_rgb(path);
    if (m.empty()) return false;
    if (m.width != cam.width || m.height != cam.height) m = resize_area(m, cam.width, cam.height);
    const size_t n = (size_t)cam.width * cam.height;
    mask.assign(n, 0);
    for (size_t i = 0; i < n; i++) {
        const bool k = m.rgb[3 * i] > 0.5f;
        mask[i] = k ? 255 : 0;
        kept += k ? 1 : 0;
    }
    return kept != (int64_t)n;  // all white: nothing masked
}

// ------------------------------------------------------------------------------- kNN grid
std::vector<float> knn_mean_distance(const float* xyz, int64_t n, int k) {
    std::vector<float> out((size_t)std::max<int64_t>(n, 0), 0.f);
    if (n <= 0 || k < 2) return out;
    // bounding box of the finite points and a grid of about two points per cell
    float lo[3] = {std::numeric_limits<float>::max(), std::numeric_limits<float>::max(), std::numeric_limits<float>::max()};
    float hi[3] = {-lo[0], -lo[1], -lo[2]};
    int64_t nf = 0;
    for (int64_t i = 0; i < n; i++) {
        const float* p = xyz + 3 * i;
        if (!(std::isfinite(p[0]) && std::isfinite(p[1]) && std::isfinite(p[2]))) continue;
        nf++;
        for (int a = 0; a < 3; a++) { lo[a] = std::min(lo[a], p[a]); hi[a] = std::max(hi[a], p[a]); }
    }
    if (nf == 0) return out;
    double ext[3], vol = 1.0;
    for (int a = 0; a < 3; a++) {
        ext[a] = std::max((double)hi[a] - lo[a], 1e-9);
        vol *= ext[a];
    }
    double cell = std::cbrt(vol * 2.0 / (double)nf);
    for (int a = 0; a < 3; a++) cell = std::max(cell, ext[a] / 1024.0);  // at most 1024 cells per axis
    int dim[3];
    for (int a = 0; a < 3; a++) dim[a] = std::max(1, std::min(1024, (int)(ext[a] / cell) + 1));
    auto cell_of = [&](const float* p, int a) {
        return std::min(dim[a] - 1, std::max(0, (int)(((double)p[a] - lo[a]) / cell)));
    };
    const int64_t ncell = (int64_t)dim[0] * dim[1] * dim[2];
    std::vector<int64_t> start((size_t)ncell + 1, 0);
    std::vector<int64_t> order((size_t)nf);
    std::vector<int64_t> cid((size_t)n, -1);
    for (int64_t i = 0; i < n; i++) {
        const float* p = xyz + 3 * i;
        if (!(std::isfinite(p[0]) && std::isfinite(p[1]) && std::isfinite(p[2]))) continue;
        cid[i] = ((int64_t)cell_of(p, 2) * dim[1] + cell_of(p, 1)) * dim[0] + cell_of(p, 0);
        start[cid[i] + 1]++;
    }
    for (int64_t c = 0; c < ncell; c++) start[c + 1] += start[c];
    {
        std::vector<int64_t> fill(start.begin(), start.end() - 1);
        for (int64_t i = 0; i < n; i++)
            if (cid[i] >= 0) order[fill[cid[i]]++] = i;
    }
    if (k > 16) bad("knn_mean_distance: k is at most 16");
    const int kk = (int)std::min<int64_t>(k, nf);
    for (int64_t i = 0; i < n; i++) {
        if (cid[i] < 0) continue;
        const float* q = xyz + 3 * i;
        int c[3] = {cell_of(q, 0), cell_of(q, 1), cell_of(q, 2)};
        float best[16];  // the kk smallest squared distances so far, ascending
        for (int j = 0; j < kk; j++) best[j] = std::numeric_limits<float>::infinity();
        for (int r = 0;; r++) {
            // the shell of cells at Chebyshev distance r from the query's cell
            for (int z = c[2] - r; z <= c[2] + r; z++) {
                if (z < 0 || z >= dim[2]) continue;
                for (int y = c[1] - r; y <= c[1] + r; y++) {
                    if (y < 0 || y >= dim[1]) continue;
                    const bool edge_zy = std::abs(z - c[2]) == r || std::abs(y - c[1]) == r;
                    for (int x = c[0] - r; x <= c[0] + r; x++) {
                        if (x < 0 || x >= dim[0]) continue;
                        if (!edge_zy && std::abs(x - c[0]) != r) continue;
                        const int64_t cc = ((int64_t)z * dim[1] + y) * dim[0] + x;
                        for (int64_t t = start[cc]; t < start[cc + 1]; t++) {
                            const float* p = xyz + 3 * order[t];
                            const float d0 = q[0] - p[0], d1 = q[1] - p[1], d2 = q[2] - p[2];
                            float d = 0.f;
                            d += d0 * d0;
                            d += d1 * d1;
                            d += d2 * d2;
                            if (d < best[kk - 1]) {
                                int j = kk - 1;
                                while (j > 0 && best[j - 1] > d) {
                                    best[j] = best[j - 1];
                                    j--;
                                }
                                best[j] = d;
                            }
                        }
                    }
                }
            }
            // every point outside the searched cube is at least this far away
            double gap = std::numeric_limits<double>::infinity();
            bool all = true;
            for (int a = 0; a < 3; a++) {
                const double face_lo = lo[a] + (double)(c[a] - r) * cell, face_hi = lo[a] + (double)(c[a] + r + 1) * cell;
                if (c[a] - r > 0) { gap = std::min(gap, (double)q[a] - face_lo); all = false; }
                if (c[a] + r + 1 < dim[a]) { gap = std::min(gap, face_hi - (double)q[a]); all = false; }
            }
            if (all) break;
            if (best[kk - 1] < std::numeric_limits<float>::infinity() && gap > 0.0 &&
                (double)best[kk - 1] <= gap * gap * (1.0 - 1e-6))
                break;
        }
        float sum = 0.f;
        for (int j = 1; j < k; j++) sum += j < kk ? std::sqrt(best[j]) : 0.f;
        out[i] = sum / (float)(k - 1);
    }
    return out;
}

void seed_gaussians(const ScenePoints& pts, int degree, float init_scale_mult, float init_opacity, float* means,
                    float* log_scales, float* quats, float* opac, float* f_dc, float* f_rest) {
    const int64_t n = pts.count();
    if (degree < 0 || degree > 3) bad("SH degree must be in [0, 3]");
    std::memcpy(means, pts.xyz.data(), (size_t)n * 3 * sizeof(float));
    const std::vector<float> d = knn_mean_distance(pts.xyz.data(), n, 4);
    for (int64_t i = 0; i < n; i++) {
        const float v = std::log(d[i] * init_scale_mult);
        log_scales[3 * i] = log_scales[3 * i + 1] = log_scales[3 * i + 2] = v;
    }
    // one 32-bit draw per uniform float: the draw converted to float, divided by 2^32
    std::mt19937 rng(42);
    auto uni = [&]() { return (float)rng() / 4294967296.0f; };
    const double two_pi = 2.0 * 3.14159265358979323846;
    for (int64_t i = 0; i < n; i++) {
        const float u = uni(), v = uni(), w = uni();
        quats[4 * i + 0] = (float)(std::sqrt(1.0f - u) * std::sin(two_pi * v));
        quats[4 * i + 1] = (float)(std::sqrt(1.0f - u) * std::cos(two_pi * v));
        quats[4 * i + 2] = (float)(std::sqrt(u) * std::sin(two_pi * w));
        quats[4 * i + 3] = (float)(std::sqrt(u) * std::cos(two_pi * w));
    }
    const double C0 = 0.28209479177387814;  // SH band 0
    for (int64_t i = 0; i < 3 * n; i++) f_dc[i] = (float)(((double)pts.rgb[i] / 255.0 - 0.5) / C0);
    const int nb = (degree + 1) * (degree + 1) - 1;
    if (nb > 0 && f_rest) std::memset(f_rest, 0, (size_t)n * nb * 3 * sizeof(float));
    const float logit = std::log(init_opacity / (1.0f - init_opacity));
    for (int64_t i = 0; i < n; i++) opac[i] = logit;
}

}  // namespace msplat_hip
