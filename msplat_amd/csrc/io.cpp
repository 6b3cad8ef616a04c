// io.cpp — the reference's on-disk formats, byte for byte.
//
// .msplat checkpoint (reference core/src/model.cpp:698-843): u32 magic 0x4C50534D ("MSPL"),
// u32 version 1, u32 step, u32 num_active, u32 sh_degree, u32 adam_step_count,
// f32 adam_lr[6], f32 means_lr_init, f32 means_lr_final, then 18 tensors (the six
// parameter groups, six exp_avg, six exp_avg_sq), each as u32 ndim, i64 shape[ndim],
// u64 byte count, raw little-endian f32 data. Shapes: means/scales/f_dc [N,3], quats [N,4],
// f_rest [N,B-1,3], opacity [N,1].
//
// Gaussian PLY (save_gaussians.cpp:10-88, 141-231): binary_little_endian 1.0 with the
// comment line "Generated by msplat at iteration <step>", per vertex x y z nx ny nz (normals
// 0), f_dc_0..2, f_rest_* channel-major (all coefficients of R, then G, then B), opacity
// (logit), scale_0..2 (log), rot_0..3 (w,x,y,z). Rows with any non-finite value are dropped.
//
// .splat (save_gaussians.cpp:90-139): 32 bytes per Gaussian, sorted by
// (sum of scales) * sigmoid(opacity) descending: f32 position[3], f32 scale[3] (linear),
// u8 rgb from the DC term, u8 alpha, u8 quaternion (q * 128 + 128).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <stdexcept>

#include "model.h"

namespace msplat_hip {

static const uint32_t kMagic = 0x4C50534Du;
static const uint32_t kVersion = 1;
static const double kSHC0 = 0.28209479177387814;

template <class T> static void put(std::ofstream& f, const T& v) { f.write(reinterpret_cast<const char*>(&v), sizeof(T)); }
template <class T> static void get(std::ifstream& f, T& v) {
    f.read(reinterpret_cast<char*>(&v), sizeof(T));
    if (!f) throw std::runtime_error("truncated file");
}

static void put_tensor(std::ofstream& f, const std::vector<int64_t>& shape, const float* data) {
    put(f, (uint32_t)shape.size());
    uint64_t count = 1;
    for (int64_t s : shape) {
        put(f, s);
        count *= (uint64_t)s;
    }
    const uint64_t bytes = count * 4;
    put(f, bytes);
    if (bytes) f.write(reinterpret_cast<const char*>(data), (std::streamsize)bytes);
}

static std::vector<float> get_tensor(std::ifstream& f, std::vector<int64_t>& shape) {
    uint32_t nd = 0;
    get(f, nd);
    if (nd == 0 || nd > 8) throw std::runtime_error("corrupt checkpoint tensor header");
    shape.assign(nd, 0);
    uint64_t count = 1;
    for (uint32_t i = 0; i < nd; i++) {
        get(f, shape[i]);
        if (shape[i] < 0) throw std::runtime_error("corrupt checkpoint tensor shape");
        count *= (uint64_t)shape[i];
    }
    uint64_t bytes = 0;
    get(f, bytes);
    if (bytes != count * 4) throw std::runtime_error("checkpoint tensor byte count does not match its shape");
    std::vector<float> out(count);
    if (bytes) {
        f.read(reinterpret_cast<char*>(out.data()), (std::streamsize)bytes);
        if (!f) throw std::runtime_error("truncated checkpoint");
    }
    return out;
}

static std::vector<int64_t> group_shape(int g, int64_t n, int degree) {
    switch (g) {
        case 0: case 1: case 3: return {n, 3};
        case 2: return {n, 4};
        case 4: return {n, (int64_t)((degree + 1) * (degree + 1) - 1), 3};
        default: return {n, 1};
    }
}

void write_checkpoint(const std::string& path, const CheckpointData& c) {
    std::ofstream f(path, std::ios::binary);
    if (!f.is_open()) throw std::runtime_error("Cannot open checkpoint file for writing: " + path);
    put(f, kMagic);
    put(f, kVersion);
    put(f, c.step);
    put(f, c.num);
    put(f, c.sh_degree);
    put(f, c.adam_steps);
    f.write(reinterpret_cast<const char*>(c.adam_lr), sizeof(c.adam_lr));
    put(f, c.lr_init);
    put(f, c.lr_final);
    const std::vector<float>* groups[6] = {&c.g.means, &c.g.scales, &c.g.quats, &c.g.f_dc, &c.g.f_rest, &c.g.opac};
    for (int g = 0; g < 6; g++) put_tensor(f, group_shape(g, c.g.n, c.g.degree), groups[g]->data());
    for (int g = 0; g < 6; g++) put_tensor(f, group_shape(g, c.g.n, c.g.degree), c.m[g].data());
    for (int g = 0; g < 6; g++) put_tensor(f, group_shape(g, c.g.n, c.g.degree), c.v[g].data());
    f.close();
    if (!f) throw std::runtime_error("failed writing checkpoint: " + path);
}

static int degree_from_bases(int64_t rest_bases) {
    const int64_t b = rest_bases + 1;
    if (b <= 1) return 0;
    if (b <= 4) return 1;
    if (b <= 9) return 2;
    if (b <= 16) return 3;
    return 4;
}

CheckpointData read_checkpoint(const std::string& path) {
    std::ifstream f(path, std::ios::binary);
    if (!f.is_open()) throw std::runtime_error("Cannot open checkpoint file: " + path);
    CheckpointData c;
    uint32_t magic = 0, version = 0;
    get(f, magic);
    get(f, version);
    if (magic != kMagic) throw std::runtime_error("Not a valid msplat checkpoint file");
    if (version != kVersion) throw std::runtime_error("Unsupported checkpoint version: " + std::to_string(version));
    get(f, c.step);
    get(f, c.num);
    get(f, c.sh_degree);
    get(f, c.adam_steps);
    f.read(reinterpret_cast<char*>(c.adam_lr), sizeof(c.adam_lr));
    get(f, c.lr_init);
    get(f, c.lr_final);
    std::vector<int64_t> shp[18];
    std::vector<float> t[18];
    for (int k = 0; k < 18; k++) t[k] = get_tensor(f, shp[k]);
    const int64_t n = shp[0][0];
    if (n != (int64_t)c.num) throw std::runtime_error("checkpoint row count disagrees with its header");
    const int degree = shp[4].size() == 3 ? degree_from_bases(shp[4][1]) : 0;
    if (degree > 3) throw std::runtime_error("checkpoint SH degree above 3 is not supported");
    if ((shp[4].size() == 3 && shp[4][1] != (int64_t)((degree + 1) * (degree + 1) - 1)))
        throw std::runtime_error("checkpoint f_rest has an unsupported basis count");
    for (int k = 0; k < 18; k++) {
        const std::vector<int64_t> want = group_shape(k % 6, n, degree);
        if (shp[k] != want) throw std::runtime_error("checkpoint tensor " + std::to_string(k) + " has an unexpected shape");
    }
    c.g.n = (int)n;
    c.g.degree = degree;
    c.g.means = std::move(t[0]);
    c.g.scales = std::move(t[1]);
    c.g.quats = std::move(t[2]);
    c.g.f_dc = std::move(t[3]);
    c.g.f_rest = std::move(t[4]);
    c.g.opac = std::move(t[5]);
    for (int g = 0; g < 6; g++) {
        c.m[g] = std::move(t[6 + g]);
        c.v[g] = std::move(t[12 + g]);
    }
    return c;
}

static bool row_finite(const HostGaussians& g, int64_t i) {
    const int nr = g.rest_len();
    if (!std::isfinite(g.opac[i])) return false;
    for (int j = 0; j < 3; j++)
        if (!std::isfinite(g.means[i * 3 + j]) || !std::isfinite(g.scales[i * 3 + j])) return false;
    for (int j = 0; j < 4; j++)
        if (!std::isfinite(g.quats[i * 4 + j])) return false;
    for (int j = 0; j < 3; j++)
        if (!std::isfinite(g.f_dc[i * 3 + j])) return false;
    for (int j = 0; j < nr; j++)
        if (!std::isfinite(g.f_rest[i * nr + j])) return false;
    return true;
}

void write_ply(const std::string& path, const HostGaussians& g, int step, bool keep_crs, float scale,
               const float translation[3]) {
    std::ofstream o(path, std::ios::binary);
    if (!o.is_open()) throw std::runtime_error("Cannot open PLY file for writing: " + path);
    const int nr = g.rest_len(), bases = nr / 3;
    std::vector<int64_t> keep;
    keep.reserve(g.n);
    for (int64_t i = 0; i < g.n; i++)
        if (row_finite(g, i)) keep.push_back(i);
    if ((int64_t)keep.size() < g.n)
        std::fprintf(stderr, "savePly: dropped %lld non-finite Gaussian(s) of %lld\n",
                     (long long)(g.n - (int64_t)keep.size()), (long long)g.n);
    std::ostringstream h;
    h << "ply\nformat binary_little_endian 1.0\n";
    h << "comment Generated by msplat at iteration " << step << "\n";
    h << "element vertex " << keep.size() << "\n";
    h << "property float x\nproperty float y\nproperty float z\n";
    h << "property float nx\nproperty float ny\nproperty float nz\n";
    for (int i = 0; i < 3; i++) h << "property float f_dc_" << i << "\n";
    for (int i = 0; i < nr; i++) h << "property float f_rest_" << i << "\n";
    h << "property float opacity\n";
    h << "property float scale_0\nproperty float scale_1\nproperty float scale_2\n";
    h << "property float rot_0\nproperty float rot_1\nproperty float rot_2\nproperty float rot_3\n";
    h << "end_header\n";
    const std::string hs = h.str();
    o.write(hs.data(), (std::streamsize)hs.size());
    const int per_row = 3 + 3 + 3 + nr + 1 + 3 + 4;
    std::vector<float> buf;
    buf.reserve((size_t)per_row * std::min<size_t>(keep.size(), 65536));
    auto flush = [&]() {
        o.write(reinterpret_cast<const char*>(buf.data()), (std::streamsize)(buf.size() * 4));
        buf.clear();
    };
    for (int64_t i : keep) {
        for (int j = 0; j < 3; j++)
            buf.push_back(keep_crs ? (g.means[i * 3 + j] / scale + translation[j]) : g.means[i * 3 + j]);
        buf.push_back(0.f);
        buf.push_back(0.f);
        buf.push_back(0.f);
        for (int j = 0; j < 3; j++) buf.push_back(g.f_dc[i * 3 + j]);
        for (int ch = 0; ch < 3; ch++)
            for (int b = 0; b < bases; b++) buf.push_back(g.f_rest[(size_t)i * nr + b * 3 + ch]);
        buf.push_back(g.opac[i]);
        for (int j = 0; j < 3; j++)
            buf.push_back(keep_crs ? std::log(std::exp(g.scales[i * 3 + j]) / scale) : g.scales[i * 3 + j]);
        for (int j = 0; j < 4; j++) buf.push_back(g.quats[i * 4 + j]);
        if (buf.size() >= (size_t)per_row * 65536) flush();
    }
    flush();
    o.close();
    if (!o) throw std::runtime_error("failed writing PLY: " + path);
}

HostGaussians read_ply(const std::string& path, int* step_out, float scale, const float translation[3], bool keep_crs) {
    std::ifstream f(path, std::ios::binary);
    if (!f.is_open()) throw std::runtime_error("Cannot open PLY file: " + path);
    std::string line;
    std::getline(f, line);
    if (line.find("ply") == std::string::npos) throw std::runtime_error("Not a PLY file: " + path);
    std::getline(f, line);
    if (line.find("binary_little_endian") == std::string::npos)
        throw std::runtime_error("PLY must be binary_little_endian: " + path);
    int num = 0, step = 0, num_dc = 0, num_rest = 0, floats = 0;
    bool ended = false;
    const std::string it_prefix = "comment Generated by msplat at iteration ";
    const std::string legacy_prefix = "comment msplat v";
    const std::string vert_prefix = "element vertex ";
    while (std::getline(f, line)) {
        if (!line.empty() && line.back() == '\r') line.pop_back();
        if (line == "end_header") { ended = true; break; }
        if (line.rfind(it_prefix, 0) == 0) step = std::atoi(line.c_str() + it_prefix.size());
        if (line.rfind(legacy_prefix, 0) == 0) step = std::atoi(line.c_str() + legacy_prefix.size());
        if (line.rfind(vert_prefix, 0) == 0) num = std::atoi(line.c_str() + vert_prefix.size());
        if (line.rfind("property float f_dc_", 0) == 0) num_dc++;
        if (line.rfind("property float f_rest_", 0) == 0) num_rest++;
        if (line.rfind("property float ", 0) == 0) floats++;
        else if (line.rfind("property ", 0) == 0) throw std::runtime_error("PLY property that is not float: " + line);
    }
    if (!ended) throw std::runtime_error("PLY header has no end_header");
    if (num <= 0) throw std::runtime_error("PLY has no vertices");
    if (num_dc != 3) throw std::runtime_error("PLY must hold f_dc_0..2");
    const int bases = num_rest / 3;
    const int degree = degree_from_bases(bases);
    if (num_rest % 3 != 0 || degree > 3 || bases != (degree + 1) * (degree + 1) - 1)
        throw std::runtime_error("PLY f_rest count is not that of SH degree 0..3");
    const int per_row = 3 + 3 + num_dc + num_rest + 1 + 3 + 4;
    if (floats != per_row) throw std::runtime_error("PLY vertex layout is not the 3DGS layout");
    HostGaussians g;
    g.resize(num, degree);
    std::vector<float> row(per_row);
    for (int i = 0; i < num; i++) {
        f.read(reinterpret_cast<char*>(row.data()), per_row * 4);
        if (!f) throw std::runtime_error("truncated PLY: " + path);
        int c = 0;
        for (int j = 0; j < 3; j++) g.means[(size_t)i * 3 + j] = row[c++];
        c += 3;
        for (int j = 0; j < 3; j++) g.f_dc[(size_t)i * 3 + j] = row[c++];
        for (int ch = 0; ch < 3; ch++)
            for (int b = 0; b < bases; b++) g.f_rest[(size_t)i * num_rest + b * 3 + ch] = row[c++];
        g.opac[i] = row[c++];
        for (int j = 0; j < 3; j++) g.scales[(size_t)i * 3 + j] = row[c++];
        for (int j = 0; j < 4; j++) g.quats[(size_t)i * 4 + j] = row[c++];
    }
    if (keep_crs) {
        for (int i = 0; i < num; i++)
            for (int j = 0; j < 3; j++) g.means[(size_t)i * 3 + j] = (g.means[(size_t)i * 3 + j] - translation[j]) * scale;
        for (size_t i = 0; i < g.scales.size(); i++) g.scales[i] = std::log(scale * std::exp(g.scales[i]));
    }
    if (step_out) *step_out = step;
    return g;
}

void write_splat(const std::string& path, const HostGaussians& g, bool keep_crs, float scale, const float translation[3]) {
    std::ofstream o(path, std::ios::binary);
    if (!o.is_open()) throw std::runtime_error("Cannot open splat file for writing: " + path);
    std::vector<float> order(g.n, 0.f);
    std::vector<int64_t> idx;
    idx.reserve(g.n);
    for (int64_t i = 0; i < g.n; i++) {
        bool fin = std::isfinite(g.opac[i]);
        for (int j = 0; fin && j < 3; j++) fin = std::isfinite(g.means[i * 3 + j]) && std::isfinite(g.scales[i * 3 + j]);
        for (int j = 0; fin && j < 4; j++) fin = std::isfinite(g.quats[i * 4 + j]);
        if (!fin) continue;
        float s = std::exp(g.scales[i * 3]) + std::exp(g.scales[i * 3 + 1]) + std::exp(g.scales[i * 3 + 2]);
        if (keep_crs) s /= scale;
        order[i] = s / (1.0f + std::exp(-g.opac[i]));
        idx.push_back(i);
    }
    // descending; ties keep index order so the file is a function of the Gaussians alone
    std::stable_sort(idx.begin(), idx.end(), [&](int64_t a, int64_t b) { return order[a] > order[b]; });
    std::vector<uint8_t> buf;
    buf.reserve(idx.size() * 32);
    for (int64_t i : idx) {
        float m[3], sc[3];
        for (int j = 0; j < 3; j++) m[j] = keep_crs ? (g.means[i * 3 + j] / scale + translation[j]) : g.means[i * 3 + j];
        for (int j = 0; j < 3; j++) sc[j] = keep_crs ? (std::exp(g.scales[i * 3 + j]) / scale) : std::exp(g.scales[i * 3 + j]);
        const uint8_t* mb = reinterpret_cast<const uint8_t*>(m);
        const uint8_t* sb = reinterpret_cast<const uint8_t*>(sc);
        buf.insert(buf.end(), mb, mb + 12);
        buf.insert(buf.end(), sb, sb + 12);
        for (int j = 0; j < 3; j++)
            buf.push_back((uint8_t)std::clamp(((double)g.f_dc[i * 3 + j] * kSHC0 + 0.5) * 255.0, 0.0, 255.0));
        const float sig = 1.0f / (1.0f + std::exp(-g.opac[i]));
        buf.push_back((uint8_t)std::clamp(sig * 255.0f, 0.0f, 255.0f));
        for (int j = 0; j < 4; j++) buf.push_back((uint8_t)std::clamp(g.quats[i * 4 + j] * 128.0f + 128.0f, 0.0f, 255.0f));
    }
    o.write(reinterpret_cast<const char*>(buf.data()), (std::streamsize)buf.size());
    o.close();
    if (!o) throw std::runtime_error("failed writing splat: " + path);
}

}  // namespace msplat_hip
