// api.cpp — the msplat C API (include/msplat_c_api.h) over the device-resident Model.
// Reference: core/src/msplat_api.mm (Dataset 20-56, Trainer 58-290, C wrappers 292-397,
// viewer 399-720). Every entry point catches C++ exceptions and reports failure through its
// return value, so nothing unwinds across the C boundary.
#include <chrono>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/msplat_api_ext.h"
#include "../../include/msplat_c_api.h"
#include "eval.h"
#include "image_io.h"
#include "ingest.h"
#include "model.h"

namespace msplat_hip {

// ---------------------------------------------------------------------------------------
// Dataset

struct DatasetImpl {
    std::vector<TrainCamera> train, test;
    std::vector<std::string> train_rel, test_rel;  // names relative to the image folder (masks)
    std::vector<std::unique_ptr<DBuf>> gt;  // owns every camera's device ground truth
    HostGaussians init;
    bool has_init = false;
    float scale = 1.0f, translation[3] = {0.f, 0.f, 0.f};
    ~DatasetImpl() {
        (void)hipDeviceSynchronize();
        for (auto& b : gt) b->release();
    }
    int add_camera(const float c2w[16], float fx, float fy, float cx, float cy, int w, int h, const uint8_t* rgb_host,
                   const uint8_t* rgb_dev, bool is_test) {
        if (w <= 0 || h <= 0 || !(fx > 0.f) || !(fy > 0.f)) throw std::invalid_argument("bad camera intrinsics");
        TrainCamera c;
        std::memcpy(c.c2w, c2w, sizeof(c.c2w));
        c.fx = fx; c.fy = fy; c.cx = cx; c.cy = cy; c.width = w; c.height = h;
        const size_t bytes = (size_t)w * h * 3;
        gt.emplace_back(new DBuf());
        gt.back()->ensure(bytes);
        if (rgb_host) hip_check(hipMemcpy(gt.back()->p, rgb_host, bytes, hipMemcpyHostToDevice), "ground truth upload");
        else hip_check(hipMemcpy(gt.back()->p, rgb_dev, bytes, hipMemcpyDeviceToDevice), "ground truth copy");
        c.gt = gt.back()->as<uint8_t>();
        auto& list = is_test ? test : train;
        list.push_back(c);
        (is_test ? test_rel : train_rel).push_back(std::string());
        return (int)list.size() - 1;
    }
    // Training mask of a camera from host uint8 [h w] values (white = use, black = ignore),
    // binarised as Camera::loadMask does (reference input_data.cpp:81-111: a value above
    // one half keeps the pixel, stored as 255 / 0); an all-white mask is dropped. Returns
    // the count of kept pixels.
    int64_t set_mask(int index, bool is_test, const uint8_t* m) {
        auto& list = is_test ? test : train;
        if (index < 0 || index >= (int)list.size()) throw std::invalid_argument("camera index out of range");
        TrainCamera& c = list[(size_t)index];
        const size_t n = (size_t)c.width * c.height;
        std::vector<uint8_t> keep(n);
        int64_t kept = 0;
        for (size_t i = 0; i < n; i++) {
            const bool k = m[i] >= 128;  // v / 255 > 0.5
            keep[i] = k ? 255 : 0;
            kept += k ? 1 : 0;
        }
        if (kept == (int64_t)n) {
            c.mask = nullptr;
            c.unmasked = -1;
            return kept;
        }
        gt.emplace_back(new DBuf());
        gt.back()->ensure(n);
        hip_check(hipMemcpy(gt.back()->p, keep.data(), n, hipMemcpyHostToDevice), "mask upload");
        c.mask = gt.back()->as<uint8_t>();
        c.unmasked = kept;
        return kept;
    }
};

static HostGaussians synth_set(int kind, uint64_t seed, int n, int degree) {
    HostGaussians g;
    g.resize(n, degree);
    if (msplat_synth_gaussians(kind, seed, n, degree, g.means.data(), g.scales.data(), g.quats.data(), g.opac.data(),
                               g.f_dc.data(), g.f_rest.data()) != 0)
        throw std::invalid_argument("bad synthetic Gaussian arguments");
    return g;
}

// "synthetic:<preset>[:key=value,...]"
static DatasetImpl* synthetic_dataset(const std::string& spec, float downscale, bool eval_mode, int test_every) {
    std::string rest = spec.substr(std::string("synthetic:").size());
    std::string preset = rest, opts;
    const size_t colon = rest.find(':');
    if (colon != std::string::npos) {
        preset = rest.substr(0, colon);
        opts = rest.substr(colon + 1);
    }
    std::map<std::string, double> kv;
    if (preset == "c0") {
        kv = {{"kind", 0}, {"n", 50000}, {"hidden", 20000}, {"cams", 16}, {"w", 256}, {"h", 256}, {"f", 256},
              {"jitter", 0.0}};
    } else if (preset == "c1") {
        kv = {{"kind", 1}, {"n", 1500000}, {"hidden", 20000}, {"cams", 64}, {"w", 1600}, {"h", 1040}, {"f", 1300},
              {"jitter", 0.15}};
    } else {
        throw std::invalid_argument("unknown synthetic preset '" + preset + "' (use c0 or c1)");
    }
    kv["seed"] = 1; kv["hidden_seed"] = 2; kv["cam_seed"] = 5; kv["radius"] = 3.0; kv["degree"] = 3;
    kv["mixed"] = 0;  // 1: each camera's size drawn from 640x480 / 1024x768 / 1600x1040 / 1920x1080
    std::stringstream ss(opts);
    std::string item;
    while (std::getline(ss, item, ',')) {
        if (item.empty()) continue;
        const size_t eq = item.find('=');
        if (eq == std::string::npos) throw std::invalid_argument("bad synthetic option '" + item + "'");
        const std::string k = item.substr(0, eq);
        if (!kv.count(k)) throw std::invalid_argument("unknown synthetic option '" + k + "'");
        kv[k] = std::stod(item.substr(eq + 1));
    }
    const int n = (int)kv["n"], hidden = (int)kv["hidden"], cams = (int)kv["cams"], kind = (int)kv["kind"];
    int W = (int)kv["w"], H = (int)kv["h"];
    float f = (float)kv["f"];
    const int degree = (int)kv["degree"];
    if (n <= 0 || hidden <= 0 || cams <= 0 || W <= 0 || H <= 0 || !(f > 0.f) || degree < 0 || degree > 3)
        throw std::invalid_argument("bad synthetic sizes");
    float cx = (float)W / 2.0f, cy = (float)H / 2.0f;
    if (downscale > 1.0f) {
        W = (int)((float)W / downscale);
        H = (int)((float)H / downscale);
        f /= downscale; cx /= downscale; cy /= downscale;
    }
    std::unique_ptr<DatasetImpl> ds(new DatasetImpl());
    std::vector<float> c2w((size_t)cams * 16);
    if (msplat_synth_cameras((uint64_t)kv["cam_seed"], cams, (float)kv["radius"], (float)kv["jitter"], c2w.data()) != 0)
        throw std::invalid_argument("bad synthetic camera arguments");
    // ground truth: the hidden set rendered at SH degree 3 over the default background
    {
        ModelOptions go;
        go.sh_degree = 3;
        go.sh_degree_interval = 1;
        go.cap_max = hidden;
        go.decimate_seeds = false;
        const float zero[3] = {0.f, 0.f, 0.f};
        Model hid(go, synth_set(kind, (uint64_t)kv["hidden_seed"], hidden, 3), 1.0f, zero);
        std::vector<float> img((size_t)W * H * 3);
        std::vector<uint8_t> q(img.size());
        static const int kSizes[4][2] = {{640, 480}, {1024, 768}, {1600, 1040}, {1920, 1080}};
        uint64_t size_rng = (uint64_t)kv["cam_seed"] * 0x2545F4914F6CDD1Dull + 17;
        for (int i = 0; i < cams; i++) {
            TrainCamera c;
            std::memcpy(c.c2w, &c2w[(size_t)i * 16], sizeof(c.c2w));
            int w = W, h = H;
            float fi = f, cxi = cx, cyi = cy;
            if (kv["mixed"] != 0.0) {
                uint64_t x = (size_rng += 0x9E3779B97F4A7C15ull);
                x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
                x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
                x ^= x >> 31;
                const int k = (int)(x % 4);
                w = (int)((float)kSizes[k][0] / std::max(downscale, 1.0f));
                h = (int)((float)kSizes[k][1] / std::max(downscale, 1.0f));
                fi = f * (float)w / (float)W;  // same horizontal field of view as the base size
                cxi = (float)w / 2.0f;
                cyi = (float)h / 2.0f;
            }
            c.fx = fi; c.fy = fi; c.cx = cxi; c.cy = cyi; c.width = w; c.height = h;
            img.resize((size_t)w * h * 3);
            q.resize(img.size());
            const float* d = hid.render(c, 1 << 20);
            hip_check(hipStreamSynchronize(hid.engine().stream()), "render");
            hip_check(hipMemcpy(img.data(), d, img.size() * 4, hipMemcpyDeviceToHost), "download");
            msplat_quantize_rgb(img.data(), (int64_t)img.size(), q.data());
            const bool is_test = eval_mode && test_every > 0 && (i % test_every == 0);
            ds->add_camera(c.c2w, fi, fi, cxi, cyi, w, h, q.data(), nullptr, is_test);
        }
    }
    ds->init = synth_set(kind, (uint64_t)kv["seed"], n, degree);
    ds->has_init = true;
    return ds.release();
}

// ---------------------------------------------------------------------------------------
// Trainer

static uint64_t split_mix(uint64_t& s) {
    uint64_t x = (s += 0x9E3779B97F4A7C15ull);
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

static HostGaussians with_degree(const HostGaussians& g, int degree) {
    if (g.degree == degree) return g;
    HostGaussians o = g;
    o.degree = degree;
    const int src_b = g.rest_len() / 3, dst_b = o.rest_len() / 3, keep = std::min(src_b, dst_b);
    o.f_rest.assign((size_t)g.n * dst_b * 3, 0.f);
    for (int i = 0; i < g.n; i++)
        std::memcpy(&o.f_rest[(size_t)i * dst_b * 3], &g.f_rest[(size_t)i * src_b * 3], (size_t)keep * 3 * sizeof(float));
    return o;
}

struct TrainerImpl {
    DatasetImpl* ds = nullptr;
    MsplatConfig cfg;
    std::unique_ptr<Model> model;
    int step = 0;
    std::vector<int> order;
    size_t pos = 0;
    uint64_t rng = 42;  // camera order: splitmix64 Fisher-Yates (reference: mt19937(42) + std::shuffle)
    int last_cam = -1;
    std::vector<int> last_batch;
    MsplatEngine handle{nullptr, false};
    DBuf rgba, partials, sums;
    // judge of the learned relocation proposal: held-out loss change over the interval that
    // follows an event, per mode (0 plain, 1 learned); the learned mode is used only while
    // its change is no worse than the plain one's, re-probed after an exponential back-off
    struct Judge {
        int force = 0;          // 0 judge, 1 always plain, 2 always learned
        int prev_mode = -1;
        double prev_loss = 0.0;
        double delta[2] = {NAN, NAN};
        int wait = 0, backoff = 1, plain_run = 0;
    } judge;

    void shuffle() {
        for (int i = (int)order.size() - 1; i > 0; i--) {
            const int j = (int)(split_mix(rng) % (uint64_t)(i + 1));
            std::swap(order[i], order[j]);
        }
        pos = 0;
    }
    int next_camera() {
        if (pos >= order.size()) shuffle();
        return order[pos++];
    }
};

static float window_1d[11];
static void ssim_window(float out[11]) {
    // createSSIMWindow's 1-D factor (reference core/src/ssim.cpp:7-24), sigma 1.5
    float sum = 0.f;
    for (int i = 0; i < 11; i++) {
        const float x = (float)(i - 5);
        out[i] = std::exp(-(x * x) / (2.0f * 1.5f * 1.5f));
        sum += out[i];
    }
    for (int i = 0; i < 11; i++) out[i] /= sum;
}

static void metrics_for(TrainerImpl* T, const TrainCamera& cam, const float* img, double out3[3]) {
    hipStream_t s = T->model->engine().stream();
    T->partials.ensure(msk::eval_partials_bytes(cam.width, cam.height));
    T->sums.ensure(64);
    ssim_window(window_1d);
    msk::eval_sums(s, img, cam.gt, cam.width, cam.height, window_1d, T->partials.as<double>(), T->sums.as<double>());
    hip_check(hipMemcpyAsync(out3, T->sums.p, 24, hipMemcpyDeviceToHost, s), "metrics");
    hip_check(hipStreamSynchronize(s), "metrics");
}

// Held-out objective: mean over the test cameras (or, without a test split, 8 evenly spaced
// training cameras) of (1 - w) L1 + w (1 - SSIM) of the current state.
static double heldout_loss(TrainerImpl* T) {
    std::vector<const TrainCamera*> cams;
    if (!T->ds->test.empty()) {
        for (const TrainCamera& c : T->ds->test) cams.push_back(&c);
    } else {
        const int n = (int)T->ds->train.size(), k = std::min(n, 8);
        for (int i = 0; i < k; i++) cams.push_back(&T->ds->train[(size_t)i * n / k]);
    }
    double acc = 0.0;
    for (const TrainCamera* c : cams) {
        const float* img = T->model->render(*c, T->step);
        double s3[3];
        metrics_for(T, *c, img, s3);
        const double n = (double)c->width * c->height * 3;
        const double w = T->cfg.ssimWeight;
        acc += (1.0 - w) * (s3[1] / n) + w * (1.0 - s3[2] / n);
    }
    return cams.empty() ? 0.0 : acc / (double)cams.size();
}

// Chooses the relocation mode of the event at T->step.
static void judge_event(TrainerImpl* T) {
    TrainerImpl::Judge& J = T->judge;
    const double L = heldout_loss(T);
    if (J.prev_mode >= 0) J.delta[J.prev_mode] = L - J.prev_loss;
    bool learned = false;
    if (J.force == 2) {
        learned = true;
    } else if (J.force == 0 && T->model->proposal_ready()) {
        if (J.wait > 0) {
            J.wait--;
        } else if (std::isnan(J.delta[1])) {
            learned = true;                          // untried (or due for a re-probe)
        } else if (std::isnan(J.delta[0]) || J.delta[1] <= J.delta[0]) {
            learned = true;                          // no worse than plain
            J.backoff = 1;
            if (++J.plain_run >= 4) {                // refresh the plain figure now and then
                learned = false;
                J.plain_run = 0;
            }
        } else {                                     // worse: plain, re-probe after a back-off
            J.wait = J.backoff;
            J.backoff = std::min(J.backoff * 2, 16);
            J.delta[1] = NAN;
        }
    }
    T->model->set_proposal_mode(learned);
    J.prev_mode = learned ? 1 : 0;
    J.prev_loss = L;
}

static MsplatPixelBuffer host_image(TrainerImpl* T, const float* dimg, int W, int H) {
    MsplatPixelBuffer b{nullptr, 0, 0};
    const size_t bytes = (size_t)W * H * 3 * sizeof(float);
    float* buf = (float*)std::malloc(bytes);
    if (!buf) throw std::bad_alloc();
    hipStream_t s = T->model->engine().stream();
    hipError_t e = hipMemcpyAsync(buf, dimg, bytes, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) {
        std::free(buf);
        hip_check(e, "image download");
    }
    b.data = buf;
    b.width = W;
    b.height = H;
    return b;
}

// ---------------------------------------------------------------------------------------
// Viewer

struct ViewerImpl {
    int width = 0, height = 0;
    float fov_y = 0.f;
    float bg[3] = {0.f, 0.f, 0.f};
    int degree = 0, n = 0;
    DBuf p[6];
    Engine eng;
    DBuf rgba;
    float bmin[3] = {0, 0, 0}, bmax[3] = {0, 0, 0}, dcenter[3] = {0, 0, 0};

    ~ViewerImpl() {
        (void)hipStreamSynchronize(eng.stream());
        for (auto& b : p) b.release();
        rgba.release();
    }
    MsplatDeviceState state() const {
        MsplatDeviceState st;
        std::memset(&st, 0, sizeof(st));
        st.N = n;
        st.degree = degree;
        st.means = p[0].as<float>(); st.scales = p[1].as<float>(); st.quats = p[2].as<float>();
        st.f_dc = p[3].as<float>(); st.f_rest = p[4].as<float>(); st.opac = p[5].as<float>();
        return st;
    }
    void bounds(const HostGaussians& g) {
        if (g.n == 0) return;
        for (int j = 0; j < 3; j++) bmin[j] = bmax[j] = g.means[j];
        for (int i = 1; i < g.n; i++)
            for (int j = 0; j < 3; j++) {
                const float v = g.means[(size_t)i * 3 + j];
                if (v < bmin[j]) bmin[j] = v;
                if (v > bmax[j]) bmax[j] = v;
            }
        // densest cell of an 8x8x8 grid (reference msplat_api.mm:453-484)
        const int G = 8;
        std::vector<int> grid(G * G * G, 0);
        float ext[3];
        for (int j = 0; j < 3; j++) ext[j] = std::max(bmax[j] - bmin[j], 1e-6f);
        for (int i = 0; i < g.n; i++) {
            int c[3];
            for (int j = 0; j < 3; j++)
                c[j] = std::min((int)((g.means[(size_t)i * 3 + j] - bmin[j]) / ext[j] * G), G - 1);
            grid[c[0] * G * G + c[1] * G + c[2]]++;
        }
        int best = 0, best_idx = 0;
        for (int i = 0; i < G * G * G; i++)
            if (grid[i] > best) { best = grid[i]; best_idx = i; }
        const int ix = best_idx / (G * G), iy = (best_idx / G) % G, iz = best_idx % G;
        dcenter[0] = bmin[0] + (ix + 0.5f) / G * ext[0];
        dcenter[1] = bmin[1] + (iy + 0.5f) / G * ext[1];
        dcenter[2] = bmin[2] + (iz + 0.5f) / G * ext[2];
    }
    const float* render_device(const float c2w[16]) {
        MsplatCamera cam;
        msplat_camera_from_fov(c2w, fov_y, width, height, &cam);
        const int rc = eng.render(state(), cam, degree, bg);
        if (rc != 0) throw std::runtime_error("viewer render rejected its inputs");
        return eng.image();
    }
};

}  // namespace msplat_hip

// ---------------------------------------------------------------------------------------
// C entry points

using namespace msplat_hip;

#define MS_GUARD_BEGIN try {
#define MS_GUARD_END(ret)                                              \
    }                                                                  \
    catch (const std::exception& ex) {                                 \
        std::fprintf(stderr, "msplat: %s: %s\n", __func__, ex.what()); \
        return ret;                                                    \
    }

static TrainerImpl* TR(MsplatTrainer t) { return static_cast<TrainerImpl*>(t); }
static DatasetImpl* DS(MsplatDataset d) { return static_cast<DatasetImpl*>(d); }
static ViewerImpl* VW(MsplatViewer v) { return static_cast<ViewerImpl*>(v); }

extern "C" {

// A scene folder (Dataset::Dataset, msplat_api.mm:25-41): the format's loader, every
// camera's image as Camera::loadImage prepares it (uint8 ground truth on the device), the
// train / held-out split (every testEvery-th camera held out in eval mode, input_data.cpp:
// 196-205) and the initial Gaussians Model::Model seeds from the point cloud (kNN scales
// x 0.1, opacity 0.5, SH degree 3: the trainer adapts the degree).
static DatasetImpl* folder_dataset(const std::string& path, float downscale, bool eval_mode, int test_every) {
    Scene sc = load_scene(path);
    if (sc.cams.empty()) throw std::runtime_error("no cameras in " + path);
    if (sc.points.count() == 0) throw std::runtime_error("no point cloud in " + path + " to seed the Gaussians from");
    if (eval_mode && test_every <= 0) throw std::invalid_argument("test_every must be positive");
    std::unique_ptr<DatasetImpl> ds(new DatasetImpl());
    for (size_t i = 0; i < sc.cams.size(); i++) {
        SceneCamera& c = sc.cams[i];
        const std::vector<uint8_t> rgb = load_camera_image(c, downscale);
        const bool held_out = eval_mode && (int)(i % (size_t)test_every) == 0;
        const int k = ds->add_camera(c.c2w, c.fx, c.fy, c.cx, c.cy, c.width, c.height, rgb.data(), nullptr, held_out);
        (held_out ? ds->test_rel : ds->train_rel)[(size_t)k] = c.relative;
    }
    const int64_t n = sc.points.count();
    if (n > INT32_MAX) throw std::runtime_error("point cloud too large");
    HostGaussians g;
    g.resize((int)n, 3);
    seed_gaussians(sc.points, 3, 0.1f, 0.5f, g.means.data(), g.scales.data(), g.quats.data(), g.opac.data(),
                   g.f_dc.data(), g.f_rest.data());
    ds->init = std::move(g);
    ds->has_init = true;
    ds->scale = sc.scale;
    for (int k = 0; k < 3; k++) ds->translation[k] = sc.translation[k];
    return ds.release();
}

MsplatDataset msplat_dataset_create(const char* path, float downscaleFactor, bool evalMode, int testEvery) {
    MS_GUARD_BEGIN
    if (!path) return nullptr;
    const std::string p(path);
    if (p.rfind("synthetic:", 0) == 0) return synthetic_dataset(p, downscaleFactor, evalMode, testEvery);
    return folder_dataset(p, downscaleFactor, evalMode, testEvery);
    MS_GUARD_END(nullptr)
}

// Camera::loadMask for every camera with a name relative to its image folder (COLMAP).
int64_t msplat_dataset_load_masks(MsplatDataset ds, const char* mask_dir) {
    MS_GUARD_BEGIN
    if (!ds || !mask_dir) return -1;
    DatasetImpl* d = DS(ds);
    int64_t masked = 0;
    for (int t = 0; t < 2; t++) {
        auto& list = t ? d->test : d->train;
        auto& rel = t ? d->test_rel : d->train_rel;
        for (size_t i = 0; i < list.size(); i++) {
            SceneCamera c;
            c.relative = rel[i];
            c.width = list[i].width;
            c.height = list[i].height;
            std::vector<uint8_t> m;
            int64_t kept = 0;
            if (!load_camera_mask(c, mask_dir, m, kept)) continue;
            if (d->set_mask((int)i, t == 1, m.data()) < (int64_t)m.size()) masked++;
        }
    }
    return masked;
    MS_GUARD_END(-1)
}

MsplatDataset msplat_dataset_create_empty(void) {
    MS_GUARD_BEGIN
    return new DatasetImpl();
    MS_GUARD_END(nullptr)
}

void msplat_dataset_destroy(MsplatDataset ds) { delete DS(ds); }

int msplat_dataset_num_train(MsplatDataset ds) { return ds ? (int)DS(ds)->train.size() : 0; }
int msplat_dataset_num_test(MsplatDataset ds) { return ds ? (int)DS(ds)->test.size() : 0; }

void msplat_dataset_camera_pose(MsplatDataset ds, int cameraIndex, float camToWorld[16]) {
    if (!ds || !camToWorld) return;
    auto& tr = DS(ds)->train;
    if (cameraIndex >= 0 && cameraIndex < (int)tr.size()) std::memcpy(camToWorld, tr[cameraIndex].c2w, 16 * sizeof(float));
}

int msplat_dataset_camera_info(MsplatDataset ds, int index, bool test, float c2w[16], float intrinsics[4], int* width,
                               int* height, uint8_t* rgb) {
    if (!ds) return -1;
    auto& list = test ? DS(ds)->test : DS(ds)->train;
    if (index < 0 || index >= (int)list.size()) return -1;
    const TrainCamera& c = list[index];
    if (c2w) std::memcpy(c2w, c.c2w, sizeof(c.c2w));
    if (intrinsics) {
        intrinsics[0] = c.fx; intrinsics[1] = c.fy; intrinsics[2] = c.cx; intrinsics[3] = c.cy;
    }
    if (width) *width = c.width;
    if (height) *height = c.height;
    if (rgb && hipMemcpy(rgb, c.gt, (size_t)c.width * c.height * 3, hipMemcpyDeviceToHost) != hipSuccess) return -1;
    return 0;
}

int msplat_dataset_add_camera(MsplatDataset ds, const float c2w[16], float fx, float fy, float cx, float cy, int width,
                              int height, const uint8_t* rgb, bool test) {
    MS_GUARD_BEGIN
    if (!ds || !c2w || !rgb) return -1;
    return DS(ds)->add_camera(c2w, fx, fy, cx, cy, width, height, rgb, nullptr, test);
    MS_GUARD_END(-1)
}

int64_t msplat_dataset_set_mask(MsplatDataset ds, int index, bool test, const uint8_t* mask) {
    MS_GUARD_BEGIN
    if (!ds || !mask) return -1;
    return DS(ds)->set_mask(index, test, mask);
    MS_GUARD_END(-1)
}

int msplat_dataset_set_init_gaussians(MsplatDataset ds, int n, int degree, const float* means, const float* log_scales,
                                      const float* quats, const float* opac_logits, const float* f_dc,
                                      const float* f_rest) {
    MS_GUARD_BEGIN
    if (!ds || n <= 0 || degree < 0 || degree > 3 || !means || !log_scales || !quats || !opac_logits || !f_dc) return -1;
    HostGaussians g;
    g.resize(n, degree);
    if (g.rest_len() > 0 && !f_rest) return -1;
    std::memcpy(g.means.data(), means, g.means.size() * 4);
    std::memcpy(g.scales.data(), log_scales, g.scales.size() * 4);
    std::memcpy(g.quats.data(), quats, g.quats.size() * 4);
    std::memcpy(g.opac.data(), opac_logits, g.opac.size() * 4);
    std::memcpy(g.f_dc.data(), f_dc, g.f_dc.size() * 4);
    if (g.rest_len() > 0) std::memcpy(g.f_rest.data(), f_rest, g.f_rest.size() * 4);
    DS(ds)->init = std::move(g);
    DS(ds)->has_init = true;
    return 0;
    MS_GUARD_END(-1)
}

MsplatTrainer msplat_trainer_create(MsplatDataset ds, MsplatConfig config) {
    MS_GUARD_BEGIN
    if (!ds) return nullptr;
    DatasetImpl* d = DS(ds);
    if (!d->has_init) throw std::runtime_error("dataset holds no initial Gaussians");
    if (d->train.empty()) throw std::runtime_error("dataset holds no training cameras");
    std::unique_ptr<TrainerImpl> T(new TrainerImpl());
    T->ds = d;
    T->cfg = config;
    ModelOptions o;
    o.sh_degree = config.shDegree;
    o.sh_degree_interval = config.shDegreeInterval;
    o.cap_max = config.capMax;
    o.max_steps = config.iterations;
    o.noise_lr = config.noiseLr;
    o.opacity_reg = config.opacityReg;
    o.scale_reg = config.scaleReg;
    o.cull_radius = config.cullRadius;
    o.keep_crs = config.keepCrs;
    for (int c = 0; c < 3; c++) o.bg[c] = config.bgColor[c];
    T->model.reset(new Model(o, with_degree(d->init, config.shDegree), d->scale, d->translation));
    T->order.resize(d->train.size());
    for (size_t i = 0; i < T->order.size(); i++) T->order[i] = (int)i;
    T->shuffle();
    T->handle.e = &T->model->engine();
    T->handle.owned = false;
    return T.release();
    MS_GUARD_END(nullptr)
}

void msplat_trainer_destroy(MsplatTrainer t) {
    if (!t) return;
    TrainerImpl* T = TR(t);
    T->rgba.release();
    T->partials.release();
    T->sums.release();
    delete T;
}

MsplatStats msplat_trainer_step(MsplatTrainer t) {
    MsplatStats s{0, 0, 0.f};
    if (!t) return s;
    TrainerImpl* T = TR(t);
    MS_GUARD_BEGIN
    T->step++;
    const int ci = T->next_camera();
    T->last_cam = ci;
    const TrainCamera& cam = T->ds->train[ci];
    const auto t0 = std::chrono::high_resolution_clock::now();
    T->model->full_iteration(cam, T->step, T->cfg.ssimWeight);
    T->model->schedulers_step(T->step);
    if (T->model->proposal_enabled() && T->model->is_mcmc_step(T->step)) judge_event(T);
    T->model->mcmc_after_train(T->step);
    const auto t1 = std::chrono::high_resolution_clock::now();
    s.iteration = T->step;
    s.splatCount = T->model->num_active();
    s.msPerStep = (float)(std::chrono::duration_cast<std::chrono::microseconds>(t1 - t0).count() / 1000.0);
    return s;
    MS_GUARD_END(s)
}

MsplatStats msplat_trainer_step_batch(MsplatTrainer t, int views) {
    MsplatStats s{0, 0, 0.f};
    if (!t || views <= 0) return s;
    TrainerImpl* T = TR(t);
    MS_GUARD_BEGIN
    const auto t0 = std::chrono::high_resolution_clock::now();
    T->step++;
    T->last_batch.clear();
    std::vector<const TrainCamera*> cams;
    for (int v = 0; v < views; v++) {
        const int ci = T->next_camera();
        T->last_batch.push_back(ci);
        cams.push_back(&T->ds->train[ci]);
    }
    T->last_cam = T->last_batch.back();
    T->model->full_iteration_batch(cams.data(), views, T->step, T->cfg.ssimWeight);
    T->model->schedulers_step(T->step);
    if (T->model->proposal_enabled() && T->model->is_mcmc_step(T->step)) judge_event(T);
    T->model->mcmc_after_train(T->step);
    const auto t1 = std::chrono::high_resolution_clock::now();
    s.iteration = T->step;
    s.splatCount = T->model->num_active();
    s.msPerStep = (float)(std::chrono::duration_cast<std::chrono::microseconds>(t1 - t0).count() / 1000.0);
    return s;
    MS_GUARD_END(s)
}

int msplat_trainer_last_batch(MsplatTrainer t, int32_t* cameras, double* losses, int max_views) {
    if (!t) return -1;
    TrainerImpl* T = TR(t);
    MS_GUARD_BEGIN
    const int n = (int)T->last_batch.size();
    std::vector<double> l;
    if (losses) l = T->model->batch_losses();
    for (int v = 0; v < n && v < max_views; v++) {
        if (cameras) cameras[v] = T->last_batch[v];
        if (losses && v < (int)l.size()) losses[v] = l[v];
    }
    return n;
    MS_GUARD_END(-1)
}

int msplat_trainer_proposal_info(MsplatTrainer t, double out[12]) {
    if (!t || !out) return -1;
    TrainerImpl* T = TR(t);
    const msplat_hip::ProposalInfo& p = T->model->proposal_info();
    const double v[12] = {(double)p.events, (double)p.learned_events, (double)p.plain_events, (double)p.samples,
                          p.val_bce, p.base_bce, p.val_off, p.base_off, (double)p.trained, T->judge.delta[0],
                          T->judge.delta[1], T->judge.prev_loss};
    for (int k = 0; k < 12; k++) out[k] = v[k];
    return T->model->proposal_enabled() ? 0 : -1;
}

int msplat_trainer_proposal_params(MsplatTrainer t, float* theta, int set) {
    if (!t || !theta) return -1;
    MS_GUARD_BEGIN
    if (set) TR(t)->model->set_proposal_params(theta);
    else TR(t)->model->proposal_params(theta);
    return 0;
    MS_GUARD_END(-1)
}

int64_t msplat_trainer_proposal_read(MsplatTrainer t, const char* name, void* host, int64_t bytes) {
    if (!t || !name) return -1;
    MS_GUARD_BEGIN
    return TR(t)->model->proposal_read(name, host, bytes);
    MS_GUARD_END(-1)
}

int msplat_trainer_batch_info(MsplatTrainer t, int64_t out[4]) {
    if (!t || !out) return -1;
    msplat_hip::BatchRunner* b = TR(t)->model->batch_runner();
    out[0] = b ? b->captures() : 0;
    out[1] = b ? b->replays() : 0;
    out[2] = b ? b->launch_geom().W : 0;
    out[3] = b ? b->launch_geom().H : 0;
    return 0;
}

int msplat_trainer_autotune_info(MsplatTrainer t, int32_t* cands, float* ms, int max, int32_t chosen[5]) {
    if (!t) return -1;
    msplat_hip::BatchRunner* b = TR(t)->model->batch_runner();
    if (chosen)
        for (int j = 0; j < 5; j++) chosen[j] = 0;
    if (!b) return 0;
    const auto& r = b->tune_results();
    for (int k = 0; k < (int)r.size() && k < max; k++) {
        if (cands) {
            cands[5 * k + 0] = r[k].c.ppl;
            cands[5 * k + 1] = r[k].c.grid;
            cands[5 * k + 2] = r[k].c.chunks;
            cands[5 * k + 3] = r[k].c.waves;
            cands[5 * k + 4] = r[k].c.graph;
        }
        if (ms) ms[k] = r[k].ms;
    }
    if (chosen && b->tuned()) {
        const auto c = b->choice();
        chosen[0] = c.ppl;
        chosen[1] = c.grid;
        chosen[2] = c.chunks;
        chosen[3] = c.waves;
        chosen[4] = c.graph;
    }
    return (int)r.size();
}

void msplat_trainer_train(MsplatTrainer t) {
    if (!t) return;
    while (TR(t)->step < TR(t)->cfg.iterations) {
        const MsplatStats s = msplat_trainer_step(t);
        if (s.iteration == 0) return;  // the step failed (reported on stderr)
    }
}

MsplatEvalMetrics msplat_trainer_evaluate(MsplatTrainer t) {
    MsplatEvalMetrics m{0.f, 0.f, 0.f, 0, 0};
    if (!t) return m;
    TrainerImpl* T = TR(t);
    MS_GUARD_BEGIN
    auto& cams = T->ds->test;
    if (cams.empty()) return m;
    double sp = 0, ss = 0, sl = 0;
    for (const TrainCamera& cam : cams) {
        const float* img = T->model->render(cam, T->cfg.iterations);
        double s3[3];
        metrics_for(T, cam, img, s3);
        const double n = (double)cam.width * cam.height * 3;
        const float psnr = 10.0f * (float)std::log10(1.0 / (s3[0] / n));
        sp += psnr;
        ss += (float)(s3[2] / n);
        sl += (float)(s3[1] / n);
    }
    const int n = (int)cams.size();
    m.psnr = (float)(sp / n);
    m.ssim = (float)(ss / n);
    m.l1 = (float)(sl / n);
    m.numTest = n;
    m.numGaussians = T->model->num_active();
    return m;
    MS_GUARD_END(m)
}

MsplatPixelBuffer msplat_trainer_render(MsplatTrainer t, int cameraIndex, bool useTest) {
    MsplatPixelBuffer b{nullptr, 0, 0};
    if (!t) return b;
    TrainerImpl* T = TR(t);
    MS_GUARD_BEGIN
    auto& cams = useTest ? T->ds->test : T->ds->train;
    if (cameraIndex < 0 || cameraIndex >= (int)cams.size()) return b;
    const TrainCamera& cam = cams[cameraIndex];
    const float* img = T->model->render(cam, T->step);
    return host_image(T, img, cam.width, cam.height);
    MS_GUARD_END(b)
}

MsplatPixelBuffer msplat_trainer_render_pose(MsplatTrainer t, const float camToWorld[16], int refCameraIndex) {
    MsplatPixelBuffer b{nullptr, 0, 0};
    if (!t || !camToWorld) return b;
    TrainerImpl* T = TR(t);
    MS_GUARD_BEGIN
    auto& cams = T->ds->train;
    if (refCameraIndex < 0 || refCameraIndex >= (int)cams.size()) return b;
    TrainCamera cam = cams[refCameraIndex];
    std::memcpy(cam.c2w, camToWorld, sizeof(cam.c2w));
    const float* img = T->model->render(cam, T->step);
    return host_image(T, img, cam.width, cam.height);
    MS_GUARD_END(b)
}

void msplat_trainer_render_pose_to_buffer(MsplatTrainer t, const float camToWorld[16], int refCameraIndex,
                                          uint8_t* outRGBA, int* outWidth, int* outHeight) {
    if (!outWidth || !outHeight) return;
    *outWidth = 0;
    *outHeight = 0;
    if (!t || !camToWorld) return;
    TrainerImpl* T = TR(t);
    try {
        auto& cams = T->ds->train;
        if (refCameraIndex < 0 || refCameraIndex >= (int)cams.size()) return;
        TrainCamera cam = cams[refCameraIndex];
        std::memcpy(cam.c2w, camToWorld, sizeof(cam.c2w));
        *outWidth = cam.width;
        *outHeight = cam.height;
        if (!outRGBA) return;
        const float* img = T->model->render(cam, T->step);
        const int n = cam.width * cam.height;
        T->rgba.ensure((size_t)n * 4);
        hipStream_t s = T->model->engine().stream();
        msk::to_rgba8(s, img, n, T->rgba.as<uint8_t>());
        hip_check(hipMemcpyAsync(outRGBA, T->rgba.p, (size_t)n * 4, hipMemcpyDeviceToHost, s), "rgba download");
        hip_check(hipStreamSynchronize(s), "rgba download");
    } catch (const std::exception& ex) {
        std::fprintf(stderr, "msplat: %s: %s\n", __func__, ex.what());
    }
}

#define MS_VOID_GUARD(body)                                                 \
    try {                                                                   \
        body;                                                               \
    } catch (const std::exception& ex) {                                    \
        std::fprintf(stderr, "msplat: %s: %s\n", __func__, ex.what());      \
    }

void msplat_trainer_export_ply(MsplatTrainer t, const char* path) {
    if (!t || !path) return;
    MS_VOID_GUARD(TR(t)->model->save_ply(path, TR(t)->step))
}

void msplat_trainer_export_splat(MsplatTrainer t, const char* path) {
    if (!t || !path) return;
    MS_VOID_GUARD(TR(t)->model->save_splat(path))
}

void msplat_trainer_save_checkpoint(MsplatTrainer t, const char* path) {
    if (!t || !path) return;
    MS_VOID_GUARD(TR(t)->model->save_checkpoint(path, TR(t)->step))
}

int msplat_trainer_load_checkpoint(MsplatTrainer t, const char* path) {
    if (!t || !path) return -1;
    TrainerImpl* T = TR(t);
    MS_GUARD_BEGIN
    T->step = T->model->load_checkpoint(path);
    T->shuffle();
    return T->step;
    MS_GUARD_END(-1)
}

int msplat_trainer_splat_count(MsplatTrainer t) { return t ? TR(t)->model->num_active() : 0; }
int msplat_trainer_iteration(MsplatTrainer t) { return t ? TR(t)->step : 0; }

int msplat_trainer_step_params(MsplatTrainer t, int cameraIndex, int step, int adamStep, MsplatStepParams* out) {
    if (!t || !out) return -1;
    TrainerImpl* T = TR(t);
    if (cameraIndex < 0 || cameraIndex >= (int)T->ds->train.size()) return -1;
    *out = T->model->step_params(T->ds->train[cameraIndex], step, T->cfg.ssimWeight, adamStep);
    return 0;
}

int msplat_trainer_last_camera(MsplatTrainer t) { return t ? TR(t)->last_cam : -1; }

double msplat_trainer_last_loss(MsplatTrainer t) { return t ? TR(t)->model->engine().loss() : 0.0; }

int64_t msplat_trainer_read_group(MsplatTrainer t, int group, int which, float* dst, int64_t count) {
    if (!t || group < 0 || group > 5 || which < 0 || which > 2) return -1;
    MS_GUARD_BEGIN
    return TR(t)->model->read_group(group, which, dst, count);
    MS_GUARD_END(-1)
}

int msplat_trainer_set_option(MsplatTrainer t, const char* name, double value) {
    if (!t || !name) return -1;
    MS_GUARD_BEGIN
    const std::string n(name);
    if (n == "sh_bf16") {
        TR(t)->model->set_sh_bf16(value != 0.0);
        return 0;
    }
    if (n == "learned_relocation") {
        TR(t)->model->enable_proposal(value != 0.0, 0x51A7C0DEull);
        return 0;
    }
    if (n == "relocation_mode") {
        if (value != 0.0 && value != 1.0 && value != 2.0) throw std::invalid_argument("relocation_mode is 0, 1 or 2");
        TR(t)->judge.force = (int)value;
        TR(t)->model->set_proposal_mode(value == 2.0);
        return 0;
    }
    if (n == "autotune") {
        TR(t)->model->set_batch_autotune(value != 0.0);
        return 0;
    }
    if (n == "batch_sequential") {
        TR(t)->model->set_batch_sequential(value != 0.0);
        return 0;
    }
    if (n == "batch_lr_scale") {
        if (!(value > 0.0) || value > 64.0) throw std::invalid_argument("batch_lr_scale must be in (0, 64]");
        TR(t)->model->set_batch_lr_scale((float)value);
        return 0;
    }
    if (n == "mcmc") {
        TR(t)->model->set_mcmc_enabled(value != 0.0);
        return 0;
    }
    if (n == "nan_inject_step") {
        if (value < -1.0 || value > 2147483647.0) throw std::invalid_argument("nan_inject_step out of range");
        TR(t)->model->set_nan_inject_step((int)value);
        return 0;
    }
    if (n == "batch_graph") {
        TR(t)->model->set_batch_graph(value != 0.0);
        return 0;
    }
    if (n == "raster_ppl") {
        if (value != 1.0 && value != 4.0) throw std::invalid_argument("raster_ppl is 1 or 4");
        TR(t)->model->engine().set_raster_ppl((int)value);
        TR(t)->model->set_raster_ppl((int)value);
        return 0;
    }
    if (n == "persist_grid") {
        if (value < 0.0 || value > 1 << 20) throw std::invalid_argument("persist_grid out of range");
        TR(t)->model->engine().set_persist_grid(value > 0.0 ? (int)value : msplat_hip::default_persist_grid());
        TR(t)->model->set_persist_grid((int)value);
        return 0;
    }
    throw std::invalid_argument("unknown option '" + n + "'");
    MS_GUARD_END(-1)
}

int msplat_trainer_run_mcmc(MsplatTrainer t, int step, int32_t info[6]) {
    if (!t) return -1;
    TrainerImpl* T = TR(t);
    MS_GUARD_BEGIN
    const msk::McmcInfo in = T->model->mcmc_after_train(step);
    if (info) {
        info[0] = in.num_active; info[1] = in.dead; info[2] = in.grown; info[3] = in.nonfinite; info[4] = in.culled;
        info[5] = (int32_t)T->model->relocation_count();
    }
    return 0;
    MS_GUARD_END(-1)
}

int msplat_trainer_optim_state(MsplatTrainer t, int* adamSteps, float lr[6]) {
    if (!t) return -1;
    OptimState& o = TR(t)->model->optim();
    if (adamSteps) *adamSteps = o.adam_step_count;
    if (lr) std::memcpy(lr, o.adam_lr, sizeof(o.adam_lr));
    return 0;
}

MsplatEngine* msplat_trainer_engine(MsplatTrainer t) { return t ? &TR(t)->handle : nullptr; }

int msplat_eq9_device(int n, const float* o_old, const int* counts, float* o_new, float* s_new) {
    if (n <= 0 || !o_old || !counts || !o_new || !s_new) return -1;
    MS_GUARD_BEGIN
    DBuf a, b, c, d;
    a.ensure((size_t)n * 4); b.ensure((size_t)n * 4); c.ensure((size_t)n * 4); d.ensure((size_t)n * 4);
    hip_check(hipMemcpy(a.p, o_old, (size_t)n * 4, hipMemcpyHostToDevice), "eq9");
    hip_check(hipMemcpy(b.p, counts, (size_t)n * 4, hipMemcpyHostToDevice), "eq9");
    msk::eq9_cases(nullptr, n, a.as<float>(), b.as<int>(), c.as<float>(), d.as<float>());
    hip_check(hipDeviceSynchronize(), "eq9");
    hip_check(hipMemcpy(o_new, c.p, (size_t)n * 4, hipMemcpyDeviceToHost), "eq9");
    hip_check(hipMemcpy(s_new, d.p, (size_t)n * 4, hipMemcpyDeviceToHost), "eq9");
    a.release(); b.release(); c.release(); d.release();
    return 0;
    MS_GUARD_END(-1)
}

void msplat_set_metallib_path(const char* path) { (void)path; }

void msplat_sync(void) {
    sync_all_engines();
    (void)hipDeviceSynchronize();
}

void msplat_cleanup(void) { trim_all_engines(); }

// ---- viewer ----------------------------------------------------------------------------

MsplatViewer msplat_viewer_create(const char* plyPath, int width, int height, float fovY) {
    MS_GUARD_BEGIN
    if (!plyPath || width <= 0 || height <= 0) return nullptr;
    std::unique_ptr<ViewerImpl> v(new ViewerImpl());
    v->width = width;
    v->height = height;
    v->fov_y = fovY;
    int step = 0;
    const float zero[3] = {0.f, 0.f, 0.f};
    HostGaussians g = read_ply(plyPath, &step, 1.0f, zero, false);
    v->degree = g.degree;
    v->n = g.n;
    const std::vector<float>* src[6] = {&g.means, &g.scales, &g.quats, &g.f_dc, &g.f_rest, &g.opac};
    for (int k = 0; k < 6; k++) {
        const size_t bytes = src[k]->size() * 4;
        v->p[k].ensure(bytes);
        if (bytes) hip_check(hipMemcpy(v->p[k].p, src[k]->data(), bytes, hipMemcpyHostToDevice), "viewer upload");
    }
    v->bounds(g);
    return v.release();
    MS_GUARD_END(nullptr)
}

void msplat_viewer_resize(MsplatViewer v, int width, int height) {
    if (!v || width <= 0 || height <= 0) return;
    VW(v)->width = width;
    VW(v)->height = height;
}

void msplat_viewer_set_fov(MsplatViewer v, float fovY) {
    if (v) VW(v)->fov_y = fovY;
}

void msplat_viewer_set_background(MsplatViewer v, float r, float g, float b) {
    if (!v) return;
    VW(v)->bg[0] = r;
    VW(v)->bg[1] = g;
    VW(v)->bg[2] = b;
}

void msplat_viewer_render(MsplatViewer v, const float camToWorld[16], uint8_t* outRGBA) {
    if (!v || !camToWorld || !outRGBA) return;
    ViewerImpl* V = VW(v);
    try {
        const float* img = V->render_device(camToWorld);
        const int n = V->width * V->height;
        V->rgba.ensure((size_t)n * 4);
        hipStream_t s = V->eng.stream();
        msk::to_rgba8(s, img, n, V->rgba.as<uint8_t>());
        hip_check(hipMemcpyAsync(outRGBA, V->rgba.p, (size_t)n * 4, hipMemcpyDeviceToHost, s), "viewer download");
        hip_check(hipStreamSynchronize(s), "viewer download");
    } catch (const std::exception& ex) {
        std::fprintf(stderr, "msplat: %s: %s\n", __func__, ex.what());
    }
}

int msplat_viewer_splat_count(MsplatViewer v) { return v ? VW(v)->n : 0; }

void msplat_viewer_get_size(MsplatViewer v, int* outWidth, int* outHeight) {
    if (!v) return;
    if (outWidth) *outWidth = VW(v)->width;
    if (outHeight) *outHeight = VW(v)->height;
}

void* msplat_viewer_render_buffer(MsplatViewer v, const float camToWorld[16], int* outWidth, int* outHeight) {
    if (!v || !camToWorld) {
        if (outWidth) *outWidth = 0;
        if (outHeight) *outHeight = 0;
        return nullptr;
    }
    ViewerImpl* V = VW(v);
    if (outWidth) *outWidth = V->width;
    if (outHeight) *outHeight = V->height;
    try {
        return (void*)V->render_device(camToWorld);
    } catch (const std::exception& ex) {
        std::fprintf(stderr, "msplat: %s: %s\n", __func__, ex.what());
        return nullptr;
    }
}

void msplat_viewer_bounds(MsplatViewer v, float boundsMin[3], float boundsMax[3]) {
    if (!v || !boundsMin || !boundsMax) return;
    std::memcpy(boundsMin, VW(v)->bmin, 12);
    std::memcpy(boundsMax, VW(v)->bmax, 12);
}

void msplat_viewer_density_center(MsplatViewer v, float center[3]) {
    if (!v || !center) return;
    std::memcpy(center, VW(v)->dcenter, 12);
}

void* msplat_viewer_command_queue(MsplatViewer v) { return v ? (void*)VW(v)->eng.stream() : nullptr; }

void msplat_viewer_destroy(MsplatViewer v) { delete VW(v); }

}  // extern "C"
