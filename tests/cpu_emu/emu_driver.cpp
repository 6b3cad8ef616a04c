// Runs the library's C API end to end on the CPU emulation of HIP (TEST INFRASTRUCTURE):
// every kernel of a training step, relocation, evaluation, the render entry points, the
// formats and the viewer, under AddressSanitizer / UBSan. Exits 0 when every scenario ran
// without a sanitizer report and the checks below hold.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "msplat_api_ext.h"
#include "msplat_c_api.h"

#define CHECK(c)                                                             \
    do {                                                                     \
        if (!(c)) {                                                          \
            std::fprintf(stderr, "CHECK failed %s:%d: %s\n", __FILE__, __LINE__, #c); \
            std::exit(1);                                                    \
        }                                                                    \
    } while (0)

struct Scene {
    int n, degree;
    std::vector<float> means, scales, quats, opac, f_dc, f_rest;
};

static Scene synth(int kind, uint64_t seed, int n, int degree) {
    Scene s;
    s.n = n;
    s.degree = degree;
    const int nr = ((degree + 1) * (degree + 1) - 1) * 3;
    s.means.resize(n * 3); s.scales.resize(n * 3); s.quats.resize(n * 4); s.opac.resize(n);
    s.f_dc.resize(n * 3); s.f_rest.resize((size_t)n * (nr ? nr : 1));
    CHECK(msplat_synth_gaussians(kind, seed, n, degree, s.means.data(), s.scales.data(), s.quats.data(),
                                 s.opac.data(), s.f_dc.data(), s.f_rest.data()) == 0);
    return s;
}

static MsplatDataset dataset(const Scene& g, int W, int H, int ncams, int test_every) {
    MsplatDataset ds = msplat_dataset_create_empty();
    CHECK(ds);
    std::vector<float> c2w((size_t)ncams * 16);
    CHECK(msplat_synth_cameras(18, ncams, 3.0f, 0.0f, c2w.data()) == 0);
    std::vector<uint8_t> img((size_t)W * H * 3);
    for (size_t i = 0; i < img.size(); i++) img[i] = (uint8_t)((i * 2654435761u) >> 24);
    for (int c = 0; c < ncams; c++) {
        const bool test = test_every > 0 && c % test_every == 0;
        CHECK(msplat_dataset_add_camera(ds, &c2w[(size_t)c * 16], (float)W, (float)W, W / 2.0f, H / 2.0f, W, H,
                                        img.data(), test) >= 0);
    }
    CHECK(msplat_dataset_set_init_gaussians(ds, g.n, g.degree, g.means.data(), g.scales.data(), g.quats.data(),
                                            g.opac.data(), g.f_dc.data(), g.f_rest.data()) == 0);
    return ds;
}

static std::vector<float> group(MsplatTrainer t, int g, int which) {
    const int64_t n = msplat_trainer_read_group(t, g, which, nullptr, 0);
    CHECK(n >= 0);
    std::vector<float> v((size_t)n);
    if (n) CHECK(msplat_trainer_read_group(t, g, which, v.data(), n) == n);
    return v;
}

static void scenario_train(const char* name, int W, int H, int n, int cap, int steps, bool bf16, const char* dir) {
    std::fprintf(stderr, "== %s: %dx%d n=%d cap=%d steps=%d bf16=%d\n", name, W, H, n, cap, steps, (int)bf16);
    Scene g = synth(0, 11, n, 3);
    MsplatDataset ds = dataset(g, W, H, 4, 0);
    MsplatConfig cfg = msplat_default_config();
    cfg.iterations = 100;
    cfg.shDegreeInterval = 1;
    cfg.capMax = cap;
    MsplatTrainer t = msplat_trainer_create(ds, cfg);
    CHECK(t);
    if (bf16) CHECK(msplat_trainer_set_option(t, "sh_bf16", 1) == 0);
    for (int s = 0; s < steps; s++) {
        MsplatStats st = msplat_trainer_step(t);
        CHECK(st.iteration == s + 1);
        const double loss = msplat_trainer_last_loss(t);
        CHECK(std::isfinite(loss) && loss > 0.0);
    }
    std::string ck = std::string(dir) + "/" + name + ".msplat";
    msplat_trainer_save_checkpoint(t, ck.c_str());
    MsplatTrainer t2 = msplat_trainer_create(ds, cfg);
    CHECK(t2);
    if (bf16) CHECK(msplat_trainer_set_option(t2, "sh_bf16", 1) == 0);
    CHECK(msplat_trainer_load_checkpoint(t2, ck.c_str()) == steps);
    for (int gi = 0; gi < 6; gi++)
        for (int w = 0; w < 3; w++) CHECK(group(t, gi, w) == group(t2, gi, w));
    MsplatStats st2 = msplat_trainer_step(t2);
    CHECK(st2.iteration == steps + 1);
    std::string ply = std::string(dir) + "/" + name + ".ply";
    msplat_trainer_export_ply(t, ply.c_str());
    msplat_trainer_export_splat(t, (std::string(dir) + "/" + name + ".splat").c_str());
    MsplatPixelBuffer pb = msplat_trainer_render(t, 1, false);
    CHECK(pb.data && pb.width == W && pb.height == H);
    std::free(pb.data);
    float pose[16];
    msplat_dataset_camera_pose(ds, 2, pose);
    int w = 0, h = 0;
    msplat_trainer_render_pose_to_buffer(t, pose, 2, nullptr, &w, &h);
    CHECK(w == W && h == H);
    std::vector<uint8_t> rgba((size_t)w * h * 4);
    msplat_trainer_render_pose_to_buffer(t, pose, 2, rgba.data(), &w, &h);
    MsplatViewer v = msplat_viewer_create(ply.c_str(), 64, 48, 0.9f);
    CHECK(v);
    std::vector<uint8_t> vb(64 * 48 * 4);
    msplat_viewer_render(v, pose, vb.data());
    msplat_viewer_resize(v, 40, 30);
    std::vector<uint8_t> vb2(40 * 30 * 4);
    msplat_viewer_render(v, pose, vb2.data());
    msplat_viewer_destroy(v);
    msplat_trainer_destroy(t2);
    msplat_trainer_destroy(t);
    msplat_dataset_destroy(ds);
}

static void scenario_mcmc(bool bf16) {
    std::fprintf(stderr, "== mcmc bf16=%d\n", (int)bf16);
    Scene g = synth(0, 5, 3000, 3);
    for (int i = 0; i < g.n; i += 5) g.opac[i] = -8.0f;
    for (int i = 1; i < g.n; i += 97) for (int k = 0; k < 3; k++) g.means[(size_t)i * 3 + k] *= 10.0f;
    g.f_rest[(size_t)7 * 45 + 4] = NAN;
    MsplatDataset ds = dataset(g, 96, 80, 2, 0);
    MsplatConfig cfg = msplat_default_config();
    cfg.iterations = 1000;
    cfg.capMax = 4000;
    MsplatTrainer t = msplat_trainer_create(ds, cfg);
    CHECK(t);
    if (bf16) CHECK(msplat_trainer_set_option(t, "sh_bf16", 1) == 0);
    msplat_trainer_step(t);
    int32_t info[6];
    CHECK(msplat_trainer_run_mcmc(t, 500, info) == 0);
    std::fprintf(stderr, "   active %d dead %d grown %d nonfinite %d culled %d count %d\n", info[0], info[1], info[2],
                 info[3], info[4], info[5]);
    CHECK(info[1] > 0 && info[2] > 0 && info[0] == 3149);
    msplat_trainer_step(t);
    CHECK(msplat_trainer_run_mcmc(t, 600, info) == 0);
    CHECK(info[0] == 3306);
    msplat_trainer_step(t);
    msplat_trainer_destroy(t);
    msplat_dataset_destroy(ds);
}

static void scenario_eval() {
    std::fprintf(stderr, "== evaluate\n");
    Scene g = synth(0, 3, 1500, 3);
    MsplatDataset ds = dataset(g, 72, 56, 4, 2);
    CHECK(msplat_dataset_num_test(ds) == 2);
    MsplatConfig cfg = msplat_default_config();
    cfg.iterations = 2;
    MsplatTrainer t = msplat_trainer_create(ds, cfg);
    CHECK(t);
    msplat_trainer_train(t);
    MsplatEvalMetrics m = msplat_trainer_evaluate(t);
    CHECK(m.numTest == 2 && std::isfinite(m.psnr) && m.ssim > -1.f && m.ssim <= 1.f);
    msplat_trainer_destroy(t);
    msplat_dataset_destroy(ds);
}

static void scenario_eq9() {
    std::fprintf(stderr, "== eq9\n");
    float o[4] = {0.05f, 0.5f, 0.9f, 0.998f};
    int c[4] = {1, 2, 7, 51};
    float on[4], sn[4];
    CHECK(msplat_eq9_device(4, o, c, on, sn) == 0);
    CHECK(on[0] == o[0] && std::fabs(sn[0] - 1.0f) < 1e-6f);
}

// A small end-to-end pass over every entry point (a few seconds under the emulation).
static void scenario_mini(const char* dir) {
    std::fprintf(stderr, "== mini\n");
    Scene g = synth(0, 21, 400, 3);
    for (int i = 0; i < g.n; i += 7) g.opac[i] = -8.0f;
    MsplatDataset ds = dataset(g, 48, 40, 4, 2);
    MsplatConfig cfg = msplat_default_config();
    cfg.iterations = 1000;
    cfg.shDegreeInterval = 1;
    cfg.capMax = 600;
    MsplatTrainer t = msplat_trainer_create(ds, cfg);
    CHECK(t);
    for (int s = 0; s < 3; s++) CHECK(msplat_trainer_step(t).iteration == s + 1);
    int32_t info[6];
    CHECK(msplat_trainer_run_mcmc(t, 500, info) == 0);
    std::fprintf(stderr, "   active %d dead %d grown %d nonfinite %d culled %d count %d\n", info[0], info[1], info[2],
                 info[3], info[4], info[5]);
    CHECK(info[0] == 419 && info[1] > 0);  // (int)(1.05f * 400) = 419
    CHECK(msplat_trainer_set_option(t, "sh_bf16", 1) == 0);
    CHECK(msplat_trainer_step(t).iteration == 4);
    CHECK(msplat_trainer_run_mcmc(t, 600, info) == 0);
    CHECK(msplat_trainer_step(t).splatCount == 439);
    MsplatEvalMetrics m = msplat_trainer_evaluate(t);
    CHECK(m.numTest == 2 && std::isfinite(m.psnr));
    const std::string ck = std::string(dir) + "/mini.msplat", ply = std::string(dir) + "/mini.ply";
    msplat_trainer_save_checkpoint(t, ck.c_str());
    CHECK(msplat_trainer_load_checkpoint(t, ck.c_str()) == 5);
    {   // a checkpoint of another SH degree is rejected and leaves the trainer as it was
        MsplatConfig c1 = cfg;
        c1.shDegree = 1;
        MsplatTrainer t1 = msplat_trainer_create(ds, c1);
        CHECK(t1);
        CHECK(msplat_trainer_step(t1).iteration == 1);
        const std::vector<float> before = group(t1, 4, 0);
        CHECK(msplat_trainer_load_checkpoint(t1, ck.c_str()) == -1);
        CHECK(msplat_trainer_iteration(t1) == 1 && msplat_trainer_splat_count(t1) == 400);
        CHECK(group(t1, 4, 0) == before);
        CHECK(msplat_trainer_step(t1).iteration == 2);
        msplat_trainer_destroy(t1);
    }
    msplat_trainer_export_ply(t, ply.c_str());
    float pose[16];
    msplat_dataset_camera_pose(ds, 1, pose);
    MsplatViewer v = msplat_viewer_create(ply.c_str(), 32, 24, 0.8f);
    CHECK(v && msplat_viewer_splat_count(v) == 439);
    std::vector<uint8_t> buf(32 * 24 * 4);
    msplat_viewer_render(v, pose, buf.data());
    msplat_viewer_destroy(v);
    msplat_trainer_destroy(t);
    msplat_dataset_destroy(ds);
}

int main(int argc, char** argv) {
    const char* dir = argc > 1 ? argv[1] : "/tmp";
    const std::string only = argc > 2 ? argv[2] : "";
    auto want = [&](const char* s) { return only.empty() || only == s; };
    if (want("mini")) scenario_mini(dir);
    if (want("chunked")) scenario_train("chunked", 96, 80, 1500, 1600, 2, false, dir);
    if (want("chunked_ragged")) scenario_train("chunked_ragged", 90, 70, 700, 4000, 2, false, dir);
    if (want("monolithic")) scenario_train("monolithic", 320, 336, 300, 4000, 2, false, dir);
    if (want("bf16")) scenario_train("bf16", 96, 80, 1500, 4000, 3, true, dir);
    if (want("mcmc")) scenario_mcmc(false);
    if (want("mcmc_bf16")) scenario_mcmc(true);
    if (want("eval")) scenario_eval();
    if (want("eq9")) scenario_eq9();
    std::fprintf(stderr, "emu driver: all scenarios passed\n");
    return 0;
}
