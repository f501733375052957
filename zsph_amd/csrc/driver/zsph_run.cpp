// zsph_run: the host driver (single_gpu_sph_src/deltasph/main.cu:8-151) over
// the C ABI.  Reads a case file (io/read_json.cu keys) and its
// <input_path>/input.vtk, or generates a dam break; steps on the GPU and
// hands output frames to a pool of writer threads (io/mul_thread.cu:3-80),
// here a blocking queue that is drained and joined before exit.  Output
// steps do not stall the solver: zsph_download_async snapshots the state on
// the device and copies it to pinned host memory on its own stream while
// the next steps run; a writer waits for the copy and unpacks it.
// Checkpoints: --checkpoint-step K writes <output_path>/zsph_ckpt<s>.bin
// after every K-th step (s = steps done); --restart FILE takes the state of
// such a file and continues the loop at its step (the start_step /
// current_step the reference's case file carries, io/read_json.cu:32-33).
// --walk, --skin and --tile-rows set the matching zsph_tuning fields before
// the first step (each left at the library default when not given).
//
//   zsph_run case.json [--checkpoint-step K] [--restart FILE] [--walk W] [--skin S] [--tile-rows R]
//   zsph_run --dambreak {2,3} [--dx DX] [--steps N] [--output-step K] [--out DIR] [--gamma G]
//            [--checkpoint-step K] [--restart FILE] [--walk W] [--skin S] [--tile-rows R]
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "zsph.h"

namespace {

struct Frame {
    int step = 0;
    int n = 0;
    std::vector<float> pos, vel, rhop, acc;
    std::vector<int> type, table, hash;
    zsph_frame *pending = nullptr;  // an asynchronous download still to be waited for (the writer does it)

    void alloc() {
        pos.resize(3 * (size_t)n); vel.resize(3 * (size_t)n); rhop.resize(2 * (size_t)n);
        acc.resize(4 * (size_t)n); type.resize(n); table.resize(n); hash.resize(n);
    }
    zsph_host_state state() {
        return zsph_host_state{pos.data(), vel.data(), rhop.data(), acc.data(), type.data(), table.data(), hash.data()};
    }
};

class Writers {
  public:
    Writers(int nthreads, std::string dir, int fields, int binary)
        : dir_(std::move(dir)), fields_(fields), binary_(binary) {
        for (int t = 0; t < (nthreads < 1 ? 1 : nthreads); ++t) pool_.emplace_back([this] { run(); });
    }
    ~Writers() {
        {
            std::lock_guard<std::mutex> g(m_);
            done_ = true;
        }
        cv_.notify_all();
        for (auto &t : pool_) t.join();
    }
    void push(Frame &&f) {
        {
            std::unique_lock<std::mutex> g(m_);
            room_.wait(g, [this] { return q_.size() < pool_.size() + 1; });
            q_.push_back(std::move(f));
        }
        cv_.notify_one();
    }
    int errors() const { return errors_; }

  private:
    void run() {
        for (;;) {
            Frame f;
            {
                std::unique_lock<std::mutex> g(m_);
                cv_.wait(g, [this] { return done_ || !q_.empty(); });
                if (q_.empty()) return;
                f = std::move(q_.front());
                q_.pop_front();
            }
            room_.notify_one();
            if (f.pending) {  // unpack the asynchronous download here, off the stepping thread
                f.alloc();
                zsph_host_state st = f.state();
                if (zsph_frame_wait(f.pending, &st) != ZSPH_OK) {
                    std::lock_guard<std::mutex> g(m_);
                    ++errors_;
                    continue;
                }
            }
            zsph_host_state st = f.state();
            const std::string path = dir_ + "/zsph" + std::to_string(f.step) + ".vtk";
            if (zsph_write_vtk(path.c_str(), &st, f.n, fields_, binary_) != ZSPH_OK) {
                std::lock_guard<std::mutex> g(m_);
                ++errors_;
            }
        }
    }
    std::string dir_;
    int fields_, binary_;
    std::vector<std::thread> pool_;
    std::deque<Frame> q_;
    std::mutex m_;
    std::condition_variable cv_, room_;
    bool done_ = false;
    int errors_ = 0;
};

int usage() {
    std::fprintf(stderr,
                 "usage: zsph_run case.json [--checkpoint-step K] [--restart FILE] [--sync-output 1] [--writers N] [--walk W]\n"
                 "                [--skin S] [--tile-rows R]\n"
                 "       zsph_run --dambreak {2,3} [--dx DX] [--steps N] [--output-step K] [--out DIR] [--gamma G] [--gsf F]\n"
                 "                [--checkpoint-step K] [--restart FILE] [--sync-output 1] [--writers N] [--walk W]\n"
                 "                [--skin S] [--tile-rows R]\n");
    return 2;
}

}  // namespace

int main(int argc, char **argv) {
    if (argc < 2) return usage();
    zsph_json j;
    zsph_json_defaults(&j);
    std::vector<float> pos, vel, rhop, acc;
    std::vector<int> type, table;
    int n = 0;
    int ckpt_every = 0;
    bool sync_output = false;
    int writers = 0;  // writer threads (0: the case's thread_num)
    int walk = -1;    // neighbour walk (zsph_tuning.walk; -1: the library default)
    float skin = -1;  // neighbour skin in units of h (zsph_tuning.skin; < 0: the library default)
    int tile_rows = -1;  // wave order of skin mode (zsph_tuning.tile_rows; -1: the library default)
    std::string restart;
    // options common to both forms; true if argv[a] was one
    auto common = [&](const std::string &k, const char *v) {
        if (k == "--checkpoint-step") ckpt_every = std::atoi(v);
        else if (k == "--restart") restart = v;
        else if (k == "--sync-output") sync_output = std::atoi(v) != 0;
        else if (k == "--writers") writers = std::atoi(v);
        else if (k == "--walk") walk = std::atoi(v);
        else if (k == "--skin") skin = (float)std::atof(v);
        else if (k == "--tile-rows") tile_rows = std::atoi(v);
        else return false;
        return true;
    };
    if (std::strcmp(argv[1], "--dambreak") == 0) {
        if (argc < 3) return usage();
        zsph_dambreak c{};
        c.dim = std::atoi(argv[2]);
        c.layers = 3;
        if (c.dim == 2) {
            c.dx = 0.005f; c.lx = 1.61f; c.wy = 0.0f; c.lz = 0.8f; c.cw = 0.4f; c.ch = 0.3f;
        } else {
            c.dx = 0.008f; c.lx = 3.22f; c.wy = 1.0f; c.lz = 1.0f; c.cw = 1.2f; c.ch = 0.55f;
        }
        j.cfl = 0.2f;
        j.gamma = 7.0f;
        j.grid_size_factor = 2.0f;  // cells of edge h, as the benchmark runs
        j.end_step = 100;
        j.output_step = 50;
        for (int a = 3; a + 1 < argc; a += 2) {
            const std::string k = argv[a];
            const char *v = argv[a + 1];
            if (k == "--dx") c.dx = (float)std::atof(v);
            else if (k == "--steps") j.end_step = std::atoi(v);
            else if (k == "--output-step") j.output_step = std::atoi(v);
            else if (k == "--out") std::snprintf(j.output_path, sizeof(j.output_path), "%s", v);
            else if (k == "--gamma") j.gamma = (float)std::atof(v);
            else if (k == "--gsf") j.grid_size_factor = (float)std::atof(v);
            else if (!common(k, v)) return usage();
        }
        if ((argc - 3) % 2) return usage();
        if (zsph_case_dambreak(&c, &j, nullptr, &n) != ZSPH_OK) return 1;
        pos.resize(3 * (size_t)n); vel.resize(3 * (size_t)n); rhop.resize(2 * (size_t)n);
        acc.resize(4 * (size_t)n); type.resize(n); table.resize(n);
        zsph_host_state st{pos.data(), vel.data(), rhop.data(), acc.data(), type.data(), table.data(), nullptr};
        if (zsph_case_dambreak(&c, &j, &st, &n) != ZSPH_OK) return 1;
    } else {
        if (zsph_read_json(argv[1], &j) != ZSPH_OK) {
            std::fprintf(stderr, "cannot read %s\n", argv[1]);
            return 1;
        }
        if (argc % 2) return usage();
        for (int a = 2; a + 1 < argc; a += 2)
            if (!common(argv[a], argv[a + 1])) return usage();
        n = j.ptc_num;
        pos.resize(3 * (size_t)n); vel.resize(3 * (size_t)n); rhop.resize(2 * (size_t)n);
        acc.resize(4 * (size_t)n); type.resize(n); table.resize(n);
        zsph_host_state st{pos.data(), vel.data(), rhop.data(), acc.data(), type.data(), table.data(), nullptr};
        const std::string vtk = std::string(j.input_path) + "/input.vtk";
        if (restart.empty() && zsph_read_vtk(vtk.c_str(), &st, n, j.rho0) != ZSPH_OK) {
            std::fprintf(stderr, "cannot read %s (ptc_num %d)\n", vtk.c_str(), n);
            return 1;
        }
    }
    zsph_cpu_param cp;
    int r = zsph_set_cpu_param(&cp, &j);
    if (r == ZSPH_E_DT) {
        std::fprintf(stderr, "* Dt is greater than the dt_max\n");
        return 1;
    }
    if (r != ZSPH_OK) {
        std::fprintf(stderr, "invalid case parameters (%d)\n", r);
        return 1;
    }
    zsph_params gp;
    zsph_set_gpu_param(&gp, &cp);
    std::printf("* ptc_num %d  grid %d x %d x %d  dt %.4g  h %.4g  c0 %.4g\n", cp.ptc_num, cp.grid_xdim,
                cp.grid_ydim, cp.grid_zdim, cp.dt, cp.h, cp.cs);
    zsph_ctx *ctx = nullptr;
    if ((r = zsph_create(&gp, cp.gpu_id, &ctx)) != ZSPH_OK) {
        std::fprintf(stderr, "zsph_create failed (%d)\n", r);
        return 1;
    }
    zsph_host_state st{pos.data(), vel.data(), rhop.data(), acc.data(), type.data(), table.data(), nullptr};
    if (walk >= 0 || skin >= 0 || tile_rows >= 0) {
        zsph_tuning t;
        zsph_get_tuning(ctx, &t);
        if (walk >= 0) t.walk = walk;
        if (skin >= 0) t.skin = skin;
        if (tile_rows >= 0) t.tile_rows = tile_rows;
        if (zsph_set_tuning(ctx, &t) != ZSPH_OK) {
            std::fprintf(stderr, "%s\n", zsph_last_error(ctx));
            zsph_destroy(ctx);
            return 1;
        }
    }
    int first = cp.start_step;
    if (restart.empty() ? zsph_upload(ctx, 0, &st) != ZSPH_OK : zsph_checkpoint_read(ctx, restart.c_str(), nullptr) != ZSPH_OK) {
        std::fprintf(stderr, "%s\n", zsph_last_error(ctx));
        zsph_destroy(ctx);
        return 1;
    }
    if (!restart.empty()) {
        zsph_checkpoint_info info;
        if (zsph_checkpoint_inspect(restart.c_str(), 0, &info) != ZSPH_OK) {
            zsph_destroy(ctx);
            return 1;
        }
        first = (int)info.step;
        std::printf("* restart from %s at step %d (t = %.6g s)\n", restart.c_str(), first, info.time);
    }
    int status = 0;
    double loop_s = 0.0;
    int loop_steps = 0;
    auto t_drain = std::chrono::steady_clock::now();
    {
        Writers w(writers > 0 ? writers : cp.thread_num, cp.output_path, cp.output_fields | ZSPH_F_POS, cp.binary);
        auto t0 = std::chrono::steady_clock::now();
        for (int step = first; step < cp.end_step; ++step) {
            if (zsph_step(ctx, 1) != ZSPH_OK) {
                std::fprintf(stderr, "step %d: %s\n", step, zsph_last_error(ctx));
                status = 1;
                break;
            }
            if (cp.output_step > 0 && step % cp.output_step == 0) {
                Frame f;
                f.step = step;
                f.n = n;
                // asynchronous: a device snapshot now, the PCIe copy overlapping the next
                // steps, unpacked by the writer; --sync-output 1: the blocking download
                const int rc = sync_output ? (f.alloc(), [&] { zsph_host_state fs = f.state();
                                                             return zsph_download(ctx, 0, &fs, ZSPH_F_ALL); }())
                                           : zsph_download_async(ctx, 0, ZSPH_F_ALL, &f.pending);
                if (rc != ZSPH_OK) {
                    std::fprintf(stderr, "download: %s\n", zsph_last_error(ctx));
                    status = 1;
                    break;
                }
                const double el = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
                std::printf("step %d  t=%.5f s  wall %.2f s\n", step, (step + 1) * cp.dt, el);
                std::fflush(stdout);
                w.push(std::move(f));
            }
            if (ckpt_every > 0 && (step + 1) % ckpt_every == 0) {
                const std::string path =
                    std::string(cp.output_path) + "/zsph_ckpt" + std::to_string(step + 1) + ".bin";
                if (zsph_checkpoint_write(ctx, path.c_str(), step + 1) != ZSPH_OK) {
                    std::fprintf(stderr, "checkpoint: %s\n", zsph_last_error(ctx));
                    status = 1;
                    break;
                }
            }
        }
        if (zsph_sync(ctx) != ZSPH_OK) {
            std::fprintf(stderr, "%s\n", zsph_last_error(ctx));
            status = 1;
        }
        loop_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        loop_steps = cp.end_step - first;
        t_drain = std::chrono::steady_clock::now();
        if (w.errors()) status = 1;
    }  // the writers drain their queue and join here
    const double drain_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_drain).count();
    std::printf("* stepping loop %.3f s for %d steps (%.4f ms/step); writers drained %.3f s later\n", loop_s,
                loop_steps, loop_steps > 0 ? 1e3 * loop_s / loop_steps : 0.0, drain_s);
    int killed = 0;
    zsph_killed(ctx, &killed);
    std::printf("* done, removed particles: %d\n", killed);
    zsph_destroy(ctx);
    return status;
}
