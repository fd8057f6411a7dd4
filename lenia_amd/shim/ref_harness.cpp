// ref_harness.cpp -- Linux stand-in for the reference's caller: main.cpp's parameter
// globals and dispatch loop (main.cpp:30-66, 173-247) and sim_thread_loop's 12-phase,
// two-thread protocol (render.cpp:357-604), driving the step through the shim.
//
//   ref_harness <config.txt> <state_in.f32> <nsteps> <out.f32>
//
// config.txt: "key value" lines (mode, border, sobel_border, gradient, growth, mu_vel,
// adv_vel, blur, dt, sigma, dd, s_norm, theta_a, gta, sta, v_damp, channels,
// asymptotic0..3, soft_clip0..3) and one "kernel c0 c1 radius beta_size r_a m s h b0 b1 .."
// line per kernel.  state_in: A as [C][z][y][x] float32.  out: A, H, U_SUM (unpadded
// [C][N] each) followed by the padded read-backs RBB, H_RB, U_SUM_RB ([C][P] each).
// The run starts at the FFT phase (the reference's first dispatch is FFT_MT,
// main.cpp:198 with current_sim_event = SIM_EVENT_FFT, main.cpp:321).
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <mutex>
#include <sstream>
#include <string>
#include <thread>
#include <vector>

#include "lenia_ref_shim.h"

// ---- main.cpp globals ----------------------------------------------------------------
float GTA = 0.1f, STA = 0.1f, V_DAMP = 0.9f, neighborhood = 50.0f, ISO_VALUE = 0.5f, N = 2;
float THETA_A = 1, timestep = 0.1f, OSC_LA_AMT = -0.001f, OSC_FLOW_AMT = -0.001f, OSC_DIV_FLOW_AMT = -0.001f;
float S_NORM = 0.0f;
int channels = 3, kernel_count = 0;
bool ADV_VEL = true, OSC_LA = false, OSC_FLOW = false, OSC_DIV_FLOW = false, OSC_VEL = false, MU_VEL = false;
bool BLUR_CENTRAL_DIFF = false;
Border SOBEL_BORDER = BORDER_NONE, BORDER = BORDER_NONE;
AlphaMode ALPHA_MODE = ALPHA_GRID;
GrowthType GROWTH_TYPE = GROWTH_RT;
ModeType MODE_TYPE = MODE_SRT;
GradientType GRADIENT_MODE = GRADIENT_SOBEL;
Kernel blur_kernel;
Kernel kernels[MAX_KERNELS];
OffsetTable off_table;
ChannelData ch[MAX_CHANNELS];

enum { FFT, FFT_MT, BLUR, TSB, SOBEL_ONE, SOBEL_TWO, F, MU, TB, REINTEGRATION, TRM, TMB, PHASES };

// One phase as sim_thread_loop dispatches it (render.cpp:376-592).
static void run_phase(int ev, TD *t) {
    switch (ev) {
    case FFT: fft_step(t); break;
    case FFT_MT: mt_fft_step(t); break;
    case BLUR: if (GRADIENT_MODE == GRADIENT_CENTRAL_DIFF) h_sum_blur(t); break;
    case TSB:
        if (GRADIENT_MODE == GRADIENT_SOBEL) {
            if (SOBEL_BORDER == BORDER_TORUS) transpose_sobel_borders(t);
            else if (SOBEL_BORDER == BORDER_WALL) wall_sobel_borders(t);
            else zero_sobel_borders(t);
        } else if (SOBEL_BORDER == BORDER_TORUS) transpose_central_diff_borders(t);
        else if (SOBEL_BORDER == BORDER_WALL) wall_central_diff_borders(t);
        break;
    case SOBEL_ONE: sobel_convolve(t); break;
    case SOBEL_TWO: sobel_convolve_h(t); break;
    case F: compute_f(t); break;
    case MU:
        if (MU_VEL) BORDER == BORDER_WALL ? compute_velocity_mu_wall(t) : compute_velocity_mu(t);
        else BORDER == BORDER_WALL ? compute_f_mu_wall(t) : compute_f_mu(t);
        break;
    case TB: transpose_borders(t); break;
    case REINTEGRATION: reintegration_step(t); break;
    case TRM: transpose_to_row_major(t); break;
    case TMB: transpose_mesh_borders(t); break;
    }
    t->is_done = true;
}

// Phase barrier: the main thread publishes a phase, both sim threads run it, the last
// one to finish signals completion (render.cpp:593-597, main.cpp:182-203).
struct Dispatcher {
    std::mutex mu;
    std::condition_variable work, done;
    int phase = -1, gen = 0, running = 0;
    bool quit = false;
};

static void sim_thread(Dispatcher *d, TD *t) {
    int seen = 0;
    for (;;) {
        int ev;
        {
            std::unique_lock<std::mutex> lk(d->mu);
            d->work.wait(lk, [&] { return d->quit || d->gen != seen; });
            if (d->quit) return;
            seen = d->gen;
            ev = d->phase;
        }
        t->is_done = false;
        run_phase(ev, t);
        std::lock_guard<std::mutex> lk(d->mu);
        if (--d->running == 0) d->done.notify_all();
    }
}

static bool read_config(const char *path) {
    std::ifstream f(path);
    if (!f) return false;
    std::string line;
    kernel_count = 0;
    while (std::getline(f, line)) {
        std::istringstream is(line);
        std::string k;
        is >> k;
        if (k.empty()) continue;
        if (k == "kernel") {
            Kernel &kr = kernels[kernel_count++];
            is >> kr.c0 >> kr.c1 >> kr.radius >> kr.beta_size >> kr.r_a >> kr.m >> kr.s >> kr.h;
            for (int b = 0; b < kr.beta_size; b++) is >> kr.b[b];
            init_kernel(&kr);
            continue;
        }
        double v;
        is >> v;
        if (k == "mode") MODE_TYPE = (ModeType)(int)v;
        else if (k == "border") BORDER = (Border)(int)v;
        else if (k == "sobel_border") SOBEL_BORDER = (Border)(int)v;
        else if (k == "gradient") GRADIENT_MODE = (GradientType)(int)v;
        else if (k == "growth") GROWTH_TYPE = (GrowthType)(int)v;
        else if (k == "mu_vel") MU_VEL = v != 0;
        else if (k == "adv_vel") ADV_VEL = v != 0;
        else if (k == "blur") BLUR_CENTRAL_DIFF = v != 0;
        else if (k == "dt") timestep = (float)v;
        else if (k == "sigma") off_table.sigma = (float)v;
        else if (k == "dd") off_table.dd = (float)v;
        else if (k == "s_norm") S_NORM = (float)v;
        else if (k == "theta_a") THETA_A = (float)v;
        else if (k == "gta") GTA = (float)v;
        else if (k == "sta") STA = (float)v;
        else if (k == "v_damp") V_DAMP = (float)v;
        else if (k == "channels") channels = (int)v;
        else if (k.rfind("asymptotic", 0) == 0) ch[k.back() - '0'].asymptotic = v != 0;
        else if (k.rfind("soft_clip", 0) == 0) ch[k.back() - '0'].soft_clip = v != 0;
        else {
            std::fprintf(stderr, "unknown key %s\n", k.c_str());
            return false;
        }
    }
    return true;
}

int main(int argc, char **argv) {
    if (argc != 5) {
        std::fprintf(stderr, "usage: %s config.txt state_in.f32 nsteps out.f32\n", argv[0]);
        return 2;
    }
    off_table.dd = 1.0f;
    off_table.sigma = 0.5f;
    if (!read_config(argv[1])) return 2;
    init_pos_array();
    init_offset_table(nullptr);
    init_bell_lut();
    init_fftw_plans();
    const int C = channels;
    {
        std::vector<float> a((size_t)C * GRID_SIZE);
        FILE *f = std::fopen(argv[2], "rb");
        if (!f || std::fread(a.data(), sizeof(float), a.size(), f) != a.size()) return 2;
        std::fclose(f);
        for (int c = 0; c < C; c++)
            for (int z = 0; z < GRID_WIDTH; z++)
                for (int y = 0; y < GRID_HEIGHT; y++)
                    for (int x = 0; x < GRID_WIDTH; x++)
                        GRID[CPOFFSET(c) + PIDX(x, y, z)] = a[(size_t)c * GRID_SIZE + IDX(x, y, z)];
        lenia_shim_state_changed();
    }
    const int nsteps = std::atoi(argv[3]);
    Dispatcher d;
    TD td[SIM_THREADS];
    std::vector<std::thread> th;
    for (int i = 0; i < SIM_THREADS; i++) {
        td[i].id = i;
        td[i].is_done = true;
        th.emplace_back(sim_thread, &d, &td[i]);
    }
    for (int s = 0; s < nsteps; s++) {
        for (int ev = 0; ev < PHASES; ev++) {
            std::unique_lock<std::mutex> lk(d.mu);
            d.phase = ev;
            d.running = SIM_THREADS;
            d.gen++;
            d.work.notify_all();
            d.done.wait(lk, [&] { return d.running == 0; });
        }
        // the main thread's copy of GRID into the mesh back buffer (main.cpp:225-236)
        for (int c = 0; c < MAX_CHANNELS; c++)
            for (int z = 0; z < GRID_WIDTH; z++)
                for (int y = 0; y < GRID_HEIGHT; y++)
                    for (int x = 0; x < GRID_WIDTH; x++) RBB[CPOFFSET(c) + PIDX(x, y, z)] = GRID[CPOFFSET(c) + PIDX(x, y, z)];
    }
    {
        std::lock_guard<std::mutex> lk(d.mu);
        d.quit = true;
        d.work.notify_all();
    }
    for (auto &t : th) t.join();
    if (*lenia_shim_last_error()) {
        std::fprintf(stderr, "shim error: %s\n", lenia_shim_last_error());
        return 1;
    }
    FILE *f = std::fopen(argv[4], "wb");
    if (!f) return 2;
    std::vector<float> a((size_t)C * GRID_SIZE);
    for (const float *src : {(const float *)GRID, (const float *)H, (const float *)U_SUM}) {
        for (int c = 0; c < C; c++)
            for (int z = 0; z < GRID_WIDTH; z++)
                for (int y = 0; y < GRID_HEIGHT; y++)
                    for (int x = 0; x < GRID_WIDTH; x++)
                        a[(size_t)c * GRID_SIZE + IDX(x, y, z)] = src[CPOFFSET(c) + PIDX(x, y, z)];
        std::fwrite(a.data(), sizeof(float), a.size(), f);
    }
    for (const float *src : {(const float *)RBB, (const float *)H_RB, (const float *)U_SUM_RB})
        std::fwrite(src, sizeof(float), (size_t)C * GRID_PADDED, f);
    std::fclose(f);
    lenia_shim_shutdown();
    return 0;
}
