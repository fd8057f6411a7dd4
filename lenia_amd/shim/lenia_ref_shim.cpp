// lenia_ref_shim.cpp -- the reference's step entry points over the liblenia_hip C-ABI.
//
// One engine holds one world of GRID_WIDTH x GRID_HEIGHT x GRID_WIDTH cells whose
// parameters come from the reference's globals (main.cpp:30-66, kernels[], ch[],
// off_table).  Parameters are compared with the last committed copy at every fft_step;
// a change rebuilds the engine world (the state is carried over).  See
// include/lenia_ref_shim.h for the phase mapping.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "lenia_hip.h"
#include "lenia_ref_shim.h"

// ---- state arrays the reference's lenia.cpp defines (the mirrored subset) ------------
float GRID[GRID_PADDED * MAX_CHANNELS];
float H[GRID_PADDED * MAX_CHANNELS];
float U_SUM[GRID_PADDED * MAX_CHANNELS];
float RBB[GRID_PADDED * MAX_CHANNELS];
float H_RB[GRID_PADDED * MAX_CHANNELS];
float U_SUM_RB[GRID_PADDED * MAX_CHANNELS];
float GA[GRID_SIZE * MAX_CHANNELS];
float SUM[GRID_PADDED];
float V_x[GRID_PADDED * MAX_CHANNELS], V_y[GRID_PADDED * MAX_CHANNELS], V_z[GRID_PADDED * MAX_CHANNELS];
float POS_ARRAY_X[GRID_SIZE], POS_ARRAY_Y[GRID_SIZE], POS_ARRAY_Z[GRID_SIZE];

namespace {

struct Snapshot {
    lh_params p;
    std::vector<lh_kernel> k;
    int C = 0;
    bool operator==(const Snapshot &o) const {
        return C == o.C && std::memcmp(&p, &o.p, sizeof(p)) == 0 && k.size() == o.k.size() &&
               (k.empty() || std::memcmp(k.data(), o.k.data(), k.size() * sizeof(lh_kernel)) == 0);
    }
};

std::mutex g_mu;
void *g_eng = nullptr;
int g_world = -1;
int g_device = 0;
// host fields the caller changed since the last upload: bit 0 A, 1 GA, 2 S, 3 V
enum { D_A = 1, D_GA = 2, D_S = 4, D_V = 8, D_ALL = 15 };
int g_dirty = D_ALL;
Snapshot g_committed;
bool g_have_committed = false;
std::string g_err;

void note(const char *what) {
    g_err = std::string(what) + ": " + lh_last_error();
    std::fprintf(stderr, "lenia_ref_shim: %s\n", g_err.c_str());
}

Snapshot current_params() {
    Snapshot s;
    std::memset(&s.p, 0, sizeof(s.p));
    lh_params &p = s.p;
    p.mode = (int)MODE_TYPE;
    p.border = (int)BORDER;
    p.sobel_border = (int)SOBEL_BORDER;
    p.gradient = (int)GRADIENT_MODE;
    p.growth_type = (int)GROWTH_TYPE;
    p.mu_vel = MU_VEL;
    p.adv_vel = ADV_VEL;
    p.blur = BLUR_CENTRAL_DIFF;
    p.dt = timestep;
    p.sigma = off_table.sigma;
    p.dd = off_table.dd;
    p.s_norm = S_NORM;
    p.theta_a = THETA_A;
    p.gta = GTA;
    p.sta = STA;
    p.v_damp = V_DAMP;
    for (int c = 0; c < MAX_CHANNELS && c < LH_MAX_C; c++) {
        p.asymptotic[c] = ch[c].asymptotic;
        p.soft_clip[c] = ch[c].soft_clip;
    }
    p.osc_la = OSC_LA;
    p.osc_flow = OSC_FLOW;
    p.osc_div_flow = OSC_DIV_FLOW;
    p.osc_vel = OSC_VEL;
    p.osc_la_amt = OSC_LA_AMT;
    p.osc_flow_amt = OSC_FLOW_AMT;
    p.osc_div_flow_amt = OSC_DIV_FLOW_AMT;
    p.quirks = LH_Q_ALL;
    s.C = channels;
    for (int i = 0; i < kernel_count && i < MAX_KERNELS; i++) {
        lh_kernel k;
        std::memset(&k, 0, sizeof(k));
        const Kernel &r = kernels[i];
        k.c0 = r.c0;
        k.c1 = r.c1;
        k.radius = r.radius;
        k.beta_size = r.beta_size;
        for (int b = 0; b < MAX_BETA && b < LH_MAX_BETA; b++) k.b[b] = r.b[b];
        k.r_a = r.r_a;
        k.m = r.m;
        k.s = r.s;
        k.h = r.h;
        s.k.push_back(k);
    }
    return s;
}

// interior of a padded channel array <-> unpadded [C][N]
void pack(const float *padded, int C, float *out) {
    for (int c = 0; c < C; c++)
        for (int z = 0; z < GRID_WIDTH; z++)
            for (int y = 0; y < GRID_HEIGHT; y++)
                for (int x = 0; x < GRID_WIDTH; x++)
                    out[(size_t)c * GRID_SIZE + IDX(x, y, z)] = padded[CPOFFSET(c) + PIDX(x, y, z)];
}
void unpack(const float *in, int C, float *padded) {
    for (int c = 0; c < C; c++)
        for (int z = 0; z < GRID_WIDTH; z++)
            for (int y = 0; y < GRID_HEIGHT; y++)
                for (int x = 0; x < GRID_WIDTH; x++)
                    padded[CPOFFSET(c) + PIDX(x, y, z)] = in[(size_t)c * GRID_SIZE + IDX(x, y, z)];
}

bool upload_state(int C, int mask) {
    std::vector<float> buf((size_t)3 * C * GRID_SIZE);
    if (mask & D_A) {
        pack(GRID, C, buf.data());
        if (lh_set_field(g_eng, g_world, LH_F_A, buf.data(), (size_t)C * GRID_SIZE) != LH_OK) return note("A"), false;
    }
    if ((mask & D_GA) && lh_set_field(g_eng, g_world, LH_F_GA, GA, (size_t)C * GRID_SIZE) != LH_OK)
        return note("GA"), false;
    if (mask & D_S) {
        pack(SUM, 1, buf.data());
        if (lh_set_field(g_eng, g_world, LH_F_S, buf.data(), GRID_SIZE) != LH_OK) return note("S"), false;
    }
    if (!(mask & D_V)) return true;
    const float *V[3] = {V_x, V_y, V_z};
    for (int c = 0; c < C; c++)
        for (int a = 0; a < 3; a++) {
            for (int z = 0; z < GRID_WIDTH; z++)
                for (int y = 0; y < GRID_HEIGHT; y++)
                    for (int x = 0; x < GRID_WIDTH; x++)
                        buf[((size_t)c * 3 + a) * GRID_SIZE + IDX(x, y, z)] = V[a][CPOFFSET(c) + PIDX(x, y, z)];
        }
    if (lh_set_field(g_eng, g_world, LH_F_V, buf.data(), (size_t)3 * C * GRID_SIZE) != LH_OK) return note("V"), false;
    return true;
}

bool download_state(int C) {
    std::vector<float> buf((size_t)3 * C * GRID_SIZE);
    if (lh_get_field(g_eng, g_world, LH_F_A, buf.data(), (size_t)C * GRID_SIZE) != LH_OK) return note("A"), false;
    unpack(buf.data(), C, GRID);
    if (lh_get_field(g_eng, g_world, LH_F_GA, GA, (size_t)C * GRID_SIZE) != LH_OK) return note("GA"), false;
    if (lh_get_field(g_eng, g_world, LH_F_S, buf.data(), GRID_SIZE) != LH_OK) return note("S"), false;
    unpack(buf.data(), 1, SUM);
    if (lh_get_field(g_eng, g_world, LH_F_V, buf.data(), (size_t)3 * C * GRID_SIZE) != LH_OK) return note("V"), false;
    float *V[3] = {V_x, V_y, V_z};
    for (int c = 0; c < C; c++)
        for (int a = 0; a < 3; a++)
            for (int z = 0; z < GRID_WIDTH; z++)
                for (int y = 0; y < GRID_HEIGHT; y++)
                    for (int x = 0; x < GRID_WIDTH; x++)
                        V[a][CPOFFSET(c) + PIDX(x, y, z)] = buf[((size_t)c * 3 + a) * GRID_SIZE + IDX(x, y, z)];
    return true;
}

// (Re)builds the engine world when parameters changed; carries the state over.
bool ensure_engine() {
    Snapshot now = current_params();
    if (g_eng && g_have_committed && now == g_committed) return true;
    // carry the device state over a rebuild (fields the caller changed keep the host copy)
    if (g_eng && g_have_committed) {
        std::vector<float> keepA(sizeof(GRID) / sizeof(float));
        if (g_dirty & D_A) std::memcpy(keepA.data(), GRID, sizeof(GRID));
        const int dirty = g_dirty;
        if (!download_state(g_committed.C)) return false;
        if (dirty & D_A) std::memcpy(GRID, keepA.data(), sizeof(GRID));
    }
    if (g_eng) {
        lh_destroy(g_eng);
        g_eng = nullptr;
    }
    if (lh_create(g_device, &g_eng) != LH_OK) return note("lh_create"), false;
    lh_set_option(g_eng, LH_OPT_KEEP_VELOCITY, 1);
    lh_set_option(g_eng, LH_OPT_KEEP_USUM, 1);
    lh_shape shape{GRID_WIDTH, GRID_HEIGHT, GRID_WIDTH, now.C, (int)now.k.size()};
    if (lh_add_world(g_eng, &shape, &now.p, now.k.data(), &g_world) != LH_OK) return note("lh_add_world"), false;
    if (lh_commit(g_eng) != LH_OK) return note("lh_commit"), false;
    g_committed = now;
    g_have_committed = true;
    g_dirty = D_ALL;
    return true;
}

}  // namespace

// ---- lenia.h ---------------------------------------------------------------------------
void init_fftw_plans() {}

void init_kernel(Kernel *) {}
void init_bell_lut() {}
void init_gaussian_kernel() {}

void init_pos_array() {
    for (int x = 0; x < GRID_WIDTH; x++)
        for (int y = 0; y < GRID_HEIGHT; y++)
            for (int z = 0; z < GRID_WIDTH; z++) {
                POS_ARRAY_X[IDX(x, y, z)] = x + 0.5f;
                POS_ARRAY_Y[IDX(x, y, z)] = y + 0.5f;
                POS_ARRAY_Z[IDX(x, y, z)] = z + 0.5f;
            }
}

void init_offset_table(Button *) { off_table.ma = off_table.dd - off_table.sigma; }

void init_conway() {
    for (int c = 0; c < channels; c++)
        for (int z = 0; z < GRID_WIDTH; z++)
            for (int y = 0; y < GRID_HEIGHT; y++)
                for (int x = 0; x < GRID_WIDTH; x++)
                    GRID[PIDX(x, y, z) + CPOFFSET(c)] = (float)(std::rand() % 1000) / 1000;
    std::lock_guard<std::mutex> lk(g_mu);
    g_dirty |= D_A;
}

// Kernel parameters as the reference's randomiser draws them (lenia.cpp:370-412) and
// its gaussian blob seeding (lenia.cpp:305-334); the kernel-mesh rebuild is render-side.
void generate_random() {
    neighborhood = 1 + (std::rand() % 8);
    kernel_count = std::rand() % 10 + 1;
    for (int i = 0; i < MAX_CHANNELS; i++) ch[i].asymptotic = ch[i].soft_clip = false;
    for (int i = 0; i < kernel_count; i++) {
        Kernel *k = &kernels[i];
        k->beta_size = std::rand() % 2 + 1;
        k->radius = (int)neighborhood;
        for (int e = 0; e < k->beta_size; e++) k->b[e] = (float)(std::rand() % 100) / 100;
        k->m = (float)(100 + std::rand() % 500) / 1000;
        k->s = (k->m / 10) + (float)(std::rand() % 200) / 1000;
        if (channels == 1 || i == 0) k->c0 = k->c1 = 0;
        else {
            k->c0 = std::rand() % channels;
            k->c1 = std::rand() % channels;
        }
        k->h = 1;
        k->r_a = 1;
    }
    std::memset(GRID, 0, sizeof(GRID));
    std::memset(GA, 0, sizeof(GA));
    std::memset(V_x, 0, sizeof(V_x));
    std::memset(V_y, 0, sizeof(V_y));
    std::memset(V_z, 0, sizeof(V_z));
    const int r = (int)neighborhood;
    for (int i = 0; i < 5; i++) {
        const int cx = std::rand() % GRID_WIDTH, cy = std::rand() % GRID_HEIGHT, cz = std::rand() % GRID_WIDTH;
        for (int c = 0; c < channels; c++)
            for (int z = 0; z < GRID_WIDTH; z++)
                for (int y = 0; y < GRID_HEIGHT; y++)
                    for (int x = 0; x < GRID_WIDTH; x++) {
                        const int dx = x - cx, dy = y - cy, dz = z - cz;
                        const float dist = (float)(dx * dx + dy * dy + dz * dz);
                        GRID[PIDX(x, y, z) + CPOFFSET(c)] += (float)std::exp(-dist / (2 * r * r));
                    }
    }
    std::lock_guard<std::mutex> lk(g_mu);
    g_dirty |= D_A | D_GA | D_V;
}

void fft_step(TD *t) {
    if (t->id == 0) {
        std::lock_guard<std::mutex> lk(g_mu);
        // mesh copies of the previous step's H and U_SUM (lenia.cpp:453-457)
        for (int c = 0; c < channels; c++)
            for (int z = 0; z < GRID_WIDTH; z++)
                for (int y = 0; y < GRID_HEIGHT; y++)
                    for (int x = 0; x < GRID_WIDTH; x++) {
                        const int q = CPOFFSET(c) + PIDX(x, y, z);
                        U_SUM_RB[q] = U_SUM[q];
                        H_RB[q] = H[q];
                    }
        if (ensure_engine()) {
            if (g_dirty && upload_state(g_committed.C, g_dirty)) g_dirty = 0;
            if (!g_dirty && lh_step(g_eng, 1) != LH_OK) note("lh_step");
        }
    }
    t->is_done = true;
}

void mt_fft_step(TD *t) { t->is_done = true; }
void sobel_convolve(TD *t) { t->is_done = true; }
void sobel_convolve_h(TD *t) { t->is_done = true; }
void compute_f(TD *t) { t->is_done = true; }
void reintegration_step(TD *t) { t->is_done = true; }
void transpose_sobel_borders(TD *t) { t->is_done = true; }
void transpose_borders(TD *t) { t->is_done = true; }

void transpose_to_row_major(TD *t) {
    if (t->id == 0) {
        std::lock_guard<std::mutex> lk(g_mu);
        if (g_eng && !g_dirty) {
            const int C = g_committed.C;
            std::vector<float> buf((size_t)C * GRID_SIZE);
            if (lh_get_field(g_eng, g_world, LH_F_A, buf.data(), buf.size()) == LH_OK) unpack(buf.data(), C, GRID);
            else note("A");
            if (lh_get_field(g_eng, g_world, LH_F_H, buf.data(), buf.size()) == LH_OK) unpack(buf.data(), C, H);
            else note("H");
            if (lh_get_field(g_eng, g_world, LH_F_USUM, buf.data(), buf.size()) == LH_OK) unpack(buf.data(), C, U_SUM);
            else note("U_SUM");
        }
    }
    t->is_done = true;
}

// transposition.cpp:881-995: torus face pads of the mesh read-backs, split by z-slab
// between the two sim threads exactly as the reference splits them.
void transpose_mesh_borders(TD *t) {
    const int section = (GRID_WIDTH / SIM_THREADS) * t->id;
    const int end = t->id == SIM_THREADS - 1 ? GRID_WIDTH : section + GRID_WIDTH / SIM_THREADS;
    auto copy = [](int dst, int src) {
        RBB[dst] = GRID[src];
        H_RB[dst] = H[src];
        U_SUM_RB[dst] = U_SUM[src];
    };
    for (int c = 0; c < channels; c++) {
        const int o = CPOFFSET(c);
        for (int z = section; z < end; z++)
            for (int y = 0; y < GRID_HEIGHT; y++) {
                copy(o + PIDX(-1, y, z), o + PIDX(GRID_WIDTH - 1, y, z));
                copy(o + PIDX(GRID_WIDTH, y, z), o + PIDX(0, y, z));
            }
        for (int z = section; z < end; z++)
            for (int x = 0; x < GRID_WIDTH; x++) {
                copy(o + PIDX(x, -1, z), o + PIDX(x, GRID_HEIGHT - 1, z));
                copy(o + PIDX(x, GRID_HEIGHT, z), o + PIDX(x, 0, z));
            }
        for (int y = 0; y < GRID_HEIGHT; y++)
            for (int x = 0; x < GRID_WIDTH; x++) {
                if (t->id == 0) copy(o + PIDX(x, y, -1), o + PIDX(x, y, GRID_WIDTH - 1));
                if (t->id == 1) copy(o + PIDX(x, y, GRID_WIDTH), o + PIDX(x, y, 0));
            }
    }
    t->is_done = true;
}

// ---- reintegration.h ----------------------------------------------------------------------
void precompute_blur_kernel(Kernel *) {}
void h_sum_blur(TD *t) { t->is_done = true; }
void compute_velocity_mu_wall(TD *t) { t->is_done = true; }
void compute_velocity_mu(TD *t) { t->is_done = true; }
void compute_f_mu_wall(TD *t) { t->is_done = true; }
void compute_f_mu(TD *t) { t->is_done = true; }
void srt(float *, int, int, int, float, float) {}
void rt_border_torus(float *, int, int, int, float, float) {}
void rt_border_other(float *, int, int, int, float, float) {}
void adv_vel(float *, int, int, int, float, float) {}
void adv_f(float *, int, int, int, float, float) {}

// ---- transposition.h -----------------------------------------------------------------------
void transpose_central_diff_borders(TD *t) { t->is_done = true; }
void wall_central_diff_borders(TD *t) { t->is_done = true; }
void zero_central_diff_borders(TD *t) { t->is_done = true; }
void wall_sobel_borders(TD *t) { t->is_done = true; }
void zero_sobel_borders(TD *t) { t->is_done = true; }

// ---- shim control ------------------------------------------------------------------------------
void lenia_shim_set_device(int device) { g_device = device; }

void lenia_shim_state_changed() {
    std::lock_guard<std::mutex> lk(g_mu);
    g_dirty = D_ALL;
}

void lenia_shim_shutdown() {
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_eng) lh_destroy(g_eng);
    g_eng = nullptr;
    g_have_committed = false;
    g_dirty = D_ALL;
}

const char *lenia_shim_last_error() { return g_err.c_str(); }
