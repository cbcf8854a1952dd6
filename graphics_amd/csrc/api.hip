// C-ABI of libsphgpu.so: handle lifecycle, scene upload, the per-step
// drivers of the three solvers (the reference's sphStep overrides) and
// readback.  All device work runs on the handle's single HIP stream.
#include "io.h"
#include "mlp.h"
#include "sph_common.h"
#include "sphgpu.h"

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <functional>
#include <numeric>
#include <string>
#include <vector>

using namespace sphgpu;

namespace {

struct SceneDesc {
    float mn[3], mx[3];
    bool set = false;
};

struct Phase {
    float rho0, visc;
};

struct GraphEntry {
    int seg, iters;
    Dev before, after;
    hipGraphExec_t exec;
    unsigned long long used;
};

struct ProfEvent {
    hipEvent_t a, b;
    double bytes;
    int iter;      // solver iteration the launch belongs to (-1 outside loops)
    bool skip;     // launched past convergence: no particle was processed
};

}  // namespace

struct sphgpu_handle {
    sphgpu_params P{};
    std::string err;
    std::vector<SceneDesc> scenes;
    std::vector<Phase> phases;
    // host staging (add order)
    std::vector<float> fx, fv;
    std::vector<float> fbeta;          // per-particle beta (IISph DF_*), empty = all 1
    std::vector<int> fphase, fscene;
    std::vector<float> bx;
    std::vector<int> bsolid, bscene;
    std::vector<float> solidVisc;
    std::vector<SolidPose> hsolids;    // per solid: pose, viscosity, dynamic flag, id range
    bool posesDirty = false;
    sphgpu_update_solids_fn updateFn = nullptr;   // host rigid-body step inside sphgpu_step
    void* updateUser = nullptr;
    bool inCallback = false;
    // frame-output capture: device copies on st, transfer to pinned memory on cs
    struct Snap {
        float4 *x = nullptr, *v = nullptr, *a = nullptr;
        float *rho = nullptr, *p = nullptr;
        int *id = nullptr, *tv = nullptr;
        void* host = nullptr;
        float4 *hx = nullptr, *hv = nullptr, *ha = nullptr;
        float *hrho = nullptr, *hp = nullptr;
        int *hid = nullptr, *htv = nullptr;
        int npad = 0;
        long long frame = 0;
        bool pending = false, valid = false;
        hipStream_t cs = nullptr;
        hipEvent_t copied = nullptr, done = nullptr;
    } snap;
    SolidPose* dSolids = nullptr;
    double* dSolidFT = nullptr;        // [nsolids][6] force, torque of the last step
    std::vector<double> hSolidFT;      // host copy of dSolidFT, read once per step
    bool ftValid = false;
    long long nextId = 0;

    bool initialized = false;
    int frame = 0;
    hipStream_t st = nullptr;
    Dev D{};
    std::vector<SceneGrid> hgrids;
    std::vector<uint32_t> hKeyBase;
    int keyBits = 1;
    long long nLive = 0;
    std::vector<void*> allocs;
    // sort
    SortScratch sort{};
    uint32_t* sk = nullptr;
    uint32_t* sv = nullptr;
    // small device buffers
    int* layoutOut = nullptr;          // [npad_new, live]
    double* statsPart = nullptr;       // per tile 4
    double* statsOut = nullptr;        // per scene 4
    // pinned host mirrors
    int* hActive = nullptr;
    int* hIters = nullptr;
    double* hErr = nullptr;
    int* hCounters = nullptr;
    double* hStats = nullptr;
    int* hLayout = nullptr;
    float* hDt = nullptr;
    std::vector<sphgpu_stats> sceneStats;
    // learned warm start of the divergence solve
    MlpDev mlp{};
    bool mlpLive = false;
    std::vector<float> mlpHost;        // parameters before init (or set by the caller)
    bool mlpPreset = false;            // caller-provided parameters: use them from the first step
    float* feat = nullptr;             // ncap x 8 features, per slot
    float* tgt = nullptr;              // ncap converged stiffness, per slot
    float* ypred = nullptr;            // ncap predictions
    float* bfeat = nullptr;            // learn_batch x 8 minibatch
    float* btgt = nullptr;
    int learned = 0;                   // training steps done
    int lastIters[2] = {0, 0};         // iterations of the previous solve (density, divergence)
    int curIter = -1;                  // iteration being launched (profiling)
    // captured step segments (use_graph)
    std::vector<GraphEntry> graphs;
    unsigned long long graphClock = 0;
    bool capturing = false;
    bool tuned = false;
    int tuneFix[3] = {-1, -1, -1};   // fixed tile map, staging depth, streaming lists (-1: timed)
    bool listsBuilt = false;   // neighbour lists built at least once
    bool needSearch = false;   // particles appended since the last neighbour search (DFSPH searches first)
    long long appended = 0;    // particles appended after init
    float tuneMs[6] = {0, 0, 0, 0, 0, 0};   // density pass: map 0, 1; list builder: 6, 8 groups;
                                            // density pass: plain, streaming list loads
    int winFix = -1;                        // windows fixed by sphgpu_set_lds_windows (-1: per pass, timed)
    float winMs[10] = {};                   // per window pass: gathers, then window (ms per launch)
    uint32_t gatherSeed = 1;
    // profiling
    std::string profName;
    std::vector<ProfEvent> prof;
    size_t profUsed = 0;
};

namespace {

int fail(sphgpu_handle* h, int code, const char* fmt, ...) {
    if (h) {
        char buf[512];
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(buf, sizeof(buf), fmt, ap);
        va_end(ap);
        h->err = buf;
    }
    return code;
}

#define HIPCHECK(h, expr)                                                                   \
    do {                                                                                    \
        hipError_t e_ = (expr);                                                             \
        if (e_ != hipSuccess)                                                               \
            return fail((h), SPHGPU_E_DEVICE, "%s failed: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

template <class T>
int dalloc(sphgpu_handle* h, T** p, size_t count) {
    void* m = nullptr;
    size_t bytes = std::max<size_t>(count * sizeof(T), 256);
    hipError_t e = hipMalloc(&m, bytes);
    if (e != hipSuccess) return fail(h, SPHGPU_E_DEVICE, "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
    h->allocs.push_back(m);
    *p = static_cast<T*>(m);
    return 0;
}

uint32_t spread2(uint32_t v) {
    v &= 0xFFFFu;
    v = (v | (v << 8)) & 0x00FF00FFu;
    v = (v | (v << 4)) & 0x0F0F0F0Fu;
    v = (v | (v << 2)) & 0x33333333u;
    v = (v | (v << 1)) & 0x55555555u;
    return v;
}

// ---- profiling helpers
struct ProfScope {
    sphgpu_handle* h;
    bool on;
    ProfScope(sphgpu_handle* hh, const char* name, double bytes) : h(hh), on(false) {
        if (h->capturing || h->profName.empty() || h->profName != name) return;
        if (h->profUsed == h->prof.size()) {
            ProfEvent e{};
            if (hipEventCreate(&e.a) != hipSuccess || hipEventCreate(&e.b) != hipSuccess) return;
            h->prof.push_back(e);
        }
        ProfEvent& e = h->prof[h->profUsed];
        e.bytes = bytes;
        e.iter = h->curIter;
        e.skip = false;
        hipEventRecord(e.a, h->st);
        // the kernel itself records both events at its dispatch boundaries
        // (hipExtLaunchKernelGGL): kernel time, without the event packets
        g_launch_timing = LaunchTiming{e.a, e.b};
        on = true;
    }
    ~ProfScope() {
        if (!on) return;
        if (g_launch_timing.start) {   // not a SPH_LAUNCH kernel: bracket with stream events
            g_launch_timing = LaunchTiming{};
            hipEventRecord(h->prof[h->profUsed].b, h->st);
        }
        ++h->profUsed;
    }
};

double nf_live(const sphgpu_handle* h) { return double(h->nLive); }

// Glorot weights with a zero output layer: the warm-start model predicts a
// change from the previous stiffness, so an untrained network predicts none.
void init_handle_mlp(std::vector<float>& p) {
    mlp_init_params(p, 0x5eedULL);
    for (int k = MLP_PARAMS - 65; k < MLP_PARAMS; ++k) p[k] = 0.0f;   // W4[1][64], b4
}

// ---- neighbour search pipeline ------------------------------------------
int run_ns(sphgpu_handle* h) {
    Dev& D = h->D;
    h->listsBuilt = true;
    h->needSearch = false;
    hipStream_t st = h->st;
    int npad = D.npad;
    {
        ProfScope ps(h, "ns_keys", nf_live(h) * 20.0);
        launch_fluid_keys_into(D, h->sk, h->sv, st);
    }
    {
        ProfScope ps(h, "ns_sort", nf_live(h) * 16.0 * ((h->keyBits + 7) / 8));
        radix_sort_pairs(h->sk, h->sv, npad, h->keyBits, h->sort, st);
    }
    launch_fill_i32(D.sceneStart, -1, D.nscenes, st);
    launch_fill_i32(D.sceneCount, 0, D.nscenes, st);
    launch_scene_bounds(D, h->sk, npad, st);
    launch_scene_layout(D, h->layoutOut, st);
    {
        ProfScope ps(h, "ns_reorder", nf_live(h) * 64.0);
        launch_reorder(D, h->sk, h->sv, npad, npad, st);
    }
    std::swap(D.x, D.x2);
    std::swap(D.v, D.v2);
    std::swap(D.id, D.id2);
    std::swap(D.ksum, D.ksum2);
    std::swap(D.kvsum, D.kvsum2);
    std::swap(D.key, D.key2);
    launch_cell_reset(D, D.key2, npad, D.cellF, st);
    launch_cell_build(D, D.key, npad, D.cellF, st);
    {
        ProfScope ps(h, "ns_neighbours", nf_live(h) * 12.0);
        launch_neighbours(D, st);
    }
    return 0;
}

// `odd`: the flags to read are those of an odd iteration (the second buffer
// of the DFSPH loops)
int sync_active(sphgpu_handle* h, bool* any, bool odd = false) {
    Dev& D = h->D;
    int ns = D.nscenes;
    // flags, iteration counts and errors: one block, same layout on both sides
    HIPCHECK(h, hipMemcpyAsync(h->hActive, D.active, 24 * (size_t)ns, hipMemcpyDeviceToHost, h->st));
    HIPCHECK(h, hipStreamSynchronize(h->st));
    const int* flags = odd ? h->hActive + ns : h->hActive;
    bool a = false;
    for (int s = 0; s < ns; ++s) a = a || flags[s] != 0;
    *any = a;
    return 0;
}

// one solver loop; records per-scene iterations/errors into field `which`
// (0 density, 1 divergence) of sceneStats.
//
// Iterations are launched speculatively: every pass and the reduction skip
// scenes whose device-side `active` flag is cleared, so iterations launched
// past convergence cost only their launches.  The host reads the flags
// after a first batch sized from the previous step's count, then after
// every second iteration.
void launch_iteration(sphgpu_handle* h, int which, int it) {
    Dev& D = h->D;
    const sphgpu_params& P = h->P;
    h->curIter = it;
    {
        if (P.solver == SPHGPU_DFSPH) {
            // iteration `it` gates on flag buffer it & 1; the velocity pass
            // tests convergence on the partials pass A left and writes the
            // other buffer for the next iteration (a separate reduction
            // kernel when a scene has no tile to host its test)
            Dev Dq = D;
            Dq.active = (it & 1) ? D.active1 : D.active;
            const bool fused = D.npad / TILE >= D.nscenes;
            Dq.activeNext = (it & 1) ? D.active : D.active1;
            if (fused) {
                Dq.redMode = which == 0 ? 0 : 1;
                Dq.redTol = which == 0 ? P.density_tol : P.divergence_tol;
                Dq.redMinIt = which == 0 ? P.min_iter_density : P.min_iter_div;
                Dq.redCap = which == 0 ? P.max_iter_density : P.max_iter_div + 1;
            }
            if (which == 0) {
                ProfScope ps(h, "df_pass_density", nf_live(h) * 36.0);
                launch_df_pass_density(Dq, h->st);
            } else {
                ProfScope ps(h, "df_pass_divergence", nf_live(h) * 36.0);
                launch_df_pass_divergence(Dq, P.warm_start != SPHGPU_WARM_NONE && it == 0, h->st);
            }
            {
                ProfScope ps(h, "df_pass_velocity", nf_live(h) * 44.0);
                launch_df_pass_velocity(Dq, h->st);
            }
            if (!fused)
                launch_scene_reduce(Dq, which == 0 ? 0 : 1, which == 0 ? P.density_tol : P.divergence_tol,
                                    which == 0 ? P.min_iter_density : P.min_iter_div,
                                    which == 0 ? P.max_iter_density : P.max_iter_div + 1, h->st);
        } else if (P.solver == SPHGPU_IISPH_DF) {
            // IISph.cpp:840-1004 (which 0) and 1062-1207 (which 1): the loops
            // break when the iteration count exceeds the maximum
            {
                ProfScope ps(h, which == 0 ? "idf_kappa_density" : "idf_kappa_divergence", nf_live(h) * 36.0);
                launch_idf_kappa(D, which == 0 ? 0 : 1, h->st);
            }
            {
                ProfScope ps(h, "idf_velocity", nf_live(h) * 44.0);
                launch_idf_velocity(D, h->st);
            }
            if (which == 0)
                launch_scene_reduce(D, 1, P.density_tol, P.min_iter_density, P.max_iter_density + 1, h->st);
            else
                launch_scene_reduce(D, 1, P.divergence_tol, P.min_iter_div, P.max_iter_div + 1, h->st);
        } else if (P.solver == SPHGPU_PCISPH) {
            {
                ProfScope ps(h, "pci_density", nf_live(h) * 44.0);
                launch_pci_density(D, h->st);
            }
            {
                ProfScope ps(h, "pci_pressure", nf_live(h) * 64.0);
                launch_pci_pressure(D, h->st);
            }
            launch_scene_reduce(D, 2, P.density_tol, P.min_iter_density,
                                P.max_iter_density, h->st);
        } else {  // IISPH
            {
                ProfScope ps(h, "ii_sumdij", nf_live(h) * 32.0);
                launch_ii_sumdij(D, h->st);
            }
            {
                ProfScope ps(h, "ii_pressure", nf_live(h) * 56.0);
                launch_ii_pressure(D, h->st);
            }
            launch_scene_reduce(D, 1, P.density_tol, P.min_iter_density,
                                P.max_iter_density, h->st);
        }
    }
}

// first batch of a solver loop: resets the per-scene flags and launches
// `batch` iterations without reading anything back (capturable)
int first_batch(const sphgpu_handle* h, int which) {
    const bool divLoop = which == 1 && (h->P.solver == SPHGPU_DFSPH || h->P.solver == SPHGPU_IISPH_DF);
    const int minIt = divLoop ? h->P.min_iter_div : h->P.min_iter_density;
    return std::max({1, minIt, h->lastIters[which] - 1});
}

void solver_loop_start(sphgpu_handle* h, int which, int batch) {
    Dev& D = h->D;
    launch_loop_start(D, h->st);
    for (int q = 0; q < batch; ++q) launch_iteration(h, which, q);
}

// rest of a solver loop whose first `launched` iterations are queued:
// reads the flags, adds iterations two at a time until every scene is done
int solver_loop_finish(sphgpu_handle* h, int which, int launched, size_t profStart, int* extra = nullptr) {
    Dev& D = h->D;
    int ns = D.nscenes;
    int maxIt = 0;
    if (extra) *extra = 0;
    for (;;) {
        bool any = false;
        // the DFSPH loops keep the flag of iteration k in buffer k & 1
        int rc = sync_active(h, &any, h->P.solver == SPHGPU_DFSPH && (launched & 1));
        if (rc) return rc;
        if (!any) break;
        for (int q = 0; q < 2; ++q) launch_iteration(h, which, launched++);
        if (extra) *extra += 2;
    }
    for (int s = 0; s < ns; ++s) maxIt = std::max(maxIt, h->hIters[s]);
    h->lastIters[which] = maxIt;
    h->curIter = -1;
    // launches past the last iteration of every scene processed nothing
    for (size_t k = profStart; k < h->profUsed; ++k) h->prof[k].skip = h->prof[k].iter >= maxIt;
    for (int s = 0; s < ns; ++s) {
        if (which == 0) {
            h->sceneStats[s].iter_density = h->hIters[s];
            h->sceneStats[s].density_err = h->hErr[s];
        } else {
            h->sceneStats[s].iter_divergence = h->hIters[s];
            h->sceneStats[s].divergence_err = h->hErr[s];
        }
    }
    return 0;
}

int solver_loop(sphgpu_handle* h, int which) {
    const size_t profStart = h->profUsed;
    const int batch = first_batch(h, which);
    solver_loop_start(h, which, batch);
    return solver_loop_finish(h, which, batch, profStart);
}

// ---- autotuner --------------------------------------------------------------
// Before the first step, on the scene itself, these choices are timed (three
// runs each, the fastest minimum wins): the tile-to-workgroup mapping
// (contiguous runs of tiles per XCD, or round robin) on the solver's density
// pass, the LDS staging depth of the list builder (6 or 8 groups) on the
// list builder, plain or streaming loads of the lists, and per solver-loop
// pass the LDS window or per-lane gathers (tune_windows).  The first three
// time kernels that recompute their outputs from positions alone; every
// setting gives the same results.  (Padding the
// lists of 16 or 32 lanes to common per-row counts, so that their gathers
// share cache lines, was measured as a third choice: the density pass took
// 0.124-0.127 ms against 0.086 unpadded on the 1M dam, so it was removed.)
// best of three: `prep` (untimed), then `reps` launches of `f`; per launch
float time_launches(sphgpu_handle* h, const std::function<void()>& f, const std::function<void()>& prep = nullptr,
                    int reps = 1, int tries = 3) {
    hipEvent_t a, b;
    if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) return 0.0f;
    float best = 1e30f;
    for (int r = 0; r < tries; ++r) {
        if (prep) prep();
        hipEventRecord(a, h->st);
        for (int q = 0; q < reps; ++q) f();
        hipEventRecord(b, h->st);
        float t = 0.0f;
        if (hipEventSynchronize(b) == hipSuccess && hipEventElapsedTime(&t, a, b) == hipSuccess)
            best = std::min(best, t / reps);
    }
    hipEventDestroy(a);
    hipEventDestroy(b);
    return best;
}

// Per solver-loop pass: stage the tile's neighbour window in LDS (fewer
// cache-line lookups, but 48-72 KB of LDS per block: two or three blocks per
// CU) or gather each neighbour per lane (five waves per SIMD).  Each pass is
// timed both ways on the scene's lists (best of two single launches: six
// forms, twelve launches a pass, about 4 ms per pass at 4M particles, so the
// re-tuning every 200 steps costs well under 0.1 ms a step) and keeps the
// faster form; both give the same bits.  The passes overwrite
// solver state, so everything they write is saved first and put back after.
static_assert(WIN_DENSITY == SPHGPU_WIN_DENSITY && WIN_DIVERGENCE == SPHGPU_WIN_DIVERGENCE &&
                  WIN_VELOCITY == SPHGPU_WIN_VELOCITY && WIN_SUMDIJ == SPHGPU_WIN_SUMDIJ &&
                  WIN_PRESSURE == SPHGPU_WIN_PRESSURE,
              "window bits of the C-ABI and the kernels");
int tune_windows(sphgpu_handle* h) {
    Dev& D = h->D;
    std::vector<std::pair<int, std::function<void()>>> passes;
    if (h->P.solver == SPHGPU_DFSPH) {
        passes.push_back({WIN_DENSITY, [h]() { launch_df_pass_density(h->D, h->st); }});
        passes.push_back({WIN_DIVERGENCE, [h]() { launch_df_pass_divergence(h->D, 0, h->st); }});
        passes.push_back({WIN_VELOCITY, [h]() { launch_df_pass_velocity(h->D, h->st); }});
    } else if (h->P.solver == SPHGPU_IISPH) {
        passes.push_back({WIN_SUMDIJ, [h]() { launch_ii_sumdij(h->D, h->st); }});
        passes.push_back({WIN_PRESSURE, [h]() { launch_ii_pressure(h->D, h->st); }});
    }
    if (passes.empty() || D.npad <= 0) return 0;
    const size_t np = (size_t)D.npad, tiles = np / TILE;
    const std::pair<void*, size_t> bufs[] = {
        {D.xk, 16 * np}, {D.vs, 16 * np}, {D.f4b, 16 * np}, {D.ksum, 4 * np}, {D.kvsum, 4 * np},
        {D.f1, 4 * np},  {D.kr, 4 * np},  {D.partial, 16 * tiles}, {D.active, 4 * (size_t)D.nscenes}};
    size_t total = 0;
    for (const auto& b : bufs) total += (b.second + 255) & ~size_t(255);
    unsigned char* save = nullptr;
    if (hipMalloc(&save, total) != hipSuccess) {   // no room to save the state: keep every window on
        (void)hipGetLastError();
        return 0;
    }
    size_t off = 0;
    hipError_t e = hipSuccess;
    for (const auto& b : bufs) {
        if (e == hipSuccess) e = hipMemcpyAsync(save + off, b.first, b.second, hipMemcpyDeviceToDevice, h->st);
        off += (b.second + 255) & ~size_t(255);
    }
    if (e != hipSuccess) {
        hipFree(save);
        return fail(h, SPHGPU_E_DEVICE, "saving the solver state for the window tuning failed: %s",
                    hipGetErrorString(e));
    }
    launch_fill_i32(D.active, 1, D.nscenes, h->st);
    // the gather build at its register-limited occupancy and capped at 6, 5,
    // 4 and 3 blocks per CU (unused dynamic LDS of 160 KB / k less a margin
    // for the kernel's own few hundred bytes)
    constexpr int kLdsPerCu = 160 * 1024;
    const int caps[5] = {0, kLdsPerCu / 6 - 1024, kLdsPerCu / 5 - 1024, kLdsPerCu / 4 - 1024,
                         kLdsPerCu / 3 - 1024};
    int mask = WIN_ALL;
    for (const auto& p : passes) {
        const int k = __builtin_ctz(p.first);
        D.winMask = WIN_ALL;
        const float tw = time_launches(h, p.second, nullptr, 1, 2);
        float tg = 1e30f;
        int best = 0;
        D.winMask = 0;
        for (int c : caps) {
            D.gatherLds[k] = c;
            const float t = time_launches(h, p.second, nullptr, 1, 2);
            if (t < tg) tg = t, best = c;
        }
        D.gatherLds[k] = best;
        h->winMs[2 * k] = tg;
        h->winMs[2 * k + 1] = tw;
        if (tg < tw) mask &= ~p.first;
    }
    D.winMask = mask;
    off = 0;
    for (const auto& b : bufs) {
        if (e == hipSuccess) e = hipMemcpyAsync(b.first, save + off, b.second, hipMemcpyDeviceToDevice, h->st);
        off += (b.second + 255) & ~size_t(255);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(h->st);
    hipFree(save);
    if (e != hipSuccess)
        return fail(h, SPHGPU_E_DEVICE, "restoring the solver state after the window tuning failed: %s",
                    hipGetErrorString(e));
    return 0;
}

int autotune(sphgpu_handle* h) {
    Dev& D = h->D;
    D.tileMap = 1;
    D.nbStage = 8;
    D.listNT = 0;
    run_ns(h);   // sorted layout and lists to time on (the first step sorts again: same order)
    auto density = [h]() {
        const Dev& D = h->D;
        if (h->P.solver == SPHGPU_DFSPH || h->P.solver == SPHGPU_IISPH_DF) launch_df_density_alpha(D, h->st);
        else if (h->P.solver == SPHGPU_IISPH) launch_ii_density(D, h->st);
        else launch_wc_pressure(D, h->st);
    };
    // each choice is timed unless sphgpu_set_tuning fixed it
    if (h->tuneFix[0] >= 0) {
        D.tileMap = h->tuneFix[0];
    } else {
        for (int m = 0; m < 2; ++m) {
            D.tileMap = m;
            h->tuneMs[m] = time_launches(h, density);
        }
        D.tileMap = h->tuneMs[0] < h->tuneMs[1] ? 0 : 1;
    }
    if (h->tuneFix[1] >= 0) {
        D.nbStage = h->tuneFix[1];
    } else {
        const int depth[2] = {6, 8};
        for (int k = 0; k < 2; ++k)
            h->tuneMs[2 + k] = time_launches(h, [h, &depth, k]() {
                launch_neighbours_variant(h->D, depth[k], h->st);
            });
        D.nbStage = h->tuneMs[2] <= h->tuneMs[3] ? 6 : 8;
    }
    launch_neighbours(D, h->st);   // lists of the chosen variant
    if (h->tuneFix[2] >= 0) {
        D.listNT = h->tuneFix[2];
    } else {
        // list loads: a fresh list build, then four passes (as in a step,
        // where whether the lists stay in the last-level cache between
        // passes decides)
        for (int k = 0; k < 2; ++k) {
            D.listNT = k;
            h->tuneMs[4 + k] = time_launches(h, density, [h]() { launch_neighbours(h->D, h->st); }, 4);
        }
        D.listNT = h->tuneMs[4] <= h->tuneMs[5] ? 0 : 1;
    }
    if (h->winFix < 0) {
        if (int rc = tune_windows(h)) return rc;
    }
    h->tuned = true;
    HIPCHECK(h, hipStreamSynchronize(h->st));
    HIPCHECK(h, hipGetLastError());
    return 0;
}

// ---- HIP graphs of whole step segments ------------------------------------
// A segment is host code that only queues kernels (no readback).  Its first
// run with a given device-state snapshot (the Dev struct: buffer pointers
// after the double-buffer swaps, sizes) and a given iteration count is
// captured and instantiated; later runs from an identical snapshot replay the
// graph and take the host state the capture ended with.  The ping-pong
// buffers repeat with period two, so a few graphs cover a run; a change of
// the particle layout (removals) makes new ones.
template <class F>
int run_segment(sphgpu_handle* h, int seg, int iters, F&& body) {
    for (auto& g : h->graphs)
        if (g.seg == seg && g.iters == iters && std::memcmp(&g.before, &h->D, sizeof(Dev)) == 0) {
            HIPCHECK(h, hipGraphLaunch(g.exec, h->st));
            h->D = g.after;
            g.used = ++h->graphClock;
            return 0;
        }
    GraphEntry e{};
    e.seg = seg;
    e.iters = iters;
    e.before = h->D;
    HIPCHECK(h, hipStreamBeginCapture(h->st, hipStreamCaptureModeThreadLocal));
    h->capturing = true;
    body();
    h->capturing = false;
    hipGraph_t graph = nullptr;
    HIPCHECK(h, hipStreamEndCapture(h->st, &graph));
    hipError_t ie = hipGraphInstantiate(&e.exec, graph, nullptr, nullptr, 0);
    hipGraphDestroy(graph);
    if (ie != hipSuccess) return fail(h, SPHGPU_E_DEVICE, "hipGraphInstantiate failed: %s", hipGetErrorString(ie));
    e.after = h->D;
    e.used = ++h->graphClock;
    if (h->graphs.size() >= 8) {   // least recently used goes
        auto old = std::min_element(h->graphs.begin(), h->graphs.end(),
                                    [](const GraphEntry& a, const GraphEntry& b) { return a.used < b.used; });
        hipGraphExecDestroy(old->exec);
        h->graphs.erase(old);
    }
    h->graphs.push_back(e);
    HIPCHECK(h, hipGraphLaunch(e.exec, h->st));
    return 0;
}

int finish_step(sphgpu_handle* h, sphgpu_stats* out) {
    Dev& D = h->D;
    int ns = D.nscenes;
    if (!out) {
        // no statistics asked for: nothing is read back and the host does not
        // wait; the fault counters keep accumulating until a step that reports
        // them, and the slot bound keeps its (larger, still valid) value
        ++h->frame;
        return 0;
    }
    launch_stats(D, h->statsPart, h->statsOut, h->st);
    HIPCHECK(h, hipMemcpyAsync(h->hStats, h->statsOut, sizeof(double) * 4 * ns, hipMemcpyDeviceToHost, h->st));
    HIPCHECK(h, hipMemcpyAsync(h->hCounters, D.counters, sizeof(int) * C_COUNT, hipMemcpyDeviceToHost, h->st));
    HIPCHECK(h, hipMemcpyAsync(h->hLayout, h->layoutOut, sizeof(int) * 2, hipMemcpyDeviceToHost, h->st));
    HIPCHECK(h, hipMemcpyAsync(h->hDt, D.dt, sizeof(float) * ns, hipMemcpyDeviceToHost, h->st));
    HIPCHECK(h, hipStreamSynchronize(h->st));
    HIPCHECK(h, hipGetLastError());
    // shrink the slot bound to the new padded layout
    if (h->hLayout[0] > 0 && h->hLayout[0] <= D.npad) D.npad = h->hLayout[0];
    sphgpu_stats agg{};
    double seSum = 0, ke = 0, emax = -1e300;
    long long nsum = 0;
    for (int s = 0; s < ns; ++s) {
        sphgpu_stats& ss = h->sceneStats[s];
        double n = h->hStats[4 * s + 3];
        ss.rho_mean_err = h->hStats[4 * s];
        ss.rho_max_err = h->hStats[4 * s + 1];
        ss.kinetic_energy = h->hStats[4 * s + 2];
        ss.dt = h->hDt[s];
        ss.num_fluid = (long long)n;
        seSum += ss.rho_mean_err * n;
        nsum += (long long)n;
        ke += ss.kinetic_energy;
        if (n > 0) emax = std::max(emax, ss.rho_max_err);
        agg.iter_density = std::max(agg.iter_density, ss.iter_density);
        agg.iter_divergence = std::max(agg.iter_divergence, ss.iter_divergence);
        agg.density_err = std::max(agg.density_err, ss.density_err);
        agg.divergence_err = std::max(agg.divergence_err, ss.divergence_err);
    }
    agg.rho_mean_err = nsum ? seSum / nsum : 0.0;
    agg.rho_max_err = nsum ? emax : 0.0;
    agg.kinetic_energy = ke;
    agg.dt = h->hDt[0];
    agg.zero_dis = h->hCounters[C_ZERO_DIS];
    agg.too_many_neigb = h->hCounters[C_TOO_MANY];
    agg.list_overflow = h->hCounters[C_LIST_OVF];
    agg.removed = h->hCounters[C_REMOVED];
    agg.nan_count = h->hCounters[C_NAN];
    agg.max_iter_hit = h->hCounters[C_COUNT - 1];
    h->nLive = (long long)nsum;
    agg.num_fluid = h->nLive;
    launch_fill_i32(D.counters, 0, C_COUNT, h->st);
    if (out) *out = agg;
    ++h->frame;
    if (agg.nan_count > 0) return fail(h, SPHGPU_E_DIVERGED, "%d non-finite fluid particles", agg.nan_count);
    return 0;
}

// Dynamic solids, at the start of a step: new poses to the device, world
// positions and velocities, boundary re-sort into cell order, cell ranges,
// Akinci volumes (SphSolid::updatePaticles, SphBase::updateSolidPartWeight).
int rigid_update(sphgpu_handle* h) {
    Dev& D = h->D;
    hipStream_t st = h->st;
    if (h->posesDirty) {
        HIPCHECK(h, hipMemcpyAsync(h->dSolids, h->hsolids.data(), sizeof(SolidPose) * D.nsolids,
                                   hipMemcpyHostToDevice, st));
        HIPCHECK(h, hipStreamSynchronize(st));   // the host table may change again right after
        h->posesDirty = false;
    }
    launch_rigid_pose(D, st);
    launch_boundary_keys_from(D, D.xb, D.bscene, h->sk, h->sv, st);
    radix_sort_pairs(h->sk, h->sv, D.nb, h->keyBits, h->sort, st);
    launch_boundary_regather(D, h->sk, h->sv, st);
    std::swap(D.xb, D.xb2); std::swap(D.vb, D.vb2); std::swap(D.bloc, D.bloc2);
    std::swap(D.bid, D.bid2); std::swap(D.bsolid, D.bsolid2); std::swap(D.bscene, D.bscene2);
    std::swap(D.bkself, D.bkself2); std::swap(D.bkey, D.bkey2);
    launch_cell_reset(D, D.bkey2, D.nb, D.cellB, st);
    launch_cell_build(D, D.bkey, D.nb, D.cellB, st);
    launch_boundary_volume_update(D, st);
    return 0;
}

// reaction forces on dynamic solids and their totals (SphWithBullet::updateSolids,
// the part before the rigid-body step); launched while x, v still hold x(t), v(t)
// and the cell ranges match them.  mode: 0 WCSPH, 1 IISPH, 2 DFSPH
void rigid_reaction(sphgpu_handle* h, int mode) {
    if (!h->D.anyDynamic) return;
    launch_rigid_reaction(h->D, mode, h->st);
    launch_solid_reduce(h->D, h->dSolidFT, h->st);
    h->ftValid = false;
}

// the point of SphWithBullet::updateSolids in the reference's step: the
// host's rigid-body step, if one is registered
int update_solids_point(sphgpu_handle* h) {
    if (!h->updateFn) return 0;
    h->inCallback = true;
    int rc = h->updateFn(h, h->updateUser);
    h->inCallback = false;
    if (rc) return fail(h, SPHGPU_E_CALLBACK, "update-solids callback returned %d", rc);
    return 0;
}

int step_wcsph(sphgpu_handle* h, sphgpu_stats* st) {
    auto body = [h]() {
        Dev& D = h->D;
        run_ns(h);
        launch_cfl(D, h->P.cfl, h->P.spacing_r, h->P.dt, h->P.dt_min, h->st);
        {
            ProfScope ps(h, "wc_pressure", nf_live(h) * 20.0);
            launch_wc_pressure(D, h->st);
        }
        {
            ProfScope ps(h, "wc_force", nf_live(h) * 56.0);
            launch_wc_force_integrate(D, h->st);
        }
        rigid_reaction(h, 0);
        std::swap(D.x, D.x2);
        std::swap(D.v, D.v2);
    };
    if (h->D.anyDynamic) {
        int rc = rigid_update(h);
        if (rc) return rc;
    }
    if (h->P.use_graph && !h->D.anyDynamic) {
        int rc = run_segment(h, 0, 0, body);
        if (rc) return rc;
    } else {
        body();
    }
    // WcSph.cpp:85; a failing callback still lets the step finish
    const int rcb = update_solids_point(h);
    const int rc = finish_step(h, st);
    return rcb ? rcb : rc;
}

// DFSPH::sphStep (DFSPH.cpp:9-75).  In graph mode (no dynamic solid, no
// update-solids callback, no online training in this step) the step is two
// captured segments, one per solver loop, each ending with enough
// iterations for the previous step's count (iterations after a scene
// converged return at once); the host reads the loop flags after each, and
// a loop that needs more iterations continues outside the capture.
int step_dfsph(sphgpu_handle* h, sphgpu_stats* st) {
    Dev& D = h->D;
    int rc;
    if (h->frame == 0 || h->needSearch) {
        // the density solve reads the lists of the previous search: after an
        // append they must cover the new particles
        run_ns(h);
        launch_df_density_alpha(D, h->st);
    }
    // divergence warm start: 0 none, 1 a scaled previous stiffness, 2 model
    const bool learnedMode = h->P.warm_start == SPHGPU_WARM_LEARNED;
    const bool training = learnedMode && !h->mlpPreset && h->learned < h->P.learn_steps;
    int mode = 0;
    if (h->frame > 0) {
        if (h->P.warm_start == SPHGPU_WARM_PREVIOUS || (learnedMode && training)) mode = 1;
        else if (learnedMode) mode = 2;
    }
    const bool graph = h->P.use_graph && !D.anyDynamic && !h->updateFn && !training;
    auto batch = [h, graph](int which) {
        return graph ? std::max({2, which ? h->P.min_iter_div : h->P.min_iter_density, (h->lastIters[which] + 2) & ~1})
                     : first_batch(h, which);
    };
    // density solve (DFSPH.cpp:195-738)
    size_t profStart = h->profUsed;
    const int b0 = batch(0);
    auto segA = [h, b0]() {
        Dev& D = h->D;
        launch_cfl(D, h->P.cfl, h->P.spacing_r, h->P.dt, h->P.dt_min, h->st);
        {
            ProfScope ps(h, "df_aadv", nf_live(h) * 40.0);
            launch_df_aadv(D, h->st);
        }
        solver_loop_start(h, 0, b0);
    };
    if (graph) {
        if ((rc = run_segment(h, 3, b0, segA))) return rc;
    } else {
        segA();
    }
    if ((rc = solver_loop_finish(h, 0, b0, profStart))) return rc;
    rigid_reaction(h, 2);
    // DFSPH.cpp:54: updateSolids between the density solve and the next
    // search; a failing callback still lets the step finish
    int rcb = 0;
    profStart = h->profUsed;
    const int b1 = batch(1);
    auto segB = [h, b1, mode, learnedMode, &rcb, &rc]() {
        Dev& D = h->D;
        launch_df_integrate(D, h->st);
        rcb = update_solids_point(h);
        if (D.anyDynamic && (rc = rigid_update(h))) return;
        run_ns(h);
        {
            ProfScope ps(h, "df_density_alpha", nf_live(h) * 20.0);
            if (mode != 0) launch_df_density_alpha_warm(D, learnedMode ? h->feat : nullptr, h->st);
            else launch_df_density_alpha(D, h->st);
        }
        launch_copy_f4(D.vs, D.v, D.npad, h->st);
        if (mode == 2) {
            ProfScope ps(h, "mlp_forward", nf_live(h) * 36.0);
            // (the warm start reads the prediction only where Drho/Dt > 0)
            mlp_forward(h->mlp, h->feat, h->ypred, D.npad, nullptr, h->st, true);
        }
        launch_df_warm_apply(D, h->ypred, h->feat, mode, h->st);
        if (mode != 0) {
            launch_fill_i32(D.active, 1, D.nscenes, h->st);
            launch_df_pass_velocity(D, h->st);
        }
        solver_loop_start(h, 1, b1);
    };
    rc = 0;
    if (graph) {
        int rg = run_segment(h, 4 + mode, b1, segB);
        if (rg) return rg;
    } else {
        segB();
    }
    if (rc) return rc;
    if ((rc = solver_loop_finish(h, 1, b1, profStart))) return rc;
    if (training && h->frame > 0) {
        // online training on this step's converged stiffness
        launch_df_targets(D, h->feat, h->tgt, h->st);
        const int b = h->P.learn_batch;
        for (int it = 0; it < h->P.learn_iters; ++it) {
            launch_mlp_gather(D, h->feat, h->tgt, h->bfeat, h->btgt, b, h->gatherSeed++, h->st);
            mlp_grad(h->mlp, h->bfeat, h->btgt, b, h->st);
            mlp_adam(h->mlp, h->P.learn_rate, h->st);
        }
        ++h->learned;
    }
    std::swap(D.v, D.vs);
    rc = finish_step(h, st);
    return rcb ? rcb : rc;
}

int step_iisph(sphgpu_handle* h, sphgpu_stats* st) {
    Dev& D = h->D;
    int rc;
    const size_t profStart = h->profUsed;
    // learned start of the pressure solve: the reference's 0.5 p(t-dt) while
    // the network trains (the first learn_steps steps), the network's
    // prediction after
    const bool learnedMode = h->P.warm_start == SPHGPU_WARM_LEARNED;
    const bool training = learnedMode && !h->mlpPreset && h->learned < h->P.learn_steps;
    const bool predict = learnedMode && !training && h->frame > 0;
    const bool graph = h->P.use_graph && !h->D.anyDynamic && !training;
    // graph mode: the whole step is one captured segment -- search, the
    // per-step passes, enough solver iterations for the previous step's
    // count plus one (iterations after a scene converged return at once),
    // and the pressure force + integration of every scene whose loop ended.
    // The host reads the flags once, after it; a scene that needed more
    // iterations continues outside the capture and is integrated then.
    const int batch = graph ? std::max({2, h->P.min_iter_density, (h->lastIters[0] + 2) & ~1}) : first_batch(h, 0);
    auto body = [h, batch, graph, learnedMode, predict]() {
        Dev& D = h->D;
        run_ns(h);
        launch_cfl(D, h->P.cfl, h->P.spacing_r, h->P.dt, h->P.dt_min, h->st);
        launch_ii_density(D, h->st);
        {
            ProfScope ps(h, "ii_aadv_dii", nf_live(h) * 64.0);
            launch_ii_aadv_dii(D, h->st);
        }
        {
            ProfScope ps(h, "ii_aii", nf_live(h) * 48.0);
            launch_ii_aii(D, learnedMode ? h->feat : nullptr, h->st);
        }
        if (predict) {
            ProfScope ps(h, "mlp_forward", nf_live(h) * 36.0);
            mlp_forward(h->mlp, h->feat, h->ypred, D.npad, nullptr, h->st);
            launch_ii_warm_apply(D, h->ypred, h->feat, h->st);
        }
        solver_loop_start(h, 0, batch);
        if (graph) launch_ii_force_integrate(h->D, 1, h->st);
    };
    if (h->D.anyDynamic && (rc = rigid_update(h))) return rc;
    if (graph) {
        if ((rc = run_segment(h, 1, batch, body))) return rc;
    } else {
        body();
    }
    int extra = 0;
    if ((rc = solver_loop_finish(h, 0, batch, profStart, &extra))) return rc;
    if (training && h->frame > 0) {
        // online training on this step's converged pressure
        launch_ii_targets(D, h->feat, h->tgt, h->st);
        const int b = h->P.learn_batch;
        for (int it = 0; it < h->P.learn_iters; ++it) {
            launch_mlp_gather(D, h->feat, h->tgt, h->bfeat, h->btgt, b, h->gatherSeed++, h->st);
            mlp_grad(h->mlp, h->bfeat, h->btgt, b, h->st);
            mlp_adam(h->mlp, h->P.learn_rate, h->st);
        }
        ++h->learned;
    }
    if (!graph || extra > 0) {
        ProfScope ps(h, "ii_force", nf_live(h) * 56.0);
        launch_ii_force_integrate(D, graph ? 2 : 0, h->st);
    }
    rigid_reaction(h, 1);
    std::swap(D.x, D.x2);
    std::swap(D.v, D.v2);
    const int rcb = update_solids_point(h);            // IISph_M.cpp:64
    rc = finish_step(h, st);
    return rcb ? rcb : rc;
}

// IISph::sphStep with the DF_* solvers in place of the VF_* ones (the
// commented-out calls at IISph.cpp:59-60): search, V_b, DF_prepareAttribute
// (IISph.cpp:100-288), the non-pressure acceleration (DFSPH's
// DF_forceExceptPressure, DFSPH.cpp:1419-1475: the IISph class computes it in
// its VF_* advection step), DF_divergenceFreeSolver (1039-1217),
// DF_constantDensitySolver (816-1037), updateFluids, updateSolids.  These
// solvers compute no reaction force on solids.
int step_iisph_df(sphgpu_handle* h, sphgpu_stats* st) {
    Dev& D = h->D;
    int rc;
    if (D.anyDynamic && (rc = rigid_update(h))) return rc;
    run_ns(h);
    launch_cfl(D, h->P.cfl, h->P.spacing_r, h->P.dt, h->P.dt_min, h->st);
    {
        ProfScope ps(h, "idf_prepare", nf_live(h) * 24.0);
        launch_idf_prepare(D, h->st);
    }
    {
        ProfScope ps(h, "df_aadv", nf_live(h) * 40.0);
        launch_df_aadv(D, h->st);
    }
    launch_idf_vadv(D, 0, h->st);
    if ((rc = solver_loop(h, 1))) return rc;
    launch_idf_vadv(D, 1, h->st);
    if ((rc = solver_loop(h, 0))) return rc;
    launch_df_integrate(D, h->st);
    const int rcb = update_solids_point(h);
    rc = finish_step(h, st);
    return rcb ? rcb : rc;
}

// PciSph::sphStep (PciSph.cpp:31-95): search, weights, pci_computeForce
// (97-126), the boundary reaction, updateFluids, updateSolids
int step_pcisph(sphgpu_handle* h, sphgpu_stats* st) {
    Dev& D = h->D;
    int rc;
    const size_t profStart = h->profUsed;
    const int batch = first_batch(h, 0);
    auto body = [h, batch]() {
        Dev& D = h->D;
        run_ns(h);
        launch_cfl(D, h->P.cfl, h->P.spacing_r, h->P.dt, h->P.dt_min, h->st);
        {
            ProfScope ps(h, "pci_aext", nf_live(h) * 56.0);
            launch_pci_aext(D, h->st);
        }
        solver_loop_start(h, 0, batch);
    };
    if (h->D.anyDynamic && (rc = rigid_update(h))) return rc;
    if (h->frame == 0) launch_pci_rest_density(D, h->st);
    if (h->P.use_graph && !h->D.anyDynamic) {
        if ((rc = run_segment(h, 2, batch, body))) return rc;
    } else {
        body();
    }
    if ((rc = solver_loop_finish(h, 0, batch, profStart))) return rc;
    rigid_reaction(h, 0);       // pci_boundPartForce_f: the WCSPH form (PciSph.cpp:440-460)
    {
        ProfScope ps(h, "pci_integrate", nf_live(h) * 84.0);
        launch_pci_integrate(D, h->st);
    }
    const int rcb = update_solids_point(h);
    rc = finish_step(h, st);
    return rcb ? rcb : rc;
}

// host: uniform grid of a scene (SphBase::SpaceGridStruct::initialize, SphBase.cpp:963-988)
void build_grids(sphgpu_handle* h, std::vector<int>& rowRank) {
    const float H = h->P.smooth_radius_h;
    uint32_t keyBase = 0;
    int rowBase = 0;
    h->hgrids.clear();
    h->hKeyBase.clear();
    rowRank.clear();
    for (int s = 0; s < h->P.num_scenes; ++s) {
        const SceneDesc& d = h->scenes[s];
        SceneGrid g{};
        volatile float twoH = H * 2.0f;
        g.minx = d.mn[0] - twoH;
        g.miny = d.mn[1] - twoH;
        g.minz = d.mn[2] - twoH;
        g.divisor = 1.0f / H;
        int num[3];
        for (int a = 0; a < 3; ++a) {
            volatile float ext = d.mx[a] - d.mn[a];
            volatile float e2 = ext + twoH;
            volatile float e3 = e2 * g.divisor;
            num[a] = int(e3) + 2;
            g.smin[a] = d.mn[a];
            g.smax[a] = d.mx[a];
        }
        g.nx = num[0]; g.ny = num[1]; g.nz = num[2];
        g.keyBase = keyBase;
        g.rowBase = rowBase;
        // rows (cy, cz) ranked by 2-D Morton code
        std::vector<std::pair<uint32_t, int>> rows;
        rows.reserve((size_t)g.ny * g.nz);
        for (int cz = 0; cz < g.nz; ++cz)
            for (int cy = 0; cy < g.ny; ++cy)
                rows.push_back({spread2(uint32_t(cy)) | (spread2(uint32_t(cz)) << 1), cz * g.ny + cy});
        std::sort(rows.begin(), rows.end());
        rowRank.resize((size_t)rowBase + (size_t)g.ny * g.nz);
        for (size_t r = 0; r < rows.size(); ++r) rowRank[(size_t)rowBase + rows[r].second] = int(r);
        h->hgrids.push_back(g);
        h->hKeyBase.push_back(keyBase);
        keyBase += uint32_t((uint64_t)g.nx * g.ny * g.nz);
        rowBase += g.ny * g.nz;
    }
    h->hKeyBase.push_back(keyBase);
}

}  // namespace

// ============================================================================
extern "C" {

int sphgpu_abi_version(void) { return SPHGPU_ABI_VERSION; }

void sphgpu_default_params(sphgpu_params* p) {
    if (!p) return;
    std::memset(p, 0, sizeof(*p));
    p->solver = SPHGPU_DFSPH;
    p->spacing_r = 0.1f;                  // SphBase.cpp:214
    p->smooth_radius_h = 0.2f;            // SphBase.cpp:215
    p->space_max[0] = p->space_max[1] = p->space_max[2] = 1.0f;
    p->dt = 0.001f;                       // SphBase.cpp:219
    p->gravity[1] = -9.8f;                // SphBase.cpp:220
    p->sound_speed = 100.0f;              // SphBase.cpp:222
    p->gamma = 7.0f;                      // WcSph.h:34
    p->nu = 0.001f;
    p->cfl = 0.0f;
    p->dt_min = 1e-6f;
    p->density_tol = 1e-4f;               // DFSPH.cpp:225
    p->divergence_tol = 2.5f;             // DFSPH.cpp:762
    p->min_iter_density = 2;              // DFSPH.cpp:229
    p->max_iter_density = 100;
    p->min_iter_div = 1;                  // DFSPH.cpp:765
    p->max_iter_div = 30;                 // DFSPH.cpp:764
    p->omega = 0.5f;                      // IISph_M.cpp:341
    p->warm_start = SPHGPU_WARM_NONE;
    p->num_scenes = 1;
    p->learn_steps = 40;
    p->learn_iters = 32;
    p->learn_batch = 16384;
    p->learn_rate = 1e-3f;
    p->learn_over_weight = 1.0f;
    p->warm_scale = 0.7f;
    p->learn_scale = 0.7f;
    p->multiphase_consistent = 0;
    p->boundary_density = 1000.0f;        // SphBase.h:576
    p->adjust_density = 0;                // SphBase.h:575
    p->fluid_capacity = 0;
}

int sphgpu_device_available(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return 0;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, 0) != hipSuccess) return 0;
    return std::strncmp(prop.gcnArchName, "gfx9", 4) == 0 ? 1 : 0;
}

int sphgpu_create(const sphgpu_params* p, sphgpu_handle** out) {
    if (!p || !out) return SPHGPU_E_ARG;
    *out = nullptr;
    static_assert(SPHGPU_IISPH_DF == SOLVER_IISPH_DF, "solver id of the IISph DF_* kernels");
    if (p->solver < SPHGPU_WCSPH || p->solver > SPHGPU_IISPH_DF) return SPHGPU_E_ARG;
    // the IISph DF_* solvers have no warm start and are not graph-captured
    if (p->solver == SPHGPU_IISPH_DF && (p->warm_start != SPHGPU_WARM_NONE || p->use_graph)) return SPHGPU_E_ARG;
    if (!(p->boundary_density > 0.0f) || !std::isfinite(p->boundary_density) || p->adjust_density < 0 ||
        p->adjust_density > 1 || p->fluid_capacity < 0)
        return SPHGPU_E_ARG;
    if (!(p->spacing_r > 0) || !(p->smooth_radius_h > 0) || !(p->dt > 0)) return SPHGPU_E_ARG;
    if (p->num_scenes < 1 || p->num_scenes > 4096) return SPHGPU_E_ARG;
    if (p->warm_start < 0 || p->warm_start > 2) return SPHGPU_E_ARG;
    if (p->warm_start == SPHGPU_WARM_LEARNED &&
        ((p->solver != SPHGPU_DFSPH && p->solver != SPHGPU_IISPH) || p->learn_steps < 0 || p->learn_iters < 0 || p->learn_batch < 16 ||
         p->learn_batch > (1 << 22) || !(p->learn_rate >= 0) || !(p->learn_over_weight > 0) ||
         !std::isfinite(p->learn_over_weight)))
        return SPHGPU_E_ARG;
    if (!(p->warm_scale >= 0.0f) || !(p->warm_scale <= 4.0f)) return SPHGPU_E_ARG;
    if (!(p->learn_scale >= 0.0f) || !(p->learn_scale <= 4.0f)) return SPHGPU_E_ARG;
    if (p->multiphase_consistent < 0 || p->multiphase_consistent > 1) return SPHGPU_E_ARG;
    if (p->use_graph < 0 || p->use_graph > 1) return SPHGPU_E_ARG;
    // every solver loop ends at a finite cap (and reports reaching it)
    if (p->max_iter_density < 1 || p->max_iter_density > SPHGPU_MAX_ITER || p->max_iter_div < 0 ||
        p->max_iter_div > SPHGPU_MAX_ITER || p->min_iter_density < 0 || p->min_iter_div < 0)
        return SPHGPU_E_ARG;
    sphgpu_handle* h = new sphgpu_handle();
    h->P = *p;
    h->scenes.resize(p->num_scenes);
    std::memcpy(h->scenes[0].mn, p->space_min, sizeof(float) * 3);
    std::memcpy(h->scenes[0].mx, p->space_max, sizeof(float) * 3);
    h->scenes[0].set = true;
    h->sceneStats.resize(p->num_scenes);
    h->D.winMask = WIN_ALL;
    *out = h;
    return SPHGPU_OK;
}

void sphgpu_destroy(sphgpu_handle* h) {
    if (!h) return;
    if (h->initialized) {
        hipStreamSynchronize(h->st);
        for (auto& e : h->prof) { hipEventDestroy(e.a); hipEventDestroy(e.b); }
        for (auto& g : h->graphs) hipGraphExecDestroy(g.exec);
        for (void* m : h->allocs) hipFree(m);
        if (h->mlpLive) mlp_free(h->mlp);
        hipHostFree(h->hActive);
        if (h->snap.cs) {
            hipStreamSynchronize(h->snap.cs);
            hipEventDestroy(h->snap.copied);
            hipEventDestroy(h->snap.done);
            hipStreamDestroy(h->snap.cs);
            hipHostFree(h->snap.host);
        }
        hipStreamDestroy(h->st);
    }
    delete h;
}

const char* sphgpu_last_error(const sphgpu_handle* h) { return h ? h->err.c_str() : "null handle"; }

int sphgpu_set_scene_domain(sphgpu_handle* h, int scene, const float mn[3], const float mx[3]) {
    if (!h || !mn || !mx) return SPHGPU_E_ARG;
    if (h->initialized) return fail(h, SPHGPU_E_STATE, "scene domain after init");
    if (scene < 0 || scene >= h->P.num_scenes) return fail(h, SPHGPU_E_ARG, "scene %d out of range", scene);
    for (int a = 0; a < 3; ++a)
        if (!(mx[a] > mn[a])) return fail(h, SPHGPU_E_ARG, "empty domain on axis %d", a);
    std::memcpy(h->scenes[scene].mn, mn, sizeof(float) * 3);
    std::memcpy(h->scenes[scene].mx, mx, sizeof(float) * 3);
    h->scenes[scene].set = true;
    if (scene == 0) {
        std::memcpy(h->P.space_min, mn, sizeof(float) * 3);
        std::memcpy(h->P.space_max, mx, sizeof(float) * 3);
    }
    return SPHGPU_OK;
}

int sphgpu_add_phase(sphgpu_handle* h, float rest_density, float visc_alpha, int* phase) {
    if (!h) return SPHGPU_E_ARG;
    if (h->initialized) return fail(h, SPHGPU_E_STATE, "phase after init");
    if (!(rest_density > 0)) return fail(h, SPHGPU_E_ARG, "rest density must be > 0");
    if ((int)h->phases.size() >= MAX_PHASES) return fail(h, SPHGPU_E_CAPACITY, "at most %d phases", MAX_PHASES);
    h->phases.push_back({rest_density, visc_alpha});
    if (phase) *phase = int(h->phases.size()) - 1;
    return SPHGPU_OK;
}

int sphgpu_add_fluid(sphgpu_handle* h, int scene, int phase, long long n, const float* xyz,
                     const float* vxyz, long long* first_id) {
    if (!h) return SPHGPU_E_ARG;
    if (h->initialized) return sphgpu_append_fluid(h, scene, phase, n, xyz, vxyz, first_id);
    if (scene < 0 || scene >= h->P.num_scenes) return fail(h, SPHGPU_E_ARG, "scene %d out of range", scene);
    if (phase < 0 || phase >= (int)h->phases.size()) return fail(h, SPHGPU_E_ARG, "phase %d not defined", phase);
    if (n < 0 || (n > 0 && !xyz)) return fail(h, SPHGPU_E_ARG, "bad particle array");
    if (h->nextId + n > (1LL << 30)) return fail(h, SPHGPU_E_CAPACITY, "too many fluid particles");
    if (first_id) *first_id = h->nextId;
    size_t o = h->fx.size();
    h->fx.insert(h->fx.end(), xyz, xyz + 3 * n);
    if (vxyz) h->fv.insert(h->fv.end(), vxyz, vxyz + 3 * n);
    else h->fv.resize(o + 3 * n, 0.0f);
    h->fphase.insert(h->fphase.end(), (size_t)n, phase);
    h->fscene.insert(h->fscene.end(), (size_t)n, scene);
    h->nextId += n;
    return SPHGPU_OK;
}

int sphgpu_append_fluid(sphgpu_handle* h, int scene, int phase, long long n, const float* xyz,
                        const float* vxyz, long long* first_id) {
    if (!h) return SPHGPU_E_ARG;
    if (!h->initialized) return sphgpu_add_fluid(h, scene, phase, n, xyz, vxyz, first_id);
    if (h->inCallback) return fail(h, SPHGPU_E_STATE, "append inside the update-solids callback");
    if (scene < 0 || scene >= h->P.num_scenes) return fail(h, SPHGPU_E_ARG, "scene %d out of range", scene);
    if (phase < 0 || phase >= (int)h->phases.size()) return fail(h, SPHGPU_E_ARG, "phase %d not defined", phase);
    if (n < 0 || (n > 0 && !xyz)) return fail(h, SPHGPU_E_ARG, "bad particle array");
    const SceneDesc& d = h->scenes[scene];
    for (long long i = 0; i < n; ++i)
        for (int a = 0; a < 3; ++a)
            if (!(xyz[3 * i + a] >= d.mn[a] && xyz[3 * i + a] <= d.mx[a]))
                return fail(h, SPHGPU_E_ARG, "appended particle %lld outside its scene domain", i);
    if (first_id) *first_id = h->nextId;
    if (n == 0) return SPHGPU_OK;
    Dev& D = h->D;
    // the new particles take fresh tiles after the current slots; the next
    // neighbour search sorts them into the scenes' layout
    const int slots = int((n + TILE - 1) / TILE * TILE);
    const long long cap = std::max<long long>(h->P.fluid_capacity, (long long)h->fscene.size() - h->appended);
    if (h->nLive + n > cap)
        return fail(h, SPHGPU_E_CAPACITY, "appending %lld particles to %lld exceeds fluid_capacity %lld", n, h->nLive,
                    cap);
    if ((long long)D.npad + slots > D.ncap)
        return fail(h, SPHGPU_E_CAPACITY, "appending %lld particles exceeds the fluid capacity (%d slots in use of %d;"
                    " set sphgpu_params.fluid_capacity)", n, D.npad, D.ncap);
    if (h->nextId + n > (1LL << 30)) return fail(h, SPHGPU_E_CAPACITY, "too many fluid particles");
    HIPCHECK(h, hipStreamSynchronize(h->st));   // a step without statistics may still run
    std::vector<float4> hx(slots, make_float4(0, 0, 0, 0)), hv(slots, make_float4(0, 0, 0, 0));
    std::vector<int> hid(slots, -1);
    std::vector<float> hk(slots, h->P.solver == SPHGPU_IISPH_DF ? 1.0f : 0.0f), hz(slots, 0.0f),
        hr(slots, h->phases[phase].rho0);
    std::vector<uint32_t> hkey(slots, KEY_INVALID);
    for (long long i = 0; i < n; ++i) {
        hx[i] = make_float4(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2], float(phase));
        if (vxyz) hv[i] = make_float4(vxyz[3 * i], vxyz[3 * i + 1], vxyz[3 * i + 2], 0.0f);
        hid[i] = int(h->nextId + i);
    }
    const size_t o = (size_t)D.npad;
    HIPCHECK(h, hipMemcpy(D.x + o, hx.data(), sizeof(float4) * slots, hipMemcpyHostToDevice));
    HIPCHECK(h, hipMemcpy(D.v + o, hv.data(), sizeof(float4) * slots, hipMemcpyHostToDevice));
    HIPCHECK(h, hipMemcpy(D.id + o, hid.data(), sizeof(int) * slots, hipMemcpyHostToDevice));
    HIPCHECK(h, hipMemcpy(D.ksum + o, hk.data(), sizeof(float) * slots, hipMemcpyHostToDevice));
    HIPCHECK(h, hipMemcpy(D.kvsum + o, hz.data(), sizeof(float) * slots, hipMemcpyHostToDevice));
    HIPCHECK(h, hipMemcpy(D.rho + o, hr.data(), sizeof(float) * slots, hipMemcpyHostToDevice));
    HIPCHECK(h, hipMemcpy(D.key + o, hkey.data(), sizeof(uint32_t) * slots, hipMemcpyHostToDevice));
    HIPCHECK(h, hipMemcpy(D.key2 + o, hkey.data(), sizeof(uint32_t) * slots, hipMemcpyHostToDevice));
    // the key pass picks each slot's grid by its tile's scene
    std::vector<int> ts(slots / TILE, scene), tv(slots / TILE, 0);
    for (int t = 0; t < slots / TILE; ++t) tv[t] = int(std::min<long long>(TILE, n - (long long)t * TILE));
    HIPCHECK(h, hipMemcpy(D.tileScene + o / TILE, ts.data(), sizeof(int) * ts.size(), hipMemcpyHostToDevice));
    HIPCHECK(h, hipMemcpy(D.tileValid + o / TILE, tv.data(), sizeof(int) * tv.size(), hipMemcpyHostToDevice));
    D.npad += slots;
    h->nLive += n;
    h->needSearch = true;
    h->fx.insert(h->fx.end(), xyz, xyz + 3 * n);
    if (vxyz) h->fv.insert(h->fv.end(), vxyz, vxyz + 3 * n);
    else h->fv.resize(h->fx.size(), 0.0f);
    h->fphase.insert(h->fphase.end(), (size_t)n, phase);
    h->fscene.insert(h->fscene.end(), (size_t)n, scene);
    if (!h->fbeta.empty()) h->fbeta.resize(h->fphase.size(), 1.0f);
    h->nextId += n;
    h->appended += n;
    return SPHGPU_OK;
}

int sphgpu_set_fluid_beta(sphgpu_handle* h, long long first, long long n, const float* beta) {
    if (!h || !beta || first < 0 || n < 0) return SPHGPU_E_ARG;
    if (h->initialized) return fail(h, SPHGPU_E_STATE, "beta is set before sphgpu_init");
    const long long nf = (long long)h->fscene.size();
    if (first + n > nf) return fail(h, SPHGPU_E_ARG, "particles [%lld, %lld) beyond the %lld added", first, first + n, nf);
    for (long long i = 0; i < n; ++i)
        if (!(beta[i] > 0.0f) || !std::isfinite(beta[i])) return fail(h, SPHGPU_E_ARG, "beta must be positive");
    if (h->fbeta.size() < (size_t)nf) h->fbeta.resize(nf, 1.0f);
    std::copy(beta, beta + n, h->fbeta.begin() + first);
    return SPHGPU_OK;
}

int sphgpu_add_solid(sphgpu_handle* h, int scene, long long n, const float* xyz, float visc_alpha,
                     int* solid) {
    if (!h) return SPHGPU_E_ARG;
    if (h->initialized) return fail(h, SPHGPU_E_STATE, "solid after init");
    if (scene < 0 || scene >= h->P.num_scenes) return fail(h, SPHGPU_E_ARG, "scene %d out of range", scene);
    if (n < 0 || (n > 0 && !xyz)) return fail(h, SPHGPU_E_ARG, "bad particle array");
    int sid = int(h->solidVisc.size());
    SolidPose sp{};
    for (int k = 0; k < 4; ++k) sp.T[5 * k] = 1.0f;   // identity
    sp.visc = visc_alpha;
    sp.dynamic = 0;
    sp.firstId = int(h->bsolid.size());
    sp.endId = int(h->bsolid.size() + n);
    h->hsolids.push_back(sp);
    h->solidVisc.push_back(visc_alpha);
    h->bx.insert(h->bx.end(), xyz, xyz + 3 * n);
    h->bsolid.insert(h->bsolid.end(), (size_t)n, sid);
    h->bscene.insert(h->bscene.end(), (size_t)n, scene);
    if (solid) *solid = sid;
    return SPHGPU_OK;
}

int sphgpu_add_rigid(sphgpu_handle* h, int scene, long long n, const float* local_xyz, float visc_alpha,
                     int* solid) {
    if (!h) return SPHGPU_E_ARG;
    int sid = -1;
    int rc = sphgpu_add_solid(h, scene, n, local_xyz, visc_alpha, &sid);
    if (rc) return rc;
    h->hsolids[sid].dynamic = 1;
    if (solid) *solid = sid;
    return SPHGPU_OK;
}

int sphgpu_set_solid_pose(sphgpu_handle* h, int solid, const float T[16], const float com[3], const float v[3],
                          const float w[3]) {
    if (!h || !T || !com || !v || !w) return SPHGPU_E_ARG;
    if (solid < 0 || solid >= (int)h->hsolids.size() || !h->hsolids[solid].dynamic)
        return fail(h, SPHGPU_E_ARG, "solid %d is not a dynamic solid", solid);
    for (int k = 0; k < 16; ++k)
        if (!std::isfinite(T[k])) return fail(h, SPHGPU_E_ARG, "non-finite transform");
    SolidPose& sp = h->hsolids[solid];
    std::memcpy(sp.T, T, sizeof(float) * 16);
    std::memcpy(sp.com, com, sizeof(float) * 3);
    std::memcpy(sp.v, v, sizeof(float) * 3);
    std::memcpy(sp.w, w, sizeof(float) * 3);
    h->posesDirty = true;
    return SPHGPU_OK;
}

int sphgpu_set_update_solids(sphgpu_handle* h, sphgpu_update_solids_fn fn, void* user) {
    if (!h) return SPHGPU_E_ARG;
    if (h->inCallback) return fail(h, SPHGPU_E_STATE, "set_update_solids inside the callback");
    h->updateFn = fn;
    h->updateUser = user;
    return SPHGPU_OK;
}

int sphgpu_get_solid_force_torque(sphgpu_handle* h, int solid, double f[3], double t[3]) {
    if (!h || !f || !t) return SPHGPU_E_ARG;
    if (solid < 0 || solid >= (int)h->hsolids.size()) return fail(h, SPHGPU_E_ARG, "solid %d out of range", solid);
    if (!h->initialized || !h->dSolidFT) return fail(h, SPHGPU_E_STATE, "no step has run");
    if (!h->ftValid) {   // all bodies in one copy: a host step reads every body
        h->hSolidFT.resize(6 * (size_t)h->D.nsolids);
        HIPCHECK(h, hipMemcpyAsync(h->hSolidFT.data(), h->dSolidFT, sizeof(double) * h->hSolidFT.size(),
                                   hipMemcpyDeviceToHost, h->st));
        HIPCHECK(h, hipStreamSynchronize(h->st));
        h->ftValid = true;
    }
    for (int q = 0; q < 3; ++q) {
        f[q] = h->hSolidFT[6 * solid + q];
        t[q] = h->hSolidFT[6 * solid + 3 + q];
    }
    return SPHGPU_OK;
}

// world position of boundary sample i at init: as given for static solids; for
// dynamic ones the body-frame sample placed by the pose set before init
static void boundary_world(const sphgpu_handle* h, size_t i, float w[3]) {
    const SolidPose& sp = h->hsolids[h->bsolid[i]];
    const float x = h->bx[3 * i], y = h->bx[3 * i + 1], z = h->bx[3 * i + 2];
    if (!sp.dynamic) {
        w[0] = x; w[1] = y; w[2] = z;
        return;
    }
    const float* T = sp.T;
    const float a = T[3] * x + T[7] * y + T[11] * z + T[15];
    w[0] = (T[0] * x + T[4] * y + T[8] * z + T[12]) / a;
    w[1] = (T[1] * x + T[5] * y + T[9] * z + T[13]) / a;
    w[2] = (T[2] * x + T[6] * y + T[10] * z + T[14]) / a;
}

int sphgpu_init(sphgpu_handle* h) {
    if (!h) return SPHGPU_E_ARG;
    if (h->initialized) return SPHGPU_OK;
    const sphgpu_params& P = h->P;
    if (h->phases.empty()) return fail(h, SPHGPU_E_STATE, "no fluid phase defined");
    for (int s = 0; s < P.num_scenes; ++s)
        if (!h->scenes[s].set) return fail(h, SPHGPU_E_STATE, "scene %d has no domain", s);
    if (!sphgpu_device_available()) return fail(h, SPHGPU_E_DEVICE, "no gfx9 HIP device available");
    // every particle must lie inside its scene domain (SphBase grows the domain to contain them)
    for (size_t i = 0; i < h->fscene.size(); ++i) {
        const SceneDesc& d = h->scenes[h->fscene[i]];
        for (int a = 0; a < 3; ++a)
            if (!(h->fx[3 * i + a] >= d.mn[a] && h->fx[3 * i + a] <= d.mx[a]))
                return fail(h, SPHGPU_E_ARG, "fluid particle %zu outside its scene domain", i);
    }
    for (size_t i = 0; i < h->bscene.size(); ++i) {
        const SceneDesc& d = h->scenes[h->bscene[i]];
        float w[3];
        boundary_world(h, i, w);
        for (int a = 0; a < 3; ++a)
            if (!(w[a] >= d.mn[a] && w[a] <= d.mx[a]))
                return fail(h, SPHGPU_E_ARG, "boundary particle %zu outside its scene domain", i);
    }
    for (float sv : h->solidVisc)
        if (sv != h->solidVisc[0]) return fail(h, SPHGPU_E_ARG, "all solids must share viscosity_alpha");

    HIPCHECK(h, hipSetDevice(0));
    HIPCHECK(h, hipStreamCreateWithFlags(&h->st, hipStreamNonBlocking));
    h->initialized = true;

    std::vector<int> rowRank;
    build_grids(h, rowRank);
    uint64_t totalKeys = h->hKeyBase.back();
    if (totalKeys >= 0xFFFFFFF0ull) return fail(h, SPHGPU_E_CAPACITY, "grid too large (%llu cells)", (unsigned long long)totalKeys);
    int bits = 1;
    while ((1ull << bits) <= totalKeys) ++bits;   // KEY_INVALID & mask must exceed every key
    h->keyBits = bits;

    const int ns = P.num_scenes;
    // initial padded layout: scenes contiguous in add order
    std::vector<int> count(ns, 0), padStart(ns + 1, 0);
    for (int s : h->fscene) ++count[s];
    for (int s = 0; s < ns; ++s) padStart[s + 1] = padStart[s] + (count[s] + TILE - 1) / TILE * TILE;
    int npad = padStart[ns];
    // room for particles appended later (fluid_capacity): every scene's
    // layout rounds up to whole tiles, so one spare tile per scene
    const long long capSlots = P.fluid_capacity > 0
                                   ? ((long long)P.fluid_capacity + TILE - 1) / TILE * TILE + (long long)ns * TILE
                                   : 0;
    if (capSlots > (1LL << 30)) return fail(h, SPHGPU_E_CAPACITY, "fluid_capacity too large");
    int ncap = std::max({npad, TILE, int(capSlots)});
    long long nb = (long long)h->bscene.size();

    Dev& D = h->D;
    D.ncap = ncap;
    D.npad = npad;
    D.nb = int(nb);
    D.nscenes = ns;
    D.nphases = int(h->phases.size());
    D.solver = P.solver;
    const float s = P.spacing_r, H = P.smooth_radius_h;
    D.kh = H / 2.0f;
    D.inv_kh = 1.0f / D.kh;
    D.sigma = 1.0f / float(M_PI) / (D.kh * D.kh * D.kh);
    D.sigma_h = 1.0f / float(M_PI) / (D.kh * D.kh * D.kh) / D.kh;
    D.W0 = D.sigma;
    {
        const double sh = double(D.sigma_h), ik = double(D.inv_kh), c = std::sqrt(0.75 * sh);
        D.gA = float(-3.0 * sh * ik);
        D.gB = float(2.25 * sh * ik * ik);
        D.gU = float(-c * ik);
        D.gV = float(2.0 * c);
    }
    {
        volatile float hh = s + s;
        volatile float h2 = hh * hh;
        D.h2 = h2;
    }
    D.Hs = H;
    D.visc_eps = 0.01f * H * H;
    D.v0 = s * s * s;
    {
        // PCISPH stiffness (PciSph::pci_updateDelta, PciSph.cpp:133-162): a
        // prototype particle with a full neighbourhood on the lattice of
        // spacing s; delta = -1 / (beta (-|sum gradW|^2 - sum |gradW|^2)),
        // beta = 2 v0^2; the passes divide by dt^2 of their scene
        const double kh = 0.5 * double(H), sig = 1.0 / M_PI / (kh * kh * kh), sd = s;
        const int n = int(std::ceil(double(H) / sd));
        double gs[3] = {0, 0, 0}, g2 = 0.0;
        for (int z = -n; z <= n; ++z)
            for (int y = -n; y <= n; ++y)
                for (int x = -n; x <= n; ++x) {
                    if (!x && !y && !z) continue;
                    const double vx = x * sd, vy = y * sd, vz = z * sd, dis = std::sqrt(vx * vx + vy * vy + vz * vz);
                    if (!(dis < double(H))) continue;
                    const double q = dis / kh;
                    const double dw = sig / kh * (q <= 1.0 ? -3.0 * q + 2.25 * q * q : -0.75 * (2 - q) * (2 - q));
                    gs[0] += vx * dw / dis; gs[1] += vy * dw / dis; gs[2] += vz * dw / dis;
                    g2 += dw * dw;
                }
        const double v0 = sd * sd * sd, beta = 2.0 * v0 * v0;
        D.pci_delta = float(-1.0 / (beta * (-(gs[0] * gs[0] + gs[1] * gs[1] + gs[2] * gs[2]) - g2)));
    }
    D.boundaryDensity = P.boundary_density;
    D.adjustDensity = P.adjust_density;
    D.spacing = s;
    D.dt_max = P.dt;
    D.warm_scale = P.warm_scale;
    D.learn_scale = P.learn_scale;
    D.mpCons = P.multiphase_consistent;
    {
        // alpha of a particle inside a lattice of the sampling spacing
        // (DFSPH.cpp:541-566 on a full neighbourhood): rho / sum (m |grad W|)^2
        const double kh = D.kh, sig = 1.0 / M_PI / (kh * kh * kh), m = double(h->phases[0].rho0) * s * s * s;
        double rho = 0.0, g2 = 0.0;
        const int R = int(std::ceil(2.0 * kh / s)) + 1;
        for (int a = -R; a <= R; ++a)
            for (int b = -R; b <= R; ++b)
                for (int c = -R; c <= R; ++c) {
                    double r = s * std::sqrt(double(a * a + b * b + c * c)), q = r / kh;
                    if (q >= 2.0) continue;
                    rho += m * sig * (q <= 1.0 ? 1.0 - 1.5 * q * q + 0.75 * q * q * q : 0.25 * (2 - q) * (2 - q) * (2 - q));
                    if (r == 0.0) continue;
                    double dw = sig / kh * (q <= 1.0 ? -3.0 * q + 2.25 * q * q : -0.75 * (2 - q) * (2 - q));
                    g2 += (m * dw) * (m * dw);
                }
        D.alpha_lat = float(rho / std::max(g2, 1e-6));
        // IISPH a_ii of the same particle at dt = 1: -(m/rho0)^2 sum |grad W|^2
        // (its d_ii vanishes on the symmetric lattice)
        D.aii_lat = float(-g2 / (double(h->phases[0].rho0) * double(h->phases[0].rho0)));
    }
    D.gx = P.gravity[0]; D.gy = P.gravity[1]; D.gz = P.gravity[2];
    D.cs = P.sound_speed;
    D.gamma = P.gamma;
    D.nu = P.nu;
    D.omega = P.omega;
    for (int k = 0; k < MAX_PHASES; ++k) {
        const Phase ph = k < (int)h->phases.size() ? h->phases[k] : h->phases[0];
        D.rho0[k] = ph.rho0;
        D.mass[k] = ph.rho0 * D.v0;
        D.visc[k] = ph.visc;
        D.tait_B[k] = ph.rho0 * P.sound_speed * P.sound_speed / P.gamma;
    }
    D.solid_visc = h->solidVisc.empty() ? 0.0f : h->solidVisc[0];
    D.totalKeys = uint32_t(totalKeys);

    int rc = 0;
#define ALLOC(ptr, n) if ((rc = dalloc(h, &(ptr), (n)))) return rc
    // the position arrays carry X_TAIL records more: the list builder's
    // candidate loads run up to that far past a cell run
    ALLOC(D.x, ncap + X_TAIL); ALLOC(D.v, ncap); ALLOC(D.id, ncap); ALLOC(D.ksum, ncap); ALLOC(D.kvsum, ncap); ALLOC(D.key, ncap);
    ALLOC(D.x2, ncap + X_TAIL); ALLOC(D.v2, ncap); ALLOC(D.id2, ncap); ALLOC(D.ksum2, ncap); ALLOC(D.kvsum2, ncap); ALLOC(D.key2, ncap);
    ALLOC(D.xk, ncap); ALLOC(D.vs, ncap); ALLOC(D.f4a, ncap); ALLOC(D.f4b, ncap);
    ALLOC(D.rho, ncap); ALLOC(D.alpha, ncap); ALLOC(D.kr, ncap); ALLOC(D.kr0, ncap); ALLOC(D.f1, ncap); ALLOC(D.f2, ncap);
    ALLOC(D.nbF, (size_t)NBF_CAP * ncap); ALLOC(D.nbB, (size_t)NBB_CAP * ncap); ALLOC(D.cnt, ncap);
    ALLOC(D.nbF16, (size_t)(NBF_CAP / 8) * ncap); ALLOC(D.nbB16, (size_t)(NBB_CAP / 8) * ncap);
    ALLOC(D.nbW16, (size_t)(NBF_CAP / 8) * ncap);
    ALLOC(D.nbTab, (size_t)NB16_TAB * (ncap / TILE));
    ALLOC(D.nbWin, (size_t)32 * (ncap / TILE));
    // boundary arrays carry one extra dummy particle at index nb (list padding)
    ALLOC(D.xb, nb + 1 + X_TAIL); ALLOC(D.vb, nb + 1); ALLOC(D.bid, nb + 1); ALLOC(D.bsolid, nb + 1);
    ALLOC(D.bscene, nb + 1); ALLOC(D.bkself, nb + 1); ALLOC(D.bkey, nb + 1);
    // solids: pose table (all solids), dynamic-solid state and re-sort targets
    D.nsolids = int(h->hsolids.size());
    D.anyDynamic = 0;
    for (const auto& sp : h->hsolids) D.anyDynamic |= sp.dynamic;
    if (D.nsolids > 0) {
        ALLOC(h->dSolids, D.nsolids); ALLOC(h->dSolidFT, 6 * (size_t)D.nsolids);
        D.solids = h->dSolids;
        HIPCHECK(h, hipMemcpy(h->dSolids, h->hsolids.data(), sizeof(SolidPose) * D.nsolids, hipMemcpyHostToDevice));
        HIPCHECK(h, hipMemset(h->dSolidFT, 0, sizeof(double) * 6 * D.nsolids));
        h->posesDirty = false;
    }
    if (D.anyDynamic) {
        ALLOC(D.bloc, nb + 1); ALLOC(D.bforce, nb + 1); ALLOC(D.brank, nb + 1);
        ALLOC(D.xb2, nb + 1 + X_TAIL); ALLOC(D.vb2, nb + 1); ALLOC(D.bloc2, nb + 1); ALLOC(D.bid2, nb + 1);
        ALLOC(D.bsolid2, nb + 1); ALLOC(D.bscene2, nb + 1); ALLOC(D.bkself2, nb + 1); ALLOC(D.bkey2, nb + 1);
        HIPCHECK(h, hipMemset(D.bforce, 0, sizeof(float4) * (nb + 1)));
    }
    SceneGrid* dgrids; int* drow; uint32_t* dkb;
    ALLOC(dgrids, ns); ALLOC(drow, rowRank.size()); ALLOC(dkb, ns + 1);
    D.grids = dgrids; D.rowRank = drow; D.sceneKeyBase = dkb;
    ALLOC(D.cellF, totalKeys); ALLOC(D.cellB, totalKeys);
    ALLOC(D.sceneStart, ns); ALLOC(D.sceneCount, ns); ALLOC(D.scenePadStart, ns + 1);
    int numTiles = ncap / TILE;
    ALLOC(D.tileScene, numTiles); ALLOC(D.tileValid, numTiles);
    ALLOC(D.dt, ns);
    {
        // solver-loop state in one block, laid out as its pinned host copy
        // (flags of even and odd iterations | iteration counts | errors, 8
        // bytes per scene each), so the host reads it with one copy
        double* loopBlk = nullptr;
        ALLOC(loopBlk, 3 * (size_t)ns);
        ALLOC(D.integ, ns);
        D.active = reinterpret_cast<int*>(loopBlk);
        D.active1 = D.active + ns;
        D.iters = reinterpret_cast<int*>(loopBlk + ns);
        D.sceneErr = loopBlk + 2 * ns;
        D.activeNext = nullptr;
        D.redMode = -1;
        D.redMinIt = D.redCap = 0;
        D.redTol = 0.0f;
    }
    ALLOC(D.partial, 2 * (size_t)numTiles); ALLOC(D.counters, C_COUNT); ALLOC(D.vmaxPartial, numTiles);
    ALLOC(h->layoutOut, 2); ALLOC(h->statsPart, 4 * (size_t)numTiles); ALLOC(h->statsOut, 4 * (size_t)ns);
    int sortItems = std::max<int>(ncap, int(nb));
    ALLOC(h->sk, sortItems); ALLOC(h->sv, sortItems);
    void* sortMem;
    HIPCHECK(h, hipMalloc(&sortMem, sort_scratch_bytes(sortItems)));
    h->allocs.push_back(sortMem);
    sort_scratch_bind(h->sort, sortMem, sortItems);
    if (P.solver == SPHGPU_IISPH_DF) {
        ALLOC(D.pm, ncap);
        ALLOC(D.brho, nb + 1); ALLOC(D.balpha, nb + 1); ALLOC(D.bk, nb + 1);
        HIPCHECK(h, hipMemset(D.bk, 0, sizeof(float) * (nb + 1)));   // the list-padding dummy stays 0
        HIPCHECK(h, hipMemset(D.pm, 0, sizeof(float) * ncap));
    }
    if ((P.solver == SPHGPU_DFSPH || P.solver == SPHGPU_IISPH) && P.warm_start == SPHGPU_WARM_LEARNED) {
        ALLOC(h->feat, 8 * (size_t)ncap); ALLOC(h->tgt, ncap); ALLOC(h->ypred, ncap);
        ALLOC(h->bfeat, 8 * (size_t)P.learn_batch); ALLOC(h->btgt, P.learn_batch);
        HIPCHECK(h, hipMemset(h->feat, 0, sizeof(float) * 8 * ncap));
        HIPCHECK(h, hipMemset(h->tgt, 0, sizeof(float) * ncap));
        if (mlp_alloc(h->mlp, P.learn_batch)) return fail(h, SPHGPU_E_DEVICE, "MLP allocation failed");
        h->mlpLive = true;
        h->mlp.overWeight = P.learn_over_weight;
        if (h->mlpHost.empty()) init_handle_mlp(h->mlpHost);
        HIPCHECK(h, hipMemcpy(h->mlp.params, h->mlpHost.data(), sizeof(float) * MLP_PARAMS, hipMemcpyHostToDevice));
        mlp_pack(h->mlp, h->st);
    }
#undef ALLOC
    // pinned host mirrors (one block)
    {
        size_t words = 4 * (size_t)ns + 2 * (size_t)ns + C_COUNT + 8 * (size_t)ns + 2 + ns + 16;
        void* hm;
        HIPCHECK(h, hipHostMalloc(&hm, words * 8));
        char* c = static_cast<char*>(hm);
        // hActive, hIters, hErr: the layout of the device block D.active
        h->hActive = reinterpret_cast<int*>(c); c += 8 * (size_t)ns;
        h->hIters = reinterpret_cast<int*>(c); c += 8 * (size_t)ns;
        h->hErr = reinterpret_cast<double*>(c); c += 8 * (size_t)ns;
        h->hStats = reinterpret_cast<double*>(c); c += 32 * (size_t)ns;
        h->hCounters = reinterpret_cast<int*>(c); c += 8 * C_COUNT;
        h->hLayout = reinterpret_cast<int*>(c); c += 16;
        h->hDt = reinterpret_cast<float*>(c);
    }

    hipStream_t st = h->st;
    HIPCHECK(h, hipMemcpyAsync(dgrids, h->hgrids.data(), sizeof(SceneGrid) * ns, hipMemcpyHostToDevice, st));
    HIPCHECK(h, hipMemcpyAsync(drow, rowRank.data(), sizeof(int) * rowRank.size(), hipMemcpyHostToDevice, st));
    HIPCHECK(h, hipMemcpyAsync(dkb, h->hKeyBase.data(), sizeof(uint32_t) * (ns + 1), hipMemcpyHostToDevice, st));
    HIPCHECK(h, hipMemsetAsync(D.cellF, 0, sizeof(int2) * totalKeys, st));
    HIPCHECK(h, hipMemsetAsync(D.cellB, 0, sizeof(int2) * totalKeys, st));
    HIPCHECK(h, hipMemsetAsync(D.counters, 0, sizeof(int) * C_COUNT, st));
    HIPCHECK(h, hipMemsetAsync(D.ksum, 0, sizeof(float) * ncap, st));
    HIPCHECK(h, hipMemsetAsync(D.kvsum, 0, sizeof(float) * ncap, st));
    HIPCHECK(h, hipMemsetAsync(D.cnt, 0, sizeof(int) * ncap, st));
    HIPCHECK(h, hipMemsetAsync(D.key, 0xFF, sizeof(uint32_t) * ncap, st));
    HIPCHECK(h, hipMemsetAsync(D.key2, 0xFF, sizeof(uint32_t) * ncap, st));

    // fluid upload in the initial padded layout
    {
        std::vector<float4> hx(ncap, make_float4(0, 0, 0, 0)), hv(ncap, make_float4(0, 0, 0, 0));
        std::vector<int> hid(ncap, -1);
        std::vector<int> fill(ns, 0);
        for (size_t i = 0; i < h->fscene.size(); ++i) {
            int sc = h->fscene[i];
            int slot = padStart[sc] + fill[sc]++;
            hx[slot] = make_float4(h->fx[3 * i], h->fx[3 * i + 1], h->fx[3 * i + 2], float(h->fphase[i]));
            hv[slot] = make_float4(h->fv[3 * i], h->fv[3 * i + 1], h->fv[3 * i + 2], 0.0f);
            hid[slot] = int(i);
        }
        std::vector<int> tileScene(numTiles, 0), tileValid(numTiles, 0);
        for (int sc = 0; sc < ns; ++sc)
            for (int t = padStart[sc] / TILE; t < padStart[sc + 1] / TILE; ++t) {
                tileScene[t] = sc;
                tileValid[t] = std::min(TILE, std::max(0, count[sc] - (t * TILE - padStart[sc])));
            }
        std::vector<int> sStart(ns, 0);
        HIPCHECK(h, hipMemcpy(D.x, hx.data(), sizeof(float4) * ncap, hipMemcpyHostToDevice));
        HIPCHECK(h, hipMemcpy(D.v, hv.data(), sizeof(float4) * ncap, hipMemcpyHostToDevice));
        HIPCHECK(h, hipMemcpy(D.id, hid.data(), sizeof(int) * ncap, hipMemcpyHostToDevice));
        HIPCHECK(h, hipMemcpy(D.tileScene, tileScene.data(), sizeof(int) * numTiles, hipMemcpyHostToDevice));
        HIPCHECK(h, hipMemcpy(D.tileValid, tileValid.data(), sizeof(int) * numTiles, hipMemcpyHostToDevice));
        HIPCHECK(h, hipMemcpy(D.sceneCount, count.data(), sizeof(int) * ns, hipMemcpyHostToDevice));
        HIPCHECK(h, hipMemcpy(D.scenePadStart, padStart.data(), sizeof(int) * (ns + 1), hipMemcpyHostToDevice));
        HIPCHECK(h, hipMemcpy(D.sceneStart, sStart.data(), sizeof(int) * ns, hipMemcpyHostToDevice));
        std::vector<float> dt0(ns, P.dt);
        HIPCHECK(h, hipMemcpy(D.dt, dt0.data(), sizeof(float) * ns, hipMemcpyHostToDevice));
        if (P.solver == SPHGPU_IISPH_DF) {
            // per-particle beta (default 1) rides in ksum
            std::vector<float> hb(ncap, 1.0f);
            for (int q = 0; q < ncap; ++q)
                if (hid[q] >= 0 && !h->fbeta.empty()) hb[q] = h->fbeta[hid[q]];
            HIPCHECK(h, hipMemcpy(D.ksum, hb.data(), sizeof(float) * ncap, hipMemcpyHostToDevice));
        }
        h->nLive = (long long)h->fscene.size();
    }
    // boundary: sort by cell once, Akinci volumes (prepareSolidPartWeight)
    if (nb > 0) {
        std::vector<float4> hb(nb);
        for (long long i = 0; i < nb; ++i) {
            float w[3];
            boundary_world(h, (size_t)i, w);
            hb[i] = make_float4(w[0], w[1], w[2], 0.0f);
        }
        float4* tx; int* tsolid; int* tscene;
        if ((rc = dalloc(h, &tx, nb))) return rc;
        if ((rc = dalloc(h, &tsolid, nb))) return rc;
        if ((rc = dalloc(h, &tscene, nb))) return rc;
        HIPCHECK(h, hipMemcpy(tx, hb.data(), sizeof(float4) * nb, hipMemcpyHostToDevice));
        HIPCHECK(h, hipMemcpy(tsolid, h->bsolid.data(), sizeof(int) * nb, hipMemcpyHostToDevice));
        HIPCHECK(h, hipMemcpy(tscene, h->bscene.data(), sizeof(int) * nb, hipMemcpyHostToDevice));
        launch_boundary_keys_from(D, tx, tscene, h->sk, h->sv, st);
        radix_sort_pairs(h->sk, h->sv, int(nb), h->keyBits, h->sort, st);
        launch_boundary_reorder(D, tx, tsolid, tscene, h->sk, h->sv, st);
        launch_cell_build(D, D.bkey, int(nb), D.cellB, st);
        launch_boundary_volume(D, st);
        if (D.anyDynamic) {
            // body-frame samples in sorted order, then the per-step update once.
            // tx is overwritten by a copy on the null stream, which does not
            // wait for the non-blocking stream: the kernels above that read
            // the world positions in tx must have finished first.
            std::vector<float4> hl(nb);
            for (long long i = 0; i < nb; ++i) hl[i] = make_float4(h->bx[3 * i], h->bx[3 * i + 1], h->bx[3 * i + 2], 0.0f);
            HIPCHECK(h, hipStreamSynchronize(st));
            HIPCHECK(h, hipMemcpy(tx, hl.data(), sizeof(float4) * nb, hipMemcpyHostToDevice));
            launch_boundary_init_local(D, tx, st);
        }
    }
    {
        // list-padding dummy at index nb (synchronous copies from the stack)
        const float4 far = make_float4(DUMMY_FAR, DUMMY_FAR, DUMMY_FAR, 0.0f), zero = make_float4(0, 0, 0, 0);
        HIPCHECK(h, hipStreamSynchronize(st));
        HIPCHECK(h, hipMemcpy(D.xb + nb, &far, sizeof(float4), hipMemcpyHostToDevice));
        HIPCHECK(h, hipMemcpy(D.vb + nb, &zero, sizeof(float4), hipMemcpyHostToDevice));
        if (D.anyDynamic) {
            HIPCHECK(h, hipMemcpy(D.xb2 + nb, &far, sizeof(float4), hipMemcpyHostToDevice));
            HIPCHECK(h, hipMemcpy(D.vb2 + nb, &zero, sizeof(float4), hipMemcpyHostToDevice));
        }
    }
    if (D.anyDynamic && (rc = rigid_update(h))) return rc;
    HIPCHECK(h, hipStreamSynchronize(st));
    HIPCHECK(h, hipGetLastError());
    return SPHGPU_OK;
}

int sphgpu_step(sphgpu_handle* h, sphgpu_stats* st) {
    if (!h) return SPHGPU_E_ARG;
    if (h->inCallback) return fail(h, SPHGPU_E_STATE, "sphgpu_step inside the update-solids callback");
    int rc = sphgpu_init(h);
    if (rc) return rc;
    if (h->nLive == 0 && h->frame > 0) return fail(h, SPHGPU_E_STATE, "no fluid particles left");
    if (!h->tuned && (rc = autotune(h))) return rc;
    // the window and occupancy choices are timed again on the flow itself,
    // after ten steps and every 200 after that (the initial lattice is not
    // the state the passes spend their time on); every choice gives the same
    // bits, so a run does not depend on the timings
    if (h->winFix < 0 && h->frame > 0 && (h->frame == 10 || h->frame % 200 == 0) && (rc = tune_windows(h)))
        return rc;
    for (auto& s : h->sceneStats) s = sphgpu_stats{};
    switch (h->P.solver) {
        case SPHGPU_WCSPH: return step_wcsph(h, st);
        case SPHGPU_DFSPH: return step_dfsph(h, st);
        case SPHGPU_PCISPH: return step_pcisph(h, st);
        case SPHGPU_IISPH_DF: return step_iisph_df(h, st);
        default: return step_iisph(h, st);
    }
}

int sphgpu_scene_stats(const sphgpu_handle* h, int scene, sphgpu_stats* st) {
    if (!h || !st || scene < 0 || scene >= (int)h->sceneStats.size()) return SPHGPU_E_ARG;
    *st = h->sceneStats[scene];
    return SPHGPU_OK;
}

int sphgpu_get_step_dt(sphgpu_handle* h, int scene, float* dt) {
    if (!h || !dt) return SPHGPU_E_ARG;
    if (!h->initialized) return fail(h, SPHGPU_E_STATE, "not initialised");
    if (scene < 0 || scene >= h->D.nscenes) return fail(h, SPHGPU_E_ARG, "scene %d out of range", scene);
    HIPCHECK(h, hipMemcpyAsync(dt, h->D.dt + scene, sizeof(float), hipMemcpyDeviceToHost, h->st));
    HIPCHECK(h, hipStreamSynchronize(h->st));
    return SPHGPU_OK;
}

int sphgpu_synchronize(sphgpu_handle* h) {
    if (!h) return SPHGPU_E_ARG;
    if (!h->initialized) return SPHGPU_OK;
    HIPCHECK(h, hipStreamSynchronize(h->st));
    return SPHGPU_OK;
}

long long sphgpu_num_fluid(const sphgpu_handle* h) {
    if (!h) return SPHGPU_E_ARG;
    return h->initialized ? h->nLive : (long long)h->fscene.size();
}

namespace {
// live slots in ascending id order: returns (slot, id) pairs
int live_slots(sphgpu_handle* h, std::vector<std::pair<int, int>>& out) {
    Dev& D = h->D;
    std::vector<int> ids(D.npad);
    HIPCHECK(h, hipStreamSynchronize(h->st));
    HIPCHECK(h, hipMemcpy(ids.data(), D.id, sizeof(int) * D.npad, hipMemcpyDeviceToHost));
    std::vector<int> tv(D.npad / TILE);
    HIPCHECK(h, hipMemcpy(tv.data(), D.tileValid, sizeof(int) * tv.size(), hipMemcpyDeviceToHost));
    out.clear();
    for (int i = 0; i < D.npad; ++i) {
        int t = i / TILE;
        if ((i - t * TILE) < tv[t] && ids[i] >= 0) out.push_back({ids[i], i});
    }
    std::sort(out.begin(), out.end());
    return 0;
}
}  // namespace

int sphgpu_get_fluid(sphgpu_handle* h, int field, float* out) {
    if (!h || !out) return SPHGPU_E_ARG;
    if (!h->initialized) return fail(h, SPHGPU_E_STATE, "not initialised");
    Dev& D = h->D;
    std::vector<std::pair<int, int>> ls;
    int rc = live_slots(h, ls);
    if (rc) return rc;
    int comp = 1;
    const void* src = nullptr;
    bool f4 = false;
    switch (field) {
        case SPHGPU_F_POS: src = D.x; f4 = true; comp = 3; break;
        case SPHGPU_F_VEL: src = D.v; f4 = true; comp = 3; break;
        case SPHGPU_F_RHO: src = D.rho; break;
        case SPHGPU_F_PRESSURE: src = (h->P.solver == SPHGPU_IISPH) ? D.kr : D.f1; break;
        case SPHGPU_F_ACC: src = (h->P.solver == SPHGPU_IISPH) ? D.f4b : D.f4a; f4 = true; comp = 3; break;
        case SPHGPU_F_ALPHA: src = D.alpha; break;
        case SPHGPU_F_KAPPA_SUM: src = D.ksum; break;
        case SPHGPU_F_PHASE: src = D.x; f4 = true; comp = 1; break;
        case SPHGPU_F_KAPPAV_SUM: src = D.kvsum; break;
        default: return fail(h, SPHGPU_E_ARG, "unknown field %d", field);
    }
    if (f4) {
        std::vector<float4> buf(D.npad);
        HIPCHECK(h, hipMemcpy(buf.data(), src, sizeof(float4) * D.npad, hipMemcpyDeviceToHost));
        for (size_t k = 0; k < ls.size(); ++k) {
            const float4& q = buf[ls[k].second];
            if (field == SPHGPU_F_PHASE) out[k] = q.w;
            else { out[3 * k] = q.x; out[3 * k + 1] = q.y; out[3 * k + 2] = q.z; }
        }
    } else {
        std::vector<float> buf(D.npad);
        HIPCHECK(h, hipMemcpy(buf.data(), src, sizeof(float) * D.npad, hipMemcpyDeviceToHost));
        for (size_t k = 0; k < ls.size(); ++k) out[k] = buf[ls[k].second];
    }
    (void)comp;
    return SPHGPU_OK;
}

int sphgpu_get_fluid_ids(sphgpu_handle* h, long long* out) {
    if (!h || !out) return SPHGPU_E_ARG;
    if (!h->initialized) return fail(h, SPHGPU_E_STATE, "not initialised");
    std::vector<std::pair<int, int>> ls;
    int rc = live_slots(h, ls);
    if (rc) return rc;
    for (size_t k = 0; k < ls.size(); ++k) out[k] = ls[k].first;
    return SPHGPU_OK;
}

// ---- frame output ---------------------------------------------------------
int sphgpu_snapshot_wait(sphgpu_handle* h) {
    if (!h) return SPHGPU_E_ARG;
    if (!h->snap.pending) return SPHGPU_OK;
    HIPCHECK(h, hipEventSynchronize(h->snap.done));
    h->snap.pending = false;
    h->snap.valid = true;
    return SPHGPU_OK;
}

int sphgpu_snapshot_begin(sphgpu_handle* h) {
    if (!h) return SPHGPU_E_ARG;
    if (h->inCallback) return fail(h, SPHGPU_E_STATE, "snapshot inside the update-solids callback");
    int rc = sphgpu_init(h);
    if (rc) return rc;
    if ((rc = sphgpu_snapshot_wait(h))) return rc;
    Dev& D = h->D;
    auto& S = h->snap;
    const size_t cap = (size_t)D.ncap, tiles = (size_t)D.ncap / TILE;
    if (!S.cs) {
        if ((rc = dalloc(h, &S.x, cap)) || (rc = dalloc(h, &S.v, cap)) || (rc = dalloc(h, &S.a, cap)) ||
            (rc = dalloc(h, &S.rho, cap)) || (rc = dalloc(h, &S.p, cap)) || (rc = dalloc(h, &S.id, cap)) ||
            (rc = dalloc(h, &S.tv, tiles)))
            return rc;
        const size_t bytes = cap * (3 * sizeof(float4) + 2 * sizeof(float) + sizeof(int)) + tiles * sizeof(int);
        HIPCHECK(h, hipHostMalloc(&S.host, bytes));
        char* c = static_cast<char*>(S.host);
        S.hx = reinterpret_cast<float4*>(c); c += cap * sizeof(float4);
        S.hv = reinterpret_cast<float4*>(c); c += cap * sizeof(float4);
        S.ha = reinterpret_cast<float4*>(c); c += cap * sizeof(float4);
        S.hrho = reinterpret_cast<float*>(c); c += cap * sizeof(float);
        S.hp = reinterpret_cast<float*>(c); c += cap * sizeof(float);
        S.hid = reinterpret_cast<int*>(c); c += cap * sizeof(int);
        S.htv = reinterpret_cast<int*>(c);
        HIPCHECK(h, hipStreamCreateWithFlags(&S.cs, hipStreamNonBlocking));
        HIPCHECK(h, hipEventCreateWithFlags(&S.copied, hipEventDisableTiming));
        HIPCHECK(h, hipEventCreateWithFlags(&S.done, hipEventDisableTiming));
    }
    const bool ii = h->P.solver == SPHGPU_IISPH;
    const size_t n = (size_t)D.npad, nt = (size_t)D.npad / TILE;
    // device-side copies in stream order (microseconds), then the PCIe
    // transfer on the second stream while later steps run
    HIPCHECK(h, hipMemcpyAsync(S.x, D.x, sizeof(float4) * n, hipMemcpyDeviceToDevice, h->st));
    HIPCHECK(h, hipMemcpyAsync(S.v, D.v, sizeof(float4) * n, hipMemcpyDeviceToDevice, h->st));
    HIPCHECK(h, hipMemcpyAsync(S.a, ii ? D.f4b : D.f4a, sizeof(float4) * n, hipMemcpyDeviceToDevice, h->st));
    HIPCHECK(h, hipMemcpyAsync(S.rho, D.rho, sizeof(float) * n, hipMemcpyDeviceToDevice, h->st));
    HIPCHECK(h, hipMemcpyAsync(S.p, ii ? D.kr : D.f1, sizeof(float) * n, hipMemcpyDeviceToDevice, h->st));
    HIPCHECK(h, hipMemcpyAsync(S.id, D.id, sizeof(int) * n, hipMemcpyDeviceToDevice, h->st));
    HIPCHECK(h, hipMemcpyAsync(S.tv, D.tileValid, sizeof(int) * nt, hipMemcpyDeviceToDevice, h->st));
    HIPCHECK(h, hipEventRecord(S.copied, h->st));
    HIPCHECK(h, hipStreamWaitEvent(S.cs, S.copied, 0));
    HIPCHECK(h, hipMemcpyAsync(S.hx, S.x, sizeof(float4) * n, hipMemcpyDeviceToHost, S.cs));
    HIPCHECK(h, hipMemcpyAsync(S.hv, S.v, sizeof(float4) * n, hipMemcpyDeviceToHost, S.cs));
    HIPCHECK(h, hipMemcpyAsync(S.ha, S.a, sizeof(float4) * n, hipMemcpyDeviceToHost, S.cs));
    HIPCHECK(h, hipMemcpyAsync(S.hrho, S.rho, sizeof(float) * n, hipMemcpyDeviceToHost, S.cs));
    HIPCHECK(h, hipMemcpyAsync(S.hp, S.p, sizeof(float) * n, hipMemcpyDeviceToHost, S.cs));
    HIPCHECK(h, hipMemcpyAsync(S.hid, S.id, sizeof(int) * n, hipMemcpyDeviceToHost, S.cs));
    HIPCHECK(h, hipMemcpyAsync(S.htv, S.tv, sizeof(int) * nt, hipMemcpyDeviceToHost, S.cs));
    HIPCHECK(h, hipEventRecord(S.done, S.cs));
    S.npad = D.npad;
    S.frame = h->frame;
    S.pending = true;
    S.valid = false;
    return SPHGPU_OK;
}

// live slots of one phase in the capture, ascending original id
static int snapshot_phase(sphgpu_handle* h, int phase, std::vector<int>& slots) {
    int rc;
    if (!h->snap.valid && !h->snap.pending && (rc = sphgpu_snapshot_begin(h))) return rc;
    if ((rc = sphgpu_snapshot_wait(h))) return rc;
    const auto& S = h->snap;
    std::vector<std::pair<int, int>> ls;
    for (int i = 0; i < S.npad; ++i) {
        const int t = i / TILE;
        if ((i - t * TILE) < S.htv[t] && S.hid[i] >= 0 && int(S.hx[i].w) == phase) ls.push_back({S.hid[i], i});
    }
    std::sort(ls.begin(), ls.end());
    slots.resize(ls.size());
    for (size_t k = 0; k < ls.size(); ++k) slots[k] = ls[k].second;
    return 0;
}

int sphgpu_write_pos(sphgpu_handle* h, const char* path, int phase, unsigned components, const char* log_name) {
    if (!h || !path) return SPHGPU_E_ARG;
    const unsigned known = SPHGPU_POS_POSITION | SPHGPU_POS_VELOCITY | SPHGPU_POS_DENSITY | SPHGPU_POS_PRESURE |
                           SPHGPU_POS_ACCELERATION;
    if (components == 0 || (components & ~known)) return fail(h, SPHGPU_E_ARG, "bad component mask %#x", components);
    if (phase < 0 || phase >= (int)h->phases.size()) return fail(h, SPHGPU_E_ARG, "phase %d out of range", phase);
    std::vector<int> slots;
    if (int rc = snapshot_phase(h, phase, slots)) return rc;
    const auto& S = h->snap;
    const size_t n = slots.size();
    std::vector<float> x(3 * n), v(3 * n), a(3 * n), rho(n), p(n);
    for (size_t k = 0; k < n; ++k) {
        const int i = slots[k];
        x[3 * k] = S.hx[i].x; x[3 * k + 1] = S.hx[i].y; x[3 * k + 2] = S.hx[i].z;
        v[3 * k] = S.hv[i].x; v[3 * k + 1] = S.hv[i].y; v[3 * k + 2] = S.hv[i].z;
        a[3 * k] = S.ha[i].x; a[3 * k + 1] = S.ha[i].y; a[3 * k + 2] = S.ha[i].z;
        rho[k] = S.hrho[i];
        p[k] = S.hp[i];
    }
    if (sphio::write_pos(path, log_name, S.frame, "fluid", phase, (long long)n, components, x.data(), v.data(),
                         rho.data(), p.data(), a.data()))
        return fail(h, SPHGPU_E_ARG, "cannot write %s", path);
    return SPHGPU_OK;
}

int sphgpu_write_json(sphgpu_handle* h, const char* path, int phase) {
    if (!h || !path) return SPHGPU_E_ARG;
    if (phase < 0 || phase >= (int)h->phases.size()) return fail(h, SPHGPU_E_ARG, "phase %d out of range", phase);
    std::vector<int> slots;
    if (int rc = snapshot_phase(h, phase, slots)) return rc;
    const auto& S = h->snap;
    const size_t n = slots.size();
    std::vector<float> x(3 * n), v(3 * n);
    for (size_t k = 0; k < n; ++k) {
        const int i = slots[k];
        x[3 * k] = S.hx[i].x; x[3 * k + 1] = S.hx[i].y; x[3 * k + 2] = S.hx[i].z;
        v[3 * k] = S.hv[i].x; v[3 * k + 1] = S.hv[i].y; v[3 * k + 2] = S.hv[i].z;
    }
    // the particle's d: half the support radius (SphBase.cpp:883)
    if (sphio::write_json(path, (long long)n, x.data(), v.data(), 0.5f * h->P.smooth_radius_h))
        return fail(h, SPHGPU_E_ARG, "cannot write %s", path);
    return SPHGPU_OK;
}

int sphgpu_write_solid_pos(sphgpu_handle* h, const char* path, int solid, const char* log_name) {
    if (!h || !path) return SPHGPU_E_ARG;
    if (solid < 0 || solid >= (int)h->hsolids.size()) return fail(h, SPHGPU_E_ARG, "solid %d out of range", solid);
    int rc = sphgpu_init(h);
    if (rc) return rc;
    Dev& D = h->D;
    std::vector<float4> xb(D.nb);
    std::vector<int> bid(D.nb);
    HIPCHECK(h, hipStreamSynchronize(h->st));
    HIPCHECK(h, hipMemcpy(xb.data(), D.xb, sizeof(float4) * D.nb, hipMemcpyDeviceToHost));
    HIPCHECK(h, hipMemcpy(bid.data(), D.bid, sizeof(int) * D.nb, hipMemcpyDeviceToHost));
    const SolidPose& sp = h->hsolids[solid];
    const int n = sp.endId - sp.firstId;
    std::vector<float> x(3 * (size_t)n);
    for (int r = 0; r < D.nb; ++r) {
        const int k = bid[r] - sp.firstId;
        if (k < 0 || k >= n) continue;
        x[3 * k] = xb[r].x; x[3 * k + 1] = xb[r].y; x[3 * k + 2] = xb[r].z;
    }
    if (sphio::write_pos(path, log_name, h->frame, "solid", solid, n, SPHGPU_POS_POSITION, x.data(), nullptr,
                         nullptr, nullptr, nullptr))
        return fail(h, SPHGPU_E_ARG, "cannot write %s", path);
    return SPHGPU_OK;
}

int sphgpu_get_boundary_volume(sphgpu_handle* h, float* out) {
    if (!h || !out) return SPHGPU_E_ARG;
    int rc = sphgpu_init(h);
    if (rc) return rc;
    Dev& D = h->D;
    std::vector<float4> xb(D.nb);
    std::vector<int> bid(D.nb);
    HIPCHECK(h, hipStreamSynchronize(h->st));
    HIPCHECK(h, hipMemcpy(xb.data(), D.xb, sizeof(float4) * D.nb, hipMemcpyDeviceToHost));
    HIPCHECK(h, hipMemcpy(bid.data(), D.bid, sizeof(int) * D.nb, hipMemcpyDeviceToHost));
    for (int i = 0; i < D.nb; ++i) out[bid[i]] = xb[i].w;
    return SPHGPU_OK;
}

int sphgpu_get_neighbours(sphgpu_handle* h, int cap, int* counts, int* kinds, long long* ids, float* dis) {
    if (!h || cap <= 0 || !counts || !kinds || !ids || !dis) return SPHGPU_E_ARG;
    int rc = sphgpu_init(h);
    if (rc) return rc;
    Dev& D = h->D;
    run_ns(h);
    if (h->P.solver == SPHGPU_DFSPH && h->frame > 0) launch_df_density_alpha(D, h->st);
    std::vector<std::pair<int, int>> ls;
    if ((rc = live_slots(h, ls))) return rc;
    std::vector<int> rankOf(D.npad, -1);
    for (size_t k = 0; k < ls.size(); ++k) rankOf[ls[k].second] = int(k);
    size_t n = ls.size();
    int *dRank, *dCounts, *dKinds;
    long long* dIds;
    float* dDis;
    HIPCHECK(h, hipMalloc(&dRank, sizeof(int) * std::max<size_t>(D.npad, 1)));
    HIPCHECK(h, hipMalloc(&dCounts, sizeof(int) * std::max<size_t>(n, 1)));
    HIPCHECK(h, hipMalloc(&dKinds, sizeof(int) * std::max<size_t>(n * cap, 1)));
    HIPCHECK(h, hipMalloc(&dIds, sizeof(long long) * std::max<size_t>(n * cap, 1)));
    HIPCHECK(h, hipMalloc(&dDis, sizeof(float) * std::max<size_t>(n * cap, 1)));
    HIPCHECK(h, hipMemcpy(dRank, rankOf.data(), sizeof(int) * D.npad, hipMemcpyHostToDevice));
    launch_neighbour_dump_ranked(D, cap, dCounts, dKinds, dIds, dDis, dRank, h->st);
    HIPCHECK(h, hipStreamSynchronize(h->st));
    HIPCHECK(h, hipGetLastError());
    std::vector<int> hk(n * cap);
    std::vector<long long> hi(n * cap);
    std::vector<float> hd(n * cap);
    HIPCHECK(h, hipMemcpy(counts, dCounts, sizeof(int) * n, hipMemcpyDeviceToHost));
    HIPCHECK(h, hipMemcpy(hk.data(), dKinds, sizeof(int) * n * cap, hipMemcpyDeviceToHost));
    HIPCHECK(h, hipMemcpy(hi.data(), dIds, sizeof(long long) * n * cap, hipMemcpyDeviceToHost));
    HIPCHECK(h, hipMemcpy(hd.data(), dDis, sizeof(float) * n * cap, hipMemcpyDeviceToHost));
    hipFree(dRank); hipFree(dCounts); hipFree(dKinds); hipFree(dIds); hipFree(dDis);
    for (size_t k = 0; k < n; ++k) {
        int c = std::min(counts[k], cap);
        std::vector<int> order(c);
        std::iota(order.begin(), order.end(), 0);
        size_t b = k * cap;
        std::sort(order.begin(), order.end(), [&](int a, int z) {
            return hk[b + a] != hk[b + z] ? hk[b + a] < hk[b + z] : hi[b + a] < hi[b + z];
        });
        for (int q = 0; q < c; ++q) {
            kinds[b + q] = hk[b + order[q]];
            ids[b + q] = hi[b + order[q]];
            dis[b + q] = hd[b + order[q]];
        }
    }
    return SPHGPU_OK;
}

int sphgpu_copy_probe(long long bytes, double* ms) {
    if (bytes < 16 || bytes > (1LL << 32) || !ms) return SPHGPU_E_ARG;
    if (!sphgpu_device_available()) return SPHGPU_E_DEVICE;
    size_t n = size_t(bytes) / 16;
    float4 *a = nullptr, *b = nullptr;
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = SPHGPU_E_DEVICE;
    if (hipMalloc(&a, n * 16) == hipSuccess && hipMalloc(&b, n * 16) == hipSuccess &&
        hipStreamCreateWithFlags(&st, hipStreamNonBlocking) == hipSuccess && hipEventCreate(&e0) == hipSuccess &&
        hipEventCreate(&e1) == hipSuccess) {
        launch_fill_i32(reinterpret_cast<int*>(a), 1, int(std::min<size_t>(n * 4, 1u << 30)), st);
        launch_copy_probe(a, b, n, st);
        hipEventRecord(e0, st);
        launch_copy_probe(a, b, n, st);
        hipEventRecord(e1, st);
        float t = 0.0f;
        if (hipEventSynchronize(e1) == hipSuccess && hipEventElapsedTime(&t, e0, e1) == hipSuccess &&
            hipGetLastError() == hipSuccess) {
            *ms = t;
            rc = SPHGPU_OK;
        }
    }
    if (e0) hipEventDestroy(e0);
    if (e1) hipEventDestroy(e1);
    if (st) hipStreamDestroy(st);
    if (a) hipFree(a);
    if (b) hipFree(b);
    return rc;
}

int sphgpu_profile_kernel(sphgpu_handle* h, const char* name) {
    if (!h) return SPHGPU_E_ARG;
    h->profName = name ? name : "";
    h->profUsed = 0;
    return SPHGPU_OK;
}

int sphgpu_kernel_timing(sphgpu_handle* h, long long* launches, double* total_ms, double* total_bytes) {
    if (!h) return SPHGPU_E_ARG;
    if (h->initialized) HIPCHECK(h, hipStreamSynchronize(h->st));
    double ms = 0, bytes = 0;
    long long n = 0;
    for (size_t k = 0; k < h->profUsed; ++k) {
        if (h->prof[k].skip) continue;
        float t = 0;
        HIPCHECK(h, hipEventElapsedTime(&t, h->prof[k].a, h->prof[k].b));
        ms += t;
        bytes += h->prof[k].bytes;
        ++n;
    }
    if (launches) *launches = n;
    if (total_ms) *total_ms = ms;
    if (total_bytes) *total_bytes = bytes;
    return SPHGPU_OK;
}

int sphgpu_get_tuning(const sphgpu_handle* h, int* choice, float* ms) {
    if (!h || !choice) return SPHGPU_E_ARG;
    if (!h->tuned) return SPHGPU_E_STATE;
    choice[0] = h->D.tileMap;
    choice[1] = h->D.nbStage;
    if (ms)
        for (int k = 0; k < 4; ++k) ms[k] = h->tuneMs[k];
    return SPHGPU_OK;
}

int sphgpu_set_tuning(sphgpu_handle* h, int tile_map, int stage_groups, int stream_lists) {
    if (!h) return SPHGPU_E_ARG;
    if (h->tuned) return fail(h, SPHGPU_E_STATE, "tuning is chosen before the first step");
    if (tile_map < -1 || tile_map > 1 || !(stage_groups == -1 || stage_groups == 6 || stage_groups == 8) ||
        stream_lists < -1 || stream_lists > 1)
        return fail(h, SPHGPU_E_ARG, "tile_map in {-1,0,1}, stage_groups in {-1,6,8}, stream_lists in {-1,0,1}");
    h->tuneFix[0] = tile_map;
    h->tuneFix[1] = stage_groups;
    h->tuneFix[2] = stream_lists;
    return SPHGPU_OK;
}

int sphgpu_set_lds_windows(sphgpu_handle* h, int on) {
    if (!h || on < 0 || on > 1) return SPHGPU_E_ARG;
    h->D.winMask = on ? WIN_ALL : 0;
    h->winFix = on;
    return SPHGPU_OK;
}

int sphgpu_get_lds_windows(const sphgpu_handle* h, int* mask, float* ms, int* blocks_per_cu) {
    if (!h || !mask) return SPHGPU_E_ARG;
    *mask = h->D.winMask;
    if (ms)
        for (int k = 0; k < 10; ++k) ms[k] = h->winMs[k];
    if (blocks_per_cu)
        for (int k = 0; k < 5; ++k) {
            const int l = h->D.gatherLds[k];
            blocks_per_cu[k] = (h->D.winMask >> k & 1) || l == 0 ? 0 : (160 * 1024) / (l + 1024);
        }
    return SPHGPU_OK;
}

int sphgpu_get_list_form(sphgpu_handle* h, int* packed_tiles, int* tiles, int* streamed) {
    if (!h || !packed_tiles || !tiles) return SPHGPU_E_ARG;
    if (streamed) *streamed = h->D.listNT;
    if (!h->initialized || !h->listsBuilt) return fail(h, SPHGPU_E_STATE, "no neighbour lists built yet");
    const int nt = h->D.npad / TILE;
    std::vector<int> flag(nt), valid(nt);
    HIPCHECK(h, hipStreamSynchronize(h->st));
    if (nt > 0) {
        HIPCHECK(h, hipMemcpy2D(flag.data(), sizeof(int), h->D.nbTab + NB16_FLAG, sizeof(int) * NB16_TAB, sizeof(int),
                                nt, hipMemcpyDeviceToHost));
        HIPCHECK(h, hipMemcpy(valid.data(), h->D.tileValid, sizeof(int) * nt, hipMemcpyDeviceToHost));
    }
    int p = 0, t = 0;
    for (int k = 0; k < nt; ++k)
        if (valid[k] > 0) {
            ++t;
            p += flag[k] != 0;
        }
    *packed_tiles = p;
    *tiles = t;
    return SPHGPU_OK;
}

int sphgpu_get_mlp_params(sphgpu_handle* h, float* params) {
    if (!h || !params) return SPHGPU_E_ARG;
    if (h->P.warm_start != SPHGPU_WARM_LEARNED) return fail(h, SPHGPU_E_STATE, "handle has no learned warm start");
    if (h->mlpLive) {
        HIPCHECK(h, hipMemcpyAsync(params, h->mlp.params, sizeof(float) * MLP_PARAMS, hipMemcpyDeviceToHost, h->st));
        HIPCHECK(h, hipStreamSynchronize(h->st));
        return SPHGPU_OK;
    }
    if (h->mlpHost.empty()) init_handle_mlp(h->mlpHost);
    std::memcpy(params, h->mlpHost.data(), sizeof(float) * MLP_PARAMS);
    return SPHGPU_OK;
}

int sphgpu_set_mlp_params(sphgpu_handle* h, const float* params) {
    if (!h || !params) return SPHGPU_E_ARG;
    if (h->P.warm_start != SPHGPU_WARM_LEARNED) return fail(h, SPHGPU_E_STATE, "handle has no learned warm start");
    for (int k = 0; k < MLP_PARAMS; ++k)
        if (!std::isfinite(params[k])) return fail(h, SPHGPU_E_ARG, "parameter %d is not finite", k);
    h->mlpHost.assign(params, params + MLP_PARAMS);
    h->mlpPreset = true;
    if (h->mlpLive) {
        HIPCHECK(h, hipMemcpyAsync(h->mlp.params, h->mlpHost.data(), sizeof(float) * MLP_PARAMS, hipMemcpyHostToDevice,
                                   h->st));
        mlp_pack(h->mlp, h->st);
        HIPCHECK(h, hipStreamSynchronize(h->st));
    }
    return SPHGPU_OK;
}

}  // extern "C"
