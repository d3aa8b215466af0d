// C ABI of libnpx_hip.so (declarations and reference citations: include/npx_esrrf.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/npx_esrrf.h"
#include "npx_kernels.h"
#include "npx_host.h"
#include "npx_plan.h"

namespace npx {
namespace host {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

int hip_fail(hipError_t e, const char* what) {
    return fail(NPX_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

// ------------------------------------------------------------------------------------------ plan cache

struct DevCopy {
    void* mem = nullptr;
    npx::DevTables t{};
};

struct Plan {
    npx::PlanHost host;
    std::map<int, DevCopy> dev;  // per device id
    TableBlob blob;              // host copy of the packed tables, for npx_esrrf_temporal_static (built on its first use)
};

using Key = std::tuple<int, int, int, uint64_t>;
std::mutex g_mu;
std::map<Key, std::unique_ptr<Plan>> g_plans;

uint64_t bits(double v) {
    uint64_t b;
    std::memcpy(&b, &v, sizeof b);
    return b;
}

// Returns the cached host plan (built on first use) or nullptr with g_err set.
Plan* get_plan(int H, int W, int M, double R) {
    const Key k{H, W, M, bits(R)};
    auto it = g_plans.find(k);
    if (it != g_plans.end()) return it->second.get();
    auto p = std::make_unique<Plan>();
    std::string err;
    if (!npx::build_plan(H, W, M, R, p->host, &err)) {
        g_err = err;
        return nullptr;
    }
    Plan* raw = p.get();
    g_plans.emplace(k, std::move(p));
    return raw;
}

size_t align256(size_t x) { return (x + 255) & ~static_cast<size_t>(255); }

// One contiguous blob of the plan's device tables (the four tap tables as float4 weights + int bases), every segment
// 256-byte aligned; tables_in(base) gives the device view of the blob copied to `base`. The ring geometry is not in
// it: the kernels form each block's from its pixel centres.
void pack_tables(const PlanHost& h, TableBlob* out) {
    auto taps = [](const std::vector<CrTap>& v, std::vector<float>& w, std::vector<int>& b) {
        w.resize(v.size() * 4);
        b.resize(v.size());
        for (size_t i = 0; i < v.size(); ++i) {
            b[i] = v[i].base;
            for (int u = 0; u < 4; ++u) w[4 * i + u] = v[i].w[u];
        }
    };
    std::vector<float> lrw, lcw, crw, ccw;
    std::vector<int> lrb, lcb, crb, ccb;
    taps(h.lat_row, lrw, lrb);
    taps(h.lat_col, lcw, lcb);
    taps(h.cr_row, crw, crb);
    taps(h.cr_col, ccw, ccb);
    const void* src[8] = {lrw.data(), lrb.data(), lcw.data(), lcb.data(), crw.data(), crb.data(), ccw.data(), ccb.data()};
    const size_t bytes[8] = {lrw.size() * 4, lrb.size() * 4, lcw.size() * 4, lcb.size() * 4,
                             crw.size() * 4, crb.size() * 4, ccw.size() * 4, ccb.size() * 4};
    size_t total = 0;
    for (int i = 0; i < 8; ++i) {
        out->off[i] = total;
        total += align256(bytes[i]);
    }
    out->bytes.assign(total, 0);
    for (int i = 0; i < 8; ++i) std::memcpy(out->bytes.data() + out->off[i], src[i], bytes[i]);
}

DevTables TableBlob::tables_in(void* mem) const {
    auto* base = static_cast<uint8_t*>(mem);
    DevTables t;
    t.lrw = reinterpret_cast<const float4*>(base + off[0]);
    t.lrb = reinterpret_cast<const int*>(base + off[1]);
    t.lcw = reinterpret_cast<const float4*>(base + off[2]);
    t.lcb = reinterpret_cast<const int*>(base + off[3]);
    t.crw = reinterpret_cast<const float4*>(base + off[4]);
    t.crb = reinterpret_cast<const int*>(base + off[5]);
    t.ccw = reinterpret_cast<const float4*>(base + off[6]);
    t.ccb = reinterpret_cast<const int*>(base + off[7]);
    return t;
}

// Uploads the plan's tables to the current device once (blocking; completes before returning).
int device_tables(Plan* p, const npx::DevTables** out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
    auto it = p->dev.find(dev);
    if (it != p->dev.end()) {
        *out = &it->second.t;
        return NPX_OK;
    }
    TableBlob blob;
    pack_tables(p->host, &blob);
    void* mem = nullptr;
    e = hipMalloc(&mem, blob.bytes.size());
    if (e != hipSuccess) return hip_fail(e, "hipMalloc(plan tables)");
    e = hipMemcpy(mem, blob.bytes.data(), blob.bytes.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e != hipSuccess) {
        (void)hipFree(mem);
        return hip_fail(e, "upload of plan tables");
    }
    DevCopy dc;
    dc.mem = mem;
    dc.t = blob.tables_in(mem);
    auto res = p->dev.emplace(dev, dc);
    *out = &res.first->second.t;
    return NPX_OK;
}

// ------------------------------------------------------------------------------------------ kernel timing

// Optional per-call HIP events around the fused kernel and the merge kernel (npx_timing / npx_timing_read).
struct CallEvents {
    hipEvent_t e0, e1, e2;
};
std::mutex g_tmu;
bool g_timing = false;
std::vector<CallEvents> g_events;

bool timing_begin(hipStream_t st, CallEvents* ev) {
    std::lock_guard<std::mutex> lk(g_tmu);
    if (!g_timing || capturing(st)) return false;
    if (hipEventCreate(&ev->e0) != hipSuccess) return false;
    if (hipEventCreate(&ev->e1) != hipSuccess) return false;
    if (hipEventCreate(&ev->e2) != hipSuccess) return false;
    return hipEventRecord(ev->e0, st) == hipSuccess;
}

void timing_end(const CallEvents& ev) {
    std::lock_guard<std::mutex> lk(g_tmu);
    g_events.push_back(ev);
}

// ------------------------------------------------------------------------------------------ launch shape

// Launch shape of the fused kernels: at most max_waves wavefronts (one output phase each) per block, and the
// number of blocks the temporal kernel aims for when it splits the frames into chunks. Defaults (npx_set_launch)
// apply unless the autotuner (npx_autotune / npx_tune_set) stored a shape for the (H, W, M, R) of the call.
struct LaunchCfg {
    int max_waves, target_blocks;
};
std::mutex g_cfg_mu;
LaunchCfg g_default_cfg{16, 2048};
std::map<Key, LaunchCfg> g_tuned;
std::map<Key, LaunchCfg> g_tuned_frames;  // shapes of the per-frame-map kernels, where tuned apart from the temporal

LaunchCfg launch_cfg(int H, int W, int M, double R) {
    std::lock_guard<std::mutex> lk(g_cfg_mu);
    auto it = g_tuned.find(Key{H, W, M, bits(R)});
    return it != g_tuned.end() ? it->second : g_default_cfg;
}

// Launch shape of the per-frame-map kernels (npx_esrrf_frames, npx_rgc's fused path): their own tuned shape, else the
// temporal kernel's (tuned or default).
LaunchCfg launch_cfg_frames(int H, int W, int M, double R) {
    std::lock_guard<std::mutex> lk(g_cfg_mu);
    const Key k{H, W, M, bits(R)};
    auto it = g_tuned_frames.find(k);
    if (it != g_tuned_frames.end()) return it->second;
    auto jt = g_tuned.find(k);
    return jt != g_tuned.end() ? jt->second : g_default_cfg;
}

// Compute units of the current device (256 on MI355X; also when no device answers, so host-side layout queries
// stay deterministic).
int device_cus() {
    int dev = 0, n = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n < 1) {
        (void)hipGetLastError();
        return 256;
    }
    return n;
}

LaunchCfg default_cfg() {
    std::lock_guard<std::mutex> lk(g_cfg_mu);
    return g_default_cfg;
}

Shape fused_shape(int H, int W, int M, const LaunchCfg& c) {
    Shape s;
    const int nphase = M * M;
    s.ngroups = (nphase + c.max_waves - 1) / c.max_waves;
    s.nwaves = (nphase + s.ngroups - 1) / s.ngroups;
    s.nregx = (W + 63) / 64;
    s.nregy = (H + npx::kQTemporal - 1) / npx::kQTemporal;
    return s;
}

// Frame chunking of the temporal kernel: about target_blocks blocks, while each chunk keeps >= 6 frames. A chunk adds
// 5 partial moments per output pixel (20 B written and read once by the merge) against the ring work of its frames
// (~700 flops per output pixel and frame), so the merge stays a small fraction of the work; images large enough to
// fill the chip take one chunk. (100 x 64^2 at M5, 16 blocks per chunk: chunks of >= 32 frames 0.37 ms, >= 8 frames
// 12 chunks 0.139 ms, >= 6 frames 15 chunks, 240 blocks for the 256 CUs, 0.119 ms.)
void temporal_chunks(int N, const Shape& s, int target_blocks, int* nchunk, int* fpc) {
    const long long spatial = static_cast<long long>(s.nregx) * s.nregy * s.ngroups;
    long long want = (target_blocks + spatial - 1) / spatial;
    want = std::min<long long>(want, std::max(1, N / 6));
    want = std::max<long long>(want, 1);
    int f = static_cast<int>((N + want - 1) / want);
    *fpc = std::max(f, 1);
    *nchunk = (N + *fpc - 1) / *fpc;
}

// Frame chunking of the per-frame-map kernel: about target_blocks blocks (per-frame maps need no merge, so the only
// cost of a chunk is the block's set-up: tables, first tile).
void frames_chunks(int N, const Shape& s, int target_blocks, int* nchunk, int* fpc) {
    const long long spatial = static_cast<long long>(s.nregx) * s.nregy * s.ngroups;
    long long want = std::max<long long>(1, (target_blocks + spatial - 1) / spatial);
    want = std::min<long long>({want, N, 65535 / s.ngroups});
    *fpc = static_cast<int>((N + want - 1) / want);
    *nchunk = (N + *fpc - 1) / *fpc;
}

// Times candidate launch shapes of one entry point on the caller's data and returns the fastest in *best (its time
// per call in *best_ms). use(c) makes shape c the one the next call takes; call() makes one call on `st`. The GPU is
// first warmed up with the first candidate (calls for >= 20 ms, so that its clocks have settled), then every
// candidate is timed in three rounds, each a batch of calls of about 2 ms or more, keeping its best batch: with the
// rounds interleaved, clock drift during tuning does not favour the candidates timed last.
int time_candidates(hipStream_t st, const std::vector<LaunchCfg>& cands,
                    const std::function<void(const LaunchCfg&)>& use, const std::function<int()>& call,
                    LaunchCfg* best, float* best_ms) {
    if (cands.empty()) return fail(NPX_ERR_ARG, "no launch shape to time");
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    if (e != hipSuccess) {
        if (e0) (void)hipEventDestroy(e0);
        return hip_fail(e, "hipEventCreate(autotune)");
    }
    auto timed = [&](int n, float* ms) -> int {
        hipError_t err = hipEventRecord(e0, st);
        if (err != hipSuccess) return hip_fail(err, "hipEventRecord(autotune)");
        for (int k = 0; k < n; ++k) {
            const int r = call();
            if (r) return r;
        }
        err = hipEventRecord(e1, st);
        if (err == hipSuccess) err = hipEventSynchronize(e1);
        if (err == hipSuccess) err = hipEventElapsedTime(ms, e0, e1);
        return err == hipSuccess ? NPX_OK : hip_fail(err, "autotune timing");
    };
    int rc = NPX_OK;
    float t = 0.f, warm = 0.f;
    use(cands[0]);
    for (int k = 0; k < 200 && rc == NPX_OK && warm < 20.f; ++k) {
        rc = timed(1, &t);
        warm += t;
    }
    const int batch = std::max(1, std::min(50, static_cast<int>(2.0f / std::max(t, 0.01f)) + 1));
    std::vector<float> per_call(cands.size(), 0.f);
    for (int round = 0; round < 3 && rc == NPX_OK; ++round) {
        for (size_t i = 0; i < cands.size() && rc == NPX_OK; ++i) {
            use(cands[i]);
            rc = timed(batch, &t);
            const float pc = t / static_cast<float>(batch);
            if (round == 0 || pc < per_call[i]) per_call[i] = pc;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc != NPX_OK) return rc;
    size_t bi = 0;
    for (size_t i = 1; i < cands.size(); ++i)
        if (per_call[i] < per_call[bi]) bi = i;
    *best = cands[bi];
    *best_ms = per_call[bi];
    return NPX_OK;
}

// Shape of the per-frame-map kernel: fused_shape with its rows per lane (frames_q).
Shape frames_shape(int H, int W, int M, const LaunchCfg& c) {
    Shape s = fused_shape(H, W, M, c);
    const int qf = npx::frames_q(s.nwaves, s.ngroups);
    s.nregy = (H + qf - 1) / qf;
    return s;
}

int pow_mode_of(double S) {
    const float s = static_cast<float>(S);
    if (!(s > 1.0f)) return 0;
    if (s == 2.0f) return 2;
    if (s == 3.0f) return 3;
    if (s == 4.0f) return 4;
    return 1;
}

int check_common(const void* d_in, int dtype, int N, int H, int W, int M) {
    if (!d_in) return fail(NPX_ERR_ARG, "input pointer is NULL");
    if (dtype != NPX_DTYPE_F32 && dtype != NPX_DTYPE_U16) return fail(NPX_ERR_ARG, "dtype must be f32 or u16");
    if (N < 1 || H < 1 || W < 1) return fail(NPX_ERR_ARG, "stack must have N, H, W >= 1");
    if (M < 1 || M > 64) return fail(NPX_ERR_ARG, "magnification must be in [1, 64]");
    return NPX_OK;
}

// Per-job fields shared by every entry point: stack, tables of one (H, W, M), launch shape.
npx::FusedArgs job_args(const void* d_in, int dtype, int N, int H, int W, int M, const npx::PlanHost& h,
                        const npx::DevTables* t, const Shape& s, int iw) {
    npx::FusedArgs a{};
    a.in = d_in;
    a.u16 = dtype == NPX_DTYPE_U16;
    a.N = N;
    a.H = H;
    a.W = W;
    a.M = M;
    a.HM = h.HM;
    a.WM = h.WM;
    a.f0 = 0;
    a.nphase = h.nphase;
    a.nwaves = s.nwaves;
    a.ngroups = s.ngroups;
    a.chunk0 = 0;
    a.direct = 1;
    a.iw = iw ? 1 : 0;
    a.lrw = t->lrw;
    a.lrb = t->lrb;
    a.lcw = t->lcw;
    a.lcb = t->lcb;
    a.crw = t->crw;
    a.crb = t->crb;
    a.ccw = t->ccw;
    a.ccb = t->ccb;
    a.nrad = 0;
    a.ns = 0;
    return a;
}

// The kernels form each block's ring geometry themselves, from tSS, tSO and the block's pixel centres.
void add_radius(npx::FusedArgs& a, const npx::PlanHost& h) {
    const int r = a.nrad++;
    a.K[r] = h.K;
    a.margin[r] = h.margin;
    a.tss[r] = h.tSS;
    a.tso[r] = h.tSO;
    a.KL = std::max(a.KL, h.K);
}

void add_sensitivity(npx::FusedArgs& a, double S) {
    const int s = a.ns++;
    a.sens[s] = static_cast<float>(S);
    a.sens_pow[s] = pow_mode_of(S);
}

npx::FusedArgs base_args(const void* d_in, int dtype, int N, int H, int W, int M, const npx::PlanHost& h,
                         const npx::DevTables* t, const Shape& s, double S, int iw) {
    npx::FusedArgs a = job_args(d_in, dtype, N, H, W, M, h, t, s, iw);
    a.KL = h.K;
    add_radius(a, h);
    add_sensitivity(a, S);
    return a;
}

// Plan + device tables under the cache lock; NPX_OK or an error code with g_err set.
// Under stream capture the tables must already be on the device (their upload blocks).
int plan_and_tables(int H, int W, int M, double R, Plan** p, const npx::DevTables** t, hipStream_t st) {
    std::lock_guard<std::mutex> lk(g_mu);
    *p = get_plan(H, W, M, R);
    if (!*p) return NPX_ERR_UNSUPPORTED;
    int dev = 0;
    if (capturing(st) && (hipGetDevice(&dev) != hipSuccess || (*p)->dev.find(dev) == (*p)->dev.end()))
        return fail(NPX_ERR_ARG, "geometry tables are uploaded by the first call: make one call with these "
                                 "parameters outside stream capture first");
    return device_tables(*p, t);
}

bool capturing(hipStream_t st) {
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    return hipStreamIsCapturing(st, &cs) == hipSuccess && cs != hipStreamCaptureStatusNone;
}

// ------------------------------------------------------------------------------------------ parameter sweep

// One job of a sweep launch: magnification mi with radii [r0, r0 + nr) and sensitivities [s0, s0 + nsn).
struct SweepJob {
    int mi, r0, nr, s0, nsn;
    int M, nregx, nregy, ngroups, kmax;
    double cost;  // estimated time of one of its blocks (ring samples of its waves + the per-frame block work)
    size_t P;
    size_t ws_off;
};

struct SweepLayout {
    std::vector<SweepJob> jobs;  // longest blocks first
    int nchunk = 1, fpc = 1;
    size_t ws_bytes = 0;
};

// Per-frame work of one block besides its ring sums (tile, gradients, lattice, finalise), in ring-sample units of
// one wave slot: each slot's ring costs (2K + 1)^2 units, the rest of the block about 750.
constexpr double kBlockFrameCost = 750.0;

// Job list, frame chunking and workspace of a sweep; deterministic in its arguments (the workspace query and
// the sweep itself agree). chunks = false: one chunk, no workspace.
//
// A magnification whose phases fill whole blocks (M * M a multiple of the 16 wave slots) gets one job per radius,
// so every block runs one ring half-width; smaller ones pack up to kMaxRad radii into the slots of one block (they
// share its lattice). Jobs run longest blocks first, so the short blocks fill the end of the launch.
int sweep_layout(int N, int H, int W, int n_mag, const int* mags, int n_rad, const double* radii, int n_sens,
                 bool chunks, SweepLayout* L) {
    if (n_mag < 1 || n_rad < 1 || n_sens < 1 || !mags || !radii) return fail(NPX_ERR_ARG, "empty sweep");
    if (N < 1 || H < 1 || W < 1) return fail(NPX_ERR_ARG, "stack must have N, H, W >= 1");
    L->jobs.clear();
    const int nw = npx::kBatchThreads / 64;
    for (int mi = 0; mi < n_mag; ++mi) {
        const int M = mags[mi];
        if (M < 1 || M > 64) return fail(NPX_ERR_ARG, "magnification must be in [1, 64]");
        std::vector<int> K(n_rad);
        for (int r = 0; r < n_rad; ++r) {
            std::lock_guard<std::mutex> lk(g_mu);
            Plan* p = get_plan(H, W, M, radii[r]);
            if (!p) return NPX_ERR_UNSUPPORTED;
            K[r] = p->host.K;
        }
        const int per_job = (M * M) % nw == 0 ? 1 : npx::kMaxRad;
        for (int r0 = 0; r0 < n_rad; r0 += per_job) {
            const int nr = std::min(per_job, n_rad - r0);
            int kmax = 0;
            double ring = 0.0;
            for (int r = r0; r < r0 + nr; ++r) {
                kmax = std::max(kmax, K[r]);
                ring += static_cast<double>(M) * M * (2 * K[r] + 1) * (2 * K[r] + 1);
            }
            const int ngroups = (M * M * nr + nw - 1) / nw;
            for (int s0 = 0; s0 < n_sens; s0 += npx::kMaxSens) {
                SweepJob j{};
                j.mi = mi;
                j.r0 = r0;
                j.nr = nr;
                j.s0 = s0;
                j.nsn = std::min(npx::kMaxSens, n_sens - s0);
                j.M = M;
                j.nregx = (W + 63) / 64;
                j.nregy = (H + npx::kQTemporal - 1) / npx::kQTemporal;
                j.ngroups = ngroups;
                j.kmax = kmax;
                j.cost = kBlockFrameCost + ring / ngroups;
                j.P = static_cast<size_t>(H) * M * static_cast<size_t>(W) * M;
                L->jobs.push_back(j);
            }
        }
    }
    std::stable_sort(L->jobs.begin(), L->jobs.end(), [](const SweepJob& a, const SweepJob& b) {
        return a.cost > b.cost;
    });
    long long spatial = 0;
    for (const auto& j : L->jobs) spatial += static_cast<long long>(j.nregx) * j.nregy * j.ngroups;
    // frame chunks: the fewest that give about four blocks per compute unit (with the jobs ordered longest first,
    // every CU then runs about as many blocks of each kind; more chunks only add per-block work), each chunk >= 32
    // frames
    long long want = 1;
    if (chunks) {
        const long long blocks = 4LL * device_cus();
        want = (blocks + spatial - 1) / spatial;
        want = std::max<long long>(1, std::min<long long>(want, N / 32));
    }
    L->fpc = static_cast<int>((N + want - 1) / want);
    L->nchunk = (N + L->fpc - 1) / L->fpc;
    L->ws_bytes = 0;
    if (L->nchunk > 1) {
        for (auto& j : L->jobs) {
            j.ws_off = L->ws_bytes;
            L->ws_bytes += align256(static_cast<size_t>(L->nchunk) * j.nr * j.nsn * j.P * sizeof(float));
        }
    }
    return NPX_OK;
}

int host_plan(int H, int W, int M, double R, const PlanHost** h) {
    std::lock_guard<std::mutex> lk(g_mu);
    Plan* p = get_plan(H, W, M, R);
    if (!p) return NPX_ERR_UNSUPPORTED;
    *h = &p->host;
    return NPX_OK;
}

int tables_for(int H, int W, int M, double R, const PlanHost** h, const DevTables** t, hipStream_t st) {
    Plan* p = nullptr;
    const int rc = plan_and_tables(H, W, M, R, &p, t, st);
    if (rc) return rc;
    *h = &p->host;
    return NPX_OK;
}

// ------------------------------------------------------------------------------------------ static scratch

// Device memory of npx_esrrf_temporal_static: its geometry tables and frame-chunk partials, for callers that cannot
// allocate. Every GPU the library runs on holds one copy until the process ends; its address on a device comes from
// hipGetSymbolAddress with that device current (in host code the array's name is not a device address).
constexpr size_t kScratchBytes = size_t(64) << 20;
__device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_scratch_mu;
std::map<int, void*> g_scratch_ptr;  // device id -> address of `scratch` on that device

// Address of the current device's scratch, looked up on the first call on each device.
int scratch_on_device(void** out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    auto it = g_scratch_ptr.find(dev);
    if (it == g_scratch_ptr.end()) {
        void* ptr = nullptr;
        e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)");
        if (!ptr) return fail(NPX_ERR_HIP, "hipGetSymbolAddress(scratch) gave a null address");
        it = g_scratch_ptr.emplace(dev, ptr).first;
    }
    *out = it->second;
    return NPX_OK;
}

}  // namespace host
}  // namespace npx

using namespace npx::host;

extern "C" {

int npx_version(void) { return 100; }

const char* npx_last_error(void) { return g_err.c_str(); }

int npx_device_info(char* buf, int len) {
    if (!buf || len <= 0) return fail(NPX_ERR_ARG, "buffer is NULL");
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
    hipDeviceProp_t p;
    e = hipGetDeviceProperties(&p, dev);
    if (e != hipSuccess) return hip_fail(e, "hipGetDeviceProperties");
    snprintf(buf, len, "%s;%s;%d;%d", p.name, p.gcnArchName, p.multiProcessorCount, p.clockRate);
    return NPX_OK;
}

int npx_esrrf_frames(const void* d_in, int dtype, int N, int H, int W, int M, double R, double S, int iw,
                     float* d_out, void* stream) {
    int rc = check_common(d_in, dtype, N, H, W, M);
    if (rc) return rc;
    if (!d_out) return fail(NPX_ERR_ARG, "output pointer is NULL");
    const npx::DevTables* t = nullptr;
    Plan* p = nullptr;
    rc = plan_and_tables(H, W, M, R, &p, &t, static_cast<hipStream_t>(stream));
    if (rc) return rc;
    const LaunchCfg c = launch_cfg_frames(H, W, M, R);
    const Shape s = frames_shape(H, W, M, c);
    npx::FusedArgs a = base_args(d_in, dtype, N, H, W, M, p->host, t, s, S, iw);
    // frames split into chunks, each block streaming one chunk
    frames_chunks(N, s, c.target_blocks, &a.nchunk, &a.fpc);
    a.out = d_out;
    dim3 grid(s.nregx, s.nregy, s.ngroups * a.nchunk);
    dim3 block(64 * s.nwaves);
    const hipStream_t st = static_cast<hipStream_t>(stream);
    CallEvents ev{};
    const bool timed = timing_begin(st, &ev);
    hipError_t e = npx::launch_fused(a, grid, block, npx::kModeFrames, st);
    if (e != hipSuccess) return hip_fail(e, "esrrf_fused(frames) launch");
    if (timed && hipEventRecord(ev.e1, st) == hipSuccess && hipEventRecord(ev.e2, st) == hipSuccess) timing_end(ev);
    return NPX_OK;
}

size_t npx_esrrf_temporal_workspace(int N, int H, int W, int M, double R) {
    if (N < 1 || H < 1 || W < 1 || M < 1) return 0;
    const LaunchCfg c = launch_cfg(H, W, M, R);
    const Shape s = fused_shape(H, W, M, c);
    int nchunk, fpc;
    temporal_chunks(N, s, c.target_blocks, &nchunk, &fpc);
    if (nchunk <= 1) return 0;
    return static_cast<size_t>(nchunk) * 5 * static_cast<size_t>(H) * M * static_cast<size_t>(W) * M * sizeof(float);
}

int npx_esrrf_temporal(const void* d_in, int dtype, int N, int H, int W, int M, double R, double S, int iw,
                       unsigned mode_mask, float* d_avg, float* d_var, float* d_tac2, void* d_work, size_t work_bytes,
                       void* stream) {
    int rc = check_common(d_in, dtype, N, H, W, M);
    if (rc) return rc;
    if ((mode_mask & (NPX_AVG | NPX_VAR | NPX_TAC2)) == 0) return fail(NPX_ERR_ARG, "empty temporal mode mask");
    if (((mode_mask & NPX_AVG) && !d_avg) || ((mode_mask & NPX_VAR) && !d_var) || ((mode_mask & NPX_TAC2) && !d_tac2))
        return fail(NPX_ERR_ARG, "output pointer of a requested temporal mode is NULL");
    const npx::DevTables* t = nullptr;
    Plan* p = nullptr;
    rc = plan_and_tables(H, W, M, R, &p, &t, static_cast<hipStream_t>(stream));
    if (rc) return rc;
    const LaunchCfg c = launch_cfg(H, W, M, R);
    const Shape s = fused_shape(H, W, M, c);
    npx::FusedArgs a = base_args(d_in, dtype, N, H, W, M, p->host, t, s, S, iw);
    int nchunk = 1, fpc = N;
    if (d_work) {
        temporal_chunks(N, s, c.target_blocks, &nchunk, &fpc);
        const size_t need = nchunk > 1 ? static_cast<size_t>(nchunk) * 5 * static_cast<size_t>(p->host.HM) *
                                             p->host.WM * sizeof(float)
                                       : 0;
        if (work_bytes < need) return fail(NPX_ERR_WORKSPACE, "workspace smaller than npx_esrrf_temporal_workspace()");
    }
    a.fpc = fpc;
    a.nchunk = nchunk;
    a.direct = nchunk == 1;
    a.avg = (mode_mask & NPX_AVG) ? d_avg : nullptr;
    a.var = (mode_mask & NPX_VAR) ? d_var : nullptr;
    a.tac2 = (mode_mask & NPX_TAC2) ? d_tac2 : nullptr;
    a.part = nchunk > 1 ? static_cast<float*>(d_work) : nullptr;
    const hipStream_t st = static_cast<hipStream_t>(stream);
    dim3 grid(s.nregx, s.nregy, s.ngroups * nchunk);
    dim3 block(64 * s.nwaves);
    CallEvents ev{};
    const bool timed = timing_begin(st, &ev);
    const int mode = a.var ? npx::kModeTemporalVar : npx::kModeTemporal;
    hipError_t e = npx::launch_fused(a, grid, block, mode, st);
    if (e != hipSuccess) return hip_fail(e, "esrrf_fused(temporal) launch");
    if (timed) (void)hipEventRecord(ev.e1, st);
    if (nchunk > 1) {
        e = npx::launch_merge(a.part, nchunk, fpc, N, static_cast<size_t>(p->host.HM) * p->host.WM, a.avg, a.var,
                              a.tac2, st);
        if (e != hipSuccess) return hip_fail(e, "moments_merge launch");
    }
    if (timed && hipEventRecord(ev.e2, st) == hipSuccess) timing_end(ev);
    return NPX_OK;
}

int npx_esrrf_temporal_static(const void* d_in, int dtype, int N, int H, int W, int M, double R, double S, int iw,
                              unsigned mode_mask, float* d_avg, float* d_var, float* d_tac2, void* stream) {
    int rc = check_common(d_in, dtype, N, H, W, M);
    if (rc) return rc;
    if ((mode_mask & (NPX_AVG | NPX_VAR | NPX_TAC2)) == 0) return fail(NPX_ERR_ARG, "empty temporal mode mask");
    if (((mode_mask & NPX_AVG) && !d_avg) || ((mode_mask & NPX_VAR) && !d_var) || ((mode_mask & NPX_TAC2) && !d_tac2))
        return fail(NPX_ERR_ARG, "output pointer of a requested temporal mode is NULL");
    // host plan and its packed tables (host memory only)
    const TableBlob* blob = nullptr;
    const npx::PlanHost* h = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        Plan* p = get_plan(H, W, M, R);
        if (!p) return NPX_ERR_UNSUPPORTED;
        if (p->blob.bytes.empty()) pack_tables(p->host, &p->blob);
        blob = &p->blob;
        h = &p->host;
    }
    const size_t tab = align256(blob->bytes.size());
    if (tab > kScratchBytes)
        return fail(NPX_ERR_ARG, "geometry tables of this image size (" + std::to_string(tab) +
                                     " bytes) exceed the static scratch of npx_esrrf_temporal_static (" +
                                     std::to_string(kScratchBytes) + " bytes)");
    // the frame chunks npx_esrrf_temporal would run, as many as the scratch holds partials for (each 5 HM WM floats)
    const LaunchCfg c = launch_cfg(H, W, M, R);
    const Shape sh = fused_shape(H, W, M, c);
    int nchunk = 1, fpc = N;
    temporal_chunks(N, sh, c.target_blocks, &nchunk, &fpc);
    const size_t P = static_cast<size_t>(h->HM) * h->WM;
    const size_t slot = 5 * P * sizeof(float);
    const long long fit = static_cast<long long>((kScratchBytes - tab) / slot);
    if (nchunk > 1 && nchunk > fit) {
        if (fit < 2) {
            nchunk = 1;
            fpc = N;
        } else {
            fpc = static_cast<int>((N + fit - 1) / fit);
            nchunk = (N + fpc - 1) / fpc;
        }
    }
    void* base = nullptr;
    rc = scratch_on_device(&base);
    if (rc) return rc;
    const hipStream_t st = static_cast<hipStream_t>(stream);
    // tables into the scratch on the call's stream (stream-ordered before the kernels that read them)
    hipError_t e = hipMemcpyAsync(base, blob->bytes.data(), blob->bytes.size(), hipMemcpyHostToDevice, st);
    if (e != hipSuccess) return hip_fail(e, "tables into the static scratch");
    const npx::DevTables t = blob->tables_in(base);
    npx::FusedArgs a = base_args(d_in, dtype, N, H, W, M, *h, &t, sh, S, iw);
    a.fpc = fpc;
    a.nchunk = nchunk;
    a.direct = nchunk == 1;
    a.avg = (mode_mask & NPX_AVG) ? d_avg : nullptr;
    a.var = (mode_mask & NPX_VAR) ? d_var : nullptr;
    a.tac2 = (mode_mask & NPX_TAC2) ? d_tac2 : nullptr;
    a.part = nchunk > 1 ? reinterpret_cast<float*>(static_cast<uint8_t*>(base) + tab) : nullptr;
    const int mode = a.var ? npx::kModeTemporalVar : npx::kModeTemporal;
    e = npx::launch_fused(a, dim3(sh.nregx, sh.nregy, sh.ngroups * nchunk), dim3(64 * sh.nwaves), mode, st);
    if (e != hipSuccess) return hip_fail(e, "esrrf_fused(temporal, static scratch) launch");
    if (nchunk > 1) {
        e = npx::launch_merge(a.part, nchunk, fpc, N, P, a.avg, a.var, a.tac2, st);
        if (e != hipSuccess) return hip_fail(e, "moments_merge launch");
    }
    return NPX_OK;
}

size_t npx_esrrf_sweep_workspace(int N, int H, int W, int n_mag, const int* mags, int n_rad, const double* radii,
                                 int n_sens) {
    SweepLayout L;
    if (sweep_layout(N, H, W, n_mag, mags, n_rad, radii, n_sens, true, &L) != NPX_OK) return 0;
    return L.ws_bytes;
}

int npx_esrrf_sweep(const void* d_in, int dtype, int N, int H, int W, int n_mag, const int* mags, int n_rad,
                    const double* radii, int n_sens, const double* sens, int iw, float* const* d_avg, void* d_work,
                    size_t work_bytes, void* stream) {
    if (!sens || !d_avg) return fail(NPX_ERR_ARG, "NULL sensitivity or output array");
    for (int mi = 0; mi < n_mag && mags; ++mi) {
        const int rc = check_common(d_in, dtype, N, H, W, mags[mi]);
        if (rc) return rc;
    }
    SweepLayout L;
    int rc = sweep_layout(N, H, W, n_mag, mags, n_rad, radii, n_sens, d_work != nullptr, &L);
    if (rc) return rc;
    if (L.ws_bytes > 0 && work_bytes < L.ws_bytes)
        return fail(NPX_ERR_WORKSPACE, "workspace smaller than npx_esrrf_sweep_workspace()");
    for (int k = 0; k < n_mag * n_rad * n_sens; ++k)
        if (!d_avg[k]) return fail(NPX_ERR_ARG, "NULL output map in the sweep output array");
    const hipStream_t st = static_cast<hipStream_t>(stream);
    const int nw = npx::kBatchThreads / 64;
    CallEvents ev{};
    const bool timed = timing_begin(st, &ev);
    std::vector<std::vector<float*>> merge_outs;
    for (size_t j0 = 0; j0 < L.jobs.size(); j0 += npx::kMaxJobs) {
        const size_t nj = std::min<size_t>(npx::kMaxJobs, L.jobs.size() - j0);
        npx::BatchArgs b{};
        b.njobs = static_cast<int>(nj);
        int blocks = 0;
        for (size_t jj = 0; jj < nj; ++jj) {
            const SweepJob& sj = L.jobs[j0 + jj];
            Shape sh{nw, sj.ngroups, sj.nregx, sj.nregy};
            npx::FusedArgs a{};
            for (int r = 0; r < sj.nr; ++r) {
                Plan* p = nullptr;
                const npx::DevTables* t = nullptr;
                rc = plan_and_tables(H, W, sj.M, radii[sj.r0 + r], &p, &t, st);
                if (rc) return rc;
                if (r == 0) a = job_args(d_in, dtype, N, H, W, sj.M, p->host, t, sh, iw);
                add_radius(a, p->host);
            }
            for (int s = 0; s < sj.nsn; ++s) add_sensitivity(a, sens[sj.s0 + s]);
            // one half-width: the body built for it; radii of different half-widths share a wider lattice
            a.vark = 0;
            for (int r = 1; r < sj.nr; ++r) a.vark |= a.K[r] != a.K[0];
            a.KL = a.vark ? npx::batched_lattice_k(sj.kmax) : a.K[0];
            a.fpc = L.fpc;
            a.nchunk = L.nchunk;
            a.direct = L.nchunk == 1;
            a.part = L.nchunk > 1 ? reinterpret_cast<float*>(static_cast<uint8_t*>(d_work) + sj.ws_off) : nullptr;
            std::vector<float*> outs;
            for (int r = 0; r < sj.nr; ++r)
                for (int s = 0; s < sj.nsn; ++s) {
                    float* o = d_avg[(static_cast<size_t>(sj.mi) * n_rad + sj.r0 + r) * n_sens + sj.s0 + s];
                    a.sweep_avg[r * npx::kMaxSens + s] = o;
                    outs.push_back(o);
                }
            merge_outs.push_back(outs);
            b.job[jj] = a;
            b.nregx[jj] = sj.nregx;
            b.nregy[jj] = sj.nregy;
            b.start[jj] = blocks;
            blocks += sj.nregx * sj.nregy * sj.ngroups * L.nchunk;
        }
        b.start[nj] = blocks;
        hipError_t e = npx::launch_batched(b, npx::kModeSweep, st);
        if (e != hipSuccess) return hip_fail(e, "esrrf_batched(sweep) launch");
    }
    if (timed) (void)hipEventRecord(ev.e1, st);
    if (L.nchunk > 1) {
        for (size_t j = 0; j < L.jobs.size(); ++j) {
            const SweepJob& sj = L.jobs[j];
            const float* part = reinterpret_cast<const float*>(static_cast<uint8_t*>(d_work) + sj.ws_off);
            hipError_t e = npx::launch_means_merge(part, L.nchunk, L.fpc, N, sj.P, sj.nr * sj.nsn,
                                                   merge_outs[j].data(), st);
            if (e != hipSuccess) return hip_fail(e, "means_merge launch");
        }
    }
    if (timed && hipEventRecord(ev.e2, st) == hipSuccess) timing_end(ev);
    return NPX_OK;
}

// ------------------------------------------------------------------------------------------ graphs

int npx_capture_begin(void** capture_stream) {
    if (!capture_stream) return fail(NPX_ERR_ARG, "NULL stream output");
    hipStream_t s = nullptr;
    hipError_t e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
    if (e != hipSuccess) return hip_fail(e, "hipStreamCreateWithFlags");
    e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) {
        (void)hipStreamDestroy(s);
        return hip_fail(e, "hipStreamBeginCapture");
    }
    *capture_stream = s;
    return NPX_OK;
}

int npx_capture_end(void* capture_stream, void** graph) {
    if (!capture_stream || !graph) return fail(NPX_ERR_ARG, "NULL stream or graph output");
    hipStream_t s = static_cast<hipStream_t>(capture_stream);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(s, &g);
    (void)hipStreamDestroy(s);
    if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
    hipGraphExec_t x = nullptr;
    e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) return hip_fail(e, "hipGraphInstantiate");
    *graph = x;
    return NPX_OK;
}

int npx_graph_launch(void* graph, void* stream) {
    if (!graph) return fail(NPX_ERR_ARG, "NULL graph");
    hipError_t e = hipGraphLaunch(static_cast<hipGraphExec_t>(graph), static_cast<hipStream_t>(stream));
    if (e != hipSuccess) return hip_fail(e, "hipGraphLaunch");
    return NPX_OK;
}

int npx_graph_destroy(void* graph) {
    if (!graph) return NPX_OK;
    hipError_t e = hipGraphExecDestroy(static_cast<hipGraphExec_t>(graph));
    if (e != hipSuccess) return hip_fail(e, "hipGraphExecDestroy");
    return NPX_OK;
}

// npx_rgc and npx_rgc_lattice: `compact` gradients hold only the lattice the RGC reads, (N, 2H + 1, 2W + 1), and
// run only on the fused kernel (NPX_ERR_UNSUPPORTED where the per-pixel kernel would be needed).
static int rgc_impl(const float* d_gx, const float* d_gy, const float* d_img, int N, int HM, int WM, int M, double R,
                    double S, int iw, int pad_edges, float* d_out, void* stream, bool compact) {
    if (!d_gx || !d_gy || !d_img || !d_out) return fail(NPX_ERR_ARG, "NULL array pointer");
    if (N < 1 || HM < 1 || WM < 1 || M < 1) return fail(NPX_ERR_ARG, "bad shape");
    const float sigma = static_cast<float>(R / 2.355);
    const float fwhm = static_cast<float>(R);
    const float tSS = static_cast<float>(2.0f * sigma) * sigma;
    const float tSO = static_cast<float>(2.0f * sigma) + 1.0f;
    const int margin = pad_edges ? 0 : static_cast<int>(static_cast<double>(fwhm) * M);
    const hipStream_t st = static_cast<hipStream_t>(stream);
    // An image of H x W input pixels magnified M times: the fused kernel runs on the gradient lattice (the caller's
    // arrays at every M-th position) with the plan's ring tables, as the eSRRF path does on its own lattice.
    // Other shapes, and radii or sizes the plan does not cover, take the per-pixel kernel below.
    if (compact && M < 2) return fail(NPX_ERR_UNSUPPORTED, "compact lattice: needs M >= 2");
    if (HM % M == 0 && WM % M == 0 && M <= 64) {
        const int H = HM / M, W = WM / M;
        const npx::DevTables* t = nullptr;
        Plan* p = nullptr;
        const std::string prev = g_err;
        const int rc = plan_and_tables(H, W, M, R, &p, &t, st);
        if (rc == NPX_OK) {
            const LaunchCfg c = launch_cfg_frames(H, W, M, R);
            const Shape s = frames_shape(H, W, M, c);
            npx::FusedArgs a = base_args(d_img, NPX_DTYPE_F32, N, H, W, M, p->host, t, s, S, iw);
            a.agx = d_gx;
            a.agy = d_gy;
            a.aimg = d_img;
            a.alat = compact ? 1 : 0;
            a.margin[0] = margin;
            const long long spatial = static_cast<long long>(s.nregx) * s.nregy * s.ngroups;
            long long want = std::max<long long>(1, (c.target_blocks + spatial - 1) / spatial);
            // at least ~4 frames per block, so a block's start-up and its per-frame lattice gathers are spread over
            // ring work (20 frames of 256 x 256 at M5: 2 -> 4 frames per block, 0.05 ms off a 2.1 ms use_fht call)
            want = std::min<long long>({want, std::max(1, N / 4), 65535 / s.ngroups});
            a.fpc = static_cast<int>((N + want - 1) / want);
            a.nchunk = (N + a.fpc - 1) / a.fpc;
            a.out = d_out;
            hipError_t e = npx::launch_fused(a, dim3(s.nregx, s.nregy, s.ngroups * a.nchunk), dim3(64 * s.nwaves),
                                             npx::kModeArrays, st);
            if (e != hipSuccess) return hip_fail(e, "esrrf_fused(arrays) launch");
            return NPX_OK;
        }
        if (compact) return rc;  // g_err holds the plan's reason
        g_err = prev;
    }
    if (compact) return fail(NPX_ERR_UNSUPPORTED, "compact lattice: needs H x W pixels magnified M <= 64 times");
    hipError_t e = npx::launch_rgc_arrays(d_gx, d_gy, d_img, N, HM, WM, M, fwhm, tSO, tSS, pow_mode_of(S),
                                          static_cast<float>(S), iw ? 1 : 0, margin, d_out, st);
    if (e != hipSuccess) return hip_fail(e, "rgc_arrays launch");
    return NPX_OK;
}

int npx_rgc(const float* d_gx, const float* d_gy, const float* d_img, int N, int HM, int WM, int M, double R,
            double S, int iw, int pad_edges, float* d_out, void* stream) {
    return rgc_impl(d_gx, d_gy, d_img, N, HM, WM, M, R, S, iw, pad_edges, d_out, stream, false);
}

int npx_rgc_lattice(const float* d_gx, const float* d_gy, const float* d_img, int N, int HM, int WM, int M,
                    double R, double S, int iw, int pad_edges, float* d_out, void* stream) {
    return rgc_impl(d_gx, d_gy, d_img, N, HM, WM, M, R, S, iw, pad_edges, d_out, stream, true);
}

int npx_temporal(const float* d_stack, int N, long long P, unsigned mode_mask, float* d_avg, float* d_var,
                 float* d_tac2, void* stream) {
    if (!d_stack || N < 1 || P < 1) return fail(NPX_ERR_ARG, "bad stack");
    if ((mode_mask & (NPX_AVG | NPX_VAR | NPX_TAC2)) == 0) return fail(NPX_ERR_ARG, "empty temporal mode mask");
    hipError_t e = npx::launch_moments_stack(d_stack, N, static_cast<size_t>(P), (mode_mask & NPX_AVG) ? d_avg : nullptr,
                                             (mode_mask & NPX_VAR) ? d_var : nullptr,
                                             (mode_mask & NPX_TAC2) ? d_tac2 : nullptr,
                                             static_cast<hipStream_t>(stream));
    if (e != hipSuccess) return hip_fail(e, "moments_stack launch");
    return NPX_OK;
}

int npx_macro_pixel(float* d_img, int S, int HM, int WM, int M, float* d_cmap, void* stream) {
    if (!d_img || !d_cmap || S < 1 || HM < 1 || WM < 1 || M < 1) return fail(NPX_ERR_ARG, "bad arguments");
    if (HM / M < 1 || WM / M < 1) return fail(NPX_ERR_ARG, "image smaller than one macro-pixel");
    hipError_t e = npx::launch_macro_pixel(d_img, S, HM, WM, M, d_cmap, static_cast<hipStream_t>(stream));
    if (e != hipSuccess) return hip_fail(e, "macro_pixel launch");
    return NPX_OK;
}

int npx_plan_info(int H, int W, int M, double R, int* info, float* finfo) {
    if (!info || !finfo) return fail(NPX_ERR_ARG, "NULL output");
    std::lock_guard<std::mutex> lk(g_mu);
    Plan* p = get_plan(H, W, M, R);
    if (!p) return NPX_ERR_UNSUPPORTED;
    const npx::PlanHost& h = p->host;
    const Shape s = fused_shape(H, W, M, launch_cfg(H, W, M, R));
    const Shape sf = fused_shape(H, W, M, launch_cfg_frames(H, W, M, R));
    const int K = h.K;
    const int HG = K / 2 + 3, HI = HG + 1;
    const int v[16] = {K, h.margin, h.HM, h.WM, h.nphase, npx::frames_q(sf.nwaves, sf.ngroups), npx::kQTemporal,
                       h.ring_stride, HG, HI,
                       64 + K, s.nwaves, s.ngroups, npx::kPrefetch, 0, 0};
    for (int i = 0; i < 16; ++i) info[i] = v[i];
    finfo[0] = h.fwhm;
    finfo[1] = h.sigma;
    finfo[2] = h.tSS;
    finfo[3] = h.tSO;
    return NPX_OK;
}

int npx_plan_tables(int H, int W, int M, double R, float* ring, float* wsum, int* lrb, float* lrw, int* lcb, float* lcw,
                    int* crb, float* crw, int* ccb, float* ccw) {
    std::lock_guard<std::mutex> lk(g_mu);
    Plan* p = get_plan(H, W, M, R);
    if (!p) return NPX_ERR_UNSUPPORTED;
    const npx::PlanHost& h = p->host;
    if (ring) std::copy(h.ring.begin(), h.ring.end(), ring);
    if (wsum) std::copy(h.wsum.begin(), h.wsum.end(), wsum);
    auto put = [](const std::vector<npx::CrTap>& v, int* b, float* w) {
        for (size_t i = 0; i < v.size(); ++i) {
            if (b) b[i] = v[i].base;
            if (w)
                for (int u = 0; u < 4; ++u) w[4 * i + u] = v[i].w[u];
        }
    };
    put(h.lat_row, lrb, lrw);
    put(h.lat_col, lcb, lcw);
    put(h.cr_row, crb, crw);
    put(h.cr_col, ccb, ccw);
    return NPX_OK;
}

int npx_set_launch(int max_waves_per_block, int target_blocks) {
    std::lock_guard<std::mutex> lk(g_cfg_mu);
    if (max_waves_per_block < 0 || max_waves_per_block > 16 || target_blocks < 0)
        return fail(NPX_ERR_ARG, "max_waves_per_block must be in [1, 16] (0 = default), target_blocks >= 0");
    g_default_cfg.max_waves = max_waves_per_block ? max_waves_per_block : 16;
    g_default_cfg.target_blocks = target_blocks ? target_blocks : 2048;
    return NPX_OK;
}

int npx_tune_set(int H, int W, int M, double R, int max_waves_per_block, int target_blocks) {
    if (max_waves_per_block < 0 || max_waves_per_block > 16 || target_blocks < 0)
        return fail(NPX_ERR_ARG, "max_waves_per_block must be in [1, 16], target_blocks >= 1 (0, 0 = forget)");
    std::lock_guard<std::mutex> lk(g_cfg_mu);
    const Key k{H, W, M, bits(R)};
    if (max_waves_per_block == 0 || target_blocks == 0) g_tuned.erase(k);
    else g_tuned[k] = LaunchCfg{max_waves_per_block, target_blocks};
    return NPX_OK;
}

int npx_tune_get(int H, int W, int M, double R, int* cfg) {
    if (!cfg) return fail(NPX_ERR_ARG, "NULL output");
    std::lock_guard<std::mutex> lk(g_cfg_mu);
    auto it = g_tuned.find(Key{H, W, M, bits(R)});
    const LaunchCfg c = it != g_tuned.end() ? it->second : g_default_cfg;
    cfg[0] = c.max_waves;
    cfg[1] = c.target_blocks;
    return it != g_tuned.end() ? 1 : 0;
}

int npx_tune_frames_set(int H, int W, int M, double R, int max_waves_per_block, int target_blocks) {
    if (max_waves_per_block < 0 || max_waves_per_block > 16 || target_blocks < 0)
        return fail(NPX_ERR_ARG, "max_waves_per_block must be in [1, 16], target_blocks >= 1 (0, 0 = forget)");
    std::lock_guard<std::mutex> lk(g_cfg_mu);
    const Key k{H, W, M, bits(R)};
    if (max_waves_per_block == 0 || target_blocks == 0) g_tuned_frames.erase(k);
    else g_tuned_frames[k] = LaunchCfg{max_waves_per_block, target_blocks};
    return NPX_OK;
}

int npx_tune_frames_get(int H, int W, int M, double R, int* cfg) {
    if (!cfg) return fail(NPX_ERR_ARG, "NULL output");
    bool own = false;
    {
        std::lock_guard<std::mutex> lk(g_cfg_mu);
        own = g_tuned_frames.count(Key{H, W, M, bits(R)}) != 0;
    }
    const LaunchCfg c = launch_cfg_frames(H, W, M, R);
    cfg[0] = c.max_waves;
    cfg[1] = c.target_blocks;
    return own ? 1 : 0;
}

// Candidate launch shapes of the autotuner.
static const int kTuneWaves[] = {16, 8, 4};
static const int kTuneBlocks[] = {256, 512, 1024, 2048, 4096, 8192};

size_t npx_autotune_workspace(int N, int H, int W, int M, double R) {
    (void)R;
    if (N < 1 || H < 1 || W < 1 || M < 1) return 0;
    size_t most = 0;
    for (int mw : kTuneWaves)
        for (int tb : kTuneBlocks) {
            const Shape s = fused_shape(H, W, M, LaunchCfg{mw, tb});
            int nchunk, fpc;
            temporal_chunks(N, s, tb, &nchunk, &fpc);
            if (nchunk > 1)
                most = std::max(most, static_cast<size_t>(nchunk) * 5 * static_cast<size_t>(H) * M *
                                          static_cast<size_t>(W) * M * sizeof(float));
        }
    return most;
}

int npx_autotune(const void* d_in, int dtype, int N, int H, int W, int M, double R, double S, int iw,
                 unsigned mode_mask, float* d_avg, float* d_var, float* d_tac2, void* d_work, size_t work_bytes,
                 void* stream, double* best_ms) {
    const hipStream_t st = static_cast<hipStream_t>(stream);
    if (capturing(st)) return fail(NPX_ERR_ARG, "npx_autotune cannot run under stream capture");
    const Key key{H, W, M, bits(R)};
    std::vector<LaunchCfg> cands;
    std::vector<std::pair<int, int>> seen;  // (waves per block, chunks): candidates that launch alike are timed once
    for (int mw : kTuneWaves) {
        for (int tb : kTuneBlocks) {
            const LaunchCfg c{mw, tb};
            const Shape s = fused_shape(H, W, M, c);
            int nchunk = 1, fpc = N;
            if (d_work) temporal_chunks(N, s, tb, &nchunk, &fpc);
            const size_t need = nchunk > 1 ? static_cast<size_t>(nchunk) * 5 * static_cast<size_t>(H) * M *
                                                 static_cast<size_t>(W) * M * sizeof(float)
                                           : 0;
            if (need > work_bytes && nchunk > 1) continue;
            const std::pair<int, int> id{s.nwaves, nchunk};
            if (std::find(seen.begin(), seen.end(), id) != seen.end()) continue;
            seen.push_back(id);
            cands.push_back(c);
        }
    }
    if (cands.empty()) return fail(NPX_ERR_ARG, "no launch shape fits the workspace");
    LaunchCfg best{0, 0};
    float best_t = 0.f;
    const int rc = time_candidates(
        st, cands,
        [&](const LaunchCfg& c) {
            std::lock_guard<std::mutex> lk(g_cfg_mu);
            g_tuned[key] = c;
        },
        [&]() {
            return npx_esrrf_temporal(d_in, dtype, N, H, W, M, R, S, iw, mode_mask, d_avg, d_var, d_tac2, d_work,
                                      work_bytes, stream);
        },
        &best, &best_t);
    std::lock_guard<std::mutex> lk(g_cfg_mu);
    if (rc != NPX_OK) {
        g_tuned.erase(key);
        return rc;
    }
    g_tuned[key] = best;
    if (best_ms) *best_ms = best_t;
    return NPX_OK;
}

int npx_autotune_frames(const void* d_in, int dtype, int N, int H, int W, int M, double R, double S, int iw,
                        float* d_out, void* stream, double* best_ms) {
    const int rc0 = check_common(d_in, dtype, N, H, W, M);
    if (rc0) return rc0;
    if (!d_out) return fail(NPX_ERR_ARG, "output pointer is NULL");
    const hipStream_t st = static_cast<hipStream_t>(stream);
    if (capturing(st)) return fail(NPX_ERR_ARG, "npx_autotune_frames cannot run under stream capture");
    const Key key{H, W, M, bits(R)};
    std::vector<LaunchCfg> cands;
    std::vector<std::pair<int, int>> seen;  // (waves per block, chunks)
    for (int mw : kTuneWaves) {
        for (int tb : kTuneBlocks) {
            const LaunchCfg c{mw, tb};
            const Shape s = frames_shape(H, W, M, c);
            int nchunk = 1, fpc = N;
            frames_chunks(N, s, tb, &nchunk, &fpc);
            const std::pair<int, int> id{s.nwaves, nchunk};
            if (std::find(seen.begin(), seen.end(), id) != seen.end()) continue;
            seen.push_back(id);
            cands.push_back(c);
        }
    }
    LaunchCfg best{0, 0};
    float best_t = 0.f;
    const int rc = time_candidates(
        st, cands,
        [&](const LaunchCfg& c) {
            std::lock_guard<std::mutex> lk(g_cfg_mu);
            g_tuned_frames[key] = c;
        },
        [&]() { return npx_esrrf_frames(d_in, dtype, N, H, W, M, R, S, iw, d_out, stream); }, &best, &best_t);
    std::lock_guard<std::mutex> lk(g_cfg_mu);
    if (rc != NPX_OK) {
        g_tuned_frames.erase(key);
        return rc;
    }
    g_tuned_frames[key] = best;
    if (best_ms) *best_ms = best_t;
    return NPX_OK;
}

int npx_timing(int on) {
    std::lock_guard<std::mutex> lk(g_tmu);
    g_timing = on != 0;
    return NPX_OK;
}

int npx_timing_read(int* n_calls, double* kernel_ms, double* merge_ms) {
    std::vector<CallEvents> evs;
    {
        std::lock_guard<std::mutex> lk(g_tmu);
        evs.swap(g_events);
    }
    double k = 0, m = 0;
    int n = 0;
    int rc = NPX_OK;
    for (auto& ev : evs) {
        float a = 0, b = 0;
        hipError_t e = hipEventSynchronize(ev.e2);
        if (e == hipSuccess) e = hipEventElapsedTime(&a, ev.e0, ev.e1);
        if (e == hipSuccess) e = hipEventElapsedTime(&b, ev.e1, ev.e2);
        if (e != hipSuccess) rc = hip_fail(e, "timing events");
        else {
            k += a;
            m += b;
            ++n;
        }
        (void)hipEventDestroy(ev.e0);
        (void)hipEventDestroy(ev.e1);
        (void)hipEventDestroy(ev.e2);
    }
    if (n_calls) *n_calls = n;
    if (kernel_ms) *kernel_ms = k;
    if (merge_ms) *merge_ms = m;
    return rc;
}

void npx_release(void) {
    std::lock_guard<std::mutex> lk(g_mu);
    int cur = 0;
    if (hipGetDevice(&cur) != hipSuccess) cur = -1;
    for (auto& kv : g_plans) {
        for (auto& d : kv.second->dev) {
            if (hipSetDevice(d.first) == hipSuccess) (void)hipFree(d.second.mem);
        }
        kv.second->dev.clear();
    }
    if (cur >= 0) (void)hipSetDevice(cur);
    npx::host::release_srrf();
    npx::host::release_esrrf3d();
    npx::host::release_fht();
    npx::host::release_interp();
}

}  // extern "C"
