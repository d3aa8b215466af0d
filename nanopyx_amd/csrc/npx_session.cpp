// Serving session of libnpx_hip.so: continuous batching of eSRRF + temporal-correlation jobs on one GPU.
//
// A session owns one device pool. A submitted job waits in a FIFO queue until its resident footprint (its stack
// when it was submitted from host memory, its geometry tables, and its running temporal moments) fits in the
// pool and fewer than max_active jobs are admitted; admission copies stack and tables on the session's copy
// stream. Every step packs one frame slice of every admitted job whose copy has landed into shared batched
// launches (npx_batched.hip, up to kMaxJobs jobs each) on the caller's stream. A slice folds its frames into the
// job's running moments (mean, M2, lag-1 co-moment, first and previous value, kept in the pool between slices);
// the last slice writes the job's AVG / VAR / TAC2 maps. Slice lengths only change where the state is stored
// between launches, not the arithmetic: a job's maps are the same alone and in any mix.
//
// Reference: there is no serving layer in NanoPyx; each job is the computation of
// src/nanopyx/methods/esrrf/eSRRF_workflow.py:98-119 (eSRRF_ST.run -> calculate_eSRRF_temporal_correlations,
// src/nanopyx/core/transform/sr_temporal_correlations.py:46-75,205-221) for one timepoint.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <vector>

#include "../../include/npx_esrrf.h"
#include "npx_host.h"
#include "npx_kernels.h"

namespace npx {
namespace serve {

using host::fail;
using host::hip_fail;

constexpr size_t kAlign = 256;

size_t align_up(size_t x) { return (x + kAlign - 1) / kAlign * kAlign; }

// First-fit allocator over one device allocation; adjacent free ranges are merged on release.
class Pool {
   public:
    void init(uint8_t* base, size_t size) {
        base_ = base;
        size_ = size;
        free_.clear();
        free_[0] = size;
    }
    bool alloc(size_t bytes, size_t* off) {
        bytes = align_up(std::max<size_t>(bytes, 1));
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second >= bytes) {
                *off = it->first;
                const size_t rest = it->second - bytes;
                const size_t at = it->first + bytes;
                free_.erase(it);
                if (rest) free_[at] = rest;
                return true;
            }
        }
        return false;
    }
    void release(size_t off, size_t bytes) {
        bytes = align_up(std::max<size_t>(bytes, 1));
        auto it = free_.emplace(off, bytes).first;
        auto next = std::next(it);
        if (next != free_.end() && it->first + it->second == next->first) {
            it->second += next->second;
            free_.erase(next);
        }
        if (it != free_.begin()) {
            auto prev = std::prev(it);
            if (prev->first + prev->second == it->first) {
                prev->second += it->second;
                free_.erase(it);
            }
        }
    }
    size_t free_bytes() const {
        size_t f = 0;
        for (const auto& kv : free_) f += kv.second;
        return f;
    }
    uint8_t* at(size_t off) const { return base_ + off; }
    size_t size() const { return size_; }

   private:
    uint8_t* base_ = nullptr;
    size_t size_ = 0;
    std::map<size_t, size_t> free_;  // offset -> bytes
};

enum JobState { kQueued = 0, kAdmitted = 1, kDone = 2 };

struct Job {
    long long id = 0;
    // request
    const void* src = nullptr;
    bool src_host = false;
    bool fed = false;           // device stack filled by the caller, frames announced by npx_session_feed
    int piece_frames = 0;       // fed jobs: frames the caller announces at a time (chunks are no longer)
    int dtype = 0, N = 0, H = 0, W = 0, M = 0;
    double R = 0, S = 0;
    int iw = 1;
    unsigned mask = 0;
    float *avg = nullptr, *var = nullptr, *tac2 = nullptr;
    const PlanHost* plan = nullptr;
    // resident resources
    size_t off = 0, bytes = 0;  // one pool block: [tables][state][chunk partials][stack if copied]
    size_t state_at = 0, part_at = 0, stack_at = 0;
    int cmax = 1;               // frame chunks one slice may run as (partials merged into the state)
    int cfpc = 0;               // cmax > 1: frames per chunk, fixed per job so chunk boundaries never depend on
                                // how the job is sliced (its maps are then bit-identical alone and in any mix)
    host::TableBlob blob;       // host copy of the tables (kept until retirement: source of the async copy)
    DevTables t{};
    const void* d_in = nullptr;
    hipEvent_t copied = nullptr, done = nullptr;  // copied: the tables are on the device
    // a host stack is uploaded in pieces, each followed by an event, so that its first slices run while the rest
    // is still being copied: pieces[k] marks frames [0, piece_end[k]) on the device
    std::vector<hipEvent_t> pieces;
    std::vector<int> piece_end;
    size_t pieces_done = 0;
    int arrived = 0;  // frames known to be on the device
    int next_frame = 0;
    bool final_launched = false;
    int state = kQueued;
};

struct Session {
    std::mutex mu;
    Pool pool;
    void* mem = nullptr;
    hipStream_t copy = nullptr;
    int max_active = 0;
    long long round_px = 0;  // output pixel-frames one step aims to process
    long long next_id = 1;
    std::deque<long long> queue;
    std::map<long long, Job> jobs;
    int active = 0;
    long long launches = 0, slices = 0, steps = 0;
    std::deque<hipEvent_t> inflight;  // end of each step's launches still running on the GPU
    std::vector<hipEvent_t> spare;
};

constexpr size_t kMaxInflight = 2;  // steps queued on the GPU ahead of the host: new jobs wait at most this many
constexpr int kMinSliceFrames = 8;  // frames per slice (and per chunk of a slice) at least
constexpr int kMaxChunks = 8;       // frame chunks of one slice at most
constexpr int kTargetBlocks = 1024; // blocks a lone job's slice aims for (chunks add blocks to small jobs)
constexpr size_t kPieceBytes = size_t(32) << 20;  // host-stack upload piece, about (whole frames, whole chunks)

// Releases a finished job's pool block and events (its maps are written and complete).
void retire(Session& s, Job& j) {
    s.pool.release(j.off, j.bytes);
    if (j.copied) (void)hipEventDestroy(j.copied);
    if (j.done) (void)hipEventDestroy(j.done);
    for (hipEvent_t ev : j.pieces) (void)hipEventDestroy(ev);
    j.pieces.clear();
    j.copied = j.done = nullptr;
    j.blob.bytes.clear();
    j.blob.bytes.shrink_to_fit();
    j.state = kDone;
    --s.active;
}

void poll_done(Session& s) {
    for (auto& kv : s.jobs) {
        Job& j = kv.second;
        if (j.state == kAdmitted && j.final_launched && hipEventQuery(j.done) == hipSuccess) retire(s, j);
    }
}

// Admits queued jobs in arrival order while the pool has room; a job that does not fit yet lets later, smaller
// jobs pass.
int admit(Session& s) {
    for (auto it = s.queue.begin(); it != s.queue.end() && s.active < s.max_active;) {
        Job& j = s.jobs[*it];
        size_t off = 0;
        if (!s.pool.alloc(j.bytes, &off)) {
            ++it;
            continue;
        }
        j.off = off;
        uint8_t* base = s.pool.at(off);
        hipError_t e = hipEventCreateWithFlags(&j.copied, hipEventDisableTiming);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&j.done, hipEventDisableTiming);
        if (e == hipSuccess)
            e = hipMemcpyAsync(base, j.blob.bytes.data(), j.blob.bytes.size(), hipMemcpyHostToDevice, s.copy);
        if (e == hipSuccess) e = hipEventRecord(j.copied, s.copy);
        if (j.fed) {
            // the caller's device stack: frames arrive as npx_session_feed announces them
        } else if (j.src_host) {
            // pieces of whole frames (and whole chunks when a slice runs as chunks)
            const size_t fb = static_cast<size_t>(j.H) * j.W * (j.dtype == NPX_DTYPE_U16 ? 2 : 4);
            int pf = static_cast<int>(std::max<size_t>(kMinSliceFrames, kPieceBytes / fb));
            if (j.cmax > 1) pf = std::max(1, pf / j.cfpc) * j.cfpc;
            for (int f0 = 0; e == hipSuccess && f0 < j.N; f0 += pf) {
                const int nf = std::min(pf, j.N - f0);
                hipEvent_t ev = nullptr;
                e = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
                if (e != hipSuccess) break;
                j.pieces.push_back(ev);
                j.piece_end.push_back(f0 + nf);
                e = hipMemcpyAsync(base + j.stack_at + f0 * fb, static_cast<const uint8_t*>(j.src) + f0 * fb,
                                   nf * fb, hipMemcpyHostToDevice, s.copy);
                if (e == hipSuccess) e = hipEventRecord(ev, s.copy);
            }
        } else {
            j.arrived = j.N;
        }
        if (e != hipSuccess) {
            // the job stays queued; nothing may still write into the block it gives back
            (void)hipStreamSynchronize(s.copy);
            if (j.copied) (void)hipEventDestroy(j.copied);
            if (j.done) (void)hipEventDestroy(j.done);
            j.copied = j.done = nullptr;
            if (!j.fed) {  // a fed job keeps the caller's announcements
                for (hipEvent_t ev : j.pieces) (void)hipEventDestroy(ev);
                j.pieces.clear();
                j.piece_end.clear();
            }
            s.pool.release(off, j.bytes);
            return hip_fail(e, "session admission copy");
        }
        j.t = j.blob.tables_in(base);
        j.d_in = j.src_host ? static_cast<const void*>(base + j.stack_at) : j.src;
        j.state = kAdmitted;
        ++s.active;
        it = s.queue.erase(it);
    }
    return NPX_OK;
}

// Frames of the next slice of job j given `left` output pixel-frames of the step's budget: at least
// kMinSliceFrames, whole chunks when the slice runs as chunks (chunk boundaries fixed per job), at most cmax chunks
// (the job's partials hold cmax chunk slots), and no more than the frames already on the device.
long long slice_frames(const Job& j, long long left) {
    const long long P = static_cast<long long>(j.plan->HM) * j.plan->WM;
    long long f = std::max<long long>(kMinSliceFrames, left / P);
    if (j.cmax > 1) {
        f = std::max<long long>(1, f / j.cfpc) * j.cfpc;
        f = std::min<long long>(f, static_cast<long long>(j.cmax) * j.cfpc);
    }
    return std::min<long long>(j.arrived - j.next_frame, f);
}

// One step: retire, admit, then one frame slice of every runnable job in batched launches on `st`.
int step(Session& s, hipStream_t st, int* nlaunch) {
    *nlaunch = 0;
    ++s.steps;
    poll_done(s);
    // back-pressure: a step only launches when at most kMaxInflight - 1 earlier steps are still on the GPU, so a
    // newly admitted job joins the launches within kMaxInflight steps
    while (!s.inflight.empty() && hipEventQuery(s.inflight.front()) == hipSuccess) {
        s.spare.push_back(s.inflight.front());
        s.inflight.pop_front();
    }
    if (s.inflight.size() >= kMaxInflight) return NPX_OK;
    int rc = admit(s);
    if (rc) return rc;
    std::vector<Job*> run;
    for (auto& kv : s.jobs) {
        Job& j = kv.second;
        if (j.state != kAdmitted || j.final_launched || hipEventQuery(j.copied) != hipSuccess) continue;
        while (j.pieces_done < j.pieces.size() && hipEventQuery(j.pieces[j.pieces_done]) == hipSuccess) {
            j.arrived = j.piece_end[j.pieces_done++];
            // slices start on the job's chunk boundaries: a caller's announcement that ends inside a chunk counts
            // only up to the chunk's start until the rest arrives
            if (j.fed && j.cmax > 1 && j.arrived < j.N) j.arrived = j.arrived / j.cfpc * j.cfpc;
        }
        if (j.arrived > j.next_frame) run.push_back(&j);
    }
    if (run.empty()) return NPX_OK;
    // Frames of each job this step. Small jobs (remaining output pixel-frames at most a quarter of the step's
    // budget) go first and finish in this step; the rest of the budget goes to the other jobs oldest first, so a
    // large job runs nearly as fast as alone while small arrivals never wait behind it. A job that gets a slice
    // gets at least kMinSliceFrames frames (or what remains); once the budget is spent, later large jobs wait.
    auto remaining = [](const Job* j) {
        return static_cast<long long>(j->N - j->next_frame) * j->plan->HM * j->plan->WM;
    };
    const long long small_cut = s.round_px / 4;
    std::sort(run.begin(), run.end(), [&](const Job* a, const Job* b) {
        const bool sa = remaining(a) <= small_cut, sb = remaining(b) <= small_cut;
        if (sa != sb) return sa;
        if (sa) return remaining(a) != remaining(b) ? remaining(a) < remaining(b) : a->id < b->id;
        return a->id < b->id;
    });
    std::vector<int> frames;
    long long left = s.round_px;
    size_t keep = 0;
    for (size_t k = 0; k < run.size(); ++k) {
        const Job& j = *run[k];
        const bool small = remaining(&j) <= small_cut;
        if (!small && left <= 0) break;
        const long long P = static_cast<long long>(j.plan->HM) * j.plan->WM;
        const long long f = slice_frames(j, left);
        frames.push_back(static_cast<int>(f));
        left -= f * P;
        keep = k + 1;
    }
    run.resize(keep);
    const int nw = kBatchThreads / 64;
    struct Merge {
        Job* job;
        int nchunk, fpc, nframes, n0;
        bool final_;
    };
    std::vector<Merge> merges;
    for (size_t j0 = 0; j0 < run.size(); j0 += kMaxJobs) {
        const size_t nj = std::min<size_t>(kMaxJobs, run.size() - j0);
        bool var = false;
        for (size_t k = j0; k < j0 + nj; ++k) var = var || (run[k]->mask & NPX_VAR);
        BatchArgs b{};
        b.njobs = static_cast<int>(nj);
        int blocks = 0;
        for (size_t k = 0; k < nj; ++k) {
            Job& j = *run[j0 + k];
            const PlanHost& h = *j.plan;
            const int nf = frames[j0 + k];
            host::Shape sh;
            sh.nwaves = nw;
            sh.ngroups = (h.nphase + nw - 1) / nw;
            sh.nregx = (j.W + 63) / 64;
            sh.nregy = (j.H + kQTemporal - 1) / kQTemporal;
            FusedArgs a = host::job_args(j.d_in, j.dtype, j.next_frame + nf, j.H, j.W, j.M, h, &j.t, sh, j.iw);
            host::add_radius(a, h);
            host::add_sensitivity(a, j.S);  // one radius: the lattice half-width is its K (add_radius)
            a.f0 = j.next_frame;
            a.avg = (j.mask & NPX_AVG) ? j.avg : nullptr;
            a.var = (j.mask & NPX_VAR) ? j.var : nullptr;
            a.tac2 = (j.mask & NPX_TAC2) ? j.tac2 : nullptr;
            Merge mg{&j, 1, nf, nf, j.next_frame, j.next_frame + nf == j.N};
            if (j.cmax > 1) {
                // frame chunks in parallel, each from a fresh start into its partial slot; merged after the launch
                mg.fpc = j.cfpc;
                mg.nchunk = (nf + mg.fpc - 1) / mg.fpc;
                if (mg.nchunk > j.cmax) return fail(NPX_ERR_INTERNAL, "session slice has more chunks than partial slots");
                a.fpc = mg.fpc;
                a.nchunk = mg.nchunk;
                a.part = reinterpret_cast<float*>(s.pool.at(j.off) + j.part_at);
                a.direct = 0;
                merges.push_back(mg);
            } else {
                a.fpc = nf;
                a.nchunk = 1;
                a.n0 = j.next_frame;
                a.state = reinterpret_cast<float*>(s.pool.at(j.off) + j.state_at);
                a.direct = mg.final_;
            }
            b.job[k] = a;
            b.nregx[k] = sh.nregx;
            b.nregy[k] = sh.nregy;
            b.start[k] = blocks;
            blocks += sh.nregx * sh.nregy * sh.ngroups * a.nchunk;
            j.next_frame += nf;
            ++s.slices;
        }
        b.start[nj] = blocks;
        hipError_t e = launch_batched(b, var ? kModeTemporalVar : kModeTemporal, st);
        if (e != hipSuccess) return hip_fail(e, "esrrf_batched(served) launch");
        ++s.launches;
        ++*nlaunch;
    }
    for (const Merge& mg : merges) {
        Job& j = *mg.job;
        const size_t P = static_cast<size_t>(j.plan->HM) * j.plan->WM;
        hipError_t e = launch_merge_state(reinterpret_cast<float*>(s.pool.at(j.off) + j.part_at), mg.nchunk, mg.fpc,
                                          mg.nframes, P, reinterpret_cast<float*>(s.pool.at(j.off) + j.state_at),
                                          mg.n0, mg.final_ ? 0 : 1, (j.mask & NPX_AVG) ? j.avg : nullptr,
                                          (j.mask & NPX_VAR) ? j.var : nullptr, (j.mask & NPX_TAC2) ? j.tac2 : nullptr,
                                          st);
        if (e != hipSuccess) return hip_fail(e, "session moments merge");
    }
    for (Job* jp : run) {
        Job& j = *jp;
        if (j.next_frame == j.N) {
            hipError_t e = hipEventRecord(j.done, st);
            if (e != hipSuccess) return hip_fail(e, "session completion event");
            j.final_launched = true;
        }
    }
    hipEvent_t ev = nullptr;
    if (!s.spare.empty()) {
        ev = s.spare.back();
        s.spare.pop_back();
    } else {
        hipError_t e = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
        if (e != hipSuccess) return hip_fail(e, "session step event");
    }
    hipError_t e = hipEventRecord(ev, st);
    if (e != hipSuccess) return hip_fail(e, "session step event");
    s.inflight.push_back(ev);
    return NPX_OK;
}

// Pool footprint of a job: [tables][running state 5P][chunk partials cmax * 5P][stack when copied from the host].
// cmax: chunks a slice may run as, enough for about kTargetBlocks blocks when the job runs alone.
void layout(Job& j, const PlanHost& plan) {
    const size_t P = static_cast<size_t>(plan.HM) * plan.WM;
    const int nw = kBatchThreads / 64;
    const long long spatial = static_cast<long long>((j.W + 63) / 64) * ((j.H + kQTemporal - 1) / kQTemporal) *
                              ((plan.nphase + nw - 1) / nw);
    j.cmax = static_cast<int>(std::max<long long>(1, std::min<long long>(kMaxChunks, (kTargetBlocks + spatial - 1) /
                                                                                         spatial)));
    j.cmax = std::min(j.cmax, std::max(1, j.N / kMinSliceFrames));
    j.cfpc = std::max(kMinSliceFrames, (j.N + j.cmax - 1) / j.cmax);
    if (j.src_host) {
        // chunks no longer than an upload piece, so that slices can follow the pieces as they arrive
        const size_t fb = static_cast<size_t>(j.H) * j.W * (j.dtype == NPX_DTYPE_U16 ? 2 : 4);
        const int pf = static_cast<int>(std::max<size_t>(kMinSliceFrames, kPieceBytes / fb));
        j.cfpc = std::max(kMinSliceFrames, std::min(j.cfpc, pf));
    } else if (j.fed) {
        j.cfpc = std::max(kMinSliceFrames, std::min(j.cfpc, j.piece_frames));
    }
    j.state_at = align_up(j.blob.bytes.size());
    j.part_at = j.state_at + align_up(5 * P * sizeof(float));
    j.stack_at = j.part_at + (j.cmax > 1 ? align_up(static_cast<size_t>(j.cmax) * 5 * P * sizeof(float)) : 0);
    j.bytes = j.stack_at + (j.src_host ? static_cast<size_t>(j.N) * j.H * j.W * (j.dtype == NPX_DTYPE_U16 ? 2 : 4) : 0);
}

}  // namespace serve
}  // namespace npx

using npx::serve::Job;
using npx::serve::Session;

extern "C" {

int npx_session_create(size_t pool_bytes, int max_active, long long px_frames_per_step, void** session) {
    if (!session || pool_bytes == 0 || max_active < 1 || px_frames_per_step < 1)
        return npx::host::fail(NPX_ERR_ARG, "session needs pool_bytes > 0, max_active >= 1, px_frames_per_step >= 1");
    auto s = std::make_unique<Session>();
    hipError_t e = hipMalloc(&s->mem, pool_bytes);
    if (e != hipSuccess) return npx::host::hip_fail(e, "hipMalloc(session pool)");
    e = hipStreamCreateWithFlags(&s->copy, hipStreamNonBlocking);
    if (e != hipSuccess) {
        (void)hipFree(s->mem);
        return npx::host::hip_fail(e, "hipStreamCreateWithFlags(session copy stream)");
    }
    s->pool.init(static_cast<uint8_t*>(s->mem), pool_bytes);
    s->max_active = max_active;
    s->round_px = px_frames_per_step;
    *session = s.release();
    return NPX_OK;
}

}  // extern "C"

namespace npx {
namespace serve {

// Validates and queues a job; fed: a device stack the caller fills, announced frames at a time (piece_frames).
int submit(void* session, const void* src, int src_is_host, int dtype, int N, int H, int W, int M, double R, double S,
           int iw, unsigned mode_mask, float* d_avg, float* d_var, float* d_tac2, bool fed, int piece_frames,
           long long* job_id) {
    if (!session || !job_id) return npx::host::fail(NPX_ERR_ARG, "NULL session or job id output");
    if (!src) return npx::host::fail(NPX_ERR_ARG, "input pointer is NULL");
    if (dtype != NPX_DTYPE_F32 && dtype != NPX_DTYPE_U16) return npx::host::fail(NPX_ERR_ARG, "dtype must be f32 or u16");
    if (N < 1 || H < 1 || W < 1) return npx::host::fail(NPX_ERR_ARG, "stack must have N, H, W >= 1");
    if (M < 1 || M > 64) return npx::host::fail(NPX_ERR_ARG, "magnification must be in [1, 64]");
    if ((mode_mask & (NPX_AVG | NPX_VAR | NPX_TAC2)) == 0) return npx::host::fail(NPX_ERR_ARG, "empty temporal mode mask");
    if (((mode_mask & NPX_AVG) && !d_avg) || ((mode_mask & NPX_VAR) && !d_var) || ((mode_mask & NPX_TAC2) && !d_tac2))
        return npx::host::fail(NPX_ERR_ARG, "output pointer of a requested temporal mode is NULL");
    Session& s = *static_cast<Session*>(session);
    const npx::PlanHost* plan = nullptr;
    int rc = npx::host::host_plan(H, W, M, R, &plan);
    if (rc) return rc;
    Job j;
    j.src = src;
    j.src_host = src_is_host != 0;
    j.dtype = dtype;
    j.N = N;
    j.H = H;
    j.W = W;
    j.M = M;
    j.R = R;
    j.S = S;
    j.iw = iw ? 1 : 0;
    j.mask = mode_mask;
    j.avg = d_avg;
    j.var = d_var;
    j.tac2 = d_tac2;
    j.plan = plan;
    j.fed = fed;
    j.piece_frames = piece_frames;
    npx::host::pack_tables(*plan, &j.blob);
    npx::serve::layout(j, *plan);
    std::lock_guard<std::mutex> lk(s.mu);
    if (npx::serve::align_up(j.bytes) > s.pool.size())
        return npx::host::fail(NPX_ERR_ARG, "job needs more device memory than the session pool holds");
    j.id = s.next_id++;
    *job_id = j.id;
    s.queue.push_back(j.id);
    s.jobs.emplace(j.id, std::move(j));
    return NPX_OK;
}

}  // namespace serve
}  // namespace npx

extern "C" {

int npx_session_submit(void* session, const void* src, int src_is_host, int dtype, int N, int H, int W, int M,
                       double R, double S, int iw, unsigned mode_mask, float* d_avg, float* d_var, float* d_tac2,
                       long long* job_id) {
    return npx::serve::submit(session, src, src_is_host, dtype, N, H, W, M, R, S, iw, mode_mask, d_avg, d_var, d_tac2,
                              false, 0, job_id);
}

int npx_session_submit_fed(void* session, const void* d_stack, int dtype, int N, int H, int W, int M, double R,
                           double S, int iw, unsigned mode_mask, float* d_avg, float* d_var, float* d_tac2,
                           int piece_frames, long long* job_id) {
    if (piece_frames < 1) return npx::host::fail(NPX_ERR_ARG, "piece_frames must be >= 1");
    return npx::serve::submit(session, d_stack, 0, dtype, N, H, W, M, R, S, iw, mode_mask, d_avg, d_var, d_tac2, true,
                              piece_frames, job_id);
}

int npx_session_feed(void* session, long long job_id, int frames, void* stream) {
    if (!session) return npx::host::fail(NPX_ERR_ARG, "NULL session");
    Session& s = *static_cast<Session*>(session);
    std::lock_guard<std::mutex> lk(s.mu);
    auto it = s.jobs.find(job_id);
    if (it == s.jobs.end()) return npx::host::fail(NPX_ERR_ARG, "unknown job id");
    Job& j = it->second;
    if (!j.fed) return npx::host::fail(NPX_ERR_ARG, "job was not submitted with npx_session_submit_fed");
    const int last = j.piece_end.empty() ? 0 : j.piece_end.back();
    if (frames <= last || frames > j.N)
        return npx::host::fail(NPX_ERR_ARG, "fed frame count must grow and stay within the stack");
    hipEvent_t ev = nullptr;
    hipError_t e = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
    if (e != hipSuccess) return npx::host::hip_fail(e, "session feed event");
    e = hipEventRecord(ev, static_cast<hipStream_t>(stream));
    if (e != hipSuccess) {
        (void)hipEventDestroy(ev);
        return npx::host::hip_fail(e, "session feed event");
    }
    j.pieces.push_back(ev);
    j.piece_end.push_back(frames);
    return NPX_OK;
}

size_t npx_session_job_bytes(int N, int H, int W, int M, double R, int dtype, int src_is_host) {
    if (N < 1 || H < 1 || W < 1 || M < 1 || M > 64) return 0;
    const npx::PlanHost* plan = nullptr;
    if (npx::host::host_plan(H, W, M, R, &plan) != NPX_OK) return 0;
    Job j;
    j.N = N;
    j.H = H;
    j.W = W;
    j.M = M;
    j.dtype = dtype;
    j.src_host = src_is_host != 0;
    npx::host::pack_tables(*plan, &j.blob);
    npx::serve::layout(j, *plan);
    return npx::serve::align_up(j.bytes);
}

int npx_session_slice_plan(int N, int H, int W, int M, double R, int dtype, int src_is_host, long long px_frames_per_step,
                           int* info) {
    if (!info || N < 1 || H < 1 || W < 1 || M < 1 || M > 64 || px_frames_per_step < 1)
        return npx::host::fail(NPX_ERR_ARG, "bad arguments");
    const npx::PlanHost* plan = nullptr;
    int rc = npx::host::host_plan(H, W, M, R, &plan);
    if (rc) return rc;
    Job j;
    j.N = N;
    j.H = H;
    j.W = W;
    j.M = M;
    j.dtype = dtype;
    j.src_host = src_is_host != 0;
    j.plan = plan;
    npx::host::pack_tables(*plan, &j.blob);
    npx::serve::layout(j, *plan);
    // a lone job whose frames arrive as admit() cuts the upload (every frame at once from device memory), given
    // the whole budget each step until done
    std::vector<int> ends;
    if (j.src_host) {
        const size_t fb = static_cast<size_t>(H) * W * (dtype == NPX_DTYPE_U16 ? 2 : 4);
        int pf = static_cast<int>(std::max<size_t>(npx::serve::kMinSliceFrames, npx::serve::kPieceBytes / fb));
        if (j.cmax > 1) pf = std::max(1, pf / j.cfpc) * j.cfpc;
        for (int f0 = 0; f0 < N; f0 += pf) ends.push_back(std::min(N, f0 + pf));
    } else {
        ends.push_back(N);
    }
    // twice: stepped as each piece lands, and stepped only once the whole upload has landed
    int slices = 0, max_chunks = 0, bad_bounds = 0;
    for (int pass = 0; pass < 2; ++pass) {
        j.next_frame = 0;
        for (size_t k = pass ? ends.size() - 1 : 0; k < ends.size(); ++k) {
            j.arrived = ends[k];
            while (j.next_frame < j.arrived) {
                const long long f = npx::serve::slice_frames(j, px_frames_per_step);
                if (f < 1) return npx::host::fail(NPX_ERR_INTERNAL, "empty slice");
                const int nchunk = j.cmax > 1 ? static_cast<int>((f + j.cfpc - 1) / j.cfpc) : 1;
                max_chunks = std::max(max_chunks, nchunk);
                if (j.cmax > 1 && j.next_frame % j.cfpc) ++bad_bounds;  // slices start on the job's chunk boundaries
                j.next_frame += static_cast<int>(f);
                ++slices;
            }
        }
    }
    info[0] = j.cmax;
    info[1] = j.cfpc;
    info[2] = max_chunks;
    info[3] = slices;
    info[4] = bad_bounds;
    info[5] = static_cast<int>(ends.size());
    return NPX_OK;
}

int npx_session_step(void* session, void* stream, int* launches) {
    if (!session) return npx::host::fail(NPX_ERR_ARG, "NULL session");
    Session& s = *static_cast<Session*>(session);
    std::lock_guard<std::mutex> lk(s.mu);
    int n = 0;
    const int rc = npx::serve::step(s, static_cast<hipStream_t>(stream), &n);
    if (launches) *launches = n;
    return rc;
}

int npx_session_poll(void* session, long long job_id) {
    if (!session) return npx::host::fail(NPX_ERR_ARG, "NULL session");
    Session& s = *static_cast<Session*>(session);
    std::lock_guard<std::mutex> lk(s.mu);
    auto it = s.jobs.find(job_id);
    if (it == s.jobs.end()) return npx::host::fail(NPX_ERR_ARG, "unknown job id");
    Job& j = it->second;
    if (j.state == npx::serve::kAdmitted && j.final_launched && hipEventQuery(j.done) == hipSuccess)
        npx::serve::retire(s, j);
    return j.state;
}

int npx_session_forget(void* session, long long job_id) {
    if (!session) return npx::host::fail(NPX_ERR_ARG, "NULL session");
    Session& s = *static_cast<Session*>(session);
    std::lock_guard<std::mutex> lk(s.mu);
    auto it = s.jobs.find(job_id);
    if (it == s.jobs.end()) return npx::host::fail(NPX_ERR_ARG, "unknown job id");
    if (it->second.state != npx::serve::kDone) return npx::host::fail(NPX_ERR_ARG, "job is not done");
    s.jobs.erase(it);
    return NPX_OK;
}

int npx_session_stats(void* session, long long* stats) {
    if (!session || !stats) return npx::host::fail(NPX_ERR_ARG, "NULL session or stats output");
    Session& s = *static_cast<Session*>(session);
    std::lock_guard<std::mutex> lk(s.mu);
    stats[0] = static_cast<long long>(s.queue.size());
    stats[1] = s.active;
    stats[2] = static_cast<long long>(s.pool.free_bytes());
    stats[3] = s.launches;
    stats[4] = s.slices;
    stats[5] = s.steps;
    return NPX_OK;
}

int npx_session_destroy(void* session) {
    if (!session) return NPX_OK;
    std::unique_ptr<Session> s(static_cast<Session*>(session));
    int rc = NPX_OK;
    hipError_t e = hipDeviceSynchronize();
    if (e != hipSuccess) rc = npx::host::hip_fail(e, "hipDeviceSynchronize(session destroy)");
    for (auto& kv : s->jobs) {
        if (kv.second.copied) (void)hipEventDestroy(kv.second.copied);
        for (hipEvent_t ev : kv.second.pieces) (void)hipEventDestroy(ev);
        if (kv.second.done) (void)hipEventDestroy(kv.second.done);
    }
    for (hipEvent_t ev : s->inflight) (void)hipEventDestroy(ev);
    for (hipEvent_t ev : s->spare) (void)hipEventDestroy(ev);
    (void)hipStreamDestroy(s->copy);
    (void)hipFree(s->mem);
    return rc;
}

}  // extern "C"
