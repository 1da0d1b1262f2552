// The serving loop: continuous batching of FDTD jobs on one device (pxfdtd_server_*).
//
// photonoxide has no counterpart (its nearest is the process-level farm, src/farm.rs:1-56):
// this is the build's own. Jobs are host-mirror simulations handed over with a step count. A
// loop thread admits them in arrival order while the jobs resident on the device fit the
// server's HBM budget and job count, and steps every admitted job together in segments of K
// steps. Each step of a segment is one launch per kernel class over the units of all jobs (the
// aligned march's units, four per workgroup, grouped by job), one tap launch over all jobs'
// monitored values, and per segment one FP64 transform accumulation over all of them. The
// segment is captured once per job set as a hipGraph and replayed; a job that finishes mid
// segment is skipped by the kernels from its last step on. Jobs admitted while others run join
// at the next segment boundary; a finished job's results go back to its simulation and its
// device memory is freed at once.
//
// Every cell is updated by the same device code as when the job runs alone (the march's unit
// body, the taps, the step-ordered FP64 sums, no atomics), so a job's fields and spectra are
// the same bits alone and in any mix. Tilings suit a full batch (2D: one row, chunks of up to 16
// rows; 3D: two rows, 8-plane chunks) or, with autotune trials asked for, come from
// pxfdtd_resident_autotune on the first 3D job of each class of grid shape (each extent's power
// of two) and are cached per class.
#pragma once
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <memory>
#include <numeric>
#include <mutex>
#include <thread>

#include "batch_launch.hpp"

// the host mirror's side of the server (simulation.cpp)
namespace pxfdtd_host {
int sim_footprint(pxfdtd_simulation *s, int32_t precision, size_t *bytes);
int sim_servable(const pxfdtd_simulation *s);
int sim_attach(pxfdtd_simulation *s, pxfdtd_device *dev, int32_t precision, pxfdtd_resident **out);
int sim_emitters(const pxfdtd_simulation *s);
int64_t sim_steps(const pxfdtd_simulation *s);
void sim_tables(const pxfdtd_simulation *s, int64_t n0, int count, int slots, double *waves, double *phases);
void sim_advanced(pxfdtd_simulation *s, int64_t count);
int sim_detach(pxfdtd_simulation *s);
int sim_fetch(pxfdtd_simulation *s);
void sim_release(pxfdtd_simulation *s);
}  // namespace pxfdtd_host

namespace pxfdtd {

template <typename T>
struct TapJob {
    const T *fields[2][6];  // copy, then Ex Ey Ez Hx Hy Hz
    T *staging;             // [K][ntaps]
    long long ntaps;
};

// after step s, every monitored value of every job that took it, from the copy it wrote
template <typename T>
__global__ void tap_batch(const TapJob<T> *tj, const JobSeg *seg, const int32_t *vjob,
                          const int64_t *vlocal, const int64_t *vindex, const int8_t *vwhich,
                          long long n, int s) {
    const long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n) return;
    const int j = vjob[v];
    const JobSeg g = seg[j];
    if (s >= g.active) return;
    const int q = (g.p0 + s + 1) & 1;
    const TapJob<T> &J = tj[j];
    J.staging[(long long)s * J.ntaps + vlocal[v]] = J.fields[q][vwhich[v]][vindex[v]];
}

template <typename T>
struct DftJob {
    const T *staged;
    long long row;
    const int32_t *series_of;
    const SeriesInfo *series;
    const double *phases;  // [K][slots] complex, this segment's
    int slots;
    double *sums;
};

// the segment's FP64 transform sums of every job's values, each in step order as
// dft_accumulate adds them (the same bits as a job stepped alone)
template <typename T>
__global__ void dft_batch(const DftJob<T> *dj, const JobSeg *seg, const int32_t *vjob,
                          const int64_t *vlocal, long long n) {
    const long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n) return;
    const int j = vjob[v];
    const int steps = seg[j].active;
    if (!steps) return;
    const DftJob<T> &D = dj[j];
    const long long lv = vlocal[v];
    const SeriesInfo S = D.series[D.series_of[lv]];
    const long long local = lv - S.base_v;
    for (int f = 0; f < S.nf; f++) {
        const long long o = S.sums_off + (long long)f * S.volume + local;
        double re = D.sums[2 * o], im = D.sums[2 * o + 1];
        for (int st = 0; st < steps; st++) {
            const double F = (double)D.staged[(long long)st * D.row + lv];
            const double pr = D.phases[2 * ((long long)st * D.slots + S.slot + f)];
            const double pi = D.phases[2 * ((long long)st * D.slots + S.slot + f) + 1];
            re += F * pr - 0.0 * pi;
            im += F * pi + 0.0 * pr;
        }
        D.sums[2 * o] = re;
        D.sums[2 * o + 1] = im;
    }
}

}  // namespace pxfdtd

// ------------------------------------------------------------------------------ the server
struct ServerJob {
    int64_t id = 0;
    pxfdtd_simulation *sim = nullptr;
    int64_t steps = 0, left = 0;
    int32_t state = PXFDTD_JOB_QUEUED;
    double t_submit = 0, t_start = 0, t_end = 0;
    size_t foot = 0;
    double work = 0;                 // cells x steps at submission
    std::atomic<bool> cancel{false};  // asked to stop: it steps no more, stops at the boundary
    pxfdtd_resident *r = nullptr;
    int emitters = 0, slots = 0;
    std::string error;
    size_t wave_off = 0, phase_off = 0;  // this job's first double in the segment arenas
};

struct pxfdtd_server {
    pxfdtd_device *dev = nullptr;
    pxfdtd_server_config cfg;
    hipStream_t stream = nullptr;
    hipEvent_t seg_start = nullptr, seg_stop = nullptr;  // a segment's span on the device
    std::thread loop;
    std::mutex mu;
    std::condition_variable wake, changed;
    std::deque<ServerJob *> queue;     // submitted, not admitted
    std::vector<ServerJob *> running;  // admitted, not finished
    std::map<int64_t, std::unique_ptr<ServerJob>> jobs;
    int64_t next_id = 1;
    bool stop = false;
    size_t used = 0;
    std::chrono::steady_clock::time_point t0;
    std::map<std::tuple<int, int, int, int, int>, std::pair<int, int>> tuned;  // shape -> tiling
    std::mutex tuned_mu;  // admissions run on several host threads
    pxfdtd_server_stats stats;
    // the tables of the job set last built, on the device
    std::vector<ServerJob *> table;  // table entry -> job (finished ones stay, with no steps)
    struct Class {
        int kind, mode, v, ax;  // v: rows per strip (the aligned march) or field set (2D)
        int first, count;       // blocks
    };
    std::vector<Class> classes;
    std::vector<BatchBlock> all_blocks;  // every table job's workgroups, class after class
    Buf d_jobs, d_colds, d_seg, d_blocks, d_counts, d_units, d_waves, d_phases;
    BatchBlock *h_blocks = nullptr;  // pinned: this segment's active workgroups, per class
    int32_t *h_counts = nullptr;     // pinned: their number per class
    size_t cap_blocks = 0, cap_counts = 0;
    Buf d_tapjobs, d_tjob, d_tlocal, d_tindex, d_twhich, d_dftjobs, d_djob, d_dlocal;
    long long ntaps = 0, nvals = 0;
    double *h_waves = nullptr, *h_phases = nullptr;  // pinned
    JobSeg *h_seg = nullptr;
    size_t cap_waves = 0, cap_phases = 0, cap_seg = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::chrono::steady_clock::time_point seg_t0;  // the segment in flight
    double seg_cells = 0;
    double t_seg_end = -1;  // when the last segment ended (-1: the loop has waited since)
    int64_t captures = 0;  // read into stats under the lock
};

static double server_now(const pxfdtd_server *S) {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - S->t0).count();
}

static void drop_graph(pxfdtd_server *S) {
    if (S->exec) (void)hipGraphExecDestroy(S->exec);
    if (S->graph) (void)hipGraphDestroy(S->graph);
    S->exec = nullptr;
    S->graph = nullptr;
}

template <typename T>
static int pinned(T *&p, size_t &cap, size_t n) {
    if (n <= cap) return PXFDTD_OK;
    if (p) (void)hipHostFree(p);
    p = nullptr;
    cap = 0;
    HIPCHK(hipHostMalloc((void **)&p, std::max<size_t>(n, 1) * sizeof(T), hipHostMallocDefault));
    cap = n;
    return PXFDTD_OK;
}

// a job on the device: its resident, the aligned march and a tiling for its grid's shape
static int server_admit(pxfdtd_server *S, ServerJob *J) {
    TRY(pxfdtd_host::sim_attach(J->sim, S->dev, PXFDTD_F32, &J->r));
    pxfdtd_resident *r = J->r;
    if (r->kernel != 3 && r->kernel != 4) TRY(pxfdtd_resident_set_kernel(r, 3));
    // tilings are cached per class of shape: 2D or 3D, each extent's power of two, the
    // coefficient storage
    auto lg = [](int n) { int b = 0; while ((2 << b) <= n) b++; return b; };
    const auto key = std::make_tuple(r->nz == 1 ? 0 : 1, lg(r->nx), lg(r->ny), lg(r->nz), r->mode);
    std::pair<int, int> known = {0, 0};
    bool have = false;
    {
        std::lock_guard<std::mutex> g(S->tuned_mu);
        auto it = S->tuned.find(key);
        have = it != S->tuned.end();
        if (have) known = it->second;
    }
    if (r->kernel == 4) {
        // 2D jobs: in a batch the device is full, so the longest chunks (the least recomputed
        // rows) win, whatever suits the job alone
        TRY(pxfdtd_resident_set_tiling(r, 1, std::min(r->ny, 16)));
    } else if (have) {
        TRY(pxfdtd_resident_set_tiling(r, known.first, known.second));
    } else {
        int32_t rows = 2, planes = 8;
        double ms = 0;
        if (S->cfg.autotune > 0) {
            TRY(pxfdtd_resident_autotune(r, S->cfg.autotune, 2, &rows, &planes, &ms));
        } else {
            // in a batch the device is full: two rows per strip and 8-plane chunks (less H
            // recomputed than the short chunks that fill the chip for one small grid alone)
            rows = 2;
            planes = std::min(r->nz, 8);
            TRY(pxfdtd_resident_set_tiling(r, rows, planes));
        }
        std::lock_guard<std::mutex> g(S->tuned_mu);
        S->tuned[key] = {rows, planes};
    }
    if (r->kernel == 3 && r->ty > 2) TRY(pxfdtd_resident_set_tiling(r, 2, 0));
    r->batched = true;
    TRY(build_units(r));
    // the staging holds a whole segment's taps (here, beside the running segment, rather than
    // in the rebuild between segments)
    if ((long long)r->staging.bytes < (long long)S->cfg.segment * r->ntaps * 4)
        TRY(r->staging.alloc((size_t)S->cfg.segment * r->ntaps * 4));
    J->emitters = std::max(1, pxfdtd_host::sim_emitters(J->sim));
    J->slots = std::max(1, (int)r->slots);
    return PXFDTD_OK;
}

static void launch_class(pxfdtd_server *S, int q, int s) {
    const pxfdtd_server::Class &c = S->classes[q];
    launch_batch_class(c.kind, c.mode, c.v, c.ax, S->d_jobs.as<JobStep<float>>(), S->d_seg.as<JobSeg>(),
                       S->d_blocks.as<BatchBlock>() + c.first, S->d_counts.as<int32_t>() + q,
                       S->d_units.as<int4>(), c.count, s, S->stream);
}

// one segment's launches on the server's stream (captured into the graph, or run directly)
static int launch_segment(pxfdtd_server *S) {
    const int K = S->cfg.segment;
    for (int s = 0; s < K; s++) {
        for (int q = 0; q < (int)S->classes.size(); q++) launch_class(S, q, s);
        if (S->ntaps) {
            const long long blocks = (S->ntaps + 255) / 256;
            hipLaunchKernelGGL((tap_batch<float>), dim3((unsigned)blocks), dim3(256), 0, S->stream,
                               S->d_tapjobs.as<TapJob<float>>(), S->d_seg.as<JobSeg>(), S->d_tjob.as<int32_t>(),
                               S->d_tlocal.as<int64_t>(), S->d_tindex.as<int64_t>(), S->d_twhich.as<int8_t>(),
                               S->ntaps, s);
        }
    }
    if (S->nvals) {
        const long long blocks = (S->nvals + 255) / 256;
        hipLaunchKernelGGL((dft_batch<float>), dim3((unsigned)blocks), dim3(256), 0, S->stream,
                           S->d_dftjobs.as<DftJob<float>>(), S->d_seg.as<JobSeg>(), S->d_djob.as<int32_t>(),
                           S->d_dlocal.as<int64_t>(), S->nvals);
    }
    HIPCHK(hipGetLastError());
    return PXFDTD_OK;
}

template <typename V>
static int upload_vec(Buf &b, const std::vector<V> &v, hipStream_t s) {
    TRY(b.reserve(std::max<size_t>(v.size(), 1) * sizeof(V)));
    if (!v.empty()) HIPCHK(hipMemcpyAsync(b.p, v.data(), v.size() * sizeof(V), hipMemcpyHostToDevice, s));
    return PXFDTD_OK;
}

// the device tables of the running jobs, and the segment's graph
static int server_rebuild(pxfdtd_server *S) {
    drop_graph(S);
    HIPCHK(hipStreamSynchronize(S->stream));
    S->table = S->running;
    const int K = S->cfg.segment;
    const int nj = (int)S->table.size();
    std::vector<JobStep<float>> steps(nj);
    std::vector<StepCold<float>> colds(2 * nj);
    std::vector<TapJob<float>> tapjobs(nj);
    std::vector<DftJob<float>> dftjobs(nj);
    std::vector<int32_t> tjob, djob;
    std::vector<int64_t> tlocal, tindex, dlocal;
    std::vector<int8_t> twhich;
    std::map<std::tuple<int, int, int, int>, std::vector<BatchBlock>> blocks;
    std::vector<int4> units;
    size_t wv = 0, ph = 0, nt = 0, nd = 0;
    for (int j = 0; j < nj; j++) {
        ServerJob *J = S->table[j];
        J->wave_off = wv;
        J->phase_off = ph;
        wv += (size_t)K * J->emitters * 2;
        ph += (size_t)K * J->slots * 2;
        nt += (size_t)J->r->ntaps;
        nd += (size_t)J->r->nvals;
    }
    // every tap and transform value of every job: whole runs per job, not value by value
    tjob.resize(nt), tlocal.resize(nt), tindex.resize(nt), twhich.resize(nt);
    djob.resize(nd), dlocal.resize(nd);
    size_t at_t = 0, at_d = 0;
    TRY(S->d_waves.reserve(std::max<size_t>(wv, 1) * 8));
    TRY(S->d_phases.reserve(std::max<size_t>(ph, 1) * 8));
    TRY(S->d_colds.reserve(std::max<size_t>(colds.size(), 1) * sizeof(StepCold<float>)));
    for (int j = 0; j < nj; j++) {
        ServerJob *J = S->table[j];
        pxfdtd_resident *r = J->r;
        if (r->units_dirty) TRY(build_units(r));
        fill_cold<float>(r, &colds[2 * j], 2);
        for (int q = 0; q < 2; q++) {
            colds[2 * j + q].waves = S->d_waves.as<double>() + J->wave_off;
            StepArgs<float> A = args_of<float>(r);
            A.ein = r->field<float>(q, 0);
            A.hin = r->field<float>(q, 3);
            A.eout = r->field<float>(1 - q, 0);
            A.hout = r->field<float>(1 - q, 3);
            A.cold = S->d_colds.as<StepCold<float>>() + 2 * j + q;
            A.units = nullptr;
            A.wave_step = 0;
            steps[j].A[q] = A;
        }
        // the units, four per workgroup, by class: no slab, x slabs only, the others (the 2D
        // march: no slab, the others)
        const bool two_d = r->kernel == 4;
        const int counts[3] = {r->n_lean, r->n_x, r->n_slab};
        const int axes[3] = {0, 1, two_d ? 3 : 7};
        const int v = two_d ? step_set(r) : r->ty;
        int at = 0;
        for (int c = 0; c < 3; c++) {
            // the aligned march's slab units may be fewer rows than its tiling's
            const int rows = !two_d && c == 2 ? r->slab_rows : v;
            auto &bl = blocks[std::make_tuple(r->kernel, r->mode, rows, axes[c])];
            for (int u = 0; u < counts[c]; u += 4) {
                bl.push_back({j, (int32_t)units.size()});
                for (int w = 0; w < 4; w++)
                    units.push_back(u + w < counts[c] ? r->units_h[at + u + w] : make_int4(0, 0, 0, 0));
            }
            at += counts[c];
        }
        // taps and transform values
        TapJob<float> &tj = tapjobs[j];
        for (int q = 0; q < 2; q++)
            for (int f = 0; f < 6; f++) tj.fields[q][f] = r->field<float>(q, f);
        tj.staging = r->staging.as<float>();
        tj.ntaps = r->ntaps;
        const size_t n_t = (size_t)r->ntaps;
        std::fill_n(tjob.begin() + at_t, n_t, j);
        std::iota(tlocal.begin() + at_t, tlocal.begin() + at_t + n_t, (int64_t)0);
        if (n_t) {
            std::memcpy(tindex.data() + at_t, r->tap_index_h.data(), n_t * sizeof(int64_t));
            std::memcpy(twhich.data() + at_t, r->tap_which_h.data(), n_t * sizeof(int8_t));
        }
        at_t += n_t;
        DftJob<float> &dj = dftjobs[j];
        dj.staged = r->staging.as<float>() + r->nprobes;
        dj.row = r->ntaps;
        dj.series_of = r->series_of.as<int32_t>();
        dj.series = r->series_info.as<SeriesInfo>();
        dj.phases = S->d_phases.as<double>() + J->phase_off;
        dj.slots = J->slots;
        dj.sums = r->sums.as<double>();
        std::fill_n(djob.begin() + at_d, (size_t)r->nvals, j);
        std::iota(dlocal.begin() + at_d, dlocal.begin() + at_d + r->nvals, (int64_t)0);
        at_d += (size_t)r->nvals;
        // the staging must hold a whole segment
        if ((long long)r->staging.bytes < (long long)K * r->ntaps * 4) {
            TRY(r->staging.alloc((size_t)K * r->ntaps * 4));
            tj.staging = r->staging.as<float>();
            dj.staged = r->staging.as<float>() + r->nprobes;
        }
    }
    std::vector<BatchBlock> all;
    S->classes.clear();
    for (auto &kv : blocks) {
        if (kv.second.empty()) continue;
        S->classes.push_back({std::get<0>(kv.first), std::get<1>(kv.first), std::get<2>(kv.first),
                              std::get<3>(kv.first), (int)all.size(), (int)kv.second.size()});
        all.insert(all.end(), kv.second.begin(), kv.second.end());
    }
    hipStream_t st = S->stream;
    TRY(upload_vec(S->d_jobs, steps, st));
    if (!colds.empty()) HIPCHK(hipMemcpyAsync(S->d_colds.p, colds.data(), colds.size() * sizeof(StepCold<float>), hipMemcpyHostToDevice, st));
    TRY(upload_vec(S->d_blocks, all, st));
    S->all_blocks = all;
    TRY(S->d_counts.reserve(std::max<size_t>(S->classes.size(), 1) * sizeof(int32_t)));
    TRY(pinned(S->h_blocks, S->cap_blocks, all.size()));
    TRY(pinned(S->h_counts, S->cap_counts, S->classes.size()));
    TRY(upload_vec(S->d_units, units, st));
    TRY(upload_vec(S->d_tapjobs, tapjobs, st));
    TRY(upload_vec(S->d_tjob, tjob, st));
    TRY(upload_vec(S->d_tlocal, tlocal, st));
    TRY(upload_vec(S->d_tindex, tindex, st));
    TRY(upload_vec(S->d_twhich, twhich, st));
    TRY(upload_vec(S->d_dftjobs, dftjobs, st));
    TRY(upload_vec(S->d_djob, djob, st));
    TRY(upload_vec(S->d_dlocal, dlocal, st));
    TRY(S->d_seg.reserve(std::max<size_t>(nj, 1) * sizeof(JobSeg)));
    S->ntaps = (long long)tjob.size();
    S->nvals = (long long)djob.size();
    TRY(pinned(S->h_waves, S->cap_waves, wv));
    TRY(pinned(S->h_phases, S->cap_phases, ph));
    TRY(pinned(S->h_seg, S->cap_seg, (size_t)nj));
    HIPCHK(hipStreamSynchronize(st));
    if (!S->cfg.direct) {
        HIPCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        const int rc = launch_segment(S);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(st, &g);
        if (rc != PXFDTD_OK) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        HIPCHK(e);
        S->graph = g;
        HIPCHK(hipGraphInstantiate(&S->exec, S->graph, nullptr, nullptr, 0));
        S->captures++;
    }
    return PXFDTD_OK;
}

// one segment: every running job's waves and phases, then the batched steps, launched on the
// server's stream without waiting (server_segment_end waits)
static int server_segment_begin(pxfdtd_server *S, double *prepare, double *gap) {
    const double t_begin = server_now(S);
    const int K = S->cfg.segment;
    const int nj = (int)S->table.size();
    double cells = 0;
    // which jobs step: all that can, or (shortest first) those with the least work left until
    // their cells fill the device
    std::vector<char> go(nj, 0);
    std::vector<int> cand;
    for (int j = 0; j < nj; j++) {
        const ServerJob *J = S->table[j];
        if (J->state == PXFDTD_JOB_RUNNING && J->left > 0 && J->r && !J->cancel.load()) cand.push_back(j);
    }
    if (S->cfg.order == 1) {
        auto rest = [&](int j) { return (double)S->table[j]->left * (double)S->table[j]->r->ncells; };
        std::stable_sort(cand.begin(), cand.end(), [&](int a, int b) { return rest(a) < rest(b); });
        const double fill = S->cfg.fill_cells > 0 ? (double)S->cfg.fill_cells : (double)(1 << 26);
        double filled = 0;
        for (int j : cand) {
            if (filled >= fill) break;
            go[j] = 1;
            filled += (double)S->table[j]->r->ncells;
        }
    } else {
        for (int j : cand) go[j] = 1;
    }
    // every job's waves and phases for the segment (host FP64, each job's own slice of the
    // tables): on up to 8 host threads, the device waiting on them
#pragma omp parallel for schedule(dynamic, 4) reduction(+ : cells) num_threads(8) if (nj > 8)
    for (int j = 0; j < nj; j++) {
        ServerJob *J = S->table[j];
        const bool on = go[j] != 0;
        const int a = on ? (int)std::min<int64_t>(K, J->left) : 0;
        S->h_seg[j] = {a, J->r ? J->r->p : 0};
        if (a) {
            pxfdtd_host::sim_tables(J->sim, pxfdtd_host::sim_steps(J->sim), a, J->slots, S->h_waves + J->wave_off,
                                    S->h_phases + J->phase_off);
            cells += (double)a * J->r->ncells;
        }
    }
    // each class's workgroups of the jobs that step, in table order
    for (size_t q = 0; q < S->classes.size(); q++) {
        const pxfdtd_server::Class &c = S->classes[q];
        int n = 0;
        for (int k = c.first; k < c.first + c.count; k++)
            if (go[S->all_blocks[k].job]) S->h_blocks[c.first + n++] = S->all_blocks[k];
        S->h_counts[q] = n;
    }
    hipStream_t st = S->stream;
    HIPCHK(hipEventRecord(S->seg_start, st));
    if (!S->all_blocks.empty())
        HIPCHK(hipMemcpyAsync(S->d_blocks.p, S->h_blocks, S->all_blocks.size() * sizeof(BatchBlock),
                              hipMemcpyHostToDevice, st));
    if (!S->classes.empty())
        HIPCHK(hipMemcpyAsync(S->d_counts.p, S->h_counts, S->classes.size() * sizeof(int32_t), hipMemcpyHostToDevice, st));
    if (S->cap_waves) HIPCHK(hipMemcpyAsync(S->d_waves.p, S->h_waves, S->cap_waves * 8, hipMemcpyHostToDevice, st));
    if (S->cap_phases) HIPCHK(hipMemcpyAsync(S->d_phases.p, S->h_phases, S->cap_phases * 8, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(S->d_seg.p, S->h_seg, (size_t)nj * sizeof(JobSeg), hipMemcpyHostToDevice, st));
    S->seg_t0 = std::chrono::steady_clock::now();
    S->seg_cells = cells;
    const double t_launch = server_now(S);
    *prepare = t_launch - t_begin;
    *gap = S->t_seg_end >= 0 ? t_launch - S->t_seg_end : 0.0;
    if (S->exec)
        HIPCHK(hipGraphLaunch(S->exec, st));
    else
        TRY(launch_segment(S));
    HIPCHK(hipEventRecord(S->seg_stop, st));
    return PXFDTD_OK;
}

static int server_segment_end(pxfdtd_server *S, double *busy, double *device) {
    HIPCHK(hipStreamSynchronize(S->stream));
    *busy = std::chrono::duration<double>(std::chrono::steady_clock::now() - S->seg_t0).count();
    S->t_seg_end = server_now(S);
    float ms = 0.0f;
    HIPCHK(hipEventElapsedTime(&ms, S->seg_start, S->seg_stop));
    *device = 1e-3 * (double)ms;
    const int nj = (int)S->table.size();
    for (int j = 0; j < nj; j++) {
        ServerJob *J = S->table[j];
        const int a = S->h_seg[j].active;
        if (!a) continue;
        J->r->p = (S->h_seg[j].p0 + a) & 1;
        J->left -= a;
        pxfdtd_host::sim_advanced(J->sim, a);
    }
    return PXFDTD_OK;
}

static void server_finish(pxfdtd_server *S, ServerJob *J, int32_t state, const std::string &why) {
    J->state = state;
    J->error = why;
    J->t_end = server_now(S);
    if (state == PXFDTD_JOB_DONE) S->stats.completed++;
    if (state == PXFDTD_JOB_FAILED) S->stats.failed++;
    S->changed.notify_all();
}

// The loop. Each turn launches a segment of the running jobs and, while the device steps it,
// puts the newly admitted jobs on the device and brings back the results of the jobs that
// finished in the previous segment; after the segment it frees those jobs' memory. Newly
// admitted jobs join at the next turn (the tables and the graph are rebuilt then).
static void server_loop(pxfdtd_server *S) {
    (void)hipSetDevice(S->dev->ordinal);
    std::vector<ServerJob *> fetch;  // finished in the last segment, results not yet back
    bool dirty = false;
    std::unique_lock<std::mutex> lk(S->mu);
    for (;;) {
        // admissions: in arrival order while the budget and the job count allow
        std::vector<ServerJob *> admit;
        while (!S->stop && !S->queue.empty() &&
               (int)(S->running.size() + admit.size() + fetch.size()) < S->cfg.max_jobs) {
            // arrival order, or (shortest first) the smallest queued job
            auto at = S->queue.begin();
            if (S->cfg.order == 1)
                at = std::min_element(S->queue.begin(), S->queue.end(),
                                      [](const ServerJob *a, const ServerJob *b) { return a->work < b->work; });
            ServerJob *J = *at;
            if (S->used + J->foot > (size_t)S->cfg.hbm_budget) break;
            S->queue.erase(at);
            J->state = PXFDTD_JOB_RUNNING;
            J->t_start = server_now(S);
            S->used += J->foot;
            admit.push_back(J);
        }
        if (S->stop && admit.empty() && fetch.empty()) {
            for (ServerJob *J : S->running) {
                lk.unlock();
                (void)pxfdtd_host::sim_detach(J->sim);
                lk.lock();
                S->used -= J->foot;
                server_finish(S, J, PXFDTD_JOB_CANCELLED, "the server stopped");
            }
            S->running.clear();
            break;
        }
        if (S->running.empty() && admit.empty() && fetch.empty()) {
            S->t_seg_end = -1;
            S->wake.wait(lk);
            continue;
        }
        lk.unlock();
        int rc = PXFDTD_OK;
        bool launched = false;
        double t_rebuild = 0, t_admit = 0, t_retire = 0, t_busy = 0, t_prepare = 0, t_gap = 0, t_device = 0;
        if (!S->running.empty() && !S->stop) {
            if (dirty) {
                const double tb = server_now(S);
                rc = server_rebuild(S);
                t_rebuild = server_now(S) - tb;
                dirty = false;
            }
            if (rc == PXFDTD_OK) rc = server_segment_begin(S, &t_prepare, &t_gap);
            launched = rc == PXFDTD_OK;
        }
        const std::string why = rc ? pxfdtd_last_error() : "";
        // while the segment runs: new jobs onto the device, finished ones' results back
        // (each job's set-up and results are its own: up to 8 host threads take them at once,
        // on the jobs' own streams)
        std::vector<std::pair<ServerJob *, std::string>> failed;
        std::vector<ServerJob *> ready;
        const double ta = server_now(S);
        {
            const int na = (int)admit.size();
            std::vector<int> arc(na, PXFDTD_OK);
            std::vector<std::string> aerr(na);
#pragma omp parallel for schedule(dynamic, 1) num_threads(std::max(1, std::min(8, na))) if (na > 1)
            for (int q = 0; q < na; q++) {
                (void)hipSetDevice(S->dev->ordinal);
                arc[q] = server_admit(S, admit[q]);
                if (arc[q] != PXFDTD_OK) {
                    aerr[q] = pxfdtd_last_error();
                    (void)pxfdtd_host::sim_detach(admit[q]->sim);
                }
            }
            for (int q = 0; q < na; q++) {
                if (arc[q] != PXFDTD_OK) failed.push_back({admit[q], aerr[q]});
                else ready.push_back(admit[q]);
            }
        }
        t_admit = server_now(S) - ta;
        const double tf = server_now(S);
        std::vector<std::pair<ServerJob *, std::string>> fetched(fetch.size());
        {
            const int nf = (int)fetch.size();
#pragma omp parallel for schedule(dynamic, 1) num_threads(std::max(1, std::min(8, nf))) if (nf > 1)
            for (int q = 0; q < nf; q++) {
                (void)hipSetDevice(S->dev->ordinal);
                const int rc3 = pxfdtd_host::sim_fetch(fetch[q]->sim);
                fetched[q] = {fetch[q], rc3 ? pxfdtd_last_error() : ""};
            }
        }
        int rc4 = PXFDTD_OK;
        if (launched) rc4 = server_segment_end(S, &t_busy, &t_device);
        const std::string why4 = rc4 ? pxfdtd_last_error() : "";
        for (auto &f : fetched) pxfdtd_host::sim_release(f.first->sim);
        t_retire = server_now(S) - tf;
        lk.lock();
        S->stats.rebuild_seconds += t_rebuild;
        S->stats.prepare_seconds += t_prepare;
        S->stats.gap_seconds += t_gap;
        S->stats.admit_seconds += t_admit;
        S->stats.retire_seconds += t_retire;
        if (launched && rc4 == PXFDTD_OK) {
            S->stats.busy_seconds += t_busy;
            S->stats.device_seconds += t_device;
            S->stats.segments++;
            S->stats.cell_updates += S->seg_cells;
        }
        for (auto &f : failed) {
            S->used -= f.first->foot;
            server_finish(S, f.first, PXFDTD_JOB_FAILED, f.second);
        }
        for (auto &f : fetched) {
            f.first->r = nullptr;
            S->used -= f.first->foot;
            if (!f.second.empty())
                server_finish(S, f.first, PXFDTD_JOB_FAILED, f.second);
            else if (f.first->left > 0)
                server_finish(S, f.first, PXFDTD_JOB_CANCELLED, "cancelled");
            else
                server_finish(S, f.first, PXFDTD_JOB_DONE, "");
        }
        fetch.clear();
        const bool device_error = rc != PXFDTD_OK || rc4 != PXFDTD_OK;
        // the jobs this segment finished (or all of them, after a device error) go to fetch
        for (auto it = S->running.begin(); it != S->running.end();) {
            ServerJob *J = *it;
            if (!device_error && J->left > 0 && !J->cancel.load()) {
                ++it;
                continue;
            }
            if (device_error) {
                lk.unlock();
                (void)pxfdtd_host::sim_detach(J->sim);
                lk.lock();
                J->r = nullptr;
                S->used -= J->foot;
                server_finish(S, J, PXFDTD_JOB_FAILED, rc ? why : why4);
            } else {
                fetch.push_back(J);
            }
            it = S->running.erase(it);
        }
        if (device_error) {
            S->table.clear();
            lk.unlock();
            drop_graph(S);
            lk.lock();
        }
        for (ServerJob *J : ready) {
            S->running.push_back(J);
            S->stats.admitted++;
            dirty = true;
        }
        S->stats.hbm_in_use = (int64_t)S->used;
        S->stats.captures = S->captures;
        S->stats.peak_jobs = std::max<int64_t>(S->stats.peak_jobs, (int64_t)S->running.size());
    }
}

extern "C" int pxfdtd_server_create(pxfdtd_device *dev, const pxfdtd_server_config *cfg, pxfdtd_server **out) {
    HIPCHK(hipSetDevice(dev->ordinal));
    auto S = std::make_unique<pxfdtd_server>();
    S->dev = dev;
    std::memset(&S->cfg, 0, sizeof S->cfg);
    if (cfg) S->cfg = *cfg;
    std::memset(&S->stats, 0, sizeof S->stats);
    if (S->cfg.segment < 0 || S->cfg.max_jobs < 0 || S->cfg.hbm_budget < 0 || S->cfg.autotune < 0)
        return px_fail(PXFDTD_ERR_INVALID, "a server setting is negative");
    if (S->cfg.order < 0 || S->cfg.order > 1 || S->cfg.fill_cells < 0)
        return px_fail(PXFDTD_ERR_INVALID, "a server's order is 0 or 1, and its fill a cell count");
    if (!S->cfg.segment) S->cfg.segment = 32;
    if (S->cfg.segment > 4096) return px_fail(PXFDTD_ERR_INVALID, "segments of at most 4096 steps");
    if (!S->cfg.max_jobs) S->cfg.max_jobs = 256;
    if (!S->cfg.hbm_budget) {
        size_t free_b = 0, total = 0;
        HIPCHK(hipMemGetInfo(&free_b, &total));
        S->cfg.hbm_budget = (int64_t)(0.9 * (double)free_b);
    }
    HIPCHK(hipStreamCreate(&S->stream));
    HIPCHK(hipEventCreate(&S->seg_start));
    HIPCHK(hipEventCreate(&S->seg_stop));
    S->t0 = std::chrono::steady_clock::now();
    pxfdtd_server *raw = S.release();
    raw->loop = std::thread(server_loop, raw);
    *out = raw;
    return PXFDTD_OK;
}

extern "C" void pxfdtd_server_destroy(pxfdtd_server *S) {
    if (!S) return;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        S->stop = true;
        for (ServerJob *J : S->queue) server_finish(S, J, PXFDTD_JOB_CANCELLED, "the server stopped");
        S->queue.clear();
    }
    S->wake.notify_all();
    if (S->loop.joinable()) S->loop.join();
    drop_graph(S);
    if (S->h_waves) (void)hipHostFree(S->h_waves);
    if (S->h_phases) (void)hipHostFree(S->h_phases);
    if (S->h_seg) (void)hipHostFree(S->h_seg);
    if (S->h_blocks) (void)hipHostFree(S->h_blocks);
    if (S->h_counts) (void)hipHostFree(S->h_counts);
    hipStream_t st = S->stream;
    hipEvent_t e0 = S->seg_start, e1 = S->seg_stop;
    delete S;  // device tables free themselves
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (st) (void)hipStreamDestroy(st);
}

extern "C" int pxfdtd_server_submit(pxfdtd_server *S, pxfdtd_simulation *sim, int64_t steps, int64_t *job) {
    if (steps < 0) return px_fail(PXFDTD_ERR_INVALID, "a negative step count");
    if (pxfdtd_simulation_resident(sim))
        return px_fail(PXFDTD_ERR_INVALID, "hand the server a simulation with no device: it sets one on admission");
    TRY(pxfdtd_host::sim_servable(sim));
    size_t foot = 0;
    TRY(pxfdtd_host::sim_footprint(sim, PXFDTD_F32, &foot));
    std::lock_guard<std::mutex> lk(S->mu);
    if (S->stop) return px_fail(PXFDTD_ERR_INVALID, "the server is stopping");
    if (foot > (size_t)S->cfg.hbm_budget)
        return px_fail(PXFDTD_ERR_MEMORY, "the job needs %zu bytes on the device, more than the server's budget of %lld",
                       foot, (long long)S->cfg.hbm_budget);
    for (auto &kv : S->jobs)
        if (kv.second->sim == sim && kv.second->state <= PXFDTD_JOB_RUNNING)
            return px_fail(PXFDTD_ERR_INVALID, "the simulation is already a queued or running job");
    auto J = std::make_unique<ServerJob>();
    J->id = S->next_id++;
    J->sim = sim;
    J->steps = J->left = steps;
    J->foot = foot;
    J->work = (double)steps * (double)pxfdtd_simulation_cells(sim);
    J->t_submit = server_now(S);
    ServerJob *raw = J.get();
    S->jobs[raw->id] = std::move(J);
    S->queue.push_back(raw);
    S->stats.submitted++;
    *job = raw->id;
    S->wake.notify_all();
    return PXFDTD_OK;
}

extern "C" int pxfdtd_server_state(pxfdtd_server *S, int64_t job, int32_t *state) {
    std::lock_guard<std::mutex> lk(S->mu);
    auto it = S->jobs.find(job);
    if (it == S->jobs.end()) return px_fail(PXFDTD_ERR_INVALID, "no job %lld", (long long)job);
    *state = it->second->state;
    return PXFDTD_OK;
}

extern "C" int pxfdtd_server_wait(pxfdtd_server *S, int64_t job, double timeout_seconds) {
    std::unique_lock<std::mutex> lk(S->mu);
    auto it = S->jobs.find(job);
    if (it == S->jobs.end()) return px_fail(PXFDTD_ERR_INVALID, "no job %lld", (long long)job);
    ServerJob *J = it->second.get();
    auto done = [&] { return J->state >= PXFDTD_JOB_DONE; };
    if (timeout_seconds > 0) {
        if (!S->changed.wait_for(lk, std::chrono::duration<double>(timeout_seconds), done))
            return px_fail(PXFDTD_ERR_TIMEOUT, "job %lld is not done after %g s", (long long)job, timeout_seconds);
    } else {
        S->changed.wait(lk, done);
    }
    if (J->state != PXFDTD_JOB_DONE)
        return px_fail(J->state == PXFDTD_JOB_FAILED ? PXFDTD_ERR_DEVICE : PXFDTD_ERR_INVALID, "job %lld %s: %s",
                       (long long)job, J->state == PXFDTD_JOB_FAILED ? "failed" : "was cancelled", J->error.c_str());
    return PXFDTD_OK;
}

extern "C" int pxfdtd_server_cancel(pxfdtd_server *S, int64_t job) {
    std::lock_guard<std::mutex> lk(S->mu);
    auto it = S->jobs.find(job);
    if (it == S->jobs.end()) return px_fail(PXFDTD_ERR_INVALID, "no job %lld", (long long)job);
    ServerJob *J = it->second.get();
    if (J->state == PXFDTD_JOB_QUEUED) {
        for (auto q = S->queue.begin(); q != S->queue.end(); ++q)
            if (*q == J) {
                S->queue.erase(q);
                break;
            }
        server_finish(S, J, PXFDTD_JOB_CANCELLED, "cancelled before it was admitted");
    } else if (J->state == PXFDTD_JOB_RUNNING) {
        J->cancel.store(true);
        S->wake.notify_all();
    }
    return PXFDTD_OK;
}

extern "C" int pxfdtd_server_times(pxfdtd_server *S, int64_t job, double *submitted, double *started,
                                   double *finished) {
    std::lock_guard<std::mutex> lk(S->mu);
    auto it = S->jobs.find(job);
    if (it == S->jobs.end()) return px_fail(PXFDTD_ERR_INVALID, "no job %lld", (long long)job);
    *submitted = it->second->t_submit;
    *started = it->second->t_start;
    *finished = it->second->t_end;
    return PXFDTD_OK;
}

extern "C" int pxfdtd_server_get_stats(pxfdtd_server *S, pxfdtd_server_stats *out) {
    std::lock_guard<std::mutex> lk(S->mu);
    *out = S->stats;
    return PXFDTD_OK;
}
