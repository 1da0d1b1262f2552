// Level 1 of the C ABI (include/pxfdtd.h): the device seam. A resident problem's buffers on
// one MI355X, its tables, and the stepping loop: the fused Yee + CPML kernel once per step,
// the taps that copy probe and transform values out, and every few steps the FP64 transform
// accumulation. It replaces photonoxide's gpu::Resident (src/fdtd/gpu/resident.rs:43-870):
// new (:154-293), upload/download (:378-413), step (:430-539) and record (:846-869).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <functional>
#include <map>
#include <mutex>
#include <set>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/pxfdtd.h"
#include "device.hpp"
#include "kernels.hpp"
#include "march64.hpp"
#include "march2d.hpp"
#include "march2s.hpp"
#include "step_launch.hpp"

using namespace pxfdtd;

// ------------------------------------------------------------------------------ errors
static thread_local std::string g_error;

extern "C" const char *pxfdtd_last_error(void) { return g_error.c_str(); }

int px_fail(int code, const char *fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_error = buf;
    return code;
}


// ------------------------------------------------------------------------------ device

extern "C" int pxfdtd_device_open(int ordinal, pxfdtd_device **out) {
    int count = 0;
    HIPCHK(hipGetDeviceCount(&count));
    if (ordinal < 0 || ordinal >= count)
        return px_fail(PXFDTD_ERR_INVALID, "no HIP device %d (%d present)", ordinal, count);
    auto *d = new pxfdtd_device();
    d->ordinal = ordinal;
    hipError_t e = hipGetDeviceProperties(&d->prop, ordinal);
    if (e != hipSuccess) {
        delete d;
        return px_fail(PXFDTD_ERR_DEVICE, "hipGetDeviceProperties: %s", hipGetErrorString(e));
    }
    const int rc = flux_scratch_setup(ordinal, &d->scratch);
    if (rc != PXFDTD_OK) {
        delete d;
        return rc;
    }
    *out = d;
    return PXFDTD_OK;
}

static size_t block_held(int ordinal);
extern "C" int pxfdtd_device_query(const pxfdtd_device *d, pxfdtd_device_info *info) {
    memset(info, 0, sizeof *info);
    snprintf(info->name, sizeof info->name, "%s", d->prop.name);
    snprintf(info->arch, sizeof info->arch, "%s", d->prop.gcnArchName);
    info->ordinal = d->ordinal;
    info->compute_units = d->prop.multiProcessorCount;
    info->hbm_bytes = (int64_t)d->prop.totalGlobalMem;
    size_t free_b = 0, total = 0;
    HIPCHK(hipSetDevice(d->ordinal));
    HIPCHK(hipMemGetInfo(&free_b, &total));
    info->free_bytes = (int64_t)(free_b + block_held(d->ordinal));  // cached blocks are free to residents
    return PXFDTD_OK;
}

extern "C" void pxfdtd_device_close(pxfdtd_device *d) { delete d; }

// Streams and events of residents that are gone, kept per device for the next ones: creating and
// destroying them costs more than the set-up of a small job (a server sets up and retires
// hundreds). Held until the process ends.
struct HipPool {
    std::mutex mu;
    std::map<int, std::vector<hipStream_t>> streams;
    std::map<std::pair<int, bool>, std::vector<hipEvent_t>> events;  // (ordinal, timing)
};
static HipPool &hip_pool() {
    static HipPool *p = new HipPool();  // never destroyed: the handles live as long as the process
    return *p;
}
static hipError_t take_stream(int ordinal, hipStream_t *s) {
    {
        HipPool &P = hip_pool();
        std::lock_guard<std::mutex> g(P.mu);
        auto &v = P.streams[ordinal];
        if (!v.empty()) {
            *s = v.back();
            v.pop_back();
            return hipSuccess;
        }
    }
    return hipStreamCreate(s);
}
static void give_stream(int ordinal, hipStream_t s) {  // s has no work left
    if (!s) return;
    HipPool &P = hip_pool();
    std::lock_guard<std::mutex> g(P.mu);
    P.streams[ordinal].push_back(s);
}
static hipError_t take_event(int ordinal, bool timing, hipEvent_t *e) {
    {
        HipPool &P = hip_pool();
        std::lock_guard<std::mutex> g(P.mu);
        auto &v = P.events[{ordinal, timing}];
        if (!v.empty()) {
            *e = v.back();
            v.pop_back();
            return hipSuccess;
        }
    }
    return hipEventCreateWithFlags(e, timing ? hipEventDefault : hipEventDisableTiming);
}
static void give_event(int ordinal, bool timing, hipEvent_t e) {
    if (!e) return;
    HipPool &P = hip_pool();
    std::lock_guard<std::mutex> g(P.mu);
    P.events[{ordinal, timing}].push_back(e);
}

// ------------------------------------------------------------------------------ buffers
// Device blocks of residents that are gone, kept per device for the next ones: a hipMalloc and
// hipFree pair (the free waits for the device) costs more than the set-up of a small job, and a
// server sets up and retires hundreds. Only a resident's destruction gives blocks back, after
// its streams are synchronized, so no work on the device still reads them; at most 8 GiB a
// device is held, and an allocation that fails returns them all first.
constexpr size_t kBlockCacheBytes = size_t(8) << 30;
struct BlockCache {
    std::mutex mu;
    std::map<int, std::multimap<size_t, void *>> free;  // ordinal -> block size -> block
    std::map<int, size_t> held;
};
static BlockCache &block_cache() {
    static BlockCache *c = new BlockCache();  // never destroyed: the blocks live as long as the process
    return *c;
}
static thread_local bool tl_give_blocks = false;  // set while a resident's buffers are released
static size_t block_size(size_t n) {               // 64 KiB steps, 2 MiB above 16 MiB
    const size_t q = n >= (size_t(16) << 20) ? (size_t(2) << 20) : (size_t(64) << 10);
    return (n + q - 1) / q * q;
}
static bool block_take(int ordinal, size_t n, void **p, size_t *size) {
    BlockCache &C = block_cache();
    std::lock_guard<std::mutex> g(C.mu);
    auto &m = C.free[ordinal];
    auto it = m.lower_bound(n);
    if (it == m.end() || it->first > n + n / 4 + (size_t(64) << 10)) return false;
    *p = it->second;
    *size = it->first;
    C.held[ordinal] -= it->first;
    m.erase(it);
    return true;
}
static bool block_give(int ordinal, void *p, size_t size) {
    BlockCache &C = block_cache();
    std::lock_guard<std::mutex> g(C.mu);
    if (C.held[ordinal] + size > kBlockCacheBytes) return false;
    C.free[ordinal].insert({size, p});
    C.held[ordinal] += size;
    return true;
}
static size_t block_held(int ordinal) {
    BlockCache &C = block_cache();
    std::lock_guard<std::mutex> g(C.mu);
    return C.held[ordinal];
}
static void block_flush(int ordinal) {
    std::multimap<size_t, void *> m;
    {
        BlockCache &C = block_cache();
        std::lock_guard<std::mutex> g(C.mu);
        m.swap(C.free[ordinal]);
        C.held[ordinal] = 0;
    }
    for (auto &b : m) (void)hipFree(b.second);
}

struct Buf {
    void *p = nullptr;
    size_t bytes = 0;   // asked for
    size_t block = 0;   // the block's own size
    int ordinal = 0;
    Buf() = default;
    Buf(const Buf &) = delete;
    Buf &operator=(const Buf &) = delete;
    ~Buf() { release(); }
    void release() {
        if (p && !(tl_give_blocks && block_give(ordinal, p, block))) (void)hipFree(p);
        p = nullptr;
        bytes = block = 0;
    }
    int alloc(size_t n) {
        release();
        if (n == 0) return PXFDTD_OK;
        (void)hipGetDevice(&ordinal);
        if (block_take(ordinal, n, &p, &block)) {
            bytes = n;
            return PXFDTD_OK;
        }
        block = block_size(n);
        hipError_t e = hipMalloc(&p, block);
        if (e != hipSuccess) {  // the cached blocks back first
            (void)hipGetLastError();
            block_flush(ordinal);
            e = hipMalloc(&p, block);
        }
        if (e != hipSuccess) {
            p = nullptr;
            block = 0;
            return px_fail(PXFDTD_ERR_MEMORY, "hipMalloc of %zu bytes: %s", n, hipGetErrorString(e));
        }
        bytes = n;
        return PXFDTD_OK;
    }
    // grows to at least n bytes, keeping nothing
    void swap(Buf &o) {
        std::swap(p, o.p), std::swap(bytes, o.bytes), std::swap(block, o.block), std::swap(ordinal, o.ordinal);
    }
    int reserve(size_t n) { return n <= bytes ? PXFDTD_OK : alloc(n); }
    template <typename T> T *as() const { return reinterpret_cast<T *>(p); }
};


struct SlabDesc {
    int field, comp, axis, side, from, count;
    long long offset, len;
};

struct pxfdtd_resident {
    pxfdtd_device *dev = nullptr;
    hipStream_t stream = nullptr;
    // a second stream: the CPML-slab launch runs beside the lean one, forked and joined by events
    hipStream_t side = nullptr;
    hipEvent_t fork_ev = nullptr, join_ev = nullptr;
    int concurrent = 0;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    int prec = PXFDTD_F32;
    size_t esz = 4;
    int nx = 0, ny = 0, nz = 0, pitch = 0;
    long long plane = 0, nval = 0, ncells = 0;
    int periodic[3] = {0, 0, 0}, low[3] = {0, 0, 0}, high[3] = {0, 0, 0};
    double dt = 0, step[3] = {0, 0, 0};
    Buf fields[2];  // per copy: Ex, Ey, Ez, Hx, Hy, Hz, nval values each
    int mode = 0;
    Buf ids, table, cacb, grpb, cold;
    bool has_grp = false;
    Buf coef;  // fe, fh per axis, then the CPML's b and a
    int fe_off[3] = {0, 0, 0}, fh_off[3] = {0, 0, 0}, cpml_off[3][2][2];
    int ncpml = 0;  // CPML coefficients in coef after the difference factors
    Buf psi[2];
    int psi_off[2][3][3][2];
    template <typename T> T *field(int copy, int q) const { return fields[copy].as<T>() + (size_t)q * nval; }
    long long psi_len = 0;
    std::vector<SlabDesc> slabs;
    // media
    Buf groups, m_before, m_cur, m_prev;
    int ngroups = 0, max_poles = 0;
    // the values in a medium, updated by media_pass after the step kernel: c * nval + r, the
    // group, cb
    Buf mvals, mgrp, mcb;
    long long nmvals = 0;
    // injections
    std::vector<pxfdtd_injection> inj_host;
    // plane waves: their auxiliary grids (FP64, E then H~ per wave), coefficients, and the
    // couplings of their boxes as injections whose emitter < 0 names the incident value
    // aux[-1 - emitter]; a coupling applies after every source on its value
    std::vector<PlaneDesc> planes_h;
    std::vector<pxfdtd_injection> coup_host;
    std::vector<char> coup_live;  // whether a coupling's incident component can be nonzero
    Buf plane_desc, aux, aux_coef;
    long long aux_len = 0, aux_max = 0;
    int emitters = 0;
    Buf inj_off, inj;
    bool have_inj = false;
    Buf waves;
    // taps, probes and transforms
    int nprobes = 0;
    long long nvals = 0, ntaps = 0;
    Buf tap_index, tap_which, tap_fields[2];
    std::vector<int64_t> tap_index_h;  // the same, on the host
    std::vector<int8_t> tap_which_h;
    Buf staging;
    int block_steps = 32;
    Buf series_of, series_info, sums, phases;
    long long sums_len = 0;
    int slots = 0;
    Buf probe_rec;
    long long probe_rows = 0, probe_cap = 0;
    // stepping
    int p = 0;
    int kernel = 3;  // 0: LDS tiles (64 x ty); 1, 2: 62-wide march (with, without prefetch); 3: aligned march; 4: 2D march
    // the 2D march's field sets with values or sources: 1 TE (Ex, Ey, Hz), 2 TM (Ez, Hx, Hy)
    int fields_set = 3, inj_set = 0;
    int ty = 4, kz = 16;
    double last_ms = 0;
    // register march: work units (strip x, strip y, first plane, end plane), those touching
    // no CPML slab first, then those that do
    Buf units;
    std::vector<int4> units_h;  // the same, on the host (the server batches them)
    int n_lean = 0, n_x = 0, n_slab = 0;  // the march's units: no slab, x slabs only, others
    int slab_rows = 2;                      // rows per unit of the others (the aligned march)
    int ordinal = 0;                        // the device's, for its pooled streams and events
    bool batched = false;                   // stepped by the server's batched kernels
    bool units_dirty = true;
    bool quiet = false;  // steps record no probe or transform values (the adjoint's recomputation)
    // launched on the stream after each step's kernels, with the step's number in the call (the
    // adjoint's record of each recomputed step); none when empty
    std::function<void(hipStream_t, uint32_t)> after_step;
    // the two-step march (kind 5): pairs of steps in one pass where the problem allows, with
    // its own units (strips of 60 x ty2 rows, chunks of kz2 planes); single steps run kind 3
    int pair = 0, ty2 = 2, kz2 = 16;
    Buf units2, inj2_off, inj2, tap2_off, tap2;
    int n2[5] = {0, 0, 0, 0, 0}, tiles2_x = 0;  // units: lean, x, y, z slabs only, the rest
    PairStreams ps = {{nullptr, nullptr, nullptr, nullptr}, nullptr, {nullptr, nullptr, nullptr, nullptr}};
    bool units2_dirty = true, have_tap2 = false;
    // per-launch timing of the fused kernel, when asked for
    bool profile = false;
    std::vector<hipEvent_t> kev;
    double kernel_ms = 0;
};

struct TapSpec;
static std::map<const pxfdtd_resident *, TapSpec> &specs();
static std::mutex &specs_mutex();

// ---------------------------------------------------------------------- problem checks
static long long slab_volume(const pxfdtd_problem *p, int axis, int count) {
    long long n[3] = {p->nx, p->ny, p->nz};
    n[axis] = count;
    return n[0] * n[1] * n[2];
}

static int check_problem(const pxfdtd_problem *p) {
    if (p->nx <= 0 || p->ny <= 0 || p->nz <= 0)
        return px_fail(PXFDTD_ERR_INVALID, "the grid needs cells, got %d x %d x %d", p->nx, p->ny, p->nz);
    if (p->precision != PXFDTD_F32 && p->precision != PXFDTD_F64)
        return px_fail(PXFDTD_ERR_INVALID, "precision must be PXFDTD_F32 or PXFDTD_F64");
    if (!(p->dt > 0) || !(p->step[0] > 0) || !(p->step[1] > 0) || !(p->step[2] > 0))
        return px_fail(PXFDTD_ERR_INVALID, "dt and the steps must be positive");
    if (!p->ca || !p->cb) return px_fail(PXFDTD_ERR_INVALID, "ca and cb are required");
    const int n[3] = {p->nx, p->ny, p->nz};
    for (int a = 0; a < 3; a++) {
        if (!p->factors_e[a] || !p->factors_h[a])
            return px_fail(PXFDTD_ERR_INVALID, "the difference factors are required");
        if (p->periodic[a] && (p->pml_low[a] || p->pml_high[a]))
            return px_fail(PXFDTD_ERR_INVALID, "a periodic axis has no CPML");
        if (p->pml_low[a] < 0 || p->pml_high[a] < 0 || p->pml_low[a] + p->pml_high[a] >= n[a])
            return px_fail(PXFDTD_ERR_INVALID, "CPMLs of %d and %d cells leave nothing of %d cells",
                           p->pml_low[a], p->pml_high[a], n[a]);
        const int cnt[2] = {p->pml_low[a], p->pml_high[a]};
        for (int s = 0; s < 2; s++)
            for (int f = 0; f < 2; f++)
                if (cnt[s] && (!p->cpml_b[a][s][f] || !p->cpml_a[a][s][f]))
                    return px_fail(PXFDTD_ERR_INVALID, "CPML coefficients missing for axis %d", a);
    }
    if (p->nmedia > 127) return px_fail(PXFDTD_ERR_UNSUPPORTED, "at most 127 dispersive media");
    for (int m = 0; m < p->nmedia; m++)
        if (p->media[m].poles < 1 || p->media[m].poles > 4)
            return px_fail(PXFDTD_ERR_UNSUPPORTED, "a dispersive medium needs 1 to 4 poles here");
    return PXFDTD_OK;
}

// The 2D march's rows per chunk alone on the device: a wavefront pays one memory round trip
// per row, so a small grid takes short chunks (more wavefronts in flight, at the cost of the
// row below each chunk recomputed); a large one takes up to 16.
static int default_rows_2d(int nx, int ny) {
    const long long strips = (nx + SW - 1) / SW;
    int rows = 16;
    while (rows > 2 && strips * ((ny + rows - 1) / rows) < 8192) rows /= 2;
    return std::max(1, std::min(ny, rows));
}

static int pitch_of(int nx) { return (nx + 63) / 64 * 64; }

// The CPML's psi values per copy: each axis side's slab holds 2 fields x 2 components.
static long long psi_values(const pxfdtd_problem *p) {
    long long n = 0;
    for (int a = 0; a < 3; a++) n += 4 * slab_volume(p, a, p->pml_low[a] + p->pml_high[a]);
    return n;
}

// Whether step kernel `kind` addresses the buffers of a problem with nval values per field
// component and psi_len psi values, esz bytes each: kinds 1 and 2 read each component through a
// buffer resource of 32-bit byte offsets, with a wall's neighbour at OUTSIDE (2^31) beyond it;
// kinds 3 and 4 read a field's three components, and psi, the same way; kind 0 indexes by
// element (int) from plain pointers. (Every kind needs nval < 2^30 and psi_len < 2^31, checked
// at creation.)
static bool kernel_fits(int kind, long long nval, long long psi_len, long long esz) {
    if (kind == 1 || kind == 2) return nval * esz < (1LL << 31);
    if (kind == 3 || kind == 4) return 3 * nval * esz < (1LL << 31) && psi_len * esz < (1LL << 31);
    return true;
}

extern "C" int pxfdtd_resident_footprint(const pxfdtd_problem *p, size_t *bytes) {
    TRY(check_problem(p));
    const size_t esz = p->precision == PXFDTD_F64 ? 8 : 4;
    const size_t nval = (size_t)pitch_of(p->nx) * p->ny * p->nz;
    size_t b = 12 * nval * esz;  // two copies of E and H
    b += 3 * nval;               // ids (the table mode; the array mode adds more, see create)
    size_t psi = 0;
    for (int a = 0; a < 3; a++) psi += 4 * (size_t)slab_volume(p, a, p->pml_low[a] + p->pml_high[a]);
    b += 2 * psi * esz;
    int maxp = 0;
    for (int m = 0; m < p->nmedia; m++) maxp = std::max(maxp, (int)p->media[m].poles);
    if (maxp) b += 3 * nval * esz * (1 + 2 * (size_t)maxp);
    *bytes = b + (1u << 20);
    return PXFDTD_OK;
}

// a component between the host layout ((k ny + j) nx + i) and the device's pitched one, in
// parallel over rows
template <typename T>
static void to_pitched(const pxfdtd_resident *r, const double *src, T *dst) {
    const int nx = r->nx, ny = r->ny, nz = r->nz;
    const long long rows = (long long)ny * nz;
#pragma omp parallel for schedule(static) if (r->ncells > (1 << 18))
    for (long long row = 0; row < rows; row++) {
        const long long k = row / ny, j = row % ny;
        const double *s = src + row * nx;
        T *d = dst + k * r->plane + j * r->pitch;
        for (int i = 0; i < nx; i++) d[i] = (T)s[i];
    }
}

template <typename T>
static void from_pitched(const pxfdtd_resident *r, const T *src, double *dst) {
    const int nx = r->nx, ny = r->ny, nz = r->nz;
    const long long rows = (long long)ny * nz;
#pragma omp parallel for schedule(static) if (r->ncells > (1 << 18))
    for (long long row = 0; row < rows; row++) {
        const long long k = row / ny, j = row % ny;
        const T *s = src + k * r->plane + j * r->pitch;
        double *d = dst + row * nx;
        for (int i = 0; i < nx; i++) d[i] = (double)s[i];
    }
}

// whether any of n values is not zero
static bool any_nonzero(const double *v, long long n) {
    int found = 0;
#pragma omp parallel for schedule(static) reduction(| : found) if (n > (1 << 20))
    for (long long q = 0; q < n; q++) found |= v[q] != 0.0;
    return found != 0;
}

template <typename T>
static std::vector<T> converted(const double *v, size_t n) {
    std::vector<T> o(n);
    for (size_t q = 0; q < n; q++) o[q] = (T)v[q];
    return o;
}

template <typename T>
static int upload_table(Buf &b, const double *v, size_t n, hipStream_t s) {
    TRY(b.alloc(std::max<size_t>(n, 1) * sizeof(T)));
    if (!n) return PXFDTD_OK;
    std::vector<T> c = converted<T>(v, n);
    HIPCHK(hipMemcpyAsync(b.p, c.data(), n * sizeof(T), hipMemcpyHostToDevice, s));
    HIPCHK(hipStreamSynchronize(s));
    return PXFDTD_OK;
}

template <typename T>
static int build(pxfdtd_resident *r, const pxfdtd_problem *p) {
    const long long N = (long long)p->nx * p->ny * p->nz;
    const long long nval = r->nval;
    hipStream_t s = r->stream;
    for (int c = 0; c < 2; c++) {
        TRY(r->fields[c].alloc(6 * nval * sizeof(T)));
        HIPCHK(hipMemsetAsync(r->fields[c].p, 0, 6 * nval * sizeof(T), s));
    }
    // media membership per value: group + 1, 0 outside
    std::vector<int8_t> member;
    if (p->nmedia) {
        member.assign(3 * N, 0);
        for (int m = 0; m < p->nmedia; m++) {
            const pxfdtd_medium &M = p->media[m];
            for (long long v = 0; v < M.count; v++) {
                const int c = M.component[v];
                const long long q = M.index[v];
                if (c < 0 || c > 2 || q < 0 || q >= N)
                    return px_fail(PXFDTD_ERR_INVALID, "a medium's value is outside the grid");
                member[c * N + q] = (int8_t)(m + 1);
            }
        }
    }
    // E's coefficients: one uint8 material id per cell if at most 256 distinct
    // (ca, cb, medium) triples over the three components occur, else whole arrays
    struct Key {
        uint64_t v[6];
        int8_t g[3];
        bool operator<(const Key &o) const {
            for (int q = 0; q < 6; q++)
                if (v[q] != o.v[q]) return v[q] < o.v[q];
            for (int q = 0; q < 3; q++)
                if (g[q] != o.g[q]) return g[q] < o.g[q];
            return false;
        }
        bool operator==(const Key &o) const { return !(*this < o) && !(o < *this); }
    };
    std::vector<uint8_t> idv(nval, 0);
    std::vector<double> tab(256 * 6, 0.0);
    std::vector<int8_t> tgrp(256 * 3, -1);
    bool table_ok = true;
    auto key_at = [&](long long q) {
        Key key;
        for (int c = 0; c < 3; c++) {
            memcpy(&key.v[2 * c], &p->ca[c * N + q], 8);
            memcpy(&key.v[2 * c + 1], &p->cb[c * N + q], 8);
            key.g[c] = member.empty() ? (int8_t)-1 : (int8_t)(member[c * N + q] - 1);
        }
        return key;
    };
    {
        // more than 256 distinct triples among 4096 evenly spaced cells settle it without the
        // serial scan of every cell (a smooth design region has thousands)
        std::set<Key> few;
        const long long stride = std::max<long long>(1, N / 4096);
        for (long long q = 0; q < N && table_ok; q += stride) {
            few.insert(key_at(q));
            if (few.size() > 256) table_ok = false;
        }
    }
    {
        std::map<Key, int> seen;
        Key last;
        int last_id = -1;
        for (long long q = 0; q < N && table_ok; q++) {
            const Key key = key_at(q);
            int id;
            if (last_id >= 0 && key == last) {
                id = last_id;
            } else {
                auto it = seen.find(key);
                if (it == seen.end()) {
                    id = (int)seen.size();
                    if (id >= 256) { table_ok = false; break; }
                    seen.emplace(key, id);
                    for (int c = 0; c < 3; c++) {
                        tab[id * 6 + 2 * c] = p->ca[c * N + q];
                        tab[id * 6 + 2 * c + 1] = p->cb[c * N + q];
                        tgrp[id * 3 + c] = key.g[c];
                    }
                } else {
                    id = it->second;
                }
                last = key;
                last_id = id;
            }
            const long long i = q % p->nx, rest = q / p->nx;
            idv[(rest / p->ny) * r->plane + (rest % p->ny) * r->pitch + i] = (uint8_t)id;
        }
    }
    // beyond 256 triples, a lossless problem without media stores cb alone: each ca is 0 where
    // its cb is 0 and 1 elsewhere (the kernels derive it), bit for bit
    bool lossless = !table_ok && member.empty();
    if (lossless) {
        const uint64_t one = 0x3ff0000000000000ull;
        int bad = 0;
#pragma omp parallel for schedule(static) reduction(| : bad)
        for (long long q = 0; q < 3 * N; q++) {
            uint64_t a;
            memcpy(&a, &p->ca[q], 8);
            bad |= a != (p->cb[q] == 0.0 ? 0ull : one);
        }
        lossless = bad == 0;
    }
    r->mode = table_ok ? 0 : (lossless ? 2 : 1);
    if (r->mode == 2) {
        std::vector<T> cbv(3 * nval, T(0));
        for (int c = 0; c < 3; c++) to_pitched<T>(r, p->cb + c * N, cbv.data() + c * nval);
        TRY(r->cacb.alloc(cbv.size() * sizeof(T)));
        HIPCHK(hipMemcpyAsync(r->cacb.p, cbv.data(), cbv.size() * sizeof(T), hipMemcpyHostToDevice, s));
        HIPCHK(hipStreamSynchronize(s));
    } else if (r->mode == 0) {
        TRY(r->ids.alloc(nval));
        HIPCHK(hipMemcpyAsync(r->ids.p, idv.data(), nval, hipMemcpyHostToDevice, s));
        HIPCHK(hipStreamSynchronize(s));
        // the table, then the media groups as int8
        std::vector<unsigned char> tb(256 * 6 * sizeof(T) + tgrp.size());
        for (size_t q = 0; q < tab.size(); q++) {
            const T v = (T)tab[q];
            memcpy(tb.data() + q * sizeof(T), &v, sizeof(T));
        }
        memcpy(tb.data() + 256 * 6 * sizeof(T), tgrp.data(), tgrp.size());
        TRY(r->table.alloc(tb.size()));
        HIPCHK(hipMemcpyAsync(r->table.p, tb.data(), tb.size(), hipMemcpyHostToDevice, s));
        HIPCHK(hipStreamSynchronize(s));
    } else {
        std::vector<T> ab(6 * nval, T(0));
        std::vector<int8_t> g(3 * nval, -1);
        for (int c = 0; c < 3; c++) {
            to_pitched<T>(r, p->ca + c * N, ab.data() + c * nval);
            to_pitched<T>(r, p->cb + c * N, ab.data() + (3 + c) * nval);
            if (!member.empty())
                for (long long q = 0; q < N; q++) {
                    const long long i = q % p->nx, rest = q / p->nx;
                    g[c * nval + (rest / p->ny) * r->plane + (rest % p->ny) * r->pitch + i] =
                        (int8_t)(member[c * N + q] - 1);
                }
        }
        TRY(r->cacb.alloc(ab.size() * sizeof(T)));
        HIPCHK(hipMemcpyAsync(r->cacb.p, ab.data(), ab.size() * sizeof(T), hipMemcpyHostToDevice, s));
        if (!member.empty()) {
            TRY(r->grpb.alloc(g.size()));
            HIPCHK(hipMemcpyAsync(r->grpb.p, g.data(), g.size(), hipMemcpyHostToDevice, s));
        }
        HIPCHK(hipStreamSynchronize(s));
        r->has_grp = !member.empty();
    }
    const int n[3] = {p->nx, p->ny, p->nz};
    {
        std::vector<double> co;
        for (int a = 0; a < 3; a++) {
            r->fe_off[a] = (int)co.size();
            co.insert(co.end(), p->factors_e[a], p->factors_e[a] + n[a]);
            r->fh_off[a] = (int)co.size();
            co.insert(co.end(), p->factors_h[a], p->factors_h[a] + n[a]);
        }
        for (int a = 0; a < 3; a++) {
            const int cnt[2] = {p->pml_low[a], p->pml_high[a]};
            for (int sd = 0; sd < 2; sd++)
                for (int f = 0; f < 2; f++) {
                    r->cpml_off[a][sd][f] = (int)co.size();
                    if (!cnt[sd]) continue;
                    co.insert(co.end(), p->cpml_b[a][sd][f], p->cpml_b[a][sd][f] + cnt[sd]);
                    co.insert(co.end(), p->cpml_a[a][sd][f], p->cpml_a[a][sd][f] + cnt[sd]);
                }
        }
        r->ncpml = (int)co.size() - 2 * (p->nx + p->ny + p->nz);
        if (r->ncpml > 2048)
            return px_fail(PXFDTD_ERR_UNSUPPORTED, "CPMLs this thick (%d coefficients) exceed the kernel's table", r->ncpml);
        TRY(upload_table<T>(r->coef, co.data(), co.size(), s));
    }
    // psi: the reference's slab order (src/fdtd.rs:838-866)
    for (int F = 0; F < 2; F++)
        for (int c = 0; c < 3; c++)
            for (int w = 0; w < 3; w++)
                for (int sd = 0; sd < 2; sd++) r->psi_off[F][c][w][sd] = -1;
    long long off = 0;
    for (int a = 0; a < 3; a++) {
        const int cnt[2] = {p->pml_low[a], p->pml_high[a]};
        const int from[2] = {0, n[a] - p->pml_high[a]};
        for (int sd = 0; sd < 2; sd++) {
            if (!cnt[sd]) continue;
            for (int F = 0; F < 2; F++)
                for (int cc = 1; cc <= 2; cc++) {
                    const int c = (a + cc) % 3;
                    const long long len = slab_volume(p, a, cnt[sd]);
                    r->slabs.push_back({F, c, a, sd, from[sd], cnt[sd], off, len});
                    r->psi_off[F][c][a][sd] = (int)off;
                    off += len;
                }
        }
    }
    r->psi_len = off;
    if (!kernel_fits(r->kernel, r->nval, off, (long long)sizeof(T)))
        return px_fail(PXFDTD_ERR_UNSUPPORTED, "kernel %d cannot address %lld psi values", r->kernel, off);
    for (int q = 0; q < 2; q++) {
        TRY(r->psi[q].alloc(std::max<long long>(off, 1) * sizeof(T)));
        HIPCHK(hipMemsetAsync(r->psi[q].p, 0, std::max<long long>(off, 1) * sizeof(T), s));
    }
    // media: groups and dense state
    if (p->nmedia) {
        std::vector<MediumGroup> gs(p->nmedia);
        for (int m = 0; m < p->nmedia; m++) {
            memset(&gs[m], 0, sizeof(MediumGroup));
            gs[m].poles = p->media[m].poles;
            gs[m].chi = p->media[m].chi;
            memcpy(gs[m].coef, p->media[m].coef, sizeof gs[m].coef);
            r->max_poles = std::max(r->max_poles, (int)p->media[m].poles);
        }
        r->ngroups = p->nmedia;
        TRY(r->groups.alloc(gs.size() * sizeof(MediumGroup)));
        HIPCHK(hipMemcpyAsync(r->groups.p, gs.data(), gs.size() * sizeof(MediumGroup),
                              hipMemcpyHostToDevice, s));
        TRY(r->m_before.alloc(3 * nval * sizeof(T)));
        TRY(r->m_cur.alloc(3 * nval * r->max_poles * sizeof(T)));
        TRY(r->m_prev.alloc(3 * nval * r->max_poles * sizeof(T)));
        HIPCHK(hipMemsetAsync(r->m_before.p, 0, r->m_before.bytes, s));
        HIPCHK(hipMemsetAsync(r->m_cur.p, 0, r->m_cur.bytes, s));
        HIPCHK(hipMemsetAsync(r->m_prev.p, 0, r->m_prev.bytes, s));
        if (3 * nval < (1LL << 32)) {
            std::vector<uint32_t> vals;
            std::vector<int8_t> vg;
            std::vector<T> vcb;
            for (int c = 0; c < 3; c++)
                for (long long q = 0; q < N; q++) {
                    if (!member[c * N + q]) continue;
                    const long long i = q % p->nx, rest = q / p->nx;
                    vals.push_back((uint32_t)(c * nval + (rest / p->ny) * r->plane + (rest % p->ny) * r->pitch + i));
                    vg.push_back((int8_t)(member[c * N + q] - 1));
                    vcb.push_back((T)p->cb[c * N + q]);
                }
            r->nmvals = (long long)vals.size();
            if (r->nmvals) {
                TRY(r->mvals.alloc(vals.size() * sizeof(uint32_t)));
                TRY(r->mgrp.alloc(vg.size()));
                TRY(r->mcb.alloc(vcb.size() * sizeof(T)));
                HIPCHK(hipMemcpyAsync(r->mvals.p, vals.data(), vals.size() * sizeof(uint32_t), hipMemcpyHostToDevice, s));
                HIPCHK(hipMemcpyAsync(r->mgrp.p, vg.data(), vg.size(), hipMemcpyHostToDevice, s));
                HIPCHK(hipMemcpyAsync(r->mcb.p, vcb.data(), vcb.size() * sizeof(T), hipMemcpyHostToDevice, s));
                HIPCHK(hipStreamSynchronize(s));
            }
        }
    }
    HIPCHK(hipStreamSynchronize(s));
    return PXFDTD_OK;
}

extern "C" int pxfdtd_resident_create(pxfdtd_device *dev, const pxfdtd_problem *p,
                                      pxfdtd_resident **out) {
    TRY(check_problem(p));
    HIPCHK(hipSetDevice(dev->ordinal));
    auto *r = new pxfdtd_resident();
    r->dev = dev;
    r->ordinal = dev->ordinal;
    r->prec = p->precision;
    r->esz = p->precision == PXFDTD_F64 ? 8 : 4;
    r->nx = p->nx;
    r->ny = p->ny;
    r->nz = p->nz;
    r->pitch = pitch_of(p->nx);
    r->plane = (long long)r->pitch * p->ny;
    r->nval = r->plane * p->nz;
    r->ncells = (long long)p->nx * p->ny * p->nz;
    for (int a = 0; a < 3; a++) {
        r->periodic[a] = p->periodic[a] ? 1 : 0;
        r->low[a] = p->pml_low[a];
        r->high[a] = p->pml_high[a];
        r->step[a] = p->step[a];
    }
    r->dt = p->dt;
    if (r->nval * 2 >= (1LL << 31)) {
        const long long nval = r->nval;
        delete r;
        return px_fail(PXFDTD_ERR_UNSUPPORTED, "a grid of %lld values per component is beyond 32-bit indexing", nval);
    }
    const long long psi_len = psi_values(p);
    if (psi_len >= (1LL << 31)) {
        delete r;
        return px_fail(PXFDTD_ERR_UNSUPPORTED, "the CPML's %lld psi values are beyond 32-bit indexing", psi_len);
    }
    // defaults: the aligned march (the 2D march for a grid one cell thick with z periodic)
    // where its 32-bit buffer offsets reach every value, else the 62-wide march, else the LDS
    // tiles; 2 rows per strip, and plane chunks short enough that the chip sees several rounds
    // of wavefronts (8 planes, fewer on small grids)
    r->ty = 2;
    const int first = p->nz == 1 && p->periodic[2] ? 4 : 3;
    r->kernel = kernel_fits(first, r->nval, psi_len, (long long)r->esz) ? first
                : (kernel_fits(1, r->nval, psi_len, (long long)r->esz) ? 1 : 0);
    if (r->kernel == 0) r->ty = 8;
    r->concurrent = r->kernel == 3 ? 1 : 0;
    if (r->kernel != 4) {
        const int width = r->kernel == 3 ? SW : (r->kernel == 0 ? TX : RW);
        const long long strips = (long long)((p->nx + width - 1) / width) * ((p->ny + r->ty - 1) / r->ty);
        int kz = std::min(p->nz, 8);
        while (kz > 2 && strips * ((p->nz + kz - 1) / kz) < 4096) kz = (kz + 1) / 2;
        r->kz = std::max(1, kz);
    } else {
        r->ty = 1;
        r->kz = default_rows_2d(p->nx, p->ny);
    }
    const int ord = dev->ordinal;
    hipError_t e = take_stream(ord, &r->stream);
    if (e != hipSuccess) {
        delete r;
        return px_fail(PXFDTD_ERR_DEVICE, "hipStreamCreate: %s", hipGetErrorString(e));
    }
    if (take_stream(ord, &r->side) != hipSuccess || take_event(ord, false, &r->fork_ev) != hipSuccess ||
        take_event(ord, false, &r->join_ev) != hipSuccess) {
        give_stream(ord, r->side);
        give_event(ord, false, r->fork_ev);
        give_event(ord, false, r->join_ev);
        give_stream(ord, r->stream);
        delete r;
        return px_fail(PXFDTD_ERR_DEVICE, "the side stream or its events could not be created");
    }
    (void)take_event(ord, true, &r->ev0);
    (void)take_event(ord, true, &r->ev1);
    int rc = r->prec == PXFDTD_F64 ? build<double>(r, p) : build<float>(r, p);
    if (rc != PXFDTD_OK) {
        std::string msg = g_error;
        pxfdtd_resident_destroy(r);
        g_error = msg;
        return rc;
    }
    *out = r;
    return PXFDTD_OK;
}

extern "C" void pxfdtd_resident_destroy(pxfdtd_resident *r) {
    if (!r) return;
    const int ord = r->ordinal;  // not r->dev: a resident may outlive its device handle
    if (r->stream) (void)hipStreamSynchronize(r->stream);
    give_event(ord, true, r->ev0);
    give_event(ord, true, r->ev1);
    for (auto e : r->kev) (void)hipEventDestroy(e);
    if (r->side) (void)hipStreamSynchronize(r->side);
    give_event(ord, false, r->fork_ev);
    give_event(ord, false, r->join_ev);
    for (int c = 0; c < 4; c++) {
        if (r->ps.s[c]) (void)hipStreamSynchronize(r->ps.s[c]);
        if (r->ps.join[c]) (void)hipEventDestroy(r->ps.join[c]);
    }
    if (r->ps.fork) (void)hipEventDestroy(r->ps.fork);
    hipStream_t s = r->stream, side = r->side;
    PairStreams ps = r->ps;
    {
        std::lock_guard<std::mutex> g(specs_mutex());
        specs().erase(r);
    }
    tl_give_blocks = true;  // its streams are synchronized: its blocks can go to the next ones
    delete r;               // buffers free themselves
    tl_give_blocks = false;
    give_stream(ord, s);
    give_stream(ord, side);
    for (int c = 0; c < 4; c++)
        if (ps.s[c]) (void)hipStreamDestroy(ps.s[c]);
}

extern "C" int64_t pxfdtd_resident_psi_len(const pxfdtd_resident *r) { return r->psi_len; }
extern "C" int32_t pxfdtd_resident_coefficient_mode(const pxfdtd_resident *r) { return r->mode; }
extern "C" double pxfdtd_resident_last_kernel_ms(const pxfdtd_resident *r) { return r->last_ms; }

extern "C" int pxfdtd_resident_sums_device(const pxfdtd_resident *r, const double **sums) {
    if (!r->sums_len) return px_fail(PXFDTD_ERR_INVALID, "the resident holds no transform sums");
    *sums = (const double *)r->sums.p;
    return PXFDTD_OK;
}

extern "C" void *pxfdtd_resident_stream(const pxfdtd_resident *r) { return (void *)r->stream; }

extern "C" int pxfdtd_resident_synchronize(pxfdtd_resident *r) {
    HIPCHK(hipStreamSynchronize(r->stream));
    return PXFDTD_OK;
}

// ------------------------------------------------------------------- upload / download
template <typename T>
static int upload_fields(pxfdtd_resident *r, const double *e, const double *h, const double *psi) {
    const long long N = r->ncells;
    // the other copy and its psi start at zero: a field set the 2D march leaves out is then
    // zero in both copies
    HIPCHK(hipMemsetAsync(r->fields[1 - r->p].p, 0, r->fields[1 - r->p].bytes, r->stream));
    if (r->psi_len) HIPCHK(hipMemsetAsync(r->psi[1 - r->p].p, 0, r->psi[1 - r->p].bytes, r->stream));
    // which sets hold values: TE is Ex, Ey, Hz (field*3 + comp 0, 1, 5), TM is Ez, Hx, Hy
    int set = 0;
    for (int q = 0; q < 6; q++) {
        const double *src = (q < 3 ? e : h);
        if (!src) continue;
        src += (q % 3) * N;
        const int bit = (q == 0 || q == 1 || q == 5) ? 1 : 2;
        if (!(set & bit) && any_nonzero(src, N)) set |= bit;
    }
    if (psi)
        for (const SlabDesc &d : r->slabs) {
            const int q = d.field * 3 + d.comp;
            const int bit = (q == 0 || q == 1 || q == 5) ? 1 : 2;
            if (set & bit) continue;
            for (long long v = 0; v < d.len; v++)
                if (psi[d.offset + v] != 0.0) {
                    set |= bit;
                    break;
                }
        }
    r->fields_set = set;
    std::vector<T> buf;
    for (int q = 0; q < 6; q++) {
        const double *src = (q < 3 ? e : h);
        if (!src) {  // zeros
            HIPCHK(hipMemsetAsync(r->field<T>(r->p, q), 0, r->nval * sizeof(T), r->stream));
            continue;
        }
        src += (q % 3) * N;
        buf.assign(r->nval, T(0));
        to_pitched<T>(r, src, buf.data());
        HIPCHK(hipMemcpyAsync(r->field<T>(r->p, q), buf.data(), r->nval * sizeof(T),
                              hipMemcpyHostToDevice, r->stream));
        HIPCHK(hipStreamSynchronize(r->stream));
    }
    if (r->psi_len && !psi) HIPCHK(hipMemsetAsync(r->psi[r->p].p, 0, r->psi_len * sizeof(T), r->stream));
    if (r->psi_len && psi) {
        std::vector<T> ps(r->psi_len);
        for (long long q = 0; q < r->psi_len; q++) ps[q] = (T)psi[q];
        HIPCHK(hipMemcpyAsync(r->psi[r->p].p, ps.data(), r->psi_len * sizeof(T),
                              hipMemcpyHostToDevice, r->stream));
        HIPCHK(hipStreamSynchronize(r->stream));
    }
    return PXFDTD_OK;
}

extern "C" int pxfdtd_resident_upload(pxfdtd_resident *r, const double *e, const double *h,
                                      const double *psi) {
    HIPCHK(hipSetDevice(r->dev->ordinal));
    return r->prec == PXFDTD_F64 ? upload_fields<double>(r, e, h, psi)
                                 : upload_fields<float>(r, e, h, psi);
}

template <typename T>
static int download_fields(pxfdtd_resident *r, double *e, double *h, double *psi) {
    const long long N = r->ncells;
    std::vector<T> buf(r->nval);
    HIPCHK(hipStreamSynchronize(r->stream));
    for (int q = 0; q < 6; q++) {
        double *dst = (q < 3 ? e : h);
        if (!dst) continue;
        dst += (q % 3) * N;
        HIPCHK(hipMemcpyAsync(buf.data(), r->field<T>(r->p, q), r->nval * sizeof(T),
                              hipMemcpyDeviceToHost, r->stream));
        HIPCHK(hipStreamSynchronize(r->stream));
        from_pitched<T>(r, buf.data(), dst);
    }
    if (psi && r->psi_len) {
        std::vector<T> ps(r->psi_len);
        HIPCHK(hipMemcpyAsync(ps.data(), r->psi[r->p].p, r->psi_len * sizeof(T),
                              hipMemcpyDeviceToHost, r->stream));
        HIPCHK(hipStreamSynchronize(r->stream));
        for (long long q = 0; q < r->psi_len; q++) psi[q] = (double)ps[q];
    }
    return PXFDTD_OK;
}

extern "C" int pxfdtd_resident_download(pxfdtd_resident *r, double *e, double *h, double *psi) {
    HIPCHK(hipSetDevice(r->dev->ordinal));
    return r->prec == PXFDTD_F64 ? download_fields<double>(r, e, h, psi)
                                 : download_fields<float>(r, e, h, psi);
}

// ------------------------------------------------------------------------- injections
// Each tile's list holds the injections on its own values and on its halo (the row below and
// the column left of it, where H is recomputed), sorted by plane, each plane's in the
// caller's order: a value's injections apply in that order.
// The tile geometry of the kernel in use: own values [i0, i0 + tw) x [j0, j0 + th), tiles at
// multiples of tw and th; H is also computed on the row below and the column left of the
// tile (the LDS kernel: the row over the tile's columns and the column over its rows; the
// register march: the whole rectangle [i0 - 1, i0 + tw) x [j0 - 1, j0 + th)).
static void tile_shape(const pxfdtd_resident *r, int &tw, int &th) {
    // kernels 1 and 2: the 62-wide march (with and without prefetch); 3: the aligned march
    tw = r->kernel == 0 ? TX : (r->kernel >= 3 ? SW : RW);
    th = r->kernel == 4 ? r->ny : r->ty;  // the 2D march: a strip's lists run over all rows
}

static bool h_region_has(const pxfdtd_resident *r, int bx, int by, int i, int j) {
    int tw, th;
    tile_shape(r, tw, th);
    const int i0 = bx * tw, j0 = by * th;
    // the tile's own and halo coordinates, unwrapped, against i or i +- n (a periodic wrap)
    auto in_x = [&](int lo, int hi) {  // [lo, hi)
        for (int d : {0, r->nx, -r->nx}) {
            if (d && !r->periodic[0]) continue;
            const int ii = i + d;
            if (ii >= lo && ii < hi) return true;
        }
        return false;
    };
    auto in_y = [&](int lo, int hi) {
        for (int d : {0, r->ny, -r->ny}) {
            if (d && !r->periodic[1]) continue;
            const int jj = j + d;
            if (jj >= lo && jj < hi) return true;
        }
        return false;
    };
    const int iw = std::min(i0 + tw, r->nx + (r->periodic[0] ? 1 : 0));
    if (r->kernel == 1 || r->kernel == 2) return in_x(i0 - 1, iw) && in_y(j0 - 1, j0 + th);
    const bool own = in_x(i0, i0 + tw) && in_y(j0, j0 + th);
    const bool row = in_x(i0, i0 + tw) && in_y(j0 - 1, j0);
    const bool col = in_x(i0 - 1, i0) && in_y(j0, j0 + th);
    return own || row || col;
}

// Each tile's list holds the injections on its own values and, for H, on the values around
// it where H is recomputed, sorted by plane, each plane's in the caller's order: a value's
// injections apply in that order.
// The 2D march's lists: per strip of 64 columns, per row. A strip holds the injections on its
// own columns and, for H, on the column left of it (where it recomputes H).
// the sources' injections, then the plane waves' couplings: the order they apply in on a value
static std::vector<pxfdtd_injection> every_injection(const pxfdtd_resident *r) {
    std::vector<pxfdtd_injection> all = r->inj_host;
    all.insert(all.end(), r->coup_host.begin(), r->coup_host.end());
    return all;
}

static int build_injection_lists_2d(pxfdtd_resident *r) {
    const int tx = (r->nx + SW - 1) / SW, ny = r->ny;
    struct Item { int strip, j; long long seq; Injection in; };
    std::vector<Item> items;
    const std::vector<pxfdtd_injection> all = every_injection(r);
    for (size_t q = 0; q < all.size(); q++) {
        const pxfdtd_injection &s = all[q];
        const long long v = s.index;
        const int i = (int)(v % r->nx), j = (int)((v / r->nx) % r->ny);
        Injection in;
        in.k = 0; in.i = i; in.j = j; in.code = s.field * 4 + s.component; in.emitter = s.emitter;
        in.pad = 0; in.a_re = s.a_re; in.a_im = s.a_im;
        items.push_back({i / SW, j, (long long)q, in});
        if (s.field == PXFDTD_H) {
            // the strip whose left column this is: i + 1 starts it (or column 0 across a
            // periodic side, for i = nx - 1)
            int right = -1;
            if ((i + 1) % SW == 0 && i + 1 < r->nx) right = (i + 1) / SW;
            if (i == r->nx - 1 && r->periodic[0]) right = 0;
            if (right >= 0 && right != i / SW) items.push_back({right, j, (long long)q, in});
        }
    }
    std::stable_sort(items.begin(), items.end(), [](const Item &a, const Item &b) {
        return a.strip != b.strip ? a.strip < b.strip : (a.j != b.j ? a.j < b.j : a.seq < b.seq);
    });
    std::vector<int32_t> off((size_t)tx * (ny + 1), 0);
    std::vector<Injection> flat(items.size());
    size_t q = 0;
    for (int t = 0; t < tx; t++)
        for (int j = 0; j <= ny; j++) {
            off[(size_t)t * (ny + 1) + j] = (int32_t)q;
            if (j == ny) break;
            while (q < items.size() && items[q].strip == t && items[q].j == j) {
                flat[q] = items[q].in;
                q++;
            }
        }
    TRY(r->inj_off.alloc(off.size() * sizeof(int32_t)));
    TRY(r->inj.alloc(std::max<size_t>(flat.size(), 1) * sizeof(Injection)));
    HIPCHK(hipMemcpyAsync(r->inj_off.p, off.data(), off.size() * sizeof(int32_t), hipMemcpyHostToDevice, r->stream));
    if (!flat.empty())
        HIPCHK(hipMemcpyAsync(r->inj.p, flat.data(), flat.size() * sizeof(Injection), hipMemcpyHostToDevice, r->stream));
    HIPCHK(hipStreamSynchronize(r->stream));
    return PXFDTD_OK;
}

static int build_injection_lists(pxfdtd_resident *r) {
    r->units2_dirty = true;  // the two-step march's lists follow
    r->have_inj = !r->inj_host.empty() || !r->coup_host.empty();
    r->inj_set = 0;
    auto mark = [&](const pxfdtd_injection &s) {
        const int q = s.field * 3 + s.component;
        r->inj_set |= (q == 0 || q == 1 || q == 5) ? 1 : 2;
    };
    for (const auto &s : r->inj_host) mark(s);
    // a coupling whose incident component stays zero changes no value of a field set at rest
    for (size_t q = 0; q < r->coup_host.size(); q++)
        if (r->coup_live[q]) mark(r->coup_host[q]);
    if (!r->have_inj) return PXFDTD_OK;
    if (r->kernel == 4) return build_injection_lists_2d(r);
    int tw, th;
    tile_shape(r, tw, th);
    const int tx = (r->nx + tw - 1) / tw, tyn = (r->ny + th - 1) / th;
    const long long tiles = (long long)tx * tyn;
    struct Item { long long tile; int k; long long seq; Injection in; };
    std::vector<Item> items;
    const std::vector<pxfdtd_injection> all = every_injection(r);
    for (size_t q = 0; q < all.size(); q++) {
        const pxfdtd_injection &s = all[q];
        const long long v = s.index;
        const int i = (int)(v % r->nx), j = (int)((v / r->nx) % r->ny), k = (int)(v / ((long long)r->nx * r->ny));
        Injection in;
        in.k = k; in.i = i; in.j = j; in.code = s.field * 4 + s.component; in.emitter = s.emitter;
        in.pad = 0; in.a_re = s.a_re; in.a_im = s.a_im;
        if (s.field == PXFDTD_E) {
            items.push_back({(long long)(j / th) * tx + i / tw, k, (long long)q, in});
            continue;
        }
        int bxs[4] = {i / tw, (i + 1) / tw, 0, tx - 1}, bys[4] = {j / th, (j + 1) / th, 0, tyn - 1};
        std::vector<long long> seen;
        for (int a = 0; a < 4; a++)
            for (int c = 0; c < 4; c++) {
                const int bx = bxs[a], by = bys[c];
                if (bx < 0 || bx >= tx || by < 0 || by >= tyn) continue;
                const long long tile = (long long)by * tx + bx;
                if (std::find(seen.begin(), seen.end(), tile) != seen.end()) continue;
                if (!h_region_has(r, bx, by, i, j)) continue;
                seen.push_back(tile);
                items.push_back({tile, k, (long long)q, in});
            }
    }
    std::stable_sort(items.begin(), items.end(), [](const Item &a, const Item &b) {
        return a.tile != b.tile ? a.tile < b.tile : (a.k != b.k ? a.k < b.k : a.seq < b.seq);
    });
    std::vector<int32_t> off(tiles * (r->nz + 1), 0);
    std::vector<Injection> flat(items.size());
    size_t q = 0;
    for (long long t = 0; t < tiles; t++)
        for (int k = 0; k <= r->nz; k++) {
            off[t * (r->nz + 1) + k] = (int32_t)q;
            if (k == r->nz) break;
            while (q < items.size() && items[q].tile == t && items[q].k == k) {
                flat[q] = items[q].in;
                q++;
            }
        }
    TRY(r->inj_off.alloc(off.size() * sizeof(int32_t)));
    TRY(r->inj.alloc(std::max<size_t>(flat.size(), 1) * sizeof(Injection)));
    HIPCHK(hipMemcpyAsync(r->inj_off.p, off.data(), off.size() * sizeof(int32_t), hipMemcpyHostToDevice, r->stream));
    if (!flat.empty())
        HIPCHK(hipMemcpyAsync(r->inj.p, flat.data(), flat.size() * sizeof(Injection), hipMemcpyHostToDevice, r->stream));
    HIPCHK(hipStreamSynchronize(r->stream));
    return PXFDTD_OK;
}

// The register march's work: plane chunks of about kz planes whose edges fall at the CPML
// slabs' along z (a chunk also reads the plane below it), crossed with the strips; a unit goes
// to the slab kernel if any value it computes H or E on lies in a CPML slab.
// The 2D march's work: strips crossed with chunks of about kz rows whose edges fall at the y
// slabs' (a chunk also reads the row below it); a unit meeting a CPML slab goes to the slab
// variant.
static int build_units_2d(pxfdtd_resident *r) {
    std::vector<int> cuts = {0};
    const int ly = r->periodic[1] ? 0 : r->low[1], hy = r->periodic[1] ? 0 : r->high[1], ny = r->ny;
    auto cut = [&](int y) {
        if (y > cuts.back() && y < ny) cuts.push_back(y);
    };
    if (ly > 0) cut(ly + 1);
    const int hi_start = hy > 0 ? ny - hy : ny;
    for (int y = cuts.back() + r->kz; y < hi_start; y += r->kz) cut(y);
    cut(hi_start);
    for (int y = cuts.back() + r->kz; y < ny; y += r->kz) cut(y);
    cuts.push_back(ny);
    const int tx = (r->nx + SW - 1) / SW;
    auto in_slab = [](int a, int b, int n, int lo, int hi) {  // [a, b) meets [0, lo) or [n-hi, n)
        return (lo > 0 && a < lo) || (hi > 0 && b > n - hi);
    };
    std::vector<int4> lean, slab;
    for (size_t c = 0; c + 1 < cuts.size(); c++) {
        const int a = cuts[c], b = cuts[c + 1];
        const bool ys = in_slab(a > 0 || r->periodic[1] ? a - 1 : a, b, ny, ly, hy);
        for (int bx = 0; bx < tx; bx++) {
            const int i0 = bx * SW;
            const bool xs = !r->periodic[0] && in_slab(std::max(i0 - 1, 0), std::min(i0 + SW, r->nx), r->nx, r->low[0], r->high[0]);
            ((xs || ys) ? slab : lean).push_back(make_int4(bx, 0, a, b));
        }
    }
    // a grid too small to fill the chip several times over steps in one launch (the slab
    // variant's psi work is behind uniform branches): its step time is one launch's latency.
    // Not in the server's batches, where the jobs fill the chip together and each class runs in
    // one launch whatever its jobs' sizes: there the lean units keep their cheaper variant.
    if ((long long)(lean.size() + slab.size()) < 4LL * 4096 && !r->batched) {
        slab.insert(slab.begin(), lean.begin(), lean.end());
        lean.clear();
    }
    r->n_lean = (int)lean.size();
    r->n_x = 0;
    r->n_slab = (int)slab.size();
    lean.insert(lean.end(), slab.begin(), slab.end());
    r->units_h = lean;
    TRY(r->units.alloc(std::max<size_t>(lean.size(), 1) * sizeof(int4)));
    HIPCHK(hipMemcpy(r->units.p, lean.data(), lean.size() * sizeof(int4), hipMemcpyHostToDevice));
    r->units_dirty = false;
    return PXFDTD_OK;
}

// lean units march this many z chunks (MMI step 0.338 -> 0.336 ms, the dispersive grid's
// 1.469 -> 1.454 ms; 4 chunks: 0.343 and 1.433)
constexpr int LEAN_CHUNKS = 2;
static int build_units(pxfdtd_resident *r) {
    if (r->kernel == 4) return build_units_2d(r);
    std::vector<int> cuts = {0};
    const int lz = r->low[2], hz = r->high[2], nz = r->nz;
    auto cut = [&](int z) {
        if (z > cuts.back() && z < nz) cuts.push_back(z);
    };
    if (lz > 0) cut(lz + 1);
    const int hi_start = hz > 0 ? nz - hz : nz;
    for (int z = cuts.back() + r->kz; z < hi_start; z += r->kz) cut(z);
    cut(hi_start);
    cuts.push_back(nz);
    int tw, th;
    tile_shape(r, tw, th);
    const int tx = (r->nx + tw - 1) / tw, ty = (r->ny + th - 1) / th;
    auto in_slab = [](int a, int b, int n, int lo, int hi) {  // [a, b) meets [0, lo) or [n-hi, n)
        return (lo > 0 && a < lo) || (hi > 0 && b > n - hi);
    };
    std::vector<int4> lean, xonly, slab;
    for (size_t c = 0; c + 1 < cuts.size(); c++) {
        const int a = cuts[c], b = cuts[c + 1];
        const bool zs = in_slab(a > 0 || r->periodic[2] ? a - 1 : a, b, nz, lz, hz) && !r->periodic[2];
        for (int by = 0; by < ty; by++) {
            const int j0 = by * th;
            const bool ys = !r->periodic[1] && in_slab(std::max(j0 - 1, 0), std::min(j0 + th, r->ny), r->ny, r->low[1], r->high[1]);
            for (int bx = 0; bx < tx; bx++) {
                const int i0 = bx * tw;
                const bool xs = !r->periodic[0] && in_slab(std::max(i0 - 1, 0), std::min(i0 + tw, r->nx), r->nx, r->low[0], r->high[0]);
                const int4 u = make_int4(bx, by, a, b);
                if (r->kernel == 3 && r->concurrent == 2)
                    slab.push_back(u);  // one launch of every unit, in the grid's order
                else if (r->kernel == 3 && xs && !ys && !zs)
                    xonly.push_back(u);
                else
                    ((xs || ys || zs) ? slab : lean).push_back(u);
            }
        }
    }
    // the lean units march LEAN_CHUNKS chunks at a time: their waves are the shortest, and a
    // longer march loads fewer prologue planes
    if (r->kernel == 3 && LEAN_CHUNKS > 1 && !lean.empty()) {
        std::sort(lean.begin(), lean.end(), [](const int4 &p, const int4 &q) {
            return std::make_tuple(p.y, p.x, p.z) < std::make_tuple(q.y, q.x, q.z);
        });
        std::vector<int4> merged;
        for (size_t q = 0; q < lean.size();) {
            int4 u = lean[q];
            size_t k = q + 1;
            while (k < lean.size() && k - q < (size_t)LEAN_CHUNKS && lean[k].x == u.x && lean[k].y == u.y &&
                   lean[k].z == u.w) {
                u.w = lean[k].w;
                k++;
            }
            merged.push_back(u);
            q = k;
        }
        std::sort(merged.begin(), merged.end(), [](const int4 &p, const int4 &q) {
            return std::make_tuple(p.z, p.y, p.x) < std::make_tuple(q.z, q.y, q.x);
        });
        lean = merged;
    }
    // the slab units of a two-row tiling march one row each, in the general variant held to four
    // waves per SIMD (at two rows its registers leave three): each unit's half as its own unit,
    // the two adjacent in the list
    r->slab_rows = th;
    if (r->kernel == 3 && r->concurrent != 2 && th == 2) {
        std::vector<int4> halves;
        halves.reserve(2 * slab.size());
        for (const int4 &u : slab)
            for (int h = 0; h < 2; h++)
                if (2 * u.y + h < r->ny) halves.push_back(make_int4(u.x, 2 * u.y + h, u.z, u.w));
        slab.swap(halves);
        r->slab_rows = 1;
    }
    r->n_lean = (int)lean.size();
    r->n_x = (int)xonly.size();
    r->n_slab = (int)slab.size();
    lean.insert(lean.end(), xonly.begin(), xonly.end());
    lean.insert(lean.end(), slab.begin(), slab.end());
    r->units_h = lean;
    TRY(r->units.alloc(std::max<size_t>(lean.size(), 1) * sizeof(int4)));
    HIPCHK(hipMemcpy(r->units.p, lean.data(), lean.size() * sizeof(int4), hipMemcpyHostToDevice));
    r->units_dirty = false;
    return PXFDTD_OK;
}

// The two-step march's tables: its units (strips of SW2 own columns by ty2 rows, chunks of kz2
// planes), each in the CPML variant of the slabs its margin meets (the values it computes reach
// 2 columns left and 2 right of its strip, 2 rows below and 2 above, 2 planes below and 2
// above); the sources, currents and couplings of each tile's margin, sorted by plane (every
// value either step computes takes its own); and each tile's monitored values, sorted by plane.
static int build_pair_tables(pxfdtd_resident *r) {
    const int nx = r->nx, ny = r->ny, nz = r->nz, th = r->ty2;
    const int tx = (nx + SW2 - 1) / SW2, tyn = (ny + th - 1) / th;
    r->tiles2_x = tx;
    std::vector<int> cuts = {0};
    const int lz = r->low[2], hz = r->high[2];
    auto cut = [&](int z) {
        if (z > cuts.back() && z < nz) cuts.push_back(z);
    };
    const int lo_end = lz > 0 ? lz + 2 : 0, hi_start = hz > 0 ? nz - hz - 2 : nz;
    for (int z = r->kz2; z < lo_end; z += r->kz2) cut(z);
    cut(lo_end);
    for (int z = cuts.back() + r->kz2; z < hi_start; z += r->kz2) cut(z);
    cut(hi_start);
    for (int z = cuts.back() + r->kz2; z < nz; z += r->kz2) cut(z);
    cuts.push_back(nz);
    auto in_slab = [](int a, int b, int n, int lo, int hi) {  // [a, b) meets [0, lo) or [n-hi, n)
        return (lo > 0 && a < lo) || (hi > 0 && b > n - hi);
    };
    std::vector<int4> lean, xonly, yonly, zonly, slab;
    for (size_t c = 0; c + 1 < cuts.size(); c++) {
        const int a = cuts[c], b = cuts[c + 1];
        const bool zs = !r->periodic[2] && in_slab(std::max(a - 2, 0), std::min(b + 2, nz), nz, lz, hz);
        for (int by = 0; by < tyn; by++) {
            const int j0 = by * th;
            const bool ys = !r->periodic[1] && in_slab(std::max(j0 - 2, 0), std::min(j0 + th + 2, ny), ny, r->low[1], r->high[1]);
            for (int bx = 0; bx < tx; bx++) {
                const int i0 = bx * SW2;
                const bool xs = !r->periodic[0] && in_slab(std::max(i0 - 2, 0), std::min(i0 + SW2 + 2, nx), nx, r->low[0], r->high[0]);
                const int4 u = make_int4(bx, by, a, b);
                const int nslabs = (int)xs + (int)ys + (int)zs;
                if (nslabs == 0)
                    lean.push_back(u);
                else if (nslabs > 1)
                    slab.push_back(u);
                else
                    (xs ? xonly : (ys ? yonly : zonly)).push_back(u);
            }
        }
    }
    r->n2[0] = (int)lean.size();
    r->n2[1] = (int)xonly.size();
    r->n2[2] = (int)yonly.size();
    r->n2[3] = (int)zonly.size();
    r->n2[4] = (int)slab.size();
    for (auto *v : {&xonly, &yonly, &zonly, &slab}) lean.insert(lean.end(), v->begin(), v->end());
    TRY(r->units2.alloc(std::max<size_t>(lean.size(), 1) * sizeof(int4)));
    HIPCHK(hipMemcpy(r->units2.p, lean.data(), lean.size() * sizeof(int4), hipMemcpyHostToDevice));

    // the sources and currents per (tile, plane, margin row, field): a tile's margin rows are
    // j0 - 2 + r, r < th + 4 (wrapped on a periodic axis), its lanes i0 - 2 .. i0 + 61
    const int NE = th + 4;
    const long long tiles = (long long)tx * tyn;
    auto xin = [&](int i, int i0) {
        for (int d : {0, nx, -nx}) {
            if (d && !r->periodic[0]) continue;
            if (i + d >= i0 - 2 && i + d < i0 + SW2 + 2) return true;
        }
        return false;
    };
    auto yrow = [&](int jr) {  // the grid row of margin row jr (unwrapped), -1 beyond a wall
        if (jr >= 0 && jr < ny) return jr;
        if (!r->periodic[1]) return -1;
        return ((jr % ny) + ny) % ny;
    };
    if (r->have_inj) {
        const std::vector<pxfdtd_injection> all = every_injection(r);
        // entries per (tile, k, margin row, field) in the list's order
        std::vector<std::vector<std::pair<long long, Injection>>> at((size_t)tiles);
        for (size_t q = 0; q < all.size(); q++) {
            const pxfdtd_injection &sj = all[q];
            const long long v = sj.index;
            const int i = (int)(v % nx), j = (int)((v / nx) % ny), k = (int)(v / ((long long)nx * ny));
            Injection in;
            in.k = k; in.i = i; in.j = j; in.code = sj.field * 4 + sj.component; in.emitter = sj.emitter;
            in.pad = 0; in.a_re = sj.a_re; in.a_im = sj.a_im;
            for (int by = 0; by < tyn; by++)
                for (int rr = 0; rr < NE; rr++) {
                    if (yrow(by * th - 2 + rr) != j) continue;
                    for (int bx = 0; bx < tx; bx++) {
                        if (!xin(i, bx * SW2)) continue;
                        const long long key = (((long long)k * NE + rr) * 2 + sj.field);
                        at[(size_t)(by * tx + bx)].push_back({key, in});
                    }
                }
        }
        const long long per = (long long)nz * NE * 2;
        std::vector<int32_t> off(tiles * per + 1, 0);
        std::vector<Injection> flat;
        for (long long t = 0; t < tiles; t++) {
            auto &L = at[(size_t)t];
            std::stable_sort(L.begin(), L.end(), [](const std::pair<long long, Injection> &a,
                                                    const std::pair<long long, Injection> &b) { return a.first < b.first; });
            size_t q = 0;
            for (long long key = 0; key < per; key++) {
                off[t * per + key] = (int32_t)flat.size();
                while (q < L.size() && L[q].first == key) flat.push_back(L[q++].second);
            }
        }
        off[tiles * per] = (int32_t)flat.size();
        TRY(r->inj2_off.alloc(off.size() * sizeof(int32_t)));
        TRY(r->inj2.alloc(std::max<size_t>(flat.size(), 1) * sizeof(Injection)));
        HIPCHK(hipMemcpy(r->inj2_off.p, off.data(), off.size() * sizeof(int32_t), hipMemcpyHostToDevice));
        if (!flat.empty()) HIPCHK(hipMemcpy(r->inj2.p, flat.data(), flat.size() * sizeof(Injection), hipMemcpyHostToDevice));
    }
    // the monitored values per (tile, plane, own row, component), each listed by its one owner
    r->have_tap2 = r->ntaps > 0;
    if (r->have_tap2) {
        const long long per = (long long)nz * th * 6;
        std::vector<std::vector<std::pair<long long, TapAt>>> at((size_t)tiles);
        for (long long q = 0; q < r->ntaps; q++) {
            const long long v = r->tap_index_h[(size_t)q];
            const int k = (int)(v / r->plane), j = (int)((v % r->plane) / r->pitch), i = (int)(v % r->pitch);
            const int c = r->tap_which_h[(size_t)q];
            const TapAt a = {i, 0, q};
            const long long key = ((long long)k * th + j % th) * 6 + c;
            at[(size_t)((j / th) * tx + i / SW2)].push_back({key, a});
        }
        std::vector<int32_t> off(tiles * per + 1, 0);
        std::vector<TapAt> flat;
        for (long long t = 0; t < tiles; t++) {
            auto &L = at[(size_t)t];
            std::stable_sort(L.begin(), L.end(), [](const std::pair<long long, TapAt> &a,
                                                    const std::pair<long long, TapAt> &b) { return a.first < b.first; });
            size_t q = 0;
            for (long long key = 0; key < per; key++) {
                off[t * per + key] = (int32_t)flat.size();
                while (q < L.size() && L[q].first == key) flat.push_back(L[q++].second);
            }
        }
        off[tiles * per] = (int32_t)flat.size();
        TRY(r->tap2_off.alloc(off.size() * sizeof(int32_t)));
        TRY(r->tap2.alloc(std::max<size_t>(flat.size(), 1) * sizeof(TapAt)));
        HIPCHK(hipMemcpy(r->tap2_off.p, off.data(), off.size() * sizeof(int32_t), hipMemcpyHostToDevice));
        if (!flat.empty()) HIPCHK(hipMemcpy(r->tap2.p, flat.data(), flat.size() * sizeof(TapAt), hipMemcpyHostToDevice));
    }
    r->units2_dirty = false;
    return PXFDTD_OK;
}

// whether the problem's next two steps can run as one pass of the two-step march: no plane
// waves (their auxiliary grids step between the grid's steps), no media (a pass of their own
// after each step), no per-step hook, and the two steps' taps in one staging block
static bool pair_ok(const pxfdtd_resident *r, int in_block) {
    return r->pair && r->nz > 1 && r->planes_h.empty() && !r->nmvals && !r->ngroups && !r->after_step &&
           (r->ntaps == 0 || r->quiet || in_block + 2 <= r->block_steps);
}

extern "C" int pxfdtd_resident_set_injections(pxfdtd_resident *r, const pxfdtd_injection *inj,
                                              int64_t n, int32_t emitters) {
    for (int64_t q = 0; q < n; q++) {
        const pxfdtd_injection &s = inj[q];
        if (s.field < 0 || s.field > 1 || s.component < 0 || s.component > 2 || s.index < 0 ||
            s.index >= r->ncells || s.emitter < 0 || s.emitter >= emitters)
            return px_fail(PXFDTD_ERR_INVALID, "injection %lld is out of range", (long long)q);
    }
    HIPCHK(hipSetDevice(r->dev->ordinal));
    r->inj_host.assign(inj, inj + n);
    r->emitters = emitters;
    return build_injection_lists(r);
}

// Which of a plane wave's auxiliary components can ever be nonzero: E's across the
// polarization, and what the curls reach from them (a difference along an axis the direction
// has no part on is exactly zero).
static void live_components(const pxfdtd_plane_wave &w, bool live[2][3]) {
    for (int c = 0; c < 3; c++) live[1][c] = false;
    for (int c = 0; c < 3; c++) {
        live[0][c] = false;
        for (int q = 0; q < w.nsheet; q++) live[0][c] = live[0][c] || w.sheet_component[q] == c;
    }
    for (bool grew = true; grew;) {
        grew = false;
        for (int f = 0; f < 2; f++)
            for (int c = 0; c < 3; c++) {
                const int a = (c + 1) % 3, b = (c + 2) % 3;
                const bool reach = (w.g[a] != 0 && live[1 - f][b]) || (w.g[b] != 0 && live[1 - f][a]);
                if (reach && !live[f][c]) live[f][c] = grew = true;
            }
    }
}

extern "C" int pxfdtd_resident_set_plane_waves(pxfdtd_resident *r, const pxfdtd_plane_wave *waves,
                                               int32_t n, const pxfdtd_coupling *couplings,
                                               int64_t count) {
    if (n < 0 || count < 0 || (n && !waves) || (count && !couplings))
        return px_fail(PXFDTD_ERR_INVALID, "plane waves: bad counts or arrays");
    std::vector<PlaneDesc> desc((size_t)n);
    long long fields = 0, coefs = 0, longest = 0;
    for (int q = 0; q < n; q++) {
        const pxfdtd_plane_wave &w = waves[q];
        long long reach = 0;
        for (int a = 0; a < 3; a++) reach = std::max<long long>(reach, std::llabs((long long)w.g[a]));
        if (w.len <= 0 || w.pad < reach || 2 * w.pad >= w.len || !w.ca || !w.cb || !w.da || !w.db ||
            w.nsheet < 0 || w.nsheet > 6 || w.emitter < 0 || w.emitter >= std::max(r->emitters, 1))
            return px_fail(PXFDTD_ERR_INVALID, "plane wave %d: its grid, pad, sheet or emitter is out of range", q);
        PlaneDesc &d = desc[q];
        memset(&d, 0, sizeof d);
        for (int a = 0; a < 3; a++) {
            d.g[a] = w.g[a];
            d.step[a] = r->step[a];
        }
        d.emitter = w.emitter;
        d.len = w.len;
        d.pad = w.pad;
        d.field_off = fields;
        d.coef_off = coefs;
        d.nsheet = w.nsheet;
        for (int s2 = 0; s2 < w.nsheet; s2++) {
            if (w.sheet_component[s2] < 0 || w.sheet_component[s2] > 2 || w.sheet_sample[s2] < w.pad ||
                w.sheet_sample[s2] >= w.len - w.pad)
                return px_fail(PXFDTD_ERR_INVALID, "plane wave %d: sheet entry %d is out of range", q, s2);
            d.sheet_c[s2] = w.sheet_component[s2];
            d.sheet_i[s2] = w.sheet_sample[s2];
            d.sheet_j[s2] = w.sheet_j[s2];
        }
        fields += 6 * w.len;
        coefs += 4 * w.len;
        longest = std::max<long long>(longest, w.len);
    }
    if (fields >= (1LL << 31))
        return px_fail(PXFDTD_ERR_INVALID, "the plane waves' auxiliary grids are too long");
    std::vector<pxfdtd_injection> coup((size_t)count);
    std::vector<char> live((size_t)count);
    std::vector<std::array<std::array<bool, 3>, 2>> lives((size_t)n);
    for (int q = 0; q < n; q++) {
        bool l[2][3];
        live_components(waves[q], l);
        for (int f = 0; f < 2; f++)
            for (int c = 0; c < 3; c++) lives[q][f][c] = l[f][c];
    }
    for (int64_t q = 0; q < count; q++) {
        const pxfdtd_coupling &c = couplings[q];
        if (c.field < 0 || c.field > 1 || c.component < 0 || c.component > 2 || c.index < 0 ||
            c.index >= r->ncells || c.plane < 0 || c.plane >= n || c.from_component < 0 ||
            c.from_component > 2 || c.sample < 0 || c.sample >= waves[c.plane].len)
            return px_fail(PXFDTD_ERR_INVALID, "coupling %lld is out of range", (long long)q);
        const PlaneDesc &d = desc[c.plane];
        // E's update reads the incident H~, H's the incident E
        const int from_field = c.field == PXFDTD_E ? 1 : 0;
        const long long at = d.field_off + (long long)(from_field * 3 + c.from_component) * d.len + c.sample;
        coup[q] = {c.field, c.component, c.index, c.weight, 0.0, (int32_t)(-1 - at)};
        live[q] = lives[c.plane][from_field][c.from_component];
    }
    HIPCHK(hipSetDevice(r->dev->ordinal));
    TRY(r->plane_desc.alloc(std::max<size_t>(desc.size(), 1) * sizeof(PlaneDesc)));
    TRY(r->aux.alloc(std::max<long long>(fields, 1) * sizeof(double)));
    TRY(r->aux_coef.alloc(std::max<long long>(coefs, 1) * sizeof(double)));
    if (n) {
        HIPCHK(hipMemcpy(r->plane_desc.p, desc.data(), desc.size() * sizeof(PlaneDesc), hipMemcpyHostToDevice));
        HIPCHK(hipMemset(r->aux.p, 0, fields * sizeof(double)));
        for (int q = 0; q < n; q++) {
            const pxfdtd_plane_wave &w = waves[q];
            const double *src[4] = {w.ca, w.cb, w.da, w.db};
            for (int t = 0; t < 4; t++)
                HIPCHK(hipMemcpy(r->aux_coef.as<double>() + desc[q].coef_off + t * w.len, src[t],
                                 w.len * sizeof(double), hipMemcpyHostToDevice));
        }
    }
    r->planes_h = desc;
    r->aux_len = fields;
    r->aux_max = longest;
    r->coup_host = coup;
    r->coup_live = live;
    return build_injection_lists(r);
}

static int plane_segment(pxfdtd_resident *r, int32_t n, long long &off, long long &len) {
    if (n < 0 || n >= (int32_t)r->planes_h.size()) return px_fail(PXFDTD_ERR_INVALID, "no plane wave %d", n);
    off = r->planes_h[n].field_off;
    len = r->planes_h[n].len;
    return PXFDTD_OK;
}

extern "C" int pxfdtd_resident_plane_state(pxfdtd_resident *r, int32_t n, double *e, double *h) {
    long long off, len;
    TRY(plane_segment(r, n, off, len));
    HIPCHK(hipSetDevice(r->dev->ordinal));
    HIPCHK(hipStreamSynchronize(r->stream));
    if (e) HIPCHK(hipMemcpy(e, r->aux.as<double>() + off, 3 * len * sizeof(double), hipMemcpyDeviceToHost));
    if (h) HIPCHK(hipMemcpy(h, r->aux.as<double>() + off + 3 * len, 3 * len * sizeof(double), hipMemcpyDeviceToHost));
    return PXFDTD_OK;
}

extern "C" int pxfdtd_resident_set_plane_state(pxfdtd_resident *r, int32_t n, const double *e,
                                               const double *h) {
    long long off, len;
    TRY(plane_segment(r, n, off, len));
    HIPCHK(hipSetDevice(r->dev->ordinal));
    HIPCHK(hipStreamSynchronize(r->stream));
    if (e) HIPCHK(hipMemcpy(r->aux.as<double>() + off, e, 3 * len * sizeof(double), hipMemcpyHostToDevice));
    if (h) HIPCHK(hipMemcpy(r->aux.as<double>() + off + 3 * len, h, 3 * len * sizeof(double), hipMemcpyHostToDevice));
    return PXFDTD_OK;
}

extern "C" int pxfdtd_resident_set_tiling(pxfdtd_resident *r, int32_t rows, int32_t planes) {
    if (r->pair) {  // the two-step march's pairs (single steps keep kind 3's tiling)
        if (rows && (rows < 2 || rows > 3))
            return px_fail(PXFDTD_ERR_INVALID, "rows per strip must be 2 or 3 for the two-step march, got %d", rows);
        if (planes < 0) return px_fail(PXFDTD_ERR_INVALID, "planes per chunk must be >= 1");
        if (rows) r->ty2 = rows;
        if (planes) r->kz2 = std::min(planes, r->nz);
        r->units2_dirty = true;
        return PXFDTD_OK;
    }
    if (rows) {
        const bool ok = r->kernel == 0 ? (rows == 2 || rows == 4 || rows == 8)
                        : r->kernel == 4 ? rows == 1  // the 2D march: one row at a time
                                         : (rows == 1 || rows == 2 || rows == 4);  // the march
        if (!ok)
            return px_fail(PXFDTD_ERR_INVALID, "rows per tile must be %s for this kernel, got %d",
                           r->kernel == 0 ? "2, 4 or 8" : "1, 2 or 4", rows);
        r->ty = rows;
    }
    if (planes) {
        if (planes < 1) return px_fail(PXFDTD_ERR_INVALID, "planes per chunk must be >= 1");
        r->kz = std::min(planes, r->kernel == 4 ? r->ny : r->nz);
    }
    r->units_dirty = true;
    HIPCHK(hipSetDevice(r->dev->ordinal));
    return build_injection_lists(r);
}

extern "C" int pxfdtd_resident_set_kernel(pxfdtd_resident *r, int32_t kind) {
    if (kind < 0 || kind > 5) return px_fail(PXFDTD_ERR_INVALID, "kernel kind must be 0 to 5");
    if (kind == 5) {
        if (r->nz == 1 || !kernel_fits(3, r->nval, r->psi_len, (long long)r->esz))
            return px_fail(PXFDTD_ERR_UNSUPPORTED, "the two-step march steps 3D grids whose buffers kernel 3 addresses");
        TRY(pxfdtd_resident_set_kernel(r, 3));
        if (!r->ps.fork) {
            HIPCHK(hipSetDevice(r->dev->ordinal));
            HIPCHK(hipEventCreateWithFlags(&r->ps.fork, hipEventDisableTiming));
            for (int c = 0; c < 4; c++) {
                HIPCHK(hipStreamCreate(&r->ps.s[c]));
                HIPCHK(hipEventCreateWithFlags(&r->ps.join[c], hipEventDisableTiming));
            }
        }
        r->pair = 1;
        r->units2_dirty = true;
        return PXFDTD_OK;
    }
    r->pair = 0;
    if (kind == 4 && !(r->nz == 1 && r->periodic[2]))
        return px_fail(PXFDTD_ERR_UNSUPPORTED, "the 2D march steps grids one cell thick with z periodic");
    if (!kernel_fits(kind, r->nval, r->psi_len, (long long)r->esz))
        return px_fail(PXFDTD_ERR_UNSUPPORTED,
                       kind >= 3 ? "kernel %d needs a field's three components and psi each within 2 GiB"
                                 : "kernel %d needs each field component within 2 GiB", kind);
    r->kernel = kind;
    r->ty = kind == 0 ? 8 : (kind == 4 ? 1 : 2);
    if (kind == 4) r->kz = default_rows_2d(r->nx, r->ny);
    r->units_dirty = true;
    HIPCHK(hipSetDevice(r->dev->ordinal));
    return build_injection_lists(r);
}

extern "C" int pxfdtd_resident_get_kernel(const pxfdtd_resident *r) { return r->pair ? 5 : r->kernel; }

extern "C" int pxfdtd_resident_set_concurrent(pxfdtd_resident *r, int32_t mode) {
    if (mode < 0 || mode > 2) return px_fail(PXFDTD_ERR_INVALID, "the launch mode must be 0, 1 or 2");
    if (mode != r->concurrent) r->units_dirty = true;
    r->concurrent = mode;
    return PXFDTD_OK;
}

extern "C" int pxfdtd_resident_get_tiling(const pxfdtd_resident *r, int32_t *rows, int32_t *planes) {
    *rows = r->pair ? r->ty2 : r->ty;
    *planes = r->pair ? r->kz2 : r->kz;
    return PXFDTD_OK;
}

// ---------------------------------------------------------------- probes and transforms
static int build_taps(pxfdtd_resident *r, const std::vector<int64_t> &index,
                      const std::vector<int8_t> &which) {
    r->units2_dirty = true;  // the two-step march's lists follow
    r->ntaps = (long long)index.size();
    r->tap_index_h = index;
    r->tap_which_h = which;
    if (!r->ntaps) return PXFDTD_OK;
    TRY(r->tap_index.alloc(index.size() * 8));
    TRY(r->tap_which.alloc(which.size()));
    HIPCHK(hipMemcpyAsync(r->tap_index.p, index.data(), index.size() * 8, hipMemcpyHostToDevice, r->stream));
    HIPCHK(hipMemcpyAsync(r->tap_which.p, which.data(), which.size(), hipMemcpyHostToDevice, r->stream));
    for (int c = 0; c < 2; c++) {
        void *ptrs[6];
        for (int q = 0; q < 6; q++) ptrs[q] = (char *)r->fields[c].p + (size_t)q * r->nval * r->esz;
        TRY(r->tap_fields[c].alloc(sizeof ptrs));
        HIPCHK(hipMemcpyAsync(r->tap_fields[c].p, ptrs, sizeof ptrs, hipMemcpyHostToDevice, r->stream));
    }
    TRY(r->staging.alloc((size_t)r->block_steps * r->ntaps * r->esz));
    HIPCHK(hipStreamSynchronize(r->stream));
    return PXFDTD_OK;
}

struct TapSpec {
    std::vector<pxfdtd_probe> probes;
    std::vector<pxfdtd_series> series;
    std::vector<int32_t> monitor_nf;
};

static int rebuild_taps(pxfdtd_resident *r, const TapSpec &t) {
    std::vector<int64_t> index;
    std::vector<int8_t> which;
    for (const auto &pr : t.probes) {
        const long long v = pr.index;
        const long long i = v % r->nx, rest = v / r->nx;
        index.push_back((rest / r->ny) * r->plane + (rest % r->ny) * r->pitch + i);
        which.push_back((int8_t)(pr.field * 3 + pr.component));
    }
    r->nprobes = (int)t.probes.size();
    std::vector<int32_t> series_of;
    std::vector<SeriesInfo> info;
    std::vector<int32_t> mbase(t.monitor_nf.size() + 1, 0);
    for (size_t m = 0; m < t.monitor_nf.size(); m++) mbase[m + 1] = mbase[m] + 2 * t.monitor_nf[m];
    r->slots = mbase.back();
    long long sums_off = 0, base_v = 0;
    for (size_t q = 0; q < t.series.size(); q++) {
        const pxfdtd_series &s = t.series[q];
        SeriesInfo in;
        in.base_v = base_v;
        in.volume = (long long)(s.hi[0] - s.lo[0]) * (s.hi[1] - s.lo[1]) * (s.hi[2] - s.lo[2]);
        in.sums_off = sums_off;
        in.nf = t.monitor_nf[s.monitor];
        in.slot = mbase[s.monitor] + (s.field == PXFDTD_H ? in.nf : 0);
        info.push_back(in);
        for (int k = s.lo[2]; k < s.hi[2]; k++)
            for (int j = s.lo[1]; j < s.hi[1]; j++)
                for (int i = s.lo[0]; i < s.hi[0]; i++) {
                    index.push_back((long long)k * r->plane + (long long)j * r->pitch + i);
                    which.push_back((int8_t)(s.field * 3 + s.component));
                    series_of.push_back((int32_t)q);
                }
        base_v += in.volume;
        sums_off += in.volume * in.nf;
    }
    r->nvals = base_v;
    r->sums_len = 2 * sums_off;
    TRY(build_taps(r, index, which));
    if (r->nvals) {
        TRY(r->series_of.alloc(series_of.size() * 4));
        TRY(r->series_info.alloc(info.size() * sizeof(SeriesInfo)));
        TRY(r->sums.alloc(r->sums_len * 8));
        HIPCHK(hipMemcpyAsync(r->series_of.p, series_of.data(), series_of.size() * 4, hipMemcpyHostToDevice, r->stream));
        HIPCHK(hipMemcpyAsync(r->series_info.p, info.data(), info.size() * sizeof(SeriesInfo), hipMemcpyHostToDevice, r->stream));
        HIPCHK(hipMemsetAsync(r->sums.p, 0, r->sums_len * 8, r->stream));
        HIPCHK(hipStreamSynchronize(r->stream));
    }
    return PXFDTD_OK;
}

// the resident keeps its tap specification to rebuild when either part changes; residents are
// made and freed on several host threads at once (the server's admissions), so the registry is
// locked (an entry, once made, stays where it is until its resident goes)
static std::mutex &specs_mutex() {
    static std::mutex m;
    return m;
}
static std::map<const pxfdtd_resident *, TapSpec> &specs() {
    static std::map<const pxfdtd_resident *, TapSpec> m;
    return m;
}
static TapSpec &spec_of(const pxfdtd_resident *r) {
    std::lock_guard<std::mutex> g(specs_mutex());
    return specs()[r];
}

extern "C" int pxfdtd_resident_set_probes(pxfdtd_resident *r, const pxfdtd_probe *probes, int32_t n) {
    for (int q = 0; q < n; q++)
        if (probes[q].index < 0 || probes[q].index >= r->ncells || probes[q].field < 0 ||
            probes[q].field > 1 || probes[q].component < 0 || probes[q].component > 2)
            return px_fail(PXFDTD_ERR_INVALID, "probe %d is out of range", q);
    HIPCHK(hipSetDevice(r->dev->ordinal));
    HIPCHK(hipStreamSynchronize(r->stream));
    TapSpec &t = spec_of(r);
    t.probes.assign(probes, probes + n);
    r->probe_rows = 0;
    return rebuild_taps(r, t);
}

extern "C" int pxfdtd_resident_set_series(pxfdtd_resident *r, const pxfdtd_series *series,
                                          int32_t n, const int32_t *monitor_nf, int32_t monitors) {
    const int dims[3] = {r->nx, r->ny, r->nz};
    for (int q = 0; q < n; q++) {
        const pxfdtd_series &s = series[q];
        if (s.monitor < 0 || s.monitor >= monitors || s.field < 0 || s.field > 1 ||
            s.component < 0 || s.component > 2)
            return px_fail(PXFDTD_ERR_INVALID, "series %d is out of range", q);
        for (int a = 0; a < 3; a++)
            if (s.lo[a] < 0 || s.hi[a] > dims[a] || s.lo[a] >= s.hi[a])
                return px_fail(PXFDTD_ERR_INVALID, "series %d's box is not on the grid", q);
    }
    for (int m = 0; m < monitors; m++)
        if (monitor_nf[m] < 1) return px_fail(PXFDTD_ERR_INVALID, "a monitor needs frequencies");
    HIPCHK(hipSetDevice(r->dev->ordinal));
    HIPCHK(hipStreamSynchronize(r->stream));
    TapSpec &t = spec_of(r);
    t.series.assign(series, series + n);
    t.monitor_nf.assign(monitor_nf, monitor_nf + monitors);
    return rebuild_taps(r, t);
}

extern "C" int32_t pxfdtd_resident_phase_slots(const pxfdtd_resident *r) { return r->slots; }
extern "C" int64_t pxfdtd_resident_sums_len(const pxfdtd_resident *r) { return r->sums_len; }

extern "C" int pxfdtd_resident_read_sums(pxfdtd_resident *r, double *out, int64_t capacity) {
    if (capacity < r->sums_len)
        return px_fail(PXFDTD_ERR_INVALID, "the sums need %lld values", (long long)r->sums_len);
    HIPCHK(hipSetDevice(r->dev->ordinal));
    if (r->sums_len)
        HIPCHK(hipMemcpyAsync(out, r->sums.p, r->sums_len * 8, hipMemcpyDeviceToHost, r->stream));
    HIPCHK(hipStreamSynchronize(r->stream));
    return PXFDTD_OK;
}

extern "C" int pxfdtd_resident_write_sums(pxfdtd_resident *r, const double *in, int64_t count) {
    if (count != r->sums_len)
        return px_fail(PXFDTD_ERR_INVALID, "the sums are %lld values, not %lld", (long long)r->sums_len,
                       (long long)count);
    HIPCHK(hipSetDevice(r->dev->ordinal));
    if (r->sums_len)
        HIPCHK(hipMemcpyAsync(r->sums.p, in, r->sums_len * 8, hipMemcpyHostToDevice, r->stream));
    HIPCHK(hipStreamSynchronize(r->stream));
    return PXFDTD_OK;
}

template <typename T>
static int read_probes_t(pxfdtd_resident *r, double *out, int64_t capacity, int64_t *steps) {
    *steps = r->probe_rows;
    if (!r->nprobes || !r->probe_rows) return PXFDTD_OK;
    if (capacity < r->probe_rows * r->nprobes)
        return px_fail(PXFDTD_ERR_INVALID, "the probes need %lld values", (long long)(r->probe_rows * r->nprobes));
    std::vector<T> buf(r->probe_rows * r->nprobes);
    HIPCHK(hipMemcpyAsync(buf.data(), r->probe_rec.p, buf.size() * sizeof(T), hipMemcpyDeviceToHost, r->stream));
    HIPCHK(hipStreamSynchronize(r->stream));
    for (size_t q = 0; q < buf.size(); q++) out[q] = (double)buf[q];
    r->probe_rows = 0;
    return PXFDTD_OK;
}

extern "C" int64_t pxfdtd_resident_probe_steps(const pxfdtd_resident *r) { return r->probe_rows; }

extern "C" int pxfdtd_resident_read_probes(pxfdtd_resident *r, double *out, int64_t capacity,
                                           int64_t *steps) {
    HIPCHK(hipSetDevice(r->dev->ordinal));
    return r->prec == PXFDTD_F64 ? read_probes_t<double>(r, out, capacity, steps)
                                 : read_probes_t<float>(r, out, capacity, steps);
}

// -------------------------------------------------------------------------- stepping
template <typename T>
static StepArgs<T> args_of(const pxfdtd_resident *r) {
    StepArgs<T> A;
    memset(&A, 0, sizeof A);
    A.nx = r->nx; A.ny = r->ny; A.nz = r->nz; A.pitch = r->pitch; A.plane = (int)r->plane;
    A.nval = (int)r->nval;
    bool slabs = false;
    for (int a = 0; a < 3; a++) {
        A.flags |= r->periodic[a] << a;
        A.step[a] = sizeof(T) == 4 ? (T)(1.0 / r->step[a]) : (T)r->step[a];  // see per_step
        slabs = slabs || r->low[a] || r->high[a];
    }
    if (slabs) A.flags |= 8;
    if (r->have_inj) A.flags |= 16;
    A.coef = r->coef.as<T>();
    A.ncpml = r->ncpml;
    int tw, th;
    tile_shape(r, tw, th);
    A.tiles_x = (r->nx + tw - 1) / tw;
    A.tiles_y = (r->ny + th - 1) / th;
    A.tile_rows = th;
    A.kz = r->kz;
    A.blocks = A.tiles_x * A.tiles_y * ((r->nz + r->kz - 1) / r->kz);
    A.per_xcd = (A.blocks + 7) / 8;
    if (r->kernel != 0) {  // the march's two unit lists: counts ride in blocks and per_xcd
        A.units = r->units.as<int4>();
        A.blocks = r->n_lean;
        A.per_xcd = r->n_slab;
    }
    A.dt = (T)r->dt;
    A.ids = r->ids.as<uint8_t>();
    A.table = r->table.as<T>();
    A.cacb = r->cacb.as<T>();
    A.grp = r->has_grp ? r->grpb.as<int8_t>() : nullptr;
    return A;
}

// the parameters slab cells, injections and media read, one set per copy parity
template <typename T>
// the step kernels' cold parameters per parity: [0, 1] the one-step kernels', [2, 3] the
// two-step march's (its tiles' source lists)
static void fill_cold(const pxfdtd_resident *r, StepCold<T> *K, int n) {
    memset(K, 0, n * sizeof(StepCold<T>));
    for (int p = 0; p < n; p++) {
        StepCold<T> &k = K[p];
        for (int a = 0; a < 3; a++) {
            k.low[a] = r->low[a];
            k.high[a] = r->high[a];
        }
        memcpy(k.cpml_off, r->cpml_off, sizeof k.cpml_off);
        memcpy(k.psi_off, r->psi_off, sizeof k.psi_off);
        k.psi_in = r->psi[p & 1].as<T>();
        k.psi_out = r->psi[1 - (p & 1)].as<T>();
        k.inj_off = !r->have_inj ? nullptr : (p < 2 ? r->inj_off.as<int32_t>() : r->inj2_off.as<int32_t>());
        k.inj = p < 2 ? r->inj.as<Injection>() : r->inj2.as<Injection>();
        k.waves = r->waves.as<double>();
        k.emitters = r->emitters;
        k.groups = r->groups.as<MediumGroup>();
        k.max_poles = r->max_poles;
        k.m_before = r->m_before.as<T>();
        k.m_cur = r->m_cur.as<T>();
        k.m_prev = r->m_prev.as<T>();
        k.psi_len = (int)r->psi_len;
        k.aux = r->aux.as<double>();
    }
}

template <typename T>
static int upload_cold(pxfdtd_resident *r) {
    StepCold<T> K[4];
    fill_cold<T>(r, K, 4);
    TRY(r->cold.reserve(sizeof K));
    HIPCHK(hipMemcpy(r->cold.p, K, sizeof K, hipMemcpyHostToDevice));
    return PXFDTD_OK;
}

// the 2D march's field sets to step: those with values or sources (TE if none has either)
static int step_set(const pxfdtd_resident *r) {
    const int set = r->fields_set | r->inj_set;
    return set ? set : 1;
}

template <typename T>
static int step_t(pxfdtd_resident *r, uint32_t count, const double *waves, const double *phases) {
    hipStream_t s = r->stream;
    if (r->emitters && (r->have_inj || !r->planes_h.empty())) {
        if (!waves) return px_fail(PXFDTD_ERR_INVALID, "the injections need their waves");
        const size_t b = (size_t)count * r->emitters * 2 * 8;
        TRY(r->waves.reserve(b));
        HIPCHK(hipMemcpyAsync(r->waves.p, waves, b, hipMemcpyHostToDevice, s));
    }
    if (r->nvals) {
        if (!phases) return px_fail(PXFDTD_ERR_INVALID, "the transforms need their phases");
        const size_t b = (size_t)count * r->slots * 2 * 8;
        TRY(r->phases.reserve(b));
        HIPCHK(hipMemcpyAsync(r->phases.p, phases, b, hipMemcpyHostToDevice, s));
    }
    if (r->nprobes) {
        const long long need = (r->probe_rows + count) * r->nprobes * (long long)sizeof(T);
        if ((long long)r->probe_rec.bytes < need) {
            Buf grown;
            TRY(grown.alloc((size_t)std::max<long long>(need, 2 * (long long)r->probe_rec.bytes)));
            if (r->probe_rows)
                HIPCHK(hipMemcpyAsync(grown.p, r->probe_rec.p, r->probe_rows * r->nprobes * sizeof(T),
                                      hipMemcpyDeviceToDevice, s));
            HIPCHK(hipStreamSynchronize(s));
            grown.swap(r->probe_rec);
        }
    }
    const size_t lds = r->mode == 0 ? (size_t)256 * (6 * sizeof(T) + 3) : 0;
    // the two-step march's lists before the cold parameters, which point at them
    if (r->pair && r->units2_dirty) TRY(build_pair_tables(r));
    TRY(upload_cold<T>(r));
    if (r->kernel != 0 && r->units_dirty) TRY(build_units(r));
    int in_block = 0;
    long long block_first = 0;
    if (r->profile)
        while (r->kev.size() < 2 * (size_t)count) {
            hipEvent_t e;
            HIPCHK(hipEventCreate(&e));
            r->kev.push_back(e);
        }
    HIPCHK(hipEventRecord(r->ev0, s));
    // the aligned march with a side stream takes each step's taps (and a block's transform
    // sums and probe rows) into the next step, on the side stream: they read the copy that step
    // reads and writes nothing it reads, and the step's join orders them before the step after
    // it, which overwrites that copy
    const bool defer_taps = r->kernel == 3 && r->concurrent == 1;
    std::function<int(hipStream_t)> pending;
    for (uint32_t n = 0; n < count; n++) {
        if (pending && (n + 1 < count && pair_ok(r, in_block))) {
            TRY(pending(s));
            pending = nullptr;
        }
        if (n + 1 < count && pair_ok(r, in_block)) {  // steps n and n + 1 in one pass
            StepArgs<T> A = args_of<T>(r);
            const int p = r->p, q = 1 - p;
            A.ein = r->field<T>(p, 0);
            A.hin = r->field<T>(p, 3);
            A.eout = r->field<T>(q, 0);
            A.hout = r->field<T>(q, 3);
            A.cold = r->cold.as<StepCold<T>>() + 2 + p;
            A.wave_step = (int)n;
            A.units = r->units2.as<int4>();
            A.tiles_x = r->tiles2_x;
            const bool tapping = r->ntaps && !r->quiet;
            PairTaps<T> TP;
            TP.off = tapping && r->have_tap2 ? r->tap2_off.as<int32_t>() : nullptr;
            TP.at = r->tap2.as<TapAt>();
            TP.stage = r->staging.as<T>() + (size_t)in_block * r->ntaps;
            if (r->profile) HIPCHK(hipEventRecord(r->kev[2 * n], s));
            const PairStreams *sp = r->concurrent == 1 && r->ps.fork ? &r->ps : nullptr;
            const PairCounts uc = {r->n2[0], r->n2[1], r->n2[2], r->n2[3], r->n2[4]};
            if (r->mode == 0)
                TRY((launch_pair<T, 0>(A, TP, r->ty2, s, lds, sp, uc)));
            else if (r->mode == 2)
                TRY((launch_pair<T, 2>(A, TP, r->ty2, s, lds, sp, uc)));
            else
                TRY((launch_pair<T, 1>(A, TP, r->ty2, s, lds, sp, uc)));
            HIPCHK(hipGetLastError());
            if (r->profile) {
                HIPCHK(hipEventRecord(r->kev[2 * n + 1], s));
                HIPCHK(hipEventRecord(r->kev[2 * n + 2], s));
                HIPCHK(hipEventRecord(r->kev[2 * n + 3], s));
            }
            r->p = q;  // the pass read copy p and wrote copy 1 - p
            if (tapping) {  // the second step's monitored values, from the fields stored
                const int threads = 256;
                const long long blocks = (r->ntaps + threads - 1) / threads;
                hipLaunchKernelGGL((tap_values<T>), dim3((unsigned)blocks), dim3(threads), 0, s,
                                   r->tap_fields[q].as<const T *const>(), r->tap_index.as<int64_t>(),
                                   r->tap_which.as<int8_t>(), r->ntaps,
                                   r->staging.as<T>() + (size_t)(in_block + 1) * r->ntaps);
                HIPCHK(hipGetLastError());
                in_block += 2;
                if (in_block == r->block_steps || n + 2 == count) {
                    if (r->nvals) {
                        const long long blocks2 = (r->nvals + threads - 1) / threads;
                        hipLaunchKernelGGL((dft_accumulate<T>), dim3((unsigned)blocks2), dim3(threads), 0, s,
                                           r->staging.as<T>() + r->nprobes, r->ntaps, r->nvals, in_block,
                                           r->series_of.as<int32_t>(), r->series_info.as<SeriesInfo>(),
                                           r->phases.as<double>() + (size_t)block_first * r->slots * 2,
                                           r->slots, r->sums.as<double>());
                        HIPCHK(hipGetLastError());
                    }
                    if (r->nprobes) {
                        HIPCHK(hipMemcpy2DAsync(r->probe_rec.as<T>() + (size_t)r->probe_rows * r->nprobes,
                                                r->nprobes * sizeof(T), r->staging.p, r->ntaps * sizeof(T),
                                                r->nprobes * sizeof(T), in_block, hipMemcpyDeviceToDevice, s));
                        r->probe_rows += in_block;
                    }
                    block_first = n + 2;
                    in_block = 0;
                }
            }
            n++;
            continue;
        }
        StepArgs<T> A = args_of<T>(r);
        const int p = r->p, q = 1 - p;
        A.ein = r->field<T>(p, 0);
        A.hin = r->field<T>(p, 3);
        A.eout = r->field<T>(q, 0);
        A.hout = r->field<T>(q, 3);
        A.cold = r->cold.as<StepCold<T>>() + p;
        A.wave_step = (int)n;
        if (r->nmvals) {  // the media after the step kernel, in a pass of their own
            A.flags |= 32;
            A.grp = nullptr;
        }
        if (r->profile) HIPCHK(hipEventRecord(r->kev[2 * n], s));
        SideStream side = {r->side, r->fork_ev, r->join_ev};
        side.first = std::move(pending);
        pending = nullptr;
        const SideStream *sp = r->concurrent == 1 ? &side : nullptr;
        if (side.first && !(r->kernel == 3 && sp)) TRY(side.first(s));
        const UnitCounts uc = {r->n_lean, r->n_x, r->n_slab, r->slab_rows};
        const int set = step_set(r);
        const dim3 aux_grid((unsigned)((r->aux_max + 255) / 256), (unsigned)r->planes_h.size());
        // the plane waves' H~ to t + dt/2 before the grid's step reads it (E's couplings);
        // their E stays at t for H's couplings until after the step
        if (!r->planes_h.empty())
            hipLaunchKernelGGL(aux_step_h, aux_grid, dim3(256), 0, s, r->plane_desc.as<PlaneDesc>(),
                               r->aux.as<double>(), r->aux_coef.as<double>());
        if (r->mode == 0)
            TRY((launch_mode<T, 0>(A, r->kernel, r->ty, s, lds, sp, uc, set)));
        else if (r->mode == 2)
            TRY((launch_mode<T, 2>(A, r->kernel, r->ty, s, lds, sp, uc, set)));
        else
            TRY((launch_mode<T, 1>(A, r->kernel, r->ty, s, lds, sp, uc, set)));
        HIPCHK(hipGetLastError());
        if (r->nmvals) {
            hipLaunchKernelGGL((media_pass<T>), dim3((unsigned)((r->nmvals + 255) / 256)), dim3(256), 0, s, A,
                               r->mvals.as<uint32_t>(), r->mgrp.as<int8_t>(), r->mcb.as<T>(), r->nmvals);
            HIPCHK(hipGetLastError());
        }
        if (r->profile) HIPCHK(hipEventRecord(r->kev[2 * n + 1], s));
        if (!r->planes_h.empty()) {
            hipLaunchKernelGGL(aux_step_e, aux_grid, dim3(256), 0, s, r->plane_desc.as<PlaneDesc>(),
                               r->aux.as<double>(), r->aux_coef.as<double>(), r->waves.as<double>(),
                               r->emitters, (int)n);
            HIPCHK(hipGetLastError());
        }
        r->p = q;
        if (r->after_step) {
            r->after_step(s, n);
            HIPCHK(hipGetLastError());
        }
        if (r->ntaps && !r->quiet) {
            const T *const *tf = r->tap_fields[q].as<const T *const>();
            T *slot = r->staging.as<T>() + (size_t)in_block * r->ntaps;
            in_block++;
            const bool block_end = in_block == r->block_steps || n + 1 == count;
            const int rows = in_block;
            const double *ph = r->phases.as<double>() + (size_t)block_first * r->slots * 2;
            T *rec = r->probe_rec.as<T>() + (size_t)r->probe_rows * r->nprobes;
            auto work = [r, tf, slot, block_end, rows, ph, rec](hipStream_t on) -> int {
                const int threads = 256;
                const long long blocks = (r->ntaps + threads - 1) / threads;
                hipLaunchKernelGGL((tap_values<T>), dim3((unsigned)blocks), dim3(threads), 0, on, tf,
                                   r->tap_index.as<int64_t>(), r->tap_which.as<int8_t>(), r->ntaps, slot);
                HIPCHK(hipGetLastError());
                if (block_end && r->nvals) {
                    const long long blocks2 = (r->nvals + threads - 1) / threads;
                    hipLaunchKernelGGL((dft_accumulate<T>), dim3((unsigned)blocks2), dim3(threads), 0, on,
                                       r->staging.as<T>() + r->nprobes, r->ntaps, r->nvals, rows,
                                       r->series_of.as<int32_t>(), r->series_info.as<SeriesInfo>(), ph,
                                       r->slots, r->sums.as<double>());
                    HIPCHK(hipGetLastError());
                }
                if (block_end && r->nprobes)
                    HIPCHK(hipMemcpy2DAsync(rec, r->nprobes * sizeof(T), r->staging.p, r->ntaps * sizeof(T),
                                            r->nprobes * sizeof(T), rows, hipMemcpyDeviceToDevice, on));
                return PXFDTD_OK;
            };
            if (defer_taps)
                pending = work;
            else
                TRY(work(s));
            if (block_end) {
                if (r->nprobes) r->probe_rows += in_block;
                block_first = n + 1;
                in_block = 0;
            }
        }
    }
    if (pending) TRY(pending(s));
    HIPCHK(hipEventRecord(r->ev1, s));
    HIPCHK(hipEventSynchronize(r->ev1));
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, r->ev0, r->ev1));
    r->last_ms = count ? ms / count : 0.0;
    if (r->profile) {
        double sum = 0;
        for (uint32_t n = 0; n < count; n++) {
            float k = 0;
            HIPCHK(hipEventElapsedTime(&k, r->kev[2 * n], r->kev[2 * n + 1]));
            sum += k;
        }
        r->kernel_ms = sum / count;
    }
    return PXFDTD_OK;
}

// ------------------------------------------------------------------------------ tuning
// `trials` launches of the step kernel that read copy p and write copy 1 - p without taking the
// step (the parity stays, and psi ping-pongs as the fields do), so the state the next real step
// reads is unchanged; injections are left out. Average ms per launch.
template <typename T>
static int trial_ms(pxfdtd_resident *r, int trials, double *ms) {
    hipStream_t s = r->stream;
    const size_t lds = r->mode == 0 ? (size_t)256 * (6 * sizeof(T) + 3) : 0;
    TRY(upload_cold<T>(r));
    if (r->kernel != 0 && r->units_dirty) TRY(build_units(r));
    StepArgs<T> A = args_of<T>(r);
    A.flags &= ~16;
    const int p = r->p, q = 1 - p;
    A.ein = r->field<T>(p, 0);
    A.hin = r->field<T>(p, 3);
    A.eout = r->field<T>(q, 0);
    A.hout = r->field<T>(q, 3);
    A.cold = r->cold.as<StepCold<T>>() + p;
    A.wave_step = 0;
    const SideStream side = {r->side, r->fork_ev, r->join_ev};
    const SideStream *sp = r->concurrent == 1 ? &side : nullptr;
    const UnitCounts uc = {r->n_lean, r->n_x, r->n_slab, r->slab_rows};
    const int set = step_set(r);
    for (int n = 0; n <= trials; n++) {
        if (n == 1) HIPCHK(hipEventRecord(r->ev0, s));
        if (r->mode == 0)
            TRY((launch_mode<T, 0>(A, r->kernel, r->ty, s, lds, sp, uc, set)));
        else if (r->mode == 2)
            TRY((launch_mode<T, 2>(A, r->kernel, r->ty, s, lds, sp, uc, set)));
        else
            TRY((launch_mode<T, 1>(A, r->kernel, r->ty, s, lds, sp, uc, set)));
        HIPCHK(hipGetLastError());
    }
    HIPCHK(hipEventRecord(r->ev1, s));
    HIPCHK(hipEventSynchronize(r->ev1));
    float t = 0;
    HIPCHK(hipEventElapsedTime(&t, r->ev0, r->ev1));
    *ms = trials ? t / trials : 0.0;
    return PXFDTD_OK;
}

extern "C" int pxfdtd_resident_autotune(pxfdtd_resident *r, int32_t trials, int32_t max_rows,
                                        int32_t *rows, int32_t *planes, double *best_ms) {
    HIPCHK(hipSetDevice(r->dev->ordinal));
    HIPCHK(hipStreamSynchronize(r->stream));
    *best_ms = 0.0;
    // dispersive values update their state in place: a trial would advance it, so those keep
    // the defaults; so do the kernels other than the aligned march
    if (r->ngroups || (r->kernel != 3 && r->kernel != 4) || trials < 1) {
        *rows = r->ty;
        *planes = r->kz;
        return PXFDTD_OK;
    }
    std::vector<std::pair<int, int>> cand;
    if (r->kernel == 4) {  // the 2D march: rows per chunk
        for (int kz : {2, 4, 8, 16, 32}) {
            const int k = std::min(kz, r->ny);
            if (std::find(cand.begin(), cand.end(), std::make_pair(1, k)) == cand.end()) cand.push_back({1, k});
        }
    }
    for (int ry : {1, 2, 4}) {
        if (r->kernel != 3 || ry > std::max(1, (int)max_rows)) continue;
        for (int kz : {4, 8, 16}) {
            const int k = std::min(kz, r->nz);
            if (std::find(cand.begin(), cand.end(), std::make_pair(ry, k)) == cand.end()) cand.push_back({ry, k});
        }
    }
    double best = 1e300;
    std::pair<int, int> pick = {r->ty, r->kz};
    for (auto c : cand) {
        TRY(pxfdtd_resident_set_tiling(r, c.first, c.second));
        double ms = 0;
        TRY(r->prec == PXFDTD_F64 ? trial_ms<double>(r, trials, &ms) : trial_ms<float>(r, trials, &ms));
        if (ms < best) {
            best = ms;
            pick = c;
        }
    }
    TRY(pxfdtd_resident_set_tiling(r, pick.first, pick.second));
    *rows = pick.first;
    *planes = pick.second;
    *best_ms = best;
    return PXFDTD_OK;
}

extern "C" int pxfdtd_resident_set_profiling(pxfdtd_resident *r, int32_t on) {
    r->profile = on != 0;
    return PXFDTD_OK;
}

extern "C" double pxfdtd_resident_kernel_ms(const pxfdtd_resident *r) { return r->kernel_ms; }

extern "C" int pxfdtd_resident_step(pxfdtd_resident *r, uint32_t count, const double *waves,
                                    const double *phases) {
    HIPCHK(hipSetDevice(r->dev->ordinal));
    if (count == 0) return PXFDTD_OK;
    return r->prec == PXFDTD_F64 ? step_t<double>(r, count, waves, phases)
                                 : step_t<float>(r, count, waves, phases);
}

// ------------------------------------------------------------------ the serving loop
#include "serving.hpp"

// ------------------------------------------------------------- the reverse-time adjoint
#include "adjoint.hpp"
#include "adjoint_host.hpp"
