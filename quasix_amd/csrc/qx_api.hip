// C ABI of the G0W0/evGW correlation path (see include/qxgw.h).
//
// Host driver for the HIP kernels: owns the stream and a grow-only device workspace,
// sequences Pi -> Cholesky -> L^-1 -> Wc_nm -> Newton for one solve, runs the numerical
// fallbacks of the reference (Cholesky -> LU -> identity, correlation_ac.rs:376-401;
// Newton -> bisection -> linearised, newton_solver.rs:262-268 and evgw.rs:766-781), and
// drives the residue evaluations of the contour deformation when a Newton iterate crosses
// an orbital energy (the only part of a solve that needs host decisions per iteration).
#include "../../include/qxgw.h"
#include "qx_internal.h"
#include "qx_newton.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <exception>
#include <thread>
#include <condition_variable>
#include <string>
#include <tuple>
#include <vector>

using namespace qx;

namespace {
struct BatchPlan;
void delete_batch_plan(BatchPlan* p);
}  // namespace

struct qx_context {
    int device = 0;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    std::string err;
    std::map<std::string, std::pair<void*, size_t>> bufs;
    // profiling: HIP events around each phase, resolved after the solve's final sync
    bool prof = false;
    double phase_ms[QX_NPHASE] = {0};
    long long phase_n[QX_NPHASE] = {0};
    std::vector<hipEvent_t> ev_free;
    std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> ev_pending;
    // autotuner: best kernel variant per (kind, precision, shape) key
    std::map<std::string, int> tuned;
    bool tune_loaded = false;
    // side streams for the grouped factorisation (created on first use)
    StreamFan fan;
    // last batched solve: work tables, workspace layout and its captured graph
    BatchPlan* batch = nullptr;
    // no-allocation mode (set for the duration of a *_noalloc call): the workspace is carved
    // from the device's static scratch region instead of being allocated
    char* arena = nullptr;
    size_t arena_size = 0, arena_used = 0;
    std::map<std::string, std::pair<void*, size_t>> arena_bufs;
    char* scratch_dev = nullptr;          // this device's copy of `scratch` (set up in qx_create)
    hipError_t scratch_status = hipErrorNotInitialized;
    // pinned host ring for the small per-solve transfers (allocated on first use, never in
    // no-allocation mode): pin_off is the next free byte, reset only after the stream has
    // been synchronised; pin_hold > 0 while a batch of downloads still holds ring space
    char* pin = nullptr;
    size_t pin_cap = 0, pin_off = 0;
    int pin_hold = 0;
    // the CD frequency grid of the last solve (buffers "cd_fr"/"cd_wt", written only by
    // run_g0w0) and the zeroed static terms ("zero_static"): re-uploaded only when the key
    // or the buffer changes; never used in no-allocation mode
    int grid_nf = 0;
    double grid_x0 = 0.0;
    const double* grid_fr = nullptr;
    const double* zero_ptr = nullptr;
    int zero_n = 0;
};

// Device scratch of the no-allocation entry point (qx_g0w0_cd_ov_noalloc_dev): every GPU the
// library runs on holds one copy for the life of the process.  Host code never uses this
// name as an address: qx_create resolves each device's copy once with hipGetSymbolAddress
// and the entry point passes that pointer (and offsets into it) to the kernels.
static __device__ __align__(256) unsigned char scratch[QXGW_NOALLOC_SCRATCH_BYTES];

namespace {

struct QxFail {
    qx_status st;
};

// Per-device address of `scratch`, resolved once per device with that device current.
struct ScratchAddr {
    std::mutex mu;
    std::map<int, std::pair<hipError_t, char*>> dev;
};
ScratchAddr& scratch_table() {
    static ScratchAddr t;
    return t;
}
std::pair<hipError_t, char*> scratch_address(int device) {
    ScratchAddr& t = scratch_table();
    std::lock_guard<std::mutex> lk(t.mu);
    auto it = t.dev.find(device);
    if (it != t.dev.end()) return it->second;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    t.dev[device] = {e, e == hipSuccess ? static_cast<char*>(p) : nullptr};
    return t.dev[device];
}

void set_err(qx_handle h, const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    h->err = buf;
}

#define HIPCK(h, expr)                                                                    \
    do {                                                                                  \
        hipError_t e_ = (expr);                                                           \
        if (e_ != hipSuccess) {                                                           \
            set_err(h, "HIP error %s at %s:%d (%s)", hipGetErrorString(e_), __FILE__,     \
                    __LINE__, #expr);                                                     \
            throw QxFail{QX_ERR_DEVICE};                                                  \
        }                                                                                 \
    } while (0)

#define REQUIRE(h, cond, code, ...)                                                       \
    do {                                                                                  \
        if (!(cond)) {                                                                    \
            set_err(h, __VA_ARGS__);                                                      \
            throw QxFail{code};                                                           \
        }                                                                                 \
    } while (0)

hipEvent_t take_event(qx_handle h) {
    if (!h->ev_free.empty()) {
        hipEvent_t e = h->ev_free.back();
        h->ev_free.pop_back();
        return e;
    }
    hipEvent_t e;
    HIPCK(h, hipEventCreate(&e));
    return e;
}

// Records start/stop events of one phase on the handle's stream while profiling is on.
struct Phase {
    qx_handle h;
    int id;
    hipEvent_t a = nullptr;
    Phase(qx_handle h_, int id_) : h(h_), id(id_) {
        if (h->prof) {
            a = take_event(h);
            HIPCK(h, hipEventRecord(a, h->stream));
        }
    }
    ~Phase() {
        if (!a) return;
        hipEvent_t b = take_event(h);
        if (hipEventRecord(b, h->stream) == hipSuccess) h->ev_pending.push_back({id, {a, b}});
    }
};

void resolve_profile(qx_handle h) {
    if (h->ev_pending.empty()) return;
    HIPCK(h, hipStreamSynchronize(h->stream));
    for (auto& p : h->ev_pending) {
        float ms = 0.f;
        HIPCK(h, hipEventElapsedTime(&ms, p.second.first, p.second.second));
        h->phase_ms[p.first] += ms;
        h->phase_n[p.first] += 1;
        h->ev_free.push_back(p.second.first);
        h->ev_free.push_back(p.second.second);
    }
    h->ev_pending.clear();
}

// ---- autotuner ----------------------------------------------------------------------------
// Kernel variants are timed once per shape key on the live operands (HIP events on the
// handle's stream; the output of a timing launch is overwritten by the real launch) and
// the fastest is cached in the handle.  QXGW_AUTOTUNE=0 disables timing (variant 0);
// QXGW_PI_VARIANT forces a variant; QXGW_AUTOTUNE_DB names a JSON-lines file that seeds
// and records the choices ({"key": ..., "variant": ..., "ms": [...]}).
const char* tune_db_path() { return std::getenv("QXGW_AUTOTUNE_DB"); }

void tune_load(qx_handle h) {
    if (h->tune_loaded) return;
    h->tune_loaded = true;
    const char* path = tune_db_path();
    if (!path) return;
    FILE* fp = std::fopen(path, "r");
    if (!fp) return;
    char line[1024];
    while (std::fgets(line, sizeof line, fp)) {
        char key[512];
        int v = -1;
        if (std::sscanf(line, "{\"key\": \"%511[^\"]\", \"variant\": %d", key, &v) == 2 && v >= 0)
            h->tuned[key] = v;
    }
    std::fclose(fp);
}

void tune_store(const std::string& key, int v, const std::vector<float>& ms) {
    const char* path = tune_db_path();
    if (!path) return;
    FILE* fp = std::fopen(path, "a");
    if (!fp) return;
    std::fprintf(fp, "{\"key\": \"%s\", \"variant\": %d, \"ms\": [", key.c_str(), v);
    for (size_t i = 0; i < ms.size(); ++i) std::fprintf(fp, i ? ", %.4f" : "%.4f", ms[i]);
    std::fprintf(fp, "]}\n");
    std::fclose(fp);
}

template <typename T>
T* ws(qx_handle h, const std::string& name, size_t count) {
    const size_t bytes = std::max<size_t>(count * sizeof(T), 256);
    if (h->arena) {
        // no-allocation mode: named buffers are carved from the fixed scratch region
        auto at = h->arena_bufs.find(name);
        if (at != h->arena_bufs.end() && at->second.second >= bytes)
            return static_cast<T*>(at->second.first);
        const size_t off = (h->arena_used + 255) / 256 * 256;
        REQUIRE(h, off <= h->arena_size && bytes <= h->arena_size - off, QX_ERR_ALLOC,
                "problem too large for the %zu-byte scratch: buffer '%s' needs %zu bytes "
                "with %zu in use",
                h->arena_size, name.c_str(), bytes, off);
        void* p = h->arena + off;
        h->arena_bufs[name] = {p, bytes};
        h->arena_used = off + bytes;
        return static_cast<T*>(p);
    }
    auto it = h->bufs.find(name);
    if (it != h->bufs.end() && it->second.second >= bytes) return static_cast<T*>(it->second.first);
    if (it != h->bufs.end()) {
        HIPCK(h, hipStreamSynchronize(h->stream));
        HIPCK(h, hipFree(it->second.first));
        h->bufs.erase(it);
    }
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) {
        set_err(h, "device allocation of %zu bytes for '%s' failed: %s", bytes, name.c_str(),
                hipGetErrorString(e));
        throw QxFail{QX_ERR_ALLOC};
    }
    h->bufs[name] = {p, bytes};
    return static_cast<T*>(p);
}

// Small transfers go through the pinned ring: a copy from pageable memory costs a host round
// trip of tens of microseconds each, during which the GPU idles between solves.  An upload is
// copied into the ring and queued without waiting; a download waits once.  The ring is
// reused from its start after a stream synchronisation, when every queued upload has been
// read.  nullptr: no ring space (no-allocation mode, too large, or a pending batch that
// holds it); the caller then copies directly.
constexpr size_t kPinBytes = 1 << 20;

char* pin_take(qx_handle h, size_t bytes) {
    if (h->arena) return nullptr;
    bytes = (bytes + 63) / 64 * 64;
    if (!h->pin) {
        void* p = nullptr;
        if (hipHostMalloc(&p, kPinBytes) != hipSuccess) return nullptr;
        h->pin = static_cast<char*>(p);
        h->pin_cap = kPinBytes;
        h->pin_off = 0;
    }
    if (bytes > h->pin_cap) return nullptr;
    if (h->pin_off + bytes > h->pin_cap) {
        if (h->pin_hold > 0) return nullptr;
        HIPCK(h, hipStreamSynchronize(h->stream));
        h->pin_off = 0;
    }
    char* p = h->pin + h->pin_off;
    h->pin_off += bytes;
    return p;
}

template <typename T>
void d2h(qx_handle h, T* dst, const T* src, size_t n) {
    if (n == 0) return;
    const size_t bytes = n * sizeof(T);
    char* p = pin_take(h, bytes);
    HIPCK(h, hipMemcpyAsync(p ? static_cast<void*>(p) : static_cast<void*>(dst), src, bytes,
                            hipMemcpyDeviceToHost, h->stream));
    HIPCK(h, hipStreamSynchronize(h->stream));
    if (p) std::memcpy(dst, p, bytes);
    if (h->pin_hold == 0) h->pin_off = 0;
}
template <typename T>
void h2d(qx_handle h, T* dst, const T* src, size_t n) {
    if (n == 0) return;
    const size_t bytes = n * sizeof(T);
    if (char* p = pin_take(h, bytes)) {
        std::memcpy(p, src, bytes);
        HIPCK(h, hipMemcpyAsync(dst, p, bytes, hipMemcpyHostToDevice, h->stream));
        return;
    }
    HIPCK(h, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, h->stream));
    HIPCK(h, hipStreamSynchronize(h->stream));
}

// Several downloads behind one synchronisation: add() queues each into the ring, run()
// waits once and copies them out (a download that finds no ring space runs on its own).
struct D2HBatch {
    qx_handle h;
    std::vector<std::tuple<void*, const char*, size_t>> items;
    explicit D2HBatch(qx_handle h_) : h(h_) { ++h->pin_hold; }
    ~D2HBatch() { --h->pin_hold; }
    template <typename T>
    void add(T* dst, const T* src, size_t n) {
        if (n == 0) return;
        const size_t bytes = n * sizeof(T);
        char* p = pin_take(h, bytes);
        if (!p) {
            d2h(h, dst, src, n);
            return;
        }
        HIPCK(h, hipMemcpyAsync(p, src, bytes, hipMemcpyDeviceToHost, h->stream));
        items.emplace_back(dst, p, bytes);
    }
    void run() {
        if (!items.empty()) HIPCK(h, hipStreamSynchronize(h->stream));
        for (auto& it : items) std::memcpy(std::get<0>(it), std::get<1>(it), std::get<2>(it));
        items.clear();
        if (h->pin_hold == 1) h->pin_off = 0;
    }
};

// ---- small utility kernels --------------------------------------------------------------
__global__ void fill_i32_kernel(int* p, int v, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        p[i] = v;
}
__global__ void fill_f64_kernel(double* p, double v, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        p[i] = v;
}
// dst[r][0:ncol] = src[r][0:ncol], dst[r][ncol:ldd] = 0
__global__ void pad_rows_kernel(const double* __restrict__ src, int nrow, int ncol,
                                double* __restrict__ dst, int ldd) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < (size_t)nrow * ldd;
         i += (size_t)gridDim.x * blockDim.x) {
        const size_t r = i / ldd, c = i % ldd;
        dst[i] = c < (size_t)ncol ? src[r * ncol + c] : 0.0;
    }
}
// eps[f][npad][npad] = I - pi[f][naux][naux], identity in the padding
__global__ void eps_from_pi_kernel(const double* __restrict__ pi, int naux, int npad, int nf,
                                   double* __restrict__ eps) {
    const size_t total = (size_t)nf * npad * npad;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < total;
         i += (size_t)gridDim.x * blockDim.x) {
        const size_t f = i / ((size_t)npad * npad), rc = i % ((size_t)npad * npad);
        const int r = rc / npad, c = rc % npad;
        double v = r == c ? 1.0 : 0.0;
        if (r < naux && c < naux) v -= pi[(f * naux + r) * naux + c];
        eps[i] = v;
    }
}

// ov-only pair rows: occupied target n pairs with every virtual a (row n*nvir + a of B_ov),
// virtual target n = nocc + a pairs with every occupied i (row i*nvir + a).
// rows: [npair][ld], columns naux .. ld - 1 zero-filled (ld = naux: unpadded).
__global__ void ov_pairs_kernel(const double* __restrict__ B, int nocc, int nvir, int naux,
                                const int* __restrict__ targets, const int* __restrict__ off,
                                int nt, int* __restrict__ pt, int* __restrict__ pm,
                                double* __restrict__ rows, int ld) {
    const int t = blockIdx.y;
    if (t >= nt) return;
    const int n = targets[t];
    const int j0 = off[t], cnt = off[t + 1] - off[t];
    for (int k = blockIdx.x; k < cnt; k += gridDim.x) {
        const int j = j0 + k;
        int r, m;
        if (n < nocc) { r = n * nvir + k; m = nocc + k; }
        else { r = k * nvir + (n - nocc); m = k; }
        if (threadIdx.x == 0) { pt[j] = t; pm[j] = m; }
        for (int c = threadIdx.x; c < ld; c += blockDim.x)
            rows[(size_t)j * ld + c] = c < naux ? B[(size_t)r * naux + c] : 0.0;
    }
}

dim3 grid1d(size_t n) {
    size_t b = (n + 255) / 256;
    if (b > 65535) b = 65535;
    if (b < 1) b = 1;
    return dim3((unsigned)b);
}

void fill_i32(qx_handle h, int* p, int v, size_t n) {
    hipLaunchKernelGGL(fill_i32_kernel, grid1d(n), dim3(256), 0, h->stream, p, v, n);
    HIPCK(h, hipGetLastError());
}
void fill_f64(qx_handle h, double* p, double v, size_t n) {
    hipLaunchKernelGGL(fill_f64_kernel, grid1d(n), dim3(256), 0, h->stream, p, v, n);
    HIPCK(h, hipGetLastError());
}

// ---- grids --------------------------------------------------------------------------------
// Gauss-Legendre by Newton iteration on P_n (the variant of
// quasix_core/src/freq/gauss_legendre.rs:24-86); nodes ascending, exact +- symmetry.
void gauss_legendre(int n, std::vector<double>& x, std::vector<double>& w) {
    x.assign(n, 0.0);
    w.assign(n, 0.0);
    if (n == 1) {
        w[0] = 2.0;
        return;
    }
    const double pi = 3.14159265358979323846;
    for (int i = 0; i < (n + 1) / 2; ++i) {
        double z = std::cos(pi * (i + 0.75) / (n + 0.5));
        double dp = 0.0;
        for (int it = 0; it < 100; ++it) {
            double p0 = 1.0, p1 = z;
            for (int k = 2; k <= n; ++k) {
                const double p2 = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p0) / k;
                p0 = p1;
                p1 = p2;
            }
            dp = n * (z * p1 - p0) / (z * z - 1.0);
            const double dz = p1 / dp;
            z -= dz;
            if (std::fabs(dz) < 1e-16) break;
        }
        {   // derivative at the converged root
            double p0 = 1.0, p1 = z;
            for (int k = 2; k <= n; ++k) {
                const double p2 = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p0) / k;
                p0 = p1;
                p1 = p2;
            }
            dp = n * (z * p1 - p0) / (z * z - 1.0);
        }
        const double wt = 2.0 / ((1.0 - z * z) * dp * dp);
        x[n - 1 - i] = z;
        x[i] = -z;
        w[n - 1 - i] = wt;
        w[i] = wt;
    }
    if (n % 2 == 1) x[n / 2] = 0.0;
}

void scaled_roots(int nw, double x0, std::vector<double>& f, std::vector<double>& w) {
    std::vector<double> t, g;
    gauss_legendre(nw, t, g);
    f.resize(nw);
    w.resize(nw);
    for (int i = 0; i < nw; ++i) {
        const double omt = 1.0 - t[i];
        f[i] = x0 * (1.0 + t[i]) / omt;
        w[i] = g[i] * 2.0 * x0 / (omt * omt);
    }
}

NewtonParams to_params(const qx_newton_config& c, int qp_solver = QX_SOLVER_NEWTON) {
    NewtonParams p;
    p.linearized = qp_solver == QX_SOLVER_LINEARIZED;
    p.max_iterations = c.max_iterations;
    p.energy_tolerance = c.energy_tolerance;
    p.initial_damping = c.initial_damping;
    p.min_damping = c.min_damping;
    p.max_damping = c.max_damping;
    p.max_step = c.max_step;
    p.derivative_delta = c.derivative_delta;
    p.z_min = c.z_min;
    p.z_max = c.z_max;
    p.use_richardson = c.use_richardson;
    p.use_bisection = c.use_bisection_fallback;
    p.max_bisection_iterations = c.max_bisection_iterations;
    return p;
}

// ---- Wc_nm(i w) for all frequencies ------------------------------------------------------
struct WnmResult {
    int chol_fallbacks = 0;
    int singular = 0;
    int pi_variant = 0;
    // deferred failure check (build_wnm with defer_check): the caller downloads fail_dev with
    // its own results and, if any frequency failed, runs fallback (returns the singular count)
    const int* fail_dev = nullptr;
    std::function<int(const std::vector<int>&)> fallback;
};

// Returns the variant to use for `kind` at this shape, timing the candidates on first use
// (every variant when `candidates` is empty; QXGW_PI_VARIANT can force any of the
// nvariants).  Unmeasured variants record -1 ms in the tuning file.
template <class Launch>
int choose_variant(qx_handle h, const char* kind, int precision, int nvariants,
                   std::initializer_list<int> shape, Launch launch,
                   const std::vector<int>& candidates = {}) {
    if (nvariants <= 1) return 0;
    if (const char* f = std::getenv("QXGW_PI_VARIANT")) {
        const int v = std::atoi(f);
        if (v >= 0 && v < nvariants) return v;
    }
    const char* at = std::getenv("QXGW_AUTOTUNE");
    if (at && std::atoi(at) == 0) return 0;
    std::string key = std::string(kind) + "/p" + std::to_string(precision);
    for (int x : shape) key += "/" + std::to_string(x);
    tune_load(h);
    auto it = h->tuned.find(key);
    if (it != h->tuned.end() && it->second < nvariants) return it->second;
    // the listed candidates that exist for this precision, else every variant
    std::vector<int> cand;
    for (int v : candidates)
        if (v >= 0 && v < nvariants) cand.push_back(v);
    if (cand.empty())
        for (int v = 0; v < nvariants; ++v) cand.push_back(v);
    // no-allocation mode times nothing (no events are created): first candidate
    if (h->arena) return cand.front();
    std::vector<float> ms(nvariants, -1.f);
    hipEvent_t a = take_event(h), b = take_event(h);
    int best = cand.front();
    for (int v : cand) {
        launch(v);                                   // warm-up (code object load, caches)
        HIPCK(h, hipEventRecord(a, h->stream));
        launch(v);
        HIPCK(h, hipEventRecord(b, h->stream));
        HIPCK(h, hipEventSynchronize(b));
        HIPCK(h, hipEventElapsedTime(&ms[v], a, b));
        if (ms[v] < ms[best] || ms[best] < 0.f) best = v;
    }
    h->ev_free.push_back(a);
    h->ev_free.push_back(b);
    h->tuned[key] = best;
    tune_store(key, best, ms);
    return best;
}

// Frequencies used when timing a Pi variant: all of them up to 8192 work items (tiles x
// frequencies), beyond that an even subset of about 4096 items (the chip filled several
// times over), so a large problem is not tuned at full cost.  Even, because the
// frequency-pair instances would otherwise time a duplicated frequency.
int tune_nf(int naux_or_npad, int nf) {
    const int T = ceil_div(naux_or_npad, 128);
    const int tiles = std::max(T * (T + 1) / 2, 1);
    if (tiles * nf <= 8192) return nf;
    const int n = ceil_div(4096, tiles);
    return std::max(1, std::min(nf, n + (n & 1)));
}

// Autotuner candidates of the BF16x3 Pi kernels: the instances that win or come within a few
// per cent at the measured shapes (the producer/consumer forms, K64 on 32x32x16 MFMAs, K64 on
// 16x16x32); the others stay selectable by QXGW_PI_VARIANT.
// FP64 has the one-frequency instance (0) and the frequency-pair instances with K chunks of
// 16 (1) and 32 (2), and with double-buffered LDS, K32 (3) and K16 (4), all timed.
std::vector<int> pi_candidates(int precision) {
    if (precision == QX_PREC_BF16X3) return {8, 5, 7, 9, 1};
    if (precision == QX_PREC_FP64) return {2, 3, 1, 4, 0};
    return {};
}

// Side streams and the group count for the factorisation step loop (QXGW_POTRF_GROUPS,
// default 3; 1 runs everything on the handle's stream).
const StreamFan* potrf_fan(qx_handle h, int& groups) {
    static const int env_groups = [] {
        const char* v = getenv("QXGW_POTRF_GROUPS");
        const int g = v ? atoi(v) : 3;
        return std::max(1, std::min(g, StreamFan::kMax + 1));
    }();
    groups = h->arena ? 1 : env_groups;   // no side streams are created in no-allocation mode
    if (groups > 1 && h->fan.n < groups - 1) {
        if (!h->fan.fork) HIPCK(h, hipEventCreateWithFlags(&h->fan.fork, hipEventDisableTiming));
        while (h->fan.n < groups - 1) {
            const int g = h->fan.n;
            HIPCK(h, hipStreamCreate(&h->fan.side[g]));
            HIPCK(h, hipEventCreateWithFlags(&h->fan.join[g], hipEventDisableTiming));
            h->fan.n = g + 1;
        }
    }
    return &h->fan;
}

// Pi operand preparation: the frequency weights s_f(r) (and their FP32 copy), and for BF16x3
// the transposed hi/lo/FP32 copies of B.  launch = false only looks the buffers up (the
// launches were queued earlier in the call).
PiOperands pi_prep(qx_handle h, int precision, int nocc, int nvir, int naux, int nf,
                   const double* eps_dev, const double* B_ov, const double* freqs_dev,
                   double** s_out, bool launch) {
    const int nov = nocc * nvir;
    const int npad = round_up(naux, NB), kpad = round_up(nov, KPAD);
    hipStream_t st = h->stream;
    PiOperands ops;
    double* s = ws<double>(h, "pi_weights", (size_t)nf * kpad);
    float* s32 = precision == QX_PREC_BF16X3 ? ws<float>(h, "pi_weights32", (size_t)nf * kpad)
                                             : nullptr;
    if (launch) HIPCK(h, launch_pi_weights(eps_dev, nocc, nvir, nf, freqs_dev, s, s32, st));
    ops.s32 = s32;
    if (precision == QX_PREC_BF16X3) {
        const int brows = round_up(npad, 256);
        uint16_t* hi = ws<uint16_t>(h, "bt_hi", (size_t)brows * kpad);
        uint16_t* lo = ws<uint16_t>(h, "bt_lo", (size_t)brows * kpad);
        float* f32 = ws<float>(h, "bt_f32", (size_t)brows * kpad);
        if (launch)
            HIPCK(h, launch_pi_prepare_mixed(B_ov, nov, naux, brows, kpad, hi, lo, f32, st));
        ops.bt_hi = hi;
        ops.bt_lo = lo;
        ops.bt_f32 = f32;
        ops.kpad = kpad;
        ops.rows = brows;
    }
    *s_out = s;
    return ops;
}

WnmResult build_wnm(qx_handle h, int precision, int nocc, int nvir, int naux,
                    const double* eps_dev, const double* B_ov, int nf, const double* freqs_dev,
                    int npair, const double* B_pair, double* wnm_out, bool defer_check = false,
                    bool prep_queued = false, bool pairs_padded = false,
                    const std::function<void()>* pair_rows = nullptr) {
    WnmResult r;
    const int nov = nocc * nvir;
    const int npad = round_up(naux, NB), nblk = npad / NB, kpad = round_up(nov, KPAD);
    hipStream_t st = h->stream;
    double* eps_mat = ws<double>(h, "eps_mat", (size_t)nf * npad * npad);
    double* linv = ws<double>(h, "linv_diag", (size_t)nf * nblk * NB * NB);
    int* fail = ws<int>(h, "chol_fail", nf);
    double* s = nullptr;
    PiOperands ops;
    {
        Phase ph(h, QX_PHASE_PREP);
        ops = pi_prep(h, precision, nocc, nvir, naux, nf, eps_dev, B_ov, freqs_dev, &s,
                      !prep_queued);
        fill_i32(h, fail, 0, nf);
    }
    double* x = ws<double>(h, "xinv", (size_t)nf * npad * npad);
    // padded pair rows already are the "bpairs" buffer
    double* bp = ws<double>(h, "bpairs", (size_t)std::max(npair, 1) * npad);
    REQUIRE(h, !pairs_padded || bp == B_pair, QX_ERR_INVALID_INPUT,
            "padded pair rows must be the pair-row workspace");
    if (npair > 0 && !pairs_padded) {
        hipLaunchKernelGGL(pad_rows_kernel, grid1d((size_t)npair * npad), dim3(256), 0, st,
                           B_pair, npair, naux, bp, npad);
        HIPCK(h, hipGetLastError());
    }
    double* wpart = ws<double>(h, "wpart", (size_t)nf * nblk * std::max(npair, 1));
    WnmMixed mx{};
    if (precision == QX_PREC_BF16X3 && npair > 0) {
        const size_t npp = (size_t)round_up(npair, 128);
        mx.bh = ws<__bf16>(h, "wnm_bh", npp * npad);
        mx.bl = ws<__bf16>(h, "wnm_bl", npp * npad);
        mx.mh = ws<__bf16>(h, "wnm_mh", (size_t)nf * npad * npad);
        mx.ml = ws<__bf16>(h, "wnm_ml", (size_t)nf * npad * npad);
    }
    const int pv = choose_variant(h, "pi", precision, pi_variant_count(precision),
                                  {naux, nov, nf}, [&](int v) {
        HIPCK(h, launch_pi_syrk(precision, B_ov, ops, s, nov, naux, npad, tune_nf(npad, nf),
                                eps_mat, PiOut::EpsLower, st, v));
    }, pi_candidates(precision));
    r.pi_variant = pv;
    {
        Phase ph(h, QX_PHASE_PI);
        HIPCK(h, launch_pi_syrk(precision, B_ov, ops, s, nov, naux, npad, nf, eps_mat,
                                PiOut::EpsLower, st, pv));
    }
    // mixed: each frequency group's share of the BF16 split of M = X - I runs on that group's
    // stream right after its factorisation, overlapping the other groups' last steps
    // (QXGW_SPLIT_EARLY=0: one split pass after the join, inside the Wc_nm launch sequence)
    const char* se = std::getenv("QXGW_SPLIT_EARLY");
    const bool split_early =
        precision == QX_PREC_BF16X3 && npair > 0 && !(se && std::atoi(se) == 0);
    const size_t mat = (size_t)npad * npad;
    const PotrfAfter split = [&](hipStream_t gs, int f0, int nfg) {
        return launch_split_m(x + f0 * mat, npad, nfg, mx.mh + f0 * mat, mx.ml + f0 * mat, gs);
    };
    // what only Wc_nm reads (the pair rows the caller queued, their BF16 split) runs after Pi,
    // on the main stream beside the other frequency groups' first steps, not in front of Pi
    const PotrfAfter lead = [&](hipStream_t ls, int, int) {
        if (pair_rows) (*pair_rows)();
        if (precision == QX_PREC_BF16X3 && npair > 0)
            return launch_wnm_split_rows(bp, npair, npad, &mx, ls);
        return hipSuccess;
    };
    {
        Phase ph(h, QX_PHASE_POTRF);
        int groups = 1;
        const StreamFan* fan = potrf_fan(h, groups);
        HIPCK(h, launch_potrf_batched(eps_mat, x, linv, npad, nf, fail, st, fan, groups,
                                      split_early ? &split : nullptr, &lead));
    }
    if (npair > 0) {
        // Wc_nm before the failure check: on the rare Cholesky failure it is recomputed below
        Phase ph(h, QX_PHASE_WNM);
        HIPCK(h, launch_wnm(precision, bp, x, npair, npad, nf, wpart, wnm_out, &mx, st, false,
                            !split_early));
    }
    // Reference chain: Cholesky failed somewhere -> LU inverse (here Gauss-Jordan with
    // partial pivoting on every frequency) -> identity if singular.  Recomputes wnm_out.
    auto fallback = [=](const std::vector<int>&) {
        double* pi = ws<double>(h, "xinv", (size_t)nf * npad * npad);
        fill_f64(h, pi, 0.0, (size_t)nf * npad * npad);
        HIPCK(h, launch_pi_syrk(precision, B_ov, ops, s, nov, naux, npad, nf, pi, PiOut::PiFull,
                                st));
        double* work = ws<double>(h, "gj_work", (size_t)nf * npad * 2 * npad + nf);
        double* wc = eps_mat;
        HIPCK(h, launch_lu_fallback_wc(pi, 0, work, wc, npad, nf, nullptr, st));
        std::vector<int> sing(nf);
        d2h(h, sing.data(), reinterpret_cast<int*>(work + (size_t)nf * npad * 2 * npad), nf);
        int nsing = 0;
        for (int f = 0; f < nf; ++f) nsing += sing[f] != 0;
        if (npair > 0)
            HIPCK(h, launch_wnm_from_wc(bp, wc, npair, npad, nf, wpart, wnm_out, st));
        return nsing;
    };
    if (defer_check) {
        r.fail_dev = fail;
        r.fallback = fallback;
        return r;
    }
    std::vector<int> fail_h(nf);
    d2h(h, fail_h.data(), fail, nf);
    for (int f = 0; f < nf; ++f) r.chol_fallbacks += fail_h[f] != 0;
    if (r.chol_fallbacks > 0) r.singular = fallback(fail_h);
    return r;
}

// ---- host view of a CD problem for the residue path ------------------------------------
struct HostCD {
    qx_handle h;
    int nocc, nvir, naux, nmo, nt, npair, nf;
    const double* B_ov;            // device
    const double* eps_g_dev;       // device
    const double* B_pair;          // device [npair][b_ld]
    int b_ld = 0;                  // row stride of B_pair: naux, or npad for padded rows
    std::vector<double> eps_g, eps_ref;
    std::vector<int> targets, pair_m, pair_off;
    double ef, eta;
    bool host_ready = false;       // eps_g, eps_ref, pair_m and ef hold the device values
    SigmaProblem prob;             // device problem (all targets)
    int residue_jobs = 0;

    // One evaluation request: Sigma_c(E) of target t at nq (<= 8) energies into out.
    struct Req {
        int t;
        const double* E;
        int nq;
        C2* out;
    };

    // Sigma_c(E) for target t at nq energies: I on the GPU quadrature kernel, residues as
    // batched real-axis solves.
    void eval(int t, const double* E, int nq, C2* out) {
        std::vector<Req> one{Req{t, E, nq, out}};
        eval_many(one);
    }

    // Several requests at once (the Newton iterations of several residue targets in
    // lockstep): the quadrature part of each request on the GPU, read back together, then
    // every residue job of every request in as few batches as the workspace allows.  Each
    // job's value does not depend on which batch it is in, and each request accumulates its
    // jobs in the order eval() would, so the results equal those of one eval() per request.
    void eval_many(std::vector<Req>& reqs) {
        const int nr = (int)reqs.size();
        double* e_dev = ws<double>(h, "hcd_E", (size_t)8 * nr);
        double* o_dev = ws<double>(h, "hcd_out", (size_t)16 * nr);
        std::vector<std::vector<double>> o(nr);
        for (int k = 0; k < nr; ++k) {
            const Req& r = reqs[k];
            SigmaProblem sub = prob;
            sub.nt = 1;
            sub.targets = prob.targets + r.t;
            sub.pair_off = prob.pair_off + r.t;
            sub.static_terms = prob.static_terms + r.t;
            h2d(h, e_dev + 8 * k, r.E, r.nq);
            HIPCK(h, launch_sigma_imag_eval(sub, e_dev + 8 * k, r.nq, o_dev + 16 * k, h->stream));
            o[k].resize(2 * r.nq);
        }
        {
            D2HBatch bt(h);
            for (int k = 0; k < nr; ++k) bt.add(o[k].data(), o_dev + 16 * k, 2 * reqs[k].nq);
            bt.run();
        }
        for (int k = 0; k < nr; ++k)
            for (int q = 0; q < reqs[k].nq; ++q) reqs[k].out[q] = C2{o[k][2 * q], o[k][2 * q + 1]};
        // residue jobs of every request, request by request in eval()'s order
        std::vector<double> om;
        std::vector<int> row, qidx, ridx;
        std::vector<double> sgn;
        for (int k = 0; k < nr; ++k) {
            const Req& r = reqs[k];
            for (int q = 0; q < r.nq; ++q)
                for (int j = pair_off[r.t]; j < pair_off[r.t + 1]; ++j) {
                    const double em = eps_g[pair_m[j]];
                    if (r.E[q] > ef && ef < em && em < r.E[q]) {
                        om.push_back(r.E[q] - em); row.push_back(j); qidx.push_back(q);
                        ridx.push_back(k); sgn.push_back(1.0);
                    } else if (r.E[q] <= ef && r.E[q] < em && em < ef) {
                        om.push_back(em - r.E[q]); row.push_back(j); qidx.push_back(q);
                        ridx.push_back(k); sgn.push_back(-1.0);
                    }
                }
        }
        const int nj = (int)om.size();
        if (nj == 0) return;
        residue_jobs += nj;
        // the complex eps(z) of a job takes two [npad][npad] planes
        const int nov = nocc * nvir, kpad = round_up(nov, KPAD), npad = round_up(naux, NB);
        const size_t per = (size_t)2 * npad * npad * sizeof(double);
        const int chunk = (int)std::max<size_t>(1, std::min<size_t>(64, (size_t)(1ull << 30) / per));
        for (int j0 = 0; j0 < nj; j0 += chunk) {
            const int cnt = std::min(chunk, nj - j0);
            double* om_dev = ws<double>(h, "res_om", chunk);
            double* rows = ws<double>(h, "res_rows", (size_t)chunk * naux);
            double* s = ws<double>(h, "res_weights", (size_t)2 * chunk * kpad);
            double* P = ws<double>(h, "res_eps", (size_t)2 * chunk * npad * npad);
            double* rhs = ws<double>(h, "res_rhs", (size_t)2 * chunk * npad);
            int* fail = ws<int>(h, "res_fail", chunk);
            double* wout = ws<double>(h, "res_out", (size_t)chunk * 2);
            h2d(h, om_dev, om.data() + j0, cnt);
            for (int k = 0; k < cnt; ++k)
                HIPCK(h, hipMemcpyAsync(rows + (size_t)k * naux, B_pair + (size_t)row[j0 + k] * b_ld,
                                        naux * sizeof(double), hipMemcpyDeviceToDevice, h->stream));
            Phase ph(h, QX_PHASE_RESIDUE);
            fill_i32(h, fail, 0, cnt);
            HIPCK(h, launch_residue_weights(eps_g_dev, nocc, nvir, cnt, om_dev, eta, kpad, s,
                                            h->stream));
            // Re Pi and Im Pi of a job are the two frequencies of one frequency-pair workgroup
            HIPCK(h, launch_pi_syrk(QX_PREC_FP64, B_ov, PiOperands{}, s, nov, naux, npad, 2 * cnt,
                                    P, PiOut::PiFull, h->stream, 2));
            HIPCK(h, launch_residue_lu_solve(P, naux, npad, cnt, rows, rhs, fail, wout,
                                             h->stream));
            std::vector<double> w(2 * cnt);
            std::vector<int> f(cnt);
            {
                D2HBatch bt(h);
                bt.add(w.data(), wout, 2 * cnt);
                bt.add(f.data(), fail, cnt);
                bt.run();
            }
            for (int k = 0; k < cnt; ++k) {
                REQUIRE(h, f[k] == 0, QX_ERR_NUMERICAL,
                        "eps(w + i eta) is singular at w = %.6f (residue of partner %d)",
                        om[j0 + k], pair_m[row[j0 + k]]);
                C2& acc = reqs[ridx[j0 + k]].out[qidx[j0 + k]];
                acc.re += sgn[j0 + k] * w[2 * k];
                acc.im += sgn[j0 + k] * w[2 * k + 1];
            }
        }
    }
};

// Newton iterations of several residue targets in lockstep: each target's solve runs on a
// host thread of its own (host arithmetic only); its Sigma evaluations are posted here, and
// the calling thread, once every live solve has posted one or finished, evaluates all of
// them with HostCD::eval_many (all device work stays on the calling thread) and wakes them.
// A failure of the device work stops every solve and is rethrown on the calling thread.
struct ResidueBatcher {
    struct Abort {};
    HostCD* cd;
    std::mutex mu;
    std::condition_variable cv_main, cv_thr;
    int live = 0;
    long long round = 0;
    std::vector<HostCD::Req> pending;
    bool failed = false;

    void eval(int t, const double* E, int nq, C2* out) {     // solve threads
        std::unique_lock<std::mutex> lk(mu);
        if (failed) throw Abort{};
        const long long my = round;
        pending.push_back(HostCD::Req{t, E, nq, out});
        cv_main.notify_one();
        cv_thr.wait(lk, [&] { return round != my; });
        if (failed) throw Abort{};
    }
    void finish() {                                           // solve threads
        std::lock_guard<std::mutex> lk(mu);
        --live;
        cv_main.notify_one();
    }
    void run() {                                              // calling thread
        std::exception_ptr err;
        std::unique_lock<std::mutex> lk(mu);
        for (;;) {
            cv_main.wait(lk, [&] { return (int)pending.size() == live; });
            if (live == 0) break;
            std::vector<HostCD::Req> batch;
            batch.swap(pending);
            if (!failed) {
                lk.unlock();
                try {
                    cd->eval_many(batch);
                } catch (...) {
                    err = std::current_exception();
                }
                lk.lock();
                if (err) failed = true;
            }
            ++round;
            cv_thr.notify_all();
        }
        if (err) std::rethrow_exception(err);
    }
};

struct HostEval {
    HostCD* cd;
    int t;
    ResidueBatcher* batch = nullptr;     // lockstep solves post their evaluations here
    bool operator()(const double* E, int nq, C2* out) {
        if (batch) batch->eval(t, E, nq, out);
        else cd->eval(t, E, nq, out);
        return false;
    }
};

// Solve target t on the host-driven path; returns the QX_QP_* state.
int host_solve(HostCD& cd, const NewtonParams& np, int t, double static_t, bool have_guess,
               double guess, double& E, double& Z, C2& s, ResidueBatcher* batch = nullptr) {
    HostEval ev{&cd, t, batch};
    const int n = cd.targets[t];
    int iters = 0;
    const int st = newton_core(ev, np, cd.eps_ref[n], static_t, cd.ef, have_guess, guess, E, Z,
                               s, iters);
    if (st == kNewtonConverged) return QX_QP_NEWTON;
    if (st == kNewtonBisection) return QX_QP_BISECTION;
    if (st == kNewtonLinearized) return QX_QP_LINEARIZED;
    // Linearised fallback at the Newton start (evgw.rs:766-781).
    double g = guess;
    const double de = 1e-6;
    if (!have_guess) {
        const double om = cd.eps_ref[n] - cd.ef;
        double Es[2] = {om + cd.ef, (om + np.derivative_delta) + cd.ef};
        C2 r[2];
        ev(Es, 2, r);
        double z = 1.0 / (1.0 - (r[1].re - r[0].re) / np.derivative_delta);
        z = std::min(std::max(z, np.z_min), np.z_max);
        g = cd.eps_ref[n] + z * (static_t + r[0].re);
    }
    const double om = g - cd.ef;
    double Es[2] = {om + cd.ef, (om + de) + cd.ef};
    C2 r[2];
    ev(Es, 2, r);
    double z = 1.0 / (1.0 - (r[1].re - r[0].re) / de);
    z = std::min(std::max(z, 0.1), 0.999);
    E = cd.eps_ref[n] + z * (static_t + r[0].re);
    Z = z;
    s = r[0];
    return QX_QP_LINEARIZED;
}

void validate_desc(qx_handle h, const qx_g0w0_desc* d) {
    REQUIRE(h, d != nullptr, QX_ERR_INVALID_INPUT, "descriptor is NULL");
    REQUIRE(h, d->nocc > 0 && d->nvir > 0, QX_ERR_INVALID_INPUT,
            "Invalid nocc=%d/nvir=%d: both must be positive", d->nocc, d->nvir);
    REQUIRE(h, d->naux > 0, QX_ERR_INVALID_INPUT, "naux must be positive (got %d)", d->naux);
    REQUIRE(h, d->nfreq > 0, QX_ERR_INVALID_INPUT, "nfreq must be positive (got %d)", d->nfreq);
    REQUIRE(h, d->x0 > 0.0, QX_ERR_INVALID_INPUT, "x0 must be positive");
    REQUIRE(h, d->eta >= 0.0, QX_ERR_INVALID_INPUT, "eta must be non-negative");
    REQUIRE(h, d->precision >= 0 && d->precision <= 2, QX_ERR_INVALID_INPUT,
            "unknown precision %d", d->precision);
    REQUIRE(h, d->ntarget > 0, QX_ERR_INVALID_INPUT, "ntarget must be positive");
    REQUIRE(h, d->npair >= 0, QX_ERR_INVALID_INPUT, "npair must be non-negative");
    REQUIRE(h, d->qp_solver == QX_SOLVER_NEWTON || d->qp_solver == QX_SOLVER_LINEARIZED,
            QX_ERR_INVALID_INPUT, "unknown qp_solver %d", d->qp_solver);
    REQUIRE(h, (long long)d->nocc * d->nvir < (1ll << 31), QX_ERR_INVALID_INPUT,
            "nocc*nvir overflows int");
}

// Builds the host+device CD problem (pair CSR, grid) shared by g0w0 and sigma_cd.
// The CD frequency grid (nodes, weights) on the device in the buffers "cd_fr" / "cd_wt",
// uploaded only when the key or the buffer changes.
void cd_grid(qx_handle h, int nf, double x0, double** fr, double** wt) {
    double* fr_dev = ws<double>(h, "cd_fr", nf);
    double* wt_dev = ws<double>(h, "cd_wt", nf);
    if (h->arena || h->grid_fr != fr_dev || h->grid_nf != nf || h->grid_x0 != x0) {
        std::vector<double> f, w;
        scaled_roots(nf, x0, f, w);
        h2d(h, fr_dev, f.data(), nf);
        h2d(h, wt_dev, w.data(), nf);
        // a buffer that grows is reallocated and both names are taken together, so an
        // unchanged fr pointer and key mean both still hold this grid
        h->grid_fr = h->arena ? nullptr : fr_dev;
        h->grid_nf = nf;
        h->grid_x0 = x0;
    }
    *fr = fr_dev;
    *wt = wt_dev;
}

// known (the ov path, whose pair list the library generated from targets it validated):
// the host targets and pair offsets are given and off_dev holds the offsets once the pair
// rows are queued (pair_rows), so nothing is downloaded here; the host copies of the
// energies and pair partners that only the residue path reads are fetched by cd_host_data
// when a target needs it.
struct KnownPairs {
    const std::vector<int>* targets;
    const std::vector<int>* off;
    int* off_dev;
    bool prep_queued;                // the Pi operand preparation is already on the stream
    bool padded;                     // B_pair is [npair][npad], zero-padded ("bpairs")
    // queues the writes of the pair rows and partners (B_pair, pair_target, pair_m, off_dev)
    // on the library's stream; build_wnm runs it after the Pi launch (nullptr: already queued)
    const std::function<void()>* pair_rows = nullptr;
};

void cd_host_data(qx_handle h, const qx_g0w0_desc* d, const double* mo_energy,
                  const double* mo_energy_ref, const int* pair_m, HostCD& cd) {
    D2HBatch bt(h);
    bt.add(cd.eps_g.data(), mo_energy, cd.nmo);
    bt.add(cd.eps_ref.data(), mo_energy_ref ? mo_energy_ref : mo_energy, cd.nmo);
    bt.add(cd.pair_m.data(), pair_m, cd.npair);
    bt.run();
    cd.ef = 0.5 * (cd.eps_g[d->nocc - 1] + cd.eps_g[d->nocc]);
    cd.prob.ef = cd.ef;
    cd.host_ready = true;
}

void setup_cd(qx_handle h, const qx_g0w0_desc* d, const double* mo_energy,
              const double* mo_energy_ref, const double* B_ov, const double* B_pair,
              const int* pair_target, const int* pair_m, const int* targets,
              const double* static_terms, const double* freqs_dev, const double* wts_dev,
              const double* wnm, HostCD& cd, const KnownPairs* known = nullptr) {
    const int nmo = d->nocc + d->nvir, nt = d->ntarget, npair = d->npair;
    cd.h = h;
    cd.nocc = d->nocc;
    cd.nvir = d->nvir;
    cd.naux = d->naux;
    cd.nmo = nmo;
    cd.nt = nt;
    cd.npair = npair;
    cd.nf = d->nfreq;
    cd.B_ov = B_ov;
    cd.eps_g_dev = mo_energy;
    cd.B_pair = B_pair;
    cd.b_ld = known && known->padded ? round_up(d->naux, NB) : d->naux;
    cd.eta = d->eta;
    cd.eps_g.resize(nmo);
    cd.eps_ref.resize(nmo);
    cd.targets.resize(nt);
    cd.pair_m.resize(npair);
    SigmaProblem& p = cd.prob;
    p.nt = nt;
    p.npair = npair;
    p.nf = d->nfreq;
    p.nocc = d->nocc;
    p.nmo = nmo;
    p.targets = targets;
    p.pair_m = pair_m;
    p.wnm = wnm;
    p.eps_g = mo_energy;
    p.eps_ref = mo_energy_ref ? mo_energy_ref : mo_energy;
    p.freqs = freqs_dev;
    p.wts = wts_dev;
    p.static_terms = static_terms;
    p.eta = d->eta;
    if (known) {
        cd.targets = *known->targets;
        cd.pair_off = *known->off;
        p.pair_off = known->off_dev;
        p.ef = 0.0;                          // formed on the device by the Newton kernel
        cd.host_ready = false;
        return;
    }
    std::vector<int> pt(npair);
    {
        D2HBatch bt(h);
        bt.add(cd.eps_g.data(), mo_energy, nmo);
        bt.add(cd.eps_ref.data(), mo_energy_ref ? mo_energy_ref : mo_energy, nmo);
        bt.add(cd.targets.data(), targets, nt);
        bt.add(pt.data(), pair_target, npair);
        bt.add(cd.pair_m.data(), pair_m, npair);
        bt.run();
    }
    for (int t = 0; t < nt; ++t)
        REQUIRE(h, cd.targets[t] >= 0 && cd.targets[t] < nmo, QX_ERR_INVALID_INPUT,
                "target %d out of range [0,%d)", cd.targets[t], nmo);
    cd.pair_off.assign(nt + 1, 0);
    for (int j = 0; j < npair; ++j) {
        REQUIRE(h, pt[j] >= 0 && pt[j] < nt, QX_ERR_INVALID_INPUT, "pair_target[%d]=%d out of range",
                j, pt[j]);
        REQUIRE(h, j == 0 || pt[j] >= pt[j - 1], QX_ERR_INVALID_INPUT,
                "pairs must be sorted by target (pair %d)", j);
        REQUIRE(h, cd.pair_m[j] >= 0 && cd.pair_m[j] < nmo, QX_ERR_INVALID_INPUT,
                "pair_m[%d]=%d out of range", j, cd.pair_m[j]);
        cd.pair_off[pt[j] + 1]++;
    }
    for (int t = 0; t < nt; ++t) cd.pair_off[t + 1] += cd.pair_off[t];
    cd.ef = 0.5 * (cd.eps_g[d->nocc - 1] + cd.eps_g[d->nocc]);
    int* off_dev = ws<int>(h, "pair_off", nt + 1);
    h2d(h, off_dev, cd.pair_off.data(), nt + 1);
    p.pair_off = off_dev;
    p.ef = cd.ef;
    cd.host_ready = true;
}

qx_status run_g0w0(qx_handle h, const qx_g0w0_desc* d, const double* mo_energy,
                   const double* mo_energy_ref, const double* B_ov, const double* B_pair,
                   const int* pair_target, const int* pair_m, const int* targets,
                   const double* static_terms, const double* guess, double* qp_energy,
                   double* z_factor, double* sigma_re, double* sigma_im, int* state,
                   qx_stats* stats, const KnownPairs* known = nullptr) {
    const auto t0 = std::chrono::steady_clock::now();
    validate_desc(h, d);
    REQUIRE(h, mo_energy && B_ov && targets && qp_energy && z_factor && sigma_re && sigma_im,
            QX_ERR_INVALID_INPUT, "required pointer is NULL");
    REQUIRE(h, d->npair == 0 || (B_pair && pair_target && pair_m), QX_ERR_INVALID_INPUT,
            "pair arrays are NULL");
    const int nf = d->nfreq, nt = d->ntarget, npair = d->npair;
    double *fr_dev, *wt_dev;
    cd_grid(h, nf, d->x0, &fr_dev, &wt_dev);
    if (!static_terms) {
        double* zero_static = ws<double>(h, "zero_static", nt);
        if (h->arena || h->zero_ptr != zero_static || h->zero_n < nt) {
            fill_f64(h, zero_static, 0.0, nt);
            h->zero_ptr = h->arena ? nullptr : zero_static;
            h->zero_n = nt;
        }
        static_terms = zero_static;
    }
    double* wnm = ws<double>(h, "wnm", (size_t)nf * std::max(npair, 1));
    HostCD cd;
    setup_cd(h, d, mo_energy, mo_energy_ref, B_ov, B_pair, pair_target, pair_m, targets,
             static_terms, fr_dev, wt_dev, wnm, cd, known);
    // the Cholesky failure flags come back with the Newton states (one synchronisation)
    WnmResult wr = build_wnm(h, d->precision, d->nocc, d->nvir, d->naux, mo_energy, B_ov, nf,
                             fr_dev, npair, B_pair, wnm, true, known && known->prep_queued,
                             known && known->padded, known ? known->pair_rows : nullptr);
    const NewtonParams np = to_params(d->newton, d->qp_solver);
    int* st_dev = ws<int>(h, "newton_state", nt);
    int* it_dev = ws<int>(h, "newton_iters", nt);
    auto newton = [&] {
        Phase ph(h, QX_PHASE_NEWTON);
        // the kernel writes the state code of every target it finished into `state`
        HIPCK(h, launch_newton_cd(cd.prob, np, guess, qp_energy, z_factor, sigma_re, sigma_im,
                                  st_dev, it_dev, h->stream, state));
    };
    newton();
    // one synchronisation ends the call unless a frequency's Cholesky failed or a target
    // needs the residue path on the host
    std::vector<int> st(nt), fail_h(nf);
    {
        D2HBatch bt(h);
        bt.add(st.data(), st_dev, nt);
        bt.add(fail_h.data(), wr.fail_dev, nf);
        bt.run();
    }
    for (int f = 0; f < nf; ++f) wr.chol_fallbacks += fail_h[f] != 0;
    if (wr.chol_fallbacks > 0) {
        // Wc_nm and the Newton results above used a failed factor: redo both
        wr.singular = wr.fallback(fail_h);
        newton();
        d2h(h, st.data(), st_dev, nt);
    }
    int residue_targets = 0, newton_fail = 0;
    for (int t = 0; t < nt; ++t)
        if (st[t] != kNewtonConverged && st[t] != kNewtonBisection && st[t] != kNewtonLinearized)
            ++residue_targets;
    if (residue_targets > 0) {
        if (!cd.host_ready) cd_host_data(h, d, mo_energy, mo_energy_ref, pair_m, cd);
        std::vector<double> guess_h;
        std::vector<double> stat_h(nt);
        {
            D2HBatch bt(h);
            if (guess) {
                guess_h.resize(nt);
                bt.add(guess_h.data(), guess, nt);
            }
            bt.add(stat_h.data(), static_terms, nt);
            bt.run();
        }
        std::vector<int> rt;                   // targets for the host residue path
        for (int t = 0; t < nt; ++t)
            if (st[t] != kNewtonConverged && st[t] != kNewtonBisection &&
                st[t] != kNewtonLinearized)
                rt.push_back(t);
        std::vector<double> Eh(nt), Zh(nt);
        std::vector<C2> Sh(nt);
        std::vector<int> code(nt);
        auto solve = [&](int t, ResidueBatcher* b) {
            code[t] = host_solve(cd, np, t, stat_h[t], guess != nullptr, guess ? guess_h[t] : 0.0,
                                 Eh[t], Zh[t], Sh[t], b);
        };
        // QXGW_RES_LOCKSTEP=0: one target after the other (each evaluation its own batch)
        const char* lv = std::getenv("QXGW_RES_LOCKSTEP");
        if (rt.size() > 1 && !(lv && std::atoi(lv) == 0)) {
            ResidueBatcher bat;
            bat.cd = &cd;
            bat.live = (int)rt.size();
            std::vector<std::thread> th;
            std::vector<char> bad(nt, 0);
            for (int t : rt)
                th.emplace_back([&, t] {
                    try {
                        solve(t, &bat);
                    } catch (const ResidueBatcher::Abort&) {   // a failed device batch
                    } catch (...) {
                        bad[t] = 1;
                    }
                    bat.finish();
                });
            try {
                bat.run();
            } catch (...) {
                for (auto& x : th) x.join();
                throw;
            }
            for (auto& x : th) x.join();
            for (int t : rt)
                REQUIRE(h, !bad[t], QX_ERR_NUMERICAL, "host residue solve of target %d failed", t);
        } else {
            for (int t : rt) solve(t, nullptr);
        }
        for (int t : rt) {
            if (code[t] == QX_QP_LINEARIZED) ++newton_fail;
            h2d(h, qp_energy + t, &Eh[t], 1);
            h2d(h, z_factor + t, &Zh[t], 1);
            h2d(h, sigma_re + t, &Sh[t].re, 1);
            h2d(h, sigma_im + t, &Sh[t].im, 1);
            if (state) h2d(h, state + t, &code[t], 1);
        }
        HIPCK(h, hipStreamSynchronize(h->stream));
    }
    resolve_profile(h);
    if (stats) {
        stats->chol_fallbacks = wr.chol_fallbacks;
        stats->singular_fallbacks = wr.singular;
        stats->residue_targets = residue_targets;
        stats->residue_jobs = cd.residue_jobs;
        stats->newton_failures = newton_fail;
        stats->pi_variant = wr.pi_variant;
        stats->seconds =
            std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
    return QX_OK;
}

template <class F>
qx_status guarded(qx_handle h, F fn) {
    if (!h) return QX_ERR_INVALID_INPUT;
    try {
        HIPCK(h, hipSetDevice(h->device));
        return fn();
    } catch (const QxFail& f) {
        return f.st;
    } catch (const std::exception& e) {
        set_err(h, "internal error: %s", e.what());
        return QX_ERR_NUMERICAL;
    }
}

}  // namespace

extern "C" {

qx_status qx_create(qx_handle* out, int device) {
    if (!out) return QX_ERR_INVALID_INPUT;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return QX_ERR_DEVICE;
    if (device < 0 || device >= n) return QX_ERR_INVALID_INPUT;
    qx_context* c = new qx_context();
    c->device = device;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreate(&c->stream) != hipSuccess) {
        delete c;
        return QX_ERR_DEVICE;
    }
    c->own_stream = true;
    // the device's copy of the no-allocation scratch, resolved once per device
    const auto sa = scratch_address(device);
    c->scratch_status = sa.first;
    c->scratch_dev = sa.second;
    *out = c;
    return QX_OK;
}

void qx_destroy(qx_handle h) {
    if (!h) return;
    (void)hipSetDevice(h->device);
    (void)hipStreamSynchronize(h->stream);
    for (auto& kv : h->bufs) (void)hipFree(kv.second.first);
    for (auto& p : h->ev_pending) {
        (void)hipEventDestroy(p.second.first);
        (void)hipEventDestroy(p.second.second);
    }
    for (hipEvent_t e : h->ev_free) (void)hipEventDestroy(e);
    delete_batch_plan(h->batch);
    for (int g = 0; g < h->fan.n; ++g) {
        (void)hipStreamSynchronize(h->fan.side[g]);
        (void)hipStreamDestroy(h->fan.side[g]);
        (void)hipEventDestroy(h->fan.join[g]);
    }
    if (h->fan.fork) (void)hipEventDestroy(h->fan.fork);
    if (h->pin) (void)hipHostFree(h->pin);
    if (h->own_stream) (void)hipStreamDestroy(h->stream);
    delete h;
}

qx_status qx_set_profiling(qx_handle h, int enable) {
    return guarded(h, [&] {
        resolve_profile(h);
        h->prof = enable != 0;
        for (int i = 0; i < QX_NPHASE; ++i) {
            h->phase_ms[i] = 0.0;
            h->phase_n[i] = 0;
        }
        return QX_OK;
    });
}

qx_status qx_get_profile(qx_handle h, double* ms, long long* counts) {
    return guarded(h, [&] {
        resolve_profile(h);
        for (int i = 0; i < QX_NPHASE; ++i) {
            if (ms) ms[i] = h->phase_ms[i];
            if (counts) counts[i] = h->phase_n[i];
        }
        return QX_OK;
    });
}

const char* qx_last_error(qx_handle h) { return h ? h->err.c_str() : "null handle"; }

qx_status qx_set_stream(qx_handle h, void* s) {
    return guarded(h, [&] {
        HIPCK(h, hipStreamSynchronize(h->stream));
        h->pin_off = 0;                   // every queued upload has been read
        if (h->own_stream) HIPCK(h, hipStreamDestroy(h->stream));
        if (s) {
            h->stream = static_cast<hipStream_t>(s);
            h->own_stream = false;
        } else {
            HIPCK(h, hipStreamCreate(&h->stream));
            h->own_stream = true;
        }
        return QX_OK;
    });
}

qx_status qx_synchronize(qx_handle h) {
    return guarded(h, [&] {
        HIPCK(h, hipStreamSynchronize(h->stream));
        return QX_OK;
    });
}

qx_status qx_release_workspace(qx_handle h) {
    return guarded(h, [&] {
        HIPCK(h, hipStreamSynchronize(h->stream));
        delete_batch_plan(h->batch);
        h->batch = nullptr;
        for (auto& kv : h->bufs) HIPCK(h, hipFree(kv.second.first));
        h->bufs.clear();
        h->grid_fr = nullptr;              // a new buffer may get a freed one's address
        h->zero_ptr = nullptr;
        return QX_OK;
    });
}

void qx_newton_config_default(qx_newton_config* c) {
    if (!c) return;
    c->max_iterations = 50;
    c->energy_tolerance = 1e-8;
    c->initial_damping = 1.0;
    c->min_damping = 0.1;
    c->max_damping = 1.0;
    c->max_step = 1.0;
    c->derivative_delta = 1e-6;
    c->use_richardson = 1;
    c->z_min = 0.1;
    c->z_max = 0.999;
    c->use_bisection_fallback = 1;
    c->max_bisection_iterations = 30;
}

void qx_g0w0_desc_default(qx_g0w0_desc* d) {
    if (!d) return;
    std::memset(d, 0, sizeof *d);
    d->nfreq = 48;
    d->x0 = 0.5;
    d->eta = 0.01;
    d->precision = QX_PREC_FP64;
    qx_newton_config_default(&d->newton);
}

qx_status qx_gauss_legendre_grid(int n, double* nodes, double* weights) {
    if (n <= 0 || !nodes || !weights) return QX_ERR_INVALID_INPUT;
    std::vector<double> x, w;
    gauss_legendre(n, x, w);
    std::copy(x.begin(), x.end(), nodes);
    std::copy(w.begin(), w.end(), weights);
    return QX_OK;
}

qx_status qx_scaled_legendre_roots(int nw, double x0, double* freqs, double* wts) {
    if (nw <= 0 || !freqs || !wts) return QX_ERR_INVALID_INPUT;
    std::vector<double> f, w;
    scaled_roots(nw, x0, f, w);
    std::copy(f.begin(), f.end(), freqs);
    std::copy(w.begin(), w.end(), wts);
    return QX_OK;
}

qx_status qx_rho_response_dev(qx_handle h, int precision, int nocc, int nvir, int naux,
                              const double* mo_energy, const double* B_ov, int nfreq,
                              const double* freqs, double* pi_out) {
    return guarded(h, [&] {
        REQUIRE(h, nocc > 0 && nvir > 0 && naux > 0 && nfreq > 0, QX_ERR_INVALID_INPUT,
                "dimensions must be positive");
        REQUIRE(h, precision >= 0 && precision <= 2, QX_ERR_INVALID_INPUT, "unknown precision");
        REQUIRE(h, mo_energy && B_ov && freqs && pi_out, QX_ERR_INVALID_INPUT, "NULL pointer");
        const int nov = nocc * nvir, kpad = round_up(nov, KPAD), npad = round_up(naux, NB);
        double* s = ws<double>(h, "pi_weights", (size_t)nfreq * kpad);
        PiOperands ops;
        float* s32 = precision == QX_PREC_BF16X3 ? ws<float>(h, "pi_weights32", (size_t)nfreq * kpad)
                                                 : nullptr;
        HIPCK(h, launch_pi_weights(mo_energy, nocc, nvir, nfreq, freqs, s, s32, h->stream));
        ops.s32 = s32;
        if (precision == QX_PREC_BF16X3) {
            const int brows = round_up(npad, 256);
            uint16_t* hi = ws<uint16_t>(h, "bt_hi", (size_t)brows * kpad);
            uint16_t* lo = ws<uint16_t>(h, "bt_lo", (size_t)brows * kpad);
            float* f32 = ws<float>(h, "bt_f32", (size_t)brows * kpad);
            HIPCK(h, launch_pi_prepare_mixed(B_ov, nov, naux, brows, kpad, hi, lo, f32, h->stream));
            ops.bt_hi = hi;
            ops.bt_lo = lo;
            ops.bt_f32 = f32;
            ops.kpad = kpad;
            ops.rows = brows;
        }
        const int pv = choose_variant(h, "pi_full", precision, pi_variant_count(precision),
                                      {naux, nov, nfreq}, [&](int v) {
            HIPCK(h, launch_pi_syrk(precision, B_ov, ops, s, nov, naux, naux,
                                    tune_nf(naux, nfreq), pi_out, PiOut::PiFull, h->stream, v));
        }, pi_candidates(precision));
        HIPCK(h, launch_pi_syrk(precision, B_ov, ops, s, nov, naux, naux, nfreq, pi_out,
                                PiOut::PiFull, h->stream, pv));
        HIPCK(h, hipStreamSynchronize(h->stream));
        return QX_OK;
    });
}

qx_status qx_screened_wc_dev(qx_handle h, int naux, int nfreq, const double* pi, double* wc_out,
                             int* fallbacks) {
    return guarded(h, [&] {
        REQUIRE(h, naux > 0 && nfreq > 0, QX_ERR_INVALID_INPUT, "dimensions must be positive");
        REQUIRE(h, pi && wc_out, QX_ERR_INVALID_INPUT, "NULL pointer");
        const int npad = round_up(naux, NB), nblk = npad / NB;
        double* eps = ws<double>(h, "eps_mat", (size_t)nfreq * npad * npad);
        hipLaunchKernelGGL(eps_from_pi_kernel, grid1d((size_t)nfreq * npad * npad), dim3(256), 0,
                           h->stream, pi, naux, npad, nfreq, eps);
        HIPCK(h, hipGetLastError());
        double* linv = ws<double>(h, "linv_diag", (size_t)nfreq * nblk * NB * NB);
        int* fail = ws<int>(h, "chol_fail", nfreq);
        fill_i32(h, fail, 0, nfreq);
        double* x = ws<double>(h, "xinv", (size_t)nfreq * npad * npad);
        int groups = 1;
        const StreamFan* fan = potrf_fan(h, groups);
        HIPCK(h, launch_potrf_batched(eps, x, linv, npad, nfreq, fail, h->stream, fan, groups));
        std::vector<int> fh(nfreq);
        d2h(h, fh.data(), fail, nfreq);
        int nfail = 0;
        for (int v : fh) nfail += v != 0;
        if (nfail == 0) {
            HIPCK(h, launch_wc_from_linv(x, wc_out, npad, naux, nfreq, h->stream));
        } else {
            double* work = ws<double>(h, "gj_work", (size_t)nfreq * naux * 2 * naux + nfreq);
            HIPCK(h, launch_lu_fallback_wc(pi, 0, work, wc_out, naux, nfreq, nullptr, h->stream));
        }
        HIPCK(h, hipStreamSynchronize(h->stream));
        if (fallbacks) *fallbacks = nfail;
        return QX_OK;
    });
}

qx_status qx_wmn_imag_dev(qx_handle h, int precision, int nocc, int nvir, int naux,
                          const double* mo_energy, const double* B_ov, int nfreq,
                          const double* freqs, int npair, const double* B_pair, double* wnm) {
    return guarded(h, [&] {
        REQUIRE(h, nocc > 0 && nvir > 0 && naux > 0 && nfreq > 0 && npair >= 0,
                QX_ERR_INVALID_INPUT, "dimensions must be positive");
        REQUIRE(h, precision >= 0 && precision <= 2, QX_ERR_INVALID_INPUT, "unknown precision");
        REQUIRE(h, mo_energy && B_ov && freqs && (npair == 0 || (B_pair && wnm)),
                QX_ERR_INVALID_INPUT, "NULL pointer");
        build_wnm(h, precision, nocc, nvir, naux, mo_energy, B_ov, nfreq, freqs, npair, B_pair, wnm);
        HIPCK(h, hipStreamSynchronize(h->stream));
        return QX_OK;
    });
}

qx_status qx_sigma_cd_dev(qx_handle h, const qx_g0w0_desc* d, const double* mo_energy,
                          const double* mo_energy_ref, const double* B_ov, const double* B_pair,
                          const int* pair_target, const int* pair_m, const int* targets,
                          const double* freqs, const double* wts, const double* wnm, int nq,
                          const double* energies, double* sigma) {
    return guarded(h, [&] {
        validate_desc(h, d);
        REQUIRE(h, nq > 0 && energies && sigma && freqs && wts && targets && mo_energy,
                QX_ERR_INVALID_INPUT, "NULL pointer or nq <= 0");
        HostCD cd;
        double* zero_static = ws<double>(h, "zero_static", d->ntarget);
        fill_f64(h, zero_static, 0.0, d->ntarget);
        setup_cd(h, d, mo_energy, mo_energy_ref, B_ov, B_pair, pair_target, pair_m, targets,
                 zero_static, freqs, wts, wnm, cd);
        std::vector<double> E((size_t)d->ntarget * nq);
        d2h(h, E.data(), energies, E.size());
        std::vector<double> out(E.size() * 2);
        for (int t = 0; t < d->ntarget; ++t)
            for (int q0 = 0; q0 < nq; q0 += 5) {
                const int cnt = std::min(5, nq - q0);
                C2 s[5];
                cd.eval(t, &E[(size_t)t * nq + q0], cnt, s);
                for (int q = 0; q < cnt; ++q) {
                    out[((size_t)t * nq + q0 + q) * 2] = s[q].re;
                    out[((size_t)t * nq + q0 + q) * 2 + 1] = s[q].im;
                }
            }
        h2d(h, sigma, out.data(), out.size());
        return QX_OK;
    });
}

namespace {

double d2h_scalar(qx_handle h, const double* p) {
    double v = 0.0;
    d2h(h, &v, p, 1);
    return v;
}

// eps[f] = L L^T with X = L^-1 in place (lower blocks), for every frequency; returns the
// number of frequencies whose Cholesky failed.  eps is [nf][npad][npad] with an identity
// padding block.
int factor_all(qx_handle h, double* eps, int npad, int nf, double** x_out) {
    const int nblk = npad / NB;
    double* x = ws<double>(h, "xinv", (size_t)nf * npad * npad);
    double* linv = ws<double>(h, "linv_diag", (size_t)nf * nblk * NB * NB);
    int* fail = ws<int>(h, "chol_fail", nf);
    fill_i32(h, fail, 0, nf);
    int groups = 1;
    const StreamFan* fan = potrf_fan(h, groups);
    HIPCK(h, launch_potrf_batched(eps, x, linv, npad, nf, fail, h->stream, fan, groups));
    std::vector<int> fh(nf);
    d2h(h, fh.data(), fail, nf);
    int nfail = 0;
    for (int v : fh) nfail += v != 0;
    *x_out = x;
    return nfail;
}

// Gauss-Jordan inverse minus identity of every eps[f] ([nf][npad][npad]) into wc; raises
// QX_ERR_NUMERICAL naming the first singular frequency.
void gj_all(qx_handle h, const double* eps, int npad, int nf, double* wc, const char* what) {
    double* work = ws<double>(h, "gj_work", (size_t)nf * npad * 2 * npad + nf);
    HIPCK(h, launch_lu_fallback_wc(eps, 1, work, wc, npad, nf, nullptr, h->stream));
    std::vector<int> sing(nf);
    d2h(h, sing.data(), reinterpret_cast<int*>(work + (size_t)nf * npad * 2 * npad), nf);
    for (int f = 0; f < nf; ++f)
        REQUIRE(h, sing[f] == 0, QX_ERR_NUMERICAL, "%s singular at frequency %d", what, f);
}

// Principal square root of the SPD metric v [n][n] into vs [npad][npad] (zero padding) by
// the coupled Newton-Schulz iteration on MFMA GEMMs: Y0 = v / c, Z0 = I,
// T = (3 I - Z Y) / 2, Y <- Y T, Z <- T Z, vs = sqrt(c) Y with c = |v|_F.  Replaces the
// eigendecomposition of compute_metric_sqrt (quasix_core/src/dielectric/screening.rs:
// 580-639) and keeps its acceptance test |vs vs^T - v|_max <= 1e-10.
void metric_sqrt(qx_handle h, const double* v, int n, int npad, double* vs) {
    hipStream_t st = h->stream;
    const size_t nn = (size_t)npad * npad;
    double* A = ws<double>(h, "ns_a", nn);
    double* Y = ws<double>(h, "ns_y", nn);
    double* Z = ws<double>(h, "ns_z", nn);
    double* Y2 = ws<double>(h, "ns_y2", nn);
    double* Z2 = ws<double>(h, "ns_z2", nn);
    double* T = ws<double>(h, "ns_t", nn);
    double* red = ws<double>(h, "red", 1025);
    HIPCK(h, launch_pad_square(v, 0, n, npad, 1, false, A, st));
    HIPCK(h, launch_reduce(A, nullptr, nn, 0, red, st));
    const double c = std::sqrt(d2h_scalar(h, red));
    REQUIRE(h, c > 0.0 && std::isfinite(c), QX_ERR_NUMERICAL,
            "metric square root: |v|_F = %g is not positive and finite", c);
    HIPCK(h, launch_pad_square(v, 0, n, npad, 1, false, Y, st));
    HIPCK(h, launch_axpby(0.0, nullptr, 1.0 / c, Y, nn, st));
    HIPCK(h, launch_pad_square(nullptr, 0, n, npad, 1, false, Z, st));
    HIPCK(h, launch_add_identity(Z, npad, n, 1, 1.0, st));
    double prev = INFINITY;
    for (int it = 0; it < 100; ++it) {
        HIPCK(h, launch_strided_gemm(false, false, Z, 0, Y, 0, T, 0, npad, 1, st));
        HIPCK(h, launch_axpby(0.0, nullptr, -0.5, T, nn, st));
        HIPCK(h, launch_add_identity(T, npad, n, 1, 1.5, st));
        HIPCK(h, launch_strided_gemm(false, false, Y, 0, T, 0, Y2, 0, npad, 1, st));
        HIPCK(h, launch_strided_gemm(false, false, T, 0, Z, 0, Z2, 0, npad, 1, st));
        HIPCK(h, launch_reduce(Y2, Y, nn, 1, red, st));
        const double diff = d2h_scalar(h, red);
        std::swap(Y, Y2);
        std::swap(Z, Z2);
        if (!(diff > 1e-15) || (diff < 1e-10 && diff >= prev)) break;   // converged or at the rounding floor
        prev = diff;
    }
    // symmetrise, scale, and check the reconstruction
    HIPCK(h, launch_sym_compact(Y, npad, n, 1, T, st));
    HIPCK(h, launch_pad_square(T, 0, n, npad, 1, false, vs, st));
    HIPCK(h, launch_axpby(0.0, nullptr, std::sqrt(c), vs, nn, st));
    HIPCK(h, launch_strided_gemm(false, true, vs, 0, vs, 0, T, 0, npad, 1, st));
    HIPCK(h, launch_reduce(T, A, nn, 1, red, st));
    const double err = d2h_scalar(h, red);
    REQUIRE(h, err <= 1e-10, QX_ERR_NUMERICAL,
            "metric square root reconstruction error: %.3e", err);
}

}  // namespace

qx_status qx_sigma_c_pyscf_cd_dev(qx_handle h, int nmo, int naux, int nfreq, double eta,
                                  const double* mo_energy, const double* p0_iw,
                                  const double* lpq, const double* freqs,
                                  const double* weights, double* sigma) {
    return guarded(h, [&] {
        REQUIRE(h, nmo > 0 && naux > 0 && nfreq > 0, QX_ERR_INVALID_INPUT,
                "dimensions must be positive (nmo=%d naux=%d nfreq=%d)", nmo, naux, nfreq);
        REQUIRE(h, mo_energy && p0_iw && lpq && freqs && weights && sigma, QX_ERR_INVALID_INPUT,
                "NULL pointer");
        REQUIRE(h, (long long)nmo * nmo * round_up(naux, NB) < (1ll << 40), QX_ERR_INVALID_INPUT,
                "nmo^2 * naux too large");
        hipStream_t st = h->stream;
        const int npad = round_up(naux, NB), nblk = npad / NB, nf = nfreq;
        const int npair = nmo * nmo;
        // eps = I - P0 (identity padding), Cholesky with X = L^-1
        double* eps = ws<double>(h, "eps_mat", (size_t)nf * npad * npad);
        hipLaunchKernelGGL(eps_from_pi_kernel, grid1d((size_t)nf * npad * npad), dim3(256), 0, st,
                           p0_iw, naux, npad, nf, eps);
        HIPCK(h, hipGetLastError());
        double* x = nullptr;
        const int nfail = factor_all(h, eps, npad, nf, &x);
        // pair rows of every (n, m): b^T Wc b' = u^T Wc u - v^T Wc v
        double* u = ws<double>(h, "cd0_u", (size_t)npair * npad);
        double* v = ws<double>(h, "cd0_v", (size_t)npair * npad);
        int* anyv = ws<int>(h, "cd0_anyv", 1);
        HIPCK(h, launch_lpq_pair_rows(lpq, naux, nmo, npad, u, v, anyv, st));
        int anyv_h = 0;
        d2h(h, &anyv_h, anyv, 1);
        double* wu = ws<double>(h, "cd0_wu", (size_t)nf * npair);
        double* wv = anyv_h ? ws<double>(h, "cd0_wv", (size_t)nf * npair) : nullptr;
        double* wpart = ws<double>(h, "wpart", (size_t)nf * nblk * npair);
        if (nfail == 0) {
            HIPCK(h, launch_wnm(QX_PREC_FP64, u, x, npair, npad, nf, wpart, wu, nullptr, st));
            if (wv) HIPCK(h, launch_wnm(QX_PREC_FP64, v, x, npair, npad, nf, wpart, wv, nullptr, st));
        } else {
            // the reference inverts eps with LU and reports a singular one as an error
            hipLaunchKernelGGL(eps_from_pi_kernel, grid1d((size_t)nf * npad * npad), dim3(256), 0,
                               st, p0_iw, naux, npad, nf, eps);
            HIPCK(h, hipGetLastError());
            double* wc = ws<double>(h, "cd0_wc", (size_t)nf * npad * npad);
            gj_all(h, eps, npad, nf, wc, "Epsilon");
            HIPCK(h, launch_wnm_from_wc(u, wc, npair, npad, nf, wpart, wu, st));
            if (wv) HIPCK(h, launch_wnm_from_wc(v, wc, npair, npad, nf, wpart, wv, st));
        }
        HIPCK(h, launch_sigma_static_cd(wu, wv, nmo, nf, mo_energy, freqs, weights, eta, sigma, st));
        HIPCK(h, hipStreamSynchronize(st));
        return QX_OK;
    });
}

qx_status qx_screening_w_dev(qx_handle h, int naux, int nfreq, const double* p0,
                             const double* metric, const double* v_sqrt, double* w_out) {
    return guarded(h, [&] {
        REQUIRE(h, naux > 0 && nfreq > 0, QX_ERR_INVALID_INPUT,
                "dimensions must be positive (naux=%d nfreq=%d)", naux, nfreq);
        REQUIRE(h, p0 && w_out, QX_ERR_INVALID_INPUT, "NULL pointer");
        hipStream_t st = h->stream;
        const int npad = round_up(naux, NB), nf = nfreq;
        const size_t nn = (size_t)npad * npad;
        double* vs = nullptr;
        if (v_sqrt || metric) {
            vs = ws<double>(h, "scr_vs", nn);
            if (v_sqrt) HIPCK(h, launch_pad_square(v_sqrt, 0, naux, npad, 1, false, vs, st));
            else metric_sqrt(h, metric, naux, npad, vs);
        }
        // M_f = vs P0_f vs^T, eps_f = I - (M_f + M_f^T) / 2 (screening.rs:334-356)
        double* P = ws<double>(h, "scr_p", nf * nn);
        HIPCK(h, launch_pad_square(p0, (size_t)naux * naux, naux, npad, nf, false, P, st));
        double* T = vs ? ws<double>(h, "scr_t", nf * nn) : nullptr;
        if (vs) {
            HIPCK(h, launch_strided_gemm(false, false, vs, 0, P, nn, T, nn, npad, nf, st));
            HIPCK(h, launch_strided_gemm(false, true, T, nn, vs, 0, P, nn, npad, nf, st));
        }
        double* eps = ws<double>(h, "eps_mat", nf * nn);
        HIPCK(h, launch_eps_sym(P, npad, nf, eps, st));
        // eps^-1: Cholesky (X^T X), LU on failure, an error if singular (screening.rs:358-383)
        double* x = nullptr;
        const int nfail = factor_all(h, eps, npad, nf, &x);
        double* einv = ws<double>(h, "scr_einv", nf * nn);
        if (nfail == 0) {
            HIPCK(h, launch_zero_upper_blocks(x, npad, nf, st));
            HIPCK(h, launch_strided_gemm(true, false, x, nn, x, nn, einv, nn, npad, nf, st));
        } else {
            HIPCK(h, launch_eps_sym(P, npad, nf, eps, st));
            gj_all(h, eps, npad, nf, einv, "Matrix inversion failed: epsilon");
            HIPCK(h, launch_add_identity(einv, npad, npad, nf, 1.0, st));
        }
        // W_f = vs eps_f^-1 vs^T, symmetrised (screening.rs:385-411)
        const double* wsrc = einv;
        if (vs) {
            HIPCK(h, launch_strided_gemm(false, false, vs, 0, einv, nn, T, nn, npad, nf, st));
            HIPCK(h, launch_strided_gemm(false, true, T, nn, vs, 0, P, nn, npad, nf, st));
            wsrc = P;
        }
        HIPCK(h, launch_sym_compact(wsrc, npad, naux, nf, w_out, st));
        double* red = ws<double>(h, "red", 1025);
        HIPCK(h, launch_reduce(w_out, nullptr, (size_t)nf * naux * naux, 2, red, st));
        const double mx = d2h_scalar(h, red);
        REQUIRE(h, std::isfinite(mx), QX_ERR_NUMERICAL, "W contains NaN or Inf values");
        return QX_OK;
    });
}

qx_status qx_assemble_blocks_dev(qx_handle h, int nocc, int nvir, int naux, const double* iaP,
                                 const double* miP, const double* abP, int ntarget,
                                 const int* targets, double* B_ov_sym, double* mi_sym,
                                 int pair_capacity, double* B_pair, int* pair_target,
                                 int* pair_m, int* npair) {
    return guarded(h, [&] {
        REQUIRE(h, nocc > 0 && nvir > 0 && naux > 0 && ntarget >= 0, QX_ERR_INVALID_INPUT,
                "Invalid dimensions nocc=%d nvir=%d naux=%d ntarget=%d", nocc, nvir, naux,
                ntarget);
        REQUIRE(h, iaP && miP, QX_ERR_INVALID_INPUT, "iaP and miP are required");
        REQUIRE(h, ntarget == 0 || (targets && abP && B_pair && pair_target && pair_m && npair),
                QX_ERR_INVALID_INPUT, "pair outputs, targets and abP are required for targets");
        const int nmo = nocc + nvir;
        hipStream_t st = h->stream;
        if (npair) *npair = 0;
        HIPCK(h, launch_sym_blocks(iaP, miP, abP, nocc, nvir, naux, B_ov_sym, mi_sym, st));
        if (ntarget > 0) {
            REQUIRE(h, (long long)ntarget * nmo < (1ll << 31), QX_ERR_INVALID_INPUT,
                    "ntarget * nmo overflows int");
            std::vector<int> tg(ntarget);
            d2h(h, tg.data(), targets, ntarget);
            for (int t = 0; t < ntarget; ++t)
                REQUIRE(h, tg[t] >= 0 && tg[t] < nmo, QX_ERR_DIMENSION_MISMATCH,
                        "target %d is outside the %d orbitals", tg[t], nmo);
            const size_t rows = (size_t)ntarget * nmo;
            int* flags = ws<int>(h, "asm_flags", rows);
            int* pos = ws<int>(h, "asm_pos", rows);
            int* total = ws<int>(h, "asm_total", 1);
            HIPCK(h, launch_assemble_pairs(iaP, miP, abP, nocc, nvir, naux, targets, ntarget,
                                           flags, pos, total, st));
            int n = 0;
            d2h(h, &n, total, 1);
            REQUIRE(h, n <= pair_capacity, QX_ERR_DIMENSION_MISMATCH,
                    "pair_capacity %d is below the %d non-zero target-partner rows",
                    pair_capacity, n);
            HIPCK(h, launch_scatter_pairs(iaP, miP, abP, nocc, nvir, naux, targets, ntarget,
                                          flags, pos, B_pair, pair_target, pair_m, st));
            *npair = n;
        }
        HIPCK(h, hipStreamSynchronize(st));
        return QX_OK;
    });
}

qx_status qx_solve_qp_linearized(int nmo, const double* mo_energy, const double* sigma_x,
                                 const double* sigma_c, const double* vxc, double* qp_energy,
                                 double* z_factor) {
    if (nmo <= 0 || !mo_energy || !sigma_x || !sigma_c || !vxc || !qp_energy || !z_factor)
        return QX_ERR_INVALID_INPUT;
    for (int i = 0; i < nmo; ++i) {
        qp_energy[i] = mo_energy[i] + sigma_x[i] + sigma_c[i] - vxc[i];
        z_factor[i] = 1.0;
        if (!std::isfinite(qp_energy[i])) return QX_ERR_NUMERICAL;
    }
    return QX_OK;
}

qx_status qx_df_mo_transform_dev(qx_handle h, int nao, int nocc, int nvir, int naux,
                                 const double* j3c_ao, const double* c_occ, const double* c_vir,
                                 double* b_ov) {
    return guarded(h, [&] {
        REQUIRE(h, nao > 0 && nocc > 0 && nvir > 0 && naux > 0, QX_ERR_INVALID_INPUT,
                "Invalid dimensions nao=%d nocc=%d nvir=%d naux=%d", nao, nocc, nvir, naux);
        REQUIRE(h, j3c_ao && c_occ && c_vir && b_ov, QX_ERR_INVALID_INPUT, "NULL pointer");
        REQUIRE(h, nocc <= 65535 && (long long)nao * naux < (1ll << 31), QX_ERR_INVALID_INPUT,
                "dimensions exceed the transform's grid limits");
        double* tmp = ws<double>(h, "mo_tmp", (size_t)nocc * nao * naux);
        HIPCK(h, launch_mo_transform(j3c_ao, c_occ, c_vir, nao, nocc, nvir, naux, tmp, b_ov,
                                     h->stream));
        HIPCK(h, hipStreamSynchronize(h->stream));
        return QX_OK;
    });
}

qx_status qx_exchange_diag_dev(qx_handle h, int nmo, int nocc, int naux, const double* mi,
                               double* sigma_x) {
    return guarded(h, [&] {
        REQUIRE(h, nmo > 0 && nocc > 0 && nocc <= nmo && naux > 0, QX_ERR_INVALID_INPUT,
                "Invalid dimensions nmo=%d nocc=%d naux=%d", nmo, nocc, naux);
        REQUIRE(h, mi && sigma_x, QX_ERR_INVALID_INPUT, "NULL pointer");
        HIPCK(h, launch_exchange_diag(mi, nmo, nocc, naux, sigma_x, h->stream));
        HIPCK(h, hipStreamSynchronize(h->stream));
        return QX_OK;
    });
}

qx_status qx_sigma_ac_dev(qx_handle h, const qx_g0w0_desc* d, const double* mo_energy,
                          const double* B_ov, const double* B_pair, const int* pair_target,
                          const int* pair_m, const int* targets, double iw_cutoff,
                          double* sigma, double* omega, int* nws_out) {
    return guarded(h, [&] {
        validate_desc(h, d);
        REQUIRE(h, mo_energy && B_ov && targets && sigma && omega && nws_out,
                QX_ERR_INVALID_INPUT, "required pointer is NULL");
        REQUIRE(h, d->npair == 0 || (B_pair && pair_target && pair_m), QX_ERR_INVALID_INPUT,
                "pair arrays are NULL");
        const int nf = d->nfreq, nt = d->ntarget;
        std::vector<double> fr, wt;
        scaled_roots(nf, d->x0, fr, wt);
        double* fr_dev = ws<double>(h, "freqs", nf);
        double* wt_dev = ws<double>(h, "wts", nf);
        h2d(h, fr_dev, fr.data(), nf);
        h2d(h, wt_dev, wt.data(), nf);
        double* wnm = ws<double>(h, "wnm", (size_t)nf * std::max(d->npair, 1));
        HostCD cd;
        setup_cd(h, d, mo_energy, nullptr, B_ov, B_pair, pair_target, pair_m, targets, nullptr,
                 fr_dev, wt_dev, wnm, cd);
        build_wnm(h, d->precision, d->nocc, d->nvir, d->naux, mo_energy, B_ov, nf, fr_dev,
                  d->npair, B_pair, wnm);
        // grid nu = {0} + {w_k < iw_cutoff} (correlation_ac.rs:270-281)
        std::vector<double> nu{0.0};
        for (int k = 0; k < nf; ++k)
            if (fr[k] < iw_cutoff) nu.push_back(fr[k]);
        const int nws = std::min((int)nu.size(), nf + 1);
        double* nu_dev = ws<double>(h, "ac_nu", nws);
        h2d(h, nu_dev, nu.data(), nws);
        {
            Phase ph(h, QX_PHASE_NEWTON);
            HIPCK(h, launch_sigma_ac(cd.prob, nu_dev, nws, sigma, h->stream));
        }
        std::vector<double> om((size_t)nt * nws * 2, 0.0);
        for (int t = 0; t < nt; ++t)
            for (int q = 0; q < nws; ++q)
                om[((size_t)t * nws + q) * 2 + 1] = (cd.targets[t] < d->nocc ? -1.0 : 1.0) * nu[q];
        h2d(h, omega, om.data(), om.size());
        HIPCK(h, hipStreamSynchronize(h->stream));
        resolve_profile(h);
        *nws_out = nws;
        return QX_OK;
    });
}

qx_status qx_g0w0_cd_dev(qx_handle h, const qx_g0w0_desc* d, const double* mo_energy,
                         const double* mo_energy_ref, const double* B_ov, const double* B_pair,
                         const int* pair_target, const int* pair_m, const int* targets,
                         const double* static_terms, const double* guess, double* qp_energy,
                         double* z_factor, double* sigma_re, double* sigma_im, int* state,
                         qx_stats* stats) {
    return guarded(h, [&] {
        return run_g0w0(h, d, mo_energy, mo_energy_ref, B_ov, B_pair, pair_target, pair_m,
                        targets, static_terms, guess, qp_energy, z_factor, sigma_re, sigma_im,
                        state, stats);
    });
}

namespace {
// Upper estimate of the workspace of one ov-only solve (the buffers build_wnm, setup_cd and
// run_g0w0 request), without the residue path's per-job buffers: the no-allocation entry
// rejects problems beyond its scratch with it before launching anything.  The pair count
// assumes every target pairs with max(nocc, nvir) partners.
size_t g0w0_ov_workspace_bytes(const qx_g0w0_desc* d) {
    const size_t nf = d->nfreq, nt = d->ntarget, naux = d->naux;
    const size_t nov = (size_t)d->nocc * d->nvir;
    const size_t npad = round_up(d->naux, NB), nblk = npad / NB;
    const size_t kpad = round_up((int)std::min<size_t>(nov, 1u << 30), KPAD);
    const size_t npair = nt * (size_t)std::max(d->nocc, d->nvir);
    const size_t sq = nf * npad * npad;
    size_t b = 0;
    auto add = [&](size_t bytes) { b += (std::max<size_t>(bytes, 256) + 255) / 256 * 256; };
    add(nf * 8); add(nf * 8);                                   // freqs, wts
    add(nt * 8); add((nt + 1) * 4); add(nt * 4); add(nt * 4);   // static, offsets, state, iters
    add((nt + 1) * 4); add(npair * 4); add(npair * 4);          // ov offsets, pair t, pair m
    add(npair * naux * 8);                                       // ov pair rows
    add(nf * npair * 8);                                         // wnm
    add(nf * kpad * 8); add(sq * 8); add(sq * 8);                // weights, eps -> L, X
    add(nf * nblk * NB * NB * 8); add(nf * 4);                   // diagonal inverses, flags
    add(npair * npad * 8); add(nf * nblk * npair * 8);           // padded rows, partials
    if (d->precision == QX_PREC_BF16X3) {
        const size_t brows = round_up((int)npad, 256), npp = round_up((int)npair, 128);
        add(nf * kpad * 4);                                      // FP32 weights
        add(brows * kpad * 2); add(brows * kpad * 2); add(brows * kpad * 4);
        add(npp * npad * 2); add(npp * npad * 2); add(sq * 2); add(sq * 2);
    }
    return b;
}

// qx_g0w0_cd_ov_dev without the error guard: gathers the ov-only pair rows on the device.
qx_status run_g0w0_ov(qx_handle h, const qx_g0w0_desc* d, const double* mo_energy,
                      const double* mo_energy_ref, const double* B_ov, const int* targets,
                      const double* static_terms, const double* guess, double* qp_energy,
                      double* z_factor, double* sigma_re, double* sigma_im, int* state,
                      qx_stats* stats) {
    REQUIRE(h, d != nullptr && targets != nullptr, QX_ERR_INVALID_INPUT, "NULL argument");
    validate_desc(h, d);
    const int nt = d->ntarget;
    REQUIRE(h, nt > 0, QX_ERR_INVALID_INPUT, "ntarget must be positive");
    REQUIRE(h, mo_energy && B_ov, QX_ERR_INVALID_INPUT, "required pointer is NULL");
    std::vector<int> tg(nt), off(nt + 1, 0);
    // the targets' download is queued first and the Pi operand preparation (which does not
    // depend on them) right behind it; the host waits for the copy alone, so the GPU works
    // while the host builds the pair offsets (with no pinned ring: a plain download)
    double *fr_dev, *wt_dev;
    cd_grid(h, d->nfreq, d->x0, &fr_dev, &wt_dev);
    char* pin = h->arena ? nullptr : pin_take(h, (size_t)nt * sizeof(int));
    hipEvent_t ev = nullptr;
    if (pin) {
        HIPCK(h, hipMemcpyAsync(pin, targets, (size_t)nt * sizeof(int), hipMemcpyDeviceToHost,
                                h->stream));
        ev = take_event(h);
        HIPCK(h, hipEventRecord(ev, h->stream));
    }
    double* s_unused = nullptr;
    {
        Phase ph(h, QX_PHASE_PREP);
        pi_prep(h, d->precision, d->nocc, d->nvir, d->naux, d->nfreq, mo_energy, B_ov, fr_dev,
                &s_unused, true);
    }
    if (ev) {
        HIPCK(h, hipEventSynchronize(ev));
        std::memcpy(tg.data(), pin, (size_t)nt * sizeof(int));
        h->ev_free.push_back(ev);
    } else {
        d2h(h, tg.data(), targets, nt);
    }
    for (int t = 0; t < nt; ++t) {
        REQUIRE(h, tg[t] >= 0 && tg[t] < d->nocc + d->nvir, QX_ERR_INVALID_INPUT,
                "target %d out of range", tg[t]);
        off[t + 1] = off[t] + (tg[t] < d->nocc ? d->nvir : d->nocc);
    }
    const int npair = off[nt];
    int* off_dev = ws<int>(h, "ov_off", nt + 1);
    int* pt = ws<int>(h, "ov_pt", npair);
    int* pm = ws<int>(h, "ov_pm", npair);
    // the pair rows go straight into the zero-padded [npair][npad] buffer the Wc_nm kernels
    // read (build_wnm's "bpairs"); only Wc_nm, the residues and Newton read them, so they are
    // queued behind the Pi launch
    const int npad = round_up(d->naux, NB);
    double* rows = ws<double>(h, "bpairs", (size_t)std::max(npair, 1) * npad);
    const int maxcnt = std::max(d->nocc, d->nvir);
    const std::function<void()> pair_rows = [&] {
        h2d(h, off_dev, off.data(), nt + 1);
        hipLaunchKernelGGL(ov_pairs_kernel, dim3(std::min(maxcnt, 1024), nt), dim3(256), 0,
                           h->stream, B_ov, d->nocc, d->nvir, d->naux, targets, off_dev, nt, pt,
                           pm, rows, npad);
        HIPCK(h, hipGetLastError());
    };
    // without the pinned ring (caller's arena) the offsets' upload synchronises, which behind
    // Pi would hold the factorisation's launches back: queue the pair rows here then
    if (h->arena) pair_rows();
    qx_g0w0_desc dd = *d;
    dd.npair = npair;
    const KnownPairs known{&tg, &off, off_dev, true, true, h->arena ? nullptr : &pair_rows};
    return run_g0w0(h, &dd, mo_energy, mo_energy_ref, B_ov, rows, pt, pm, targets, static_terms,
                    guess, qp_energy, z_factor, sigma_re, sigma_im, state, stats, &known);
}
}  // namespace

qx_status qx_g0w0_cd_ov_dev(qx_handle h, const qx_g0w0_desc* d, const double* mo_energy,
                            const double* mo_energy_ref, const double* B_ov, const int* targets,
                            const double* static_terms, const double* guess, double* qp_energy,
                            double* z_factor, double* sigma_re, double* sigma_im, int* state,
                            qx_stats* stats) {
    return guarded(h, [&] {
        return run_g0w0_ov(h, d, mo_energy, mo_energy_ref, B_ov, targets, static_terms, guess,
                           qp_energy, z_factor, sigma_re, sigma_im, state, stats);
    });
}

qx_status qx_g0w0_cd_ov_noalloc_dev(qx_handle h, const qx_g0w0_desc* d,
                                    const double* mo_energy, const double* mo_energy_ref,
                                    const double* B_ov, const int* targets,
                                    const double* static_terms, const double* guess,
                                    double* qp_energy, double* z_factor, double* sigma_re,
                                    double* sigma_im, int* state, qx_stats* stats) {
    return guarded(h, [&] {
        REQUIRE(h, h->scratch_status == hipSuccess && h->scratch_dev, QX_ERR_DEVICE,
                "scratch address of device %d unavailable: %s", h->device,
                hipGetErrorString(h->scratch_status));
        REQUIRE(h, !h->prof, QX_ERR_INVALID_INPUT,
                "profiling creates events; disable it for the no-allocation entry point");
        // carve every workspace buffer of this call from the scratch; cleared on exit so
        // the handle's other entry points allocate as before
        struct Arena {
            qx_handle h;
            Arena(qx_handle h_) : h(h_) {
                h->arena = h->scratch_dev;
                h->arena_size = QXGW_NOALLOC_SCRATCH_BYTES;
                h->arena_used = 0;
                h->arena_bufs.clear();
            }
            ~Arena() {
                h->arena = nullptr;
                h->arena_size = h->arena_used = 0;
                h->arena_bufs.clear();
            }
        } arena(h);
        validate_desc(h, d);
        const size_t need = g0w0_ov_workspace_bytes(d);
        REQUIRE(h, need <= QXGW_NOALLOC_SCRATCH_BYTES, QX_ERR_INVALID_INPUT,
                "problem too large for the no-allocation entry point: about %zu bytes of "
                "workspace needed, the scratch holds %llu (nocc=%d nvir=%d naux=%d nfreq=%d "
                "ntarget=%d)",
                need, (unsigned long long)QXGW_NOALLOC_SCRATCH_BYTES, d->nocc, d->nvir, d->naux,
                d->nfreq, d->ntarget);
        return run_g0w0_ov(h, d, mo_energy, mo_energy_ref, B_ov, targets, static_terms, guess,
                           qp_energy, z_factor, sigma_re, sigma_im, state, stats);
    });
}

qx_status qx_g0w0_adjoint_ov_dev(qx_handle h, const qx_g0w0_desc* d, const double* mo_energy,
                                 const double* B_ov, const int* targets,
                                 const double* qp_energy, const double* grad_qp,
                                 double* grad_B, double* grad_eps) {
    return guarded(h, [&] {
        validate_desc(h, d);
        REQUIRE(h, mo_energy && B_ov && targets && qp_energy && grad_qp && grad_B && grad_eps,
                QX_ERR_INVALID_INPUT, "NULL pointer");
        const int nocc = d->nocc, nvir = d->nvir, naux = d->naux, nf = d->nfreq;
        const int nt = d->ntarget, nmo = nocc + nvir, nov = nocc * nvir;
        const int npad = round_up(naux, NB);
        hipStream_t st = h->stream;
        // ov pairs (as the forward)
        std::vector<int> tg(nt), off(nt + 1, 0);
        d2h(h, tg.data(), targets, nt);
        for (int t = 0; t < nt; ++t) {
            REQUIRE(h, tg[t] >= 0 && tg[t] < nmo, QX_ERR_INVALID_INPUT, "target out of range");
            off[t + 1] = off[t] + (tg[t] < nocc ? nvir : nocc);
        }
        const int npair = off[nt];
        int* off_dev = ws<int>(h, "adj_off", nt + 1);
        h2d(h, off_dev, off.data(), nt + 1);
        int* pt = ws<int>(h, "adj_pt", npair);
        int* pm = ws<int>(h, "adj_pm", npair);
        double* rows = ws<double>(h, "adj_rows", (size_t)npair * naux);
        hipLaunchKernelGGL(ov_pairs_kernel, dim3(std::min(std::max(nocc, nvir), 1024), nt),
                           dim3(256), 0, st, B_ov, nocc, nvir, naux, targets, off_dev, nt, pt, pm,
                           rows, naux);
        HIPCK(h, hipGetLastError());
        // forward checkpoints in FP64: X_k = L_k^-1 for every frequency, Wc_nm
        std::vector<double> fr, wt;
        scaled_roots(nf, d->x0, fr, wt);
        double* fr_dev = ws<double>(h, "freqs", nf);
        double* wt_dev = ws<double>(h, "wts", nf);
        h2d(h, fr_dev, fr.data(), nf);
        h2d(h, wt_dev, wt.data(), nf);
        double* wnm = ws<double>(h, "wnm", (size_t)nf * npair);
        WnmResult wr = build_wnm(h, QX_PREC_FP64, nocc, nvir, naux, mo_energy, B_ov, nf, fr_dev,
                                 npair, rows, wnm);
        REQUIRE(h, wr.chol_fallbacks == 0, QX_ERR_NOT_IMPLEMENTED,
                "adjoint needs SPD dielectric matrices (Cholesky failed at %d frequencies)",
                wr.chol_fallbacks);
        double* X = ws<double>(h, "xinv", (size_t)nf * npad * npad);
        HIPCK(h, launch_zero_upper_blocks(X, npad, nf, st));
        fill_f64(h, grad_B, 0.0, (size_t)nov * naux);
        fill_f64(h, grad_eps, 0.0, nmo);
        // adjoint of the quadrature + Newton
        const int sldp = round_up(npair, KPAD);
        double* abar = ws<double>(h, "adj_abar", (size_t)nf * sldp);
        fill_f64(h, abar, 0.0, (size_t)nf * sldp);
        std::vector<double> eps_h(nmo);
        d2h(h, eps_h.data(), mo_energy, nmo);
        HIPCK(h, launch_adjoint_quad(targets, off_dev, pm, wnm, nt, npair, nf, mo_energy, fr_dev,
                                     wt_dev, d->eta, qp_energy, grad_qp, sldp, abar, grad_eps, st));
        // Wbar_k = sum_j abar_jk b_j b_j^T ;  Pibar_k = eps_k^-1 Wbar_k eps_k^-1
        const size_t mat = (size_t)nf * npad * npad;
        double* wbar = ws<double>(h, "adj_wbar", mat);
        double* E = ws<double>(h, "adj_einv", mat);
        double* T = ws<double>(h, "adj_tmp", mat);
        double* pibar = ws<double>(h, "adj_pibar", mat);
        fill_f64(h, wbar, 0.0, mat);
        PiOperands none;
        HIPCK(h, launch_pi_syrk(QX_PREC_FP64, rows, none, abar, npair, naux, npad, nf, wbar,
                                PiOut::PiFull, st));
        HIPCK(h, launch_batched_gemm(true, false, X, X, E, npad, nf, st));        // X^T X
        HIPCK(h, launch_batched_gemm(false, false, E, wbar, T, npad, nf, st));
        HIPCK(h, launch_batched_gemm(false, false, T, E, pibar, npad, nf, st));
        // dL/dB through Pi and dL/de_ia
        const int sldb = round_up(nov, KPAD);
        double* s = ws<double>(h, "adj_s", (size_t)nf * sldb);
        double* ds = ws<double>(h, "adj_ds", (size_t)nf * sldb);
        HIPCK(h, launch_adjoint_weights(mo_energy, nocc, nvir, nf, fr_dev, sldb, s, ds, st));
        const int nparts = ceil_div(naux, 64);
        double* part = ws<double>(h, "adj_part", (size_t)nparts * nov);
        HIPCK(h, launch_scaled_gemm(B_ov, nov, naux, s, ds, sldb, pibar, npad, nf, 2.0, grad_B,
                                    part, st));
        HIPCK(h, launch_adjoint_eps_from_ebar(part, nparts, nocc, nvir, grad_eps, st));
        // dL/db_j through w_j = b_j^T Wc_k b_j (pair rows are rows of B_ov)
        double* gbp = ws<double>(h, "adj_gbp", (size_t)npair * naux);
        fill_f64(h, gbp, 0.0, (size_t)npair * naux);
        HIPCK(h, launch_scaled_gemm(rows, npair, naux, abar, nullptr, sldp, E, npad, nf, 2.0, gbp,
                                    nullptr, st));
        HIPCK(h, launch_adjoint_pair_scatter(gbp, rows, abar, sldp, nf, targets, pt, pm, npair,
                                             nocc, nvir, naux, grad_B, st));
        HIPCK(h, hipStreamSynchronize(st));
        return QX_OK;
    });
}

qx_status qx_g0w0_cd(qx_handle h, const qx_g0w0_desc* d, const double* mo_energy,
                     const double* mo_energy_ref, const double* B_ov, const double* B_pair,
                     const int* pair_target, const int* pair_m, const int* targets,
                     const double* static_terms, const double* guess, double* qp_energy,
                     double* z_factor, double* sigma_re, double* sigma_im, int* state,
                     qx_stats* stats) {
    return guarded(h, [&] {
        validate_desc(h, d);
        const size_t nmo = d->nocc + d->nvir, nov = (size_t)d->nocc * d->nvir;
        const size_t nt = d->ntarget, np = d->npair;
        std::vector<void*> owned;
        auto up = [&](const void* src, size_t bytes) -> void* {
            if (!src) return nullptr;
            void* p = nullptr;
            if (hipMalloc(&p, std::max<size_t>(bytes, 8)) != hipSuccess) {
                set_err(h, "device allocation of %zu bytes failed", bytes);
                throw QxFail{QX_ERR_ALLOC};
            }
            owned.push_back(p);
            HIPCK(h, hipMemcpy(p, src, bytes, hipMemcpyHostToDevice));
            return p;
        };
        auto down = [&](size_t bytes) -> void* {
            void* p = nullptr;
            if (hipMalloc(&p, std::max<size_t>(bytes, 8)) != hipSuccess) {
                set_err(h, "device allocation of %zu bytes failed", bytes);
                throw QxFail{QX_ERR_ALLOC};
            }
            owned.push_back(p);
            return p;
        };
        qx_status rc = QX_OK;
        try {
            auto* e = (double*)up(mo_energy, nmo * 8);
            auto* er = (double*)up(mo_energy_ref, nmo * 8);
            auto* b = (double*)up(B_ov, nov * d->naux * 8);
            auto* bp = (double*)up(B_pair, np * d->naux * 8);
            auto* pt = (int*)up(pair_target, np * 4);
            auto* pm = (int*)up(pair_m, np * 4);
            auto* tg = (int*)up(targets, nt * 4);
            auto* sx = (double*)up(static_terms, nt * 8);
            auto* gs = (double*)up(guess, nt * 8);
            auto* oe = (double*)down(nt * 8);
            auto* oz = (double*)down(nt * 8);
            auto* osr = (double*)down(nt * 8);
            auto* osi = (double*)down(nt * 8);
            auto* ost = (int*)down(nt * 4);
            rc = run_g0w0(h, d, e, er, b, bp, pt, pm, tg, sx, gs, oe, oz, osr, osi, ost, stats);
            if (rc == QX_OK) {
                HIPCK(h, hipMemcpy(qp_energy, oe, nt * 8, hipMemcpyDeviceToHost));
                HIPCK(h, hipMemcpy(z_factor, oz, nt * 8, hipMemcpyDeviceToHost));
                HIPCK(h, hipMemcpy(sigma_re, osr, nt * 8, hipMemcpyDeviceToHost));
                HIPCK(h, hipMemcpy(sigma_im, osi, nt * 8, hipMemcpyDeviceToHost));
                if (state) HIPCK(h, hipMemcpy(state, ost, nt * 4, hipMemcpyDeviceToHost));
            }
        } catch (...) {
            (void)hipStreamSynchronize(h->stream);
            for (void* p : owned) (void)hipFree(p);
            throw;
        }
        (void)hipStreamSynchronize(h->stream);
        for (void* p : owned) (void)hipFree(p);
        return rc;
    });
}

}  // extern "C"

// ---- Batched solves ----------------------------------------------------------------------
// Many independent instances of different sizes share one frequency grid and one set of
// options.  Instances are sorted by padded size (descending) and packed into waves whose
// workspace fits a budget; each wave is a fixed sequence of launches: per-instance operand
// preparation (weights, BF16 splits, pair rows), ONE persistent Pi launch over every tile
// of every instance, one launch per factorisation kernel and step over the instances that
// still have that block column, one Wc_nm launch pair and one Newton launch.  The sequence
// reads its work tables from device memory, so it is captured once as a HIP graph and
// replayed by later calls with the same batch.
namespace {

struct BatchInst {
    int idx = 0;                          // position in the caller's array
    int nocc = 0, nvir = 0, naux = 0, nov = 0, nt = 0, npair = 0;
    int npad = 0, nblk = 0, kpad = 0, rows = 0, out_off = 0;
    const double* mo_energy = nullptr;
    const double* B_ov = nullptr;
    const double* static_terms = nullptr; // caller's array or the zero block
    std::vector<int> targets, pair_off;
    // device memory (wave workspace / tables / persistent block)
    double *s = nullptr, *eps = nullptr, *x = nullptr, *linv = nullptr, *rows_naux = nullptr,
           *bp = nullptr, *wpart = nullptr, *wnm = nullptr;
    float *s32 = nullptr, *f32 = nullptr;
    uint16_t *hi = nullptr, *lo = nullptr;
    __bf16 *bh = nullptr, *bl = nullptr;  // mixed: BF16 hi/lo of the padded pair rows
    int npp = 0;                          // pair rows rounded up to the 128-row tile
    int *pt = nullptr, *pm = nullptr, *fail = nullptr, *targets_dev = nullptr,
        *pair_off_dev = nullptr;
};

struct BatchWave {
    int b = 0, e = 0;                     // [b, e) in size order
    int pi_total = 0, pi_variant = 0, pi_grid = 8, wn_total = 0, red_total = 0, t_total = 0;
    PiInst* pi = nullptr;
    int* pi_off = nullptr;
    PotrfInst* po = nullptr;
    int *cu_off = nullptr, *rp_off = nullptr;
    std::vector<int> nblk, cu_h, rp_h;
    WnmInst* wn = nullptr;
    int *wn_off = nullptr, *red_off = nullptr;
    NewtonInst* nw = nullptr;
    int* t_off = nullptr;
};

struct BatchPlan {
    std::vector<unsigned long long> key;
    std::vector<BatchInst> inst;          // size order
    std::vector<BatchWave> waves;
    int nf = 0, precision = 0, ntot = 0, pi_grid = 0;
    double eta = 0.0;
    NewtonParams np{};
    double *freqs = nullptr, *wts = nullptr;
    int *state = nullptr, *iters = nullptr;
    double *e_out = nullptr, *z_out = nullptr, *sre_out = nullptr, *sim_out = nullptr;
    bool ran = false;                     // the direct (tuning) run has happened
    hipGraphExec_t exec = nullptr;
    ~BatchPlan() {
        if (exec) (void)hipGraphExecDestroy(exec);
    }
};

void delete_batch_plan(BatchPlan* p) { delete p; }

// Bump allocator over a device region; with base == nullptr it only measures.
struct Bump {
    char* base = nullptr;
    size_t off = 0;
    template <typename T>
    T* take(size_t n) {
        off = (off + 255) / 256 * 256;
        T* p = base ? reinterpret_cast<T*>(base + off) : nullptr;
        off += std::max<size_t>(n, 1) * sizeof(T);
        return p;
    }
};

void carve_instance(Bump& b, BatchInst& in, int nf, bool mixed) {
    const size_t mat = (size_t)in.npad * in.npad;
    in.s = b.take<double>((size_t)nf * in.kpad);
    in.s32 = mixed ? b.take<float>((size_t)nf * in.kpad) : nullptr;
    in.hi = mixed ? b.take<uint16_t>((size_t)in.rows * in.kpad) : nullptr;
    in.lo = mixed ? b.take<uint16_t>((size_t)in.rows * in.kpad) : nullptr;
    in.f32 = mixed ? b.take<float>((size_t)in.rows * in.kpad) : nullptr;
    in.eps = b.take<double>((size_t)nf * mat);
    in.x = b.take<double>((size_t)nf * mat);
    in.linv = b.take<double>((size_t)nf * in.nblk * NB * NB);
    in.rows_naux = b.take<double>((size_t)in.npair * in.naux);
    in.bp = b.take<double>((size_t)in.npair * in.npad);
    in.pt = b.take<int>(in.npair);
    in.pm = b.take<int>(in.npair);
    in.wpart = b.take<double>((size_t)nf * in.nblk * in.npair);
    in.wnm = b.take<double>((size_t)nf * in.npair);
    in.npp = round_up(std::max(in.npair, 1), 128);
    in.bh = mixed ? b.take<__bf16>((size_t)in.npp * in.npad) : nullptr;
    in.bl = mixed ? b.take<__bf16>((size_t)in.npp * in.npad) : nullptr;
}

size_t batch_budget(qx_handle h) {
    if (const char* v = std::getenv("QXGW_BATCH_GB")) {
        const double gb = std::atof(v);
        if (gb > 0) return (size_t)(gb * (1ull << 30));
    }
    size_t fr = 0, tot = 0;
    HIPCK(h, hipMemGetInfo(&fr, &tot));
    return tot / 10 * 4;
}

std::vector<unsigned long long> batch_key(const qx_g0w0_desc* d, int ninst,
                                          const qx_batch_instance* inst, const void* const* outs,
                                          int nouts) {
    std::vector<unsigned long long> k;
    auto put = [&](const void* p, size_t n) {
        const unsigned char* c = static_cast<const unsigned char*>(p);
        for (size_t i = 0; i < n; i += 8) {
            unsigned long long v = 0;
            std::memcpy(&v, c + i, std::min<size_t>(8, n - i));
            k.push_back(v);
        }
    };
    qx_g0w0_desc dd = *d;
    dd.nocc = dd.nvir = dd.naux = dd.ntarget = dd.npair = 0;
    put(&dd, sizeof dd);
    k.push_back((unsigned long long)ninst);
    for (int q = 0; q < ninst; ++q) {
        const qx_batch_instance& in = inst[q];
        const long long dims[4] = {in.nocc, in.nvir, in.naux, in.ntarget};
        put(dims, sizeof dims);
        const void* ptrs[3] = {in.mo_energy, in.B_ov, in.static_terms};
        put(ptrs, sizeof ptrs);
        put(in.targets, sizeof(int) * (size_t)in.ntarget);
    }
    put(outs, sizeof(void*) * (size_t)nouts);
    return k;
}

BatchPlan* build_batch_plan(qx_handle h, const qx_g0w0_desc* d, int ninst,
                            const qx_batch_instance* cin) {
    const bool mixed = d->precision == QX_PREC_BF16X3;
    const int nf = d->nfreq;
    auto plan = std::make_unique<BatchPlan>();
    plan->nf = nf;
    plan->precision = d->precision;
    plan->eta = d->eta;
    plan->np = to_params(d->newton, d->qp_solver);
    // instances, sorted by padded size then by N_ov (largest first)
    std::vector<BatchInst>& v = plan->inst;
    v.resize(ninst);
    int off = 0;
    for (int q = 0; q < ninst; ++q) {
        const qx_batch_instance& c = cin[q];
        BatchInst& in = v[q];
        in.idx = q;
        in.nocc = c.nocc;
        in.nvir = c.nvir;
        in.naux = c.naux;
        in.nov = c.nocc * c.nvir;
        in.nt = c.ntarget;
        in.mo_energy = c.mo_energy;
        in.B_ov = c.B_ov;
        in.static_terms = c.static_terms;
        in.targets.assign(c.targets, c.targets + c.ntarget);
        in.pair_off.assign(in.nt + 1, 0);
        for (int t = 0; t < in.nt; ++t)
            in.pair_off[t + 1] = in.pair_off[t] + (in.targets[t] < in.nocc ? in.nvir : in.nocc);
        in.npair = in.pair_off[in.nt];
        in.npad = round_up(in.naux, NB);
        in.nblk = in.npad / NB;
        in.kpad = round_up(in.nov, KPAD);
        in.rows = round_up(in.npad, 256);
        in.out_off = off;
        off += in.nt;
    }
    plan->ntot = off;
    std::stable_sort(v.begin(), v.end(), [](const BatchInst& a, const BatchInst& b) {
        return a.npad != b.npad ? a.npad > b.npad : a.nov > b.nov;
    });
    // waves under the workspace budget
    const size_t budget = batch_budget(h);
    size_t wave_max = 0;
    {
        int b = 0;
        while (b < ninst) {
            Bump m;
            int e = b;
            while (e < ninst) {
                Bump t = m;
                carve_instance(t, v[e], nf, mixed);
                if (e > b && t.off > budget) break;
                m = t;
                ++e;
            }
            BatchWave w;
            w.b = b;
            w.e = e;
            plan->waves.push_back(std::move(w));
            wave_max = std::max(wave_max, m.off);
            b = e;
        }
    }
    // persistent block: grid, fail flags, zero statics, Newton state
    {
        Bump m;
        m.take<double>(nf);
        m.take<double>(nf);
        m.take<int>((size_t)ninst * nf);
        m.take<double>(plan->ntot);
        m.take<int>(plan->ntot);
        m.take<int>(plan->ntot);
        Bump k{static_cast<char*>(static_cast<void*>(ws<char>(h, "batch_keep", m.off)))};
        plan->freqs = k.take<double>(nf);
        plan->wts = k.take<double>(nf);
        int* fail = k.take<int>((size_t)ninst * nf);
        double* zero = k.take<double>(plan->ntot);
        plan->state = k.take<int>(plan->ntot);
        plan->iters = k.take<int>(plan->ntot);
        std::vector<double> fr, wt;
        scaled_roots(nf, d->x0, fr, wt);
        h2d(h, plan->freqs, fr.data(), nf);
        h2d(h, plan->wts, wt.data(), nf);
        fill_f64(h, zero, 0.0, plan->ntot);
        for (int q = 0; q < ninst; ++q) {
            v[q].fail = fail + (size_t)q * nf;
            if (!v[q].static_terms) v[q].static_terms = zero + v[q].out_off;
        }
    }
    // wave workspace (shared by all waves, which run one after another on the stream)
    char* wsb = ws<char>(h, "batch_ws", wave_max);
    for (BatchWave& w : plan->waves) {
        Bump m{wsb};
        for (int q = w.b; q < w.e; ++q) carve_instance(m, v[q], nf, mixed);
    }
    // work tables: measure, then fill a host image and copy it once
    auto layout = [&](Bump& m) {
        for (BatchInst& in : v) {
            in.targets_dev = m.take<int>(in.nt);
            in.pair_off_dev = m.take<int>(in.nt + 1);
        }
        for (BatchWave& w : plan->waves) {
            const int n = w.e - w.b;
            w.pi = m.take<PiInst>(n);
            w.pi_off = m.take<int>(n + 1);
            w.po = m.take<PotrfInst>(n);
            w.cu_off = m.take<int>(n + 1);
            w.rp_off = m.take<int>(n + 1);
            w.wn = m.take<WnmInst>(n);
            w.wn_off = m.take<int>(n + 1);
            w.red_off = m.take<int>(n + 1);
            w.nw = m.take<NewtonInst>(n);
            w.t_off = m.take<int>(n + 1);
        }
    };
    Bump meas;
    layout(meas);
    char* tab = ws<char>(h, "batch_tab", meas.off);
    Bump real{tab};
    layout(real);
    std::vector<char> img(meas.off, 0);
    auto host = [&](const void* dev) { return img.data() + (static_cast<const char*>(dev) - tab); };
    auto put = [&](const void* dev, const void* src, size_t bytes) {
        std::memcpy(host(dev), src, bytes);
    };
    for (BatchInst& in : v) {
        put(in.targets_dev, in.targets.data(), sizeof(int) * in.nt);
        put(in.pair_off_dev, in.pair_off.data(), sizeof(int) * (in.nt + 1));
    }
    for (BatchWave& w : plan->waves) {
        const int n = w.e - w.b;
        std::vector<PiInst> pi(n);
        std::vector<PotrfInst> po(n);
        std::vector<WnmInst> wn(n);
        std::vector<NewtonInst> nw(n);
        std::vector<int> pio(n + 1, 0), cu(n + 1, 0), rp(n + 1, 0), wno(n + 1, 0), red(n + 1, 0),
            to(n + 1, 0);
        w.nblk.resize(n);
        for (int i = 0; i < n; ++i) {
            const BatchInst& in = v[w.b + i];
            const int T = ceil_div(in.npad, 128);
            pi[i] = PiInst{in.hi, in.lo, in.f32, in.s32, in.B_ov, in.s, in.eps, in.nov,
                           in.kpad, in.naux, in.npad};
            pio[i + 1] = pio[i] + T * (T + 1) / 2 * nf;
            po[i] = PotrfInst{in.eps, in.x, in.linv, in.fail, in.npad};
            cu[i + 1] = cu[i] + in.nblk * nf;
            rp[i + 1] = rp[i] + (in.nblk - 1) * nf;
            w.nblk[i] = in.nblk;
            wn[i] = WnmInst{in.bp, in.x, in.wpart, in.wnm, in.npair, in.npad};
            if (mixed) {
                // the BF16 split of M = X - I goes into the eps buffer, free once X is formed
                // (4 of its 8 bytes per element)
                const size_t mat = (size_t)in.npad * in.npad;
                wn[i].bh = in.bh;
                wn[i].bl = in.bl;
                wn[i].mh = reinterpret_cast<__bf16*>(in.eps);
                wn[i].ml = wn[i].mh + (size_t)nf * mat;
                wn[i].npp = in.npp;
                wno[i + 1] = wno[i] + in.npp / 128 * ((nf + 1) / 2) * ceil_div(in.npad, 128);
            } else {
                wno[i + 1] = wno[i] + ceil_div(std::max(in.npair, 1), NB) * in.nblk * nf;
            }
            red[i + 1] = red[i] + nf * in.npair;
            SigmaProblem& p = nw[i].p;
            p.nt = in.nt;
            p.npair = in.npair;
            p.nf = nf;
            p.nocc = in.nocc;
            p.nmo = in.nocc + in.nvir;
            p.targets = in.targets_dev;
            p.pair_off = in.pair_off_dev;
            p.pair_m = in.pm;
            p.wnm = in.wnm;
            p.eps_g = in.mo_energy;
            p.eps_ref = in.mo_energy;
            p.freqs = plan->freqs;
            p.wts = plan->wts;
            p.static_terms = in.static_terms;
            p.eta = d->eta;
            p.ef = 0.0;                      // recomputed on the device
            nw[i].out_off = in.out_off;
            to[i + 1] = to[i] + in.nt;
        }
        put(w.pi, pi.data(), sizeof(PiInst) * n);
        put(w.pi_off, pio.data(), sizeof(int) * (n + 1));
        put(w.po, po.data(), sizeof(PotrfInst) * n);
        put(w.cu_off, cu.data(), sizeof(int) * (n + 1));
        put(w.rp_off, rp.data(), sizeof(int) * (n + 1));
        put(w.wn, wn.data(), sizeof(WnmInst) * n);
        put(w.wn_off, wno.data(), sizeof(int) * (n + 1));
        put(w.red_off, red.data(), sizeof(int) * (n + 1));
        put(w.nw, nw.data(), sizeof(NewtonInst) * n);
        put(w.t_off, to.data(), sizeof(int) * (n + 1));
        w.pi_total = pio[n];
        w.cu_h = cu;
        w.rp_h = rp;
        w.wn_total = wno[n];
        w.red_total = red[n];
        w.t_total = to[n];
    }
    h2d(h, tab, img.data(), img.size());
    int dev = 0, ncu = 0;
    HIPCK(h, hipGetDevice(&dev));
    HIPCK(h, hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    plan->pi_grid = ncu;                     // times the occupancy of the chosen variant
    return plan.release();
}

// Enqueues the whole batch on the handle's stream.  direct: the first run, which tunes the
// Pi variant of each wave and records phase times; otherwise the launches only (capture).
void enqueue_batch(qx_handle h, BatchPlan& P, bool direct) {
    hipStream_t st = h->stream;
    const bool mixed = P.precision == QX_PREC_BF16X3;
    const int nf = P.nf;
    auto phase = [&](int id, auto&& fn) {
        if (direct) {
            Phase ph(h, id);
            fn();
        } else {
            fn();
        }
    };
    for (BatchWave& w : P.waves) {
        const int n = w.e - w.b;
        phase(QX_PHASE_PREP, [&] {
            for (int q = w.b; q < w.e; ++q) {
                BatchInst& in = P.inst[q];
                HIPCK(h, launch_pi_weights(in.mo_energy, in.nocc, in.nvir, nf, P.freqs, in.s,
                                           mixed ? in.s32 : nullptr, st));
                if (mixed)
                    HIPCK(h, launch_pi_prepare_mixed(in.B_ov, in.nov, in.naux, in.rows, in.kpad,
                                                     in.hi, in.lo, in.f32, st));
                hipLaunchKernelGGL(fill_i32_kernel, dim3(1), dim3(256), 0, st, in.fail, 0,
                                   (size_t)nf);
                hipLaunchKernelGGL(ov_pairs_kernel,
                                   dim3(std::min(std::max(in.nocc, in.nvir), 1024), in.nt),
                                   dim3(256), 0, st, in.B_ov, in.nocc, in.nvir, in.naux,
                                   in.targets_dev, in.pair_off_dev, in.nt, in.pt, in.pm,
                                   in.rows_naux, in.naux);
                hipLaunchKernelGGL(pad_rows_kernel, grid1d((size_t)in.npair * in.npad), dim3(256),
                                   0, st, in.rows_naux, in.npair, in.naux, in.bp, in.npad);
                HIPCK(h, hipGetLastError());
            }
        });
        const int prec = mixed ? 2 : 0;
        auto pi_launch = [&](int var) {
            HIPCK(h, launch_pi_group(prec, w.pi, w.pi_off, n, nf, w.pi_total,
                                     P.pi_grid * pi_group_occupancy(prec, var), st, var));
        };
        if (direct && prec == 0) {
            // FP64: the frequency-pair instance (1) for even nf, without a timing run (one
            // FP64 wave of a large batch takes seconds); QXGW_PI_VARIANT=0 forces the other
            const char* fv = std::getenv("QXGW_PI_VARIANT");
            w.pi_variant = nf % 2 ? 0 : (fv && std::atoi(fv) == 0 ? 0 : 1);
            w.pi_grid = P.pi_grid * pi_group_occupancy(prec, w.pi_variant);
        } else if (direct) {
            // the two-frequency instance (4) needs an even frequency count
            w.pi_variant =
                nf % 2 ? choose_variant(h, "pi_group", prec, pi_group_variant_count(prec),
                                        {n, w.pi_total, P.inst[w.b].kpad}, pi_launch,
                                        {0, 1, 2, 3})
                       : choose_variant(h, "pi_group", prec, pi_group_variant_count(prec),
                                        {n, w.pi_total, P.inst[w.b].kpad}, pi_launch);
            w.pi_grid = P.pi_grid * pi_group_occupancy(prec, w.pi_variant);
        }
        phase(QX_PHASE_PI, [&] {
            HIPCK(h, launch_pi_group(prec, w.pi, w.pi_off, n, nf, w.pi_total, w.pi_grid, st,
                                     w.pi_variant));
        });
        phase(QX_PHASE_POTRF, [&] {
            HIPCK(h, launch_potrf_group(w.po, w.nblk.data(), n, nf, w.cu_off, w.rp_off,
                                        w.cu_h.data(), w.rp_h.data(), st));
        });
        phase(QX_PHASE_WNM, [&] {
            HIPCK(h, launch_wnm_group(w.wn, w.wn_off, w.red_off, n, nf, w.wn_total, w.red_total,
                                      st, mixed));
        });
        phase(QX_PHASE_NEWTON, [&] {
            HIPCK(h, launch_newton_group(w.nw, w.t_off, n, w.t_total, P.np, P.e_out, P.z_out,
                                         P.sre_out, P.sim_out, P.state, P.iters, st));
        });
    }
}

}  // namespace

qx_status qx_g0w0_batch_ov_dev(qx_handle h, const qx_g0w0_desc* d, int ninst,
                               const qx_batch_instance* inst, double* qp_energy,
                               double* z_factor, double* sigma_re, double* sigma_im, int* state,
                               qx_stats* stats) {
    return guarded(h, [&] {
        const auto t0 = std::chrono::steady_clock::now();
        REQUIRE(h, d != nullptr && inst != nullptr, QX_ERR_INVALID_INPUT, "NULL argument");
        REQUIRE(h, ninst > 0, QX_ERR_INVALID_INPUT, "ninst must be positive (got %d)", ninst);
        REQUIRE(h, qp_energy && z_factor && sigma_re && sigma_im, QX_ERR_INVALID_INPUT,
                "output pointer is NULL");
        REQUIRE(h, d->precision == QX_PREC_FP64 || d->precision == QX_PREC_BF16X3,
                QX_ERR_INVALID_INPUT, "batched solves support FP64 and BF16X3 (got %d)",
                d->precision);
        long long ntot = 0;
        for (int q = 0; q < ninst; ++q) {
            const qx_batch_instance& c = inst[q];
            qx_g0w0_desc dq = *d;
            dq.nocc = c.nocc;
            dq.nvir = c.nvir;
            dq.naux = c.naux;
            dq.ntarget = c.ntarget;
            dq.npair = 0;
            validate_desc(h, &dq);
            REQUIRE(h, c.mo_energy && c.B_ov && c.targets, QX_ERR_INVALID_INPUT,
                    "instance %d: NULL pointer", q);
            for (int t = 0; t < c.ntarget; ++t)
                REQUIRE(h, c.targets[t] >= 0 && c.targets[t] < c.nocc + c.nvir,
                        QX_ERR_INVALID_INPUT, "instance %d: target %d out of range", q,
                        c.targets[t]);
            ntot += c.ntarget;
        }
        REQUIRE(h, ntot < (1ll << 31), QX_ERR_INVALID_INPUT, "too many targets");
        const void* outs[5] = {qp_energy, z_factor, sigma_re, sigma_im, state};
        std::vector<unsigned long long> key = batch_key(d, ninst, inst, outs, 5);
        if (!h->batch || h->batch->key != key) {
            HIPCK(h, hipStreamSynchronize(h->stream));
            delete_batch_plan(h->batch);
            h->batch = nullptr;
            h->batch = build_batch_plan(h, d, ninst, inst);
            h->batch->key = std::move(key);
            h->batch->e_out = qp_energy;
            h->batch->z_out = z_factor;
            h->batch->sre_out = sigma_re;
            h->batch->sim_out = sigma_im;
        }
        BatchPlan& P = *h->batch;
        const char* gv = std::getenv("QXGW_BATCH_GRAPH");
        const bool use_graph = !(gv && std::atoi(gv) == 0) && h->stream != nullptr;
        int replayed = 0;
        if (!P.ran || !use_graph) {
            enqueue_batch(h, P, true);
            P.ran = true;
        } else {
            if (!P.exec) {
                hipGraph_t g = nullptr;
                HIPCK(h, hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
                try {
                    enqueue_batch(h, P, false);
                } catch (...) {
                    (void)hipStreamEndCapture(h->stream, &g);
                    if (g) (void)hipGraphDestroy(g);
                    throw;
                }
                HIPCK(h, hipStreamEndCapture(h->stream, &g));
                const hipError_t e = hipGraphInstantiate(&P.exec, g, nullptr, nullptr, 0);
                (void)hipGraphDestroy(g);
                HIPCK(h, e);
            }
            HIPCK(h, hipGraphLaunch(P.exec, h->stream));
            replayed = 1;
        }
        // results that need the single-instance path: a failed Cholesky somewhere, or a
        // target whose Newton iterates reached a residue pole or failed
        std::vector<int> fail((size_t)ninst * P.nf), st(P.ntot);
        d2h(h, fail.data(), P.inst[0].fail, fail.size());
        d2h(h, st.data(), P.state, st.size());
        std::vector<int> redo;
        for (int q = 0; q < ninst; ++q) {
            const BatchInst& in = P.inst[q];
            bool bad = false;
            for (int f = 0; f < P.nf; ++f) bad |= fail[(size_t)q * P.nf + f] != 0;
            for (int t = 0; t < in.nt; ++t)
                bad |= st[in.out_off + t] != kNewtonConverged &&
                       st[in.out_off + t] != kNewtonBisection &&
                       st[in.out_off + t] != kNewtonLinearized;
            if (bad) redo.push_back(q);
        }
        if (state) {
            for (int& v : st)
                v = v == kNewtonBisection ? QX_QP_BISECTION
                    : v == kNewtonLinearized ? QX_QP_LINEARIZED : QX_QP_NEWTON;
            h2d(h, state, st.data(), st.size());
        }
        qx_stats acc{};
        for (int q : redo) {
            const BatchInst& in = P.inst[q];
            const qx_batch_instance& c = inst[in.idx];
            qx_g0w0_desc dq = *d;
            dq.nocc = in.nocc;
            dq.nvir = in.nvir;
            dq.naux = in.naux;
            dq.ntarget = in.nt;
            qx_stats sq{};
            const qx_status r = run_g0w0_ov(
                h, &dq, c.mo_energy, nullptr, c.B_ov, in.targets_dev, c.static_terms, nullptr,
                qp_energy + in.out_off, z_factor + in.out_off, sigma_re + in.out_off,
                sigma_im + in.out_off, state ? state + in.out_off : nullptr, &sq);
            if (r != QX_OK) return r;
            acc.chol_fallbacks += sq.chol_fallbacks;
            acc.singular_fallbacks += sq.singular_fallbacks;
            acc.residue_targets += sq.residue_targets;
            acc.residue_jobs += sq.residue_jobs;
            acc.newton_failures += sq.newton_failures;
        }
        HIPCK(h, hipStreamSynchronize(h->stream));
        resolve_profile(h);
        if (stats) {
            *stats = acc;
            stats->pi_variant = P.waves[0].pi_variant;
            stats->waves = (int)P.waves.size();
            stats->graph_replay = replayed;
            stats->resolved_instances = (int)redo.size();
            stats->seconds =
                std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        }
        return QX_OK;
    });
}
