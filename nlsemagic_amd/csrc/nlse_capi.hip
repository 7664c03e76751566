// nlse_capi.hip -- the C ABI (include/nlsemagic_hip.h): solver objects with
// device-resident state, the RK4 chunk loop, hipGraph replay, and the
// NLSE{1,2,3}D_TAKE_STEPS_HIP[_F] drop-in entry points.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "nlse_host.h"

using namespace nlse;

namespace {
thread_local std::string g_err;
}

namespace nlse {
int set_error(int code, const std::string& msg)
{
    g_err = msg;
    return code;
}

std::atomic<long long> g_allocs{0};

hipError_t dev_malloc_raw(void** p, size_t bytes)
{
    ++g_allocs;
    return hipMalloc(p, bytes);
}

namespace {

// host threads that split one memcpy between them
class CopyPool {
public:
    explicit CopyPool(int n) : n_(n)
    {
        for (int i = 0; i < n; ++i) th_.emplace_back([this, i] { run(i); });
    }
    // the calling thread copies the last of n + 1 parts
    void copy(void* dst, const void* src, size_t len)
    {
        std::unique_lock<std::mutex> lk(m_);
        dst_ = (char*)dst;
        src_ = (const char*)src;
        len_ = len;
        pending_ = n_;
        ++gen_;
        go_.notify_all();
        lk.unlock();
        part(n_, (char*)dst, (const char*)src, len);
        lk.lock();
        done_.wait(lk, [&] { return pending_ == 0; });
    }

private:
    void part(int id, char* dst, const char* src, size_t len) const
    {
        const size_t per = ((len + n_) / (n_ + 1) + 63) / 64 * 64;  // whole cache lines per part
        const size_t a = std::min(len, per * id), b = std::min(len, a + per);
        if (b > a) std::memcpy(dst + a, src + a, b - a);
    }
    void run(int id)
    {
        unsigned seen = 0;
        std::unique_lock<std::mutex> lk(m_);
        for (;;) {
            go_.wait(lk, [&] { return gen_ != seen; });
            seen = gen_;
            char* d = dst_;
            const char* sp = src_;
            const size_t len = len_;
            lk.unlock();
            part(id, d, sp, len);
            lk.lock();
            if (--pending_ == 0) done_.notify_one();
        }
    }
    int n_;
    std::vector<std::thread> th_;
    std::mutex m_;
    std::condition_variable go_, done_;
    char* dst_ = nullptr;
    const char* src_ = nullptr;
    size_t len_ = 0;
    int pending_ = 0;
    unsigned gen_ = 0;
};

// pinned chunks of the pageable copies, one set per device, shared by every
// solver on it (one copy at a time); created on the device's first large
// copy and kept to the end of the process, with the pool's threads
#ifndef NLSE_XFER_CHUNK_MB
#define NLSE_XFER_CHUNK_MB 4
#endif
#ifndef NLSE_XFER_SLOTS
#define NLSE_XFER_SLOTS 3
#endif
#ifndef NLSE_XFER_THREADS
#define NLSE_XFER_THREADS 8
#endif
constexpr size_t XFER_CHUNK = (size_t)NLSE_XFER_CHUNK_MB << 20;  // bytes per pinned chunk
constexpr int XFER_SLOTS = NLSE_XFER_SLOTS;
constexpr size_t XFER_DIRECT = 1u << 20;  // smaller copies: one plain hipMemcpyAsync
// The DMAs run on a copy stream of the staging's own, which lives as long
// as the process: every event the host waits on is recorded there, never on
// a caller's stream (which may be destroyed, or captured, by the time a
// later copy reuses the chunk).  The caller's stream and the copy stream
// are ordered by event waits both ways.
struct Staging {
    std::mutex m;
    bool tried = false, ok = false;
    char* pin[XFER_SLOTS] = {};
    hipEvent_t ev[XFER_SLOTS] = {};
    bool used[XFER_SLOTS] = {};
    hipStream_t xs = nullptr;             // copy stream
    hipEvent_t after = nullptr;           // on the caller's stream: its earlier work
    hipEvent_t done = nullptr;            // on xs: the copy
    CopyPool* pool = nullptr;
};

constexpr int XFER_MAX_DEVICES = 64;

// the staging of the device `st` belongs to (its events must be that device's)
Staging* staging(hipStream_t st, int& dev)
{
    static std::mutex mu;
    static Staging* per[XFER_MAX_DEVICES] = {};
    dev = 0;
    if (hipStreamGetDevice(st, &dev) != hipSuccess || dev < 0 || dev >= XFER_MAX_DEVICES) return nullptr;
    std::lock_guard<std::mutex> lk(mu);
    if (!per[dev]) per[dev] = new Staging();
    return per[dev];
}

// under S.m; events are created on `dev`
bool staging_ready(Staging& S, int dev)
{
    if (S.tried) return S.ok;
    S.tried = true;
    int cur = 0;
    if (hipGetDevice(&cur) != hipSuccess || hipSetDevice(dev) != hipSuccess) return false;
    bool ok = hipStreamCreateWithFlags(&S.xs, hipStreamNonBlocking) == hipSuccess &&
              hipEventCreateWithFlags(&S.after, hipEventDisableTiming) == hipSuccess &&
              hipEventCreateWithFlags(&S.done, hipEventDisableTiming) == hipSuccess;
    for (int i = 0; i < XFER_SLOTS && ok; ++i) {
        ok = hipHostMalloc((void**)&S.pin[i], XFER_CHUNK, hipHostMallocDefault) == hipSuccess &&
             hipEventCreateWithFlags(&S.ev[i], hipEventDisableTiming) == hipSuccess;
    }
    if (hipSetDevice(cur) != hipSuccess || !ok) return false;
    const unsigned hc = std::thread::hardware_concurrency();
    // pool threads besides the caller: at most NLSE_XFER_THREADS - 1, half the cores
    S.pool = new CopyPool((int)std::max(1u, std::min((unsigned)NLSE_XFER_THREADS - 1, hc / 2)));
    S.ok = true;
    return true;
}

}  // namespace

int h2d_pageable(void* dst, const void* src, size_t bytes, hipStream_t st)
{
    int dev = 0;
    Staging* Sp = bytes < XFER_DIRECT ? nullptr : staging(st, dev);
    std::unique_lock<std::mutex> lk;
    if (Sp) lk = std::unique_lock<std::mutex>(Sp->m);
    if (!Sp || !staging_ready(*Sp, dev)) {
        HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, st));
        return NLSE_OK;
    }
    Staging& S = *Sp;
    // the copy writes dst after the caller's earlier work on st, and st's
    // later work reads it after the copy
    HIPCHK(hipEventRecord(S.after, st));
    HIPCHK(hipStreamWaitEvent(S.xs, S.after, 0));
    int sl = 0;
    for (size_t off = 0; off < bytes; off += XFER_CHUNK, sl = (sl + 1) % XFER_SLOTS) {
        const size_t len = std::min(XFER_CHUNK, bytes - off);
        if (S.used[sl]) HIPCHK(hipEventSynchronize(S.ev[sl]));  // its previous copy has left the chunk
        S.pool->copy(S.pin[sl], (const char*)src + off, len);
        HIPCHK(hipMemcpyAsync((char*)dst + off, S.pin[sl], len, hipMemcpyHostToDevice, S.xs));
        HIPCHK(hipEventRecord(S.ev[sl], S.xs));
        S.used[sl] = true;
    }
    HIPCHK(hipEventRecord(S.done, S.xs));
    HIPCHK(hipStreamWaitEvent(st, S.done, 0));
    return NLSE_OK;
}

int d2h_pageable(void* dst, const void* src, size_t bytes, hipStream_t st)
{
    int dev = 0;
    Staging* Sp = bytes < XFER_DIRECT ? nullptr : staging(st, dev);
    std::unique_lock<std::mutex> lk;
    if (Sp) lk = std::unique_lock<std::mutex>(Sp->m);
    if (!Sp || !staging_ready(*Sp, dev)) {
        HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        return NLSE_OK;
    }
    Staging& S = *Sp;
    // the copy reads src after the caller's earlier work on st (it has
    // finished, chunk by chunk, when this returns)
    HIPCHK(hipEventRecord(S.after, st));
    HIPCHK(hipStreamWaitEvent(S.xs, S.after, 0));
    const size_t nch = (bytes + XFER_CHUNK - 1) / XFER_CHUNK;
    // the DMA of chunk c into slot c mod XFER_SLOTS
    auto issue = [&](size_t c) -> int {
        const int sl = (int)(c % XFER_SLOTS);
        if (S.used[sl]) HIPCHK(hipEventSynchronize(S.ev[sl]));
        const size_t off = c * XFER_CHUNK, len = std::min(XFER_CHUNK, bytes - off);
        HIPCHK(hipMemcpyAsync(S.pin[sl], (const char*)src + off, len, hipMemcpyDeviceToHost, S.xs));
        HIPCHK(hipEventRecord(S.ev[sl], S.xs));
        S.used[sl] = true;
        return NLSE_OK;
    };
    for (size_t c = 0; c < nch && c < (size_t)XFER_SLOTS; ++c) {
        const int rc = issue(c);
        if (rc != NLSE_OK) return rc;
    }
    for (size_t c = 0; c < nch; ++c) {
        const int sl = (int)(c % XFER_SLOTS);
        HIPCHK(hipEventSynchronize(S.ev[sl]));
        const size_t off = c * XFER_CHUNK, len = std::min(XFER_CHUNK, bytes - off);
        S.pool->copy((char*)dst + off, S.pin[sl], len);
        if (c + XFER_SLOTS < nch) {
            const int rc = issue(c + XFER_SLOTS);
            if (rc != NLSE_OK) return rc;
        }
    }
    return NLSE_OK;
}

void split_hilo(const double* x, float* hi, float* lo, long n)
{
    for (long p = 0; p < n; ++p) {
        hi[p] = (float)x[p];
        lo[p] = (float)(x[p] - (double)hi[p]);
    }
}
}  // namespace nlse

namespace {

int fail(int code, const std::string& msg) { return set_error(code, msg); }

// plain streaming copy (16 B per lane), used to calibrate HBM byte counters
__global__ __launch_bounds__(256) void copy_probe_kernel(const double2* __restrict__ src, double2* __restrict__ dst,
                                                         long n)
{
    for (long p = (long)blockIdx.x * 256 + threadIdx.x; p < n; p += (long)gridDim.x * 256) dst[p] = src[p];
}

}  // namespace

struct nlse_solver {
    int dim, method, prec, BC;
    int N, M, L;
    long n;
    double s, a, h2, k;
    Coefs<double> cd;
    Coefs<float> cf;
    void *U, *Y0, *Y1, *acc, *D, *V;
    void* Ulo;  // precision 3: low part of the compensated state
    hipStream_t stream;
    hipGraphExec_t gexec[2];  // per buffer holding U when the graph starts
    int gsteps;               // steps per graph replay (0: plain launches)
    int gcaptured[2];         // steps held by gexec[i]
    int fuse;                 // RK4 stages per launch requested (2D 1, 2 or 4; 3D 1 or 2)
    int kern;                 // kernel family (KERNEL_TILES, KERNEL_ROWS, KERNEL_STRIPS)
    int tile;                 // 2D rows per tile / strip segment; 3D march rows; 1D steps per launch
    void* ubase;              // the buffer U was allocated in (U swaps with Y1 when fuse > 1)
    hipEvent_t ev0, ev1;
    // reduction workspace (block partials and results of the diagnostics,
    // energy, kfull and error-norm reductions, the vortex generators' radial
    // profile), allocated with the solver: those calls allocate nothing
    double* ws;
    long ws_len;  // doubles
};

namespace {

size_t elem_bytes(int prec) { return prec == 2 ? 16 : 8; }
// workspace of a solver, in doubles: the widest reduction (6 values per
// block, 1024 blocks, plus its result) and a vortex profile (at most twice
// the grid's diagonal in points)
long solver_ws_len(int N, int M, int L) { return 6 * 1024 + 16 + 4L * (N + M + L); }
// precision codes: 2 double, 1 single, 3 single with a compensated (hi + lo) state
bool is_double(const nlse_solver* h) { return h->prec == 2; }

int validate(int dim, int N, int M, int L, int method, int prec, int BC)
{
    if (dim < 1 || dim > 3) return fail(NLSE_EINVAL, "dim must be 1, 2 or 3");
    if (method != 1 && method != 2) return fail(NLSE_EINVAL, "method must be 1 (CD) or 2 (2SHOC)");
    if (prec < 1 || prec > 3)
        return fail(NLSE_EINVAL, "precision must be 1 (single), 2 (double) or 3 (single, compensated)");
    if (N < 3 || (dim >= 2 && M < 3) || (dim == 3 && L < 3))
        return fail(NLSE_EINVAL, "every grid dimension must have at least 3 points");
    if (dim == 1 && BC == BC_ONESIDED && N < 4) return fail(NLSE_EINVAL, "BC=4 needs N >= 4");
    const long long n = (long long)N * (dim >= 2 ? M : 1) * (dim == 3 ? L : 1);
    if (n > (1LL << 31) - 1) return fail(NLSE_EINVAL, "grid larger than 2^31-1 points");
    return NLSE_OK;
}

void drop_graphs(nlse_solver* h)
{
    for (int i = 0; i < 2; ++i) {
        if (h->gexec[i]) (void)hipGraphExecDestroy(h->gexec[i]);
        h->gexec[i] = nullptr;
        h->gcaptured[i] = 0;
    }
}

// RK4 stages per launch actually used: 2D 1, 2 or 4; 3D pairs of stages
// (not MSD, whose partner may lie in another plane, nor planes of 2^31 points
// or more, beyond the pair kernels' 32-bit offsets); 1 otherwise
int fusion(const nlse_solver* h)
{
    if (h->dim == 2) return h->fuse;
    if (h->dim == 3 && h->BC != BC_MSD && h->fuse == 2 && (long)h->N * h->M <= 0x7fffffffL) return 2;
    return 1;
}

void set_params(nlse_solver* h, int BC, double s, double a, double h2, double k)
{
    h->BC = BC;
    h->s = s; h->a = a; h->h2 = h2; h->k = k;
    const float fs = (float)s, fa = (float)a, fh2 = (float)h2, fk = (float)k;
    if (h->dim == 1) { coefs1d<double>(h->cd, s, a, h2, k); coefs1d<float>(h->cf, fs, fa, fh2, fk); }
    else if (h->dim == 2) { coefs2d(h->cd, s, a, h2, k); coefs2d(h->cf, fs, fa, fh2, fk); }
    else { coefs3d<double>(h->cd, s, a, h2, k); coefs3d<float>(h->cf, fs, fa, fh2, fk); }
    drop_graphs(h);
}

// 1D: time-blocked windows of h->tile steps per launch (kind 1, default), or
// (kind 0) the whole chunk in one workgroup for N <= 4096, per-stage
// launches above
bool blocked1d(const nlse_solver* h) { return h->dim == 1 && h->kern == KERNEL_ROWS; }

bool resident1d(const nlse_solver* h)
{
    return h->dim == 1 && !blocked1d(h) && h->N <= chunk1d_resident_max_points((int)elem_bytes(h->prec));
}

template <typename T>
int enqueue_stage_t(nlse_solver* h, const Coefs<T>& c, int stage)
{
    using C = c2<T>;
    C* const U = (C*)h->U;
    C* const Y0 = (C*)h->Y0;
    C* const Y1 = (C*)h->Y1;
    // ping-pong of the stage input: U -> Y0 -> Y1 -> Y0 -> U
    const C* yin[4] = {U, Y0, Y1, Y0};
    C* yout[4] = {Y0, Y1, Y0, U};
    StageArgs<T> a;
    a.u = U;
    a.acc = (C*)h->acc;
    a.V = (const T*)h->V;
    a.dbuf = (C*)h->D;
    a.ulo = (C*)h->Ulo;
    a.c = c;
    a.N = h->N; a.M = h->M; a.L = h->L;
    a.BC = h->BC; a.method = h->method;
    a.kernel = h->kern;
    a.tile = h->tile;
    a.stage = stage;
    a.yin = yin[stage - 1];
    a.yout = yout[stage - 1];
    hipError_t e;
    if (h->dim == 1) e = launch_stage1d<T>(a, h->stream);
    else if (h->dim == 2) e = launch_stage2d<T>(a, h->stream);
    else e = launch_stage3d<T>(a, h->stream);
    if (e != hipSuccess) return fail(NLSE_EHIP, std::string("stage launch: ") + hipGetErrorString(e));
    return NLSE_OK;
}

int enqueue_stage(nlse_solver* h, int stage)
{
    return is_double(h) ? enqueue_stage_t<double>(h, h->cd, stage) : enqueue_stage_t<float>(h, h->cf, stage);
}

// launch l (0 <= l < 4/fusion) of one RK4 step; a fused step leaves the new
// state in Y1 and swaps it with U
template <typename T>
int enqueue_launch_t(nlse_solver* h, const Coefs<T>& c, int l)
{
    using C = c2<T>;
    const int fz = fusion(h);
    if (fz == 1) return enqueue_stage_t<T>(h, c, l + 1);
    StageArgs<T> a;
    a.acc = (C*)h->acc;
    a.V = (const T*)h->V;
    a.dbuf = nullptr;
    a.ulo = (C*)h->Ulo;
    a.c = c;
    a.N = h->N; a.M = h->M; a.L = h->L;
    a.BC = h->BC; a.method = h->method;
    a.kernel = h->kern;
    a.tile = h->tile;
    a.stage = 0;
    a.u = (const C*)h->U;
    hipError_t e;
    bool last;
    if (fz == 2 && l == 0) {
        a.yin = (const C*)h->U;
        a.yout = (C*)h->Y0;  // Y_2 (and ktot) after stages 1-2
        e = (h->dim == 3) ? launch_pair3d<T>(a, true, h->stream) : launch_pair2d<T>(a, true, h->stream);
        last = false;
    } else if (fz == 2) {
        a.yin = (const C*)h->Y0;
        a.yout = (C*)h->Y1;
        e = (h->dim == 3) ? launch_pair3d<T>(a, false, h->stream) : launch_pair2d<T>(a, false, h->stream);
        last = true;
    } else {
        a.yin = (const C*)h->U;
        a.yout = (C*)h->Y1;
        e = launch_step2d<T>(a, h->stream);
        last = true;
    }
    if (e != hipSuccess) return fail(NLSE_EHIP, std::string("fused launch: ") + hipGetErrorString(e));
    if (last) std::swap(h->U, h->Y1);
    return NLSE_OK;
}

int enqueue_launch(nlse_solver* h, int l)
{
    return is_double(h) ? enqueue_launch_t<double>(h, h->cd, l) : enqueue_launch_t<float>(h, h->cf, l);
}

template <typename T>
int enqueue_steps_t(nlse_solver* h, const Coefs<T>& c, int steps)
{
    using C = c2<T>;
    if (steps <= 0) return NLSE_OK;
    if (blocked1d(h)) {
        // an even number of launches (for two or more steps), so U ends in
        // the buffer it started in and graph replays need no pointer fix-up
        const int S = h->tile;
        int nl = (steps + S - 1) / S;
        if (nl > 1 && (nl & 1)) ++nl;
        if (nl == 1 && steps > 1) nl = 2;
        for (int l = 0; l < nl; ++l) {
            const int n = steps / nl + (l < steps % nl ? 1 : 0);
            HIPCHK(launch_block1d<T>((const C*)h->U, (C*)h->Y1, (const C*)h->Ulo, (C*)h->Y0, (const T*)h->V, c, h->N,
                                     h->BC, h->method, n, S, h->stream));
            std::swap(h->U, h->Y1);
            if (h->Ulo) std::swap(h->Ulo, h->Y0);
        }
        return NLSE_OK;
    }
    if (resident1d(h)) {
        HIPCHK(launch_chunk1d_resident<T>((C*)h->U, (C*)h->Ulo, (const T*)h->V, c, h->N, h->BC, h->method, steps,
                                          h->stream));
        return NLSE_OK;
    }
    const int nl = 4 / fusion(h);
    for (int st = 0; st < steps; ++st)
        for (int l = 0; l < nl; ++l) {
            int rc = enqueue_launch_t<T>(h, c, l);
            if (rc != NLSE_OK) return rc;
        }
    return NLSE_OK;
}

int enqueue_steps(nlse_solver* h, int steps)
{
    return is_double(h) ? enqueue_steps_t<double>(h, h->cd, steps) : enqueue_steps_t<float>(h, h->cf, steps);
}

int step_impl(nlse_solver* h, int steps)
{
    if (steps < 0) return fail(NLSE_EINVAL, "steps must be >= 0");
    // a fused step swaps U with Y1: replay graphs of an even number of steps,
    // one per buffer holding U when the replay starts
    int gs = h->gsteps;
    if (fusion(h) > 1 && (gs & 1)) gs += 1;
    if (blocked1d(h) && gs == 1) gs = 2;  // one launch would leave U in the other buffer
    if (gs <= 0 || resident1d(h) || steps < gs) return enqueue_steps(h, steps);
    const int slot = (h->U == h->ubase) ? 0 : 1;
    if (!h->gexec[slot] || h->gcaptured[slot] != gs) {
        if (h->gexec[slot]) (void)hipGraphExecDestroy(h->gexec[slot]);
        h->gexec[slot] = nullptr;
        hipGraph_t g;
        void* u_before = h->U;
        void* y1_before = h->Y1;
        void* lo_before = h->Ulo;
        void* y0_before = h->Y0;
        HIPCHK(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
        int rc = enqueue_steps(h, gs);
        hipError_t e = hipStreamEndCapture(h->stream, &g);
        h->U = u_before;  // capture records launches without running them
        h->Y1 = y1_before;
        h->Ulo = lo_before;
        h->Y0 = y0_before;
        if (rc != NLSE_OK) return rc;
        if (e != hipSuccess) return fail(NLSE_EHIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
        e = hipGraphInstantiate(&h->gexec[slot], g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e != hipSuccess) return fail(NLSE_EHIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
        h->gcaptured[slot] = gs;
    }
    int left = steps;
    while (left >= gs) {
        HIPCHK(hipGraphLaunch(h->gexec[slot], h->stream));
        left -= gs;
    }
    return enqueue_steps(h, left);
}

template <typename T>
int upload(nlse_solver* h, const T* Ur, const T* Ui, const T* V)
{
    if (V) {
        const int rc = h2d_pageable(h->V, V, sizeof(T) * (size_t)h->n, h->stream);
        if (rc != NLSE_OK) return rc;
    }
    int rc = upload_planes<T>(Ur, Ui, (c2<T>*)h->U, (T*)h->Y0, h->n, h->stream);
    if (rc == NLSE_OK && h->Ulo) HIPCHK(hipMemsetAsync(h->Ulo, 0, sizeof(c2<T>) * (size_t)h->n, h->stream));
    if (rc == NLSE_OK) HIPCHK(hipStreamSynchronize(h->stream));
    return rc;
}

template <typename T>
int download(nlse_solver* h, T* Ur, T* Ui)
{
    return download_planes<T>((const c2<T>*)h->U, Ur, Ui, (T*)h->Y0, h->n, h->stream);
}

// double planes -> compensated single-precision state (hi, lo)
int upload_comp(nlse_solver* h, const double* Ur, const double* Ui, const double* V)
{
    const long n = h->n;
    std::vector<float> hr(n), hi(n), lr(n), li(n, 0.f), zero;
    split_hilo(Ur, hr.data(), lr.data(), n);
    if (Ui) split_hilo(Ui, hi.data(), li.data(), n);
    else std::fill(hi.begin(), hi.end(), 0.f);
    if (V) {
        std::vector<float> v(n);
        for (long p = 0; p < n; ++p) v[p] = (float)V[p];
        const int rv = h2d_pageable(h->V, v.data(), sizeof(float) * (size_t)n, h->stream);
        if (rv != NLSE_OK) return rv;
        HIPCHK(hipStreamSynchronize(h->stream));
    }
    int rc = upload_planes<float>(hr.data(), hi.data(), (float2*)h->U, (float*)h->Y0, n, h->stream);
    if (rc == NLSE_OK) rc = upload_planes<float>(lr.data(), li.data(), (float2*)h->Ulo, (float*)h->Y0, n, h->stream);
    return rc;
}

int download_comp(nlse_solver* h, double* Ur, double* Ui)
{
    const long n = h->n;
    std::vector<float> hr(n), hi(n), lr(n), li(n);
    int rc = download_planes<float>((const float2*)h->U, hr.data(), hi.data(), (float*)h->Y0, n, h->stream);
    if (rc == NLSE_OK) rc = download_planes<float>((const float2*)h->Ulo, lr.data(), li.data(), (float*)h->Y0, n, h->stream);
    if (rc != NLSE_OK) return rc;
    for (long p = 0; p < n; ++p) {
        Ur[p] = (double)hr[p] + (double)lr[p];
        Ui[p] = (double)hi[p] + (double)li[p];
    }
    return NLSE_OK;
}

IcGrid ic_grid(const nlse_solver* h, double xmin, double ymin, double zmin, double hg)
{
    IcGrid q;
    q.dim = h->dim;
    q.N = h->N;
    q.M = h->M;
    q.L = h->L;
    q.xmin = xmin;
    q.ymin = ymin;
    q.zmin = zmin;
    q.h = hg;
    q.ws = h->ws;
    q.ws_len = h->ws_len;
    return q;
}

// ---- drop-in entry points: cached solvers of the most recently used shapes

// Up to DROPIN_SLOTS solvers (any shape, method and precision) keep their
// device buffers between drop-in calls, so a caller that alternates between a
// few problems does not reallocate; a new shape replaces the least recently
// used solver.
constexpr int DROPIN_SLOTS = 4;
std::mutex g_cache_mu;
nlse_solver* g_cache[DROPIN_SLOTS] = {};
unsigned long g_cache_use[DROPIN_SLOTS] = {};
unsigned long g_cache_clock = 0;

nlse_solver* cached_solver(int dim, int N, int M, int L, int method, int prec, int BC, double s, double a,
                           double h2, double k, int* rc)
{
    if (dim == 1) M = L = 1;
    if (dim == 2) L = 1;
    *rc = validate(dim, N, M, L, method, prec, BC);  // a bad call evicts nothing
    if (*rc != NLSE_OK) return nullptr;
    int slot = -1;
    for (int i = 0; i < DROPIN_SLOTS && slot < 0; ++i) {
        const nlse_solver* c = g_cache[i];
        if (c && c->dim == dim && c->N == N && c->M == M && c->L == L && c->method == method && c->prec == prec)
            slot = i;
    }
    if (slot >= 0) {
        set_params(g_cache[slot], BC, s, a, h2, k);
    } else {
        slot = 0;  // an empty slot, else the least recently used one
        for (int i = 0; i < DROPIN_SLOTS; ++i) {
            if (!g_cache[i]) {
                slot = i;
                break;
            }
            if (g_cache_use[i] < g_cache_use[slot]) slot = i;
        }
        if (g_cache[slot]) {
            nlse_destroy(g_cache[slot]);
            g_cache[slot] = nullptr;
        }
        *rc = nlse_create(&g_cache[slot], dim, N, M, L, method, prec, BC, s, a, h2, k);
        if (*rc != NLSE_OK) {
            g_cache[slot] = nullptr;
            return nullptr;
        }
    }
    g_cache_use[slot] = ++g_cache_clock;
    return g_cache[slot];
}

template <typename T>
int take_steps(int dim, const T* Ur, const T* Ui, const T* V, T* Unr, T* Uni, double s, double a, double h2, int BC,
               int chunk, double k, int N, int M, int L, int method)
{
    if (!Ur || !V || !Unr || !Uni) return fail(NLSE_EINVAL, "null array argument");
    if (chunk < 0) return fail(NLSE_EINVAL, "chunk_size must be >= 0");
    std::lock_guard<std::mutex> lock(g_cache_mu);
    int rc = NLSE_OK;
    const int prec = sizeof(T) == 8 ? 2 : 1;
    nlse_solver* h = cached_solver(dim, N, M, L, method, prec, BC, s, a, h2, k, &rc);
    if (!h) return rc;
    if ((rc = upload<T>(h, Ur, Ui, V)) != NLSE_OK) return rc;
    if ((rc = step_impl(h, chunk)) != NLSE_OK) return rc;
    return download<T>(h, Unr, Uni);
}

}  // namespace

extern "C" {

const char* nlse_last_error(void) { return g_err.c_str(); }

long long nlse_device_allocations(void) { return g_allocs.load(); }

int nlse_device_count(void)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

long long nlse_footprint_bytes(int dim, int N, int M, int L, int method, int precision)
{
    const long long n = (long long)N * (dim >= 2 ? M : 1) * (dim == 3 ? L : 1);
    const long long e = precision == 2 ? 16 : 8;
    long long b = 4 * n * e + n * (e / 2);  // U, Y0, Y1, ktot, V
    if (dim == 3 && method == 2) b += n * e;
    if (precision == 3) b += n * e;  // low part of the compensated state
    b += 8LL * solver_ws_len(N, M, L);  // reduction workspace
    return b;
}

int nlse_create(nlse_solver** out, int dim, int N, int M, int L, int method, int precision, int BC, double s,
                double a, double h2, double k)
{
    if (!out) return fail(NLSE_EINVAL, "null solver pointer");
    *out = nullptr;
    if (dim == 1) M = L = 1;
    if (dim == 2) L = 1;
    int rc = validate(dim, N, M, L, method, precision, BC);
    if (rc != NLSE_OK) return rc;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(NLSE_EHIP, "no HIP device available");
    nlse_solver* h = new nlse_solver();
    std::memset(h, 0, sizeof(*h));
    h->dim = dim; h->N = N; h->M = M; h->L = L;
    h->method = method; h->prec = precision;
    h->n = (long)N * M * L;
    const size_t eb = elem_bytes(precision) * (size_t)h->n;
    void** bufs[4] = {&h->U, &h->Y0, &h->Y1, &h->acc};
    for (void** b : bufs)
        if (dev_malloc(b, eb) != hipSuccess) { nlse_destroy(h); return fail(NLSE_ENOMEM, "hipMalloc failed"); }
    if (dev_malloc(&h->V, eb / 2) != hipSuccess) { nlse_destroy(h); return fail(NLSE_ENOMEM, "hipMalloc failed"); }
    if (dim == 3 && method == 2 && dev_malloc(&h->D, eb) != hipSuccess) {
        nlse_destroy(h);
        return fail(NLSE_ENOMEM, "hipMalloc failed");
    }
    if (precision == 3 && dev_malloc(&h->Ulo, eb) != hipSuccess) {
        nlse_destroy(h);
        return fail(NLSE_ENOMEM, "hipMalloc failed");
    }
    h->ws_len = solver_ws_len(N, M, L);
    if (dev_malloc((void**)&h->ws, sizeof(double) * (size_t)h->ws_len) != hipSuccess) {
        nlse_destroy(h);
        return fail(NLSE_ENOMEM, "hipMalloc failed");
    }
    if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&h->ev0) != hipSuccess || hipEventCreate(&h->ev1) != hipSuccess) {
        nlse_destroy(h);
        return fail(NLSE_EHIP, "stream/event creation failed");
    }
    h->ubase = h->U;
    // 2D: a whole RK4 step per launch on row-sweep tiles of 16 rows (on a
    // 1024^2 MSD grid 8-55 % faster than pairs of stages on 8-row tiles,
    // across methods and precisions: half the launches, each with a fixed
    // start-up and drain); 3D: pairs of stages where the BC allows
    h->fuse = (dim == 1) ? 1 : (dim == 2 ? 4 : 2);
    h->kern = KERNEL_ROWS;         // the fastest family on a 1024^2 grid and for 3D 2SHOC
    // 3D: plane-march rows; 1D: steps per launch (the fastest on N = 2048)
    h->tile = (dim == 3) ? 4 : (dim == 1 ? (method == 2 ? 16 : 32) : 16);
    set_params(h, BC, s, a, h2, k);
    // a defined initial state: zero field and potential, cleared on the
    // solver's own stream (a null-stream memset is not ordered with it and
    // could land after the caller's first writes)
    hipError_t e = hipMemsetAsync(h->U, 0, eb, h->stream);
    if (e == hipSuccess) e = hipMemsetAsync(h->V, 0, eb / 2, h->stream);
    if (e == hipSuccess && h->Ulo) e = hipMemsetAsync(h->Ulo, 0, eb, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (e != hipSuccess) {
        nlse_destroy(h);
        return fail(NLSE_EHIP, std::string("initial clear: ") + hipGetErrorString(e));
    }
    *out = h;
    return NLSE_OK;
}

void nlse_destroy(nlse_solver* h)
{
    if (!h) return;
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    drop_graphs(h);
    void* bufs[8] = {h->U, h->Y0, h->Y1, h->acc, h->D, h->V, h->Ulo, h->ws};
    for (void* b : bufs)
        if (b) (void)hipFree(b);
    if (h->ev0) (void)hipEventDestroy(h->ev0);
    if (h->ev1) (void)hipEventDestroy(h->ev1);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
}

int nlse_set_state(nlse_solver* h, const double* Ur, const double* Ui, const double* V)
{
    if (!h || !Ur) return fail(NLSE_EINVAL, "null argument");
    if (h->prec == 2) return upload<double>(h, Ur, Ui, V);
    if (h->prec == 3) return upload_comp(h, Ur, Ui, V);
    std::vector<float> r(h->n), i(h->n, 0.0f), v;
    for (long p = 0; p < h->n; ++p) r[p] = (float)Ur[p];
    if (Ui) for (long p = 0; p < h->n; ++p) i[p] = (float)Ui[p];
    if (V) { v.resize(h->n); for (long p = 0; p < h->n; ++p) v[p] = (float)V[p]; }
    return upload<float>(h, r.data(), i.data(), V ? v.data() : nullptr);
}

int nlse_set_state_f(nlse_solver* h, const float* Ur, const float* Ui, const float* V)
{
    if (!h || !Ur) return fail(NLSE_EINVAL, "null argument");
    if (h->prec != 1) return fail(NLSE_EINVAL, "nlse_set_state_f needs a single-precision (1) solver");
    return upload<float>(h, Ur, Ui, V);
}

int nlse_get_state(nlse_solver* h, double* Ur, double* Ui)
{
    if (!h || !Ur || !Ui) return fail(NLSE_EINVAL, "null argument");
    if (h->prec == 2) return download<double>(h, Ur, Ui);
    if (h->prec == 3) return download_comp(h, Ur, Ui);
    std::vector<float> r(h->n), i(h->n);
    int rc = download<float>(h, r.data(), i.data());
    if (rc != NLSE_OK) return rc;
    for (long p = 0; p < h->n; ++p) { Ur[p] = r[p]; Ui[p] = i[p]; }
    return NLSE_OK;
}

int nlse_get_state_f(nlse_solver* h, float* Ur, float* Ui)
{
    if (!h || !Ur || !Ui) return fail(NLSE_EINVAL, "null argument");
    if (h->prec != 1) return fail(NLSE_EINVAL, "nlse_get_state_f needs a single-precision (1) solver");
    return download<float>(h, Ur, Ui);
}

int nlse_step(nlse_solver* h, int steps)
{
    if (!h) return fail(NLSE_EINVAL, "null solver");
    return step_impl(h, steps);
}

int nlse_step_timed(nlse_solver* h, int steps, double* ms)
{
    if (!h || !ms) return fail(NLSE_EINVAL, "null argument");
    HIPCHK(hipEventRecord(h->ev0, h->stream));
    int rc = step_impl(h, steps);
    if (rc != NLSE_OK) return rc;
    HIPCHK(hipEventRecord(h->ev1, h->stream));
    HIPCHK(hipEventSynchronize(h->ev1));
    float f = 0.f;
    HIPCHK(hipEventElapsedTime(&f, h->ev0, h->ev1));
    *ms = f;
    return NLSE_OK;
}

int nlse_profile_launches(nlse_solver* h, int steps, double* ms_launch, int* nlaunch)
{
    if (!h || !ms_launch || !nlaunch || steps <= 0) return fail(NLSE_EINVAL, "bad argument");
    if (resident1d(h)) return fail(NLSE_EINVAL, "small 1D problems run whole chunks in one kernel");
    if (blocked1d(h)) return fail(NLSE_EINVAL, "1D time-blocked windows run several steps per launch");
    const int nl = 4 / fusion(h);
    hipEvent_t ev[5];
    for (int i = 0; i <= nl; ++i) HIPCHK(hipEventCreate(&ev[i]));
    double acc[4] = {0, 0, 0, 0};
    int rc = NLSE_OK;
    for (int st = 0; st < steps && rc == NLSE_OK; ++st) {
        if (hipEventRecord(ev[0], h->stream) != hipSuccess) { rc = fail(NLSE_EHIP, "hipEventRecord"); break; }
        for (int l = 0; l < nl && rc == NLSE_OK; ++l) {
            rc = enqueue_launch(h, l);
            if (rc == NLSE_OK && hipEventRecord(ev[l + 1], h->stream) != hipSuccess)
                rc = fail(NLSE_EHIP, "hipEventRecord");
        }
        if (rc != NLSE_OK) break;
        if (hipEventSynchronize(ev[nl]) != hipSuccess) { rc = fail(NLSE_EHIP, "hipEventSynchronize"); break; }
        for (int l = 0; l < nl; ++l) {
            float f = 0.f;
            (void)hipEventElapsedTime(&f, ev[l], ev[l + 1]);
            acc[l] += f;
        }
    }
    for (int i = 0; i <= nl; ++i) (void)hipEventDestroy(ev[i]);
    if (rc != NLSE_OK) return rc;
    for (int l = 0; l < nl; ++l) ms_launch[l] = acc[l] / steps;
    *nlaunch = nl;
    return NLSE_OK;
}

int nlse_diagnostics(nlse_solver* h, double* out3)
{
    if (!h || !out3) return fail(NLSE_EINVAL, "null argument");
    double* d = h->ws;  // result, then 3 * 1024 block partials
    hipError_t e = launch_diag(h->U, h->Ulo, is_double(h), h->n, d + 3, d, h->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(out3, d, 3 * sizeof(double), hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (e != hipSuccess) return fail(NLSE_EHIP, std::string("diagnostics: ") + hipGetErrorString(e));
    return NLSE_OK;
}

int nlse_energy(nlse_solver* h, double* out4)
{
    if (!h || !out4) return fail(NLSE_EINVAL, "null argument");
    EnGrid g;
    g.n = h->n;
    g.dim = h->dim;
    g.ext[0] = h->dim == 3 ? h->L : h->N;
    g.ext[1] = h->dim == 3 ? h->N : h->M;
    g.ext[2] = h->M;
    g.s = h->s;
    double* d = h->ws;  // result, then 3 * 1024 block partials
    hipError_t e = launch_energy(h->U, h->Ulo, h->V, is_double(h), g, d + 3, d, h->stream);
    double r[3];
    if (e == hipSuccess) e = hipMemcpyAsync(r, d, sizeof(r), hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (e != hipSuccess) return fail(NLSE_EHIP, std::string("energy: ") + hipGetErrorString(e));
    const double hd = std::pow(h->h2, 0.5 * h->dim);  // the cell volume h^d
    out4[1] = hd * (h->a / h->h2) * r[0];
    out4[2] = hd * r[1];
    out4[0] = out4[1] + out4[2];
    out4[3] = hd * r[2];
    return NLSE_OK;
}

int nlse_set_step(nlse_solver* h, double k)
{
    if (!h) return fail(NLSE_EINVAL, "null solver");
    if (!(k > 0) || !std::isfinite(k)) return fail(NLSE_EINVAL, "time step k must be positive and finite");
    HIPCHK(hipStreamSynchronize(h->stream));  // steps already enqueued keep their k
    set_params(h, h->BC, h->s, h->a, h->h2, k);
    return NLSE_OK;
}

int nlse_kfull(nlse_solver* h, double* out3)
{
    if (!h || !out3) return fail(NLSE_EINVAL, "null argument");
    // eigenvalue offsets G of the scheme's stability analysis, per dimension
    // and method (NLSEmagic1D_FRS.m:236-242, NLSE2D.m:430-436, NLSE3D.m:447-453)
    static const double g1cd[] = {4, 3, 1, 0};
    static const double g1sh[] = {64, 63, 46, 12, -3, -4};
    static const double g2cd[] = {8, 7, 6, 2, 1, 0};
    static const double g2sh[] = {128, 127, 126, 110, 109, 92, 24, 9, 8, -6, -7, -8};
    static const double g3cd[] = {12, 11, 10, 9, 3, 2, 1, 0};
    static const double g3sh[] = {192, 191, 190, 189, 174, 173, 172, 156, 155, 138, 36, 21, 20, 6, 5, 4, -9, -10, -11, -12};
    const double* G;
    int ng;
    double gs;  // the 2SHOC tables are given in twelfths
    if (h->method == 1) {
        G = h->dim == 1 ? g1cd : (h->dim == 2 ? g2cd : g3cd);
        ng = h->dim == 1 ? 4 : (h->dim == 2 ? 6 : 8);
        gs = 1.0;
    } else {
        G = h->dim == 1 ? g1sh : (h->dim == 2 ? g2sh : g3sh);
        ng = h->dim == 1 ? 6 : (h->dim == 2 ? 12 : 20);
        gs = 1.0 / 12;
    }
    // MSD needs F on the first interior layer: stage 1 leaves it in ktot
    // (it also overwrites the stage buffer Y0; neither holds data between steps)
    if (h->BC == BC_MSD) {
        const int rc = enqueue_stage(h, 1);
        if (rc != NLSE_OK) return rc;
    }
    KfGrid g;
    g.n = h->n;
    g.dim = h->dim;
    g.ext[0] = h->dim == 3 ? h->L : h->N;
    g.ext[1] = h->dim == 3 ? h->N : h->M;
    g.ext[2] = h->M;
    g.c = h->h2 / h->a;
    g.s = h->s;
    g.BC = h->BC;
    double* d = h->ws;  // result, then 6 * 1024 block partials
    hipError_t e = launch_kfull(h->U, h->Ulo, h->V, h->BC == BC_MSD ? h->acc : nullptr, is_double(h), g, d + 6, d,
                                h->stream);
    double r[6];
    if (e == hipSuccess) e = hipMemcpyAsync(r, d, sizeof(r), hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (e != hipSuccess) return fail(NLSE_EHIP, std::string("kfull: ") + hipGetErrorString(e));
    // Lmax = max over G of max over points |x - G|; Bmax by the BC
    double lmax = 0;
    for (int i = 0; i < ng; ++i) {
        const double gv = G[i] * gs;
        lmax = std::max(lmax, std::max(std::fabs(r[0] - gv), std::fabs(r[1] - gv)));
    }
    double bmax = 0;
    if (h->BC == BC_MSD) bmax = r[3] >= 0 ? r[5] : 0.0;
    else if (h->BC == BC_LAP0) bmax = r[2];
    const double kfull = h->h2 / h->a * std::sqrt(8.0) / std::max(std::fabs(bmax), lmax);
    // the linear bound and the step the drivers take (0.9 of the smaller in
    // 1D and 2D, 0.6 in 3D)
    const double klin = (h->method == 1 ? 1.0 : 0.75) * h->h2 / (h->dim * std::sqrt(2.0) * h->a);
    const double f = h->dim == 3 ? 0.6 : 0.9;
    out3[0] = kfull;
    out3[1] = klin;
    out3[2] = (kfull < klin) ? f * kfull : f * klin;
    return NLSE_OK;
}

int nlse_vorticity_device(nlse_solver* h, double lh2, double veps, double* wx, double* wy, double* wz, double* W)
{
    if (!h) return fail(NLSE_EINVAL, "null solver");
    if (h->dim != 3) return fail(NLSE_EINVAL, "vorticity is defined for 3D solvers");
    if (h->N < 3 || h->M < 3 || h->L < 3) return fail(NLSE_EINVAL, "vorticity needs >= 3 points along every axis");
    hipError_t e = is_double(h)
        ? launch_vorticity3d<double>((const double2*)h->U, nullptr, h->L, h->N, h->M, lh2, veps, wx, wy, wz, W,
                                     h->stream)
        : launch_vorticity3d<float>((const float2*)h->U, (const float2*)h->Ulo, h->L, h->N, h->M, lh2, veps, wx, wy,
                                    wz, W, h->stream);
    if (e != hipSuccess) return fail(NLSE_EHIP, std::string("vorticity: ") + hipGetErrorString(e));
    return NLSE_OK;
}

int nlse_vorticity_timed(nlse_solver* h, double lh2, double veps, double* wx, double* wy, double* wz, double* W,
                         int reps, double* ms)
{
    if (!h || !ms || reps < 1) return fail(NLSE_EINVAL, "null argument or reps < 1");
    HIPCHK(hipEventRecord(h->ev0, h->stream));
    for (int r = 0; r < reps; ++r) {
        const int rc = nlse_vorticity_device(h, lh2, veps, wx, wy, wz, W);
        if (rc != NLSE_OK) return rc;
    }
    HIPCHK(hipEventRecord(h->ev1, h->stream));
    HIPCHK(hipEventSynchronize(h->ev1));
    float f = 0.f;
    HIPCHK(hipEventElapsedTime(&f, h->ev0, h->ev1));
    *ms = (double)f / reps;
    return NLSE_OK;
}

int nlse_vorticity(nlse_solver* h, double lh2, double veps, double* wx, double* wy, double* wz, double* W)
{
    if (!h) return fail(NLSE_EINVAL, "null solver");
    double* host[4] = {wx, wy, wz, W};
    double* dev[4] = {nullptr, nullptr, nullptr, nullptr};
    double* buf = nullptr;
    const size_t b = sizeof(double) * (size_t)h->n;
    int nout = 0;
    for (double* p : host) nout += p != nullptr;
    if (nout && dev_malloc((void**)&buf, b * nout) != hipSuccess) return fail(NLSE_ENOMEM, "hipMalloc failed");
    for (int o = 0, q = 0; o < 4; ++o)
        if (host[o]) dev[o] = buf + (size_t)h->n * q++;
    int rc = nlse_vorticity_device(h, lh2, veps, dev[0], dev[1], dev[2], dev[3]);
    hipError_t e = hipSuccess;
    for (int o = 0; o < 4 && rc == NLSE_OK && e == hipSuccess; ++o)
        if (host[o]) e = hipMemcpyAsync(host[o], dev[o], b, hipMemcpyDeviceToHost, h->stream);
    if (rc == NLSE_OK && e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (buf) (void)hipFree(buf);
    if (rc == NLSE_OK && e != hipSuccess) rc = fail(NLSE_EHIP, std::string("vorticity: ") + hipGetErrorString(e));
    return rc;
}

int nlse_set_exact(nlse_solver* h, int kind, const double* par, double t, double xmin, double ymin, double zmin,
                   double hgrid)
{
    if (!h || !par) return fail(NLSE_EINVAL, "null argument");
    const IcGrid q = ic_grid(h, xmin, ymin, zmin, hgrid);
    return is_double(h) ? ic_set_exact<double>((double2*)h->U, nullptr, (double*)h->V, q, kind, par, t, h->stream)
                        : ic_set_exact<float>((float2*)h->U, (float2*)h->Ulo, (float*)h->V, q, kind, par, t, h->stream);
}

int nlse_fill_state(nlse_solver* h, double re, double im)
{
    if (!h) return fail(NLSE_EINVAL, "null solver");
    return is_double(h) ? ic_fill<double>((double2*)h->U, nullptr, h->n, re, im, h->stream)
                        : ic_fill<float>((float2*)h->U, (float2*)h->Ulo, h->n, re, im, h->stream);
}

int nlse_exact_error(nlse_solver* h, int kind, const double* par, double t, double xmin, double ymin, double zmin,
                     double hgrid, double* out6)
{
    if (!h || !par || !out6) return fail(NLSE_EINVAL, "null argument");
    const IcGrid q = ic_grid(h, xmin, ymin, zmin, hgrid);
    return is_double(h) ? ic_exact_error<double>((const double2*)h->U, nullptr, q, kind, par, t, out6, h->stream)
                        : ic_exact_error<float>((const float2*)h->U, (const float2*)h->Ulo, q, kind, par, t, out6,
                                                h->stream);
}

int nlse_add_vortex_2d(nlse_solver* h, int charge, double om, double x0, double y0, double vx, double vy,
                       double phase, double hgrid, double a, double s, double amp, double xmin, double ymin, int opt)
{
    if (!h) return fail(NLSE_EINVAL, "null solver");
    const IcGrid q = ic_grid(h, xmin, ymin, 0, hgrid);
    return is_double(h) ? ic_vortex_2d<double>((double2*)h->U, nullptr, q, charge, om, x0, y0, vx, vy, phase, a, s,
                                               amp, opt, h->stream)
                        : ic_vortex_2d<float>((float2*)h->U, (float2*)h->Ulo, q, charge, om, x0, y0, vx, vy, phase,
                                              a, s, amp, opt, h->stream);
}

int nlse_add_vortex_line(nlse_solver* h, int m, double OM, double x0, double y0, double z0, double vx, double vy,
                         double vz, double phase, double hgrid, double a, double s, double xmin, double ymin,
                         double zmin)
{
    if (!h) return fail(NLSE_EINVAL, "null solver");
    const IcGrid q = ic_grid(h, xmin, ymin, zmin, hgrid);
    return is_double(h) ? ic_vortex_line<double>((double2*)h->U, nullptr, q, m, OM, x0, y0, z0, vx, vy, vz, phase,
                                                 a, s, h->stream)
                        : ic_vortex_line<float>((float2*)h->U, (float2*)h->Ulo, q, m, OM, x0, y0, z0, vx, vy, vz,
                                                phase, a, s, h->stream);
}

int nlse_add_vortex_ring(nlse_solver* h, int m, double OM, double x0, double y0, double z0, double thetax,
                         double thetay, double thetaz, double vx, double vy, double vz, double phase, double hgrid,
                         double a, double s, double xmin, double ymin, double zmin, double d, int opt_2D, double ea,
                         double eb)
{
    if (!h) return fail(NLSE_EINVAL, "null solver");
    const IcGrid q = ic_grid(h, xmin, ymin, zmin, hgrid);
    return is_double(h) ? ic_vortex_ring<double>((double2*)h->U, nullptr, q, m, OM, x0, y0, z0, thetax, thetay,
                                                 thetaz, vx, vy, vz, phase, a, s, d, opt_2D, ea, eb, h->stream)
                        : ic_vortex_ring<float>((float2*)h->U, (float2*)h->Ulo, q, m, OM, x0, y0, z0, thetax, thetay,
                                                thetaz, vx, vy, vz, phase, a, s, d, opt_2D, ea, eb, h->stream);
}

int nlse_vortex_profile(int m, double OM, double a, double s, double hr, int n, int ansatz, int optimize, double tol,
                        double* f)
{
    if (!f) return fail(NLSE_EINVAL, "null argument");
    if (ansatz < 0 || ansatz > 2) return fail(NLSE_EINVAL, "ansatz must be 0 (dark sqrt), 1 (dark tanh) or 2 (bright sech)");
    if (s == 0) return fail(NLSE_EINVAL, "s must be nonzero");
    return vortex_profile(m, OM, a, s, hr, n, ansatz, optimize, tol, f);
}

int nlse_query(nlse_solver* h)
{
    if (!h) return fail(NLSE_EINVAL, "null solver");
    const hipError_t e = hipStreamQuery(h->stream);
    if (e == hipSuccess) return 1;
    if (e == hipErrorNotReady) return 0;
    return -fail(NLSE_EHIP, std::string("hipStreamQuery: ") + hipGetErrorString(e));
}

int nlse_sync(nlse_solver* h)
{
    if (!h) return fail(NLSE_EINVAL, "null solver");
    HIPCHK(hipStreamSynchronize(h->stream));
    return NLSE_OK;
}

// waits for everything enqueued on `st` so far with the calling thread
// asleep (a blocking-sync event) rather than spinning
static int wait_blocking(hipStream_t st)
{
    hipEvent_t ev;
    HIPCHK(hipEventCreateWithFlags(&ev, hipEventBlockingSync | hipEventDisableTiming));
    hipError_t e = hipEventRecord(ev, st);
    if (e == hipSuccess) e = hipEventSynchronize(ev);
    (void)hipEventDestroy(ev);
    if (e != hipSuccess) return fail(NLSE_EHIP, std::string("blocking wait: ") + hipGetErrorString(e));
    return NLSE_OK;
}

int nlse_wait(nlse_solver* h)
{
    if (!h) return fail(NLSE_EINVAL, "null solver");
    return wait_blocking(h->stream);
}

int nlse_set_fusion(nlse_solver* h, int stages_per_launch)
{
    if (!h) return fail(NLSE_EINVAL, "null solver");
    if (stages_per_launch != 1 && stages_per_launch != 2 && stages_per_launch != 4)
        return fail(NLSE_EINVAL, "stages per launch must be 1, 2 or 4");
    if (h->dim == 1 && stages_per_launch != 1) return fail(NLSE_EINVAL, "stage fusion is implemented for 2D and 3D");
    if (h->dim == 3 && stages_per_launch == 4) return fail(NLSE_EINVAL, "3D fuses at most two stages per launch");
    h->fuse = stages_per_launch;
    drop_graphs(h);
    return NLSE_OK;
}

int nlse_set_kernel(nlse_solver* h, int kind, int rows)
{
    if (!h) return fail(NLSE_EINVAL, "null solver");
    if (h->dim == 1) {
        if (kind == KERNEL_ROWS && (rows < 1 || rows > block1d_max_steps(h->method)))
            return fail(NLSE_EINVAL, "1D time-blocked windows take 1.." + std::to_string(block1d_max_steps(h->method)) +
                                         " steps per launch");
        if (kind != KERNEL_ROWS && kind != KERNEL_TILES)
            return fail(NLSE_EINVAL, "1D kernel kind must be 0 (one workgroup / per-stage) or 1 (time-blocked)");
    } else if (h->dim == 3) {
        if (kind == KERNEL_ROWS && rows != 4 && rows != 8)
            return fail(NLSE_EINVAL, "3D plane-marching tile rows must be 4 or 8");
        if (kind != KERNEL_ROWS && kind != KERNEL_TILES)
            return fail(NLSE_EINVAL, "3D kernel kind must be 0 (per-point passes) or 1 (plane-marching 2SHOC)");
    } else if (kind == KERNEL_TILES) {
        if (rows != 8 && rows != 16 && rows != 32) return fail(NLSE_EINVAL, "LDS tile rows must be 8, 16 or 32");
    } else if (kind == KERNEL_ROWS) {
        if (rows != 8 && rows != 12 && rows != 16) return fail(NLSE_EINVAL, "row-sweep tile rows must be 8, 12 or 16");
    } else if (kind == KERNEL_STRIPS) {
        // 3 rows at least: with the origin shift every segment then owns two
        // rows, so an MSD boundary row and its partner row are in one wave
        if (rows < 3 || rows > 256) return fail(NLSE_EINVAL, "strip segment rows must be 3..256");
    } else {
        return fail(NLSE_EINVAL, "kernel kind must be 0 (LDS tiles), 1 (row-sweep tiles) or 2 (register strips)");
    }
    h->kern = kind;
    h->tile = rows;
    drop_graphs(h);
    return NLSE_OK;
}

int nlse_set_tile(nlse_solver* h, int tile_rows) { return nlse_set_kernel(h, KERNEL_TILES, tile_rows); }

int nlse_get_config(const nlse_solver* h, int* fusion_out, int* kind, int* rows)
{
    if (!h) return fail(NLSE_EINVAL, "null solver");
    if (fusion_out) *fusion_out = fusion(h);
    if (kind) *kind = h->kern;
    if (rows) *rows = h->tile;
    return NLSE_OK;
}

int nlse_autotune(nlse_solver* h, int steps, int* best_fusion, int* best_tile, double* best_ms)
{
    if (!h || steps <= 0) return fail(NLSE_EINVAL, "bad argument");
    if (h->dim == 1 || (h->dim == 3 && h->BC == BC_MSD)) {  // nothing to tune; report what is in use
        if (best_fusion) *best_fusion = fusion(h);
        if (best_tile) *best_tile = 0;
        if (best_ms) *best_ms = 0;
        return NLSE_OK;
    }
    const size_t bytes = elem_bytes(h->prec) * (size_t)h->n;
    void* saved = nullptr;  // U, then the compensated low part when there is one
    const size_t sbytes = h->Ulo ? 2 * bytes : bytes;
    if (dev_malloc(&saved, sbytes) != hipSuccess) return fail(NLSE_ENOMEM, "hipMalloc failed");
    HIPCHK(hipMemcpyAsync(saved, h->U, bytes, hipMemcpyDeviceToDevice, h->stream));
    if (h->Ulo) HIPCHK(hipMemcpyAsync((char*)saved + bytes, h->Ulo, bytes, hipMemcpyDeviceToDevice, h->stream));
    // candidates: 2D every kernel family and tile height for 1, 2 and 4
    // stages per launch; 3D one stage (plane marching) or two per launch
    struct Cand { int fuse, kern, tile; };
    const Cand cand3[] = {{1, KERNEL_ROWS, 4}, {1, KERNEL_ROWS, 8}, {2, KERNEL_ROWS, 4}};
    const Cand cand2[] = {{1, KERNEL_TILES, 8},  {1, KERNEL_TILES, 16}, {1, KERNEL_TILES, 32}, {2, KERNEL_TILES, 8},
                         {2, KERNEL_TILES, 16}, {2, KERNEL_TILES, 32}, {4, KERNEL_TILES, 8},  {4, KERNEL_TILES, 16},
                         {4, KERNEL_TILES, 32}, {1, KERNEL_ROWS, 8},   {1, KERNEL_ROWS, 12},  {1, KERNEL_ROWS, 16},
                         {2, KERNEL_ROWS, 8},   {2, KERNEL_ROWS, 12},  {2, KERNEL_ROWS, 16},  {4, KERNEL_ROWS, 8},
                         {4, KERNEL_ROWS, 12},  {4, KERNEL_ROWS, 16},  {1, KERNEL_STRIPS, 10}, {1, KERNEL_STRIPS, 16},
                         {2, KERNEL_STRIPS, 10}, {2, KERNEL_STRIPS, 16}, {4, KERNEL_STRIPS, 32},
                         {4, KERNEL_STRIPS, 64}};
    const int gsave = h->gsteps;
    h->gsteps = 0;
    int bf = h->fuse, bk = h->kern, bt = h->tile, rc = NLSE_OK;
    double bms = 1e30;
    const Cand* cand = h->dim == 3 ? cand3 : cand2;
    const int ncand = h->dim == 3 ? (int)(sizeof(cand3) / sizeof(cand3[0])) : (int)(sizeof(cand2) / sizeof(cand2[0]));
    for (int ci = 0; ci < ncand; ++ci) {
        const Cand& cd = cand[ci];
        if (rc != NLSE_OK) break;
        h->fuse = cd.fuse;
        h->kern = cd.kern;
        h->tile = cd.tile;
        rc = enqueue_steps(h, 1);  // warm-up (first launch of this kernel)
        double ms = 0;
        if (rc == NLSE_OK) rc = nlse_step_timed(h, steps, &ms);
        if (rc == NLSE_OK && ms < bms) { bms = ms; bf = cd.fuse; bk = cd.kern; bt = cd.tile; }
        if (hipMemcpyAsync(h->U, saved, bytes, hipMemcpyDeviceToDevice, h->stream) != hipSuccess && rc == NLSE_OK)
            rc = fail(NLSE_EHIP, "restore after tuning trial");
        if (h->Ulo && hipMemcpyAsync(h->Ulo, (char*)saved + bytes, bytes, hipMemcpyDeviceToDevice, h->stream) !=
                          hipSuccess && rc == NLSE_OK)
            rc = fail(NLSE_EHIP, "restore after tuning trial");
    }
    (void)hipStreamSynchronize(h->stream);
    (void)hipFree(saved);
    h->gsteps = gsave;
    h->fuse = bf;
    h->kern = bk;
    h->tile = bt;
    drop_graphs(h);
    if (rc != NLSE_OK) return rc;
    if (best_fusion) *best_fusion = bf;
    if (best_tile) *best_tile = bt;
    if (best_ms) *best_ms = bms / steps;
    return NLSE_OK;
}

int nlse_set_graph(nlse_solver* h, int steps_per_graph)
{
    if (!h || steps_per_graph < 0) return fail(NLSE_EINVAL, "bad argument");
    h->gsteps = steps_per_graph;
    return NLSE_OK;
}

int nlse_copy_probe(long long bytes, int reps, double* ms)
{
    if (bytes <= 0 || bytes % 16 != 0 || reps <= 0 || !ms) return fail(NLSE_EINVAL, "bad argument");
    void *a = nullptr, *b = nullptr;
    hipEvent_t e0, e1;
    HIPCHK(dev_malloc(&a, (size_t)bytes));
    if (dev_malloc(&b, (size_t)bytes) != hipSuccess) {
        (void)hipFree(a);
        return fail(NLSE_ENOMEM, "hipMalloc failed");
    }
    (void)hipMemset(a, 0, (size_t)bytes);
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    const long n = bytes / 16;
    (void)hipEventRecord(e0, 0);
    for (int r = 0; r < reps; ++r)
        hipLaunchKernelGGL(copy_probe_kernel, dim3(8192), dim3(256), 0, 0, (const double2*)a, (double2*)b, n);
    (void)hipEventRecord(e1, 0);
    hipError_t e = hipEventSynchronize(e1);
    float f = 0.f;
    (void)hipEventElapsedTime(&f, e0, e1);
    *ms = f / reps;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(a);
    (void)hipFree(b);
    if (e != hipSuccess) return fail(NLSE_EHIP, hipGetErrorString(e));
    return NLSE_OK;
}

int NLSE1D_TAKE_STEPS_HIP(const double* Uoldr, const double* Uoldi, const double* V, double* Unewr, double* Unewi,
                          double s, double a, double h2, int BC, int chunk_size, double k, int N, int method)
{
    return take_steps<double>(1, Uoldr, Uoldi, V, Unewr, Unewi, s, a, h2, BC, chunk_size, k, N, 1, 1, method);
}

int NLSE1D_TAKE_STEPS_HIP_F(const float* Uoldr, const float* Uoldi, const float* V, float* Unewr, float* Unewi,
                            float s, float a, float h2, int BC, int chunk_size, float k, int N, int method)
{
    return take_steps<float>(1, Uoldr, Uoldi, V, Unewr, Unewi, s, a, h2, BC, chunk_size, k, N, 1, 1, method);
}

int NLSE2D_TAKE_STEPS_HIP(const double* Uoldr, const double* Uoldi, const double* V, double* Unewr, double* Unewi,
                          double s, double a, double h2, int BC, int chunk_size, double k, int N, int M, int method)
{
    return take_steps<double>(2, Uoldr, Uoldi, V, Unewr, Unewi, s, a, h2, BC, chunk_size, k, N, M, 1, method);
}

int NLSE2D_TAKE_STEPS_HIP_F(const float* Uoldr, const float* Uoldi, const float* V, float* Unewr, float* Unewi,
                            float s, float a, float h2, int BC, int chunk_size, float k, int N, int M, int method)
{
    return take_steps<float>(2, Uoldr, Uoldi, V, Unewr, Unewi, s, a, h2, BC, chunk_size, k, N, M, 1, method);
}

int NLSE3D_TAKE_STEPS_HIP(const double* Uoldr, const double* Uoldi, const double* V, double* Unewr, double* Unewi,
                          double s, double a, double h2, int BC, int chunk_size, double k, int N, int M, int L,
                          int method)
{
    return take_steps<double>(3, Uoldr, Uoldi, V, Unewr, Unewi, s, a, h2, BC, chunk_size, k, N, M, L, method);
}

int NLSE3D_TAKE_STEPS_HIP_F(const float* Uoldr, const float* Uoldi, const float* V, float* Unewr, float* Unewi,
                            float s, float a, float h2, int BC, int chunk_size, float k, int N, int M, int L,
                            int method)
{
    return take_steps<float>(3, Uoldr, Uoldi, V, Unewr, Unewi, s, a, h2, BC, chunk_size, k, N, M, L, method);
}

int NLSE3D_VORTICITY_HIP(const double* Ur, const double* Ui, double lh2, double veps, int M, int N, int L,
                         double* wx, double* wy, double* wz, double* W)
{
    if (!Ur || !wx || !wy || !wz || !W) return fail(NLSE_EINVAL, "null array argument");
    if (N < 3 || M < 3 || L < 3) return fail(NLSE_EINVAL, "vorticity needs >= 3 points along every axis");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(NLSE_EHIP, "no HIP device available");
    const long n = (long)N * M * L;
    const size_t b = sizeof(double) * (size_t)n;
    // U (2n), wx, wy, wz, W (4n); the outputs double as the upload staging area
    double* buf = nullptr;
    if (dev_malloc((void**)&buf, b * 6) != hipSuccess) return fail(NLSE_ENOMEM, "hipMalloc failed");
    hipStream_t st = nullptr;
    if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) {
        (void)hipFree(buf);
        return fail(NLSE_EHIP, "stream creation failed");
    }
    double2* u = (double2*)buf;
    double* w = buf + 2 * n;
    int rc = upload_planes<double>(Ur, Ui, u, w, n, st);
    hipError_t e = hipSuccess;
    if (rc == NLSE_OK) e = launch_vorticity3d<double>(u, nullptr, L, N, M, lh2, veps, w, w + n, w + 2 * n, w + 3 * n, st);
    double* host[4] = {wx, wy, wz, W};
    for (int o = 0; o < 4 && rc == NLSE_OK && e == hipSuccess; ++o)
        e = hipMemcpyAsync(host[o], w + (size_t)o * n, b, hipMemcpyDeviceToHost, st);
    if (rc == NLSE_OK && e == hipSuccess) e = hipStreamSynchronize(st);
    (void)hipStreamSynchronize(st);
    (void)hipStreamDestroy(st);
    (void)hipFree(buf);
    if (rc == NLSE_OK && e != hipSuccess) rc = fail(NLSE_EHIP, std::string("vorticity: ") + hipGetErrorString(e));
    return rc;
}

}  // extern "C"

// ---- groups of resident 2D solvers advanced by shared launches
//
// The solvers keep their own buffers; a problem table points the batched 2D
// kernels (nlse2d_batch.hip) at them, so one persistent launch per fused pair
// of RK4 stages walks the tiles of every member.  The tile code and each
// member's coefficients are the solver's own: a member's result is bit for
// bit the result of stepping it alone.

struct nlse_group {
    std::vector<nlse_solver*> mem;
    int method, prec;
    void* probs;     // device BatchProb2<T>[n]
    int* tile_prob;  // device: member of each tile
    int ntiles, grid;
    int parity;      // fused steps taken mod 2: buffers 0 and 2 of the table have swapped roles
    hipStream_t stream;
};

namespace {

constexpr int GROUP_ROWS = 16;  // row-sweep tiles of 16 rows, a whole RK4 step per launch

template <typename T>
int group_table_t(nlse_group* g)
{
    const int R = g->method == 2 ? 2 : 1, TW = 64 - 8 * R;  // owned columns of a four-stage tile
    std::vector<BatchProb2<T>> P(g->mem.size());
    std::vector<int> tp;
    int t0 = 0;
    for (size_t p = 0; p < g->mem.size(); ++p) {
        const nlse_solver* h = g->mem[p];
        BatchProb2<T>& q = P[p];
        q.N = h->N;
        q.M = h->M;
        q.BC = h->BC;
        q.ioff = (q.N % GROUP_ROWS == 1) ? 1 : 0;
        q.joff = (q.M % TW == 1) ? 1 : 0;
        q.tiles_i = (q.N + q.ioff + GROUP_ROWS - 1) / GROUP_ROWS;
        q.tiles_j = (q.M + q.joff + TW - 1) / TW;
        q.tile0 = t0;
        q.buf[0] = (c2<T>*)h->U;  // the state when the group forms
        q.buf[1] = (c2<T>*)h->Y0;
        q.buf[2] = (c2<T>*)h->Y1;
        q.acc = (c2<T>*)h->acc;
        q.V = (const T*)h->V;
        q.ulo = (c2<T>*)h->Ulo;
        if constexpr (std::is_same<T, double>::value) q.c = h->cd;
        else q.c = h->cf;
        const int nt = q.tiles_i * q.tiles_j;
        tp.insert(tp.end(), nt, (int)p);
        t0 += nt;
    }
    if (dev_malloc(&g->probs, sizeof(BatchProb2<T>) * P.size()) != hipSuccess ||
        dev_malloc((void**)&g->tile_prob, sizeof(int) * tp.size()) != hipSuccess)
        return fail(NLSE_ENOMEM, "hipMalloc failed");
    HIPCHK(hipMemcpyAsync(g->probs, P.data(), sizeof(BatchProb2<T>) * P.size(), hipMemcpyHostToDevice, g->stream));
    HIPCHK(hipMemcpyAsync(g->tile_prob, tp.data(), sizeof(int) * tp.size(), hipMemcpyHostToDevice, g->stream));
    HIPCHK(hipStreamSynchronize(g->stream));
    g->ntiles = t0;
    return NLSE_OK;
}

template <typename T>
int group_launch_t(nlse_group* g, int first, int last, int r_yin, int r_out)
{
    BatchArgs<T> a;
    a.probs = (const BatchProb2<T>*)g->probs;
    a.tile_prob = g->tile_prob;
    a.ntiles = g->ntiles;
    a.r_yin = r_yin;
    a.r_u = g->parity ? 2 : 0;
    a.r_out = r_out;
    a.comp = g->prec == 3;
    a.method = g->method;
    a.first = first;
    a.last = last;
    a.tile = GROUP_ROWS;
    a.grid = g->grid;
    a.kernel = KERNEL_ROWS;
    const hipError_t e = launch_batch2d<T>(a, g->stream);
    if (e != hipSuccess) return fail(NLSE_EHIP, std::string("group launch: ") + hipGetErrorString(e));
    return NLSE_OK;
}

// one RK4 step of every member, one launch: from the buffer holding U into
// the one that does not, which then holds U
int group_step1(nlse_group* g)
{
    const int cur = g->parity ? 2 : 0, other = 2 - cur;
    const bool d = g->prec == 2;
    const int rc = d ? group_launch_t<double>(g, 1, 4, cur, other) : group_launch_t<float>(g, 1, 4, cur, other);
    if (rc == NLSE_OK) g->parity ^= 1;
    return rc;
}

}  // namespace

extern "C" {

int nlse_group_create(nlse_group** out, nlse_solver* const* members, int n)
{
    if (!out) return fail(NLSE_EINVAL, "null group pointer");
    *out = nullptr;
    if (!members || n < 1) return fail(NLSE_EINVAL, "a group needs at least one member");
    for (int i = 0; i < n; ++i) {
        const nlse_solver* h = members[i];
        if (!h) return fail(NLSE_EINVAL, "null member");
        if (h->dim != 2) return fail(NLSE_EINVAL, "group members must be 2D solvers");
        if (h->method != members[0]->method || h->prec != members[0]->prec)
            return fail(NLSE_EINVAL, "group members must share the method and the precision");
        for (int j = 0; j < i; ++j)
            if (members[j] == h) return fail(NLSE_EINVAL, "a solver appears twice in the group");
    }
    nlse_group* g = new nlse_group();
    g->mem.assign(members, members + n);
    g->method = members[0]->method;
    g->prec = members[0]->prec;
    g->probs = nullptr;
    g->tile_prob = nullptr;
    g->parity = 0;
    g->stream = nullptr;
    int cus = 256;
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, 0);
    g->grid = 8 * ((cus * 4 + 7) / 8);
    // the members' queued work finishes before the group touches their buffers
    for (nlse_solver* h : g->mem)
        if (hipStreamSynchronize(h->stream) != hipSuccess) {
            delete g;
            return fail(NLSE_EHIP, "member stream synchronisation failed");
        }
    if (hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking) != hipSuccess) {
        delete g;
        return fail(NLSE_EHIP, "stream creation failed");
    }
    const int rc = g->prec == 2 ? group_table_t<double>(g) : group_table_t<float>(g);
    if (rc != NLSE_OK) {
        nlse_group_destroy(g);
        return rc;
    }
    *out = g;
    return NLSE_OK;
}

int nlse_group_step(nlse_group* g, int steps)
{
    if (!g) return fail(NLSE_EINVAL, "null group");
    if (steps < 0) return fail(NLSE_EINVAL, "steps must be >= 0");
    for (int i = 0; i < steps; ++i) {
        const int rc = group_step1(g);
        if (rc != NLSE_OK) return rc;
    }
    return NLSE_OK;
}

int nlse_group_done(nlse_group* g)
{
    if (!g) return fail(NLSE_EINVAL, "null group");
    const hipError_t e = hipStreamQuery(g->stream);
    if (e == hipSuccess) return 1;
    if (e == hipErrorNotReady) return 0;
    return -fail(NLSE_EHIP, std::string("hipStreamQuery: ") + hipGetErrorString(e));
}

int nlse_group_sync(nlse_group* g)
{
    if (!g) return fail(NLSE_EINVAL, "null group");
    return wait_blocking(g->stream);
}

void nlse_group_destroy(nlse_group* g)
{
    if (!g) return;
    if (g->stream) (void)hipStreamSynchronize(g->stream);
    // an odd number of steps leaves each member's U in its Y1 buffer: the
    // solvers swap, as their own fused steps do
    if (g->parity)
        for (nlse_solver* h : g->mem) std::swap(h->U, h->Y1);
    if (g->probs) (void)hipFree(g->probs);
    if (g->tile_prob) (void)hipFree(g->tile_prob);
    if (g->stream) (void)hipStreamDestroy(g->stream);
    delete g;
}

}  // extern "C"
