// nlse_batch.hip -- C ABI of the batched 2D integrator (include/nlsemagic_hip.h,
// nlse_batch_*): many independent 2D problems of uneven size, each with its
// own s, a, h^2, k and potential, advanced together by one persistent launch
// per fused group of RK4 stages.  The state of every problem stays resident;
// problem p's fields live at element offset off[p] of shared buffers.
#include <algorithm>
#include <cstring>
#include <vector>

#include "nlse_host.h"

using namespace nlse;

struct nlse_batch {
    int nprob, method, prec, BC;
    std::vector<int> N, M;
    std::vector<long> off;
    std::vector<double> s, a, h2, k;
    long ntot;
    void *U, *Ulo, *Y0, *Y1, *acc, *V;
    void* ubase;     // the allocations U and Y1 started in (they swap after fused steps):
    void* ybase;     // the problem table's buffers 0 and 2
    double* ws;      // diagnostics: 3 results per problem, then 3 * 1024 block partials
    void* probs;     // device BatchProb2<T>[nprob]
    int* tile_prob;  // device: problem of each tile
    int ntiles;
    int fuse, kern, tile, grid;
    hipStream_t stream;
    hipGraphExec_t gexec[2];
    int gsteps, gcaptured[2];
    hipEvent_t ev0, ev1;
};

namespace {

int fail(int code, const std::string& msg) { return set_error(code, msg); }

bool dbl(const nlse_batch* b) { return b->prec == 2; }

void drop_graphs(nlse_batch* b)
{
    for (int i = 0; i < 2; ++i) {
        if (b->gexec[i]) (void)hipGraphExecDestroy(b->gexec[i]);
        b->gexec[i] = nullptr;
        b->gcaptured[i] = 0;
    }
}

// tile width of the current kernel family: 64 columns, or for row-sweep tiles
// and strips 64 minus the halo of the fused stages on both sides
int tile_width(const nlse_batch* b)
{
    if (b->kern == KERNEL_TILES) return 64;
    return 64 - 2 * b->fuse * (b->method == 2 ? 2 : 1);
}

// rows per tile (strips: per segment) accepted by each kernel family
int check_rows(int kind, int rows)
{
    if (kind == KERNEL_TILES) {
        if (rows != 8 && rows != 16 && rows != 32) return fail(NLSE_EINVAL, "LDS tile rows must be 8, 16 or 32");
    } else if (kind == KERNEL_ROWS) {
        if (rows != 8 && rows != 12 && rows != 16) return fail(NLSE_EINVAL, "row-sweep tile rows must be 8, 12 or 16");
    } else if (kind == KERNEL_STRIPS) {
        if (rows < 3 || rows > 256) return fail(NLSE_EINVAL, "strip segment rows must be 3..256");
    } else {
        return fail(NLSE_EINVAL,
                    "batch kernel kind must be 0 (LDS tiles), 1 (row-sweep tiles) or 2 (register strips)");
    }
    return NLSE_OK;
}

// (re)build the tile list and per-problem descriptors for the current kernel
template <typename T>
int build_tiles_t(nlse_batch* b)
{
    const int TI = b->tile, TW = tile_width(b);
    std::vector<BatchProb2<T>> P(b->nprob);
    std::vector<int> tp;
    int t0 = 0;
    for (int p = 0; p < b->nprob; ++p) {
        BatchProb2<T>& q = P[p];
        q.N = b->N[p];
        q.M = b->M[p];
        q.BC = b->BC;
        q.ioff = (q.N % TI == 1) ? 1 : 0;
        q.joff = (q.M % TW == 1) ? 1 : 0;
        q.tiles_i = (q.N + q.ioff + TI - 1) / TI;
        q.tiles_j = (q.M + q.joff + TW - 1) / TW;
        q.tile0 = t0;
        const long o = b->off[p];
        q.buf[0] = (c2<T>*)b->ubase + o;
        q.buf[1] = (c2<T>*)b->Y0 + o;
        q.buf[2] = (c2<T>*)b->ybase + o;
        q.acc = (c2<T>*)b->acc + o;
        q.V = (const T*)b->V + o;
        q.ulo = b->Ulo ? (c2<T>*)b->Ulo + o : nullptr;
        coefs2d(q.c, (T)b->s[p], (T)b->a[p], (T)b->h2[p], (T)b->k[p]);
        const int nt = q.tiles_i * q.tiles_j;
        tp.insert(tp.end(), nt, p);
        t0 += nt;
    }
    // steps already enqueued read the old tables: let them finish first
    if (b->stream) HIPCHK(hipStreamSynchronize(b->stream));
    if (b->probs) (void)hipFree(b->probs);
    if (b->tile_prob) (void)hipFree(b->tile_prob);
    b->probs = nullptr;
    b->tile_prob = nullptr;
    if (dev_malloc(&b->probs, sizeof(BatchProb2<T>) * P.size()) != hipSuccess ||
        dev_malloc((void**)&b->tile_prob, sizeof(int) * tp.size()) != hipSuccess)
        return fail(NLSE_ENOMEM, "hipMalloc failed");
    HIPCHK(hipMemcpy(b->probs, P.data(), sizeof(BatchProb2<T>) * P.size(), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(b->tile_prob, tp.data(), sizeof(int) * tp.size(), hipMemcpyHostToDevice));
    b->ntiles = t0;
    drop_graphs(b);
    return NLSE_OK;
}

int build_tiles(nlse_batch* b) { return dbl(b) ? build_tiles_t<double>(b) : build_tiles_t<float>(b); }

// the role (table buffer) of one of the batch's state buffers
int role(const nlse_batch* b, const void* p) { return p == b->ubase ? 0 : (p == b->Y0 ? 1 : 2); }

template <typename T>
int launch_t(nlse_batch* b, int first, int last, const void* yin, void* out)
{
    BatchArgs<T> a;
    a.r_yin = role(b, yin);
    a.r_u = role(b, b->U);
    a.r_out = role(b, out);
    a.comp = b->Ulo != nullptr;
    a.probs = (const BatchProb2<T>*)b->probs;
    a.tile_prob = b->tile_prob;
    a.ntiles = b->ntiles;
    a.method = b->method;
    a.first = first;
    a.last = last;
    a.tile = b->tile;
    a.grid = b->grid;
    a.kernel = b->kern;
    hipError_t e = launch_batch2d<T>(a, b->stream);
    if (e != hipSuccess) return fail(NLSE_EHIP, std::string("batch launch: ") + hipGetErrorString(e));
    return NLSE_OK;
}

int launch(nlse_batch* b, int first, int last, const void* yin, void* out)
{
    return dbl(b) ? launch_t<double>(b, first, last, yin, out) : launch_t<float>(b, first, last, yin, out);
}

// one RK4 step: 4, 2 or 1 launches; a fused step leaves the new U in Y1 and swaps
int enqueue_step(nlse_batch* b)
{
    int rc;
    if (b->fuse == 1) {
        const void* yin[4] = {b->U, b->Y0, b->Y1, b->Y0};
        void* yout[4] = {b->Y0, b->Y1, b->Y0, b->U};
        for (int s = 0; s < 4; ++s)
            if ((rc = launch(b, s + 1, s + 1, yin[s], yout[s])) != NLSE_OK) return rc;
        return NLSE_OK;
    }
    if (b->fuse == 2) {
        if ((rc = launch(b, 1, 2, b->U, b->Y0)) != NLSE_OK) return rc;
        if ((rc = launch(b, 3, 4, b->Y0, b->Y1)) != NLSE_OK) return rc;
    } else {
        if ((rc = launch(b, 1, 4, b->U, b->Y1)) != NLSE_OK) return rc;
    }
    std::swap(b->U, b->Y1);
    return NLSE_OK;
}

int enqueue_steps(nlse_batch* b, int steps)
{
    for (int i = 0; i < steps; ++i) {
        int rc = enqueue_step(b);
        if (rc != NLSE_OK) return rc;
    }
    return NLSE_OK;
}

int step_impl(nlse_batch* b, int steps)
{
    if (steps < 0) return fail(NLSE_EINVAL, "steps must be >= 0");
    int gs = b->gsteps;
    if (b->fuse > 1 && (gs & 1)) gs += 1;
    if (gs <= 0 || steps < gs) return enqueue_steps(b, steps);
    const int slot = (b->U == b->ubase) ? 0 : 1;
    if (!b->gexec[slot] || b->gcaptured[slot] != gs) {
        if (b->gexec[slot]) (void)hipGraphExecDestroy(b->gexec[slot]);
        b->gexec[slot] = nullptr;
        hipGraph_t g;
        void *u0 = b->U, *y0 = b->Y1;
        HIPCHK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
        int rc = enqueue_steps(b, gs);
        hipError_t e = hipStreamEndCapture(b->stream, &g);
        b->U = u0;
        b->Y1 = y0;
        if (rc != NLSE_OK) return rc;
        if (e != hipSuccess) return fail(NLSE_EHIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
        e = hipGraphInstantiate(&b->gexec[slot], g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e != hipSuccess) return fail(NLSE_EHIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
        b->gcaptured[slot] = gs;
    }
    int left = steps;
    while (left >= gs) {
        HIPCHK(hipGraphLaunch(b->gexec[slot], b->stream));
        left -= gs;
    }
    return enqueue_steps(b, left);
}

}  // namespace

extern "C" {

int nlse_batch_create(nlse_batch** out, int nprob, const int* N, const int* M, const double* s, const double* a,
                      const double* h2, const double* k, int method, int precision, int BC)
{
    if (!out) return fail(NLSE_EINVAL, "null batch pointer");
    *out = nullptr;
    if (nprob < 1 || !N || !M || !s || !a || !h2 || !k) return fail(NLSE_EINVAL, "bad problem list");
    if (method != 1 && method != 2) return fail(NLSE_EINVAL, "method must be 1 (CD) or 2 (2SHOC)");
    if (precision != 2 && precision != 3)
        return fail(NLSE_EINVAL, "batch precision must be 2 (double) or 3 (single, compensated)");
    long ntot = 0;
    for (int p = 0; p < nprob; ++p) {
        if (N[p] < 3 || M[p] < 3) return fail(NLSE_EINVAL, "every grid dimension must have at least 3 points");
        ntot += (long)N[p] * M[p];
    }
    if (ntot > (1L << 31) - 1) return fail(NLSE_EINVAL, "batch larger than 2^31-1 points");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(NLSE_EHIP, "no HIP device available");
    nlse_batch* b = new nlse_batch();
    b->nprob = nprob;
    b->method = method;
    b->prec = precision;
    b->BC = BC;
    b->N.assign(N, N + nprob);
    b->M.assign(M, M + nprob);
    b->s.assign(s, s + nprob);
    b->a.assign(a, a + nprob);
    b->h2.assign(h2, h2 + nprob);
    b->k.assign(k, k + nprob);
    b->off.resize(nprob);
    long o = 0;
    for (int p = 0; p < nprob; ++p) {
        b->off[p] = o;
        o += (long)N[p] * M[p];
    }
    b->ntot = ntot;
    b->fuse = 2;
    b->kern = KERNEL_ROWS;
    b->tile = 8;
    int cus = 256;
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, 0);
    b->grid = 8 * ((cus * 4 + 7) / 8);
    const size_t eb = (precision == 2 ? 16 : 8) * (size_t)ntot;
    void** bufs[5] = {&b->U, &b->Y0, &b->Y1, &b->acc, &b->V};
    for (int i = 0; i < 5; ++i)
        if (dev_malloc(bufs[i], i == 4 ? eb / 2 : eb) != hipSuccess) {
            nlse_batch_destroy(b);
            return fail(NLSE_ENOMEM, "hipMalloc failed");
        }
    if (precision == 3 && dev_malloc(&b->Ulo, eb) != hipSuccess) {
        nlse_batch_destroy(b);
        return fail(NLSE_ENOMEM, "hipMalloc failed");
    }
    b->ubase = b->U;
    b->ybase = b->Y1;
    if (dev_malloc((void**)&b->ws, sizeof(double) * (3 * 1024 + 3 * (size_t)nprob)) != hipSuccess) {
        nlse_batch_destroy(b);
        return fail(NLSE_ENOMEM, "hipMalloc failed");
    }
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&b->ev0) != hipSuccess || hipEventCreate(&b->ev1) != hipSuccess) {
        nlse_batch_destroy(b);
        return fail(NLSE_EHIP, "stream/event creation failed");
    }
    // cleared on the batch's own stream (a null-stream memset is not ordered
    // with it and could land after the caller's first uploads)
    hipError_t e = hipMemsetAsync(b->U, 0, eb, b->stream);
    if (e == hipSuccess) e = hipMemsetAsync(b->V, 0, eb / 2, b->stream);
    if (e == hipSuccess && b->Ulo) e = hipMemsetAsync(b->Ulo, 0, eb, b->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
    if (e != hipSuccess) {
        nlse_batch_destroy(b);
        return fail(NLSE_EHIP, std::string("initial clear: ") + hipGetErrorString(e));
    }
    int rc = build_tiles(b);
    if (rc != NLSE_OK) {
        nlse_batch_destroy(b);
        return rc;
    }
    *out = b;
    return NLSE_OK;
}

void nlse_batch_destroy(nlse_batch* b)
{
    if (!b) return;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    drop_graphs(b);
    void* bufs[9] = {b->U, b->Ulo, b->Y0, b->Y1, b->acc, b->V, b->probs, b->tile_prob, b->ws};
    for (void* p : bufs)
        if (p) (void)hipFree(p);
    if (b->ev0) (void)hipEventDestroy(b->ev0);
    if (b->ev1) (void)hipEventDestroy(b->ev1);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
}

long long nlse_batch_points(const nlse_batch* b) { return b ? b->ntot : 0; }

int nlse_batch_set_state(nlse_batch* b, int p, const double* Ur, const double* Ui, const double* V)
{
    if (!b || !Ur || p < 0 || p >= b->nprob) return fail(NLSE_EINVAL, "bad argument");
    const long n = (long)b->N[p] * b->M[p], o = b->off[p];
    if (b->prec == 2) {
        if (V) HIPCHK(hipMemcpyAsync((double*)b->V + o, V, sizeof(double) * n, hipMemcpyHostToDevice, b->stream));
        return upload_planes<double>(Ur, Ui, (double2*)b->U + o, (double*)b->Y0, n, b->stream);
    }
    std::vector<float> hr(n), hi(n, 0.f), lr(n), li(n, 0.f);
    split_hilo(Ur, hr.data(), lr.data(), n);
    if (Ui) split_hilo(Ui, hi.data(), li.data(), n);
    if (V) {
        std::vector<float> v(n);
        for (long q = 0; q < n; ++q) v[q] = (float)V[q];
        HIPCHK(hipMemcpyAsync((float*)b->V + o, v.data(), sizeof(float) * n, hipMemcpyHostToDevice, b->stream));
        HIPCHK(hipStreamSynchronize(b->stream));
    }
    int rc = upload_planes<float>(hr.data(), hi.data(), (float2*)b->U + o, (float*)b->Y0, n, b->stream);
    if (rc == NLSE_OK) rc = upload_planes<float>(lr.data(), li.data(), (float2*)b->Ulo + o, (float*)b->Y0, n, b->stream);
    return rc;
}

int nlse_batch_get_state(nlse_batch* b, int p, double* Ur, double* Ui)
{
    if (!b || !Ur || !Ui || p < 0 || p >= b->nprob) return fail(NLSE_EINVAL, "bad argument");
    const long n = (long)b->N[p] * b->M[p], o = b->off[p];
    if (b->prec == 2) return download_planes<double>((const double2*)b->U + o, Ur, Ui, (double*)b->Y0, n, b->stream);
    std::vector<float> hr(n), hi(n), lr(n), li(n);
    int rc = download_planes<float>((const float2*)b->U + o, hr.data(), hi.data(), (float*)b->Y0, n, b->stream);
    if (rc == NLSE_OK)
        rc = download_planes<float>((const float2*)b->Ulo + o, lr.data(), li.data(), (float*)b->Y0, n, b->stream);
    if (rc != NLSE_OK) return rc;
    for (long q = 0; q < n; ++q) {
        Ur[q] = (double)hr[q] + (double)lr[q];
        Ui[q] = (double)hi[q] + (double)li[q];
    }
    return NLSE_OK;
}

int nlse_batch_step(nlse_batch* b, int steps)
{
    if (!b) return fail(NLSE_EINVAL, "null batch");
    return step_impl(b, steps);
}

int nlse_batch_step_timed(nlse_batch* b, int steps, double* ms)
{
    if (!b || !ms) return fail(NLSE_EINVAL, "null argument");
    HIPCHK(hipEventRecord(b->ev0, b->stream));
    int rc = step_impl(b, steps);
    if (rc != NLSE_OK) return rc;
    HIPCHK(hipEventRecord(b->ev1, b->stream));
    HIPCHK(hipEventSynchronize(b->ev1));
    float f = 0.f;
    HIPCHK(hipEventElapsedTime(&f, b->ev0, b->ev1));
    *ms = f;
    return NLSE_OK;
}

int nlse_batch_diagnostics(nlse_batch* b, double* out)
{
    if (!b || !out) return fail(NLSE_EINVAL, "null argument");
    double* d = b->ws;
    double* part = d + 3 * (size_t)b->nprob;
    const size_t eb = b->prec == 2 ? 16 : 8;
    hipError_t e = hipSuccess;
    for (int p = 0; p < b->nprob && e == hipSuccess; ++p) {
        const long n = (long)b->N[p] * b->M[p];
        const char* u = (const char*)b->U + eb * b->off[p];
        const char* lo = b->Ulo ? (const char*)b->Ulo + eb * b->off[p] : nullptr;
        e = launch_diag(u, lo, b->prec == 2, n, part, d + 3 * p, b->stream);
    }
    if (e == hipSuccess)
        e = hipMemcpyAsync(out, d, 3 * sizeof(double) * b->nprob, hipMemcpyDeviceToHost, b->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
    if (e != hipSuccess) return fail(NLSE_EHIP, std::string("diagnostics: ") + hipGetErrorString(e));
    return NLSE_OK;
}

int nlse_batch_sync(nlse_batch* b)
{
    if (!b) return fail(NLSE_EINVAL, "null batch");
    HIPCHK(hipStreamSynchronize(b->stream));
    return NLSE_OK;
}

int nlse_batch_set_kernel(nlse_batch* b, int kind, int rows)
{
    if (!b) return fail(NLSE_EINVAL, "null batch");
    const int rc = check_rows(kind, rows);
    if (rc != NLSE_OK) return rc;
    b->kern = kind;
    b->tile = rows;
    return build_tiles(b);
}

int nlse_batch_get_config(const nlse_batch* b, int* fusion, int* kind, int* rows, int* grid)
{
    if (!b) return fail(NLSE_EINVAL, "null batch");
    if (fusion) *fusion = b->fuse;
    if (kind) *kind = b->kern;
    if (rows) *rows = b->tile;
    if (grid) *grid = b->grid;
    return NLSE_OK;
}

int nlse_batch_configure(nlse_batch* b, int stages_per_launch, int tile_rows, int grid, int steps_per_graph)
{
    if (!b) return fail(NLSE_EINVAL, "null batch");
    if (stages_per_launch != 1 && stages_per_launch != 2 && stages_per_launch != 4)
        return fail(NLSE_EINVAL, "stages per launch must be 1, 2 or 4");
    const int rc = check_rows(b->kern, tile_rows);
    if (rc != NLSE_OK) return rc;
    if (grid < 8 || grid % 8 != 0) return fail(NLSE_EINVAL, "grid must be a positive multiple of 8");
    if (steps_per_graph < 0) return fail(NLSE_EINVAL, "steps per graph must be >= 0");
    b->grid = grid;
    b->gsteps = steps_per_graph;
    drop_graphs(b);
    if (tile_rows != b->tile || stages_per_launch != b->fuse) {
        b->fuse = stages_per_launch;
        b->tile = tile_rows;
        return build_tiles(b);  // row-sweep tile and strip widths depend on the stages per launch
    }
    return NLSE_OK;
}

int nlse_batch_autotune(nlse_batch* b, int steps, int* best_fusion, int* best_tile, int* best_grid, double* best_ms)
{
    if (!b || steps <= 0) return fail(NLSE_EINVAL, "bad argument");
    const size_t eb = (b->prec == 2 ? 16 : 8) * (size_t)b->ntot;
    void *saved = nullptr, *saved_lo = nullptr;
    if (dev_malloc(&saved, eb) != hipSuccess) return fail(NLSE_ENOMEM, "hipMalloc failed");
    if (b->Ulo && dev_malloc(&saved_lo, eb) != hipSuccess) {
        (void)hipFree(saved);
        return fail(NLSE_ENOMEM, "hipMalloc failed");
    }
    HIPCHK(hipMemcpyAsync(saved, b->U, eb, hipMemcpyDeviceToDevice, b->stream));
    if (saved_lo) HIPCHK(hipMemcpyAsync(saved_lo, b->Ulo, eb, hipMemcpyDeviceToDevice, b->stream));
    int cus = 256;
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, 0);
    const int fz[3] = {1, 2, 4}, occ[3] = {2, 4, 8};
    // rows per tile (strips: per segment) of each kernel family
    const int tl[3][3] = {{8, 16, 32}, {8, 12, 16}, {16, 32, 64}};
    const int gsave = b->gsteps;
    const int of = b->fuse, ok = b->kern, ot = b->tile, og = b->grid;  // restored if a trial fails
    int bf = of, bk = ok, bt = ot, bg = og, rc = NLSE_OK;
    double bms = 1e30;
    for (int kind = 0; kind < 3 && rc == NLSE_OK; ++kind)
        for (int x = 0; x < 3 && rc == NLSE_OK; ++x)
            for (int y = 0; y < 3 && rc == NLSE_OK; ++y)
                for (int z = 0; z < 3 && rc == NLSE_OK; ++z) {
                    rc = nlse_batch_set_kernel(b, kind, tl[kind][y]);
                    if (rc == NLSE_OK) rc = nlse_batch_configure(b, fz[x], tl[kind][y], 8 * ((cus * occ[z] + 7) / 8), 0);
                    if (rc == NLSE_OK) rc = enqueue_steps(b, 1);
                    double ms = 0;
                    if (rc == NLSE_OK) rc = nlse_batch_step_timed(b, steps, &ms);
                    if (rc == NLSE_OK && ms < bms) {
                        bms = ms;
                        bf = fz[x];
                        bk = kind;
                        bt = tl[kind][y];
                        bg = b->grid;
                    }
                    if (hipMemcpyAsync(b->U, saved, eb, hipMemcpyDeviceToDevice, b->stream) != hipSuccess ||
                        (saved_lo &&
                         hipMemcpyAsync(b->Ulo, saved_lo, eb, hipMemcpyDeviceToDevice, b->stream) != hipSuccess))
                        if (rc == NLSE_OK) rc = fail(NLSE_EHIP, "restore after tuning trial");
                }
    (void)hipStreamSynchronize(b->stream);
    (void)hipFree(saved);
    if (saved_lo) (void)hipFree(saved_lo);
    if (rc != NLSE_OK) {  // the batch keeps the configuration it had; the trial's error is reported
        if (nlse_batch_set_kernel(b, ok, ot) == NLSE_OK) (void)nlse_batch_configure(b, of, ot, og, gsave);
        return rc;
    }
    rc = nlse_batch_set_kernel(b, bk, bt);
    if (rc == NLSE_OK) rc = nlse_batch_configure(b, bf, bt, bg, gsave);
    if (best_fusion) *best_fusion = bf;
    if (best_tile) *best_tile = bt;
    if (best_grid) *best_grid = bg;
    if (best_ms) *best_ms = bms / steps;
    return rc;
}

}  // extern "C"
