// Host side of the drop-in C ABI (nln/gravcuda.h) on HIP.
//
// State mirrors the reference module (nln/pgrav.cu:803-876): a device copy of
// the sink positions (devpos), a device accumulator (devacc), and a queue of
// NQ = 32 slots, each with its own stream and pinned staging buffer so that a
// host walk can keep up to 32 evaluations in flight.  Differences, all on the
// MI355X side of the boundary: slot streams are created once and reused (the
// reference creates and destroys one per call), staging buffers grow on demand
// instead of failing above 1 MB, and sources are converted on the device into
// prepared records (pgrav_device.h) before the evaluation kernel runs.
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/pgrav.h"
#include "pgrav_device.h"
#include "pgrav_internal.h"

namespace pgrav {

static int g_error_mode = 0;
static thread_local std::string g_error;
static thread_local bool g_failed = false;
static int g_qsign = 0;  // 0 = not read yet

void fail(const char *fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    if (g_error_mode == 0) {
        fprintf(stderr, "libpgrav error: %s\n", buf);
        fflush(stderr);
        abort();
    }
    if (!g_failed) {  // keep the first error
        g_error = buf;
        g_failed = true;
    }
}
bool failed() { return g_failed; }
void clear_error() {
    g_failed = false;
    g_error.clear();
}
bool check(hipError_t e, const char *what) {
    if (e == hipSuccess)
        return true;
    fail("%s failed: %d %s", what, (int)e, hipGetErrorString(e));
    return false;
}

int quad_phi_sign() {
    if (g_qsign == 0) {
        const char *s = getenv("PGRAV_QUAD_PHI_SIGN");
        g_qsign = (s && atoi(s) < 0) ? -1 : 1;
    }
    return g_qsign;
}

// ---------------------------------------------------------------- buffers
struct DevBuf {
    void *p = nullptr;
    size_t cap = 0;
    bool ensure(size_t bytes) {
        if (bytes <= cap)
            return true;
        if (p && !check(hipFree(p), "hipFree"))
            return false;
        p = nullptr;
        cap = 0;
        size_t want = bytes < 4096 ? 4096 : bytes + bytes / 4;
        if (!check(hipMalloc(&p, want), "hipMalloc"))
            return false;
        cap = want;
        return true;
    }
    void release() {
        if (p)
            (void)hipFree(p);
        p = nullptr;
        cap = 0;
    }
};

struct HostBuf {
    void *p = nullptr;
    size_t cap = 0;
    bool ensure(size_t bytes) {
        if (bytes <= cap)
            return true;
        if (p && !check(hipHostFree(p), "hipHostFree"))
            return false;
        p = nullptr;
        cap = 0;
        size_t want = bytes < 4096 ? 4096 : bytes + bytes / 4;
        if (!check(hipHostMalloc(&p, want, hipHostMallocDefault), "hipHostMalloc"))
            return false;
        cap = want;
        return true;
    }
    void release() {
        if (p)
            (void)hipHostFree(p);
        p = nullptr;
        cap = 0;
    }
};

static const int NQ = 32;  // nln/pgrav.cu:811

struct Slot {
    bool inuse = false;
    hipStream_t stream = nullptr;
    hipEvent_t staged = nullptr;  // recorded after the last copy out of `host`
    bool staged_pending = false;
    HostBuf host;
    DevBuf dev;   // raw copy of the caller's list
    DevBuf pos4;  // prepared source positions
    DevBuf mom;   // prepared source records
};

// The per-sink-group entry points (grav_mn_CUDA, pinteractCUDA, grav_mns_CUDA,
// grav_qns_CUDA) append their lists to a pinned batch instead of launching: the
// reference's one copy and one launch per call cost ~20 us of host API time each,
// more than most of these calls' kernels take.  A batch is uploaded in one copy
// and evaluated by one prep and one evaluation kernel per kind once its lists
// reach BATCH_BYTES, and before anything reads the results or replaces a table
// the calls use (WalkTerminateSinkCUDA, WalkInitSinkCUDA, WalkInitSrcCUDA,
// WalkTerminateSrcCUDA, pgrav_synchronize, pgrav_set_quad_phi_sign,
// CUDA_Finalize).  Two batches alternate, so the host fills one while the
// other's copy runs.  The queue entry points (grav_mnss_CUDA, grav_qnss_CUDA)
// carry thousands of sinks per call and keep their own launches.
enum { BK_M, BK_MK1, BK_Q, BK_H, BK_MNS, BK_QNS, BK_KINDS };
static const size_t BATCH_BYTES = (size_t)16 << 20;
static const int CHUNK[BK_KINDS] = {512, 512, 256, 128, 256, 256};  // sources per task

struct Batch {
    HostBuf host;
    DevBuf dev;
    hipEvent_t copied = nullptr;  // recorded after the copy out of `host`
    bool pending = false;
    size_t used = 0;         // bytes of list data at the front of `host`
    size_t table_bytes = 0;  // bytes the call, task and block tables will add
    int ncalls = 0;
    std::vector<DeferredCall> calls[BK_KINDS];
    std::vector<DeferredTask> tasks[BK_KINDS];
    std::vector<int> blocks[BK_KINDS];  // M x N kinds: the call of each prep block
    int prep[BK_KINDS] = {};            // M x N kinds: prepared records so far
};

struct State {
    bool init = false;
    int device = 0;
    Slot slot[NQ];
    float *btab = nullptr;
    int64_t nobj = 0;
    int seg_mismatch = 0;  // pMM_sK1 calls whose segments did not sum to source_n
    int stride = 0;
    DevBuf devpos, devacc, errflag;
    HostBuf staging;
    // WalkInitSrcCUDA table (prepared order-2 records)
    DevBuf qraw, qpos4, qmom;
    int64_t qcells = 0;
    // deferred calls: two batches filled in turn, their stream, the prepared
    // records of the batch being evaluated
    bool defer = true;
    Batch batch[2];
    int bcur = 0;
    hipStream_t bstream = nullptr;
    DevBuf bpos4, bmom;
};

static State S;
static std::mutex g_mutex;

static bool ensure_init() {
    if (S.init)
        return true;
    int n = 0;
    if (!check(hipGetDeviceCount(&n), "hipGetDeviceCount"))
        return false;
    if (n < 1) {
        fail("no HIP device visible");
        return false;
    }
    S.device = 0;  // nln/cuda.cu:20
    if (!check(hipSetDevice(S.device), "hipSetDevice"))
        return false;
    for (int i = 0; i < NQ; i++)
        if (!check(hipStreamCreateWithFlags(&S.slot[i].stream, hipStreamNonBlocking),
                   "hipStreamCreate") ||
            !check(hipEventCreateWithFlags(&S.slot[i].staged, hipEventDisableTiming),
                   "hipEventCreate"))
            return false;
    if (!check(hipStreamCreateWithFlags(&S.bstream, hipStreamNonBlocking), "hipStreamCreate"))
        return false;
    for (Batch &b : S.batch)
        if (!check(hipEventCreateWithFlags(&b.copied, hipEventDisableTiming), "hipEventCreate"))
            return false;
    if (!S.errflag.ensure(sizeof(int)) || !check(hipMemset(S.errflag.p, 0, sizeof(int)), "hipMemset"))
        return false;
    S.init = true;
    return true;
}

static bool slot_ok(int q) {
    if (!ensure_init())
        return false;
    if (q < 0 || q >= NQ) {
        fail("bad queue slot %d", q);
        return false;
    }
    // work already queued on the slot stays queued: the stream orders it before
    // this call's, and stage() guards the host-side buffers
    S.slot[q].inuse = true;
    return true;
}

// Before a call rewrites a slot's pinned staging buffer or grows its buffers.
// A slot can carry several calls before qallocCUDA sees it idle (the cell and
// body lists of one sink group): the pinned buffer is rewritten once the copy
// out of it has run (an event after that copy), and buffers grow only once the
// slot's stream has drained.
static bool stage(Slot &s, size_t host_bytes, size_t dev_bytes, size_t pos4_bytes = 0,
                  size_t mom_bytes = 0) {
    const bool grow = host_bytes > s.host.cap || dev_bytes > s.dev.cap ||
                      pos4_bytes > s.pos4.cap || mom_bytes > s.mom.cap;
    if (grow) {
        if (!check(hipStreamSynchronize(s.stream), "hipStreamSynchronize"))
            return false;
        s.staged_pending = false;
    } else if (s.staged_pending) {
        if (!check(hipEventSynchronize(s.staged), "hipEventSynchronize"))
            return false;
        s.staged_pending = false;
    }
    return s.host.ensure(host_bytes) && s.dev.ensure(dev_bytes) &&
           (!pos4_bytes || s.pos4.ensure(pos4_bytes)) && (!mom_bytes || s.mom.ensure(mom_bytes));
}

// queue the copy of the staged bytes to the slot's device buffer
static bool upload(Slot &s, size_t bytes) {
    if (!check(hipMemcpyAsync(s.dev.p, s.host.p, bytes, hipMemcpyHostToDevice, s.stream),
               "hipMemcpyAsync") ||
        !check(hipEventRecord(s.staged, s.stream), "hipEventRecord"))
        return false;
    s.staged_pending = true;
    return true;
}

static bool sinks_ready(int64_t first, int64_t m) {
    if (!S.devpos.p) {
        fail("WalkInitSinkCUDA has not been called");
        return false;
    }
    if (first < 0 || m < 0 || first + m > S.nobj) {
        fail("sink range [%lld, %lld) outside the body table of %lld", (long long)first,
             (long long)(first + m), (long long)S.nobj);
        return false;
    }
    return true;
}

static bool flush_batch();

// Runs the deferred calls and waits for every queued evaluation.
// A batch that fails to run has its error recorded; every stream is still
// drained, so that no kernel reads a buffer the caller goes on to release.
static void wait_all() {
    if (S.init) {
        flush_batch();
        if (!check(hipStreamSynchronize(S.bstream), "hipStreamSynchronize"))
            return;
    }
    for (int i = 0; i < NQ; i++) {
        Slot &s = S.slot[i];
        if (s.inuse) {
            if (!check(hipStreamSynchronize(s.stream), "hipStreamSynchronize"))
                return;
            s.inuse = false;
        }
    }
}

static void release_slots() {
    for (int i = 0; i < NQ; i++) {
        S.slot[i].host.release();
        S.slot[i].dev.release();
        S.slot[i].pos4.release();
        S.slot[i].mom.release();
    }
}

static size_t align128(size_t x) { return (x + 127) & ~(size_t)127; }

// Uploads the current batch and queues its kernels on the batch stream, then
// switches to the other batch.
static bool flush_batch() {
    Batch &b = S.batch[S.bcur];
    if (!b.ncalls)
        return true;
    // the tables go behind the lists, into the same copy
    char *h = (char *)b.host.p;
    size_t off = align128(b.used);
    size_t end = off;
    for (int k = 0; k < BK_KINDS; k++)
        end += align128(b.calls[k].size() * sizeof(DeferredCall)) +
               align128(b.tasks[k].size() * sizeof(DeferredTask)) +
               align128(b.blocks[k].size() * sizeof(int));
    if (end > b.host.cap) {  // batch_room reserves this room; checked before any write
        fail("deferred batch overran its staging buffer (%zu > %zu bytes)", end, b.host.cap);
        return false;
    }
    size_t o_calls[BK_KINDS], o_tasks[BK_KINDS], o_blocks[BK_KINDS];
    auto put = [&](const void *src, size_t bytes) {
        const size_t at = off;
        if (bytes)
            memcpy(h + at, src, bytes);
        off = align128(off + bytes);
        return at;
    };
    for (int k = 0; k < BK_KINDS; k++) {
        o_calls[k] = put(b.calls[k].data(), b.calls[k].size() * sizeof(DeferredCall));
        o_tasks[k] = put(b.tasks[k].data(), b.tasks[k].size() * sizeof(DeferredTask));
        o_blocks[k] = put(b.blocks[k].data(), b.blocks[k].size() * sizeof(int));
    }
    // a buffer still read by queued kernels is replaced only after they ran
    const size_t npos = (size_t)b.prep[BK_M] + b.prep[BK_MK1] + b.prep[BK_Q] + b.prep[BK_H];
    const size_t nmom = (size_t)b.prep[BK_Q] * PREC2 + (size_t)b.prep[BK_H] * PREC4;
    if (off > b.dev.cap || 16 * npos > S.bpos4.cap || 4 * nmom > S.bmom.cap) {
        if (!check(hipStreamSynchronize(S.bstream), "hipStreamSynchronize"))
            return false;
        if (!b.dev.ensure(off) || !S.bpos4.ensure(16 * npos + 16) ||
            !S.bmom.ensure(4 * nmom + 16))
            return false;
    }
    if (!check(hipMemcpyAsync(b.dev.p, b.host.p, off, hipMemcpyHostToDevice, S.bstream),
               "hipMemcpyAsync") ||
        !check(hipEventRecord(b.copied, S.bstream), "hipEventRecord"))
        return false;
    b.pending = true;
    const char *d = (const char *)b.dev.p;
    const float *pos = (const float *)S.devpos.p;
    float *acc = (float *)S.devacc.p;
    const int qs = quad_phi_sign();
    size_t p4 = 0, pm = 0;
    static const int order_of[4] = {0, 0, 2, 4};
    bool ok = true;
    for (int k = BK_M; k <= BK_H && ok; k++) {
        ok = check(launch_mn_batch(order_of[k], k == BK_MK1 ? SOFT_K1 : SOFT_NONE, qs,
                                   (const float *)d, (const DeferredCall *)(d + o_calls[k]),
                                   (const int *)(d + o_blocks[k]), (int)b.blocks[k].size(),
                                   (const DeferredTask *)(d + o_tasks[k]), (int)b.tasks[k].size(),
                                   (float4 *)S.bpos4.p + p4, (float *)S.bmom.p + pm, pos, acc,
                                   S.bstream),
                   "deferred M x N kernels");
        p4 += b.prep[k];
        pm += (size_t)b.prep[k] * (k == BK_H ? PREC4 : (k == BK_Q ? PREC2 : 0));
    }
    ok = ok &&
         check(launch_mns_batch((const uint32_t *)d, (const DeferredCall *)(d + o_calls[BK_MNS]),
                                (const DeferredTask *)(d + o_tasks[BK_MNS]),
                                (int)b.tasks[BK_MNS].size(), pos, acc, S.bstream),
               "deferred pMM_sK1 kernel") &&
         check(launch_qns_batch(qs, (const int *)d, (const DeferredCall *)(d + o_calls[BK_QNS]),
                                (const DeferredTask *)(d + o_tasks[BK_QNS]),
                                (int)b.tasks[BK_QNS].size(), (const float4 *)S.qpos4.p,
                                (const float *)S.qmom.p, pos, acc, S.bstream),
               "deferred pQQ1 kernel");
    for (int k = 0; k < BK_KINDS; k++) {
        b.calls[k].clear();
        b.tasks[k].clear();
        b.blocks[k].clear();
        b.prep[k] = 0;
    }
    b.used = 0;
    b.table_bytes = 0;
    b.ncalls = 0;
    S.bcur ^= 1;
    return ok;
}

// Room for `data` more bytes of lists and `tables` more bytes of tables in the
// current batch; flushes it first when it is full.  Returns the batch, or null.
static Batch *batch_room(size_t data, size_t tables) {
    for (int pass = 0; pass < 2; pass++) {
        Batch &b = S.batch[S.bcur];
        if (b.pending) {  // the copy out of this buffer from its last flush
            if (!check(hipEventSynchronize(b.copied), "hipEventSynchronize"))
                return nullptr;
            b.pending = false;
        }
        // lists, tables and the 128-B alignment of 3 tables per kind
        const size_t need = align128(b.used + data) + b.table_bytes + tables + 3 * 128 * BK_KINDS;
        if (b.ncalls && (b.used + data > BATCH_BYTES || need > b.host.cap)) {
            if (!flush_batch())
                return nullptr;
            continue;
        }
        if (!b.host.ensure(std::max(need, BATCH_BYTES + ((size_t)1 << 20))))
            return nullptr;
        return &b;
    }
    return nullptr;
}

// Appends one call's tasks (64 sinks x CHUNK[kind] sources each) and table bytes.
static void batch_tasks(Batch &b, int kind, int m, int n) {
    const int call = (int)b.calls[kind].size() - 1;
    const int ch = CHUNK[kind];
    for (int s0 = 0; s0 < m; s0 += 64)
        for (int i0 = 0; i0 < n; i0 += ch) b.tasks[kind].push_back(make_int4(call, s0, i0, std::min(n, i0 + ch)));
    b.ncalls++;
}

static int64_t task_count(int kind, int64_t m, int64_t n) {
    return ((m + 63) / 64) * ((n + CHUNK[kind] - 1) / CHUNK[kind]);
}

static size_t task_bytes(int kind, int m, int n) {
    return sizeof(DeferredCall) + sizeof(DeferredTask) * (size_t)task_count(kind, m, n);
}

// A call with more tasks than this fills the GPU by itself and runs at once on
// its slot (its task table alone would take more than 1 MB of the batch), as
// does a call whose list is larger than a batch: the batches keep their size.
static const int64_t DEFER_MAX_TASKS = 65536;

static bool deferrable(int kind, int64_t m, int64_t n, size_t bytes) {
    return S.defer && bytes <= BATCH_BYTES && task_count(kind, m, n) <= DEFER_MAX_TASKS;
}

static int mn_kind(int order, int soft) {
    return order == 0 ? (soft == SOFT_K1 ? BK_MK1 : BK_M) : (order == 2 ? BK_Q : BK_H);
}

}  // namespace pgrav

using namespace pgrav;

// ====================================================================== C ABI
extern "C" {

int pgrav_set_error_mode(int mode) {
    int old = g_error_mode;
    g_error_mode = mode ? 1 : 0;
    return old;
}
const char *pgrav_last_error(void) { return g_failed ? g_error.c_str() : nullptr; }
void pgrav_clear_error(void) { clear_error(); }
void pgrav_set_quad_phi_sign(int sign) {
    std::lock_guard<std::mutex> lk(g_mutex);
    if (S.init)
        flush_batch();  // calls already made keep the sign they were made with
    g_qsign = sign < 0 ? -1 : 1;
}

int pgrav_set_dropin_deferral(int enable) {
    std::lock_guard<std::mutex> lk(g_mutex);
    const int old = S.defer ? 1 : 0;
    if (S.init && S.defer && !enable)
        flush_batch();
    S.defer = enable != 0;
    return old;
}

int pgrav_synchronize(void) {
    std::lock_guard<std::mutex> lk(g_mutex);
    if (!ensure_init() || !flush_batch())
        return -1;
    return check(hipDeviceSynchronize(), "hipDeviceSynchronize") ? 0 : -1;
}

void CUDA_Init(void *device_prop) {
    (void)device_prop;
    std::lock_guard<std::mutex> lk(g_mutex);
    if (!ensure_init())
        return;
    if (getenv("PGRAV_VERBOSE")) {
        hipDeviceProp_t pr;
        if (check(hipGetDeviceProperties(&pr, S.device), "hipGetDeviceProperties"))
            fprintf(stderr, "libpgrav: device %d [%s] %s, %d CUs, %.1f GB\n", S.device, pr.name,
                    pr.gcnArchName, pr.multiProcessorCount, pr.totalGlobalMem / 1073741824.0);
    }
}

void CUDA_Finalize(void) {
    std::lock_guard<std::mutex> lk(g_mutex);
    if (!S.init)
        return;
    wait_all();
    release_slots();
    S.devpos.release();
    S.devacc.release();
    S.errflag.release();
    S.staging.release();
    S.qraw.release();
    S.qpos4.release();
    S.qmom.release();
    for (Batch &b : S.batch) {
        b.host.release();
        b.dev.release();
        if (b.copied)
            (void)hipEventDestroy(b.copied);
        b.copied = nullptr;
        b.pending = false;
    }
    S.bpos4.release();
    S.bmom.release();
    if (S.bstream)
        (void)hipStreamDestroy(S.bstream);
    S.bstream = nullptr;
    for (int i = 0; i < NQ; i++) {
        if (S.slot[i].stream)
            (void)hipStreamDestroy(S.slot[i].stream);
        if (S.slot[i].staged)
            (void)hipEventDestroy(S.slot[i].staged);
        S.slot[i].stream = nullptr;
        S.slot[i].staged = nullptr;
        S.slot[i].staged_pending = false;
        S.slot[i].inuse = false;
    }
    S.btab = nullptr;
    S.nobj = 0;
    S.init = false;
}

int vecwidthCUDA(void) { return 4; }

int qallocCUDA(int *inuse) {
    std::lock_guard<std::mutex> lk(g_mutex);
    if (!ensure_init())
        return -1;
    if (!inuse) {
        // nln/mac.c's callers pass NULL: hand out slots round robin and stop at the
        // first idle one -- the slot used NQ calls ago has usually finished, so one
        // stream query instead of one per busy slot
        static int next = 0;
        for (int k = 0; k < NQ; k++) {
            const int i = (next + k) % NQ;
            Slot &s = S.slot[i];
            if (s.inuse) {
                const hipError_t e = hipStreamQuery(s.stream);
                if (e == hipErrorNotReady)
                    continue;
                if (e != hipSuccess) {
                    check(e, "hipStreamQuery");
                    return -1;
                }
                s.inuse = false;
            }
            next = (i + 1) % NQ;
            return i;
        }
        return -1;
    }
    int busy = 0;
    for (int i = 0; i < NQ; i++) {
        Slot &s = S.slot[i];
        if (!s.inuse)
            continue;
        hipError_t e = hipStreamQuery(s.stream);
        if (e == hipSuccess)
            s.inuse = false;
        else if (e == hipErrorNotReady)
            busy++;
        else {
            check(e, "hipStreamQuery");
            return -1;
        }
    }
    if (inuse)
        *inuse = busy;
    if (busy == NQ)
        return -1;
    for (int i = 0; i < NQ; i++)
        if (!S.slot[i].inuse)
            return i;
    return -1;
}

void WalkInitSinkCUDA(float *btab, int stride, int64_t nobj) {
    std::lock_guard<std::mutex> lk(g_mutex);
    if (!ensure_init())
        return;
    if (nobj < 0 || stride < 4) {
        fail("WalkInitSinkCUDA: bad nobj %lld / stride %d", (long long)nobj, stride);
        return;
    }
    wait_all();
    S.btab = btab;
    S.nobj = nobj;
    S.seg_mismatch = 0;
    S.stride = stride;
    size_t pos_bytes = 3 * sizeof(float) * (size_t)(nobj ? nobj : 1);
    size_t acc_bytes = 4 * sizeof(float) * (size_t)(nobj ? nobj : 1);
    if (!S.devpos.ensure(pos_bytes) || !S.devacc.ensure(acc_bytes) || !S.staging.ensure(pos_bytes))
        return;
    float *h = (float *)S.staging.p;
    for (int64_t i = 0; i < nobj; i++) {
        h[3 * i + 0] = btab[i * stride + 1];
        h[3 * i + 1] = btab[i * stride + 2];
        h[3 * i + 2] = btab[i * stride + 3];
    }
    if (!check(hipMemcpy(S.devpos.p, h, 3 * sizeof(float) * nobj, hipMemcpyHostToDevice),
               "hipMemcpy"))
        return;
    // cleared before any slot stream may add into the table
    check(hipMemset(S.devacc.p, 0, acc_bytes), "hipMemset") &&
        check(hipMemset(S.errflag.p, 0, sizeof(int)), "hipMemset") &&
        check(hipDeviceSynchronize(), "hipDeviceSynchronize");
}

void WalkTerminateSinkCUDA(float *btab, int stride, int64_t nobj) {
    std::lock_guard<std::mutex> lk(g_mutex);
    if (!ensure_init())
        return;
    wait_all();
    if (!S.devacc.p || nobj > S.nobj) {
        fail("WalkTerminateSinkCUDA without a matching WalkInitSinkCUDA");
        return;
    }
    size_t acc_bytes = 4 * sizeof(float) * (size_t)nobj;
    if (!S.staging.ensure(acc_bytes ? acc_bytes : 16))
        return;
    float *h = (float *)S.staging.p;
    if (!check(hipMemcpy(h, S.devacc.p, acc_bytes, hipMemcpyDeviceToHost), "hipMemcpy"))
        return;
    int err = 0;
    if (!check(hipMemcpy(&err, S.errflag.p, sizeof(int), hipMemcpyDeviceToHost), "hipMemcpy"))
        return;
    err += S.seg_mismatch;
    S.seg_mismatch = 0;
    for (int64_t i = 0; i < nobj; i++) {
        btab[i * stride + 7] += h[4 * i + 0];
        btab[i * stride + 8] += h[4 * i + 1];
        btab[i * stride + 9] += h[4 * i + 2];
        btab[i * stride + 10] += h[4 * i + 3];
    }
    // the reference frees its slot buffers and device tables here
    release_slots();
    S.devpos.release();
    S.devacc.release();
    S.btab = nullptr;
    S.nobj = 0;
    if (err)
        fail("%d segment lists did not sum to their source_n (nln/pgrav.cu:779)", err);
}

void WalkInitSrcCUDA(float *qtab, int stride, int64_t nobj) {
    std::lock_guard<std::mutex> lk(g_mutex);
    if (!ensure_init())
        return;
    if (stride <= 0 || stride % 4 || stride < 80) {
        fail("WalkInitSrcCUDA: stride %d bytes is not a hexacell/quadcell record", stride);
        return;
    }
    wait_all();  // queued calls may still read the previous cell table
    size_t bytes = (size_t)stride * (size_t)nobj;
    if (!S.qraw.ensure(bytes ? bytes : 16) || !S.qpos4.ensure(16 * (size_t)(nobj ? nobj : 1)) ||
        !S.qmom.ensure(PREC2 * 4 * (size_t)(nobj ? nobj : 1)))
        return;
    if (!check(hipMemcpy(S.qraw.p, qtab, bytes, hipMemcpyHostToDevice), "hipMemcpy"))
        return;
    if (!check(launch_prep_hexacell((const float *)S.qraw.p, stride / 4, nobj,
                                    (float4 *)S.qpos4.p, (float *)S.qmom.p, 0),
               "prep kernel"))
        return;
    check(hipDeviceSynchronize(), "hipDeviceSynchronize");
    S.qcells = nobj;
}

void WalkTerminateSrcCUDA(void) {
    std::lock_guard<std::mutex> lk(g_mutex);
    wait_all();
    S.qraw.release();
    S.qpos4.release();
    S.qmom.release();
    S.qcells = 0;
}

static void mn_common(int order, int soft, const float *p, int m, int stride, const float *f,
                      int source_n, int sz, float e, int q) {
    if (!slot_ok(q))
        return;
    if (stride <= 0) {
        fail("bad stride %d", stride);
        return;
    }
    int64_t off = (int64_t)(p - S.btab);
    if (off % stride) {
        fail("sink pointer is not at a body boundary");
        return;
    }
    off /= stride;
    if (!sinks_ready(off, m))
        return;
    const int need = order == 0 ? 4 : (order == 2 ? 13 : 29);
    if (sz < need) {
        fail("record size %d too small for order %d (needs %d floats)", sz, order, need);
        return;
    }
    if (source_n < 0 || (source_n % 8)) {
        fail("source_n %d is not a multiple of NSSE = 8", source_n);
        return;
    }
    size_t bytes = (size_t)source_n * sz * sizeof(float);
    if (deferrable(mn_kind(order, soft), m, source_n, bytes)) {
        if (m <= 0 || source_n <= 0)
            return;
        const int kind = mn_kind(order, soft);
        const int nblk = (source_n + 127) / 128;
        const size_t tb = task_bytes(kind, m, source_n) + sizeof(int) * (size_t)nblk;
        Batch *b = batch_room(bytes + 64, tb);
        if (!b)
            return;
        const size_t at = (b->used + 63) & ~(size_t)63;
        memcpy((char *)b->host.p + at, f, bytes);  // the caller may reuse f on return
        b->used = at + bytes;
        DeferredCall c = {off, (int64_t)(at / 4), m, source_n, sz, b->prep[kind], e, 0.0f};
        b->calls[kind].push_back(c);
        const int call = (int)b->calls[kind].size() - 1;
        for (int k = 0; k < nblk; k++) b->blocks[kind].push_back(call);
        b->prep[kind] += 128 * nblk;
        b->table_bytes += tb;
        batch_tasks(*b, kind, m, source_n);
        return;
    }
    Slot &s = S.slot[q];
    const int mstride = order == 4 ? PREC4 : PREC2;
    if (!stage(s, bytes + 16, bytes + 16, 16 * (size_t)(source_n + 1),
               4 * (size_t)mstride * (source_n + 1)))
        return;
    memcpy(s.host.p, f, bytes);  // the caller may reuse f on return (nln/pgrav.cu:1053)
    if (!upload(s, bytes))
        return;
    if (!check(launch_prep_soa8((const float *)s.dev.p, source_n, sz, order, (float4 *)s.pos4.p,
                                (float *)s.mom.p, s.stream),
               "prep kernel"))
        return;
    check(launch_mn(order, soft, quad_phi_sign(), (const float *)S.devpos.p + 3 * off,
                    (float *)S.devacc.p + 4 * off, m, (const float4 *)s.pos4.p,
                    (const float *)s.mom.p, source_n, e, s.stream),
          "grav_mn kernel");
}

void grav_mn_CUDA(const char *routine, const float *p, float *accp, const int m,
                  const int stride, const float *f, const int source_n, const int sz,
                  const float e, int *ncut, int q) {
    (void)accp;
    (void)ncut;
    std::lock_guard<std::mutex> lk(g_mutex);
    int order, soft = SOFT_NONE;
    if (!routine)
        return;
    if (!strcmp(routine, "pM"))
        order = 0;
    else if (!strcmp(routine, "pQ"))
        order = 2;
    else if (!strcmp(routine, "pH"))
        order = 4;
    else if (!strcmp(routine, "pM_sK1")) {
        order = 0;
        soft = SOFT_K1;
    } else
        return;  // unknown names are ignored (nln/pgrav.cu:1074-1081)
    mn_common(order, soft, p, m, stride, f, source_n, sz, e, q);
}

void pinteractCUDA(const float *p, float *accp, const int m, const int stride, const float *f,
                   const int source_n, const int sz, int q) {
    (void)accp;
    std::lock_guard<std::mutex> lk(g_mutex);
    int order;
    if (sz == 4)
        order = 0;
    else if (sz == 13)
        order = 2;
    else if (sz == 29)
        order = 4;
    else {
        fail("Unknown size %d", sz);  // nln/py_pgrav_cu/pgrav.cu:563
        return;
    }
    mn_common(order, SOFT_NONE, p, m, stride, f, source_n, sz, 0.0f, q);
}

void grav_mns_CUDA(const char *routine, const int base, const int m, const segment *seg,
                   const int seg_n, const int source_n, const float mmass, float e, int *ncut,
                   int q) {
    (void)ncut;
    std::lock_guard<std::mutex> lk(g_mutex);
    if (!routine || strcmp(routine, "pMM_sK1")) {
        fail("routine %s not found", routine ? routine : "(null)");
        return;
    }
    if (!slot_ok(q) || !sinks_ready(base, m))
        return;
    int64_t total = 0;
    for (int g = 0; g < seg_n; g++) {
        if ((int64_t)seg[g].base + seg[g].length > S.nobj) {
            fail("segment %d [%u,+%u) outside the body table", g, seg[g].base, seg[g].length);
            return;
        }
        total += seg[g].length;
    }
    if (total > INT32_MAX) {
        fail("pMM_sK1 segments hold %lld bodies (at most %d)", (long long)total, INT32_MAX);
        return;
    }
    if (total != source_n)
        S.seg_mismatch++;  // assert(sum_n == source_n), nln/pgrav.cu:636; reported at terminate
    // the segments expanded into body indices: one coalesced load per 64 bodies
    size_t bytes = (size_t)total * sizeof(uint32_t);
    if (deferrable(BK_MNS, m, total, bytes)) {
        if (m <= 0 || total <= 0)
            return;
        const size_t tb = task_bytes(BK_MNS, m, (int)total);
        Batch *b = batch_room(bytes + 64, tb);
        if (!b)
            return;
        const size_t at = (b->used + 63) & ~(size_t)63;
        uint32_t *h = (uint32_t *)((char *)b->host.p + at);
        for (int g = 0; g < seg_n; g++)
            for (uint32_t j = 0; j < seg[g].length; j++) *h++ = seg[g].base + j;
        b->used = at + bytes;
        DeferredCall c = {(int64_t)base, (int64_t)(at / 4), m, (int)total, 0, 0, e, mmass};
        b->calls[BK_MNS].push_back(c);
        b->table_bytes += tb;
        batch_tasks(*b, BK_MNS, m, (int)total);
        return;
    }
    Slot &s = S.slot[q];
    if (!stage(s, bytes + 16, bytes + 16))
        return;
    uint32_t *h = (uint32_t *)s.host.p;
    for (int g = 0; g < seg_n; g++)
        for (uint32_t j = 0; j < seg[g].length; j++) *h++ = seg[g].base + j;
    if (total && !upload(s, bytes))
        return;
    check(launch_mns((const float *)S.devpos.p + 3 * (size_t)base,
                     (float *)S.devacc.p + 4 * (size_t)base, m, (const uint32_t *)s.dev.p,
                     (int)total, (const float *)S.devpos.p, mmass, e, s.stream),
          "pMM_sK1 kernel");
}

void grav_mnss_CUDA(const char *routine, const int sink_base, const int m,
                    const uint16_t *ss_index, const segment *ss_seg, const int ss_len,
                    const segment *seg, const int seg_len, const int source_base,
                    const int *source_n, const float mmass, float e, int *ncut, int q) {
    (void)ncut;
    std::lock_guard<std::mutex> lk(g_mutex);
    if (!routine || strcmp(routine, "pMM_mnss_sK1")) {
        fail("routine %s not found", routine ? routine : "(null)");
        return;
    }
    if (!slot_ok(q) || !sinks_ready(sink_base, m))
        return;
    if (source_base < 0 || source_base > S.nobj) {
        fail("source_base %d outside the body table", source_base);
        return;
    }
    for (int g = 0; g < seg_len; g++)
        if ((int64_t)source_base + seg[g].base + seg[g].length > S.nobj) {
            fail("segment %d outside the body table", g);
            return;
        }
    // one packed, 128-B aligned upload (nln/pgrav.cu:1218-1250)
    size_t o_ss_seg = align128((size_t)m * 2);
    size_t o_srcn = align128(o_ss_seg + (size_t)ss_len * 8);
    size_t o_seg = align128(o_srcn + (size_t)ss_len * 4);
    size_t total = o_seg + (size_t)seg_len * 8;
    Slot &s = S.slot[q];
    if (!stage(s, total + 16, total + 16))
        return;
    char *h = (char *)s.host.p;
    memcpy(h, ss_index, (size_t)m * 2);
    memcpy(h + o_ss_seg, ss_seg, (size_t)ss_len * 8);
    memcpy(h + o_srcn, source_n, (size_t)ss_len * 4);
    memcpy(h + o_seg, seg, (size_t)seg_len * 8);
    if (!upload(s, total))
        return;
    char *d = (char *)s.dev.p;
    int64_t max_source_n = 0;
    for (int k = 0; k < ss_len; k++) max_source_n = std::max<int64_t>(max_source_n, source_n[k]);
    check(launch_mnss((const float *)S.devpos.p + 3 * (size_t)sink_base,
                      (float *)S.devacc.p + 4 * (size_t)sink_base, m, (const uint16_t *)d,
                      (const uint32_t *)(d + o_ss_seg), ss_len, (const uint32_t *)(d + o_seg),
                      seg_len, (const float *)S.devpos.p + 3 * (size_t)source_base,
                      (const int *)(d + o_srcn), max_source_n, mmass, e, (int *)S.errflag.p,
                      s.stream),
          "pMM_mnss_sK1 kernel");
}

void grav_qns_CUDA(const char *routine, const int base, const int m, const int *source_list,
                   const int source_n, int q) {
    std::lock_guard<std::mutex> lk(g_mutex);
    if (!routine || (strcmp(routine, "pQQ1") && strcmp(routine, "pQ1"))) {
        fail("routine %s not found", routine ? routine : "(null)");
        return;
    }
    if (!slot_ok(q) || !sinks_ready(base, m))
        return;
    if (!S.qpos4.p || S.qcells <= 0) {
        fail("WalkInitSrcCUDA has not been called");
        return;
    }
    for (int i = 0; i < source_n; i++)
        if (source_list[i] < 0 || source_list[i] >= S.qcells) {
            fail("source_list[%d] = %d outside the cell table", i, source_list[i]);
            return;
        }
    size_t bytes = (size_t)source_n * sizeof(int);
    if (deferrable(BK_QNS, m, source_n, bytes)) {
        if (m <= 0 || source_n <= 0)
            return;
        const size_t tb = task_bytes(BK_QNS, m, source_n);
        Batch *b = batch_room(bytes + 64, tb);
        if (!b)
            return;
        const size_t at = (b->used + 63) & ~(size_t)63;
        memcpy((char *)b->host.p + at, source_list, bytes);
        b->used = at + bytes;
        DeferredCall c = {(int64_t)base, (int64_t)(at / 4), m, source_n, 0, 0, 0.0f, 0.0f};
        b->calls[BK_QNS].push_back(c);
        b->table_bytes += tb;
        batch_tasks(*b, BK_QNS, m, source_n);
        return;
    }
    Slot &s = S.slot[q];
    if (!stage(s, bytes + 16, bytes + 16))
        return;
    memcpy(s.host.p, source_list, bytes);
    if (!upload(s, bytes))
        return;
    check(launch_qns(quad_phi_sign(), (const float *)S.devpos.p + 3 * (size_t)base,
                     (float *)S.devacc.p + 4 * (size_t)base, m, (const int *)s.dev.p, source_n,
                     (const float4 *)S.qpos4.p, (const float *)S.qmom.p, s.stream),
          "pQQ1 kernel");
}

void grav_qnss_CUDA(const char *routine, const int base, const int m, const segment *seg,
                    const float *source, const int source_n, int q) {
    std::lock_guard<std::mutex> lk(g_mutex);
    if (!routine || strcmp(routine, "pQ1_qnss")) {
        fail("routine %s not found", routine ? routine : "(null)");
        return;
    }
    if (!slot_ok(q) || !sinks_ready(base, m))
        return;
    if (source_n < 0 || source_n % 8) {
        fail("source_n %d is not a multiple of NSSE = 8", source_n);
        return;
    }
    int64_t max_len = 0;
    for (int i = 0; i < m; i++) {
        if ((int64_t)seg[i].base + seg[i].length > source_n) {
            fail("seg[%d] runs past the %d queued sources", i, source_n);
            return;
        }
        max_len = std::max<int64_t>(max_len, seg[i].length);
    }
    size_t o_src = align128((size_t)m * 8);
    size_t bytes = (size_t)source_n * 13 * sizeof(float);
    size_t total = o_src + bytes;
    Slot &s = S.slot[q];
    if (!stage(s, total + 16, total + 16, 16 * (size_t)(source_n + 1),
               4 * (size_t)PREC2 * (source_n + 1)))
        return;
    memcpy(s.host.p, seg, (size_t)m * 8);
    memcpy((char *)s.host.p + o_src, source, bytes);
    if (!upload(s, total))
        return;
    if (!check(launch_prep_soa8((const float *)((char *)s.dev.p + o_src), source_n, 13, 2,
                                (float4 *)s.pos4.p, (float *)s.mom.p, s.stream),
               "prep kernel"))
        return;
    check(launch_qnss(quad_phi_sign(), (const float *)S.devpos.p + 3 * (size_t)base,
                      (float *)S.devacc.p + 4 * (size_t)base, m, (const uint32_t *)s.dev.p,
                      max_len, (const float4 *)s.pos4.p, (const float *)s.mom.p, s.stream),
          "pQ1_qnss kernel");
}

int pgrav_mn_host(int kind, const float *p, float *accp, int m, int stride, const float *f,
                  int source_n, int sz, int layout, float e) {
    std::lock_guard<std::mutex> lk(g_mutex);
    clear_error();
    if (!ensure_init())
        return -1;
    int order, soft = SOFT_NONE;
    switch (kind) {
    case 0: order = 0; break;
    case 2: order = 2; break;
    case 4: order = 4; break;
    case 5: order = 0; soft = SOFT_K1; break;
    default: fail("pgrav_mn_host: unknown kind %d", kind); return -1;
    }
    const int need = order == 0 ? 4 : (order == 2 ? 13 : 29);
    if (m < 0 || source_n < 0 || stride < 4 || sz < need) {
        fail("pgrav_mn_host: bad sizes m=%d n=%d stride=%d sz=%d", m, source_n, stride, sz);
        return -1;
    }
    if (m == 0 || source_n == 0)
        return 0;
    // SoA-8 staging of the sources (layout 0 converts AoS on the host)
    const int nb = (source_n + 7) / 8;
    std::vector<float> soa((size_t)nb * 8 * sz, 0.0f);
    if (layout == 0) {
        for (int i = 0; i < source_n; i++)
            for (int k = 0; k < sz; k++)
                soa[(size_t)(i / 8) * 8 * sz + k * 8 + i % 8] = f[(size_t)i * sz + k];
    } else {
        memcpy(soa.data(), f, soa.size() * sizeof(float));
    }
    std::vector<float> sink((size_t)3 * m);
    for (int i = 0; i < m; i++)
        for (int k = 0; k < 3; k++) sink[3 * (size_t)i + k] = p[(size_t)i * stride + 1 + k];
    const int ns = nb * 8;
    const int mstride = order == 4 ? PREC4 : PREC2;
    void *d_src = nullptr, *d_sink = nullptr, *d_acc = nullptr, *d_pos4 = nullptr, *d_mom = nullptr;
    hipStream_t st = S.slot[0].stream;
    bool ok = check(hipMalloc(&d_src, soa.size() * 4), "hipMalloc") &&
              check(hipMalloc(&d_sink, sink.size() * 4), "hipMalloc") &&
              check(hipMalloc(&d_acc, (size_t)16 * m), "hipMalloc") &&
              check(hipMalloc(&d_pos4, (size_t)16 * ns), "hipMalloc") &&
              check(hipMalloc(&d_mom, (size_t)4 * mstride * ns), "hipMalloc");
    std::vector<float> acc((size_t)4 * m);
    if (ok)
        ok = check(hipStreamSynchronize(st), "hipStreamSynchronize") &&
             check(hipMemcpyAsync(d_src, soa.data(), soa.size() * 4, hipMemcpyHostToDevice, st), "hipMemcpyAsync") &&
             check(hipMemcpyAsync(d_sink, sink.data(), sink.size() * 4, hipMemcpyHostToDevice, st), "hipMemcpyAsync") &&
             check(hipMemsetAsync(d_acc, 0, (size_t)16 * m, st), "hipMemsetAsync") &&
             check(launch_prep_soa8((const float *)d_src, ns, sz, order, (float4 *)d_pos4,
                                    (float *)d_mom, st), "prep kernel") &&
             check(launch_mn(order, soft, quad_phi_sign(), (const float *)d_sink, (float *)d_acc, m,
                             (const float4 *)d_pos4, (const float *)d_mom, ns, e, st), "mn kernel") &&
             check(hipMemcpyAsync(acc.data(), d_acc, (size_t)16 * m, hipMemcpyDeviceToHost, st), "hipMemcpyAsync") &&
             check(hipStreamSynchronize(st), "hipStreamSynchronize");
    for (void *ptr : {d_src, d_sink, d_acc, d_pos4, d_mom})
        if (ptr)
            (void)hipFree(ptr);
    if (!ok)
        return -1;
    for (int i = 0; i < m; i++)
        for (int k = 0; k < 4; k++) accp[(size_t)i * stride + k] += acc[4 * (size_t)i + k];
    return 0;
}

}  // extern "C"
