// libpysweep_hip.so -- HIP runtime, hiprtc and native phase sequencing for
// pysweep_amd on MI355X (gfx950).  The C ABI is declared in
// include/pysweep_hip.h; every entry returns 0 or an error code and keeps the
// message for ps_last_error().
#include <hip/hip_runtime.h>
#include <climits>
#include <hip/hiprtc.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "pysweep_hip.h"

namespace {

thread_local std::string g_error;

int fail(const std::string &what)
{
    g_error = what;
    return 1;
}

int fail_hip(hipError_t e, const char *call, int line)
{
    char buf[512];
    std::snprintf(buf, sizeof(buf), "%s failed at pysweep_hip.cpp:%d: %s (%d)", call, line,
                  hipGetErrorString(e), static_cast<int>(e));
    g_error = buf;
    return static_cast<int>(e) ? static_cast<int>(e) : 1;
}

#define PS_HIP(call)                                              \
    do {                                                          \
        hipError_t e_ = (call);                                   \
        if (e_ != hipSuccess) return fail_hip(e_, #call, __LINE__); \
    } while (0)

// Optional per-kernel timing: an event pair around each sampled launch,
// resolved after the run (ps_run_stats in the header).
struct Timer {
    ps_run_stats *stats = nullptr;
    std::vector<hipEvent_t> ev;   // 2 per timed launch
    std::vector<int> kind;
    size_t used = 0;
    long long seen[PS_K_COUNT] = {};  // launches of each kind in this run call
};

// A device context: its device, its work stream and its launch timer.  The
// process context (ps_init) serves every thread that has no context of its
// own current; ps_context_push makes one current on the calling thread.
struct Ctx {
    int device = -1;
    hipStream_t stream = nullptr;
    Timer timer;
    double *sum_part = nullptr;  // partial sums of ps_level_sums (allocated on first use)
};

Ctx *g_process = nullptr;
thread_local std::vector<Ctx *> t_stack;

int scratch_setup(int device);

Ctx *current() { return t_stack.empty() ? g_process : t_stack.back(); }
hipStream_t stream() { return current()->stream; }

int create_ctx(int device, Ctx **out)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return fail("no HIP device is visible");
    if (device < 0 || device >= n) return fail("device ordinal out of range");
    PS_HIP(hipSetDevice(device));
    if (int rc = scratch_setup(device)) return rc;
    Ctx *c = new Ctx;
    c->device = device;
    const hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete c;
        return fail_hip(e, "hipStreamCreateWithFlags", __LINE__);
    }
    *out = c;
    return 0;
}

// Drain and release a context's stream and events.
int destroy_ctx(Ctx *c)
{
    int rc = 0;
    if (c->stream) {
        hipError_t e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) rc = fail_hip(e, "hipStreamSynchronize", __LINE__);
        e = hipStreamDestroy(c->stream);
        if (e != hipSuccess && !rc) rc = fail_hip(e, "hipStreamDestroy", __LINE__);
    }
    for (hipEvent_t ev : c->timer.ev) (void)hipEventDestroy(ev);
    if (c->sum_part) (void)hipFree(c->sum_part);
    delete c;
    return rc;
}

#define PS_NEED_INIT()                                                                            \
    do {                                                                                          \
        if (!current()) return fail("ps_init() has not been called and no context is current"); \
    } while (0)

char *dup_host(const std::string &s)
{
    char *p = static_cast<char *>(std::malloc(s.size() + 1));
    if (p) std::memcpy(p, s.c_str(), s.size() + 1);
    return p;
}

int launch(ps_function f, unsigned gx, unsigned threads, void **args)
{
    PS_HIP(hipModuleLaunchKernel(static_cast<hipFunction_t>(f), gx, 1, 1, threads, 1, 1, 0, stream(),
                                 args, nullptr));
    return 0;
}

int timer_begin(ps_run_stats *stats)
{
    Timer &g_timer = current()->timer;
    g_timer.stats = stats;
    g_timer.used = 0;
    g_timer.kind.clear();
    for (long long &n : g_timer.seen) n = 0;
    return 0;
}

// whether this launch of `kind` is one the caller asked to time
bool timer_wants(int kind)
{
    Timer &g_timer = current()->timer;
    const ps_run_stats *st = g_timer.stats;
    if (!st) return false;
    const long long n = g_timer.seen[kind]++;
    if (st->kinds && !(st->kinds & (1 << kind))) return false;
    return st->every <= 1 || n % st->every == 0;
}

int timer_mark(bool start, int kind)
{
    Timer &g_timer = current()->timer;
    if (start) {
        if (g_timer.used + 2 > g_timer.ev.size()) {
            for (int i = 0; i < 64; ++i) {
                hipEvent_t e;
                PS_HIP(hipEventCreate(&e));
                g_timer.ev.push_back(e);
            }
        }
        g_timer.kind.push_back(kind);
    }
    PS_HIP(hipEventRecord(g_timer.ev[g_timer.used++], stream()));
    return 0;
}

int timer_end()
{
    Timer &g_timer = current()->timer;
    ps_run_stats *st = g_timer.stats;
    g_timer.stats = nullptr;
    if (!st) return 0;
    PS_HIP(hipStreamSynchronize(stream()));
    for (size_t i = 0; i < g_timer.kind.size(); ++i) {
        float ms = 0.f;
        PS_HIP(hipEventElapsedTime(&ms, g_timer.ev[2 * i], g_timer.ev[2 * i + 1]));
        const int k = g_timer.kind[i];
        st->ms[k] += ms;
        st->launches[k] += 1;
    }
    return 0;
}

int launch_kind(ps_function f, unsigned gx, unsigned threads, void **args, int kind)
{
    const bool timed = timer_wants(kind);
    int rc = timed ? timer_mark(true, kind) : 0;
    if (rc) return rc;
    if ((rc = launch(f, gx, threads, args))) return rc;
    return timed ? timer_mark(false, kind) : 0;
}

__global__ void ps_copy_kernel(double *__restrict__ dst, const double *__restrict__ src, size_t n)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) dst[i] = src[i];
}

// One phase launch: K(state, gts, A0, grid).
int launch_phase(const ps_swept_plan *p, ps_function f, int gts, int A0, int xoff, int keep_from, int kind)
{
    if (!f) return fail("swept plan has a null phase kernel");
    ps_grid g;
    g.nx = p->nx;
    g.ny = p->ny;
    g.T = p->T;
    g.xoff = xoff;
    g.tiles_x = p->tile_rows > 0 ? p->tile_rows : p->nx / p->b;
    g.tiles_y = p->ny / p->b;
    g.keep_from = keep_from;
    void *state = p->state;
    void *args[] = {&state, &gts, &A0, &g};
    return launch_kind(f, static_cast<unsigned>(g.tiles_x * g.tiles_y), static_cast<unsigned>(p->threads), args,
                       kind);
}

// ---- FP64 sums of a level's variables (conserved-quantity checks) ----
// The grid shape is fixed, so the order of the additions, and with it the
// result, does not depend on the device or the call.
constexpr int PS_SUM_BLOCKS = 512;
constexpr int PS_SUM_THREADS = 256;
constexpr int PS_SUM_MAX_NV = 8;
constexpr size_t PS_SUM_PARTIALS = 2 * PS_SUM_MAX_NV * PS_SUM_BLOCKS;  // doubles

__device__ __forceinline__ void ps_sum_tree(double *red, double &a, double &b)
{
    // pairwise over the workgroup; red holds 2 * PS_SUM_THREADS doubles
    red[threadIdx.x] = a;
    red[PS_SUM_THREADS + threadIdx.x] = b;
    __syncthreads();
    for (int w = PS_SUM_THREADS / 2; w > 0; w >>= 1) {
        if (static_cast<int>(threadIdx.x) < w) {
            red[threadIdx.x] += red[threadIdx.x + w];
            red[PS_SUM_THREADS + threadIdx.x] += red[PS_SUM_THREADS + threadIdx.x + w];
        }
        __syncthreads();
    }
    a = red[0];
    b = red[PS_SUM_THREADS];
}

// Block (i, v): the i-th of PS_SUM_BLOCKS contiguous chunks of variable v;
// part[v][i] = sum, part[nv + v][i] = sum of magnitudes, both in FP64.
template <class R>
__global__ void __launch_bounds__(PS_SUM_THREADS)
    ps_sum_partial(const R *__restrict__ lvl, size_t cells, double *__restrict__ part)
{
    __shared__ double red[2 * PS_SUM_THREADS];
    const int v = blockIdx.y, nv = gridDim.y;
    const R *__restrict__ q = lvl + v * cells;
    const size_t chunk = (cells + PS_SUM_BLOCKS - 1) / PS_SUM_BLOCKS;
    const size_t lo = blockIdx.x * chunk;
    const size_t hi = lo + chunk < cells ? lo + chunk : cells;
    double s = 0.0, m = 0.0;
    for (size_t i = lo + threadIdx.x; i < hi; i += PS_SUM_THREADS) {
        const double x = static_cast<double>(q[i]);
        s += x;
        m += fabs(x);
    }
    ps_sum_tree(red, s, m);
    if (threadIdx.x == 0) {
        part[v * PS_SUM_BLOCKS + blockIdx.x] = s;
        part[(nv + v) * PS_SUM_BLOCKS + blockIdx.x] = m;
    }
}

// Block v: sums[v] and sums[nv + v] from the partials of variable v.
__global__ void __launch_bounds__(PS_SUM_THREADS)
    ps_sum_final(const double *__restrict__ part, int nv, double *__restrict__ sums)
{
    __shared__ double red[2 * PS_SUM_THREADS];
    const int v = blockIdx.x;
    double s = 0.0, m = 0.0;
    for (int i = threadIdx.x; i < PS_SUM_BLOCKS; i += PS_SUM_THREADS) {
        s += part[v * PS_SUM_BLOCKS + i];
        m += part[(nv + v) * PS_SUM_BLOCKS + i];
    }
    ps_sum_tree(red, s, m);
    if (threadIdx.x == 0) {
        sums[v] = s;
        sums[nv + v] = m;
    }
}

// Scratch of ps_level_sums_noalloc: the partial sums in a file-scope device
// array instead of a buffer the context allocates.  Every GPU the library
// runs on holds its own copy until the process ends.  In host code the
// array's name is not a device address: the address of each device's copy
// comes from hipGetSymbolAddress, taken once per device (with that device
// current) when the first context on it is made, and only that address is
// handed to kernels.  Every block of ps_sum_partial writes all of its
// partial sums, so nothing in the array needs to start at zero.
static __device__ __align__(256) double scratch[PS_SUM_PARTIALS];
constexpr int PS_MAX_DEVICES = 64;
double *g_scratch[PS_MAX_DEVICES] = {};
std::mutex g_scratch_mutex;

// Address of `device`'s copy of the scratch array; `device` is current.
int scratch_setup(int device)
{
    if (device < 0 || device >= PS_MAX_DEVICES) return fail("device ordinal beyond the scratch table");
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (g_scratch[device]) return 0;
    void *p = nullptr;
    PS_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_scratch[device] = static_cast<double *>(p);
    return 0;
}

int level_sums(const void *level, int fp32, int nv, size_t cells, double *sums, double *part)
{
    if (!level || !sums) return fail("level sums: null level or output");
    if (nv < 1 || nv > PS_SUM_MAX_NV) return fail("level sums: nv must be 1..8");
    (void)hipGetLastError();
    const dim3 grid(PS_SUM_BLOCKS, static_cast<unsigned>(nv));
    if (fp32)
        hipLaunchKernelGGL(ps_sum_partial<float>, grid, dim3(PS_SUM_THREADS), 0, stream(),
                           static_cast<const float *>(level), cells, part);
    else
        hipLaunchKernelGGL(ps_sum_partial<double>, grid, dim3(PS_SUM_THREADS), 0, stream(),
                           static_cast<const double *>(level), cells, part);
    PS_HIP(hipGetLastError());
    hipLaunchKernelGGL(ps_sum_final, dim3(static_cast<unsigned>(nv)), dim3(PS_SUM_THREADS), 0, stream(),
                       static_cast<const double *>(part), nv, sums);
    PS_HIP(hipGetLastError());
    return 0;
}

}  // namespace

extern "C" {

int ps_version(void) { return 6; }

const char *ps_last_error(void) { return g_error.c_str(); }

int ps_device_count(int *count)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
    *count = n;
    return 0;
}

int ps_init(int device)
{
    if (g_process && g_process->device == device) return 0;
    Ctx *c = nullptr;
    int rc = create_ctx(device, &c);
    if (rc) return rc;
    if (g_process) {
        rc = destroy_ctx(g_process);
        g_process = nullptr;
        if (rc) {
            destroy_ctx(c);
            return rc;
        }
        PS_HIP(hipSetDevice(device));
    }
    g_process = c;
    return 0;
}

int ps_shutdown(void)
{
    if (!g_process) return 0;
    Ctx *c = g_process;
    g_process = nullptr;
    return destroy_ctx(c);
}

int ps_context_create(int device, ps_context *ctx)
{
    if (!ctx) return fail("ps_context_create: null output");
    *ctx = nullptr;
    const Ctx *prev = current();
    Ctx *c = nullptr;
    int rc = create_ctx(device, &c);
    // creating a context must not change the calling thread's device
    if (prev && prev->device != device) (void)hipSetDevice(prev->device);
    if (rc) return rc;
    *ctx = c;
    return 0;
}

int ps_context_destroy(ps_context ctx)
{
    if (!ctx) return 0;
    Ctx *c = static_cast<Ctx *>(ctx);
    for (Ctx *t : t_stack)
        if (t == c) return fail("ps_context_destroy: the context is current on the calling thread");
    if (c == g_process) return fail("ps_context_destroy: the process context is released by ps_shutdown");
    return destroy_ctx(c);
}

int ps_context_push(ps_context ctx)
{
    if (!ctx) return fail("ps_context_push: null context");
    Ctx *c = static_cast<Ctx *>(ctx);
    PS_HIP(hipSetDevice(c->device));
    t_stack.push_back(c);
    return 0;
}

int ps_context_pop(void)
{
    if (t_stack.empty()) return fail("ps_context_pop: no context is current on this thread");
    t_stack.pop_back();
    const Ctx *c = current();
    if (c) PS_HIP(hipSetDevice(c->device));
    return 0;
}

int ps_context_current(ps_context *ctx)
{
    if (!ctx) return fail("ps_context_current: null output");
    *ctx = t_stack.empty() ? nullptr : t_stack.back();
    return 0;
}

int ps_context_device(int *device)
{
    PS_NEED_INIT();
    if (!device) return fail("ps_context_device: null output");
    *device = current()->device;
    return 0;
}

int ps_compile(const char *source, const char *name, const char *const *options, int noptions, char **code,
               size_t *size, char **log)
{
    *code = nullptr;
    *size = 0;
    if (log) *log = nullptr;
    hiprtcProgram prog;
    hiprtcResult r = hiprtcCreateProgram(&prog, source, name ? name : "pysweep.hip", 0, nullptr, nullptr);
    if (r != HIPRTC_SUCCESS) return fail(std::string("hiprtcCreateProgram: ") + hiprtcGetErrorString(r));
    r = hiprtcCompileProgram(prog, noptions, const_cast<const char **>(options));
    size_t logsz = 0;
    std::string logtxt;
    if (hiprtcGetProgramLogSize(prog, &logsz) == HIPRTC_SUCCESS && logsz > 1) {
        logtxt.resize(logsz);
        hiprtcGetProgramLog(prog, &logtxt[0]);
        while (!logtxt.empty() && logtxt.back() == '\0') logtxt.pop_back();
    }
    if (log && !logtxt.empty()) *log = dup_host(logtxt);
    if (r != HIPRTC_SUCCESS) {
        hiprtcDestroyProgram(&prog);
        return fail(std::string("hiprtc compile of ") + (name ? name : "source") + " failed: " +
                    hiprtcGetErrorString(r) + "\n" + logtxt);
    }
    size_t n = 0;
    if (hiprtcGetCodeSize(prog, &n) != HIPRTC_SUCCESS || n == 0) {
        hiprtcDestroyProgram(&prog);
        return fail("hiprtc produced no code object");
    }
    char *buf = static_cast<char *>(std::malloc(n));
    if (!buf) {
        hiprtcDestroyProgram(&prog);
        return fail("out of host memory for the code object");
    }
    hiprtcGetCode(prog, buf);
    hiprtcDestroyProgram(&prog);
    *code = buf;
    *size = n;
    return 0;
}

void ps_free_host(void *p) { std::free(p); }

int ps_compiler_version(int *major, int *minor)
{
    const hiprtcResult r = hiprtcVersion(major, minor);
    if (r != HIPRTC_SUCCESS) return fail(std::string("hiprtcVersion: ") + hiprtcGetErrorString(r));
    return 0;
}

int ps_module_load(const void *code, size_t size, ps_module *module)
{
    PS_NEED_INIT();
    (void)size;
    hipModule_t m = nullptr;
    PS_HIP(hipModuleLoadData(&m, code));
    *module = m;
    return 0;
}

int ps_module_unload(ps_module module)
{
    if (module) PS_HIP(hipModuleUnload(static_cast<hipModule_t>(module)));
    return 0;
}

int ps_module_get_global(ps_module module, const char *name, void **dptr, size_t *bytes)
{
    hipDeviceptr_t p = nullptr;
    size_t n = 0;
    hipError_t e = hipModuleGetGlobal(&p, &n, static_cast<hipModule_t>(module), name);
    if (e != hipSuccess) {
        (void)hipGetLastError();  // a failed lookup must not poison later calls
        return fail(std::string("no device global named '") + name + "' in the module");
    }
    *dptr = p;
    *bytes = n;
    return 0;
}

int ps_module_get_function(ps_module module, const char *name, ps_function *function)
{
    hipFunction_t f = nullptr;
    hipError_t e = hipModuleGetFunction(&f, static_cast<hipModule_t>(module), name);
    if (e != hipSuccess) {
        (void)hipGetLastError();  // a failed lookup must not poison later calls
        return fail(std::string("no kernel named '") + name + "' in the module");
    }
    *function = f;
    return 0;
}

int ps_function_info(ps_function function, int *regs, int *lds_bytes, int *max_threads)
{
    hipFunction_t f = static_cast<hipFunction_t>(function);
    int v = 0;
    PS_HIP(hipFuncGetAttribute(&v, HIP_FUNC_ATTRIBUTE_NUM_REGS, f));
    *regs = v;
    PS_HIP(hipFuncGetAttribute(&v, HIP_FUNC_ATTRIBUTE_SHARED_SIZE_BYTES, f));
    *lds_bytes = v;
    PS_HIP(hipFuncGetAttribute(&v, HIP_FUNC_ATTRIBUTE_MAX_THREADS_PER_BLOCK, f));
    *max_threads = v;
    return 0;
}

int ps_malloc(void **dptr, size_t bytes)
{
    PS_NEED_INIT();
    PS_HIP(hipMalloc(dptr, bytes));
    return 0;
}

int ps_free(void *dptr)
{
    if (dptr) PS_HIP(hipFree(dptr));
    return 0;
}

int ps_memcpy_htod(void *dst, const void *src, size_t bytes)
{
    PS_NEED_INIT();
    PS_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, stream()));
    PS_HIP(hipStreamSynchronize(stream()));
    return 0;
}

int ps_memcpy_dtoh(void *dst, const void *src, size_t bytes)
{
    PS_NEED_INIT();
    PS_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, stream()));
    PS_HIP(hipStreamSynchronize(stream()));
    return 0;
}

int ps_memcpy_dtod(void *dst, const void *src, size_t bytes)
{
    PS_NEED_INIT();
    PS_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, stream()));
    PS_HIP(hipStreamSynchronize(stream()));
    return 0;
}

int ps_host_malloc(void **ptr, size_t bytes)
{
    PS_NEED_INIT();
    PS_HIP(hipHostMalloc(ptr, bytes, hipHostMallocDefault));
    return 0;
}

int ps_host_free(void *ptr)
{
    if (ptr) PS_HIP(hipHostFree(ptr));
    return 0;
}

int ps_memcpy_htod_async(void *dst, const void *src, size_t bytes)
{
    PS_NEED_INIT();
    PS_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, stream()));
    return 0;
}

int ps_memcpy_dtoh_async(void *dst, const void *src, size_t bytes)
{
    PS_NEED_INIT();
    PS_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, stream()));
    return 0;
}

int ps_memcpy_dtod_async(void *dst, const void *src, size_t bytes)
{
    PS_NEED_INIT();
    PS_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, stream()));
    return 0;
}

int ps_memset(void *dst, int value, size_t bytes)
{
    PS_NEED_INIT();
    PS_HIP(hipMemsetAsync(dst, value, bytes, stream()));
    return 0;
}

int ps_launch(ps_function function, unsigned gx, unsigned gy, unsigned gz, unsigned bx, unsigned by, unsigned bz,
              unsigned shared_bytes, void **args)
{
    PS_NEED_INIT();
    PS_HIP(hipModuleLaunchKernel(static_cast<hipFunction_t>(function), gx, gy, gz, bx, by, bz, shared_bytes,
                                 stream(), args, nullptr));
    return 0;
}

int ps_synchronize(void)
{
    PS_NEED_INIT();
    PS_HIP(hipStreamSynchronize(stream()));
    return 0;
}

int ps_event_create(ps_event *event)
{
    hipEvent_t e = nullptr;
    PS_HIP(hipEventCreate(&e));
    *event = e;
    return 0;
}

int ps_event_record(ps_event event)
{
    PS_NEED_INIT();
    PS_HIP(hipEventRecord(static_cast<hipEvent_t>(event), stream()));
    return 0;
}

int ps_event_elapsed_ms(ps_event start, ps_event stop, float *ms)
{
    PS_HIP(hipEventSynchronize(static_cast<hipEvent_t>(stop)));
    PS_HIP(hipEventElapsedTime(ms, static_cast<hipEvent_t>(start), static_cast<hipEvent_t>(stop)));
    return 0;
}

int ps_event_synchronize(ps_event event)
{
    PS_HIP(hipEventSynchronize(static_cast<hipEvent_t>(event)));
    return 0;
}

int ps_event_query(ps_event event, int *done)
{
    const hipError_t e = hipEventQuery(static_cast<hipEvent_t>(event));
    if (e == hipErrorNotReady) {
        *done = 0;
        return 0;
    }
    PS_HIP(e);
    *done = 1;
    return 0;
}

int ps_event_destroy(ps_event event)
{
    if (event) PS_HIP(hipEventDestroy(static_cast<hipEvent_t>(event)));
    return 0;
}

int ps_stream_wait_event(ps_event event)
{
    PS_NEED_INIT();
    if (!event) return fail("ps_stream_wait_event: null event");
    PS_HIP(hipStreamWaitEvent(stream(), static_cast<hipEvent_t>(event), 0));
    return 0;
}

static int swept_run(const ps_swept_plan *p, int G, int c_begin, int c_end)
{
    if (!p || !p->state) return fail("swept plan without state");
    const int b = p->b, k = p->ops, s = p->its;
    if (b <= 0 || k <= 0 || s <= 0 || b % (2 * k) != 0) return fail("invalid block size / operating points");
    const int P = b / (2 * k) - 1;
    if (P < 1) return fail("block size too small: b/(2k)-1 must be >= 1");
    if (p->nx % b != 0 || p->ny % b != 0) return fail("nx and ny must be multiples of the block size");
    if (p->T != 2 * P + s) return fail("ring must hold 2P+s levels");
    if (p->tile_rows < 0 || p->tile_rows > p->nx / b) return fail("tile_rows out of range");
    if (G < 0 || c_begin < 0 || c_end > G + 2 || c_begin > c_end) return fail("cycle range out of bounds");
    const int h = b / 2, base = s - 1;
    // The final s levels (s-1+(G+1)P-s+1 ..) are always written in full.
    const int keep = p->store_all ? INT_MIN : (G + 1) * P;
    for (int c = c_begin; c < c_end; ++c) {
        const int xoff = (c % 2) * h;
        int rc = 0;
        if (c == 0) {
            // FirstPrism: stages 1..P
            if ((rc = launch_phase(p, p->up, 1, base, 0, keep, PS_K_UP))) return rc;
            if ((rc = launch_phase(p, p->ybridge, 1, base, 0, keep, PS_K_YBRIDGE))) return rc;
        } else if (c <= G) {
            // UpPrism c: stages (c-1)P+1 .. (c+1)P
            const int gts = (c - 1) * P + 1;
            const int A = base + (c - 1) * P;
            if ((rc = launch_phase(p, p->xbridge, gts, A, xoff, keep, PS_K_XBRIDGE))) return rc;
            if ((rc = launch_phase(p, p->octahedron, gts, A, xoff, keep, PS_K_OCTAHEDRON))) return rc;
            if ((rc = launch_phase(p, p->ybridge, gts + P, A + P, xoff, keep, PS_K_YBRIDGE))) return rc;
        } else {
            // LastPrism: stages GP+1 .. (G+1)P
            const int gts = G * P + 1;
            const int A = base + G * P;
            if ((rc = launch_phase(p, p->xbridge, gts, A, xoff, keep, PS_K_XBRIDGE))) return rc;
            if ((rc = launch_phase(p, p->down, gts, A, xoff, keep, PS_K_DOWN))) return rc;
        }
    }
    return 0;
}

int ps_swept_run(const ps_swept_plan *p, int G, int c_begin, int c_end)
{
    PS_NEED_INIT();
    return swept_run(p, G, c_begin, c_end);
}

int ps_swept_run_timed(const ps_swept_plan *p, int G, int c_begin, int c_end, ps_run_stats *stats)
{
    PS_NEED_INIT();
    timer_begin(stats);
    int rc = swept_run(p, G, c_begin, c_end);
    int rc2 = timer_end();
    return rc ? rc : rc2;
}

static int standard_run(const ps_standard_plan *p, int gts_begin, int gts_end)
{
    if (!p || !p->state || !p->standard) return fail("incomplete standard plan");
    if (p->T < p->its + 1) return fail("standard ring must hold at least its+1 levels");
    if (p->bx <= 0 || p->by <= 0 || p->bx > p->nx || p->by > p->ny) return fail("invalid standard tile");
    ps_grid g;
    g.nx = p->nx;
    g.ny = p->ny;
    g.T = p->T;
    g.xoff = 0;
    g.tiles_x = (p->nx + p->bx - 1) / p->bx;
    g.tiles_y = (p->ny + p->by - 1) / p->by;
    g.keep_from = 0;
    void *state = p->state;
    for (int gts = gts_begin; gts < gts_end; ++gts) {
        int A = p->its - 2 + gts;
        void *args[] = {&state, &gts, &A, &g};
        int rc = launch_kind(p->standard, static_cast<unsigned>(g.tiles_x * g.tiles_y),
                             static_cast<unsigned>(p->threads), args, PS_K_STANDARD);
        if (rc) return rc;
    }
    return 0;
}

int ps_standard_run(const ps_standard_plan *p, int gts_begin, int gts_end)
{
    PS_NEED_INIT();
    return standard_run(p, gts_begin, gts_end);
}

int ps_standard_run_timed(const ps_standard_plan *p, int gts_begin, int gts_end, ps_run_stats *stats)
{
    PS_NEED_INIT();
    timer_begin(stats);
    int rc = standard_run(p, gts_begin, gts_end);
    int rc2 = timer_end();
    return rc ? rc : rc2;
}

static int band_run(const ps_band_plan *p, int A, int stages)
{
    if (!p || !p->state || !p->band) return fail("incomplete band plan");
    const int H = p->levels, s = p->its;
    if (s < 1 || s > 2) return fail("band solve: intermediate steps must be 1 or 2");
    if (H < 2 || H > 16 || H % 2 || H % s) return fail("band solve: levels must be even, a multiple of s, 2..16");
    if (p->T < H + 1) return fail("band ring must hold at least levels + 1 levels");
    if (p->nx < 1 || p->ny < 1 || p->rows < 1) return fail("band solve: empty grid or row segment");
    if (A < s - 1 || (A - (s - 1)) % s) return fail("band solve: must start on a full-step level");
    if (stages < 0) return fail("band solve: negative stage count");
    const int wout = p->wout;
    if (wout < 1 || wout > 128) return fail("band solve: output columns per band must be 1..128");
    ps_band_grid g;
    g.nx = p->nx;
    g.ny = p->ny;
    g.T = p->T;
    g.rows = p->rows;
    g.bands = (p->ny + wout - 1) / wout;
    g.segs = (p->nx + p->rows - 1) / p->rows;
    const long long waves = static_cast<long long>(g.bands) * g.segs;
    if (waves > (1LL << 30)) return fail("band solve: grid too large");
    const unsigned groups = static_cast<unsigned>((waves + 3) / 4);
    void *state = p->state;
    for (int done = 0; done < stages;) {
        g.hh = stages - done < H ? stages - done : H;
        // the next launch reads only its first level (a full step); the
        // last launch leaves its last s levels complete
        const bool last = done + g.hh >= stages;
        g.wlo = p->store_all ? 1 : !last ? g.hh : (g.hh - s + 1 > 1 ? g.hh - s + 1 : 1);
        int gts = A - (s - 1) + 1;
        void *args[] = {&state, &gts, &A, &g};
        if (int rc = launch_kind(p->band, groups, 256, args, PS_K_BAND)) return rc;
        A += g.hh;
        done += g.hh;
    }
    return 0;
}

int ps_band_run(const ps_band_plan *p, int A, int stages)
{
    PS_NEED_INIT();
    return band_run(p, A, stages);
}

int ps_band_run_timed(const ps_band_plan *p, int A, int stages, ps_run_stats *stats)
{
    PS_NEED_INIT();
    timer_begin(stats);
    int rc = band_run(p, A, stages);
    int rc2 = timer_end();
    return rc ? rc : rc2;
}

int ps_copy_f64(void *dst, const void *src, size_t n)
{
    PS_NEED_INIT();
    const unsigned threads = 256;
    size_t blocks = (n + threads - 1) / threads;
    if (blocks > 8192) blocks = 8192;
    if (blocks == 0) return 0;
    (void)hipGetLastError();
    hipLaunchKernelGGL(ps_copy_kernel, dim3(static_cast<unsigned>(blocks)), dim3(threads), 0, stream(),
                       static_cast<double *>(dst), static_cast<const double *>(src), n);
    PS_HIP(hipGetLastError());
    return 0;
}

__global__ void ps_owner_kernel(const int *__restrict__ prefix, int n, int *__restrict__ owner)
{
    const int i = blockIdx.x;
    if (i >= n) return;
    const int lo = prefix[i], hi = prefix[i + 1];
    for (int t = lo + threadIdx.x; t < hi; t += blockDim.x) owner[t] = i;
}

static int multi_round(const ps_multi_plan *p)
{
    void *insts = p->insts, *prefix = p->prefix, *sprefix = p->sprefix, *owner = p->owner, *sowner = p->sowner;
    int n = p->ninst;
    void *args[] = {&insts, &prefix, &owner, &n};
    void *sargs[] = {&insts, &sprefix, &sowner, &n};
    int rc = 0;
    if (p->total_tiles > 0) {
        const unsigned g = static_cast<unsigned>(p->total_tiles);  // one workgroup per work item
        const unsigned t = static_cast<unsigned>(p->threads);
        // Within a round every problem keeps its own phase order:
        // FirstPrism Up<Yb, UpPrism Xb<Oct<Yb, LastPrism Xb<Down.
        if ((rc = launch_kind(p->up, g, t, args, PS_K_UP))) return rc;
        if ((rc = launch_kind(p->xbridge, g, t, args, PS_K_XBRIDGE))) return rc;
        if ((rc = launch_kind(p->octahedron, g, t, args, PS_K_OCTAHEDRON))) return rc;
        if ((rc = launch_kind(p->down, g, t, args, PS_K_DOWN))) return rc;
        if ((rc = launch_kind(p->ybridge, g, t, args, PS_K_YBRIDGE))) return rc;
    }
    if (p->total_stiles > 0 && p->standard) {
        const unsigned g = static_cast<unsigned>(p->total_stiles);
        if ((rc = launch_kind(p->standard, g, static_cast<unsigned>(p->sthreads), sargs, PS_K_STANDARD))) return rc;
    }
    if (p->total_bwaves > 0 && p->band) {
        void *bprefix = p->bprefix, *bowner = p->bowner;
        void *bargs[] = {&insts, &bprefix, &bowner, &n};
        // one wavefront per workgroup: the workgroup's LDS holds its problem's constants
        if ((rc = launch_kind(p->band, static_cast<unsigned>(p->total_bwaves), 64, bargs, PS_K_BAND))) return rc;
    }
    void *aargs[] = {&insts, &n};
    return launch(p->advance, static_cast<unsigned>((n + 255) / 256), 256, aargs);
}

int ps_level_sums(const void *level, int fp32, int nv, size_t cells, double *sums)
{
    PS_NEED_INIT();
    Ctx *c = current();
    if (!c->sum_part) PS_HIP(hipMalloc(&c->sum_part, PS_SUM_PARTIALS * sizeof(double)));
    return level_sums(level, fp32, nv, cells, sums, c->sum_part);
}

int ps_level_sums_noalloc(const void *level, int fp32, int nv, size_t cells, double *sums)
{
    PS_NEED_INIT();
    if (nv < 1 || nv > PS_SUM_MAX_NV)
        return fail("ps_level_sums_noalloc: nv must be 1..8 (its scratch holds the partial sums of 8 variables)");
    const int dev = current()->device;
    double *part = dev >= 0 && dev < PS_MAX_DEVICES ? g_scratch[dev] : nullptr;
    if (!part) return fail("ps_level_sums_noalloc: no scratch address for the current device");
    return level_sums(level, fp32, nv, cells, sums, part);
}

int ps_multi_index(const ps_multi_plan *plan)
{
    PS_NEED_INIT();
    if (!plan || plan->ninst <= 0) return fail("empty multi plan");
    if (plan->total_tiles > 0) {
        if (!plan->owner || !plan->prefix) return fail("multi plan without swept owner table");
        hipLaunchKernelGGL(ps_owner_kernel, dim3(plan->ninst), dim3(256), 0, stream(),
                           static_cast<const int *>(plan->prefix), plan->ninst, static_cast<int *>(plan->owner));
        PS_HIP(hipGetLastError());
    }
    if (plan->total_stiles > 0) {
        if (!plan->sowner || !plan->sprefix) return fail("multi plan without standard owner table");
        hipLaunchKernelGGL(ps_owner_kernel, dim3(plan->ninst), dim3(256), 0, stream(),
                           static_cast<const int *>(plan->sprefix), plan->ninst, static_cast<int *>(plan->sowner));
        PS_HIP(hipGetLastError());
    }
    if (plan->total_bwaves > 0) {
        if (!plan->bowner || !plan->bprefix) return fail("multi plan without band owner table");
        hipLaunchKernelGGL(ps_owner_kernel, dim3(plan->ninst), dim3(256), 0, stream(),
                           static_cast<const int *>(plan->bprefix), plan->ninst, static_cast<int *>(plan->bowner));
        PS_HIP(hipGetLastError());
    }
    return 0;
}

int ps_multi_run(const ps_multi_plan *plan, int rounds)
{
    PS_NEED_INIT();
    if (!plan || !plan->insts || !plan->prefix || !plan->advance) return fail("incomplete multi plan");
    if ((plan->total_tiles > 0 && !plan->owner) || (plan->total_stiles > 0 && !plan->sowner) ||
        (plan->total_bwaves > 0 && !plan->bowner))
        return fail("multi plan without owner tables (ps_multi_index)");
    if (plan->total_bwaves > 0 && (!plan->band || !plan->bprefix))
        return fail("multi plan with band work items but no BandMulti kernel");
    if (plan->total_tiles > 0 && (!plan->up || !plan->ybridge || !plan->xbridge || !plan->octahedron || !plan->down))
        return fail("multi plan without phase kernels");
    for (int r = 0; r < rounds; ++r) {
        int rc = multi_round(plan);
        if (rc) return rc;
    }
    return 0;
}

int ps_multi_run_timed(const ps_multi_plan *plan, int rounds, ps_run_stats *stats)
{
    PS_NEED_INIT();
    timer_begin(stats);
    int rc = ps_multi_run(plan, rounds);
    int rc2 = timer_end();
    return rc ? rc : rc2;
}

int ps_graph_begin(void)
{
    PS_NEED_INIT();
    PS_HIP(hipStreamBeginCapture(stream(), hipStreamCaptureModeThreadLocal));
    return 0;
}

int ps_graph_end(ps_graph *graph)
{
    PS_NEED_INIT();
    hipGraph_t g = nullptr;
    PS_HIP(hipStreamEndCapture(stream(), &g));
    hipGraphExec_t ge = nullptr;
    hipError_t e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
    hipGraphDestroy(g);
    if (e != hipSuccess) return fail_hip(e, "hipGraphInstantiate", __LINE__);
    *graph = ge;
    return 0;
}

int ps_graph_launch(ps_graph graph)
{
    PS_NEED_INIT();
    PS_HIP(hipGraphLaunch(static_cast<hipGraphExec_t>(graph), stream()));
    return 0;
}

int ps_graph_destroy(ps_graph graph)
{
    if (graph) PS_HIP(hipGraphExecDestroy(static_cast<hipGraphExec_t>(graph)));
    return 0;
}

}  // extern "C"
