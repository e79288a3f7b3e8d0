/*
 * pysweep_hip.h -- C ABI of libpysweep_hip.so, the MI355X (gfx950) runtime
 * under pysweep_amd.
 *
 * The reference drives its GPU through PyCUDA.  Each entry point below
 * replaces one PyCUDA call site of the reference (cited as file:line into
 * the reference tree) or one host loop of its solver.
 *
 * Conventions
 *   - every entry returns int: 0 on success, nonzero on failure; the
 *     message of the last failure on the calling thread is ps_last_error();
 *   - no entry throws across the ABI;
 *   - host buffers are caller-owned and only borrowed for the call;
 *   - device buffers and modules are owned by the caller through the
 *     handles returned here and released with ps_free / ps_module_unload;
 *   - device work is ordered on the work stream of the calling thread's
 *     current context: the one made current by ps_context_push, else the
 *     process context of ps_init.  Each context has its own stream (default
 *     priority) and launch timer.  One context must not be used by two
 *     threads at once; distinct contexts may be driven from distinct threads
 *     (a server solving several problems at once).
 */
#ifndef PYSWEEP_HIP_H
#define PYSWEEP_HIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef void *ps_module;   /* loaded code object (hipModule_t)   */
typedef void *ps_function; /* kernel handle (hipFunction_t)       */
typedef void *ps_event;    /* timing event (hipEvent_t)           */
typedef void *ps_graph;    /* instantiated graph (hipGraphExec_t) */
typedef void *ps_context;  /* device + work stream + launch timer */

/* ABI version of this header. */
int ps_version(void);

/* Message of the last failure on this thread ("" if none). */
const char *ps_last_error(void);

/* Number of visible GPUs; 0 without a GPU (never an error). */
int ps_device_count(int *count);

/* Select the device and create the work stream.
 * Replaces cuda.init(); cuda.Device(rank).make_context() (core/block.py:36-38). */
int ps_init(int device);

/* Release the stream.  Replaces cuda_context.pop() (core/process.py:31-32). */
int ps_shutdown(void);

/* Contexts of their own: a device and a work stream per handle, made current
 * per thread.  Replace cuda.Device(rank).make_context() / context.push() /
 * context.pop() (core/block.py:36-38; core/process.py:31-32), which the
 * reference holds one per MPI rank.  ps_context_create does not make the
 * context current; ps_context_push does (on the calling thread, over any
 * context already current there) and ps_context_pop restores the previous
 * one.  Modules loaded while a context is current belong to its device.
 * ps_context_destroy drains the stream first and fails for a context current
 * on the calling thread. */
int ps_context_create(int device, ps_context *ctx);
int ps_context_destroy(ps_context ctx);
int ps_context_push(ps_context ctx);
int ps_context_pop(void);
/* The context current on the calling thread, NULL when it uses the process
 * context. */
int ps_context_current(ps_context *ctx);
/* Device of the calling thread's current context (a second context for
 * its copies is made on the same one). */
int ps_context_device(int *device);

/* Compile HIP source text for gfx950 with hiprtc; no device is needed.
 * On success *code (size *size) is a code object the caller frees with
 * ps_free_host; *log (may be NULL on success) is the compiler log, also
 * freed with ps_free_host.  Replaces SourceModule(...) in
 * io.buildGPUSource (core/io.py:146-155). */
int ps_compile(const char *source, const char *name, const char *const *options, int noptions,
               char **code, size_t *size, char **log);
void ps_free_host(void *p);

/* Version of the hiprtc library that ps_compile uses (part of the key under
 * which compiled code objects are cached; no device is needed). */
int ps_compiler_version(int *major, int *minor);

/* Load / unload a code object on the current device. */
int ps_module_load(const void *code, size_t size, ps_module *module);
int ps_module_unload(ps_module module);

/* Device address and size of a __device__ / __constant__ global.
 * Replaces source_mod.get_global(name) (core/io.py:178, equations/heat.py:43). */
int ps_module_get_global(ps_module module, const char *name, void **dptr, size_t *bytes);

/* Kernel handle.  Replaces source_mod.get_function(name) (core/block.py:92-96, 213). */
int ps_module_get_function(ps_module module, const char *name, ps_function *function);

/* Registers, LDS bytes, max threads per block of a kernel (for the autotuner). */
int ps_function_info(ps_function function, int *vgprs_or_regs, int *lds_bytes, int *max_threads);

/* Device memory.  Replaces cuda.mem_alloc (core/block.py:99-104). */
int ps_malloc(void **dptr, size_t bytes);
int ps_free(void *dptr);

/* Copies (synchronous with respect to the host).
 * Replace cuda.memcpy_htod / memcpy_dtoh (core/functions.py:22, 32; core/io.py:184). */
int ps_memcpy_htod(void *dst, const void *src, size_t bytes);
int ps_memcpy_dtoh(void *dst, const void *src, size_t bytes);
int ps_memcpy_dtod(void *dst, const void *src, size_t bytes);
/* Fill bytes (stream ordered). */
int ps_memset(void *dst, int value, size_t bytes);

/* Page-locked host memory: the host side of the asynchronous copies below
 * (no counterpart in the reference, whose copies all wait). */
int ps_host_malloc(void **ptr, size_t bytes);
int ps_host_free(void *ptr);

/* Copies ordered on the work stream that return without waiting for it.
 * The host buffer must come from ps_host_malloc and stay untouched until an
 * event recorded after the copy has completed (ps_event_query /
 * ps_event_synchronize).  They let a serving loop fetch a finished
 * problem's result, or stage the next one's initial condition, while the
 * rounds queued before them still run (no counterpart in the reference). */
int ps_memcpy_htod_async(void *dst, const void *src, size_t bytes);
int ps_memcpy_dtoh_async(void *dst, const void *src, size_t bytes);
int ps_memcpy_dtod_async(void *dst, const void *src, size_t bytes);

/* Launch a kernel on the work stream.  args[i] points at the i-th kernel
 * argument.  Replaces the PyCUDA Function call in Geometry.callGPU /
 * callStandardGPU (core/geometry.py:38-46). */
int ps_launch(ps_function function, unsigned gx, unsigned gy, unsigned gz, unsigned bx, unsigned by,
              unsigned bz, unsigned shared_bytes, void **args);

/* Wait for the current context's work stream.  Replaces cuda.Context.synchronize()
 * (core/functions.py:31, 60, 84, 142). */
int ps_synchronize(void);

/* Timing events on the work stream. */
int ps_event_create(ps_event *event);
int ps_event_record(ps_event event);
int ps_event_elapsed_ms(ps_event start, ps_event stop, float *ms);
/* Block the host until the work recorded before the event is done (bounds
 * how far a host loop runs ahead of the GPU). */
int ps_event_synchronize(ps_event event);
/* *done = 1 when the work recorded before the event has completed, else 0;
 * never blocks. */
int ps_event_query(ps_event event, int *done);
int ps_event_destroy(ps_event event);
/* The current context's work stream waits, on the device, for the work
 * recorded before `event` on any stream of the same device (another
 * context's); the host does not wait.  Lets a second context's stream
 * carry copies beside the work stream's kernels, each kernel that needs a
 * copy waiting for that copy alone (no counterpart in the reference, whose
 * copies and launches share one stream). */
int ps_stream_wait_event(ps_event event);

/* ---------------------------------------------------------------------- */
/* Native phase sequencing.                                                */
/* ---------------------------------------------------------------------- */

/* Geometry argument every phase / standard kernel receives by value. */
typedef struct ps_grid {
    int nx, ny;   /* periodic domain (rows x, columns y; y contiguous)   */
    int T;        /* ring levels: absolute level A lives in slot A % T    */
    int xoff;     /* x lattice offset of this cycle (0 or b/2)            */
    int tiles_x;  /* tile rows                                            */
    int tiles_y;  /* tile columns                                         */
    int keep_from; /* swept: levels >= keep_from are written in full; below
                      it a phase skips the cells no later phase reads      */
} ps_grid;

/* One swept solve on a device-resident ring state[T][nv][nx][ny].
 * Phase kernels have the reference's names and leading arguments
 * K(double *state, int globalTimeStep, int sweptStep, ps_grid g)
 * (core/pysweep.cu:64-249), sweptStep being the absolute level read by the
 * first sub-step. */
typedef struct ps_swept_plan {
    void *state;
    int nx, ny;
    int b;          /* block size                           */
    int ops;        /* operating points k                   */
    int its;        /* intermediate steps s                 */
    int T;          /* 2*P + s                              */
    int threads;    /* threads per workgroup                */
    int store_all;  /* 1: write every computed cell to the ring (every level
                       complete after its cycle, e.g. to record history);
                       0: write only the cells a later phase reads and the
                       final s levels                                      */
    int tile_rows;  /* tile rows from x = 0 (plus the cycle's offset) that
                       the GPU computes; 0: all (the rest of the rows are a
                       host's share, core/process.py:138-176)              */
    ps_function up, ybridge, xbridge, octahedron, down;
} ps_swept_plan;

/* Run cycles [c_begin, c_end) of the swept schedule with G UpPrism cycles:
 * cycle 0 = FirstPrism (UpPyramid + YBridge), 1..G = UpPrism (XBridge,
 * Octahedron, YBridge), G+1 = LastPrism (XBridge + DownPyramid).
 * Replaces Solver.sweptSolve and functions.FirstPrism/UpPrism/LastPrism
 * (core/solver.py:158-176; core/functions.py:12-88); the window roll and
 * shift (functions.py:90-123, io.py:134-144) become the lattice offset and
 * ring indexing. */
int ps_swept_run(const ps_swept_plan *plan, int G, int c_begin, int c_end);

/* Standard decomposition: one launch per stage. */
typedef struct ps_standard_plan {
    void *state;
    int nx, ny;
    int its;
    int T;          /* ring levels, at least its + 1        */
    int bx, by;     /* tile shape                           */
    int threads;
    ps_function standard;
} ps_standard_plan;

/* Run stages gts = gts_begin .. gts_end-1; stage gts reads absolute level
 * its-2+gts and writes its-1+gts.  Replaces Solver.standardSolve and
 * functions.StandardFunction / sendEdges (core/solver.py:178-189;
 * core/functions.py:126-161). */
int ps_standard_run(const ps_standard_plan *plan, int gts_begin, int gts_end);

/* Per-kernel timing of a run: HIP events around launches on the work
 * stream, resolved after the run (no counterpart in the reference, whose
 * only timing is wall clock around the whole solve, core/solver.py:13-68).
 * Index: PS_K_* below; ms are summed over the timed launches, which are
 * counted in `launches` (both accumulate).  `kinds` (a bit mask of
 * 1 << PS_K_*, 0 = every kind) and `every` (time one launch in `every` of
 * each timed kind, counted per run call; 0 or 1 = each) sample the
 * launches, so that timing a long run adds few event records to it. */
enum {
    PS_K_UP = 0,
    PS_K_YBRIDGE = 1,
    PS_K_XBRIDGE = 2,
    PS_K_OCTAHEDRON = 3,
    PS_K_DOWN = 4,
    PS_K_STANDARD = 5,
    PS_K_BAND = 6
};
#define PS_K_COUNT 7
typedef struct ps_run_stats {
    double ms[PS_K_COUNT];
    long long launches[PS_K_COUNT];
    int kinds;
    int every;
} ps_run_stats;
int ps_swept_run_timed(const ps_swept_plan *plan, int G, int c_begin, int c_end, ps_run_stats *stats);
int ps_standard_run_timed(const ps_standard_plan *plan, int gts_begin, int gts_end, ps_run_stats *stats);

/* Band solve: the swept solve of a plugin that provides a point update
 * (PS_HAS_POINT: single variable, radius-1 cross stencil, e.g. heat) or a
 * band sweep of its own (PS_HAS_BAND, e.g. Euler).  Each launch of the
 * `Band` kernel advances every cell `levels` levels (the last launch fewer)
 * with all intermediate levels in registers, one HBM read and one HBM write
 * per cell: bands of columns, each swept along x by one wavefront, whose
 * sloped edges are recomputed from a halo of k columns per level instead of
 * being exchanged through bridge phases.  Kernel
 * Band(double *state, int globalTimeStep, int level, ps_band_grid g).
 * Replaces Solver.sweptSolve with functions.FirstPrism/UpPrism/LastPrism
 * (core/solver.py:158-176; core/functions.py:12-88) for such plugins: the
 * same stages on the same ring, with results bitwise those of the phase
 * kernels and the standard kernel. */
typedef struct ps_band_grid {
    int nx, ny;
    int T;          /* ring levels                                         */
    int hh;         /* levels of this launch                               */
    int wlo;        /* levels hh' >= wlo of the launch are written         */
    int rows;       /* output rows per wavefront                           */
    int bands;      /* ceil(ny / (128 - 2 * levels))                       */
    int segs;       /* ceil(nx / rows)                                     */
} ps_band_grid;

typedef struct ps_band_plan {
    void *state;    /* ring state[T][nx][ny]                               */
    int nx, ny;
    int its;        /* intermediate steps s (1 or 2)                       */
    int T;          /* at least levels + 1 (a launch reads level A only)   */
    int levels;     /* levels per launch compiled into the kernel (even,
                       a multiple of s, 2..16)                              */
    int wout;       /* output columns per band (the kernel's band width less
                       its halo columns: 128 - 2*levels for a point-update
                       plugin, 64 - 2*k*levels for a plugin sweep)          */
    int rows;       /* output rows per wavefront (> 0)                     */
    int store_all;  /* 1: write every level; 0: the last s of each launch  */
    ps_function band;
} ps_band_plan;

/* Run stages from absolute level A (which must be complete in the ring and
 * be a full-step level) for `stages` stages, in launches of plan->levels
 * stages (the last one fewer); the last s levels reached are complete in the
 * ring afterwards, and every level when store_all is set. */
int ps_band_run(const ps_band_plan *plan, int A, int stages);
int ps_band_run_timed(const ps_band_plan *plan, int A, int stages, ps_run_stats *stats);

/* ---------------------------------------------------------------------- */
/* Instance tables: many problems of one plugin class per launch.          */
/* (No counterpart in the reference, which solves one problem per run.)    */
/* ---------------------------------------------------------------------- */

/* One problem (layout shared with the kernels). */
typedef struct ps_inst {
    void *ring;          /* state ring [T][nv][nx][ny] of the class's type  */
    const void *consts;  /* the plugin's constants for this problem         */
    int nx, ny, T;
    int tiles_x, tiles_y;    /* swept tiles (b x b)                          */
    int stiles_x, stiles_y;  /* standard tiles                               */
    int G;               /* UpPrism cycles, -1 if none                       */
    int tail;            /* standard stages after LastPrism                  */
    int c;               /* round counter: 0 FirstPrism, 1..G UpPrism,        */
                         /* G+1 LastPrism, G+2.. tail stages                 */
    int active;
    int brows, bands, bsegs; /* band-method problems (G = -1, stages on       */
                         /* BandMulti): output rows per wavefront, bands     */
                         /* across y, row segments; else unused              */
} ps_inst;

typedef struct ps_multi_plan {
    void *insts;        /* device ps_inst[ninst]                            */
    void *prefix;       /* device int[ninst+1]: prefix sums of swept tiles  */
    void *sprefix;      /* device int[ninst+1]: prefix sums of std tiles    */
    void *owner;        /* device int[total_tiles]: problem of each tile    */
    void *sowner;       /* device int[total_stiles]                         */
    int ninst;
    int total_tiles, total_stiles;  /* work items: one b x b swept tile, or one
                                       standard tile, each; one workgroup each */
    int threads, sthreads;
    ps_function up, ybridge, xbridge, octahedron, down, standard, advance;
    void *bprefix;      /* device int[ninst+1]: prefix sums of band         */
                        /* wavefronts (bands x bsegs of each band-method    */
                        /* problem), or null                                */
    void *bowner;       /* device int[total_bwaves]                         */
    int total_bwaves;   /* band work items: one wavefront (a workgroup) each */
    ps_function band;   /* BandMulti of the class's band module, or null    */
} ps_multi_plan;

/* Fill the owner tables of a plan from its prefix sums on the work stream
 * (owner[t] = i for prefix[i] <= t < prefix[i+1]). */
int ps_multi_index(const ps_multi_plan *plan);

/* Run `rounds` rounds: each round launches UpPyramidMulti, XBridgeMulti,
 * OctahedronMulti, DownPyramidMulti, YBridgeMulti, StandardMulti, BandMulti
 * and ps_advance; every problem executes the phases of its own cycle c (a
 * band-method problem one whole step on the rounds whose stage begins one,
 * nothing on the others). */
int ps_multi_run(const ps_multi_plan *plan, int rounds);
int ps_multi_run_timed(const ps_multi_plan *plan, int rounds, ps_run_stats *stats);

/* Device-to-device copy of n doubles by a plain copy kernel on the work
 * stream (ring resets, and the reference copy of known size that calibrates
 * HBM byte counters). */
int ps_copy_f64(void *dst, const void *src, size_t n);

/* FP64 sums of each variable of one level level[nv][cells] (FP64 storage,
 * or FP32 when fp32 != 0), stream ordered: sums[v] = sum of variable v and
 * sums[nv + v] = sum of its magnitudes, for v < nv <= 8, into device memory
 * sums[2 * nv].  The conserved integrals (mass, momentum, energy; the heat
 * content) of the FP32-state mode are summed in FP64 here.  The order of
 * the additions is fixed, so a level gives the same bits on every call.
 * Partial sums go to a buffer the context allocates on first use.  No
 * counterpart in the reference, which checks no conserved quantity. */
int ps_level_sums(const void *level, int fp32, int nv, size_t cells, double *sums);

/* ps_level_sums without any allocation: the partial sums go to a fixed
 * device array of the library (2 x 8 variables x 512 blocks of doubles,
 * 64 KiB per GPU, held until the process ends) instead of a buffer the
 * context allocates; nv above 8 is an error.  Same results, bit for bit.
 * No counterpart in the reference: it is for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops).  All
 * calls on one GPU share the array: calls that overlap on one GPU (two
 * threads, or two contexts' streams) race. */
int ps_level_sums_noalloc(const void *level, int fp32, int nv, size_t cells, double *sums);

/* ---------------------------------------------------------------------- */
/* Graph capture of the work stream (no counterpart in the reference).     */
/* ---------------------------------------------------------------------- */
int ps_graph_begin(void);
int ps_graph_end(ps_graph *graph);
int ps_graph_launch(ps_graph graph);
int ps_graph_destroy(ps_graph graph);

#ifdef __cplusplus
}
#endif

#endif /* PYSWEEP_HIP_H */
