/*
 * fftoc.h -- C ABI of the MI355X (gfx950) batched small complex FFT library
 * (libfftoc.so).
 *
 * The reference (gprabowski/fft_onchip) exposes its FFTs as compile-time
 * device functors run in place on shared memory by a test kernel:
 *   template <typename CT, int Size, int UPB, int FPU> struct tensor_fft_N
 *     { __device__ tensor_fft_N(CT *smem); __device__ void operator()(); }
 *   (includes/tensor_fft_{8,64,128,256,512,4096}.cuh, launched by
 *    testing::fft_tester / run_fft_kernel, includes/testing.cuh:16-92),
 *   template <typename CT, int Size> struct legacy_fft      (radix 8,
 *    includes/legacy8_fft.cuh:11-120) and legacy16_fft      (radix 16,
 *    includes/legacy16_fft.cuh:17-173).
 * A caller binding that path over an FFI would bind one launch per size
 * family: "run the size-N FFT over `batch` contiguous signals in this device
 * buffer".  The entry points below are exactly that, with the reference's
 * data layout (interleaved complex<double>, signal-contiguous, batch-major;
 * includes/config.hpp:6) and its in-place mode (in == out) kept.
 *
 * Conventions for every transform entry point:
 *   in, out   device pointers (hipMalloc'd or equivalent), 16-byte aligned;
 *             in == out is allowed (in-place, as the reference runs); other
 *             overlaps are rejected with FFTOC_ERROR_OVERLAP.
 *   batch     number of signals; 0 is a no-op, < 0 is an error.
 *   n         transform length (power of two; per-family limits below).
 *   flags     OR of the FFTOC_* flag bits below; 0 = forward, FP64, C128.
 *   stream    a hipStream_t (NULL = the legacy default stream); the call is
 *             asynchronous on that stream and never synchronises the device.
 * Errors are returned as fftoc_status codes (the reference's gpuErrchk
 * printed and exit()ed, includes/common.cuh:3-13); nothing here exits.
 * Calls are thread-safe; twiddle tables are built once per device (lazily or
 * by fftoc_init) and are read-only afterwards.
 */
#ifndef FFTOC_H
#define FFTOC_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum fftoc_status {
  FFTOC_SUCCESS = 0,
  FFTOC_ERROR_INVALID_SIZE = 1,   /* n not supported by this entry point   */
  FFTOC_ERROR_INVALID_BATCH = 2,  /* batch < 0                              */
  FFTOC_ERROR_NULL_POINTER = 3,   /* in or out is NULL with batch > 0       */
  FFTOC_ERROR_INVALID_FLAGS = 4,  /* unknown or contradictory flag bits     */
  FFTOC_ERROR_OVERLAP = 5,        /* in and out overlap but are not equal   */
  FFTOC_ERROR_UNSUPPORTED = 6,    /* valid request this path does not offer */
  FFTOC_ERROR_DEVICE = 7,         /* HIP runtime error / no gfx950 device   */
  FFTOC_ERROR_LAUNCH = 8,         /* kernel launch failed                   */
  FFTOC_ERROR_ALLOC = 9,          /* device or host allocation failed       */
  FFTOC_ERROR_MISALIGNED = 10,    /* in/out not 16-byte aligned             */
  FFTOC_ERROR_BUDGET = 11,        /* request exceeds a memory budget        */
  FFTOC_ERROR_QUEUE_FULL = 12,    /* no free request slot                   */
  FFTOC_ERROR_INVALID_HANDLE = 13 /* unknown plan / server / request id     */
} fftoc_status;

/* ---- flag bits ---------------------------------------------------------- */
#define FFTOC_FORWARD 0x0u        /* exp(-2 pi i nk/N) (reference direction) */
#define FFTOC_INVERSE 0x1u        /* exp(+2 pi i nk/N), unnormalised         */
#define FFTOC_NORMALIZE 0x2u      /* multiply the result by 1/N (exact)      */
/* arithmetic of the DFT stages of fftoc_tensor_fft (n = 64, 256, 4096 for
 * the two split forms): data stay complex128 in memory, are rounded to FP32
 * for the GEMM stages, twiddles are the FP64 table rounded to FP32, results
 * are widened to FP64 and FFTOC_NORMALIZE is applied in FP64.  At n = 4096
 * the first 16-point stage is an FP64 butterfly (as in the FP64 default). */
#define FFTOC_PREC_FP64 0x00u     /* FP64 everywhere (default)               */
#define FFTOC_PREC_FP32 0x10u     /* FP32 MFMA stage GEMMs, FP32 accumulate  */
#define FFTOC_PREC_BF16X3 0x20u   /* 3-term split-BF16 MFMA, FP32 accumulate */
#define FFTOC_PREC_MASK 0xF0u
/* data type: complex64 (float2) input and output with FP32 arithmetic and
 * FP32 twiddles rounded from the FP64 table (default: complex128) */
#define FFTOC_DATA_C64 0x100u
/* signals per workgroup (the reference's UPB x FPU): selects the entry
 * point's kernel variant with that tile (FFTOC_ERROR_UNSUPPORTED if the
 * length has none; fftoc_variant_tile lists them).  0 = the tuned choice. */
#define FFTOC_TILE_SHIFT 16
#define FFTOC_TILE_MASK (0xFFu << FFTOC_TILE_SHIFT)
#define FFTOC_TILE(s) (((uint32_t)(s) & 0xFFu) << FFTOC_TILE_SHIFT)
/* at most k workgroups resident per CU (the reference's blocks-per-SM sweep,
 * src/saturation.cu:15-62), k = 1..63: a persistent kernel launches k per CU;
 * a kernel with one workgroup per tile requests enough (unused) LDS that no
 * more than k fit on a CU.  0 = the tuned choice (the defaults cap a few
 * lengths at large batches; fftoc_autotune sweeps the caps; otherwise the
 * kernel's occupancy).  Never changes results. */
#define FFTOC_WGS_PER_CU_SHIFT 24
#define FFTOC_WGS_PER_CU_MASK (0x3Fu << FFTOC_WGS_PER_CU_SHIFT)
#define FFTOC_WGS_PER_CU(k) (((uint32_t)(k) & 0x3Fu) << FFTOC_WGS_PER_CU_SHIFT)

/* ---- library ------------------------------------------------------------ */
/* Builds the FP64 twiddle / DFT-matrix tables on the current device.
 * Optional: every entry point initialises its device on first use. */
int fftoc_init(void);
/* Frees the tables of every device.  Not thread-safe against running calls.
 * FFTOC_ERROR_UNSUPPORTED while plans or servers exist (their kernels read the
 * tables): destroy them first.  Tables from fftoc_twiddle_table are invalid
 * afterwards. */
int fftoc_shutdown(void);
/* Static description of a status code (never NULL). */
const char *fftoc_strerror(int status);
/* Library version as MAJOR*10000 + MINOR*100 + PATCH. */
int fftoc_version(void);
/* Device addresses of the current device's twiddle tables W_4096^j,
 * j = 0..4095 (double2 and float2), for kernels that call the in-LDS
 * functors of include/fftoc_lds.hpp.  Either pointer may be NULL.  Replaces
 * the per-functor twiddle setup of the reference (pow_theta,
 * includes/tensor_fft_64.cuh:33-39; the table of includes/constants.cuh:7-110). */
int fftoc_twiddle_table(const void **fp64, const void **fp32);

/* ---- tensor path: FP64 MFMA Karatsuba DFT stages ------------------------ */
/* Replaces fft::tensor_fft_N<CT,N,UPB,FPU> (includes/tensor_fft_8.cuh:12-66,
 * tensor_fft_16.cuh:9-63, tensor_fft_64.cuh:12-77, tensor_fft_128.cuh:12-101,
 * tensor_fft_256.cuh:10-123, tensor_fft_512.cuh:12-122,
 * tensor_fft_4096.cuh:12-98) as launched by testing::run_fft_kernel
 * (includes/testing.cuh:59-92).  n in {8,16,32,...,4096}.  The UPB/FPU tile
 * template parameters become FFTOC_TILE(s) in flags. */
int fftoc_tensor_fft(const void *in, void *out, int64_t batch, int n,
                     uint32_t flags, void *stream);
int fftoc_tensor_fft_8(const void *in, void *out, int64_t batch, uint32_t flags, void *stream);
int fftoc_tensor_fft_16(const void *in, void *out, int64_t batch, uint32_t flags, void *stream);
int fftoc_tensor_fft_64(const void *in, void *out, int64_t batch, uint32_t flags, void *stream);
int fftoc_tensor_fft_128(const void *in, void *out, int64_t batch, uint32_t flags, void *stream);
int fftoc_tensor_fft_256(const void *in, void *out, int64_t batch, uint32_t flags, void *stream);
int fftoc_tensor_fft_512(const void *in, void *out, int64_t batch, uint32_t flags, void *stream);
int fftoc_tensor_fft_4096(const void *in, void *out, int64_t batch, uint32_t flags, void *stream);

/* ---- legacy radix paths: register butterflies on the vector ALUs -------- */
/* Replaces fft::legacy_fft<CT,N> (radix-8 Stockham, includes/legacy8_fft.cuh:
 * 11-120): n in {8, 64, 512, 4096}.  FP64 only. */
int fftoc_legacy8_fft(const void *in, void *out, int64_t batch, int n,
                      uint32_t flags, void *stream);
/* Replaces fft::legacy16_fft<CT,N> (radix-16 Stockham,
 * includes/legacy16_fft.cuh:17-173): n in {16, 256, 4096}.  FP64 only. */
int fftoc_legacy16_fft(const void *in, void *out, int64_t batch, int n,
                       uint32_t flags, void *stream);

/* ---- best path ----------------------------------------------------------- */
/* Any power of two 8..4096; picks the fastest FP64 kernel for n from the
 * tuning table (the role of the reference's UPB/FPU sweep,
 * scripts/generate_bench.py:93-107). */
int fftoc_fft(const void *in, void *out, int64_t batch, int n, uint32_t flags,
              void *stream);

/* ---- kernel variants and tuning (no counterpart in the reference, whose
 * UPB/FPU choice was a compile-time sweep, scripts/generate_bench.py:93-107,
 * with results kept in scripts/results/configs.csv) ------------------------
 * Each length has one variant list per data type: `flags` selects it by its
 * FFTOC_DATA_C64 bit (complex64 data) and is otherwise ignored by the
 * enumeration calls; a variant index is an index into the list of that data
 * type, as fftoc_fft_variant and FFTOC_TILE use it. */
/* Number of kernel variants for length n and data type (0 if n is
 * unsupported). */
int fftoc_variant_count(int n, uint32_t flags);
/* Stable name of variant v of length n and data type ("tensor:..." or
 * "radix:..."), NULL for an unknown variant. */
const char *fftoc_variant_name(int n, uint32_t flags, int variant);
/* Runs one specific variant (same conventions as fftoc_fft). */
int fftoc_fft_variant(const void *in, void *out, int64_t batch, int n,
                      uint32_t flags, int variant, void *stream);
/* Variant fftoc_fft uses for length n and data type at this batch size (the
 * choice is kept per floor(log2(batch)) bucket), and its override (batch < 0:
 * every bucket). */
int fftoc_get_variant(int n, uint32_t flags, int64_t batch);
int fftoc_set_variant(int n, uint32_t flags, int64_t batch, int variant);
/* Times every variant of length n and data type over `batch` signals, at
 * its occupancy and at each smaller workgroups-per-CU cap up to 16
 * (allocates and frees two device scratch buffers, synchronises `stream`),
 * and makes the fastest (variant, cap) of each entry-point family its
 * default for that batch bucket; a cap must beat the occupancy by 1 %. */
int fftoc_autotune(int n, uint32_t flags, int64_t batch, int reps, void *stream);
/* Median time in microseconds of a variant in the last autotune of its
 * length and data type, at its best cap (<0: none). */
double fftoc_variant_time_us(int n, uint32_t flags, int variant);
/* That best cap (0: the occupancy; -1: not tuned). */
int fftoc_variant_per_cu(int n, uint32_t flags, int variant);
/* Resident workgroups per CU of a variant's kernel on the current device:
 * the upper end of its FFTOC_WGS_PER_CU range (-1 for an unknown variant). */
int fftoc_variant_occupancy(int n, uint32_t flags, int variant);
/* Workgroups-per-CU cap fftoc_fft uses with its chosen variant for length n
 * and data type at this batch size (0: the occupancy), and its override
 * (batch < 0: every bucket).  fftoc_set_variant resets it to 0. */
int fftoc_get_workgroups_per_cu(int n, uint32_t flags, int64_t batch);
int fftoc_set_workgroups_per_cu(int n, uint32_t flags, int64_t batch, int per_cu);
/* Signals per workgroup of a variant (the FFTOC_TILE value that selects it),
 * or -1 for an unknown variant. */
int fftoc_variant_tile(int n, uint32_t flags, int variant);
/* Persisted tuning table, one line per (entry point, length, data type,
 * batch bucket): "<family> <n> <c128|c64> <log2 batch> [wgs=<k>] <variant
 * name>", wgs=<k> for a workgroups-per-CU cap. */
int fftoc_tuning_save(const char *path);
int fftoc_tuning_load(const char *path);

/* ---- mixed-size batches: one persistent launch over many transforms -----
 * No counterpart in the reference, whose harness launches one size per
 * kernel (includes/testing.cuh:59-92).  Each job is one call's worth of work
 * (same conventions as fftoc_fft; flags: FFTOC_INVERSE, FFTOC_NORMALIZE,
 * FFTOC_DATA_C64).  A plan fixes the buffers; executing it transforms their
 * current contents.  A job's output must not overlap any other job's input
 * or output (FFTOC_ERROR_OVERLAP); a job may run in place.  Executions of one
 * plan must not overlap (they share the plan's work counter); different
 * plans are independent. */
typedef struct fftoc_job {
  const void *in;
  void *out;
  int64_t batch;
  int n;
  uint32_t flags;
} fftoc_job;
typedef struct fftoc_plan_t *fftoc_plan;

/* Validates the jobs and uploads their descriptors (allocates device memory). */
int fftoc_plan_create(const fftoc_job *jobs, int njobs, fftoc_plan *plan);
/* One persistent kernel launch over every job, asynchronous on `stream`. */
int fftoc_plan_execute(fftoc_plan plan, void *stream);
/* The same launch replayed from a hipGraph captured on first use. */
int fftoc_plan_execute_graph(fftoc_plan plan, void *stream);
/* Times the candidate tile routines of every (length, data type) in the plan
 * on its own buffers -- `reps` back-to-back launches each, after one warm-up
 * launch -- and keeps the fastest (overwrites the outputs); then does the
 * same for the workgroups per CU of the launch (1 up to the kernel's
 * occupancy).  A plan whose jobs share one routine, data type and direction
 * runs a kernel specialized to them, and its candidates include the
 * 512-thread routines. */
int fftoc_plan_autotune(fftoc_plan plan, int reps, void *stream);
/* Workgroups per CU of the plan's launch (0 = the default: two, or the
 * kernel's occupancy if lower), and its override, k = 0..63 (FFTOC_ERROR_
 * INVALID_FLAGS outside; values above the occupancy act as the occupancy). */
int fftoc_plan_workgroups_per_cu(fftoc_plan plan);
int fftoc_plan_set_workgroups_per_cu(fftoc_plan plan, int per_cu);
/* Name of the tile routine the plan uses for (n, data type of flags). */
int fftoc_plan_routine(fftoc_plan plan, int n, uint32_t flags, const char **name);
/* Pins the plan's tile routine for (n, data type of flags) to the named one
 * (FFTOC_ERROR_UNSUPPORTED for a name that is not a routine of length n), as
 * a persisted choice of fftoc_plan_autotune would be restored.  The
 * reference fixes one configuration per compiled kernel
 * (scripts/generate_bench.py:93-107); here the choice is made at run time. */
int fftoc_plan_set_routine(fftoc_plan plan, int n, uint32_t flags, const char *name);
/* index-th tile routine of length n, in table order (FFTOC_ERROR_INVALID_HANDLE
 * past the last one). */
int fftoc_batch_routine_name(int n, int index, const char **name);
int fftoc_plan_destroy(fftoc_plan plan);

/* ---- allocation-free batch ------------------------------------------------
 * No counterpart in the reference (nor in the plan calls above, which allocate
 * at create).  For callers that cannot allocate device memory: code under
 * hipGraph stream capture, tight serving loops, harnesses that forbid
 * hipMalloc.  One launch over `njobs` jobs, with the same validation, results
 * and ordering as fftoc_plan_create + fftoc_plan_execute of the same jobs
 * (untuned tile routines), asynchronous on `stream`.  It calls no allocating
 * routine: the job descriptors and the work counter live in a fixed 4 KiB
 * scratch array per device that the library holds until the process ends
 * (the device's first use of the library locates it; twiddle tables are
 * module globals too).  At most FFTOC_NOALLOC_MAX_JOBS jobs per call
 * (FFTOC_ERROR_BUDGET beyond).  Every call on a device shares that device's
 * array: calls that overlap in time on one device race and give wrong results
 * -- order them (one stream, or events). */
#define FFTOC_NOALLOC_MAX_JOBS 32
int fftoc_batch_fft_noalloc(const fftoc_job *jobs, int njobs, void *stream);

/* ---- serving: device-resident request loop with HBM-budget admission -----
 * No counterpart in the reference.  fftoc_server_create launches a
 * persistent kernel that serves requests posted through a ring in pinned
 * host memory until fftoc_server_destroy.  It seats as many workgroups as fit
 * (two per CU): the first on each CU always takes work, the second joins only
 * while a backlog waits.  The environment variable
 * FFTOC_SERVER_WORKERS_PER_CU=1 keeps one workgroup per CU, which lowers the
 * latency at light load and the throughput at saturation.  A request is
 * admitted when a ring slot is free and its output fits in the free part of
 * the HBM budget (the server's output pool); requests that arrive while others
 * run start on idle workgroups at once.
 * The loop leaves after an idle timeout without requests (20 ms by default,
 * fftoc_server_set_idle_timeout) and is relaunched by the next submit or
 * poll.  While it runs, work that waits for the whole device waits for it to
 * go idle and leave: hipDeviceSynchronize(), and any stream HIP places on the
 * loop's hardware queue (HIP maps streams onto GPU_MAX_HW_QUEUES queues, 4 by
 * default), including synchronous hipMemcpy on the default stream.  Such a
 * wait therefore lasts until the idle timeout after the last request. */
typedef struct fftoc_server_t *fftoc_server;

/* hbm_budget_bytes: size of the output pool (rounded down to a multiple of
 * 256 bytes, the pool's granule); max_inflight: ring slots. */
int fftoc_server_create(size_t hbm_budget_bytes, int max_inflight, fftoc_server *server);
/* Admits one request (flags: FFTOC_INVERSE, FFTOC_NORMALIZE, FFTOC_DATA_C64).
 * FFTOC_ERROR_QUEUE_FULL: no slot or budget free now (retry later);
 * FFTOC_ERROR_BUDGET: the output can never fit the budget. */
int fftoc_server_submit(fftoc_server server, const void *in, int64_t batch, int n,
                        uint32_t flags, uint64_t *request_id);
/* *done = 1 and *out = device pointer of the result once complete. */
int fftoc_server_poll(fftoc_server server, uint64_t request_id, int *done, void **out);
/* Returns a completed request's output memory and ring slot. */
int fftoc_server_release(fftoc_server server, uint64_t request_id);
/* Device clock (100 MHz ticks) at which a completed request finished. */
int fftoc_server_gpu_time(fftoc_server server, uint64_t request_id, uint64_t *ticks);
/* Replays a trace: request i (in, batch, n, flags of reqs[i]; out ignored)
 * arrives arrival_s[i] seconds after the call; admission in arrival order;
 * latency_s[i] = completion seen by the host - arrival (NaN if not served);
 * *span_s = wall time.  If admission stays refused for 1 s while none of the
 * trace's requests is in flight (slots or budget held by requests submitted
 * outside the trace), returns FFTOC_ERROR_QUEUE_FULL.  On any error the
 * requests the replay admitted are completed and released before it returns. */
int fftoc_server_replay(fftoc_server server, const fftoc_job *reqs, int count,
                        const double *arrival_s, double *latency_s, double *span_s);
int fftoc_server_stats(fftoc_server server, size_t *budget_used, int *inflight);
/* Idle time after which the device loop leaves (1e-4 .. 60 s); takes effect
 * from the loop's next launch. */
int fftoc_server_set_idle_timeout(fftoc_server server, double seconds);
/* Stops the loop (waits up to 10 s for it to leave) and frees everything.  A
 * loop that has not left by then keeps its memory (it is not freed under a
 * running kernel) and FFTOC_ERROR_DEVICE is returned. */
int fftoc_server_destroy(fftoc_server server);

#ifdef __cplusplus
}
#endif

#endif /* FFTOC_H */
