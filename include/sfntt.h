/*
 * sfntt.h -- C ABI of the MI355X (gfx950) small-field NTT / coset-LDE engine.
 *
 * This library replaces the reference's device runtime `src/gpu.rs`
 * (MetalContext, src/gpu.rs:26-417) and the GPU entry points of its variant
 * backends.  Every entry point below names the reference interface it stands
 * in for.  Plain pointers and sizes only; no HIP or torch types.
 *
 * Conventions (follow the reference):
 *  - All values are u32.  M31 values are canonical in [0, 2^31-1); BabyBear
 *    values are the raw Montgomery form (R = 2^32), as returned by
 *    `Field::raw()` (src/field/mod.rs:37-47).
 *  - Forward = evaluation (M31: CT-DIT circle transform; BabyBear: GS-DIF,
 *    natural-order input, bit-reversed output).  Inverse = interpolation with
 *    the 1/n normalisation (src/ntt/m31/cpu_reference.rs, src/ntt/babybear/cpu_reference.rs).
 *  - Batches: SFN_LAYOUT_SOA is the reference's `data[col*n + row]`
 *    (src/ntt/babybear/batch.rs:74); SFN_LAYOUT_ROW_MAJOR is `data[row*batch + col]`.
 *  - Errors mirror NttError (src/ntt/mod.rs:51-69).  Size validation happens
 *    before any GPU work; n == 1 (log_n == 0) is the identity.
 *  - A context is externally synchronised (one host thread at a time), like
 *    the reference's single-threaded backends.  The server object is
 *    internally synchronised; its workers share the context's device tables,
 *    tuning entries and error message with the owning thread, and the library
 *    guards those itself (a table that grows keeps its old buffer until the
 *    context is destroyed, so a launch another thread is issuing stays valid).
 */
#ifndef SFNTT_H
#define SFNTT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SFNTT_ABI_VERSION 1

typedef enum {
    SFN_OK = 0,
    SFN_ERR_INVALID_SIZE = 1,        /* NttError::InvalidSize */
    SFN_ERR_DEVICE_NOT_FOUND = 2,    /* NttError::DeviceNotFound */
    SFN_ERR_BUFFER_ALLOC_FAILED = 3, /* NttError::BufferAllocFailed */
    SFN_ERR_KERNEL_LOAD = 4,         /* NttError::ShaderCompileError */
    SFN_ERR_GPU_EXECUTION = 5,       /* NttError::GpuExecutionError */
    SFN_ERR_CORRECTNESS = 6,         /* NttError::CorrectnessMismatch */
    SFN_ERR_UNSUPPORTED = 7,         /* variant/field/direction combination not offered */
    SFN_ERR_INVALID_ARGUMENT = 8,    /* null pointer, bad enum, bad handle */
    SFN_ERR_BUSY = 9,                /* server: admission refused (queue full / over capacity) */
    SFN_ERR_NOT_READY = 10           /* server: job not finished yet */
} sfn_status;

/* KoalaBear (p = 2^31 - 2^24 + 1) is on the reference's roadmap ("standard
 * NTT", README.md:20) without code; it follows the BabyBear conventions
 * exactly (Montgomery raw values, R = 2^32; forward = GS from natural order to
 * bit-reversed evaluations; inverse scales by 1/n), with two-adicity 24 and the
 * 2^24-th root 3^127 = 0x6ac49f88.  NTT only: the coset LDE is BabyBear's. */
typedef enum { SFN_FIELD_M31 = 0, SFN_FIELD_BABYBEAR = 1, SFN_FIELD_KOALABEAR = 2 } sfn_field;

/* The reference's four variants (src/ntt/mod.rs:19-25) plus SFN_VARIANT_AUTO,
 * which takes the tuned plan.  All variants return identical output. */
typedef enum {
    SFN_CT_DIT_R2 = 0,    /* V1: one launch per layer (MetalCtDitR2 / BbMetalCtDitR2) */
    SFN_CT_GS_R2 = 1,     /* V2: in-place r2, LDS-fused layer groups (MetalCtGsR2 / BbMetalR2) */
    SFN_STOCKHAM_R2 = 2,  /* V3: out-of-place ping-pong r2 passes (MetalStockhamR2) */
    SFN_CT_GS_R4 = 3,     /* V4: radix-4 register blocks (MetalCtGsR4 / BbMetalCtGsR4) */
    SFN_VARIANT_AUTO = 4  /* tuned: radix-2^k register blocks, plan from the autotuner */
} sfn_variant;

typedef enum { SFN_FORWARD = 0, SFN_INVERSE = 1 } sfn_direction;
typedef enum { SFN_LAYOUT_SOA = 0, SFN_LAYOUT_ROW_MAJOR = 1 } sfn_layout;

typedef struct sfn_ctx sfn_ctx;

typedef struct {
    char name[256];
    char gcn_arch[64];
    int compute_units;
    int lds_bytes_per_cu;
    size_t hbm_bytes;
    int clock_khz;
    int mem_clock_khz;
    int mem_bus_width;
} sfn_device_info;

/* ---- context: replaces MetalContext::new / device_info (src/gpu.rs:38-64, :137-150) ---- */
sfn_status sfn_ctx_create(int device, sfn_ctx **out);
void sfn_ctx_destroy(sfn_ctx *ctx);
/* Last error message of ctx ("" when none): a copy owned by the library for the
 * calling thread, valid until that thread calls sfn_last_error again. */
const char *sfn_last_error(const sfn_ctx *ctx);
const char *sfn_status_string(sfn_status s);
sfn_status sfn_device_info_get(sfn_ctx *ctx, sfn_device_info *info);

/* ---- host-pointer synchronous calls ----
 * sfn_ntt replaces `forward_ntt_gpu` / `inverse_ntt_gpu(&[u32], log_n)`
 * (src/ntt/metal_r2.rs:69-173, src/ntt/m31/metal_*.rs, src/ntt/babybear/metal_*.rs)
 * for batch == 1, and `forward_ntt_batch_gpu` / `inverse_ntt_batch_gpu`
 * (src/ntt/babybear/batch.rs:76-180) for batch > 1.  `in` and `out` hold
 * batch << log_n values and may alias.  device_ns (nullable) receives the
 * device time of the launch sequence (first kernel to last kernel), the
 * analogue of the reference's commit->wait time (src/gpu.rs:225-242). */
sfn_status sfn_ntt(sfn_ctx *ctx, sfn_field field, sfn_variant variant, sfn_direction dir,
                   uint32_t log_n, uint32_t batch, sfn_layout layout, const uint32_t *in,
                   uint32_t *out, uint64_t *device_ns);

/* Replaces CosetLdeBatch::execute (src/lde/mod.rs:43-107): per column,
 * iDFT (bit-reversed evaluations over H_n -> coefficients), multiply
 * coefficient j by n^-1 * shift^j, zero-pad to N = n << added_bits, forward
 * DFT (bit-reversed evaluations over shift*H_N).  BabyBear only.
 * shift_raw = 0 selects the reference shift two_adic_generator(log_n +
 * added_bits + 1) (src/lde/mod.rs:59); otherwise it is the Montgomery raw
 * shift (e.g. to_monty(31) for the field generator).  `out` holds
 * batch << (log_n + added_bits) values and must not overlap `in`. */
sfn_status sfn_coset_lde_batch(sfn_ctx *ctx, uint32_t log_n, uint32_t batch, uint32_t added_bits,
                               uint32_t shift_raw, sfn_layout layout, const uint32_t *in,
                               uint32_t *out, uint64_t *device_ns);

/* The same coset LDE with natural row order on both sides: input row i is the
 * evaluation at g_n^i, output row k the evaluation at shift * g_N^k (the
 * order of Plonky3's coset_lde_batch, which the reference converts to with
 * bit_reverse_permutation in tests/ntt_cross_validation.rs:21-36, :64-66).
 * No reference entry point does this; with SFN_LAYOUT_ROW_MAJOR and
 * shift_raw = to_monty(31) it is the row-major matrix LDE of the benchmark the
 * reference targets (benches/ntt_benchmark.rs:442). */
sfn_status sfn_coset_lde_natural(sfn_ctx *ctx, uint32_t log_n, uint32_t batch, uint32_t added_bits,
                                 uint32_t shift_raw, sfn_layout layout, const uint32_t *in,
                                 uint32_t *out, uint64_t *device_ns);

/* Element-wise product (NttBackend::pointwise_mul, src/ntt/mod.rs:46-48):
 * out[i] = a[i] * b[i] in the field; len may be any value >= 0. */
sfn_status sfn_pointwise_mul(sfn_ctx *ctx, sfn_field field, const uint32_t *a, const uint32_t *b,
                             uint32_t *out, size_t len);

/* ---- device-resident tier (no counterpart in the reference: it re-uploads
 * everything per call, src/gpu.rs:256-260).  Pointers are device pointers
 * from sfn_buf_alloc; `stream` is a hipStream_t passed as void* (NULL = the
 * context's own stream).  Calls are asynchronous on that stream. ---- */
sfn_status sfn_buf_alloc(sfn_ctx *ctx, size_t bytes, void **dptr);
sfn_status sfn_buf_free(sfn_ctx *ctx, void *dptr);
sfn_status sfn_upload(sfn_ctx *ctx, void *dst, const void *src, size_t bytes, void *stream);
sfn_status sfn_download(sfn_ctx *ctx, void *dst, const void *src, size_t bytes, void *stream);
sfn_status sfn_synchronize(sfn_ctx *ctx, void *stream);

/* Twiddle cache: replaces TwiddleCache<F> (src/ntt/twiddle_cache.rs:17-82),
 * which keeps only the last log_n on the host.  Tables stay device-resident,
 * keyed by (field, log_n); this call builds them ahead of time. */
sfn_status sfn_twiddles_prepare(sfn_ctx *ctx, sfn_field field, uint32_t log_n);
/* Bytes of device tables the context holds now. */
sfn_status sfn_twiddles_bytes(sfn_ctx *ctx, size_t *bytes);
/* Byte budget of the per-shape tables (four-step twists, LDE factor tables
 * keyed by the shift, M31 block twiddles); 0 restores the default, 4 GiB.
 * Past it the least recently used tables are freed (after waiting for the
 * device), except those an unfinished call or a live plan holds; a table
 * larger than the budget is still admitted alone.  The reference's cache is
 * bounded the same way in spirit: it keeps one log_n
 * (src/ntt/twiddle_cache.rs:29-82). */
sfn_status sfn_twiddles_limit(sfn_ctx *ctx, size_t max_bytes);

sfn_status sfn_ntt_device(sfn_ctx *ctx, sfn_field field, sfn_variant variant, sfn_direction dir,
                          uint32_t log_n, uint32_t batch, sfn_layout layout, const uint32_t *d_in,
                          uint32_t *d_out, void *stream);
sfn_status sfn_coset_lde_batch_device(sfn_ctx *ctx, uint32_t log_n, uint32_t batch,
                                      uint32_t added_bits, uint32_t shift_raw, sfn_layout layout,
                                      const uint32_t *d_in, uint32_t *d_out, void *stream);

sfn_status sfn_coset_lde_natural_device(sfn_ctx *ctx, uint32_t log_n, uint32_t batch,
                                        uint32_t added_bits, uint32_t shift_raw, sfn_layout layout,
                                        const uint32_t *d_in, uint32_t *d_out, void *stream);

/* Allocation-free variant of sfn_coset_lde_natural_device.  No counterpart in
 * the reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  It calls no hipMalloc / hipFree: the intermediate array of the
 * row-major natural-order LDE is a static device array of 64 MiB that the
 * library holds on every GPU it runs on until the process ends, so outputs
 * above 64 MiB (batch << (log_n + added_bits) u32) are refused with
 * SFN_ERR_INVALID_SIZE, and every table the call needs must already be
 * resident: run sfn_coset_lde_natural_device once for the shape first, else
 * SFN_ERR_NOT_READY (also after the table cache evicted them).  Calls that
 * overlap on one GPU (other streams, other contexts, other threads) share that
 * array and race; nothing in it needs clearing between calls (every word is
 * written before it is read).  Same arguments and results as
 * sfn_coset_lde_natural_device. */
sfn_status sfn_coset_lde_natural_device_noalloc(sfn_ctx *ctx, uint32_t log_n, uint32_t batch,
                                                uint32_t added_bits, uint32_t shift_raw,
                                                sfn_layout layout, const uint32_t *d_in,
                                                uint32_t *d_out, void *stream);

/* Device time of the last launch sequence issued by this context on its own
 * stream (hipEvent pair), in nanoseconds; waits for it to finish.  Host-pointer
 * calls are always timed; *_device calls only while sfn_call_timing is on
 * (otherwise SFN_ERR_INVALID_ARGUMENT). */
sfn_status sfn_last_device_ns(sfn_ctx *ctx, uint64_t *ns);

/* Per-call timing of the *_device entry points (no reference counterpart; the
 * reference times its calls on the host).  Off by default: the event pair
 * around each call costs ~8-10 us of device time, which dominates launch-bound
 * calls (one 2^8-point column: 17 us per call with it, 8 us without). */
sfn_status sfn_call_timing(sfn_ctx *ctx, int enable);

/* Per-kernel timing (no reference counterpart).  While enabled, every kernel
 * the context launches on its own stream is bracketed by a hipEvent pair;
 * sfn_pass_times waits for the last launch sequence and returns its kernel
 * durations in launch order, sfn_pass_label names kernel i of that sequence. */
sfn_status sfn_pass_timing(sfn_ctx *ctx, int enable);
sfn_status sfn_pass_times(sfn_ctx *ctx, uint64_t *ns, uint32_t cap, uint32_t *count);
const char *sfn_pass_label(sfn_ctx *ctx, uint32_t index);

/* ---- batched plans: many transforms of uneven sizes and both fields in one
 * launch per pass, captured in a hipGraph (no reference counterpart; the
 * reference runs one column size per call). ---- */
typedef struct {
    sfn_field field;
    sfn_direction dir;
    uint32_t log_n;
    uint32_t batch;          /* columns of this job (SoA, data[col*n+row]) */
    const uint32_t *d_in;    /* device pointers; in may equal out */
    uint32_t *d_out;
} sfn_job_desc;

typedef struct sfn_plan sfn_plan;
sfn_status sfn_plan_create(sfn_ctx *ctx, const sfn_job_desc *jobs, uint32_t count, int use_graph,
                           sfn_plan **out);
sfn_status sfn_plan_launch(sfn_plan *plan, void *stream);
sfn_status sfn_plan_info(sfn_plan *plan, uint32_t *launches, uint32_t *tiles);
void sfn_plan_destroy(sfn_plan *plan);

/* ---- autotuner: measures candidate schedules for (field, log_n, batch,
 * layout) on the device and stores the winner in the context's plan table.
 * The table can be saved/loaded as JSON text. ---- */
/* field SFN_TUNE_LDE tunes the BabyBear coset LDE at 2x instead of an NTT. */
#define SFN_TUNE_LDE 16
sfn_status sfn_autotune(sfn_ctx *ctx, sfn_field field, uint32_t log_n, uint32_t batch,
                        sfn_layout layout, uint32_t *best_log_tile);
sfn_status sfn_tuning_save(sfn_ctx *ctx, char *buf, size_t cap, size_t *len);
sfn_status sfn_tuning_load(sfn_ctx *ctx, const char *json);

/* ---- serving: continuous batching of NTT / LDE jobs (no reference
 * counterpart).  Jobs are admitted by device-memory footprint, run on a pool
 * of streams, and report their own latency.  Thread-safe. ---- */
typedef enum { SFN_JOB_NTT = 0, SFN_JOB_LDE = 1 } sfn_job_kind;

typedef struct {
    sfn_job_kind kind;
    sfn_field field;          /* SFN_JOB_LDE requires SFN_FIELD_BABYBEAR */
    sfn_direction dir;        /* SFN_JOB_NTT only */
    uint32_t log_n;
    uint32_t batch;           /* columns, SoA (data[col*n + row]) */
    uint32_t added_bits;      /* SFN_JOB_LDE only */
    uint32_t shift_raw;       /* SFN_JOB_LDE only; 0 = reference shift */
    int device_ptrs;          /* 0: `in`/`out` are host memory (staged over PCIe);
                                 1: they are device memory that stays resident in HBM */
    const uint32_t *in;       /* batch << log_n values */
    uint32_t *out;            /* NTT: batch << log_n; LDE: batch << (log_n+added_bits); must not overlap `in` */
} sfn_request;

typedef struct {
    size_t mem_budget_bytes;  /* device bytes the server may hold for job buffers
                                 (0 = a quarter of the device's memory, at least 8 GiB) */
    uint32_t max_inflight;    /* worker streams, i.e. launch sequences on the device at once (0 = 6) */
    uint32_t max_queue;       /* admission queue bound (0 = 4096) */
    uint32_t max_group;       /* most queued jobs fused into one launch set (0 = 128; 1 = no grouping) */
    uint32_t max_group_cols;  /* column cap of one fused launch (0 = 512) */
} sfn_server_config;

typedef struct {
    uint64_t submit_ns;   /* host steady clock at submit */
    uint64_t start_ns;    /* when a worker took the job */
    uint64_t end_ns;      /* when its result was complete in `out` */
    uint32_t group_size;  /* jobs fused into the launch that ran it */
    sfn_status status;
} sfn_job_timing;

/* Pinned host memory for request buffers (hipHostMalloc). */
sfn_status sfn_host_alloc(sfn_ctx *ctx, size_t bytes, void **ptr);
sfn_status sfn_host_free(sfn_ctx *ctx, void *ptr);
/* Host steady clock in nanoseconds (the clock of sfn_job_timing). */
uint64_t sfn_now(void);

typedef struct sfn_server sfn_server;
sfn_status sfn_server_create(sfn_ctx *ctx, const sfn_server_config *cfg, sfn_server **out);
sfn_status sfn_server_submit(sfn_server *srv, const sfn_request *req, uint64_t *ticket);
/* poll: SFN_ERR_NOT_READY while running; once a finished job is reported by
 * poll or wait its ticket is released. */
sfn_status sfn_server_poll(sfn_server *srv, uint64_t ticket, sfn_job_timing *timing);
sfn_status sfn_server_wait(sfn_server *srv, uint64_t ticket, sfn_job_timing *timing);
sfn_status sfn_server_drain(sfn_server *srv);
void sfn_server_destroy(sfn_server *srv);

#ifdef __cplusplus
}
#endif
#endif /* SFNTT_H */
