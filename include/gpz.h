/* gpz.h -- C ABI of the MI355X-native GPZ particle compressor.
 *
 * Plain C: fixed-width integers, device pointers as void*, hipStream_t as an
 * opaque pointer.  Every call is asynchronous on the given stream unless it
 * says otherwise; no call allocates device memory (the caller passes a
 * workspace sized by the *_workspace_bytes queries).  Errors come back as a
 * negative status instead of the reference's process exit
 * (include/utils.cuh:37-42, src/cudaWarpCompress.cu:1547-1551).
 *
 * The byte stream is the reference's, byte for byte: header (24 B FP32,
 * 32 B FP64), u64 payload offsets[G], per-group meta[G] (36 B / 64 B), and
 * the concatenated group payloads, G = ceil(N / 1024)
 * (src/cudaWarpCompress.cu:1668-1702, src/cudaWarpCompressDouble.cu:1667-1702).
 */
#ifndef GPZ_H_
#define GPZ_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ihipStream_t* gpz_stream_t; /* == hipStream_t */

enum {
  GPZ_OK = 0,
  GPZ_ERR_INVALID_ARG = -1,   /* bad dim / bits / n / error bound / mode   */
  GPZ_ERR_WORKSPACE = -2,     /* workspace missing or too small           */
  GPZ_ERR_CAPACITY = -3,      /* output buffer smaller than the bound     */
  GPZ_ERR_CORRUPT = -4,       /* stream header inconsistent               */
  GPZ_ERR_SEG_TOO_LARGE = -5, /* auto segment size would exceed 1024      */
  GPZ_ERR_HIP = -6            /* a HIP runtime call failed                */
};

enum { GPZ_MODE_ABS = 0, GPZ_MODE_REL = 1 };

/* Compression parameters: the fields of the reference's CudaConfig
 * (include/cudaCompress.cuh:16-40) that the compressor reads. */
typedef struct gpz_params {
  uint32_t dim;        /* 2 or 3                                            */
  uint32_t float_bits; /* 32 or 64                                          */
  uint64_t n;          /* number of points (>= 1)                           */
  double eb;           /* absolute error bound (FP32: used as (float)eb)    */
  double rel_eb;       /* relative bound as given (drives the auto seg)     */
  uint32_t mode;       /* GPZ_MODE_ABS or GPZ_MODE_REL                      */
  uint32_t seg_size;   /* 0 = auto (rel mode only)                          */
} gpz_params;

/* Parsed stream header (src/CompressorManager.cpp:24-34 reads dim, N,
 * float_bits; src/cudaWarpCompress.cu:1802-1810 reads eb and seg). */
typedef struct gpz_header {
  uint32_t dim;
  uint32_t float_bits;
  uint64_t n;
  double eb;
  uint32_t seg_size;
  uint32_t header_bytes; /* 24 or 32                                        */
  uint64_t groups;       /* ceil(n / 1024)                                  */
  uint64_t min_stream_bytes; /* header + offsets + meta                     */
} gpz_header;

/* Result of a compression, written to device memory by gpz_compress and
 * read back with gpz_compress_result.  seg_size/adjusted mirror the
 * CudaConfig fields computeSegSize updates (src/cudaWarpCompress.cu:1513,
 * :1552-1553). */
typedef struct gpz_result {
  uint64_t stream_bytes;
  uint32_t seg_size;
  uint32_t adjusted;
  int32_t status;
  uint32_t reserved;
} gpz_result;

/* Library version and a self-description of the build target. */
int gpz_version(void);

/* Upper bound of the stream size for these parameters. */
uint64_t gpz_stream_bound(const gpz_params* p);

/* Workspace the compressor needs (replaces the caller scratch of 4*N*s
 * bytes, src/main.cpp:289-291, and the per-call cudaMallocs of
 * src/cudaWarpCompress.cu:1628-1634,1658-1666). */
uint64_t gpz_compress_workspace_bytes(const gpz_params* p);

/* Compress SoA planes d_x, d_y[, d_z] (device, N elements each of
 * float/double; d_z ignored for dim 2) into d_stream (device, >= cap bytes).
 * Replaces ICudaCompressor::compress / CudaWarpCompress{,Double}::compress
 * (include/ICudaCompressor.cuh:10-11, src/cudaWarpCompress.cu:1559-1563,
 * src/cudaWarpCompressDouble.cu:1557-1561).  Unlike the reference, the stream
 * does not overwrite the input: d_stream must not alias the planes.
 * Asynchronous; the size, the segment size used and a device-side status are
 * written into the workspace (read them with gpz_compress_result). */
int gpz_compress(const gpz_params* p, const void* d_x, const void* d_y,
                 const void* d_z, void* d_stream, uint64_t cap, void* d_ws,
                 uint64_t ws_bytes, gpz_stream_t stream);

/* Copy the gpz_result of the last gpz_compress on this workspace to the host
 * (synchronises the stream).  Mirrors the D2H size read of
 * src/cudaWarpCompress.cu:1697-1702. */
int gpz_compress_result(const void* d_ws, gpz_result* out, gpz_stream_t stream);

/* Parse a stream header from host memory.  Mirrors
 * CompressorManager::decompressLoadParas (src/CompressorManager.cpp:24-34). */
int gpz_parse_header(const void* h_stream, uint64_t bytes, gpz_header* out);

/* Read and parse the header of a device-resident stream (synchronises the
 * stream; the reference does the same four D2H reads,
 * src/cudaWarpCompress.cu:1807-1810). */
int gpz_read_header(const void* d_stream, gpz_header* out, gpz_stream_t stream);

/* Total size of a device-resident stream, from its header, the last
 * group's offset and meta (synchronises the stream). */
int gpz_stream_bytes(const gpz_header* h, const void* d_stream,
                     uint64_t* bytes, gpz_stream_t stream);

/* Decompress a device stream into SoA planes d_x, d_y[, d_z] (N elements
 * each; d_z untouched for dim 2).  Point 1024*c + r receives the point of
 * sorted rank r of group c, as in the reference decoder
 * (src/cudaWarpCompress.cu:971-1007).  Replaces ICudaCompressor::decompress
 * (include/ICudaCompressor.cuh:13-14, src/cudaWarpCompress.cu:1565-1571).
 * Asynchronous. */
int gpz_decompress(const gpz_header* h, const void* d_stream,
                   uint64_t stream_bytes, void* d_x, void* d_y, void* d_z,
                   gpz_stream_t stream);

/* Full-data per-axis extrema {min_x, min_y, min_z, max_x, max_y, max_z} of
 * device planes, written to host memory mm (float or double per float_bits;
 * z extent 0 in 2D).  Replaces the host findMinimumValues
 * (include/utils.cuh:44-65) used for rel -> abs; synchronises the stream.
 * d_ws needs gpz_min_max_workspace_bytes(). */
uint64_t gpz_min_max_workspace_bytes(void);
int gpz_min_max(uint32_t dim, uint32_t float_bits, uint64_t n,
                const void* d_x, const void* d_y, const void* d_z, void* d_ws,
                void* mm, gpz_stream_t stream);

/* rel -> abs error bound exactly as src/main.cpp:35-55 (eb = max_range * e,
 * in float for FP32 and in double for FP64). */
double gpz_rel_to_abs(uint32_t float_bits, double e, const void* mm);

/* Per-kernel timing for benchmarks: while enabled, each kernel launch is
 * bracketed by hipEvents on its stream.  gpz_timing_read returns the summed
 * elapsed milliseconds and launch count of one kernel ("seg_sample",
 * "seg_decide", "encode", "scan", "compact", "decode"), waiting for pending
 * events.  The reference marks the same regions with NVTX ranges
 * (src/cudaWarpCompress.cu:1577,1626,1747,1817). */
int gpz_timing_enable(int on);
int gpz_timing_reset(void);
int gpz_timing_read(const char* kernel, double* total_ms, uint64_t* launches);

/* ---------------------------------------------------------------- batches
 * Many snapshots of uneven size, dimension and precision in one launch per
 * (precision, dimension) class.  No counterpart in the reference, whose
 * compressor handles one snapshot per call (src/cudaWarpCompress.cu:1573-1741,
 * 1743-1856).  Each snapshot's stream is byte-identical to what
 * gpz_compress writes for it alone.
 *
 * Use: gpz_batch_plan_* uploads the snapshot table and the segment queues
 * into the workspace (synchronous); gpz_batch_run then only launches
 * kernels, so it can be captured in a hipGraph (gpz_batch_graph_create)
 * and replayed while the planned buffers stay valid.  Calls that overlap
 * on one workspace race. */
typedef struct gpz_batch_item {   /* one snapshot to compress */
  gpz_params p;                   /* as for gpz_compress */
  const void* x;                  /* device planes */
  const void* y;
  const void* z;                  /* NULL in 2D */
  void* stream;                   /* device output, >= gpz_stream_bound(&p) bytes */
  uint64_t cap;
} gpz_batch_item;

typedef struct gpz_batch_ditem {  /* one stream to decompress */
  gpz_header h;                   /* from gpz_parse_header / gpz_read_header */
  const void* stream;             /* device */
  uint64_t bytes;
  void* x;                        /* device outputs, h.n elements each */
  void* y;
  void* z;                        /* NULL in 2D */
} gpz_batch_ditem;

typedef struct gpz_batch_tuning {
  uint32_t seg_groups;            /* 1024-point groups per queue segment: 1, 2, 4, 8 */
  uint32_t waves_per_wg;          /* waves per persistent workgroup: 1, 2, 4, 8 */
} gpz_batch_tuning;

typedef struct gpz_batch_plan {   /* filled by the planners; pass back unchanged */
  uint32_t count;
  uint32_t kind;                  /* 0 compress, 1 decompress */
  uint32_t seg_groups;
  uint32_t waves_per_wg;
  uint32_t cls_items[4];          /* classes: f32 2D, f32 3D, f64 2D, f64 3D */
  uint32_t cls_first[4];
  uint32_t cls_segs[4];
  uint32_t cls_seg_first[4];
  uint32_t grid[4];               /* persistent workgroups per class */
  uint64_t total_groups;
  uint64_t ws_bytes;
} gpz_batch_plan;

uint64_t gpz_batch_compress_workspace_bytes(const gpz_batch_item* items, uint32_t count);
uint64_t gpz_batch_decompress_workspace_bytes(const gpz_batch_ditem* items, uint32_t count);
/* tune NULL: one group per segment, one wave per workgroup. */
int gpz_batch_plan_compress(const gpz_batch_item* items, uint32_t count,
                            const gpz_batch_tuning* tune, void* d_ws, uint64_t ws_bytes,
                            gpz_batch_plan* plan, gpz_stream_t stream);
int gpz_batch_plan_decompress(const gpz_batch_ditem* items, uint32_t count,
                              const gpz_batch_tuning* tune, void* d_ws, uint64_t ws_bytes,
                              gpz_batch_plan* plan, gpz_stream_t stream);
/* Asynchronous: launches the planned batch on `stream`. */
int gpz_batch_run(const gpz_batch_plan* plan, void* d_ws, gpz_stream_t stream);
/* Synchronises: out[i] = result of compressed snapshot i. */
int gpz_batch_compress_results(const gpz_batch_plan* plan, const void* d_ws,
                               gpz_result* out, gpz_stream_t stream);
/* A hipGraphExec_t of gpz_batch_run(plan, d_ws), as an opaque pointer. */
int gpz_batch_graph_create(const gpz_batch_plan* plan, void* d_ws, void** graph_exec);
int gpz_batch_graph_launch(void* graph_exec, gpz_stream_t stream);
int gpz_batch_graph_destroy(void* graph_exec);
/* Autotuner: times every (seg_groups, waves_per_wg) pair over `reps` runs
 * of the batch and returns the fastest; table (16 floats, may be NULL)
 * receives ms per run at [4*log2(seg_groups) + log2(waves_per_wg)].
 * Leaves the workspace planned with the last pair tried. */
int gpz_batch_autotune_compress(const gpz_batch_item* items, uint32_t count, void* d_ws,
                                uint64_t ws_bytes, gpz_stream_t stream, uint32_t reps,
                                gpz_batch_tuning* best, float* table);
int gpz_batch_autotune_decompress(const gpz_batch_ditem* items, uint32_t count, void* d_ws,
                                  uint64_t ws_bytes, gpz_stream_t stream, uint32_t reps,
                                  gpz_batch_tuning* best, float* table);

#ifdef __cplusplus
}
#endif

#endif /* GPZ_H_ */
