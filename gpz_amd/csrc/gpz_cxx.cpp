// gpz_cxx.cpp -- the reference's C++ interface (ICudaCompressor,
// CudaWarpCompress, CudaWarpCompressDouble, CompressorManager) implemented on
// the C ABI of include/gpz.h.
//
// Semantics kept from the reference:
//   * compress() leaves the stream in d_xyz (src/cudaWarpCompress.cu:1669,
//     read back by src/main.cpp:348) and its size in *comp_data_size;
//   * the segment size chosen on the first call sticks to the object
//     (CudaConfig is updated, src/cudaWarpCompress.cu:1513,1552-1553);
//   * both calls block until the device work is done;
//   * a non-"none" nvcomp type is refused (src/cudaWarpCompress.cu:1614-1617).
// Differences: errors throw gpz::Error instead of exit(); device workspace
// is cached in the object instead of cudaMalloc/cudaFree in every call.

#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <string>

#include "gpz.h"
#include "gpz/CompressorManager.h"
#include "gpz/ICudaCompressor.h"
#include "gpz/cudaWarpCompress.h"

namespace {

void hip_check(hipError_t e, const char* what) {
  if (e != hipSuccess)
    throw gpz::Error(GPZ_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

void gpz_check(int rc, const char* what) {
  if (rc == GPZ_OK) return;
  const char* msg = rc == GPZ_ERR_INVALID_ARG     ? "invalid argument"
                    : rc == GPZ_ERR_WORKSPACE     ? "workspace too small"
                    : rc == GPZ_ERR_CAPACITY      ? "output buffer too small"
                    : rc == GPZ_ERR_CORRUPT       ? "corrupt stream"
                    : rc == GPZ_ERR_SEG_TOO_LARGE ? "The auto changed segment size is too large, please adjust your error bound"
                    : rc == GPZ_ERR_HIP           ? "HIP runtime error"
                                                  : "error";
  throw gpz::Error(rc, std::string(what) + ": " + msg);
}

void check_type(const std::string& t) {
  if (!t.empty() && t != "none")
    throw gpz::Error(GPZ_ERR_INVALID_ARG,
                     "nvcomp lossless post-pass '" + t + "' is not supported by this build");
}

template <typename T>
void compress_common(CudaConfig& cfg, gpz::DeviceCache& cache, void* d_xyz,
                     unsigned char* comp_data, uint64_t* comp_data_size,
                     const std::string& comp_type) {
  check_type(comp_type);
  if (!d_xyz || !comp_data || !comp_data_size)
    throw gpz::Error(GPZ_ERR_INVALID_ARG, "null buffer");
  gpz_params p;
  p.dim = cfg.dim;
  p.float_bits = sizeof(T) * 8;
  p.n = cfg.coordinate_num;
  p.eb = sizeof(T) == 4 ? static_cast<double>(cfg.eb) : cfg.eb_dbl;
  p.rel_eb = sizeof(T) == 4 ? static_cast<double>(cfg.rel_eb) : cfg.rel_eb_dbl;
  p.mode = cfg.error_mode == "rel" ? GPZ_MODE_REL : GPZ_MODE_ABS;
  p.seg_size = cfg.seg_size;
  const uint64_t N = cfg.coordinate_num;
  T* x = static_cast<T*>(d_xyz);
  const uint64_t bound = gpz_stream_bound(&p);
  // comp_data (>= 4 N sizeof(T) bytes, src/main.cpp:289-291) holds the
  // stream, which is larger than the planes' 3 N sizeof(T) only if it
  // exceeds the worst-case bound; a cached buffer is used then.
  void* out = bound <= 4 * N * sizeof(T) ? static_cast<void*>(comp_data)
                                         : cache.output(bound);
  void* ws = cache.workspace(gpz_compress_workspace_bytes(&p));
  gpz_check(gpz_compress(&p, x, x + N, x + 2 * N, out, bound, ws,
                         gpz_compress_workspace_bytes(&p), nullptr),
            "compress");
  gpz_result r;
  gpz_check(gpz_compress_result(ws, &r, nullptr), "compress");
  gpz_check(r.status, "compress");
  if (r.stream_bytes > 3 * N * sizeof(T))
    throw gpz::Error(GPZ_ERR_CAPACITY, "compress: stream larger than the input planes");
  hip_check(hipMemcpy(d_xyz, out, r.stream_bytes, hipMemcpyDeviceToDevice),
            "compress: stream copy");
  cfg.seg_size = r.seg_size;
  if (r.adjusted) cfg.adjusted_seg = true;
  *comp_data_size = r.stream_bytes;
}

template <typename T>
void decompress_common(CudaConfig& cfg, void* x, void* y, void* z,
                       unsigned char* d_stream, const std::string& type) {
  check_type(type);
  gpz_header h;
  gpz_check(gpz_read_header(d_stream, &h, nullptr), "decompress");
  if (h.float_bits != sizeof(T) * 8)
    throw gpz::Error(GPZ_ERR_CORRUPT, "decompress: precision of the stream does not match");
  uint64_t bytes = 0;
  gpz_check(gpz_stream_bytes(&h, d_stream, &bytes, nullptr), "decompress");
  gpz_check(gpz_decompress(&h, d_stream, bytes, x, y, z, nullptr), "decompress");
  hip_check(hipDeviceSynchronize(), "decompress");
  cfg.dim = h.dim;
  cfg.coordinate_num = h.n;
  cfg.seg_size = h.seg_size;
  cfg.float_bits = h.float_bits;
  if (sizeof(T) == 4) cfg.eb = static_cast<float>(h.eb);
  else cfg.eb_dbl = h.eb;
}

}  // namespace

namespace gpz {
DeviceCache::~DeviceCache() {
  if (ws) (void)hipFree(ws);
  if (out) (void)hipFree(out);
}
void* DeviceCache::workspace(uint64_t bytes) {
  if (bytes > ws_bytes) {
    if (ws) (void)hipFree(ws);
    ws = nullptr;
    ws_bytes = 0;
    hip_check(hipMalloc(&ws, bytes), "workspace");
    ws_bytes = bytes;
  }
  return ws;
}
void* DeviceCache::output(uint64_t bytes) {
  if (bytes > out_bytes) {
    if (out) (void)hipFree(out);
    out = nullptr;
    out_bytes = 0;
    hip_check(hipMalloc(&out, bytes), "output");
    out_bytes = bytes;
  }
  return out;
}
}  // namespace gpz

CudaWarpCompress::CudaWarpCompress(CudaConfig c) : cuda_config(std::move(c)) {}
CudaWarpCompress::CudaWarpCompress() = default;
CudaWarpCompress::~CudaWarpCompress() = default;

void CudaWarpCompress::compress(void* d_xyz, unsigned char* comp_data,
                                uint64_t* comp_data_size,
                                const std::string& comp_type) {
  compress_common<float>(cuda_config, cache, d_xyz, comp_data, comp_data_size,
                         comp_type);
}

void CudaWarpCompress::decompress(void* x, void* y, void* z,
                                  unsigned char* d_compressed_data,
                                  const std::string& nvcomp_type) {
  decompress_common<float>(cuda_config, x, y, z, d_compressed_data, nvcomp_type);
}

CudaWarpCompressDouble::CudaWarpCompressDouble(CudaConfig c)
    : cuda_config(std::move(c)) {}
CudaWarpCompressDouble::CudaWarpCompressDouble() = default;
CudaWarpCompressDouble::~CudaWarpCompressDouble() = default;

void CudaWarpCompressDouble::compress(void* d_xyz, unsigned char* comp_data,
                                      uint64_t* comp_data_size,
                                      const std::string& comp_type) {
  compress_common<double>(cuda_config, cache, d_xyz, comp_data, comp_data_size,
                          comp_type);
}

void CudaWarpCompressDouble::decompress(void* x, void* y, void* z,
                                        unsigned char* d_compressed_data,
                                        const std::string& nvcomp_type) {
  decompress_common<double>(cuda_config, x, y, z, d_compressed_data, nvcomp_type);
}

CompressorManager::CompressorManager() = default;
CompressorManager::~CompressorManager() = default;

void CompressorManager::compressWarp(void* d_xyz, unsigned char* d_comp_data,
                                     uint64_t* comp_size, std::string nvcomp_type,
                                     ICudaCompressor* compressor) {
  compressor->compress(d_xyz, d_comp_data, comp_size, nvcomp_type);
}

void CompressorManager::decompressWarp(void* d_dec_data_x, void* d_dec_data_y,
                                       void* d_dec_data_z,
                                       unsigned char* d_compressed_data,
                                       ICudaCompressor* decompressor,
                                       const std::string& nvcomp_type) {
  decompressor->decompress(d_dec_data_x, d_dec_data_y, d_dec_data_z,
                           d_compressed_data, nvcomp_type);
}

void CompressorManager::decompressLoadParas(
    uint32_t* dim, uint64_t* numPoints, uint32_t* float_bits,
    const std::string& nvcomp_type, uint64_t nvcomp_num, uint32_t nvcomp_dim,
    uint32_t float_bits_num, unsigned char* h_compressed_data) {
  if (!nvcomp_type.empty() && nvcomp_type != "none") {
    *dim = nvcomp_dim;
    *float_bits = float_bits_num;
    *numPoints = nvcomp_num;
  } else {
    memcpy(dim, h_compressed_data, sizeof(uint32_t));
    memcpy(float_bits, h_compressed_data + 12, sizeof(uint32_t));
    memcpy(numPoints, h_compressed_data + 4, sizeof(uint64_t));
  }
}
