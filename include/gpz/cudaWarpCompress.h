// cudaWarpCompress.h -- FP32 and FP64 compressors
// (include/cudaWarpCompress.cuh:11-38, include/cudaWarpCompressDouble.cuh),
// implemented on the MI355X kernels of libgpz_amd through include/gpz.h.
#ifndef GPZ_CUDA_WARP_COMPRESS_H_
#define GPZ_CUDA_WARP_COMPRESS_H_

#include <cstdint>
#include <string>

#include "ICudaCompressor.h"

namespace gpz {
// Device buffers the compressor keeps between calls (the reference
// cudaMallocs and frees them inside every call,
// src/cudaWarpCompress.cu:1628-1634,1658-1666).
struct DeviceCache {
  void* ws = nullptr;
  uint64_t ws_bytes = 0;
  void* out = nullptr;
  uint64_t out_bytes = 0;
  ~DeviceCache();
  void* workspace(uint64_t bytes);
  void* output(uint64_t bytes);
};
}  // namespace gpz

class CudaWarpCompress : public ICudaCompressor {
 public:
  explicit CudaWarpCompress(CudaConfig cuda_config);
  CudaWarpCompress();
  ~CudaWarpCompress() override;

  void compress(void* d_xyz, unsigned char* comp_data,
                uint64_t* comp_data_size, const std::string& comp_type) override;
  void decompress(void* decData_x, void* decData_y, void* decData_z,
                  unsigned char* d_compressed_data,
                  const std::string& nvcomp_type) override;
  const CudaConfig& getConfig() const override { return cuda_config; }

 private:
  CudaConfig cuda_config;
  gpz::DeviceCache cache;
};

class CudaWarpCompressDouble : public ICudaCompressor {
 public:
  explicit CudaWarpCompressDouble(CudaConfig cuda_config);
  CudaWarpCompressDouble();
  ~CudaWarpCompressDouble() override;

  void compress(void* d_xyz, unsigned char* comp_data,
                uint64_t* comp_data_size, const std::string& comp_type) override;
  void decompress(void* decData_x, void* decData_y, void* decData_z,
                  unsigned char* d_compressed_data,
                  const std::string& nvcomp_type) override;
  const CudaConfig& getConfig() const override { return cuda_config; }

 private:
  CudaConfig cuda_config;
  gpz::DeviceCache cache;
};

#endif  // GPZ_CUDA_WARP_COMPRESS_H_
