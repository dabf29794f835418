// CompressorManager.h -- the reference's facade
// (include/CompressorManager.h:9-28, src/CompressorManager.cpp:12-34).
#ifndef GPZ_COMPRESSOR_MANAGER_H_
#define GPZ_COMPRESSOR_MANAGER_H_

#include <cstdint>
#include <string>

#include "ICudaCompressor.h"
#include "cudaCompress.h"
#include "cudaWarpCompress.h"

class CompressorManager {
 public:
  CompressorManager();
  ~CompressorManager();

  void compressWarp(void* d_xyz, unsigned char* d_comp_data, uint64_t* comp_size,
                    std::string nvcomp_type, ICudaCompressor* compressor);

  // Reads dim (byte 0), N (byte 4) and float_bits (byte 12) from a host copy
  // of the stream; with an nvcomp type the values come from the arguments.
  void decompressLoadParas(uint32_t* dim, uint64_t* numPoints,
                           uint32_t* float_bits, const std::string& nvcomp_type,
                           uint64_t nvcomp_num, uint32_t nvcomp_dim,
                           uint32_t float_bits_num,
                           unsigned char* h_compressed_data);

  void decompressWarp(void* d_dec_data_x, void* d_dec_data_y,
                      void* d_dec_data_z, unsigned char* d_compressed_data,
                      ICudaCompressor* decompressor,
                      const std::string& nvcomp_type);
};

#endif  // GPZ_COMPRESSOR_MANAGER_H_
