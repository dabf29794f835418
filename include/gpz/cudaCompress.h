// cudaCompress.h -- group constants and CudaConfig, the configuration the
// reference's compressors carry (include/cudaCompress.cuh:9-40).  Same names,
// fields and constructors, so code written against the reference compiles
// unchanged.
#ifndef GPZ_CUDA_COMPRESS_H_
#define GPZ_CUDA_COMPRESS_H_

#include <cstdint>
#include <string>
#include <utility>

static const uint32_t CHUNK_SIZE = 32;
static const uint32_t BLOCK_SIZE = 32;
// 1024 points per group: the on-disk granularity of the stream.
const int ELEMS_PER_BLOCK = BLOCK_SIZE * CHUNK_SIZE;

class CudaConfig {
 public:
  uint32_t seg_size = 0;     // 0 = auto (rel mode); updated by compress()
  bool adjusted_seg = false; // set when the block-count guard doubled seg
  uint32_t dim = 3;
  uint64_t coordinate_num = 0;
  float eb = 0.0f;           // absolute bound, FP32 compressor
  float rel_eb = 0.0f;       // bound as given, FP32
  double eb_dbl = 0.0;       // absolute bound, FP64 compressor
  double rel_eb_dbl = 0.0;   // bound as given, FP64
  std::string error_mode;    // "rel" or "abs"
  uint32_t float_bits = 32;

  CudaConfig() {}

  CudaConfig(uint32_t seg_size, bool adjusted_seg, uint32_t dim,
             uint64_t coordinate_num, std::string error_mode, float eb,
             float rel_eb, uint32_t float_bits)
      : seg_size(seg_size), adjusted_seg(adjusted_seg), dim(dim),
        coordinate_num(coordinate_num), eb(eb), rel_eb(rel_eb),
        error_mode(std::move(error_mode)), float_bits(float_bits) {}

  CudaConfig(uint32_t seg_size, bool adjusted_seg, uint32_t dim,
             uint64_t coordinate_num, std::string error_mode, double eb,
             double rel_eb, uint32_t float_bits)
      : seg_size(seg_size), adjusted_seg(adjusted_seg), dim(dim),
        coordinate_num(coordinate_num), eb_dbl(eb), rel_eb_dbl(rel_eb),
        error_mode(std::move(error_mode)), float_bits(float_bits) {}
};

#endif  // GPZ_CUDA_COMPRESS_H_
