// ICudaCompressor.h -- the compressor interface of the reference
// (include/ICudaCompressor.cuh:6-17), kept name for name.
//
// compress(d_xyz, comp_data, comp_data_size, comp_type)
//   d_xyz          device SoA planes [x(N) | y(N) | z(N)] of float or double
//                  (always three planes; z is ignored for dim 2).  As in the
//                  reference, the stream is written back over d_xyz.
//   comp_data      device scratch of at least 4*N*sizeof(T) bytes
//                  (src/main.cpp:289-291).
//   comp_data_size host out: stream bytes.
//   comp_type      "none" or "" (lossless post-passes are not built).
// decompress(x, y, z, d_compressed_data, nvcomp_type)
//   d_compressed_data  device stream, header at offset 0; x, y, z device
//                      planes of N elements (z untouched for dim 2).
// Both calls are blocking, like the reference's (src/cudaWarpCompress.cu:1693,
// :1849); errors throw gpz::Error instead of exiting the process.
#ifndef GPZ_ICUDA_COMPRESSOR_H_
#define GPZ_ICUDA_COMPRESSOR_H_

#include <cstdint>
#include <stdexcept>
#include <string>

#include "cudaCompress.h"

namespace gpz {
class Error : public std::runtime_error {
 public:
  Error(int code, const std::string& what)
      : std::runtime_error(what), code_(code) {}
  int code() const { return code_; }

 private:
  int code_;
};
}  // namespace gpz

class ICudaCompressor {
 public:
  virtual ~ICudaCompressor() = default;

  virtual void compress(void* d_xyz, unsigned char* comp_data,
                        uint64_t* comp_data_size,
                        const std::string& comp_type) = 0;

  virtual void decompress(void* decData_x, void* decData_y, void* decData_z,
                          unsigned char* d_compressed_data,
                          const std::string& nvcomp_type) = 0;

  virtual const CudaConfig& getConfig() const = 0;
};

#endif  // GPZ_ICUDA_COMPRESSOR_H_
