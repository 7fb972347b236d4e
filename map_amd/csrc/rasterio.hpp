// Block-tiled raster files.
//
// Both formats hold a D2 raster as its blocks in block order (bx + by*NB0),
// every block a full BS0 x BS1 samples (edge blocks padded), row-major with x
// fastest — the layout of a resident raster in HBM, so a file is streamed to
// or from the device as one run of payload bytes:
//  - .bin: a 4096-byte header page of little-endian int32 fields (data type,
//    num dim, data size x/y, mem order, block size x/y), then the payload
//    (file/binary.cpp:221-254,344-390);
//  - .tif / .tiff: a tiled TIFF or BigTIFF, one uncompressed single-sample
//    tile per block (file/tiff.cpp:184-300 read, 284-356 write, 499-515 tiles).
//    Files are written as little-endian BigTIFF, as the reference's "w8" mode
//    does; classic TIFF and either byte order are read.
#pragma once

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "dag.hpp"

namespace mapx {

enum class FileFormat { Bin, Tiff };

// Format of a path by its extension (file/File.cpp:25-45): "bin", "tif",
// "tiff"; "/dev/null" is a .bin sink. Other extensions are an error.
FileFormat format_of(const std::string& path);

// Reader: the header is parsed on construction; read() copies payload bytes.
class RasterFile {
 public:
  explicit RasterFile(const std::string& path);
  ~RasterFile();
  RasterFile(const RasterFile&) = delete;
  RasterFile& operator=(const RasterFile&) = delete;

  const Meta& meta() const { return meta_; }
  size_t payload_bytes() const { return payload_; }
  // Payload bytes [off, off + len) into dst.
  void read(size_t off, size_t len, void* dst) const;

 private:
  std::string path_;
  int fd_ = -1;
  FileFormat fmt_;
  Meta meta_;
  size_t payload_ = 0, block_bytes_ = 0;
  std::vector<uint64_t> tile_off_;  // TIFF: file offset of each block's tile
  int swap_ = 0;                    // TIFF in the other byte order: sample size to swap (0: none)
};

// Writer: the payload is appended in order; finish() completes the file
// (a TIFF's directory follows its tiles). A writer destroyed before finish()
// removes the partial file.
class RasterWriter {
 public:
  RasterWriter(const std::string& path, const Meta& m);
  ~RasterWriter();
  RasterWriter(const RasterWriter&) = delete;
  RasterWriter& operator=(const RasterWriter&) = delete;

  size_t payload_bytes() const { return payload_; }
  void append(const void* src, size_t len);
  void finish();

 private:
  void put(const void* p, size_t n);
  std::string path_;
  int fd_ = -1;
  FileFormat fmt_;
  Meta meta_;
  size_t payload_ = 0, written_ = 0;
  uint64_t pos_ = 0;  // file position
  bool done_ = false;
};

// Raster meta of a file without reading its payload.
Meta raster_file_meta(const std::string& path);

}  // namespace mapx
