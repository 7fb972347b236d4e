// Block-tiled raster files: .bin and tiled TIFF / BigTIFF (see rasterio.hpp).
#include "rasterio.hpp"

#include <fcntl.h>
#include <unistd.h>

#include <algorithm>
#include <cerrno>
#include <cstring>
#include <stdexcept>

namespace mapx {

static void fail(const std::string& m) { throw std::runtime_error(m); }

static const size_t kPage = 4096;  // .bin header page (file/binary.cpp:24)

FileFormat format_of(const std::string& path) {
  if (path == "/dev/null") return FileFormat::Bin;
  const size_t dot = path.find_last_of('.');
  const std::string ext = dot == std::string::npos ? "" : path.substr(dot + 1);
  if (ext == "bin") return FileFormat::Bin;
  if (ext == "tif" || ext == "tiff") return FileFormat::Tiff;
  fail("unknown raster file format (extension .bin, .tif or .tiff): " + path);
  return FileFormat::Bin;
}

static size_t block_bytes(const Meta& m) { return (size_t)m.bs.arr[0] * m.bs.arr[1] * sizeof_dt(m.dt); }
static size_t payload_of(const Meta& m) { return (size_t)m.nb(0) * m.nb(1) * block_bytes(m); }

static void pread_all(int fd, void* dst, size_t len, uint64_t off, const std::string& path) {
  char* p = (char*)dst;
  while (len) {
    const ssize_t r = ::pread(fd, p, len, (off_t)off);
    if (r < 0 && errno == EINTR) continue;
    if (r <= 0) fail("read: short or failed read of " + path);
    p += r; len -= (size_t)r; off += (uint64_t)r;
  }
}

// ---- TIFF --------------------------------------------------------------------------
//
// TIFF 6.0 and BigTIFF: an 8-byte (16-byte) header giving the byte order
// ("II" little, "MM" big), the version (42 classic, 43 BigTIFF) and the offset
// of the first image file directory (IFD); an IFD is a count of 12-byte
// (20-byte) entries {tag, type, count, value or offset of the values} sorted by
// tag, then the offset of the next IFD. Values that fit the entry's value
// field (4 bytes, 8 for BigTIFF) are stored in it.

namespace tiff {
enum Tag : uint16_t {
  ImageWidth = 256, ImageLength = 257, BitsPerSample = 258, Compression = 259, Photometric = 262,
  StripOffsets = 273, Orientation = 274, SamplesPerPixel = 277, PlanarConfig = 284, TileWidth = 322,
  TileLength = 323, TileOffsets = 324, TileByteCounts = 325, SampleFormat = 339
};
enum Type : uint16_t { BYTE = 1, SHORT = 3, LONG = 4, LONG8 = 16 };
enum SampleFmt { UINT = 1, INT = 2, IEEEFP = 3 };

static int type_size(int t) {
  switch (t) {
    case 1: case 2: case 6: case 7: return 1;
    case 3: case 8: return 2;
    case 4: case 9: case 11: case 13: return 4;
    case 5: case 10: case 12: case 16: case 17: case 18: return 8;
    default: return 0;
  }
}

struct Reader {
  int fd;
  const std::string& path;
  bool be = false, big = false;
  uint64_t get(const unsigned char* p, int n) const {
    uint64_t v = 0;
    for (int i = 0; i < n; i++) v |= (uint64_t)p[be ? n - 1 - i : i] << (8 * i);
    return v;
  }
  uint64_t at(uint64_t off, int n) const {
    unsigned char b[8];
    pread_all(fd, b, (size_t)n, off, path);
    return get(b, n);
  }
};

struct Entry {
  uint16_t tag = 0, type = 0;
  uint64_t count = 0;
  unsigned char field[8] = {};  // value field as stored
};

// Unsigned integer values of an entry (BYTE, SHORT, LONG, LONG8).
static std::vector<uint64_t> values(const Reader& rd, const Entry& e) {
  const int ts = type_size(e.type);
  if (e.type != BYTE && e.type != SHORT && e.type != LONG && e.type != LONG8)
    fail("read: TIFF tag " + std::to_string(e.tag) + " has a non-integer type in " + rd.path);
  if (e.count > ((uint64_t)1 << 32)) fail("read: TIFF tag " + std::to_string(e.tag) + " count too large in " + rd.path);
  std::vector<uint64_t> out((size_t)e.count);
  const uint64_t bytes = e.count * (uint64_t)ts;
  const size_t field = rd.big ? 8 : 4;
  std::vector<unsigned char> buf((size_t)bytes);
  if (bytes <= field) std::memcpy(buf.data(), e.field, (size_t)bytes);
  else pread_all(rd.fd, buf.data(), (size_t)bytes, rd.get(e.field, (int)field), rd.path);
  for (size_t i = 0; i < out.size(); i++) out[i] = rd.get(buf.data() + i * ts, ts);
  return out;
}
}  // namespace tiff

RasterFile::RasterFile(const std::string& path) : path_(path), fmt_(format_of(path)) {
  fd_ = ::open(path.c_str(), O_RDONLY | O_CLOEXEC);
  if (fd_ < 0) fail("read: cannot open " + path);
  try {
    if (fmt_ == FileFormat::Bin) {
      int32_t h[7];
      pread_all(fd_, h, sizeof h, 0, path);
      meta_.dt = h[0]; meta_.nd = h[1]; meta_.ds = dim2(h[2], h[3]); meta_.mo = h[4]; meta_.bs = dim2(h[5], h[6]);
      if (meta_.dt <= NONE_DATATYPE || meta_.dt >= N_DATATYPE) fail("read: bad data type in " + path);
      if (meta_.nd != D2) fail("read: only D2 rasters are supported: " + path);
      if (meta_.mo != (ROW + BLK)) fail("read: memory order must be ROW+BLK: " + path);
      if (h[2] <= 0 || h[3] <= 0 || h[5] <= 0 || h[6] <= 0) fail("read: bad sizes in " + path);
    } else {
      tiff::Reader rd{fd_, path_};
      unsigned char h[16];
      pread_all(fd_, h, 8, 0, path);
      if (h[0] == 'I' && h[1] == 'I') rd.be = false;
      else if (h[0] == 'M' && h[1] == 'M') rd.be = true;
      else fail("read: not a TIFF file: " + path);
      const uint64_t ver = rd.get(h + 2, 2);
      uint64_t ifd;
      if (ver == 42) ifd = rd.get(h + 4, 4);
      else if (ver == 43) {
        rd.big = true;
        pread_all(fd_, h + 8, 8, 8, path);
        if (rd.get(h + 4, 2) != 8) fail("read: unsupported BigTIFF offset size in " + path);
        ifd = rd.get(h + 8, 8);
      } else fail("read: unknown TIFF version in " + path);
      const uint64_t n = rd.big ? rd.at(ifd, 8) : rd.at(ifd, 2);
      if (n == 0 || n > 4096) fail("read: bad TIFF directory in " + path);
      const size_t esz = rd.big ? 20 : 12, hdr = rd.big ? 8 : 2;
      std::vector<unsigned char> raw((size_t)n * esz);
      pread_all(fd_, raw.data(), raw.size(), ifd + hdr, path);
      std::vector<tiff::Entry> es((size_t)n);
      for (size_t i = 0; i < es.size(); i++) {
        const unsigned char* p = raw.data() + i * esz;
        es[i].tag = (uint16_t)rd.get(p, 2);
        es[i].type = (uint16_t)rd.get(p + 2, 2);
        es[i].count = rd.big ? rd.get(p + 4, 8) : rd.get(p + 4, 4);
        std::memcpy(es[i].field, p + (rd.big ? 12 : 8), rd.big ? 8 : 4);
      }
      auto find = [&](uint16_t t) -> const tiff::Entry* {
        for (auto& e : es) if (e.tag == t) return &e;
        return nullptr;
      };
      auto one = [&](uint16_t t, uint64_t dflt, bool required) -> uint64_t {
        const tiff::Entry* e = find(t);
        if (!e) {
          if (required) fail("read: TIFF tag " + std::to_string(t) + " missing in " + path);
          return dflt;
        }
        std::vector<uint64_t> v = tiff::values(rd, *e);
        if (v.empty()) fail("read: empty TIFF tag " + std::to_string(t) + " in " + path);
        for (uint64_t x : v)
          if (x != v[0]) fail("read: TIFF tag " + std::to_string(t) + " differs between samples in " + path);
        return v[0];
      };
      // file/tiff.cpp:184-282: one sample per pixel, tiles (strips are not supported)
      if (one(tiff::SamplesPerPixel, 1, false) != 1) fail("read: SamplesPerPixel must be 1: " + path);
      if (one(tiff::Compression, 1, false) != 1) fail("read: compressed TIFF is not supported: " + path);
      if (one(tiff::PlanarConfig, 1, false) != 1) fail("read: PlanarConfiguration must be 1: " + path);
      if (!find(tiff::TileOffsets)) fail("read: only tiled TIFF is supported (strips found): " + path);
      const uint64_t fmt = one(tiff::SampleFormat, tiff::UINT, false), bps = one(tiff::BitsPerSample, 1, true);
      int dt = NONE_DATATYPE;
      if (fmt == tiff::UINT) dt = bps == 8 ? U8 : bps == 16 ? U16 : bps == 32 ? U32 : NONE_DATATYPE;
      else if (fmt == tiff::INT) dt = bps == 8 ? S8 : bps == 16 ? S16 : bps == 32 ? S32 : NONE_DATATYPE;
      else if (fmt == tiff::IEEEFP) dt = bps == 32 ? F32 : bps == 64 ? F64 : NONE_DATATYPE;
      if (dt == NONE_DATATYPE) fail("read: unsupported TIFF sample format/bits in " + path);
      const uint64_t w = one(tiff::ImageWidth, 0, true), l = one(tiff::ImageLength, 0, true);
      const uint64_t tw = one(tiff::TileWidth, 0, true), tl = one(tiff::TileLength, 0, true);
      if (w == 0 || l == 0 || tw == 0 || tl == 0 || w > 0x7fffffff || l > 0x7fffffff || tw > 0x7fffffff || tl > 0x7fffffff)
        fail("read: bad TIFF sizes in " + path);
      meta_.dt = dt; meta_.nd = D2; meta_.mo = ROW + BLK;
      meta_.ds = dim2((int)w, (int)l);
      meta_.bs = dim2((int)tw, (int)tl);
      tile_off_ = tiff::values(rd, *find(tiff::TileOffsets));
      const size_t ntiles = (size_t)meta_.nb(0) * meta_.nb(1);
      if (tile_off_.size() != ntiles) fail("read: TIFF tile count does not match its sizes in " + path);
      const size_t tb = block_bytes(meta_);
      if (const tiff::Entry* bc = find(tiff::TileByteCounts)) {
        for (uint64_t c : tiff::values(rd, *bc))
          if (c != tb) fail("read: TIFF tile byte count differs from an uncompressed tile in " + path);
      }
      swap_ = rd.be ? sizeof_dt(dt) : 0;
      if (swap_ == 1) swap_ = 0;
    }
    block_bytes_ = block_bytes(meta_);
    payload_ = payload_of(meta_);
  } catch (...) {
    ::close(fd_);
    fd_ = -1;
    throw;
  }
}

RasterFile::~RasterFile() {
  if (fd_ >= 0) ::close(fd_);
}

void RasterFile::read(size_t off, size_t len, void* dst) const {
  if (off + len > payload_) fail("read: range past the payload of " + path_);
  if (fmt_ == FileFormat::Bin) {
    pread_all(fd_, dst, len, kPage + off, path_);
    return;
  }
  // tiles hold whole blocks: a payload range maps to runs inside consecutive tiles
  char* p = (char*)dst;
  size_t o = off, left = len;
  while (left) {
    const size_t t = o / block_bytes_, in = o % block_bytes_;
    const size_t n = std::min(left, block_bytes_ - in);
    pread_all(fd_, p, n, tile_off_[t] + in, path_);
    if (swap_) {
      if (in % (size_t)swap_) fail("read: unaligned range in " + path_);
      for (size_t i = 0; i + (size_t)swap_ <= n; i += (size_t)swap_) std::reverse(p + i, p + i + swap_);
    }
    p += n; o += n; left -= n;
  }
}

Meta raster_file_meta(const std::string& path) {
  RasterFile f(path);
  return f.meta();
}

// ---- writer ------------------------------------------------------------------------

RasterWriter::RasterWriter(const std::string& path, const Meta& m) : path_(path), fmt_(format_of(path)), meta_(m) {
  if (m.nd != D2) fail("write: needs a D2 raster");
  if (fmt_ == FileFormat::Tiff) {
    // file/tiff.cpp:294-305: the sample types a TIFF holds
    const int dt = m.dt;
    if (!(dt == U8 || dt == U16 || dt == U32 || dt == S8 || dt == S16 || dt == S32 || dt == F32 || dt == F64))
      fail("write: data type not supported by TIFF: " + path);
  }
  payload_ = payload_of(m);
  fd_ = ::open(path.c_str(), O_WRONLY | O_CREAT | O_TRUNC | O_CLOEXEC, 0644);
  if (fd_ < 0) fail("write: cannot open " + path);
  if (fmt_ == FileFormat::Bin) {
    std::vector<char> page(kPage, 0);
    const int32_t h[7] = {m.dt, m.nd, m.ds.arr[0], m.ds.arr[1], m.mo, m.bs.arr[0], m.bs.arr[1]};
    std::memcpy(page.data(), h, sizeof h);
    put(page.data(), page.size());
  } else {
    // BigTIFF header; the first IFD follows the tiles (its offset is patched in finish())
    unsigned char h[16] = {'I', 'I', 43, 0, 8, 0, 0, 0};
    put(h, sizeof h);
  }
}

RasterWriter::~RasterWriter() {
  if (fd_ >= 0) {
    ::close(fd_);
    if (!done_ && path_ != "/dev/null") ::unlink(path_.c_str());
  }
}

void RasterWriter::put(const void* p, size_t n) {
  const char* c = (const char*)p;
  while (n) {
    const ssize_t r = ::write(fd_, c, n);
    if (r < 0 && errno == EINTR) continue;
    if (r <= 0) fail("write: I/O error on " + path_);
    c += r; n -= (size_t)r; pos_ += (uint64_t)r;
  }
}

void RasterWriter::append(const void* src, size_t len) {
  if (written_ + len > payload_) fail("write: more payload than the raster holds: " + path_);
  put(src, len);
  written_ += len;
}

void RasterWriter::finish() {
  if (done_) return;
  if (written_ != payload_) fail("write: incomplete payload for " + path_);
  if (fmt_ == FileFormat::Tiff) {
    const Meta& m = meta_;
    const uint64_t ntiles = (uint64_t)m.nb(0) * m.nb(1), tb = block_bytes(m), data0 = 16;
    const int dt = m.dt;
    const uint64_t fmt = (dt == F32 || dt == F64) ? tiff::IEEEFP : (dt == S8 || dt == S16 || dt == S32) ? tiff::INT : tiff::UINT;
    // out-of-line arrays (tile offsets and byte counts, LONG8) after the tiles, then the IFD
    if (pos_ % 8) {
      const unsigned char z[8] = {};
      put(z, 8 - pos_ % 8);
    }
    const uint64_t offs_at = pos_;
    std::vector<uint64_t> arr((size_t)ntiles);
    for (uint64_t t = 0; t < ntiles; t++) arr[(size_t)t] = data0 + t * tb;
    if (ntiles > 1) put(arr.data(), arr.size() * 8);
    const uint64_t cnts_at = pos_;
    std::fill(arr.begin(), arr.end(), tb);
    if (ntiles > 1) put(arr.data(), arr.size() * 8);
    const uint64_t ifd_at = pos_;
    struct E { uint16_t tag, type; uint64_t count, value; };
    const E es[] = {
        {tiff::ImageWidth, tiff::LONG, 1, (uint64_t)m.ds.arr[0]},
        {tiff::ImageLength, tiff::LONG, 1, (uint64_t)m.ds.arr[1]},
        {tiff::BitsPerSample, tiff::SHORT, 1, (uint64_t)sizeof_dt(dt) * 8},
        {tiff::Compression, tiff::SHORT, 1, 1},       // none
        {tiff::Photometric, tiff::SHORT, 1, 1},       // min-is-black
        {tiff::Orientation, tiff::SHORT, 1, 1},       // top-left
        {tiff::SamplesPerPixel, tiff::SHORT, 1, 1},
        {tiff::PlanarConfig, tiff::SHORT, 1, 1},      // contiguous
        {tiff::TileWidth, tiff::LONG, 1, (uint64_t)m.bs.arr[0]},
        {tiff::TileLength, tiff::LONG, 1, (uint64_t)m.bs.arr[1]},
        {tiff::TileOffsets, tiff::LONG8, ntiles, ntiles > 1 ? offs_at : data0},
        {tiff::TileByteCounts, tiff::LONG8, ntiles, ntiles > 1 ? cnts_at : tb},
        {tiff::SampleFormat, tiff::SHORT, 1, fmt},
    };
    const uint64_t n = sizeof es / sizeof es[0];
    std::vector<unsigned char> ifd(8 + n * 20 + 8, 0);
    std::memcpy(ifd.data(), &n, 8);
    for (uint64_t i = 0; i < n; i++) {
      unsigned char* p = ifd.data() + 8 + i * 20;
      std::memcpy(p, &es[i].tag, 2);
      std::memcpy(p + 2, &es[i].type, 2);
      std::memcpy(p + 4, &es[i].count, 8);
      std::memcpy(p + 12, &es[i].value, 8);  // little-endian: small values sit in the low bytes
    }
    put(ifd.data(), ifd.size());  // next IFD offset 0 closes the chain
    if (::pwrite(fd_, &ifd_at, 8, 8) != 8) fail("write: I/O error on " + path_);
  }
  if (::close(fd_) != 0) {
    fd_ = -1;
    fail("write: I/O error on " + path_);
  }
  fd_ = -1;
  done_ = true;
}

}  // namespace mapx
