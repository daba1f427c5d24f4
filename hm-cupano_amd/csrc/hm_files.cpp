// hm_files.cpp -- TIFF and PNG decoding, the TIFF writer of pyramid dumps,
// and the loaders of a rig's control directory.  Host C++ over zlib; no GPU
// code.
//
// The loaders follow the reference's rules:
//   ControlMasks::load   src/pano/controlMasks.cpp:273-326 (maps, seam
//                        inversion :225-251, placements :96-118, :42-57)
//   ControlMasksN::load  src/pano/controlMasksN.cpp:125-193 (unique-label LUT)
//   ControlMasks3::load  src/pano/controlMasks3.cpp:228-295 (paletted seam)
// The reference reads the files with libtiff, libpng and OpenCV; the decoders
// here are written from the TIFF 6.0 and PNG specifications.
#include "hm_files.h"

#include <sys/stat.h>
#include <zlib.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <string>
#include <vector>

namespace {

thread_local std::string g_error;

int fail(int code, const std::string& msg) {
  g_error = msg;
  return code;
}

bool file_exists(const std::string& path) {
  struct stat st;
  return stat(path.c_str(), &st) == 0 && S_ISREG(st.st_mode);
}

int read_whole(const std::string& path, std::vector<uint8_t>& out) {
  FILE* f = fopen(path.c_str(), "rb");
  if (!f) return fail(HM_FILE_NOT_FOUND, "cannot open " + path);
  std::vector<uint8_t> buf;
  uint8_t chunk[1 << 16];
  size_t got;
  while ((got = fread(chunk, 1, sizeof(chunk), f)) > 0) buf.insert(buf.end(), chunk, chunk + got);
  const bool err = ferror(f) != 0;
  fclose(f);
  if (err) return fail(HM_FILE_NOT_FOUND, "read error on " + path);
  out.swap(buf);
  return HM_FILE_OK;
}

// ===========================================================================
// TIFF
// ===========================================================================
enum : uint16_t {
  kTagWidth = 256,
  kTagHeight = 257,
  kTagBits = 258,
  kTagCompression = 259,
  kTagPhotometric = 262,
  kTagFillOrder = 266,
  kTagStripOffsets = 273,
  kTagSamples = 277,
  kTagRowsPerStrip = 278,
  kTagStripCounts = 279,
  kTagXRes = 282,
  kTagYRes = 283,
  kTagPlanar = 284,
  kTagXPos = 286,
  kTagYPos = 287,
  kTagResUnit = 296,
  kTagPredictor = 317,
  kTagTileWidth = 322,
  kTagTileLength = 323,
  kTagTileOffsets = 324,
  kTagTileCounts = 325,
  kTagExtraSamples = 338,
  kTagSampleFormat = 339,
};

struct Bytes {
  const uint8_t* p = nullptr;
  size_t n = 0;
  bool be = false;
  bool u16(size_t off, uint32_t& v) const {
    if (off > n || n - off < 2) return false;
    v = be ? (uint32_t)p[off] << 8 | p[off + 1] : (uint32_t)p[off + 1] << 8 | p[off];
    return true;
  }
  bool u32(size_t off, uint32_t& v) const {
    if (off > n || n - off < 4) return false;
    const uint8_t* q = p + off;
    v = be ? (uint32_t)q[0] << 24 | (uint32_t)q[1] << 16 | (uint32_t)q[2] << 8 | q[3]
           : (uint32_t)q[3] << 24 | (uint32_t)q[2] << 16 | (uint32_t)q[1] << 8 | q[0];
    return true;
  }
};

struct Entry {
  uint16_t tag = 0, type = 0;
  uint32_t count = 0;
  size_t value_off = 0;  // file offset of the first value (inline or not)
};

size_t type_size(uint16_t t) {
  switch (t) {
    case 1: case 2: case 6: case 7: return 1;  // BYTE ASCII SBYTE UNDEFINED
    case 3: case 8: return 2;                  // SHORT SSHORT
    case 4: case 9: case 11: return 4;         // LONG SLONG FLOAT
    case 5: case 10: case 12: return 8;        // RATIONAL SRATIONAL DOUBLE
    default: return 0;
  }
}

struct Ifd {
  Bytes b;
  std::vector<Entry> e;
  const Entry* find(uint16_t tag) const {
    for (const Entry& x : e)
      if (x.tag == tag) return &x;
    return nullptr;
  }
  // Value i of an entry as double; rationals as num / den (0 when either is 0).
  bool value(const Entry& x, uint32_t i, double& v) const {
    if (i >= x.count) return false;
    const size_t sz = type_size(x.type);
    const size_t off = x.value_off + (size_t)i * sz;
    uint32_t a, c;
    switch (x.type) {
      case 1: case 7:
        if (off >= b.n) return false;
        v = b.p[off];
        return true;
      case 6:
        if (off >= b.n) return false;
        v = (int8_t)b.p[off];
        return true;
      case 3:
        if (!b.u16(off, a)) return false;
        v = a;
        return true;
      case 8:
        if (!b.u16(off, a)) return false;
        v = (int16_t)a;
        return true;
      case 4:
        if (!b.u32(off, a)) return false;
        v = a;
        return true;
      case 9:
        if (!b.u32(off, a)) return false;
        v = (int32_t)a;
        return true;
      case 5:
        if (!b.u32(off, a) || !b.u32(off + 4, c)) return false;
        v = (a == 0 || c == 0) ? 0.0 : (double)a / (double)c;
        return true;
      case 10:
        if (!b.u32(off, a) || !b.u32(off + 4, c)) return false;
        v = (a == 0 || c == 0) ? 0.0 : (double)(int32_t)a / (double)(int32_t)c;
        return true;
      case 11: {
        if (!b.u32(off, a)) return false;
        float f;
        memcpy(&f, &a, 4);
        v = f;
        return true;
      }
      case 12: {
        if (!b.u32(off, a) || !b.u32(off + 4, c)) return false;
        const uint64_t u = b.be ? ((uint64_t)a << 32 | c) : ((uint64_t)c << 32 | a);
        double d;
        memcpy(&d, &u, 8);
        v = d;
        return true;
      }
      default: return false;
    }
  }
  bool uint_at(uint16_t tag, uint32_t i, uint64_t& v) const {
    const Entry* x = find(tag);
    double d;
    if (!x || !value(*x, i, d) || d < 0) return false;
    v = (uint64_t)d;
    return true;
  }
  uint64_t uint_or(uint16_t tag, uint64_t dflt) const {
    uint64_t v;
    return uint_at(tag, 0, v) ? v : dflt;
  }
  // libtiff's float view of a rational tag: float(double(num) / double(den)).
  bool float_tag(uint16_t tag, float& v) const {
    const Entry* x = find(tag);
    double d;
    if (!x || !value(*x, 0, d)) return false;
    v = (float)d;
    return true;
  }
};

int parse_ifd0(const std::vector<uint8_t>& file, const std::string& path, Ifd& ifd) {
  if (file.size() < 8) return fail(HM_FILE_BAD_FORMAT, path + ": too short for a TIFF");
  Bytes b{file.data(), file.size(), false};
  if (file[0] == 'I' && file[1] == 'I') b.be = false;
  else if (file[0] == 'M' && file[1] == 'M') b.be = true;
  else return fail(HM_FILE_BAD_FORMAT, path + ": not a TIFF");
  uint32_t magic = 0, off = 0, cnt = 0;
  b.u16(2, magic);
  if (magic == 43) return fail(HM_FILE_UNSUPPORTED, path + ": BigTIFF is not supported");
  if (magic != 42) return fail(HM_FILE_BAD_FORMAT, path + ": not a TIFF");
  if (!b.u32(4, off) || !b.u16(off, cnt) || (size_t)off + 2 + (size_t)cnt * 12 > file.size())
    return fail(HM_FILE_BAD_FORMAT, path + ": truncated IFD");
  ifd.b = b;
  ifd.e.clear();
  for (uint32_t i = 0; i < cnt; ++i) {
    const size_t at = (size_t)off + 2 + (size_t)i * 12;
    uint32_t tag = 0, type = 0, count = 0, vo = 0;
    b.u16(at, tag);
    b.u16(at + 2, type);
    b.u32(at + 4, count);
    const size_t sz = type_size((uint16_t)type);
    if (sz == 0) continue;  // unknown type: skip, as readers must
    Entry x;
    x.tag = (uint16_t)tag;
    x.type = (uint16_t)type;
    x.count = count;
    if ((uint64_t)sz * count <= 4) {
      x.value_off = at + 8;
    } else {
      b.u32(at + 8, vo);
      if ((uint64_t)vo + (uint64_t)sz * count > file.size())
        return fail(HM_FILE_BAD_FORMAT, path + ": tag " + std::to_string(tag) + " points past the end");
      x.value_off = vo;
    }
    ifd.e.push_back(x);
  }
  return HM_FILE_OK;
}

// --- decompressors --------------------------------------------------------
// TIFF LZW (TIFF 6.0 section 13): MSB-first codes of 9..12 bits, Clear 256,
// EndOfInformation 257, the width grows one code early (at 511, 1023, 2047).
bool lzw_decode(const uint8_t* in, size_t n, uint8_t* out, size_t out_n, std::string& err) {
  if (n >= 2 && in[0] == 0 && (in[1] & 1)) {
    err = "old-style (pre-6.0) LZW is not supported";
    return false;
  }
  const int kClear = 256, kEoi = 257, kMax = 4096;
  std::vector<uint16_t> prefix(kMax), len(kMax);
  std::vector<uint8_t> suffix(kMax), first(kMax);
  for (int i = 0; i < 256; ++i) {
    prefix[i] = 0xFFFF;
    suffix[i] = first[i] = (uint8_t)i;
    len[i] = 1;
  }
  int next = 258, width = 9, prev = -1;
  uint64_t bits = 0;
  int nbits = 0;
  size_t ip = 0, op = 0;
  auto emit = [&](int code) {
    const size_t l = len[code];
    const size_t take = std::min(l, out_n - op);
    // walk from the last byte of the string back to its first
    size_t k = l;
    int c = code;
    while (k > 0) {
      --k;
      if (k < take) out[op + k] = suffix[c];
      c = prefix[c];
    }
    op += take;
  };
  while (op < out_n) {
    while (nbits < width) {
      if (ip >= n) {
        nbits = -1;
        break;
      }
      bits = (bits << 8) | in[ip++];
      nbits += 8;
    }
    if (nbits < 0) break;  // data ended without EndOfInformation
    const int code = (int)((bits >> (nbits - width)) & ((1u << width) - 1));
    nbits -= width;
    if (code == kEoi) break;
    if (code == kClear) {
      next = 258;
      width = 9;
      prev = -1;
      continue;
    }
    if (prev < 0) {
      if (code > 255) {
        err = "LZW: first code after Clear is not a literal";
        return false;
      }
      emit(code);
      prev = code;
      continue;
    }
    if (code > next || (code == next && next >= kMax)) {
      err = "LZW: code out of sequence";
      return false;
    }
    if (next < kMax) {
      const int f = code < next ? first[code] : first[prev];
      prefix[next] = (uint16_t)prev;
      suffix[next] = (uint8_t)f;
      first[next] = first[prev];
      len[next] = (uint16_t)(len[prev] + 1);
      ++next;
    }
    emit(code);
    prev = code;
    if (next + 1 >= (1 << width) && width < 12) ++width;
  }
  if (op < out_n) {
    err = "LZW: data ends " + std::to_string(out_n - op) + " bytes early";
    return false;
  }
  return true;
}

bool packbits_decode(const uint8_t* in, size_t n, uint8_t* out, size_t out_n, std::string& err) {
  size_t ip = 0, op = 0;
  while (op < out_n && ip < n) {
    const int8_t c = (int8_t)in[ip++];
    if (c >= 0) {
      const size_t k = (size_t)c + 1;
      if (ip + k > n || op + k > out_n) {
        err = "PackBits: literal run overflows";
        return false;
      }
      memcpy(out + op, in + ip, k);
      ip += k;
      op += k;
    } else if (c != -128) {
      const size_t k = (size_t)(1 - c);
      if (ip >= n || op + k > out_n) {
        err = "PackBits: repeat run overflows";
        return false;
      }
      memset(out + op, in[ip++], k);
      op += k;
    }
  }
  if (op < out_n) {
    err = "PackBits: data ends early";
    return false;
  }
  return true;
}

// zlib-wrapped deflate into exactly out_n bytes.
bool inflate_exact(const uint8_t* in, size_t n, uint8_t* out, size_t out_n, std::string& err) {
  z_stream z;
  memset(&z, 0, sizeof(z));
  if (inflateInit(&z) != Z_OK) {
    err = "inflateInit failed";
    return false;
  }
  size_t ip = 0, op = 0;
  int rc = Z_OK;
  while (op < out_n && rc != Z_STREAM_END) {
    const uInt in_step = (uInt)std::min<size_t>(n - ip, 1u << 30);
    const uInt out_step = (uInt)std::min<size_t>(out_n - op, 1u << 30);
    z.next_in = const_cast<Bytef*>(in + ip);
    z.avail_in = in_step;
    z.next_out = out + op;
    z.avail_out = out_step;
    rc = inflate(&z, Z_NO_FLUSH);
    ip += in_step - z.avail_in;
    op += out_step - z.avail_out;
    if (rc == Z_BUF_ERROR && ip >= n) break;
    if (rc != Z_OK && rc != Z_STREAM_END && rc != Z_BUF_ERROR) break;
  }
  inflateEnd(&z);
  if (op < out_n) {
    err = std::string("deflate: ") + (rc == Z_DATA_ERROR ? "corrupt stream" : "data ends early");
    return false;
  }
  return true;
}

void swap_samples(uint8_t* p, size_t n_samples, int bytes) {
  for (size_t i = 0; i < n_samples; ++i, p += bytes) std::reverse(p, p + bytes);
}

// Predictor 2: each sample adds the same sample of the pixel to its left,
// modulo 2^bits.  `row` holds `px` pixels of `spp` samples in host order.
void undo_horizontal(uint8_t* row, size_t px, int spp, int bytes) {
  const size_t n = px * spp;
  switch (bytes) {
    case 1:
      for (size_t i = spp; i < n; ++i) row[i] = (uint8_t)(row[i] + row[i - spp]);
      break;
    case 2: {
      uint16_t* r = reinterpret_cast<uint16_t*>(row);
      for (size_t i = spp; i < n; ++i) r[i] = (uint16_t)(r[i] + r[i - spp]);
      break;
    }
    case 4: {
      uint32_t* r = reinterpret_cast<uint32_t*>(row);
      for (size_t i = spp; i < n; ++i) r[i] = r[i] + r[i - spp];
      break;
    }
    default: {
      uint64_t* r = reinterpret_cast<uint64_t*>(row);
      for (size_t i = spp; i < n; ++i) r[i] = r[i] + r[i - spp];
      break;
    }
  }
}

// Predictor 3 (Adobe floating point): the row's bytes are differenced with a
// stride of spp, and stored as byte planes, most significant plane first.
void undo_float_predictor(uint8_t* row, size_t px, int spp, int bytes, std::vector<uint8_t>& tmp) {
  const size_t wc = px * spp, n = wc * bytes;
  for (size_t i = spp; i < n; ++i) row[i] = (uint8_t)(row[i] + row[i - spp]);
  tmp.assign(row, row + n);
  for (size_t k = 0; k < wc; ++k)
    for (int j = 0; j < bytes; ++j) row[k * bytes + j] = tmp[(size_t)(bytes - 1 - j) * wc + k];  // to little endian
}

struct TiffImage {
  uint32_t w = 0, h = 0, spp = 0, bits = 0, fmt = 1, compression = 1, predictor = 1, photometric = 1;
  std::vector<uint8_t> data;  // interleaved, host order
};

int tiff_decode(const std::vector<uint8_t>& file, const std::string& path, const Ifd& ifd, TiffImage& img) {
  const Bytes& b = ifd.b;
  uint64_t w = ifd.uint_or(kTagWidth, 0), h = ifd.uint_or(kTagHeight, 0);
  if (w == 0 || h == 0 || w > (1u << 30) || h > (1u << 30)) return fail(HM_FILE_BAD_FORMAT, path + ": bad image size");
  const uint64_t spp = ifd.uint_or(kTagSamples, 1);
  if (spp < 1 || spp > HM_CONTROL_FILES_MAX)
    return fail(HM_FILE_UNSUPPORTED, path + ": " + std::to_string(spp) + " samples per pixel");
  const uint64_t bits = ifd.uint_or(kTagBits, 1);
  const uint64_t fmt = ifd.uint_or(kTagSampleFormat, 1);
  for (uint32_t i = 1; i < spp; ++i) {
    uint64_t v;
    if (ifd.uint_at(kTagBits, i, v) && v != bits) return fail(HM_FILE_UNSUPPORTED, path + ": mixed sample sizes");
    if (ifd.uint_at(kTagSampleFormat, i, v) && v != fmt) return fail(HM_FILE_UNSUPPORTED, path + ": mixed formats");
  }
  const bool int_ok = (fmt == 1 || fmt == 2) && (bits == 8 || bits == 16 || bits == 32);
  const bool flt_ok = fmt == 3 && (bits == 32 || bits == 64);
  if (!int_ok && !flt_ok)
    return fail(HM_FILE_UNSUPPORTED,
                path + ": " + std::to_string(bits) + "-bit samples of format " + std::to_string(fmt));
  const uint64_t compression = ifd.uint_or(kTagCompression, 1);
  if (compression != 1 && compression != 5 && compression != 8 && compression != 32946 && compression != 32773)
    return fail(HM_FILE_UNSUPPORTED, path + ": compression " + std::to_string(compression));
  const uint64_t predictor = ifd.uint_or(kTagPredictor, 1);
  if (predictor < 1 || predictor > 3) return fail(HM_FILE_UNSUPPORTED, path + ": predictor " + std::to_string(predictor));
  if (ifd.uint_or(kTagFillOrder, 1) != 1) return fail(HM_FILE_UNSUPPORTED, path + ": reversed fill order");
  const uint64_t planar = ifd.uint_or(kTagPlanar, 1);
  if (planar != 1 && planar != 2) return fail(HM_FILE_BAD_FORMAT, path + ": bad planar configuration");
  const int bytes = (int)(bits / 8);
  const int cspp = planar == 2 ? 1 : (int)spp;  // samples per pixel inside one chunk
  const uint64_t planes = planar == 2 ? spp : 1;

  const bool tiled = ifd.find(kTagTileWidth) != nullptr;
  uint64_t cw, chh, across, down;
  const Entry *offs, *cnts;
  if (tiled) {
    cw = ifd.uint_or(kTagTileWidth, 0);
    chh = ifd.uint_or(kTagTileLength, 0);
    if (cw == 0 || chh == 0 || cw > (1u << 20) || chh > (1u << 20))
      return fail(HM_FILE_BAD_FORMAT, path + ": bad tile size");
    across = (w + cw - 1) / cw;
    down = (h + chh - 1) / chh;
    offs = ifd.find(kTagTileOffsets);
    cnts = ifd.find(kTagTileCounts);
  } else {
    cw = w;
    chh = std::min<uint64_t>(ifd.uint_or(kTagRowsPerStrip, h), h);
    if (chh == 0) chh = h;
    across = 1;
    down = (h + chh - 1) / chh;
    offs = ifd.find(kTagStripOffsets);
    cnts = ifd.find(kTagStripCounts);
  }
  const uint64_t chunks = across * down * planes;
  if (!offs || offs->count < chunks) return fail(HM_FILE_BAD_FORMAT, path + ": missing strip/tile offsets");
  if (cnts && cnts->count < chunks) return fail(HM_FILE_BAD_FORMAT, path + ": missing strip/tile byte counts");
  if (!cnts && compression != 1) return fail(HM_FILE_BAD_FORMAT, path + ": missing strip/tile byte counts");

  img.w = (uint32_t)w;
  img.h = (uint32_t)h;
  img.spp = (uint32_t)spp;
  img.bits = (uint32_t)bits;
  img.fmt = (uint32_t)fmt;
  img.compression = (uint32_t)compression;
  img.predictor = (uint32_t)predictor;
  img.photometric = (uint32_t)ifd.uint_or(kTagPhotometric, 1);
  const size_t px_bytes = (size_t)spp * bytes;
  img.data.assign((size_t)w * h * px_bytes, 0);

  std::vector<uint8_t> chunk, tmp;
  std::string err;
  for (uint64_t pl = 0; pl < planes; ++pl) {
    for (uint64_t ty = 0; ty < down; ++ty) {
      for (uint64_t tx = 0; tx < across; ++tx) {
        const uint64_t idx = pl * across * down + ty * across + tx;
        double dv;
        ifd.value(*offs, (uint32_t)idx, dv);
        const uint64_t off = (uint64_t)dv;
        const uint64_t y0 = ty * chh, x0 = tx * cw;
        // strips end at the image; tiles are always whole
        const uint64_t rows = tiled ? chh : std::min(chh, h - y0);
        const size_t row_bytes = (size_t)cw * cspp * bytes;
        const size_t need = row_bytes * rows;
        uint64_t cnt = need;
        if (cnts) {
          ifd.value(*cnts, (uint32_t)idx, dv);
          cnt = (uint64_t)dv;
        }
        if (off > file.size() || cnt > file.size() - off)
          return fail(HM_FILE_BAD_FORMAT, path + ": strip/tile " + std::to_string(idx) + " lies past the end");
        chunk.assign(need, 0);
        const uint8_t* src = file.data() + off;
        bool ok = true;
        switch (compression) {
          case 1:
            if (cnt < need) {
              err = "uncompressed strip/tile is short";
              ok = false;
            } else {
              memcpy(chunk.data(), src, need);
            }
            break;
          case 5: ok = lzw_decode(src, cnt, chunk.data(), need, err); break;
          case 8:
          case 32946: ok = inflate_exact(src, cnt, chunk.data(), need, err); break;
          default: ok = packbits_decode(src, cnt, chunk.data(), need, err); break;
        }
        if (!ok) return fail(HM_FILE_BAD_FORMAT, path + ": " + err);
        if (predictor == 3) {
          for (uint64_t r = 0; r < rows; ++r) undo_float_predictor(chunk.data() + r * row_bytes, cw, cspp, bytes, tmp);
        } else {
          if (b.be && bytes > 1) swap_samples(chunk.data(), need / bytes, bytes);
          if (predictor == 2)
            for (uint64_t r = 0; r < rows; ++r) undo_horizontal(chunk.data() + r * row_bytes, cw, cspp, bytes);
        }
        // place the chunk
        const uint64_t vr = std::min<uint64_t>(rows, h - y0), vc = std::min<uint64_t>(cw, w - x0);
        for (uint64_t r = 0; r < vr; ++r) {
          const uint8_t* s = chunk.data() + r * row_bytes;
          uint8_t* d = img.data.data() + ((y0 + r) * w + x0) * px_bytes;
          if (planar == 1) {
            memcpy(d, s, vc * px_bytes);
          } else {
            for (uint64_t x = 0; x < vc; ++x) memcpy(d + x * px_bytes + pl * bytes, s + x * bytes, bytes);
          }
        }
      }
    }
  }
  return HM_FILE_OK;
}

int tiff_read(const std::string& path, TiffImage& img) {
  std::vector<uint8_t> file;
  int rc = read_whole(path, file);
  if (rc) return rc;
  Ifd ifd;
  if ((rc = parse_ifd0(file, path, ifd))) return rc;
  return tiff_decode(file, path, ifd, img);
}

void fill_tags(const Ifd& ifd, hmTiffTags& t) {
  memset(&t, 0, sizeof(t));
  t.has_xres = ifd.float_tag(kTagXRes, t.xres);
  t.has_yres = ifd.float_tag(kTagYRes, t.yres);
  t.has_resolution = t.has_xres && t.has_yres;
  t.has_xpos = ifd.float_tag(kTagXPos, t.xpos);
  t.has_ypos = ifd.float_tag(kTagYPos, t.ypos);
  t.width = (uint32_t)ifd.uint_or(kTagWidth, 0);
  t.height = (uint32_t)ifd.uint_or(kTagHeight, 0);
  t.channels = (uint32_t)ifd.uint_or(kTagSamples, 1);
  t.bits = (uint32_t)ifd.uint_or(kTagBits, 1);
  t.sample_format = (uint32_t)ifd.uint_or(kTagSampleFormat, 1);
  t.compression = (uint32_t)ifd.uint_or(kTagCompression, 1);
  t.predictor = (uint32_t)ifd.uint_or(kTagPredictor, 1);
  t.photometric = (uint32_t)ifd.uint_or(kTagPhotometric, 1);
}

int tiff_tags(const std::string& path, hmTiffTags& t) {
  std::vector<uint8_t> file;
  int rc = read_whole(path, file);
  if (rc) return rc;
  Ifd ifd;
  if ((rc = parse_ifd0(file, path, ifd))) return rc;
  fill_tags(ifd, t);
  return HM_FILE_OK;
}

// --- writer ---------------------------------------------------------------
void rational_of(double x, uint32_t& num, uint32_t& den) {
  // best continued-fraction approximation with both terms in 32 bits
  if (!(x > 0)) {
    num = 0;
    den = 1;
    return;
  }
  double h0 = 0, h1 = 1, k0 = 1, k1 = 0, r = x;
  for (int it = 0; it < 64; ++it) {
    const double a = std::floor(r);
    const double h2 = a * h1 + h0, k2 = a * k1 + k0;
    if (h2 > 4294967295.0 || k2 > 4294967295.0) break;
    h0 = h1;
    h1 = h2;
    k0 = k1;
    k1 = k2;
    if (h1 / k1 == x) break;
    const double f = r - a;
    if (f < 1e-12) break;
    r = 1.0 / f;
  }
  if (k1 == 0) {  // x beyond 2^32
    num = 0xFFFFFFFFu;
    den = 1;
    return;
  }
  num = (uint32_t)h1;
  den = (uint32_t)k1;
}

struct TiffWriter {
  std::vector<uint8_t> out;
  void u16(uint32_t v) {
    out.push_back((uint8_t)v);
    out.push_back((uint8_t)(v >> 8));
  }
  void u32(uint32_t v) {
    for (int i = 0; i < 4; ++i) out.push_back((uint8_t)(v >> (8 * i)));
  }
  void put32(size_t at, uint32_t v) {
    for (int i = 0; i < 4; ++i) out[at + i] = (uint8_t)(v >> (8 * i));
  }
};

struct OutTag {
  uint16_t tag, type;
  std::vector<uint32_t> v;  // SHORT/LONG values, or RATIONAL num,den pairs
};

}  // namespace

// ===========================================================================
// PNG
// ===========================================================================
namespace {

uint32_t be32(const uint8_t* p) { return (uint32_t)p[0] << 24 | (uint32_t)p[1] << 16 | (uint32_t)p[2] << 8 | p[3]; }

int paeth(int a, int b, int c) {
  const int p = a + b - c, pa = std::abs(p - a), pb = std::abs(p - b), pc = std::abs(p - c);
  if (pa <= pb && pa <= pc) return a;
  return pb <= pc ? b : c;
}

// Reverses the five PNG filters in place over `rows` scanlines of `stride`
// bytes, each preceded by its filter byte; bpp = bytes per complete pixel
// (at least 1).
bool png_unfilter(uint8_t* data, size_t rows, size_t stride, size_t bpp, std::vector<uint8_t>& out) {
  out.assign(rows * stride, 0);
  for (size_t y = 0; y < rows; ++y) {
    const uint8_t f = data[y * (stride + 1)];
    const uint8_t* in = data + y * (stride + 1) + 1;
    uint8_t* cur = out.data() + y * stride;
    const uint8_t* up = y ? cur - stride : nullptr;
    for (size_t i = 0; i < stride; ++i) {
      const int a = i >= bpp ? cur[i - bpp] : 0;
      const int b = up ? up[i] : 0;
      const int c = (up && i >= bpp) ? up[i - bpp] : 0;
      int v;
      switch (f) {
        case 0: v = in[i]; break;
        case 1: v = in[i] + a; break;
        case 2: v = in[i] + b; break;
        case 3: v = in[i] + ((a + b) >> 1); break;
        case 4: v = in[i] + paeth(a, b, c); break;
        default: return false;
      }
      cur[i] = (uint8_t)v;
    }
  }
  return true;
}

int png_read_seam(const std::string& path, std::vector<uint8_t>& pixels, uint32_t& W, uint32_t& H, int& paletted) {
  std::vector<uint8_t> file;
  int rc = read_whole(path, file);
  if (rc) return rc;
  static const uint8_t sig[8] = {137, 80, 78, 71, 13, 10, 26, 10};
  if (file.size() < 8 || memcmp(file.data(), sig, 8) != 0) return fail(HM_FILE_BAD_FORMAT, path + ": not a PNG");
  size_t at = 8;
  bool have_hdr = false, ended = false;
  uint32_t w = 0, h = 0;
  int depth = 0, ctype = 0, interlace = 0;
  std::vector<uint8_t> idat;
  while (at + 12 <= file.size()) {
    const uint32_t len = be32(&file[at]);
    if (len > file.size() - at - 12) return fail(HM_FILE_BAD_FORMAT, path + ": truncated chunk");
    const uint8_t* type = &file[at + 4];
    const uint8_t* data = &file[at + 8];
    const uint32_t crc = be32(&file[at + 8 + len]);
    if ((uint32_t)crc32(crc32(0L, Z_NULL, 0), type, len + 4) != crc)
      return fail(HM_FILE_BAD_FORMAT, path + ": chunk CRC mismatch");
    const std::string t(reinterpret_cast<const char*>(type), 4);
    if (t == "IHDR") {
      if (len != 13) return fail(HM_FILE_BAD_FORMAT, path + ": bad IHDR");
      w = be32(data);
      h = be32(data + 4);
      depth = data[8];
      ctype = data[9];
      interlace = data[12];
      if (data[10] != 0 || data[11] != 0 || interlace > 1)
        return fail(HM_FILE_BAD_FORMAT, path + ": bad IHDR method fields");
      have_hdr = true;
    } else if (t == "IDAT") {
      if (!have_hdr) return fail(HM_FILE_BAD_FORMAT, path + ": IDAT before IHDR");
      idat.insert(idat.end(), data, data + len);
    } else if (t == "IEND") {
      ended = true;
      break;
    }
    at += 12 + (size_t)len;
  }
  if (!have_hdr || !ended) return fail(HM_FILE_BAD_FORMAT, path + ": missing IHDR or IEND");
  if (w == 0 || h == 0 || w > (1u << 30) || h > (1u << 30)) return fail(HM_FILE_BAD_FORMAT, path + ": bad size");
  const bool valid = (ctype == 0 && (depth == 1 || depth == 2 || depth == 4 || depth == 8 || depth == 16)) ||
                     (ctype == 3 && (depth == 1 || depth == 2 || depth == 4 || depth == 8)) ||
                     ((ctype == 2 || ctype == 4 || ctype == 6) && (depth == 8 || depth == 16));
  if (!valid) return fail(HM_FILE_BAD_FORMAT, path + ": invalid colour type / bit depth");
  if (ctype == 2 || ctype == 6 || (ctype == 3 && depth != 8))
    return fail(HM_FILE_UNSUPPORTED, path + ": colour seam PNG (colour type " + std::to_string(ctype) + ", " +
                                         std::to_string(depth) + " bits); use 8-bit paletted or grayscale");
  const int channels = ctype == 4 ? 2 : 1;
  const size_t bits_pp = (size_t)channels * depth;
  const size_t bpp = std::max<size_t>(1, bits_pp / 8);
  struct Pass {
    int x0, y0, dx, dy;
  };
  static const Pass adam7[7] = {{0, 0, 8, 8}, {4, 0, 8, 8}, {0, 4, 4, 8}, {2, 0, 4, 4},
                                {0, 2, 2, 4}, {1, 0, 2, 2}, {0, 1, 1, 2}};
  static const Pass whole[1] = {{0, 0, 1, 1}};
  const Pass* passes = interlace ? adam7 : whole;
  const int n_passes = interlace ? 7 : 1;
  size_t total = 0;
  size_t pw[7], ph[7];
  for (int p = 0; p < n_passes; ++p) {
    pw[p] = w > (uint32_t)passes[p].x0 ? (w - passes[p].x0 + passes[p].dx - 1) / passes[p].dx : 0;
    ph[p] = h > (uint32_t)passes[p].y0 ? (h - passes[p].y0 + passes[p].dy - 1) / passes[p].dy : 0;
    if (pw[p] && ph[p]) total += ph[p] * (1 + (pw[p] * bits_pp + 7) / 8);
  }
  std::vector<uint8_t> raw(total);
  std::string err;
  if (!inflate_exact(idat.data(), idat.size(), raw.data(), total, err)) return fail(HM_FILE_BAD_FORMAT, path + ": " + err);
  pixels.assign((size_t)w * h, 0);
  std::vector<uint8_t> rows;
  size_t pos = 0;
  const int maxv = (1 << std::min(depth, 8)) - 1;
  for (int p = 0; p < n_passes; ++p) {
    if (!pw[p] || !ph[p]) continue;
    const size_t stride = (pw[p] * bits_pp + 7) / 8;
    if (!png_unfilter(raw.data() + pos, ph[p], stride, bpp, rows))
      return fail(HM_FILE_BAD_FORMAT, path + ": bad filter type");
    pos += ph[p] * (stride + 1);
    for (size_t y = 0; y < ph[p]; ++y) {
      const uint8_t* r = rows.data() + y * stride;
      uint8_t* dst = pixels.data() + (size_t)(passes[p].y0 + y * passes[p].dy) * w;
      for (size_t x = 0; x < pw[p]; ++x) {
        int v;
        if (depth < 8) {
          const size_t bit = x * depth;
          v = (r[bit >> 3] >> (8 - depth - (bit & 7))) & maxv;
          if (ctype == 0) v = v * 255 / maxv;  // 1, 2, 4 bits -> 0..255 (bit replication)
        } else {
          v = r[x * channels * (depth / 8)];  // first (gray) sample; 16 bits: its high byte
        }
        dst[passes[p].x0 + x * passes[p].dx] = (uint8_t)v;
      }
    }
  }
  W = w;
  H = h;
  paletted = ctype == 3;
  return HM_FILE_OK;
}

}  // namespace

// ===========================================================================
// Rig loaders
// ===========================================================================
struct hmControlFiles {
  int n = 0;
  std::vector<uint16_t> mx[HM_CONTROL_FILES_MAX], my[HM_CONTROL_FILES_MAX];
  uint32_t w[HM_CONTROL_FILES_MAX] = {}, h[HM_CONTROL_FILES_MAX] = {};
  float xpos[HM_CONTROL_FILES_MAX] = {}, ypos[HM_CONTROL_FILES_MAX] = {};
  uint32_t found = 0;
  std::vector<uint8_t> seam;
  uint32_t sw = 0, sh = 0;
  int paletted = 0;
};

namespace {

std::string idx_name(const std::string& dir, int i, const char* suffix) {
  char buf[64];
  snprintf(buf, sizeof(buf), "mapping_%04d%s.tif", i, suffix);
  return dir + buf;
}

int load_map(const std::string& path, std::vector<uint16_t>& out, uint32_t& w, uint32_t& h) {
  if (!file_exists(path)) return fail(HM_FILE_NOT_FOUND, "unable to load remap file " + path);
  TiffImage img;
  const int rc = tiff_read(path, img);
  if (rc) return rc;
  if (img.spp != 1 || img.bits != 16 || img.fmt != 1)
    return fail(HM_FILE_WRONG_TYPE, path + ": remap tables must be one-sample uint16 (found " +
                                        std::to_string(img.spp) + " x " + std::to_string(img.bits) + "-bit, format " +
                                        std::to_string(img.fmt) + ")");
  out.resize((size_t)img.w * img.h);
  memcpy(out.data(), img.data.data(), out.size() * 2);
  w = img.w;
  h = img.h;
  return HM_FILE_OK;
}

// Placement tags of mapping_%04d.tif: position x resolution in float.  The
// 2- and 3-image loaders take the resolutions only when both are present
// (controlMasks.cpp:75-80, controlMasks3.cpp:37-42); the N-image loader
// takes each on its own (controlMasksN.cpp:29-32).  An unreadable file gives
// (0, 0).
bool load_position(const std::string& path, bool both_resolutions, float& x, float& y) {
  x = y = 0.0f;
  std::vector<uint8_t> file;
  if (read_whole(path, file) != HM_FILE_OK) return false;
  Ifd ifd;
  if (parse_ifd0(file, path, ifd) != HM_FILE_OK) return false;
  hmTiffTags t;
  fill_tags(ifd, t);
  float xres = 0.0f, yres = 0.0f;
  if (both_resolutions ? t.has_resolution : true) {
    if (t.has_xres) xres = t.xres;
    if (t.has_yres) yres = t.yres;
  }
  const float xp = t.has_xpos ? t.xpos : 0.0f, yp = t.has_ypos ? t.ypos : 0.0f;
  x = xp * xres;
  y = yp * yres;
  return true;
}

void normalize_positions(hmControlFiles& f) {  // controlMasks.cpp:42-57
  float mx = std::numeric_limits<float>::max(), my = std::numeric_limits<float>::max();
  for (int i = 0; i < f.n; ++i) {
    mx = std::min(mx, f.xpos[i]);
    my = std::min(my, f.ypos[i]);
  }
  for (int i = 0; i < f.n; ++i) {
    f.xpos[i] -= mx;
    f.ypos[i] -= my;
  }
}

int load_seam(const std::string& path, hmControlFiles& f) {
  if (!file_exists(path)) return fail(HM_FILE_NOT_FOUND, "could not find seam file " + path);
  return png_read_seam(path, f.seam, f.sw, f.sh, f.paletted);
}

// controlMasks.cpp:225-251: the maximum value becomes 0, then the minimum
// value becomes 1 (both located on the file's values; a constant seam ends
// all 1).
void invert_two_image_seam(std::vector<uint8_t>& s) {
  if (s.empty()) return;
  const auto mm = std::minmax_element(s.begin(), s.end());
  const uint8_t lo = *mm.first, hi = *mm.second;
  for (uint8_t& v : s) {
    const uint8_t o = v;
    if (o == hi) v = 0;
    if (o == lo) v = 1;
  }
}

// controlMasksN.cpp:162-190: exactly n distinct values, mapped to their rank
// when they are not already 0..n-1.
int relabel_n(hmControlFiles& f, const std::string& path) {
  bool present[256] = {};
  for (uint8_t v : f.seam) present[v] = true;
  int lut[256];
  int count = 0, lo = -1, hi = -1;
  for (int v = 0; v < 256; ++v)
    if (present[v]) {
      lut[v] = count++;
      if (lo < 0) lo = v;
      hi = v;
    }
  if (count == 0 || count != f.n)
    return fail(HM_FILE_SEAM_CLASSES, path + ": seam has " + std::to_string(count) + " classes, expected " +
                                          std::to_string(f.n));
  if (lo != 0 || hi != f.n - 1)
    for (uint8_t& v : f.seam) v = (uint8_t)lut[v];
  return HM_FILE_OK;
}

int load_rig(std::string dir, int kind, int n_images, hmControlFiles& f) {
  if (!dir.empty() && dir.back() != '/') dir += '/';
  const int n = kind == HM_CONTROL_TWO ? 2 : kind == HM_CONTROL_THREE ? 3 : n_images;
  if (n < 1 || n > HM_CONTROL_FILES_MAX) return fail(HM_FILE_INVALID_ARGUMENT, "n_images out of range");
  f.n = n;
  const bool both_res = kind != HM_CONTROL_N;
  for (int i = 0; i < n; ++i) {
    int rc = load_map(idx_name(dir, i, "_x"), f.mx[i], f.w[i], f.h[i]);
    if (rc) return rc;
    uint32_t w2, h2;
    rc = load_map(idx_name(dir, i, "_y"), f.my[i], w2, h2);
    if (rc) return rc;
    if (w2 != f.w[i] || h2 != f.h[i])
      return fail(HM_FILE_MISMATCH, "remap tables of image " + std::to_string(i) + " differ in size");
  }
  const std::string seam_path = dir + "seam_file.png";
  int rc = load_seam(seam_path, f);
  if (rc) return rc;
  if (kind == HM_CONTROL_TWO) {
    invert_two_image_seam(f.seam);
  } else if (kind == HM_CONTROL_THREE) {
    if (!f.paletted) return fail(HM_FILE_UNSUPPORTED, seam_path + ": the 3-image seam must be 8-bit paletted");
  } else {
    if ((rc = relabel_n(f, seam_path))) return rc;
  }
  for (int i = 0; i < n; ++i)
    if (load_position(idx_name(dir, i, ""), both_res, f.xpos[i], f.ypos[i])) f.found |= 1u << i;
  normalize_positions(f);
  return HM_FILE_OK;
}

void to_raster(std::vector<uint8_t>&& bytes, uint32_t w, uint32_t h, uint32_t c, uint32_t bits, uint32_t fmt,
               hmRaster* out) {
  out->bytes = bytes.size();
  out->data = malloc(bytes.size() ? bytes.size() : 1);
  if (out->data && !bytes.empty()) memcpy(out->data, bytes.data(), bytes.size());
  out->width = w;
  out->height = h;
  out->channels = c;
  out->bits = bits;
  out->sample_format = fmt;
}

}  // namespace

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" const char* hmFilesLastError(void) { return g_error.c_str(); }

extern "C" int hmControlFilesLoad(const char* dir, int kind, int n_images, hmControlFiles** out) {
  if (!dir || !out) return fail(HM_FILE_INVALID_ARGUMENT, "null argument");
  *out = nullptr;
  if (kind != HM_CONTROL_TWO && kind != HM_CONTROL_THREE && kind != HM_CONTROL_N)
    return fail(HM_FILE_INVALID_ARGUMENT, "unknown loader kind " + std::to_string(kind));
  hmControlFiles* f = new hmControlFiles;
  const int rc = load_rig(dir, kind, n_images, *f);
  if (rc) {
    delete f;
    return rc;
  }
  *out = f;
  return HM_FILE_OK;
}

extern "C" int hmControlFilesGetInfo(const hmControlFiles* f, hmControlFilesInfo* info) {
  if (!f || !info) return fail(HM_FILE_INVALID_ARGUMENT, "null argument");
  memset(info, 0, sizeof(*info));
  info->n_images = f->n;
  for (int i = 0; i < f->n; ++i) {
    info->map_w[i] = f->w[i];
    info->map_h[i] = f->h[i];
    info->map_x[i] = f->mx[i].data();
    info->map_y[i] = f->my[i].data();
    info->xpos[i] = f->xpos[i];
    info->ypos[i] = f->ypos[i];
  }
  info->position_files_found = f->found;
  info->seam_w = f->sw;
  info->seam_h = f->sh;
  info->seam = f->seam.data();
  info->seam_paletted = f->paletted;
  return HM_FILE_OK;
}

extern "C" void hmControlFilesDestroy(hmControlFiles* f) { delete f; }

extern "C" int hmTiffRead(const char* path, int opencv_order, hmRaster* out) {
  if (!path || !out) return fail(HM_FILE_INVALID_ARGUMENT, "null argument");
  memset(out, 0, sizeof(*out));
  TiffImage img;
  const int rc = tiff_read(path, img);
  if (rc) return rc;
  if (opencv_order && (img.spp == 3 || img.spp == 4)) {
    const size_t b = img.bits / 8, px = (size_t)img.spp * b;
    for (size_t i = 0; i < img.data.size(); i += px)
      for (size_t k = 0; k < b; ++k) std::swap(img.data[i + k], img.data[i + 2 * b + k]);
  }
  to_raster(std::move(img.data), img.w, img.h, img.spp, img.bits, img.fmt, out);
  return out->data ? HM_FILE_OK : fail(HM_FILE_INVALID_ARGUMENT, "out of host memory");
}

extern "C" int hmTiffReadTags(const char* path, hmTiffTags* tags) {
  if (!path || !tags) return fail(HM_FILE_INVALID_ARGUMENT, "null argument");
  return tiff_tags(path, *tags);
}

extern "C" int hmPngReadSeam(const char* path, hmRaster* out, int* paletted) {
  if (!path || !out) return fail(HM_FILE_INVALID_ARGUMENT, "null argument");
  memset(out, 0, sizeof(*out));
  std::vector<uint8_t> px;
  uint32_t w = 0, h = 0;
  int pal = 0;
  const int rc = png_read_seam(path, px, w, h, pal);
  if (rc) return rc;
  if (paletted) *paletted = pal;
  to_raster(std::move(px), w, h, 1, 8, HM_SAMPLE_UINT, out);
  return out->data ? HM_FILE_OK : fail(HM_FILE_INVALID_ARGUMENT, "out of host memory");
}

extern "C" void hmRasterFree(hmRaster* r) {
  if (!r) return;
  free(r->data);
  memset(r, 0, sizeof(*r));
}

extern "C" int hmTiffWrite(const char* path, const hmRaster* r, const hmTiffTags* tags, int opencv_order) {
  if (!path || !r || !r->data) return fail(HM_FILE_INVALID_ARGUMENT, "null argument");
  const uint32_t c = r->channels, bits = r->bits, fmt = r->sample_format;
  const bool int_ok = (fmt == HM_SAMPLE_UINT || fmt == HM_SAMPLE_INT) && (bits == 8 || bits == 16 || bits == 32);
  const bool flt_ok = fmt == HM_SAMPLE_FLOAT && (bits == 32 || bits == 64);
  if (r->width == 0 || r->height == 0 || c < 1 || c > HM_CONTROL_FILES_MAX || (!int_ok && !flt_ok))
    return fail(HM_FILE_INVALID_ARGUMENT, "unsupported raster for TIFF");
  const size_t bps = bits / 8, px = (size_t)c * bps, row = px * r->width, total = row * r->height;
  if (r->bytes < total) return fail(HM_FILE_INVALID_ARGUMENT, "raster smaller than its size");
  if (tags) {
    if ((tags->has_xres && !(tags->xres >= 0)) || (tags->has_yres && !(tags->yres >= 0)) ||
        (tags->has_xpos && !(tags->xpos >= 0)) || (tags->has_ypos && !(tags->ypos >= 0)))
      return fail(HM_FILE_INVALID_ARGUMENT, "TIFF position/resolution tags must be non-negative");
  }
  const uint32_t rps = (uint32_t)std::max<size_t>(1, std::min<size_t>(r->height, (1u << 16) / std::max<size_t>(row, 1)));
  const uint32_t strips = (r->height + rps - 1) / rps;
  if (total > 0xFFFF0000u) return fail(HM_FILE_INVALID_ARGUMENT, "raster too large for a classic TIFF");

  TiffWriter w;
  w.out.reserve(total + 4096);
  w.out.insert(w.out.end(), {'I', 'I', 42, 0});
  w.u32(0);  // IFD offset, patched below
  std::vector<uint32_t> offsets(strips), counts(strips);
  const uint8_t* src = static_cast<const uint8_t*>(r->data);
  const bool swap = opencv_order && (c == 3 || c == 4);
  for (uint32_t s = 0; s < strips; ++s) {
    const uint32_t y0 = s * rps, y1 = std::min(r->height, y0 + rps);
    offsets[s] = (uint32_t)w.out.size();
    counts[s] = (uint32_t)((y1 - y0) * row);
    const size_t at = w.out.size();
    w.out.insert(w.out.end(), src + y0 * row, src + y1 * row);
    if (swap)
      for (size_t i = at; i < w.out.size(); i += px)
        for (size_t k = 0; k < bps; ++k) std::swap(w.out[i + k], w.out[i + 2 * bps + k]);
  }
  if (w.out.size() & 1) w.out.push_back(0);

  const bool rgb = c == 3 || c == 4;
  std::vector<OutTag> t;
  t.push_back({kTagWidth, 4, {r->width}});
  t.push_back({kTagHeight, 4, {r->height}});
  t.push_back({kTagBits, 3, std::vector<uint32_t>(c, bits)});
  t.push_back({kTagCompression, 3, {1}});
  t.push_back({kTagPhotometric, 3, {rgb ? 2u : 1u}});
  t.push_back({kTagStripOffsets, 4, offsets});
  t.push_back({kTagSamples, 3, {c}});
  t.push_back({kTagRowsPerStrip, 4, {rps}});
  t.push_back({kTagStripCounts, 4, counts});
  uint32_t num, den;
  if (tags && tags->has_xres) {
    rational_of(tags->xres, num, den);
    t.push_back({kTagXRes, 5, {num, den}});
  }
  if (tags && tags->has_yres) {
    rational_of(tags->yres, num, den);
    t.push_back({kTagYRes, 5, {num, den}});
  }
  t.push_back({kTagPlanar, 3, {1}});
  if (tags && tags->has_xpos) {
    rational_of(tags->xpos, num, den);
    t.push_back({kTagXPos, 5, {num, den}});
  }
  if (tags && tags->has_ypos) {
    rational_of(tags->ypos, num, den);
    t.push_back({kTagYPos, 5, {num, den}});
  }
  if (tags && (tags->has_xres || tags->has_yres)) t.push_back({kTagResUnit, 3, {2}});
  const uint32_t extras = rgb ? c - 3 : c - 1;
  if (extras) t.push_back({kTagExtraSamples, 3, std::vector<uint32_t>(extras, rgb ? 2u : 0u)});
  t.push_back({kTagSampleFormat, 3, std::vector<uint32_t>(c, fmt)});

  const size_t ifd_at = w.out.size();
  w.put32(4, (uint32_t)ifd_at);
  size_t extra_at = ifd_at + 2 + t.size() * 12 + 4;
  std::vector<uint8_t> extra;
  w.u16((uint32_t)t.size());
  for (const OutTag& x : t) {
    const uint32_t count = x.type == 5 ? (uint32_t)x.v.size() / 2 : (uint32_t)x.v.size();
    const size_t sz = x.type == 3 ? 2 : x.type == 4 ? 4 : 8;
    w.u16(x.tag);
    w.u16(x.type);
    w.u32(count);
    std::vector<uint8_t> bytes;
    for (uint32_t v : x.v) {
      if (x.type == 3) {
        bytes.push_back((uint8_t)v);
        bytes.push_back((uint8_t)(v >> 8));
      } else {
        for (int i = 0; i < 4; ++i) bytes.push_back((uint8_t)(v >> (8 * i)));
      }
    }
    if (sz * count <= 4) {
      bytes.resize(4, 0);
      w.out.insert(w.out.end(), bytes.begin(), bytes.end());
    } else {
      w.u32((uint32_t)(extra_at + extra.size()));
      extra.insert(extra.end(), bytes.begin(), bytes.end());
      if (extra.size() & 1) extra.push_back(0);
    }
  }
  w.u32(0);  // no next IFD
  w.out.insert(w.out.end(), extra.begin(), extra.end());

  FILE* f = fopen(path, "wb");
  if (!f) return fail(HM_FILE_WRITE_FAILED, std::string("cannot create ") + path);
  const bool ok = fwrite(w.out.data(), 1, w.out.size(), f) == w.out.size();
  const bool closed = fclose(f) == 0;
  if (!ok || !closed) return fail(HM_FILE_WRITE_FAILED, std::string("write failed on ") + path);
  return HM_FILE_OK;
}
