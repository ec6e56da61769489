// Image output as the reference's save_image writes it (src/util.cpp:111-150):
// pixel (x, y) of the row-major float RGB buffer (y = 0 the first row the
// renderer produced) goes to image position (w-1-x, h-1-y), i.e. both axes
// flipped.
//   PNG: 8-bit RGB, each channel clamp(int(255 * v), 0, 255) -- no gamma, the
//        float-to-int conversion truncates, as png::rgb_pixel(clamp(int(...)))
//        does.  Written with zlib (deflate, crc32).
//   EXR: RGBA half (A = 1), as Imf::RgbaOutputFile(WRITE_RGBA) writes it;
//        here uncompressed scanlines (a valid OpenEXR 2 single-part file).
//        Float to half rounds to nearest even.
#include <zlib.h>

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/gi_hip.h"

namespace {

void put_be32(std::vector<unsigned char> &b, uint32_t v) {
  b.push_back((unsigned char)(v >> 24));
  b.push_back((unsigned char)(v >> 16));
  b.push_back((unsigned char)(v >> 8));
  b.push_back((unsigned char)v);
}

void png_chunk(std::vector<unsigned char> &out, const char *type, const std::vector<unsigned char> &data) {
  put_be32(out, (uint32_t)data.size());
  const size_t start = out.size();
  out.insert(out.end(), type, type + 4);
  out.insert(out.end(), data.begin(), data.end());
  const uLong crc = crc32(0L, out.data() + start, (uInt)(out.size() - start));
  put_be32(out, (uint32_t)crc);
}

int clamp255(float v) {
  // int(255*v): truncation toward zero (NaN -> 0), then clamp to [0, 255]
  const float s = 255.0f * v;
  if (!(s > 0.0f)) return 0;
  if (s >= 255.0f) return 255;
  return (int)s;
}

// float -> IEEE half, round to nearest even (overflow -> inf, NaN kept).
uint16_t to_half(float f) {
  uint32_t x;
  std::memcpy(&x, &f, 4);
  const uint32_t sign = (x >> 16) & 0x8000u;
  const uint32_t absx = x & 0x7fffffffu;
  if (absx >= 0x7f800000u) return (uint16_t)(sign | 0x7c00u | (absx > 0x7f800000u ? 0x200u : 0u));
  if (absx >= 0x477ff000u) return (uint16_t)(sign | 0x7c00u);  // rounds to >= 65520: inf
  if (absx < 0x38800000u) {  // below the smallest normal half 2^-14: subnormal or zero
    if (absx < 0x33000000u) return (uint16_t)sign;  // < 2^-25: rounds to 0
    const uint32_t e = absx >> 23, m = (absx & 0x7fffffu) | 0x800000u;
    const uint32_t shift = 126u - e;  // value = m * 2^(e-150); half subnormal unit 2^-24
    uint32_t h = m >> shift;
    const uint32_t rem = m & ((1u << shift) - 1u), halfway = 1u << (shift - 1u);
    if (rem > halfway || (rem == halfway && (h & 1u))) ++h;
    return (uint16_t)(sign | h);
  }
  uint32_t h = ((absx - 0x38000000u) >> 13);  // rebias exponent 127 -> 15
  const uint32_t rem = absx & 0x1fffu;
  if (rem > 0x1000u || (rem == 0x1000u && (h & 1u))) ++h;
  return (uint16_t)(sign | h);
}

template <class T>
void put_le(std::vector<unsigned char> &b, T v) {
  unsigned char raw[sizeof(T)];
  std::memcpy(raw, &v, sizeof(T));
  b.insert(b.end(), raw, raw + sizeof(T));
}

void put_str(std::vector<unsigned char> &b, const char *s) { b.insert(b.end(), s, s + std::strlen(s) + 1); }

void exr_attr(std::vector<unsigned char> &b, const char *name, const char *type, const std::vector<unsigned char> &v) {
  put_str(b, name);
  put_str(b, type);
  put_le<int32_t>(b, (int32_t)v.size());
  b.insert(b.end(), v.begin(), v.end());
}

bool write_file(const char *path, const std::vector<unsigned char> &b) {
  std::FILE *f = std::fopen(path, "wb");
  if (!f) return false;
  const bool ok = std::fwrite(b.data(), 1, b.size(), f) == b.size();
  return std::fclose(f) == 0 && ok;
}

}  // namespace

extern "C" {

int gi_write_png(const char *path, const float *rgb, int32_t w, int32_t h) {
  if (!path || !rgb || w <= 0 || h <= 0) return GI_ERR_INVALID_ARG;
  // filtered scanlines: filter byte 0 + RGB
  std::vector<unsigned char> raw((size_t)h * (1 + 3 * (size_t)w));
  for (int32_t yo = 0; yo < h; ++yo) {
    unsigned char *row = raw.data() + (size_t)yo * (1 + 3 * (size_t)w);
    row[0] = 0;
    const int32_t y = h - 1 - yo;
    for (int32_t xo = 0; xo < w; ++xo) {
      const int32_t x = w - 1 - xo;
      const float *p = rgb + 3 * ((size_t)y * w + x);
      for (int c = 0; c < 3; ++c) row[1 + 3 * xo + c] = (unsigned char)clamp255(p[c]);
    }
  }
  uLongf zlen = compressBound((uLong)raw.size());
  std::vector<unsigned char> z(zlen);
  if (compress2(z.data(), &zlen, raw.data(), (uLong)raw.size(), 6) != Z_OK) return GI_ERR_OUT_OF_MEMORY;
  z.resize(zlen);
  std::vector<unsigned char> out = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1a, '\n'};
  std::vector<unsigned char> ihdr;
  put_be32(ihdr, (uint32_t)w);
  put_be32(ihdr, (uint32_t)h);
  ihdr.insert(ihdr.end(), {8, 2, 0, 0, 0});  // 8-bit, truecolour, deflate, no filter method, no interlace
  png_chunk(out, "IHDR", ihdr);
  png_chunk(out, "IDAT", z);
  png_chunk(out, "IEND", {});
  return write_file(path, out) ? GI_OK : GI_ERR_INVALID_ARG;
}

int gi_write_exr(const char *path, const float *rgb, int32_t w, int32_t h) {
  if (!path || !rgb || w <= 0 || h <= 0) return GI_ERR_INVALID_ARG;
  std::vector<unsigned char> b;
  put_le<uint32_t>(b, 20000630u);  // magic 0x762f3101
  put_le<uint32_t>(b, 2u);         // version 2, single-part scanline
  std::vector<unsigned char> ch;
  for (const char *name : {"A", "B", "G", "R"}) {  // alphabetical
    put_str(ch, name);
    put_le<int32_t>(ch, 1);  // HALF
    ch.insert(ch.end(), {0, 0, 0, 0});  // pLinear + reserved
    put_le<int32_t>(ch, 1);
    put_le<int32_t>(ch, 1);
  }
  ch.push_back(0);
  exr_attr(b, "channels", "chlist", ch);
  exr_attr(b, "compression", "compression", {0});
  std::vector<unsigned char> box;
  put_le<int32_t>(box, 0);
  put_le<int32_t>(box, 0);
  put_le<int32_t>(box, w - 1);
  put_le<int32_t>(box, h - 1);
  exr_attr(b, "dataWindow", "box2i", box);
  exr_attr(b, "displayWindow", "box2i", box);
  exr_attr(b, "lineOrder", "lineOrder", {0});
  std::vector<unsigned char> one;
  put_le<float>(one, 1.0f);
  exr_attr(b, "pixelAspectRatio", "float", one);
  std::vector<unsigned char> v2;
  put_le<float>(v2, 0.0f);
  put_le<float>(v2, 0.0f);
  exr_attr(b, "screenWindowCenter", "v2f", v2);
  exr_attr(b, "screenWindowWidth", "float", one);
  b.push_back(0);  // end of header
  const size_t table = b.size();
  const size_t line_bytes = 4 * 2 * (size_t)w;
  b.resize(table + 8 * (size_t)h);
  for (int32_t yo = 0; yo < h; ++yo) {
    const uint64_t off = b.size();
    std::memcpy(b.data() + table + 8 * (size_t)yo, &off, 8);
    put_le<int32_t>(b, yo);
    put_le<int32_t>(b, (int32_t)line_bytes);
    const int32_t y = h - 1 - yo;
    for (int c : {3, 2, 1, 0}) {  // A, B, G, R
      for (int32_t xo = 0; xo < w; ++xo) {
        const int32_t x = w - 1 - xo;
        const float v = c == 3 ? 1.0f : rgb[3 * ((size_t)y * w + x) + c];
        put_le<uint16_t>(b, to_half(v));
      }
    }
  }
  return write_file(path, b) ? GI_OK : GI_ERR_INVALID_ARG;
}

}  // extern "C"

// ---------------------------------------------------------------- reading --
// Texture images for map_Kd (the reference decodes them with rta's image
// library, which is external): PNG (ISO/IEC 15948: grey, grey+alpha, RGB,
// RGBA at 8 or 16 bits, palette at 1/2/4/8 bits with tRNS, not interlaced;
// the five scanline filters) and binary PPM (P6, maxval < 65536).  Output
// RGBA8, row 0 = the file's first (top) row, 16-bit samples keep their high
// byte, missing alpha is 255.
namespace {

uint32_t be32(const unsigned char *p) {
  return ((uint32_t)p[0] << 24) | ((uint32_t)p[1] << 16) | ((uint32_t)p[2] << 8) | (uint32_t)p[3];
}

bool read_file(const char *path, std::vector<unsigned char> &b) {
  std::FILE *f = std::fopen(path, "rb");
  if (!f) return false;
  unsigned char buf[65536];
  size_t n;
  while ((n = std::fread(buf, 1, sizeof(buf), f)) > 0) b.insert(b.end(), buf, buf + n);
  std::fclose(f);
  return true;
}

int paeth(int a, int b, int c) {
  const int p = a + b - c, pa = std::abs(p - a), pb = std::abs(p - b), pc = std::abs(p - c);
  if (pa <= pb && pa <= pc) return a;
  return pb <= pc ? b : c;
}

// decode == false: only the size
int decode_png(const std::vector<unsigned char> &b, int32_t &w, int32_t &h, uint8_t *rgba, bool decode) {
  static const unsigned char sig[8] = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1a, '\n'};
  if (b.size() < 8 || std::memcmp(b.data(), sig, 8) != 0) return GI_ERR_INVALID_ARG;
  size_t pos = 8;
  int depth = 0, ctype = -1, interlace = 0;
  std::vector<unsigned char> idat, plte, trns;
  bool ihdr = false;
  while (pos + 12 <= b.size()) {
    const uint32_t len = be32(&b[pos]);
    if (len > b.size() - pos - 12) return GI_ERR_INVALID_ARG;
    const unsigned char *type = &b[pos + 4], *data = &b[pos + 8];
    if (!std::memcmp(type, "IHDR", 4) && len >= 13) {
      w = (int32_t)be32(data);
      h = (int32_t)be32(data + 4);
      depth = data[8];
      ctype = data[9];
      interlace = data[12];
      ihdr = true;
    } else if (!std::memcmp(type, "PLTE", 4)) {
      plte.assign(data, data + len);
    } else if (!std::memcmp(type, "tRNS", 4)) {
      trns.assign(data, data + len);
    } else if (!std::memcmp(type, "IDAT", 4)) {
      idat.insert(idat.end(), data, data + len);
    } else if (!std::memcmp(type, "IEND", 4)) {
      break;
    }
    pos += 12 + (size_t)len;
  }
  if (!ihdr || w <= 0 || h <= 0 || (int64_t)w * h > ((int64_t)1 << 28)) return GI_ERR_INVALID_ARG;
  int channels;
  switch (ctype) {
    case 0: channels = 1; break;
    case 2: channels = 3; break;
    case 3: channels = 1; break;
    case 4: channels = 2; break;
    case 6: channels = 4; break;
    default: return GI_ERR_INVALID_ARG;
  }
  const bool depth_ok = (ctype == 3) ? (depth == 1 || depth == 2 || depth == 4 || depth == 8)
                                     : (depth == 8 || depth == 16 || (ctype == 0 && depth < 8 && (depth == 1 || depth == 2 || depth == 4)));
  if (!depth_ok || interlace != 0) return GI_ERR_INVALID_ARG;
  if (!decode) return GI_OK;
  if (ctype == 3 && plte.size() < 3) return GI_ERR_INVALID_ARG;
  const size_t bits_pp = (size_t)channels * depth;
  const size_t row_bytes = ((size_t)w * bits_pp + 7) / 8;
  const size_t bpp = std::max<size_t>(1, bits_pp / 8);  // filter unit
  std::vector<unsigned char> raw((size_t)h * (row_bytes + 1));
  uLongf rl = (uLongf)raw.size();
  if (uncompress(raw.data(), &rl, idat.data(), (uLong)idat.size()) != Z_OK || rl != raw.size())
    return GI_ERR_INVALID_ARG;
  std::vector<unsigned char> prev(row_bytes, 0), cur(row_bytes);
  for (int32_t y = 0; y < h; ++y) {
    const unsigned char *r = raw.data() + (size_t)y * (row_bytes + 1);
    const int ft = r[0];
    for (size_t i = 0; i < row_bytes; ++i) {
      const int x = r[1 + i], a = i >= bpp ? cur[i - bpp] : 0, up = prev[i], c = i >= bpp ? prev[i - bpp] : 0;
      int v;
      switch (ft) {
        case 0: v = x; break;
        case 1: v = x + a; break;
        case 2: v = x + up; break;
        case 3: v = x + ((a + up) >> 1); break;
        case 4: v = x + paeth(a, up, c); break;
        default: return GI_ERR_INVALID_ARG;
      }
      cur[i] = (unsigned char)v;
    }
    uint8_t *o = rgba + 4 * (size_t)y * w;
    for (int32_t x = 0; x < w; ++x) {
      auto sample = [&](int ch) -> int {  // 8-bit value of channel ch of pixel x
        if (depth == 16) return cur[((size_t)x * channels + ch) * 2];
        if (depth == 8) return cur[(size_t)x * channels + ch];
        const size_t bit = (size_t)x * depth;  // sub-byte: one channel
        const int v = (cur[bit / 8] >> (8 - depth - (int)(bit % 8))) & ((1 << depth) - 1);
        return ctype == 3 ? v : v * 255 / ((1 << depth) - 1);
      };
      int rr, gg, bb, aa = 255;
      if (ctype == 3) {
        const int k = sample(0);
        if ((size_t)(3 * k + 2) >= plte.size()) return GI_ERR_INVALID_ARG;
        rr = plte[3 * k];
        gg = plte[3 * k + 1];
        bb = plte[3 * k + 2];
        if ((size_t)k < trns.size()) aa = trns[(size_t)k];
      } else if (channels <= 2) {
        rr = gg = bb = sample(0);
        if (channels == 2) aa = sample(1);
      } else {
        rr = sample(0);
        gg = sample(1);
        bb = sample(2);
        if (channels == 4) aa = sample(3);
      }
      o[4 * x + 0] = (uint8_t)rr;
      o[4 * x + 1] = (uint8_t)gg;
      o[4 * x + 2] = (uint8_t)bb;
      o[4 * x + 3] = (uint8_t)aa;
    }
    std::swap(prev, cur);
  }
  return GI_OK;
}

int decode_ppm(const std::vector<unsigned char> &b, int32_t &w, int32_t &h, uint8_t *rgba, bool decode) {
  if (b.size() < 2 || b[0] != 'P' || b[1] != '6') return GI_ERR_INVALID_ARG;
  size_t pos = 2;
  long v[3];
  for (int k = 0; k < 3; ++k) {
    for (;;) {  // whitespace and comments
      while (pos < b.size() && std::isspace(b[pos])) ++pos;
      if (pos < b.size() && b[pos] == '#') {
        while (pos < b.size() && b[pos] != '\n') ++pos;
        continue;
      }
      break;
    }
    if (pos >= b.size() || !std::isdigit(b[pos])) return GI_ERR_INVALID_ARG;
    v[k] = 0;
    while (pos < b.size() && std::isdigit(b[pos]) && v[k] < (1L << 30)) v[k] = 10 * v[k] + (b[pos++] - '0');
  }
  if (pos >= b.size() || !std::isspace(b[pos])) return GI_ERR_INVALID_ARG;
  ++pos;  // one whitespace byte before the samples
  const long maxval = v[2];
  if (v[0] <= 0 || v[1] <= 0 || v[0] * v[1] > (1L << 28) || maxval <= 0 || maxval > 65535) return GI_ERR_INVALID_ARG;
  w = (int32_t)v[0];
  h = (int32_t)v[1];
  if (!decode) return GI_OK;
  const size_t sb = maxval < 256 ? 1 : 2, n = (size_t)w * h;
  if (b.size() - pos < 3 * sb * n) return GI_ERR_INVALID_ARG;
  for (size_t i = 0; i < n; ++i) {
    for (int c = 0; c < 3; ++c) {
      const unsigned char *q = &b[pos + (3 * i + c) * sb];
      const long s = sb == 1 ? q[0] : ((long)q[0] << 8 | q[1]);
      rgba[4 * i + c] = (uint8_t)(maxval == 255 ? s : (s * 255 + maxval / 2) / maxval);
    }
    rgba[4 * i + 3] = 255;
  }
  return GI_OK;
}

int decode_image(const char *path, int32_t &w, int32_t &h, uint8_t *rgba, bool decode) {
  if (!path) return GI_ERR_INVALID_ARG;
  std::vector<unsigned char> b;
  if (!read_file(path, b)) return GI_ERR_INVALID_ARG;
  if (b.size() >= 8 && b[0] == 0x89 && b[1] == 'P') return decode_png(b, w, h, rgba, decode);
  if (b.size() >= 2 && b[0] == 'P' && b[1] == '6') return decode_ppm(b, w, h, rgba, decode);
  return GI_ERR_INVALID_ARG;
}

}  // namespace

extern "C" {

int gi_image_size(const char *path, int32_t *w, int32_t *h) {
  if (!w || !h) return GI_ERR_INVALID_ARG;
  return decode_image(path, *w, *h, nullptr, false);
}

int gi_image_read(const char *path, uint8_t *rgba) {
  if (!rgba) return GI_ERR_INVALID_ARG;
  int32_t w = 0, h = 0;
  return decode_image(path, w, h, rgba, true);
}

}  // extern "C"
