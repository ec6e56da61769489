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

#include <cmath>
#include <cstdint>
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
