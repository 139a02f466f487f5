// Film output: PNGFilm::Output semantics (src/cpu/films/PNGFilm.cpp:20-58) —
// each channel of the averaged film clamped at 255 and truncated to 8 bits,
// alpha 255, written as an 8-bit RGBA PNG — and a float PFM of the same film.
//
// The PNG is self-contained: its zlib stream uses stored (uncompressed)
// deflate blocks, so no compression library is needed; any PNG reader decodes
// it to the same pixels lodepng's compressed file holds.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "ptx_internal.h"

namespace ptx {
namespace {

uint32_t crc32_of(const uint8_t *p, size_t n, uint32_t crc = 0) {
   static uint32_t table[256];
   static bool init = false;
   if (!init) {
      for (uint32_t i = 0; i < 256; i++) {
         uint32_t c = i;
         for (int k = 0; k < 8; k++) c = c & 1 ? 0xEDB88320u ^ (c >> 1) : c >> 1;
         table[i] = c;
      }
      init = true;
   }
   crc = ~crc;
   for (size_t i = 0; i < n; i++) crc = table[(crc ^ p[i]) & 0xFF] ^ (crc >> 8);
   return ~crc;
}

void put_be32(std::vector<uint8_t> &v, uint32_t x) {
   v.push_back(uint8_t(x >> 24)); v.push_back(uint8_t(x >> 16)); v.push_back(uint8_t(x >> 8)); v.push_back(uint8_t(x));
}

void chunk(std::vector<uint8_t> &out, const char *type, const std::vector<uint8_t> &data) {
   put_be32(out, (uint32_t)data.size());
   const size_t start = out.size();
   out.insert(out.end(), type, type + 4);
   out.insert(out.end(), data.begin(), data.end());
   put_be32(out, crc32_of(out.data() + start, out.size() - start));
}

// PNGFilm: static_cast<unsigned char>(v > 255 ? 255 : v).  A NaN channel
// (undefined in the reference's conversion) becomes 0.
uint8_t to_u8(float v) {
   if (!(v == v)) return 0;
   const float c = v > 255 ? 255.f : v;
   return c <= 0.f ? 0 : (uint8_t)c;
}

}  // namespace
}  // namespace ptx

using namespace ptx;

extern "C" {

int ptx_film_to_rgba8(const float *film_rgb, uint32_t width, uint32_t height, uint8_t *rgba) {
   if (!film_rgb || !rgba) return set_error(PTX_ERR_INVALID, "NULL argument");
   const size_t n = (size_t)width * height;
   for (size_t i = 0; i < n; i++) {
      rgba[4 * i + 0] = to_u8(film_rgb[3 * i + 0]);
      rgba[4 * i + 1] = to_u8(film_rgb[3 * i + 1]);
      rgba[4 * i + 2] = to_u8(film_rgb[3 * i + 2]);
      rgba[4 * i + 3] = 255;
   }
   return PTX_OK;
}

int ptx_film_write_png(const float *film_rgb, uint32_t width, uint32_t height, const char *path) {
   if (!film_rgb || !path) return set_error(PTX_ERR_INVALID, "NULL argument");
   if (width == 0 || height == 0) return set_error(PTX_ERR_INVALID, "empty film");
   std::vector<uint8_t> rgba((size_t)width * height * 4);
   ptx_film_to_rgba8(film_rgb, width, height, rgba.data());
   // raw scanlines, filter type 0 (None)
   const size_t row = (size_t)width * 4 + 1;
   std::vector<uint8_t> raw(row * height);
   for (uint32_t y = 0; y < height; y++) {
      raw[y * row] = 0;
      std::memcpy(&raw[y * row + 1], &rgba[(size_t)y * width * 4], (size_t)width * 4);
   }
   // zlib stream of stored deflate blocks (at most 65535 bytes each) + Adler-32
   std::vector<uint8_t> z = {0x78, 0x01};
   uint32_t a = 1, b = 0;
   for (size_t off = 0; off < raw.size() || off == 0;) {
      const size_t len = std::min<size_t>(65535, raw.size() - off);
      const bool last = off + len >= raw.size();
      z.push_back(last ? 1 : 0);
      z.push_back(uint8_t(len)); z.push_back(uint8_t(len >> 8));
      z.push_back(uint8_t(~len)); z.push_back(uint8_t((~len) >> 8));
      z.insert(z.end(), raw.begin() + off, raw.begin() + off + len);
      off += len;
      if (last) break;
   }
   for (uint8_t c : raw) {
      a = (a + c) % 65521u;
      b = (b + a) % 65521u;
   }
   put_be32(z, (b << 16) | a);
   std::vector<uint8_t> png = {0x89, 'P', 'N', 'G', 0x0D, 0x0A, 0x1A, 0x0A};
   std::vector<uint8_t> ihdr;
   put_be32(ihdr, width);
   put_be32(ihdr, height);
   ihdr.insert(ihdr.end(), {8, 6, 0, 0, 0});   // 8-bit RGBA, deflate, adaptive filtering, no interlace
   chunk(png, "IHDR", ihdr);
   chunk(png, "IDAT", z);
   chunk(png, "IEND", {});
   FILE *f = std::fopen(path, "wb");
   if (!f) return set_error(PTX_ERR_INVALID, std::string("cannot write ") + path);
   const bool ok = std::fwrite(png.data(), 1, png.size(), f) == png.size();
   std::fclose(f);
   return ok ? PTX_OK : set_error(PTX_ERR_INVALID, std::string("short write to ") + path);
}

int ptx_film_write_pfm(const float *film_rgb, uint32_t width, uint32_t height, const char *path) {
   if (!film_rgb || !path) return set_error(PTX_ERR_INVALID, "NULL argument");
   FILE *f = std::fopen(path, "wb");
   if (!f) return set_error(PTX_ERR_INVALID, std::string("cannot write ") + path);
   std::fprintf(f, "PF\n%u %u\n-1.0\n", width, height);   // little-endian, rows bottom to top
   bool ok = true;
   for (uint32_t y = height; y-- > 0;)
      ok &= std::fwrite(film_rgb + (size_t)y * width * 3, sizeof(float), (size_t)width * 3, f) == (size_t)width * 3;
   std::fclose(f);
   return ok ? PTX_OK : set_error(PTX_ERR_INVALID, std::string("short write to ") + path);
}

}  // extern "C"
