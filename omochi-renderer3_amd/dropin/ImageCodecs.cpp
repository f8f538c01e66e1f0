// ImageCodecs.cpp -- PNG and baseline JPEG decoding for textures.
//
// Written from the format specifications: PNG (ISO/IEC 15948, W3C PNG 2nd
// edition: chunk layout, the five scanline filters, Adam7) on top of zlib's
// inflate, and JPEG (ITU-T T.81: Huffman entropy coding, zig-zag order,
// 8x8 inverse DCT; JFIF 1.02 YCbCr conversion).  Chroma is upsampled with the
// triangle ("fancy") filter described in the IJG libjpeg documentation.
#include "ImageCodecs.h"

#include <zlib.h>

#include <cmath>
#include <cstring>
#include <fstream>
#include <iterator>

namespace OmochiRenderer {
namespace codec {
namespace {

bool fail(std::string &err, const char *msg) {
  err = msg;
  return false;
}
uint32_t be32(const uint8_t *p) { return (uint32_t)p[0] << 24 | (uint32_t)p[1] << 16 | (uint32_t)p[2] << 8 | p[3]; }
uint32_t be16(const uint8_t *p) { return (uint32_t)p[0] << 8 | p[1]; }

// ------------------------------------------------------------------ PNG
struct Pass {
  uint32_t x0, y0, dx, dy;
};
const Pass kAdam7[7] = {{0, 0, 8, 8}, {4, 0, 8, 8}, {0, 4, 4, 8}, {2, 0, 4, 4},
                        {0, 2, 2, 4}, {1, 0, 2, 2}, {0, 1, 1, 2}};

uint8_t paeth(int a, int b, int c) {
  int p = a + b - c;
  int pa = std::abs(p - a), pb = std::abs(p - b), pc = std::abs(p - c);
  if (pa <= pb && pa <= pc) return (uint8_t)a;
  return (uint8_t)(pb <= pc ? b : c);
}

// Reverses the filter of each scanline in place; rows of `len` bytes, each
// preceded by its filter byte.
bool unfilter(uint8_t *raw, uint32_t rows, size_t len, uint32_t bpp, std::string &err) {
  const uint8_t *prev = nullptr;
  for (uint32_t y = 0; y < rows; y++) {
    uint8_t f = raw[0];
    uint8_t *x = raw + 1;
    for (size_t i = 0; i < len; i++) {
      int a = i >= bpp ? x[i - bpp] : 0;
      int b = prev ? prev[i] : 0;
      int c = (prev && i >= bpp) ? prev[i - bpp] : 0;
      switch (f) {
        case 0: break;
        case 1: x[i] = (uint8_t)(x[i] + a); break;
        case 2: x[i] = (uint8_t)(x[i] + b); break;
        case 3: x[i] = (uint8_t)(x[i] + ((a + b) >> 1)); break;
        case 4: x[i] = (uint8_t)(x[i] + paeth(a, b, c)); break;
        default: return fail(err, "invalid PNG filter type");
      }
    }
    prev = x;
    raw += 1 + len;
  }
  return true;
}

bool inflate_exact(const std::vector<uint8_t> &in, std::vector<uint8_t> &out, std::string &err) {
  z_stream zs;
  std::memset(&zs, 0, sizeof(zs));
  if (inflateInit(&zs) != Z_OK) return fail(err, "zlib initialisation failed");
  zs.next_in = const_cast<Bytef *>(in.data());
  zs.avail_in = (uInt)in.size();
  zs.next_out = out.data();
  zs.avail_out = (uInt)out.size();
  int rc = inflate(&zs, Z_FINISH);
  size_t got = zs.total_out;
  inflateEnd(&zs);
  // data beyond the image (rc = Z_BUF_ERROR with a full buffer) is ignored
  if (rc != Z_STREAM_END && !(rc == Z_BUF_ERROR && zs.avail_out == 0)) {
    if (rc == Z_DATA_ERROR) return fail(err, "corrupt PNG image data");
    if (got < out.size()) return fail(err, "PNG image data too short");
  }
  if (got < out.size()) return fail(err, "PNG image data too short");
  return true;
}

}  // namespace

bool decode_png(const uint8_t *d, size_t n, Pixels8 &out, std::string &err) {
  static const uint8_t sig[8] = {137, 80, 78, 71, 13, 10, 26, 10};
  if (n < 8 || std::memcmp(d, sig, 8) != 0) return fail(err, "not a PNG file");
  size_t pos = 8;
  uint32_t W = 0, H = 0;
  int depth = 0, ctype = -1, interlace = 0;
  bool ihdr = false;
  std::vector<uint8_t> pal, trns, idat;
  while (pos + 8 <= n) {
    uint32_t len = be32(d + pos);
    const uint8_t *type = d + pos + 4;
    pos += 8;
    if (len > n - pos || n - pos - len < 4) return fail(err, "truncated PNG chunk");
    const uint8_t *c = d + pos;
    if (!std::memcmp(type, "IHDR", 4)) {
      if (len != 13 || ihdr) return fail(err, "bad PNG header");
      W = be32(c);
      H = be32(c + 4);
      depth = c[8];
      ctype = c[9];
      if (c[10] != 0 || c[11] != 0) return fail(err, "unknown PNG compression or filter method");
      interlace = c[12];
      ihdr = true;
    } else if (!std::memcmp(type, "PLTE", 4)) {
      if (len == 0 || len % 3 != 0 || len > 768) return fail(err, "bad PNG palette");
      pal.assign(c, c + len);
    } else if (!std::memcmp(type, "tRNS", 4)) {
      trns.assign(c, c + len);
    } else if (!std::memcmp(type, "IDAT", 4)) {
      idat.insert(idat.end(), c, c + len);
    } else if (!std::memcmp(type, "IEND", 4)) {
      break;
    } else if (!(type[0] & 0x20)) {
      return fail(err, "unknown critical PNG chunk");
    }
    pos += len + 4;  // data + CRC
  }
  if (!ihdr) return fail(err, "PNG without IHDR");
  if (depth != 8) return fail(err, "PNG: only 8-bit samples are supported");
  uint32_t spp;
  switch (ctype) {
    case 0: spp = 1; break;
    case 2: spp = 3; break;
    case 3: spp = 1; break;
    case 4: spp = 2; break;
    case 6: spp = 4; break;
    default: return fail(err, "bad PNG colour type");
  }
  if (interlace > 1) return fail(err, "bad PNG interlace method");
  if (W == 0 || H == 0 || W > (1u << 24) || H > (1u << 24)) return fail(err, "bad PNG size");
  if ((uint64_t)W * H * 4 > (1ull << 31)) return fail(err, "PNG too large");
  if (ctype == 3 && pal.empty()) return fail(err, "PNG palette image without PLTE");
  if (idat.empty()) return fail(err, "PNG without image data");

  Pass whole = {0, 0, 1, 1};
  const Pass *passes = interlace ? kAdam7 : &whole;
  const int npass = interlace ? 7 : 1;
  size_t raw_len = 0;
  for (int p = 0; p < npass; p++) {
    const Pass &q = passes[p];
    uint64_t pw = W > q.x0 ? (W - q.x0 + q.dx - 1) / q.dx : 0;
    uint64_t ph = H > q.y0 ? (H - q.y0 + q.dy - 1) / q.dy : 0;
    if (pw && ph) raw_len += ph * (1 + pw * spp);
  }
  std::vector<uint8_t> raw(raw_len);
  if (!inflate_exact(idat, raw, err)) return false;

  std::vector<uint8_t> img((size_t)W * H * spp);
  uint8_t *r = raw.data();
  for (int p = 0; p < npass; p++) {
    const Pass &q = passes[p];
    uint32_t pw = W > q.x0 ? (W - q.x0 + q.dx - 1) / q.dx : 0;
    uint32_t ph = H > q.y0 ? (H - q.y0 + q.dy - 1) / q.dy : 0;
    if (!pw || !ph) continue;
    size_t len = (size_t)pw * spp;
    if (!unfilter(r, ph, len, spp, err)) return false;
    for (uint32_t y = 0; y < ph; y++) {
      const uint8_t *row = r + (size_t)y * (1 + len) + 1;
      for (uint32_t x = 0; x < pw; x++) {
        size_t o = ((size_t)(q.y0 + y * q.dy) * W + q.x0 + x * q.dx) * spp;
        std::memcpy(&img[o], row + (size_t)x * spp, spp);
      }
    }
    r += (size_t)ph * (1 + len);
  }

  out.width = W;
  out.height = H;
  if (ctype == 3) {
    out.channels = trns.empty() ? 3 : 4;
    out.px.resize((size_t)W * H * out.channels);
    const size_t npal = pal.size() / 3;
    for (size_t i = 0; i < (size_t)W * H; i++) {
      size_t k = img[i];
      uint8_t *o = &out.px[i * out.channels];
      if (k < npal) {
        o[0] = pal[3 * k];
        o[1] = pal[3 * k + 1];
        o[2] = pal[3 * k + 2];
      } else {
        o[0] = o[1] = o[2] = 0;
      }
      if (out.channels == 4) o[3] = k < trns.size() ? trns[k] : 255;
    }
  } else {
    out.channels = spp;
    out.px.swap(img);
  }
  return true;
}

// ------------------------------------------------------------------ JPEG
namespace {

const uint8_t kZigZag[64] = {0,  1,  8,  16, 9,  2,  3,  10, 17, 24, 32, 25, 18, 11, 4,  5,
                             12, 19, 26, 33, 40, 48, 41, 34, 27, 20, 13, 6,  7,  14, 21, 28,
                             35, 42, 49, 56, 57, 50, 43, 36, 29, 22, 15, 23, 30, 37, 44, 51,
                             58, 59, 52, 45, 38, 31, 39, 46, 53, 60, 61, 54, 47, 55, 62, 63};

// Canonical Huffman table (T.81 Annex C): codes of each length are
// consecutive; maxcode[l] is the largest code of length l (-1 if none).
struct Huff {
  bool present = false;
  int32_t mincode[17], maxcode[17], valptr[17];
  uint8_t vals[256];
  bool build(const uint8_t *counts, const uint8_t *symbols, int nsym) {
    int code = 0, k = 0;
    for (int l = 1; l <= 16; l++) {
      valptr[l] = k;
      mincode[l] = code;
      code += counts[l - 1];
      k += counts[l - 1];
      maxcode[l] = counts[l - 1] ? code - 1 : -1;
      if (code > (1 << l)) return false;
      code <<= 1;
    }
    std::memcpy(vals, symbols, (size_t)nsym);
    present = true;
    return true;
  }
};

// Entropy-coded segment reader: removes stuffed zero bytes, stops at a
// marker (then yields zero bits).
struct BitReader {
  const uint8_t *d;
  size_t n, pos;
  uint32_t acc = 0;
  int nbits = 0;
  bool at_marker = false;
  void fill() {
    while (nbits <= 24) {
      uint32_t b = 0;
      if (!at_marker && pos < n) {
        b = d[pos];
        if (b == 0xFF) {
          uint8_t nx = pos + 1 < n ? d[pos + 1] : 0xD9;
          if (nx == 0x00) {
            pos += 2;
          } else {
            at_marker = true;
            b = 0;
          }
        } else {
          pos++;
        }
      }
      acc |= b << (24 - nbits);
      nbits += 8;
    }
  }
  int bit() {
    if (nbits < 1) fill();
    int v = (int)(acc >> 31);
    acc <<= 1;
    nbits--;
    return v;
  }
  int bits(int s) {
    if (s == 0) return 0;
    if (nbits < s) fill();
    int v = (int)(acc >> (32 - s));
    acc <<= s;
    nbits -= s;
    return v;
  }
  void reset() {
    acc = 0;
    nbits = 0;
  }
};

int decode_sym(BitReader &br, const Huff &h) {
  int code = br.bit();
  for (int l = 1; l <= 16; l++) {
    if (h.maxcode[l] >= 0 && code <= h.maxcode[l] && code >= h.mincode[l]) return h.vals[h.valptr[l] + code - h.mincode[l]];
    code = (code << 1) | br.bit();
  }
  return -1;
}

// EXTEND (T.81 F.2.2.1): s-bit magnitude category to a signed value.
int extend(int v, int s) { return v < (1 << (s - 1)) ? v - (1 << s) + 1 : v; }

struct Comp {
  int id, h, v, tq;
  int td = 0, ta = 0;
  int bw = 0, bh = 0;  // blocks per row / column in the plane (MCU padded)
  int dc = 0;
  std::vector<uint8_t> plane;
};

struct IDCT {
  float c[8][8];  // c[x][u] = C(u)/2 cos((2x+1) u pi / 16)
  IDCT() {
    for (int x = 0; x < 8; x++)
      for (int u = 0; u < 8; u++)
        c[x][u] = (float)((u == 0 ? std::sqrt(0.5) : 1.0) * 0.5 * std::cos((2 * x + 1) * u * M_PI / 16.0));
  }
  void run(const int32_t *F, uint8_t *out, int stride) const {
    float t[8][8];
    for (int v = 0; v < 8; v++)
      for (int x = 0; x < 8; x++) {
        float s = 0.0f;
        for (int u = 0; u < 8; u++) s += c[x][u] * (float)F[v * 8 + u];
        t[v][x] = s;
      }
    for (int y = 0; y < 8; y++)
      for (int x = 0; x < 8; x++) {
        float s = 0.0f;
        for (int v = 0; v < 8; v++) s += c[y][v] * t[v][x];
        int p = (int)std::floor(s + 128.5f);
        out[y * stride + x] = (uint8_t)(p < 0 ? 0 : (p > 255 ? 255 : p));
      }
  }
};

uint8_t clamp8(float v) {
  int p = (int)std::floor(v + 0.5f);
  return (uint8_t)(p < 0 ? 0 : (p > 255 ? 255 : p));
}

// Triangle-filter upsampling of one plane by (fx, fy) in {1, 2}: each output
// sample weighs the nearer input sample 3/4 and the farther 1/4 along each
// doubled axis (edges repeat).  Other factors replicate.
void upsample(const Comp &c, int fx, int fy, int W, int H, std::vector<uint8_t> &dst) {
  const int pw = c.bw * 8;
  const int cw = (W + fx - 1) / fx, ch = (H + fy - 1) / fy;  // samples that cover the image
  dst.resize((size_t)W * H);
  auto at = [&](int x, int y) -> int {
    x = x < 0 ? 0 : (x >= cw ? cw - 1 : x);
    y = y < 0 ? 0 : (y >= ch ? ch - 1 : y);
    return c.plane[(size_t)y * pw + x];
  };
  if ((fx == 1 || fx == 2) && (fy == 1 || fy == 2)) {
    for (int y = 0; y < H; y++) {
      int sy = y / fy, ny = sy;
      if (fy == 2) ny = (y & 1) ? sy + 1 : sy - 1;
      for (int x = 0; x < W; x++) {
        int sx = x / fx, nx = sx;
        if (fx == 2) nx = (x & 1) ? sx + 1 : sx - 1;
        int v;
        if (fx == 2 && fy == 2) {
          int near = 3 * at(sx, sy) + at(sx, ny);
          int far = 3 * at(nx, sy) + at(nx, ny);
          v = (3 * near + far + 8) >> 4;
        } else if (fx == 2) {
          v = (3 * at(sx, sy) + at(nx, sy) + 2) >> 2;
        } else if (fy == 2) {
          v = (3 * at(sx, sy) + at(sx, ny) + 2) >> 2;
        } else {
          v = at(sx, sy);
        }
        dst[(size_t)y * W + x] = (uint8_t)v;
      }
    }
  } else {
    for (int y = 0; y < H; y++)
      for (int x = 0; x < W; x++) dst[(size_t)y * W + x] = (uint8_t)at(x / fx, y / fy);
  }
}

}  // namespace

bool decode_jpeg(const uint8_t *d, size_t n, Pixels8 &out, std::string &err) {
  if (n < 4 || d[0] != 0xFF || d[1] != 0xD8) return fail(err, "not a JPEG file");
  uint16_t qt[4][64];
  bool qt_ok[4] = {false, false, false, false};
  Huff hdc[4], hac[4];
  std::vector<Comp> comps;
  int W = 0, H = 0, hmax = 1, vmax = 1, mcux = 0, mcuy = 0;
  uint32_t restart = 0;
  bool frame = false, any_scan = false;
  static const IDCT idct;
  size_t pos = 2;
  while (true) {
    // next marker (fill bytes 0xFF allowed)
    while (pos < n && d[pos] != 0xFF) pos++;
    while (pos < n && d[pos] == 0xFF) pos++;
    if (pos >= n) break;
    uint8_t m = d[pos++];
    if (m == 0xD9) break;                       // EOI
    if (m >= 0xD0 && m <= 0xD7) continue;       // stray RSTn
    if (m == 0x01) continue;                    // TEM
    if (pos + 2 > n) return fail(err, "truncated JPEG marker");
    uint32_t L = be16(d + pos);
    if (L < 2 || pos + L > n) return fail(err, "truncated JPEG segment");
    const uint8_t *s = d + pos + 2;
    const uint32_t len = L - 2;
    if (m == 0xDB) {  // DQT
      uint32_t i = 0;
      while (i < len) {
        int pq = s[i] >> 4, tq = s[i] & 15;
        if (tq > 3 || pq > 1) return fail(err, "bad JPEG quantization table");
        i++;
        if (i + (pq ? 128u : 64u) > len) return fail(err, "truncated JPEG quantization table");
        for (int k = 0; k < 64; k++) qt[tq][k] = pq ? (uint16_t)be16(s + i + 2 * k) : s[i + k];
        i += pq ? 128 : 64;
        qt_ok[tq] = true;
      }
    } else if (m == 0xC4) {  // DHT
      uint32_t i = 0;
      while (i < len) {
        if (i + 17 > len) return fail(err, "truncated JPEG Huffman table");
        int tc = s[i] >> 4, th = s[i] & 15;
        if (tc > 1 || th > 3) return fail(err, "bad JPEG Huffman table");
        int total = 0;
        for (int k = 0; k < 16; k++) total += s[i + 1 + k];
        if (total > 256 || i + 17 + total > len) return fail(err, "bad JPEG Huffman table");
        Huff &h = tc ? hac[th] : hdc[th];
        if (!h.build(s + i + 1, s + i + 17, total)) return fail(err, "bad JPEG Huffman code lengths");
        i += 17 + total;
      }
    } else if (m == 0xDD) {  // DRI
      if (len < 2) return fail(err, "bad JPEG restart interval");
      restart = be16(s);
    } else if (m == 0xC0 || m == 0xC1) {  // SOF0 / SOF1
      if (frame) return fail(err, "JPEG with two frames");
      if (len < 6 || s[0] != 8) return fail(err, "JPEG: only 8-bit samples are supported");
      H = (int)be16(s + 1);
      W = (int)be16(s + 3);
      int nc = s[5];
      if (W == 0 || H == 0) return fail(err, "JPEG without a frame size");
      if (nc != 1 && nc != 3) return fail(err, "JPEG: 1 or 3 components are supported");
      if (len != 6u + 3u * nc) return fail(err, "bad JPEG frame header");
      if ((uint64_t)W * H * 3 > (1ull << 31)) return fail(err, "JPEG too large");
      comps.resize(nc);
      for (int k = 0; k < nc; k++) {
        comps[k].id = s[6 + 3 * k];
        comps[k].h = s[7 + 3 * k] >> 4;
        comps[k].v = s[7 + 3 * k] & 15;
        comps[k].tq = s[8 + 3 * k];
        if (comps[k].h < 1 || comps[k].h > 4 || comps[k].v < 1 || comps[k].v > 4 || comps[k].tq > 3)
          return fail(err, "bad JPEG component");
        hmax = std::max(hmax, comps[k].h);
        vmax = std::max(vmax, comps[k].v);
      }
      mcux = (W + 8 * hmax - 1) / (8 * hmax);
      mcuy = (H + 8 * vmax - 1) / (8 * vmax);
      for (Comp &c : comps) {
        c.bw = mcux * c.h;
        c.bh = mcuy * c.v;
        c.plane.assign((size_t)c.bw * 8 * c.bh * 8, 0);
      }
      frame = true;
    } else if (m == 0xC2 || m == 0xC3 || (m >= 0xC5 && m <= 0xC7) || (m >= 0xC9 && m <= 0xCB) ||
               (m >= 0xCD && m <= 0xCF)) {
      return fail(err, "JPEG: progressive, lossless and arithmetic-coded files are not supported");
    } else if (m == 0xDA) {  // SOS
      if (!frame) return fail(err, "JPEG scan before the frame header");
      int ns = len ? s[0] : 0;
      if (ns < 1 || ns > 4 || len != 4u + 2u * ns) return fail(err, "bad JPEG scan header");
      std::vector<Comp *> sc;
      for (int k = 0; k < ns; k++) {
        int id = s[1 + 2 * k];
        Comp *c = nullptr;
        for (Comp &q : comps)
          if (q.id == id) c = &q;
        if (!c) return fail(err, "JPEG scan names an unknown component");
        c->td = s[2 + 2 * k] >> 4;
        c->ta = s[2 + 2 * k] & 15;
        if (c->td > 3 || c->ta > 3 || !hdc[c->td].present || !hac[c->ta].present)
          return fail(err, "JPEG scan uses an undefined Huffman table");
        if (!qt_ok[c->tq]) return fail(err, "JPEG component uses an undefined quantization table");
        c->dc = 0;
        sc.push_back(c);
      }
      BitReader br{d, n, pos + L};
      int32_t F[64];
      auto block = [&](Comp &c, int bx, int by) -> bool {
        std::memset(F, 0, sizeof(F));
        int t = decode_sym(br, hdc[c.td]);
        if (t < 0 || t > 11) return false;
        c.dc += t ? extend(br.bits(t), t) : 0;
        F[0] = c.dc * (int32_t)qt[c.tq][0];
        for (int k = 1; k < 64;) {
          int rs = decode_sym(br, hac[c.ta]);
          if (rs < 0) return false;
          int r = rs >> 4, sz = rs & 15;
          if (sz == 0) {
            if (r != 15) break;  // end of block
            k += 16;
            continue;
          }
          k += r;
          if (k > 63) return false;
          F[kZigZag[k]] = extend(br.bits(sz), sz) * (int32_t)qt[c.tq][k];
          k++;
        }
        idct.run(F, &c.plane[((size_t)by * 8) * (c.bw * 8) + (size_t)bx * 8], c.bw * 8);
        return true;
      };
      // a single-component scan codes the component's own block grid, an
      // interleaved scan whole MCUs (T.81 A.2)
      const bool single = ns == 1;
      const int cols = single ? (((W * sc[0]->h + hmax - 1) / hmax) + 7) / 8 : mcux;
      const int rows = single ? (((H * sc[0]->v + vmax - 1) / vmax) + 7) / 8 : mcuy;
      uint32_t todo = restart;
      for (int my = 0; my < rows; my++)
        for (int mx = 0; mx < cols; mx++) {
          if (restart && todo == 0) {
            // RSTn: byte-align, skip the marker, reset the predictors
            br.reset();
            size_t p = br.pos;
            while (p + 1 < n && !(d[p] == 0xFF && d[p + 1] >= 0xD0 && d[p + 1] <= 0xD7)) p++;
            br.pos = p + 1 < n ? p + 2 : n;
            br.at_marker = false;
            for (Comp *c : sc) c->dc = 0;
            todo = restart;
          }
          if (single) {
            if (!block(*sc[0], mx, my)) return fail(err, "corrupt JPEG data");
          } else {
            for (Comp *c : sc)
              for (int v = 0; v < c->v; v++)
                for (int h = 0; h < c->h; h++)
                  if (!block(*c, mx * c->h + h, my * c->v + v)) return fail(err, "corrupt JPEG data");
          }
          if (restart) todo--;
        }
      any_scan = true;
      // continue after the entropy-coded data (at the next marker)
      size_t p = br.pos;
      while (p + 1 < n && !(d[p] == 0xFF && d[p + 1] != 0x00 && !(d[p + 1] >= 0xD0 && d[p + 1] <= 0xD7))) p++;
      pos = p;
      continue;
    }
    pos += L;
  }
  if (!frame || !any_scan) return fail(err, "JPEG without image data");

  out.width = (uint32_t)W;
  out.height = (uint32_t)H;
  out.channels = comps.size() == 3 ? 3 : 1;
  std::vector<std::vector<uint8_t>> full(comps.size());
  for (size_t k = 0; k < comps.size(); k++) {
    if (hmax % comps[k].h || vmax % comps[k].v) return fail(err, "JPEG sampling factors do not divide");
    upsample(comps[k], hmax / comps[k].h, vmax / comps[k].v, W, H, full[k]);
  }
  out.px.resize((size_t)W * H * out.channels);
  if (out.channels == 1) {
    out.px = full[0];
  } else {
    for (size_t i = 0; i < (size_t)W * H; i++) {
      float y = full[0][i], cb = (float)full[1][i] - 128.0f, cr = (float)full[2][i] - 128.0f;
      out.px[3 * i] = clamp8(y + 1.402f * cr);
      out.px[3 * i + 1] = clamp8(y - 0.344136f * cb - 0.714136f * cr);
      out.px[3 * i + 2] = clamp8(y + 1.772f * cb);
    }
  }
  return true;
}

bool decode_file(const std::string &path, Pixels8 &out, std::string &err) {
  std::ifstream f(path.c_str(), std::ios::binary);
  if (!f) return fail(err, "cannot open image file");
  std::vector<uint8_t> data((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
  if (data.size() >= 8 && data[0] == 137 && data[1] == 'P' && data[2] == 'N' && data[3] == 'G')
    return decode_png(data.data(), data.size(), out, err);
  if (data.size() >= 2 && data[0] == 0xFF && data[1] == 0xD8) return decode_jpeg(data.data(), data.size(), out, err);
  return fail(err, "unsupported image format (PNG and baseline JPEG are read)");
}

}  // namespace codec
}  // namespace OmochiRenderer
