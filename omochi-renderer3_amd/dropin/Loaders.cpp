// Loaders.cpp -- the reference's file formats on the host: Radiance RGBE
// images, Wavefront OBJ/MTL meshes and the reference's .scene text format.
//
// Parsing follows the reference's observable behaviour (same accepted
// syntax, same materials and geometry, same results for its own assets);
// reference citations are `path:line` in the reference checkout.  Where the
// reference misbehaves (reading < 8-pixel-wide RGBE images, a .mtl next to an
// .obj given without a directory, duplicated polygons for materials of equal
// value, an empty material list), this code does the evident intended thing
// or reports an error instead.
#include <cassert>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <sstream>
#include <unordered_map>

#include "ImageCodecs.h"
#include "OmochiRenderer.h"

namespace OmochiRenderer {
namespace {

constexpr double kPi = 3.14159265358979323846264338327950288419716939;  // src/tools/Constant.h:5

// Utils::split / ltrim / rtrim / trim (src/tools/Utils.h:13-55): empty fields
// between delimiters are kept, no trailing empty field after a final
// delimiter.
std::vector<std::string> split(const std::string &s, char delim) {
  std::vector<std::string> out;
  size_t next = 0;
  while (next < s.size()) {
    size_t pos = s.find(delim, next);
    if (pos == std::string::npos) {
      out.push_back(s.substr(next));
      break;
    }
    out.push_back(s.substr(next, pos - next));
    next = pos + 1;
  }
  return out;
}
std::string ltrim(const std::string &s, const char *set = " \t") {
  size_t i = s.find_first_not_of(set);
  return i == std::string::npos ? std::string() : s.substr(i);
}
std::string rtrim(const std::string &s, const char *set = " \t") {
  size_t i = s.find_last_not_of(set);
  return i == std::string::npos ? std::string() : s.substr(0, i + 1);
}
std::string trim(const std::string &s) { return ltrim(rtrim(s)); }
bool starts(const std::string &s, const char *p) { return s.compare(0, std::strlen(p), p) == 0; }

}  // namespace

// ------------------------------------------------------------ Matrix
Matrix::Matrix() { std::memset(m, 0, sizeof(m)); }
Matrix::Matrix(const double m_[4][4]) { std::memcpy(m, m_, sizeof(m)); }
const Matrix &Matrix::Identity() {
  static const double id[4][4] = {{1, 0, 0, 0}, {0, 1, 0, 0}, {0, 0, 1, 0}, {0, 0, 0, 1}};
  static const Matrix i(id);
  return i;
}
// Rodrigues' rotation matrix, terms in the order of Matrix.h:18-32 so the
// binary64 entries agree with the reference's.
Matrix Matrix::RotateAroundVector(const Vector3 &axis, double a) {
  Matrix r;
  const double c = std::cos(a), s = std::sin(a), k = 1 - c;
  Vector3 v(axis);
  v.normalize();
  r.m[0][0] = v.x * v.x * k + c;
  r.m[0][1] = v.x * v.y * k - v.z * s;
  r.m[0][2] = v.z * v.x * k + v.y * s;
  r.m[1][0] = v.x * v.y * k + v.z * s;
  r.m[1][1] = v.y * v.y * k + c;
  r.m[1][2] = v.z * v.y * k - v.x * s;
  r.m[2][0] = v.x * v.z * k - v.y * s;
  r.m[2][1] = v.z * v.y * k + v.x * s;
  r.m[2][2] = v.z * v.z * k + c;
  r.m[3][3] = 1.0;
  return r;
}
Vector3 Matrix::Apply(const Vector3 &v) const {
  return Vector3(m[0][0] * v.x + m[0][1] * v.y + m[0][2] * v.z, m[1][0] * v.x + m[1][1] * v.y + m[1][2] * v.z,
                 m[2][0] * v.x + m[2][1] * v.y + m[2][2] * v.z);
}
Matrix Matrix::operator*(const Matrix &b) const {
  Matrix r;
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 4; j++) {
      double acc = 0;
      for (int k = 0; k < 4; k++) acc += m[i][k] * b.m[k][j];
      r.m[i][j] = acc;
    }
  return r;
}

// ------------------------------------------------------------ Image
const Color &Image::GetPixelByUV(double u, double v) const {
  auto clamp01 = [](double x) { return x < 0.0 ? 0.0 : (x > 1.0 ? 1.0 : x); };  // Utils::Clamp
  u = clamp01(u) * 0.99999;
  v = clamp01(v) * 0.99999;
  size_t x = (size_t)(int)(u * (double)m_width), y = (size_t)(int)(v * (double)m_height);
  return m_image[x + y * m_width];
}

// ------------------------------------------------------------ ImageHandler
ImageHandler &ImageHandler::GetInstance() {
  static ImageHandler h;
  return h;
}

// ImageHandler::LoadFromFile (src/tools/ImageHandler.cpp:20-90).  A name
// ending in ".hdr" is read as a Radiance file every time (linear, no gamma);
// any other name that is already loaded returns its id, else the file is
// decoded as PNG or baseline JPEG and each 8-bit sample b becomes b / 255
// (binary32 division, as the reference) and, with doReverseGamma,
// pow(b / 255, 2.2) (Utils::InvGammaRev, src/tools/Utils.h:103-105).
ImageHandler::IMAGE_ID ImageHandler::LoadFromFile(const std::string &fname, bool doReverseGamma) {
  std::unique_ptr<Image> image;
  const bool hdr = fname.size() >= 4 && fname.compare(fname.size() - 4, 4, ".hdr") == 0;
  if (hdr) {
    std::unique_ptr<HDRImage> h(new HDRImage);
    if (!h->ReadFromRadianceFile(fname)) {
      m_error = fname + ": not a readable Radiance file";
      return INVALID_IMAGE_ID;
    }
    image = std::move(h);
  } else {
    for (size_t i = m_files.size(); i-- > 0;)
      if (m_files[i] == fname && m_images[i]) return (IMAGE_ID)i;
    codec::Pixels8 px;
    std::string err;
    if (!codec::decode_file(fname, px, err)) {
      m_error = fname + ": " + err;
      return INVALID_IMAGE_ID;
    }
    // the reference requires 3 or 4 samples per pixel (ImageHandler.cpp:59)
    if (px.channels < 3) {
      m_error = fname + ": a texture needs 3 or 4 channels";
      return INVALID_IMAGE_ID;
    }
    image.reset(new Image);
    image->m_width = px.width;
    image->m_height = px.height;
    image->m_image.resize((size_t)px.width * px.height);
    auto level = [doReverseGamma](uint8_t b) {
      double x = b / 255.0f;
      if (doReverseGamma) x = std::pow(x < 0.0 ? 0.0 : (x > 1.0 ? 1.0 : x), 2.2);
      return x;
    };
    for (size_t i = 0; i < image->m_image.size(); i++) {
      const uint8_t *p = &px.px[i * px.channels];
      image->m_image[i] = Color(level(p[0]), level(p[1]), level(p[2]));
    }
  }
  m_images.push_back(std::move(image));
  m_files.push_back(fname);
  return (IMAGE_ID)(m_images.size() - 1);
}

ImageHandler::IMAGE_ID ImageHandler::CreateImage(size_t width, size_t height) {
  std::unique_ptr<Image> img(new Image);
  img->m_width = width;
  img->m_height = height;
  img->m_image.resize(width * height);
  m_images.push_back(std::move(img));
  m_files.emplace_back();
  return (IMAGE_ID)(m_images.size() - 1);
}

Image *ImageHandler::GetImage(IMAGE_ID id) {
  return id >= 0 && (size_t)id < m_images.size() ? m_images[(size_t)id].get() : nullptr;
}
const Image *ImageHandler::GetImage(IMAGE_ID id) const {
  return id >= 0 && (size_t)id < m_images.size() ? m_images[(size_t)id].get() : nullptr;
}
void ImageHandler::ReleaseImage(IMAGE_ID id) {
  if (id >= 0 && (size_t)id < m_images.size()) m_images[(size_t)id].reset();
}
const std::string &ImageHandler::GetFileName(IMAGE_ID id) const {
  static const std::string none;
  return id >= 0 && (size_t)id < m_files.size() ? m_files[(size_t)id] : none;
}

// ------------------------------------------------------------ HDRImage
// RGBE shared-exponent pixels (Greg Ward's Radiance format).
void HDRImage::Color2RGBE(unsigned char rgbe[4], const Color &c) {
  float v = (float)std::max(std::max(c.x, c.y), c.z);
  if (v < 1e-32) {
    rgbe[0] = rgbe[1] = rgbe[2] = rgbe[3] = 0;
    return;
  }
  int e;
  v = std::frexp(v, &e) * 256.0f / v;
  rgbe[0] = (unsigned char)(c.x * v);
  rgbe[1] = (unsigned char)(c.y * v);
  rgbe[2] = (unsigned char)(c.z * v);
  rgbe[3] = (unsigned char)(e + 128);
}
void HDRImage::RGBE2Color(const unsigned char rgbe[4], Color &c) {
  if (rgbe[3] == 0) {
    c = Color();
    return;
  }
  float f = (float)std::ldexp(1.0, rgbe[3] - (128 + 8));
  c.x = rgbe[0] * f;
  c.y = rgbe[1] * f;
  c.z = rgbe[2] * f;
}

void HDRImage::SetImage(size_t width, size_t height, std::vector<Color> pixels) {
  m_width = width;
  m_height = height;
  m_image = std::move(pixels);
  m_image.resize(width * height);
}

bool HDRImage::ReadFromRadianceFile(const std::string &file) {
  std::ifstream in(file.c_str(), std::ios::binary);
  if (!in || in.bad()) return false;
  // header (HDRImage.cpp:51-95): "#?program", variable lines up to a blank
  // line, then "-Y height +X width"
  Header h;
  h.gamma = h.exposure = 1.0f;
  std::string line;
  std::getline(in, line);
  if (line.compare(0, 2, "#?") != 0) return false;
  h.valid |= 1;
  h.programtype = line.substr(2, line.find(' ') - 2);
  std::getline(in, line);
  while (!in.eof()) {
    if (line.empty() || line[0] == '\n') break;
    if (starts(line, "GAMMA=")) {
      std::istringstream iss(line);
      std::string key;
      iss >> key >> h.gamma;
      h.valid |= 2;
    } else if (starts(line, "EXPOSURE=")) {
      std::istringstream iss(line);
      std::string key;
      iss >> key >> h.exposure;
      h.valid |= 4;
    }
    std::getline(in, line);
  }
  std::getline(in, line);
  if (in.eof()) return false;
  {
    std::istringstream iss(line);
    std::string sy, sx;
    iss >> sy >> h.height >> sx >> h.width;
  }
  if (h.width <= 0 || h.height <= 0) return false;
  const size_t W = (size_t)h.width, H = (size_t)h.height;
  std::vector<Color> img(W * H);
  // flat pixels from index `from` on
  auto read_flat = [&](size_t from) -> bool {
    unsigned char p[4];
    for (size_t i = from; i < W * H; i++) {
      if (in.eof()) return false;
      in.read(reinterpret_cast<char *>(p), 4);
      RGBE2Color(p, img[i]);
    }
    return true;
  };
  bool ok = true;
  if (W < 8 || W > 0x7fff) {
    ok = read_flat(0);  // run-length encoding is not allowed at these widths
  } else {
    std::vector<unsigned char> scan(4 * W);
    size_t idx = 0;
    for (size_t y = 0; y < H && ok; y++) {
      unsigned char p[4];
      in.read(reinterpret_cast<char *>(p), 4);
      if (in.eof()) return false;
      if (p[0] != 2 || p[1] != 2 || (p[2] & 0x80)) {  // not run-length encoded: flat from here
        RGBE2Color(p, img[idx]);
        ok = read_flat(idx + 1);
        break;
      }
      if ((((size_t)p[2] << 8) | p[3]) != W) return false;
      // four channel planes, each a sequence of runs (count > 128: count-128
      // copies of one byte) and literal stretches (count bytes)
      unsigned char *b = scan.data();
      for (int ch = 0; ch < 4; ch++) {
        unsigned char *end = scan.data() + (size_t)(ch + 1) * W;
        while (b < end) {
          unsigned char rc[2];
          in.read(reinterpret_cast<char *>(rc), 2);
          if (in.eof()) return false;
          int count = rc[0] > 128 ? rc[0] - 128 : rc[0];
          if (count == 0 || count > end - b) return false;
          if (rc[0] > 128) {
            std::memset(b, rc[1], (size_t)count);
            b += count;
          } else {
            *b++ = rc[1];
            if (--count > 0) {
              in.read(reinterpret_cast<char *>(b), count);
              b += count;
            }
          }
        }
      }
      for (size_t x = 0; x < W; x++) {
        unsigned char q[4] = {scan[x], scan[x + W], scan[x + 2 * W], scan[x + 3 * W]};
        RGBE2Color(q, img[idx++]);
      }
    }
  }
  if (!ok) return false;
  m_info = h;
  m_width = W;
  m_height = H;
  m_image = std::move(img);
  return true;
}

namespace {
// One channel plane of a scanline, run-length encoded as HDRImage::WriteBytes_RLE
// (HDRImage.cpp:224-272) does: runs of >= 4 equal bytes (at most 127) as
// (128 + n, byte), everything before a run as literal stretches of <= 128.
void write_rle(std::ostream &out, const unsigned char *d, int n) {
  const int kMinRun = 4;
  int cur = 0;
  while (cur < n) {
    int beg = cur, run = 0;
    while (run < kMinRun && beg < n) {
      beg += run;
      run = 1;
      while (beg + run < n && d[beg] == d[beg + run] && run < 127) run++;
    }
    while (cur < beg && cur < n) {
      int lit = std::min(beg - cur, 128);
      unsigned char c = (unsigned char)lit;
      out.write(reinterpret_cast<const char *>(&c), 1);
      out.write(reinterpret_cast<const char *>(d + cur), lit);
      cur += lit;
    }
    if (run >= kMinRun && beg < n) {
      unsigned char rb[2] = {(unsigned char)(128 + run), d[beg]};
      out.write(reinterpret_cast<const char *>(rb), 2);
      cur += run;
    }
  }
}
}  // namespace

bool HDRImage::WriteToRadianceFile(const std::string &file) {
  std::ofstream out(file.c_str(), std::ios::binary);
  if (!out || out.bad()) return false;
  m_info.width = (int)m_width;
  m_info.height = (int)m_height;
  m_info.programtype = "RADIANCE";
  m_info.valid |= 1;
  out << "#?" << m_info.programtype << "\n";
  if (m_info.valid & 2) out << "GAMMA=" << m_info.gamma << "\n";
  if (m_info.valid & 4) out << "EXPOSURE=" << m_info.exposure << "\n";
  out << "FORMAT=32-bit_rle_rgbe\n\n";
  out << "-Y " << m_info.height << " +X " << m_info.width << "\n";
  const int W = (int)m_width, H = (int)m_height;
  const Color *px = m_image.data();
  if (W < 8 || W > 0x7fff) {  // flat
    for (size_t i = 0; i < m_width * m_height; i++) {
      unsigned char p[4];
      Color2RGBE(p, px[i]);
      out.write(reinterpret_cast<const char *>(p), 4);
    }
    return (bool)out;
  }
  std::vector<unsigned char> plane(4 * (size_t)W);
  for (int y = 0; y < H; y++) {
    unsigned char hdr[4] = {2, 2, (unsigned char)(W >> 8), (unsigned char)(W & 0xFF)};
    out.write(reinterpret_cast<const char *>(hdr), 4);
    for (int x = 0; x < W; x++) {
      unsigned char p[4];
      Color2RGBE(p, *px++);
      for (int c = 0; c < 4; c++) plane[(size_t)c * W + x] = p[c];
    }
    for (int c = 0; c < 4; c++) write_rle(out, plane.data() + (size_t)c * W, W);
  }
  return (bool)out;
}

// ------------------------------------------------------------ IBL
IBL::IBL(const std::string &hdr_filename) : m_w(0), m_h(0) {
  HDRImage img;
  if (!img.ReadFromRadianceFile(hdr_filename)) {
    std::cerr << "failed to load HDR file: " << hdr_filename << std::endl;
    return;
  }
  m_w = (uint32_t)img.GetWidth();
  m_h = (uint32_t)img.GetHeight();
  m_rgb.resize((size_t)m_w * m_h * 3);
  for (uint32_t y = 0; y < m_h; y++)
    for (uint32_t x = 0; x < m_w; x++) {
      const Color &c = img.GetPixel(x, y);
      float *o = &m_rgb[((size_t)y * m_w + x) * 3];
      o[0] = (float)c.x;
      o[1] = (float)c.y;
      o[2] = (float)c.z;
    }
}

// ------------------------------------------------------------ Model
namespace {

// MTL file (Model::LoadMaterialFile, Model.cpp:184-304).  The reference keeps
// the reflection type and the Kd/Ks colours of one material as the starting
// point of the next, sets no emission, and attaches a map_Kd texture only to
// the last material of the file; all kept here.
bool load_mtl(const std::string &filename, std::vector<std::pair<std::string, Material>> &out) {
  std::ifstream in(filename.c_str());
  std::string fname(filename);
  for (char &ch : fname)
    if (ch == '\\') ch = '/';
  std::string basedir;
  size_t p = fname.find_last_of('/');
  if (p != std::string::npos && p > 0) basedir = fname.substr(0, p);
  if (!in) return false;
  std::string name;
  Material cur;
  Color diffuse, specular;
  ImageHandler::IMAGE_ID tex = ImageHandler::INVALID_IMAGE_ID;
  auto reg = [&](bool with_texture) {
    cur.color = cur.reflection_type == Material::REFLECTION_TYPE_SPECULAR ? specular : diffuse;
    if (with_texture) cur.texture_id = tex;
    for (auto &e : out)
      if (e.first == name) {
        e.second = cur;
        return;
      }
    out.emplace_back(name, cur);
  };
  auto triple = [](const std::string &s, Color &c) {
    std::vector<std::string> v = split(s, ' ');
    if (v.size() < 3) return false;
    c = Color(std::atof(v[0].c_str()), std::atof(v[1].c_str()), std::atof(v[2].c_str()));
    return true;
  };
  std::string line;
  while (!in.eof()) {
    std::getline(in, line);
    if (starts(line, "#")) {
    } else if (starts(line, "newmtl ")) {
      if (!name.empty()) reg(false);
      name = line.substr(7);
    } else if (starts(line, "illum ")) {
      std::vector<std::string> v = split(line.substr(6), ' ');
      int mode = v.empty() ? 0 : std::atoi(v[0].c_str());
      if (mode >= 0 && mode <= 2) {
        cur.reflection_type = Material::REFLECTION_TYPE_LAMBERT;
      } else if (mode == 3) {
        if (v.size() != 2) {
          std::cerr << "illum 3 needs a refractive index: " << line << std::endl;
          return false;
        }
        cur.reflection_type = Material::REFLECTION_TYPE_REFRACTION;
        cur.refraction_rate = std::atof(v[1].c_str());
      }
    } else if (starts(line, "Kd ")) {
      if (!triple(line.substr(3), diffuse)) return false;
    } else if (starts(line, "Ks ")) {
      if (!triple(line.substr(3), specular)) return false;
    } else if (starts(line, "map_Kd ")) {
      std::string file = line.substr(7);
      tex = ImageHandler::GetInstance().LoadFromFile(basedir.empty() ? file : basedir + "/" + file);
    }
  }
  reg(true);
  return true;
}

}  // namespace

Model::~Model() { Clear(); }

void Model::Clear() {
  for (auto &l : m_meshes)
    for (Polygon *p : l) delete p;
  m_meshes.clear();
  m_list_keys.clear();
  m_materials.clear();
}

const Model::PolygonList &Model::GetPolygonList(const Material &mat) const {
  static const PolygonList none;
  for (size_t i = 0; i < m_list_keys.size(); i++)
    if (m_list_keys[i] == mat) return m_meshes[i];
  return none;
}

void Model::Transform(const Vector3 &pos, const Vector3 &scale, const Matrix &rot) {
  m_position = pos;
  for (auto &l : m_meshes)
    for (Polygon *p : l) p->Transform(pos, scale, rot);
}

// Name of the default material in the reference's name table (Model.cpp:64).
const char *const kDefaultMaterialName = "__default__material__";

// Model::ReadFromObj (Model.cpp:50-182) with Load3verticesFace (:321-367).
bool Model::ReadFromObj(const std::string &filename, bool flipV_of_UV) {
  std::ifstream in(filename.c_str());
  std::string baseDir;
  size_t p = filename.find_last_of('/');
  baseDir = p != std::string::npos ? filename.substr(0, p) : std::string(".");
  if (!in) return false;
  Clear();
  const Material deflt(Material::REFLECTION_TYPE_LAMBERT, Vector3(0, 0, 0), Vector3(0.99, 0.99, 0.99));
  std::vector<std::pair<std::string, Material>> named;  // material library
  Material current = deflt;
  auto list_of = [&](const Material &m) -> PolygonList & {
    for (size_t i = 0; i < m_list_keys.size(); i++)
      if (m_list_keys[i] == m) return m_meshes[i];
    m_list_keys.push_back(m);
    m_meshes.emplace_back();
    return m_meshes.back();
  };
  list_of(deflt);
  std::vector<Vector3> verts, norms, uvs;
  auto bad = [&](const std::string &why, const std::string &l) {
    std::cerr << "not correct .obj file!!: " << why << ": " << l << std::endl;
    Clear();
    return false;
  };
  // one triangle from three "v[/vt[/vn]]" references
  auto face3 = [&](const std::string &a, const std::string &b, const std::string &c, std::string &why) -> bool {
    const std::string *f[3] = {&a, &b, &c};
    Vector3 v[3], n[3], t[3];
    bool has_n = false;
    for (int i = 0; i < 3; i++) {
      std::vector<std::string> d = split(*f[i], '/');
      long vi = -1, ti = -1, ni = -1;
      if (d.size() >= 1 && d.size() <= 3) vi = std::atoi(d[0].c_str()) - 1;
      if (d.size() >= 2 && !d[1].empty()) ti = std::atoi(d[1].c_str()) - 1;
      if (d.size() == 3) ni = std::atoi(d[2].c_str()) - 1;
      if (vi < 0 || (size_t)vi >= verts.size()) { why = "vertex index"; return false; }
      if (ti >= (long)uvs.size() || ni >= (long)norms.size() || (d.size() >= 2 && !d[1].empty() && ti < 0) ||
          (d.size() == 3 && ni < 0)) { why = "uv or normal index"; return false; }
      v[i] = verts[(size_t)vi];
      if (ni >= 0) { n[i] = norms[(size_t)ni]; has_n = true; }
      if (ti >= 0) t[i] = uvs[(size_t)ti];
    }
    if (!has_n) n[0] = n[1] = n[2] = Polygon::CalculateNormal(v[0], v[1], v[2]);
    list_of(current).push_back(new Polygon(v[0], v[1], v[2], t[0], t[1], t[2], n[0], n[1], n[2], current,
                                           Vector3(0, 0, 0)));
    return true;
  };
  std::string line;
  while (!in.eof()) {
    std::getline(in, line);
    if (starts(line, "#")) {
    } else if (starts(line, "mtllib ")) {
      std::string lib = baseDir + "/" + line.substr(7);
      if (!load_mtl(lib, named)) {
        std::cerr << "failed to load material file: " << line.substr(7) << std::endl;
        Clear();
        return false;
      }
    } else if (starts(line, "g ")) {
      current = deflt;  // a group starts with the default material (Model.cpp:97-104)
    } else if (starts(line, "usemtl ")) {
      std::string name = line.substr(7);
      current = deflt;
      for (auto &e : named)
        if (e.first == name) current = e.second;
      list_of(current);
    } else if (starts(line, "v ")) {
      std::vector<std::string> t = split(line.substr(2), ' ');
      if (t.size() != 3) return bad("vertex", line);
      verts.emplace_back(std::atof(t[0].c_str()), std::atof(t[1].c_str()), std::atof(t[2].c_str()));
    } else if (starts(line, "vt ")) {
      std::vector<std::string> t = split(line.substr(3), ' ');
      if (t.size() < 2) return bad("texture coordinate", line);
      Vector3 vt(std::atof(t[0].c_str()), std::atof(t[1].c_str()), 0.0);
      if (flipV_of_UV) vt.y = 1 - vt.y;
      uvs.push_back(vt);
    } else if (starts(line, "vn ")) {
      std::vector<std::string> t = split(line.substr(3), ' ');
      if (t.size() != 3) return bad("normal", line);
      norms.emplace_back(std::atof(t[0].c_str()), std::atof(t[1].c_str()), std::atof(t[2].c_str()));
    } else if (starts(line, "f ")) {
      std::vector<std::string> f = split(line.substr(2), ' ');
      std::string why;
      // a face of n >= 3 vertices as the fan (0, i, i+1) (Model.cpp:148-171)
      for (size_t i = 1; f.size() >= 3 && i + 1 < f.size(); i++)
        if (!face3(f[0], f[i], f[i + 1], why)) return bad(why, line);
    }
  }
  // The material list is the reference's name table walked in its own
  // order: an std::unordered_map<std::string, Material> filled with the
  // default name first and then the library names in file order
  // (Model.cpp:64-67, :231, :301, :175-179).  The same container with the
  // same insertion sequence walks in the same order, which fixes the order
  // in which AddModel lists the polygons.  Library materials no face uses
  // get an empty list.
  std::unordered_map<std::string, Material> table;
  table[kDefaultMaterialName] = deflt;
  for (auto &e : named) table[e.first] = e.second;
  for (auto &e : table) {
    m_materials.push_back(e.second);
    list_of(e.second);
  }
  return true;
}

// ------------------------------------------------------------ SceneFromExternalFile
namespace {
// parseVector3 (SceneFromExternalFile.cpp:15-44): "x, y, z", each optionally
// followed by " deg" (degrees to radians).
Vector3 parse_vec3(const std::string &s) {
  std::vector<std::string> p = split(s, ',');
  double c[3] = {0, 0, 0};
  for (size_t i = 0; i < 3 && i < p.size(); i++) {
    std::vector<std::string> w = split(trim(p[i]), ' ');
    if (w.empty()) continue;
    c[i] = std::atof(trim(w[0]).c_str());
    if (w.size() >= 2 && w[1] == "deg") c[i] = c[i] * kPi / 180;
  }
  return Vector3(c[0], c[1], c[2]);
}
bool parse_xyz(const std::string &s, Vector3 &v) {
  std::vector<std::string> p = split(s, ',');
  if (p.size() != 3) return false;
  v = Vector3(std::atof(trim(p[0]).c_str()), std::atof(trim(p[1]).c_str()), std::atof(trim(p[2]).c_str()));
  return true;
}
void parse_rates(const std::string &s, std::vector<double> &r) {
  for (const std::string &x : split(s, ',')) r.push_back(std::atof(x.c_str()));
}
}  // namespace

SceneFromExternalFile::SceneFromExternalFile(const std::string &file) { m_isValid = ReadFromFile(file); }

bool SceneFromExternalFile::ReadFromFile(const std::string &file) {
  std::ifstream in(file.c_str());
  if (!in || in.bad() || in.eof()) return false;
  int line_number = 1;
  std::vector<LinePair> lines;
  auto pair_of = [&](const std::string &l) {
    std::vector<std::string> d = split(l, ':');
    if (d.size() == 1) lines.emplace_back(trim(d[0]), "");
    else if (d.size() >= 2) lines.emplace_back(trim(d[0]), trim(d[1]));
  };
  // header up to "Objects:"
  do {
    std::string line;
    std::getline(in, line);
    line = ltrim(line);
    if (line == "Objects:") break;
    if (!line.empty() && line.find('#') != 0) pair_of(line);
    line_number++;
  } while (!in.eof());
  if (in.eof()) return false;
  if (!ReadHeader(lines)) return false;
  // object blocks "Type {" ... "}"
  enum { NONE, MESH, SPHERE, FLOOR, SPHERE_LIGHT, INFINITE_FLOOR } type = NONE;
  bool in_block = false;
  lines.clear();
  do {
    std::string line;
    std::getline(in, line);
    line = ltrim(line);
    if (line.empty() || line[0] == '#') {
      line_number++;
      continue;
    }
    if (line.find('{') != std::string::npos) {
      if (in_block) {
        std::cerr << "Nested Object Definition is not allowded!!: line " << line_number << std::endl;
        return false;
      }
      std::string t = rtrim(rtrim(line, "{"));
      type = t == "Obj Mesh" ? MESH : t == "Sphere" ? SPHERE : t == "Floor" ? FLOOR
           : t == "SphereLight" ? SPHERE_LIGHT : t == "InfiniteFloor" ? INFINITE_FLOOR : NONE;
      if (type == NONE) {
        std::cerr << "Undefined object type: " << t << " line " << line_number << std::endl;
        return false;
      }
      in_block = true;
    } else if (line.find('}') == 0) {
      if (!in_block) {
        std::cerr << "Inconsistent delimiter: line " << line_number << std::endl;
        return false;
      }
      bool ok = type == MESH ? ReadMesh(lines) : type == FLOOR ? ReadFloor(lines)
              : type == SPHERE || type == SPHERE_LIGHT ? ReadSphere(lines, type == SPHERE_LIGHT)
              : type == INFINITE_FLOOR ? ReadInfiniteFloor(lines) : false;
      if (!ok) {
        std::cerr << "failed to load object: line " << line_number << std::endl;
        return false;
      }
      type = NONE;
      in_block = false;
      lines.clear();
    } else if (in_block) {
      pair_of(line);
    }
    line_number++;
  } while (!in.eof());
  return true;
}

bool SceneFromExternalFile::ReadHeader(const std::vector<LinePair> &lines) {
  m_iblFileName = m_spacePartitioningMethod = "";
  for (const LinePair &l : lines) {
    if (l.first == "Base Dir" || l.first == "BaseDir") {
      m_baseDir = l.second;
      if (!m_baseDir.empty() && m_baseDir.back() != '/') m_baseDir.push_back('/');
    } else if (l.first == "IBL") {
      m_iblFileName = m_baseDir + l.second;
      std::unique_ptr<IBL> ibl(new IBL(m_iblFileName));
      if (ibl->IsValid()) SetIBL(std::move(ibl));
    } else if (l.first == "Space Partitioning") {
      m_spacePartitioningMethod = l.second;  // the BVH is built on upload whatever the value
    } else {
      std::cerr << "undefined header information: " << l.first << std::endl;
      return false;
    }
  }
  return true;
}

// "emission r, g, b, TYPE, color r, g, b[, ior][, texture]" (ParseMaterial,
// SceneFromExternalFile.cpp:214-274)
bool SceneFromExternalFile::ParseMaterial(const std::string &s, Material &mat) {
  std::vector<std::string> d = split(s, ',');
  if (d.size() < 7) {
    std::cerr << "insufficient material definition: " << s << std::endl;
    return false;
  }
  Vector3 em(std::atof(trim(d[0]).c_str()), std::atof(trim(d[1]).c_str()), std::atof(trim(d[2]).c_str()));
  std::string t = trim(d[3]);
  Material::REFLECTION_TYPE rt;
  if (t == "LAMBERT") rt = Material::REFLECTION_TYPE_LAMBERT;
  else if (t == "SPECULAR") rt = Material::REFLECTION_TYPE_SPECULAR;
  else if (t == "REFRACTION") rt = Material::REFLECTION_TYPE_REFRACTION;
  else {
    std::cerr << "undefined reflection type: " << t << std::endl;
    return false;
  }
  Vector3 col(std::atof(trim(d[4]).c_str()), std::atof(trim(d[5]).c_str()), std::atof(trim(d[6]).c_str()));
  double ior = 0.0;
  const bool refr = rt == Material::REFLECTION_TYPE_REFRACTION;
  if (refr) {
    if (d.size() < 8) {
      std::cerr << "reflectance is not defined: " << s << std::endl;
      return false;
    }
    ior = std::atof(trim(d[7]).c_str());
  }
  mat = Material(rt, em, col, ior);
  if (d.size() >= (refr ? 9u : 8u))
    mat.texture_id = ImageHandler::GetInstance().LoadFromFile(m_baseDir + trim(d[refr ? 8 : 7]));
  return true;
}

// ProcessMaterials (SceneFromExternalFile.cpp:279-298): the first material
// takes rate[0], the others are appended and the rates normalized once.
bool SceneFromExternalFile::ProcessMaterials(SceneObject *obj, std::vector<Material> &materials,
                                             std::vector<double> &rates) {
  if (materials.size() == 1) rates.assign(1, 1.0);
  if (materials.size() != rates.size() || materials.empty()) return false;
  obj->GetMaterial(0)->rate_ = rates[0];
  for (size_t i = 1; i < materials.size(); i++) obj->AddMaterial(materials[i], rates[i], i + 1 == materials.size());
  return true;
}

bool SceneFromExternalFile::ReadSphere(const std::vector<LinePair> &lines, bool isLight) {
  double radius = 0.0;
  Vector3 pos;
  bool valid = true, in_bvh = true;
  std::vector<Material> mats;
  std::vector<double> rates;
  for (const LinePair &l : lines) {
    if (l.first == "Radius") radius = std::atof(l.second.c_str());
    else if (l.first == "Position") { if (!parse_xyz(l.second, pos)) valid = false; }
    else if (l.first == "Material") {
      Material m;
      if (ParseMaterial(l.second, m)) mats.push_back(m);
      else valid = false;
    } else if (l.first == "MaterialRates") parse_rates(l.second, rates);
    else if (l.first == "Space Partitioning") {
      if (l.second == "True") in_bvh = true;
      else if (l.second == "False") in_bvh = false;
    }
  }
  if (!valid || mats.empty()) return false;
  SceneObject *obj = isLight ? (SceneObject *)new SphereLight(radius, pos, mats[0]) : new Sphere(radius, pos, mats[0]);
  AddObject(obj, true, in_bvh);
  return ProcessMaterials(obj, mats, rates);
}

bool SceneFromExternalFile::ReadFloor(const std::vector<LinePair> &lines) {
  double sx = 0, sz = 0;
  bool valid = true;
  Vector3 pos;
  std::vector<Material> mats;
  std::vector<double> rates;
  enum { XZ_YUP, XZ_YDOWN, XY_ZUP, XY_ZDOWN, YZ_XUP, YZ_XDOWN } type = XZ_YUP;
  for (const LinePair &l : lines) {
    if (l.first == "Size") {
      std::vector<std::string> p = split(l.second, ',');
      if (p.size() != 2) { valid = false; continue; }
      sx = std::atof(trim(p[0]).c_str());
      sz = std::atof(trim(p[1]).c_str());
    } else if (l.first == "Position") { if (!parse_xyz(l.second, pos)) valid = false; }
    else if (l.first == "Material") {
      Material m;
      if (ParseMaterial(l.second, m)) mats.push_back(m);
      else valid = false;
    } else if (l.first == "MaterialRates") parse_rates(l.second, rates);
    else if (l.first == "Type") {
      const std::string &t = l.second;
      if (t == "XZ_yUp") type = XZ_YUP;
      else if (t == "XZ_yDown") type = XZ_YDOWN;
      else if (t == "XY_zUp") type = XY_ZUP;
      else if (t == "XY_zDown") type = XY_ZDOWN;
      else if (t == "YZ_xUp") type = YZ_XUP;
      else if (t == "YZ_xDown") type = YZ_XDOWN;
    }
  }
  if (!valid || mats.empty()) return false;
  std::vector<SceneObject *> objs;
  switch (type) {
    case XZ_YUP: objs = AddFloorXZ_yUp(sx, sz, pos, mats[0]); break;
    case XZ_YDOWN: objs = AddFloorXZ_yDown(sx, sz, pos, mats[0]); break;
    case XY_ZUP: objs = AddFloorXY_zUp(sx, sz, pos, mats[0]); break;
    case XY_ZDOWN: objs = AddFloorXY_zDown(sx, sz, pos, mats[0]); break;
    case YZ_XUP: objs = AddFloorYZ_xUp(sx, sz, pos, mats[0]); break;
    case YZ_XDOWN: objs = AddFloorYZ_xDown(sx, sz, pos, mats[0]); break;
  }
  for (SceneObject *o : objs) {
    std::vector<double> r(rates);
    if (!ProcessMaterials(o, mats, r)) return false;
  }
  return true;
}

bool SceneFromExternalFile::ReadInfiniteFloor(const std::vector<LinePair> &lines) {
  double y = 0;
  std::vector<Material> mats;
  std::vector<double> rates;
  for (const LinePair &l : lines) {
    if (l.first == "Y") y = std::atof(l.second.c_str());
    else if (l.first == "Material") {
      Material m;
      if (ParseMaterial(l.second, m)) mats.push_back(m);
      else return false;
    } else if (l.first == "MaterialRates") parse_rates(l.second, rates);
  }
  if (mats.empty()) return false;
  SceneObject *o = AddInfiniteFLoor(y, mats[0]);
  ProcessMaterials(o, mats, rates);
  return true;
}

bool SceneFromExternalFile::ReadMesh(const std::vector<LinePair> &lines) {
  std::string name;
  Vector3 pos, scale(1, 1, 1), rot;
  bool in_bvh = true;
  for (const LinePair &l : lines) {
    if (l.first == "FileName") name = m_baseDir + l.second;
    else if (l.first == "Position") pos = parse_vec3(l.second);
    else if (l.first == "Scaling") scale = parse_vec3(l.second);
    else if (l.first == "Rotation") rot = parse_vec3(l.second);
    else if (l.first == "Space Partitioning") {
      if (l.second == "True") in_bvh = true;
      else if (l.second == "False") in_bvh = false;
    }
  }
  std::unique_ptr<Model> model(new Model);
  if (!model->ReadFromObj(name)) return false;
  model->Transform(pos, scale,
                   Matrix::RotateAroundVector(Vector3(0, 0, 1), rot.z) * Matrix::RotateAroundVector(Vector3(0, 1, 0), rot.y) *
                       Matrix::RotateAroundVector(Vector3(1, 0, 0), rot.x));
  AddModel(model.release(), true, in_bvh);
  return true;
}

}  // namespace OmochiRenderer
