// Bake output of one tank's current frame, in the layout of the reference's Blender cache
// (blender_addon/suiri_lab/bake.py:226-249, file names bake.py:24-30):
//   frame_%05d.npz    per-frame arrays (numpy .npz: a stored zip of .npy files)
//   surface_%05d.exr  maps over the water footprint: R top height, G ripple slope d/dX, B d/dY
//                     (exr.py:1-41: uncompressed scanlines of 32-bit float B, G, R)
// Positions and velocities are in Blender axes (X, Y, Z) = (x, -z, y) of the solver's
// (space.py:1-15). The cache's surface mesh (meshing.surface) and its floor-light map belong to the
// renderer and are not produced: the npz carries the live particles instead.
//
// Host code only: it reads the tank through the library's C ABI.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "suiri.h"

namespace {

uint32_t crc_table[256];
bool crc_ready = false;
uint32_t crc32(const uint8_t *p, size_t n) {
  if (!crc_ready) {
    for (uint32_t i = 0; i < 256; i++) {
      uint32_t c = i;
      for (int k = 0; k < 8; k++) c = (c & 1u) ? 0xedb88320u ^ (c >> 1) : c >> 1;
      crc_table[i] = c;
    }
    crc_ready = true;
  }
  uint32_t c = 0xffffffffu;
  for (size_t i = 0; i < n; i++) c = crc_table[(c ^ p[i]) & 0xffu] ^ (c >> 8);
  return c ^ 0xffffffffu;
}

struct Npy {
  std::string name;
  std::vector<uint8_t> bytes;  // the whole .npy file
};

// .npy version 1.0: magic, version, header length, a dict literal padded with spaces to a multiple
// of 64 bytes including the terminating newline, then the little-endian data in C order
Npy npy(const std::string &name, const char *descr, const std::vector<size_t> &shape, const void *data, size_t bytes) {
  std::string dims;  // "(N, 3)", "(N,)" or "()"
  for (size_t i = 0; i < shape.size(); i++) dims += (i ? ", " : "") + std::to_string(shape[i]);
  if (shape.size() == 1) dims += ",";
  std::string hdr = std::string("{'descr': '") + descr + "', 'fortran_order': False, 'shape': (" + dims + "), }";
  size_t total = 10 + hdr.size() + 1;
  hdr.append((64 - total % 64) % 64, ' ');
  hdr.push_back('\n');
  Npy out;
  out.name = name + ".npy";
  const uint8_t magic[8] = {0x93, 'N', 'U', 'M', 'P', 'Y', 1, 0};
  out.bytes.insert(out.bytes.end(), magic, magic + 8);
  const uint16_t hl = (uint16_t)hdr.size();
  out.bytes.push_back((uint8_t)(hl & 0xff));
  out.bytes.push_back((uint8_t)(hl >> 8));
  out.bytes.insert(out.bytes.end(), hdr.begin(), hdr.end());
  const uint8_t *d = (const uint8_t *)data;
  out.bytes.insert(out.bytes.end(), d, d + bytes);
  return out;
}

void put16(std::vector<uint8_t> &b, uint32_t v) { b.push_back(v & 0xff); b.push_back((v >> 8) & 0xff); }
void put32(std::vector<uint8_t> &b, uint32_t v) { put16(b, v & 0xffff); put16(b, v >> 16); }

// a zip archive with stored (uncompressed) members, what numpy.savez writes
bool write_npz(const std::string &path, const std::vector<Npy> &members) {
  FILE *f = fopen(path.c_str(), "wb");
  if (!f) return false;
  std::vector<uint8_t> central;
  uint32_t offset = 0;
  bool ok = true;
  for (const Npy &m : members) {
    if (m.bytes.size() >= 0xffffffffull) { ok = false; break; }  // no zip64 here
    const uint32_t crc = crc32(m.bytes.data(), m.bytes.size()), size = (uint32_t)m.bytes.size();
    std::vector<uint8_t> lh;
    put32(lh, 0x04034b50u); put16(lh, 20); put16(lh, 0); put16(lh, 0); put16(lh, 0); put16(lh, 0x21);
    put32(lh, crc); put32(lh, size); put32(lh, size); put16(lh, (uint32_t)m.name.size()); put16(lh, 0);
    lh.insert(lh.end(), m.name.begin(), m.name.end());
    ok &= fwrite(lh.data(), 1, lh.size(), f) == lh.size();
    ok &= fwrite(m.bytes.data(), 1, m.bytes.size(), f) == m.bytes.size();
    put32(central, 0x02014b50u); put16(central, 20); put16(central, 20); put16(central, 0); put16(central, 0);
    put16(central, 0); put16(central, 0x21); put32(central, crc); put32(central, size); put32(central, size);
    put16(central, (uint32_t)m.name.size()); put16(central, 0); put16(central, 0); put16(central, 0);
    put16(central, 0); put32(central, 0); put32(central, offset);
    central.insert(central.end(), m.name.begin(), m.name.end());
    offset += (uint32_t)(lh.size() + m.bytes.size());
  }
  std::vector<uint8_t> end;
  put32(end, 0x06054b50u); put16(end, 0); put16(end, 0); put16(end, (uint32_t)members.size());
  put16(end, (uint32_t)members.size()); put32(end, (uint32_t)central.size()); put32(end, offset); put16(end, 0);
  ok &= fwrite(central.data(), 1, central.size(), f) == central.size();
  ok &= fwrite(end.data(), 1, end.size(), f) == end.size();
  ok &= fclose(f) == 0;
  return ok;
}

void attr(std::vector<uint8_t> &h, const char *name, const char *type, const std::vector<uint8_t> &data) {
  h.insert(h.end(), name, name + strlen(name) + 1);
  h.insert(h.end(), type, type + strlen(type) + 1);
  put32(h, (uint32_t)data.size());
  h.insert(h.end(), data.begin(), data.end());
}
std::vector<uint8_t> f32le(std::initializer_list<float> v) {
  std::vector<uint8_t> b;
  for (float x : v) { uint32_t u; memcpy(&u, &x, 4); put32(b, u); }
  return b;
}

// OpenEXR, single part, scanline, no compression, three FLOAT channels B, G, R (alphabetical), one
// scanline per block; r, g, b are H x W row-major with row 0 at the top of the picture
bool write_exr(const std::string &path, uint32_t W, uint32_t H, const float *r, const float *g, const float *b) {
  std::vector<uint8_t> h = {0x76, 0x2f, 0x31, 0x01};
  put32(h, 2);
  std::vector<uint8_t> ch;
  for (const char *c : {"B", "G", "R"}) {
    ch.push_back((uint8_t)c[0]); ch.push_back(0);
    put32(ch, 2);                                   // FLOAT
    ch.push_back(0); ch.push_back(0); ch.push_back(0); ch.push_back(0);  // pLinear + reserved
    put32(ch, 1); put32(ch, 1);                     // x, y sampling
  }
  ch.push_back(0);
  attr(h, "channels", "chlist", ch);
  attr(h, "compression", "compression", {0});
  std::vector<uint8_t> box;
  put32(box, 0); put32(box, 0); put32(box, W - 1); put32(box, H - 1);
  attr(h, "dataWindow", "box2i", box);
  attr(h, "displayWindow", "box2i", box);
  attr(h, "lineOrder", "lineOrder", {0});
  attr(h, "pixelAspectRatio", "float", f32le({1.0f}));
  attr(h, "screenWindowCenter", "v2f", f32le({0.0f, 0.0f}));
  attr(h, "screenWindowWidth", "float", f32le({1.0f}));
  h.push_back(0);
  const uint64_t line = 8 + 12ull * W, start = h.size() + 8ull * H;
  for (uint32_t y = 0; y < H; y++) {
    const uint64_t o = start + y * line;
    put32(h, (uint32_t)(o & 0xffffffffu));
    put32(h, (uint32_t)(o >> 32));
  }
  FILE *f = fopen(path.c_str(), "wb");
  if (!f) return false;
  bool ok = fwrite(h.data(), 1, h.size(), f) == h.size();
  std::vector<uint8_t> row;
  for (uint32_t y = 0; y < H && ok; y++) {
    row.clear();
    put32(row, y);
    put32(row, 12 * W);
    for (const float *c : {b, g, r}) {
      const uint8_t *p = (const uint8_t *)(c + (size_t)y * W);
      row.insert(row.end(), p, p + 4ull * W);
    }
    ok &= fwrite(row.data(), 1, row.size(), f) == row.size();
  }
  ok &= fclose(f) == 0;
  return ok;
}

}  // namespace

extern "C" int sl_write_exr_rgb(const char *path, uint32_t W, uint32_t H, const float *r, const float *g,
                                const float *b) {
  if (!path || !W || !H || !r || !g || !b) return SL_EARG;
  return write_exr(path, W, H, r, g, b) ? SL_OK : SL_EARG;
}

extern "C" int sl_bake_frame(sl_sim *s, const char *dir, uint32_t frame) {
  if (!s || !dir) return SL_EARG;
  sl_grid_t g;
  sl_stats_t st;
  int r;
  if ((r = sl_grid_info(s, &g)) || (r = sl_stats(s, &st))) return r;
  const std::string base = std::string(dir) + "/";
  char name[64];
  // particles (reference layout: pos.xyz, alive, vel.xyz, mark) -> live positions and velocities
  std::vector<float> parts(sl_field_bytes(s, SL_F_PARTICLES) / 4);
  if ((r = sl_readback(s, SL_F_PARTICLES, parts.data(), parts.size() * 4))) return r;
  std::vector<float> pos, vel;
  pos.reserve(3ull * st.count);
  vel.reserve(3ull * st.count);
  for (size_t i = 0; i < st.count && 8 * i + 7 < parts.size(); i++) {
    const float *p = &parts[8 * i];
    if (p[3] < 0.5f) continue;
    pos.insert(pos.end(), {p[0], -p[2], p[1]});
    vel.insert(vel.end(), {p[4], -p[6], p[5]});
  }
  std::vector<float>().swap(parts);
  // whitewater (pos.xyz + life, vel.xyz + kind) by kind: 0 spray, 1 foam, 2 bubble
  std::vector<float> ww(sl_field_bytes(s, SL_F_WW) / 4), kinds[3];
  if (!ww.empty() && (r = sl_readback(s, SL_F_WW, ww.data(), ww.size() * 4))) return r;
  for (size_t i = 0; i < st.ww_count && 8 * i + 7 < ww.size(); i++) {
    const float *p = &ww[8 * i];
    const int k = (int)lrintf(p[7]);
    if (k >= 0 && k < 3) kinds[k].insert(kinds[k].end(), {p[0], -p[2], p[1]});
  }
  const float liters = (float)(st.count * g.vp * 1000.0), time = (float)st.time;
  const int32_t count = (int32_t)st.count;
  std::vector<Npy> m;
  m.push_back(npy("pos", "<f4", {pos.size() / 3, 3}, pos.data(), pos.size() * 4));
  m.push_back(npy("vel", "<f4", {vel.size() / 3, 3}, vel.data(), vel.size() * 4));
  const char *kn[3] = {"spray", "foam", "bubble"};
  for (int k = 0; k < 3; k++) m.push_back(npy(kn[k], "<f4", {kinds[k].size() / 3, 3}, kinds[k].data(), kinds[k].size() * 4));
  m.push_back(npy("liters", "<f4", {}, &liters, 4));
  m.push_back(npy("particles", "<i4", {}, &count, 4));
  m.push_back(npy("time", "<f4", {}, &time, 4));
  snprintf(name, sizeof name, "frame_%05u.npz", frame);
  if (!write_npz(base + name, m)) return SL_EARG;
  // maps over the footprint: the top-height map and the ripple slopes at its texel centres
  // (bilinear in the ripple vertices, bake.py:240-246 with solver.py:909-919)
  const size_t tb = sl_field_bytes(s, SL_F_TOPH);
  if (!g.R || !tb) return SL_OK;
  const uint32_t W = (uint32_t)llround(sqrt((double)(tb / 4)));
  std::vector<float> top(tb / 4), tex(sl_field_bytes(s, SL_F_RTEX) / 4);
  if ((r = sl_readback(s, SL_F_TOPH, top.data(), tb)) || (r = sl_readback(s, SL_F_RTEX, tex.data(), tex.size() * 4)))
    return r;
  const uint32_t R = g.R, R1 = R + 1;
  std::vector<float> sx((size_t)W * W), sz((size_t)W * W);
  for (uint32_t j = 0; j < W; j++)
    for (uint32_t i = 0; i < W; i++) {
      const double qx = std::min(std::max((i + 0.5) / W * R, 0.0), (double)R);
      const double qz = std::min(std::max((j + 0.5) / W * R, 0.0), (double)R);
      const uint32_t bx = std::min((uint32_t)floor(qx), R - 1), bz = std::min((uint32_t)floor(qz), R - 1);
      const double fx = qx - bx, fz = qz - bz;
      double out[2];
      for (int c = 0; c < 2; c++) {
        auto F = [&](uint32_t zz, uint32_t xx) { return (double)tex[4 * ((size_t)zz * R1 + xx) + 1 + c]; };
        out[c] = (F(bz, bx) * (1 - fx) + F(bz, bx + 1) * fx) * (1 - fz) + (F(bz + 1, bx) * (1 - fx) + F(bz + 1, bx + 1) * fx) * fz;
      }
      sx[(size_t)j * W + i] = (float)out[0];
      sz[(size_t)j * W + i] = (float)(-out[1]);
    }
  snprintf(name, sizeof name, "surface_%05u.exr", frame);
  return write_exr(base + name, W, W, top.data(), sx.data(), sz.data()) ? SL_OK : SL_EARG;
}
