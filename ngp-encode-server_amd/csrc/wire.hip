// wire.hip -- renderer endpoint speaking neserver's own protocol.
//
// neserver connects to each `-r ip:port` renderer and, per iteration, sends
// an 8-byte host-endian length followed by a serialized nesproto::FrameRequest
// and blocks for a length-prefixed nesproto::RenderedFrame
// (src/server.cpp:23-112 framing, :140-194 loop; schema proto/nes.proto:5-25).
// This server accepts those connections and answers each request with a frame
// rendered by this library, so an unmodified neserver can use it in place of
// the external instant-ngp renderer.
//
// The proto3 wire codec is written out here (no protoc in this toolchain):
//   Camera        2: is_left (varint)  3: width  4: height  12: matrix
//                 (repeated float, packed or not)  5: fov_y_deg (float; an
//                 additive field the reference does not send -- absent means
//                 the server's default)
//   FrameRequest  1: index  2: camera (message)  3: is_left
//   RenderedFrame 1: index  2: camera  3: is_left  6: frame (RGB24 bytes)
//                 7: depth (GRAY8 bytes)
// Unknown fields are skipped.  GRAY8 depth = clamp(floor(fma(d, 255 / far,
// 0.5))) in FP32, computed on the device by the render (nes_view_out
// depth_gray8), with a configurable far distance (the reference's renderer
// quantises depth itself; its rule is not in the reference).
#include <arpa/inet.h>
#include <netinet/in.h>
#include <sys/socket.h>
#include <unistd.h>

#include <atomic>
#include <cmath>
#include <cstring>
#include <thread>

#include "render.h"

namespace nes {

namespace {

struct Reader {
  const uint8_t *p, *end;
  bool ok = true;
  uint64_t varint() {
    uint64_t v = 0;
    for (int shift = 0; shift < 64; shift += 7) {
      if (p >= end) {
        ok = false;
        return 0;
      }
      const uint8_t b = *p++;
      v |= (uint64_t)(b & 0x7f) << shift;
      if (!(b & 0x80)) return v;
    }
    ok = false;
    return 0;
  }
  uint32_t fixed32() {
    if (end - p < 4) {
      ok = false;
      return 0;
    }
    uint32_t v;
    std::memcpy(&v, p, 4);
    p += 4;
    return v;
  }
  bool skip(int wt) {
    switch (wt) {
      case 0: varint(); break;
      case 1:
        if (end - p < 8) ok = false; else p += 8;
        break;
      case 2: {
        const uint64_t n = varint();
        if (!ok || (uint64_t)(end - p) < n) ok = false; else p += n;
        break;
      }
      case 5: fixed32(); break;
      default: ok = false;
    }
    return ok;
  }
};

float bits_to_float(uint32_t b) {
  float f;
  std::memcpy(&f, &b, 4);
  return f;
}

bool decode_camera(const uint8_t *p, size_t n, nes_camera *cam, bool *has_fov) {
  Reader r{p, p + n};
  int mi = 0;
  while (r.ok && r.p < r.end) {
    const uint64_t key = r.varint();
    if (!r.ok) break;
    const int field = (int)(key >> 3), wt = (int)(key & 7);
    if (field == 0 || (key >> 32)) return false;  // invalid tag, as the protobuf runtime
    if (field == 2 && wt == 0) {
      cam->is_left = r.varint() != 0;
    } else if (field == 3 && wt == 0) {
      cam->width = (uint32_t)r.varint();
    } else if (field == 4 && wt == 0) {
      cam->height = (uint32_t)r.varint();
    } else if (field == 5 && wt == 5) {
      cam->fov_y_deg = bits_to_float(r.fixed32());
      *has_fov = true;
    } else if (field == 12 && wt == 2) {
      const uint64_t len = r.varint();
      if (!r.ok || (uint64_t)(r.end - r.p) < len || len % 4) return false;
      for (uint64_t k = 0; k < len; k += 4) {
        const uint32_t b = r.fixed32();
        if (mi < 12) cam->c2w[mi] = bits_to_float(b);
        ++mi;
      }
    } else if (field == 12 && wt == 5) {
      const uint32_t b = r.fixed32();
      if (mi < 12) cam->c2w[mi] = bits_to_float(b);
      ++mi;
    } else if (!r.skip(wt)) {
      return false;
    }
  }
  return r.ok && mi == 12;
}

void put_varint(std::string &o, uint64_t v) {
  while (v >= 0x80) {
    o.push_back((char)((v & 0x7f) | 0x80));
    v >>= 7;
  }
  o.push_back((char)v);
}

void put_key(std::string &o, int field, int wt) { put_varint(o, ((uint64_t)field << 3) | (uint64_t)wt); }

void put_bytes(std::string &o, int field, const void *data, size_t n) {
  put_key(o, field, 2);
  put_varint(o, n);
  o.append(reinterpret_cast<const char *>(data), n);
}

std::string encode_camera(const nes_camera &c, bool with_fov) {
  std::string o;
  if (c.is_left) {
    put_key(o, 2, 0);
    put_varint(o, 1);
  }
  if (c.width) {
    put_key(o, 3, 0);
    put_varint(o, c.width);
  }
  if (c.height) {
    put_key(o, 4, 0);
    put_varint(o, c.height);
  }
  if (with_fov) {
    put_key(o, 5, 5);
    o.append(reinterpret_cast<const char *>(&c.fov_y_deg), 4);
  }
  put_bytes(o, 12, c.c2w, sizeof(c.c2w));
  return o;
}

}  // namespace

}  // namespace nes

using namespace nes;

extern "C" {

int nes_wire_decode_request(const uint8_t *msg, size_t len, float default_fov,
                            nes_camera *cam) {
  if (!msg || !cam) return NES_EINVAL;
  std::memset(cam, 0, sizeof(*cam));
  cam->fov_y_deg = default_fov;
  Reader r{msg, msg + len};
  bool have_cam = false, has_fov = false, top_left = false;
  while (r.ok && r.p < r.end) {
    const uint64_t key = r.varint();
    if (!r.ok) break;
    const int field = (int)(key >> 3), wt = (int)(key & 7);
    if (field == 0 || (key >> 32)) return NES_EINVAL;  // invalid tag, as the protobuf runtime
    if (field == 1 && wt == 0) {
      cam->index = r.varint();
    } else if (field == 2 && wt == 2) {
      const uint64_t n = r.varint();
      if (!r.ok || (uint64_t)(r.end - r.p) < n) return NES_EINVAL;
      // a second camera would be merged by the protobuf runtime (its matrix
      // then has 24 floats); refused here
      if (have_cam) return NES_EINVAL;
      const float keep_fov = cam->fov_y_deg;
      const uint64_t idx = cam->index;
      if (!decode_camera(r.p, n, cam, &has_fov)) return NES_EINVAL;
      cam->index = idx;
      if (!has_fov) cam->fov_y_deg = keep_fov;
      r.p += n;
      have_cam = true;
    } else if (field == 3 && wt == 0) {
      top_left = r.varint() != 0;
    } else if (!r.skip(wt)) {
      return NES_EINVAL;
    }
  }
  if (!r.ok || !have_cam) return NES_EINVAL;
  // FrameRequest.is_left is the one the reference sets per request
  // (src/server.cpp:144-145) and routes frames by (rendered_frame.h:43)
  cam->is_left = top_left ? 1u : 0u;
  return NES_OK;
}

// Serialize a RenderedFrame; *len receives the size.  Pass out = NULL to
// query the size.
int nes_wire_encode_frame(const nes_camera *cam, const uint8_t *rgb24, const uint8_t *gray8,
                          uint8_t *out, size_t cap, size_t *len) {
  if (!cam || !len) return NES_EINVAL;
  const size_t px = (size_t)cam->width * cam->height;
  std::string o;
  o.reserve(4 * px + 128);
  if (cam->index) {
    put_key(o, 1, 0);
    put_varint(o, cam->index);
  }
  const std::string c = encode_camera(*cam, false);
  put_bytes(o, 2, c.data(), c.size());
  if (cam->is_left) {
    put_key(o, 3, 0);
    put_varint(o, 1);
  }
  if (rgb24) put_bytes(o, 6, rgb24, 3 * px);
  if (gray8) put_bytes(o, 7, gray8, px);
  *len = o.size();
  if (!out) return NES_OK;
  if (cap < o.size()) return NES_ERANGE;
  std::memcpy(out, o.data(), o.size());
  return NES_OK;
}

}  // extern "C"

struct nes_wire_server {
  nes_ctx *ctx = nullptr;
  int precision = NES_PRECISION_MIXED;
  float default_fov = 60.0f, depth_far = 4.0f;
  int listen_fd = -1;
  uint16_t port = 0;
  std::atomic<bool> stop{false};
  std::atomic<uint64_t> served{0};
  std::thread acceptor;
  std::mutex conn_mu;
  std::vector<std::thread> conns;
  std::vector<int> fds;
};

namespace nes {
namespace {

bool read_full(int fd, void *buf, size_t n) {
  uint8_t *p = static_cast<uint8_t *>(buf);
  while (n) {
    const ssize_t r = ::read(fd, p, n);
    if (r < 0 && errno == EINTR) continue;
    if (r <= 0) return false;  // EOF mid-message is an error (src/server.cpp:78-83)
    p += r;
    n -= (size_t)r;
  }
  return true;
}

bool write_full(int fd, const void *buf, size_t n) {
  const uint8_t *p = static_cast<const uint8_t *>(buf);
  while (n) {
    const ssize_t r = ::send(fd, p, n, MSG_NOSIGNAL);
    if (r < 0 && (errno == EINTR || errno == EAGAIN)) continue;
    if (r <= 0) return false;
    p += r;
    n -= (size_t)r;
  }
  return true;
}

void serve_connection(nes_wire_server *s, int fd) {
  std::vector<uint8_t> msg, rgb, gray, out;
  while (!s->stop) {
    uint64_t len = 0;  // host-endian size_t, as src/server.cpp:50
    if (!read_full(fd, &len, sizeof(len))) break;
    if (len > (64u << 20)) break;
    msg.resize(len);
    if (!read_full(fd, msg.data(), len)) break;
    nes_camera cam;
    if (nes_wire_decode_request(msg.data(), len, s->default_fov, &cam) != NES_OK) break;
    if (cam.width < 1 || cam.height < 1 || cam.width > 16384 || cam.height > 16384) break;
    const size_t px = (size_t)cam.width * cam.height;
    rgb.resize(3 * px);
    gray.resize(px);
    // RGB24 and the GRAY8 depth plane both come from the device
    nes_view_out vo{rgb.data(), nullptr, nullptr, nullptr, gray.data(), s->depth_far};
    if (nes_render(s->ctx, s->precision, &cam, 1, &vo, 0, nullptr) != NES_OK) break;
    size_t n = 0;
    nes_wire_encode_frame(&cam, rgb.data(), gray.data(), nullptr, 0, &n);
    out.resize(n);
    nes_wire_encode_frame(&cam, rgb.data(), gray.data(), out.data(), out.size(), &n);
    const uint64_t n64 = n;
    if (!write_full(fd, &n64, sizeof(n64)) || !write_full(fd, out.data(), n)) break;
    s->served++;
  }
  {
    // forget the descriptor before closing it, so nes_wire_server_stop never
    // shuts down a number the process has since reused
    std::lock_guard<std::mutex> lk(s->conn_mu);
    for (int &f : s->fds)
      if (f == fd) f = -1;
  }
  ::close(fd);
}

}  // namespace
}  // namespace nes

extern "C" {

// Listen on addr:port (port 0 picks a free port, reported by
// nes_wire_server_port) and serve every connection on its own thread.
int nes_wire_server_start(nes_ctx *ctx, const char *addr, uint16_t port, int precision,
                          float default_fov, float depth_far, nes_wire_server **out) {
  if (!ctx || !out || !(default_fov > 0.0f) || !(depth_far > 0.0f)) return NES_EINVAL;
  auto s = new nes_wire_server();
  s->ctx = ctx;
  s->precision = precision;
  s->default_fov = default_fov;
  s->depth_far = depth_far;
  s->listen_fd = ::socket(AF_INET, SOCK_STREAM, 0);
  if (s->listen_fd < 0) {
    delete s;
    return NES_EIO;
  }
  int one = 1;
  ::setsockopt(s->listen_fd, SOL_SOCKET, SO_REUSEADDR, &one, sizeof(one));
  sockaddr_in sa{};
  sa.sin_family = AF_INET;
  sa.sin_port = htons(port);
  sa.sin_addr.s_addr = inet_addr(addr ? addr : "127.0.0.1");
  socklen_t sl = sizeof(sa);
  if (::bind(s->listen_fd, (sockaddr *)&sa, sizeof(sa)) < 0 || ::listen(s->listen_fd, 16) < 0 ||
      ::getsockname(s->listen_fd, (sockaddr *)&sa, &sl) < 0) {
    ::close(s->listen_fd);
    delete s;
    return NES_EIO;
  }
  s->port = ntohs(sa.sin_port);
  s->acceptor = std::thread([s] {
    while (!s->stop) {
      const int fd = ::accept(s->listen_fd, nullptr, nullptr);
      if (fd < 0) {
        if (s->stop) break;
        continue;
      }
      std::lock_guard<std::mutex> lk(s->conn_mu);
      s->fds.push_back(fd);
      s->conns.emplace_back(nes::serve_connection, s, fd);
    }
  });
  *out = s;
  return NES_OK;
}

int nes_wire_server_port(nes_wire_server *s) { return s ? (int)s->port : NES_EINVAL; }

uint64_t nes_wire_server_served(nes_wire_server *s) { return s ? s->served.load() : 0; }

int nes_wire_server_stop(nes_wire_server *s) {
  if (!s) return NES_EINVAL;
  s->stop = true;
  ::shutdown(s->listen_fd, SHUT_RDWR);
  ::close(s->listen_fd);
  if (s->acceptor.joinable()) s->acceptor.join();
  {
    std::lock_guard<std::mutex> lk(s->conn_mu);
    for (int fd : s->fds)
      if (fd >= 0) ::shutdown(fd, SHUT_RDWR);
  }
  for (auto &t : s->conns)
    if (t.joinable()) t.join();
  delete s;
  return NES_OK;
}

}  // extern "C"
