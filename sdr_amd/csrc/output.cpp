// output.cpp — frame output of the C ABI: the reference's particle writers and reader
// (output.rs:23-122), its mesh writers (mesh.rs:356-456), and an asynchronous frame
// writer that overlaps a solver's substeps with the download and the file write.
//
// Text coordinates follow Rust's f64 Display: the shortest decimal digit string that
// parses back to the same double (std::to_chars), laid out positionally without an
// exponent ("0", "-0", "0.00001", "100000000000000000000").
#include <hip/hip_runtime.h>

#include <sys/stat.h>

#include <algorithm>
#include <cerrno>
#include <charconv>
#include <cmath>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/sdr_amd.h"
#include "host.hpp"

namespace sdr {
namespace {

// Rust's `{}` for f64 (core::fmt::float, shortest representation, no exponent).
void put_f64(std::string& out, double v) {
  if (std::isnan(v)) {
    out += "NaN";
    return;
  }
  if (std::isinf(v)) {
    out += v < 0 ? "-inf" : "inf";
    return;
  }
  if (v == 0.0) {
    out += std::signbit(v) ? "-0" : "0";
    return;
  }
  char sci[40];
  const auto res = std::to_chars(sci, sci + sizeof(sci), v, std::chars_format::scientific);
  const char* p = sci;
  const char* end = res.ptr;
  if (*p == '-') {
    out += '-';
    ++p;
  }
  char digits[24];
  int nd = 0;
  while (p < end && *p != 'e') {
    if (*p != '.') digits[nd++] = *p;
    ++p;
  }
  // value = d0.d1d2... x 10^exp10 (to_chars writes no terminator: parse [p + 1, end))
  const char* e = p + 1;
  if (e < end && *e == '+') ++e;
  int exp10 = 0;
  std::from_chars(e, end, exp10);
  const int point = exp10 + 1;          // digits before the decimal point
  if (point <= 0) {
    out += "0.";
    out.append((size_t)(-point), '0');
    out.append(digits, (size_t)nd);
  } else if (point >= nd) {
    out.append(digits, (size_t)nd);
    out.append((size_t)(point - nd), '0');
  } else {
    out.append(digits, (size_t)point);
    out += '.';
    out.append(digits + point, (size_t)(nd - point));
  }
}

void put_u64(std::string& out, uint64_t v) {
  char b[24];
  const auto r = std::to_chars(b, b + sizeof(b), v);
  out.append(b, (size_t)(r.ptr - b));
}

struct File {
  FILE* f = nullptr;
  std::string path;
  File(const char* p, const char* mode) : path(p) {
    f = std::fopen(p, mode);
    if (!f) throw Error(SDR_E_IO, "creating " + path + ": " + std::strerror(errno));
  }
  ~File() {
    if (f) std::fclose(f);
  }
  void write(const void* d, size_t n) {
    if (n && std::fwrite(d, 1, n, f) != n) throw Error(SDR_E_IO, "writing " + path + ": " + std::strerror(errno));
  }
  void write(const std::string& s) { write(s.data(), s.size()); }
  void close() {
    const int rc = std::fclose(f);
    f = nullptr;
    if (rc != 0) throw Error(SDR_E_IO, "writing " + path + ": " + std::strerror(errno));
  }
};

// `create_dir_all(parent)` of save_obj / save_stl (mesh.rs:358-360); failures ignored.
void make_parent_dirs(const std::string& path) {
  for (size_t i = 1; i < path.size(); ++i)
    if (path[i] == '/') (void)mkdir(path.substr(0, i).c_str(), 0777);
}

// "x y z\n" lines for points [lo, hi).
void format_points(const double* pos, uint64_t lo, uint64_t hi, std::string& out) {
  out.clear();
  out.reserve((size_t)(hi - lo) * 60);
  for (uint64_t q = lo; q < hi; ++q) {
    put_f64(out, pos[3 * q]);
    out += ' ';
    put_f64(out, pos[3 * q + 1]);
    out += ' ';
    put_f64(out, pos[3 * q + 2]);
    out += '\n';
  }
}

// output.rs:23-47.  The POINTS block is formatted in chunks on up to 16 threads and
// written in order.
void write_vtk(const char* path, const double* pos, uint64_t n) {
  File f(path, "wb");
  std::string s;
  s += "# vtk DataFile Version 4.2\nsdr irrigation particles\nASCII\nDATASET UNSTRUCTURED_GRID\nPOINTS ";
  put_u64(s, n);
  s += " double\n";
  f.write(s);
  constexpr uint64_t kChunk = 1 << 18;
  const uint64_t nchunks = (n + kChunk - 1) / kChunk;
  unsigned hw = std::thread::hardware_concurrency();
  const unsigned nthreads = (unsigned)std::min<uint64_t>(std::max(1u, std::min(hw, 16u)), std::max<uint64_t>(nchunks, 1));
  std::vector<std::string> buf(nthreads);
  for (uint64_t c0 = 0; c0 < nchunks; c0 += nthreads) {
    const uint64_t batch = std::min<uint64_t>(nthreads, nchunks - c0);
    std::vector<std::thread> th;
    for (uint64_t b = 1; b < batch; ++b)
      th.emplace_back([&, b] {
        const uint64_t lo = (c0 + b) * kChunk;
        format_points(pos, lo, std::min(n, lo + kChunk), buf[b]);
      });
    format_points(pos, c0 * kChunk, std::min(n, c0 * kChunk + kChunk), buf[0]);
    for (auto& t : th) t.join();
    for (uint64_t b = 0; b < batch; ++b) f.write(buf[b]);
  }
  // one VTK_VERTEX (type 1) cell per point: "1 <index>"
  s.clear();
  s += "CELLS ";
  put_u64(s, n);
  s += ' ';
  put_u64(s, n * 2);
  s += '\n';
  for (uint64_t q = 0; q < n; ++q) {
    s += "1 ";
    put_u64(s, q);
    s += '\n';
    if (s.size() > (8u << 20)) f.write(s), s.clear();
  }
  s += "CELL_TYPES ";
  put_u64(s, n);
  s += '\n';
  for (uint64_t q = 0; q < n; ++q) {
    s += "1\n";
    if (s.size() > (8u << 20)) f.write(s), s.clear();
  }
  f.write(s);
  f.close();
}

// output.rs:99-112, from f64 positions (`p.x as f32`) or from f32 triples already
// converted on the device.
void write_xyz(const char* path, const double* pos64, const float* pos32, uint64_t n) {
  File f(path, "wb");
  if (pos32) {
    f.write(pos32, (size_t)n * 12);
  } else {
    std::vector<float> b;
    constexpr uint64_t kChunk = 1 << 20;
    for (uint64_t lo = 0; lo < n; lo += kChunk) {
      const uint64_t hi = std::min(n, lo + kChunk);
      b.resize((size_t)(hi - lo) * 3);
      for (uint64_t q = lo; q < hi; ++q)
        for (int a = 0; a < 3; ++a) b[(size_t)(3 * (q - lo) + a)] = (float)pos64[3 * q + a];
      f.write(b.data(), b.size() * 4);
    }
  }
  f.close();
}

// Rust's `str::parse::<f64>` on one whitespace-free token: decimal or exponent form,
// an optional sign, "inf" / "infinity" / "nan" in any case.
bool parse_f64(const char* b, const char* e, double* out) {
  bool neg = false;
  if (b < e && (*b == '+' || *b == '-')) neg = *b == '-', ++b;
  if (b == e || *b == '+' || *b == '-') return false;
  std::string low;
  for (const char* p = b; p < e && low.size() < 9; ++p) low += (char)std::tolower((unsigned char)*p);
  if ((size_t)(e - b) == low.size() && (low == "inf" || low == "infinity" || low == "nan")) {
    *out = low == "nan" ? NAN : (neg ? -INFINITY : INFINITY);
    return true;
  }
  for (const char* p = b; p < e; ++p)
    if (!(std::isdigit((unsigned char)*p) || *p == '.' || *p == 'e' || *p == 'E' || *p == '+' || *p == '-'))
      return false;
  double v = 0;
  const auto r = std::from_chars(b, e, v, std::chars_format::general);
  if (r.ec == std::errc::result_out_of_range) {
    // Rust rounds an out-of-range literal to inf / 0 instead of failing
    v = std::strtod(std::string(b, e).c_str(), nullptr);
  } else if (r.ec != std::errc() || r.ptr != e) {
    return false;
  }
  *out = neg ? -v : v;
  return true;
}

// output.rs:56-95
uint64_t read_vtk(const char* path, double* pos, uint64_t cap) {
  FILE* f = std::fopen(path, "rb");
  if (!f) throw Error(SDR_E_IO, std::string("opening ") + path + ": " + std::strerror(errno));
  std::string text;
  char chunk[1 << 16];
  size_t got;
  while ((got = std::fread(chunk, 1, sizeof(chunk), f)) > 0) text.append(chunk, got);
  const bool bad = std::ferror(f) != 0;
  std::fclose(f);
  if (bad) throw Error(SDR_E_IO, std::string("reading ") + path);
  auto is_ws = [](char c) { return c == ' ' || c == '\t' || c == '\n' || c == '\r' || c == '\f' || c == '\v'; };
  const char* p = text.data();
  const char* end = p + text.size();
  // find the "POINTS <n> <type>" header: the first line whose first token is POINTS
  uint64_t n = 0;
  bool found = false;
  while (p < end && !found) {
    const char* eol = (const char*)std::memchr(p, '\n', (size_t)(end - p));
    if (!eol) eol = end;
    const char* a = p;
    while (a < eol && is_ws(*a)) ++a;
    const char* b = a;
    while (b < eol && !is_ws(*b)) ++b;
    if (b - a == 6 && std::memcmp(a, "POINTS", 6) == 0) {
      const char* c = b;
      while (c < eol && is_ws(*c)) ++c;
      const char* d = c;
      while (d < eol && !is_ws(*d)) ++d;
      const auto r = std::from_chars(c, d, n);
      if (c == d || r.ec != std::errc() || r.ptr != d)
        throw Error(SDR_E_INVALID, std::string(path) + ": malformed POINTS header");
      found = true;
    }
    p = eol < end ? eol + 1 : end;
  }
  if (!found) throw Error(SDR_E_INVALID, std::string(path) + ": no POINTS block");
  if (!pos) return n;
  if (cap < n) throw Error(SDR_E_INVALID, "output holds fewer points than the file");
  // the next 3 n whitespace-separated tokens are the coordinates
  for (uint64_t t = 0; t < 3 * n; ++t) {
    while (p < end && is_ws(*p)) ++p;
    const char* b = p;
    while (p < end && !is_ws(*p)) ++p;
    if (b == p || !parse_f64(b, p, &pos[t])) throw Error(SDR_E_INVALID, std::string(path) + ": truncated POINTS data");
  }
  return n;
}

std::string frame_name(const char* dir, const char* prefix, uint64_t frame, const char* ext) {
  char num[32];
  std::snprintf(num, sizeof(num), "%06llu", (unsigned long long)frame);
  std::string name = std::string(prefix) + "_" + num + "." + ext;
  // Path::join: an empty base gives the name itself, a base ending in '/' gets no
  // second separator
  std::string d = dir ? dir : "";
  if (d.empty()) return name;
  if (d.back() == '/') return d + name;
  return d + "/" + name;
}

void check_mesh(const double* v, uint64_t nv, const uint32_t* t, uint64_t nt) {
  if ((!v && nv) || (!t && nt)) throw Error(SDR_E_INVALID, "null argument");
  for (uint64_t q = 0; q < 3 * nt; ++q)
    if (t[q] >= nv) throw Error(SDR_E_INVALID, "triangle index out of range");
}

}  // namespace

// Implemented with the solver (solver.cpp): the particle count, and the scatter of the
// positions into `dst` (device; f64 or f32 triples, reference order) on the solver's
// stream followed by `ev` recorded there.
uint64_t solver_particle_total(sdr_solver* s);
void solver_stage_positions(sdr_solver* s, void* dst, bool as_f32, hipEvent_t ev, int* device);

}  // namespace sdr

using sdr::Error;
using sdr::guard;

struct sdr_frame_writer {
  struct Slot {
    sdr::Buf dev;
    sdr::PinnedBuf host;
    hipEvent_t staged = nullptr, copied = nullptr;
    uint64_t n = 0, frame = 0;
    bool busy = false;
  };
  std::string dir, prefix;
  int format = SDR_FORMAT_VTK;
  std::vector<std::unique_ptr<Slot>> slots;
  int device = -1;
  hipStream_t copy = nullptr;
  std::mutex mu;
  std::condition_variable cv;
  std::deque<int> queue;  // slots with a queued copy, in submit order
  uint64_t submitted = 0, written = 0;
  int err_code = SDR_OK;
  std::string err;
  bool stop = false;
  std::thread worker;

  void run() {
    for (;;) {
      int idx;
      {
        std::unique_lock<std::mutex> lk(mu);
        cv.wait(lk, [&] { return stop || !queue.empty(); });
        if (queue.empty()) return;
        idx = queue.front();
        queue.pop_front();
      }
      Slot& sl = *slots[(size_t)idx];
      int code = SDR_OK;
      std::string msg;
      try {
        HIPCHK(hipSetDevice(device));
        HIPCHK(hipEventSynchronize(sl.copied));
        const std::string path =
            sdr::frame_name(dir.c_str(), prefix.c_str(), sl.frame, format == SDR_FORMAT_XYZ ? "xyz" : "vtk");
        if (format == SDR_FORMAT_XYZ) sdr::write_xyz(path.c_str(), nullptr, (const float*)sl.host.p, sl.n);
        else sdr::write_vtk(path.c_str(), (const double*)sl.host.p, sl.n);
      } catch (const Error& e) {
        code = e.code, msg = e.what();
      } catch (const std::exception& e) {
        code = SDR_E_IO, msg = e.what();
      }
      std::lock_guard<std::mutex> lk(mu);
      if (code != SDR_OK && err_code == SDR_OK) err_code = code, err = msg;
      if (code == SDR_OK) ++written;
      sl.busy = false;
      cv.notify_all();
    }
  }
  // Returns (and clears) the first error a write reported.
  void raise_pending() {
    if (err_code == SDR_OK) return;
    const int c = err_code;
    const std::string m = err;
    err_code = SDR_OK, err.clear();
    throw Error(c, m);
  }
};

extern "C" {

int sdr_write_points_vtk(const char* path, const double* pos, uint64_t n) {
  return guard([&] {
    if (!path || (!pos && n)) throw Error(SDR_E_INVALID, "null argument");
    sdr::write_vtk(path, pos, n);
  });
}

int sdr_read_points_vtk(const char* path, double* pos, uint64_t cap, uint64_t* n) {
  return guard([&] {
    if (!path || !n) throw Error(SDR_E_INVALID, "null argument");
    *n = sdr::read_vtk(path, pos, cap);
  });
}

int sdr_write_points_xyz(const char* path, const double* pos, uint64_t n) {
  return guard([&] {
    if (!path || (!pos && n)) throw Error(SDR_E_INVALID, "null argument");
    sdr::write_xyz(path, pos, nullptr, n);
  });
}

int sdr_frame_path(const char* dir, const char* prefix, uint64_t frame, const char* ext, char* buf, size_t len) {
  return guard([&] {
    if (!prefix || !ext || !buf) throw Error(SDR_E_INVALID, "null argument");
    const std::string p = sdr::frame_name(dir, prefix, frame, ext);
    if (p.size() + 1 > len) throw Error(SDR_E_INVALID, "buffer too small for the frame path");
    std::memcpy(buf, p.c_str(), p.size() + 1);
  });
}

// mesh.rs:356-379
int sdr_write_mesh_obj(const char* path, const double* v, uint64_t nv, const uint32_t* t, uint64_t nt) {
  return guard([&] {
    if (!path) throw Error(SDR_E_INVALID, "null argument");
    sdr::check_mesh(v, nv, t, nt);
    sdr::make_parent_dirs(path);
    sdr::File f(path, "wb");
    std::string s = "# sdr parametric sinus mesh\n# vertices: ";
    sdr::put_u64(s, nv);
    s += " faces: ";
    sdr::put_u64(s, nt);
    s += '\n';
    for (uint64_t q = 0; q < nv; ++q) {
      s += "v ";
      sdr::put_f64(s, v[3 * q]);
      s += ' ';
      sdr::put_f64(s, v[3 * q + 1]);
      s += ' ';
      sdr::put_f64(s, v[3 * q + 2]);
      s += '\n';
      if (s.size() > (8u << 20)) f.write(s), s.clear();
    }
    for (uint64_t q = 0; q < nt; ++q) {  // OBJ is 1-based
      s += "f ";
      sdr::put_u64(s, (uint64_t)t[3 * q] + 1);
      s += ' ';
      sdr::put_u64(s, (uint64_t)t[3 * q + 1] + 1);
      s += ' ';
      sdr::put_u64(s, (uint64_t)t[3 * q + 2] + 1);
      s += '\n';
      if (s.size() > (8u << 20)) f.write(s), s.clear();
    }
    f.write(s);
    f.close();
  });
}

// mesh.rs:427-456
int sdr_write_mesh_stl(const char* path, const double* v, uint64_t nv, const uint32_t* t, uint64_t nt) {
  return guard([&] {
    if (!path) throw Error(SDR_E_INVALID, "null argument");
    sdr::check_mesh(v, nv, t, nt);
    if (nt > 0xffffffffull) throw Error(SDR_E_INVALID, "too many triangles for STL");
    sdr::make_parent_dirs(path);
    sdr::File f(path, "wb");
    std::vector<unsigned char> b(84, 0);
    const uint32_t cnt = (uint32_t)nt;
    std::memcpy(&b[80], &cnt, 4);
    f.write(b.data(), b.size());
    std::vector<unsigned char> rec;
    rec.reserve(1 << 20);
    for (uint64_t q = 0; q < nt; ++q) {
      const double* a = v + 3 * (uint64_t)t[3 * q];
      const double* bb = v + 3 * (uint64_t)t[3 * q + 1];
      const double* c = v + 3 * (uint64_t)t[3 * q + 2];
      const double e1[3] = {bb[0] - a[0], bb[1] - a[1], bb[2] - a[2]};
      const double e2[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
      double nrm[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
      const double len = std::sqrt(nrm[0] * nrm[0] + nrm[1] * nrm[1] + nrm[2] * nrm[2]);  // normalize_or_zero
      for (int k = 0; k < 3; ++k) nrm[k] = len > 1e-12 ? nrm[k] / len : 0.0;
      float w[12];
      for (int k = 0; k < 3; ++k) w[k] = (float)nrm[k];
      for (int k = 0; k < 3; ++k) w[3 + k] = (float)a[k], w[6 + k] = (float)bb[k], w[9 + k] = (float)c[k];
      const size_t o = rec.size();
      rec.resize(o + 50, 0);
      std::memcpy(&rec[o], w, 48);  // the 2 attribute bytes stay 0
      if (rec.size() > (8u << 20)) f.write(rec.data(), rec.size()), rec.clear();
    }
    f.write(rec.data(), rec.size());
    f.close();
  });
}

int sdr_frame_writer_new(const char* dir, const char* prefix, int32_t format, uint32_t slots,
                         sdr_frame_writer** out) {
  return guard([&] {
    if (!dir || !prefix || !out) throw Error(SDR_E_INVALID, "null argument");
    if (format != SDR_FORMAT_VTK && format != SDR_FORMAT_XYZ) throw Error(SDR_E_INVALID, "unknown frame format");
    if (slots < 1 || slots > 8) throw Error(SDR_E_INVALID, "slots must be 1..8");
    auto w = std::make_unique<sdr_frame_writer>();
    w->dir = dir, w->prefix = prefix, w->format = format;
    for (uint32_t q = 0; q < slots; ++q) w->slots.push_back(std::make_unique<sdr_frame_writer::Slot>());
    w->worker = std::thread([p = w.get()] { p->run(); });
    *out = w.release();
  });
}

int sdr_frame_writer_submit(sdr_frame_writer* w, sdr_solver* s, uint64_t frame) {
  return guard([&] {
    if (!w || !s) throw Error(SDR_E_INVALID, "null argument");
    std::unique_lock<std::mutex> lk(w->mu);
    w->raise_pending();
    w->cv.wait(lk, [&] {
      for (auto& sl : w->slots)
        if (!sl->busy) return true;
      return false;
    });
    size_t idx = 0;
    while (w->slots[idx]->busy) ++idx;
    auto& sl = *w->slots[idx];
    sl.busy = true;  // reserved; released below if staging fails
    lk.unlock();
    struct Release {
      sdr_frame_writer* w;
      sdr_frame_writer::Slot* sl;
      bool armed = true;
      ~Release() {
        if (!armed) return;
        std::lock_guard<std::mutex> g(w->mu);
        sl->busy = false;
        w->cv.notify_all();
      }
    } release{w, &sl};
    const uint64_t n = sdr::solver_particle_total(s);
    const size_t bytes = (size_t)n * 3 * (w->format == SDR_FORMAT_XYZ ? 4 : 8);
    int dev = -1;
    sl.dev.ensure(bytes);
    sl.host.ensure(bytes ? bytes : 256);
    if (!sl.staged) {
      HIPCHK(hipEventCreateWithFlags(&sl.staged, hipEventDisableTiming));
      HIPCHK(hipEventCreateWithFlags(&sl.copied, hipEventDisableTiming));
    }
    sdr::solver_stage_positions(s, sl.dev.p, w->format == SDR_FORMAT_XYZ, sl.staged, &dev);
    if (w->device < 0) {
      w->device = dev;
      HIPCHK(hipStreamCreateWithFlags(&w->copy, hipStreamNonBlocking));
    } else if (w->device != dev) {
      throw Error(SDR_E_INVALID, "a frame writer serves solvers of one device");
    }
    HIPCHK(hipStreamWaitEvent(w->copy, sl.staged, 0));
    if (bytes) HIPCHK(hipMemcpyAsync(sl.host.p, sl.dev.p, bytes, hipMemcpyDeviceToHost, w->copy));
    HIPCHK(hipEventRecord(sl.copied, w->copy));
    lk.lock();
    release.armed = false;
    sl.n = n, sl.frame = frame;
    w->queue.push_back((int)idx);
    ++w->submitted;
    w->cv.notify_all();
  });
}

int sdr_frame_writer_flush(sdr_frame_writer* w, uint64_t* written) {
  return guard([&] {
    if (!w) throw Error(SDR_E_INVALID, "null argument");
    std::unique_lock<std::mutex> lk(w->mu);
    w->cv.wait(lk, [&] {
      if (!w->queue.empty()) return false;
      for (auto& sl : w->slots)
        if (sl->busy) return false;
      return true;
    });
    if (written) *written = w->written;
    w->raise_pending();
  });
}

int sdr_frame_writer_free(sdr_frame_writer* w) {
  if (!w) return SDR_OK;
  const int rc = sdr_frame_writer_flush(w, nullptr);
  {
    std::lock_guard<std::mutex> lk(w->mu);
    w->stop = true;
    w->cv.notify_all();
  }
  w->worker.join();
  if (w->device >= 0) (void)hipSetDevice(w->device);
  for (auto& sl : w->slots) {
    if (sl->staged) (void)hipEventDestroy(sl->staged);
    if (sl->copied) (void)hipEventDestroy(sl->copied);
  }
  if (w->copy) (void)hipStreamDestroy(w->copy);
  delete w;
  return rc;
}

}  // extern "C"
