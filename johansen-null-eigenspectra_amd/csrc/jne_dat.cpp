// jne_dat.cpp -- the append-only .dat V6 store, its reader, resume and
// percentiles, plus the display helpers the simulation driver and the CLI print with.
//
// Byte format (DATA_FORMAT.md, src/data_storage/file_format.rs:8-9,
// writer.rs:61-64,241-249,262-289):
//   header  "EIGENVALS_V6" | u8 model | u8 dim | u32 LE steps           (18 B)
//   record  ULEB128(seed) | u8 nF | nF x f64 LE
//   trailer "EOF_MARK" | u64 LE record count | u8 nF                     (17 B)
// Files written here are byte-identical to the reference writer's for the
// same records in the same order.
#include <errno.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <string>
#include <thread>
#include <vector>

#include "../../include/jne.h"
#include "jne_host.h"

namespace jne_host {

namespace {

constexpr char kMagic[12] = {'E', 'I', 'G', 'E', 'N', 'V', 'A', 'L', 'S', '_', 'V', '6'};
constexpr char kEofMark[8] = {'E', 'O', 'F', '_', 'M', 'A', 'R', 'K'};
constexpr uint64_t kHeaderBytes = 18;
constexpr uint64_t kTrailerBytes = 17;
constexpr uint64_t kProgressInterval = 10000;  // src/data_storage/config.rs:4

thread_local std::string g_err;

std::string fmt(const char* f, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, f);
  vsnprintf(buf, sizeof buf, f, ap);
  va_end(ap);
  return buf;
}

// The text Rust's std::io::Error shows for an OS error.
std::string os_error(int e) { return fmt("%s (os error %d)", strerror(e), e); }

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

// ULEB128 of a u32 (src/data_storage/uleb128.rs:68-86); returns the length.
inline int uleb128_put(uint32_t v, uint8_t* out) {
  int n = 0;
  do {
    uint8_t b = v & 0x7F;
    v >>= 7;
    if (v) b |= 0x80;
    out[n++] = b;
  } while (v);
  return n;
}

// Sequential reader over a FILE* with a large buffer.
class ByteReader {
 public:
  ByteReader(FILE* f, size_t cap) : f_(f), buf_(cap) {}
  // true and the byte, or false at end of file
  bool get(uint8_t* b) {
    if (pos_ == len_ && !fill()) return false;
    *b = buf_[pos_++];
    return true;
  }
  bool read(void* dst, size_t n) {
    uint8_t* d = static_cast<uint8_t*>(dst);
    while (n) {
      if (pos_ == len_ && !fill()) return false;
      const size_t k = std::min(n, len_ - pos_);
      memcpy(d, buf_.data() + pos_, k);
      pos_ += k;
      d += k;
      n -= k;
    }
    return true;
  }
  uint64_t offset() const { return base_ + pos_; }
  void seek(uint64_t off) {
    fseeko(f_, (off_t)off, SEEK_SET);
    base_ = off;
    pos_ = len_ = 0;
  }

 private:
  bool fill() {
    base_ += len_;
    len_ = fread(buf_.data(), 1, buf_.size(), f_);
    pos_ = 0;
    return len_ > 0;
  }
  FILE* f_;
  std::vector<uint8_t> buf_;
  size_t pos_ = 0, len_ = 0;
  uint64_t base_ = 0;
};

// uleb128::read_from_reader: 0 ok, 1 end of file / short read, 2 malformed
int read_uleb128(ByteReader& r, uint32_t* out, std::string* why) {
  uint32_t result = 0;
  int shift = 0;
  for (int i = 1;; ++i) {
    uint8_t b;
    if (!r.get(&b)) {
      *why = "IO error: failed to fill whole buffer";
      return 1;
    }
    if (i > 5) {
      *why = "ULEB128 encoding too long";
      return 2;
    }
    const uint32_t bits = b & 0x7F;
    if (shift == 28 && bits > 0x0F) {
      *why = "ULEB128 value too large for u32";
      return 2;
    }
    result |= bits << shift;
    shift += 7;
    if (!(b & 0x80)) {
      *out = result;
      return 0;
    }
  }
}

// read buffer size by file size (src/data_storage/file_format.rs:80-99)
size_t read_buffer_size(uint64_t file_size) {
  constexpr uint64_t kMB = 1024 * 1024;
  if (file_size < kMB) return 64 * 1024;
  if (file_size < 100 * kMB) return (size_t)std::min<uint64_t>(std::max<uint64_t>(file_size / 8, 64 * 1024), 4 * kMB);
  return 16 * kMB;
}

}  // namespace

void set_error(const std::string& msg) { g_err = msg; }
const std::string& last_error() { return g_err; }

std::string commas(uint64_t n) {
  const std::string s = std::to_string(n);
  std::string r;
  for (size_t i = 0; i < s.size(); ++i) {
    if (i > 0 && (s.size() - i) % 3 == 0) r.push_back(',');
    r.push_back(s[i]);
  }
  return r;
}

std::string format_duration(double total) {
  const uint64_t days = (uint64_t)(total / 86400.0);
  const uint64_t hours = (uint64_t)(fmod(total, 86400.0) / 3600.0);
  const uint64_t minutes = (uint64_t)(fmod(total, 3600.0) / 60.0);
  const double seconds = fmod(total, 60.0);
  auto pl = [](uint64_t v) { return v == 1 ? "" : "s"; };
  const char* ps = seconds == 1.0 ? "" : "s";
  if (days > 0)
    return fmt("%llu day%s %llu hour%s %llu minute%s %.2f second%s", (unsigned long long)days, pl(days),
               (unsigned long long)hours, pl(hours), (unsigned long long)minutes, pl(minutes), seconds, ps);
  if (hours > 0)
    return fmt("%llu hour%s %llu minute%s %.2f second%s", (unsigned long long)hours, pl(hours),
               (unsigned long long)minutes, pl(minutes), seconds, ps);
  if (minutes > 0) return fmt("%llu minute%s %.2f second%s", (unsigned long long)minutes, pl(minutes), seconds, ps);
  return fmt("%.2f second%s", seconds, ps);
}

std::string format_remaining_time(double elapsed_s, uint64_t completed, uint64_t total) {
  if (completed == 0 || completed >= total) return "unknown";
  const double ratio = (double)completed / (double)total;
  const double remaining = elapsed_s / ratio - elapsed_s;
  if (remaining <= 0.0) return "completing soon";
  return "estimated remaining: " + format_duration(remaining);
}

void print_percentiles_table(const std::string& model_name, const std::string& statistic, const double* p,
                             const double* v, size_t n, uint64_t total_count) {
  printf("%s for model %s:\n", statistic.c_str(), model_name.c_str());
  printf("Total calculated %s values\n", commas(total_count).c_str());
  size_t pw = 0, vw = 0;
  for (size_t i = 0; i < n; ++i) {
    pw = std::max(pw, fmt("%.1fth", p[i] * 100.0).size());
    vw = std::max(vw, fmt("%.6f", v[i]).size());
  }
  if (n == 0) pw = 10, vw = 12;
  pw = std::max(pw, strlen("Percentile"));
  vw = std::max(vw, strlen("Value"));
  printf("%-*s %*s\n", (int)pw, "Percentile", (int)vw, "Value");
  printf("%s\n", std::string(pw + vw + 1, '-').c_str());
  for (size_t i = 0; i < n; ++i) printf("%-*s %*.6f\n", (int)pw, fmt("%.1fth", p[i] * 100.0).c_str(), (int)vw, v[i]);
}

const char* model_description(int model) {
  static const char* const kDesc[5] = {
      "No intercept, no trend",
      "Intercept, no trend, intercept in cointegration",
      "Intercept, no trend, intercept not fully explained by cointegration",
      "Intercept, trend, trend in cointegration",
      "Intercept, trend, intercept and trend not fully explained by cointegration",
  };
  return (model >= 0 && model < 5) ? kDesc[model] : "";
}

std::string model_display(int model) { return fmt("Model %d: %s", model, model_description(model)); }

std::string dat_filename(const std::string& data_dir, int model, int dim, uint64_t steps, bool create_dir) {
  if (create_dir && !data_dir.empty()) mkdir(data_dir.c_str(), 0777);
  std::string name = fmt("eigenvalues_model%d_dim%d_steps%llu.dat", model, dim, (unsigned long long)steps);
  if (data_dir.empty()) return name;
  return data_dir.back() == '/' ? data_dir + name : data_dir + "/" + name;
}

double percentile_sorted(const double* s, size_t n, double p) {
  if (n == 0) return NAN;
  const double rank = p * (double)(n - 1);
  const size_t lo = (size_t)floor(rank), hi = (size_t)ceil(rank);
  if (lo == hi) return s[lo];
  const double w = rank - (double)lo;
  return s[lo] * (1.0 - w) + s[hi] * w;
}

int read_dat(const char* path, bool want_data, DatContents* out, bool force_scan) {
  *out = DatContents();
  FILE* f = fopen(path, "rb");
  if (!f) return fail(JNE_ERR_IO, os_error(errno));
  struct stat st;
  if (fstat(fileno(f), &st) != 0) {
    const int e = errno;
    fclose(f);
    return fail(JNE_ERR_IO, os_error(e));
  }
  const uint64_t len = (uint64_t)st.st_size;
  ByteReader r(f, read_buffer_size(len));
  auto done = [&](int code, const std::string& msg) {
    fclose(f);
    return code ? fail(code, msg) : 0;
  };
  char magic[12];
  if (!r.read(magic, 12)) return done(JNE_ERR_IO, "failed to fill whole buffer");
  if (memcmp(magic, kMagic, 12) != 0) return done(JNE_ERR_FORMAT, "File format error: magic header mismatch");
  uint8_t hdr[6];
  if (!r.read(hdr, 6)) return done(JNE_ERR_IO, "failed to fill whole buffer");
  out->model = hdr[0];
  out->dim = hdr[1];
  out->steps = (uint32_t)hdr[2] | ((uint32_t)hdr[3] << 8) | ((uint32_t)hdr[4] << 16) | ((uint32_t)hdr[5] << 24);
  out->data_end = kHeaderBytes;
  if (len < kHeaderBytes + kTrailerBytes) return done(0, "");  // too small for records + trailer

  // trailer: EOF_MARK, u64 count, u8 nF at the end (reader.rs:80-104)
  uint8_t tail[17];
  r.seek(len - kTrailerBytes);
  if (!r.read(tail, 17)) return done(JNE_ERR_IO, "failed to fill whole buffer");
  r.seek(kHeaderBytes);
  std::string why;
  if (!force_scan && memcmp(tail, kEofMark, 8) == 0) {
    uint64_t count = 0;
    for (int i = 0; i < 8; ++i) count |= (uint64_t)tail[8 + i] << (8 * i);
    const int nf = tail[16];
    out->has_trailer = true;
    out->nf = nf;
    out->seeds.reserve((size_t)std::min<uint64_t>(count, len / 10));
    if (want_data) out->eigs.reserve(out->seeds.capacity() * (size_t)nf);
    for (uint64_t k = 0; k < count; ++k) {  // reader.rs:107-150
      uint32_t seed;
      if (read_uleb128(r, &seed, &why) != 0) return done(JNE_ERR_FORMAT, why);
      uint8_t c;
      if (!r.get(&c)) return done(JNE_ERR_IO, "failed to fill whole buffer");
      if (c == 0) return done(JNE_ERR_FORMAT, "Invalid eigenvalue count: cannot be zero");
      if (c != nf)
        return done(JNE_ERR_FORMAT, fmt("Eigenvalue count mismatch: expected %s, actual %s", commas(nf).c_str(),
                                        commas(c).c_str()));
      out->seeds.push_back(seed);
      if (want_data) {
        const size_t at = out->eigs.size();
        out->eigs.resize(at + c);
        if (!r.read(&out->eigs[at], 8 * (size_t)c)) return done(JNE_ERR_IO, "failed to fill whole buffer");
      } else {
        double tmp[255];
        if (!r.read(tmp, 8 * (size_t)c)) return done(JNE_ERR_IO, "failed to fill whole buffer");
      }
    }
    out->data_end = r.offset();
    return done(0, "");
  }

  // no trailer: scan whole records until the data runs out (reader.rs:153-216).
  // A record whose eigenvalue count differs from the first record's ends the
  // scan as well (the writer never produces one).
  for (;;) {
    uint32_t seed;
    if (read_uleb128(r, &seed, &why) != 0) break;
    uint8_t c;
    if (!r.get(&c)) break;
    if (c == 0) {
      const uint64_t at = r.offset();
      char mark[8];
      if (r.read(mark, 8) && memcmp(mark, kEofMark, 8) == 0) break;
      r.seek(at);
    }
    if (!out->seeds.empty() && c != out->nf) break;
    double vals[255];
    if (!r.read(vals, 8 * (size_t)c)) break;
    if (out->seeds.empty()) out->nf = c;
    out->seeds.push_back(seed);
    if (want_data) out->eigs.insert(out->eigs.end(), vals, vals + c);
    out->data_end = r.offset();
  }
  return done(0, "");
}

}  // namespace jne_host

using namespace jne_host;

struct jne_dat_writer {
  FILE* f = nullptr;
  std::string path;
  uint8_t model = 0, dim = 0;
  uint32_t steps = 0;
  int nf = -1;          // eigenvalues per record, fixed by the first record
  uint64_t count = 0;   // records in the file
  uint64_t initial = 0; // records found when opened
  int quiet = 1;
  int threads = 1;
  // progress lines (src/data_storage/thread_manager.rs:62-83)
  uint64_t progress_total = 0;
  Clock::time_point t0;
  std::vector<std::vector<uint8_t>> bufs;
};

namespace {

int write_header(FILE* f, uint8_t model, uint8_t dim, uint32_t steps) {
  uint8_t h[18];
  memcpy(h, kMagic, 12);
  h[12] = model;
  h[13] = dim;
  for (int i = 0; i < 4; ++i) h[14 + i] = (uint8_t)(steps >> (8 * i));
  if (fwrite(h, 1, 18, f) != 18 || fflush(f) != 0) return fail(JNE_ERR_IO, os_error(errno));
  return 0;
}

// Encodes records [lo, hi) into out.
void encode_range(const uint32_t* seeds, const double* eigs, size_t lo, size_t hi, int nf, std::vector<uint8_t>* out) {
  const size_t rec_max = 5 + 1 + 8 * (size_t)nf;
  out->resize((hi - lo) * rec_max);
  uint8_t* p = out->data();
  for (size_t i = lo; i < hi; ++i) {
    p += uleb128_put(seeds[i], p);
    *p++ = (uint8_t)nf;
    memcpy(p, eigs + i * (size_t)nf, 8 * (size_t)nf);  // little-endian host
    p += 8 * (size_t)nf;
  }
  out->resize((size_t)(p - out->data()));
}

}  // namespace

extern "C" {

const char* jne_host_last_error(void) { return g_err.c_str(); }

int jne_dat_writer_open(const char* path, int model, int dim, uint32_t steps, uint64_t expected_records, int quiet,
                        jne_dat_writer** out) {
  if (!path || !out) return fail(JNE_ERR_INVALID_ARG, "jne_dat_writer_open: path/out must not be NULL");
  *out = nullptr;
  if (model < 0 || model > 255 || dim < 0) return fail(JNE_ERR_INVALID_ARG, "jne_dat_writer_open: invalid model/dim");
  const uint8_t m8 = (uint8_t)model, d8 = (uint8_t)dim;  // dim is stored as u8 (thread_manager.rs:52)
  jne_dat_writer* w = new jne_dat_writer();
  w->path = path;
  w->model = m8;
  w->dim = d8;
  w->steps = steps;
  w->quiet = quiet;
  const unsigned hc = std::thread::hardware_concurrency();
  w->threads = (int)std::max(1u, std::min(hc ? hc : 1u, 8u));
  w->t0 = Clock::now();

  struct stat st;
  const bool exists = stat(path, &st) == 0;
  auto fresh = [&]() -> int {
    w->f = fopen(path, "w+b");
    if (!w->f) return fail(JNE_ERR_IO, os_error(errno));
    setvbuf(w->f, nullptr, _IOFBF, 2 << 20);  // WRITE_BUFFER_CAPACITY (config.rs:9)
    return write_header(w->f, m8, d8, steps);
  };
  int rc = 0;
  if (!exists) {
    if (expected_records && !quiet)
      printf("File pre-allocation disabled to ensure resume functionality works correctly\n");
    rc = fresh();
  } else {
    DatContents c;
    rc = read_dat(path, false, &c);
    uint64_t keep = 0;  // bytes to keep: the header and every whole record
    if (rc == 0) {
      if (c.model != m8)
        rc = fail(JNE_ERR_FORMAT, fmt("Model mismatch: file has model %u, expected %u", c.model, m8));
      else if (c.dim != d8)
        rc = fail(JNE_ERR_FORMAT, fmt("Dimension mismatch: file has dim %u, expected %u", c.dim, d8));
      else if (c.steps != steps)
        rc = fail(JNE_ERR_FORMAT, fmt("Steps mismatch: file has steps %u, expected %u", c.steps, steps));
      if (rc == 0) {
        w->count = c.seeds.size();
        if (!c.seeds.empty()) w->nf = c.nf;
        keep = c.data_end;
        if (!quiet) printf("Detected existing file with %s data records\n", commas(w->count).c_str());
      }
    } else if (last_error() == "File format error: magic header mismatch") {
      if (!quiet) printf("WARNING: Incompatible file format detected, recreating file...\n");
      if (unlink(path) != 0) rc = fail(JNE_ERR_IO, os_error(errno));
      else rc = fresh();
      if (rc) {
        if (w->f) fclose(w->f);
        delete w;
        return rc;
      }
      *out = w;
      if (!quiet) fflush(stdout);
      return 0;
    } else {
      if (!quiet) printf("WARNING: Could not read existing file, will attempt to append...\n");
      // keep the whole records a scan recovers (a trailer that disagrees with
      // the data is not trusted); a file too damaged to scan is rewritten
      DatContents sc;
      if (read_dat(path, false, &sc, true) == 0 && sc.model == m8 && sc.dim == d8 && sc.steps == steps) {
        rc = 0;
        w->count = sc.seeds.size();
        if (!sc.seeds.empty()) w->nf = sc.nf;
        keep = sc.data_end;
      } else {
        rc = fresh();
        if (rc) {
          if (w->f) fclose(w->f);
          delete w;
          return rc;
        }
        *out = w;
        if (!quiet) fflush(stdout);
        return 0;
      }
    }
    if (rc) {
      delete w;
      return rc;
    }
    // Drop the trailer (writer.rs:170-195) and any incomplete trailing
    // record: appending after a partial record would corrupt every record
    // that follows it.
    FILE* f = fopen(path, "r+b");
    if (!f) {
      delete w;
      return fail(JNE_ERR_IO, os_error(errno));
    }
    setvbuf(f, nullptr, _IOFBF, 2 << 20);  // before any other operation on the stream
    fseeko(f, 0, SEEK_END);
    const uint64_t flen = (uint64_t)ftello(f);
    if (flen > keep) {
      bool trailer_only = false;
      if (flen - keep == kTrailerBytes) {
        char mark[8];
        fseeko(f, (off_t)keep, SEEK_SET);
        trailer_only = fread(mark, 1, 8, f) == 8 && memcmp(mark, kEofMark, 8) == 0;
      }
      if (ftruncate(fileno(f), (off_t)keep) != 0) {
        const int e = errno;
        fclose(f);
        delete w;
        return fail(JNE_ERR_IO, os_error(e));
      }
      if (!quiet) {
        if (trailer_only)
          printf("Removed EOF marker to enable append mode\n");
        else
          printf("Truncated %s bytes after the last whole record to enable append mode\n",
                 commas(flen - keep).c_str());
      }
    }
    fseeko(f, 0, SEEK_END);
    w->f = f;
  }
  if (rc) {
    if (w->f) fclose(w->f);
    delete w;
    return rc;
  }
  w->initial = w->count;
  *out = w;
  if (!quiet) fflush(stdout);
  return 0;
}

uint64_t jne_dat_writer_count(const jne_dat_writer* w) { return w ? w->count : 0; }

int jne_dat_writer_set_progress(jne_dat_writer* w, uint64_t total_runs) {
  if (!w) return JNE_ERR_INVALID_ARG;
  w->progress_total = total_runs;
  w->t0 = Clock::now();
  return 0;
}

int jne_dat_writer_append(jne_dat_writer* w, const uint32_t* seeds, const double* eigs, size_t n, int nf) {
  if (!w || (n && (!seeds || !eigs))) return fail(JNE_ERR_INVALID_ARG, "jne_dat_writer_append: NULL argument");
  if (n == 0) return 0;
  if (nf < 1) return fail(JNE_ERR_FORMAT, "Invalid eigenvalue count: cannot be zero");
  if (nf > 255)
    return fail(JNE_ERR_FORMAT, fmt("Too many eigenvalues: %s exceeds maximum of 255", commas((uint64_t)nf).c_str()));
  if (w->nf < 0) w->nf = nf;
  if (nf != w->nf)
    return fail(JNE_ERR_FORMAT, fmt("Eigenvalue count mismatch: expected %s, actual %s (model %u, dim %u, steps %u)",
                                    commas(w->nf).c_str(), commas(nf).c_str(), w->model, w->dim, w->steps));
  // encode in parallel slices, write in order
  const int k = (int)std::min<size_t>((size_t)w->threads, std::max<size_t>(1, n / 16384));
  if ((int)w->bufs.size() < k) w->bufs.resize(k);
  if (k == 1) {
    encode_range(seeds, eigs, 0, n, nf, &w->bufs[0]);
  } else {
    std::vector<std::thread> th;
    for (int t = 0; t < k; ++t) {
      const size_t lo = n * t / k, hi = n * (t + 1) / k;
      th.emplace_back(encode_range, seeds, eigs, lo, hi, nf, &w->bufs[t]);
    }
    for (auto& x : th) x.join();
  }
  for (int t = 0; t < k; ++t)
    if (fwrite(w->bufs[t].data(), 1, w->bufs[t].size(), w->f) != w->bufs[t].size())
      return fail(JNE_ERR_IO, os_error(errno));
  const uint64_t before = w->count;
  w->count += n;
  if (w->progress_total && !w->quiet) {
    const uint64_t total = w->progress_total;
    const double elapsed = std::chrono::duration<double>(Clock::now() - w->t0).count();
    for (uint64_t c = (before / kProgressInterval + 1) * kProgressInterval; c <= w->count; c += kProgressInterval) {
      const double ratio = (double)c / (double)total;
      const uint64_t done_here = c - w->initial;
      printf("Simulation progress: %s/%s (%.2f%%) - %s\n", commas(c).c_str(), commas(total).c_str(), ratio * 100.0,
             format_remaining_time(elapsed, done_here, total - w->initial).c_str());
    }
  }
  if (w->count / kProgressInterval != before / kProgressInterval) {
    if (fflush(w->f) != 0) return fail(JNE_ERR_IO, os_error(errno));
    if (w->progress_total && !w->quiet) fflush(stdout);
  }
  return 0;
}

int jne_dat_writer_finish(jne_dat_writer* w) {
  if (!w) return fail(JNE_ERR_INVALID_ARG, "jne_dat_writer_finish: NULL writer");
  uint8_t t[17];
  memcpy(t, kEofMark, 8);
  for (int i = 0; i < 8; ++i) t[8 + i] = (uint8_t)(w->count >> (8 * i));
  t[16] = (uint8_t)(w->nf < 0 ? 0 : w->nf);
  int rc = 0;
  if (fwrite(t, 1, 17, w->f) != 17 || fflush(w->f) != 0) rc = fail(JNE_ERR_IO, os_error(errno));
  if (fclose(w->f) != 0 && rc == 0) rc = fail(JNE_ERR_IO, os_error(errno));
  if (rc == 0 && !w->quiet)
    printf("SUCCESS: append write completed, wrote %s data records for model %u, dim %u, steps %u\n",
           commas(w->count).c_str(), w->model, w->dim, w->steps);
  if (!w->quiet) fflush(stdout);
  delete w;
  return rc;
}

void jne_dat_writer_abandon(jne_dat_writer* w) {
  if (!w) return;
  fflush(w->f);
  fclose(w->f);
  delete w;
}

int jne_dat_read_info(const char* path, int* model, int* dim, uint32_t* steps, uint64_t* records, int* nf,
                      int* has_trailer) {
  if (!path) return fail(JNE_ERR_INVALID_ARG, "jne_dat_read_info: path is NULL");
  DatContents c;
  const int rc = read_dat(path, false, &c);
  if (rc) return rc;
  if (model) *model = c.model;
  if (dim) *dim = c.dim;
  if (steps) *steps = c.steps;
  if (records) *records = c.seeds.size();
  if (nf) *nf = c.nf;
  if (has_trailer) *has_trailer = c.has_trailer ? 1 : 0;
  return 0;
}

int jne_dat_read(const char* path, uint32_t* seeds, double* eigs, uint64_t capacity, uint64_t* n_read, int* nf) {
  if (!path || !n_read) return fail(JNE_ERR_INVALID_ARG, "jne_dat_read: NULL argument");
  DatContents c;
  const int rc = read_dat(path, eigs != nullptr, &c);
  if (rc) return rc;
  const uint64_t n = c.seeds.size();
  if (nf) *nf = c.nf;
  *n_read = n;
  if (n > capacity) return fail(JNE_ERR_INVALID_ARG, fmt("jne_dat_read: %llu records exceed capacity %llu",
                                                          (unsigned long long)n, (unsigned long long)capacity));
  if (seeds && n) memcpy(seeds, c.seeds.data(), n * sizeof(uint32_t));
  if (eigs && !c.eigs.empty()) memcpy(eigs, c.eigs.data(), c.eigs.size() * sizeof(double));
  return 0;
}

int jne_dat_remaining_seeds(const char* path, int model, int dim, uint32_t steps, uint64_t runs, uint32_t* out,
                            uint64_t capacity, uint64_t* n_out, uint64_t* completed) {
  if (!path || !n_out) return fail(JNE_ERR_INVALID_ARG, "jne_dat_remaining_seeds: NULL argument");
  if (runs > 0xFFFFFFFFull) return fail(JNE_ERR_INVALID_ARG, "runs exceeds the u32 seed range");
  std::vector<uint8_t> done((size_t)runs + 1, 0);
  uint64_t ncomp = 0;
  struct stat st;
  if (stat(path, &st) == 0) {  // progress.rs:10-54
    DatContents c;
    if (read_dat(path, false, &c) == 0) {
      if (c.model != (uint8_t)model)
        return fail(JNE_ERR_FORMAT, fmt("Model mismatch: file has model %u, expected %u", c.model, (uint8_t)model));
      if (c.dim != (uint8_t)dim)
        return fail(JNE_ERR_FORMAT, fmt("Dimension mismatch: file has dim %u, expected %u", c.dim, (uint8_t)dim));
      if (c.steps != steps)
        return fail(JNE_ERR_FORMAT, fmt("Steps mismatch: file has steps %u, expected %u", c.steps, steps));
      ncomp = c.seeds.size();
      for (uint32_t s : c.seeds)
        if (s >= 1 && s <= runs) done[s] = 1;
    }  // an unreadable file counts as empty (progress.rs:52)
  }
  uint64_t k = 0;
  for (uint64_t s = 1; s <= runs; ++s) {  // progress.rs:56-61
    if (done[s]) continue;
    if (out && k < capacity) out[k] = (uint32_t)s;
    ++k;
  }
  *n_out = k;
  if (completed) *completed = ncomp;
  if (out && k > capacity) return fail(JNE_ERR_INVALID_ARG, "jne_dat_remaining_seeds: capacity too small");
  return 0;
}

int jne_percentiles(const double* values, size_t n, const double* p, size_t np, double* out) {
  if ((n && !values) || (np && (!p || !out))) return fail(JNE_ERR_INVALID_ARG, "jne_percentiles: NULL argument");
  std::vector<double> v(values, values + n);
  for (double x : v)
    if (isnan(x)) return fail(JNE_ERR_NUMERIC, "jne_percentiles: NaN value");
  std::sort(v.begin(), v.end());
  for (size_t i = 0; i < np; ++i) out[i] = percentile_sorted(v.data(), n, p[i]);
  return 0;
}

int jne_dat_statistics(const char* path, uint64_t runs, int statistic, double* out, uint64_t capacity,
                       uint64_t* n_out) {
  if (!path || !n_out) return fail(JNE_ERR_INVALID_ARG, "jne_dat_statistics: NULL argument");
  if (statistic != JNE_STAT_TRACE && statistic != JNE_STAT_MAXEIG)
    return fail(JNE_ERR_INVALID_ARG, fmt("invalid statistic %d", statistic));
  DatContents c;
  const int rc = read_dat(path, true, &c);
  if (rc) return rc;
  uint64_t k = 0;
  for (size_t i = 0; i < c.seeds.size(); ++i) {
    if (runs && c.seeds[i] > runs) continue;  // read_data keeps seed <= num_runs (simulation.rs:62-67)
    // a scanned file may hold zero-count records (reader.rs:180-191): then the
    // sum is 0 and the maximum stays at the fold's start
    const double* e = c.eigs.data() + i * (size_t)c.nf;
    double v;
    if (statistic == JNE_STAT_TRACE) {  // simulation_analyzers.rs:26-31
      v = 0.0;
      for (int j = 0; j < c.nf; ++j) v += e[j];
    } else {  // simulation_analyzers.rs:34-39
      v = -1.7976931348623157e308;
      for (int j = 0; j < c.nf; ++j) v = std::max(v, e[j]);
    }
    if (out && k < capacity) out[k] = v;
    ++k;
  }
  *n_out = k;
  if (out && k > capacity) return fail(JNE_ERR_INVALID_ARG, "jne_dat_statistics: capacity too small");
  return 0;
}

}  // extern "C"
