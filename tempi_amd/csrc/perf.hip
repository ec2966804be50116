// perf.hip -- measured system performance and the send-method model.
//
// The reference keeps a table-driven model of the machine
// (include/measure_system.hpp:26-106): contiguous transfer curves indexed
// by log2 bytes, and 9x9 pack/unpack tables indexed by (2^(2i+6) bytes,
// blocks of 2^j bytes) measured at a 512-B stride, for packing into device
// memory and straight into mapped host memory.  A time for any (bytes,
// block) comes from linear interpolation between the measured points
// (src/internal/measure_system.cpp:182-297), and a non-contiguous send is
// priced as pack + send + unpack for each method
// (src/internal/sender.cpp:128-137, 180-191, 238-248); the cheaper of
// DEVICE and ONE_SHOT is taken (sender.cpp:251-273).  The tables live in
// perf.json in the cache directory, with the reference's keys, and are
// loaded at MPI_Init.
//
// Here the tables are measured with this library's own kernels on this GPU
// (one process: the ping-pong curves need a second rank, so they are only
// imported or set).  The interpolation follows the reference's piecewise
// definition, with two places where the reference's is undefined given a
// meaning: past the last row, the column clamp uses the last row's width
// (the reference reads the row after the table), and sizes below 32 B are
// priced as 32 B (the reference's row index underflows there).  Arithmetic
// is double throughout (the reference mixes in float).
#include "internal.hpp"

#include <hip/hip_runtime.h>

#include <sys/stat.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <limits>
#include <atomic>
#include <map>
#include <memory>
#include <sstream>
#include <tuple>

namespace tempi {

namespace {

constexpr double kInf = std::numeric_limits<double>::infinity();

struct IidTime {
  double time = 0;
  bool iid = false;
};

struct Table {
  int rows = 0;
  int cols = 0;
  std::vector<IidTime> v; // row-major
};

struct Performance {
  double kernel_launch = 0;
  std::vector<IidTime> curves[TEMPI_NUM_CURVES];
  Table tables[TEMPI_NUM_TABLES];
};

const char *const kCurveKeys[TEMPI_NUM_CURVES] = {
    "d2h",
    "h2d",
    "intraNodeCpuCpuPingpong",
    "intraNodeGpuGpuPingpong",
    "interNodeCpuCpuPingpong",
    "interNodeGpuGpuPingpong"};
const char *const kTableKeys[TEMPI_NUM_TABLES] = {"packDevice", "unpackDevice",
                                                  "packHost", "unpackHost"};
const char *const kLaunchKey = "cudaKernelLaunch";

constexpr int kMaxCurve = 62;
constexpr int kMaxRows = 16;
constexpr int kMaxCols = 32;

std::mutex g_perf_mu;
Performance g_perf;
std::map<std::tuple<int, int64_t, int64_t>, int> g_method_cache;
// bumped whenever the tables change, so that choices cached elsewhere
// (the point-to-point method cache) are dropped
std::atomic<uint64_t> g_perf_generation{0};

int log2_floor(int64_t x) { return x <= 0 ? 0 : 63 - __builtin_clzll(x); }
int log2_ceil(int64_t x) {
  return log2_floor(x) + ((x > 0 && (x & (x - 1))) ? 1 : 0);
}

// ---------------------------------------------------------------------------
// interpolation (reference: src/internal/measure_system.cpp:182-297)
// ---------------------------------------------------------------------------
template <class Get>
double interp_curve(Get at, int n, int64_t bytes) {
  if (n <= 0 || bytes < 0) {
    return kInf;
  }
  const int lb = log2_floor(bytes);
  const int ub = log2_ceil(bytes);
  if (ub >= n) { // past the last point: scale the largest time
    return at(n - 1) * (double)bytes / std::ldexp(1.0, n - 1);
  }
  if (lb == ub) {
    return at(lb);
  }
  const double sf =
      (double)(bytes - (1ll << lb)) / (double)((1ll << ub) - (1ll << lb));
  return at(lb) * (1 - sf) + at(ub) * sf;
}

template <class Get>
double interp_table(Get at, int rows, int cols, int64_t bytes, int64_t x) {
  if (rows <= 0 || cols <= 0 || bytes < 0 || x < 0) {
    return kInf;
  }
  bytes = std::max<int64_t>(bytes, 32);
  // rows: y1 = 2^(2 yi1 + 6) <= bytes, the next row above unless exact;
  // between 32 and 63 B this is row 0 with a negative weight, as in the
  // reference
  const int yi1 = std::max(0, (log2_floor(bytes) - 6) / 2);
  const double y1 = std::ldexp(1.0, 2 * yi1 + 6);
  const int yi2 = (y1 == (double)bytes) ? yi1 : yi1 + 1;
  const double y2 = std::ldexp(1.0, 2 * yi2 + 6);
  // columns: powers of two around x, clamped to the table's width
  int xi1 = log2_floor(x);
  int xi2 = log2_ceil(x);
  const double x1 = std::ldexp(1.0, xi1);
  const double x2 = std::ldexp(1.0, xi2);
  if (xi2 >= cols) {
    xi2 = cols - 1;
  }
  xi1 = std::min(xi1, xi2);
  const double sfx = (xi2 == xi1) ? 0.0 : ((double)x - x1) / (x2 - x1);
  if (yi2 >= rows) { // past the last row: scale the last row's value
    const int r = rows - 1;
    const double base = (1 - sfx) * at(r, xi1) + sfx * at(r, xi2);
    return base / std::ldexp(1.0, r * 2 + 6) * (double)bytes;
  }
  const double sfy = (yi2 == yi1) ? 0.0 : ((double)bytes - y1) / (y2 - y1);
  const double f1 = (1 - sfx) * at(yi1, xi1) + sfx * at(yi1, xi2);
  const double f2 = (1 - sfx) * at(yi2, xi1) + sfx * at(yi2, xi2);
  return (1 - sfy) * f1 + sfy * f2;
}

double curve_time(const std::vector<IidTime> &c, int64_t bytes) {
  return interp_curve([&](int i) { return c[i].time; }, (int)c.size(), bytes);
}

double table_time(const Table &t, int64_t bytes, int64_t bl) {
  return interp_table(
      [&](int i, int j) { return t.v[(size_t)i * t.cols + j].time; }, t.rows,
      t.cols, bytes, bl);
}

// modelled seconds of one method; +inf when an input is missing
double model_locked(int method, bool colocated, int64_t bytes, int64_t bl) {
  const Performance &p = g_perf;
  const auto &cpu = p.curves[colocated ? TEMPI_CURVE_INTRA_CPU_CPU
                                       : TEMPI_CURVE_INTER_CPU_CPU];
  const auto &gpu = p.curves[colocated ? TEMPI_CURVE_INTRA_GPU_GPU
                                       : TEMPI_CURVE_INTER_GPU_GPU];
  switch (method) {
  case TEMPI_METHOD_DEVICE: // sender.cpp:180-191
    return table_time(p.tables[TEMPI_TABLE_PACK_DEVICE], bytes, bl) +
           curve_time(gpu, bytes) +
           table_time(p.tables[TEMPI_TABLE_UNPACK_DEVICE], bytes, bl);
  case TEMPI_METHOD_ONESHOT: // sender.cpp:128-137
    return table_time(p.tables[TEMPI_TABLE_PACK_HOST], bytes, bl) +
           curve_time(cpu, bytes) +
           table_time(p.tables[TEMPI_TABLE_UNPACK_HOST], bytes, bl);
  case TEMPI_METHOD_STAGED: // sender.cpp:238-248
    return table_time(p.tables[TEMPI_TABLE_PACK_DEVICE], bytes, bl) +
           curve_time(p.curves[TEMPI_CURVE_D2H], bytes) +
           curve_time(cpu, bytes) +
           curve_time(p.curves[TEMPI_CURVE_H2D], bytes) +
           table_time(p.tables[TEMPI_TABLE_UNPACK_DEVICE], bytes, bl);
  default:
    return kInf;
  }
}

// ---------------------------------------------------------------------------
// perf.json: a small JSON reader and writer for the reference's schema
// ---------------------------------------------------------------------------
struct JVal {
  enum Kind { NUL, BOOL, NUM, STR, ARR, OBJ } kind = NUL;
  bool b = false;
  double num = 0;
  std::string str;
  std::vector<JVal> arr;
  std::map<std::string, JVal> obj;
};

struct JParser {
  const char *p;
  const char *end;
  std::string err;

  void ws() {
    while (p < end && (*p == ' ' || *p == '\n' || *p == '\r' || *p == '\t')) {
      ++p;
    }
  }
  bool fail(const char *m) {
    if (err.empty()) {
      err = m;
    }
    return false;
  }
  bool lit(const char *s) {
    const size_t n = strlen(s);
    if ((size_t)(end - p) < n || strncmp(p, s, n) != 0) {
      return fail("bad literal");
    }
    p += n;
    return true;
  }
  bool string(std::string &out) {
    if (p >= end || *p != '"') {
      return fail("expected string");
    }
    ++p;
    while (p < end && *p != '"') {
      if (*p == '\\') {
        ++p;
        if (p >= end) {
          return fail("bad escape");
        }
        const char c = *p;
        out += c == 'n' ? '\n' : c == 't' ? '\t' : c == 'r' ? '\r' : c;
        if (c == 'u') { // keys here are ASCII; keep \uXXXX verbatim
          out.pop_back();
          out += "\\u";
        }
      } else {
        out += *p;
      }
      ++p;
    }
    if (p >= end) {
      return fail("unterminated string");
    }
    ++p;
    return true;
  }
  bool value(JVal &v, int depth) {
    if (depth > 16) {
      return fail("nesting too deep");
    }
    ws();
    if (p >= end) {
      return fail("unexpected end");
    }
    if (*p == '{') {
      v.kind = JVal::OBJ;
      ++p;
      ws();
      if (p < end && *p == '}') {
        ++p;
        return true;
      }
      while (true) {
        ws();
        std::string k;
        if (!string(k)) {
          return false;
        }
        ws();
        if (p >= end || *p != ':') {
          return fail("expected ':'");
        }
        ++p;
        if (!value(v.obj[k], depth + 1)) {
          return false;
        }
        ws();
        if (p < end && *p == ',') {
          ++p;
          continue;
        }
        if (p < end && *p == '}') {
          ++p;
          return true;
        }
        return fail("expected ',' or '}'");
      }
    }
    if (*p == '[') {
      v.kind = JVal::ARR;
      ++p;
      ws();
      if (p < end && *p == ']') {
        ++p;
        return true;
      }
      while (true) {
        v.arr.emplace_back();
        if (!value(v.arr.back(), depth + 1)) {
          return false;
        }
        ws();
        if (p < end && *p == ',') {
          ++p;
          continue;
        }
        if (p < end && *p == ']') {
          ++p;
          return true;
        }
        return fail("expected ',' or ']'");
      }
    }
    if (*p == '"') {
      v.kind = JVal::STR;
      return string(v.str);
    }
    if (*p == 't') {
      v.kind = JVal::BOOL;
      v.b = true;
      return lit("true");
    }
    if (*p == 'f') {
      v.kind = JVal::BOOL;
      return lit("false");
    }
    if (*p == 'n') {
      v.kind = JVal::NUL;
      return lit("null");
    }
    std::string num;
    while (p < end && (isdigit((unsigned char)*p) || *p == '-' || *p == '+' ||
                       *p == '.' || *p == 'e' || *p == 'E')) {
      num += *p++;
    }
    if (num.empty()) {
      return fail("unexpected character");
    }
    v.kind = JVal::NUM;
    if (!parse_number(num.data(), num.data() + num.size(), &v.num)) {
      return fail("bad number");
    }
    return true;
  }
};

// {"time": t, "iid": b}; nlohmann writes a non-finite time as null
bool iid_from(const JVal &j, IidTime *out) {
  if (j.kind != JVal::OBJ) {
    return false;
  }
  auto t = j.obj.find("time");
  auto i = j.obj.find("iid");
  if (t == j.obj.end() || i == j.obj.end() || i->second.kind != JVal::BOOL) {
    return false;
  }
  if (t->second.kind == JVal::NUM) {
    out->time = t->second.num;
  } else if (t->second.kind == JVal::NUL) {
    out->time = kInf;
  } else {
    return false;
  }
  out->iid = i->second.b;
  return true;
}

bool curve_from(const JVal &j, std::vector<IidTime> *out) {
  if (j.kind != JVal::ARR || (int)j.arr.size() > kMaxCurve) {
    return false;
  }
  out->resize(j.arr.size());
  for (size_t k = 0; k < j.arr.size(); ++k) {
    if (!iid_from(j.arr[k], &(*out)[k])) {
      return false;
    }
  }
  return true;
}

bool table_from(const JVal &j, Table *out) {
  if (j.kind != JVal::ARR || (int)j.arr.size() > kMaxRows) {
    return false;
  }
  Table t;
  t.rows = (int)j.arr.size();
  for (const JVal &row : j.arr) {
    std::vector<IidTime> r;
    if (!curve_from(row, &r) || r.empty() || (int)r.size() > kMaxCols) {
      return false;
    }
    if (t.cols == 0) {
      t.cols = (int)r.size();
    } else if (t.cols != (int)r.size()) {
      return false; // ragged rows: the interpolation assumes one width
    }
    t.v.insert(t.v.end(), r.begin(), r.end());
  }
  *out = std::move(t);
  return true;
}

void put_num(std::ostream &os, double v) {
  if (std::isfinite(v)) {
    os << format_number(v);
  } else {
    os << "null";
  }
}

void put_iid(std::ostream &os, const IidTime &t) {
  os << "{\"iid\": " << (t.iid ? "true" : "false") << ", \"time\": ";
  put_num(os, t.time);
  os << "}";
}

void put_curve(std::ostream &os, const std::vector<IidTime> &c,
               const char *indent) {
  if (c.empty()) {
    os << "[]";
    return;
  }
  os << "[";
  for (size_t k = 0; k < c.size(); ++k) {
    os << (k ? ",\n" : "\n") << indent << "  ";
    put_iid(os, c[k]);
  }
  os << "\n" << indent << "]";
}

std::string to_json(const Performance &p) {
  // keys in sorted order, as the reference's writer emits them
  std::map<std::string, int> keys;
  for (int c = 0; c < TEMPI_NUM_CURVES; ++c) {
    keys[kCurveKeys[c]] = c;
  }
  for (int t = 0; t < TEMPI_NUM_TABLES; ++t) {
    keys[kTableKeys[t]] = 100 + t;
  }
  keys[kLaunchKey] = 200;
  std::ostringstream os;
  os << "{";
  bool first = true;
  for (const auto &kv : keys) {
    os << (first ? "\n" : ",\n") << "  \"" << kv.first << "\": ";
    first = false;
    if (kv.second == 200) {
      put_num(os, p.kernel_launch);
    } else if (kv.second >= 100) {
      const Table &t = p.tables[kv.second - 100];
      os << "[";
      for (int r = 0; r < t.rows; ++r) {
        std::vector<IidTime> row(t.v.begin() + (size_t)r * t.cols,
                                 t.v.begin() + (size_t)(r + 1) * t.cols);
        os << (r ? ",\n" : "\n") << "    ";
        put_curve(os, row, "    ");
      }
      os << (t.rows ? "\n  ]" : "]");
    } else {
      put_curve(os, p.curves[kv.second], "  ");
    }
  }
  os << "\n}\n";
  return os.str();
}

// all keys required, as the reference's from_json (measure_system.cpp:49-61)
bool from_json(const std::string &text, Performance *out, std::string *why) {
  JParser ps{text.data(), text.data() + text.size(), ""};
  JVal root;
  if (!ps.value(root, 0)) {
    *why = ps.err;
    return false;
  }
  ps.ws();
  if (ps.p != ps.end) {
    *why = "trailing characters";
    return false;
  }
  if (root.kind != JVal::OBJ) {
    *why = "not an object";
    return false;
  }
  Performance p;
  for (int c = 0; c < TEMPI_NUM_CURVES; ++c) {
    auto it = root.obj.find(kCurveKeys[c]);
    if (it == root.obj.end() || !curve_from(it->second, &p.curves[c])) {
      *why = std::string("missing or malformed ") + kCurveKeys[c];
      return false;
    }
  }
  for (int t = 0; t < TEMPI_NUM_TABLES; ++t) {
    auto it = root.obj.find(kTableKeys[t]);
    if (it == root.obj.end() || !table_from(it->second, &p.tables[t])) {
      *why = std::string("missing or malformed ") + kTableKeys[t];
      return false;
    }
  }
  auto it = root.obj.find(kLaunchKey);
  if (it == root.obj.end() || it->second.kind != JVal::NUM) {
    *why = std::string("missing or malformed ") + kLaunchKey;
    return false;
  }
  p.kernel_launch = it->second.num;
  *out = std::move(p);
  return true;
}

std::string default_path() {
  ensure_env();
  return env().cache_dir + "/perf.json";
}

// ---------------------------------------------------------------------------
// measurement (reference: src/internal/measure_system.cu:36-609 and the
// Benchmark loop, src/internal/benchmark.cpp:44-89)
// ---------------------------------------------------------------------------
typedef std::chrono::steady_clock Clock;

double seconds_since(Clock::time_point t0) {
  return std::chrono::duration<double>(Clock::now() - t0).count();
}

// statistics.cpp:30-39
double trimean(std::vector<double> x) {
  if (x.empty()) {
    return std::nan("");
  }
  std::sort(x.begin(), x.end());
  const size_t q = x.size() / 4;
  return (x[q] + 2 * x[2 * q] + x[3 * q]) / 4;
}

// How each point is sampled: trials of at least 7 and at most 500 samples
// or `max_s` seconds, repeated until the samples pass the SP 800-90B
// permutation test (iid.cpp) or `trials` trials have run (reference:
// Benchmark::run, src/internal/benchmark.cpp:44-83).
struct Sampling {
  double max_s;
  int trials;
};

// seeds of successive IID tests: a fixed sequence, so that a run's verdicts
// depend only on its samples
static std::atomic<uint64_t> g_iid_calls{0};

// One point: a warm-up sample, then trials as above.  The time is the last
// trial's trimean; `iid` records whether that trial passed the test.
template <class F> int sample_point(F &&iter, const Sampling &sp, IidTime *out) {
  double t = 0;
  int err = iter(&t);
  std::vector<double> xs;
  for (int trial = 1; !err; ++trial) {
    xs.clear();
    const auto t0 = Clock::now();
    while (!err && (xs.size() < 7 ||
                    (seconds_since(t0) < sp.max_s && xs.size() < 500))) {
      err = iter(&t);
      xs.push_back(t);
    }
    if (err) {
      break;
    }
    const uint64_t seed = 0x243F6A8885A308D3ull +
                          g_iid_calls.fetch_add(1) * 0x9E3779B97F4A7C15ull;
    int iid = 0;
    try {
      iid = iid_test(xs.data(), (int64_t)xs.size(), seed, nullptr);
    } catch (...) {
      iid = set_error(TEMPI_ERR_NOMEM, "IID test of %zu samples", xs.size());
    }
    if (iid < 0) {
      return iid;
    }
    if (iid == 1 || trial >= sp.trials) {
      out->time = trimean(xs);
      out->iid = iid == 1;
      return TEMPI_SUCCESS;
    }
  }
  return err;
}

__global__ void k_perf_empty(int *p) {
  if (p) {
    *p = threadIdx.x;
  }
}

struct DevBuf {
  void *p = nullptr;
  ~DevBuf() {
    if (p) {
      (void)hipFree(p);
    }
  }
};
struct HostBuf {
  void *p = nullptr;
  ~HostBuf() {
    if (p) {
      (void)hipHostFree(p);
    }
  }
};

int measure_launch(const Sampling &sp, double *out) {
  hipStream_t s = library_stream();
  IidTime r;
  int err = sample_point(
      [&](double *t) -> int {
        const auto t0 = Clock::now();
        for (int i = 0; i < 32; ++i) {
          hipLaunchKernelGGL(k_perf_empty, dim3(1), dim3(1), 0, s, nullptr);
        }
        *t = seconds_since(t0) / 32.0;
        TEMPI_HIP_CHECK(hipGetLastError());
        TEMPI_HIP_CHECK(hipStreamSynchronize(s));
        return TEMPI_SUCCESS;
      },
      sp, &r);
  if (!err) {
    *out = r.time;
  }
  return err;
}

// d2h / h2d of 1 B .. 8 MiB between device memory and pinned host memory,
// each sample repeated to last about 200 us (measure_system.cu:97-170)
int measure_copies(int which, const Sampling &sp, std::vector<IidTime> *out) {
  const size_t maxb = (size_t)1 << (TEMPI_PERF_CURVE_LEN - 1);
  DevBuf d;
  HostBuf h;
  TEMPI_HIP_CHECK(hipMalloc(&d.p, maxb));
  TEMPI_HIP_CHECK(hipHostMalloc(&h.p, maxb, hipHostMallocDefault));
  hipStream_t s = library_stream();
  std::vector<IidTime> c;
  for (int i = 0; i < TEMPI_PERF_CURVE_LEN; ++i) {
    const size_t n = (size_t)1 << i;
    void *dst = which == TEMPI_CURVE_D2H ? h.p : d.p;
    const void *src = which == TEMPI_CURVE_D2H ? d.p : h.p;
    const hipMemcpyKind kind = which == TEMPI_CURVE_D2H
                                   ? hipMemcpyDeviceToHost
                                   : hipMemcpyHostToDevice;
    int reps = 1;
    auto once = [&](double *t) -> int {
      const auto t0 = Clock::now();
      for (int r = 0; r < reps; ++r) {
        TEMPI_HIP_CHECK(hipMemcpyAsync(dst, src, n, kind, s));
        TEMPI_HIP_CHECK(hipStreamSynchronize(s));
      }
      *t = seconds_since(t0) / reps;
      return TEMPI_SUCCESS;
    };
    double t1 = 0;
    int err = once(&t1);
    if (!err) {
      err = once(&t1);
    }
    if (err) {
      return err;
    }
    reps = std::max(1, (int)(200e-6 / std::max(t1, 1e-9)));
    IidTime r;
    err = sample_point(once, sp, &r);
    if (err) {
      return err;
    }
    c.push_back(r);
  }
  *out = std::move(c);
  return TEMPI_SUCCESS;
}

// one-way CPU-CPU message time between world ranks 0 and 1, collective over
// MPI_COMM_WORLD (reference: bin/measure_system.cpp, the CPU-CPU ping-pong);
// rank 0 picks each size's repetition count so that every rank runs the same
// loop, and broadcasts the curve
int measure_pingpong(int *curve, std::vector<IidTime> *out) {
  int rank = 0, size = 1, up = 0, down = 0;
  MPI_Initialized(&up);
  MPI_Finalized(&down);
  if (!up || down) {
    return set_error(TEMPI_ERR_UNSUPPORTED, "ping-pong needs MPI initialized");
  }
  MPI_Comm_rank(MPI_COMM_WORLD, &rank);
  MPI_Comm_size(MPI_COMM_WORLD, &size);
  if (size < 2) {
    return set_error(TEMPI_ERR_UNSUPPORTED, "ping-pong needs two ranks");
  }
  *curve = world_ranks_share_node(0, 1) ? TEMPI_CURVE_INTRA_CPU_CPU
                                        : TEMPI_CURVE_INTER_CPU_CPU;
  const int n_max = 1 << (TEMPI_PERF_CURVE_LEN - 1);
  std::vector<char> buf((size_t)n_max, 1);
  const RealMpi &mpi = real_mpi();
  constexpr int kTag = 27717, kSamples = 7;
  auto bounce = [&](int n, int reps) {
    for (int r = 0; r < reps; ++r) {
      if (rank == 0) {
        mpi.Send(buf.data(), n, MPI_BYTE, 1, kTag, MPI_COMM_WORLD);
        mpi.Recv(buf.data(), n, MPI_BYTE, 1, kTag, MPI_COMM_WORLD,
                 MPI_STATUS_IGNORE);
      } else if (rank == 1) {
        mpi.Recv(buf.data(), n, MPI_BYTE, 0, kTag, MPI_COMM_WORLD,
                 MPI_STATUS_IGNORE);
        mpi.Send(buf.data(), n, MPI_BYTE, 0, kTag, MPI_COMM_WORLD);
      }
    }
  };
  std::vector<double> one_way(TEMPI_PERF_CURVE_LEN, 0.0);
  for (int i = 0; i < TEMPI_PERF_CURVE_LEN; ++i) {
    const int n = 1 << i;
    int reps = 1;
    const auto w = Clock::now();
    bounce(n, 1); // warm-up, and the round trip that sizes the samples
    if (rank == 0) {
      reps = std::max(1, std::min(10000, (int)(200e-6 / std::max(
                                              seconds_since(w), 1e-9))));
    }
    MPI_Bcast(&reps, 1, MPI_INT, 0, MPI_COMM_WORLD);
    std::vector<double> xs;
    for (int k = 0; k < kSamples; ++k) {
      const auto t0 = Clock::now();
      bounce(n, reps);
      xs.push_back(seconds_since(t0) / reps / 2);
    }
    one_way[i] = trimean(xs);
  }
  MPI_Bcast(one_way.data(), TEMPI_PERF_CURVE_LEN, MPI_DOUBLE, 0,
            MPI_COMM_WORLD);
  out->clear();
  for (double t : one_way) {
    out->push_back(IidTime{t, false});
  }
  return TEMPI_SUCCESS;
}

// a synchronous pack or unpack of 2^(2i+6) bytes in blocks of
// min(bytes, 2^j) bytes at a 512-B stride, as MPI_Pack runs it (the variant
// the library's selector picks, then a stream synchronization)
int measure_table(int table, const Sampling &sampling, Table *out) {
  TrialScope trial; // the table's copies are measurement, not the caller's
  const int64_t max_bytes = 1ll << (2 * (TEMPI_PERF_TABLE_DIM - 1) + 6);
  const int64_t stride = 512;
  const bool host = table == TEMPI_TABLE_PACK_HOST ||
                    table == TEMPI_TABLE_UNPACK_HOST;
  const int dir = (table == TEMPI_TABLE_PACK_DEVICE ||
                   table == TEMPI_TABLE_PACK_HOST)
                      ? TEMPI_DIR_PACK
                      : TEMPI_DIR_UNPACK;
  DevBuf strided, dpacked;
  TEMPI_HIP_CHECK(hipMalloc(&strided.p, (size_t)(max_bytes * stride)));
  void *packed = nullptr;
  if (host) {
    packed = tempi_alloc_host_mapped(max_bytes);
    if (!packed) {
      return TEMPI_ERR_NOMEM;
    }
  } else {
    TEMPI_HIP_CHECK(hipMalloc(&dpacked.p, (size_t)max_bytes));
    packed = dpacked.p;
  }
  std::unique_ptr<void, void (*)(void *)> host_guard(
      host ? packed : nullptr, [](void *p) {
        if (p) {
          tempi_free_host_mapped(p);
        }
      });
  void *packed_dev = nullptr;
  if (!device_address(packed, &packed_dev)) {
    return set_error(TEMPI_ERR_HIP, "no device address for the packed buffer");
  }
  hipStream_t s = library_stream();
  Table t;
  t.rows = t.cols = TEMPI_PERF_TABLE_DIM;
  for (int i = 0; i < TEMPI_PERF_TABLE_DIM; ++i) {
    for (int j = 0; j < TEMPI_PERF_TABLE_DIM; ++j) {
      const int64_t bytes = 1ll << (2 * i + 6);
      const int64_t bl = std::min<int64_t>(bytes, 1ll << j);
      const int64_t nb = bytes / bl;
      tempi_strided_block sb{};
      sb.start = 0;
      sb.extent = nb * stride;
      sb.ndims = 2;
      sb.counts[0] = bl;
      sb.counts[1] = nb;
      sb.strides[0] = 1;
      sb.strides[1] = stride;
      Geom g;
      int err = make_geom(sb, 1, &g);
      if (err) {
        return err;
      }
      const char *sp = static_cast<const char *>(strided.p);
      char *pp = static_cast<char *>(packed_dev);
      const Choice ch = choose(g, sp, pp, dir, env().select, s,
                               /*in_place=*/true);
      IidTime r;
      err = sample_point(
          [&](double *tm) -> int {
            const auto t0 = Clock::now();
            int e = launch_copy(g, sp, pp, dir, ch.variant, s, ch.max_blocks);
            if (e) {
              return e;
            }
            TEMPI_HIP_CHECK(hipStreamSynchronize(s));
            *tm = seconds_since(t0);
            return TEMPI_SUCCESS;
          },
          sampling, &r);
      if (err) {
        return err;
      }
      t.v.push_back(r);
    }
  }
  *out = std::move(t);
  return TEMPI_SUCCESS;
}

int check_curve_id(int c) {
  return (c < 0 || c >= TEMPI_NUM_CURVES)
             ? set_error(TEMPI_ERR_ARG, "curve id %d", c)
             : TEMPI_SUCCESS;
}
int check_table_id(int t) {
  return (t < 0 || t >= TEMPI_NUM_TABLES)
             ? set_error(TEMPI_ERR_ARG, "table id %d", t)
             : TEMPI_SUCCESS;
}

} // namespace

uint64_t perf_generation() { return g_perf_generation.load(); }

double perf_model_seconds(int method, bool colocated, int64_t bytes,
                          int64_t block_length) {
  std::lock_guard<std::mutex> lk(g_perf_mu);
  return model_locked(method, colocated, bytes, block_length);
}

// load <cache dir>/perf.json if present (reference: system::init at
// MPI_Init, measure_system.cu:32)
void perf_init() {
  std::ifstream f(default_path());
  if (!f) {
    TEMPI_LOG(LOG_INFO, "no %s: the method model has no tables",
              default_path().c_str());
    return;
  }
  (void)tempi_perf_load(nullptr);
}

} // namespace tempi

using namespace tempi;

TEMPI_EXPORT double tempi_interp_time(const double *times, int n,
                                      int64_t bytes) {
  if (!times || n <= 0) {
    return kInf;
  }
  n = std::min(n, kMaxCurve);
  return interp_curve([&](int i) { return times[i]; }, n, bytes);
}

TEMPI_EXPORT double tempi_interp_2d(const double *table, int rows, int cols,
                                    int64_t bytes, int64_t block_length) {
  if (!table || rows <= 0 || cols <= 0 || rows > kMaxRows || cols > kMaxCols) {
    return kInf;
  }
  return interp_table(
      [&](int i, int j) { return table[(size_t)i * cols + j]; }, rows, cols,
      bytes, block_length);
}

TEMPI_EXPORT int tempi_perf_clear(void) {
  std::lock_guard<std::mutex> lk(g_perf_mu);
  g_perf = Performance();
  g_method_cache.clear();
  ++g_perf_generation;
  return TEMPI_SUCCESS;
}

TEMPI_EXPORT int tempi_perf_get_curve(int curve, double *times, int cap,
                                      int *n) {
  int err = check_curve_id(curve);
  if (err) {
    return err;
  }
  std::lock_guard<std::mutex> lk(g_perf_mu);
  const auto &c = g_perf.curves[curve];
  if (n) {
    *n = (int)c.size();
  }
  for (int i = 0; times && i < cap && i < (int)c.size(); ++i) {
    times[i] = c[i].time;
  }
  return TEMPI_SUCCESS;
}

TEMPI_EXPORT int tempi_perf_set_curve(int curve, const double *times, int n) {
  int err = check_curve_id(curve);
  if (err) {
    return err;
  }
  if (n < 0 || n > kMaxCurve || (n > 0 && !times)) {
    return set_error(TEMPI_ERR_ARG, "curve length %d", n);
  }
  std::lock_guard<std::mutex> lk(g_perf_mu);
  auto &c = g_perf.curves[curve];
  c.assign((size_t)n, IidTime());
  for (int i = 0; i < n; ++i) {
    c[i].time = times[i];
  }
  g_method_cache.clear();
  ++g_perf_generation;
  return TEMPI_SUCCESS;
}

TEMPI_EXPORT int tempi_perf_get_table(int table, double *times, int cap,
                                      int *rows, int *cols) {
  int err = check_table_id(table);
  if (err) {
    return err;
  }
  std::lock_guard<std::mutex> lk(g_perf_mu);
  const Table &t = g_perf.tables[table];
  if (rows) {
    *rows = t.rows;
  }
  if (cols) {
    *cols = t.cols;
  }
  for (int i = 0; times && i < cap && i < (int)t.v.size(); ++i) {
    times[i] = t.v[i].time;
  }
  return TEMPI_SUCCESS;
}

TEMPI_EXPORT int tempi_perf_set_table(int table, const double *times, int rows,
                                      int cols) {
  int err = check_table_id(table);
  if (err) {
    return err;
  }
  const bool empty = rows == 0 || cols == 0;
  if (rows < 0 || cols < 0 || rows > kMaxRows || cols > kMaxCols ||
      (!empty && !times)) {
    return set_error(TEMPI_ERR_ARG, "table shape %d x %d", rows, cols);
  }
  std::lock_guard<std::mutex> lk(g_perf_mu);
  Table &t = g_perf.tables[table];
  t = Table();
  if (!empty) {
    t.rows = rows;
    t.cols = cols;
    t.v.resize((size_t)rows * cols);
    for (size_t i = 0; i < t.v.size(); ++i) {
      t.v[i].time = times[i];
    }
  }
  g_method_cache.clear();
  ++g_perf_generation;
  return TEMPI_SUCCESS;
}

TEMPI_EXPORT double tempi_perf_kernel_launch(void) {
  std::lock_guard<std::mutex> lk(g_perf_mu);
  return g_perf.kernel_launch;
}

TEMPI_EXPORT int tempi_perf_set_kernel_launch(double seconds) {
  std::lock_guard<std::mutex> lk(g_perf_mu);
  g_perf.kernel_launch = seconds;
  return TEMPI_SUCCESS;
}

TEMPI_EXPORT int tempi_perf_load(const char *path) {
  const std::string p = path ? std::string(path) : default_path();
  std::ifstream f(p);
  if (!f) {
    return set_error(TEMPI_ERR_IO, "cannot open %s", p.c_str());
  }
  std::stringstream ss;
  ss << f.rdbuf();
  Performance perf;
  std::string why;
  if (!from_json(ss.str(), &perf, &why)) {
    return set_error(TEMPI_ERR_IO, "%s: %s", p.c_str(), why.c_str());
  }
  std::lock_guard<std::mutex> lk(g_perf_mu);
  g_perf = std::move(perf);
  g_method_cache.clear();
  ++g_perf_generation;
  TEMPI_LOG(LOG_INFO, "loaded system performance from %s", p.c_str());
  return TEMPI_SUCCESS;
}

TEMPI_EXPORT int tempi_perf_save(const char *path) {
  std::string p;
  if (path) {
    p = path;
  } else {
    p = default_path();
    (void)mkdir(env().cache_dir.c_str(), 0755); // as create_directory
  }
  std::string text;
  {
    std::lock_guard<std::mutex> lk(g_perf_mu);
    text = to_json(g_perf);
  }
  std::ofstream f(p);
  if (!f) {
    return set_error(TEMPI_ERR_IO, "cannot write %s", p.c_str());
  }
  f << text;
  f.close();
  return f ? TEMPI_SUCCESS : set_error(TEMPI_ERR_IO, "write %s failed", p.c_str());
}

TEMPI_EXPORT int tempi_perf_measure(int what, double seconds_per_point) {
  if (what & ~0x1F) {
    return set_error(TEMPI_ERR_ARG, "measure mask 0x%x", what);
  }
  int err = TEMPI_SUCCESS;
  if ((what & 0xF) && (err = ensure_runtime())) {
    return err;
  }
  Sampling sp{seconds_per_point > 0 ? seconds_per_point : 0.05, 10};
  if (const char *e = std::getenv("TEMPI_MEASURE_TRIALS")) {
    int v = 0;
    if (!parse_number(e, e + std::strlen(e), &v) || v < 1) {
      return set_error(TEMPI_ERR_ARG, "TEMPI_MEASURE_TRIALS=%s", e);
    }
    sp.trials = v;
  }
  // measure outside the lock, fill only what is still empty
  Performance cur;
  {
    std::lock_guard<std::mutex> lk(g_perf_mu);
    cur = g_perf;
  }
  if (what & TEMPI_MEASURE_PINGPONG) {
    // collective: every rank measures, whatever its tables hold
    int c = 0;
    std::vector<IidTime> v;
    if ((err = measure_pingpong(&c, &v))) {
      return err;
    }
    std::lock_guard<std::mutex> lk(g_perf_mu);
    if (g_perf.curves[c].empty()) {
      g_perf.curves[c] = std::move(v);
    }
  }
  if ((what & 1) && cur.kernel_launch == 0) {
    double t = 0;
    if ((err = measure_launch(sp, &t))) {
      return err;
    }
    std::lock_guard<std::mutex> lk(g_perf_mu);
    g_perf.kernel_launch = t;
  }
  if (what & 2) {
    for (int c : {TEMPI_CURVE_D2H, TEMPI_CURVE_H2D}) {
      if (!cur.curves[c].empty()) {
        continue;
      }
      std::vector<IidTime> v;
      if ((err = measure_copies(c, sp, &v))) {
        return err;
      }
      std::lock_guard<std::mutex> lk(g_perf_mu);
      g_perf.curves[c] = std::move(v);
    }
  }
  for (int t = 0; t < TEMPI_NUM_TABLES; ++t) {
    const bool host = t == TEMPI_TABLE_PACK_HOST || t == TEMPI_TABLE_UNPACK_HOST;
    if (!(what & (host ? 8 : 4)) || cur.tables[t].rows) {
      continue;
    }
    Table tb;
    if ((err = measure_table(t, sp, &tb))) {
      return err;
    }
    std::lock_guard<std::mutex> lk(g_perf_mu);
    g_perf.tables[t] = std::move(tb);
  }
  std::lock_guard<std::mutex> lk(g_perf_mu);
  g_method_cache.clear();
  ++g_perf_generation;
  return TEMPI_SUCCESS;
}

TEMPI_EXPORT int tempi_perf_model(int method, int colocated, int64_t bytes,
                                  int64_t block_length, double *seconds) {
  if (method != TEMPI_METHOD_DEVICE && method != TEMPI_METHOD_ONESHOT &&
      method != TEMPI_METHOD_STAGED) {
    return set_error(TEMPI_ERR_ARG, "method %d", method);
  }
  if (bytes < 0 || block_length < 0 || !seconds) {
    return set_error(TEMPI_ERR_ARG, "bytes %lld, block %lld",
                     (long long)bytes, (long long)block_length);
  }
  std::lock_guard<std::mutex> lk(g_perf_mu);
  const double t = model_locked(method, colocated != 0, bytes, block_length);
  if (!std::isfinite(t)) {
    return set_error(TEMPI_ERR_UNSUPPORTED, "no measurements for method %d",
                     method);
  }
  *seconds = t;
  return TEMPI_SUCCESS;
}

TEMPI_EXPORT int tempi_choose_method(int colocated, int64_t bytes,
                                     int64_t block_length) {
  if (bytes < 0 || block_length < 0) {
    return TEMPI_METHOD_UNKNOWN;
  }
  std::lock_guard<std::mutex> lk(g_perf_mu);
  const auto key = std::make_tuple(colocated != 0, bytes, block_length);
  auto it = g_method_cache.find(key);
  if (it != g_method_cache.end()) {
    return it->second;
  }
  const double o =
      model_locked(TEMPI_METHOD_ONESHOT, colocated != 0, bytes, block_length);
  const double d =
      model_locked(TEMPI_METHOD_DEVICE, colocated != 0, bytes, block_length);
  const int m = o < d ? TEMPI_METHOD_ONESHOT : TEMPI_METHOD_DEVICE;
  g_method_cache[key] = m;
  return m;
}
