// kspn_manifest.hip -- graphs stored as a manifest: a text file of
// "key value" lines naming four binary arrays of dense unsigned integers
// (row offsets and adjacency of the CSR and of the reverse CSR), each
// element 1..8 bytes wide, little or big endian.
//
//   kaspan/include/kaspan/scc/adapter/manifest.hpp:19-113   key/value parsing
//   kaspan/include/kaspan/scc/adapter/manifest.hpp:117-197  manifest::load
//   kaspan/include/kaspan/scc/adapter/manifest.hpp:200-250  load_graph_from_manifest
//   kaspan/include/kaspan/memory/dense_unsigned_ops.hpp:14-44  element decoding
//   kaspan/include/kaspan/memory/file_buffer.hpp:63-83     exact file sizes
//
// The reference maps the files and widens every element on one CPU core.
// Here the packed bytes stream through two pinned host buffers into HBM
// (file reads overlap the DMA and the kernels of the previous chunk) and a
// kernel widens them in place to the solver's layout, u64 offsets and u32
// ids, checking every value's range on the way. Offsets are also checked to
// start at 0, end at m and never decrease, which the solver relies on.
#include <fcntl.h>
#include <sys/stat.h>
#include <unistd.h>

#include <cerrno>
#include <climits>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <thread>
#include <unordered_map>

#include "kspn_internal.h"

namespace kspn {

namespace {

// ------------------------------------------------------------------ device

// out[i] = element i of `src` (w bytes, little or big endian); err[0] is set
// when an element exceeds `limit`
__global__ __launch_bounds__(WG) void k_unpack(const uint8_t* __restrict__ src, uint64_t count, uint32_t w,
                                               uint32_t big, uint64_t limit, void* __restrict__ out, uint32_t out_w,
                                               uint32_t* err) {
  bool bad = false;
  for (uint64_t i = blockIdx.x * (uint64_t)WG + threadIdx.x; i < count; i += (uint64_t)gridDim.x * WG) {
    const uint8_t* p = src + i * w;
    uint64_t v = 0;
    if (big) {
      for (uint32_t j = 0; j < w; ++j) v = (v << 8) | p[j];
    } else {
      for (uint32_t j = 0; j < w; ++j) v |= (uint64_t)p[j] << (8 * j);
    }
    bad |= v > limit;
    if (out_w == 8) ((uint64_t*)out)[i] = v;
    else ((uint32_t*)out)[i] = (uint32_t)v;
  }
  if (__any(bad) && lane_id() == 0) atomicOr(&err[0], 1u);
}

// off[0] == 0, off[n] == m, off[i] <= off[i+1]; err[1] is set otherwise
__global__ __launch_bounds__(WG) void k_check_offsets(const uint64_t* __restrict__ off, uint32_t n, uint64_t m,
                                                      uint32_t* err) {
  bool bad = false;
  for (uint64_t i = blockIdx.x * (uint64_t)WG + threadIdx.x; i < n; i += (uint64_t)gridDim.x * WG)
    bad |= off[i] > off[i + 1];
  if (blockIdx.x == 0 && threadIdx.x == 0) bad |= off[0] != 0 || off[n] != m;
  if (__any(bad) && lane_id() == 0) atomicOr(&err[1], 1u);
}

uint32_t unpack_grid(uint64_t items) {
  uint64_t g = (items + WG - 1) / WG;
  if (g < 1) g = 1;
  if (g > 16384) g = 16384;
  return (uint32_t)g;
}

// -------------------------------------------------------------------- host

int fail(int code, const std::string& what) {
  note_error("manifest", what.c_str(), __FILE__, __LINE__);
  return code;
}

bool copy_text(char* dst, size_t cap, const std::string& s) {
  if (s.size() + 1 > cap) return false;
  std::memcpy(dst, s.c_str(), s.size() + 1);
  return true;
}

// std::from_chars for an unsigned integer: decimal digits only, whole string
bool parse_u64(const std::string& s, uint64_t* out) {
  if (s.empty()) return false;
  uint64_t v = 0;
  for (char ch : s) {
    if (ch < '0' || ch > '9') return false;
    uint64_t d = (uint64_t)(ch - '0');
    if (v > (UINT64_MAX - d) / 10) return false;
    v = v * 10 + d;
  }
  *out = v;
  return true;
}

// Two pinned host buffers and two device buffers; chunk k uses pair k % 2,
// and done[k % 2] marks the end of that chunk's copy and kernel.
struct Stager {
  size_t cap = 0;
  uint8_t* host[2] = {nullptr, nullptr};
  uint8_t* dev[2] = {nullptr, nullptr};
  hipEvent_t done[2] = {nullptr, nullptr};
  uint32_t* err = nullptr;  // device: [0] value out of range, [1] offsets malformed
  ~Stager() {
    for (int b = 0; b < 2; ++b) {
      if (done[b]) {
        (void)hipEventSynchronize(done[b]);
        (void)hipEventDestroy(done[b]);
      }
      if (host[b]) (void)hipHostFree(host[b]);
      if (dev[b]) (void)hipFree(dev[b]);
    }
    if (err) (void)hipFree(err);
  }
  int init(size_t bytes, hipStream_t s) {
    cap = bytes;
    for (int b = 0; b < 2; ++b) {
      if (hipHostMalloc((void**)&host[b], cap, hipHostMallocDefault) != hipSuccess) {
        host[b] = nullptr;
        return KSPN_E_OOM;
      }
      if (hipMalloc((void**)&dev[b], cap) != hipSuccess) {
        dev[b] = nullptr;
        return KSPN_E_OOM;
      }
      KSPN_HIP(hipEventCreateWithFlags(&done[b], hipEventDisableTiming));
      KSPN_HIP(hipEventRecord(done[b], s));
    }
    if (hipMalloc((void**)&err, 4 * sizeof(uint32_t)) != hipSuccess) {
      err = nullptr;
      return KSPN_E_OOM;
    }
    KSPN_HIP(hipMemsetAsync(err, 0, 4 * sizeof(uint32_t), s));
    return KSPN_OK;
  }
};

constexpr size_t kStageBytes = 64ull << 20;

// reads `bytes` bytes at `offset` of fd into dst
bool read_span(int fd, uint8_t* dst, size_t bytes, uint64_t offset) {
  while (bytes > 0) {
    ssize_t r = pread(fd, dst, bytes, (off_t)offset);
    if (r < 0 && errno == EINTR) continue;
    if (r <= 0) return false;
    dst += r;
    bytes -= (size_t)r;
    offset += (uint64_t)r;
  }
  return true;
}

// A chunk is read by several threads at once: one thread copying out of the
// page cache reaches about 10 GB/s, well below the link to the device.
constexpr int kReadThreads = 4;
constexpr size_t kReadSplit = 8ull << 20;  // smallest part worth a thread

bool read_full(int fd, uint8_t* dst, size_t bytes, uint64_t offset) {
  const int parts = (int)std::min<size_t>(kReadThreads, std::max<size_t>(1, bytes / kReadSplit));
  if (parts == 1) return read_span(fd, dst, bytes, offset);
  const size_t per = (bytes + parts - 1) / parts;
  bool ok[kReadThreads];
  std::vector<std::thread> th;
  for (int i = 1; i < parts; ++i) {
    const size_t b = (size_t)i * per, e = std::min(bytes, b + per);
    th.emplace_back([=, &ok] { ok[i] = read_span(fd, dst + b, e - b, offset + b); });
  }
  ok[0] = read_span(fd, dst, std::min(bytes, per), offset);
  for (auto& t : th) t.join();
  for (int i = 0; i < parts; ++i)
    if (!ok[i]) return false;
  return true;
}

// widens the `count` elements of file `path` (w bytes each) into out (out_w bytes each)
int stream_array(Stager& S, const char* path, uint64_t count, uint32_t w, bool big, uint64_t limit, void* out,
                 uint32_t out_w, hipStream_t s) {
  const uint64_t bytes = count * w;
  if (bytes == 0) return KSPN_OK;  // the reference opens no file for an empty array
  int fd = open(path, O_RDONLY | O_CLOEXEC | O_NOFOLLOW);
  if (fd < 0) return fail(KSPN_E_IO, std::string("cannot open ") + path);
  struct stat st {};
  if (fstat(fd, &st) != 0 || !S_ISREG(st.st_mode)) {
    close(fd);
    return fail(KSPN_E_IO, std::string("not a regular file: ") + path);
  }
  if ((uint64_t)st.st_size != bytes) {
    close(fd);
    return fail(KSPN_E_IO, std::string(path) + " holds " + std::to_string((uint64_t)st.st_size) +
                               " bytes, the manifest asks for " + std::to_string(bytes));
  }
  const uint64_t per = S.cap / w;  // elements per chunk
  int rc = KSPN_OK;
  uint64_t k = 0;
  for (uint64_t e0 = 0; e0 < count && rc == KSPN_OK; e0 += per, ++k) {
    const uint64_t ne = std::min<uint64_t>(per, count - e0);
    const int b = (int)(k & 1);
    if (hipEventSynchronize(S.done[b]) != hipSuccess) {
      rc = fail(KSPN_E_HIP, "staging event");
      break;
    }
    if (!read_full(fd, S.host[b], ne * w, e0 * w)) {
      rc = fail(KSPN_E_IO, std::string("short read from ") + path);
      break;
    }
    if (hipMemcpyAsync(S.dev[b], S.host[b], ne * w, hipMemcpyHostToDevice, s) != hipSuccess) {
      rc = fail(KSPN_E_HIP, "staging copy");
      break;
    }
    k_unpack<<<unpack_grid(ne), WG, 0, s>>>(S.dev[b], ne, w, big ? 1u : 0u, limit, (uint8_t*)out + e0 * out_w, out_w,
                                           S.err);
    if (hipGetLastError() != hipSuccess || hipEventRecord(S.done[b], s) != hipSuccess) {
      rc = fail(KSPN_E_HIP, "unpack launch");
      break;
    }
  }
  close(fd);
  return rc;
}

}  // namespace

}  // namespace kspn

using namespace kspn;

extern "C" {

int kspn_manifest_parse(const char* path, kspn_manifest* out) {
  clear_error();
  if (!path || !out) return KSPN_E_INVALID;
  std::memset(out, 0, sizeof(*out));
  FILE* f = std::fopen(path, "rb");
  if (!f) return fail(KSPN_E_IO, std::string("cannot open ") + path);
  // parse_kv_map (manifest.hpp:20-51): '%' starts a comment line, the key
  // ends at the first space, the value starts after the spaces that follow
  std::unordered_map<std::string, std::string> kv;
  std::string line;
  int ch;
  bool more = true;
  int rc = KSPN_OK;
  while (more && rc == KSPN_OK) {
    line.clear();
    while ((ch = std::fgetc(f)) != EOF && ch != '\n') line.push_back((char)ch);
    more = ch != EOF;
    if (!more && line.empty()) break;
    if (line.empty() || line[0] == '%') continue;
    size_t key_end = line.find(' ');
    if (key_end == std::string::npos) {
      rc = fail(KSPN_E_INVALID, "line without a value: " + line);
      break;
    }
    size_t vs = key_end;
    while (vs < line.size() && line[vs] == ' ') ++vs;
    if (!kv.emplace(line.substr(0, key_end), vs < line.size() ? line.substr(vs) : std::string()).second)
      rc = fail(KSPN_E_INVALID, "repeated key " + line.substr(0, key_end));
  }
  std::fclose(f);
  if (rc != KSPN_OK) return rc;

  auto get = [&](const char* key, std::string* v) -> int {
    auto it = kv.find(key);
    if (it == kv.end()) return fail(KSPN_E_INVALID, std::string("missing key ") + key);
    *v = it->second;
    return KSPN_OK;
  };
  auto get_u64 = [&](const char* key, uint64_t* v) -> int {
    std::string s;
    KSPN_TRY(get(key, &s));
    if (!parse_u64(s, v)) return fail(KSPN_E_INVALID, std::string("not an unsigned integer: ") + key + " " + s);
    return KSPN_OK;
  };
  auto get_bool = [&](const char* key, uint32_t* v) -> int {
    std::string s;
    KSPN_TRY(get(key, &s));
    if (s == "true") *v = 1;
    else if (s == "false") *v = 0;
    else return fail(KSPN_E_INVALID, std::string("not true/false: ") + key + " " + s);
    return KSPN_OK;
  };
  // paths are relative to the manifest's directory and must name regular files
  std::string base(path);
  size_t slash = base.find_last_of('/');
  base = slash == std::string::npos ? std::string(".") : base.substr(0, slash == 0 ? 1 : slash);
  auto get_path = [&](const char* key, char* dst) -> int {
    std::string p;
    KSPN_TRY(get(key, &p));
    std::string full = (!p.empty() && p[0] == '/') ? p : base + "/" + p;
    char resolved[PATH_MAX];
    struct stat st {};
    if (!realpath(full.c_str(), resolved) || stat(resolved, &st) != 0 || !S_ISREG(st.st_mode))
      return fail(KSPN_E_IO, std::string("no such file: ") + key + " " + full);
    if (!copy_text(dst, KSPN_MANIFEST_PATH, resolved)) return fail(KSPN_E_INVALID, std::string("path too long: ") + key);
    return KSPN_OK;
  };

  uint64_t v = 0;
  std::string s;
  KSPN_TRY(get_u64("schema.version", &v));
  if (v > UINT32_MAX) return fail(KSPN_E_INVALID, "schema.version out of range");
  out->schema_version = (uint32_t)v;
  KSPN_TRY(get("graph.code", &s));
  if (!copy_text(out->code, KSPN_MANIFEST_TEXT, s)) return fail(KSPN_E_INVALID, "graph.code too long");
  KSPN_TRY(get("graph.name", &s));
  if (!copy_text(out->name, KSPN_MANIFEST_TEXT, s)) return fail(KSPN_E_INVALID, "graph.name too long");
  KSPN_TRY(get("graph.endian", &s));
  if (s == "little") out->big_endian = 0;
  else if (s == "big") out->big_endian = 1;
  else return fail(KSPN_E_INVALID, "graph.endian is neither little nor big: " + s);
  KSPN_TRY(get_u64("graph.node_count", &out->node_count));
  KSPN_TRY(get_u64("graph.edge_count", &out->edge_count));
  KSPN_TRY(get_bool("graph.contains_self_loops", &out->contains_self_loops));
  KSPN_TRY(get_bool("graph.contains_duplicate_edges", &out->contains_duplicate_edges));
  KSPN_TRY(get_u64("graph.head.bytes", &v));
  if (v > 255) return fail(KSPN_E_INVALID, "graph.head.bytes out of range");
  out->head_bytes = (uint32_t)v;
  KSPN_TRY(get_u64("graph.csr.bytes", &v));
  if (v > 255) return fail(KSPN_E_INVALID, "graph.csr.bytes out of range");
  out->csr_bytes = (uint32_t)v;
  KSPN_TRY(get_path("fw.head.path", out->fw_head_path));
  KSPN_TRY(get_path("fw.csr.path", out->fw_csr_path));
  KSPN_TRY(get_path("bw.head.path", out->bw_head_path));
  KSPN_TRY(get_path("bw.csr.path", out->bw_csr_path));
  return KSPN_OK;
}

int kspn_manifest_load(kspn_ctx* k, const kspn_manifest* mf, uint64_t* fw_off, uint32_t* fw_adj, uint64_t* bw_off,
                       uint32_t* bw_adj, void* stream) {
  clear_error();
  if (!k || !mf) return KSPN_E_INVALID;
  const uint64_t n = mf->node_count, m = mf->edge_count;
  // load_graph_from_manifest's bounds (manifest.hpp:214-223); ids are u32
  // with 0xFFFFFFFF reserved, offsets u64
  if (n + 1 > 0xFFFFFFFFull) return fail(KSPN_E_INVALID, "graph.node_count does not fit 32-bit vertex ids");
  if (mf->head_bytes < 1 || mf->head_bytes > 8 || mf->csr_bytes < 1 || mf->csr_bytes > 8)
    return fail(KSPN_E_INVALID, "element widths must be 1..8 bytes");
  if (m > 0 && n == 0) return fail(KSPN_E_INVALID, "edges without vertices");
  if (!fw_off || !bw_off || (m > 0 && (!fw_adj || !bw_adj))) return KSPN_E_INVALID;
  if (hipSetDevice(k->c.device) != hipSuccess) return KSPN_E_HIP;
  hipStream_t s = (hipStream_t)stream;
  Stager S;
  KSPN_TRY(S.init(kStageBytes, s));
  const bool big = mf->big_endian != 0;
  KSPN_TRY(stream_array(S, mf->fw_head_path, n + 1, mf->head_bytes, big, m, fw_off, 8, s));
  KSPN_TRY(stream_array(S, mf->fw_csr_path, m, mf->csr_bytes, big, n ? n - 1 : 0, fw_adj, 4, s));
  KSPN_TRY(stream_array(S, mf->bw_head_path, n + 1, mf->head_bytes, big, m, bw_off, 8, s));
  KSPN_TRY(stream_array(S, mf->bw_csr_path, m, mf->csr_bytes, big, n ? n - 1 : 0, bw_adj, 4, s));
  k_check_offsets<<<unpack_grid(n), WG, 0, s>>>(fw_off, (uint32_t)n, m, S.err);
  k_check_offsets<<<unpack_grid(n), WG, 0, s>>>(bw_off, (uint32_t)n, m, S.err + 2);
  KSPN_HIP(hipGetLastError());
  uint32_t h_err[4] = {0, 0, 0, 0};
  KSPN_HIP(hipMemcpyAsync(h_err, S.err, sizeof(h_err), hipMemcpyDeviceToHost, s));
  KSPN_HIP(hipStreamSynchronize(s));
  if (h_err[0]) return fail(KSPN_E_INVALID, "an offset exceeds graph.edge_count or an id graph.node_count - 1");
  if (h_err[1]) return fail(KSPN_E_INVALID, "fw head: offsets must start at 0, end at graph.edge_count and not decrease");
  if (h_err[3]) return fail(KSPN_E_INVALID, "bw head: offsets must start at 0, end at graph.edge_count and not decrease");
  return KSPN_OK;
}

int kspn_manifest_graph(kspn_ctx* k, const kspn_manifest* mf, uint32_t* n, uint64_t* m, uint64_t** fw_off,
                        uint32_t** fw_adj, uint64_t** bw_off, uint32_t** bw_adj, void* stream) {
  clear_error();
  if (!k || !mf || !n || !m || !fw_off || !fw_adj || !bw_off || !bw_adj) return KSPN_E_INVALID;
  if (mf->node_count + 1 > 0xFFFFFFFFull) return fail(KSPN_E_INVALID, "graph.node_count does not fit 32-bit vertex ids");
  if (hipSetDevice(k->c.device) != hipSuccess) return KSPN_E_HIP;
  const uint64_t N = mf->node_count, M = mf->edge_count;
  void* p[4] = {nullptr, nullptr, nullptr, nullptr};
  const uint64_t bytes[4] = {8 * (N + 1), 4 * (M ? M : 1), 8 * (N + 1), 4 * (M ? M : 1)};
  int rc = KSPN_OK;
  for (int i = 0; i < 4 && rc == KSPN_OK; ++i)
    if (hipMalloc(&p[i], bytes[i]) != hipSuccess) {
      p[i] = nullptr;
      rc = KSPN_E_OOM;
    }
  if (rc == KSPN_OK)
    rc = kspn_manifest_load(k, mf, (uint64_t*)p[0], (uint32_t*)p[1], (uint64_t*)p[2], (uint32_t*)p[3], stream);
  if (rc != KSPN_OK) {
    for (void* q : p)
      if (q) (void)hipFree(q);
    return rc;
  }
  *n = (uint32_t)N;
  *m = M;
  *fw_off = (uint64_t*)p[0];
  *fw_adj = (uint32_t*)p[1];
  *bw_off = (uint64_t*)p[2];
  *bw_adj = (uint32_t*)p[3];
  return KSPN_OK;
}

int kspn_unpack_dense(kspn_ctx* k, const void* packed, uint64_t count, uint32_t elem_bytes, uint32_t big_endian,
                      uint64_t limit, void* out, uint32_t out_bytes, void* stream) {
  clear_error();
  if (!k || elem_bytes < 1 || elem_bytes > 8 || (out_bytes != 4 && out_bytes != 8)) return KSPN_E_INVALID;
  if (count > 0 && (!packed || !out)) return KSPN_E_INVALID;
  if (out_bytes == 4 && limit > 0xFFFFFFFFull) return KSPN_E_INVALID;
  if (hipSetDevice(k->c.device) != hipSuccess) return KSPN_E_HIP;
  hipStream_t s = (hipStream_t)stream;
  uint32_t* err = nullptr;
  if (hipMalloc((void**)&err, 4 * sizeof(uint32_t)) != hipSuccess) return KSPN_E_OOM;
  uint32_t h_err[4] = {0, 0, 0, 0};
  hipError_t e = hipMemsetAsync(err, 0, 4 * sizeof(uint32_t), s);
  if (e == hipSuccess && count > 0) {
    k_unpack<<<unpack_grid(count), WG, 0, s>>>((const uint8_t*)packed, count, elem_bytes, big_endian ? 1u : 0u, limit,
                                              out, out_bytes, err);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpyAsync(h_err, err, sizeof(h_err), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  (void)hipFree(err);
  if (e != hipSuccess) return fail(KSPN_E_HIP, hipGetErrorName(e));
  if (h_err[0]) return fail(KSPN_E_INVALID, "a value exceeds the limit");
  return KSPN_OK;
}

}  // extern "C"
