// kspn/kaspan.hpp -- C++ host interface mirroring rawbby/KaSpan's SCC entry
// points over libkspn's C ABI (kspn/scc.h). Header-only; link libkspn.so.
//
//   kaspan::scc(graph, scc_id)                 scc/scc.hpp:21-87
//   kaspan::scc_trim_ex(graph, scc_id)         scc/variant/scc_trim_ex.hpp:16-100
//   kaspan::scc_trim_ex_residual(graph, ...)   scc/variant/scc_trim_ex_residual.hpp:17-128
//   kaspan::scc_trim_ex_light_residual(...)    scc/variant/scc_trim_ex_light_residual.hpp:16-131
//   kaspan::scc_hpc_like(graph, scc_id)        scc/variant/scc_hpc_like.hpp:16-69
//   kaspan::scc_hpc_trim_ex(graph, scc_id)     scc/variant/scc_hpc_trim_ex.hpp:16-76
//   kaspan::scc_ispan_like(graph, scc_id)      scc/variant/scc_ispan_like.hpp:18-87
//   kaspan::async::scc(graph, scc_id)          scc/async/scc.hpp:22-90 (same stages as scc)
//   kaspan::manifest::load(path)               scc/adapter/manifest.hpp:136-197
//   kaspan::load_graph_from_manifest(mf)       scc/adapter/manifest.hpp:200-250
//   kaspan::backward_complement_graph(g)       scc/backward_complement.hpp:22-68
//   kaspan::normalise_scc_id(n, scc_id)        test/normalise_scc.hpp:10-25
//   kaspan::scc_batch(batch, scc_id)           one scc() per graph of a packed batch
//
// The graph is a device-resident bidi_graph_view (graph/bidi_graph.hpp:250):
// n, m, fw.head/fw.csr, bw.head/bw.csr with uint64 heads and uint32 ids
// (the reference built with KASPAN_INDEX64). scc_id is a device array of n
// entries and receives canonical ids (minimum vertex id per component).
// Where the reference aborts, these functions throw kaspan::error.
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>

#include "scc.h"

namespace kaspan {

using vertex_t = uint32_t;
using index_t = uint64_t;

struct error : std::runtime_error {
  int code;
  error(int c, const char* where)
      : std::runtime_error(std::string(where) + ": " + kspn_strerror(c) +
                           (c == KSPN_E_HIP ? std::string(" [") + kspn_last_error() + "]" : std::string())),
        code(c) {}
};

inline void check(int rc, const char* where) {
  if (rc != KSPN_OK) throw error(rc, where);
}

struct csr_view {
  index_t* head = nullptr;  // n + 1 entries, device
  vertex_t* csr = nullptr;  // m entries, device
};

struct bidi_graph_view {
  vertex_t n = 0;
  index_t m = 0;
  csr_view fw, bw;

  kspn_graph c_view() const { return kspn_graph{n, m, fw.head, fw.csr, bw.head, bw.csr}; }
};

// one context per thread of use and device; holds the solver's scratch
class context {
 public:
  explicit context(int device = 0) { check(kspn_ctx_create(&ctx_, device), "kspn_ctx_create"); }
  ~context() { kspn_ctx_destroy(ctx_); }
  context(const context&) = delete;
  context& operator=(const context&) = delete;
  kspn_ctx* get() const { return ctx_; }

  static context& thread_default() {
    thread_local context c(0);
    return c;
  }

 private:
  kspn_ctx* ctx_ = nullptr;
};

inline kspn_opts default_opts() {
  kspn_opts o;
  kspn_opts_default(&o);
  return o;
}

inline void scc(bidi_graph_view graph, vertex_t* scc_id, const kspn_opts& opts = default_opts(),
                kspn_stats* stats = nullptr, void* stream = nullptr, context& ctx = context::thread_default()) {
  kspn_graph g = graph.c_view();
  check(kspn_scc(ctx.get(), &g, scc_id, &opts, stats, stream), "kaspan::scc");
}

inline void scc_method(uint32_t method, bidi_graph_view graph, vertex_t* scc_id, kspn_stats* stats,
                       void* stream) {
  kspn_opts o = default_opts();
  o.method = method;
  scc(graph, scc_id, o, stats, stream);
}

inline void scc_trim_ex(bidi_graph_view graph, vertex_t* scc_id, kspn_stats* stats = nullptr, void* stream = nullptr) {
  scc_method(KSPN_METHOD_TRIM_EX, graph, scc_id, stats, stream);
}

// one rank: the decided threshold n - n/1 is 0, so everything after the
// exhaustive trim is the residual
inline void scc_trim_ex_residual(bidi_graph_view graph, vertex_t* scc_id, kspn_stats* stats = nullptr,
                                 void* stream = nullptr) {
  scc_method(KSPN_METHOD_TRIM_EX_RESIDUAL, graph, scc_id, stats, stream);
}

inline void scc_trim_ex_light_residual(bidi_graph_view graph, vertex_t* scc_id, kspn_stats* stats = nullptr,
                                       void* stream = nullptr) {
  scc_method(KSPN_METHOD_TRIM_EX_LIGHT_RESIDUAL, graph, scc_id, stats, stream);
}

inline void scc_hpc_like(bidi_graph_view graph, vertex_t* scc_id, kspn_stats* stats = nullptr, void* stream = nullptr) {
  scc_method(KSPN_METHOD_HPC_LIKE, graph, scc_id, stats, stream);
}

inline void scc_hpc_trim_ex(bidi_graph_view graph, vertex_t* scc_id, kspn_stats* stats = nullptr,
                            void* stream = nullptr) {
  scc_method(KSPN_METHOD_HPC_TRIM_EX, graph, scc_id, stats, stream);
}

inline void scc_ispan_like(bidi_graph_view graph, vertex_t* scc_id, kspn_stats* stats = nullptr,
                           void* stream = nullptr) {
  scc_method(KSPN_METHOD_ISPAN_LIKE, graph, scc_id, stats, stream);
}

namespace async {
// the asynchronous variant differs from scc() only in how ranks exchange
// frontiers; at one device its stages are scc()'s
inline void scc(bidi_graph_view graph, vertex_t* scc_id, kspn_stats* stats = nullptr, void* stream = nullptr) {
  scc_method(KSPN_METHOD_KASPAN, graph, scc_id, stats, stream);
}
}  // namespace async

// every graph of a packed batch (kspn_batch: union graph, batch-global ids);
// scc_id[vtx_off[i] + v] receives v's canonical id inside graph i
inline void scc_batch(const kspn_batch& batch, vertex_t* scc_id, const kspn_opts& opts = default_opts(),
                      kspn_stats* stats = nullptr, void* stream = nullptr,
                      context& ctx = context::thread_default()) {
  check(kspn_scc_batch(ctx.get(), &batch, scc_id, &opts, stats, stream), "kaspan::scc_batch");
}

// fills g.bw from g.fw (both device)
inline void backward_complement_graph(bidi_graph_view g, void* stream = nullptr,
                                      context& ctx = context::thread_default()) {
  check(kspn_transpose(ctx.get(), g.n, g.m, g.fw.head, g.fw.csr, g.bw.head, g.bw.csr, stream),
        "kaspan::backward_complement_graph");
}

inline void normalise_scc_id(vertex_t n, vertex_t* scc_id, void* stream = nullptr,
                             context& ctx = context::thread_default()) {
  check(kspn_canonicalize(ctx.get(), n, scc_id, stream), "kaspan::normalise_scc_id");
}

// a device-resident graph whose four arrays the library allocated
class bidi_graph {
 public:
  bidi_graph() = default;
  bidi_graph(const bidi_graph&) = delete;
  bidi_graph& operator=(const bidi_graph&) = delete;
  bidi_graph(bidi_graph&& o) noexcept : v_(o.v_) { o.v_ = bidi_graph_view{}; }
  bidi_graph& operator=(bidi_graph&& o) noexcept {
    if (this != &o) {
      release();
      v_ = o.v_;
      o.v_ = bidi_graph_view{};
    }
    return *this;
  }
  ~bidi_graph() { release(); }
  bidi_graph_view view() const { return v_; }
  vertex_t n() const { return v_.n; }
  index_t m() const { return v_.m; }

  // takes ownership of arrays allocated by the library (kspn_generate, kspn_manifest_graph)
  static bidi_graph adopt(vertex_t n, index_t m, index_t* fw_head, vertex_t* fw_csr, index_t* bw_head,
                          vertex_t* bw_csr) {
    bidi_graph g;
    g.v_ = bidi_graph_view{n, m, {fw_head, fw_csr}, {bw_head, bw_csr}};
    return g;
  }

 private:
  void release() {
    for (void* p : {(void*)v_.fw.head, (void*)v_.fw.csr, (void*)v_.bw.head, (void*)v_.bw.csr})
      if (p) kspn_free(p);
    v_ = bidi_graph_view{};
  }
  bidi_graph_view v_{};
};

// the manifest of a graph converted to dense binary arrays
struct manifest : kspn_manifest {
  static manifest load(const std::string& file) {
    manifest mf;
    int rc = kspn_manifest_parse(file.c_str(), &mf);
    if (rc != KSPN_OK) throw error(rc, ("kaspan::manifest::load: " + std::string(kspn_last_error())).c_str());
    return mf;
  }
};

inline bidi_graph load_graph_from_manifest(const manifest& mf, void* stream = nullptr,
                                           context& ctx = context::thread_default()) {
  uint32_t n = 0;
  uint64_t m = 0;
  uint64_t *fo = nullptr, *bo = nullptr;
  uint32_t *fa = nullptr, *ba = nullptr;
  int rc = kspn_manifest_graph(ctx.get(), &mf, &n, &m, &fo, &fa, &bo, &ba, stream);
  if (rc != KSPN_OK) throw error(rc, ("kaspan::load_graph_from_manifest: " + std::string(kspn_last_error())).c_str());
  return bidi_graph::adopt(n, m, fo, fa, bo, ba);
}

}  // namespace kaspan
