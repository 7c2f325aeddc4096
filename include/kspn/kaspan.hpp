// kspn/kaspan.hpp -- C++ host interface mirroring rawbby/KaSpan's SCC entry
// points over libkspn's C ABI (kspn/scc.h). Header-only; link libkspn.so.
//
//   kaspan::scc(graph, scc_id)                 scc/scc.hpp:21-87
//   kaspan::scc_trim_ex(graph, scc_id)         scc/variant/scc_trim_ex.hpp:16-100
//   kaspan::scc_trim_ex_residual(graph, ...)   scc/variant/scc_trim_ex_residual.hpp:17-128
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

inline void scc_trim_ex(bidi_graph_view graph, vertex_t* scc_id, kspn_stats* stats = nullptr, void* stream = nullptr) {
  kspn_opts o = default_opts();
  o.method = KSPN_METHOD_TRIM_EX;
  scc(graph, scc_id, o, stats, stream);
}

inline void scc_trim_ex_residual(bidi_graph_view graph, vertex_t* scc_id, kspn_stats* stats = nullptr,
                                 void* stream = nullptr) {
  kspn_opts o = default_opts();
  o.method = KSPN_METHOD_TRIM_EX;
  o.residual_threshold = 0xFFFFFFFFu;  // one rank: everything after the trim goes to Tarjan
  scc(graph, scc_id, o, stats, stream);
}

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

}  // namespace kaspan
