// C++ caller of the kaspan:: mirror over the C ABI: the iSpan tiny graph
// (extern/ispan_test/test_ispan_tiny.cpp:21-90) and a directed cycle under
// every variant, and (with a manifest path as argument) the tiny graph
// loaded through kaspan::manifest::load / load_graph_from_manifest.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <vector>

#include "kspn/kaspan.hpp"

#define HIP_OK(x) do { if ((x) != hipSuccess) { std::printf("HIP error line %d\n", __LINE__); return 2; } } while (0)

static int run(std::vector<uint64_t> fo, std::vector<uint32_t> fa, const std::vector<uint32_t>& expect, int method) {
  uint32_t n = (uint32_t)fo.size() - 1;
  uint64_t m = fa.size();
  kaspan::bidi_graph_view g;
  g.n = n;
  g.m = m;
  uint32_t* d_ids = nullptr;
  HIP_OK(hipMalloc(&g.fw.head, 8 * (n + 1)));
  HIP_OK(hipMalloc(&g.bw.head, 8 * (n + 1)));
  HIP_OK(hipMalloc(&g.fw.csr, 4 * m));
  HIP_OK(hipMalloc(&g.bw.csr, 4 * m));
  HIP_OK(hipMalloc(&d_ids, 4 * n));
  HIP_OK(hipMemcpy(g.fw.head, fo.data(), 8 * (n + 1), hipMemcpyHostToDevice));
  HIP_OK(hipMemcpy(g.fw.csr, fa.data(), 4 * m, hipMemcpyHostToDevice));
  kaspan::backward_complement_graph(g);
  kspn_opts o = kaspan::default_opts();
  o.method = method;
  kaspan::scc(g, d_ids, o);
  std::vector<uint32_t> got(n);
  HIP_OK(hipMemcpy(got.data(), d_ids, 4 * n, hipMemcpyDeviceToHost));
  for (void* p : {(void*)g.fw.head, (void*)g.bw.head, (void*)g.fw.csr, (void*)g.bw.csr, (void*)d_ids}) (void)hipFree(p);
  for (uint32_t i = 0; i < n; ++i)
    if (got[i] != expect[i]) {
      std::printf("mismatch at %u: %u != %u\n", i, got[i], expect[i]);
      return 1;
    }
  return 0;
}

// argv[1]: manifest of the iSpan tiny graph
static int run_manifest(const char* path) {
  const std::vector<uint32_t> expect{0, 1, 0, 1, 1, 5, 5};
  auto mf = kaspan::manifest::load(path);
  auto g = kaspan::load_graph_from_manifest(mf);
  if (g.n() != 7 || g.m() != 11) {
    std::printf("manifest graph has n=%u m=%llu\n", g.n(), (unsigned long long)g.m());
    return 1;
  }
  uint32_t* d_ids = nullptr;
  HIP_OK(hipMalloc(&d_ids, 4 * 7));
  kspn_stats st;
  kaspan::scc_ispan_like(g.view(), d_ids, &st);
  std::vector<uint32_t> got(7);
  HIP_OK(hipMemcpy(got.data(), d_ids, 4 * 7, hipMemcpyDeviceToHost));
  (void)hipFree(d_ids);
  if (got != expect) {
    std::printf("manifest graph: wrong labels\n");
    return 1;
  }
  // a manifest naming a file that does not exist
  try {
    auto bad = kaspan::manifest::load(std::string(path) + ".missing");
    std::printf("expected an exception\n");
    return 1;
  } catch (const kaspan::error& e) {
    if (e.code != KSPN_E_IO) return 1;
  }
  return 0;
}

int main(int argc, char** argv) {
  try {
    if (argc > 1 && run_manifest(argv[1])) return 1;
    for (int method : {KSPN_METHOD_KASPAN, KSPN_METHOD_TRIM_EX, KSPN_METHOD_HPC_LIKE, KSPN_METHOD_ISPAN_LIKE,
                       KSPN_METHOD_HPC_TRIM_EX, KSPN_METHOD_TRIM_EX_RESIDUAL, KSPN_METHOD_TRIM_EX_LIGHT_RESIDUAL}) {
      if (run({0, 1, 4, 5, 7, 9, 10, 11}, {2, 0, 2, 3, 0, 2, 4, 1, 3, 6, 5}, {0, 1, 0, 1, 1, 5, 5}, method)) return 1;
      std::vector<uint64_t> off(1001);
      std::vector<uint32_t> adj(1000), zero(1000, 0);
      for (uint32_t i = 0; i <= 1000; ++i) off[i] = i;
      for (uint32_t i = 0; i < 1000; ++i) adj[i] = (i + 1) % 1000;
      if (run(off, adj, zero, method)) return 1;
    }
    // errors surface as exceptions
    kaspan::bidi_graph_view bad;
    bad.n = 3;
    try {
      kaspan::scc(bad, nullptr);
      std::printf("expected an exception\n");
      return 1;
    } catch (const kaspan::error& e) {
      if (e.code != KSPN_E_INVALID) return 1;
    }
  } catch (const std::exception& e) {
    std::printf("exception: %s\n", e.what());
    return 1;
  }
  std::printf("cpp ok\n");
  return 0;
}
