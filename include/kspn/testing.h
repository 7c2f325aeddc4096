/*
 * kspn/testing.h -- test hooks of libkspn. They have no counterpart in the
 * reference; tests use them to drive internal guards with inputs that the
 * public entry points never produce.
 */
#ifndef KSPN_TESTING_H
#define KSPN_TESTING_H

#include "kspn/scc.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Runs the tail solver's trim split, degree pass and sub-CSR fill, one after
 * the other, over `list` (count <= 4095 device ids, every vertex of g
 * taken as live) as if the graph had only n_limit <= g->n vertices. Bit k
 * of *flagged is set when kernel k (1 split, 2 degree pass, 4 fill)
 * reported a listed id >= n_limit instead of reading with it. The arrays of
 * g hold every id < g->n, so the call stays in bounds whatever the guards
 * do. Synchronises the stream. */
int kspn_test_tail_guard(kspn_ctx* ctx, const kspn_graph* g, uint32_t n_limit, const uint32_t* list, uint32_t count,
                         uint32_t* flagged, void* stream);

#ifdef __cplusplus
}
#endif
#endif
