// Node-centred gather of per-element contributions, tiled through LDS
// (body for gfx950 and the CPU emulator).
//
// The reference adds each element's contributions to its nodes with
// VecSetValues(ADD_VALUES) element after element (assembling_rhs.F90:124,
// fem_rhs_axi.F90:982).  Here a node's value is the sum of its (element,
// local node) entries taken in element order from the node's CSR list: the
// same order, deterministic, without atomics.
//
// A work-group handles GATHER_NODES consecutive nodes and a chunk of
// GATHER_MCHUNK modes at a time.  Reads: a lane takes one mode of one node,
// so 16 lanes read the 16 consecutive modes of one entry and component, one
// whole 128-byte line (elo layout: entry, component, mode).  Writes: the
// sums go to an LDS tile and leave it with the nodes fastest, so a wave
// writes runs of GATHER_NODES consecutive nodes of one output plane.  (32
// nodes x 16 modes: 1.345 ms for the 6-component gather at 200,704
// triangles, M = 64, against 1.393 ms with 16 nodes x 32 modes, 1.45 ms
// with 16 x 16, 1.49-1.53 ms with 32 x 32 and 64 x 16.)
#pragma once
#include "sfx_portable.h"

namespace sfx {

constexpr int GATHER_NODES = 32;
constexpr int GATHER_THREADS = 256;
constexpr int GATHER_MCHUNK = 16;

// LDS doubles of the tile for NK components.
template <int NK>
constexpr int gather_tile_doubles() { return NK * GATHER_MCHUNK * (GATHER_NODES + 1); }

// Sum the entries of nodes [n0, n0 + GATHER_NODES) for modes [m0, m0 + mc)
// into tile[(k * mc + mi) * (GATHER_NODES + 1) + node].
template <int NK>
SFX_DEV void gather_sum(const int64_t* adj_ptr, const int32_t* adj_ent, int64_t n_adj, const double* elo,
                        int64_t nelo, int64_t np, int M, int64_t n0, int m0, int mc, int tid, double* tile) {
  constexpr int ROW = GATHER_NODES + 1;
  const int items = GATHER_NODES * mc;
  for (int it = tid; it < items; it += GATHER_THREADS) {
    const int mi = it % mc, node = it / mc;
    const int64_t n = n0 + node;
    double acc[NK];
#pragma unroll
    for (int k = 0; k < NK; ++k) acc[k] = 0.0;
    if (n < np) {
      for (int64_t q = adj_ptr[n]; q < adj_ptr[n + 1]; ++q) {
        SFX_CHECK("adj", q, n_adj);
        const int64_t ent = adj_ent[q];
#pragma unroll
        for (int k = 0; k < NK; ++k) {
          const int64_t i = (ent * NK + k) * M + m0 + mi;
          SFX_CHECK("elo load", i, nelo);
          acc[k] += elo[i];
        }
      }
    }
#pragma unroll
    for (int k = 0; k < NK; ++k) {
      SFX_CHECK("gather tile", (k * mc + mi) * ROW + node, gather_tile_doubles<NK>());
      tile[(k * mc + mi) * ROW + node] = acc[k];
    }
  }
}

// Work-group body: for each chunk of modes, sum into the tile, then call
// store(k, m, n, value) with the nodes fastest.
template <int NK, typename Store>
SFX_DEV void gather_body(const int64_t* adj_ptr, const int32_t* adj_ent, int64_t n_adj, const double* elo,
                         int64_t nelo, int64_t np, int M, int64_t bid, int tid, double* tile, Store store) {
  constexpr int ROW = GATHER_NODES + 1;
  const int64_t n0 = bid * GATHER_NODES;
  for (int m0 = 0; m0 < M; m0 += GATHER_MCHUNK) {
    const int mc = (M - m0 < GATHER_MCHUNK) ? M - m0 : GATHER_MCHUNK;
    SFX_SYNC();
    gather_sum<NK>(adj_ptr, adj_ent, n_adj, elo, nelo, np, M, n0, m0, mc, tid, tile);
    SFX_SYNC();
    const int outs = NK * mc * GATHER_NODES;
    for (int it = tid; it < outs; it += GATHER_THREADS) {
      const int node = it % GATHER_NODES;
      const int mi = (it / GATHER_NODES) % mc;
      const int k = it / (GATHER_NODES * mc);
      const int64_t n = n0 + node;
      if (n < np) store(k, m0 + mi, n, tile[(k * mc + mi) * ROW + node]);
    }
  }
}

}  // namespace sfx
