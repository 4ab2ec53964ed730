// Helpers shared by the C-ABI translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "sfx_kernels.h"

// Opaque mesh handle: device copies of the connectivity and coordinates plus
// the reference-element tables.
struct sfx_mesh_s {
  int device = 0;
  int64_t me = 0, np = 0;
  int32_t* jj = nullptr;    // (6, me)
  double* rr = nullptr;     // (2, np)
  double* tab = nullptr;    // ww(6,7), dd(2,6,7)
  // deterministic node-centred gather for the right-hand-side assembly
  int64_t* adj_ptr = nullptr;   // (np+1) CSR over nodes
  int32_t* adj_ent = nullptr;   // entries e*6 + local node, element order
  int64_t n_adj = 0;
  double* wgt = nullptr;        // (7, me) rj * ray
  double* mdinv = nullptr;      // (np) 1 / diagonal of the r-weighted mass matrix
  // boundary sides (mesh%jjs) for the boundary term of rhs_3x3; mes = -1: none registered
  int64_t mes = -1, nb = 0, n_bent = 0;
  int64_t* bptr = nullptr;      // (nb+1) CSR over boundary nodes
  int32_t* bent = nullptr;      // 3*side + local node, side order
  int64_t* bnode = nullptr;     // (nb) node index
  double* bw = nullptr;         // (3, mes) ray * rjs per side Gauss point
};

namespace sfx {

int fail(int code, const char* fmt, ...);
int hip_fail(hipError_t e, const char* where);
void clear_err();
int get_twiddles(int MP, Twiddles* tw);
int staging(int slot, size_t bytes, void** out);
int stream_scratch(int slot, size_t bytes, hipStream_t st, void** out);
int release_stream_scratch(hipStream_t st);
double now_s();
int check_modes(const char* name, int64_t np, int32_t m_max_c, int32_t m_max_pad);
int check_procs(const char* name, int64_t np, int32_t nb_procs, int64_t bloc_size);
// Exact-grid products behind FFT_SCALAR_VECT_DCL (pb = 2) and FFT_TENSOR_DCL (opt_tension).
int exact_sv2_dev(const double* V1, const double* V2, double* out, int64_t np, int M, int mp, hipStream_t st);
int exact_tension_dev(const double* V1, const double* V2, double* out, int64_t np, int M, int mp, hipStream_t st);
int exact_binary_host(int op, const double* V1, const double* V2, double* out, int64_t np, int M, int mp,
                      double* temps);

}  // namespace sfx
