// twoldg_impl.h -- kernels, launch templates and the problem record shared by
// the translation units of libtwoldg (one per arithmetic type and dimension).
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <type_traits>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <utility>
#include <vector>

#include "../../include/twoldg.h"
#include "element_tables.h"

namespace twoldg_detail {
int fail(int code, const std::string& msg);
// The library's device memory (host = false) and fine-grained host staging
// (host = true).  Freed blocks are kept, per device and size class, and handed
// out again: a virtual address range the library has used stays mapped to the
// same memory for the life of the process.  Giving ranges back to the driver
// and having new allocations map them to other pages let workgroups of a new
// problem's first stages read old contents of some pages (wrong values in a
// few blocks of the first RK4 step after a problem was created, never in a
// repeat of that step on the same problem).
hipError_t pool_alloc(void** ptr, size_t bytes, bool host);
void pool_free(void* ptr);
// A two-dimensional problem's own batch of one (batch.hip): whether a
// description is one the batch stage takes, and the batch on the problem's
// buffers and stream.
bool batch_desc_ok(const twoldg_desc* d, std::string* why);
int batch_attach(const twoldg_desc* d, void* yn, void* ys, void* acc, void* lo,
                 hipStream_t stream, twoldg_batch** out);
void batch_detach(twoldg_batch* B);
int batch_attached_steps(twoldg_batch* B, double dt, int64_t nsteps);
// The batch's working arrays (its face slots), as (address, bytes).
void batch_work(const twoldg_batch* B, std::vector<std::pair<void*, size_t>>& out);
}  // namespace twoldg_detail
using twoldg_detail::fail;
using twoldg_detail::pool_alloc;
using twoldg_detail::pool_free;

// The modes a modal indicator reads (regrid.h): per mode q its number m[q], the
// weight wt[q] and the power pw[q], and the factor post.  Outside the unnamed
// namespace because it crosses from the ABI unit to the per-type units.
struct twoldg_mode_args {
  int nq;
  int m[8];
  double wt[8], pw[8];
  double post;
};

namespace {

#define HIP_TRY(expr)                                                                       \
  do {                                                                                      \
    hipError_t e_ = (expr);                                                                 \
    if (e_ != hipSuccess)                                                                   \
      return fail(TWOLDG_EDEVICE, std::string(#expr) + ": " + hipGetErrorString(e_));       \
  } while (0)

constexpr int kMaxD = 3;

// ---------------------------------------------------------------------------
// Device-side tables of one element order, in the problem's precision.
// ---------------------------------------------------------------------------
template <class T>
struct Tables {
  T W[9][11];
  T mode[9];
  T dual[9];
  double wq[9];  // quadrature weights on [-1,1], FP64 for the reductions
};

// Everything a kernel needs to know about the problem's shape.
struct Shape {
  int N;        // nodes per block per dimension
  int64_t nn;   // N^D
  int64_t Nt;   // N^(D-1), nodes per face
  int64_t nb;   // blocks
};

template <int D>
__device__ __forceinline__ void unravel(int64_t l, int N, int I[kMaxD]) {
  I[0] = (int)(l % N);
  if (D >= 2) I[1] = (int)((l / N) % N);
  if (D >= 3) I[2] = (int)(l / ((int64_t)N * N));
}

// A node index inside a block (below 2^31): 32-bit divisions.
template <int D>
__device__ __forceinline__ void unravel(int l, int N, int I[kMaxD]) {
  const unsigned ul = (unsigned)l, uN = (unsigned)N;
  I[0] = (int)(ul % uN);
  if (D >= 2) I[1] = (int)((ul / uN) % uN);
  if (D >= 3) I[2] = (int)(ul / (uN * uN));
}

// Block-node launches: a 1D grid whose workgroup serves one chunk of 256 nodes
// of one block.  Workgroups go to the eight XCDs in turn (the workgroup index
// mod 8); the mapping gives every chunk of a block the same XCD and dispatches
// them close together (groups of 8 blocks x their chunks), so that the lines
// of a block, which each chunk reads across its planes in the wide stencils,
// come into one XCD's L2 once.  No thread divides a 64-bit global index.
constexpr int kXcds = 8;
inline dim3 block_node_grid(const Shape& sh) {
  const int64_t nc = (sh.nn + 255) / 256;
  const int64_t groups = (sh.nb + kXcds - 1) / kXcds;
  return dim3((unsigned)(groups * kXcds * nc));
}
// This workgroup's block b and this thread's node l; false past the last block
// or node.  256 threads per workgroup.
__device__ __forceinline__ bool block_node(const Shape& sh, int64_t* b, int* l) {
  const unsigned nc = (unsigned)((sh.nn + 255) / 256);
  const unsigned gsz = kXcds * nc;
  const unsigned w = blockIdx.x, g = w / gsz, r = w % gsz;
  *b = (int64_t)g * kXcds + r % kXcds;
  *l = (int)((r / kXcds) * blockDim.x + threadIdx.x);
  return *b < sh.nb && *l < sh.nn;
}

template <int D>
__device__ __forceinline__ int64_t ravel(const int I[kMaxD], int N) {
  int64_t l = I[0];
  if (D >= 2) l += (int64_t)N * I[1];
  if (D >= 3) l += (int64_t)N * N * I[2];
  return l;
}

// Tangential linear index of node I on a face normal to d (other dims in order).
template <int D>
__device__ __forceinline__ int64_t tlin(const int I[kMaxD], int d, int N) {
  int64_t t = 0, s = 1;
#pragma unroll
  for (int e = 0; e < D; ++e)
    if (e != d) {
      t += s * I[e];
      s *= N;
    }
  return t;
}

// ---------------------------------------------------------------------------
// Exchange: ghost rows (b, d, s, t) <- sum_k coef_k * src(block_k, node_k).
// MODE 0: u from the state (stride 2 nn per block)
// MODE 1: flat gradient g_d from grads (nb, D, nn), component d of the row
// MODE 2: Cartesian gradient, all D components, into (rows/Nt, D, Nt), from the
//         face-compact gradient (nb, 2D, D, Nt) the gradient pass writes at the
//         face nodes; cf gives each entry's face and tangential position there
// ---------------------------------------------------------------------------
struct ExchangeRows {
  int64_t nrows;
  const int32_t* rowptr;
  const int32_t* cb;
  const int32_t* cl;
  const int32_t* cf;  // MODE 2: ((block 2D + face) D) Nt + tangential index
};
template <class T, int MODE, int D>
__device__ __forceinline__ void exchange_row(int64_t r, const Shape& sh, const ExchangeRows& X,
                                             const T* __restrict__ coef,
                                             const T* __restrict__ src, T* __restrict__ ghost) {
  const int32_t* __restrict__ cb = X.cb;
  const int32_t* __restrict__ cl = X.cl;
  const int k0 = X.rowptr[r], k1 = X.rowptr[r + 1];
  if (k0 == k1) return;  // a Dirichlet row: written by dirichlet_row
  if (MODE == 0) {
    T acc = 0;
    for (int k = k0; k < k1; ++k) acc += coef[k] * src[(int64_t)cb[k] * 2 * sh.nn + cl[k]];
    ghost[r] = acc;
  } else if (MODE == 1) {
    const int d = (int)((r / (2 * sh.Nt)) % D);
    T acc = 0;
    for (int k = k0; k < k1; ++k)
      acc += coef[k] * src[((int64_t)cb[k] * D + d) * sh.nn + cl[k]];
    ghost[r] = acc;
  } else {
    const int64_t face = r / sh.Nt, t = r % sh.Nt;
    const int32_t* __restrict__ cf = X.cf;
#pragma unroll
    for (int i = 0; i < D; ++i) {
      T acc = 0;
      for (int k = k0; k < k1; ++k) acc += coef[k] * src[(int64_t)cf[k] + (int64_t)i * sh.Nt];
      ghost[(face * D + i) * sh.Nt + t] = acc;
    }
  }
}
template <class T, int MODE, int D>
__global__ void k_exchange(Shape sh, ExchangeRows X, const T* __restrict__ coef,
                           const T* __restrict__ src, T* __restrict__ ghost) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r < X.nrows) exchange_row<T, MODE, D>(r, sh, X, coef, src, ghost);
}

// Dirichlet ghosts of a plane wave sin(k.x - w t + phi): u (MODE 0) or the
// Cartesian gradient k cos(...) (MODE 2).  rows: u-shaped ghost rows, whose
// exchange rows are empty.  The time is read from the problem's device clock
// (t0, s): t = (t0 + s dt) + cdt, cdt the stage's offset, so that a captured
// RK4 step replays at every step's own time.
struct DirichletRows {
  int64_t nrows;
  const int64_t* rows;
  const double* x;  // (nrows, D) node positions
  double k0, k1, k2, w, phi;
  const double* clock;  // device: (t0, s)
  double dt;
};
template <class T, int MODE, int D>
__device__ __forceinline__ void dirichlet_row(int64_t q, const Shape& sh, const DirichletRows& R,
                                              double t, T* __restrict__ ghost) {
  const double k0 = R.k0, k1 = R.k1, k2 = R.k2, w = R.w, phi = R.phi;
  const int64_t r = R.rows[q];
  const double* xp = R.x + q * D;
  double sx = k0 * xp[0];
  if (D >= 2) sx += k1 * xp[1];
  if (D >= 3) sx += k2 * xp[2];
  const double th = sx - w * t + phi;
  if (MODE == 0) {
    ghost[r] = (T)sin(th);
  } else {
    const double c = cos(th);
    const int64_t face = r / sh.Nt, tt = r % sh.Nt;
    const double kk[3] = {k0, k1, k2};
#pragma unroll
    for (int i = 0; i < D; ++i) ghost[(face * D + i) * sh.Nt + tt] = (T)(kk[i] * c);
  }
}
// The exchange rows and the Dirichlet rows (disjoint ghosts) in one launch.
template <class T, int MODE, int D>
__global__ void k_exchange_dirichlet(Shape sh, ExchangeRows X, const T* __restrict__ coef,
                                     const T* __restrict__ src, DirichletRows R, double cdt,
                                     T* __restrict__ ghost) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r < X.nrows) {
    exchange_row<T, MODE, D>(r, sh, X, coef, src, ghost);
  } else if (r - X.nrows < R.nrows) {
    // rounded as the host's t0 + s dt, then + cdt (no fused multiply-add)
    const double t = __dadd_rn(__dadd_rn(R.clock[0], __dmul_rn(R.clock[1], R.dt)), cdt);
    dirichlet_row<T, MODE, D>(r - X.nrows, sh, R, t, ghost);
  }
}

// The device clock of a problem: (t0, step count).
__global__ void k_set_clock(double* __restrict__ clock, double t0) {
  clock[0] = t0;
  clock[1] = 0.0;
}
__global__ void k_tick(double* __restrict__ clock) { clock[1] += 1.0; }

// A load of a table the host wrote when the problem was built (neighbour codes,
// spacings, exchange rows and weights): relaxed, at agent scope, so that it
// observes the value last written at device level whatever a cache of the
// workgroup's XCD still holds for that address.  The tables live in recycled
// device memory; plain loads of them have been seen to return the previous
// occupant's contents in a few workgroups of a new problem's first stages.
template <class T>
__device__ __forceinline__ T ld_table(const T* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// The wide derivative of `read` along d at owned node I, in units of 1/h
// (evolution.jl:25-35): window from one node before the element to one after.
template <class T, int D, int P1, class Read>
__device__ __forceinline__ T wide(const Tables<T>& tb, const int I[kMaxD], int d, Read read) {
  const int i = I[d] % P1;
  const int start = I[d] - i - 1;
  T acc = 0;
#pragma unroll
  for (int k = 0; k < P1 + 2; ++k) {
    const T w = tb.W[i][k];
    // rows other than the face rows have zeros in the outer columns; skip them
    if ((k == 0 && i != 0) || (k == P1 + 1 && i != P1 - 1)) continue;
    acc += w * read(start + k);
  }
  return acc;
}

// ---------------------------------------------------------------------------
// Gradient pass (flat): grads[b][d][l] = (W u)_d / h.
// ---------------------------------------------------------------------------
template <class T, int D, int P1>
__global__ void k_gradient(Shape sh, const Tables<T>* __restrict__ tbp,
                           const double* __restrict__ spacing, const T* __restrict__ y,
                           const T* __restrict__ ug, const T* __restrict__ dcoef,
                           T* __restrict__ grads) {
  int64_t b;
  int l;
  if (!block_node(sh, &b, &l)) return;
  const Tables<T>& tb = *tbp;
  const int N = sh.N;
  int I[kMaxD] = {0, 0, 0};
  unravel<D>(l, N, I);
  const T h = (T)spacing[b];
  const T* u = y + b * 2 * sh.nn;
#pragma unroll
  for (int d = 0; d < D; ++d) {
    const int stride = d == 0 ? 1 : (d == 1 ? N : N * N);
    const int base = l - I[d] * stride;
    const int tl = (int)tlin<D>(I, d, N);
    auto read = [&](int o) -> T {
      if (o < 0) return ug[((b * D + d) * 2 + 0) * sh.Nt + tl];
      if (o >= N) return ug[((b * D + d) * 2 + 1) * sh.Nt + tl];
      return u[base + o * stride];
    };
    const T g = wide<T, D, P1>(tb, I, d, read) / h;
    // on a diagonal metric the flux A_d D_d u (evolution.jl:67-70)
    grads[(b * D + d) * sh.nn + l] = dcoef ? dcoef[(b * (1 + D) + 1 + d) * sh.nn + l] * g : g;
  }
}

// Cartesian gradient (full metric): V_i = sum_e S[e][i] D_e u / sqrt(g), or on a
// flat block S[i][i] D_i u / J from the block's first node (evolution.jl:144-164).
// FACES (the stages): V at the face nodes into the face-compact cart (nb, 2D, D,
// Nt), which the exchange reads, and the fluxes; otherwise (the energy) V at every
// node into cart (nb, D, nn) and no fluxes.
template <class T, int D, int P1, bool FACES>
__global__ void k_gradient_cart(Shape sh, const Tables<T>* __restrict__ tbp,
                                const double* __restrict__ spacing, const T* __restrict__ y,
                                const T* __restrict__ ug, const T* __restrict__ coef,
                                const uint8_t* __restrict__ flat, T* __restrict__ cart,
                                T* __restrict__ flux) {
  int64_t b;
  int l;
  if (!block_node(sh, &b, &l)) return;
  const Tables<T>& tb = *tbp;
  const int N = sh.N;
  int I[kMaxD] = {0, 0, 0};
  unravel<D>(l, N, I);
  const T h = (T)spacing[b];
  const T* u = y + b * 2 * sh.nn;
  T Du[kMaxD];
#pragma unroll
  for (int d = 0; d < D; ++d) {
    const int stride = d == 0 ? 1 : (d == 1 ? N : N * N);
    const int base = l - I[d] * stride;
    const int tl = (int)tlin<D>(I, d, N);
    auto read = [&](int o) -> T {
      if (o < 0) return ug[((b * D + d) * 2 + 0) * sh.Nt + tl];
      if (o >= N) return ug[((b * D + d) * 2 + 1) * sh.Nt + tl];
      return u[base + o * stride];
    };
    Du[d] = wide<T, D, P1>(tb, I, d, read) / h;
  }
  const int nv = 1 + D * D;
  const T* c = coef + b * nv * sh.nn;
  // the Cartesian gradient V, and each dimension's flux F_d = S[d] . V at the
  // node, which the stage's wide derivative along d reads (evolution.jl:136-220)
  T V[kMaxD];
  if (flat[b]) {
    const T J = c[0];
#pragma unroll
    for (int i = 0; i < D; ++i) V[i] = c[(1 + i * D + i) * sh.nn] * Du[i] / J;
    if (FACES) {
#pragma unroll
      for (int d = 0; d < D; ++d)
        flux[(b * D + d) * sh.nn + l] = c[(1 + d * D + d) * sh.nn] * V[d];
    }
  } else {
#pragma unroll
    for (int i = 0; i < D; ++i) {
      T s = c[(1 + 0 * D + i) * sh.nn + l] * Du[0];
#pragma unroll
      for (int e = 1; e < D; ++e) s += c[(1 + e * D + i) * sh.nn + l] * Du[e];
      V[i] = s / c[l];
    }
    if (FACES) {
#pragma unroll
      for (int d = 0; d < D; ++d) {
        T s = c[(1 + d * D + 0) * sh.nn + l] * V[0];
#pragma unroll
        for (int j = 1; j < D; ++j) s += c[(1 + d * D + j) * sh.nn + l] * V[j];
        flux[(b * D + d) * sh.nn + l] = s;
      }
    }
  }
  if (FACES) {
    // a node on face (d, s) writes V at its tangential position there (a node
    // on an edge or corner, to each of its faces)
#pragma unroll
    for (int d = 0; d < D; ++d)
#pragma unroll
      for (int s = 0; s < 2; ++s)
        if (I[d] == (s ? N - 1 : 0)) {
          const int64_t base = ((b * 2 * D + 2 * d + s) * D) * sh.Nt + tlin<D>(I, d, N);
#pragma unroll
          for (int i = 0; i < D; ++i) cart[base + i * sh.Nt] = V[i];
        }
  } else {
#pragma unroll
    for (int i = 0; i < D; ++i) cart[(b * D + i) * sh.nn + l] = V[i];
  }
}

// ---------------------------------------------------------------------------
// Stage kernels: the Laplacian and the RK4 combination in one pass.
// k = (v_s, lap); by stage:
//   MODE 0 (plain RHS):  du = k
// FP32c keeps three registers:
//   MODE 1 (stage 1):    acc = k;          yo = yn + a k
//   MODE 2 (stages 2,3): acc = acc + 2 k;  yo = yn + a k
//   MODE 4 (stage 4, compensated): yn = yn + a (acc + k) (a = dt/6) as a Kahan
//                        sum into (yn, lo)
// FP64 keeps no accumulator (the k's are (y_2 - y_n) 2/dt and (y_3 - y_n) 2/dt):
//   MODE 5 (stages 1, 2): yo = yn + a k
//   MODE 7 (stage 3):    yo = yn + a k (y_4, over y_3 = ys); acc (y_2 in) takes
//                        z = (y_2 + 2 y_3)/3 + a2 k = y_n + dt/6 (k1 + 2 k2 + 2 k3)
//   MODE 8 (stage 4):    yo = acc (z) + a k              (a = dt/6)
// ---------------------------------------------------------------------------
template <class T>
struct StageIO {
  const T* ys;   // stage input state (its v is k_u)
  const T* yn;   // step start
  T* acc;
  T* yo;         // output
  T* lo;         // compensation of yn (MODE 4)
  T a;
  T a2;          // MODE 7: dt/3
};

template <class T, int MODE>
__device__ __forceinline__ void stage_store(const StageIO<T>& io, int64_t iu, int64_t iv, T ku,
                                            T kv) {
  if (MODE == 0) {
    io.yo[iu] = ku;
    io.yo[iv] = kv;
  } else if (MODE == 1) {
    io.acc[iu] = ku;
    io.acc[iv] = kv;
    io.yo[iu] = io.yn[iu] + io.a * ku;
    io.yo[iv] = io.yn[iv] + io.a * kv;
  } else if (MODE == 2) {
    io.acc[iu] = io.acc[iu] + T(2) * ku;
    io.acc[iv] = io.acc[iv] + T(2) * kv;
    io.yo[iu] = io.yn[iu] + io.a * ku;
    io.yo[iv] = io.yn[iv] + io.a * kv;
  } else if (MODE == 5) {
    io.yo[iu] = io.yn[iu] + io.a * ku;
    io.yo[iv] = io.yn[iv] + io.a * kv;
  } else if (MODE == 7) {
    constexpr T third = T(1) / T(3);
    const T su = io.ys[iu];  // y_3's u (its v is ku)
    const T zu = (io.acc[iu] + T(2) * su) * third + io.a2 * ku;
    const T zv = (io.acc[iv] + T(2) * ku) * third + io.a2 * kv;
    io.yo[iu] = io.yn[iu] + io.a * ku;
    io.yo[iv] = io.yn[iv] + io.a * kv;
    io.acc[iu] = zu;
    io.acc[iv] = zv;
  } else if (MODE == 8) {
    io.yo[iu] = io.acc[iu] + io.a * ku;
    io.yo[iv] = io.acc[iv] + io.a * kv;
  } else {
    // Kahan: (hi, c) += incr, c holds minus the lost low part
    const T incr_u = io.a * (io.acc[iu] + ku), incr_v = io.a * (io.acc[iv] + kv);
    T yu = incr_u - io.lo[iu], hu = io.yn[iu], tu = hu + yu;
    io.lo[iu] = (tu - hu) - yu;
    io.yo[iu] = tu;
    T yv = incr_v - io.lo[iv], hv = io.yn[iv], tv = hv + yv;
    io.lo[iv] = (tv - hv) - yv;
    io.yo[iv] = tv;
  }
}

template <class T, int D, int P1, int MODE>
__global__ void k_stage_flat(Shape sh, const Tables<T>* __restrict__ tbp,
                             const double* __restrict__ spacing, const T* __restrict__ grads,
                             const T* __restrict__ gg, const T* __restrict__ dcoef,
                             StageIO<T> io) {
  int64_t b;
  int l;
  if (!block_node(sh, &b, &l)) return;
  const Tables<T>& tb = *tbp;
  const int N = sh.N;
  int I[kMaxD] = {0, 0, 0};
  unravel<D>(l, N, I);
  T acc = 0;
#pragma unroll
  for (int d = 0; d < D; ++d) {
    const int stride = d == 0 ? 1 : (d == 1 ? N : N * N);
    const int base = l - I[d] * stride;
    const int tl = (int)tlin<D>(I, d, N);
    const T* gd = grads + (b * D + d) * sh.nn;
    auto read = [&](int o) -> T {
      if (o < 0) return gg[((b * D + d) * 2 + 0) * sh.Nt + tl];
      if (o >= N) return gg[((b * D + d) * 2 + 1) * sh.Nt + tl];
      return gd[base + o * stride];
    };
    acc += wide<T, D, P1>(tb, I, d, read);
  }
  const int64_t iu = b * 2 * sh.nn + l, iv = iu + sh.nn;
  const T ku = io.ys[iv];
  const T lap = acc / (T)spacing[b];
  // on a diagonal metric the divergence over sqrt(g) (evolution.jl:69-70)
  const T kv = dcoef ? lap / dcoef[b * (1 + D) * sh.nn + l] : lap;
  stage_store<T, MODE>(io, iu, iv, ku, kv);
}

// Laplacian on a full metric: sqrt(g) dv/dt = sum_d D_d (S[d] . V), the frame at
// a ghost being the adjacent face node's (evolution.jl:169-220).
template <class T, int D, int P1, int MODE>
__global__ void k_stage_cart(Shape sh, const Tables<T>* __restrict__ tbp,
                             const double* __restrict__ spacing, const T* __restrict__ flux,
                             const T* __restrict__ cg, const T* __restrict__ coef,
                             const uint8_t* __restrict__ flat, StageIO<T> io) {
  int64_t b;
  int l;
  if (!block_node(sh, &b, &l)) return;
  const Tables<T>& tb = *tbp;
  const int N = sh.N;
  int I[kMaxD] = {0, 0, 0};
  unravel<D>(l, N, I);
  const int nv = 1 + D * D;
  const T* c = coef + b * nv * sh.nn;
  const bool isflat = flat[b] != 0;
  T acc = 0;
#pragma unroll
  for (int d = 0; d < D; ++d) {
    const int stride = d == 0 ? 1 : (d == 1 ? N : N * N);
    const int base = l - I[d] * stride;
    const int tl = (int)tlin<D>(I, d, N);
    const T Sdd = c[(1 + d * D + d) * sh.nn];
    auto read = [&](int o) -> T {
      // inside the block the flux the gradient pass formed at the node; across
      // a face the ghost's flux with the face node's frame
      if (o >= 0 && o < N) return flux[(b * D + d) * sh.nn + base + o * stride];
      const int lf = base + (o < 0 ? 0 : N - 1) * stride;  // the face node
      auto V = [&](int j) -> T {
        return cg[((((b * D + d) * 2 + (o < 0 ? 0 : 1)) * D) + j) * sh.Nt + tl];
      };
      if (isflat) return Sdd * V(d);
      T s = c[(1 + d * D + 0) * sh.nn + lf] * V(0);
#pragma unroll
      for (int j = 1; j < D; ++j) s += c[(1 + d * D + j) * sh.nn + lf] * V(j);
      return s;
    };
    acc += wide<T, D, P1>(tb, I, d, read);
  }
  const int64_t iu = b * 2 * sh.nn + l, iv = iu + sh.nn;
  const T ku = io.ys[iv];
  const T kv = acc / (T)spacing[b] / c[l];
  stage_store<T, MODE>(io, iu, iv, ku, kv);
}

// ---------------------------------------------------------------------------
// Truncation: per element, for d = 1..D and every line along d,
// c = dual . a, a -= mode c (evolution.jl:238-264).  v of the state.
// One workgroup per block, a thread per line of the current direction (the
// block's lines along d, all elements), a barrier between directions.
// ---------------------------------------------------------------------------
constexpr int kTruncThreads = 256;

template <class T, int D, int P1>
__global__ void __launch_bounds__(kTruncThreads)
k_truncate(Shape sh, int E, const Tables<T>* __restrict__ tbp,
           const int32_t* __restrict__ tslot, const T* __restrict__ lmode,
           const T* __restrict__ ldual, T* __restrict__ y) {
  const int64_t b = blockIdx.x;
  if (tslot && tslot[b] >= 0) return;  // per-element tables (k_truncate_elem)
  const Tables<T>& tb = *tbp;
  const int N = sh.N;
  T* a = y + b * 2 * sh.nn + sh.nn;
  const int64_t strides[kMaxD] = {1, N, (int64_t)N * N};
  const int64_t nlines = sh.nn / P1;  // lines along one direction: N^(D-1) E
#pragma unroll
  for (int d = 0; d < D; ++d) {
    for (int64_t l = threadIdx.x; l < nlines; l += blockDim.x) {
      // line l: element index along d (l % E), then the tangential node indices
      const int ed = (int)(l % E);
      int64_t rest = l / E, base = (int64_t)ed * P1 * strides[d];
#pragma unroll
      for (int q = 0; q < D; ++q)
        if (q != d) {
          base += (rest % N) * strides[q];
          rest /= N;
        }
      T* line = a + base;
      T vals[P1];
      T c = 0;
      if (lmode) {  // the line's own pair on a diagonal metric (evolution.jl:556-575)
        const int64_t t = ((b * D + d) * E + ed) * P1;
#pragma unroll
        for (int j = 0; j < P1; ++j) {
          vals[j] = line[j * strides[d]];
          c += ldual[t + j] * vals[j];
        }
#pragma unroll
        for (int i = 0; i < P1; ++i) line[i * strides[d]] = vals[i] - lmode[t + i] * c;
        continue;
      }
#pragma unroll
      for (int j = 0; j < P1; ++j) {
        vals[j] = line[j * strides[d]];
        c += tb.dual[j] * vals[j];
      }
#pragma unroll
      for (int i = 0; i < P1; ++i) line[i * strides[d]] = vals[i] - tb.mode[i] * c;
    }
    if (d + 1 < D) __syncthreads();
  }
}

// A workgroup barrier that orders LDS only.  __syncthreads() also fences global
// memory, which makes every wave wait for its outstanding global loads at each
// barrier and so serializes the plane's loads with the LDS phases; the kernels
// that use it share nothing through global memory inside a launch.
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Per-element sqrt(g)-orthogonal projection (evolution.jl:384-425):
// c = Psi^T (W_e a); z = G^-1 c with G = Psi^T W_e Psi = U^T U (the element's
// Gram matrix, whose Cholesky factor U the caller supplies; its inverse is
// formed once per problem on the host, so the two triangular solves become one
// mat-vec); a = Psi z.  D = 3.  One workgroup per element, a thread per node
// (and per mode), three LDS barriers.  The elements of blocks without tables
// (tslot < 0) get the tensor projector here too, direction by direction with
// k_truncate's arithmetic, so one launch covers the forest.
// The factored truncation's tables (device): per block its factored index f or
// -1; per f the radial dimension, the radial pairs (f, e_k, i) and the angular
// table A; the plane basis Psi ((P+1)^2, P^2); per (A, element across k) the
// plane's weights and the inverse of its Gram matrix Psi^T W Psi.
// One factored block of a truncation launch: the block, its table index, its
// radial dimension and angular table, in one load.
struct FBlock {
  int32_t b, f, k, A;
};
template <class T>
struct FactoredTabs {
  const FBlock* blk;  // the factored blocks (a launch of the factored kind)
  const int32_t* oblk;  // the other blocks (the other kind), or null: every block
  const int32_t* slot;
  const int32_t* radial;
  const T* rmodes;
  const T* rduals;
  const int32_t* aslot;
  const T* aproj;  // per (angular table, element position): Psi G^-1 Psi^T W on a plane
};

// EPW elements of one block per workgroup, each on its own (P+1)^3 threads
// rounded up to whole waves, so that the many small elements do not each cost a
// workgroup dispatch; all elements of a workgroup take their block's branch.
template <int P1>
constexpr int trunc_elem_threads() { return (P1 * P1 * P1 + 63) / 64 * 64; }
template <int P1>
constexpr int trunc_elem_epw() { return trunc_elem_threads<P1>() <= 64 ? 4 : 2; }

// KIND 1 covers the factored blocks (their list F.blk), KIND 0 the others (the
// list F.oblk, or every block); the two are launched apart, so that the
// factored kind's workgroups hold only the small angular tables in LDS.
template <class T, int P1, int KIND>
__global__ void __launch_bounds__(trunc_elem_threads<P1>() * trunc_elem_epw<P1>())
k_truncate_elem(Shape sh, int E, const Tables<T>* __restrict__ tbp,
                const int32_t* __restrict__ tslot, const T* __restrict__ tmodes,
                const T* __restrict__ tweights, const T* __restrict__ tginv,
                FactoredTabs<T> F, T* __restrict__ y) {
  constexpr int P = P1 - 1;
  constexpr int n = P1 * P1 * P1, m = P * P * P;
  constexpr int NTE = trunc_elem_threads<P1>(), EPW = trunc_elem_epw<P1>();
  constexpr int GS = KIND == 1 ? (P1 * P1) * (P1 * P1) : m * m;
  constexpr int PS = KIND == 1 ? 1 : n * m;
  __shared__ T sWA_[EPW][n];
  __shared__ T sc_[EPW][P1 * P * P];  // m <= (P+1) P^2: the element's, or the planes' coefficients
  __shared__ T sz_[EPW][P1 * P * P];
  __shared__ T sG_[EPW][GS];  // G^-1, column-major (symmetric)
  // the Legendre basis Psi (the element's, or a plane's) in LDS where it fits,
  // shared by the workgroup's elements
  constexpr bool psi_lds = (size_t)PS * sizeof(T) <= 16384;
  __shared__ T sPsi[psi_lds ? PS : 1];
  const int q0 = threadIdx.x / NTE;  // this thread's element of the workgroup
  T* sWA = sWA_[q0];
  T* sc = sc_[q0];
  T* sz = sz_[q0];
  T* sG = sG_[q0];
  const int64_t ne = (int64_t)E * E * E, per = ne / EPW;
  FBlock fb{0, -1, 0, 0};
  if (KIND == 1) fb = F.blk[blockIdx.x / per];
  const int64_t b = KIND == 1 ? (int64_t)fb.b : (F.oblk ? (int64_t)F.oblk[blockIdx.x / per]
                                                        : (int64_t)(blockIdx.x / per));
  const int64_t e = (blockIdx.x % per) * EPW + q0;
  const int s = (KIND == 0 && tslot) ? tslot[b] : -1;
  const int k = threadIdx.x % NTE;
  const int N = sh.N;
  const int J0 = (int)(e % E), J1 = (int)((e / E) % E), J2 = (int)(e / ((int64_t)E * E));
  T* a = y + b * 2 * sh.nn + sh.nn;
  const int64_t base = (int64_t)J0 * P1 + (int64_t)N * (J1 * P1) + (int64_t)N * N * (J2 * P1);
  auto at = [&](int q) -> int64_t {
    const int k0 = q % P1, k1 = (q / P1) % P1, k2 = q / (P1 * P1);
    return base + k0 + (int64_t)N * k1 + (int64_t)N * N * k2;
  };
  if constexpr (KIND == 1) {
    // evolution.jl:306-380: along k, a <- a - mode (dual . a) on every line;
    // then on each plane across k, a <- M a with M = Psi G^-1 Psi^T W, the
    // plane's projector formed on the host for the element's angular table
    constexpr int n2 = P1 * P1;
    const int fs = fb.f, kr = fb.k, A = fb.A;
    const int t0 = kr == 0 ? 1 : 0, t1 = kr == 2 ? 1 : 2;
    const int J[3] = {J0, J1, J2};
    const int64_t ea = J[t0] + (int64_t)E * J[t1];
    const int st[3] = {1, P1, P1 * P1};
    const T* rm = F.rmodes + ((int64_t)fs * E + J[kr]) * P1;
    const T* rd = F.rduals + ((int64_t)fs * E + J[kr]) * P1;
    const T* M = F.aproj + ((int64_t)A * E * E + ea) * n2 * n2;
    for (int i = k; i < n2 * n2; i += NTE) sG[i] = M[i];
    int kk[3] = {0, 0, 0};
    if (k < n) {
      kk[0] = k % P1;
      kk[1] = (k / P1) % P1;
      kk[2] = k / (P1 * P1);
      sWA[k] = a[at(k)];
    }
    __syncthreads();
    T v = 0;
    if (k < n) {
      const int l0 = k - kk[kr] * st[kr];
      T c = 0;
#pragma unroll
      for (int j = 0; j < P1; ++j) c += rd[j] * sWA[l0 + j * st[kr]];
      v = sWA[k] - rm[kk[kr]] * c;
    }
    lds_barrier();
    if (k < n) sWA[k] = v;
    lds_barrier();
    if (k < n) {
      const int pn = kk[t0] + P1 * kk[t1];  // the node's index in its plane
      const int plane = kk[kr] * st[kr];
      T z = 0;
#pragma unroll
      for (int r = 0; r < n2; ++r)
        z += sG[pn * n2 + r] * sWA[plane + (r % P1) * st[t0] + (r / P1) * st[t1]];
      a[at(k)] = z;
    }
    return;
  } else {
  if (s < 0) {  // the tensor projector (evolution.jl:238-264), x, y, then z
    const Tables<T>& tb = *tbp;
    if (k < n) sWA[k] = a[at(k)];
    T v = 0;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      lds_barrier();
      if (k < n) {
        const int st = d == 0 ? 1 : (d == 1 ? P1 : P1 * P1);
        const int i = (k / st) % P1;
        const int l0 = k - i * st;
        T c = 0;
#pragma unroll
        for (int j = 0; j < P1; ++j) c += tb.dual[j] * sWA[l0 + j * st];
        v = sWA[k] - tb.mode[i] * c;
      }
      lds_barrier();
      if (k < n) sWA[k] = v;
    }
    if (k < n) a[at(k)] = v;
    return;
  }
  const T* w = tweights + ((int64_t)s * ne + e) * n;
  const T* G = tginv + ((int64_t)s * ne + e) * m * m;
  for (int i = k; i < m * m; i += NTE) sG[i] = G[i];
  if (psi_lds)
    for (int i = threadIdx.x; i < n * m; i += blockDim.x) sPsi[i] = tmodes[i];
  const T* Psi = psi_lds ? sPsi : tmodes;
  if (k < n) sWA[k] = w[k] * a[at(k)];
  __syncthreads();
  if (k < m) {
    T cq = 0;
    for (int i = 0; i < n; ++i) cq += Psi[i * m + k] * sWA[i];
    sc[k] = cq;
  }
  lds_barrier();
  if (k < m) {
    T zq = 0;
    for (int q = 0; q < m; ++q) zq += sG[k + m * q] * sc[q];
    sz[k] = zq;
  }
  lds_barrier();
  if (k < n) {
    T z = 0;
    for (int q = 0; q < m; ++q) z += Psi[k * m + q] * sz[q];
    a[at(k)] = z;
  }
  }
}

// ---------------------------------------------------------------------------
// Reductions (FP64).  One workgroup per block writes its partial sums; one
// workgroup then adds the partials in a fixed order.
// ---------------------------------------------------------------------------
constexpr int kRedThreads = 256;

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  return v;
}
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_down(v, o, 64));
  return v;
}

// Sum (and max) over the workgroup; thread 0 gets the result.
template <int NV>
__device__ void group_reduce(double* v, bool maxlast) {
  __shared__ double part[NV][kRedThreads / 64];
  const int lane = threadIdx.x % 64, wid = threadIdx.x / 64;
#pragma unroll
  for (int q = 0; q < NV; ++q) {
    double x = (maxlast && q == NV - 1) ? wave_max(v[q]) : wave_sum(v[q]);
    if (lane == 0) part[q][wid] = x;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int q = 0; q < NV; ++q) {
      double x = part[q][0];
      for (int w = 1; w < kRedThreads / 64; ++w)
        x = (maxlast && q == NV - 1) ? fmax(x, part[q][w]) : x + part[q][w];
      v[q] = x;
    }
  }
}

template <int D>
__device__ __forceinline__ double node_weight(const int I[kMaxD], double h, int P1,
                                              const double* wq) {
  double w = 1;
#pragma unroll
  for (int d = 0; d < D; ++d) w *= h * P1 / 2 * wq[I[d] % P1];
  return w;
}

// Energy partials per block: (kinetic, potential) (evolution.jl:1472-1489).
// flat: sum w v^2 and sum_d w g_d^2; diagonal metric (dcoef): w sqrt(g) v^2 and
// w F_d^2 / A_d; full metric: weights times sqrt(g), |V|^2.
template <class T, int D, int P1, bool FULL>
__global__ void k_energy_part(Shape sh, const Tables<T>* __restrict__ tbp,
                              const double* __restrict__ spacing, const T* __restrict__ y,
                              const T* __restrict__ grads, const T* __restrict__ coef,
                              const T* __restrict__ dcoef, double* __restrict__ part) {
  const int64_t b = blockIdx.x;
  const double h = spacing[b];
  double v[2] = {0, 0};
  for (int64_t l = threadIdx.x; l < sh.nn; l += blockDim.x) {
    int I[kMaxD] = {0, 0, 0};
    unravel<D>(l, sh.N, I);
    double w = node_weight<D>(I, h, P1, tbp->wq);
    const double vv = (double)y[b * 2 * sh.nn + sh.nn + l];
    double pot = 0;
    if (dcoef) {
      const T* c = dcoef + b * (1 + D) * sh.nn;
#pragma unroll
      for (int d = 0; d < D; ++d) {
        const double g = (double)grads[(b * D + d) * sh.nn + l];
        pot += w * (1.0 / (double)c[(1 + d) * sh.nn + l]) * (g * g);
      }
      v[0] += w * (double)c[l] * (vv * vv);
      v[1] += pot;
      continue;
    }
#pragma unroll
    for (int d = 0; d < D; ++d) {
      const double g = (double)grads[(b * D + d) * sh.nn + l];
      pot += g * g;
    }
    if (FULL) w *= (double)coef[b * (1 + D * D) * sh.nn + l];
    v[0] += w * (vv * vv);
    v[1] += w * pot;
  }
  group_reduce<2>(v, false);
  if (threadIdx.x == 0) {
    part[2 * b] = v[0];
    part[2 * b + 1] = v[1];
  }
}

// Norm partials per block: (sum w sum_vars |x|^p, sum w nvars, max |x|).
template <class T, int D, int P1, bool FULL>
__global__ void k_norm_part(Shape sh, const Tables<T>* __restrict__ tbp,
                            const double* __restrict__ spacing, const T* __restrict__ a,
                            const T* __restrict__ alo, const T* __restrict__ bb,
                            const T* __restrict__ blo, const T* __restrict__ coef,
                            const T* __restrict__ dcoef, int mask, int pnorm,
                            double* __restrict__ part) {
  const int64_t b = blockIdx.x;
  const double h = spacing[b];
  double v[3] = {0, 0, 0};
  for (int64_t l = threadIdx.x; l < sh.nn; l += blockDim.x) {
    int I[kMaxD] = {0, 0, 0};
    unravel<D>(l, sh.N, I);
    double w = node_weight<D>(I, h, P1, tbp->wq);
    if (FULL) w *= (double)coef[b * (1 + D * D) * sh.nn + l];
    if (dcoef) w *= (double)dcoef[b * (1 + D) * sh.nn + l];  // sqrt(g)
    double s = 0;
    int nv = 0;
    for (int q = 0; q < 2; ++q) {
      if (!(mask & (1 << q))) continue;
      const int64_t i = b * 2 * sh.nn + q * sh.nn + l;
      double x = (double)a[i];
      if (alo) x -= (double)alo[i];
      if (bb) {
        double z = (double)bb[i];
        if (blo) z -= (double)blo[i];
        x -= z;
      }
      x = fabs(x);
      s += pnorm == 2 ? x * x : pow(x, (double)pnorm);
      v[2] = fmax(v[2], x);
      ++nv;
    }
    v[0] += w * s;
    v[1] += w * nv;
  }
  group_reduce<3>(v, true);
  if (threadIdx.x == 0) {
    part[3 * b] = v[0];
    part[3 * b + 1] = v[1];
    part[3 * b + 2] = v[2];
  }
}

// Fixed-order sum of per-block partials: out[q] = sum_b part[NV b + q]
// (max for q == NV-1 when maxlast).
template <int NV>
__global__ void k_final(int64_t nb, const double* __restrict__ part, bool maxlast,
                        double* __restrict__ out) {
  double v[NV];
#pragma unroll
  for (int q = 0; q < NV; ++q) v[q] = 0;
  for (int64_t b = threadIdx.x; b < nb; b += blockDim.x)
#pragma unroll
    for (int q = 0; q < NV; ++q)
      v[q] = (maxlast && q == NV - 1) ? fmax(v[q], part[NV * b + q]) : v[q] + part[NV * b + q];
  group_reduce<NV>(v, maxlast);
  if (threadIdx.x == 0)
#pragma unroll
    for (int q = 0; q < NV; ++q) out[q] = v[q];
}

// ---------------------------------------------------------------------------
// The radiative condition's gradient ghosts (evolution.jl:1236-1270): on each
// radiative row the ghost holds the face node's Cartesian gradient V (the
// exchange copies it); its component along S[d] is changed so that
// S[d] . V_g = 2 F* - F, F = S[d] . V the face node's flux and
// F* = -sigma |S[d]| (kappa v + invR u), sigma = +1 on a high face.  The
// Laplacian averages the two, so the face's flux is F*.
// ---------------------------------------------------------------------------
template <int D>
__device__ __forceinline__ int64_t face_node(int64_t face, int64_t t, int N, int* d_out,
                                             int* s_out) {
  const int s = (int)(face % 2), d = (int)((face / 2) % D);
  int I[kMaxD] = {0, 0, 0};
  I[d] = s ? N - 1 : 0;
  int64_t rest = t;
#pragma unroll
  for (int e = 0; e < D; ++e)
    if (e != d) {
      I[e] = (int)(rest % N);
      rest /= N;
    }
  *d_out = d;
  *s_out = s;
  return I[0] + (int64_t)N * (D >= 2 ? I[1] : 0) + (int64_t)N * N * (D >= 3 ? I[2] : 0);
}

template <class T, int D>
__global__ void k_radiative_grad(Shape sh, int64_t nrad, const int64_t* __restrict__ rows,
                                 const T* __restrict__ coef, const T* __restrict__ y,
                                 double kappa, double invR, T* __restrict__ ghost) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nrad) return;
  const int64_t r = rows[q];
  const int64_t face = r / sh.Nt, t = r % sh.Nt;
  const int64_t b = face / (2 * D);
  int d, s;
  const int64_t l = face_node<D>(face, t, sh.N, &d, &s);
  double S[kMaxD], V[kMaxD], S2 = 0, F = 0;
#pragma unroll
  for (int i = 0; i < D; ++i) {
    S[i] = (double)coef[(b * (1 + D * D) + 1 + d * D + i) * sh.nn + l];
    V[i] = (double)ghost[(face * D + i) * sh.Nt + t];
    S2 += S[i] * S[i];
    F += S[i] * V[i];
  }
  const double u = (double)y[b * 2 * sh.nn + l], v = (double)y[b * 2 * sh.nn + sh.nn + l];
  const double Fs = -(s ? 1.0 : -1.0) * sqrt(S2) * (kappa * v + invR * u);
  const double c = 2 * (Fs - F) / S2;
#pragma unroll
  for (int i = 0; i < D; ++i) ghost[(face * D + i) * sh.Nt + t] = (T)(V[i] + c * S[i]);
}

// The radiative boundary's own energy (evolution.jl:1276-1299): per radiative
// face, sum w |S[d]| u^2 over its nodes, w the face's quadrature weight; one
// workgroup per face (a face's rows are consecutive).
template <class T, int D, int P1>
__global__ void k_rad_energy(Shape sh, const Tables<T>* __restrict__ tbp,
                             const double* __restrict__ spacing, const int64_t* __restrict__ rows,
                             const T* __restrict__ coef, const T* __restrict__ y,
                             const T* __restrict__ ylo, double* __restrict__ part) {
  const int64_t face = rows[(int64_t)blockIdx.x * sh.Nt] / sh.Nt;
  const int64_t b = face / (2 * D);
  const double h = spacing[b];
  double acc[1] = {0};
  for (int64_t t = threadIdx.x; t < sh.Nt; t += blockDim.x) {
    int d, s;
    const int64_t l = face_node<D>(face, t, sh.N, &d, &s);
    double w = 1;
    int64_t rest = t;
#pragma unroll
    for (int e = 0; e < D - 1; ++e) {
      w *= h * P1 / 2 * tbp->wq[(rest % sh.N) % P1];
      rest /= sh.N;
    }
    double S2 = 0;
#pragma unroll
    for (int i = 0; i < D; ++i) {
      const double Si = (double)coef[(b * (1 + D * D) + 1 + d * D + i) * sh.nn + l];
      S2 += Si * Si;
    }
    double u = (double)y[b * 2 * sh.nn + l];
    if (ylo) u -= (double)ylo[b * 2 * sh.nn + l];
    acc[0] += w * sqrt(S2) * u * u;
  }
  group_reduce<1>(acc, false);
  if (threadIdx.x == 0) part[blockIdx.x] = acc[0];
}

// y (double, host layout) <-> device precision
template <class T>
__global__ void k_from_double(int64_t n, const double* __restrict__ src, T* __restrict__ hi,
                              T* __restrict__ lo) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double x = src[i];
  const T h = (T)x;
  hi[i] = h;
  if (lo) lo[i] = (T)((double)h - x);  // minus the low part, as the Kahan sum keeps it
}

template <class T>
__global__ void k_to_double(int64_t n, const T* __restrict__ hi, const T* __restrict__ lo,
                            double* __restrict__ dst) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double x = (double)hi[i];
  if (lo) x -= (double)lo[i];
  dst[i] = x;
}

// The compensated value hi - lo rounded to T (a copy out of the compensated state).
template <class T>
__global__ void k_compensated_copy(int64_t n, const T* __restrict__ hi, const T* __restrict__ lo,
                                   T* __restrict__ dst) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  dst[i] = (T)((double)hi[i] - (double)lo[i]);
}

#include "fused3.h"
#include "regrid.h"

}  // namespace

// ===========================================================================
// Host side.
// ===========================================================================
struct twoldg_problem {
  int D = 0, P = 0, E = 0, N = 0, P1 = 0;
  int precision = TWOLDG_F64;
  bool truncate = true;
  bool full = false;
  Shape sh{};
  int64_t state_len = 0;
  int64_t nrows = 0;
  size_t elt = 8;
  int device = 0;  // the device current at create: the stream's and every buffer's
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  // device buffers (void*: element type follows precision)
  void* slot[TWOLDG_NSLOTS] = {nullptr, nullptr, nullptr, nullptr};  // STATE at create, others on first use
  void* lo = nullptr;       // compensation of the state (F32C)
  void* ys = nullptr;       // stage state
  void* acc = nullptr;      // RK accumulator
  void* ug = nullptr;       // u ghosts (rows)
  void* gg = nullptr;       // gradient ghosts (rows, or rows x D on a full metric)
  void* grads = nullptr;    // (nb, D, nn)
  void* flux = nullptr;     // full metric: (nb, D, nn) F_d = S[d] . V at every node
  void* cartf = nullptr;    // full metric: (nb, 2D, D, Nt) V at the face nodes
  int32_t* cf = nullptr;    // full metric: each exchange entry's position in cartf
  void* tables = nullptr;   // Tables<T>
  void* coef = nullptr;     // (nb, 1+D*D, nn) in T
  void* dcoef = nullptr;    // diagonal metric (nb, 1+D, nn): sqrt(g), A_d
  void* lmode = nullptr;    // per-line truncation pairs (nb, D, E, P1)
  void* ldual = nullptr;
  uint8_t* flat = nullptr;
  int32_t* tslot = nullptr;
  void* tmodes = nullptr;
  void* tweights = nullptr;
  void* tginv = nullptr;  // per element: the inverse of the truncation's Gram matrix
  // factored truncation (outer shells): radial pairs, angular tables and the
  // inverses of their Gram matrices
  int32_t* fslot = nullptr;
  int32_t* fradial = nullptr;
  FBlock* fblk = nullptr;   // the factored blocks, and the others, for the truncation
  int32_t* oblk = nullptr;
  int64_t nfblk = 0, noblk = 0;
  void* frmodes = nullptr;
  void* frduals = nullptr;
  int32_t* faslot = nullptr;
  void* faproj = nullptr;  // (nangular, E^2, (P+1)^2, (P+1)^2): each plane's projector
  // radiative faces: their exchange rows (u-shaped) and the condition's constants
  int64_t nrad = 0;
  int64_t* rad_rows = nullptr;
  double rad_kappa = 0, rad_invR = 0;
  double* rpart = nullptr;  // per radiative face: the boundary energy's partial sum
  double* spacing = nullptr;
  int32_t* rowptr = nullptr;
  int32_t* cb = nullptr;
  int32_t* cl = nullptr;
  void* cval_u = nullptr;   // coefficients for u (mirror +1)
  void* cval_g = nullptr;   // coefficients for flat gradients (mirror -1)
  int64_t ndir = 0;
  int64_t* dir_rows = nullptr;
  double* dir_x = nullptr;
  double wave_k[3] = {0, 0, 0}, wave_w = 0, wave_phi = 0;
  double* clock = nullptr;  // device (t0, s): the Dirichlet data's time is t0 + s dt
  double clock_dt = 0;      // the dt of the clock (a kernel argument)
  double* part = nullptr;   // per-block reduction partials (3 nb)
  double* red = nullptr;    // final reduction (4)
  double hmin = 0, hspeed = 0;
  bool has_tables = false;
  int64_t device_bytes = 0;
  // the working arrays (address, bytes): stage states, ghosts, gradients, face
  // buffers and reduction partials, which every call writes before it reads them
  std::vector<std::pair<void*, size_t>> work;
  bool uploaded[TWOLDG_NSLOTS] = {false, false, false, false};
  // host <-> device staging of state transfers, allocated on the first
  // transfer and kept: fine-grained host doubles of one state, which the
  // conversion kernels read (upload) and write (download) directly
  double* pinned = nullptr;
  // fused one-launch stages (3D, faces all same-level or mirror, flat metric)
  bool fused = false;
  bool warming = false;     // the throwaway step at creation (fused forms)
  bool no_fused = false;    // the caller asked for the generic stages
  int32_t* nbr6 = nullptr;  // (nb, 6)
  void* uf[2] = {nullptr, nullptr};
  void* pf[2] = {nullptr, nullptr};
  // ghosted fused stages (coarse-fine faces): the (fine face, coarse quarter)
  // pairs, and the complete face ghosts (nb, 6, N^2) formed from them
  bool ghosted = false;
  int64_t ncf = 0;
  CFJob* cfjobs = nullptr;
  void* fgu = nullptr;
  void* fgg = nullptr;
  // split form: v of the step start as an FP32 (hi, lo) pair, v of the stage
  // state and of the accumulator in FP32, (nb, N^3) each
  float* svn = nullptr;
  float* svlo = nullptr;
  float* svs = nullptr;
  float* svacc = nullptr;
  // the fused RK4 step captured as a graph (replayed for every step of a call)
  hipGraph_t step_graph = nullptr;
  hipGraphExec_t step_exec = nullptr;
  double step_dt = 0;
  // two-dimensional uniform flat problems: RK4 steps by the batch stage
  twoldg_batch* b2 = nullptr;
  // per-launch timing of the stage kernels (bench): event pairs
  bool timing = false;
  std::vector<hipEvent_t> tev;
  size_t ntev = 0;
};

namespace {

template <class T>
int dalloc(twoldg_problem* p, T** ptr, size_t count) {
  *ptr = nullptr;
  if (count == 0) return 0;
  hipError_t e = pool_alloc((void**)ptr, count * sizeof(T), false);
  if (e != hipSuccess)
    return fail(TWOLDG_ENOMEM, std::string("hipMalloc of ") + std::to_string(count * sizeof(T)) +
                                   " bytes: " + hipGetErrorString(e));
  p->device_bytes += (int64_t)(count * sizeof(T));
  return 0;
}

int dalloc_bytes(twoldg_problem* p, void** ptr, size_t bytes) {
  *ptr = nullptr;
  if (bytes == 0) return 0;
  hipError_t e = pool_alloc(ptr, bytes, false);
  if (e != hipSuccess)
    return fail(TWOLDG_ENOMEM, std::string("hipMalloc of ") + std::to_string(bytes) +
                                   " bytes: " + hipGetErrorString(e));
  p->device_bytes += (int64_t)bytes;
  return 0;
}

// Host memory the GPU reads or writes directly (staging of tables and states).
// It is allocated fine-grained (coherent): the GPU does not cache it, so a
// kernel always sees what the host wrote last, and the host what a kernel
// wrote once the stream is synchronized.  The default hipHostMalloc memory is
// coarse-grained and may be cached by the GPU; a new allocation can reuse the
// host pages of a freed one whose lines are still cached, and then part of a
// fresh upload read as the previous one's contents (seen on the first problem
// built on a new forest in a process, whose state and tables differ from the
// last problem's; a rebuilt problem, uploading the same data to the same pages,
// read right).
inline hipError_t host_stage_alloc(void** ptr, size_t bytes) {
  return pool_alloc(ptr, bytes, true);
}

// Host data into device memory: staged in fine-grained host memory and written
// by a kernel on the problem's stream (no DMA engine touches the device copy).
__global__ void k_copy_words(int64_t n, const uint32_t* __restrict__ src, uint32_t* __restrict__ dst) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[i];
}

int h2d(twoldg_problem* p, void* dst, const void* src, size_t bytes) {
  if (bytes == 0) return 0;
  const size_t words = (bytes + 3) / 4;
  void* pinned = nullptr;
  HIP_TRY(host_stage_alloc(&pinned, words * 4));
  std::memcpy(pinned, src, bytes);
  if (bytes % 4) std::memset((char*)pinned + bytes, 0, words * 4 - bytes);
  // whole words only: the device allocations are at least 4-byte granular
  hipLaunchKernelGGL(k_copy_words, dim3((unsigned)((words + 255) / 256)), dim3(256), 0, p->stream,
                     (int64_t)words, (const uint32_t*)pinned, (uint32_t*)dst);
  hipError_t e = hipStreamSynchronize(p->stream);
  pool_free(pinned);
  if (e != hipSuccess) return fail(TWOLDG_EDEVICE, std::string("table upload: ") + hipGetErrorString(e));
  return 0;
}

void dfree(void* ptr) {
  if (ptr) pool_free(ptr);
}

inline unsigned grid_for(int64_t n, int threads) { return (unsigned)((n + threads - 1) / threads); }

// The exchange operator as CSR rows over (block, dim, side, tangential node).
struct Csr {
  std::vector<int32_t> rowptr, cb, cl;
  std::vector<int64_t> cf;  // each entry's source in the face-compact layout
  std::vector<double> vu, vg;
  std::vector<int64_t> dir_rows;
  std::vector<int64_t> rad_rows;
};

int64_t ipow(int64_t a, int b) {
  int64_t r = 1;
  for (int i = 0; i < b; ++i) r *= a;
  return r;
}

// Build the exchange rows from the face descriptors.  Element-local operators
// (TreeAMR Operators(family=ElementLocal), CODE.md "The coupling"): injection
// normal to the face; tangentially I (coarse polynomial at fine nodes) for a
// fine ghost and R = M_c^-1 I^T M_f for a coarse ghost, tensor over the D-1
// tangential dimensions.
int build_csr(const twoldg_desc* d, Csr& csr) {
  const int D = d->D, P1 = d->P + 1, N = d->E * P1, E = d->E;
  const int64_t Nt = ipow(N, D - 1), nn = ipow(N, D), nb = d->nblocks;
  const twoldg::ElementTables& et = twoldg::kElementTables[d->P - 1];
  const int64_t nrows = nb * 2 * D * Nt;
  csr.rowptr.assign(nrows + 1, 0);
  csr.cb.reserve(nrows);
  csr.cl.reserve(nrows);
  csr.vu.reserve(nrows);
  csr.vg.reserve(nrows);
  bool bad = false;
  // an entry: source node I of block blk, which lies on that block's face
  // fface = 2 d + s (the face that touches the row's, or the row's own face)
  auto push = [&](int64_t blk, const int I[3], double vu, double vg, int fface) {
    int64_t l = I[0] + (int64_t)N * (D >= 2 ? I[1] : 0) + (int64_t)N * N * (D >= 3 ? I[2] : 0);
    for (int e = 0; e < D; ++e)
      if (I[e] < 0 || I[e] >= N) bad = true;
    if (blk < 0 || blk >= nb || l < 0 || l >= nn) bad = true;
    const int fd = fface / 2;
    if (I[fd] != (fface % 2 ? N - 1 : 0)) bad = true;
    int64_t tl = 0, st = 1;
    for (int e = 0; e < D; ++e)
      if (e != fd) {
        tl += st * I[e];
        st *= N;
      }
    csr.cb.push_back((int32_t)blk);
    csr.cl.push_back((int32_t)l);
    csr.cf.push_back(((blk * 2 * D + fface) * D) * Nt + tl);
    csr.vu.push_back(vu);
    csr.vg.push_back(vg);
  };
  int64_t r = 0;
  for (int64_t b = 0; b < nb; ++b)
    for (int dd = 0; dd < D; ++dd)
      for (int s = 0; s < 2; ++s) {
        const twoldg_face& F = d->faces[(b * D + dd) * 2 + s];
        int tdim[2] = {0, 0}, nt = 0;
        for (int e = 0; e < D; ++e)
          if (e != dd) tdim[nt++] = e;
        for (int64_t t = 0; t < Nt; ++t, ++r) {
          int T[2] = {(int)(t % N), (int)(nt >= 2 ? t / N : 0)};
          int I[3] = {0, 0, 0};
          switch (F.kind) {
            case TWOLDG_FACE_SAME: {
              if (F.nb[0] < 0 || F.nb[0] >= nb) return fail(TWOLDG_EINVAL, "face neighbour out of range");
              I[dd] = s ? 0 : N - 1;
              for (int k = 0; k < nt; ++k) I[tdim[k]] = T[k];
              push(F.nb[0], I, 1.0, 1.0, 2 * dd + 1 - s);
            } break;
            case TWOLDG_FACE_MIRROR: {
              I[dd] = s ? N - 1 : 0;
              for (int k = 0; k < nt; ++k) I[tdim[k]] = T[k];
              push(b, I, 1.0, -1.0, 2 * dd + s);
            } break;
            case TWOLDG_FACE_GLUE: {
              if (F.nb[0] < 0 || F.nb[0] >= nb) return fail(TWOLDG_EINVAL, "glue neighbour out of range");
              const int nd = F.nface / 2, ns = F.nface % 2;
              int ntd[2] = {0, 0}, k2 = 0;
              for (int e = 0; e < D; ++e)
                if (e != nd) ntd[k2++] = e;
              I[nd] = ns ? N - 1 : 0;
              for (int j = 0; j < nt; ++j) {
                const int m = F.map[j];
                const int k = (m > 0 ? m : -m) - 1;
                if (k < 0 || k >= nt) return fail(TWOLDG_EINVAL, "glue map entry out of range");
                I[ntd[j]] = m > 0 ? T[k] : N - 1 - T[k];
              }
              push(F.nb[0], I, 1.0, 1.0, F.nface);
            } break;
            case TWOLDG_FACE_COARSER: {
              if (F.nb[0] < 0 || F.nb[0] >= nb) return fail(TWOLDG_EINVAL, "coarse neighbour out of range");
              int qc[2] = {0, 0}, fn[2] = {0, 0};
              for (int k = 0; k < nt; ++k) {
                const int qf = T[k] / P1, jj = T[k] % P1, f = qf % 2;
                qc[k] = F.off[tdim[k]] * (E / 2) + qf / 2;
                fn[k] = f * P1 + jj;  // fine node index within the coarse element's pair
              }
              const int64_t cnt = ipow(P1, nt);
              for (int64_t a = 0; a < cnt; ++a) {
                double w = 1;
                int64_t rem = a;
                I[dd] = s ? 0 : N - 1;
                for (int k = 0; k < nt; ++k) {
                  const int ic = (int)(rem % P1);
                  rem /= P1;
                  I[tdim[k]] = qc[k] * P1 + ic;
                  w *= et.interp[fn[k]][ic];
                }
                push(F.nb[0], I, w, w, 2 * dd + 1 - s);
              }
            } break;
            case TWOLDG_FACE_FINER: {
              int qc[2] = {0, 0}, ic[2] = {0, 0}, o[2] = {0, 0};
              int ch = 0;
              for (int k = 0; k < nt; ++k) {
                qc[k] = T[k] / P1;
                ic[k] = T[k] % P1;
                o[k] = (qc[k] * P1 >= N / 2) ? 1 : 0;
                ch |= o[k] << k;
              }
              const int64_t fb = F.nb[ch];
              if (fb < 0 || fb >= nb) return fail(TWOLDG_EINVAL, "fine neighbour out of range");
              const int64_t cnt = ipow(2 * P1, nt);
              for (int64_t a = 0; a < cnt; ++a) {
                double w = 1;
                int64_t rem = a;
                I[dd] = s ? 0 : N - 1;
                for (int k = 0; k < nt; ++k) {
                  const int fj = (int)(rem % (2 * P1));
                  rem /= 2 * P1;
                  const int f = fj / P1, jj = fj % P1;
                  const int qfine = 2 * (qc[k] - o[k] * (E / 2)) + f;
                  I[tdim[k]] = qfine * P1 + jj;
                  w *= et.restrict_[ic[k]][fj];
                }
                push(fb, I, w, w, 2 * dd + 1 - s);
              }
            } break;
            case TWOLDG_FACE_DIRICHLET:
              csr.dir_rows.push_back(r);
              break;
            case TWOLDG_FACE_RADIATIVE: {
              // u's ghost is the face node's own u; the gradient's starts as the
              // face node's and is then corrected by k_radiative_grad
              I[dd] = s ? N - 1 : 0;
              for (int k = 0; k < nt; ++k) I[tdim[k]] = T[k];
              push(b, I, 1.0, 1.0, 2 * dd + s);
              csr.rad_rows.push_back(r);
            } break;
            default:
              return fail(TWOLDG_EINVAL, "block " + std::to_string(b) + " face (" +
                                             std::to_string(dd) + "," + std::to_string(s) +
                                             ") has no valid kind");
          }
          if (bad) return fail(TWOLDG_EINVAL, "an exchange source lies outside its block");
          if (csr.cb.size() > (size_t)INT32_MAX) return fail(TWOLDG_EINVAL, "exchange too large");
          csr.rowptr[r + 1] = (int32_t)csr.cb.size();
        }
      }
  return 0;
}

template <class T>
int upload_as(twoldg_problem* p, void** dst, const double* src, size_t n) {
  std::vector<T> tmp(n);
  for (size_t i = 0; i < n; ++i) tmp[i] = (T)src[i];
  int rc = dalloc_bytes(p, dst, n * sizeof(T));
  if (rc) return rc;
  return h2d(p, *dst, tmp.data(), n * sizeof(T));
  return 0;
}

template <class T>
int upload_tables(twoldg_problem* p) {
  const twoldg::ElementTables& et = twoldg::kElementTables[p->P - 1];
  Tables<T> tb;
  std::memset(&tb, 0, sizeof tb);
  for (int i = 0; i < p->P1; ++i) {
    for (int j = 0; j < p->P1 + 2; ++j) tb.W[i][j] = (T)et.stencil[i][j];
    tb.mode[i] = (T)et.topmode[i];
    tb.dual[i] = (T)et.topdual[i];
    tb.wq[i] = et.weights[i];
  }
  int rc = dalloc_bytes(p, &p->tables, sizeof tb);
  if (rc) return rc;
  if ((rc = h2d(p, p->tables, &tb, sizeof tb))) return rc;
  return 0;
}

// Stage timing for the benchmark: an event pair around each fused stage launch.
inline void tmark(twoldg_problem* p) {
  if (!p->timing) return;
  if (p->ntev >= p->tev.size()) {
    hipEvent_t e;
    (void)hipEventCreate(&e);
    p->tev.push_back(e);
  }
  (void)hipEventRecord(p->tev[p->ntev++], p->stream);
}

// The problem's clock at t0 (step count 0), for steps of dt.
inline void set_clock(twoldg_problem* p, double t0, double dt) {
  p->clock_dt = dt;
  hipLaunchKernelGGL(k_set_clock, dim3(1), dim3(1), 0, p->stream, p->clock, t0);
}

// ---- launch helpers, dispatched on (D, P1) at run time ----------------------
template <class T, int D, int P1>
struct Launch {
  static ExchangeRows rows(twoldg_problem* p) {
    return {p->nrows, p->rowptr, p->cb, p->cl, p->cf};
  }
  static DirichletRows dir(twoldg_problem* p) {
    return {p->ndir,     p->dir_rows,  p->dir_x,     p->wave_k[0], p->wave_k[1],
            p->wave_k[2], p->wave_w,   p->wave_phi,  p->clock,     p->clock_dt};
  }
  // the ghosts of src: exchange rows, and the Dirichlet rows at the clock's
  // time plus cdt
  template <int MODE>
  static void ghosts(twoldg_problem* p, const T* coef, const T* src, double cdt, T* ghost) {
    if (p->nrows == 0) return;
    if constexpr (MODE != 1) {  // Dirichlet data come with the Cartesian gradients
      if (p->ndir) {
        hipLaunchKernelGGL((k_exchange_dirichlet<T, MODE, D>),
                           dim3(grid_for(p->nrows + p->ndir, 256)), dim3(256), 0, p->stream,
                           p->sh, rows(p), coef, src, dir(p), cdt, ghost);
        return;
      }
    }
    hipLaunchKernelGGL((k_exchange<T, MODE, D>), dim3(grid_for(p->nrows, 256)), dim3(256), 0,
                       p->stream, p->sh, rows(p), coef, src, ghost);
  }
  static void exchange_u(twoldg_problem* p, const T* y, double cdt) {
    ghosts<0>(p, (const T*)p->cval_u, y, cdt, (T*)p->ug);
  }
  // the gradient (flat, or Cartesian on a full metric) of y and its ghosts, the
  // Dirichlet data at the clock's time plus cdt; for the energy, the gradient at
  // every node (in grads) and no ghosts
  static void gradient(twoldg_problem* p, const T* y, double cdt, bool for_energy = false) {
    exchange_u(p, y, cdt);
    if (!p->full) {
      hipLaunchKernelGGL((k_gradient<T, D, P1>), block_node_grid(p->sh), dim3(256), 0, p->stream,
                         p->sh, (const Tables<T>*)p->tables, p->spacing, y, (const T*)p->ug,
                         (const T*)p->dcoef, (T*)p->grads);
      if (!for_energy) ghosts<1>(p, (const T*)p->cval_g, (const T*)p->grads, cdt, (T*)p->gg);
    } else if (for_energy) {
      hipLaunchKernelGGL((k_gradient_cart<T, D, P1, false>), block_node_grid(p->sh), dim3(256),
                         0, p->stream, p->sh, (const Tables<T>*)p->tables, p->spacing, y,
                         (const T*)p->ug, (const T*)p->coef, p->flat, (T*)p->grads,
                         (T*)nullptr);
    } else {
      hipLaunchKernelGGL((k_gradient_cart<T, D, P1, true>), block_node_grid(p->sh), dim3(256), 0,
                         p->stream, p->sh, (const Tables<T>*)p->tables, p->spacing, y,
                         (const T*)p->ug, (const T*)p->coef, p->flat, (T*)p->cartf,
                         (T*)p->flux);
      ghosts<2>(p, (const T*)p->cval_u, (const T*)p->cartf, cdt, (T*)p->gg);
      if (p->nrad)
        hipLaunchKernelGGL((k_radiative_grad<T, D>), dim3(grid_for(p->nrad, 256)), dim3(256), 0,
                           p->stream, p->sh, p->nrad, p->rad_rows, (const T*)p->coef, y,
                           p->rad_kappa, p->rad_invR, (T*)p->gg);
    }
  }
  static FactoredTabs<T> factored(twoldg_problem* p) {
    return {p->fblk,            p->nfblk ? p->oblk : nullptr,
            p->fslot,           p->fradial,          (const T*)p->frmodes,
            (const T*)p->frduals, p->faslot,         (const T*)p->faproj};
  }
  template <int MODE>
  static void stage(twoldg_problem* p, const StageIO<T>& io) {
    const int64_t n = p->sh.nb * p->sh.nn;
    tmark(p);
    stage_launch<MODE>(p, io, n);
    tmark(p);
  }
  template <int MODE>
  static void stage_launch(twoldg_problem* p, const StageIO<T>& io, int64_t n) {
    if (!p->full)
      hipLaunchKernelGGL((k_stage_flat<T, D, P1, MODE>), block_node_grid(p->sh), dim3(256), 0,
                         p->stream, p->sh, (const Tables<T>*)p->tables, p->spacing,
                         (const T*)p->grads, (const T*)p->gg, (const T*)p->dcoef, io);
    else
      hipLaunchKernelGGL((k_stage_cart<T, D, P1, MODE>), block_node_grid(p->sh), dim3(256), 0,
                         p->stream, p->sh, (const Tables<T>*)p->tables, p->spacing,
                         (const T*)p->flux, (const T*)p->gg, (const T*)p->coef, p->flat, io);
  }
  static void truncate(twoldg_problem* p, T* y) {
    if (!p->truncate) return;
    const int64_t ne = ipow(p->E, D);
    const int64_t n = p->sh.nb * ne;
    if constexpr (D == 3 && P1 <= 5) {
      if (p->has_tables) {  // every block: per-element tables or the tensor projector
        constexpr int epw = trunc_elem_epw<P1>(), thr = trunc_elem_threads<P1>() * epw;
        const FactoredTabs<T> F = factored(p);
        if (p->nfblk)  // the factored blocks, then the others
          hipLaunchKernelGGL((k_truncate_elem<T, P1, 1>), dim3((unsigned)(p->nfblk * ne / epw)),
                             dim3(thr), 0, p->stream, p->sh, p->E, (const Tables<T>*)p->tables,
                             p->tslot, (const T*)p->tmodes, (const T*)p->tweights,
                             (const T*)p->tginv, F, y);
        const int64_t no = p->nfblk ? p->noblk : p->sh.nb;
        if (no)
          hipLaunchKernelGGL((k_truncate_elem<T, P1, 0>), dim3((unsigned)(no * ne / epw)),
                             dim3(thr), 0, p->stream, p->sh, p->E, (const Tables<T>*)p->tables,
                             p->tslot, (const T*)p->tmodes, (const T*)p->tweights,
                             (const T*)p->tginv, F, y);
        return;
      }
    }
    hipLaunchKernelGGL((k_truncate<T, D, P1>), dim3((unsigned)p->sh.nb), dim3(kTruncThreads), 0,
                       p->stream, p->sh, p->E, (const Tables<T>*)p->tables, p->tslot,
                       (const T*)p->lmode, (const T*)p->ldual, y);
  }
  static void energy(twoldg_problem* p, const T* y) {
    if (!p->full)
      hipLaunchKernelGGL((k_energy_part<T, D, P1, false>), dim3((unsigned)p->sh.nb),
                         dim3(kRedThreads), 0, p->stream, p->sh, (const Tables<T>*)p->tables,
                         p->spacing, y, (const T*)p->grads, (const T*)nullptr,
                         (const T*)p->dcoef, p->part);
    else
      hipLaunchKernelGGL((k_energy_part<T, D, P1, true>), dim3((unsigned)p->sh.nb),
                         dim3(kRedThreads), 0, p->stream, p->sh, (const Tables<T>*)p->tables,
                         p->spacing, y, (const T*)p->grads, (const T*)p->coef,
                         (const T*)nullptr, p->part);
    hipLaunchKernelGGL((k_final<2>), dim3(1), dim3(kRedThreads), 0, p->stream, p->sh.nb, p->part,
                       false, p->red);
    if (p->nrad && p->rad_invR != 0) {  // the radiative boundary's own energy
      const int64_t nf = p->nrad / p->sh.Nt;
      hipLaunchKernelGGL((k_rad_energy<T, D, P1>), dim3((unsigned)nf), dim3(kRedThreads), 0,
                         p->stream, p->sh, (const Tables<T>*)p->tables, p->spacing,
                         p->rad_rows, (const T*)p->coef, y, (const T*)nullptr, p->rpart);
      hipLaunchKernelGGL((k_final<1>), dim3(1), dim3(kRedThreads), 0, p->stream, nf, p->rpart,
                         false, p->red + 2);
    }
  }
  static void norms(twoldg_problem* p, const T* a, const T* alo, const T* b, const T* blo,
                    int mask, int pnorm, double* part, double* red) {
    if (!p->full)
      hipLaunchKernelGGL((k_norm_part<T, D, P1, false>), dim3((unsigned)p->sh.nb),
                         dim3(kRedThreads), 0, p->stream, p->sh, (const Tables<T>*)p->tables,
                         p->spacing, a, alo, b, blo, (const T*)nullptr, (const T*)p->dcoef,
                         mask, pnorm, part);
    else
      hipLaunchKernelGGL((k_norm_part<T, D, P1, true>), dim3((unsigned)p->sh.nb),
                         dim3(kRedThreads), 0, p->stream, p->sh, (const Tables<T>*)p->tables,
                         p->spacing, a, alo, b, blo, (const T*)p->coef, (const T*)nullptr,
                         mask, pnorm, part);
    hipLaunchKernelGGL((k_final<3>), dim3(1), dim3(kRedThreads), 0, p->stream, p->sh.nb, part,
                       true, red);
  }
};

// The split form's v between the FP64 state (v at (2b + 1) N^3 + l) and the FP32
// pair (hi, lo) with v = hi - lo: hi = fl(v), lo = fl(hi - v).
__global__ void k_split_v(int64_t nb, int64_t nn, const double* __restrict__ y,
                          float* __restrict__ hi, float* __restrict__ lo) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nb * nn) return;
  const double v = y[(i / nn * 2 + 1) * nn + i % nn];
  const float h = (float)v;
  hi[i] = h;
  lo[i] = (float)((double)h - v);
}
__global__ void k_join_v(int64_t nb, int64_t nn, const float* __restrict__ hi,
                         const float* __restrict__ lo, double* __restrict__ y) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nb * nn) return;
  y[(i / nn * 2 + 1) * nn + i % nn] = (double)hi[i] - (double)lo[i];
}

// The fused RK4 steps of a forest, v in TV: T for the FP64 and compensated
// forms, float under FP64 u for the split form (whose v lives in the FP32 pair
// (svn, svlo) between the split and join kernels around the steps).
template <class T, class TV, int P1, int E, bool GHOST>
int fused_run(twoldg_problem* p, double dt, int64_t nsteps) {
  constexpr bool split = !std::is_same<T, TV>::value;
  constexpr int N = P1 * E;
  constexpr int threads = (N * N + 63) / 64 * 64;
  const int64_t nb = p->sh.nb, nn = p->sh.nn;
  // the warm-up at creation steps the stage buffer in place of the state
  T* yn = (T*)(p->warming ? p->ys : p->slot[TWOLDG_SLOT_STATE]);
  {
    const int64_t n = nb * 6 * (int64_t)N * N;
    hipLaunchKernelGGL((k_faces3<T, P1, E>), dim3(grid_for(n, 256)), dim3(256), 0, p->stream, nb,
                       yn, (T*)p->uf[0], (T*)p->pf[0]);
  }
  if constexpr (split)
    hipLaunchKernelGGL(k_split_v, dim3(grid_for(nb * nn, 256)), dim3(256), 0, p->stream, nb, nn,
                       (const double*)yn, p->svn, p->svlo);
  FusedArgs<T, TV> A{};
  A.yn = yn;
  A.acc = (T*)p->acc;
  A.lo = split ? nullptr : (T*)p->lo;
  A.nbr = p->nbr6;
  A.spacing = p->spacing;
  A.tables = (const Tables<T>*)p->tables;
  A.truncate = p->truncate ? 1 : 0;
  A.gu = (const T*)p->fgu;
  A.gg = (const T*)p->fgg;
  A.nfb = nb * 6 * (int64_t)N * N;
  A.vstride = split ? nn : 2 * nn;
  // v of a stage: the v halves of the state arrays, or the split form's own
  auto vptrs = [&](int stage) {
    if constexpr (split) {
      A.vn = (const TV*)p->svn;
      A.vacc = (TV*)p->svacc;
      A.vlo = (TV*)p->svlo;
      A.vs = stage == 1 ? (const TV*)p->svn : (const TV*)p->svs;
      A.vo = stage == 4 ? (TV*)p->svn : (TV*)p->svs;
    } else {
      A.vs = (const TV*)(A.ys + nn);
      A.vn = (const TV*)(A.yn + nn);
      A.vacc = A.acc ? (TV*)(A.acc + nn) : nullptr;
      A.vo = (TV*)(A.yo + nn);
      A.vlo = A.lo ? (TV*)(A.lo + nn) : nullptr;
    }
  };
  // the complete ghosts of the coarse-fine faces from the stage input's face
  // buffers (ghosted form): one launch, a workgroup per face pair
  auto ghosts = [&](int in) {
    if constexpr (GHOST) {
      if (p->ncf == 0) return;
      hipLaunchKernelGGL((k_cf_ghosts<T, P1, E>), dim3((unsigned)p->ncf), dim3(256), 0, p->stream,
                         p->cfjobs, (const T*)p->uf[in], (const T*)p->pf[in], p->spacing,
                         (T*)p->fgu, (T*)p->fgg);
    }
  };
  auto io = [&](int in) {
    A.uf_in = (const T*)p->uf[in];
    A.pf_in = (const T*)p->pf[in];
    A.uf_out = (T*)p->uf[1 - in];
    A.pf_out = (T*)p->pf[1 - in];
  };
  // FP64 and the split form's u: the stage states in the two buffers by turns,
  // no accumulator: y_2 in ys; y_3 in acc; y_4 over y_3 and z over y_2 (FP64
  // modes 11, 12, 13, 8; split 1, 14, 15, 10, whose v keeps its accumulator)
  auto noacc_step = [&]() {
    if constexpr (!std::is_same<T, float>::value) {
      A.ys = yn;  // stage 1 relies on ys = yn
      A.yo = (T*)p->ys;
      A.acc = nullptr;
      A.a = (T)(dt / 2);
      vptrs(1);
      io(0);
      tmark(p);
      ghosts(0);
      hipLaunchKernelGGL((k_fused3<T, P1, E, split ? 1 : 11, GHOST, TV>), dim3((unsigned)nb),
                         dim3(threads), 0, p->stream, A);
      tmark(p);
      A.ys = (T*)p->ys;
      A.yo = (T*)p->acc;
      vptrs(2);
      io(1);
      tmark(p);
      ghosts(1);
      hipLaunchKernelGGL((k_fused3<T, P1, E, split ? 14 : 12, GHOST, TV>), dim3((unsigned)nb),
                         dim3(threads), 0, p->stream, A);
      tmark(p);
      A.ys = (T*)p->acc;
      A.yo = (T*)p->acc;
      A.acc = (T*)p->ys;
      A.a = (T)dt;
      A.a2 = (T)(dt / 3);
      vptrs(3);
      io(0);
      tmark(p);
      ghosts(0);
      hipLaunchKernelGGL((k_fused3<T, P1, E, split ? 15 : 13, GHOST, TV>), dim3((unsigned)nb),
                         dim3(threads), 0, p->stream, A);
      tmark(p);
      A.a = (T)(dt / 6);
      A.yo = yn;
      vptrs(4);
      io(1);
      tmark(p);
      ghosts(1);
      hipLaunchKernelGGL((k_fused3<T, P1, E, split ? 10 : 8, GHOST, TV>), dim3((unsigned)nb),
                         dim3(threads), 0, p->stream, A);
      tmark(p);
    }
  };
  // FP32c: the accumulator's three registers (modes 1, 6, 2, 4)
  auto acc_step = [&]() {
    if constexpr (std::is_same<T, float>::value) {
      A.ys = yn;  // stage 1 (mode 1) relies on ys = yn
      A.yo = (T*)p->ys;
      A.a = (T)(dt / 2);
      vptrs(1);
      io(0);
      tmark(p);
      ghosts(0);
      hipLaunchKernelGGL((k_fused3<T, P1, E, 1, GHOST, TV>), dim3((unsigned)nb), dim3(threads), 0,
                         p->stream, A);
      tmark(p);
      A.ys = (T*)p->ys;
      vptrs(2);
      io(1);
      tmark(p);
      ghosts(1);
      hipLaunchKernelGGL((k_fused3<T, P1, E, 6, GHOST, TV>), dim3((unsigned)nb), dim3(threads), 0,
                         p->stream, A);
      tmark(p);
      A.a = (T)dt;
      io(0);
      tmark(p);
      ghosts(0);
      hipLaunchKernelGGL((k_fused3<T, P1, E, 2, GHOST, TV>), dim3((unsigned)nb), dim3(threads), 0,
                         p->stream, A);
      tmark(p);
      A.a = (T)(dt / 6);
      A.yo = yn;
      vptrs(4);
      io(1);
      tmark(p);
      ghosts(1);
      hipLaunchKernelGGL((k_fused3<T, P1, E, 4, GHOST, TV>), dim3((unsigned)nb), dim3(threads), 0,
                         p->stream, A);
      tmark(p);
    }
  };
  auto one_step = [&]() {
    if constexpr (std::is_same<T, float>::value)
      acc_step();
    else
      noacc_step();
  };
  if (p->timing || p->warming || nsteps < 2) {
    for (int64_t s = 0; s < nsteps; ++s) one_step();
  } else {
    // the step's launches replayed as one graph: they are the same every step
    // (the stage buffers ping-pong within the step); rebuilt when dt changes
    if (!p->step_exec || p->step_dt != dt) {
      if (p->step_exec) (void)hipGraphExecDestroy(p->step_exec);
      if (p->step_graph) (void)hipGraphDestroy(p->step_graph);
      p->step_exec = nullptr;
      p->step_graph = nullptr;
      HIP_TRY(hipStreamBeginCapture(p->stream, hipStreamCaptureModeThreadLocal));
      one_step();
      HIP_TRY(hipStreamEndCapture(p->stream, &p->step_graph));
      HIP_TRY(hipGraphInstantiate(&p->step_exec, p->step_graph, nullptr, nullptr, 0));
      p->step_dt = dt;
    }
    for (int64_t s = 0; s < nsteps; ++s) HIP_TRY(hipGraphLaunch(p->step_exec, p->stream));
  }
  if constexpr (split)
    hipLaunchKernelGGL(k_join_v, dim3(grid_for(nb * nn, 256)), dim3(256), 0, p->stream, nb, nn,
                       (const float*)p->svn, (const float*)p->svlo, (double*)yn);
  HIP_TRY(hipGetLastError());
  return 0;
}

template <class T, int P1, int E, bool GHOST>
int fused_steps(twoldg_problem* p, double dt, int64_t nsteps) {
  if constexpr (std::is_same<T, double>::value)
    if (p->precision == TWOLDG_SPLIT) return fused_run<double, float, P1, E, GHOST>(p, dt, nsteps);
  return fused_run<T, T, P1, E, GHOST>(p, dt, nsteps);
}

// (P+1, E) pairs the fused stage is compiled for.
template <class T, int D, int P1>
int fused_dispatch(twoldg_problem* p, double dt, int64_t nsteps) {
  if constexpr (D == 3) {
    switch (p->E) {
      case 2:
        if constexpr (P1 * 2 <= 32)
          return p->ghosted ? fused_steps<T, P1, 2, true>(p, dt, nsteps)
                            : fused_steps<T, P1, 2, false>(p, dt, nsteps);
        break;
      case 4:
        if constexpr (P1 * 4 <= 24)
          return p->ghosted ? fused_steps<T, P1, 4, true>(p, dt, nsteps)
                            : fused_steps<T, P1, 4, false>(p, dt, nsteps);
        break;
    }
  }
  return fail(TWOLDG_EINVAL, "no fused stage for this shape");
}

inline bool fused_supported(int D, int P1, int E) {
  return D == 3 && ((E == 2 && P1 * 2 <= 32) || (E == 4 && P1 * 4 <= 24));
}

template <class T, int D, int P1>
int rk4_steps(twoldg_problem* p, double t0, double dt, int64_t nsteps) {
  if (p->fused) return fused_dispatch<T, D, P1>(p, dt, nsteps);
  if (p->b2) return twoldg_detail::batch_attached_steps(p->b2, dt, nsteps);
  using L = Launch<T, D, P1>;
  T* yn = (T*)p->slot[TWOLDG_SLOT_STATE];
  T* ys = (T*)p->ys;
  T* acc = (T*)p->acc;
  T* lo = (T*)p->lo;
  // one step at the clock's time t0 + s dt, then the clock's tick
  auto one_step = [&]() {
    if constexpr (std::is_same<T, float>::value) {  // FP32c: three registers
      StageIO<T> io{yn, yn, acc, ys, lo, (T)(dt / 2), (T)0};
      L::gradient(p, yn, 0.0);
      L::template stage<1>(p, io);
      L::truncate(p, ys);
      io.ys = ys;
      L::gradient(p, ys, dt / 2);
      L::template stage<2>(p, io);
      L::truncate(p, ys);
      io.a = (T)dt;
      L::gradient(p, ys, dt / 2);
      L::template stage<2>(p, io);
      L::truncate(p, ys);
      io.a = (T)(dt / 6);
      io.yo = yn;
      L::gradient(p, ys, dt);
      L::template stage<4>(p, io);
      L::truncate(p, yn);
      L::truncate(p, lo);
    } else {
      // FP64 without the accumulator: y_2 in ys, y_3 in acc, then y_4 over
      // y_3 and z over y_2
      StageIO<T> io{yn, yn, nullptr, ys, nullptr, (T)(dt / 2), (T)0};
      L::gradient(p, yn, 0.0);
      L::template stage<5>(p, io);
      L::truncate(p, ys);
      io.ys = ys;
      io.yo = acc;
      L::gradient(p, ys, dt / 2);
      L::template stage<5>(p, io);
      L::truncate(p, acc);
      io.ys = acc;
      io.acc = ys;
      io.yo = acc;
      io.a = (T)dt;
      io.a2 = (T)(dt / 3);
      L::gradient(p, acc, dt / 2);
      L::template stage<7>(p, io);
      L::truncate(p, acc);
      io.a = (T)(dt / 6);
      io.yo = yn;
      L::gradient(p, acc, dt);
      L::template stage<8>(p, io);
      L::truncate(p, yn);
    }
    hipLaunchKernelGGL(k_tick, dim3(1), dim3(1), 0, p->stream, p->clock);
  };
  set_clock(p, t0, dt);
  if (p->timing || nsteps < 2) {
    for (int64_t s = 0; s < nsteps; ++s) one_step();
  } else {
    // the step's launches replayed as one graph; the Dirichlet data read the
    // step's time from the device clock, which the step's last launch advances
    if (!p->step_exec || p->step_dt != dt) {
      if (p->step_exec) (void)hipGraphExecDestroy(p->step_exec);
      if (p->step_graph) (void)hipGraphDestroy(p->step_graph);
      p->step_exec = nullptr;
      p->step_graph = nullptr;
      HIP_TRY(hipStreamBeginCapture(p->stream, hipStreamCaptureModeThreadLocal));
      one_step();
      HIP_TRY(hipStreamEndCapture(p->stream, &p->step_graph));
      HIP_TRY(hipGraphInstantiate(&p->step_exec, p->step_graph, nullptr, nullptr, 0));
      p->step_dt = dt;
    }
    for (int64_t s = 0; s < nsteps; ++s) HIP_TRY(hipGraphLaunch(p->step_exec, p->stream));
  }
  HIP_TRY(hipGetLastError());
  return 0;
}

template <class T, int D, int P1>
int rhs_impl(twoldg_problem* p, int src, int dst, double t) {
  using L = Launch<T, D, P1>;
  const T* y = (const T*)p->slot[src];
  set_clock(p, t, 0.0);
  L::gradient(p, y, 0.0);
  StageIO<T> io{y, y, nullptr, (T*)p->slot[dst], nullptr, (T)0, (T)0};
  L::template stage<0>(p, io);
  HIP_TRY(hipGetLastError());
  return 0;
}

template <class T, int D, int P1>
int truncate_impl(twoldg_problem* p, int slot) {
  using L = Launch<T, D, P1>;
  L::truncate(p, (T*)p->slot[slot]);
  if (slot == TWOLDG_SLOT_STATE && p->lo) L::truncate(p, (T*)p->lo);
  HIP_TRY(hipGetLastError());
  return 0;
}

template <class T, int D, int P1>
int energy_impl(twoldg_problem* p, int slot, double t, double* out) {
  using L = Launch<T, D, P1>;
  const T* y = (const T*)p->slot[slot];
  set_clock(p, t, 0.0);
  L::gradient(p, y, 0.0, true);
  L::energy(p, y);
  double h[3] = {0, 0, 0};
  const bool bnd = p->nrad && p->rad_invR != 0;
  HIP_TRY(hipMemcpyAsync(h, p->red, (bnd ? 3 : 2) * sizeof(double), hipMemcpyDeviceToHost,
                         p->stream));
  HIP_TRY(hipStreamSynchronize(p->stream));
  *out = (h[0] + h[1]) / 2;
  if (bnd) *out += h[2] * p->rad_invR / 2;  // boundary_energy, evolution.jl:1276-1299
  return 0;
}

// part: 3 nb doubles of per-block partials, red: 3 doubles of final sums
// (the problem's own buffers, or the no-allocation entry point's scratch).
template <class T, int D, int P1>
int norms_impl(twoldg_problem* p, int a, int b, int mask, int pnorm, double* lp, double* linf,
               double* part, double* red) {
  using L = Launch<T, D, P1>;
  const T* alo = (a == TWOLDG_SLOT_STATE) ? (const T*)p->lo : nullptr;
  const T* blo = (b == TWOLDG_SLOT_STATE) ? (const T*)p->lo : nullptr;
  L::norms(p, (const T*)p->slot[a], alo, b >= 0 ? (const T*)p->slot[b] : nullptr, blo, mask,
           pnorm, part, red);
  double h[3];
  HIP_TRY(hipMemcpyAsync(h, red, 3 * sizeof(double), hipMemcpyDeviceToHost, p->stream));
  HIP_TRY(hipStreamSynchronize(p->stream));
  *lp = h[1] == 0 ? 0 : std::pow(h[0] / h[1], 1.0 / pnorm);
  *linf = h[2];
  return 0;
}

// The modal amplitudes of variable var of a slot, one per element, into the host
// array tau (blocks slowest, then e_D .. e_1).  The peak A = max |var| is the
// norms' reduction, left on the device for the element kernel to read.
template <class T, int D, int P1>
int modes_impl(twoldg_problem* p, int slot, int var, ModeArgs M, double* tau) {
  using L = Launch<T, D, P1>;
  const T* y = (const T*)p->slot[slot];
  const T* lo = slot == TWOLDG_SLOT_STATE ? (const T*)p->lo : nullptr;
  L::norms(p, y, lo, nullptr, nullptr, 1 << var, 2, p->part, p->red);
  int64_t ne = p->sh.nb;
  for (int d = 0; d < D; ++d) ne *= p->E;
  double* dtau = nullptr;
  int rc = dalloc(p, &dtau, (size_t)ne);
  if (rc) return rc;
  hipLaunchKernelGGL((k_modes<T, D, P1>), dim3(grid_for(ne, 256)), dim3(256), 0, p->stream,
                     p->sh, p->E, y, lo, var, M, (const double*)p->red + 2, dtau);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(tau, dtau, ne * sizeof(double), hipMemcpyDeviceToHost,
                                          p->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(p->stream);
  dfree(dtau);
  if (e != hipSuccess) return fail(TWOLDG_EDEVICE, std::string("modes: ") + hipGetErrorString(e));
  return 0;
}

// Each block of dst from its origin in src (copy, refine, coarsen): map is a
// device array of dst->sh.nb entries, checked on the host.
template <class T, int D, int P1>
int transfer_impl(twoldg_problem* dst, int dslot, const twoldg_problem* src, int sslot,
                  const TransferBlock* map) {
  const T* s = (const T*)src->slot[sslot];
  const T* slo = sslot == TWOLDG_SLOT_STATE ? (const T*)src->lo : nullptr;
  T* lo = dslot == TWOLDG_SLOT_STATE ? (T*)dst->lo : nullptr;
  hipLaunchKernelGGL((k_transfer<T, D, P1>), dim3((unsigned)dst->sh.nb), dim3(256), 0,
                     dst->stream, dst->sh, dst->E, map, s, slo, (T*)dst->slot[dslot], lo);
  HIP_TRY(hipGetLastError());
  return 0;
}

// Run-time dispatch on (precision, D, P1).

#define TWOLDG_DISPATCH_FN(name, impl)                                                     \
  template <class T, int D, int P1>                                                        \
  struct name {                                                                            \
    template <class... A>                                                                  \
    static int run(twoldg_problem* p, A... args) {                                         \
      return impl<T, D, P1>(p, args...);                                                   \
    }                                                                                      \
  };
TWOLDG_DISPATCH_FN(RhsF, rhs_impl)
TWOLDG_DISPATCH_FN(Rk4F, rk4_steps)
TWOLDG_DISPATCH_FN(TruncF, truncate_impl)
TWOLDG_DISPATCH_FN(EnergyF, energy_impl)
TWOLDG_DISPATCH_FN(NormsF, norms_impl)
TWOLDG_DISPATCH_FN(ModesF, modes_impl)
TWOLDG_DISPATCH_FN(TransferF, transfer_impl)

template <template <class, int, int> class F, class T, int D, class... A>
int dispatch_p(twoldg_problem* p, A... args) {
  switch (p->P1) {
    case 2: return F<T, D, 2>::run(p, args...);
    case 3: return F<T, D, 3>::run(p, args...);
    case 4: return F<T, D, 4>::run(p, args...);
    case 5: return F<T, D, 5>::run(p, args...);
    case 6: return F<T, D, 6>::run(p, args...);
    case 7: return F<T, D, 7>::run(p, args...);
    case 8: return F<T, D, 8>::run(p, args...);
    case 9: return F<T, D, 9>::run(p, args...);
  }
  return fail(TWOLDG_EINVAL, "unsupported P");
}

}  // namespace
