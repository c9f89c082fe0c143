// fused2.h -- one-launch RK4 stage over a batch of independent 2D problems of
// different element orders.
//
// The blocks of every problem are laid out one after another; a device-resident
// work list gives each workgroup one entry (P+1, first block, block count), so
// one launch advances every block of every problem by one stage.  Inside a
// workgroup one thread owns one node; a block lives in LDS with its x and y
// ghosts: u, then the two gradients, then the Laplacian, the RK4 combination
// (stepping.jl:72-78) and v's truncation (evolution.jl:238-264, x then y).
// Ghosts come from face buffers written by the previous stage (the face plane
// of u and the block-local part of the face node's gradient), as in the 3D
// fused stage.  Each problem has its own spacing and time step.

#pragma once

constexpr int kBatchThreads = 512;
constexpr int kBatchLds = 4096;  // values of T per workgroup

struct BatchItem {
  int32_t P1;      // element nodes of this entry's blocks
  int32_t N;       // nodes per block side, E (P+1)
  int32_t count;   // blocks in this entry
  int32_t pad;
  int64_t first;   // first block
};

template <class T>
struct BatchArgs {
  const BatchItem* items;
  const int64_t* off;      // (nb) state offset of block b (2 N^2 values per block)
  const int64_t* foff;     // (nb) face-buffer offset (4 N values per block)
  const int32_t* nbr;      // (nb, 4) neighbour across face 2d + s, -1 for a mirror
  const double* spacing;   // (nb)
  const double* dt;        // (nb) the block's problem's time step
  const Tables<T>* tables; // [P1 - 2]: tables of order P1 - 1
  const T* ys;
  const T* yn;
  T* acc;
  T* yo;
  T* lo;
  const T* uf_in;
  const T* pf_in;
  T* uf_out;
  T* pf_out;
  double coef;             // stage coefficient in units of dt: 1/2, 1 or 1/6
  int truncate;
};

// The general body: N = E (P+1) nodes per block side, E even, N^2 <= 256.
template <class T, int P1, int MODE>
__device__ void batch_blocks(const BatchArgs<T>& A, const BatchItem& it, int N, T* lds, T* ltab) {
  constexpr int PW = P1 + 2;
  constexpr int P = P1 - 1;
  const int NT = N * N, NP = N + 2;
  const int tid = threadIdx.x;
  const int bl = tid / NT;            // block within the entry
  const bool act = bl < it.count;
  const int node = tid - bl * NT;
  const int x = act ? node % N : 0, y = act ? node / N : 0;
  const int nbk = it.count;
  // per-node rows of W and the top mode from LDS; the uniform entries are
  // compile-time constants
  constexpr FusedTab<T, P1> tb = fused_tab<T, P1>();
  {
    const Tables<T>& gt = A.tables[P1 - 2];
    if (tid < P1 * PW) ltab[tid] = gt.W[tid / PW][tid % PW];
    if (tid < P1) ltab[P1 * PW + tid] = gt.mode[tid];
  }
  // LDS carve: per block U (NP^2), GX (N NP), GY (NP N), VA, VB, UN (N^2)
  const int per = NP * NP + 2 * N * NP + 3 * NT;
  T* base = lds + (act ? bl : 0) * per;
  T* U = base;
  T* GX = U + NP * NP;
  T* GY = GX + N * NP;
  T* VA = GY + NP * N;
  T* VB = VA + NT;
  T* UN = VB + NT;
  const int64_t b = it.first + (act ? bl : 0);
  const T h = (T)A.spacing[b], rh = T(1) / h;
  const T a = (T)(A.coef * A.dt[b]);
  const int32_t* nb4 = A.nbr + b * 4;
  const int nxl = nb4[0], nxh = nb4[1], nyl = nb4[2], nyh = nb4[3];
  const int64_t o = A.off[b];
  const int64_t nn = NT;
  const int64_t iu = o + node, iv = iu + nn;
  const T Wlo = tb.W[0][0], Whi = tb.W[P][P1 + 1];
  const int ix = x % P1, iy = y % P1;
  const int sx = x - ix - 1, sy = y - iy - 1;
  auto fidx = [&](int64_t blk, int f, int t) { return A.foff[blk] + (int64_t)f * N + t; };

  // (1) u and its ghosts; every global operand of the stage, so that no later
  // wait for a load also waits for the stage's stores (the vector memory
  // counter covers both)
  T uz = 0, ku = 0, yu = 0, yv = 0, au = 0, av = 0, lu = 0, lv = 0;
  // an edge node's neighbour across its x (y) face: partial face gradient, 1/h
  T pgx = 0, rgx = 0, pgy = 0, rgy = 0;
  if (act) {
    uz = A.ys[iu];
    ku = A.ys[iv];
    if (MODE == 1) {  // stage 1: the stage input is the step start
      yu = uz;
      yv = ku;
    } else {  // the compensated form: the high words
      yu = A.yn[iu];
      yv = A.yn[iv];
    }
    if (MODE == 4) {
      lu = A.lo[iu];
      lv = A.lo[iv];
    }
    if (MODE >= 2) {
      au = A.acc[iu];
      av = A.acc[iv];
    }
    const int fx = x == 0 ? nxl : (x == N - 1 ? nxh : -1);
    const int fy = y == 0 ? nyl : (y == N - 1 ? nyh : -1);
    if (fx >= 0) {
      pgx = A.pf_in[fidx(fx, x == 0 ? 1 : 0, y)];
      rgx = T(1) / (T)A.spacing[fx];
    }
    if (fy >= 0) {
      pgy = A.pf_in[fidx(fy, y == 0 ? 3 : 2, x)];
      rgy = T(1) / (T)A.spacing[fy];
    }
    U[(y + 1) * NP + x + 1] = uz;
    if (x == 0) U[(y + 1) * NP] = nxl >= 0 ? A.uf_in[fidx(nxl, 1, y)] : uz;
    if (x == N - 1) U[(y + 1) * NP + N + 1] = nxh >= 0 ? A.uf_in[fidx(nxh, 0, y)] : uz;
    if (y == 0) U[x + 1] = nyl >= 0 ? A.uf_in[fidx(nyl, 3, x)] : uz;
    if (y == N - 1) U[(N + 1) * NP + x + 1] = nyh >= 0 ? A.uf_in[fidx(nyh, 2, x)] : uz;
  }
  lds_barrier();
  // (2) gradients and their ghosts
  if (act) {
    T gx = 0, gy = 0;
#pragma unroll
    for (int k = 0; k < PW; ++k) {
      gx += ltab[ix * PW + k] * U[(y + 1) * NP + (sx + k + 1)];
      gy += ltab[iy * PW + k] * U[(sy + k + 1) * NP + x + 1];
    }
    gx *= rh;
    gy *= rh;
    GX[y * NP + x + 1] = gx;
    GY[(y + 1) * N + x] = gy;
    if (x == 0) GX[y * NP] = nxl >= 0 ? (pgx + Whi * uz) * rgx : -gx;
    if (x == N - 1) GX[y * NP + N + 1] = nxh >= 0 ? (pgx + Wlo * uz) * rgx : -gx;
    if (y == 0) GY[x] = nyl >= 0 ? (pgy + Whi * uz) * rgy : -gy;
    if (y == N - 1) GY[(N + 1) * N + x] = nyh >= 0 ? (pgy + Wlo * uz) * rgy : -gy;
  }
  lds_barrier();
  // (3) the Laplacian and the stage combination
  T unext = 0, w = 0;
  if (act) {
    T lx = 0, ly = 0;
#pragma unroll
    for (int k = 0; k < PW; ++k) {
      lx += ltab[ix * PW + k] * GX[y * NP + (sx + k + 1)];
      ly += ltab[iy * PW + k] * GY[(sy + k + 1) * N + x];
    }
    const T kv = ((T(0) + lx) + ly) * rh;
    if (MODE == 1) {
      A.acc[iu] = ku;
      A.acc[iv] = kv;
      unext = yu + a * ku;
      w = yv + a * kv;
    } else if (MODE == 2) {
      A.acc[iu] = au + T(2) * ku;
      A.acc[iv] = av + T(2) * kv;
      unext = yu + a * ku;
      w = yv + a * kv;
    } else if (MODE == 3) {
      unext = yu + a * (au + ku);
      w = yv + a * (av + kv);
    } else {
      const T du = a * (au + ku);
      const T yk = du - lu, tu = yu + yk;
      A.lo[iu] = (tu - yu) - yk;
      unext = tu;
      w = a * (av + kv);
    }
    A.yo[iu] = unext;
    VA[node] = w;
    UN[node] = unext;
  }
  lds_barrier();
  // (4) faces of the next u; truncation along x
  if (act) {
    if (x == 0) {
      T s = 0;
#pragma unroll
      for (int j = 0; j < P1; ++j) s += tb.W[0][j + 1] * UN[y * N + j];
      A.uf_out[fidx(b, 0, y)] = unext;
      A.pf_out[fidx(b, 0, y)] = s;
    }
    if (x == N - 1) {
      T s = 0;
#pragma unroll
      for (int j = 0; j < P1; ++j) s += tb.W[P][j + 1] * UN[y * N + N - P1 + j];
      A.uf_out[fidx(b, 1, y)] = unext;
      A.pf_out[fidx(b, 1, y)] = s;
    }
    if (y == 0) {
      T s = 0;
#pragma unroll
      for (int j = 0; j < P1; ++j) s += tb.W[0][j + 1] * UN[j * N + x];
      A.uf_out[fidx(b, 2, x)] = unext;
      A.pf_out[fidx(b, 2, x)] = s;
    }
    if (y == N - 1) {
      T s = 0;
#pragma unroll
      for (int j = 0; j < P1; ++j) s += tb.W[P][j + 1] * UN[(N - P1 + j) * N + x];
      A.uf_out[fidx(b, 3, x)] = unext;
      A.pf_out[fidx(b, 3, x)] = s;
    }
    if (A.truncate) {
      T c = 0;
#pragma unroll
      for (int j = 0; j < P1; ++j) c += tb.dual[j] * VA[y * N + (x - ix) + j];
      VB[node] = w - ltab[P1 * PW + ix] * c;
    } else {
      VB[node] = w;
    }
  }
  lds_barrier();
  // (5) truncation along y; v out
  if (act) {
    T vv = VB[node];
    if (A.truncate) {
      T c = 0;
#pragma unroll
      for (int j = 0; j < P1; ++j) c += tb.dual[j] * VB[((y - iy) + j) * N + x];
      vv -= ltab[P1 * PW + iy] * c;
    }
    if (MODE == 4) {
      const T yk = vv - lv, tv = yv + yk;
      A.lo[iv] = (tv - yv) - yk;
      A.yo[iv] = tv;
    } else {
      A.yo[iv] = vv;
    }
  }
  (void)nbk;
}

template <class T, int MODE>
__global__ void __launch_bounds__(kBatchThreads) k_batch2(BatchArgs<T> A) {
  __shared__ T lds[kBatchLds];
  __shared__ T ltab[8 * 10 + 8];  // W rows and the top mode of the entry's order
  const BatchItem it = A.items[blockIdx.x];
  const int N = it.N;
  switch (it.P1) {
    case 2: batch_blocks<T, 2, MODE>(A, it, N, lds, ltab); break;
    case 3: batch_blocks<T, 3, MODE>(A, it, N, lds, ltab); break;
    case 4: batch_blocks<T, 4, MODE>(A, it, N, lds, ltab); break;
    case 5: batch_blocks<T, 5, MODE>(A, it, N, lds, ltab); break;
    case 6: batch_blocks<T, 6, MODE>(A, it, N, lds, ltab); break;
    case 7: batch_blocks<T, 7, MODE>(A, it, N, lds, ltab); break;
    case 8: batch_blocks<T, 8, MODE>(A, it, N, lds, ltab); break;
    default: break;
  }
}

// The face buffers of a state: per block face, the face plane of u and the
// block-local part of the face gradient.  One thread per (block, face, node).
template <class T>
__global__ void k_batch_faces(int64_t nb, const int32_t* __restrict__ P1s,
                              const int32_t* __restrict__ Ns, const int64_t* __restrict__ off,
                              const int64_t* __restrict__ foff,
                              const Tables<T>* __restrict__ tables, const T* __restrict__ y,
                              T* __restrict__ uf, T* __restrict__ pf) {
  const int64_t b = blockIdx.x;
  if (b >= nb) return;
  const int P1 = P1s[b], N = Ns[b], P = P1 - 1;
  const Tables<T>& tb = tables[P1 - 2];
  for (int q = threadIdx.x; q < 4 * N; q += blockDim.x) {
    const int f = q / N, t = q % N, d = f / 2, s = f % 2;
    const int first = s ? N - P1 : 0, row = s ? P : 0;
    T acc = 0;
    for (int j = 0; j < P1; ++j) {
      const int i = first + j;
      const int64_t l = d == 0 ? (int64_t)t * N + i : (int64_t)i * N + t;
      acc += tb.W[row][j + 1] * y[off[b] + l];
    }
    const int i = s ? N - 1 : 0;
    const int64_t l = d == 0 ? (int64_t)t * N + i : (int64_t)i * N + t;
    uf[foff[b] + q] = y[off[b] + l];
    pf[foff[b] + q] = acc;
  }
}
