// fused2.h -- one-launch RK4 stage over a batch of independent 2D problems of
// different element orders.
//
// The blocks of every problem are laid out one after another; a device-resident
// work list gives each workgroup one entry (P+1, first block, block count), so
// one launch advances every block of every problem by one stage.  Inside a
// workgroup each thread owns one or more nodes (all their global operands are
// loaded before the first barrier); a block lives in LDS with its x and y
// ghosts: u, then the two gradients, then the Laplacian, the RK4 combination
// (stepping.jl:72-78) and v's truncation (evolution.jl:238-264, x then y).
// Ghosts come from face buffers written by the previous stage (the face plane
// of u and the block-local part of the face node's gradient), as in the 3D
// fused stage.  Each problem has its own spacing and time step.

#pragma once

// Launch shape of the batch stage per precision, from timing the 512-problem
// batch (orders 2..6): FP32 one node per thread, 512 threads, compiled for 8
// waves per SIMD; FP64 two nodes per thread, 256 threads, the compiler's
// register budget.  K nodes per thread put K times the loads in flight
// between two barriers; more waves per SIMD do the same across workgroups.
#ifndef TWOLDG_BATCH_THREADS_F32
#define TWOLDG_BATCH_THREADS_F32 512
#endif
#ifndef TWOLDG_BATCH_THREADS_F64
#define TWOLDG_BATCH_THREADS_F64 256
#endif
#ifndef TWOLDG_BATCH_K_F32
#define TWOLDG_BATCH_K_F32 1
#endif
#ifndef TWOLDG_BATCH_K_F64
#define TWOLDG_BATCH_K_F64 2
#endif
#ifndef TWOLDG_BATCH_WAVES_F32
#define TWOLDG_BATCH_WAVES_F32 8
#endif
#ifndef TWOLDG_BATCH_WAVES_F64
#define TWOLDG_BATCH_WAVES_F64 0
#endif
template <class T>
constexpr int batch_threads() {
  return sizeof(T) == 4 ? TWOLDG_BATCH_THREADS_F32 : TWOLDG_BATCH_THREADS_F64;
}
template <class T>
constexpr int batch_k() {
  return sizeof(T) == 4 ? TWOLDG_BATCH_K_F32 : TWOLDG_BATCH_K_F64;
}
// waves per SIMD (0: the compiler's choice)
template <class T>
constexpr int batch_waves() {
  return sizeof(T) == 4 ? TWOLDG_BATCH_WAVES_F32 : TWOLDG_BATCH_WAVES_F64;
}
// LDS values of T per workgroup: room for the tiles of up to TH K nodes at
// N >= 6 (23 values per 4 nodes at most, at N = 6).
template <class T>
constexpr int batch_lds() {
  return (batch_threads<T>() * batch_k<T>() * 23 / 4 + 63) / 64 * 64;
}
// LDS values per block: U (the u plane with its ghost ring, later the next u),
// GX and GY (the gradients with their ghosts along their axis; GX later the
// x-truncated v), VA (the stage's v).
inline __host__ __device__ int batch_per_block(int N) {
  const int NP = N + 2;
  return NP * NP + 2 * N * NP + N * N;
}

struct BatchItem {
  int32_t P1;      // element nodes of this entry's blocks
  int32_t N;       // nodes per block side, E (P+1)
  int32_t count;   // blocks in this entry (consecutive, stored one after another)
  int32_t pad;
  int64_t first;   // first block
  int64_t off0;    // state offset of the first block (2 N^2 values per block)
  int64_t foff0;   // face-buffer offset of the first block (4 N values per block)
};

template <class T>
struct BatchArgs {
  const BatchItem* items;
  const int64_t* off;      // (nb) state offset of block b (2 N^2 values per block)
  const int64_t* foff;     // (nb) face-buffer offset (4 N values per block)
  // (nb, 4) across face 2d + s: the face-buffer index of the neighbour's facing
  // face (its tangential node 0), or -1 for a mirror; and the neighbour's
  // reciprocal spacing -- so an edge node's ghost is one load away
  const int64_t* nfo;
  const T* nrh;
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

// The general body: N = E (P+1) nodes per block side, E even.  Thread t owns
// the entry's nodes t, t + TH, ... (K of them, TH threads), node s being node
// s % N^2 of the entry's block s / N^2.
template <class T, int P1, int MODE>
__device__ void batch_blocks(const BatchArgs<T>& A, const BatchItem& it, T* lds, T* ltab) {
  constexpr int K = batch_k<T>();
  constexpr int TH = batch_threads<T>();
  constexpr int PW = P1 + 2;
  constexpr int P = P1 - 1;
  constexpr FusedTab<T, P1> tb = fused_tab<T, P1>();
  const int N = it.N, NT = N * N, NP = N + 2;
  const int per = batch_per_block(N);
  const int nslots = it.count * NT;
  const int tid = threadIdx.x;
  // the rows of W and the top mode of the entry's order, for the per-node rows
  {
    const Tables<T>& gt = A.tables[P1 - 2];
    if (tid < P1 * PW) ltab[tid] = gt.W[tid / PW][tid % PW];
    if (tid < P1) ltab[P1 * PW + tid] = gt.mode[tid];
  }
  // the entry's blocks are consecutive in the state and the face buffers
  const int64_t off0 = it.off0, foff0 = it.foff0;
  const T Wlo = tb.W[0][0], Whi = tb.W[P][P1 + 1];

  int bl[K], x[K], y[K];
  bool act[K];
  T uz[K], ku[K], yu[K], yv[K], au[K], av[K], lu[K], lv[K];
  T pgx[K], rgx[K], pgy[K], rgy[K], rh[K], ak[K];
  int64_t nbx[K], nby[K];  // the facing face of the neighbour across the node's x (y)
                           // face, -1 none or a mirror

  // (1) every global operand of the thread's nodes, then u and its ghosts into
  // LDS: no wait for a load in a later phase also waits for the stage's
  // stores (the vector memory counter covers both)
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int s = tid + k * TH;
    act[k] = s < nslots;
    const int b0 = act[k] ? s / NT : 0;
    const int node = act[k] ? s - b0 * NT : 0;
    bl[k] = b0;
    x[k] = node % N;
    y[k] = node / N;
    uz[k] = ku[k] = yu[k] = yv[k] = au[k] = av[k] = lu[k] = lv[k] = T(0);
    pgx[k] = rgx[k] = pgy[k] = rgy[k] = T(0);
    rh[k] = ak[k] = T(0);
    nbx[k] = nby[k] = -1;
    if (act[k]) {
      const int64_t b = it.first + b0;
      const int64_t iu = off0 + (int64_t)b0 * 2 * NT + node, iv = iu + NT;
      uz[k] = A.ys[iu];
      ku[k] = A.ys[iv];
      if (MODE != 1) {  // the compensated form: the high words
        yu[k] = A.yn[iu];
        yv[k] = A.yn[iv];
      }
      if (MODE == 4) {
        lu[k] = A.lo[iu];
        lv[k] = A.lo[iv];
      }
      if (MODE >= 2) {
        au[k] = A.acc[iu];
        av[k] = A.acc[iv];
      }
      rh[k] = T(1) / (T)A.spacing[b];
      ak[k] = (T)(A.coef * A.dt[b]);
      const int xk = x[k], yk = y[k];
      if (xk == 0 || xk == N - 1) {
        const int64_t f = b * 4 + (xk == 0 ? 0 : 1);
        nbx[k] = A.nfo[f];
        rgx[k] = A.nrh[f];
      }
      if (yk == 0 || yk == N - 1) {
        const int64_t f = b * 4 + (yk == 0 ? 2 : 3);
        nby[k] = A.nfo[f];
        rgy[k] = A.nrh[f];
      }
    }
  }
  T gux[K], guy[K];  // the ghost u across an edge node's x (y) face
#pragma unroll
  for (int k = 0; k < K; ++k) {
    gux[k] = guy[k] = uz[k];  // a mirror's ghost u is the own face value
    if (!act[k]) continue;
    if (nbx[k] >= 0) {
      pgx[k] = A.pf_in[nbx[k] + y[k]];
      gux[k] = A.uf_in[nbx[k] + y[k]];
    }
    if (nby[k] >= 0) {
      pgy[k] = A.pf_in[nby[k] + x[k]];
      guy[k] = A.uf_in[nby[k] + x[k]];
    }
  }
#pragma unroll
  for (int k = 0; k < K; ++k) {
    if (!act[k]) continue;
    const int xk = x[k], yk = y[k];
    T* U = lds + bl[k] * per;
    U[(yk + 1) * NP + xk + 1] = uz[k];
    if (xk == 0) U[(yk + 1) * NP] = gux[k];
    if (xk == N - 1) U[(yk + 1) * NP + N + 1] = gux[k];
    if (yk == 0) U[xk + 1] = guy[k];
    if (yk == N - 1) U[(N + 1) * NP + xk + 1] = guy[k];
  }
  lds_barrier();
  // (2) gradients and their ghosts
#pragma unroll
  for (int k = 0; k < K; ++k) {
    if (!act[k]) continue;
    const int xk = x[k], yk = y[k];
    const T* U = lds + bl[k] * per;
    T* GX = lds + bl[k] * per + NP * NP;
    T* GY = GX + N * NP;
    const int ix = xk % P1, iy = yk % P1;
    const int sx = xk - ix - 1, sy = yk - iy - 1;
    T gx = 0, gy = 0;
#pragma unroll
    for (int q = 0; q < PW; ++q) {
      gx += ltab[ix * PW + q] * U[(yk + 1) * NP + (sx + q + 1)];
      gy += ltab[iy * PW + q] * U[(sy + q + 1) * NP + xk + 1];
    }
    gx *= rh[k];
    gy *= rh[k];
    GX[yk * NP + xk + 1] = gx;
    GY[(yk + 1) * N + xk] = gy;
    // a mirror's ghost gradient is minus the own face gradient
    if (xk == 0) GX[yk * NP] = nbx[k] >= 0 ? (pgx[k] + Whi * uz[k]) * rgx[k] : -gx;
    if (xk == N - 1) GX[yk * NP + N + 1] = nbx[k] >= 0 ? (pgx[k] + Wlo * uz[k]) * rgx[k] : -gx;
    if (yk == 0) GY[xk] = nby[k] >= 0 ? (pgy[k] + Whi * uz[k]) * rgy[k] : -gy;
    if (yk == N - 1) GY[(N + 1) * N + xk] = nby[k] >= 0 ? (pgy[k] + Wlo * uz[k]) * rgy[k] : -gy;
  }
  lds_barrier();
  // (3) the Laplacian and the stage combination; the next u into the u plane's
  // place (U is dead), the stage's v into VA
  T unext[K], w[K];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    unext[k] = w[k] = T(0);
    if (!act[k]) continue;
    const int xk = x[k], yk = y[k];
    const T* GX = lds + bl[k] * per + NP * NP;
    const T* GY = GX + N * NP;
    T* VA = lds + bl[k] * per + NP * NP + 2 * N * NP;
    const int ix = xk % P1, iy = yk % P1;
    const int sx = xk - ix - 1, sy = yk - iy - 1;
    T lx = 0, ly = 0;
#pragma unroll
    for (int q = 0; q < PW; ++q) {
      lx += ltab[ix * PW + q] * GX[yk * NP + (sx + q + 1)];
      ly += ltab[iy * PW + q] * GY[(sy + q + 1) * N + xk];
    }
    const T kv = ((T(0) + lx) + ly) * rh[k];
    const int64_t iu = off0 + (int64_t)bl[k] * 2 * NT + yk * N + xk, iv = iu + NT;
    const T a = ak[k];
    if (MODE == 1) {  // stage 1: the stage input is the step start
      A.acc[iu] = ku[k];
      A.acc[iv] = kv;
      unext[k] = uz[k] + a * ku[k];
      w[k] = ku[k] + a * kv;
    } else if (MODE == 2) {
      A.acc[iu] = au[k] + T(2) * ku[k];
      A.acc[iv] = av[k] + T(2) * kv;
      unext[k] = yu[k] + a * ku[k];
      w[k] = yv[k] + a * kv;
    } else if (MODE == 3) {
      unext[k] = yu[k] + a * (au[k] + ku[k]);
      w[k] = yv[k] + a * (av[k] + kv);
    } else {  // compensated: u by Kahan now, v's increment truncated first
      const T du = a * (au[k] + ku[k]);
      const T yk2 = du - lu[k], tu = yu[k] + yk2;
      A.lo[iu] = (tu - yu[k]) - yk2;
      unext[k] = tu;
      w[k] = a * (av[k] + kv);
    }
    A.yo[iu] = unext[k];
    VA[yk * N + xk] = w[k];
  }
  // the next u goes where U was: every thread is past its reads of U (barrier
  // 2), and nobody reads UN before barrier 3
#pragma unroll
  for (int k = 0; k < K; ++k)
    if (act[k]) lds[bl[k] * per + y[k] * N + x[k]] = unext[k];
  lds_barrier();
  // (4) faces of the next u; truncation along x into VB (where GX was)
#pragma unroll
  for (int k = 0; k < K; ++k) {
    if (!act[k]) continue;
    const int xk = x[k], yk = y[k];
    const T* UN = lds + bl[k] * per;
    const T* VA = lds + bl[k] * per + NP * NP + 2 * N * NP;
    T* VB = lds + bl[k] * per + NP * NP;
    const int64_t fb = foff0 + (int64_t)bl[k] * 4 * N;
    if (xk == 0) {
      T s = 0;
#pragma unroll
      for (int j = 0; j < P1; ++j) s += tb.W[0][j + 1] * UN[yk * N + j];
      A.uf_out[fb + 0 * N + yk] = unext[k];
      A.pf_out[fb + 0 * N + yk] = s;
    }
    if (xk == N - 1) {
      T s = 0;
#pragma unroll
      for (int j = 0; j < P1; ++j) s += tb.W[P][j + 1] * UN[yk * N + N - P1 + j];
      A.uf_out[fb + 1 * N + yk] = unext[k];
      A.pf_out[fb + 1 * N + yk] = s;
    }
    if (yk == 0) {
      T s = 0;
#pragma unroll
      for (int j = 0; j < P1; ++j) s += tb.W[0][j + 1] * UN[j * N + xk];
      A.uf_out[fb + 2 * N + xk] = unext[k];
      A.pf_out[fb + 2 * N + xk] = s;
    }
    if (yk == N - 1) {
      T s = 0;
#pragma unroll
      for (int j = 0; j < P1; ++j) s += tb.W[P][j + 1] * UN[(N - P1 + j) * N + xk];
      A.uf_out[fb + 3 * N + xk] = unext[k];
      A.pf_out[fb + 3 * N + xk] = s;
    }
    const int ix = xk % P1;
    if (A.truncate) {
      T c = 0;
#pragma unroll
      for (int j = 0; j < P1; ++j) c += tb.dual[j] * VA[yk * N + (xk - ix) + j];
      VB[yk * N + xk] = w[k] - ltab[P1 * PW + ix] * c;
    } else {
      VB[yk * N + xk] = w[k];
    }
  }
  lds_barrier();
  // (5) truncation along y; v out
#pragma unroll
  for (int k = 0; k < K; ++k) {
    if (!act[k]) continue;
    const int xk = x[k], yk = y[k];
    const T* VB = lds + bl[k] * per + NP * NP;
    const int iy = yk % P1;
    T vv = VB[yk * N + xk];
    if (A.truncate) {
      T c = 0;
#pragma unroll
      for (int j = 0; j < P1; ++j) c += tb.dual[j] * VB[((yk - iy) + j) * N + xk];
      vv -= ltab[P1 * PW + iy] * c;
    }
    const int64_t iv = off0 + (int64_t)bl[k] * 2 * NT + NT + yk * N + xk;
    if (MODE == 4) {
      const T yk2 = vv - lv[k], tv = yv[k] + yk2;
      A.lo[iv] = (tv - yv[k]) - yk2;
      A.yo[iv] = tv;
    } else {
      A.yo[iv] = vv;
    }
  }
}

template <class T, int MODE>
__global__ void __launch_bounds__(batch_threads<T>(), batch_waves<T>()) k_batch2(BatchArgs<T> A) {
  __shared__ T lds[batch_lds<T>()];
  __shared__ T ltab[8 * 10 + 8];  // W rows and the top mode of the entry's order
  const BatchItem it = A.items[blockIdx.x];
  switch (it.P1) {
    case 2: batch_blocks<T, 2, MODE>(A, it, lds, ltab); break;
    case 3: batch_blocks<T, 3, MODE>(A, it, lds, ltab); break;
    case 4: batch_blocks<T, 4, MODE>(A, it, lds, ltab); break;
    case 5: batch_blocks<T, 5, MODE>(A, it, lds, ltab); break;
    case 6: batch_blocks<T, 6, MODE>(A, it, lds, ltab); break;
    case 7: batch_blocks<T, 7, MODE>(A, it, lds, ltab); break;
    case 8: batch_blocks<T, 8, MODE>(A, it, lds, ltab); break;
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
