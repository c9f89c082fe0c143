// fused2.h -- one-launch RK4 stage over a batch of independent 2D problems of
// different element orders.
//
// The blocks of every problem are laid out one after another, grouped in runs
// of equal order and block size; each run is cut into tiles of a fixed number
// of blocks, one tile per workgroup, and the workgroup finds its tile from the
// run table in the kernel arguments (no load), so one launch advances every
// block of every problem by one stage.  Inside a workgroup each thread owns one
// or more nodes (all their global operands are loaded before the first
// barrier); a block lives in LDS with its x and y ghosts: u, then the two
// gradients, then the Laplacian, the RK4 combination (stepping.jl:72-78) and
// v's truncation (evolution.jl:238-264, x then y).
//
// Ghosts are pushed: at the end of a stage a block writes each of its faces
// (the face plane of the new u and the block-local part of the face node's
// gradient) into the ghost slots of the neighbour across that face, in the
// other of two ghost buffers; the next stage reads its own slots, at an offset
// fixed by its block index.  So every global load of a stage depends only on
// the workgroup's index: one round of loads, not a chain.  Each problem has its
// own spacing and time step.
#pragma once

// Launch shape of the batch stage: threads per workgroup and waves per SIMD
// (0: the compiler's register budget) per precision.
#ifndef TWOLDG_BATCH_THREADS_F32
#define TWOLDG_BATCH_THREADS_F32 256
#endif
#ifndef TWOLDG_BATCH_THREADS_F64
#define TWOLDG_BATCH_THREADS_F64 256
#endif
#ifndef TWOLDG_BATCH_WAVES_F32
#define TWOLDG_BATCH_WAVES_F32 0
#endif
#ifndef TWOLDG_BATCH_WAVES_F64
#define TWOLDG_BATCH_WAVES_F64 0
#endif
template <class T>
constexpr int batch_threads() {
  return sizeof(T) == 4 ? TWOLDG_BATCH_THREADS_F32 : TWOLDG_BATCH_THREADS_F64;
}
template <class T>
constexpr int batch_waves() {
  return sizeof(T) == 4 ? TWOLDG_BATCH_WAVES_F32 : TWOLDG_BATCH_WAVES_F64;
}
// The element counts per block side a batch takes.
constexpr bool batch_e_ok(int E) { return E == 2 || E == 4; }
// Blocks of N^2 nodes (E elements per side) per workgroup at most: a thread
// per element line along x and one along y, E N of each per block.
inline __host__ __device__ int batch_max_tile(int N, int E, int threads) {
  return threads / (E * N);
}
// LDS values of T per workgroup: two planes per block (u, then the next u;
// the x-Laplacian sums, then v for its truncation), at most TH / (E N) blocks
// of 2 N^2 values: 2 TH P1 values, P1 <= 8.
template <class T>
constexpr int batch_lds() {
  return batch_threads<T>() * 16;
}

// One workgroup's tile: consecutive blocks of one order and block size.
struct BatchItem {
  int32_t P1;      // element nodes of the tile's blocks
  int32_t N;       // nodes per block side, E (P+1)
  int32_t count;   // blocks in the tile
  int64_t first;   // first block
  int64_t off0;    // state offset of the first block (2 N^2 values per block)
  int64_t foff0;   // ghost-slot offset of the first block (4 N values per block)
};

// A run of consecutive blocks of one order and block size, cut into tiles of
// `tile` blocks; its tiles are the workgroups wg0, wg0 + 1, ...  Unused entries
// of the table have wg0 beyond every workgroup.
struct BatchRun {
  int32_t P1, N, tile, pad;
  int64_t wg0;     // the run's first workgroup
  int64_t first;   // first block
  int64_t end;     // one past the last block
  int64_t off0;    // state offset of the first block
  int64_t foff0;   // ghost-slot offset of the first block
};
constexpr int kBatchMaxRuns = 16;

template <class T>
struct BatchArgs {
  BatchRun runs[kBatchMaxRuns];
  // (nb, 4) across face 2d + s: the ghost slots (tangential node 0) of the
  // neighbour's facing face, where this block's face goes, or -1 for a mirror;
  // and the neighbour's reciprocal spacing, 0 for a mirror
  const int32_t* nfo;
  const T* nrh;
  const double* spacing;   // (nb)
  const double* dt;        // (nb) the block's problem's time step
  const Tables<T>* tables; // [P1 - 2]: tables of order P1 - 1
  const T* ys;
  const T* yn;
  T* acc;
  T* yo;
  T* lo;
  const T* uf_in;          // this stage's ghosts: per block 4 N slots, face 2d + s
  const T* pf_in;          //   at 2d + s N (u on the face plane, gradient part)
  T* uf_out;               // the next stage's ghosts (the neighbours' slots)
  T* pf_out;
  double coef;             // stage coefficient in units of dt: 1/2, 1 or 1/6
  double coef2;            // mode 13: z's coefficient of k3, 1/3
  int truncate;
};

// The derivative along one element line of a block (x or y): the gradient at
// the element's P1 nodes and at the two nodes beyond it (the neighbour
// elements' face nodes, or the block face's ghost gradient), then the
// Laplacian sums lap[j] (evolution.jl:72-110 along one dimension).  rd(c) reads
// u at node c - 1 along the line (c = 1..N); ulo/uhi give the ghost u beyond
// the block's faces, glo/ghi the ghost gradient there from the own face value
// and gradient.
template <class T, int P1, int E, class RD, class UL, class UH, class GL, class GH>
__device__ __forceinline__ void batch_line_lap(int el, RD rd, UL ulo, UH uhi, GL glo, GH ghi,
                                               T rh, T* lap) {
  constexpr FusedTab<T, P1> tb = fused_tab<T, P1>();
  constexpr int PW = P1 + 2, P = P1 - 1;
  T uw[PW];
  uw[0] = el == 0 ? ulo() : rd(el * P1);
#pragma unroll
  for (int k = 1; k < PW - 1; ++k) uw[k] = rd(el * P1 + k);
  uw[PW - 1] = el == E - 1 ? uhi() : rd(el * P1 + PW - 1);
  T g[PW];
#pragma unroll
  for (int j = 0; j < P1; ++j) {
    T s = 0;
#pragma unroll
    for (int k = 0; k < PW; ++k) {
      if ((k == 0 && j != 0) || (k == PW - 1 && j != P)) continue;
      s += tb.W[j][k] * uw[k];
    }
    g[j + 1] = s * rh;
  }
  if (el > 0) {  // row P of the element below: its nodes, then our first node
    T s = 0;
#pragma unroll
    for (int k = 1; k < PW - 1; ++k) s += tb.W[P][k] * rd((el - 1) * P1 + k);
    s += tb.W[P][PW - 1] * uw[1];
    g[0] = s * rh;
  } else {
    g[0] = glo(uw[1], g[1]);
  }
  if (el < E - 1) {  // row 0 of the element above: our last node, then its nodes
    T s = tb.W[0][0] * uw[P1];
#pragma unroll
    for (int k = 1; k < PW - 1; ++k) s += tb.W[0][k] * rd((el + 1) * P1 + k);
    g[PW - 1] = s * rh;
  } else {
    g[PW - 1] = ghi(uw[P1], g[P1]);
  }
#pragma unroll
  for (int j = 0; j < P1; ++j) {
    T s = 0;
#pragma unroll
    for (int k = 0; k < PW; ++k) {
      if ((k == 0 && j != 0) || (k == PW - 1 && j != P)) continue;
      s += tb.W[j][k] * g[k];
    }
    lap[j] = s;
  }
}

// One tile: blocks of N = E P1 nodes per side.  Thread t of the workgroup
// serves block t / (E N) of the tile in two roles: the y line (column xy,
// element ey), whose P1 nodes it loads, combines and stores, and the x line
// (element ex, row yx), whose x-derivative and x-truncation it forms in LDS.
// Every global load is issued in phase 1; phase 3 stores as it goes (nothing
// is loaded after it, so no wait covers a store).
template <class T, int P1, int E, int MODE>
__device__ void batch_lines(const BatchArgs<T>& A, const BatchItem& it, T* lds) {
  constexpr int N = P1 * E, NT = N * N, LB = E * N;
  constexpr int P = P1 - 1;
  constexpr FusedTab<T, P1> tb = fused_tab<T, P1>();
  static_assert(batch_threads<T>() / LB * 2 * NT <= batch_lds<T>(), "tile exceeds the LDS");
  const int tid = threadIdx.x;
  const int b0 = tid / LB, r = tid - b0 * LB;
  const bool act = b0 < it.count;
  const int xy = r % N, ey = r / N;  // y-line role
  const int ex = r % E, yx = r / E;  // x-line role
  T* U = lds + b0 * 2 * NT;  // u, then the next u
  T* L = U + NT;             // the x-Laplacian sums, then v for its x-truncation
  // the tile's arrays from uniform bases, per-thread 32-bit offsets
  const T* ys = A.ys + it.off0;
  const T* yn = A.yn + it.off0;
  T* acc = A.acc + it.off0;
  T* yo = A.yo + it.off0;
  T* lo = A.lo ? A.lo + it.off0 : nullptr;
  const T* ufi = A.uf_in + it.foff0;
  const T* pfi = A.pf_in + it.foff0;
  const int32_t* nfo = A.nfo + it.first * 4;
  const T* nrh = A.nrh + it.first * 4;
  const unsigned ob = (unsigned)b0 * 2 * NT;  // the block's state
  const unsigned gb = (unsigned)b0 * 4 * N;   // its ghost slots
  auto node = [&](int j) { return (ey * P1 + j) * N + xy; };  // y line's node j in the plane
  const T Wlo = tb.W[0][0], Whi = tb.W[P][P1 + 1];

  // (1) every global operand: the y line's u, v, the step start, the
  // accumulator (and compensation words), the block's spacing and time step,
  // and the ghosts of the faces the two lines end on
  T rh = 0, ak = 0, ak2 = 0;
  T uy[P1], ku[P1], yu[P1], yv[P1], au[P1], av[P1], lu[P1], lv[P1];
  T gyu = 0, gyp = 0, gyr = 0, gxu = 0, gxp = 0, gxr = 0;
  int32_t nby = -1, nbx = -1;  // where the y (x) line's face goes, -1 a mirror
#pragma unroll
  for (int j = 0; j < P1; ++j) uy[j] = ku[j] = yu[j] = yv[j] = au[j] = av[j] = lu[j] = lv[j] = T(0);
  if (act) {
    rh = T(1) / (T)A.spacing[it.first + b0];
    ak = (T)(A.coef * A.dt[it.first + b0]);
    ak2 = (T)(A.coef2 * A.dt[it.first + b0]);
#pragma unroll
    for (int j = 0; j < P1; ++j) {
      const unsigned o = ob + node(j);
      uy[j] = ys[o];
      ku[j] = ys[o + NT];
      if (MODE != 1 && MODE != 8 && MODE != 11) {  // the compensated form: the high words
        yu[j] = yn[o];
        yv[j] = yn[o + NT];
      }
      if (MODE == 4) {
        lu[j] = lo[o];
        lv[j] = lo[o + NT];
      }
      // stage 2 (MODE 6): the accumulator's u is y_n's v; MODE 13: y_2
      if (MODE == 2 || MODE == 4 || MODE == 6 || MODE == 8 || MODE == 13) {
        au[j] = MODE == 6 ? yv[j] : acc[o];
        av[j] = acc[o + NT];
      }
    }
    // an end face's ghosts sit in the block's own slots (read whatever the
    // face, so that no load waits for another; a mirror's are not used)
    if (ey == 0 || ey == E - 1) {
      const int f = ey == 0 ? 2 : 3;
      nby = nfo[b0 * 4 + f];
      gyr = nrh[b0 * 4 + f];
      gyu = ufi[gb + f * N + xy];
      gyp = pfi[gb + f * N + xy];
    }
    if (ex == 0 || ex == E - 1) {
      const int f = ex == 0 ? 0 : 1;
      nbx = nfo[b0 * 4 + f];
      gxr = nrh[b0 * 4 + f];
      gxu = ufi[gb + f * N + yx];
      gxp = pfi[gb + f * N + yx];
    }
#pragma unroll
    for (int j = 0; j < P1; ++j) U[node(j)] = uy[j];
  }
  lds_barrier();
  // (2) x line: the x-Laplacian sums to L; y line: its own, kept
  T ly[P1];
  if (act) {
    {
      const T* Ur = U + yx * N - 1;
      T lx[P1];
      batch_line_lap<T, P1, E>(
          ex, [&](int c) { return Ur[c]; },
          [&]() { return gxr != T(0) ? gxu : Ur[1]; },  // a mirror's ghost u is the face value
          [&]() { return gxr != T(0) ? gxu : Ur[N]; },
          [&](T uf, T gf) { return gxr != T(0) ? (gxp + Whi * uf) * gxr : -gf; },
          [&](T uf, T gf) { return gxr != T(0) ? (gxp + Wlo * uf) * gxr : -gf; }, rh, lx);
#pragma unroll
      for (int j = 0; j < P1; ++j) L[yx * N + ex * P1 + j] = lx[j];
    }
    const T* Uc = U - N + xy;
    batch_line_lap<T, P1, E>(
        ey, [&](int c) { return Uc[c * N]; },
        [&]() { return gyr != T(0) ? gyu : Uc[N]; },
        [&]() { return gyr != T(0) ? gyu : Uc[N * N]; },
        [&](T uf, T gf) { return gyr != T(0) ? (gyp + Whi * uf) * gyr : -gf; },
        [&](T uf, T gf) { return gyr != T(0) ? (gyp + Wlo * uf) * gyr : -gf; }, rh, ly);
  }
  lds_barrier();
  // (3) y line: the Laplacian, the stage combination (stepping.jl:72-78) and its
  // stores; the next u into U (this thread's nodes only), v truncated along y
  // into L, the y faces of the next u into the neighbours' ghost slots
  if (act) {
    T w[P1];
    T plo = 0, phi = 0, ulo = 0, uhi = 0;
#pragma unroll
    for (int j = 0; j < P1; ++j) {
      const int l = node(j);
      const unsigned o = ob + l;
      const T kv = ((T(0) + L[l]) + ly[j]) * rh;
      T unext;
      if (MODE == 1) {  // stage 1: the stage input is the step start; the
        // accumulator's u would be k1's u = y_n's v, which stage 2 reads itself
        acc[o + NT] = kv;
        unext = uy[j] + ak * ku[j];
        w[j] = ku[j] + ak * kv;
      } else if (MODE == 2 || MODE == 6) {
        acc[o] = au[j] + T(2) * ku[j];
        acc[o + NT] = av[j] + T(2) * kv;
        unext = yu[j] + ak * ku[j];
        w[j] = yv[j] + ak * kv;
      } else if (MODE == 11) {  // FP64 stage 1: y_2 only
        unext = uy[j] + ak * ku[j];
        w[j] = ku[j] + ak * kv;
      } else if (MODE == 12) {  // FP64 stage 2: y_3 only
        unext = yu[j] + ak * ku[j];
        w[j] = yv[j] + ak * kv;
      } else if (MODE == 13) {
        // FP64 stage 3: y_4, and z = (y_2 + 2 y_3)/3 + dt/3 k3 over y_2 (in acc)
        constexpr T third = T(1) / T(3);
        acc[o] = (au[j] + T(2) * uy[j]) * third + ak2 * ku[j];
        acc[o + NT] = (av[j] + T(2) * ku[j]) * third + ak2 * kv;
        unext = yu[j] + ak * ku[j];
        w[j] = yv[j] + ak * kv;
      } else if (MODE == 8) {  // the final stage from z (no y_n)
        unext = au[j] + ak * ku[j];
        w[j] = av[j] + ak * kv;
      } else {  // compensated: u by Kahan now, v's increment truncated first
        const T du = ak * (au[j] + ku[j]);
        const T yk = du - lu[j], tu = yu[j] + yk;
        lo[o] = (tu - yu[j]) - yk;
        unext = tu;
        w[j] = ak * (av[j] + kv);
      }
      yo[o] = unext;
      U[l] = unext;
      if (j == 0) ulo = unext;
      if (j == P) uhi = unext;
      plo += tb.W[0][j + 1] * unext;
      phi += tb.W[P][j + 1] * unext;
    }
    if (nby >= 0) {
      A.uf_out[nby + xy] = ey == 0 ? ulo : uhi;
      A.pf_out[nby + xy] = ey == 0 ? plo : phi;
    }
    if (A.truncate) {
      T c = 0;
#pragma unroll
      for (int j = 0; j < P1; ++j) c += tb.dual[j] * w[j];
#pragma unroll
      for (int j = 0; j < P1; ++j) w[j] -= tb.mode[j] * c;
    }
#pragma unroll
    for (int j = 0; j < P1; ++j) L[node(j)] = w[j];  // L at these nodes was read above, by us
  }
  lds_barrier();
  // (4) x line: v's truncation along x; the x faces of the next u
  if (act) {
    const int l = yx * N + ex * P1;
    if (A.truncate) {
      T w[P1];
      T c = 0;
#pragma unroll
      for (int j = 0; j < P1; ++j) {
        w[j] = L[l + j];
        c += tb.dual[j] * w[j];
      }
#pragma unroll
      for (int j = 0; j < P1; ++j) L[l + j] = w[j] - tb.mode[j] * c;
    }
    if (nbx >= 0) {
      T s = 0;
#pragma unroll
      for (int j = 0; j < P1; ++j) s += (ex == 0 ? tb.W[0][j + 1] : tb.W[P][j + 1]) * U[l + j];
      A.uf_out[nbx + yx] = U[l + (ex == 0 ? 0 : P)];
      A.pf_out[nbx + yx] = s;
    }
  }
  lds_barrier();
  // (5) y line: v out
  if (act) {
#pragma unroll
    for (int j = 0; j < P1; ++j) {
      const unsigned o = ob + node(j) + NT;
      const T vv = L[node(j)];
      if (MODE == 4) {
        const T yk = vv - lv[j], tv = yv[j] + yk;
        lo[o] = (tv - yv[j]) - yk;
        yo[o] = tv;
      } else {
        yo[o] = vv;
      }
    }
  }
}

template <class T, int MODE>
__global__ void __launch_bounds__(batch_threads<T>(), batch_waves<T>()) k_batch2(BatchArgs<T> A) {
  __shared__ T lds[batch_lds<T>()];
  // the workgroup's run: the last that starts at or before it
  BatchRun R = A.runs[0];
#pragma unroll
  for (int r = 1; r < kBatchMaxRuns; ++r)
    if ((int64_t)blockIdx.x >= A.runs[r].wg0) R = A.runs[r];
  const int64_t j = (int64_t)blockIdx.x - R.wg0;
  BatchItem it;
  it.P1 = R.P1;
  it.N = R.N;
  it.first = R.first + j * R.tile;
  it.count = (int32_t)min((int64_t)R.tile, R.end - it.first);
  it.off0 = R.off0 + j * R.tile * 2 * (int64_t)R.N * R.N;
  it.foff0 = R.foff0 + j * R.tile * 4 * (int64_t)R.N;
  const int E = R.N / R.P1;
#define TW_BATCH_CASE(P1_)                                                 \
  case P1_:                                                                \
    if (E == 2)                                                            \
      batch_lines<T, P1_, 2, MODE>(A, it, lds);                            \
    else                                                                   \
      batch_lines<T, P1_, 4, MODE>(A, it, lds);                            \
    break;
  switch (it.P1) {
    TW_BATCH_CASE(2)
    TW_BATCH_CASE(3)
    TW_BATCH_CASE(4)
    TW_BATCH_CASE(5)
    TW_BATCH_CASE(6)
    TW_BATCH_CASE(7)
    TW_BATCH_CASE(8)
    default: break;
  }
#undef TW_BATCH_CASE
}

// The ghosts of a state, pushed as the stage does: per block face, the face
// plane of u and the block-local part of the face gradient, into the ghost
// slots of the neighbour across the face (nfo; none across a mirror).  One
// thread per (block, face, node).
template <class T>
__global__ void k_batch_faces(int64_t nb, const int32_t* __restrict__ P1s,
                              const int32_t* __restrict__ Ns, const int64_t* __restrict__ off,
                              const int32_t* __restrict__ nfo,
                              const Tables<T>* __restrict__ tables, const T* __restrict__ y,
                              T* __restrict__ uf, T* __restrict__ pf) {
  const int64_t b = blockIdx.x;
  if (b >= nb) return;
  const int P1 = P1s[b], N = Ns[b], P = P1 - 1;
  const Tables<T>& tb = tables[P1 - 2];
  for (int q = threadIdx.x; q < 4 * N; q += blockDim.x) {
    const int f = q / N, t = q % N, d = f / 2, s = f % 2;
    const int32_t dst = nfo[b * 4 + f];
    if (dst < 0) continue;
    const int first = s ? N - P1 : 0, row = s ? P : 0;
    T acc = 0;
    for (int j = 0; j < P1; ++j) {
      const int i = first + j;
      const int64_t l = d == 0 ? (int64_t)t * N + i : (int64_t)i * N + t;
      acc += tb.W[row][j + 1] * y[off[b] + l];
    }
    const int i = s ? N - 1 : 0;
    const int64_t l = d == 0 ? (int64_t)t * N + i : (int64_t)i * N + t;
    uf[dst + t] = y[off[b] + l];
    pf[dst + t] = acc;
  }
}
