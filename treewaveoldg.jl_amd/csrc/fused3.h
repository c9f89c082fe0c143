// fused3.h -- one-launch RK4 stage for 3D forests whose faces are all same-level
// (periodic included) or reflecting walls, flat metric.
//
// One workgroup per block, one thread per (x, y) column of the block:
//   * the column's u is held in registers; its z-gradient and z-Laplacian are
//     formed there (evolution.jl:72-110 along d = 3);
//   * plane by plane, u goes through LDS for the x- and y-gradients and
//     Laplacians;
//   * the RK4 combination (stepping.jl:72-78) is applied per node, and v's
//     truncation (evolution.jl:238-264, x then y then z) per element layer;
//   * the ghosts come from compact face buffers the previous stage wrote: per
//     block face, the face plane of u and the block-local part of the face
//     node's gradient (the wide derivative's row 0 or P without the ghost
//     term).  A neighbour's face gradient is that part plus the ghost term
//     formed with our own face value.
// So one stage reads each of its operands (the stage state, y_n, the accumulator
// or the other stage state) once and writes each result once (26 values per
// node and step in FP64, 36 in FP32c), plus faces, O(1/N) of that.

#pragma once

// Waves per SIMD the fused stage is compiled for: four (128 VGPRs) lets two
// workgroups of the 20^3-node blocks share a CU, which hides the latency of
// one's barriers and loads behind the other's work.  Both types load the stage
// operands of a y line one node ahead (NB 1 for FP64).
#ifndef TWOLDG_FUSED_WAVES_F64
#define TWOLDG_FUSED_WAVES_F64 4
#endif
#ifndef TWOLDG_FUSED_WAVES_F32
#define TWOLDG_FUSED_WAVES_F32 4
#endif
#ifndef TWOLDG_FUSED_NB_F64
#define TWOLDG_FUSED_NB_F64 1
#endif
// Waves per SIMD that the workgroups the LDS admits on one CU can bring (the
// most loaded SIMD), capped by the above.
template <class T, int P1, int E>
constexpr int fused_waves_per_eu() {
  constexpr int N = P1 * E;
  constexpr long lds = (long)sizeof(T) * (4 * P1 * N * N + 8 * P1 * N);
  constexpr long wgs = (160L * 1024) / lds;
  constexpr long by_lds = (wgs * ((N * N + 63) / 64) + 3) / 4;
  constexpr int want = sizeof(T) == 8 ? TWOLDG_FUSED_WAVES_F64 : TWOLDG_FUSED_WAVES_F32;
  return by_lds < 1 ? 1 : (by_lds < want ? (int)by_lds : want);
}

// Whether the next element layer's column stays in registers from its first
// read, or is read again.
#ifndef TWOLDG_FUSED_CARRY_F64
#define TWOLDG_FUSED_CARRY_F64 1
#endif
#ifndef TWOLDG_FUSED_CARRY_F32
#define TWOLDG_FUSED_CARRY_F32 1
#endif
// Whether the stage update's outputs wait in LDS (in the slots of the Laplacian
// sums it has just consumed) and are stored by the columns in the last phase.
// The vector memory counter covers stores as well as loads, so stores issued
// between the stage operand loads make every wait for an operand wait for them
// too.
#ifndef TWOLDG_FUSED_DEFER
#define TWOLDG_FUSED_DEFER 1
#endif
#ifndef TWOLDG_FUSED_NT
#define TWOLDG_FUSED_NT 0
#endif
// Streaming access for the values a stage touches once (the stage operands and
// its outputs): marked nontemporal, they leave the L2 to the data that is read
// twice (the next element layer of u, the face buffers).
template <class T>
__device__ __forceinline__ T ld_once(const T* p) {
  if (TWOLDG_FUSED_NT) return __builtin_nontemporal_load(p);
  return *p;
}
template <class T>
__device__ __forceinline__ void st_once(T* p, T v) {
  if (TWOLDG_FUSED_NT)
    __builtin_nontemporal_store(v, p);
  else
    *p = v;
}

// The v arrays may be of another type TV and layout than u's: the split form
// (u in FP64, v in FP32) keeps v in separate (nb, N^3) arrays (vstride N^3);
// the other forms point them at the v halves of the state arrays (vstride 2 N^3).
template <class T, class TV = T>
struct FusedArgs {
  const T* ys;       // stage input (u, v)
  const T* yn;       // step start (the same as ys in mode 1)
  T* acc;            // RK accumulator
  T* yo;             // output state
  T* lo;             // compensation of yn's u (final stage, compensated form)
  const TV* vs;      // v of ys, of yn (its high word in the split form), of the
  const TV* vn;      //   accumulator and of the output, block 0's; block b at
  TV* vacc;          //   + b vstride
  TV* vo;
  TV* vlo;           // compensation of yn's v (final stage, compensated and split forms)
  int64_t vstride;
  const T* uf_in;    // faces of ys: (nb, 6, N^2) u on the face plane
  const T* pf_in;    //              block-local part of the face gradient (= uf_in + nfb)
  T* uf_out;         // faces of yo
  T* pf_out;
  const int32_t* nbr;       // (nb, 6) same-level neighbour, -1 a mirror, -2 coarse-fine
  const T* gu;              // complete ghosts of the coarse-fine faces (u, gradient),
  const T* gg;              //   (nb, 6, N^2) each, gg = gu + nfb
  int64_t nfb;              // nb 6 N^2: pf_in - uf_in, gg - gu
  const double* spacing;    // (nb)
  const Tables<T>* tables;
  T a;               // stage coefficient (dt/2, dt or dt/6)
  T a2;              // modes 13, 15: dt/3
  int truncate;
};

// Face index f = 2d + s; tangential index t of node (x, y, z):
// d = 0: y + N z; d = 1: x + N z; d = 2: x + N y.

// The element's operators as compile-time constants: after unrolling, every
// coefficient is an immediate that the compiler materializes in scalar
// registers where it is used (and skips where it is zero), so the tables take
// neither vector registers nor LDS.
template <class T, int P1>
struct FusedTab {
  T W[P1][P1 + 2];
  T mode[P1];
  T dual[P1];
};
template <class T, int P1>
constexpr FusedTab<T, P1> fused_tab() {
  FusedTab<T, P1> r{};
  const twoldg::ElementTables& et = twoldg::kElementTables[P1 - 2];
  for (int j = 0; j < P1; ++j) {
    for (int k = 0; k < P1 + 2; ++k) r.W[j][k] = (T)et.stencil[j][k];
    r.mode[j] = (T)et.topmode[j];
    r.dual[j] = (T)et.topdual[j];
  }
  return r;
}

// Face ghosts are formed here from a same-level neighbour's face buffers (nbr
// >= 0) or from the own face on a mirror (-1); with GHOST, a coarse-fine face
// (-2) takes them complete from the ghost slots k_cf_ghosts fills.
// MODE selects the stage.  The compensated form (FP32c) keeps three registers:
// 1, stage 1 (ys = yn); 6, stage 2; 2, stage 3; 4, the final stage (u by a
// Kahan sum into (yn, lo), v's truncated increment likewise into (vn, vlo)).
// The accumulator's u after stage 1 would be k1's u, which is y_n's v: stage 1
// stores only the accumulator's v, and stage 2 forms acc_u = v_n + 2 k2_u from
// the y_n it reads anyway.
// FP64 keeps no accumulator: 11, stage 1, and 12, stage 2, store only the next
// stage state (y_2, then y_3 beside it); 13, stage 3, forms y_4 and, from y_2
// and y_3, z = (y_2 + 2 y_3)/3 + dt/3 k3 (= y_n + dt/6 (k1 + 2 k2 + 2 k3), the
// k's being (y_2 - y_n) 2/dt and (y_3 - y_n) 2/dt) in y_2's place; 8, the final
// stage, forms y_n+1 = z + dt/6 k4: 26 values per node and step.
// The split form does so for u and keeps the accumulator for v: 1, stage 1;
// 14, stage 2 (y_3 beside y_2, v's accumulator); 15, stage 3 (u as in 13, v's
// accumulator); 10, the final stage (u from z, v's increment added to the FP32
// pair (vn, vlo) in FP64).
template <class T, int P1, int E, int MODE, bool GHOST, class TV = T>
__global__ void __launch_bounds__(((P1 * E) * (P1 * E) + 63) / 64 * 64, (fused_waves_per_eu<T, P1, E>()))
k_fused3(FusedArgs<T, TV> A) {
  constexpr FusedTab<T, P1> tb = fused_tab<T, P1>();
  constexpr int N = P1 * E;
  constexpr int NT = N * N;  // = E * N * P1: one element line per thread and role
  constexpr int NTH = (NT + 63) / 64 * 64;
  constexpr int PW = P1 + 2;
  constexpr int P = P1 - 1;
  constexpr int NG = 8 * P1 * N;             // x/y ghost values of one layer
  constexpr int GPT = (NG + NTH - 1) / NTH;  // per thread
  constexpr int N2 = N * N, nn = N2 * N;
  // One element layer (P1 planes) at a time, each plane row-major with row
  // stride N (conflict-free banks for the line roles).  RA holds the u planes,
  // later the next u; LZ the z-Laplacian sums, LX the x-Laplacian sums, VB the
  // values to truncate; GH the x/y ghosts of the layer: per block face (x low,
  // x high, y low, y high), the ghost u, then the neighbour's partial face
  // gradient, at [plane * N + tangential index].
  __shared__ T RA[P1 * NT];
  __shared__ T LZ[P1 * NT];
  __shared__ T LX[P1 * NT];
  __shared__ T VB[P1 * NT];
  __shared__ T GH[8][P1 * N];
  T* U = RA;
  T* UN = RA;

  const int64_t b = blockIdx.x;
  const int tid = threadIdx.x;
  const bool act = tid < NT;
  // column role: (x, y); x-line role: element ex of row yx in plane px;
  // y-line role: element ey of column xy in plane py
  const int x = act ? tid % N : 0, y = act ? tid / N : 0;
  // The line roles' indices are formed in each phase from an opaque copy of
  // the thread index: hoisted out of the layer loop they would hold registers
  // for its whole length.
  auto opaque = [](int v) {
    asm volatile("" : "+v"(v));
    return v;
  };

  const T h = (T)ld_table(A.spacing + b);
  const T rh = T(1) / h;
  const int32_t* nb6 = A.nbr + b * 6;
  // the block's arrays: uniform bases, per-thread 32-bit offsets (z stride N2,
  // v at +nn)
  const int col = y * N + x;
  const T* us = A.ys + b * 2 * nn;
  const T* un_ = A.yn + b * 2 * nn;
  T* acc = A.acc + b * 2 * nn;
  T* yo = A.yo + b * 2 * nn;
  T* lo = A.lo ? A.lo + b * 2 * nn : nullptr;
  // v at the same per-thread offsets as u
  const TV* vs = A.vs + b * A.vstride;
  const TV* vn = A.vn + b * A.vstride;
  TV* vacc = A.vacc + b * A.vstride;
  TV* vo = A.vo + b * A.vstride;
  TV* vlo = A.vlo ? A.vlo + b * A.vstride : nullptr;
  T* ufo = A.uf_out + b * 6 * N2;
  T* pfo = A.pf_out + b * 6 * N2;
  const int nxl = ld_table(nb6 + 0), nxh = ld_table(nb6 + 1), nyl = ld_table(nb6 + 2),
            nyh = ld_table(nb6 + 3), nzl = ld_table(nb6 + 4), nzh = ld_table(nb6 + 5);
  // the reciprocal spacing of a same-level neighbour, read where it is used
  // the reciprocal spacings of the same-level neighbours, per face, through LDS:
  // read there, they wait on the LDS counter, not on the stage's stores
  __shared__ T RSP[6];
  auto rsp = [&](int f) { return RSP[f]; };
  constexpr int CF = -2;  // a coarse-fine face

  auto load_layer = [&](int e, T* dst) {
#pragma unroll
    for (int j = 0; j < P1; ++j) dst[j] = act ? us[col + (e * P1 + j) * N2] : T(0);
  };
  // the ghost terms of rows 0 and P
  const T Wlo = tb.W[0][0], Whi = tb.W[P][P1 + 1];

  // The derivative along a line of one element (x or y): the gradient at the
  // element's P1 nodes and at the two nodes beyond it (the neighbour elements'
  // face nodes, or the block face's ghost gradient), then the Laplacian sums.
  // rd(c) reads u at node c - 1 along the line (c = 1..N); ulo/uhi give the
  // ghost u beyond the block's faces, gface_lo/hi the ghost gradient there from
  // the own face value and gradient.
  auto line_laplacian = [&](int el, auto rd, auto ulo, auto uhi, auto gface_lo, auto gface_hi,
                            T* lap) {
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
      g[0] = gface_lo(uw[1], g[1]);
    }
    if (el < E - 1) {  // row 0 of the element above: our last node, then its nodes
      T s = tb.W[0][0] * uw[P1];
#pragma unroll
      for (int k = 1; k < PW - 1; ++k) s += tb.W[0][k] * rd((el + 1) * P1 + k);
      g[PW - 1] = s * rh;
    } else {
      g[PW - 1] = gface_hi(uw[P1], g[P1]);
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
  };
  // one element line's truncation (evolution.jl:238-264 along one dimension)
  auto truncate_line = [&](T* v) {
    T c = 0;
#pragma unroll
    for (int j = 0; j < P1; ++j) c += tb.dual[j] * v[j];
#pragma unroll
    for (int j = 0; j < P1; ++j) v[j] -= tb.mode[j] * c;
  };

  // a layer's x/y ghosts: the neighbours' face values (two per thread)
  auto load_ghosts = [&](int e, T* gv) {
#pragma unroll
    for (int k = 0; k < GPT; ++k) {
      const int idx = tid + k * NTH;
      gv[k] = T(0);
      if (idx < NG) {
        const int q = idx / (4 * P1 * N), r = idx % (4 * P1 * N), f = r / (P1 * N),
                  t = r % (P1 * N);  // t = (plane in layer) * N + tangential index
        const int nbq = f == 0 ? nxl : (f == 1 ? nxh : (f == 2 ? nyl : nyh));
        if (nbq >= 0)
          gv[k] = A.uf_in[q * A.nfb + ((int64_t)nbq * 6 + (f ^ 1)) * N2 + e * P1 * N + t];
        else if (GHOST && nbq == CF)
          gv[k] = A.gu[q * A.nfb + ((int64_t)b * 6 + f) * N2 + e * P1 * N + t];
      }
    }
  };
  constexpr bool carry = sizeof(T) == 8 ? TWOLDG_FUSED_CARRY_F64 : TWOLDG_FUSED_CARRY_F32;
  constexpr bool defer = TWOLDG_FUSED_DEFER;

  T uc[P1];   // the layer's column
  T un2[P1];  // the next layer's column
  T gv[GPT];  // the layer's x/y ghosts
  T prev_last = 0, prev_part = 0;  // previous layer: last u, row-P partial sum
  load_layer(0, uc);
  if (tid < 6) {
    const int nb = ld_table(nb6 + tid);
    RSP[tid] = nb >= 0 ? T(1) / (T)ld_table(A.spacing + nb) : T(0);
  }
  lds_barrier();

#pragma unroll 1
  for (int e = 0; e < E; ++e) {
    // ---- the layer's x/y ghosts, issued first ----
    load_ghosts(e, gv);

    // ---- (1) columns: z-gradients (one node either side) and z-Laplacian ----
    {
      // the z ghosts: the next layer's first u and its row-0 partial sum, or
      // on the block's z faces the neighbour's face u and face gradient (a
      // mirror's ghost u is the own face u, its gradient minus the own face
      // gradient)
      T before = prev_last, after, gbefore = 0, gafter = 0, next_part = 0;
      if (GHOST && e == 0 && nzl == CF) {
        before = act ? A.gu[((int64_t)b * 6 + 4) * N2 + opaque(col)] : T(0);
        gbefore = act ? A.gg[((int64_t)b * 6 + 4) * N2 + opaque(col)] : T(0);
      } else if (e == 0) {
        const int nlo = nzl;
        before = (act && nlo >= 0) ? A.uf_in[((int64_t)nlo * 6 + 5) * N2 + opaque(col)] : uc[0];
        if (act && nlo >= 0) {
          gbefore = (A.pf_in[((int64_t)nlo * 6 + 5) * N2 + opaque(col)] + Whi * uc[0]) * rsp(4);
        } else {
          T g0 = Wlo * before;
#pragma unroll
          for (int j = 0; j < P1; ++j) g0 += tb.W[0][j + 1] * uc[j];
          gbefore = -(g0 * rh);
        }
      }
      if (e < E - 1) {
        load_layer(e + 1, un2);
        after = un2[0];
#pragma unroll
        for (int j = 0; j < P1; ++j) next_part += tb.W[0][j + 1] * un2[j];
      } else if (GHOST && nzh == CF) {
        after = act ? A.gu[((int64_t)b * 6 + 5) * N2 + opaque(col)] : T(0);
        gafter = act ? A.gg[((int64_t)b * 6 + 5) * N2 + opaque(col)] : T(0);
      } else {
        const int nhi = nzh;
        after = (act && nhi >= 0) ? A.uf_in[((int64_t)nhi * 6 + 4) * N2 + opaque(col)] : uc[P];
        if (act && nhi >= 0) {
          gafter = (A.pf_in[((int64_t)nhi * 6 + 4) * N2 + opaque(col)] + Wlo * uc[P]) * rsp(5);
        } else {
          T gN = Whi * after;
#pragma unroll
          for (int j = 0; j < P1; ++j) gN += tb.W[P][j + 1] * uc[j];
          gafter = -(gN * rh);
        }
      }
      T gl[PW];
      gl[0] = e == 0 ? gbefore : (prev_part + Whi * uc[0]) * rh;
#pragma unroll
      for (int i = 0; i < P1; ++i) {
        T s = 0;
#pragma unroll
        for (int k = 0; k < PW; ++k) {
          if ((k == 0 && i != 0) || (k == PW - 1 && i != P)) continue;
          s += tb.W[i][k] * (k == 0 ? before : (k == PW - 1 ? after : uc[k - 1]));
        }
        gl[i + 1] = s * rh;
      }
      gl[PW - 1] = e == E - 1 ? gafter : (Wlo * uc[P] + next_part) * rh;
      if (act) {
#pragma unroll
        for (int i = 0; i < P1; ++i) {
          T s = 0;
#pragma unroll
          for (int k = 0; k < PW; ++k) {
            if ((k == 0 && i != 0) || (k == PW - 1 && i != P)) continue;
            s += tb.W[i][k] * gl[k];
          }
          LZ[i * NT + col] = s;
        }
      }
    }
    prev_last = uc[P];
    prev_part = 0;
#pragma unroll
    for (int j = 0; j < P1; ++j) prev_part += tb.W[P][j + 1] * uc[j];

    // the layer's u planes, and the x/y ghosts: the neighbours' face values
    // into GH; a mirror's ghost u is the own face value
#pragma unroll
    for (int k = 0; k < GPT; ++k) {
      const int idx = tid + k * NTH;
      if (idx < NG) {
        const int q = idx / (4 * P1 * N), r = idx % (4 * P1 * N), f = r / (P1 * N),
                  t = r % (P1 * N);
        const int nbq = f == 0 ? nxl : (f == 1 ? nxh : (f == 2 ? nyl : nyh));
        if (nbq >= 0 || (GHOST && nbq == CF)) GH[4 * q + f][t] = gv[k];
      }
    }
    if (act) {
#pragma unroll
      for (int i = 0; i < P1; ++i) {
        const T uz = uc[i];
        U[i * NT + col] = uz;
        if (x == 0 && nxl == -1) GH[0][i * N + y] = uz;
        if (x == N - 1 && nxh == -1) GH[1][i * N + y] = uz;
        if (y == 0 && nyl == -1) GH[2][i * N + x] = uz;
        if (y == N - 1 && nyh == -1) GH[3][i * N + x] = uz;
      }
    }
    lds_barrier();
    // ---- (2) element lines: x-Laplacian sums to LX; y-Laplacian sums kept ----
    T ly[P1];
    if (act) {
      const int t2 = opaque(tid);
      const int ex = t2 % E, yx = (t2 / E) % N, px = t2 / (E * N);
      const int xy = t2 % N, ey = (t2 / N) % E, py = t2 / (N * E);
      {
        const T* Ur = U + px * NT + yx * N - 1;
        const int g = px * N + yx;
        T lx[P1];
        line_laplacian(
            ex, [&](int c) { return Ur[c]; }, [&]() { return GH[0][g]; }, [&]() { return GH[1][g]; },
            [&](T uf, T gf) {
              if (GHOST && nxl == CF) return GH[4][g];
              return nxl >= 0 ? (GH[4][g] + Whi * uf) * rsp(0) : -gf;
            },
            [&](T uf, T gf) {
              if (GHOST && nxh == CF) return GH[5][g];
              return nxh >= 0 ? (GH[5][g] + Wlo * uf) * rsp(1) : -gf;
            },
            lx);
#pragma unroll
        for (int j = 0; j < P1; ++j) LX[px * NT + yx * N + ex * P1 + j] = lx[j];
      }
      const T* Uc = U + py * NT - N + xy;
      const int g = py * N + xy;
      line_laplacian(
          ey, [&](int c) { return Uc[c * N]; }, [&]() { return GH[2][g]; }, [&]() { return GH[3][g]; },
          [&](T uf, T gf) {
            if (GHOST && nyl == CF) return GH[6][g];
            return nyl >= 0 ? (GH[6][g] + Whi * uf) * rsp(2) : -gf;
          },
          [&](T uf, T gf) {
            if (GHOST && nyh == CF) return GH[7][g];
            return nyh >= 0 ? (GH[7][g] + Wlo * uf) * rsp(3) : -gf;
          },
          ly);
    }
    lds_barrier();
    // ---- (3) y lines: the Laplacian, the stage combination, the next u, and
    // v's truncation along y; U is dead, UN takes its place ----
    if (act) {
      const int t3 = opaque(tid);
      const int xy = t3 % N, ey = (t3 / N) % E, py = t3 / (N * E);
      const int z = e * P1 + py;
      const int base = xy + ey * P1 * N + z * N2;  // node j of the line at base + j N
      T w[P1];
      T plo = 0, phi = 0;  // partial y-face gradients of the next u
      // modes whose accumulator operand is read (modes 8, 10: z; 13: y_2)
      constexpr bool acc_in = MODE == 2 || MODE == 4 || MODE == 8 || MODE == 10 || MODE == 13 ||
                              MODE == 15;
      constexpr T third = T(1) / T(3);
      auto ops = [&](int j, T& ku, T& yu, T& yv, T& au, T& av) {
        const int iu = base + j * N;
        ku = (T)ld_once(vs + iu);
        if (MODE == 1 || MODE == 11) {
          // stage 1: the stage input is the step start (ys = yn), whose u is
          // the layer's plane in LDS (this thread's node, not yet overwritten)
          yu = U[py * NT + (ey * P1 + j) * N + xy];
          yv = ku;
        } else if (MODE == 8 || MODE == 10) {  // the final stage from z: no y_n (u)
          yu = yv = T(0);
        } else {
          yu = ld_once(un_ + iu);  // compensated form: the high word
          yv = MODE == 4 ? ld_once(lo + iu) : (T)ld_once(vn + iu);  // compensated: u's low word
        }
        au = MODE == 6 ? yv : (acc_in ? ld_once(acc + iu) : T(0));
        av = (MODE == 6 || MODE == 14 || acc_in) ? (T)ld_once(vacc + iu) : T(0);
      };
      // the stage operands one node ahead where registers allow
      constexpr bool ahead = sizeof(T) == 4 || TWOLDG_FUSED_NB_F64 > 0;
      T ku, yu, yv, au, av;
      if (ahead) ops(0, ku, yu, yv, au, av);
#pragma unroll
      for (int j = 0; j < P1; ++j) {
        T ku1 = 0, yu1 = 0, yv1 = 0, au1 = 0, av1 = 0;
        if (!ahead) ops(j, ku, yu, yv, au, av);
        if (ahead && j + 1 < P1) ops(j + 1, ku1, yu1, yv1, au1, av1);
        const int iu = base + j * N;
        const int l = py * NT + (ey * P1 + j) * N + xy;
        const T lap = ((T(0) + LX[l]) + ly[j]) + LZ[l];
        const T kv = lap * rh;
        T unext;
        // the accumulator (modes 1, 2) or the compensation of u (mode 4): stored
        // now, or left in LX and LZ at l (read above by this thread only)
        auto out = [&](T* pu, bool with_v, T ou, T ov) {
          if (defer) {
            LX[l] = ou;
            if (with_v) LZ[l] = ov;
          } else {
            if (MODE != 1 && MODE != 14) st_once(pu + iu, ou);
            if (with_v) st_once(vacc + iu, (TV)ov);
          }
        };
        if (MODE == 1) {
          out(acc, true, ku, kv);
          unext = yu + A.a * ku;
          w[j] = yv + A.a * kv;
        } else if (MODE == 2 || MODE == 6) {
          out(acc, true, au + T(2) * ku, av + T(2) * kv);
          unext = yu + A.a * ku;
          w[j] = yv + A.a * kv;
        } else if (MODE == 11 || MODE == 12) {  // FP64 stages 1 and 2: no accumulator
          unext = yu + A.a * ku;
          w[j] = yv + A.a * kv;
        } else if (MODE == 13) {
          // FP64 stage 3: y_4, and z = y_n + dt/6 (k1 + 2 k2 + 2 k3) formed as
          // (y_2 + 2 y_3)/3 + dt/3 k3 from y_2 (in au, av) and the stage input
          // y_3 (u: this thread's node of the LDS plane, v: ku)
          const T su = U[l];
          out(acc, true, (au + T(2) * su) * third + A.a2 * ku,
              (av + T(2) * ku) * third + A.a2 * kv);
          unext = yu + A.a * ku;
          w[j] = yv + A.a * kv;
        } else if (MODE == 14) {  // split, stage 2: y_3, the accumulator for v only
          out(acc, true, T(0), av + T(2) * kv);
          unext = yu + A.a * ku;
          w[j] = yv + A.a * kv;
        } else if (MODE == 15) {
          // split, stage 3: u as in mode 13 (z over y_2), the accumulator for v
          const T su = U[l];
          out(acc, true, (au + T(2) * su) * third + A.a2 * ku, av + T(2) * kv);
          unext = yu + A.a * ku;
          w[j] = yv + A.a * kv;
        } else if (MODE == 8) {  // the final stage from z (in au, av)
          unext = au + A.a * ku;
          w[j] = av + A.a * kv;
        } else if (MODE == 10) {  // split, the final stage: u from z, v's increment
          unext = au + A.a * ku;
          w[j] = A.a * (av + kv);
        } else if (MODE == 4) {  // compensated: u by Kahan now, v's increment truncated first
          const T du = A.a * (au + ku);
          const T yk = du - yv, tu = yu + yk;
          out(lo, false, (tu - yu) - yk, T(0));
          unext = tu;
          w[j] = A.a * (av + kv);
        } else {  // split: u plain, v's increment truncated first
          unext = yu + A.a * (au + ku);
          w[j] = A.a * (av + kv);
        }
        if (!defer) st_once(yo + iu, unext);
        UN[l] = unext;
        // faces of the next u along y
        if (ey == 0) {
          plo += tb.W[0][j + 1] * unext;
          if (!defer && j == 0) ufo[2 * N2 + xy + N * z] = unext;
        }
        if (ey == E - 1) {
          phi += tb.W[P][j + 1] * unext;
          if (!defer && j == P) ufo[3 * N2 + xy + N * z] = unext;
        }
        if (ahead) {
          ku = ku1; yu = yu1; yv = yv1; au = au1; av = av1;
        }
      }
      if (defer && ey == 0) ufo[2 * N2 + xy + N * z] = UN[py * NT + xy];
      if (defer && ey == E - 1) ufo[3 * N2 + xy + N * z] = UN[py * NT + (N - 1) * N + xy];
      if (ey == 0) pfo[2 * N2 + xy + N * z] = plo;
      if (ey == E - 1) pfo[3 * N2 + xy + N * z] = phi;
      if (A.truncate) truncate_line(w);
#pragma unroll
      for (int j = 0; j < P1; ++j) VB[py * NT + (ey * P1 + j) * N + xy] = w[j];
    }
    lds_barrier();
    // ---- (4) x lines: v's truncation along x, faces of the next u along x;
    // columns: faces of the next u along z ----
    TV lov[P1], hv[P1];  // the final stage's v pair, ahead of this phase's stores
    if (act && (MODE == 4 || MODE == 10)) {
#pragma unroll
      for (int j = 0; j < P1; ++j) {
        const int iv = col + (e * P1 + j) * N2;
        lov[j] = ld_once(vlo + iv);
        hv[j] = ld_once(vn + iv);
      }
    }
    if (act) {
      const int t4 = opaque(tid);
      const int ex = t4 % E, yx = (t4 / E) % N, px = t4 / (E * N);
      const int z = e * P1 + px;
      const int l = px * NT + yx * N + ex * P1;
      if (A.truncate) {
        T w[P1];
#pragma unroll
        for (int j = 0; j < P1; ++j) w[j] = VB[l + j];
        truncate_line(w);
#pragma unroll
        for (int j = 0; j < P1; ++j) VB[l + j] = w[j];
      }
      if (ex == 0) {
        T s = 0;
#pragma unroll
        for (int j = 0; j < P1; ++j) s += tb.W[0][j + 1] * UN[l + j];
        ufo[0 * N2 + yx + N * z] = UN[l];
        pfo[0 * N2 + yx + N * z] = s;
      }
      if (ex == E - 1) {
        T s = 0;
#pragma unroll
        for (int j = 0; j < P1; ++j) s += tb.W[P][j + 1] * UN[l + j];
        ufo[1 * N2 + yx + N * z] = UN[l + P];
        pfo[1 * N2 + yx + N * z] = s;
      }
      if (e == 0) {
        T s = 0;
#pragma unroll
        for (int i = 0; i < P1; ++i) s += tb.W[0][i + 1] * UN[i * NT + col];
        ufo[4 * N2 + opaque(col)] = UN[col];
        pfo[4 * N2 + opaque(col)] = s;
      }
      if (e == E - 1) {
        T s = 0;
#pragma unroll
        for (int i = 0; i < P1; ++i) s += tb.W[P][i + 1] * UN[i * NT + col];
        ufo[5 * N2 + opaque(col)] = UN[P * NT + col];
        pfo[5 * N2 + opaque(col)] = s;
      }
    }
    lds_barrier();
    // ---- (5) columns: truncation along z in registers; v out ----
    // The next layer's column is read again (registers allowing no carry) ahead
    // of this phase's stores, so that waiting for it does not wait for them.
    if (!carry && e + 1 < E) load_layer(e + 1, uc);
    if (act) {
      T wl[P1];
#pragma unroll
      for (int i = 0; i < P1; ++i) wl[i] = VB[i * NT + col];
      if (A.truncate) truncate_line(wl);
#pragma unroll
      for (int j = 0; j < P1; ++j) {
        const int iu = col + (e * P1 + j) * N2;
        if (defer) {  // the stage update's outputs of the column
          const int li = j * NT + col;
          st_once(yo + iu, UN[li]);
          if (MODE == 2 || MODE == 6 || MODE == 13 || MODE == 15) st_once(acc + iu, LX[li]);
          if (MODE == 1 || MODE == 2 || MODE == 6 || MODE == 13 || MODE == 14 || MODE == 15)
            st_once(vacc + iu, (TV)LZ[li]);
          if (MODE == 4) st_once(lo + iu, LX[li]);
        }
        if (MODE == 4) {
          const T yk = wl[j] - (T)lov[j], tv = (T)hv[j] + yk;
          st_once(vlo + iu, (TV)((tv - (T)hv[j]) - yk));
          st_once(vo + iu, (TV)tv);
        } else if (MODE == 10) {
          // v_n = hi - lo exactly in T; the new pair rounds v_n + increment
          const T vt = ((T)hv[j] - (T)lov[j]) + wl[j];
          const TV hi = (TV)vt;
          st_once(vlo + iu, (TV)((T)hi - vt));
          st_once(vo + iu, hi);
        } else {
          st_once(vo + iu, (TV)wl[j]);
        }
      }
    }
    // slide the column window
    if (carry && e + 1 < E) {
#pragma unroll
      for (int j = 0; j < P1; ++j) uc[j] = un2[j];
    }
  }
}

// The coarse-fine faces, one pair per job: a fine block's face F and the
// quarter of its coarser neighbour's face C that F covers (the quarter's coarse
// elements are the parents of F's elements: E is even).  Face planes share the
// tangential axes t0 < t1, node t = i0 + N i1.
struct CFJob {
  int32_t bf, ff;  // fine block, its face 2d + s
  int32_t bc, fc;  // coarse block, its face 2d + 1 - s
  int32_t q;       // C's first quarter node: o0 N/2 + N o1 N/2, (o0, o1) F's half along t0, t1
  int32_t pad;
};

// The complete ghosts of both sides of a coarse-fine face pair, from the face
// buffers the previous stage wrote (u on the face plane, pf the block-local part
// of the face gradient); injection along the normal, tensor products of the
// element's 1D operators along the two tangential axes:
//   F's ghost u = I (C's u on the quarter), C's ghost u = R (F's u), with I the
//   coarse element's polynomial at its children's nodes and R = M_c^-1 I^T M_f;
//   a face's complete gradient = (pf + W_ghost ghost u) / h (evolution.jl:72-110
//   at a face node), and the ghost gradients the same I and R of the other
//   side's complete gradients.
// One workgroup per pair; every value stays in LDS between the passes.
template <class T, int P1, int E>
__global__ void __launch_bounds__(256) k_cf_ghosts(const CFJob* __restrict__ jobs,
                                                   const T* __restrict__ uf,
                                                   const T* __restrict__ pf,
                                                   const double* __restrict__ spacing,
                                                   T* __restrict__ gu, T* __restrict__ gg) {
  constexpr int N = P1 * E, N2 = N * N, H = N / 2, H2 = H * H, P = P1 - 1;
  constexpr FusedTab<T, P1> tb = fused_tab<T, P1>();
  static_assert(E % 2 == 0, "a coarse element's children pair up inside a block");
  __shared__ T sI[2 * P1 * P1];  // I[h P1 + a][c]: child h's node a from coarse node c
  __shared__ T sR[P1 * 2 * P1];  // R[c][h P1 + a]
  __shared__ T sUF[N2], sPF[N2], sUC[H2], sPC[H2];
  __shared__ T sA[N * H], sB[H * N];  // first-axis passes: I along t0 (N x H), R along t0 (H x N)
  __shared__ T sGF[N2], sGC[H2];      // complete face gradients
  const int tid = threadIdx.x;
  // the job's fields and the spacings with the device-level loads of the other
  // per-block tables (ld_table): written at creation into recycled memory
  const int32_t* jp = (const int32_t*)(jobs + blockIdx.x);
  CFJob J;
  J.bf = ld_table(jp + 0);
  J.ff = ld_table(jp + 1);
  J.bc = ld_table(jp + 2);
  J.fc = ld_table(jp + 3);
  J.q = ld_table(jp + 4);
  const twoldg::ElementTables& et = twoldg::kElementTables[P1 - 2];
  for (int k = tid; k < 2 * P1 * P1; k += 256) {
    sI[k] = (T)et.interp[k / P1][k % P1];
    sR[k] = (T)et.restrict_[k / (2 * P1)][k % (2 * P1)];
  }
  const int64_t of = ((int64_t)J.bf * 6 + J.ff) * N2;
  const int64_t oc = ((int64_t)J.bc * 6 + J.fc) * N2 + J.q;
  for (int k = tid; k < N2; k += 256) {
    sUF[k] = uf[of + k];
    sPF[k] = pf[of + k];
  }
  for (int k = tid; k < H2; k += 256) {
    const int64_t c = oc + (k % H) + (int64_t)N * (k / H);
    sUC[k] = uf[c];
    sPC[k] = pf[c];
  }
  const T rhf = T(1) / (T)ld_table(spacing + J.bf), rhc = T(1) / (T)ld_table(spacing + J.bc);
  // the ghost term of each side's own face row (row 0 on a low face, P on a high one)
  const T wf = (J.ff % 2) ? tb.W[P][P1 + 1] : tb.W[0][0];
  const T wc = (J.fc % 2) ? tb.W[P][P1 + 1] : tb.W[0][0];
  __syncthreads();
  // I along one axis: fine node i of element fe = i / P1 from coarse element fe / 2
  auto interp = [&](const T* src, int i, int stride) {
    const int fe = i / P1, row = (fe % 2) * P1 + i % P1;
    const T* s = src + (fe / 2) * P1 * stride;
    T acc = 0;
#pragma unroll
    for (int c = 0; c < P1; ++c) acc += sI[row * P1 + c] * s[c * stride];
    return acc;
  };
  // R along one axis: coarse node c of element ce = c / P1 from fine elements 2 ce, 2 ce + 1
  auto restrict1 = [&](const T* src, int c, int stride) {
    const int ce = c / P1, row = c % P1;
    const T* s = src + 2 * ce * P1 * stride;
    T acc = 0;
#pragma unroll
    for (int f = 0; f < 2 * P1; ++f) acc += sR[row * 2 * P1 + f] * s[f * stride];
    return acc;
  };
  // two passes of the same operators: on u (then the complete gradients), on those
#pragma unroll 1
  for (int pass = 0; pass < 2; ++pass) {
    const T* uc = pass ? sGC : sUC;  // C's quarter, H x H
    const T* ufn = pass ? sGF : sUF; // F's face, N x N
    for (int k = tid; k < N * H; k += 256) {
      const int i = k % N, cj = k / N;  // sA[cj][i]: I along t0 of C's row cj
      sA[k] = interp(uc + cj * H, i, 1);
      const int ci = k % H, j = k / H;  // sB[j][ci]: R along t0 of F's row j
      sB[k] = restrict1(ufn + j * N, ci, 1);
    }
    __syncthreads();
    T* out = pass ? gg : gu;
    for (int k = tid; k < N2; k += 256) {
      const int i = k % N, j = k / N;
      const T g = interp(sA + i, j, N);  // along t1: column i of sA
      out[of + k] = g;
      if (pass == 0) sGF[k] = (sPF[k] + wf * g) * rhf;
    }
    for (int k = tid; k < H2; k += 256) {
      const int ci = k % H, cj = k / H;
      const T g = restrict1(sB + ci, cj, H);  // along t1: column ci of sB
      out[oc + ci + (int64_t)N * cj] = g;
      if (pass == 0) sGC[k] = (sPC[k] + wc * g) * rhc;
    }
    __syncthreads();
  }
}

// The face buffers of a state (before the first fused stage): per block face,
// the face plane of u and the block-local part of the face gradient.
template <class T, int P1, int E>
__global__ void k_faces3(int64_t nb, const T* __restrict__ y, T* __restrict__ uf,
                         T* __restrict__ pf) {
  constexpr FusedTab<T, P1> tb = fused_tab<T, P1>();  // as the stage kernel's
  constexpr int N = P1 * E, P = P1 - 1;
  const int64_t N2 = (int64_t)N * N, nn = N2 * N;
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= nb * 6 * N2) return;
  const int64_t b = g / (6 * N2), f = (g / N2) % 6, t = g % N2;
  const int d = (int)(f / 2), s = (int)(f % 2);
  const int t0 = (int)(t % N), t1 = (int)(t / N);
  int I[3];
  I[d] = 0;
  if (d == 0) { I[1] = t0; I[2] = t1; }
  if (d == 1) { I[0] = t0; I[2] = t1; }
  if (d == 2) { I[0] = t0; I[1] = t1; }
  const int64_t stride = d == 0 ? 1 : (d == 1 ? N : N2);
  const T* u = y + b * 2 * nn;
  const int64_t base = I[0] + (int64_t)N * I[1] + N2 * I[2];  // with I[d] = 0
  const int first = s ? N - P1 : 0;
  const int row = s ? P : 0;
  T acc = 0;
#pragma unroll
  for (int j = 0; j < P1; ++j) acc += tb.W[row][j + 1] * u[base + (first + j) * stride];
  uf[g] = u[base + (s ? N - 1 : 0) * stride];
  pf[g] = acc;
}
