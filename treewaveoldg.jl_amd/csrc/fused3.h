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
// So one stage reads u, v of the stage state, y_n and the accumulator once and
// writes the accumulator and the next state once (34 values per node and step
// in FP64), plus faces, O(1/N) of that.

#pragma once

// Waves per SIMD the fused stage is compiled for.  FP64 needs about 250 VGPRs
// without spilling (two waves per SIMD); FP32 fits three.
template <class T>
constexpr int kFusedWavesPerEU = sizeof(T) == 8 ? 2 : 3;

// A workgroup barrier that orders LDS only.  __syncthreads() also fences global
// memory, which makes every wave wait for its outstanding global loads at each
// barrier and so serializes the plane's loads with the LDS phases; the fused
// stage shares nothing through global memory inside a launch.
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

template <class T>
struct FusedArgs {
  const T* ys;       // stage input (u, v)
  const T* yn;       // step start
  T* acc;            // RK accumulator
  T* yo;             // output state
  T* lo;             // compensation of yn (final stage, compensated form)
  const T* uf_in;    // faces of ys: (nb, 6, N^2) u on the face plane
  const T* pf_in;    //              block-local part of the face gradient
  T* uf_out;         // faces of yo
  T* pf_out;
  const int32_t* nbr;       // (nb, 6) same-level neighbour, or -1 for a mirror
  const double* spacing;    // (nb)
  const Tables<T>* tables;
  T a;               // stage coefficient (dt/2, dt or dt/6)
  int truncate;
};

// Face index f = 2d + s; tangential index t of node (x, y, z):
// d = 0: y + N z; d = 1: x + N z; d = 2: x + N y.

template <class T, int P1, int E, int MODE>
__global__ void __launch_bounds__(((P1 * E) * (P1 * E) + 63) / 64 * 64, kFusedWavesPerEU<T>)
k_fused3(FusedArgs<T> A) {
  constexpr int N = P1 * E;
  constexpr int NT = N * N;
  constexpr int PW = P1 + 2;
  constexpr int P = P1 - 1;
  constexpr int NP = N + 2;
  __shared__ T sW[P1][PW];
  __shared__ T sMode[P1], sDual[P1];
  __shared__ T U[NP * NP];         // u plane with its x and y ghosts
  __shared__ T GX[N * NP];         // g_x plane, x ghosts
  __shared__ T GY[NP * N];         // g_y plane, y ghosts
  __shared__ T VA[NT], VB[NT];     // truncation passes
  __shared__ T UN[NT];             // next u plane (face parts)
  __shared__ T EG[8][NT];          // x/y neighbours' face values: u (xl xh yl yh), then p

  const int64_t b = blockIdx.x;
  const int tid = threadIdx.x;
  const bool act = tid < NT;
  const int x = act ? tid % N : 0, y = act ? tid / N : 0;
  const int64_t nn = (int64_t)N * N * N;
  const int64_t N2 = (int64_t)N * N;

  if (tid < P1 * PW) sW[tid / PW][tid % PW] = A.tables->W[tid / PW][tid % PW];
  if (tid < P1) {
    sMode[tid] = A.tables->mode[tid];
    sDual[tid] = A.tables->dual[tid];
  }

  const T h = (T)A.spacing[b];
  const T rh = T(1) / h;
  const int32_t* nb6 = A.nbr + b * 6;
  const int64_t col = (int64_t)y * N + x;
  const T* us = A.ys + b * 2 * nn + col;          // this column's u (stride N2)
  const T* vs = us + nn;
  const T* un_ = A.yn + b * 2 * nn + col;
  T* acc = A.acc + b * 2 * nn + col;
  T* yo = A.yo + b * 2 * nn + col;
  T* lo = A.lo ? A.lo + b * 2 * nn + col : nullptr;
  lds_barrier();

  auto load_layer = [&](int e, T* dst) {
#pragma unroll
    for (int j = 0; j < P1; ++j) dst[j] = act ? us[(int64_t)(e * P1 + j) * N2] : T(0);
  };
  // the ghost terms of rows 0 and P
  const T Wlo = sW[0][0], Whi = sW[P][P1 + 1];

  // ---- z ghosts: u and the gradient on both z faces -------------------------
  T uc[P1], up[P1];
  T prev_last = 0, prev_part = 0;  // previous layer: last u, row-P partial sum
  load_layer(0, uc);
  T ug_lo, ug_hi, gg_lo, gg_hi;
  {
    const int nlo = nb6[4], nhi = nb6[5];
    T ul[P1];
    load_layer(E - 1, ul);
    ug_lo = (act && nlo >= 0) ? A.uf_in[((int64_t)nlo * 6 + 5) * N2 + col] : uc[0];
    ug_hi = (act && nhi >= 0) ? A.uf_in[((int64_t)nhi * 6 + 4) * N2 + col] : ul[P];
    if (act && nlo >= 0) {
      gg_lo = (A.pf_in[((int64_t)nlo * 6 + 5) * N2 + col] + Whi * uc[0]) * (T(1) / (T)A.spacing[nlo]);
    } else {  // mirror: minus the own face gradient (row 0 with ghost = own face)
      T g0 = Wlo * ug_lo;
#pragma unroll
      for (int j = 0; j < P1; ++j) g0 += sW[0][j + 1] * uc[j];
      gg_lo = -(g0 * rh);
    }
    if (act && nhi >= 0) {
      gg_hi = (A.pf_in[((int64_t)nhi * 6 + 4) * N2 + col] + Wlo * ul[P]) * (T(1) / (T)A.spacing[nhi]);
    } else {
      T gN = 0;
#pragma unroll
      for (int j = 0; j < P1; ++j) gN += sW[P][j + 1] * ul[j];
      gN += Whi * ug_hi;
      gg_hi = -(gN * rh);
    }
  }
  if (E > 1) load_layer(1, up);

  // x and y neighbours
  const int nxl = nb6[0], nxh = nb6[1], nyl = nb6[2], nyh = nb6[3];
  const T rxl = nxl >= 0 ? T(1) / (T)A.spacing[nxl] : rh, rxh = nxh >= 0 ? T(1) / (T)A.spacing[nxh] : rh;
  const T ryl = nyl >= 0 ? T(1) / (T)A.spacing[nyl] : rh, ryh = nyh >= 0 ? T(1) / (T)A.spacing[nyh] : rh;
  const int ix = x % P1, iy = y % P1;
  const int sx = x - ix - 1, sy = y - iy - 1;  // window starts (owned index)

  T pz_lo = 0, pz_hi = 0;  // partial z-face gradients of the next u
  T uz0 = 0, uzN = 0;      // the next u on the z faces

  // Every global operand of plane z: the stage arithmetic's, and on the block's
  // x and y faces the neighbours' face values.  They are loaded one plane ahead
  // so that a plane's loads are in flight while the previous plane's LDS phases
  // run.
  struct Ops {
    T ku, yu, yv, au, av;
  };
  auto load_ops = [&](int z) {
    Ops o{};
    if (!act) return o;
    const int64_t iu = (int64_t)z * N2, iv = iu + nn;
    o.ku = vs[iu];
    if (MODE != 4) {
      o.yu = un_[iu];
      o.yv = un_[iv];
    }
    if (MODE >= 2) {
      o.au = acc[iu];
      o.av = acc[iv];
    }
    return o;
  };
  // the x and y neighbours' face planes, once per block
  {
    const int nbs[4] = {nxl, nxh, nyl, nyh};
    const int opp[4] = {1, 0, 3, 2};  // the neighbour's face that touches ours
    for (int idx = tid; idx < 8 * NT; idx += blockDim.x) {
      const int q = idx / NT, t = idx % NT, f = q % 4;
      const int nbq = nbs[f];
      if (nbq >= 0) {
        const T* src = q < 4 ? A.uf_in : A.pf_in;
        EG[q][t] = src[((int64_t)nbq * 6 + opp[f]) * N2 + t];
      }
    }
  }
  Ops cur = load_ops(0);
  lds_barrier();

#pragma unroll 1
  for (int e = 0; e < E; ++e) {
    // ---- the layer's z-gradients (one node either side) and z-Laplacian ----
    T gl[PW];
    {
      // node e*P1 - 1: row P of the previous layer, its ghost after = uc[0]
      if (e == 0) {
        gl[0] = gg_lo;
      } else {
        gl[0] = (prev_part + sW[P][P1 + 1] * uc[0]) * rh;
      }
      const T before = e == 0 ? ug_lo : prev_last;
      const T after = e == E - 1 ? ug_hi : up[0];
#pragma unroll
      for (int i = 0; i < P1; ++i) {
        T s = 0;
#pragma unroll
        for (int k = 0; k < PW; ++k) {
          if ((k == 0 && i != 0) || (k == PW - 1 && i != P)) continue;
          s += sW[i][k] * (k == 0 ? before : (k == PW - 1 ? after : uc[k - 1]));
        }
        gl[i + 1] = s * rh;
      }
      // node e*P1 + P1: row 0 of the next layer, its ghost before = uc[P]
      if (e == E - 1) {
        gl[PW - 1] = gg_hi;
      } else {
        T s = sW[0][0] * uc[P];
#pragma unroll
        for (int j = 0; j < P1; ++j) s += sW[0][j + 1] * up[j];
        gl[PW - 1] = s * rh;
      }
    }
    T lapl[P1];
#pragma unroll
    for (int i = 0; i < P1; ++i) {
      T s = 0;
#pragma unroll
      for (int k = 0; k < PW; ++k) {
        if ((k == 0 && i != 0) || (k == PW - 1 && i != P)) continue;
        s += sW[i][k] * gl[k];
      }
      lapl[i] = s;
    }

    T wl[P1];  // the layer's values to truncate (v_next, or its increment)
#pragma unroll
    for (int i = 0; i < P1; ++i) {
      const int z = e * P1 + i;
      const T uz = uc[i];
      const int64_t iu = (int64_t)z * N2, iv = iu + nn;
      // the next plane's operands go out first
      Ops nxt{};
      if (z + 1 < N) nxt = load_ops(z + 1);
      const T ku = cur.ku, yu = cur.yu, yv = cur.yv, au = cur.au, av = cur.av;
      // (1) the u plane and its x/y ghosts
      if (act) {
        U[(y + 1) * NP + x + 1] = uz;
        if (x == 0) U[(y + 1) * NP] = nxl >= 0 ? EG[0][y + N * z] : uz;
        if (x == N - 1) U[(y + 1) * NP + N + 1] = nxh >= 0 ? EG[1][y + N * z] : uz;
        if (y == 0) U[x + 1] = nyl >= 0 ? EG[2][x + N * z] : uz;
        if (y == N - 1) U[(N + 1) * NP + x + 1] = nyh >= 0 ? EG[3][x + N * z] : uz;
      }
      lds_barrier();
      // (2) g_x and g_y at the node, and the ghosts of the gradients
      if (act) {
        T gx = 0, gy = 0;
#pragma unroll
        for (int k = 0; k < PW; ++k) {
          gx += sW[ix][k] * U[(y + 1) * NP + (sx + k + 1)];
          gy += sW[iy][k] * U[(sy + k + 1) * NP + x + 1];
        }
        gx *= rh;
        gy *= rh;
        GX[y * NP + x + 1] = gx;
        GY[(y + 1) * N + x] = gy;
        if (x == 0)
          GX[y * NP] = nxl >= 0 ? (EG[4][y + N * z] + Whi * uz) * rxl : -gx;
        if (x == N - 1)
          GX[y * NP + N + 1] = nxh >= 0 ? (EG[5][y + N * z] + Wlo * uz) * rxh : -gx;
        if (y == 0)
          GY[x] = nyl >= 0 ? (EG[6][x + N * z] + Whi * uz) * ryl : -gy;
        if (y == N - 1)
          GY[(N + 1) * N + x] = nyh >= 0 ? (EG[7][x + N * z] + Wlo * uz) * ryh : -gy;
      }
      lds_barrier();
      // (3) the Laplacian and the stage combination
      T unext = 0, w = 0;
      if (act) {
        T lx = 0, ly = 0;
#pragma unroll
        for (int k = 0; k < PW; ++k) {
          lx += sW[ix][k] * GX[y * NP + (sx + k + 1)];
          ly += sW[iy][k] * GY[(sy + k + 1) * N + x];
        }
        const T lap = ((T(0) + lx) + ly) + lapl[i];
        const T kv = lap * rh;
        if (MODE == 1) {
          acc[iu] = ku;
          acc[iv] = kv;
          unext = yu + A.a * ku;
          w = yv + A.a * kv;
        } else if (MODE == 2) {
          acc[iu] = au + T(2) * ku;
          acc[iv] = av + T(2) * kv;
          unext = yu + A.a * ku;
          w = yv + A.a * kv;
        } else if (MODE == 3) {
          unext = yu + A.a * (au + ku);
          w = yv + A.a * (av + kv);
        } else {  // compensated: u by Kahan now, v's increment truncated first
          const T du = A.a * (au + ku);
          const T yk = du - lo[iu], hu = un_[iu], tu = hu + yk;
          lo[iu] = (tu - hu) - yk;
          unext = tu;
          w = A.a * (av + kv);
        }
        yo[iu] = unext;
        VA[tid] = w;
        UN[tid] = unext;
      }
      // faces of the next u along z
      if (e == 0) pz_lo += sW[0][i + 1] * unext;
      if (e == E - 1) pz_hi += sW[P][i + 1] * unext;
      if (z == 0) uz0 = unext;
      if (z == N - 1) uzN = unext;
      lds_barrier();
      // (4) faces of the next u along x and y; truncation along x
      if (act) {
        if (x == 0) {
          T s = 0;
#pragma unroll
          for (int j = 0; j < P1; ++j) s += sW[0][j + 1] * UN[y * N + j];
          A.uf_out[(b * 6 + 0) * N2 + y + N * z] = unext;
          A.pf_out[(b * 6 + 0) * N2 + y + N * z] = s;
        }
        if (x == N - 1) {
          T s = 0;
#pragma unroll
          for (int j = 0; j < P1; ++j) s += sW[P][j + 1] * UN[y * N + N - P1 + j];
          A.uf_out[(b * 6 + 1) * N2 + y + N * z] = unext;
          A.pf_out[(b * 6 + 1) * N2 + y + N * z] = s;
        }
        if (y == 0) {
          T s = 0;
#pragma unroll
          for (int j = 0; j < P1; ++j) s += sW[0][j + 1] * UN[j * N + x];
          A.uf_out[(b * 6 + 2) * N2 + x + N * z] = unext;
          A.pf_out[(b * 6 + 2) * N2 + x + N * z] = s;
        }
        if (y == N - 1) {
          T s = 0;
#pragma unroll
          for (int j = 0; j < P1; ++j) s += sW[P][j + 1] * UN[(N - P1 + j) * N + x];
          A.uf_out[(b * 6 + 3) * N2 + x + N * z] = unext;
          A.pf_out[(b * 6 + 3) * N2 + x + N * z] = s;
        }
        if (A.truncate) {
          T c = 0;
#pragma unroll
          for (int j = 0; j < P1; ++j) c += sDual[j] * VA[y * N + (x - ix) + j];
          VB[tid] = w - sMode[ix] * c;
        } else {
          VB[tid] = w;
        }
      }
      lds_barrier();
      // (5) truncation along y: the layer's value
      if (act) {
        if (A.truncate) {
          T c = 0;
#pragma unroll
          for (int j = 0; j < P1; ++j) c += sDual[j] * VB[((y - iy) + j) * N + x];
          wl[i] = VB[tid] - sMode[iy] * c;
        } else {
          wl[i] = VB[tid];
        }
      }
      cur = nxt;
    }
    // end of the layer: truncation along z, then v out
    if (act) {
      if (A.truncate) {
        T c = 0;
#pragma unroll
        for (int j = 0; j < P1; ++j) c += sDual[j] * wl[j];
#pragma unroll
        for (int j = 0; j < P1; ++j) wl[j] -= sMode[j] * c;
      }
#pragma unroll
      for (int j = 0; j < P1; ++j) {
        const int64_t iv = (int64_t)(e * P1 + j) * N2 + nn;
        if (MODE == 4) {
          const T yk = wl[j] - lo[iv], hv = un_[iv], tv = hv + yk;
          lo[iv] = (tv - hv) - yk;
          yo[iv] = tv;
        } else {
          yo[iv] = wl[j];
        }
      }
    }
    // slide the column window
    prev_last = uc[P];
    prev_part = 0;
#pragma unroll
    for (int j = 0; j < P1; ++j) prev_part += sW[P][j + 1] * uc[j];
#pragma unroll
    for (int j = 0; j < P1; ++j) uc[j] = up[j];
    if (e + 2 < E) load_layer(e + 2, up);
  }
  if (act) {
    A.uf_out[(b * 6 + 4) * N2 + col] = uz0;
    A.uf_out[(b * 6 + 5) * N2 + col] = uzN;
    A.pf_out[(b * 6 + 4) * N2 + col] = pz_lo;
    A.pf_out[(b * 6 + 5) * N2 + col] = pz_hi;
  }
}

// The face buffers of a state (before the first fused stage): per block face,
// the face plane of u and the block-local part of the face gradient.
template <class T, int P1, int E>
__global__ void k_faces3(int64_t nb, const T* __restrict__ y, const Tables<T>* __restrict__ tb,
                         T* __restrict__ uf, T* __restrict__ pf) {
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
  for (int j = 0; j < P1; ++j) acc += tb->W[row][j + 1] * u[base + (first + j) * stride];
  uf[g] = u[base + (s ? N - 1 : 0) * stride];
  pf[g] = acc;
}
