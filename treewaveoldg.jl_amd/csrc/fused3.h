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
#ifndef TWOLDG_FUSED_WAVES_F32
#define TWOLDG_FUSED_WAVES_F32 3
#endif
template <class T>
constexpr int kFusedWavesPerEU = sizeof(T) == 8 ? 2 : TWOLDG_FUSED_WAVES_F32;

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
  constexpr int SU = NP * NP, SG = N * NP;  // plane strides of U and of GX, GY
  // One element layer (P1 planes) per phase.  R1 holds the u planes, later the
  // values to truncate; R2 the x-gradients, later the x-truncated values.
  __shared__ T sW[P1][PW];
  __shared__ T sMode[P1], sDual[P1];
  __shared__ T R1[P1 * SU];
  __shared__ T R2[P1 * SG];
  __shared__ T GY[P1 * SG];
  __shared__ T UN[P1 * NT];
  __shared__ T EG[8][NT];          // x/y neighbours' face values: u (xl xh yl yh), then p
  T* U = R1;
  T* VA = R1;
  T* GX = R2;
  T* VB = R2;

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
  const int nxl = nb6[0], nxh = nb6[1], nyl = nb6[2], nyh = nb6[3];

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

  const T rxl = nxl >= 0 ? T(1) / (T)A.spacing[nxl] : rh, rxh = nxh >= 0 ? T(1) / (T)A.spacing[nxh] : rh;
  const T ryl = nyl >= 0 ? T(1) / (T)A.spacing[nyl] : rh, ryh = nyh >= 0 ? T(1) / (T)A.spacing[nyh] : rh;
  const int ix = x % P1, iy = y % P1;
  const int sx = x - ix - 1, sy = y - iy - 1;  // window starts (owned index)

  T pz_lo = 0, pz_hi = 0;  // partial z-face gradients of the next u
  T uz0 = 0, uzN = 0;      // the next u on the z faces

#pragma unroll 1
  for (int e = 0; e < E; ++e) {
    // ---- the layer's global operands: issued first, used in phase (3) ----
    T ku[P1], yu[P1], yv[P1], au[P1], av[P1];
#pragma unroll
    for (int i = 0; i < P1; ++i) {
      ku[i] = yu[i] = yv[i] = au[i] = av[i] = T(0);
      if (act) {
        const int64_t iu = (int64_t)(e * P1 + i) * N2, iv = iu + nn;
        ku[i] = vs[iu];
        yu[i] = un_[iu];
        yv[i] = un_[iv];
        if (MODE >= 2) {
          au[i] = acc[iu];
          av[i] = acc[iv];
        }
      }
    }
    // the column two layers on, needed when the window slides
    T un2[P1];
    if (e + 2 < E) load_layer(e + 2, un2);

    // ---- the layer's z-gradients (one node either side) and z-Laplacian ----
    T lapl[P1];
    {
      T gl[PW];
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
      if (e == E - 1) {
        gl[PW - 1] = gg_hi;
      } else {
        T s = sW[0][0] * uc[P];
#pragma unroll
        for (int j = 0; j < P1; ++j) s += sW[0][j + 1] * up[j];
        gl[PW - 1] = s * rh;
      }
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
    }

    // (1) the layer's u planes and their x/y ghosts
    if (act) {
#pragma unroll
      for (int i = 0; i < P1; ++i) {
        const int z = e * P1 + i;
        const T uz = uc[i];
        T* Ui = U + i * SU;
        Ui[(y + 1) * NP + x + 1] = uz;
        if (x == 0) Ui[(y + 1) * NP] = nxl >= 0 ? EG[0][y + N * z] : uz;
        if (x == N - 1) Ui[(y + 1) * NP + N + 1] = nxh >= 0 ? EG[1][y + N * z] : uz;
        if (y == 0) Ui[x + 1] = nyl >= 0 ? EG[2][x + N * z] : uz;
        if (y == N - 1) Ui[(N + 1) * NP + x + 1] = nyh >= 0 ? EG[3][x + N * z] : uz;
      }
    }
    lds_barrier();
    // (2) g_x and g_y, and the ghosts of the gradients
    if (act) {
#pragma unroll
      for (int i = 0; i < P1; ++i) {
        const int z = e * P1 + i;
        const T uz = uc[i];
        const T* Ui = U + i * SU;
        T gx = 0, gy = 0;
#pragma unroll
        for (int k = 0; k < PW; ++k) {
          gx += sW[ix][k] * Ui[(y + 1) * NP + (sx + k + 1)];
          gy += sW[iy][k] * Ui[(sy + k + 1) * NP + x + 1];
        }
        gx *= rh;
        gy *= rh;
        T* GXi = GX + i * SG;
        T* GYi = GY + i * SG;
        GXi[y * NP + x + 1] = gx;
        GYi[(y + 1) * N + x] = gy;
        if (x == 0) GXi[y * NP] = nxl >= 0 ? (EG[4][y + N * z] + Whi * uz) * rxl : -gx;
        if (x == N - 1) GXi[y * NP + N + 1] = nxh >= 0 ? (EG[5][y + N * z] + Wlo * uz) * rxh : -gx;
        if (y == 0) GYi[x] = nyl >= 0 ? (EG[6][x + N * z] + Whi * uz) * ryl : -gy;
        if (y == N - 1) GYi[(N + 1) * N + x] = nyh >= 0 ? (EG[7][x + N * z] + Wlo * uz) * ryh : -gy;
      }
    }
    lds_barrier();
    // (3) the Laplacian and the stage combination
    T w[P1];
#pragma unroll
    for (int i = 0; i < P1; ++i) {
      const int z = e * P1 + i;
      T unext = 0;
      w[i] = 0;
      if (act) {
        const T* GXi = GX + i * SG;
        const T* GYi = GY + i * SG;
        T lx = 0, ly = 0;
#pragma unroll
        for (int k = 0; k < PW; ++k) {
          lx += sW[ix][k] * GXi[y * NP + (sx + k + 1)];
          ly += sW[iy][k] * GYi[(sy + k + 1) * N + x];
        }
        const T lap = ((T(0) + lx) + ly) + lapl[i];
        const T kv = lap * rh;
        const int64_t iu = (int64_t)z * N2, iv = iu + nn;
        if (MODE == 1) {
          acc[iu] = ku[i];
          acc[iv] = kv;
          unext = yu[i] + A.a * ku[i];
          w[i] = yv[i] + A.a * kv;
        } else if (MODE == 2) {
          acc[iu] = au[i] + T(2) * ku[i];
          acc[iv] = av[i] + T(2) * kv;
          unext = yu[i] + A.a * ku[i];
          w[i] = yv[i] + A.a * kv;
        } else if (MODE == 3) {
          unext = yu[i] + A.a * (au[i] + ku[i]);
          w[i] = yv[i] + A.a * (av[i] + kv);
        } else {  // compensated: u by Kahan now, v's increment truncated first
          const T du = A.a * (au[i] + ku[i]);
          const T yk = du - lo[iu], hu = yu[i], tu = hu + yk;
          lo[iu] = (tu - hu) - yk;
          unext = tu;
          w[i] = A.a * (av[i] + kv);
        }
        yo[iu] = unext;
        UN[i * NT + tid] = unext;
      }
      // faces of the next u along z
      if (e == 0) pz_lo += sW[0][i + 1] * unext;
      if (e == E - 1) pz_hi += sW[P][i + 1] * unext;
      if (z == 0) uz0 = unext;
      if (z == N - 1) uzN = unext;
    }
    lds_barrier();
    // U is dead: the values to truncate go to VA (aliasing R1)
    if (act) {
#pragma unroll
      for (int i = 0; i < P1; ++i) VA[i * NT + tid] = w[i];
    }
    lds_barrier();
    // (4) faces of the next u along x and y; truncation along x (GX is dead: VB)
    if (act) {
#pragma unroll
      for (int i = 0; i < P1; ++i) {
        const int z = e * P1 + i;
        const T* UNi = UN + i * NT;
        if (x == 0) {
          T s = 0;
#pragma unroll
          for (int j = 0; j < P1; ++j) s += sW[0][j + 1] * UNi[y * N + j];
          A.uf_out[(b * 6 + 0) * N2 + y + N * z] = UNi[tid];
          A.pf_out[(b * 6 + 0) * N2 + y + N * z] = s;
        }
        if (x == N - 1) {
          T s = 0;
#pragma unroll
          for (int j = 0; j < P1; ++j) s += sW[P][j + 1] * UNi[y * N + N - P1 + j];
          A.uf_out[(b * 6 + 1) * N2 + y + N * z] = UNi[tid];
          A.pf_out[(b * 6 + 1) * N2 + y + N * z] = s;
        }
        if (y == 0) {
          T s = 0;
#pragma unroll
          for (int j = 0; j < P1; ++j) s += sW[0][j + 1] * UNi[j * N + x];
          A.uf_out[(b * 6 + 2) * N2 + x + N * z] = UNi[tid];
          A.pf_out[(b * 6 + 2) * N2 + x + N * z] = s;
        }
        if (y == N - 1) {
          T s = 0;
#pragma unroll
          for (int j = 0; j < P1; ++j) s += sW[P][j + 1] * UNi[(N - P1 + j) * N + x];
          A.uf_out[(b * 6 + 3) * N2 + x + N * z] = UNi[tid];
          A.pf_out[(b * 6 + 3) * N2 + x + N * z] = s;
        }
        const T* VAi = VA + i * NT;
        if (A.truncate) {
          T c = 0;
#pragma unroll
          for (int j = 0; j < P1; ++j) c += sDual[j] * VAi[y * N + (x - ix) + j];
          VB[i * NT + tid] = w[i] - sMode[ix] * c;
        } else {
          VB[i * NT + tid] = w[i];
        }
      }
    }
    lds_barrier();
    // (5) truncation along y, then along z in registers; v out
    if (act) {
      T wl[P1];
#pragma unroll
      for (int i = 0; i < P1; ++i) {
        const T* VBi = VB + i * NT;
        wl[i] = VBi[tid];
        if (A.truncate) {
          T c = 0;
#pragma unroll
          for (int j = 0; j < P1; ++j) c += sDual[j] * VBi[((y - iy) + j) * N + x];
          wl[i] -= sMode[iy] * c;
        }
      }
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
          const T yk = wl[j] - lo[iv], hv = yv[j], tv = hv + yk;
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
    for (int j = 0; j < P1; ++j) {
      uc[j] = up[j];
      up[j] = un2[j];
    }
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
