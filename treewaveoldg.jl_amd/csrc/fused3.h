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

// Waves per SIMD the fused stage is compiled for: four (128 VGPRs) lets two
// workgroups of the 20^3-node blocks share a CU, which hides the latency of
// one's barriers and loads behind the other's work.  FP64 then keeps the stage
// operands one plane at a time (NB 0).
#ifndef TWOLDG_FUSED_WAVES_F64
#define TWOLDG_FUSED_WAVES_F64 4
#endif
#ifndef TWOLDG_FUSED_WAVES_F32
#define TWOLDG_FUSED_WAVES_F32 4
#endif
#ifndef TWOLDG_FUSED_NB_F64
#define TWOLDG_FUSED_NB_F64 0
#endif
// Waves per SIMD that the workgroups the LDS admits on one CU can bring (the
// most loaded SIMD), capped by the above.
template <class T, int P1, int E>
constexpr int fused_waves_per_eu() {
  constexpr int N = P1 * E, NP = N + 2;
  constexpr long lds = (long)sizeof(T) *
      (P1 * (NP * NP + 2 * N * NP + N * N) + 4 * P1 * N + P1 * (P1 + 2) + P1);
  constexpr long wgs = (160L * 1024) / lds;
  constexpr long by_lds = (wgs * ((N * N + 63) / 64) + 3) / 4;
  constexpr int want = sizeof(T) == 8 ? TWOLDG_FUSED_WAVES_F64 : TWOLDG_FUSED_WAVES_F32;
  return by_lds < 1 ? 1 : (by_lds < want ? (int)by_lds : want);
}

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

// The element's operators by value: a kernel argument lives in the constant
// kernarg segment, so the uniform reads of the z-direction go to scalar
// registers instead of occupying vector registers.
template <class T, int P1>
struct FusedTab {
  T W[P1][P1 + 2];
  T mode[P1];
  T dual[P1];
};

template <class T, int P1, int E, int MODE>
__global__ void __launch_bounds__(((P1 * E) * (P1 * E) + 63) / 64 * 64, (fused_waves_per_eu<T, P1, E>()))
k_fused3(FusedArgs<T> A, FusedTab<T, P1> tb) {
  constexpr int N = P1 * E;
  constexpr int NT = N * N;
  constexpr int NTH = (NT + 63) / 64 * 64;
  constexpr int PW = P1 + 2;
  constexpr int P = P1 - 1;
  constexpr int NP = N + 2;
  constexpr int SU = NP * NP, SG = N * NP;  // plane strides of U and of GX, GY
  constexpr int NG = 8 * P1 * N;             // x/y ghost values of one layer
  constexpr int GPT = (NG + NTH - 1) / NTH;  // per thread
  // One element layer (P1 planes) per phase.  R1 holds the u planes, later the
  // values to truncate; R2 the x-gradients, later the x-truncated values.
  __shared__ T sW[P1 * PW];
  __shared__ T sMode[P1];
  __shared__ T R1[P1 * SU];
  __shared__ T R2[P1 * SG];
  __shared__ T GY[P1 * SG];
  __shared__ T UN[P1 * NT];
  __shared__ T EP[4][P1 * N];  // the x/y neighbours' partial face gradients
  T* U = R1;
  T* VA = R1;
  T* GX = R2;
  T* VB = R2;

  const int64_t b = blockIdx.x;
  const int tid = threadIdx.x;
  const bool act = tid < NT;
  const int x = act ? tid % N : 0, y = act ? tid / N : 0;
  constexpr int N2 = N * N, nn = N2 * N;

  if (tid < P1 * PW) sW[tid] = tb.W[tid / PW][tid % PW];
  if (tid < P1) sMode[tid] = tb.mode[tid];

  const T h = (T)A.spacing[b];
  const T rh = T(1) / h;
  const int32_t* nb6 = A.nbr + b * 6;
  // the block's arrays: uniform bases, per-thread 32-bit offsets from this
  // column (stride N2 along z; v at +nn)
  const int col = y * N + x;
  const T* us = A.ys + b * 2 * nn;
  const T* vs = us + nn;
  const T* un_ = A.yn + b * 2 * nn;
  T* acc = A.acc + b * 2 * nn;
  T* yo = A.yo + b * 2 * nn;
  T* lo = A.lo ? A.lo + b * 2 * nn : nullptr;
  T* ufo = A.uf_out + b * 6 * N2;
  T* pfo = A.pf_out + b * 6 * N2;
  const int nxl = nb6[0], nxh = nb6[1], nyl = nb6[2], nyh = nb6[3];

  auto load_layer = [&](int e, T* dst) {
#pragma unroll
    for (int j = 0; j < P1; ++j) dst[j] = act ? us[col + (e * P1 + j) * N2] : T(0);
  };
  // the ghost terms of rows 0 and P
  const T Wlo = tb.W[0][0], Whi = tb.W[P][P1 + 1];

  T uc[P1];
  T prev_last = 0, prev_part = 0;  // previous layer: last u, row-P partial sum
  load_layer(0, uc);

  const int ix = x % P1, iy = y % P1;
  const int sx = x - ix - 1, sy = y - iy - 1;  // window starts (owned index)

#pragma unroll 1
  for (int e = 0; e < E; ++e) {
    // ---- the layer's x/y ghosts: the neighbours' face values, issued first ----
    T gv[GPT];
#pragma unroll
    for (int k = 0; k < GPT; ++k) {
      const int idx = tid + k * NTH;
      gv[k] = T(0);
      if (idx < NG) {
        const int q = idx / (4 * P1 * N), r = idx % (4 * P1 * N), f = r / (P1 * N),
                  t = r % (P1 * N);  // t = (plane in layer) * N + tangential index
        const int nbq = f == 0 ? nxl : (f == 1 ? nxh : (f == 2 ? nyl : nyh));
        if (nbq >= 0)
          gv[k] = (q ? A.pf_in : A.uf_in)[((int64_t)nbq * 6 + (f ^ 1)) * N2 + e * P1 * N + t];
      }
    }

    // ---- the layer's z-gradients (one node either side) and z-Laplacian ----
    T lapl[P1];
    {
      // the z ghosts: the next layer's first u and its row-0 partial sum, or
      // on the block's z faces the neighbour's face u and face gradient (a
      // mirror's ghost u is the own face u, its gradient minus the own face
      // gradient)
      T before = prev_last, after, gbefore = 0, gafter = 0, next_part = 0;
      if (e == 0) {
        const int nlo = nb6[4];
        before = (act && nlo >= 0) ? A.uf_in[((int64_t)nlo * 6 + 5) * N2 + col] : uc[0];
        if (act && nlo >= 0) {
          gbefore = (A.pf_in[((int64_t)nlo * 6 + 5) * N2 + col] + Whi * uc[0]) * (T(1) / (T)A.spacing[nlo]);
        } else {
          T g0 = Wlo * before;
#pragma unroll
          for (int j = 0; j < P1; ++j) g0 += tb.W[0][j + 1] * uc[j];
          gbefore = -(g0 * rh);
        }
      }
      if (e < E - 1) {
        T un2[P1];
        load_layer(e + 1, un2);
        after = un2[0];
#pragma unroll
        for (int j = 0; j < P1; ++j) next_part += tb.W[0][j + 1] * un2[j];
      } else {
        const int nhi = nb6[5];
        after = (act && nhi >= 0) ? A.uf_in[((int64_t)nhi * 6 + 4) * N2 + col] : uc[P];
        if (act && nhi >= 0) {
          gafter = (A.pf_in[((int64_t)nhi * 6 + 4) * N2 + col] + Wlo * uc[P]) * (T(1) / (T)A.spacing[nhi]);
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
#pragma unroll
      for (int i = 0; i < P1; ++i) {
        T s = 0;
#pragma unroll
        for (int k = 0; k < PW; ++k) {
          if ((k == 0 && i != 0) || (k == PW - 1 && i != P)) continue;
          s += tb.W[i][k] * gl[k];
        }
        lapl[i] = s;
      }
    }
    prev_last = uc[P];
    prev_part = 0;
#pragma unroll
    for (int j = 0; j < P1; ++j) prev_part += tb.W[P][j + 1] * uc[j];

    // (1) the layer's u planes, and the x/y ghosts: neighbours' u into the
    // ghost ring of U, their partial gradients into EP; a mirror's ghost u is
    // the own face value
#pragma unroll
    for (int k = 0; k < GPT; ++k) {
      const int idx = tid + k * NTH;
      if (idx < NG) {
        const int q = idx / (4 * P1 * N), r = idx % (4 * P1 * N), f = r / (P1 * N),
                  t = r % (P1 * N), i = t / N, c = t % N;
        const int nbq = f == 0 ? nxl : (f == 1 ? nxh : (f == 2 ? nyl : nyh));
        if (nbq >= 0) {
          if (q) {
            EP[f][t] = gv[k];
          } else {
            T* Ui = U + i * SU;
            const int at = f == 0 ? (c + 1) * NP : f == 1 ? (c + 1) * NP + N + 1
                         : f == 2 ? c + 1 : (N + 1) * NP + c + 1;
            Ui[at] = gv[k];
          }
        }
      }
    }
    if (act) {
#pragma unroll
      for (int i = 0; i < P1; ++i) {
        const T uz = uc[i];
        T* Ui = U + i * SU;
        Ui[(y + 1) * NP + x + 1] = uz;
        if (x == 0 && nxl < 0) Ui[(y + 1) * NP] = uz;
        if (x == N - 1 && nxh < 0) Ui[(y + 1) * NP + N + 1] = uz;
        if (y == 0 && nyl < 0) Ui[x + 1] = uz;
        if (y == N - 1 && nyh < 0) Ui[(N + 1) * NP + x + 1] = uz;
      }
    }
    lds_barrier();
    // (2) g_x and g_y, and the ghosts of the gradients
    if (act) {
      const T* wx = sW + ix * PW;
      const T* wy = sW + iy * PW;
#pragma unroll 1
      for (int i = 0; i < P1; ++i) {
        const T* Ui = U + i * SU;
        const T uz = Ui[(y + 1) * NP + x + 1];
        T gx = 0, gy = 0;
#pragma unroll
        for (int k = 0; k < PW; ++k) {
          gx += wx[k] * Ui[(y + 1) * NP + (sx + k + 1)];
          gy += wy[k] * Ui[(sy + k + 1) * NP + x + 1];
        }
        gx *= rh;
        gy *= rh;
        T* GXi = GX + i * SG;
        T* GYi = GY + i * SG;
        GXi[y * NP + x + 1] = gx;
        GYi[(y + 1) * N + x] = gy;
        if (x == 0)
          GXi[y * NP] = nxl >= 0 ? (EP[0][i * N + y] + Whi * uz) * (T(1) / (T)A.spacing[nxl]) : -gx;
        if (x == N - 1)
          GXi[y * NP + N + 1] = nxh >= 0 ? (EP[1][i * N + y] + Wlo * uz) * (T(1) / (T)A.spacing[nxh]) : -gx;
        if (y == 0)
          GYi[x] = nyl >= 0 ? (EP[2][i * N + x] + Whi * uz) * (T(1) / (T)A.spacing[nyl]) : -gy;
        if (y == N - 1)
          GYi[(N + 1) * N + x] = nyh >= 0 ? (EP[3][i * N + x] + Wlo * uz) * (T(1) / (T)A.spacing[nyh]) : -gy;
      }
    }
    lds_barrier();
    // (3) the Laplacian and the stage combination; U is dead, VA takes its place
    {
      // the stage operands: the whole layer at once where registers allow
      // (FP32), plane by plane otherwise
      constexpr int NB = sizeof(T) == 8 ? TWOLDG_FUSED_NB_F64 : P1;
      T ku[P1], yu[P1], yv[P1], au[P1], av[P1];
      auto load_ops = [&](int i) {
        ku[i] = yu[i] = yv[i] = au[i] = av[i] = T(0);
        if (act) {
          const int iu = col + (e * P1 + i) * N2, iv = iu + nn;
          ku[i] = vs[iu];
          if (MODE != 4) {
            yu[i] = un_[iu];
            yv[i] = un_[iv];
          }
          if (MODE >= 2) {
            au[i] = acc[iu];
            av[i] = acc[iv];
          }
        }
      };
#pragma unroll
      for (int i = 0; i < NB; ++i) load_ops(i);
      T pzl = 0, pzh = 0;  // partial z-face gradients of the next u
      const T* wx = sW + ix * PW;
      const T* wy = sW + iy * PW;
#pragma unroll
      for (int i = 0; i < P1; ++i) {
        const int z = e * P1 + i;
        if (i + NB < P1) load_ops(i + NB);
        T unext = 0;
        if (act) {
          const int iu = col + z * N2, iv = iu + nn;
          const T* GXi = GX + i * SG;
          const T* GYi = GY + i * SG;
          T lx = 0, ly = 0;
#pragma unroll
          for (int k = 0; k < PW; ++k) {
            lx += wx[k] * GXi[y * NP + (sx + k + 1)];
            ly += wy[k] * GYi[(sy + k + 1) * N + x];
          }
          const T lap = ((T(0) + lx) + ly) + lapl[i];
          const T kv = lap * rh;
          T w;
          if (MODE == 1) {
            acc[iu] = ku[i];
            acc[iv] = kv;
            unext = yu[i] + A.a * ku[i];
            w = yv[i] + A.a * kv;
          } else if (MODE == 2) {
            acc[iu] = au[i] + T(2) * ku[i];
            acc[iv] = av[i] + T(2) * kv;
            unext = yu[i] + A.a * ku[i];
            w = yv[i] + A.a * kv;
          } else if (MODE == 3) {
            unext = yu[i] + A.a * (au[i] + ku[i]);
            w = yv[i] + A.a * (av[i] + kv);
          } else {  // compensated: u by Kahan now, v's increment truncated first
            const T du = A.a * (au[i] + ku[i]);
            const T yk = du - lo[iu], hu = un_[iu], tu = hu + yk;
            lo[iu] = (tu - hu) - yk;
            unext = tu;
            w = A.a * (av[i] + kv);
          }
          yo[iu] = unext;
          UN[i * NT + tid] = unext;
          VA[i * NT + tid] = w;
        }
        // faces of the next u along z
        if (e == 0) pzl += tb.W[0][i + 1] * unext;
        if (e == E - 1) pzh += tb.W[P][i + 1] * unext;
        if (act && z == 0) ufo[4 * N2 + col] = unext;
        if (act && z == N - 1) ufo[5 * N2 + col] = unext;
      }
      if (act && e == 0) pfo[4 * N2 + col] = pzl;
      if (act && e == E - 1) pfo[5 * N2 + col] = pzh;
    }
    lds_barrier();
    // (4) faces of the next u along x and y; truncation along x (GX is dead: VB)
    if (act) {
#pragma unroll 1
      for (int i = 0; i < P1; ++i) {
        const int z = e * P1 + i;
        const T* UNi = UN + i * NT;
        if (x == 0) {
          T s = 0;
#pragma unroll
          for (int j = 0; j < P1; ++j) s += tb.W[0][j + 1] * UNi[y * N + j];
          ufo[0 * N2 + y + N * z] = UNi[tid];
          pfo[0 * N2 + y + N * z] = s;
        }
        if (x == N - 1) {
          T s = 0;
#pragma unroll
          for (int j = 0; j < P1; ++j) s += tb.W[P][j + 1] * UNi[y * N + N - P1 + j];
          ufo[1 * N2 + y + N * z] = UNi[tid];
          pfo[1 * N2 + y + N * z] = s;
        }
        if (y == 0) {
          T s = 0;
#pragma unroll
          for (int j = 0; j < P1; ++j) s += tb.W[0][j + 1] * UNi[j * N + x];
          ufo[2 * N2 + x + N * z] = UNi[tid];
          pfo[2 * N2 + x + N * z] = s;
        }
        if (y == N - 1) {
          T s = 0;
#pragma unroll
          for (int j = 0; j < P1; ++j) s += tb.W[P][j + 1] * UNi[(N - P1 + j) * N + x];
          ufo[3 * N2 + x + N * z] = UNi[tid];
          pfo[3 * N2 + x + N * z] = s;
        }
        const T* VAi = VA + i * NT;
        if (A.truncate) {
          T c = 0;
#pragma unroll
          for (int j = 0; j < P1; ++j) c += tb.dual[j] * VAi[y * N + (x - ix) + j];
          VB[i * NT + tid] = VAi[tid] - sMode[ix] * c;
        } else {
          VB[i * NT + tid] = VAi[tid];
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
          for (int j = 0; j < P1; ++j) c += tb.dual[j] * VBi[((y - iy) + j) * N + x];
          wl[i] -= sMode[iy] * c;
        }
      }
      if (A.truncate) {
        T c = 0;
#pragma unroll
        for (int j = 0; j < P1; ++j) c += tb.dual[j] * wl[j];
#pragma unroll
        for (int j = 0; j < P1; ++j) wl[j] -= tb.mode[j] * c;
      }
#pragma unroll
      for (int j = 0; j < P1; ++j) {
        const int iv = col + (e * P1 + j) * N2 + nn;
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
    if (e + 1 < E) load_layer(e + 1, uc);
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
