// regrid.h -- the modal refinement indicator and the transfer of a state from
// one forest to its regridded successor (TreeWaveOLDG.jl src/refinement.jl,
// src/blast.jl:336-348; TreeAMR's regrid! with Operators(family=ElementLocal)).
//
// Both are element-local and need no ghosts.  The indicator reads, along every
// line of an element in every dimension, the Legendre coefficients
// c_m = sum_j duals[m][j] a_j; the transfer refines a block by evaluating the
// parent's element polynomials at the child's nodes (the tabulated I, tensor
// over the dimensions) and coarsens 2^D children by the exact L2 projection
// R = M_c^-1 I^T M_f, both of which keep the quadrature integral of the data.
#pragma once

// The element's value is, with A the global peak of the variable (*peak, from
// the norms' reduction),
//   post * max over its lines and q of ((wt[q] / A) |c_m[q]|)^pw[q]
// (refinement.jl:56-75, 82-101).
using ModeArgs = twoldg_mode_args;

template <class T, int D, int P1>
__global__ void k_modes(Shape sh, int E, const T* __restrict__ y, const T* __restrict__ lo,
                        int var, ModeArgs M, const double* __restrict__ peak,
                        double* __restrict__ tau) {
  const twoldg::ElementTables& et = twoldg::kElementTables[P1 - 2];
  int64_t ne = E;
  if (D >= 2) ne *= E;
  if (D >= 3) ne *= E;
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= sh.nb * ne) return;
  const int64_t b = g / ne, e = g % ne;
  const int N = sh.N;
  const int J[3] = {(int)(e % E), D >= 2 ? (int)((e / E) % E) : 0,
                    D >= 3 ? (int)(e / ((int64_t)E * E)) : 0};
  const T* a = y + (b * 2 + var) * sh.nn;
  const T* al = lo ? lo + (b * 2 + var) * sh.nn : nullptr;
  auto val = [&](const int I[3]) -> double {
    const int64_t l = I[0] + (int64_t)N * (I[1] + (int64_t)N * I[2]);
    return al ? (double)a[l] - (double)al[l] : (double)a[l];
  };
  // A field that vanishes everywhere scores zero rather than NaN.
  const double pk = *peak, scale = pk == 0 ? 0.0 : 1.0 / pk;
  double mv = 0;
#pragma unroll
  for (int d = 0; d < D; ++d) {
    const int t0 = d == 0 ? 1 : 0, t1 = d == 2 ? 1 : 2;
    const int n0 = D >= 2 ? P1 : 1, n1 = D >= 3 ? P1 : 1;
    for (int k1 = 0; k1 < n1; ++k1)
      for (int k0 = 0; k0 < n0; ++k0) {
        int I[3] = {J[0] * P1, J[1] * P1, J[2] * P1};
        if (D >= 2) I[t0] += k0;
        if (D >= 3) I[t1] += k1;
        const int i0 = I[d];
        for (int q = 0; q < M.nq; ++q) {
          const double* du = et.duals[M.m[q]];
          double c = 0;
          for (int j = 0; j < P1; ++j) {
            I[d] = i0 + j;
            c += du[j] * val(I);
          }
          I[d] = i0;
          const double x = (M.wt[q] * scale) * fabs(c);
          mv = fmax(mv, M.pw[q] == 1.0 ? x : pow(x, M.pw[q]));
        }
      }
  }
  tau[g] = M.post * mv;
}

// One destination block's origin in the source forest: the ABI's
// twoldg_transfer_block (kind copy, refine from the parent src[0] with the
// block's offset `child` in it, bit d per dimension, or coarsen from the 2^D
// children src[c]).
using TransferBlock = twoldg_transfer_block;

template <class T, int D, int P1>
__global__ void k_transfer(Shape sh, int E, const TransferBlock* __restrict__ map,
                           const T* __restrict__ src, const T* __restrict__ slo,
                           T* __restrict__ dst, T* __restrict__ dlo) {
  const twoldg::ElementTables& et = twoldg::kElementTables[P1 - 2];
  const int64_t b = blockIdx.x;
  const TransferBlock tb = map[b];
  const int N = sh.N, H = E / 2;
  auto at = [&](int64_t blk, int var, const int I[3]) -> double {
    const int64_t i = (blk * 2 + var) * sh.nn + I[0] + (int64_t)N * (I[1] + (int64_t)N * I[2]);
    return slo ? (double)src[i] - (double)slo[i] : (double)src[i];
  };
  for (int64_t l = threadIdx.x; l < sh.nn; l += blockDim.x) {
    int I[3] = {0, 0, 0};
    unravel<D>(l, N, I);
    for (int var = 0; var < 2; ++var) {
      double v = 0;
      if (tb.kind == 0) {
        v = at(tb.src[0], var, I);
      } else if (tb.kind == 1) {
        // child node i along d: its element ec = i / P1 lies in the parent's
        // element ep = o H + ec / 2 as fine half f = ec % 2 (fine row f P1 + j)
        int fn[3] = {0, 0, 0}, ep[3] = {0, 0, 0};
        for (int d = 0; d < D; ++d) {
          const int ec = I[d] / P1, j = I[d] % P1, o = (tb.child >> d) & 1;
          ep[d] = o * H + ec / 2;
          fn[d] = (ec % 2) * P1 + j;
        }
        const int nk = D == 1 ? P1 : (D == 2 ? P1 * P1 : P1 * P1 * P1);
        for (int k = 0; k < nk; ++k) {
          const int kk[3] = {k % P1, (k / P1) % P1, k / (P1 * P1)};
          double w = 1;
          int S[3] = {0, 0, 0};
          for (int d = 0; d < D; ++d) {
            w *= et.interp[fn[d]][kk[d]];
            S[d] = ep[d] * P1 + kk[d];
          }
          v += w * at(tb.src[0], var, S);
        }
      } else {
        // parent node i along d: element ep = i / P1, coarse node ic; its fine
        // data come from the child o = ep / H, whose elements 2 (ep - o H) + f
        int ic[3] = {0, 0, 0}, eb[3] = {0, 0, 0}, ch = 0;
        for (int d = 0; d < D; ++d) {
          const int e = I[d] / P1, o = e / H;
          ic[d] = I[d] % P1;
          eb[d] = 2 * (e - o * H);
          ch |= o << d;
        }
        const int F = 2 * P1;
        const int nk = D == 1 ? F : (D == 2 ? F * F : F * F * F);
        for (int k = 0; k < nk; ++k) {
          const int kk[3] = {k % F, (k / F) % F, k / (F * F)};
          double w = 1;
          int S[3] = {0, 0, 0};
          for (int d = 0; d < D; ++d) {
            w *= et.restrict_[ic[d]][kk[d]];
            S[d] = (eb[d] + kk[d] / P1) * P1 + kk[d] % P1;
          }
          v += w * at(tb.src[ch], var, S);
        }
      }
      const int64_t o = (b * 2 + var) * sh.nn + l;
      const T h = (T)v;
      dst[o] = h;
      if (dlo) dlo[o] = (T)((double)h - v);  // minus the low part, as an upload sets it
    }
  }
}
