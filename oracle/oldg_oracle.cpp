// OLDG oracle: a CPU FP64 restatement of TreeWaveOLDG.jl's field-evolution path.
//
// TEST INFRASTRUCTURE ONLY.  This file is the checker the parity tests, the
// smoke check and the benchmark's CPU-baseline leg compare the HIP library
// against.  Nothing in the product (the HIP library or its Python host) links,
// loads or calls it.
//
// It restates, with its own code and its own tables, the reference's
//   * element tables            src/element.jl:72-170  (built here in
//                               binary128 and rounded once to binary64)
//   * forests                   src/sinewave.jl:114-199 (brick, refined
//                               regions level by level, face 2:1 balance),
//                               src/geometry.jl:305-553 (seven-patch ball)
//   * ghost fill                TreeAMR fill_ghosts! as used by
//                               src/evolution.jl:1303-1348: same level,
//                               periodic, mirror (parity), coarse-fine
//                               ElementLocal I and R = M_c^-1 I^T M_f, glue,
//                               Dirichlet
//   * gradient / Laplacian      src/evolution.jl:21-110 (flat) and 136-220
//                               (full metric, Cartesian gradient)
//   * truncation                src/evolution.jl:222-264 (tensor) and
//                               279-425 (per-element Cholesky projection)
//   * RK4 with limiters         src/stepping.jl:72-78 (classic RK4, stage and
//                               step limiter = truncation)
//   * energy and norms          src/evolution.jl:1472-1512, src/layout.jl:95-235
//   * time step                 src/evolution.jl:1528-1531
//
// Pinning: see tests/test_oracle_*.py.  The element tables are checked
// against the paper's printed matrices and closed-form invariants; the RHS
// against a dense global operator and against polynomial exactness; the
// coarse-fine coupling against skew-adjointness and dE/dt = 0; the glue
// against dE/dt = 0 across every glue; the whole path against the
// reference's published convergence tables.  The ghost sources on glued faces
// are found by physical-coordinate coincidence, independent of any glue-map
// convention.
//
// Build: make -C oracle   (g++ -O2 -fopenmp, links libquadmath)

#include <quadmath.h>
#include <omp.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <tuple>
#include <unordered_map>
#include <vector>

typedef __float128 qf;

namespace {

// ---------------------------------------------------------------------------
// Element tables in binary128, rounded once (element.jl:72-170).
// ---------------------------------------------------------------------------
struct Element {
  int P = 0, P1 = 0;
  double nodes[9], weights[9], offsets[9];
  double W[9][11];
  double trunc[9][9];
  double mode[9], dual[9];
  double leg[9][9];  // L_k(xi_i): [i][k]
  double duals[9][9];  // [m][j]: c_m = sum_j duals[m][j] u_j along a line
};

static void legendre_q(int P, qf x, qf* L) {
  L[0] = 1;
  if (P >= 1) L[1] = x;
  for (int k = 2; k <= P; ++k) L[k] = ((2 * k - 1) * x * L[k - 1] - (k - 1) * L[k - 2]) / k;
}

static void build_element(int P, Element& e) {
  e.P = P;
  e.P1 = P + 1;
  const int P1 = P + 1;
  qf x[9], w[9], L[9];
  for (int i = 0; i <= P; ++i) x[i] = -cosq(M_PIq * i / P);
  for (int it = 0; it < 200; ++it) {
    qf dmax = 0, dl[9];
    for (int i = 0; i <= P; ++i) {
      legendre_q(P, x[i], L);
      dl[i] = (x[i] * L[P] - L[P - 1]) / ((P + 1) * L[P]);
      dmax = fmaxq(dmax, fabsq(dl[i]));
    }
    for (int i = 0; i <= P; ++i) x[i] -= dl[i];
    if (dmax <= ldexpq(1, -106)) break;
  }
  qf xs[9];
  for (int i = 0; i <= P; ++i) xs[i] = (x[i] - x[P - i]) / 2;
  xs[0] = -1;
  xs[P] = 1;
  qf LP[9], Lall[9][9];
  for (int i = 0; i <= P; ++i) {
    legendre_q(P, xs[i], Lall[i]);
    LP[i] = Lall[i][P];
    w[i] = 2 / (qf(P) * (P + 1) * LP[i] * LP[i]);
  }
  qf d[9][9];
  for (int i = 0; i < P1; ++i)
    for (int j = 0; j < P1; ++j) d[i][j] = (i == j) ? 0 : LP[i] / (LP[j] * (xs[i] - xs[j]));
  d[0][0] = -qf(P) * (P + 1) / 4;
  d[P][P] = qf(P) * (P + 1) / 4;
  qf Wq[9][11];
  for (int i = 0; i < P1; ++i)
    for (int j = 0; j < P1 + 2; ++j) Wq[i][j] = 0;
  for (int i = 0; i < P1; ++i)
    for (int j = 0; j < P1; ++j) Wq[i][j + 1] = 2 * d[i][j] / (P + 1);
  qf face = qf(P) / 2;
  Wq[0][1] += face;
  Wq[0][0] -= face;
  Wq[P][P1] -= face;
  Wq[P][P1 + 1] += face;
  qf gamma = 0;
  for (int i = 0; i < P1; ++i) gamma += w[i] * LP[i] * LP[i];
  for (int i = 0; i < P1; ++i) {
    e.nodes[i] = (double)xs[i];
    e.weights[i] = (double)w[i];
    e.offsets[i] = (double)((P + 1) * (1 + xs[i]) / 2);
    e.mode[i] = (double)LP[i];
    e.dual[i] = (double)(w[i] * LP[i] / gamma);
    for (int j = 0; j < P1 + 2; ++j) e.W[i][j] = (double)Wq[i][j];
    for (int j = 0; j < P1; ++j)
      e.trunc[i][j] = (double)((i == j ? qf(1) : qf(0)) - LP[i] * w[j] * LP[j] / gamma);
    for (int k = 0; k < P1; ++k) e.leg[i][k] = (double)Lall[i][k];
  }
  // The Legendre coefficients by discrete orthogonality under the
  // Gauss-Lobatto rule, sum_i w_i L_m L_k = 0 for m != k <= P
  // (refinement.jl:45-54): c_m = sum_j w_j L_m(xi_j) u_j / sum_i w_i L_m(xi_i)^2.
  for (int m = 0; m < P1; ++m) {
    qf nrm = 0;
    for (int i = 0; i < P1; ++i) nrm += w[i] * Lall[i][m] * Lall[i][m];
    for (int j = 0; j < P1; ++j) e.duals[m][j] = (double)(w[j] * Lall[j][m] / nrm);
  }
  e.offsets[0] = 0;
  e.offsets[P] = P1;
}

// Lagrange basis polynomial i of the element's nodes at reference coordinate x.
static double lagrange(const Element& el, int i, double x) {
  double r = 1;
  for (int j = 0; j < el.P1; ++j)
    if (j != i) r *= (x - el.nodes[j]) / (el.nodes[i] - el.nodes[j]);
  return r;
}

static double cospi_(double z) {
  double r = std::fmod(std::fabs(z), 2.0);
  if (r > 1) r = 2 - r;
  if (r <= 0.25) return std::cos(M_PI * r);
  if (r <= 0.75) return std::sin(M_PI * (0.5 - r));
  return -std::cos(M_PI * (1 - r));
}

// ---------------------------------------------------------------------------
// Seven-patch ball map (geometry.jl:270-400) and its frame.
// ---------------------------------------------------------------------------
struct Ball {
  double Lc = 1, R1 = 3;
  int n = 2, m = 2;
  double grading = 0;
  // the thirteen-patch ball (geometry.jl:555-749): inner shells out to Rm,
  // exactly radial outer shells (patches 7..12) of mo roots out to R1
  bool thirteen = false;
  double Rm = 0;
  int mo = 0;
};

static inline double exp_(double a) { return std::exp(a); }
static inline double sqrt_(double a) { return std::sqrt(a); }
static inline qf exp_(qf a) { return expq(a); }
static inline qf sqrt_(qf a) { return sqrtq(a); }

// Dual numbers for the Jacobian (forward mode, three directions).
struct Dual {
  double v;
  double g[3];
  Dual(double a = 0) : v(a) { g[0] = g[1] = g[2] = 0; }
};
static inline Dual operator+(Dual a, Dual b) { Dual r(a.v + b.v); for (int i = 0; i < 3; ++i) r.g[i] = a.g[i] + b.g[i]; return r; }
static inline Dual operator-(Dual a, Dual b) { Dual r(a.v - b.v); for (int i = 0; i < 3; ++i) r.g[i] = a.g[i] - b.g[i]; return r; }
static inline Dual operator*(Dual a, Dual b) { Dual r(a.v * b.v); for (int i = 0; i < 3; ++i) r.g[i] = a.g[i] * b.v + a.v * b.g[i]; return r; }
static inline Dual operator/(Dual a, Dual b) { Dual r(a.v / b.v); for (int i = 0; i < 3; ++i) r.g[i] = (a.g[i] * b.v - a.v * b.g[i]) / (b.v * b.v); return r; }
static inline Dual exp_(Dual a) { Dual r(std::exp(a.v)); for (int i = 0; i < 3; ++i) r.g[i] = r.v * a.g[i]; return r; }
static inline Dual sqrt_(Dual a) { Dual r(std::sqrt(a.v)); for (int i = 0; i < 3; ++i) r.g[i] = a.g[i] / (2 * r.v); return r; }

template <class R>
static void ball_position(const Ball& g, int p, const R xi[3], R x[3]) {
  if (p == 0) {  // cube: (2Lc/n) xi - Lc
    for (int d = 0; d < 3; ++d) x[d] = (R(2 * g.Lc) / R(g.n)) * xi[d] - R(g.Lc);
    return;
  }
  if (g.thirteen && p >= 7) {
    // outer shell over (k, s): x = (Rm + (R1 - Rm) rho) c/|c|, c_k = s,
    // c_j = 2 xi_j/n - 1, rho = xi_k/mo (s = 1), 1 - xi_k/mo (s = -1)
    int k = (p - 7) / 2, s = ((p - 7) % 2 == 0) ? -1 : 1;
    R c[3];
    for (int j = 0; j < 3; ++j) c[j] = (j == k) ? R(s) : R(2) * xi[j] / R(g.n) - R(1);
    R rho = (s == 1) ? xi[k] / R(g.mo) : R(1) - xi[k] / R(g.mo);
    R r = R(g.Rm) + (R(g.R1) - R(g.Rm)) * rho;
    R nc = sqrt_(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
    for (int j = 0; j < 3; ++j) x[j] = (r / nc) * c[j];
    return;
  }
  const double Rout = g.thirteen ? g.Rm : g.R1;  // the inner shells end on Rm
  int k = (p - 1) / 2, s = ((p - 1) % 2 == 0) ? -1 : 1;
  R f[3];
  for (int j = 0; j < 3; ++j)
    f[j] = (j == k) ? R(s) * R(g.Lc) : R(g.Lc) * (R(2) * xi[j] / R(g.n) - R(1));
  R eta = (s == 1) ? xi[k] / R(g.m) : R(1) - xi[k] / R(g.m);
  R rho = eta;
  if (g.grading != 0) {
    R a = R(g.grading);
    rho = (exp_(a * eta) - R(1)) / (exp_(a) - R(1));
  }
  R nf = sqrt_(f[0] * f[0] + f[1] * f[1] + f[2] * f[2]);
  for (int j = 0; j < 3; ++j) x[j] = (R(1) - rho) * f[j] + (rho * R(Rout) / nf) * f[j];
}


// (J, S) with S[d][i] = J * (dxi^d/dx^i), J = |det dx/dxi|  (geometry.jl:402-409).
static void ball_frame(const Ball& g, int p, const double xi[3], double& J, double S[3][3]) {
  Dual X[3], Y[3];
  for (int d = 0; d < 3; ++d) {
    X[d] = Dual(xi[d]);
    X[d].g[d] = 1;
  }
  ball_position<Dual>(g, p, X, Y);
  double M[3][3];  // M[i][j] = dx^i/dxi^j
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) M[i][j] = Y[i].g[j];
  double det = M[0][0] * (M[1][1] * M[2][2] - M[1][2] * M[2][1]) -
               M[0][1] * (M[1][0] * M[2][2] - M[1][2] * M[2][0]) +
               M[0][2] * (M[1][0] * M[2][1] - M[1][1] * M[2][0]);
  J = std::fabs(det);
  // inverse: Minv[d][i] = cofactor(i,d)/det
  double C[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      int i1 = (i + 1) % 3, i2 = (i + 2) % 3, j1 = (j + 1) % 3, j2 = (j + 2) % 3;
      C[i][j] = M[i1][j1] * M[i2][j2] - M[i1][j2] * M[i2][j1];
    }
  for (int d = 0; d < 3; ++d)
    for (int i = 0; i < 3; ++i) S[d][i] = J * (C[i][d] / det);
}

static double ball_volume_q(const Ball& g, int p, const qf xi[3]) {
  // sqrt(g) in binary128 by central differences would lose digits; use the
  // analytic Jacobian of the map in binary128 through hyper-dual-free
  // finite differences is avoided: evaluate in double with duals instead.
  double x[3] = {(double)xi[0], (double)xi[1], (double)xi[2]};
  double J, S[3][3];
  ball_frame(g, p, x, J, S);
  return J;
}

// ---------------------------------------------------------------------------
// Forest.
// ---------------------------------------------------------------------------
struct Leaf {
  int patch, level;
  int64_t c[3];
};

struct Key {
  int patch, level;
  int64_t c[3];
  bool operator==(const Key& o) const {
    return patch == o.patch && level == o.level && c[0] == o.c[0] && c[1] == o.c[1] && c[2] == o.c[2];
  }
};
struct KeyHash {
  size_t operator()(const Key& k) const {
    uint64_t h = (uint64_t)k.patch * 1000003u + (uint64_t)k.level * 7919u;
    for (int d = 0; d < 3; ++d) h = h * 0x9E3779B97F4A7C15ull + (uint64_t)k.c[d];
    return (size_t)(h ^ (h >> 29));
  }
};

enum GhostKind { G_NONE = 0, G_SAME, G_COARSER, G_FINER, G_MIRROR, G_GLUE, G_DIRICHLET,
                 G_RADIATIVE };

struct Face {
  int kind = G_NONE;
  int64_t nb = -1;              // same / coarser / glue: neighbour block
  int64_t fine[4] = {-1, -1, -1, -1};  // finer: children by tangential offset
  int64_t off[3] = {0, 0, 0};   // coarser: our offset inside the coarse block (0/1 per dim)
  // glue: per ghost node the neighbour's owned node index (linear, block nb)
  std::vector<int64_t> glue_src;
  int glue_plane = 0;
};

struct Problem {
  int D = 2, N = 0, P = 0, E = 0, P1 = 0;
  Element el;
  int geometry = 0;  // 0 brick, 1 seven-patch ball, 2 annulus (diagonal metric)
  Ball ball;
  int npatch = 1;
  int proots[13][3];
  double plo[13][3], phi_[13][3];
  bool periodic[3] = {true, true, true};
  bool reflecting = false;
  bool truncate = true;
  std::vector<Leaf> leaves;
  std::vector<double> h;  // per block
  std::vector<Face> faces;  // nb * 2D
  int64_t nb = 0, nn = 0;   // blocks, nodes per block
  std::unordered_map<Key, int64_t, KeyHash> index;
  std::unordered_map<Key, int, KeyHash> nodes_exist;  // every tree node (leaf or not)
  // full metric
  std::vector<double> coef;    // (nb, 1 + D*D, nn)
  std::vector<uint8_t> flat;   // per block
  // element truncation tables (per shell block): weights (n, E^D) and chol (m, m, E^D)
  std::vector<int64_t> slot;
  std::vector<double> tw, tchol, tmodes;
  int tn = 0, tm = 0;
  double hspeed = 0;
  // the radiative boundary (evolution.jl:1129-1299) on the outer sphere
  bool radiative = false;
  double rad_kappa = 0, rad_invR = 0;
  const double* ycur = nullptr;  // the state whose gradient is being formed
  // Dirichlet plane wave
  double kvec[3] = {0, 0, 0}, omega = 0, phase = 0;
  // annulus (geometry 2): r0 <= r <= r1, four patches of (n1, n2) roots
  double an_r0 = 1, an_r1 = 2, an_dth = 0, an_dr = 0;
  // diagonal metric: (nb, 1 + D, nn) sqrt(g), A_1..A_D; per-line truncation
  // pairs (nb, D, E, P1)
  std::vector<double> dcoef, lmode, ldual;
  // work arrays
  std::vector<double> ufs;            // (nb, (N+2)^D)
  std::vector<double> grads[3];       // (nb, prod(N + 2 delta_d))
  std::vector<double> cart;           // (nb, D, (N+2)^D)
  int64_t ufs_n = 0, grad_n[3] = {0, 0, 0};
};

static int64_t ipow(int64_t a, int b) {
  int64_t r = 1;
  for (int i = 0; i < b; ++i) r *= a;
  return r;
}

static void pdims(const Problem& pr, int p, int level, int64_t n[3]) {
  for (int d = 0; d < 3; ++d) n[d] = d < pr.D ? (int64_t)pr.proots[p][d] << level : 1;
}

static void leaf_extent(const Problem& pr, const Leaf& L, double lo[3], double hi[3]) {
  for (int d = 0; d < pr.D; ++d) {
    double w = (pr.phi_[L.patch][d] - pr.plo[L.patch][d]) / pr.proots[L.patch][d] / (double)(1LL << L.level);
    lo[d] = pr.plo[L.patch][d] + w * (double)L.c[d];
    hi[d] = lo[d] + w;
  }
}

static double leaf_spacing(const Problem& pr, const Leaf& L) {
  double w = (pr.phi_[L.patch][0] - pr.plo[L.patch][0]) / pr.proots[L.patch][0] / (double)(1LL << L.level);
  return w / pr.N;
}

// Morton order inside a patch's root grid: patch, root (x fastest), then the
// interleaved path bits from the root down (x lowest in each group).
static bool leaf_less(const Problem& pr, const Leaf& a, const Leaf& b, int maxlev) {
  if (a.patch != b.patch) return a.patch < b.patch;
  auto rootlin = [&](const Leaf& L) {
    int64_t r = 0, s = 1;
    for (int d = 0; d < pr.D; ++d) {
      r += (L.c[d] >> L.level) * s;
      s *= pr.proots[L.patch][d];
    }
    return r;
  };
  int64_t ra = rootlin(a), rb = rootlin(b);
  if (ra != rb) return ra < rb;
  for (int l = 1; l <= maxlev; ++l) {
    int ca = 0, cb = 0;
    for (int d = pr.D - 1; d >= 0; --d) {
      int ba = (l <= a.level) ? (int)((a.c[d] >> (a.level - l)) & 1) : 0;
      int bb = (l <= b.level) ? (int)((b.c[d] >> (b.level - l)) & 1) : 0;
      ca = ca * 2 + ba;
      cb = cb * 2 + bb;
    }
    if (ca != cb) return ca < cb;
  }
  return a.level < b.level;
}

static void finalize_leaves(Problem& pr) {
  int maxlev = 0;
  for (auto& L : pr.leaves) maxlev = std::max(maxlev, L.level);
  std::sort(pr.leaves.begin(), pr.leaves.end(),
            [&](const Leaf& a, const Leaf& b) { return leaf_less(pr, a, b, maxlev); });
  pr.nb = (int64_t)pr.leaves.size();
  pr.index.clear();
  for (int64_t b = 0; b < pr.nb; ++b) {
    const Leaf& L = pr.leaves[b];
    Key k{L.patch, L.level, {L.c[0], L.c[1], L.c[2]}};
    pr.index[k] = b;
  }
  pr.h.resize(pr.nb);
  for (int64_t b = 0; b < pr.nb; ++b) pr.h[b] = leaf_spacing(pr, pr.leaves[b]);
}

static bool is_leaf(const Problem& pr, int patch, int level, const int64_t c[3], int64_t* out = nullptr) {
  Key k{patch, level, {c[0], c[1], c[2]}};
  auto it = pr.index.find(k);
  if (it == pr.index.end()) return false;
  if (out) *out = it->second;
  return true;
}

// Refine leaves level by level (sinewave.jl:141-160); balance faces 2:1.
static void refine_leaf(std::vector<Leaf>& leaves, size_t i, int D) {
  Leaf L = leaves[i];
  leaves.erase(leaves.begin() + i);
  for (int ch = 0; ch < (1 << D); ++ch) {
    Leaf C = L;
    C.level = L.level + 1;
    for (int d = 0; d < 3; ++d) C.c[d] = (d < D) ? 2 * L.c[d] + ((ch >> d) & 1) : 0;
    leaves.push_back(C);
  }
}

static void rebuild_index(Problem& pr) {
  pr.index.clear();
  for (int64_t b = 0; b < (int64_t)pr.leaves.size(); ++b) {
    const Leaf& L = pr.leaves[b];
    pr.index[Key{L.patch, L.level, {L.c[0], L.c[1], L.c[2]}}] = b;
  }
}

// The leaf at or above the cell (patch, level, c), or -1 if the cell is refined.
static int64_t covering_leaf(const Problem& pr, int patch, int level, const int64_t c[3], int* lev_out) {
  int64_t cc[3] = {c[0], c[1], c[2]};
  for (int l = level; l >= 0; --l) {
    int64_t b;
    if (is_leaf(pr, patch, l, cc, &b)) {
      *lev_out = l;
      return b;
    }
    for (int d = 0; d < 3; ++d) cc[d] >>= 1;
  }
  return -1;
}

static void balance(Problem& pr) {
  for (;;) {
    rebuild_index(pr);
    std::vector<size_t> todo;
    for (size_t i = 0; i < pr.leaves.size(); ++i) {
      const Leaf& L = pr.leaves[i];
      int64_t n[3];
      pdims(pr, L.patch, L.level, n);
      bool need = false;
      for (int d = 0; d < pr.D && !need; ++d)
        for (int s = -1; s <= 1 && !need; s += 2) {
          int64_t c[3] = {L.c[0], L.c[1], L.c[2]};
          c[d] += s;
          if (c[d] < 0 || c[d] >= n[d]) {
            if (pr.geometry == 0 && pr.periodic[d]) c[d] = (c[d] + n[d]) % n[d];
            else continue;
          }
          int lev;
          if (covering_leaf(pr, L.patch, L.level, c, &lev) >= 0) continue;
          // the neighbour cell is refined: its children next to our face must be leaves
          for (int ch = 0; ch < (1 << (pr.D - 1)) && !need; ++ch) {
            int64_t cc[3];
            int t = 0;
            for (int e = 0; e < 3; ++e) {
              if (e >= pr.D) { cc[e] = 0; continue; }
              if (e == d) cc[e] = 2 * c[e] + (s > 0 ? 0 : 1);
              else cc[e] = 2 * c[e] + ((ch >> (t++)) & 1);
            }
            int64_t b;
            if (!is_leaf(pr, L.patch, L.level + 1, cc, &b)) need = true;
          }
        }
      if (need) todo.push_back(i);
    }
    if (todo.empty()) break;
    std::sort(todo.rbegin(), todo.rend());
    for (size_t i : todo) refine_leaf(pr.leaves, i, pr.D);
  }
}

static void refine_regions(Problem& pr, int nreg, const double* regions, int patch_only) {
  for (int k = 0; k < nreg; ++k) {
    const double* R = regions + (size_t)k * pr.D * 2;
    std::vector<size_t> targets;
    for (size_t i = 0; i < pr.leaves.size(); ++i) {
      const Leaf& L = pr.leaves[i];
      if (L.level != k) continue;
      if (patch_only >= 0 && L.patch != patch_only) continue;
      double lo[3], hi[3];
      leaf_extent(pr, L, lo, hi);
      bool in = true;
      for (int d = 0; d < pr.D; ++d) {
        double c = (lo[d] + hi[d]) / 2;
        if (!(R[2 * d] < c && c < R[2 * d + 1])) in = false;
      }
      if (in) targets.push_back(i);
    }
    std::sort(targets.rbegin(), targets.rend());
    for (size_t i : targets) refine_leaf(pr.leaves, i, pr.D);
    balance(pr);
  }
}

// Physical / logical coordinates of owned node I (0-based) of block b.
static void node_logical(const Problem& pr, int64_t b, const int64_t I[3], double x[3]) {
  const Leaf& L = pr.leaves[b];
  double h = pr.h[b];
  for (int d = 0; d < pr.D; ++d) {
    int64_t q = I[d] / pr.P1, i = I[d] % pr.P1;
    double B = (double)(L.c[d] * pr.N);
    x[d] = pr.plo[L.patch][d] + h * (B + (double)(q * pr.P1) + pr.el.offsets[i]);
  }
}

static void node_physical(const Problem& pr, int64_t b, const int64_t I[3], double x[3]) {
  double xl[3] = {0, 0, 0};
  node_logical(pr, b, I, xl);
  if (pr.geometry == 0) {
    for (int d = 0; d < 3; ++d) x[d] = xl[d];
    return;
  }
  if (pr.geometry == 2) {  // annulus_radius, annulus_angle (geometry.jl:211-212)
    const double r = pr.an_r0 + pr.an_dr * xl[1];
    const double th = pr.leaves[b].patch * (M_PI / 2) + pr.an_dth * xl[0];
    x[0] = r * std::cos(th);
    x[1] = r * std::sin(th);
    x[2] = 0;
    return;
  }
  ball_position<double>(pr.ball, pr.leaves[b].patch, xl, x);
}

// ---------------------------------------------------------------------------
// Face resolution.
// ---------------------------------------------------------------------------
static inline int64_t lin(const Problem& pr, const int64_t I[3]) {
  return I[0] + pr.N * (I[1] + pr.N * I[2]);
}

// Tangential dimensions of a face normal to d, in increasing order.
static int tangential(int D, int d, int t[2]) {
  int k = 0;
  for (int e = 0; e < D; ++e)
    if (e != d) t[k++] = e;
  return k;
}

static void resolve_glue(Problem& pr, int64_t b, int d, int s, Face& F);

static void resolve_faces(Problem& pr) {
  const int D = pr.D;
  pr.faces.assign(pr.nb * 2 * D, Face());
  for (int64_t b = 0; b < pr.nb; ++b) {
    const Leaf& L = pr.leaves[b];
    int64_t n[3];
    pdims(pr, L.patch, L.level, n);
    for (int d = 0; d < D; ++d)
      for (int s = 0; s < 2; ++s) {
        Face& F = pr.faces[(b * D + d) * 2 + s];
        int64_t c[3] = {L.c[0], L.c[1], L.c[2]};
        c[d] += s ? 1 : -1;
        int np = L.patch;
        if (pr.geometry == 2 && (c[d] < 0 || c[d] >= n[d])) {
          // the annulus: the angular faces glued by the identity to the next
          // (previous) patch, the radial faces reflecting walls
          if (d == 1) {
            F.kind = G_MIRROR;
            continue;
          }
          np = (L.patch + (s ? 1 : 3)) % 4;
          c[d] = s ? 0 : n[d] - 1;
        } else if (c[d] < 0 || c[d] >= n[d]) {
          if (pr.geometry == 0) {
            if (pr.periodic[d]) c[d] = (c[d] + n[d]) % n[d];
            else {
              F.kind = G_MIRROR;
              continue;
            }
          } else {
            resolve_glue(pr, b, d, s, F);
            continue;
          }
        }
        int lev;
        int64_t nb = covering_leaf(pr, np, L.level, c, &lev);
        if (nb >= 0 && lev == L.level) {
          F.kind = G_SAME;
          F.nb = nb;
        } else if (nb >= 0) {
          F.kind = G_COARSER;
          F.nb = nb;
          for (int e = 0; e < 3; ++e) F.off[e] = c[e] & 1;
        } else {
          F.kind = G_FINER;
          int t[2];
          int nt = tangential(D, d, t);
          for (int ch = 0; ch < (1 << nt); ++ch) {
            int64_t cc[3] = {0, 0, 0};
            cc[d] = 2 * c[d] + (s ? 0 : 1);
            for (int k = 0; k < nt; ++k) cc[t[k]] = 2 * c[t[k]] + ((ch >> k) & 1);
            int64_t fb;
            if (!is_leaf(pr, L.patch, L.level + 1, cc, &fb)) {
              fprintf(stderr, "oracle: forest is not 2:1 face balanced\n");
              abort();
            }
            F.fine[ch] = fb;
          }
        }
      }
  }
}

// Glued and outer faces of the ball.  The ghost of an element set sits at the
// face node itself, so the source is the neighbour patch's owned node at the
// same physical point, in the element whose face centre is at the same
// physical point (the layout duplicates element face nodes).
struct PtKey {
  int64_t q[6];
  bool operator==(const PtKey& o) const { return std::memcmp(q, o.q, sizeof q) == 0; }
};
struct PtHash {
  size_t operator()(const PtKey& k) const {
    uint64_t h = 1469598103934665603ull;
    for (int i = 0; i < 6; ++i) h = (h ^ (uint64_t)k.q[i]) * 1099511628211ull;
    return (size_t)h;
  }
};

static std::unordered_map<PtKey, std::vector<std::pair<int64_t, int64_t>>, PtHash>* g_faceindex = nullptr;

static void face_points(const Problem& pr, int64_t b, int d, int s, const int64_t I[3], double x[3],
                        double c[3]) {
  node_physical(pr, b, I, x);
  // element face centre: tangential element centre, normal at the face
  const Leaf& L = pr.leaves[b];
  double xl[3] = {0, 0, 0};
  double h = pr.h[b];
  for (int e = 0; e < 3; ++e) {
    if (e == d) {
      xl[e] = pr.plo[L.patch][e] + h * (double)(L.c[e] * pr.N + (s ? pr.N : 0));
    } else {
      int64_t q = I[e] / pr.P1;
      xl[e] = pr.plo[L.patch][e] + h * ((double)(L.c[e] * pr.N + q * pr.P1) + 0.5 * pr.P1);
    }
  }
  ball_position<double>(pr.ball, L.patch, xl, c);
}

static PtKey ptkey(const double x[3], const double c[3], double scale) {
  PtKey k;
  for (int i = 0; i < 3; ++i) {
    k.q[i] = (int64_t)std::llround(x[i] * scale);
    k.q[3 + i] = (int64_t)std::llround(c[i] * scale);
  }
  return k;
}

static void build_face_index(Problem& pr) {
  delete g_faceindex;
  g_faceindex = new std::unordered_map<PtKey, std::vector<std::pair<int64_t, int64_t>>, PtHash>();
  const double scale = 1e7;
  for (int64_t b = 0; b < pr.nb; ++b) {
    const Leaf& L = pr.leaves[b];
    int64_t n[3];
    pdims(pr, L.patch, L.level, n);
    for (int d = 0; d < 3; ++d)
      for (int s = 0; s < 2; ++s) {
        int64_t c = L.c[d] + (s ? 1 : -1);
        if (c >= 0 && c < n[d]) continue;
        int t[2];
        tangential(3, d, t);
        for (int64_t j = 0; j < pr.N; ++j)
          for (int64_t i = 0; i < pr.N; ++i) {
            int64_t I[3];
            I[d] = s ? pr.N - 1 : 0;
            I[t[0]] = i;
            I[t[1]] = j;
            double x[3], cc[3];
            face_points(pr, b, d, s, I, x, cc);
            (*g_faceindex)[ptkey(x, cc, scale)].push_back({b, lin(pr, I)});
          }
      }
  }
}

static void resolve_glue(Problem& pr, int64_t b, int d, int s, Face& F) {
  const Leaf& L = pr.leaves[b];
  // the outer sphere: a shell's face away from the cube along its radial dim
  const int first = pr.ball.thirteen ? 7 : 1;  // the shells that reach the sphere
  if (L.patch >= first) {
    int k = (L.patch - first) / 2, sg = ((L.patch - first) % 2 == 0) ? -1 : 1;
    if (d == k && ((sg == 1 && s == 1) || (sg == -1 && s == 0))) {
      F.kind = pr.radiative ? G_RADIATIVE : G_DIRICHLET;
      return;
    }
  }
  F.kind = G_GLUE;
  F.glue_src.assign(pr.N * pr.N, -1);
  const double scale = 1e7;
  int t[2];
  tangential(3, d, t);
  for (int64_t j = 0; j < pr.N; ++j)
    for (int64_t i = 0; i < pr.N; ++i) {
      int64_t I[3];
      I[d] = s ? pr.N - 1 : 0;
      I[t[0]] = i;
      I[t[1]] = j;
      double x[3], cc[3];
      face_points(pr, b, d, s, I, x, cc);
      // the quantized key of a coordinate that lies near a cell boundary may
      // round either way on the two sides: try both roundings there
      int64_t base[6], alt[6];
      int amb[6], na = 0;
      for (int q = 0; q < 6; ++q) {
        double v = (q < 3 ? x[q] : cc[q - 3]) * scale;
        base[q] = (int64_t)std::llround(v);
        double fr = v - std::floor(v);
        alt[q] = (fr < 0.5) ? base[q] + 1 : base[q] - 1;
        if (std::fabs(fr - 0.5) < 1e-3) amb[na++] = q;
      }
      int64_t src = -1, srcb = -1;
      for (int a = 0; a < (1 << na) && src < 0; ++a) {
        PtKey k2;
        for (int q = 0; q < 6; ++q) k2.q[q] = base[q];
        for (int r = 0; r < na; ++r)
          if ((a >> r) & 1) k2.q[amb[r]] = alt[amb[r]];
        auto it = g_faceindex->find(k2);
        if (it == g_faceindex->end()) continue;
        for (auto& pr2 : it->second)
          if (pr.leaves[pr2.first].patch != L.patch) {
            src = pr2.second;
            srcb = pr2.first;
          }
      }
      if (src < 0) {
        fprintf(stderr, "oracle: no glued source for block %lld face %d%d\n", (long long)b, d, s);
        abort();
      }
      if (F.nb < 0) F.nb = srcb;
      if (F.nb != srcb) {
        fprintf(stderr, "oracle: glued face spans two blocks\n");
        abort();
      }
      F.glue_src[i + pr.N * j] = src;
    }
}

// ---------------------------------------------------------------------------
// Ghost fill.  `src(b, I)` reads an owned value; `dst` writes the ghost at
// tangential position (i, j) of face (d, s).  par = parity under a mirror.
// ---------------------------------------------------------------------------
template <class Src, class Dst>
static void fill_face(const Problem& pr, int64_t b, int d, int s, double par, Src src, Dst dst,
                      bool dirichlet_ok) {
  const int D = pr.D, N = pr.N, P1 = pr.P1;
  const Face& F = pr.faces[(b * D + d) * 2 + s];
  int t[2] = {0, 0};
  int nt = tangential(D, d, t);
  int64_t nj = nt >= 2 ? N : 1, ni = nt >= 1 ? N : 1;
  for (int64_t j = 0; j < nj; ++j)
    for (int64_t i = 0; i < ni; ++i) {
      int64_t T[2] = {i, j};
      double val = 0;
      switch (F.kind) {
        case G_SAME: {
          int64_t I[3] = {0, 0, 0};
          I[d] = s ? 0 : N - 1;
          for (int k = 0; k < nt; ++k) I[t[k]] = T[k];
          val = src(F.nb, I);
        } break;
        case G_MIRROR: {
          int64_t I[3] = {0, 0, 0};
          I[d] = s ? N - 1 : 0;
          for (int k = 0; k < nt; ++k) I[t[k]] = T[k];
          val = par * src(b, I);
        } break;
        case G_RADIATIVE: {  // the face node's own value (radiative_u_kernel!)
          int64_t I[3] = {0, 0, 0};
          I[d] = s ? N - 1 : 0;
          for (int k = 0; k < nt; ++k) I[t[k]] = T[k];
          val = src(b, I);
        } break;
        case G_GLUE: {
          int64_t l = F.glue_src[i + N * j];
          int64_t I[3] = {l % N, (l / N) % N, l / ((int64_t)N * N)};
          val = src(F.nb, I);
        } break;
        case G_COARSER: {
          // fine ghost: the coarse element's polynomial at our fine nodes
          int64_t qc[2];
          double xr[2];
          for (int k = 0; k < nt; ++k) {
            int64_t qf_ = T[k] / P1, jj = T[k] % P1;
            int f = (int)(qf_ % 2);
            qc[k] = F.off[t[k]] * (pr.E / 2) + qf_ / 2;
            xr[k] = (pr.el.nodes[jj] + 2 * f - 1) / 2;
          }
          int64_t cnt = ipow(P1, nt);
          for (int64_t a = 0; a < cnt; ++a) {
            int64_t I[3] = {0, 0, 0};
            I[d] = s ? 0 : N - 1;
            double wgt = 1;
            int64_t r = a;
            for (int k = 0; k < nt; ++k) {
              int64_t ic = r % P1;
              r /= P1;
              I[t[k]] = qc[k] * P1 + ic;
              wgt *= lagrange(pr.el, (int)ic, xr[k]);
            }
            val += wgt * src(F.nb, I);
          }
        } break;
        case G_FINER: {
          // coarse ghost: R = M_c^-1 I^T M_f over the fine nodes of our element
          int64_t qc[2], ic[2], o[2];
          for (int k = 0; k < nt; ++k) {
            qc[k] = T[k] / P1;
            ic[k] = T[k] % P1;
            o[k] = (qc[k] * P1 >= N / 2) ? 1 : 0;
          }
          int ch = 0;
          for (int k = 0; k < nt; ++k) ch |= (int)o[k] << k;
          int64_t fb = F.fine[ch];
          int64_t cnt = ipow(2 * P1, nt);
          for (int64_t a = 0; a < cnt; ++a) {
            int64_t I[3] = {0, 0, 0};
            I[d] = s ? 0 : N - 1;
            double wgt = 1;
            int64_t r = a;
            for (int k = 0; k < nt; ++k) {
              int64_t fj = r % (2 * P1);
              r /= 2 * P1;
              int f = (int)(fj / P1);
              int jj = (int)(fj % P1);
              int64_t qfine = 2 * (qc[k] - o[k] * (pr.E / 2)) + f;
              I[t[k]] = qfine * P1 + jj;
              double xr = (pr.el.nodes[jj] + 2 * f - 1) / 2;
              wgt *= lagrange(pr.el, (int)ic[k], xr) * pr.el.weights[jj] / (2 * pr.el.weights[ic[k]]);
            }
            val += wgt * src(fb, I);
          }
        } break;
        case G_DIRICHLET:
          if (!dirichlet_ok) abort();
          val = 0;  // written by the caller
          break;
        default:
          abort();
      }
      dst(i, j, val);
    }
}

// ---------------------------------------------------------------------------
// The RHS (evolution.jl:1318-1371).
// ---------------------------------------------------------------------------
static inline int64_t hidx(const Problem& pr, const int64_t I[3], const int64_t n[3]) {
  return I[0] + n[0] * (I[1] + n[1] * I[2]);
}

static void plane_wave(const Problem& pr, const double x[3], double t, double out[4]) {
  // plane_phase: sum(k .* x) - w t + phi (sevenpatch.jl:29-30)
  double sx = 0;
  for (int d = 0; d < 3; ++d) sx += pr.kvec[d] * x[d];
  double th = sx - pr.omega * t + pr.phase;
  double sn = std::sin(th), cs = std::cos(th);
  out[0] = sn;
  for (int d = 0; d < 3; ++d) out[1 + d] = pr.kvec[d] * cs;
}

static void scatter_exchange_u(Problem& pr, const double* y, double t) {
  const int D = pr.D, N = pr.N;
  const int64_t nn = pr.nn;
  int64_t n[3] = {1, 1, 1};
  for (int d = 0; d < D; ++d) n[d] = N + 2;
  const int64_t hn = pr.ufs_n;
  double* ufs = pr.ufs.data();
#pragma omp parallel for schedule(static)
  for (int64_t b = 0; b < pr.nb; ++b) {
    for (int64_t l = 0; l < nn; ++l) {
      int64_t I[3] = {l % N, (l / N) % N, l / ((int64_t)N * N)};
      int64_t S[3] = {0, 0, 0};
      for (int d = 0; d < D; ++d) S[d] = I[d] + 1;
      ufs[b * hn + hidx(pr, S, n)] = y[b * 2 * nn + l];
    }
  }
  auto src = [&](int64_t bb, const int64_t I[3]) { return y[bb * 2 * nn + lin(pr, I)]; };
#pragma omp parallel for schedule(static)
  for (int64_t b = 0; b < pr.nb; ++b)
    for (int d = 0; d < D; ++d)
      for (int s = 0; s < 2; ++s) {
        int tt[2];
        int nt = tangential(D, d, tt);
        const Face& F = pr.faces[(b * D + d) * 2 + s];
        auto dst = [&](int64_t i, int64_t j, double v) {
          int64_t S[3] = {0, 0, 0};
          S[d] = s ? N + 1 : 0;
          int64_t T[2] = {i, j};
          for (int k = 0; k < nt; ++k) S[tt[k]] = T[k] + 1;
          if (F.kind == G_DIRICHLET) {
            int64_t I[3] = {0, 0, 0};
            I[d] = s ? N - 1 : 0;
            for (int k = 0; k < nt; ++k) I[tt[k]] = T[k];
            double x[3], dv[4];
            node_physical(pr, b, I, x);
            plane_wave(pr, x, t, dv);
            v = dv[0];
          }
          ufs[b * hn + hidx(pr, S, n)] = v;
        };
        fill_face(pr, b, d, s, 1.0, src, dst, true);
      }
}

static void gradient_flat(Problem& pr) {
  const int D = pr.D, N = pr.N, P1 = pr.P1, PW = pr.P1 + 2;
  const int64_t nn = pr.nn;
  int64_t n[3] = {1, 1, 1};
  for (int e = 0; e < D; ++e) n[e] = N + 2;
  const double* ufs = pr.ufs.data();
#pragma omp parallel for schedule(static)
  for (int64_t b = 0; b < pr.nb; ++b) {
    const double h = pr.h[b];
    for (int64_t l = 0; l < nn; ++l) {
      int64_t I[3] = {l % N, (l / N) % N, l / ((int64_t)N * N)};
      for (int d = 0; d < D; ++d) {
        int64_t C[3] = {0, 0, 0};
        for (int e = 0; e < D; ++e) C[e] = I[e] + 1;
        int i = (int)(I[d] % P1);
        int64_t start = C[d] - i - 1;
        double acc = 0;
        for (int k = 0; k < PW; ++k) {
          int64_t W[3] = {C[0], C[1], C[2]};
          W[d] = start + k;
          acc += pr.el.W[i][k] * ufs[b * pr.ufs_n + hidx(pr, W, n)];
        }
        int64_t gn[3] = {1, 1, 1};
        for (int e = 0; e < D; ++e) gn[e] = N + (e == d ? 2 : 0);
        int64_t G[3] = {I[0], I[1], I[2]};
        G[d] += 1;
        double g = acc / h;
        if (pr.geometry == 2) g = pr.dcoef[(b * (1 + D) + 1 + d) * nn + l] * g;  // flux A_d x
        pr.grads[d][b * pr.grad_n[d] + hidx(pr, G, gn)] = g;
      }
    }
  }
}

static void exchange_grads(Problem& pr) {
  const int D = pr.D, N = pr.N;
#pragma omp parallel for schedule(static)
  for (int64_t b = 0; b < pr.nb; ++b)
    for (int d = 0; d < D; ++d) {
      int64_t gn[3] = {1, 1, 1};
      for (int e = 0; e < D; ++e) gn[e] = N + (e == d ? 2 : 0);
      double* g = pr.grads[d].data();
      const int64_t gsz = pr.grad_n[d];
      auto src = [&](int64_t bb, const int64_t I[3]) {
        int64_t G[3] = {I[0], I[1], I[2]};
        G[d] += 1;
        return g[bb * gsz + hidx(pr, G, gn)];
      };
      int tt[2];
      int nt = tangential(D, d, tt);
      for (int s = 0; s < 2; ++s) {
        auto dst = [&](int64_t i, int64_t j, double v) {
          int64_t G[3] = {0, 0, 0};
          G[d] = s ? N + 1 : 0;
          int64_t T[2] = {i, j};
          for (int k = 0; k < nt; ++k) G[tt[k]] = T[k];
          g[b * gsz + hidx(pr, G, gn)] = v;
        };
        fill_face(pr, b, d, s, -1.0, src, dst, false);
      }
    }
}

static void laplacian_flat(Problem& pr, const double* y, double* du) {
  const int D = pr.D, N = pr.N, P1 = pr.P1, PW = pr.P1 + 2;
  const int64_t nn = pr.nn;
#pragma omp parallel for schedule(static)
  for (int64_t b = 0; b < pr.nb; ++b) {
    const double h = pr.h[b];
    for (int64_t l = 0; l < nn; ++l) {
      int64_t I[3] = {l % N, (l / N) % N, l / ((int64_t)N * N)};
      double acc = 0;
      for (int d = 0; d < D; ++d) {
        int64_t gn[3] = {1, 1, 1};
        for (int e = 0; e < D; ++e) gn[e] = N + (e == d ? 2 : 0);
        int i = (int)(I[d] % P1);
        int64_t cd = I[d] + 1;
        int64_t start = cd - i - 1;
        double term = 0;
        for (int k = 0; k < PW; ++k) {
          int64_t G[3] = {I[0], I[1], I[2]};
          G[d] = start + k;
          term += pr.el.W[i][k] * pr.grads[d][b * pr.grad_n[d] + hidx(pr, G, gn)];
        }
        acc += term;
      }
      du[b * 2 * nn + l] = y[b * 2 * nn + nn + l];
      double lap = acc / h;
      if (pr.geometry == 2) lap = lap / pr.dcoef[(b * (1 + D)) * nn + l];  // divergence / sqrt(g)
      du[b * 2 * nn + nn + l] = lap;
    }
  }
}

// --- full metric: Cartesian gradient (evolution.jl:136-220) ---------------
static inline double coefv(const Problem& pr, int64_t b, int var, int64_t l) {
  return pr.coef[(b * (1 + pr.D * pr.D) + var) * pr.nn + l];
}

static void gradient_cart(Problem& pr) {
  const int D = pr.D, N = pr.N, P1 = pr.P1, PW = pr.P1 + 2;
  const int64_t nn = pr.nn;
  int64_t n[3] = {1, 1, 1};
  for (int e = 0; e < D; ++e) n[e] = N + 2;
  const int64_t hn = pr.ufs_n;
#pragma omp parallel for schedule(static)
  for (int64_t b = 0; b < pr.nb; ++b) {
    const double h = pr.h[b];
    for (int64_t l = 0; l < nn; ++l) {
      int64_t I[3] = {l % N, (l / N) % N, l / ((int64_t)N * N)};
      int64_t C[3] = {0, 0, 0};
      for (int e = 0; e < D; ++e) C[e] = I[e] + 1;
      double Du[3];
      for (int e = 0; e < D; ++e) {
        int i = (int)(I[e] % P1);
        int64_t start = C[e] - i - 1;
        double acc = 0;
        for (int k = 0; k < PW; ++k) {
          int64_t W[3] = {C[0], C[1], C[2]};
          W[e] = start + k;
          acc += pr.el.W[i][k] * pr.ufs[b * hn + hidx(pr, W, n)];
        }
        Du[e] = acc / h;
      }
      for (int i = 0; i < D; ++i) {
        double v;
        if (pr.flat[b]) {
          double J = coefv(pr, b, 0, 0);
          v = coefv(pr, b, 1 + i * D + i, 0) * Du[i] / J;
        } else {
          double s = 0;
          for (int e = 0; e < D; ++e) {
            double term = coefv(pr, b, 1 + e * D + i, l) * Du[e];
            s = (e == 0) ? term : s + term;
          }
          v = s / coefv(pr, b, 0, l);
        }
        pr.cart[(b * D + i) * hn + hidx(pr, C, n)] = v;
      }
    }
  }
}

static void exchange_cart(Problem& pr, double t) {
  const int D = pr.D, N = pr.N;
  int64_t n[3] = {1, 1, 1};
  for (int e = 0; e < D; ++e) n[e] = N + 2;
  const int64_t hn = pr.ufs_n;
#pragma omp parallel for schedule(static)
  for (int64_t b = 0; b < pr.nb; ++b)
    for (int comp = 0; comp < D; ++comp) {
      double* g = pr.cart.data();
      auto src = [&](int64_t bb, const int64_t I[3]) {
        int64_t S[3] = {0, 0, 0};
        for (int e = 0; e < D; ++e) S[e] = I[e] + 1;
        return g[(bb * D + comp) * hn + hidx(pr, S, n)];
      };
      for (int d = 0; d < D; ++d) {
        int tt[2];
        int nt = tangential(D, d, tt);
        for (int s = 0; s < 2; ++s) {
          const Face& F = pr.faces[(b * D + d) * 2 + s];
          auto dst = [&](int64_t i, int64_t j, double v) {
            int64_t S[3] = {0, 0, 0};
            S[d] = s ? N + 1 : 0;
            int64_t T[2] = {i, j};
            for (int k = 0; k < nt; ++k) S[tt[k]] = T[k] + 1;
            if (F.kind == G_DIRICHLET) {
              int64_t I[3] = {0, 0, 0};
              I[d] = s ? N - 1 : 0;
              for (int k = 0; k < nt; ++k) I[tt[k]] = T[k];
              double x[3], dv[4];
              node_physical(pr, b, I, x);
              plane_wave(pr, x, t, dv);
              v = dv[1 + comp];
            } else if (F.kind == G_RADIATIVE) {
              // V_g = V + c S, S . V_g = 2 F* - S . V, F* = -sigma |S| (kappa v + u/R)
              int64_t I[3] = {0, 0, 0};
              I[d] = s ? N - 1 : 0;
              for (int k = 0; k < nt; ++k) I[tt[k]] = T[k];
              const int64_t lf = lin(pr, I);
              int64_t Sf[3] = {0, 0, 0};
              for (int e = 0; e < D; ++e) Sf[e] = I[e] + 1;
              double Sv[3], V[3], S2 = 0, Fl = 0;
              for (int q = 0; q < D; ++q) {
                Sv[q] = coefv(pr, b, 1 + d * D + q, lf);
                V[q] = g[(b * D + q) * hn + hidx(pr, Sf, n)];
                S2 += Sv[q] * Sv[q];
                Fl += Sv[q] * V[q];
              }
              const double uu = pr.ycur[b * 2 * pr.nn + lf], vv = pr.ycur[b * 2 * pr.nn + pr.nn + lf];
              const double Fs = -(s ? 1.0 : -1.0) * std::sqrt(S2) * (pr.rad_kappa * vv + pr.rad_invR * uu);
              v = V[comp] + 2 * (Fs - Fl) / S2 * Sv[comp];
            }
            g[(b * D + comp) * hn + hidx(pr, S, n)] = v;
          };
          fill_face(pr, b, d, s, 1.0, src, dst, true);
        }
      }
    }
}

static void laplacian_cart(Problem& pr, const double* y, double* du) {
  const int D = pr.D, N = pr.N, P1 = pr.P1, PW = pr.P1 + 2;
  const int64_t nn = pr.nn;
  int64_t n[3] = {1, 1, 1};
  for (int e = 0; e < D; ++e) n[e] = N + 2;
  const int64_t hn = pr.ufs_n;
#pragma omp parallel for schedule(static)
  for (int64_t b = 0; b < pr.nb; ++b) {
    const double h = pr.h[b];
    for (int64_t l = 0; l < nn; ++l) {
      int64_t I[3] = {l % N, (l / N) % N, l / ((int64_t)N * N)};
      double acc = 0;
      for (int d = 0; d < D; ++d) {
        int i = (int)(I[d] % P1);
        int64_t start = I[d] - i - 1;  // owned index of window point k = 0
        double term = 0;
        for (int k = 0; k < PW; ++k) {
          int64_t o = start + k;  // owned index along d, -1 .. N
          int64_t St[3] = {0, 0, 0};
          for (int e = 0; e < D; ++e) St[e] = I[e] + 1;
          St[d] = o + 1;
          double F;
          if (pr.flat[b]) {
            F = coefv(pr, b, 1 + d * D + d, 0) * pr.cart[(b * D + d) * hn + hidx(pr, St, n)];
          } else {
            int64_t Os[3] = {I[0], I[1], I[2]};
            Os[d] = std::min<int64_t>(std::max<int64_t>(o, 0), N - 1);
            int64_t lo = lin(pr, Os);
            double s = 0;
            for (int j = 0; j < D; ++j) {
              double tj = coefv(pr, b, 1 + d * D + j, lo) * pr.cart[(b * D + j) * hn + hidx(pr, St, n)];
              s = (j == 0) ? tj : s + tj;
            }
            F = s;
          }
          term += pr.el.W[i][k] * F;
        }
        acc += term;
      }
      du[b * 2 * nn + l] = y[b * 2 * nn + nn + l];
      du[b * 2 * nn + nn + l] = acc / h / coefv(pr, b, 0, l);
    }
  }
}

static void fill_gradient(Problem& pr, const double* y, double t) {
  pr.ycur = y;
  scatter_exchange_u(pr, y, t);
  if (pr.geometry != 1) {
    gradient_flat(pr);
    exchange_grads(pr);
  } else {
    gradient_cart(pr);
    exchange_cart(pr, t);
  }
}

static void rhs(Problem& pr, const double* y, double* du, double t) {
  fill_gradient(pr, y, t);
  if (pr.geometry != 1) laplacian_flat(pr, y, du);
  else laplacian_cart(pr, y, du);
}

// --- truncation (evolution.jl:222-425) --------------------------------------
static void truncate_tensor_elem(const Problem& pr, double* a, const int64_t base[3]) {
  const int D = pr.D, P1 = pr.P1, N = pr.N;
  for (int d = 0; d < D; ++d) {
    int64_t K1 = (D >= 2) ? P1 : 1, K2 = (D >= 3) ? P1 : 1;
    int t[2];
    int nt = tangential(D, d, t);
    for (int64_t k2 = 0; k2 < (nt >= 2 ? P1 : 1); ++k2)
      for (int64_t k1 = 0; k1 < (nt >= 1 ? P1 : 1); ++k1) {
        int64_t K[2] = {k1, k2};
        double c = 0;
        for (int j = 0; j < P1; ++j) {
          int64_t I[3] = {base[0], base[1], base[2]};
          I[d] += j;
          for (int q = 0; q < nt; ++q) I[t[q]] += K[q];
          c += pr.el.dual[j] * a[I[0] + N * (I[1] + N * I[2])];
        }
        for (int i = 0; i < P1; ++i) {
          int64_t I[3] = {base[0], base[1], base[2]};
          I[d] += i;
          for (int q = 0; q < nt; ++q) I[t[q]] += K[q];
          a[I[0] + N * (I[1] + N * I[2])] -= pr.el.mode[i] * c;
        }
      }
    (void)K1;
    (void)K2;
  }
}

// The tensor projector with each line's own rank-one pair (MetricTruncation,
// evolution.jl:556-575): line along d of element J of block b.
static void truncate_metric_elem(const Problem& pr, double* a, const int64_t base[3], int64_t b,
                                 const int64_t J[3]) {
  const int D = pr.D, P1 = pr.P1, N = pr.N, E = pr.E;
  for (int d = 0; d < D; ++d) {
    const double* md = &pr.lmode[((b * D + d) * E + J[d]) * P1];
    const double* dd = &pr.ldual[((b * D + d) * E + J[d]) * P1];
    int t[2];
    int nt = tangential(D, d, t);
    for (int64_t k2 = 0; k2 < (nt >= 2 ? P1 : 1); ++k2)
      for (int64_t k1 = 0; k1 < (nt >= 1 ? P1 : 1); ++k1) {
        int64_t K[2] = {k1, k2};
        double c = 0;
        for (int j = 0; j < P1; ++j) {
          int64_t I[3] = {base[0], base[1], base[2]};
          I[d] += j;
          for (int q = 0; q < nt; ++q) I[t[q]] += K[q];
          c += dd[j] * a[I[0] + N * (I[1] + N * I[2])];
        }
        for (int i = 0; i < P1; ++i) {
          int64_t I[3] = {base[0], base[1], base[2]};
          I[d] += i;
          for (int q = 0; q < nt; ++q) I[t[q]] += K[q];
          a[I[0] + N * (I[1] + N * I[2])] -= md[i] * c;
        }
      }
  }
}

static void truncate_element_chol(const Problem& pr, double* a, const int64_t base[3], int64_t s, int64_t e) {
  const int n = pr.tn, m = pr.tm, P1 = pr.P1, N = pr.N;
  const int64_t ne = ipow(pr.E, 3);
  double c[512];
  const double* w = &pr.tw[(s * ne + e) * n];
  const double* U = &pr.tchol[(s * ne + e) * (int64_t)m * m];
  auto at = [&](int node) {
    int64_t i0 = node % P1, i1 = (node / P1) % P1, i2 = node / (P1 * P1);
    return (base[0] + i0) + N * ((base[1] + i1) + N * (base[2] + i2));
  };
  for (int q = 0; q < m; ++q) {
    double acc = 0;
    for (int k = 0; k < n; ++k) acc += pr.tmodes[k * m + q] * w[k] * a[at(k)];
    c[q] = acc;
  }
  for (int q = 0; q < m; ++q) {  // U^T y = c  (U upper, column-major U[r + m q])
    double acc = c[q];
    for (int r = 0; r < q; ++r) acc -= U[r + (int64_t)m * q] * c[r];
    c[q] = acc / U[q + (int64_t)m * q];
  }
  for (int q = m - 1; q >= 0; --q) {  // U z = y
    double acc = c[q];
    for (int r = q + 1; r < m; ++r) acc -= U[q + (int64_t)m * r] * c[r];
    c[q] = acc / U[q + (int64_t)m * q];
  }
  for (int k = 0; k < n; ++k) {
    double z = 0;
    for (int q = 0; q < m; ++q) z += pr.tmodes[k * m + q] * c[q];
    a[at(k)] = z;
  }
}

static void truncate_state(Problem& pr, double* y) {
  if (!pr.truncate) return;
  const int D = pr.D, E = pr.E, P1 = pr.P1;
  const int64_t nn = pr.nn;
  const int64_t ne = ipow(E, D);
#pragma omp parallel for schedule(static)
  for (int64_t b = 0; b < pr.nb; ++b) {
    double* a = y + b * 2 * nn + nn;
    for (int64_t e = 0; e < ne; ++e) {
      int64_t J[3] = {e % E, (e / E) % (D >= 2 ? E : 1), 0};
      if (D == 3) J[2] = e / ((int64_t)E * E);
      int64_t base[3] = {0, 0, 0};
      for (int d = 0; d < D; ++d) base[d] = J[d] * P1;
      if (pr.geometry == 1 && pr.slot[b] >= 0) truncate_element_chol(pr, a, base, pr.slot[b], e);
      else if (pr.geometry == 2) truncate_metric_elem(pr, a, base, b, J);
      else truncate_tensor_elem(pr, a, base);
    }
  }
}

// --- energy and norms (evolution.jl:1472-1512, layout.jl:95-235) -----------
static inline double node_weight(const Problem& pr, const int64_t I[3], double h) {
  double w = 1;
  for (int d = 0; d < pr.D; ++d) w *= h * pr.P1 / 2 * pr.el.weights[I[d] % pr.P1];
  return w;
}

static double energy(Problem& pr, const double* y, double t) {
  fill_gradient(pr, y, t);
  const int D = pr.D, N = pr.N;
  const int64_t nn = pr.nn;
  std::vector<double> kin(pr.nb), pot(pr.nb * 3);
#pragma omp parallel for schedule(static)
  for (int64_t b = 0; b < pr.nb; ++b) {
    double k = 0, p[3] = {0, 0, 0};
    for (int64_t l = 0; l < nn; ++l) {
      int64_t I[3] = {l % N, (l / N) % N, l / ((int64_t)N * N)};
      double w = node_weight(pr, I, pr.h[b]);
      double v = y[b * 2 * nn + nn + l];
      if (pr.geometry != 1) {
        // flat: w v^2 + w g_d^2; diagonal metric: w sqrt(g) v^2 + w F_d^2 / A_d
        // (evolution.jl:1483-1487, layout.jl:135-161)
        const double* dc = pr.geometry == 2 ? &pr.dcoef[(b * (1 + D)) * nn] : nullptr;
        k += (dc ? w * dc[l] : w) * (v * v);
        for (int d = 0; d < D; ++d) {
          int64_t gn[3] = {1, 1, 1};
          for (int e = 0; e < D; ++e) gn[e] = N + (e == d ? 2 : 0);
          int64_t G[3] = {I[0], I[1], I[2]};
          G[d] += 1;
          double g = pr.grads[d][b * pr.grad_n[d] + hidx(pr, G, gn)];
          p[d] += (dc ? w * (1.0 / dc[(1 + d) * nn + l]) : w) * (g * g);
        }
      } else {
        double c = coefv(pr, b, 0, l);
        k += w * c * (v * v);
        int64_t n3[3] = {1, 1, 1};
        for (int e = 0; e < D; ++e) n3[e] = N + 2;
        int64_t S[3] = {0, 0, 0};
        for (int e = 0; e < D; ++e) S[e] = I[e] + 1;
        double s = 0;
        for (int i = 0; i < D; ++i) {
          double g = pr.cart[(b * D + i) * pr.ufs_n + hidx(pr, S, n3)];
          s += g * g;
        }
        p[0] += w * c * s;
      }
    }
    kin[b] = k;
    for (int d = 0; d < 3; ++d) pot[b * 3 + d] = p[d];
  }
  double K = 0, Pd[3] = {0, 0, 0};
  for (int64_t b = 0; b < pr.nb; ++b) {
    K += kin[b];
    for (int d = 0; d < 3; ++d) Pd[d] += pot[b * 3 + d];
  }
  double Pot = Pd[0] + Pd[1] + Pd[2];
  double bnd = 0;  // the radiative boundary's 1/2 sum w |S[d]| u^2 / R (evolution.jl:1276-1299)
  if (pr.radiative && pr.rad_invR != 0)
    for (int64_t b = 0; b < pr.nb; ++b)
      for (int d = 0; d < D; ++d)
        for (int s = 0; s < 2; ++s) {
          if (pr.faces[(b * D + d) * 2 + s].kind != G_RADIATIVE) continue;
          for (int64_t l = 0; l < nn; ++l) {
            int64_t I[3] = {l % N, (l / N) % N, l / ((int64_t)N * N)};
            if (I[d] != (s ? N - 1 : 0)) continue;
            double w = 1, S2 = 0;
            for (int e = 0; e < D; ++e)
              if (e != d) w *= pr.h[b] * pr.P1 / 2 * pr.el.weights[I[e] % pr.P1];
            for (int q = 0; q < D; ++q) S2 += coefv(pr, b, 1 + d * D + q, l) * coefv(pr, b, 1 + d * D + q, l);
            const double u = y[b * 2 * nn + l];
            bnd += w * std::sqrt(S2) * u * u;
          }
        }
  return (K + Pot) / 2 + bnd * pr.rad_invR / 2;
}

static void norms(const Problem& pr, const double* err, int vars_mask, int p, double* lp, double* linf) {
  const int N = pr.N;
  const int64_t nn = pr.nn;
  std::vector<double> tot(pr.nb), vol(pr.nb), mx(pr.nb);
#pragma omp parallel for schedule(static)
  for (int64_t b = 0; b < pr.nb; ++b) {
    double T = 0, V = 0, M = 0;
    for (int64_t l = 0; l < nn; ++l) {
      int64_t I[3] = {l % N, (l / N) % N, l / ((int64_t)N * N)};
      double w = node_weight(pr, I, pr.h[b]);
      if (pr.geometry == 1) w *= coefv(pr, b, 0, l);
      if (pr.geometry == 2) w *= pr.dcoef[(b * (1 + pr.D)) * nn + l];
      double s = 0;
      int nv = 0;
      for (int v = 0; v < 2; ++v)
        if (vars_mask & (1 << v)) {
          double x = std::fabs(err[b * 2 * nn + v * nn + l]);
          s += std::pow(x, p);
          M = std::max(M, x);
          ++nv;
        }
      T += w * s;
      V += w * nv;
    }
    tot[b] = T;
    vol[b] = V;
    mx[b] = M;
  }
  double T = 0, V = 0, M = 0;
  for (int64_t b = 0; b < pr.nb; ++b) {
    T += tot[b];
    V += vol[b];
    M = std::max(M, mx[b]);
  }
  *lp = V == 0 ? 0 : std::pow(T / V, 1.0 / p);
  *linf = M;
}

static void allocate_work(Problem& pr) {
  const int D = pr.D, N = pr.N;
  pr.nn = ipow(N, D);
  pr.ufs_n = ipow(N + 2, D);
  pr.ufs.assign(pr.nb * pr.ufs_n, 0.0);
  for (int d = 0; d < D; ++d) {
    pr.grad_n[d] = ipow(N, D - 1) * (N + 2);
    if (pr.geometry != 1) pr.grads[d].assign(pr.nb * pr.grad_n[d], 0.0);
  }
  if (pr.geometry == 1) pr.cart.assign(pr.nb * D * pr.ufs_n, 0.0);
}

// --- seven-patch metric and tables (evolution.jl:594-1018) -----------------
static void cholesky_upper(int m, std::vector<qf>& G, std::vector<qf>& U) {
  // G symmetric (m x m, column-major); U upper with U^T U = G
  U.assign((size_t)m * m, 0);
  for (int j = 0; j < m; ++j) {
    for (int i = 0; i <= j; ++i) {
      qf s = G[i + (size_t)m * j];
      for (int k = 0; k < i; ++k) s -= U[k + (size_t)m * i] * U[k + (size_t)m * j];
      if (i == j) U[i + (size_t)m * j] = sqrtq(s);
      else U[i + (size_t)m * j] = s / U[i + (size_t)m * i];
    }
  }
}

static void build_ball_metric(Problem& pr) {
  const int D = 3, N = pr.N, P1 = pr.P1, P = pr.P, E = pr.E;
  const int64_t nn = pr.nn;
  const int nv = 1 + D * D;
  pr.coef.assign(pr.nb * nv * nn, 0.0);
#pragma omp parallel for schedule(static)
  for (int64_t b = 0; b < pr.nb; ++b)
    for (int64_t l = 0; l < nn; ++l) {
      int64_t I[3] = {l % N, (l / N) % N, l / ((int64_t)N * N)};
      double xl[3];
      node_logical(pr, b, I, xl);
      double J, S[3][3];
      ball_frame(pr.ball, pr.leaves[b].patch, xl, J, S);
      if (pr.leaves[b].patch == 0) {
        // the cube is affine: J and S are constants; evaluate them exactly
        double a = 2 * pr.ball.Lc / pr.ball.n;
        J = a * a * a;
        for (int d = 0; d < 3; ++d)
          for (int i = 0; i < 3; ++i) S[d][i] = (d == i) ? a * a : 0.0;
      }
      pr.coef[(b * nv + 0) * nn + l] = J;
      for (int d = 0; d < 3; ++d)
        for (int i = 0; i < 3; ++i) pr.coef[(b * nv + 1 + d * 3 + i) * nn + l] = S[d][i];
    }
  pr.flat.assign(pr.nb, 0);
  for (int64_t b = 0; b < pr.nb; ++b) {
    bool f = true;
    for (int v = 0; v < nv && f; ++v) {
      double c0 = pr.coef[(b * nv + v) * nn];
      for (int64_t l = 0; l < nn && f; ++l)
        if (pr.coef[(b * nv + v) * nn + l] != c0) f = false;
      int dd = (v - 1) / 3, ii = (v - 1) % 3;
      if (v >= 1 && dd != ii && c0 != 0) f = false;
    }
    pr.flat[b] = f;
  }
  // time step: min over blocks of h_b / max speed, speed = sqrt(lambda_max(g^{de}))
  double hs = INFINITY;
  for (int64_t b = 0; b < pr.nb; ++b) {
    double sb = 0;
    for (int64_t l = 0; l < nn; ++l) {
      double J = pr.coef[(b * nv) * nn + l];
      double S[3][3];
      for (int d = 0; d < 3; ++d)
        for (int i = 0; i < 3; ++i) S[d][i] = pr.coef[(b * nv + 1 + d * 3 + i) * nn + l];
      double g[3][3];
      for (int a = 0; a < 3; ++a)
        for (int c = 0; c < 3; ++c) g[a][c] = (S[a][0] * S[c][0] + S[a][1] * S[c][1] + S[a][2] * S[c][2]) / (J * J);
      // largest eigenvalue of a symmetric 3x3 by power-free closed form
      double p1 = g[0][1] * g[0][1] + g[0][2] * g[0][2] + g[1][2] * g[1][2];
      double q = (g[0][0] + g[1][1] + g[2][2]) / 3;
      double p2 = (g[0][0] - q) * (g[0][0] - q) + (g[1][1] - q) * (g[1][1] - q) + (g[2][2] - q) * (g[2][2] - q) + 2 * p1;
      double pp = std::sqrt(p2 / 6);
      double lam;
      if (pp == 0) lam = q;
      else {
        double B[3][3];
        for (int a = 0; a < 3; ++a)
          for (int c = 0; c < 3; ++c) B[a][c] = (g[a][c] - (a == c ? q : 0)) / pp;
        double r = (B[0][0] * (B[1][1] * B[2][2] - B[1][2] * B[2][1]) - B[0][1] * (B[1][0] * B[2][2] - B[1][2] * B[2][0]) +
                    B[0][2] * (B[1][0] * B[2][1] - B[1][1] * B[2][0])) / 2;
        r = std::min(1.0, std::max(-1.0, r));
        double phi = std::acos(r) / 3;
        lam = q + 2 * pp * std::cos(phi);
      }
      sb = std::max(sb, std::sqrt(lam));
    }
    hs = std::min(hs, pr.h[b] / sb);
  }
  pr.hspeed = hs;
  // per-element truncation on the shells: Psi G^-1 Psi^T W, G = U^T U (evolution.jl:844-899)
  const int n = P1 * P1 * P1, m = P * P * P;
  pr.tn = n;
  pr.tm = m;
  pr.tmodes.assign((size_t)n * m, 0.0);
  qf xq[9], wq[9], Lq[9][9];
  {
    // binary128 nodes/weights again (element.jl's BigFloat analogue)
    qf x[9], L[9];
    for (int i = 0; i <= P; ++i) x[i] = -cosq(M_PIq * i / P);
    for (int it = 0; it < 200; ++it) {
      qf dmax = 0, dl[9];
      for (int i = 0; i <= P; ++i) {
        legendre_q(P, x[i], L);
        dl[i] = (x[i] * L[P] - L[P - 1]) / ((P + 1) * L[P]);
        dmax = fmaxq(dmax, fabsq(dl[i]));
      }
      for (int i = 0; i <= P; ++i) x[i] -= dl[i];
      if (dmax <= ldexpq(1, -106)) break;
    }
    for (int i = 0; i <= P; ++i) xq[i] = (x[i] - x[P - i]) / 2;
    xq[0] = -1;
    xq[P] = 1;
    for (int i = 0; i <= P; ++i) {
      legendre_q(P, xq[i], Lq[i]);
      wq[i] = 2 / (qf(P) * (P + 1) * Lq[i][P] * Lq[i][P]);
    }
  }
  qf offq[9];
  for (int i = 0; i <= P; ++i) offq[i] = (P + 1) * (1 + xq[i]) / 2;
  offq[0] = 0;
  offq[P] = P1;
  std::vector<qf> Psi((size_t)n * m);
  for (int k = 0; k < n; ++k)
    for (int q = 0; q < m; ++q) {
      int k0 = k % P1, k1 = (k / P1) % P1, k2 = k / (P1 * P1);
      int q0 = q % P, q1 = (q / P) % P, q2 = q / (P * P);
      Psi[(size_t)k * m + q] = Lq[k0][q0] * Lq[k1][q1] * Lq[k2][q2];
      pr.tmodes[(size_t)k * m + q] = (double)Psi[(size_t)k * m + q];
    }
  pr.slot.assign(pr.nb, -1);
  int64_t ns = 0;
  for (int64_t b = 0; b < pr.nb; ++b)
    if (pr.leaves[b].patch != 0) pr.slot[b] = ns++;
  const int64_t ne = (int64_t)E * E * E;
  pr.tw.assign((size_t)ns * ne * n, 0.0);
  pr.tchol.assign((size_t)ns * ne * m * m, 0.0);
#pragma omp parallel for schedule(dynamic)
  for (int64_t b = 0; b < pr.nb; ++b) {
    if (pr.slot[b] < 0) continue;
    const Leaf& L = pr.leaves[b];
    int64_t s = pr.slot[b];
    std::vector<qf> G((size_t)m * m), U, w(n);
    for (int64_t e = 0; e < ne; ++e) {
      int64_t J[3] = {e % E, (e / E) % E, e / ((int64_t)E * E)};
      for (int k = 0; k < n; ++k) {
        int kk[3] = {k % P1, (k / P1) % P1, k / (P1 * P1)};
        qf xi[3];
        for (int d = 0; d < 3; ++d)
          xi[d] = qf(pr.plo[L.patch][d]) + qf(pr.h[b]) * (qf((double)(L.c[d] * N)) + qf(J[d] * P1) + offq[kk[d]]);
        double vol = ball_volume_q(pr.ball, L.patch, xi);
        w[k] = wq[kk[0]] * wq[kk[1]] * wq[kk[2]] * qf(vol);
        pr.tw[(s * ne + e) * n + k] = (double)w[k];
      }
      for (int q = 0; q < m; ++q)
        for (int r = 0; r < m; ++r) {
          qf acc = 0;
          for (int k = 0; k < n; ++k) acc += Psi[(size_t)k * m + q] * w[k] * Psi[(size_t)k * m + r];
          G[q + (size_t)m * r] = acc;
        }
      cholesky_upper(m, G, U);
      for (size_t i = 0; i < (size_t)m * m; ++i) pr.tchol[(s * ne + e) * (size_t)m * m + i] = (double)U[i];
    }
  }
}

}  // namespace

// --- annulus: diagonal metric and per-line truncation (geometry.jl:176-268,
// evolution.jl:576-622, 698-708) ----------------------------------------------
// The degree-P polynomial at the element's nodes orthogonal in sum w_i a_i f_i g_i
// to every lower one: L_P less its projection onto L_0..L_{P-1}, by the normal
// equations solved in binary128.
static void projector_pair(const Problem& pr, const qf* a, double* mode, double* dual) {
  const int P = pr.P, P1 = pr.P1;
  qf L[9][9];  // L[i][k] = L_k(xi_i)
  for (int i = 0; i < P1; ++i) legendre_q(P, (qf)pr.el.nodes[i], L[i]);
  auto inner = [&](int k, int l) {
    qf s = 0;
    for (int i = 0; i < P1; ++i) s += (qf)pr.el.weights[i] * a[i] * L[i][k] * L[i][l];
    return s;
  };
  qf G[8][9];
  for (int k = 0; k < P; ++k) {
    for (int l = 0; l < P; ++l) G[k][l] = inner(k, l);
    G[k][P] = inner(P, k);
  }
  for (int c = 0; c < P; ++c)  // Gauss-Jordan, the Gram matrix is positive definite
    for (int r = 0; r < P; ++r) {
      if (r == c) continue;
      const qf f = G[r][c] / G[c][c];
      for (int q = c; q <= P; ++q) G[r][q] -= f * G[c][q];
    }
  qf phi[9], gam = 0;
  for (int i = 0; i < P1; ++i) {
    phi[i] = L[i][P];
    for (int k = 0; k < P; ++k) phi[i] -= (G[k][P] / G[k][k]) * L[i][k];
    gam += (qf)pr.el.weights[i] * a[i] * phi[i] * phi[i];
  }
  for (int i = 0; i < P1; ++i) {
    mode[i] = (double)phi[i];
    dual[i] = (double)((qf)pr.el.weights[i] * a[i] * phi[i] / gam);
  }
}

static void build_annulus_metric(Problem& pr) {
  const int D = 2, N = pr.N, P1 = pr.P1, E = pr.E;
  const int64_t nn = pr.nn;
  pr.dcoef.assign(pr.nb * (1 + D) * nn, 0.0);
  pr.lmode.assign(pr.nb * D * E * P1, 0.0);
  pr.ldual.assign(pr.nb * D * E * P1, 0.0);
  double smax = 0;
  pr.hspeed = INFINITY;
  for (int64_t b = 0; b < pr.nb; ++b) {
    double sb = 0;
    for (int64_t l = 0; l < nn; ++l) {
      int64_t I[3] = {l % N, l / N, 0};
      double x[3];
      node_logical(pr, b, I, x);
      // geometry_terms: sqrt(g) = r th' r', A_1 = r'/(r th'), A_2 = r th'/r'
      const double r = pr.an_r0 + pr.an_dr * x[1];
      const double J = r * pr.an_dth * pr.an_dr;
      const double A1 = pr.an_dr / (r * pr.an_dth), A2 = r * pr.an_dth / pr.an_dr;
      double* c = &pr.dcoef[b * (1 + D) * nn];
      c[l] = J;
      c[nn + l] = A1;
      c[2 * nn + l] = A2;
      sb = std::max(sb, std::max(std::sqrt(A1 / J), std::sqrt(A2 / J)));
    }
    smax = std::max(smax, sb);
    pr.hspeed = std::min(pr.hspeed, pr.h[b] / sb);
    // the lines' pairs, from the factors of sqrt(g) along each dimension:
    // th' along d = 0, r(x_2) r' along d = 1 (jacobian_factor, geometry.jl:220-222)
    const Leaf& Lf = pr.leaves[b];
    for (int d = 0; d < D; ++d)
      for (int e = 0; e < E; ++e) {
        qf a[9];
        for (int i = 0; i < P1; ++i) {
          const qf xi = (qf)pr.plo[Lf.patch][d] +
                        (qf)pr.h[b] * ((qf)(Lf.c[d] * N + (int64_t)e * P1) + (qf)pr.el.offsets[i]);
          const qf th = (qf)M_PI / 2 / (qf)pr.proots[Lf.patch][0];
          const qf dr = ((qf)pr.an_r1 - (qf)pr.an_r0) / (qf)pr.proots[Lf.patch][1];
          a[i] = d == 0 ? th : ((qf)pr.an_r0 + dr * xi) * dr;
        }
        projector_pair(pr, a, &pr.lmode[((b * D + d) * E + e) * P1],
                       &pr.ldual[((b * D + d) * E + e) * P1]);
      }
  }
  (void)smax;
}

// ===========================================================================
// C API (ctypes).
// ===========================================================================
extern "C" {

int orc_element(int P, double* out) {
  // out: nodes(P1) weights(P1) offsets(P1) stencil(P1*(P1+2)) trunc(P1*P1) mode(P1) dual(P1)
  if (P < 1 || P > 8) return -1;
  Element e;
  build_element(P, e);
  int P1 = P + 1, k = 0;
  for (int i = 0; i < P1; ++i) out[k++] = e.nodes[i];
  for (int i = 0; i < P1; ++i) out[k++] = e.weights[i];
  for (int i = 0; i < P1; ++i) out[k++] = e.offsets[i];
  for (int i = 0; i < P1; ++i)
    for (int j = 0; j < P1 + 2; ++j) out[k++] = e.W[i][j];
  for (int i = 0; i < P1; ++i)
    for (int j = 0; j < P1; ++j) out[k++] = e.trunc[i][j];
  for (int i = 0; i < P1; ++i) out[k++] = e.mode[i];
  for (int i = 0; i < P1; ++i) out[k++] = e.dual[i];
  return k;
}

void* orc_brick(int D, int P, int E, int roots, double L, int reflecting, int truncate, int nreg,
                const double* regions) {
  auto* pr = new Problem();
  pr->D = D;
  pr->P = P;
  pr->P1 = P + 1;
  pr->E = E;
  pr->N = E * (P + 1);
  build_element(P, pr->el);
  pr->geometry = 0;
  pr->npatch = 1;
  for (int d = 0; d < 3; ++d) {
    pr->proots[0][d] = d < D ? roots : 1;
    pr->plo[0][d] = 0;
    pr->phi_[0][d] = d < D ? L : 1;
    pr->periodic[d] = !reflecting;
  }
  pr->reflecting = reflecting != 0;
  pr->truncate = truncate != 0;
  int64_t nr = ipow(roots, D);
  for (int64_t r = 0; r < nr; ++r) {
    Leaf Lf{0, 0, {0, 0, 0}};
    int64_t x = r;
    for (int d = 0; d < D; ++d) {
      Lf.c[d] = x % roots;
      x /= roots;
    }
    pr->leaves.push_back(Lf);
  }
  refine_regions(*pr, nreg, regions, -1);
  finalize_leaves(*pr);
  resolve_faces(*pr);
  allocate_work(*pr);
  return pr;
}

void* orc_ball(int P, int E, int n, int m, double Lc, double R1, double grading, int truncate,
               const double* kvec, double phase, int nreg, const double* regions) {
  auto* pr = new Problem();
  pr->D = 3;
  pr->P = P;
  pr->P1 = P + 1;
  pr->E = E;
  pr->N = E * (P + 1);
  build_element(P, pr->el);
  pr->geometry = 1;
  pr->ball.Lc = Lc;
  pr->ball.R1 = R1;
  pr->ball.n = n;
  pr->ball.m = m;
  pr->ball.grading = grading;
  pr->truncate = truncate != 0;
  pr->npatch = 7;
  for (int p = 0; p < 7; ++p)
    for (int d = 0; d < 3; ++d) {
      int k = (p - 1) / 2;
      pr->proots[p][d] = (p > 0 && d == k) ? m : n;
      pr->plo[p][d] = 0;
      pr->phi_[p][d] = pr->proots[p][d];
      pr->periodic[d] = false;
    }
  for (int d = 0; d < 3; ++d) pr->kvec[d] = kvec[d];
  pr->omega = std::sqrt(kvec[0] * kvec[0] + kvec[1] * kvec[1] + kvec[2] * kvec[2]);
  pr->phase = phase;
  for (int p = 0; p < 7; ++p) {
    int64_t nr = (int64_t)pr->proots[p][0] * pr->proots[p][1] * pr->proots[p][2];
    for (int64_t r = 0; r < nr; ++r) {
      Leaf Lf{p, 0, {0, 0, 0}};
      int64_t x = r;
      for (int d = 0; d < 3; ++d) {
        Lf.c[d] = x % pr->proots[p][d];
        x /= pr->proots[p][d];
      }
      pr->leaves.push_back(Lf);
    }
  }
  refine_regions(*pr, nreg, regions, 0);
  finalize_leaves(*pr);
  build_face_index(*pr);
  resolve_faces(*pr);
  delete g_faceindex;
  g_faceindex = nullptr;
  allocate_work(*pr);
  build_ball_metric(*pr);
  return pr;
}

// The four-patch annulus r0 <= r <= r1, each patch (n1, n2) roots, the angular
// faces glued by the identity and the radial ones reflecting (annulus_forest,
// geometry.jl:240-268).
// The seven-patch ball (mo = 0) or the thirteen-patch ball (mo > 0, inner
// shells to Rm), with Dirichlet data of a plane wave or (radiative != 0) the
// radiative boundary of kappa and invR on the outer sphere.
void* orc_ball_ext(int P, int E, int n, int m, int mo, double Lc, double Rm, double R1,
                   int truncate, const double* kvec, double phase, int radiative, double kappa,
                   double invR) {
  auto* pr = new Problem();
  pr->D = 3;
  pr->P = P;
  pr->P1 = P + 1;
  pr->E = E;
  pr->N = E * (P + 1);
  build_element(P, pr->el);
  pr->geometry = 1;
  pr->ball.Lc = Lc;
  pr->ball.R1 = R1;
  pr->ball.n = n;
  pr->ball.m = m;
  pr->ball.thirteen = mo > 0;
  pr->ball.Rm = Rm;
  pr->ball.mo = mo;
  pr->truncate = truncate != 0;
  pr->radiative = radiative != 0;
  pr->rad_kappa = kappa;
  pr->rad_invR = invR;
  pr->npatch = mo > 0 ? 13 : 7;
  for (int p = 0; p < pr->npatch; ++p)
    for (int d = 0; d < 3; ++d) {
      const bool outer = p >= 7;
      const int k = p == 0 ? -1 : ((outer ? p - 7 : p - 1) / 2);
      pr->proots[p][d] = (d == k) ? (outer ? mo : m) : n;
      pr->plo[p][d] = 0;
      pr->phi_[p][d] = pr->proots[p][d];
      pr->periodic[d] = false;
    }
  for (int d = 0; d < 3; ++d) pr->kvec[d] = kvec[d];
  pr->omega = std::sqrt(kvec[0] * kvec[0] + kvec[1] * kvec[1] + kvec[2] * kvec[2]);
  pr->phase = phase;
  for (int p = 0; p < pr->npatch; ++p) {
    int64_t nr = (int64_t)pr->proots[p][0] * pr->proots[p][1] * pr->proots[p][2];
    for (int64_t r = 0; r < nr; ++r) {
      Leaf Lf{p, 0, {0, 0, 0}};
      int64_t x = r;
      for (int d = 0; d < 3; ++d) {
        Lf.c[d] = x % pr->proots[p][d];
        x /= pr->proots[p][d];
      }
      pr->leaves.push_back(Lf);
    }
  }
  finalize_leaves(*pr);
  build_face_index(*pr);
  resolve_faces(*pr);
  delete g_faceindex;
  g_faceindex = nullptr;
  allocate_work(*pr);
  build_ball_metric(*pr);
  return pr;
}

void* orc_annulus(int P, int E, int n1, int n2, double r0, double r1, int truncate) {
  auto* pr = new Problem();
  pr->D = 2;
  pr->P = P;
  pr->P1 = P + 1;
  pr->E = E;
  pr->N = E * (P + 1);
  build_element(P, pr->el);
  pr->geometry = 2;
  pr->truncate = truncate != 0;
  pr->npatch = 4;
  pr->an_r0 = r0;
  pr->an_r1 = r1;
  pr->an_dth = (double)((qf)M_PI / 2 / (qf)n1);
  pr->an_dr = (double)(((qf)r1 - (qf)r0) / (qf)n2);
  for (int p = 0; p < 4; ++p) {
    const int nr[3] = {n1, n2, 1};
    for (int d = 0; d < 3; ++d) {
      pr->proots[p][d] = nr[d];
      pr->plo[p][d] = 0;
      pr->phi_[p][d] = nr[d];
      pr->periodic[d] = false;
    }
    for (int64_t r = 0; r < (int64_t)n1 * n2; ++r)
      pr->leaves.push_back(Leaf{p, 0, {r % n1, r / n1, 0}});
  }
  finalize_leaves(*pr);
  resolve_faces(*pr);
  allocate_work(*pr);
  build_annulus_metric(*pr);
  return pr;
}

// The diagonal metric's coefficients (nb, 1 + D, nn) and the truncation pairs
// (nb, D, E, P1) of an annulus; 0 for other geometries.
int orc_dcoef(void* p, double* coef, double* lmode, double* ldual) {
  Problem* pr = (Problem*)p;
  if (pr->geometry != 2) return 0;
  std::memcpy(coef, pr->dcoef.data(), pr->dcoef.size() * sizeof(double));
  std::memcpy(lmode, pr->lmode.data(), pr->lmode.size() * sizeof(double));
  std::memcpy(ldual, pr->ldual.data(), pr->ldual.size() * sizeof(double));
  return 1;
}

void orc_free(void* p) { delete (Problem*)p; }
int64_t orc_nblocks(void* p) { return ((Problem*)p)->nb; }
int orc_N(void* p) { return ((Problem*)p)->N; }
double orc_hspeed(void* p) { return ((Problem*)p)->hspeed; }

void orc_block(void* p, int64_t b, int64_t* info, double* h) {
  Problem* pr = (Problem*)p;
  const Leaf& L = pr->leaves[b];
  info[0] = L.patch;
  info[1] = L.level;
  info[2] = L.c[0];
  info[3] = L.c[1];
  info[4] = L.c[2];
  *h = pr->h[b];
}

// Face resolution per (block, dim, side): kind and the first neighbour.
void orc_faces(void* p, int64_t* kind, int64_t* nb) {
  Problem* pr = (Problem*)p;
  for (size_t i = 0; i < pr->faces.size(); ++i) {
    kind[i] = pr->faces[i].kind;
    nb[i] = pr->faces[i].kind == G_FINER ? pr->faces[i].fine[0] : pr->faces[i].nb;
  }
}

// Physical node coordinates, layout (nb, D, N^D).
void orc_coords(void* p, double* x) {
  Problem* pr = (Problem*)p;
  const int D = pr->D, N = pr->N;
#pragma omp parallel for schedule(static)
  for (int64_t b = 0; b < pr->nb; ++b)
    for (int64_t l = 0; l < pr->nn; ++l) {
      int64_t I[3] = {l % N, (l / N) % N, l / ((int64_t)N * N)};
      double xx[3];
      node_physical(*pr, b, I, xx);
      for (int d = 0; d < D; ++d) x[(b * D + d) * pr->nn + l] = xx[d];
    }
}

// Coefficients of a full geometry, layout (nb, 1 + D*D, N^D); returns 0 if none.
int orc_coef(void* p, double* out, uint8_t* flat) {
  Problem* pr = (Problem*)p;
  if (pr->geometry != 1) return 0;
  std::memcpy(out, pr->coef.data(), pr->coef.size() * sizeof(double));
  std::memcpy(flat, pr->flat.data(), pr->flat.size());
  return 1;
}

// Exact data.  kind 0: standing sine (wave_exact, sinewave.jl:28-39);
// 1: reflecting cos mode (reflect_exact, :64-76); 2: plane wave (sevenpatch.jl:40-49);
// 3: Gaussian pulse u = exp(-|x-c|^2/w^2), v = 0 (prm: c..., w).
void orc_fill_exact(void* p, int kind, double t, const double* prm, double* y) {
  Problem* pr = (Problem*)p;
  const int D = pr->D, N = pr->N;
  const int64_t nn = pr->nn;
#pragma omp parallel for schedule(static)
  for (int64_t b = 0; b < pr->nb; ++b)
    for (int64_t l = 0; l < nn; ++l) {
      int64_t I[3] = {l % N, (l / N) % N, l / ((int64_t)N * N)};
      double x[3];
      node_physical(*pr, b, I, x);
      double u = 0, v = 0;
      if (kind == 0 || kind == 1) {
        double L = prm[0], m = prm[1];
        double om = (kind == 0 ? 2 * M_PI : M_PI) * m * std::sqrt((double)D) / L;
        double shape = 1;
        for (int d = 0; d < D; ++d)
          shape *= (kind == 0) ? std::sin(2 * M_PI * m * x[d] / L) : cospi_(m * x[d] / L);
        u = std::cos(om * t) * shape;
        v = -om * std::sin(om * t) * shape;
      } else if (kind == 2) {
        double dv[4];
        plane_wave(*pr, x, t, dv);
        double sx = 0;
        for (int d = 0; d < 3; ++d) sx += pr->kvec[d] * x[d];
        double th = sx - pr->omega * t + pr->phase;
        u = std::sin(th);
        v = -pr->omega * std::cos(th);
      } else if (kind == 4) {
        // the rotating Neumann mode u = R(kr) cos(m th - k t), v = k R sin(...),
        // R(z) = s (J_m(z) Y_m'(k r0) - Y_m(z) J_m'(k r0)), s = pi k r0 / 2
        // (annulus.jl:95-135); prm: m, k
        const int m = (int)prm[0];
        const double k = prm[1];
        double xl[3];
        node_logical(*pr, b, I, xl);
        const double r = pr->an_r0 + pr->an_dr * xl[1];
        const double th = pr->leaves[b].patch * (M_PI / 2) + pr->an_dth * xl[0];
        const double kr0 = k * pr->an_r0;
        const double jr0 = (jn(m - 1, kr0) - jn(m + 1, kr0)) / 2;
        const double yr0 = (yn(m - 1, kr0) - yn(m + 1, kr0)) / 2;
        const double sc = M_PI * k * pr->an_r0 / 2;
        const double R = sc * (jn(m, k * r) * yr0 - yn(m, k * r) * jr0);
        const double ph = m * th - k * t;
        u = R * std::cos(ph);
        v = k * R * std::sin(ph);
      } else {
        double r2 = 0;
        for (int d = 0; d < D; ++d) r2 += (x[d] - prm[d]) * (x[d] - prm[d]);
        u = std::exp(-r2 / (prm[D] * prm[D]));
        v = 0;
      }
      y[b * 2 * nn + l] = u;
      y[b * 2 * nn + nn + l] = v;
    }
}

void orc_rhs(void* p, const double* y, double* du, double t) { rhs(*(Problem*)p, y, du, t); }
void orc_truncate(void* p, double* y) { truncate_state(*(Problem*)p, y); }
double orc_energy(void* p, const double* y, double t) { return energy(*(Problem*)p, y, t); }

void orc_gradients(void* p, const double* y, double t, double* g) {
  // owned gradients: flat (D, nb, N^D) logical components; ball (nb, D, N^D) Cartesian
  Problem* pr = (Problem*)p;
  fill_gradient(*pr, y, t);
  const int D = pr->D, N = pr->N;
  for (int64_t b = 0; b < pr->nb; ++b)
    for (int64_t l = 0; l < pr->nn; ++l) {
      int64_t I[3] = {l % N, (l / N) % N, l / ((int64_t)N * N)};
      for (int d = 0; d < D; ++d) {
        if (pr->geometry != 1) {
          int64_t gn[3] = {1, 1, 1};
          for (int e = 0; e < D; ++e) gn[e] = N + (e == d ? 2 : 0);
          int64_t G[3] = {I[0], I[1], I[2]};
          G[d] += 1;
          g[(d * pr->nb + b) * pr->nn + l] = pr->grads[d][b * pr->grad_n[d] + hidx(*pr, G, gn)];
        } else {
          int64_t n3[3] = {1, 1, 1};
          for (int e = 0; e < D; ++e) n3[e] = N + 2;
          int64_t S[3] = {0, 0, 0};
          for (int e = 0; e < D; ++e) S[e] = I[e] + 1;
          g[(b * D + d) * pr->nn + l] = pr->cart[(b * D + d) * pr->ufs_n + hidx(*pr, S, n3)];
        }
      }
    }
}

// Ghost planes after fill_gradient, rows (b, d, s, t): u, and the flat
// gradient d on faces normal to d (ball: Cartesian components, rows x D).
void orc_ghosts(void* p, const double* y, double t, double* ug, double* gg) {
  Problem* pr = (Problem*)p;
  fill_gradient(*pr, y, t);
  const int D = pr->D, N = pr->N;
  const int64_t Nt = ipow(N, D - 1);
  int64_t n[3] = {1, 1, 1};
  for (int e = 0; e < D; ++e) n[e] = N + 2;
  for (int64_t b = 0; b < pr->nb; ++b)
    for (int d = 0; d < D; ++d)
      for (int s = 0; s < 2; ++s) {
        int tt[2];
        int nt = tangential(D, d, tt);
        for (int64_t t2 = 0; t2 < Nt; ++t2) {
          int64_t T[2] = {t2 % N, t2 / N};
          const int64_t row = ((b * D + d) * 2 + s) * Nt + t2;
          int64_t S[3] = {0, 0, 0};
          S[d] = s ? N + 1 : 0;
          for (int k = 0; k < nt; ++k) S[tt[k]] = T[k] + 1;
          ug[row] = pr->ufs[b * pr->ufs_n + hidx(*pr, S, n)];
          if (pr->geometry != 1) {
            int64_t gn[3] = {1, 1, 1};
            for (int e = 0; e < D; ++e) gn[e] = N + (e == d ? 2 : 0);
            int64_t G[3] = {0, 0, 0};
            G[d] = s ? N + 1 : 0;
            for (int k = 0; k < nt; ++k) G[tt[k]] = T[k];
            gg[row] = pr->grads[d][b * pr->grad_n[d] + hidx(*pr, G, gn)];
          } else {
            for (int i = 0; i < D; ++i)
              gg[(((b * D + d) * 2 + s) * D + i) * Nt + t2] = pr->cart[(b * D + i) * pr->ufs_n + hidx(*pr, S, n)];
          }
        }
      }
}

// D_d g with g exchanged as the gradient's component d (odd under a mirror in
// d): the Laplacian pass with every other gradient zero (refined_tests.jl:54-68).
void orc_divergence(void* p, const double* g, int d, double* out) {
  Problem* pr = (Problem*)p;
  const int D = pr->D, N = pr->N;
  for (int e = 0; e < D; ++e) std::fill(pr->grads[e].begin(), pr->grads[e].end(), 0.0);
  int64_t gn[3] = {1, 1, 1};
  for (int e = 0; e < D; ++e) gn[e] = N + (e == d ? 2 : 0);
  for (int64_t b = 0; b < pr->nb; ++b)
    for (int64_t l = 0; l < pr->nn; ++l) {
      int64_t I[3] = {l % N, (l / N) % N, l / ((int64_t)N * N)};
      int64_t G[3] = {I[0], I[1], I[2]};
      G[d] += 1;
      pr->grads[d][b * pr->grad_n[d] + hidx(*pr, G, gn)] = g[b * pr->nn + l];
    }
  exchange_grads(*pr);
  std::vector<double> y(pr->nb * 2 * pr->nn, 0.0), du(pr->nb * 2 * pr->nn);
  laplacian_flat(*pr, y.data(), du.data());
  for (int64_t b = 0; b < pr->nb; ++b)
    for (int64_t l = 0; l < pr->nn; ++l) out[b * pr->nn + l] = du[b * 2 * pr->nn + pr->nn + l];
}

void orc_norms(void* p, const double* err, int vars_mask, int pnorm, double* lp, double* linf) {
  norms(*(Problem*)p, err, vars_mask, pnorm, lp, linf);
}

double orc_min_spacing(void* p) {
  Problem* pr = (Problem*)p;
  double h = INFINITY;
  for (double x : pr->h) h = std::min(h, x);
  return h;
}

double orc_timestep(void* p, double cfl) {
  Problem* pr = (Problem*)p;
  double h = pr->geometry == 0 ? orc_min_spacing(p) : pr->hspeed;  // metric: min h_b / speed_b
  return cfl * h / (pr->P + 1);
}

// Classic RK4 with the truncation as stage and step limiter (stepping.jl:72-78):
// y_i = L(y_n + dt sum_j a_ij k_j), y_{n+1} = L(y_n + dt sum_i b_i k_i).
void orc_rk4(void* p, double* y, double t0, double dt, int64_t nsteps) {
  Problem* pr = (Problem*)p;
  const int64_t n = pr->nb * 2 * pr->nn;
  std::vector<double> k1(n), k2(n), k3(n), k4(n), ys(n);
  const double b1 = 1.0 / 6, b2 = 1.0 / 3, b3 = 1.0 / 3, b4 = 1.0 / 6;
  for (int64_t s = 0; s < nsteps; ++s) {
    double t = t0 + s * dt;
    rhs(*pr, y, k1.data(), t);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) ys[i] = y[i] + dt * (0.5 * k1[i]);
    truncate_state(*pr, ys.data());
    rhs(*pr, ys.data(), k2.data(), t + dt / 2);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) ys[i] = y[i] + dt * (0.5 * k2[i]);
    truncate_state(*pr, ys.data());
    rhs(*pr, ys.data(), k3.data(), t + dt / 2);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) ys[i] = y[i] + dt * k3[i];
    truncate_state(*pr, ys.data());
    rhs(*pr, ys.data(), k4.data(), t + dt);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) y[i] = y[i] + dt * (b1 * k1[i] + b2 * k2[i] + b3 * k3[i] + b4 * k4[i]);
    truncate_state(*pr, y);
  }
}


// A brick of explicit leaves (level, c0, c1, c2 each), in any order: the forest
// after a regrid, built here from its keys alone.
void* orc_brick_leaves(int D, int P, int E, int roots, double L, int reflecting, int truncate,
                       int64_t nleaves, const int64_t* leaves) {
  auto* pr = new Problem();
  pr->D = D;
  pr->P = P;
  pr->P1 = P + 1;
  pr->E = E;
  pr->N = E * (P + 1);
  build_element(P, pr->el);
  pr->geometry = 0;
  pr->npatch = 1;
  for (int d = 0; d < 3; ++d) {
    pr->proots[0][d] = d < D ? roots : 1;
    pr->plo[0][d] = 0;
    pr->phi_[0][d] = d < D ? L : 1;
    pr->periodic[d] = !reflecting;
  }
  pr->reflecting = reflecting != 0;
  pr->truncate = truncate != 0;
  for (int64_t i = 0; i < nleaves; ++i) {
    const int64_t* q = leaves + 4 * i;
    pr->leaves.push_back(Leaf{0, (int)q[0], {q[1], q[2], q[3]}});
  }
  finalize_leaves(*pr);
  resolve_faces(*pr);
  allocate_work(*pr);
  return pr;
}

// The modal amplitudes of variable var of y (refinement.jl:56-101): per element
// post * max over its lines and the modes q of ((weights[q]/A) |c_modes[q]|)^pows[q],
// A = max |var| over the forest, zero when A is.  tau: e_1 fastest, block slowest.
void orc_modes(void* pv, const double* y, int var, int nq, const int* modes,
               const double* weights, const double* pows, double post, double* tau) {
  const Problem& pr = *(const Problem*)pv;
  const int D = pr.D, N = pr.N, P1 = pr.P1, E = pr.E;
  double peak = 0;
  for (int64_t b = 0; b < pr.nb; ++b)
    for (int64_t l = 0; l < pr.nn; ++l) peak = std::max(peak, std::fabs(y[(b * 2 + var) * pr.nn + l]));
  const double scale = peak == 0 ? 0 : 1 / peak;
  const int64_t ne = ipow(E, D);
  for (int64_t b = 0; b < pr.nb; ++b) {
    const double* a = y + (b * 2 + var) * pr.nn;
    for (int64_t e = 0; e < ne; ++e) {
      int64_t J[3] = {0, 0, 0};
      for (int d = 0, r = (int)e; d < D; ++d, r /= E) J[d] = r % E;
      double best = 0;
      for (int d = 0; d < D; ++d) {
        // every line of the element along d: the other coordinates over its nodes
        const int64_t nlines = ipow(P1, D - 1);
        for (int64_t t = 0; t < nlines; ++t) {
          int64_t I[3] = {0, 0, 0};
          for (int k = 0, r = (int)t; k < D; ++k) {
            if (k == d) continue;
            I[k] = J[k] * P1 + r % P1;
            r /= P1;
          }
          for (int q = 0; q < nq; ++q) {
            double c = 0;
            for (int j = 0; j < P1; ++j) {
              I[d] = J[d] * P1 + j;
              c += pr.el.duals[modes[q]][j] * a[I[0] + N * (I[1] + (int64_t)N * I[2])];
            }
            best = std::max(best, std::pow(weights[q] * scale * std::fabs(c), pows[q]));
          }
        }
      }
      tau[b * ne + e] = post * best;
    }
  }
}

// The state of a regrid, by keys: every leaf of dst from src, a copy where src
// holds the same leaf, the parent's polynomial at the leaf's nodes where src
// holds the parent, and the L2 projection R = M_c^-1 I^T M_f of the children
// where src holds those (TreeAMR regrid! with ElementLocal operators).
// Returns the number of dst leaves of each kind in counts[3], or -1 where a
// leaf has no origin.
int orc_transfer(void* dv, double* ydst, void* sv, const double* ysrc, int64_t* counts) {
  const Problem& pd = *(const Problem*)dv;
  const Problem& ps = *(const Problem*)sv;
  const int D = pd.D, N = pd.N, P1 = pd.P1, H = pd.E / 2;
  const int64_t nn = pd.nn;
  counts[0] = counts[1] = counts[2] = 0;
  for (int64_t b = 0; b < pd.nb; ++b) {
    const Leaf& L = pd.leaves[b];
    int64_t sb;
    auto node = [&](int64_t l, int64_t I[3]) {
      I[0] = l % N;
      I[1] = D >= 2 ? (l / N) % N : 0;
      I[2] = D >= 3 ? l / ((int64_t)N * N) : 0;
    };
    if (is_leaf(ps, L.patch, L.level, L.c, &sb)) {
      ++counts[0];
      for (int v = 0; v < 2; ++v)
        for (int64_t l = 0; l < nn; ++l) ydst[(b * 2 + v) * nn + l] = ysrc[(sb * 2 + v) * nn + l];
      continue;
    }
    int64_t pc[3] = {L.c[0] >> 1, L.c[1] >> 1, L.c[2] >> 1};
    if (L.level > 0 && is_leaf(ps, L.patch, L.level - 1, pc, &sb)) {
      ++counts[1];
      for (int64_t l = 0; l < nn; ++l) {
        int64_t I[3];
        node(l, I);
        int64_t ep[3] = {0, 0, 0};
        double xr[3] = {0, 0, 0};
        for (int d = 0; d < D; ++d) {
          const int64_t ec = I[d] / P1, j = I[d] % P1, o = L.c[d] & 1;
          ep[d] = o * H + ec / 2;
          xr[d] = (pd.el.nodes[j] + 2 * (ec % 2) - 1) / 2;
        }
        for (int v = 0; v < 2; ++v) {
          double val = 0;
          for (int64_t k = 0; k < ipow(P1, D); ++k) {
            double w = 1;
            int64_t S[3] = {0, 0, 0};
            for (int d = 0, r = (int)k; d < D; ++d, r /= P1) {
              w *= lagrange(pd.el, r % P1, xr[d]);
              S[d] = ep[d] * P1 + r % P1;
            }
            val += w * ysrc[(sb * 2 + v) * nn + S[0] + N * (S[1] + (int64_t)N * S[2])];
          }
          ydst[(b * 2 + v) * nn + l] = val;
        }
      }
      continue;
    }
    int64_t ch[8];
    bool all = true;
    for (int c = 0; c < (1 << D) && all; ++c) {
      int64_t cc[3] = {0, 0, 0};
      for (int d = 0; d < D; ++d) cc[d] = 2 * L.c[d] + ((c >> d) & 1);
      all = is_leaf(ps, L.patch, L.level + 1, cc, &ch[c]);
    }
    if (!all) return -1;
    ++counts[2];
    for (int64_t l = 0; l < nn; ++l) {
      int64_t I[3];
      node(l, I);
      int64_t ic[3] = {0, 0, 0}, eb[3] = {0, 0, 0};
      int c = 0;
      for (int d = 0; d < D; ++d) {
        const int64_t e = I[d] / P1, o = e / H;
        ic[d] = I[d] % P1;
        eb[d] = 2 * (e - o * H);
        c |= (int)o << d;
      }
      for (int v = 0; v < 2; ++v) {
        double val = 0;
        for (int64_t k = 0; k < ipow(2 * P1, D); ++k) {
          double w = 1;
          int64_t S[3] = {0, 0, 0};
          for (int d = 0, r = (int)k; d < D; ++d, r /= 2 * P1) {
            const int fj = r % (2 * P1), f = fj / P1, jj = fj % P1;
            const double x = (pd.el.nodes[jj] + 2 * f - 1) / 2;
            w *= lagrange(pd.el, (int)ic[d], x) * pd.el.weights[jj] / (2 * pd.el.weights[ic[d]]);
            S[d] = (eb[d] + f) * P1 + jj;
          }
          val += w * ysrc[(ch[c] * 2 + v) * nn + S[0] + N * (S[1] + (int64_t)N * S[2])];
        }
        ydst[(b * 2 + v) * nn + l] = val;
      }
    }
  }
  return 0;
}

// The Legendre duals of the element, (P+1, P+1) [mode][node].
int orc_duals(int P, double* out) {
  if (P < 1 || P > 8) return -1;
  Element e;
  build_element(P, e);
  for (int m = 0; m <= P; ++m)
    for (int j = 0; j <= P; ++j) out[m * (P + 1) + j] = e.duals[m][j];
  return 0;
}

int orc_threads(void) { return omp_get_max_threads(); }

}  // extern "C"
