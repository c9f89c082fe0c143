/*
 * twoldg.h -- C ABI of the MI355X (gfx950) OLDG field-evolution library.
 *
 * This is the drop-in boundary for TreeWaveOLDG.jl's field-evolution hot path.
 * A Julia shim (ccall), the Python host package `twoldg` (ctypes) or any C
 * caller binds these symbols; no Torch or HIP type crosses the boundary.
 *
 * Conventions
 *   - Every entry point returns an int status: 0 on success, a negative
 *     TWOLDG_E* code on failure.  twoldg_last_error() then holds the reason
 *     (thread-local), the text the reference raises as ArgumentError.
 *   - State vectors are in the reference's statearray layout: for block b,
 *     variable v (0 = u, 1 = v) and owned node I (x fastest), entry
 *     (b * 2 + v) * N^D + lin(I); src/stepping.jl:36-41.
 *   - The library owns every device buffer of a problem.  One HIP stream per
 *     problem; calls on one problem are serialized, distinct problems may be
 *     used from distinct threads.
 *   - Forest description: blocks in the caller's order, each with its spacing
 *     h, and one face descriptor per (block, dimension, side) saying where its
 *     ghost plane comes from (TreeAMR's ghost schedule, evolution.jl:1093-1097).
 */
#ifndef TWOLDG_H
#define TWOLDG_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TWOLDG_ABI_VERSION 4

/* status codes */
#define TWOLDG_OK 0
#define TWOLDG_EINVAL (-1)   /* an argument the reference refuses (ArgumentError) */
#define TWOLDG_EDEVICE (-2)  /* a HIP runtime failure */
#define TWOLDG_ENOMEM (-3)
#define TWOLDG_ESTATE (-4)   /* call out of order (e.g. RHS before upload) */

/* arithmetic forms */
#define TWOLDG_F64 0   /* FP64 throughout, the reference's structure */
#define TWOLDG_F32C 1  /* FP32 stencil and storage, FP32 hi+lo (Kahan) RK step, FP64 reductions */
#define TWOLDG_SPLIT 2 /* u and its stencil in FP64, v in FP32 with its step update added to an
                          FP32 (hi, lo) pair: the fused 3D stages only (the state crosses the
                          boundary and every other entry point as FP64) */

/* face kinds: where the ghost plane of a block face comes from */
#define TWOLDG_FACE_SAME 1      /* same-level neighbour (periodic wrap included); nb[0] */
#define TWOLDG_FACE_COARSER 2   /* coarser neighbour nb[0]; off[] = our half inside it */
#define TWOLDG_FACE_FINER 3     /* 2^(D-1) finer neighbours nb[] by tangential half bits */
#define TWOLDG_FACE_MIRROR 4    /* reflecting wall: parity-signed own face plane */
#define TWOLDG_FACE_GLUE 5      /* glued patch face: nb[0], its face nface, tangential map */
#define TWOLDG_FACE_DIRICHLET 6 /* outer face with Dirichlet data */
#define TWOLDG_FACE_RADIATIVE 7 /* outer face under the radiative condition
                                   n . grad u = -(rad_kappa du/dt + rad_invR u):
                                   RadiativeBoundary, evolution.jl:1129-1270 */

/* slots of a problem's device state buffers.  Only twoldg_rk4 advances STATE;
   the pure entry points (rhs, truncate, energy, norms) read and write the slots
   they are given, so a caller that keeps an integrator's state in STATE passes
   its own vectors through ARG (and DU, AUX) without disturbing it.  Slots other
   than STATE are allocated on their first use. */
#define TWOLDG_SLOT_STATE 0  /* the integrator's state y_n (FP32C: with its compensation) */
#define TWOLDG_SLOT_DU 1     /* right-hand side output */
#define TWOLDG_SLOT_AUX 2    /* caller data, e.g. an exact solution for error norms */
#define TWOLDG_SLOT_ARG 3    /* the argument of a pure entry point */
#define TWOLDG_NSLOTS 4

typedef struct twoldg_face {
  int32_t kind;
  int32_t nface;   /* GLUE: the neighbour's face, 2*d' + s' (s' = 0 low, 1 high) */
  int32_t map[2];  /* GLUE: for the neighbour's tangential axis j (in increasing order),
                      +/-(1 + k): it runs along our tangential axis k, reversed if < 0 */
  int32_t off[3];  /* COARSER: our offset (0/1) inside the coarse block, per dimension */
  int32_t pad;
  int64_t nb[4];   /* neighbour block(s) */
} twoldg_face;

typedef struct twoldg_desc {
  int32_t abi_version;   /* TWOLDG_ABI_VERSION */
  int32_t D, P, E;       /* N = E (P+1) nodes per block per dimension, E even */
  int32_t precision;     /* TWOLDG_F64, TWOLDG_F32C or TWOLDG_SPLIT */
  int32_t truncate;      /* 1: v held at degree P-1 (the paper's scheme), 0: plain */
  int64_t nblocks;
  const double* spacing;        /* [nblocks] h per block (block extent / N) */
  const twoldg_face* faces;     /* [nblocks * D * 2], index (b*D + d)*2 + s */
  /* full metric (a FullGeometry such as the seven-patch ball); NULL when flat */
  const double* coef;           /* (nblocks, 1 + D*D, N^D): sqrt(g), S[d][i] at 1 + d*D + i */
  const uint8_t* flat;          /* [nblocks] frame is one constant diagonal */
  /* per-element sqrt(g)-orthogonal truncation tables, NULL when tensor-only */
  const int32_t* tslot;         /* [nblocks] table index or -1 (tensor projector) */
  int64_t ntables;
  const double* tmodes;         /* (n, m) Legendre basis of Q_{P-1}: n=(P+1)^D, m=P^D */
  const double* tweights;       /* (ntables, E^D, n) quadrature weights times sqrt(g) */
  const double* tchol;          /* (ntables, E^D, m, m) upper Cholesky factor, column-major */
  /* Dirichlet data on DIRICHLET faces: a plane wave sin(k.x - |k| t + phi) */
  const double* dirichlet_x;    /* (nblocks, D, 2, D, N^(D-1)) physical position of every face
                                   ghost (only DIRICHLET faces read), may be NULL */
  double wave_k[3];
  double wave_phi;
  double hspeed;                /* min over blocks of h_b / max logical speed (0: flat) */
  int32_t flags;                /* TWOLDG_FLAG_* */
  int32_t pad;
  /* diagonal metric (a Geometry such as the annulus, evolution.jl:67-70); NULL when
     flat or full.  The gradient sets hold the fluxes F_d = A_d D_d u, the
     Laplacian divides by sqrt(g), energies and norms weigh by sqrt(g). */
  const double* dcoef;          /* (nblocks, 1 + D, N^D): sqrt(g), A_1 .. A_D */
  const double* lmode;          /* (nblocks, D, E, P+1): each line's rank-one truncation */
  const double* ldual;          /*   pair (MetricTruncation, evolution.jl:556-575) */
  /* factored truncation (FactoredTruncation, evolution.jl:296-380, 788-798, 904-984):
     on blocks whose sqrt(g) is a radial factor along one dimension k times an
     angular one across it (the thirteen-patch ball's outer shells), a rank-one
     pair along k times a sqrt(g)-orthogonal projection on each plane across k.
     NULL fslot: none.  Needs a full metric, D = 3 and P <= 4. */
  const int32_t* fslot;         /* [nblocks] factored-table index f, or -1 */
  int64_t nfactored;
  const int32_t* fradial;       /* [nf] the radial dimension k, 0-based */
  const double* frmodes;        /* (nf, E, P+1): the radial pair of each element */
  const double* frduals;        /*   along k */
  const int32_t* faslot;        /* [nf] angular table index A */
  int64_t nangular;
  const double* famodes;        /* ((P+1)^2, P^2): Legendre basis of a plane, first
                                   tangential index fastest */
  const double* faweights;      /* (nangular, E^2, (P+1)^2): quadrature weight times the
                                   angular factor */
  const double* fachol;         /* (nangular, E^2, P^4): upper Cholesky factor of
                                   Psi^T W Psi, column-major */
  /* TWOLDG_FACE_RADIATIVE faces (RadiativeFaces, evolution.jl:1166-1203) */
  double rad_kappa;             /* kappa >= 0 */
  double rad_invR;              /* 1/R of the u/R term, 0 without it */
} twoldg_desc;

/* flags */
#define TWOLDG_FLAG_GENERIC 1  /* never use the fused one-launch stages */

typedef struct twoldg_problem twoldg_problem;

/* --- library ------------------------------------------------------------ */
int twoldg_abi_version(void);
const char* twoldg_last_error(void);
int twoldg_device_count(int* out);

/* --- problem (WaveProblem, src/evolution.jl:1043-1108) ------------------- */
int twoldg_problem_create(const twoldg_desc* desc, twoldg_problem** out);
int twoldg_problem_destroy(twoldg_problem* p);
/* state vector length, device bytes held, and which stages RK4 runs: the fused
   one-launch 3D stages (1), the batch stage on a two-dimensional uniform flat
   problem (2: one launch per stage), or the generic exchange + gradient + stage
   sequence (0) */
int twoldg_problem_info(const twoldg_problem* p, int64_t* state_len, int64_t* device_bytes,
                        int32_t* fused);

/* --- state transfer (statevector / gather!; layout above) --------------- */
int twoldg_state_upload(twoldg_problem* p, int slot, const double* host);
int twoldg_state_download(twoldg_problem* p, int slot, double* host);
/* copy one slot into another on the device.  FP32C: a copy out of STATE holds
   its compensated value rounded to FP32 (hi - lo), a copy into STATE starts a
   zero compensation. */
int twoldg_state_copy(twoldg_problem* p, int dst, int src);

/* --- the hot path -------------------------------------------------------- */
/* slot dst <- f(slot src, t): wave_rhs!(du, u, p, t), src/evolution.jl:1367-1371
   (pure: never writes src, keeps nothing between calls; dst != src) */
int twoldg_rhs(twoldg_problem* p, int src, int dst, double t);
/* v <- (Pi x ... x Pi) v in place: truncate_state!, src/evolution.jl:1383-1400 */
int twoldg_truncate(twoldg_problem* p, int slot);
/* nsteps classic RK4 steps of the state slot, truncation as stage and step
   limiter: wave_integrator + solve!, src/stepping.jl:72-78 */
int twoldg_rk4(twoldg_problem* p, double t0, double dt, int64_t nsteps);
/* discrete energy: wave_energy, src/evolution.jl:1472-1489; under radiative faces
   with a u/R term it includes the boundary's 1/2 sum w |S[d]| u^2/R over the outer
   faces' nodes (boundary_energy, evolution.jl:1276-1299) */
int twoldg_energy(twoldg_problem* p, int slot, double t, double* out);
/* p-norm (pnorm >= 1) and max norm of slot a, or of a - b when b >= 0:
   quadrature_norm / wave_norm, src/layout.jl:223-235, src/evolution.jl:1498-1512.
   vars_mask: bit 0 = u, bit 1 = v. */
int twoldg_norms(twoldg_problem* p, int a, int b, int vars_mask, int pnorm, double* lp,
                 double* linf);
/* twoldg_norms without any allocation: no counterpart in the reference.  For
   callers that cannot allocate device memory while they run (code under HIP
   graph capture, tight serving loops, harnesses that forbid hipMalloc): the
   per-block partial sums go to a scratch array of fixed size that the library
   holds on each GPU until the process ends (3 MiB), not to the problem's
   buffers.  Takes problems of at most 131072 blocks (larger: TWOLDG_EINVAL,
   use twoldg_norms).  Every call on one GPU shares that GPU's scratch: calls
   that overlap in time on one GPU (from several threads, or on several
   problems' streams) race and must be serialized by the caller. */
int twoldg_norms_noalloc(twoldg_problem* p, int a, int b, int vars_mask, int pnorm, double* lp,
                         double* linf);
/* cfl * h / (P+1), on a metric (full or diagonal) h = min over blocks of h_b
   over the block's largest logical speed: oldg_timestep, src/evolution.jl:1528-1531 */
int twoldg_timestep(const twoldg_problem* p, double cfl, double* out);
/* wait for the problem's stream */
int twoldg_synchronize(twoldg_problem* p);
/* Fill every working array of p (stage states, ghost rows, gradient sets, face
   buffers and ghost slots, reduction partials; not the slots) with all-ones
   bytes, a NaN in FP32 and FP64, and wait.  Every entry point writes what it
   reads of these arrays first, so any call after this one returns what it
   would have returned without it.  No counterpart among the reference's entry
   points: its tests poison the working arrays the same way
   (test/refined_tests.jl:145-161) to show that no kernel reads an entry it has
   not written. */
int twoldg_problem_poison(twoldg_problem* p);

/* --- adaptive refinement ----------------------------------------------------- */
/* The Legendre-mode amplitudes of variable var (0 = u, 1 = v) of a slot, one
   value per element into tau (nblocks * E^D doubles; e_1 fastest, then e_2 ..
   e_D, the block slowest): with c_m the coefficient of L_m along a line of the
   element and A the largest |var| over the problem,
       tau = post * max over the element's lines in every dimension and
             q < nmodes of ((weights[q] / A) |c_modes[q]|)^pows[q],
   zero where A is zero.  mode_kernel! and scaled_modes, src/refinement.jl:56-101;
   modal_indicator (refinement.jl:148-155) is modes 1..M, weights (2m-1)!!,
   pows M/m and post 1/(2M-1)!!; mode_amplitude (refinement.jl:168-174) is the
   single mode with weight, pow and post 1.  1 <= nmodes <= 8, 0 <= modes[q] <= P. */
int twoldg_modes(twoldg_problem* p, int slot, int var, int nmodes, const int32_t* modes,
                 const double* weights, const double* pows, double post, double* tau);

#define TWOLDG_XFER_COPY 0
#define TWOLDG_XFER_REFINE 1
#define TWOLDG_XFER_COARSEN 2
/* Where one block of a regridded forest takes its data from in the forest
   before the regrid. */
typedef struct twoldg_transfer_block {
  int32_t kind;    /* TWOLDG_XFER_* */
  int32_t child;   /* REFINE: the block's offset in its parent, bit d for dimension d */
  int64_t src[8];  /* COPY, REFINE: the source block src[0]; COARSEN: the 2^D children,
                      the child at offset c (bit d for dimension d) in src[c] */
} twoldg_transfer_block;
/* The state of a regrid: every block of dst's slot dst_slot from src's slot
   src_slot, map[b] for dst block b (dst's nblocks entries).  A copy keeps the
   block; a refined block evaluates its parent's element polynomials at its own
   nodes; a coarsened block is the exact L2 projection R = M_c^-1 I^T M_f of its
   children.  Both keep the quadrature integral sum w u of each variable.  The
   element-local operators of TreeAMR's regrid! that blast.jl:336-348 calls;
   the truncation the reference applies after it is twoldg_truncate.  dst and
   src must share D, P, E, precision and device; src is not written. */
int twoldg_transfer(twoldg_problem* dst, int dst_slot, twoldg_problem* src, int src_slot,
                    const twoldg_transfer_block* map);

/* --- host-only inspection --------------------------------------------------- */
/* The exchange operator the problem would build, on the host, without touching
   a GPU: rows (b, d, s, t) in order (b*D + d)*2 + s then t; CSR with entries
   (block, node, coefficient for u and the state, coefficient for gradients).
   Call once with NULL arrays to get *nrows and *nnz, then with arrays of those
   sizes (rowptr has nrows + 1 entries). */
int twoldg_exchange_plan(const twoldg_desc* desc, int64_t* nrows, int64_t* nnz,
                         int32_t* rowptr, int32_t* cb, int32_t* cl, double* vu, double* vg);

/* Logical positions of every owned node of a flat forest, on the host:
   x[(b*D + d)*N^D + node] = lo[b*D + d] + h_b ((c_d N + q (P+1)) + offsets[i])
   with node = (i_1, ..., i_D) (first index fastest), i_d = q (P+1) + i, c the
   block's integer position at its level (corner[b*D + d]) and lo its patch's
   origin (lo[b*D + d]), in this order of rounding.  Replaces the positions the
   reference's layout computes (src/layout.jl:59-76, TreeAMR's nodeposition)
   that the Julia side needs to evaluate initial data f(x, v). */
int twoldg_node_coordinates(const twoldg_desc* desc, const double* lo, const int64_t* corner,
                            double* x);

/* --- measurement ---------------------------------------------------------- */
/* Run nsteps RK4 steps like twoldg_rk4, timed with HIP events on the problem's
   stream: out_ms = the whole run, out_kernel_ms = the summed time of the stage
   kernels (the Laplacian + RK combination + truncation launches), measured by
   an event pair around each of them, out_launches = their count.  which is
   reserved (0). */
int twoldg_rk4_timed(twoldg_problem* p, double t0, double dt, int64_t nsteps, int which,
                     double* out_ms, double* out_kernel_ms, int64_t* out_launches);

/* --- batches of independent problems --------------------------------------- */
/* Many independent flat 2D problems (uniform: same-level or mirror faces, P <= 7,
   (E(P+1))^2 <= 256 nodes per block), each with its own order, spacing and time
   step, advanced together: one launch per RK4 stage over a device-resident work
   list of (order, first block, blocks per workgroup), the four stage launches of
   a step captured once in a hipGraph and replayed.  The blocks per workgroup of
   each order are chosen at creation by timing the candidates.  All problems
   share precision and truncation.  There is no batch in the reference: this is
   its wave_errors driver run over many problems at once (sinewave.jl:228-268). */
typedef struct twoldg_batch twoldg_batch;
int twoldg_batch_create(const twoldg_desc* const* descs, int64_t n, const double* dt,
                        twoldg_batch** out);
int twoldg_batch_destroy(twoldg_batch* b);
/* problem i's state, statearray layout of that problem */
int twoldg_batch_upload(twoldg_batch* b, int64_t i, const double* host);
int twoldg_batch_download(twoldg_batch* b, int64_t i, double* host);
/* nsteps RK4 steps of every problem (graph replays); out_ms: HIP-event time */
int twoldg_batch_rk4(twoldg_batch* b, int64_t nsteps, double* out_ms);
/* nodes per field over all problems, workgroups per stage launch, and per order
   P = 1..8 the chosen blocks per workgroup (tiles[8]; 0: order absent) and the
   autotune's timings (tune_ms[16]): [P-1] the stage launch of that order's
   blocks alone at the first pass's tile, [8+P-1] the full stage launch with the
   chosen tile (second pass), both in ms per launch */
int twoldg_batch_info(const twoldg_batch* b, int64_t* nodes, int64_t* workgroups,
                      int32_t* tiles, double* tune_ms);

#ifdef __cplusplus
}
#endif

#endif /* TWOLDG_H */
