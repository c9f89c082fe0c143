// batch.hip -- many independent 2D problems advanced together: one launch per
// RK4 stage over every block (a run table in the kernel arguments maps each
// workgroup to its tile), a hipGraph per step, and a blocks-per-workgroup tile
// per element order chosen by timing.
#include "twoldg_impl.h"
#include "fused2.h"

struct twoldg_batch {
  int precision = TWOLDG_F64;
  bool truncate = true;
  size_t elt = 8;
  int64_t nprob = 0, nb = 0, total = 0, nfaces = 0;
  std::vector<int64_t> prob_off, prob_len;  // state slice of each problem
  std::vector<int32_t> P1s_host;
  std::vector<int64_t> off_host, foff_host;  // per block: state and face-buffer offsets
  int tiles[10] = {0};                      // blocks per workgroup by P1
  double tune_ms[10] = {0};                 // first autotune pass: the order's launch alone
  double tune_full_ms[10] = {0};            // second pass: the full launch with the chosen tile
  // device
  int64_t *off = nullptr, *foff = nullptr;
  int32_t* nfo = nullptr;
  void* nrh = nullptr;
  int32_t *P1s = nullptr, *Ns = nullptr;
  double *spacing = nullptr, *dt = nullptr;
  void* tables = nullptr;
  std::vector<BatchRun> runs;  // the launch's runs (tiles[P1] blocks per workgroup)
  int64_t nwg = 0;             // its workgroups
  void *yn = nullptr, *ys = nullptr, *acc = nullptr, *lo = nullptr;
  void *uf[2] = {nullptr, nullptr}, *pf[2] = {nullptr, nullptr};
  hipStream_t stream = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  int64_t device_bytes = 0;
  // a problem's own batch of one: the state, stage, accumulator and
  // compensation buffers and the stream are the problem's
  bool borrowed = false;
  double cur_dt = 0;  // the time step in the device table (borrowed form)
};

namespace {

int balloc(twoldg_batch* B, void** p, size_t bytes) {
  hipError_t e = pool_alloc(p, bytes ? bytes : 8, false);
  if (e != hipSuccess)
    return fail(TWOLDG_ENOMEM, std::string("hipMalloc of ") + std::to_string(bytes) +
                                   " bytes: " + hipGetErrorString(e));
  B->device_bytes += (int64_t)bytes;
  return 0;
}

// Host tables into device memory through fine-grained host staging and a copy
// kernel on the batch's stream (see host_stage_alloc and h2d in twoldg_impl.h).
int bh2d(twoldg_batch* B, void* dst, const void* src, size_t bytes) {
  if (bytes == 0) return 0;
  const size_t words = (bytes + 3) / 4;
  void* pinned = nullptr;
  HIP_TRY(host_stage_alloc(&pinned, words * 4));
  std::memcpy(pinned, src, bytes);
  if (bytes % 4) std::memset((char*)pinned + bytes, 0, words * 4 - bytes);
  hipLaunchKernelGGL(k_copy_words, dim3((unsigned)((words + 255) / 256)), dim3(256), 0, B->stream,
                     (int64_t)words, (const uint32_t*)pinned, (uint32_t*)dst);
  hipError_t e = hipStreamSynchronize(B->stream);
  pool_free(pinned);
  if (e != hipSuccess) return fail(TWOLDG_EDEVICE, std::string("batch table upload: ") + hipGetErrorString(e));
  return 0;
}

template <class T>
int batch_tables(twoldg_batch* B) {
  std::vector<Tables<T>> tb(8);
  std::memset(tb.data(), 0, tb.size() * sizeof(Tables<T>));
  for (int P = 1; P <= 8; ++P) {
    const twoldg::ElementTables& et = twoldg::kElementTables[P - 1];
    for (int i = 0; i <= P; ++i) {
      for (int j = 0; j < P + 3; ++j) tb[P - 1].W[i][j] = (T)et.stencil[i][j];
      tb[P - 1].mode[i] = (T)et.topmode[i];
      tb[P - 1].dual[i] = (T)et.topdual[i];
      tb[P - 1].wq[i] = et.weights[i];
    }
  }
  int rc = balloc(B, &B->tables, tb.size() * sizeof(Tables<T>));
  if (rc) return rc;
  if ((rc = bh2d(B, B->tables, tb.data(), tb.size() * sizeof(Tables<T>)))) return rc;
  return 0;
}

// Blocks per workgroup at most (elt: bytes per value).
int max_tile(int N, int E, size_t elt) {
  return batch_max_tile(N, E, elt == 4 ? batch_threads<float>() : batch_threads<double>());
}

// The runs with tile[P1] blocks per workgroup; only orders in `only` (0: all).
// Returns the number of workgroups.
int64_t make_runs(const twoldg_batch* B, const std::vector<int32_t>& Nh, const int* tile, int only,
                  std::vector<BatchRun>& runs) {
  runs.clear();
  int64_t b = 0, wg = 0;
  while (b < B->nb) {
    const int P1 = B->P1s_host[b], N = Nh[b];
    int64_t e = b;
    while (e < B->nb && B->P1s_host[e] == P1 && Nh[e] == N) ++e;
    if (only == 0 || only == P1) {
      // an order's tile, within what this run's block size allows
      const int t = std::max(1, std::min(tile[P1], max_tile(N, N / P1, B->elt)));
      runs.push_back(BatchRun{P1, N, t, 0, wg, b, e, B->off_host[b], B->foff_host[b]});
      wg += (e - b + t - 1) / t;
    }
    b = e;
  }
  return wg;
}

template <class T, int MODE>
void launch_stage(twoldg_batch* B, const std::vector<BatchRun>& runs, int64_t nwg, const T* ys,
                  T* yo, double coef, int in, double coef2 = 0, T* acc = nullptr) {
  BatchArgs<T> A{};
  for (int r = 0; r < kBatchMaxRuns; ++r) {
    if (r < (int)runs.size()) {
      A.runs[r] = runs[r];
    } else {
      A.runs[r] = runs[0];
      A.runs[r].wg0 = INT64_MAX;
    }
  }
  A.nfo = B->nfo;
  A.nrh = (const T*)B->nrh;
  A.spacing = B->spacing;
  A.dt = B->dt;
  A.tables = (const Tables<T>*)B->tables;
  A.ys = ys;
  A.yn = (const T*)B->yn;
  A.acc = acc ? acc : (T*)B->acc;
  A.yo = yo;
  A.lo = (T*)B->lo;
  A.uf_in = (const T*)B->uf[in];
  A.pf_in = (const T*)B->pf[in];
  A.uf_out = (T*)B->uf[1 - in];
  A.pf_out = (T*)B->pf[1 - in];
  A.coef = coef;
  A.coef2 = coef2;
  A.truncate = B->truncate ? 1 : 0;
  hipLaunchKernelGGL((k_batch2<T, MODE>), dim3((unsigned)nwg), dim3(batch_threads<T>()), 0,
                     B->stream, A);
}

template <class T>
void enqueue_step(twoldg_batch* B) {
  T* yn = (T*)B->yn;
  T* ys = (T*)B->ys;
  if constexpr (std::is_same<T, float>::value) {  // FP32c: the accumulator's three registers
    launch_stage<T, 1>(B, B->runs, B->nwg, yn, ys, 0.5, 0);
    launch_stage<T, 6>(B, B->runs, B->nwg, ys, ys, 0.5, 1);
    launch_stage<T, 2>(B, B->runs, B->nwg, ys, ys, 1.0, 0);
    launch_stage<T, 4>(B, B->runs, B->nwg, ys, yn, 1.0 / 6.0, 1);
  } else {
    // FP64, no accumulator: y_2 in ys, y_3 in acc; stage 3 writes y_4 over y_3
    // and z = (y_2 + 2 y_3)/3 + dt/3 k3 over y_2; the last stage adds dt/6 k4
    T* y3 = (T*)B->acc;
    launch_stage<T, 11>(B, B->runs, B->nwg, yn, ys, 0.5, 0);
    launch_stage<T, 12>(B, B->runs, B->nwg, ys, y3, 0.5, 1);
    launch_stage<T, 13>(B, B->runs, B->nwg, y3, y3, 1.0, 0, 1.0 / 3.0, ys);
    launch_stage<T, 8>(B, B->runs, B->nwg, y3, yn, 1.0 / 6.0, 1, 0.0, ys);
  }
}

// The stage the autotune times: stage 3, which moves the most values (FP32c
// mode 2, FP64 mode 13).
template <class T>
constexpr int kTuneMode = std::is_same<T, float>::value ? 2 : 13;

// Time stage 3 for each order and candidate tile; keep the fastest.
// First each order alone, then each order again inside the full launch with
// the other orders at their chosen tiles: two tiles that time alike alone can
// differ by 5% once every order shares the GPU.
template <class T>
int autotune(twoldg_batch* B, const std::vector<int32_t>& Nh) {
  std::vector<int> orders;
  for (int64_t b = 0; b < B->nb; ++b)
    if (std::find(orders.begin(), orders.end(), B->P1s_host[b]) == orders.end())
      orders.push_back(B->P1s_host[b]);
  constexpr int kReps = 10;
  std::vector<BatchRun> runs;
  for (int P1 : orders) {
    int N = 0;
    for (int64_t b = 0; b < B->nb; ++b)
      if (B->P1s_host[b] == P1) N = std::max<int>(N, Nh[b]);
    const int tmax = max_tile(N, N / P1, sizeof(T));
    float best = 1e30f;
    int besttile = 1;
    for (int t = 1; t <= tmax; t = (t == tmax) ? tmax + 1 : std::min(2 * t, tmax)) {
      int tile[10];
      for (int i = 0; i < 10; ++i) tile[i] = t;
      const int64_t nwg = make_runs(B, Nh, tile, P1, runs);
      // stage 3 writes the accumulator and the stage buffer only
      launch_stage<T, kTuneMode<T>>(B, runs, nwg, (const T*)B->yn, (T*)B->ys, 0.5, 0);
      // the faster of two timings of kReps launches each: one order's launch
      // is short, so a few launches leave the choice to timer noise
      float ms = 1e30f;
      for (int round = 0; round < 2; ++round) {
        HIP_TRY(hipEventRecord(B->ev0, B->stream));
        for (int r = 0; r < kReps; ++r)
          launch_stage<T, kTuneMode<T>>(B, runs, nwg, (const T*)B->yn, (T*)B->ys, 0.5, 0);
        HIP_TRY(hipEventRecord(B->ev1, B->stream));
        HIP_TRY(hipEventSynchronize(B->ev1));
        float m = 0;
        HIP_TRY(hipEventElapsedTime(&m, B->ev0, B->ev1));
        ms = std::min(ms, m);
      }
      if (ms < best) {
        best = ms;
        besttile = t;
      }
    }
    B->tiles[P1] = besttile;
    B->tune_ms[P1] = best / kReps;
  }
  for (int P1 : orders) {
    int N = 0;
    for (int64_t b = 0; b < B->nb; ++b)
      if (B->P1s_host[b] == P1) N = std::max<int>(N, Nh[b]);
    const int tmax = max_tile(N, N / P1, sizeof(T));
    float best = 1e30f;
    int besttile = B->tiles[P1];
    for (int t = 1; t <= tmax; t = (t == tmax) ? tmax + 1 : std::min(2 * t, tmax)) {
      int tile[10];
      for (int i = 0; i < 10; ++i) tile[i] = B->tiles[i];
      tile[P1] = t;
      const int64_t nwg = make_runs(B, Nh, tile, 0, runs);
      launch_stage<T, kTuneMode<T>>(B, runs, nwg, (const T*)B->yn, (T*)B->ys, 0.5, 0);
      float ms = 1e30f;
      for (int round = 0; round < 2; ++round) {
        HIP_TRY(hipEventRecord(B->ev0, B->stream));
        for (int r = 0; r < kReps; ++r)
          launch_stage<T, kTuneMode<T>>(B, runs, nwg, (const T*)B->yn, (T*)B->ys, 0.5, 0);
        HIP_TRY(hipEventRecord(B->ev1, B->stream));
        HIP_TRY(hipEventSynchronize(B->ev1));
        float m = 0;
        HIP_TRY(hipEventElapsedTime(&m, B->ev0, B->ev1));
        ms = std::min(ms, m);
      }
      if (ms < best) {
        best = ms;
        besttile = t;
      }
    }
    B->tiles[P1] = besttile;
    B->tune_full_ms[P1] = best / kReps;
  }
  return 0;
}

template <class T>
int batch_build(twoldg_batch* B, const twoldg_desc* const* d, int64_t n, const double* dt) {
  // order problems by element order and block size (then as given): the blocks
  // of each (order, size) are one run of the block list, so one tile applies
  // to the run
  std::vector<int64_t> order(n);
  for (int64_t i = 0; i < n; ++i) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](int64_t a, int64_t b) {
    return d[a]->P != d[b]->P ? d[a]->P < d[b]->P : d[a]->E < d[b]->E;
  });
  {
    int nruns = 0;
    for (int64_t oi = 0; oi < n; ++oi)
      if (oi == 0 || d[order[oi]]->P != d[order[oi - 1]]->P || d[order[oi]]->E != d[order[oi - 1]]->E)
        ++nruns;
    if (nruns > kBatchMaxRuns)
      return fail(TWOLDG_EINVAL, "a batch takes at most " + std::to_string(kBatchMaxRuns) +
                                     " distinct (P, E) pairs");
  }
  B->nprob = n;
  B->prob_off.assign(n, 0);
  B->prob_len.assign(n, 0);
  std::vector<int64_t> first(n, 0), offh, foffh;
  std::vector<int32_t> Nh;
  std::vector<double> hh, dth;
  int64_t nb = 0, total = 0, fsum = 0;
  for (int64_t oi = 0; oi < n; ++oi) {
    const int64_t i = order[oi];
    const twoldg_desc* p = d[i];
    const int N = p->E * (p->P + 1);
    first[i] = nb;
    B->prob_off[i] = total;
    B->prob_len[i] = p->nblocks * 2 * (int64_t)N * N;
    for (int64_t b = 0; b < p->nblocks; ++b) {
      offh.push_back(total);
      foffh.push_back(fsum);
      total += 2 * (int64_t)N * N;
      fsum += 4 * (int64_t)N;
      B->P1s_host.push_back(p->P + 1);
      Nh.push_back(N);
      hh.push_back(p->spacing[b]);
      dth.push_back(dt[i]);
    }
    nb += p->nblocks;
  }
  if (fsum > INT32_MAX)
    return fail(TWOLDG_EINVAL, "a batch's ghost slots exceed 2^31 values");
  // per block face: the ghost slots of the neighbour's facing face (where the
  // face goes) and the neighbour's reciprocal spacing (-1 and 0 for a mirror)
  std::vector<int32_t> nfoh;
  std::vector<T> nrhh;
  for (int64_t oi = 0; oi < n; ++oi) {
    const int64_t i = order[oi];
    const twoldg_desc* p = d[i];
    const int N = p->E * (p->P + 1);
    for (int64_t b = 0; b < p->nblocks; ++b)
      for (int f = 0; f < 4; ++f) {
        const twoldg_face& F = p->faces[b * 4 + f];
        if (F.kind == TWOLDG_FACE_SAME) {
          const int64_t nb_ = first[i] + F.nb[0];
          nfoh.push_back((int32_t)(foffh[nb_] + (int64_t)(f ^ 1) * N));
          nrhh.push_back((T)(1.0 / p->spacing[F.nb[0]]));
        } else {
          nfoh.push_back(-1);
          nrhh.push_back(T(0));
        }
      }
  }
  B->nb = nb;
  B->total = total;
  B->nfaces = fsum;
  int rc;
  B->off_host = offh;
  B->foff_host = foffh;
  if ((rc = balloc(B, (void**)&B->off, nb * 8)) || (rc = balloc(B, (void**)&B->foff, nb * 8)) ||
      (rc = balloc(B, (void**)&B->nfo, nb * 16)) || (rc = balloc(B, &B->nrh, nb * 4 * sizeof(T))) ||
      (rc = balloc(B, (void**)&B->P1s, nb * 4)) ||
      (rc = balloc(B, (void**)&B->Ns, nb * 4)) || (rc = balloc(B, (void**)&B->spacing, nb * 8)) ||
      (rc = balloc(B, (void**)&B->dt, nb * 8)))
    return rc;
  if ((rc = bh2d(B, B->off, offh.data(), nb * 8))) return rc;
  if ((rc = bh2d(B, B->foff, foffh.data(), nb * 8))) return rc;
  if ((rc = bh2d(B, B->nfo, nfoh.data(), nb * 16))) return rc;
  if ((rc = bh2d(B, B->nrh, nrhh.data(), nb * 4 * sizeof(T)))) return rc;
  if ((rc = bh2d(B, B->P1s, B->P1s_host.data(), nb * 4))) return rc;
  if ((rc = bh2d(B, B->Ns, Nh.data(), nb * 4))) return rc;
  if ((rc = bh2d(B, B->spacing, hh.data(), nb * 8))) return rc;
  if ((rc = bh2d(B, B->dt, dth.data(), nb * 8))) return rc;
  if ((rc = batch_tables<T>(B))) return rc;
  if (!B->borrowed) {
    for (void** p : {&B->yn, &B->ys, &B->acc})
      if ((rc = balloc(B, p, total * sizeof(T)))) return rc;
    if (B->precision == TWOLDG_F32C && (rc = balloc(B, &B->lo, total * sizeof(T)))) return rc;
    for (void* p : {B->yn, B->ys, B->acc, B->lo})
      if (p) HIP_TRY(hipMemsetAsync(p, 0, total * sizeof(T), B->stream));
  }
  for (int k = 0; k < 2; ++k)
    if ((rc = balloc(B, &B->uf[k], fsum * sizeof(T))) || (rc = balloc(B, &B->pf[k], fsum * sizeof(T))))
      return rc;
  for (void* p : {B->uf[0], B->uf[1], B->pf[0], B->pf[1]})
    HIP_TRY(hipMemsetAsync(p, 0, fsum * sizeof(T), B->stream));
  // tiles per order (timed; a problem's own batch of one takes the largest
  // tile that still leaves two workgroups per CU, so a small problem spreads
  // over the chip), then the work list, then the step graph
  if (B->borrowed) {
    const int64_t nb1 = B->nb;
    const int P1 = B->P1s_host[0], N = Nh[0];
    B->tiles[P1] = (int)std::max<int64_t>(1, std::min<int64_t>(max_tile(N, N / P1, sizeof(T)),
                                                               nb1 / 512));
  } else if ((rc = autotune<T>(B, Nh))) {
    return rc;
  }
  B->nwg = make_runs(B, Nh, B->tiles, 0, B->runs);
  HIP_TRY(hipStreamSynchronize(B->stream));
  HIP_TRY(hipStreamBeginCapture(B->stream, hipStreamCaptureModeThreadLocal));
  enqueue_step<T>(B);
  HIP_TRY(hipStreamEndCapture(B->stream, &B->graph));
  HIP_TRY(hipGraphInstantiate(&B->exec, B->graph, nullptr, nullptr, 0));
  return 0;
}

void batch_destroy(twoldg_batch* B) {
  if (!B) return;
  if (B->stream) (void)hipStreamSynchronize(B->stream);
  if (B->exec) (void)hipGraphExecDestroy(B->exec);
  if (B->graph) (void)hipGraphDestroy(B->graph);
  for (void* p : {(void*)B->off, (void*)B->foff, (void*)B->nfo, B->nrh, (void*)B->P1s, (void*)B->Ns,
                  (void*)B->spacing, (void*)B->dt, B->tables, B->uf[0], B->uf[1], B->pf[0], B->pf[1]})
    if (p) pool_free(p);
  if (!B->borrowed)
    for (void* p : {B->yn, B->ys, B->acc, B->lo})
      if (p) pool_free(p);
  if (B->ev0) (void)hipEventDestroy(B->ev0);
  if (B->ev1) (void)hipEventDestroy(B->ev1);
  if (B->stream && !B->borrowed) (void)hipStreamDestroy(B->stream);
  delete B;
}

template <class T>
int batch_steps(twoldg_batch* B, int64_t nsteps) {
  hipLaunchKernelGGL((k_batch_faces<T>), dim3((unsigned)B->nb), dim3(64), 0, B->stream, B->nb,
                     B->P1s, B->Ns, B->off, B->nfo, (const Tables<T>*)B->tables,
                     (const T*)B->yn, (T*)B->uf[0], (T*)B->pf[0]);
  for (int64_t s = 0; s < nsteps; ++s) HIP_TRY(hipGraphLaunch(B->exec, B->stream));
  HIP_TRY(hipGetLastError());
  return 0;
}

}  // namespace

namespace twoldg_detail {

// Whether a description is one a batch takes; why not, if `why`.
bool batch_desc_ok(const twoldg_desc* d, std::string* why) {
  auto no = [&](const std::string& w) {
    if (why) *why = w;
    return false;
  };
  if (!d) return no("no description");
  if (d->abi_version != TWOLDG_ABI_VERSION)
    return no("ABI version " + std::to_string(d->abi_version) + " does not match the library's " +
              std::to_string(TWOLDG_ABI_VERSION));
  if (d->D != 2 || d->coef || d->dcoef) return no("a batch takes flat two-dimensional problems");
  if (d->precision != TWOLDG_F64 && d->precision != TWOLDG_F32C)
    return no("unknown precision " + std::to_string(d->precision));
  if (d->P < 1 || d->P > 7 || !batch_e_ok(d->E)) return no("P must be 1..7 and E 2 or 4");
  if (d->nblocks < 1 || !d->spacing || !d->faces)
    return no("a problem needs at least one block, its spacings and faces");
  const int N = d->E * (d->P + 1);
  if (max_tile(N, d->E, d->precision == TWOLDG_F32C ? 4 : 8) < 1)
    return no("blocks of " + std::to_string(N) + "^2 nodes exceed a workgroup");
  for (int64_t b = 0; b < d->nblocks; ++b)
    if (!(d->spacing[b] > 0)) return no("block spacing must be positive");
  for (int64_t f = 0; f < d->nblocks * 4; ++f) {
    const int k = d->faces[f].kind;
    if (k != TWOLDG_FACE_SAME && k != TWOLDG_FACE_MIRROR)
      return no("a batch takes uniform problems (same-level or mirror faces)");
    if (k == TWOLDG_FACE_SAME && (d->faces[f].nb[0] < 0 || d->faces[f].nb[0] >= d->nblocks))
      return no("face neighbour out of range");
  }
  // a block's faces go into its neighbours' ghost slots, so the neighbour
  // across face 2d + s must face back across 2d + 1 - s
  for (int64_t f = 0; f < d->nblocks * 4; ++f) {
    if (d->faces[f].kind != TWOLDG_FACE_SAME) continue;
    const twoldg_face& back = d->faces[d->faces[f].nb[0] * 4 + ((f % 4) ^ 1)];
    if (back.kind != TWOLDG_FACE_SAME || back.nb[0] != f / 4)
      return no("block " + std::to_string(f / 4) + " face " + std::to_string(f % 4) +
                ": its neighbour does not face back");
  }
  return true;
}

// A two-dimensional problem's own batch of one: the batch stage (one launch
// per RK4 stage, a graph per step) on the problem's state (yn, with its
// compensation lo in FP32C), stage and accumulator buffers, on its stream.
int batch_attach(const twoldg_desc* d, void* yn, void* ys, void* acc, void* lo,
                 hipStream_t stream, twoldg_batch** out) {
  *out = nullptr;
  if (!batch_desc_ok(d, nullptr)) return fail(TWOLDG_EINVAL, "not a batch problem");
  auto* B = new twoldg_batch();
  B->borrowed = true;
  B->precision = d->precision;
  B->truncate = d->truncate != 0;
  B->elt = d->precision == TWOLDG_F64 ? 8 : 4;
  B->stream = stream;
  B->yn = yn;
  B->ys = ys;
  B->acc = acc;
  B->lo = lo;
  (void)hipEventCreate(&B->ev0);
  (void)hipEventCreate(&B->ev1);
  const double dt1 = 1.0;  // the table is written before the first steps
  const twoldg_desc* const ds[1] = {d};
  int rc = d->precision == TWOLDG_F64 ? batch_build<double>(B, ds, 1, &dt1)
                                      : batch_build<float>(B, ds, 1, &dt1);
  if (rc) {
    batch_destroy(B);
    return rc;
  }
  B->cur_dt = NAN;
  *out = B;
  return 0;
}

void batch_detach(twoldg_batch* B) { batch_destroy(B); }

void batch_work(const twoldg_batch* B, std::vector<std::pair<void*, size_t>>& out) {
  for (int k = 0; k < 2; ++k) {
    out.emplace_back(B->uf[k], (size_t)B->nfaces * B->elt);
    out.emplace_back(B->pf[k], (size_t)B->nfaces * B->elt);
  }
}

// nsteps RK4 steps of the attached problem with time step dt.
int batch_attached_steps(twoldg_batch* B, double dt, int64_t nsteps) {
  if (nsteps <= 0) return 0;
  if (!(dt == B->cur_dt)) {
    std::vector<double> dts((size_t)B->nb, dt);
    int rc = bh2d(B, B->dt, dts.data(), dts.size() * sizeof(double));
    if (rc) return rc;
    B->cur_dt = dt;
  }
  return B->precision == TWOLDG_F64 ? batch_steps<double>(B, nsteps) : batch_steps<float>(B, nsteps);
}

}  // namespace twoldg_detail
using twoldg_detail::batch_desc_ok;

extern "C" {

int twoldg_batch_create(const twoldg_desc* const* descs, int64_t n, const double* dt,
                        twoldg_batch** out) {
  if (!out) return fail(TWOLDG_EINVAL, "no output handle");
  *out = nullptr;
  if (!descs || n < 1 || !dt) return fail(TWOLDG_EINVAL, "a batch needs problems and time steps");
  for (int64_t i = 0; i < n; ++i) {
    const twoldg_desc* d = descs[i];
    const std::string who = "problem " + std::to_string(i) + ": ";
    std::string why;
    if (!batch_desc_ok(d, &why)) return fail(TWOLDG_EINVAL, who + why);
    if (d->precision != descs[0]->precision || d->truncate != descs[0]->truncate)
      return fail(TWOLDG_EINVAL, "every problem of a batch shares its precision and truncation");
    if (!(dt[i] > 0)) return fail(TWOLDG_EINVAL, who + "time steps must be positive");
  }
  const int prec = descs[0]->precision;
  const int trunc = descs[0]->truncate;
  auto* B = new twoldg_batch();
  B->precision = prec;
  B->truncate = trunc != 0;
  B->elt = prec == TWOLDG_F64 ? 8 : 4;
  hipError_t e = hipStreamCreate(&B->stream);
  if (e != hipSuccess) {
    delete B;
    return fail(TWOLDG_EDEVICE, std::string("hipStreamCreate: ") + hipGetErrorString(e));
  }
  (void)hipEventCreate(&B->ev0);
  (void)hipEventCreate(&B->ev1);
  int rc = prec == TWOLDG_F64 ? batch_build<double>(B, descs, n, dt)
                              : batch_build<float>(B, descs, n, dt);
  if (rc) {
    batch_destroy(B);
    return rc;
  }
  *out = B;
  return 0;
}

int twoldg_batch_destroy(twoldg_batch* B) {
  batch_destroy(B);
  return 0;
}

int twoldg_batch_upload(twoldg_batch* B, int64_t i, const double* host) {
  if (!B || !host || i < 0 || i >= B->nprob) return fail(TWOLDG_EINVAL, "bad batch upload");
  const int64_t n = B->prob_len[i], o = B->prob_off[i];
  // staged in fine-grained host memory and converted by a kernel that reads it there
  double* tmp = nullptr;
  HIP_TRY(host_stage_alloc((void**)&tmp, n * 8));
  std::memcpy(tmp, host, n * 8);
  if (B->precision == TWOLDG_F64)
    hipLaunchKernelGGL((k_from_double<double>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                       B->stream, n, (const double*)tmp, (double*)B->yn + o, (double*)nullptr);
  else
    hipLaunchKernelGGL((k_from_double<float>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                       B->stream, n, (const double*)tmp, (float*)B->yn + o,
                       B->lo ? (float*)B->lo + o : nullptr);
  hipError_t e = hipStreamSynchronize(B->stream);
  pool_free(tmp);
  if (e != hipSuccess) return fail(TWOLDG_EDEVICE, std::string("batch upload: ") + hipGetErrorString(e));
  return 0;
}

int twoldg_batch_download(twoldg_batch* B, int64_t i, double* host) {
  if (!B || !host || i < 0 || i >= B->nprob) return fail(TWOLDG_EINVAL, "bad batch download");
  const int64_t n = B->prob_len[i], o = B->prob_off[i];
  // converted by a kernel that writes fine-grained host staging directly
  double* tmp = nullptr;
  HIP_TRY(host_stage_alloc((void**)&tmp, n * 8));
  if (B->precision == TWOLDG_F64)
    hipLaunchKernelGGL((k_to_double<double>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                       B->stream, n, (const double*)B->yn + o, (const double*)nullptr, tmp);
  else
    hipLaunchKernelGGL((k_to_double<float>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                       B->stream, n, (const float*)B->yn + o,
                       B->lo ? (const float*)B->lo + o : nullptr, tmp);
  hipError_t e = hipStreamSynchronize(B->stream);
  if (e == hipSuccess) std::memcpy(host, tmp, n * 8);
  pool_free(tmp);
  if (e != hipSuccess) return fail(TWOLDG_EDEVICE, std::string("batch download: ") + hipGetErrorString(e));
  return 0;
}

int twoldg_batch_rk4(twoldg_batch* B, int64_t nsteps, double* out_ms) {
  if (!B || nsteps < 0) return fail(TWOLDG_EINVAL, "bad batch step call");
  HIP_TRY(hipEventRecord(B->ev0, B->stream));
  int rc = B->precision == TWOLDG_F64 ? batch_steps<double>(B, nsteps) : batch_steps<float>(B, nsteps);
  if (rc) return rc;
  HIP_TRY(hipEventRecord(B->ev1, B->stream));
  HIP_TRY(hipEventSynchronize(B->ev1));
  float ms = 0;
  HIP_TRY(hipEventElapsedTime(&ms, B->ev0, B->ev1));
  if (out_ms) *out_ms = ms;
  return 0;
}

int twoldg_batch_info(const twoldg_batch* B, int64_t* nodes, int64_t* workgroups, int32_t* tiles,
                      double* tune_ms) {
  if (!B) return fail(TWOLDG_EINVAL, "no batch");
  if (nodes) *nodes = B->total / 2;
  if (workgroups) *workgroups = B->nwg;
  for (int P = 1; P <= 8; ++P) {
    if (tiles) tiles[P - 1] = B->tiles[P + 1];
    if (tune_ms) {
      tune_ms[P - 1] = B->tune_ms[P + 1];
      tune_ms[8 + P - 1] = B->tune_full_ms[P + 1];
    }
  }
  return 0;
}

}  // extern "C"
