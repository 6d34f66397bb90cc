// Host side of the batched path: builds every box once through the plain
// path (initial energies, structure factors, COMs), packs all boxes into
// flat HBM arrays, and launches the persistent MC kernel.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>

#include "batch.hpp"
#include "box.hpp"

using namespace gomc;

namespace {

int BFail(gomc_batch *bt, int code, const std::string &msg) {
  if (bt) bt->err = msg;
  return code;
}

#define BHIP(bt, call)                                                                   \
  do {                                                                                   \
    hipError_t e_ = (call);                                                              \
    if (e_ != hipSuccess) return BFail(bt, GOMC_ERR_HIP, std::string(#call ": ") + hipGetErrorString(e_)); \
  } while (0)

void FreeBatch(gomc_batch *bt) {
  void *ptrs[] = {bt->boxes, bt->state, bt->x,    bt->y,    bt->z,    bt->cx,         bt->cy,
                  bt->cz,    bt->kn,    bt->pref, bt->sumR, bt->sumI, bt->dR, bt->dI, bt->moves,      bt->trace.acc,
                  bt->trace.dlj, bt->trace.dreal, bt->trace.drecip, bt->tasks, bt->wgTask,
                  bt->xchg, bt->errFlag, bt->boxTicks};
  for (void *p : ptrs)
    if (p) (void)hipFree(p);
  if (bt->graphExec) (void)hipGraphExecDestroy(bt->graphExec);
  if (bt->graph) (void)hipGraphDestroy(bt->graph);
  if (bt->ev0) (void)hipEventDestroy(bt->ev0);
  if (bt->ev1) (void)hipEventDestroy(bt->ev1);
  if (bt->stream) (void)hipStreamDestroy(bt->stream);
}

}  // namespace

extern "C" {

const char *gomc_batch_last_error(const gomc_batch *bt) { return bt ? bt->err.c_str() : "null batch"; }

int gomc_batch_destroy(gomc_batch *bt) {
  if (!bt) return GOMC_ERR_ARG;
  (void)hipSetDevice(bt->device);
  FreeBatch(bt);
  delete bt;
  return GOMC_OK;
}

}  // extern "C"

namespace gomc {

int PackBox(gomc_batch *bt, const gomc_system_desc &d, int64_t atomOff, int64_t molOff,
            int64_t kOff, PackedBox &out) {
  if (d.n_kinds != 1) return BFail(bt, GOMC_ERR_ARG, "batched boxes hold one molecule kind");
  if (d.ensemble != 0) return BFail(bt, GOMC_ERR_ARG, "batched boxes are NVT");
  const int len = d.kind_natoms[0];
  if (len <= 0 || len > kBatchMaxAtoms) return BFail(bt, GOMC_ERR_ARG, "molecule size out of range");
  if (d.n_atom_types > kBatchMaxTypes) return BFail(bt, GOMC_ERR_ARG, "too many atom types");
  gomc_box *b = nullptr;
  int rc = gomc_box_create(&d, bt->device, &b);
  if (rc) return BFail(bt, rc, "box setup failed");
  struct Guard {
    gomc_box *b;
    ~Guard() { gomc_box_destroy(b); }
  } guard{b};
  BatchBox &P = out.P;
  std::memset(&P, 0, sizeof(P));
  P.team = 1;
  for (int a = 0; a < 3; ++a) {
    P.L[a] = b->geom.L[a];
    P.half[a] = b->geom.half[a];
    P.kc[a] = (1.0 / b->geom.L[a]) * (2.0 * M_PI);
    P.nmax[a] = int(b->recipRcut * b->geom.L[a] / (2.0 * M_PI)) + 1;
    if (P.nmax[a] > kBatchMaxN) return BFail(bt, GOMC_ERR_ARG, "box too large for the batched phase tables");
  }
  // largest atom-to-centre distance of the kind (its first molecule), for
  // the centre cull of the pair phase
  double rmax = 0.0;
  for (int p = 0; p < len; ++p) {
    double d2 = 0.0;
    const double *P3[3] = {b->hx.data(), b->hy.data(), b->hz.data()};
    const double c3[3] = {b->comx[0], b->comy[0], b->comz[0]};
    for (int a = 0; a < 3; ++a) {
      double v = P3[a][p] - c3[a];
      if (v > b->geom.half[a]) v -= b->geom.L[a];
      else if (v < -b->geom.half[a]) v += b->geom.L[a];
      d2 += v * v;
    }
    rmax = std::max(rmax, std::sqrt(d2));
  }
  const double nearCut = std::sqrt(b->geom.rCutSq) + 2.0 * rmax + 1e-6;
  P.nearCutSq = nearCut * nearCut;
  // the cull pays when the near sphere is well inside the box
  P.cull = (b->nmols <= kBatchNearCap &&
            (4.0 / 3.0) * M_PI * nearCut * nearCut * nearCut < 0.5 * b->volume) ? 1 : 0;
  {
    // Waves for the pair phase; the rest run the reciprocal phase beside
    // it.  Measured on MI355X (water, 200-4000 molecules): 12 pair waves
    // are fastest while a single workgroup holds at most 2048 k-vectors
    // (8 per reciprocal thread), 8 beyond that and in teams.
    const int T = bt->teams ? TeamSize(bt, b->nmols) : 1;
    const int kper = b->ewald ? b->K : 0;
    P.pairWaves = (T == 1 && kper <= bt->tune.pair12_max_k) ? 12 : 8;
    P.kWidth = kBatchThreads - kWave * P.pairWaves;
  }
  P.rCutSq = b->geom.rCutSq;
  P.rCutLJSq = b->geom.rCutLJSq;
  P.rCutCoulSq = b->geom.rCutCoulSq;
  P.rCutLowSq = b->geom.rCutLowSq;
  P.alpha = b->alpha;
  P.beta = b->beta;
  P.pDisp = b->movePerc[0] / (b->movePerc[0] + b->movePerc[1]);
  for (int p = 0; p < len; ++p) {
    P.charge[p] = b->kindCharge[0][p];
    for (int q = 0; q < len; ++q) {
      const int idx = b->kindType[0][p] + b->kindType[0][q] * b->ntypes;
      const int pq = p * kBatchMaxAtoms + q;
      P.sigSq[pq] = b->sigSq[idx];
      P.epsCn[pq] = b->epsCn[idx];
      P.nHalf[pq] = b->nHalf[idx];
      if (b->epsCn[idx] != 0.0 && b->nHalf[idx] != 6.0)
        return BFail(bt, GOMC_ERR_ARG, "batched boxes use the 12-6 Lennard-Jones form");
      P.qq[pq] = b->kindCharge[0][p] * b->kindCharge[0][q] * kQQFact;
      if (P.qq[pq] != 0.0 || P.epsCn[pq] != 0.0) P.live |= 1 << pq;
    }
  }
  P.atomOff = atomOff;
  P.molOff = molOff;
  P.kOff = kOff;
  P.perAdjust = int64_t(b->perAdjust);
  P.seed = d.seed;
  P.natoms = b->natoms;
  P.nmols = b->nmols;
  P.molLen = len;
  // exact floor division by len and len^2 for item indices < 2^32 / len^2
  P.magicLen = ((uint64_t(1) << 32) + uint64_t(len) - 1) / uint64_t(len);
  P.magicLen2 = ((uint64_t(1) << 32) + uint64_t(len * len) - 1) / uint64_t(len * len);
  if (uint64_t(b->nmols) * len * len >= (uint64_t(1) << 32) / uint64_t(len * len))
    return BFail(bt, GOMC_ERR_ARG, "box too large for the batched item indexing");
  P.ewald = b->ewald;
  P.elect = b->elect;
  out.x = b->hx;
  out.y = b->hy;
  out.z = b->hz;
  out.cx = b->comx;
  out.cy = b->comy;
  out.cz = b->comz;
  out.pref.clear();
  out.kn.clear();
  out.sr.clear();
  out.si.clear();
  if (b->ewald) {
    std::vector<double> kx, ky, kz, pref;
    std::vector<int32_t> packed;
    KVectorsOrth(b->geom.L, b->alpha * b->alpha, b->recipRcut, b->volume, kx, ky, kz, pref, &packed);
    std::vector<double> sr(b->K), si(b->K);
    rc = gomc_recip_get(b, nullptr, nullptr, nullptr, nullptr, sr.data(), si.data());
    if (rc || int(pref.size()) != b->K) return BFail(bt, GOMC_ERR_STATE, "k-vector setup mismatch");
    // Segments: each run of consecutive nz at fixed (nx, ny) (contiguous
    // in the reference order) is cut into pieces of at most S k-vectors;
    // column t of the W-wide layout takes segment r * W + t in round r, its
    // step j at slot (r * S + j) * W + t.  Unused slots carry kn -1.
    std::vector<int> runStart, runLen;
    for (int k = 0; k < b->K; ++k) {
      if (k > 0 && packed[k] == packed[k - 1] + (1 << 16)) {
        ++runLen.back();
      } else {
        runStart.push_back(k);
        runLen.push_back(1);
      }
    }
    const int W = P.kWidth;
    // per round a thread pays about three table products to start a
    // segment and one complex product per step; a team of T splits the
    // rounds
    const int64_t T = bt->teams ? TeamSize(bt, b->nmols) : 1;
    int bestS = 1;
    int64_t bestCost = -1;
    for (int S : {1, 2, 4, 8}) {
      int64_t segs = 0;
      for (int L : runLen) segs += (L + S - 1) / S;
      const int64_t rounds = (segs + W - 1) / W;
      const int64_t cost = ((rounds + T - 1) / T) * (S + 3);
      if (bestCost < 0 || cost < bestCost) {
        bestCost = cost;
        bestS = S;
      }
    }
    const int S = bestS;
    int64_t segs = 0;
    for (int L : runLen) segs += (L + S - 1) / S;
    const int64_t rounds = (segs + W - 1) / W;
    if (rounds * S * W > (int64_t(1) << 30))
      return BFail(bt, GOMC_ERR_ARG, "too many k-vectors for the batched layout");
    P.kSeg = S;
    P.kChunk = int(rounds);
    P.K = int(rounds * S * W);
    out.pref.assign(P.K, 0.0);
    out.kn.assign(P.K, -1);
    out.sr.assign(P.K, 0.0);
    out.si.assign(P.K, 0.0);
    int64_t seg = 0;
    for (size_t r = 0; r < runLen.size(); ++r) {
      for (int off = 0; off < runLen[r]; off += S, ++seg) {
        const int64_t round = seg / W, t = seg % W;
        for (int j = 0; j < S && off + j < runLen[r]; ++j) {
          const int k = runStart[r] + off + j;
          const int64_t slot = (round * S + j) * W + t;
          out.pref[slot] = pref[k];
          out.kn[slot] = packed[k];
          out.sr[slot] = sr[k];
          out.si[slot] = si[k];
        }
      }
    }
  }
  BatchState &S = out.S;
  std::memset(&S, 0, sizeof(S));
  S.energy[0] = b->pot.inter;
  S.energy[1] = b->pot.tc;
  S.energy[2] = b->pot.real;
  S.energy[3] = b->pot.recip;
  S.energy[4] = b->pot.self;
  S.energy[5] = b->pot.corr;
  S.energy[6] = b->pot.total;
  S.scale[0] = b->moves.scale[0][0];
  S.scale[1] = b->moves.scale[1][0];
  return GOMC_OK;
}

BatchArgs MakeArgs(const gomc_batch *bt) {
  BatchArgs a;
  a.boxes = bt->boxes;
  a.state = bt->state;
  a.x = bt->x;
  a.y = bt->y;
  a.z = bt->z;
  a.cx = bt->cx;
  a.cy = bt->cy;
  a.cz = bt->cz;
  a.kn = bt->kn;
  a.pref = bt->pref;
  a.sumR = bt->sumR;
  a.sumI = bt->sumI;
  a.dR = bt->dR;
  a.dI = bt->dI;
  a.moves = bt->moves;
  a.tasks = bt->tasks;
  a.wgTask = bt->wgTask;
  a.xchg = bt->xchg;
  a.err = bt->errFlag;
  a.boxTicks = bt->boxTicks;
  a.ldsMols = bt->ldsMols;
  a.trace = bt->trace;
  return a;
}

int AllocBatch(gomc_batch *bt, int nboxes, int64_t atoms, int64_t mols, int64_t kvec) {
  BHIP(bt, hipSetDevice(bt->device));
  auto alloc = [&](void **p, size_t bytes) { return hipMalloc(p, std::max<size_t>(bytes, 8)) == hipSuccess; };
  bool ok = hipStreamCreateWithFlags(&bt->stream, hipStreamNonBlocking) == hipSuccess;
  ok &= hipEventCreate(&bt->ev0) == hipSuccess;
  ok &= hipEventCreate(&bt->ev1) == hipSuccess;
  ok &= alloc((void **)&bt->boxes, sizeof(BatchBox) * nboxes);
  ok &= alloc((void **)&bt->state, sizeof(BatchState) * nboxes);
  ok &= alloc((void **)&bt->moves, sizeof(int64_t) * nboxes);
  ok &= alloc((void **)&bt->x, sizeof(double) * atoms);
  ok &= alloc((void **)&bt->y, sizeof(double) * atoms);
  ok &= alloc((void **)&bt->z, sizeof(double) * atoms);
  ok &= alloc((void **)&bt->cx, sizeof(double) * mols);
  ok &= alloc((void **)&bt->cy, sizeof(double) * mols);
  ok &= alloc((void **)&bt->cz, sizeof(double) * mols);
  ok &= alloc((void **)&bt->kn, sizeof(int32_t) * kvec);
  ok &= alloc((void **)&bt->pref, sizeof(double) * kvec);
  ok &= alloc((void **)&bt->sumR, sizeof(double) * kvec);
  ok &= alloc((void **)&bt->sumI, sizeof(double) * kvec);
  ok &= alloc((void **)&bt->dR, sizeof(double) * kvec);
  ok &= alloc((void **)&bt->dI, sizeof(double) * kvec);
  // one workgroup per compute unit: the kernel's 1024 threads at 128
  // registers fill a CU, and team members must all be resident at once
  int cus = 0;
  ok &= hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, bt->device) == hipSuccess;
  bt->gridCap = std::max(8, cus - cus % 8);
  bt->grid = bt->gridCap;
  ok &= alloc((void **)&bt->tasks, sizeof(BatchTask) * size_t(nboxes) * kTeamMax);
  ok &= alloc((void **)&bt->wgTask, sizeof(int) * size_t(bt->gridCap + 1));
  ok &= alloc((void **)&bt->xchg, sizeof(uint64_t) * size_t(nboxes) * kXchgWords);
  ok &= alloc((void **)&bt->errFlag, sizeof(int));
  ok &= alloc((void **)&bt->boxTicks, sizeof(uint64_t) * size_t(nboxes));
  if (!ok) return BFail(bt, GOMC_ERR_HIP, "batch allocation failed");
  BHIP(bt, hipMemsetAsync(bt->xchg, 0, sizeof(uint64_t) * size_t(nboxes) * kXchgWords, bt->stream));
  BHIP(bt, hipMemsetAsync(bt->errFlag, 0, sizeof(int), bt->stream));
  // the structure-factor changes start defined (unused k slots are read)
  if (kvec > 0) {
    BHIP(bt, hipMemsetAsync(bt->dR, 0, sizeof(double) * kvec, bt->stream));
    BHIP(bt, hipMemsetAsync(bt->dI, 0, sizeof(double) * kvec, bt->stream));
  }
  bt->totalAtoms = atoms;
  bt->totalMols = mols;
  bt->totalK = kvec;
  return GOMC_OK;
}

int UploadBox(gomc_batch *bt, int slot, const PackedBox &pb) {
  const BatchBox &P = pb.P;
  hipStream_t s = bt->stream;
  auto up = [&](void *dst, const void *src, size_t bytes) -> hipError_t {
    return bytes ? hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s) : hipSuccess;
  };
  BHIP(bt, up(bt->boxes + slot, &pb.P, sizeof(BatchBox)));
  BHIP(bt, up(bt->state + slot, &pb.S, sizeof(BatchState)));
  {
    const int64_t ao = P.atomOff, mo = P.molOff;
    BHIP(bt, up(bt->x + ao, pb.x.data(), sizeof(double) * pb.x.size()));
    BHIP(bt, up(bt->y + ao, pb.y.data(), sizeof(double) * pb.y.size()));
    BHIP(bt, up(bt->z + ao, pb.z.data(), sizeof(double) * pb.z.size()));
    BHIP(bt, up(bt->cx + mo, pb.cx.data(), sizeof(double) * pb.cx.size()));
    BHIP(bt, up(bt->cy + mo, pb.cy.data(), sizeof(double) * pb.cy.size()));
    BHIP(bt, up(bt->cz + mo, pb.cz.data(), sizeof(double) * pb.cz.size()));
  }
  BHIP(bt, up(bt->kn + P.kOff, pb.kn.data(), sizeof(int32_t) * pb.kn.size()));
  BHIP(bt, up(bt->pref + P.kOff, pb.pref.data(), sizeof(double) * pb.pref.size()));
  BHIP(bt, up(bt->sumR + P.kOff, pb.sr.data(), sizeof(double) * pb.sr.size()));
  BHIP(bt, up(bt->sumI + P.kOff, pb.si.data(), sizeof(double) * pb.si.size()));
  BHIP(bt, hipStreamSynchronize(s));
  bt->hboxes[slot] = pb.P;
  return GOMC_OK;
}

int TeamSize(const gomc_batch *bt, int nmols) {
  // Per-move cost of one workgroup grows about 10 us + 6 ns per molecule
  // (measured on MI355X, water); a team of T repeats ~7 us of it per member
  // (draws, trial coordinates, tables, the exchange) and splits the rest.
  // The default thresholds {1800, 2700, 3500} keep a 256-box cycle near 49 ms.
  const int32_t *t = bt->tune.team_min;
  if (t[2] > 0 && nmols >= t[2]) return 8;
  if (t[1] > 0 && nmols >= t[1]) return 4;
  if (t[0] > 0 && nmols >= t[0]) return 2;
  return 1;
}

int BuildTasks(gomc_batch *bt, const std::vector<int64_t> &moves) {
  const int nb = bt->nboxes, grid = bt->grid, cols = 8, perCol = grid / cols;
  std::vector<std::vector<BatchTask>> lists(grid);
  // teams first, largest first; members of a team share one column of
  // workgroups (index = 8 k + column), which the hardware places on one XCD
  std::vector<int> order;
  for (int b = 0; b < nb; ++b)
    if (bt->teams && bt->hboxes[b].team > 1) order.push_back(b);
  std::stable_sort(order.begin(), order.end(),
                   [&](int p, int q) { return bt->hboxes[p].team > bt->hboxes[q].team; });
  std::vector<int> used(cols, 0);
  for (int b : order) {
    const int T = bt->hboxes[b].team;
    int best = 0;
    for (int c = 1; c < cols; ++c)
      if (used[c] < used[best]) best = c;
    if (perCol - used[best] < T) return BFail(bt, GOMC_ERR_STATE, "team does not fit the device");
    for (int r = 0; r < T; ++r) lists[size_t(8) * (used[best] + r) + best].push_back(BatchTask{b, r});
    used[best] += T;
  }
  // single-workgroup boxes: longest first onto the least-loaded free
  // workgroup
  std::vector<int> solo;
  for (int b = 0; b < nb; ++b)
    if (!bt->teams || bt->hboxes[b].team <= 1) solo.push_back(b);
  // per-move cost: measured in the last launch when there was one, else the
  // model 10 us + 6 ns per molecule
  auto cost = [&](int b) {
    const double us = (b < int(bt->usPerMove.size()) && bt->usPerMove[b] > 0.0)
                          ? bt->usPerMove[b]
                          : 10.0 + 0.006 * bt->hboxes[b].nmols;
    return double(moves[b]) * us;
  };
  std::stable_sort(solo.begin(), solo.end(), [&](int p, int q) { return cost(p) > cost(q); });
  std::vector<int> freeWg;
  for (int g = 0; g < grid; ++g)
    if (lists[g].empty()) freeWg.push_back(g);
  if (!solo.empty() && freeWg.empty()) return BFail(bt, GOMC_ERR_STATE, "no workgroup left for single boxes");
  std::vector<double> load(freeWg.size(), 0.0);
  for (int b : solo) {
    size_t best = 0;
    for (size_t i = 1; i < load.size(); ++i)
      if (load[i] < load[best]) best = i;
    lists[freeWg[best]].push_back(BatchTask{b, 0});
    load[best] += cost(b);
  }
  std::vector<BatchTask> flat;
  std::vector<int> start(grid + 1, 0);
  for (int g = 0; g < grid; ++g) {
    start[g] = int(flat.size());
    flat.insert(flat.end(), lists[g].begin(), lists[g].end());
  }
  start[grid] = int(flat.size());
  if (flat.size() > size_t(nb) * kTeamMax) return BFail(bt, GOMC_ERR_STATE, "task list overflow");
  if (!flat.empty())
    BHIP(bt, hipMemcpyAsync(bt->tasks, flat.data(), sizeof(BatchTask) * flat.size(),
                            hipMemcpyHostToDevice, bt->stream));
  BHIP(bt, hipMemcpyAsync(bt->wgTask, start.data(), sizeof(int) * start.size(), hipMemcpyHostToDevice,
                          bt->stream));
  BHIP(bt, hipStreamSynchronize(bt->stream));
  return GOMC_OK;
}

// After a launch: a team member that waited out its limit set the flag;
// the boxes' times become the cost model of the next task lists.
int CheckTeams(gomc_batch *bt, const std::vector<int64_t> &moves) {
  int flag = 0;
  BHIP(bt, hipMemcpy(&flag, bt->errFlag, sizeof(int), hipMemcpyDeviceToHost));
  if (flag) return BFail(bt, GOMC_ERR_STATE, "a team member timed out waiting for its partners");
  std::vector<uint64_t> ticks(bt->nboxes);
  BHIP(bt, hipMemcpy(ticks.data(), bt->boxTicks, sizeof(uint64_t) * bt->nboxes,
                     hipMemcpyDeviceToHost));
  int rate = 0;  // device wall clock, kHz
  BHIP(bt, hipDeviceGetAttribute(&rate, hipDeviceAttributeWallClockRate, bt->device));
  bt->usPerMove.assign(bt->nboxes, 0.0);
  bt->lastUs.assign(bt->nboxes, 0.0);
  for (int b = 0; b < bt->nboxes; ++b) {
    const double us = rate > 0 ? double(ticks[b]) * 1e3 / rate : 0.0;
    bt->lastUs[b] = us;
    if (moves[b] > 0) bt->usPerMove[b] = us / double(moves[b]);
  }
  return GOMC_OK;
}

}  // namespace gomc

extern "C" {

int gomc_batch_default_tuning(gomc_batch_tuning *out) {
  if (!out) return GOMC_ERR_ARG;
  *out = gomc_batch().tune;
  return GOMC_OK;
}

int gomc_batch_create(const gomc_system_desc *descs, int nboxes, int device, int precision,
                      gomc_batch **out) {
  return gomc_batch_create_tuned(descs, nboxes, device, precision, nullptr, out);
}

int gomc_batch_box_times(const gomc_batch *bt, double *us) {
  if (!bt || !us) return GOMC_ERR_ARG;
  for (int b = 0; b < bt->nboxes; ++b) us[b] = b < int(bt->lastUs.size()) ? bt->lastUs[b] : 0.0;
  return GOMC_OK;
}

int gomc_batch_create_tuned(const gomc_system_desc *descs, int nboxes, int device, int precision,
                            const gomc_batch_tuning *tuning, gomc_batch **out) {
  if (!descs || nboxes <= 0 || !out || precision < 0 || precision > 1) return GOMC_ERR_ARG;
  *out = nullptr;
  gomc_batch *bt = new gomc_batch();
  if (tuning) bt->tune = *tuning;
  bt->device = device;
  bt->precision = precision;
  bt->nboxes = nboxes;
  auto fail = [&](int code) {
    gomc_batch_destroy(bt);
    return code;
  };
  if (hipSetDevice(device) != hipSuccess) return fail(GOMC_ERR_NODEV);
  int cus = 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess)
    return fail(GOMC_ERR_NODEV);
  const int gridCap = std::max(8, cus - cus % 8);
  std::vector<PackedBox> packed(nboxes);
  for (int i = 0; i < nboxes; ++i) {
    int rc = PackBox(bt, descs[i], 0, 0, 0, packed[i]);
    if (rc) return fail(rc);
    packed[i].P.team = TeamSize(bt, packed[i].P.nmols);
  }
  // Teams need their members resident together: when the batch holds more
  // large boxes than the device fits (a column of gridCap / 8 workgroups per
  // team, and up to a quarter of the workgroups left for the other boxes),
  // the smallest teams are halved.
  {
    int nsolo = 0;
    for (auto &pb : packed) nsolo += pb.P.team <= 1;
    const int perCol = gridCap / 8;
    auto fits = [&]() {
      std::vector<int> t;
      int sum = 0;
      for (auto &pb : packed)
        if (pb.P.team > 1) {
          t.push_back(pb.P.team);
          sum += pb.P.team;
        }
      if (sum + std::min(nsolo, gridCap / 4) > gridCap) return false;
      std::sort(t.rbegin(), t.rend());
      std::vector<int> used(8, 0);
      for (int T : t) {
        int best = int(std::min_element(used.begin(), used.end()) - used.begin());
        if (perCol - used[best] < T) return false;
        used[best] += T;
      }
      return true;
    };
    while (!fits()) {
      int pick = -1;
      for (int i = 0; i < nboxes; ++i)
        if (packed[i].P.team > 1 &&
            (pick < 0 || packed[i].P.nmols < packed[pick].P.nmols ||
             (packed[i].P.nmols == packed[pick].P.nmols && packed[i].P.team > packed[pick].P.team)))
          pick = i;
      packed[pick].P.team /= 2;
      if (packed[pick].P.team <= 1) ++nsolo;
    }
  }
  int64_t atoms = 0, mols = 0, kvec = 0;
  int maxMols = 0;
  for (int i = 0; i < nboxes; ++i) {
    BatchBox &P = packed[i].P;
    P.atomOff = atoms;
    P.molOff = mols;
    P.kOff = kvec;
    atoms += P.natoms;
    mols += P.nmols;
    kvec += P.K;
    maxMols = std::max(maxMols, P.nmols);
  }
  // Boxes up to the cap keep their molecule centres in LDS.
  bt->ldsMols = BatchLdsMols(maxMols);
  // Team members wait for one another, so every workgroup of a launch must be
  // resident at once: the launch has at most gridCap <= CU-count workgroups,
  // which needs one resident workgroup per CU on an otherwise idle device.
  // (Another process's kernels on the same device can still hold CUs; a
  // team wait is bounded and then reports an error instead of hanging.)
  {
    int perCU = 0;
    if (BatchResidentPerCU(precision, bt->ldsMols, &perCU) != hipSuccess)
      return fail(GOMC_ERR_NODEV);
    bool teams = false;
    for (auto &pb : packed) teams = teams || pb.P.team > 1;
    if (teams && int64_t(perCU) * cus < gridCap) return fail(GOMC_ERR_STATE);
  }
  int rc = AllocBatch(bt, nboxes, atoms, mols, kvec);
  if (rc) return fail(rc);
  bt->hboxes.resize(nboxes);
  for (int i = 0; i < nboxes; ++i) {
    rc = UploadBox(bt, i, packed[i]);
    if (rc) return fail(rc);
  }
  *out = bt;
  return GOMC_OK;
}

int gomc_batch_run(gomc_batch *bt, const int64_t *moves) {
  if (!bt) return GOMC_ERR_ARG;
  BHIP(bt, hipSetDevice(bt->device));
  std::vector<int64_t> mv(bt->nboxes);
  for (int i = 0; i < bt->nboxes; ++i) {
    mv[i] = moves ? moves[i] : int64_t(bt->hboxes[i].nmols);
    if (mv[i] < 0) return BFail(bt, GOMC_ERR_ARG, "negative move count");
  }
  BHIP(bt, hipMemcpyAsync(bt->moves, mv.data(), sizeof(int64_t) * bt->nboxes, hipMemcpyHostToDevice,
                          bt->stream));
  int rc = BuildTasks(bt, mv);
  if (rc) return rc;
  const BatchArgs a = MakeArgs(bt);
  BHIP(bt, hipEventRecord(bt->ev0, bt->stream));
  BHIP(bt, LaunchBatch(bt->precision, bt->grid, a, bt->stream));
  BHIP(bt, hipEventRecord(bt->ev1, bt->stream));
  BHIP(bt, hipStreamSynchronize(bt->stream));
  float ms = 0.0f;
  BHIP(bt, hipEventElapsedTime(&ms, bt->ev0, bt->ev1));
  bt->lastKernelMs = ms;
  if (bt->trace.box >= 0) bt->traceLen = std::min<int64_t>(bt->trace.cap, mv[bt->trace.box]);
  return CheckTeams(bt, mv);
}

int gomc_batch_run_cycles(gomc_batch *bt, int ncycles, int use_graph) {
  if (!bt || ncycles < 0) return GOMC_ERR_ARG;
  BHIP(bt, hipSetDevice(bt->device));
  std::vector<int64_t> mv(bt->nboxes);
  for (int i = 0; i < bt->nboxes; ++i) mv[i] = int64_t(bt->hboxes[i].nmols);
  BHIP(bt, hipMemcpyAsync(bt->moves, mv.data(), sizeof(int64_t) * bt->nboxes, hipMemcpyHostToDevice,
                          bt->stream));
  BHIP(bt, hipStreamSynchronize(bt->stream));
  int rc = BuildTasks(bt, mv);
  if (rc) return rc;
  const BatchArgs a = MakeArgs(bt);
  if (use_graph && !bt->graphExec) {
    // one cycle = one kernel node; the move counts are read from device
    // memory when a replay runs, so the graph stays valid while they hold
    BHIP(bt, PrepareBatch(bt->precision, bt->ldsMols));
    BHIP(bt, hipStreamBeginCapture(bt->stream, hipStreamCaptureModeThreadLocal));
    hipError_t e = LaunchBatch(bt->precision, bt->grid, a, bt->stream, false);
    hipGraph_t g = nullptr;
    hipError_t e2 = hipStreamEndCapture(bt->stream, &g);
    if (e != hipSuccess || e2 != hipSuccess) {
      if (g) (void)hipGraphDestroy(g);
      return BFail(bt, GOMC_ERR_HIP, "batch graph capture failed");
    }
    bt->graph = g;
    BHIP(bt, hipGraphInstantiate(&bt->graphExec, bt->graph, nullptr, nullptr, 0));
  }
  BHIP(bt, hipEventRecord(bt->ev0, bt->stream));
  for (int c = 0; c < ncycles; ++c) {
    if (use_graph)
      BHIP(bt, hipGraphLaunch(bt->graphExec, bt->stream));
    else
      BHIP(bt, LaunchBatch(bt->precision, bt->grid, a, bt->stream));
  }
  BHIP(bt, hipEventRecord(bt->ev1, bt->stream));
  BHIP(bt, hipStreamSynchronize(bt->stream));
  float ms = 0.0f;
  BHIP(bt, hipEventElapsedTime(&ms, bt->ev0, bt->ev1));
  bt->lastKernelMs = ms;
  return CheckTeams(bt, mv);
}

int gomc_batch_stats(const gomc_batch *bt_, double *energies, int64_t *accepted, int64_t *moves_done) {
  gomc_batch *bt = const_cast<gomc_batch *>(bt_);
  if (!bt) return GOMC_ERR_ARG;
  BHIP(bt, hipSetDevice(bt->device));
  std::vector<BatchState> st(bt->nboxes);
  BHIP(bt, hipMemcpy(st.data(), bt->state, sizeof(BatchState) * bt->nboxes, hipMemcpyDeviceToHost));
  for (int i = 0; i < bt->nboxes; ++i) {
    if (energies)
      for (int e = 0; e < 7; ++e) energies[i * 7 + e] = st[i].energy[e];
    if (accepted) accepted[i] = st[i].accepted[0] + st[i].accepted[1];
    if (moves_done) moves_done[i] = st[i].step;
  }
  return GOMC_OK;
}

// Replica exchange between boxes of one batch: the neighbour exchange of
// ParallelTemperingUtilities::evaluateExchangeCriteria
// (src/ParallelTemperingUtilities.cpp:108-240, the non-multi branch), the
// MPI all-reduce of energies and the coordinate/centre exchange
// (exchangePositions/COMsNonBlocking, 406-448) replaced by an exchange of
// the temperature-bound state between the two boxes: beta, the move scales
// and counters and the Philox stream, so each configuration keeps its
// coordinates, structure factors and energies in HBM and moves to the other
// temperature without a copy.  ladder[s] is the box running the s-th
// temperature; pairs (s-1, s) with s % 2 == attempt % 2 are tried, each with
// one draw u < exp((beta_s - beta_{s-1}) (E_s - E_{s-1})), E the total
// energy of the configuration at that temperature.
int gomc_batch_exchange(gomc_batch *bt, int32_t *ladder, int n, int64_t attempt, uint32_t seed,
                        int32_t *accepted) {
  if (!bt || !ladder || n < 2 || attempt < 0) return GOMC_ERR_ARG;
  BHIP(bt, hipSetDevice(bt->device));
  std::vector<char> seen(bt->nboxes, 0);
  for (int s = 0; s < n; ++s) {
    const int b = ladder[s];
    if (b < 0 || b >= bt->nboxes || seen[b]) return BFail(bt, GOMC_ERR_ARG, "ladder must list distinct boxes");
    seen[b] = 1;
    const gomc::BatchBox &P = bt->hboxes[b], &Q = bt->hboxes[ladder[0]];
    if (P.natoms != Q.natoms || P.nmols != Q.nmols || P.L[0] != Q.L[0] || P.L[1] != Q.L[1] ||
        P.L[2] != Q.L[2])
      return BFail(bt, GOMC_ERR_ARG, "replicas must have the same molecules and box");
  }
  if (attempt == 0) bt->rexMt.Seed(seed);
  std::vector<gomc::BatchState> st(bt->nboxes);
  BHIP(bt, hipMemcpy(st.data(), bt->state, sizeof(gomc::BatchState) * bt->nboxes,
                     hipMemcpyDeviceToHost));
  std::vector<double> beta(n), energy(n);
  for (int s = 0; s < n; ++s) {
    beta[s] = bt->hboxes[ladder[s]].beta;
    energy[s] = st[ladder[s]].energy[6];
  }
  const int parity = int(attempt % 2);
  bool any = false;
  for (int s = 1; s < n; ++s) {
    int ok = 0;
    if (s % 2 == parity) {
      const double uBoltz = std::exp((beta[s] - beta[s - 1]) * (energy[s] - energy[s - 1]));
      ok = bt->rexMt.Closed() < uBoltz;
    }
    if (accepted) accepted[s] = ok;
    if (!ok) continue;
    any = true;
    gomc::BatchBox &A = bt->hboxes[ladder[s - 1]], &B = bt->hboxes[ladder[s]];
    std::swap(A.beta, B.beta);
    std::swap(A.seed, B.seed);
    gomc::BatchState &SA = st[ladder[s - 1]], &SB = st[ladder[s]];
    std::swap(SA.scale, SB.scale);
    std::swap(SA.tries, SB.tries);
    std::swap(SA.accepted, SB.accepted);
    std::swap(SA.tempTries, SB.tempTries);
    std::swap(SA.tempAccepted, SB.tempAccepted);
    std::swap(ladder[s - 1], ladder[s]);
  }
  if (accepted) accepted[0] = 0;
  if (any) {
    BHIP(bt, hipMemcpy(bt->boxes, bt->hboxes.data(), sizeof(gomc::BatchBox) * bt->nboxes,
                       hipMemcpyHostToDevice));
    BHIP(bt, hipMemcpy(bt->state, st.data(), sizeof(gomc::BatchState) * bt->nboxes,
                       hipMemcpyHostToDevice));
  }
  return GOMC_OK;
}

int gomc_batch_get_coords(const gomc_batch *bt_, int box, double *x, double *y, double *z) {
  gomc_batch *bt = const_cast<gomc_batch *>(bt_);
  if (!bt || box < 0 || box >= bt->nboxes) return GOMC_ERR_ARG;
  BHIP(bt, hipSetDevice(bt->device));
  const BatchBox &P = bt->hboxes[box];
  const size_t n = sizeof(double) * P.natoms;
  BHIP(bt, hipMemcpy(x, bt->x + P.atomOff, n, hipMemcpyDeviceToHost));
  BHIP(bt, hipMemcpy(y, bt->y + P.atomOff, n, hipMemcpyDeviceToHost));
  BHIP(bt, hipMemcpy(z, bt->z + P.atomOff, n, hipMemcpyDeviceToHost));
  return GOMC_OK;
}

int gomc_batch_box_layout(const gomc_batch *bt, int box, int32_t out[4]) {
  if (!bt || !out || box < 0 || box >= bt->nboxes) return GOMC_ERR_ARG;
  const BatchBox &P = bt->hboxes[box];
  out[0] = std::max(1, P.team);
  out[1] = P.kSeg;
  out[2] = P.kChunk;
  out[3] = P.cull;
  return GOMC_OK;
}

int gomc_batch_set_trace(gomc_batch *bt, int box, int64_t cap) {
  if (!bt || box >= bt->nboxes || cap < 0) return GOMC_ERR_ARG;
  BHIP(bt, hipSetDevice(bt->device));
  void *old[] = {bt->trace.acc, bt->trace.dlj, bt->trace.dreal, bt->trace.drecip};
  for (void *p : old)
    if (p) (void)hipFree(p);
  bt->trace = BatchTrace{-1, 0, nullptr, nullptr, nullptr, nullptr};
  bt->traceLen = 0;
  // a captured cycle holds the old trace pointers
  if (bt->graphExec) (void)hipGraphExecDestroy(bt->graphExec);
  if (bt->graph) (void)hipGraphDestroy(bt->graph);
  bt->graphExec = nullptr;
  bt->graph = nullptr;
  if (box < 0 || cap == 0) return GOMC_OK;
  BHIP(bt, hipMalloc(&bt->trace.acc, sizeof(int32_t) * cap));
  BHIP(bt, hipMalloc(&bt->trace.dlj, sizeof(double) * cap));
  BHIP(bt, hipMalloc(&bt->trace.dreal, sizeof(double) * cap));
  BHIP(bt, hipMalloc(&bt->trace.drecip, sizeof(double) * cap));
  bt->trace.box = box;
  bt->trace.cap = cap;
  return GOMC_OK;
}

int gomc_batch_get_trace(const gomc_batch *bt_, int32_t *acc, double *dlj, double *dreal,
                         double *drecip, int64_t *n) {
  gomc_batch *bt = const_cast<gomc_batch *>(bt_);
  if (!bt) return GOMC_ERR_ARG;
  BHIP(bt, hipSetDevice(bt->device));
  const int64_t len = bt->traceLen;
  if (n) *n = len;
  if (len == 0) return GOMC_OK;
  if (acc) BHIP(bt, hipMemcpy(acc, bt->trace.acc, sizeof(int32_t) * len, hipMemcpyDeviceToHost));
  if (dlj) BHIP(bt, hipMemcpy(dlj, bt->trace.dlj, sizeof(double) * len, hipMemcpyDeviceToHost));
  if (dreal) BHIP(bt, hipMemcpy(dreal, bt->trace.dreal, sizeof(double) * len, hipMemcpyDeviceToHost));
  if (drecip) BHIP(bt, hipMemcpy(drecip, bt->trace.drecip, sizeof(double) * len, hipMemcpyDeviceToHost));
  return GOMC_OK;
}

int gomc_batch_last_kernel_ms(const gomc_batch *bt, double *ms) {
  if (!bt || !ms) return GOMC_ERR_ARG;
  *ms = bt->lastKernelMs;
  return GOMC_OK;
}

}  // extern "C"
