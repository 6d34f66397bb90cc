// Serving engine: a request loop that admits state-point jobs into one
// resident batch (the persistent MC kernel) as slots and device memory
// free up, advances every resident job in each launch, and retires jobs as
// they finish -- continuous batching.  A job's coordinates, centres, k
// vectors and structure factors stay in HBM from admission to completion.
//
// Memory: the flat atom / molecule / k-vector arrays of the batch are
// arenas carved out of the byte budget; a job is admitted (in arrival
// order) when a slot is free and first-fit extents of all three fit.
// Time: each launch gives every resident job a number of moves that costs
// about `tick_ms` at the measured per-move cost of its size, so one launch
// keeps every slot busy for the same time and small jobs are not held
// behind large ones.
//
// A job's moves draw from its own Philox stream indexed by its move
// number, so its result does not depend on which other jobs share the
// launches or how its moves are split between launches.
//
// GCMC jobs (ensemble 1: insertion/deletion with CBMC growth, two molecule
// kinds) run in a second lane: builder threads take queued GCMC jobs while
// the budget holds their bytes and build each as a device box; a runner
// advances all built jobs together, one launch of the device move loop
// (the reference move sequence, MT19937 draws continued on the device, a
// workgroup per job) after another, sampling the molecules of each kind in
// the box after every MC cycle.
// The lanes share the byte budget (the batched lane's arenas are reserved
// from it at creation; a GCMC job's device box is allocated on admission);
// jobs of both lanes overlap on the device.
#include <array>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>

#include "batch.hpp"
#include "box.hpp"

using namespace gomc;

namespace {

// First-fit extents over [0, cap) with coalescing on release.
class Arena {
 public:
  explicit Arena(int64_t cap = 0) {
    if (cap > 0) free_[0] = cap;
  }
  int64_t Alloc(int64_t n) {
    if (n <= 0) return 0;
    for (auto it = free_.begin(); it != free_.end(); ++it)
      if (it->second >= n) {
        const int64_t off = it->first, len = it->second;
        free_.erase(it);
        if (len > n) free_[off + n] = len - n;
        return off;
      }
    return -1;
  }
  void Release(int64_t off, int64_t n) {
    if (n <= 0) return;
    auto it = free_.emplace(off, n).first;
    auto next = std::next(it);
    if (next != free_.end() && it->first + it->second == next->first) {
      it->second += next->second;
      free_.erase(next);
    }
    if (it != free_.begin()) {
      auto prev = std::prev(it);
      if (prev->first + prev->second == it->first) {
        prev->second += it->second;
        free_.erase(it);
      }
    }
  }

 private:
  std::map<int64_t, int64_t> free_;
};

constexpr int kGcmcBuilders = 4;  // threads building GCMC jobs' device boxes
constexpr int64_t kAtomBytes = 3 * sizeof(double);                          // x, y, z
constexpr int64_t kMolBytes = 3 * sizeof(double);                           // centres
constexpr int64_t kKBytes = sizeof(int32_t) + 5 * sizeof(double);           // kn, pref, S, dS

// A descriptor deep-copied at submission, for jobs that are built later
// (the caller's arrays may be gone by then).
struct DescCopy {
  gomc_system_desc d{};
  std::vector<double> sigma, eps, nrep, kcharge, x, y, z, fug;
  std::vector<int32_t> knat, ktype, molKind, molBox;
};

std::unique_ptr<DescCopy> CopyDesc(const gomc_system_desc &d) {
  auto c = std::make_unique<DescCopy>();
  c->d = d;
  const int nt = d.n_atom_types, nk = d.n_kinds, nm = d.n_mols;
  auto cp = [](auto *src, size_t n, auto &dst) {
    dst.assign(src, src + n);
    return dst.data();
  };
  c->d.sigma = cp(d.sigma, nt, c->sigma);
  c->d.epsilon = cp(d.epsilon, nt, c->eps);
  c->d.n_rep = cp(d.n_rep, nt, c->nrep);
  c->d.kind_natoms = cp(d.kind_natoms, nk, c->knat);
  int64_t kindAtoms = 0;
  for (int k = 0; k < nk; ++k) kindAtoms += d.kind_natoms[k];
  c->d.kind_atom_type = cp(d.kind_atom_type, kindAtoms, c->ktype);
  c->d.kind_charge = cp(d.kind_charge, kindAtoms, c->kcharge);
  c->d.mol_kind = cp(d.mol_kind, nm, c->molKind);
  int64_t natoms = 0;
  for (int m = 0; m < nm; ++m) natoms += d.kind_natoms[d.mol_kind[m]];
  c->d.x = cp(d.x, natoms, c->x);
  c->d.y = cp(d.y, natoms, c->y);
  c->d.z = cp(d.z, natoms, c->z);
  c->d.mol_box = d.mol_box ? cp(d.mol_box, nm, c->molBox) : nullptr;
  c->d.kind_fugacity = d.kind_fugacity ? cp(d.kind_fugacity, nk, c->fug) : nullptr;
  return c;
}

struct Job {
  int64_t id = 0, totalMoves = 0, done = 0;
  int status = 0;  // 0 queued, 1 running, 2 done, 3 failed
  int slot = -1;
  int64_t atomOff = -1, molOff = -1, kOff = -1;
  double tSubmit = 0, tAdmit = 0, tDone = 0;
  double energy[7] = {0};
  int64_t accepted = 0;
  double costUs = 1.0;  // per-move cost estimate
  PackedBox pb;
  // GCMC lane: the job runs on its own device box and stream (the plain
  // path), from a deep copy of its descriptor
  bool gcmc = false;
  std::unique_ptr<DescCopy> desc;
  gomc_box *box = nullptr;
  int64_t gcmcBytes = 0;
  double sumN[2] = {0, 0};  // molecules of kinds 0, 1 in the box, summed over samples
  int64_t samples = 0;
};

double Now(const std::chrono::steady_clock::time_point &t0) {
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace

struct gomc_server {
  int device = 0, precision = 0, nslots = 0;
  double tickMs = 20.0;
  int64_t budget = 0, atomCap = 0, molCap = 0, kCap = 0, bytesUsed = 0, bytesPeak = 0;
  gomc_batch *bt = nullptr;
  Arena atoms, mols, kvec;
  std::mutex mu;
  std::condition_variable cv, doneCv;
  std::deque<Job *> queue;
  std::map<int64_t, std::unique_ptr<Job>> jobs;
  std::vector<Job *> slotJob;
  int64_t nextId = 1, movesDone = 0, ticks = 0, outstanding = 0;
  double kernelSec = 0.0;
  bool stop = false;
  std::string err;
  std::thread loop;
  std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
  // GCMC lane
  std::deque<Job *> gcmcQueue, gcmcReady;
  std::deque<std::pair<gomc_box *, int64_t>> gcmcRetire;  // boxes to destroy, their bytes
  std::vector<std::thread> workers;
  std::thread gcmcRunner;
  int gcmcBuilding = 0;
  int64_t gcmcLaunches = 0;
  double gcmcSec = 0.0;
};

namespace {

int64_t JobBytes(const Job &j) {
  return j.pb.P.natoms * kAtomBytes + j.pb.P.nmols * kMolBytes + int64_t(j.pb.P.K) * kKBytes;
}

// Admits queued jobs in arrival order while a slot and memory are free.
// Called with the lock held; uploads run with it released.
void Admit(gomc_server *sv, std::unique_lock<std::mutex> &lk) {
  while (!sv->queue.empty()) {
    int slot = -1;
    for (int s = 0; s < sv->nslots; ++s)
      if (!sv->slotJob[s]) {
        slot = s;
        break;
      }
    if (slot < 0) return;
    Job *j = sv->queue.front();
    const int64_t a = sv->atoms.Alloc(j->pb.P.natoms);
    const int64_t m = a < 0 ? -1 : sv->mols.Alloc(j->pb.P.nmols);
    const int64_t k = m < 0 ? -1 : sv->kvec.Alloc(j->pb.P.K);
    if (k < 0) {  // does not fit yet: give back what was taken, wait for a retire
      if (a >= 0) sv->atoms.Release(a, j->pb.P.natoms);
      if (m >= 0) sv->mols.Release(m, j->pb.P.nmols);
      return;
    }
    sv->queue.pop_front();
    j->atomOff = a;
    j->molOff = m;
    j->kOff = k;
    j->slot = slot;
    j->pb.P.atomOff = a;
    j->pb.P.molOff = m;
    j->pb.P.kOff = k;
    sv->slotJob[slot] = j;
    sv->bytesUsed += JobBytes(*j);
    sv->bytesPeak = std::max(sv->bytesPeak, sv->bytesUsed);
    lk.unlock();
    int rc = UploadBox(sv->bt, slot, j->pb);
    lk.lock();
    if (rc) {
      j->status = 3;
      sv->err = sv->bt->err;
      sv->slotJob[slot] = nullptr;
      --sv->outstanding;
      sv->doneCv.notify_all();
      continue;
    }
    j->status = 1;
    j->tAdmit = Now(sv->t0);
    // the host copies are on the device now
    j->pb.x.clear();
    j->pb.y.clear();
    j->pb.z.clear();
    j->pb.cx.clear();
    j->pb.cy.clear();
    j->pb.cz.clear();
    j->pb.pref.clear();
    j->pb.kn.clear();
    j->pb.sr.clear();
    j->pb.si.clear();
  }
}

void Loop(gomc_server *sv) {
  (void)hipSetDevice(sv->device);
  std::vector<int64_t> mv(sv->nslots, 0);
  std::vector<BatchState> st(sv->nslots);
  std::unique_lock<std::mutex> lk(sv->mu);
  while (true) {
    Admit(sv, lk);
    int running = 0;
    for (Job *j : sv->slotJob) running += j != nullptr;
    if (running == 0) {
      if (sv->stop) break;
      sv->cv.wait_for(lk, std::chrono::milliseconds(2));
      continue;
    }
    for (int s = 0; s < sv->nslots; ++s) {
      Job *j = sv->slotJob[s];
      if (!j) {
        mv[s] = 0;
        continue;
      }
      const int64_t quota = std::max<int64_t>(1, int64_t(sv->tickMs * 1000.0 / j->costUs));
      mv[s] = std::min(quota, j->totalMoves - j->done);
    }
    lk.unlock();
    int rc = gomc_batch_run(sv->bt, mv.data());
    double ms = sv->bt->lastKernelMs;
    hipError_t e = rc ? hipSuccess
                      : hipMemcpy(st.data(), sv->bt->state, sizeof(BatchState) * sv->nslots,
                                  hipMemcpyDeviceToHost);
    lk.lock();
    if (rc || e != hipSuccess) {
      sv->err = rc ? sv->bt->err : std::string("state copy: ") + hipGetErrorString(e);
      for (int s = 0; s < sv->nslots; ++s)
        if (Job *j = sv->slotJob[s]) {
          j->status = 3;
          sv->atoms.Release(j->atomOff, j->pb.P.natoms);
          sv->mols.Release(j->molOff, j->pb.P.nmols);
          sv->kvec.Release(j->kOff, j->pb.P.K);
          sv->bytesUsed -= JobBytes(*j);
          sv->slotJob[s] = nullptr;
          --sv->outstanding;
        }
      sv->doneCv.notify_all();
      continue;
    }
    sv->kernelSec += ms * 1e-3;
    ++sv->ticks;
    const double now = Now(sv->t0);
    for (int s = 0; s < sv->nslots; ++s) {
      Job *j = sv->slotJob[s];
      if (!j) continue;
      j->done += mv[s];
      sv->movesDone += mv[s];
      if (j->done < j->totalMoves) continue;
      for (int i = 0; i < 7; ++i) j->energy[i] = st[s].energy[i];
      j->accepted = st[s].accepted[0] + st[s].accepted[1];
      j->status = 2;
      j->tDone = now;
      sv->atoms.Release(j->atomOff, j->pb.P.natoms);
      sv->mols.Release(j->molOff, j->pb.P.nmols);
      sv->kvec.Release(j->kOff, j->pb.P.K);
      sv->bytesUsed -= JobBytes(*j);
      sv->slotJob[s] = nullptr;
      // an empty slot makes no moves and reads nothing
      BatchBox empty;
      std::memset(&empty, 0, sizeof(empty));
      sv->bt->hboxes[s] = empty;
      (void)hipMemcpy(sv->bt->boxes + s, &empty, sizeof(empty), hipMemcpyHostToDevice);
      --sv->outstanding;
    }
    sv->doneCv.notify_all();
  }
}

// Device bytes of a GCMC job on the plain path: coordinates, charges,
// kinds, cell list, forces are not built; k-vector arrays (index triple,
// prefactor, two structure-factor pairs) and per-molecule data.
int64_t GcmcBytes(const gomc_system_desc &d) {
  int64_t natoms = 0;
  for (int m = 0; m < d.n_mols; ++m) natoms += d.kind_natoms[d.mol_kind[m]];
  const double a = std::sqrt(-std::log(d.tolerance)) / d.rcut_coulomb;
  const double kc = -2.0 * std::log(d.tolerance) / d.rcut_coulomb;
  (void)a;
  // k-vectors in the half sphere |k| < kc: (2/3) pi (kc L / 2 pi)^3
  double kv = 2.0 / 3.0 * M_PI;
  for (int i = 0; i < 3; ++i) kv *= kc * d.box[i] / (2.0 * M_PI);
  return natoms * 96 + int64_t(d.n_mols) * 64 + int64_t(kv) * 64 + (int64_t(1) << 20);
}

// Builder of the GCMC lane: takes queued GCMC jobs in arrival order while
// the byte budget allows, builds each as a device box (initial energies,
// structure factors) and hands it to the runner; destroys the boxes of
// finished jobs and returns their bytes.
void GcmcBuilder(gomc_server *sv) {
  (void)hipSetDevice(sv->device);
  std::unique_lock<std::mutex> lk(sv->mu);
  while (true) {
    Job *j = nullptr;
    gomc_box *dead = nullptr;
    int64_t deadBytes = 0;
    while (true) {
      if (!sv->gcmcRetire.empty()) {
        dead = sv->gcmcRetire.front().first;
        deadBytes = sv->gcmcRetire.front().second;
        sv->gcmcRetire.pop_front();
        break;
      }
      if (sv->stop) return;
      if (!sv->gcmcQueue.empty() &&
          sv->bytesUsed + sv->gcmcQueue.front()->gcmcBytes <= sv->budget) {
        j = sv->gcmcQueue.front();
        sv->gcmcQueue.pop_front();
        break;
      }
      sv->cv.wait_for(lk, std::chrono::milliseconds(2));
    }
    if (dead) {
      lk.unlock();
      gomc_box_destroy(dead);
      lk.lock();
      sv->bytesUsed -= deadBytes;
      sv->cv.notify_all();
      continue;
    }
    sv->bytesUsed += j->gcmcBytes;
    sv->bytesPeak = std::max(sv->bytesPeak, sv->bytesUsed);
    ++sv->gcmcBuilding;
    j->status = 1;
    j->tAdmit = Now(sv->t0);
    lk.unlock();
    gomc_box *b = nullptr;
    int rc = gomc_box_create(&j->desc->d, sv->device, &b);
    if (!rc && sv->precision == 1) rc = gomc_set_precision(b, 1);
    lk.lock();
    --sv->gcmcBuilding;
    if (rc) {
      if (b) {
        sv->gcmcRetire.emplace_back(b, j->gcmcBytes);
      } else {
        sv->bytesUsed -= j->gcmcBytes;
      }
      j->status = 3;
      sv->err = "GCMC job setup failed";
      j->tDone = Now(sv->t0);
      j->desc.reset();
      --sv->outstanding;
      sv->doneCv.notify_all();
    } else {
      j->box = b;
      sv->gcmcReady.push_back(j);
    }
    sv->cv.notify_all();
  }
}

// Runner of the GCMC lane: every launch of the device move loop advances
// all running GCMC jobs together (one workgroup each), each by whole MC
// cycles of N_box moves -- about as many moves per job as the largest
// job's one cycle, so small jobs are not held behind large ones -- and
// samples the molecules of each kind in the box after every cycle.  Jobs
// join at the next launch after their box is built and leave when their
// moves are done.
void GcmcRunner(gomc_server *sv) {
  (void)hipSetDevice(sv->device);
  std::vector<Job *> run;
  std::unique_lock<std::mutex> lk(sv->mu);
  while (true) {
    while (!sv->gcmcReady.empty()) {
      run.push_back(sv->gcmcReady.front());
      sv->gcmcReady.pop_front();
    }
    if (sv->stop) {  // running jobs are dropped after the launch in flight
      for (Job *j : run) {
        sv->gcmcRetire.emplace_back(j->box, j->gcmcBytes);
        j->box = nullptr;
      }
      return;
    }
    if (run.empty()) {
      if (sv->stop) return;
      sv->cv.wait_for(lk, std::chrono::milliseconds(2));
      continue;
    }
    lk.unlock();
    const int n = int(run.size());
    std::vector<gomc_box *> boxes(n);
    std::vector<int64_t> start(n), steps(n), moves(n), acc(n), samples(n), nIn(n);
    std::vector<int32_t> cycles(n), status(n);
    std::vector<double> sums(2 * n);
    std::vector<int32_t> where;
    int64_t nMax = 1;
    for (int i = 0; i < n; ++i) {
      Job *j = run[i];
      boxes[i] = j->box;
      nIn[i] = std::max<int64_t>(1, j->box->bkStart[j->box->nkinds] - j->box->bkStart[0]);
      nMax = std::max(nMax, nIn[i]);
    }
    for (int i = 0; i < n; ++i) {
      start[i] = run[i]->done;
      steps[i] = run[i]->totalMoves - run[i]->done;
      cycles[i] = int32_t(std::max<int64_t>(1, nMax / nIn[i]));
    }
    const auto t0 = std::chrono::steady_clock::now();
    const int rc = gomc_run_moves_device_multi(boxes.data(), n, start.data(), steps.data(),
                                               cycles.data(), moves.data(), acc.data(),
                                               sums.data(), samples.data(), status.data());
    const double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    std::vector<std::array<double, 7>> energy(n);
    for (int i = 0; i < n; ++i) {
      const Job *j = run[i];
      if (!status[i] && j->done + moves[i] >= j->totalMoves) gomc_energies(j->box, energy[i].data());
    }
    lk.lock();
    sv->gcmcSec += sec;
    ++sv->gcmcLaunches;
    std::vector<Job *> keep;
    for (int i = 0; i < n; ++i) {
      Job *j = run[i];
      const bool failed = rc != GOMC_OK && status[i] != GOMC_OK;
      if (!failed) {
        j->done += moves[i];
        j->accepted += acc[i];
        j->sumN[0] += sums[2 * i];
        j->sumN[1] += sums[2 * i + 1];
        j->samples += samples[i];
        sv->movesDone += moves[i];
      }
      if (failed || j->done >= j->totalMoves) {
        if (failed) {
          j->status = 3;
          sv->err = gomc_last_error(j->box);
        } else {
          for (int e = 0; e < 7; ++e) j->energy[e] = energy[i][e];
          j->status = 2;
        }
        j->tDone = Now(sv->t0);
        sv->gcmcRetire.emplace_back(j->box, j->gcmcBytes);
        j->box = nullptr;
        j->desc.reset();
        --sv->outstanding;
        sv->doneCv.notify_all();
      } else {
        keep.push_back(j);
      }
    }
    run.swap(keep);
    sv->cv.notify_all();
  }
}

}  // namespace

extern "C" {

const char *gomc_server_last_error(const gomc_server *sv) {
  return sv ? sv->err.c_str() : "null server";
}

int gomc_server_create(int device, int precision, int slots, int64_t budget_bytes, double tick_ms,
                       gomc_server **out) {
  if (!out || slots <= 0 || budget_bytes <= 0 || !(tick_ms > 0) || precision < 0 || precision > 1)
    return GOMC_ERR_ARG;
  *out = nullptr;
  auto sv = std::make_unique<gomc_server>();
  sv->device = device;
  sv->precision = precision;
  sv->nslots = slots;
  sv->tickMs = tick_ms;
  sv->budget = budget_bytes;
  // The batched lane's arenas take the budget, up to 64 MiB per slot (a
  // water job of about 20,000 molecules), so that a large budget meant for
  // GCMC jobs does not reserve device memory the NVT lane cannot use.  They
  // are split in the proportions of a water box: 3 atoms and about 3.1
  // k-vectors per molecule, k-vector slots rounded up to whole chunks of
  // 1024 (about 1.4 slots per k-vector over 200-2000 molecules).
  const double arena = double(std::min<int64_t>(budget_bytes, int64_t(slots) * (int64_t(64) << 20)));
  sv->atomCap = int64_t(0.25 * arena) / kAtomBytes;
  sv->molCap = int64_t(0.085 * arena) / kMolBytes;
  sv->kCap = int64_t(0.665 * arena) / kKBytes;
  sv->atoms = Arena(sv->atomCap);
  sv->mols = Arena(sv->molCap);
  sv->kvec = Arena(sv->kCap);
  if (hipSetDevice(device) != hipSuccess) return GOMC_ERR_NODEV;
  gomc_batch *bt = new gomc_batch();
  bt->device = device;
  bt->precision = precision;
  bt->nboxes = slots;
  bt->teams = false;  // a job's result must not depend on the jobs beside it
  bt->hboxes.assign(slots, BatchBox());
  std::memset(bt->hboxes.data(), 0, sizeof(BatchBox) * slots);
  int rc = AllocBatch(bt, slots, sv->atomCap, sv->molCap, sv->kCap);
  if (!rc) {
    std::vector<BatchState> zs(slots);
    std::memset(zs.data(), 0, sizeof(BatchState) * slots);
    if (hipMemcpy(bt->boxes, bt->hboxes.data(), sizeof(BatchBox) * slots, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(bt->state, zs.data(), sizeof(BatchState) * slots, hipMemcpyHostToDevice) != hipSuccess)
      rc = GOMC_ERR_HIP;
  }
  if (rc) {
    gomc_batch_destroy(bt);
    return rc;
  }
  bt->ldsMols = BatchLdsMols(kBatchLdsMolCap);
  sv->bt = bt;
  sv->slotJob.assign(slots, nullptr);
  gomc_server *raw = sv.release();
  raw->loop = std::thread(Loop, raw);
  for (int w = 0; w < kGcmcBuilders; ++w) raw->workers.emplace_back(GcmcBuilder, raw);
  raw->gcmcRunner = std::thread(GcmcRunner, raw);
  *out = raw;
  return GOMC_OK;
}

int gomc_server_submit(gomc_server *sv, const gomc_system_desc *d, int64_t moves, int64_t *job_id) {
  if (!sv || !d || moves < 0 || !job_id) return GOMC_ERR_ARG;
  auto j = std::make_unique<Job>();
  j->totalMoves = moves;
  if (d->ensemble == 1) {  // GCMC lane
    j->gcmc = true;
    j->desc = CopyDesc(*d);
    j->gcmcBytes = GcmcBytes(*d);
    std::lock_guard<std::mutex> g(sv->mu);
    if (j->gcmcBytes > sv->budget) {
      sv->err = "job larger than the memory budget";
      return GOMC_ERR_ARG;
    }
    j->id = sv->nextId++;
    j->tSubmit = Now(sv->t0);
    *job_id = j->id;
    Job *raw = j.get();
    sv->jobs[j->id] = std::move(j);
    ++sv->outstanding;
    sv->gcmcQueue.push_back(raw);
    sv->cv.notify_all();
    return GOMC_OK;
  }
  // build and pack the job (initial energies, structure factors) now, on
  // the caller's thread, while the loop keeps the device busy
  gomc_batch scratch;  // holds the packing error, not shared with the loop
  scratch.device = sv->device;
  scratch.teams = false;
  int rc = PackBox(&scratch, *d, 0, 0, 0, j->pb);
  if (rc) {
    std::lock_guard<std::mutex> g(sv->mu);
    sv->err = scratch.err;
    return rc;
  }
  if (j->pb.P.natoms > sv->atomCap || j->pb.P.nmols > sv->molCap || j->pb.P.K > sv->kCap) {
    std::lock_guard<std::mutex> g(sv->mu);
    sv->err = "job larger than the memory budget";
    return GOMC_ERR_ARG;
  }
  // per-move cost of the batched kernel at this size, FP64, microseconds
  j->costUs = 10.0 + 0.006 * double(j->pb.P.nmols);
  std::lock_guard<std::mutex> g(sv->mu);
  j->id = sv->nextId++;
  j->tSubmit = Now(sv->t0);
  *job_id = j->id;
  Job *raw = j.get();
  sv->jobs[j->id] = std::move(j);
  if (raw->totalMoves == 0) {
    raw->status = 2;
    raw->tAdmit = raw->tDone = raw->tSubmit;
    for (int i = 0; i < 7; ++i) raw->energy[i] = raw->pb.S.energy[i];
    return GOMC_OK;
  }
  ++sv->outstanding;
  sv->queue.push_back(raw);
  sv->cv.notify_all();
  return GOMC_OK;
}

int gomc_server_poll(gomc_server *sv, int64_t job_id, int32_t *status, double times[3],
                     double energies[7], int64_t *accepted, int64_t *moves_done) {
  if (!sv) return GOMC_ERR_ARG;
  std::lock_guard<std::mutex> g(sv->mu);
  auto it = sv->jobs.find(job_id);
  if (it == sv->jobs.end()) return GOMC_ERR_ARG;
  const Job &j = *it->second;
  if (status) *status = j.status;
  if (times) {
    times[0] = j.tSubmit;
    times[1] = j.tAdmit;
    times[2] = j.tDone;
  }
  if (energies)
    for (int i = 0; i < 7; ++i) energies[i] = j.energy[i];
  if (accepted) *accepted = j.accepted;
  if (moves_done) *moves_done = j.done;
  return GOMC_OK;
}

int gomc_server_poll_counts(gomc_server *sv, int64_t job_id, double mean_n[2]) {
  if (!sv || !mean_n) return GOMC_ERR_ARG;
  std::lock_guard<std::mutex> g(sv->mu);
  auto it = sv->jobs.find(job_id);
  if (it == sv->jobs.end()) return GOMC_ERR_ARG;
  const Job &j = *it->second;
  if (j.gcmc) {
    const double n = j.samples > 0 ? double(j.samples) : 1.0;
    mean_n[0] = j.sumN[0] / n;
    mean_n[1] = j.sumN[1] / n;
  } else {
    mean_n[0] = j.pb.P.nmols;
    mean_n[1] = 0.0;
  }
  return GOMC_OK;
}

int gomc_server_drain(gomc_server *sv, double timeout_s) {
  if (!sv) return GOMC_ERR_ARG;
  std::unique_lock<std::mutex> lk(sv->mu);
  const bool ok = sv->doneCv.wait_for(lk, std::chrono::duration<double>(timeout_s),
                                      [&] { return sv->outstanding == 0; });
  return ok ? GOMC_OK : GOMC_ERR_STATE;
}

int gomc_server_stats(gomc_server *sv, double out[8]) {
  if (!sv || !out) return GOMC_ERR_ARG;
  std::lock_guard<std::mutex> g(sv->mu);
  out[0] = double(sv->movesDone);
  out[1] = sv->kernelSec;
  out[2] = double(sv->ticks);
  out[3] = double(sv->bytesPeak);
  out[4] = double(sv->budget);
  out[5] = Now(sv->t0);
  out[6] = double(sv->queue.size() + sv->gcmcQueue.size());
  out[7] = double(sv->outstanding);
  return GOMC_OK;
}

int gomc_server_destroy(gomc_server *sv) {
  if (!sv) return GOMC_ERR_ARG;
  {
    std::lock_guard<std::mutex> g(sv->mu);
    sv->stop = true;
    // queued jobs are dropped; running ones finish their current launch
    sv->queue.clear();
    sv->gcmcQueue.clear();
    for (int s = 0; s < sv->nslots; ++s) sv->slotJob[s] = nullptr;
    sv->cv.notify_all();
  }
  if (sv->loop.joinable()) sv->loop.join();
  if (sv->gcmcRunner.joinable()) sv->gcmcRunner.join();  // its current launch finishes first
  for (auto &w : sv->workers)
    if (w.joinable()) w.join();
  for (auto &r : sv->gcmcRetire) gomc_box_destroy(r.first);
  for (Job *j : sv->gcmcReady)
    if (j->box) gomc_box_destroy(j->box);
  if (sv->bt) gomc_batch_destroy(sv->bt);
  delete sv;
  return GOMC_OK;
}

}  // extern "C"
