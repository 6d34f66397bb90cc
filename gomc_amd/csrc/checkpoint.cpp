// Checkpoint and restore of one box's Monte Carlo state: what GOMC's
// Checkpoint (src/Checkpoint.h:118-167) keeps -- step, the MT19937 state
// (saveArray/seedLocation), the MoveSettings counters and scales, the
// MultiParticle settings, the MoleculeLookup and the coordinates -- plus
// the box geometry (NPT), the molecule centres, the running energies and the
// Ewald structure factors, so that a restored box continues the run it was
// taken from move for move (same draws, same energy changes, not merely a
// recomputation from coordinates as the reference's restart does).
//
// The format is this library's own (little-endian, versioned), not cereal's
// binary archive: the reference archive also carries MolSetup and PDB atom
// records of its own classes, which the energy path does not hold.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "box.hpp"

namespace gomc {
namespace {

constexpr char kMagic[8] = {'G', 'O', 'M', 'C', 'H', 'C', 'K', 'P'};
constexpr uint32_t kVersion = 2;

// FNV-1a over the bytes of what fixes the box's Hamiltonian and ensemble:
// cutoffs, Ewald splitting, temperature, the switches, the pair tables, the
// atoms' types and charges, the molecule layout, fugacities, pressure and
// the CBMC trial counts.  A checkpoint restores running energies, so one
// taken under another Hamiltonian must be refused, not continued.
struct Fnv {
  uint64_t h = 1469598103934665603ull;
  void Bytes(const void *p, size_t n) {
    const unsigned char *c = static_cast<const unsigned char *>(p);
    for (size_t i = 0; i < n; ++i) h = (h ^ c[i]) * 1099511628211ull;
  }
  template <class T> void Put(const T &v) { Bytes(&v, sizeof(T)); }
  template <class T> void PutVec(const std::vector<T> &v) {
    Put<uint64_t>(v.size());
    Bytes(v.data(), sizeof(T) * v.size());
  }
};

uint64_t Fingerprint(const gomc_box *b) {
  Fnv f;
  for (double v : {b->rCutLJ, b->rCutCoul, b->rCutLow, b->tol, b->alpha, b->recipRcut, b->T,
                   b->pressure})
    f.Put(v);
  for (bool v : {b->ewald, b->elect, b->lrc}) f.Put<uint8_t>(v);
  f.Put<int32_t>(b->ntypes);
  f.PutVec(b->sigSq);
  f.PutVec(b->epsCn);
  f.PutVec(b->nRep);
  f.PutVec(b->atomType);
  f.PutVec(b->atomCharge);
  f.PutVec(b->molStart);
  f.PutVec(b->fugacity);
  for (int v : {b->cbmcFirst, b->cbmcNth, b->cbmcAng}) f.Put<int32_t>(v);
  return f.h;
}

// The MoleculeLookup read from a checkpoint must describe the system's
// molecules: section starts from 0 to n_mols in order, every molecule listed
// once, each in the section of its box and kind, box 0 or 1 (1 only in
// GCMC), swappable counts within their sections and box-0 kind counts that
// agree.  Later moves index the box's arrays with these values.
bool LookupConsistent(const gomc_box *b, const std::vector<int> &molBox,
                      const std::vector<uint32_t> &lookup, const std::vector<uint32_t> &bkStart,
                      const std::vector<uint32_t> &swapCount, const std::vector<int> &kindCount) {
  const int K = b->nkinds, N = b->nmols;
  if (bkStart.empty() || bkStart.front() != 0 || bkStart.back() != uint32_t(N)) return false;
  if (lookup.size() != size_t(N) || molBox.size() != size_t(N)) return false;
  if (bkStart.size() != size_t(2 * K + 1) || swapCount.size() != size_t(2 * K) ||
      kindCount.size() != size_t(K))
    return false;
  for (size_t s = 1; s < bkStart.size(); ++s)
    if (bkStart[s] < bkStart[s - 1]) return false;
  std::vector<char> seen(size_t(N), 0);
  std::vector<int> box0(size_t(K), 0);
  for (int m = 0; m < N; ++m) {
    if (molBox[m] != 0 && molBox[m] != 1) return false;
    if (molBox[m] == 1 && b->ensemble != 1) return false;
    if (molBox[m] == 0) box0[b->molKind[m]]++;
  }
  for (int s = 0; s < 2 * K; ++s) {
    const int bx = s / K, k = s % K;
    if (swapCount[s] > bkStart[s + 1] - bkStart[s]) return false;
    for (uint32_t i = bkStart[s]; i < bkStart[s + 1]; ++i) {
      const uint32_t m = lookup[i];
      if (m >= uint32_t(N) || seen[m] || molBox[m] != bx || b->molKind[m] != k) return false;
      seen[m] = 1;
    }
  }
  for (int k = 0; k < K; ++k)
    if (kindCount[k] != box0[k]) return false;
  return true;
}

struct Writer {
  std::vector<unsigned char> buf;
  template <class T> void Put(const T &v) {
    const unsigned char *p = reinterpret_cast<const unsigned char *>(&v);
    buf.insert(buf.end(), p, p + sizeof(T));
  }
  template <class T> void PutVec(const std::vector<T> &v) {
    Put<uint64_t>(v.size());
    const unsigned char *p = reinterpret_cast<const unsigned char *>(v.data());
    buf.insert(buf.end(), p, p + sizeof(T) * v.size());
  }
};

struct Reader {
  const unsigned char *p, *end;
  bool ok = true;
  template <class T> T Get() {
    T v{};
    if (size_t(end - p) < sizeof(T)) {
      ok = false;
      return v;
    }
    std::memcpy(&v, p, sizeof(T));
    p += sizeof(T);
    return v;
  }
  template <class T> std::vector<T> GetVec(uint64_t maxLen) {
    const uint64_t n = Get<uint64_t>();
    if (!ok || n > maxLen || size_t(end - p) < n * sizeof(T)) {
      ok = false;
      return {};
    }
    std::vector<T> v(n);
    std::memcpy(v.data(), p, n * sizeof(T));
    p += n * sizeof(T);
    return v;
  }
};

int Serialize(gomc_box *b, int64_t step, Writer &w) {
  int rc = FlushPending(b);
  if (rc) return rc;
  for (char c : kMagic) w.Put(c);
  w.Put(kVersion);
  w.Put<int64_t>(step);
  // identity of the system the state belongs to
  w.Put<int32_t>(b->natoms);
  w.Put<int32_t>(b->nmols);
  w.Put<int32_t>(b->nkinds);
  w.Put<int32_t>(b->ensemble);
  w.PutVec(b->molKind);
  w.Put<uint64_t>(Fingerprint(b));
  // geometry: edges, half edges and volume as the volume moves left them
  // (VolumeTransfer scales the half edges and adds dV to the volume, so
  // neither is recomputed from the edges)
  for (int i = 0; i < 3; ++i) w.Put(b->geom.L[i]);
  for (int i = 0; i < 3; ++i) w.Put(b->geom.half[i]);
  w.Put(b->volume);
  w.Put(b->volInv);
  // coordinates and centres
  w.PutVec(b->hx);
  w.PutVec(b->hy);
  w.PutVec(b->hz);
  w.PutVec(b->comx);
  w.PutVec(b->comy);
  w.PutVec(b->comz);
  // MT19937 state
  std::vector<uint32_t> words(624);
  int idx = 0;
  b->mt.GetState(words.data(), &idx);
  w.PutVec(words);
  w.Put<int32_t>(idx);
  // MoveSettings
  for (int mv = 0; mv < 2; ++mv) {
    w.PutVec(b->moves.scale[mv]);
    w.PutVec(b->moves.tries[mv]);
    w.PutVec(b->moves.accepted[mv]);
    w.PutVec(b->moves.tempTries[mv]);
    w.PutVec(b->moves.tempAccepted[mv]);
  }
  w.Put(b->moves.volScale);
  w.Put(b->moves.volTries);
  w.Put(b->moves.volAccepted);
  w.Put(b->moves.volTempTries);
  w.Put(b->moves.volTempAccepted);
  // MultiParticle settings
  w.Put(b->mp.tMax);
  w.Put(b->mp.rMax);
  for (int i = 0; i < 2; ++i) {
    w.Put(b->mp.tries[i]);
    w.Put(b->mp.acc[i]);
    w.Put(b->mp.itries[i]);
    w.Put(b->mp.iacc[i]);
  }
  // MoleculeLookup and transfer counters
  w.PutVec(b->molBox);
  w.PutVec(b->lookup);
  w.PutVec(b->bkStart);
  w.PutVec(b->swapCount);
  w.PutVec(b->kindCount);
  w.PutVec(b->swapTries);
  w.PutVec(b->swapAcc);
  // running energies
  const double pot[7] = {b->pot.inter, b->pot.tc, b->pot.real, b->pot.recip,
                         b->pot.self,  b->pot.corr, b->pot.total};
  for (double e : pot) w.Put(e);
  // structure factors of the current configuration
  std::vector<double> sr(size_t(b->ewald ? b->K : 0)), si(sr.size());
  if (!sr.empty()) {
    HIP_TRY(b, hipMemcpyAsync(sr.data(), b->sumRref, sizeof(double) * sr.size(),
                              hipMemcpyDeviceToHost, b->stream));
    HIP_TRY(b, hipMemcpyAsync(si.data(), b->sumIref, sizeof(double) * si.size(),
                              hipMemcpyDeviceToHost, b->stream));
    HIP_TRY(b, hipStreamSynchronize(b->stream));
  }
  w.PutVec(sr);
  w.PutVec(si);
  return GOMC_OK;
}

}  // namespace
}  // namespace gomc

using namespace gomc;

extern "C" {

int64_t gomc_box_checkpoint(gomc_box *b, int64_t step, void *buf, int64_t cap) {
  if (!b || cap < 0 || (cap > 0 && !buf)) return GOMC_ERR_ARG;
  (void)hipSetDevice(b->device);
  Writer w;
  int rc = Serialize(b, step, w);
  if (rc) return rc;
  const int64_t n = int64_t(w.buf.size());
  if (buf && cap >= n) std::memcpy(buf, w.buf.data(), size_t(n));
  return n;
}

int gomc_box_restore(gomc_box *b, const void *buf, int64_t size, int64_t *step) {
  if (!b || !buf || size <= 0) return GOMC_ERR_ARG;
  (void)hipSetDevice(b->device);
  Reader r{static_cast<const unsigned char *>(buf), static_cast<const unsigned char *>(buf) + size};
  for (char c : kMagic)
    if (r.Get<char>() != c || !r.ok) return Fail(b, GOMC_ERR_ARG, "not a box checkpoint");
  if (r.Get<uint32_t>() != kVersion) return Fail(b, GOMC_ERR_ARG, "unknown checkpoint version");
  const int64_t st = r.Get<int64_t>();
  const int32_t natoms = r.Get<int32_t>(), nmols = r.Get<int32_t>(), nkinds = r.Get<int32_t>(),
                ens = r.Get<int32_t>();
  const std::vector<int> molKind = r.GetVec<int>(uint64_t(b->nmols));
  const uint64_t print = r.Get<uint64_t>();
  if (!r.ok || natoms != b->natoms || nmols != b->nmols || nkinds != b->nkinds ||
      ens != b->ensemble || molKind != b->molKind)
    return Fail(b, GOMC_ERR_ARG, "checkpoint belongs to a different system");
  if (print != Fingerprint(b))
    return Fail(b, GOMC_ERR_ARG,
                "checkpoint was taken under another force field, cutoff, Ewald, charge or "
                "ensemble setting");
  double L[3], H[3];
  for (double &v : L) v = r.Get<double>();
  for (double &v : H) v = r.Get<double>();
  const double vol = r.Get<double>(), volInv = r.Get<double>();
  const uint64_t na = uint64_t(natoms), nm = uint64_t(nmols), nk = uint64_t(nkinds);
  auto hx = r.GetVec<double>(na), hy = r.GetVec<double>(na), hz = r.GetVec<double>(na);
  auto cx = r.GetVec<double>(nm), cy = r.GetVec<double>(nm), cz = r.GetVec<double>(nm);
  auto words = r.GetVec<uint32_t>(624);
  const int32_t idx = r.Get<int32_t>();
  DriverMoves mv = b->moves;
  for (int m = 0; m < 2; ++m) {
    mv.scale[m] = r.GetVec<double>(nk);
    mv.tries[m] = r.GetVec<uint64_t>(nk);
    mv.accepted[m] = r.GetVec<uint64_t>(nk);
    mv.tempTries[m] = r.GetVec<uint64_t>(nk);
    mv.tempAccepted[m] = r.GetVec<uint64_t>(nk);
  }
  mv.volScale = r.Get<double>();
  mv.volTries = r.Get<uint64_t>();
  mv.volAccepted = r.Get<uint64_t>();
  mv.volTempTries = r.Get<uint64_t>();
  mv.volTempAccepted = r.Get<uint64_t>();
  MPState mp;
  mp.tMax = r.Get<double>();
  mp.rMax = r.Get<double>();
  for (int i = 0; i < 2; ++i) {
    mp.tries[i] = r.Get<uint64_t>();
    mp.acc[i] = r.Get<uint64_t>();
    mp.itries[i] = r.Get<uint64_t>();
    mp.iacc[i] = r.Get<uint64_t>();
  }
  auto molBox = r.GetVec<int>(nm);
  auto lookup = r.GetVec<uint32_t>(nm);
  auto bkStart = r.GetVec<uint32_t>(2 * nk + 1);
  auto swapCount = r.GetVec<uint32_t>(2 * nk);
  auto kindCount = r.GetVec<int>(nk);
  auto swapTries = r.GetVec<uint64_t>(2 * nk);
  auto swapAcc = r.GetVec<uint64_t>(2 * nk);
  double pot[7];
  for (double &e : pot) e = r.Get<double>();
  auto sr = r.GetVec<double>(uint64_t(1) << 32), si = r.GetVec<double>(uint64_t(1) << 32);
  if (!r.ok || r.p != r.end) return Fail(b, GOMC_ERR_ARG, "truncated or malformed checkpoint");
  if (hx.size() != na || hy.size() != na || hz.size() != na || cx.size() != nm ||
      cy.size() != nm || cz.size() != nm || words.size() != 624 || idx < 0 || idx > 624 ||
      molBox.size() != b->molBox.size() || lookup.size() != b->lookup.size() ||
      bkStart.size() != b->bkStart.size() || swapCount.size() != b->swapCount.size() ||
      kindCount.size() != b->kindCount.size() || swapTries.size() != b->swapTries.size() ||
      swapAcc.size() != b->swapAcc.size() || sr.size() != si.size())
    return Fail(b, GOMC_ERR_ARG, "checkpoint arrays do not match the system");
  for (int m = 0; m < 2; ++m)
    if (mv.scale[m].size() != nk || mv.tries[m].size() != nk || mv.accepted[m].size() != nk ||
        mv.tempTries[m].size() != nk || mv.tempAccepted[m].size() != nk)
      return Fail(b, GOMC_ERR_ARG, "checkpoint move settings do not match the system");
  for (int i = 0; i < 3; ++i)
    if (!(L[i] > 0.0) || !(H[i] > 0.0) || !(vol > 0.0))
      return Fail(b, GOMC_ERR_ARG, "checkpoint box length not positive");
  if (2.0 * std::max(b->rCutLJ, b->rCutCoul) > std::min(L[0], std::min(L[1], L[2])))
    return Fail(b, GOMC_ERR_ARG, "checkpoint box is smaller than twice the cutoff");
  for (int i = 0; i < 3; ++i)
    if (!(H[i] == 0.5 * L[i] || std::abs(H[i] - 0.5 * L[i]) <= 1e-9 * L[i]))
      return Fail(b, GOMC_ERR_ARG, "checkpoint half edges do not match its edges");
  for (const auto *v : {&hx, &hy, &hz, &cx, &cy, &cz})
    for (double c : *v)
      if (!std::isfinite(c)) return Fail(b, GOMC_ERR_ARG, "checkpoint coordinate not finite");
  for (int m = 0; m < 2; ++m)
    for (double sc : mv.scale[m])
      if (!(sc > 0.0) || !std::isfinite(sc))
        return Fail(b, GOMC_ERR_ARG, "checkpoint move scale not positive");
  if (!LookupConsistent(b, molBox, lookup, bkStart, swapCount, kindCount))
    return Fail(b, GOMC_ERR_ARG, "checkpoint molecule lookup does not describe the system");
  // the k-vector count of the checkpoint's geometry, computed before the box
  // is touched
  if (b->ewald) {
    std::vector<double> kx, ky, kz, pref;
    KVectorsOrth(L, b->alpha * b->alpha, b->recipRcut, vol, kx, ky, kz, pref, nullptr);
    if (sr.size() != kx.size())
      return Fail(b, GOMC_ERR_ARG, "checkpoint k-vector count does not match the box");
  } else if (!sr.empty()) {
    return Fail(b, GOMC_ERR_ARG, "checkpoint has structure factors but Ewald is off");
  }

  // A device move loop keeps a copy of the state: drop it, the next device
  // run uploads the restored one.
  DevFree(b);
  b->pendingMol = -1;
  b->stageMol = -1;
  TrialCellsInvalidate(b);
  int rc = GOMC_OK;
  bool newGeom = vol != b->volume || volInv != b->volInv;
  for (int i = 0; i < 3; ++i) newGeom = newGeom || L[i] != b->geom.L[i] || H[i] != b->geom.half[i];
  if (newGeom) {
    for (int i = 0; i < 3; ++i) {
      b->geom.L[i] = L[i];
      b->geom.half[i] = H[i];
    }
    b->volume = vol;
    b->volInv = volInv;
    rc = Regrid(b);
    if (rc) return rc;
  }
  b->hx = std::move(hx);
  b->hy = std::move(hy);
  b->hz = std::move(hz);
  b->comx = std::move(cx);
  b->comy = std::move(cy);
  b->comz = std::move(cz);
  b->mt.SetState(words.data(), idx);
  b->moves = std::move(mv);
  b->mp.tMax = mp.tMax;
  b->mp.rMax = mp.rMax;
  for (int i = 0; i < 2; ++i) {
    b->mp.tries[i] = mp.tries[i];
    b->mp.acc[i] = mp.acc[i];
    b->mp.itries[i] = mp.itries[i];
    b->mp.iacc[i] = mp.iacc[i];
  }
  b->mp.singleAccepted = true;  // reference forces are recomputed by the next MP move
  b->molBox = std::move(molBox);
  b->lookup = std::move(lookup);
  b->bkStart = std::move(bkStart);
  b->swapCount = std::move(swapCount);
  b->kindCount = std::move(kindCount);
  b->swapTries = std::move(swapTries);
  b->swapAcc = std::move(swapAcc);
  b->pot.inter = pot[0];
  b->pot.tc = pot[1];
  b->pot.real = pot[2];
  b->pot.recip = pot[3];
  b->pot.self = pot[4];
  b->pot.corr = pot[5];
  b->pot.total = pot[6];

  // device copies: coordinates, live molecule ids, cells, structure factors
  const size_t bytes = sizeof(double) * b->natoms;
  std::vector<int> ids(b->natoms);
  for (int m = 0; m < b->nmols; ++m) {
    const bool live = b->molBox.empty() || b->molBox[m] == 0;
    for (int p = b->molStart[m]; p < b->molStart[m] + b->molLen[m]; ++p) ids[p] = live ? m : -1;
  }
  HIP_TRY(b, hipMemcpyAsync(b->x, b->hx.data(), bytes, hipMemcpyHostToDevice, b->stream));
  HIP_TRY(b, hipMemcpyAsync(b->y, b->hy.data(), bytes, hipMemcpyHostToDevice, b->stream));
  HIP_TRY(b, hipMemcpyAsync(b->z, b->hz.data(), bytes, hipMemcpyHostToDevice, b->stream));
  HIP_TRY(b, hipMemcpyAsync(b->mol, ids.data(), sizeof(int) * ids.size(), hipMemcpyHostToDevice,
                            b->stream));
  rc = Sync(b);
  if (!rc) rc = BuildCells(b);
  if (rc) return rc;
  if (b->ewald) {
    double erecip = 0.0;
    rc = newGeom ? RecipInitImpl(b, &erecip) : GOMC_OK;
    if (rc) return rc;
    if (sr.size() != size_t(b->K))  // checked above against the same geometry
      return Fail(b, GOMC_ERR_STATE, "k-vector count changed during restore");
    const size_t kb = sizeof(double) * sr.size();
    HIP_TRY(b, hipMemcpyAsync(b->sumRref, sr.data(), kb, hipMemcpyHostToDevice, b->stream));
    HIP_TRY(b, hipMemcpyAsync(b->sumIref, si.data(), kb, hipMemcpyHostToDevice, b->stream));
    HIP_TRY(b, hipMemcpyAsync(b->sumRnew, sr.data(), kb, hipMemcpyHostToDevice, b->stream));
    HIP_TRY(b, hipMemcpyAsync(b->sumInew, si.data(), kb, hipMemcpyHostToDevice, b->stream));
    rc = Sync(b);
    if (rc) return rc;
    if (b->cached) {
      rc = CacheBuild(b);
      if (rc) return rc;
    }
  }
  if (step) *step = st;
  return GOMC_OK;
}

}  // extern "C"
