// mcfl_engine.cpp -- batched launches and the continuous-batching server
// (include/mcfl.h, mcfl_batch_* and mcfl_server_*).
//
// No counterpart in the reference: its multi-source scans are one process per
// .mci file (ScriptGenerator-CeCAR-Translation.ipynb:48-55) and its emission
// stage is a host loop of one full simulation per voxel (CUDAMCFLmain.cu:
// 471-594).  Here every simulation of a batch or of a serving epoch is a job
// of a single persistent launch (mcfl_transport.hip, MCFL_WORK_JOBS).
//
// Photon k of a job draws from the stream keyed by (seed, phase) -- phase 0
// excitation, phase 1 emission -- whatever launch, epoch or lane runs it, so a
// job's tallies equal those of the same job run alone.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <new>
#include <vector>

#include "mcfl_internal.h"
#include "mcfl_mwc.h"

using namespace mcfl_detail;

namespace {

// A medium resident in device memory: voxel labels (bulk method 2) or the
// layer bounds of a layered slab with its inclusion (bulk method 1).
struct Medium {
  mcfl_sim sim{};                 // scalar fields (pointers below)
  std::vector<mcfl_bulk> bulks;
  std::vector<mcfl_layer> layers;
  std::vector<uint8_t> labels;
  std::vector<MatDev> mat_ex, mat_em;
  DevBuf d_labels, d_tpos, d_zb;
  size_t nvox = 0, npix = 0;
  bool layered() const { return sim.bulk_method == 1; }
  ~Medium() {
    release(d_labels);
    release(d_tpos);
    release(d_zb);
  }
};

// The kernel's view of a resident medium.
void medium_dev(const Medium& m, MediumDev* md) {
  fill_medium(&m.sim, m.d_labels.as<const uint8_t>(), md, m.d_zb.as<const float>());
}

// The layer holding depth zi of a layered medium (0 when none does).
uint32_t layer_at(const Medium& m, float zi) {
  mcfl_sim t = m.sim;
  t.zi = zi;
  return start_layer(&t);
}

mcfl_status medium_make(const mcfl_sim* s, Medium* m, bool allow_layered) {
  MCFL_STCHK(validate(s));
  if (s->bulk_method == 1 && !allow_layered) return MCFL_ERR_UNSUPPORTED;  // serving takes voxel media
  m->sim = *s;
  m->nvox = nvox_of(s);
  m->npix = npix_of(s);
  m->sim.bulks = nullptr;
  m->sim.bulk_info = nullptr;
  m->sim.layers = nullptr;
  if (s->bulk_method == 1) {
    m->layers.assign(s->layers, s->layers + s->n_layers + 2);
    m->sim.layers = m->layers.data();
    std::vector<float> zb;
    layer_bounds(&m->sim, zb);
    MCFL_STCHK(ensure(m->d_zb, zb.size() * sizeof(float)));
    MCFL_HIPCHK(hipMemcpy(m->d_zb.p, zb.data(), zb.size() * sizeof(float), hipMemcpyHostToDevice));
  } else {
    m->bulks.assign(s->bulks, s->bulks + s->n_bulks + 2);
    m->labels.assign(s->bulk_info, s->bulk_info + m->nvox);
    m->sim.bulks = m->bulks.data();
    m->sim.bulk_info = m->labels.data();
    MCFL_STCHK(ensure(m->d_labels, m->nvox));
    MCFL_HIPCHK(hipMemcpy(m->d_labels.p, m->labels.data(), m->nvox, hipMemcpyHostToDevice));
  }
  build_mats(&m->sim, false, m->mat_ex);
  build_mats(&m->sim, true, m->mat_em);
  if (s->do_temp_sim == 1) {
    std::vector<float> pos;
    tdet_positions(s, pos);
    MCFL_STCHK(ensure(m->d_tpos, pos.size() * sizeof(float)));
    MCFL_HIPCHK(hipMemcpy(m->d_tpos.p, pos.data(), pos.size() * sizeof(float), hipMemcpyHostToDevice));
  }
  return MCFL_OK;
}

// Walker/Vose alias table; column i accepts itself when a uniform 32-bit
// draw u < prob[i], else takes alias[i].
void build_alias(const std::vector<double>& w, std::vector<uint32_t>& prob, std::vector<uint32_t>& alias) {
  const size_t n = w.size();
  prob.assign(n, 0xFFFFFFFFu);
  alias.resize(n);
  double sum = 0.0;
  for (double v : w) sum += v;
  std::vector<double> p(n);
  std::vector<uint32_t> small, large;
  for (size_t i = 0; i < n; ++i) {
    p[i] = w[i] * (double)n / sum;
    alias[i] = (uint32_t)i;
    (p[i] < 1.0 ? small : large).push_back((uint32_t)i);
  }
  while (!small.empty() && !large.empty()) {
    uint32_t s = small.back(), l = large.back();
    small.pop_back();
    alias[s] = l;
    prob[s] = (uint32_t)std::min(4294967295.0, std::floor(p[s] * 4294967296.0));
    p[l] = (p[l] + p[s]) - 1.0;
    if (p[l] < 1.0) {
      large.pop_back();
      small.push_back(l);
    }
  }
  // leftovers are full columns (rounding residue)
}

// Device tallies of one job (or one phase of a served request).
// Device buffers recycled between served requests.  hipFree waits for the
// whole device (both epochs in flight), so retired requests hand their
// buffers to the pool and admitted ones take them back (best fit, at most
// twice the size asked for); only what exceeds `limit` bytes is freed.
struct BufPool {
  std::multimap<size_t, void*> free_;
  size_t held = 0, limit = 0;
  ~BufPool() {
    for (auto& kv : free_) (void)hipFree(kv.second);
  }
  mcfl_status get(DevBuf& b, size_t bytes) {
    if (bytes == 0) bytes = 8;
    if (b.cap >= bytes) return MCFL_OK;
    put(b);
    auto it = free_.lower_bound(bytes);
    if (it != free_.end() && it->first / 2 <= bytes) {
      b.p = it->second;
      b.cap = it->first;
      held -= it->first;
      free_.erase(it);
      return MCFL_OK;
    }
    return ensure(b, bytes);
  }
  void put(DevBuf& b) {
    if (!b.p) return;
    free_.emplace(b.cap, b.p);
    held += b.cap;
    b.p = nullptr;
    b.cap = 0;
    while (held > limit && !free_.empty()) {
      auto it = std::prev(free_.end());
      held -= it->first;
      (void)hipFree(it->second);
      free_.erase(it);
    }
  }
};

// Pinned host staging buffers recycled between served requests (the same
// reason: freeing pinned memory is not free either); at most `limit` bytes
// are kept.
struct HostPool {
  std::multimap<size_t, uint8_t*> free_;
  size_t held = 0, limit = size_t(1) << 30;
  ~HostPool() {
    for (auto& kv : free_) (void)hipHostFree(kv.second);
  }
  mcfl_status get(uint8_t** p, size_t* cap, size_t bytes) {
    auto it = free_.lower_bound(bytes);
    if (it != free_.end() && it->first / 2 <= bytes) {
      *p = it->second;
      *cap = it->first;
      held -= it->first;
      free_.erase(it);
      return MCFL_OK;
    }
    if (hipHostMalloc((void**)p, bytes, hipHostMallocDefault) != hipSuccess) return MCFL_ERR_NOMEM;
    *cap = bytes;
    return MCFL_OK;
  }
  void put(uint8_t** p, size_t* cap) {
    if (!*p) return;
    free_.emplace(*cap, *p);
    held += *cap;
    *p = nullptr;
    *cap = 0;
    while (held > limit && !free_.empty()) {
      auto it = std::prev(free_.end());
      held -= it->first;
      (void)hipHostFree(it->second);
      free_.erase(it);
    }
  }
};

struct Tallies {
  DevBuf Rd, Tt, fhd, axyz, Rd2, Tt2, time, tot, al_prob, al_alias, al_vox;
  // Volume tallies of a job run in slices: each slice's per-XCD copies are
  // summed into these (saturating at 2^64-1) and cleared for the next slice.
  DevBuf fhd_acc, axyz_acc;
  uint32_t flags = 0;
  uint32_t copies = 1; /* per-XCD copies of fhd / axyz */
  uint32_t al_n = 0;
  double fl_unit = 0.0;
  size_t bytes = 0;
  BufPool* pool = nullptr; /* served requests: buffers come from and go back to the server's pool */
  mcfl_status get(DevBuf& b, size_t n) { return pool ? pool->get(b, n) : ensure(b, n); }
  ~Tallies() {
    for (DevBuf* b : {&Rd, &Tt, &fhd, &axyz, &Rd2, &Tt2, &time, &tot, &al_prob, &al_alias, &al_vox, &fhd_acc,
                      &axyz_acc}) {
      if (pool)
        pool->put(*b);
      else
        release(*b);
    }
  }
};

mcfl_status alloc_tallies(const Medium& m, uint32_t tallies, bool emission, uint32_t copies, Tallies* t) {
  const size_t npix = m.npix * 8, nvox = m.nvox * 8 * copies;
  t->copies = copies;
  MCFL_STCHK(t->get(t->Rd, npix));
  MCFL_STCHK(t->get(t->Tt, npix));
  MCFL_STCHK(t->get(t->tot, MCFL_NTOT * 8));
  t->bytes = 2 * npix + MCFL_NTOT * 8;
  t->flags = 0;
  if ((tallies & MCFL_TALLY_FHD) && !emission) {
    MCFL_STCHK(t->get(t->fhd, nvox));
    t->flags |= MCFL_F_FHD;
    t->bytes += nvox;
  }
  if (tallies & MCFL_TALLY_AXYZ) {
    MCFL_STCHK(t->get(t->axyz, nvox));
    t->flags |= MCFL_F_AXYZ;
    t->bytes += nvox;
  }
  if (tallies & MCFL_TALLY_SQ) {
    MCFL_STCHK(t->get(t->Rd2, npix));
    MCFL_STCHK(t->get(t->Tt2, npix));
    t->flags |= MCFL_F_SQ;
    t->bytes += 2 * npix;
  }
  if ((tallies & MCFL_TALLY_TEMP) && m.sim.do_temp_sim == 1 && !emission) {
    MCFL_STCHK(t->get(t->time, ntime_of(&m.sim) * 8));
    t->flags |= MCFL_F_TEMP;
    t->bytes += ntime_of(&m.sim) * 8;
  }
  return MCFL_OK;
}

// Queue clearing of every tally buffer with the fill kernel (graph safe).
mcfl_status clear_tallies(const Medium& m, Tallies* t, hipStream_t st) {
  const uint64_t npix = m.npix, nvox = m.nvox * t->copies;
  auto fill = [&](DevBuf& b, uint64_t n) { return b.p ? mcfl_fill_u64(b.as<unsigned long long>(), n, st) : 0; };
  if (fill(t->Rd, npix) || fill(t->Tt, npix) || fill(t->tot, MCFL_NTOT) || fill(t->fhd, nvox) ||
      fill(t->axyz, nvox) || fill(t->Rd2, npix) || fill(t->Tt2, npix) || fill(t->time, ntime_of(&m.sim)))
    return MCFL_ERR_DEVICE;
  return MCFL_OK;
}

// Sum the per-XCD copies of the volume tallies into copy 0.
int reduce_tallies(const Medium& m, Tallies* t, hipStream_t st) {
  int e = 0;
  if (t->fhd.p) e |= mcfl_reduce_copies(t->fhd.as<unsigned long long>(), m.nvox, t->copies, st);
  if (t->axyz.p) e |= mcfl_reduce_copies(t->axyz.as<unsigned long long>(), m.nvox, t->copies, st);
  return e;
}

// Volume tallies run in slices: allocate the accumulators (cleared on st).
mcfl_status alloc_accumulators(const Medium& m, Tallies* t, hipStream_t st) {
  if (t->fhd.p) {
    MCFL_STCHK(t->get(t->fhd_acc, m.nvox * 8));
    if (mcfl_fill_u64(t->fhd_acc.as<unsigned long long>(), m.nvox, st) != 0) return MCFL_ERR_DEVICE;
    t->bytes += m.nvox * 8;
  }
  if (t->axyz.p) {
    MCFL_STCHK(t->get(t->axyz_acc, m.nvox * 8));
    if (mcfl_fill_u64(t->axyz_acc.as<unsigned long long>(), m.nvox, st) != 0) return MCFL_ERR_DEVICE;
    t->bytes += m.nvox * 8;
  }
  return MCFL_OK;
}

// A slice's per-XCD copies into the accumulators (saturating), copies cleared.
int accumulate_tallies(const Medium& m, Tallies* t, hipStream_t st) {
  int e = 0;
  if (t->fhd_acc.p)
    e |= mcfl_accumulate_copies(t->fhd.as<unsigned long long>(), m.nvox, t->copies,
                                t->fhd_acc.as<unsigned long long>(), st);
  if (t->axyz_acc.p)
    e |= mcfl_accumulate_copies(t->axyz.as<unsigned long long>(), m.nvox, t->copies,
                                t->axyz_acc.as<unsigned long long>(), st);
  return e;
}

// Emission sources of a medium given its normalised fhd: weights
// eY (1 - albedof) Fx over voxels whose material has albedof < 1
// (CUDAMCFLmain.cu:550-556); returns the sum.
// stage (optional): pinned host memory of at least 12 bytes per voxel; the
// tables are then uploaded asynchronously on st from it (stage must stay
// untouched until st has passed the copies).
mcfl_status build_emission(const Medium& m, const double* Fx, uint64_t photons, Tallies* t, hipStream_t st = nullptr,
                           uint8_t* stage = nullptr) {
  std::vector<double> w;
  std::vector<uint32_t> vox;
  w.reserve(m.nvox);
  vox.reserve(m.nvox);
  double C = 0.0;
  for (size_t v = 0; v < m.nvox; ++v) {
    const mcfl_bulk& b = m.bulks[m.labels[v]];
    if (!(b.albedof < 1)) continue;
    const double c = (double)b.eY * (double)(1 - b.albedof) * Fx[v];
    if (!(c > 0.0)) continue;
    w.push_back(c);
    vox.push_back((uint32_t)v);
    C += c;
  }
  t->al_n = (uint32_t)vox.size();
  t->fl_unit = photons ? C / ((double)photons * 4294967295.0) : 0.0;
  if (vox.empty()) return MCFL_OK;  // nothing emits: the job launches no photon
  std::vector<uint32_t> prob, alias;
  build_alias(w, prob, alias);
  const size_t n = vox.size() * 4;
  MCFL_STCHK(t->get(t->al_prob, n));
  MCFL_STCHK(t->get(t->al_alias, n));
  MCFL_STCHK(t->get(t->al_vox, n));
  if (stage) {
    memcpy(stage, prob.data(), n);
    memcpy(stage + n, alias.data(), n);
    memcpy(stage + 2 * n, vox.data(), n);
    MCFL_HIPCHK(hipMemcpyAsync(t->al_prob.p, stage, n, hipMemcpyHostToDevice, st));
    MCFL_HIPCHK(hipMemcpyAsync(t->al_alias.p, stage + n, n, hipMemcpyHostToDevice, st));
    MCFL_HIPCHK(hipMemcpyAsync(t->al_vox.p, stage + 2 * n, n, hipMemcpyHostToDevice, st));
  } else {
    MCFL_HIPCHK(hipMemcpy(t->al_prob.p, prob.data(), n, hipMemcpyHostToDevice));
    MCFL_HIPCHK(hipMemcpy(t->al_alias.p, alias.data(), n, hipMemcpyHostToDevice));
    MCFL_HIPCHK(hipMemcpy(t->al_vox.p, vox.data(), n, hipMemcpyHostToDevice));
  }
  t->bytes += 3 * n;
  return MCFL_OK;
}

int src_kind_dev(int32_t k) {
  switch (k) {
    case MCFL_SOURCE_ISOTROPIC: return MCFL_SRC_ISOTROPIC;
    case MCFL_SOURCE_DISK: return MCFL_SRC_DISK;
    case MCFL_SOURCE_EMISSION: return MCFL_SRC_EMISSION;
    default: return MCFL_SRC_PENCIL;
  }
}

void fill_job(JobDev& J, const Medium& m, Tallies& t, int src_kind, float xi, float yi, float zi, float src_r,
              uint64_t seed, uint64_t begin, uint64_t count, uint64_t k0, uint32_t mat_base) {
  memset(&J, 0, sizeof J);
  const bool em = src_kind == MCFL_SOURCE_EMISSION;
  J.begin = begin;
  J.count = count;
  J.key = mcfl_key(seed, em ? 1u : 0u);
  J.k0 = k0;
  J.src_kind = src_kind_dev(src_kind);
  J.xi = xi;
  J.yi = yi;
  J.zi = zi;
  J.src_r = src_r;
  J.start_weight = m.sim.start_weight;
  J.mat_base = mat_base;
  J.flags = t.flags;
  J.Rd = t.Rd.as<unsigned long long>();
  J.Tt = t.Tt.as<unsigned long long>();
  J.fhd = t.fhd.as<unsigned long long>();
  J.axyz = t.axyz.as<unsigned long long>();
  J.time = t.time.as<unsigned long long>();
  J.Rd2 = t.Rd2.as<unsigned long long>();
  J.Tt2 = t.Tt2.as<unsigned long long>();
  J.tot = t.tot.as<unsigned long long>();
  J.copy_stride = t.copies > 1 ? m.nvox : 0;
  if (m.layered()) J.start_lab = layer_at(m, zi);
  fill_det(&m.sim, &J.det, m.d_tpos.as<const float>());
  medium_dev(m, &J.med);
  J.al_prob = t.al_prob.as<const uint32_t>();
  J.al_alias = t.al_alias.as<const uint32_t>();
  J.al_vox = t.al_vox.as<const uint32_t>();
  J.al_n = t.al_n;
}

// Concatenated material tables of the (medium, emission) pairs a launch uses.
struct MatTable {
  std::vector<MatDev> mats;
  std::vector<std::pair<const Medium*, bool>> keys;
  std::vector<uint32_t> bases;
  uint32_t base_of(const Medium* m, bool em) {
    for (size_t i = 0; i < keys.size(); ++i)
      if (keys[i].first == m && keys[i].second == em) return bases[i];
    const std::vector<MatDev>& src = em ? m->mat_em : m->mat_ex;
    uint32_t b = (uint32_t)mats.size();
    mats.insert(mats.end(), src.begin(), src.end());
    keys.emplace_back(m, em);
    bases.push_back(b);
    return b;
  }
};

// Download one job's tallies into caller buffers.
mcfl_status download(const Medium& m, const Tallies& t, uint64_t* Rd, uint64_t* Tt, uint64_t* fhd, uint64_t* axyz,
                     uint64_t* Rd2, uint64_t* Tt2, uint64_t* tim, uint64_t* totals) {
  const size_t npix = m.npix * 8, nvox = m.nvox * 8;
  if (Rd) MCFL_HIPCHK(hipMemcpy(Rd, t.Rd.p, npix, hipMemcpyDeviceToHost));
  if (Tt) MCFL_HIPCHK(hipMemcpy(Tt, t.Tt.p, npix, hipMemcpyDeviceToHost));
  if (fhd && t.fhd.p) MCFL_HIPCHK(hipMemcpy(fhd, (t.fhd_acc.p ? t.fhd_acc : t.fhd).p, nvox, hipMemcpyDeviceToHost));
  if (axyz && t.axyz.p)
    MCFL_HIPCHK(hipMemcpy(axyz, (t.axyz_acc.p ? t.axyz_acc : t.axyz).p, nvox, hipMemcpyDeviceToHost));
  if (Rd2 && t.Rd2.p) MCFL_HIPCHK(hipMemcpy(Rd2, t.Rd2.p, npix, hipMemcpyDeviceToHost));
  if (Tt2 && t.Tt2.p) MCFL_HIPCHK(hipMemcpy(Tt2, t.Tt2.p, npix, hipMemcpyDeviceToHost));
  if (tim && t.time.p) MCFL_HIPCHK(hipMemcpy(tim, t.time.p, ntime_of(&m.sim) * 8, hipMemcpyDeviceToHost));
  if (totals) MCFL_HIPCHK(hipMemcpy(totals, t.tot.p, MCFL_NTOT * 8, hipMemcpyDeviceToHost));
  return MCFL_OK;
}

// Shared launch plumbing: job table + material table upload, args.
struct Launcher {
  DevBuf d_jobs, d_mats, d_counter;
  ~Launcher() {
    release(d_jobs);
    release(d_mats);
    release(d_counter);
  }
  mcfl_status prepare(mcfl_ctx* c, const std::vector<JobDev>& jd, const MatTable& mt, bool perjob,
                      const Medium* uniform, LaunchArgs* A, hipStream_t st) {
    if (mt.mats.size() > MCFL_MAX_MATERIALS) return MCFL_ERR_CAPACITY;
    if (jd.size() > MCFL_MAX_JOBS) return MCFL_ERR_CAPACITY;
    MCFL_STCHK(ensure(d_jobs, jd.size() * sizeof(JobDev)));
    MCFL_STCHK(ensure(d_mats, mt.mats.size() * sizeof(MatDev)));
    MCFL_STCHK(ensure(d_counter, 8 * MCFL_MAX_JOBS));
    MCFL_HIPCHK(hipMemcpyAsync(d_mats.p, mt.mats.data(), mt.mats.size() * sizeof(MatDev), hipMemcpyHostToDevice, st));
    memset(A, 0, sizeof *A);
    if (!perjob) medium_dev(*uniform, &A->med);
    A->mats = d_mats.as<const MatDev>();
    A->n_mats = (uint32_t)mt.mats.size();
    A->work = MCFL_WORK_JOBS;
    uint64_t total = 0;
    uint32_t flags = 0;
    for (const JobDev& J : jd) {
      total += J.count;
      flags |= J.flags;
    }
    A->total = total;
    A->vol = mcfl_vol_of(flags);
    A->counter = d_counter.as<unsigned long long>();
    A->a_tab = c->d_atab;
    A->a_mask = kTablePhotons - 1;
    A->jobs = d_jobs.as<const JobDev>();
    A->njobs = (uint32_t)jd.size();
    A->chunk_max = c->chunk_max;
    A->refill_min = c->refill_min;
    A->max_steps = c->max_steps;
    return MCFL_OK;
  }
  // Job table upload, after the tiles are placed.
  mcfl_status upload(const std::vector<JobDev>& jd, hipStream_t st) {
    MCFL_HIPCHK(hipMemcpyAsync(d_jobs.p, jd.data(), jd.size() * sizeof(JobDev), hipMemcpyHostToDevice, st));
    return MCFL_OK;
  }
};

}  // namespace

// ================================================================== batch

struct mcfl_batch {
  mcfl_ctx* c = nullptr;
  std::vector<std::unique_ptr<Medium>> media;
  std::vector<std::unique_ptr<Tallies>> tal;
  std::vector<mcfl_job> spec;
  std::vector<JobDev> jd;
  MatTable mt;
  Launcher L;
  LaunchArgs A{};
  int temporal = 0, perjob = 0, nblocks = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  // A job above the context's photon split (mcfl_ctx::run_photons) makes the
  // batch run in slices: slice r runs photons [r S, (r+1) S) of every job
  // that has them, the volume tallies summed into saturating accumulators
  // between slices and the totals on the host, as mcfl_run_excitation
  // splits a long run.
  bool sliced = false;
  bool layered = false;  // one layered medium: the 768-thread kernel only
};

namespace {
// (Re)captures the batch: clear every tally and the work counter, the
// transport launch with b->A, the copy reductions.
mcfl_status capture(mcfl_batch* b) {
  mcfl_ctx* c = b->c;
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
  MCFL_HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  mcfl_status st = MCFL_OK;
  const size_t n_jobs = b->spec.size();
  for (size_t j = 0; j < n_jobs && st == MCFL_OK; ++j)
    st = clear_tallies(*b->media[b->spec[j].medium], b->tal[j].get(), c->stream);
  if (st == MCFL_OK && mcfl_fill_u64(b->A.counter, b->A.njobs, c->stream) != 0) st = MCFL_ERR_DEVICE;
  if (st == MCFL_OK && b->A.total > 0 &&
      mcfl_transport_launch(&b->A, b->temporal, b->perjob, b->nblocks, c->stream) != 0)
    st = MCFL_ERR_DEVICE;
  for (size_t j = 0; j < n_jobs && st == MCFL_OK; ++j)
    if (reduce_tallies(*b->media[b->spec[j].medium], b->tal[j].get(), c->stream) != 0) st = MCFL_ERR_DEVICE;
  hipError_t e = hipStreamEndCapture(c->stream, &b->graph);
  if (st != MCFL_OK) return st;
  if (e != hipSuccess) return MCFL_ERR_DEVICE;
  MCFL_HIPCHK(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
  return MCFL_OK;
}
}  // namespace

extern "C" mcfl_status mcfl_batch_create(mcfl_ctx* c, const mcfl_sim* const* media, int n_media, const mcfl_job* jobs,
                                         int n_jobs, mcfl_batch** out) {
  if (!c || !media || n_media <= 0 || !jobs || n_jobs <= 0 || !out) return MCFL_ERR_ARG;
  if (n_jobs > MCFL_MAX_BATCH_JOBS) return MCFL_ERR_CAPACITY;
  *out = nullptr;
  if (hipSetDevice(c->device) != hipSuccess) return MCFL_ERR_DEVICE;
  std::unique_ptr<mcfl_batch> b(new (std::nothrow) mcfl_batch());
  if (!b) return MCFL_ERR_NOMEM;
  b->c = c;
  for (int i = 0; i < n_media; ++i) {
    b->media.emplace_back(new Medium());
    MCFL_STCHK(medium_make(media[i], b->media.back().get(), true));
  }
  uint64_t begin = 0;
  std::vector<bool> used(n_media, false);
  for (int j = 0; j < n_jobs; ++j) {
    const mcfl_job& s = jobs[j];
    if (s.medium < 0 || s.medium >= n_media || s.photons == 0) return MCFL_ERR_ARG;
    if (s.src_kind == MCFL_SOURCE_EMISSION && !s.Fx) return MCFL_ERR_ARG;
    if (s.photons > 0xFFFFFFFFull) return MCFL_ERR_CAPACITY;  // images stay exact below 2^32 photons
    if (s.photons > c->run_photons) b->sliced = true;
    const Medium& m = *b->media[s.medium];
    used[s.medium] = true;
    const bool em = s.src_kind == MCFL_SOURCE_EMISSION;
    if (m.layered()) {
      // a layered slab emits nothing (emitters need albedof < 0, CUDAMCFLmain.cu:520,535), and a source
      // outside every layer has no starting layer (the reference leaves it undefined)
      if (em) return MCFL_ERR_UNSUPPORTED;
      if (layer_at(m, s.zi) == 0) return MCFL_ERR_ARG;
    }
    b->tal.emplace_back(new Tallies());
    Tallies& t = *b->tal.back();
    MCFL_STCHK(alloc_tallies(m, s.tallies, em, c->copies, &t));
    if (em) MCFL_STCHK(build_emission(m, s.Fx, s.photons, &t));
    if (t.flags & MCFL_F_TEMP) b->temporal = 1;
    const uint64_t count = (em && t.al_n == 0) ? 0 : s.photons;
    JobDev J;
    fill_job(J, m, t, s.src_kind, s.xi, s.yi, s.zi, s.src_r, s.seed, begin, count, 0, b->mt.base_of(&m, em));
    b->jd.push_back(J);
    b->spec.push_back(s);
    begin += count;
  }
  int n_used = 0;
  const Medium* uni = nullptr;
  for (int i = 0; i < n_media; ++i)
    if (used[i]) {
      ++n_used;
      uni = b->media[i].get();
    }
  b->perjob = n_used > 1;
  for (int i = 0; i < n_media; ++i)  // the layered kernel runs one medium per launch
    if (used[i] && b->media[i]->layered() && b->perjob) return MCFL_ERR_UNSUPPORTED;
  b->layered = uni->layered();
  MCFL_STCHK(b->L.prepare(c, b->jd, b->mt, b->perjob, uni, &b->A, c->stream));
  plan_tiles(c, b->A, b->jd.data(), b->jd.size(), b->temporal, b->perjob);  // an LDS fhd tile per fhd job
  MCFL_STCHK(b->L.upload(b->jd, c->stream));
  b->nblocks = grid_blocks(c, b->A, b->temporal, b->perjob);
  b->A.n_waves = (uint32_t)b->nblocks * (mcfl_wg_of(&b->A) / 64u);
  if (b->sliced)
    for (size_t j = 0; j < b->jd.size(); ++j)
      MCFL_STCHK(alloc_accumulators(*b->media[b->spec[j].medium], b->tal[j].get(), c->stream));
  MCFL_HIPCHK(hipStreamSynchronize(c->stream));
  MCFL_STCHK(capture(b.get()));
  *out = b.release();
  return MCFL_OK;
}

namespace {
// Launch settings of the batch: workgroup size, fhd tile side (0: LDS budget
// rule), refill threshold, chunk cap; re-places the tiles, uploads the job
// table and re-captures the graph.
mcfl_status apply_settings(mcfl_batch* b, uint32_t wg, int side, uint32_t refill, uint32_t chunk) {
  mcfl_ctx* c = b->c;
  LaunchArgs T = b->A;
  T.wg = wg == MCFL_WG ? 0u : wg;
  plan_tiles(c, T, b->jd.data(), b->jd.size(), b->temporal, b->perjob, side);
  T.refill_min = refill;
  T.chunk_max = chunk;
  const int nblocks = grid_blocks(c, T, b->temporal, b->perjob);
  T.n_waves = (uint32_t)nblocks * (mcfl_wg_of(&T) / 64u);
  MCFL_STCHK(b->L.upload(b->jd, c->stream));
  b->A = T;
  b->nblocks = nblocks;
  return capture(b);
}
}  // namespace

extern "C" mcfl_status mcfl_batch_configure(mcfl_batch* b, const mcfl_tune* in) {
  if (!b || !in) return MCFL_ERR_ARG;
  const uint32_t wg = in->wg ? in->wg : (uint32_t)MCFL_WG;
  if (wg != MCFL_WG && wg != MCFL_WG_ALT1 && wg != MCFL_WG_ALT2) return MCFL_ERR_ARG;
  if (b->layered && wg != MCFL_WG) return MCFL_ERR_UNSUPPORTED;
  if (in->tile_side < 0 || in->refill_min < 1 || in->refill_min > 64 || in->chunk_max < 64 || in->chunk_max > (1u << 20))
    return MCFL_ERR_ARG;
  if (hipSetDevice(b->c->device) != hipSuccess) return MCFL_ERR_DEVICE;
  return apply_settings(b, wg, in->tile_side, in->refill_min, in->chunk_max);
}

extern "C" mcfl_status mcfl_batch_autotune(mcfl_batch* b, double fraction, mcfl_tune* out) {
  if (!b || !(fraction > 0.0) || fraction > 1.0) return MCFL_ERR_ARG;
  mcfl_ctx* c = b->c;
  if (hipSetDevice(c->device) != hipSuccess) return MCFL_ERR_DEVICE;
  // the scaled-down copy of the batch: every job keeps its source, medium and
  // tallies, with fraction of its budget (at least 4096 photons)
  std::vector<JobDev> jd = b->jd;
  uint64_t begin = 0;
  for (JobDev& J : jd) {
    const uint64_t n = std::max<uint64_t>(std::min<uint64_t>(J.count, 4096), (uint64_t)((double)J.count * fraction));
    J.begin = begin;
    J.count = n;
    begin += n;
  }
  DevBuf d_trial;
  struct Release {
    DevBuf& b;
    ~Release() { release(b); }
  } rel{d_trial};
  MCFL_STCHK(ensure(d_trial, jd.size() * sizeof(JobDev)));
  int fhd_job = -1;
  for (size_t j = 0; j < b->jd.size(); ++j)
    if (b->jd[j].fhd != nullptr) {
      fhd_job = (int)j;
      break;
    }
  struct Cand {
    uint32_t wg;  // 0: MCFL_WG
    int side;
    uint32_t refill, chunk;
  };
  auto setup = [&](const Cand& cd, LaunchArgs& T, std::vector<JobDev>& tj, int* nblocks) {
    T = b->A;
    T.wg = cd.wg;
    plan_tiles(c, T, tj.data(), tj.size(), b->temporal, b->perjob, cd.side);
    T.refill_min = cd.refill;
    T.chunk_max = cd.chunk;
    *nblocks = grid_blocks(c, T, b->temporal, b->perjob);
    T.n_waves = (uint32_t)*nblocks * (mcfl_wg_of(&T) / 64u);
  };
  int32_t trials = 0;
  auto trial = [&](const Cand& cd, double* ms) -> mcfl_status {
    LaunchArgs T;
    int nblocks = 0;
    setup(cd, T, jd, &nblocks);
    MCFL_HIPCHK(hipMemcpyAsync(d_trial.p, jd.data(), jd.size() * sizeof(JobDev), hipMemcpyHostToDevice, c->stream));
    T.jobs = d_trial.as<const JobDev>();
    T.total = begin;
    MCFL_HIPCHK(hipMemsetAsync(T.counter, 0, 8 * (size_t)T.njobs, c->stream));
    MCFL_HIPCHK(hipEventRecord(c->ev0, c->stream));
    if (mcfl_transport_launch(&T, b->temporal, b->perjob, nblocks, c->stream) != 0) return MCFL_ERR_DEVICE;
    MCFL_HIPCHK(hipEventRecord(c->ev1, c->stream));
    MCFL_HIPCHK(hipEventSynchronize(c->ev1));
    float f = 0.f;
    MCFL_HIPCHK(hipEventElapsedTime(&f, c->ev0, c->ev1));
    *ms = f;
    ++trials;
    return MCFL_OK;
  };
  double best = 1e300, dflt = 0.0;
  Cand bc{0, 0, c->refill_min, c->chunk_max};
  auto consider = [&](const Cand& cd) -> mcfl_status {  // best of two launches
    double t0 = 0.0, t1 = 0.0;
    MCFL_STCHK(trial(cd, &t0));
    MCFL_STCHK(trial(cd, &t1));
    const double t = std::min(t0, t1);
    if (t < best) {
      best = t;
      bc = cd;
    }
    return MCFL_OK;
  };
  double warm = 0.0;
  MCFL_STCHK(trial(bc, &warm));  // untimed warm-up of the default
  // stage 1: workgroup size, at the default tile rule, refill threshold and chunk cap
  const Cand dc = bc;
  MCFL_STCHK(consider(dc));
  dflt = best;
  if (!b->layered)
    for (uint32_t wg : {(uint32_t)MCFL_WG_ALT1, (uint32_t)MCFL_WG_ALT2})
      MCFL_STCHK(consider({wg, dc.side, dc.refill, dc.chunk}));
  // stage 2: tile side x refill threshold x chunk cap at that workgroup size
  const uint32_t wg = bc.wg;
  const int sides_fhd[] = {0, 16, 32};
  const uint32_t refills[] = {4, 8}, chunks[] = {1024, 256};
  for (int sd : sides_fhd) {
    if (fhd_job < 0 && sd != 0) continue;  // the tile side only matters with an fhd tally
    for (uint32_t r : refills)
      for (uint32_t k : chunks) {
        if (sd == dc.side && r == dc.refill && k == dc.chunk) continue;  // measured in stage 1
        MCFL_STCHK(consider({wg, sd, r, k}));
      }
  }
  // keep the fastest settings and capture the batch with them (trial tallies
  // are cleared by the graph itself)
  MCFL_STCHK(apply_settings(b, bc.wg, bc.side, bc.refill, bc.chunk));
  if (out) {
    out->tile_side = fhd_job >= 0 ? (int32_t)b->A.tnx : 0;
    out->tile_voxels = b->A.tile_n;
    out->refill_min = b->A.refill_min;
    out->chunk_max = b->A.chunk_max;
    out->trial_ms = best;
    out->default_ms = dflt;
    out->trials = trials;
    out->wg = mcfl_wg_of(&b->A);
  }
  return MCFL_OK;
}

namespace {
// A batch in slices of at most S photons per job (see mcfl_batch::sliced):
// the same photon streams as one launch, so the same tallies, except that
// volume tallies and totals saturate at 2^64-1 instead of wrapping.
mcfl_status run_sliced(mcfl_batch* b, mcfl_job* jobs, double* kernel_ms) {
  mcfl_ctx* c = b->c;
  const uint64_t S = c->run_photons;
  const size_t n = b->jd.size();
  uint64_t rounds = 0;
  for (const JobDev& J : b->jd) rounds = std::max<uint64_t>(rounds, (J.count + S - 1) / S);
  for (size_t j = 0; j < n; ++j) {
    const Medium& m = *b->media[b->spec[j].medium];
    Tallies* t = b->tal[j].get();
    MCFL_STCHK(clear_tallies(m, t, c->stream));
    if ((t->fhd_acc.p && mcfl_fill_u64(t->fhd_acc.as<unsigned long long>(), m.nvox, c->stream) != 0) ||
        (t->axyz_acc.p && mcfl_fill_u64(t->axyz_acc.as<unsigned long long>(), m.nvox, c->stream) != 0))
      return MCFL_ERR_DEVICE;
  }
  std::vector<std::vector<uint64_t>> tot(n, std::vector<uint64_t>(MCFL_NTOT, 0));
  std::vector<uint64_t> part(MCFL_NTOT);
  std::vector<JobDev> jr(b->jd);
  double ms_all = 0.0;
  for (uint64_t r = 0; r < rounds; ++r) {
    uint64_t begin = 0;
    for (size_t j = 0; j < n; ++j) {
      const uint64_t full = b->jd[j].count, k0 = r * S;
      jr[j] = b->jd[j];
      jr[j].k0 = k0;
      jr[j].count = full > k0 ? std::min(S, full - k0) : 0;
      jr[j].begin = begin;
      begin += jr[j].count;
    }
    LaunchArgs A = b->A;
    A.total = begin;
    MCFL_HIPCHK(hipMemcpyAsync(b->L.d_jobs.p, jr.data(), n * sizeof(JobDev), hipMemcpyHostToDevice, c->stream));
    if (mcfl_fill_u64(A.counter, A.njobs, c->stream) != 0) return MCFL_ERR_DEVICE;
    MCFL_HIPCHK(hipEventRecord(c->ev0, c->stream));
    if (mcfl_transport_launch(&A, b->temporal, b->perjob, b->nblocks, c->stream) != 0) return MCFL_ERR_DEVICE;
    for (size_t j = 0; j < n; ++j)
      if (jr[j].count > 0 && accumulate_tallies(*b->media[b->spec[j].medium], b->tal[j].get(), c->stream) != 0)
        return MCFL_ERR_DEVICE;
    MCFL_HIPCHK(hipEventRecord(c->ev1, c->stream));
    MCFL_HIPCHK(hipEventSynchronize(c->ev1));
    float ms = 0.f;
    MCFL_HIPCHK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    ms_all += ms;
    for (size_t j = 0; j < n; ++j) {  // totals of the slice, added saturating
      if (jr[j].count == 0) continue;
      Tallies* t = b->tal[j].get();
      MCFL_HIPCHK(hipMemcpy(part.data(), t->tot.p, MCFL_NTOT * 8, hipMemcpyDeviceToHost));
      for (int i = 0; i < MCFL_NTOT; ++i) {
        const uint64_t a = tot[j][i] + part[i];
        tot[j][i] = a < part[i] ? ~0ull : a;
      }
      if (mcfl_fill_u64(t->tot.as<unsigned long long>(), MCFL_NTOT, c->stream) != 0) return MCFL_ERR_DEVICE;
    }
  }
  MCFL_HIPCHK(hipMemcpyAsync(b->L.d_jobs.p, b->jd.data(), n * sizeof(JobDev), hipMemcpyHostToDevice, c->stream));
  MCFL_HIPCHK(hipStreamSynchronize(c->stream));
  if (kernel_ms) *kernel_ms = ms_all;
  for (size_t j = 0; j < n; ++j) {
    mcfl_job& o = jobs[j];
    const Medium& m = *b->media[b->spec[j].medium];
    MCFL_STCHK(download(m, *b->tal[j], o.Rd_xy, o.Tt_xy, o.fhd, o.A_xyz, o.Rd2_xy, o.Tt2_xy, o.time_xyt, nullptr));
    memcpy(o.totals, tot[j].data(), sizeof o.totals);
    o.fl_unit = b->tal[j]->fl_unit;
  }
  return MCFL_OK;
}
}  // namespace

extern "C" mcfl_status mcfl_batch_run(mcfl_batch* b, mcfl_job* jobs, double* kernel_ms) {
  if (!b || !jobs) return MCFL_ERR_ARG;
  mcfl_ctx* c = b->c;
  if (hipSetDevice(c->device) != hipSuccess) return MCFL_ERR_DEVICE;
  bool changed = false;
  for (size_t j = 0; j < b->jd.size(); ++j) {
    const uint64_t key = mcfl_key(jobs[j].seed, b->spec[j].src_kind == MCFL_SOURCE_EMISSION ? 1u : 0u);
    if (key != b->jd[j].key) {
      b->jd[j].key = key;
      changed = true;
    }
  }
  if (b->sliced) return run_sliced(b, jobs, kernel_ms);
  if (changed)
    MCFL_HIPCHK(hipMemcpyAsync(b->L.d_jobs.p, b->jd.data(), b->jd.size() * sizeof(JobDev), hipMemcpyHostToDevice,
                               c->stream));
  MCFL_HIPCHK(hipEventRecord(c->ev0, c->stream));
  MCFL_HIPCHK(hipGraphLaunch(b->exec, c->stream));
  MCFL_HIPCHK(hipEventRecord(c->ev1, c->stream));
  MCFL_HIPCHK(hipEventSynchronize(c->ev1));
  float ms = 0.f;
  MCFL_HIPCHK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
  if (kernel_ms) *kernel_ms = ms;
  for (size_t j = 0; j < b->jd.size(); ++j) {
    mcfl_job& o = jobs[j];
    const Medium& m = *b->media[b->spec[j].medium];
    MCFL_STCHK(download(m, *b->tal[j], o.Rd_xy, o.Tt_xy, o.fhd, o.A_xyz, o.Rd2_xy, o.Tt2_xy, o.time_xyt, o.totals));
    o.fl_unit = b->tal[j]->fl_unit;
  }
  return MCFL_OK;
}

extern "C" void mcfl_batch_destroy(mcfl_batch* b) {
  if (!b) return;
  (void)hipSetDevice(b->c->device);
  (void)hipStreamSynchronize(b->c->stream);
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  delete b;
}

extern "C" mcfl_status mcfl_run_batch(mcfl_ctx* c, const mcfl_sim* const* media, int n_media, mcfl_job* jobs,
                                      int n_jobs, double* kernel_ms) {
  mcfl_batch* b = nullptr;
  MCFL_STCHK(mcfl_batch_create(c, media, n_media, jobs, n_jobs, &b));
  mcfl_status st = mcfl_batch_run(b, jobs, kernel_ms);
  mcfl_batch_destroy(b);
  return st;
}

// ================================================================= server

namespace {

struct Request {
  int64_t id = 0;
  mcfl_request r{};
  int phase = 0;       // 0 queued, 1 excitation, 2 emission, 3 done
  uint64_t done = 0;   // photons of the current phase handed to epochs
  uint64_t finished = 0;  // photons of the current phase whose epochs completed
  uint64_t epochs = 0;
  size_t footprint = 0;
  Medium* m = nullptr;
  std::unique_ptr<Tallies> ex, em;
  std::vector<uint64_t> Rd, Tt, FlR, FlT;
  uint64_t tot[MCFL_NTOT] = {}, fl_tot[MCFL_NTOT] = {};
  double fl_unit = 0.0;
  // A finished phase's tallies are copied out asynchronously (a synchronous
  // copy or reduction would wait for the epoch kernel occupying the GPU):
  // pending until `ev` has passed; pinned staging [images | totals | fhd or
  // emission tables].
  bool pending = false;
  hipEvent_t ev = nullptr;
  uint8_t* stage = nullptr;
  size_t stage_bytes = 0;
  // Excitation of a fluorescence request above the photon split: its fhd
  // copies take at most S photons (mcfl_ctx::run_photons) between two sums
  // into the saturating accumulator ex->fhd_acc.  since_acc counts the
  // photons handed out since the last sum; acc_ev marks the sum, which the
  // next epoch carrying the request waits for (acc_wait).
  uint64_t since_acc = 0;
  hipEvent_t acc_ev = nullptr;
  bool acc_wait = false;
  ~Request() {
    if (ev) (void)hipEventDestroy(ev);
    if (acc_ev) (void)hipEventDestroy(acc_ev);
    if (stage) (void)hipHostFree(stage);  // normally handed back to the server's pool at retirement
  }
  size_t img_bytes() const { return m->npix * 8; }
  uint8_t* stage_tot() const { return stage + 2 * img_bytes(); }
  uint8_t* stage_vol() const { return stage_tot() + MCFL_NTOT * 8; }
};

size_t footprint_of(const Medium& m, const mcfl_request& r, uint32_t copies, uint64_t split) {
  size_t b = 2 * m.npix * 8 + MCFL_NTOT * 8;
  if (r.fluorescence) b += m.nvox * 8 * copies + 2 * m.npix * 8 + MCFL_NTOT * 8 + 3 * 4 * m.nvox;
  if (r.fluorescence && r.photons > split) b += m.nvox * 8;  // the fhd accumulator
  return b;
}

}  // namespace

// One launch of the server: its own job table, counter and stream, and the
// request slices it carries.
struct Epoch {
  Launcher L;
  hipStream_t stream = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  bool busy = false;
  std::vector<std::pair<int64_t, uint64_t>> parts;  // request id, photons
  uint64_t photons = 0;
};

struct mcfl_server {
  mcfl_ctx* c = nullptr;
  uint64_t budget = 0, epoch_photons = 0, in_use = 0;
  BufPool pool;    // declared before the requests, so it outlives their tallies
  HostPool hpool;  // pinned staging of the copy-outs
  uint64_t epoch_target = 20000000;  // photons per epoch to aim for (env MCFL_EPOCH_TARGET)
  double oldest_share = 0.75;        // of an epoch's photons above the floors, the oldest request's (env MCFL_OLDEST_SHARE)
  int reserve = 8;                   // workgroup slots an epoch leaves free (env MCFL_SERVER_RESERVE)
  std::vector<std::unique_ptr<Medium>> media;
  std::deque<std::unique_ptr<Request>> queue;
  std::vector<std::unique_ptr<Request>> active;
  std::map<int64_t, std::unique_ptr<Request>> done;
  int64_t next_id = 1;
  // Two epochs in flight on two streams: the workgroups of the newer launch
  // take the CUs that the older one's last photons leave idle.
  Epoch ep[2];
  int next = 0;              // slot of the next launch
  std::deque<int> inflight;  // slots in launch order
  // cumulative counters; kernel_ms is the device time during which an epoch
  // ran, overlapping epochs counted once: the union of the epochs' [start,
  // end) intervals, measured from the event `origin` (ends at busy_end)
  uint64_t photons = 0, epochs = 0;
  double kernel_ms = 0.0, host_ms = 0.0, busy_end = 0.0;
  hipEvent_t origin = nullptr;
  // host time by part (ms): admission, epoch assembly, launch calls, phase
  // changes to emission, retirements; printed at destruction when the
  // environment sets MCFL_SERVER_PROFILE
  double t_admit = 0.0, t_build = 0.0, t_launch = 0.0, t_emission = 0.0, t_retire = 0.0;
  ~mcfl_server() {
    if (getenv("MCFL_SERVER_PROFILE"))
      fprintf(stderr,
              "mcfl_server: %llu epochs, host %.1f ms (admit %.1f, assemble %.1f, launch %.1f, to emission %.1f, "
              "retire %.1f)\n",
              (unsigned long long)epochs, host_ms, t_admit, t_build, t_launch, t_emission, t_retire);
    for (Epoch& e : ep) {
      if (e.stream) (void)hipStreamSynchronize(e.stream);
      if (e.e0) (void)hipEventDestroy(e.e0);
      if (e.e1) (void)hipEventDestroy(e.e1);
      if (e.stream) (void)hipStreamDestroy(e.stream);
    }
    if (origin) (void)hipEventDestroy(origin);
  }
};

extern "C" mcfl_status mcfl_server_create(mcfl_ctx* c, uint64_t budget, uint64_t epoch_photons, mcfl_server** out) {
  if (!c || !out || epoch_photons == 0) return MCFL_ERR_ARG;
  mcfl_server* s = new (std::nothrow) mcfl_server();
  if (!s) return MCFL_ERR_NOMEM;
  s->c = c;
  s->budget = budget;
  s->pool.limit = budget;  // recycled buffers kept: at most the admission budget again
  s->epoch_photons = epoch_photons;
  if (const char* e = getenv("MCFL_EPOCH_TARGET")) s->epoch_target = strtoull(e, nullptr, 10);
  if (const char* e = getenv("MCFL_OLDEST_SHARE")) s->oldest_share = std::min(1.0, std::max(0.0, atof(e)));
  if (const char* e = getenv("MCFL_SERVER_RESERVE")) s->reserve = std::max(0, atoi(e));
  if (hipSetDevice(c->device) != hipSuccess) {
    delete s;
    return MCFL_ERR_DEVICE;
  }
  for (Epoch& e : s->ep)
    if (hipStreamCreateWithFlags(&e.stream, hipStreamNonBlocking) != hipSuccess || hipEventCreate(&e.e0) != hipSuccess ||
        hipEventCreate(&e.e1) != hipSuccess) {
      delete s;
      return MCFL_ERR_DEVICE;
    }
  if (hipEventCreate(&s->origin) != hipSuccess || hipEventRecord(s->origin, s->ep[0].stream) != hipSuccess ||
      hipEventSynchronize(s->origin) != hipSuccess) {
    delete s;
    return MCFL_ERR_DEVICE;
  }
  *out = s;
  return MCFL_OK;
}

extern "C" mcfl_status mcfl_server_add_medium(mcfl_server* s, const mcfl_sim* sim, int32_t* id) {
  if (!s || !sim || !id) return MCFL_ERR_ARG;
  if (hipSetDevice(s->c->device) != hipSuccess) return MCFL_ERR_DEVICE;
  std::unique_ptr<Medium> m(new (std::nothrow) Medium());
  if (!m) return MCFL_ERR_NOMEM;
  MCFL_STCHK(medium_make(sim, m.get(), false));
  if (m->mat_ex.size() > MCFL_MAX_MATERIALS / 2) return MCFL_ERR_CAPACITY;
  s->media.push_back(std::move(m));
  *id = (int32_t)s->media.size() - 1;
  return MCFL_OK;
}

extern "C" mcfl_status mcfl_server_submit(mcfl_server* s, const mcfl_request* r, int64_t* id) {
  if (!s || !r || !id || r->photons == 0 || r->medium < 0 || r->medium >= (int32_t)s->media.size())
    return MCFL_ERR_ARG;
  if (r->src_kind == MCFL_SOURCE_EMISSION) return MCFL_ERR_ARG;
  std::unique_ptr<Request> q(new (std::nothrow) Request());
  if (!q) return MCFL_ERR_NOMEM;
  q->id = s->next_id++;
  q->r = *r;
  q->m = s->media[r->medium].get();
  if (r->photons > 0xFFFFFFFFull) return MCFL_ERR_CAPACITY;  // images and totals stay exact below 2^32 photons
  q->footprint = footprint_of(*q->m, *r, s->c->copies, s->c->run_photons);
  if (q->footprint > s->budget) return MCFL_ERR_CAPACITY;
  *id = q->id;
  s->queue.push_back(std::move(q));
  return MCFL_OK;
}

namespace {

mcfl_status admit(mcfl_server* s) {
  while (!s->queue.empty() && s->active.size() < MCFL_MAX_JOBS) {
    Request& q = *s->queue.front();
    if (s->in_use + q.footprint > s->budget) break;  // admission by memory footprint, FIFO
    q.ex.reset(new Tallies());
    q.ex->pool = &s->pool;
    // cleared on the stream of the next launch, the first one that can carry the request
    const hipStream_t st = s->ep[s->next].stream;
    MCFL_STCHK(alloc_tallies(*q.m, q.r.fluorescence ? (uint32_t)MCFL_TALLY_FHD : 0u, false, s->c->copies, q.ex.get()));
    MCFL_STCHK(clear_tallies(*q.m, q.ex.get(), st));
    if (q.r.fluorescence && q.r.photons > s->c->run_photons) MCFL_STCHK(alloc_accumulators(*q.m, q.ex.get(), st));
    if (q.r.fluorescence) {
      q.em.reset(new Tallies());
      q.em->pool = &s->pool;
      MCFL_STCHK(alloc_tallies(*q.m, 0u, true, s->c->copies, q.em.get()));
      MCFL_STCHK(clear_tallies(*q.m, q.em.get(), st));
    }
    q.phase = 1;
    s->in_use += q.footprint;
    s->active.push_back(std::move(s->queue.front()));
    s->queue.pop_front();
  }
  return MCFL_OK;
}

// A phase's photons are all done: queue the copy-out of its tallies (after
// summing the fhd copies when the emission pass follows) on the stream of the
// epoch that finished it, and mark the request pending.
mcfl_status start_copy_out(HostPool& pool, Request& q, hipStream_t st) {
  const Medium& m = *q.m;
  const bool to_emission = q.phase == 1 && q.r.fluorescence;
  Tallies& t = q.phase == 1 ? *q.ex : *q.em;
  const size_t need = 2 * q.img_bytes() + MCFL_NTOT * 8 + (q.r.fluorescence ? std::max(m.nvox * 8, m.nvox * 12) : 0);
  if (q.stage_bytes < need) {
    pool.put(&q.stage, &q.stage_bytes);
    MCFL_STCHK(pool.get(&q.stage, &q.stage_bytes, need));
  }
  if (!q.ev) MCFL_HIPCHK(hipEventCreateWithFlags(&q.ev, hipEventDisableTiming));
  if (to_emission) {
    if (t.fhd_acc.p) {  // the last slice into the accumulator
      if (accumulate_tallies(m, &t, st) != 0) return MCFL_ERR_DEVICE;
      MCFL_HIPCHK(hipMemcpyAsync(q.stage_vol(), t.fhd_acc.p, m.nvox * 8, hipMemcpyDeviceToHost, st));
    } else {
      if (reduce_tallies(m, &t, st) != 0) return MCFL_ERR_DEVICE;
      MCFL_HIPCHK(hipMemcpyAsync(q.stage_vol(), t.fhd.p, m.nvox * 8, hipMemcpyDeviceToHost, st));
    }
  }
  MCFL_HIPCHK(hipMemcpyAsync(q.stage, t.Rd.p, q.img_bytes(), hipMemcpyDeviceToHost, st));
  MCFL_HIPCHK(hipMemcpyAsync(q.stage + q.img_bytes(), t.Tt.p, q.img_bytes(), hipMemcpyDeviceToHost, st));
  MCFL_HIPCHK(hipMemcpyAsync(q.stage_tot(), t.tot.p, MCFL_NTOT * 8, hipMemcpyDeviceToHost, st));
  MCFL_HIPCHK(hipEventRecord(q.ev, st));
  q.pending = true;
  return MCFL_OK;
}

// The copy-out has landed: keep the phase's results, then start the emission
// pass (its sampling tables uploaded on `up`, the stream of the next launch)
// or retire the request.
mcfl_status finish_copy_out(mcfl_server* s, Request& q, hipStream_t up) {
  const Medium& m = *q.m;
  q.pending = false;
  const size_t npix = m.npix;
  if (q.phase == 1) {
    q.Rd.assign((const uint64_t*)q.stage, (const uint64_t*)q.stage + npix);
    q.Tt.assign((const uint64_t*)(q.stage + q.img_bytes()), (const uint64_t*)(q.stage + q.img_bytes()) + npix);
    memcpy(q.tot, q.stage_tot(), sizeof q.tot);
    if (q.r.fluorescence) {
      const uint64_t* fhd = (const uint64_t*)q.stage_vol();
      std::vector<double> Fx(m.nvox);
      const double norm = 4294967295.0 * (double)q.tot[MCFL_TOT_PHOTONS];  // CUDAMCFLmain.cu:181-183
      for (size_t v = 0; v < m.nvox; ++v) Fx[v] = (double)fhd[v] / norm;
      MCFL_STCHK(build_emission(m, Fx.data(), q.r.photons, q.em.get(), up, q.stage_vol()));
      q.fl_unit = q.em->fl_unit;
      q.phase = 2;
      q.done = 0;
      q.finished = 0;
      if (q.em->al_n > 0) return MCFL_OK;
    }
  } else {
    q.FlR.assign((const uint64_t*)q.stage, (const uint64_t*)q.stage + npix);
    q.FlT.assign((const uint64_t*)(q.stage + q.img_bytes()), (const uint64_t*)(q.stage + q.img_bytes()) + npix);
    memcpy(q.fl_tot, q.stage_tot(), sizeof q.fl_tot);
  }
  q.phase = 3;
  q.ex.reset();
  q.em.reset();
  s->hpool.put(&q.stage, &q.stage_bytes);
  s->in_use -= q.footprint;
  return MCFL_OK;
}

}  // namespace

extern "C" mcfl_status mcfl_server_step(mcfl_server* s, int64_t* done_ids, int32_t max_done, int32_t* n_done,
                                        double* kernel_ms) {
  if (!s) return MCFL_ERR_ARG;
  mcfl_ctx* c = s->c;
  if (hipSetDevice(c->device) != hipSuccess) return MCFL_ERR_DEVICE;
  if (n_done) *n_done = 0;
  if (kernel_ms) *kernel_ms = 0.0;
  const auto t_host0 = std::chrono::steady_clock::now();
  double t_wait = 0.0;
  auto since = [](std::chrono::steady_clock::time_point t) {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t).count();
  };
  MCFL_STCHK(admit(s));
  s->t_admit += since(t_host0);
  // 0. Copy-outs that have landed move their requests on (emission or retirement).
  auto finish_landed = [&](bool wait) -> mcfl_status {
    for (std::unique_ptr<Request>& q : s->active) {
      if (!q->pending) continue;
      if (wait) MCFL_HIPCHK(hipEventSynchronize(q->ev));
      const hipError_t e = hipEventQuery(q->ev);
      if (e == hipErrorNotReady) continue;
      if (e != hipSuccess) return MCFL_ERR_DEVICE;
      const auto ta = std::chrono::steady_clock::now();
      const bool to_emission = q->phase == 1 && q->r.fluorescence;
      MCFL_STCHK(finish_copy_out(s, *q, s->ep[s->next].stream));
      (to_emission ? s->t_emission : s->t_retire) += since(ta);
    }
    return MCFL_OK;
  };
  MCFL_STCHK(finish_landed(false));
  const auto t_build0 = std::chrono::steady_clock::now();

  // 1. Launch the next epoch if its slot is free.  An epoch carries about
  // epoch_target photons (tens of milliseconds of device time): every active request
  // with photons left in its phase gets at least epoch_photons of them; of
  // the rest of the target the oldest request (first submitted) takes the
  // fraction oldest_share and the remainder goes to the requests with the
  // least work left first (shortest remaining work first).  Small requests
  // finish in the next epoch while the oldest one runs at most of the GPU's
  // rate, so a large request is not starved behind a stream of small ones
  // (shortest-first alone leaves it last whenever anything smaller is open).
  bool launched = false;
  Epoch& E = s->ep[s->next];
  std::vector<Request*> open;
  for (const std::unique_ptr<Request>& qp : s->active)
    if ((qp->phase == 1 || qp->phase == 2) && qp->done < qp->r.photons) open.push_back(qp.get());
  auto left_of = [](const Request* q) {  // photons left in the whole request
    return (q->r.photons - q->done) + (q->phase == 1 && q->r.fluorescence ? q->r.photons : 0);
  };
  std::stable_sort(open.begin(), open.end(), [&](const Request* a, const Request* b) { return left_of(a) < left_of(b); });
  std::map<int64_t, uint64_t> share;
  {
    uint64_t floors = 0;
    const Request* oldest = nullptr;
    for (const Request* q : open) {
      const uint64_t fl = std::min<uint64_t>(q->r.photons - q->done, s->epoch_photons);
      floors += fl;
      share[q->id] = fl;
      if (!oldest || q->id < oldest->id) oldest = q;
    }
    uint64_t extra = s->epoch_target > floors ? s->epoch_target - floors : 0;
    if (oldest) {
      const uint64_t rem = oldest->r.photons - oldest->done - share[oldest->id];
      const uint64_t add = std::min<uint64_t>(rem, (uint64_t)(s->oldest_share * (double)extra));
      share[oldest->id] += add;
      extra -= add;
    }
    for (const Request* q : open) {
      const uint64_t rem = q->r.photons - q->done - share[q->id];
      const uint64_t add = std::min<uint64_t>(rem, extra);
      extra -= add;
      share[q->id] += add;
    }
  }
  if (!E.busy) {
    std::vector<JobDev> jd;
    MatTable mt;
    E.parts.clear();
    const Medium* uni = nullptr;
    bool perjob = false;
    uint64_t begin = 0;
    std::vector<hipEvent_t> waits;
    for (std::unique_ptr<Request>& qp : s->active) {
      Request& q = *qp;
      if ((q.phase != 1 && q.phase != 2) || q.done >= q.r.photons) continue;
      uint64_t take = share[q.id];
      const bool acc = q.phase == 1 && q.ex->fhd_acc.p;
      if (acc) {  // at most S photons into the fhd copies between two sums
        take = std::min(take, c->run_photons - std::min(c->run_photons, q.since_acc));
        if (take == 0) continue;  // waits for the sum queued when its last epoch completes
        q.since_acc += take;
        if (q.acc_wait) {
          waits.push_back(q.acc_ev);
          q.acc_wait = false;
        }
      }
      if (!uni) uni = q.m;
      if (q.m != uni) perjob = true;
      const bool em = q.phase == 2;
      Tallies& t = em ? *q.em : *q.ex;
      JobDev J;
      fill_job(J, *q.m, t, em ? MCFL_SOURCE_EMISSION : q.r.src_kind, q.r.xi, q.r.yi, q.r.zi, q.r.src_r, q.r.seed,
               begin, take, q.done, mt.base_of(q.m, em));
      jd.push_back(J);
      E.parts.emplace_back(q.id, take);
      begin += take;
      q.done += take;
      ++q.epochs;
    }
    if (!jd.empty()) {
      LaunchArgs A;
      MCFL_STCHK(E.L.prepare(c, jd, mt, perjob, uni, &A, E.stream));
      s->t_build += since(t_build0);
      const auto t_launch0 = std::chrono::steady_clock::now();
      plan_tiles(c, A, jd.data(), jd.size(), 0, perjob ? 1 : 0);  // an LDS fhd tile per request tallying fhd
      MCFL_STCHK(E.L.upload(jd, E.stream));
      // leave `reserve` workgroup slots free, so the copy-outs and sums queued
      // behind a finished epoch run beside the one still in flight
      const int nblocks = std::max(1, grid_blocks(c, A, 0, perjob ? 1 : 0) - s->reserve);
      A.n_waves = (uint32_t)nblocks * (mcfl_wg_of(&A) / 64u);
      MCFL_HIPCHK(hipMemsetAsync(A.counter, 0, 8 * jd.size(), E.stream));
      for (hipEvent_t w : waits) MCFL_HIPCHK(hipStreamWaitEvent(E.stream, w, 0));
      MCFL_HIPCHK(hipEventRecord(E.e0, E.stream));
      if (mcfl_transport_launch(&A, 0, perjob ? 1 : 0, nblocks, E.stream) != 0) return MCFL_ERR_DEVICE;
      MCFL_HIPCHK(hipEventRecord(E.e1, E.stream));
      s->t_launch += since(t_launch0);
      E.photons = A.total;
      E.busy = true;
      s->inflight.push_back(s->next);
      s->next ^= 1;
      launched = true;
    }
  }

  // 2. Complete the oldest epoch when both slots are busy or nothing new
  // could start, then move finished phases on and retire finished requests.
  int32_t nd = 0;
  if (!s->inflight.empty() && (s->inflight.size() == 2 || !launched)) {
    Epoch& O = s->ep[s->inflight.front()];
    s->inflight.pop_front();
    const auto tw0 = std::chrono::steady_clock::now();
    MCFL_HIPCHK(hipEventSynchronize(O.e1));
    t_wait = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - tw0).count();
    float ms = 0.f, t_start = 0.f, t_end = 0.f;
    MCFL_HIPCHK(hipEventElapsedTime(&ms, O.e0, O.e1));
    MCFL_HIPCHK(hipEventElapsedTime(&t_start, s->origin, O.e0));
    MCFL_HIPCHK(hipEventElapsedTime(&t_end, s->origin, O.e1));
    if (kernel_ms) *kernel_ms = ms;
    s->photons += O.photons;
    s->epochs += 1;
    // epochs complete in launch order; only the part after the busy time so far counts
    s->kernel_ms += std::max(0.0, (double)t_end - std::max((double)t_start, s->busy_end));
    s->busy_end = std::max(s->busy_end, (double)t_end);
    O.busy = false;
    for (const auto& pt : O.parts)
      for (std::unique_ptr<Request>& qp : s->active)
        if (qp->id == pt.first) qp->finished += pt.second;
    // requests whose fhd copies hold S photons since their last sum, with no
    // epoch in flight still adding to them and photons left to hand out:
    // sum the copies into the accumulator on this epoch's stream
    for (const auto& pt : O.parts)
      for (std::unique_ptr<Request>& qp : s->active) {
        Request& q = *qp;
        if (q.id != pt.first || q.phase != 1 || !q.ex->fhd_acc.p || q.since_acc < c->run_photons ||
            q.done >= q.r.photons)
          continue;
        bool in_flight = false;
        for (int slot : s->inflight)
          for (const auto& op : s->ep[slot].parts) in_flight |= op.first == q.id;
        if (in_flight) continue;
        if (accumulate_tallies(*q.m, q.ex.get(), O.stream) != 0) return MCFL_ERR_DEVICE;
        if (!q.acc_ev) MCFL_HIPCHK(hipEventCreateWithFlags(&q.acc_ev, hipEventDisableTiming));
        MCFL_HIPCHK(hipEventRecord(q.acc_ev, O.stream));
        q.acc_wait = true;
        q.since_acc = 0;
      }
    for (std::unique_ptr<Request>& q : s->active)
      if ((q->phase == 1 || q->phase == 2) && !q->pending && q->finished >= q->r.photons) {
        const auto ta = std::chrono::steady_clock::now();
        MCFL_STCHK(start_copy_out(s->hpool, *q, O.stream));
        s->t_retire += since(ta);
      }
    // the copies just queued normally land within a millisecond (they run
    // beside the epoch in flight): give them that long, so this step can
    // already report the requests they finish
    const auto tq = std::chrono::steady_clock::now();
    for (;;) {
      bool waiting = false;
      for (const std::unique_ptr<Request>& q : s->active)
        if (q->pending && hipEventQuery(q->ev) == hipErrorNotReady) waiting = true;
      if (!waiting || since(tq) > 2.0) break;
    }
    MCFL_STCHK(finish_landed(false));
  }
  // Nothing on the device but copy-outs: wait for them rather than spin.
  if (s->inflight.empty()) MCFL_STCHK(finish_landed(true));
  std::vector<std::unique_ptr<Request>> still;
  for (std::unique_ptr<Request>& q : s->active) {
    if (q->phase == 3) {
      if (done_ids && nd < max_done) done_ids[nd] = q->id;
      ++nd;
      s->done[q->id] = std::move(q);
    } else {
      still.push_back(std::move(q));
    }
  }
  s->active.swap(still);
  if (n_done) *n_done = nd;
  s->host_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_host0).count() - t_wait;
  return MCFL_OK;
}

extern "C" mcfl_status mcfl_server_counters(mcfl_server* s, uint64_t* photons, uint64_t* epochs, double* kernel_ms,
                                            double* host_ms) {
  if (!s) return MCFL_ERR_ARG;
  if (photons) *photons = s->photons;
  if (epochs) *epochs = s->epochs;
  if (kernel_ms) *kernel_ms = s->kernel_ms;
  if (host_ms) *host_ms = s->host_ms;
  return MCFL_OK;
}

extern "C" mcfl_status mcfl_server_collect(mcfl_server* s, int64_t id, mcfl_request_result* out) {
  if (!s || !out) return MCFL_ERR_ARG;
  auto it = s->done.find(id);
  if (it == s->done.end()) return MCFL_ERR_ARG;
  Request& q = *it->second;
  const size_t npix = q.m->npix * 8;
  if (out->Rd_xy) memcpy(out->Rd_xy, q.Rd.data(), npix);
  if (out->Tt_xy) memcpy(out->Tt_xy, q.Tt.data(), npix);
  if (out->FlR_raw && !q.FlR.empty()) memcpy(out->FlR_raw, q.FlR.data(), npix);
  if (out->FlT_raw && !q.FlT.empty()) memcpy(out->FlT_raw, q.FlT.data(), npix);
  out->fl_unit = q.fl_unit;
  memcpy(out->totals, q.tot, sizeof q.tot);
  memcpy(out->fl_totals, q.fl_tot, sizeof q.fl_tot);
  out->epochs = q.epochs;
  s->done.erase(it);
  return MCFL_OK;
}

extern "C" mcfl_status mcfl_server_stats(mcfl_server* s, int32_t* queued, int32_t* active, uint64_t* bytes) {
  if (!s) return MCFL_ERR_ARG;
  if (queued) *queued = (int32_t)s->queue.size();
  if (active) *active = (int32_t)s->active.size();
  if (bytes) *bytes = s->in_use;
  return MCFL_OK;
}

extern "C" void mcfl_server_destroy(mcfl_server* s) {
  if (!s) return;
  (void)hipSetDevice(s->c->device);
  (void)hipStreamSynchronize(s->c->stream);
  delete s;  // waits for both epoch streams
}
