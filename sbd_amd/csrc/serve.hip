// Serving engine: many independent TPB solves on one GPU with continuous
// batching.
//
// Requests arrive at any time (sbd_server_submit).  An admission thread takes
// them in arrival order (or smallest first), reserves their device footprint
// against the budget and builds their context (tables on the device).
//
//   * Single-root FP64 solves are batched: admission also makes their
//     diagonal, start vector and Davidson buffers and hands each to one of a
//     few engine lanes (a thread with its own stream), balancing the lanes by
//     estimated work per step so that light solves do not wait each round for
//     a heavy one.  A lane advances every solve it holds one Davidson pass at
//     a time (DavidsonJob, davidson_job.hpp) and runs each kind of pass for
//     all of them in one launch: one sigma launch (k_sigma_batch, or the
//     solve's own launch where the batched kernel does not apply), one Ritz
//     launch, one Gram-Schmidt launch per round; the lanes' launches overlap
//     on the device.  Each solve takes the same kernels, grids and fixed-order
//     sums in a batch as alone, so its iterates -- and its energy -- are
//     bit-identical alone, in any mix, and to sbd_diag with the same options.
//   * Requests the batched form does not take (precision 1: FP32 Davidson +
//     FP64 sweep; nroots > 1: block Davidson) run whole on solo workers, each
//     a thread with its own stream set as the context's stream, through
//     sbd_diag / sbd_diag_roots: concurrent with the batches and with each
//     other, and equal to those calls.
//
// The batched solves' Davidson vectors are blocks of one device arena
// reserved when the server is created (up to the memory budget), so admission
// and retirement make no hipMalloc / hipFree for them.  Finished solves
// publish their result at once and hand their block and context to a reaper
// thread (the context's tables are freed with hipFree, which waits for the
// device; doing it on the lanes would stall every lane behind it).  The memory
// returns to the budget when it has been freed, and queued requests are
// admitted as memory, arena space and slots allow, so their work joins the
// next round.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <iterator>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "davidson_job.hpp"
#include "sbd_internal.hpp"

using namespace sbdx;

namespace {

constexpr int kServerLanes = 8;  // engine lanes (threads + streams) of batched solves
constexpr int kSoloWorkers = 4;  // threads + streams of whole (FP32 / multi-root) solves

double clock_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// The device memory of the batched solves' Davidson vectors: one allocation
// made when the server is created, handed out first-fit in 2 MiB-aligned
// blocks (one block per solve) and returned on completion, so admitting and
// retiring a solve calls neither hipMalloc nor hipFree for its vectors.
struct Arena {
  static constexpr size_t kAlign = size_t(1) << 21;
  char* base = nullptr;
  size_t cap = 0;
  std::map<size_t, size_t> free_;  // offset -> bytes, coalesced
  std::mutex m;

  static size_t round_up(size_t b) { return (b + kAlign - 1) / kAlign * kAlign; }
  // reserve up to want bytes (less when the device has less free memory)
  void init(size_t want) {
    size_t fr = 0, tot = 0;
    if (hipMemGetInfo(&fr, &tot) == hipSuccess) want = std::min(want, fr / 10 * 9);
    for (size_t b = round_up(want); b >= (size_t(1) << 30); b = round_up(b / 2)) {
      if (hipMalloc(&base, b) == hipSuccess) {
        cap = b;
        free_[0] = b;
        return;
      }
      (void)hipGetLastError();
      base = nullptr;
    }
  }
  size_t largest() {
    std::lock_guard<std::mutex> g(m);
    size_t mx = 0;
    for (auto& f : free_) mx = std::max(mx, f.second);
    return mx;
  }
  void* alloc(size_t bytes) {
    bytes = round_up(bytes);
    std::lock_guard<std::mutex> g(m);
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second < bytes) continue;
      const size_t off = it->first, rest = it->second - bytes;
      free_.erase(it);
      if (rest) free_[off + bytes] = rest;
      return base + off;
    }
    return nullptr;
  }
  void release(void* p, size_t bytes) {
    bytes = round_up(bytes);
    std::lock_guard<std::mutex> g(m);
    size_t off = (size_t)((char*)p - base);
    auto next = free_.lower_bound(off);
    if (next != free_.end() && off + bytes == next->first) {
      bytes += next->second;
      next = free_.erase(next);
    }
    if (next != free_.begin()) {
      auto prev = std::prev(next);
      if (prev->first + prev->second == off) {
        off = prev->first;
        bytes += prev->second;
        free_.erase(prev);
      }
    }
    free_[off] = bytes;
  }
  void destroy() {
    if (base) (void)hipFree(base);
    base = nullptr;
    cap = 0;
    free_.clear();
  }
};

// bytes of one batched solve's vectors (C, HC, spare, hii) in the arena
size_t solve_block_bytes(uint64_t N, int nb) {
  const size_t v = (size_t)(8 * N + 255) / 256 * 256;
  return v * (size_t)(2 * nb + 2);
}

struct Request {
  uint64_t id = 0;
  int norb = 0;
  double I0 = 0.0;
  std::vector<double> h1, eri;
  std::vector<uint64_t> a, b;
  sbd_options opt{};
  int nroots = 1;
  double t_submit = 0.0;
  uint64_t bytes = 0;  // reserved device footprint
  bool batched() const { return opt.precision == 0 && nroots == 1; }
};

// One admitted solve with its device buffers (batched solves only).
struct Solve {
  Request req;
  sbd_ctx* ctx = nullptr;
  std::vector<double*> C, HC;
  double* spare = nullptr;
  double* hii = nullptr;
  Arena* arena = nullptr;  // set when the vectors are a block of the arena
  void* block = nullptr;
  size_t block_bytes = 0;
  DavidsonJob<double> job;
  double t_admit = 0.0;
  double cost = 0.0;  // estimated work of one step (terms), for lane balancing
  int lane = 0;
  int status = SBD_OK;

  void release() {
    if (block) {
      arena->release(block, block_bytes);
      block = nullptr;
    } else {
      for (double* p : C)
        if (p) (void)hipFree(p);
      for (double* p : HC)
        if (p) (void)hipFree(p);
      if (spare) (void)hipFree(spare);
      if (hii) (void)hipFree(hii);
    }
    C.clear();
    HC.clear();
    spare = hii = nullptr;
    if (ctx) sbd_ctx_destroy(ctx);
    ctx = nullptr;
  }
};

// Vectors a request's solve holds at its peak, in units of N FP64 elements:
// batched and single-root solves 2 max_nb + 2 (C, HC, spare, hii); block
// Davidson 2 M + 5 R + 2 (C, HC, restart buffers, the root vectors, hii and
// H W), M = the subspace bound of sbd_diag_roots.
double vectors_of(const Request& r) {
  const double nb = (double)r.opt.max_nb;
  if (r.nroots <= 1) return 2 * nb + 2;
  const double R = (double)r.nroots;
  const double M = std::min(std::max(nb, 3 * R), (double)kMaxVecs);
  return 2 * M + 5 * R + 2;
}

// Device bytes of one solve before its tables exist: its vectors, integrals,
// and tables and caches estimated for random strings (each possible
// excitation partner present with probability strings / C(norb, n)).
// Corrected to the exact figure once the tables are built.
uint64_t footprint(const Request& r) {
  const uint64_t N = (uint64_t)r.a.size() * r.b.size();
  const uint64_t L = (uint64_t)r.norb;
  double bytes = vectors_of(r) * N * 8;
  bytes += 8.0 * (4 * L * L + r.eri.size() + 2 * L * L + 4 * L * L * L + L * L * L * L);
  auto est = [&](const std::vector<uint64_t>& s, double& singles) {
    const double n = s.empty() ? 0 : (double)__builtin_popcountll(s[0]);
    const double v1 = n * (L - n), v2 = n * (n - 1) / 2 * (L - n) * (L - n - 1) / 2;
    double comb = 1.0;
    for (double k = 0; k < n; k++) comb = comb * (L - k) / (k + 1);
    const double fill = std::min(1.0, (double)s.size() / comb);
    singles = s.size() * v1 * fill;
    return s.size() * (v1 + v2) * fill;
  };
  double sa = 0, sb = 0;
  const double recs = est(r.a, sa) + est(r.b, sb);
  bytes += 64.0 * recs + 16.0 * (r.a.size() + r.b.size() + 2);
  bytes += 8.0 * (sa * r.b.size() + r.a.size() * sb);
  return (uint64_t)bytes;
}

// The same after the tables are built, from the actual counts.
uint64_t footprint_built(const sbd_ctx* c, const Request& r) {
  const uint64_t N = c->N, L = (uint64_t)c->norb;
  uint64_t bytes = (uint64_t)(vectors_of(r) * (double)N * 8);
  bytes += 8 * (4 * L * L + c->eri.size() + 2 * L * L + 4 * L * L * L);
  if (c->d_Vab) bytes += 8 * L * L * L * L;
  bytes += 16 * (c->ta.n_s + c->ta.n_d + c->tb.n_s + c->tb.n_d);  // records
  bytes += 2 * 36 * (c->tb.n_d + c->nb);                          // ELL copies (bounded)
  bytes += 8 * 8 * (c->na + c->nb + 2);                           // offsets, orders, strings
  bytes += 4 * (uint64_t)c->tb.max_s * c->nb;                     // dense-pass records
  if (c->d_Ea) bytes += 8 * c->ta.n_s * c->nb;
  if (c->d_Eb) bytes += 8 * c->na * c->tb.n_s;
  return bytes;
}

}  // namespace

// One engine lane: its own thread, stream and scratch, advancing the batched
// solves assigned to it.
struct Lane {
  std::deque<Solve*> ready;
  hipStream_t st = nullptr;
  sbdx::BatchScratch sc;
  std::thread th;
  double load = 0.0;
  std::vector<double> costs;  // step cost of each solve the lane holds
  double max_cost() const {
    double m = 0.0;
    for (double c : costs) m = std::max(m, c);
    return m;
  }
};

// One solo worker: a thread and a stream running whole solves.
struct Solo {
  hipStream_t st = nullptr;
  std::thread th;
};

struct sbd_server {
  int device = 0;
  uint64_t budget = 0, used = 0;
  int max_active = 0;
  mutable std::mutex mu;
  std::condition_variable cv_admit, cv_engine, cv_done, cv_solo, cv_reap;
  std::deque<Request> pending;
  std::vector<Lane> lanes;
  std::vector<Solo> solos;
  std::deque<Solve*> solo_ready;  // prepared whole solves (context only)
  std::deque<Solve*> reap;        // finished solves whose memory is to be freed
  int preparing = 0, running = 0, queued = 0;  // queued: prepared, not yet running
  int reaping = 0;
  std::map<uint64_t, sbd_result> results;
  std::map<uint64_t, int> state;  // 0 queued, 1 admitted, 2 done
  bool stopping = false;
  uint64_t next_id = 1;
  std::thread admitter, reaper;
  hipStream_t st_admit = nullptr;
  Arena arena;
  double t0 = 0.0;
  uint64_t rounds = 0, peak_active = 0;
  std::string last_error;
  // admission order: 0 = arrival order; 1 = smallest footprint first, except
  // that a request waiting age_limit seconds or more goes first (no request
  // waits behind smaller ones for ever)
  int policy = 0;
  double age_limit = 30.0;

  double now() const { return clock_s() - t0; }

  // index of the pending request to admit next (mu held, pending not empty;
  // pending is in arrival order)
  size_t pick_pending() const {
    if (policy == 0 || now() - pending.front().t_submit >= age_limit) return 0;
    size_t best = 0;
    for (size_t i = 1; i < pending.size(); i++)
      if (pending[i].bytes < pending[best].bytes) best = i;
    return best;
  }

  // Publish the result now; the memory goes back to the budget once the
  // reaper has freed it.
  void finish(Solve* s, const sbd_result& res) {
    sbd_result r = res;
    r.t_submit = s->req.t_submit;
    r.t_admit = s->t_admit;
    r.t_done = now();
    r.status = s->status;
    std::lock_guard<std::mutex> g(mu);
    if (s->req.batched()) {
      lanes[s->lane].load -= s->cost;
      auto& cs = lanes[s->lane].costs;
      auto it = std::find(cs.begin(), cs.end(), s->cost);
      if (it != cs.end()) cs.erase(it);
    }
    results[s->req.id] = r;
    state[s->req.id] = 2;
    running--;
    reaping++;
    reap.push_back(s);
    cv_done.notify_all();
    cv_reap.notify_all();
  }

  void reap_loop() {
    (void)hipSetDevice(device);
    for (;;) {
      Solve* s = nullptr;
      {
        std::unique_lock<std::mutex> g(mu);
        cv_reap.wait(g, [&] { return !reap.empty() || (stopping && reaping == 0); });
        if (reap.empty()) return;
        s = reap.front();
        reap.pop_front();
      }
      s->release();
      std::lock_guard<std::mutex> g(mu);
      used -= s->req.bytes;
      reaping--;
      delete s;
      cv_admit.notify_all();
      cv_reap.notify_all();
    }
  }

  // ---- admission: context and tables; for batched solves also the diagonal,
  // start vector and Davidson buffers
  int prepare(Solve* s) {
    const Request& q = s->req;
    int st0 = 0;
    s->ctx = sbd_ctx_create(q.norb, q.I0, q.h1.data(), q.eri.data(), q.a.data(), q.a.size(),
                            q.b.data(), q.b.size(), &st0);
    if (!s->ctx) return st0 ? st0 : SBD_ERR_DEVICE;
    int rc = sbd_make_helpers(s->ctx);
    if (rc) return rc;
    if (!q.batched()) return SBD_OK;
    const uint64_t N = s->ctx->N;
    const int nb = q.opt.max_nb;
    s->C.assign(nb, nullptr);
    s->HC.assign(nb, nullptr);
    const size_t bytes = solve_block_bytes(N, nb);
    if (arena.base) {
      s->block = arena.alloc(bytes);
      if (!s->block) return SBD_ERR_MEMORY;
      s->arena = &arena;
      s->block_bytes = bytes;
      const size_t v = bytes / (size_t)(2 * nb + 2);
      char* b = (char*)s->block;
      for (int k = 0; k < nb; k++) {
        s->C[k] = (double*)(b + (size_t)(2 * k) * v);
        s->HC[k] = (double*)(b + (size_t)(2 * k + 1) * v);
      }
      s->spare = (double*)(b + (size_t)(2 * nb) * v);
      s->hii = (double*)(b + (size_t)(2 * nb + 1) * v);
    } else {
      for (int k = 0; k < nb; k++) {
        if (hipMalloc(&s->C[k], 8 * N) != hipSuccess || hipMalloc(&s->HC[k], 8 * N) != hipSuccess)
          return SBD_ERR_MEMORY;
      }
      if (hipMalloc(&s->spare, 8 * N) != hipSuccess || hipMalloc(&s->hii, 8 * N) != hipSuccess)
        return SBD_ERR_MEMORY;
    }
    // the diagonal and the start vector as sbd_diag makes them
    // (BasisInitVector, davidson.h:144-184), on the admission stream
    s->ctx->stream = st_admit;
    rc = launch_hdiag(s->ctx, s->hii, st_admit);
    if (!rc) rc = start_vector<double>(s->ctx, &q.opt, nullptr, s->C[0]);
    if (rc) return rc;
    s->job.init(s->C, s->HC, s->spare, s->hii, N, q.opt.max_it, q.opt.eps);
    const sbd_ctx* c = s->ctx;
    const double na = (double)c->na, nbv = (double)c->nb;
    s->cost = (double)N * (1.0 + (c->ta.n_s + c->ta.n_d) / na + (c->tb.n_s + c->tb.n_d) / nbv +
                           (c->ta.n_s / na) * (c->tb.n_s / nbv));
    return SBD_OK;
  }

  void admit_loop() {
    (void)hipSetDevice(device);
    for (;;) {
      Request q;
      {
        std::unique_lock<std::mutex> g(mu);
        size_t k = 0;
        cv_admit.wait(g, [&] {
          if (stopping && pending.empty()) return true;
          if (pending.empty()) return false;
          k = pick_pending();
          const uint64_t need = pending[k].bytes;
          const int active = running + preparing + queued;
          const bool slot = active < max_active;
          const bool idle = active == 0 && reaping == 0;
          bool mem = used + need <= budget || idle;
          // a batched solve's vectors must also find a block of the arena
          if (mem && !idle && arena.base && pending[k].batched())
            mem = arena.largest() >=
                  solve_block_bytes(pending[k].a.size() * pending[k].b.size(), pending[k].opt.max_nb);
          return slot && mem;
        });
        if (pending.empty()) return;  // stopping
        // admit exactly the request the predicate checked
        q = std::move(pending[k]);
        pending.erase(pending.begin() + (std::ptrdiff_t)k);
        used += q.bytes;
        preparing++;
        state[q.id] = 1;
      }
      Solve* s = new Solve();
      s->req = std::move(q);
      const int rc = prepare(s);
      std::unique_lock<std::mutex> g(mu);
      preparing--;
      if (!rc) {
        // replace the estimate by the actual footprint
        const uint64_t actual = footprint_built(s->ctx, s->req);
        used = used - s->req.bytes + actual;
        s->req.bytes = actual;
        cv_admit.notify_all();
      }
      if (rc) {
        s->status = rc;
        last_error = s->ctx ? s->ctx->last_error : "request preparation failed";
        running++;  // finish() counts it out
        g.unlock();
        finish(s, sbd_result{});
        continue;
      }
      queued++;
      if (!s->req.batched()) {
        solo_ready.push_back(s);
        cv_solo.notify_one();
        continue;
      }
      // A lane's round lasts about as long as its heaviest solve's step, so a
      // solve goes to the lane where it raises that the least (ties: the
      // lane with the least work): light solves stay away from heavy ones.
      int best = 0;
      double best_round = 0.0;
      for (int l = 0; l < (int)lanes.size(); l++) {
        const double r = std::max(lanes[l].max_cost(), s->cost);
        if (l == 0 || r < best_round || (r == best_round && lanes[l].load < lanes[best].load)) {
          best = l;
          best_round = r;
        }
      }
      s->lane = best;
      lanes[best].load += s->cost;
      lanes[best].costs.push_back(s->cost);
      lanes[best].ready.push_back(s);
      cv_engine.notify_all();
    }
  }

  // ---- solo workers: whole solves on their own stream
  void solo_loop(int wi) {
    (void)hipSetDevice(device);
    Solo& w = solos[wi];
    for (;;) {
      Solve* s = nullptr;
      {
        std::unique_lock<std::mutex> g(mu);
        cv_solo.wait(g, [&] {
          return !solo_ready.empty() || (stopping && pending.empty() && preparing == 0);
        });
        if (solo_ready.empty()) return;
        s = solo_ready.front();
        solo_ready.pop_front();
        s->t_admit = now();
        queued--;
        running++;
        peak_active = std::max<uint64_t>(peak_active, (uint64_t)running);
      }
      s->ctx->stream = w.st;
      sbd_result r{};
      sbd_report rep{};
      const Request& q = s->req;
      int rc;
      if (q.nroots > 1) {
        double res[SBD_MAX_ROOTS];
        rc = sbd_diag_roots(s->ctx, &q.opt, q.nroots, r.energies, res, nullptr, nullptr, 0, &rep);
        r.energy = r.energies[0];
      } else {
        rc = sbd_diag(s->ctx, &q.opt, &r.energy, nullptr, nullptr, 0, &rep);
        r.energies[0] = r.energy;
      }
      r.nroots = q.nroots;
      r.residual = rep.residual;
      r.converged = rep.converged;
      r.steps = rep.steps;
      r.steps_fp32 = rep.steps_fp32;
      if (rc && rc != SBD_ERR_NOT_CONVERGED) s->status = rc;
      finish(s, r);
    }
  }

  // ---- engine: one round advances every active batched solve through its passes
  void engine_loop(int li) {
    (void)hipSetDevice(device);
    Lane& ln = lanes[li];
    std::vector<Solve*> active;
    for (;;) {
      {
        std::unique_lock<std::mutex> g(mu);
        cv_engine.wait(g, [&] {
          return !ln.ready.empty() || !active.empty() ||
                 (stopping && pending.empty() && preparing == 0);
        });
        while (!ln.ready.empty()) {
          Solve* s = ln.ready.front();
          ln.ready.pop_front();
          s->t_admit = now();
          s->ctx->stream = ln.st;
          queued--;
          running++;
          active.push_back(s);
        }
        if (active.empty() && stopping && pending.empty() && preparing == 0) return;
        peak_active = std::max<uint64_t>(peak_active, (uint64_t)running);
        rounds++;
      }
      round(active, ln);
      // completed solves leave the batch
      std::vector<Solve*> keep;
      for (Solve* s : active) {
        if (s->job.next() == DavidsonJob<double>::kDone || s->status) complete(s, ln.st);
        else keep.push_back(s);
      }
      active.swap(keep);
    }
  }

  void fail(Solve* s, int rc) {
    s->status = rc;
    s->job.stop();
  }

  void round(std::vector<Solve*>& active, Lane& ln) {
    using DJ = DavidsonJob<double>;
    hipStream_t st = ln.st;
    BatchScratch& sc = ln.sc;
    // sigma for every solve that needs one, in one launch
    {
      std::vector<SigmaItem> items;
      std::vector<Solve*> who;
      for (Solve* s : active)
        if (s->job.next() == DJ::kSigma) {
          items.push_back(SigmaItem{s->ctx, s->hii, s->job.sigma_in(), s->job.sigma_out()});
          who.push_back(s);
        }
      if (!items.empty()) {
        std::vector<double> tw(items.size(), 0.0);
        const int rc = launch_sigma_batch(sc, items.data(), (int)items.size(), tw.data(), st);
        for (size_t q = 0; q < who.size(); q++) {
          if (rc) fail(who[q], rc);
          else who[q]->job.sigma_done(tw[q]);
        }
      }
    }
    // Ritz pass
    {
      std::vector<RitzArgs<double>> args;
      std::vector<std::vector<const double*>> cp, hp;
      std::vector<double*> outs;
      std::vector<Solve*> who;
      for (Solve* s : active)
        if (s->job.next() == DJ::kRitz) who.push_back(s);
      cp.resize(who.size());
      hp.resize(who.size());
      for (size_t q = 0; q < who.size(); q++) {
        Solve* s = who[q];
        const int k = s->job.k();
        cp[q].assign(s->job.C.begin(), s->job.C.begin() + k);
        hp[q].assign(s->job.HC.begin(), s->job.HC.begin() + k);
        args.push_back(RitzArgs<double>{cp[q].data(), hp[q].data(), k, s->job.ritz_u(), s->job.E0,
                                        s->hii, s->job.t(), s->ctx->N, s->job.ritz_restart()});
        outs.push_back(s->job.ritz_sums());
      }
      if (!who.empty()) {
        const int rc = vec_ritz_batch<double>(sc, args.data(), (int)args.size(), outs.data(), st);
        for (Solve* s : who) {
          if (rc) fail(s, rc);
          else s->job.ritz_done();
        }
      }
    }
    // Gram-Schmidt
    {
      std::vector<GsArgs<double>> args;
      std::vector<std::vector<const double*>> cp;
      std::vector<Solve*> who;
      for (Solve* s : active)
        if (s->job.next() == DJ::kGs) who.push_back(s);
      cp.resize(who.size());
      for (size_t q = 0; q < who.size(); q++) {
        Solve* s = who[q];
        const int k = s->job.k();
        cp[q].assign(s->job.C.begin(), s->job.C.begin() + k);
        args.push_back(GsArgs<double>{cp[q].data(), k, s->job.gs_coef(), s->job.t(),
                                      s->job.gs_scale(), s->ctx->N});
      }
      if (!who.empty()) {
        std::vector<double> n2(who.size(), 0.0);
        const int rc = vec_gs_batch<double>(sc, args.data(), (int)args.size(), n2.data(), st);
        for (size_t q = 0; q < who.size(); q++) {
          if (rc) fail(who[q], rc);
          else who[q]->job.gs_done(n2[q]);
        }
      }
    }
    // the rare passes (re-orthogonalisation, restart) one solve at a time,
    // until each solve waits for its next sigma or is done
    for (Solve* s : active) {
      for (;;) {
        const auto op = s->job.next();
        if (op == DJ::kSigma || op == DJ::kDone || op == DJ::kRitz || op == DJ::kGs) break;
        int rc = SBD_OK;
        const uint64_t N = s->ctx->N;
        if (op == DJ::kDots) {
          const int k = s->job.k();
          std::vector<const double*> cp(s->job.C.begin(), s->job.C.begin() + k);
          rc = vec_multidot<double>(s->ctx, cp.data(), k, s->job.t(), N, s->job.dots_out(), st);
          if (!rc) s->job.dots_done();
        } else if (op == DJ::kLincomb) {
          const int k = s->job.lincomb_k();
          std::vector<const double*> cp(s->job.C.begin(), s->job.C.begin() + k);
          double n2 = 0.0;
          rc = vec_lincomb<double>(s->ctx, cp.data(), k, s->job.ritz_u(), 1.0,
                                   s->job.lincomb_out(), N, &n2, st);
          if (!rc) s->job.lincomb_done(n2);
        } else if (op == DJ::kScale) {
          rc = vec_scale<double>(s->ctx, s->job.lincomb_out(), s->job.scale_factor(), N, st);
          if (!rc) s->job.scale_done();
        }
        if (rc) {
          fail(s, rc);
          break;
        }
      }
    }
  }

  // final vector and E = <W, H W> with a fresh sigma and its fused inner
  // product, as sbd_diag does
  void complete(Solve* s, hipStream_t st) {
    sbd_result r{};
    if (!s->status) {
      const uint64_t N = s->ctx->N;
      auto& j = s->job;
      std::vector<const double*> cp(j.C.begin(), j.C.begin() + j.last_k);
      double n2 = 0.0;
      double* W = j.spare;
      int rc = vec_lincomb<double>(s->ctx, cp.data(), j.last_k, j.ritz_u(), 1.0, W, N, &n2, st);
      if (!rc) rc = vec_scale<double>(s->ctx, W, 1.0 / std::sqrt(n2), N, st);
      double* HW = j.HC[0];
      if (!rc) rc = launch_sigma<double>(s->ctx, s->hii, W, HW, 0, st, &r.energy);
      if (rc) s->status = rc;
    }
    r.nroots = 1;
    r.energies[0] = r.energy;
    r.converged = s->job.converged ? 1 : 0;
    r.steps = s->job.steps;
    r.residual = s->job.normR;
    // buffers may have been rotated by restarts
    s->C = s->job.C;
    s->spare = s->job.spare;
    finish(s, r);
  }
};

extern "C" {

sbd_server* sbd_server_create(uint64_t mem_budget, int max_active, int* status) {
  if (max_active < 1 || mem_budget == 0) {
    if (status) *status = SBD_ERR_ARG;
    return nullptr;
  }
  sbd_server* sv = new sbd_server();
  sv->budget = mem_budget;
  sv->max_active = max_active;
  sv->t0 = clock_s();
  sv->lanes = std::vector<Lane>(std::min(kServerLanes, max_active));
  sv->solos = std::vector<Solo>(std::min(kSoloWorkers, max_active));
  bool ok = hipGetDevice(&sv->device) == hipSuccess;
  ok = ok && hipStreamCreateWithFlags(&sv->st_admit, hipStreamNonBlocking) == hipSuccess;
  for (Lane& l : sv->lanes)
    ok = ok && hipStreamCreateWithFlags(&l.st, hipStreamNonBlocking) == hipSuccess;
  for (Solo& w : sv->solos)
    ok = ok && hipStreamCreateWithFlags(&w.st, hipStreamNonBlocking) == hipSuccess;
  // the batched solves' vectors come from one arena of three quarters of the
  // budget (the rest, counted in the same budget, is the contexts' tables;
  // less when the device has less free memory; without an arena the vectors
  // are allocated per solve)
  if (ok) sv->arena.init(mem_budget / 4 * 3);
  if (!ok) {
    for (Lane& l : sv->lanes)
      if (l.st) (void)hipStreamDestroy(l.st);
    for (Solo& w : sv->solos)
      if (w.st) (void)hipStreamDestroy(w.st);
    if (sv->st_admit) (void)hipStreamDestroy(sv->st_admit);
    if (status) *status = SBD_ERR_DEVICE;
    delete sv;
    return nullptr;
  }
  sv->admitter = std::thread([sv] { sv->admit_loop(); });
  sv->reaper = std::thread([sv] { sv->reap_loop(); });
  for (int l = 0; l < (int)sv->lanes.size(); l++)
    sv->lanes[l].th = std::thread([sv, l] { sv->engine_loop(l); });
  for (int w = 0; w < (int)sv->solos.size(); w++)
    sv->solos[w].th = std::thread([sv, w] { sv->solo_loop(w); });
  if (status) *status = SBD_OK;
  return sv;
}

int sbd_server_submit(sbd_server* sv, const sbd_request* rq, uint64_t* id) {
  if (!sv || !rq || !rq->h1 || !rq->eri_packed || !rq->a_strings || !rq->b_strings ||
      rq->norb < 1 || rq->norb > 64 || rq->na == 0 || rq->nb == 0)
    return SBD_ERR_ARG;
  const int nroots = rq->nroots < 1 ? 1 : rq->nroots;
  if (rq->opt.method != 0 || (rq->opt.precision != 0 && rq->opt.precision != 1) ||
      rq->opt.max_nb < 1 || rq->opt.max_nb > kMaxVecs || rq->opt.max_it < 1 ||
      (rq->opt.init != 0 && rq->opt.init != 1) || nroots > SBD_MAX_ROOTS ||
      nroots > kMaxVecs / 3 || (uint64_t)nroots > (uint64_t)rq->na * rq->nb)
    return SBD_ERR_ARG;
  Request q;
  q.norb = rq->norb;
  q.I0 = rq->I0;
  const size_t L = (size_t)rq->norb, B = L * (L + 1) / 2;
  q.h1.assign(rq->h1, rq->h1 + L * L);
  q.eri.assign(rq->eri_packed, rq->eri_packed + B * (B + 1) / 2);
  q.a.assign(rq->a_strings, rq->a_strings + rq->na);
  q.b.assign(rq->b_strings, rq->b_strings + rq->nb);
  q.opt = rq->opt;
  q.nroots = nroots;
  q.bytes = footprint(q);
  std::lock_guard<std::mutex> g(sv->mu);
  if (sv->stopping) return SBD_ERR_STATE;
  q.id = sv->next_id++;
  q.t_submit = sv->now();
  if (id) *id = q.id;
  sv->state[q.id] = 0;
  sv->pending.push_back(std::move(q));
  sv->cv_admit.notify_all();
  return SBD_OK;
}

int sbd_server_wait(sbd_server* sv, uint64_t id, sbd_result* out) {
  if (!sv) return SBD_ERR_ARG;
  std::unique_lock<std::mutex> g(sv->mu);
  if (!sv->state.count(id)) return SBD_ERR_ARG;
  sv->cv_done.wait(g, [&] { return sv->state[id] == 2; });
  if (out) *out = sv->results[id];
  return SBD_OK;
}

int sbd_server_poll(sbd_server* sv, uint64_t id, sbd_result* out, int* done) {
  if (!sv || !done) return SBD_ERR_ARG;
  std::lock_guard<std::mutex> g(sv->mu);
  if (!sv->state.count(id)) return SBD_ERR_ARG;
  *done = sv->state[id] == 2;
  if (*done && out) *out = sv->results[id];
  return SBD_OK;
}

double sbd_server_now(const sbd_server* sv) { return sv ? sv->now() : 0.0; }

int sbd_server_set_policy(sbd_server* sv, int policy, double age_limit_s) {
  if (!sv || (policy != 0 && policy != 1) || !(age_limit_s > 0.0)) return SBD_ERR_ARG;
  std::lock_guard<std::mutex> g(sv->mu);
  sv->policy = policy;
  sv->age_limit = age_limit_s;
  sv->cv_admit.notify_all();
  return SBD_OK;
}

int sbd_server_stats(const sbd_server* sv, uint64_t* out) {
  if (!sv || !out) return SBD_ERR_ARG;
  std::lock_guard<std::mutex> g(sv->mu);
  out[0] = sv->rounds;
  out[1] = sv->peak_active;
  return SBD_OK;
}

void sbd_server_destroy(sbd_server* sv) {
  if (!sv) return;
  {
    std::lock_guard<std::mutex> g(sv->mu);
    sv->stopping = true;
    sv->cv_admit.notify_all();
    sv->cv_engine.notify_all();
    sv->cv_solo.notify_all();
  }
  if (sv->admitter.joinable()) sv->admitter.join();
  {
    std::lock_guard<std::mutex> g(sv->mu);
    sv->cv_engine.notify_all();
    sv->cv_solo.notify_all();
  }
  for (Lane& l : sv->lanes)
    if (l.th.joinable()) l.th.join();
  for (Solo& w : sv->solos)
    if (w.th.joinable()) w.th.join();
  {
    std::lock_guard<std::mutex> g(sv->mu);
    sv->cv_reap.notify_all();
  }
  if (sv->reaper.joinable()) sv->reaper.join();
  for (Lane& l : sv->lanes) {
    batch_free(l.sc);
    if (l.st) (void)hipStreamDestroy(l.st);
  }
  for (Solo& w : sv->solos)
    if (w.st) (void)hipStreamDestroy(w.st);
  if (sv->st_admit) (void)hipStreamDestroy(sv->st_admit);
  sv->arena.destroy();
  delete sv;
}

}  // extern "C"
