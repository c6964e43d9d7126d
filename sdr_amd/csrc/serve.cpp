// serve.cpp — serving loop for many irrigation scenarios on one GPU (sdr_serve).
//
// The reference runs one scene per process (main.rs:215-386).  Here requests arrive over
// time and share the device:
//   * admission: requests are taken first come, first served; the head of the queue is
//     admitted when it has arrived, a slot is free and its estimated footprint
//     (sdr_solver_footprint) fits in what the budget has left.  A request larger than
//     the whole budget is rejected (SDR_E_OOM) instead of blocking the queue;
//   * execution: every admitted request gets its own host thread, solver and HIP
//     stream (sdr_solver_new), so the kernels of all requests in flight interleave on
//     the device; the thread sets the scenario up (SDF upload, cavity mask, seeded
//     fill), runs its substeps and reads back its particles;
//   * release: when a request finishes, its slot and footprint return to the pool and
//     the queue is re-examined at once;
//   * resident pools (cfg->resident_pools): each slot keeps a memory pool (sdr_pool_*)
//     across the requests it serves, so after the first requests a solver's buffers come
//     from blocks the slot already holds instead of hipMalloc / hipFree.
// The loop only uses the public C ABI, so every scenario runs exactly the code a
// standalone caller would run: its result is the same alone and in the mix.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <mutex>
#include <numeric>
#include <string>
#include <thread>
#include <vector>

#include "../../include/sdr_amd.h"

namespace {

using Clock = std::chrono::steady_clock;

uint64_t fnv1a(uint64_t h, const void* p, size_t n) {
  const unsigned char* b = (const unsigned char*)p;
  for (size_t i = 0; i < n; ++i) {
    h ^= b[i];
    h *= 1099511628211ull;
  }
  return h;
}

// Particles a scenario holds at most: the seeded fill (every cavity cell of the fill
// band, bounded by the whole grid) plus what the needle emits over its substeps.
uint64_t particle_bound(const sdr_scenario& s, uint64_t nx, uint64_t ny, uint64_t nz) {
  const double cell_vol = s.sim_dx * s.sim_dx * s.sim_dx;
  const double ppc = (double)(s.fluid.particles_per_cell ? s.fluid.particles_per_cell : 1);
  const double emit = s.needle.flow_rate * s.dt * (double)s.substeps / (cell_vol / ppc);
  return (uint64_t)nx * ny * nz * s.fill_per_cell + (uint64_t)emit + 1024;
}

int run_one(const sdr_scenario& s, sdr_scenario_result* r, sdr_pool* pool) {
  // the slot's pool, if any, serves every allocation of the request on this thread
  struct Bind {
    sdr_pool* p;
    uint64_t misses0 = 0;
    sdr_scenario_result* r;
    Bind(sdr_pool* p_, sdr_scenario_result* r_) : p(p_), r(r_) {
      if (!p) return;
      sdr_pool_stats ps{};
      sdr_pool_get_stats(p, &ps);
      misses0 = ps.misses;
      sdr_pool_bind(p);
    }
    ~Bind() {
      if (!p) return;
      sdr_pool_bind(nullptr);
      sdr_pool_stats ps{};
      sdr_pool_get_stats(p, &ps);
      r->device_allocs = ps.misses - misses0;
    }
  } bind{pool, r};
  sdr_solver* h = nullptr;
  int rc = sdr_solver_new(&s.solid, s.sim_dx, &s.fluid, &s.needle, &s.opts, &h);
  if (rc != SDR_OK) return rc;
  struct Free {
    sdr_solver* h;
    ~Free() { sdr_solver_free(h); }
  } guard{h};
  if ((rc = sdr_solver_set_outlet(h, s.has_outlet ? &s.outlet : nullptr)) != SDR_OK) return rc;
  if ((rc = sdr_solver_set_solve_params(h, &s.solve)) != SDR_OK) return rc;
  const sdr_vec3 v0{0.0, 0.0, 0.0};
  uint64_t n = 0;
  if ((rc = sdr_solver_seed_fill(h, s.fill_per_cell, s.fill_y_lo, s.fill_y_hi, s.fill_seed, &v0, &n)) != SDR_OK)
    return rc;
  sdr_state st{};
  for (uint32_t k = 0; k < s.substeps; ++k) {
    if ((rc = sdr_solver_particle_count(h, &n)) != SDR_OK) return rc;
    r->particle_substeps += n;
    if ((rc = sdr_solver_substep(h, s.dt)) != SDR_OK) return rc;
    if ((rc = sdr_solver_get_state(h, &st)) != SDR_OK) return rc;
    r->pcg_iters += st.last_pcg_iters;
  }
  if ((rc = sdr_solver_particle_count(h, &n)) != SDR_OK) return rc;
  r->particles = n;
  std::vector<double> pos(3 * n), vel(3 * n);
  if ((rc = sdr_solver_get_particles(h, pos.data(), vel.data(), n)) != SDR_OK) return rc;
  uint64_t c = 1469598103934665603ull;
  c = fnv1a(c, pos.data(), pos.size() * sizeof(double));
  c = fnv1a(c, vel.data(), vel.size() * sizeof(double));
  r->checksum = c;
  return SDR_OK;
}

}  // namespace

extern "C" int sdr_serve(const sdr_scenario* sc, uint32_t n, const sdr_serve_config* cfg,
                         sdr_scenario_result* out) {
  if (!sc || !out || !cfg || cfg->max_slots == 0 || !(cfg->hbm_budget_bytes > 0)) return SDR_E_INVALID;
  std::memset(out, 0, sizeof(sdr_scenario_result) * n);
  // footprints and arrival order
  std::vector<uint32_t> order(n);
  std::iota(order.begin(), order.end(), 0u);
  std::stable_sort(order.begin(), order.end(),
                   [&](uint32_t a, uint32_t b) { return sc[a].arrival_s < sc[b].arrival_s; });
  for (uint32_t i = 0; i < n; ++i) {
    const sdr_scenario& s = sc[i];
    // grid dims as Solver::new derives them (solver.rs:185-190)
    auto dim = [&](uint64_t nodes) {
      const double size = (double)(nodes - 1) * s.solid.dx;
      const double d = std::max(std::ceil(size / s.sim_dx), 1.0);
      return (uint64_t)d;
    };
    const uint64_t nx = dim(s.solid.nx), ny = dim(s.solid.ny), nz = dim(s.solid.nz);
    uint64_t fp = 0;
    int rc = sdr_solver_footprint(nx, ny, nz, particle_bound(s, nx, ny, nz), s.opts.precision, &fp);
    out[i].footprint_bytes = fp;
    out[i].arrival_s = s.arrival_s;
    out[i].status = rc;
  }
  std::mutex mu;
  std::condition_variable cv;
  double used = 0.0;
  uint32_t inflight = 0, finished = 0;
  std::vector<uint32_t> free_slots;
  for (uint32_t k = cfg->max_slots; k > 0; --k) free_slots.push_back(k - 1);
  std::vector<std::thread> workers;
  // one pool per slot, kept across the requests the slot serves
  std::vector<sdr_pool*> pools(cfg->max_slots, nullptr);
  struct FreePools {
    std::vector<sdr_pool*>& v;
    ~FreePools() {
      for (sdr_pool* p : v)
        if (p) sdr_pool_free(p);
    }
  } free_pools{pools};
  if (cfg->resident_pools) {
    int32_t dev = n > 0 ? sc[0].opts.device : 0;
    for (auto& p : pools)
      if (sdr_pool_new(dev, &p) != SDR_OK) p = nullptr;
  }
  const Clock::time_point t0 = Clock::now();
  auto now_s = [&] { return std::chrono::duration<double>(Clock::now() - t0).count(); };
  size_t head = 0;
  std::unique_lock<std::mutex> lk(mu);
  while (finished < n) {
    // admit from the head of the queue
    while (head < order.size()) {
      const uint32_t i = order[head];
      if (out[i].status != SDR_OK || (double)out[i].footprint_bytes > cfg->hbm_budget_bytes) {
        if (out[i].status == SDR_OK) out[i].status = SDR_E_OOM;  // can never fit
        out[i].admit_s = out[i].finish_s = std::max(now_s(), sc[i].arrival_s);
        ++finished;
        ++head;
        continue;
      }
      if (sc[i].arrival_s > now_s() || free_slots.empty() ||
          used + (double)out[i].footprint_bytes > cfg->hbm_budget_bytes)
        break;
      const uint32_t slot = free_slots.back();
      free_slots.pop_back();
      used += (double)out[i].footprint_bytes;
      ++inflight;
      out[i].slot = slot;
      out[i].admit_s = now_s();
      ++head;
      workers.emplace_back([&, i, slot] {
        sdr_scenario_result r = out[i];
        const int rc = run_one(sc[i], &r, pools[slot]);
        std::lock_guard<std::mutex> g(mu);
        r.status = rc;
        r.finish_s = now_s();
        out[i] = r;
        used -= (double)r.footprint_bytes;
        --inflight;
        ++finished;
        free_slots.push_back(slot);
        cv.notify_all();
      });
    }
    if (finished >= n) break;
    // sleep until the next arrival or a completion
    if (head < order.size() && sc[order[head]].arrival_s > now_s()) {
      const double wait = sc[order[head]].arrival_s - now_s();
      cv.wait_for(lk, std::chrono::duration<double>(wait));
    } else {
      cv.wait(lk);
    }
  }
  lk.unlock();
  for (auto& w : workers) w.join();
  return SDR_OK;
}
