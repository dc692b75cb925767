// b2_serve.cpp -- the serving pattern: a continuous-batching request loop over
// a pool of worlds that stay resident in device memory.
//
// A request asks for `substeps` steps of one world at its own dt. Every tick:
//   1. requests whose arrival time has passed join the arrival queue;
//   2. queued requests are admitted in arrival order while the device working
//      memory of the admitted worlds stays within the budget (a request whose
//      world already has one in flight waits behind it; the first request
//      that does not fit stops admission, so large requests are not starved);
//   3. exactly the worlds with a request in flight are marked active, each
//      with its request's dt and its own dtRatio (inv_dt0 * dt, as
//      b2World::Step computes it, src/dynamics/b2_world.cpp:1130);
//   4. the pool takes one step: every in-flight world advances one substep in
//      the same set of launches, whatever tick its request was admitted at;
//   5. requests whose last substep just ran complete.
// Requests arriving while a tick runs are admitted at the next tick, so work
// from different arrivals overlaps on the device. Every stage of a step works
// per world (inactive worlds are left untouched; the colouring rank is
// world-local), so a world's result does not depend on which other worlds
// share its ticks.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <deque>
#include <thread>
#include <vector>

#include "b2g.h"

namespace {

using Clock = std::chrono::steady_clock;

double seconds_since(Clock::time_point t0) {
  return std::chrono::duration<double>(Clock::now() - t0).count();
}

struct Flight {
  int32_t request;
  int32_t left;
};

}  // namespace

extern "C" {

int b2s_run(b2g_batch* pool, int32_t world_count, const int32_t* world_bodies, const int64_t* world_bytes,
            const float* gravity, float* inv_dt0, int32_t request_count, const b2s_request* req,
            const b2s_config* cfg, b2s_result* res, b2s_stats* stats) {
  if (!pool || world_count < 1 || !world_bodies || !world_bytes || !gravity || !inv_dt0 || request_count < 0 ||
      (request_count && (!req || !res)) || !cfg || !stats)
    return B2G_ERR_INVALID;
  for (int i = 0; i < request_count; ++i) {
    if (req[i].world < 0 || req[i].world >= world_count || req[i].substeps < 1 || !(req[i].dt > 0.0f))
      return B2G_ERR_INVALID;
    if (i && req[i].arrival < req[i - 1].arrival) return B2G_ERR_INVALID;  // sorted by arrival
    if (world_bytes[req[i].world] > cfg->memory_budget) return B2G_ERR_CAPACITY;  // could never be admitted
  }
  const unsigned step_flags = cfg->step_flags & (B2G_STEP_GRAPH);
  std::vector<std::deque<int32_t>> world_queue(world_count);  // arrived, waiting for their world
  std::deque<int32_t> arrived;                                // arrival order
  std::vector<char> busy(world_count, 0);
  std::vector<float> dt_of(world_count, 1.0f / 60.0f);
  std::vector<Flight> flights;
  std::vector<b2g_world_params> params(world_count);
  for (int w = 0; w < world_count; ++w)
    params[w] = b2g_world_params{{gravity[2 * w], gravity[2 * w + 1]}, 1.0f / 60.0f, 0.0f, 1, 0};
  int64_t bytes_in_flight = 0;
  int32_t next = 0, done = 0;
  int64_t tick = 0;
  *stats = b2s_stats{};
  const Clock::time_point t0 = Clock::now();
  const bool realtime = cfg->realtime != 0;
  double in_flight_sum = 0.0;
  while (done < request_count) {
    // 1. arrivals (real time: seconds since start; otherwise the tick number)
    const double now = realtime ? seconds_since(t0) : (double)tick;
    while (next < request_count && req[next].arrival <= now) {
      res[next] = b2s_result{req[next].arrival, -1.0, -1.0, 0, 0};
      arrived.push_back(next++);
    }
    // 2. admission in arrival order, bounded by the working-memory budget
    for (auto it = arrived.begin(); it != arrived.end();) {
      const int32_t r = *it;
      const int32_t w = req[r].world;
      if (busy[w]) {
        ++it;
        continue;
      }
      if (bytes_in_flight + world_bytes[w] > cfg->memory_budget) break;
      busy[w] = 1;
      bytes_in_flight += world_bytes[w];
      flights.push_back(Flight{r, req[r].substeps});
      res[r].admitted = realtime ? seconds_since(t0) : (double)tick;
      it = arrived.erase(it);
    }
    if (flights.empty()) {
      if (next >= request_count) break;  // nothing queued can ever run (cannot happen after the checks)
      if (realtime) {
        const double wait = req[next].arrival - seconds_since(t0);
        if (wait > 2e-4) std::this_thread::sleep_for(std::chrono::microseconds((int64_t)(wait * 0.5e6)));
      } else {
        tick = (int64_t)std::ceil(req[next].arrival);
      }
      continue;
    }
    // 3. activate exactly the in-flight worlds
    for (int w = 0; w < world_count; ++w) params[w].active = 0;
    for (const Flight& f : flights) {
      const int32_t w = req[f.request].world;
      const float dt = req[f.request].dt;
      params[w].active = 1;
      params[w].dt = dt;
      params[w].dt_ratio = inv_dt0[w] * dt;
      dt_of[w] = dt;
    }
    int rc = b2g_set_world_params(pool, 0, world_count, params.data());
    if (rc) return rc;
    // 4. one step of every in-flight world (an eager step also times itself
    // on the device)
    const bool timed = (step_flags & B2G_STEP_GRAPH) == 0;
    const Clock::time_point s0 = Clock::now();
    rc = b2g_batch_step(pool, cfg->velocity_iterations, cfg->position_iterations,
                        step_flags | (timed ? B2G_STEP_PROFILE : 0u));
    if (rc) return rc;
    stats->busy_s += seconds_since(s0);
    b2g_profile prof{};
    if (timed && b2g_get_profile(pool, &prof) == B2G_OK) stats->device_s += 1e-3 * (double)prof.step;
    ++tick;
    ++stats->ticks;
    in_flight_sum += (double)flights.size();
    stats->max_in_flight = std::max<int32_t>(stats->max_in_flight, (int32_t)flights.size());
    // 5. retire finished requests
    const double t_done = realtime ? seconds_since(t0) : (double)tick;
    for (size_t k = 0; k < flights.size();) {
      Flight& f = flights[k];
      const int32_t w = req[f.request].world;
      inv_dt0[w] = 1.0f / dt_of[w];
      stats->body_steps += world_bodies[w];
      stats->world_steps += 1;
      ++res[f.request].ticks;
      if (--f.left == 0) {
        res[f.request].completed = t_done;
        busy[w] = 0;
        bytes_in_flight -= world_bytes[w];
        ++done;
        f = flights.back();
        flights.pop_back();
      } else {
        ++k;
      }
    }
  }
  stats->wall_s = seconds_since(t0);
  stats->mean_in_flight = stats->ticks ? in_flight_sum / (double)stats->ticks : 0.0;
  stats->completed = done;
  return B2G_OK;
}

}  // extern "C"
