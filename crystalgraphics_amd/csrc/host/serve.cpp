// serve.cpp — the serving form over the batch engine of the HIP backend
// (cg_batch_*): one thread owns the backend env, keeps the scenes resident,
// admits requests in arrival order as slots and output memory allow (a
// later request that fits may pass an earlier one that does not), launches
// rounds while anything is in flight, and hands finished results back.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <future>
#include <map>
#include <mutex>
#include <thread>
#include <utility>

#include "../../../include/CrystalGraphics/serve.h"
#include "../../../include/cg_device.h"
#include "scene_impl.h"

namespace crystal::graphics {

namespace {

using Clock = std::chrono::steady_clock;
double seconds_since(Clock::time_point t0, Clock::time_point t) {
  return std::chrono::duration<double>(t - t0).count();
}

// First-fit allocator over one device allocation (request outputs), so the
// serving loop never calls the HIP allocator.
class Arena {
 public:
  void reset(char* base, size_t bytes) {
    base_ = base;
    free_.clear();
    free_.push_back({0, bytes});
  }
  char* alloc(size_t bytes) {
    bytes = (bytes + 255) & ~size_t{255};
    for (size_t i = 0; i < free_.size(); ++i) {
      if (free_[i].second < bytes) continue;
      const size_t off = free_[i].first;
      free_[i].first += bytes;
      free_[i].second -= bytes;
      if (free_[i].second == 0) free_.erase(free_.begin() + static_cast<long>(i));
      sizes_[off] = bytes;
      return base_ + off;
    }
    return nullptr;
  }
  void release(char* p) {
    const size_t off = static_cast<size_t>(p - base_);
    auto s = sizes_.find(off);
    if (s == sizes_.end()) return;
    const size_t bytes = s->second;
    sizes_.erase(s);
    auto it = std::lower_bound(free_.begin(), free_.end(), std::make_pair(off, size_t{0}));
    it = free_.insert(it, {off, bytes});
    size_t i = static_cast<size_t>(it - free_.begin());
    if (i + 1 < free_.size() && free_[i].first + free_[i].second == free_[i + 1].first) {
      free_[i].second += free_[i + 1].second;
      free_.erase(free_.begin() + static_cast<long>(i) + 1);
    }
    if (i > 0 && free_[i - 1].first + free_[i - 1].second == free_[i].first) {
      free_[i - 1].second += free_[i].second;
      free_.erase(free_.begin() + static_cast<long>(i));
    }
  }

 private:
  char* base_ = nullptr;
  std::vector<std::pair<size_t, size_t>> free_;  // (offset, bytes), sorted
  std::map<size_t, size_t> sizes_;
};

cg_camera to_cg_camera(const Camera& c) {
  cg_camera o{};
  o.position[0] = c.position.x;
  o.position[1] = c.position.y;
  o.position[2] = c.position.z;
  o.direction[0] = c.direction.x;
  o.direction[1] = c.direction.y;
  o.direction[2] = c.direction.z;
  o.viewport[0] = c.viewport.x;
  o.viewport[1] = c.viewport.y;
  return o;
}
cg_conf to_cg_conf(const PathTraceConf& c) {
  return cg_conf{c.window_width, c.window_height, c.render_width, c.render_height, c.sample_count, c.trace_depth,
                 c.max_transport_sample_count, c.max_diffuse_sample_count, c.max_rough_sample_count,
                 c.max_emission_sample_count};
}

struct Request {
  uint64_t id = 0;
  RenderRequest req;
  Clock::time_point arrival, admitted;
  char* out = nullptr;  // device history (W*H cg_history_pixel) in the arena
  bool alone = false;
  std::promise<expected<RenderResult, Error>>* done_alone = nullptr;
};

}  // namespace

struct Server::Impl {
  cg_env* env = nullptr;
  cg_batch* batch = nullptr;
  std::vector<cg_scene*> scenes;
  void* arena_mem = nullptr;
  Arena arena;
  ServeConf conf;
  uint64_t slot_capacity = 0, out_bytes = 0;

  std::mutex mu;
  std::condition_variable cv_in, cv_out;
  std::deque<Request> incoming;
  std::deque<RenderResult> results;
  bool stop = false;
  std::string error;
  ServeStats stats;
  uint64_t waiting = 0;  // submitted, not yet admitted (guarded by mu)
  std::atomic<uint64_t> next_id{1};
  std::thread worker;

  // server-thread state
  std::deque<Request> pending;
  std::map<uint32_t, Request> active;  // batch job id -> request (wavefront)
  struct Launch {
    cg_vbatch* vb = nullptr;
    std::vector<Request> reqs;
  };
  std::deque<Launch> launches;  // in flight (kLaunches)
  Clock::time_point busy_since;

  ~Impl() {
    {
      std::lock_guard<std::mutex> g(mu);
      stop = true;
    }
    cv_in.notify_all();
    if (worker.joinable()) worker.join();
    for (Launch& L : launches) cg_view_batch_release(L.vb);  // waits: nothing in flight reads what is freed below
    launches.clear();
    if (batch) cg_batch_destroy(batch);
    if (env) {
      if (arena_mem) cg_device_free(env, arena_mem);
      for (cg_scene* s : scenes) cg_scene_release(env, s);
      cg_env_destroy(env);
    }
  }

  bool fail_locked(const std::string& e) {
    error = e;
    stop = true;
    cv_out.notify_all();
    return false;
  }

  // The worker ran into an error: every RenderAlone still waiting (in the
  // worker's queue or not yet taken from the incoming one) gets it.
  void fail_waiting_alone() {
    std::lock_guard<std::mutex> g(mu);
    for (std::deque<Request>* q : {&pending, &incoming})
      for (Request& r : *q)
        if (r.alone && r.done_alone) {
          r.done_alone->set_value(unexpected<Error>(Error{error.empty() ? "Server: stopped" : error}));
          r.done_alone = nullptr;
        }
  }

  void run() {
    loop();
    fail_waiting_alone();
  }

  RenderResult finish(const Request& r, const std::vector<cg_history_pixel>& h) {
    RenderResult res;
    res.id = r.id;
    res.width = r.req.conf.render_width;
    res.height = r.req.conf.render_height;
    res.xyz.resize(h.size() * 3);
    res.seeds.resize(h.size());
    for (size_t i = 0; i < h.size(); ++i) {
      std::memcpy(&res.xyz[3 * i], h[i].xyz, 3 * sizeof(float));
      res.seeds[i] = h[i].rand_seed;
    }
    const auto now = Clock::now();
    res.queued_s = seconds_since(r.arrival, r.admitted);
    res.latency_s = seconds_since(r.arrival, now);
    return res;
  }

  cg_job job_of(const Request& r) {
    cg_job j{};
    j.scene = scenes[r.req.scene];
    j.camera = to_cg_camera(r.req.camera);
    j.conf = to_cg_conf(r.req.conf);
    j.iter_count = 0;
    j.rng_mode = conf.rng_mode;
    j.history = reinterpret_cast<cg_history_pixel*>(r.out);
    return j;
  }

  std::string last() { return cg_env_last_error(env); }

  // One request alone with the plain kernel (nothing else in flight).
  void run_alone(Request& r) {
    const size_t px = static_cast<size_t>(r.req.conf.render_width) * r.req.conf.render_height;
    r.admitted = Clock::now();
    r.out = arena.alloc(px * sizeof(cg_history_pixel));
    if (!r.out) {
      r.done_alone->set_value(unexpected<Error>(Error{"request output does not fit the output arena"}));
      return;
    }
    cg_job j = job_of(r);
    std::vector<cg_history_pixel> h(px);
    cg_stats before{};  // the env's query counters, cumulative over its kernels
    const bool counted = cg_stats_read(env, &before) == CG_OK;
    if (cg_view_job(env, &j) != CG_OK || cg_memcpy_d2h(env, h.data(), r.out, px * sizeof(cg_history_pixel)) != CG_OK) {
      arena.release(r.out);
      r.done_alone->set_value(unexpected<Error>(Error{last()}));
      return;
    }
    arena.release(r.out);
    cg_stats after{};
    if (counted && cg_stats_read(env, &after) == CG_OK) {  // the alone render's queries count too
      std::lock_guard<std::mutex> g(mu);
      stats.rays += (after.rays + after.shadow_rays) - (before.rays + before.shadow_rays);
    }
    r.done_alone->set_value(finish(r, h));
  }

  // kLaunches: admission windows become persistent launches of the per-pixel
  // kernel (cg_view_batch_launch), up to max_launches_in_flight at once.
  void loop_launches() {
    for (;;) {
      {
        std::unique_lock<std::mutex> lk(mu);
        if (pending.empty() && launches.empty())
          cv_in.wait(lk, [&] { return stop || !incoming.empty(); });
        else if (pending.empty() || launches.size() >= conf.max_launches_in_flight)
          cv_in.wait_for(lk, std::chrono::microseconds(200), [&] { return stop || !incoming.empty(); });
        if (stop && incoming.empty() && pending.empty() && launches.empty()) return;
        while (!incoming.empty()) {
          pending.push_back(std::move(incoming.front()));
          incoming.pop_front();
        }
        if (stop && !error.empty()) return;
      }
      // finished launches, in any order
      std::vector<RenderResult> out;
      for (auto it = launches.begin(); it != launches.end();) {
        int done = 0;
        if (cg_view_batch_query(it->vb, &done) != CG_OK) {
          std::lock_guard<std::mutex> g(mu);
          fail_locked(last());
          return;
        }
        if (!done) {
          ++it;
          continue;
        }
        cg_view_batch_release(it->vb);
        for (Request& r : it->reqs) {
          const size_t px = static_cast<size_t>(r.req.conf.render_width) * r.req.conf.render_height;
          std::vector<cg_history_pixel> h(px);
          if (cg_memcpy_d2h(env, h.data(), r.out, px * sizeof(cg_history_pixel)) != CG_OK) {
            std::lock_guard<std::mutex> g(mu);
            fail_locked(last());
            return;
          }
          arena.release(r.out);
          out.push_back(finish(r, h));
        }
        it = launches.erase(it);
      }
      if (!out.empty()) {
        cg_stats st{};
        cg_stats_read(env, &st);  // cumulative over every launch so far
        std::lock_guard<std::mutex> g(mu);
        stats.rays = st.rays + st.shadow_rays;
        if (launches.empty()) stats.gpu_busy_s += seconds_since(busy_since, Clock::now());
        for (auto& r : out) {
          ++stats.completed;
          results.push_back(std::move(r));
        }
        cv_out.notify_all();
      }
      // requests for the plain kernel alone, when nothing is in flight
      if (launches.empty()) {
        for (auto it = pending.begin(); it != pending.end();) {
          if (!it->alone) {
            ++it;
            continue;
          }
          run_alone(*it);
          it = pending.erase(it);
        }
      }
      if (pending.empty() || launches.size() >= conf.max_launches_in_flight) continue;
      // the next launch: every waiting request whose output fits, in arrival order
      Launch L;
      std::vector<cg_job> jobs;
      for (auto it = pending.begin(); it != pending.end();) {
        if (it->alone) {
          ++it;
          continue;
        }
        const size_t px = static_cast<size_t>(it->req.conf.render_width) * it->req.conf.render_height;
        char* o = arena.alloc(px * sizeof(cg_history_pixel));
        if (!o) {
          ++it;
          continue;
        }
        it->out = o;
        it->admitted = Clock::now();
        jobs.push_back(job_of(*it));
        L.reqs.push_back(std::move(*it));
        it = pending.erase(it);
      }
      if (jobs.empty()) continue;
      if (cg_view_batch_launch(env, jobs.data(), static_cast<uint32_t>(jobs.size()), &L.vb) != CG_OK) {
        std::lock_guard<std::mutex> g(mu);
        fail_locked(last());
        return;
      }
      if (launches.empty()) busy_since = Clock::now();
      launches.push_back(std::move(L));
      std::lock_guard<std::mutex> g(mu);
      stats.admitted += jobs.size();
      waiting -= jobs.size();
    }
  }

  void loop() {
    if (conf.engine == ServeConf::Engine::kLaunches) {
      loop_launches();
      return;
    }
    std::vector<uint32_t> done(4096);
    for (;;) {
      {
        std::unique_lock<std::mutex> lk(mu);
        if (pending.empty() && active.empty())
          cv_in.wait(lk, [&] { return stop || !incoming.empty(); });
        if (stop && incoming.empty() && pending.empty() && active.empty()) return;
        while (!incoming.empty()) {
          pending.push_back(std::move(incoming.front()));
          incoming.pop_front();
        }
        if (stop && !error.empty()) return;
      }
      // Admission, in arrival order; a request that does not fit now waits.
      for (auto it = pending.begin(); it != pending.end();) {
        if (it->alone) {
          if (active.empty()) {
            run_alone(*it);
            it = pending.erase(it);
            continue;
          }
          ++it;
          continue;
        }
        const size_t px = static_cast<size_t>(it->req.conf.render_width) * it->req.conf.render_height;
        char* out = arena.alloc(px * sizeof(cg_history_pixel));
        if (!out) {
          ++it;
          continue;
        }
        it->out = out;
        cg_job j = job_of(*it);
        uint32_t jid = 0;
        const int rc = cg_batch_submit(batch, &j, &jid);
        if (rc == CG_ERR_FULL) {
          arena.release(out);
          it->out = nullptr;
          ++it;
          continue;
        }
        if (rc != CG_OK) {
          std::lock_guard<std::mutex> g(mu);
          fail_locked(last());
          return;
        }
        it->admitted = Clock::now();
        active.emplace(jid, std::move(*it));
        it = pending.erase(it);
        std::lock_guard<std::mutex> g(mu);
        ++stats.admitted;
        --waiting;
      }
      if (active.empty()) continue;
      const auto t0 = Clock::now();
      uint32_t nd = 0, na = 0;
      cg_stats st{};
      if (cg_batch_launch(batch, 1) != CG_OK ||
          cg_batch_poll(batch, done.data(), static_cast<uint32_t>(done.size()), &nd, &na) != CG_OK ||
          cg_batch_stats(batch, &st, 1) != CG_OK) {
        std::lock_guard<std::mutex> g(mu);
        fail_locked(last());
        return;
      }
      std::vector<RenderResult> out;
      for (uint32_t k = 0; k < nd; ++k) {
        auto a = active.find(done[k]);
        if (a == active.end()) continue;
        Request& r = a->second;
        const size_t px = static_cast<size_t>(r.req.conf.render_width) * r.req.conf.render_height;
        std::vector<cg_history_pixel> h(px);
        if (cg_memcpy_d2h(env, h.data(), r.out, px * sizeof(cg_history_pixel)) != CG_OK) {
          std::lock_guard<std::mutex> g(mu);
          fail_locked(last());
          return;
        }
        arena.release(r.out);
        out.push_back(finish(r, h));
        active.erase(a);
      }
      std::lock_guard<std::mutex> g(mu);
      stats.rays += st.rays + st.shadow_rays;
      stats.gpu_busy_s += seconds_since(t0, Clock::now());
      for (auto& r : out) {
        ++stats.completed;
        results.push_back(std::move(r));
      }
      if (!out.empty()) cv_out.notify_all();
    }
  }
};

Server::Server(std::unique_ptr<Impl> impl) : impl_(std::move(impl)) {}
Server::~Server() = default;

expected<std::unique_ptr<Server>, Error> Server::Create(const ServeConf& conf, const std::vector<const Scene*>& scenes) {
  if (scenes.empty()) return unexpected<Error>(Error{"Server: no scenes"});
  if (conf.max_trace_depth < 1 || conf.max_trace_depth > 64)
    return unexpected<Error>(Error{"Server: max_trace_depth must be in [1, 64]"});
  if (!(conf.output_fraction > 0.0 && conf.output_fraction < 1.0))
    return unexpected<Error>(Error{"Server: output_fraction must be in (0, 1)"});
  if (conf.rng_mode > 1) return unexpected<Error>(Error{"Server: rng_mode must be 0 or 1"});
  if (conf.rng_mode == 1 && conf.engine == ServeConf::Engine::kWavefront)
    return unexpected<Error>(Error{"Server: the wavefront engine renders in reference order only (rng_mode 0)"});
  auto impl = std::make_unique<Impl>();
  impl->conf = conf;
  cg_conf c{8, 8, 8, 8, 1, conf.max_trace_depth, 3, 1, 1, 1};
  if (cg_env_create(&c, 0, &impl->env) != CG_OK) return unexpected<Error>(Error{cg_last_error()});
  uint32_t stack = 0;
  const int threads = static_cast<int>(std::max(1u, std::min(16u, std::thread::hardware_concurrency())));
  for (const Scene* s : scenes) {
    SceneData sd = BuildSceneData(s->impl(), true, threads);
    cg_scene_data v = sd.View();
    cg_scene* ds = nullptr;
    if (cg_scene_upload(impl->env, &v, &ds) != CG_OK) return unexpected<Error>(Error{impl->last()});
    impl->scenes.push_back(ds);
    uint32_t depth = 0;
    cg_scene_footprint(ds, nullptr, &depth);
    stack = std::max(stack, depth);
  }
  const bool wavefront = conf.engine == ServeConf::Engine::kWavefront;
  const size_t out_bytes =
      wavefront ? static_cast<size_t>(static_cast<double>(conf.hbm_budget_bytes) * conf.output_fraction)
                : conf.hbm_budget_bytes;
  if (cg_device_alloc(impl->env, out_bytes, &impl->arena_mem) != CG_OK) return unexpected<Error>(Error{impl->last()});
  impl->arena.reset(static_cast<char*>(impl->arena_mem), out_bytes);
  impl->out_bytes = out_bytes;
  impl->slot_capacity = ~uint64_t{0};
  if (wavefront) {
    const size_t slot_bytes = 64 + 32 + 4 + 8 + 128 * static_cast<size_t>(std::max(1u, conf.max_trace_depth - 1));
    const size_t slots = (conf.hbm_budget_bytes - out_bytes) / slot_bytes;
    if (slots < 64) return unexpected<Error>(Error{"Server: budget too small"});
    impl->slot_capacity = std::min<size_t>(slots, 1u << 28) & ~size_t{63};
    cg_batch_conf bc{static_cast<uint32_t>(std::min<size_t>(slots, 1u << 28)), conf.max_trace_depth,
                     conf.rounds_per_launch, 1u, 1024u, std::max(stack, 16u), conf.block_size, conf.short_stack};
    if (cg_batch_create(impl->env, &bc, &impl->batch) != CG_OK) return unexpected<Error>(Error{impl->last()});
  } else {
    if (conf.max_launches_in_flight == 0) return unexpected<Error>(Error{"Server: max_launches_in_flight must be > 0"});
    cg_env_set_counting(impl->env, 1);
  }
  Impl* p = impl.get();
  impl->worker = std::thread([p] { p->run(); });
  return std::unique_ptr<Server>(new Server(std::move(impl)));
}

expected<std::uint64_t, Error> Server::Submit(const RenderRequest& request) {
  Impl& m = *impl_;
  if (request.scene >= m.scenes.size()) return unexpected<Error>(Error{"Submit: unknown scene"});
  if (request.conf.trace_depth > m.conf.max_trace_depth)
    return unexpected<Error>(Error{"Submit: trace_depth beyond the server's max_trace_depth"});
  if (request.conf.render_width == 0 || request.conf.render_height == 0)
    return unexpected<Error>(Error{"Submit: render size must be positive"});
  // A request must fit the empty server (then it is admitted once everything
  // ahead of it has finished).
  const uint64_t px = static_cast<uint64_t>(request.conf.render_width) * request.conf.render_height;
  const uint64_t slots = static_cast<uint64_t>((request.conf.render_width + 7) / 8) *
                         ((request.conf.render_height + 7) / 8) * 64u;
  if (slots > m.slot_capacity || px * sizeof(cg_history_pixel) > m.out_bytes)
    return unexpected<Error>(Error{"Submit: request larger than the server's budget allows"});
  Request r;
  const uint64_t id = m.next_id++;
  r.id = id;
  r.req = request;
  r.arrival = Clock::now();
  {
    std::lock_guard<std::mutex> g(m.mu);
    if (!m.error.empty()) return unexpected<Error>(Error{m.error});
    ++m.stats.submitted;
    if (m.conf.max_waiting > 0 && m.waiting >= m.conf.max_waiting) {
      ++m.stats.refused;
      return unexpected<Error>(Error{"Submit: queue full (" + std::to_string(m.waiting) + " requests waiting)"});
    }
    m.incoming.push_back(std::move(r));
    ++m.waiting;
  }
  m.cv_in.notify_all();
  return id;
}

expected<RenderResult, Error> Server::Next(double timeout_s) {
  Impl& m = *impl_;
  std::unique_lock<std::mutex> lk(m.mu);
  m.cv_out.wait_for(lk, std::chrono::duration<double>(timeout_s),
                    [&] { return !m.results.empty() || !m.error.empty(); });
  if (!m.results.empty()) {
    RenderResult r = std::move(m.results.front());
    m.results.pop_front();
    return r;
  }
  if (!m.error.empty()) return unexpected<Error>(Error{m.error});
  return RenderResult{};
}

expected<RenderResult, Error> Server::RenderAlone(const RenderRequest& request) {
  Impl& m = *impl_;
  if (request.scene >= m.scenes.size()) return unexpected<Error>(Error{"RenderAlone: unknown scene"});
  std::promise<expected<RenderResult, Error>> done;
  auto fut = done.get_future();
  Request r;
  r.id = m.next_id++;
  r.req = request;
  r.arrival = Clock::now();
  r.alone = true;
  r.done_alone = &done;
  {
    std::lock_guard<std::mutex> g(m.mu);
    if (!m.error.empty()) return unexpected<Error>(Error{m.error});
    m.incoming.push_back(std::move(r));
  }
  m.cv_in.notify_all();
  return fut.get();
}

ServeStats Server::Stats() const {
  std::lock_guard<std::mutex> g(impl_->mu);
  return impl_->stats;
}

}  // namespace crystal::graphics

// ===========================================================================
// C exports (bindings, benchmarks).
using namespace crystal::graphics;

namespace {
thread_local std::string t_serve_err;
int serve_fail(const std::string& e) {
  t_serve_err = e;
  return CG_ERR_INVALID;
}
}  // namespace

extern "C" {

typedef struct cgh_server cgh_server;
typedef struct cgh_scene cgh_scene;
typedef struct cgh_result_info {
  uint64_t id;
  uint32_t width, height;
  double queued_s, latency_s;
} cgh_result_info;
typedef struct cgh_serve_stats {
  uint64_t submitted, admitted, completed, rays;
  double gpu_busy_s;
  uint64_t refused;
} cgh_serve_stats;

const char* cgh_serve_last_error(void) { return t_serve_err.c_str(); }

int cgh_server_create(const cgh_scene* const* scenes, uint32_t n, uint32_t engine, uint64_t budget_bytes,
                      double output_fraction, uint32_t max_trace_depth, uint32_t in_flight, uint32_t rounds_per_launch,
                      uint32_t block_size, uint32_t short_stack, uint32_t max_waiting, uint32_t rng_mode,
                      cgh_server** out) {
  if (!out || (n && !scenes)) return serve_fail("cgh_server_create: null argument");
  *out = nullptr;
  std::vector<const Scene*> sc;
  for (uint32_t i = 0; i < n; ++i) sc.push_back(reinterpret_cast<const Scene*>(scenes[i]));
  ServeConf c;
  c.engine = engine == 1 ? ServeConf::Engine::kWavefront : ServeConf::Engine::kLaunches;
  c.max_launches_in_flight = in_flight;
  c.hbm_budget_bytes = budget_bytes;
  c.output_fraction = output_fraction;
  c.max_trace_depth = max_trace_depth;
  c.rounds_per_launch = rounds_per_launch;
  c.block_size = block_size;
  c.short_stack = short_stack;
  c.max_waiting = max_waiting;
  c.rng_mode = rng_mode;
  auto s = Server::Create(c, sc);
  if (!s) return serve_fail(s.error());
  *out = reinterpret_cast<cgh_server*>(std::move(*s).release());
  return CG_OK;
}

void cgh_server_free(cgh_server* s) { delete reinterpret_cast<Server*>(s); }

static RenderRequest make_request(uint32_t scene, const cg_camera* cam, const cg_conf* conf) {
  RenderRequest r;
  r.scene = scene;
  r.camera.position = {cam->position[0], cam->position[1], cam->position[2]};
  r.camera.direction = {cam->direction[0], cam->direction[1], cam->direction[2]};
  r.camera.viewport = {cam->viewport[0], cam->viewport[1]};
  r.conf = PathTraceConf{conf->window_width, conf->window_height, conf->render_width, conf->render_height,
                         conf->sample_count, conf->trace_depth, conf->max_transport_sample_count,
                         conf->max_diffuse_sample_count, conf->max_rough_sample_count, conf->max_emission_sample_count};
  return r;
}

int cgh_server_submit(cgh_server* s, uint32_t scene, const cg_camera* cam, const cg_conf* conf, uint64_t* id) {
  if (!s || !cam || !conf || !id) return serve_fail("cgh_server_submit: null argument");
  auto r = reinterpret_cast<Server*>(s)->Submit(make_request(scene, cam, conf));
  if (!r) {
    serve_fail(r.error());
    // a refusal by the overload bound is "no room now", not a failure
    return r.error().rfind("Submit: queue full", 0) == 0 ? CG_ERR_FULL : CG_ERR_INVALID;
  }
  *id = *r;
  return CG_OK;
}

static void copy_result(const RenderResult& r, cgh_result_info* info, cg_history_pixel* out) {
  info->id = r.id;
  info->width = r.width;
  info->height = r.height;
  info->queued_s = r.queued_s;
  info->latency_s = r.latency_s;
  if (out)
    for (size_t i = 0; i < r.seeds.size(); ++i) {
      std::memcpy(out[i].xyz, &r.xyz[3 * i], 3 * sizeof(float));
      out[i].rand_seed = r.seeds[i];
    }
}

// The next finished request (info->id == 0: none within the timeout).  `out`
// (nullable) receives W*H history records; `cap` is its length.
int cgh_server_next(cgh_server* s, double timeout_s, cgh_result_info* info, cg_history_pixel* out, size_t cap) {
  if (!s || !info) return serve_fail("cgh_server_next: null argument");
  auto r = reinterpret_cast<Server*>(s)->Next(timeout_s);
  if (!r) return serve_fail(r.error());
  if (out && r->seeds.size() > cap) return serve_fail("cgh_server_next: output buffer too small");
  copy_result(*r, info, out);
  return CG_OK;
}

int cgh_server_render_alone(cgh_server* s, uint32_t scene, const cg_camera* cam, const cg_conf* conf,
                            cgh_result_info* info, cg_history_pixel* out, size_t cap) {
  if (!s || !cam || !conf || !info) return serve_fail("cgh_server_render_alone: null argument");
  auto r = reinterpret_cast<Server*>(s)->RenderAlone(make_request(scene, cam, conf));
  if (!r) return serve_fail(r.error());
  if (out && r->seeds.size() > cap) return serve_fail("cgh_server_render_alone: output buffer too small");
  copy_result(*r, info, out);
  return CG_OK;
}

int cgh_server_stats(cgh_server* s, cgh_serve_stats* out) {
  ServeStats st = reinterpret_cast<Server*>(s)->Stats();
  out->submitted = st.submitted;
  out->admitted = st.admitted;
  out->completed = st.completed;
  out->rays = st.rays;
  out->gpu_busy_s = st.gpu_busy_s;
  out->refused = st.refused;
  return CG_OK;
}

}  // extern "C"
