// Serving form (no counterpart in the reference, which renders one View at a
// time): a request loop that keeps several scenes resident in HBM, admits
// render requests as path-state slots and output memory within a fixed HBM
// budget free up, batches every admitted request into the wavefront rounds
// as it arrives (continuous batching), and returns each request's running-
// mean CIE XYZ with its latency.  A request's result equals a plain View of
// the same scene, camera and conf, bit for bit, whatever else is in flight.
#ifndef CRYSTALGRAPHICS_SERVE_H_
#define CRYSTALGRAPHICS_SERVE_H_

#include <cstddef>
#include <cstdint>
#include <memory>
#include <vector>

#include "camera.h"
#include "environment.h"
#include "error.h"
#include "scene.h"

namespace crystal::graphics {

struct ServeConf {
  // kLaunches: each admission window becomes one persistent launch of the
  // per-pixel kernel over the admitted requests' tiles, several in flight on
  // their own streams.  kWavefront: requests join the wavefront rounds of one
  // batch engine (per-request completion, path state in a slot pool).
  enum class Engine { kLaunches, kWavefront };
  Engine engine = Engine::kLaunches;
  std::size_t hbm_budget_bytes = std::size_t{16} << 30;  // request outputs (+ the slot pool, wavefront)
  double output_fraction = 0.25;                         // wavefront: share of the budget for outputs
  std::uint32_t max_trace_depth = 8;
  std::uint32_t max_launches_in_flight = 4;              // kLaunches
  std::uint32_t rounds_per_launch = 16;                  // kWavefront
  std::uint32_t block_size = 0;   // kWavefront kernel shape, 0: default
  std::uint32_t short_stack = 0;
  // Overload policy: Submit refuses a request (error "Submit: queue full")
  // while this many requests wait for admission; 0 leaves the queue unbounded.
  std::uint32_t max_waiting = 0;
  // Random-number streams of every render (cg_device.h): 0 the reference's
  // one stream per pixel, 1 one stream per sample (kLaunches only: a pixel's
  // samples on separate waves).  RenderAlone uses the same mode.
  std::uint32_t rng_mode = 0;
};

struct RenderRequest {
  std::size_t scene = 0;  // index into the server's scenes
  Camera camera;
  PathTraceConf conf;     // render size, sample_count, trace_depth, lobe caps
};

struct RenderResult {
  std::uint64_t id = 0;
  std::uint32_t width = 0, height = 0;
  std::vector<float> xyz;             // width * height * 3, running-mean CIE XYZ
  std::vector<std::uint32_t> seeds;   // width * height carried RNG states
  double queued_s = 0;                // arrival -> admission
  double latency_s = 0;               // arrival -> result ready
};

struct ServeStats {
  std::uint64_t submitted = 0, admitted = 0, completed = 0;
  std::uint64_t refused = 0;          // turned away by the max_waiting bound
  std::uint64_t rays = 0;             // closest-hit + shadow queries traced
  double gpu_busy_s = 0;              // time with at least one request in flight
};

// Batch form (a one-shot call; no counterpart in the reference, which renders one View at a time):
// every request rendered in one launch of the per-pixel kernel through the
// environment of EnvInit (each request at its own size, spp and depth, a new
// accumulation), results in request order.  `scenes` are uploaded as needed
// and released afterwards.  latency_s is the batch's wall time.
expected<std::vector<RenderResult>, Error> SubmitJobs(const std::vector<const Scene*>& scenes,
                                                      const std::vector<RenderRequest>& requests);

class Server {
 public:
  // Uploads every scene (they stay resident) and sizes the slot pool and the
  // output arena from the budget.
  static expected<std::unique_ptr<Server>, Error> Create(const ServeConf& conf, const std::vector<const Scene*>& scenes);
  ~Server();
  // Thread-safe; the request arrives now.  Returns its id, or the error
  // "Submit: queue full" when max_waiting requests already wait.
  expected<std::uint64_t, Error> Submit(const RenderRequest& request);
  // The next finished request, waiting up to timeout_s; an empty result
  // (id 0) when none finished in time.
  expected<RenderResult, Error> Next(double timeout_s);
  // Render a request alone with the plain kernel (after the in-flight ones
  // finish): the reference result a served request must equal.
  expected<RenderResult, Error> RenderAlone(const RenderRequest& request);
  ServeStats Stats() const;

  struct Impl;

 private:
  explicit Server(std::unique_ptr<Impl> impl);
  std::unique_ptr<Impl> impl_;
};

}  // namespace crystal::graphics

#endif
