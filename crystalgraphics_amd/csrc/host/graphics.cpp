// graphics.cpp — the crystal::graphics API over the HIP backend's C-ABI, and
// C exports of the host layer for bindings.
//
// Mirrors the reference's facade (src/environment.cpp:12-74, src/view.cpp:
// 12-16): one process-global environment; View builds the SceneData and hands
// it to the backend.  Where the reference rebuilds SceneData on every View
// (src/environment.cpp:69-74) this layer caches it per scene tag, and uploads
// a scene once per tag as the reference's CacheTag does
// (src/wgpu/view.cpp:25-55).  FrameHistory (src/wgpu/frame_history.h:12-47)
// restarts accumulation when the scene or any camera float changes.
#include <chrono>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>

#include "../../../include/CrystalGraphics/graphics.h"
#include "../../../include/cg_device.h"
#include "scene_impl.h"

namespace crystal::graphics {

Scene::Scene(const std::filesystem::path& filepath, std::unique_ptr<Impl> impl)
    : filepath_(filepath), impl_(std::move(impl)) {}
Scene::~Scene() = default;
Scene::Scene(Scene&& other) noexcept = default;
Scene MakeScene(const std::filesystem::path& p, std::unique_ptr<Scene::Impl> impl) {
  return Scene(p, std::move(impl));
}

namespace {

struct FrameHistory {  // src/wgpu/frame_history.h:12-47
  bool has_previous = false;
  uint64_t tag = 0;
  cg_camera camera{};
  uint32_t frames = 0;
  static bool same(const cg_camera& a, const cg_camera& b) {
    for (int k = 0; k < 3; ++k)
      if (a.position[k] != b.position[k] || a.direction[k] != b.direction[k]) return false;
    return a.viewport[0] == b.viewport[0] && a.viewport[1] == b.viewport[1];
  }
  uint32_t next(uint64_t scene_tag, const cg_camera& cam) {
    if (!has_previous || tag != scene_tag || !same(camera, cam)) {
      has_previous = true;
      tag = scene_tag;
      camera = cam;
      frames = 1;
      return 0;
    }
    return frames++;
  }
};

struct Env {
  cg_env* backend = nullptr;
  PathTraceConf conf;
  FrameHistory history;
  uint64_t uploaded_tag = 0;
  cg_scene* uploaded = nullptr;
  std::map<uint64_t, std::unique_ptr<SceneData>> scene_data;
};
std::unique_ptr<Env> g_env;  // src/global.h:14

cg_camera to_cg(const Camera& c) {
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
int default_threads() { return static_cast<int>(std::max(1u, std::min(16u, std::thread::hardware_concurrency()))); }

expected<void, Error> view_impl(const Scene& scene, const cg_camera& cam) {
  if (!g_env) return unexpected<Error>(Error{"Environment not initialized."});  // src/view.cpp:13-14
  Env& e = *g_env;
  const Scene::Impl& impl = scene.impl();
  auto it = e.scene_data.find(impl.tag);
  if (it == e.scene_data.end())
    it = e.scene_data.emplace(impl.tag, std::make_unique<SceneData>(BuildSceneData(impl, true, default_threads())))
             .first;
  if (!e.uploaded || e.uploaded_tag != impl.tag) {
    if (e.uploaded) cg_scene_release(e.backend, e.uploaded);
    e.uploaded = nullptr;
    cg_scene_data view = it->second->View();
    if (cg_scene_upload(e.backend, &view, &e.uploaded) != CG_OK)
      return unexpected<Error>(Error{cg_env_last_error(e.backend)});
    e.uploaded_tag = impl.tag;
  }
  const uint32_t iter = e.history.next(impl.tag, cam);
  if (cg_view(e.backend, e.uploaded, &cam, iter) != CG_OK) return unexpected<Error>(Error{cg_env_last_error(e.backend)});
  return {};
}

}  // namespace

expected<void, Error> EnvInit(PathTraceConf conf) {  // src/environment.cpp:12-17
  if (g_env) return unexpected<Error>(Error{"Environment already initialized."});
  cg_conf c{conf.window_width,
            conf.window_height,
            conf.render_width,
            conf.render_height,
            conf.sample_count,
            conf.trace_depth,
            conf.max_transport_sample_count,
            conf.max_diffuse_sample_count,
            conf.max_rough_sample_count,
            conf.max_emission_sample_count};
  auto env = std::make_unique<Env>();
  if (cg_env_create(&c, 0, &env->backend) != CG_OK) return unexpected<Error>(Error{cg_last_error()});
  env->conf = conf;
  g_env = std::move(env);
  return {};
}

expected<void, Error> EnvTerm() {  // src/environment.cpp:18-23
  if (!g_env) return unexpected<Error>(Error{"Environment already terminated."});
  cg_env_destroy(g_env->backend);
  g_env.reset();
  return {};
}

expected<void, Error> EnvPoll() {
  if (!g_env) return unexpected<Error>(Error{"Environment not initialized."});
  if (cg_env_sync(g_env->backend) != CG_OK) return unexpected<Error>(Error{cg_env_last_error(g_env->backend)});
  return {};
}

expected<void, Error> EnvPresent() {
  if (!g_env) return unexpected<Error>(Error{"Environment not initialized."});
  return {};
}

bool EnvStatus() { return static_cast<bool>(g_env); }

expected<Scene, Error> LoadScene(std::filesystem::path file) {
  const std::string s = file.string();
  const std::string prefix = "procedural:";
  if (s.rfind(prefix, 0) != 0) {
    const std::string ext = file.extension().string();
    if (ext != ".gltf" && ext != ".glb") return unexpected<Error>(Error{"scene file not supported: " + s});
    auto impl = std::make_unique<Scene::Impl>();
    std::string err = LoadGltf(s, *impl);
    if (!err.empty()) return unexpected<Error>(Error{"glTF " + s + ": " + err});
    return MakeScene(file, std::move(impl));
  }
  std::string name = s.substr(prefix.size());
  uint64_t seed = 1;
  auto q = name.find('?');
  if (q != std::string::npos) {
    std::string args = name.substr(q + 1);
    name = name.substr(0, q);
    if (args.rfind("seed=", 0) == 0) seed = std::stoull(args.substr(5));
  }
  auto impl = std::make_unique<Scene::Impl>();
  std::string err = GenerateScene(name, seed, nullptr, 0, *impl);
  if (!err.empty()) return unexpected<Error>(Error{err});
  return MakeScene(file, std::move(impl));
}

expected<void, Error> View(const Scene& scene, const Camera& camera) { return view_impl(scene, to_cg(camera)); }

expected<void, Error> ViewToBuffer(const Scene& scene, const Camera& camera, float* xyz, std::size_t count) {
  if (!g_env) return unexpected<Error>(Error{"Environment not initialized."});
  const size_t px = static_cast<size_t>(g_env->conf.render_width) * g_env->conf.render_height;
  if (!xyz || count < px * 3) return unexpected<Error>(Error{"ViewToBuffer: buffer smaller than render_width * render_height * 3"});
  auto r = view_impl(scene, to_cg(camera));
  if (!r) return r;
  std::unique_ptr<cg_history_pixel[]> h(new cg_history_pixel[px]);
  if (cg_read_target(g_env->backend, h.get(), nullptr) != CG_OK)
    return unexpected<Error>(Error{cg_env_last_error(g_env->backend)});
  for (size_t i = 0; i < px; ++i) std::memcpy(xyz + 3 * i, h[i].xyz, 3 * sizeof(float));
  return {};
}

expected<std::vector<RenderResult>, Error> SubmitJobs(const std::vector<const Scene*>& scenes,
                                                      const std::vector<RenderRequest>& requests) {
  if (!g_env) return unexpected<Error>(Error{"Environment not initialized."});
  Env& e = *g_env;
  std::vector<RenderResult> out(requests.size());
  if (requests.empty()) return out;
  std::map<size_t, cg_scene*> uploaded;  // scene index -> device scene (released at the end unless it is e.uploaded)
  std::vector<void*> buffers;
  auto cleanup = [&] {
    for (void* b : buffers) cg_device_free(e.backend, b);
    for (auto& kv : uploaded)
      if (kv.second != e.uploaded) cg_scene_release(e.backend, kv.second);
  };
  auto fail = [&](const std::string& msg) {
    cleanup();
    return unexpected<Error>(Error{msg});
  };
  std::vector<cg_job> jobs(requests.size());
  for (size_t k = 0; k < requests.size(); ++k) {
    const RenderRequest& r = requests[k];
    if (r.scene >= scenes.size() || !scenes[r.scene]) return fail("SubmitJobs: request " + std::to_string(k) + " names no scene");
    const PathTraceConf& c = r.conf;
    if (c.render_width == 0 || c.render_height == 0) return fail("SubmitJobs: request " + std::to_string(k) + " has an empty size");
    if (!uploaded.count(r.scene)) {
      const Scene::Impl& impl = scenes[r.scene]->impl();
      if (e.uploaded && e.uploaded_tag == impl.tag) {
        uploaded[r.scene] = e.uploaded;
      } else {
        auto it = e.scene_data.find(impl.tag);
        if (it == e.scene_data.end())
          it = e.scene_data.emplace(impl.tag, std::make_unique<SceneData>(BuildSceneData(impl, true, default_threads())))
                   .first;
        cg_scene_data view = it->second->View();
        cg_scene* ds = nullptr;
        if (cg_scene_upload(e.backend, &view, &ds) != CG_OK) return fail(cg_env_last_error(e.backend));
        uploaded[r.scene] = ds;
      }
    }
    const size_t px = static_cast<size_t>(c.render_width) * c.render_height;
    void* h = nullptr;
    if (cg_device_alloc(e.backend, px * sizeof(cg_history_pixel), &h) != CG_OK) return fail(cg_env_last_error(e.backend));
    buffers.push_back(h);
    cg_job& j = jobs[k];
    j = cg_job{};
    j.scene = uploaded[r.scene];
    j.camera = to_cg(r.camera);
    j.conf = cg_conf{c.window_width, c.window_height, c.render_width, c.render_height, c.sample_count, c.trace_depth,
                     c.max_transport_sample_count, c.max_diffuse_sample_count, c.max_rough_sample_count,
                     c.max_emission_sample_count};
    j.iter_count = 0;
    j.history = static_cast<cg_history_pixel*>(h);
  }
  const auto t0 = std::chrono::steady_clock::now();
  cg_vbatch* vb = nullptr;
  if (cg_view_batch_launch(e.backend, jobs.data(), static_cast<uint32_t>(jobs.size()), &vb) != CG_OK)
    return fail(cg_env_last_error(e.backend));
  if (cg_view_batch_release(vb) != CG_OK) return fail(cg_env_last_error(e.backend));
  const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  for (size_t k = 0; k < requests.size(); ++k) {
    const uint32_t w = requests[k].conf.render_width, hgt = requests[k].conf.render_height;
    std::vector<cg_history_pixel> hist(static_cast<size_t>(w) * hgt);
    if (cg_memcpy_d2h(e.backend, hist.data(), jobs[k].history, hist.size() * sizeof(cg_history_pixel)) != CG_OK)
      return fail(cg_env_last_error(e.backend));
    RenderResult& o = out[k];
    o.id = k + 1;
    o.width = w;
    o.height = hgt;
    o.xyz.resize(hist.size() * 3);
    o.seeds.resize(hist.size());
    for (size_t i = 0; i < hist.size(); ++i) {
      std::memcpy(&o.xyz[3 * i], hist[i].xyz, 3 * sizeof(float));
      o.seeds[i] = hist[i].rand_seed;
    }
    o.queued_s = 0.0;
    o.latency_s = secs;
  }
  cleanup();
  return out;
}

}  // namespace crystal::graphics

// ===========================================================================
// C exports of the host layer (bindings, tests, benchmarks).
using namespace crystal::graphics;

namespace {
thread_local std::string t_err;
}

extern "C" {

typedef struct cgh_scene cgh_scene;          // a Scene
typedef struct cgh_scene_data cgh_scene_data;  // a SceneData

const char* cgh_last_error(void) { return t_err.c_str(); }

static cgh_scene* wrap_scene(const char* path, std::unique_ptr<Scene::Impl> impl) {
  return reinterpret_cast<cgh_scene*>(new Scene(MakeScene(path, std::move(impl))));
}
static const Scene::Impl& impl_of(const cgh_scene* s) { return reinterpret_cast<const Scene*>(s)->impl(); }

int cgh_scene_generate(const char* name, uint64_t seed, const double* params, int nparams, cgh_scene** out) {
  auto impl = std::make_unique<Scene::Impl>();
  std::string err = GenerateScene(name ? name : "", seed, params, nparams, *impl);
  *out = nullptr;
  if (!err.empty()) {
    t_err = err;
    return CG_ERR_INVALID;
  }
  *out = wrap_scene((std::string("procedural:") + name).c_str(), std::move(impl));
  return CG_OK;
}

int cgh_scene_create(const cg_vertex* v, uint32_t nv, const uint32_t* idx, uint32_t ni, const cg_material* m,
                     uint32_t nm, const cg_primitive* p, uint32_t np, const cg_texture* tex, uint32_t ntex,
                     const float* env, uint32_t ew, uint32_t eh, uint64_t tag, cgh_scene** out) {
  *out = nullptr;
  auto impl = std::make_unique<Scene::Impl>();
  for (uint32_t i = 0; i < ntex; ++i) {
    if (!tex[i].rgba8 || tex[i].width == 0 || tex[i].height == 0) {
      t_err = "texture " + std::to_string(i) + " has no image data";
      return CG_ERR_INVALID;
    }
    HostTexture t;
    t.width = tex[i].width;
    t.height = tex[i].height;
    t.wrap_s = tex[i].wrap_s;
    t.wrap_t = tex[i].wrap_t;
    t.rgba8.assign(tex[i].rgba8, tex[i].rgba8 + static_cast<size_t>(t.width) * t.height * 4);
    impl->textures.push_back(std::move(t));
  }
  impl->vertices.assign(v, v + nv);
  impl->indices.assign(idx, idx + ni);
  impl->materials.assign(m, m + nm);
  impl->prims.assign(p, p + np);
  if (env && ew && eh) {
    impl->env_rgba32f.assign(env, env + static_cast<size_t>(ew) * eh * 4);
    impl->env_w = ew;
    impl->env_h = eh;
  }
  impl->tag = tag;
  *out = wrap_scene("", std::move(impl));
  return CG_OK;
}

void cgh_scene_free(cgh_scene* s) { delete reinterpret_cast<Scene*>(s); }

typedef struct cgh_scene_desc {
  const cg_vertex* vertices;
  uint32_t vertex_count;
  const uint32_t* indices;
  uint32_t index_count;
  const cg_material* materials;
  uint32_t material_count;
  const cg_primitive* prims;
  uint32_t prim_count;
  const cg_texture* textures;  // valid while the scene lives
  uint32_t texture_count;
  const float* env;
  uint32_t env_w, env_h;
  cg_camera camera;
  cg_conf conf;
  uint64_t tag;
} cgh_scene_desc;

int cgh_scene_describe(const cgh_scene* s, cgh_scene_desc* d) {
  const auto* impl = &impl_of(s);
  impl->texture_views.clear();
  for (const auto& t : impl->textures)
    impl->texture_views.push_back(cg_texture{t.rgba8.data(), t.width, t.height, t.wrap_s, t.wrap_t});
  d->vertices = impl->vertices.data();
  d->vertex_count = static_cast<uint32_t>(impl->vertices.size());
  d->indices = impl->indices.data();
  d->index_count = static_cast<uint32_t>(impl->indices.size());
  d->materials = impl->materials.data();
  d->material_count = static_cast<uint32_t>(impl->materials.size());
  d->prims = impl->prims.data();
  d->prim_count = static_cast<uint32_t>(impl->prims.size());
  d->textures = impl->texture_views.empty() ? nullptr : impl->texture_views.data();
  d->texture_count = static_cast<uint32_t>(impl->texture_views.size());
  d->env = impl->env_rgba32f.empty() ? nullptr : impl->env_rgba32f.data();
  d->env_w = impl->env_w;
  d->env_h = impl->env_h;
  d->camera = impl->camera;
  d->conf = impl->conf;
  d->tag = impl->tag;
  return CG_OK;
}

int cgh_rays_generate(const cgh_scene* s, uint64_t seed, uint32_t n, cg_ray* out) {
  const auto* impl = &impl_of(s);
  if (impl->indices.empty()) {
    t_err = "scene has no triangles";
    return CG_ERR_INVALID;
  }
  GenerateRays(*impl, seed, n, out);
  return CG_OK;
}

int cgh_scene_data_build(const cgh_scene* s, int share_blas, int threads, cgh_scene_data** out) {
  *out = reinterpret_cast<cgh_scene_data*>(new SceneData(BuildSceneData(impl_of(s), share_blas != 0, threads)));
  return CG_OK;
}

int cgh_scene_data_view(const cgh_scene_data* sd, cg_scene_data* out) {
  *out = reinterpret_cast<const SceneData*>(sd)->View();
  return CG_OK;
}

void cgh_scene_data_free(cgh_scene_data* sd) { delete reinterpret_cast<SceneData*>(sd); }

int cgh_alias_build(const double* w, uint32_t n, cg_alias_entry* out) {
  std::vector<double> v(w, w + n);
  auto t = BuildAliasTable(v);
  std::memcpy(out, t.data(), n * sizeof(cg_alias_entry));
  return CG_OK;
}

// The crystal::graphics API, for bindings.
int cgh_env_init(const cg_conf* c) {
  PathTraceConf conf;
  conf.window_width = c->window_width;
  conf.window_height = c->window_height;
  conf.render_width = c->render_width;
  conf.render_height = c->render_height;
  conf.sample_count = c->sample_count;
  conf.trace_depth = c->trace_depth;
  conf.max_transport_sample_count = c->max_transport_sample_count;
  conf.max_diffuse_sample_count = c->max_diffuse_sample_count;
  conf.max_rough_sample_count = c->max_rough_sample_count;
  conf.max_emission_sample_count = c->max_emission_sample_count;
  auto r = EnvInit(conf);
  if (!r) {
    t_err = r.error();
    return CG_ERR_STATE;
  }
  return CG_OK;
}
int cgh_env_term(void) {
  auto r = EnvTerm();
  if (!r) {
    t_err = r.error();
    return CG_ERR_STATE;
  }
  return CG_OK;
}
int cgh_env_status(void) { return EnvStatus() ? 1 : 0; }
int cgh_env_poll(void) {
  auto r = EnvPoll();
  if (!r) {
    t_err = r.error();
    return CG_ERR_STATE;
  }
  return CG_OK;
}
static int wrap(expected<void, Error> r) {
  if (!r) {
    t_err = r.error();
    return CG_ERR_STATE;
  }
  return CG_OK;
}
static Camera from_cg(const cg_camera* c) {
  Camera cam;
  cam.position = vec3{c->position[0], c->position[1], c->position[2]};
  cam.direction = vec3{c->direction[0], c->direction[1], c->direction[2]};
  cam.viewport = vec2{c->viewport[0], c->viewport[1]};
  return cam;
}
int cgh_view(const cgh_scene* s, const cg_camera* cam) {
  return wrap(View(*reinterpret_cast<const Scene*>(s), from_cg(cam)));
}
int cgh_view_to_buffer(const cgh_scene* s, const cg_camera* cam, float* xyz, size_t count) {
  return wrap(ViewToBuffer(*reinterpret_cast<const Scene*>(s), from_cg(cam), xyz, count));
}
// SubmitJobs for bindings: requests as (scene index, camera, conf) arrays;
// xyz[k] receives W*H*3 floats of request k.
int cgh_submit_jobs(const cgh_scene* const* scenes, uint32_t nscenes, const uint32_t* scene_of, const cg_camera* cams,
                    const cg_conf* confs, uint32_t n, float* const* xyz, double* seconds) {
  if ((nscenes && !scenes) || (n && (!scene_of || !cams || !confs || !xyz))) {
    t_err = "cgh_submit_jobs: null argument";
    return CG_ERR_INVALID;
  }
  for (uint32_t k = 0; k < n; ++k)
    if (!xyz[k]) {
      t_err = "cgh_submit_jobs: null output for request " + std::to_string(k);
      return CG_ERR_INVALID;
    }
  std::vector<const Scene*> sc(nscenes);
  for (uint32_t i = 0; i < nscenes; ++i) sc[i] = reinterpret_cast<const Scene*>(scenes[i]);
  std::vector<RenderRequest> reqs(n);
  for (uint32_t k = 0; k < n; ++k) {
    reqs[k].scene = scene_of[k];
    reqs[k].camera = from_cg(&cams[k]);
    const cg_conf& c = confs[k];
    PathTraceConf p;
    p.window_width = c.window_width;
    p.window_height = c.window_height;
    p.render_width = c.render_width;
    p.render_height = c.render_height;
    p.sample_count = c.sample_count;
    p.trace_depth = c.trace_depth;
    p.max_transport_sample_count = c.max_transport_sample_count;
    p.max_diffuse_sample_count = c.max_diffuse_sample_count;
    p.max_rough_sample_count = c.max_rough_sample_count;
    p.max_emission_sample_count = c.max_emission_sample_count;
    reqs[k].conf = p;
  }
  auto r = SubmitJobs(sc, reqs);
  if (!r) {
    t_err = r.error();
    return CG_ERR_STATE;
  }
  for (uint32_t k = 0; k < n; ++k) std::memcpy(xyz[k], (*r)[k].xyz.data(), (*r)[k].xyz.size() * sizeof(float));
  if (seconds) *seconds = n ? (*r)[0].latency_s : 0.0;
  return CG_OK;
}

int cgh_load_scene(const char* path, cgh_scene** out) {
  *out = nullptr;
  auto r = LoadScene(path);
  if (!r) {
    t_err = r.error();
    return CG_ERR_INVALID;
  }
  *out = reinterpret_cast<cgh_scene*>(new Scene(std::move(*r)));
  return CG_OK;
}

}  // extern "C"
