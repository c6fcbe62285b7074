"""The crystal::graphics C++ API (include/CrystalGraphics/*.h) as a C++ caller
of the reference's API would use it: a program using EnvInit, LoadScene,
ViewToBuffer, EnvPoll, EnvTerm and SubmitJobs compiles against the public
headers alone (C++20 and, where the compiler has it, C++23 with
std::expected) and links against the built host library.  Running it needs a
GPU; the facade tests drive the same functions through the C exports."""
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
LIB = ROOT / "crystalgraphics_amd" / "lib"

PROGRAM = r"""
#include <CrystalGraphics/graphics.h>
#include <CrystalGraphics/environment.h>
#include <CrystalGraphics/serve.h>
#include <cstdio>
#include <vector>

namespace cg = crystal::graphics;

int main(int argc, char** argv) {
  cg::PathTraceConf conf;
  conf.render_width = 64;
  conf.render_height = 36;
  conf.sample_count = 2;
  conf.trace_depth = 4;
  if (auto r = cg::EnvInit(conf); !r) { std::printf("EnvInit: %s\n", r.error().c_str()); return 1; }
  auto scene = cg::LoadScene(argc > 1 ? argv[1] : "procedural:spheres?seed=4");
  if (!scene) { std::printf("LoadScene: %s\n", scene.error().c_str()); return 1; }
  cg::Camera cam;
  cam.position = {0.0f, -6.0f, 1.0f};
  cam.direction = {0.0f, 1.0f, 0.0f};
  std::vector<float> xyz(64 * 36 * 3);
  if (auto r = cg::ViewToBuffer(*scene, cam, xyz.data(), xyz.size()); !r) return 1;
  (void)cg::EnvPoll();
  std::vector<const cg::Scene*> scenes{&*scene};
  std::vector<cg::RenderRequest> reqs(1);
  reqs[0].camera = cam;
  reqs[0].conf = conf;
  auto out = cg::SubmitJobs(scenes, reqs);
  (void)cg::EnvTerm();
  return out ? 0 : 1;
}
"""


def _std_supported(std: str) -> bool:
    r = subprocess.run(["g++", f"-std={std}", "-x", "c++", "-fsyntax-only", "-"], input="int main(){}",
                       capture_output=True, text=True)
    return r.returncode == 0


GLM_CALLER = r"""
// A stand-in with glm's vector shape (glm itself is not installed here): the
// reference's callers assign glm::vec3 / glm::vec2 into Camera and back.
namespace glm {
struct vec2 { float x, y; vec2(float a, float b) : x(a), y(b) {} };
struct vec3 { float x, y, z; vec3(float a, float b, float c) : x(a), y(b), z(c) {} };
}
#include <CrystalGraphics/camera.h>
#include <cstdio>

int main() {
  crystal::graphics::Camera cam;
  cam.position = glm::vec3(1.0f, 2.0f, 3.0f);
  cam.direction = glm::vec3{0.0f, 1.0f, 0.0f};
  cam.viewport = glm::vec2(0.25f, 0.75f);
  glm::vec3 back = cam.position;
  glm::vec2 vp = cam.viewport;
  crystal::graphics::Camera other{glm::vec3(4, 5, 6), {0, 0, 1}, {0.5f, 0.5f}};
  static_assert(alignof(crystal::graphics::Camera) == 16 && sizeof(crystal::graphics::Camera) == 48);
  std::printf("%g %g %g %g %g %g %g\n", back.x, back.y, back.z, vp.x, vp.y, other.position.z, cam.direction.y);
  return 0;
}
"""


@pytest.mark.parametrize("std", ["c++17", "c++20"])
def test_camera_takes_glm_shaped_vectors(tmp_path, std):
    src = tmp_path / "glm_caller.cpp"
    src.write_text(GLM_CALLER)
    exe = tmp_path / "glm_caller"
    r = subprocess.run(["g++", f"-std={std}", "-Wall", "-Wextra", "-Werror", f"-I{ROOT / 'include'}", str(src),
                        "-o", str(exe)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    out = subprocess.run([str(exe)], capture_output=True, text=True)
    assert out.stdout.split() == ["1", "2", "3", "0.25", "0.75", "6", "1"]


@pytest.mark.parametrize("std", ["c++20", "c++23"])
def test_api_program_compiles_and_links(tmp_path, std):
    if not _std_supported(std):
        pytest.skip(f"g++ has no -std={std}")
    if not (LIB / "libcrystal_graphics.so").exists():
        pytest.skip("host library not built")
    src = tmp_path / "api_demo.cpp"
    src.write_text(PROGRAM)
    exe = tmp_path / "api_demo"
    r = subprocess.run(["g++", f"-std={std}", "-Wall", "-Wextra", "-Werror", f"-I{ROOT / 'include'}", str(src),
                        "-o", str(exe), f"-L{LIB}", "-lcrystal_graphics", f"-Wl,-rpath,{LIB}"],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    assert exe.exists()
