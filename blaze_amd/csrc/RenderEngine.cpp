// RenderEngine.cpp — see RenderEngine.h.
#include "RenderEngine.h"

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <sys/stat.h>

#include "FluidSolver.h"
#include "Parser.h"

namespace {
void check(int rc, const char* what) {
    if (rc != BLZ_OK) {
        std::fprintf(stderr, "blaze: %s failed: %s\n", what, blz_last_error());
        std::exit(EXIT_FAILURE);
    }
}
} // namespace

RenderEngine::RenderEngine(Timer* tmr, const SceneSettings& scn)
    : mTmr(tmr), mSettings(scn.render), mRenderFile(scn.renderFile), mSceneDir(scn.sceneDir),
      mFrame((size_t)scn.render.width * scn.render.height * 3, 0.0f) {}

void RenderEngine::render(FluidSolver* solver) {
    const auto t0 = std::chrono::steady_clock::now();
    check(blz_render(solver->handle(), &mSettings, mFrame.data()), "render");
    mTmr->render_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

// src/RenderEngine.cu:367-382: name = renderFile minus its last 9
// characters ("_0001.ext"), then _%04d of the frame and the extension
void RenderEngine::writeToDisk() {
    const auto t0 = std::chrono::steady_clock::now();
    const std::string f(mRenderFile);
    if (f.size() < 9) {
        std::fprintf(stderr, "blaze: renderFile %s is not <name>_0001.<ext>\n", f.c_str());
        std::exit(EXIT_FAILURE);
    }
    const std::string stem = f.substr(0, f.size() - 9), ext = f.substr(f.size() - 3);
    char name[1024];
    std::snprintf(name, sizeof(name), "%s/%s_%04d.%s", mSceneDir.c_str(), stem.c_str(), mTmr->iter, ext.c_str());
    const std::string dir = Parser::getDirPath(name);
    mkdir(dir.c_str(), S_IRWXU | S_IRWXG | S_IROTH | S_IXOTH);
    check(blz_write_image(name, mFrame.data(), mSettings.width, mSettings.height), "writeToDisk");
    mTmr->write_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}
