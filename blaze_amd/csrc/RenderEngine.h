// RenderEngine.h — C++ drop-in for the reference renderer class.
//
// Reference: class RenderEngine (asiroisvigneux/Blaze src/RenderEngine.h:40-95,
// src/RenderEngine.cu:253-476). Same construction from the scene and the
// same render()/writeToDisk() pair; render() takes the solver (its handle
// carries the temperature grid and the domain) instead of a device-heap
// Grid. Frames are written as <sceneDir>/<renderFile stem>_%04d.<ext>
// (ppm, png or exr), the reference's naming.
#pragma once

#include <string>
#include <vector>

#include "../../include/blaze.h"
#include "SceneSettings.h"
#include "Timer.h"

class FluidSolver;

class RenderEngine {
  public:
    RenderEngine(Timer* tmr, const SceneSettings& scn);
    void render(FluidSolver* solver);
    void writeToDisk();
    const std::vector<float>& frame() const { return mFrame; }

  private:
    Timer* mTmr;
    blz_render_settings mSettings;
    std::string mRenderFile;
    std::string mSceneDir;
    std::vector<float> mFrame;
};
