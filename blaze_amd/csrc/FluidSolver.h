// FluidSolver.h — C++ drop-in for the reference solver class.
//
// Reference: class FluidSolver (asiroisvigneux/Blaze src/FluidSolver.h:38-110).
// Same constructor and methods; Timer and SceneSettings are borrowed and
// must outlive the solver. All work runs on the MI355X through the C ABI
// (include/blaze.h). getTempGrid() returns a device view of the temperature
// front buffer instead of a device-heap Grid object.
#pragma once

#include <future>
#include <vector>

#include "../../include/blaze.h"
#include "SceneSettings.h"
#include "Timer.h"

struct GridHandle {
    const float* data; // device pointer, x-fastest i + j*W + k*W*H
    int width, height, depth;
    float ox, oy, oz; // origin offset in cells
    float dx;
};

class FluidSolver {
  public:
    FluidSolver(Timer* tmr, SceneSettings* scn);
    ~FluidSolver();
    void addSource();
    void step();
    GridHandle getTempGrid();
    blz_solver* handle() { return mH; }

  private:
    SceneSettings* mScn;
    Timer* mTmr;
    blz_solver* mH = nullptr;
    bool mSingleFrameSourceInit = false;
    int mParticleCount = 0;
    std::vector<float> mPart;
    // animated .part sources: the next frame is read and its upload started
    // on a worker thread while the current frame steps
    std::future<int> mPrefetch;
    int mPrefetchFrame = -1;
};
