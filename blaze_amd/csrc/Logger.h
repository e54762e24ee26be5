// Logger.h — run log of the `blaze` command (-l).
//
// Reference: Logger (asiroisvigneux/Blaze src/Logger.h:31-63): the log goes
// to <sceneDir>/../logs/log_<scene>_<YYYY-MM-DD_HH:MM:SS>.txt (directory
// created if missing), opened for append on every message; the run starts
// with the scene summary, then one stats block per frame.
#pragma once

#include <cstdio>
#include <ctime>
#include <string>
#include <sys/stat.h>

#include "SceneSettings.h"

class Logger {
  public:
    Logger(const std::string& sceneDir, const std::string& sceneName) {
        const std::string dir = sceneDir + "/../logs/";
        std::time_t now = std::time(nullptr);
        std::tm tm{};
        localtime_r(&now, &tm);
        char stamp[64];
        std::strftime(stamp, sizeof(stamp), "%Y-%m-%d_%X", &tm);
        mFile = dir + "log_" + sceneName + "_" + stamp + ".txt";
        mkdir(dir.c_str(), S_IRWXU | S_IRWXG | S_IROTH | S_IXOTH);
    }
    void append(const std::string& msg) const {
        if (std::FILE* f = std::fopen(mFile.c_str(), "a")) {
            std::fputs(msg.c_str(), f);
            std::fclose(f);
        }
    }
    const std::string& path() const { return mFile; }

    // scene summary (the reference's SceneSettings::getSceneInfo lists the
    // same quantities)
    static std::string sceneInfo(const SceneSettings& s) {
        std::string m;
        char b[1024];
        std::snprintf(b, sizeof(b), "sceneName: %s\nsceneDir: %s\nparticleFile: %s\n", s.sceneName, s.sceneDir,
                      s.particleFile);
        m += b;
        std::snprintf(b, sizeof(b), "sourceAnimated: %d\nsourceRange: (%d, %d)\nrenderFile: %s\nrenderRange: (%d, %d)\n\n",
                      (int)s.sourceAnimated, s.sourceRange.x, s.sourceRange.y, s.renderFile, s.renderRange.x,
                      s.renderRange.y);
        m += b;
        std::snprintf(b, sizeof(b), "gridRes: (%d, %d, %d)\ndx: %4.4f\ndt: %4.4f\nmaxIterSolve: %d\nrenderRes: (%d, %d)\n\n",
                      s.gridRes.x, s.gridRes.y, s.gridRes.z, s.dx, s.dt, s.maxIterSolve, s.renderRes.x, s.renderRes.y);
        m += b;
        std::snprintf(b, sizeof(b),
                      "density: %4.4f\ngravity: %4.4f\nbuoyancy: %4.4f\ncoolingRate: %4.4f\nvorticityConf: %4.4f\n"
                      "drag: %4.4f\nturbulenceAmp: %4.4f\nturbulenceScale: %4.4f\n\n",
                      s.density, s.gravity, s.buoyancy, s.coolingRate, s.vorticityConf, s.drag, s.turbulence_amp,
                      s.turbulence_scale);
        m += b;
        std::snprintf(b, sizeof(b), "residualTol: %g\nresidualCheckEvery: %d\ncflNumber: %4.4f\nwarmStart: %d\n\n",
                      s.residualTol, s.residualCheckEvery, s.cflNumber, s.warmStart);
        m += b;
        return m;
    }

  private:
    std::string mFile;
};
