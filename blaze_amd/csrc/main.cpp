// main.cpp — `blaze [-l] scene.blz`: runs the solver frame loop of a scene.
//
// Reference: src/main.cu:32-117 (argParse, frame loop). Frames inside the
// render range are rendered and written (RenderEngine) when the scene has
// the render keys; each frame prints the stats table of the reference Timer,
// and -l appends the scene summary and the tables to the timestamped log of
// the reference Logger.
#include <cstdio>
#include <cstring>
#include <fstream>
#include <string>

#include "FluidSolver.h"
#include "Logger.h"
#include "Parser.h"
#include "RenderEngine.h"

static void printHelp() {
    std::printf("blaze (MI355X) solver\n"
                "Usage: blaze [options] <blz scene file>\n"
                " -l             : log runtime performance data to a txt file\n"
                " -h             : print usage information\n");
}

static bool argParse(int argc, char* argv[], std::string& scenePath, bool& writeLog) {
    if (argc == 1 || (argc == 2 && std::strcmp(argv[1], "-l") == 0)) {
        std::printf("No scene file provided, blaze will be terminated\n");
        return false;
    }
    if (argc > 3 || (argc == 3 && std::strcmp(argv[1], "-l") != 0)) {
        std::printf("Only a single scene file should be provided, blaze will be terminated\n");
        return false;
    }
    if (!std::strcmp(argv[1], "-h") || !std::strcmp(argv[1], "-help") || !std::strcmp(argv[1], "--help")) {
        printHelp();
        return false;
    }
    writeLog = std::strcmp(argv[1], "-l") == 0;
    std::string fn = argv[writeLog ? 2 : 1];
    if (fn.substr(fn.find_last_of('.') + 1) != "blz") {
        std::printf("The provided scene file should have a .blz extention, blaze will be terminated\n");
        return false;
    }
    std::ifstream f(fn);
    if (!f.good()) {
        std::printf("Scene file %s does not exist on disk\n", fn.c_str());
        return false;
    }
    scenePath = fn;
    return true;
}

int main(int argc, char* argv[]) {
    std::string scenePath;
    bool writeLog = false;
    if (!argParse(argc, argv, scenePath, writeLog)) return 1;
    SceneSettings scn;
    Parser::sceneParse(&scn, scenePath.c_str());
    Logger* log = nullptr;
    if (writeLog) {
        log = new Logger(scn.sceneDir, scn.sceneName);
        log->append(Logger::sceneInfo(scn));
    }
    Timer tmr;
    FluidSolver solver(&tmr, &scn);
    RenderEngine* renderer = scn.hasRender ? new RenderEngine(&tmr, scn) : nullptr;
    if (!renderer) std::printf("scene has no render keys: frames are simulated, not rendered\n");
    tmr.time = 0.0f;
    tmr.iter = scn.sourceRange.x;
    const int last = scn.renderRange.y > 0 ? scn.renderRange.y : scn.sourceRange.y;
    while (tmr.iter <= last) {
        if (tmr.iter <= scn.sourceRange.y) solver.addSource();
        solver.step();
        if (renderer && tmr.iter >= scn.renderRange.x && tmr.iter <= scn.renderRange.y) {
            renderer->render(&solver);
            renderer->writeToDisk();
        }
        const std::string msg = tmr.getStats();
        std::fputs(msg.c_str(), stdout);
        if (log) log->append(msg);
        tmr.iter++;
        tmr.time += scn.dt;
    }
    delete renderer;
    delete log;
    return 0;
}
