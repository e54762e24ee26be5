// Parser.h — .blz scene and .part particle readers of the C++ drop-in API.
//
// Reference: Parser (asiroisvigneux/Blaze src/Parser.h:38-221). Same entry
// points and argument meaning; on a missing file or key the reference
// asserts, this version prints the reason and exits with EXIT_FAILURE.
#pragma once

#include <string>
#include <vector>

#include "SceneSettings.h"

class Parser {
  public:
    static void sceneParse(SceneSettings* scene, const char* sceneFile);
    // Fills part8 (8 floats per particle, the .part record layout) and
    // pointCount; file name rule of src/Parser.h:181-188.
    static void sourceParticleParse(const SceneSettings* scn, std::vector<float>& part8, int& pointCount,
                                    int frameNum);
    static std::string getDirPath(const std::string& fname);
};
