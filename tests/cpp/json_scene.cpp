// buildFromJSON + advance + exportToOBJ through the drop-in C++ header
// (jsonBuilder.cpp:8-170, fileExport.cpp:63-100): the test compares the OBJ
// with the same scene stepped through the Python mirror.
#include <cstdio>
#include <cstdlib>

#include "YarnBall/YarnBall.h"

int main(int argc, char** argv) {
    if (argc < 4) {
        fprintf(stderr, "usage: json_scene scene.json seconds out.obj\n");
        return 2;
    }
    YarnBall::Sim* sim = YarnBall::buildFromJSON(argv[1]);
    sim->printErrors = false;
    sim->advance((float)atof(argv[2]));
    sim->exportToOBJ(argv[3]);
    printf("verts %d steps %zu\n", sim->meta.numVerts, sim->stepCount());
    delete sim;
    return 0;
}
