// A segment crossed by 200 others (more than MAX_COLLISIONS_PER_SEGMENT = 128
// contacts), built against the drop-in header: advance() must print the
// reference's error line and throw (YarnBall.cu:283-286).
#include <cstdio>
#include <stdexcept>

#include "YarnBall/YarnBall.h"

using namespace YarnBall;

int main() {
    constexpr int m = 200;
    constexpr int numVerts = 2 + 2 * m;
    YarnBall::Sim sim(numVerts);
    sim.verts[0].pos = vec3(-1e-2f, 0, 0);
    sim.verts[1].pos = vec3(1e-2f, 0, 0);
    sim.verts[1].flags = 0;
    for (int k = 0; k < m; k++) {
        const float x = -9e-3f + k * 9e-5f;
        sim.verts[2 + 2 * k].pos = vec3(x, -5e-4f, 3e-4f);
        sim.verts[3 + 2 * k].pos = vec3(x, 5e-4f, 3e-4f);
        sim.verts[3 + 2 * k].flags = 0;
    }
    sim.verts[0].invMass = sim.verts[1].invMass = 0;
    sim.meta.gravity = vec3(0);
    sim.configure();
    sim.setKBend(1e-8f);
    sim.setKStretch(2e-2f);
    sim.maxH = 1e-3f;
    sim.upload();
    try {
        sim.advance(1e-3f);
        printf("no exception\n");
        return 1;
    } catch (const std::runtime_error& e) {
        printf("caught: %s\n", e.what());
    }
    printf("lastErrorCode %d\n", sim.lastErrorCode);
    return sim.lastErrorCode == YarnBall::Sim::ERROR_MAX_COLLISIONS_PER_SEGMENT_EXCEEDED ? 0 : 1;
}
