// The reference's README example (README.md:104-126) compiled against the
// drop-in header, unchanged apart from the include line.  Prints the final
// vertex positions so the test can compare them with the Python mirror.
#include <cstdio>

#include "YarnBall/YarnBall.h"

using namespace YarnBall;

int main(int argc, char** argv) {
    float seconds = argc > 1 ? (float)atof(argv[1]) : 5.f;
    constexpr int numVerts = 64;
    auto sim = new YarnBall::Sim(numVerts);
    const float segLen = 0.002f;

    for (int i = 0; i < 32; i++) {
        vec3 pos = vec3(0.00002f * i * (i - 16) + segLen * 1, -segLen * i, 0);
        sim->verts[i].pos = pos;
        pos.x *= -1;
        sim->verts[i + 32].pos = pos;
    }

    sim->verts[0].invMass = sim->verts[32].invMass = sim->verts[63].invMass = 0;
    sim->verts[31].flags = 0;
    sim->meta.kCollision = 1e-7;
    sim->configure();
    sim->setKBend(1e-8);
    sim->setKStretch(2e-2);
    sim->maxH = 1e-3;
    sim->upload();

    sim->advance(seconds);
    sim->download();
    printf("steps %zu\n", sim->stepCount());
    for (int i = 0; i < numVerts; i++)
        printf("%.9e %.9e %.9e\n", sim->verts[i].pos.x, sim->verts[i].pos.y, sim->verts[i].pos.z);
    sim->printCollisionStats();
    delete sim;

    // Error behaviour of the reference (YarnBall.cu:12, 203).
    int errors = 0;
    try {
        YarnBall::Sim bad(2);
    } catch (const std::runtime_error&) {
        errors++;
    }
    try {
        YarnBall::Sim s(8);
        s.setKBend(1e-8);
    } catch (const std::runtime_error&) {
        errors++;
    }
    printf("errors %d\n", errors);
    return errors == 2 ? 0 : 1;
}
