// facade.cpp — YarnBall::Sim implemented on the C ABI (drop-in for
// KittenEngine/YarnBall/YarnBall.cu and sim/step.cpp, sim/statistics.cpp).
#include "../../include/YarnBall/YarnBall.h"

#include <cstdio>
#include <cstring>

namespace YarnBall {

static_assert(sizeof(Kit::Rotor) == 16, "Rotor is 4 floats");
static_assert(sizeof(vec4) == 16 && sizeof(vec3) == 12, "glm-compatible sizes");

void Sim::raise(int code, const char* where) {
    std::string msg = yb_last_message();
    if (msg.empty()) msg = where;
    throw std::runtime_error(msg);
}

// Sim::Sim (YarnBall.cu:11-51).
Sim::Sim(int numVerts) {
    if (numVerts < 3) throw std::runtime_error("Too little vertices");
    yb_default_params(reinterpret_cast<yb_params*>(&meta), numVerts, maxH);
    verts = new Vertex[numVerts];
    vels = new vec3[numVerts];
    qs = new Kit::Rotor[numVerts];
    qRests = new vec4[numVerts];
    yb_default_vertices(numVerts, reinterpret_cast<yb_vertex*>(verts), &vels[0].x, &qs[0].x,
                        &qRests[0].x);
}

Sim::~Sim() {
    if (handle) yb_destroy(handle);
    delete[] verts;
    delete[] vels;
    delete[] qs;
    delete[] qRests;
}

Kit::LBVH::aabb Sim::bounds() {
    Kit::LBVH::aabb b;
    if (!handle) return b;
    float out[6];
    if (yb_bounds(handle, out) == YB_OK) {
        b.min = vec3(out[0], out[1], out[2]);
        b.max = vec3(out[3], out[4], out[5]);
    }
    return b;
}

// Sim::configure (YarnBall.cu:91-200).
void Sim::configure(float density) {
    const int n = meta.numVerts;
    initialInvMasses.resize(n);
    yb_params* p = reinterpret_cast<yb_params*>(&meta);
    int rc = yb_configure_host(n, reinterpret_cast<yb_vertex*>(verts), &qs[0].x, &qRests[0].x,
                               density, initialInvMasses.data(), p);
    if (rc != YB_OK) raise(rc, "configure");
    if (!handle) {
        rc = yb_create(n, &handle);
        if (rc != YB_OK) raise(rc, "configure");
    }
    rc = yb_validate_params(p);  // uploadMeta (YarnBall.cu:197)
    if (rc != YB_OK) raise(rc, "uploadMeta");
    upload();
}

void Sim::setKStretch(float k) {
    if (!handle) throw std::runtime_error("No rest length. Must call configure()");
    yb_set_kstretch_host(meta.numVerts, reinterpret_cast<yb_vertex*>(verts), k);
}

void Sim::setKBend(float k) {
    if (!handle) throw std::runtime_error("No rest length. Must call configure()");
    yb_set_kbend_host(meta.numVerts, reinterpret_cast<const yb_vertex*>(verts), &qRests[0].x, k);
}

void Sim::upload() {
    if (!handle) throw std::runtime_error("No rest length. Must call configure()");
    int rc = yb_upload(handle, reinterpret_cast<const yb_vertex*>(verts), &vels[0].x, &qs[0].x,
                       &qRests[0].x);
    if (rc != YB_OK) raise(rc, "upload");
}

void Sim::download() {
    if (!handle) return;
    int rc = yb_download(handle, reinterpret_cast<yb_vertex*>(verts), &vels[0].x, &qs[0].x);
    if (rc != YB_OK) raise(rc, "download");
}

void Sim::zeroVelocities() {
    if (!handle) return;
    int rc = yb_zero_velocities(handle);
    if (rc != YB_OK) raise(rc, "zeroVelocities");
}

// Sim::advance (step.cpp:36-55) + checkErrors (YarnBall.cu:276-304).
// Returns the simulated time advanced.
float Sim::advance(float h) {
    if (h <= 0) return 0;
    if (!handle) throw std::runtime_error("No rest length. Must call configure()");
    int warning = 0;
    int rc = yb_advance(handle, reinterpret_cast<yb_params*>(&meta), h, maxH, &warning);
    if (warning == YB_WARNING_SEGMENT_INTERPENETRATION) {
        lastWarningCode = warning;
        if (printErrors)
            fprintf(stderr,
                    "WARNING: Interpenetration detection. This can be due to unstable contacts\n");
    }
    if (rc == YB_ERROR_MAX_COLLISIONS_PER_SEGMENT_EXCEEDED) {
        lastErrorCode = rc;
        fprintf(stderr,
                "ERROR: MAX_COLLISIONS_PER_SEGMENT exceeded. Current simulation state may be "
                "corrupted!\n");
        throw std::runtime_error("MAX_COLLISIONS_PER_SEGMENT exceeded");
    }
    if (rc == YB_ERROR_TRAVERSAL_STACK_EXCEEDED) {
        lastErrorCode = rc;
        throw std::runtime_error("broadphase traversal stack exceeded");
    }
    if (rc != YB_OK) raise(rc, "advance");
    float out[6];
    if (yb_bounds(handle, out) == YB_OK) {
        currentBounds.min = vec3(out[0], out[1], out[2]);
        currentBounds.max = vec3(out[3], out[4], out[5]);
    }
    return h;
}

// Sim::step (step.cpp:57-59): one maxH step; dt is ignored as in the reference.
void Sim::step(float) { advance(maxH); }

// Sim::printCollisionStats (statistics.cpp:5-24).
void Sim::printCollisionStats() {
    if (!handle) return;
    std::vector<int32_t> numCols(meta.numVerts);
    if (yb_collision_counts(handle, numCols.data()) != YB_OK) return;
    size_t totCols = 0;
    int numSegWithCols = 0;
    for (int i = 0; i < meta.numVerts; i++)
        if (numCols[i] > 0) {
            totCols += numCols[i];
            numSegWithCols++;
        }
    totCols /= 2;
    printf("Total collisions: %zd\n", totCols);
    printf("Segments with collisions: %d/%d\n", numSegWithCols, meta.numVerts);
}

// ---- scene files ----------------------------------------------------------------
void Sim::exportToBCC(std::string path, bool exportAsPolyline) {
    download();
    if (yb_export_bcc(path.c_str(), meta.numVerts, reinterpret_cast<const yb_vertex*>(verts),
                      exportAsPolyline ? 1 : 0) != YB_OK)
        fprintf(stderr, "Error opening file\n");
}

void Sim::exportToOBJ(std::string path) {
    download();
    if (yb_export_obj(path.c_str(), meta.numVerts, reinterpret_cast<const yb_vertex*>(verts)) !=
        YB_OK)
        fprintf(stderr, "Error opening file\n");
}

void Sim::glueEndpoints(float searchRadius) {
    yb_glue_endpoints(meta.numVerts, reinterpret_cast<yb_vertex*>(verts), searchRadius);
}

Sim* Sim::fromScene(yb_scene* scene, bool configured) {
    int n = 0;
    yb_scene_size(scene, &n);
    Sim* sim = new Sim(n);
    sim->initialInvMasses.resize(n);
    yb_scene_get(scene, reinterpret_cast<yb_params*>(&sim->meta), &sim->maxH,
                 reinterpret_cast<yb_vertex*>(sim->verts), &sim->vels[0].x, &sim->qs[0].x,
                 &sim->qRests[0].x, sim->initialInvMasses.data());
    yb_scene_free(scene);
    if (configured) {
        int rc = yb_create(n, &sim->handle);
        if (rc == YB_OK) rc = yb_validate_params(reinterpret_cast<yb_params*>(&sim->meta));
        if (rc != YB_OK) {
            std::string msg = yb_last_message();
            delete sim;
            throw std::runtime_error(msg);
        }
        sim->upload();
    } else {
        sim->initialInvMasses.clear();
    }
    return sim;
}

static Sim* read_curves(const std::string& path, float segLen, const mat4& t, bool breakUp,
                        bool resample) {
    float tf[12];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 4; j++) tf[4 * i + j] = t[j][i];
    yb_scene* sc = nullptr;
    if (yb_read_curves(path.c_str(), segLen, tf, breakUp ? 1 : 0, resample ? 1 : 0, &sc) != YB_OK)
        throw std::runtime_error(yb_last_message());
    return Sim::fromScene(sc, false);
}

Sim* readFromOBJ(std::string path, float targetSegLen, mat4 transform, bool breakUpClosedCurves,
                 bool allowResample) {
    return read_curves(path, targetSegLen, transform, breakUpClosedCurves, allowResample);
}
Sim* readFromBCC(std::string path, float targetSegLen, mat4 transform, bool breakUpClosedCurves,
                 bool allowResample) {
    return read_curves(path, targetSegLen, transform, breakUpClosedCurves, allowResample);
}
Sim* readFromPoly(std::string path, float targetSegLen, mat4 transform, bool breakUpClosedCurves,
                  bool allowResample) {
    return read_curves(path, targetSegLen, transform, breakUpClosedCurves, allowResample);
}

Sim* buildFromJSON(std::string path) {
    yb_scene* sc = nullptr;
    if (yb_build_from_json(path.c_str(), &sc) != YB_OK) throw std::runtime_error(yb_last_message());
    return Sim::fromScene(sc, true);
}

}  // namespace YarnBall
